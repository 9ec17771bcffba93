"""GPU parity of the HIP library (through the C ABI) against the oracle.

Bar: bit-exact for cell keys, the (cell, id) order, cell ranges and the
neighbour pair sets; FP64 fields within 1e-10 of the oracle (the two sum the
same terms in the same order, differing only by the last bit of q = r/h);
FP32-form fields within the FP32 tolerances stated per test.
"""
import numpy as np
import pytest

import armando_amd
from armando_amd import cases
from oracle import orc

pytestmark = pytest.mark.gpu

FIELDS_IN = ("x", "y", "vx", "vy", "rho", "u")


def banner(msg):
    print(f"[gpu] {msg}", flush=True)


def oracle_state(case):
    st = orc.OracleState(case["n"], case["np"])
    st.set(**{k: case[k] for k in FIELDS_IN})
    return st


def relmax(a, b):
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / scale


def small_cases():
    return {
        "dam_fp64": cases.dam_break_small(scale=0.05),
        "box_fp64": cases.box_lattice(vel=0.5),
    }


@pytest.mark.parametrize("name", ["dam_fp64", "box_fp64"])
def test_link_list_cells_and_pairs_fp64(name):
    case = small_cases()[name]
    banner(f"link_list fp64 {name} n={case['n']}")
    sim = armando_amd.Simulation(case)
    niac = sim.link_list(count=True)
    P = orc.params_from(case)
    key, bad = orc.keys(P, case["x"], case["y"])
    assert bad == 0
    ncell = case["nx"] * case["ny"]
    sid, start = orc.sort_cells(key, ncell)
    gk, gs, gc = sim.cells()
    assert np.array_equal(gs, sid), "sort permutation"
    assert np.array_equal(gk, key[sid]), "sorted keys"
    assert np.array_equal(gc, start), "cell ranges"
    oi, oj = orc.pairs_f64(P, case["x"], case["y"])
    fi, fj = orc.link_list(P, case["x"], case["y"])  # link_list.f restated
    order = np.lexsort((fj, fi))
    assert np.array_equal(oi, fi[order]) and np.array_equal(oj, fj[order])
    gi, gj = sim.pairs()
    assert niac == len(oi)
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj), "neighbour pairs"
    sim.close()


def test_link_list_fp32_form_pairs():
    case = cases.dam_break_small(scale=0.05, precision=cases.FP32)
    banner(f"link_list fp32 n={case['n']}")
    sim = armando_amd.Simulation(case)
    niac = sim.link_list(count=True)
    P = orc.params_from(case)
    cx, cy, ox, oy = orc.split_f32(P, case["x"], case["y"])
    oi, oj = orc.pairs_f32(P, cx, cy, ox, oy)
    gi, gj = sim.pairs()
    assert niac == len(oi)
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    sim.close()


@pytest.mark.parametrize("name", ["dam_fp64", "box_fp64"])
def test_routines_fp64(name):
    """density -> eos -> internal_force(+av, +ah) -> av_vel, one routine at a time."""
    case = small_cases()[name]
    banner(f"routines fp64 {name}")
    sim = armando_amd.Simulation(case)
    sim.link_list()
    sim.density(0)
    sim.eos()
    sim.internal_force(case["flags"])
    sim.av_vel()
    g = sim.rates()
    s = sim.state()
    P = orc.params_from(case)
    st = oracle_state(case)
    orc.single_step_gather(P, st)
    for k in ("drho", "ax", "ay", "du", "avx", "avy", "divv"):
        assert relmax(g[k], st.field(k)) < 1e-10, k
    assert relmax(s["p"], st.field("p")) < 1e-12
    assert relmax(s["rho"], st.field("rho")) < 1e-14
    sim.close()


def test_single_step_fp64_matches_faithful_half_list():
    """The gathered single step against the serial half-list restatement of
    single_step.f (different summation order)."""
    case = cases.dam_break_small(scale=0.05)
    banner("single_step vs faithful")
    sim = armando_amd.Simulation(case)
    sim.link_list()
    sim.single_step()
    g = sim.rates()
    P = orc.params_from(case)
    st = oracle_state(case)
    pi, pj = orc.link_list(P, case["x"], case["y"])
    import ctypes as C
    rc = orc.lib().orc_single_step(C.byref(P), st.ptr, len(pi), pi, pj)
    assert rc == 0
    for k in ("drho", "ax", "ay", "du", "avx", "avy"):
        assert relmax(g[k], st.field(k)) < 1e-9, k
    sim.close()


@pytest.mark.parametrize("nsteps", [1, 12])
def test_time_integration_fp64(nsteps):
    case = cases.dam_break_small(scale=0.05)
    banner(f"time_integration fp64 steps={nsteps}")
    sim = armando_amd.Simulation(case)
    sim.time_integration(1, nsteps)
    s = sim.state()
    P = orc.params_from(case)
    st = oracle_state(case)
    orc.run(P, st, 1, nsteps, gather=True)
    for k in ("x", "y", "rho", "u"):
        assert relmax(s[k], st.field(k)) < 1e-10, k
    for k in ("vx", "vy"):
        assert np.max(np.abs(s[k] - st.field(k))) < 1e-10 * max(1.0, np.max(np.abs(st.field(k))))
    sim.close()


def test_per_routine_loop_equals_fused_loop():
    """predictor -> link_list -> single_step -> corrector, driven routine by
    routine, must reproduce the fused time_integration bit for bit."""
    case = cases.dam_break_small(scale=0.05)
    banner("per-routine vs fused")
    a = armando_amd.Simulation(case)
    a.time_integration(1, 6)
    sa = a.state()
    b = armando_amd.Simulation(case)
    for its in range(1, 7):
        if its != 1:
            b.predictor()
        b.link_list()
        b.single_step()
        b.corrector(its)
    sb = b.state()
    for k in ("x", "y", "vx", "vy", "rho", "u", "p"):
        assert np.array_equal(sa[k], sb[k]), k
    a.close()
    b.close()


def test_fp32_form_close_to_fp64():
    c64 = cases.dam_break_small(scale=0.05)
    c32 = cases.dam_break_small(scale=0.05, precision=cases.FP32)
    banner("fp32 vs fp64, 50 steps")
    a = armando_amd.Simulation(c64)
    b = armando_amd.Simulation(c32)
    a.time_integration(1, 50)
    b.time_integration(1, 50)
    sa, sb = a.state(), b.state()
    for k in ("rho", "p", "vx", "vy", "x", "y"):
        rel = np.linalg.norm(sb[k] - sa[k]) / np.linalg.norm(sa[k])
        banner(f"  {k} rel-L2 {rel:.3e}")
        assert rel < 1e-4, k
    a.close()
    b.close()


def test_domain_error_is_reported():
    case = cases.box_lattice()
    case = dict(case)
    x = case["x"].copy()
    x[5] = case["ox"] - 1.0  # left of the hash grid
    case["x"] = x
    banner("domain error")
    with pytest.raises(armando_amd.ArmandoError) as e:
        armando_amd.Simulation(case).link_list()
    assert e.value.code == -4


@pytest.mark.parametrize("prec", [cases.FP64, cases.FP32])
def test_tiled_gather_equals_direct(prec):
    """LDS-tiled and direct gathers visit the same neighbours in the same
    order with the same arithmetic: identical bits after several steps."""
    case = cases.dam_break_small(scale=0.05, precision=prec)
    banner(f"tiled vs direct prec={prec}")
    a = armando_amd.Simulation(case)
    b = armando_amd.Simulation(case)
    b.set_gather_path(False)
    a.time_integration(1, 8)
    b.time_integration(1, 8)
    sa, sb = a.state(), b.state()
    ra, rb = a.rates(), b.rates()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    for k in ra:
        assert np.array_equal(ra[k], rb[k]), k
    a.close()
    b.close()


@pytest.mark.parametrize("prec", [cases.FP64, cases.FP32])
def test_dense_tiles_overflow_fall_back(prec):
    """h = 5 dx: ~300 neighbours per particle, so the staged rows exceed the
    LDS capacity (direct fallback) and neighbour lists flush many times."""
    case = cases.box_lattice(nx=48, ny=40, vel=0.3, precision=prec, h_fac=5.0)
    banner(f"dense tiles prec={prec} n={case['n']}")
    a = armando_amd.Simulation(case)
    b = armando_amd.Simulation(case)
    b.set_gather_path(False)
    for s in (a, b):
        s.link_list()
        s.single_step()
    ra, rb = a.rates(), b.rates()
    for k in ra:
        assert np.array_equal(ra[k], rb[k]), k
    if prec == cases.FP64:
        P = orc.params_from(case)
        st = oracle_state(case)
        orc.single_step_gather(P, st)
        for k in ("drho", "ax", "ay", "du", "avx", "avy"):
            assert relmax(ra[k], st.field(k)) < 1e-10, k
    a.close()
    b.close()


def test_medium_tiles_flush_lists():
    """h = 2 dx: ~50 neighbours, more than one list chunk per particle."""
    case = cases.box_lattice(nx=40, ny=50, vel=0.3, h_fac=2.0)
    banner("medium tiles")
    a = armando_amd.Simulation(case)
    a.link_list()
    a.single_step()
    ra = a.rates()
    P = orc.params_from(case)
    st = oracle_state(case)
    orc.single_step_gather(P, st)
    for k in ("drho", "ax", "ay", "du", "avx", "avy"):
        assert relmax(ra[k], st.field(k)) < 1e-10, k
    a.close()


def test_radix_sort_path_on_a_large_grid():
    """Grids above 2048 x 4096 cells take the radix sort instead of the
    counting sort: same keys, permutation, ranges and pairs as the oracle."""
    case = dict(cases.box_lattice(nx=30, ny=20, vel=0.3))
    case["nx"], case["ny"] = 4200, 2100  # 8.8M cells, particles near the origin corner
    banner(f"radix path grid {case['nx']}x{case['ny']}")
    sim = armando_amd.Simulation(case)
    niac = sim.link_list(count=True)
    P = orc.params_from(case)
    key, bad = orc.keys(P, case["x"], case["y"])
    assert bad == 0
    sid, start = orc.sort_cells(key, case["nx"] * case["ny"])
    gk, gs, gc = sim.cells()
    assert np.array_equal(gs, sid) and np.array_equal(gk, key[sid]) and np.array_equal(gc, start)
    oi, oj = orc.pairs_f64(P, case["x"], case["y"])
    gi, gj = sim.pairs()
    assert niac == len(oi) and np.array_equal(gi, oi) and np.array_equal(gj, oj)
    sim.time_integration(1, 3)
    st = oracle_state(case)
    orc.run(P, st, 1, 3, gather=True)
    s = sim.state()
    for k in ("x", "y", "rho", "u"):
        assert relmax(s[k], st.field(k)) < 1e-10, k
    sim.close()
