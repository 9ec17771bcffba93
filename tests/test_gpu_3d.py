"""GPU parity of the 3D step (prm.dim = 3) through the C ABI against the
oracle's 3D restatement (oracle/armando_oracle3d.c, itself pinned against the
reference's dim = 3 routines).

Bar: bit-exact for the 3D cell keys, the (cell, id) order, cell ranges and
the neighbour pair sets (FP64 criterion and the FP32 cell + offset
criterion); FP64 fields within 1e-10 of the oracle after one step and after
20 steps; FP32-form fields within 1e-4 relative L2 of the FP64 form after
50 steps.
"""
import numpy as np
import pytest

import armando_amd
from armando_amd import cases
from oracle import orc

pytestmark = pytest.mark.gpu

FIELDS_IN = ("x", "y", "z", "vx", "vy", "vz", "rho", "u")


def banner(msg):
    print(f"[gpu] {msg}", flush=True)


def oracle_state(case):
    st = orc.OracleState3(case["n"], case["np"])
    st.set(**{k: case[k] for k in FIELDS_IN})
    return st


def relmax(a, b):
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / scale


def dam(prec=cases.FP64, scale=0.1):
    return cases.dam_break3d_small(scale=scale, precision=prec)


def box(prec=cases.FP64):
    return cases.box_lattice3(nx=14, ny=12, nz=10, jitter=0.45, vel=0.5, seed=17,
                              precision=prec)


@pytest.mark.parametrize("name", ["dam", "box"])
def test_link_list3_cells_and_pairs_fp64(name):
    case = dam() if name == "dam" else box()
    banner(f"3D link_list fp64 {name} n={case['n']}")
    sim = armando_amd.Simulation(case)
    niac = sim.link_list(count=True)
    P = orc.params_from(case)
    key, bad = orc.keys3(P, case["x"], case["y"], case["z"])
    assert bad == 0
    ncell = case["nx"] * case["ny"] * case["nz"]
    sid, start = orc.sort_cells(key, ncell)
    gk, gs, gc = sim.cells()
    assert np.array_equal(gs, sid), "sort permutation"
    assert np.array_equal(gk, key[sid]), "sorted keys"
    assert np.array_equal(gc, start), "cell ranges"
    oi, oj = orc.pairs3_f64(P, case["x"], case["y"], case["z"])
    gi, gj = sim.pairs()
    assert niac == len(oi) and niac > 10 * case["n"] // 2
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj), "neighbour pairs"
    sim.close()


def test_link_list3_fp32_form_pairs():
    case = box(cases.FP32)
    banner(f"3D link_list fp32 n={case['n']}")
    sim = armando_amd.Simulation(case)
    niac = sim.link_list(count=True)
    P = orc.params_from(case)
    cells, offs = orc.split3_f32(P, case["x"], case["y"], case["z"])
    oi, oj = orc.pairs3_f32(P, cells, offs)
    gi, gj = sim.pairs()
    assert niac == len(oi)
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    sim.close()


@pytest.mark.parametrize("name", ["dam", "box"])
def test_routines3_fp64(name):
    """density -> eos -> internal_force(+av, +ah) -> av_vel, one routine at a
    time, against the oracle's 3D single step."""
    case = dam() if name == "dam" else box()
    banner(f"3D routines fp64 {name}")
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
    orc.single_step3(P, st)
    for k in ("drho", "ax", "ay", "az", "du", "avx", "avy", "avz", "divv"):
        ref = st.field(k)
        if not np.any(ref):
            assert not np.any(g[k]), k
            continue
        assert relmax(g[k], ref) < 1e-10, (k, relmax(g[k], ref))
    assert relmax(s["p"], st.field("p")) < 1e-12
    sim.close()


def test_summation_density3_fp64():
    case = dict(box())
    case["density_method"] = 1
    banner("3D sum_density fp64")
    sim = armando_amd.Simulation(case)
    sim.link_list()
    sim.density(1)
    s = sim.state()
    P = orc.params_from(case)
    o = orc.routines3(P, case["x"], case["y"], case["z"], case["vx"], case["vy"], case["vz"],
                      case["rho"], np.zeros(case["n"]), case["u"], np.full(case["n"], case["c0"]))
    assert relmax(s["rho"], o["rho_sum"]) < 1e-13
    sim.close()


def test_time_integration3_fp64():
    case = dam()
    nsteps = 20
    banner(f"3D time_integration fp64 n={case['n']} steps={nsteps}")
    sim = armando_amd.Simulation(case)
    sim.time_integration(1, nsteps)
    s = sim.state()
    P = orc.params_from(case)
    st = oracle_state(case)
    orc.run3(P, st, 1, nsteps)
    for k in ("x", "y", "z", "rho", "u"):
        assert relmax(s[k], st.field(k)) < 1e-10, k
    for k in ("vx", "vy", "vz"):
        ref = st.field(k)
        assert np.max(np.abs(s[k] - ref)) < 1e-10 * max(1.0, np.max(np.abs(ref))), k
    assert np.max(np.abs(s["vz"][:case["np"]])) > 0
    sim.close()


def test_per_routine_loop3_equals_fused_loop():
    case = dam()
    banner("3D per-routine vs fused")
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
    for k in ("x", "y", "z", "vx", "vy", "vz", "rho", "u", "p"):
        assert np.array_equal(sa[k], sb[k]), k
    a.close()
    b.close()


def test_fp32_form3_close_to_fp64():
    banner("3D fp32 vs fp64, 50 steps")
    a = armando_amd.Simulation(dam())
    b = armando_amd.Simulation(dam(cases.FP32))
    a.time_integration(1, 50)
    b.time_integration(1, 50)
    sa, sb = a.state(), b.state()
    for k in ("rho", "p", "vx", "vy", "vz", "x", "y", "z"):
        rel = np.linalg.norm(sb[k] - sa[k]) / np.linalg.norm(sa[k])
        banner(f"  {k} rel-L2 {rel:.3e}")
        assert rel < 1e-4, k
    a.close()
    b.close()


def test_momentum3_conserved_without_gravity():
    """No walls, no gravity: pairwise-antisymmetric forces and XSPH keep the
    total momentum at its initial value to round-off (FP64 sums)."""
    case = box()
    banner("3D momentum conservation")
    sim = armando_amd.Simulation(case)
    c0 = sim.conservation3()
    sim.time_integration(1, 20)
    c1 = sim.conservation3()
    p_scale = c0[0] * 0.5
    assert c1[0] == c0[0]
    for k in (1, 2, 3):
        assert abs(c1[k] - c0[k]) < 1e-10 * p_scale, (k, c0[k], c1[k])
    c5 = sim.conservation()
    assert np.allclose(c5, [c1[0], c1[1], c1[2], c1[4], c1[5]], rtol=1e-15, atol=0)
    sim.close()


def test_3d_rejects_2d_only_options():
    case = dict(box())
    case["flags"] = case["flags"] | cases.VISC
    case["visc_eta"] = 1e-3
    with pytest.raises(armando_amd.ArmandoError) as e:
        armando_amd.Simulation(case)
    assert e.value.code == -5


def test_dense3_rows_overflow_masks():
    """h = 2 dx: rows of more than 96 candidates, so the density pass marks
    the masks incomplete and the force pass tests the candidates again; the
    FP64 step still matches the oracle and the FP32 form the FP64 form."""
    for shape in ((40, 4, 4), (16, 14, 12)):
        got = {}
        for prec in (cases.FP64, cases.FP32):
            case = cases.box_lattice3(nx=shape[0], ny=shape[1], nz=shape[2], jitter=0.3,
                                      vel=0.4, seed=29, h_fac=2.0, precision=prec)
            banner(f"3D dense rows {shape} prec={prec} n={case['n']}")
            sim = armando_amd.Simulation(case)
            sim.link_list()
            sim.single_step()
            got[prec] = sim.rates()
            sim.close()
        P = orc.params_from(case)
        st = oracle_state(case)
        orc.single_step3(P, st)
        for k in ("drho", "ax", "ay", "az", "du", "avx", "avy", "avz"):
            assert relmax(got[cases.FP64][k], st.field(k)) < 1e-10, k
            assert relmax(got[cases.FP32][k], st.field(k)) < 1e-3, k
