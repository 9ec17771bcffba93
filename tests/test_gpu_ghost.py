"""GPU parity of the mirror ghosts (ghost.f) against the oracle.

Bar: FP64 the ghost set (count, parents) identical and the ghosts' positions,
velocities, rho, p, u within 1e-12 after one step; 20 steps of the fused
loop within 1e-10 of the oracle (states and ghosts); the per-routine
sequence predictor -> ghost_particles -> link_list -> single_step ->
corrector bit-identical to the fused loop; LDS-tiled and direct gathers
bit-identical; FP32 within 1e-4 relative L2 of FP64 after 50 steps.
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


def tank(prec=cases.FP64):
    return cases.ghost_tank(col_w=0.02, col_h=0.03, tank_w=0.05, precision=prec)


def oracle_state(c):
    st = orc.GhostState(c["np"], c["nv"])
    st.set(**{k: c[k] for k in FIELDS_IN})
    st.set(c=np.full(c["n"], c["c0"]), p=np.zeros(c["n"]))
    return st


def relmax(a, b):
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / scale


def compare(sim, st, c, tol, gtol):
    s = sim.state()
    for k in ("x", "y", "rho", "u", "p"):
        err = relmax(s[k], st.field(k)[:c["n"]])
        banner(f"  {k} err {err:.2e}")
        assert err < tol, k
    for k in ("vx", "vy"):
        ref = st.field(k)[:c["n"]]
        assert np.max(np.abs(s[k] - ref)) < tol * max(1.0, np.max(np.abs(ref))), k
    g = sim.ghosts()
    nb = c["n"]
    assert len(g["x"]) == st.ng
    assert np.array_equal(g["parent"], st.parents())
    for k in ("x", "y", "vx", "vy", "rho", "p", "u"):
        ref = st.field(k)[nb:]
        err = float(np.max(np.abs(g[k] - ref))) / max(float(np.max(np.abs(ref))), 1.0)
        banner(f"  ghost {k} err {err:.2e}")
        assert err < gtol, k


def test_ghost_first_step_fp64():
    c = tank()
    banner(f"ghost tank step 1 fp64 n={c['n']}")
    sim = armando_amd.Simulation(c)
    sim.time_integration(1, 1)
    P = orc.params_from(c)
    st = oracle_state(c)
    orc.run(P, st, 1, 1, gather=True)
    assert st.ng > 50 and sim.ghost_count() == st.ng
    compare(sim, st, c, 1e-12, 1e-12)
    sim.close()


def test_ghost_time_integration_fp64():
    c = tank()
    nsteps = 20
    banner(f"ghost tank time_integration fp64 steps={nsteps}")
    sim = armando_amd.Simulation(c)
    sim.time_integration(1, 10)
    sim.time_integration(11, nsteps - 10)  # resumes with the stored search positions
    P = orc.params_from(c)
    st = oracle_state(c)
    orc.run(P, st, 1, nsteps, gather=True)
    compare(sim, st, c, 1e-10, 1e-10)
    sim.close()


def test_ghost_routines_equal_fused():
    c = tank()
    banner("ghost per-routine sequence vs fused loop, 6 steps")
    a = armando_amd.Simulation(c)
    b = armando_amd.Simulation(c)
    a.time_integration(1, 6)
    b.link_list()
    for its in range(1, 7):
        if its != 1:
            b.predictor()
        ng = b.ghost_particles()
        assert ng > 0
        b.link_list()
        b.single_step()
        b.corrector(its)
    sa, sb = a.state(), b.state()
    for k in ("x", "y", "vx", "vy", "rho", "u", "p"):
        assert np.array_equal(sa[k], sb[k]), k
    ga, gb = a.ghosts(), b.ghosts()
    assert len(ga["x"]) > 0 and np.array_equal(ga["parent"], gb["parent"])
    for k in ("x", "y", "vx", "vy", "rho", "p", "u"):
        assert np.array_equal(ga[k], gb[k]), k
    b.link_list()  # without a new set the ghosts drop out
    assert b.ghost_count() == 0
    a.close()
    b.close()


@pytest.mark.parametrize("prec", [cases.FP64, cases.FP32])
def test_ghost_tiled_equals_direct(prec):
    c = tank(prec)
    banner(f"ghost tiled vs direct prec={prec}")
    a = armando_amd.Simulation(c)
    b = armando_amd.Simulation(c)
    b.set_gather_path(False)
    a.time_integration(1, 10)
    b.time_integration(1, 10)
    sa, sb = a.state(), b.state()
    for k in ("x", "y", "vx", "vy", "rho", "u", "p"):
        assert np.array_equal(sa[k], sb[k]), k
    a.close()
    b.close()


def test_ghost_fp32_close_to_fp64():
    banner("ghost tank fp32 vs fp64, 50 steps")
    a = armando_amd.Simulation(tank())
    b = armando_amd.Simulation(tank(cases.FP32))
    a.time_integration(1, 50)
    b.time_integration(1, 50)
    sa, sb = a.state(), b.state()
    sa["v"] = np.concatenate([sa["vx"], sa["vy"]])
    sb["v"] = np.concatenate([sb["vx"], sb["vy"]])
    for k in ("rho", "v", "x", "y"):
        rel = np.linalg.norm(sb[k] - sa[k]) / np.linalg.norm(sa[k])
        banner(f"  {k} rel-L2 {rel:.3e}")
        assert rel < 1e-4, k
    assert a.ghost_count() == b.ghost_count()
    a.close()
    b.close()


def test_ghost_rejections():
    c = dict(tank())
    c["flags"] = c["flags"] | cases.DBC
    with pytest.raises(armando_amd.ArmandoError) as e:
        armando_amd.Simulation(c)
    assert e.value.code == -5
    sim = armando_amd.Simulation(tank())
    with pytest.raises(armando_amd.ArmandoError) as e:
        sim.adjoint_rollout(2, 2, np.array([0], np.int32))
    assert e.value.code == -5
    with pytest.raises(armando_amd.ArmandoError):
        sim.ghost_particles()  # no search yet
    sim.close()


def test_cells_with_ghosts_hold_the_live_set():
    """arm_get_cells after ghost_particles -> link_list: keys and ids of the
    np + nv + ng live particles (ghost ids np + nv + g) in (cell, id) order,
    and the cell ranges, against the oracle's keys and stable sort of the
    same positions (the ghosts' as downloaded)."""
    c = tank()
    banner(f"cells with ghosts n={c['n']}")
    sim = armando_amd.Simulation(c)
    sim.link_list()
    ng = sim.ghost_particles()
    sim.link_list()
    assert ng > 0 and sim.ghost_count() == ng
    k, ids, cs = sim.cells()
    assert len(k) == c["n"] + ng and len(ids) == c["n"] + ng
    s, g = sim.state(), sim.ghosts()
    sim.close()
    x = np.concatenate([s["x"], g["x"]])
    y = np.concatenate([s["y"], g["y"]])
    P = orc.params_from(c)
    key, bad = orc.keys(P, x, y)
    assert bad == 0
    sid, start = orc.sort_cells(key, int(c["nx"]) * int(c["ny"]))
    assert np.array_equal(ids, sid)
    assert np.array_equal(k, key[sid])
    assert np.array_equal(cs, start)
