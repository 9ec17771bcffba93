"""Edge cases of the GPU step against the oracle: a single particle, walls
only, no walls, pairs at exactly the support 2h, particles exactly on cell
faces, coincident particles, and the capacity errors of the C ABI.

Bar: pair sets and cell keys bit-exact; FP64 rates within 1e-12 of the
oracle's gather twin (the same sums), no NaN anywhere.
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


def subset(c, idx, np_new):
    """The particles idx of a case (the first np_new of them real)."""
    d = dict(c)
    for k in FIELDS_IN:
        d[k] = np.ascontiguousarray(c[k][idx])
    d["n"] = len(idx)
    d["np"] = np_new
    d["nv"] = len(idx) - np_new
    return d


def step_both(c):
    sim = armando_amd.Simulation(c)
    sim.link_list()
    sim.single_step()
    g = sim.rates()
    s = sim.state()
    sim.close()
    P = orc.params_from(c)
    st = orc.OracleState(c["n"], c["np"])
    st.set(**{k: c[k] for k in FIELDS_IN})
    orc.single_step_gather(P, st)
    return g, s, st


def check_rates(g, st, tol=1e-12):
    for k, ok in (("drho", "drho"), ("ax", "ax"), ("ay", "ay"), ("du", "du"), ("avx", "avx"),
                  ("avy", "avy")):
        ref = st.field(ok)
        assert np.all(np.isfinite(g[k])), k
        scale = max(float(np.max(np.abs(ref))), 1e-300)
        assert float(np.max(np.abs(g[k] - ref))) <= tol * scale, k


def test_single_particle():
    c = subset(cases.box_lattice(nx=4, ny=4), np.array([5]), 1)
    c["flags"] = c["flags"] | cases.GRAVITY
    c["gy"] = -9.81
    banner("single particle")
    g, s, st = step_both(c)
    check_rates(g, st)
    assert g["drho"][0] == 0.0 and g["ax"][0] == 0.0 and g["ay"][0] == -9.81
    assert np.isfinite(s["p"]).all()


def test_walls_only_and_no_walls():
    dam = cases.dam_break_small(scale=0.05)
    nf = dam["np"]
    banner(f"walls only n={dam['nv']}")
    walls = subset(dam, np.arange(nf, dam["n"]), 0)
    g, s, st = step_both(walls)
    check_rates(g, st)
    sim = armando_amd.Simulation(walls)
    x0 = walls["x"].copy()
    sim.time_integration(1, 5)
    assert np.array_equal(sim.state()["x"], x0)  # no real particle, nothing moves
    sim.close()
    banner(f"fluid only n={nf}")
    fluid = subset(dam, np.arange(nf), nf)
    g, s, st = step_both(fluid)
    check_rates(g, st)


@pytest.mark.parametrize("prec", [cases.FP64, cases.FP32])
def test_pairs_at_exactly_the_support(prec):
    """Pairs at r = 2h exactly (r^2 = (2h)^2 in FP64) are not neighbours;
    one ulp inside they are."""
    h = 0.25
    c = cases.box_lattice(nx=3, ny=3)
    x = np.array([0.0, 0.5, 0.0, np.nextafter(0.5, 0.0), 2.0, 2.0])
    y = np.array([0.0, 0.0, 1.0, 1.0, 0.0, 0.5])
    c.update(x=x, y=y, vx=np.zeros(6), vy=np.zeros(6), rho=np.full(6, 1000.0),
             u=np.ones(6), n=6, np=6, nv=0, h=h, cell=2 * h, precision=prec)
    ox, oy, nxg, nyg = cases.grid_for(x, y, 2 * h)
    c.update(ox=ox, oy=oy, nx=nxg, ny=nyg)
    sim = armando_amd.Simulation(c)
    sim.link_list()
    gi, gj = sim.pairs()
    sim.close()
    P = orc.params_from(c)
    if prec == cases.FP64:
        oi, oj = orc.pairs_f64(P, x, y)
        assert (0, 1) not in set(zip(oi.tolist(), oj.tolist()))
        assert (2, 3) in set(zip(oi.tolist(), oj.tolist()))
        assert (4, 5) not in set(zip(oi.tolist(), oj.tolist()))
    else:
        cx, cy, fx, fy = orc.split_f32(P, x, y)
        oi, oj = orc.pairs_f32(P, cx, cy, fx, fy)
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)


def test_positions_on_cell_faces():
    """Particles exactly on cell faces hash like the oracle (floor of the
    FP64 quotient) and give the same pairs and rates."""
    base = cases.box_lattice(nx=12, ny=10, jitter=0.0)
    cell = base["cell"]
    rng = np.random.default_rng(3)
    k = rng.integers(0, 6, size=(base["n"], 2))
    x = base["ox"] + (k[:, 0] + 2) * cell + np.where(rng.random(base["n"]) < 0.5, 0.0,
                                                    0.37 * cell)
    y = base["oy"] + (k[:, 1] + 2) * cell + np.where(rng.random(base["n"]) < 0.5, 0.0,
                                                    0.61 * cell)
    # distinct positions
    _, first = np.unique(np.stack([x, y], 1), axis=0, return_index=True)
    idx = np.sort(first)
    c = subset(dict(base, x=x, y=y), idx, len(idx))
    banner(f"cell faces n={c['n']}")
    sim = armando_amd.Simulation(c)
    sim.link_list()
    key, sid, cs = sim.cells()
    gi, gj = sim.pairs()
    sim.close()
    P = orc.params_from(c)
    okey, bad = orc.keys(P, c["x"], c["y"])
    assert bad == 0
    osid, ocs = orc.sort_cells(okey, c["nx"] * c["ny"])
    assert np.array_equal(sid, osid) and np.array_equal(cs, ocs)
    oi, oj = orc.pairs_f64(P, c["x"], c["y"])
    assert np.array_equal(gi, oi) and np.array_equal(gj, oj)
    g, s, st = step_both(c)
    check_rates(g, st)


def test_coincident_particles():
    """Two particles at one position: r = 0 contributes nothing (grad W = 0,
    the regularised terms stay finite) and no NaN appears."""
    c = cases.box_lattice(nx=6, ny=5, vel=0.5)
    idx = np.concatenate([np.arange(c["n"]), [7]])
    d = subset(c, idx, len(idx))
    d["vx"] = d["vx"].copy()
    d["vx"][-1] = -0.3
    banner("coincident particles")
    g, s, st = step_both(d)
    check_rates(g, st)


def test_capacity_errors():
    c = cases.box_lattice(nx=10, ny=10)
    with pytest.raises(armando_amd.ArmandoError) as e:
        armando_amd.Simulation(c, capacity=c["n"] - 1)
    assert e.value.code == -3
    sim = armando_amd.Simulation(c)
    niac = sim.link_list(count=True)
    assert niac > 0
    with pytest.raises(armando_amd.ArmandoError) as e:
        sim.pairs(cap=niac - 1)
    assert e.value.code == -3
    gi, gj = sim.pairs(cap=niac)
    assert len(gi) == niac
    sim.close()
