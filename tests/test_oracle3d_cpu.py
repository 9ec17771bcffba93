"""CPU tests of the oracle's 3D step (oracle/armando_oracle3d.c).

  * oracle/_ref (the reference's CERNsource routines, compiled unmodified
    with the as-shipped dim = 3) with real z against the oracle's 3D gather:
    pair sets of link_list.f and the per-routine outputs of kernel.f,
    density.f, internal_force.f, art_visc.f, art_heat.f and av_vel.f;
  * brute force for the pair sets, the float32 restatement of the FP32
    criterion, the definition of the 3D cell keys;
  * closed forms and invariants: Wendland C2 3D normalisation, momentum
    conservation of the step, free fall under gravity along z.
"""
import numpy as np
import pytest

from armando_amd import cases
from oracle import orc

from . import refbind

need_ref = pytest.mark.skipif(not refbind.available(), reason="oracle/_ref not built here")


def lattice(seed=3, vel=0.4, **kw):
    return cases.box_lattice3(seed=seed, vel=vel, **kw)


def brute_pairs(x, y, z, h):
    n = len(x)
    d2 = (x[:, None] - x[None, :]) ** 2 + (y[:, None] - y[None, :]) ** 2 + \
        (z[:, None] - z[None, :]) ** 2
    i, j = np.nonzero(np.triu(d2 < (2 * h) ** 2, 1))
    return i.astype(np.int32), j.astype(np.int32)


def pair_set(i, j):
    return set(zip(i.tolist(), j.tolist()))


# ---------------------------------------------------------------- reference
@need_ref
def test_link_list3_equals_reference_fortran():
    c = lattice()
    P = orc.params_from(c)
    ri, rj = refbind.link_list(c["x"], c["y"], c["h"], z=c["z"])
    oi, oj = orc.pairs3_f64(P, c["x"], c["y"], c["z"])
    assert len(ri) > 10 * c["n"]
    lo = np.minimum(ri, rj)
    hi = np.maximum(ri, rj)
    assert pair_set(lo, hi) == pair_set(oi, oj)
    assert len(lo) == len(oi)


@need_ref
@pytest.mark.parametrize("seed", [3, 11])
def test_routines3_equal_reference_fortran(seed):
    """sum_density, con_density, int_force (pa_sph 0 and 1), art_visc,
    art_heat and av_vel of the reference with dim = 3 and real z against the
    oracle's 3D gather form with kernel.f's cubic spline."""
    c = lattice(seed=seed)
    n = c["n"]
    rng = np.random.default_rng(seed)
    rho = 1000.0 * (1 + 0.02 * rng.standard_normal(n))
    p = 1e3 * rng.standard_normal(n)
    u = 1.0 + rng.random(n)
    cs = np.full(n, 5.0)
    P = orc.params_from(c)
    P.kernel = orc.K_CUBIC3D
    P.av_alpha, P.av_beta, P.av_eta = 1.0, 1.0, 0.1   # art_visc.f:38-44
    P.ah_g1, P.ah_g2, P.ah_eta = 0.1, 1.0, 0.1         # art_heat.f:38-40
    P.xsph_eps = 0.3                                   # av_vel.f:35
    pi, pj = orc.pairs3_f64(P, c["x"], c["y"], c["z"])
    ref = refbind.pair_routines(c["x"], c["y"], c["vx"], c["vy"], rho, p, u, cs, c["h"],
                                c["mass"], pi, pj, z=c["z"], vz=c["vz"])
    got = {}
    for pa in (0, 1):
        P.pa_sph = pa
        o = orc.routines3(P, c["x"], c["y"], c["z"], c["vx"], c["vy"], c["vz"], rho, p, u, cs)
        got[f"if{pa}_ax"], got[f"if{pa}_ay"], got[f"if{pa}_az"] = o["if_a"].T
        got[f"if{pa}_du"] = o["if_du"]
    got["sum_rho"], got["drho"] = o["rho_sum"], o["drho"]
    got["av_ax"], got["av_ay"], got["av_az"] = o["av_a"].T
    got["av_du"], got["ah_du"] = o["av_du"], o["ah_du"]
    got["xs_x"], got["xs_y"], got["xs_z"] = o["xs"].T
    assert set(ref) == set(got)
    for k, r in ref.items():
        scale = np.max(np.abs(r))
        assert scale > 0, k
        err = np.max(np.abs(got[k] - r)) / scale
        # the gather adds the same terms in another order than the pair loop
        assert err < 1e-12, (k, err)


# ---------------------------------------------------------------- definitions
def test_pairs3_match_brute_force_and_keys_by_definition():
    c = lattice(nx=8, ny=7, nz=6, seed=5)
    P = orc.params_from(c)
    oi, oj = orc.pairs3_f64(P, c["x"], c["y"], c["z"])
    bi, bj = brute_pairs(c["x"], c["y"], c["z"], c["h"])
    assert np.array_equal(oi, bi) and np.array_equal(oj, bj)
    key, bad = orc.keys3(P, c["x"], c["y"], c["z"])
    assert bad == 0
    ix = np.floor((c["x"] - c["ox"]) / c["cell"]).astype(int)
    iy = np.floor((c["y"] - c["oy"]) / c["cell"]).astype(int)
    iz = np.floor((c["z"] - c["oz"]) / c["cell"]).astype(int)
    assert np.array_equal(key, ix + c["nx"] * (iy + c["ny"] * iz))
    # outside the grid: counted, key -1
    x = c["x"].copy()
    x[0] = c["ox"] - 1.0
    key, bad = orc.keys3(P, x, c["y"], c["z"])
    assert bad == 1 and key[0] == -1


def test_fp32_criterion3_matches_float32_restatement():
    c = lattice(nx=9, ny=8, nz=7, seed=9, jitter=0.45)
    P = orc.params_from(c)
    cells, offs = orc.split3_f32(P, c["x"], c["y"], c["z"])
    oi, oj = orc.pairs3_f32(P, cells, offs)
    f = np.float32
    cellf = f(c["cell"])
    cut2 = f((2.0 * c["h"]) ** 2)
    d = []
    for a in range(3):
        t = offs[a][:, None] - offs[a][None, :]
        u = (cells[a][:, None] - cells[a][None, :]).astype(f) * cellf
        d.append((t + u).astype(f))
    r2 = ((d[0] * d[0] + d[1] * d[1]).astype(f) + d[2] * d[2]).astype(f)
    bi, bj = np.nonzero(np.triu(r2 < cut2, 1))
    assert np.array_equal(oi, bi) and np.array_equal(oj, bj)
    # the split reconstructs the FP64 positions to FP32 rounding
    for a, (key, o) in enumerate((("x", "ox"), ("y", "oy"), ("z", "oz"))):
        back = c[o] + cells[a] * c["cell"] + offs[a].astype(np.float64)
        assert np.max(np.abs(back - c[key])) <= 2e-7 * c["cell"]


def test_wendland3d_normalisation():
    """W(0) = 21/(16 pi h^3), W = 0 from 2h on, and int W dV = 1 by the
    radial integral 4 pi int_0^2h W r^2 dr (Simpson, 4001 points).  W(r) is
    read off the summation density of a unit-mass particle with one
    neighbour at distance r (self term W(0) plus W(r))."""
    h = 0.013
    c = lattice(nx=2, ny=1, nz=1, dx=0.01, jitter=0.0)
    P = orc.params_from(c)
    P.kernel = orc.K_WENDLAND3D
    P.mass, P.h = 1.0, h
    P.ox, P.oy, P.oz, P.cell, P.nx, P.ny, P.nz = -0.1, -0.1, -0.1, 2 * h, 64, 64, 64
    z0 = np.zeros(2)
    one = np.ones(2)

    def W(r):
        o = orc.routines3(P, np.array([0.0, r]), z0, z0, z0, z0, z0, one, z0, z0, z0)
        return o["rho_sum"][0]

    W0 = 21.0 / (16.0 * np.pi * h ** 3)
    assert abs(W(0.0) - 2 * W0) < 1e-12 * W0
    assert W(2 * h) == W(2.5 * h) and abs(W(2 * h) - W0) < 1e-12 * W0
    r = np.linspace(0.0, 2 * h, 4001)
    Wr = np.array([W(rr) for rr in r]) - W0
    from scipy.integrate import simpson
    total = 4 * np.pi * simpson(Wr * r * r, x=r)
    assert abs(total - 1.0) < 1e-9


# ---------------------------------------------------------------- step
def state_from(c):
    st = orc.OracleState3(c["n"], c["np"])
    st.set(**{k: c[k] for k in ("x", "y", "z", "vx", "vy", "vz", "rho", "u")})
    return st


def test_step3_conserves_momentum_without_gravity():
    """Pressure, artificial viscosity and heat are pairwise antisymmetric:
    sum m a = 0 to round-off, sum m du (heat part) = 0; XSPH sums to zero."""
    c = lattice(nx=9, ny=8, nz=7, seed=13, vel=0.5)
    c["flags"] = cases.ART_VISC | cases.ART_HEAT | cases.AV_VEL
    P = orc.params_from(c)
    P.kernel = orc.K_WENDLAND3D
    st = state_from(c)
    orc.single_step3(P, st)
    a = np.stack([st.field("ax"), st.field("ay"), st.field("az")])
    scale = np.max(np.abs(a))
    assert scale > 0
    assert np.max(np.abs(a.sum(1))) < 1e-11 * scale * c["n"]
    xs = np.stack([st.field("avx"), st.field("avy"), st.field("avz")])
    assert np.max(np.abs(xs.sum(1))) < 1e-11 * np.max(np.abs(xs)) * c["n"]


def test_step3_gather_equals_routines():
    """The step's gather (continuity density, Tait EOS, forces) against the
    per-routine gather outputs with the same p and c."""
    c = lattice(nx=8, ny=8, nz=8, seed=21, vel=0.3)
    c["flags"] = cases.ART_VISC | cases.ART_HEAT | cases.AV_VEL
    P = orc.params_from(c)
    P.kernel = orc.K_WENDLAND3D
    P.pmin = -1e30
    st = state_from(c)
    orc.single_step3(P, st)
    rho, p, cs = st.field("rho").copy(), st.field("p").copy(), st.field("c").copy()
    o = orc.routines3(P, c["x"], c["y"], c["z"], c["vx"], c["vy"], c["vz"], rho, p, c["u"], cs)
    B = P.rho0 * P.c0 ** 2 / 7
    assert np.allclose(p, B * ((rho / P.rho0) ** 7 - 1), rtol=1e-12, atol=1e-9)
    ref_a = o["if_a"] + o["av_a"]
    got_a = np.stack([st.field("ax"), st.field("ay"), st.field("az")], 1)
    assert np.max(np.abs(got_a - ref_a)) < 1e-12 * np.max(np.abs(ref_a))
    ref_du = o["if_du"] + o["av_du"] + o["ah_du"]
    assert np.max(np.abs(st.field("du") - ref_du)) < 1e-12 * np.max(np.abs(ref_du))
    assert np.max(np.abs(st.field("drho") - o["drho"])) < 1e-12 * np.max(np.abs(o["drho"]))
    got_xs = np.stack([st.field("avx"), st.field("avy"), st.field("avz")], 1)
    assert np.max(np.abs(got_xs - o["xs"])) < 1e-12 * np.max(np.abs(o["xs"]))


def test_free_fall3_closed_form():
    """Isolated particles (no neighbours) under gravity along z: the
    leapfrog gives z(t) = z0 + v0 t - g t^2/2 exactly up to round-off."""
    c = lattice(nx=3, ny=3, nz=3, dx=0.1, jitter=0.0, vel=0.0, h_fac=0.3)
    c["flags"] = cases.GRAVITY
    c["gz"] = -9.81
    c["dt"] = 1e-3
    P = orc.params_from(c)
    P.kernel = orc.K_WENDLAND3D
    P.pmin = -1e30
    st = state_from(c)
    z0 = c["z"].copy()
    nsteps = 200
    orc.run3(P, st, 1, nsteps)
    t = nsteps * c["dt"]
    assert np.allclose(st.field("z"), z0 - 0.5 * 9.81 * t * t, rtol=0, atol=1e-12)
    assert np.allclose(st.field("vz"), -9.81 * (t - 0.5 * c["dt"]), rtol=1e-12)
    assert np.allclose(st.field("x"), c["x"], rtol=0, atol=0)


def test_dam_break3d_layout():
    c = cases.dam_break3d_small(scale=0.1)
    assert c["np"] == 8 * 8 * 17
    P = orc.params_from(c)
    key, bad = orc.keys3(P, c["x"], c["y"], c["z"])
    assert bad == 0
    # walls enclose the fluid: no wall particle inside the tank
    nf = c["np"]
    w = slice(nf, None)
    inside = (c["x"][w] > 0) & (c["x"][w] < 0.32) & (c["y"][w] > 0) & (c["y"][w] < 0.1) & \
        (c["z"][w] > 0)
    assert not inside.any()
