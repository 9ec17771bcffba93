"""CPU tests of the mirror ghosts (ghost.f) in the oracle.

Pinned against the reference's own ghost_particles (CERNsource/ghost.f,
compiled as shipped) on the same pair list, and the step with ghosts against
the reference's pair routines (density.f, internal_force.f, art_visc.f,
art_heat.f, av_vel.f) on the particle set extended by the ghosts, with the
ghosts' rho, p, u taken from their parents after the EOS
(single_step.f:188-196).  The mirror geometry is checked on its own, and the
gather twin against the faithful half-list step over several steps.
"""
import numpy as np
import pytest

from armando_amd import cases
from oracle import orc

from . import refbind

need_ref = pytest.mark.skipif(not refbind.available(), reason="oracle/_ref not built here")

FIELDS = ("x", "y", "vx", "vy", "rho", "u")


def tank(seed=12345, **kw):
    c = cases.ghost_tank(col_w=0.02, col_h=0.03, tank_w=0.05, seed=seed, **kw)
    return c


def ghost_state(c):
    st = orc.GhostState(c["np"], c["nv"])
    st.set(**{k: c[k] for k in FIELDS})
    st.set(c=np.full(c["n"], c["c0"]), p=np.zeros(c["n"]))
    return st


def moved(c, seed=4, amp=0.3):
    """Real particles displaced by up to amp dx and given velocities; returns
    the displaced case and the previous positions."""
    rng = np.random.default_rng(seed)
    d = c["h"] / 1.3
    c2 = dict(c)
    nf = c["np"]
    x, y = c["x"].copy(), c["y"].copy()
    x[:nf] += amp * d * (2 * rng.random(nf) - 1)
    y[:nf] += amp * d * (2 * rng.random(nf) - 1)
    vx, vy = c["vx"].copy(), c["vy"].copy()
    vx[:nf] = 0.5 * (2 * rng.random(nf) - 1)
    vy[:nf] = 0.5 * (2 * rng.random(nf) - 1)
    c2.update(x=x, y=y, vx=vx, vy=vy)
    return c2, c["x"], c["y"]


def brute_pairs(x, y, h):
    d2 = (x[:, None] - x[None, :]) ** 2 + (y[:, None] - y[None, :]) ** 2
    i, j = np.nonzero(np.triu(d2 < (2 * h) ** 2, 1))
    return i.astype(np.int32), j.astype(np.int32)


@need_ref
@pytest.mark.parametrize("seed", [12345, 99])
def test_ghost_particles_equal_reference_fortran(seed):
    c0 = tank(seed)
    c, xp, yp = moved(c0, seed)
    P = orc.params_from(c)
    st = ghost_state(c)
    st.set_prev(xp, yp)
    ng = orc.ghost_particles(P, st)
    assert ng > 50
    # the reference walks the pair list of the last search (positions xp)
    pi, pj = brute_pairs(xp, yp, c["h"])
    n, nf = c["n"], c["np"]
    ref = refbind.ghost_particles(c["x"], c["y"], c["vx"], c["vy"], c["h"], c["mass"], c["rho"],
                                  np.zeros(n), nf, c["nv"], pi, pj)
    assert ref["ng"] == ng
    par = st.parents()
    want = np.full(nf, -1)
    want[par] = np.arange(ng)
    assert np.array_equal(ref["index"] - n, np.where(want >= 0, want, -1 - n))
    for k in ("x", "y", "vx", "vy", "rho"):
        got = st.field(k)[n:]
        assert np.max(np.abs(got - ref[k][n:])) <= 1e-15 * np.max(np.abs(ref[k][n:])), k


def test_mirror_geometry():
    c, xp, yp = moved(tank(), 7)
    P = orc.params_from(c)
    st = ghost_state(c)
    st.set_prev(xp, yp)
    ng = orc.ghost_particles(P, st)
    n = c["n"]
    par = st.parents()
    x, y, vx, vy = (st.field(k) for k in ("x", "y", "vx", "vy"))
    gx, gy = x[n:], y[n:]
    px, py = x[par], y[par]
    # the nearest virtual particle among the candidates of xp
    wx, wy = c["x"][c["np"]:], c["y"][c["np"]:]
    nx_, ny_ = c["vx"][c["np"]:], c["vy"][c["np"]:]
    cand = ((xp[par, None] - wx[None, :]) ** 2 + (yp[par, None] - wy[None, :]) ** 2
            < (2 * c["h"]) ** 2)
    d2 = (wx[None, :] - px[:, None]) ** 2 + (wy[None, :] - py[:, None]) ** 2
    j = np.argmin(np.where(cand, d2, np.inf), axis=1)
    nxj, nyj = nx_[j], ny_[j]
    # the midpoint lies on the wall line, the offset is along the normal
    mid = ((gx + px) / 2 - wx[j]) * nxj + ((gy + py) / 2 - wy[j]) * nyj
    assert np.max(np.abs(mid)) < 1e-15
    assert np.max(np.abs((gx - px) * nyj - (gy - py) * nxj)) < 1e-15
    # tangential velocity kept, normal velocity reversed
    vn_p = vx[par] * nxj + vy[par] * nyj
    vn_g = vx[n:] * nxj + vy[n:] * nyj
    assert np.allclose(vn_g, -vn_p, rtol=0, atol=1e-15)
    assert np.allclose(vx[n:] * nyj - vy[n:] * nxj, vx[par] * nyj - vy[par] * nxj, rtol=0,
                       atol=1e-15)
    # real particles without candidates have no ghost
    allc = ((xp[:c["np"], None] - wx[None, :]) ** 2 + (yp[:c["np"], None] - wy[None, :]) ** 2
            < (2 * c["h"]) ** 2).any(axis=1)
    assert set(par.tolist()) == set(np.nonzero(allc)[0].tolist())
    assert ng == int(allc.sum())


def test_candidates_come_from_the_previous_search():
    """A particle that reached a wall since the last search has no ghost yet
    (time_integration.f:157-161 runs ghost_particles before link_list), and
    one that left it still has one."""
    c = tank()
    P = orc.params_from(c)
    h = c["h"]
    nf = c["np"]
    y0 = c["y"].copy()
    # particle a: previous position far from the bottom, now close to it
    a = int(np.argmin(np.abs(c["x"][:nf] - 0.01) + np.abs(c["y"][:nf] - 0.015)))
    xp, yp = c["x"].copy(), c["y"].copy()
    yp[a] = 0.02
    c2 = dict(c)
    y = y0.copy()
    y[a] = 0.5 * h
    c2["y"] = y
    st = ghost_state(c2)
    st.set_prev(xp, yp)
    orc.ghost_particles(P, st)
    assert a not in set(st.parents().tolist())
    # the next step sees it
    orc.ghost_particles(P, st)
    assert a in set(st.parents().tolist())


@need_ref
def test_ghost_step_equals_reference_routines():
    c, xp, yp = moved(tank(), 11, amp=0.1)
    P = orc.params_from(c)
    st = ghost_state(c)
    st.set_prev(xp, yp)
    orc.ghost_particles(P, st)
    orc.single_step_gather(P, st)
    n, ntot = c["n"], st.n
    par = st.parents()
    for k in ("rho", "p", "u"):
        f = st.field(k)
        assert np.array_equal(f[n:], f[par]), k
    x, y = st.field("x").copy(), st.field("y").copy()
    pi, pj = brute_pairs(x, y, c["h"])
    # the reference build has dim = 3 and the cubic spline, and art_visc.f
    # and av_vel.f hard-code alpha = beta = 1 and eps = 0.3
    P3 = orc.params_from(c)
    P3.kernel = orc.K_CUBIC3D
    P3.av_alpha, P3.av_beta, P3.xsph_eps = 1.0, 1.0, 0.3
    st3 = ghost_state(c)
    st3.set_prev(xp, yp)
    orc.ghost_particles(P3, st3)
    orc.single_step_gather(P3, st3)
    ref = refbind.pair_routines(*(st3.field(k).copy() for k in ("x", "y", "vx", "vy", "rho", "p",
                                                                "u")),
                                np.full(ntot, c["c0"]), c["h"], c["mass"], pi, pj)
    nf = c["np"]
    got = {"drho": st3.field("drho"), "ax": st3.field("ax") - P3.gx,
           "ay": st3.field("ay") - P3.gy, "du": st3.field("du"),
           "xs_x": st3.field("avx"), "xs_y": st3.field("avy")}
    want = {"drho": ref["drho"], "ax": ref["if0_ax"] + ref["av_ax"],
            "ay": ref["if0_ay"] + ref["av_ay"], "du": ref["if0_du"] + ref["av_du"] + ref["ah_du"],
            "xs_x": ref["xs_x"], "xs_y": ref["xs_y"]}
    for k in want:
        scale = np.max(np.abs(want[k][:nf]))
        assert scale > 0, k
        err = np.max(np.abs(got[k][:nf] - want[k][:nf])) / scale
        assert err < 1e-12, (k, err)


def test_gather_equals_faithful_with_ghosts():
    c = tank()
    P = orc.params_from(c)
    a = ghost_state(c)
    b = ghost_state(c)
    orc.run(P, a, 1, 8, gather=True)
    orc.run(P, b, 1, 8, gather=False)
    assert a.n == b.n and a.n > c["n"]
    for k in ("x", "y", "vx", "vy", "rho", "u", "p"):
        fa, fb = a.field(k), b.field(k)
        scale = max(np.max(np.abs(fb)), 1e-300)
        assert np.max(np.abs(fa - fb)) < 1e-10 * scale, k
