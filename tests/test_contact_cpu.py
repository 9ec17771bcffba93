"""The contact-displacement wall model of the other Fortran tree
(FORTRANsource/external_force.f90:191-339, applied in
FORTRANsource/time_integration.f90:128-142): the oracle's restatement
against that tree's own routines, compiled unmodified into
oracle/_ref/libarmref90.so (its options.inc has dim = 3; the tests run it
with z = 0), on the pair list of the reference's own link_list; closed forms
of the displacement at a flat wall; and the oracle's faithful step against
its gather twin with the model on.
"""
import numpy as np
import pytest

from armando_amd import cases
from oracle import orc
from tests import refbind

needs_ref90 = pytest.mark.skipif(not refbind.available90(),
                                 reason="oracle/_ref/libarmref90.so not built")
needs_ref = pytest.mark.skipif(not refbind.available(),
                               reason="oracle/_ref/libarmref.so not built")


def tank():
    return cases.contact_tank()


def oracle_params(c, kernel=orc.K_WENDLAND2D):
    P = orc.params_from(c)
    P.kernel = kernel
    P.contact_hdr = 1.2
    return P


def contact(P, c, pi, pj):
    dx = np.zeros(c["n"])
    dy = np.zeros(c["n"])
    orc.lib().orc_contact_displacement(P, c["np"], c["nv"], c["x"], c["y"], c["vx"], c["vy"],
                                       len(pi), pi, pj, dx, dy)
    return dx, dy


@needs_ref90
def test_f90_cubic_kernel_is_the_oracles():
    """kernel.f90 (skf 0, dim 3) equals the oracle's cubic spline in 3D,
    the kernel the smoothing is pinned with below."""
    h = 1.3e-3
    P = orc.params(kernel=orc.K_CUBIC3D, h=h, mass=1.0)
    import ctypes

    F = ctypes.c_double(0.0)
    for r in np.linspace(0.0, 2.2 * h, 97):
        assert orc.lib().orc_kernel(ctypes.byref(P), float(r), ctypes.byref(F)) == \
            refbind.kernel90(float(r), h), r


@needs_ref
@needs_ref90
def test_contact_displacement_matches_reference():
    c = tank()
    P = oracle_params(c)
    pi, pj = refbind.link_list(c["x"], c["y"], c["h"])
    # the tank's walls push: some particles lie inside the contact distance
    dx, dy = contact(P, c, pi, pj)
    assert np.count_nonzero(dx[:c["np"]]) > 10 and np.count_nonzero(dy[:c["np"]]) > 10
    rx, ry = refbind.contact_displacement(c["np"], c["nv"], c["x"], c["y"], c["vx"], c["vy"],
                                          c["h"], pi, pj)
    assert np.array_equal(dx, rx) and np.array_equal(dy, ry)
    # from a non-zero start (the routine accumulates into d)
    rng = np.random.default_rng(3)
    d0x, d0y = 1e-4 * rng.normal(size=c["n"]), 1e-4 * rng.normal(size=c["n"])
    ox, oy = d0x.copy(), d0y.copy()
    orc.lib().orc_contact_displacement(P, c["np"], c["nv"], c["x"], c["y"], c["vx"], c["vy"],
                                       len(pi), pi, pj, ox, oy)
    rx, ry = refbind.contact_displacement(c["np"], c["nv"], c["x"], c["y"], c["vx"], c["vy"],
                                          c["h"], pi, pj, d0x, d0y)
    assert np.array_equal(ox, rx) and np.array_equal(oy, ry)


@needs_ref
@needs_ref90
def test_smooth_vector_matches_reference():
    c = tank()
    P = oracle_params(c, kernel=orc.K_CUBIC3D)
    pi, pj = refbind.link_list(c["x"], c["y"], c["h"])
    rng = np.random.default_rng(5)
    n, npf = c["n"], c["np"]
    vx, vy = rng.normal(size=n), rng.normal(size=n)
    rho = 1000.0 * (1.0 + 0.01 * rng.uniform(-1, 1, n))
    ox, oy = vx.copy(), vy.copy()
    orc.lib().orc_smooth_vector(P, npf, c["x"], c["y"], rho, len(pi), pi, pj, ox, oy)
    rx, ry = refbind.smooth_vector(npf, c["x"], c["y"], c["h"], c["mass"], rho, vx, vy, pi, pj)
    np.testing.assert_allclose(ox[:npf], rx[:npf], rtol=1e-14, atol=1e-14 * np.abs(rx).max())
    np.testing.assert_allclose(oy[:npf], ry[:npf], rtol=1e-14, atol=1e-14 * np.abs(ry).max())
    # the virtual particles' entries are left alone
    assert np.array_equal(ox[npf:], vx[npf:])


def test_displacement_at_a_flat_wall_closed_form():
    """A particle at distance s from a row of wall particles with normal n
    moves to the contact distance: d = (h/hdr - s) n when s < h/hdr, else 0;
    a second, already displaced pass adds nothing."""
    h, hdr, dxw = 1.3e-3, 1.2, 0.5e-3
    wx = np.arange(-10, 11) * dxw
    for s in (0.2e-3, 0.7e-3, 1.0e-3, 1.2e-3):
        x = np.concatenate([[0.13e-3], wx])
        y = np.concatenate([[s], np.zeros(wx.size)])
        n = x.size
        vx = np.zeros(n)
        vy = np.concatenate([[0.0], np.ones(wx.size)])
        c = dict(n=n, np=1, nv=n - 1, x=x, y=y, vx=vx, vy=vy, h=h)
        P = orc.params(kernel=orc.K_WENDLAND2D, h=h, mass=1.0, contact_hdr=hdr)
        pi = np.zeros(n - 1, np.int32)
        pj = np.arange(1, n, dtype=np.int32)
        keep = np.hypot(x[0] - x[1:], y[0] - y[1:]) < 2 * h
        pi, pj = pi[keep], pj[keep]
        dx, dy = contact(P, c, pi, pj)
        want = max(h / hdr - s, 0.0)
        assert dx[0] == 0.0
        assert dy[0] == pytest.approx(want, rel=1e-12, abs=1e-18)
        dx2, dy2 = dx.copy(), dy.copy()
        orc.lib().orc_contact_displacement(P, 1, n - 1, x, y, vx, vy, len(pi), pi, pj, dx2, dy2)
        assert dy2[0] == dy[0]


def test_faithful_step_equals_gather_twin_with_contact():
    c = tank()
    P = orc.params_from(c)
    P.contact_hdr = 1.2
    a = orc.OracleState(c["n"], c["np"])
    b = orc.OracleState(c["n"], c["np"])
    for st in (a, b):
        st.set(**{k: c[k] for k in ("x", "y", "vx", "vy", "rho", "u")})
        st.set(c=np.full(c["n"], c["c0"]), p=np.zeros(c["n"]))
    pi, pj = orc.link_list(P, c["x"], c["y"])
    rc = orc.lib().orc_single_step(P, a.ptr, len(pi), pi, pj)
    assert rc == 0
    orc.single_step_gather(P, b)
    P2 = orc.params_from(c)
    P2.flags = c["flags"] & ~orc.CONTACT
    base = orc.OracleState(c["n"], c["np"])
    base.set(**{k: c[k] for k in ("x", "y", "vx", "vy", "rho", "u")})
    orc.single_step_gather(P2, base)
    push = np.abs(b.field("ay") - base.field("ay"))[:c["np"]]
    assert push.max() > 10.0 * np.abs(base.field("ay")).max()  # the walls do push
    for k in ("ax", "ay"):
        ra, rb = a.field(k)[:c["np"]], b.field(k)[:c["np"]]
        np.testing.assert_allclose(ra, rb, rtol=0, atol=1e-12 * np.abs(rb).max())
