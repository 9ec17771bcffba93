"""Closed-form pins of the oracle's leapfrog legs and conservation sums.

The integrator of CERNsource/time_integration.f:136-217 (orc_predictor,
orc_corrector) is driven here with prescribed constant rates on isolated
particles, so that every leg has a closed form:
  u_n   = u_0 + (n - 1/2) dt du                     (u >= 0 throughout)
  rho_n = rho_0 + (n - 1/2) dt drho                 (density_method 0 only)
  v_n   = v_0 + (n - 1/2) dt a + (n - 1) av         (av enters from step 2 on)
  x_n   = x_0 + dt (n v_0 + dt a n^2 / 2 + av n (n - 1) / 2)
Step 1 is a half step without av (time_integration.f:203-217); every later
step a predictor half step from the saved values followed by a full step
from them.  The u < 0 -> 0 clamp acts in both halves.  Walls (particles
np .. n-1) stay frozen unless DBC is set, in which case they integrate rho
and u and keep v and x (ArmandoList/armando2Dcpu.cu:879-882).

orc_conservation follows the budget of FORTRANsource/time_integration.f90:
185-203 (sums over the real particles 1..np of m u, m v and m v.v / 2),
plus the mass; it is pinned by states whose sums are known in closed form.
"""
import numpy as np
import pytest

from oracle import orc

NP, NV = 4, 2
N = NP + NV
DT = 1e-3


def _params(flags=0, density_method=0):
    return orc.params(kernel=0, flags=flags, density_method=density_method, h=0.01, mass=0.25,
                      rho0=1000.0, c0=10.0, pmin=-1e30, dt=DT, ox=-10.0, oy=-10.0, cell=1.0,
                      nx=20, ny=20)


def _state(rng):
    st = orc.OracleState(N, NP)
    init = dict(
        x=rng.uniform(-1, 1, N), y=rng.uniform(-1, 1, N),
        vx=rng.uniform(-2, 2, N), vy=rng.uniform(-2, 2, N),
        rho=rng.uniform(990, 1010, N), u=rng.uniform(5, 10, N))
    st.set(**init)
    rates = dict(
        du=rng.uniform(-50, 50, N), drho=rng.uniform(-300, 300, N),
        ax=rng.uniform(-20, 20, N), ay=rng.uniform(-20, 20, N),
        avx=rng.uniform(-1e-3, 1e-3, N), avy=rng.uniform(-1e-3, 1e-3, N))
    return st, init, rates


def _drive(P, st, rates, nsteps):
    """The step loop of time_integration.f with the rates held fixed (the
    predictor does not touch them; they are set again before every
    corrector, where single_step would have written them)."""
    for its in range(1, nsteps + 1):
        if its != 1:
            orc.predictor(P, st)
        st.set(**rates)
        orc.corrector(P, st, its)


@pytest.mark.parametrize("nsteps", [1, 2, 7, 40])
@pytest.mark.parametrize("dbc", [False, True])
def test_leapfrog_prescribed_rates(nsteps, dbc):
    rng = np.random.default_rng(100 + nsteps)
    P = _params(orc.DBC if dbc else 0)
    st, s0, r = _state(rng)
    st.set(**r)
    _drive(P, st, r, nsteps)
    n = nsteps
    half = (n - 0.5) * DT
    got = st.get("x", "y", "vx", "vy", "rho", "u")
    fl = slice(0, NP)
    wl = slice(NP, N)
    # real particles: every leg
    np.testing.assert_allclose(got["u"][fl], s0["u"][fl] + half * r["du"][fl], rtol=1e-13)
    np.testing.assert_allclose(got["rho"][fl], s0["rho"][fl] + half * r["drho"][fl], rtol=1e-14)
    for c, a, av in (("x", "ax", "avx"), ("y", "ay", "avy")):
        v0 = s0["v" + c][fl]
        vn = v0 + half * r[a][fl] + (n - 1) * r[av][fl]
        np.testing.assert_allclose(got["v" + c][fl], vn, rtol=1e-12, atol=1e-13)
        xn = s0[c][fl] + DT * (n * v0 + DT * r[a][fl] * n * n / 2 + r[av][fl] * n * (n - 1) / 2)
        np.testing.assert_allclose(got[c][fl], xn, rtol=0, atol=1e-13)
    # walls: position and velocity held always; rho and u integrate only with DBC
    for k in ("x", "y", "vx", "vy"):
        assert np.array_equal(got[k][wl], s0[k][wl]), k
    if dbc:
        np.testing.assert_allclose(got["u"][wl], s0["u"][wl] + half * r["du"][wl], rtol=1e-13)
        np.testing.assert_allclose(got["rho"][wl], s0["rho"][wl] + half * r["drho"][wl],
                                   rtol=1e-14)
    else:
        assert np.array_equal(got["u"][wl], s0["u"][wl])
        assert np.array_equal(got["rho"][wl], s0["rho"][wl])


def test_leapfrog_first_step_takes_no_xsph():
    """Step 1 adds no av (the `! + av(d, i)` of time_integration.f:212);
    step 2 adds it once."""
    rng = np.random.default_rng(5)
    P = _params()
    st, s0, r = _state(rng)
    r["ax"][:] = 0.0
    r["ay"][:] = 0.0
    st.set(**r)
    orc.corrector(P, st, 1)
    assert np.array_equal(st.field("vx")[:NP], s0["vx"][:NP])
    assert np.array_equal(st.field("vy")[:NP], s0["vy"][:NP])
    orc.predictor(P, st)
    st.set(**r)
    orc.corrector(P, st, 2)
    np.testing.assert_allclose(st.field("vx")[:NP], s0["vx"][:NP] + r["avx"][:NP], rtol=1e-14)
    np.testing.assert_allclose(st.field("vy")[:NP], s0["vy"][:NP] + r["avy"][:NP], rtol=1e-14)


def test_summation_density_is_not_integrated():
    """density_method 1: rho is recomputed by sum_density, never integrated."""
    rng = np.random.default_rng(6)
    P = _params(orc.DBC, density_method=1)
    st, s0, r = _state(rng)
    st.set(**r)
    _drive(P, st, r, 5)
    assert np.array_equal(st.field("rho"), s0["rho"])


def test_energy_clamp_in_both_halves():
    """u < 0 -> 0 after the predictor's half step and after the corrector
    (time_integration.f:139-141, 189-190, 205-206), with u_min the unclamped
    value saved before the half step."""
    P = _params(orc.DBC)
    st = orc.OracleState(N, NP)
    du = -100.0
    q = DT * abs(du)  # one full step's decrease
    u0 = np.full(N, 2.3 * q)
    st.set(x=np.zeros(N), y=np.zeros(N), vx=np.zeros(N), vy=np.zeros(N),
           rho=np.full(N, 1000.0), u=u0)
    rates = dict(du=np.full(N, du), drho=np.zeros(N), ax=np.zeros(N), ay=np.zeros(N),
                 avx=np.zeros(N), avy=np.zeros(N))
    st.set(**rates)
    orc.corrector(P, st, 1)                       # 2.3 q - q/2
    np.testing.assert_allclose(st.field("u"), 1.8 * q, rtol=1e-14)
    orc.predictor(P, st)                          # 1.8 q - q/2 = 1.3 q (no clamp)
    np.testing.assert_allclose(st.field("u"), 1.3 * q, rtol=1e-14)
    np.testing.assert_allclose(st.field("u_min"), 1.8 * q, rtol=1e-14)
    st.set(**rates)
    orc.corrector(P, st, 2)                       # 1.8 q - q = 0.8 q
    np.testing.assert_allclose(st.field("u"), 0.8 * q, rtol=1e-13)
    orc.predictor(P, st)                          # 0.8 q - q/2 = 0.3 q
    np.testing.assert_allclose(st.field("u"), 0.3 * q, rtol=1e-12)
    st.set(**rates)
    orc.corrector(P, st, 3)                       # 0.8 q - q < 0 -> 0
    assert np.all(st.field("u") == 0.0)
    orc.predictor(P, st)                          # 0 - q/2 < 0 -> 0, u_min = 0
    assert np.all(st.field("u") == 0.0)
    assert np.all(st.field("u_min") == 0.0)
    st.set(**rates)
    orc.corrector(P, st, 4)
    assert np.all(st.field("u") == 0.0)
    # the first step's half step clamps too
    st2 = orc.OracleState(N, NP)
    st2.set(x=np.zeros(N), y=np.zeros(N), vx=np.zeros(N), vy=np.zeros(N),
            rho=np.full(N, 1000.0), u=np.full(N, 0.2 * q), **rates)
    orc.corrector(P, st2, 1)
    assert np.all(st2.field("u") == 0.0)


def test_conservation_sums_closed_form():
    """Sums over the real particles only; walls carry large values that a
    sum over every particle would pick up."""
    P = _params(orc.DBC)
    n, npf = 50, 37
    st = orc.OracleState(n, npf)
    vx = np.where(np.arange(n) < npf, 1.5, 1e6)
    vy = np.where(np.arange(n) < npf, -2.0, 1e6)
    u = np.where(np.arange(n) < npf, 3.0, 1e6)
    st.set(x=np.zeros(n), y=np.zeros(n), vx=vx, vy=vy, rho=np.full(n, 1000.0), u=u)
    m = P.mass
    out = orc.conservation(P, st)
    expect = [npf * m, npf * m * 1.5, -npf * m * 2.0, npf * m * 0.5 * (1.5 ** 2 + 2.0 ** 2),
              npf * m * 3.0]
    np.testing.assert_allclose(out, expect, rtol=1e-15)


def test_conservation_sums_of_a_random_state():
    """The same sums against exactly rounded ones (math.fsum) on a random state."""
    import math

    P = _params()
    rng = np.random.default_rng(11)
    n, npf = 400, 333
    st = orc.OracleState(n, npf)
    vx, vy, u = rng.normal(size=n), rng.normal(size=n), rng.uniform(0, 5, n)
    st.set(x=np.zeros(n), y=np.zeros(n), vx=vx, vy=vy, rho=np.full(n, 1000.0), u=u)
    m = P.mass
    out = orc.conservation(P, st)
    f = slice(0, npf)
    ke = math.fsum(0.5 * m * vx[f] ** 2) + math.fsum(0.5 * m * vy[f] ** 2)
    expect = [npf * m, m * math.fsum(vx[f]), m * math.fsum(vy[f]), ke, m * math.fsum(u[f])]
    np.testing.assert_allclose(out, expect, rtol=1e-13)


def test_conservation_of_momentum_and_energy_over_steps():
    """Over a run of the faithful step without gravity, walls or XSPH, the
    pressure and artificial-viscosity forces are pairwise antisymmetric, so
    the momentum sums stay at their initial values to round-off, and the
    total energy KE + IE drifts only at the integrator's O(dt^2) level."""
    from armando_amd import cases

    c = cases.box_lattice(nx=18, ny=14, vel=0.5)
    c["flags"] = orc.ART_VISC
    P = orc.params_from(c)
    st = orc.OracleState(c["n"], c["np"])
    st.set(**{k: c[k] for k in ("x", "y", "vx", "vy", "rho", "u")})
    b0 = orc.conservation(P, st)
    orc.run(P, st, 1, 20, gather=True)
    b1 = orc.conservation(P, st)
    scale = P.mass * c["np"] * 0.5
    assert abs(b1[1] - b0[1]) < 1e-13 * scale
    assert abs(b1[2] - b0[2]) < 1e-13 * scale
    assert b1[0] == b0[0]
    e0, e1 = b0[3] + b0[4], b1[3] + b1[4]
    assert abs(e1 - e0) < 1e-6 * e0
