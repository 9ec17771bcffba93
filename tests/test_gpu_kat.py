"""The reference's known-answer test run through the HIP gathers.

check() of ArmandoList/armando2Dlist.cu:2451-2548 on the lattice of
initFree (:1834-1891): 100 x 100 particles at dr = 0.01, h = 1.2 dr,
rho = 1000, with the manufactured fields
    p = 1000 (sin(10 pi x) + y^2),  v = (sin(10 pi x), y^2)
whose continuity and pressure-gradient rates have the closed forms
    drho/dt = -1000 (10 pi cos(10 pi x) + 2 y),
    dv/dt   = (-10 pi cos(10 pi x), -2 y).
Particles within one cell (2h) of the lattice edge take the closed form, and
the error is check()'s: 100 * mean |closed form - computed| / max |closed
form| per column.  The reference's own GPU result (old.txt) scores 2.868069 %,
8.042519 % and 1.729856 %.

Through the C ABI: arm_link_list, arm_density(0) (con_density,
density.f:166-223), arm_eos, arm_internal_force(0) (pa_sph 0 pressure term of
internal_force.f).  The ABI derives p from rho, so the state carries
rho = rho0 + p / c0^2 under the linear EOS p = c0^2 (rho - rho0) with
c0 = 1e4 m/s: rho stays within 2e-5 of rho0 = 1000, which moves the pressure
term by at most 4e-5 of itself.  Artificial viscosity, art_heat and XSPH are
off (check() adds the 0.03 shock correction and density diffusion of the C
path, which the CERN step does not have).  The rows are matched to the
fixture's by the lattice coordinates recovered from its closed-form columns.
"""
import os

import numpy as np
import pytest

import armando_amd
from armando_amd import cases
from tests.test_oracle_cpu import GOLDEN, recover_kat_order

pytestmark = pytest.mark.gpu

REF_ERR = np.array([2.868, 8.043, 1.730])  # check() on old.txt, percent


def banner(msg):
    print(f"[gpu-kat] {msg}", flush=True)


def kat_case(precision):
    fx = np.load(os.path.join(GOLDEN, "kat_check.npz"))
    ii, jj = recover_kat_order(fx["analytic"])
    x = ii * 0.01
    y = jj * 0.01
    n = len(x)
    c0, rho0 = 1.0e4, 1000.0
    p = 1000.0 * (np.sin(10 * np.pi * x) + y * y)
    h = 1.2 * 0.01
    cell = 2.0 * h
    nxg = int(2.0 / cell) + 1
    case = dict(
        name="kat", n=n, np=n, nv=0, x=x, y=y, vx=np.sin(10 * np.pi * x), vy=y * y,
        rho=rho0 + p / (c0 * c0), u=np.zeros(n), precision=precision, flags=0, pa_sph=0,
        density_method=0, h=h, mass=rho0 * 0.01 * 0.01, rho0=rho0, c0=c0, pmin=-1e30,
        av_alpha=0.0, av_beta=0.0, av_eta=0.1, ah_g1=0.0, ah_g2=0.0, ah_eta=0.1, xsph_eps=0.0,
        gx=0.0, gy=0.0, dt=1e-6, ox=-0.5, oy=-0.5, cell=cell, nx=nxg, ny=nxg,
        eos=armando_amd.EOS_LINEAR, eos_p0=-c0 * c0 * rho0)
    return case, fx["gpu"].astype(np.float64), fx["analytic"].astype(np.float64), p


@pytest.mark.parametrize("precision", [cases.FP64, cases.FP32])
def test_known_answer_test_through_the_hip_gathers(precision):
    case, ref_gpu, ana, p = kat_case(precision)
    banner(f"KAT lattice n={case['n']} precision={precision}")
    sim = armando_amd.Simulation(case)
    sim.link_list()
    sim.density(0)
    sim.eos()
    st = sim.state()
    assert np.max(np.abs(st["p"] - p)) < 1e-6 * 2000.0
    sim.internal_force(0)
    r = sim.rates()
    sim.close()
    mine = np.stack([r["drho"], r["ax"], r["ay"]], 1)
    x, y = case["x"], case["y"]
    size = case["cell"]
    edge = (x < size) | (y < size) | (x > 1.0 - size) | (y > 1.0 - size)
    mine[edge] = ana[edge]
    scale = np.abs(ana).max(axis=0)
    err = 100 * np.abs(mine - ana).mean(axis=0) / scale
    err_ref = 100 * np.abs(ref_gpu - ana).mean(axis=0) / scale
    banner(f"  errors (%): drho {err[0]:.6f} dvx {err[1]:.6f} dvy {err[2]:.6f}; reference "
           f"GPU {err_ref[0]:.6f} {err_ref[1]:.6f} {err_ref[2]:.6f}")
    assert np.allclose(err_ref, REF_ERR, atol=1e-3)
    # below the reference GPU's own score on every column (drho/dt by a hair:
    # the continuity sum is the same operator, summed with less rounding)
    assert np.all(err < err_ref), (err, err_ref)
