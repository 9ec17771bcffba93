"""Physical viscosity (internal_force.f's `visc` branches) and the linear
EOS on the GPU, through the C ABI, against the oracle.

Bar: FP64 within 1e-10 of the oracle's gather twin (same neighbour order,
different last bits of 1/rho); the tiled and direct gathers bit-identical;
the FP32 form within 1e-4 relative L2 of FP64.
"""
import numpy as np
import pytest

import armando_amd
from armando_amd import cases
from oracle import orc

pytestmark = pytest.mark.gpu

FIELDS_IN = ("x", "y", "vx", "vy", "rho", "u")


def banner(msg):
    print(f"[gpu-visc] {msg}", flush=True)


def oracle_state(case):
    st = orc.OracleState(case["n"], case["np"])
    st.set(**{k: case[k] for k in FIELDS_IN})
    return st


def relmax(a, b):
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / scale


def visc_box(pa, eos, precision=cases.FP64):
    c = dict(cases.box_lattice(nx=30, ny=24, vel=0.5, precision=precision))
    c["flags"] = c["flags"] | cases.VISC
    c["visc_eta"] = 5.0
    c["pa_sph"] = pa
    c["eos"] = eos
    c["eos_p0"] = -5e4 if eos == cases.EOS_LINEAR else 0.0
    return c


@pytest.mark.parametrize("pa", [0, 1])
@pytest.mark.parametrize("eos", [cases.EOS_TAIT, cases.EOS_LINEAR])
def test_routines_with_viscosity_fp64(pa, eos):
    c = visc_box(pa, eos)
    banner(f"routines visc pa_sph={pa} eos={eos} n={c['n']}")
    sim = armando_amd.Simulation(c)
    sim.link_list()
    sim.density(0)
    sim.eos()
    sim.internal_force(c["flags"])
    sim.av_vel()
    g = sim.rates()
    s = sim.state()
    P = orc.params_from(c)
    st = oracle_state(c)
    orc.single_step_gather(P, st)
    for k in ("drho", "ax", "ay", "du", "avx", "avy"):
        assert relmax(g[k], st.field(k)) < 1e-10, k
    assert relmax(s["p"], st.field("p")) < 1e-12
    # the viscous terms matter in this case
    c0 = dict(c)
    c0["flags"] = c["flags"] & ~cases.VISC
    st0 = oracle_state(c0)
    orc.single_step_gather(orc.params_from(c0), st0)
    assert relmax(st.field("ax"), st0.field("ax")) > 1e-3
    sim.close()


def test_shear_cavity_3000_steps_fp64():
    """The lid-driven cavity (1681 fluid + 320 wall particles, 3000 steps of
    5e-5 s) against the oracle's gather twin."""
    c = cases.shear_cavity()
    banner(f"shear cavity n={c['n']} 3000 steps")
    sim = armando_amd.Simulation(c)
    sim.time_integration(1, 3000)
    s = sim.state()
    sim.close()
    P = orc.params_from(c)
    st = oracle_state(c)
    orc.run(P, st, 1, 3000, gather=True)
    nf = c["np"]
    for k in ("x", "y", "rho", "u"):
        assert relmax(s[k], st.field(k)) < 1e-10, k
    for k in ("vx", "vy"):
        assert relmax(s[k][:nf], st.field(k)[:nf]) < 1e-8, k
    # the lid has set up the circulation: fluid under the lid moves with it
    top = (c["y"][:nf] > 0.9 * np.max(c["y"]))
    assert np.mean(s["vx"][:nf][top]) < -1e-4


@pytest.mark.parametrize("prec", [cases.FP64, cases.FP32])
def test_tiled_equals_direct_with_viscosity(prec):
    c = cases.shear_cavity(precision=prec)
    banner(f"visc tiled vs direct prec={prec}")
    a = armando_amd.Simulation(c)
    b = armando_amd.Simulation(c)
    b.set_gather_path(False)
    a.time_integration(1, 20)
    b.time_integration(1, 20)
    sa, sb = a.state(), b.state()
    ra, rb = a.rates(), b.rates()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    for k in ra:
        assert np.array_equal(ra[k], rb[k]), k
    a.close()
    b.close()


def test_fp32_form_close_to_fp64_with_viscosity():
    a = armando_amd.Simulation(cases.shear_cavity())
    b = armando_amd.Simulation(cases.shear_cavity(precision=cases.FP32))
    banner("visc fp32 vs fp64, 500 steps")
    a.time_integration(1, 500)
    b.time_integration(1, 500)
    sa, sb = a.state(), b.state()
    for k in ("rho", "p", "vx", "vy", "x", "y"):
        rel = np.linalg.norm(sb[k] - sa[k]) / np.linalg.norm(sa[k])
        banner(f"  {k} rel-L2 {rel:.3e}")
        assert rel < 1e-4, k
    a.close()
    b.close()


def test_adjoint_rejects_viscosity():
    c = cases.shear_cavity()
    sim = armando_amd.Simulation(c)
    with pytest.raises(armando_amd.ArmandoError) as e:
        sim.rollout_forward(4, 2, np.array([0], np.int32))
    assert e.value.code == -5
    sim.close()
