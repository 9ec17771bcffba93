"""The FP64 conservation reductions (arm_conservation, arm_conservation3)
against the oracle's sums (orc_conservation: mass, momentum and the kinetic
and internal energy over the real particles, the budget of
FORTRANsource/time_integration.f90:185-203) after 50 steps of the small dam
break, in both forms.  The FP64 form steps the oracle's arithmetic, so its
sums agree to round-off; the FP32 form's sums carry its FP32 velocities, so
they agree to the FP32 form's field error.
"""
import numpy as np
import pytest

import armando_amd
from armando_amd import cases
from oracle import orc

pytestmark = pytest.mark.gpu

STEPS = 50


def banner(msg):
    print(f"[gpu-cons] {msg}", flush=True)


@pytest.fixture(scope="module")
def oracle_sums():
    case = cases.dam_break_small(scale=0.05)
    P = orc.params_from(case)
    st = orc.OracleState(case["n"], case["np"])
    st.set(**{k: case[k] for k in ("x", "y", "vx", "vy", "rho", "u")})
    b0 = orc.conservation(P, st)
    orc.run(P, st, 1, STEPS, gather=True)
    return case, b0, orc.conservation(P, st)


@pytest.mark.parametrize("precision,tol", [(cases.FP64, 1e-10), (cases.FP32, 1e-4)])
def test_conservation_sums_match_oracle(oracle_sums, precision, tol):
    case0, b0, b1 = oracle_sums
    case = cases.dam_break_small(scale=0.05, precision=precision)
    banner(f"conservation sums n={case['n']} precision={precision} steps={STEPS}")
    sim = armando_amd.Simulation(case)
    g0 = sim.conservation()
    assert np.array_equal(g0[1:], b0[1:])  # at rest, u = 0: every sum is an exact 0
    assert abs(g0[0] - b0[0]) <= 1e-13 * b0[0]
    sim.time_integration(1, STEPS)
    g1 = sim.conservation()
    g3 = sim.conservation3()
    sim.close()
    assert abs(g1[0] - b1[0]) <= 1e-13 * b1[0]  # the same particle count
    # momentum against the flow's momentum scale sqrt(2 M KE) >= |sum m v|
    pscale = np.sqrt(2.0 * b1[0] * b1[3])
    for k, name in ((1, "mx"), (2, "my")):
        e = abs(g1[k] - b1[k]) / pscale
        banner(f"  {name} {g1[k]:+.6e} oracle {b1[k]:+.6e} rel {e:.1e}")
        assert e < tol, name
    for k, name in ((3, "ke"), (4, "ie")):
        e = abs(g1[k] - b1[k]) / max(abs(b1[k]), 1e-300)
        banner(f"  {name} {g1[k]:.6e} oracle {b1[k]:.6e} rel {e:.1e}")
        assert e < tol, name
    # the 3-component form carries the same sums with a zero z momentum
    assert np.array_equal(g3[[0, 1, 2, 4, 5]], g1) and g3[3] == 0.0
