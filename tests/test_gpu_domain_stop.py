"""arm_time_integration stops at the first step that puts a particle outside
the hash grid, as link_list.f:63-67 stops the reference run.

One fluid particle at the top of the column is shot upwards so that it
leaves the grid in the middle of a 20-step call.  The reference point is the
same state stepped one step per call: the call of step s fails, which fixes
s and the state that step leaves.  The 20-step call must report that step,
run none of the later ones, and leave that same state bit for bit.
"""
import numpy as np
import pytest

import armando_amd
from armando_amd import ArmandoError, cases

pytestmark = pytest.mark.gpu

STEPS = 20


def banner(msg):
    print(f"[gpu-domain] {msg}", flush=True)


def shot_case(case, axis):
    """The highest fluid particle along `axis` gets the velocity that takes it
    past the grid's far face in about 7.3 steps."""
    c = dict(case)
    nf = c["np"]
    top = int(np.argmax(c[axis][:nf]))
    o, nc = {"y": ("oy", "ny"), "z": ("oz", "nz")}[axis]
    face = c[o] + c[nc] * c["cell"]
    vk = "v" + axis
    c[vk] = c[vk].copy()
    c[vk][top] = (face - c[axis][top]) / (7.3 * c["dt"])
    return c, top, face


def stepwise_failure(c):
    sim = armando_amd.Simulation(c)
    for its in range(1, STEPS + 1):
        try:
            sim.time_integration(its, 1)
        except ArmandoError as e:
            assert e.code == -4, e
            st = sim.state()
            sim.close()
            return its, st
    sim.close()
    raise AssertionError("the particle did not leave the grid")


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("precision", [cases.FP64, cases.FP32])
def test_time_integration_stops_at_the_first_domain_error(dim, precision):
    if dim == 2:
        base, axis = cases.dam_break_small(scale=0.05, precision=precision), "y"
    else:
        base, axis = cases.dam_break3d_small(scale=0.1, precision=precision), "z"
    c, top, face = shot_case(base, axis)
    banner(f"dim {dim} precision {precision}: n={c['n']}, particle {top} shot along {axis}")
    s_fail, ref = stepwise_failure(c)
    banner(f"  one step per call: fails at step {s_fail}")
    assert 3 < s_fail < STEPS
    assert ref[axis][top] >= face
    sim = armando_amd.Simulation(c)
    with pytest.raises(ArmandoError) as ei:
        sim.time_integration(1, STEPS)
    msg = str(ei.value)
    banner(f"  {STEPS}-step call: {msg}")
    assert ei.value.code == -4
    assert f"after step {s_fail} " in msg
    got = sim.state()
    sim.close()
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k
