"""The FP32 neighbour pre-test of the 3D density pass (x = cx cell + off
rounded once per particle, against a bound on (2h)^2 through that rounding)
followed by the exact test: it must find exactly the pairs of the exact
test.  Lattices with h = dx put every axis-aligned second neighbour at
r = 2h, jittered by 1e-6 dx, so that thousands of pairs straddle the
cut-off.

Bar: bit-identical rates of the force passes driven by the masks the
density pass recorded and by the exact walk (3D, and 2D where the density
pass tests exactly), and of the LDS-tiled and direct 2D density passes on
the same straddling lattice."""
import numpy as np
import pytest

import armando_amd
from armando_amd import cases

pytestmark = pytest.mark.gpu

RATES_F = ("ax", "ay", "du", "avx", "avy")


def banner(msg):
    print(f"[gpu] {msg}", flush=True)


def straddling2d():
    return cases.box_lattice(nx=48, ny=40, dx=0.01, jitter=1e-6, h_fac=1.0,
                             precision=cases.FP32, vel=0.1)


def straddling3d():
    return cases.box_lattice3(nx=16, ny=15, nz=14, dx=0.01, jitter=1e-6, h_fac=1.0,
                              precision=cases.FP32, vel=0.1)


def rates_masked_and_exact(case):
    """internal_force after density (recorded masks) and after a second
    link_list (same order, masks dropped: the exact walk)."""
    a = armando_amd.Simulation(case)
    a.link_list()
    a.density(0)
    a.eos()
    a.internal_force(case["flags"])
    ra = a.rates()
    b = armando_amd.Simulation(case)
    b.link_list()
    b.density(0)
    b.link_list()
    b.eos()
    b.internal_force(case["flags"])
    rb = b.rates()
    a.close()
    b.close()
    return ra, rb


def test_straddling2d_tiled_density_equals_direct():
    c = straddling2d()
    banner(f"pre-test 2D tiled vs direct density n={c['n']}")
    a = armando_amd.Simulation(c)
    b = armando_amd.Simulation(c)
    b.set_gather_path(False)
    for s in (a, b):
        s.link_list()
        s.density(0)
        s.eos()
        s.internal_force(c["flags"])
    ra, rb = a.rates(), b.rates()
    a.close()
    b.close()
    assert np.count_nonzero(ra["drho"]) > 0.9 * c["n"]
    for k in ("drho",) + RATES_F:
        assert np.array_equal(ra[k], rb[k]), k


@pytest.mark.parametrize("dim", [2, 3])
def test_pretest_masks_equal_exact_walk(dim):
    c = straddling2d() if dim == 2 else straddling3d()
    banner(f"pre-test masks vs exact walk dim={dim} n={c['n']}")
    ra, rb = rates_masked_and_exact(c)
    axes = RATES_F if dim == 2 else RATES_F + ("az", "avz")
    for k in axes:
        assert np.array_equal(ra[k], rb[k]), k
