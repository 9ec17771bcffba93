"""Restart through the reference's Fortran state files: a state written
with 17 digits and read back continues bit for bit."""
import numpy as np
import pytest

import armando_amd
from armando_amd import cases
from armando_amd import io as aio

pytestmark = pytest.mark.gpu


def test_restart_through_fortran_files_is_exact(tmp_path):
    c = cases.dam_break_small(scale=0.03)
    print(f"[gpu-io] restart n={c['n']}", flush=True)
    a = armando_amd.Simulation(c)
    a.time_integration(1, 6)
    s = a.state()
    aio.write_fortran(tmp_path, s, c["np"], h=c["h"], mass=c["mass"], precise=True)
    f = aio.read_fortran(tmp_path, "f")
    v = aio.read_fortran(tmp_path, "v")
    st = {k: np.concatenate([f[k], v[k]]) for k in ("x", "y", "vx", "vy", "rho", "u")}
    b = armando_amd.Simulation(aio.case_from_state(c, st, c["np"]))
    sb = b.state()
    for k in ("x", "y", "vx", "vy", "rho", "u"):
        assert np.array_equal(sb[k], s[k]), k
    # both restart with the leapfrog's first step (its = 1), as the reference does
    a.time_integration(1, 5)
    b.time_integration(1, 5)
    sa, sb = a.state(), b.state()
    for k in ("x", "y", "vx", "vy", "rho", "u", "p"):
        assert np.array_equal(sa[k], sb[k]), k
    a.close()
    b.close()
