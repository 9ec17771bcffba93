"""The reference's own 85,773-particle dam-break state, stepped on the GPU
against the oracle.

tests/golden/damstate85773.npz holds ArmandoList/new_{pos,vel,info}.txt, the
state outputData (armando2Dcpu.cu:1542-1566) wrote for the 85,773-particle
dam break of THINS Tab. 1 (case 3): 80,601 fluid and 5,172 wall particles,
listed in the C path's sorted order, the surge front already at the far
wall.  The test writes the three files back in the reference's text format,
reads them with armando_amd.io (read_c, walls_last), and steps the state 20
steps in the FP64 form on the GPU and in the oracle's gather form (same
physics, armando_amd.cases.c_path_dam_state): every field within 1e-10.  The
FP32 form is compared with the FP64 form.
"""
import os

import numpy as np
import pytest

import armando_amd
from armando_amd import cases
from armando_amd import io as aio
from oracle import orc
from tests.test_oracle_cpu import GOLDEN

pytestmark = pytest.mark.gpu

STEPS = 20


def banner(msg):
    print(f"[gpu-refstate] {msg}", flush=True)


def reference_state(tmp_path):
    fx = np.load(os.path.join(GOLDEN, "damstate85773.npz"))
    pos = fx["pos"].astype(np.float64)
    vel = fx["vel"].astype(np.float64)
    mat = fx["mat"].astype(int)
    mass, h = (float(v) for v in fx["mass_h"].astype(np.float64))
    d = str(tmp_path)
    np.savetxt(os.path.join(d, "new_pos.txt"), pos, fmt="%+14.8e")
    np.savetxt(os.path.join(d, "new_vel.txt"), vel, fmt="%+14.8e", delimiter=" ", newline=" \n")
    with open(os.path.join(d, "new_info.txt"), "w") as fh:
        for m in mat:
            fh.write(f"{m} {mass:+14.8e} {h:+14.8e} \n")
    st = aio.read_c(d, prefix="new_")
    st, npf = aio.walls_last(st)
    return st, npf


def relmax(a, b):
    return float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)


def test_reference_state_fp64_against_oracle(tmp_path):
    st, npf = reference_state(tmp_path)
    assert len(st["x"]) == 85773 and npf == 80601
    case = cases.c_path_dam_state(st, npf)
    banner(f"reference state n={case['n']} np={npf} grid {case['nx']} x {case['ny']}, "
           f"{STEPS} steps fp64")
    sim = armando_amd.Simulation(case)
    sim.time_integration(1, STEPS)
    g = sim.state()
    sim.close()
    P = orc.params_from(case)
    o = orc.OracleState(case["n"], case["np"])
    o.set(**{k: case[k] for k in ("x", "y", "vx", "vy", "rho", "u")})
    orc.run(P, o, 1, STEPS, gather=True)
    for k in ("x", "y", "rho", "p", "u"):
        e = relmax(g[k], o.field(k))
        banner(f"  {k} relmax {e:.2e}")
        assert e < 1e-10, k
    for k in ("vx", "vy"):
        e = float(np.max(np.abs(g[k] - o.field(k))))
        banner(f"  {k} max abs {e:.2e}")
        assert e < 1e-10 * 10.0, k


def test_reference_state_fp32_close_to_fp64(tmp_path):
    st, npf = reference_state(tmp_path)
    out = {}
    for prec in (cases.FP64, cases.FP32):
        case = cases.c_path_dam_state(st, npf, precision=prec)
        sim = armando_amd.Simulation(case)
        sim.time_integration(1, STEPS)
        out[prec] = sim.state()
        sim.close()
    a, b = out[cases.FP32], out[cases.FP64]
    for k, bound in (("rho", 1e-6), ("p", 1e-4), ("vx", 1e-4), ("vy", 1e-4), ("x", 1e-7),
                     ("y", 1e-7)):
        e = float(np.linalg.norm(a[k] - b[k]) / np.linalg.norm(b[k]))
        banner(f"  fp32 vs fp64 {k} rel-L2 {e:.2e}")
        assert e < bound, k
