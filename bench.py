"""Throughput of the WCSPH time step on one MI355X.

Workload: the 2D dam break with 720,000 fluid particles (0.6 x 1.2 m column
in a 3.2 x 1.6 m tank, dx = 1e-3 m, seeded jitter +-0.01 dx) plus 19,218 wall
particles, Tait EOS, Monaghan viscosity (alpha 0.01, beta 0), artificial heat,
XSPH (eps 0.5), leapfrog at dt = 0.2 h / c0.  One step is one pass of
time_integration: keys, sort, reorder + predictor, density, EOS, force
(+ artificial viscosity, artificial heat, XSPH), corrector.

The headline `value` is the FP32 chosen form in particle-steps/s (walls
included); the FP64 plain form runs the same steps for the FP32/FP64 speedup
and for the accuracy comparison.  Inputs are resident in HBM before the timed
region starts; the timed region is K whole steps.

The adjoint leg reports the adjoint / forward device-time ratio of the
rollout on the 100k dam break (400 steps, checkpoint every 20), FP32 form
(forward and backward) first, FP64 beside it.

The 3D leg runs the 3D dam break (0.6 x 0.6 x 1.2 m column in a
3.2 x 1.0 x 1.6 m tank, dx = 7.1e-3 m, three wall layers: 2,229,862
particles, Wendland C2 3D, 27-cell hash) in both forms.

The channel leg runs the 2D heated channel (0.4 x 0.05 m, periodic in x,
dx = 1e-4 m: 2,000,000 fluid + 24,000 wall particles; Poiseuille flow at
0.5 m/s peak, physical viscosity, Cleary-Monaghan conduction, artificial
heat, Gaussian heat source) for 2000 steps in both forms.

The ghost leg runs the 720k dam-break column against one row of virtual
particles carrying the wall normals, with mirror ghosts rebuilt every step
(ghost.f), for 500 steps in both forms.

The reference-state leg steps the reference's own 85,773-particle dam-break
state (ArmandoList/new_{pos,vel,info}.txt, committed as
tests/golden/damstate85773.npz) for 2000 steps in both forms, beside the
particle-steps/s of THINS Tab. 1 case 3 (85,773 particles, Tesla C1060).

The oracle leg steps both forms and the oracle (OpenMP on the host cores)
1000 steps from the same seeded 720k state and reports the errors of every
field and the conservation sums against the oracle.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-fp64]
                  [--no-adjoint] [--no-3d] [--no-channel] [--no-ghosts]
                  [--no-refstate] [--oracle-steps K]
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s, 720k 2D dam break, FP32 vs FP64; adjoint/forward time ratio"

# Algorithmic bytes per particle per launch: every particle-indexed input the
# pass needs read once plus every output written once (neighbour re-reads,
# cell-range arrays and uniform constants are not counted).
BYTES = {
    # force pass (internal_force + art_visc + art_heat fused): reads cell 4 +
    # offset 8 + v 8 + rho 4 + p 4 + u 4 + divv 4 + type 1, writes a 8 + du 4
    ("force", 1): 37 + 12,
    # FP64: reads x 16 + v 16 + rho 8 + p 8 + u 8 + divv 8 + type 1, writes a 16 + du 8
    ("force", 0): 65 + 24,
    # density + divv pass (FP32 / FP64)
    ("density", 1): 29,
    ("density", 0): 49,
    # 3D force gather: reads key, cell + offset (float4), v (float4), rho', p, psi, u,
    # id; writes a (float4), du, av (float4)
    ("force3", 1): 4 + 16 + 16 + 4 + 4 + 4 + 4 + 4 + (16 + 4 + 16),
    ("force3", 0): 4 + 32 + 32 + 8 + 8 + 8 + 8 + 4 + (32 + 8 + 32),
}
# What the FP32 force gather actually stages and writes per particle: key,
# cell, offset, v, rho', p, psi, u, id read; a, du, av written (the XSPH
# correction and the art_heat coefficient travel with it).
STAGED_BYTES = {1: 4 + 4 + 8 + 8 + 4 + 4 + 4 + 4 + 4 + (8 + 4 + 8),
                0: 4 + 16 + 16 + 8 + 8 + 8 + 8 + 4 + (16 + 8 + 16)}
HBM_PEAK_GBS = 8000.0
SIMDS, CLOCK_HZ = 1024, 2.4e9
PROFILE_PMC = os.path.join(ROOT, "profiles", "dam720k_pmc.md")


def profile_figures(kernel_prefix, path=PROFILE_PMC):
    """Per-launch figures of a kernel from the committed counter profile: the
    ```bench-figures``` block of JSON lines {"kernel", "traffic_bytes",
    "valu_insts", "commit", ...} that the profile records.  None when the
    file or the kernel is missing."""
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"```bench-figures\n(.*?)```", text, re.S)
    if not m:
        return None
    for line in m.group(1).splitlines():
        line = line.strip()
        if not line:
            continue
        rec = json.loads(line)
        if rec.get("kernel", "").startswith(kernel_prefix):
            return rec
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 form")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-3d", action="store_true", help="skip the 3D dam break leg")
    ap.add_argument("--steps-3d", type=int, default=100)
    ap.add_argument("--steps-3d-adjoint", type=int, default=500,
                    help="steps of the 3D FP32 adjoint rollout (0: skip)")
    ap.add_argument("--no-channel", action="store_true", help="skip the heated channel leg")
    ap.add_argument("--steps-channel", type=int, default=2000)
    ap.add_argument("--no-ghosts", action="store_true", help="skip the mirror-ghost leg")
    ap.add_argument("--steps-ghosts", type=int, default=500)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="approximate CPU work for the baseline sample")
    ap.add_argument("--scale", type=float, default=1.0,
                    help="dam-break size factor (1.0 = the 720k workload)")
    ap.add_argument("--oracle-steps", type=int, default=1000,
                    help="step both forms and the oracle (OpenMP, host cores) this many steps "
                         "from the initial state and report the errors and conservation sums "
                         "against the oracle (1000 steps: about 45 s of CPU work on 16 "
                         "threads; 0 skips)")
    ap.add_argument("--no-refstate", action="store_true",
                    help="skip the leg on the reference's 85,773-particle state")
    ap.add_argument("--steps-refstate", type=int, default=2000)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB three times (a kernel of "
                         "known traffic for scaling the memory counters of a profiler)")
    return ap.parse_args()


def copy_probe(mib):
    import torch

    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def run_form(precision, args, case_fn):
    import armando_amd

    case = case_fn(precision)
    sim = armando_amd.Simulation(case)
    niac = sim.link_list(count=True)
    # warmup steps (untimed), starting from the initial state
    sim.time_integration(1, args.warmup)
    its = 1 + args.warmup
    # timed region: whole steps, no per-pass events
    sim.set_timing(False)
    t0 = time.perf_counter()
    sim.time_integration(its, args.steps)  # returns after the stream finished
    wall = time.perf_counter() - t0
    its += args.steps
    # per-pass device times from HIP events on the library stream, over a
    # second window of the same length (events only bracket launches)
    sim.set_timing(True)
    sim.reset_pass_times()
    nt = max(20, min(args.steps, 200))
    sim.time_integration(its, nt)
    passes = sim.pass_times()
    sim.set_timing(False)
    state = sim.state()
    sim.close()
    return dict(case=case, niac=niac, wall=wall, passes=passes, timed_steps=nt, state=state)


def adjoint_ratio(steps=400, ckpt=20):
    """Adjoint / forward device time of the rollout on the 100k dam break
    (400 steps, checkpoints every 20 steps): the FP32 chosen form (forward
    and backward) as the headline, the FP64 plain form beside it."""
    import armando_amd
    from armando_amd import cases

    out = {}
    for name, prec in (("fp32", cases.FP32), ("fp64", cases.FP64)):
        case, obj = cases.dam_break_adjoint_100k(precision=prec)
        sim = armando_amd.Simulation(case)
        sim.adjoint_rollout(5, 5, obj)  # warm up the kernels
        sim.close()
        sim = armando_amd.Simulation(case)
        J, da, g, (t_fwd, t_bwd) = sim.adjoint_rollout(steps, ckpt, obj)
        sim.close()
        plain = armando_amd.Simulation(case)
        plain.time_integration(1, 5)
        plain.close()
        plain = armando_amd.Simulation(case)
        t0 = time.perf_counter()
        plain.time_integration(1, steps)
        t_plain = 1e3 * (time.perf_counter() - t0)
        plain.close()
        out[name] = {
            "ratio": t_bwd / t_fwd,
            "forward_ms": t_fwd,
            "backward_ms": t_bwd,
            "plain_forward_ms": t_plain,
            "ratio_vs_plain_forward": t_bwd / t_plain,
            "J": J,
            "dJ_dalpha": da,
            "dJ_dv0_norm": float(np.linalg.norm(g)),
        }
    return {
        "workload": f"dam break {case['np']} fluid + {case['nv']} wall particles, "
                    f"{steps} steps, checkpoint every {ckpt}",
        "ratio": out["fp32"]["ratio"],
        "dtype": "f32",
        **out,
    }


def dam3d_adjoint(steps, ckpt=20, seed=4243, vstd=0.01):
    """The adjoint rollout of the 3D dam break, FP32 chosen form: initial
    velocities N(0, vstd) per component on the fluid (seeded), objective the
    time-integrated p^2 of the left-wall particles below 0.48 m, checkpoints
    every `ckpt` steps.  Returns the adjoint/forward device-time ratio."""
    import armando_amd
    from armando_amd import cases

    c = cases.dam_break3d(precision=cases.FP32)
    rng = np.random.default_rng(seed)
    nf = c["np"]
    for k in ("vx", "vy", "vz"):
        c[k] = c[k].copy()
        c[k][:nf] = rng.normal(0.0, vstd, nf)
    obj = np.nonzero((c["x"] < 0) & (c["z"] < 0.48))[0].astype(np.int32)
    sim = armando_amd.Simulation(c)
    sim.adjoint_rollout(3, 3, obj)  # warm up the kernels
    sim.close()
    sim = armando_amd.Simulation(c)
    J, da, g, (t_fwd, t_bwd) = sim.adjoint_rollout(steps, ckpt, obj)
    sim.close()
    # memory of the rollout: one tape per step of a segment (the state after
    # the force pass, 165 bytes per particle, and the neighbour masks, 112),
    # one checkpoint per segment boundary (92 bytes per particle)
    tape_b, ck_b = 165 + 112, 92
    mem = c["n"] * (ckpt * tape_b + (steps // ckpt + 2) * ck_b)
    return {"steps": steps, "checkpoint_every": ckpt, "ratio": t_bwd / t_fwd,
            "forward_ms": t_fwd, "backward_ms": t_bwd, "J": J, "dJ_dalpha": da,
            "dJ_dv0_norm": float(np.linalg.norm(g)), "objective_particles": int(len(obj)),
            "rollout_bytes": mem}


def dam3d_leg(steps, warmup=5, adjoint_steps=500):
    """The 3D dam break in the FP32 chosen and the FP64 plain form: whole
    steps of time_integration, per-pass device times, FP32-vs-FP64 accuracy;
    then the FP32 adjoint rollout."""
    import armando_amd
    from armando_amd import cases

    res = {}
    for name, prec in (("fp32", cases.FP32), ("fp64", cases.FP64)):
        case = cases.dam_break3d(precision=prec)
        sim = armando_amd.Simulation(case)
        niac = sim.link_list(count=True)
        sim.time_integration(1, warmup)
        t0 = time.perf_counter()
        sim.time_integration(1 + warmup, steps)
        wall = time.perf_counter() - t0
        sim.set_timing(True)
        sim.reset_pass_times()
        sim.time_integration(1 + warmup + steps, 20)
        passes = {k: ms / max(cnt, 1) for k, (ms, cnt) in sim.pass_times().items()}
        sim.set_timing(False)
        res[name] = dict(case=case, niac=niac, wall=wall, passes=passes, state=sim.state())
        sim.close()
    r32, r64 = res["fp32"], res["fp64"]
    case = r32["case"]
    n = case["n"]
    t_force = r32["passes"]["force"] * 1e-3
    achieved = n * BYTES[("force3", 1)] / t_force / 1e9
    acc = {}
    for k in ("rho", "p", "vx", "vy", "vz", "x", "y", "z"):
        d = r32["state"][k] - r64["state"][k]
        nrm = np.linalg.norm(r64["state"][k])
        acc[k] = float(np.linalg.norm(d) / nrm) if nrm > 0 else float(np.linalg.norm(d))
    v32 = n * steps / r32["wall"]
    v64 = n * steps / r64["wall"]
    return {
        "workload": (f"3D dam break, {case['np']} fluid + {case['nv']} wall particles, "
                     f"dx 7.1e-3 m, grid {case['nx']}x{case['ny']}x{case['nz']}"),
        "value": v32,
        "unit": "particle-steps/s",
        "dtype": "f32",
        "ms_per_step": 1e3 * r32["wall"] / steps,
        "steps": steps,
        "pairs_i_lt_j": r32["niac"],
        "pair_interactions_per_s": 2 * r32["niac"] / t_force,
        "passes_ms": r32["passes"],
        "fp64": {"value": v64, "ms_per_step": 1e3 * r64["wall"] / steps,
                 "passes_ms": r64["passes"]},
        "speedup_fp32_over_fp64": v32 / v64,
        "roofline_force": {"kernel": "k3_force<float,...>", "bound": "hbm",
                           "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBS,
                           "bytes_per_particle": BYTES[("force3", 1)],
                           "avg_launch_ms": t_force * 1e3},
        "accuracy_fp32_vs_fp64_rel_l2": {"after_steps": warmup + steps + 20, **acc},
        "adjoint": dam3d_adjoint(adjoint_steps) if adjoint_steps else None,
    }


def channel_leg(steps, warmup=5):
    """The heated channel in the FP32 chosen and the FP64 plain form: whole
    steps of time_integration, per-pass device times, FP32-vs-FP64 accuracy
    and the temperature rise under the source."""
    import armando_amd
    from armando_amd import cases

    res = {}
    for name, prec in (("fp32", cases.FP32), ("fp64", cases.FP64)):
        case = cases.heated_channel(precision=prec)
        sim = armando_amd.Simulation(case)
        niac = sim.link_list(count=True)
        sim.time_integration(1, warmup)
        t0 = time.perf_counter()
        sim.time_integration(1 + warmup, steps)
        wall = time.perf_counter() - t0
        sim.set_timing(True)
        sim.reset_pass_times()
        sim.time_integration(1 + warmup + steps, 20)
        passes = {k: ms / max(cnt, 1) for k, (ms, cnt) in sim.pass_times().items()}
        sim.set_timing(False)
        res[name] = dict(case=case, niac=niac, wall=wall, passes=passes, state=sim.state())
        sim.close()
    r32, r64 = res["fp32"], res["fp64"]
    case = r32["case"]
    n, nf = case["n"], case["np"]
    s32, s64 = r32["state"], r64["state"]
    acc = {}
    for k in ("rho", "u", "x", "y"):
        acc[k] = float(np.linalg.norm(s32[k] - s64[k]) / np.linalg.norm(s64[k]))
    v32 = np.concatenate([s32["vx"], s32["vy"]])
    v64 = np.concatenate([s64["vx"], s64["vy"]])
    acc["v"] = float(np.linalg.norm(v32 - v64) / np.linalg.norm(v64))
    # per component: the flow is along x, so vy is a small residue whose own
    # relative error is large; its error against |v| is the one that matters
    for k in ("vx", "vy"):
        acc[k] = float(np.linalg.norm(s32[k] - s64[k]) / np.linalg.norm(s64[k]))
    acc["vy_err_over_v"] = float(np.linalg.norm(s32["vy"] - s64["vy"]) / np.linalg.norm(v64))
    dT = (s32["u"][:nf] - case["u"][:nf]) / case["cp"]
    return {
        "workload": (f"2D heated channel, {nf} fluid + {case['nv']} wall particles, periodic x, "
                     f"Q0 {case['hs_q0']:.4g} W/m^3"),
        "value": n * steps / r32["wall"],
        "unit": "particle-steps/s",
        "dtype": "f32",
        "ms_per_step": 1e3 * r32["wall"] / steps,
        "steps": steps,
        "pairs_i_lt_j": r32["niac"],
        "passes_ms": r32["passes"],
        "fp64": {"value": n * steps / r64["wall"], "ms_per_step": 1e3 * r64["wall"] / steps,
                 "passes_ms": r64["passes"]},
        "speedup_fp32_over_fp64": r64["wall"] / r32["wall"],
        "max_temperature_rise_K": float(np.max(dT)),
        "accuracy_fp32_vs_fp64_rel_l2": {"after_steps": warmup + steps + 20, **acc},
    }


def ghost_leg(steps, warmup=5):
    """The 720k dam-break column against virtual walls with mirror ghosts,
    both forms: whole steps of time_integration (the ghost count is read
    back every step), per-pass device times and the ghost count."""
    import armando_amd
    from armando_amd import cases

    res = {}
    for name, prec in (("fp32", cases.FP32), ("fp64", cases.FP64)):
        case = cases.ghost_tank(col_w=0.6, col_h=1.2, tank_w=3.2, precision=prec)
        sim = armando_amd.Simulation(case)
        sim.time_integration(1, warmup)
        t0 = time.perf_counter()
        sim.time_integration(1 + warmup, steps)
        wall = time.perf_counter() - t0
        ng = sim.ghost_count()
        sim.set_timing(True)
        sim.reset_pass_times()
        sim.time_integration(1 + warmup + steps, 20)
        passes = {k: ms / max(cnt, 1) for k, (ms, cnt) in sim.pass_times().items()}
        sim.set_timing(False)
        res[name] = dict(case=case, wall=wall, passes=passes, ng=ng)
        sim.close()
    r32, r64 = res["fp32"], res["fp64"]
    case = r32["case"]
    n = case["n"]
    return {
        "workload": (f"2D dam break, {case['np']} fluid + {case['nv']} virtual wall particles "
                     f"with mirror ghosts (ghost.f)"),
        "value": n * steps / r32["wall"],
        "unit": "particle-steps/s",
        "dtype": "f32",
        "ms_per_step": 1e3 * r32["wall"] / steps,
        "steps": steps,
        "ghosts_at_end": r32["ng"],
        "passes_ms": r32["passes"],
        "fp64": {"value": n * steps / r64["wall"], "ms_per_step": 1e3 * r64["wall"] / steps,
                 "ghosts_at_end": r64["ng"], "passes_ms": r64["passes"]},
        "speedup_fp32_over_fp64": r64["wall"] / r32["wall"],
    }


def oracle_leg(steps, case_fn):
    """Both forms and the oracle's gather form (time_integration.f:131-235,
    on the host's cores) stepped `steps` steps from the same seeded state;
    rel-L2, max-abs and max-relative errors of every field against the
    oracle."""
    import armando_amd
    from armando_amd import cases
    from oracle import orc

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    orc.lib().orc_set_threads(threads)
    case = case_fn(cases.FP64)
    P = orc.params_from(case)
    st = orc.OracleState(case["n"], case["np"])
    st.set(**{k: case[k] for k in ("x", "y", "vx", "vy", "rho", "u")})
    t0 = time.perf_counter()
    orc.run(P, st, 1, steps, gather=True)
    t_orc = time.perf_counter() - t0
    names = ("rho", "p", "vx", "vy", "u", "x", "y")
    ref = {k: st.field(k).copy() for k in names}
    st0 = orc.OracleState(case["n"], case["np"])
    st0.set(**{k: case[k] for k in ("x", "y", "vx", "vy", "rho", "u")})
    cons_keys = ("mass", "momentum_x", "momentum_y", "kinetic", "internal")
    out = {"after_steps": steps, "oracle_threads": threads, "oracle_seconds": t_orc,
           "conservation": {"sums": list(cons_keys),
                            "oracle": {"start": orc.conservation(P, st0).tolist(),
                                       "end": orc.conservation(P, st).tolist()}}}
    for tag, prec in (("fp64", cases.FP64), ("fp32", cases.FP32)):
        sim = armando_amd.Simulation(case_fn(prec))
        c0 = sim.conservation()
        sim.time_integration(1, steps)
        s = sim.state()
        c1 = sim.conservation()
        sim.close()
        ref_end = orc.conservation(P, st)
        out["conservation"][tag] = {
            "start": c0.tolist(), "end": c1.tolist(),
            # against the oracle's sums after the same steps: the mass exactly,
            # momenta on the flow's scale sqrt(2 M KE), energies relative
            "vs_oracle": {
                "mass_rel": float(abs(c1[0] - ref_end[0]) / ref_end[0]),
                "momentum_rel": float(max(abs(c1[1] - ref_end[1]), abs(c1[2] - ref_end[2]))
                                      / np.sqrt(2.0 * ref_end[0] * max(ref_end[3], 1e-300))),
                "kinetic_rel": float(abs(c1[3] - ref_end[3]) / max(abs(ref_end[3]), 1e-300)),
                "internal_rel": float(abs(c1[4] - ref_end[4]) / max(abs(ref_end[4]), 1e-300)),
            }}
        acc = {}
        for k in names:
            d = s[k] - ref[k]
            scale = float(np.max(np.abs(ref[k])))
            acc[k] = {"rel_l2": float(np.linalg.norm(d) / max(np.linalg.norm(ref[k]), 1e-300)),
                      "max_abs": float(np.max(np.abs(d))),
                      "max_rel": float(np.max(np.abs(d))) / max(scale, 1e-300)}
        out[tag] = acc
    return out


def refstate_leg(steps, warmup=5):
    """The reference's 85,773-particle dam-break state (outputData of
    armando2Dcpu.cu, committed as tests/golden/damstate85773.npz) stepped by
    both forms; THINS Tab. 1 case 3 (85,773 particles) ran 40,000 steps in
    8 m 34 s on a Tesla C1060 and 784 m on one 2.67 GHz Xeon core."""
    import armando_amd
    from armando_amd import cases
    from armando_amd import io as aio

    fx = np.load(os.path.join(ROOT, "tests", "golden", "damstate85773.npz"))
    pos, vel = fx["pos"].astype(np.float64), fx["vel"].astype(np.float64)
    mass, h = (float(v) for v in fx["mass_h"].astype(np.float64))
    n = len(pos)
    st = {"x": pos[:, 0], "y": pos[:, 1], "vx": vel[:, 0], "vy": vel[:, 1],
          "mat": fx["mat"].astype(np.int64), "mass": np.full(n, mass), "h": np.full(n, h)}
    st, npf = aio.walls_last(st)
    res = {}
    for name, prec in (("fp32", cases.FP32), ("fp64", cases.FP64)):
        case = cases.c_path_dam_state(st, npf, precision=prec)
        sim = armando_amd.Simulation(case)
        sim.time_integration(1, warmup)
        t0 = time.perf_counter()
        sim.time_integration(1 + warmup, steps)
        res[name] = time.perf_counter() - t0
        sim.close()
    thins_gpu = 85773 * 40000 / (8 * 60 + 34)
    thins_cpu = 85773 * 40000 / (784 * 60)
    v32 = n * steps / res["fp32"]
    return {
        "workload": (f"the reference's dam-break state, {npf} fluid + {n - npf} wall particles "
                     f"(ArmandoList/new_*.txt), grid {case['nx']}x{case['ny']}"),
        "value": v32, "unit": "particle-steps/s", "dtype": "f32", "steps": steps,
        "ms_per_step": 1e3 * res["fp32"] / steps,
        "fp64": {"value": n * steps / res["fp64"], "ms_per_step": 1e3 * res["fp64"] / steps},
        "thins_tab1_case3": {
            "gpu_c1060_particle_steps_per_s": thins_gpu,
            "cpu_1core_particle_steps_per_s": thins_cpu,
            "note": "the reference step is TVD-RK3 (three force evaluations); one leapfrog "
                    "step here is one",
            "ratio_fp32_vs_c1060": v32 / thins_gpu},
    }


def cpu_baseline(case, seconds):
    """The oracle's faithful serial restatement of the reference loop
    (half pair list, scatter; one core) on the same seeded input."""
    from oracle import orc

    P = orc.params_from(case)
    st = orc.OracleState(case["n"], case["np"])
    st.set(**{k: case[k] for k in ("x", "y", "vx", "vy", "rho", "u")})
    t0 = time.perf_counter()
    orc.run(P, st, 1, 1, gather=False)
    one = time.perf_counter() - t0
    k = max(1, int(seconds / max(one, 1e-3)) - 1)
    t0 = time.perf_counter()
    orc.run(P, st, 2, k, gather=False)
    dt = time.perf_counter() - t0
    out = dict(value=case["n"] * k / dt, unit="particle-steps/s", cores=1, kind="port",
               sample=f"steps 2..{k + 1} of the same 720k dam break, oracle half-list "
                      f"restatement of CERNsource time_integration.f, 1 thread, "
                      f"{dt:.1f} s")
    # the oracle's gather form on every host core the job may use (OpenMP)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    orc.lib().orc_set_threads(threads)
    t0 = time.perf_counter()
    orc.run(P, st, k + 2, 1, gather=True)
    one = time.perf_counter() - t0
    kk = max(1, int(0.75 * seconds / max(one, 1e-3)))
    t0 = time.perf_counter()
    orc.run(P, st, k + 3, kk, gather=True)
    dt2 = time.perf_counter() - t0
    out["all_cores"] = dict(value=case["n"] * kk / dt2, unit="particle-steps/s", cores=threads,
                            kind="port",
                            sample=f"{kk} further steps, oracle gather form (cell order), "
                                   f"OpenMP {threads} threads, {dt2:.1f} s")
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the step is not "
              f"sharded across GPUs in this build", file=sys.stderr)
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    from armando_amd import cases

    if args.copy_probe:
        copy_probe(args.copy_probe)
    if args.scale == 1.0:
        case_fn = lambda prec: cases.dam_break_720k(precision=prec)  # noqa: E731
    else:
        case_fn = lambda prec: cases.dam_break_small(scale=args.scale, precision=prec)  # noqa: E731

    r32 = run_form(cases.FP32, args, case_fn)
    r64 = None if args.no_fp64 else run_form(cases.FP64, args, case_fn)
    case = r32["case"]
    n = case["n"]
    ms_step = 1e3 * r32["wall"] / args.steps
    value = n * args.steps / r32["wall"]

    def pass_ms(r, name):
        ms, cnt = r["passes"][name]
        return ms / max(cnt, 1)

    t_force = pass_ms(r32, "force") * 1e-3
    bytes_force = n * BYTES[("force", 1)]
    achieved = bytes_force / t_force / 1e9
    prof = profile_figures("arm::k_force_hot<") if args.scale == 1.0 else None
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"2D dam break, {case['np']} fluid + {case['nv']} wall particles, "
                         f"FP32 chosen form (FP64 position/energy accumulators)"),
            "n_particles": n,
            "pairs_i_lt_j": r32["niac"],
            "dt": case["dt"],
            "h": case["h"],
            "grid": [case["nx"], case["ny"]],
            "parallelism": "single GPU",
        },
        "pair_interactions_per_s": 2 * r32["niac"] / t_force,
        "passes_ms": {k: pass_ms(r32, k) for k in r32["passes"]},
        "roofline": {
            "kernel": "k_force_hot<...> (internal_force + art_visc + art_heat + av_vel)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": prof["traffic_bytes"] if prof else None,
            "bytes_per_particle": BYTES[("force", 1)],
            "avg_launch_ms": t_force * 1e3,
            "staged_bytes_per_particle": STAGED_BYTES[1],
            "staged_frac": n * STAGED_BYTES[1] / t_force / 1e9 / HBM_PEAK_GBS,
            "valu_issue_frac": (prof["valu_insts"] * 4 / SIMDS / (t_force * CLOCK_HZ)
                                if prof else None),
            "counters_source": (f"profiles/dam720k_pmc.md, kernel {prof['kernel']}, "
                                f"commit {prof['commit']}" if prof else None),
        },
    }
    if r64 is not None:
        v64 = n * args.steps / r64["wall"]
        s32, s64 = r32["state"], r64["state"]
        acc = {}
        for k in ("rho", "p", "vx", "vy", "u", "x", "y"):
            d = s32[k] - s64[k]
            nrm = np.linalg.norm(s64[k])
            acc[k] = {"rel_l2": float(np.linalg.norm(d) / nrm) if nrm > 0 else float(np.linalg.norm(d)),
                      "max_abs": float(np.max(np.abs(d)))}
        out["fp64"] = {
            "value": v64,
            "ms_per_step": 1e3 * r64["wall"] / args.steps,
            "speedup_fp32_over_fp64": value / v64,
            "passes_ms": {k: pass_ms(r64, k) for k in r64["passes"]},
            "force_achieved_GBs": n * BYTES[("force", 0)] / (pass_ms(r64, "force") * 1e-3) / 1e9,
        }
        out["accuracy_fp32_vs_fp64"] = {
            "after_steps": args.warmup + args.steps + r32["timed_steps"], **acc}
    if not args.no_adjoint:
        out["adjoint"] = adjoint_ratio()
    if not args.no_3d:
        out["dam_break_3d"] = dam3d_leg(args.steps_3d, adjoint_steps=args.steps_3d_adjoint)
    if not args.no_channel:
        out["heated_channel"] = channel_leg(args.steps_channel)
    if not args.no_ghosts:
        out["ghost_walls"] = ghost_leg(args.steps_ghosts)
    if not args.no_refstate:
        out["reference_state"] = refstate_leg(args.steps_refstate)
    if args.oracle_steps > 0:
        out["accuracy_vs_oracle"] = oracle_leg(args.oracle_steps, case_fn)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
