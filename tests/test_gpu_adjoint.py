"""GPU discrete adjoint (arm_adjoint_rollout) against the oracle's adjoint,
against central finite differences of the GPU forward, and invariance to the
checkpoint interval.

Objective: J = sum over steps of dt * sum p^2 over the left-wall particles
below 0.2 of the column height.  Initial velocities: N(0, 0.01 m/s) from a
seeded generator on the fluid particles.
"""
import numpy as np
import pytest

import armando_amd
from armando_amd import cases
from oracle import orc

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-adj] {msg}", flush=True)


def adjoint_case(scale=0.02, seed=777, precision=cases.FP64):
    c = cases.dam_break_small(scale=scale, precision=precision)
    rng = np.random.default_rng(seed)
    nf = c["np"]
    c["vx"] = c["vx"].copy()
    c["vy"] = c["vy"].copy()
    c["vx"][:nf] = rng.normal(0.0, 0.01, nf)
    c["vy"][:nf] = rng.normal(0.0, 0.01, nf)
    obj = np.nonzero((c["x"] < 0) & (c["y"] < 0.2 * 1.2 * scale))[0].astype(np.int32)
    return c, obj


def gpu_rollout(c, obj, nsteps, ckpt, alpha=None):
    c2 = dict(c)
    if alpha is not None:
        c2["av_alpha"] = alpha
    sim = armando_amd.Simulation(c2)
    out = sim.adjoint_rollout(nsteps, ckpt, obj)
    sim.close()
    return out


def gpu_objective(c, obj, nsteps, ckpt, alpha=None):
    c2 = dict(c)
    if alpha is not None:
        c2["av_alpha"] = alpha
    sim = armando_amd.Simulation(c2)
    J = sim.rollout_forward(nsteps, ckpt, obj)[0]
    sim.close()
    return J


def test_adjoint_matches_oracle():
    c, obj = adjoint_case()
    steps = 20
    banner(f"gpu vs oracle adjoint n={c['n']} steps={steps}")
    J, da, g, _ = gpu_rollout(c, obj, steps, 6)
    P = orc.params_from(c)
    st = orc.OracleState(c["n"], c["np"])
    st.set(**{k: c[k] for k in ("x", "y", "vx", "vy", "rho", "u")})
    mask = np.zeros(c["n"], np.int32)
    mask[obj] = 1
    Jo, dao, gxo, gyo = orc.adjoint(P, st, steps, mask)
    assert abs(J - Jo) <= 1e-10 * abs(Jo)
    assert abs(da - dao) <= 1e-8 * abs(dao)
    nf = c["np"]
    go = np.stack([gxo, gyo], 1)[:nf]
    assert np.linalg.norm(g[:nf] - go) <= 1e-8 * np.linalg.norm(go)


def test_adjoint_matches_central_differences():
    c, obj = adjoint_case()
    steps = 16
    banner("gpu adjoint vs central FD")
    J, da, g, _ = gpu_rollout(c, obj, steps, 4)
    a0 = c["av_alpha"]
    eps = 1e-4 * a0
    Jp = gpu_rollout(c, obj, steps, 4, alpha=a0 + eps)[0]
    Jm = gpu_rollout(c, obj, steps, 4, alpha=a0 - eps)[0]
    fd = (Jp - Jm) / (2 * eps)
    assert abs(fd - da) <= 1e-5 * abs(fd), (fd, da)
    rng = np.random.default_rng(1001)
    nf = c["np"]
    for k in range(3):
        d = np.zeros((c["n"], 2))
        d[:nf] = rng.normal(size=(nf, 2))
        e = 1e-6
        cp, cm = dict(c), dict(c)
        cp["vx"], cp["vy"] = c["vx"] + e * d[:, 0], c["vy"] + e * d[:, 1]
        cm["vx"], cm["vy"] = c["vx"] - e * d[:, 0], c["vy"] - e * d[:, 1]
        fd = (gpu_rollout(cp, obj, steps, 4)[0] - gpu_rollout(cm, obj, steps, 4)[0]) / (2 * e)
        ad = float(np.sum(g * d))
        assert abs(fd - ad) <= 1e-5 * abs(fd), (k, fd, ad)


@pytest.mark.parametrize("prec", [cases.FP64, cases.FP32])
def test_adjoint_independent_of_checkpoint_interval(prec):
    c, obj = adjoint_case(precision=prec)
    steps = 12
    banner(f"checkpoint interval prec={prec}")
    ref = gpu_rollout(c, obj, steps, 1)
    for ck in (5, 12):
        out = gpu_rollout(c, obj, steps, ck)
        assert out[0] == ref[0] and out[1] == ref[1]
        assert np.array_equal(out[2], ref[2])


def test_fp32_adjoint_against_fp64_differences():
    """The FP32 chosen form's adjoint (FP32 pair arithmetic, FP64 adjoint
    accumulators) against FP64 central differences."""
    c, obj = adjoint_case()
    c32, _ = adjoint_case(precision=cases.FP32)
    steps = 16
    banner("fp32 adjoint vs fp64 FD")
    J, da, g, _ = gpu_rollout(c32, obj, steps, 4)
    a0 = c["av_alpha"]
    e = 1e-4 * a0
    fd = (gpu_rollout(c, obj, steps, 4, alpha=a0 + e)[0]
          - gpu_rollout(c, obj, steps, 4, alpha=a0 - e)[0]) / (2 * e)
    assert abs(fd - da) <= 1e-3 * abs(fd), (fd, da)
    nf = c["np"]
    z = cases.normal01(1001, 2 * nf)
    d = np.zeros((c["n"], 2))
    d[:nf, 0], d[:nf, 1] = z[:nf], z[nf:]
    ev = 1e-6
    cp, cm = dict(c), dict(c)
    cp["vx"], cp["vy"] = c["vx"] + ev * d[:, 0], c["vy"] + ev * d[:, 1]
    cm["vx"], cm["vy"] = c["vx"] - ev * d[:, 0], c["vy"] - ev * d[:, 1]
    fdv = (gpu_rollout(cp, obj, steps, 4)[0] - gpu_rollout(cm, obj, steps, 4)[0]) / (2 * ev)
    ad = float(np.sum(g * d))
    banner(f"  alpha fd {fd:.6e} fp32 adj {da:.6e}; v0 fd {fdv:.6e} fp32 adj {ad:.6e}")
    assert abs(fdv - ad) <= 1e-3 * abs(fdv)


def test_adjoint_leaves_the_final_state():
    c, obj = adjoint_case()
    banner("final state")
    a = armando_amd.Simulation(c)
    a.adjoint_rollout(9, 4, obj)
    sa = a.state()
    b = armando_amd.Simulation(c)
    b.time_integration(1, 9)
    sb = b.state()
    for k in ("x", "y", "vx", "vy", "rho", "u"):
        assert np.array_equal(sa[k], sb[k]), k


def test_torch_autograd_function_backward_is_the_adjoint():
    import torch

    from armando_amd.autograd import RolloutLoss

    c, obj = adjoint_case()
    steps = 10
    banner("torch autograd")
    ref = gpu_rollout(c, obj, steps, 4)
    loss = RolloutLoss(c, obj, steps, 4)
    v0 = torch.tensor(np.stack([c["vx"], c["vy"]], 1), dtype=torch.float64, requires_grad=True)
    alpha = torch.tensor(c["av_alpha"], dtype=torch.float64, requires_grad=True)
    J = loss(v0, alpha)
    assert J.item() == ref[0]
    (3.0 * J).backward()
    assert alpha.grad.item() == pytest.approx(3.0 * ref[1], rel=1e-12)
    assert np.allclose(v0.grad.numpy(), 3.0 * ref[2], rtol=1e-12, atol=0)
    loss.close()


def test_torch_backward_of_an_earlier_forward():
    """Two forwards on one RolloutLoss, then the backward of the first: its
    gradients are those of the first forward's inputs, not of the second's
    (the context keeps one forward's checkpoints; the first is run again)."""
    import torch

    from armando_amd.autograd import RolloutLoss

    c, obj = adjoint_case()
    steps = 10
    banner("torch autograd: backward of an earlier forward")
    a1, a2 = c["av_alpha"], 3.0 * c["av_alpha"]
    ref1 = gpu_rollout(c, obj, steps, 4, alpha=a1)
    ref2 = gpu_rollout(c, obj, steps, 4, alpha=a2)
    assert ref1[1] != ref2[1]
    loss = RolloutLoss(c, obj, steps, 4)
    v0 = np.stack([c["vx"], c["vy"]], 1)
    va = torch.tensor(v0, dtype=torch.float64, requires_grad=True)
    vb = torch.tensor(1.5 * v0, dtype=torch.float64, requires_grad=True)
    al1 = torch.tensor(a1, dtype=torch.float64, requires_grad=True)
    al2 = torch.tensor(a2, dtype=torch.float64, requires_grad=True)
    J1 = loss(va, al1)
    J2 = loss(vb, al2)
    assert J1.item() == ref1[0]
    J1.backward()
    assert al1.grad.item() == pytest.approx(ref1[1], rel=1e-12)
    assert np.allclose(va.grad.numpy(), ref1[2], rtol=1e-12, atol=0)
    J2.backward()
    c2 = dict(c)
    c2["vx"], c2["vy"] = 1.5 * c["vx"], 1.5 * c["vy"]
    ref2b = gpu_rollout(c2, obj, steps, 4, alpha=a2)
    assert J2.item() == ref2b[0]
    assert al2.grad.item() == pytest.approx(ref2b[1], rel=1e-12)
    assert np.allclose(vb.grad.numpy(), ref2b[2], rtol=1e-12, atol=0)
    loss.close()


def test_adjoint_100k_400_steps_against_central_differences_on_16_directions():
    """The 100k adjoint workload at full size: 400 steps, checkpoint every 20.
    dJ/dalpha and dJ/dv0 . d for 16 directions d ~ N(0, 1) per fluid particle
    component (normal01 seeds 1001..1016) against central differences of the
    GPU forward with eps = 1e-4 (relative for alpha, m/s for v0); tolerance
    1e-3 relative."""
    c, obj = cases.dam_break_adjoint_100k()
    steps, ck = 400, 20
    banner(f"100k adjoint n={c['n']} steps={steps} ckpt={ck}")
    J, da, g, (tf, tb) = gpu_rollout(c, obj, steps, ck)
    assert J > 0 and tb <= 3.0 * tf, (J, tf, tb)
    # the FP32 chosen form, forward and backward, against the same FP64 differences
    c32, _ = cases.dam_break_adjoint_100k(precision=cases.FP32)
    J32, da32, g32, (tf32, tb32) = gpu_rollout(c32, obj, steps, ck)
    banner(f"  fp32: J {J32:.6e} (fp64 {J:.6e}), fwd {tf32:.1f} ms bwd {tb32:.1f} ms")
    assert abs(J32 - J) <= 1e-4 * J and tb32 <= 3.0 * tf32
    a0 = c["av_alpha"]
    e = 1e-4 * a0
    fd = (gpu_objective(c, obj, steps, ck, alpha=a0 + e)
          - gpu_objective(c, obj, steps, ck, alpha=a0 - e)) / (2 * e)
    assert abs(fd - da) <= 1e-3 * abs(fd), (fd, da)
    assert abs(fd - da32) <= 1e-3 * abs(fd), (fd, da32)
    nf = c["np"]
    e = 1e-4
    worst = worst32 = 0.0
    for seed in range(1001, 1017):
        z = cases.normal01(seed, 2 * nf)
        d = np.zeros((c["n"], 2))
        d[:nf, 0], d[:nf, 1] = z[:nf], z[nf:]
        cp, cm = dict(c), dict(c)
        cp["vx"], cp["vy"] = c["vx"] + e * d[:, 0], c["vy"] + e * d[:, 1]
        cm["vx"], cm["vy"] = c["vx"] - e * d[:, 0], c["vy"] - e * d[:, 1]
        fdv = (gpu_objective(cp, obj, steps, ck) - gpu_objective(cm, obj, steps, ck)) / (2 * e)
        ad = float(np.sum(g * d))
        rel = abs(fdv - ad) / abs(fdv)
        worst = max(worst, rel)
        assert rel <= 1e-3, (seed, fdv, ad)
        rel32 = abs(fdv - float(np.sum(g32 * d))) / abs(fdv)
        worst32 = max(worst32, rel32)
        assert rel32 <= 1e-3, (seed, fdv, float(np.sum(g32 * d)))
    banner(f"100k adjoint: alpha fd {fd:.6e} adj {da:.6e} fp32 {da32:.6e}; worst direction "
           f"rel {worst:.2e} (fp32 {worst32:.2e})")


@pytest.mark.parametrize("h_fac", [1.6, 2.6])
def test_adjoint_with_incomplete_masks_matches_oracle(h_fac):
    """Rows of more than 32 candidates leave the forward's neighbour masks
    incomplete: the reverse gathers then test the candidates again (h = 1.6
    dx mixes both walks, h = 2.6 dx takes the second everywhere)."""
    s = 0.02
    c = cases.dam_break(col_w=0.6 * s, col_h=1.2 * s, tank_w=3.2 * s, tank_h=1.6 * s, dx=1e-3,
                        h_fac=h_fac)
    rng = np.random.default_rng(5)
    nf = c["np"]
    c["vx"] = c["vx"].copy()
    c["vy"] = c["vy"].copy()
    c["vx"][:nf] = rng.normal(0.0, 0.01, nf)
    c["vy"][:nf] = rng.normal(0.0, 0.01, nf)
    obj = np.nonzero((c["x"] < 0) & (c["y"] < 0.2 * 1.2 * s))[0].astype(np.int32)
    steps = 12
    banner(f"gpu vs oracle adjoint, h = {h_fac} dx, n={c['n']}")
    J, da, g, _ = gpu_rollout(c, obj, steps, 5)
    P = orc.params_from(c)
    st = orc.OracleState(c["n"], c["np"])
    st.set(**{k: c[k] for k in ("x", "y", "vx", "vy", "rho", "u")})
    mask = np.zeros(c["n"], np.int32)
    mask[obj] = 1
    Jo, dao, gxo, gyo = orc.adjoint(P, st, steps, mask)
    assert abs(J - Jo) <= 1e-10 * abs(Jo)
    assert abs(da - dao) <= 1e-8 * abs(dao)
    go = np.stack([gxo, gyo], 1)[:nf]
    assert np.linalg.norm(g[:nf] - go) <= 1e-8 * np.linalg.norm(go)
