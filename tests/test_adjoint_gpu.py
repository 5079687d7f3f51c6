"""GPU: the differentiable form. The device rollout (world forward with
recorded contact sets, checkpoints every 16 steps, recomputed segments and
the Dy adjoint in reverse) behind a torch.autograd.Function, against the
FP64 CPU reference gradient of the same frozen-contact function (built
from the device's recorded contact sets), whose own adjoint is checked
against central differences on the CPU.

The reference freezes each step's rows along the FP32 restatement's
trajectory from the same initial records, which is the device's trajectory
bit for bit (the recorded inputs replay the device world exactly), so both
sides differentiate the same function. An FP64 nominal trajectory is not a
usable reference here: the Dy step keeps or resets friction anchors on
pose-dependent tests, and on these three stacks one such decision goes the
other way in FP64 within 20 steps (final positions then differ by 1e-3).

Bars: final positions within 1e-4 of the FP64 evaluation (FP32 PGS and
integration rounding over 20 steps); gradient vector (v0, w0 of every body
and the per-scene friction) within 1e-3 relative, a row whose clamp sits
within rounding of its threshold may take the other branch in FP64 and the
vector norm absorbs such isolated rows."""
import numpy as np
import pytest

from fyzz_amd import scenes

pytestmark = pytest.mark.gpu


def test_rollout_gradient_matches_fp64_reference(ctx, orc, torch_cuda):
    import torch
    from fyzz_amd.diff import Rollout, rollout_function
    from fyzz_amd.dy import DY_BODY, DY_MANAGER, DY_PATCH, from_device
    from oracle_bind import OracleDiff
    counts, half, pos, quat, v0, w0, mu, target = scenes.diff_scenes(3)
    steps = 20
    ro = Rollout(ctx, half, pos, quat, counts, steps, checkpoint_every=16)
    fn = rollout_function()
    dev = torch.device("cuda")
    tv = torch.tensor(v0, device=dev, requires_grad=True)
    tw = torch.tensor(w0, device=dev, requires_grad=True)
    tm = torch.tensor(mu, device=dev, requires_grad=True)
    tgt = torch.tensor(target, dtype=torch.float32, device=dev)
    p = fn.apply(tv, tw, tm, ro)
    scene = torch.as_tensor(ro.scene_of_body, device=dev)
    com = torch.zeros((len(counts), 3), device=dev).index_add_(0, scene, p) / 64.0
    d = (com - tgt)[:, [0, 2]]
    loss = (d * d).sum()
    loss.backward()
    gv, gw, gm = tv.grad.cpu().numpy(), tw.grad.cpu().numpy(), tm.grad.cpu().numpy()
    # the FP64 reference over the same recorded contact sets
    b0 = from_device(ro.checkpoints[0], DY_BODY)
    mgs = [from_device(m, DY_MANAGER) if len(m) else np.zeros(0, DY_MANAGER) for m, _ in ro.records]
    hs = [from_device(h, DY_PATCH) if len(h) else np.zeros(0, DY_PATCH) for _, h in ro.records]
    st = np.zeros(1, [("position", "<f4", 3), ("orientation", "<f4", 4), ("static_friction", "<f4"),
                      ("dynamic_friction", "<f4"), ("restitution", "<f4")])
    st["orientation"][0, 3] = 1.0
    st["static_friction"] = st["dynamic_friction"] = 0.5
    ref = OracleDiff(orc.lib, b0, st, mgs, hs, 1 / 120, nominal_fp32=True)
    mus = np.repeat(mu.astype(np.float64), counts)
    pf = ref.eval(v0.astype(np.float64), w0.astype(np.float64), mus)
    np.testing.assert_allclose(p.detach().cpu().numpy(), pf, rtol=0, atol=1e-4)
    com_r = np.stack([pf[ro.scene_of_body == k].mean(0) for k in range(len(counts))])
    dr = com_r - target
    dr[:, 1] = 0.0
    adj = 2.0 * dr[ro.scene_of_body] / 64.0
    rv, rw, rm, _ = ref.grad(v0.astype(np.float64), w0.astype(np.float64), mus, adj)
    rm_scene = np.array([rm[ro.scene_of_body == k].sum() for k in range(len(counts))])
    g = np.concatenate([gv.ravel(), gw.ravel(), gm])
    r = np.concatenate([rv.ravel(), rw.ravel(), rm_scene])
    rel = np.linalg.norm(g - r) / np.linalg.norm(r)
    print(f"loss {loss.item():.5g}; gradient norm {np.linalg.norm(r):.4g}; device vs FP64 relative {rel:.3g}; "
          f"d/dmu {gm} vs {rm_scene}")
    assert rel < 1e-3


def _gradients(ctx, n_scenes, steps):
    """Loss on the stacks' final COM x/z against their targets; returns the
    positions and the gradients w.r.t. v0, w0 and the per-scene friction."""
    import torch
    from fyzz_amd.diff import Rollout, rollout_function
    counts, half, pos, quat, v0, w0, mu, target = scenes.diff_scenes(n_scenes)
    ro = Rollout(ctx, half, pos, quat, counts, steps, checkpoint_every=16)
    try:
        dev = torch.device("cuda")
        tv = torch.tensor(v0, device=dev, requires_grad=True)
        tw = torch.tensor(w0, device=dev, requires_grad=True)
        tm = torch.tensor(mu, device=dev, requires_grad=True)
        tgt = torch.tensor(target, dtype=torch.float32, device=dev)
        p = rollout_function().apply(tv, tw, tm, ro)
        com = p.view(n_scenes, 64, 3).sum(1) / 64.0  # a fixed-order reduction
        d = (com - tgt)[:, [0, 2]]
        (d * d).sum().backward()
        return (p.detach().cpu().numpy(), tv.grad.cpu().numpy(), tw.grad.cpu().numpy(), tm.grad.cpu().numpy())
    finally:
        ro.close()


def test_full_size_rollout_gradients_are_reproducible(ctx, torch_cuda):
    """The differentiable workload at full size (1,024 stacks of 64 boxes,
    120 steps, checkpoints every 16): two runs give bit-identical positions
    and gradients (the adjoint sums each body's row terms in a fixed order),
    the gradients are finite, and the first three stacks' positions and
    gradients equal, bit for bit, those of a rollout of those three stacks
    alone (stacks never interact; the three-stack form is checked against
    the FP64 reference above)."""
    a = _gradients(ctx, 1024, 120)
    b = _gradients(ctx, 1024, 120)
    for x, y, name in zip(a, b, ("positions", "d/dv0", "d/dw0", "d/dmu")):
        assert np.isfinite(x).all(), name
        np.testing.assert_array_equal(x, y, err_msg=f"{name} differ between two runs")
    small = _gradients(ctx, 3, 120)
    k = 3 * 64
    for x, y, name in zip(a[:3], small[:3], ("positions", "d/dv0", "d/dw0")):
        np.testing.assert_array_equal(x[:k], y, err_msg=f"{name}: first three stacks of the full batch")
    np.testing.assert_array_equal(a[3][:3], small[3], err_msg="d/dmu: first three stacks of the full batch")
    assert np.linalg.norm(a[1]) > 0 and np.linalg.norm(a[3]) > 0
