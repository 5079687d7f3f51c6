"""CPU: the FP64 frozen-row rollout of the differentiable form (the
reference gradient for the device adjoint) against central finite
differences of itself.

Setup (the differentiable workload, one scene): a 4x4x4 stack of 64 boxes
(h = 0.5, spacing 1.05, bottom layer on the plane), seeded v0, w0 ~
N(0, 0.5), friction mu ~ U(0.2, 0.8) for the scene's boxes (floor 0.5),
20 steps at 1/120 s. The contact sets are those of the oracle pipeline's
own run (frozen per step: rows fixed, velocities, PGS and positions free).
Loss = ||(COM - target)_xz||^2. The frozen-row function is piecewise
linear (the PGS clamps), so a finite difference straddling a clamp's kink
differs locally from the one-sided derivative. Bars: with eps 1e-3, the
gradient vector within 1e-3 relative (norm); with eps 1e-6 (inside the
linear pieces), every component above 1% of the largest within 1e-4."""
import numpy as np

from fyzz_amd import scenes
from oracle_bind import OracleDiff, OracleWorld


def stack_case(steps=20, seed=scenes.SEED_BASE + 4):
    sc = scenes.stack_scene(seed)
    n = sc.n
    v0 = np.stack([scenes.normal(seed, 10 + a, n) * 0.5 for a in range(3)], 1).astype(np.float32)
    w0 = np.stack([scenes.normal(seed, 13 + a, n) * 0.5 for a in range(3)], 1).astype(np.float32)
    mu = np.float32(0.2 + 0.6 * scenes.uniform(seed, 16, 1)[0])
    target = np.array([(scenes.uniform(seed, 17, 1)[0] - 0.5) * 4.0, 0.0, (scenes.uniform(seed, 18, 1)[0] - 0.5) * 4.0])
    sc.vel, sc.ang = v0, w0
    return sc, mu, target, steps


def record(orc, sc, mu, steps):
    ow = OracleWorld(orc, sc)
    ow.bodies["static_friction"] = mu
    ow.bodies["dynamic_friction"] = mu
    b0 = ow.bodies.copy()
    mgs, hists = [], []
    for _ in range(steps):
        ow.step(1 / 120)
        _, mg, hist = ow.last_dy_inputs
        mgs.append(mg)
        hists.append(hist)
    b0["counted_interactions"] = 0
    return b0, ow.statics, mgs, hists


def loss_and_adj(p, target):
    com = p.mean(axis=0)
    d = com - target
    d[1] = 0.0
    adj = np.tile(2.0 * d / len(p), (len(p), 1))
    return float(d @ d), adj


def test_adjoint_matches_central_differences(orc):
    sc, mu, target, steps = stack_case()
    b0, statics, mgs, hists = record(orc, sc, mu, steps)
    assert sum(len(m) for m in mgs) > 0
    diff = OracleDiff(orc.lib, b0, statics, mgs, hists, 1 / 120)
    v0 = b0["linear_velocity"].astype(np.float64)
    w0 = b0["angular_velocity"].astype(np.float64)
    mus = np.full(sc.n, float(mu))
    p = diff.eval(v0, w0, mus)
    _, adj = loss_and_adj(p, target)
    gv, gw, gm, pf = diff.grad(v0, w0, mus, adj)
    np.testing.assert_array_equal(pf, p)

    def L(v, w, m):
        return loss_and_adj(diff.eval(v, w, m), target)[0]

    def central(eps):
        fd_v = np.zeros_like(v0)
        fd_w = np.zeros_like(w0)
        for i in range(sc.n):
            for k in range(3):
                e = np.zeros_like(v0)
                e[i, k] = eps
                fd_v[i, k] = (L(v0 + e, w0, mus) - L(v0 - e, w0, mus)) / (2 * eps)
                fd_w[i, k] = (L(v0, w0 + e, mus) - L(v0, w0 - e, mus)) / (2 * eps)
        fd_mu = (L(v0, w0, mus + eps) - L(v0, w0, mus - eps)) / (2 * eps)
        return np.concatenate([fd_v.ravel(), fd_w.ravel(), [fd_mu]])

    g = np.concatenate([gv.ravel(), gw.ravel(), [gm.sum()]])
    f3 = central(1e-3)
    rel = np.linalg.norm(g - f3) / np.linalg.norm(f3)
    f6 = central(1e-6)
    big = np.abs(f6) > 0.01 * np.abs(f6).max()
    comp = np.abs(g[big] - f6[big]) / np.abs(f6[big])
    print(f"gradient norm {np.linalg.norm(f3):.4g}; eps 1e-3: relative error {rel:.3g}; eps 1e-6: worst component "
          f"{comp.max():.3g} over {int(big.sum())}; d/dmu {gm.sum():.6g} vs {f6[-1]:.6g}")
    assert rel < 1e-3
    assert comp.max() < 1e-4


def test_frozen_rollout_reproduces_its_nominal_trajectory(orc):
    """At the recorded inputs the frozen-row function is the nominal
    trajectory itself: rows frozen along the FP32 restatement's run give
    that run's final positions up to FP64-vs-FP32 arithmetic in the PGS and
    integration (a dropped row term, a wrong lever arm or a widened field
    left out moves them by far more than the 1e-5 bar); rows frozen along
    the FP64 twin's run give the same to the same bar here (this scene
    takes no pose-dependent anchor decision differently in the two
    precisions)."""
    sc, mu, _, steps = stack_case()
    ow = OracleWorld(orc, sc)
    ow.bodies["static_friction"] = mu
    ow.bodies["dynamic_friction"] = mu
    b0 = ow.bodies.copy()
    mgs, hists = [], []
    for _ in range(steps):
        ow.step(1 / 120)
        _, mg, hist = ow.last_dy_inputs
        mgs.append(mg)
        hists.append(hist)
    p32 = ow.bodies["position"].astype(np.float64)
    b0["counted_interactions"] = 0
    args = (b0["linear_velocity"].astype(np.float64), b0["angular_velocity"].astype(np.float64),
            np.full(sc.n, float(mu)))
    for fp32 in (True, False):
        diff = OracleDiff(orc.lib, b0, ow.statics, mgs, hists, 1 / 120, nominal_fp32=fp32)
        gap = np.abs(diff.eval(*args) - p32).max()
        print(f"nominal {'FP32' if fp32 else 'FP64'}: final positions within {gap:.3g} of the FP32 run")
        assert gap < 1e-5
