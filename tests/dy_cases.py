"""Seeded synthetic Dy inputs for parity tests (not physical scenes: random
contact sets that exercise every branch the admitted class has)."""
from __future__ import annotations

import numpy as np

from fyzz_amd import scenes
from fyzz_amd.dy import DY_MANAGER, DY_STATIC, default_bodies


def random_case(seed: int, nb: int = 2048, n_dd: int = 4000, hub: int = 70, n_static: int = 1500,
                hubs: int = 1):
    """Bodies with random boxes, poses and velocities; n_dd dynamic/dynamic
    managers (body pairs drawn at random), `hub` extra managers on each of
    bodies 0..hubs-1 (more than 32: forces overflow waves) and n_static
    dynamic/static managers against two statics. Each manager has 1-4 contacts sharing
    one normal; separations in [-0.03, 0.03]."""
    rng = np.random.default_rng(seed)
    half = (0.2 + 0.8 * scenes.uniform(seed, 0, 3 * nb).reshape(nb, 3)).astype(np.float32)
    inv_mass, inv_inertia = scenes.box_mass_properties(half, 1000.0)
    b = default_bodies(nb)
    b["position"] = (rng.random((nb, 3)) * 40.0).astype(np.float32)
    q = scenes.normalize_quat32(scenes.random_quat(rng.random(nb), rng.random(nb), rng.random(nb)))
    b["orientation"] = q
    b["linear_velocity"] = (rng.normal(size=(nb, 3)) * 0.5).astype(np.float32)
    b["angular_velocity"] = (rng.normal(size=(nb, 3)) * 0.5).astype(np.float32)
    b["inverse_mass"] = inv_mass
    b["inverse_inertia"] = inv_inertia
    b["restitution"] = np.where(rng.random(nb) < 0.5, 0.0, 0.3).astype(np.float32)
    b["counted_interactions"] = rng.integers(0, 8, nb).astype(np.uint32)
    b["wake_counter"] = np.where(rng.random(nb) < 0.3, 0.1, 0.4).astype(np.float32)
    st = np.zeros(2, DY_STATIC)
    st["orientation"][:, 3] = 1.0
    st["static_friction"] = [0.5, 0.8]
    st["dynamic_friction"] = [0.5, 0.6]
    st["position"][1] = [5.0, -1.0, 5.0]
    pairs = []
    while len(pairs) < n_dd:
        a, c = rng.integers(hubs, nb, 2)
        if a != c:
            pairs.append((int(a), int(c)))
    for h in range(hubs):
        pairs += [(h, int(k)) for k in rng.choice(np.arange(hubs, nb), hub, replace=False)]
    pairs += [(int(rng.integers(0, nb)), -1 - int(rng.integers(0, 2))) for _ in range(n_static)]
    order = rng.permutation(len(pairs))
    mg = np.zeros(len(pairs), DY_MANAGER)
    for i, k in enumerate(order):
        b0, b1 = pairs[k]
        if rng.random() < 0.5 and b1 >= 0:
            b0, b1 = b1, b0
        cnt = int(rng.integers(1, 5))
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        base = b["position"][b0] if b0 >= 0 else b["position"][b1]
        mg[i]["body0"], mg[i]["body1"], mg[i]["count"], mg[i]["history"] = b0, b1, cnt, 0xFFFFFFFF
        for c in range(cnt):
            mg[i]["contacts"][c]["position"] = base + rng.normal(size=3) * 0.4
            mg[i]["contacts"][c]["normal"] = n
            mg[i]["contacts"][c]["separation"] = rng.uniform(-0.03, 0.03)
    return b, st, mg


def with_history(mg, patches):
    """The same managers pointing at last step's patch of each manager."""
    m2 = mg.copy()
    m2["history"] = np.arange(len(mg), dtype=np.uint32)
    return m2, patches
