"""Seeded synthetic workloads for the FyzzXYZZY step pipeline.

One counter-based generator feeds every workload, so the device path, the
CPU oracle and the compiled reference receive identical inputs:

    draw(seed, stream, i) = splitmix64(seed * 0x9E3779B97F4A7C15
                                       + stream * 0xD1B54A32D192ED03 + i)
    uniform in [0,1)      = (draw >> 11) * 2**-53
    normal                = Box-Muller on two uniforms of consecutive
                            streams (2s, 2s+1) with the same index

Values are computed in float64 and rounded once to float32. Seeds are
0xF7220000 + config index. Each function documents which stream holds
which draw.

Workloads (sizes as the reference's scene authoring calls take them):
  grid_scene      1,000 boxes on a 10x10x10 grid over a floor plane
  sap_scene       N axis-aligned boxes in a cube, lognormal half extents,
                  plus per-frame moves of a seeded 10% subset
  pile_scene      N boxes with uniform half extents dropped in a column
  stack_scene     4x4x4 stacks of 64 boxes per scene
"""
from __future__ import annotations

import math

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
SEED_BASE = 0xF7220000


def _splitmix64(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.uint64, copy=True)
    with np.errstate(over="ignore"):
        x += np.uint64(0x9E3779B97F4A7C15)
        z = x
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def draw(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n raw 64-bit draws of one stream."""
    with np.errstate(over="ignore"):
        base = (np.uint64(seed % 2**64) * np.uint64(0x9E3779B97F4A7C15)
                + np.uint64(stream % 2**64) * np.uint64(0xD1B54A32D192ED03))
        idx = np.arange(offset, offset + n, dtype=np.uint64) + base
    return _splitmix64(idx)


def uniform(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    return (draw(seed, stream, n, offset) >> np.uint64(11)).astype(np.float64) * 2.0**-53


def normal(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    u1 = uniform(seed, 2 * stream, n, offset)
    u2 = uniform(seed, 2 * stream + 1, n, offset)
    r = np.sqrt(-2.0 * np.log1p(-u1))  # 1-u1 in (0,1]
    return r * np.cos(2.0 * math.pi * u2)


def f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).astype(np.float32))


def yaw_quat(theta: np.ndarray) -> np.ndarray:
    """Quaternions (x,y,z,w) of a rotation by theta about +y."""
    h = 0.5 * theta
    q = np.zeros((theta.size, 4))
    q[:, 1] = np.sin(h)
    q[:, 3] = np.cos(h)
    return q


def random_quat(u1, u2, u3) -> np.ndarray:
    """Shoemake's uniform random rotation from three uniforms, (x,y,z,w)."""
    a = np.sqrt(1.0 - u1)
    b = np.sqrt(u1)
    t1 = 2.0 * math.pi * u2
    t2 = 2.0 * math.pi * u3
    q = np.stack([a * np.sin(t1), a * np.cos(t1), b * np.sin(t2), b * np.cos(t2)], axis=1)
    return q


def normalize_quat32(q: np.ndarray) -> np.ndarray:
    """Round to float32 and renormalise the way World::set_body_pose does
    (divide each component by the float32 length)."""
    q32 = f32(q)
    ln = np.sqrt((q32.astype(np.float32) ** 2).sum(axis=1, dtype=np.float32)).astype(np.float32)
    return np.ascontiguousarray(q32 / ln[:, None])


class Scene:
    """Authoring record of one scene: a floor (optional) then dynamic boxes
    in body-id order, exactly as add_static_floor / add_dynamic_box /
    set_body_pose / set_body_velocity / set_body_material would author it."""

    def __init__(self, half, pos, quat, vel=None, ang=None, density=1000.0,
                 floor=True, floor_height=0.0, friction=0.5, restitution=0.0,
                 gravity=True, sphere=None):
        n = len(half)
        self.n = n
        # per body: a sphere of radius half[i, 0] instead of a box
        self.sphere = np.zeros(n, bool) if sphere is None else np.asarray(sphere, bool).reshape(n)
        # static bodies created after the dynamic ones: (half extents or
        # radius, position, orientation, is_sphere)
        self.statics = []
        self.half = f32(half).reshape(n, 3)
        self.pos = f32(pos).reshape(n, 3)
        self.quat = np.ascontiguousarray(np.asarray(quat, dtype=np.float32).reshape(n, 4))
        self.vel = f32(np.zeros((n, 3)) if vel is None else vel).reshape(n, 3)
        self.ang = f32(np.zeros((n, 3)) if ang is None else ang).reshape(n, 3)
        self.density = float(density)
        self.floor = bool(floor)
        self.floor_height = float(floor_height)
        self.friction = float(friction)
        self.restitution = float(restitution)
        self.gravity = bool(gravity)

    @property
    def first_box_id(self) -> int:
        return 1 if self.floor else 0


def grid_scene(seed: int = SEED_BASE + 0) -> Scene:
    """1,000 boxes, h=0.5, on a 10x10x10 grid with 1.2 m spacing from y=1.
    Box (i,j,k) -> x=(i-4.5)*1.2, y=1+1.2j, z=(k-4.5)*1.2, body order
    i + 10k + 100j. Stream 0: yaw uniforms; streams 1-3: velocity
    normals (sigma 0.1) per axis."""
    n = 1000
    idx = np.arange(n)
    i = idx % 10
    k = (idx // 10) % 10
    j = idx // 100
    pos = np.stack([(i - 4.5) * 1.2, 1.0 + j * 1.2, (k - 4.5) * 1.2], axis=1)
    yaw = uniform(seed, 0, n) * 2.0 * math.pi
    quat = normalize_quat32(yaw_quat(yaw))
    vel = np.stack([normal(seed, 1 + a, n) * 0.1 for a in range(3)], axis=1)
    half = np.full((n, 3), 0.5)
    return Scene(half, pos, quat, vel=vel, friction=0.5, restitution=0.0)


def sap_scene(n: int, seed: int = SEED_BASE + 1, extent: float | None = None):
    """Axis-aligned boxes for the broadphase-only workload.
    Centers uniform in a cube of side `extent` (1000 m at n=1e6; scaled by
    (n/1e6)^(1/3) otherwise so the density matches). Half extents per axis
    lognormal(ln 0.5, 0.5). Streams 0-2: center uniforms; 3-5: half-extent
    normals. Returns (half[n,3], pos[n,3]) float32."""
    if extent is None:
        extent = 1000.0 * (n / 1.0e6) ** (1.0 / 3.0)
    pos = np.stack([uniform(seed, a, n) * extent for a in range(3)], axis=1)
    half = np.stack([np.exp(math.log(0.5) + 0.5 * normal(seed, 3 + a, n)) for a in range(3)], axis=1)
    return f32(half), f32(pos)


def sap_moves(n: int, frame: int, seed: int = SEED_BASE + 1, fraction: float = 0.1):
    """The seeded subset moved in `frame` (1-based) and its displacements.
    The subset is the n*fraction slots with the smallest draws of stream
    100+frame; displacements are normals (sigma 0.05 m) of streams
    (1000+3*frame+axis), indexed by subset rank. Returns (slots ascending,
    delta[m,3] float64)."""
    m = int(n * fraction)
    keys = draw(seed, 100 + frame, n)
    slots = np.sort(np.argpartition(keys, m - 1)[:m]) if m < n else np.arange(n)
    delta = np.stack([normal(seed, 1000 + 3 * frame + a, m) * 0.05 for a in range(3)], axis=1)
    return slots.astype(np.int64), delta


def _pile_centers_py(seed, radius, footprint, height):
    """The placement loop of pile_scene in Python for radii in placement
    order (small n; the native helper reproduces it for large n)."""
    n = len(radius)
    cell = 2.0 * float(np.sqrt(3.0))  # >= 2 * max radius
    grid: dict = {}
    pos = np.zeros((n, 3))
    attempt = 0
    batch = 4096
    cache = None
    cache_base = -1
    for b in range(n):
        while True:
            if cache is None or attempt - cache_base >= batch:
                cache_base = attempt
                cache = np.stack([uniform(seed, 3 + a, batch, attempt) for a in range(3)], axis=1)
            u = cache[attempt - cache_base]
            attempt += 1
            c = np.array([(u[0] - 0.5) * footprint, 1.0 + u[1] * height, (u[2] - 0.5) * footprint])
            key = tuple(np.floor(c / cell).astype(int))
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        for o in grid.get((key[0] + dx, key[1] + dy, key[2] + dz), ()):
                            d = c - pos[o]
                            if d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < (radius[b] + radius[o]) ** 2:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok and c[1] - radius[b] > 0.0:
                pos[b] = c
                grid.setdefault(key, []).append(b)
                break
    return pos


def _pile_centers_native(seed, radius, footprint, height):
    import ctypes as C
    from pathlib import Path
    lib = C.CDLL(str(Path(__file__).resolve().parent / "libfyzz_scenes.so"))
    n = len(radius)
    r = np.ascontiguousarray(radius, np.float64)
    pos = np.zeros((n, 3))
    used = C.c_uint64()
    rc = lib.fyzz_scene_pile_centers(C.c_uint64(seed), C.c_uint32(n), r.ctypes.data_as(C.c_void_p),
                                     C.c_double(footprint), C.c_double(height), pos.ctypes.data_as(C.c_void_p),
                                     C.c_uint64(1 << 31), C.byref(used))
    if rc != 0:
        raise RuntimeError("pile placement did not finish")
    return pos


def pile_scene(n: int, seed: int = SEED_BASE + 2, footprint: float = 100.0,
               height: float = 200.0, native: bool | None = None) -> Scene:
    """Boxes with half extents U[0.2,1.0] per axis dropped from a
    footprint x footprint x height column above the floor, uniform random
    orientation (Shoemake). Boxes are placed in order of decreasing
    bounding-sphere radius (stable in body order); each draws centers until
    its bounding sphere overlaps no placed one (a cell grid holds the placed
    spheres) and stays above the floor. Largest-first placement reaches the
    30% sphere volume fraction of 100k boxes in this column in under a
    million draws (body order would need over 10^8). Streams 0-2 half
    extents; 3-5 center uniforms with the global attempt counter as index;
    6-8 orientation uniforms. The placement loop runs natively
    (libfyzz_scenes.so) unless native=False or the helper is not built."""
    half = np.stack([0.2 + 0.8 * uniform(seed, a, n) for a in range(3)], axis=1)
    radius = np.sqrt((half ** 2).sum(axis=1))
    order = np.argsort(-radius, kind="stable")
    if native is None:
        from pathlib import Path
        native = (Path(__file__).resolve().parent / "libfyzz_scenes.so").exists()
    placed = (_pile_centers_native if native else _pile_centers_py)(seed, radius[order], footprint, height)
    pos = np.zeros((n, 3))
    pos[order] = placed
    quat = normalize_quat32(random_quat(uniform(seed, 6, n), uniform(seed, 7, n), uniform(seed, 8, n)))
    return Scene(half, pos, quat, friction=0.5, restitution=0.0)


def stack_scene(seed: int = SEED_BASE + 4, spacing: float = 1.05):
    """One 4x4x4 stack of 64 boxes (h=0.5), bottom layer resting on the
    plane (center y=0.5), layer pitch `spacing`. Index i + 4k + 16j."""
    idx = np.arange(64)
    i = idx % 4
    k = (idx // 4) % 4
    j = idx // 16
    pos = np.stack([(i - 1.5) * spacing, 0.5 + j * spacing, (k - 1.5) * spacing], axis=1)
    half = np.full((64, 3), 0.5)
    quat = np.tile(np.array([0.0, 0.0, 0.0, 1.0]), (64, 1))
    return Scene(half, pos, quat, friction=0.5, restitution=0.0)


def box_mass_properties(half: np.ndarray, density: float):
    """Inverse mass and local inverse inertia of boxes, in the float32
    operation order of Scene::add_dynamic_box (reference
    src/native/scene.cpp:408-475): mass in double, then the unit-density
    PxMassProperties tensor scaled by density in float32."""
    h = np.asarray(half, np.float32).reshape(-1, 3)
    hx, hy, hz = h[:, 0], h[:, 1], h[:, 2]
    mass = 8.0 * hx.astype(np.float64) * hy * hz * float(density)
    inv_mass = (1.0 / mass).astype(np.float32)
    f = np.float32
    unit_mass = (hx * hy * hz) * f(8.0)
    ex, ey, ez = hx * hx, hy * hy, hz * hz
    unit_scale = (unit_mass * f(1.0)) / f(3.0)
    d = f(density)
    inertia = np.stack([((ey + ez) * unit_scale) * d, ((ex + ez) * unit_scale) * d, ((ex + ey) * unit_scale) * d], 1)
    inv_inertia = (f(1.0) / inertia).astype(np.float32)
    return inv_mass, inv_inertia


def batch_scenes(n_scenes: int = 4096, seed: int = SEED_BASE + 3, min_boxes: int = 8, max_boxes: int = 512,
                 height: float = 200.0, footprint_per_100k: float = 100.0):
    """Independent piles for the batched form: scene k has
    n_k = min_boxes + floor(u * (max_boxes - min_boxes + 1)) boxes (stream 0
    of `seed`, index k), each a pile_scene of seed (seed << 20) + k with the
    footprint scaled by sqrt(n_k / 100,000) and the column height kept.
    Returns (counts, half, pos, quat) concatenated in scene order."""
    u = uniform(seed, 0, n_scenes)
    counts = (min_boxes + np.floor(u * (max_boxes - min_boxes + 1))).astype(np.int64)
    halves, poss, quats = [], [], []
    for k, n in enumerate(counts):
        fp = footprint_per_100k * math.sqrt(n / 100000.0)
        sc = pile_scene(int(n), seed=(seed << 20) + k, footprint=fp, height=height)
        halves.append(sc.half)
        poss.append(sc.pos)
        quats.append(sc.quat)
    return (counts.astype(np.uint32), np.concatenate(halves), np.concatenate(poss), np.concatenate(quats))


def diff_scenes(n_scenes: int = 1024, seed: int = SEED_BASE + 4):
    """The differentiable workload: n_scenes independent 4x4x4 stacks of 64
    boxes (h = 0.5, spacing 1.05, bottom layer on the plane; scenes never
    interact, so every stack sits at the origin). Per body v0, w0 ~
    N(0, 0.5) per axis (streams 10-12 and 13-15, index = global body);
    per scene friction mu ~ U(0.2, 0.8) (stream 16) and target (x, z) ~
    U(-2, 2) (streams 17, 18). Returns (counts, half, pos, quat, v0, w0,
    mu, target[n_scenes, 3] with y = 0)."""
    st = stack_scene(seed)
    n = n_scenes * st.n
    counts = np.full(n_scenes, st.n, np.uint32)
    half = np.tile(st.half, (n_scenes, 1))
    pos = np.tile(st.pos, (n_scenes, 1))
    quat = np.tile(st.quat, (n_scenes, 1))
    v0 = f32(np.stack([normal(seed, 10 + a, n) * 0.5 for a in range(3)], 1))
    w0 = f32(np.stack([normal(seed, 13 + a, n) * 0.5 for a in range(3)], 1))
    mu = f32(0.2 + 0.6 * uniform(seed, 16, n_scenes))
    target = np.zeros((n_scenes, 3))
    target[:, 0] = (uniform(seed, 17, n_scenes) - 0.5) * 4.0
    target[:, 2] = (uniform(seed, 18, n_scenes) - 0.5) * 4.0
    return counts, half, pos, quat, v0, w0, mu, target
