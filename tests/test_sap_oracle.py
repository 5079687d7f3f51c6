"""CPU: the SAP oracle against the reference's golden vectors and against
what the mathematics fixes (order preservation of the key encoding,
brute-force O(n^2) overlap, set differences)."""
from pathlib import Path

import numpy as np
import pytest

from fyzz_amd import scenes

GOLDEN = Path(__file__).resolve().parent / "golden"


def _f32bits(x):
    return np.frombuffer(np.float32(x).tobytes(), np.uint32)[0]


def test_encoding_order_and_nibbles(orc):
    # Boxes of zero rotation and tiny half extents around chosen centres:
    # min key nibble 0, max key nibble 4 (scene.cpp:29-34), and the keys
    # are monotone in the float value, across the sign boundary and zeros.
    centers = np.array([-1e6, -3.5, -1.0, -1e-30, -0.0, 0.0, 1e-30, 1.0, 3.5, 1e6], np.float64)
    n = len(centers)
    pos = np.zeros((n, 3), np.float32)
    pos[:, 0] = centers
    half = np.full((n, 3), 1e-3, np.float32)
    quat = np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1))
    off = np.zeros(n, np.float32)
    b = orc.box_bounds(pos, quat, half, off)
    assert np.all(b[0:3] & 0xF == 0)
    assert np.all(b[3:6] & 0xF == 4)
    assert np.all(b[0] < b[3])
    order = np.argsort(pos[:, 0] - half[:, 0], kind="stable")
    assert np.all(np.diff(b[0][order].astype(np.int64)) >= 0)
    # the encoding of a positive float sets the sign bit; a negative one is
    # bit-inverted (order-preserving): check one value of each by hand
    one = _f32bits(1.0 - 1e-3)
    assert b[0][7] == ((((one | 0x80000000) >> 4) - 1) << 4) & 0xFFFFFFFF
    neg = _f32bits(-1.0 - 1e-3)
    assert b[0][2] == ((((~neg & 0xFFFFFFFF) >> 4) - 1) << 4) & 0xFFFFFFFF


def test_nonfinite_bounds_rejected(orc):
    pos = np.array([[np.inf, 0, 0]], np.float32)
    with pytest.raises(OverflowError):
        orc.box_bounds(pos, np.array([[0, 0, 0, 1]], np.float32), np.ones((1, 3), np.float32),
                       np.zeros(1, np.float32))


def _brute_pairs(bounds, body, dynamic):
    n = bounds.shape[1]
    b = bounds.astype(np.int64)
    out = []
    for i in range(n):
        ov = ~((b[0, i] > b[3]) | (b[0] > b[3, i]) | (b[1, i] > b[4]) | (b[1] > b[4, i]) |
               (b[2, i] > b[5]) | (b[2] > b[5, i]))
        for j in np.nonzero(ov)[0]:
            if j > i and body[i] != body[j] and (dynamic[i] or dynamic[j]):
                out.append((i, j))
    return np.array(sorted(out), np.uint32).reshape(-1, 2)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_sweep_equals_bruteforce(orc, seed):
    n = 400
    half, pos = scenes.sap_scene(n, seed=seed, extent=2.0 * n ** (1 / 3))
    quat = scenes.normalize_quat32(scenes.random_quat(scenes.uniform(seed, 20, n), scenes.uniform(seed, 21, n),
                                                      scenes.uniform(seed, 22, n)))
    b = orc.box_bounds(pos, quat, half, np.full(n, 0.02, np.float32))
    body = np.arange(n, dtype=np.uint32)
    body[1::7] = body[0::7][: len(body[1::7])]  # some shapes share a body
    dyn = (scenes.uniform(seed, 30, n) < 0.7).astype(np.uint8)
    got = orc.sap_pairs(b, body_id=body, dynamic=dyn)
    want = _brute_pairs(b, body, dyn)
    assert len(want) > 50
    np.testing.assert_array_equal(got, want)


def test_found_lost_is_set_difference(orc):
    rng = np.random.default_rng(5)
    a = np.unique(rng.integers(0, 50, (300, 2)).astype(np.uint32), axis=0)
    b = np.unique(rng.integers(0, 50, (300, 2)).astype(np.uint32), axis=0)
    f, l = orc.found_lost(a, b)
    sa, sb = set(map(tuple, a)), set(map(tuple, b))
    assert set(map(tuple, f)) == sb - sa and set(map(tuple, l)) == sa - sb
    assert np.all(np.diff(f[:, 0].astype(np.int64) * 1000 + f[:, 1]) > 0)


@pytest.mark.parametrize("name", ["sap_aabb.npz", "sap_rotated_floor.npz"])
def test_oracle_matches_reference_golden(orc, name):
    g = np.load(GOLDEN / name)
    half, quat, ids = g["half"], g["quat"], g["ids"]
    n = len(ids)
    floor = int(g["floor_id"][0])
    pos = g["pos0"].copy()
    off = np.full(n, 0.02, np.float32)
    prev = np.zeros((0, 2), np.uint32)
    for f in range(int(g["frames"][0])):
        if f:
            pos[g[f"moved_{f}"]] = g[f"pos_{f}"]
        b = orc.box_bounds(pos, quat, half, off)
        vals, pays = orc.sort_endpoints(b)
        np.testing.assert_array_equal(vals, g[f"ep_value_{f}"])
        np.testing.assert_array_equal(ids[pays >> 1], g[f"ep_id_{f}"])
        np.testing.assert_array_equal((pays & 1).astype(np.uint8), g[f"ep_max_{f}"])
        pairs = orc.sap_pairs(b, shape_id=ids, plane_ids=[] if floor < 0 else [floor])
        np.testing.assert_array_equal(pairs, g[f"pairs_{f}"])
        found, lost = orc.found_lost(prev, pairs)
        np.testing.assert_array_equal(found, g[f"found_{f}"])
        np.testing.assert_array_equal(lost, g[f"lost_{f}"])
        prev = pairs
