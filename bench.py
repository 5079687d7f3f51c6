"""Throughput benchmark of the FyzzXYZZY step pipeline on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload pile|sap]

Workload `pile` (default): the full step on the 100,000-box pile. Half
extents U[0.2, 1.0] m per axis, density 1000 kg/m^3, seeded
non-overlapping centers in a 100 x 100 x 200 m column above the y = 0
plane, seeded uniform orientations, friction 0.5, restitution 0,
dt = 1/120 s, 4 position + 1 velocity iterations. One step = the device
world's whole step (pose gather, SAP advance, plane x box pairs, box/box +
plane/box PCM with GJK/EPA continuation, touching-manager list, Dy:
partitions, contact prep, PGS, integration). Steps 0..W-1 are warm-up;
steps W..W+K-1 are timed (the drop: boxes land and pile up). `value` is
body-steps/s (bodies x timed steps / time; every body counts every step).

Workload `batch`: 4,096 independent piles (n_s ~ U{8..512} boxes, each the
pile setup with its footprint scaled by sqrt(n_s / 100,000)) in one device
world, every stage one launch over all scenes, each step one replay of a
captured HIP graph; launch shapes picked by the autotuner during warm-up.
Compared with the same scenes stepped as plain per-scene worlds (a sample
of 32 scenes, per-scene launches) and with the reference NativeCpu.

Workload `diff`: the differentiable form, 1,024 stacks of 64 boxes, 120
steps; one step = forward (device world, contact sets recorded,
checkpoints every 16 steps) plus backward (segments recomputed, Dy adjoint)
of one rollout step on average; value = body-steps/s of forward + backward.

Workload `sap`: broadphase only, 1,000,000 axis-aligned boxes (centers
uniform in a 1000 m cube, lognormal half extents median 0.5 m, sigma 0.5).
Each frame a seeded 10% of the boxes move by N(0, 0.05 m) per axis; one
step = apply that frame's moves + one incremental SAP frame. Frame 0 (the
full build) is set-up, not timed.

Prints one JSON line; inputs are resident in HBM before the timed region.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "fyzzxyzzy-sdk_amd"))
sys.path.insert(0, str(ROOT / "tests"))

METRIC = "body-steps/s at 100k-box pile + 4096-scene batch vs plain; SAP/radix HBM GB/s"
HBM_PEAK_GBS = 8000.0
# HBM bytes moved by one frame's endpoint radix sort (4 passes of tile
# histogram + tile scan + scatter over 3 x 2n keys, n = 1e6), from the two
# rocprofv3 PMC passes (FETCH_SIZE x 2.0,
# WRITE_SIZE x 1.0 after the copy calibration) summarised in
# profiles/sap_1m_incremental.md.
SAP_SORT_TRAFFIC_BYTES = 611_000_000
# HBM bytes of one k_pgs launch on the formed pile (3 steps after the state
# at step 610 is loaded: 224k managers, 601k contact rows, 789k friction
# rows, 5 iterations), same passes, profiles/pile_100k_full_step.md; the
# algorithmic bytes of those launches by the byte model below.
PILE_PGS_TRAFFIC = {"bytes": 2_319_000_000, "algorithmic_bytes": 488_600_000,
                    "measured_on": "formed pile: 3 steps after loading the state at step 610 (224k managers)"}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=None)
    p.add_argument("--warmup", type=int, default=None)
    p.add_argument("--workload", default="pile", choices=["pile", "sap", "batch", "diff"])
    p.add_argument("--cpu-seconds", type=float, default=12.0,
                   help="wall-clock budget of the timed CPU-baseline sample")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-frames", type=int, default=1)
    p.add_argument("--save-state", default=None,
                   help="pile: save the body states after the timed region (.npz)")
    p.add_argument("--load-state", default=None,
                   help="pile: start from saved body states instead of the drop (counter passes over a few "
                        "steps of the formed pile)")
    p.add_argument("--graph", action="store_true",
                   help="pile: time replays of captured step graphs (no per-stage timers in that region)")
    p.add_argument("--plain-launch", action="store_true",
                   help="plain instead of cooperative launches of the Dy grid-barrier kernels (counter profilers)")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="run one plain 256 MiB device copy first (PMC byte-counter calibration)")
    return p.parse_args()


def cpu_baseline_sap(n, frames, half, pos0, moves):
    """The compiled reference NativeCpu scene (oracle/_ref) when present,
    else the oracle restatement: SAP frames of the same workload on one
    host thread. Returns the bench.py cpu_baseline object."""
    import ctypes as C
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    ref_path = ROOT / "oracle" / "_ref" / "libfyzzref.so"
    if ref_path.exists():
        ref = C.CDLL(str(ref_path))
        ref.ref_scene_create.restype = C.c_void_p
        s = C.c_void_p(ref.ref_scene_create())
        ids = np.zeros(n, np.uint32)
        tmp = C.c_uint32()
        for i in range(n):
            if ref.ref_add_dynamic_box(s, P(half[i]), P(pos0[i]), C.c_float(1000.0), C.byref(tmp)) != 0:
                raise RuntimeError("reference rejected a box")
            ids[i] = tmp.value
        ref.ref_advance_collision(s, 1)  # frame 0: full build (not timed)
        # One incremental frame of the reference's own SAP: set_position of
        # the moved subset, then Scene::update_bounds + update_pairs (the
        # NativeCpu broadphase, without narrowphase or island work). Only
        # one frame: the pair publication that a second frame would need
        # belongs to advance_collision, whose Sc/island work is not SAP.
        frames = 1
        slots, newpos = moves[0]
        t0 = time.perf_counter()
        for j, sl in enumerate(slots):
            ref.ref_set_position(s, C.c_uint32(int(ids[sl])), P(newpos[j]))
        if ref.ref_sap_update(s) != 0:
            raise RuntimeError("reference SAP update failed")
        dt = time.perf_counter() - t0
        ref.ref_scene_destroy(s)
        kind = "reference"
        sample = (f"compiled reference NativeCpu Scene, {n} boxes, 1 incremental frame: set_position x "
                  f"{len(slots)} + Scene::update_bounds + Scene::update_pairs; frame 0 untimed")
    else:
        from oracle_bind import Oracle
        orc = Oracle()
        quat = np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1))
        off = np.full(n, 0.02, np.float32)
        pos = pos0.copy()
        prev = orc.sap_pairs(orc.box_bounds(pos, quat, half, off))
        t0 = time.perf_counter()
        for f in range(1, frames + 1):
            slots, newpos = moves[f - 1]
            pos[slots] = newpos
            b = orc.box_bounds(pos, quat, half, off)
            orc.sort_endpoints(b)
            cur = orc.sap_pairs(b)
            orc.found_lost(prev, cur)
            prev = cur
        dt = time.perf_counter() - t0
        kind = "port"
        sample = f"oracle restatement (full sort + sweep per frame), {n} boxes, {frames} frames"
    return {"value": n * frames / dt, "unit": "body-steps/s", "cores": 1, "kind": kind, "sample": sample,
            "seconds": round(dt, 3)}


def bench_sap(args):
    import torch
    from fyzz_amd import Context, scenes
    from fyzz_amd.sap import SapBroadphase

    n = 1_000_000
    frames = args.warmup + args.steps
    half, pos0 = scenes.sap_scene(n)
    quat = np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1))
    off = np.full(n, 0.02, np.float32)
    # per-frame moved subsets and their authored positions (host prepass)
    moves = []
    pos = pos0.copy()
    for f in range(1, frames + 1):
        slots, delta = scenes.sap_moves(n, f)
        pos[slots] = (pos[slots].astype(np.float64) + delta).astype(np.float32)
        moves.append((slots, pos[slots].copy()))

    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    ctx = Context(0, stream=stream.cuda_stream)
    d_pos = torch.from_numpy(pos0).to(dev)
    d_quat = torch.from_numpy(quat).to(dev)
    d_half = torch.from_numpy(half).to(dev)
    d_off = torch.from_numpy(off).to(dev)
    d_moves = [(torch.from_numpy(s.astype(np.int64)).to(dev), torch.from_numpy(s.astype(np.int32)).to(dev),
                torch.from_numpy(p).to(dev)) for s, p in moves]
    if args.calibrate_copy:
        calibrate_copy(dev)
    bp = SapBroadphase(ctx, np.arange(n, dtype=np.uint32), np.ones(n, np.uint8))
    bp.advance(d_pos, d_quat, d_half, d_off)  # frame 0: full build
    torch.cuda.synchronize()

    def frame(f):
        idx64, idx32, p = d_moves[f]
        d_pos.index_copy_(0, idx64, p)
        bp.advance(d_pos, d_quat, d_half, d_off, dirty=idx32)

    for f in range(args.warmup):
        frame(f)
    torch.cuda.synchronize()
    ctx.profile(True)
    for name in ("sap.bounds", "sap.sort", "sap.sort.scatter", "sap.pairs", "sap.lists"):
        ctx.profile_read(name)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for f in range(args.warmup, frames):
        frame(f)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    stages = {}
    for name in ("sap.bounds", "sap.sort", "sap.sort.scatter", "sap.pairs", "sap.lists"):
        t, c = ctx.profile_read(name)
        stages[name] = {"ms_total": round(t, 4), "launches": c, "ms_avg": round(t / c, 5) if c else None}
    ctx.profile(False)
    n_pairs, _, _, flags = bp.counts()
    if flags:
        raise RuntimeError(f"SAP status flags {flags}")
    ms_step = ms / args.steps
    # algorithmic bytes of the endpoint radix sort per frame: 3 axes x 2n
    # keys; per pass the tile histogram reads the 4-B key; pass 0's scatter
    # reads the plane key and writes key+payload (12 B); passes 1-3 read and
    # write key+payload (16 B). (The moved boxes' endpoint positions are a
    # separate pass of the walks.)
    keys = 3 * 2 * n
    sort_bytes = keys * (4 * 4 + 12 + 3 * 16)
    sort_ms = stages["sap.sort"]["ms_avg"]
    achieved = sort_bytes / (sort_ms * 1e-3) / 1e9
    result = {
        "metric": METRIC,
        "value": n * args.steps / (ms * 1e-3),
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "SAP broadphase only: 1,000,000 AABBs, 10% moved per frame, incremental SAP "
                               "with 32-bit endpoint keys (found/lost reports)",
                   "boxes": n, "moved_per_frame": len(moves[0][0]), "live_pairs": int(n_pairs),
                   "radix_digit_bits": 8, "radix_passes": 4,
                   "endpoint_keys_per_s": keys * args.steps / (ms * 1e-3)},
        "roofline": {"kernel": "endpoint radix sort (3 axes, 4 passes x tile histogram + tile scan + scatter)",
                     "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": SAP_SORT_TRAFFIC_BYTES,
                     "algorithmic_bytes_per_launch": sort_bytes},
        "stages": stages,
    }
    return result, (n, half, pos0, moves)


PILE_STAGES = ("sap.bounds", "sap.sort", "sap.sort.scatter", "sap.pairs", "sap.lists", "pcm.box_box",
               "pcm.plane_box", "world.managers", "dy.bodies", "dy.partition", "dy.prep", "dy.pgs",
               "dy.integrate")


def pile_algorithmic_bytes(stage, tot, n):
    """Algorithmic HBM bytes of one stage summed over the counted steps
    (tot: world totals over the same steps). Per-unit figures:
      dy.pgs     per iteration: 68 B per contact row (64-B row read +
                 applied impulse write), 60 B per friction row, 96 B per body
                 (velocity deltas read + written)
      pcm.box_box 644 B per broadphase pair (2 x 40 B poses, 224-B cache
                 read + write, 112-B result + count)
      dy.prep    per manager 128-B record + 96-B history read, per contact
                 row 64 B and per friction row 64 B written, 96-B patch write
      dy.partition per manager 128-B record read twice + 12 B of incidence
                 and links + 8 B partition/order written."""
    steps = max(tot["steps"], 1)
    iters_per_step = tot["iterations"] / steps
    if stage == "dy.pgs":
        return iters_per_step * (tot["contact_rows"] * 68 + tot["friction_rows"] * 60) + \
            tot["iterations"] * n * 96
    if stage == "pcm.box_box":
        return tot["pairs"] * 644
    if stage == "dy.prep":
        return tot["managers"] * (128 + 96 + 96) + (tot["contact_rows"] + tot["friction_rows"]) * 64
    if stage == "dy.partition":
        return tot["managers"] * (2 * 128 + 12 + 8)
    return None


def cpu_baseline_pile(records, half, budget_s):
    """The compiled reference NativeCpu scene (oracle/_ref) loaded with the
    device world's body states: 1 untimed step, then timed steps until the
    budget is spent (at least one)."""
    import ctypes as C
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    ref_path = ROOT / "oracle" / "_ref" / "libfyzzref.so"
    n = len(records)
    if not ref_path.exists():
        return {"value": None, "unit": "body-steps/s", "cores": 0, "kind": "reference",
                "sample": "unavailable: oracle/_ref not built"}
    ref = C.CDLL(str(ref_path))
    ref.ref_scene_create.restype = C.c_void_p
    s = C.c_void_p(ref.ref_scene_create())
    fid = C.c_uint32()
    ref.ref_add_floor(s, C.c_float(0.0), C.byref(fid))
    ref.ref_set_material(s, fid, C.c_float(0.5), C.c_float(0.5), C.c_float(0.0))
    ids = np.zeros(n, np.uint32)
    pos = np.ascontiguousarray(records["position"])
    quat = np.ascontiguousarray(records["orientation"])
    vel = np.ascontiguousarray(records["linear_velocity"])
    ang = np.ascontiguousarray(records["angular_velocity"])
    if ref.ref_add_boxes(s, C.c_uint32(n), P(np.ascontiguousarray(half, np.float32)), P(pos), P(quat), P(vel),
                         P(ang), C.c_float(1000.0), C.c_float(0.5), C.c_float(0.5), C.c_float(0.0), P(ids)) != 0:
        raise RuntimeError("reference rejected the scene")
    if ref.ref_step(s, C.c_float(1.0 / 120.0)) != 0:
        raise RuntimeError("reference step failed")
    steps = 0
    t0 = time.perf_counter()
    while True:
        if ref.ref_step(s, C.c_float(1.0 / 120.0)) != 0:
            raise RuntimeError("reference step failed")
        steps += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s or steps >= 20:
            break
    ref.ref_scene_destroy(s)
    return {"value": n * steps / dt, "unit": "body-steps/s", "cores": 2, "kind": "reference",
            "sample": (f"compiled reference NativeCpu Scene (caller + one solver worker thread) loaded with the "
                       f"device world's {n} body states at the end of the timed region (floor + boxes, same "
                       f"materials); 1 untimed step (fresh caches), then {steps} timed steps"),
            "seconds": round(dt, 3), "steps": steps}


def calibrate_copy(dev):
    """One plain copy of known size (256 MiB read + 256 MiB written), the
    reference dispatch for scaling the PMC byte counters."""
    import torch
    src = torch.empty(64 << 20, dtype=torch.int32, device=dev).fill_(1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def bench_pile(args):
    import torch
    from fyzz_amd import Context, scenes
    from fyzz_amd.world import World

    n = 100_000
    dt = 1.0 / 120.0
    sc = scenes.pile_scene(n)
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    ctx = Context(0, stream=stream.cuda_stream)
    if args.plain_launch:
        ctx.set_cooperative(False)
    if args.calibrate_copy:
        calibrate_copy(dev)
    w = World(ctx, n + 1)
    fid = w.add_static_floor(0.0)
    w.set_body_material(fid, 0.5, 0.5, 0.0)
    if args.load_state:
        s0 = np.load(args.load_state)
        w.add_dynamic_boxes(sc.half, s0["position"], s0["orientation"], linear=s0["linear_velocity"],
                            angular=s0["angular_velocity"], friction=0.5, restitution=0.0)
    else:
        w.add_dynamic_boxes(sc.half, sc.pos, sc.quat, friction=0.5, restitution=0.0)
    step = w.step_graph if args.graph else w.step
    for _ in range(args.warmup):
        step(dt)
    torch.cuda.synchronize()
    w.totals(reset=True)
    if not args.graph:
        # stage timers bracket the launches; a captured graph records none
        ctx.profile(True)
        for name in PILE_STAGES:
            ctx.profile_read(name)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        step(dt)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    stages = {}
    for name in PILE_STAGES:
        t, c = ctx.profile_read(name)
        stages[name] = {"ms_total": round(t, 3), "launches": c, "ms_avg": round(t / c, 5) if c else None}
    ctx.profile(False)
    tot = w.totals()
    st = w.stats()
    if st["flags"]:
        raise RuntimeError(f"world status flags {st['flags']:#x}")
    if args.save_state:
        r = w.body_records()
        np.savez(args.save_state, **{f: np.ascontiguousarray(r[f]) for f in
                                     ("position", "orientation", "linear_velocity", "angular_velocity")})
    timed = [k for k in stages if stages[k]["launches"] and pile_algorithmic_bytes(k, tot, n) is not None]
    dominant = max(timed, key=lambda k: stages[k]["ms_total"]) if timed else None
    alg = pile_algorithmic_bytes(dominant, tot, n) if dominant else 0.0
    achieved = alg / (stages[dominant]["ms_total"] * 1e-3) / 1e9 if dominant else 0.0
    result = {
        "metric": METRIC,
        "value": n * args.steps / (ms * 1e-3),
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "100,000-box pile, full step (SAP, PCM box/box + plane/box with GJK/EPA, Dy "
                               "partitions + PGS 4+1 + integration), dt 1/120 s",
                   "boxes": n, "timed_steps": f"{args.warmup}..{args.warmup + args.steps - 1}",
                   "pairs_at_end": st["pairs"], "managers_at_end": st["managers"],
                   "partitions_at_end": st["partitions"],
                   "contact_rows_per_step": tot["contact_rows"] / max(tot["steps"], 1),
                   "friction_rows_per_step": tot["friction_rows"] / max(tot["steps"], 1),
                   "partitions_per_step": tot["partitions"] / max(tot["steps"], 1),
                   "partition_rounds_per_step": tot["rounds"] / max(tot["steps"], 1)},
        "roofline": {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": PILE_PGS_TRAFFIC["bytes"] if dominant == "dy.pgs" else None,
                     "algorithmic_bytes_per_launch": alg / max(stages[dominant]["launches"], 1) if dominant else None}
        if dominant else None,
        "stages": stages,
    }
    if dominant == "dy.pgs":
        result["roofline"]["traffic_measured_on"] = PILE_PGS_TRAFFIC["measured_on"]
        result["roofline"]["traffic_algorithmic_bytes"] = PILE_PGS_TRAFFIC["algorithmic_bytes"]
    if not args.graph and not args.load_state:
        # the same workload in a fresh world, stepped by graph replay (the
        # headline keeps direct launches: its stage timers need them)
        g = World(ctx, n + 1)
        gf = g.add_static_floor(0.0)
        g.set_body_material(gf, 0.5, 0.5, 0.0)
        g.add_dynamic_boxes(sc.half, sc.pos, sc.quat, friction=0.5, restitution=0.0)
        for _ in range(args.warmup):
            g.step_graph(dt)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(args.steps):
            g.step_graph(dt)
        e1.record(stream)
        torch.cuda.synchronize()
        gms = e0.elapsed_time(e1)
        gst = g.stats()
        if gst["flags"]:
            raise RuntimeError(f"world status flags {gst['flags']:#x}")
        result["config"]["graph_replay"] = {"value": n * args.steps / (gms * 1e-3), "unit": "body-steps/s",
                                            "ms_per_step": gms / args.steps,
                                            "same_end_state": gst == st}
        g.close()
    return result, (w, sc)


def cpu_baseline_scenes(records_list, halves, budget_s, max_rounds):
    """The compiled reference NativeCpu on a sample of scenes, each loaded
    with its device end state: 1 untimed step, then rounds of one step per
    scene until the budget is spent or max_rounds (the device's timed step
    count) are done."""
    import ctypes as C
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    ref_path = ROOT / "oracle" / "_ref" / "libfyzzref.so"
    if not ref_path.exists():
        return {"value": None, "unit": "body-steps/s", "cores": 0, "kind": "reference",
                "sample": "unavailable: oracle/_ref not built"}
    ref = C.CDLL(str(ref_path))
    ref.ref_scene_create.restype = C.c_void_p
    handles = []
    for rec, half in zip(records_list, halves):
        sc = C.c_void_p(ref.ref_scene_create())
        fid = C.c_uint32()
        ref.ref_add_floor(sc, C.c_float(0.0), C.byref(fid))
        ref.ref_set_material(sc, fid, C.c_float(0.5), C.c_float(0.5), C.c_float(0.0))
        ids = np.zeros(len(rec), np.uint32)
        arrs = [np.ascontiguousarray(rec[f]) for f in ("position", "orientation", "linear_velocity",
                                                       "angular_velocity")]
        if ref.ref_add_boxes(sc, C.c_uint32(len(rec)), P(np.ascontiguousarray(half, np.float32)), P(arrs[0]),
                             P(arrs[1]), P(arrs[2]), P(arrs[3]), C.c_float(1000.0), C.c_float(0.5), C.c_float(0.5),
                             C.c_float(0.0), P(ids)) != 0:
            raise RuntimeError("reference rejected a scene")
        ref.ref_step(sc, C.c_float(1.0 / 120.0))
        handles.append((sc, len(rec)))
    body_steps = 0
    rounds = 0
    t0 = time.perf_counter()
    while True:
        for sc, nb in handles:
            if ref.ref_step(sc, C.c_float(1.0 / 120.0)) != 0:
                raise RuntimeError("reference step failed")
            body_steps += nb
        rounds += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s or rounds >= max_rounds:
            break
    # all-core: the same scenes stepped concurrently, one host thread per
    # group of scenes (ctypes releases the interpreter lock during the call)
    import concurrent.futures as cf
    import os
    cores = min(16, os.cpu_count() or 1)
    groups = [handles[k::cores] for k in range(cores)]

    def run_group(g):
        done = 0
        for _ in range(rounds):
            for sc, nb in g:
                if ref.ref_step(sc, C.c_float(1.0 / 120.0)) != 0:
                    raise RuntimeError("reference step failed")
                done += nb
        return done

    t1 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=cores) as ex:
        all_steps = sum(ex.map(run_group, groups))
    dt_all = time.perf_counter() - t1
    for sc, _ in handles:
        ref.ref_scene_destroy(sc)
    return {"value": body_steps / dt, "unit": "body-steps/s", "cores": 2, "kind": "reference",
            "sample": (f"compiled reference NativeCpu Scene per scene (caller + one solver worker thread), "
                       f"{len(handles)} of the scenes loaded with their device end states; 1 untimed step, then "
                       f"{rounds} timed steps each"),
            "seconds": round(dt, 3),
            "all_core": {"value": all_steps / dt_all, "threads": cores, "host_cores": os.cpu_count(),
                         "sample": f"the same scenes, {rounds} more steps each, split over {cores} host threads",
                         "seconds": round(dt_all, 3)}}


def bench_batch(args):
    import torch
    from fyzz_amd import Context, scenes
    from fyzz_amd.world import World

    dt = 1.0 / 120.0
    n_scenes = 4096
    counts, half, pos, quat = scenes.batch_scenes(n_scenes)
    n = int(counts.sum())
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    ctx = Context(0, stream=stream.cuda_stream)
    w = World(ctx, n + 1)
    fid = w.add_static_floor(0.0)
    w.set_body_material(fid, 0.5, 0.5, 0.0)
    w.add_dynamic_boxes(half, pos, quat, friction=0.5, restitution=0.0)
    w.set_scenes(counts)
    tuned = w.autotune(dt, trial_steps=3)
    tuned_steps = w.totals()["steps"]
    for _ in range(max(args.warmup - tuned_steps, 0)):
        w.step_graph(dt)
    torch.cuda.synchronize()
    warm = w.totals()["steps"]
    w.totals(reset=True)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        w.step_graph(dt)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    st = w.stats()
    if st["flags"]:
        raise RuntimeError(f"world status flags {st['flags']:#x}")
    tot = w.totals()
    # stage shares from a few profiled eager steps after the timed region
    ctx.profile(True)
    for name in PILE_STAGES + ("world.scene_pairs",):
        ctx.profile_read(name)
    prof_steps = 5
    w.totals(reset=True)
    for _ in range(prof_steps):
        w.step(dt)
    stages = {}
    for name in PILE_STAGES + ("world.scene_pairs",):
        t, c = ctx.profile_read(name)
        if c:
            stages[name] = {"ms_total": round(t, 3), "launches": c, "ms_avg": round(t / c, 5)}
    ctx.profile(False)
    ptot = w.totals()
    dominant = max((k for k in stages if pile_algorithmic_bytes(k, ptot, n) is not None),
                   key=lambda k: stages[k]["ms_total"])
    alg = pile_algorithmic_bytes(dominant, ptot, n)
    achieved = alg / (stages[dominant]["ms_total"] * 1e-3) / 1e9
    # plain per-scene launches: a sample of scenes, one world each
    first = np.concatenate([[0], np.cumsum(counts.astype(np.int64))]).astype(np.int64)
    sample = np.linspace(0, n_scenes - 1, 32).astype(np.int64)
    worlds = []
    for k in sample:
        a, b = first[k], first[k + 1]
        pw = World(ctx, int(b - a) + 1)
        f2 = pw.add_static_floor(0.0)
        pw.set_body_material(f2, 0.5, 0.5, 0.0)
        pw.add_dynamic_boxes(half[a:b], pos[a:b], quat[a:b], friction=0.5, restitution=0.0)
        worlds.append((pw, int(b - a)))
    plain_steps = 20
    for _ in range(warm):
        for pw, _nb in worlds:
            pw.step(dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        for pw, _nb in worlds:
            pw.step(dt)
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    plain_value = sum(nb for _, nb in worlds) * plain_steps / plain_s
    value = n * args.steps / (ms * 1e-3)
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warm,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched: 4,096 independent piles (U{8..512} boxes, footprint scaled by "
                               "sqrt(n/1e5)) in one world, one launch per stage, one HIP graph replay per step",
                   "scenes": n_scenes, "boxes": n, "timed_steps": f"{warm}..{warm + args.steps - 1}",
                   "autotuned_workgroups": {"partition": tuned[0], "pgs": tuned[1]},
                   "managers_at_end": st["managers"], "pairs_at_end": st["pairs"],
                   "plain_per_scene_launches": {"value": plain_value, "unit": "body-steps/s",
                                                "sample": f"{len(worlds)} scenes, one world each, {plain_steps} "
                                                          f"steps after the same warm-up"},
                   "speedup_vs_plain": value / plain_value},
        "roofline": {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "measured_on": f"{prof_steps} profiled eager steps after the timed region",
                     "algorithmic_bytes_per_launch": alg / max(stages[dominant]["launches"], 1)},
        "stages": stages,
    }
    recs = w.body_records()
    cpu_sample = np.linspace(0, n_scenes - 1, 128).astype(np.int64)
    sample_recs = [recs[first[k]:first[k + 1]] for k in cpu_sample]
    sample_half = [half[first[k]:first[k + 1]] for k in cpu_sample]
    return result, (sample_recs, sample_half)


def bench_diff(args):
    import torch
    from fyzz_amd import Context, scenes
    from fyzz_amd.diff import Rollout, rollout_function

    n_scenes, steps = 1024, 120
    counts, half, pos, quat, v0, w0, mu, target = scenes.diff_scenes(n_scenes)
    n = int(counts.sum())
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    ctx = Context(0, stream=stream.cuda_stream)
    ro = Rollout(ctx, half, pos, quat, counts, steps, checkpoint_every=16)
    fn = rollout_function()
    scene = torch.as_tensor(ro.scene_of_body, device=dev)
    tgt = torch.tensor(target, dtype=torch.float32, device=dev)

    def once():
        tv = torch.tensor(v0, device=dev, requires_grad=True)
        tw = torch.tensor(w0, device=dev, requires_grad=True)
        tm = torch.tensor(mu, device=dev, requires_grad=True)
        t0 = time.perf_counter()
        p = fn.apply(tv, tw, tm, ro)
        com = torch.zeros((n_scenes, 3), device=dev).index_add_(0, scene, p) / 64.0
        d = (com - tgt)[:, [0, 2]]
        loss = (d * d).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return loss.item(), t1 - t0, t2 - t1, tv.grad, tm.grad

    once()  # warm-up rollout (allocations, first-use paths)
    loss, tf, tb, gv, gm = once()
    total = tf + tb
    result = {
        "metric": METRIC,
        "value": n * steps / total,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": steps,
        "ms_per_step": total * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "differentiable: 1,024 stacks of 64 boxes, 120 steps, loss on the stacks' final "
                               "COM x/z; gradients w.r.t. v0, w0 per body and friction per scene (frozen contacts, "
                               "checkpoint every 16 steps)",
                   "bodies": n, "forward_s": round(tf, 4), "backward_s": round(tb, 4), "loss": loss,
                   "grad_v0_norm": float(gv.norm()), "grad_mu_norm": float(gm.norm())},
        "roofline": None,
    }
    return result, (ro, counts, v0, w0, mu)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; --gpus must be 1 "
                                   "(independent scenes or slabs would shard across GPUs; see DESIGN.md)"}))
        sys.exit(2)
    if os.environ.get("WORLD_SIZE", "1") not in ("1", ""):
        print("bench.py runs on one GPU; WORLD_SIZE>1 is not supported", file=sys.stderr)
        sys.exit(2)
    if args.workload == "sap":
        args.steps = 50 if args.steps is None else args.steps
        args.warmup = 9 if args.warmup is None else args.warmup
        result, (n, half, pos0, moves) = bench_sap(args)
        if not args.no_cpu_baseline:
            result["cpu_baseline"] = cpu_baseline_sap(n, min(args.cpu_frames, len(moves)), half, pos0, moves)
        else:
            result["cpu_baseline"] = None
    elif args.workload == "diff":
        result, _ = bench_diff(args)
        result["cpu_baseline"] = None
    elif args.workload == "batch":
        args.steps = 240 if args.steps is None else args.steps
        args.warmup = 40 if args.warmup is None else args.warmup
        result, (recs, halves) = bench_batch(args)
        if not args.no_cpu_baseline:
            result["cpu_baseline"] = cpu_baseline_scenes(recs, halves, args.cpu_seconds, args.steps)
        else:
            result["cpu_baseline"] = None
    else:
        args.steps = 600 if args.steps is None else args.steps
        args.warmup = 10 if args.warmup is None else args.warmup
        result, (w, sc) = bench_pile(args)
        if not args.no_cpu_baseline:
            result["cpu_baseline"] = cpu_baseline_pile(w.body_records(), sc.half, args.cpu_seconds)
        else:
            result["cpu_baseline"] = None
    print(json.dumps(result))


if __name__ == "__main__":
    main()
