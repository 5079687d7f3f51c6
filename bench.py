"""Throughput benchmark of the FyzzXYZZY step pipeline on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload pile|sap|batch|diff]

Workload `pile` (default): the full step on the 100,000-box pile. Half
extents U[0.2, 1.0] m per axis, density 1000 kg/m^3, seeded
non-overlapping centers in a 100 x 100 x 200 m column above the y = 0
plane, seeded uniform orientations, friction 0.5, restitution 0,
dt = 1/120 s, 4 position + 1 velocity iterations. One step = the device
world's whole step (pose gather, SAP advance, plane x box pairs, box/box +
plane/box PCM with GJK/EPA continuation, touching-manager list, Dy:
partitions, contact prep, PGS, integration), one replay of the captured
step graph. The pile is first dropped and settled for --settle steps (600
by default: the boxes land and form the pile, ~2e5 touching managers),
untimed for the headline but clocked as a whole ("drop" in the line);
with --autotune the library's autotuner then picks the partition and PGS
workgroup counts on the formed pile; then W warm-up steps; then K timed
steps of the formed pile. `value` is
body-steps/s (bodies x timed steps / time; every body counts every step).
K more steps of the same pile run with direct launches and stage timers
(HIP events on the launch stream): their per-stage times give the
roofline of the dominant stage.

Three more measurements are made in the same run and attached to the line
(skipped with --no-extras):
  "pile_native_order"  the pile in the native solver order (World::step
                with the host island authority, the NativeCpu order), 10
                eager steps from the timed region's end state.
  "sap_1m"      broadphase only, 1,000,000 axis-aligned boxes (centers
                uniform in a 1000 m cube, lognormal half extents median
                0.5 m, sigma 0.5); each frame a seeded 10% moves by
                N(0, 0.05 m) per axis; 60 frames (frame 0 the full build,
                9 warm-up, 50 timed incremental frames with found/lost
                reports); roofline of the endpoint radix sort.
  "batch_4096"  4,096 independent piles (n_s ~ U{8..512} boxes, each the
                pile setup with its footprint scaled by sqrt(n_s/100,000))
                in one device world, each step one graph replay; 40
                warm-up steps, then the launch shapes autotuned on the
                formed piles (100 trial steps), then 200 timed steps; against
                the same scenes stepped as plain per-scene worlds and the
                reference NativeCpu.

Workload `diff`: the differentiable form, 1,024 stacks of 64 boxes, 120
steps of forward + backward.

Prints one JSON line; inputs are resident in HBM before each timed region.
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
DT = 1.0 / 120.0

# HBM bytes per launch from rocprofv3 PMC passes (FETCH_SIZE x 2.0 and
# WRITE_SIZE x 1.0 after the plain-copy calibration), keyed by stage; the
# passes and the launches they cover are in profiles/pile_100k_counters.md
# and profiles/sap_1m_incremental.md.
PILE_TRAFFIC = {
    "dy.pgs": {"bytes": 2_235_000_000, "algorithmic_bytes": 543_700_000,
               "measured_on": "formed pile: 3 steps after loading the state saved at step 645 (248k managers)"},
}
SAP_SORT_TRAFFIC_BYTES = 611_000_000


_T0 = time.perf_counter()


def progress(msg):
    """A progress line on stderr (the result is the one JSON line on stdout)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=None)
    p.add_argument("--warmup", type=int, default=None)
    p.add_argument("--settle", type=int, default=600,
                   help="pile: untimed steps from the drop before the warm-up (the pile forms)")
    p.add_argument("--workload", default="pile", choices=["pile", "sap", "batch", "diff"])
    p.add_argument("--autotune", action="store_true",
                   help="pile: let the autotuner pick the partition/PGS workgroup counts after settling")
    p.add_argument("--no-extras", action="store_true",
                   help="pile: skip the attached native-order, 1M-box SAP and 4,096-scene batch measurements")
    p.add_argument("--cpu-seconds", type=float, default=12.0,
                   help="wall-clock budget of a timed CPU-baseline sample")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--save-state", default=None,
                   help="pile: save the body states after the timed region (.npz)")
    p.add_argument("--load-state", default=None,
                   help="pile: start from saved body states instead of the drop (no settle; counter passes "
                        "over a few steps of the formed pile)")
    p.add_argument("--eager", action="store_true",
                   help="pile: direct launches only (no graph capture; for counter passes)")
    p.add_argument("--plain-launch", action="store_true",
                   help="plain instead of cooperative launches of the Dy grid-barrier kernels (counter profilers)")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="run one plain 256 MiB device copy first (PMC byte-counter calibration)")
    return p.parse_args()


# ---- algorithmic bytes ------------------------------------------------------

def stage_bytes(stage, tot, n):
    """Algorithmic HBM bytes of one stage summed over the counted steps
    (tot: world totals over the same steps; n: boxes). Per-unit figures:
      sap.bounds   68 B per box (pose 28 + half 12 + offset 4 in, 24 out)
      sap.sort     76 B per endpoint key, 3 axes x 2n keys (8-bit digits,
                   4 passes: 4 x 4-B histogram read, 12 B pass 0, 3 x 16 B)
      sap.pairs    3 x 2n x 12 B (rank, payload, prefix per endpoint) + 48 B
                   per box touched by a crossing (<= n: every box moves in a
                   stepped world) + 8 B per found/lost pair
      sap.lists    16 B per live pair (old list read, new list written) +
                   8 B per found/lost pair
      pcm.box_box  644 B per broadphase pair (2 x 40 B poses, 224-B cache
                   read + write, 112-B result + count)
      pcm.plane_box 604 B per box (40-B pose, 224-B cache read + write,
                   112-B result + count)
      world.managers 4 B per candidate (pairs + boxes) + 244 B per touching
                   manager (112-B contacts read, 128-B record + slot written)
      world.scene_pairs 80 B per box (40-B pose read in each of the two
                   phases) + 8 B per pair written
      dy.bodies    400 B per body (144-B record read and written, 112 B of
                   solver rows)
      dy.partition 276 B per manager (128-B record read twice, 12 B of
                   incidence and links, 8 B partition/order written)
      dy.prep      320 B per manager (record, history, patch) + 64 B per
                   contact row and per friction row written
      dy.pgs       per iteration: 68 B per contact row (64-B row read +
                   applied impulse write), 60 B per friction row, 96 B per
                   body (velocity deltas read + written)
      dy.integrate 236 B per body (152-B input, 84-B pose/sleep output)
      world.islands 24 B per live pair (pair, counts, source, touch flag
                   read once) + 32 B per body (flags, two parent words and
                   three root flags written, flags read)"""
    steps = max(tot["steps"], 1)
    pairs, managers, fl = tot["pairs"], tot["managers"], tot.get("found_lost", 0)
    body_steps = n * steps
    if stage == "sap.bounds":
        return 68 * body_steps
    if stage == "sap.sort":
        return 76 * 6 * body_steps
    if stage == "sap.pairs":
        return (3 * 2 * 12 + 48) * body_steps + 8 * fl
    if stage == "sap.lists":
        return 16 * pairs + 8 * fl
    if stage == "pcm.box_box":
        return 644 * pairs
    if stage == "pcm.plane_box":
        return 604 * body_steps
    if stage == "world.managers":
        return 4 * (pairs + body_steps) + 244 * managers
    if stage == "world.scene_pairs":
        return 80 * body_steps + 8 * pairs
    if stage == "dy.bodies":
        return 400 * body_steps
    if stage == "dy.partition":
        return 276 * managers
    if stage == "dy.prep":
        return 320 * managers + 64 * (tot["contact_rows"] + tot["friction_rows"])
    if stage == "dy.pgs":
        iters_per_step = tot["iterations"] / steps
        return iters_per_step * (tot["contact_rows"] * 68 + tot["friction_rows"] * 60) + tot["iterations"] * n * 96
    if stage == "dy.integrate":
        return 236 * body_steps
    if stage == "world.islands":
        return 24 * pairs + 32 * body_steps
    return None


def roofline_of(stages, tot, n, traffic=None):
    """The dominant stage (largest summed time among the stages that own
    launches; the radix scatter is timed inside sap.sort and not ranked on
    its own) with its algorithmic bytes per launch over its average launch
    time, and every stage's own fraction of the HBM peak."""
    ranked = {k: v for k, v in stages.items() if v["launches"] and stage_bytes(k, tot, n) is not None}
    for k, v in ranked.items():
        alg = stage_bytes(k, tot, n)
        v["algorithmic_bytes_per_launch"] = round(alg / v["launches"])
        v["GBps"] = round(alg / (v["ms_total"] * 1e-3) / 1e9, 1) if v["ms_total"] else None
        v["frac"] = round(v["GBps"] / HBM_PEAK_GBS, 4) if v["GBps"] else None
    if not ranked:
        return None
    dom = max(ranked, key=lambda k: ranked[k]["ms_total"])
    v = ranked[dom]
    out = {"kernel": dom, "bound": "hbm", "achieved": v["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": v["frac"], "traffic": None, "algorithmic_bytes_per_launch": v["algorithmic_bytes_per_launch"],
           "ms_per_launch": round(v["ms_total"] / v["launches"], 5)}
    if traffic and dom in traffic:
        out["traffic"] = traffic[dom]["bytes"]
        out["traffic_measured_on"] = traffic[dom]["measured_on"]
        out["traffic_algorithmic_bytes"] = traffic[dom]["algorithmic_bytes"]
    return out


# ---- CPU baselines (the compiled reference NativeCpu, else skipped) ---------

def _ref():
    import ctypes as C
    ref_path = ROOT / "oracle" / "_ref" / "libfyzzref.so"
    if not ref_path.exists():
        return None
    ref = C.CDLL(str(ref_path))
    ref.ref_scene_create.restype = C.c_void_p
    return ref


def cpu_baseline_sap(n, half, pos0, moves):
    """The compiled reference NativeCpu scene (oracle/_ref) when present,
    else the oracle restatement: one incremental SAP frame of the same
    workload on one host thread."""
    progress("sap: CPU baseline")
    import ctypes as C
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    ref = _ref()
    if ref is not None:
        s = C.c_void_p(ref.ref_scene_create())
        ids = np.zeros(n, np.uint32)
        if ref.ref_add_aabbs(s, C.c_uint32(n), P(np.ascontiguousarray(half, np.float32)),
                             P(np.ascontiguousarray(pos0, np.float32)), C.c_float(1000.0), P(ids)) != 0:
            raise RuntimeError("reference rejected the boxes")
        ref.ref_advance_collision(s, 1)  # frame 0: full build (not timed)
        # One incremental frame of the reference's own SAP: set_position of
        # the moved subset, then Scene::update_bounds + update_pairs (the
        # NativeCpu broadphase, without narrowphase or island work).
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
        slots, newpos = moves[0]
        t0 = time.perf_counter()
        pos[slots] = newpos
        b = orc.box_bounds(pos, quat, half, off)
        orc.sort_endpoints(b)
        orc.found_lost(prev, orc.sap_pairs(b))
        dt = time.perf_counter() - t0
        kind = "port"
        sample = f"oracle restatement (full sort + sweep), {n} boxes, 1 frame"
    return {"value": n / dt, "unit": "box-frames/s", "cores": 1, "kind": kind, "sample": sample,
            "seconds": round(dt, 3)}


def _ref_scene(ref, rec, half):
    import ctypes as C
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    s = C.c_void_p(ref.ref_scene_create())
    fid = C.c_uint32()
    ref.ref_add_floor(s, C.c_float(0.0), C.byref(fid))
    ref.ref_set_material(s, fid, C.c_float(0.5), C.c_float(0.5), C.c_float(0.0))
    ids = np.zeros(len(rec), np.uint32)
    arrs = [np.ascontiguousarray(rec[f]) for f in ("position", "orientation", "linear_velocity", "angular_velocity")]
    if ref.ref_add_boxes(s, C.c_uint32(len(rec)), P(np.ascontiguousarray(half, np.float32)), P(arrs[0]), P(arrs[1]),
                         P(arrs[2]), P(arrs[3]), C.c_float(1000.0), C.c_float(0.5), C.c_float(0.5), C.c_float(0.0),
                         P(ids)) != 0:
        raise RuntimeError("reference rejected the scene")
    return s


def cpu_baseline_pile(records, half, budget_s):
    """The compiled reference NativeCpu scene (oracle/_ref) loaded with the
    device world's body states: 1 untimed step, then timed steps until the
    budget is spent (at least one)."""
    progress("pile: CPU baseline")
    import ctypes as C
    ref = _ref()
    n = len(records)
    if ref is None:
        return {"value": None, "unit": "body-steps/s", "cores": 0, "kind": "reference",
                "sample": "unavailable: oracle/_ref not built"}
    s = _ref_scene(ref, records, half)
    if ref.ref_step(s, C.c_float(DT)) != 0:
        raise RuntimeError("reference step failed")
    steps = 0
    t0 = time.perf_counter()
    while True:
        if ref.ref_step(s, C.c_float(DT)) != 0:
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


def cpu_baseline_scenes(records_list, halves, budget_s, max_rounds):
    """The compiled reference NativeCpu on a sample of scenes, each loaded
    with its device end state: 1 untimed step, then rounds of one step per
    scene until the budget is spent or max_rounds are done; then the same
    scenes split over host threads."""
    import concurrent.futures as cf
    import ctypes as C
    progress("batch: CPU baseline")
    ref = _ref()
    if ref is None:
        return {"value": None, "unit": "body-steps/s", "cores": 0, "kind": "reference",
                "sample": "unavailable: oracle/_ref not built"}
    handles = []
    for rec, half in zip(records_list, halves):
        sc = _ref_scene(ref, rec, half)
        ref.ref_step(sc, C.c_float(DT))
        handles.append((sc, len(rec)))
    body_steps = 0
    rounds = 0
    t0 = time.perf_counter()
    while True:
        for sc, nb in handles:
            if ref.ref_step(sc, C.c_float(DT)) != 0:
                raise RuntimeError("reference step failed")
            body_steps += nb
        rounds += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s or rounds >= max_rounds:
            break
    # all-core: the same scenes stepped concurrently, one host thread per
    # group of scenes (ctypes releases the interpreter lock during the call)
    cores = min(16, os.cpu_count() or 1)
    groups = [handles[k::cores] for k in range(cores)]

    def run_group(g):
        done = 0
        for _ in range(rounds):
            for sc, nb in g:
                if ref.ref_step(sc, C.c_float(DT)) != 0:
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


class Clock:
    """HIP events on one stream around a region."""

    def __init__(self, stream):
        import torch
        self.stream = stream
        self.a = torch.cuda.Event(enable_timing=True)
        self.b = torch.cuda.Event(enable_timing=True)

    def __enter__(self):
        import torch
        torch.cuda.synchronize()
        self.a.record(self.stream)
        return self

    def __exit__(self, *exc):
        import torch
        self.b.record(self.stream)
        torch.cuda.synchronize()
        self.ms = self.a.elapsed_time(self.b)


PILE_STAGES = ("sap.bounds", "sap.sort", "sap.pairs", "sap.lists", "pcm.box_box", "pcm.plane_box",
               "world.islands", "world.managers", "dy.bodies", "dy.partition", "dy.prep", "dy.pgs",
               "dy.integrate")


def profiled_steps(ctx, w, steps, names, step):
    """`steps` steps with the stage timers on; returns (stages, totals)."""
    w.totals(reset=True)
    ctx.profile(True)
    for name in names:
        ctx.profile_read(name)
    for _ in range(steps):
        step(DT)
    stages = {}
    for name in names:
        t, c = ctx.profile_read(name)
        stages[name] = {"ms_total": round(t, 4), "launches": c, "ms_avg": round(t / c, 5) if c else None}
    ctx.profile(False)
    return stages, w.totals()


# ---- workloads -----------------------------------------------------------------

def bench_pile(args, ctx, stream):
    import torch
    from fyzz_amd import scenes
    from fyzz_amd.world import World

    n = 100_000
    sc = scenes.pile_scene(n)
    w = World(ctx, n + 1)
    fid = w.add_static_floor(0.0)
    w.set_body_material(fid, 0.5, 0.5, 0.0)
    if args.load_state:
        s0 = np.load(args.load_state)
        w.add_dynamic_boxes(sc.half, s0["position"], s0["orientation"], linear=s0["linear_velocity"],
                            angular=s0["angular_velocity"], friction=0.5, restitution=0.0)
    else:
        w.add_dynamic_boxes(sc.half, sc.pos, sc.quat, friction=0.5, restitution=0.0)
    step = w.step if args.eager else w.step_graph
    settle = 0 if args.load_state else args.settle
    drop = None
    if settle:
        progress(f"pile: dropping and settling {n} boxes for {settle} steps")
        w.totals(reset=True)
        with Clock(stream) as c:
            for _ in range(settle):
                step(DT)
        t = w.totals()
        drop = {"steps": settle, "ms_per_step": c.ms / settle, "value": n * settle / (c.ms * 1e-3),
                "unit": "body-steps/s", "contact_rows_per_step": t["contact_rows"] / settle,
                "managers_per_step": t["managers"] / settle}
    # optionally: launch shapes of the partition and PGS kernels timed on
    # the formed pile by the library's autotuner (its trial steps advance
    # the pile; whole-step trials on a changing pile are noisy, so the
    # default keeps the measured defaults)
    tuned = w.autotune(DT, trial_steps=2) if args.autotune else None
    progress(f"pile: {args.warmup} warm-up and {args.steps} timed steps")
    for _ in range(args.warmup):
        step(DT)
    w.totals(reset=True)
    with Clock(stream) as c:
        for _ in range(args.steps):
            step(DT)
    ms = c.ms
    tot = w.totals()
    st = w.stats()
    if st["flags"]:
        raise RuntimeError(f"world status flags {st['flags']:#x}")
    # the same pile, K more steps with direct launches and stage timers
    progress(f"pile: {ms / args.steps:.3f} ms per step; {args.steps} steps with stage timers")
    stages, ptot = profiled_steps(ctx, w, args.steps, PILE_STAGES, w.step)
    pst = w.stats()
    if pst["flags"]:
        raise RuntimeError(f"world status flags {pst['flags']:#x}")
    if args.save_state:
        r = w.body_records()
        np.savez(args.save_state, **{f: np.ascontiguousarray(r[f]) for f in
                                     ("position", "orientation", "linear_velocity", "angular_velocity")})
    per = max(tot["steps"], 1)
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
                               "partitions + PGS 4+1 + integration), dt 1/120 s, timed on the formed pile after "
                               f"{settle} settle steps",
                   "boxes": n, "launch": "direct" if args.eager else "graph replay",
                   "timed_steps": f"{settle + args.warmup}..{settle + args.warmup + args.steps - 1}",
                   "pairs_per_step": tot["pairs"] / per, "managers_per_step": tot["managers"] / per,
                   "contact_rows_per_step": tot["contact_rows"] / per,
                   "friction_rows_per_step": tot["friction_rows"] / per,
                   "partitions_per_step": tot["partitions"] / per,
                   "found_lost_per_step": tot["found_lost"] / per,
                   "pairs_at_end": st["pairs"], "managers_at_end": st["managers"],
                   "partitions_at_end": st["partitions"],
                   "drop": drop,
                   "autotuned_workgroups": ({"partition": tuned[0], "pgs": tuned[1]} if tuned else None),
                   "direct_launch_profiled": {"steps": args.steps,
                                              "ms_per_step": sum(v["ms_total"] for v in stages.values())
                                              / max(args.steps, 1),
                                              "note": "sum of stage timers (HIP events), direct launches"}},
        "roofline": roofline_of(stages, ptot, n, PILE_TRAFFIC),
        "stages": stages,
    }
    return result, w, sc


def bench_pile_native(ctx, stream, recs, half, steps=10, warmup=2):
    """The pile in the native solver order (World::step with the host island
    authority, the NativeCpu order): a world loaded with the timed region's
    end state, one untimed first step (uploads, full broadphase build, the
    authority's admission), `warmup` more, then `steps` timed eager steps."""
    from fyzz_amd.world import World
    progress(f"pile: {steps} steps in the native solver order")
    n = len(recs)
    w = World(ctx, n + 1)
    try:
        fid = w.add_static_floor(0.0)
        w.set_body_material(fid, 0.5, 0.5, 0.0)
        w.add_dynamic_boxes(half, recs["position"], recs["orientation"], linear=recs["linear_velocity"],
                            angular=recs["angular_velocity"], friction=0.5, restitution=0.0)
        w.set_solver_order(True)
        for _ in range(1 + warmup):
            w.step(DT)
        with Clock(stream) as c:
            for _ in range(steps):
                w.step(DT)
        st = w.stats()
        if st["flags"]:
            raise RuntimeError(f"world status flags {st['flags']:#x}")
        return {"value": n * steps / (c.ms * 1e-3), "unit": "body-steps/s", "steps": steps,
                "ms_per_step": c.ms / steps, "managers_at_end": st["managers"],
                "note": "World::step in the native solver order (the host island authority feeding the device "
                        "solver list; two synchronizations per step), eager, from the timed region's end state"}
    finally:
        w.close()


def bench_sap(args, ctx, stream, frames=60, warmup=9):
    """Broadphase-only: frame 0 the full build, `warmup` incremental frames,
    then frames-1-warmup timed incremental frames."""
    progress(f"sap: 1,000,000 boxes, {frames} frames")
    import torch
    from fyzz_amd import scenes
    from fyzz_amd.sap import SapBroadphase

    n = 1_000_000
    timed = frames - 1 - warmup
    half, pos0 = scenes.sap_scene(n)
    quat = np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1))
    off = np.full(n, 0.02, np.float32)
    moves = []
    pos = pos0.copy()
    for f in range(1, frames):
        slots, delta = scenes.sap_moves(n, f)
        pos[slots] = (pos[slots].astype(np.float64) + delta).astype(np.float32)
        moves.append((slots, pos[slots].copy()))
    dev = torch.device("cuda", 0)
    d_pos = torch.from_numpy(pos0).to(dev)
    d_quat = torch.from_numpy(quat).to(dev)
    d_half = torch.from_numpy(half).to(dev)
    d_off = torch.from_numpy(off).to(dev)
    d_moves = [(torch.from_numpy(s.astype(np.int64)).to(dev), torch.from_numpy(s.astype(np.int32)).to(dev),
                torch.from_numpy(p).to(dev)) for s, p in moves]
    bp = SapBroadphase(ctx, np.arange(n, dtype=np.uint32), np.ones(n, np.uint8))
    try:
        bp.advance(d_pos, d_quat, d_half, d_off)  # frame 0: full build
        torch.cuda.synchronize()

        def frame(f):
            idx64, idx32, p = d_moves[f]
            d_pos.index_copy_(0, idx64, p)
            bp.advance(d_pos, d_quat, d_half, d_off, dirty=idx32)

        for f in range(warmup):
            frame(f)
        names = ("sap.bounds", "sap.sort", "sap.pairs", "sap.lists")
        ctx.profile(True)
        for name in names:
            ctx.profile_read(name)
        with Clock(stream) as c:
            for f in range(warmup, frames - 1):
                frame(f)
        stages = {}
        for name in names:
            t, k = ctx.profile_read(name)
            stages[name] = {"ms_total": round(t, 4), "launches": k, "ms_avg": round(t / k, 5) if k else None}
        ctx.profile(False)
        n_pairs, _, _, flags = bp.counts()
        if flags:
            raise RuntimeError(f"SAP status flags {flags}")
    finally:
        bp.close()
    keys = 3 * 2 * n
    sort_bytes = 76 * keys
    sort_ms = stages["sap.sort"]["ms_avg"]
    achieved = sort_bytes / (sort_ms * 1e-3) / 1e9
    frame_bytes = (68 + 76 * 6 + 3 * 2 * 12) * n + 48 * len(moves[0][0]) + 16 * int(n_pairs)
    result = {
        "value": n * timed / (c.ms * 1e-3),
        "unit": "box-frames/s",
        "frames": timed,
        "ms_per_frame": c.ms / timed,
        "dtype": "u32",
        "config": {"workload": "SAP broadphase only: 1,000,000 AABBs, 10% moved per frame, incremental SAP with "
                               "32-bit endpoint keys (found/lost reports)",
                   "boxes": n, "moved_per_frame": len(moves[0][0]), "live_pairs": int(n_pairs),
                   "radix_digit_bits": 8, "radix_passes": 4,
                   "endpoint_keys_per_s": keys * timed / (c.ms * 1e-3),
                   "frame_algorithmic_GBps": round(frame_bytes / (c.ms / timed * 1e-3) / 1e9, 1),
                   "frame_algorithmic_bytes": frame_bytes},
        "roofline": {"kernel": "sap.sort: endpoint radix sort (3 axes, 4 passes x tile histogram + tile scan + "
                               "scatter)",
                     "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": SAP_SORT_TRAFFIC_BYTES,
                     "algorithmic_bytes_per_launch": sort_bytes},
        "stages": stages,
    }
    return result, (n, half, pos0, moves)


def bench_batch(args, ctx, stream, steps=200, warmup=40):
    import torch
    from fyzz_amd import scenes
    from fyzz_amd.world import World

    n_scenes = 4096
    counts, half, pos, quat = scenes.batch_scenes(n_scenes)
    n = int(counts.sum())
    w = World(ctx, n + 1)
    worlds = []
    try:
        fid = w.add_static_floor(0.0)
        w.set_body_material(fid, 0.5, 0.5, 0.0)
        w.add_dynamic_boxes(half, pos, quat, friction=0.5, restitution=0.0)
        w.set_scenes(counts)
        # the warm-up first, so the autotuner times launch shapes on the
        # formed piles the timed steps run (its trials advance the world)
        progress(f"batch: {n_scenes} scenes, {n} boxes; {warmup} warm-up steps")
        for _ in range(warmup):
            w.step_graph(DT)
        progress("batch: autotuning launch shapes")
        tuned = w.autotune(DT, trial_steps=3)
        progress(f"batch: {steps} timed steps")
        warm = w.totals()["steps"]
        w.totals(reset=True)
        with Clock(stream) as c:
            for _ in range(steps):
                w.step_graph(DT)
        ms = c.ms
        st = w.stats()
        if st["flags"]:
            raise RuntimeError(f"world status flags {st['flags']:#x}")
        tot = w.totals()
        stages, ptot = profiled_steps(ctx, w, 5, PILE_STAGES + ("world.scene_pairs",), w.step)
        stages = {k: v for k, v in stages.items() if v["launches"]}
        # plain per-scene launches: a sample of scenes, one world each
        progress("batch: the same scenes as plain per-scene worlds")
        first = np.concatenate([[0], np.cumsum(counts.astype(np.int64))]).astype(np.int64)
        sample = np.linspace(0, n_scenes - 1, 32).astype(np.int64)
        for k in sample:
            a, b = first[k], first[k + 1]
            pw = World(ctx, int(b - a) + 1)
            worlds.append((pw, int(b - a)))
            f2 = pw.add_static_floor(0.0)
            pw.set_body_material(f2, 0.5, 0.5, 0.0)
            pw.add_dynamic_boxes(half[a:b], pos[a:b], quat[a:b], friction=0.5, restitution=0.0)
        plain_steps = 20
        for _ in range(warm):
            for pw, _nb in worlds:
                pw.step(DT)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(plain_steps):
            for pw, _nb in worlds:
                pw.step(DT)
        torch.cuda.synchronize()
        plain_s = time.perf_counter() - t0
        plain_value = sum(nb for _, nb in worlds) * plain_steps / plain_s
        value = n * steps / (ms * 1e-3)
        recs = w.body_records()
    finally:
        for pw, _ in worlds:
            pw.close()
        w.close()
    per = max(tot["steps"], 1)
    result = {
        "value": value,
        "unit": "body-steps/s",
        "steps": steps,
        "warmup": warm,
        "ms_per_step": ms / steps,
        "dtype": "f32",
        "config": {"workload": "batched: 4,096 independent piles (U{8..512} boxes, footprint scaled by "
                               "sqrt(n/1e5)) in one world, one launch per stage, one HIP graph replay per step",
                   "scenes": n_scenes, "boxes": n, "timed_steps": f"{warm}..{warm + steps - 1}",
                   "autotuned_workgroups": {"partition": tuned[0], "pgs": tuned[1]},
                   "managers_per_step": tot["managers"] / per, "pairs_per_step": tot["pairs"] / per,
                   "plain_per_scene_launches": {"value": plain_value, "unit": "body-steps/s",
                                                "sample": f"{len(worlds)} scenes, one world each, {plain_steps} "
                                                          f"steps after the same warm-up"},
                   "speedup_vs_plain": value / plain_value},
        "roofline": roofline_of(stages, ptot, n),
        "stages": stages,
    }
    cpu_sample = np.linspace(0, n_scenes - 1, 128).astype(np.int64)
    sample_recs = [recs[first[k]:first[k + 1]] for k in cpu_sample]
    sample_half = [half[first[k]:first[k + 1]] for k in cpu_sample]
    return result, (sample_recs, sample_half)


def bench_diff(args, ctx, stream):
    import torch
    from fyzz_amd import scenes
    from fyzz_amd.diff import Rollout, rollout_function

    n_scenes, steps = 1024, 120
    counts, half, pos, quat, v0, w0, mu, target = scenes.diff_scenes(n_scenes)
    n = int(counts.sum())
    dev = torch.device("cuda", 0)
    ro = Rollout(ctx, half, pos, quat, counts, steps, checkpoint_every=16)
    fn = rollout_function()
    tgt = torch.tensor(target, dtype=torch.float32, device=dev)

    def once():
        tv = torch.tensor(v0, device=dev, requires_grad=True)
        tw = torch.tensor(w0, device=dev, requires_grad=True)
        tm = torch.tensor(mu, device=dev, requires_grad=True)
        t0 = time.perf_counter()
        p = fn.apply(tv, tw, tm, ro)
        com = p.view(n_scenes, 64, 3).sum(1) / 64.0  # stacks of 64 boxes, fixed-order sums
        d = (com - tgt)[:, [0, 2]]
        loss = (d * d).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return loss.item(), t1 - t0, t2 - t1, tv.grad, tm.grad

    try:
        once()  # warm-up rollout (allocations, first-use paths)
        loss, tf, tb, gv, gm = once()
    finally:
        ro.close()
    total = tf + tb
    return {
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
        "cpu_baseline": None,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU only; --gpus must be 1 "
                                   "(independent scenes would shard by scene, one pile by spatial slabs)"}))
        sys.exit(2)
    if os.environ.get("WORLD_SIZE", "1") not in ("1", ""):
        print("bench.py runs on one GPU; WORLD_SIZE>1 is not supported", file=sys.stderr)
        sys.exit(2)
    import torch
    from fyzz_amd import Context

    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    ctx = Context(0, stream=stream.cuda_stream)
    try:
        if args.plain_launch:
            ctx.set_cooperative(False)
        if args.calibrate_copy:
            calibrate_copy(dev)
        if args.workload == "sap":
            frames = 60 if args.steps is None else args.steps + 1 + (9 if args.warmup is None else args.warmup)
            warm = 9 if args.warmup is None else args.warmup
            sub, (n, half, pos0, moves) = bench_sap(args, ctx, stream, frames=frames, warmup=warm)
            result = {"metric": METRIC, "value": sub["value"], "unit": sub["unit"], "n_gpus": 1,
                      "steps": sub["frames"], "warmup": warm, "ms_per_step": sub["ms_per_frame"],
                      "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
                      "data": "synthetic", "config": sub["config"], "roofline": sub["roofline"],
                      "stages": sub["stages"],
                      "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline_sap(n, half, pos0, moves)}
        elif args.workload == "diff":
            result = bench_diff(args, ctx, stream)
        elif args.workload == "batch":
            steps = 200 if args.steps is None else args.steps
            warm = 40 if args.warmup is None else args.warmup
            sub, (recs, halves) = bench_batch(args, ctx, stream, steps=steps, warmup=warm)
            result = {"metric": METRIC, "value": sub["value"], "unit": sub["unit"], "n_gpus": 1, "steps": steps,
                      "warmup": sub["warmup"], "ms_per_step": sub["ms_per_step"], "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                      "config": sub["config"], "roofline": sub["roofline"], "stages": sub["stages"],
                      "cpu_baseline": None if args.no_cpu_baseline else
                      cpu_baseline_scenes(recs, halves, args.cpu_seconds, steps)}
        else:
            args.steps = 100 if args.steps is None else args.steps
            args.warmup = 10 if args.warmup is None else args.warmup
            result, w, sc = bench_pile(args, ctx, stream)
            recs = w.body_records()
            w.close()
            if not args.no_extras:
                result["pile_native_order"] = bench_pile_native(ctx, stream, recs, sc.half)
                sub, (n, half, pos0, moves) = bench_sap(args, ctx, stream)
                sub["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline_sap(n, half, pos0, moves)
                result["sap_1m"] = sub
                sub, (brecs, bhalves) = bench_batch(args, ctx, stream)
                sub["cpu_baseline"] = None if args.no_cpu_baseline else \
                    cpu_baseline_scenes(brecs, bhalves, args.cpu_seconds, sub["steps"])
                result["batch_4096"] = sub
            result["cpu_baseline"] = None if args.no_cpu_baseline else \
                cpu_baseline_pile(recs, sc.half, args.cpu_seconds)
    finally:
        torch.cuda.synchronize()
        ctx.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
