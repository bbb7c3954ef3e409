"""Benchmark of the kNN hot path on a 2M-point map, on one MI355X, plus the
per-scan latency of the incremental mapping pipeline, a many-tree batching
workload and a many-robot serving workload.

Scene: the LiDAR-like mapping workload (ground + 300 boxes in a 400 x 400 m
area, 32 beams x 1800 azimuths per scan, sensor moving 1 m per scan on a
circle of radius 80 m, seed 1002, ~37-42k points per scan).

Phase A (the JSON value).  The map is a Build of raw scans from every 7th pose
along the path until it holds >= --map-points points (2.0M by default; run as
specified, 200 downsampled scans of this scene only reach ~0.2M points, so the
seed Build supplies the 2M-point map the metric is quoted on).  The queries are
the scans of the poses in between (each within 3 m of a mapped pose).  One
step = Nearest_Search (k=5, max_dist=5 m) of every point of --batch-scans such
scans against that map in one launch ("batched").  The same scans are also searched with one
launch per scan ("plain"); both forms must give identical results.  Queries
and outputs are resident in HBM before the timed region.  value = kNN
queries/s of the batched form over the K timed steps.

Roofline: the kNN kernel's algorithmic bytes per query
12 + 16 k_found + 48 V + 24 C (V nodes visited, C child boxes evaluated but not
visited, counted by the instrumented one-lane depth-first walk on the same
queries and map) over its average launch time from HIP events on the tree's
stream.

Phase B.  The mapping pipeline as specified: Build of scan 0, then for scans
1..--frames-1 kNN (k=5, max_dist=5 m) of the scan followed by Add_Points with
0.5 m voxel downsample.  p50/p99 per-scan latency (kNN + insert, synchronous).

Phase C.  --trees trees of log-uniform 10k-1M points with log-uniform 1k-200k
queries each, mixed k and max_dist, 20% radius queries: one multi-tree launch
against one call per tree and (k, max_dist) group.

Phase D.  --robots maps of 0.5-3M points served by the request loop with
Poisson scan arrivals at 10 Hz per robot: p50/p99 of completion - arrival.
Subtrees above --serve-defer live points are rebuilt while no request
waits, --serve-chunk points at a time; each finished request carries the
map's event counts (rebuilds, rebuilt points, allocations, host waits), and
the report gives p50/p99 per 5 s of arrivals.

The CPU baseline is the oracle (CPU restatement of the reference) on a bounded
sample of the same queries and map.
"""
from __future__ import annotations

import argparse
import faulthandler
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "kNN queries/s (k=5, 2M-pt map) + p99 scan latency, batched vs plain GPU"
SEED_STRIDE = 7  # map scans every 7th pose (1 m apart); queries from the poses in between
HBM_PEAK_GBS = 8000.0


def main():
    faulthandler.enable()  # a native crash prints the Python stack
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--map-points", type=int, default=2_000_000)
    ap.add_argument("--k", type=int, default=5)
    ap.add_argument("--max-dist", type=float, default=5.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-scans", type=int, default=2, help="scans of queries in the CPU sample")
    ap.add_argument("--count-scans", type=int, default=16,
                    help="scans whose reference-order traversal counts give the algorithmic bytes")
    ap.add_argument("--batch-scans", type=int, default=32, help="scans per launch for the batched form")
    ap.add_argument("--frames", type=int, default=200, help="scans in the mapping-pipeline phase")
    ap.add_argument("--trees", type=int, default=64, help="trees in the multi-tree batching phase (0 skips it)")
    ap.add_argument("--robots", type=int, default=16, help="maps in the serving phase (0 skips it)")
    ap.add_argument("--box-points", type=int, default=5_000_000,
                    help="map points of the box/radius/trim phase (0 skips it)")
    ap.add_argument("--serve-seconds", type=float, default=30.0, help="simulated arrival window of the serving phase")
    ap.add_argument("--serve-defer", type=int, default=128,
                    help="serving: subtrees above this many live points are rebuilt between requests")
    ap.add_argument("--serve-slots", type=int, default=8, help="serving: requests on the device at once")
    ap.add_argument("--serve-chunk", type=int, default=100_000,
                    help="serving: deferred-rebuild points per maintenance step (-1: all at once)")
    ap.add_argument("--lanes", type=int, default=0, help="force lanes per query (0: autotuned)")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default launch shape or --lanes)")
    ap.add_argument("--emu-lib", default=None, help=argparse.SUPPRESS)  # debugging: a CPU emulation build
    ap.add_argument("--scale", type=float, default=1.0, help=argparse.SUPPRESS)  # debugging: shrink phases C/D
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (10 launches), to calibrate HBM byte counters")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (maps would shard by tree / robot, "
              "one process per GPU). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_calibration:
        copy_calibration(torch, args.copy_calibration)
        return

    ikd = importlib.import_module("ikd-tree_amd")
    scenes = importlib.import_module("ikd-tree_amd.scenes")
    autotune = importlib.import_module("ikd-tree_amd.autotune")
    global SCALE
    SCALE = args.scale
    if args.emu_lib:
        # host-side debugging only: the same flow against the CPU emulation
        # build of the library, tensors on the host, wall-clock timing
        ikd.LIB_PATH = os.path.abspath(args.emu_lib)
        dev = torch.device("cpu")
        _emulate_cuda(torch)
    else:
        dev = torch.device("cuda:0")
        torch.cuda.set_device(dev)

    scene = scenes.MappingScene(1002)
    k = args.k

    # ================================================ phase A: 2M-point map
    from concurrent.futures import ThreadPoolExecutor

    B = max(1, args.batch_scans)
    total_steps = args.warmup + args.steps
    with ThreadPoolExecutor(max_workers=16) as pool:  # scan synthesis releases the GIL
        # the map: raw scans from every SEED_STRIDE-th pose until it holds
        # --map-points points; the queries: the scans of the poses in between,
        # each within SEED_STRIDE/2 m of a mapped pose, as a mapping pipeline
        # queries each new scan against the map of the scans before it
        seed, f, npts = [], 0, 0
        while npts < args.map_points:
            frames = list(range(f, f + 8 * SEED_STRIDE, SEED_STRIDE))
            for s_ in pool.map(scene.scan, frames):
                if npts < args.map_points:
                    seed.append(s_)
                    npts += s_.shape[0]
                    f += SEED_STRIDE
        inmap = [g for g in range(f) if g % SEED_STRIDE]
        # more scans than mapped poses in between: the steps reuse them in turn
        qframes = [inmap[i % len(inmap)] for i in range(total_steps * B)]
        uniq = sorted(set(qframes))
        by_frame = dict(zip(uniq, pool.map(scene.scan, uniq)))
        scans = [by_frame[g] for g in qframes]
    seed_pts = np.concatenate(seed)

    tree = ikd.KD_TREE(0.5, 0.6, 0.5)
    t0 = time.perf_counter()
    tree.Build(seed_pts)
    build_s = time.perf_counter() - t0
    map_points = tree.validnum()

    def outs_for(n):
        return (torch.empty((n, k, 3), dtype=torch.float32, device=dev),
                torch.empty((n, k), dtype=torch.float32, device=dev),
                torch.empty((n, k), dtype=torch.int32, device=dev),
                torch.empty((n,), dtype=torch.int32, device=dev))

    qd = [torch.from_numpy(x).to(dev) for x in scans]
    po = [outs_for(x.shape[0]) for x in scans]  # plain: one launch per scan
    bq = [torch.cat(qd[i * B:(i + 1) * B]) for i in range(total_steps)]  # batched: B scans per launch
    bo = [outs_for(x.shape[0]) for x in bq]
    ext = torch.cuda.ExternalStream(tree.stream())
    torch.cuda.synchronize()

    def tune(q):
        if args.no_tune:
            return (args.lanes if args.lanes else -1, 0, 0, 0, 1), {}
        best, ms = autotune.autotune_knn(tree, q, k, args.max_dist)
        return best, ms

    def apply(cfg):
        lanes = args.lanes if args.lanes else cfg[0]
        tree.set_tuning(ikd.IKD_TUNE_LANES_PER_QUERY, lanes)
        tree.set_tuning(ikd.IKD_TUNE_KNN_STACK, cfg[1])
        tree.set_tuning(ikd.IKD_TUNE_QUERY_ORDER, cfg[2])
        tree.set_tuning(ikd.IKD_TUNE_KNN_STAGE, cfg[3])
        tree.set_tuning(ikd.IKD_TUNE_KNN_COMPACT, cfg[4])
        return (lanes,) + tuple(cfg[1:])

    def timed_launches(queries, outs, idx):
        """Launch idx in order, each bracketed by HIP events on the tree stream;
        returns (wall seconds, per-launch ms)."""
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in idx]
        tree.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for e, i in zip(evs, idx):
            e[0].record(ext)
            tree.knn_device(queries[i], k, args.max_dist, *outs[i])
            e[1].record(ext)
        tree.synchronize()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return wall, [e[0].elapsed_time(e[1]) for e in evs]

    warm = list(range(args.warmup))
    timed = list(range(args.warmup, total_steps))
    timed_scans = [j for i in timed for j in range(i * B, (i + 1) * B)]

    # batched form (the value): B scans per launch
    b_tune, b_sweep = tune(bq[0])
    b_cfg = apply(b_tune)
    timed_launches(bq, bo, warm)
    b_wall, b_ms = timed_launches(bq, bo, timed)
    # plain form: one launch per scan, same scans
    p_tune, p_sweep = tune(qd[0])
    p_cfg = apply(p_tune)
    timed_launches(qd, po, list(range(B * args.warmup)))
    p_wall, p_ms = timed_launches(qd, po, timed_scans)
    identical = all(torch.equal(bo[i][m][off:off + qd[j].shape[0]], po[j][m])
                    for i in timed for m in (1, 2, 3)
                    for off, j in zip(np.cumsum([0] + [qd[j].shape[0] for j in range(i * B, (i + 1) * B)])[:-1],
                                      range(i * B, (i + 1) * B)))
    map_pool = tree.pool_stats()

    # algorithmic bytes: the one-lane depth-first walk (the reference's
    # traversal) of the same queries on the same map, counted on the device
    nq_total = sum(scans[j].shape[0] for j in timed_scans)
    found = sum(int(po[j][3].sum().item()) for j in timed_scans)
    vis_total = box_total = 0
    wave_max = 0
    for j in timed_scans:
        v, b_ = tree.knn_counters(scans[j], k, args.max_dist)
        v64 = v.astype(np.int64)
        vis_total += int(v64.sum())
        box_total += int(b_.astype(np.int64).sum())
        pad = (-v64.shape[0]) % 64
        wave_max += int(np.pad(v64, (0, pad)).reshape(-1, 64).max(axis=1).sum()) * 64
    unvisited = box_total - vis_total + nq_total  # C = E - (V - 1) per query
    alg_bytes = 12 * nq_total + 16 * found + 48 * vis_total + 24 * unvisited

    def roof(ms_list, launches):
        avg_s = (sum(ms_list) / len(ms_list)) / 1e3
        per_launch = alg_bytes / launches
        gbs = per_launch / avg_s / 1e9
        return gbs, per_launch, avg_s

    b_gbs, b_bpl, b_avg = roof(b_ms, len(timed))
    p_gbs, p_bpl, p_avg = roof(p_ms, len(timed_scans))
    plain = {
        "qps": nq_total / p_wall,
        "kernel_qps": nq_total / (sum(p_ms) / 1e3),
        "avg_launch_ms": p_avg * 1e3,
        "launches": len(timed_scans),
        "roofline_achieved_gbs": p_gbs,
        "roofline_frac": p_gbs / HBM_PEAK_GBS,
        "kernel": knn_kernel_name(k, p_cfg[0], p_cfg[3], p_cfg[4]),
        "tuned": {"lanes_per_query": p_cfg[0], "stack": p_cfg[1], "morton": p_cfg[2], "lds_levels": p_cfg[3],
                  "compact_stack": p_cfg[4]},
        "sweep_ms": {f"lanes{a}_stack{b}_order{c}_lds{d}_compact{e}": v for (a, b, c, d, e), v in p_sweep.items()},
    }

    progress("phase A done")
    # ================================================ phase B: mapping pipeline
    pipe = pipeline_latency(ikd, torch, dev, scene, args.frames, k, args.max_dist) if args.frames > 1 else None
    del tree
    torch.cuda.empty_cache()

    progress("phase B done")
    # ================================================ phase C: many trees, one launch
    batching = batching_phase(ikd, torch, dev, scenes, args.trees) if args.trees > 0 else None

    progress("phase C done")
    # ================================================ phase D: serving many robots
    serving = (serving_phase(ikd, scenes, args.robots, args.serve_seconds, args.serve_defer, args.serve_chunk,
                             args.serve_slots)
               if args.robots > 0 else None)

    progress("phase D done")
    # ================================================ phase E: box/radius/trims on a 5M map
    box_radius = box_radius_phase(ikd, scenes, args.box_points, cpu=not args.no_cpu) if args.box_points > 0 else None

    result = {
        "metric": METRIC,
        "value": nq_total / b_wall,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b_wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"batched Nearest_Search k={k}, max_dist={args.max_dist:g} m, of every point of {B} "
                         "LiDAR-like scans per step (one launch) against a 2M-point map (mapping scene seed 1002, "
                         f"32x1800-ray scans; map = raw scans of every {SEED_STRIDE}th pose, queries = scans of the "
                         "poses in between); the plain form (one launch per scan, same scans, identical results) "
                         "is reported under 'plain'"),
            "distinct_query_scans": len(set(qframes[j] for j in timed_scans)),
            "timed_query_scans": len(timed_scans),
            "map_points": int(map_points),
            "scans_per_step": B,
            "queries_per_step": nq_total / args.steps,
            "k": k,
            "max_dist_m": args.max_dist,
            "tuned": {"lanes_per_query": b_cfg[0], "stack": b_cfg[1], "morton": b_cfg[2], "lds_levels": b_cfg[3],
                      "compact_stack": b_cfg[4]},
            "sweep_ms": {f"lanes{a}_stack{b}_order{c}_lds{d}_compact{e}": v
                         for (a, b, c, d, e), v in b_sweep.items()},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": b_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": b_gbs / HBM_PEAK_GBS,
            "traffic": measured_traffic(),
            "kernel": knn_kernel_name(k, b_cfg[0], b_cfg[3], b_cfg[4]),
            "algorithmic_bytes_per_launch": b_bpl,
            "avg_launch_ms": b_avg * 1e3,
            "bytes_per_query": alg_bytes / nq_total,
            "visits_per_query": vis_total / nq_total,
            "unvisited_boxes_per_query": unvisited / nq_total,
            "simt_efficiency_one_lane": vis_total / max(wave_max, 1),
        },
        "knn_kernel_qps": nq_total / (sum(b_ms) / 1e3),
        "plain": plain,
        "identical_batched_plain": bool(identical),
        "pipeline": pipe,
        "batching": batching,
        "serving": serving,
        "box_radius_5M": box_radius,
        "build_2M_s": build_s,
        "pool": map_pool,
    }

    progress("phase E done")
    if not args.no_cpu:
        cpu_idx = timed_scans[: args.cpu_scans]
        gpu_out = tuple(np.concatenate([po[j][m].cpu().numpy() for j in cpu_idx]) for m in (1, 2, 3))
        gpu_out = (gpu_out[0], gpu_out[1].view(np.uint32), gpu_out[2])
        # reference-order V and E on a spread of the timed scans (every
        # count_stride-th distinct scan), scaled to the timed set per query
        distinct = sorted(set(qframes[j] for j in timed_scans))
        stride = max(1, len(distinct) // args.count_scans)
        count_frames = distinct[::stride][: args.count_scans]
        count_scans = [by_frame[g] for g in count_frames]
        cb = cpu_baseline(seed_pts, [scans[j] for j in cpu_idx], gpu_out, count_scans, k, args.max_dist)
        result["cpu_baseline"] = cb
        rw = cb["reference_walk"]
        v_q = rw["visits"] / rw["queries"]
        c_q = (rw["child_boxes"] - rw["visits"] + rw["queries"]) / rw["queries"]
        bpq = 12 + 16 * found / nq_total + 48 * v_q + 24 * c_q
        roof_r = result["roofline"]
        roof_r["device_walk"] = {k_: roof_r.pop(k_) for k_ in ("achieved", "frac", "algorithmic_bytes_per_launch",
                                                                  "bytes_per_query", "visits_per_query",
                                                                  "unvisited_boxes_per_query")}
        per_launch = bpq * nq_total / len(timed)
        roof_r.update({
            "achieved": per_launch / b_avg / 1e9,
            "frac": per_launch / b_avg / 1e9 / HBM_PEAK_GBS,
            "algorithmic_bytes_per_launch": per_launch,
            "bytes_per_query": bpq,
            "visits_per_query": v_q,
            "unvisited_boxes_per_query": c_q,
            "counts": (f"V and C from the oracle's reference-order walk on {rw['queries']} queries "
                       f"({rw['scans']} of the {len(distinct)} distinct timed scans, evenly spaced); "
                       "k_found from the device results of every timed query"),
        })
        result["value_check"] = {"identical_batched_plain": bool(identical),
                                 "gpu_vs_oracle_mismatches": cb["mismatches"]}
    print(json.dumps(result))


SCALE = 1.0


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _emulate_cuda(torch):
    """Stand-ins for the few torch.cuda calls of this script on a host-only run."""
    class _Ev:
        def __init__(self, enable_timing=True):
            self.t = 0.0

        def record(self, stream=None):
            self.t = time.perf_counter()

        def elapsed_time(self, other):
            return (other.t - self.t) * 1e3

    torch.cuda.Event = _Ev
    torch.cuda.ExternalStream = lambda s: None
    torch.cuda.synchronize = lambda *a: None
    torch.cuda.empty_cache = lambda *a: None


def _scene_points(scenes, seed, first_frame, n, pool):
    """The first n points of consecutive raw scans of scene `seed` from first_frame."""
    sc = scenes.MappingScene(seed)
    out, have, f = [], 0, first_frame
    while have < n:
        batch = list(pool.map(sc.scan, range(f, f + 4)))
        for b in batch:
            out.append(b)
            have += b.shape[0]
        f += 4
    return np.concatenate(out)[:n], f


def batching_phase(ikd, torch, dev, scenes, n_trees, seed=1004):
    """Many trees of uneven size (log-uniform 10k-1M points, each its own
    mapping scene), uneven query sets (log-uniform 1k-200k points of fresh
    scans), mixed k in {1,5,10,20} and max_dist in {1,5,inf} m, plus 20% radius
    queries (r 0.2-2 m).  Batched: one multi-tree kNN launch and one multi-tree
    radius launch.  Plain: one Nearest_Search call per tree and (k, max_dist)
    group.  Both give identical kNN results; the wall time of each is the
    median of 5 runs with all inputs on the device."""
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(seed)
    sizes = (SCALE * np.exp(rng.uniform(np.log(1e4), np.log(1e6), n_trees))).astype(np.int64) + 1
    nqs = (SCALE * np.exp(rng.uniform(np.log(1e3), np.log(2e5), n_trees))).astype(np.int64) + 1
    trees, qparts = [], []
    with ThreadPoolExecutor(max_workers=16) as pool:
        for t in range(n_trees):
            pts, f = _scene_points(scenes, seed * 100 + t, 0, int(sizes[t]), pool)
            q, _ = _scene_points(scenes, seed * 100 + t, f, int(nqs[t]), pool)
            tr = ikd.KD_TREE(0.5, 0.6, 0.5)
            tr.Build(pts)
            trees.append(tr)
            qparts.append(q)
    qt = np.concatenate([np.full(len(q), t, np.int32) for t, q in enumerate(qparts)])
    q = np.concatenate(qparts)
    n = q.shape[0]
    is_rad = rng.random(n) < 0.2
    ks = rng.choice(np.array([1, 5, 10, 20], np.int32), n)
    mds = rng.choice(np.array([1.0, 5.0, np.inf]), n)
    rad = (0.2 + 1.8 * rng.random(n)).astype(np.float32)
    kn = np.nonzero(~is_rad)[0]
    rq = np.nonzero(is_rad)[0]
    kmax = 20
    d_qt = torch.from_numpy(qt[kn]).to(dev)
    d_q = torch.from_numpy(q[kn]).to(dev)
    d_k = torch.from_numpy(ks[kn]).to(dev)
    d_md = torch.from_numpy(mds[kn]).to(dev)
    nk = kn.size
    bo = (torch.empty((nk, kmax, 3), dtype=torch.float32, device=dev),
          torch.empty((nk, kmax), dtype=torch.float32, device=dev),
          torch.empty((nk, kmax), dtype=torch.int32, device=dev),
          torch.empty((nk,), dtype=torch.int32, device=dev))
    batch = ikd.Batch()

    def run_batched():
        batch.nearest_search_device(trees, d_qt, d_q, kmax, *bo, k=d_k, max_dist=d_md)
        batch.synchronize()

    sweep = {}
    for lanes in (1, 2, 4, 0):
        batch.set_tuning(ikd.IKD_TUNE_LANES_PER_QUERY, lanes)
        run_batched()
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            run_batched()
            ts.append(time.perf_counter() - t0)
        sweep[lanes] = float(np.median(ts))
    best = min(sweep, key=sweep.get)
    batch.set_tuning(ikd.IKD_TUNE_LANES_PER_QUERY, best)
    bt = []
    for _ in range(5):
        t0 = time.perf_counter()
        run_batched()
        bt.append(time.perf_counter() - t0)
    # the same launch recorded into a HIP graph and replayed
    go_ = tuple(torch.empty_like(x) for x in bo)
    graph = batch.knn_graph(trees, d_qt, d_q, kmax, *go_, k=d_k, max_dist=d_md)
    graph.launch()
    batch.synchronize()
    gt = []
    for _ in range(5):
        t0 = time.perf_counter()
        graph.launch()
        batch.synchronize()
        gt.append(time.perf_counter() - t0)
    graph_same = all(torch.equal(a_, b_) for a_, b_ in zip(go_, bo))
    graph.close()

    # plain: one call per tree and (k, max_dist) group, on each tree's stream
    groups = []
    kq = qt[kn]
    for t in range(n_trees):
        for kk in (1, 5, 10, 20):
            for md in (1.0, 5.0, np.inf):
                sel = np.nonzero((kq == t) & (ks[kn] == kk) & (mds[kn] == md))[0]
                if sel.size == 0:
                    continue
                gq = d_q[torch.from_numpy(sel).to(dev)].contiguous()
                go = (torch.empty((sel.size, kk, 3), dtype=torch.float32, device=dev),
                      torch.empty((sel.size, kk), dtype=torch.float32, device=dev),
                      torch.empty((sel.size, kk), dtype=torch.int32, device=dev),
                      torch.empty((sel.size,), dtype=torch.int32, device=dev))
                groups.append((t, kk, md, sel, gq, go))
    torch.cuda.synchronize()

    def run_plain():
        for (t, kk, md, sel, gq, go) in groups:
            trees[t].knn_device(gq, kk, md, *go)
        for tr in trees:
            tr.synchronize()

    run_plain()
    pt = []
    for _ in range(5):
        t0 = time.perf_counter()
        run_plain()
        pt.append(time.perf_counter() - t0)
    same = True
    bd, bi, bc = bo[1].cpu(), bo[2].cpu(), bo[3].cpu()
    for (t, kk, md, sel, gq, go) in groups:
        s_ = torch.from_numpy(sel)
        same &= torch.equal(bc[s_], go[3].cpu()) and torch.equal(bd[s_][:, :kk], go[1].cpu()) and \
            torch.equal(bi[s_][:, :kk], go[2].cpu())

    # radius queries of the same mix, one multi-tree launch
    d_rq = torch.from_numpy(qt[rq]).to(dev)
    d_rc = torch.from_numpy(q[rq]).to(dev)
    d_rr = torch.from_numpy(rad[rq]).to(dev)
    off = torch.empty(rq.size + 1, dtype=torch.int64, device=dev)
    total = batch.radius_search_device(trees, d_rq, d_rc, d_rr, off)
    rx = torch.empty((max(total, 1), 3), dtype=torch.float32, device=dev)
    ri = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
    rt = []
    for _ in range(3):
        t0 = time.perf_counter()
        batch.radius_search_device(trees, d_rq, d_rc, d_rr, off, rx, ri)
        batch.synchronize()
        rt.append(time.perf_counter() - t0)
    b_s, p_s = float(np.median(bt)), float(np.median(pt))
    out = {
        "trees": n_trees,
        "tree_points_total": int(sizes.sum()),
        "knn_queries": int(nk),
        "radius_queries": int(rq.size),
        "radius_results": int(total),
        "knn_launches_plain": len(groups),
        "batched_knn_s": b_s,
        "graph_replay_knn_s": float(np.median(gt)),
        "graph_replay_identical": bool(graph_same),
        "plain_knn_s": p_s,
        "batched_knn_qps": nk / b_s,
        "plain_knn_qps": nk / p_s,
        "speedup": p_s / b_s,
        "identical_to_plain": bool(same),
        "batched_lanes": best,
        "lanes_sweep_s": {str(k_): v for k_, v in sweep.items()},
        "batched_radius_s": float(np.median(rt)),
    }
    for tr in trees:
        tr.close()
    batch.close()
    return out


def serving_phase(ikd, scenes, n_robots, seconds, defer_points, chunk_points, slots, seed=1005):
    """Many robots, each with its own device-resident map (its own mapping
    scene, 0.5-3M points), sending scans (10k-60k points) as Poisson arrivals at
    10 Hz; each request is kNN (k=5, max_dist 5 m) + Add_Points (0.5 m voxels),
    every 10th also a sliding-window trim.  p50/p99 of completion - arrival."""
    from concurrent.futures import ThreadPoolExecutor

    serving = importlib.import_module("ikd-tree_amd.serving")
    rng = np.random.default_rng(seed)
    sizes = (SCALE * rng.uniform(0.5e6, 3e6, n_robots)).astype(np.int64) + 1
    scs = [scenes.MappingScene(seed * 100 + r) for r in range(n_robots)]
    trees, first = [], []
    with ThreadPoolExecutor(max_workers=16) as pool:
        for r in range(n_robots):
            pts, f = _scene_points(scenes, seed * 100 + r, 0, int(sizes[r]), pool)
            tr = ikd.KD_TREE(0.5, 0.6, 0.5)
            tr.Build(pts)
            trees.append(tr)
            first.append(f)
    scan_len = rng.integers(10_000, 60_001, size=(n_robots, 4096))
    srv = serving.Server(slots=slots, hbm_budget_bytes=64 << 30)
    # rebuilds of subtrees above defer_points run between a map's requests,
    # chunk_points at a time
    ids = [srv.add_map(t, defer_rebuild_points=defer_points) for t in trees]
    srv.set_maintenance(chunk_points)

    # one full scan per robot before the clock starts: every map's scratch
    # buffers reach their working size (allocation is not part of serving)
    for r in range(n_robots):
        srv.submit(ids[r], scs[r].scan(first[r]), 5, 5.0, 1, None, tag=r)
    srv.drain()

    def scan_fn(r, j):
        s = scs[r].scan(first[r] + 1 + j)
        return s[: scan_len[r, j % 4096]]

    def trim_fn(r, j):
        p = scs[r].pose(first[r] + 1 + j)
        big, h = 1e7, 100.0
        lo, hi = p - h, p + h
        return np.array([[-big, -big, -big, lo[0], big, big], [hi[0], -big, -big, big, big, big],
                         [-big, -big, -big, big, lo[1], big], [-big, hi[1], -big, big, big, big],
                         [-big, -big, -big, big, big, lo[2]], [-big, -big, hi[2], big, big, big]], np.float32)

    res = serving.simulate(srv, ids, scan_fn, rate_hz=10.0, duration_s=seconds, k=5, max_dist=5.0,
                           trim_fn=trim_fn, trim_every=10, seed=seed)
    res["deferred_rebuilds_run"] = srv.maintenance_rebuilds()
    srv.close()
    for t in trees:
        t.close()
    res.update({"robots": n_robots, "map_points_total": int(sizes.sum()), "slots": slots, "hbm_budget_gb": 64,
                "rate_hz_per_robot": 10.0, "defer_rebuild_points": defer_points,
                "maintenance_chunk_points": chunk_points})
    return res


def _trim_slabs(pose, half):
    """Six half-open slabs: everything outside a cube of side 2*half around pose."""
    big = 1e7
    lo, hi = pose - half, pose + half
    return np.array([[-big, -big, -big, lo[0], big, big], [hi[0], -big, -big, big, big, big],
                     [-big, -big, -big, big, lo[1], big], [-big, hi[1], -big, big, big, big],
                     [-big, -big, -big, big, big, lo[2]], [-big, -big, hi[2], big, big, big]], np.float32)


def box_radius_phase(ikd, scenes, n_points, frames=20, trim_every=10, seed=1003, reps=5, cpu=True):
    """Box/radius and deletion stress on one map: raw scans of mapping scene
    `seed` up to n_points points, Build, then 1000 Box_Search boxes (centred on
    map points, sides log-uniform 1-20 m) and 1000 Radius_Search balls (centred
    on map points, r uniform 0.2-3 m), each set in one call (count pass, scan,
    fill pass, results copied back to the host): median wall time of `reps`
    calls.  Then `frames` mapping frames (kNN k=5, max_dist 5 m of the scan,
    Add_Points with 0.5 m voxels) with Delete_Point_Boxes of everything outside
    a 200 m cube around the sensor every trim_every frames: wall time per frame
    and per trim.  With cpu, the oracle runs a sample of the same boxes and
    balls (cpu_range_baseline)."""
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(seed)
    sc = scenes.MappingScene(seed)
    with ThreadPoolExecutor(max_workers=16) as pool:
        pts, f = _scene_points(scenes, seed, 0, int(SCALE * n_points), pool)
        frame_scans = list(pool.map(sc.scan, range(f, f + frames)))
    n = pts.shape[0]
    g = ikd.KD_TREE(0.5, 0.6, 0.5)
    t0 = time.perf_counter()
    g.Build(pts)
    build_s = time.perf_counter() - t0
    c = pts[rng.integers(0, n, 1000)]
    side = np.exp(rng.uniform(0.0, np.log(20.0), (1000, 3))).astype(np.float32)
    boxes = np.concatenate([c - side / 2, c + side / 2], axis=1).astype(np.float32)
    cen = pts[rng.integers(0, n, 1000)]
    rad = (0.2 + 2.8 * rng.random(1000)).astype(np.float32)

    def timed(fn):
        out = fn()  # warm-up: scratch reaches its working size
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            out = fn()
            ts.append(time.perf_counter() - t)
        return float(np.median(ts)), out

    box_s, box_out = timed(lambda: g.box_search_batch(boxes))
    rad_s, rad_out = timed(lambda: g.radius_search_batch(cen, rad))
    n_box, n_rad = int(box_out[0][-1]), int(rad_out[0][-1])
    res = {
        "map_points": int(n),
        "build_s": build_s,
        "box": {"queries": 1000, "results": n_box, "call_s": box_s, "queries_per_s": 1000 / box_s,
                "result_points_per_s": n_box / box_s},
        "radius": {"queries": 1000, "results": n_rad, "call_s": rad_s, "queries_per_s": 1000 / rad_s,
                   "result_points_per_s": n_rad / rad_s},
    }
    if cpu:
        res["cpu_baseline"] = cpu_range_baseline(pts, boxes, cen, rad, box_out, rad_out)
    frame_ms, trim_ms, trimmed = [], [], []
    rb0 = g.pool_stats()["rebuilds"]
    for i, s in enumerate(frame_scans):
        t = time.perf_counter()
        g.knn(s, 5, 5.0)
        g.Add_Points(s, True)
        frame_ms.append((time.perf_counter() - t) * 1e3)
        if i % trim_every == trim_every - 1:
            slabs = _trim_slabs(sc.pose(f + i), 100.0)
            t = time.perf_counter()
            trimmed.append(int(g.Delete_Point_Boxes(slabs)))
            trim_ms.append((time.perf_counter() - t) * 1e3)
    res["mapping"] = {"frames": frames, "frame_ms": {"p50": float(np.median(frame_ms)), "max": float(max(frame_ms))},
                      "trim_ms": trim_ms, "trimmed_points": trimmed,
                      "rebuilds": int(g.pool_stats()["rebuilds"] - rb0), "validnum_end": int(g.validnum())}
    g.close()
    return res


def cpu_range_baseline(pts, boxes, cen, rad, box_out, rad_out, sample=100):
    """The oracle on the host: Build of the same map (not timed), then the
    first `sample` boxes and balls one Box_Search / Radius_Search call each
    (ikd_Tree.cpp:399-411), timed, and their results compared with the GPU's
    (ids as sets; radius ids within the reference's sqrt rounding band of the
    radius may differ, counted separately)."""
    from oracle import OracleTree

    o = OracleTree(0.5, 0.6, 0.5)
    ids = np.arange(pts.shape[0], dtype=np.uint32)
    o.Build(pts, ids)
    boff, _, bids = box_out
    roff, _, rids = rad_out
    t = time.perf_counter()
    ob = [o.Box_Search(boxes[i]) for i in range(sample)]
    box_s = time.perf_counter() - t
    t = time.perf_counter()
    orr = [o.Radius_Search(cen[i], float(rad[i])) for i in range(sample)]
    rad_s = time.perf_counter() - t
    box_bad = sum(set(ob[i][1].tolist()) != set(bids[boff[i]:boff[i + 1]].tolist()) for i in range(sample))
    rad_bad = band = 0
    for i in range(sample):
        a, b = set(orr[i][1].tolist()), set(rids[roff[i]:roff[i + 1]].tolist())
        if a != b:
            d = np.sqrt(((pts[list(a ^ b)].astype(np.float64) - cen[i]) ** 2).sum(axis=1))
            if np.all(np.abs(d - rad[i]) <= 1e-5 * max(1.0, float(rad[i]))):
                band += 1
            else:
                rad_bad += 1
    return {"kind": "port", "cores": 1, "sample": f"first {sample} boxes and {sample} balls of the phase",
            "box_queries_per_s": sample / box_s, "radius_queries_per_s": sample / rad_s,
            "box_mismatches": int(box_bad), "radius_mismatches": int(rad_bad),
            "radius_band_differences": int(band)}


def copy_calibration(torch, mib):
    """A plain device-to-device copy of known size: its FETCH_SIZE / WRITE_SIZE
    counters per dispatch give the scale factor from counter units to bytes."""
    n = mib * (1 << 20) // 4
    a = torch.arange(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(10):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 10}))


def measured_traffic():
    """HBM bytes per batched launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (profiles/knn_2M_map_hbm_traffic.md),
    or None when that summary is absent."""
    p = os.path.join(ROOT, "profiles", "knn_2M_map_hbm_traffic.md")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def knn_kernel_name(k, lanes, lds_levels=0, compact=1):
    """The kernel symbol a launch shape runs (as rocprofv3 names it); the
    compact-stack form applies on maps of at most 2^24 node slots."""
    if lanes == 1:
        if lds_levels > 0:
            return f"ikd::k_knn_staged<{knn_bucket(k)}>"
        return f"ikd::k_knn<{knn_bucket(k)}, {'true' if compact else 'false'}>"
    if lanes == 0:
        return f"ikd::k_knn_persist<{knn_bucket(k)}>"
    if lanes < 0:
        return f"ikd::k_knn*<{knn_bucket(k)}> (lanes per query from the batch size)"
    return f"ikd::k_knn_team<{knn_bucket(k)}, {lanes}, {'true' if compact else 'false'}>"


def knn_bucket(k):
    for b in (1, 2, 4, 5, 8, 10, 16, 20, 32, 64):
        if k <= b:
            return b
    return k


def pipeline_latency(ikd, torch, dev, scene, frames, k, max_dist, defer_points=1500):
    """Mapping pipeline as specified: Build(scan 0), then per scan kNN + Add_Points(0.5 m).

    Two forms on the same scans.  'inline' rebuilds every unbalanced subtree
    inside the Add_Points call.  'deferred' is the reference's two-thread split
    (ikd_Tree.cpp:627-633: subtrees of >= Multi_Thread_Rebuild_Point_Num = 1500
    points go to Rebuild_Thread): subtrees above defer_points live points are
    listed instead and rebuilt by rebuild_deferred between scans, timed apart
    as 'maintenance'.  The kNN results of every scan and the final live point
    sets of the two forms are compared."""
    scans = [scene.scan(i) for i in range(frames)]
    qd = [torch.from_numpy(x).to(dev) for x in scans]
    n_max = max(x.shape[0] for x in scans)
    outs = [tuple(torch.empty(shape, dtype=dt, device=dev) for shape, dt in
                  (((n_max, k, 3), torch.float32), ((n_max, k), torch.float32), ((n_max, k), torch.int32),
                   ((n_max,), torch.int32))) for _ in range(2)]
    digests = {}

    def run(form):
        tree = ikd.KD_TREE(0.5, 0.6, 0.5)
        tree.set_tuning(ikd.IKD_TUNE_QUERY_ORDER, 0)
        if form == "deferred":
            tree.set_tuning(ikd.IKD_TUNE_DEFER_REBUILD, defer_points)
        tree.Build(scans[0])
        o = outs[0 if form == "inline" else 1]
        torch.cuda.synchronize()
        lat, knn_lat, ins_lat, maint, added, nq, n_defer = [], [], [], [], 0, 0, 0
        dig = []
        t_start = time.perf_counter()
        for i in range(1, frames):
            n = qd[i].shape[0]
            ts = time.perf_counter()
            tree.knn_device(qd[i], k, max_dist, o[0][:n], o[1][:n], o[2][:n], o[3][:n])
            tree.synchronize()
            tk = time.perf_counter()
            added += tree.add_points_device(qd[i], True)
            te = time.perf_counter()
            if form == "deferred":
                n_defer += tree.rebuild_deferred(0)
                maint.append(time.perf_counter() - te)
            lat.append(te - ts)
            knn_lat.append(tk - ts)
            ins_lat.append(te - tk)
            nq += n
            # a digest of this scan's kNN distances and counts (ids differ between forms: rebuilds reorder nodes)
            dig.append(torch.stack([o[1][:n].double().sum(), o[3][:n].double().sum()]))
        total = time.perf_counter() - t_start
        digests[form] = torch.stack(dig).cpu().numpy()
        live = tree.flatten()[0]
        live = live[np.lexsort(live.T[::-1])]
        lat_ms = np.array(lat) * 1e3
        r = {
            "map_points_end": int(tree.validnum()),
            "scan_latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99)),
                                "max": float(lat_ms.max()), "mean": float(lat_ms.mean())},
            "knn_ms_p50": float(np.percentile(np.array(knn_lat) * 1e3, 50)),
            "insert_ms_p50": float(np.percentile(np.array(ins_lat) * 1e3, 50)),
            "insert_ms_p99": float(np.percentile(np.array(ins_lat) * 1e3, 99)),
            "queries_per_s": nq / total,
            "inserted_points_per_s": nq / sum(ins_lat),
            "add_points_returned": int(added),
            "pool": tree.pool_stats(),
            "counters": tree.counters(),
        }
        if form == "deferred":
            m_ms = np.array(maint) * 1e3
            r["maintenance_ms"] = {"p50": float(np.percentile(m_ms, 50)), "p99": float(np.percentile(m_ms, 99)),
                                   "max": float(m_ms.max()), "mean": float(m_ms.mean())}
            r["deferred_rebuilds"] = int(n_defer)
            r["defer_above_live_points"] = int(defer_points)
        tree.close()
        return r, live

    inline, live_a = run("inline")
    deferred, live_b = run("deferred")
    res = {"frames": frames, **inline, "deferred": deferred,
           "identical_live_sets": bool(live_a.shape == live_b.shape and np.array_equal(live_a, live_b)),
           "identical_knn_digests": bool(np.array_equal(digests["inline"], digests["deferred"]))}
    return res


def cpu_baseline(seed_pts, cpu_scans, gpu_out, count_scans, k, max_dist, min_seconds=10.0):
    """The oracle (CPU restatement of the reference) on the host cores: Build of
    the same seed map (not timed), then kNN of a bounded sample of scans,
    repeated until min_seconds of single-thread work.  The same leg checks the
    GPU's results for that sample against the oracle's (squared distances and
    ids; entries within the reference comparator's 1e-10 window of the k-th
    distance may be exchanged, ikd_Tree.h:101) and counts, with the oracle's
    reference-order walk (Search, ikd_Tree.cpp:868-1013), the nodes visited (V)
    and child boxes evaluated (E) for count_scans, which the roofline's
    algorithmic bytes use."""
    from oracle import OracleTree

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from common import knn_equivalent

    ora = OracleTree(0.5, 0.6, 0.5)
    ora.Build(seed_pts)
    q = np.concatenate(cpu_scans)
    ncpu = min(os.cpu_count() or 1, 16)  # one GPU's share of the host
    # parity of the timed GPU results on the sample
    _, od, oi, oc = ora.knn(q, k, max_dist, threads=ncpu)
    gd, gi, gc = gpu_out
    loose = knn_equivalent(gd, gi, gc, od, oi, oc)
    strict = knn_equivalent(gd, gi, gc, od, oi, oc, strict_ties=True)
    # reference-order traversal counts
    vis = box = nq = found = 0
    for sc in count_scans:
        _, _, _, c_, v_, b_ = ora.knn(sc, k, max_dist, counters=True, threads=ncpu)
        vis += int(v_.sum())
        box += int(b_.sum())
        nq += sc.shape[0]
        found += int(c_.sum())

    def timed(threads, seconds):
        n, t0 = 0, time.perf_counter()
        while True:
            ora.knn(q, k, max_dist, threads=threads)
            n += q.shape[0]
            dt = time.perf_counter() - t0
            if dt >= seconds:
                return n / dt, n

    v1, n1 = timed(1, min_seconds)
    vn, _ = timed(ncpu, min_seconds / 2)
    return {
        "value": v1,
        "unit": "queries/s",
        "cores": 1,
        "kind": "port",
        "sample": f"kNN k={k} max_dist={max_dist} m of the {q.shape[0]} points of {len(cpu_scans)} scans, repeated "
                  f"({n1} queries, >= {min_seconds:.0f} s), against the {seed_pts.shape[0]}-point seed map; oracle "
                  f"on 1 host thread",
        "all_cores_value": vn,
        "all_cores": ncpu,
        "parity_queries": int(q.shape[0]),
        "mismatches": len(loose),
        "near_tie_differences": len(strict) - len(loose),
        "reference_walk": {"queries": nq, "scans": len(count_scans), "visits": vis, "child_boxes": box,
                           "found": found},
    }


if __name__ == "__main__":
    main()
