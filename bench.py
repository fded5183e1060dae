#!/usr/bin/env python3
"""Benchmark of the reconstruct + CCL hot path on one MI355X.

Workload (default): one 4096x4096 u8 tile from the disk generator (3000 disks,
seed 0x5EED0001), marker = mask - 30, inputs resident in HBM.  One step =
    imreconstruct (8-conn) -> imfillHoles (grayscale, 4-conn) -> (> 128)
    -> bwlabel (8-conn) -> bwareaopen [11, 1000) -> canonical relabel
through the C ABI (nsx_segment_chain).  Reported value: Mpixel/s over the
timed steps.  The CPU baseline leg runs the oracle's restatement of the
reference chain on the same input and also compares it bit-exactly with the
GPU's first step.

Prints ONE JSON line.  --gpus N > 1 is refused: this path runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "emory-cci-fast-ia-gsoc_amd"))
sys.path.insert(0, str(REPO))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--size", type=int, default=4096)
    p.add_argument("--disks", type=int, default=3000)
    p.add_argument("--no-cpu", "--no-check", dest="no_cpu", action="store_true",
                   help="skip the CPU baseline leg (the oracle run that also checks the GPU results)")
    p.add_argument("--cpu-runs", type=int, default=20, help="oracle chain runs in the CPU sample")
    p.add_argument("--workload", default="chain4096",
                   choices=["chain4096", "batch", "maze", "maze_levels", "serve", "copy", "ops"],
                   help="chain4096: the headline 4096^2 chain (default); batch: 256 u8 + 32 f32 uneven tiles "
                        "batched vs per-tile; maze: 16384^2 maze stress; maze_levels: the same maze with "
                        "corridors at 150..220 plus noise (propagation's worst case); serve: continuous-batching "
                        "server; "
                        "ops: seeded imfill, bwselect, imhmin, localMaxima/Minima, boundingBox and "
                        "distanceTransform, each on the 4096^2 disk image")
    p.add_argument("--drain", type=int, default=0, choices=[0, 1],
                   help="chain: run each reconstruction's propagation rounds in one persistent launch")
    p.add_argument("--tiles-u8", type=int, default=256)
    p.add_argument("--tiles-f32", type=int, default=32)
    p.add_argument("--tune-reps", type=int, default=2,
                   help="batch: graph replays per setting tried by the autotuner (0: no autotuning)")
    p.add_argument("--check-tiles", type=int, default=288,
                   help="batch: tiles run by the CPU baseline (all host threads) and checked against the oracle")
    p.add_argument("--maze-size", type=int, default=16384, help="maze: side of the maze image")
    p.add_argument("--maze-cpu-size", type=int, default=16384,
                   help="maze: side of the maze the CPU baseline runs (default: the whole workload)")
    p.add_argument("--maze-levels-cpu-size", type=int, default=4096,
                   help="maze_levels: side of the many-valued maze the CPU baseline runs")
    p.add_argument("--requests", type=int, default=200, help="serve: requests in the trace")
    p.add_argument("--slots", type=int, default=4, help="serve: worker slots of the server")
    p.add_argument("--serve-sides", default="256,4096", help="serve: smallest,largest request side")
    p.add_argument("--serve-reps", type=int, default=5, help="serve: saturation bursts per multi-slot server")
    p.add_argument("--share-max-px", type=int, default=0,
                   help="serve: largest request that joins a shared set (0: the server's default)")
    p.add_argument("--share-set-px", type=int, default=0, help="serve: pixels per shared set (0: default)")
    p.add_argument("--max-batch", type=int, default=16,
                   help="serve: requests a slot takes at once (those of one kind share launches)")
    p.add_argument("--load", type=float, default=1.0,
                   help="serve: Poisson arrival rate as a multiple of the sequential capacity")
    p.add_argument("--budget-gib", type=float, default=4.0, help="serve: HBM arena of the server (GiB)")
    p.add_argument("--check-max-side", type=int, default=4096,
                   help="serve: requests up to this side are also checked against the oracle")
    return p.parse_args()


def batch_tiles(nh, n_u8: int, n_f32: int):
    """The batch workload's tiles (nscale_hip.workloads.batch_tiles, seed
    0x5EED0002 + index)."""
    from nscale_hip.workloads import batch_tiles as gen
    return gen(n_u8, n_f32)


def bench_batch(args):
    """All tiles through one set of launches (graph replay) vs one call per tile."""
    import numpy as np
    import torch

    import nscale_hip as nh
    t0 = time.perf_counter()
    tiles = batch_tiles(nh, args.tiles_u8, args.tiles_f32)
    log(f"generated {len(tiles)} tiles in {time.perf_counter() - t0:.1f} s")
    dev = torch.device("cuda:0")
    dtiles = []
    for marker, mask, thr in tiles:
        dm = torch.from_numpy(marker).to(dev)
        dk = torch.from_numpy(mask).to(dev)
        dtiles.append((dm, dk, torch.empty_like(dk), torch.empty(dk.shape, dtype=torch.int32, device=dev),
                       torch.empty_like(dk), torch.empty(dk.shape, dtype=torch.int32, device=dev), thr))
    total_px = sum(m.size for _, m, _ in tiles)
    ctx = nh.default_context(0)
    # per-tile launches, with the same algorithm as the batch: the level-set
    # reconstruction (whose value-set scan costs a synchronisation per call)
    # is switched off, as the batched path never takes it
    ctx_tile = nh.Context(0)
    ctx_tile.set_tuning(nh._lib.NSX_TUNE_LEVELSET_LEVELS, 0)

    def per_tile():
        return [nh.recon_label(dm, dk, 8, thr, 8, recon_out=r2, labels_out=l2, ctx=ctx_tile)[2]
                for dm, dk, _, _, r2, l2, thr in dtiles]
    per_tile()
    torch.cuda.synchronize()
    reps = max(1, args.steps // 4)
    t0 = time.perf_counter()
    for _ in range(reps):
        n_tile = per_tile()
    torch.cuda.synchronize()
    t_tile = (time.perf_counter() - t0) / reps
    # batched: one set of launches, captured once
    b = nh.Batch(ctx)
    for dm, dk, r, lab, _, _, thr in dtiles:
        b.add(dm, dk, r, lab, 8, thr, 8)
    t0 = time.perf_counter()
    b.finalize(use_graph=True)
    t_fin = time.perf_counter() - t0

    def replay_ms(reps):
        b.run()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            b.run()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps * 1e3
    # the fixed launch shapes before tuning: every group drained in one
    # persistent launch; one launch per round with the groups in turn; and
    # with the odd group on a second stream (the default)
    b.set_drain(True)
    shape_ms = {"drain": replay_ms(max(3, args.steps // 4))}
    b.set_drain(False)
    b.set_overlap(False)
    shape_ms["per_round_serial"] = replay_ms(max(3, args.steps // 4))
    b.set_overlap(True)
    shape_ms["per_round_overlap"] = replay_ms(max(3, args.steps // 4))
    # per tile group, a coordinate descent over the launch shape
    t0 = time.perf_counter()
    tuned = b.autotune(reps=args.tune_reps) if args.tune_reps > 0 else b.tuning()
    t_tune = time.perf_counter() - t0
    log(f"fixed shapes (ms per replay): {shape_ms}; autotune: {tuned} in {t_tune:.2f} s")
    for _ in range(args.warmup):
        b.run()
    counts = b.counts()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        b.run()
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - s0) * 1e3)
    t_batch = (time.perf_counter() - t0) / args.steps
    counts = b.counts()
    same = all(torch.equal(r, r2) and torch.equal(lab, l2) for _, _, r, lab, r2, l2, _ in dtiles) and counts == n_tile
    roof = batch_roofline(nh, dtiles, args, drain=bool(tuned and tuned[0]["drain"]))
    cpu, exact = (None, None) if args.no_cpu else cpu_baseline_batch(tiles, dtiles, counts, args)
    if exact is False:
        sys.exit(1)
    lat.sort()
    out = {
        "metric": "Mpix/s recon+bwlabel, uneven tiles batched in one graph vs per-tile launches",
        "value": round(total_px / t_batch / 1e6, 2), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(t_batch * 1e3, 3),
        "p50_ms": round(statistics.median(lat), 3), "p99_ms": round(lat[min(len(lat) - 1, int(0.99 * len(lat)))], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8+f32",
        "data": "synthetic", "bitexact": exact, "batched_equals_per_tile": bool(same),
        "per_tile_mpix_s": round(total_px / t_tile / 1e6, 2), "per_tile_ms": round(t_tile * 1e3, 3),
        "speedup_vs_per_tile": round(t_tile / t_batch, 3), "finalize_s": round(t_fin, 3),
        "autotune": {"groups": tuned, "s": round(t_tune, 3),
                     "fixed_shape_ms": {k: round(v, 3) for k, v in shape_ms.items()}},
        "config": {"workload": f"{args.tiles_u8} u8 + {args.tiles_f32} f32 tiles, sides 512-4096 step 16: "
                               "recon8 -> >128 (f32: >128/255) -> bwlabel8", "pixels": total_px,
                   "tiles": len(tiles), "step_alg_bytes": batch_alg_bytes(tiles),
                   "step_hbm_frac": round(batch_alg_bytes(tiles) / t_batch / (HBM_PEAK_GBS * 1e9), 5)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    b.close()
    print(json.dumps(out))


def batch_alg_bytes(tiles) -> int:
    """Algorithmic bytes of one batch step: per u8 tile 8 B/px (marker, mask in,
    recon out: 3; thresholded labelling: 1 in, 4 out), per f32 tile 16 B/px
    (12 + 4: the labelling reads the f32 reconstruction)."""
    return int(sum(m.size * (8 if m.dtype.itemsize == 1 else 16) for _, m, _ in tiles))


def batch_roofline(nh, dtiles, args, drain: bool = True) -> dict | None:
    """Roofline of the batch's dominant kernel kind.  Graph replays record no
    per-kernel events, so the same tiles run as an eager batch (finalized
    without a graph: the same launches, drained or per round as the tuned
    graph's first group) with HIP events on its private context; its
    statistics give the tile visits of each group's propagation rounds (a
    drain launch covers all of them).  Algorithmic bytes per launch: reconstruction rounds 3 B (u8) or
    12 B (f32) per visited tile pixel (62 x 62), the first pass that per pixel
    of every tile, labelling local 1 B (u8) or 4 B (f32) per pixel, output 5 B
    or 8 B per pixel."""
    import torch
    be = nh.Batch(nh.Context(0))
    u8px = f32px = 0
    for dm, dk, r, lab, _, _, thr in dtiles:
        be.add(dm, dk, r, lab, 8, thr, 8)
        if dk.dtype == torch.uint8:
            u8px += dk.numel()
        else:
            f32px += dk.numel()
    be.finalize(use_graph=False)
    be.set_drain(drain)
    bctx = be.context()
    steps = max(2, args.steps // 2)
    be.run()
    be.counts()
    bctx.set_timing(True)
    for _ in range(steps):
        be.run()
    torch.cuda.synchronize()
    kt = bctx.read_timing()
    bctx.set_timing(False)
    vis = [bctx.recon_stats(g)["tile_visits"] for g in (0, 1)]
    be.close()
    alg_step = {
        "recon_first": 3 * u8px + 12 * f32px,
        "recon_round": (3 * vis[0] + 12 * vis[1]) * TILE_PX,
        "recon_drain": (3 * vis[0] + 12 * vis[1]) * TILE_PX,
        "ccl_local": u8px + 4 * f32px,
        "ccl_final": 5 * u8px + 8 * f32px,
        "ccl_merge": 0,
        "ccl_small": 0,
        "recon_scan": 0,   # batches are multi-image: no image-wide scans
    }
    kind = max(kt, key=lambda k: kt[k][0])
    ms_step = kt[kind][0] / steps
    launches_step = kt[kind][1] / steps
    if not launches_step:
        return None
    alg_launch = alg_step[kind] / launches_step
    avg_launch_s = ms_step / launches_step / 1e3
    achieved = alg_launch / avg_launch_s / 1e9
    log("batch kernel time per step (ms, eager): " +
        ", ".join(f"{k}={v[0] / steps:.3f} ({v[1] / steps:.1f} launches)" for k, v in kt.items()))
    return {"bound": "hbm", "kernel": kind, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "alg_bytes_per_launch": int(alg_launch),
            "avg_launch_us": round(avg_launch_s * 1e6, 3), "launches_per_step": round(launches_step, 2),
            "per_kind_ms_per_step": {k: round(v[0] / steps, 4) for k, v in kt.items()},
            "timed_as": "eager batch (same launches as the graph) with HIP events"}


def host_threads() -> int:
    """Host threads for the multi-core CPU baselines: the CPUs this process may
    run on, at most 16 (one GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_batch(tiles, dtiles, counts, args):
    """The oracle's imreconstruct + bwlabel2 on every tile of the batch, one
    tile per host thread on all host CPUs (the oracle's C calls release the
    GIL); its outputs are also the exactness check of every batched tile.
    Returns (cpu_baseline, exact)."""
    import concurrent.futures as cf

    import numpy as np

    from oracle import oracle as orc
    idx = list(range(len(tiles)))[: max(0, args.check_tiles)] if args.check_tiles < len(tiles) else \
        list(range(len(tiles)))
    if not idx:
        return None, None

    def one(i):
        marker, mask, thr = tiles[i]
        wr = orc.imreconstruct(marker, mask, 8)
        bw = np.where(wr > (np.float32(thr) if wr.dtype == np.float32 else thr), 255, 0).astype(np.uint8)
        wl, wn = orc.bwlabel2(bw, 8, relab=True)
        return i, wr, wl, wn

    threads = host_threads()
    exact, px = True, 0
    t0 = time.perf_counter()
    results = []
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        for res in ex.map(one, idx):
            results.append(res)
    dt = time.perf_counter() - t0
    for i, wr, wl, wn in results:
        px += tiles[i][1].size
        _, _, r, lab, _, _, _ = dtiles[i]
        if not (np.array_equal(r.cpu().numpy(), wr) and np.array_equal(lab.cpu().numpy(), wl) and counts[i] == wn):
            exact = False
            log(f"tile {i} differs from the oracle")
    cpu = {"value": round(px / dt / 1e6, 3), "unit": "Mpix/s", "cores": threads, "kind": "port",
           "sample": f"{len(idx)} tiles ({px / 1e6:.1f} Mpx) through imreconstruct + bwlabel2, one tile per thread "
                     f"on {threads} threads, {dt:.1f} s wall; every GPU tile checked against it",
           "host_cpus": os.cpu_count(), "cpu_model": cpu_model()}
    return cpu, exact


def pct(xs, q):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, int(round(q * (len(xs) - 1))))]


def bench_serve(args):
    """Serving loop: a trace of requests (log-uniform sides 256..4096, 20% int32)
    through nsx_server with `slots` workers, each taking up to `max_batch`
    queued requests at once (those of one kind share launches), vs one worker
    serving one request at a time (sequential) and vs the same slots serving
    every request on its own (unshared).

    Reported value: Mpix/s of the server with every request queued at once
    (saturation).  Latency: Poisson arrivals at `load` x the sequential
    capacity (requests/s of one-at-a-time service), p50/p99 of submit -> done,
    for the server and for the sequential one-slot server on the same trace.
    Inputs and label outputs are in pinned host memory: each request's time
    includes its upload and the label download."""
    import threading
    import zlib

    import numpy as np
    import torch

    import nscale_hip as nh
    from nscale_hip.workloads import SERVE_SEED, request_images, serve_trace

    def pinned(a):
        t = torch.empty(a.shape, dtype={np.uint8: torch.uint8, np.int32: torch.int32}[a.dtype.type],
                        pin_memory=True)
        t.numpy()[...] = a
        return t.numpy()

    t0 = time.perf_counter()
    lo_side, hi_side = (int(v) for v in args.serve_sides.split(","))
    trace = serve_trace(args.requests, SERVE_SEED, lo_side, hi_side)
    reqs = []
    for side, kind, seed in trace:
        marker, mask, thr = request_images(side, kind, seed)
        reqs.append((pinned(marker), pinned(mask), thr,
                     torch.empty((side, side), dtype=torch.int32, pin_memory=True).numpy()))
    px = [m.size for _, m, _, _ in reqs]
    total_px = sum(px)
    log(f"generated {len(reqs)} requests ({total_px / 1e6:.1f} Mpx, "
        f"{sum(k == 'i32' for _, k, _ in trace)} int32) in {time.perf_counter() - t0:.1f} s")
    budget = int(args.budget_gib * (1 << 30))
    # one request at a time; the server as configured; the same slots serving
    # every request on its own (no shared launches)
    s1 = nh.Server(0, budget, slots=1, max_batch=1)
    sk = nh.Server(0, budget, slots=args.slots)
    sk.set_sharing(args.max_batch, args.share_max_px, args.share_set_px)
    sn = nh.Server(0, budget, slots=args.slots, max_batch=1)

    def submit(srv, i):
        marker, mask, thr, lab = reqs[i]
        return srv.submit(marker, mask, 8, 4, thr, 8, 11, 1000, labels_out=lab)

    # Every run writes the same label buffers; after each run (outside its
    # timed region) all outputs are hashed and compared with the first run's.
    digests = {}
    counts = [0] * len(reqs)

    def verify(tag):
        for i in range(len(reqs)):
            d = (zlib.crc32(reqs[i][3]), counts[i])
            if digests.setdefault(i, d) != d:
                raise SystemExit(f"request {i}: result differs in run '{tag}'")

    def burst(srv, tag):
        ts = time.perf_counter()
        tickets = [submit(srv, i) for i in range(len(reqs))]
        lat = []
        for i, t in enumerate(tickets):
            _, _, _, counts[i], tm = srv.wait(t)
            lat.append(tm[2])
        dt = time.perf_counter() - ts
        verify(tag)
        return dt, lat

    def closed_loop(srv, tag):
        ts = time.perf_counter()
        svc = []
        for i in range(len(reqs)):
            _, _, _, counts[i], tm = srv.wait(submit(srv, i))
            svc.append(tm[1])
        dt = time.perf_counter() - ts
        verify(tag)
        return dt, svc

    def poisson(srv, gaps, tag):
        tickets = [None] * len(reqs)
        lat = [0.0] * len(reqs)
        ready = threading.Semaphore(0)
        done_at = [0.0]

        def waiter():
            for i in range(len(reqs)):
                ready.acquire()
                _, _, _, counts[i], tm = srv.wait(tickets[i])
                lat[i] = tm[2]
            done_at[0] = time.perf_counter()

        th = threading.Thread(target=waiter)
        th.start()
        ts = time.perf_counter()
        due = ts
        for i in range(len(reqs)):
            due += gaps[i]
            while True:
                now = time.perf_counter()
                if now >= due:
                    break
                time.sleep(min(due - now, 0.0005))
            tickets[i] = submit(srv, i)
            ready.release()
        th.join()
        verify(tag)
        return done_at[0] - ts, lat

    # warm both servers (workspace sizes, round predictions per size bucket)
    for _ in range(max(1, args.warmup)):
        burst(s1, "warmup-1")
        burst(sk, "warmup-k")
        burst(sn, "warmup-n")
    t_seq, svc = closed_loop(s1, "sequential")
    t_sat1, _ = burst(s1, "saturation-1")
    # the two multi-slot servers alternate over a few bursts (the order flips
    # each time) and each keeps its median, so that drift over the run does
    # not favour either
    tn, tk = [], []
    for rep in range(args.serve_reps):
        order = ((sn, tn, "unshared"), (sk, tk, "shared"))
        for srv, acc, tag in (order if rep % 2 == 0 else order[::-1]):
            acc.append(burst(srv, f"saturation-{args.slots}-{tag}")[0])
    t_satn, t_satk = statistics.median(tn), statistics.median(tk)
    lam = args.load * len(reqs) / t_seq
    gaps = np.random.default_rng(SERVE_SEED ^ 0xA5A5).exponential(1.0 / lam, len(reqs))
    t_p1, lat_p1 = poisson(s1, gaps, "poisson-1")
    t_pk, lat_pk = poisson(sk, gaps, f"poisson-{args.slots}")
    roof = serve_roofline(sk, burst, reqs, args.slots)
    stk = sk.stats()
    s1.close()
    sk.close()
    sn.close()

    cpu, exact, checked = (None, None, 0) if args.no_cpu else cpu_baseline_serve(reqs, trace, digests,
                                                                                   args.check_max_side)
    if exact is False:
        sys.exit(1)
    mp = total_px / 1e6
    out = {
        "metric": "Mpix/s served segmentation chain (continuous admission, HBM budget), p50/p99 ms",
        "value": round(mp / t_satk, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": args.warmup,
        "ms_per_step": round(t_satk * 1e3, 3),
        "p50_ms": round(statistics.median(lat_pk), 3),
        "p99_ms": round(pct(lat_pk, 0.99), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8+i32",
        "data": "synthetic",
        "bitexact": exact,
        "results_same_alone_and_in_mix": True,
        "oracle_checked_requests": checked,
        "saturation": {"server_mpix_s": round(mp / t_satk, 2), "sequential_mpix_s": round(mp / t_sat1, 2),
                       "unshared_mpix_s": round(mp / t_satn, 2),
                       "closed_loop_sequential_mpix_s": round(mp / t_seq, 2),
                       "speedup": round(t_sat1 / t_satk, 3), "speedup_vs_unshared": round(t_satn / t_satk, 3),
                       "bursts_ms": {"shared": [round(t * 1e3, 2) for t in tk],
                                     "unshared": [round(t * 1e3, 2) for t in tn]}},
        "poisson": {"lambda_req_s": round(lam, 2), "load_vs_sequential_capacity": args.load,
                    "server": {"p50_ms": round(statistics.median(lat_pk), 3), "p99_ms": round(pct(lat_pk, 0.99), 3),
                               "mpix_s": round(mp / t_pk, 2)},
                    "sequential": {"p50_ms": round(statistics.median(lat_p1), 3),
                                   "p99_ms": round(pct(lat_p1, 0.99), 3), "mpix_s": round(mp / t_p1, 2)}},
        "service_ms_alone": {"p50": round(statistics.median(svc), 3), "p99": round(pct(svc, 0.99), 3)},
        "server": stk,
        "config": {"workload": f"{len(reqs)} requests, square sides log-uniform {lo_side}-{hi_side}, 80% u8 (>128) / 20% "
                               "int32 (x257, >32768): recon8 -> fillholes4 -> threshold -> bwlabel8 -> "
                               "bwareaopen[11,1000) -> relabel; pinned host in/out",
                   "pixels": total_px, "slots": args.slots, "max_batch": args.max_batch,
                   "share_max_request_px": stk["share_max_request_px"], "share_max_set_px": stk["share_max_set_px"],
                   "hbm_budget_bytes": budget, "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES")},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def serve_roofline(srv, burst, reqs, slots) -> dict | None:
    """Roofline of the serving loop's dominant kernel kind.  One more
    saturation burst runs with HIP-event timing on every slot's context (events
    add gaps between launches, so the timed bursts run without them); the
    slots' kernel times and tile counts are summed.  Kernels of different slots
    overlap on the device, so a launch's duration here is its duration while
    sharing the GPU.  Algorithmic bytes: reconstruction 3 B per pixel per
    element byte (marker, mask in; state out) and hole filling 2 B (image in,
    state out), over the tiles (62 x 62) each first pass and round processed;
    labelling local: the fill read once (1 or 4 B/px); label output: that read
    plus 4 B written."""
    ctxs = [srv.context(i) for i in range(slots)]
    for c in ctxs:
        c.set_timing(True)
    burst(srv, "kernel-timing")
    kt, tiles = {}, {}
    for c in ctxs:
        for k, (ms, n) in c.read_timing().items():
            a = kt.get(k, (0.0, 0))
            kt[k] = (a[0] + ms, a[1] + n)
        for k, (f, r) in c.read_tile_counts().items():
            a = tiles.get(k, (0, 0))
            tiles[k] = (a[0] + f, a[1] + r)
        c.set_timing(False)
    per_px = {"recon": 3, "fill": 2, "binary": 3, "fillbin": 2}
    es = {"u8": 1, "i32": 4, "f32": 4}
    first_b = sum(f * TILE_PX * per_px[m] * es[d] for (m, d), (f, _) in tiles.items())
    round_b = sum(r * TILE_PX * per_px[m] * es[d] for (m, d), (_, r) in tiles.items())
    in_b = sum(mask.size * mask.itemsize for _, mask, _, _ in reqs)
    alg = {"recon_first": first_b, "recon_round": round_b, "ccl_local": in_b, "ccl_final": in_b + 4 * sum(
        mask.size for _, mask, _, _ in reqs), "ccl_merge": 0, "ccl_small": 0, "recon_scan": 0,
           "recon_drain": round_b}
    kind = max(kt, key=lambda k: kt[k][0])
    ms, n = kt[kind]
    if not n or not alg[kind]:
        return None
    alg_launch = alg[kind] / n
    avg_s = ms / n / 1e3
    achieved = alg_launch / avg_s / 1e9
    log("serve kernel time per burst (ms, summed over slots): " +
        ", ".join(f"{k}={v[0]:.3f} ({v[1]} launches)" for k, v in kt.items()))
    log(f"serve tile counts (first pass, rounds): {tiles}")
    return {"bound": "hbm", "kernel": kind, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "alg_bytes_per_launch": int(alg_launch),
            "avg_launch_us": round(avg_s * 1e6, 3), "launches_per_burst": n,
            "per_kind_ms_per_burst": {k: round(v[0], 4) for k, v in kt.items()},
            "timed_as": f"one saturation burst with HIP events on all {slots} slot streams (kernels of different "
                        "slots overlap)"}


def cpu_baseline_serve(reqs, trace, digests, max_side):
    """The oracle's chain on the trace's requests of side <= max_side (a
    bounded sample), one request per host thread on all host CPUs (the
    oracle's C calls release the GIL); its outputs double as the exactness
    check of the served labels.  Returns (cpu_baseline, exact, requests
    checked)."""
    import concurrent.futures as cf
    import zlib

    import numpy as np

    from oracle import oracle as orc
    idx = [i for i, (side, _, _) in enumerate(trace) if side <= max_side]

    def one(i):
        marker, mask, thr, _ = reqs[i]
        if trace[i][1] == "u8":
            _, _, want, wn = orc.chain_u8(marker, mask, 8, 4, int(thr), 8, 11, 1000)
        else:
            r = orc.imreconstruct(marker, mask, 8)
            f = orc.fill_holes(r, False, 4)
            b = np.where(f > thr, 255, 0).astype(np.uint8)
            want, wn = orc.relabel(orc.bwareaopen2(b, False, False, 11, 1000, 8)[0], -1)
        return i, want, wn

    threads = host_threads()
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        results = list(ex.map(one, idx))
    dt = time.perf_counter() - t0
    exact, px = True, 0
    for i, want, wn in results:
        lab = reqs[i][3]
        px += lab.size
        if (zlib.crc32(np.ascontiguousarray(want)), wn) != digests[i] or not np.array_equal(lab, want):
            exact = False
            log(f"request {i} ({trace[i][0]}x{trace[i][0]} {trace[i][1]}) differs from the oracle")
    cpu = {"value": round(px / dt / 1e6, 3) if dt > 0 else None, "unit": "Mpix/s", "cores": threads,
           "kind": "port",
           "sample": f"{len(idx)} requests of side <= {max_side} ({px / 1e6:.1f} Mpx) through the chain, one "
                     f"request per thread on {threads} threads, {dt:.1f} s wall; every one checked against the "
                     "served labels",
           "host_cpus": os.cpu_count(), "cpu_model": cpu_model()}
    return cpu, exact, len(idx)


def bench_maze(args):
    """16384x16384 u8 maze (6-px cells, 3-px corridors at 200, walls 20, seed
    0x5EED0003), marker = 5x5 block of 200 at the entrance.  One step =
    imreconstruct (4-conn) -> (> 100) -> bwlabel (4-conn, canonical) through
    nsx_recon_label on device-resident buffers.  Checked against the closed
    form (200 on the corridors, 20 elsewhere; exactly one label, equal to the
    corridor set)."""
    import numpy as np
    import torch

    import nscale_hip as nh
    n = args.maze_size
    t0 = time.perf_counter()
    mask, marker = nh.gen_maze_u8(n, n, 0x5EED0003)
    log(f"generated {n}x{n} maze in {time.perf_counter() - t0:.1f} s")
    dev = torch.device("cuda:0")
    dm, dk = torch.from_numpy(marker).to(dev), torch.from_numpy(mask).to(dev)
    dr = torch.empty_like(dk)
    dl = torch.empty((n, n), dtype=torch.int32, device=dev)
    ctx = nh.default_context(0)

    def step():
        return nh.recon_label(dm, dk, 4, 100.0, 4, recon_out=dr, labels_out=dl, ctx=ctx)[2]

    ncomp = step()
    torch.cuda.synchronize()
    corridors = torch.from_numpy(mask == 200).to(dev)
    exact = bool(ncomp == 1 and torch.equal(dr, torch.where(corridors, 200, 20).to(torch.uint8))
                 and torch.equal(dl, corridors.to(torch.int32)))
    del corridors
    log(f"closed form: {exact} (components {ncomp})")
    if not exact:
        sys.exit(1)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        step()
        lat.append((time.perf_counter() - s0) * 1e3)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ctx.set_timing(True)
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    kt = ctx.read_timing()
    ctx.set_timing(False)
    px = n * n
    # Algorithmic bytes per pixel of each launch kind over one step: one
    # labelling of {mask >= 200} reads mask + marker (local); its output pass
    # writes the reconstruction and the int32 labels of {recon > 100}, which
    # are that labelling's components (final).  Step: 2 B in, 5 B out.
    alg = {"ccl_local": 2 * px, "ccl_final": 5 * px}
    kind = max(kt, key=lambda k: kt[k][0])
    roof = None
    if kind in alg and kt[kind][1]:
        ms = kt[kind][0] / kt[kind][1]
        achieved = alg[kind] / (kt[kind][1] / args.steps) / (ms / 1e3) / 1e9
        roof = {"bound": "hbm", "kernel": kind, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "alg_bytes_per_launch": int(alg[kind] / (kt[kind][1] / args.steps)), "avg_launch_us": round(ms * 1e3, 3),
                "launches_per_step": round(kt[kind][1] / args.steps, 2),
                "per_kind_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in kt.items()}}
    cpu = None if args.no_cpu else cpu_baseline_maze(nh, args.maze_cpu_size, ctx)
    value = args.steps * px / elapsed / 1e6
    out = {
        "metric": "Mpix/s maze reconstruct+CCL @16384^2 u8 bit-exact",
        "value": round(value, 2), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4), "p50_ms": round(statistics.median(lat), 4),
        "p99_ms": round(pct(lat, 0.99), 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8", "data": "synthetic", "bitexact": exact,
        "config": {"workload": f"{n}x{n} u8 maze (6-px cells, 3-px corridors 200 / walls 20), 5x5 seed: recon4 "
                               "-> >100 -> bwlabel4", "pixels": px, "components": ncomp,
                   "step_alg_bytes": 7 * px,
                   "step_hbm_frac": round(7 * px / (elapsed / args.steps) / (HBM_PEAK_GBS * 1e9), 5)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def bench_maze_levels(args):
    """The many-valued maze (nscale_hip.workloads.maze_levels_images: the 16384^2
    maze with corridor pixels uniform in 150..220, noise -8..8 on every pixel,
    a 5x5 seed of 255 at the entrance).  The level-set path declines it (over
    100 distinct values) and propagation needs one round per tile crossing of
    paths millions of pixels long.  One step = imreconstruct (4-conn) -> (> 100)
    -> bwlabel (4-conn) through nsx_recon_label with the default engine
    (NSX_ENGINE_AUTO: rounds, switching to the bit planes past 256 rounds
    and starting there on later calls of the size).  Checked: every corridor
    pixel is reached above 100 and every wall stays below it, so the labels are
    exactly one component, the corridor set; the reconstruction equals the one
    propagation rounds alone compute (timed once, reported as rounds_ms) and,
    on a sample of the same construction, the oracle's."""
    import numpy as np
    import torch

    import nscale_hip as nh
    from nscale_hip.workloads import maze_levels_images
    n = args.maze_size
    t0 = time.perf_counter()
    marker, mask = maze_levels_images(n)
    log(f"generated {n}x{n} many-valued maze in {time.perf_counter() - t0:.1f} s "
        f"({len(np.unique(mask))} distinct values)")
    dev = torch.device("cuda:0")
    dm, dk = torch.from_numpy(marker).to(dev), torch.from_numpy(mask).to(dev)
    corridors = torch.from_numpy(mask > 100).to(dev)
    del marker, mask
    dr = torch.empty_like(dk)
    dl = torch.empty((n, n), dtype=torch.int32, device=dev)
    ctx = nh.Context(0)

    def step(c=ctx, r=dr, lab=dl):
        return nh.recon_label(dm, dk, 4, 100.0, 4, recon_out=r, labels_out=lab, ctx=c)[2]

    t0 = time.perf_counter()
    ncomp = step()
    torch.cuda.synchronize()
    first_ms = (time.perf_counter() - t0) * 1e3
    exact = bool(ncomp == 1 and torch.equal(dl, corridors.to(torch.int32)))
    # propagation rounds alone, once: its time and its (identical) result
    rctx = nh.Context(0)
    rctx.set_tuning(nh._lib.NSX_TUNE_ENGINE, nh._lib.NSX_ENGINE_ROUNDS)
    rr = torch.empty_like(dk)
    rl = torch.empty_like(dl)
    t0 = time.perf_counter()
    step(rctx, rr, rl)
    torch.cuda.synchronize()
    rounds_ms = (time.perf_counter() - t0) * 1e3
    rounds = rctx.recon_stats(0)["rounds"]
    exact = exact and bool(torch.equal(rr, dr) and torch.equal(rl, dl))
    del rr, rl, corridors
    rctx.close()
    log(f"closed form and rounds agree: {exact}; first call {first_ms:.1f} ms; rounds alone {rounds_ms:.1f} ms "
        f"({rounds} rounds)")
    if not exact:
        sys.exit(1)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - s0) * 1e3)
    elapsed = time.perf_counter() - t0
    px = n * n
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        cn = min(n, args.maze_levels_cpu_size)
        cm, ck = maze_levels_images(cn)
        t0 = time.perf_counter()
        wr = orc.imreconstruct(cm, ck, 4)
        wl, wn = orc.bwlabel2(np.where(wr > 100, 255, 0).astype(np.uint8), 4, relab=True)
        dt = time.perf_counter() - t0
        gr, gl, gn = nh.recon_label(torch.from_numpy(cm).cuda(), torch.from_numpy(ck).cuda(), 4, 100.0, 4, ctx=ctx)
        same = bool(np.array_equal(gr.cpu().numpy(), wr) and np.array_equal(gl.cpu().numpy(), wl) and gn == wn)
        if not same:
            log("many-valued maze sample differs from the oracle")
            sys.exit(1)
        cpu = {"value": round(cn * cn / dt / 1e6, 3), "unit": "Mpix/s", "cores": 1, "kind": "port",
               "sample": f"one {cn}x{cn} many-valued maze of the same construction (recon4 + bwlabel4), single "
                         f"thread, {dt:.1f} s; GPU output on it identical", "host_cpus": os.cpu_count(), "cpu_model": cpu_model()}
    # per-kind kernel times of a few more steps with HIP events: under the
    # adaptive engine each step is the bit-plane reconstruction (8 passes of
    # prep + labelling with a fold) and one labelling of {recon > 100}
    ctx.set_timing(True)
    ksteps = max(1, args.steps)
    for _ in range(ksteps):
        step()
    torch.cuda.synchronize()
    kt = ctx.read_timing()
    ctx.set_timing(False)
    log("maze_levels kernel time per step (ms): " +
        ", ".join(f"{k}={v[0] / ksteps:.3f} ({v[1] / ksteps:.1f} launches)" for k, v in kt.items() if v[1]))
    roof = None
    if kt.get("ccl_local", (0, 0))[1]:
        # labelling local pass: 2 B/px per bit-plane pass (class labels and
        # seed weights read), 1 B/px for the final labelling (recon read)
        ms_l, n_l = kt["ccl_local"]
        passes = n_l / ksteps - 1
        alg_launch = (passes * 2 * px + px) / (passes + 1)
        avg_s = ms_l / n_l / 1e3
        achieved = alg_launch / avg_s / 1e9
        kind = max(kt, key=lambda k: kt[k][0])
        roof = {"bound": "hbm", "kernel": "ccl_local", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "alg_bytes_per_launch": int(alg_launch), "avg_launch_us": round(avg_s * 1e6, 3),
                "launches_per_step": round(n_l / ksteps, 2), "dominant_kind": kind,
                "per_kind_ms_per_step": {k: round(v[0] / ksteps, 4) for k, v in kt.items() if v[1]}}
    ctx.close()
    out = {
        "metric": f"Mpix/s many-valued maze reconstruct+CCL @{n}^2 u8 bit-exact",
        "value": round(args.steps * px / elapsed / 1e6, 2), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "p50_ms": round(statistics.median(lat), 3), "p99_ms": round(pct(lat, 0.99), 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic", "bitexact": exact,
        "first_call_ms": round(first_ms, 1), "rounds_ms": round(rounds_ms, 1), "rounds": rounds,
        "config": {"workload": f"{n}x{n} u8 maze, corridors 150..220 + noise, 5x5 seed: recon4 -> >100 -> bwlabel4 "
                               "(engine auto)", "pixels": px},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def bench_copy(args):
    """Calibration workload for the HBM traffic counters: one elementwise
    kernel (int32 negation, torch) that reads and writes exactly 256 MiB per
    launch.  rocprofv3's FETCH_SIZE / WRITE_SIZE for this kernel, divided into
    268435456 bytes, give the factor that converts those counters into bytes."""
    import torch
    n = 64 << 20
    x = torch.ones(n, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    gbs = 2 * 4 * n / dt / 1e9
    print(json.dumps({"metric": "GB/s device copy (int32 negation, 256 MiB read + 256 MiB written per launch)",
                      "value": round(gbs, 1), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
                      "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 4), "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
                      "config": {"workload": "copy calibration", "bytes_read": 4 * n, "bytes_written": 4 * n},
                      "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None},
                      "cpu_baseline": None}))


def bench_ops(args):
    """The operators around the chain, each timed alone on the 4096^2 disk
    image (inputs resident in HBM, outputs to HBM; each call ends with the
    stream synchronised, as the entry points do), each checked against the
    oracle, whose single-thread time is the CPU baseline."""
    import numpy as np
    import torch

    import nscale_hip as nh
    n = args.size
    mask = nh.gen_disks_u8(n, n, args.disks, 0x5EED0001)
    binm = np.where(mask > 128, 255, 0).astype(np.uint8)
    rng = np.random.default_rng(0x5EED0003)
    seeds = np.where(rng.random((n, n)) < 1 / 2000, 255, 0).astype(np.uint8)
    dev = torch.device("cuda:0")
    d_mask, d_bin, d_seeds = (torch.from_numpy(a).to(dev) for a in (mask, binm, seeds))
    d_labels = nh.bwlabel2(d_bin, 8, True)
    labels = d_labels.cpu().numpy()
    orc = None
    if not args.no_cpu:
        from oracle import oracle as orc
    npx = n * n
    # name, GPU call, oracle call, HBM bytes the operator must move at least
    ops = [
        ("imfill gray 4-conn", lambda: nh.imfill(d_mask, d_seeds, False, 4),
         lambda: orc.imfill(mask, seeds, False, 4), 3 * npx),
        ("bwselect 8-conn", lambda: nh.bwselect(d_bin, d_seeds, 8), lambda: orc.bwselect(binm, seeds, 8), 3 * npx),
        ("imhmin h=10 8-conn", lambda: nh.imhmin(d_mask, 10, 8), lambda: orc.imhmin(mask, 10, 8), 2 * npx),
        ("localMaxima 8-conn", lambda: nh.localMaxima(d_mask, 8), lambda: orc.local_extrema(mask, 8), 2 * npx),
        ("localMinima 8-conn", lambda: nh.localMinima(d_mask, 8), lambda: orc.local_extrema(mask, 8, True),
         2 * npx),
        ("boundingBox", lambda: nh.bounding_box(d_labels), lambda: orc.bounding_box(labels), 4 * npx),
        ("distanceTransform", lambda: nh.distanceTransform(d_bin), lambda: orc.distance_transform(binm)[1],
         5 * npx),
    ]
    reps = max(1, min(args.steps, 10))
    res = {}
    total_ms = 0.0
    for name, gpu, cpu, alg in ops:
        for _ in range(max(1, args.warmup)):
            got = gpu()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            got = gpu()
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        ms = statistics.median(ts)
        total_ms += ms
        r = {"ms": round(ms, 4), "mpix_s": round(npx / ms / 1e3, 1),
             "hbm_frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        tile_got = None
        if name == "distanceTransform":
            # the same rule in tile order (a fixed point of the rule, not the
            # reference's queue order), for comparison
            tctx = nh.Context(0)
            tctx.set_option(nh.OPT_DT_ORDER, nh.DT_ORDER_TILES)
            tile_got = nh.distanceTransform(d_bin, ctx=tctx)
            torch.cuda.synchronize()
            tts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                tile_got = nh.distanceTransform(d_bin, ctx=tctx)
                torch.cuda.synchronize()
                tts.append((time.perf_counter() - t0) * 1e3)
            tctx.close()
            r["tile_order_ms"] = round(statistics.median(tts), 4)
        if orc is not None:
            t0 = time.perf_counter()
            want = cpu()
            cpu_ms = (time.perf_counter() - t0) * 1e3
            if name == "boundingBox":
                same = got[0] == want[0] and np.array_equal(got[1], want[1])
            else:
                g = got.cpu().numpy()
                same = bool(np.array_equal(g, want))
                if name == "distanceTransform":
                    # the propagation can stop short of the exact distance at
                    # a few pixels, at places that depend on the order of the
                    # offers: count them on both sides against scipy's EDT,
                    # for the queue order and the tile order
                    from scipy import ndimage
                    yy, xx = np.mgrid[0:n, 0:n]
                    d2 = np.rint(ndimage.distance_transform_edt(binm != 0) ** 2).astype(np.int64)
                    d2 = np.minimum(d2, xx.astype(np.int64) ** 2 + (3 * n - yy.astype(np.int64)) ** 2)
                    exact = np.sqrt(d2.astype(np.float64)).astype(np.float32)
                    tg = tile_got.cpu().numpy()
                    r.update({"pixels_differing": int((g != want).sum()),
                              "oracle_not_exact": int((want != exact).sum()),
                              "gpu_not_exact": int((g != exact).sum()),
                              "gpu_below_exact": int((g < exact).sum()),
                              "tile_order_pixels_differing": int((tg != want).sum()),
                              "tile_order_not_exact": int((tg != exact).sum())})
            r.update({"cpu_ms": round(cpu_ms, 1), "speedup": round(cpu_ms / ms, 1), "bitexact": bool(same)})
        res[name] = r
        log(name, r)
    value = len(ops) * npx / (total_ms * 1e-3) / 1e6
    print(json.dumps({
        "metric": "Mpix/s over the operators around the chain @4096^2 (each run once)", "value": round(value, 1),
        "unit": "Mpix/s", "n_gpus": 1, "steps": reps, "warmup": args.warmup, "ms_per_step": round(total_ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": f"{n}x{n} u8 disks({args.disks}); binary = >128; seeds 1 in 2000 pixels; "
                               "labels = bwlabel8", "ops": res},
        "roofline": None,
        "cpu_baseline": None if orc is None else {
            "value": round(len(ops) * npx / (sum(r["cpu_ms"] for r in res.values()) * 1e-3) / 1e6, 3),
            "unit": "Mpix/s", "cores": 1, "kind": "port", "sample": "each operator once on the same image",
            "host_cpus": os.cpu_count(), "cpu_model": cpu_model()},
    }))


def cpu_baseline_maze(nh, n, ctx):
    """The oracle's imreconstruct + bwlabel2 (single thread) on an n x n maze of
    the same construction (by default the whole 16384^2 workload, a few seconds
    of CPU work); the GPU result on the same input is compared with it.
    Returns cpu_baseline."""
    import numpy as np
    import torch

    from oracle import oracle as orc
    mask, marker = nh.gen_maze_u8(n, n, 0x5EED0003)
    t0 = time.perf_counter()
    wr = orc.imreconstruct(marker, mask, 4)
    wl, wn = orc.bwlabel2(np.where(wr > 100, 255, 0).astype(np.uint8), 4, relab=True)
    dt = time.perf_counter() - t0
    r, lab, gn = nh.recon_label(torch.from_numpy(marker).cuda(), torch.from_numpy(mask).cuda(), 4, 100.0, 4, ctx=ctx)
    same = bool(np.array_equal(r.cpu().numpy(), wr) and np.array_equal(lab.cpu().numpy(), wl) and gn == wn)
    if not same:
        log("maze sample differs from the oracle")
        sys.exit(1)
    return {"value": round(n * n / dt / 1e6, 3), "unit": "Mpix/s", "cores": 1, "kind": "port",
            "sample": f"one {n}x{n} maze of the same construction (recon4 + bwlabel4), single thread, {dt:.1f} s; "
                      "GPU output on it identical", "host_cpus": os.cpu_count(), "cpu_model": cpu_model()}


TILE_PX = 62 * 62   # interior pixels of one reconstruction tile visit

# HBM bytes per launch of each kernel kind of the default 4096x4096 chain,
# measured by two rocprofv3 runs of this bench (--steps 3 --warmup 1 --no-cpu,
# one with --pmc FETCH_SIZE, one with --pmc WRITE_SIZE), averaged over every
# launch of the kind.  FETCH_SIZE is scaled by 2.0 and WRITE_SIZE by 1.0: the
# factors that make `--workload copy` (256 MiB read + 256 MiB written per
# launch) read true.  Commands and raw figures: profiles/traffic_chain4096.md.
TRAFFIC_CHAIN4096 = {"recon_round": 13146098, "recon_first": 27326407, "ccl_local": 50972672,
                     "ccl_final": 101438818, "ccl_merge": 36901740}


def border_tiles(rows: int, cols: int) -> int:
    """Reconstruction tiles (62 x 62) that touch the image border."""
    ntx, nty = -(-cols // 62), -(-rows // 62)
    return ntx * nty if ntx <= 2 or nty <= 2 else 2 * ntx + 2 * (nty - 2)


def roofline(kt: dict, px: int, tiles: dict, steps: int, traffic: dict | None = None, scan_pairs: int = 0) -> dict:
    """HBM roofline of the kernel kind with the most time per step.

    Algorithmic bytes per pixel (compulsory I/O of the operation a launch
    performs): reconstruction 3 B (marker, mask in; state out), fillHoles 2 B
    (image in; state out), labelling read 1 B, label output 1 B in + 4 B out.
    First passes and propagation rounds process the tiles (62 x 62 pixels)
    the context counted over the timed steps (`tiles`, from
    Context.read_tile_counts).  A pair of image-wide scans moves 6 B/px (row
    scan: image and state in, state out; column scan the same), 5 B/px for the
    first pair, whose row scan starts from the initial state.
    """
    per_px = {"recon": 3, "fill": 2, "binary": 3, "fillbin": 2}
    es = {"u8": 1, "i32": 4, "f32": 4}
    first_b = sum(f * TILE_PX * per_px[m] * es[d] for (m, d), (f, _) in tiles.items())
    round_b = sum(r * TILE_PX * per_px[m] * es[d] for (m, d), (_, r) in tiles.items())
    alg_total = {
        "recon_first": first_b,
        "recon_round": round_b,
        "recon_drain": round_b,
        "ccl_local": 1 * px * steps,
        "ccl_final": 5 * px * steps,
        "ccl_merge": 0,
        "ccl_small": 0,
        "recon_scan": (6 * scan_pairs - 1) * px * steps if scan_pairs else 0,
    }
    kind = max(kt, key=lambda k: kt[k][0])
    ms_total, launches = kt[kind]
    alg_launch = alg_total[kind] / launches if launches else 0.0
    avg_launch_s = ms_total / launches / 1e3 if launches else float("nan")
    achieved = alg_launch / avg_launch_s / 1e9 if launches else 0.0
    return {"bound": "hbm", "kernel": kind, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": (traffic or {}).get(kind),
            "alg_bytes_per_launch": int(alg_launch), "avg_launch_us": round(avg_launch_s * 1e6, 3),
            "launches_per_step": round(launches / steps, 2),
            "per_kind_ms_per_step": {k: round(v[0] / steps, 4) for k, v in kt.items()}}


def changed_input_ms(nh, ctx, n, args, d_marker, d_mask, d_recon, d_fill, d_lab, stream) -> dict:
    """Steps whose input changes every time: the chain alternates between the
    workload and a second image of the same size and construction (another
    generator seed), so the round prediction carried from the previous call
    is for other content (a short prediction costs a host round trip and the
    stages after it again).  Reports ms per step over the alternation.  Runs
    after the timed steps."""
    import torch
    mask2 = nh.gen_disks_u8(n, n, args.disks, 0x5EED0000 + 77)
    marker2 = nh.gen_marker_u8(mask2, 30)
    pairs = [(d_marker, d_mask), (torch.from_numpy(marker2).to(d_mask.device), torch.from_numpy(mask2).to(d_mask.device))]
    steps = max(4, args.steps // 2) // 2 * 2
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        dm, dk = pairs[i % 2]
        nh.segment_chain(dm, dk, 8, 4, 128.0, 8, 11, 1000, recon_out=d_recon, fill_out=d_fill, labels_out=d_lab,
                         ctx=ctx, stream=stream)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    return {"ms_per_step": round(ms, 4), "steps": steps, "inputs": 2,
            "note": f"alternating two {n}^2 disk images (generator seeds 0x5EED0001, 0x5EED004D)"}


def cpu_baseline(marker, mask, runs: int, gpu_first) -> tuple:
    """The oracle's single-threaded restatement of the reference chain, timed
    on this host on `runs` repetitions of the same 4096x4096 workload.  The
    first repetition's outputs double as the exactness check of the GPU's first
    step.  Returns (cpu_baseline, exact)."""
    import numpy as np

    from oracle import oracle as orc
    want = orc.chain_u8(marker, mask, 8, 4, 128, 8, 11, 1000)
    exact = bool(all(np.array_equal(a, b) for a, b in zip(gpu_first[:3], want[:3])) and gpu_first[3] == want[3])
    log(f"bit-exact vs oracle: {exact} (components {gpu_first[3]} / {want[3]})")
    t0 = time.perf_counter()
    for _ in range(runs):
        orc.chain_u8(marker, mask, 8, 4, 128, 8, 11, 1000)
    dt = time.perf_counter() - t0
    return {"value": round(runs * mask.size / dt / 1e6, 3), "unit": "Mpix/s", "cores": 1, "kind": "port",
            "sample": f"{runs} runs of the full {mask.shape[0]}x{mask.shape[1]} chain, single thread, "
                      f"{dt:.1f} s", "host_cpus": os.cpu_count(), "cpu_model": cpu_model()}, exact


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this hot path runs on one GPU; multi-GPU sharding is a design note only "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "serve":
        # one hardware queue per stream (the slots of three servers, torch's
        # and the default context's): with HIP's default of 4 queues per
        # process, streams created later share queues and serialise, which
        # would favour whichever server got its own queues
        if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 16:
            os.environ["GPU_MAX_HW_QUEUES"] = "16"
    import numpy as np
    import torch

    import nscale_hip as nh

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "maze":
        return bench_maze(args)
    if args.workload == "maze_levels":
        return bench_maze_levels(args)
    if args.workload == "copy":
        return bench_copy(args)
    if args.workload == "ops":
        return bench_ops(args)
    n = args.size
    seed = 0x5EED0001
    mask = nh.gen_disks_u8(n, n, args.disks, seed)
    marker = nh.gen_marker_u8(mask, 30)
    dev = torch.device("cuda:0")
    d_mask = torch.from_numpy(mask).to(dev)
    d_marker = torch.from_numpy(marker).to(dev)
    d_recon = torch.empty_like(d_mask)
    d_fill = torch.empty_like(d_mask)
    d_lab = torch.empty((n, n), dtype=torch.int32, device=dev)
    ctx = nh.default_context(0)
    ctx.set_tuning(nh._lib.NSX_TUNE_DRAIN, int(args.drain))
    stream = torch.cuda.current_stream(dev)

    def step():
        return nh.segment_chain(d_marker, d_mask, 8, 4, 128.0, 8, 11, 1000, recon_out=d_recon, fill_out=d_fill,
                                labels_out=d_lab, ctx=ctx, stream=stream)[3]

    # a cold call: a fresh context has no round prediction for this size, so
    # both propagations run past the default and are completed by the host
    torch.cuda.synchronize()
    c0 = time.perf_counter()
    ncomp = step()
    torch.cuda.synchronize()
    cold_ms = (time.perf_counter() - c0) * 1e3
    first = (d_recon.cpu().numpy(), d_fill.cpu().numpy(), d_lab.cpu().numpy(), ncomp)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        step()
        lat.append((time.perf_counter() - s0) * 1e3)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    stats = ctx.recon_stats(0)
    fstats = ctx.recon_stats(1)
    # work-list occupancy of every propagation round of the last timed step
    # (tiles of 62 x 62 pixels queued per round; entry 0 is the first pass)
    work_lists = {"recon": ctx.round_log(0), "fill": ctx.round_log(1)}
    changed = changed_input_ms(nh, ctx, n, args, d_marker, d_mask, d_recon, d_fill, d_lab, stream)
    # Per-kernel durations come from a second set of steps with HIP events
    # around every launch (events add a gap between launches, so the timed
    # steps above run without them).
    ctx.set_timing(True)
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    kt = ctx.read_timing()
    tiles = ctx.read_tile_counts()
    ctx.set_timing(False)
    log("kernel time per step (ms): " + ", ".join(f"{k}={v[0] / args.steps:.4f} ({v[1] / args.steps:.1f} launches)"
                                                  for k, v in kt.items()))
    log(f"recon stage {stats}; fill stage {fstats}; tiles over the timed steps {tiles}")
    default_chain = n == 4096 and args.disks == 3000
    scan_pairs = ctx.get_tuning(nh._lib.NSX_TUNE_FILL_SCANS) if kt.get("recon_scan", (0, 0))[1] else 0
    roof = roofline(kt, n * n, tiles, args.steps, TRAFFIC_CHAIN4096 if default_chain else None, scan_pairs)
    cpu, exact = (None, None) if args.no_cpu else cpu_baseline(marker, mask, args.cpu_runs, first)
    if exact is False:
        sys.exit(1)
    px = n * n
    value = args.steps * px / elapsed / 1e6
    lat_sorted = sorted(lat)
    p99 = lat_sorted[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))]
    out = {
        "metric": "Mpix/s reconstruct+CCL @4096^2 u8 bit-exact",
        "value": round(value, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "p50_ms": round(statistics.median(lat), 4),
        "p99_ms": round(p99, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "bitexact": exact,
        "cold_call_ms": round(cold_ms, 3),
        "changed_input": changed,
        "work_lists": work_lists,
        "config": {"workload": f"{n}x{n} u8 disks({args.disks})+noise: recon8 -> fillholes4 -> >128 -> "
                               "bwlabel8 -> bwareaopen[11,1000) -> relabel",
                   "global_batch": 1, "pixels": px, "components": ncomp,
                   "recon_rounds": stats["rounds"], "recon_tile_visits": stats["tile_visits"],
                   "fill_rounds": fstats["rounds"], "fill_tile_visits": fstats["tile_visits"],
                   "tiles": stats["tiles"], "step_alg_bytes": 10 * px,
                   "step_hbm_frac": round(10 * px / (elapsed / args.steps) / (HBM_PEAK_GBS * 1e9), 5)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
