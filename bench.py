#!/usr/bin/env python3
"""Benchmark of the reconstruct + CCL hot path on one MI355X.

Workload (default): one 4096x4096 u8 tile from the disk generator (3000 disks,
seed 0x5EED0001), marker = mask - 30, inputs resident in HBM.  One step =
    imreconstruct (8-conn) -> imfillHoles (grayscale, 4-conn) -> (> 128)
    -> bwlabel (8-conn) -> bwareaopen [11, 1000) -> canonical relabel
through the C ABI (nsx_segment_chain).  Reported value: Mpixel/s over the
timed steps.  The first step's outputs are checked bit-exactly against the
oracle before timing.

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
    p.add_argument("--no-check", action="store_true", help="skip the oracle comparison")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--cpu-runs", type=int, default=10, help="oracle chain runs in the CPU sample")
    p.add_argument("--workload", default="chain4096",
                   choices=["chain4096", "batch", "maze", "serve"],
                   help="chain4096: the headline 4096^2 chain (default); batch: 256 u8 + 32 f32 uneven tiles "
                        "batched vs per-tile; maze: 16384^2 maze stress; serve: continuous-batching server")
    p.add_argument("--tiles-u8", type=int, default=256)
    p.add_argument("--tiles-f32", type=int, default=32)
    p.add_argument("--check-tiles", type=int, default=6, help="batch tiles also checked against the oracle")
    return p.parse_args()


def batch_tiles(nh, n_u8: int, n_f32: int):
    """Uneven tiles: sides uniform in 512..4096 (multiples of 16), disks in
    proportion to area (200 per 512^2), seed 0x5EED0002 + index; float tiles are
    the same construction divided by 255."""
    import numpy as np
    rng = np.random.default_rng(0x5EED0002)
    out = []
    for k in range(n_u8 + n_f32):
        h = 512 + 16 * int(rng.integers(0, 225))
        w = 512 + 16 * int(rng.integers(0, 225))
        mask = nh.gen_disks_u8(h, w, round(200 * h * w / 512 ** 2), 0x5EED0002 + k)
        marker = nh.gen_marker_u8(mask, 30)
        if k >= n_u8:
            out.append((nh.scale_f32(marker), nh.scale_f32(mask), float(np.float32(128) / np.float32(255))))
        else:
            out.append((marker, mask, 128.0))
    return out


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
    # per-tile launches
    def per_tile():
        return [nh.recon_label(dm, dk, 8, thr, 8, recon_out=r2, labels_out=l2, ctx=ctx)[2]
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
    exact = None
    if args.check_tiles > 0:
        from oracle import oracle as orc
        idx = list(range(min(args.check_tiles // 2, args.tiles_u8))) + \
            list(range(args.tiles_u8, args.tiles_u8 + min(args.check_tiles - args.check_tiles // 2, args.tiles_f32)))
        exact = True
        for i in idx:
            marker, mask, thr = tiles[i]
            wr = orc.imreconstruct(marker, mask, 8)
            bw = np.where(wr > (np.float32(thr) if wr.dtype == np.float32 else thr), 255, 0).astype(np.uint8)
            wl, wn = orc.bwlabel2(bw, 8, relab=True)
            _, _, r, lab, _, _, _ = dtiles[i]
            exact &= bool(np.array_equal(r.cpu().numpy(), wr) and np.array_equal(lab.cpu().numpy(), wl)
                          and counts[i] == wn)
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
        "config": {"workload": f"{args.tiles_u8} u8 + {args.tiles_f32} f32 tiles, sides 512-4096 step 16: "
                               "recon8 -> >128 (f32: >128/255) -> bwlabel8", "pixels": total_px,
                   "tiles": len(tiles)},
    }
    b.close()
    print(json.dumps(out))


TILE_PX = 62 * 62   # interior pixels of one reconstruction tile visit


def roofline(kt: dict, px: int, visits: tuple, steps: int) -> dict:
    """HBM roofline of the kernel kind with the most time per step.

    Algorithmic bytes per pixel (compulsory I/O of the operation a launch
    performs): reconstruction 3 B (marker, mask in; state out), fillHoles 2 B
    (image in; state out), labelling read 1 B, label output 1 B in + 4 B out.
    A reconstruction first pass processes every pixel of the image; a
    propagation round processes the tiles it visits (62x62 pixels each).
    """
    v_recon, v_fill = visits
    alg_step = {
        "recon_first": (3 + 2) * px,
        "recon_round": (3 * v_recon + 2 * v_fill) * TILE_PX,
        "ccl_local": 1 * px,
        "ccl_final": 5 * px,
        "ccl_merge": 0,
        "ccl_small": 0,
    }
    kind = max(kt, key=lambda k: kt[k][0])
    ms_step = kt[kind][0] / steps
    launches_step = kt[kind][1] / steps
    alg_launch = alg_step[kind] / launches_step if launches_step else 0.0
    avg_launch_s = ms_step / launches_step / 1e3 if launches_step else float("nan")
    achieved = alg_launch / avg_launch_s / 1e9 if launches_step else 0.0
    return {"bound": "hbm", "kernel": kind, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
            "alg_bytes_per_launch": int(alg_launch), "avg_launch_us": round(avg_launch_s * 1e6, 3),
            "launches_per_step": round(launches_step, 2),
            "per_kind_ms_per_step": {k: round(v[0] / steps, 4) for k, v in kt.items()}}


def cpu_baseline(marker, mask, runs: int) -> dict:
    """The oracle's single-threaded restatement of the reference chain, timed
    on this host on `runs` repetitions of the same 4096x4096 workload."""
    from oracle import oracle as orc
    orc.chain_u8(marker, mask)   # page in
    t0 = time.perf_counter()
    for _ in range(runs):
        orc.chain_u8(marker, mask, 8, 4, 128, 8, 11, 1000)
    dt = time.perf_counter() - t0
    return {"value": round(runs * mask.size / dt / 1e6, 3), "unit": "Mpix/s", "cores": 1, "kind": "port",
            "sample": f"{runs} runs of the full {mask.shape[0]}x{mask.shape[1]} chain, single thread, "
                      f"{dt:.1f} s", "host_cpus": os.cpu_count()}


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this hot path runs on one GPU; multi-GPU sharding is a design note only "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import nscale_hip as nh

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.workload == "batch":
        return bench_batch(args)
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
    stream = torch.cuda.current_stream(dev)

    def step():
        return nh.segment_chain(d_marker, d_mask, 8, 4, 128.0, 8, 11, 1000, recon_out=d_recon, fill_out=d_fill,
                                labels_out=d_lab, ctx=ctx, stream=stream)[3]

    ncomp = step()
    torch.cuda.synchronize()
    exact = None
    if not args.no_check:
        from oracle import oracle as orc
        er, ef, elab, en = orc.chain_u8(marker, mask, 8, 4, 128, 8, 11, 1000)
        exact = bool(np.array_equal(d_recon.cpu().numpy(), er) and np.array_equal(d_fill.cpu().numpy(), ef)
                     and np.array_equal(d_lab.cpu().numpy(), elab) and ncomp == en)
        log(f"bit-exact vs oracle: {exact} (components {ncomp} / {en})")
        if not exact:
            sys.exit(1)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.set_timing(True)
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        step()
        lat.append((time.perf_counter() - s0) * 1e3)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kt = ctx.read_timing()
    ctx.set_timing(False)
    stats = ctx.recon_stats(0)
    fstats = ctx.recon_stats(1)
    log("kernel time per step (ms): " + ", ".join(f"{k}={v[0] / args.steps:.4f} ({v[1] / args.steps:.1f} launches)"
                                                  for k, v in kt.items()))
    log(f"recon stage {stats}; fill stage {fstats}")
    roof = roofline(kt, n * n, (stats["tile_visits"], fstats["tile_visits"]), args.steps)
    cpu = None if args.no_cpu else cpu_baseline(marker, mask, args.cpu_runs)
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
