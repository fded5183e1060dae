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
    return p.parse_args()


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
