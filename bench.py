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
    stats = ctx.recon_stats()
    log("kernel time per step (ms): " + ", ".join(f"{k}={v[0] / args.steps:.4f} ({v[1] / args.steps:.1f} launches)"
                                                  for k, v in kt.items()))
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
                   "tiles": stats["tiles"]},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
