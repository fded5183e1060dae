"""Synthetic request images of the serving workload (host numpy arrays).

A request is a square image of the disks generator (200 disks per 512x512 of
area, intensities 50..220 plus noise, marker = mask - 30 saturating), either
uint8 with threshold 128 or int32 with the uint8 values scaled by 257 (0..65535)
and threshold 32768.  Sides are log-uniform in [256, 4096]; 20% of requests are
int32.  Request i of a trace with base seed S uses generator seed S + i."""
from __future__ import annotations

import numpy as np

from . import gen_disks_u8, gen_marker_u8, scale_i32

SERVE_SEED = 0x5EED0004


def request_images(side: int, kind: str, seed: int):
    mask = gen_disks_u8(side, side, max(1, round(200 * side * side / 512 ** 2)), seed)
    marker = gen_marker_u8(mask, 30)
    if kind == "i32":
        return scale_i32(marker, 257), scale_i32(mask, 257), 32768.0
    return marker, mask, 128.0


def serve_trace(n: int, seed: int = SERVE_SEED, min_side: int = 256, max_side: int = 4096, i32_frac: float = 0.2):
    """(side, kind, seed) of n requests."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        side = int(round(float(np.exp(rng.uniform(np.log(min_side), np.log(max_side))))))
        kind = "i32" if rng.random() < i32_frac else "u8"
        out.append((side, kind, seed + 1 + i))
    return out
