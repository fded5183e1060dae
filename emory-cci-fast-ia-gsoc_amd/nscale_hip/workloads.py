"""Synthetic request images of the serving workload (host numpy arrays).

A request is a square image of the disks generator (200 disks per 512x512 of
area, intensities 50..220 plus noise, marker = mask - 30 saturating), either
uint8 with threshold 128 or int32 with the uint8 values scaled by 257 (0..65535)
and threshold 32768.  Sides are log-uniform in [256, 4096]; 20% of requests are
int32.  Request i of a trace with base seed S uses generator seed S + i."""
from __future__ import annotations

import numpy as np

from . import gen_disks_u8, gen_marker_u8, gen_maze_u8, scale_f32, scale_i32

SERVE_SEED = 0x5EED0004
BATCH_SEED = 0x5EED0002
CHAIN_SEED = 0x5EED0001
MAZE_SEED = 0x5EED0003


def chain_images(n: int = 4096, ndisks: int = 3000, seed: int = CHAIN_SEED):
    """The single-tile chain workload: an n x n disks image and its marker."""
    mask = gen_disks_u8(n, n, ndisks, seed)
    return gen_marker_u8(mask, 30), mask


def batch_tiles(n_u8: int = 256, n_f32: int = 32, seed: int = BATCH_SEED):
    """Uneven batch tiles: sides uniform in 512..4096 (multiples of 16), disks
    in proportion to area (200 per 512^2), generator seed `seed` + index; the
    float tiles are the same construction divided by 255 (threshold 128/255).
    Returns [(marker, mask, threshold)]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_u8 + n_f32):
        h = 512 + 16 * int(rng.integers(0, 225))
        w = 512 + 16 * int(rng.integers(0, 225))
        mask = gen_disks_u8(h, w, round(200 * h * w / 512 ** 2), seed + k)
        marker = gen_marker_u8(mask, 30)
        if k >= n_u8:
            out.append((scale_f32(marker), scale_f32(mask), float(np.float32(128) / np.float32(255))))
        else:
            out.append((marker, mask, 128.0))
    return out


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


def maze_levels_images(n: int = 16384, seed: int = MAZE_SEED, lo: int = 150, hi: int = 220):
    """A many-valued maze: the maze of gen_maze_u8 (6-px cells, 3-px corridors,
    seed `seed`) with every corridor pixel at an intensity uniform in lo..hi
    instead of 200, then integer noise uniform in -8..8 on every pixel
    (clamped to 0..255), as in the disks construction.  The marker is 0 except
    a 5x5 block at the entrance at 255 (clamped by the mask).  Its many distinct
    values keep the level-set path out, and its corridors carry values over
    paths millions of pixels long.  Returns (marker, mask)."""
    mask, _ = gen_maze_u8(n, n, seed)
    rng = np.random.default_rng(seed ^ 0x5A5A)
    corr = mask == 200
    out = np.where(corr, rng.integers(lo, hi + 1, size=mask.shape, dtype=np.int16), mask.astype(np.int16))
    out += rng.integers(-8, 9, size=mask.shape, dtype=np.int16)
    np.clip(out, 0, 255, out=out)
    mask = out.astype(np.uint8)
    marker = np.zeros_like(mask)
    marker[:5, :5] = 255
    return marker, mask
