"""Writes the golden fixtures under tests/golden/ (run from the repo root).

c0_mask.sha256      SHA-256 of the 512x512 disk generator output at seed
                    0x5EED0000 (200 disks), freezing the generator.
small_cases.npz     small seeded inputs with the oracle's outputs for
                    imreconstruct (u8/i32/f32, 4/8), imreconstructBinary,
                    imfillHoles, bwlabel2 and bwareaopen2; the oracle itself is
                    pinned by tests/test_oracle.py before these are written.
"""
import hashlib
import sys
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(REPO))
sys.path.insert(0, str(REPO / "emory-cci-fast-ia-gsoc_amd"))

import nscale_hip as nh  # noqa: E402
from oracle import oracle as orc  # noqa: E402


def main():
    out = Path(__file__).resolve().parent
    a = nh.gen_disks_u8(512, 512, 200, 0x5EED0000)
    (out / "c0_mask.sha256").write_text(hashlib.sha256(a.tobytes()).hexdigest() + "  c0 512x512 disks seed 0x5EED0000\n")
    rng = np.random.default_rng(20240601)
    d = {}
    mask = nh.gen_disks_u8(96, 130, 12, 0x5EED00AA)
    marker = nh.gen_marker_u8(mask, 30)
    d["disk_mask"], d["disk_marker"] = mask, marker
    for conn in (4, 8):
        d[f"recon_u8_c{conn}"] = orc.imreconstruct(marker, mask, conn)
        d[f"recon_i32_c{conn}"] = orc.imreconstruct(marker.astype(np.int32) * 257, mask.astype(np.int32) * 257, conn)
        d[f"recon_f32_c{conn}"] = orc.imreconstruct(marker.astype(np.float32) / np.float32(255),
                                                    mask.astype(np.float32) / np.float32(255), conn)
        d[f"fill_gray_c{conn}"] = orc.fill_holes(d[f"recon_u8_c{conn}"], False, conn)
        d[f"fill_bin_c{conn}"] = orc.fill_holes(d[f"recon_u8_c{conn}"], True, conn)
    bmask = ((rng.random((70, 90)) < 0.55) * 255).astype(np.uint8)
    bseed = ((rng.random((70, 90)) < 0.02) * 255).astype(np.uint8)
    d["bin_mask"], d["bin_seed"] = bmask, bseed
    for conn in (4, 8):
        d[f"recon_bin_c{conn}"] = orc.imreconstruct_binary(bseed, bmask, conn)
        d[f"label_root_c{conn}"], _ = orc.bwlabel2(bmask, conn, relab=False)
        d[f"label_canon_c{conn}"], _ = orc.bwlabel2(bmask, conn, relab=True)
        d[f"areaopen_c{conn}"], _ = orc.bwareaopen2(bmask, False, False, 3, 12, conn)
    np.savez_compressed(out / "small_cases.npz", **d)
    print("wrote", out)


if __name__ == "__main__":
    main()
