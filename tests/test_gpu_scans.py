"""Image-wide row and column scans before a gray hole filling's propagation
rounds (NSX_TUNE_FILL_SCANS): the fill equals the oracle's with 0, 1, 3 and 8
scan pairs, u8 and int32 (saturating negation, whose fill is not constant), 4-
and 8-connected, on images at the scans' size threshold (24 tiles of 62 px on
each side), odd sizes, the widest rows scanned (4096) and one past it (no
scans)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(1488, 1488), (1501, 1700), (2049, 1535), (1490, 4096), (1600, 4097)]


def _images(nh, shape, seed):
    rows, cols = shape
    disks = nh.gen_disks_u8(rows, cols, max(20, rows * cols // 6000), seed)
    rng = np.random.default_rng(seed)
    yield "disks", disks
    yield "recon", nh.imreconstruct(nh.gen_marker_u8(disks, 30), disks, 8)
    yield "noise", rng.integers(0, 256, size=shape).astype(np.uint8)


@pytest.mark.parametrize("conn", [4, 8])
def test_fill_scans_u8_vs_oracle(nh, orc, gpu, conn):
    ctxs = {k: nh.Context(0) for k in (0, 1, 3, 8)}
    for k, c in ctxs.items():
        c.set_tuning(nh._lib.NSX_TUNE_FILL_SCANS, k)
    try:
        for i, shape in enumerate(SHAPES):
            for name, img in _images(nh, shape, 0x5EED0800 + i):
                want = orc.fill_holes(img, False, conn)
                for k, c in ctxs.items():
                    print(f"u8 conn={conn} shape={shape} image={name} scans={k}", flush=True)
                    got = nh.imfillHoles(img, False, conn, ctx=c)
                    assert np.array_equal(got, want), (shape, name, k)
    finally:
        for c in ctxs.values():
            c.close()


def test_fill_scans_i32_saturate_vs_oracle(nh, orc, gpu):
    ctxs = {k: nh.Context(0) for k in (0, 3)}
    for k, c in ctxs.items():
        c.set_tuning(nh._lib.NSX_TUNE_FILL_SCANS, k)
        c.set_option(nh.OPT_I32_INVERT, nh.I32_INVERT_SATURATE)
    try:
        for i, shape in enumerate(SHAPES[:3]):
            for name, img in _images(nh, shape, 0x5EED0810 + i):
                im32 = img.astype(np.int32) * 257 - 30000
                want = orc.fill_holes(im32, False, 4, i32_saturate=True)
                for k, c in ctxs.items():
                    print(f"i32 shape={shape} image={name} scans={k}", flush=True)
                    got = nh.imfillHoles(im32, False, 4, ctx=c)
                    assert np.array_equal(got, want), (shape, name, k)
    finally:
        for c in ctxs.values():
            c.close()


def test_fill_scans_tuning_bounds(nh, gpu):
    c = nh.Context(0)
    try:
        assert c.get_tuning(nh._lib.NSX_TUNE_FILL_SCANS) == 0
        for bad in (-1, 9):
            with pytest.raises(nh.NsxError):
                c.set_tuning(nh._lib.NSX_TUNE_FILL_SCANS, bad)
        c.set_tuning(nh._lib.NSX_TUNE_FILL_SCANS, 8)
        assert c.get_tuning(nh._lib.NSX_TUNE_FILL_SCANS) == 8
    finally:
        c.close()
