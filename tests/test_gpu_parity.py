"""HIP path vs the oracle, bit-exact, through the C ABI (nscale_hip)."""
import numpy as np
import pytest

import brute

pytestmark = pytest.mark.gpu

# tile edges of the reconstruction (62) and labelling (64 x 32) kernels
SHAPES = [(1, 1), (1, 130), (130, 1), (3, 5), (61, 62), (62, 62), (63, 64), (64, 65), (96, 33), (124, 125),
          (130, 300), (257, 190)]


def rand_case(rng, shape, dtype, hi=12, p=0.1):
    mask = rng.integers(0, hi, size=shape).astype(dtype)
    marker = np.where(rng.random(shape) < p, rng.integers(0, hi, size=shape), 0).astype(dtype)
    return marker, mask


@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("shape", SHAPES)
def test_recon_u8_random(nh, orc, gpu, shape, conn):
    rng = np.random.default_rng(hash((shape, conn)) & 0xFFFFF)
    for _ in range(3):
        marker, mask = rand_case(rng, shape, np.uint8)
        got = nh.imreconstruct(marker, mask, conn)
        assert np.array_equal(got, orc.imreconstruct(marker, mask, conn))


@pytest.mark.parametrize("conn", [4, 8])
def test_recon_u8_c0_disks(nh, orc, gpu, conn):
    """512x512, 200 disks + noise, marker = mask - 30."""
    mask = nh.gen_disks_u8(512, 512, 200, 0x5EED0000)
    marker = nh.gen_marker_u8(mask, 30)
    got = nh.imreconstruct(marker, mask, conn)
    assert np.array_equal(got, orc.imreconstruct(marker, mask, conn))


def test_recon_u8_marker_above_mask(nh, orc, gpu):
    rng = np.random.default_rng(1)
    mask = rng.integers(0, 256, size=(200, 150)).astype(np.uint8)
    marker = rng.integers(0, 256, size=(200, 150)).astype(np.uint8)
    for conn in (4, 8):
        assert np.array_equal(nh.imreconstruct(marker, mask, conn), orc.imreconstruct(marker, mask, conn))


@pytest.mark.parametrize("conn", [4, 8])
def test_recon_i32_f32(nh, orc, gpu, conn):
    rng = np.random.default_rng(2 + conn)
    for shape in [(63, 64), (130, 257)]:
        m = rng.integers(-3000, 3000, size=shape).astype(np.int32)
        k = np.where(rng.random(shape) < 0.2, rng.integers(-5000, 3000, size=shape), -5000).astype(np.int32)
        assert np.array_equal(nh.imreconstruct(k, m, conn), orc.imreconstruct(k, m, conn))
        mf = (m.astype(np.float32) / np.float32(7.0))
        kf = (k.astype(np.float32) / np.float32(7.0))
        mf[mf == 0] = np.float32(0.0)
        kf[kf == 0] = np.float32(0.0)
        assert np.array_equal(nh.imreconstruct(kf, mf, conn), orc.imreconstruct(kf, mf, conn))


def test_known_answer_32x32(nh, gpu):
    from test_oracle import fixture_32x32
    img, seeds = fixture_32x32()
    out = nh.imreconstruct(seeds, img, 4)
    expect = img.copy()
    expect[20:, :] = 0
    assert np.array_equal(out, expect)


@pytest.mark.parametrize("dtype", [np.uint8, np.int32])
@pytest.mark.parametrize("conn", [4, 8])
def test_recon_binary(nh, orc, gpu, dtype, conn):
    rng = np.random.default_rng(5 + conn)
    for shape in [(1, 1), (62, 63), (150, 131)]:
        mask = ((rng.random(shape) < 0.55) * rng.integers(1, 4, size=shape)).astype(dtype)
        marker = ((rng.random(shape) < 0.03) * rng.integers(1, 9, size=shape)).astype(dtype)
        got = nh.imreconstructBinary(marker, mask, conn)
        assert np.array_equal(got, orc.imreconstruct_binary(marker, mask, conn))


@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("binary", [False, True])
def test_fill_holes_u8(nh, orc, gpu, conn, binary):
    rng = np.random.default_rng(7 + conn + 2 * binary)
    # tile edges: 62 (8-connectivity) and 126 (4-connectivity) interior pixels
    for shape in [(1, 1), (5, 70), (62, 62), (128, 190), (126, 126), (127, 253), (253, 127), (300, 380)]:
        bw = np.where(rng.random(shape) < 0.6, 255, 0).astype(np.uint8)
        assert np.array_equal(nh.imfillHoles(bw, binary, conn), orc.fill_holes(bw, binary, conn))
        g = rng.integers(0, 256, size=shape).astype(np.uint8)
        assert np.array_equal(nh.imfillHoles(g, binary, conn), orc.fill_holes(g, binary, conn))


@pytest.mark.parametrize("conn", [4, 8])
def test_fill_holes_i32(nh, orc, gpu, conn):
    """Both readings of invert<int> at INT_MIN: the default (wrap) and
    saturation, each against the oracle restated with the same reading."""
    rng = np.random.default_rng(17 + conn)
    g = rng.integers(0, 65536, size=(100, 77)).astype(np.int32)
    assert np.array_equal(nh.imfillHoles(g, False, conn), orc.fill_holes(g, False, conn))
    assert np.array_equal(nh.imfillHoles(g, True, conn), orc.fill_holes(g, True, conn))
    sat = nh.Context(0)
    try:
        assert sat.get_option(nh.OPT_I32_INVERT) == nh.I32_INVERT_WRAP
        sat.set_option(nh.OPT_I32_INVERT, nh.I32_INVERT_SATURATE)
        for shape in [(100, 77), (1, 1), (130, 200)]:
            g = rng.integers(0, 65536, size=shape).astype(np.int32)
            for binary in (False, True):
                assert np.array_equal(nh.imfillHoles(g, binary, conn, ctx=sat),
                                      orc.fill_holes(g, binary, conn, i32_saturate=True))
        with pytest.raises(nh.NsxError):
            sat.set_option(nh.OPT_I32_INVERT, 7)
    finally:
        sat.close()


def test_chain_i32_both_invert_readings(nh, orc, gpu):
    """The chain on an int32 image (a serving request's construction) under
    both readings of invert<int>, against the oracle's composed chain."""
    from chain_expect import expected_chain, request_images
    marker, mask, thr = request_images(nh, 300, "i32", 0x5EED0204)
    for sat in (False, True):
        c = nh.Context(0)
        try:
            if sat:
                c.set_option(nh.OPT_I32_INVERT, nh.I32_INVERT_SATURATE)
            r, f, lab, n = nh.segment_chain(marker, mask, 8, 4, thr, 8, 11, 1000, ctx=c)
        finally:
            c.close()
        er, ef, elab, en = expected_chain(orc, marker, mask, thr, i32_saturate=sat)
        assert np.array_equal(r, er) and np.array_equal(f, ef)
        assert np.array_equal(lab, elab) and n == en
        if not sat:
            assert n == 0 and (f == np.iinfo(np.int32).min).all()


@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("shape", [(1, 1), (1, 200), (200, 1), (31, 63), (32, 64), (33, 65), (97, 130), (300, 257)])
def test_bwlabel(nh, orc, gpu, shape, conn):
    rng = np.random.default_rng(hash((shape, conn, 1)) & 0xFFFFF)
    for p in (0.3, 0.6, 0.9):
        img = np.where(rng.random(shape) < p, rng.integers(1, 3, size=shape), 0).astype(np.uint8)
        for relab in (False, True):
            got, n = nh.bwlabel2(img, conn, relab, return_count=True)
            want, wn = orc.bwlabel2(img, conn, relab)
            assert np.array_equal(got, want)
            assert n == wn


@pytest.mark.parametrize("conn", [4, 8])
def test_bwareaopen(nh, orc, gpu, conn):
    rng = np.random.default_rng(23 + conn)
    for shape in [(64, 64), (130, 77)]:
        img = np.where(rng.random(shape) < 0.5, 255, 0).astype(np.uint8)
        for lo, hi in [(1, 1 << 30), (3, 12), (11, 1000)]:
            for flatten in (False, True):
                got, n = nh.bwareaopen2(img, False, flatten, lo, hi, conn)
                want, wn = orc.bwareaopen2(img, False, flatten, lo, hi, conn)
                assert np.array_equal(got, want) and n == wn
            lab, _ = orc.bwlabel2(img, conn, relab=False)
            got, n = nh.bwareaopen2(lab, True, False, lo, hi, conn)
            want, wn = orc.bwareaopen2(lab, True, False, lo, hi, conn)
            assert np.array_equal(got, want) and n == wn


def test_relabel(nh, orc, gpu):
    rng = np.random.default_rng(29)
    img = np.where(rng.random((90, 140)) < 0.5, 255, 0).astype(np.uint8)
    lab, _ = orc.bwlabel2(img, 8, relab=False)
    got, n = nh.relabel(lab.copy(), -1)
    want, wn = orc.relabel(lab, -1)
    assert np.array_equal(got, want) and n == wn


def test_threshold(nh, gpu):
    rng = np.random.default_rng(31)
    a = rng.integers(0, 256, size=(70, 90)).astype(np.uint8)
    assert np.array_equal(nh.threshold_gt(a, 128), np.where(a > 128, 255, 0).astype(np.uint8))


@pytest.mark.parametrize("shape,nd", [((256, 320), 60), ((512, 512), 200), ((700, 530), 300)])
def test_chain_device(nh, orc, gpu, shape, nd):
    import torch
    mask = nh.gen_disks_u8(shape[0], shape[1], nd, 0x5EED0000 + nd)
    marker = nh.gen_marker_u8(mask, 30)
    r, f, lab, n = nh.segment_chain(torch.from_numpy(marker).cuda(), torch.from_numpy(mask).cuda())
    er, ef, elab, en = orc.chain_u8(marker, mask)
    assert np.array_equal(r.cpu().numpy(), er)
    assert np.array_equal(f.cpu().numpy(), ef)
    assert np.array_equal(lab.cpu().numpy(), elab)
    assert n == en


def test_device_tensors_and_pitch(nh, orc, gpu):
    """Non-continuous device rows (a column slice keeps the parent's pitch)."""
    import torch
    mask = nh.gen_disks_u8(300, 400, 80, 99)
    marker = nh.gen_marker_u8(mask, 30)
    big_m = torch.zeros((300, 512), dtype=torch.uint8, device="cuda")
    big_k = torch.zeros((300, 512), dtype=torch.uint8, device="cuda")
    big_m[:, 50:450] = torch.from_numpy(mask).cuda()
    big_k[:, 50:450] = torch.from_numpy(marker).cuda()
    out = torch.zeros((300, 600), dtype=torch.uint8, device="cuda")[:, 100:500]
    nh.imreconstruct(big_k[:, 50:450], big_m[:, 50:450], 8, out=out)
    assert np.array_equal(out.cpu().numpy(), orc.imreconstruct(marker, mask, 8))


def test_errors(nh, gpu):
    with pytest.raises(nh.NsxError):
        nh.imreconstruct(np.zeros((3, 4), np.uint8), np.zeros((4, 3), np.uint8))
    with pytest.raises(nh.NsxError):
        nh.bwlabel2(np.zeros((3, 4), np.int32))
    with pytest.raises(nh.NsxError):
        nh.imfillHoles(np.zeros((3, 4), np.float32))


def test_golden_fixtures(nh, gpu):
    from pathlib import Path
    g = np.load(Path(__file__).parent / "golden" / "small_cases.npz")
    mask, marker = g["disk_mask"], g["disk_marker"]
    for conn in (4, 8):
        r = nh.imreconstruct(marker, mask, conn)
        assert np.array_equal(r, g[f"recon_u8_c{conn}"])
        assert np.array_equal(nh.imreconstruct(marker.astype(np.int32) * 257, mask.astype(np.int32) * 257, conn),
                              g[f"recon_i32_c{conn}"])
        assert np.array_equal(nh.imreconstruct(marker.astype(np.float32) / np.float32(255),
                                               mask.astype(np.float32) / np.float32(255), conn),
                              g[f"recon_f32_c{conn}"])
        assert np.array_equal(nh.imfillHoles(r, False, conn), g[f"fill_gray_c{conn}"])
        assert np.array_equal(nh.imfillHoles(r, True, conn), g[f"fill_bin_c{conn}"])
        assert np.array_equal(nh.imreconstructBinary(g["bin_seed"], g["bin_mask"], conn), g[f"recon_bin_c{conn}"])
        assert np.array_equal(nh.bwlabel2(g["bin_mask"], conn, False), g[f"label_root_c{conn}"])
        assert np.array_equal(nh.bwlabel2(g["bin_mask"], conn, True), g[f"label_canon_c{conn}"])
        got, _ = nh.bwareaopen2(g["bin_mask"], False, False, 3, 12, conn)
        assert np.array_equal(got, g[f"areaopen_c{conn}"])


def test_bwlabel_large_repeated(nh, orc, gpu):
    """Labelling a 4096x4096 image several times gives the oracle's labels
    every time (a race in the tile-local flatten once left single pixels
    pointing at a non-root, about once per few hundred megapixels)."""
    import torch
    mask = nh.gen_disks_u8(4096, 4096, 3000, 0x5EED0101)
    binimg = np.where(mask > 128, 255, 0).astype(np.uint8)
    want, wn = orc.bwlabel2(binimg, 8, relab=True)
    d = torch.from_numpy(binimg).cuda()
    out = torch.empty(d.shape, dtype=torch.int32, device="cuda")
    for rep in range(6):
        out.fill_(-7)
        _, n = nh.bwlabel2(d, 8, True, out=out, return_count=True)
        assert n == wn and np.array_equal(out.cpu().numpy(), want), f"repetition {rep}"


def test_empty_images(nh, gpu):
    """0-pixel images are accepted by every entry point and give 0 components,
    also right after a call that found some (no stale count)."""
    img = np.where(np.random.default_rng(3).random((40, 50)) < 0.5, 255, 0).astype(np.uint8)
    for e in (np.zeros((0, 7), np.uint8), np.zeros((5, 0), np.uint8)):
        _, n = nh.bwlabel2(img, 8, True, return_count=True)
        assert n > 0
        assert nh.imreconstruct(e, e, 8).shape == e.shape
        assert nh.imreconstructBinary(e, e, 4).shape == e.shape
        assert nh.imfillHoles(e, False, 4).shape == e.shape
        lab, n = nh.bwlabel2(e, 8, True, return_count=True)
        assert lab.shape == e.shape and n == 0
        out, n = nh.bwareaopen2(e, False, False, 1, 10, 8)
        assert out.shape == e.shape and n == 0
        _, _, lab, n = nh.segment_chain(e, e)
        assert lab.shape == e.shape and n == 0
        _, lab, n = nh.recon_label(e, e)
        assert n == 0


def test_more_than_int32_pixels_rejected(nh, gpu):
    """Images of 2^31-1 pixels or more are refused by the host checks before
    any device work (labels and tile lists index pixels with int32).  The
    oversized descriptors name host memory (a small numpy buffer), so nothing
    could reach a kernel even if the check regressed; a regression would fail
    on the host."""
    import ctypes as C
    buf = np.zeros(64, np.uint8)
    lab = np.zeros(64, np.int32)
    L = nh._lib
    big = L.NsxImg(buf.ctypes.data, 50000, 50000, 50000, L.NSX_U8, L.NSX_MEM_HOST)
    bigl = L.NsxImg(lab.ctypes.data, 200000, 50000, 50000, L.NSX_I32, L.NSX_MEM_HOST)
    ctx = nh.default_context(0)
    n = C.c_int32(0)
    assert ctx.lib.nsx_bwlabel(ctx.h, C.byref(big), C.byref(bigl), 8, 1, C.byref(n), None) == L.NSX_ERR_INVALID
    assert ctx.lib.nsx_imreconstruct(ctx.h, C.byref(big), C.byref(big), C.byref(big), 8, None) == L.NSX_ERR_INVALID
    assert b"2^31" in ctx.lib.nsx_last_error(ctx.h)


@pytest.mark.parametrize("cols", [7936, 7937])
def test_recon_fill_large(nh, orc, gpu, cols):
    """3968 x 7936 is exactly 8192 reconstruction tiles (62 x 62); one more
    column adds a partial tile column (8256 tiles).  Both give the oracle's
    recon and fillHoles."""
    import torch
    mask = nh.gen_disks_u8(3968, cols, 2000, 0x5EED0202)
    marker = nh.gen_marker_u8(mask, 30)
    r = nh.imreconstruct(torch.from_numpy(marker).cuda(), torch.from_numpy(mask).cuda(), 8).cpu().numpy()
    want = orc.imreconstruct(marker, mask, 8)
    assert np.array_equal(r, want)
    f = nh.imfillHoles(torch.from_numpy(want).cuda(), False, 4).cpu().numpy()
    assert np.array_equal(f, orc.fill_holes(want, False, 4))


def test_default_stream_ordering(nh, orc, gpu):
    """Device tensors on torch's default stream: inputs written by torch just
    before the call (behind a spin kernel that keeps the stream busy) are seen
    by the call, and outputs are complete for torch's next reads."""
    import torch
    assert torch.cuda.current_stream().cuda_stream == 0
    mask = nh.gen_disks_u8(300, 333, 80, 0x5EED0404)
    marker = nh.gen_marker_u8(mask, 30)
    dmask, dmarker = torch.from_numpy(mask).cuda(), torch.from_numpy(marker).cuda()
    rng = np.random.default_rng(5)
    few = rng.choice(np.array([5, 90, 200], np.uint8), size=(211, 517))
    fmark = np.where(rng.random(few.shape) < 0.02, rng.integers(0, 256, few.shape), 0).astype(np.uint8)
    dfew, dfmark = torch.from_numpy(few).cuda(), torch.from_numpy(fmark).cuda()
    torch.cuda.synchronize()

    def busy_then(*xs):
        torch.cuda._sleep(20_000_000)   # the default stream stays busy ...
        return [x.clone() for x in xs]  # ... and these copies land after it

    k, m = busy_then(dmarker, dmask)
    assert np.array_equal(nh.imreconstruct(k, m, 8).cpu().numpy(), orc.imreconstruct(marker, mask, 8))
    k, m = busy_then(dfmark, dfew)   # few values: the level-set path
    assert np.array_equal(nh.imreconstruct(k, m, 8).cpu().numpy(), orc.imreconstruct(fmark, few, 8))
    (m,) = busy_then(dmask)
    assert np.array_equal(nh.imfillHoles(m, False, 4).cpu().numpy(), orc.fill_holes(mask, False, 4))
    bw = np.where(mask > 128, 255, 0).astype(np.uint8)
    (b,) = busy_then(torch.from_numpy(bw).cuda())
    assert np.array_equal(nh.bwlabel2(b, 8, True).cpu().numpy(), orc.bwlabel2(bw, 8, True)[0])
    k, m = busy_then(dmarker, dmask)
    r, f, lab, n = nh.segment_chain(k, m)
    er, ef, elab, en = orc.chain_u8(marker, mask)
    assert np.array_equal(lab.cpu().numpy(), elab) and n == en


@pytest.mark.parametrize("conn", [4, 8])
def test_imfill_seeded_and_bwselect(nh, orc, gpu, conn):
    """Seeded imfill (binary and gray) and bwselect against the oracle, host
    arrays and device tensors, at reconstruction tile edges."""
    import torch
    rng = np.random.default_rng(71 + conn)
    # (40, 512), (3, 1040): rows 16-byte aligned, 16 pixels per thread
    for shape in [(1, 1), (5, 70), (62, 63), (127, 130), (300, 257), (40, 512), (3, 1040)]:
        bw = np.where(rng.random(shape) < 0.55, 255, 0).astype(np.uint8)
        g = rng.integers(0, 256, size=shape).astype(np.uint8)
        seeds = np.where(rng.random(shape) < 0.01, rng.integers(1, 256, shape), 0).astype(np.uint8)
        for img in (bw, g):
            for binary in (False, True):
                want = orc.imfill(img, seeds, binary, conn)
                assert np.array_equal(nh.imfill(img, seeds, binary, conn), want), (shape, binary)
                got = nh.imfill(torch.from_numpy(img).cuda(), torch.from_numpy(seeds).cuda(), binary, conn)
                assert np.array_equal(got.cpu().numpy(), want)
            want = orc.bwselect(img, seeds, conn)
            assert np.array_equal(nh.bwselect(img, seeds, conn), want), shape
            got = nh.bwselect(torch.from_numpy(img).cuda(), torch.from_numpy(seeds).cuda(), conn)
            assert np.array_equal(got.cpu().numpy(), want)


def test_imfill_bwselect_checks(nh, gpu):
    """Types the reference does not instantiate are refused; empty images are fine."""
    a = np.zeros((4, 5), np.int32)
    s = np.zeros((4, 5), np.uint8)
    with pytest.raises(nh.NsxError):
        nh.imfill(a, s)
    with pytest.raises(nh.NsxError):
        nh.bwselect(np.zeros((4, 5), np.uint8), np.zeros((4, 6), np.uint8))
    e = np.zeros((0, 7), np.uint8)
    assert nh.imfill(e, e).shape == e.shape
    assert nh.bwselect(e, e).shape == e.shape


@pytest.mark.parametrize("conn", [4, 8])
def test_imhmin(nh, orc, gpu, conn):
    """imhmin (uchar and float, negative floats included) against the oracle,
    host arrays and device tensors; uchar h must be an integer in 0..255."""
    import torch
    rng = np.random.default_rng(91 + conn)
    for shape in [(1, 1), (61, 62), (130, 257), (40, 512)]:
        u = rng.integers(0, 256, size=shape).astype(np.uint8)
        for h in (0, 9, 255):
            want = orc.imhmin(u, h, conn)
            assert np.array_equal(nh.imhmin(u, h, conn), want), (shape, h)
            assert np.array_equal(nh.imhmin(torch.from_numpy(u).cuda(), h, conn).cpu().numpy(), want)
        f = (rng.integers(-400, 400, size=shape) / np.float32(64)).astype(np.float32)
        for h in (0.5, 2.0):
            want = orc.imhmin(f, h, conn)
            assert np.array_equal(nh.imhmin(f, h, conn), want), (shape, h)
    with pytest.raises(nh.NsxError):
        nh.imhmin(np.zeros((3, 3), np.uint8), 1.5)
    with pytest.raises(nh.NsxError):
        nh.imhmin(np.zeros((3, 3), np.int32), 1)


def test_bounding_box(nh, orc, gpu):
    """boundingBox against the oracle (pinned to the reference's own
    ConnComponents.cpp): canonical labels (background 0), root-index labels
    (background -1), host and device, a stray 0 label that ends the scan, and
    labels outside [-1, rows*cols] refused."""
    import torch
    rng = np.random.default_rng(101)
    for shape in [(1, 1), (33, 65), (257, 300), (700, 530)]:
        img = np.where(rng.random(shape) < 0.45, 255, 0).astype(np.uint8)
        canon, _ = orc.bwlabel2(img, 8, True)
        roots, _ = orc.bwlabel2(img, 8, False)
        for labels, bg in ((canon, 0), (roots, -1)):
            wn, wb = orc.bounding_box(labels, bg)
            n, b = nh.bounding_box(labels, bg)
            assert n == wn and np.array_equal(b, wb), (shape, bg)
            n, b = nh.bounding_box(torch.from_numpy(labels).cuda(), bg)
            assert n == wn and np.array_equal(b, wb)
        stray = roots.copy()
        stray[shape[0] // 2, shape[1] // 2] = 0 if roots[shape[0] // 2, shape[1] // 2] != 0 else 5
        wn, wb = orc.bounding_box(stray, -1)
        n, b = nh.bounding_box(stray, -1)
        assert n == wn and np.array_equal(b, wb)
    # isolated pixels: 1024 labels per 64x64 tile, more than a tile's table holds
    dots = np.zeros((130, 200), np.uint8)
    dots[::2, ::2] = 255
    canon, _ = orc.bwlabel2(dots, 8, True)
    for labels in (canon, torch.from_numpy(canon).cuda()):
        wn, wb = orc.bounding_box(canon, 0)
        n, b = nh.bounding_box(labels, 0)
        assert n == wn and np.array_equal(b, wb)
    bad = np.full((4, 4), 7, np.int32)
    bad[1, 1] = 100000
    with pytest.raises(nh.NsxError):
        nh.bounding_box(bad, 0)
    assert nh.bounding_box(np.zeros((0, 3), np.int32))[0] == 0


def _plateau_images(rng, shape):
    """Few-valued noise (many plateaus), a blocky smooth field and a constant."""
    few = rng.integers(0, 4, size=shape).astype(np.uint8) * 60
    by, bx = max(1, shape[0] // 7), max(1, shape[1] // 9)
    coarse = rng.integers(0, 6, size=(shape[0] // by + 1, shape[1] // bx + 1)).astype(np.uint8) * 40
    blocky = np.kron(coarse, np.ones((by, bx), np.uint8))[: shape[0], : shape[1]].copy()
    return [few, blocky, np.full(shape, 17, np.uint8)]


@pytest.mark.parametrize("conn", [4, 8])
def test_local_extrema(nh, orc, gpu, conn):
    """localMaxima / localMinima (uchar and float) against the oracle: random
    and plateau-heavy images, constants, a serpentine plateau that needs many
    propagation steps, host arrays and device tensors."""
    import torch
    rng = np.random.default_rng(131 + conn)
    for shape in [(1, 1), (1, 40), (61, 62), (130, 257), (300, 517)]:
        imgs = [rng.integers(0, 256, size=shape).astype(np.uint8)] + _plateau_images(rng, shape)
        for u in imgs:
            f = (u.astype(np.float32) - 100.0) / np.float32(8)
            for img in (u, f):
                for minima in (False, True):
                    want = orc.local_extrema(img, conn, minima)
                    fn = nh.localMinima if minima else nh.localMaxima
                    assert np.array_equal(fn(img, conn), want), (shape, img.dtype, minima)
            want = orc.local_extrema(u, conn, False)
            assert np.array_equal(nh.localMaxima(torch.from_numpy(u).cuda(), conn).cpu().numpy(), want)
    # a serpentine plateau at 1e9 (where v - 1 rounds back to v in float, so
    # the whole snake stays a candidate) between walls at -1: with a higher
    # pixel at its far end the whole snake is dropped, one mark step at a time
    snake = np.full((99, 101), -1, np.float32)
    snake[1::4, 1:-1] = 1e9
    for k in range(24):
        snake[4 * k + 2 : 4 * k + 5, -2 if k % 2 == 0 else 1] = 1e9
    for end in (None, 2e9):
        img = snake.copy()
        if end is not None:
            img[97, -2] = end
        for minima in (False, True):
            want = orc.local_extrema(img, conn, minima)
            fn = nh.localMinima if minima else nh.localMaxima
            assert np.array_equal(fn(img, conn), want), (end, minima)
    # u8 images whose rows are 16-byte aligned (whole 16-pixel groups per
    # thread), strided device views with a partial last group, and the 0 / 255
    # plateaus that the u8 path settles without a flood
    for shape in [(40, 512), (3, 4096), (17, 1040)]:
        u = rng.integers(0, 256, size=shape).astype(np.uint8)
        z = np.where(rng.random(shape) < 0.8, 0, rng.integers(1, 4, size=shape)).astype(np.uint8)
        one = np.zeros(shape, np.uint8)
        one[shape[0] // 2, shape[1] // 3] = 200
        for img in (u, z, 255 - z, one, 255 - one, np.zeros(shape, np.uint8), np.full(shape, 255, np.uint8)):
            for minima in (False, True):
                fn = nh.localMinima if minima else nh.localMaxima
                want = orc.local_extrema(img, conn, minima)
                assert np.array_equal(fn(img, conn), want), (shape, minima)
                assert np.array_equal(fn(torch.from_numpy(img).cuda(), conn).cpu().numpy(), want), (shape, minima)
    base = torch.from_numpy(rng.integers(0, 3, size=(21, 1024)).astype(np.uint8) * 50).cuda()
    obase = torch.zeros((21, 1024), dtype=torch.uint8, device="cuda")
    for w in (1000, 1017, 1024):
        img = base[:, :w]
        for minima in (False, True):
            fn = nh.localMinima if minima else nh.localMaxima
            o = obase[:, :w]
            fn(img, conn, out=o)
            want = orc.local_extrema(np.ascontiguousarray(img.cpu().numpy()), conn, minima)
            assert np.array_equal(o.cpu().numpy(), want), (w, minima)
    with pytest.raises(nh.NsxError):
        nh.localMaxima(np.zeros((3, 3), np.int32))
    with pytest.raises(nh.NsxError):
        nh.localMaxima(np.zeros((3, 3), np.uint8), out=np.zeros((3, 3), np.float32))
    e = np.zeros((0, 7), np.uint8)
    assert nh.localMaxima(e).shape == e.shape


def _dt_gpu_masks(rng):
    from scipy import ndimage
    out = []
    for shape, p in [((1, 1), 0.5), ((1, 300), 0.01), ((300, 1), 0.01), ((63, 65), 0.02), ((130, 257), 0.003),
                     ((257, 300), 0.0005)]:
        out.append(np.where(rng.random(shape) < p, 0, 255).astype(np.uint8))
    blob = ndimage.gaussian_filter(rng.random((300, 410)), 4) > 0.5
    out.append(np.where(blob, 255, 0).astype(np.uint8))
    out.append(np.full((70, 130), 255, np.uint8))
    out.append(np.zeros((65, 66), np.uint8))
    wide = np.full((2, 400), 255, np.uint8)
    wide[0, 399] = 0
    out.append(wide)
    # a front of more than a quarter of the pixels: the queue's entry buffers
    # start smaller and the call starts over with more room
    out.append(np.where(rng.random((200, 300)) < 0.6, 0, 255).astype(np.uint8))
    # distances above 1024, where float rounding makes ties between squares
    corner = np.full((1300, 1100), 255, np.uint8)
    corner[0, 0] = 0
    corner[1299, 1099] = 0
    out.append(corner)
    return out


def test_distance_transform(nh, orc, gpu):
    """distanceTransform in the reference's queue order (the default): float
    distances and int32 labels equal the oracle's bit for bit, ties between
    equally near backgrounds included; every label is a background pixel (or
    rows*cols*3) at that distance, and no neighbour's label is strictly closer;
    host and device inputs."""
    import torch
    rng = np.random.default_rng(151)
    for m in _dt_gpu_masks(rng):
        want_nn, want = orc.distance_transform(m)
        got = nh.distanceTransform(m)
        assert got.dtype == np.float32
        assert np.array_equal(got, want), m.shape
        nn = nh.distanceTransform(m, False)
        assert nn.dtype == np.int32
        assert np.array_equal(nn, want_nn.reshape(m.shape)), m.shape
        real = nn != 3 * m.size
        assert (m.reshape(-1)[nn[real]] == 0).all()
        assert np.array_equal(brute.label_distance(nn, m.shape), want)
        if m.size <= 200_000:
            assert brute.distance_fixed_point_violations(nn, m.shape) == 0
    m = _dt_gpu_masks(np.random.default_rng(151))[6]
    _, want = orc.distance_transform(m)
    assert np.array_equal(nh.distanceTransform(torch.from_numpy(m).cuda()).cpu().numpy(), want)
    with pytest.raises(nh.NsxError):
        nh.distanceTransform(np.zeros((3, 3), np.int32))
    with pytest.raises(nh.NsxError):
        nh.distanceTransform(np.zeros((3, 3), np.uint8), out=np.zeros((3, 3), np.int32))
    with pytest.raises(nh.NsxError):
        nh.distanceTransform(np.zeros((1, 65535), np.uint8))   # packed labels need cols < 65535
    e = np.zeros((0, 7), np.uint8)
    assert nh.distanceTransform(e).shape == e.shape


def _dt_tiles_ctx(nh):
    c = nh.Context(0)
    c.set_option(nh.OPT_DT_ORDER, nh.DT_ORDER_TILES)
    assert c.get_option(nh.OPT_DT_ORDER) == nh.DT_ORDER_TILES
    return c


def test_distance_transform_tiles(nh, orc, gpu):
    """The tile order (DT_ORDER_TILES) reaches a fixed point of the same rule:
    on these fixtures its distances equal the oracle's, every label is a
    background pixel at that distance and no neighbour's label is closer."""
    c = _dt_tiles_ctx(nh)
    try:
        for m in _dt_gpu_masks(np.random.default_rng(151)):
            _, want = orc.distance_transform(m)
            assert np.array_equal(nh.distanceTransform(m, ctx=c), want), m.shape
            nn = nh.distanceTransform(m, False, ctx=c)
            assert np.array_equal(brute.label_distance(nn, m.shape), want)
            if m.size <= 200_000:
                assert brute.distance_fixed_point_violations(nn, m.shape) == 0
    finally:
        c.close()


def _stuck_mask():
    rng = np.random.default_rng(2)
    m = None
    for p in (0.00002, 0.0001, 0.0005):
        m = np.where(rng.random((1000, 1000)) < p, 0, 255).astype(np.uint8)
    return m


def test_distance_transform_stuck_pixel(nh, orc, gpu):
    """A sparse 1000x1000 background where the oracle's FIFO propagation stops
    short of the exact distance at one pixel.  The queue order stops there too:
    labels and distances equal the oracle's everywhere, on repeated calls."""
    m = _stuck_mask()
    want_nn, want = orc.distance_transform(m)
    exact, _ = brute.exact_distance_with_far(m)
    assert int((want != exact).sum()) >= 1
    for _ in range(2):
        assert np.array_equal(nh.distanceTransform(m), want)
        assert np.array_equal(nh.distanceTransform(m, False), want_nn.reshape(m.shape))


def test_distance_transform_stuck_pixel_tiles(nh, orc, gpu):
    """The same image in tile order.  Which fixed point a stuck pixel reaches
    depends on the order of the offers, so this result is held to what the rule
    fixes: the same distances elsewhere, never nearer than the exact distance,
    and no neighbour's label strictly closer."""
    ctx = _dt_tiles_ctx(nh)
    try:
        _check_stuck_tiles(nh, orc, ctx)
    finally:
        ctx.close()


def _check_stuck_tiles(nh, orc, ctx):
    m = _stuck_mask()
    _, want = orc.distance_transform(m)
    exact, _ = brute.exact_distance_with_far(m)
    got = nh.distanceTransform(m, ctx=ctx)
    differ = got != want
    assert int(differ.sum()) <= 10
    assert (got >= exact).all()
    assert (want[differ] > exact[differ]).all() or (got[differ] > exact[differ]).all()
    nn = nh.distanceTransform(m, False, ctx=ctx)
    assert np.array_equal(brute.label_distance(nn, m.shape), got)
    assert brute.distance_fixed_point_violations(nn, m.shape) == 0
    # the same fixed point on every call
    assert np.array_equal(nh.distanceTransform(m, False, ctx=ctx), nn)
    assert np.array_equal(nh.distanceTransform(m, ctx=ctx), got)


def test_distance_transform_random(nh, orc, gpu):
    """Queue order against the oracle on 30 random masks: shapes from 1 to 320
    pixels a side, background densities from 1e-4 to 0.9, blobs, diagonal
    stripes, rings (many equally near backgrounds) and a few isolated
    backgrounds; labels and distances bit for bit.  The tile order, a fixed
    point of the same rule, is never nearer than the exact distance."""
    from scipy import ndimage
    rng = np.random.default_rng(0x5EED0D7)
    tiles = _dt_tiles_ctx(nh)
    try:
        for case in range(30):
            h, w = int(rng.integers(1, 321)), int(rng.integers(1, 321))
            kind = case % 5
            if kind == 0:
                m = np.where(rng.random((h, w)) < 10 ** rng.uniform(-4, -0.05), 0, 255)
            elif kind == 1:
                m = np.where(ndimage.gaussian_filter(rng.random((h, w)), 2.5) > 0.5, 255, 0)
            elif kind == 2:
                yy, xx = np.mgrid[0:h, 0:w]
                m = np.where(((xx + 2 * yy) // int(rng.integers(2, 9))) % 3 == 0, 0, 255)
            elif kind == 3:
                yy, xx = np.mgrid[0:h, 0:w]
                r = np.hypot(yy - h / 2, xx - w / 2)
                m = np.where((r.astype(np.int64) % int(rng.integers(3, 12))) == 0, 0, 255)
            else:
                m = np.full((h, w), 255)
                m.reshape(-1)[rng.integers(0, h * w, size=int(rng.integers(1, 4)))] = 0
            m = m.astype(np.uint8)
            want_nn, want = orc.distance_transform(m)
            assert np.array_equal(nh.distanceTransform(m, False), want_nn), (case, m.shape)
            assert np.array_equal(nh.distanceTransform(m), want), (case, m.shape)
            got_t = nh.distanceTransform(m, ctx=tiles)
            exact, _ = brute.exact_distance_with_far(m)
            assert (got_t >= exact).all(), (case, m.shape)
    finally:
        tiles.close()


@pytest.mark.parametrize("solve", [0, 8])
def test_distance_transform_solve_passes(nh, orc, gpu, solve):
    """The queue order with no solve pass launched per generation (every
    generation whose labels change is finished after a host round trip) and
    with eight: the same labels as the oracle."""
    ctx = nh.Context(0)
    try:
        ctx.set_tuning(nh._lib.NSX_TUNE_DT_SOLVE, solve)
        assert ctx.get_tuning(nh._lib.NSX_TUNE_DT_SOLVE) == solve
        rng = np.random.default_rng(152)
        for m in _dt_gpu_masks(rng)[4:7] + [_stuck_mask()[:400, :600]]:
            want_nn, want = orc.distance_transform(m)
            assert np.array_equal(nh.distanceTransform(m, False, ctx=ctx), want_nn.reshape(m.shape)), m.shape
            assert np.array_equal(nh.distanceTransform(m, ctx=ctx), want), m.shape
        if solve == 0:
            assert ctx.dt_resumes() > 0
    finally:
        ctx.close()
    with pytest.raises(nh.NsxError):
        nh.default_context().set_tuning(nh._lib.NSX_TUNE_DT_SOLVE, 9)


def test_distance_transform_4096(nh, orc, gpu):
    """At 4096x4096 (the disks image of the chain workload thresholded at 128)
    the reference's queue order stops short of the exact distance at 1812 stuck
    pixels; the queue order here gives the same labels and distances."""
    import torch
    from nscale_hip.workloads import chain_images
    _, mask = chain_images()
    binm = np.where(mask > 128, 255, 0).astype(np.uint8)
    want_nn, want = orc.distance_transform(binm)
    d = torch.from_numpy(binm).cuda()
    assert np.array_equal(nh.distanceTransform(d).cpu().numpy(), want)
    assert np.array_equal(nh.distanceTransform(d, False).cpu().numpy(), want_nn.reshape(binm.shape))


def test_distance_transform_4096_bounds_tiles(nh, orc, gpu):
    """The same image in tile order, which stops short at a few stuck pixels
    too, at different ones.  Pinned: never nearer than the exact Euclidean
    distance, a fixed point of the rule, exact at least as often as the
    reference's FIFO order, and different from it at fewer than 1 pixel in
    10 000 (1397 of 16.7 M when measured)."""
    import torch
    from nscale_hip.workloads import chain_images
    _, mask = chain_images()
    binm = np.where(mask > 128, 255, 0).astype(np.uint8)
    _, want = orc.distance_transform(binm)
    exact, _ = brute.exact_distance_with_far(binm)
    d = torch.from_numpy(binm).cuda()
    ctx = _dt_tiles_ctx(nh)
    try:
        got = nh.distanceTransform(d, ctx=ctx).cpu().numpy()
        nn = nh.distanceTransform(d, False, ctx=ctx).cpu().numpy()
    finally:
        ctx.close()
    assert (got >= exact).all()
    assert int((got != want).sum()) < binm.size // 10_000
    assert int((got != exact).sum()) <= int((want != exact).sum())
    assert np.array_equal(brute.label_distance(nn, binm.shape), got)
    assert brute.distance_fixed_point_violations(nn, binm.shape) == 0


def test_round_log(nh, orc, gpu):
    """The per-round work-list log of the chain's two propagations: entry 0 is
    the first pass (every tile for the reconstruction, the border tiles for the
    hole filling), and the later entries add up to the tile visits the
    statistics count; the same holds under the persistent drain."""
    mask = nh.gen_disks_u8(1000, 1300, 250, 0x5EED0031)
    marker = nh.gen_marker_u8(mask, 30)
    ntx, nty = -(-1300 // 62), -(-1000 // 62)
    for drain in (0, 1):
        ctx = nh.Context(0)
        ctx.set_tuning(nh._lib.NSX_TUNE_DRAIN, drain)
        try:
            r, f, lab, n = nh.segment_chain(marker, mask, 8, 4, 128.0, 8, 11, 1000, ctx=ctx)
            er, ef, elab, en = orc.chain_u8(marker, mask, 8, 4, 128, 8, 11, 1000)
            assert np.array_equal(f, ef) and np.array_equal(lab, elab) and n == en
            for stage, first in ((0, ntx * nty), (1, 2 * ntx + 2 * (nty - 2))):
                st = ctx.recon_stats(stage)
                log = ctx.round_log(stage)
                assert len(log) == st["rounds"] + 1, (drain, stage)
                assert log[0] == first, (drain, stage)
                assert sum(log[1:]) == st["tile_visits"], (drain, stage)
                assert all(v > 0 for v in log), (drain, stage)
        finally:
            ctx.close()
