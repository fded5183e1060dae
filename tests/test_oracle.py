"""Pins the oracle (oracle/nscale_oracle.cpp) before anything is checked
against it: brute-force fixed points, textbook flood fills, scipy labelling,
the reference's own compiled ConnComponents.cpp, the 32x32 known-answer
fixture of test/segment/cuda/globalQueueTest.cu:446-583, and closed forms."""
import numpy as np
import pytest

import brute

SIZES = [(1, 1), (1, 7), (9, 1), (2, 2), (5, 8), (13, 11), (31, 33), (40, 64)]


def rand_pair(rng, shape, dtype, lo=0, hi=12):
    mask = rng.integers(lo, hi, size=shape).astype(dtype)
    marker = rng.integers(lo, hi, size=shape).astype(dtype)
    marker = np.where(rng.random(shape) < 0.15, marker, lo).astype(dtype)
    return marker, mask


@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("shape", SIZES)
def test_recon_u8_vs_fixpoint(orc, conn, shape):
    rng = np.random.default_rng(hash((conn,) + shape) & 0xFFFF)
    for _ in range(6):
        marker, mask = rand_pair(rng, shape, np.uint8)
        got = orc.imreconstruct(marker, mask, conn)
        assert np.array_equal(got, brute.recon_fixpoint(marker, mask, conn))


@pytest.mark.parametrize("conn", [4, 8])
def test_recon_marker_above_mask_and_noise(orc, conn):
    rng = np.random.default_rng(7)
    mask = rng.integers(0, 256, size=(37, 29)).astype(np.uint8)
    marker = rng.integers(0, 256, size=(37, 29)).astype(np.uint8)  # often > mask
    assert np.array_equal(orc.imreconstruct(marker, mask, conn), brute.recon_fixpoint(marker, mask, conn))


@pytest.mark.parametrize("conn", [4, 8])
def test_recon_i32_negative_values_see_zero_frame(orc, conn):
    # negative masks: the constant 0 frame raises border pixels to their mask
    rng = np.random.default_rng(11)
    mask = rng.integers(-50, 50, size=(23, 19)).astype(np.int32)
    marker = rng.integers(-80, 50, size=(23, 19)).astype(np.int32)
    got = orc.imreconstruct(marker, mask, conn)
    assert np.array_equal(got, brute.recon_fixpoint(marker, mask, conn, frame=0))
    # a wrong frame value would give a different answer
    assert not np.array_equal(got, brute.recon_fixpoint(marker, mask, conn, frame=np.int32(-1000)))


@pytest.mark.parametrize("conn", [4, 8])
def test_recon_order_isomorphism(orc, conn):
    """R(f(m), f(M)) = f(R(m, M)) for strictly increasing f with f(0)=0: the
    u8 result maps exactly onto i32 (x257) and f32 (/255) results."""
    mask = np.zeros((64, 80), np.uint8)
    rng = np.random.default_rng(3)
    mask[:] = rng.integers(0, 30, size=mask.shape)
    marker = np.clip(mask.astype(int) - 10, 0, 255).astype(np.uint8)
    r8 = orc.imreconstruct(marker, mask, conn)
    r32 = orc.imreconstruct(marker.astype(np.int32) * 257, mask.astype(np.int32) * 257, conn)
    rf = orc.imreconstruct(marker.astype(np.float32) / np.float32(255), mask.astype(np.float32) / np.float32(255), conn)
    assert np.array_equal(r32, r8.astype(np.int32) * 257)
    assert np.array_equal(rf, r8.astype(np.float32) / np.float32(255))


def test_recon_closed_forms(orc):
    rng = np.random.default_rng(5)
    mask = rng.integers(0, 200, size=(30, 40)).astype(np.uint8)
    for conn in (4, 8):
        assert np.array_equal(orc.imreconstruct(mask, mask, conn), mask)          # marker == mask
        assert np.array_equal(orc.imreconstruct(np.zeros_like(mask), mask, conn), np.zeros_like(mask))
        r = orc.imreconstruct(np.clip(mask.astype(int) - 20, 0, 255).astype(np.uint8), mask, conn)
        assert np.array_equal(orc.imreconstruct(r, mask, conn), r)                # idempotent


def fixture_32x32():
    """test/segment/cuda/globalQueueTest.cu:446-583: a 0/1 mask with three blobs
    and seeds at (8,8) and (4,24); the blob at rows 22-25 has no seed."""
    img = np.zeros((32, 32), np.uint8)
    rows = {4: (22, 29), 5: (22, 29), 6: [(7, 10), (22, 24), (28, 29)], 7: [(8, 11), (22, 24), (28, 29)],
            8: [(7, 10), (22, 24), (27, 29)], 9: [(6, 9), (22, 29)], 10: [(8, 11), (22, 29)], 11: (10, 13),
            22: (14, 17), 23: (15, 18), 24: (13, 16), 25: (15, 18)}
    for r, spans in rows.items():
        if isinstance(spans, tuple):
            spans = [spans]
        for a, b in spans:
            img[r, a:b + 1] = 1
    seeds = np.zeros((32, 32), np.uint8)
    seeds[4, 24] = 1
    seeds[8, 8] = 1
    return img, seeds


def test_known_answer_32x32(orc):
    img, seeds = fixture_32x32()
    out = orc.imreconstruct(seeds, img, 4)
    expect = img.copy()
    expect[20:, :] = 0          # unseeded blob stays 0
    assert np.array_equal(out, expect)
    assert out[22:26].sum() == 0 and out[4:12].sum() == img[4:12].sum()


@pytest.mark.parametrize("dtype", [np.uint8, np.int32])
@pytest.mark.parametrize("conn", [4, 8])
def test_recon_binary_vs_bfs(orc, dtype, conn):
    rng = np.random.default_rng(21 + conn)
    for shape in [(1, 1), (6, 9), (33, 31), (50, 64)]:
        for _ in range(4):
            mask = (rng.random(shape) < 0.55).astype(dtype) * rng.integers(1, 4, size=shape).astype(dtype)
            marker = ((rng.random(shape) < 0.05) * rng.integers(1, 9, size=shape)).astype(dtype)
            got = orc.imreconstruct_binary(marker, mask, conn)
            assert np.array_equal(got, brute.recon_binary_bfs(marker, mask, conn))


@pytest.mark.parametrize("conn", [4, 8])
def test_fill_holes_binary_images(orc, conn):
    rng = np.random.default_rng(31 + conn)
    for shape in [(1, 1), (3, 3), (20, 17), (48, 70)]:
        for _ in range(4):
            bw = np.where(rng.random(shape) < 0.6, 255, 0).astype(np.uint8)
            want = brute.fill_holes_binary_scipy(bw, conn)
            assert np.array_equal(orc.fill_holes(bw, False, conn), want)
            assert np.array_equal(orc.fill_holes(bw, True, conn), want)


@pytest.mark.parametrize("conn", [4, 8])
def test_fill_holes_gray_u8(orc, conn):
    rng = np.random.default_rng(41 + conn)
    for shape in [(1, 1), (7, 5), (30, 41)]:
        img = rng.integers(0, 256, size=shape).astype(np.uint8)
        got = orc.fill_holes(img, False, conn)
        assert np.array_equal(got, brute.fill_holes_gray_u8(img, conn))
        assert (got >= img).all()
        assert np.array_equal(orc.fill_holes(got, False, conn), got)


def test_fill_holes_binary_treats_only_255_as_wall(orc):
    img = np.zeros((7, 7), np.uint8)
    img[1:6, 1:6] = 255
    img[2:5, 2:5] = 128          # non-255 inside a 255 ring: a hole
    img[3, 3] = 0
    got = orc.fill_holes(img, True, 4)
    assert (got[1:6, 1:6] == 255).all() and got[0].sum() == 0


def _fill_i32_fixpoint(img, conn, saturate):
    """imfillHoles<int> from its definition: pad with INT_MIN, negate (wrapping
    or saturating at INT_MIN), reconstruct the frame marker by brute-force
    geodesic dilation, negate back, crop."""
    lo, hi = np.iinfo(np.int32).min, np.iinfo(np.int32).max

    def neg(a):
        out = np.negative(a)                 # numpy int32 negation wraps
        if saturate:
            out[a == lo] = hi
        return out

    mk = neg(np.pad(img, 1, constant_values=lo))
    mr = np.full_like(mk, lo)
    mr[0, :] = mr[-1, :] = mr[:, 0] = mr[:, -1] = hi
    r = brute.recon_fixpoint(mr, mk, conn, frame=0)
    return neg(r)[1:-1, 1:-1]


def test_fill_holes_i32(orc):
    """Both readings of invert<int> at INT_MIN, each against the brute-force
    fixed point built with that negation and against its closed form:
    wrapping gives the constant INT_MIN (gray and binary), saturation the
    textbook fill (gray) and the constant INT_MAX (binary)."""
    lo, hi = np.iinfo(np.int32).min, np.iinfo(np.int32).max
    rng = np.random.default_rng(9)
    img = rng.integers(0, 65536, size=(25, 31)).astype(np.int32)
    for conn in (4, 8):
        wrap = orc.fill_holes(img, False, conn)
        assert np.array_equal(wrap, _fill_i32_fixpoint(img, conn, False))
        assert (wrap == lo).all()
        assert (orc.fill_holes(img, True, conn) == lo).all()
        sat = orc.fill_holes(img, False, conn, i32_saturate=True)
        assert np.array_equal(sat, _fill_i32_fixpoint(img, conn, True))
        assert (orc.fill_holes(img, True, conn, i32_saturate=True) == hi).all()


def test_fill_holes_i32_tiny_fixture(orc):
    """A 5x5 ring of 9 around a 0 hole, with a 0 channel to the border on
    one side in the second image.  Saturation: the enclosed hole is filled to
    9, the open one is not.  Wrap: INT_MIN everywhere."""
    lo = np.iinfo(np.int32).min
    ring = np.zeros((5, 5), np.int32)
    ring[1:4, 1:4] = 9
    ring[2, 2] = 0
    filled = ring.copy()
    filled[2, 2] = 9
    opened = ring.copy()
    opened[2, 3] = 0          # hole joined to the background on the right
    for conn in (4, 8):
        assert np.array_equal(orc.fill_holes(ring, False, conn, i32_saturate=True), filled)
        assert np.array_equal(orc.fill_holes(opened, False, 4, i32_saturate=True), opened)
        assert (orc.fill_holes(ring, False, conn) == lo).all()
        assert (orc.fill_holes(opened, True, conn) == lo).all()
    # wrap and saturate agree everywhere but at INT_MIN
    a = np.array([lo, lo + 1, -1, 0, 1, np.iinfo(np.int32).max], np.int32)
    w, t = np.empty_like(a), np.empty_like(a)
    orc.lib().or_invert_i32(orc._p(a), orc._p(w), a.size, 0)
    orc.lib().or_invert_i32(orc._p(a), orc._p(t), a.size, 1)
    assert w.tolist() == [lo, -(lo + 1), 1, 0, -1, -np.iinfo(np.int32).max]
    assert t.tolist() == [np.iinfo(np.int32).max] + w.tolist()[1:]


def test_bwlabel_vs_reference_and_scipy(orc):
    if orc.ref() is None:
        pytest.skip("reference ConnComponents not built here")
    rng = np.random.default_rng(51)
    for shape in [(1, 1), (1, 9), (8, 1), (17, 23), (64, 64), (65, 130)]:
        for p in (0.3, 0.55, 0.8):
            for conn in (4, 8):
                img = np.where(rng.random(shape) < p, rng.integers(1, 3, size=shape), 0).astype(np.uint8)
                lab, _ = orc.bwlabel2(img, conn, relab=False)
                assert np.array_equal(lab, orc.ref_label(img, conn))
                can, n = orc.bwlabel2(img, conn, relab=True)
                rcan, rn = orc.ref_relabel(orc.ref_label(img, conn))
                assert np.array_equal(can, rcan) and n == rn
                assert np.array_equal(can, brute.canonical_labels(img, conn))
                assert np.array_equal(lab, brute.root_labels_from_canonical(can))


def test_bwareaopen_vs_reference(orc):
    if orc.ref() is None:
        pytest.skip("reference ConnComponents not built here")
    rng = np.random.default_rng(61)
    for shape in [(10, 10), (48, 77)]:
        for conn in (4, 8):
            img = np.where(rng.random(shape) < 0.5, 255, 0).astype(np.uint8)
            for lo, hi in [(1, 1 << 30), (3, 12), (11, 1000), (2, 3)]:
                out, n = orc.bwareaopen2(img, False, False, lo, hi, conn)
                ref_out, ref_n = orc.ref_area_threshold_labeled(orc.ref_label(img, conn), lo, hi)
                assert np.array_equal(out, ref_out) and n == ref_n
                # areaThreshold (the other reference variant) agrees
                rr, rn = orc.ref_area_threshold(img, lo, hi, conn)
                assert np.array_equal(out, rr) and n == rn
                flat, n2 = orc.bwareaopen2(img, False, True, lo, hi, conn)
                assert np.array_equal(flat, np.where(out > -1, 255, 0)) and n2 == n
                lab_in, _ = orc.bwlabel2(img, conn, relab=False)
                out2, n3 = orc.bwareaopen2(lab_in, True, False, lo, hi, conn)
                assert np.array_equal(out2, out) and n3 == n


def test_chain_composition(orc):
    mask = np.zeros((60, 70), np.uint8)
    rng = np.random.default_rng(71)
    mask[:] = rng.integers(0, 9, size=mask.shape)
    mask[10:30, 10:30] = 200
    mask[15:20, 15:20] = 40       # a hole
    mask[40:45, 50:53] = 180
    marker = np.clip(mask.astype(int) - 30, 0, 255).astype(np.uint8)
    r, f, lab, n = orc.chain_u8(marker, mask)
    assert np.array_equal(r, orc.imreconstruct(marker, mask, 8))
    assert np.array_equal(f, orc.fill_holes(r, False, 4))
    b = np.where(f > 128, 255, 0).astype(np.uint8)
    filt, _ = orc.bwareaopen2(b, False, False, 11, 1000, 8)
    can, n2 = orc.relabel(filt)
    assert np.array_equal(lab, can) and n == n2 == 2


def test_golden_fixtures(orc):
    """The committed fixtures are reproduced by the oracle."""
    from pathlib import Path
    g = np.load(Path(__file__).parent / "golden" / "small_cases.npz")
    mask, marker = g["disk_mask"], g["disk_marker"]
    for conn in (4, 8):
        r = orc.imreconstruct(marker, mask, conn)
        assert np.array_equal(r, g[f"recon_u8_c{conn}"])
        assert np.array_equal(r, brute.recon_fixpoint(marker, mask, conn))
        assert np.array_equal(orc.fill_holes(r, False, conn), g[f"fill_gray_c{conn}"])
        assert np.array_equal(orc.fill_holes(r, True, conn), g[f"fill_bin_c{conn}"])
        assert np.array_equal(orc.imreconstruct_binary(g["bin_seed"], g["bin_mask"], conn), g[f"recon_bin_c{conn}"])
        assert np.array_equal(orc.bwlabel2(g["bin_mask"], conn, False)[0], g[f"label_root_c{conn}"])
        assert np.array_equal(orc.bwlabel2(g["bin_mask"], conn, True)[0], g[f"label_canon_c{conn}"])


def test_chain_composition_matches_chain_u8(orc):
    """The per-operation composition used for the int32 serving requests equals
    the oracle's fused u8 chain."""
    import numpy as np
    from chain_expect import expected_chain
    from nscale_hip import gen_disks_u8, gen_marker_u8
    mask = gen_disks_u8(150, 170, 30, 0x5EED0004)
    marker = gen_marker_u8(mask, 30)
    want = orc.chain_u8(marker, mask, 8, 4, 128, 8, 11, 1000)
    got = expected_chain(orc, marker, mask, 128)
    for a, b in zip(want[:3], got[:3]):
        assert np.array_equal(a, b)
    assert want[3] == got[3]


def test_chain_i32_scaling_is_order_isomorphic(orc):
    """int32 requests are the u8 image scaled by 257 with threshold 32768.
    Under saturating negation reconstruction and fill commute with the
    scaling, and 257*v > 32768 holds exactly when v > 127, so the labels
    match the u8 chain thresholded at 127.  Under wrapping negation (the
    default) the reconstruction still commutes, the fill is INT_MIN
    everywhere and no component survives the threshold."""
    import numpy as np
    from chain_expect import expected_chain
    from nscale_hip import gen_disks_u8, gen_marker_u8, scale_i32
    mask = gen_disks_u8(120, 140, 25, 0x5EED0005)
    marker = gen_marker_u8(mask, 30)
    r8, f8, l8, n8 = expected_chain(orc, marker, mask, 127)
    m32, k32 = scale_i32(marker, 257), scale_i32(mask, 257)
    r32, f32, l32, n32 = expected_chain(orc, m32, k32, 32768, i32_saturate=True)
    assert np.array_equal(r32, r8.astype(np.int32) * 257)
    assert np.array_equal(f32, f8.astype(np.int32) * 257)
    assert np.array_equal(l32, l8) and n32 == n8
    r32, f32, l32, n32 = expected_chain(orc, m32, k32, 32768)
    assert np.array_equal(r32, r8.astype(np.int32) * 257)
    assert (f32 == np.iinfo(np.int32).min).all()
    assert not l32.any() and n32 == 0


@pytest.mark.parametrize("conn", [4, 8])
def test_bwselect_vs_scipy(orc, conn):
    """bwselect keeps the components (of the nonzero pixels) that hold a seed,
    with their values, for binary and multi-valued images."""
    rng = np.random.default_rng(51 + conn)
    for shape in [(1, 1), (7, 9), (40, 33)]:
        for vals in ([0, 255], [0, 3, 200, 255]):
            img = rng.choice(np.array(vals, np.uint8), size=shape).astype(np.uint8)
            seeds = np.where(rng.random(shape) < 0.06, rng.integers(1, 256, shape), 0).astype(np.uint8)
            assert np.array_equal(orc.bwselect(img, seeds, conn), brute.bwselect_scipy(img, seeds, conn))


@pytest.mark.parametrize("conn", [4, 8])
def test_imfill_seeded_closed_forms(orc, conn):
    """Seeded imfill: binary = flood from the seeds through the pixels below 255;
    gray = image | brute-force reconstruction of the seeded complement."""
    rng = np.random.default_rng(61 + conn)
    for shape in [(1, 1), (6, 11), (31, 29)]:
        bw = np.where(rng.random(shape) < 0.4, 255, 0).astype(np.uint8)
        g = rng.integers(0, 256, size=shape).astype(np.uint8)
        seeds = np.where(rng.random(shape) < 0.05, rng.integers(1, 256, shape), 0).astype(np.uint8)
        for img in (bw, g):
            assert np.array_equal(orc.imfill(img, seeds, True, conn), brute.imfill_binary_flood(img, seeds, conn))
            assert np.array_equal(orc.imfill(img, seeds, False, conn), brute.imfill_gray_fixpoint(img, seeds, conn))


@pytest.mark.parametrize("conn", [4, 8])
def test_imhmin_closed_form(orc, conn):
    """imhmin against complement / subtract / brute-force reconstruction /
    complement, for u8 and for float images with negative values (where the
    reconstruction's 0 frame shows)."""
    rng = np.random.default_rng(81 + conn)
    for shape in [(1, 1), (9, 7), (27, 31)]:
        u = rng.integers(0, 256, size=shape).astype(np.uint8)
        for h in (0, 1, 17, 255):
            assert np.array_equal(orc.imhmin(u, h, conn), brute.imhmin_fixpoint(u, h, conn))
        f = (rng.integers(-400, 400, size=shape) / np.float32(64)).astype(np.float32)
        for h in (0.25, 3.0):
            got = orc.imhmin(f, h, conn)
            assert np.array_equal(got, brute.imhmin_fixpoint(f, h, conn))
    # the 0 frame matters for negative floats: a frame of -1000 changes the answer
    f = -(rng.integers(1, 400, size=(15, 13)) / np.float32(8)).astype(np.float32)
    comp = (np.float32(0) - f).astype(np.float32)
    other = np.float32(0) - brute.recon_fixpoint((comp - np.float32(2)).astype(np.float32), comp, conn,
                                                  frame=np.float32(1000))
    assert not np.array_equal(orc.imhmin(f, 2.0, conn), other)


def test_bounding_box_vs_reference(orc):
    """boundingBox against the reference's own ConnComponents.cpp: canonical
    labels (background 0), root-index labels (background -1), and a stray 0
    label that ends the reference's scan."""
    if orc.ref() is None:
        pytest.skip("reference ConnComponents not built here")
    rng = np.random.default_rng(97)
    for shape in [(1, 1), (7, 9), (40, 57)]:
        img = np.where(rng.random(shape) < 0.5, 255, 0).astype(np.uint8)
        canon, _ = orc.bwlabel2(img, 8, True)
        roots, _ = orc.bwlabel2(img, 8, False)
        for labels, bg in ((canon, 0), (roots, -1)):
            assert [np.array_equal(a, b) if isinstance(a, np.ndarray) else a == b
                    for a, b in zip(orc.bounding_box(labels, bg), orc.ref_bounding_box(labels, bg))] == [True, True]
        stray = canon.copy()
        stray[shape[0] // 2, shape[1] // 2] = 0
        stray[stray == 0] = np.where(img[stray == 0] > 0, 0, -1)
        got, want = orc.bounding_box(stray, -1), orc.ref_bounding_box(stray, -1)
        assert got[0] == want[0] and np.array_equal(got[1], want[1])


@pytest.mark.parametrize("conn", [4, 8])
def test_local_extrema_closed_form(orc, conn):
    """localMaxima / localMinima against the plateau definition, for u8 and for
    float images with plateaus (few distinct values) and with fine values."""
    rng = np.random.default_rng(111 + conn)
    for shape in [(1, 1), (8, 9), (30, 27)]:
        u = rng.integers(0, 6, size=shape).astype(np.uint8) * 40
        f1 = (rng.integers(0, 5, size=shape) * np.float32(0.5)).astype(np.float32)   # plateaus, steps < 1
        f2 = (rng.integers(-300, 300, size=shape) / np.float32(37)).astype(np.float32)
        for img in (u, f1, f2, np.zeros(shape, np.uint8)):
            assert np.array_equal(orc.local_extrema(img, conn), brute.local_maxima_closed_form(img, conn))
            if img.dtype == np.uint8:
                inv = (255 - img.astype(np.int32)).astype(np.uint8)
            else:
                inv = (np.float32(0) - img).astype(np.float32)
            assert np.array_equal(orc.local_extrema(img, conn, minima=True),
                                  brute.local_maxima_closed_form(inv, conn))


def _u8_extrema_no_flood(img, conn):
    """u8 localMaxima without the marks' flood: 255 at pixels of value v >= 1
    whose reconstruction of image - 1 stays at v - 1, and at the 0 pixels
    exactly when the whole image is 0."""
    v = img.astype(np.int32)
    marker = np.maximum(v - 1, 0).astype(np.uint8)
    rec = brute.recon_fixpoint(marker, img, conn).astype(np.int32)
    keep = (v >= 1) & (rec == v - 1)
    if not v.any():
        keep = np.ones(img.shape, bool)
    return np.where(keep, 255, 0).astype(np.uint8)


def _u8_extrema_plateaus(img, conn):
    """The same through plateaus, without a reconstruction: a pixel of value
    v >= 1 reaches a strictly higher pixel through pixels >= v exactly when its
    plateau (the conn-connected pixels of value v around it) has a strictly
    higher neighbour.  255 on the nonzero plateaus with no such neighbour, and
    on the 0 pixels when the whole image is 0."""
    from scipy import ndimage
    v = img.astype(np.int32)
    pad = np.pad(v, 1, constant_values=-1)
    H, W = v.shape
    offs = [(-1, 0), (1, 0), (0, -1), (0, 1)] + ([(-1, -1), (-1, 1), (1, -1), (1, 1)] if conn == 8 else [])
    higher = np.zeros(v.shape, bool)
    for dy, dx in offs:
        higher |= pad[1 + dy:1 + dy + H, 1 + dx:1 + dx + W] > v
    keep = np.zeros(v.shape, bool)
    for val in np.unique(v[v >= 1]):
        lab, _ = ndimage.label(v == val, structure=brute.structure(conn))
        bad = np.unique(lab[higher & (lab > 0)])
        keep |= (lab > 0) & ~np.isin(lab, bad)
    if not v.any():
        keep = np.ones(v.shape, bool)
    return np.where(keep, 255, 0).astype(np.uint8)


@pytest.mark.parametrize("conn", [4, 8])
def test_local_extrema_u8_without_flood(orc, conn):
    """For u8 images the flood of marks over equal-valued candidates only ever
    drops 0-valued pixels: a candidate of value v >= 1 has no strictly higher
    neighbour, and a connected set of 0 pixels other than the whole image
    touches a nonzero pixel.  The GPU's u8 path relies on this, in the plateau
    form (one labelling of equal nonzero values); here both forms are checked
    against the oracle's flood on random, zero-heavy, plateau and constant
    images, for maxima and minima."""
    rng = np.random.default_rng(29 + conn)
    for shape in [(1, 1), (1, 7), (7, 1), (5, 6), (12, 31), (33, 40)]:
        imgs = [rng.integers(0, 256, size=shape).astype(np.uint8),
                (rng.integers(0, 3, size=shape) * 60).astype(np.uint8),
                np.where(rng.random(shape) < 0.9, 0, rng.integers(1, 4, size=shape)).astype(np.uint8),
                np.where(rng.random(shape) < 0.9, 255, rng.integers(250, 255, size=shape)).astype(np.uint8),
                np.zeros(shape, np.uint8), np.full(shape, 255, np.uint8), np.full(shape, 1, np.uint8)]
        one = np.zeros(shape, np.uint8)
        one[shape[0] // 2, shape[1] // 2] = 9
        imgs.append(one)
        for img in imgs:
            want = orc.local_extrema(img, conn)
            assert np.array_equal(want, _u8_extrema_no_flood(img, conn)), (shape, img)
            assert np.array_equal(want, _u8_extrema_plateaus(img, conn)), (shape, img)
            inv = (255 - img.astype(np.int32)).astype(np.uint8)
            want = orc.local_extrema(img, conn, minima=True)
            assert np.array_equal(want, _u8_extrema_no_flood(inv, conn))
            assert np.array_equal(want, _u8_extrema_plateaus(inv, conn))


@pytest.mark.parametrize("conn", [4, 8])
def test_bwselect_as_components(orc, conn):
    """bwselect keeps, with their values, the conn-connected components of the
    nonzero pixels that hold a pixel with a nonzero seed: the reconstruction of
    the seeded marker reaches exactly those.  The GPU's bwselect labels the
    components once on this ground; here it is checked against the oracle's
    reconstruction, on 0/255 and many-valued images."""
    from scipy import ndimage
    rng = np.random.default_rng(47 + conn)
    for shape in [(1, 1), (1, 9), (9, 1), (17, 23), (40, 61)]:
        for p in (0.3, 0.6):
            for img in (np.where(rng.random(shape) < p, 255, 0).astype(np.uint8),
                        np.where(rng.random(shape) < p, rng.integers(1, 256, size=shape), 0).astype(np.uint8)):
                seeds = np.where(rng.random(shape) < 0.05, rng.integers(1, 256, size=shape), 0).astype(np.uint8)
                lab, _ = ndimage.label(img != 0, structure=brute.structure(conn))
                hit = np.unique(lab[(seeds != 0) & (img != 0)])
                want = np.where(np.isin(lab, hit[hit > 0]), img, 0).astype(np.uint8)
                assert np.array_equal(orc.bwselect(img, seeds, conn), want), (shape, p)


@pytest.mark.parametrize("conn", [4, 8])
def test_imreconstruct_binary_as_components(orc, conn):
    """imreconstructBinary through components: the pixels reached are the
    conn-connected components of the nonzero mask pixels that hold a pixel
    with a nonzero marker; the output is the marker's value wherever the
    marker is nonzero (inside the mask or not), the type's maximum at the other
    reached pixels and 0 elsewhere.  Checked against the oracle for u8 and
    int32."""
    from scipy import ndimage
    rng = np.random.default_rng(53 + conn)
    for shape in [(1, 1), (1, 9), (9, 1), (17, 23), (40, 61)]:
        for dt, top in ((np.uint8, 255), (np.int32, np.iinfo(np.int32).max)):
            mask = np.where(rng.random(shape) < 0.55, rng.integers(1, 200, size=shape), 0).astype(dt)
            marker = np.where(rng.random(shape) < 0.06, rng.integers(1, 200, size=shape), 0).astype(dt)
            lab, _ = ndimage.label(mask != 0, structure=brute.structure(conn))
            hit = np.unique(lab[(marker != 0) & (mask != 0)])
            reached = np.isin(lab, hit[hit > 0])
            want = np.where(marker != 0, marker, np.where(reached, top, 0)).astype(dt)
            assert np.array_equal(orc.imreconstruct_binary(marker, mask, conn), want), (shape, dt)


def _dt_masks(rng):
    from scipy import ndimage
    out = []
    for shape, p in [((1, 1), 0.5), ((1, 40), 0.1), ((40, 1), 0.1), ((23, 37), 0.05), ((64, 65), 0.01),
                     ((130, 97), 0.002)]:
        out.append(np.where(rng.random(shape) < p, 0, 255).astype(np.uint8))
    blob = ndimage.gaussian_filter(rng.random((90, 110)), 3) > 0.5
    out.append(np.where(blob, 255, 0).astype(np.uint8))
    out.append(np.full((17, 19), 255, np.uint8))            # no background: the start point stays
    out.append(np.zeros((5, 6), np.uint8))                  # all background
    wide = np.full((2, 40), 255, np.uint8)
    wide[0, 39] = 0                                         # the start point (0, 6) beats it on the left
    out.append(wide)
    return out


def test_distance_transform_exact(orc):
    """distanceTransform: on these inputs the propagation finds the exact
    nearest background (or the start point (0, 3*rows) where that is nearer):
    distances equal scipy's EDT, every label is a background pixel (or
    rows*cols*3) at the reported distance, and no neighbour's label is closer."""
    rng = np.random.default_rng(141)
    for m in _dt_masks(rng):
        nn, dist = orc.distance_transform(m)
        want, _ = brute.exact_distance_with_far(m)
        assert np.array_equal(dist, want), m.shape
        n = m.size
        real = nn != 3 * n
        assert (m.reshape(-1)[nn[real]] == 0).all()
        assert np.array_equal(brute.label_distance(nn, m.shape), dist)
        assert brute.distance_fixed_point_violations(nn, m.shape) == 0
    wide = _dt_masks(np.random.default_rng(141))[-1]
    nn, _ = orc.distance_transform(wide)
    assert nn[0, 0] == 3 * wide.size and nn[0, 38] == 39


def test_distance_transform_stuck_pixel(orc):
    """Far from a sparse background the 8-neighbour propagation can stop short of
    the true nearest background: the result is still a fixed point of the rule,
    never nearer than the exact distance, and equals it almost everywhere."""
    rng = np.random.default_rng(2)
    m = None
    for _ in range(3):
        m = np.where(rng.random((1000, 1000)) < [0.00002, 0.0001, 0.0005][_], 0, 255).astype(np.uint8)
    nn, dist = orc.distance_transform(m)
    want, _ = brute.exact_distance_with_far(m)
    assert (dist >= want).all()
    assert 1 <= int((dist != want).sum()) <= 10
    assert brute.distance_fixed_point_violations(nn, m.shape) == 0
