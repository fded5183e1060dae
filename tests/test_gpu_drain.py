"""Persistent drain of the propagation rounds (NSX_TUNE_DRAIN = 1) vs the oracle
and vs per-round launches, bit-exact, through the C ABI.  Each test also checks
that the drain ran (a recon_drain launch was timed) and never gave up."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (1, 130), (130, 1), (61, 62), (62, 62), (63, 64), (124, 125), (130, 300), (257, 190)]


@pytest.fixture(scope="module")
def dctx(nh, gpu):
    c = nh.Context(0)
    c.set_tuning(nh._lib.NSX_TUNE_DRAIN, 1)
    c.set_tuning(nh._lib.NSX_TUNE_LEVELSET_LEVELS, 0)   # keep every u8 case on the propagation path
    assert c.get_tuning(nh._lib.NSX_TUNE_DRAIN) == 1
    yield c
    c.close()


def drained(nh, ctx, fn):
    """Runs fn() on ctx with timing on; returns its result and the drain launches."""
    ctx.set_timing(True)
    out = fn()
    ctx.synchronize()
    launches = ctx.read_timing()["recon_drain"][1]
    ctx.set_timing(False)
    assert ctx.drain_failures() == 0
    return out, launches


def rand_case(rng, shape, dtype, hi=12, p=0.1):
    mask = rng.integers(0, hi, size=shape).astype(dtype)
    marker = np.where(rng.random(shape) < p, rng.integers(0, hi, size=shape), 0).astype(dtype)
    return marker, mask


@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("shape", SHAPES)
def test_drain_recon_u8_random(nh, orc, dctx, shape, conn):
    rng = np.random.default_rng(hash((shape, conn, "drain")) & 0xFFFFF)
    for _ in range(2):
        marker, mask = rand_case(rng, shape, np.uint8)
        got, n = drained(nh, dctx, lambda: nh.imreconstruct(marker, mask, conn, ctx=dctx))
        assert n >= 1
        assert np.array_equal(got, orc.imreconstruct(marker, mask, conn))


@pytest.mark.parametrize("conn", [4, 8])
def test_drain_recon_disks(nh, orc, dctx, conn):
    mask = nh.gen_disks_u8(700, 530, 300, 0x5EED0007)
    marker = nh.gen_marker_u8(mask, 30)
    got, n = drained(nh, dctx, lambda: nh.imreconstruct(marker, mask, conn, ctx=dctx))
    assert n == 1
    assert np.array_equal(got, orc.imreconstruct(marker, mask, conn))
    assert np.array_equal(got, nh.imreconstruct(marker, mask, conn))   # per-round launches


@pytest.mark.parametrize("conn", [4, 8])
def test_drain_recon_i32_f32(nh, orc, dctx, conn):
    rng = np.random.default_rng(40 + conn)
    for shape in [(63, 64), (130, 257), (300, 200)]:
        m = rng.integers(-3000, 3000, size=shape).astype(np.int32)
        k = np.where(rng.random(shape) < 0.2, rng.integers(-5000, 3000, size=shape), -5000).astype(np.int32)
        got, n = drained(nh, dctx, lambda: nh.imreconstruct(k, m, conn, ctx=dctx))
        assert n == 1
        assert np.array_equal(got, orc.imreconstruct(k, m, conn))
        mf = m.astype(np.float32) / np.float32(7.0)
        kf = k.astype(np.float32) / np.float32(7.0)
        mf[mf == 0] = np.float32(0.0)
        kf[kf == 0] = np.float32(0.0)
        got, n = drained(nh, dctx, lambda: nh.imreconstruct(kf, mf, conn, ctx=dctx))
        assert n == 1
        assert np.array_equal(got, orc.imreconstruct(kf, mf, conn))


@pytest.mark.parametrize("conn", [4, 8])
def test_drain_binary(nh, orc, dctx, conn):
    """Binary reconstruction rounds under the drain: the binary seeded imfill
    (imreconstructBinary itself is one labelling and runs no rounds)."""
    rng = np.random.default_rng(50 + conn)
    for shape in [(62, 63), (150, 131), (400, 90)]:
        img = np.where(rng.random(shape) < 0.45, 255, 0).astype(np.uint8)
        seeds = np.where(rng.random(shape) < 0.03, rng.integers(1, 9, size=shape), 0).astype(np.uint8)
        got, n = drained(nh, dctx, lambda: nh.imfill(img, seeds, True, conn, ctx=dctx))
        assert n == 1
        assert np.array_equal(got, orc.imfill(img, seeds, True, conn))
        mask = ((rng.random(shape) < 0.55) * rng.integers(1, 4, size=shape)).astype(np.int32)
        marker = ((rng.random(shape) < 0.03) * rng.integers(1, 9, size=shape)).astype(np.int32)
        assert np.array_equal(nh.imreconstructBinary(marker, mask, conn, ctx=dctx),
                              orc.imreconstruct_binary(marker, mask, conn))


@pytest.mark.parametrize("conn", [4, 8])
def test_drain_fill(nh, orc, dctx, conn):
    """Gray and binary hole filling: conn 4 drains on four-wave tiles."""
    rng = np.random.default_rng(60 + conn)
    for shape in [(62, 62), (200, 333), (517, 260)]:
        img = rng.integers(0, 40, size=shape).astype(np.uint8)
        got, n = drained(nh, dctx, lambda: nh.imfillHoles(img, False, conn, ctx=dctx))
        assert n == 1
        assert np.array_equal(got, orc.fill_holes(img, False, conn))
        b = (rng.random(shape) < 0.45).astype(np.uint8) * 255
        got, n = drained(nh, dctx, lambda: nh.imfillHoles(b, True, conn, ctx=dctx))
        assert n == 1
        assert np.array_equal(got, orc.fill_holes(b, True, conn))
        # int32: the saturating invert (under the default wrap every int32
        # fill is the constant INT_MIN and runs no propagation)
        i = rng.integers(-500, 500, size=shape).astype(np.int32)
        dctx.set_option(nh.OPT_I32_INVERT, nh.I32_INVERT_SATURATE)
        try:
            got, n = drained(nh, dctx, lambda: nh.imfillHoles(i, False, conn, ctx=dctx))
        finally:
            dctx.set_option(nh.OPT_I32_INVERT, nh.I32_INVERT_WRAP)
        assert n == 1
        assert np.array_equal(got, orc.fill_holes(i, False, conn, i32_saturate=True))


def test_drain_chain_1024(nh, orc, dctx):
    """The segmentation chain with both propagations drained, twice in a row."""
    mask = nh.gen_disks_u8(1024, 1024, 190, 0x5EED0011)
    marker = nh.gen_marker_u8(mask, 30)
    er, ef, elab, en = orc.chain_u8(marker, mask, 8, 4, 128, 8, 11, 1000)
    for _ in range(2):
        (r, f, lab, n), k = drained(nh, dctx, lambda: nh.segment_chain(marker, mask, 8, 4, 128.0, 8, 11, 1000, ctx=dctx))
        assert k == 2
        assert np.array_equal(r, er) and np.array_equal(f, ef) and np.array_equal(lab, elab) and n == en
