"""u8 reconstruction and hole filling through the fixed-launch engines, the
component tree (NSX_ENGINE_TREE) and the bit planes (NSX_ENGINE_PLANES),
against the oracle: tile-edge shapes, plateaus, few-valued and noisy masks,
both connectivities; the same results as the propagation rounds
(NSX_ENGINE_ROUNDS)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (1, 300), (300, 1), (63, 64), (64, 64), (65, 65), (64, 129), (128, 128), (129, 200),
          (257, 300), (700, 530)]


def _ctx(nh, engine, levelset=0):
    c = nh.Context(0)
    c.set_tuning(nh._lib.NSX_TUNE_ENGINE, engine)
    c.set_tuning(nh._lib.NSX_TUNE_LEVELSET_LEVELS, levelset)
    return c


def _masks(rng, shape):
    yield "noise", rng.integers(0, 256, size=shape).astype(np.uint8)
    yield "few", (rng.integers(0, 4, size=shape) * 60).astype(np.uint8)
    yield "flat", np.full(shape, 77, np.uint8)
    yield "zeros", np.zeros(shape, np.uint8)
    smooth = np.cumsum(np.cumsum(rng.integers(-1, 2, size=shape), 0), 1)
    yield "smooth", np.clip(smooth - smooth.min(), 0, 255).astype(np.uint8)


ENGINES = ["tree", "planes"]


def _engine(nh, name):
    return {"tree": nh._lib.NSX_ENGINE_TREE, "planes": nh._lib.NSX_ENGINE_PLANES}[name]


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("conn", [4, 8])
def test_mtree_recon_vs_oracle(nh, orc, gpu, conn, engine):
    ctx = _ctx(nh, _engine(nh, engine))
    rng = np.random.default_rng(300 + conn)
    try:
        for shape in SHAPES:
            for name, mask in _masks(rng, shape):
                for mk in ("sparse", "sub30"):
                    if mk == "sparse":
                        marker = np.where(rng.random(shape) < 0.02, rng.integers(0, 256, size=shape), 0).astype(np.uint8)
                    else:
                        marker = np.clip(mask.astype(int) - 30, 0, 255).astype(np.uint8)
                    print(f"recon conn={conn} shape={shape} mask={name} marker={mk}", flush=True)
                    got = nh.imreconstruct(marker, mask, conn, ctx=ctx)
                    assert np.array_equal(got, orc.imreconstruct(marker, mask, conn)), (shape, name, mk)
    finally:
        ctx.close()


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("conn", [4, 8])
def test_mtree_fill_vs_oracle(nh, orc, gpu, conn, engine):
    ctx = _ctx(nh, _engine(nh, engine))
    rng = np.random.default_rng(400 + conn)
    try:
        for shape in SHAPES:
            for name, img in _masks(rng, shape):
                print(f"fill conn={conn} shape={shape} image={name}", flush=True)
                got = nh.imfillHoles(img, False, conn, ctx=ctx)
                assert np.array_equal(got, orc.fill_holes(img, False, conn)), (shape, name)
    finally:
        ctx.close()


@pytest.mark.parametrize("engine", ENGINES)
def test_mtree_disks_match_rounds(nh, orc, gpu, engine):
    """The disks construction at 1500 x 1700: fixed-launch engine == rounds ==
    oracle for recon (8) and fill (4), device buffers."""
    import torch
    mask = nh.gen_disks_u8(1500, 1700, 500, 0x5EED0601)
    marker = nh.gen_marker_u8(mask, 30)
    auto, rounds = _ctx(nh, _engine(nh, engine), 3), _ctx(nh, nh._lib.NSX_ENGINE_ROUNDS, 3)
    try:
        dm, dk = torch.from_numpy(marker).cuda(), torch.from_numpy(mask).cuda()
        a = nh.imreconstruct(dm, dk, 8, ctx=auto)
        b = nh.imreconstruct(dm, dk, 8, ctx=rounds)
        want = orc.imreconstruct(marker, mask, 8)
        assert np.array_equal(a.cpu().numpy(), want) and torch.equal(a, b)
        fa = nh.imfillHoles(a, False, 4, ctx=auto)
        fb = nh.imfillHoles(a, False, 4, ctx=rounds)
        assert np.array_equal(fa.cpu().numpy(), orc.fill_holes(want, False, 4)) and torch.equal(fa, fb)
    finally:
        auto.close()
        rounds.close()


def test_auto_engine_switches_on_long_paths(nh, orc, gpu):
    """NSX_ENGINE_AUTO (the default) on a many-valued maze (1024^2, about 1500
    propagation rounds): the first call passes 256 rounds and finishes through
    the bit planes, the next ones start with them; every result equals the
    oracle, for imreconstruct, recon_label and the chain's fill."""
    import torch
    from nscale_hip.workloads import maze_levels_images
    marker, mask = maze_levels_images(1024)
    want = orc.imreconstruct(marker, mask, 4)
    ctx = nh.Context(0)
    try:
        assert ctx.get_tuning(nh._lib.NSX_TUNE_ENGINE) == nh._lib.NSX_ENGINE_AUTO
        dm, dk = torch.from_numpy(marker).cuda(), torch.from_numpy(mask).cuda()
        for rep in range(3):
            got = nh.imreconstruct(dm, dk, 4, ctx=ctx)
            assert np.array_equal(got.cpu().numpy(), want), f"call {rep}"
        r, lab, n = nh.recon_label(dm, dk, 4, 100.0, 4, ctx=ctx)
        assert np.array_equal(r.cpu().numpy(), want)
        assert n == 1 and np.array_equal(lab.cpu().numpy(), (mask > 100).astype(np.int32))
        for rep in range(2):
            rr, ff, ll, nn = nh.segment_chain(dm, dk, 4, 4, 100.0, 4, 11, 1 << 30, ctx=ctx)
            assert np.array_equal(rr.cpu().numpy(), want)
            assert np.array_equal(ff.cpu().numpy(), orc.fill_holes(want, False, 4)), f"chain {rep}"
    finally:
        ctx.close()


def test_planes_reject_output_aliasing_an_input(nh, gpu):
    """As with every engine, an output that aliases an input is rejected at
    the boundary (nothing runs)."""
    import torch
    mask = nh.gen_disks_u8(300, 200, 40, 0x5EED0602)
    marker = nh.gen_marker_u8(mask, 30)
    ctx = _ctx(nh, nh._lib.NSX_ENGINE_PLANES)
    try:
        dm, dk = torch.from_numpy(marker).cuda(), torch.from_numpy(mask).cuda()
        with pytest.raises(nh.NsxError):
            nh.imreconstruct(dm, dk, 8, out=dm, ctx=ctx)
        with pytest.raises(nh.NsxError):
            nh.imfillHoles(dk, False, 4, out=dk, ctx=ctx)
    finally:
        ctx.close()


def test_planes_chain_matches_oracle(nh, orc, gpu):
    """segment_chain under NSX_ENGINE_PLANES on a 1030 x 1500 disks image."""
    import torch
    mask = nh.gen_disks_u8(1030, 1500, 300, 0x5EED0603)
    marker = nh.gen_marker_u8(mask, 30)
    er, ef, elab, en = orc.chain_u8(marker, mask, 8, 4, 128, 8, 11, 1000)
    ctx = _ctx(nh, nh._lib.NSX_ENGINE_PLANES)
    try:
        dm, dk = torch.from_numpy(marker).cuda(), torch.from_numpy(mask).cuda()
        r, f, lab, n = nh.segment_chain(dm, dk, 8, 4, 128.0, 8, 11, 1000, ctx=ctx)
        assert np.array_equal(r.cpu().numpy(), er) and np.array_equal(f.cpu().numpy(), ef)
        assert np.array_equal(lab.cpu().numpy(), elab) and n == en
    finally:
        ctx.close()
