"""nsx_imreconstruct_noalloc (the entry point that allocates nothing) through
the C ABI: the same output as nsx_imreconstruct on two calls in a row, the
second with a smaller image (the shared scratch array keeps the first call's
contents), for every dtype and connectivity; sizes past its limit are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _img(nh, t):
    return nh.describe(t)


def _noalloc(nh, ctx, marker, mask, conn):
    import torch
    out = torch.full_like(mask, 7)
    st = ctx.lib.nsx_imreconstruct_noalloc(ctx.h, C.byref(_img(nh, marker)), C.byref(_img(nh, mask)),
                                           C.byref(_img(nh, out)), conn, None)
    assert st == nh._lib.NSX_OK, ctx.lib.nsx_last_error(ctx.h)
    torch.cuda.synchronize()
    return out


def _reference_call(nh, ctx, marker, mask, conn):
    import torch
    out = torch.full_like(mask, 9)
    st = ctx.lib.nsx_imreconstruct(ctx.h, C.byref(_img(nh, marker)), C.byref(_img(nh, mask)),
                                   C.byref(_img(nh, out)), conn, None)
    assert st == nh._lib.NSX_OK, ctx.lib.nsx_last_error(ctx.h)
    torch.cuda.synchronize()
    return out


def _pair(nh, rows, cols, seed, dtype):
    import torch
    mask = nh.gen_disks_u8(rows, cols, max(1, round(200 * rows * cols / 512 ** 2)), seed)
    marker = nh.gen_marker_u8(mask, 30)
    if dtype == "i32":
        marker, mask = nh.scale_i32(marker, 257) - 20000, nh.scale_i32(mask, 257) - 20000
    elif dtype == "f32":
        marker, mask = nh.scale_f32(marker), nh.scale_f32(mask)
    return torch.from_numpy(marker).cuda(), torch.from_numpy(mask).cuda()


@pytest.mark.parametrize("dtype", ["u8", "i32", "f32"])
@pytest.mark.parametrize("conn", [4, 8])
def test_noalloc_matches_imreconstruct_two_calls(nh, gpu, dtype, conn):
    import torch
    ctx = nh.Context(0)
    try:
        for rows, cols, seed in [(1500, 1700, 0x5EED0501), (300, 260, 0x5EED0502)]:
            marker, mask = _pair(nh, rows, cols, seed, dtype)
            got = _noalloc(nh, ctx, marker, mask, conn)
            want = _reference_call(nh, ctx, marker, mask, conn)
            assert torch.equal(got, want), (rows, cols)
    finally:
        ctx.close()


def test_noalloc_u8_levels_and_workspace(nh, orc, gpu):
    """A two-valued u8 mask (which nsx_imreconstruct labels through its level
    sets) agrees too, against the oracle; the call grows no workspace."""
    import torch
    mask, marker = nh.gen_maze_u8(700, 900, 0x5EED0503)
    ctx = nh.Context(0)
    try:
        before = ctx.workspace_bytes()
        got = _noalloc(nh, ctx, torch.from_numpy(marker).cuda(), torch.from_numpy(mask).cuda(), 4)
        assert ctx.workspace_bytes() == before
        assert np.array_equal(got.cpu().numpy(), orc.imreconstruct(marker, mask, 4))
    finally:
        ctx.close()


def test_noalloc_limits(nh, gpu):
    import torch
    ctx = nh.Context(0)
    try:
        big = torch.zeros((8000, 8000), dtype=torch.uint8, device="cuda")   # 130 x 130 tiles > 16384
        out = torch.empty_like(big)
        st = ctx.lib.nsx_imreconstruct_noalloc(ctx.h, C.byref(_img(nh, big)), C.byref(_img(nh, big)),
                                               C.byref(_img(nh, out)), 8, None)
        assert st == nh._lib.NSX_ERR_INVALID and b"16384" in ctx.lib.nsx_last_error(ctx.h)
        host = np.zeros((8, 8), np.uint8)
        st = ctx.lib.nsx_imreconstruct_noalloc(ctx.h, C.byref(_img(nh, host)), C.byref(_img(nh, host)),
                                               C.byref(_img(nh, np.zeros((8, 8), np.uint8))), 8, None)
        assert st == nh._lib.NSX_ERR_INVALID
        # the largest accepted size: 7936 x 7936 = 128 x 128 tiles
        mask = torch.zeros((7936, 7936), dtype=torch.uint8, device="cuda")
        mask[::2, :] = 50
        marker = torch.zeros_like(mask)
        marker[0, 0] = 40
        got = _noalloc(nh, ctx, marker, mask, 8)
        want = _reference_call(nh, ctx, marker, mask, 8)
        assert torch.equal(got, want)
    finally:
        ctx.close()
