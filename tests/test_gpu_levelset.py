"""u8 reconstruction through the mask's level sets (few distinct mask values)
against the oracle and against the propagation-round path on the same input."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def few_level_case(rng, shape, levels, p_marker=0.02):
    mask = rng.choice(np.array(levels, np.uint8), size=shape).astype(np.uint8)
    marker = np.zeros(shape, np.uint8)
    sel = rng.random(shape) < p_marker
    marker[sel] = rng.integers(0, 256, int(sel.sum()))
    return marker, mask


def contexts(nh):
    ls = nh.Context(0)
    ls.set_tuning(1, 3)
    sweep = nh.Context(0)
    sweep.set_tuning(1, 0)
    return ls, sweep


@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("levels,shape", [((0, 200), (97, 131)), ((20, 200), (256, 300)),
                                          ((0, 90, 255), (200, 170)), ((10, 60, 61, 250), (64, 640)),
                                          ((3, 7, 100, 101, 180), (150, 150))])
def test_levelset_matches_oracle(nh, orc, gpu, conn, levels, shape):
    rng = np.random.default_rng(sum(levels) + shape[0] * conn)
    marker, mask = few_level_case(rng, shape, levels)
    want = orc.imreconstruct(marker, mask, conn)
    ls, sweep = contexts(nh)
    got_ls = nh.imreconstruct(marker, mask, conn, ctx=ls)
    got_sw = nh.imreconstruct(marker, mask, conn, ctx=sweep)
    assert np.array_equal(got_ls, want)
    assert np.array_equal(got_sw, want)
    ls.close()
    sweep.close()


def test_levelset_marker_above_mask_and_uniform(nh, orc, gpu):
    ls, _ = contexts(nh)
    rng = np.random.default_rng(7)
    mask = np.where(rng.random((90, 77)) < 0.55, 180, 40).astype(np.uint8)
    marker = rng.integers(0, 256, mask.shape).astype(np.uint8)   # often above the mask
    assert np.array_equal(nh.imreconstruct(marker, mask, 4, ctx=ls), orc.imreconstruct(marker, mask, 4))
    flat = np.full((33, 70), 77, np.uint8)                          # one level: floor only
    mk = np.zeros_like(flat)
    mk[5, 5] = 30
    assert np.array_equal(nh.imreconstruct(mk, flat, 8, ctx=ls), orc.imreconstruct(mk, flat, 8))
    ls.close()


@pytest.mark.parametrize("shape", [(1030, 1500), (611, 389)])
def test_maze_recon_label_closed_form(nh, orc, gpu, shape):
    """Maze: reconstruction from the entrance seed is 200 on the corridors
    (one perfect maze, all reachable), 20 elsewhere; > 100 then 4-connected
    labelling gives exactly one component."""
    import torch
    mask, marker = nh.gen_maze_u8(shape[0], shape[1], 0x5EED0003)
    corridors = mask == 200
    r, lab, n = nh.recon_label(torch.from_numpy(marker).cuda(), torch.from_numpy(mask).cuda(), 4, 100.0, 4)
    r = r.cpu().numpy()
    lab = lab.cpu().numpy()
    assert n == 1
    assert np.array_equal(r, np.where(corridors, 200, 20).astype(np.uint8))
    assert np.array_equal(lab, corridors.astype(np.int32))
    assert np.array_equal(r, orc.imreconstruct(marker, mask, 4))


def test_levelset_values_split_between_waves(nh, orc, gpu):
    """Each half of every row holds four values, the image seven: no single
    scan wave sees more than the limit, the union exceeds it, and the rounds
    path runs.  Both contexts give the oracle's result."""
    rng = np.random.default_rng(23)
    shape = (64, 4096)
    mask = np.empty(shape, np.uint8)
    mask[:, :2048] = rng.choice(np.array([0, 10, 20, 30], np.uint8), size=(64, 2048))
    mask[:, 2048:] = rng.choice(np.array([0, 40, 50, 60], np.uint8), size=(64, 2048))
    marker = np.where(rng.random(shape) < 0.02, rng.integers(0, 256, shape), 0).astype(np.uint8)
    ls, sweep = contexts(nh)
    want = orc.imreconstruct(marker, mask, 8)
    assert np.array_equal(nh.imreconstruct(marker, mask, 8, ctx=ls), want)
    assert np.array_equal(nh.imreconstruct(marker, mask, 8, ctx=sweep), want)
    ls.close()
    sweep.close()


@pytest.mark.parametrize("x0,cols", [(3, 517), (16, 1024), (7, 15), (1, 1)])
def test_levelset_unaligned_rows(nh, orc, gpu, x0, cols):
    """Device views whose rows start and end off 16-byte boundaries (column
    slices of wider, odd-pitched parents)."""
    import torch
    rng = np.random.default_rng(x0 * 1000 + cols)
    rows = 211
    marker, mask = few_level_case(rng, (rows, cols), (5, 90, 200))
    pitch_m, pitch_k = x0 + cols + 13, x0 + cols + 29
    big_m = torch.full((rows, pitch_m), 77, dtype=torch.uint8, device="cuda")   # 77: a value outside the image
    big_k = torch.full((rows, pitch_k), 250, dtype=torch.uint8, device="cuda")  # above every marker value
    big_m[:, x0:x0 + cols] = torch.from_numpy(mask).cuda()
    big_k[:, x0:x0 + cols] = torch.from_numpy(marker).cuda()
    ls, _ = contexts(nh)
    for conn in (4, 8):
        got = nh.imreconstruct(big_k[:, x0:x0 + cols], big_m[:, x0:x0 + cols], conn, ctx=ls).cpu().numpy()
        assert np.array_equal(got, orc.imreconstruct(marker, mask, conn))
    ls.close()


def test_levelset_many_levels_setting(nh, orc, gpu):
    """A level limit above the per-wave list (16 values) takes the bitmap scan."""
    rng = np.random.default_rng(31)
    levels = tuple(range(0, 200, 10))   # 20 values
    marker, mask = few_level_case(rng, (150, 333), levels)
    ctx = nh.Context(0)
    ctx.set_tuning(1, 40)
    for conn in (4, 8):
        assert np.array_equal(nh.imreconstruct(marker, mask, conn, ctx=ctx), orc.imreconstruct(marker, mask, conn))
    ctx.close()


@pytest.mark.parametrize("rconn,lconn", [(4, 4), (8, 8), (4, 8), (8, 4)])
def test_recon_label_levels(nh, orc, gpu, rconn, lconn):
    """nsx_recon_label on few-level masks: with equal connectivities the labels
    come from the level sets' own labelling (threshold below every level, between
    levels, at and above the largest; largest marker value below the threshold),
    otherwise from a separate labelling.  Both must equal the oracle's
    bwlabel2(recon > thresh)."""
    import torch
    rng = np.random.default_rng(41 + 10 * rconn + lconn)
    cases = []
    for levels, shape in [((5, 90, 200), (211, 517)), ((20, 200), (300, 300)), ((0, 60, 61, 250), (97, 640))]:
        marker, mask = few_level_case(rng, shape, levels)
        for t in (-3.0, levels[0] - 1.0, levels[0] + 0.5, float(levels[1]), levels[-1] - 0.5, float(levels[-1]), 300.0):
            cases.append((marker, mask, t))
    marker, mask = few_level_case(rng, (150, 170), (10, 100))
    cases.append((np.minimum(marker, 3).astype(np.uint8), mask, 5.0))   # every marker value below the threshold
    ls, sweep = contexts(nh)
    for marker, mask, t in cases:
        r = orc.imreconstruct(marker, mask, rconn)
        want, wn = orc.bwlabel2(np.where(r > t, 255, 0).astype(np.uint8), lconn, True)
        for c in (ls, sweep):
            gr, gl, n = nh.recon_label(torch.from_numpy(marker).cuda(), torch.from_numpy(mask).cuda(), rconn, t, lconn,
                                       ctx=c)
            assert np.array_equal(gr.cpu().numpy(), r)
            assert np.array_equal(gl.cpu().numpy(), want), (mask.shape, t)
            assert n == wn
    ls.close()
    sweep.close()
