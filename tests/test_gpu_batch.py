"""Batched tiles (one set of launches, hipGraph) and the per-tile path vs the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F_THR = float(np.float32(128) / np.float32(255))


def make_tiles(nh, n, seed, dtypes=("u8", "f32", "i32")):
    rng = np.random.default_rng(seed)
    tiles = []
    for k in range(n):
        h = int(rng.integers(4, 44)) * 16
        w = int(rng.integers(4, 44)) * 16
        mask = nh.gen_disks_u8(h, w, max(1, round(200 * h * w / 512 ** 2)), 0x5EED0002 + seed * 1000 + k)
        marker = nh.gen_marker_u8(mask, 30)
        kind = dtypes[k % len(dtypes)]
        if kind == "f32":
            tiles.append((nh.scale_f32(marker), nh.scale_f32(mask), F_THR))
        elif kind == "i32":
            tiles.append((nh.scale_i32(marker), nh.scale_i32(mask), 128.0 * 257))
        else:
            tiles.append((marker, mask, 128.0))
    return tiles


def expected(orc, marker, mask, thr):
    r = orc.imreconstruct(marker, mask, 8)
    b = np.where(r > (np.float32(thr) if r.dtype == np.float32 else thr), 255, 0).astype(np.uint8)
    lab, n = orc.bwlabel2(b, 8, relab=True)
    return r, lab, n


@pytest.mark.parametrize("shape", ["drain", "serial", "overlap"])
@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_matches_oracle_every_replay(nh, orc, gpu, use_graph, shape):
    """Three tile groups (u8, f32, i32) under each launch shape: drained,
    per-round launches in turn, and per-round with the odd group on a second
    stream (the default)."""
    import torch
    tiles = make_tiles(nh, 9, 1)
    want = [expected(orc, m, k, t) for m, k, t in tiles]
    batch = nh.Batch()
    batch.set_drain(shape == "drain")
    batch.set_overlap(shape == "overlap")
    outs = []
    for marker, mask, thr in tiles:
        dm = torch.from_numpy(marker).cuda()
        dk = torch.from_numpy(mask).cuda()
        r = torch.empty_like(dk)
        lab = torch.empty(dk.shape, dtype=torch.int32, device="cuda")
        batch.add(dm, dk, r, lab, 8, thr, 8)
        outs.append((r, lab))
    batch.finalize(use_graph=use_graph)
    for replay in range(3):
        for r, lab in outs:          # poison the outputs between replays
            r.fill_(7)
            lab.fill_(-5)
        junk = torch.randn(256, 256, device="cuda") @ torch.randn(256, 256, device="cuda")
        torch.cuda.synchronize()
        if shape == "overlap" and not use_graph:
            # an eager run forks only from a stream object (not torch's default stream)
            side = torch.cuda.Stream()
            with torch.cuda.stream(side):
                batch.run()
            side.synchronize()
        else:
            batch.run()
        counts = batch.counts()
        for i, ((r, lab), (wr, wlab, wn)) in enumerate(zip(outs, want)):
            assert np.array_equal(r.cpu().numpy(), wr), f"replay {replay} tile {i} recon"
            assert np.array_equal(lab.cpu().numpy(), wlab), f"replay {replay} tile {i} labels"
            assert counts[i] == wn, f"replay {replay} tile {i} count"
        del junk
    batch.close()


def test_recon_label_per_tile(nh, orc, gpu):
    import torch
    for marker, mask, thr in make_tiles(nh, 6, 2):
        r, lab, n = nh.recon_label(torch.from_numpy(marker).cuda(), torch.from_numpy(mask).cuda(), 8, thr, 8)
        wr, wlab, wn = expected(orc, marker, mask, thr)
        assert np.array_equal(r.cpu().numpy(), wr)
        assert np.array_equal(lab.cpu().numpy(), wlab)
        assert n == wn


def test_batch_rejects_host_images(nh, gpu):
    b = nh.Batch()
    a = np.zeros((16, 16), np.uint8)
    with pytest.raises(nh.NsxError):
        b.add(a, a, a, np.zeros((16, 16), np.int32))
    b.close()


def test_batch_autotune_keeps_results(nh, orc, gpu):
    """nsx_batch_autotune re-captures the graph per tried setting and leaves the
    fastest one captured: the choices are in 3..6 per group, and replays after
    it still give the oracle's results.  A batch without a graph is refused."""
    import torch
    tiles = make_tiles(nh, 6, 3)
    want = [expected(orc, m, k, t) for m, k, t in tiles]
    batch = nh.Batch()
    outs = []
    for marker, mask, thr in tiles:
        dm = torch.from_numpy(marker).cuda()
        dk = torch.from_numpy(mask).cuda()
        r = torch.empty_like(dk)
        lab = torch.empty(dk.shape, dtype=torch.int32, device="cuda")
        batch.add(dm, dk, r, lab, 8, thr, 8)
        outs.append((r, lab))
    batch.finalize(use_graph=True)
    tuned = batch.autotune(reps=1)
    assert 1 <= len(tuned) <= 3 and tuned == batch.tuning()
    for g in tuned:
        assert 3 <= g["visit_pairs"] <= 6 and g["drain"] in (0, 1) and g["xcd_split"] in (0, 1)
        assert g["drain_blocks_per_cu"] in (0, 8, 4) and g["overlap"] in (0, 1)
    for replay in range(2):
        for r, lab in outs:
            r.fill_(7)
            lab.fill_(-5)
        torch.cuda.synchronize()
        batch.run()
        counts = batch.counts()
        for i, ((r, lab), (wr, wlab, wn)) in enumerate(zip(outs, want)):
            assert np.array_equal(r.cpu().numpy(), wr), f"replay {replay} tile {i} recon"
            assert np.array_equal(lab.cpu().numpy(), wlab), f"replay {replay} tile {i} labels"
            assert counts[i] == wn
    batch.close()
    eager = nh.Batch()
    marker, mask, thr = tiles[0]
    dk = torch.from_numpy(mask).cuda()
    eager.add(torch.from_numpy(marker).cuda(), dk, torch.empty_like(dk),
              torch.empty(dk.shape, dtype=torch.int32, device="cuda"), 8, thr, 8)
    eager.finalize(use_graph=False)
    with pytest.raises(nh.NsxError):
        eager.autotune(reps=1)
    eager.close()


@pytest.mark.parametrize("drain", [True, False])
@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_contents_changed_after_finalize(nh, orc, gpu, use_graph, drain):
    """Tile contents rewritten in place after finalize so that a replay needs
    more propagation rounds than were captured (a maze where disks were).
    Drained: the run itself is final (checked before nsx_batch_counts).  Per
    round: nsx_batch_counts finishes the short group.  Either way recon, labels
    and counts match the oracle, on that replay and on the next one."""
    import torch
    tiles = make_tiles(nh, 4, 4, dtypes=("u8",))
    batch = nh.Batch()
    batch.set_drain(drain)
    ins, outs = [], []
    for marker, mask, thr in tiles:
        dm = torch.from_numpy(marker).cuda()
        dk = torch.from_numpy(mask).cuda()
        r = torch.empty_like(dk)
        lab = torch.empty(dk.shape, dtype=torch.int32, device="cuda")
        batch.add(dm, dk, r, lab, 8, 128.0, 8)
        ins.append((dm, dk))
        outs.append((r, lab))
    batch.finalize(use_graph=use_graph)
    batch.counts()
    new = []
    for i, (dm, dk) in enumerate(ins):
        if i % 2 == 0:
            mask, marker = nh.gen_maze_u8(dk.shape[0], dk.shape[1], 0x5EED0302 + i)
        else:
            marker, mask = tiles[i][0], tiles[i][1]
        dm.copy_(torch.from_numpy(marker))
        dk.copy_(torch.from_numpy(mask))
        new.append(expected(orc, marker, mask, 128.0))
    torch.cuda.synchronize()
    for replay in range(2):
        for r, lab in outs:
            r.fill_(7)
            lab.fill_(-5)
        torch.cuda.synchronize()
        batch.run()
        if drain:
            torch.cuda.synchronize()
            for i, ((r, lab), (wr, wlab, wn)) in enumerate(zip(outs, new)):
                assert np.array_equal(r.cpu().numpy(), wr), f"replay {replay} tile {i} recon before counts"
                assert np.array_equal(lab.cpu().numpy(), wlab), f"replay {replay} tile {i} labels before counts"
        counts = batch.counts()
        for i, ((r, lab), (wr, wlab, wn)) in enumerate(zip(outs, new)):
            assert np.array_equal(r.cpu().numpy(), wr), f"replay {replay} tile {i} recon"
            assert np.array_equal(lab.cpu().numpy(), wlab), f"replay {replay} tile {i} labels"
            assert counts[i] == wn, f"replay {replay} tile {i} count"
    batch.close()
