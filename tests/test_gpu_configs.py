"""The benchmark workloads at their stated sizes, through the C ABI, against
the oracle or the workload's closed form:

* the 4096x4096 u8 chain (3000 disks, seed 0x5EED0001): recon8 -> fill4 ->
  >128 -> bwareaopen8 [11, 1000) -> relabel, on a fresh context (the round
  prediction starts low, so the first call finishes its propagation by
  re-launching), after an easy input of the same size (prediction lowered),
  and once more with the learned prediction;
* the batch of 256 u8 + 32 f32 uneven tiles (sides 512..4096): batched ==
  per-tile on every tile, and the oracle on a sample holding the largest u8
  and f32 tiles;
* the 16384x16384 maze: recon = 200 on the corridor set, 20 elsewhere, and
  exactly one label equal to the corridor set;
* the serving trace (sides log-uniform 256..4096, 20 % int32) with Poisson
  arrivals replayed: each result identical alone and in the mix, and equal to
  the oracle's chain for sides up to 1024.
"""
import time

import numpy as np
import pytest

from chain_expect import expected_chain

pytestmark = pytest.mark.gpu


def _chain_dev(nh, ctx, dm, dk):
    import torch
    r = torch.empty_like(dk)
    f = torch.empty_like(dk)
    lab = torch.empty(dk.shape, dtype=torch.int32, device=dk.device)
    _, _, _, n = nh.segment_chain(dm, dk, 8, 4, 128.0, 8, 11, 1000, recon_out=r, fill_out=f, labels_out=lab, ctx=ctx)
    torch.cuda.synchronize()
    return r.cpu().numpy(), f.cpu().numpy(), lab.cpu().numpy(), n


@pytest.mark.parametrize("engine", ["rounds", "tree", "drain"])
def test_chain_4096_full(nh, orc, gpu, engine):
    """The default engine (propagation rounds), the component tree, and the
    rounds in one persistent drain launch per propagation."""
    import torch
    from nscale_hip.workloads import chain_images
    marker, mask = chain_images()
    er, ef, elab, en = orc.chain_u8(marker, mask, 8, 4, 128, 8, 11, 1000)
    assert en == 629
    dm, dk = torch.from_numpy(marker).cuda(), torch.from_numpy(mask).cuda()
    ctx = nh.Context(0)
    if engine == "tree":
        ctx.set_tuning(nh._lib.NSX_TUNE_ENGINE, nh._lib.NSX_ENGINE_TREE)
    if engine == "drain":
        ctx.set_tuning(nh._lib.NSX_TUNE_DRAIN, 1)
    try:
        # fresh context: with rounds, the default prediction (4 rounds) is
        # short of the 10 reconstruction and 36 fill rounds, so the call
        # re-launches both
        r, f, lab, n = _chain_dev(nh, ctx, dm, dk)
        assert np.array_equal(r, er) and np.array_equal(f, ef)
        assert np.array_equal(lab, elab) and n == en
        if engine != "tree":
            assert ctx.recon_stats(1)["rounds"] > 4
        if engine == "drain":
            assert ctx.drain_failures() == 0
        for rep in range(2):
            r, f, lab, n = _chain_dev(nh, ctx, dm, dk)
            assert np.array_equal(r, er) and np.array_equal(f, ef), f"repetition {rep}"
            assert np.array_equal(lab, elab) and n == en, f"repetition {rep}"
    finally:
        ctx.close()


@pytest.mark.parametrize("engine", ["rounds", "tree", "drain"])
def test_chain_4096_after_easy_input(nh, orc, gpu, engine):
    """An easy input of the same size (marker == mask: nothing propagates)
    sets a low round prediction for the size; the full chain that follows
    needs more rounds than predicted and must still be exact (with rounds:
    the re-launch branch; the component tree has no prediction)."""
    import torch
    from nscale_hip.workloads import chain_images
    marker, mask = chain_images()
    er, ef, elab, en = orc.chain_u8(marker, mask, 8, 4, 128, 8, 11, 1000)
    easy = np.zeros_like(mask)
    ctx = nh.Context(0)
    if engine == "tree":
        ctx.set_tuning(nh._lib.NSX_TUNE_ENGINE, nh._lib.NSX_ENGINE_TREE)
    if engine == "drain":
        ctx.set_tuning(nh._lib.NSX_TUNE_DRAIN, 1)
    try:
        de = torch.from_numpy(easy).cuda()
        r, f, lab, n = _chain_dev(nh, ctx, de, de)
        assert not r.any() and not f.any() and n == 0
        r, f, lab, n = _chain_dev(nh, ctx, torch.from_numpy(marker).cuda(), torch.from_numpy(mask).cuda())
        assert np.array_equal(r, er) and np.array_equal(f, ef)
        assert np.array_equal(lab, elab) and n == en
    finally:
        ctx.close()


def test_batch_full(nh, orc, gpu):
    import torch
    from nscale_hip.workloads import batch_tiles
    tiles = batch_tiles(256, 32)
    dev = torch.device("cuda:0")
    dt = []
    for marker, mask, thr in tiles:
        dk = torch.from_numpy(mask).to(dev)
        dt.append((torch.from_numpy(marker).to(dev), dk, torch.empty_like(dk),
                   torch.empty(dk.shape, dtype=torch.int32, device=dev), thr))
    per_tile = nh.Context(0)
    per_tile.set_tuning(nh._lib.NSX_TUNE_LEVELSET_LEVELS, 0)
    b = nh.Batch(nh.Context(0))
    try:
        for dm, dk, r, lab, thr in dt:
            b.add(dm, dk, r, lab, 8, thr, 8)
        b.finalize(use_graph=True)
        for rep in range(2):
            b.run()
            counts = b.counts()
            for i, (dm, dk, r, lab, thr) in enumerate(dt):
                r2, lab2, n2 = nh.recon_label(dm, dk, 8, thr, 8, ctx=per_tile)
                if not (torch.equal(r, r2) and torch.equal(lab, lab2) and counts[i] == n2):
                    # name the side that disagrees with the oracle
                    rn = r.cpu().numpy()
                    bw = np.where(rn > (np.float32(thr) if rn.dtype == np.float32 else thr), 255, 0).astype(np.uint8)
                    wl, wn = orc.bwlabel2(bw, 8, relab=True)
                    a, c = lab.cpu().numpy(), lab2.cpu().numpy()
                    raise AssertionError(
                        f"replay {rep}: tile {i} ({tuple(dk.shape)}, {dk.dtype}) batched != per-tile: recon equal "
                        f"{torch.equal(r, r2)}; labels vs oracle: batched {int((a != wl).sum())} px, per-tile "
                        f"{int((c != wl).sum())} px (first batched {np.argwhere(a != wl)[:3].tolist()}, per-tile "
                        f"{np.argwhere(c != wl)[:3].tolist()}); counts {counts[i]} {n2} oracle {wn}")
        big_u8 = max(range(256), key=lambda i: tiles[i][1].size)
        big_f32 = max(range(256, 288), key=lambda i: tiles[i][1].size)
        for i in sorted({0, 1, 255, 256, 287, big_u8, big_f32}):
            marker, mask, thr = tiles[i]
            wr = orc.imreconstruct(marker, mask, 8)
            bw = np.where(wr > (np.float32(thr) if wr.dtype == np.float32 else thr), 255, 0).astype(np.uint8)
            wl, wn = orc.bwlabel2(bw, 8, relab=True)
            assert np.array_equal(dt[i][2].cpu().numpy(), wr), f"tile {i} recon vs oracle"
            assert np.array_equal(dt[i][3].cpu().numpy(), wl) and counts[i] == wn, f"tile {i} labels vs oracle"
    finally:
        b.close()
        per_tile.close()


def test_maze_16384_closed_form(nh, gpu):
    import torch
    from nscale_hip.workloads import MAZE_SEED
    n = 16384
    mask, marker = nh.gen_maze_u8(n, n, MAZE_SEED)
    dev = torch.device("cuda:0")
    dk, dm = torch.from_numpy(mask).to(dev), torch.from_numpy(marker).to(dev)
    del mask, marker
    r, lab, ncomp = nh.recon_label(dm, dk, 4, 100.0, 4)
    corridors = dk == 200
    assert ncomp == 1
    assert torch.equal(r, corridors.to(torch.uint8) * 180 + 20)
    assert torch.equal(lab, corridors.to(torch.int32))


def test_maze_levels_16384_closed_form(nh, gpu):
    """The many-valued maze at 16384^2 (corridors 150..220 plus noise; the
    level-set path declines it) under the default engine: rounds give way to
    the bit planes past 256 rounds.  Every corridor pixel is reached above 100
    and every wall stays below it, so the labels are exactly one component, the
    corridor set; the second call starts with the bit planes and agrees."""
    import torch
    from nscale_hip.workloads import maze_levels_images
    marker, mask = maze_levels_images(16384)
    dev = torch.device("cuda:0")
    dm, dk = torch.from_numpy(marker).to(dev), torch.from_numpy(mask).to(dev)
    del marker, mask
    corridors = (dk > 100).to(torch.int32)
    ctx = nh.Context(0)
    try:
        r, lab, ncomp = nh.recon_label(dm, dk, 4, 100.0, 4, ctx=ctx)
        assert ncomp == 1 and torch.equal(lab, corridors)
        assert bool(((r > 100) == (dk > 100)).all()) and bool((r <= dk).all())
        r2, lab2, n2 = nh.recon_label(dm, dk, 4, 100.0, 4, ctx=ctx)
        assert n2 == 1 and torch.equal(r2, r) and torch.equal(lab2, lab)
    finally:
        ctx.close()


def test_serve_trace_poisson(nh, orc, gpu):
    """40 requests of the serving trace; arrivals replayed with exponential
    gaps at about the sequential service rate, on a 4-slot server whose
    arena holds about three of the largest requests."""
    from nscale_hip.workloads import SERVE_SEED, request_images, serve_trace
    trace = serve_trace(40, SERVE_SEED)
    reqs = [request_images(side, kind, seed) for side, kind, seed in trace]
    alone = []
    srv = nh.Server(0, 1 << 30, slots=1)
    t0 = time.perf_counter()
    for marker, mask, thr in reqs:
        r, f = np.empty_like(mask), np.empty_like(mask)
        t = srv.submit(marker, mask, 8, 4, thr, 8, 11, 1000, recon_out=r, fill_out=f)
        lab, _, _, n, _ = srv.wait(t)
        alone.append((r, f, lab, n))
    seq_s = time.perf_counter() - t0
    srv.close()
    biggest = max(m.size * m.itemsize for _, m, _ in reqs)
    srv = nh.Server(0, int(biggest * 8 * 3) + (1 << 20), slots=4)
    rng = np.random.default_rng(SERVE_SEED)
    gaps = rng.exponential(seq_s / len(reqs), size=len(reqs))
    tickets, outs = [], []
    for (marker, mask, thr), gap in zip(reqs, gaps):
        time.sleep(float(gap))
        r, f = np.full_like(mask, 7), np.full_like(mask, 7)
        tickets.append(srv.submit(marker, mask, 8, 4, thr, 8, 11, 1000, recon_out=r, fill_out=f))
        outs.append((r, f))
    for i, t in enumerate(tickets):
        lab, r, f, n, _ = srv.wait(t)
        ar, af, alab, an = alone[i]
        side, kind, _ = trace[i]
        assert np.array_equal(r, ar) and np.array_equal(f, af), f"request {i} ({side}, {kind}) differs in the mix"
        assert np.array_equal(lab, alab) and n == an, f"request {i} ({side}, {kind}) labels differ in the mix"
        if side <= 1024:
            er, ef, elab, en = expected_chain(orc, reqs[i][0], reqs[i][1], reqs[i][2])
            assert np.array_equal(r, er) and np.array_equal(f, ef), f"request {i} vs oracle"
            assert np.array_equal(lab, elab) and n == en, f"request {i} labels vs oracle"
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["in_flight"] == 0
    srv.close()
