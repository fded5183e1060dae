"""Serving loop: every request's result is the same alone and in the mix, and
equals the oracle's chain."""
import numpy as np
import pytest

from chain_expect import expected_chain, request_images

pytestmark = pytest.mark.gpu


def make_requests(nh, n, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        side = int(round(np.exp(rng.uniform(np.log(48), np.log(600)))))
        kind = "i32" if rng.random() < 0.3 else "u8"
        reqs.append(request_images(nh, side, kind, 0x5EED0004 + seed * 1000 + i))
    return reqs


def run_alone(nh, reqs):
    srv = nh.Server(0, 256 << 20, slots=1)
    out = []
    for marker, mask, thr in reqs:
        r = np.empty_like(mask)
        f = np.empty_like(mask)
        t = srv.submit(marker, mask, 8, 4, thr, 8, 11, 1000, recon_out=r, fill_out=f)
        lab, _, _, n, _ = srv.wait(t)
        out.append((r, f, lab, n))
    srv.close()
    return out


def test_server_alone_equals_mix_and_oracle(nh, orc, gpu):
    reqs = make_requests(nh, 14, 1)
    alone = run_alone(nh, reqs)
    # a budget of ~2.5 of the largest requests forces admission to wait on memory
    biggest = max(m.size * m.itemsize for _, m, _ in reqs)
    srv = nh.Server(0, int(biggest * 8 * 2.5) + (1 << 20), slots=4)
    tickets, outs = [], []
    for marker, mask, thr in reqs:
        r = np.full_like(mask, 7)
        f = np.full_like(mask, 7)
        tickets.append(srv.submit(marker, mask, 8, 4, thr, 8, 11, 1000, recon_out=r, fill_out=f))
        outs.append((r, f))
    for i, t in enumerate(tickets):
        lab, r, f, n, times = srv.wait(t)
        ar, af, alab, an = alone[i]
        assert np.array_equal(r, ar) and np.array_equal(f, af), f"request {i} recon/fill differ in the mix"
        assert np.array_equal(lab, alab) and n == an, f"request {i} labels differ in the mix"
        er, ef, elab, en = expected_chain(orc, reqs[i][0], reqs[i][1], reqs[i][2])
        assert np.array_equal(r, er) and np.array_equal(f, ef), f"request {i} vs oracle"
        assert np.array_equal(lab, elab) and n == en, f"request {i} labels vs oracle"
        assert times[2] >= times[1] >= 0.0
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["in_flight"] == 0 and st["queued"] == 0
    assert st["arena_peak_bytes"] <= st["arena_bytes"]
    srv.close()


def test_server_device_buffers(nh, orc, gpu):
    import torch
    marker, mask, thr = request_images(nh, 300, "u8", 0x5EED0104)
    srv = nh.Server(0, 64 << 20, slots=2)
    dm, dk = torch.from_numpy(marker).cuda(), torch.from_numpy(mask).cuda()
    lab = torch.empty(mask.shape, dtype=torch.int32, device="cuda")
    rec = torch.empty_like(dk)
    torch.cuda.synchronize()
    t = srv.submit(dm, dk, 8, 4, thr, 8, 11, 1000, labels_out=lab, recon_out=rec)
    _, _, _, n, _ = srv.wait(t)
    er, _, elab, en = expected_chain(orc, marker, mask, thr)
    assert np.array_equal(rec.cpu().numpy(), er)
    assert np.array_equal(lab.cpu().numpy(), elab) and n == en
    srv.close()


def test_server_rejects_over_budget_and_bad_input(nh, gpu):
    srv = nh.Server(0, 1 << 20, slots=1)
    big = np.zeros((1024, 1024), np.uint8)
    with pytest.raises(nh.NsxError) as e:
        srv.submit(big, big, 8, 4, 128.0)
    assert e.value.status == 5   # NSX_ERR_BUDGET
    f = np.zeros((16, 16), np.float32)
    with pytest.raises(nh.NsxError):
        srv.submit(f, f, 8, 4, 0.5)
    srv.close()


def _shared_mix(nh):
    """A long request first (it keeps the one slot busy while the rest queue),
    then small requests of several kinds: u8 and int32, two thresholds, 4- and
    8-connected reconstruction, 1-pixel and 1-row images, and a many-valued maze
    whose reconstruction passes the shared set's round limit."""
    from nscale_hip.workloads import maze_levels_images
    rng = np.random.default_rng(77)
    reqs = []
    m, k, t = request_images(nh, 2048, "u8", 0x5EED0701)
    reqs.append((m, k, t, 8))
    for i in range(22):
        side = int(round(np.exp(rng.uniform(np.log(40), np.log(500)))))
        kind = "i32" if i % 4 == 3 else "u8"
        m, k, t = request_images(nh, side, kind, 0x5EED0710 + i)
        if i % 5 == 1:
            t = t + 17
        reqs.append((m, k, t, 4 if i % 3 == 2 else 8))
    for shape in ((1, 1), (1, 300), (300, 1)):
        m, k, t = request_images(nh, 300, "u8", 0x5EED0799)
        reqs.append((np.ascontiguousarray(m[:shape[0], :shape[1]]), np.ascontiguousarray(k[:shape[0], :shape[1]]), t, 8))
    mm, mk = maze_levels_images(512)
    reqs.insert(5, (mm, mk, 128.0, 4))
    return reqs


def _serve(nh, reqs, slots, max_batch):
    srv = nh.Server(0, 1 << 30, slots=slots, max_batch=max_batch)
    tickets, outs = [], []
    for marker, mask, thr, rc in reqs:
        r, f = np.full_like(mask, 7), np.full_like(mask, 7)
        tickets.append(srv.submit(marker, mask, rc, 4, thr, 8, 11, 1000, recon_out=r, fill_out=f))
        outs.append((r, f))
    res = []
    for (r, f), t in zip(outs, tickets):
        lab, _, _, n, _ = srv.wait(t)
        res.append((r, f, lab, n))
    st = srv.stats()
    srv.close()
    return res, st


@pytest.mark.parametrize("slots", [1, 2])
def test_server_shared_launches_equal_alone_and_oracle(nh, orc, gpu, slots):
    """Requests that queue behind a busy slot are served together (shared
    reconstruction, fill and labelling launches per kind); every result equals
    the request served alone and the oracle's chain."""
    reqs = _shared_mix(nh)
    alone = []
    srv = nh.Server(0, 1 << 30, slots=1, max_batch=1)
    for marker, mask, thr, rc in reqs:
        r, f = np.empty_like(mask), np.empty_like(mask)
        lab, _, _, n, _ = srv.wait(srv.submit(marker, mask, rc, 4, thr, 8, 11, 1000, recon_out=r, fill_out=f))
        alone.append((r, f, lab, n))
    assert srv.stats()["shared_sets"] == 0
    srv.close()
    for rep in range(2):   # the second pass starts from the first one's round predictions
        mixed, st = _serve(nh, reqs, slots, 64)
        print(f"slots={slots} pass={rep} stats={st}", flush=True)
        assert st["shared_sets"] >= 1 and st["shared_requests"] >= 2 and st["completed"] == len(reqs)
        for i, (got, want) in enumerate(zip(mixed, alone)):
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), f"request {i} recon/fill"
            assert np.array_equal(got[2], want[2]) and got[3] == want[3], f"request {i} labels"
    for i, (marker, mask, thr, rc) in enumerate(reqs):
        er, ef, elab, en = expected_chain(orc, marker, mask, thr, recon_conn=rc)
        r, f, lab, n = alone[i]
        assert np.array_equal(r, er) and np.array_equal(f, ef), f"request {i} recon/fill vs oracle"
        assert np.array_equal(lab, elab) and n == en, f"request {i} labels vs oracle"


def test_server_sharing_bounds(nh, gpu):
    srv = nh.Server(0, 1 << 20, slots=1)
    st = srv.stats()
    assert (st["max_batch"], st["share_max_request_px"], st["share_max_set_px"]) == (16, 1 << 20, 1 << 22)
    for bad in ((0, 0, 0), (65, 0, 0), (4, -1, 0), (4, 0, -1)):
        with pytest.raises(nh.NsxError):
            srv.set_sharing(*bad)
    srv.set_sharing(64, 1 << 24, 1 << 26)
    st = srv.stats()
    assert (st["max_batch"], st["share_max_request_px"], st["share_max_set_px"]) == (64, 1 << 24, 1 << 26)
    srv.set_sharing(8)
    assert srv.stats()["share_max_request_px"] == 1 << 20
    srv.close()


def test_slot_timing_and_tile_counts(nh, gpu):
    """Kernel timing and tile counts through the slot contexts: one 300x300 u8
    request on its own (its round count already learned) counts 25 first-pass reconstruction tiles (5 x 5 of 62 x
    62) and 16 first-pass fill tiles (the border ring), and some round visits;
    switching timing clears the counts."""
    marker, mask, thr = request_images(nh, 300, "u8", 0x5EED0104)
    srv = nh.Server(0, 64 << 20, slots=1, max_batch=1)
    c = srv.context(0)
    with pytest.raises(nh.NsxError):
        srv.context(1)
    # the first call learns the round count (a short prediction is finished
    # and the stages after it run again); the timed one launches enough rounds
    srv.wait(srv.submit(marker, mask, 8, 4, thr, 8, 11, 1000))
    c.set_timing(True)
    srv.wait(srv.submit(marker, mask, 8, 4, thr, 8, 11, 1000))
    kt = c.read_timing()
    counts = c.read_tile_counts()
    assert kt["recon_first"][1] == 2 and kt["recon_round"][1] >= 2 and kt["ccl_local"][1] >= 1
    assert counts[("recon", "u8")][0] == 25 and counts[("fill", "u8")][0] == 16
    assert counts[("recon", "u8")][1] > 0
    assert set(counts) == {("recon", "u8"), ("fill", "u8")}
    c.set_timing(False)
    assert c.read_tile_counts() == {}
    srv.close()
