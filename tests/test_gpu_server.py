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
