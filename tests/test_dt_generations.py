"""The generation decomposition behind the queue-order distanceTransform,
restated in plain Python on small masks and checked against the oracle's FIFO
(oracle/nscale_oracle.cpp, a restatement of MorphologicOperations.cpp:197-275):

* generation 0 is the initial front in raster order, generation g+1 what
  generation g pushed, in push order (entry by entry, each entry's neighbours
  in the order up-left, up-right, up, down-left, down-right, down, left,
  right);
* an entry's label when popped is the earliest smallest (in float distance)
  of its pixel's generation-start label and the labels earlier entries of its
  generation offered it; these form a triangular system, solved here by
  Jacobi passes until nothing changes;
* entry k pushes neighbour n when its label is strictly closer to n than n's
  generation-start label and every offer of an earlier entry at a neighbour
  of n; n then holds the offer of the last entry that pushed it.

Also the float comparison shortcut the kernels use: for integer squares
a < b with a <= 2^20, float(sqrt(a)) < float(sqrt(b)).  Runs on the CPU."""
import math

import numpy as np
import pytest

ORDER = [(-1, -1), (1, -1), (0, -1), (-1, 1), (1, 1), (0, 1), (-1, 0), (1, 0)]


def _fdist(x, y, lab, cols):
    lx, ly = lab % cols, lab // cols
    return np.float32(math.sqrt(float((x - lx) ** 2 + (y - ly) ** 2)))


def generations(mask):
    rows, cols = mask.shape
    n = rows * cols
    S = [3 * n] * n
    E = []
    for y in range(rows):
        for x in range(cols):
            i = y * cols + x
            if mask[y, x]:
                continue
            S[i] = i
            if any(0 <= y + dy < rows and 0 <= x + dx < cols and mask[y + dy, x + dx]
                   for dy in (-1, 0, 1) for dx in (-1, 0, 1) if dx or dy):
                E.append(i)
    gens = 0
    while E:
        gens += 1
        at = {}
        for k, p in enumerate(E):
            at.setdefault(p, []).append(k)

        def offers_to(p, k, c):
            """(j, label) offered to pixel p by entries j < k at its neighbours."""
            y, x = divmod(p, cols)
            out = []
            for dx, dy in ORDER:
                if 0 <= x + dx < cols and 0 <= y + dy < rows:
                    out += [(j, c[j]) for j in at.get((y + dy) * cols + x + dx, []) if j < k]
            return out

        c = [S[p] for p in E]
        while True:
            nc = []
            for k, p in enumerate(E):
                y, x = divmod(p, cols)
                best, bd, bj = S[p], _fdist(x, y, S[p], cols), -1
                for j, v in offers_to(p, k, c):
                    d = _fdist(x, y, v, cols)
                    if d < bd or (d == bd and bj >= 0 and j < bj):
                        best, bd, bj = v, d, j
                nc.append(best)
            if nc == c:
                break
            c = nc
        nxt, upd = [], {}
        for k, p in enumerate(E):
            y, x = divmod(p, cols)
            for dx, dy in ORDER:
                tx, ty = x + dx, y + dy
                if not (0 <= tx < cols and 0 <= ty < rows):
                    continue
                t = ty * cols + tx
                dk = _fdist(tx, ty, c[k], cols)
                if dk < _fdist(tx, ty, S[t], cols) and all(dk < _fdist(tx, ty, v, cols) for _, v in offers_to(t, k, c)):
                    nxt.append(t)
                    upd[t] = c[k]
        for t, v in upd.items():
            S[t] = v
        E = nxt
    return np.array(S, np.int32).reshape(rows, cols), gens


def _masks():
    rng = np.random.default_rng(17)
    out = [np.where(rng.random((h, w)) < p, 0, 255).astype(np.uint8)
           for (h, w), p in [((1, 1), 0.5), ((3, 9), 0.2), ((12, 10), 0.05), ((20, 23), 0.02), ((24, 31), 0.3),
                             ((17, 40), 0.01)]]
    ring = np.full((21, 21), 255, np.uint8)
    ring[0, :] = ring[-1, :] = ring[:, 0] = ring[:, -1] = 0
    out.append(ring)
    return out


@pytest.mark.parametrize("i", range(7))
def test_generations_equal_fifo(orc, i):
    m = _masks()[i]
    want, _ = orc.distance_transform(m)
    got, gens = generations(m)
    assert np.array_equal(got, want)


def test_float_sqrt_order_below_2_20():
    a = np.arange(0, (1 << 20) + 1, dtype=np.float64)
    fa = np.sqrt(a).astype(np.float32)
    assert (fa[1:] > fa[:-1]).all()
    # above the boundary every root rounds higher (roots are monotone)
    assert np.float32(np.sqrt(np.float64((1 << 20) + 1))) > fa[-1]
    # and the shortcut is not true above it: two squares there can tie
    b = np.arange(1 << 22, (1 << 22) + 4096, dtype=np.float64)
    fb = np.sqrt(b).astype(np.float32)
    assert (fb[1:] == fb[:-1]).any()
