"""Independent restatements used to pin the oracle (test helpers).

Each one computes the same mathematical object as the reference by a
different, textbook method, so that a mistake in the oracle's line-by-line
restatement (a dropped neighbour, a wrong frame, an off-by-one in a scan)
shows up as a mismatch.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage


def _dilate(x: np.ndarray, conn: int) -> np.ndarray:
    """Max over the conn-neighbourhood (self included) of a padded array; the
    outermost ring of the result is not used."""
    d = x.copy()
    d[1:, :] = np.maximum(d[1:, :], x[:-1, :])
    d[:-1, :] = np.maximum(d[:-1, :], x[1:, :])
    d[:, 1:] = np.maximum(d[:, 1:], x[:, :-1])
    d[:, :-1] = np.maximum(d[:, :-1], x[:, 1:])
    if conn == 8:
        d[1:, 1:] = np.maximum(d[1:, 1:], x[:-1, :-1])
        d[1:, :-1] = np.maximum(d[1:, :-1], x[:-1, 1:])
        d[:-1, 1:] = np.maximum(d[:-1, 1:], x[1:, :-1])
        d[:-1, :-1] = np.maximum(d[:-1, :-1], x[1:, 1:])
    return d


def recon_fixpoint(marker: np.ndarray, mask: np.ndarray, conn: int, frame=0) -> np.ndarray:
    """Geodesic dilation iterated to its fixed point with a constant frame:
    X0 = min(marker, mask); X <- min(dilate(X), mask) until stable."""
    dt = mask.dtype
    mp = np.pad(mask, 1, constant_values=frame).astype(dt)
    xp = np.pad(np.minimum(marker, mask), 1, constant_values=frame).astype(dt)
    while True:
        nx = np.minimum(_dilate(xp, conn), mp)
        nx[0, :] = frame
        nx[-1, :] = frame
        nx[:, 0] = frame
        nx[:, -1] = frame
        if np.array_equal(nx, xp):
            return xp[1:-1, 1:-1]
        xp = nx


def structure(conn: int) -> np.ndarray:
    return np.ones((3, 3), bool) if conn == 8 else ndimage.generate_binary_structure(2, 1)


def recon_binary_bfs(marker: np.ndarray, mask: np.ndarray, conn: int) -> np.ndarray:
    """Binary reconstruction by flood fill: components of mask != 0 that hold a
    pixel with marker != 0 and mask != 0 become max(T); nonzero marker pixels
    keep their value."""
    s = mask != 0
    lab, _ = ndimage.label(s, structure=structure(conn))
    seeds = np.unique(lab[(marker != 0) & s])
    seeds = seeds[seeds > 0]
    reached = np.isin(lab, seeds) & s
    tmax = np.iinfo(mask.dtype).max
    out = np.where(reached, tmax, 0).astype(mask.dtype)
    return np.where(marker != 0, marker, out).astype(mask.dtype)


def fill_holes_binary_scipy(bw: np.ndarray, conn: int) -> np.ndarray:
    """MATLAB imfill(BW, 'holes') with background connectivity conn."""
    filled = ndimage.binary_fill_holes(bw != 0, structure=structure(conn))
    return np.where(filled, 255, 0).astype(np.uint8)


def fill_holes_gray_u8(img: np.ndarray, conn: int) -> np.ndarray:
    """255 - R(marker = 255 on a frame ring, mask = 255 - img) by fixed point."""
    inv = (255 - img.astype(np.int32)).astype(np.uint8)
    mk = np.pad(inv, 1, constant_values=255)
    mr = np.zeros_like(mk)
    mr[0, :] = mr[-1, :] = mr[:, 0] = mr[:, -1] = 255
    r = recon_fixpoint(mr, mk, conn, frame=0)
    return (255 - r[1:-1, 1:-1].astype(np.int32)).astype(np.uint8)


def canonical_labels(img: np.ndarray, conn: int) -> np.ndarray:
    """Union of equal-valued nonzero neighbours, numbered 1.. by the raster
    order of each component's first pixel (scipy labels per value, then
    renumbered)."""
    out = np.zeros(img.shape, np.int64)
    nxt = 0
    for v in np.unique(img):
        if v == 0:
            continue
        lab, n = ndimage.label(img == v, structure=structure(conn))
        out[lab > 0] = lab[lab > 0] + nxt
        nxt += n
    flat = out.ravel()
    order = {}
    res = np.zeros(flat.shape, np.int32)
    k = 0
    for i, l in enumerate(flat):
        if l == 0:
            continue
        if l not in order:
            k += 1
            order[l] = k
        res[i] = order[l]
    return res.reshape(img.shape)


def root_labels_from_canonical(canon: np.ndarray) -> np.ndarray:
    """Root-index labels (raster index of each component's first pixel, -1 for
    background) from canonical labels."""
    flat = canon.ravel()
    first = {}
    for i, l in enumerate(flat):
        if l and l not in first:
            first[l] = i
    out = np.full(flat.shape, -1, np.int32)
    for i, l in enumerate(flat):
        if l:
            out[i] = first[l]
    return out.reshape(canon.shape)


def bwselect_scipy(img: np.ndarray, seeds: np.ndarray, conn: int) -> np.ndarray:
    """The components of img != 0 that hold a pixel with seeds != 0, with the
    image's values; 0 elsewhere."""
    s = img != 0
    lab, _ = ndimage.label(s, structure=structure(conn))
    sel = np.unique(lab[(seeds != 0) & s])
    sel = sel[sel > 0]
    return np.where(np.isin(lab, sel) & s, img, 0).astype(img.dtype)


def imfill_binary_flood(img: np.ndarray, seeds: np.ndarray, conn: int) -> np.ndarray:
    """Seeded binary fill of a u8 image: flood from the seeds through the pixels
    below 255 (the complement's nonzero pixels); reached pixels become 255."""
    s = img != 255
    lab, _ = ndimage.label(s, structure=structure(conn))
    sel = np.unique(lab[(seeds != 0) & s])
    sel = sel[sel > 0]
    return np.where(np.isin(lab, sel) & s, 255, img).astype(np.uint8)


def imfill_gray_fixpoint(img: np.ndarray, seeds: np.ndarray, conn: int) -> np.ndarray:
    """Seeded grayscale fill of a u8 image: img | R, R the geodesic
    reconstruction of the complement from the complement at the seeds."""
    comp = (255 - img.astype(np.int32)).astype(np.uint8)
    marker = np.where(seeds != 0, comp, 0).astype(np.uint8)
    return (img | recon_fixpoint(marker, comp, conn)).astype(np.uint8)


def imhmin_fixpoint(img: np.ndarray, h, conn: int) -> np.ndarray:
    """imhmin from the textbook form: complement, subtract h (saturating for
    u8), geodesic reconstruction with a 0 frame, complement back."""
    if img.dtype == np.uint8:
        comp = (255 - img.astype(np.int32)).astype(np.uint8)
        marker = np.maximum(comp.astype(np.int32) - int(h), 0).astype(np.uint8)
        return (255 - recon_fixpoint(marker, comp, conn).astype(np.int32)).astype(np.uint8)
    comp = (np.float32(0) - img).astype(np.float32)
    marker = (comp - np.float32(h)).astype(np.float32)
    return (np.float32(0) - recon_fixpoint(marker, comp, conn, frame=np.float32(0))).astype(np.float32)


def local_maxima_closed_form(img: np.ndarray, conn: int) -> np.ndarray:
    """Candidates: the reconstruction of image - 1 (u8 saturating) under the
    image stays at the marker.  The result drops every candidate plateau (a
    conn-connected set of equal-valued candidates) that holds a pixel with a
    strictly higher neighbour; pixels outside the image never count as higher."""
    if img.dtype == np.uint8:
        marker = np.maximum(img.astype(np.int32) - 1, 0).astype(np.uint8)
    else:
        marker = (img - np.float32(1)).astype(np.float32)
    frame = np.float32(0) if img.dtype == np.float32 else 0
    cand = ~(marker < recon_fixpoint(marker, img, conn, frame=frame))
    lo = np.iinfo(np.uint8).min if img.dtype == np.uint8 else -np.inf
    pad = np.pad(img.astype(np.float64), 1, constant_values=lo)
    higher = np.zeros(img.shape, bool)
    offs = [(-1, 0), (1, 0), (0, -1), (0, 1)] + ([(-1, -1), (-1, 1), (1, -1), (1, 1)] if conn == 8 else [])
    H, W = img.shape
    for dy, dx in offs:
        higher |= pad[1 + dy:1 + dy + H, 1 + dx:1 + dx + W] > img
    out = cand.copy()
    for v in np.unique(img[cand]):
        lab, _ = ndimage.label(cand & (img == v), structure=structure(conn))
        bad = np.unique(lab[higher & (lab > 0)])
        out &= ~np.isin(lab, bad[bad > 0])
    return np.where(out, 255, 0).astype(np.uint8)


def exact_distance_with_far(mask):
    """Exact Euclidean distance of every pixel to the nearest zero pixel of a
    u8 mask (scipy's separable EDT), or to the point (0, 3*rows) when that is
    nearer, as float32(sqrt(double(d^2))).  Also returns the squared distances."""
    from scipy import ndimage
    r, c = mask.shape
    yy, xx = np.mgrid[0:r, 0:c]
    far = (xx.astype(np.int64) ** 2 + (3 * r - yy.astype(np.int64)) ** 2)
    if (mask == 0).any():
        d2 = np.rint(ndimage.distance_transform_edt(mask != 0) ** 2).astype(np.int64)
        d2 = np.minimum(d2, far)
    else:
        d2 = far
    return np.sqrt(d2.astype(np.float64)).astype(np.float32), d2


def label_distance(nn, shape):
    """float32 distance from every pixel to the point its raster-index label
    names (rows*cols*3 names (0, 3*rows))."""
    r, c = shape
    yy, xx = np.mgrid[0:r, 0:c]
    lx = (nn.astype(np.int64) % c)
    ly = (nn.astype(np.int64) // c)
    d2 = (xx - lx) ** 2 + (yy - ly) ** 2
    return np.sqrt(d2.astype(np.float64)).astype(np.float32)


def distance_fixed_point_violations(nn, shape):
    """Pixels p with an 8-neighbour q whose label is strictly closer to p than
    p's own, comparing float32-rounded distances (the propagation's rule)."""
    r, c = shape
    yy, xx = np.mgrid[0:r, 0:c]

    def dist_to(lab):
        lx, ly = lab.astype(np.int64) % c, lab.astype(np.int64) // c
        return np.sqrt(((xx - lx) ** 2 + (yy - ly) ** 2).astype(np.float64)).astype(np.float32)

    own = dist_to(nn)
    bad = np.zeros(shape, bool)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if dy == 0 and dx == 0:
                continue
            shifted = np.full(shape, r * c * 3, np.int64)
            ys, yd = (slice(max(dy, 0), r + min(dy, 0)), slice(max(-dy, 0), r + min(-dy, 0)))
            xs, xd = (slice(max(dx, 0), c + min(dx, 0)), slice(max(-dx, 0), c + min(-dx, 0)))
            shifted[yd, xd] = nn[ys, xs]
            bad |= dist_to(shifted) < own
    return int(bad.sum())
