"""Python access to the parity oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/liboracle.so (our C++ restatement of the reference CPU path,
nscale_oracle.cpp) and, when built, oracle/_ref/libccref.so (the reference's
own ConnComponents.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module; the product never does.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF = HERE / "_ref" / "libccref.so"

_lib = None
_ref = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        _lib = C.CDLL(str(LIB))
    return _lib


def ref():
    """The compiled reference ConnComponents (None when not built here)."""
    global _ref
    if _ref is None and REF.exists():
        _ref = C.CDLL(str(REF))
    return _ref


def _c(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a


def _p(a):
    return C.c_void_p(a.ctypes.data)


def imreconstruct(marker, mask, conn=4):
    mask = np.ascontiguousarray(mask)
    marker = _c(marker, mask.dtype)
    out = np.empty_like(mask)
    fn = {np.uint8: "or_imreconstruct_u8", np.int32: "or_imreconstruct_i32",
          np.float32: "or_imreconstruct_f32"}[mask.dtype.type]
    getattr(lib(), fn)(_p(marker), _p(mask), _p(out), mask.shape[0], mask.shape[1], conn)
    return out


def imreconstruct_binary(marker, mask, conn=4):
    mask = np.ascontiguousarray(mask)
    marker = _c(marker, mask.dtype)
    out = np.empty_like(mask)
    fn = {np.uint8: "or_imreconstruct_binary_u8", np.int32: "or_imreconstruct_binary_i32"}[mask.dtype.type]
    getattr(lib(), fn)(_p(marker), _p(mask), _p(out), mask.shape[0], mask.shape[1], conn)
    return out


def fill_holes(img, binary=False, conn=4, i32_saturate=False):
    """imfillHoles<T>.  For int32, invert<int> wraps at INT_MIN unless
    i32_saturate (the two readings of PixelOperations.cpp:24-27)."""
    img = np.ascontiguousarray(img)
    out = np.empty_like(img)
    if img.dtype == np.int32:
        lib().or_fill_holes_i32(_p(img), _p(out), img.shape[0], img.shape[1], int(bool(binary)), conn,
                                int(bool(i32_saturate)))
    else:
        lib().or_fill_holes_u8(_p(img), _p(out), img.shape[0], img.shape[1], int(bool(binary)), conn)
    return out


def imfill(img, seeds, binary=False, conn=4):
    """imfill<uchar>(image, seeds, binary, conn) with seeds (any nonzero selects)."""
    img = _c(img, np.uint8)
    seeds = _c(seeds, np.uint8)
    out = np.empty_like(img)
    lib().or_imfill_u8(_p(img), _p(seeds), _p(out), img.shape[0], img.shape[1], int(bool(binary)), conn)
    return out


def bwselect(img, seeds, conn=8):
    """bwselect<uchar>(binaryImage, seeds, conn)."""
    img = _c(img, np.uint8)
    seeds = _c(seeds, np.uint8)
    out = np.empty_like(img)
    lib().or_bwselect_u8(_p(img), _p(seeds), _p(out), img.shape[0], img.shape[1], conn)
    return out


def imhmin(img, h, conn=8):
    """imhmin<T>(image, h, conn) for uchar (integer h in 0..255) and float."""
    img = np.ascontiguousarray(img)
    out = np.empty_like(img)
    if img.dtype == np.uint8:
        lib().or_imhmin_u8(_p(img), int(h), _p(out), img.shape[0], img.shape[1], conn)
    else:
        img = _c(img, np.float32)
        out = np.empty_like(img)
        lib().or_imhmin_f32(_p(img), C.c_float(h), _p(out), img.shape[0], img.shape[1], conn)
    return out


def local_extrema(img, conn=8, minima=False):
    """localMaxima / localMinima<T> (uchar or float): u8 255 at the extrema."""
    img = np.ascontiguousarray(img)
    out = np.empty(img.shape, np.uint8)
    kind = "minima" if minima else "maxima"
    if img.dtype == np.uint8:
        getattr(lib(), f"or_local_{kind}_u8")(_p(img), _p(out), img.shape[0], img.shape[1], conn)
    else:
        img = _c(img, np.float32)
        getattr(lib(), f"or_local_{kind}_f32")(_p(img), _p(out), img.shape[0], img.shape[1], conn)
    return out


def distance_transform(mask):
    """distanceTransform (uchar mask): (nearest-background index map int32, float distances)."""
    m = _c(mask, np.uint8)
    nn = np.empty(m.shape, np.int32)
    dist = np.empty(m.shape, np.float32)
    lib().or_distance_transform(_p(m), m.shape[0], m.shape[1], _p(nn), _p(dist))
    return nn, dist


def bounding_box(labels, bgval=0):
    """ConnComponents::boundingBox: (n, [labels, minx, maxx, miny, maxy] as a 5 x n array)."""
    a = _c(labels, np.int32)
    cap = 5 * a.size + 5
    out = np.empty(cap, np.int32)
    n = lib().or_bounding_box(_p(a), a.shape[1], a.shape[0], int(bgval), _p(out), cap)
    return int(n), out[:5 * n].reshape(5, n).copy()


def ref_bounding_box(labels, bgval=0):
    r = ref()
    a = _c(labels, np.int32)
    cap = 5 * a.size + 5
    out = np.empty(cap, np.int32)
    n = r.ref_bounding_box(_p(a), a.shape[1], a.shape[0], int(bgval), _p(out), cap)
    return int(n), out[:5 * n].reshape(5, n).copy()


def bwlabel2(binimg, conn=8, relab=True):
    b = _c(binimg, np.uint8)
    out = np.empty(b.shape, np.int32)
    n = lib().or_bwlabel2(_p(b), _p(out), b.shape[0], b.shape[1], conn, int(bool(relab)))
    return out, int(n)


def bwareaopen2(img, labeled, flatten, min_size, max_size, conn=8):
    a = np.ascontiguousarray(img)
    rows, cols = a.shape
    out_i = np.empty((rows, cols), np.int32)
    out_u = np.empty((rows, cols), np.uint8)
    n = lib().or_bwareaopen2(_p(a), int(bool(labeled)), int(bool(flatten)), min_size, max_size, conn, rows, cols,
                             _p(out_i), _p(out_u))
    return (out_u if flatten else out_i), int(n)


def relabel(labels, bgval=-1):
    a = np.array(labels, dtype=np.int32, copy=True, order="C")
    n = lib().or_relabel(_p(a), a.shape[0], a.shape[1], bgval)
    return a, int(n)


def chain_u8(marker, mask, recon_conn=8, fill_conn=4, thresh=128, ccl_conn=8, min_size=11, max_size=1000):
    mask = _c(mask, np.uint8)
    marker = _c(marker, np.uint8)
    r = np.empty_like(mask)
    f = np.empty_like(mask)
    lab = np.empty(mask.shape, np.int32)
    n = lib().or_chain_u8(_p(marker), _p(mask), mask.shape[0], mask.shape[1], recon_conn, fill_conn, int(thresh),
                          ccl_conn, min_size, max_size, _p(r), _p(f), _p(lab))
    return r, f, lab, int(n)


# ---- the reference's own ConnComponents.cpp -------------------------------------------
def ref_label(binimg, conn=8, bgval=-1):
    r = ref()
    b = _c(binimg, np.uint8).copy()
    out = np.empty(b.shape, np.int32)
    r.ref_label(_p(b), b.shape[1], b.shape[0], _p(out), bgval, conn)
    return out


def ref_relabel(labels, bgval=-1):
    r = ref()
    a = np.array(labels, dtype=np.int32, copy=True, order="C")
    n = r.ref_relabel(a.shape[1], a.shape[0], _p(a), bgval)
    return a, int(n)


def ref_area_threshold_labeled(labels, lower, upper, bgval=-1):
    r = ref()
    a = _c(labels, np.int32)
    out = np.empty_like(a)
    n = r.ref_area_threshold_labeled(_p(a), a.shape[1], a.shape[0], _p(out), bgval, lower, upper)
    return out, int(n)


def ref_area_threshold(binimg, lower, upper, conn=8, bgval=-1):
    r = ref()
    b = _c(binimg, np.uint8).copy()
    out = np.empty(b.shape, np.int32)
    n = r.ref_area_threshold(_p(b), b.shape[1], b.shape[0], _p(out), bgval, lower, upper, conn)
    return out, int(n)
