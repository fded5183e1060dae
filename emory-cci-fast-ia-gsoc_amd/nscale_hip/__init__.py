"""nscale_hip -- Python mirror of the nscale:: morphology / CCL API over the
gfx950 C ABI (include/nscale_hip.h).

Function names, argument meaning and error behaviour follow the reference's
C++ API (src/segment/MorphologicOperations.h):

    imreconstruct(seeds, image, connectivity)          (:21-22)
    imreconstructBinary(seeds, binaryImage, connectivity)  (:43-44)
    imfillHoles(image, binary, connectivity)            (:50-51)
    bwlabel2(binaryImage, connectivity, relab)          (:59)
    bwareaopen2(image, labeled, flatten, minSize, maxSize, connectivity) -> (out, count)   (:74)

Inputs are numpy arrays (host; copied to the GPU and back, like the
reference's synchronous CPU calls) or torch tensors on a ROCm device (used in
place, outputs allocated on the same device).  Failures raise NsxError, the
counterpart of the reference's cv::Exception from CV_Assert.  There is no CPU
fallback: every call runs the HIP kernels of libnscale_hip.so.
"""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from . import _lib
from ._lib import NsxImg
from ._lib import NSX_OPT_I32_INVERT as OPT_I32_INVERT
from ._lib import NSX_I32_INVERT_WRAP as I32_INVERT_WRAP
from ._lib import NSX_I32_INVERT_SATURATE as I32_INVERT_SATURATE
from ._lib import NSX_OPT_DT_ORDER as OPT_DT_ORDER
from ._lib import NSX_DT_ORDER_QUEUE as DT_ORDER_QUEUE
from ._lib import NSX_DT_ORDER_TILES as DT_ORDER_TILES

__all__ = [
    "NsxError", "Context", "default_context", "OPT_I32_INVERT", "I32_INVERT_WRAP", "I32_INVERT_SATURATE",
    "OPT_DT_ORDER", "DT_ORDER_QUEUE", "DT_ORDER_TILES",
    "imreconstruct", "imreconstructBinary", "imfillHoles", "fillHoles", "imfill", "bwselect", "imhmin", "bounding_box", "localMaxima", "localMinima", "distanceTransform",
    "bwlabel2", "bwlabel", "bwareaopen2", "relabel", "threshold_gt", "segment_chain",
    "gen_disks_u8", "gen_marker_u8", "gen_maze_u8", "scale_i32", "scale_f32",
]

_DT = {np.dtype(np.uint8): _lib.NSX_U8, np.dtype(np.int32): _lib.NSX_I32, np.dtype(np.float32): _lib.NSX_F32}


class NsxError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"nsx status {status}: {msg}")
        self.status = status


def _torch():
    try:
        import torch  # noqa: F401
        return torch
    except Exception:  # pragma: no cover
        return None


def _is_device_tensor(x) -> bool:
    t = _torch()
    return t is not None and isinstance(x, t.Tensor) and x.is_cuda


def _torch_dt(x) -> int:
    t = _torch()
    m = {t.uint8: _lib.NSX_U8, t.int32: _lib.NSX_I32, t.float32: _lib.NSX_F32}
    if x.dtype not in m:
        raise NsxError(_lib.NSX_ERR_INVALID, f"unsupported dtype {x.dtype}")
    return m[x.dtype]


def describe(x) -> NsxImg:
    """nsx_img of a 2-D numpy array or device tensor (rows, cols, pitch)."""
    if _is_device_tensor(x):
        if x.dim() != 2 or x.stride(1) != 1:
            raise NsxError(_lib.NSX_ERR_INVALID, "device image must be 2-D with unit column stride")
        return NsxImg(x.data_ptr(), x.stride(0) * x.element_size(), x.shape[0], x.shape[1], _torch_dt(x),
                      _lib.NSX_MEM_DEVICE)
    a = x
    if not isinstance(a, np.ndarray) or a.ndim != 2:
        raise NsxError(_lib.NSX_ERR_INVALID, "image must be a 2-D numpy array or device tensor (one channel)")
    if a.dtype not in _DT:
        raise NsxError(_lib.NSX_ERR_INVALID, f"unsupported dtype {a.dtype}")
    if a.size == 0:  # numpy gives empty arrays zero strides
        return NsxImg(a.ctypes.data, a.shape[1] * a.itemsize, a.shape[0], a.shape[1], _DT[a.dtype],
                      _lib.NSX_MEM_HOST)
    if a.strides[1] != a.itemsize:
        raise NsxError(_lib.NSX_ERR_INVALID, "host image must have unit column stride")
    return NsxImg(a.ctypes.data, a.strides[0], a.shape[0], a.shape[1], _DT[a.dtype], _lib.NSX_MEM_HOST)


def _like(x, dtype_np):
    if _is_device_tensor(x):
        t = _torch()
        tdt = {np.uint8: t.uint8, np.int32: t.int32, np.float32: t.float32}[dtype_np]
        return t.empty(tuple(x.shape), dtype=tdt, device=x.device)
    return np.empty(x.shape, dtype=dtype_np)


def _np_dtype(x):
    if _is_device_tensor(x):
        t = _torch()
        return {t.uint8: np.uint8, t.int32: np.int32, t.float32: np.float32}[x.dtype]
    return x.dtype.type


class Context:
    """Owns an nsx_ctx (device workspace + HBM budget).  One per thread/stream."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        st = self.lib.nsx_ctx_create(C.byref(h), device, hbm_budget_bytes)
        if st != _lib.NSX_OK:
            raise NsxError(st, f"nsx_ctx_create(device={device}) failed")
        self.h = h
        self.device = device

    def close(self) -> None:
        if getattr(self, "h", None) and not getattr(self, "_borrowed", False):
            self.lib.nsx_ctx_destroy(self.h)
        self.h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def check(self, st: int) -> None:
        if st != _lib.NSX_OK:
            msg = self.lib.nsx_last_error(self.h)
            raise NsxError(st, msg.decode() if msg else "")

    def synchronize(self) -> None:
        self.check(self.lib.nsx_ctx_synchronize(self.h))

    def workspace_bytes(self) -> int:
        return int(self.lib.nsx_ctx_workspace_bytes(self.h))

    TIMED_KINDS = ("recon_first", "recon_round", "ccl_local", "ccl_merge", "ccl_final", "ccl_small", "recon_scan",
                   "recon_drain")

    def set_timing(self, enable: bool) -> None:
        """Start (and clear) or stop HIP-event timing of the library's kernels."""
        self.check(self.lib.nsx_ctx_set_timing(self.h, int(bool(enable))))

    def read_timing(self) -> dict:
        """{kind: (total_ms, launches)} for the kernel kinds in TIMED_KINDS."""
        k = len(self.TIMED_KINDS)
        ms = (C.c_double * k)()
        n = (C.c_int64 * k)()
        self.check(self.lib.nsx_ctx_read_timing(self.h, ms, n, k))
        return {name: (float(ms[i]), int(n[i])) for i, name in enumerate(self.TIMED_KINDS)}

    def read_tile_counts(self) -> dict:
        """Tiles processed by propagation reconstructions since timing was last
        switched: {(mode, dtype): (first_pass_tiles, round_tile_visits)} for the
        kinds with any, mode in ("recon", "binary", "fill", "fillbin"), dtype in
        ("u8", "i32", "f32")."""
        first = (C.c_int64 * 12)()
        rounds = (C.c_int64 * 12)()
        self.check(self.lib.nsx_ctx_read_tile_counts(self.h, first, rounds, 12))
        out = {}
        for k in range(12):
            if first[k] or rounds[k]:
                out[(("recon", "binary", "fill", "fillbin")[k // 3], ("u8", "i32", "f32")[k % 3])] = \
                    (int(first[k]), int(rounds[k]))
        return out

    def set_tuning(self, param: int, value: int) -> None:
        self.check(self.lib.nsx_ctx_set_tuning(self.h, param, value))

    def get_tuning(self, param: int) -> int:
        v = C.c_int(0)
        self.check(self.lib.nsx_ctx_get_tuning(self.h, param, C.byref(v)))
        return int(v.value)

    def drain_failures(self) -> int:
        """Persistent drains (TUNE_DRAIN) that gave up and restarted with per-round launches."""
        v = C.c_int64(0)
        self.check(self.lib.nsx_ctx_drain_failures(self.h, C.byref(v)))
        return int(v.value)

    def dt_resumes(self) -> int:
        """distanceTransform queue generations finished after a host round trip
        (more solve passes than TUNE_DT_SOLVE launched)."""
        v = C.c_int64(0)
        self.check(self.lib.nsx_ctx_dt_resumes(self.h, C.byref(v)))
        return int(v.value)

    def set_option(self, option: int, value: int) -> None:
        """Semantic options (these change results), e.g.
        set_option(OPT_I32_INVERT, I32_INVERT_SATURATE)."""
        self.check(self.lib.nsx_ctx_set_option(self.h, option, value))

    def get_option(self, option: int) -> int:
        v = C.c_int(0)
        self.check(self.lib.nsx_ctx_get_option(self.h, option, C.byref(v)))
        return int(v.value)

    def round_log(self, stage: int = 0) -> list:
        """Tiles in each propagation round's work list of the last reconstruction
        (stage 1: the chain's fill); entry 0 is the first pass."""
        buf = (C.c_int32 * 4096)()
        n = C.c_int(0)
        self.check(self.lib.nsx_last_round_log(self.h, stage, buf, 4096, C.byref(n)))
        return [int(buf[i]) for i in range(n.value)]

    def recon_stats(self, stage: int = 0) -> dict:
        """Statistics of the last reconstruction (stage 1: the chain's fill)."""
        s = (C.c_int64 * 4)()
        self.check(self.lib.nsx_last_recon_stats(self.h, stage, s))
        return {"rounds": int(s[0]), "tile_visits": int(s[1]), "tiles": int(s[2]), "half_passes": int(s[3])}


_tls = threading.local()


def default_context(device: int = 0) -> Context:
    ctxs = getattr(_tls, "ctxs", None)
    if ctxs is None:
        ctxs = _tls.ctxs = {}
    if device not in ctxs:
        ctxs[device] = Context(device)
    return ctxs[device]


def _ctx(x, ctx):
    if ctx is not None:
        return ctx
    dev = x.device.index if _is_device_tensor(x) and x.device.index is not None else 0
    return default_context(dev)


# HIP's handle of the legacy default stream (hipStreamLegacy).  torch reports
# its default stream as handle 0, which the C library reads as "the context's
# own stream" (a non-blocking stream, unordered with torch's work), so a torch
# stream with handle 0 is passed as this handle instead.
HIP_STREAM_LEGACY = 1


def _torch_handle(ts) -> C.c_void_p:
    h = ts.cuda_stream
    return C.c_void_p(h if h else HIP_STREAM_LEGACY)


def _stream_arg(stream):
    """A torch stream object, a raw hipStream_t handle, or None (the context's stream)."""
    if stream is None:
        return None
    if hasattr(stream, "cuda_stream"):
        return _torch_handle(stream)
    return C.c_void_p(stream)


def _torch_stream(x, stream):
    """Device tensors run on torch's current stream unless a stream is given."""
    if stream is not None:
        return _stream_arg(stream)
    if _is_device_tensor(x):
        return _torch_handle(_torch().cuda.current_stream(x.device))
    return None


def imreconstruct(seeds, image, connectivity: int = 4, *, out=None, ctx=None, stream=None):
    """Grayscale reconstruction by dilation (nscale::imreconstruct<T>)."""
    c = _ctx(image, ctx)
    out = _like(image, _np_dtype(image)) if out is None else out
    c.check(c.lib.nsx_imreconstruct(c.h, C.byref(describe(seeds)), C.byref(describe(image)),
                                    C.byref(describe(out)), connectivity, _torch_stream(image, stream)))
    return out


def imreconstructBinary(seeds, binaryImage, connectivity: int = 4, *, out=None, ctx=None, stream=None):
    """Binary reconstruction (nscale::imreconstructBinary<T>)."""
    c = _ctx(binaryImage, ctx)
    out = _like(binaryImage, _np_dtype(binaryImage)) if out is None else out
    c.check(c.lib.nsx_imreconstruct_binary(c.h, C.byref(describe(seeds)), C.byref(describe(binaryImage)),
                                           C.byref(describe(out)), connectivity,
                                           _torch_stream(binaryImage, stream)))
    return out


def imfillHoles(image, binary: bool = False, connectivity: int = 4, *, out=None, ctx=None, stream=None):
    """Hole filling (nscale::imfillHoles<T>)."""
    c = _ctx(image, ctx)
    out = _like(image, _np_dtype(image)) if out is None else out
    c.check(c.lib.nsx_fill_holes(c.h, C.byref(describe(image)), C.byref(describe(out)), int(bool(binary)),
                                 connectivity, _torch_stream(image, stream)))
    return out


fillHoles = imfillHoles


def imfill(image, seeds, binary: bool = False, connectivity: int = 4, *, out=None, ctx=None, stream=None):
    """Seeded fill (nscale::imfill<uchar>): image | reconstruction of the
    complement from the seeds (u8 image and seeds)."""
    c = _ctx(image, ctx)
    out = _like(image, np.uint8) if out is None else out
    c.check(c.lib.nsx_imfill(c.h, C.byref(describe(image)), C.byref(describe(seeds)), C.byref(describe(out)),
                             int(bool(binary)), connectivity, _torch_stream(image, stream)))
    return out


def localMaxima(image, connectivity: int = 8, *, out=None, ctx=None, stream=None):
    """nscale::localMaxima<T> (uchar or float): u8, 255 at the local maxima."""
    c = _ctx(image, ctx)
    out = _like(image, np.uint8) if out is None else out
    c.check(c.lib.nsx_local_maxima(c.h, C.byref(describe(image)), C.byref(describe(out)), connectivity,
                                   _torch_stream(image, stream)))
    return out


def localMinima(image, connectivity: int = 8, *, out=None, ctx=None, stream=None):
    """nscale::localMinima<T> (uchar or float): u8, 255 at the local minima."""
    c = _ctx(image, ctx)
    out = _like(image, np.uint8) if out is None else out
    c.check(c.lib.nsx_local_minima(c.h, C.byref(describe(image)), C.byref(describe(out)), connectivity,
                                   _torch_stream(image, stream)))
    return out


def distanceTransform(mask, calcDist: bool = True, *, out=None, ctx=None, stream=None):
    """nscale::distanceTransform (uchar mask): float32 distances to the nearest
    background, or with calcDist=False int32 nearest-background raster indices."""
    c = _ctx(mask, ctx)
    out = _like(mask, np.float32 if calcDist else np.int32) if out is None else out
    c.check(c.lib.nsx_distance_transform(c.h, C.byref(describe(mask)), C.byref(describe(out)), 1 if calcDist else 0,
                                         _torch_stream(mask, stream)))
    return out


def bounding_box(labels, bgval: int = 0, *, ctx=None, stream=None):
    """ConnComponents::boundingBox of an int32 label image: (count, 5 x count
    int32 array of labels, min x, max x, min y, max y in order of first
    appearance in raster order)."""
    c = _ctx(labels, ctx)
    n = C.c_int32(0)
    img = describe(labels)
    c.check(c.lib.nsx_bounding_box(c.h, C.byref(img), int(bgval), None, 0, C.byref(n), _torch_stream(labels, stream)))
    boxes = np.empty(5 * max(1, n.value), np.int32)
    if n.value:
        c.check(c.lib.nsx_bounding_box(c.h, C.byref(img), int(bgval), boxes.ctypes.data_as(_lib._I32P), n.value,
                                       C.byref(n), _torch_stream(labels, stream)))
    return int(n.value), boxes[:5 * n.value].reshape(5, n.value)


def imhmin(image, h, connectivity: int = 8, *, out=None, ctx=None, stream=None):
    """Suppression of minima shallower than h (nscale::imhmin<T>, uchar or float)."""
    c = _ctx(image, ctx)
    out = _like(image, _np_dtype(image)) if out is None else out
    c.check(c.lib.nsx_imhmin(c.h, C.byref(describe(image)), float(h), C.byref(describe(out)), connectivity,
                             _torch_stream(image, stream)))
    return out


def bwselect(binaryImage, seeds, connectivity: int = 8, *, out=None, ctx=None, stream=None):
    """Components holding a seed (nscale::bwselect<uchar>), with their values."""
    c = _ctx(binaryImage, ctx)
    out = _like(binaryImage, np.uint8) if out is None else out
    c.check(c.lib.nsx_bwselect(c.h, C.byref(describe(binaryImage)), C.byref(describe(seeds)), C.byref(describe(out)),
                               connectivity, _torch_stream(binaryImage, stream)))
    return out


def bwlabel2(binaryImage, connectivity: int = 8, relab: bool = True, *, out=None, ctx=None, stream=None,
             return_count: bool = False):
    """Union-find labelling (nscale::bwlabel2): int32 labels."""
    c = _ctx(binaryImage, ctx)
    out = _like(binaryImage, np.int32) if out is None else out
    n = C.c_int32(0)
    c.check(c.lib.nsx_bwlabel(c.h, C.byref(describe(binaryImage)), C.byref(describe(out)), connectivity,
                              int(bool(relab)), C.byref(n), _torch_stream(binaryImage, stream)))
    return (out, int(n.value)) if return_count else out


def bwlabel(binaryImage, connectivity: int = 8, relab: bool = True, **kw):
    """Union-find labelling under the name of the reference's contour bwlabel."""
    return bwlabel2(binaryImage, connectivity, relab, **kw)


def bwareaopen2(image, labeled: bool, flatten: bool, minSize: int, maxSize: int, connectivity: int = 8, *,
                out=None, ctx=None, stream=None):
    """Area filter (nscale::bwareaopen2), inclusive min / exclusive max.  Returns (out, count)."""
    c = _ctx(image, ctx)
    out = _like(image, np.uint8 if flatten else np.int32) if out is None else out
    n = C.c_int32(0)
    c.check(c.lib.nsx_bwareaopen(c.h, C.byref(describe(image)), int(bool(labeled)), int(bool(flatten)), minSize,
                                 maxSize, connectivity, C.byref(describe(out)), C.byref(n),
                                 _torch_stream(image, stream)))
    return out, int(n.value)


def relabel(labels, bgval: int = -1, *, ctx=None, stream=None):
    """ConnComponents::relabel in place; returns (labels, count)."""
    c = _ctx(labels, ctx)
    n = C.c_int32(0)
    c.check(c.lib.nsx_relabel(c.h, C.byref(describe(labels)), bgval, C.byref(n), _torch_stream(labels, stream)))
    return labels, int(n.value)


def threshold_gt(image, k: float, *, out=None, ctx=None, stream=None):
    """`image > k` as u8 255/0."""
    c = _ctx(image, ctx)
    out = _like(image, np.uint8) if out is None else out
    c.check(c.lib.nsx_threshold_gt(c.h, C.byref(describe(image)), float(k), C.byref(describe(out)),
                                   _torch_stream(image, stream)))
    return out


def segment_chain(marker, mask, recon_conn=8, fill_conn=4, thresh=128.0, ccl_conn=8, min_size=11, max_size=1000,
                  *, recon_out=None, fill_out=None, labels_out=None, want_intermediates=True, ctx=None,
                  stream=None):
    """imreconstruct -> imfillHoles -> (> thresh) -> bwareaopen2 -> relabel, fused.

    Returns (recon, filled, labels, ncomp); recon/filled are None when
    want_intermediates is False and no output buffer was given."""
    c = _ctx(mask, ctx)
    dt = _np_dtype(mask)
    if want_intermediates:
        recon_out = _like(mask, dt) if recon_out is None else recon_out
        fill_out = _like(mask, dt) if fill_out is None else fill_out
    labels_out = _like(mask, np.int32) if labels_out is None else labels_out
    n = C.c_int32(0)
    r = C.byref(describe(recon_out)) if recon_out is not None else None
    f = C.byref(describe(fill_out)) if fill_out is not None else None
    c.check(c.lib.nsx_segment_chain(c.h, C.byref(describe(marker)), C.byref(describe(mask)), recon_conn, fill_conn,
                                    float(thresh), ccl_conn, min_size, max_size, r, f,
                                    C.byref(describe(labels_out)), C.byref(n), _torch_stream(mask, stream)))
    return recon_out, fill_out, labels_out, int(n.value)


def recon_label(marker, mask, conn=8, thresh=128.0, ccl_conn=8, *, recon_out=None, labels_out=None, ctx=None,
                stream=None):
    """One tile of the batched workload: imreconstruct then bwlabel(recon > thresh).

    Returns (recon, labels, ncomp)."""
    c = _ctx(mask, ctx)
    recon_out = _like(mask, _np_dtype(mask)) if recon_out is None else recon_out
    labels_out = _like(mask, np.int32) if labels_out is None else labels_out
    n = C.c_int32(0)
    c.check(c.lib.nsx_recon_label(c.h, C.byref(describe(marker)), C.byref(describe(mask)), conn, float(thresh),
                                  ccl_conn, C.byref(describe(recon_out)), C.byref(describe(labels_out)), C.byref(n),
                                  _torch_stream(mask, stream)))
    return recon_out, labels_out, int(n.value)


class Batch:
    """Many device-resident tiles of uneven size in one set of launches,
    captured as a hipGraph (nsx_batch_*).  Per tile: recon =
    imreconstruct(marker, mask, conn); labels = bwlabel(recon > thresh)."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or default_context(0)
        self.lib = self.ctx.lib
        h = C.c_void_p()
        self.ctx.check(self.lib.nsx_batch_create(self.ctx.h, C.byref(h)))
        self.h = h
        self.n = 0
        self._keep = []   # tensors referenced by the batch

    def add(self, marker, mask, recon_out, labels_out, conn=8, thresh=128.0, ccl_conn=8) -> int:
        for t in (marker, mask, recon_out, labels_out):
            if not _is_device_tensor(t):
                raise NsxError(_lib.NSX_ERR_INVALID, "batch tiles must be device tensors")
        self.ctx.check(self.lib.nsx_batch_add(self.h, C.byref(describe(marker)), C.byref(describe(mask)),
                                              C.byref(describe(recon_out)), C.byref(describe(labels_out)), conn,
                                              float(thresh), ccl_conn))
        self._keep.append((marker, mask, recon_out, labels_out))
        self.n += 1
        return self.n - 1

    def finalize(self, use_graph: bool = True) -> None:
        self.ctx.check(self.lib.nsx_batch_finalize(self.h, int(bool(use_graph))))

    def run(self, stream=None) -> None:
        s = _stream_arg(stream) if stream is not None else _torch_handle(_torch().cuda.current_stream())
        self.ctx.check(self.lib.nsx_batch_run(self.h, s))

    def autotune(self, reps: int = 3, stream=None) -> list:
        """Coordinate descent over each tile group's launch shape (drain or
        per-round launches, sweep pairs per visit, XCD split, drain blocks per
        CU) timed by graph replays (nsx_batch_autotune); returns tuning() of the
        kept settings.  Results do not depend on it."""
        s = _stream_arg(stream) if stream is not None else _torch_handle(_torch().cuda.current_stream())
        pairs = (C.c_int32 * 4)()
        self.ctx.check(self.lib.nsx_batch_autotune(self.h, s, int(reps), pairs, 4))
        return self.tuning()

    def tuning(self) -> list:
        """Per tile group: {"visit_pairs", "drain", "xcd_split", "drain_blocks_per_cu",
        "overlap"} (overlap: odd groups on a second stream; batch-wide)."""
        out = []
        g = 0
        while True:
            v = (C.c_int32 * 5)()
            if self.lib.nsx_batch_get_tuning(self.h, g, v, 5) != _lib.NSX_OK:
                return out
            out.append({"visit_pairs": int(v[0]), "drain": int(v[1]), "xcd_split": int(v[2]),
                        "drain_blocks_per_cu": int(v[3]), "overlap": int(v[4])})
            g += 1

    def set_overlap(self, on: bool) -> None:
        """Odd tile groups on a second stream beside the even ones (default), or all in turn."""
        self.ctx.check(self.lib.nsx_batch_set_overlap(self.h, int(bool(on))))

    def set_drain(self, on: bool) -> None:
        """Every group to the persistent drain (default) or to per-round launches."""
        self.ctx.check(self.lib.nsx_batch_set_drain(self.h, int(bool(on))))

    def context(self) -> "Context":
        """The batch's private context (timing of eager runs, statistics); a
        borrowed handle, valid while the batch lives."""
        c = Context.__new__(Context)
        c.lib = self.lib
        c.h = C.c_void_p(self.lib.nsx_batch_context(self.h))
        c.device = self.ctx.device
        c._borrowed = True
        return c

    def counts(self) -> list:
        arr = (C.c_int32 * max(1, self.n))()
        self.ctx.check(self.lib.nsx_batch_counts(self.h, arr, self.n))
        return [int(arr[i]) for i in range(self.n)]

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.nsx_batch_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving loop for the segmentation chain (nsx_server_*): requests are
    admitted in arrival order as worker slots and HBM arena space free up, and
    the work of admitted requests overlaps on the GPU.  Each request computes
    segment_chain(marker, mask, ...) on its own buffers."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 1 << 30, slots: int = 4,
                 max_batch: "int | None" = None):
        self.lib = _lib.load()
        h = C.c_void_p()
        st = self.lib.nsx_server_create(device, hbm_budget_bytes, slots, C.byref(h))
        if st != _lib.NSX_OK:
            raise NsxError(st, f"nsx_server_create(device={device}, budget={hbm_budget_bytes}, slots={slots})")
        self.h = h
        self.device = device
        self._keep = {}   # ticket -> buffers the worker may still read or write
        if max_batch is not None:
            self.set_sharing(max_batch)

    def set_sharing(self, max_requests: int = 16, max_request_pixels: int = 0, max_set_pixels: int = 0) -> None:
        """Sharing policy: an idle slot takes the oldest request and, when it
        has at most max_request_pixels pixels, the small ones queued behind it
        (at most max_requests, max_set_pixels in all); those of one kind share
        their launches.  max_requests=1 serves every request on its own; a
        pixel limit of 0 keeps the default (2^20, 2^22)."""
        self._check(self.lib.nsx_server_set_sharing(self.h, int(max_requests), int(max_request_pixels),
                                                    int(max_set_pixels)))

    def _check(self, st: int) -> None:
        if st != _lib.NSX_OK:
            raise NsxError(st, (self.lib.nsx_server_last_error(self.h) or b"").decode())

    def submit(self, marker, mask, recon_conn=8, fill_conn=4, thresh=128.0, ccl_conn=8, min_size=11,
               max_size=1000, *, labels_out=None, recon_out=None, fill_out=None) -> int:
        """Queues one request; returns its ticket.  Host (numpy) or device
        images; outputs are allocated like the mask when not given (labels
        always, recon/fill only when passed)."""
        labels_out = _like(mask, np.int32) if labels_out is None else labels_out
        t = C.c_uint64(0)
        r = C.byref(describe(recon_out)) if recon_out is not None else None
        f = C.byref(describe(fill_out)) if fill_out is not None else None
        self._check(self.lib.nsx_server_submit(self.h, C.byref(describe(marker)), C.byref(describe(mask)),
                                               recon_conn, fill_conn, float(thresh), ccl_conn, min_size, max_size,
                                               r, f, C.byref(describe(labels_out)), C.byref(t)))
        self._keep[t.value] = (marker, mask, labels_out, recon_out, fill_out)
        return t.value

    def wait(self, ticket: int):
        """Blocks until the request is done.  Returns (labels, recon, fill,
        ncomp, times_ms) with times_ms = (queued, service, total)."""
        n = C.c_int32(0)
        tm = (C.c_double * 3)()
        st = self.lib.nsx_server_wait(self.h, ticket, C.byref(n), tm)
        bufs = self._keep.pop(ticket, None)
        self._check(st)
        return bufs[2], bufs[3], bufs[4], int(n.value), (tm[0], tm[1], tm[2])

    def context(self, slot: int) -> "Context":
        """Worker slot `slot`'s context (timing, tile counts); a borrowed
        handle, valid while the server lives.  Switch its timing only while no
        request is in flight."""
        h = self.lib.nsx_server_context(self.h, int(slot))
        if not h:
            raise NsxError(_lib.NSX_ERR_INVALID, f"no worker slot {slot}")
        c = Context.__new__(Context)
        c.lib = self.lib
        c.h = C.c_void_p(h)
        c.device = self.device
        c._borrowed = True
        return c

    def stats(self) -> dict:
        a = (C.c_int64 * 8)()
        self._check(self.lib.nsx_server_stats(self.h, a))
        keys = ("completed", "in_flight", "queued", "peak_in_flight", "arena_bytes", "arena_peak_bytes",
                "workspace_bytes", "slots")
        out = dict(zip(keys, (int(v) for v in a)))
        b = (C.c_int64 * 7)()
        self._check(self.lib.nsx_server_batch_stats(self.h, b))
        out.update(zip(("shared_sets", "shared_requests", "largest_set", "max_batch", "shared_short_rounds",
                        "share_max_request_px", "share_max_set_px"), (int(v) for v in b)))
        return out

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.nsx_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


# ---- synthetic inputs ------------------------------------------------------------
def gen_disks_u8(rows: int, cols: int, ndisks: int, seed: int) -> np.ndarray:
    lib = _lib.load()
    a = np.empty((rows, cols), np.uint8)
    if lib.nsx_gen_disks_u8(a.ctypes.data, rows, cols, cols, ndisks, seed) != 0:
        raise NsxError(_lib.NSX_ERR_INVALID, "gen_disks_u8")
    return a


def gen_marker_u8(mask: np.ndarray, sub: int = 30) -> np.ndarray:
    lib = _lib.load()
    mask = np.ascontiguousarray(mask)
    out = np.empty_like(mask)
    lib.nsx_gen_marker_u8(mask.ctypes.data, out.ctypes.data, mask.size, sub)
    return out


def gen_maze_u8(rows: int, cols: int, seed: int):
    lib = _lib.load()
    mask = np.empty((rows, cols), np.uint8)
    marker = np.empty((rows, cols), np.uint8)
    if lib.nsx_gen_maze_u8(mask.ctypes.data, marker.ctypes.data, rows, cols, seed) != 0:
        raise NsxError(_lib.NSX_ERR_INVALID, "gen_maze_u8")
    return mask, marker


def scale_i32(a: np.ndarray, scale: int = 257) -> np.ndarray:
    lib = _lib.load()
    a = np.ascontiguousarray(a)
    out = np.empty(a.shape, np.int32)
    lib.nsx_gen_scale_i32(a.ctypes.data, out.ctypes.data, a.size, scale)
    return out


def scale_f32(a: np.ndarray) -> np.ndarray:
    lib = _lib.load()
    a = np.ascontiguousarray(a)
    out = np.empty(a.shape, np.float32)
    lib.nsx_gen_scale_f32(a.ctypes.data, out.ctypes.data, a.size)
    return out
