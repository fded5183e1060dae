// nsx_internal.h -- shared declarations of the gfx950 morphology / CCL library.
#pragma once

#include "round_hints.h"
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>
#include <mutex>
#include <unordered_map>
#include <cstring>

#include "../../include/nscale_hip.h"

namespace nsx {

// ---- errors ------------------------------------------------------------------
struct Error {
    nsx_status code;
    std::string msg;
};

#define NSX_HIP_TRY(ctx, expr)                                                            \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            return (ctx)->fail(NSX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
        }                                                                                \
    } while (0)

#define NSX_TRY(expr)                    \
    do {                                 \
        nsx_status s_ = (expr);          \
        if (s_ != NSX_OK) return s_;     \
    } while (0)

// ---- device workspace ----------------------------------------------------------
// A named, grow-only device buffer held by a context.  Every buffer counts
// against the context's HBM budget.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    uint64_t serial = 0;   // changes whenever the slot is (re)allocated
};

}  // namespace nsx

namespace nsx {
// Kernel kinds timed with HIP events when a context's timing is on.
enum TimedKind {
    TK_RECON_FIRST = 0,   // first full pass of a reconstruction (all tiles)
    TK_RECON_ROUND,       // one later propagation round / drain launch
    TK_CCL_LOCAL,         // tile-local labelling
    TK_CCL_MERGE,         // tile-edge union
    TK_CCL_FINAL,         // relabel + output write
    TK_CCL_SMALL,         // area / keep / scan kernels (one event pair per call)
    TK_RECON_SCAN,        // image-wide row / column scans before a hole filling's rounds
    TK_RECON_DRAIN,       // the persistent drain of a reconstruction's rounds (one launch)
    TK_NKINDS
};
}  // namespace nsx

struct nsx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t budget = 0;       // 0 = unlimited
    size_t held = 0;         // bytes of workspace held
    std::string last_error;
    std::vector<nsx::DevBuf> bufs;   // indexed by workspace slot
    int64_t* d_stats = nullptr;      // [rounds, tile passes, ntiles, spare] ..., see kCumWord
    // Tiles processed by reconstructions since timing was last enabled, per
    // (mode, dtype) at index mode * 3 + dtype: first passes counted here on
    // the host, propagation-round visits in d_stats[kCumWord + index].
    static constexpr int kCumWord = 32, kCumKinds = 12;
    int64_t first_tiles[kCumKinds] = {};
    int32_t* h_pinned = nullptr;     // small pinned readback area
    int cu_count = 256;
    // Work-list length of every propagation round of the last reconstruction
    // per stats stage (0, 1): kRoundLog ints each, round r at [r] (r = 0 the
    // first pass), written by the round kernels.
    static constexpr int kRoundLog = 4096;
    int32_t* d_round_log = nullptr;

    // Tuning (nsx_ctx_set_tuning).
    int visit_sweeps = 4;
    int levelset_levels = 3;
    int engine = NSX_ENGINE_AUTO;     // NSX_TUNE_ENGINE
    int fill_scans = 0;               // NSX_TUNE_FILL_SCANS
    int drain = 0;                    // NSX_TUNE_DRAIN
    int dt_solve = 2;                 // NSX_TUNE_DT_SOLVE
    int64_t drain_failures = 0;       // drains that gave up (restarted with per-round launches)

    // Options (nsx_ctx_set_option).
    int i32_invert = NSX_I32_INVERT_WRAP;
    int dt_order = NSX_DT_ORDER_QUEUE;

    // distanceTransform in queue order: the next generation tag, the
    // allocation of WS_DT_TAGS the tags were last cleared for, and how many
    // generations needed more solve passes than launched (a host round trip).
    uint32_t dtq_tag = 0;
    uint64_t dtq_tags_serial = 0;
    int64_t dtq_resumes = 0;
    uint64_t ws_serial = 0;

    // Propagation rounds the last reconstruction of a given size/mode needed.
    nsx::RoundHints round_hint;

    // HIP-event kernel timing (off by default).
    bool timing = false;
    struct EvPair { hipEvent_t a, b; int kind; };
    std::vector<EvPair> ev_pending;
    std::vector<hipEvent_t> ev_pool;
    double t_ms[nsx::TK_NKINDS] = {};
    int64_t t_n[nsx::TK_NKINDS] = {};

    nsx_status fail(nsx_status code, const std::string& msg) {
        last_error = msg;
        return code;
    }
    // Returns a device buffer of at least `bytes` in workspace slot `slot`.
    nsx_status ws(int slot, size_t bytes, void** out);

    // Event bracket around launches of `kind` on stream s (no-op when timing off).
    int tic(int kind, hipStream_t s);
    void toc(int handle, hipStream_t s);
    // Folds completed event pairs into t_ms / t_n (call after a stream sync).
    void collect_timing();
};

namespace nsx {

// Workspace slots.
enum WsSlot {
    WS_RECON_STAMP = 0,  // per-tile queue stamps
    WS_RECON_LIST,       // 2 segmented tile lists
    WS_RECON_COUNTS,     // 3 rounds x 8 segment counters (spaced) + a sum word
    WS_CCL_LABEL,        // int32 per pixel (when the caller's output cannot hold it)
    WS_CCL_AREA,         // int32 per pixel (roots only are touched)
    WS_CCL_ROOTS,        // compact local-root table: (global index, count) pairs
    WS_CCL_BITS,         // kept-root bitmap
    WS_CCL_SCAN,         // per-word prefix + block sums
    WS_CCL_MISC,         // counters
    WS_CHAIN_RECON,      // chain intermediates
    WS_CHAIN_FILL,
    WS_FILL_STAMP,       // a second reconstruction in flight (the chain's fill)
    WS_FILL_LIST,
    WS_FILL_COUNTS,
    WS_STAGE_A,          // host<->device staging
    WS_STAGE_B,
    WS_STAGE_C,
    WS_CCL_LROOT,        // u16 per pixel index: tile-local root (0xFFFF background)
    WS_MT_PAR,           // component-tree reconstruction: parent of each node (int32)
    WS_MT_M,             // subtree maximum of the marker per node (u32)
    WS_MT_G,             // level per node (u8)
    WS_MT_R,             // result per node (u8)
    WS_MT_LNODE,         // node of each pixel within its tile (u16)
    WS_MT_TCOUNT,        // nodes per tile
    WS_MT_WALK,          // nodes whose parent changed in the tile-edge merge
    WS_MT_CTL,           // counters and error bits
    WS_MT2_PAR,          // a second component-tree reconstruction in flight (the chain's fill)
    WS_MT2_M,
    WS_MT2_G,
    WS_MT2_R,
    WS_MT2_LNODE,
    WS_MT2_TCOUNT,
    WS_MT2_WALK,
    WS_MT2_CTL,
    WS_SERVE_TABLES,     // serving: descriptors, tile tables and counts of a shared launch set
    WS_SCAN_SEG,         // image-wide column scans: per-segment clamp compositions
    WS_BP_L,             // bit-plane reconstruction: class labels of a pass (u8)
    WS_BP_W,             //   and its seed weights (u8)
    WS_BP_COPY,          //   an input that is also the output, copied first
    WS_CCL2_LABEL,       // a second labelling in flight (a batch's odd tile groups, on their own
    WS_CCL2_AREA,        //   stream): the same six slots as WS_CCL_LABEL .. WS_CCL_MISC
    WS_CCL2_ROOTS,
    WS_CCL2_BITS,
    WS_CCL2_SCAN,
    WS_CCL2_MISC,
    WS_CCL2_LROOT,
    WS_DT_TAGS,          // distanceTransform queue order: per pixel two tagged words (list head, last offer)
    WS_DT_ENT,           //   per queue entry: positions and labels (two generations), list links, push masks
    WS_DT_PART,          //   per chunk of entries: counts, then offsets
    WS_NSLOTS
};

// ---- image descriptors ------------------------------------------------------------
struct View {
    char* data;
    int64_t pitch;  // bytes
    int rows, cols;
    int dtype;
};

inline size_t dtype_size(int dt) { return dt == NSX_U8 ? 1 : 4; }

// Host<->device copy of `rows` rows of `rowb` bytes.  When both sides are
// contiguous (pitch == row width) it is one linear copy: a strided 2-D copy
// between pinned host memory and the device is split into per-row transfers.
inline hipError_t copy_rows(void* dst, size_t dpitch, const void* src, size_t spitch, size_t rowb, size_t rows,
                            hipMemcpyKind kind, hipStream_t s) {
    if (rows == 0 || rowb == 0) return hipSuccess;
    if (dpitch == rowb && spitch == rowb) return hipMemcpyAsync(dst, src, rowb * rows, kind, s);
    return hipMemcpy2DAsync(dst, dpitch, src, spitch, rowb, rows, kind, s);
}

// ---- reconstruction ---------------------------------------------------------------
enum ReconMode { RM_RECON = 0, RM_BINARY = 1, RM_FILL = 2, RM_FILLBIN = 3 };

// Runs reconstruction in `mode` on device views (marker may be null for the
// fill modes) to convergence.  `state` is the output buffer and holds the
// propagation state.  Synchronises the stream at least once.
struct LevelLabelReq;
// `stats`: the device statistics words to fill (default ctx->d_stats, stage 0;
// a chain's fill passes stage 1, ctx->d_stats + 8).
nsx_status recon_run(nsx_ctx* ctx, int mode, const View* marker, const View& mask, const View& state,
                     int conn, hipStream_t s, LevelLabelReq* req = nullptr, int64_t* stats = nullptr);

// Several images of one dtype/mode/connectivity in one set of launches.
struct ReconBatch {
    int nimg = 0;
    int ntiles = 0;                 // total tiles over images
    const void* host_desc = nullptr;  // the descriptor when nimg == 1
    const void* dev_desc = nullptr;   // device array of nimg descriptors
    const int32_t* tile_img = nullptr;  // device: tile -> image index (nimg > 1)
    int ws_base = 0;                // first of 3 consecutive workspace slots
    int64_t* stats = nullptr;       // device stats words (default ctx->d_stats)
    int visit_sweeps = 0;           // sweep pairs per visit; 0: the context's setting
    int drain = 0;                  // rounds in one persistent launch: 1 on, -1 off, 0 the context's setting
    int drain_cap = 0;              // the drain stops before this round (0: no cap); the host goes on from there
    int drain_occ = 0;              // drain blocks per CU (0: as many as stay resident)
    int xcd_split = -1;             // XCD-contiguous work split: 1 on, 0 off, -1 on for one image only
};
struct ReconRunImpl;
// A launched reconstruction whose convergence has not been checked yet.
struct ReconRun {
    ReconRunImpl* impl = nullptr;
    ~ReconRun() { release(); }
    void release();
};
int recon_tiles(int rows, int cols);
size_t recon_desc_bytes();
// Bytes of the two round lists (second of a reconstruction's ws slots).
size_t recon_list_bytes(int ntiles);
void recon_fill_desc(void* out, const View* marker, const View& mask, const View& state, int tile_base);
// Enqueues round 0 and the predicted number of further rounds (no host sync).
// When pending_dev is given, the number of tiles still queued after the last
// launched round is copied there (0 = converged) for a later batched check.
nsx_status recon_launch(nsx_ctx* ctx, int mode, int dtype, int conn, const ReconBatch& B, hipStream_t s,
                        ReconRun* run, int32_t* pending_dev);
// Synchronises, launches further rounds until the work list is empty, and
// updates the round prediction.  *extended tells whether rounds were added.
// max_rounds > 0: once the run passes that many rounds unconverged, it is
// released and *gave_up set (the caller finishes another way; the round
// prediction for the size is raised so the next call can skip the rounds).
nsx_status recon_complete(nsx_ctx* ctx, ReconRun* run, hipStream_t s, bool* extended, int max_rounds = 0,
                          bool* gave_up = nullptr);
// Rounds the last reconstruction of this kind and size needed (0 = none yet).
int recon_predicted_rounds(nsx_ctx* ctx, int mode, int dtype, int conn, int ntiles, int visit_sweeps);
// For a run already known to have converged: records how many rounds had work
// (read by the caller from the run's stats word 0) and releases the run.
// One kernel that gathers a chain's status for a single copy: out[0] = *ncomp,
// out[1], out[2] = tiles still queued by rr and rf (0 when their predicted
// rounds sufficed), int64 out[2], out[3] = rstats[0], fstats[0].  Both runs
// must hold a launch (non-empty images).
void recon_status_gather(const ReconRun* rr, const ReconRun* rf, const int32_t* ncomp, const int64_t* rstats,
                         const int64_t* fstats, int32_t* out, hipStream_t s);
void recon_note_rounds(nsx_ctx* ctx, ReconRun* run, int64_t rounds_with_work);

// The allocation-free reconstruction (nsx_imreconstruct_noalloc): work lists
// in a file-scope device array of recon_noalloc_scratch_bytes(), shared by
// every call on a device.  recon_noalloc_init looks up the array's device
// address for `device` (which must be current); recon_run_noalloc runs
// imreconstruct (propagation rounds only) on device views with at most
// recon_noalloc_max_tiles() tiles, allocating nothing.
nsx_status recon_noalloc_init(int device);
size_t recon_noalloc_scratch_bytes();
int recon_noalloc_max_tiles();
nsx_status recon_run_noalloc(nsx_ctx* ctx, const View& marker, const View& mask, const View& state, int conn,
                             hipStream_t s);

// ---- reconstruction through the component tree (u8) ------------------------------------
// Enqueues the component-tree reconstruction of a u8 image (mode RM_RECON:
// marker under mask; RM_FILL: hole filling of `mask`, marker unused) into
// `out`, a fixed number of launches.  *err_dev receives the device word that
// is nonzero when a bound was hit (the result is then invalid and must be
// recomputed by propagation); it stays valid until the next launch on the same
// workspace slots.  ws_base: WS_MT_PAR or WS_MT2_PAR.
nsx_status mtree_launch(nsx_ctx* ctx, int mode, const View* marker, const View& mask, const View& out, int conn,
                        hipStream_t s, uint32_t** err_dev, int ws_base = WS_MT_PAR);
int mtree_tiles(int rows, int cols);

// ---- reconstruction through bit planes (u8) ---------------------------------------------
// u8 imreconstruct (RM_RECON) or gray hole filling (RM_FILL, marker unused)
// into `out` in eight labelling passes, one per bit from the top; enqueues
// work only.  Any image content, a fixed number of launches.
nsx_status recon_bitplane_u8(nsx_ctx* ctx, int mode, const View* marker, const View& mask, const View& out, int conn,
                             hipStream_t s);

// ---- connected components ---------------------------------------------------------
enum CclOut {
    CO_ROOT = 0,          // int32: root raster index, background -1
    CO_ROOT_FILTERED = 1, // int32: root index if kept by the area filter, else -1
    CO_CANON = 2,         // int32: 1..N in raster order of kept roots, background 0
    CO_FLAT_U8 = 3,       // u8: 255 if kept, else 0
    CO_LEVEL = 4          // u8: out = max(out, min(level, component's maximum weight)) on the foreground
};

struct CclParams {
    int conn = 8;
    int out_mode = CO_CANON;
    bool threshold = false;   // input is thresholded on the fly: fg = in > thr
    double thr = 0.0;
    int min_size = 0;         // area filter (inclusive) ; filter on when use_area
    int max_size = 0;         // (exclusive)
    bool use_area = false;
    int level = 0;            // CO_LEVEL
    // CO_LEVEL: with floor_mmax, every pixel is written (no read of `out`):
    // background min(floor_level, *floor_mmax), foreground the larger of that
    // and the level value
    int floor_level = 0;
    const uint32_t* floor_mmax = nullptr;
};

// Labels the image `in` (u8; or u8/i32/f32 with threshold) into `out`.
// `ncomp_dev` (may be null) receives the kept component count on the device.
// Enqueues work only.
nsx_status ccl_run(nsx_ctx* ctx, const View& in, const View& out, const CclParams& p,
                   int32_t* ncomp_dev, hipStream_t s);

// One level of the level-set reconstruction of a u8 image: components of
// {mask >= level} (conn), each folded into `out` as min(level, its maximum of
// `weight`).  With `labels`, the same pass also writes the canonical labels of
// the components whose maximum exceeds `thresh` (0 elsewhere) and their count
// to *ncomp_dev.  With floor_mmax, `out` is written everywhere instead of
// folded (see CclParams).  Enqueues work only.
nsx_status ccl_level_run(nsx_ctx* ctx, const View& mask, const View& weight, const View& out, int conn, int level,
                         hipStream_t s, const View* labels = nullptr, double thresh = 0.0,
                         int32_t* ncomp_dev = nullptr, int floor_level = 0, const uint32_t* floor_mmax = nullptr);

// Reconstruction by dilation of a u8 image through its level sets when the
// mask holds at most max_levels + 1 distinct values: R = max over mask values
// L of min(L, largest marker value in the component of {mask >= L}), one
// labelling per level above the smallest.  Reads the value set back (one
// synchronisation); *used = false (nothing written) when there are more levels.
// A labelling of {out > thresh} wanted together with a u8 reconstruction
// (nsx_recon_label, same connectivity).  {R > t} is the union of the
// components of {mask >= L*} whose largest marker value exceeds t, L* the
// smallest mask value above t (none: empty; the smallest mask value: the whole
// image), so the level-set path labels it from its own labelling of L*.
struct LevelLabelReq {
    View labels;              // int32 output
    double thresh = 0.0;
    int32_t* ncomp_dev = nullptr;
    bool done = false;        // set when the labels (and the count) were enqueued
};
nsx_status recon_levelset_u8(nsx_ctx* ctx, const View& marker, const View& mask, const View& out, int conn,
                             int max_levels, bool* used, hipStream_t s, LevelLabelReq* req = nullptr);

// Several images labelled by one set of launches.  Image b's pixels occupy a
// global index range starting at its pix_base (a multiple of 32768, see
// ccl_pix_span); tiles are numbered across images.
struct CclBatch {
    int nimg = 0;
    int ntiles = 0;
    int64_t pix_total = 0;            // end of the last image's index range
    const void* host_desc = nullptr;  // the descriptor when nimg == 1
    const void* dev_desc = nullptr;   // device array of nimg descriptors
    const int32_t* tile_img = nullptr;  // device: tile -> image (nimg > 1)
    int in_dtype = NSX_U8;
    int ws_base = 0;                  // first of 6 consecutive workspace slots
    int lroot_slot = WS_CCL_LROOT;    // slot of the u16 tile-local roots
    int32_t* ncomp_dev = nullptr;     // device: kept components per image
};
size_t ccl_desc_bytes();
int ccl_tiles(int rows, int cols);
int64_t ccl_pix_span(int rows, int cols);
void ccl_fill_desc(void* out, const View& in, const View& o, int tile_base, int64_t pix_base);
nsx_status ccl_launch(nsx_ctx* ctx, const CclBatch& B, const CclParams& p, hipStream_t s);

// Components of equal nonzero values of the u8 image `classes`; each folds its
// largest `weight` into the u8 image `out` (out = max(out, weight max) on
// the nonzero pixels).  Enqueues work only.
nsx_status ccl_class_fold_run(nsx_ctx* ctx, const View& classes, const View& weight, const View& out, int conn,
                              hipStream_t s);

// imreconstructBinary (u8, int32): the marker where it is nonzero, the type's
// maximum at the other pixels of the conn-connected components of the nonzero
// mask pixels that hold a nonzero marker pixel, 0 elsewhere; one labelling.
// Enqueues work only.
nsx_status binrecon_run(nsx_ctx* ctx, const View& marker, const View& mask, const View& out, int conn, hipStream_t s);

// bwselect (u8): `out` = the image on the conn-connected components of its
// nonzero pixels that hold a nonzero seed, 0 elsewhere; one labelling.
// Enqueues work only.
nsx_status bwselect_run(nsx_ctx* ctx, const View& img, const View& seeds, const View& out, int conn, hipStream_t s);

// Area filter on int32 root labels (background -1), out int32 or u8.
nsx_status ccl_area_labeled(nsx_ctx* ctx, const View& in, const View& out, int min_size, int max_size,
                            bool flatten, int32_t* count_dev, hipStream_t s);

// Canonical relabel of int32 labels (roots are label[i]==i), in place.
nsx_status ccl_relabel(nsx_ctx* ctx, const View& lab, int bgval, int32_t* ncomp_dev, hipStream_t s);

// Pointwise `in > k` -> u8 255/0.
nsx_status threshold_run(nsx_ctx* ctx, const View& in, double k, const View& out, hipStream_t s);

// imfill / bwselect markers (u8): marker = (mask ? 255 - img : img) where
// seeds != 0, else 0; with `mask`, 255 - img is also written there.
nsx_status seeded_marker_run(nsx_ctx* ctx, const View& img, const View& seeds, const View& marker, const View* mask,
                             hipStream_t s);
// imhmin's mask = invert(img) and marker = mask - h (u8 saturating, or f32).
nsx_status hmin_prep_run(nsx_ctx* ctx, const View& img, double h, const View& mask, const View& marker,
                         hipStream_t s);
// ConnComponents::boundingBox of int32 labels into host `boxes` (5 x count
// ints: labels, min x, max x, min y, max y in order of first appearance);
// boxes == nullptr: count only.  Synchronises.
nsx_status bbox_run(nsx_ctx* ctx, const View& lab, int bgval, int32_t* boxes, int capacity, int32_t* count,
                    hipStream_t s);
// localMaxima / localMinima of a u8 or f32 image into u8 out (255 at the
// extrema).  Synchronises.
nsx_status local_extrema_run(nsx_ctx* ctx, const View& img, const View& out, int conn, bool minima, hipStream_t s);
// out = invert(out) in place (u8 or f32).
nsx_status invert_run(nsx_ctx* ctx, const View& out, hipStream_t s);
// distanceTransform of a u8 mask into out: f32 distances (calc_dist) or int32
// nearest-background raster indices.  Synchronises.
nsx_status dist_run(nsx_ctx* ctx, const View& mask, const View& out, bool calc_dist, hipStream_t s);
// out = img | out, or img & out (and_op), u8.
nsx_status bitwise_u8_run(nsx_ctx* ctx, const View& img, const View& out, bool and_op, hipStream_t s);

// Pointwise i32 fill of a constant.
nsx_status fill_const_i32(nsx_ctx* ctx, const View& out, int32_t v, hipStream_t s);

}  // namespace nsx
