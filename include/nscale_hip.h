/* nscale_hip.h -- C ABI of the MI355X (gfx950) morphology + CCL library.
 *
 * Drop-in boundary for the nscale:: morphological-reconstruction and
 * connected-component path of fastIA (paths below are relative to
 * fastIA-master/).  Every entry point takes plain pointers and sizes, returns an
 * nsx_status and never throws; reference CV_Assert failures become
 * NSX_ERR_INVALID with a message from nsx_last_error().
 *
 * Images are described by nsx_img (what cv::Mat exposes: data, step, rows,
 * cols, type).  Buffers are caller-owned.  NSX_MEM_DEVICE images live in HBM
 * (hipMalloc'd, on the context's device); NSX_MEM_HOST images are copied in
 * and out by the call (synchronously, like the reference CPU functions).
 * Output images must not alias inputs.
 *
 * A context owns the device workspace (tile work lists, union-find tables)
 * and its HBM budget.  Contexts are independent: there is no module-global
 * device state, so one context per host thread or stream is reentrant.
 * Entry points are ordered on the stream passed in (a hipStream_t, or NULL for
 * the context's stream); calls that return a count (ncomp, count) synchronise
 * that stream before returning.
 */
#ifndef NSCALE_HIP_H_
#define NSCALE_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    NSX_OK = 0,
    NSX_ERR_INVALID = 1,      /* bad argument (reference: CV_Assert) */
    NSX_ERR_UNSUPPORTED = 2,  /* dtype/mode the reference does not instantiate */
    NSX_ERR_HIP = 3,          /* HIP runtime error */
    NSX_ERR_NOMEM = 4,        /* device allocation failed */
    NSX_ERR_BUDGET = 5,       /* request exceeds the context's HBM budget */
    NSX_ERR_NOT_CONVERGED = 6 /* propagation hit its round limit */
} nsx_status;

typedef enum { NSX_U8 = 0, NSX_I32 = 1, NSX_F32 = 2 } nsx_dtype;
typedef enum { NSX_MEM_DEVICE = 0, NSX_MEM_HOST = 1 } nsx_mem;

typedef struct {
    void* data;
    int64_t pitch_bytes; /* row step in bytes (cv::Mat::step) */
    int32_t rows;
    int32_t cols;
    int32_t dtype;       /* nsx_dtype */
    int32_t mem;         /* nsx_mem */
} nsx_img;

typedef struct nsx_ctx nsx_ctx;

/* ---- context -------------------------------------------------------------- */
/* Creates a context on `device`.  hbm_budget_bytes bounds the workspace the
 * context may hold (0 = no limit).  No reference counterpart: the reference
 * keeps its queue state in module globals (src/segment/cuda/global_queue.cu:9-22). */
nsx_status nsx_ctx_create(nsx_ctx** ctx, int device, size_t hbm_budget_bytes);
void nsx_ctx_destroy(nsx_ctx* ctx);
const char* nsx_last_error(const nsx_ctx* ctx);
nsx_status nsx_ctx_synchronize(nsx_ctx* ctx);
/* Device bytes currently held by the context's workspace. */
size_t nsx_ctx_workspace_bytes(const nsx_ctx* ctx);
/* Reads the statistics of the last reconstruction run on this context;
 * stage 0 = the call's (first) reconstruction, stage 1 = the fill stage of
 * nsx_segment_chain.  stats[0] = propagation rounds that had work after the
 * first full pass, stats[1] = tile visits in those rounds, stats[2] = tiles,
 * stats[3] = sweep half-passes summed over all visits.  Synchronises the device. */
nsx_status nsx_last_recon_stats(nsx_ctx* ctx, int stage, int64_t stats[4]);
/* Work-list occupancy of the last propagation reconstruction of a stage (as
 * nsx_last_recon_stats): counts[r] = tiles (62 x 62 pixels) in round r's device
 * work list, r = 0 the first pass, for the rounds that had work (at most n
 * entries, 4096 logged); *rounds receives how many are valid.  Reconstructions
 * that did not run by rounds (level sets, bit planes, component tree) leave the
 * log of the rounds before they switched, if any.  Synchronises the device.  No
 * reference counterpart (the reference's queue kernels report nothing). */
nsx_status nsx_last_round_log(nsx_ctx* ctx, int stage, int32_t* counts, int n, int* rounds);

/* Kernel timing with HIP events recorded on the stream each kernel runs on.
 * nsx_ctx_set_timing(ctx, 1) starts recording (and clears the totals);
 * nsx_ctx_read_timing fills ms[k] (summed kernel time) and launches[k] for the
 * kernel kinds k = 0 reconstruction first pass, 1 reconstruction propagation
 * launch, 2 tile-local labelling, 3 tile-edge union, 4 label output, 5 area /
 * keep / scan kernels (per call), 6 a hole filling's image-wide row and column
 * scans (per call), 7 a persistent drain of a reconstruction's rounds
 * (NSX_TUNE_DRAIN; one launch).  Synchronises the context.  No reference
 * counterpart (the reference's CUDA event timing is commented out,
 * src/segment/cuda/ccl_uf.cu:28-31). */
nsx_status nsx_ctx_set_timing(nsx_ctx* ctx, int enable);
nsx_status nsx_ctx_read_timing(nsx_ctx* ctx, double* ms, int64_t* launches, int nkinds);
/* Tiles (62 x 62 pixels) processed by the context's propagation
 * reconstructions since nsx_ctx_set_timing last started or stopped timing
 * (counted whether timing is on or not): first_pass[k] by first passes,
 * rounds[k] by propagation rounds, for k = mode * 3 + dtype (mode 0
 * imreconstruct, 1 imreconstructBinary, 2 grayscale fill, 3 binary fill;
 * dtype as nsx_dtype), n <= 12 entries.  With the kernel times these give the
 * algorithmic bytes per launch of a mixed workload.  Synchronises the device.
 * No reference counterpart. */
nsx_status nsx_ctx_read_tile_counts(nsx_ctx* ctx, int64_t* first_pass, int64_t* rounds, int n);

/* Tuning parameters of the context's kernels (results never depend on them).
 * NSX_TUNE_VISIT_SWEEPS: sweep pairs (column + row half-pass) a reconstruction
 * tile runs per visit before it re-queues itself unfinished, 1..64 (default
 * 4; 4-connected hole filling, whose tiles run on four waves with shorter
 * pairs, runs one pair more); small caps keep one slow tile from stretching a
 * propagation round.
 * NSX_TUNE_LEVELSET_LEVELS: a u8 reconstruction (nsx_imreconstruct,
 * nsx_recon_label) whose mask holds at most this many distinct values above
 * its smallest one is computed through the mask's level sets, one
 * union-find labelling per level, instead of by propagation rounds; 0..255
 * (default 3, 0 = never).  Long thin propagation paths (mazes, binary-like
 * masks) then cost a few labellings instead of one round per tile crossing.
 * No reference counterpart. */
/* NSX_TUNE_ENGINE chooses how u8 imreconstruct and gray imfillHoles run:
 * NSX_ENGINE_ROUNDS: propagation rounds (and the level-set path above);
 * NSX_ENGINE_TREE: through the mask's component tree, a fixed number of
 * launches whatever the image holds, so its time does not grow with the
 * distances values travel (long many-valued paths), at a higher cost on
 * ordinary images;
 * NSX_ENGINE_PLANES: through the bit planes of the result, eight labelling
 * passes (one per bit, from the top) whatever the image holds;
 * NSX_ENGINE_AUTO (default): rounds, switching to the bit planes when a
 * reconstruction passes 256 rounds, and starting with them when the last
 * reconstruction of the same kind and size did.
 * Other types and modes always use rounds.  Results are the same. */
/* NSX_TUNE_FILL_SCANS: image-wide row + column scan pairs run before the
 * propagation rounds of a gray hole filling (u8, int32) of one image at most
 * 4096 columns wide whose tile grid is at least 24 tiles (of 62 pixels) on
 * each side, 0..8 (default 0 = never).  A row or column scan is the exact
 * reconstruction along every line of the image, so the frame's value crosses
 * the open parts of the image at once, and the rounds, which move values one
 * tile per round, are left with the twisted paths.  Off by default: on noisy
 * images the rounds that remain are set by how many sweep pairs a tile visit
 * runs, and the rounds saved do not repay the scans.  Counterpart of the
 * reference's 1-D passes (cuda/imreconstruct_int_kernel.cu:217-585), which
 * alternate until nothing changes. */
/* NSX_TUNE_DRAIN: 1 runs every propagation round after a reconstruction's
 * first pass in one persistent launch (blocks sized to stay resident on the
 * whole GPU drain the device tile lists round by round, separated by
 * grid-wide barriers), 0 (default) launches one kernel per round.  The drain
 * needs no round prediction or host check; a barrier that waits 0.5 s without
 * progress (a GPU shared with other work) makes the call start again with
 * per-round launches.  The serving loop's slots always use per-round
 * launches; nsx_batch_autotune picks per tile group.  Counterpart of the
 * reference's persistent queue kernel (global_queue.cu:746-857). */
/* NSX_TUNE_DT_SOLVE: relaxation passes nsx_distance_transform launches per
 * queue generation (NSX_DT_ORDER_QUEUE) before checking, 0..8, default 2 (the
 * second pass only revisits entries that read labels of other chunks).  A
 * generation whose labels the passes did not settle is found by the check and
 * finished with more passes after a host round trip; results do not depend on
 * this value. */
enum { NSX_TUNE_VISIT_SWEEPS = 0, NSX_TUNE_LEVELSET_LEVELS = 1, NSX_TUNE_ENGINE = 2, NSX_TUNE_FILL_SCANS = 3,
       NSX_TUNE_DRAIN = 4, NSX_TUNE_DT_SOLVE = 5, NSX_TUNE_NPARAMS = 6 };
enum { NSX_ENGINE_ROUNDS = 0, NSX_ENGINE_TREE = 1, NSX_ENGINE_AUTO = 2, NSX_ENGINE_PLANES = 3 };
nsx_status nsx_ctx_set_tuning(nsx_ctx* ctx, int param, int value);
nsx_status nsx_ctx_get_tuning(const nsx_ctx* ctx, int param, int* value);
/* Persistent drains (NSX_TUNE_DRAIN) on this context that gave up and were
 * restarted with per-round launches, since the context was created.  No
 * reference counterpart. */
nsx_status nsx_ctx_drain_failures(const nsx_ctx* ctx, int64_t* n);

/* Queue generations of nsx_distance_transform (NSX_DT_ORDER_QUEUE) on this
 * context that the launched solve passes (NSX_TUNE_DT_SOLVE) did not settle
 * and that were finished after a host round trip.  No reference counterpart. */
nsx_status nsx_ctx_dt_resumes(const nsx_ctx* ctx, int64_t* n);

/* Semantic options of a context (these change results).
 * NSX_OPT_I32_INVERT: how PixelOperations::invert<int> (PixelOperations.cpp:
 * 17-39: bitwise_not(x) + 1 through OpenCV's CV_32S Mat + Scalar) treats
 * INT_MIN, the frame value imfillHoles<int> pads with
 * (MorphologicOperations.cpp:1889-1903).
 *   NSX_I32_INVERT_WRAP (default): two's-complement wrap, INT_MIN -> INT_MIN,
 *     as the reference's own GPU kernel computes it (~x + 1 in int,
 *     src/segment/cuda/pixel-ops.cu:39-48, instantiated for imfillholes,
 *     PixelOperationsGPU.cpp:301) and as OpenCV documents CV_32S arithmetic
 *     (no saturation).  The inverted mask's frame is then INT_MIN, so the
 *     reconstruction is INT_MIN everywhere and every int32 imfillHoles result,
 *     gray or binary, is the constant INT_MIN.
 *   NSX_I32_INVERT_SATURATE: INT_MIN -> INT_MAX.  Gray int32 imfillHoles is
 *     then the textbook fill (MATLAB imfill 'holes'); binary int32
 *     imfillHoles is the constant INT_MAX (the marker's interior min<int> is
 *     nonzero and is kept).
 * Applies to nsx_fill_holes and nsx_segment_chain on int32 images.  No
 * reference counterpart (OpenCV 2.x is not available to settle the
 * reference's CPU behaviour, so the default follows its GPU code). */
enum { NSX_OPT_I32_INVERT = 0, NSX_OPT_DT_ORDER = 1 };
enum { NSX_I32_INVERT_WRAP = 0, NSX_I32_INVERT_SATURATE = 1 };
/* NSX_OPT_DT_ORDER: the order in which nsx_distance_transform applies the
 * reference's propagation rule (MorphologicOperations.cpp:90-119).
 *   NSX_DT_ORDER_QUEUE (default): the reference's FIFO queue order
 *     (:224-245), reproduced generation by generation, so labels and
 *     distances equal the reference's bit for bit.
 *   NSX_DT_ORDER_TILES: synchronous passes in 64x64 tiles to a fixed point of
 *     the same rule.  Where the rule has several fixed points (pixels the
 *     propagation leaves short of the nearest background, and equally near
 *     backgrounds) the result can differ from the reference's; it is never
 *     nearer than the exact Euclidean distance.  No reference counterpart. */
enum { NSX_DT_ORDER_QUEUE = 0, NSX_DT_ORDER_TILES = 1 };
nsx_status nsx_ctx_set_option(nsx_ctx* ctx, int option, int value);
nsx_status nsx_ctx_get_option(const nsx_ctx* ctx, int option, int* value);

/* ---- morphology ----------------------------------------------------------- */
/* Grayscale reconstruction by dilation of `marker` under `mask`; conn 8 gives
 * 8-connectivity, anything else 4.  u8, i32, f32.
 * Replaces nscale::imreconstruct<T>(const Mat& seeds, const Mat& image, int)
 * (src/segment/MorphologicOperations.h:21-22, .cpp:665-838) and
 * nscale::gpu::imreconstruct / imreconstructQueueSpeedup (.h:101-130,
 * MorphologicOperationsGPU.cpp:172-200,286-441). */
nsx_status nsx_imreconstruct(nsx_ctx* ctx, const nsx_img* marker, const nsx_img* mask,
                             nsx_img* out, int conn, void* stream);

/* nsx_imreconstruct without allocating: no hipMalloc, hipFree or allocating
 * library call.  No counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  Device images only (marker, mask, out);
 * u8, i32, f32; conn 8 or 4; the same result as nsx_imreconstruct.  The tile
 * work lists live in a static device array of the library (about 1.1 MB per
 * GPU, held until the process ends), which limits images to 16384
 * reconstruction tiles of 62 x 62 pixels (e.g. 7936 x 7936); larger images are
 * refused with NSX_ERR_INVALID.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (other threads, other streams) race and must be
 * serialised by the caller.  The context supplies the stream and the pinned
 * word the convergence check reads; it must have been created (which looks up
 * the array's device address) before the call.  Synchronises the stream. */
nsx_status nsx_imreconstruct_noalloc(nsx_ctx* ctx, const nsx_img* marker, const nsx_img* mask,
                                     nsx_img* out, int conn, void* stream);

/* Binary reconstruction: pixels of `mask` != 0 connected to a seed (marker != 0
 * and mask != 0) become max(T); nonzero marker pixels are kept as given.
 * u8, i32.  Replaces nscale::imreconstructBinary<T> (MorphologicOperations.h:43-44,
 * .cpp:1688-1829) and nscale::gpu::imreconstructBinary (MorphologicOperationsGPU.cpp:686-721). */
nsx_status nsx_imreconstruct_binary(nsx_ctx* ctx, const nsx_img* marker, const nsx_img* mask,
                                    nsx_img* out, int conn, void* stream);

/* Hole filling.  u8, i32.  Replaces nscale::imfillHoles<T>(const Mat&, bool
 * binary, int) (MorphologicOperations.h:50-51, .cpp:1859-1926) and
 * nscale::gpu::imfillHoles (MorphologicOperationsGPU.cpp:751-836). */
nsx_status nsx_fill_holes(nsx_ctx* ctx, const nsx_img* in, nsx_img* out, int binary, int conn,
                          void* stream);

/* imfill with seeds (MorphologicOperations.cpp:1833-1857, uchar only, :2516):
 * mask = 255 - image; marker = mask where seeds != 0, else 0; marker =
 * imreconstructBinary (binary) or imreconstruct (marker, mask, conn); out =
 * image | marker.  image, seeds (any nonzero selects) and out are u8. */
nsx_status nsx_imfill(nsx_ctx* ctx, const nsx_img* image, const nsx_img* seeds, nsx_img* out, int binary,
                      int conn, void* stream);

/* bwselect (MorphologicOperations.cpp:1929-1951, uchar only, :2522): marker =
 * binary_image where seeds != 0, else 0; out = imreconstructBinary(marker,
 * binary_image, conn) & binary_image, i.e. the components of the nonzero
 * pixels that hold a seed, with their values.  All three images u8. */
nsx_status nsx_bwselect(nsx_ctx* ctx, const nsx_img* binary_image, const nsx_img* seeds, nsx_img* out, int conn,
                        void* stream);

/* imhmin (MorphologicOperations.cpp:2260-2282, uchar and float, :2527-2528):
 * suppresses the minima shallower than h.  mask = invert(image) (uchar 255 - x,
 * float -x); out = invert(imreconstruct(mask - h, mask, conn)), with uchar
 * mask - h saturating at 0 (uchar h: an integer in 0..255).  Float images may
 * be negative; the reconstruction's constant 0 frame applies as in the
 * reference. */
nsx_status nsx_imhmin(nsx_ctx* ctx, const nsx_img* image, double h, nsx_img* out, int conn, void* stream);

/* localMaxima / localMinima (MorphologicOperations.cpp:2374-2456, uchar and
 * float, :2529-2532): u8 out, 255 at the local maxima (minima: of the
 * complement).  Candidates are the pixels where imreconstruct(image - 1,
 * image) stays at image - 1 (uchar saturating); a connected region of equal-
 * valued candidates that touches a strictly higher neighbour is dropped (the
 * reference's cv::floodFill, fixed range, mask only). */
nsx_status nsx_local_maxima(nsx_ctx* ctx, const nsx_img* image, nsx_img* out, int conn, void* stream);
nsx_status nsx_local_minima(nsx_ctx* ctx, const nsx_img* image, nsx_img* out, int conn, void* stream);

/* distanceTransform (MorphologicOperations.cpp:197-275): u8 mask, zero pixels
 * are background.  calc_dist != 0: out is f32, the Euclidean distance to the
 * nearest background found by 8-neighbour propagation of nearest-background
 * labels (a neighbour's label is taken when strictly closer after rounding the
 * distance to float); calc_dist == 0: out is int32, those labels as raster
 * indices (rows*cols*3 where no background was found, the reference's initial
 * value).  In the default queue order (NSX_OPT_DT_ORDER) labels and
 * distances are the reference's, ties and pixels the propagation leaves short
 * of the nearest background included.  rows and cols must be below 65535 and
 * rows*cols*3 must fit int32. */
nsx_status nsx_distance_transform(nsx_ctx* ctx, const nsx_img* mask, nsx_img* out, int calc_dist, void* stream);

/* ConnComponents::boundingBox (ConnComponents.cpp:241-284): for the labels of
 * an int32 image other than bgval, in order of first appearance in raster
 * order, the bounding box.  boxes (host memory, room for 5 x capacity ints)
 * receives the reference's layout: count labels, then their min x, max x, min
 * y and max y.  A label -1 or 0 that is not bgval ends the scan, as in the
 * reference.  Labels must lie in [-1, rows*cols] (the reference's hash map
 * takes any int; every label bwlabel2 writes is in that range).  boxes NULL:
 * *count only.  *count is set also when capacity is too small (error).
 * Synchronises. */
nsx_status nsx_bounding_box(nsx_ctx* ctx, const nsx_img* labels, int bgval, int32_t* boxes, int capacity,
                            int32_t* count, void* stream);

/* ---- connected components ------------------------------------------------- */
/* Union-find labelling of a u8 image: 4/8-neighbours of equal nonzero value are
 * joined.  relabel=0: label = raster index of the component's first pixel,
 * background -1.  relabel=1: labels 1..N in raster order of those first pixels,
 * background 0; *ncomp = N (may be NULL).
 * Replaces nscale::bwlabel2(const Mat&, int, bool) (MorphologicOperations.h:59,
 * .cpp:2012-2035; ConnComponents.cpp:52-118) and nscale::gpu::bwlabel
 * (MorphologicOperationsGPU.cpp:878-900). */
nsx_status nsx_bwlabel(nsx_ctx* ctx, const nsx_img* bin, nsx_img* labels, int conn, int relabel,
                       int32_t* ncomp, void* stream);

/* Area filter, inclusive min, exclusive max.  labeled=0: `in` is u8 and is
 * labelled first (conn); labeled=1: `in` is int32 root labels, background -1.
 * flatten=0: out is int32 root labels of the kept components, background -1;
 * flatten=1: out is u8, 255 on kept components, else 0.  *count = number kept.
 * Replaces nscale::bwareaopen2 (MorphologicOperations.h:74, .cpp:2199-2232;
 * ConnComponents.cpp:205-239) and nscale::gpu::bwareaopen2
 * (MorphologicOperationsGPU.cpp:1024-1057). */
nsx_status nsx_bwareaopen(nsx_ctx* ctx, const nsx_img* in, int labeled, int flatten, int min_size,
                          int max_size, int conn, nsx_img* out, int32_t* count, void* stream);

/* Canonical relabel of int32 labels in place: pixels whose label equals their
 * own raster index are roots, numbered 1.. in raster order; background (bgval)
 * and everything else map to their root's number, background to 0.
 * Replaces ConnComponents::relabel (ConnComponents.cpp:89-118) and
 * nscale::gpu::relabel (src/segment/cuda/ccl_uf.cu:285-479). */
nsx_status nsx_relabel(nsx_ctx* ctx, nsx_img* labels, int bgval, int32_t* ncomp, void* stream);

/* `Mat > k` glue on u8/i32/f32 input: out (u8) = 255 where in > k, else 0.
 * Replaces the comparisons at HistologicalEntities.cpp:214,261 and
 * nscale::gpu::PixelOperations::threshold (src/segment/cuda/pixel-ops.cu:348-362). */
nsx_status nsx_threshold_gt(nsx_ctx* ctx, const nsx_img* in, double k, nsx_img* out, void* stream);

/* ---- fused chain ---------------------------------------------------------- */
/* The nuclei-candidate chain (HistologicalEntities.cpp:195,219,246,567):
 *   recon  = imreconstruct(marker, mask, recon_conn)
 *   filled = imfillHoles(recon, binary=0, fill_conn)
 *   labels = relabel(bwareaopen2(filled > thresh, min_size, max_size, ccl_conn))
 * recon_out / fill_out may be NULL when only labels are wanted (then the
 * context holds the intermediates).  u8 and i32 marker/mask; labels int32 with
 * background 0.  *ncomp = number of labels (may be NULL). */
nsx_status nsx_segment_chain(nsx_ctx* ctx, const nsx_img* marker, const nsx_img* mask,
                             int recon_conn, int fill_conn, double thresh, int ccl_conn,
                             int min_size, int max_size, nsx_img* recon_out, nsx_img* fill_out,
                             nsx_img* labels_out, int32_t* ncomp, void* stream);

/* One tile of the batched workload on its own: recon = imreconstruct(marker,
 * mask, conn); labels = bwlabel(recon > thresh, ccl_conn, relabel=1); *ncomp =
 * number of labels.  u8, i32, f32 marker/mask; device or host images.  The
 * per-tile launch path that nsx_batch_* is compared against (the reference's
 * per-image gpu::imreconstruct + gpu::bwlabel, MorphologicOperationsGPU.cpp:
 * 172-200,878-900). */
nsx_status nsx_recon_label(nsx_ctx* ctx, const nsx_img* marker, const nsx_img* mask, int conn, double thresh,
                           int ccl_conn, nsx_img* recon_out, nsx_img* labels_out, int32_t* ncomp,
                           void* stream);

/* ---- batched tiles -------------------------------------------------------- */
/* Many independent device-resident tiles of uneven size run reconstruction +
 * threshold + labelling together, captured once in a hipGraph and replayed.
 * Replaces nscale::gpu::imreconstructQueueThroughput (MorphologicOperationsGPU.cpp:447-545,
 * global_queue.cu:903-968), which gives each image one thread block.
 * Per tile: out_recon = imreconstruct(marker, mask, conn); labels =
 * bwlabel(out_recon > thresh, ccl_conn, relabel=1). */
typedef struct nsx_batch nsx_batch;
nsx_status nsx_batch_create(nsx_ctx* ctx, nsx_batch** batch);
nsx_status nsx_batch_add(nsx_batch* batch, const nsx_img* marker, const nsx_img* mask,
                         nsx_img* recon_out, nsx_img* labels_out, int conn, double thresh,
                         int ccl_conn);
/* Builds the work tables and (use_graph=1) captures the graph.  Call once after
 * the last add. */
nsx_status nsx_batch_finalize(nsx_batch* batch, int use_graph);
/* Runs the whole batch on the stream (graph replay when captured).  A tile
 * group set to the persistent drain (nsx_batch_set_drain(b, 1) or the tuner's
 * choice; NSX_TUNE_DRAIN) runs its propagation rounds in one launch that
 * drains the group's device tile lists to convergence, so the graph's launches
 * do not depend on the content and every run is final.  A group on per-round
 * launches (the default: faster on MI355X) runs the number of rounds it needed
 * when last completed; its outputs are final only after nsx_batch_counts,
 * which checks every group's convergence (and a drain's failure) and, where
 * tile contents changed since and need more rounds, finishes that group
 * eagerly, relabels it and re-captures the graph.  A caller that rewrites
 * tile contents between runs of such a group must therefore call
 * nsx_batch_counts before reading the outputs. */
nsx_status nsx_batch_run(nsx_batch* batch, void* stream);
/* Completes the last run (see nsx_batch_run) and returns the per-tile
 * component counts (array of n tiles).  Synchronises. */
nsx_status nsx_batch_counts(nsx_batch* batch, int32_t* counts, int n);
/* Tunes a batch finalized with a graph: for each group of tiles with the same
 * (dtype, conn, ccl_conn, thresh), in the order groups first appear, a
 * coordinate descent over the group's launch shape: persistent drain or one
 * launch per round, 3..6 sweep pairs per tile visit (NSX_TUNE_VISIT_SWEEPS for
 * that group only), the split of a round's tiles over the XCDs (plain stride
 * or a contiguous share per XCD) and the drain's blocks per CU (all that stay
 * resident, 8 or 4); before each pass, whether odd groups run on a second
 * stream, overlapping the even groups' launches (the default), or every group
 * runs on the caller's stream.  Each setting is timed as the mean of `reps` graph
 * replays on `stream` after re-learning the round counts and re-capturing;
 * the fastest is kept and left captured (a second pass runs if the first
 * changed anything).  pairs[g] (g < npairs) receives the sweep pairs, 0 past
 * the last group; nsx_batch_get_tuning reads every choice.  Results never
 * depend on it; it runs the batch about 10 x groups x (reps + 2) times.  No
 * counterpart in the reference (its batched path has fixed launch shapes). */
nsx_status nsx_batch_autotune(nsx_batch* batch, void* stream, int reps, int32_t* pairs, int npairs);
/* Launch shape of tile group `group` (groups in order of first appearance):
 * out[0] sweep pairs per visit, out[1] 1 persistent drain / 0 per-round
 * launches, out[2] 1 XCD-contiguous split / 0 plain stride, out[3] drain
 * blocks per CU (0: all that stay resident), out[4] 1 when odd groups run on
 * a second stream beside the even ones (the batch-wide overlap setting);
 * n <= 5 entries.  NSX_ERR_INVALID past the last group.  No reference
 * counterpart. */
nsx_status nsx_batch_get_tuning(nsx_batch* batch, int group, int32_t* out, int n);
/* Sets every group to the persistent drain (drain = 1) or to per-round
 * launches (0, the default), dropping the tuner's choice of it; re-learns
 * round counts and re-captures a finalized batch.  No reference counterpart. */
nsx_status nsx_batch_set_drain(nsx_batch* batch, int drain);
/* overlap = 1 (default): odd tile groups run on a second stream beside the
 * even ones (forked from and joined back to the run's stream, also inside the
 * graph), so one group's late rounds, which hold few tiles, overlap the other
 * group's launches; 0: every group on the run's stream in turn.  Groups on the
 * persistent drain always run in turn.  Re-captures a finalized batch.  No
 * reference counterpart. */
nsx_status nsx_batch_set_overlap(nsx_batch* batch, int overlap);
/* The batch's private context (owned by the batch; valid until
 * nsx_batch_destroy): its kernel timing (nsx_ctx_set_timing /
 * nsx_ctx_read_timing, recorded by eager runs of a batch finalized without a
 * graph) and its reconstruction statistics (nsx_last_recon_stats, stage g =
 * tile group g mod 2).  No reference counterpart. */
nsx_ctx* nsx_batch_context(nsx_batch* batch);
void nsx_batch_destroy(nsx_batch* batch);

/* ---- serving loop --------------------------------------------------------- */
/* Continuous admission of segmentation-chain requests (the nsx_segment_chain
 * pipeline) on one GPU.  `slots` worker threads each own a context and a
 * stream; requests are admitted when a slot is idle and their device buffers
 * fit in a fixed HBM arena of `hbm_budget_bytes` (allocated once; no
 * per-request hipMalloc).  Requests are admitted in arrival order; an idle
 * slot takes the oldest one and, when it is small, the small ones queued
 * behind it, which then share their launches (nsx_server_set_sharing).  The
 * work of different slots overlaps on the device: give the process at least
 * slots + 2 hardware queues (environment variable GPU_MAX_HW_QUEUES, read when
 * HIP initialises; HIP's default is 4), or streams share queues and serialise.
 * Each request's inputs are copied into its arena block, all stages run
 * there, and the outputs the caller asked for are copied back before the
 * request completes.  Results are those of nsx_segment_chain on the same
 * inputs, however many requests are in flight or share launches.
 * No reference counterpart: the reference's GPU functions are synchronous and
 * not reentrant (global queue state in global_queue.cu:9-22); the request
 * processing matches the chain in HistologicalEntities.cpp:195,219,246,567. */
typedef struct nsx_server nsx_server;
nsx_status nsx_server_create(int device, size_t hbm_budget_bytes, int slots, nsx_server** server);
/* Queues one request and returns its ticket.  marker/mask: u8 or i32, host or
 * device; labels_out int32 (required), recon_out/fill_out optional (NULL), host
 * or device.  Host buffers must stay valid until nsx_server_wait returns for the
 * ticket.  NSX_ERR_BUDGET when the request alone exceeds the arena. */
nsx_status nsx_server_submit(nsx_server* server, const nsx_img* marker, const nsx_img* mask, int recon_conn,
                             int fill_conn, double thresh, int ccl_conn, int min_size, int max_size,
                             nsx_img* recon_out, nsx_img* fill_out, nsx_img* labels_out, uint64_t* ticket);
/* Blocks until the request completes and releases the ticket.  Returns the
 * request's status; *ncomp = components kept; times_ms (optional, 3 entries) =
 * queueing (submit -> admit), service (admit -> done), total (submit -> done). */
nsx_status nsx_server_wait(nsx_server* server, uint64_t ticket, int32_t* ncomp, double* times_ms);
/* stats (8 entries): completed, in flight, queued, peak in flight, arena bytes,
 * peak arena bytes in use, workspace bytes held by the slot contexts, slots. */
nsx_status nsx_server_stats(nsx_server* server, int64_t* stats);
/* Sharing policy.  An idle slot takes the oldest queued request and, when it
 * has at most max_request_pixels pixels, the requests of at most that size
 * queued behind it, in order, while they fit in the arena, the set holds at
 * most max_set_pixels pixels and max_requests requests (1..64).  Those of one
 * kind (element type, connectivities, threshold, area bounds) share their
 * launches (one reconstruction, fill and labelling over all their tiles).
 * Defaults 16 requests, 2^20 and 2^22 pixels (0 selects a default);
 * max_requests = 1 serves every request on its own.  Results do not depend on
 * it.  No reference counterpart (the reference's queue-throughput entry
 * point, MorphologicOperationsGPU.cpp:447-545, batches reconstructions only
 * and has no admission). */
nsx_status nsx_server_set_sharing(nsx_server* server, int max_requests, int64_t max_request_pixels,
                                  int64_t max_set_pixels);
/* batch stats (7 entries): shared launch sets run, requests served in them,
 * largest set, max_requests, sets whose predicted propagation rounds fell
 * short (finished with a host check; the stages after it ran again),
 * max_request_pixels, max_set_pixels. */
nsx_status nsx_server_batch_stats(nsx_server* server, int64_t* stats);
/* The context of worker slot `slot` (0..slots-1), owned by the server, for
 * kernel timing and tile counts (nsx_ctx_set_timing / nsx_ctx_read_timing /
 * nsx_ctx_read_tile_counts).  Change its timing only while no request is in
 * flight.  NULL for a bad slot. */
nsx_ctx* nsx_server_context(nsx_server* server, int slot);
/* Message of the last failed server call or request: a copy private to the
 * calling thread, valid until that thread calls this function again. */
const char* nsx_server_last_error(nsx_server* server);
/* Finishes every queued request, then releases the workers and the arena. */
void nsx_server_destroy(nsx_server* server);

/* ---- synthetic inputs (host) ---------------------------------------------- */
uint64_t nsx_splitmix64_next(uint64_t* state);
int nsx_gen_disks_u8(uint8_t* mask, int rows, int cols, long pitch, int ndisks, uint64_t seed);
void nsx_gen_marker_u8(const uint8_t* mask, uint8_t* marker, long n, int sub);
void nsx_gen_scale_i32(const uint8_t* in, int32_t* out, long n, int scale);
void nsx_gen_scale_f32(const uint8_t* in, float* out, long n);
int nsx_gen_maze_u8(uint8_t* mask, uint8_t* marker, int rows, int cols, uint64_t seed);

/* The descriptor checks every entry point applies to its images (NULL data,
 * unknown dtype or memory kind, pitch below a row, 2^31-1 pixels or more),
 * without a context or a device: NSX_OK, or NSX_ERR_INVALID with *why (may be
 * NULL) set to a static message.  No reference counterpart (the reference's
 * CV_Assert sites check channels and types inside each call). */
nsx_status nsx_image_check(const nsx_img* img, const char** why);

/* Library version string. */
const char* nsx_version(void);

#ifdef __cplusplus
}
#endif

#endif /* NSCALE_HIP_H_ */
