// api.hip -- extern "C" entry points of libnscale_hip.so (declared in
// include/nscale_hip.h).  Argument checks mirror the reference's CV_Assert
// sites; failures return a status and leave a message in nsx_last_error().
#include "nsx_internal.h"

#include <climits>
#include <cmath>
#include <cstring>
#include <new>

using namespace nsx;

// d_stats: per-run statistics and status words (int64 0..31), then the
// cumulative tile counters (int64 kCumWord..).
constexpr size_t kStatsBytes = 512;
static_assert(sizeof(int64_t) * (nsx_ctx::kCumWord + nsx_ctx::kCumKinds) <= kStatsBytes, "stats words");

nsx_status nsx_ctx::ws(int slot, size_t bytes, void** out) {
    if (slot < 0 || slot >= WS_NSLOTS) return fail(NSX_ERR_INVALID, "bad workspace slot");
    if ((int)bufs.size() < WS_NSLOTS) bufs.resize(WS_NSLOTS);
    DevBuf& b = bufs[slot];
    if (bytes == 0) bytes = 1;
    if (b.bytes >= bytes) { *out = b.ptr; return NSX_OK; }
    // grow by 1/8 to avoid reallocating on slowly growing inputs
    size_t want = bytes + bytes / 8;
    want = (want + 255) & ~(size_t)255;
    if (budget && held - b.bytes + want > budget) {
        want = bytes;
        if (held - b.bytes + want > budget)
            return fail(NSX_ERR_BUDGET, "workspace request exceeds the context's HBM budget");
    }
    if (b.ptr) {
        // Work enqueued by earlier calls (on whatever stream the caller passed)
        // may still read this buffer: wait for the whole device before freeing.
        hipError_t e = hipDeviceSynchronize();
        (void)e;
        hipFree(b.ptr);
        held -= b.bytes;
        b.ptr = nullptr;
        b.bytes = 0;
    }
    void* p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        return fail(NSX_ERR_NOMEM, "hipMalloc of workspace failed");
    }
    b.ptr = p;
    b.bytes = want;
    b.serial = ++ws_serial;
    held += want;
    *out = p;
    return NSX_OK;
}

int nsx_ctx::tic(int kind, hipStream_t s) {
    if (!timing) return -1;
    auto take = [&]() {
        hipEvent_t e;
        if (!ev_pool.empty()) { e = ev_pool.back(); ev_pool.pop_back(); return e; }
        if (hipEventCreate(&e) != hipSuccess) return (hipEvent_t) nullptr;
        return e;
    };
    EvPair p{take(), take(), kind};
    if (!p.a || !p.b) return -1;
    (void)hipEventRecord(p.a, s);
    ev_pending.push_back(p);
    return (int)ev_pending.size() - 1;
}

void nsx_ctx::toc(int handle, hipStream_t s) {
    if (handle < 0 || handle >= (int)ev_pending.size()) return;
    (void)hipEventRecord(ev_pending[handle].b, s);
}

void nsx_ctx::collect_timing() {
    for (auto& p : ev_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            t_ms[p.kind] += ms;
            t_n[p.kind] += 1;
        }
        ev_pool.push_back(p.a);
        ev_pool.push_back(p.b);
    }
    ev_pending.clear();
}

namespace {

const char kVersion[] = "nscale-hip 0.1 (gfx950)";

hipStream_t pick_stream(nsx_ctx* ctx, void* stream) {
    return stream ? (hipStream_t)stream : ctx->stream;
}

// What is wrong with an image descriptor (nullptr: nothing).  Host-side only:
// no device call, so it also runs where there is no GPU.
const char* img_problem(const nsx_img* im) {
    if (!im) return "is NULL";
    if (im->rows < 0 || im->cols < 0) return "negative size";
    if (im->dtype != NSX_U8 && im->dtype != NSX_I32 && im->dtype != NSX_F32) return "unknown dtype";
    if (im->mem != NSX_MEM_DEVICE && im->mem != NSX_MEM_HOST) return "unknown memory kind";
    if (im->rows > 0 && im->cols > 0) {
        if (!im->data) return "data is NULL";
        if (im->pitch_bytes < (int64_t)im->cols * (int64_t)dtype_size(im->dtype)) return "pitch smaller than a row";
        if ((int64_t)im->rows * im->cols >= (int64_t)INT32_MAX) return "more than 2^31-1 pixels";
    }
    return nullptr;
}

nsx_status check_img(nsx_ctx* ctx, const nsx_img* im, const char* name) {
    const char* why = img_problem(im);
    if (why) return ctx->fail(NSX_ERR_INVALID, std::string(name) + (im ? ": " : " ") + why);
    return NSX_OK;
}

bool same_shape(const nsx_img* a, const nsx_img* b) { return a->rows == b->rows && a->cols == b->cols; }

// A 0-pixel image: entry points that report a component count return 0
// without touching the device.
bool is_empty(const nsx_img* im) { return (int64_t)im->rows * (int64_t)im->cols == 0; }

// Device view of an image: device images are used in place; host images are
// staged through a workspace slot (copied in when `copy_in`).
nsx_status to_device(nsx_ctx* ctx, const nsx_img* im, int slot, bool copy_in, hipStream_t s, View* v) {
    v->rows = im->rows;
    v->cols = im->cols;
    v->dtype = im->dtype;
    if (im->mem == NSX_MEM_DEVICE) {
        v->data = (char*)im->data;
        v->pitch = im->pitch_bytes;
        return NSX_OK;
    }
    // staged planes are packed (pitch = row width) so contiguous host images
    // move in one linear copy
    const size_t rowb = (size_t)im->cols * dtype_size(im->dtype);
    const size_t pitch = rowb > 0 ? rowb : 1;
    void* p;
    NSX_TRY(ctx->ws(slot, pitch * (size_t)(im->rows > 0 ? im->rows : 1), &p));
    v->data = (char*)p;
    v->pitch = (int64_t)pitch;
    if (copy_in && im->rows > 0)
        NSX_HIP_TRY(ctx, copy_rows(p, pitch, im->data, im->pitch_bytes, rowb, im->rows, hipMemcpyHostToDevice, s));
    return NSX_OK;
}

nsx_status from_device(nsx_ctx* ctx, const View& v, const nsx_img* im, hipStream_t s) {
    if (im->mem == NSX_MEM_DEVICE) return NSX_OK;
    const size_t rowb = (size_t)im->cols * dtype_size(im->dtype);
    if (im->rows > 0)
        NSX_HIP_TRY(ctx, copy_rows(im->data, im->pitch_bytes, v.data, v.pitch, rowb, im->rows, hipMemcpyDeviceToHost, s));
    NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
    return NSX_OK;
}

bool overlaps(const nsx_img* a, const nsx_img* b) {
    if (!a || !b || !a->data || !b->data || a->mem != b->mem) return false;
    const char* a0 = (const char*)a->data;
    const char* a1 = a0 + (a->rows > 0 ? (a->rows - 1) * a->pitch_bytes + (int64_t)a->cols * dtype_size(a->dtype) : 0);
    const char* b0 = (const char*)b->data;
    const char* b1 = b0 + (b->rows > 0 ? (b->rows - 1) * b->pitch_bytes + (int64_t)b->cols * dtype_size(b->dtype) : 0);
    return a0 < b1 && b0 < a1;
}

nsx_status read_count(nsx_ctx* ctx, const int32_t* dev, int32_t* host, hipStream_t s) {
    NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned + 8, dev, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
    if (host) *host = ctx->h_pinned[8];
    return NSX_OK;
}

nsx_status finish_recon(nsx_ctx* ctx, const View& v, const nsx_img* out, hipStream_t s) {
    return from_device(ctx, v, out, s);
}

// imfill with seeds and bwselect share their checks and staging: u8 image and
// seeds of one size, u8 output that aliases neither.
nsx_status seeded_setup(nsx_ctx* ctx, const nsx_img* img, const nsx_img* seeds, const nsx_img* out,
                               const char* what) {
    NSX_TRY(check_img(ctx, img, "image"));
    NSX_TRY(check_img(ctx, seeds, "seeds"));
    NSX_TRY(check_img(ctx, out, "out"));
    if (img->dtype != NSX_U8 || seeds->dtype != NSX_U8 || out->dtype != NSX_U8)
        return ctx->fail(NSX_ERR_UNSUPPORTED, std::string(what) + " is instantiated for uchar only (seeds: CV_8U)");
    if (!same_shape(img, seeds) || !same_shape(img, out))
        return ctx->fail(NSX_ERR_INVALID, "image, seeds and out must have the same size");
    if (overlaps(out, img) || overlaps(out, seeds)) return ctx->fail(NSX_ERR_INVALID, "out aliases an input");
    return NSX_OK;
}

// The constant every int32 imfillHoles result takes, except the gray fill
// under saturating negation: INT_MIN under wrapping negation, INT_MAX (binary
// only) under saturation.
int32_t i32_fill_const(const nsx_ctx* ctx) {
    return ctx->i32_invert == NSX_I32_INVERT_WRAP ? INT_MIN : INT_MAX;
}

nsx_status scratch_like(nsx_ctx* ctx, int slot, const View& like, View* v) {
    void* p;
    const size_t pitch = like.cols > 0 ? (size_t)like.cols * dtype_size(like.dtype) : 1;
    NSX_TRY(ctx->ws(slot, pitch * (size_t)(like.rows > 0 ? like.rows : 1), &p));
    *v = View{(char*)p, (int64_t)pitch, like.rows, like.cols, like.dtype};
    return NSX_OK;
}

}  // namespace

extern "C" {

const char* nsx_version(void) { return kVersion; }

nsx_status nsx_image_check(const nsx_img* img, const char** why) {
    const char* w = img_problem(img);
    if (why) *why = w ? w : "";
    return w ? NSX_ERR_INVALID : NSX_OK;
}

nsx_status nsx_ctx_create(nsx_ctx** out, int device, size_t hbm_budget_bytes) {
    if (!out) return NSX_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return NSX_ERR_HIP;
    }
    if (device < 0 || device >= ndev) return NSX_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return NSX_ERR_HIP;
    nsx_ctx* c = new (std::nothrow) nsx_ctx();
    if (!c) return NSX_ERR_NOMEM;
    c->device = device;
    c->budget = hbm_budget_bytes;
    c->bufs.resize(WS_NSLOTS);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        c->cu_count = cus;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void**)&c->d_stats, kStatsBytes) != hipSuccess ||
        hipMalloc((void**)&c->d_round_log, 2 * sizeof(int32_t) * nsx_ctx::kRoundLog) != hipSuccess ||
        hipHostMalloc((void**)&c->h_pinned, 64 * sizeof(int32_t)) != hipSuccess) {
        nsx_ctx_destroy(c);
        return NSX_ERR_HIP;
    }
    hipMemset(c->d_stats, 0, kStatsBytes);
    hipMemset(c->d_round_log, 0, 2 * sizeof(int32_t) * nsx_ctx::kRoundLog);
    c->held = kStatsBytes + 2 * sizeof(int32_t) * nsx_ctx::kRoundLog;
    // the allocation-free entry point's scratch array: its device address,
    // once per device, with the device current
    if (recon_noalloc_init(device) != NSX_OK) {
        nsx_ctx_destroy(c);
        return NSX_ERR_HIP;
    }
    *out = c;
    return NSX_OK;
}

void nsx_ctx_destroy(nsx_ctx* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    for (auto& b : ctx->bufs)
        if (b.ptr) hipFree(b.ptr);
    if (ctx->d_stats) hipFree(ctx->d_stats);
    if (ctx->d_round_log) hipFree(ctx->d_round_log);
    if (ctx->h_pinned) hipHostFree(ctx->h_pinned);
    ctx->collect_timing();
    for (hipEvent_t e : ctx->ev_pool) hipEventDestroy(e);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* nsx_last_error(const nsx_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

nsx_status nsx_ctx_synchronize(nsx_ctx* ctx) {
    if (!ctx) return NSX_ERR_INVALID;
    NSX_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return NSX_OK;
}

size_t nsx_ctx_workspace_bytes(const nsx_ctx* ctx) { return ctx ? ctx->held : 0; }

nsx_status nsx_ctx_set_tuning(nsx_ctx* ctx, int param, int value) {
    if (!ctx) return NSX_ERR_INVALID;
    if (param == NSX_TUNE_VISIT_SWEEPS) {
        if (value < 1 || value > 64) return ctx->fail(NSX_ERR_INVALID, "visit sweeps must be in 1..64");
        ctx->visit_sweeps = value;
        return NSX_OK;
    }
    if (param == NSX_TUNE_LEVELSET_LEVELS) {
        if (value < 0 || value > 255) return ctx->fail(NSX_ERR_INVALID, "level-set levels must be in 0..255");
        ctx->levelset_levels = value;
        return NSX_OK;
    }
    if (param == NSX_TUNE_FILL_SCANS) {
        if (value < 0 || value > 8) return ctx->fail(NSX_ERR_INVALID, "fill scans must be in 0..8");
        ctx->fill_scans = value;
        return NSX_OK;
    }
    if (param == NSX_TUNE_DRAIN) {
        if (value != 0 && value != 1) return ctx->fail(NSX_ERR_INVALID, "drain must be 0 or 1");
        ctx->drain = value;
        return NSX_OK;
    }
    if (param == NSX_TUNE_DT_SOLVE) {
        if (value < 0 || value > 8) return ctx->fail(NSX_ERR_INVALID, "distanceTransform solve passes must be in 0..8");
        ctx->dt_solve = value;
        return NSX_OK;
    }
    if (param == NSX_TUNE_ENGINE) {
        if (value != NSX_ENGINE_ROUNDS && value != NSX_ENGINE_TREE && value != NSX_ENGINE_AUTO &&
            value != NSX_ENGINE_PLANES)
            return ctx->fail(NSX_ERR_INVALID,
                             "engine must be NSX_ENGINE_ROUNDS, NSX_ENGINE_TREE, NSX_ENGINE_PLANES or NSX_ENGINE_AUTO");
        ctx->engine = value;
        return NSX_OK;
    }
    return ctx->fail(NSX_ERR_INVALID, "unknown tuning parameter");
}

nsx_status nsx_ctx_get_tuning(const nsx_ctx* ctx, int param, int* value) {
    if (!ctx || !value) return NSX_ERR_INVALID;
    if (param == NSX_TUNE_VISIT_SWEEPS) {
        *value = ctx->visit_sweeps;
        return NSX_OK;
    }
    if (param == NSX_TUNE_LEVELSET_LEVELS) {
        *value = ctx->levelset_levels;
        return NSX_OK;
    }
    if (param == NSX_TUNE_ENGINE) {
        *value = ctx->engine;
        return NSX_OK;
    }
    if (param == NSX_TUNE_FILL_SCANS) {
        *value = ctx->fill_scans;
        return NSX_OK;
    }
    if (param == NSX_TUNE_DRAIN) {
        *value = ctx->drain;
        return NSX_OK;
    }
    if (param == NSX_TUNE_DT_SOLVE) {
        *value = ctx->dt_solve;
        return NSX_OK;
    }
    return NSX_ERR_INVALID;
}

nsx_status nsx_ctx_drain_failures(const nsx_ctx* ctx, int64_t* n) {
    if (!ctx || !n) return NSX_ERR_INVALID;
    *n = ctx->drain_failures;
    return NSX_OK;
}

nsx_status nsx_ctx_dt_resumes(const nsx_ctx* ctx, int64_t* n) {
    if (!ctx || !n) return NSX_ERR_INVALID;
    *n = ctx->dtq_resumes;
    return NSX_OK;
}

nsx_status nsx_ctx_set_option(nsx_ctx* ctx, int option, int value) {
    if (!ctx) return NSX_ERR_INVALID;
    if (option == NSX_OPT_I32_INVERT) {
        if (value != NSX_I32_INVERT_WRAP && value != NSX_I32_INVERT_SATURATE)
            return ctx->fail(NSX_ERR_INVALID, "NSX_OPT_I32_INVERT takes NSX_I32_INVERT_WRAP or NSX_I32_INVERT_SATURATE");
        ctx->i32_invert = value;
        return NSX_OK;
    }
    if (option == NSX_OPT_DT_ORDER) {
        if (value != NSX_DT_ORDER_QUEUE && value != NSX_DT_ORDER_TILES)
            return ctx->fail(NSX_ERR_INVALID, "NSX_OPT_DT_ORDER takes NSX_DT_ORDER_QUEUE or NSX_DT_ORDER_TILES");
        ctx->dt_order = value;
        return NSX_OK;
    }
    return ctx->fail(NSX_ERR_INVALID, "unknown option");
}

nsx_status nsx_ctx_get_option(const nsx_ctx* ctx, int option, int* value) {
    if (!ctx || !value) return NSX_ERR_INVALID;
    if (option == NSX_OPT_I32_INVERT) {
        *value = ctx->i32_invert;
        return NSX_OK;
    }
    if (option == NSX_OPT_DT_ORDER) {
        *value = ctx->dt_order;
        return NSX_OK;
    }
    return NSX_ERR_INVALID;
}

nsx_status nsx_ctx_set_timing(nsx_ctx* ctx, int enable) {
    if (!ctx) return NSX_ERR_INVALID;
    NSX_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->collect_timing();
    for (int k = 0; k < TK_NKINDS; ++k) { ctx->t_ms[k] = 0.0; ctx->t_n[k] = 0; }
    for (int k = 0; k < nsx_ctx::kCumKinds; ++k) ctx->first_tiles[k] = 0;
    NSX_HIP_TRY(ctx, hipMemset(ctx->d_stats + nsx_ctx::kCumWord, 0, sizeof(int64_t) * nsx_ctx::kCumKinds));
    ctx->timing = enable != 0;
    return NSX_OK;
}

nsx_status nsx_ctx_read_tile_counts(nsx_ctx* ctx, int64_t* first_pass, int64_t* rounds, int n) {
    if (!ctx || n < 0) return NSX_ERR_INVALID;
    int64_t h[nsx_ctx::kCumKinds];
    NSX_HIP_TRY(ctx, hipDeviceSynchronize());
    NSX_HIP_TRY(ctx, hipMemcpy(h, ctx->d_stats + nsx_ctx::kCumWord, sizeof(h), hipMemcpyDeviceToHost));
    for (int k = 0; k < n && k < nsx_ctx::kCumKinds; ++k) {
        if (first_pass) first_pass[k] = ctx->first_tiles[k];
        if (rounds) rounds[k] = h[k];
    }
    return NSX_OK;
}

nsx_status nsx_ctx_read_timing(nsx_ctx* ctx, double* ms, int64_t* launches, int nkinds) {
    if (!ctx || nkinds < 0) return NSX_ERR_INVALID;
    NSX_HIP_TRY(ctx, hipDeviceSynchronize());
    ctx->collect_timing();
    for (int k = 0; k < nkinds && k < TK_NKINDS; ++k) {
        if (ms) ms[k] = ctx->t_ms[k];
        if (launches) launches[k] = ctx->t_n[k];
    }
    return NSX_OK;
}

nsx_status nsx_last_recon_stats(nsx_ctx* ctx, int stage, int64_t stats[4]) {
    if (!ctx || !stats || stage < 0 || stage > 1) return NSX_ERR_INVALID;
    int64_t h[4];
    NSX_HIP_TRY(ctx, hipDeviceSynchronize());
    NSX_HIP_TRY(ctx, hipMemcpy(h, ctx->d_stats + 8 * stage, sizeof(h), hipMemcpyDeviceToHost));
    stats[0] = h[0];
    stats[1] = h[1];
    stats[2] = h[2];
    stats[3] = h[3];
    return NSX_OK;
}

nsx_status nsx_last_round_log(nsx_ctx* ctx, int stage, int32_t* counts, int n, int* rounds) {
    if (!ctx || !counts || !rounds || n < 0 || stage < 0 || stage > 1) return NSX_ERR_INVALID;
    int64_t h[4];
    NSX_HIP_TRY(ctx, hipDeviceSynchronize());
    NSX_HIP_TRY(ctx, hipMemcpy(h, ctx->d_stats + 8 * stage, sizeof(h), hipMemcpyDeviceToHost));
    const int r = (int)std::min<int64_t>(h[0] + 1, nsx_ctx::kRoundLog);
    *rounds = r;
    const int k = std::min(n, r);
    if (k > 0)
        NSX_HIP_TRY(ctx, hipMemcpy(counts, ctx->d_round_log + (size_t)stage * nsx_ctx::kRoundLog,
                                   sizeof(int32_t) * (size_t)k, hipMemcpyDeviceToHost));
    return NSX_OK;
}

nsx_status nsx_imreconstruct(nsx_ctx* ctx, const nsx_img* marker, const nsx_img* mask, nsx_img* out, int conn,
                             void* stream) {
    if (!ctx) return NSX_ERR_INVALID;
    NSX_TRY(check_img(ctx, marker, "marker"));
    NSX_TRY(check_img(ctx, mask, "mask"));
    NSX_TRY(check_img(ctx, out, "out"));
    if (!same_shape(marker, mask) || !same_shape(mask, out))
        return ctx->fail(NSX_ERR_INVALID, "marker, mask and out must have the same size");
    if (marker->dtype != mask->dtype || out->dtype != mask->dtype)
        return ctx->fail(NSX_ERR_INVALID, "marker, mask and out must have the same type");
    if (overlaps(out, marker) || overlaps(out, mask)) return ctx->fail(NSX_ERR_INVALID, "out aliases an input");
    hipStream_t s = pick_stream(ctx, stream);
    View vm, vk, vo;
    NSX_TRY(to_device(ctx, marker, WS_STAGE_A, true, s, &vm));
    NSX_TRY(to_device(ctx, mask, WS_STAGE_B, true, s, &vk));
    NSX_TRY(to_device(ctx, out, WS_STAGE_C, false, s, &vo));
    NSX_TRY(recon_run(ctx, RM_RECON, &vm, vk, vo, conn == 8 ? 8 : 4, s));
    return finish_recon(ctx, vo, out, s);
}

nsx_status nsx_imreconstruct_noalloc(nsx_ctx* ctx, const nsx_img* marker, const nsx_img* mask, nsx_img* out,
                                     int conn, void* stream) {
    if (!ctx) return NSX_ERR_INVALID;
    NSX_TRY(check_img(ctx, marker, "marker"));
    NSX_TRY(check_img(ctx, mask, "mask"));
    NSX_TRY(check_img(ctx, out, "out"));
    if (!same_shape(marker, mask) || !same_shape(mask, out))
        return ctx->fail(NSX_ERR_INVALID, "marker, mask and out must have the same size");
    if (marker->dtype != mask->dtype || out->dtype != mask->dtype)
        return ctx->fail(NSX_ERR_INVALID, "marker, mask and out must have the same type");
    if (marker->mem != NSX_MEM_DEVICE || mask->mem != NSX_MEM_DEVICE || out->mem != NSX_MEM_DEVICE)
        return ctx->fail(NSX_ERR_INVALID, "nsx_imreconstruct_noalloc takes device images only");
    if (overlaps(out, marker) || overlaps(out, mask)) return ctx->fail(NSX_ERR_INVALID, "out aliases an input");
    const View vm{(char*)marker->data, marker->pitch_bytes, marker->rows, marker->cols, marker->dtype};
    const View vk{(char*)mask->data, mask->pitch_bytes, mask->rows, mask->cols, mask->dtype};
    const View vo{(char*)out->data, out->pitch_bytes, out->rows, out->cols, out->dtype};
    return recon_run_noalloc(ctx, vm, vk, vo, conn == 8 ? 8 : 4, pick_stream(ctx, stream));
}

nsx_status nsx_imreconstruct_binary(nsx_ctx* ctx, const nsx_img* marker, const nsx_img* mask, nsx_img* out,
                                    int conn, void* stream) {
    if (!ctx) return NSX_ERR_INVALID;
    NSX_TRY(check_img(ctx, marker, "marker"));
    NSX_TRY(check_img(ctx, mask, "mask"));
    NSX_TRY(check_img(ctx, out, "out"));
    if (!same_shape(marker, mask) || !same_shape(mask, out))
        return ctx->fail(NSX_ERR_INVALID, "marker, mask and out must have the same size");
    if (marker->dtype != mask->dtype || out->dtype != mask->dtype)
        return ctx->fail(NSX_ERR_INVALID, "marker, mask and out must have the same type");
    if (mask->dtype == NSX_F32)
        return ctx->fail(NSX_ERR_UNSUPPORTED, "imreconstructBinary is instantiated for uchar and int only");
    if (overlaps(out, marker) || overlaps(out, mask)) return ctx->fail(NSX_ERR_INVALID, "out aliases an input");
    hipStream_t s = pick_stream(ctx, stream);
    View vm, vk, vo;
    NSX_TRY(to_device(ctx, marker, WS_STAGE_A, true, s, &vm));
    NSX_TRY(to_device(ctx, mask, WS_STAGE_B, true, s, &vk));
    NSX_TRY(to_device(ctx, out, WS_STAGE_C, false, s, &vo));
    // the reached pixels are the components of the nonzero mask pixels that
    // hold a nonzero marker pixel: one labelling finds them
    NSX_TRY(binrecon_run(ctx, vm, vk, vo, conn == 8 ? 8 : 4, s));
    return finish_recon(ctx, vo, out, s);
}

nsx_status nsx_fill_holes(nsx_ctx* ctx, const nsx_img* in, nsx_img* out, int binary, int conn, void* stream) {
    if (!ctx) return NSX_ERR_INVALID;
    NSX_TRY(check_img(ctx, in, "in"));
    NSX_TRY(check_img(ctx, out, "out"));
    if (!same_shape(in, out) || in->dtype != out->dtype)
        return ctx->fail(NSX_ERR_INVALID, "in and out must have the same size and type");
    if (in->dtype == NSX_F32) return ctx->fail(NSX_ERR_UNSUPPORTED, "imfillHoles is instantiated for uchar and int only");
    if (overlaps(out, in)) return ctx->fail(NSX_ERR_INVALID, "out aliases in");
    hipStream_t s = pick_stream(ctx, stream);
    View vi, vo;
    NSX_TRY(to_device(ctx, in, WS_STAGE_B, true, s, &vi));
    NSX_TRY(to_device(ctx, out, WS_STAGE_C, false, s, &vo));
    if (in->dtype == NSX_I32 && (binary || ctx->i32_invert == NSX_I32_INVERT_WRAP)) {
        // int32 closed forms (MorphologicOperations.cpp:1889-1925): with
        // wrapping negation the inverted mask's frame is INT_MIN, the
        // reconstruction stays INT_MIN everywhere and so does its inverse.
        // With saturation, binary: the marker's interior min<int> is nonzero,
        // imreconstructBinary keeps it everywhere, and its inverse is max<int>.
        NSX_TRY(fill_const_i32(ctx, vo, i32_fill_const(ctx), s));
    } else {
        NSX_TRY(recon_run(ctx, binary ? RM_FILLBIN : RM_FILL, nullptr, vi, vo, conn == 8 ? 8 : 4, s));
    }
    return finish_recon(ctx, vo, out, s);
}

nsx_status nsx_imfill(nsx_ctx* ctx, const nsx_img* image, const nsx_img* seeds, nsx_img* out, int binary, int conn,
                      void* stream) {
    if (!ctx) return NSX_ERR_INVALID;
    NSX_TRY(seeded_setup(ctx, image, seeds, out, "imfill"));
    if (is_empty(image)) return NSX_OK;
    hipStream_t s = pick_stream(ctx, stream);
    View vi, vs, vo, vm, vk;
    NSX_TRY(to_device(ctx, image, WS_STAGE_B, true, s, &vi));
    NSX_TRY(to_device(ctx, seeds, WS_STAGE_A, true, s, &vs));
    NSX_TRY(to_device(ctx, out, WS_STAGE_C, false, s, &vo));
    NSX_TRY(scratch_like(ctx, WS_CHAIN_RECON, vi, &vm));
    NSX_TRY(scratch_like(ctx, WS_CHAIN_FILL, vi, &vk));
    NSX_TRY(seeded_marker_run(ctx, vi, vs, vm, &vk, s));
    NSX_TRY(recon_run(ctx, binary ? RM_BINARY : RM_RECON, &vm, vk, vo, conn == 8 ? 8 : 4, s));
    NSX_TRY(bitwise_u8_run(ctx, vi, vo, false, s));
    return finish_recon(ctx, vo, out, s);
}

nsx_status nsx_imhmin(nsx_ctx* ctx, const nsx_img* image, double h, nsx_img* out, int conn, void* stream) {
    if (!ctx) return NSX_ERR_INVALID;
    NSX_TRY(check_img(ctx, image, "image"));
    NSX_TRY(check_img(ctx, out, "out"));
    if (image->dtype != NSX_U8 && image->dtype != NSX_F32)
        return ctx->fail(NSX_ERR_UNSUPPORTED, "imhmin is instantiated for uchar and float only");
    if (!same_shape(image, out) || image->dtype != out->dtype)
        return ctx->fail(NSX_ERR_INVALID, "image and out must have the same size and type");
    if (image->dtype == NSX_U8 && !(h >= 0.0 && h <= 255.0 && h == std::floor(h)))
        return ctx->fail(NSX_ERR_INVALID, "uchar imhmin takes an integer h in 0..255");
    if (overlaps(out, image)) return ctx->fail(NSX_ERR_INVALID, "out aliases image");
    if (is_empty(image)) return NSX_OK;
    hipStream_t s = pick_stream(ctx, stream);
    View vi, vo, vk, vm;
    NSX_TRY(to_device(ctx, image, WS_STAGE_B, true, s, &vi));
    NSX_TRY(to_device(ctx, out, WS_STAGE_C, false, s, &vo));
    NSX_TRY(scratch_like(ctx, WS_CHAIN_FILL, vi, &vk));
    NSX_TRY(scratch_like(ctx, WS_CHAIN_RECON, vi, &vm));
    NSX_TRY(hmin_prep_run(ctx, vi, h, vk, vm, s));
    NSX_TRY(recon_run(ctx, RM_RECON, &vm, vk, vo, conn == 8 ? 8 : 4, s));
    NSX_TRY(invert_run(ctx, vo, s));
    return finish_recon(ctx, vo, out, s);
}

static nsx_status local_extrema(nsx_ctx* ctx, const nsx_img* image, nsx_img* out, int conn, bool minima,
                                void* stream) {
    if (!ctx) return NSX_ERR_INVALID;
    NSX_TRY(check_img(ctx, image, "image"));
    NSX_TRY(check_img(ctx, out, "out"));
    if (image->dtype != NSX_U8 && image->dtype != NSX_F32)
        return ctx->fail(NSX_ERR_UNSUPPORTED, "localMaxima/localMinima are instantiated for uchar and float only");
    if (out->dtype != NSX_U8) return ctx->fail(NSX_ERR_INVALID, "the extrema image is u8");
    if (!same_shape(image, out)) return ctx->fail(NSX_ERR_INVALID, "image and out must have the same size");
    if (overlaps(out, image)) return ctx->fail(NSX_ERR_INVALID, "out aliases image");
    if (is_empty(image)) return NSX_OK;
    hipStream_t s = pick_stream(ctx, stream);
    View vi, vo;
    NSX_TRY(to_device(ctx, image, WS_STAGE_B, true, s, &vi));
    NSX_TRY(to_device(ctx, out, WS_STAGE_C, false, s, &vo));
    NSX_TRY(local_extrema_run(ctx, vi, vo, conn == 8 ? 8 : 4, minima, s));
    return from_device(ctx, vo, out, s);
}

nsx_status nsx_local_maxima(nsx_ctx* ctx, const nsx_img* image, nsx_img* out, int conn, void* stream) {
    return local_extrema(ctx, image, out, conn, false, stream);
}

nsx_status nsx_local_minima(nsx_ctx* ctx, const nsx_img* image, nsx_img* out, int conn, void* stream) {
    return local_extrema(ctx, image, out, conn, true, stream);
}

nsx_status nsx_distance_transform(nsx_ctx* ctx, const nsx_img* mask, nsx_img* out, int calc_dist, void* stream) {
    if (!ctx) return NSX_ERR_INVALID;
    NSX_TRY(check_img(ctx, mask, "mask"));
    NSX_TRY(check_img(ctx, out, "out"));
    if (mask->dtype != NSX_U8) return ctx->fail(NSX_ERR_INVALID, "distanceTransform takes a uchar mask");
    if (out->dtype != (calc_dist ? NSX_F32 : NSX_I32))
        return ctx->fail(NSX_ERR_INVALID, "out is float32 with calc_dist, int32 without");
    if (!same_shape(mask, out)) return ctx->fail(NSX_ERR_INVALID, "mask and out must have the same size");
    if (overlaps(out, mask)) return ctx->fail(NSX_ERR_INVALID, "out aliases mask");
    if (is_empty(mask)) return NSX_OK;
    hipStream_t s = pick_stream(ctx, stream);
    View vm, vo;
    NSX_TRY(to_device(ctx, mask, WS_STAGE_B, true, s, &vm));
    NSX_TRY(to_device(ctx, out, WS_STAGE_C, false, s, &vo));
    NSX_TRY(dist_run(ctx, vm, vo, calc_dist != 0, s));
    return from_device(ctx, vo, out, s);
}

nsx_status nsx_bounding_box(nsx_ctx* ctx, const nsx_img* labels, int bgval, int32_t* boxes, int capacity,
                            int32_t* count, void* stream) {
    if (!ctx || !count) return NSX_ERR_INVALID;
    NSX_TRY(check_img(ctx, labels, "labels"));
    if (labels->dtype != NSX_I32) return ctx->fail(NSX_ERR_INVALID, "labels must be int32");
    if (boxes && capacity < 0) return ctx->fail(NSX_ERR_INVALID, "negative capacity");
    *count = 0;
    if (is_empty(labels)) return NSX_OK;
    hipStream_t s = pick_stream(ctx, stream);
    View vl;
    NSX_TRY(to_device(ctx, labels, WS_STAGE_B, true, s, &vl));
    return bbox_run(ctx, vl, bgval, boxes, capacity, count, s);
}

nsx_status nsx_bwselect(nsx_ctx* ctx, const nsx_img* binary_image, const nsx_img* seeds, nsx_img* out, int conn,
                        void* stream) {
    if (!ctx) return NSX_ERR_INVALID;
    NSX_TRY(seeded_setup(ctx, binary_image, seeds, out, "bwselect"));
    if (is_empty(binary_image)) return NSX_OK;
    hipStream_t s = pick_stream(ctx, stream);
    View vi, vs, vo;
    NSX_TRY(to_device(ctx, binary_image, WS_STAGE_B, true, s, &vi));
    NSX_TRY(to_device(ctx, seeds, WS_STAGE_A, true, s, &vs));
    NSX_TRY(to_device(ctx, out, WS_STAGE_C, false, s, &vo));
    // the reconstruction of the seeded marker reaches exactly the components
    // of the nonzero pixels that hold a seed: one labelling finds them
    NSX_TRY(bwselect_run(ctx, vi, vs, vo, conn == 8 ? 8 : 4, s));
    return finish_recon(ctx, vo, out, s);
}

nsx_status nsx_bwlabel(nsx_ctx* ctx, const nsx_img* bin, nsx_img* labels, int conn, int relabel, int32_t* ncomp,
                       void* stream) {
    if (!ctx) return NSX_ERR_INVALID;
    NSX_TRY(check_img(ctx, bin, "bin"));
    NSX_TRY(check_img(ctx, labels, "labels"));
    if (bin->dtype != NSX_U8) return ctx->fail(NSX_ERR_INVALID, "bwlabel2 requires CV_8U input");
    if (labels->dtype != NSX_I32) return ctx->fail(NSX_ERR_INVALID, "labels must be int32");
    if (!same_shape(bin, labels)) return ctx->fail(NSX_ERR_INVALID, "bin and labels must have the same size");
    if (overlaps(bin, labels)) return ctx->fail(NSX_ERR_INVALID, "labels aliases bin");
    if (is_empty(bin)) {
        if (ncomp) *ncomp = 0;
        return NSX_OK;
    }
    hipStream_t s = pick_stream(ctx, stream);
    View vi, vo;
    NSX_TRY(to_device(ctx, bin, WS_STAGE_B, true, s, &vi));
    NSX_TRY(to_device(ctx, labels, WS_STAGE_C, false, s, &vo));
    CclParams p;
    p.conn = conn == 8 ? 8 : 4;
    p.out_mode = relabel ? CO_CANON : CO_ROOT;
    int32_t* dcount = (int32_t*)ctx->d_stats + 12;
    NSX_TRY(ccl_run(ctx, vi, vo, p, dcount, s));
    NSX_TRY(from_device(ctx, vo, labels, s));
    int32_t n = 0;
    NSX_TRY(read_count(ctx, dcount, &n, s));
    if (ncomp) *ncomp = relabel ? n : 0;
    return NSX_OK;
}

nsx_status nsx_bwareaopen(nsx_ctx* ctx, const nsx_img* in, int labeled, int flatten, int min_size, int max_size,
                          int conn, nsx_img* out, int32_t* count, void* stream) {
    if (!ctx) return NSX_ERR_INVALID;
    NSX_TRY(check_img(ctx, in, "in"));
    NSX_TRY(check_img(ctx, out, "out"));
    if (!labeled && in->dtype != NSX_U8) return ctx->fail(NSX_ERR_INVALID, "bwareaopen2 requires CV_8U input");
    if (labeled && in->dtype != NSX_I32) return ctx->fail(NSX_ERR_INVALID, "labeled bwareaopen2 requires CV_32S input");
    if (out->dtype != (flatten ? NSX_U8 : NSX_I32))
        return ctx->fail(NSX_ERR_INVALID, flatten ? "flattened output must be u8" : "output must be int32");
    if (!same_shape(in, out)) return ctx->fail(NSX_ERR_INVALID, "in and out must have the same size");
    if (overlaps(in, out)) return ctx->fail(NSX_ERR_INVALID, "out aliases in");
    if (is_empty(in)) {
        if (count) *count = 0;
        return NSX_OK;
    }
    hipStream_t s = pick_stream(ctx, stream);
    View vi, vo;
    NSX_TRY(to_device(ctx, in, WS_STAGE_B, true, s, &vi));
    NSX_TRY(to_device(ctx, out, WS_STAGE_C, false, s, &vo));
    int32_t* dcount = (int32_t*)ctx->d_stats + 12;
    if (!labeled) {
        CclParams p;
        p.conn = conn == 8 ? 8 : 4;
        p.out_mode = flatten ? CO_FLAT_U8 : CO_ROOT_FILTERED;
        p.use_area = true;
        p.min_size = min_size;
        p.max_size = max_size;
        NSX_TRY(ccl_run(ctx, vi, vo, p, dcount, s));
    } else {
        NSX_TRY(ccl_area_labeled(ctx, vi, vo, min_size, max_size, flatten != 0, dcount, s));
    }
    NSX_TRY(from_device(ctx, vo, out, s));
    int32_t n = 0;
    NSX_TRY(read_count(ctx, dcount, &n, s));
    if (count) *count = n;
    return NSX_OK;
}

nsx_status nsx_relabel(nsx_ctx* ctx, nsx_img* labels, int bgval, int32_t* ncomp, void* stream) {
    if (!ctx) return NSX_ERR_INVALID;
    NSX_TRY(check_img(ctx, labels, "labels"));
    if (labels->dtype != NSX_I32) return ctx->fail(NSX_ERR_INVALID, "labels must be int32");
    if (is_empty(labels)) {
        if (ncomp) *ncomp = 0;
        return NSX_OK;
    }
    hipStream_t s = pick_stream(ctx, stream);
    View v;
    NSX_TRY(to_device(ctx, labels, WS_STAGE_C, true, s, &v));
    int32_t* dcount = (int32_t*)ctx->d_stats + 12;
    NSX_TRY(ccl_relabel(ctx, v, bgval, dcount, s));
    NSX_TRY(from_device(ctx, v, labels, s));
    int32_t n = 0;
    NSX_TRY(read_count(ctx, dcount, &n, s));
    if (ncomp) *ncomp = n;
    return NSX_OK;
}

nsx_status nsx_threshold_gt(nsx_ctx* ctx, const nsx_img* in, double k, nsx_img* out, void* stream) {
    if (!ctx) return NSX_ERR_INVALID;
    NSX_TRY(check_img(ctx, in, "in"));
    NSX_TRY(check_img(ctx, out, "out"));
    if (out->dtype != NSX_U8) return ctx->fail(NSX_ERR_INVALID, "threshold output must be u8");
    if (!same_shape(in, out)) return ctx->fail(NSX_ERR_INVALID, "in and out must have the same size");
    hipStream_t s = pick_stream(ctx, stream);
    View vi, vo;
    NSX_TRY(to_device(ctx, in, WS_STAGE_B, true, s, &vi));
    NSX_TRY(to_device(ctx, out, WS_STAGE_C, false, s, &vo));
    NSX_TRY(threshold_run(ctx, vi, k, vo, s));
    return from_device(ctx, vo, out, s);
}

nsx_status nsx_recon_label(nsx_ctx* ctx, const nsx_img* marker, const nsx_img* mask, int conn, double thresh,
                           int ccl_conn, nsx_img* recon_out, nsx_img* labels_out, int32_t* ncomp, void* stream) {
    if (!ctx) return NSX_ERR_INVALID;
    NSX_TRY(check_img(ctx, marker, "marker"));
    NSX_TRY(check_img(ctx, mask, "mask"));
    NSX_TRY(check_img(ctx, recon_out, "recon_out"));
    NSX_TRY(check_img(ctx, labels_out, "labels_out"));
    if (!same_shape(marker, mask) || !same_shape(mask, recon_out) || !same_shape(mask, labels_out))
        return ctx->fail(NSX_ERR_INVALID, "all images must have the same size");
    if (marker->dtype != mask->dtype || recon_out->dtype != mask->dtype)
        return ctx->fail(NSX_ERR_INVALID, "marker, mask and recon_out must have the same type");
    if (labels_out->dtype != NSX_I32) return ctx->fail(NSX_ERR_INVALID, "labels must be int32");
    if (overlaps(recon_out, marker) || overlaps(recon_out, mask) || overlaps(labels_out, recon_out))
        return ctx->fail(NSX_ERR_INVALID, "outputs alias inputs");
    if (is_empty(mask)) {
        if (ncomp) *ncomp = 0;
        return NSX_OK;
    }
    hipStream_t s = pick_stream(ctx, stream);
    View vm, vk, vr, vl;
    NSX_TRY(to_device(ctx, marker, WS_STAGE_A, true, s, &vm));
    NSX_TRY(to_device(ctx, mask, WS_STAGE_B, true, s, &vk));
    NSX_TRY(to_device(ctx, recon_out, WS_CHAIN_RECON, false, s, &vr));
    NSX_TRY(to_device(ctx, labels_out, WS_STAGE_C, false, s, &vl));
    int32_t* dcount = (int32_t*)ctx->d_stats + 12;
    const int rconn = conn == 8 ? 8 : 4, lconn = ccl_conn == 8 ? 8 : 4;
    // a u8 reconstruction through its level sets labels {recon > thresh} from
    // its own labelling when both connectivities agree
    LevelLabelReq req;
    req.labels = vl;
    req.thresh = thresh;
    req.ncomp_dev = dcount;
    NSX_TRY(recon_run(ctx, RM_RECON, &vm, vk, vr, rconn, s, rconn == lconn ? &req : nullptr));
    if (!req.done) {
        CclParams p;
        p.conn = lconn;
        p.out_mode = CO_CANON;
        p.threshold = true;
        p.thr = thresh;
        NSX_TRY(ccl_run(ctx, vr, vl, p, dcount, s));
    }
    if (recon_out->mem == NSX_MEM_HOST) NSX_TRY(from_device(ctx, vr, recon_out, s));
    NSX_TRY(from_device(ctx, vl, labels_out, s));
    int32_t n = 0;
    NSX_TRY(read_count(ctx, dcount, &n, s));
    if (ncomp) *ncomp = n;
    return NSX_OK;
}

nsx_status nsx_segment_chain(nsx_ctx* ctx, const nsx_img* marker, const nsx_img* mask, int recon_conn, int fill_conn,
                             double thresh, int ccl_conn, int min_size, int max_size, nsx_img* recon_out,
                             nsx_img* fill_out, nsx_img* labels_out, int32_t* ncomp, void* stream) {
    if (!ctx) return NSX_ERR_INVALID;
    NSX_TRY(check_img(ctx, marker, "marker"));
    NSX_TRY(check_img(ctx, mask, "mask"));
    NSX_TRY(check_img(ctx, labels_out, "labels_out"));
    if (!same_shape(marker, mask) || !same_shape(mask, labels_out))
        return ctx->fail(NSX_ERR_INVALID, "marker, mask and labels must have the same size");
    if (marker->dtype != mask->dtype) return ctx->fail(NSX_ERR_INVALID, "marker and mask must have the same type");
    if (mask->dtype == NSX_F32) return ctx->fail(NSX_ERR_UNSUPPORTED, "imfillHoles is instantiated for uchar and int only");
    if (labels_out->dtype != NSX_I32) return ctx->fail(NSX_ERR_INVALID, "labels must be int32");
    for (const nsx_img* o : {recon_out, fill_out}) {
        if (!o) continue;
        NSX_TRY(check_img(ctx, o, "chain output"));
        if (!same_shape(o, mask) || o->dtype != mask->dtype)
            return ctx->fail(NSX_ERR_INVALID, "chain outputs must match the mask's size and type");
    }
    if (is_empty(mask)) {
        if (ncomp) *ncomp = 0;
        return NSX_OK;
    }
    hipStream_t s = pick_stream(ctx, stream);
    View vm, vk, vr, vf, vl;
    NSX_TRY(to_device(ctx, marker, WS_STAGE_A, true, s, &vm));
    NSX_TRY(to_device(ctx, mask, WS_STAGE_B, true, s, &vk));
    const size_t es = dtype_size(mask->dtype);
    auto inter = [&](nsx_img* o, int slot, View* v) -> nsx_status {
        if (o && o->mem == NSX_MEM_DEVICE) return to_device(ctx, o, slot, false, s, v);
        const size_t pitch = std::max<size_t>(1, (size_t)mask->cols * es);
        void* p;
        NSX_TRY(ctx->ws(slot, pitch * (size_t)(mask->rows > 0 ? mask->rows : 1), &p));
        v->data = (char*)p; v->pitch = (int64_t)pitch; v->rows = mask->rows; v->cols = mask->cols; v->dtype = mask->dtype;
        return NSX_OK;
    };
    NSX_TRY(inter(recon_out, WS_CHAIN_RECON, &vr));
    NSX_TRY(inter(fill_out, WS_CHAIN_FILL, &vf));
    NSX_TRY(to_device(ctx, labels_out, WS_STAGE_C, false, s, &vl));
    CclParams p;
    p.conn = ccl_conn == 8 ? 8 : 4;
    p.out_mode = CO_CANON;
    p.threshold = true;
    p.thr = thresh;
    p.use_area = true;
    p.min_size = min_size;
    p.max_size = max_size;
    // d_stats words: int32 [12] components, int64 [8..11] the fill stage's
    // statistics, int32 [32..39] the status gathered for the host.
    int32_t* dw = (int32_t*)ctx->d_stats;
    const int rconn = recon_conn == 8 ? 8 : 4, fconn = fill_conn == 8 ? 8 : 4;
    if (mask->dtype == NSX_U8 && ctx->engine == NSX_ENGINE_TREE) {
        // NSX_ENGINE_TREE: both reconstructions through the component tree
        // (fixed launches), then the labelling; one synchronisation reads the
        // count and both error words.  A bound hit in either falls back to the
        // rounds below.
        uint32_t *er = nullptr, *ef = nullptr;
        NSX_TRY(mtree_launch(ctx, RM_RECON, &vm, vk, vr, rconn, s, &er, WS_MT_PAR));
        NSX_TRY(mtree_launch(ctx, RM_FILL, nullptr, vr, vf, fconn, s, &ef, WS_MT2_PAR));
        NSX_TRY(ccl_run(ctx, vf, vl, p, dw + 12, s));
        NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned + 12, dw + 12, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned + 56, er, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned + 57, ef, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
        if (ctx->h_pinned[56] == 0 && ctx->h_pinned[57] == 0) {
            const int32_t n = ctx->h_pinned[12];
            if (recon_out && recon_out->mem == NSX_MEM_HOST) NSX_TRY(from_device(ctx, vr, recon_out, s));
            if (fill_out && fill_out->mem == NSX_MEM_HOST) NSX_TRY(from_device(ctx, vf, fill_out, s));
            NSX_TRY(from_device(ctx, vl, labels_out, s));
            if (ncomp) *ncomp = n;
            return NSX_OK;
        }
    }
    if (mask->dtype == NSX_U8 && ctx->engine == NSX_ENGINE_PLANES) {
        // NSX_ENGINE_PLANES: both reconstructions through the bit planes
        // (fixed launches), then the labelling; one synchronisation
        NSX_TRY(recon_bitplane_u8(ctx, RM_RECON, &vm, vk, vr, rconn, s));
        NSX_TRY(recon_bitplane_u8(ctx, RM_FILL, nullptr, vr, vf, fconn, s));
        NSX_TRY(ccl_run(ctx, vf, vl, p, dw + 12, s));
        int32_t n = 0;
        NSX_TRY(read_count(ctx, dw + 12, &n, s));
        if (recon_out && recon_out->mem == NSX_MEM_HOST) NSX_TRY(from_device(ctx, vr, recon_out, s));
        if (fill_out && fill_out->mem == NSX_MEM_HOST) NSX_TRY(from_device(ctx, vf, fill_out, s));
        NSX_TRY(from_device(ctx, vl, labels_out, s));
        if (ncomp) *ncomp = n;
        return NSX_OK;
    }
    if (mask->dtype == NSX_I32 && ctx->i32_invert == NSX_I32_INVERT_WRAP) {
        // int32 fill under wrapping negation is the constant INT_MIN (see
        // nsx_fill_holes); the labelling still runs on it
        NSX_TRY(recon_run(ctx, RM_RECON, &vm, vk, vr, rconn, s));
        NSX_TRY(fill_const_i32(ctx, vf, INT_MIN, s));
        NSX_TRY(ccl_run(ctx, vf, vl, p, dw + 12, s));
        int32_t n = 0;
        NSX_TRY(read_count(ctx, dw + 12, &n, s));
        if (recon_out && recon_out->mem == NSX_MEM_HOST) NSX_TRY(from_device(ctx, vr, recon_out, s));
        if (fill_out && fill_out->mem == NSX_MEM_HOST) NSX_TRY(from_device(ctx, vf, fill_out, s));
        NSX_TRY(from_device(ctx, vl, labels_out, s));
        if (ncomp) *ncomp = n;
        return NSX_OK;
    }
    std::vector<uint8_t> dr(recon_desc_bytes()), df(recon_desc_bytes());
    recon_fill_desc(dr.data(), &vm, vk, vr, 0);
    recon_fill_desc(df.data(), nullptr, vr, vf, 0);
    ReconBatch br{}, bf{};
    br.nimg = bf.nimg = 1;
    br.ntiles = bf.ntiles = recon_tiles(mask->rows, mask->cols);
    br.host_desc = dr.data();
    bf.host_desc = df.data();
    br.ws_base = WS_RECON_STAMP;
    bf.ws_base = WS_FILL_STAMP;
    bf.stats = ctx->d_stats + 8;
    ReconRun rr, rf;
    // a drain stops where NSX_ENGINE_AUTO would switch to the bit planes
    // NSX_ENGINE_AUTO on u8: stages whose last run of this size passed the
    // tree threshold go stage by stage through recon_run, which picks the tree
    const bool auto_u8 = mask->dtype == NSX_U8 && ctx->engine == NSX_ENGINE_AUTO;
    if (auto_u8) br.drain_cap = bf.drain_cap = 256 + 1;
    if (auto_u8 && (recon_predicted_rounds(ctx, RM_RECON, NSX_U8, rconn, br.ntiles, 0) > 256 ||
                    recon_predicted_rounds(ctx, RM_FILL, NSX_U8, fconn, bf.ntiles, 0) > 256)) {
        NSX_TRY(recon_run(ctx, RM_RECON, &vm, vk, vr, rconn, s));
        NSX_TRY(recon_run(ctx, RM_FILL, nullptr, vr, vf, fconn, s, nullptr, ctx->d_stats + 8));
        NSX_TRY(ccl_run(ctx, vf, vl, p, dw + 12, s));
        int32_t n = 0;
        NSX_TRY(read_count(ctx, dw + 12, &n, s));
        if (recon_out && recon_out->mem == NSX_MEM_HOST) NSX_TRY(from_device(ctx, vr, recon_out, s));
        if (fill_out && fill_out->mem == NSX_MEM_HOST) NSX_TRY(from_device(ctx, vf, fill_out, s));
        NSX_TRY(from_device(ctx, vl, labels_out, s));
        if (ncomp) *ncomp = n;
        return NSX_OK;
    }
    // All stages are enqueued back to back; convergence of both propagations
    // is checked once, together with the component count.
    NSX_TRY(recon_launch(ctx, RM_RECON, mask->dtype, rconn, br, s, &rr, nullptr));
    NSX_TRY(recon_launch(ctx, RM_FILL, mask->dtype, fconn, bf, s, &rf, nullptr));
    NSX_TRY(ccl_run(ctx, vf, vl, p, dw + 12, s));
    recon_status_gather(&rr, &rf, dw + 12, ctx->d_stats, ctx->d_stats + 8, dw + 32, s);
    NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned + 40, dw + 32, 8 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
    int32_t status[8];
    std::memcpy(status, ctx->h_pinned + 40, sizeof(status));
    ctx->h_pinned[12] = status[0];
    if (status[1] != 0 || status[2] != 0) {
        // a propagation needed more rounds than predicted: finish it and redo
        // the stages downstream of it
        if (status[1] != 0) {
            if (auto_u8) {
                rr.release();
                rf.release();
                NSX_TRY(recon_run(ctx, RM_RECON, &vm, vk, vr, rconn, s));
                NSX_TRY(recon_run(ctx, RM_FILL, nullptr, vr, vf, fconn, s, nullptr, ctx->d_stats + 8));
            } else {
                NSX_TRY(recon_complete(ctx, &rr, s, nullptr));
                NSX_TRY(recon_launch(ctx, RM_FILL, mask->dtype, fconn, bf, s, &rf, nullptr));
                NSX_TRY(recon_complete(ctx, &rf, s, nullptr));
            }
        } else if (auto_u8) {
            rf.release();
            NSX_TRY(recon_run(ctx, RM_FILL, nullptr, vr, vf, fconn, s, nullptr, ctx->d_stats + 8));
        } else {
            NSX_TRY(recon_complete(ctx, &rf, s, nullptr));
        }
        NSX_TRY(ccl_run(ctx, vf, vl, p, dw + 12, s));
        NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned + 12, dw + 12, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
    } else {
        int64_t wr, wf;   // rounds that had work, for the next call's prediction
        std::memcpy(&wr, status + 4, sizeof(wr));
        std::memcpy(&wf, status + 6, sizeof(wf));
        recon_note_rounds(ctx, &rr, wr);
        recon_note_rounds(ctx, &rf, wf);
    }
    const int32_t n = ctx->h_pinned[12];
    if (recon_out && recon_out->mem == NSX_MEM_HOST) NSX_TRY(from_device(ctx, vr, recon_out, s));
    if (fill_out && fill_out->mem == NSX_MEM_HOST) NSX_TRY(from_device(ctx, vf, fill_out, s));
    NSX_TRY(from_device(ctx, vl, labels_out, s));
    if (ncomp) *ncomp = n;
    return NSX_OK;
}

}  // extern "C"
