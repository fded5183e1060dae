// server.hip -- serving loop for the segmentation chain.
//
// Requests (marker, mask -> recon, fill, labels) arrive on any host thread and
// are admitted in arrival order when a worker slot is idle and the request's
// device buffers fit in a fixed HBM arena.  Each slot is a host thread with its
// own nsx_ctx (workspace, stream, round predictions), so the kernels of all
// admitted requests overlap on the device while each request's stages stay in
// order on its slot's stream.  The arena is one hipMalloc made at creation;
// requests take and return blocks of it through a first-fit free list, so the
// loop itself never allocates device memory.  Everything a request computes
// comes from nsx_segment_chain on its own buffers, so its result does not
// depend on what else is in flight.
#include "nsx_internal.h"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <thread>

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over one device block; offsets are kAlign-aligned.
struct Arena {
    char* base = nullptr;
    size_t size = 0, used = 0, peak = 0;
    std::map<size_t, size_t> free_blocks;   // offset -> length, coalesced

    bool alloc(size_t bytes, size_t* off) {
        bytes = align_up(bytes);
        for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t o = it->first, rest = it->second - bytes;
            free_blocks.erase(it);
            if (rest) free_blocks.emplace(o + bytes, rest);
            used += bytes;
            peak = std::max(peak, used);
            *off = o;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = align_up(bytes);
        used -= bytes;
        auto it = free_blocks.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_blocks.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_blocks.erase(nx);
        }
        if (it != free_blocks.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_blocks.erase(it);
            }
        }
    }
};

size_t elem_size(int dtype) { return dtype == NSX_U8 ? 1 : 4; }

struct Request {
    uint64_t ticket = 0;
    nsx_img marker{}, mask{}, labels{}, recon{}, fill{};
    bool want_recon = false, want_fill = false;
    int recon_conn = 8, fill_conn = 4, ccl_conn = 8, min_size = 0, max_size = 0;
    double thresh = 0;
    size_t bytes = 0, off = 0;
    bool done = false;
    nsx_status status = NSX_OK;
    std::string error;
    int32_t ncomp = 0;
    Clock::time_point t_submit, t_admit, t_done;
};

}  // namespace

struct nsx_server {
    int device = 0;
    std::vector<nsx_ctx*> ctxs;
    std::vector<std::thread> workers;
    Arena arena;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::unordered_map<uint64_t, Request*> reqs;
    uint64_t next_ticket = 1;
    bool stopping = false;
    int inflight = 0, peak_inflight = 0;
    int64_t completed = 0;
    std::string last_error;

    nsx_status fail(nsx_status st, const std::string& msg) {
        std::lock_guard<std::mutex> lk(mu);
        last_error = msg;
        return st;
    }
    void worker(int slot);
    void run(nsx_ctx* ctx, Request* r);
};

namespace {

// Device layout of one request inside its arena block: marker, mask, recon,
// fill (element type of the mask) and labels (int32), each plane packed (pitch
// = row width, so contiguous host images move in one linear copy) and starting
// on a 256-byte boundary.  Inputs or outputs the caller already holds on the
// device are used in place and take no arena space.
struct Layout {
    size_t pitch_t = 0, pitch_l = 0, plane_t = 0, plane_l = 0;
    size_t off_marker = 0, off_mask = 0, off_recon = 0, off_fill = 0, off_labels = 0, total = 0;
};

Layout layout_of(const Request& r) {
    Layout L;
    const size_t rows = (size_t)std::max(1, r.mask.rows);
    L.pitch_t = std::max<size_t>(1, (size_t)r.mask.cols * elem_size(r.mask.dtype));
    L.pitch_l = std::max<size_t>(4, (size_t)r.mask.cols * 4);
    L.plane_t = align_up(L.pitch_t * rows);
    L.plane_l = align_up(L.pitch_l * rows);
    size_t o = 0;
    auto take = [&](bool need, size_t plane, size_t* at) { if (need) { *at = o; o += plane; } };
    take(r.marker.mem == NSX_MEM_HOST, L.plane_t, &L.off_marker);
    take(r.mask.mem == NSX_MEM_HOST, L.plane_t, &L.off_mask);
    take(!(r.want_recon && r.recon.mem == NSX_MEM_DEVICE), L.plane_t, &L.off_recon);
    take(!(r.want_fill && r.fill.mem == NSX_MEM_DEVICE), L.plane_t, &L.off_fill);
    take(r.labels.mem == NSX_MEM_HOST, L.plane_l, &L.off_labels);
    L.total = std::max<size_t>(o, kAlign);
    return L;
}

nsx_img device_img(char* base, size_t off, size_t pitch, const nsx_img& like, int dtype) {
    nsx_img d = like;
    d.data = base + off;
    d.pitch_bytes = (int64_t)pitch;
    d.dtype = dtype;
    d.mem = NSX_MEM_DEVICE;
    return d;
}

nsx_status valid_img(const nsx_img* im) {
    if (!im || (!im->data && (size_t)im->rows * (size_t)im->cols > 0)) return NSX_ERR_INVALID;
    if (im->rows < 0 || im->cols < 0) return NSX_ERR_INVALID;
    if (im->dtype != NSX_U8 && im->dtype != NSX_I32 && im->dtype != NSX_F32) return NSX_ERR_INVALID;
    if (im->mem != NSX_MEM_HOST && im->mem != NSX_MEM_DEVICE) return NSX_ERR_INVALID;
    if (im->pitch_bytes < (int64_t)im->cols * (int64_t)elem_size(im->dtype)) return NSX_ERR_INVALID;
    return NSX_OK;
}

}  // namespace

void nsx_server::run(nsx_ctx* ctx, Request* r) {
    const Layout L = layout_of(*r);
    char* base = arena.base + r->off;
    const int dt = r->mask.dtype;
    const size_t row_t = (size_t)r->mask.cols * elem_size(dt), row_l = (size_t)r->mask.cols * 4;
    const size_t rows = (size_t)r->mask.rows;
    hipStream_t s = ctx->stream;
    nsx_status st = NSX_OK;
    auto hip = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && st == NSX_OK) {
            st = NSX_ERR_HIP;
            r->error = std::string(what) + ": " + hipGetErrorString(e);
        }
    };
    nsx_img dm = r->marker, dk = r->mask, dr, df, dl = r->labels;
    if (r->marker.mem == NSX_MEM_HOST) {
        dm = device_img(base, L.off_marker, L.pitch_t, r->marker, dt);
        hip(nsx::copy_rows(dm.data, L.pitch_t, r->marker.data, r->marker.pitch_bytes, row_t, rows,
                           hipMemcpyHostToDevice, s), "marker upload");
    }
    if (r->mask.mem == NSX_MEM_HOST) {
        dk = device_img(base, L.off_mask, L.pitch_t, r->mask, dt);
        hip(nsx::copy_rows(dk.data, L.pitch_t, r->mask.data, r->mask.pitch_bytes, row_t, rows,
                           hipMemcpyHostToDevice, s), "mask upload");
    }
    dr = (r->want_recon && r->recon.mem == NSX_MEM_DEVICE) ? r->recon
                                                           : device_img(base, L.off_recon, L.pitch_t, r->mask, dt);
    df = (r->want_fill && r->fill.mem == NSX_MEM_DEVICE) ? r->fill
                                                         : device_img(base, L.off_fill, L.pitch_t, r->mask, dt);
    if (r->labels.mem == NSX_MEM_HOST) dl = device_img(base, L.off_labels, L.pitch_l, r->labels, NSX_I32);
    if (st == NSX_OK) {
        int32_t n = 0;
        st = nsx_segment_chain(ctx, &dm, &dk, r->recon_conn, r->fill_conn, r->thresh, r->ccl_conn, r->min_size,
                               r->max_size, &dr, &df, &dl, &n, s);
        if (st != NSX_OK) r->error = nsx_last_error(ctx);
        r->ncomp = n;
    }
    if (st == NSX_OK && rows) {
        if (r->want_recon && r->recon.mem == NSX_MEM_HOST)
            hip(nsx::copy_rows(r->recon.data, r->recon.pitch_bytes, dr.data, L.pitch_t, row_t, rows,
                               hipMemcpyDeviceToHost, s), "recon download");
        if (r->want_fill && r->fill.mem == NSX_MEM_HOST)
            hip(nsx::copy_rows(r->fill.data, r->fill.pitch_bytes, df.data, L.pitch_t, row_t, rows,
                               hipMemcpyDeviceToHost, s), "fill download");
        if (r->labels.mem == NSX_MEM_HOST)
            hip(nsx::copy_rows(r->labels.data, r->labels.pitch_bytes, dl.data, L.pitch_l, row_l, rows,
                               hipMemcpyDeviceToHost, s), "labels download");
    }
    // the arena block is reused as soon as it is released: drain the stream first
    hip(hipStreamSynchronize(s), "request completion");
    r->status = st;
}

void nsx_server::worker(int slot) {
    (void)hipSetDevice(device);
    nsx_ctx* ctx = ctxs[slot];
    for (;;) {
        Request* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu);
            // Admission: the oldest queued request, once its block fits.  Later
            // requests wait behind it, so a large request cannot be starved.
            cv_work.wait(lk, [&] {
                if (queue.empty()) return stopping;
                size_t off;
                if (!arena.alloc(queue.front()->bytes, &off)) return false;
                queue.front()->off = off;
                return true;
            });
            if (queue.empty()) return;
            r = queue.front();
            queue.pop_front();
            r->t_admit = Clock::now();
            peak_inflight = std::max(peak_inflight, ++inflight);
        }
        run(ctx, r);
        {
            std::lock_guard<std::mutex> lk(mu);
            arena.release(r->off, r->bytes);
            r->t_done = Clock::now();
            r->done = true;
            --inflight;
            ++completed;
            if (r->status != NSX_OK) last_error = r->error;
        }
        cv_work.notify_all();
        cv_done.notify_all();
    }
}

extern "C" {

nsx_status nsx_server_create(int device, size_t hbm_budget_bytes, int slots, nsx_server** out) {
    if (!out) return NSX_ERR_INVALID;
    *out = nullptr;
    if (slots < 1 || slots > 32 || hbm_budget_bytes < kAlign) return NSX_ERR_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return NSX_ERR_HIP;
    }
    if (device < 0 || device >= ndev) return NSX_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return NSX_ERR_HIP;
    nsx_server* S = new (std::nothrow) nsx_server();
    if (!S) return NSX_ERR_NOMEM;
    S->device = device;
    S->arena.size = align_up(hbm_budget_bytes);
    if (hipMalloc((void**)&S->arena.base, S->arena.size) != hipSuccess) {
        (void)hipGetLastError();
        delete S;
        return NSX_ERR_NOMEM;
    }
    S->arena.free_blocks.emplace(0, S->arena.size);
    for (int i = 0; i < slots; ++i) {
        nsx_ctx* c = nullptr;
        nsx_status st = nsx_ctx_create(&c, device, 0);
        if (st != NSX_OK) {
            for (nsx_ctx* x : S->ctxs) nsx_ctx_destroy(x);
            hipFree(S->arena.base);
            delete S;
            return st;
        }
        S->ctxs.push_back(c);
    }
    for (int i = 0; i < slots; ++i) S->workers.emplace_back(&nsx_server::worker, S, i);
    *out = S;
    return NSX_OK;
}

nsx_status nsx_server_submit(nsx_server* S, const nsx_img* marker, const nsx_img* mask, int recon_conn,
                             int fill_conn, double thresh, int ccl_conn, int min_size, int max_size,
                             nsx_img* recon_out, nsx_img* fill_out, nsx_img* labels_out, uint64_t* ticket) {
    if (!S || !ticket) return NSX_ERR_INVALID;
    for (const nsx_img* im : {marker, mask, (const nsx_img*)labels_out})
        if (valid_img(im) != NSX_OK) return S->fail(NSX_ERR_INVALID, "request image is NULL or malformed");
    for (const nsx_img* im : {(const nsx_img*)recon_out, (const nsx_img*)fill_out})
        if (im && valid_img(im) != NSX_OK) return S->fail(NSX_ERR_INVALID, "request output is malformed");
    auto same = [&](const nsx_img* a) { return a->rows == mask->rows && a->cols == mask->cols; };
    if (!same(marker) || !same(labels_out) || (recon_out && !same(recon_out)) || (fill_out && !same(fill_out)))
        return S->fail(NSX_ERR_INVALID, "request images must all have the mask's size");
    if (marker->dtype != mask->dtype) return S->fail(NSX_ERR_INVALID, "marker and mask must have the same type");
    if (mask->dtype == NSX_F32) return S->fail(NSX_ERR_UNSUPPORTED, "imfillHoles is instantiated for uchar and int only");
    if (labels_out->dtype != NSX_I32) return S->fail(NSX_ERR_INVALID, "labels must be int32");
    if ((recon_out && recon_out->dtype != mask->dtype) || (fill_out && fill_out->dtype != mask->dtype))
        return S->fail(NSX_ERR_INVALID, "recon/fill outputs must have the mask's type");
    Request* r = new (std::nothrow) Request();
    if (!r) return NSX_ERR_NOMEM;
    r->marker = *marker;
    r->mask = *mask;
    r->labels = *labels_out;
    if (recon_out) { r->recon = *recon_out; r->want_recon = true; }
    if (fill_out) { r->fill = *fill_out; r->want_fill = true; }
    r->recon_conn = recon_conn;
    r->fill_conn = fill_conn;
    r->ccl_conn = ccl_conn;
    r->thresh = thresh;
    r->min_size = min_size;
    r->max_size = max_size;
    r->bytes = layout_of(*r).total;
    if (align_up(r->bytes) > S->arena.size) {
        delete r;
        return S->fail(NSX_ERR_BUDGET, "request needs more device memory than the server's HBM budget");
    }
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stopping) {
            delete r;
            S->last_error = "server is shutting down";
            return NSX_ERR_INVALID;
        }
        r->ticket = S->next_ticket++;
        S->reqs.emplace(r->ticket, r);
        S->queue.push_back(r);
        *ticket = r->ticket;
    }
    S->cv_work.notify_all();
    return NSX_OK;
}

nsx_status nsx_server_wait(nsx_server* S, uint64_t ticket, int32_t* ncomp, double* times_ms) {
    if (!S) return NSX_ERR_INVALID;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) {
        S->last_error = "unknown ticket";
        return NSX_ERR_INVALID;
    }
    Request* r = it->second;
    S->cv_done.wait(lk, [&] { return r->done; });
    S->reqs.erase(it);
    lk.unlock();
    const nsx_status st = r->status;
    if (ncomp) *ncomp = r->ncomp;
    if (times_ms) {
        times_ms[0] = ms_between(r->t_submit, r->t_admit);
        times_ms[1] = ms_between(r->t_admit, r->t_done);
        times_ms[2] = ms_between(r->t_submit, r->t_done);
    }
    delete r;
    return st;
}

nsx_status nsx_server_stats(nsx_server* S, int64_t* stats) {
    if (!S || !stats) return NSX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    stats[0] = S->completed;
    stats[1] = S->inflight;
    stats[2] = (int64_t)S->queue.size();
    stats[3] = S->peak_inflight;
    stats[4] = (int64_t)S->arena.size;
    stats[5] = (int64_t)S->arena.peak;
    int64_t ws = 0;
    for (nsx_ctx* c : S->ctxs) ws += (int64_t)nsx_ctx_workspace_bytes(c);
    stats[6] = ws;
    stats[7] = (int64_t)S->ctxs.size();
    return NSX_OK;
}

// The message is copied under the lock into the calling thread's own buffer:
// workers may replace last_error at any time, so a pointer into it could
// dangle.  The returned pointer stays valid until this thread's next call.
const char* nsx_server_last_error(nsx_server* S) {
    if (!S) return "null server";
    thread_local std::string copy;
    std::lock_guard<std::mutex> lk(S->mu);
    copy = S->last_error;
    return copy.c_str();
}

void nsx_server_destroy(nsx_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stopping = true;
    }
    S->cv_work.notify_all();
    for (std::thread& t : S->workers) t.join();
    for (auto& kv : S->reqs) delete kv.second;   // completed but never waited for
    (void)hipSetDevice(S->device);
    for (nsx_ctx* c : S->ctxs) nsx_ctx_destroy(c);
    hipFree(S->arena.base);
    delete S;
}

}  // extern "C"
