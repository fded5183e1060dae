// server.hip -- serving loop for the segmentation chain.
//
// Requests (marker, mask -> recon, fill, labels) arrive on any host thread and
// are admitted in arrival order when a worker slot is idle and the request's
// device buffers fit in a fixed HBM arena.  Each slot is a host thread with its
// own nsx_ctx (workspace, stream, round predictions).  An idle slot takes the
// oldest queued request and, when it is small, the small requests queued
// behind it that fit in the arena (up to the server's batch limit), so small
// requests that arrive while the device is busy are served together.  Requests
// of one kind (element type, connectivities, threshold, area bounds) then
// share their launches: one reconstruction over the tiles of all their images,
// one hole filling, one labelling, with a tile -> request table selecting each
// tile's image.  The
// arena is one hipMalloc made at creation; requests take and return blocks of
// it through a first-fit free list, so the loop itself never allocates device
// memory per request.  Every stage's result is unique (a fixpoint, or labels
// numbered in raster order within each image), so a request's output does not
// depend on what else it was served with.
#include "nsx_internal.h"
#include "server_core.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <tuple>

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

using nsx_srv::align_up;
using nsx_srv::kAlign;

constexpr int kMaxBatch = 64;        // requests one slot may take at once
// Per slot, pinned host memory: readback words of a shared set (6 + n) in
// the first kPinnedWordBytes, then the staging of its tables.
constexpr size_t kPinnedBytes = 512 << 10, kPinnedWordBytes = 4096;

size_t elem_size(int dtype) { return dtype == NSX_U8 ? 1 : 4; }

struct Request {
    uint64_t ticket = 0;
    nsx_img marker{}, mask{}, labels{}, recon{}, fill{};
    bool want_recon = false, want_fill = false;
    int recon_conn = 8, fill_conn = 4, ccl_conn = 8, min_size = 0, max_size = 0;
    double thresh = 0;
    size_t bytes = 0, off = 0;   // arena block
    int64_t px = 0;              // pixels of the mask
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
    nsx_srv::Arena arena;                // offsets into arena_base
    char* arena_base = nullptr;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::unordered_map<uint64_t, Request*> reqs;
    uint64_t next_ticket = 1;
    bool stopping = false;
    int inflight = 0, peak_inflight = 0;
    int64_t completed = 0;
    nsx_srv::SharingPolicy policy;
    int64_t shared_sets = 0, shared_reqs = 0, largest_set = 0, shared_short = 0;
    std::vector<char*> pinned;           // per slot: kPinnedBytes of pinned host memory
    std::string last_error;

    nsx_status fail(nsx_status st, const std::string& msg) {
        std::lock_guard<std::mutex> lk(mu);
        last_error = msg;
        return st;
    }
    void worker(int slot);
    void serve(int slot, std::vector<Request*>& rs);
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

namespace {

using nsx::View;

// A request's device images while it is being served.
struct Job {
    Request* r = nullptr;
    Layout L;
    nsx_img dm{}, dk{}, dr{}, df{}, dl{};
    nsx_status st = NSX_OK;

    View view(const nsx_img& im) const { return View{(char*)im.data, im.pitch_bytes, im.rows, im.cols, im.dtype}; }
    void hip(hipError_t e, const char* what) {
        if (e != hipSuccess && st == NSX_OK) {
            st = NSX_ERR_HIP;
            r->error = std::string(what) + ": " + hipGetErrorString(e);
        }
    }
    void fail(nsx_status s, const std::string& msg) {
        if (st == NSX_OK) { st = s; r->error = msg; }
    }
};

// Places the request's planes in its arena block and enqueues the uploads of
// its host inputs.
void stage_in(char* arena_base, Job& j, hipStream_t s) {
    Request& r = *j.r;
    j.L = layout_of(r);
    const Layout& L = j.L;
    char* base = arena_base + r.off;
    const int dt = r.mask.dtype;
    const size_t row_t = (size_t)r.mask.cols * elem_size(dt), rows = (size_t)r.mask.rows;
    j.dm = r.marker;
    j.dk = r.mask;
    j.dl = r.labels;
    if (r.marker.mem == NSX_MEM_HOST) {
        j.dm = device_img(base, L.off_marker, L.pitch_t, r.marker, dt);
        j.hip(nsx::copy_rows(j.dm.data, L.pitch_t, r.marker.data, r.marker.pitch_bytes, row_t, rows,
                             hipMemcpyHostToDevice, s), "marker upload");
    }
    if (r.mask.mem == NSX_MEM_HOST) {
        j.dk = device_img(base, L.off_mask, L.pitch_t, r.mask, dt);
        j.hip(nsx::copy_rows(j.dk.data, L.pitch_t, r.mask.data, r.mask.pitch_bytes, row_t, rows,
                             hipMemcpyHostToDevice, s), "mask upload");
    }
    j.dr = (r.want_recon && r.recon.mem == NSX_MEM_DEVICE) ? r.recon : device_img(base, L.off_recon, L.pitch_t, r.mask, dt);
    j.df = (r.want_fill && r.fill.mem == NSX_MEM_DEVICE) ? r.fill : device_img(base, L.off_fill, L.pitch_t, r.mask, dt);
    if (r.labels.mem == NSX_MEM_HOST) j.dl = device_img(base, L.off_labels, L.pitch_l, r.labels, NSX_I32);
}

// Enqueues the downloads of the outputs the caller holds in host memory.
void stage_out(Job& j, hipStream_t s) {
    Request& r = *j.r;
    if (j.st != NSX_OK || r.mask.rows <= 0) return;
    const size_t row_t = (size_t)r.mask.cols * elem_size(r.mask.dtype), row_l = (size_t)r.mask.cols * 4;
    const size_t rows = (size_t)r.mask.rows;
    if (r.want_recon && r.recon.mem == NSX_MEM_HOST)
        j.hip(nsx::copy_rows(r.recon.data, r.recon.pitch_bytes, j.dr.data, j.L.pitch_t, row_t, rows,
                             hipMemcpyDeviceToHost, s), "recon download");
    if (r.want_fill && r.fill.mem == NSX_MEM_HOST)
        j.hip(nsx::copy_rows(r.fill.data, r.fill.pitch_bytes, j.df.data, j.L.pitch_t, row_t, rows,
                             hipMemcpyDeviceToHost, s), "fill download");
    if (r.labels.mem == NSX_MEM_HOST)
        j.hip(nsx::copy_rows(r.labels.data, r.labels.pitch_bytes, j.dl.data, j.L.pitch_l, row_l, rows,
                             hipMemcpyDeviceToHost, s), "labels download");
}

// One request on its own: the chain entry point on its device images.
void run_one(nsx_ctx* ctx, Job& j) {
    if (j.st != NSX_OK) return;
    Request& r = *j.r;
    int32_t n = 0;
    const nsx_status st = nsx_segment_chain(ctx, &j.dm, &j.dk, r.recon_conn, r.fill_conn, r.thresh, r.ccl_conn,
                                            r.min_size, r.max_size, &j.dr, &j.df, &j.dl, &n, ctx->stream);
    if (st != NSX_OK) j.fail(st, nsx_last_error(ctx));
    r.ncomp = n;
}

// Requests that may share launches: same element type, connectivities,
// threshold and area bounds.
using Kind = std::tuple<int, int, int, int, double, int, int>;
Kind kind_of(const Request& r) {
    return Kind(r.mask.dtype, r.recon_conn == 8 ? 8 : 4, r.fill_conn == 8 ? 8 : 4, r.ccl_conn == 8 ? 8 : 4, r.thresh,
                r.min_size, r.max_size);
}

// u8 reconstructions past this many propagation rounds are finished per
// request (where the adaptive engine can take the component tree).
constexpr int kSharedMaxRounds = 256;

size_t up256(size_t v) { return (v + 255) & ~(size_t)255; }

__global__ void clear_words_kernel(int32_t* w, int first, int second) {
    if (threadIdx.x == 0) {
        if (first) w[0] = 0;
        if (second) w[1] = 0;
    }
}

// Several requests of one kind through one set of launches: reconstruction,
// hole filling and labelling each over the tiles of every image.  The
// propagations launch their predicted rounds without a host check; one
// readback then gives both convergence words and every request's component
// count.  A propagation that needed more rounds is finished and the stages
// after it are run again.  `pin` = the slot's pinned block (kPinnedBytes);
// *short_rounds tells whether a prediction fell short.
void run_shared(nsx_ctx* ctx, std::vector<Job*>& g, char* pin, bool* short_rounds) {
    using namespace nsx;
    hipStream_t s = ctx->stream;
    const int n = (int)g.size();
    int32_t* hw = (int32_t*)pin;
    *short_rounds = false;
    const Request& k = *g[0]->r;
    const int dt = k.mask.dtype;
    const int rconn = k.recon_conn == 8 ? 8 : 4, fconn = k.fill_conn == 8 ? 8 : 4;
    auto fail_all = [&](nsx_status st, const std::string& msg) {
        for (Job* j : g) j->fail(st, msg);
    };
    // host tables: reconstruction and fill descriptors, labelling descriptors,
    // tile -> request tables
    const size_t rdb = recon_desc_bytes(), cdb = ccl_desc_bytes();
    std::vector<uint8_t> rd(rdb * n), fd(rdb * n), cd(cdb * n);
    std::vector<int32_t> rtile, ctile;
    int rbase = 0, cbase = 0;
    int64_t pbase = 0;
    for (int i = 0; i < n; ++i) {
        Job& j = *g[i];
        const View vm = j.view(j.dm), vk = j.view(j.dk), vr = j.view(j.dr), vf = j.view(j.df), vl = j.view(j.dl);
        recon_fill_desc(rd.data() + rdb * i, &vm, vk, vr, rbase);
        recon_fill_desc(fd.data() + rdb * i, nullptr, vr, vf, rbase);
        const int nr = recon_tiles(vk.rows, vk.cols);
        rtile.insert(rtile.end(), nr, i);
        rbase += nr;
        ccl_fill_desc(cd.data() + cdb * i, vf, vl, cbase, pbase);
        const int nc = ccl_tiles(vk.rows, vk.cols);
        ctile.insert(ctile.end(), nc, i);
        cbase += nc;
        pbase += ccl_pix_span(vk.rows, vk.cols);
    }
    // one device block: [rd | fd | cd | rtile | ctile | words: pending r, f, ncomp[n]]
    const size_t o_fd = up256(rd.size()), o_cd = o_fd + up256(fd.size()), o_rt = o_cd + up256(cd.size());
    const size_t o_ct = o_rt + up256(sizeof(int32_t) * rtile.size()), o_w = o_ct + up256(sizeof(int32_t) * ctile.size());
    const size_t total = o_w + sizeof(int32_t) * (2 + (size_t)n);
    // staged in pinned memory (a copy from pageable memory may wait for the
    // stream); a set whose tables do not fit there goes through a host vector
    std::vector<uint8_t> big;
    uint8_t* blob = (uint8_t*)pin + kPinnedWordBytes;
    if (o_w > kPinnedBytes - kPinnedWordBytes) {
        big.assign(o_w, 0);
        blob = big.data();
    }
    std::memcpy(blob, rd.data(), rd.size());
    std::memcpy(blob + o_fd, fd.data(), fd.size());
    std::memcpy(blob + o_cd, cd.data(), cd.size());
    std::memcpy(blob + o_rt, rtile.data(), sizeof(int32_t) * rtile.size());
    std::memcpy(blob + o_ct, ctile.data(), sizeof(int32_t) * ctile.size());
    void* dev = nullptr;
    nsx_status st = ctx->ws(WS_SERVE_TABLES, total, &dev);
    if (st != NSX_OK) return fail_all(st, nsx_last_error(ctx));
    char* d = (char*)dev;
    int32_t* words = (int32_t*)(d + o_w);
    // `blob` outlives the copy: the stream is drained before this returns,
    // and the slot's next set is staged only after that
    hipError_t he = hipMemcpyAsync(d, blob, o_w, hipMemcpyHostToDevice, s);
    if (he != hipSuccess) return fail_all(NSX_ERR_HIP, std::string("table upload: ") + hipGetErrorString(he));

    ReconBatch br{}, bf{};
    br.nimg = bf.nimg = n;
    br.ntiles = bf.ntiles = rbase;
    br.dev_desc = d;
    bf.dev_desc = d + o_fd;
    br.tile_img = bf.tile_img = (const int32_t*)(d + o_rt);
    br.ws_base = WS_RECON_STAMP;
    bf.ws_base = WS_FILL_STAMP;
    br.stats = ctx->d_stats;
    bf.stats = ctx->d_stats + 8;
    // slots run side by side on their own streams: a persistent drain sized
    // for the whole GPU would wait for blocks held by another slot's kernels
    br.drain = bf.drain = -1;
    CclBatch cb{};
    cb.nimg = n;
    cb.ntiles = cbase;
    cb.pix_total = pbase;
    cb.dev_desc = d + o_cd;
    cb.tile_img = (const int32_t*)(d + o_ct);
    cb.in_dtype = dt;
    cb.ws_base = WS_CCL_LABEL;
    cb.ncomp_dev = words + 2;
    CclParams p;
    p.conn = k.ccl_conn == 8 ? 8 : 4;
    p.out_mode = CO_CANON;
    p.threshold = true;
    p.thr = k.thresh;
    p.use_area = true;
    p.min_size = k.min_size;
    p.max_size = k.max_size;

    // int32 hole filling under wrapping negation is the constant INT_MIN
    const bool fill_const = dt == NSX_I32 && ctx->i32_invert == NSX_I32_INVERT_WRAP;
    const bool auto_u8 = dt == NSX_U8 && ctx->engine == NSX_ENGINE_AUTO;
    auto fill_constant = [&]() -> nsx_status {
        for (Job* j : g) NSX_TRY(fill_const_i32(ctx, j->view(j->df), INT_MIN, s));
        return NSX_OK;
    };
    // per request, stage by stage (the adaptive engine picks rounds or tree)
    auto recon_each = [&](int mode) -> nsx_status {
        for (Job* j : g) {
            const View vm = j->view(j->dm);
            NSX_TRY(mode == RM_RECON ? recon_run(ctx, RM_RECON, &vm, j->view(j->dk), j->view(j->dr), rconn, s)
                                     : recon_run(ctx, RM_FILL, nullptr, j->view(j->dr), j->view(j->df), fconn, s, nullptr,
                                                 ctx->d_stats + 8));
        }
        return NSX_OK;
    };
    auto read_words = [&]() -> nsx_status {
        NSX_HIP_TRY(ctx, hipMemcpyAsync(hw, words, sizeof(int32_t) * (2 + (size_t)n), hipMemcpyDeviceToHost, s));
        NSX_HIP_TRY(ctx, hipMemcpyAsync(hw + 2 + n, ctx->d_stats, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        NSX_HIP_TRY(ctx, hipMemcpyAsync(hw + 4 + n, ctx->d_stats + 8, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
        return NSX_OK;
    };
    auto body = [&]() -> nsx_status {
        const bool long_r = auto_u8 && recon_predicted_rounds(ctx, RM_RECON, dt, rconn, rbase, 0) > kSharedMaxRounds;
        const bool long_f = auto_u8 && recon_predicted_rounds(ctx, RM_FILL, dt, fconn, rbase, 0) > kSharedMaxRounds;
        ReconRun rr, rf;
        if (long_r) NSX_TRY(recon_each(RM_RECON));
        else NSX_TRY(recon_launch(ctx, RM_RECON, dt, rconn, br, s, &rr, words));
        if (fill_const) {
            NSX_TRY(fill_constant());
        } else if (long_r || long_f) {
            // per-request fills read the reconstruction: complete it first
            bool gave_up = false;
            if (rr.impl) NSX_TRY(recon_complete(ctx, &rr, s, nullptr, kSharedMaxRounds, &gave_up));
            if (gave_up) NSX_TRY(recon_each(RM_RECON));
            NSX_TRY(recon_each(RM_FILL));
            NSX_TRY(ccl_launch(ctx, cb, p, s));
            return read_words();
        } else {
            NSX_TRY(recon_launch(ctx, RM_FILL, dt, fconn, bf, s, &rf, words + 1));
        }
        NSX_TRY(ccl_launch(ctx, cb, p, s));
        if (!rr.impl || !rf.impl)   // a stage that was not launched has nothing pending
            hipLaunchKernelGGL(clear_words_kernel, dim3(1), dim3(64), 0, s, words, !rr.impl, !rf.impl);
        NSX_HIP_TRY(ctx, hipGetLastError());
        NSX_TRY(read_words());
        const int32_t pend_r = hw[0], pend_f = hw[1];
        *short_rounds = pend_r != 0 || pend_f != 0;
        if (pend_r == 0 && pend_f == 0) {
            int64_t wr, wf;   // rounds that had work, for the next prediction
            std::memcpy(&wr, hw + 2 + n, sizeof(wr));
            std::memcpy(&wf, hw + 4 + n, sizeof(wf));
            recon_note_rounds(ctx, &rr, wr);
            recon_note_rounds(ctx, &rf, wf);
            return NSX_OK;
        }
        // a propagation needed more rounds than predicted: finish it, then
        // run the stages after it again
        bool gave_up = false;
        if (pend_r != 0) {
            rf.release();
            NSX_TRY(recon_complete(ctx, &rr, s, nullptr, auto_u8 ? kSharedMaxRounds : 0, &gave_up));
            if (gave_up) NSX_TRY(recon_each(RM_RECON));
            if (!fill_const) {
                if (gave_up) NSX_TRY(recon_each(RM_FILL));
                else {
                    NSX_TRY(recon_launch(ctx, RM_FILL, dt, fconn, bf, s, &rf, nullptr));
                    NSX_TRY(recon_complete(ctx, &rf, s, nullptr, auto_u8 ? kSharedMaxRounds : 0, &gave_up));
                    if (gave_up) NSX_TRY(recon_each(RM_FILL));
                }
            }
        } else {
            rr.release();
            NSX_TRY(recon_complete(ctx, &rf, s, nullptr, auto_u8 ? kSharedMaxRounds : 0, &gave_up));
            if (gave_up) NSX_TRY(recon_each(RM_FILL));
        }
        NSX_TRY(ccl_launch(ctx, cb, p, s));
        return read_words();
    };
    st = body();
    if (st != NSX_OK) return fail_all(st, nsx_last_error(ctx));
    for (int i = 0; i < n; ++i) g[i]->r->ncomp = hw[2 + i];
}

}  // namespace

void nsx_server::serve(int slot, std::vector<Request*>& rs) {
    nsx_ctx* ctx = ctxs[slot];
    hipStream_t s = ctx->stream;
    std::vector<Job> jobs(rs.size());
    for (size_t i = 0; i < rs.size(); ++i) {
        jobs[i].r = rs[i];
        stage_in(arena_base, jobs[i], s);
    }
    // group by kind, in arrival order; empty images and kinds with one
    // request run on their own
    std::vector<std::pair<Kind, std::vector<Job*>>> groups;
    for (Job& j : jobs) {
        if (j.r->mask.rows <= 0 || j.r->mask.cols <= 0) {
            groups.push_back({kind_of(*j.r), {&j}});
            continue;
        }
        const Kind kd = kind_of(*j.r);
        auto it = std::find_if(groups.begin(), groups.end(), [&](auto& x) {
            return x.first == kd && x.second[0]->r->mask.rows > 0 && x.second[0]->r->mask.cols > 0;
        });
        if (it == groups.end()) groups.push_back({kd, {&j}});
        else it->second.push_back(&j);
    }
    int64_t sets = 0, shared = 0, largest = 0, shorts = 0;
    for (auto& grp : groups) {
        std::vector<Job*> live;
        for (Job* j : grp.second)
            if (j->st == NSX_OK) live.push_back(j);
        if (live.size() >= 2 && (ctx->engine == NSX_ENGINE_ROUNDS || ctx->engine == NSX_ENGINE_AUTO)) {
            bool short_rounds = false;
            run_shared(ctx, live, pinned[slot], &short_rounds);
            ++sets;
            shorts += short_rounds ? 1 : 0;
            shared += (int64_t)live.size();
            largest = std::max<int64_t>(largest, (int64_t)live.size());
        } else {
            for (Job* j : live) run_one(ctx, *j);
        }
    }
    for (Job& j : jobs) stage_out(j, s);
    // the arena blocks are reused as soon as they are released: drain the stream first
    const hipError_t e = hipStreamSynchronize(s);
    for (Job& j : jobs) {
        j.hip(e, "request completion");
        j.r->status = j.st;
    }
    std::lock_guard<std::mutex> lk(mu);
    shared_sets += sets;
    shared_reqs += shared;
    shared_short += shorts;
    largest_set = std::max(largest_set, largest);
}

void nsx_server::worker(int slot) {
    (void)hipSetDevice(device);
    std::vector<Request*> rs;
    for (;;) {
        rs.clear();
        {
            std::unique_lock<std::mutex> lk(mu);
            // Admission: the oldest queued request, once its block fits.  Later
            // requests wait behind it, so a large request cannot be starved.
            cv_work.wait(lk, [&] { return queue.empty() ? stopping : nsx_srv::admit_front(queue, arena); });
            if (queue.empty()) return;
            // then, behind a small request, the small requests that follow it
            // in order while they fit (large ones fill the device by themselves)
            nsx_srv::take_set(queue, arena, policy, rs);
            const Clock::time_point now = Clock::now();
            for (Request* r : rs) r->t_admit = now;
            inflight += (int)rs.size();
            peak_inflight = std::max(peak_inflight, inflight);
        }
        serve(slot, rs);
        {
            std::lock_guard<std::mutex> lk(mu);
            for (Request* r : rs) {
                arena.release(r->off, r->bytes);
                r->t_done = Clock::now();
                r->done = true;
                --inflight;
                ++completed;
                if (r->status != NSX_OK) last_error = r->error;
            }
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
    S->arena.reset(hbm_budget_bytes);
    if (hipMalloc((void**)&S->arena_base, S->arena.size) != hipSuccess) {
        (void)hipGetLastError();
        delete S;
        return NSX_ERR_NOMEM;
    }
    auto undo = [&](nsx_status st) {
        for (nsx_ctx* x : S->ctxs) nsx_ctx_destroy(x);
        for (char* w : S->pinned) hipHostFree(w);
        hipFree(S->arena_base);
        delete S;
        return st;
    };
    for (int i = 0; i < slots; ++i) {
        nsx_ctx* c = nullptr;
        nsx_status st = nsx_ctx_create(&c, device, 0);
        if (st != NSX_OK) return undo(st);
        S->ctxs.push_back(c);
        char* w = nullptr;
        if (hipHostMalloc((void**)&w, kPinnedBytes) != hipSuccess) {
            (void)hipGetLastError();
            return undo(NSX_ERR_NOMEM);
        }
        S->pinned.push_back(w);
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
    r->px = (int64_t)mask->rows * (int64_t)mask->cols;
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
    for (char* w : S->pinned) hipHostFree(w);
    hipFree(S->arena_base);
    delete S;
}

nsx_status nsx_server_set_sharing(nsx_server* S, int max_requests, int64_t max_request_pixels,
                                  int64_t max_set_pixels) {
    if (!S) return NSX_ERR_INVALID;
    if (max_requests < 1 || max_requests > kMaxBatch)
        return S->fail(NSX_ERR_INVALID, "max_requests must be 1.." + std::to_string(kMaxBatch));
    if (max_request_pixels < 0 || max_set_pixels < 0)
        return S->fail(NSX_ERR_INVALID, "pixel limits must be >= 0");
    std::lock_guard<std::mutex> lk(S->mu);
    const nsx_srv::SharingPolicy def;
    S->policy.max_requests = max_requests;
    S->policy.max_request_px = max_request_pixels ? max_request_pixels : def.max_request_px;
    S->policy.max_set_px = max_set_pixels ? max_set_pixels : def.max_set_px;
    return NSX_OK;
}

nsx_ctx* nsx_server_context(nsx_server* S, int slot) {
    if (!S || slot < 0 || slot >= (int)S->ctxs.size()) return nullptr;
    return S->ctxs[slot];
}

nsx_status nsx_server_batch_stats(nsx_server* S, int64_t* stats) {
    if (!S || !stats) return NSX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    stats[0] = S->shared_sets;
    stats[1] = S->shared_reqs;
    stats[2] = S->largest_set;
    stats[3] = S->policy.max_requests;
    stats[4] = S->shared_short;
    stats[5] = S->policy.max_request_px;
    stats[6] = S->policy.max_set_px;
    return NSX_OK;
}

}  // extern "C"
