// batch.hip -- many uneven tiles per launch, captured as one hipGraph.
//
// Replaces nscale::gpu::imreconstructQueueThroughput
// (MorphologicOperationsGPU.cpp:447-545, global_queue.cu:903-968), which gives
// each image one thread block and at most 70 images.  Here every tile of every
// image is a work item of the same launches: the reconstruction's first pass
// and propagation rounds run over a tile list spanning all images (a tile ->
// image table selects each tile's descriptor), and one set of labelling
// launches covers all images.  Items are grouped by (dtype, connectivity,
// threshold); each group is one set of launches.
//
// The batch owns a private workspace (a context of its own on the caller's
// device), so the device addresses captured in the graph stay valid for the
// batch's lifetime whatever else runs on the caller's context.  finalize()
// runs the batch once eagerly (allocating workspace and learning how many
// propagation rounds each group needs), then captures the launches.  A replay
// that needed more rounds than captured is detected by nsx_batch_counts(),
// which completes the affected groups eagerly and re-captures.
#include "nsx_internal.h"

#include <map>
#include <new>
#include <tuple>

using namespace nsx;

namespace {

__global__ void copy_i32_kernel(int32_t* dst, const int32_t* src, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

struct Group {
    int dtype, conn, ccl_conn;
    double thresh;
    std::vector<int> items;
    // reconstruction tables
    std::vector<uint8_t> rdesc;      // host copy of the descriptors
    void* d_rdesc = nullptr;
    int32_t* d_rtile = nullptr;
    int rtiles = 0;
    // labelling tables
    std::vector<uint8_t> cdesc;
    void* d_cdesc = nullptr;
    int32_t* d_ctile = nullptr;
    int ctiles = 0;
    int64_t pix_total = 0;
    int32_t* d_ncomp = nullptr;      // per item of the group
    int32_t* d_pending = nullptr;    // tiles still queued after the captured rounds
    int ws_recon = 0, ws_ccl = 0;    // workspace slot bases in the batch context
    int lroot_slot = WS_CCL_LROOT;
    int visit_sweeps = 0;            // sweep pairs per visit (0: the context's), see autotune
    int drain = 0;                   // rounds in one persistent launch: 1 on, -1 off, 0 the batch's default
    int drain_occ = 0;               // drain blocks per CU (0: as many as stay resident)
    int xcd_split = 0;               // 1: each XCD takes a contiguous share of a round's tiles; 0: plain stride
};

}  // namespace

struct nsx_batch {
    nsx_ctx* owner = nullptr;   // caller's context (errors are reported there)
    nsx_ctx* ws = nullptr;      // private workspace context
    struct Item { nsx_img marker, mask, recon, labels; int conn; double thresh; int ccl_conn; };
    std::vector<Item> items;
    std::vector<Group> groups;
    int32_t* d_counts = nullptr;     // per item, in add order
    bool finalized = false;
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap_stream = nullptr;
    int32_t* h_pending = nullptr;    // pinned, one per group
    // Odd groups run on a second stream, forked from the caller's stream and
    // joined back (also inside the graph), so that one group's late rounds,
    // which hold few tiles, overlap the other group's launches.  Even and odd
    // groups use separate reconstruction and labelling workspace slots.
    bool overlap = true;
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // Groups run one launch per propagation round unless set or tuned to the
    // persistent drain: on MI355X a kernel boundary (~2.5 us) is a cheaper
    // grid-wide barrier than the drain's (~100-290 us with every block resident).
    int drain_default = -1;
};

namespace {

View view_of(const nsx_img& im) {
    return View{(char*)im.data, im.pitch_bytes, im.rows, im.cols, im.dtype};
}

// The reconstruction of group gi as one batch of images.
ReconBatch recon_batch(nsx_batch* b, size_t gi) {
    const Group& g = b->groups[gi];
    ReconBatch rb{};
    rb.nimg = (int)g.items.size();
    rb.ntiles = g.rtiles;
    rb.host_desc = g.rdesc.data();
    rb.dev_desc = g.d_rdesc;
    rb.tile_img = g.d_rtile;
    rb.ws_base = g.ws_recon;
    rb.stats = b->ws->d_stats + 8 * (gi % 2);
    rb.visit_sweeps = g.visit_sweeps;
    rb.drain = g.drain != 0 ? g.drain : b->drain_default;
    rb.drain_occ = g.drain_occ;
    rb.xcd_split = g.xcd_split;
    return rb;
}

// Enqueues every group's reconstruction (first pass + predicted rounds) and
// labelling, then gathers the per-item counts.  No host synchronisation.
nsx_status enqueue(nsx_batch* b, hipStream_t s0, std::vector<ReconRun>* runs) {
    nsx_ctx* c = b->ws;
    runs->clear();
    runs->resize(b->groups.size());
    // A persistent drain sizes its grid to the whole GPU: it runs with nothing
    // beside it.  The fork needs a stream object to wait on: the special
    // handles (NULL, hipStreamLegacy, hipStreamPerThread) are not accepted by
    // hipStreamWaitEvent (an eager run from torch's default stream crashed in
    // it), so eager runs on them take the groups in turn; graph captures run
    // on the batch's own stream and always fork.
    bool fork = b->overlap && b->groups.size() > 1 && s0 != nullptr && s0 != hipStreamLegacy &&
                s0 != hipStreamPerThread;
    for (const auto& g : b->groups) fork = fork && (g.drain != 0 ? g.drain : b->drain_default) < 0;
    if (fork) {
        NSX_HIP_TRY(b->owner, hipEventRecord(b->ev_fork, s0));
        NSX_HIP_TRY(b->owner, hipStreamWaitEvent(b->side, b->ev_fork, 0));
    }
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        Group& g = b->groups[gi];
        hipStream_t s = fork && (gi % 2) ? b->side : s0;
        const ReconBatch rb = recon_batch(b, gi);
        NSX_TRY(recon_launch(c, RM_RECON, g.dtype, g.conn, rb, s, &(*runs)[gi], g.d_pending));
        CclBatch cb{};
        cb.nimg = (int)g.items.size();
        cb.ntiles = g.ctiles;
        cb.pix_total = g.pix_total;
        cb.host_desc = g.cdesc.data();
        cb.dev_desc = g.d_cdesc;
        cb.tile_img = g.d_ctile;
        cb.in_dtype = g.dtype;
        cb.ws_base = g.ws_ccl;
        cb.lroot_slot = g.lroot_slot;
        cb.ncomp_dev = g.d_ncomp;
        CclParams p;
        p.conn = g.ccl_conn;
        p.out_mode = CO_CANON;
        p.threshold = true;
        p.thr = g.thresh;
        NSX_TRY(ccl_launch(c, cb, p, s));
    }
    if (fork) {
        NSX_HIP_TRY(b->owner, hipEventRecord(b->ev_join, b->side));
        NSX_HIP_TRY(b->owner, hipStreamWaitEvent(s0, b->ev_join, 0));
    }
    return NSX_OK;
}

// Copies group-ordered counts back to add order.
nsx_status gather_counts(nsx_batch* b, hipStream_t s) {
    for (auto& g : b->groups)
        for (size_t k = 0; k < g.items.size(); ++k)
            hipLaunchKernelGGL(copy_i32_kernel, dim3(1), dim3(64), 0, s, b->d_counts + g.items[k], g.d_ncomp + k, 1);
    NSX_HIP_TRY(b->owner, hipGetLastError());
    return NSX_OK;
}

nsx_status fail_to_owner(nsx_batch* b, nsx_status st) {
    if (st != NSX_OK && b->ws && b->owner != b->ws) b->owner->last_error = b->ws->last_error;
    return st;
}

// Runs every group's reconstruction eagerly to convergence, which records the
// rounds each one needs (the next enqueue launches that many).
nsx_status learn_rounds(nsx_batch* b, hipStream_t s) {
    nsx_ctx* c = b->ws;
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        const Group& g = b->groups[gi];
        const ReconBatch rb = recon_batch(b, gi);
        ReconRun run;
        nsx_status st = recon_launch(c, RM_RECON, g.dtype, g.conn, rb, s, &run, nullptr);
        if (st == NSX_OK) st = recon_complete(c, &run, s, nullptr);
        if (st != NSX_OK) return fail_to_owner(b, st);
    }
    return NSX_OK;
}

// (Re-)captures the batch's launches into its graph.
nsx_status capture(nsx_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    if (!b->cap_stream) NSX_HIP_TRY(b->owner, hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
    std::vector<ReconRun> runs;
    NSX_HIP_TRY(b->owner, hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeRelaxed));
    nsx_status st = enqueue(b, b->cap_stream, &runs);
    if (st == NSX_OK) st = gather_counts(b, b->cap_stream);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(b->cap_stream, &gr);
    for (auto& r : runs) r.release();
    if (st != NSX_OK) return fail_to_owner(b, st);
    if (e != hipSuccess) return b->owner->fail(NSX_ERR_HIP, "batch graph capture failed");
    b->graph = gr;
    NSX_HIP_TRY(b->owner, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return NSX_OK;
}

}  // namespace

extern "C" {

nsx_status nsx_batch_create(nsx_ctx* ctx, nsx_batch** batch) {
    if (!ctx || !batch) return NSX_ERR_INVALID;
    nsx_batch* b = new (std::nothrow) nsx_batch();
    if (!b) return NSX_ERR_NOMEM;
    b->owner = ctx;
    nsx_status st = nsx_ctx_create(&b->ws, ctx->device, ctx->budget);
    if (st != NSX_OK) { delete b; return ctx->fail(st, "batch workspace context"); }
    *batch = b;
    return NSX_OK;
}

nsx_status nsx_batch_add(nsx_batch* b, const nsx_img* marker, const nsx_img* mask, nsx_img* recon_out,
                         nsx_img* labels_out, int conn, double thresh, int ccl_conn) {
    if (!b || !marker || !mask || !recon_out || !labels_out) return NSX_ERR_INVALID;
    if (b->finalized) return b->owner->fail(NSX_ERR_INVALID, "batch already finalized");
    for (const nsx_img* im : {marker, mask, (const nsx_img*)recon_out, (const nsx_img*)labels_out})
        if (im->mem != NSX_MEM_DEVICE || !im->data || im->rows <= 0 || im->cols <= 0)
            return b->owner->fail(NSX_ERR_INVALID, "batch images must be non-empty and device-resident");
    if (marker->rows != mask->rows || marker->cols != mask->cols || recon_out->rows != mask->rows ||
        recon_out->cols != mask->cols || labels_out->rows != mask->rows || labels_out->cols != mask->cols)
        return b->owner->fail(NSX_ERR_INVALID, "batch item images must have the same size");
    if (marker->dtype != mask->dtype || recon_out->dtype != mask->dtype || labels_out->dtype != NSX_I32)
        return b->owner->fail(NSX_ERR_INVALID, "marker/mask/recon share a dtype; labels are int32");
    b->items.push_back({*marker, *mask, *recon_out, *labels_out, conn == 8 ? 8 : 4, thresh, ccl_conn == 8 ? 8 : 4});
    return NSX_OK;
}

nsx_status nsx_batch_finalize(nsx_batch* b, int use_graph) {
    if (!b) return NSX_ERR_INVALID;
    if (b->finalized) return b->owner->fail(NSX_ERR_INVALID, "batch already finalized");
    if (b->items.empty()) return b->owner->fail(NSX_ERR_INVALID, "empty batch");
    nsx_ctx* c = b->ws;
    // group items
    std::map<std::tuple<int, int, int, double>, int> key2g;
    for (size_t i = 0; i < b->items.size(); ++i) {
        const auto& it = b->items[i];
        auto key = std::make_tuple(it.mask.dtype, it.conn, it.ccl_conn, it.thresh);
        auto f = key2g.find(key);
        if (f == key2g.end()) {
            key2g[key] = (int)b->groups.size();
            Group g{};
            g.dtype = it.mask.dtype; g.conn = it.conn; g.ccl_conn = it.ccl_conn; g.thresh = it.thresh;
            b->groups.push_back(g);
            f = key2g.find(key);
        }
        b->groups[f->second].items.push_back((int)i);
    }
    if (b->groups.size() > 4) return b->owner->fail(NSX_ERR_UNSUPPORTED, "at most 4 (dtype, conn, threshold) groups");
    if (hipMalloc((void**)&b->d_counts, sizeof(int32_t) * b->items.size()) != hipSuccess)
        return b->owner->fail(NSX_ERR_NOMEM, "batch counts");
    if (hipHostMalloc((void**)&b->h_pending, sizeof(int32_t) * 8) != hipSuccess)
        return b->owner->fail(NSX_ERR_NOMEM, "batch pinned words");
    const size_t rdb = recon_desc_bytes(), cdb = ccl_desc_bytes();
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        Group& g = b->groups[gi];
        const int n = (int)g.items.size();
        g.rdesc.assign(rdb * n, 0);
        g.cdesc.assign(cdb * n, 0);
        std::vector<int32_t> rt, ct;
        int rbase = 0, cbase = 0;
        int64_t pbase = 0;
        for (int k = 0; k < n; ++k) {
            const auto& it = b->items[g.items[k]];
            View vm = view_of(it.marker), vk = view_of(it.mask), vr = view_of(it.recon), vl = view_of(it.labels);
            recon_fill_desc(g.rdesc.data() + rdb * k, &vm, vk, vr, rbase);
            const int nr = recon_tiles(it.mask.rows, it.mask.cols);
            rt.insert(rt.end(), nr, k);
            rbase += nr;
            ccl_fill_desc(g.cdesc.data() + cdb * k, vr, vl, cbase, pbase);
            const int nc = ccl_tiles(it.mask.rows, it.mask.cols);
            ct.insert(ct.end(), nc, k);
            cbase += nc;
            pbase += ccl_pix_span(it.mask.rows, it.mask.cols);
        }
        g.rtiles = rbase;
        g.ctiles = cbase;
        g.pix_total = pbase;
        if (pbase >= (int64_t)INT32_MAX) return b->owner->fail(NSX_ERR_INVALID, "batch group exceeds 2^31 pixels");
        bool ok = hipMalloc(&g.d_rdesc, g.rdesc.size()) == hipSuccess &&
                  hipMalloc((void**)&g.d_rtile, sizeof(int32_t) * rt.size()) == hipSuccess &&
                  hipMalloc(&g.d_cdesc, g.cdesc.size()) == hipSuccess &&
                  hipMalloc((void**)&g.d_ctile, sizeof(int32_t) * ct.size()) == hipSuccess &&
                  hipMalloc((void**)&g.d_ncomp, sizeof(int32_t) * n) == hipSuccess &&
                  hipMalloc((void**)&g.d_pending, sizeof(int32_t)) == hipSuccess;
        if (!ok) return b->owner->fail(NSX_ERR_NOMEM, "batch tables");
        NSX_HIP_TRY(b->owner, hipMemcpy(g.d_rdesc, g.rdesc.data(), g.rdesc.size(), hipMemcpyHostToDevice));
        NSX_HIP_TRY(b->owner, hipMemcpy(g.d_rtile, rt.data(), sizeof(int32_t) * rt.size(), hipMemcpyHostToDevice));
        NSX_HIP_TRY(b->owner, hipMemcpy(g.d_cdesc, g.cdesc.data(), g.cdesc.size(), hipMemcpyHostToDevice));
        NSX_HIP_TRY(b->owner, hipMemcpy(g.d_ctile, ct.data(), sizeof(int32_t) * ct.size(), hipMemcpyHostToDevice));
        // even and odd groups get their own workspace slots in the private context
        g.ws_recon = gi % 2 == 0 ? WS_RECON_STAMP : WS_FILL_STAMP;
        g.ws_ccl = gi % 2 == 0 ? WS_CCL_LABEL : WS_CCL2_LABEL;
        g.lroot_slot = gi % 2 == 0 ? WS_CCL_LROOT : WS_CCL2_LROOT;
    }
    // groups 2 and 3 reuse the slots of 0 and 1: they run after them on the
    // same stream
    NSX_HIP_TRY(b->owner, hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
    NSX_HIP_TRY(b->owner, hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
    NSX_HIP_TRY(b->owner, hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming));
    b->finalized = true;
    b->use_graph = use_graph != 0;
    // warm-up: learn the round counts of each group, then one full eager pass
    // that sizes every workspace slot for the largest group, so that nothing
    // is reallocated while launches are captured
    hipStream_t s = c->stream;
    NSX_TRY(learn_rounds(b, s));
    {
        // size every workspace slot for the largest group of its parity, so
        // that nothing is reallocated while launches are captured or while
        // the other parity's launches run
        int64_t max_pix[2] = {0, 0};
        int max_ct[2] = {0, 0}, max_rt[2] = {0, 0};
        for (size_t gi = 0; gi < b->groups.size(); ++gi) {
            const int q = (int)(gi % 2);
            max_pix[q] = std::max(max_pix[q], b->groups[gi].pix_total);
            max_ct[q] = std::max(max_ct[q], b->groups[gi].ctiles);
            max_rt[q] = std::max(max_rt[q], b->groups[gi].rtiles);
        }
        void* p;
        nsx_status st = NSX_OK;
        for (int q = 0; q < 2 && st == NSX_OK; ++q) {
            if (!max_pix[q]) continue;
            const int base = q ? WS_CCL2_LABEL : WS_CCL_LABEL;
            const int64_t nw = max_pix[q] / 32, nb = max_pix[q] / 32768;
            if (st == NSX_OK) st = c->ws(base + 0, sizeof(int32_t) * (size_t)max_pix[q], &p);
            if (st == NSX_OK) st = c->ws(base + 1, sizeof(int32_t) * (size_t)max_pix[q], &p);
            if (st == NSX_OK) st = c->ws(base + 2, sizeof(uint2) * (size_t)max_ct[q] * 2048, &p);
            if (st == NSX_OK) st = c->ws(base + 3, sizeof(int32_t) * (size_t)max_ct[q], &p);
            if (st == NSX_OK) st = c->ws(base + 4, sizeof(uint32_t) * (size_t)nw, &p);
            if (st == NSX_OK) st = c->ws(base + 5, sizeof(uint32_t) * (size_t)(nw + nb + 1), &p);
            if (st == NSX_OK)
                st = c->ws(q ? WS_CCL2_LROOT : WS_CCL_LROOT, sizeof(uint16_t) * (size_t)max_pix[q], &p);
            const int rbase = q ? WS_FILL_STAMP : WS_RECON_STAMP;
            if (st == NSX_OK) st = c->ws(rbase + 0, sizeof(uint32_t) * (size_t)max_rt[q], &p);
            if (st == NSX_OK) st = c->ws(rbase + 1, recon_list_bytes(max_rt[q]), &p);
        }
        if (st != NSX_OK) return fail_to_owner(b, st);
        std::vector<ReconRun> runs;
        st = enqueue(b, s, &runs);
        for (auto& r : runs) r.release();
        if (st == NSX_OK) st = gather_counts(b, s);
        if (st != NSX_OK) return fail_to_owner(b, st);
        NSX_HIP_TRY(b->owner, hipStreamSynchronize(s));
    }
    if (b->use_graph) NSX_TRY(capture(b));
    return NSX_OK;
}

nsx_status nsx_batch_run(nsx_batch* b, void* stream) {
    if (!b || !b->finalized) return NSX_ERR_INVALID;
    hipStream_t s = stream ? (hipStream_t)stream : b->owner->stream;
    if (b->exec) {
        NSX_HIP_TRY(b->owner, hipGraphLaunch(b->exec, s));
        return NSX_OK;
    }
    std::vector<ReconRun> runs;
    nsx_status st = enqueue(b, s, &runs);
    for (auto& r : runs) r.release();
    if (st != NSX_OK) return fail_to_owner(b, st);
    return gather_counts(b, s);
}

nsx_status nsx_batch_counts(nsx_batch* b, int32_t* counts, int n) {
    if (!b || !counts || n < 0 || (size_t)n > b->items.size()) return NSX_ERR_INVALID;
    NSX_HIP_TRY(b->owner, hipDeviceSynchronize());
    // a group whose propagation needed more rounds than launched: finish it
    // eagerly, relabel, and re-capture the graph with the new round counts
    bool redo = false;
    nsx_ctx* c = b->ws;
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        Group& g = b->groups[gi];
        NSX_HIP_TRY(b->owner, hipMemcpy(b->h_pending, g.d_pending, sizeof(int32_t), hipMemcpyDeviceToHost));
        if (b->h_pending[0] == 0) continue;
        redo = true;
        const ReconBatch rb = recon_batch(b, gi);
        ReconRun run;
        nsx_status st = recon_launch(c, RM_RECON, g.dtype, g.conn, rb, c->stream, &run, nullptr);
        if (st == NSX_OK) st = recon_complete(c, &run, c->stream, nullptr);
        CclBatch cb{};
        cb.nimg = (int)g.items.size(); cb.ntiles = g.ctiles; cb.pix_total = g.pix_total;
        cb.host_desc = g.cdesc.data(); cb.dev_desc = g.d_cdesc; cb.tile_img = g.d_ctile; cb.in_dtype = g.dtype;
        cb.ws_base = g.ws_ccl; cb.lroot_slot = g.lroot_slot; cb.ncomp_dev = g.d_ncomp;
        CclParams p;
        p.conn = g.ccl_conn; p.out_mode = CO_CANON; p.threshold = true; p.thr = g.thresh;
        if (st == NSX_OK) st = ccl_launch(c, cb, p, c->stream);
        if (st != NSX_OK) return fail_to_owner(b, st);
    }
    if (redo) {
        NSX_TRY(gather_counts(b, c->stream));
        NSX_HIP_TRY(b->owner, hipStreamSynchronize(c->stream));
        if (b->exec) NSX_TRY(capture(b));
    }
    NSX_HIP_TRY(b->owner, hipMemcpy(counts, b->d_counts, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    return NSX_OK;
}

nsx_status nsx_batch_autotune(nsx_batch* b, void* stream, int reps, int32_t* pairs, int npairs) {
    if (!b || !b->finalized || reps < 1 || npairs < 0 || (npairs > 0 && !pairs)) return NSX_ERR_INVALID;
    if (!b->exec) return b->owner->fail(NSX_ERR_INVALID, "autotune needs a batch finalized with a graph");
    hipStream_t s = stream ? (hipStream_t)stream : b->owner->stream;
    hipEvent_t e0, e1;
    NSX_HIP_TRY(b->owner, hipEventCreate(&e0));
    NSX_HIP_TRY(b->owner, hipEventCreate(&e1));
    // one setting: learn its round counts, capture, one replay to warm up,
    // then the mean of `reps` timed replays
    auto trial = [&](float* ms) -> nsx_status {
        NSX_TRY(learn_rounds(b, b->ws->stream));
        NSX_HIP_TRY(b->owner, hipStreamSynchronize(b->ws->stream));
        NSX_TRY(capture(b));
        NSX_HIP_TRY(b->owner, hipGraphLaunch(b->exec, s));
        NSX_HIP_TRY(b->owner, hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) NSX_HIP_TRY(b->owner, hipGraphLaunch(b->exec, s));
        NSX_HIP_TRY(b->owner, hipEventRecord(e1, s));
        NSX_HIP_TRY(b->owner, hipEventSynchronize(e1));
        NSX_HIP_TRY(b->owner, hipEventElapsedTime(ms, e0, e1));
        *ms /= (float)reps;
        return NSX_OK;
    };
    // Coordinate descent, group by group (each group is its own set of
    // launches), one parameter at a time with the others held: the drain (one
    // persistent launch, or one launch per round), the sweep pairs per tile
    // visit, the work split over the XCDs, and the drain's blocks per CU.
    // A second pass runs only if the first changed something.
    static const int kSweeps[] = {3, 4, 5, 6};
    static const int kOcc[] = {0, 8, 4};
    nsx_status st = NSX_OK;
    float cur_ms = 0.f;
    for (auto& g : b->groups) {
        if (g.visit_sweeps <= 0) g.visit_sweeps = b->ws->visit_sweeps;
        if (g.drain == 0) g.drain = b->drain_default;
    }
    st = trial(&cur_ms);
    for (int pass = 0; pass < 2 && st == NSX_OK; ++pass) {
        bool changed = false;
        if (b->groups.size() > 1 && st == NSX_OK) {
            // odd groups on their own stream, or every group on one stream
            b->overlap = !b->overlap;
            float ms = 0.f;
            st = trial(&ms);
            if (st == NSX_OK && ms < cur_ms) {
                cur_ms = ms;
                changed = true;
            } else {
                b->overlap = !b->overlap;
            }
        }
        for (size_t gi = 0; gi < b->groups.size() && st == NSX_OK; ++gi) {
            Group& g = b->groups[gi];
            // try one value of one parameter; keep it if faster
            auto attempt = [&](int* field, int value) {
                if (st != NSX_OK || *field == value) return;
                const int old = *field;
                *field = value;
                float ms = 0.f;
                st = trial(&ms);
                if (st == NSX_OK && ms < cur_ms) {
                    cur_ms = ms;
                    changed = true;
                } else {
                    *field = old;
                }
            };
            attempt(&g.drain, -g.drain);
            for (int v : kSweeps) attempt(&g.visit_sweeps, v);
            attempt(&g.xcd_split, 1 - g.xcd_split);
            if (g.drain > 0)
                for (int v : kOcc) attempt(&g.drain_occ, v);
        }
        if (!changed) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    NSX_TRY(st);
    // leave the graph captured with the chosen settings
    NSX_TRY(learn_rounds(b, b->ws->stream));
    NSX_HIP_TRY(b->owner, hipStreamSynchronize(b->ws->stream));
    NSX_TRY(capture(b));
    for (int i = 0; i < npairs; ++i)
        pairs[i] = (size_t)i < b->groups.size() ? b->groups[i].visit_sweeps : 0;
    return NSX_OK;
}

nsx_status nsx_batch_get_tuning(nsx_batch* b, int group, int32_t* out, int n) {
    if (!b || !b->finalized || !out || n < 0 || group < 0) return NSX_ERR_INVALID;
    if ((size_t)group >= b->groups.size()) return NSX_ERR_INVALID;
    const Group& g = b->groups[group];
    const int32_t v[5] = {g.visit_sweeps > 0 ? g.visit_sweeps : b->ws->visit_sweeps,
                          (g.drain != 0 ? g.drain : b->drain_default) > 0 ? 1 : 0, g.xcd_split, g.drain_occ,
                          b->overlap && b->groups.size() > 1 ? 1 : 0};
    for (int i = 0; i < n && i < 5; ++i) out[i] = v[i];
    return NSX_OK;
}

nsx_status nsx_batch_set_drain(nsx_batch* b, int drain) {
    if (!b || (drain != 0 && drain != 1)) return NSX_ERR_INVALID;
    b->drain_default = drain ? 1 : -1;
    for (auto& g : b->groups) g.drain = 0;
    if (b->finalized) {
        // per-round launches replay the learned round counts
        NSX_TRY(learn_rounds(b, b->ws->stream));
        NSX_HIP_TRY(b->owner, hipStreamSynchronize(b->ws->stream));
    }
    if (b->exec) NSX_TRY(capture(b));
    return NSX_OK;
}

nsx_status nsx_batch_set_overlap(nsx_batch* b, int overlap) {
    if (!b || (overlap != 0 && overlap != 1)) return NSX_ERR_INVALID;
    b->overlap = overlap != 0;
    if (b->exec) NSX_TRY(capture(b));
    return NSX_OK;
}

nsx_ctx* nsx_batch_context(nsx_batch* b) { return b ? b->ws : nullptr; }

void nsx_batch_destroy(nsx_batch* b) {
    if (!b) return;
    (void)hipDeviceSynchronize();
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_join) (void)hipEventDestroy(b->ev_join);
    for (auto& g : b->groups) {
        for (void* p : {g.d_rdesc, (void*)g.d_rtile, g.d_cdesc, (void*)g.d_ctile, (void*)g.d_ncomp, (void*)g.d_pending})
            if (p) (void)hipFree(p);
    }
    if (b->d_counts) (void)hipFree(b->d_counts);
    if (b->h_pending) (void)hipHostFree(b->h_pending);
    if (b->ws) nsx_ctx_destroy(b->ws);
    delete b;
}

}  // extern "C"
