// batch.hip -- many independent tiles per call (placeholder: sequential).
#include "nsx_internal.h"

#include <new>

using namespace nsx;

struct nsx_batch {
    nsx_ctx* ctx;
    struct Item { nsx_img marker, mask, recon, labels; int conn; double thresh; int ccl_conn; };
    std::vector<Item> items;
    int32_t* d_counts = nullptr;
    bool finalized = false;
};

extern "C" {

nsx_status nsx_batch_create(nsx_ctx* ctx, nsx_batch** batch) {
    if (!ctx || !batch) return NSX_ERR_INVALID;
    nsx_batch* b = new (std::nothrow) nsx_batch();
    if (!b) return NSX_ERR_NOMEM;
    b->ctx = ctx;
    *batch = b;
    return NSX_OK;
}

nsx_status nsx_batch_add(nsx_batch* b, const nsx_img* marker, const nsx_img* mask, nsx_img* recon_out,
                         nsx_img* labels_out, int conn, double thresh, int ccl_conn) {
    if (!b || !marker || !mask || !recon_out || !labels_out) return NSX_ERR_INVALID;
    if (b->finalized) return b->ctx->fail(NSX_ERR_INVALID, "batch already finalized");
    for (const nsx_img* im : {marker, mask, (const nsx_img*)recon_out, (const nsx_img*)labels_out})
        if (im->mem != NSX_MEM_DEVICE) return b->ctx->fail(NSX_ERR_INVALID, "batch images must be device-resident");
    b->items.push_back({*marker, *mask, *recon_out, *labels_out, conn, thresh, ccl_conn});
    return NSX_OK;
}

nsx_status nsx_batch_finalize(nsx_batch* b, int use_graph) {
    (void)use_graph;
    if (!b) return NSX_ERR_INVALID;
    if (hipMalloc((void**)&b->d_counts, sizeof(int32_t) * (b->items.size() + 1)) != hipSuccess)
        return b->ctx->fail(NSX_ERR_NOMEM, "batch counts");
    b->finalized = true;
    return NSX_OK;
}

nsx_status nsx_batch_run(nsx_batch* b, void* stream) {
    if (!b || !b->finalized) return NSX_ERR_INVALID;
    nsx_ctx* ctx = b->ctx;
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    for (size_t i = 0; i < b->items.size(); ++i) {
        auto& it = b->items[i];
        View vm{(char*)it.marker.data, it.marker.pitch_bytes, it.marker.rows, it.marker.cols, it.marker.dtype};
        View vk{(char*)it.mask.data, it.mask.pitch_bytes, it.mask.rows, it.mask.cols, it.mask.dtype};
        View vr{(char*)it.recon.data, it.recon.pitch_bytes, it.recon.rows, it.recon.cols, it.recon.dtype};
        View vl{(char*)it.labels.data, it.labels.pitch_bytes, it.labels.rows, it.labels.cols, it.labels.dtype};
        NSX_TRY(recon_run(ctx, RM_RECON, &vm, vk, vr, it.conn == 8 ? 8 : 4, s));
        CclParams p;
        p.conn = it.ccl_conn == 8 ? 8 : 4;
        p.out_mode = CO_CANON;
        p.threshold = true;
        p.thr = it.thresh;
        NSX_TRY(ccl_run(ctx, vr, vl, p, b->d_counts + i, s));
    }
    return NSX_OK;
}

nsx_status nsx_batch_counts(nsx_batch* b, int32_t* counts, int n) {
    if (!b || !counts || n < 0 || (size_t)n > b->items.size()) return NSX_ERR_INVALID;
    NSX_HIP_TRY(b->ctx, hipDeviceSynchronize());
    NSX_HIP_TRY(b->ctx, hipMemcpy(counts, b->d_counts, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    return NSX_OK;
}

void nsx_batch_destroy(nsx_batch* b) {
    if (!b) return;
    if (b->d_counts) hipFree(b->d_counts);
    delete b;
}

}  // extern "C"
