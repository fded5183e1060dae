// bitplane.hip -- u8 reconstruction through its bit planes: eight labelling
// passes, whatever distances the values travel.
//
// R, the reconstruction of f = min(marker, mask) under g = mask, is found one
// bit at a time from the top.  Before the pass for bit b every pixel holds the
// prefix P of R above b.  R(p) >= t = P(p) | b exactly when p reaches, through
// {g >= t}, a marker pixel with f >= t.  Along such a path every pixel q has
// R(q) >= t, hence P(q) >= P(p): the path stays among pixels of p's class
// (same P) until it meets a marker pixel with f >= t, or a pixel of a larger
// class (whose R is above t, and which reaches a marker through {g >= t}).  So
// a pass labels the components of {g >= t} among equal classes, marks those
// holding a seed (f >= t, or a neighbour of a larger class), and sets bit b on
// them.  Hole filling is the reconstruction of the frame under g = 255 - image
// (the frame holds 255, so the image's border pixels are seeds), output
// 255 - R.  Each pass is one elementwise kernel (class labels L = P + 1 on the
// foreground, 0 elsewhere, and seed weights) and one union-find labelling of L
// (equal-valued neighbours) that folds each component's largest weight into P.
// No reference counterpart: the reference propagates
// (MorphologicOperations.cpp:665-838, 1859-1926); the result is the same.
#include "nsx_internal.h"

namespace nsx {
namespace {

// One pass's labelling input and seeds.  FILL: g = 255 - mask, border
// pixels are seeds; RECON: g = mask, seeds where min(marker, mask) >= t.
template <bool FILL, int CONN>
__global__ __launch_bounds__(256) void bp_prep_kernel(const char* marker, int64_t pm, const char* mask, int64_t pk,
                                                      const char* P, int64_t pp, char* L, int64_t pl, char* W,
                                                      int64_t pw, int rows, int cols, uint32_t bit) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= cols || y >= rows) return;
    const uint32_t m = ((const uint8_t*)(mask + (int64_t)y * pk))[x];
    const uint32_t g = FILL ? 255u - m : m;
    const uint32_t p = ((const uint8_t*)(P + (int64_t)y * pp))[x];
    const uint32_t t = p | bit;
    const bool fg = g >= t;
    bool seed = false;
    if (fg) {
        if (FILL) {
            seed = x == 0 || y == 0 || x == cols - 1 || y == rows - 1;
        } else {
            const uint32_t f = min(m, (uint32_t)((const uint8_t*)(marker + (int64_t)y * pm))[x]);
            seed = f >= t;
        }
        // a neighbour of a larger class: its value is above t and it reaches a
        // marker through {g >= t}
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                if ((dy == 0 && dx == 0) || (CONN == 4 && dy != 0 && dx != 0)) continue;
                const int yy = y + dy, xx = x + dx;
                if (yy < 0 || yy >= rows || xx < 0 || xx >= cols) continue;
                seed = seed || ((const uint8_t*)(P + (int64_t)yy * pp))[xx] > p;
            }
        }
    }
    ((uint8_t*)(L + (int64_t)y * pl))[x] = (uint8_t)(fg ? p + 1u : 0u);
    ((uint8_t*)(W + (int64_t)y * pw))[x] = (uint8_t)(seed ? t : 0u);
}

// The same for four pixels per thread, with 4-byte loads and stores: rows of
// every plane 4-byte aligned and cols a multiple of 4.  The neighbours left of
// x0 and right of x0 + 3 come from single bytes.
template <bool FILL, int CONN>
__global__ __launch_bounds__(256) void bp_prep4_kernel(const char* marker, int64_t pm, const char* mask, int64_t pk,
                                                       const char* P, int64_t pp, char* L, int64_t pl, char* W,
                                                       int64_t pw, int rows, int cols, uint32_t bit) {
    const int x0 = (blockIdx.x * 64 + (threadIdx.x & 63)) * 4;
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x0 >= cols || y >= rows) return;
    const uint32_t mw = *(const uint32_t*)(mask + (int64_t)y * pk + x0);
    const uint32_t kw = FILL ? 0u : *(const uint32_t*)(marker + (int64_t)y * pm + x0);
    // P around the four pixels: rows y-1, y, y+1, columns x0-1 .. x0+4
    // (outside the image: 0, which no class exceeds... a prefix 0 never counts
    // as larger, so it stands for "no neighbour")
    uint32_t pr[3][6];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int yy = y + r - 1;
        const bool in = yy >= 0 && yy < rows;
        const uint8_t* row = (const uint8_t*)(P + (int64_t)(in ? yy : y) * pp);
        const uint32_t w = in ? *(const uint32_t*)(row + x0) : 0u;
        pr[r][0] = (in && x0 > 0) ? row[x0 - 1] : 0u;
#pragma unroll
        for (int j = 0; j < 4; ++j) pr[r][1 + j] = (w >> (8 * j)) & 0xFFu;
        pr[r][5] = (in && x0 + 4 < cols) ? row[x0 + 4] : 0u;
    }
    uint32_t lw = 0u, ww = 0u;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int x = x0 + j;
        const uint32_t m = (mw >> (8 * j)) & 0xFFu;
        const uint32_t g = FILL ? 255u - m : m;
        const uint32_t p = pr[1][1 + j];
        const uint32_t t = p | bit;
        const bool fg = g >= t;
        bool seed;
        if (FILL) {
            seed = x == 0 || y == 0 || x == cols - 1 || y == rows - 1;
        } else {
            seed = min(m, (kw >> (8 * j)) & 0xFFu) >= t;
        }
        // neighbours outside the image read 0 above: never larger than p
        seed = seed || pr[1][j] > p || pr[1][j + 2] > p || pr[0][1 + j] > p || pr[2][1 + j] > p;
        if (CONN == 8) seed = seed || pr[0][j] > p || pr[0][j + 2] > p || pr[2][j] > p || pr[2][j + 2] > p;
        seed = seed && fg;
        lw |= (fg ? p + 1u : 0u) << (8 * j);
        ww |= (seed ? t : 0u) << (8 * j);
    }
    *(uint32_t*)(L + (int64_t)y * pl + x0) = lw;
    *(uint32_t*)(W + (int64_t)y * pw + x0) = ww;
}

__global__ void bp_zero_kernel(char* P, int64_t pp, int rows, int cols) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x < cols && y < rows) ((uint8_t*)(P + (int64_t)y * pp))[x] = 0;
}

}  // namespace

nsx_status recon_bitplane_u8(nsx_ctx* ctx, int mode, const View* marker, const View& mask, const View& out, int conn,
                             hipStream_t s) {
    if (mask.rows <= 0 || mask.cols <= 0) return NSX_OK;
    if (mask.dtype != NSX_U8 || out.dtype != NSX_U8 || (mode != RM_RECON && mode != RM_FILL) ||
        (mode == RM_RECON && !marker))
        return ctx->fail(NSX_ERR_UNSUPPORTED, "bit-plane reconstruction is for u8 imreconstruct and gray fill");
    conn = conn == 8 ? 8 : 4;
    const size_t pitch = (size_t)mask.cols;
    void *pl, *pw;
    NSX_TRY(ctx->ws(WS_BP_L, pitch * (size_t)mask.rows, &pl));
    NSX_TRY(ctx->ws(WS_BP_W, pitch * (size_t)mask.rows, &pw));
    const View L{(char*)pl, (int64_t)pitch, mask.rows, mask.cols, NSX_U8};
    const View W{(char*)pw, (int64_t)pitch, mask.rows, mask.cols, NSX_U8};
    const dim3 grid((mask.cols + 63) / 64, (mask.rows + 3) / 4), block(256);
    const bool fill = mode == RM_FILL;
    const char* mk = marker ? marker->data : nullptr;
    int64_t pm = marker ? marker->pitch : 0;
    View g = mask;
    // `out` is cleared before the passes read the inputs: an input that is
    // also the output is read from a copy
    const bool alias_m = mk && mk == out.data, alias_k = mask.data == out.data;
    if (alias_m || alias_k) {
        void* pc;
        NSX_TRY(ctx->ws(WS_BP_COPY, pitch * (size_t)mask.rows, &pc));
        NSX_HIP_TRY(ctx, copy_rows(pc, pitch, out.data, out.pitch, pitch, (size_t)mask.rows, hipMemcpyDeviceToDevice, s));
        if (alias_m) { mk = (const char*)pc; pm = (int64_t)pitch; }
        if (alias_k) { g.data = (char*)pc; g.pitch = (int64_t)pitch; }
    }
    hipLaunchKernelGGL(bp_zero_kernel, grid, block, 0, s, out.data, out.pitch, out.rows, out.cols);
    // four pixels per thread when every plane's rows are 4-byte aligned
    const auto a4 = [](const char* p, int64_t pitch) { return ((uintptr_t)p % 4 == 0) && pitch % 4 == 0; };
    const bool wide = mask.cols % 4 == 0 && a4(g.data, g.pitch) && a4(out.data, out.pitch) && a4(L.data, L.pitch) &&
                      a4(W.data, W.pitch) && (fill || a4(mk, pm));
    const dim3 grid4((mask.cols / 4 + 63) / 64, (mask.rows + 3) / 4);
    for (int b = 7; b >= 0; --b) {
        const uint32_t bit = 1u << b;
        if (wide) {
            if (fill && conn == 8)
                hipLaunchKernelGGL((bp_prep4_kernel<true, 8>), grid4, block, 0, s, mk, pm, g.data, g.pitch, out.data,
                                   out.pitch, L.data, L.pitch, W.data, W.pitch, mask.rows, mask.cols, bit);
            else if (fill)
                hipLaunchKernelGGL((bp_prep4_kernel<true, 4>), grid4, block, 0, s, mk, pm, g.data, g.pitch, out.data,
                                   out.pitch, L.data, L.pitch, W.data, W.pitch, mask.rows, mask.cols, bit);
            else if (conn == 8)
                hipLaunchKernelGGL((bp_prep4_kernel<false, 8>), grid4, block, 0, s, mk, pm, g.data, g.pitch, out.data,
                                   out.pitch, L.data, L.pitch, W.data, W.pitch, mask.rows, mask.cols, bit);
            else
                hipLaunchKernelGGL((bp_prep4_kernel<false, 4>), grid4, block, 0, s, mk, pm, g.data, g.pitch, out.data,
                                   out.pitch, L.data, L.pitch, W.data, W.pitch, mask.rows, mask.cols, bit);
        } else if (fill && conn == 8)
            hipLaunchKernelGGL((bp_prep_kernel<true, 8>), grid, block, 0, s, mk, pm, g.data, g.pitch, out.data,
                               out.pitch, L.data, L.pitch, W.data, W.pitch, mask.rows, mask.cols, bit);
        else if (fill)
            hipLaunchKernelGGL((bp_prep_kernel<true, 4>), grid, block, 0, s, mk, pm, g.data, g.pitch, out.data,
                               out.pitch, L.data, L.pitch, W.data, W.pitch, mask.rows, mask.cols, bit);
        else if (conn == 8)
            hipLaunchKernelGGL((bp_prep_kernel<false, 8>), grid, block, 0, s, mk, pm, g.data, g.pitch, out.data,
                               out.pitch, L.data, L.pitch, W.data, W.pitch, mask.rows, mask.cols, bit);
        else
            hipLaunchKernelGGL((bp_prep_kernel<false, 4>), grid, block, 0, s, mk, pm, g.data, g.pitch, out.data,
                               out.pitch, L.data, L.pitch, W.data, W.pitch, mask.rows, mask.cols, bit);
        NSX_HIP_TRY(ctx, hipGetLastError());
        NSX_TRY(ccl_class_fold_run(ctx, L, W, out, conn, s));
    }
    if (fill) NSX_TRY(invert_run(ctx, out, s));
    return NSX_OK;
}

}  // namespace nsx
