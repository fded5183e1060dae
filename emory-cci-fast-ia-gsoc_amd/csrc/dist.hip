// dist.hip -- distanceTransform (MorphologicOperations.cpp:197-275) on gfx950.
//
// The reference propagates nearest-background labels through a FIFO queue:
// every pixel starts at the point (0, 3*rows) below the image (:214), a
// background pixel holds itself (:218), and a queued pixel offers its label to
// its 8 neighbours, which take it when the distance to it, sqrt in double
// rounded to float, is strictly below the distance to their own (:90-119).  It
// ends at a fixed point of that rule: no pixel is strictly closer to a
// neighbour's label than to its own.
//
// Here the same rule runs to a fixed point in 64x64 tiles.  A tile loads its
// labels and a one-pixel halo into LDS, applies the rule in synchronous passes
// (read one LDS buffer, write the other) until no label in it changes, and
// writes the tile to its other global buffer; rounds of tile launches repeat
// until a round changes nothing.  A tile runs in a round only when it or one of
// its 8 neighbour tiles changed in the previous round.  Each tile's labels live
// in one of two global buffers, named by a per-tile index (one byte per tile
// and round parity): a tile that changes writes the buffer its index does not
// name and flips the index for the next round; the others keep theirs, so
// nothing is copied.  Every read is of the round-start state (the neighbours'
// indices of this round), so the result is deterministic.  Rounds are
// launched in batches with one host check of the last round's change word.
//
// Labels are packed (y << 16) | x inside the kernels, with DT_FAR for the
// starting point, so no integer division is needed per offer; the output turns
// them into the reference's raster index (far = rows*cols*3) or its float
// distance (:254-266).
//
// Parity is partial.  The rule has many fixed points: a pixel can be left with
// a label that no neighbour can improve although a nearer background exists
// (the "stuck" pixels of propagation distance transforms), and which fixed
// point is reached depends on the order of the offers.  The reference's
// sequential FIFO order cannot be reproduced by a parallel schedule, so at
// stuck pixels the distances of the two may differ; on the 4096x4096 disks
// image (3000 disks) 1397 of 16.7 M pixels differ, the reference's order is
// off the exact Euclidean distance at 1812 pixels and this one at 577, never
// below it.  Where several backgrounds are equally far the two also keep
// different ones (first offered by the queue vs by the sweep), so index maps
// differ there with equal distances.
#include "nsx_internal.h"

#include <cmath>
#include <utility>

namespace nsx {
namespace {

constexpr int DT = 64;
constexpr int DH = DT + 2;
constexpr uint32_t DT_FAR = 0xFFFFFFFFu;

// One tile per block, blocks dispatched to the 8 XCDs in turn; renumbered
// XCD-major, each XCD takes consecutive tiles (one region of the image).
__device__ __forceinline__ int xcd_tile(int b, int G) {
    const int x = b & 7;
    return x * (G >> 3) + min(x, G & 7) + (b >> 3);
}

__device__ __forceinline__ uint64_t dt_d2(int x, int y, uint32_t lab, int far_y) {
    const int lx = lab == DT_FAR ? 0 : (int)(lab & 0xFFFFu);
    const int ly = lab == DT_FAR ? far_y : (int)(lab >> 16);
    const int64_t dx = x - lx, dy = y - ly;
    return (uint64_t)(dx * dx + dy * dy);
}

// The reference's test: float(sqrt(double(a))) < float(sqrt(double(b))).  Below
// 2^20 distinct squares stay distinct after rounding to float (their square
// roots differ by more than 1/2048, a float ulp there is 2^-14), so the integer
// test decides; above, ties from rounding are possible and the roots are taken.
__device__ __forceinline__ bool dt_closer(uint64_t a, uint64_t b) {
    if (a >= b) return false;
    if (b <= (1ull << 20)) return true;
    return (float)sqrt((double)a) < (float)sqrt((double)b);
}

__global__ void dt_init_kernel(const char* mask, int64_t pm, uint32_t* lab, int rows, int cols, uint8_t* flags,
                               int ntiles) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (y == 0)
        for (int t = x; t < ntiles; t += gridDim.x * blockDim.x) flags[t] = 1;
    if (x >= cols) return;
    const uint8_t m = ((const uint8_t*)(mask + (int64_t)y * pm))[x];
    lab[(int64_t)y * cols + x] = m == 0 ? ((uint32_t)y << 16) | (uint32_t)x : DT_FAR;
}

// buf[b] is label buffer b; idx_cur[t] names the buffer holding tile t's
// labels at the start of this round, idx_next[t] receives it for the next.
__global__ __launch_bounds__(256) void dt_tile_kernel(uint32_t* buf0, uint32_t* buf1, int rows, int cols,
                                                      int tiles_x, int tiles_y, int far_y, const uint8_t* prev,
                                                      uint8_t* next, const uint8_t* idx_cur, uint8_t* idx_next,
                                                      int32_t* changed) {
    __shared__ uint32_t s[2][DH * DH];
    __shared__ int active;
    __shared__ uint8_t nidx[9];   // buffer index of the 3x3 tiles around this one
    const int tile = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int tx = tile % tiles_x, ty = tile / tiles_x;
    if (threadIdx.x < 9) {
        const int nx = tx + (int)threadIdx.x % 3 - 1, ny = ty + (int)threadIdx.x / 3 - 1;
        const bool in = nx >= 0 && nx < tiles_x && ny >= 0 && ny < tiles_y;
        nidx[threadIdx.x] = in ? idx_cur[ny * tiles_x + nx] : 0;
    }
    if (threadIdx.x == 0) {
        int a = 0;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int nx = tx + dx, ny = ty + dy;
                if (nx >= 0 && nx < tiles_x && ny >= 0 && ny < tiles_y && prev[ny * tiles_x + nx]) a = 1;
            }
        active = a;
    }
    __syncthreads();
    const uint8_t own = nidx[4];
    if (!active) {
        // the labels stay where they are
        if (threadIdx.x == 0) {
            next[tile] = 0;
            idx_next[tile] = own;
        }
        return;
    }
    const int x0 = tx * DT - 1, y0 = ty * DT - 1;
    // each pixel from the buffer its tile's index names; halo outside the
    // image reads as the starting point, which never wins
    for (int i = threadIdx.x; i < DH * DH; i += 256) {
        const int hx = i % DH, hy = i / DH;
        const int gx = x0 + hx, gy = y0 + hy;
        const int q = (hy == 0 ? 0 : hy == DH - 1 ? 2 : 1) * 3 + (hx == 0 ? 0 : hx == DH - 1 ? 2 : 1);
        const uint32_t* src = nidx[q] ? buf1 : buf0;
        const uint32_t v = (gx >= 0 && gx < cols && gy >= 0 && gy < rows) ? src[(int64_t)gy * cols + gx] : DT_FAR;
        s[0][i] = v;
        s[1][i] = v;
    }
    __syncthreads();
    // offers in the reference's order: up-left, up-right, up, down-left,
    // down-right, down, left, right.  Passes read one buffer and write the
    // other, so the result does not depend on the order threads run in.  A
    // pixel none of whose neighbours changed in the previous pass would find
    // the same best label as then, so it is skipped: rm[r] holds, per tile row,
    // the lanes (columns) that changed (the wave of a step owns one row).
    __shared__ uint64_t rm[2][DT + 2];
    if (threadIdx.x < DT + 2) {
        rm[0][threadIdx.x] = 0;
        rm[1][threadIdx.x] = 0;
    }
    __syncthreads();
    const int off[8] = {-DH - 1, -DH + 1, -DH, DH - 1, DH + 1, DH, -1, 1};
    const int lane = threadIdx.x & 63;
    const uint64_t near = lane == 0 ? 3ull : (7ull << (lane - 1));
    bool any = false;
    int cur = 0;
    for (int pass = 0;; ++pass) {
        const uint32_t* rd = s[cur];
        uint32_t* wr = s[cur ^ 1];
        const uint64_t* rmr = rm[cur];
        uint64_t* rmw = rm[cur ^ 1];
        int ch = 0;
        for (int k = 0; k < DT * DT / 256; ++k) {
            const int p = threadIdx.x + 256 * k;
            const int px = p & (DT - 1), py = p / DT;
            const int gx = tx * DT + px, gy = ty * DT + py;
            const int c = (py + 1) * DH + px + 1;
            const uint32_t own = rd[c];
            uint32_t best = own;
            // rows py, py+1, py+2 of rm are tile rows py-1, py, py+1; the halo
            // rows and columns only count in the first pass
            const bool look = pass == 0 || ((rmr[py] | rmr[py + 1] | rmr[py + 2]) & near) != 0;
            if (look && gx < cols && gy < rows) {
                uint64_t bd = dt_d2(gx, gy, own, far_y);
                if (bd != 0) {
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const uint32_t cand = rd[c + off[j]];
                        if (cand == best) continue;
                        const uint64_t d = dt_d2(gx, gy, cand, far_y);
                        if (dt_closer(d, bd)) {
                            best = cand;
                            bd = d;
                        }
                    }
                }
            }
            wr[c] = best;
            const uint64_t moved = __ballot(best != own);
            if (lane == 0) rmw[py + 1] = moved;
            ch |= moved != 0;
        }
        cur ^= 1;
        if (!__syncthreads_or(ch)) break;
        any = true;
    }
    // a tile that changed writes its other buffer (which no tile reads this
    // round) and names it for the next round; an unchanged one keeps its own
    if (any) {
        uint32_t* dst = own ? buf0 : buf1;
        for (int i = threadIdx.x; i < DT * DT; i += 256) {
            const int px = i & (DT - 1), py = i / DT;
            const int gx = tx * DT + px, gy = ty * DT + py;
            if (gx < cols && gy < rows) dst[(int64_t)gy * cols + gx] = s[cur][(py + 1) * DH + px + 1];
        }
    }
    if (threadIdx.x == 0) {
        next[tile] = any ? 1 : 0;
        idx_next[tile] = any ? (uint8_t)(own ^ 1) : own;
        if (any) *changed = 1;
    }
}

__global__ void dt_zero_kernel(int32_t* w, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) w[i] = 0;
}

// Every tile's labels start in buffer 0.
__global__ void dt_idx_init_kernel(uint8_t* idx, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) idx[i] = 0;
}

__global__ void dt_out_kernel(const uint32_t* buf0, const uint32_t* buf1, const uint8_t* idx, int tiles_x, char* out,
                              int64_t po, int rows, int cols, int far_y, int calc_dist) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= cols) return;
    const uint32_t* lab = idx[(y / DT) * tiles_x + x / DT] ? buf1 : buf0;
    const uint32_t l = lab[(int64_t)y * cols + x];
    char* row = out + (int64_t)y * po;
    if (calc_dist) {
        ((float*)row)[x] = (float)sqrt((double)dt_d2(x, y, l, far_y));
    } else {
        ((int32_t*)row)[x] =
            l == DT_FAR ? (int32_t)((int64_t)rows * cols * 3) : (int32_t)((l >> 16) * (uint32_t)cols + (l & 0xFFFFu));
    }
}

}  // namespace

nsx_status dist_run(nsx_ctx* ctx, const View& mask, const View& out, bool calc_dist, hipStream_t s) {
    const int rows = mask.rows, cols = mask.cols;
    if (rows == 0 || cols == 0) return NSX_OK;
    if (rows >= 65535 || cols >= 65535 || (int64_t)rows * cols * 3 > INT32_MAX)
        return ctx->fail(NSX_ERR_INVALID, "distanceTransform: rows, cols < 65535 and rows*cols*3 within int32");
    const int64_t npix = (int64_t)rows * cols;
    const int tiles_x = (cols + DT - 1) / DT, tiles_y = (rows + DT - 1) / DT;
    const int ntiles = tiles_x * tiles_y;
    // change words of a batch of rounds, at most kBatch rounds per host check
    constexpr int kBatch = 16;
    void *lab_p, *lab2_p, *flags_p, *misc_p;
    NSX_TRY(ctx->ws(WS_CCL_LABEL, sizeof(uint32_t) * (size_t)npix, &lab_p));
    NSX_TRY(ctx->ws(WS_CCL_AREA, sizeof(uint32_t) * (size_t)npix, &lab2_p));
    NSX_TRY(ctx->ws(WS_CCL_BITS, 4 * (size_t)ntiles, &flags_p));
    NSX_TRY(ctx->ws(WS_CCL_MISC, sizeof(int32_t) * kBatch, &misc_p));
    uint32_t* buf0 = (uint32_t*)lab_p;
    uint32_t* buf1 = (uint32_t*)lab2_p;
    // per tile: active flags and buffer indices, each for two round parities
    uint8_t* fl[2] = {(uint8_t*)flags_p, (uint8_t*)flags_p + ntiles};
    uint8_t* ix[2] = {(uint8_t*)flags_p + 2 * (size_t)ntiles, (uint8_t*)flags_p + 3 * (size_t)ntiles};
    int32_t* changed = (int32_t*)misc_p;
    const dim3 grid((cols + 255) / 256, rows);
    hipLaunchKernelGGL(dt_init_kernel, grid, dim3(256), 0, s, mask.data, mask.pitch, buf0, rows, cols, fl[0], ntiles);
    hipLaunchKernelGGL(dt_idx_init_kernel, dim3(std::min(1024, (ntiles + 255) / 256)), dim3(256), 0, s, ix[0],
                       ntiles);
    NSX_HIP_TRY(ctx, hipGetLastError());
    // Each round that changes something moves a label strictly closer, so the
    // rounds end (the bound only guards the loop).  Rounds after the one that
    // changed nothing find no active tile and change nothing either, so a
    // batch may overshoot; batches grow 2, 4, 8, 16.
    int64_t round = 0;
    int batch = 2;
    for (;;) {
        hipLaunchKernelGGL(dt_zero_kernel, dim3(1), dim3(64), 0, s, changed, batch);
        for (int k = 0; k < batch; ++k, ++round) {
            const int p = (int)(round & 1);
            hipLaunchKernelGGL(dt_tile_kernel, dim3(ntiles), dim3(256), 0, s, buf0, buf1, rows, cols, tiles_x,
                               tiles_y, 3 * rows, (const uint8_t*)fl[p], fl[p ^ 1], (const uint8_t*)ix[p], ix[p ^ 1],
                               changed + k);
        }
        NSX_HIP_TRY(ctx, hipGetLastError());
        NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned + 40, changed + batch - 1, sizeof(int32_t),
                                        hipMemcpyDeviceToHost, s));
        NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
        if (ctx->h_pinned[40] == 0) break;
        if (round > npix) return ctx->fail(NSX_ERR_NOT_CONVERGED, "distanceTransform did not converge");
        batch = std::min(2 * batch, kBatch);
    }
    // after `round` rounds the indices of parity round & 1 are current
    hipLaunchKernelGGL(dt_out_kernel, grid, dim3(256), 0, s, (const uint32_t*)buf0, (const uint32_t*)buf1,
                       (const uint8_t*)ix[round & 1], tiles_x, out.data, out.pitch, rows, cols, 3 * rows,
                       calc_dist ? 1 : 0);
    NSX_HIP_TRY(ctx, hipGetLastError());
    return NSX_OK;
}

}  // namespace nsx
