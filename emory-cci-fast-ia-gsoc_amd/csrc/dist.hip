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
// writes the tile to the other global buffer; rounds of tile launches repeat
// until a round changes nothing.  A tile runs in a round only when it or one of
// its 8 neighbour tiles changed in the previous round (the others are copied).
// Every read is of a finished buffer, so the result is deterministic.
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

__global__ __launch_bounds__(256) void dt_tile_kernel(const uint32_t* src, uint32_t* dst, int rows, int cols,
                                                      int tiles_x, int tiles_y, int far_y, const uint8_t* prev,
                                                      uint8_t* next, int32_t* changed) {
    __shared__ uint32_t s[2][DH * DH];
    __shared__ int active;
    const int tile = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int tx = tile % tiles_x, ty = tile / tiles_x;
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
    if (!active) {
        // carry the tile into the other buffer unchanged
        for (int i = threadIdx.x; i < DT * DT; i += 256) {
            const int gx = tx * DT + (i & (DT - 1)), gy = ty * DT + i / DT;
            if (gx < cols && gy < rows) dst[(int64_t)gy * cols + gx] = src[(int64_t)gy * cols + gx];
        }
        if (threadIdx.x == 0) next[tile] = 0;
        return;
    }
    const int x0 = tx * DT - 1, y0 = ty * DT - 1;
    // halo outside the image reads as the starting point, which never wins
    for (int i = threadIdx.x; i < DH * DH; i += 256) {
        const int gx = x0 + i % DH, gy = y0 + i / DH;
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
    for (int i = threadIdx.x; i < DT * DT; i += 256) {
        const int px = i & (DT - 1), py = i / DT;
        const int gx = tx * DT + px, gy = ty * DT + py;
        if (gx < cols && gy < rows) dst[(int64_t)gy * cols + gx] = s[cur][(py + 1) * DH + px + 1];
    }
    if (threadIdx.x == 0) {
        next[tile] = any ? 1 : 0;
        if (any) *changed = 1;
    }
}

__global__ void dt_zero_kernel(int32_t* w) {
    if (threadIdx.x == 0) *w = 0;
}

__global__ void dt_out_kernel(const uint32_t* lab, char* out, int64_t po, int rows, int cols, int far_y,
                              int calc_dist) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= cols) return;
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
    void *lab_p, *lab2_p, *flags_p, *misc_p;
    NSX_TRY(ctx->ws(WS_CCL_LABEL, sizeof(uint32_t) * (size_t)npix, &lab_p));
    NSX_TRY(ctx->ws(WS_CCL_AREA, sizeof(uint32_t) * (size_t)npix, &lab2_p));
    NSX_TRY(ctx->ws(WS_CCL_BITS, 2 * (size_t)ntiles, &flags_p));
    NSX_TRY(ctx->ws(WS_CCL_MISC, 64, &misc_p));
    uint32_t* lab = (uint32_t*)lab_p;
    uint32_t* lab2 = (uint32_t*)lab2_p;
    uint8_t* fa = (uint8_t*)flags_p;
    uint8_t* fb = fa + ntiles;
    int32_t* changed = (int32_t*)misc_p;
    const dim3 grid((cols + 255) / 256, rows);
    hipLaunchKernelGGL(dt_init_kernel, grid, dim3(256), 0, s, mask.data, mask.pitch, lab, rows, cols, fa, ntiles);
    NSX_HIP_TRY(ctx, hipGetLastError());
    // rounds read one label buffer and write the other; each round that
    // changes something moves a label strictly closer, so the rounds end (the
    // bound only guards the loop)
    for (int64_t round = 0; round <= npix; ++round) {
        hipLaunchKernelGGL(dt_zero_kernel, dim3(1), dim3(64), 0, s, changed);
        hipLaunchKernelGGL(dt_tile_kernel, dim3(ntiles), dim3(256), 0, s, (const uint32_t*)lab, lab2, rows, cols,
                           tiles_x, tiles_y, 3 * rows, (const uint8_t*)fa, fb, changed);
        NSX_HIP_TRY(ctx, hipGetLastError());
        NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned + 40, changed, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
        std::swap(lab, lab2);
        if (ctx->h_pinned[40] == 0) break;
        std::swap(fa, fb);
    }
    hipLaunchKernelGGL(dt_out_kernel, grid, dim3(256), 0, s, (const uint32_t*)lab, out.data, out.pitch, rows, cols,
                       3 * rows, calc_dist ? 1 : 0);
    NSX_HIP_TRY(ctx, hipGetLastError());
    return NSX_OK;
}

}  // namespace nsx
