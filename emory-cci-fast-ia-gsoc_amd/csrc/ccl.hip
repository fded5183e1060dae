// ccl.hip -- union-find connected-component labelling, area filter and
// canonical relabel on gfx950.
//
// Semantics follow ConnComponents.cpp:52-118,205-239 (bwlabel2 / bwareaopen2):
// 4/8-neighbours of EQUAL nonzero value are joined; each component's label is
// the raster index of its first pixel (the union-find root is always the
// smaller index, ConnComponents.cpp:34-43, ccl_uf.cu:40-62); the canonical
// relabel numbers those roots 1.. in raster order with background 0.
//
// Kernels (tile = 64 columns x 32 rows, 256 threads, 8 pixels per thread):
//   K1 local    union-find in LDS with ds_min; per local component append
//               (root raster index, pixel count) to a compact table; write the
//               local root index only for tile-boundary pixels and for the
//               roots themselves (one sparse int32 array `lab`).
//   K2 merge    pixel pairs that cross a tile edge are joined with global
//               atomicMin union-find over `lab`.
//   K3 area     per table entry: atomicAdd its count into area[find(root)].
//   K4 keep     per table entry that is a global root: keep it if the area is
//               in [min, max); set its bit in a raster-order bitmap.
//   K5 scan     popcount prefix over the bitmap -> rank of every kept root.
//   K6 final    recompute the tile-local labelling in LDS from the input (1 B
//               per pixel instead of re-reading a 4 B label array), map each
//               local root to its global root once, write the output.
// HBM bytes per pixel: K1 reads 1, K6 reads 1 and writes 4 (u8 input, int32
// output); everything else touches only boundary pixels and roots.
#include "nsx_internal.h"

#include <algorithm>
#include <cmath>

namespace nsx {
namespace {

constexpr int CW = 64;               // tile columns
constexpr int CH = 32;               // tile rows
constexpr int CN = CW * CH;          // pixels per tile
constexpr int CTHREADS = 256;
constexpr int CPT = CN / CTHREADS;   // pixels per thread
constexpr uint32_t BGL = 0xFFFFFFFFu;

struct CclArgs {
    const char* in; int64_t pin; int in_dtype;
    bool thresh; int64_t thr_i; float thr_f;
    int rows, cols;
    int64_t npix;
    int32_t* lab;        // sparse: boundary pixels + local roots
    int32_t* area;       // sparse: roots
    uint2* entries;      // (root raster index, count)
    int32_t* nentries;   // [0] entries, [1] kept components
    uint32_t* bits;      // kept-root bitmap
    uint32_t* wordpre;   // exclusive popcount prefix inside each 1024-word block
    uint32_t* blockpre;  // exclusive prefix of block totals
    int64_t nwords;
    int min_size, max_size;
    bool use_area;
    char* out; int64_t pout; int out_mode;
};

// Pixel value used for connectivity: raw u8, or (in > thr) for the fused
// threshold (OpenCV `Mat > k`, 255/0 -> here 1/0).  The threshold is
// pre-converted on the host to the input type's comparison (thr_i for
// integers: v > k  <=>  v > floor(k); thr_f for float).
template <typename InT, bool THR>
__device__ __forceinline__ uint32_t pix(const CclArgs& A, int y, int x) {
    const InT v = reinterpret_cast<const InT*>(A.in + (int64_t)y * A.pin)[x];
    if constexpr (!THR) {
        return (uint32_t)v;
    } else if constexpr (sizeof(InT) == 4 && !__is_same(InT, int32_t)) {
        return v > A.thr_f ? 1u : 0u;
    } else {
        return (int64_t)v > A.thr_i ? 1u : 0u;
    }
}

// Find with path halving: each visited node is re-pointed to its grandparent
// (an ancestor), which concurrent atomicMin unions tolerate.
__device__ __forceinline__ uint32_t lds_find(uint32_t* par, uint32_t x) {
    uint32_t p = par[x];
    while (p != x) {
        const uint32_t g = par[p];
        if (g != p) par[x] = g;
        x = p;
        p = g;
    }
    return x;
}

__device__ __forceinline__ void lds_union(uint32_t* par, uint32_t a, uint32_t b) {
    while (true) {
        a = lds_find(par, a);
        b = lds_find(par, b);
        if (a == b) return;
        if (a < b) { uint32_t t = a; a = b; b = t; }
        const uint32_t old = atomicMin(&par[a], b);
        if (old == a) return;
        a = old;
    }
}

__device__ __forceinline__ int32_t g_find(const int32_t* lab, int32_t x) {
    int32_t p = lab[x];
    while (p != x) { x = p; p = lab[x]; }
    return x;
}

__device__ __forceinline__ void g_union(int32_t* lab, int32_t a, int32_t b) {
    while (true) {
        a = g_find(lab, a);
        b = g_find(lab, b);
        if (a == b) return;
        if (a < b) { int32_t t = a; a = b; b = t; }
        const int32_t old = atomicMin(&lab[a], b);
        if (old == a) return;
        a = old;
    }
}

// Tile-local labelling shared by K1 and K6.  On return val[] holds the pixel
// values and par[i] the local root of pixel i (BGL for background).
template <int CONN, typename InT, bool THR>
__device__ __forceinline__ void local_label(const CclArgs& A, int x0, int y0, uint32_t* val, uint32_t* par) {
    const int tid = threadIdx.x;
    const int lx = tid & 63, w = tid >> 6;
    const int gx = x0 + lx;
    // all loads first (independent, in flight together), then the LDS writes
    uint32_t v[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int gy = y0 + w + 4 * j;
        v[j] = (gx < A.cols && gy < A.rows) ? pix<InT, THR>(A, gy, gx) : 0u;
    }
    // Each wave holds whole tile rows (lane = column): a pixel's initial parent
    // is the first pixel of its horizontal run of equal values, found with a
    // ballot of run starts, so the unions below only join runs of adjacent rows.
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int i = (w + 4 * j) * CW + lx;
        const uint32_t left = __shfl_up(v[j], 1, 64);
        const bool start = v[j] != 0u && (lx == 0 || left != v[j]);
        const uint64_t starts = __ballot(start);
        const uint64_t upto = lx == 63 ? ~0ull : ((2ull << lx) - 1ull);
        const int rs = 63 - __clzll(starts & upto);
        val[i] = v[j];
        par[i] = v[j] ? (uint32_t)((w + 4 * j) * CW + rs) : BGL;
    }
    __syncthreads();
    // Join each pair of overlapping runs once, at the leftmost column where they
    // touch: a vertical (or diagonal) link that a left neighbour in the same run
    // already makes through the same run above is skipped.
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int ly = w + 4 * j;
        const int i = ly * CW + lx;
        const uint32_t vv = val[i];
        if (!vv || ly == 0) continue;
        const int u = i - CW;
        const uint32_t vu = val[u];
        const uint32_t vl = lx > 0 ? val[i - 1] : 0u;
        const uint32_t vul = lx > 0 ? val[u - 1] : 0u;
        const uint32_t vur = lx < CW - 1 ? val[u + 1] : 0u;
        const uint32_t vr = lx < CW - 1 ? val[i + 1] : 0u;
        if (vu == vv && !(vl == vv && vul == vv)) lds_union(par, i, u);
        if (CONN == 8) {
            if (vul == vv && vu != vv && vl != vv) lds_union(par, i, u - 1);
            if (vur == vv && vu != vv && vr != vv) lds_union(par, i, u + 1);
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int i = (w + 4 * j) * CW + lx;
        if (par[i] != BGL) par[i] = lds_find(par, i);
    }
    __syncthreads();
}

__device__ __forceinline__ int64_t gidx(const CclArgs& A, int y0, int x0, uint32_t li) {
    return (int64_t)(y0 + (int)(li >> 6)) * A.cols + x0 + (int)(li & 63);
}

template <int CONN, typename InT, bool THR>
__global__ __launch_bounds__(CTHREADS) void ccl_local_kernel(CclArgs A) {
    __shared__ uint32_t val[CN];
    __shared__ uint32_t par[CN];
    __shared__ uint32_t cnt[CN];
    __shared__ uint32_t nroots, base;
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * CW, y0 = blockIdx.y * CH;
    for (int i = tid; i < CN; i += CTHREADS) cnt[i] = 0;
    if (tid == 0) nroots = 0;
    local_label<CONN, InT, THR>(A, x0, y0, val, par);
    const int lx = tid & 63, w = tid >> 6;
    const int gx = x0 + lx;
    // pixel counts per local root, bitmap words of this tile cleared
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int ly = w + 4 * j;
        const int i = ly * CW + lx;
        const int gy = y0 + ly;
        if (gx < A.cols && gy < A.rows) {
            const int64_t g = (int64_t)gy * A.cols + gx;
            if (A.bits && ((g & 31) == 0 || lx == 0)) A.bits[g >> 5] = 0u;
        }
        if (par[i] != BGL) atomicAdd(&cnt[par[i]], 1u);
    }
    __syncthreads();
    uint32_t slot[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int i = (w + 4 * j) * CW + lx;
        slot[j] = (par[i] == (uint32_t)i) ? atomicAdd(&nroots, 1u) : 0xFFFFFFFFu;
    }
    __syncthreads();
    if (tid == 0) base = nroots ? (uint32_t)atomicAdd(&A.nentries[0], (int)nroots) : 0u;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int ly = w + 4 * j;
        const int i = ly * CW + lx;
        const uint32_t r = par[i];
        if (r == BGL) continue;
        const int64_t g = (int64_t)(y0 + ly) * A.cols + gx;
        if (slot[j] != 0xFFFFFFFFu) {
            A.entries[base + slot[j]] = make_uint2((uint32_t)g, cnt[i]);
            A.lab[g] = (int32_t)g;
            if (A.use_area) A.area[g] = 0;
        } else if (ly == 0 || ly == CH - 1 || lx == 0 || lx == CW - 1) {
            A.lab[g] = (int32_t)gidx(A, y0, x0, r);
        }
    }
}

// K2: one thread per tile-edge pixel whose backward neighbours cross the edge:
// threads 0..63 the top row (up, up-left, up-right), 64..95 the left column
// (left, and up-left below the top row), 96..127 the right column (up-right
// below the top row).
template <int CONN, typename InT, bool THR>
__global__ __launch_bounds__(128) void ccl_merge_kernel(CclArgs A) {
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * CW, y0 = blockIdx.y * CH;
    int y, x;
    if (tid < 64) { y = y0; x = x0 + tid; }
    else if (tid < 96) { y = y0 + (tid - 64); x = x0; }
    else { y = y0 + 1 + (tid - 96); x = x0 + CW - 1; if (tid == 127) return; }
    if (x >= A.cols || y >= A.rows) return;
    const uint32_t v = pix<InT, THR>(A,y, x);
    if (!v) return;
    const int64_t g = (int64_t)y * A.cols + x;
    auto join = [&](int qy, int qx) {
        if (qy < 0 || qx < 0 || qx >= A.cols) return;
        if (pix<InT, THR>(A,qy, qx) != v) return;
        const int64_t q = (int64_t)qy * A.cols + qx;
        g_union(A.lab, A.lab[g], A.lab[q]);
    };
    if (tid < 64) {
        if (y > 0) {
            join(y - 1, x);
            if (CONN == 8) { join(y - 1, x - 1); join(y - 1, x + 1); }
        }
    } else if (tid < 96) {
        if (x > 0) {
            join(y, x - 1);
            if (CONN == 8 && y > y0) join(y - 1, x - 1);
        }
    } else {
        if (CONN == 8 && x + 1 < A.cols) join(y - 1, x + 1);
    }
}

__global__ void ccl_area_kernel(CclArgs A) {
    const int n = A.nentries[0];
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const uint2 en = A.entries[e];
        const int32_t r = g_find(A.lab, (int32_t)en.x);
        atomicAdd(&A.area[r], (int)en.y);
    }
}

__global__ void ccl_keep_kernel(CclArgs A) {
    const int n = A.nentries[0];
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const int32_t g = (int32_t)A.entries[e].x;
        if (A.lab[g] != g) continue;
        bool keep = true;
        if (A.use_area) {
            const int a = A.area[g];
            keep = a >= A.min_size && a < A.max_size;
        }
        if (keep) {
            if (A.bits) atomicOr(&A.bits[g >> 5], 1u << (g & 31));
            atomicAdd(&A.nentries[1], 1);
        }
    }
}

// K5a: 256 threads x 4 words; exclusive popcount prefix per 1024-word block.
__global__ __launch_bounds__(256) void scan_words_kernel(const uint32_t* bits, int64_t nwords,
                                                        uint32_t* wordpre, uint32_t* blocksum) {
    __shared__ uint32_t wsum[4];
    const int tid = threadIdx.x;
    const int64_t w0 = (int64_t)blockIdx.x * 1024 + tid * 4;
    uint32_t c[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        c[k] = (w0 + k < nwords) ? (uint32_t)__popc(bits[w0 + k]) : 0u;
        s += c[k];
    }
    // inclusive wave scan of s
    const int lane = tid & 63, wv = tid >> 6;
    uint32_t inc = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) wsum[wv] = inc;
    __syncthreads();
    uint32_t off = 0;
    for (int k = 0; k < wv; ++k) off += wsum[k];
    uint32_t ex = off + inc - s;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (w0 + k < nwords) wordpre[w0 + k] = ex;
        ex += c[k];
    }
    if (tid == 255) blocksum[blockIdx.x] = off + inc;
}

// K5b: one block of 1024 threads scans the block totals in place (exclusive).
__global__ __launch_bounds__(1024) void scan_blocks_kernel(uint32_t* blocksum, int nb) {
    __shared__ uint32_t part[1024];
    const int tid = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int b0 = tid * per;
    uint32_t s = 0;
    for (int k = 0; k < per; ++k) if (b0 + k < nb) s += blocksum[b0 + k];
    part[tid] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const uint32_t t = tid >= o ? part[tid - o] : 0u;
        __syncthreads();
        part[tid] += t;
        __syncthreads();
    }
    uint32_t ex = part[tid] - s;
    for (int k = 0; k < per; ++k) {
        if (b0 + k < nb) {
            const uint32_t v = blocksum[b0 + k];
            blocksum[b0 + k] = ex;
            ex += v;
        }
    }
}

__device__ __forceinline__ bool kept(const CclArgs& A, int32_t r) {
    return (A.bits[r >> 5] >> (r & 31)) & 1u;
}
__device__ __forceinline__ int32_t rank1(const CclArgs& A, int32_t r) {
    const uint32_t w = A.bits[r >> 5];
    return (int32_t)(A.blockpre[r >> 15] + A.wordpre[r >> 5] + __popc(w & ((1u << (r & 31)) - 1u))) + 1;
}

template <int CONN, typename InT, bool THR>
__global__ __launch_bounds__(CTHREADS) void ccl_final_kernel(CclArgs A) {
    __shared__ uint32_t val[CN];
    __shared__ uint32_t par[CN];
    __shared__ int32_t outv[CN];
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * CW, y0 = blockIdx.y * CH;
    local_label<CONN, InT, THR>(A, x0, y0, val, par);
    const int lx = tid & 63, w = tid >> 6;
    const int gx = x0 + lx;
    // each local root resolves its global root once
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int ly = w + 4 * j;
        const int i = ly * CW + lx;
        if (par[i] != (uint32_t)i) continue;
        const int32_t r = g_find(A.lab, (int32_t)((int64_t)(y0 + ly) * A.cols + gx));
        int32_t o;
        switch (A.out_mode) {
            case CO_ROOT: o = r; break;
            case CO_ROOT_FILTERED: o = kept(A, r) ? r : -1; break;
            case CO_CANON: o = kept(A, r) ? rank1(A, r) : 0; break;
            default: o = kept(A, r) ? 255 : 0; break;
        }
        outv[i] = o;
    }
    __syncthreads();
    const int32_t bg = (A.out_mode == CO_ROOT || A.out_mode == CO_ROOT_FILTERED) ? -1 : 0;
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int ly = w + 4 * j;
        const int gy = y0 + ly;
        if (gx >= A.cols || gy >= A.rows) continue;
        const uint32_t r = par[ly * CW + lx];
        const int32_t o = (r == BGL) ? bg : outv[r];
        char* row = A.out + (int64_t)gy * A.pout;
        if (A.out_mode == CO_FLAT_U8) ((uint8_t*)row)[gx] = (uint8_t)o;
        else ((int32_t*)row)[gx] = o;
    }
}

__global__ void zero_i32_kernel(int32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

template <int CONN, typename InT, bool THR>
nsx_status ccl_run_conn(nsx_ctx* ctx, CclArgs A, hipStream_t s) {
    const dim3 tiles((A.cols + CW - 1) / CW, (A.rows + CH - 1) / CH);
    const int stream_grid = 4 * ctx->cu_count;
    hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(64), 0, s, A.nentries, 2);
    int h = ctx->tic(TK_CCL_LOCAL, s);
    hipLaunchKernelGGL((ccl_local_kernel<CONN, InT, THR>), tiles, dim3(CTHREADS), 0, s, A);
    ctx->toc(h, s);
    h = ctx->tic(TK_CCL_MERGE, s);
    hipLaunchKernelGGL((ccl_merge_kernel<CONN, InT, THR>), tiles, dim3(128), 0, s, A);
    ctx->toc(h, s);
    h = ctx->tic(TK_CCL_SMALL, s);
    if (A.use_area) hipLaunchKernelGGL(ccl_area_kernel, dim3(stream_grid), dim3(256), 0, s, A);
    hipLaunchKernelGGL(ccl_keep_kernel, dim3(stream_grid), dim3(256), 0, s, A);
    if (A.out_mode == CO_CANON) {
        const int nb = (int)((A.nwords + 1023) / 1024);
        hipLaunchKernelGGL(scan_words_kernel, dim3(nb), dim3(256), 0, s, A.bits, A.nwords, A.wordpre, A.blockpre);
        hipLaunchKernelGGL(scan_blocks_kernel, dim3(1), dim3(1024), 0, s, A.blockpre, nb);
    }
    ctx->toc(h, s);
    h = ctx->tic(TK_CCL_FINAL, s);
    hipLaunchKernelGGL((ccl_final_kernel<CONN, InT, THR>), tiles, dim3(CTHREADS), 0, s, A);
    ctx->toc(h, s);
    NSX_HIP_TRY(ctx, hipGetLastError());
    return NSX_OK;
}

// ---- labelled-input area filter (bwareaopen2 with labeled=true) ---------------------
__global__ void area_count_kernel(const char* in, int64_t pin, int rows, int cols, int32_t* area, int64_t npix,
                                  int32_t* bad) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= cols) return;
    const int32_t l = ((const int32_t*)(in + (int64_t)y * pin))[x];
    if (l == -1) return;
    if (l < 0 || (int64_t)l >= npix) { atomicOr(bad, 1); return; }
    atomicAdd(&area[l], 1);
}

__global__ void area_keep_count_kernel(const int32_t* area, int64_t npix, int minv, int maxv, int32_t* count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += (int64_t)gridDim.x * blockDim.x) {
        const int a = area[i];
        if (a > 0 && a >= minv && a < maxv) atomicAdd(count, 1);
    }
}

__global__ void area_apply_kernel(const char* in, int64_t pin, int rows, int cols, const int32_t* area,
                                  int64_t npix, int minv, int maxv, char* out, int64_t pout, int flatten) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= cols) return;
    const int32_t l = ((const int32_t*)(in + (int64_t)y * pin))[x];
    bool keep = false;
    if (l >= 0 && (int64_t)l < npix) {
        const int a = area[l];
        keep = a >= minv && a < maxv;
    }
    if (flatten) ((uint8_t*)(out + (int64_t)y * pout))[x] = keep ? 255 : 0;
    else ((int32_t*)(out + (int64_t)y * pout))[x] = keep ? l : -1;
}

__global__ void zero_words_kernel(uint32_t* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// ---- relabel of an arbitrary int32 label image ---------------------------------------
__global__ void root_bits_kernel(const char* lab, int64_t pl, int rows, int cols, uint32_t* bits, int32_t* count) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= cols) return;
    const int64_t g = (int64_t)y * cols + x;
    const int32_t l = ((const int32_t*)(lab + (int64_t)y * pl))[x];
    if ((int64_t)l == g) {
        atomicOr(&bits[g >> 5], 1u << (g & 31));
        atomicAdd(count, 1);
    }
}

__global__ void relabel_apply_kernel(char* lab, int64_t pl, int rows, int cols, int64_t npix, const uint32_t* bits,
                                     const uint32_t* wordpre, const uint32_t* blockpre) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= cols) return;
    int32_t* p = ((int32_t*)(lab + (int64_t)y * pl)) + x;
    const int32_t l = *p;
    int32_t o = 0;
    if (l >= 0 && (int64_t)l < npix && ((bits[l >> 5] >> (l & 31)) & 1u))
        o = (int32_t)(blockpre[l >> 15] + wordpre[l >> 5] + __popc(bits[l >> 5] & ((1u << (l & 31)) - 1u))) + 1;
    *p = o;
}

__global__ void threshold_kernel(const char* in, int64_t pin, int dtype, int rows, int cols, double k, char* out,
                                 int64_t pout) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= cols) return;
    const char* row = in + (int64_t)y * pin;
    double v;
    if (dtype == NSX_U8) v = ((const uint8_t*)row)[x];
    else if (dtype == NSX_I32) v = ((const int32_t*)row)[x];
    else v = ((const float*)row)[x];
    ((uint8_t*)(out + (int64_t)y * pout))[x] = v > k ? 255 : 0;
}

__global__ void fill_i32_kernel(char* out, int64_t pout, int rows, int cols, int32_t v) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= cols) return;
    ((int32_t*)(out + (int64_t)y * pout))[x] = v;
}

}  // namespace

nsx_status ccl_run(nsx_ctx* ctx, const View& in, const View& out, const CclParams& p, int32_t* ncomp_dev,
                   hipStream_t s) {
    CclArgs A{};
    A.in = in.data; A.pin = in.pitch; A.in_dtype = in.dtype;
    A.thresh = p.threshold;
    // v > k for an integer v  <=>  v > floor(k); for float, compare in float
    // (the reference compares a float Mat with a scalar converted to float).
    A.thr_i = (int64_t)std::floor(std::max(-9.0e18, std::min(9.0e18, p.thr)));
    A.thr_f = (float)p.thr;
    A.rows = in.rows; A.cols = in.cols;
    A.npix = (int64_t)in.rows * in.cols;
    if (A.npix == 0) return NSX_OK;
    if (A.npix >= (int64_t)INT32_MAX) return ctx->fail(NSX_ERR_INVALID, "image too large for int32 labels");
    A.min_size = p.min_size; A.max_size = p.max_size; A.use_area = p.use_area;
    A.out = out.data; A.pout = out.pitch; A.out_mode = p.out_mode;
    const int64_t ntile = (int64_t)((A.cols + CW - 1) / CW) * ((A.rows + CH - 1) / CH);
    A.nwords = (A.npix + 31) / 32;
    const int64_t nb = (A.nwords + 1023) / 1024;
    void *lab, *area = nullptr, *ent, *misc, *bits = nullptr, *scan = nullptr;
    NSX_TRY(ctx->ws(WS_CCL_LABEL, sizeof(int32_t) * (size_t)A.npix, &lab));
    if (A.use_area) NSX_TRY(ctx->ws(WS_CCL_AREA, sizeof(int32_t) * (size_t)A.npix, &area));
    const int64_t max_entries = std::min<int64_t>(A.npix, A.npix / 2 + ntile + 64);
    NSX_TRY(ctx->ws(WS_CCL_ROOTS, sizeof(uint2) * (size_t)max_entries, &ent));
    NSX_TRY(ctx->ws(WS_CCL_MISC, 64, &misc));
    const bool need_bits = p.out_mode != CO_ROOT;
    if (need_bits) NSX_TRY(ctx->ws(WS_CCL_BITS, sizeof(uint32_t) * (size_t)A.nwords, &bits));
    if (p.out_mode == CO_CANON) NSX_TRY(ctx->ws(WS_CCL_SCAN, sizeof(uint32_t) * (size_t)(A.nwords + nb + 1), &scan));
    A.lab = (int32_t*)lab; A.area = (int32_t*)area; A.entries = (uint2*)ent;
    A.nentries = (int32_t*)misc;
    A.bits = (uint32_t*)bits;
    A.wordpre = (uint32_t*)scan;
    A.blockpre = scan ? (uint32_t*)scan + A.nwords : nullptr;
    nsx_status st;
    if (!p.threshold) {
        if (in.dtype != NSX_U8) return ctx->fail(NSX_ERR_INVALID, "labelling needs u8 input or a threshold");
        st = p.conn == 8 ? ccl_run_conn<8, uint8_t, false>(ctx, A, s) : ccl_run_conn<4, uint8_t, false>(ctx, A, s);
    } else if (in.dtype == NSX_U8) {
        st = p.conn == 8 ? ccl_run_conn<8, uint8_t, true>(ctx, A, s) : ccl_run_conn<4, uint8_t, true>(ctx, A, s);
    } else if (in.dtype == NSX_I32) {
        st = p.conn == 8 ? ccl_run_conn<8, int32_t, true>(ctx, A, s) : ccl_run_conn<4, int32_t, true>(ctx, A, s);
    } else {
        st = p.conn == 8 ? ccl_run_conn<8, float, true>(ctx, A, s) : ccl_run_conn<4, float, true>(ctx, A, s);
    }
    if (st != NSX_OK) return st;
    if (ncomp_dev) NSX_HIP_TRY(ctx, hipMemcpyAsync(ncomp_dev, A.nentries + 1, sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    return NSX_OK;
}

nsx_status ccl_area_labeled(nsx_ctx* ctx, const View& in, const View& out, int min_size, int max_size,
                            bool flatten, int32_t* count_dev, hipStream_t s) {
    const int64_t npix = (int64_t)in.rows * in.cols;
    if (npix == 0) return NSX_OK;
    void *area, *misc;
    NSX_TRY(ctx->ws(WS_CCL_AREA, sizeof(int32_t) * (size_t)npix, &area));
    NSX_TRY(ctx->ws(WS_CCL_MISC, 64, &misc));
    int32_t* m = (int32_t*)misc;
    const int g = 4 * ctx->cu_count;
    hipLaunchKernelGGL(zero_words_kernel, dim3(g), dim3(256), 0, s, (uint32_t*)area, npix);
    hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(64), 0, s, m, 4);
    const dim3 grid((in.cols + 255) / 256, in.rows);
    hipLaunchKernelGGL(area_count_kernel, grid, dim3(256), 0, s, in.data, in.pitch, in.rows, in.cols, (int32_t*)area,
                       npix, m + 2);
    hipLaunchKernelGGL(area_keep_count_kernel, dim3(g), dim3(256), 0, s, (const int32_t*)area, npix, min_size,
                       max_size, m + 1);
    hipLaunchKernelGGL(area_apply_kernel, grid, dim3(256), 0, s, in.data, in.pitch, in.rows, in.cols,
                       (const int32_t*)area, npix, min_size, max_size, out.data, out.pitch, flatten ? 1 : 0);
    NSX_HIP_TRY(ctx, hipGetLastError());
    NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned, m + 1, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
    if (ctx->h_pinned[1] != 0) return ctx->fail(NSX_ERR_INVALID, "labels must be -1 or a raster index of the image");
    if (count_dev) NSX_HIP_TRY(ctx, hipMemcpyAsync(count_dev, m + 1, sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    return NSX_OK;
}

nsx_status ccl_relabel(nsx_ctx* ctx, const View& lab, int bgval, int32_t* ncomp_dev, hipStream_t s) {
    (void)bgval;  // background maps to 0 like every non-root label value
    const int64_t npix = (int64_t)lab.rows * lab.cols;
    if (npix == 0) return NSX_OK;
    const int64_t nwords = (npix + 31) / 32;
    const int64_t nb = (nwords + 1023) / 1024;
    void *bits, *scan, *misc;
    NSX_TRY(ctx->ws(WS_CCL_BITS, sizeof(uint32_t) * (size_t)nwords, &bits));
    NSX_TRY(ctx->ws(WS_CCL_SCAN, sizeof(uint32_t) * (size_t)(nwords + nb + 1), &scan));
    NSX_TRY(ctx->ws(WS_CCL_MISC, 64, &misc));
    const int g = 4 * ctx->cu_count;
    hipLaunchKernelGGL(zero_words_kernel, dim3(g), dim3(256), 0, s, (uint32_t*)bits, nwords);
    hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(64), 0, s, (int32_t*)misc, 4);
    const dim3 grid((lab.cols + 255) / 256, lab.rows);
    hipLaunchKernelGGL(root_bits_kernel, grid, dim3(256), 0, s, lab.data, lab.pitch, lab.rows, lab.cols,
                       (uint32_t*)bits, (int32_t*)misc + 1);
    uint32_t* wordpre = (uint32_t*)scan;
    uint32_t* blockpre = wordpre + nwords;
    hipLaunchKernelGGL(scan_words_kernel, dim3(nb), dim3(256), 0, s, (const uint32_t*)bits, nwords, wordpre, blockpre);
    hipLaunchKernelGGL(scan_blocks_kernel, dim3(1), dim3(1024), 0, s, blockpre, (int)nb);
    hipLaunchKernelGGL(relabel_apply_kernel, grid, dim3(256), 0, s, lab.data, lab.pitch, lab.rows, lab.cols, npix,
                       (const uint32_t*)bits, (const uint32_t*)wordpre, (const uint32_t*)blockpre);
    NSX_HIP_TRY(ctx, hipGetLastError());
    if (ncomp_dev) NSX_HIP_TRY(ctx, hipMemcpyAsync(ncomp_dev, (int32_t*)misc + 1, sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    return NSX_OK;
}

nsx_status threshold_run(nsx_ctx* ctx, const View& in, double k, const View& out, hipStream_t s) {
    if (in.rows == 0 || in.cols == 0) return NSX_OK;
    const dim3 grid((in.cols + 255) / 256, in.rows);
    hipLaunchKernelGGL(threshold_kernel, grid, dim3(256), 0, s, in.data, in.pitch, in.dtype, in.rows, in.cols, k,
                       out.data, out.pitch);
    NSX_HIP_TRY(ctx, hipGetLastError());
    return NSX_OK;
}

nsx_status fill_const_i32(nsx_ctx* ctx, const View& out, int32_t v, hipStream_t s) {
    if (out.rows == 0 || out.cols == 0) return NSX_OK;
    const dim3 grid((out.cols + 255) / 256, out.rows);
    hipLaunchKernelGGL(fill_i32_kernel, grid, dim3(256), 0, s, out.data, out.pitch, out.rows, out.cols, v);
    NSX_HIP_TRY(ctx, hipGetLastError());
    return NSX_OK;
}

}  // namespace nsx
