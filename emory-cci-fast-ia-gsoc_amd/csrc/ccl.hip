// ccl.hip -- union-find connected-component labelling, area filter and
// canonical relabel on gfx950.
//
// Semantics follow ConnComponents.cpp:52-118,205-239 (bwlabel2 / bwareaopen2):
// 4/8-neighbours of EQUAL nonzero value are joined; each component's label is
// the raster index of its first pixel (the union-find root is always the
// smaller index, ConnComponents.cpp:34-43, ccl_uf.cu:40-62); the canonical
// relabel numbers those roots 1.. in raster order with background 0.
//
// One launch labels a batch of images (a single image is a batch of one).
// Each image owns a range of a global pixel index space that starts at a
// multiple of 32768, so the root bitmap's scan blocks never straddle images.
// Kernels (tile = 64 columns x 32 rows, 256 threads, 8 pixels per thread):
//   K1 local    tile-local union-find in LDS: each wave holds a band of rows;
//               each row's runs of equal values get their first pixel as
//               parent (wave ballot), then overlapping runs of adjacent rows are
//               joined once with ds_min; finds halve paths.  Per local
//               component (root global index, pixel count) goes to the tile's
//               own slots of a component table; every pixel's tile-local root
//               goes to a u16 array; `lab` (int32, sparse) holds the roots.
//   K2 merge    pixel pairs that cross a tile edge unite their local roots
//               with global atomicMin union-find over `lab`.
//   K3 area     per table entry: atomicAdd its count into area[find(root)].
//   K4 keep     per table entry that is a global root: keep it if the area is
//               in [min, max); set its bit in the root bitmap.
//   K5 scan     popcount prefix over the bitmap -> rank of every kept root and
//               the kept count of every image.
//   K6 final    read every pixel's tile-local root (2 B), map each local root
//               to its global root once, write the output.
// HBM bytes per pixel: K1 reads 1, K6 reads 1 and writes 4 (u8 input, int32
// output); everything else touches only tile edges, roots and the bitmap.
#include "nsx_internal.h"

#include <algorithm>
#include <initializer_list>
#include <cmath>

namespace nsx {
namespace {

// One tile per block, blocks dispatched to the 8 XCDs in turn (block b on XCD
// b mod 8).  Renumbered XCD-major, the blocks of one XCD take consecutive tiles,
// so each XCD works on one region of the image and neighbouring tiles, whose
// rows share cache lines, meet in its L2.  A bijection of [0, G).
__device__ __forceinline__ int xcd_tile(int b, int G) {
    const int x = b & 7;
    return x * (G >> 3) + min(x, G & 7) + (b >> 3);
}

constexpr int CW = 64;               // tile columns
constexpr int CH = 32;               // tile rows
constexpr int CN = CW * CH;          // pixels per tile (and table slots per tile)
constexpr int CTHREADS = 256;
static_assert(CN <= 65536 && CTHREADS == 256, "u16 local indices; four waves per tile");
constexpr int CPT = CN / CTHREADS;   // pixels per thread
constexpr uint32_t BGL = 0xFFFFFFFFu;
constexpr int BLOCK_BITS = 32768;    // bitmap scan block (1024 words)

struct CclImg {
    const char* in; int64_t pin;
    char* out; int64_t pout;
    const char* wgt; int64_t pw;   // u8 weight image (per-component maximum mode)
    char* lout; int64_t plout;     // int32 labels written with a CO_LEVEL output
    int rows, cols, ntx, nty;
    int tile_base;     // first tile of the image in the launch
    int pix_base;      // global index of pixel (0,0); multiple of BLOCK_BITS
};

struct CclArgs {
    CclImg img0;               // used when nimg == 1
    const CclImg* imgs;        // device table when nimg > 1
    const int32_t* tile_img;   // tile -> image when nimg > 1
    int nimg, ntiles;
    int64_t thr_i; float thr_f;
    int32_t* lab;              // sparse: tile-edge pixels + local roots
    uint16_t* lroot;           // every pixel: tile-local root index, 0xFFFF background
    int32_t* area;             // sparse: roots
    uint2* entries;            // tile t's components at [t*CN, t*CN + tcount[t])
    int32_t* tcount;           // components per tile
    uint32_t* bits;            // kept-root bitmap over the global index space
    uint32_t* wordpre;         // exclusive popcount prefix inside each 1024-word block
    uint32_t* blockpre;        // exclusive prefix of block totals (+ total at [nblocks])
    int64_t nwords; int nblocks;
    int32_t* ncomp;            // kept components per image
    int min_size, max_size;
    bool use_area;
    bool wmax;                 // `area` holds each component's maximum weight, not its size
    int level;                 // CO_LEVEL: out = max(out, min(level, maximum weight))
    int out_mode;
    bool level_labels;         // CO_LEVEL also writes canonical labels (kept: area filter on the maxima)
    int floor_level;           // CO_LEVEL with floor_mmax: writes every pixel (see CclParams)
    const uint32_t* floor_mmax;
};

// The tile -> image entry and the image descriptor are read through the
// constant address space with a wave-uniform index, so they are scalar loads
// into SGPRs and every row address computed from them stays on the scalar unit.
template <typename T> using cp = const __attribute__((address_space(4))) T*;

__device__ __forceinline__ void tile_of(const CclArgs& A, int t, CclImg& D, int& tx, int& ty) {
    t = __builtin_amdgcn_readfirstlane(t);
    if (A.nimg == 1) {
        D = A.img0;
    } else {
        const int idx = __builtin_amdgcn_readfirstlane(((cp<int32_t>)A.tile_img)[t]);
        static_assert(sizeof(CclImg) % 4 == 0, "descriptor copied as words");
        const cp<uint32_t> src = (cp<uint32_t>)(A.imgs + idx);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&D);
#pragma unroll
        for (int k = 0; k < (int)(sizeof(CclImg) / 4); ++k) dst[k] = src[k];
    }
    const int lt = t - D.tile_base;
    ty = lt / D.ntx;
    tx = lt - ty * D.ntx;
}

// Loads and stores through global (not flat) pointers.
template <typename T> using gp = __attribute__((address_space(1))) T*;
template <typename T> __device__ __forceinline__ T gload(const T* p) { return *(gp<const T>)p; }
template <typename T> __device__ __forceinline__ void gstore(T* p, T v) { *(gp<T>)p = v; }

// Pixel value used for connectivity: raw u8, or (in > thr) for the fused
// threshold (OpenCV `Mat > k`, 255/0 -> here 1/0).  The threshold is
// pre-converted on the host to the input type's comparison (thr_i for
// integers: v > k  <=>  v > floor(k); thr_f for float).
template <typename InT, bool THR>
__device__ __forceinline__ uint32_t pix(const CclArgs& A, const CclImg& D, int y, int x) {
    const InT v = gload(reinterpret_cast<const InT*>(D.in + (int64_t)y * D.pin) + x);
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

// Root of x without modifying the forest.
__device__ __forceinline__ uint32_t lds_root(const uint32_t* par, uint32_t x) {
    uint32_t p = par[x];
    while (p != x) {
        x = p;
        p = par[x];
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

// Find with path halving for the merge kernel.  Parents only ever decrease
// (every link points to a smaller index), so a halving write made from a stale
// read still points at an ancestor; a union that lands on a non-root carries
// on from the parent its atomicMin returned.
__device__ __forceinline__ int32_t g_find_halve(int32_t* lab, int32_t x) {
    int32_t p = lab[x];
    while (p != x) {
        const int32_t g = lab[p];
        if (g != p) lab[x] = g;
        x = p;
        p = g;
    }
    return x;
}

__device__ __forceinline__ void g_union(int32_t* lab, int32_t a, int32_t b) {
    while (true) {
        a = g_find_halve(lab, a);
        b = g_find_halve(lab, b);
        if (a == b) return;
        if (a < b) { int32_t t = a; a = b; b = t; }
        const int32_t old = atomicMin(&lab[a], b);
        if (old == a) return;
        a = old;
    }
}

// Tile-local labelling shared by K1 and K6.  On return val[] holds the pixel
// values and par[i] the local root of pixel i (BGL for background).
// Wave w holds the band of tile rows w*CPT .. w*CPT+CPT-1, lane = column, one
// register per row.
__device__ __forceinline__ int band_row(int w, int j) { return w * CPT + j; }

// lane i <- lane i-1 (lane 0 <- 0), lane i <- lane i+1 (lane 63 <- 0): the
// neighbour value across a tile edge counts as background.
__device__ __forceinline__ uint32_t from_left(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xF, 0xF, true);  // wave_shr:1
}
__device__ __forceinline__ uint32_t from_right(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xF, 0xF, true);  // wave_shl:1
}

// Tile-local labelling: v[j] receives the thread's pixel values (band rows),
// par[i] the local root of pixel i (BGL for background); edge[] (4 x CW) is
// scratch for the last row of each band.
template <int CONN, typename InT, bool THR>
__device__ __forceinline__ void local_label(const CclArgs& A, const CclImg& D, int x0, int y0, uint32_t (&v)[CPT],
                                            uint32_t* edge, uint32_t* par) {
    const int tid = threadIdx.x;
    // w is wave-uniform: readfirstlane lets the row arithmetic run on the scalar unit
    const int lx = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int gx = x0 + lx;
    // all loads first (independent, in flight together), then the LDS writes
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int gy = y0 + band_row(w, j);
        v[j] = (gx < D.cols && gy < D.rows) ? pix<InT, THR>(A, D, gy, gx) : 0u;
    }
    const uint64_t upto = lx == 63 ? ~0ull : ((2ull << lx) - 1ull);   // lanes 0..lx
    if constexpr (THR) {
        // 0/1 values: each row is one 64-bit foreground mask (bit = lane), and
        // the run starts and the links between rows below are bit operations
        // on those wave-uniform masks instead of per-lane neighbour moves
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int i = band_row(w, j) * CW + lx;
            const uint64_t M = __ballot(v[j] != 0u);
            const uint64_t S = M & ~(M << 1);   // first pixel of each run
            const int rs = 63 - __clzll(S & upto);
            par[i] = v[j] ? (uint32_t)(band_row(w, j) * CW + rs) : BGL;
        }
    } else {
        // A pixel's initial parent is the first pixel of its horizontal run of
        // equal values, found with a ballot of run starts, so the unions below
        // only join runs of adjacent rows.
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int i = band_row(w, j) * CW + lx;
            const bool start = v[j] != 0u && from_left(v[j]) != v[j];
            const uint64_t starts = __ballot(start);
            const int rs = 63 - __clzll(starts & upto);
            par[i] = v[j] ? (uint32_t)(band_row(w, j) * CW + rs) : BGL;
        }
    }
    edge[w * CW + lx] = v[CPT - 1];
    __syncthreads();
    // Join each pair of overlapping runs once, at the leftmost column where they
    // touch: a vertical (or diagonal) link that a left neighbour in the same run
    // already makes through the same run above is skipped.  The row above is the
    // previous register, or for a band's first row the band above's last row.
    if constexpr (THR) {
        // the same links as below, for 0/1 values: with M the row's mask and U
        // the row above's (bit b = lane b; << 1 moves a lane's left neighbour
        // onto it, >> 1 its right one)
        //   up:         M & U & ~(left & up-left)
        //   up-left:    M & (U << 1) & ~U & ~(M << 1)          (8-conn)
        //   up-right:   M & (U >> 1) & ~U & ~(M >> 1)          (8-conn)
        uint64_t U = w > 0 ? __ballot(edge[(w - 1) * CW + lx] != 0u) : 0ull;
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int ly = band_row(w, j);
            const uint64_t M = __ballot(v[j] != 0u);
            const uint64_t up = M & U & ~((M << 1) & (U << 1));
            const uint64_t ul = CONN == 8 ? (M & (U << 1) & ~U & ~(M << 1)) : 0ull;
            const uint64_t ur = CONN == 8 ? (M & (U >> 1) & ~U & ~(M >> 1)) : 0ull;
            U = M;
            if (ly == 0 || !(up | ul | ur)) continue;   // wave-uniform
            const int i = ly * CW + lx;
            const int u = i - CW;
            if ((up >> lx) & 1ull) lds_union(par, i, u);
            if (CONN == 8) {
                if ((ul >> lx) & 1ull) lds_union(par, i, u - 1);
                if ((ur >> lx) & 1ull) lds_union(par, i, u + 1);
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int ly = band_row(w, j);
            const int i = ly * CW + lx;
            const uint32_t vv = v[j];
            const uint32_t vu = j > 0 ? v[j - 1] : (w > 0 ? edge[(w - 1) * CW + lx] : 0u);
            const uint32_t vl = from_left(vv), vr = from_right(vv);
            const uint32_t vul = from_left(vu), vur = from_right(vu);
            if (!vv || ly == 0) continue;
            const int u = i - CW;
            if (vu == vv && !(vl == vv && vul == vv)) lds_union(par, i, u);
            if (CONN == 8) {
                if (vul == vv && vu != vv && vl != vv) lds_union(par, i, u - 1);
                if (vur == vv && vu != vv && vr != vv) lds_union(par, i, u + 1);
            }
        }
    }
    __syncthreads();
    // Flatten: every pixel's root is found with reads only, and written after a
    // barrier.  (A halving write made while other threads store their roots can
    // land after one of them and leave a pixel pointing at a non-root.)
    uint32_t root[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int i = band_row(w, j) * CW + lx;
        root[j] = par[i] != BGL ? lds_root(par, i) : BGL;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CPT; ++j) par[band_row(w, j) * CW + lx] = root[j];
    __syncthreads();
}

// WMAX: each local component's slot carries the maximum of the u8 weight image
// over its pixels instead of its pixel count.
template <int CONN, typename InT, bool THR, bool WMAX>
__global__ __launch_bounds__(CTHREADS) void ccl_local_kernel(CclArgs A) {
    __shared__ uint32_t edge[(CTHREADS / 64) * CW];
    __shared__ uint32_t par[CN];
    __shared__ uint32_t cnt[CN];
    __shared__ uint32_t nroots;
    const int tid = threadIdx.x;
    CclImg D;
    int tx, ty;
    const int tb = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    tile_of(A, tb, D, tx, ty);
    const int x0 = tx * CW, y0 = ty * CH;
    for (int i = tid; i < CN; i += CTHREADS) cnt[i] = 0;
    if (tid == 0) nroots = 0;
    const int lx = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int gx = x0 + lx;
    uint32_t wv[WMAX ? CPT : 1];
    if (WMAX) {
        // issued before the labelling so they arrive while it runs
        const int cx = min(gx, D.cols - 1);
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int cy = min(y0 + band_row(w, j), D.rows - 1);
            wv[j] = gload((const uint8_t*)(D.wgt + (int64_t)cy * D.pw) + cx);
        }
    }
    uint32_t v[CPT];
    local_label<CONN, InT, THR>(A, D, x0, y0, v, edge, par);
    // Sizes: a run's pixels share a root, so the lane at a run's start adds the
    // run's length (first end at or after it, from a ballot of run ends).
    // Maximum weights: every pixel.
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int i = band_row(w, j) * CW + lx;
        if (WMAX) {
            // a weight that does not exceed the current maximum (0 included)
            // changes nothing: most pixels of a component skip the atomic
            if (par[i] != BGL && wv[j] > cnt[par[i]]) atomicMax(&cnt[par[i]], wv[j]);
        } else if constexpr (THR) {
            const uint64_t M = __ballot(v[j] != 0u);
            const uint64_t ends = M & ~(M >> 1);   // last pixel of each run
            if (((M & ~(M << 1)) >> lx) & 1ull) atomicAdd(&cnt[par[i]], (uint32_t)__builtin_ctzll(ends >> lx) + 1u);
        } else {
            const bool start = v[j] != 0u && from_left(v[j]) != v[j];
            const bool end = v[j] != 0u && from_right(v[j]) != v[j];
            const uint64_t ends = __ballot(end);
            if (start) atomicAdd(&cnt[par[i]], (uint32_t)__builtin_ctzll(ends >> lx) + 1u);
        }
    }
    __syncthreads();
    uint2* slots = A.entries + (int64_t)tb * CN;
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int ly = band_row(w, j);
        const int i = ly * CW + lx;
        const uint32_t r = par[i];
        const int32_t g = D.pix_base + (y0 + ly) * D.cols + gx;
        // the output kernel reads each pixel's local root back (2 B) instead of
        // labelling the tile again
        if (gx < D.cols && y0 + ly < D.rows) gstore(A.lroot + g, r == BGL ? (uint16_t)0xFFFF : (uint16_t)r);
        if (r == BGL) continue;
        if (r == (uint32_t)i) {
            const uint32_t k = atomicAdd(&nroots, 1u);
            slots[k] = make_uint2((uint32_t)g, cnt[i]);
            A.lab[g] = g;
            if (A.use_area || WMAX) A.area[g] = 0;
        }
    }
    __syncthreads();
    if (tid == 0) A.tcount[tb] = (int32_t)nroots;
}

// Global index of the tile-local root lr of the tile holding pixel (qy, qx).
__device__ __forceinline__ int32_t root_gidx(const CclImg& D, int qy, int qx, uint32_t lr) {
    const int ty0 = qy - qy % CH, tx0 = qx - qx % CW;
    return D.pix_base + (ty0 + (int)(lr >> 6)) * D.cols + tx0 + (int)(lr & 63u);
}

// K2: one thread per tile-edge pixel whose backward neighbours cross the edge:
// threads 0..63 (wave 0) the top row (up, up-left, up-right), 64..95 (wave 1,
// lanes 0..31) the left column (left, and up-left below the top row), 96..126
// (lanes 32..62) the right column (up-right below the top row).  Pixels and
// their neighbours are compared through the local roots K1 stored (a root
// means foreground; raw u8 input also compares values) and the two local
// roots are united.  Redundant unions are skipped:
//   * the previous pixel along the same edge joined the same pair of roots;
//   * a diagonal whose straight neighbour across the edge already joins: the
//     two neighbours are adjacent inside one tile, hence one local component
//     (except at the tile corners, where the diagonal lies in a third tile);
//   * the right column's up-right, when the pixel above is in the same local
//     component: the right tile's left column joins that pair.
template <int CONN, typename InT, bool THR>
__global__ __launch_bounds__(256) void ccl_merge_kernel(CclArgs A) {
    // two tiles per workgroup, 128 threads (two whole waves) each
    const int t = (int)blockIdx.x * 2 + (int)(threadIdx.x >> 7);
    if (t >= A.ntiles) return;
    const int tid = threadIdx.x & 127;
    const int lane = tid & 63;
    CclImg D;
    int tx, ty;
    tile_of(A, t, D, tx, ty);
    const int x0 = tx * CW, y0 = ty * CH;
    const int kind = tid < 64 ? 0 : tid < 96 ? 1 : 2;   // top, left, right
    int y, x;
    if (kind == 0) { y = y0; x = x0 + tid; }
    else if (kind == 1) { y = y0 + (tid - 64); x = x0; }
    else { y = y0 + 1 + (tid - 96); x = x0 + CW - 1; }
    const bool valid = tid != 127 && x < D.cols && y < D.rows && y < y0 + CH;
    auto root_at = [&](int qy, int qx) -> uint32_t {
        if (qy < 0 || qx < 0 || qx >= D.cols || qy >= D.rows) return 0xFFFFu;
        return A.lroot[D.pix_base + qy * D.cols + qx];
    };
    auto same = [&](int qy, int qx, uint32_t qr, uint32_t v) -> bool {
        // fg pixel q with the value of this pixel
        if (qr == 0xFFFFu) return false;
        return THR || pix<InT, THR>(A, D, qy, qx) == v;
    };
    const uint32_t a = valid ? root_at(y, x) : 0xFFFFu;
    const uint32_t v = (!THR && a != 0xFFFFu) ? pix<InT, THR>(A, D, y, x) : 1u;
    // the straight neighbour across the edge (up for the top row, left for the
    // left column) and its root, for the previous-lane test
    int sy = y - 1, sx = x;
    if (kind == 1) { sy = y; sx = x - 1; }
    const uint32_t b = (a != 0xFFFFu && kind != 2) ? root_at(sy, sx) : 0xFFFFu;
    const bool straight = kind != 2 && a != 0xFFFFu && same(sy, sx, b, v);
    // previous pixel along the edge (lane - 1, same edge) with the same pair
    const uint32_t pa = from_left(a), pb = from_left(b);
    const bool prev_same_edge = kind == 0 ? lane > 0 : (kind == 1 ? lane > 0 : false);
    const bool dup = prev_same_edge && pa == a && pb == b;
    if (a == 0xFFFFu) return;
    const int32_t ga = root_gidx(D, y, x, a);
    if (straight && !dup) g_union(A.lab, ga, root_gidx(D, sy, sx, b));
    if (CONN != 8) return;
    auto diag = [&](int qy, int qx) {
        const uint32_t q = root_at(qy, qx);
        if (same(qy, qx, q, v)) g_union(A.lab, ga, root_gidx(D, qy, qx, q));
    };
    if (kind == 0) {
        if (y > 0) {
            if (!straight || x == x0) diag(y - 1, x - 1);
            if (!straight || x == x0 + CW - 1) diag(y - 1, x + 1);
        }
    } else if (kind == 1) {
        if (x > 0 && y > y0 && !straight) diag(y - 1, x - 1);
    } else {
        if (x + 1 < D.cols) {
            const uint32_t u = root_at(y - 1, x);   // inside this tile (y > y0)
            if (u != a) diag(y - 1, x + 1);
        }
    }
}

// K3 / K4: over each tile's component slots.
// SLOT_LANES lanes per tile, SLOT_TILES tiles per workgroup: tiles hold few
// components, so a workgroup per tile would leave most lanes idle and cost a
// dispatch per tile.
constexpr int SLOT_LANES = 16;
constexpr int SLOT_TILES = 256 / SLOT_LANES;
inline int slot_grid(int ntiles) { return (ntiles + SLOT_TILES - 1) / SLOT_TILES; }

__global__ __launch_bounds__(256) void ccl_area_kernel(CclArgs A) {
    const int t = (int)blockIdx.x * SLOT_TILES + (int)(threadIdx.x / SLOT_LANES);
    if (t >= A.ntiles) return;
    const int n = A.tcount[t];
    const uint2* slots = A.entries + (int64_t)t * CN;
    for (int e = (int)(threadIdx.x % SLOT_LANES); e < n; e += SLOT_LANES) {
        const uint2 en = slots[e];
        const int v = (int)en.y;
        const int32_t r = g_find(A.lab, (int32_t)en.x);
        // the output kernel finds the root of this same local root again
        if (r != (int32_t)en.x) A.lab[en.x] = r;
        if (A.wmax) {
            // a large component has one slot per tile it crosses: only a slot
            // that can raise the maximum pays for the atomic on its root
            if (v > A.area[r]) atomicMax(&A.area[r], v);
        } else {
            atomicAdd(&A.area[r], v);
        }
    }
}

__global__ __launch_bounds__(256) void ccl_keep_kernel(CclArgs A) {
    const int t = (int)blockIdx.x * SLOT_TILES + (int)(threadIdx.x / SLOT_LANES);
    if (t >= A.ntiles) return;
    const int n = A.tcount[t];
    const uint2* slots = A.entries + (int64_t)t * CN;
    for (int e = (int)(threadIdx.x % SLOT_LANES); e < n; e += SLOT_LANES) {
        const int32_t g = (int32_t)slots[e].x;
        const int32_t p = A.lab[g];
        if (p != g) {
            // the output kernel finds this local root's root again: shorten it
            // to one hop (the root is final: all unions are done)
            const int32_t r = g_find(A.lab, p);
            if (r != p) A.lab[g] = r;
            continue;
        }
        bool keep = true;
        if (A.use_area) {
            const int a = A.area[g];
            keep = a >= A.min_size && a < A.max_size;
        }
        if (keep) atomicOr(&A.bits[g >> 5], 1u << (g & 31));
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

// K5b: one block of 1024 threads scans the block totals in place (exclusive)
// and stores the grand total at blocksum[nb].
__device__ __forceinline__ void scan_blocks(uint32_t* blocksum, int nb) {
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
    if (tid == 1023) blocksum[nb] = part[1023];
}

__global__ __launch_bounds__(1024) void scan_blocks_kernel(uint32_t* blocksum, int nb) { scan_blocks(blocksum, nb); }

// K5b + K5c: the block scan, then the kept components per image from the
// block prefix (one block: its own writes are visible after the barrier).
__global__ __launch_bounds__(1024) void scan_blocks_count_kernel(CclArgs A) {
    scan_blocks(A.blockpre, A.nblocks);
    __syncthreads();
    for (int b = threadIdx.x; b < A.nimg; b += blockDim.x) {
        const CclImg D = A.nimg == 1 ? A.img0 : A.imgs[b];
        const int n = D.rows * D.cols;
        const int blk0 = D.pix_base / BLOCK_BITS;
        const int blk1 = (int)(((int64_t)D.pix_base + n + BLOCK_BITS - 1) / BLOCK_BITS);
        A.ncomp[b] = (int32_t)(A.blockpre[blk1] - A.blockpre[blk0]);
    }
}

__device__ __forceinline__ bool kept(const CclArgs& A, int32_t r) {
    return (A.bits[r >> 5] >> (r & 31)) & 1u;
}
__device__ __forceinline__ int32_t rank1(const CclArgs& A, const CclImg& D, int32_t r) {
    const uint32_t w = A.bits[r >> 5];
    return (int32_t)(A.blockpre[r >> 15] - A.blockpre[D.pix_base >> 15] + A.wordpre[r >> 5] +
                     __popc(w & ((1u << (r & 31)) - 1u))) + 1;
}

// LL: CO_LEVEL that also writes the canonical labels of the kept components
// to D.lout (0 for the others and the background).
template <int CONN, typename InT, bool THR, bool LL = false>
__global__ __launch_bounds__(CTHREADS) void ccl_final_kernel(CclArgs A) {
    __shared__ int32_t outv[CN];
    __shared__ int32_t outl[LL ? CN : 1];
    __shared__ uint16_t rlist[CN];   // local indices of the local roots, CN/4 per wave
    const int tid = threadIdx.x;
    CclImg D;
    int tx, ty;
    const int tb = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    tile_of(A, tb, D, tx, ty);
    const int x0 = tx * CW, y0 = ty * CH;
    // w stays a per-lane value here: made wave-uniform (as in the local kernel)
    // the rows' loads and stores split into scalar branches and this kernel
    // ran 12-25 % slower (same-session A/B on the 4096^2 chain, the batch and
    // the many-valued maze)
    const int lx = tid & 63, w = tid >> 6;
    const int gx = x0 + lx;
    // each pixel's tile-local root, as the local kernel left it, and for a
    // level output the value so far (loaded now, used after the barrier)
    const bool level = A.out_mode == CO_LEVEL;
    const bool fold = level && !A.floor_mmax;   // level value folded into `out`
    uint32_t par[CPT], prev[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int gy = y0 + w + 4 * j;
        const bool in = gx < D.cols && gy < D.rows;
        const uint16_t r = in ? gload(A.lroot + D.pix_base + gy * D.cols + gx) : (uint16_t)0xFFFF;
        par[j] = r == 0xFFFF ? BGL : (uint32_t)r;
        prev[j] = (fold && in) ? gload((const uint8_t*)(D.out + (int64_t)gy * D.pout) + gx) : 0u;
    }
    // Each local root resolves its output once, one lane per local root: each
    // wave compacts the roots among its pixels into its own LDS list (ballot
    // per row) and resolves them with its lanes, so roots in several rows of a
    // wave do not chain their global lookups row after row.  The area kernel
    // (level mode) or the keep kernel (otherwise) left every local root's `lab`
    // entry pointing at its global root (a root points at itself).
    static_assert(CPT * 64 * (CTHREADS / 64) == CN, "one list region per wave");
    uint16_t* wl = rlist + w * (CPT * 64);
    uint32_t nw = 0;   // this wave's roots so far (wave-uniform)
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int i = (w + 4 * j) * CW + lx;
        const uint64_t m = __ballot(par[j] == (uint32_t)i);
        if (par[j] == (uint32_t)i) wl[nw + __builtin_popcountll(m & ((1ull << lx) - 1ull))] = (uint16_t)i;
        nw += (uint32_t)__builtin_popcountll(m);
    }
    // the list is read back by other lanes of the same wave: LDS accesses of a
    // wave complete in order, so a wave-scope fence suffices
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int k = lx; k < (int)nw; k += 64) {
        const int i = wl[k];
        const int32_t r = A.lab[D.pix_base + (y0 + i / CW) * D.cols + x0 + (i % CW)];
        int32_t o;
        switch (A.out_mode) {
            case CO_ROOT: o = r - D.pix_base; break;
            case CO_ROOT_FILTERED: o = kept(A, r) ? r - D.pix_base : -1; break;
            case CO_CANON: o = kept(A, r) ? rank1(A, D, r) : 0; break;
            case CO_LEVEL: o = min(A.level, A.area[r]); break;
            default: o = kept(A, r) ? 255 : 0; break;
        }
        outv[i] = o;
        if (LL) outl[i] = kept(A, r) ? rank1(A, D, r) : 0;
    }
    __syncthreads();
    const int32_t bg = (A.out_mode == CO_ROOT || A.out_mode == CO_ROOT_FILTERED) ? -1 : 0;
    const int32_t floorv = (level && !fold) ? min(A.floor_level, (int32_t)*A.floor_mmax) : 0;
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int ly = w + 4 * j;
        const int gy = y0 + ly;
        if (gx >= D.cols || gy >= D.rows) continue;
        const uint32_t r = par[j];
        char* row = D.out + (int64_t)gy * D.pout;
        if (A.out_mode == CO_LEVEL) {
            // u8 reconstruction accumulated over levels: foreground only
            if (!fold) gstore((uint8_t*)row + gx, (uint8_t)(r == BGL ? floorv : max(floorv, outv[r])));
            else if (r != BGL) gstore((uint8_t*)row + gx, (uint8_t)max((int32_t)prev[j], outv[r]));
            if (LL) gstore((int32_t*)(D.lout + (int64_t)gy * D.plout) + gx, r == BGL ? 0 : outl[r]);
            continue;
        }
        const int32_t o = (r == BGL) ? bg : outv[r];
        if (A.out_mode == CO_FLAT_U8) gstore((uint8_t*)row + gx, (uint8_t)o);
        else gstore((int32_t*)row + gx, o);
    }
}

__global__ void zero_words_kernel(uint32_t* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

template <int CONN, typename InT, bool THR>
void ccl_launch_conn(nsx_ctx* ctx, const CclArgs& A, hipStream_t s) {
    if (A.out_mode == CO_LEVEL) {
        // per-component maximum weight folded into a u8 image: no bitmap or ranks
        int h = ctx->tic(TK_CCL_LOCAL, s);
        hipLaunchKernelGGL((ccl_local_kernel<CONN, InT, THR, true>), dim3(A.ntiles), dim3(CTHREADS), 0, s, A);
        ctx->toc(h, s);
        h = ctx->tic(TK_CCL_MERGE, s);
        hipLaunchKernelGGL((ccl_merge_kernel<CONN, InT, THR>), dim3((A.ntiles + 1) / 2), dim3(256), 0, s, A);
        ctx->toc(h, s);
        h = ctx->tic(TK_CCL_SMALL, s);
        hipLaunchKernelGGL(ccl_area_kernel, dim3(slot_grid(A.ntiles)), dim3(256), 0, s, A);
        if (A.level_labels) {
            // keep and rank the components whose maximum passes the filter
            hipLaunchKernelGGL(zero_words_kernel, dim3(4 * ctx->cu_count), dim3(256), 0, s, A.bits, A.nwords);
            hipLaunchKernelGGL(ccl_keep_kernel, dim3(slot_grid(A.ntiles)), dim3(256), 0, s, A);
            hipLaunchKernelGGL(scan_words_kernel, dim3(A.nblocks), dim3(256), 0, s, A.bits, A.nwords, A.wordpre,
                               A.blockpre);
            hipLaunchKernelGGL(scan_blocks_count_kernel, dim3(1), dim3(1024), 0, s, A);
        }
        ctx->toc(h, s);
        h = ctx->tic(TK_CCL_FINAL, s);
        if (A.level_labels)
            hipLaunchKernelGGL((ccl_final_kernel<CONN, InT, THR, true>), dim3(A.ntiles), dim3(CTHREADS), 0, s, A);
        else
            hipLaunchKernelGGL((ccl_final_kernel<CONN, InT, THR>), dim3(A.ntiles), dim3(CTHREADS), 0, s, A);
        ctx->toc(h, s);
        return;
    }
    const int stream_grid = 4 * ctx->cu_count;
    hipLaunchKernelGGL(zero_words_kernel, dim3(stream_grid), dim3(256), 0, s, A.bits, A.nwords);
    int h = ctx->tic(TK_CCL_LOCAL, s);
    hipLaunchKernelGGL((ccl_local_kernel<CONN, InT, THR, false>), dim3(A.ntiles), dim3(CTHREADS), 0, s, A);
    ctx->toc(h, s);
    h = ctx->tic(TK_CCL_MERGE, s);
    hipLaunchKernelGGL((ccl_merge_kernel<CONN, InT, THR>), dim3((A.ntiles + 1) / 2), dim3(256), 0, s, A);
    ctx->toc(h, s);
    h = ctx->tic(TK_CCL_SMALL, s);
    if (A.use_area) hipLaunchKernelGGL(ccl_area_kernel, dim3(slot_grid(A.ntiles)), dim3(256), 0, s, A);
    hipLaunchKernelGGL(ccl_keep_kernel, dim3(slot_grid(A.ntiles)), dim3(256), 0, s, A);
    hipLaunchKernelGGL(scan_words_kernel, dim3(A.nblocks), dim3(256), 0, s, A.bits, A.nwords, A.wordpre, A.blockpre);
    hipLaunchKernelGGL(scan_blocks_count_kernel, dim3(1), dim3(1024), 0, s, A);
    ctx->toc(h, s);
    h = ctx->tic(TK_CCL_FINAL, s);
    hipLaunchKernelGGL((ccl_final_kernel<CONN, InT, THR>), dim3(A.ntiles), dim3(CTHREADS), 0, s, A);
    ctx->toc(h, s);
}

template <typename InT, bool THR>
void ccl_launch_typed(nsx_ctx* ctx, const CclArgs& A, int conn, hipStream_t s) {
    if (conn == 8) ccl_launch_conn<8, InT, THR>(ctx, A, s);
    else ccl_launch_conn<4, InT, THR>(ctx, A, s);
}

}  // namespace

size_t ccl_desc_bytes() { return sizeof(CclImg); }

int ccl_tiles(int rows, int cols) {
    if (rows <= 0 || cols <= 0) return 0;
    return ((cols + CW - 1) / CW) * ((rows + CH - 1) / CH);
}

int64_t ccl_pix_span(int rows, int cols) {
    const int64_t n = (int64_t)rows * cols;
    return (n + BLOCK_BITS - 1) / BLOCK_BITS * BLOCK_BITS;
}

void ccl_fill_desc(void* out, const View& in, const View& o, int tile_base, int64_t pix_base) {
    CclImg D{};
    D.in = in.data; D.pin = in.pitch;
    D.out = o.data; D.pout = o.pitch;
    D.rows = in.rows; D.cols = in.cols;
    D.ntx = (in.cols + CW - 1) / CW;
    D.nty = (in.rows + CH - 1) / CH;
    D.tile_base = tile_base;
    D.pix_base = (int)pix_base;
    *reinterpret_cast<CclImg*>(out) = D;
}

nsx_status ccl_launch(nsx_ctx* ctx, const CclBatch& B, const CclParams& p, hipStream_t s) {
    if (B.ntiles <= 0) return NSX_OK;
    if (B.pix_total >= (int64_t)INT32_MAX) return ctx->fail(NSX_ERR_INVALID, "batch too large for int32 labels");
    if (!p.threshold && B.in_dtype != NSX_U8) return ctx->fail(NSX_ERR_INVALID, "labelling needs u8 input or a threshold");
    CclArgs A{};
    A.nimg = B.nimg;
    A.ntiles = B.ntiles;
    if (B.nimg == 1) A.img0 = *reinterpret_cast<const CclImg*>(B.host_desc);
    A.imgs = reinterpret_cast<const CclImg*>(B.dev_desc);
    A.tile_img = B.tile_img;
    A.thr_i = (int64_t)std::floor(std::max(-9.0e18, std::min(9.0e18, p.thr)));
    A.thr_f = (float)p.thr;
    A.min_size = p.min_size; A.max_size = p.max_size; A.use_area = p.use_area;
    A.wmax = p.out_mode == CO_LEVEL;
    A.level = p.level;
    A.level_labels = A.wmax && A.use_area;
    A.floor_level = p.floor_level;
    A.floor_mmax = A.wmax ? p.floor_mmax : nullptr;
    if (A.level_labels && !A.img0.lout)
        return ctx->fail(NSX_ERR_INVALID, "level labels need a labels image");
    if (A.wmax && (B.in_dtype != NSX_U8 || B.nimg != 1))
        return ctx->fail(NSX_ERR_INVALID, "level labelling needs one u8 image");
    A.out_mode = p.out_mode;
    A.nwords = B.pix_total / 32;
    A.nblocks = (int)(B.pix_total / BLOCK_BITS);
    void *lab, *area = nullptr, *ent, *tcnt, *bits, *scan;
    NSX_TRY(ctx->ws(B.ws_base + 0, sizeof(int32_t) * (size_t)B.pix_total, &lab));
    if (A.use_area || A.wmax) NSX_TRY(ctx->ws(B.ws_base + 1, sizeof(int32_t) * (size_t)B.pix_total, &area));
    NSX_TRY(ctx->ws(B.ws_base + 2, sizeof(uint2) * (size_t)B.ntiles * CN, &ent));
    NSX_TRY(ctx->ws(B.ws_base + 3, sizeof(int32_t) * (size_t)B.ntiles, &tcnt));
    NSX_TRY(ctx->ws(B.ws_base + 4, sizeof(uint32_t) * (size_t)A.nwords, &bits));
    NSX_TRY(ctx->ws(B.ws_base + 5, sizeof(uint32_t) * (size_t)(A.nwords + A.nblocks + 1), &scan));
    void* lroot;
    NSX_TRY(ctx->ws(B.lroot_slot, sizeof(uint16_t) * (size_t)B.pix_total, &lroot));
    A.lroot = (uint16_t*)lroot;
    A.lab = (int32_t*)lab; A.area = (int32_t*)area; A.entries = (uint2*)ent; A.tcount = (int32_t*)tcnt;
    A.bits = (uint32_t*)bits;
    A.wordpre = (uint32_t*)scan;
    A.blockpre = (uint32_t*)scan + A.nwords;
    A.ncomp = B.ncomp_dev;
    if (!p.threshold) ccl_launch_typed<uint8_t, false>(ctx, A, p.conn, s);
    else if (B.in_dtype == NSX_U8) ccl_launch_typed<uint8_t, true>(ctx, A, p.conn, s);
    else if (B.in_dtype == NSX_I32) ccl_launch_typed<int32_t, true>(ctx, A, p.conn, s);
    else ccl_launch_typed<float, true>(ctx, A, p.conn, s);
    NSX_HIP_TRY(ctx, hipGetLastError());
    return NSX_OK;
}

nsx_status ccl_run(nsx_ctx* ctx, const View& in, const View& out, const CclParams& p, int32_t* ncomp_dev,
                   hipStream_t s) {
    if ((int64_t)in.rows * in.cols == 0) return NSX_OK;
    CclImg D;
    ccl_fill_desc(&D, in, out, 0, 0);
    CclBatch B{};
    B.nimg = 1;
    B.ntiles = ccl_tiles(in.rows, in.cols);
    B.pix_total = ccl_pix_span(in.rows, in.cols);
    B.host_desc = &D;
    B.in_dtype = in.dtype;
    B.ws_base = WS_CCL_LABEL;
    B.ncomp_dev = ncomp_dev ? ncomp_dev : (int32_t*)ctx->d_stats + 30;
    return ccl_launch(ctx, B, p, s);
}

nsx_status ccl_level_run(nsx_ctx* ctx, const View& mask, const View& weight, const View& out, int conn, int level,
                         hipStream_t s, const View* labels, double thresh, int32_t* ncomp_dev, int floor_level,
                         const uint32_t* floor_mmax) {
    if ((int64_t)mask.rows * mask.cols == 0) return NSX_OK;
    CclImg D;
    ccl_fill_desc(&D, mask, out, 0, 0);
    D.wgt = weight.data;
    D.pw = weight.pitch;
    if (labels) {
        D.lout = labels->data;
        D.plout = labels->pitch;
    }
    CclBatch B{};
    B.nimg = 1;
    B.ntiles = ccl_tiles(mask.rows, mask.cols);
    B.pix_total = ccl_pix_span(mask.rows, mask.cols);
    B.host_desc = &D;
    B.in_dtype = mask.dtype;
    B.ws_base = WS_CCL_LABEL;
    B.ncomp_dev = (int32_t*)ctx->d_stats + 30;
    CclParams p;
    p.conn = conn;
    p.out_mode = CO_LEVEL;
    p.threshold = true;
    p.thr = level - 1;   // foreground: mask > level - 1, i.e. mask >= level
    p.level = level;
    p.floor_level = floor_level;
    p.floor_mmax = floor_mmax;
    if (labels) {
        // kept: maximum > thresh, i.e. >= floor(thresh) + 1
        p.use_area = true;
        p.min_size = (int)std::floor(std::max(-1.0, std::min(256.0, thresh))) + 1;
        p.max_size = INT32_MAX;
        B.ncomp_dev = ncomp_dev;
    }
    return ccl_launch(ctx, B, p, s);
}

nsx_status ccl_class_fold_run(nsx_ctx* ctx, const View& classes, const View& weight, const View& out, int conn,
                              hipStream_t s) {
    if ((int64_t)classes.rows * classes.cols == 0) return NSX_OK;
    CclImg D;
    ccl_fill_desc(&D, classes, out, 0, 0);
    D.wgt = weight.data;
    D.pw = weight.pitch;
    CclBatch B{};
    B.nimg = 1;
    B.ntiles = ccl_tiles(classes.rows, classes.cols);
    B.pix_total = ccl_pix_span(classes.rows, classes.cols);
    B.host_desc = &D;
    B.in_dtype = NSX_U8;
    B.ws_base = WS_CCL_LABEL;
    B.ncomp_dev = (int32_t*)ctx->d_stats + 30;
    CclParams p;
    p.conn = conn;
    p.out_mode = CO_LEVEL;
    p.threshold = false;   // components: equal-valued nonzero neighbours
    p.level = 255;
    return ccl_launch(ctx, B, p, s);
}

namespace {
// Which u8 values occur in `mask` (bit v of present[v >> 5]) and the largest
// marker value.  The LDS bitmap is read before it is set, so a value costs one
// atomic per block, not one per pixel.
__global__ __launch_bounds__(256) void levels_u8_kernel(const char* mask, int64_t pk, const char* marker,
                                                        int64_t pm, int rows, int cols, uint32_t* present,
                                                        uint32_t* mmax) {
    __shared__ uint32_t lp[8];
    __shared__ uint32_t lm;
    if (threadIdx.x < 8) lp[threadIdx.x] = 0u;
    if (threadIdx.x == 0) lm = 0u;
    __syncthreads();
    // per-thread value set: quarter q of the byte range in q[...] (64 bits each)
    uint64_t q0 = 0, q1 = 0, q2 = 0, q3 = 0;
    uint32_t m = 0;
    auto add = [&](uint32_t v) {
        const uint64_t b = 1ull << (v & 63u);
        const uint32_t q = v >> 6;
        q0 |= q == 0 ? b : 0ull;
        q1 |= q == 1 ? b : 0ull;
        q2 |= q == 2 ? b : 0ull;
        q3 |= q == 3 ? b : 0ull;
    };
    for (int y = blockIdx.x; y < rows; y += gridDim.x) {
        const uint8_t* mr = (const uint8_t*)(mask + (int64_t)y * pk);
        const uint8_t* ar = (const uint8_t*)(marker + (int64_t)y * pm);
        for (int x0 = 4 * (int)threadIdx.x; x0 < cols; x0 += 4 * (int)blockDim.x) {
            uint32_t v[4], a[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int x = min(x0 + k, cols - 1);   // repeats the last pixel: same value set
                v[k] = mr[x];
                a[k] = ar[x];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                add(v[k]);
                m = max(m, a[k]);
            }
        }
    }
    uint32_t w[8] = {(uint32_t)q0, (uint32_t)(q0 >> 32), (uint32_t)q1, (uint32_t)(q1 >> 32),
                     (uint32_t)q2, (uint32_t)(q2 >> 32), (uint32_t)q3, (uint32_t)(q3 >> 32)};
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
#pragma unroll
        for (int k = 0; k < 8; ++k) w[k] |= (uint32_t)__shfl_xor((int)w[k], o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&lm, m);
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (w[k]) atomicOr(&lp[k], w[k]);
    }
    __syncthreads();
    if (threadIdx.x < 8 && lp[threadIdx.x]) atomicOr(&present[threadIdx.x], lp[threadIdx.x]);
    if (threadIdx.x == 0) atomicMax(mmax, lm);
}

// The same scan for a small value limit: each wave keeps the list of values it
// has seen (wave-uniform, at most LV_KCAP) and checks every byte against it,
// a few compares per byte instead of a 256-bit bitmap update.  A wave that
// meets more than `kmax` distinct values raises *overflow and every wave stops
// at its next row: the mask then has too many values for the level-set path.
// Rows are read as 16-byte aligned blocks; bytes of an edge block outside the
// row are replaced by a value already listed (mask) or by 0 (marker).  An
// aligned 16-byte block holding a byte of the row lies in that byte's page.
constexpr int LV_KCAP = 16;
typedef uint32_t lv_u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short lv_u16x2 __attribute__((ext_vector_type(2)));

// 0xFF in each byte k of word j (block bytes 4j..4j+3) with lo <= 4j+k < hi
__device__ __forceinline__ uint32_t lv_keep(int lo, int hi, int j) {
    uint32_t m = 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k) m |= (4 * j + k >= lo && 4 * j + k < hi) ? (0xFFu << (8 * k)) : 0u;
    return m;
}

// Loads the block of row [start, end) with index i, zero for i >= nblk; the
// bytes outside the row become `fill4`'s byte.
__device__ __forceinline__ void lv_block(uintptr_t start, uintptr_t end, int i, int nblk, uint32_t fill4,
                                         uint32_t (&w)[4]) {
    const uintptr_t a = (start & ~(uintptr_t)15) + 16 * (uintptr_t)i;
    if (i < nblk) {
        const lv_u32x4 v = *(const lv_u32x4*)a;
        w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
        if (a < start || a + 16 > end) {
            const int lo = a < start ? (int)(start - a) : 0;
            const int hi = a + 16 > end ? (int)(end - a) : 16;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t km = lv_keep(lo, hi, j);
                w[j] = (w[j] & km) | (fill4 & ~km);
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) w[j] = fill4;
    }
}

__global__ __launch_bounds__(256) void levels_few_u8_kernel(const char* mask, int64_t pk, const char* marker,
                                                            int64_t pm, int rows, int cols, int kmax,
                                                            uint32_t* present, uint32_t* mmax, uint32_t* overflow) {
    __shared__ uint32_t lp[8];
    __shared__ uint32_t lm;
    if (threadIdx.x < 8) lp[threadIdx.x] = 0u;
    if (threadIdx.x == 0) lm = 0u;
    __syncthreads();
    // the list starts with the image's first pixel, so a listed value exists
    // to stand in for bytes outside the row
    uint32_t sv[LV_KCAP];
    const uint32_t first = *(const uint8_t*)mask;
#pragma unroll
    for (int j = 0; j < LV_KCAP; ++j) sv[j] = first;
    int cnt = 1;
    bool over = false, seen = false;   // seen: another block (or pass) set the overflow word
    lv_u16x2 ml = {0, 0}, mh = {0, 0};
    for (int y = blockIdx.x; y < rows; y += gridDim.x) {
        if (__builtin_amdgcn_readfirstlane(*(volatile uint32_t*)overflow)) {
            over = seen = true;
            break;
        }
        const uintptr_t ms = (uintptr_t)(mask + (int64_t)y * pk), me = ms + (uintptr_t)cols;
        const int mblk = (int)((me - (ms & ~(uintptr_t)15) + 15) >> 4);
        for (int b0 = 0; b0 < mblk; b0 += (int)blockDim.x) {   // uniform trip count: every lane takes part
            uint32_t w[4];
            lv_block(ms, me, b0 + (int)threadIdx.x, mblk, sv[0] * 0x01010101u, w);
            // high bit of each byte of um: the byte differs from every listed
            // value.  (x & 0x7F..) + 0x7F.. sets a byte's high bit iff its low
            // seven bits are not all zero, with no carry into the next byte.
            uint32_t um[4] = {~0u, ~0u, ~0u, ~0u};
#pragma unroll
            for (int j = 0; j < LV_KCAP; ++j) {
                if (j < cnt) {   // uniform
                    const uint32_t s4 = sv[j] * 0x01010101u;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const uint32_t x = w[q] ^ s4;
                        um[q] &= ((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x;
                    }
                }
            }
            const bool any = ((um[0] | um[1] | um[2] | um[3]) & 0x80808080u) != 0u;
            if (__ballot(any) == 0ull) continue;
            // rare: add the new values one at a time
            for (int b = 0; b < 16 && !over; ++b) {
                const uint32_t x = (w[b >> 2] >> (8 * (b & 3))) & 0xFFu;
                while (true) {
                    bool in = false;
#pragma unroll
                    for (int j = 0; j < LV_KCAP; ++j)
                        if (j < cnt) in = in || x == sv[j];
                    const uint64_t m = __ballot(!in);
                    if (m == 0ull) break;
                    if (cnt >= kmax) {
                        over = true;
                        break;
                    }
                    const uint32_t nv = (uint32_t)__builtin_amdgcn_readlane((int)x, (int)__builtin_ctzll(m));
#pragma unroll
                    for (int j = 0; j < LV_KCAP; ++j)
                        if (j == cnt) sv[j] = nv;
                    ++cnt;
                }
            }
            if (over) break;
        }
        if (over) break;
        const uintptr_t as = (uintptr_t)(marker + (int64_t)y * pm), ae = as + (uintptr_t)cols;
        const int ablk = (int)((ae - (as & ~(uintptr_t)15) + 15) >> 4);
        for (int i = (int)threadIdx.x; i < ablk; i += (int)blockDim.x) {
            uint32_t w[4];
            lv_block(as, ae, i, ablk, 0u, w);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                ml = __builtin_elementwise_max(ml, __builtin_bit_cast(lv_u16x2, w[j] & 0x00FF00FFu));
                mh = __builtin_elementwise_max(mh, __builtin_bit_cast(lv_u16x2, (w[j] >> 8) & 0x00FF00FFu));
            }
        }
    }
    uint32_t m = max(max((uint32_t)ml.x, (uint32_t)ml.y), max((uint32_t)mh.x, (uint32_t)mh.y));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
    if ((threadIdx.x & 63) == 0) {
        if (over) {
            // one atomic per block that found the overflow itself
            if (!seen && threadIdx.x == 0) atomicOr(overflow, 1u);
        } else {
            atomicMax(&lm, m);
            for (int j = 0; j < cnt; ++j) atomicOr(&lp[sv[j] >> 5], 1u << (sv[j] & 31u));
        }
    }
    __syncthreads();
    if (threadIdx.x < 8 && lp[threadIdx.x]) atomicOr(&present[threadIdx.x], lp[threadIdx.x]);
    if (threadIdx.x == 0) atomicMax(mmax, lm);
}

// labels = v everywhere, *count = v (0 or 1 component)
__global__ void fill_labels_kernel(char* out, int64_t po, int rows, int cols, int v, int32_t* count) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *count = v;
    for (int y = blockIdx.x; y < rows; y += gridDim.x) {
        int32_t* r = (int32_t*)(out + (int64_t)y * po);
        for (int x = threadIdx.x; x < cols; x += blockDim.x) r[x] = v;
    }
}

__global__ void zero_u32_kernel(uint32_t* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

// out = min(level, *mmax) everywhere: the reconstruction at the image's
// smallest mask value, whose level set is the whole (connected) image.
__global__ void level_floor_kernel(char* out, int64_t po, int rows, int cols, int level, const uint32_t* mmax) {
    const uint8_t v = (uint8_t)min((uint32_t)level, *mmax);
    for (int y = blockIdx.x; y < rows; y += gridDim.x) {
        uint8_t* r = (uint8_t*)(out + (int64_t)y * po);
        for (int x = threadIdx.x; x < cols; x += blockDim.x) r[x] = v;
    }
}
}  // namespace

constexpr int kLvSample = 32;
nsx_status recon_levelset_u8(nsx_ctx* ctx, const View& marker, const View& mask, const View& out, int conn,
                             int max_levels, bool* used, hipStream_t s, LevelLabelReq* req) {
    *used = false;
    if (mask.dtype != NSX_U8 || max_levels <= 0 || (int64_t)mask.rows * mask.cols == 0) return NSX_OK;
    // d_stats int32 words 40..47: value bitmap, 48: largest marker value,
    // 49: more distinct values than the level-set path takes
    uint32_t* d = (uint32_t*)ctx->d_stats + 40;
    const int grid = std::max(1, std::min(mask.rows, 4 * ctx->cu_count));
    const uint32_t* h = (const uint32_t*)ctx->h_pinned + 40;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, d, 10);
    if (max_levels + 1 <= LV_KCAP && mask.rows >= 8 * kLvSample) {
        // every kLvSample-th row first: on an image with many values (most)
        // it sets the overflow word after a short pass, and the full pass then
        // stops at its first row.  Otherwise the full pass adds to the same
        // words (the sampled values are among the image's).
        const int rs = (mask.rows - 1) / kLvSample + 1;
        hipLaunchKernelGGL(levels_few_u8_kernel, dim3(std::min(rs, 4 * ctx->cu_count)), dim3(256), 0, s, mask.data,
                           mask.pitch * kLvSample, marker.data, marker.pitch * kLvSample, rs, mask.cols,
                           max_levels + 1, d, d + 8, d + 9);
    }
    if (max_levels + 1 <= LV_KCAP)
        hipLaunchKernelGGL(levels_few_u8_kernel, dim3(grid), dim3(256), 0, s, mask.data, mask.pitch, marker.data,
                           marker.pitch, mask.rows, mask.cols, max_levels + 1, d, d + 8, d + 9);
    else
        hipLaunchKernelGGL(levels_u8_kernel, dim3(grid), dim3(256), 0, s, mask.data, mask.pitch, marker.data,
                           marker.pitch, mask.rows, mask.cols, d, d + 8);
    NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned + 40, d, 10 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
    if (h[9] != 0u) return NSX_OK;
    std::vector<int> levels;   // descending
    for (int v = 255; v >= 0; --v)
        if ((h[v >> 5] >> (v & 31)) & 1u) levels.push_back(v);
    if (levels.empty() || (int)levels.size() - 1 > max_levels) return NSX_OK;
    // the smallest level above the labelling threshold (-1: none) runs last
    // and also writes the labels, from its own components
    const int nl = (int)levels.size();
    int lstar = -1;
    if (req)
        for (int j = nl - 1; j >= 0 && lstar < 0; --j)
            if ((double)levels[j] > req->thresh) lstar = j;
    std::vector<int> order;
    for (int j = 0; j + 1 < nl; ++j)
        if (j != lstar) order.push_back(j);
    if (lstar >= 0 && lstar + 1 < nl) order.push_back(lstar);
    // the smallest level's value min(level, largest marker value) holds
    // everywhere: the first level run writes it with its own values (every
    // pixel, no read), the others fold into the result; no run: a fill
    if (order.empty())
        hipLaunchKernelGGL(level_floor_kernel, dim3(grid), dim3(256), 0, s, out.data, out.pitch, out.rows, out.cols,
                           levels.back(), d + 8);
    for (size_t o = 0; o < order.size(); ++o) {
        const int j = order[o];
        const bool lab = req && j == lstar;
        NSX_TRY(ccl_level_run(ctx, mask, marker, out, conn, levels[j], s, lab ? &req->labels : nullptr,
                              req ? req->thresh : 0.0, lab ? req->ncomp_dev : nullptr, levels.back(),
                              o == 0 ? d + 8 : nullptr));
    }
    if (req) {
        const View& L = req->labels;
        if (lstar < 0 || lstar + 1 >= nl) {
            // none above the threshold: no foreground; the smallest level: the
            // whole (connected) image, foreground iff the largest marker value
            // is above the threshold (h[8])
            const bool all = lstar == nl - 1 && (double)h[8] > req->thresh;
            hipLaunchKernelGGL(fill_labels_kernel, dim3(grid), dim3(256), 0, s, L.data, L.pitch, L.rows, L.cols,
                               all ? 1 : 0, req->ncomp_dev);
        }
        req->done = true;
    }
    NSX_HIP_TRY(ctx, hipGetLastError());
    *used = true;
    return NSX_OK;
}

namespace {
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

__global__ void zero_i32_kernel(int32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
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

// ---- u8 elementwise kernels: 16 pixels per thread along a row ------------------
// `vec`: every row of every image involved starts on a 16-byte boundary (one
// 16-byte access per row group); otherwise, and for a row's last partial
// group, byte accesses.
constexpr int U8V = 16;
struct U8x16 {
    uint32_t w[4];
};
__device__ __forceinline__ U8x16 u8x16_load(const char* row, int x, int cols, bool vec) {
    U8x16 b;
    if (vec && x + U8V <= cols) {
        const uint4 v = *reinterpret_cast<const uint4*>(row + x);
        b.w[0] = v.x;
        b.w[1] = v.y;
        b.w[2] = v.z;
        b.w[3] = v.w;
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            uint32_t w = 0u;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int xx = x + 4 * i + j;
                if (xx < cols) w |= (uint32_t)((const uint8_t*)row)[xx] << (8 * j);
            }
            b.w[i] = w;
        }
    }
    return b;
}
__device__ __forceinline__ void u8x16_store(char* row, int x, int cols, bool vec, const U8x16& b) {
    if (vec && x + U8V <= cols) {
        *reinterpret_cast<uint4*>(row + x) = make_uint4(b.w[0], b.w[1], b.w[2], b.w[3]);
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int xx = x + 4 * i + j;
                if (xx < cols) ((uint8_t*)row)[xx] = (uint8_t)(b.w[i] >> (8 * j));
            }
    }
}
inline bool u8_vec(std::initializer_list<uintptr_t> v) {
    uintptr_t o = 0;
    for (uintptr_t x : v) o |= x;
    return (o & 15u) == 0u;
}
inline dim3 u8x16_grid(int rows, int cols) {
    return dim3((unsigned)((cols + 256 * U8V - 1) / (256 * U8V)), (unsigned)rows);
}

// Seeded markers of imfill / bwselect (u8): marker = image where seeds != 0,
// else 0; with `invert`, the image is first complemented (255 - x) and the
// complement is also written to `mask` (imfill's reconstruction mask).
__global__ __launch_bounds__(256) void seeded_marker16_kernel(const char* img, int64_t pi, const char* seeds, int64_t ps,
                                                              char* marker, int64_t pm, char* mask, int64_t pk,
                                                              int cols, int invert, int vec) {
    const int x = (blockIdx.x * 256 + (int)threadIdx.x) * U8V, y = blockIdx.y;
    if (x >= cols) return;
    const bool vc = vec != 0;
    U8x16 v = u8x16_load(img + (int64_t)y * pi, x, cols, vc);
    const U8x16 sd = u8x16_load(seeds + (int64_t)y * ps, x, cols, vc);
    U8x16 m;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (invert) v.w[i] = ~v.w[i];
        const uint32_t nz = (sd.w[i] | ((sd.w[i] & 0x7F7F7F7Fu) + 0x7F7F7F7Fu)) & 0x80808080u;
        m.w[i] = v.w[i] & ((nz >> 7) * 0xFFu);
    }
    if (invert) u8x16_store(mask + (int64_t)y * pk, x, cols, vc, v);
    u8x16_store(marker + (int64_t)y * pm, x, cols, vc, m);
}

// ---- local maxima / minima (localMaxima, localMinima) ---------------------------
__device__ __forceinline__ float lx_get(const char* p, int64_t pitch, int y, int x, int dtype) {
    const char* row = p + (int64_t)y * pitch;
    return dtype == NSX_U8 ? (float)((const uint8_t*)row)[x] : ((const float*)row)[x];
}
__device__ __forceinline__ void lx_put(char* p, int64_t pitch, int y, int x, int dtype, float v) {
    char* row = p + (int64_t)y * pitch;
    if (dtype == NSX_U8) ((uint8_t*)row)[x] = (uint8_t)v;
    else ((float*)row)[x] = v;
}

// work = image (minima: its complement, u8 255 - x, f32 0 - x); marker = work - 1
// (u8 saturating).  u8 values are exact in float.
__global__ void lx_prep_kernel(const char* img, int64_t pi, char* work, int64_t pw, char* marker, int64_t pm,
                               int rows, int cols, int dtype, int minima) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= cols) return;
    float v = lx_get(img, pi, y, x, dtype);
    if (minima) v = dtype == NSX_U8 ? 255.0f - v : 0.0f - v;
    lx_put(work, pw, y, x, dtype, v);
    lx_put(marker, pm, y, x, dtype, dtype == NSX_U8 ? fmaxf(v - 1.0f, 0.0f) : v - 1.0f);
}

// flags: 0 not a candidate (the reconstruction rose above the marker), 1 a
// candidate, 2 a candidate with a strictly higher neighbour (out-of-image
// neighbours are never higher).
__global__ void lx_mark_kernel(const char* work, int64_t pw, const char* marker, int64_t pm, const char* rec,
                               int64_t pr, uint8_t* flags, int rows, int cols, int dtype, int conn) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= cols) return;
    uint8_t f = 0;
    if (!(lx_get(marker, pm, y, x, dtype) < lx_get(rec, pr, y, x, dtype))) {
        f = 1;
        const float v = lx_get(work, pw, y, x, dtype);
        for (int k = 0; k < (conn == 8 ? 8 : 4); ++k) {
            const int dy = k < 4 ? (k == 0 ? -1 : k == 1 ? 1 : 0) : (k < 6 ? -1 : 1);
            const int dx = k < 4 ? (k == 2 ? -1 : k == 3 ? 1 : 0) : ((k & 1) ? 1 : -1);
            const int qy = y + dy, qx = x + dx;
            if (qy >= 0 && qy < rows && qx >= 0 && qx < cols && v < lx_get(work, pw, qy, qx, dtype)) f = 2;
        }
    }
    flags[(int64_t)y * cols + x] = f;
}

// A candidate next to a marked candidate of the same value is marked (the
// reference's flood of equal-valued candidates).  One 64x64 tile per block:
// values and marks with a one-pixel halo in LDS, the rule applied until the
// tile is stable, the marks written back.  A tile runs only when it or one of
// its 8 neighbour tiles changed in the previous round (prev); marks only grow,
// so reading a halo that a neighbour is writing is harmless.
constexpr int LX_T = 64;
constexpr int LX_H = LX_T + 2;

__global__ __launch_bounds__(256) void lx_tile_kernel(const char* work, int64_t pw, uint8_t* flags, int rows,
                                                      int cols, int dtype, int conn, int tiles_x, int tiles_y,
                                                      const uint8_t* prev, uint8_t* next, int32_t* changed) {
    __shared__ float sv[LX_H * LX_H];
    __shared__ uint8_t sf[LX_H * LX_H];
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
        if (threadIdx.x == 0) next[tile] = 0;
        return;
    }
    const int x0 = tx * LX_T - 1, y0 = ty * LX_T - 1;
    for (int i = threadIdx.x; i < LX_H * LX_H; i += 256) {
        const int gx = x0 + i % LX_H, gy = y0 + i / LX_H;
        const bool in = gx >= 0 && gx < cols && gy >= 0 && gy < rows;
        sf[i] = in ? flags[(int64_t)gy * cols + gx] : 0;
        sv[i] = in ? lx_get(work, pw, gy, gx, dtype) : 0.0f;
    }
    __syncthreads();
    const int off[8] = {-LX_H, LX_H, -1, 1, -LX_H - 1, -LX_H + 1, LX_H - 1, LX_H + 1};
    const int nn = conn == 8 ? 8 : 4;
    bool any = false;
    for (;;) {
        int ch = 0;
        for (int k = 0; k < LX_T * LX_T / 256; ++k) {
            const int p = threadIdx.x + 256 * k;
            const int c = (p / LX_T + 1) * LX_H + (p & (LX_T - 1)) + 1;
            if (sf[c] != 1) continue;
            const float v = sv[c];
            for (int j = 0; j < nn; ++j)
                if (sf[c + off[j]] == 2 && sv[c + off[j]] == v) {
                    sf[c] = 2;
                    ch = 1;
                    break;
                }
        }
        if (!__syncthreads_or(ch)) break;
        any = true;
    }
    if (any) {
        for (int i = threadIdx.x; i < LX_T * LX_T; i += 256) {
            const int gx = tx * LX_T + (i & (LX_T - 1)), gy = ty * LX_T + i / LX_T;
            if (gx < cols && gy < rows) flags[(int64_t)gy * cols + gx] = sf[(i / LX_T + 1) * LX_H + (i & (LX_T - 1)) + 1];
        }
    }
    if (threadIdx.x == 0) {
        next[tile] = any ? 1 : 0;
        if (any) *changed = 1;
    }
}

__global__ void lx_tiles_init_kernel(uint8_t* f, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) f[i] = 1;
}

__global__ void lx_out_kernel(const uint8_t* flags, char* out, int64_t po, int rows, int cols) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= cols) return;
    ((uint8_t*)(out + (int64_t)y * po))[x] = flags[(int64_t)y * cols + x] == 1 ? 255 : 0;
}

__global__ void lx_zero_kernel(int32_t* w) {
    if (threadIdx.x == 0) *w = 0;
}

// u8 images.  A candidate of value v >= 1 (the reconstruction of work - 1
// stays at v - 1) has no strictly higher neighbour q: the reconstruction would
// reach min(v, v(q) - 1) = v there.  Marks therefore only ever flood value-0
// candidates, and every 0 pixel is a candidate (its marker saturates at 0).
// A conn-connected set of 0 pixels that is not the whole image has a nonzero
// neighbour, so it is dropped whole.  The result: 255 at the candidates of
// value >= 1, and at the 0 pixels exactly when every pixel is 0.  A pixel of
// value v >= 1 is a candidate exactly when no strictly higher pixel is reached
// through pixels >= v, i.e. when its plateau (the conn-connected pixels of
// value v around it) has no strictly higher neighbour: one labelling of equal
// nonzero values, each plateau folding the largest of its pixels' "has a
// higher neighbour" weights, replaces the reconstruction.
constexpr int LX_V = U8V;

// per byte: 0x80 where the byte is nonzero
__device__ __forceinline__ uint32_t lx_nz_bytes(uint32_t w) {
    return (w | ((w & 0x7F7F7F7Fu) + 0x7F7F7F7Fu)) & 0x80808080u;
}

// Work value of byte i of a 16-pixel group, i = -1 and 16 the bytes beside it
// (0 outside the row: pixels outside the image are never higher).
struct LxRow {
    uint32_t w[4];
    uint32_t l, r;
    __device__ __forceinline__ uint32_t at(int i) const {
        return i < 0 ? l : i >= LX_V ? r : (w[i >> 2] >> (8 * (i & 3))) & 0xFFu;
    }
};
__device__ __forceinline__ LxRow lx_row(const char* img, int64_t pi, int y, int x, int rows, int cols, int minima,
                                        bool vec) {
    LxRow R;
    if (y < 0 || y >= rows) {
        R.w[0] = R.w[1] = R.w[2] = R.w[3] = R.l = R.r = 0u;
        return R;
    }
    const char* row = img + (int64_t)y * pi;
    const U8x16 b = u8x16_load(row, x, cols, vec);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int left = cols - x - 4 * i;
        const uint32_t valid = left >= 4 ? 0xFFFFFFFFu : left <= 0 ? 0u : (1u << (8 * left)) - 1u;
        R.w[i] = (minima ? ~b.w[i] : b.w[i]) & valid;
    }
    const uint32_t fl = minima ? 0xFFu : 0u;
    R.l = x > 0 ? (((const uint8_t*)row)[x - 1] ^ fl) : 0u;
    R.r = x + LX_V < cols ? (((const uint8_t*)row)[x + LX_V] ^ fl) : 0u;
    return R;
}

// Per group of 16 pixels of row y: weight = 1 where a work pixel of value >= 1
// has a strictly higher conn-neighbour; the classes to label (the work value;
// 0 for 0 and 255, since a 255 plateau never has a higher neighbour, and for a
// pixel with no equal-valued neighbour, a plateau of its own); the fold plane
// (0 where the labelling folds the plateau's weight in, the weight itself for
// the single-pixel plateaus); *nonzero |= 1 if some work pixel is not 0.
__global__ __launch_bounds__(256) void lx_prep_u8_kernel(const char* img, int64_t pi, char* cls, int64_t pc,
                                                         char* weight, int64_t pwt, char* fold, int64_t pf, int rows,
                                                         int cols, int minima, int conn, int vec, uint32_t* nonzero) {
    const int x = (blockIdx.x * 256 + (int)threadIdx.x) * LX_V, y = blockIdx.y;
    bool any = false;
    if (x < cols) {
        const bool vc = vec != 0;
        const LxRow up = lx_row(img, pi, y - 1, x, rows, cols, minima, vc);
        const LxRow mid = lx_row(img, pi, y, x, rows, cols, minima, vc);
        const LxRow dn = lx_row(img, pi, y + 1, x, rows, cols, minima, vc);
        U8x16 wt, cb, fb;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            uint32_t ww = 0u, cw = 0u, fw = 0u;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = 4 * i + j;
                const uint32_t v = mid.at(k);
                uint32_t h = max(max(up.at(k), dn.at(k)), max(mid.at(k - 1), mid.at(k + 1)));
                bool eq = up.at(k) == v || dn.at(k) == v || mid.at(k - 1) == v || mid.at(k + 1) == v;
                if (conn == 8) {
                    h = max(h, max(max(up.at(k - 1), up.at(k + 1)), max(dn.at(k - 1), dn.at(k + 1))));
                    eq = eq || up.at(k - 1) == v || up.at(k + 1) == v || dn.at(k - 1) == v || dn.at(k + 1) == v;
                }
                const uint32_t wgt = (v != 0u && h > v) ? 1u : 0u;
                const bool lab = eq && v != 0u && v != 0xFFu;
                ww |= wgt << (8 * j);
                cw |= (lab ? v : 0u) << (8 * j);
                fw |= (lab ? 0u : wgt) << (8 * j);
            }
            wt.w[i] = ww;
            cb.w[i] = cw;
            fb.w[i] = fw;
            any = any || mid.w[i] != 0u;
        }
        u8x16_store(weight + (int64_t)y * pwt, x, cols, vc, wt);
        u8x16_store(cls + (int64_t)y * pc, x, cols, vc, cb);
        u8x16_store(fold + (int64_t)y * pf, x, cols, vc, fb);
    }
    if (__syncthreads_or(any) && threadIdx.x == 0 && __hip_atomic_load(nonzero, __ATOMIC_RELAXED,
                                                                        __HIP_MEMORY_SCOPE_AGENT) == 0u)
        atomicOr(nonzero, 1u);
}

// out = 255 at work pixels of value 255, at those of value 1..254 whose plateau
// took no weight (fold == 0), and at 0 pixels when no work pixel is nonzero;
// 0 elsewhere.
__global__ __launch_bounds__(256) void lx_out_u8_kernel(const char* img, int64_t pi, const char* rec, int64_t pr,
                                                        char* out, int64_t po, int rows, int cols, int minima,
                                                        int vec, const uint32_t* nonzero) {
    const int x = (blockIdx.x * 256 + (int)threadIdx.x) * LX_V, y = blockIdx.y;
    if (x >= cols) return;
    const uint32_t zero_out = *nonzero ? 0u : 0xFFu;
    const U8x16 b = u8x16_load(img + (int64_t)y * pi, x, cols, vec != 0);
    const U8x16 r = u8x16_load(rec + (int64_t)y * pr, x, cols, vec != 0);
    U8x16 o;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t w = minima ? ~b.w[i] : b.w[i];
        uint32_t ow = 0u;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t v = (w >> (8 * j)) & 0xFFu, rv = (r.w[i] >> (8 * j)) & 0xFFu;
            const uint32_t ob = v == 0xFFu ? 0xFFu : v != 0u ? (rv == 0u ? 0xFFu : 0u) : zero_out;
            ow |= ob << (8 * j);
        }
        o.w[i] = ow;
    }
    u8x16_store(out + (int64_t)y * po, x, cols, vec != 0, o);
}

// ---- bounding boxes (ConnComponents::boundingBox) ------------------------------
// misc: [0] raster index of the first label -1 or 0 that is not the background
// (the reference stops there), [1] first raster index of a label outside
// [-1, rows*cols], [2] the largest label.  Blocks stride over rows; one
// atomic per block and word.
__global__ __launch_bounds__(256) void bbox_scan_kernel(const char* lab, int64_t pl, int rows, int cols, int bgval,
                                                        int vec, int32_t* misc) {
    __shared__ int32_t red[3][4];
    int32_t stop = INT_MAX, bad = INT_MAX, mx = 0;
    const int64_t lim = (int64_t)rows * cols;
    auto see = [&](int32_t v, int32_t i) {
        if (v == bgval) return;
        if (v == -1 || v == 0) stop = min(stop, i);
        else if (v < -1 || (int64_t)v > lim) bad = min(bad, i);
        else mx = max(mx, v);
    };
    for (int y = blockIdx.x; y < rows; y += gridDim.x) {
        const int32_t* row = (const int32_t*)(lab + (int64_t)y * pl);
        // four labels per thread (one 16-byte load when rows are 16-byte aligned)
        for (int x = 4 * (int)threadIdx.x; x < cols; x += 1024) {
            const int32_t i = y * cols + x;
            if (vec && x + 4 <= cols) {
                const int4 q = *reinterpret_cast<const int4*>(row + x);
                see(q.x, i);
                see(q.y, i + 1);
                see(q.z, i + 2);
                see(q.w, i + 3);
            } else {
                for (int k = 0; k < 4 && x + k < cols; ++k) see(row[x + k], i + k);
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        stop = min(stop, __shfl_xor(stop, o, 64));
        bad = min(bad, __shfl_xor(bad, o, 64));
        mx = max(mx, __shfl_xor(mx, o, 64));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = stop;
        red[1][w] = bad;
        red[2][w] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; ++k) {
            stop = min(stop, red[0][k]);
            bad = min(bad, red[1][k]);
            mx = max(mx, red[2][k]);
        }
        // the words only move one way: a block that cannot move one skips its
        // atomic (every block has a largest label, so most would contend)
        volatile int32_t* vm = misc;
        if (stop != INT_MAX && stop < vm[0]) atomicMin(&misc[0], stop);
        if (bad != INT_MAX && bad < vm[1]) atomicMin(&misc[1], bad);
        if (mx > 0 && mx > vm[2]) atomicMax(&misc[2], mx);
    }
}

__global__ void bbox_misc_init_kernel(int32_t* misc) {
    if (threadIdx.x == 0) { misc[0] = INT_MAX; misc[1] = INT_MAX; misc[2] = 0; }
}

// tab[5][cap]: first raster index, min x, max x, min y, max y per label value
__global__ void bbox_init_kernel(int32_t* tab, int64_t cap) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 5 * cap; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = i / cap;
        tab[i] = (c == 2 || c == 4) ? -1 : INT_MAX;
    }
}

// Runs of one label along a row, within each wave's 64 columns; pixels from
// the stop index on count as background.  A block takes a tile of 64 columns
// and BB_ROWS rows: the run's first lane folds the run into the tile's LDS
// table (open addressing on the label), and the table's entries update the
// global boxes once per label and tile.  A run whose label finds no free slot
// within BB_PROBE probes updates the global boxes itself.
constexpr int BB_ROWS = 64;
constexpr int BB_SLOTS = 256;
constexpr int BB_PROBE = 16;
__global__ __launch_bounds__(256) void bbox_runs_kernel(const char* lab, int64_t pl, int rows, int cols, int bgval,
                                                        const int32_t* misc, int32_t* tab, int64_t cap) {
    __shared__ int32_t key[BB_SLOTS];
    __shared__ int32_t sb[5][BB_SLOTS];   // first raster index, min x, max x, min y, max y
    constexpr int32_t kEmpty = -1;        // labels folded here are >= 1
    for (int k = threadIdx.x; k < BB_SLOTS; k += 256) {
        key[k] = kEmpty;
        sb[0][k] = INT_MAX;
        sb[1][k] = INT_MAX;
        sb[2][k] = -1;
        sb[3][k] = INT_MAX;
        sb[4][k] = -1;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int x = blockIdx.x * 64 + lane;
    const int32_t stop = misc[0];
    for (int r = w; r < BB_ROWS; r += 4) {
        const int y = blockIdx.y * BB_ROWS + r;
        if (y >= rows) break;
        const int32_t i = y * cols + x;
        int32_t v = bgval;
        if (x < cols && i < stop) v = ((const int32_t*)(lab + (int64_t)y * pl))[x];
        const int32_t left = __shfl_up(v, 1, 64), right = __shfl_down(v, 1, 64);
        const bool fg = v != bgval;
        const bool start = fg && (lane == 0 || left != v);
        const bool end = fg && (lane == 63 || right != v);
        const uint64_t ends = __ballot(end);
        if (start) {
            const int xe = x + __builtin_ctzll(ends >> lane);
            int slot = -1;
            uint32_t h = ((uint32_t)v * 2654435761u) >> 24;   // 8 bits: BB_SLOTS
            for (int p = 0; p < BB_PROBE; ++p, h = (h + 1) & (BB_SLOTS - 1)) {
                int32_t k = key[h];
                if (k == kEmpty) k = atomicCAS(&key[h], kEmpty, v);
                if (k == kEmpty || k == v) {
                    slot = (int)h;
                    break;
                }
            }
            if (slot >= 0) {
                atomicMin(&sb[0][slot], i);
                atomicMin(&sb[1][slot], x);
                atomicMax(&sb[2][slot], xe);
                atomicMin(&sb[3][slot], y);
                atomicMax(&sb[4][slot], y);
            } else {
                atomicMin(&tab[v], i);
                atomicMin(&tab[cap + v], x);
                atomicMax(&tab[2 * cap + v], xe);
                atomicMin(&tab[3 * cap + v], y);
                atomicMax(&tab[4 * cap + v], y);
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < BB_SLOTS; k += 256) {
        const int32_t v = key[k];
        if (v == kEmpty) continue;
        atomicMin(&tab[v], sb[0][k]);
        atomicMin(&tab[cap + v], sb[1][k]);
        atomicMax(&tab[2 * cap + v], sb[2][k]);
        atomicMin(&tab[3 * cap + v], sb[3][k]);
        atomicMax(&tab[4 * cap + v], sb[4][k]);
    }
}

// Marks each present label's first raster index in a bitmap over positions.
__global__ void bbox_mark_kernel(const int32_t* tab, int64_t cap, uint32_t* bits) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < cap; v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t f = tab[v];
        if (f != INT_MAX) atomicOr(&bits[f >> 5], 1u << (f & 31));
    }
}

// Writes the boxes in order of first appearance: the rank of a label's first
// position among all first positions is its slot; n = blockpre[nb].
__global__ void bbox_scatter_kernel(const int32_t* tab, int64_t cap, const uint32_t* bits, const uint32_t* wordpre,
                                    const uint32_t* blockpre, int nb, int32_t* out) {
    const int n = (int)blockpre[nb];
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < cap; v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t f = tab[v];
        if (f == INT_MAX) continue;
        const int r = (int)(blockpre[f >> 15] + wordpre[f >> 5] + __popc(bits[f >> 5] & ((1u << (f & 31)) - 1u)));
        out[r] = (int32_t)v;
        out[n + r] = tab[cap + v];
        out[2 * n + r] = tab[2 * cap + v];
        out[3 * n + r] = tab[3 * cap + v];
        out[4 * n + r] = tab[4 * cap + v];
    }
}

// imhmin's inputs: mask = invert(image) (u8 255 - x, f32 0 - x), marker =
// mask - h (saturating for u8).  dtype NSX_U8 or NSX_F32.
__global__ void hmin_prep_kernel(const char* img, int64_t pi, char* mask, int64_t pk, char* marker, int64_t pm,
                                 int rows, int cols, int dtype, int hu, float hf) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= cols) return;
    if (dtype == NSX_U8) {
        const int v = 255 - ((const uint8_t*)(img + (int64_t)y * pi))[x];
        ((uint8_t*)(mask + (int64_t)y * pk))[x] = (uint8_t)v;
        ((uint8_t*)(marker + (int64_t)y * pm))[x] = (uint8_t)(v > hu ? v - hu : 0);
    } else {
        const float v = 0.0f - ((const float*)(img + (int64_t)y * pi))[x];
        ((float*)(mask + (int64_t)y * pk))[x] = v;
        ((float*)(marker + (int64_t)y * pm))[x] = v - hf;
    }
}

// imhmin's u8 inputs, 16 pixels per thread.
__global__ __launch_bounds__(256) void hmin_prep16_kernel(const char* img, int64_t pi, char* mask, int64_t pk,
                                                          char* marker, int64_t pm, int cols, int hu, int vec) {
    const int x = (blockIdx.x * 256 + (int)threadIdx.x) * U8V, y = blockIdx.y;
    if (x >= cols) return;
    const bool vc = vec != 0;
    U8x16 v = u8x16_load(img + (int64_t)y * pi, x, cols, vc);
    U8x16 m;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        v.w[i] = ~v.w[i];
        uint32_t mw = 0u;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int b = (int)((v.w[i] >> (8 * j)) & 0xFFu);
            mw |= (uint32_t)(b > hu ? b - hu : 0) << (8 * j);
        }
        m.w[i] = mw;
    }
    u8x16_store(mask + (int64_t)y * pk, x, cols, vc, v);
    u8x16_store(marker + (int64_t)y * pm, x, cols, vc, m);
}

// bwselect through a labelling: classes = 1 on the nonzero pixels, weight = 1
// where a nonzero pixel holds a nonzero seed, the fold plane zeroed; 16 pixels
// per thread.
__global__ __launch_bounds__(256) void bwsel_prep16_kernel(const char* img, int64_t pi, const char* seeds, int64_t ps,
                                                           char* cls, int64_t pc, char* wgt, int64_t pw, char* fold,
                                                           int64_t pf, int cols, int vec) {
    const int x = (blockIdx.x * 256 + (int)threadIdx.x) * U8V, y = blockIdx.y;
    if (x >= cols) return;
    const bool vc = vec != 0;
    const U8x16 a = u8x16_load(img + (int64_t)y * pi, x, cols, vc);
    const U8x16 sd = u8x16_load(seeds + (int64_t)y * ps, x, cols, vc);
    U8x16 c, w;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t na = ((a.w[i] | ((a.w[i] & 0x7F7F7F7Fu) + 0x7F7F7F7Fu)) & 0x80808080u) >> 7;
        const uint32_t ns = ((sd.w[i] | ((sd.w[i] & 0x7F7F7F7Fu) + 0x7F7F7F7Fu)) & 0x80808080u) >> 7;
        c.w[i] = na;
        w.w[i] = na & ns;
    }
    u8x16_store(cls + (int64_t)y * pc, x, cols, vc, c);
    u8x16_store(wgt + (int64_t)y * pw, x, cols, vc, w);
    const U8x16 z{{0u, 0u, 0u, 0u}};
    u8x16_store(fold + (int64_t)y * pf, x, cols, vc, z);
}

// u8 out = 255 - out in place (op 2), out = image | out (op 0), image & out
// (op 1), or the image where out is nonzero and 0 elsewhere (op 3); 16 pixels
// per thread.
__global__ __launch_bounds__(256) void u8_op16_kernel(const char* img, int64_t pi, char* out, int64_t po, int cols,
                                                      int op, int vec) {
    const int x = (blockIdx.x * 256 + (int)threadIdx.x) * U8V, y = blockIdx.y;
    if (x >= cols) return;
    const bool vc = vec != 0;
    char* orow = out + (int64_t)y * po;
    U8x16 o = u8x16_load(orow, x, cols, vc);
    if (op == 2) {
#pragma unroll
        for (int i = 0; i < 4; ++i) o.w[i] = ~o.w[i];
    } else {
        const U8x16 a = u8x16_load(img + (int64_t)y * pi, x, cols, vc);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t nz = ((o.w[i] | ((o.w[i] & 0x7F7F7F7Fu) + 0x7F7F7F7Fu)) & 0x80808080u) >> 7;
            o.w[i] = op == 0 ? (a.w[i] | o.w[i]) : op == 1 ? (a.w[i] & o.w[i]) : (a.w[i] & (nz * 0xFFu));
        }
    }
    u8x16_store(orow, x, cols, vc, o);
}

// out = invert(out) in place (u8 255 - x, f32 0 - x).
__global__ void invert_kernel(char* out, int64_t po, int rows, int cols, int dtype) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= cols) return;
    char* row = out + (int64_t)y * po;
    if (dtype == NSX_U8) ((uint8_t*)row)[x] = (uint8_t)(255 - ((uint8_t*)row)[x]);
    else ((float*)row)[x] = 0.0f - ((float*)row)[x];
}

__global__ void fill_i32_kernel(char* out, int64_t pout, int rows, int cols, int32_t v) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= cols) return;
    ((int32_t*)(out + (int64_t)y * pout))[x] = v;
}

}  // namespace

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

nsx_status seeded_marker_run(nsx_ctx* ctx, const View& img, const View& seeds, const View& marker, const View* mask,
                             hipStream_t s) {
    if (img.rows == 0 || img.cols == 0) return NSX_OK;
    const bool vec = u8_vec({(uintptr_t)img.data, (uintptr_t)img.pitch, (uintptr_t)seeds.data, (uintptr_t)seeds.pitch,
                             (uintptr_t)marker.data, (uintptr_t)marker.pitch, mask ? (uintptr_t)mask->data : 0u,
                             mask ? (uintptr_t)mask->pitch : 0u});
    hipLaunchKernelGGL(seeded_marker16_kernel, u8x16_grid(img.rows, img.cols), dim3(256), 0, s, img.data, img.pitch,
                       seeds.data, seeds.pitch, marker.data, marker.pitch, mask ? mask->data : nullptr,
                       mask ? mask->pitch : 0, img.cols, mask ? 1 : 0, vec ? 1 : 0);
    NSX_HIP_TRY(ctx, hipGetLastError());
    return NSX_OK;
}

nsx_status hmin_prep_run(nsx_ctx* ctx, const View& img, double h, const View& mask, const View& marker,
                         hipStream_t s) {
    if (img.rows == 0 || img.cols == 0) return NSX_OK;
    if (img.dtype == NSX_U8) {
        const bool vec = u8_vec({(uintptr_t)img.data, (uintptr_t)img.pitch, (uintptr_t)mask.data,
                                 (uintptr_t)mask.pitch, (uintptr_t)marker.data, (uintptr_t)marker.pitch});
        hipLaunchKernelGGL(hmin_prep16_kernel, u8x16_grid(img.rows, img.cols), dim3(256), 0, s, img.data, img.pitch,
                           mask.data, mask.pitch, marker.data, marker.pitch, img.cols, (int)h, vec ? 1 : 0);
        NSX_HIP_TRY(ctx, hipGetLastError());
        return NSX_OK;
    }
    const dim3 grid((img.cols + 255) / 256, img.rows);
    hipLaunchKernelGGL(hmin_prep_kernel, grid, dim3(256), 0, s, img.data, img.pitch, mask.data, mask.pitch,
                       marker.data, marker.pitch, img.rows, img.cols, img.dtype, (int)h, (float)h);
    NSX_HIP_TRY(ctx, hipGetLastError());
    return NSX_OK;
}

nsx_status bbox_run(nsx_ctx* ctx, const View& lab, int bgval, int32_t* boxes, int capacity, int32_t* count,
                    hipStream_t s) {
    *count = 0;
    const int64_t npix = (int64_t)lab.rows * lab.cols;
    if (npix == 0) return NSX_OK;
    void *misc_p, *tab_p, *bits_p, *scan_p, *out_p;
    NSX_TRY(ctx->ws(WS_CCL_MISC, 64, &misc_p));
    int32_t* misc = (int32_t*)misc_p;
    const dim3 grid((lab.cols + 63) / 64, (lab.rows + BB_ROWS - 1) / BB_ROWS);
    const int vec = ((((uintptr_t)lab.data | (uintptr_t)lab.pitch) & 15u) == 0u) ? 1 : 0;
    hipLaunchKernelGGL(bbox_misc_init_kernel, dim3(1), dim3(64), 0, s, misc);
    hipLaunchKernelGGL(bbox_scan_kernel, dim3(std::min(lab.rows, 8 * ctx->cu_count)), dim3(256), 0, s, lab.data,
                       lab.pitch, lab.rows, lab.cols, bgval, vec, misc);
    NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned + 40, misc, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
    const int32_t stop = ctx->h_pinned[40], bad = ctx->h_pinned[41], maxv = ctx->h_pinned[42];
    if (bad < stop)
        return ctx->fail(NSX_ERR_UNSUPPORTED, "boundingBox labels must lie in [-1, rows*cols] (or be the background)");
    const int64_t cap = (int64_t)maxv + 1;
    const int64_t nwords = (npix + 31) / 32, nb = (nwords + 1023) / 1024;
    NSX_TRY(ctx->ws(WS_CCL_LABEL, sizeof(int32_t) * 5 * (size_t)cap, &tab_p));
    NSX_TRY(ctx->ws(WS_CCL_BITS, sizeof(uint32_t) * (size_t)nwords, &bits_p));
    NSX_TRY(ctx->ws(WS_CCL_SCAN, sizeof(uint32_t) * (size_t)(nwords + nb + 1), &scan_p));
    NSX_TRY(ctx->ws(WS_CCL_AREA, sizeof(int32_t) * 5 * (size_t)cap, &out_p));
    int32_t* tab = (int32_t*)tab_p;
    uint32_t* bits = (uint32_t*)bits_p;
    uint32_t* wordpre = (uint32_t*)scan_p;
    uint32_t* blockpre = wordpre + nwords;
    const int g = 4 * ctx->cu_count;
    hipLaunchKernelGGL(bbox_init_kernel, dim3(g), dim3(256), 0, s, tab, cap);
    hipLaunchKernelGGL(bbox_runs_kernel, grid, dim3(256), 0, s, lab.data, lab.pitch, lab.rows, lab.cols, bgval,
                       (const int32_t*)misc, tab, cap);
    hipLaunchKernelGGL(zero_words_kernel, dim3(g), dim3(256), 0, s, bits, nwords);
    hipLaunchKernelGGL(bbox_mark_kernel, dim3(g), dim3(256), 0, s, (const int32_t*)tab, cap, bits);
    hipLaunchKernelGGL(scan_words_kernel, dim3(nb), dim3(256), 0, s, (const uint32_t*)bits, nwords, wordpre, blockpre);
    hipLaunchKernelGGL(scan_blocks_kernel, dim3(1), dim3(1024), 0, s, blockpre, (int)nb);
    NSX_HIP_TRY(ctx, hipGetLastError());
    NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned + 40, blockpre + nb, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
    const int32_t n = ctx->h_pinned[40];
    *count = n;
    if (!boxes || n == 0) return NSX_OK;
    if (n > capacity) return ctx->fail(NSX_ERR_INVALID, "boxes holds fewer components than the image has");
    int32_t* out = (int32_t*)out_p;
    hipLaunchKernelGGL(bbox_scatter_kernel, dim3(g), dim3(256), 0, s, (const int32_t*)tab, cap,
                       (const uint32_t*)bits, (const uint32_t*)wordpre, (const uint32_t*)blockpre, (int)nb, out);
    NSX_HIP_TRY(ctx, hipGetLastError());
    NSX_HIP_TRY(ctx, hipMemcpyAsync(boxes, out, sizeof(int32_t) * 5 * (size_t)n, hipMemcpyDeviceToHost, s));
    NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
    return NSX_OK;
}

nsx_status local_extrema_run(nsx_ctx* ctx, const View& img, const View& out, int conn, bool minima, hipStream_t s) {
    if (img.rows == 0 || img.cols == 0) return NSX_OK;
    const size_t es = img.dtype == NSX_U8 ? 1 : 4;
    const size_t plane = (size_t)img.rows * img.cols * es;
    void *work_p, *marker_p, *rec_p;
    NSX_TRY(ctx->ws(WS_CHAIN_FILL, plane, &work_p));
    NSX_TRY(ctx->ws(WS_CHAIN_RECON, plane, &marker_p));
    NSX_TRY(ctx->ws(WS_STAGE_A, plane, &rec_p));
    const int64_t pitch = (int64_t)img.cols * (int64_t)es;
    const View work{(char*)work_p, pitch, img.rows, img.cols, img.dtype};
    const View marker{(char*)marker_p, pitch, img.rows, img.cols, img.dtype};
    const View rec{(char*)rec_p, pitch, img.rows, img.cols, img.dtype};
    if (img.dtype == NSX_U8) {
        // u8: plateaus without a higher neighbour (see lx_prep_u8_kernel);
        // d_stats int32 word 56 holds "some work pixel is nonzero"
        uint32_t* nonzero = (uint32_t*)ctx->d_stats + 56;
        const View& cls = work;        // plateau classes
        const View& weight = marker;   // 0/1 per pixel
        const View& fold = rec;        // per plateau: largest weight
        const bool vec = ((((uintptr_t)img.data | (uintptr_t)img.pitch | (uintptr_t)work.data | (uintptr_t)pitch |
                            (uintptr_t)weight.data | (uintptr_t)fold.data | (uintptr_t)out.data |
                            (uintptr_t)out.pitch) & 15u) == 0u);
        const dim3 vgrid = u8x16_grid(img.rows, img.cols);
        hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, nonzero, 1);
        hipLaunchKernelGGL(lx_prep_u8_kernel, vgrid, dim3(256), 0, s, img.data, img.pitch, work.data, work.pitch,
                           weight.data, weight.pitch, fold.data, fold.pitch, img.rows, img.cols, minima ? 1 : 0, conn,
                           vec ? 1 : 0, nonzero);
        NSX_HIP_TRY(ctx, hipGetLastError());
        NSX_TRY(ccl_class_fold_run(ctx, cls, weight, fold, conn, s));
        hipLaunchKernelGGL(lx_out_u8_kernel, vgrid, dim3(256), 0, s, img.data, img.pitch, fold.data, fold.pitch,
                           out.data, out.pitch, img.rows, img.cols, minima ? 1 : 0, vec ? 1 : 0,
                           (const uint32_t*)nonzero);
        NSX_HIP_TRY(ctx, hipGetLastError());
        return NSX_OK;
    }
    const dim3 grid((img.cols + 255) / 256, img.rows);
    hipLaunchKernelGGL(lx_prep_kernel, grid, dim3(256), 0, s, img.data, img.pitch, work.data, work.pitch, marker.data,
                       marker.pitch, img.rows, img.cols, img.dtype, minima ? 1 : 0);
    NSX_HIP_TRY(ctx, hipGetLastError());
    NSX_TRY(recon_run(ctx, RM_RECON, &marker, work, rec, conn, s));
    // after the reconstruction, whose level-set path may use the labelling slots
    void *flags_p, *misc_p;
    NSX_TRY(ctx->ws(WS_CCL_AREA, (size_t)img.rows * img.cols, &flags_p));
    NSX_TRY(ctx->ws(WS_CCL_MISC, 64, &misc_p));
    uint8_t* flags = (uint8_t*)flags_p;
    int32_t* changed = (int32_t*)misc_p;
    hipLaunchKernelGGL(lx_mark_kernel, grid, dim3(256), 0, s, work.data, work.pitch, marker.data, marker.pitch,
                       rec.data, rec.pitch, flags, img.rows, img.cols, img.dtype, conn);
    // spread the marks over equal-valued candidates, tile by tile; each round
    // with a change marks at least one more pixel, so the rounds end
    const int64_t npix = (int64_t)img.rows * img.cols;
    const int tiles_x = (img.cols + LX_T - 1) / LX_T, tiles_y = (img.rows + LX_T - 1) / LX_T;
    const int ntiles = tiles_x * tiles_y;
    void* tf_p;
    NSX_TRY(ctx->ws(WS_CCL_BITS, 2 * (size_t)ntiles, &tf_p));
    uint8_t* fa = (uint8_t*)tf_p;
    uint8_t* fb = fa + ntiles;
    hipLaunchKernelGGL(lx_tiles_init_kernel, dim3(std::min(1024, (ntiles + 255) / 256)), dim3(256), 0, s, fa, ntiles);
    for (int64_t round = 0; round <= npix; ++round) {
        hipLaunchKernelGGL(lx_zero_kernel, dim3(1), dim3(64), 0, s, changed);
        hipLaunchKernelGGL(lx_tile_kernel, dim3(ntiles), dim3(256), 0, s, work.data, work.pitch, flags, img.rows,
                           img.cols, img.dtype, conn, tiles_x, tiles_y, (const uint8_t*)fa, fb, changed);
        NSX_HIP_TRY(ctx, hipGetLastError());
        NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned + 40, changed, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
        if (ctx->h_pinned[40] == 0) break;
        uint8_t* t = fa;
        fa = fb;
        fb = t;
    }
    hipLaunchKernelGGL(lx_out_kernel, grid, dim3(256), 0, s, (const uint8_t*)flags, out.data, out.pitch, img.rows,
                       img.cols);
    NSX_HIP_TRY(ctx, hipGetLastError());
    return NSX_OK;
}

nsx_status invert_run(nsx_ctx* ctx, const View& out, hipStream_t s) {
    if (out.rows == 0 || out.cols == 0) return NSX_OK;
    if (out.dtype == NSX_U8) {
        const bool vec = u8_vec({(uintptr_t)out.data, (uintptr_t)out.pitch});
        hipLaunchKernelGGL(u8_op16_kernel, u8x16_grid(out.rows, out.cols), dim3(256), 0, s, (const char*)out.data,
                           out.pitch, out.data, out.pitch, out.cols, 2, vec ? 1 : 0);
        NSX_HIP_TRY(ctx, hipGetLastError());
        return NSX_OK;
    }
    const dim3 grid((out.cols + 255) / 256, out.rows);
    hipLaunchKernelGGL(invert_kernel, grid, dim3(256), 0, s, out.data, out.pitch, out.rows, out.cols, out.dtype);
    NSX_HIP_TRY(ctx, hipGetLastError());
    return NSX_OK;
}

nsx_status bitwise_u8_run(nsx_ctx* ctx, const View& img, const View& out, bool and_op, hipStream_t s) {
    if (img.rows == 0 || img.cols == 0) return NSX_OK;
    const bool vec = u8_vec({(uintptr_t)img.data, (uintptr_t)img.pitch, (uintptr_t)out.data, (uintptr_t)out.pitch});
    hipLaunchKernelGGL(u8_op16_kernel, u8x16_grid(img.rows, img.cols), dim3(256), 0, s, img.data, img.pitch, out.data,
                       out.pitch, img.cols, and_op ? 1 : 0, vec ? 1 : 0);
    NSX_HIP_TRY(ctx, hipGetLastError());
    return NSX_OK;
}

// imreconstructBinary through a labelling: classes = 1 on the nonzero mask
// pixels, weight = 1 where such a pixel has a nonzero marker, the fold plane
// zeroed; then out = the marker where it is nonzero, the type's maximum at the
// other pixels whose component took a weight, 0 elsewhere.  Four pixels per
// thread along a row.
template <typename T>
__global__ __launch_bounds__(256) void binrec_prep_kernel(const char* marker, int64_t pm, const char* mask, int64_t pk,
                                                          uint8_t* cls, uint8_t* wgt, uint8_t* fold, int cols) {
    const int x0 = 4 * (blockIdx.x * 256 + (int)threadIdx.x), y = blockIdx.y;
    const T* a = (const T*)(marker + (int64_t)y * pm);
    const T* m = (const T*)(mask + (int64_t)y * pk);
    const int64_t row = (int64_t)y * cols;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int x = x0 + k;
        if (x >= cols) break;
        const bool in = m[x] != T(0);
        cls[row + x] = in ? 1 : 0;
        wgt[row + x] = (in && a[x] != T(0)) ? 1 : 0;
        fold[row + x] = 0;
    }
}
template <typename T>
__global__ __launch_bounds__(256) void binrec_out_kernel(const char* marker, int64_t pm, const uint8_t* fold, char* out,
                                                         int64_t po, int cols, T top) {
    const int x0 = 4 * (blockIdx.x * 256 + (int)threadIdx.x), y = blockIdx.y;
    const T* a = (const T*)(marker + (int64_t)y * pm);
    T* o = (T*)(out + (int64_t)y * po);
    const int64_t row = (int64_t)y * cols;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int x = x0 + k;
        if (x >= cols) break;
        const T av = a[x];
        o[x] = av != T(0) ? av : (fold[row + x] ? top : T(0));
    }
}

nsx_status binrecon_run(nsx_ctx* ctx, const View& marker, const View& mask, const View& out, int conn, hipStream_t s) {
    if (mask.rows == 0 || mask.cols == 0) return NSX_OK;
    const size_t plane = (size_t)mask.rows * mask.cols;
    void *cls_p, *wgt_p, *fold_p;
    NSX_TRY(ctx->ws(WS_CHAIN_RECON, plane, &cls_p));
    NSX_TRY(ctx->ws(WS_CHAIN_FILL, plane, &wgt_p));
    NSX_TRY(ctx->ws(WS_BP_L, plane, &fold_p));
    const View cls{(char*)cls_p, (int64_t)mask.cols, mask.rows, mask.cols, NSX_U8};
    const View wgt{(char*)wgt_p, (int64_t)mask.cols, mask.rows, mask.cols, NSX_U8};
    const View fold{(char*)fold_p, (int64_t)mask.cols, mask.rows, mask.cols, NSX_U8};
    const dim3 grid((unsigned)((mask.cols + 1023) / 1024), (unsigned)mask.rows);
    if (mask.dtype == NSX_U8)
        hipLaunchKernelGGL(binrec_prep_kernel<uint8_t>, grid, dim3(256), 0, s, marker.data, marker.pitch, mask.data,
                           mask.pitch, (uint8_t*)cls_p, (uint8_t*)wgt_p, (uint8_t*)fold_p, mask.cols);
    else
        hipLaunchKernelGGL(binrec_prep_kernel<int32_t>, grid, dim3(256), 0, s, marker.data, marker.pitch, mask.data,
                           mask.pitch, (uint8_t*)cls_p, (uint8_t*)wgt_p, (uint8_t*)fold_p, mask.cols);
    NSX_HIP_TRY(ctx, hipGetLastError());
    NSX_TRY(ccl_class_fold_run(ctx, cls, wgt, fold, conn, s));
    if (mask.dtype == NSX_U8)
        hipLaunchKernelGGL(binrec_out_kernel<uint8_t>, grid, dim3(256), 0, s, marker.data, marker.pitch,
                           (const uint8_t*)fold_p, out.data, out.pitch, mask.cols, (uint8_t)255);
    else
        hipLaunchKernelGGL(binrec_out_kernel<int32_t>, grid, dim3(256), 0, s, marker.data, marker.pitch,
                           (const uint8_t*)fold_p, out.data, out.pitch, mask.cols, (int32_t)INT32_MAX);
    NSX_HIP_TRY(ctx, hipGetLastError());
    return NSX_OK;
}

nsx_status bwselect_run(nsx_ctx* ctx, const View& img, const View& seeds, const View& out, int conn, hipStream_t s) {
    if (img.rows == 0 || img.cols == 0) return NSX_OK;
    const size_t plane = (size_t)img.rows * img.cols;
    void *cls_p, *wgt_p;
    NSX_TRY(ctx->ws(WS_CHAIN_RECON, plane, &cls_p));
    NSX_TRY(ctx->ws(WS_CHAIN_FILL, plane, &wgt_p));
    const View cls{(char*)cls_p, (int64_t)img.cols, img.rows, img.cols, NSX_U8};
    const View wgt{(char*)wgt_p, (int64_t)img.cols, img.rows, img.cols, NSX_U8};
    const bool vec = u8_vec({(uintptr_t)img.data, (uintptr_t)img.pitch, (uintptr_t)seeds.data, (uintptr_t)seeds.pitch,
                             (uintptr_t)cls.data, (uintptr_t)wgt.data, (uintptr_t)img.cols, (uintptr_t)out.data,
                             (uintptr_t)out.pitch});
    const dim3 grid = u8x16_grid(img.rows, img.cols);
    hipLaunchKernelGGL(bwsel_prep16_kernel, grid, dim3(256), 0, s, img.data, img.pitch, seeds.data, seeds.pitch,
                       cls.data, cls.pitch, wgt.data, wgt.pitch, out.data, out.pitch, img.cols, vec ? 1 : 0);
    NSX_HIP_TRY(ctx, hipGetLastError());
    NSX_TRY(ccl_class_fold_run(ctx, cls, wgt, out, conn, s));
    hipLaunchKernelGGL(u8_op16_kernel, grid, dim3(256), 0, s, img.data, img.pitch, out.data, out.pitch, img.cols, 3,
                       vec ? 1 : 0);
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
