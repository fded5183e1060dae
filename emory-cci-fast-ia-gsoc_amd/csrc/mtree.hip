// mtree.hip -- u8 reconstruction through the mask's component tree (max-tree).
//
// Reconstruction by dilation is a statement about the upper level sets of the
// mask g: the result at x is
//     R(x) = max over levels L <= g(x) of min(L, largest marker value in the
//            connected component of {g >= L} that holds x)
// (threshold decomposition of MorphologicOperations.cpp:665-838's fixed
// point).  Those components, nested over all levels, are the nodes of the max-
// tree of g.  With M(N) the largest (clamped) marker value in node N's subtree:
//     R(N) = max over the ancestors A of N (N included) of min(level(A), M(A)).
// Along the path from N to the root level falls and M rises, so the maximum
// is found at the first ancestor A* with M(A*) >= level(A*):
//     R(N) = max(M of the node before A*, level(A*))   (the root if none).
// This costs a fixed number of launches whatever the image holds, unlike
// propagation, whose rounds grow with the distances values travel.
//
// Nodes are identified by keys ordered by (level, id): a node's parent is
// always lower.  Kernels:
//   MA tile   64x64 tile per 512-thread block.  The tile's max-tree is built in
//             LDS from singletons by merging the trees across every in-tile edge
//             (connect: walk the higher chain down to the insertion point of the
//             lower node, insert it with a CAS, continue merging below; the
//             standard merge of component trees, Wilkinson et al., IEEE TPAMI
//             2008).  Equal-level chains are shortened by CAS path halving.  Then
//             canonical parents (every pixel at its level root, every level root
//             at its parent's level root), subtree maxima of the marker (walk up
//             with atomicMax, stop where the value is already there), level
//             roots numbered in raster order, and the tile's nodes exported:
//             parent, level, subtree maximum; every pixel's node (u16).
//   MB merge  one thread per tile-border pixel: the same connect, over the
//             exported nodes in HBM (global CAS), for every edge that crosses
//             a tile border.  Nodes whose parent changed are listed.
//   MC walk   every listed node carries its subtree maximum up to its new
//             ancestors (atomicMax, stop where the value is already there).
//   MD rnode  every node: R(N) by the walk to its first ancestor with M >= level.
//   ME out    every pixel: R of its node (hole filling: 255 - R).
// Hole filling (imfillHoles, :1859-1926) is the same computation on
// g = 255 - image with the frame as the only marker: every pixel on the image
// border is seeded with its own value g (the padded marker's max<uchar> frame,
// clamped by the mask), all others with 0.
// Every loop is bounded; a bound that is hit sets an error bit, and the caller
// then computes the result by propagation instead.
#include "nsx_internal.h"

#include <algorithm>
#include <climits>

namespace nsx {
namespace {

constexpr int MT = 64;                 // tile side
constexpr int MN = MT * MT;            // pixels (and node slots) per tile
constexpr int MTHREADS = 512;
constexpr int MWAVES = MTHREADS / 64;
constexpr int MPT = MN / MTHREADS;     // pixels per thread
constexpr uint32_t KINV = 0xFFFFFFFFu; // pixel outside the image
constexpr uint32_t LIMIT_LDS = 1u << 16;
constexpr uint32_t LIMIT_G = 1u << 20;
constexpr int MB_SLOTS = 192;          // border pixels examined per tile

enum { MTE_TILE = 1, MTE_MERGE = 2, MTE_WALK = 4, MTE_RNODE = 8, MTE_WLIST = 16 };
enum { MT_RECON = 0, MT_FILL = 1 };

struct MtImg {
    const uint8_t* mask; int64_t pk;     // RECON: the mask; FILL: the image whose holes are filled
    const uint8_t* marker; int64_t pm;   // RECON only
    uint8_t* out; int64_t po;
    int rows, cols, ntx, nty;
    int tile_base;
    int pad;
};

struct MtArgs {
    MtImg img0;
    const MtImg* imgs;
    const int32_t* tile_img;
    int nimg, ntiles;
    int32_t* gpar;      // [ntiles * MN] parent node of each node
    uint32_t* gM;       // [ntiles * MN] largest marker value in the node's subtree
    uint8_t* gG;        // [ntiles * MN] level of the node
    uint8_t* gR;        // [ntiles * MN] reconstruction value of the node
    uint16_t* lnode;    // [ntiles * MN] node (in its tile) of each pixel, tile-major
    int32_t* tcount;    // [ntiles] nodes per tile
    int32_t* troot;     // [ntiles] root node of each tile's tree
    int32_t* walk;      // nodes whose parent changed in the merge
    uint32_t* ctl;      // [0] walk count, [1] error bits
    uint32_t walk_cap;
};

__device__ __forceinline__ int xcd_tile(int b, int G) {
    const int x = b & 7;
    return x * (G >> 3) + min(x, G & 7) + (b >> 3);
}

__device__ __forceinline__ void tile_of(const MtArgs& A, int t, MtImg& D, int& tx, int& ty) {
    t = __builtin_amdgcn_readfirstlane(t);
    if (A.nimg == 1) {
        D = A.img0;
    } else {
        const int idx = __builtin_amdgcn_readfirstlane(A.tile_img[t]);
        D = A.imgs[idx];
    }
    const int lt = t - D.tile_base;
    ty = lt / D.ntx;
    tx = lt - ty * D.ntx;
}

// ---- tile build in LDS: keys are (level << 12) | pixel index -------------------------
__device__ __forceinline__ uint32_t lds_levroot(uint32_t* par, uint32_t v) {
    uint32_t p = par[v & 0xFFFu];
    // parents are always lower keys, so the walk ends
    while (p != v && (p >> 12) == (v >> 12)) {
        const uint32_t pp = par[p & 0xFFFu];
        if (pp != p && (pp >> 12) == (v >> 12)) atomicCAS(&par[v & 0xFFFu], p, pp);
        v = p;
        p = pp;
    }
    return v;
}

__device__ __forceinline__ uint32_t lds_levroot_ro(const uint32_t* par, uint32_t v) {
    uint32_t p = par[v & 0xFFFu];
    while (p != v && (p >> 12) == (v >> 12)) {
        v = p;
        p = par[v & 0xFFFu];
    }
    return v;
}

// Merges the component trees holding a and b (adjacent pixels).  False when
// the iteration bound is hit.
__device__ bool lds_connect(uint32_t* par, uint32_t a, uint32_t b) {
    uint32_t x = lds_levroot(par, a), y = lds_levroot(par, b);
    for (uint32_t it = 0; it < LIMIT_LDS; ++it) {
        if (x == y) return true;
        if (x < y) { const uint32_t t = x; x = y; y = t; }   // x the higher node
        const uint32_t px = par[x & 0xFFFu];
        if (px == x) {                                        // x a root: hang it under y
            if (atomicCAS(&par[x & 0xFFFu], x, y) == x) return true;
            x = lds_levroot(par, x);
            y = lds_levroot(par, y);
            continue;
        }
        const uint32_t z = lds_levroot(par, px);
        if (z == y) return true;
        if (z > y) { x = z; continue; }                       // y belongs further down
        if (atomicCAS(&par[x & 0xFFFu], px, y) == px) {       // insert y between x and z
            x = y;
            y = z;
            continue;
        }
        x = lds_levroot(par, x);
        y = lds_levroot(par, y);
    }
    return false;
}

template <int MODE, int CONN>
__global__ __launch_bounds__(MTHREADS) void mt_tile_kernel(MtArgs A) {
    __shared__ uint32_t par[MN];
    __shared__ uint32_t M[MN];
    __shared__ uint16_t nid[MN];
    __shared__ uint8_t gl[MN];
    __shared__ uint32_t rowcnt[MT];
    __shared__ uint32_t fail;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int T = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    MtImg D;
    int tx, ty;
    tile_of(A, T, D, tx, ty);
    const int x0 = tx * MT, y0 = ty * MT;
    const int gx = x0 + lane;
    if (tid == 0) fail = 0;
    uint32_t key[MPT];
    // loads first (all in flight together), then the LDS writes
    uint32_t gv[MPT], fv[MPT];
#pragma unroll
    for (int j = 0; j < MPT; ++j) {
        const int gy = y0 + w + MWAVES * j;
        const bool valid = gy < D.rows && gx < D.cols;
        uint32_t g = 0, f = 0;
        if (valid) {
            const uint32_t m = D.mask[(int64_t)gy * D.pk + gx];
            if (MODE == MT_FILL) {
                g = 255u - m;
                f = (gy == 0 || gx == 0 || gy == D.rows - 1 || gx == D.cols - 1) ? g : 0u;
            } else {
                g = m;
                f = min((uint32_t)D.marker[(int64_t)gy * D.pm + gx], m);
            }
        }
        gv[j] = valid ? g : 256u;
        fv[j] = f;
    }
#pragma unroll
    for (int j = 0; j < MPT; ++j) {
        const int i = (w + MWAVES * j) * MT + lane;
        key[j] = gv[j] < 256u ? (gv[j] << 12) | (uint32_t)i : KINV;
        par[i] = key[j];
        M[i] = fv[j];
        gl[i] = (uint8_t)(gv[j] < 256u ? gv[j] : 0u);
    }
    __syncthreads();
    // merge the trees across every in-tile edge (right, down; 8-conn also the
    // two lower diagonals)
    bool ok = true;
#pragma unroll 1
    for (int j = 0; j < MPT; ++j) {
        if (key[j] == KINV) continue;
        const int ly = w + MWAVES * j;
        const int i = ly * MT + lane;
        const bool r_ok = lane + 1 < MT && gx + 1 < D.cols;
        const bool d_ok = ly + 1 < MT && y0 + ly + 1 < D.rows;
        if (r_ok) ok &= lds_connect(par, key[j], ((uint32_t)gl[i + 1] << 12) | (uint32_t)(i + 1));
        if (d_ok) ok &= lds_connect(par, key[j], ((uint32_t)gl[i + MT] << 12) | (uint32_t)(i + MT));
        if (CONN == 8 && d_ok) {
            if (lane > 0) ok &= lds_connect(par, key[j], ((uint32_t)gl[i + MT - 1] << 12) | (uint32_t)(i + MT - 1));
            if (r_ok) ok &= lds_connect(par, key[j], ((uint32_t)gl[i + MT + 1] << 12) | (uint32_t)(i + MT + 1));
        }
    }
    if (!ok) fail = 1;
    __syncthreads();
    // canonical parents: found read-only, written after a barrier
    uint32_t lr[MPT], cp[MPT];
#pragma unroll
    for (int j = 0; j < MPT; ++j) {
        lr[j] = cp[j] = KINV;
        if (key[j] == KINV) continue;
        lr[j] = lds_levroot_ro(par, key[j]);
        if (lr[j] == key[j]) {
            const uint32_t p = par[key[j] & 0xFFFu];
            cp[j] = p == key[j] ? p : lds_levroot_ro(par, p);
        } else {
            cp[j] = lr[j];
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < MPT; ++j)
        if (key[j] != KINV) par[key[j] & 0xFFFu] = cp[j];
    __syncthreads();
    // subtree maxima: every pixel's marker value into its level root, then each
    // level root carries its node's maximum up until an ancestor holds it
    // already (whoever wrote that value carries it further)
#pragma unroll
    for (int j = 0; j < MPT; ++j)
        if (key[j] != KINV && lr[j] != key[j] && fv[j] > 0u) atomicMax(&M[lr[j] & 0xFFFu], fv[j]);
    __syncthreads();
#pragma unroll 1
    for (int j = 0; j < MPT; ++j) {
        if (key[j] == KINV || lr[j] != key[j]) continue;
        uint32_t a = key[j];
        const uint32_t m = M[a & 0xFFFu];
        for (uint32_t it = 0; it < MN; ++it) {
            const uint32_t p = par[a & 0xFFFu];
            if (p == a) break;
            if (atomicMax(&M[p & 0xFFFu], m) >= m) break;
            a = p;
        }
    }
    __syncthreads();
    // number the level roots in raster order
#pragma unroll
    for (int j = 0; j < MPT; ++j) {
        const uint64_t b = __ballot(key[j] != KINV && lr[j] == key[j]);
        if (lane == 0) rowcnt[w + MWAVES * j] = (uint32_t)__builtin_popcountll(b);
    }
    __syncthreads();
    if (w == 0) {
        const uint32_t c = rowcnt[lane];
        uint32_t incl = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        rowcnt[lane] = incl - c;
        if (lane == 63) A.tcount[T] = (int32_t)incl;
    }
    __syncthreads();
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int j = 0; j < MPT; ++j) {
        const bool isr = key[j] != KINV && lr[j] == key[j];
        const uint64_t b = __ballot(isr);
        if (isr) nid[key[j] & 0xFFFu] = (uint16_t)(rowcnt[w + MWAVES * j] + (uint32_t)__builtin_popcountll(b & below));
    }
    __syncthreads();
    // export the nodes and every pixel's node
    const int64_t base = (int64_t)T * MN;
#pragma unroll
    for (int j = 0; j < MPT; ++j) {
        if (key[j] == KINV) continue;
        const int i = (w + MWAVES * j) * MT + lane;
        const uint32_t n = nid[lr[j] & 0xFFFu];
        A.lnode[base + i] = (uint16_t)n;
        if (lr[j] == key[j]) {
            const int64_t g = base + n;
            A.gpar[g] = (int32_t)(cp[j] == key[j] ? g : base + nid[cp[j] & 0xFFFu]);
            if (cp[j] == key[j]) A.troot[T] = (int32_t)g;
            A.gG[g] = (uint8_t)(key[j] >> 12);
            A.gM[g] = M[i];
        }
    }
    if (tid == 0 && fail) atomicOr(&A.ctl[1], (uint32_t)MTE_TILE);
}

// ---- global nodes --------------------------------------------------------------------
__device__ __forceinline__ int32_t ld_par(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// (level, id) order of nodes
__device__ __forceinline__ bool g_lower(int32_t a, uint32_t ga, int32_t b, uint32_t gb) {
    return ga < gb || (ga == gb && a < b);
}

// Level root of node v (level gv), with CAS path halving along the level.
__device__ __forceinline__ int32_t g_levroot(int32_t* gpar, const uint8_t* gG, int32_t v, uint32_t gv) {
    int32_t p = ld_par(&gpar[v]);
    for (uint32_t it = 0; it < LIMIT_G && p != v && (uint32_t)gG[p] == gv; ++it) {
        const int32_t pp = ld_par(&gpar[p]);
        if (pp != p && (uint32_t)gG[pp] == gv) atomicCAS(&gpar[v], p, pp);
        v = p;
        p = pp;
    }
    return v;
}

__device__ __forceinline__ int32_t g_levroot_ro(const int32_t* gpar, const uint8_t* gG, int32_t v) {
    const uint32_t gv = gG[v];
    int32_t p = gpar[v];
    for (uint32_t it = 0; it < LIMIT_G && p != v && (uint32_t)gG[p] == gv; ++it) {
        v = p;
        p = gpar[v];
    }
    return v;
}

__device__ __forceinline__ void push_walker(const MtArgs& A, int32_t x) {
    const uint32_t k = atomicAdd(&A.ctl[0], 1u);
    if (k < A.walk_cap) A.walk[k] = x;
    else atomicOr(&A.ctl[1], (uint32_t)MTE_WLIST);
}

// The in-tile connect over the exported nodes.  A failed CAS returns the
// current parent, from which the walk resumes (no reload that could keep
// returning a stale value).
// Waits about 2^nfail x 64 clocks (capped at 2^7), scaled by a hash of the
// node so that the threads that lost one CAS do not return together.
__device__ __forceinline__ void backoff(uint32_t& nfail, int32_t x) {
    nfail = nfail < 7u ? nfail + 1u : 7u;
    const uint32_t h = ((uint32_t)x * 2654435761u + threadIdx.x * 40503u) >> 25;   // 0..127
    const uint32_t n = 1u + ((h & ((1u << nfail) - 1u)));
    for (uint32_t i = 0; i < n; ++i) __builtin_amdgcn_s_sleep(1);
}

__device__ bool g_connect(const MtArgs& A, int32_t a, int32_t b) {
    int32_t* gpar = A.gpar;
    const uint8_t* gG = A.gG;
    uint32_t nfail = 0;
    uint32_t gx = gG[a], gy = gG[b];
    int32_t x = g_levroot(gpar, gG, a, gx), y = g_levroot(gpar, gG, b, gy);
    for (uint32_t it = 0; it < LIMIT_G; ++it) {
        if (x == y) return true;
        if (g_lower(x, gx, y, gy)) {
            const int32_t t = x; x = y; y = t;
            const uint32_t tg = gx; gx = gy; gy = tg;
        }
        int32_t px = ld_par(&gpar[x]);
        if (px == x) {
            const int32_t old = atomicCAS(&gpar[x], x, y);
            if (old == x) { push_walker(A, x); return true; }
            backoff(nfail, x);
            px = old;   // x has a parent now: continue as below
        }
        const uint32_t gp = gG[px];
        const int32_t z = g_levroot(gpar, gG, px, gp);
        const uint32_t gz = gp;
        if (z == y) return true;
        if (g_lower(y, gy, z, gz)) { x = z; gx = gz; continue; }
        const int32_t old = atomicCAS(&gpar[x], px, y);
        if (old == px) {
            push_walker(A, x);
            x = y; gx = gy;
            y = z; gy = gz;
            continue;
        }
        // x's parent changed under us: retry from x with the current state,
        // after a randomised exponential wait so that connects racing on one
        // chain do not all retry at once
        backoff(nfail, x);
    }
    return false;
}

// Global node id of the pixel (gy, gx) of image D.
__device__ __forceinline__ int32_t node_of(const MtArgs& A, const MtImg& D, int gy, int gx) {
    const int tx = gx / MT, ty = gy / MT;
    const int t = D.tile_base + ty * D.ntx + tx;
    const int i = (gy - ty * MT) * MT + (gx - tx * MT);
    return t * MN + (int32_t)A.lnode[(int64_t)t * MN + i];
}

// Before the tile-edge merge: the roots of every two horizontally or
// vertically adjacent tiles of an image are joined (two adjacent tiles form a
// connected set, whose lowest level is the lower root's).  The lowest chains of
// all tiles are then one chain before the edge merge starts, so its connects
// do not all race to merge them.
__global__ __launch_bounds__(256) void mt_roots_kernel(MtArgs A) {
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int T = (int)(gtid >> 1), dir = (int)(gtid & 1);
    if (T >= A.ntiles) return;
    const MtImg D = A.nimg == 1 ? A.img0 : A.imgs[A.tile_img[T]];
    const int lt = T - D.tile_base, ty = lt / D.ntx, tx = lt - ty * D.ntx;
    int U;
    if (dir == 0) { if (tx + 1 >= D.ntx) return; U = T + 1; }
    else { if (ty + 1 >= D.nty) return; U = T + D.ntx; }
    if (!g_connect(A, A.troot[T], A.troot[U])) atomicOr(&A.ctl[1], (uint32_t)MTE_MERGE);
}

template <int CONN>
__global__ __launch_bounds__(256) void mt_merge_kernel(MtArgs A) {
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int T = (int)(gtid / MB_SLOTS), slot = (int)(gtid - (int64_t)T * MB_SLOTS);
    if (T >= A.ntiles) return;
    MtImg D;
    int tx, ty;
    tile_of(A, T, D, tx, ty);
    int ly, lx;
    if (slot < 64) { ly = slot; lx = MT - 1; }                       // right column
    else if (slot < 127) { ly = MT - 1; lx = slot - 64; }            // bottom row (corner in the column)
    else if (CONN == 8 && slot >= 128 && slot < 191) { ly = slot - 128; lx = 0; }   // left column, rows 0..62
    else return;
    const int gy = ty * MT + ly, gx = tx * MT + lx;
    if (gy >= D.rows || gx >= D.cols) return;
    const int32_t p = node_of(A, D, gy, gx);
    bool ok = true;
    // forward neighbours: (0,+1), (+1,0), and for 8-conn (+1,-1), (+1,+1)
    const int ndir = CONN == 8 ? 4 : 2;
    for (int d = 0; d < ndir; ++d) {
        const int dy = d == 0 ? 0 : 1;
        const int dx = d == 0 ? 1 : (d == 1 ? 0 : (d == 2 ? -1 : 1));
        const int qy = gy + dy, qx = gx + dx;
        if (qy >= D.rows || qx < 0 || qx >= D.cols) continue;
        if (qy / MT == gy / MT && qx / MT == gx / MT) continue;      // same tile: done in the tile kernel
        ok &= g_connect(A, p, node_of(A, D, qy, qx));
    }
    if (!ok) atomicOr(&A.ctl[1], (uint32_t)MTE_MERGE);
}

__global__ __launch_bounds__(256) void mt_walk_kernel(MtArgs A) {
    const uint32_t n = min(A.ctl[0], A.walk_cap);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        int32_t a = A.walk[k];
        const uint32_t m = A.gM[a];
        bool done = false;
        for (uint32_t it = 0; it < LIMIT_G; ++it) {
            const int32_t p = A.gpar[a];
            if (p == a) { done = true; break; }
            const int32_t q = g_levroot_ro(A.gpar, A.gG, p);
            if (atomicMax(&A.gM[q], m) >= m) { done = true; break; }
            a = q;
        }
        if (!done) atomicOr(&A.ctl[1], (uint32_t)MTE_WALK);
    }
}

__global__ __launch_bounds__(256) void mt_rnode_kernel(MtArgs A) {
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int T = (int)(gtid / MN), n = (int)(gtid - (int64_t)T * MN);
    if (T >= A.ntiles || n >= A.tcount[T]) return;
    const int32_t id = T * MN + n;
    int32_t a = g_levroot_ro(A.gpar, A.gG, id);
    uint32_t best = 0;
    bool done = false;
    for (uint32_t it = 0; it < LIMIT_G; ++it) {
        const uint32_t m = A.gM[a], lv = A.gG[a];
        if (m >= lv) { best = max(best, lv); done = true; break; }
        best = max(best, m);
        const int32_t p = A.gpar[a];
        if (p == a) { done = true; break; }
        a = g_levroot_ro(A.gpar, A.gG, p);
    }
    if (!done) atomicOr(&A.ctl[1], (uint32_t)MTE_RNODE);
    A.gR[id] = (uint8_t)best;
}

template <int MODE>
__global__ __launch_bounds__(256) void mt_out_kernel(MtArgs A) {
    // 256 threads: 4 rows of 64 per pass, 16 passes per tile
    const int T = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    MtImg D;
    int tx, ty;
    tile_of(A, T, D, tx, ty);
    const int lane = threadIdx.x & 63, r0 = threadIdx.x >> 6;
    const int gx = tx * MT + lane;
    if (gx >= D.cols) return;
    const int64_t base = (int64_t)T * MN;
#pragma unroll 4
    for (int ly = r0; ly < MT; ly += 4) {
        const int gy = ty * MT + ly;
        if (gy >= D.rows) break;
        const uint32_t r = A.gR[base + A.lnode[base + ly * MT + lane]];
        D.out[(int64_t)gy * D.po + gx] = (uint8_t)(MODE == MT_FILL ? 255u - r : r);
    }
}

__global__ void mt_clear_kernel(uint32_t* ctl) {
    if (threadIdx.x < 4) ctl[threadIdx.x] = 0u;
}

MtImg make_img(const View* marker, const View& mask, const View& out, int tile_base) {
    MtImg D{};
    D.mask = (const uint8_t*)mask.data;
    D.pk = mask.pitch;
    D.marker = marker ? (const uint8_t*)marker->data : nullptr;
    D.pm = marker ? marker->pitch : 0;
    D.out = (uint8_t*)out.data;
    D.po = out.pitch;
    D.rows = mask.rows;
    D.cols = mask.cols;
    D.ntx = (mask.cols + MT - 1) / MT;
    D.nty = (mask.rows + MT - 1) / MT;
    D.tile_base = tile_base;
    return D;
}

}  // namespace

int mtree_tiles(int rows, int cols) {
    if (rows <= 0 || cols <= 0) return 0;
    return ((rows + MT - 1) / MT) * ((cols + MT - 1) / MT);
}

nsx_status mtree_launch(nsx_ctx* ctx, int mode, const View* marker, const View& mask, const View& out, int conn,
                        hipStream_t s, uint32_t** err_dev, int ws_base) {
    *err_dev = nullptr;
    if (mask.rows <= 0 || mask.cols <= 0) return NSX_OK;
    if (mask.dtype != NSX_U8) return ctx->fail(NSX_ERR_UNSUPPORTED, "component-tree reconstruction is u8 only");
    const int ntiles = mtree_tiles(mask.rows, mask.cols);
    if ((int64_t)ntiles * MN >= (int64_t)INT32_MAX)
        return ctx->fail(NSX_ERR_INVALID, "image too large for the component-tree reconstruction");
    const size_t nn = (size_t)ntiles * MN;
    MtArgs A{};
    A.img0 = make_img(marker, mask, out, 0);
    A.nimg = 1;
    A.ntiles = ntiles;
    void *p_par, *p_m, *p_g, *p_r, *p_ln, *p_tc, *p_walk, *p_ctl;
    const uint32_t walk_cap = (uint32_t)std::min<size_t>((size_t)ntiles * MB_SLOTS * 4 + 1024, (size_t)1 << 30);
    NSX_TRY(ctx->ws(ws_base + 0, sizeof(int32_t) * nn, &p_par));
    NSX_TRY(ctx->ws(ws_base + 1, sizeof(uint32_t) * nn, &p_m));
    NSX_TRY(ctx->ws(ws_base + 2, nn, &p_g));
    NSX_TRY(ctx->ws(ws_base + 3, nn, &p_r));
    NSX_TRY(ctx->ws(ws_base + 4, sizeof(uint16_t) * nn, &p_ln));
    NSX_TRY(ctx->ws(ws_base + 5, 2 * sizeof(int32_t) * (size_t)ntiles, &p_tc));
    NSX_TRY(ctx->ws(ws_base + 6, sizeof(int32_t) * (size_t)walk_cap, &p_walk));
    NSX_TRY(ctx->ws(ws_base + 7, 256, &p_ctl));
    A.gpar = (int32_t*)p_par;
    A.gM = (uint32_t*)p_m;
    A.gG = (uint8_t*)p_g;
    A.gR = (uint8_t*)p_r;
    A.lnode = (uint16_t*)p_ln;
    A.tcount = (int32_t*)p_tc;
    A.troot = (int32_t*)p_tc + ntiles;
    A.walk = (int32_t*)p_walk;
    A.ctl = (uint32_t*)p_ctl;
    A.walk_cap = walk_cap;
    const bool fill = mode == RM_FILL;
    const bool c8 = conn == 8;
    hipLaunchKernelGGL(mt_clear_kernel, dim3(1), dim3(64), 0, s, A.ctl);
    const int h0 = ctx->tic(TK_RECON_FIRST, s);
    if (fill) {
        if (c8) hipLaunchKernelGGL((mt_tile_kernel<MT_FILL, 8>), dim3(ntiles), dim3(MTHREADS), 0, s, A);
        else hipLaunchKernelGGL((mt_tile_kernel<MT_FILL, 4>), dim3(ntiles), dim3(MTHREADS), 0, s, A);
    } else {
        if (c8) hipLaunchKernelGGL((mt_tile_kernel<MT_RECON, 8>), dim3(ntiles), dim3(MTHREADS), 0, s, A);
        else hipLaunchKernelGGL((mt_tile_kernel<MT_RECON, 4>), dim3(ntiles), dim3(MTHREADS), 0, s, A);
    }
    ctx->toc(h0, s);
    const int h1 = ctx->tic(TK_RECON_ROUND, s);
    hipLaunchKernelGGL(mt_roots_kernel, dim3((unsigned)((2 * (int64_t)ntiles + 255) / 256)), dim3(256), 0, s, A);
    const int64_t nmb = (int64_t)ntiles * MB_SLOTS;
    const int gmb = (int)((nmb + 255) / 256);
    if (c8) hipLaunchKernelGGL((mt_merge_kernel<8>), dim3(gmb), dim3(256), 0, s, A);
    else hipLaunchKernelGGL((mt_merge_kernel<4>), dim3(gmb), dim3(256), 0, s, A);
    hipLaunchKernelGGL(mt_walk_kernel, dim3(std::max(1, std::min(4096, (int)((walk_cap + 255) / 256)))), dim3(256), 0, s,
                       A);
    hipLaunchKernelGGL(mt_rnode_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, A);
    if (fill) hipLaunchKernelGGL(mt_out_kernel<MT_FILL>, dim3(ntiles), dim3(256), 0, s, A);
    else hipLaunchKernelGGL(mt_out_kernel<MT_RECON>, dim3(ntiles), dim3(256), 0, s, A);
    ctx->toc(h1, s);
    NSX_HIP_TRY(ctx, hipGetLastError());
    *err_dev = A.ctl + 1;
    return NSX_OK;
}

}  // namespace nsx
