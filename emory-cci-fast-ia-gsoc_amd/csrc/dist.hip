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
// The rule has many fixed points: a pixel can be left with a label that no
// neighbour can improve although a nearer background exists (the "stuck"
// pixels of propagation distance transforms), and where several backgrounds
// are equally far, which one a pixel keeps depends on the order of the offers.
// Two orders are built here.
//
// Queue order (NSX_DT_ORDER_QUEUE, the default) reproduces the reference's
// FIFO order generation by generation (see the section below) and so its
// labels and distances, bit for bit.
//
// Tile order (NSX_DT_ORDER_TILES) runs the same rule to a fixed point in 64x64
// tiles.  A tile loads its labels and a one-pixel halo into LDS, applies the
// rule in synchronous passes (read one LDS buffer, write the other) until no
// label in it changes, and writes the tile to its other global buffer; rounds
// of tile launches repeat until a round changes nothing.  A tile runs in a
// round only when it or one of its 8 neighbour tiles changed in the previous
// round.  Each tile's labels live in one of two global buffers, named by a
// per-tile index (one byte per tile and round parity): a tile that changes
// writes the buffer its index does not name and flips the index for the next
// round; the others keep theirs, so nothing is copied.  Every read is of the
// round-start state (the neighbours' indices of this round), so the result is
// deterministic.  Rounds are launched in batches with one host check of the
// last round's change word.  At stuck pixels and ties its result can differ
// from the reference's: on the 4096x4096 disks image (3000 disks) 1397 of
// 16.7 M distances differ; the reference's order is off the exact Euclidean
// distance at 1812 pixels and this one at 577, never below it.
//
// Labels are packed (y << 16) | x inside the kernels, with DT_FAR for the
// starting point, so no integer division is needed per offer; the output turns
// them into the reference's raster index (far = rows*cols*3) or its float
// distance (:254-266).
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

// The reference's test: float(sqrt(double(a))) < float(sqrt(double(b))).  For
// a < b with a <= 2^20 the rounded roots stay apart: the roots differ by more
// than 1/2048 and a float ulp up to 1024 is at most 2^-13 (and
// float(sqrt(2^20 + 1)) > 1024), so the integer test decides.  Above, ties
// from rounding are possible and the roots are taken.  (The far starting
// point's squared distance is above 2^20, so the common test of an offer
// against it stays integer.)
__device__ __forceinline__ bool dt_closer(uint64_t a, uint64_t b) {
    if (a >= b) return false;
    if (a <= (1ull << 20)) return true;
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

// Four pixels per thread (x .. x+3, one tile since DT is a multiple of 4); with
// `vec` (rows of labels and output 16-byte aligned) one 16-byte load and store.
__global__ __launch_bounds__(256) void dt_out_kernel(const uint32_t* buf0, const uint32_t* buf1, const uint8_t* idx,
                                                     int tiles_x, char* out, int64_t po, int rows, int cols, int far_y,
                                                     int calc_dist, int vec) {
    const int x = 4 * (blockIdx.x * blockDim.x + threadIdx.x), y = blockIdx.y;
    if (x >= cols) return;
    const uint32_t* lab = (idx && idx[(y / DT) * tiles_x + x / DT]) ? buf1 : buf0;
    const uint32_t* lrow = lab + (int64_t)y * cols;
    char* row = out + (int64_t)y * po;
    auto value = [&](int xx, uint32_t l) -> uint32_t {
        if (calc_dist) return __float_as_uint((float)sqrt((double)dt_d2(xx, y, l, far_y)));
        return l == DT_FAR ? (uint32_t)(int32_t)((int64_t)rows * cols * 3)
                           : (l >> 16) * (uint32_t)cols + (l & 0xFFFFu);
    };
    if (vec && x + 4 <= cols) {
        const uint4 l = *reinterpret_cast<const uint4*>(lrow + x);
        *reinterpret_cast<uint4*>(row + 4 * (int64_t)x) =
            make_uint4(value(x, l.x), value(x + 1, l.y), value(x + 2, l.z), value(x + 3, l.w));
    } else {
        for (int k = 0; k < 4 && x + k < cols; ++k) ((uint32_t*)row)[x + k] = value(x + k, lrow[x + k]);
    }
}
inline dim3 dt_out_grid(int rows, int cols) { return dim3((unsigned)((cols + 1023) / 1024), (unsigned)rows); }
inline int dt_out_vec(const void* lab, int cols, const View& out) {
    return ((((uintptr_t)lab | (uintptr_t)cols * 4u | (uintptr_t)out.data | (uintptr_t)out.pitch) & 15u) == 0u) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Queue order (the default).
//
// The reference pops its FIFO queue in generations: generation 0 is the
// initial front in raster order (:205-222), generation g+1 what generation g
// pushed, in push order: entry by entry, each entry's neighbours in the order
// up-left, up-right, up, down-left, down-right, down, left, right (:122-193).
// Within a generation the sequential run is fixed by two facts:
//  * a pixel's label after a sequence of offers is the earliest offer with the
//    smallest float distance (every offer taken is strictly closer than the
//    label it replaces), the generation-start label counting as the first;
//  * the label entry k offers, c_k, is its pixel's label when it is popped:
//    by the first fact, from the pixel's generation-start label and the labels
//    c_j that earlier entries j < k of the generation at its 8 neighbours
//    offered it.
// So the c_k solve a triangular system (c_k depends on c_j, j < k), whose one
// solution is reached by relaxation passes: a pass in which nothing changes
// has found it.  Entry k pushes neighbour n when c_k is strictly closer to n
// than n's generation-start label and than every offer of an entry j < k at a
// neighbour of n; after the generation n holds the offer of the last entry that
// pushed it.  Each generation is then: link (the entries at each pixel, in a
// tagged list), solve passes, push masks, an exclusive scan of their counts,
// and the scatter of the next generation in push order.

constexpr int QCH = 1024;   // entries per solve chunk, one per thread of a block
constexpr int QPC = 32;     // entries per push chunk (push counts, scan): 8 lanes per entry
constexpr int QNB = 512;    // blocks of the solve kernel, looping over chunks
constexpr int QNP = 2048;   // blocks of the push and scatter kernels, looping over chunks
constexpr int QMAXSOLVE = 16;
// control words
// QC_NEED: entries a generation that did not fit needs
enum { QC_LEN0 = 0, QC_LEN1 = 1, QC_STALL = 2, QC_ERR = 3, QC_NEED = 4, QC_CHG = 8 };   // QC_CHG + solve pass

struct Dtq {
    uint32_t* S;       // label of every pixel
    // The entries of a generation at one pixel: head holds (tag << 32) | the
    // first to arrive; head2 (tag << 32) | the latest of the others, with bit
    // 31 set when there is a third, the ones before it linked through nxt.
    uint64_t* head;
    uint64_t* head2;
    uint64_t* win;     // per pixel: (tag << 32) | last entry that pushed it
    uint32_t* E[2];    // entry positions (y << 16 | x), by generation parity
    uint32_t* c[2];    // entry labels, by generation parity
    uint64_t* nxt;     // per later entry: the head2 it replaced
    uint8_t* mask;     // per entry: the neighbours it pushes (bit o = offer o)
    uint8_t* ext;      // per entry: 1 when an offer to it comes from an earlier chunk
    uint32_t* part;    // per push chunk (or front chunk): count, then exclusive offset
    int32_t* ctl;
    int rows, cols, far_y;
    uint32_t tag0;     // generation g is tagged tag0 + g
    uint32_t cap;      // entries a generation buffer holds
};

// the reference's offer order: up-left, up-right, up, down-left, down-right,
// down, left, right
constexpr int kQdx[8] = {-1, 1, 0, -1, 1, 0, -1, 1};
constexpr int kQdy[8] = {-1, -1, -1, 1, 1, 1, 0, 0};

__device__ __forceinline__ int64_t qlin(const Dtq& Q, int x, int y) { return (int64_t)y * Q.cols + x; }
__device__ __forceinline__ uint32_t qpack(int x, int y) { return ((uint32_t)y << 16) | (uint32_t)x; }
__device__ __forceinline__ bool qin(const Dtq& Q, int x, int y) {
    return x >= 0 && y >= 0 && x < Q.cols && y < Q.rows;
}

// Offer j (label v, squared distance q) against the best so far (label best,
// squared distance qb, entry jb; jb = -1 for the generation-start label, which
// comes first): the smaller float distance wins, a tie goes to the earlier.
__device__ __forceinline__ void dtq_take(uint32_t v, uint64_t q, int j, uint32_t& best, uint64_t& qb, int& jb) {
    if (dt_closer(q, qb) || (jb >= 0 && j < jb && !dt_closer(qb, q))) {
        best = v;
        qb = q;
        jb = j;
    }
}

constexpr uint32_t QMORE = 0x80000000u;   // head2 flag: a third entry exists

// Clamped position (always in the image), so that loads need no branch.
__device__ __forceinline__ int64_t qclamp(const Dtq& Q, int x, int y) {
    return qlin(Q, min(max(x, 0), Q.cols - 1), min(max(y, 0), Q.rows - 1));
}

// The entries of this generation at the 8 neighbours of (x, y): the first
// (j, -1 for none), the latest later one (j2) and whether more exist (more
// bit o), all loads independent.
__device__ __forceinline__ void dtq_near(const Dtq& Q, uint32_t tag, int x, int y, int (&j)[8], int (&j2)[8],
                                         uint32_t& more) {
    uint64_t h[8], h2[8];
#pragma unroll
    for (int o = 0; o < 8; ++o) {
        const int64_t li = qclamp(Q, x + kQdx[o], y + kQdy[o]);
        h[o] = Q.head[li];
        h2[o] = Q.head2[li];
    }
    more = 0;
#pragma unroll
    for (int o = 0; o < 8; ++o) {
        const bool in = qin(Q, x + kQdx[o], y + kQdy[o]);
        j[o] = (in && (uint32_t)(h[o] >> 32) == tag) ? (int)(uint32_t)h[o] : -1;
        const bool in2 = in && (uint32_t)(h2[o] >> 32) == tag;
        j2[o] = in2 ? (int)((uint32_t)h2[o] & ~QMORE) : -1;
        more |= (in2 && ((uint32_t)h2[o] & QMORE)) ? 1u << o : 0u;
    }
}

// Calls f(j) for the entries of a list past its latest later one j2.
template <typename F>
__device__ __forceinline__ void dtq_rest(const Dtq& Q, uint32_t tag, int j2, F f) {
    uint64_t h = Q.nxt[j2];
    while ((uint32_t)(h >> 32) == tag) {
        const int j = (int)((uint32_t)h & ~QMORE);
        f(j);
        h = Q.nxt[j];
    }
}

// c_k: the best of the generation-start label s0 and the offers of entries
// j < k at the 8 neighbours (val1(o, j) / val2(o, j): the label of a first /
// latest later entry of list o, val(j): any entry's).
template <typename V1, typename V2, typename Val>
__device__ __forceinline__ uint32_t dtq_popped_near(const Dtq& Q, uint32_t tag, int x, int y, int k, uint32_t s0,
                                                    const int (&j)[8], const int (&j2)[8], uint32_t more, V1 val1,
                                                    V2 val2, Val val) {
    uint32_t best = s0;
    uint64_t qb = dt_d2(x, y, best, Q.far_y);
    if (qb == 0) return best;
    int jb = -1;
#pragma unroll
    for (int o = 0; o < 8; ++o) {
        if (j[o] >= 0 && j[o] < k) {
            const uint32_t v = val1(o, j[o]);
            dtq_take(v, dt_d2(x, y, v, Q.far_y), j[o], best, qb, jb);
        }
        if (j2[o] >= 0 && j2[o] < k) {
            const uint32_t v = val2(o, j2[o]);
            dtq_take(v, dt_d2(x, y, v, Q.far_y), j2[o], best, qb, jb);
        }
    }
    if (more)
#pragma unroll
        for (int o = 0; o < 8; ++o)
            if ((more >> o) & 1u)
                dtq_rest(Q, tag, j2[o], [&](int jj) {
                    if (jj < k) {
                        const uint32_t v = val(jj);
                        dtq_take(v, dt_d2(x, y, v, Q.far_y), jj, best, qb, jb);
                    }
                });
    return best;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Mask bytes x-1 .. x+16 of row y (0 outside the image), for the front test of
// a group of 16 pixels.
struct DtqRow {
    uint32_t w[4];
    uint32_t l, r;
    __device__ __forceinline__ uint32_t at(int i) const {
        return i < 0 ? l : i >= 16 ? r : (w[i >> 2] >> (8 * (i & 3))) & 0xFFu;
    }
};
__device__ __forceinline__ DtqRow dtq_row(const char* mask, int64_t pm, int rows, int cols, int x, int y, bool vec) {
    DtqRow R;
    R.w[0] = R.w[1] = R.w[2] = R.w[3] = R.l = R.r = 0u;
    if (y < 0 || y >= rows || x >= cols) return R;
    const uint8_t* row = (const uint8_t*)(mask + (int64_t)y * pm);
    if (vec && x + 16 <= cols) {
        const uint4 v = *reinterpret_cast<const uint4*>(row + x);
        R.w[0] = v.x;
        R.w[1] = v.y;
        R.w[2] = v.z;
        R.w[3] = v.w;
    } else {
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if (x + k < cols) R.w[k >> 2] |= (uint32_t)row[x + k] << (8 * (k & 3));
    }
    R.l = x > 0 ? row[x - 1] : 0u;
    R.r = x + 16 < cols ? row[x + 16] : 0u;
    return R;
}
// checkDistNeighbors8 (:33-88) for pixels x .. x+15 of row y: bit k set for a
// background pixel with a nonzero 8-neighbour in the image.
__device__ __forceinline__ uint32_t dtq_front16(const char* mask, int64_t pm, int rows, int cols, int x, int y,
                                                bool vec) {
    const DtqRow up = dtq_row(mask, pm, rows, cols, x, y - 1, vec), mid = dtq_row(mask, pm, rows, cols, x, y, vec),
                 dn = dtq_row(mask, pm, rows, cols, x, y + 1, vec);
    uint32_t fm = 0u;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint32_t nb = up.at(k - 1) | up.at(k) | up.at(k + 1) | mid.at(k - 1) | mid.at(k + 1) | dn.at(k - 1) |
                            dn.at(k) | dn.at(k + 1);
        if (x + k < cols && mid.at(k) == 0u && nb != 0u) fm |= 1u << k;
    }
    return fm;
}

// Generation 0, count pass: every label initialised (a background pixel holds
// itself, the others the far point) and per chunk (1024 columns of a row,
// chunks in raster order) the number of front pixels.  A wave takes a chunk,
// 16 pixels per lane; a block takes 4 consecutive chunks.
__global__ __launch_bounds__(256) void dtq_front_count_kernel(Dtq Q, const char* mask, int64_t pm, int nseg,
                                                              int nfront, int vec) {
    const int lane = threadIdx.x & 63;
    const int ci = blockIdx.x * 4 + ((int)threadIdx.x >> 6);
    if (ci >= nfront) return;
    const int y = ci / nseg, x0 = (ci - y * nseg) * QCH;
    const uint32_t fm = dtq_front16(mask, pm, Q.rows, Q.cols, x0 + 16 * lane, y, vec != 0);
    const uint32_t t = wave_sum((uint32_t)__builtin_popcount(fm));
    if (lane == 0) Q.part[ci] = t;
    // labels, lane-contiguous
    const uint8_t* row = (const uint8_t*)(mask + (int64_t)y * pm);
#pragma unroll 4
    for (int k = 0; k < QCH / 64; ++k) {
        const int x = x0 + 64 * k + lane;
        if (x < Q.cols) Q.S[qlin(Q, x, y)] = row[x] == 0 ? qpack(x, y) : DT_FAR;
    }
}

// Generation 0, scatter pass: the front pixels in raster order.
__global__ __launch_bounds__(256) void dtq_front_scatter_kernel(Dtq Q, const char* mask, int64_t pm, int nseg,
                                                                int nfront, int vec) {
    if (Q.ctl[QC_STALL]) return;
    const int lane = threadIdx.x & 63;
    const int ci = blockIdx.x * 4 + ((int)threadIdx.x >> 6);
    if (ci >= nfront) return;
    const int y = ci / nseg, x = (ci - y * nseg) * QCH + 16 * lane;
    uint32_t fm = dtq_front16(mask, pm, Q.rows, Q.cols, x, y, vec != 0);
    const uint32_t cnt = (uint32_t)__builtin_popcount(fm);
    uint32_t pos = Q.part[ci] + wave_incl_scan(cnt) - cnt;
    while (fm) {
        const int k = __builtin_ctz(fm);
        fm &= fm - 1u;
        Q.E[0][pos++] = qpack(x + k, y);
    }
}

// Exclusive prefix of part[0..n) in place, one block of 1024 threads; n is the
// chunk count of the generation length at ctl[len_slot] (or nfix when
// len_slot < 0).  The total, the next generation's length, goes to
// ctl[out_slot]; above the buffers' capacity it sets the error and stall words.
__global__ __launch_bounds__(1024) void dtq_scan_kernel(Dtq Q, int len_slot, int nfix, int out_slot) {
    if (Q.ctl[QC_STALL]) return;
    const int n = len_slot >= 0 ? (Q.ctl[len_slot] + QPC - 1) / QPC : nfix;
    const int per = (n + 1023) / 1024;
    const int b = (int)threadIdx.x * per, e = min(b + per, n);
    uint64_t s = 0;
    for (int i = b; i < e; ++i) s += Q.part[i];
    // block scan of the per-thread sums (16 waves)
    __shared__ uint64_t wt[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t inc = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t t = __shfl_up(inc, d, 64);
        if (lane >= d) inc += t;
    }
    if (lane == 63) wt[w] = inc;
    __syncthreads();
    uint64_t before = 0;
    for (int q = 0; q < w; ++q) before += wt[q];
    uint64_t run = before + inc - s;
    for (int i = b; i < e; ++i) {
        const uint32_t v = Q.part[i];
        Q.part[i] = (uint32_t)run;
        run += v;
    }
    if (threadIdx.x == 1023) {
        if (run > Q.cap) {
            Q.ctl[QC_NEED] = (int32_t)min(run, (uint64_t)INT32_MAX);
            Q.ctl[QC_ERR] = 1;
            Q.ctl[QC_STALL] = -1;
        } else {
            Q.ctl[out_slot] = (int32_t)run;
        }
    }
}

// Generation g's entries: each takes its pixel's label (for g > 0 the offer of
// the last entry of generation g - 1 that pushed it, which becomes the pixel's
// label) and is linked into its pixel's list.
__global__ __launch_bounds__(256) void dtq_link_kernel(Dtq Q, int g) {
    if (Q.ctl[QC_STALL]) return;
    if (blockIdx.x == 0 && threadIdx.x < QMAXSOLVE) Q.ctl[QC_CHG + threadIdx.x] = 0;
    const int L = Q.ctl[g & 1];
    const uint32_t tag = Q.tag0 + (uint32_t)g;
    const uint32_t* E = Q.E[g & 1];
    uint32_t* cv = Q.c[g & 1];
    const uint32_t* cprev = Q.c[(g & 1) ^ 1];
    for (int k = blockIdx.x * 256 + threadIdx.x; k < L; k += gridDim.x * 256) {
        const uint32_t p = E[k];
        const int64_t li = qlin(Q, (int)(p & 0xFFFFu), (int)(p >> 16));
        uint32_t v = p;   // a front pixel is background: it holds itself
        if (g > 0) {
            const uint64_t w = Q.win[li];
            if ((uint32_t)(w >> 32) != tag - 1u) {
                Q.ctl[QC_ERR] = 2;   // not pushed by the previous generation: an internal error
                Q.ctl[QC_STALL] = -1;
                return;
            }
            v = cprev[(uint32_t)w];
            Q.S[li] = v;
        }
        cv[k] = v;
        Q.ext[k] = 1;   // checked by the push kernel unless a solve pass clears it
        // the first entry at the pixel takes head; the others go to head2
        const unsigned long long mine = ((unsigned long long)tag << 32) | (uint32_t)k;
        unsigned long long cur = Q.head[li];
        bool first = false;
        while ((uint32_t)(cur >> 32) != tag) {   // at most twice: old tags only turn into this one
            const unsigned long long got = atomicCAS((unsigned long long*)&Q.head[li], cur, mine);
            if (got == cur) {
                first = true;
                break;
            }
            cur = got;
        }
        if (!first) {
            unsigned long long old = Q.head2[li];
            for (;;) {
                const unsigned long long want = mine | ((uint32_t)(old >> 32) == tag ? QMORE : 0u);
                const unsigned long long got = atomicCAS((unsigned long long*)&Q.head2[li], old, want);
                if (got == old) break;
                old = got;
            }
            Q.nxt[k] = old;
        }
    }
}

// One relaxation pass of generation g's labels, in place.  A chunk's labels
// are relaxed in LDS until they stop changing (offers from earlier chunks read
// as they are in global memory).  Pass 0 relaxes every entry and records which
// entries take offers from earlier chunks; a later pass (run only if the pass
// before it changed something) recomputes those first and relaxes the whole
// chunk again only if one of them changed.
__global__ __launch_bounds__(QCH) void dtq_solve_kernel(Dtq Q, int g, int pass) {
    if (Q.ctl[QC_STALL]) return;
    if (pass > 0 && Q.ctl[QC_CHG + pass - 1] == 0) return;
    const int L = Q.ctl[g & 1];
    const uint32_t tag = Q.tag0 + (uint32_t)g;
    const uint32_t* E = Q.E[g & 1];
    uint32_t* cv = Q.c[g & 1];
    __shared__ uint32_t lc[QCH];
    for (int ci = blockIdx.x; ci * QCH < L; ci += gridDim.x) {
        const int base = ci * QCH;
        const int k = base + (int)threadIdx.x;
        const bool live = k < L;
        int x = 0, y = 0, jn[8], jn2[8];
        uint32_t more = 0;
        uint32_t ev[8], ev2[8];   // labels of entries in earlier chunks, read once per pass
        bool ext = false;
        uint32_t s0 = 0;
        bool load = live && (pass == 0 || Q.ext[k]);
        auto fetch = [&]() {
            const uint32_t p = E[k];
            x = (int)(p & 0xFFFFu);
            y = (int)(p >> 16);
            s0 = Q.S[qlin(Q, x, y)];
            dtq_near(Q, tag, x, y, jn, jn2, more);
#pragma unroll
            for (int o = 0; o < 8; ++o) {
                ev[o] = cv[(jn[o] >= 0 && jn[o] < base) ? jn[o] : k];
                ev2[o] = cv[(jn2[o] >= 0 && jn2[o] < base) ? jn2[o] : k];
            }
        };
        if (load) fetch();
        lc[threadIdx.x] = live ? cv[k] : 0u;
        __syncthreads();
        auto val = [&](int j) { return j >= base ? lc[j - base] : cv[j]; };
        auto val1 = [&](int o, int j) { return j >= base ? lc[j - base] : ev[o]; };
        auto val2 = [&](int o, int j) { return j >= base ? lc[j - base] : ev2[o]; };
        bool any = false, settled = false;
        for (int sweep = 0; sweep < 64; ++sweep) {
            if (sweep == 1 && live && !load) {
                load = true;
                fetch();
            }
            bool ch = false;
            if (load) {
                const uint32_t v = dtq_popped_near(Q, tag, x, y, k, s0, jn, jn2, more, val1, val2, val);
                if (v != lc[threadIdx.x]) {
                    lc[threadIdx.x] = v;
                    ch = true;
                }
            }
            if (!__syncthreads_or(ch)) {
                settled = true;
                break;
            }
            any = true;
        }
        if (live && (pass == 0 || !settled)) {
            // offers from earlier chunks (a list with a second entry counts as
            // one); a chunk that did not settle within the sweeps marks all
            // its entries, so that the push kernel checks every one
#pragma unroll
            for (int o = 0; o < 8; ++o) ext |= (jn[o] >= 0 && jn[o] < base) || (jn2[o] >= 0 && jn2[o] < base);
            Q.ext[k] = (uint8_t)(ext || more || !settled);
        }
        if (any) {
            if (live) cv[k] = lc[threadIdx.x];
            if (threadIdx.x == 0) Q.ctl[QC_CHG + pass] = 1;
        }
        __syncthreads();
    }
}

// Generation g's push masks and per-chunk push counts, 8 lanes per entry.
// The offers that decide an entry's pushes come from the entries at its 5x5
// neighbourhood: the 8 lanes load its 25 pixels' first and latest later
// entries and their labels together (3 or 4 pixels each) into LDS.  Then lane o
// decides whether the entry pushes its neighbour o (the reference's offer
// order): when its label is strictly closer to it than the neighbour's
// generation-start label and than every offer of an earlier entry at one of
// the neighbour's own 8 neighbours.  For g > 0 an entry that takes offers
// from earlier chunks first checks that its label solves its equation, lane o
// contributing the offers of the entries at neighbour o, reduced over the 8
// lanes (the other entries were relaxed to it within their chunk).  If one
// does not, the solve passes were too few: the stall word names the
// generation and every later kernel returns at once, so the host can run more
// passes and go on.
__global__ __launch_bounds__(256) void dtq_push_kernel(Dtq Q, int g) {
    if (Q.ctl[QC_STALL]) return;
    const int L = Q.ctl[g & 1];
    const uint32_t tag = Q.tag0 + (uint32_t)g;
    const uint32_t* E = Q.E[g & 1];
    const uint32_t* cv = Q.c[g & 1];
    const int o = threadIdx.x & 7, lane = threadIdx.x & 63, e = threadIdx.x >> 3;
    // offer o's direction: up-left, up-right, up, down-left, down-right, down, left, right
    const int odx = o < 6 ? (o % 3 == 0 ? -1 : o % 3 == 1 ? 1 : 0) : (o == 6 ? -1 : 1);
    const int ody = o < 3 ? -1 : (o < 6 ? 1 : 0);
    // per entry of the block and pixel t = (dy + 2) * 5 + dx + 2 of its 5x5
    // neighbourhood: first entry, latest later entry (-1: none; bit 30 of the
    // later one: more entries follow it), and their labels
    __shared__ int sj1[QPC][25], sj2[QPC][25];
    __shared__ uint32_t sc1[QPC][25], sc2[QPC][25];
    __shared__ uint32_t red[4];
    constexpr int kMore = 1 << 30;
    for (int ci = blockIdx.x; ci * QPC < L; ci += gridDim.x) {
        const int k = ci * QPC + e;
        const bool live = k < L;
        int x = 0, y = 0;
        uint32_t ck = 0;
        if (live) {
            const uint32_t p = E[k];
            x = (int)(p & 0xFFFFu);
            y = (int)(p >> 16);
            ck = cv[k];
            uint64_t h1[4], h2[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int t = o + 8 * i;
                const int64_t li = qclamp(Q, x + t % 5 - 2, y + t / 5 - 2);
                h1[i] = t < 25 ? Q.head[li] : 0ull;
                h2[i] = t < 25 ? Q.head2[li] : 0ull;
            }
            int j1[4], j2[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int t = o + 8 * i;
                const bool in = t < 25 && qin(Q, x + t % 5 - 2, y + t / 5 - 2);
                j1[i] = (in && (uint32_t)(h1[i] >> 32) == tag) ? (int)(uint32_t)h1[i] : -1;
                j2[i] = (in && (uint32_t)(h2[i] >> 32) == tag) ? (int)((uint32_t)h2[i] & ~QMORE) : -1;
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int t = o + 8 * i;
                const uint32_t c1 = cv[j1[i] >= 0 ? j1[i] : k], c2 = cv[j2[i] >= 0 ? j2[i] : k];
                if (t < 25) {
                    sj1[e][t] = j1[i];
                    sj2[e][t] = j2[i] >= 0 && ((uint32_t)h2[i] & QMORE) ? (j2[i] | kMore) : j2[i];
                    sc1[e][t] = c1;
                    sc2[e][t] = c2;
                }
            }
        }
        __syncthreads();
        bool ok = false;
        if (live) {
            const int nx = x + odx, ny = y + ody;
            const uint32_t sn = Q.S[qclamp(Q, nx, ny)];
            const uint64_t qk = dt_d2(nx, ny, ck, Q.far_y);
            ok = qin(Q, nx, ny) && dt_closer(qk, dt_d2(nx, ny, sn, Q.far_y));
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int t = (ody + kQdy[q] + 2) * 5 + (odx + kQdx[q] + 2);
                const int a = sj1[e][t], b2 = sj2[e][t];
                const int b = b2 & ~kMore;
                if (a >= 0 && a < k && !dt_closer(qk, dt_d2(nx, ny, sc1[e][t], Q.far_y))) ok = false;
                if (b2 >= 0 && b < k && !dt_closer(qk, dt_d2(nx, ny, sc2[e][t], Q.far_y))) ok = false;
                if (ok && b2 >= 0 && (b2 & kMore))
                    dtq_rest(Q, tag, b, [&](int jj) {
                        if (jj < k && !dt_closer(qk, dt_d2(nx, ny, cv[jj], Q.far_y))) ok = false;
                    });
            }
            if (g > 0 && Q.ext[k]) {
                // the offers to (x, y) of the entries at neighbour o
                uint64_t qb = ~0ull;
                int jb = INT_MAX;
                uint32_t vb = 0;
                auto offer = [&](int jj, uint32_t v) {
                    if (jj >= k) return;
                    const uint64_t qv = dt_d2(x, y, v, Q.far_y);
                    if (dt_closer(qv, qb) || (jj < jb && !dt_closer(qb, qv))) {
                        qb = qv;
                        jb = jj;
                        vb = v;
                    }
                };
                const int t = (ody + 2) * 5 + (odx + 2);
                const int a = sj1[e][t], b2 = sj2[e][t];
                if (a >= 0) offer(a, sc1[e][t]);
                if (b2 >= 0) {
                    offer(b2 & ~kMore, sc2[e][t]);
                    if (b2 & kMore) dtq_rest(Q, tag, b2 & ~kMore, [&](int jj) { offer(jj, cv[jj]); });
                }
#pragma unroll
                for (int d = 1; d < 8; d <<= 1) {
                    const uint64_t q2 = __shfl_xor(qb, d, 64);
                    const int jo = __shfl_xor(jb, d, 64);
                    const uint32_t v2 = __shfl_xor(vb, d, 64);
                    if (dt_closer(q2, qb) || (jo < jb && !dt_closer(qb, q2))) {
                        qb = q2;
                        jb = jo;
                        vb = v2;
                    }
                }
                if (o == 0) {
                    // the generation-start label comes before every offer
                    const uint32_t s0 = Q.S[qlin(Q, x, y)];
                    const uint64_t q0 = dt_d2(x, y, s0, Q.far_y);
                    const uint32_t want = (q0 != 0 && jb != INT_MAX && dt_closer(qb, q0)) ? vb : s0;
                    if (want != ck) Q.ctl[QC_STALL] = g + 1;
                }
            }
        }
        const uint64_t bal = __ballot(ok);
        const uint32_t m = (uint32_t)(bal >> (lane & ~7)) & 0xFFu;
        if (o == 0 && live) Q.mask[k] = (uint8_t)m;
        uint32_t c = o == 0 ? (uint32_t)__builtin_popcount(m) : 0u;
        c = wave_sum(c);
        if (lane == 0) red[threadIdx.x >> 6] = c;
        __syncthreads();
        if (threadIdx.x == 0) Q.part[ci] = red[0] + red[1] + red[2] + red[3];
    }
}

// Generation g + 1 in push order, and each pushed pixel's last pusher.  A
// block takes 256 entries (8 push chunks): a wave's two halves are two chunks.
// The chunks' offsets come from their push counts directly: each block sums
// the counts of the chunks before its own and of all chunks (the latter is the
// next generation's length; above the buffers' capacity every block returns
// before writing and the error and stall words are set).
__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}
__global__ __launch_bounds__(256) void dtq_scatter_kernel(Dtq Q, int g) {
    if (Q.ctl[QC_STALL]) return;
    const int L = Q.ctl[g & 1];
    if (L == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) Q.ctl[(g & 1) ^ 1] = 0;
        return;
    }
    const int n = (L + QPC - 1) / QPC;
    const uint64_t tagw = (uint64_t)(Q.tag0 + (uint32_t)g) << 32;
    const uint32_t* E = Q.E[g & 1];
    uint32_t* En = Q.E[(g & 1) ^ 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __shared__ uint64_t red[2][4];
    __shared__ uint32_t coff[8];
    for (int b = blockIdx.x; b * 256 < L; b += gridDim.x) {
        uint64_t sb = 0, st = 0;
        for (int i = (int)threadIdx.x; i < n; i += 256) {
            const uint32_t v = Q.part[i];
            st += v;
            if (i < 8 * b) sb += v;
        }
        sb = wave_sum64(sb);
        st = wave_sum64(st);
        if (lane == 0) {
            red[0][w] = sb;
            red[1][w] = st;
        }
        __syncthreads();
        const uint64_t before = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        const uint64_t total = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        if (total > Q.cap) {
            if (threadIdx.x == 0) {
                Q.ctl[QC_NEED] = (int32_t)min(total, (uint64_t)INT32_MAX);
                Q.ctl[QC_ERR] = 1;
                Q.ctl[QC_STALL] = -1;
            }
            return;
        }
        if (b == 0 && threadIdx.x == 0) Q.ctl[(g & 1) ^ 1] = (int32_t)total;
        if (threadIdx.x < 8) {
            uint32_t o = (uint32_t)before;
            for (int j = 0; j < (int)threadIdx.x && 8 * b + j < n; ++j) o += Q.part[8 * b + j];
            coff[threadIdx.x] = o;
        }
        __syncthreads();
        const int k = b * 256 + (int)threadIdx.x;
        const uint32_t m = k < L ? Q.mask[k] : 0u;
        const uint32_t cnt = (uint32_t)__builtin_popcount(m);
        const uint32_t inc = wave_incl_scan(cnt);
        const uint32_t half = __shfl(inc, 31, 64);
        const uint32_t ex = inc - cnt - (lane >= 32 ? half : 0u);
        if (m) {
            const uint32_t p = E[k];
            const int x = (int)(p & 0xFFFFu), y = (int)(p >> 16);
            uint32_t pos = coff[threadIdx.x >> 5] + ex;
            for (int o = 0; o < 8; ++o) {
                if (!((m >> o) & 1u)) continue;
                const int nx = x + kQdx[o], ny = y + kQdy[o];
                En[pos++] = qpack(nx, ny);
                atomicMax((unsigned long long*)&Q.win[qlin(Q, nx, ny)], (unsigned long long)(tagw | (uint32_t)k));
            }
        }
        __syncthreads();   // red and coff are reused by the next 256 entries
    }
}

// Queue order on the host: generation 0, then batches of generations with one
// host check each (the next length, the stall and error words).
nsx_status dist_queue_run(nsx_ctx* ctx, const View& mask, const View& out, bool calc_dist, hipStream_t s) {
    const int rows = mask.rows, cols = mask.cols;
    const int64_t npix = (int64_t)rows * cols;
    const int nseg = (cols + QCH - 1) / QCH;
    const int64_t nfront = (int64_t)rows * nseg;
    // Solve passes launched per generation, and per resumed generation.
    const int kSolve = std::min(std::max(ctx->dt_solve, 0), 8);
    constexpr int kResumeSolve = 8;
    static_assert(kResumeSolve <= QMAXSOLVE, "change words");
    // Entries a generation buffer holds (at least what the workspace already
    // has room for); a longer generation (a dense front) starts the call over
    // with room for it.
    constexpr size_t kEntryBytes = 4 * sizeof(uint32_t) + sizeof(uint64_t) + 2;
    int64_t cap = npix / 4 + 4096;
    if ((int)ctx->bufs.size() > WS_DT_ENT)
        cap = std::min<int64_t>(std::max<int64_t>(cap, (int64_t)(ctx->bufs[WS_DT_ENT].bytes / kEntryBytes)),
                                (int64_t)1 << 30);
    for (;;) {
        const int64_t nchunk = std::max<int64_t>(nfront, (cap + QPC - 1) / QPC);
        void *lab_p, *tags_p, *ent_p, *part_p, *misc_p;
        NSX_TRY(ctx->ws(WS_CCL_LABEL, sizeof(uint32_t) * (size_t)npix, &lab_p));
        NSX_TRY(ctx->ws(WS_DT_TAGS, 3 * sizeof(uint64_t) * (size_t)npix, &tags_p));
        NSX_TRY(ctx->ws(WS_DT_ENT, (size_t)cap * kEntryBytes, &ent_p));
        NSX_TRY(ctx->ws(WS_DT_PART, sizeof(uint32_t) * (size_t)nchunk, &part_p));
        NSX_TRY(ctx->ws(WS_CCL_MISC, sizeof(int32_t) * 64, &misc_p));
        // Tags only ever grow, so list heads and pushers left by earlier calls
        // never match; a fresh allocation (or a tag near wrapping) is cleared.
        const nsx::DevBuf& tb = ctx->bufs[WS_DT_TAGS];
        if (tb.serial != ctx->dtq_tags_serial || ctx->dtq_tag == 0 || ctx->dtq_tag > 0xF0000000u) {
            NSX_HIP_TRY(ctx, hipMemsetAsync(tb.ptr, 0, tb.bytes, s));
            ctx->dtq_tags_serial = tb.serial;
            ctx->dtq_tag = 1;
        }
        Dtq Q;
        Q.S = (uint32_t*)lab_p;
        Q.head = (uint64_t*)tags_p;
        Q.head2 = Q.head + npix;
        Q.win = Q.head2 + npix;
        Q.nxt = (uint64_t*)ent_p;
        Q.E[0] = (uint32_t*)(Q.nxt + cap);
        Q.E[1] = Q.E[0] + cap;
        Q.c[0] = Q.E[1] + cap;
        Q.c[1] = Q.c[0] + cap;
        Q.mask = (uint8_t*)(Q.c[1] + cap);
        Q.ext = Q.mask + cap;
        Q.part = (uint32_t*)part_p;
        Q.ctl = (int32_t*)misc_p;
        Q.rows = rows;
        Q.cols = cols;
        Q.far_y = 3 * rows;
        Q.tag0 = ctx->dtq_tag;
        Q.cap = (uint32_t)cap;
        NSX_HIP_TRY(ctx, hipMemsetAsync(Q.ctl, 0, sizeof(int32_t) * 64, s));
        const unsigned fblocks = (unsigned)((nfront + 3) / 4);
        const int fvec = ((((uintptr_t)mask.data | (uintptr_t)mask.pitch) & 15u) == 0u) ? 1 : 0;
        hipLaunchKernelGGL(dtq_front_count_kernel, dim3(fblocks), dim3(256), 0, s, Q, mask.data, mask.pitch, nseg,
                           (int)nfront, fvec);
        hipLaunchKernelGGL(dtq_scan_kernel, dim3(1), dim3(1024), 0, s, Q, -1, (int)nfront, (int)QC_LEN0);
        hipLaunchKernelGGL(dtq_front_scatter_kernel, dim3(fblocks), dim3(256), 0, s, Q, mask.data, mask.pitch,
                           nseg, (int)nfront, fvec);
        NSX_HIP_TRY(ctx, hipGetLastError());
        auto generation = [&](int g, int solve, bool resumed) {
            if (!resumed) hipLaunchKernelGGL(dtq_link_kernel, dim3(QNB), dim3(256), 0, s, Q, g);
            if (g > 0)
                for (int k = 0; k < solve; ++k)
                    hipLaunchKernelGGL(dtq_solve_kernel, dim3(QNB), dim3(QCH), 0, s, Q, g, k);
            hipLaunchKernelGGL(dtq_push_kernel, dim3(QNP), dim3(256), 0, s, Q, g);
            hipLaunchKernelGGL(dtq_scatter_kernel, dim3(QNP), dim3(256), 0, s, Q, g);
        };
        // Every generation moves some label strictly closer, so the number of
        // generations is bounded; the tag space bounds the loop here.
        int g = 0, batch = 4;
        bool grow = false;
        int64_t need = 0;
        for (;;) {
            for (int b = 0; b < batch; ++b, ++g) generation(g, kSolve, false);
            NSX_HIP_TRY(ctx, hipGetLastError());
            NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned + 40, Q.ctl, sizeof(int32_t) * 5, hipMemcpyDeviceToHost, s));
            NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
            int32_t len = ctx->h_pinned[40 + (g & 1)], stall = ctx->h_pinned[42], err = ctx->h_pinned[43];
            while (stall > 0 && err == 0) {
                // generation stall - 1 needs more solve passes; the kernels
                // after it returned at once, so it resumes where it stopped
                ++ctx->dtq_resumes;
                g = stall - 1;
                NSX_HIP_TRY(ctx, hipMemsetAsync(Q.ctl + QC_STALL, 0, sizeof(int32_t), s));
                NSX_HIP_TRY(ctx, hipMemsetAsync(Q.ctl + QC_CHG, 0, sizeof(int32_t) * QMAXSOLVE, s));
                generation(g, kResumeSolve, true);
                ++g;
                NSX_HIP_TRY(ctx, hipGetLastError());
                NSX_HIP_TRY(ctx,
                            hipMemcpyAsync(ctx->h_pinned + 40, Q.ctl, sizeof(int32_t) * 5, hipMemcpyDeviceToHost, s));
                NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
                len = ctx->h_pinned[40 + (g & 1)];
                stall = ctx->h_pinned[42];
                err = ctx->h_pinned[43];
            }
            if (err == 1) {
                grow = true;
                need = ctx->h_pinned[40 + QC_NEED];
                break;
            }
            if (err != 0) return ctx->fail(NSX_ERR_NOT_CONVERGED, "distanceTransform: queue generations inconsistent");
            if (len == 0) break;
            if ((uint64_t)ctx->dtq_tag + (uint64_t)g + 64 > 0xFFFFFFF0ull)
                return ctx->fail(NSX_ERR_NOT_CONVERGED, "distanceTransform: generation tags exhausted");
            batch = 8;
        }
        ctx->dtq_tag += (uint32_t)g + 1;
        if (grow) {
            // a generation longer than the entry buffers: start over with room
            // for it
            const int64_t want = std::max<int64_t>(2 * cap, need + need / 4);
            if (want > (int64_t)1 << 30)
                return ctx->fail(NSX_ERR_NOMEM, "distanceTransform: queue generation beyond 2^30 entries");
            cap = want;
            continue;
        }
        hipLaunchKernelGGL(dt_out_kernel, dt_out_grid(rows, cols), dim3(256), 0, s, (const uint32_t*)Q.S,
                           (const uint32_t*)Q.S, (const uint8_t*)nullptr, 1, out.data, out.pitch, rows, cols, 3 * rows,
                           calc_dist ? 1 : 0, dt_out_vec(Q.S, cols, out));
        NSX_HIP_TRY(ctx, hipGetLastError());
        return NSX_OK;
    }
}

}  // namespace

nsx_status dist_run(nsx_ctx* ctx, const View& mask, const View& out, bool calc_dist, hipStream_t s) {
    const int rows = mask.rows, cols = mask.cols;
    if (rows == 0 || cols == 0) return NSX_OK;
    if (rows >= 65535 || cols >= 65535 || (int64_t)rows * cols * 3 > INT32_MAX)
        return ctx->fail(NSX_ERR_INVALID, "distanceTransform: rows, cols < 65535 and rows*cols*3 within int32");
    if (ctx->dt_order == NSX_DT_ORDER_QUEUE) return dist_queue_run(ctx, mask, out, calc_dist, s);
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
    hipLaunchKernelGGL(dt_out_kernel, dt_out_grid(rows, cols), dim3(256), 0, s, (const uint32_t*)buf0,
                       (const uint32_t*)buf1, (const uint8_t*)ix[round & 1], tiles_x, out.data, out.pitch, rows, cols,
                       3 * rows, calc_dist ? 1 : 0, dt_out_vec(buf0, cols, out) & dt_out_vec(buf1, cols, out));
    NSX_HIP_TRY(ctx, hipGetLastError());
    return NSX_OK;
}

}  // namespace nsx
