// recon.hip -- morphological reconstruction by dilation on gfx950.
//
// Replaces the reference's CPU hybrid algorithm (MorphologicOperations.cpp:
// 665-838 imreconstruct, 1688-1829 imreconstructBinary, 1859-1926 imfillHoles)
// and its GPU scan + queue variants (imreconstruct_int_kernel.cu:589-712,
// global_queue.cu:746-857).  The result is the unique fixed point
//     R = min(max(R, dilate_conn(R)), mask),   R >= min(marker, mask),
// with a constant frame around the image, so any propagation order that reaches
// the fixed point is bit-exact with the reference's sequential order.
//
// Design (one wave64 per 64x64 register tile):
//   * The image is cut into 62x62 tiles; each wave loads its tile plus a 1-px
//     halo (64x64) into VGPRs, lane = column, register = row, as 32-bit order-
//     preserving keys (u8 zero-extended, i32 sign-flipped, f32 ordered-int).
//   * Pixels outside the image take the frame value with mask == value, so they
//     never change: the frame needs no special case in the sweeps.
//   * Propagation inside the tile runs to the tile's fixed point by repeated
//     sweeps: a downward and an upward sweep (each row takes the clamped max of
//     its own value and the previous row's 1- or 3-neighbourhood, the latter
//     through DPP wave_shr/wave_shl), then the tile is transposed through LDS
//     and the same two sweeps run along the other axis.  v_med3_u32 does the
//     clamp: new = med3(neighbour max, cur, mask) since cur <= mask.
//   * Halo pixels are computed with the tile (their raised values are valid
//     lower bounds) but never written; when one rises, the tile that owns it is
//     activated.  Interior pixels are written back.
//   * One launch visits every tile once (state = min(marker, mask), then the
//     local fixed point).  Activated tiles go to a device-side FIFO of tile ids
//     (atomic append into a power-of-two ring; a per-tile state word keeps a tile
//     in the ring at most once).  A persistent drain kernel, one wave per
//     worker, pops tiles until no tile is queued or being visited, so a whole
//     reconstruction is two launches with no host round trip.  Values only
//     rise, so reading a neighbour mid-update gives a valid lower bound; the
//     activation protocol re-visits any tile whose halo rose after it was read.
//   * Every spin in the queue is bounded; a stuck queue sets an error bit and
//     all waves exit (reported as NSX_ERR_NOT_CONVERGED), so the kernel cannot
//     hang the device.
//   * Several images (the batched form) share one launch: tiles are numbered
//     across images and a tile -> image table selects each tile's descriptor.
#include "nsx_internal.h"

#include <climits>
#include <cfloat>
#include <algorithm>

namespace nsx {
namespace {

constexpr int TS = 64;          // register tile side, halo included
constexpr int TI = TS - 2;      // interior side written back
constexpr int LDS_PITCH = TS + 1;
constexpr int MAX_IT = 64;      // sweep iterations per visit before re-queueing itself

// ---- order-preserving 32-bit keys -------------------------------------------------
template <typename T> struct Key;
template <> struct Key<uint8_t> {
    __device__ static uint32_t to(uint8_t v) { return v; }
    __device__ static uint8_t from(uint32_t k) { return (uint8_t)k; }
};
template <> struct Key<int32_t> {
    __device__ static uint32_t to(int32_t v) { return (uint32_t)v ^ 0x80000000u; }
    __device__ static int32_t from(uint32_t k) { return (int32_t)(k ^ 0x80000000u); }
};
template <> struct Key<float> {
    // Compare-select order of finite floats; -0.0 sorts just below +0.0.
    __device__ static uint32_t to(float v) {
        uint32_t b = __float_as_uint(v);
        return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    }
    __device__ static float from(uint32_t k) {
        uint32_t b = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
        return __uint_as_float(b);
    }
};

// PixelOperations::invert<T> (PixelOperations.cpp:17-39): 255-x, saturating
// two's-complement negation, -x.
__device__ inline uint8_t inv(uint8_t v) { return (uint8_t)(255 - v); }
__device__ inline int32_t inv(int32_t v) { return v == INT_MIN ? INT_MAX : -v; }
__device__ inline float inv(float v) { return -v; }

template <typename T> __device__ inline T tmax();
template <> __device__ inline uint8_t tmax<uint8_t>() { return 255; }
template <> __device__ inline int32_t tmax<int32_t>() { return INT_MAX; }
template <> __device__ inline float tmax<float>() { return FLT_MAX; }
// cci::common::type::min<T> (TypeUtils.h:33-41)
template <typename T> __device__ inline T tmin();
template <> __device__ inline uint8_t tmin<uint8_t>() { return 0; }
template <> __device__ inline int32_t tmin<int32_t>() { return INT_MIN; }
template <> __device__ inline float tmin<float>() { return -FLT_MAX; }

// ---- per-mode codecs ------------------------------------------------------------
// frame(): key of the constant ring around the image.
// mkey(a, m): mask key.   first(a, m): initial key.
// dec(s, a, m): key of the stored state.   enc(k, a): stored value of a key.
template <typename T, int MODE> struct Codec;

// imreconstruct: state = R; frame 0 (MorphologicOperations.cpp:671-674).
template <typename T> struct Codec<T, RM_RECON> {
    static constexpr bool kMarkerFirst = true, kMarkerLater = false;
    __device__ static uint32_t frame() { return Key<T>::to(T(0)); }
    __device__ static uint32_t mkey(T, T m) { return Key<T>::to(m); }
    __device__ static uint32_t first(T a, T m) { return min(Key<T>::to(a), Key<T>::to(m)); }
    __device__ static uint32_t dec(T s, T, T) { return Key<T>::to(s); }
    __device__ static T enc(uint32_t k, T) { return Key<T>::from(k); }
};

// imreconstructBinary: key = reached (0/1) in the set mask != 0, seeded by
// marker != 0 && mask != 0; nonzero marker pixels keep their value
// (MorphologicOperations.cpp:1665-1672,1712-1756).
template <typename T> struct Codec<T, RM_BINARY> {
    static constexpr bool kMarkerFirst = true, kMarkerLater = true;
    __device__ static uint32_t frame() { return 0u; }
    __device__ static uint32_t mkey(T, T m) { return m != T(0) ? 1u : 0u; }
    __device__ static uint32_t first(T a, T m) { return (a != T(0) && m != T(0)) ? 1u : 0u; }
    __device__ static uint32_t dec(T s, T a, T m) {
        return a != T(0) ? (m != T(0) ? 1u : 0u) : (s != T(0) ? 1u : 0u);
    }
    __device__ static T enc(uint32_t k, T a) { return a != T(0) ? a : (k ? tmax<T>() : T(0)); }
};

// imfillHoles, grayscale: R = reconstruction of marker (min<T> inside, max<T>
// on the frame) under invert(pad(image, min<T>)); output invert(R)
// (MorphologicOperations.cpp:1889-1925).  The frame ring holds
// min(max<T>, invert(min<T>)) = max<T>.  State stored = invert(R) = output.
template <typename T> struct Codec<T, RM_FILL> {
    static constexpr bool kMarkerFirst = false, kMarkerLater = false;
    __device__ static uint32_t frame() { return Key<T>::to(tmax<T>()); }
    __device__ static uint32_t mkey(T, T m) { return Key<T>::to(inv(m)); }
    __device__ static uint32_t first(T, T) { return Key<T>::to(tmin<T>()); }
    __device__ static uint32_t dec(T s, T, T) { return Key<T>::to(inv(s)); }
    __device__ static T enc(uint32_t k, T) { return inv(Key<T>::from(k)); }
};

// imfillHoles, binary, u8: reached set of invert(image) != 0 (image != 255)
// from the frame; output invert(reached ? 255 : 0).
template <typename T> struct Codec<T, RM_FILLBIN> {
    static constexpr bool kMarkerFirst = false, kMarkerLater = false;
    __device__ static uint32_t frame() { return 1u; }
    __device__ static uint32_t mkey(T, T m) { return inv(m) != T(0) ? 1u : 0u; }
    __device__ static uint32_t first(T, T) { return 0u; }
    __device__ static uint32_t dec(T s, T, T) { return s == inv(tmax<T>()) ? 1u : 0u; }
    __device__ static T enc(uint32_t k, T) { return k ? inv(tmax<T>()) : inv(T(0)); }
};

template <typename T>
__device__ __forceinline__ T ld(const char* base, int64_t pitch, int y, int x) {
    return *reinterpret_cast<const T*>(base + (int64_t)y * pitch + (int64_t)x * (int64_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st(char* base, int64_t pitch, int y, int x, T v) {
    *reinterpret_cast<T*>(base + (int64_t)y * pitch + (int64_t)x * (int64_t)sizeof(T)) = v;
}
// Device-coherent (agent-scope) access to the state buffer while tiles are
// visited concurrently from different XCDs: the loads and stores carry sc1 and
// do not rely on another XCD's L2 being written back or invalidated.
template <typename T>
__device__ __forceinline__ T ldc(const char* base, int64_t pitch, int y, int x) {
    const T* p = reinterpret_cast<const T*>(base + (int64_t)y * pitch + (int64_t)x * (int64_t)sizeof(T));
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void stc(char* base, int64_t pitch, int y, int x, T v) {
    T* p = reinterpret_cast<T*>(base + (int64_t)y * pitch + (int64_t)x * (int64_t)sizeof(T));
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// lane i <- lane i-1 (lane 0 <- 0), lane i <- lane i+1 (lane 63 <- 0); 0 is the
// identity of max over keys.
__device__ __forceinline__ uint32_t from_left(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xF, 0xF, true);  // wave_shr:1
}
__device__ __forceinline__ uint32_t from_right(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xF, 0xF, true);  // wave_shl:1
}

__device__ __forceinline__ uint32_t med3u(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

template <int CONN>
__device__ __forceinline__ uint32_t nbr(uint32_t p) {
    if constexpr (CONN == 8) return max(p, max(from_left(p), from_right(p)));
    else return p;
}

// Forward sweep over registers 1..TS-1 then backward over TS-2..0.
template <int CONN>
__device__ __forceinline__ void sweep_pair(uint32_t (&c)[TS], const uint32_t (&m)[TS]) {
    uint32_t p = c[0];
#pragma unroll
    for (int k = 1; k < TS; ++k) {
        c[k] = med3u(nbr<CONN>(p), c[k], m[k]);
        p = c[k];
    }
    p = c[TS - 1];
#pragma unroll
    for (int k = TS - 2; k >= 0; --k) {
        c[k] = med3u(nbr<CONN>(p), c[k], m[k]);
        p = c[k];
    }
}

// Register tile transpose through LDS: lane l, register k <-> lane k, register l.
__device__ __forceinline__ void transpose(uint32_t (&c)[TS], uint32_t* lds) {
    const int lane = threadIdx.x;
#pragma unroll
    for (int k = 0; k < TS; ++k) lds[k * LDS_PITCH + lane] = c[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TS; ++k) c[k] = lds[lane * LDS_PITCH + k];
    __syncthreads();
}

template <typename T> struct SumT { typedef uint64_t type; };
template <> struct SumT<uint8_t> { typedef uint32_t type; };

template <typename S>
__device__ __forceinline__ S wave_sum(S v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <typename S>
__device__ __forceinline__ S tile_sum(const uint32_t (&c)[TS]) {
    S s = 0;
#pragma unroll
    for (int k = 0; k < TS; ++k) s += (S)c[k];
    return wave_sum<S>(s);
}

// ---- device tile queue ------------------------------------------------------------
// Per-tile state bits: TQ_DIRTY = a visit is owed, TQ_RUN = a wave is visiting it.
// All transitions are RMWs, so a release fence before an activation and an
// acquire fence after the visit's own RMW order the tile data between waves.
constexpr uint32_t TQ_DIRTY = 1u, TQ_RUN = 2u;
constexpr int32_t RING_EMPTY = -1;
// A wait with no queue progress for this long (s_memrealtime runs at 100 MHz)
// gives up and sets an error bit, so the kernel always exits.
constexpr uint64_t STALL_TICKS = 200000000ull;   // 2 s

struct Queue {
    uint32_t* tstate;   // per tile
    int32_t* ring;      // tile ids, capacity a power of two >= tiles
    uint32_t mask;      // capacity - 1
    uint32_t* ctl;      // [0] head [1] tail [2] pending [3] error [4] visits
    uint32_t max_visits;  // visits after which the drain gives up (livelock guard)
};

__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t now_ticks() { return __builtin_amdgcn_s_memrealtime(); }

__device__ void q_push(const Queue& Q, int32_t t) {
    atomicAdd(&Q.ctl[2], 1u);
    const uint32_t pos = atomicAdd(&Q.ctl[1], 1u);
    int32_t* slot = &Q.ring[pos & Q.mask];
    const uint64_t t0 = now_ticks();
    while (atomicCAS(slot, RING_EMPTY, t) != RING_EMPTY) {
        __builtin_amdgcn_s_sleep(2);
        if (now_ticks() - t0 > STALL_TICKS) { atomicOr(&Q.ctl[3], 1u); return; }
    }
}

// Marks tile t as owing a visit; the caller's tile data must be released first.
__device__ __forceinline__ void q_activate(const Queue& Q, int32_t t) {
    if (atomicOr(&Q.tstate[t], TQ_DIRTY) == 0u) q_push(Q, t);
}

// Returns the next queued tile, or -1 when no tile is queued or being visited.
__device__ int32_t q_pop(const Queue& Q) {
    uint32_t last_visits = ld_relaxed(&Q.ctl[4]);
    if (last_visits > Q.max_visits) { atomicOr(&Q.ctl[3], 8u); return -1; }
    uint64_t t0 = now_ticks();
    while (true) {
        const uint32_t h = ld_relaxed(&Q.ctl[0]);
        const uint32_t t = ld_relaxed(&Q.ctl[1]);
        if (h < t) {
            if (atomicCAS(&Q.ctl[0], h, h + 1u) == h) {
                int32_t* slot = &Q.ring[h & Q.mask];
                const uint64_t s0 = now_ticks();
                while (true) {
                    const int32_t v = atomicExch(slot, RING_EMPTY);
                    if (v != RING_EMPTY) return v;
                    __builtin_amdgcn_s_sleep(1);
                    if (now_ticks() - s0 > STALL_TICKS) { atomicOr(&Q.ctl[3], 2u); return -1; }
                }
            }
            continue;
        }
        if (ld_relaxed(&Q.ctl[2]) == 0u || ld_relaxed(&Q.ctl[3]) != 0u) return -1;
        __builtin_amdgcn_s_sleep(8);
        const uint32_t v = ld_relaxed(&Q.ctl[4]);
        if (v != last_visits) { last_visits = v; t0 = now_ticks(); }
        else if (now_ticks() - t0 > STALL_TICKS) { atomicOr(&Q.ctl[3], 4u); return -1; }
    }
}

// One image of a reconstruction launch.
struct ImgDesc {
    const char* marker; int64_t pm;
    const char* mask;   int64_t pk;
    char* state;        int64_t ps;
    int rows, cols, ntx, nty;
    int tile_base;      // global id of the image's first tile
    int pad;
};

struct ReconArgs {
    ImgDesc img0;              // used when nimg == 1
    const ImgDesc* imgs;       // device table when nimg > 1
    const int32_t* tile_img;   // global tile -> image index when nimg > 1
    int nimg, ntiles;
    Queue Q;
};

__device__ __forceinline__ void tile_of(const ReconArgs& A, int32_t t, ImgDesc& D, int& lt) {
    if (A.nimg == 1) { D = A.img0; lt = t; return; }
    D = A.imgs[A.tile_img[t]];
    lt = t - D.tile_base;
}

__device__ __forceinline__ void activate(const ReconArgs& A, const ImgDesc& D, int ty, int tx) {
    if (ty < 0 || tx < 0 || ty >= D.nty || tx >= D.ntx) return;
    q_activate(A.Q, D.tile_base + ty * D.ntx + tx);
}

// Loads a tile (+ halo), runs it to its local fixed point, writes the interior
// back and activates the tiles whose pixels in the halo rose.
template <typename T, int MODE, int CONN, bool FIRST>
__device__ __forceinline__ void process_tile(const ReconArgs& A, const ImgDesc& D, int lt, uint32_t* lds) {
    using C = Codec<T, MODE>;
    using S = typename SumT<T>::type;
    const int lane = threadIdx.x;
    const int ty = lt / D.ntx, tx = lt - ty * D.ntx;
    const int x0 = tx * TI - 1, y0 = ty * TI - 1;
    const uint32_t F = C::frame();
    const bool needA = FIRST ? C::kMarkerFirst : C::kMarkerLater;

    auto key_at = [&](int gy, int gx, uint32_t& mk) -> uint32_t {
        if (gx < 0 || gx >= D.cols || gy < 0 || gy >= D.rows) { mk = F; return F; }
        const T mv = ld<T>(D.mask, D.pk, gy, gx);
        const T av = needA ? ld<T>(D.marker, D.pm, gy, gx) : T(0);
        mk = C::mkey(av, mv);
        if (FIRST) return C::first(av, mv);
        return C::dec(ldc<T>(D.state, D.ps, gy, gx), av, mv);
    };

    uint32_t c[TS], m[TS];
    const int gx = x0 + lane;
#pragma unroll
    for (int k = 0; k < TS; ++k) c[k] = key_at(y0 + k, gx, m[k]);
    uint32_t dummy;
    const uint32_t initL = key_at(y0 + lane, x0, dummy);
    const uint32_t initR = key_at(y0 + lane, x0 + TS - 1, dummy);
    const uint32_t initT = c[0], initB = c[TS - 1];

    // transposed mask for the horizontal sweeps
    uint32_t mt[TS];
#pragma unroll
    for (int k = 0; k < TS; ++k) mt[k] = m[k];
    transpose(mt, lds);

    const S tot_in = tile_sum<S>(c);
    S tot = tot_in;
    bool converged = false;
    for (int it = 0; it < MAX_IT; ++it) {
        sweep_pair<CONN>(c, m);        // along y
        transpose(c, lds);
        sweep_pair<CONN>(c, mt);       // along x
        const S t1 = tile_sum<S>(c);
        if (t1 == tot) { converged = true; break; }
        tot = t1;
        if (it + 1 == MAX_IT) break;
        transpose(c, lds);
    }
    // c is in transposed orientation here: register 0 / TS-1 = left / right halo column.
    const uint32_t finL = c[0], finR = c[TS - 1];
    transpose(c, lds);

    if (FIRST || tot != tot_in) {
        if (lane >= 1 && lane <= TI && gx < D.cols) {
#pragma unroll
            for (int k = 1; k <= TI; ++k) {
                const int gy = y0 + k;
                if (gy < D.rows) {
                    const T av = (MODE == RM_BINARY) ? ld<T>(D.marker, D.pm, gy, gx) : T(0);
                    if (FIRST) st<T>(D.state, D.ps, gy, gx, C::enc(c[k], av));
                    else stc<T>(D.state, D.ps, gy, gx, C::enc(c[k], av));
                }
            }
        }
    }

    // Halo pixels that rose above what was read: their owners must re-visit.
    const uint64_t topm = __ballot(c[0] > initT);
    const uint64_t botm = __ballot(c[TS - 1] > initB);
    const uint64_t lefm = __ballot(finL > initL);
    const uint64_t rigm = __ballot(finR > initR);
    const uint64_t mid = 0x7FFFFFFFFFFFFFFEull;
    const bool any = (topm | botm | ((lefm | rigm) & mid)) != 0ull || !converged;
    if (any) {
        __threadfence();   // release this tile's writes before anyone is told to read them
        if (lane == 0) {
            if (topm & mid) activate(A, D, ty - 1, tx);
            if (topm & 1ull) activate(A, D, ty - 1, tx - 1);
            if (topm >> 63) activate(A, D, ty - 1, tx + 1);
            if (botm & mid) activate(A, D, ty + 1, tx);
            if (botm & 1ull) activate(A, D, ty + 1, tx - 1);
            if (botm >> 63) activate(A, D, ty + 1, tx + 1);
            if (lefm & mid) activate(A, D, ty, tx - 1);
            if (rigm & mid) activate(A, D, ty, tx + 1);
            if (!converged) activate(A, D, ty, tx);
        }
    }
}

// First visit of every tile: state = clamped marker, then the local fixed point.
template <typename T, int MODE, int CONN>
__global__ __launch_bounds__(64) void recon_first_kernel(ReconArgs A) {
    __shared__ uint32_t lds[TS * LDS_PITCH];
    ImgDesc D;
    int lt;
    tile_of(A, (int32_t)blockIdx.x, D, lt);
    process_tile<T, MODE, CONN, true>(A, D, lt, lds);
}

// Persistent drain: each wave pops queued tiles until none is queued or running.
template <typename T, int MODE, int CONN>
__global__ __launch_bounds__(64) void recon_drain_kernel(ReconArgs A) {
    __shared__ uint32_t lds[TS * LDS_PITCH];
    const int lane = threadIdx.x;
    while (true) {
        int32_t t = 0;
        if (lane == 0) {
            t = q_pop(A.Q);
            if (t >= 0) atomicExch(&A.Q.tstate[t], TQ_RUN);   // clears TQ_DIRTY
        }
        t = __shfl(t, 0, 64);
        if (t < 0) break;
        __threadfence();   // acquire the data released by the tile's activators
        ImgDesc D;
        int lt;
        tile_of(A, t, D, lt);
        process_tile<T, MODE, CONN, false>(A, D, lt, lds);
        if (lane == 0) {
            if (atomicCAS(&A.Q.tstate[t], TQ_RUN, 0u) != TQ_RUN) {
                // activated while running: owe another visit
                atomicExch(&A.Q.tstate[t], TQ_DIRTY);
                q_push(A.Q, t);
            }
            atomicAdd(&A.Q.ctl[4], 1u);
            atomicSub(&A.Q.ctl[2], 1u);
        }
    }
}

__global__ void recon_reset_kernel(Queue Q, int ntiles, int cap) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    for (int j = i; j < ntiles; j += stride) Q.tstate[j] = 0u;
    for (int j = i; j < cap; j += stride) Q.ring[j] = RING_EMPTY;
    if (i < 8) Q.ctl[i] = 0u;
}

__global__ void recon_stats_kernel(Queue Q, int ntiles, int64_t* stats) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        stats[0] = Q.ctl[3];            // error bits (0 = converged)
        stats[1] = (int64_t)Q.ctl[4];   // drain visits
        stats[2] = ntiles;
        stats[3] |= (int64_t)Q.ctl[3];  // sticky until the host reads it
    }
}

template <typename T, int MODE, int CONN>
nsx_status run_typed(nsx_ctx* ctx, const ReconArgs& A, int cap, int workers, hipStream_t s) {
    const int grid_reset = std::max(1, std::min(1024, (std::max(A.ntiles, cap) + 255) / 256));
    hipLaunchKernelGGL(recon_reset_kernel, dim3(grid_reset), dim3(256), 0, s, A.Q, A.ntiles, cap);
    const int h0 = ctx->tic(TK_RECON_FIRST, s);
    hipLaunchKernelGGL((recon_first_kernel<T, MODE, CONN>), dim3(A.ntiles), dim3(64), 0, s, A);
    ctx->toc(h0, s);
    const int h1 = ctx->tic(TK_RECON_ROUND, s);
    hipLaunchKernelGGL((recon_drain_kernel<T, MODE, CONN>), dim3(workers), dim3(64), 0, s, A);
    ctx->toc(h1, s);
    hipLaunchKernelGGL(recon_stats_kernel, dim3(1), dim3(64), 0, s, A.Q, A.ntiles, ctx->d_stats);
    NSX_HIP_TRY(ctx, hipGetLastError());
    return NSX_OK;
}

template <typename T, int MODE>
nsx_status run_conn(nsx_ctx* ctx, const ReconArgs& A, int conn, int cap, int workers, hipStream_t s) {
    if (conn == 8) return run_typed<T, MODE, 8>(ctx, A, cap, workers, s);
    return run_typed<T, MODE, 4>(ctx, A, cap, workers, s);
}

ImgDesc make_desc(const View* marker, const View& mask, const View& state, int tile_base) {
    ImgDesc D{};
    D.marker = marker ? marker->data : nullptr;
    D.pm = marker ? marker->pitch : 0;
    D.mask = mask.data;
    D.pk = mask.pitch;
    D.state = state.data;
    D.ps = state.pitch;
    D.rows = mask.rows;
    D.cols = mask.cols;
    D.ntx = (D.cols + TI - 1) / TI;
    D.nty = (D.rows + TI - 1) / TI;
    D.tile_base = tile_base;
    return D;
}

}  // namespace

int recon_tiles(int rows, int cols) {
    if (rows <= 0 || cols <= 0) return 0;
    return ((cols + TI - 1) / TI) * ((rows + TI - 1) / TI);
}

nsx_status recon_launch(nsx_ctx* ctx, int mode, int dtype, int conn, const ReconBatch& B, hipStream_t s) {
    ReconArgs A{};
    A.nimg = B.nimg;
    A.ntiles = B.ntiles;
    if (A.ntiles <= 0) return NSX_OK;
    if (B.nimg == 1) A.img0 = *reinterpret_cast<const ImgDesc*>(B.host_desc);
    A.imgs = reinterpret_cast<const ImgDesc*>(B.dev_desc);
    A.tile_img = B.tile_img;
    int cap = 1;
    while (cap < A.ntiles) cap <<= 1;
    void *tstate, *ring, *ctl;
    NSX_TRY(ctx->ws(B.ws_base + 0, sizeof(uint32_t) * (size_t)A.ntiles, &tstate));
    NSX_TRY(ctx->ws(B.ws_base + 1, sizeof(int32_t) * (size_t)cap, &ring));
    NSX_TRY(ctx->ws(B.ws_base + 2, 64, &ctl));
    A.Q.tstate = (uint32_t*)tstate;
    A.Q.ring = (int32_t*)ring;
    A.Q.mask = (uint32_t)(cap - 1);
    A.Q.ctl = (uint32_t*)ctl;
    // Every visit raises at least one pixel or ends a tile's chain, so the visit
    // count is bounded by the value range; this generous cap only catches a
    // drain that stopped making progress while still visiting.
    A.Q.max_visits = (uint32_t)std::min<int64_t>(0x7FFFFFFF, std::max<int64_t>(1 << 20, (int64_t)A.ntiles * 512));
    // resident waves: the drain kernel holds one wave per SIMD (>256 VGPR+AGPR)
    const int per_cu = 4;
    const int workers = std::max(1, std::min(A.ntiles, per_cu * ctx->cu_count));
    switch (mode) {
        case RM_RECON:
            if (dtype == NSX_U8) return run_conn<uint8_t, RM_RECON>(ctx, A, conn, cap, workers, s);
            if (dtype == NSX_I32) return run_conn<int32_t, RM_RECON>(ctx, A, conn, cap, workers, s);
            return run_conn<float, RM_RECON>(ctx, A, conn, cap, workers, s);
        case RM_BINARY:
            if (dtype == NSX_U8) return run_conn<uint8_t, RM_BINARY>(ctx, A, conn, cap, workers, s);
            if (dtype == NSX_I32) return run_conn<int32_t, RM_BINARY>(ctx, A, conn, cap, workers, s);
            break;
        case RM_FILL:
            if (dtype == NSX_U8) return run_conn<uint8_t, RM_FILL>(ctx, A, conn, cap, workers, s);
            if (dtype == NSX_I32) return run_conn<int32_t, RM_FILL>(ctx, A, conn, cap, workers, s);
            break;
        case RM_FILLBIN:
            if (dtype == NSX_U8) return run_conn<uint8_t, RM_FILLBIN>(ctx, A, conn, cap, workers, s);
            break;
        default:
            break;
    }
    return ctx->fail(NSX_ERR_UNSUPPORTED, "reconstruction mode/dtype not instantiated by the reference");
}

size_t recon_desc_bytes() { return sizeof(ImgDesc); }

void recon_fill_desc(void* out, const View* marker, const View& mask, const View& state, int tile_base) {
    *reinterpret_cast<ImgDesc*>(out) = make_desc(marker, mask, state, tile_base);
}

nsx_status recon_run(nsx_ctx* ctx, int mode, const View* marker, const View& mask, const View& state,
                     int conn, hipStream_t s) {
    ImgDesc D = make_desc(marker, mask, state, 0);
    ReconBatch B{};
    B.nimg = 1;
    B.ntiles = D.ntx * D.nty;
    B.host_desc = &D;
    B.ws_base = WS_RECON_STAMP;
    if (mask.rows <= 0 || mask.cols <= 0) return NSX_OK;
    return recon_launch(ctx, mode, mask.dtype, conn == 8 ? 8 : 4, B, s);
}

}  // namespace nsx
