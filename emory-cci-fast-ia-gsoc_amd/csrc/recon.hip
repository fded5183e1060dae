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
//   * Round 0 visits every tile (state = min(marker, mask), then the local
//     fixed point).  Activated tiles are appended to a device-side work list
//     (atomicAdd on the round's counter, a per-tile stamp keeps a tile in a list
//     at most once); each later round is one launch over the previous round's
//     list, and the list length lives only on the device, so rounds are launched
//     back to back with no host round trip.  The host predicts the round count
//     from the previous call of the same size and checks convergence once, at
//     the call's synchronisation point, resuming if the prediction fell short.
//   * Several images (the batched form) share one launch: tiles are numbered
//     across images and a tile -> image table selects each tile's descriptor.
#include "nsx_internal.h"

#include <climits>
#include <cfloat>
#include <cmath>
#include <algorithm>
#include <type_traits>
#include <mutex>

namespace nsx {
namespace {

constexpr int TS = 64;          // register tile side, halo included
constexpr int TI = TS - 2;      // interior side written back
constexpr int LDS_PITCH = TS + 1;
constexpr int MAX_IT = 64;      // upper bound of the per-visit sweep-pair cap

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

// PixelOperations::invert<T> (PixelOperations.cpp:17-39): 255-x, -x, and for
// int32 saturating negation.  The int32 form is only reached by gray
// imfillHoles under NSX_I32_INVERT_SATURATE; under the default wrapping
// negation the int32 fill is a constant written by the caller.
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
// (MorphologicOperations.cpp:1665-1672,1712-1756).  The first pass stores
// the marker's value at its nonzero pixels and max<T> at the others it
// reached; a stored nonzero value inside the mask is then a reached pixel, so
// later rounds decode the key from state and mask alone, and write max<T> at
// the pixels a visit newly reached and nothing elsewhere (marker values stay).
template <typename T> struct Codec<T, RM_BINARY> {
    static constexpr bool kMarkerFirst = true, kMarkerLater = false;
    __device__ static uint32_t frame() { return 0u; }
    __device__ static uint32_t mkey(T, T m) { return m != T(0) ? 1u : 0u; }
    __device__ static uint32_t first(T a, T m) { return (a != T(0) && m != T(0)) ? 1u : 0u; }
    __device__ static uint32_t dec(T s, T, T m) { return (s != T(0) && m != T(0)) ? 1u : 0u; }
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

// Image accesses go through global-address-space pointers: global_load/store
// count only against vmcnt, so they do not serialise with the LDS traffic the
// way flat accesses (vmcnt and lgkmcnt) do.
template <typename T> using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ T ld(const char* base, int64_t pitch, int y, int x) {
    return *(gptr<const T>)(base + (int64_t)y * pitch + (int64_t)x * (int64_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st(char* base, int64_t pitch, int y, int x, T v) {
    *(gptr<T>)(base + (int64_t)y * pitch + (int64_t)x * (int64_t)sizeof(T)) = v;
}

// 32-bit register image of a pixel value and back.
template <typename T> __device__ __forceinline__ uint32_t bits(T v);
template <> __device__ __forceinline__ uint32_t bits<uint8_t>(uint8_t v) { return v; }
template <> __device__ __forceinline__ uint32_t bits<int32_t>(int32_t v) { return (uint32_t)v; }
template <> __device__ __forceinline__ uint32_t bits<float>(float v) { return __float_as_uint(v); }
template <typename T> __device__ __forceinline__ T unbits(uint32_t b);
template <> __device__ __forceinline__ uint8_t unbits<uint8_t>(uint32_t b) { return (uint8_t)b; }
template <> __device__ __forceinline__ int32_t unbits<int32_t>(uint32_t b) { return (int32_t)b; }
template <> __device__ __forceinline__ float unbits<float>(uint32_t b) { return __uint_as_float(b); }

// Row loads are issued in chunks of LOAD_CH; an empty asm that takes a chunk's
// registers as read-write operands marks the point where that chunk must have
// arrived.  Without it the compiler sinks each load next to its first use and
// waits for every row in turn.
constexpr int LOAD_CH = 16;
__device__ __forceinline__ void arrived(uint32_t (&r)[LOAD_CH]) {
    asm volatile("" : "+v"(r[0]), "+v"(r[1]), "+v"(r[2]), "+v"(r[3]), "+v"(r[4]), "+v"(r[5]), "+v"(r[6]),
                      "+v"(r[7]), "+v"(r[8]), "+v"(r[9]), "+v"(r[10]), "+v"(r[11]), "+v"(r[12]), "+v"(r[13]),
                      "+v"(r[14]), "+v"(r[15]));
}

// lane i <- lane i-1 (lane 0 <- 0), lane i <- lane i+1 (lane 63 <- 0); 0 is the
// identity of max over keys.
__device__ __forceinline__ uint32_t from_left(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xF, 0xF, true);  // wave_shr:1
}
__device__ __forceinline__ uint32_t from_right(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xF, 0xF, true);  // wave_shl:1
}

// Median of three.  Kept opaque to the scheduler: written as min/max (which
// the compiler also turns into v_med3_u32) it interleaves the sweep steps
// until the tile no longer fits the 168 VGPRs of three waves per SIMD.
__device__ __forceinline__ uint32_t med3u(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Maximum over the previous row's 1- or 3-neighbourhood.  For 8-connectivity
// the two maxima are kept apart (an empty asm holds the first result) so that
// each becomes a v_max_u32 with its DPP shift folded in, two instructions
// instead of two DPP moves and a v_max3.
template <int CONN>
__device__ __forceinline__ uint32_t nbr(uint32_t p) {
    if constexpr (CONN == 8) {
        uint32_t t = max(p, from_left(p));
        asm volatile("" : "+v"(t));
        return max(t, from_right(p));
    } else {
        return p;
    }
}

// The mask lives in LDS as mb[row * LDS_PITCH + col] (u8 keys for u8 images,
// u32 otherwise).  Orientation A (lane = column, register = row) reads
// mb[k * P + lane]; orientation B (lane = row, register = column) reads
// mb[lane * P + k]; the odd pitch keeps both conflict-free.
template <bool B, typename LT>
__device__ __forceinline__ uint32_t mask_at(const LT* mb, int k) {
    const int lane = threadIdx.x;
    return (uint32_t)(B ? mb[lane * LDS_PITCH + k] : mb[k * LDS_PITCH + lane]);
}

// One sweep step, r = med3(a, b, c) with b the register's old value, plus the
// change bookkeeping in the same asm statement (so each term is taken while the
// value is in a register and the scheduler cannot stretch live ranges):
//   sum_old: s += b (old value)      sum_new: s += r (new value)
//   chg:     s |= r ^ b
__device__ __forceinline__ uint32_t med3_sum_old(uint32_t a, uint32_t b, uint32_t c, uint32_t& s) {
    uint32_t r;
    asm("v_med3_u32 %0, %2, %3, %4\n\tv_add_u32 %1, %1, %3" : "=&v"(r), "+v"(s) : "v"(a), "v"(b), "v"(c));
    return r;
}
__device__ __forceinline__ uint32_t med3_sum_new(uint32_t a, uint32_t b, uint32_t c, uint32_t& s) {
    uint32_t r;
    asm("v_med3_u32 %0, %2, %3, %4\n\tv_add_u32 %1, %1, %0" : "=&v"(r), "+v"(s) : "v"(a), "v"(b), "v"(c));
    return r;
}
__device__ __forceinline__ uint32_t med3_chg(uint32_t a, uint32_t b, uint32_t c, uint32_t& s) {
    uint32_t r, t;
    asm("v_med3_u32 %0, %3, %4, %5\n\tv_xor_b32 %2, %0, %4\n\tv_or_b32 %1, %1, %2"
        : "=&v"(r), "+v"(s), "=&v"(t) : "v"(a), "v"(b), "v"(c));
    return r;
}

// A downward and an upward sweep as two independent chains of TS/2 steps each
// (a visit's time is its dependent chains, not VALU issue): first the top half
// downwards with the bottom half upwards, then each on through the other half.
// Every row is revisited with its neighbour's newer value, so a pair that
// changes nothing still means the tile is a fixed point along this axis; a
// pair moves a value across the tile as before, but a path that turns back
// inside the bottom half waits for the next pair.  Returns whether any of the
// lane's registers changed.  Values only rise, so for u8 keys a lane changed
// iff the sum of its registers rose (old values summed in the first half-step
// of each row, new ones in the second; 64 x 255 cannot wrap); 32-bit keys OR
// the differences instead.  The lane's mask column (row in orientation B) is
// read from LDS once, before both sweeps.
template <int CONN, bool B, typename LT>
__device__ __forceinline__ bool sweep_pair(uint32_t (&c)[TS], const LT* mb) {
    uint32_t m[TS];
#pragma unroll
    for (int k = 0; k < TS; ++k) m[k] = mask_at<B>(mb, k);
#pragma unroll
    for (int q = 0; q < TS / LOAD_CH; ++q) arrived(*reinterpret_cast<uint32_t(*)[LOAD_CH]>(&m[q * LOAD_CH]));
    constexpr bool kSum = sizeof(LT) == 1;
    constexpr int H = TS / 2;
    // Two independent chains at a time: first the top half downwards and the
    // bottom half upwards, then each on through the other half (down through
    // the bottom half from the top half's last row, up through the top half
    // from the bottom half's first row as the first phase left it).
    uint32_t before = kSum ? c[0] + c[TS - 1] : 0u, after = 0;
    uint32_t pd = c[0], pu = c[TS - 1];
#pragma unroll
    for (int s = 1; s < H; ++s) {
        const int kd = s, ku = TS - 1 - s;
        c[kd] = kSum ? med3_sum_old(nbr<CONN>(pd), c[kd], m[kd], before) : med3_chg(nbr<CONN>(pd), c[kd], m[kd], before);
        c[ku] = kSum ? med3_sum_old(nbr<CONN>(pu), c[ku], m[ku], before) : med3_chg(nbr<CONN>(pu), c[ku], m[ku], before);
        pd = c[kd];
        pu = c[ku];
    }
    pd = c[H - 1];
    pu = c[H];
#pragma unroll
    for (int s = 0; s < H; ++s) {
        const int kd = H + s, ku = H - 1 - s;
        c[kd] = kSum ? med3_sum_new(nbr<CONN>(pd), c[kd], m[kd], after) : med3_chg(nbr<CONN>(pd), c[kd], m[kd], after);
        c[ku] = kSum ? med3_sum_new(nbr<CONN>(pu), c[ku], m[ku], after) : med3_chg(nbr<CONN>(pu), c[ku], m[ku], after);
        pd = c[kd];
        pu = c[ku];
    }
    return kSum ? after != before : (before | after) != 0u;
}

// 32-bit keys with the lane's mask column (row in orientation B) held in
// registers for the whole visit (no LDS mask reads).
template <int CONN>
__device__ __forceinline__ bool sweep_pair_reg(uint32_t (&c)[TS], const uint32_t (&m)[TS]) {
    constexpr int H = TS / 2;
    uint32_t chg = 0;
    uint32_t pd = c[0], pu = c[TS - 1];
#pragma unroll
    for (int s = 1; s < H; ++s) {
        const int kd = s, ku = TS - 1 - s;
        c[kd] = med3_chg(nbr<CONN>(pd), c[kd], m[kd], chg);
        c[ku] = med3_chg(nbr<CONN>(pu), c[ku], m[ku], chg);
        pd = c[kd];
        pu = c[ku];
    }
    pd = c[H - 1];
    pu = c[H];
#pragma unroll
    for (int s = 0; s < H; ++s) {
        const int kd = H + s, ku = H - 1 - s;
        c[kd] = med3_chg(nbr<CONN>(pd), c[kd], m[kd], chg);
        c[ku] = med3_chg(nbr<CONN>(pu), c[ku], m[ku], chg);
        pd = c[kd];
        pu = c[ku];
    }
    return chg != 0u;
}

// Register tile transpose through LDS: lane l, register k <-> lane k, register l.
template <typename LT>
__device__ __forceinline__ void transpose(uint32_t (&c)[TS], LT* lds) {
    const int lane = threadIdx.x;
#pragma unroll
    for (int k = 0; k < TS; ++k) lds[k * LDS_PITCH + lane] = (LT)c[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TS; ++k) c[k] = lds[lane * LDS_PITCH + k];
    __syncthreads();
}

// LDS element of the mask / transpose buffers: every key of a u8 image (value,
// inverted value, 0/1 reached flag, frame) fits in a byte.
template <typename T> struct LdsT { typedef uint32_t type; };
template <> struct LdsT<uint8_t> { typedef uint8_t type; };

// ---- u8 tile loads in 16-byte blocks ----------------------------------------------
// A u8 tile row is 64 bytes at an arbitrary offset, so one load per row (lane =
// column) moves one byte per lane: 64 load instructions per image and tile.
// Instead each row's window is read as five 16-byte-aligned blocks (320 blocks,
// five load instructions per image), staged in LDS at WPITCH bytes per row, and
// each lane then picks its column's byte.  Row y's window starts at the block
// holding column xs = max(x0, 0) and needs the blocks up to the one holding
// column xe - 1 = min(x0 + 64, cols) - 1; the remaining block slots read that
// last block again (their bytes are never picked).  So every block read holds a
// byte of the row, and an aligned 16-byte block never straddles a page.
constexpr int WBLK = 5;
constexpr int WPITCH = 16 * WBLK;
static_assert(TS * WBLK == 5 * 64, "one block per lane per load instruction");

template <typename T> struct Stage { static constexpr int elems = TS * LDS_PITCH; };
template <> struct Stage<uint8_t> { static constexpr int elems = TS * WPITCH > TS * LDS_PITCH ? TS * WPITCH : TS * LDS_PITCH; };

__device__ __forceinline__ uintptr_t row_addr(const char* base, int64_t pitch, int y) {
    return (uintptr_t)base + (uintptr_t)((int64_t)y * pitch);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Block j (0..WBLK-1) of the window of row y.
__device__ __forceinline__ u32x4 ld_block(const char* base, int64_t pitch, int y, int xs, int xe, int j) {
    const uintptr_t row = row_addr(base, pitch, y);
    const uintptr_t first = (row + (uintptr_t)xs) & ~(uintptr_t)15;
    const uintptr_t last = (row + (uintptr_t)xe - 1u) & ~(uintptr_t)15;
    const uintptr_t a = min(first + 16u * (uintptr_t)j, last);
    return *(gptr<const u32x4>)a;
}

// Byte offset, in row y's staged window, of column xs.
__device__ __forceinline__ int win_off(const char* base, int64_t pitch, int y, int xs) {
    return (int)((row_addr(base, pitch, y) + (uintptr_t)xs) & 15u);
}

// Inclusive prefix sum over the wave: Hillis-Steele within each row of 16
// (DPP row_shr 1, 2, 4, 8), then row_bcast15 / row_bcast31 add the totals of
// the rows below.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return v;
}

// ---- propagation rounds over a device tile list ---------------------------------------
// Round 0 visits every tile.  A visit that raises a halo pixel appends the
// owning tile to the next round's list (a per-tile stamp keeps each tile at
// most once per round).  Round r reads list[r&1] and its count counts[r%3],
// appends to list[(r+1)&1] / counts[(r+1)%3] and clears counts[(r+2)%3] for
// the round after.  Kernel boundaries order the rounds, so a tile sees every
// write of the previous round; visits inside a round only read neighbours,
// whose values are valid lower bounds whether or not they are mid-update.
//
// Each list is split into NSEG segments with a counter each (256 bytes apart):
// a workgroup appends to segment blockIdx % NSEG, so appends from the XCDs
// (workgroups are dealt to them round-robin) go to different counters instead
// of queueing on one.  The next round reads the segments as one list, through
// the prefix sum of the counters.
constexpr int NSEG = 8;
constexpr int CSTRIDE = 64;   // ints between counters
__host__ __device__ inline int cnt_idx(int round, int seg) { return ((round % 3) * NSEG + seg) * CSTRIDE; }
constexpr int CNT_SUM = 3 * NSEG * CSTRIDE;   // scratch word: sum of one round's counters

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
    uint32_t* stamp;           // per tile: last round (+2) it was queued for
    int32_t* list0;            // [NSEG][ntiles]
    int32_t* list1;
    int32_t* counts;           // [3][NSEG] rotating round counters, CSTRIDE ints apart
    int64_t* stats;            // [0] rounds with work [1] tile visits after round 0 [2] tiles
    unsigned long long* cum;   // tile visits summed over runs (may be null)
    int visit_sweeps;          // sweep pairs per visit before re-queueing (1..MAX_IT)
    int first_list;            // round 0 takes its tiles from the list (fill: the border tiles)
    int xcd_split;             // each XCD takes a contiguous share of the work items (else a plain stride)
    int32_t* round_log;        // work-list length of each round, nsx_ctx::kRoundLog entries (may be null)
};

// The tile id is wave-uniform; readfirstlane makes that visible to the
// compiler, and the tile -> image entry and the descriptor are read through
// the constant address space, so they are scalar loads into SGPRs.
template <typename T> using cptr = const __attribute__((address_space(4))) T*;
__device__ __forceinline__ void tile_of(const ReconArgs& A, int32_t t, ImgDesc& D, int& lt) {
    t = __builtin_amdgcn_readfirstlane(t);
    if (A.nimg == 1) { D = A.img0; lt = t; return; }
    const int idx = __builtin_amdgcn_readfirstlane(((cptr<int32_t>)A.tile_img)[t]);
    static_assert(sizeof(ImgDesc) % 4 == 0, "descriptor copied as words");
    const cptr<uint32_t> src = (cptr<uint32_t>)(A.imgs + idx);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&D);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(ImgDesc) / 4); ++k) dst[k] = src[k];
    lt = t - D.tile_base;
}

// Queues up to nine tiles for the next round at once: lane j < 9 handles
// candidate j (the eight neighbours, then the tile itself) when bit j of
// `want` is set.  Each lane claims its tile through the per-tile stamp
// (atomicMax; a tile is queued at most once per round), and one atomicAdd
// reserves list slots for every lane that won a claim.
__device__ __forceinline__ void activate_all(const ReconArgs& A, const ImgDesc& D, int ty, int tx, int round,
                                             uint32_t want) {
    const int lane = threadIdx.x;
    // candidate j: up, up-left, up-right, down, down-left, down-right, left,
    // right, self; dx + 1 of candidate j is held in bits 2j..2j+1 of 0x18861
    const int dy = lane < 3 ? -1 : lane < 6 ? 1 : 0;
    const int dx = lane < 9 ? (int)((0x18861u >> (2 * lane)) & 3u) - 1 : 0;
    const int ny = ty + dy, nx = tx + dx;
    bool claim = false;
    int t = 0;
    if (lane < 9 && ((want >> lane) & 1u) && ny >= 0 && nx >= 0 && ny < D.nty && nx < D.ntx) {
        t = D.tile_base + ny * D.ntx + nx;
        const uint32_t stamp = (uint32_t)round + 2u;   // "queued for round+1"
        claim = atomicMax(&A.stamp[t], stamp) < stamp;
    }
    const uint64_t won = __ballot(claim);
    if (!won) return;
    const int first = __builtin_ctzll(won);
    int base = 0;
    const int seg = (int)(blockIdx.x % NSEG);
    if (lane == first) base = atomicAdd(&A.counts[cnt_idx(round + 1, seg)], __builtin_popcountll(won));
    base = __shfl(base, first, 64);
    if (claim) {
        int32_t* nl = (((round + 1) & 1) ? A.list1 : A.list0) + (int64_t)seg * A.ntiles;
        nl[base + __builtin_popcountll(won & ((1ull << lane) - 1ull))] = t;
    }
}

// Loads a tile (+ halo), runs it to its local fixed point, writes the interior
// back and queues the tiles whose pixels in the halo rose.
template <typename T, int MODE, int CONN, bool FIRST>
__device__ __forceinline__ void process_tile(const ReconArgs& A, const ImgDesc& D, int lt, int round,
                                             typename LdsT<T>::type* mb, typename LdsT<T>::type* tb,
                                             uint32_t* hs) {
    using LT = typename LdsT<T>::type;
    using C = Codec<T, MODE>;
    const int lane = threadIdx.x;
    const int ty = lt / D.ntx, tx = lt - ty * D.ntx;
    const int x0 = tx * TI - 1, y0 = ty * TI - 1;
    const uint32_t F = C::frame();
    const bool needA = FIRST ? C::kMarkerFirst : C::kMarkerLater;
    // 32-bit keys: the mask keys stay in registers, column and row orientation
    constexpr bool kRegMask = sizeof(T) == 4;
    uint32_t mA[kRegMask ? TS : 1], mB[kRegMask ? TS : 1];

    // Pixels outside the image read a clamped (valid) address and take the
    // frame value, so no load depends on a branch.
    auto key_of = [&](bool in, uint32_t m, uint32_t a, uint32_t st, uint32_t& mk) -> uint32_t {
        const T mv = unbits<T>(m), av = unbits<T>(a);
        const uint32_t key = FIRST ? C::first(av, mv) : C::dec(unbits<T>(st), av, mv);
        mk = in ? C::mkey(av, mv) : F;
        return in ? key : F;
    };
    uint32_t c[TS];
    const int gx = x0 + lane;
    const bool col_in = (unsigned)gx < (unsigned)D.cols;
    const int cx = min(max(gx, 0), D.cols - 1);
    // u8 images other than imreconstructBinary load in 16-byte blocks through
    // the staging buffer tb.  (Binary u8 through the blocks measured slower:
    // its later rounds spill 82 registers at 3 waves per SIMD.)
    constexpr bool kWide = sizeof(T) == 1 && MODE != RM_BINARY;
    if constexpr (kWide) {
        const bool kSec = FIRST ? C::kMarkerFirst : true;   // marker (first pass) or state
        const char* sec = FIRST ? D.marker : D.state;
        const int64_t psec = FIRST ? D.pm : D.ps;
        const int xs = max(x0, 0), xe = min(x0 + TS, D.cols);
        u32x4 bm[WBLK], bs[WBLK];
#pragma unroll
        for (int i = 0; i < WBLK; ++i) {
            const int q = i * TS + lane;
            const int k = q / WBLK, j = q - k * WBLK;
            const int cy = min(max(y0 + k, 0), D.rows - 1);
            bm[i] = ld_block(D.mask, D.pk, cy, xs, xe, j);
            if (kSec) bs[i] = ld_block(sec, psec, cy, xs, xe, j);
        }
        u32x4* t4 = reinterpret_cast<u32x4*>(tb);
        const int lo = cx - xs;
        // Row k's window offset: lane k computes it once.  When both pitches
        // are multiples of 16 every row has the same offset.
        const int cyl = min(max(y0 + lane, 0), D.rows - 1);
        const int offm_l = win_off(D.mask, D.pk, cyl, xs);
        const int offs_l = kSec ? win_off(sec, psec, cyl, xs) : 0;
        // Staged bytes are read 16 rows at a time with no condition (so the
        // reads issue back to back); the frame is applied afterwards, in tiles
        // that reach past the image only.
        auto decode = [&](auto off_of) {
#pragma unroll
            for (int i = 0; i < WBLK; ++i) t4[i * TS + lane] = bm[i];
            __syncthreads();
#pragma unroll
            for (int ch = 0; ch < TS / LOAD_CH; ++ch) {
                uint32_t v[LOAD_CH];
#pragma unroll
                for (int j = 0; j < LOAD_CH; ++j) {
                    const int k = ch * LOAD_CH + j;
                    v[j] = tb[k * WPITCH + off_of(offm_l, k) + lo];
                }
                arrived(v);
#pragma unroll
                for (int j = 0; j < LOAD_CH; ++j) {
                    const int k = ch * LOAD_CH + j;
                    mb[k * LDS_PITCH + lane] = (LT)C::mkey(T(0), (T)v[j]);
                    // imreconstruct's first key is min(marker, mask): keep the mask value
                    if (FIRST && MODE == RM_RECON) c[k] = v[j];
                }
            }
            __syncthreads();
            if (kSec) {
#pragma unroll
                for (int i = 0; i < WBLK; ++i) t4[i * TS + lane] = bs[i];
                __syncthreads();
            }
#pragma unroll
            for (int ch = 0; ch < TS / LOAD_CH; ++ch) {
                uint32_t v[LOAD_CH];
#pragma unroll
                for (int j = 0; j < LOAD_CH; ++j) {
                    const int k = ch * LOAD_CH + j;
                    v[j] = kSec ? (uint32_t)tb[k * WPITCH + off_of(offs_l, k) + lo] : 0u;
                }
                if (kSec) arrived(v);
#pragma unroll
                for (int j = 0; j < LOAD_CH; ++j) {
                    const int k = ch * LOAD_CH + j;
                    if constexpr (FIRST) c[k] = C::first((T)v[j], MODE == RM_RECON ? (T)c[k] : T(0));
                    else c[k] = C::dec((T)v[j], T(0), T(0));
                }
            }
            __syncthreads();   // tb is the transpose buffer from here on
        };
        if (((D.pk | psec) & 15) == 0)
            decode([](int v, int) { return __builtin_amdgcn_readfirstlane(v); });
        else
            decode([](int v, int k) { return __builtin_amdgcn_readlane(v, k); });
        if (x0 < 0 || y0 < 0 || x0 + TS > D.cols || y0 + TS > D.rows) {
#pragma unroll
            for (int k = 0; k < TS; ++k) {
                const bool in = col_in && (unsigned)(y0 + k) < (unsigned)D.rows;
                c[k] = in ? c[k] : F;
                if (!in) mb[k * LDS_PITCH + lane] = (LT)F;
            }
        }
    } else {
    // The tile's rows: chunk j+1 is in flight while chunk j is decoded.
    uint32_t rm[2][LOAD_CH], ra[2][LOAD_CH], rs[2][LOAD_CH];
    auto issue = [&](int b, int k0) {
#pragma unroll
        for (int j = 0; j < LOAD_CH; ++j) {
            const int cy = min(max(y0 + k0 + j, 0), D.rows - 1);
            rm[b][j] = bits<T>(ld<T>(D.mask, D.pk, cy, cx));
            ra[b][j] = needA ? bits<T>(ld<T>(D.marker, D.pm, cy, cx)) : 0u;
            rs[b][j] = FIRST ? 0u : bits<T>(ld<T>(D.state, D.ps, cy, cx));
        }
    };
    issue(0, 0);
    // hs[0..255]: the halo's keys before the sweeps (top row, bottom row, left
    // column, right column), parked in LDS while the sweeps hold the registers
#pragma unroll
    for (int ch = 0; ch < TS / LOAD_CH; ++ch) {
        const int b = ch & 1;
        if (ch + 1 < TS / LOAD_CH) issue(b ^ 1, (ch + 1) * LOAD_CH);
        arrived(rm[b]);
        if (needA) arrived(ra[b]);
        if (!FIRST) arrived(rs[b]);
#pragma unroll
        for (int j = 0; j < LOAD_CH; ++j) {
            const int k = ch * LOAD_CH + j;
            const bool in = col_in && (unsigned)(y0 + k) < (unsigned)D.rows;
            uint32_t mk;
            c[k] = key_of(in, rm[b][j], ra[b][j], rs[b][j], mk);
            if constexpr (kRegMask) mA[k] = mk;
            else mb[k * LDS_PITCH + lane] = (LT)mk;
        }
    }
    }
    // imreconstructBinary, later rounds: the pixels reached before the visit
    uint64_t oldb = 0;
    if constexpr (MODE == RM_BINARY && !FIRST) {
#pragma unroll
        for (int k = 1; k <= TI; ++k) oldb |= (uint64_t)(c[k] != 0u) << k;
    }
    if constexpr (kRegMask) {
#pragma unroll
        for (int k = 0; k < TS; ++k) mB[k] = mA[k];
        transpose(mB, tb);
    }
    hs[lane] = c[0];
    hs[TS + lane] = c[TS - 1];
    // left / right halo columns are lanes 0 and TS-1: they write them out so
    // that lane l later finds row l's two values (no column loads from HBM)
    if (lane == 0 || lane == TS - 1) {
        const int base = lane == 0 ? 2 * TS : 3 * TS;
#pragma unroll
        for (int k = 0; k < TS; ++k) hs[base + k] = c[k];
    }
    __syncthreads();

    // Half-passes alternate between the column sweeps (orientation A: lane =
    // column) and the row sweeps (after a transpose, orientation B: lane = row).
    // A half-pass changed the tile iff some lane's sweep saw a register rise.
    // 4-connectivity: a half-pass is an exact 1-D reconstruction of every
    // column (row) and so idempotent; one quiet half-pass after the other
    // orientation ran proves the fixed point.  8-connectivity: the diagonal
    // terms break idempotence, so two consecutive quiet half-passes are needed.
    bool converged = false;
    bool orientB = false;
    bool changed = false;
    int quiet = 0;
    int half = 0;
    auto settled = [&](bool lane_moved) {
        const bool moved = __ballot(lane_moved) != 0ull;
        changed = changed || moved;
        quiet = moved ? 0 : quiet + 1;
        return CONN == 4 ? (quiet >= 1 && half >= 1) : (quiet >= 2);
    };
    for (int it = 0; it < A.visit_sweeps; ++it) {
        bool done;   // columns (orientation A)
        if constexpr (kRegMask) done = settled(sweep_pair_reg<CONN>(c, mA));
        else done = settled(sweep_pair<CONN, false>(c, mb));
        if (done) { converged = true; break; }
        ++half;
        transpose(c, tb);
        // rows (orientation B)
        if constexpr (kRegMask) done = settled(sweep_pair_reg<CONN>(c, mB));
        else done = settled(sweep_pair<CONN, true>(c, mb));
        if (done) { converged = true; orientB = true; break; }
        ++half;
        transpose(c, tb);
    }
    if (lane == 0) atomicAdd((unsigned long long*)&A.stats[3], (unsigned long long)(half + 1));
    uint32_t finL, finR;
    if (orientB) {
        // register 0 / TS-1 = left / right halo column
        finL = c[0];
        finR = c[TS - 1];
        transpose(c, tb);
    } else {
        transpose(c, tb);
        finL = c[0];
        finR = c[TS - 1];
        transpose(c, tb);
    }

    if (FIRST || changed) {
        if (lane >= 1 && lane <= TI && gx < D.cols) {
#pragma unroll
            for (int k = 1; k <= TI; ++k) {
                const int gy = y0 + k;
                if (gy < D.rows) {
                    if constexpr (MODE == RM_BINARY && !FIRST) {
                        if (c[k] != 0u && !((oldb >> k) & 1u)) st<T>(D.state, D.ps, gy, gx, tmax<T>());
                    } else {
                        const T av = (MODE == RM_BINARY) ? ld<T>(D.marker, D.pm, gy, gx) : T(0);
                        st<T>(D.state, D.ps, gy, gx, C::enc(c[k], av));
                    }
                }
            }
        }
    }

    // Halo pixels that rose above what was read: their owners must re-visit.
    const uint64_t topm = __ballot(c[0] > hs[lane]);
    const uint64_t botm = __ballot(c[TS - 1] > hs[TS + lane]);
    const uint64_t lefm = __ballot(finL > hs[2 * TS + lane]);
    const uint64_t rigm = __ballot(finR > hs[3 * TS + lane]);
    const uint64_t mid = 0x7FFFFFFFFFFFFFFEull;
    const uint32_t want = ((topm & mid) ? 1u : 0u) | ((topm & 1ull) ? 2u : 0u) | ((topm >> 63) ? 4u : 0u) |
                          ((botm & mid) ? 8u : 0u) | ((botm & 1ull) ? 16u : 0u) | ((botm >> 63) ? 32u : 0u) |
                          ((lefm & mid) ? 64u : 0u) | ((rigm & mid) ? 128u : 0u) | (converged ? 0u : 256u);
    if (want) activate_all(A, D, ty, tx, round, want);
}

// u8 tiles keep their mask, block staging / transpose buffer and halo keys in
// 10.3 KB of LDS and run 3 waves per SIMD (at most 168 VGPRs); 32-bit tiles
// run 2 (at most 256 VGPRs).  f32 later rounds run 1 wave per SIMD: at 2 they
// spilled 43 registers, at 1 they take 296 with none (batch 29.37 -> 28.91 ms
// in a same-session A/B).  Binary reconstruction's later rounds also keep a
// bit mask of the pixels reached before the visit; they run 2 (u8: no spills
// instead of 7) or 1 (int32: none instead of 7): a 2048^2 maze, 8-connected,
// 0.92 ms against 0.95 at 3 and 2.
template <typename T, int MODE, int CONN, bool FIRST>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(MODE == RM_BINARY && !FIRST ? (sizeof(T) == 1 ? 2 : 1)
                                    : sizeof(T) == 1 ? 3 : (std::is_same<T, float>::value && !FIRST ? 1 : 2))))
void recon_round_kernel(ReconArgs A, int round) {
    using LT = typename LdsT<T>::type;
    __shared__ LT mb[sizeof(T) == 1 ? TS * LDS_PITCH : 1];   // u8 mask keys of the tile
    __shared__ __align__(16) LT tb[Stage<T>::elems];      // u8 block staging, then transposes
    __shared__ uint32_t hs[4 * TS];     // halo keys before the sweeps
    const int lane = threadIdx.x;
    if (blockIdx.x == 0) {
        if (lane < NSEG) A.counts[cnt_idx(round + 2, lane)] = 0;
        if (!FIRST && lane == 0) {
            int n = 0;
            for (int q = 0; q < NSEG; ++q) n += A.counts[cnt_idx(round, q)];
            if (n > 0) {
                A.stats[0] += 1;
                A.stats[1] += n;
                if (A.cum) atomicAdd(A.cum, (unsigned long long)n);
            }
            if (A.round_log && round < nsx_ctx::kRoundLog) A.round_log[round] = n;
        }
    }
    // Blocks go to the XCDs in turn (block b to XCD b mod 8).  With xcd_split
    // XCD x takes the x-th eighth of the work items, so each XCD works on one
    // region of the image and neighbouring tiles meet in its L2 (the default
    // for one image).  A batch defaults to the plain stride: its tiles differ
    // in cost, and contiguous chunks can pile the costly ones onto one XCD.
    const int nx = !A.xcd_split ? 1 : (gridDim.x < 8 ? (int)gridDim.x : 8);
    const int xcd = (int)blockIdx.x % nx, xb = (int)blockIdx.x / nx;
    const int xstride = ((int)gridDim.x - xcd + nx - 1) / nx;
    if (FIRST && !A.first_list) {
        const int chunk = (A.ntiles + nx - 1) / nx;
        const int lo = min(A.ntiles, xcd * chunk), hi = min(A.ntiles, lo + chunk);
        for (int i = lo + xb; i < hi; i += xstride) {
            ImgDesc D;
            int lt;
            tile_of(A, i, D, lt);
            process_tile<T, MODE, CONN, FIRST>(A, D, lt, round, mb, tb, hs);
            __syncthreads();   // the next tile overwrites mb
        }
        return;
    }
    // lane q < NSEG: segment q's count and the inclusive prefix over segments
    const uint32_t cq = lane < NSEG ? (uint32_t)A.counts[cnt_idx(round, lane)] : 0u;
    const uint32_t incl = wave_incl_scan(cq);
    const int n = __builtin_amdgcn_readlane((int)incl, NSEG - 1);
    const int32_t* list = (round & 1) ? A.list1 : A.list0;
    const int chunk = (n + nx - 1) / nx;
    const int lo = min(n, xcd * chunk), hi = min(n, lo + chunk);
    for (int i = lo + xb; i < hi; i += xstride) {
        const int q = __builtin_ctzll(__ballot(lane < NSEG && incl > (uint32_t)i));
        const int off = i - (__builtin_amdgcn_readlane((int)incl, q) - __builtin_amdgcn_readlane((int)cq, q));
        const int t = __builtin_amdgcn_readfirstlane(list[(int64_t)q * A.ntiles + off]);
        ImgDesc D;
        int lt;
        tile_of(A, t, D, lt);
        process_tile<T, MODE, CONN, FIRST>(A, D, lt, round, mb, tb, hs);
        __syncthreads();   // the next tile overwrites mb
    }
}

// ---- persistent drain of the round lists -------------------------------------------
// One launch runs every propagation round after the first pass: its blocks stay
// resident and go through the rounds' device lists in the same order as the
// per-round launches, with a grid-wide barrier where a kernel boundary was.  The
// grid is the number of blocks the GPU holds at once (occupancy x CUs), so every
// block is running when the first one waits.  The launch sequence no longer
// depends on the content: it needs no round prediction and no host check, and
// a graph of it has the same nodes however many rounds a replay takes.
//
// A barrier: the block's first thread writes its XCD's L2 back (release, so
// that the tiles it stored and the list entries it appended reach memory),
// adds one to the arrival counter, polls it with coherent loads until every
// block of this barrier has arrived, then invalidates its L2 (acquire) before
// the block reads the next round's counters, list and tiles.  Arrivals only
// grow: barrier k waits for k x grid.  Every wait is bounded: a barrier that
// sees no new arrival for `stall` ticks of the 100 MHz real-time counter (a
// block that never became resident, on a GPU shared with other work) sets the
// error word, every block leaves at its next barrier, and the host restarts
// the reconstruction with per-round launches.  Rounds that end with an empty
// list make every block leave together (all read the same final counters).
struct Drain {
    uint32_t* ctl;      // [0] barrier arrivals, [1] error, [2] the round the drain stopped at
    int max_round;      // rounds below this one run here; the host continues from there
    uint32_t stall;     // real-time ticks a barrier waits without a new arrival
};
// The drain's control words sit in the counters' buffer, 128 bytes past the
// sum word; the reset kernel clears them with the counters.
constexpr int CTL_OFF = CNT_SUM + 32;
constexpr int CTL_WORDS = 4;

__device__ __forceinline__ uint32_t ld_agent(const void* p) {
    return __hip_atomic_load((const uint32_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(void* p, uint32_t v) {
    __hip_atomic_store((uint32_t*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier number target / gridDim.x; false (in every thread of the
// block) when the drain has failed and the block must leave.
__device__ __forceinline__ bool drain_barrier(const Drain& P, uint32_t target, uint32_t* okw) {
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ok = 1u;
        __threadfence();
        atomicAdd(&P.ctl[0], 1u);
        uint32_t last = 0u;
        uint64_t since = __builtin_amdgcn_s_memrealtime();
        for (;;) {
            const uint32_t v = ld_agent(&P.ctl[0]);
            if (v >= target) break;
            if (ld_agent(&P.ctl[1]) != 0u) { ok = 0u; break; }
            const uint64_t now = __builtin_amdgcn_s_memrealtime();
            if (v != last) {
                last = v;
                since = now;
            } else if (now - since > (uint64_t)P.stall) {
                atomicOr(&P.ctl[1], 1u);
                ok = 0u;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __threadfence();
        *okw = ok;
    }
    __syncthreads();
    return *okw != 0u;
}

// Round header of the drain, by block 0: clears the counters of round + 2 and
// counts the round's work into the statistics.
__device__ __forceinline__ void drain_header(const ReconArgs& A, int round, int n, int lane) {
    if (lane < NSEG) st_agent(&A.counts[cnt_idx(round + 2, lane)], 0u);
    if (lane == 0) {
        A.stats[0] += 1;
        A.stats[1] += n;
        if (A.cum) atomicAdd(A.cum, (unsigned long long)n);
        if (A.round_log && round < nsx_ctx::kRoundLog) A.round_log[round] = n;
    }
}

// The work items of one round: the same XCD-aware split of the list as the
// per-round kernels.  `excl` (LDS) holds the segments' exclusive prefix over
// the counters, NSEG + 1 words.  visit(t, round) processes global tile t.
template <typename Visit>
__device__ __forceinline__ void drain_round_items(const ReconArgs& A, int round, const uint32_t* excl, int n,
                                                  Visit visit) {
    const int nx = !A.xcd_split ? 1 : (gridDim.x < 8 ? (int)gridDim.x : 8);
    const int xcd = (int)blockIdx.x % nx, xb = (int)blockIdx.x / nx;
    const int xstride = ((int)gridDim.x - xcd + nx - 1) / nx;
    const int32_t* list = (round & 1) ? A.list1 : A.list0;
    const int chunk = (n + nx - 1) / nx;
    const int lo = min(n, xcd * chunk), hi = min(n, lo + chunk);
    for (int i = lo + xb; i < hi; i += xstride) {
        int q = 0;
#pragma unroll
        for (int k = 1; k < NSEG; ++k) q += excl[k] <= (uint32_t)i ? 1 : 0;
        q = __builtin_amdgcn_readfirstlane(q);
        const int off = i - (int)excl[q];
        visit(__builtin_amdgcn_readfirstlane(list[(int64_t)q * A.ntiles + off]), round);
        __syncthreads();   // the next tile overwrites the tile's LDS
    }
}

// The drain loop shared by the one-wave and four-wave tiles.  excl: NSEG + 1
// words of LDS.
template <typename Visit>
__device__ __forceinline__ void drain_rounds(const ReconArgs& A, const Drain& P, uint32_t* okw, uint32_t* excl,
                                             Visit visit) {
    const int lane = threadIdx.x & 63;
    uint32_t target = 0u;
    for (int round = 1; round < P.max_round; ++round) {
        // the counters of this round are final: their appends came before the barrier
        const uint32_t cq = lane < NSEG ? ld_agent(&A.counts[cnt_idx(round, lane)]) : 0u;
        const uint32_t incl = wave_incl_scan(cq);
        const int n = __builtin_amdgcn_readlane((int)incl, NSEG - 1);
        if (threadIdx.x < NSEG) excl[threadIdx.x] = incl - cq;
        __syncthreads();
        if (n == 0) {
            // done; the previous round's counters are the only non-zero ones
            // left: clear them, so every counter reads zero for the host's check
            if (blockIdx.x == 0 && threadIdx.x < 64) {
                if (lane < NSEG) st_agent(&A.counts[cnt_idx(round + 2, lane)], 0u);
                if (lane == 0) st_agent(&P.ctl[2], (uint32_t)round);
            }
            return;
        }
        if (blockIdx.x == 0 && threadIdx.x < 64) drain_header(A, round, n, lane);
        drain_round_items(A, round, excl, n, visit);
        target += gridDim.x;
        if (!drain_barrier(P, target, okw)) return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) st_agent(&P.ctl[2], (uint32_t)P.max_round);
}

template <typename T, int MODE, int CONN>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 1 ? 3 : 2)))
void recon_drain_kernel(ReconArgs A, Drain P) {
    using LT = typename LdsT<T>::type;
    __shared__ LT mb[sizeof(T) == 1 ? TS * LDS_PITCH : 1];
    __shared__ __align__(16) LT tb[Stage<T>::elems];
    __shared__ uint32_t hs[4 * TS];
    __shared__ uint32_t okw, excl[NSEG];
    drain_rounds(A, P, &okw, excl, [&](int t, int round) {
        ImgDesc D;
        int lt;
        tile_of(A, t, D, lt);
        process_tile<T, MODE, CONN, false>(A, D, lt, round, mb, tb, hs);
    });
}

// dst = sum of one round's NSEG counters (zero iff the round has no work), or
// -1 when the drain whose control words are `ctl` gave up.
__global__ void sum_counts_kernel(int32_t* dst, const int32_t* round_counts, const uint32_t* ctl) {
    if (threadIdx.x == 0) {
        int n = 0;
        for (int q = 0; q < NSEG; ++q) n += round_counts[q * CSTRIDE];
        *dst = (ctl && ctl[1] != 0u) ? -1 : n;
    }
}

// A barrier of the drain gives up after 0.5 s without a new arrival.
constexpr uint32_t DRAIN_STALL_TICKS = 50u * 1000u * 1000u;

// Number of tiles on the border of an ntx x nty tile grid.
__host__ __device__ inline int border_tiles(int ntx, int nty) {
    return (ntx <= 2 || nty <= 2) ? ntx * nty : 2 * ntx + 2 * (nty - 2);
}

// The initial state written into every pixel when a hole filling starts from
// its border tiles: invert(min) = max, one byte (u8) or one int32.
struct StateFill {
    char* state; int64_t ps;
    int rows, cols;   // rows == 0: nothing to write
    int esize;        // 1 or 4
    uint32_t value;
};

// Clears the stamps, the round counters and the stats.  With ntx > 0 it also
// lists the border tiles of a single ntx x nty image as round 0's work, and
// writes the initial state F; with all_tiles it lists every tile instead.
__global__ void recon_reset_kernel(uint32_t* stamp, int ntiles, int32_t* counts, int64_t* stats, int32_t* list0,
                                   int ntx, int nty, StateFill F, int all_tiles, int32_t* round_log) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    // the initial state, a row per block: 16-byte stores between the row's first
    // and last 16-byte boundaries, element stores before and after (and for a
    // whole int32 row that is not 4-byte aligned, where the pattern would be
    // out of phase)
    const int64_t rowbytes = (int64_t)F.cols * F.esize;
    const uint32_t pat = F.esize == 1 ? (F.value & 0xFFu) * 0x01010101u : F.value;
    for (int y = blockIdx.x; y < F.rows; y += gridDim.x) {
        char* p = F.state + (int64_t)y * F.ps;
        const uintptr_t a = (uintptr_t)p;
        int64_t h = std::min((int64_t)((16 - (a & 15)) & 15), rowbytes);
        if (F.esize == 4 && (a & 3)) h = rowbytes;
        const int64_t body = (rowbytes - h) / 16 * 16;
        for (int part = 0; part < 2; ++part) {   // before and after the 16-byte body
            const int64_t e0 = part == 0 ? 0 : h + body, e1 = part == 0 ? h : rowbytes;
            for (int64_t e = e0 + (int64_t)threadIdx.x * F.esize; e < e1; e += (int64_t)blockDim.x * F.esize) {
                if (F.esize == 1) *(uint8_t*)(p + e) = (uint8_t)F.value;
                else *(int32_t*)(p + e) = (int32_t)F.value;
            }
        }
        const u32x4 v4 = {pat, pat, pat, pat};
        for (int64_t o = h + 16 * (int64_t)threadIdx.x; o < h + body; o += 16 * (int64_t)blockDim.x)
            *(u32x4*)(p + o) = v4;
    }
    for (int j = i; j < ntiles; j += gridDim.x * blockDim.x) stamp[j] = 0u;
    const int nb = all_tiles ? ntiles : ntx > 0 ? border_tiles(ntx, nty) : 0;
    for (int b = i; b < nb; b += gridDim.x * blockDim.x) {
        int t;
        if (all_tiles || ntx <= 2 || nty <= 2) t = b;
        else if (b < ntx) t = b;                                   // top row
        else if (b < 2 * ntx) t = (nty - 1) * ntx + (b - ntx);     // bottom row
        else {                                                     // left / right column
            const int q = b - 2 * ntx;
            t = (1 + q / 2) * ntx + ((q & 1) ? ntx - 1 : 0);
        }
        list0[b] = t;
    }
    if (i < 3 * NSEG) counts[i * CSTRIDE] = i == 0 ? nb : 0;
    if (i < CTL_WORDS) counts[CTL_OFF + i] = 0;
    if (i == 0) { stats[0] = 0; stats[1] = 0; stats[2] = ntiles; stats[3] = 0; }
    if (i == 0 && round_log) round_log[0] = nb > 0 ? nb : ntiles;   // round 0: the first pass
}

// ---- image-wide row and column scans (hole filling's start) -------------------
// Along one line (a row or a column) the reconstruction of state s under mask
// m is exact in two scans: the forward value f(x) = med3(f(x-1), s(x), m(x))
// (keys; f(-1) = the frame) carries every source on the left, the backward
// value b(x) likewise from the right, and the line's result is max(f, b), both
// computed from the same s.  Each element is the clamp v -> med3(v, s, m)
// (s <= m), and clamps compose into clamps, so a line splits into pieces whose
// compositions are combined first.  Rows: one wave per row, lane = column
// within 64-column pieces, the pieces' compositions combined by a wave scan.
// Columns: one wave per 64-column strip and 32-row segment; the segments'
// compositions go through global memory (scan_cols_seg_kernel), a pass over
// them gives each segment's carries from above and below
// (scan_cols_carry_kernel), and each segment then runs its rows from those
// (scan_cols_apply_kernel).  Alternating row and column scans carry the
// frame's value across the open parts of the image before the tile rounds
// start, so the rounds are left with the twisted paths.  Results are the same
// with or without scans: every scan value is a lower bound of the
// reconstruction and at least the state it started from.
constexpr int SCAN_SEG = 32;   // rows per column segment
constexpr uint32_t KEY_ALL = 0xFFFFFFFFu;

// (lo, hi) <- (lo, hi) after (plo, phi): the clamp composition v -> (lo,hi)((plo,phi)(v)).
__device__ __forceinline__ void clamp_after(uint32_t& lo, uint32_t& hi, uint32_t plo, uint32_t phi) {
    const uint32_t nlo = med3u(plo, lo, hi), nhi = med3u(phi, lo, hi);
    lo = nlo;
    hi = nhi;
}

// Keys of a pixel: its mask key and its state key (INIT: the state before any
// propagation, Codec::first).
template <typename T, int MODE, bool INIT>
__device__ __forceinline__ void scan_keys(const ImgDesc& D, int y, int x, uint32_t& ks, uint32_t& km) {
    using C = Codec<T, MODE>;
    const T m = ld<T>(D.mask, D.pk, y, x);
    km = C::mkey(T(0), m);
    ks = INIT ? C::first(T(0), m) : C::dec(ld<T>(D.state, D.ps, y, x), T(0), m);
}

// One wave per row; NS 64-column pieces per row at most (cols <= 64 NS).
template <typename T, int MODE, bool INIT, int NS>
__global__ __launch_bounds__(256) void scan_rows_kernel(ImgDesc D) {
    using C = Codec<T, MODE>;
    const int lane = threadIdx.x & 63;
    const int y = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (y >= D.rows) return;
    const int npc = (D.cols + 63) >> 6;
    const uint32_t F = C::frame();
    uint32_t bw[NS];
    // backward: pieces right to left; lane l's composition covers lanes l..63
    uint32_t v = F;
#pragma unroll
    for (int i = NS - 1; i >= 0; --i) {
        if (i < npc) {
            const int x = i * 64 + lane;
            uint32_t lo = 0u, hi = KEY_ALL;   // identity past the row's end
            if (x < D.cols) scan_keys<T, MODE, INIT>(D, y, x, lo, hi);
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t plo = __shfl_down(lo, d, 64), phi = __shfl_down(hi, d, 64);
                if (lane + d < 64) clamp_after(lo, hi, plo, phi);
            }
            const uint32_t out = med3u(v, lo, hi);
            bw[i] = out;
            v = __builtin_amdgcn_readlane(out, 0);
        }
    }
    // forward: pieces left to right; the result is the larger of both values
    v = F;
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        if (i < npc) {
            const int x = i * 64 + lane;
            uint32_t lo = 0u, hi = KEY_ALL;
            if (x < D.cols) scan_keys<T, MODE, INIT>(D, y, x, lo, hi);
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t plo = __shfl_up(lo, d, 64), phi = __shfl_up(hi, d, 64);
                if (lane >= d) clamp_after(lo, hi, plo, phi);
            }
            const uint32_t out = med3u(v, lo, hi);
            if (x < D.cols) st<T>(D.state, D.ps, y, x, C::enc(max(out, bw[i]), T(0)));
            v = __builtin_amdgcn_readlane(out, 63);
        }
    }
}

// Column segments: the forward (top to bottom) and backward compositions of
// each SCAN_SEG-row segment of each column, as four planes [nseg][colsp] of
// segf: forward lo, forward hi, backward lo, backward hi.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void scan_cols_seg_kernel(ImgDesc D, uint32_t* segf, int nseg, int colsp) {
    const int lane = threadIdx.x & 63;
    const int wid = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int strips = colsp >> 6;
    if (wid >= strips * nseg) return;
    const int strip = wid % strips, seg = wid / strips;
    const int x = strip * 64 + lane, cx = min(x, D.cols - 1);
    const int y0 = seg * SCAN_SEG;
    uint32_t ks[SCAN_SEG], km[SCAN_SEG];
#pragma unroll
    for (int k = 0; k < SCAN_SEG; ++k) {
        ks[k] = 0u;
        km[k] = KEY_ALL;   // identity below the image
        if (y0 + k < D.rows) scan_keys<T, MODE, false>(D, y0 + k, cx, ks[k], km[k]);
    }
    uint32_t flo = 0u, fhi = KEY_ALL, blo = 0u, bhi = KEY_ALL;
#pragma unroll
    for (int k = 0; k < SCAN_SEG; ++k) {
        uint32_t lo = ks[k], hi = km[k];   // row k after the rows above it
        clamp_after(lo, hi, flo, fhi);
        flo = lo;
        fhi = hi;
        uint32_t lo2 = ks[SCAN_SEG - 1 - k], hi2 = km[SCAN_SEG - 1 - k];   // and after the rows below
        clamp_after(lo2, hi2, blo, bhi);
        blo = lo2;
        bhi = hi2;
    }
    const size_t plane = (size_t)nseg * colsp, at = (size_t)seg * colsp + x;
    segf[at] = flo;
    segf[plane + at] = fhi;
    segf[2 * plane + at] = blo;
    segf[3 * plane + at] = bhi;
}

// The value entering each segment from above (forward) and from below
// (backward), planes 4 and 5 of segf.  A block of four waves takes 64
// columns (lane = column); wave w composes the segments of its quarter of the
// column, the quarters' compositions meet in LDS, and each wave then walks its
// quarter again from its carry.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void scan_cols_carry_kernel(uint32_t* segf, int nseg, int colsp) {
    using C = Codec<T, MODE>;
    __shared__ uint32_t part[4][4][64];   // [wave][flo fhi blo bhi][lane]
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int x = blockIdx.x * 64 + lane;   // colsp is a multiple of 64: always a column of segf
    const size_t plane = (size_t)nseg * colsp;
    const int ch = (nseg + 3) / 4, q0 = min(nseg, w * ch), q1 = min(nseg, q0 + ch);
    uint32_t flo = 0u, fhi = KEY_ALL, blo = 0u, bhi = KEY_ALL;
#pragma unroll 8
    for (int q = q0; q < q1; ++q) {   // this quarter's forward composition, top down
        const size_t at = (size_t)q * colsp + x;
        uint32_t lo = segf[at], hi = segf[plane + at];
        clamp_after(lo, hi, flo, fhi);
        flo = lo;
        fhi = hi;
    }
#pragma unroll 8
    for (int q = q1 - 1; q >= q0; --q) {   // and its backward composition, bottom up
        const size_t at = (size_t)q * colsp + x;
        uint32_t lo = segf[2 * plane + at], hi = segf[3 * plane + at];
        clamp_after(lo, hi, blo, bhi);
        blo = lo;
        bhi = hi;
    }
    part[w][0][lane] = flo;
    part[w][1][lane] = fhi;
    part[w][2][lane] = blo;
    part[w][3][lane] = bhi;
    __syncthreads();
    uint32_t vf = C::frame(), vb = C::frame();
    for (int u = 0; u < w; ++u) vf = med3u(vf, part[u][0][lane], part[u][1][lane]);
    for (int u = 3; u > w; --u) vb = med3u(vb, part[u][2][lane], part[u][3][lane]);
#pragma unroll 8
    for (int q = q0; q < q1; ++q) {
        const size_t at = (size_t)q * colsp + x;
        segf[4 * plane + at] = vf;
        vf = med3u(vf, segf[at], segf[plane + at]);
    }
#pragma unroll 8
    for (int q = q1 - 1; q >= q0; --q) {
        const size_t at = (size_t)q * colsp + x;
        segf[5 * plane + at] = vb;
        vb = med3u(vb, segf[2 * plane + at], segf[3 * plane + at]);
    }
}

// Column segments, last pass: each segment's rows from its two carries.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void scan_cols_apply_kernel(ImgDesc D, const uint32_t* segf, int nseg, int colsp) {
    using C = Codec<T, MODE>;
    const int lane = threadIdx.x & 63;
    const int wid = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int strips = colsp >> 6;
    if (wid >= strips * nseg) return;
    const int strip = wid % strips, seg = wid / strips;
    const int x = strip * 64 + lane, cx = min(x, D.cols - 1);
    const int y0 = seg * SCAN_SEG;
    const size_t plane = (size_t)nseg * colsp, at = (size_t)seg * colsp + x;
    uint32_t vf = segf[4 * plane + at], vb = segf[5 * plane + at];
    uint32_t ks[SCAN_SEG], km[SCAN_SEG], bw[SCAN_SEG];
#pragma unroll
    for (int k = 0; k < SCAN_SEG; ++k) {
        ks[k] = 0u;
        km[k] = KEY_ALL;
        if (y0 + k < D.rows) scan_keys<T, MODE, false>(D, y0 + k, cx, ks[k], km[k]);
    }
#pragma unroll
    for (int k = SCAN_SEG - 1; k >= 0; --k) {
        vb = med3u(vb, ks[k], km[k]);
        bw[k] = vb;
    }
#pragma unroll
    for (int k = 0; k < SCAN_SEG; ++k) {
        vf = med3u(vf, ks[k], km[k]);
        if (y0 + k < D.rows && x < D.cols) st<T>(D.state, D.ps, y0 + k, x, C::enc(max(vf, bw[k]), T(0)));
    }
}

template <typename T, int MODE, bool INIT>
void launch_scan_rows(const ImgDesc& D, hipStream_t s) {
    const dim3 grid((D.rows + 3) / 4), block(256);
    if (D.cols <= 1024) hipLaunchKernelGGL((scan_rows_kernel<T, MODE, INIT, 16>), grid, block, 0, s, D);
    else if (D.cols <= 2048) hipLaunchKernelGGL((scan_rows_kernel<T, MODE, INIT, 32>), grid, block, 0, s, D);
    else hipLaunchKernelGGL((scan_rows_kernel<T, MODE, INIT, 64>), grid, block, 0, s, D);
}

// `pairs` row + column scan pairs over one image, the first row scan starting
// from the initial state (the state buffer is written everywhere).  segf:
// scan_segf_words() words.  cols <= SCAN_MAX_COLS.
constexpr int SCAN_MAX_COLS = 4096;
// Images whose tile grid is narrower than this on either side keep the
// border start: their rounds are few, and the scans' launches would not pay.
constexpr int SCAN_MIN_TILES = 24;
template <typename T, int MODE>
void run_scans(const ImgDesc& D, int pairs, uint32_t* segf, hipStream_t s) {
    const int nseg = (D.rows + SCAN_SEG - 1) / SCAN_SEG, colsp = ((D.cols + 63) / 64) * 64;
    const int waves = (colsp / 64) * nseg;
    for (int p = 0; p < pairs; ++p) {
        if (p == 0) launch_scan_rows<T, MODE, true>(D, s);
        else launch_scan_rows<T, MODE, false>(D, s);
        hipLaunchKernelGGL((scan_cols_seg_kernel<T, MODE>), dim3((waves + 3) / 4), dim3(256), 0, s, D, segf, nseg, colsp);
        hipLaunchKernelGGL((scan_cols_carry_kernel<T, MODE>), dim3(colsp / 64), dim3(256), 0, s, segf, nseg, colsp);
        hipLaunchKernelGGL((scan_cols_apply_kernel<T, MODE>), dim3((waves + 3) / 4), dim3(256), 0, s, D,
                           (const uint32_t*)segf, nseg, colsp);
    }
}

size_t scan_segf_words(int rows, int cols) {
    return (size_t)6 * ((rows + SCAN_SEG - 1) / SCAN_SEG) * (((size_t)cols + 63) / 64) * 64;
}

// out[0] = *ncomp; out[1], out[2] = tiles queued after the launched rounds of
// two reconstructions (the sums of their next round's counters); int64 out[2],
// out[3] = their rounds with work.
__global__ void chain_status_kernel(int32_t* out, const int32_t* ncomp, const int32_t* rc, const int32_t* fc,
                                    const int64_t* rs, const int64_t* fs, const uint32_t* rctl, const uint32_t* fctl) {
    if (threadIdx.x == 0) {
        int a = 0, b = 0;
        for (int q = 0; q < NSEG; ++q) {
            a += rc[q * CSTRIDE];
            b += fc[q * CSTRIDE];
        }
        out[0] = *ncomp;
        out[1] = (rctl && rctl[1] != 0u) ? -1 : a;   // -1: the drain gave up
        out[2] = (fctl && fctl[1] != 0u) ? -1 : b;
        out[3] = 0;
        int64_t* o64 = reinterpret_cast<int64_t*>(out);
        o64[2] = rs[0];
        o64[3] = fs[0];
    }
}



// ---- 4-connectivity tiles on four waves -------------------------------------------
// For 4-connectivity a sweep step sets a register to med3(previous, cur, mask)
// = clamp(previous, cur, mask) (cur <= mask), and clamps compose into clamps:
//     clamp(clamp(x, a1, b1), a2, b2) = clamp(x, med3(a1, a2, b2), med3(b1, a2, b2)).
// A tile is split over four waves of QR registers each (rows in orientation A,
// columns in orientation B).  In a sweep the wave holding the starting halo
// line runs its registers while the other three compose their QR clamps (two
// independent chains); after one barrier each of them clamps its entry value
// through the composed blocks on the way and reruns its registers.  A 63-step
// dependent chain becomes 16 + up to 3 + 16 steps on four SIMDs.  Hole filling
// with 4-connectivity uses this, and 4-connected reconstruction for its rounds
// after the first pass: such rounds hold few tiles each, so a round lasts as
// long as one visit's dependent chain.
constexpr int QW = 4;
constexpr int QR = TS / QW;

struct QShared {
    uint32_t xch[2][(2 * QW - 1) * TS];   // per sweep direction: slot 0 the starting wave's exit,
                                          // slots 2p-1, 2p the composed clamp at position p
    uint32_t hs[4 * TS];               // halo keys before the sweeps
    uint32_t flag[QW];                 // per wave: a register rose in this half-pass
    uint32_t want[QW];                 // per wave: neighbour tiles to queue
};

// One sweep direction over the four waves.  DOWN: registers 0..QR-1 of waves
// 0..3 in order, starting from wave 0's register 0; otherwise the reverse,
// starting from wave 3's register QR-1.  One barrier, between the composing
// and the rerun; each direction has its own exchange slots, and the next
// writes to them come after the other direction's barrier.
template <bool DOWN>
__device__ __forceinline__ void qsweep(uint32_t (&c)[QR], const uint32_t (&m)[QR], int w, uint32_t* xch) {
    const int lane = threadIdx.x & 63;
    const int ws = DOWN ? 0 : QW - 1;
    const int pos = DOWN ? w : QW - 1 - w;   // position along the sweep, 0 = starting wave
    if (w == ws) {
        uint32_t p = DOWN ? c[0] : c[QR - 1];
#pragma unroll
        for (int s = 1; s < QR; ++s) {
            const int k = DOWN ? s : QR - 1 - s;
            c[k] = med3u(p, c[k], m[k]);
            p = c[k];
        }
        xch[lane] = p;
    } else {
        const int k0 = DOWN ? 0 : QR - 1;
        uint32_t lo = c[k0], hi = m[k0];
#pragma unroll
        for (int s = 1; s < QR; ++s) {
            const int k = DOWN ? s : QR - 1 - s;
            lo = med3u(lo, c[k], m[k]);
            hi = med3u(hi, c[k], m[k]);
        }
        xch[(2 * pos - 1) * TS + lane] = lo;
        xch[(2 * pos) * TS + lane] = hi;
    }
    __syncthreads();
    if (w != ws) {
        uint32_t x = xch[lane];
        for (int r = 1; r < pos; ++r) x = med3u(x, xch[(2 * r - 1) * TS + lane], xch[(2 * r) * TS + lane]);
#pragma unroll
        for (int s = 0; s < QR; ++s) {
            const int k = DOWN ? s : QR - 1 - s;
            c[k] = med3u(x, c[k], m[k]);
            x = c[k];
        }
    }
}

// Orientation A (wave w: rows QR*w.., lane = column) <-> orientation B (wave
// w: columns QR*w.., lane = row) through LDS.  TAIL: a closing barrier, for
// when tb is written again before another barrier.
template <bool TO_B, bool TAIL, typename LT>
__device__ __forceinline__ void qtranspose(uint32_t (&c)[QR], LT* tb, int w) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < QR; ++k) {
        const int posA = (QR * w + k) * LDS_PITCH + lane, posB = lane * LDS_PITCH + QR * w + k;
        tb[TO_B ? posA : posB] = (LT)c[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < QR; ++k) {
        const int posA = (QR * w + k) * LDS_PITCH + lane, posB = lane * LDS_PITCH + QR * w + k;
        c[k] = tb[TO_B ? posB : posA];
    }
    if (TAIL) __syncthreads();
}

// A visit of one tile by four waves (4-connectivity; every mode reads mask
// and state only after the first pass).
template <typename T, int MODE, bool FIRST>
__device__ __forceinline__ void process_tile_q(const ReconArgs& A, const ImgDesc& D, int lt, int round,
                                               typename LdsT<T>::type* tb, QShared& Q) {
    using C = Codec<T, MODE>;
    static_assert(!C::kMarkerLater, "four-wave tiles read mask and state only after round 0");
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ty = lt / D.ntx, tx = lt - ty * D.ntx;
    const int x0 = tx * TI - 1, y0 = ty * TI - 1;
    const uint32_t F = C::frame();
    constexpr bool needA = FIRST && C::kMarkerFirst;
    const int gx = x0 + lane;
    const bool col_in = (unsigned)gx < (unsigned)D.cols;
    const int cx = min(max(gx, 0), D.cols - 1);
    // this wave's QR rows, every load in flight at once
    uint32_t rm[QR], ra[QR], rs[QR];
#pragma unroll
    for (int k = 0; k < QR; ++k) {
        const int cy = min(max(y0 + QR * w + k, 0), D.rows - 1);
        rm[k] = bits<T>(ld<T>(D.mask, D.pk, cy, cx));
        ra[k] = needA ? bits<T>(ld<T>(D.marker, D.pm, cy, cx)) : 0u;
        rs[k] = FIRST ? 0u : bits<T>(ld<T>(D.state, D.ps, cy, cx));
    }
    uint32_t c[QR], mA[QR], mB[QR];
#pragma unroll
    for (int k = 0; k < QR; ++k) {
        const bool in = col_in && (unsigned)(y0 + QR * w + k) < (unsigned)D.rows;
        const T mv = unbits<T>(rm[k]), av = unbits<T>(ra[k]);
        const uint32_t key = FIRST ? C::first(av, mv) : C::dec(unbits<T>(rs[k]), av, mv);
        c[k] = in ? key : F;
        mA[k] = in ? C::mkey(av, mv) : F;
    }
    // imreconstructBinary, later rounds: the pixels reached before the visit
    uint32_t oldq = 0;
    if constexpr (MODE == RM_BINARY && !FIRST) {
#pragma unroll
        for (int k = 0; k < QR; ++k) oldq |= (uint32_t)(c[k] != 0u) << k;
    }
#pragma unroll
    for (int k = 0; k < QR; ++k) mB[k] = mA[k];
    qtranspose<true, false>(mB, tb, w);
    if (w == 0) Q.hs[lane] = c[0];
    if (w == QW - 1) Q.hs[TS + lane] = c[QR - 1];
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < QR; ++k) Q.hs[2 * TS + QR * w + k] = c[k];
    }
    if (lane == TS - 1) {
#pragma unroll
        for (int k = 0; k < QR; ++k) Q.hs[3 * TS + QR * w + k] = c[k];
    }
    __syncthreads();

    // As in the one-wave visit: 4-connectivity half-passes are idempotent, so
    // one quiet half-pass after the other orientation ran is the fixed point.
    bool converged = false, orientB = false, changed = false;
    int quiet = 0, half = 0;
    // A half-pass: both sweeps; then each wave writes its "a register rose"
    // flag and its registers for the transpose, and one barrier serves both.
    // A converged visit keeps its registers and skips the transpose's reads.
    // (Flags and tb are written again only after the next half-pass's two
    // sweep barriers, which every reader passes first.)
    // u8 keys: values only rise and 16 x 255 cannot wrap, so a lane changed iff
    // its sum rose; 32-bit keys compare against a copy.
    auto half_pass = [&](const uint32_t (&m)[QR], bool toB) {
        constexpr bool kSum = sizeof(T) == 1;
        uint32_t before[kSum ? 1 : QR], s0 = 0;
#pragma unroll
        for (int k = 0; k < QR; ++k) {
            if (kSum) s0 += c[k];
            else before[kSum ? 0 : k] = c[k];
        }
        qsweep<true>(c, m, w, Q.xch[0]);
        qsweep<false>(c, m, w, Q.xch[1]);
        uint32_t d = 0, s1 = 0;
#pragma unroll
        for (int k = 0; k < QR; ++k) {
            if (kSum) s1 += c[k];
            else d |= c[k] ^ before[kSum ? 0 : k];
        }
        if (kSum) d = s1 ^ s0;
        const bool mine = __ballot(d != 0u) != 0ull;   // every lane votes
        if (lane == 0) Q.flag[w] = mine ? 1u : 0u;
#pragma unroll
        for (int k = 0; k < QR; ++k) {
            const int posA = (QR * w + k) * LDS_PITCH + lane, posB = lane * LDS_PITCH + QR * w + k;
            tb[toB ? posA : posB] = (typename LdsT<T>::type)c[k];
        }
        __syncthreads();
        uint32_t fl = 0u;
#pragma unroll
        for (int q = 0; q < QW; ++q) fl |= Q.flag[q];
        const bool moved = fl != 0u;
        changed = changed || moved;
        quiet = moved ? 0 : quiet + 1;
        const bool done = quiet >= 1 && half >= 1;
        if (!done) {
#pragma unroll
            for (int k = 0; k < QR; ++k) {
                const int posA = (QR * w + k) * LDS_PITCH + lane, posB = lane * LDS_PITCH + QR * w + k;
                c[k] = tb[toB ? posB : posA];
            }
        }
        return done;
    };
    for (int it = 0; it < A.visit_sweeps; ++it) {
        if (half_pass(mA, true)) { converged = true; break; }   // columns (orientation A), then to B
        ++half;
        if (half_pass(mB, false)) { converged = true; orientB = true; break; }   // rows, then back to A
        ++half;
    }
    if (threadIdx.x == 0) atomicAdd((unsigned long long*)&A.stats[3], (unsigned long long)(half + 1));
    // left / right halo columns: orientation B's register 0 of wave 0 and
    // register QR-1 of the last wave
    uint32_t finL, finR;
    if (orientB) {
        finL = c[0];
        finR = c[QR - 1];
        qtranspose<false, true>(c, tb, w);
    } else {
        qtranspose<true, true>(c, tb, w);
        finL = c[0];
        finR = c[QR - 1];
        qtranspose<false, true>(c, tb, w);
    }

    if (FIRST || changed) {
        if (lane >= 1 && lane <= TI && gx < D.cols) {
#pragma unroll
            for (int k = 0; k < QR; ++k) {
                const int r = QR * w + k, gy = y0 + r;
                if (r >= 1 && r <= TI && gy < D.rows) {
                    if constexpr (MODE == RM_BINARY && !FIRST) {
                        if (c[k] != 0u && !((oldq >> k) & 1u)) st<T>(D.state, D.ps, gy, gx, tmax<T>());
                    } else {
                        st<T>(D.state, D.ps, gy, gx, C::enc(c[k], T(0)));
                    }
                }
            }
        }
    }

    // halo pixels that rose above what was read: their owners must re-visit
    const uint64_t mid = 0x7FFFFFFFFFFFFFFEull;
    uint32_t wb = 0;
    if (w == 0) {
        const uint64_t topm = __ballot(c[0] > Q.hs[lane]);
        const uint64_t lefm = __ballot(finL > Q.hs[2 * TS + lane]);
        wb = ((topm & mid) ? 1u : 0u) | ((topm & 1ull) ? 2u : 0u) | ((topm >> 63) ? 4u : 0u) | ((lefm & mid) ? 64u : 0u);
    } else if (w == QW - 1) {
        const uint64_t botm = __ballot(c[QR - 1] > Q.hs[TS + lane]);
        const uint64_t rigm = __ballot(finR > Q.hs[3 * TS + lane]);
        wb = ((botm & mid) ? 8u : 0u) | ((botm & 1ull) ? 16u : 0u) | ((botm >> 63) ? 32u : 0u) |
             ((rigm & mid) ? 128u : 0u);
    }
    if (lane == 0) Q.want[w] = wb;
    __syncthreads();
    if (w == 0) {
        const uint32_t want = Q.want[0] | Q.want[QW - 1] | (converged ? 0u : 256u);
        if (want) activate_all(A, D, ty, tx, round, want);
    }
}

// u8: at most 128 VGPRs, so four tiles (16 waves) fit on a CU.
template <typename T, int MODE, bool FIRST>
__global__ __launch_bounds__(QW * 64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 1 ? 4 : 3)))
void recon_quad_kernel(ReconArgs A, int round) {
    using LT = typename LdsT<T>::type;
    __shared__ LT tb[TS * LDS_PITCH];   // transposes
    __shared__ QShared Q;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (blockIdx.x == 0 && w == 0) {
        if (lane < NSEG) A.counts[cnt_idx(round + 2, lane)] = 0;
        if (!FIRST && lane == 0) {
            int n = 0;
            for (int q = 0; q < NSEG; ++q) n += A.counts[cnt_idx(round, q)];
            if (n > 0) {
                A.stats[0] += 1;
                A.stats[1] += n;
                if (A.cum) atomicAdd(A.cum, (unsigned long long)n);
            }
            if (A.round_log && round < nsx_ctx::kRoundLog) A.round_log[round] = n;
        }
    }
    // Blocks go to the XCDs in turn (block b to XCD b mod 8).  With xcd_split
    // XCD x takes the x-th eighth of the work items, so each XCD works on one
    // region of the image and neighbouring tiles meet in its L2 (the default
    // for one image).  A batch defaults to the plain stride: its tiles differ
    // in cost, and contiguous chunks can pile the costly ones onto one XCD.
    const int nx = !A.xcd_split ? 1 : (gridDim.x < 8 ? (int)gridDim.x : 8);
    const int xcd = (int)blockIdx.x % nx, xb = (int)blockIdx.x / nx;
    const int xstride = ((int)gridDim.x - xcd + nx - 1) / nx;
    if (FIRST && !A.first_list) {
        const int chunk = (A.ntiles + nx - 1) / nx;
        const int lo = min(A.ntiles, xcd * chunk), hi = min(A.ntiles, lo + chunk);
        for (int i = lo + xb; i < hi; i += xstride) {
            ImgDesc D;
            int lt;
            tile_of(A, i, D, lt);
            process_tile_q<T, MODE, FIRST>(A, D, lt, round, tb, Q);
            __syncthreads();
        }
        return;
    }
    const uint32_t cq = lane < NSEG ? (uint32_t)A.counts[cnt_idx(round, lane)] : 0u;
    const uint32_t incl = wave_incl_scan(cq);
    const int n = __builtin_amdgcn_readlane((int)incl, NSEG - 1);
    const int32_t* list = (round & 1) ? A.list1 : A.list0;
    const int chunk = (n + nx - 1) / nx;
    const int lo = min(n, xcd * chunk), hi = min(n, lo + chunk);
    for (int i = lo + xb; i < hi; i += xstride) {
        const int q = __builtin_ctzll(__ballot(lane < NSEG && incl > (uint32_t)i));
        const int off = i - (__builtin_amdgcn_readlane((int)incl, q) - __builtin_amdgcn_readlane((int)cq, q));
        const int t = __builtin_amdgcn_readfirstlane(list[(int64_t)q * A.ntiles + off]);
        ImgDesc D;
        int lt;
        tile_of(A, t, D, lt);
        process_tile_q<T, MODE, FIRST>(A, D, lt, round, tb, Q);
        __syncthreads();
    }
}

// The persistent drain on four-wave tiles (4-connected hole filling).
template <typename T, int MODE>
__global__ __launch_bounds__(QW * 64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 1 ? 4 : 3)))
void recon_quad_drain_kernel(ReconArgs A, Drain P) {
    using LT = typename LdsT<T>::type;
    __shared__ LT tb[TS * LDS_PITCH];
    __shared__ QShared Q;
    __shared__ uint32_t okw, excl[NSEG];
    drain_rounds(A, P, &okw, excl, [&](int t, int round) {
        ImgDesc D;
        int lt;
        tile_of(A, t, D, lt);
        process_tile_q<T, MODE, false>(A, D, lt, round, tb, Q);
    });
}

typedef void (*RoundFn)(const ReconArgs&, int, int, hipStream_t);
// Launches a persistent drain; returns the grid it used (0: not launched).
typedef int (*DrainFn)(const ReconArgs&, const Drain&, int cu_count, int occ, hipStream_t);

// Blocks of `kernel` one CU holds at once (at most `design`), from the
// compiled kernel's registers and LDS.
template <typename K>
int resident_blocks(K kernel, int block, int design) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block, 0) != hipSuccess) {
        (void)hipGetLastError();
        nb = 0;
    }
    return std::max(0, std::min(nb, design));
}

// occ > 0 asks for fewer blocks per CU than stay resident.
template <typename T, int MODE, int CONN>
int launch_drain(const ReconArgs& A, const Drain& P, int cu_count, int occ, hipStream_t s) {
    static const int per_cu = resident_blocks(recon_drain_kernel<T, MODE, CONN>, 64, sizeof(T) == 1 ? 12 : 8);
    if (per_cu <= 0) return 0;
    const int grid = (occ > 0 ? std::min(occ, per_cu) : per_cu) * cu_count;
    hipLaunchKernelGGL((recon_drain_kernel<T, MODE, CONN>), dim3(grid), dim3(64), 0, s, A, P);
    return grid;
}

template <typename T, int MODE>
int launch_quad_drain(const ReconArgs& A, const Drain& P, int cu_count, int occ, hipStream_t s) {
    static const int per_cu = resident_blocks(recon_quad_drain_kernel<T, MODE>, QW * 64, sizeof(T) == 1 ? 4 : 3);
    if (per_cu <= 0) return 0;
    const int grid = (occ > 0 ? std::min(occ, per_cu) : per_cu) * cu_count;
    hipLaunchKernelGGL((recon_quad_drain_kernel<T, MODE>), dim3(grid), dim3(QW * 64), 0, s, A, P);
    return grid;
}


template <typename T, int MODE, int CONN, bool FIRST>
void launch_round(const ReconArgs& A, int round, int grid, hipStream_t s) {
    hipLaunchKernelGGL((recon_round_kernel<T, MODE, CONN, FIRST>), dim3(grid), dim3(64), 0, s, A, round);
}

template <typename T, int MODE, bool FIRST>
void launch_quad(const ReconArgs& A, int round, int grid, hipStream_t s) {
    hipLaunchKernelGGL((recon_quad_kernel<T, MODE, FIRST>), dim3(grid), dim3(QW * 64), 0, s, A, round);
}

struct RoundFns {
    RoundFn first = nullptr, later = nullptr;
    DrainFn drain = nullptr;
};

// Hole filling with 4-connectivity runs on four-wave tiles.
template <typename T, int MODE>
bool pick_fill(int conn, RoundFns* f) {
    if (conn == 8) {
        f->first = launch_round<T, MODE, 8, true>;
        f->later = launch_round<T, MODE, 8, false>;
        f->drain = launch_drain<T, MODE, 8>;
    } else {
        f->first = launch_quad<T, MODE, true>;
        f->later = launch_quad<T, MODE, false>;
        f->drain = launch_quad_drain<T, MODE>;
    }
    return true;
}

template <typename T, int MODE>
bool pick(int conn, RoundFns* f) {
    if (conn == 8) {
        f->first = launch_round<T, MODE, 8, true>;
        f->later = launch_round<T, MODE, 8, false>;
        f->drain = launch_drain<T, MODE, 8>;
    } else {
        // 4-connected reconstruction: the first pass (every tile) on one-wave
        // tiles, the later rounds (few tiles, a visit's chain sets the time)
        // on four-wave tiles, as the hole filling's
        f->first = launch_round<T, MODE, 4, true>;
        f->later = launch_quad<T, MODE, false>;
        f->drain = launch_quad_drain<T, MODE>;
    }
    return true;
}

bool select_fns(int mode, int dtype, int conn, RoundFns* f) {
    switch (mode) {
        case RM_RECON:
            if (dtype == NSX_U8) return pick<uint8_t, RM_RECON>(conn, f);
            if (dtype == NSX_I32) return pick<int32_t, RM_RECON>(conn, f);
            return pick<float, RM_RECON>(conn, f);
        case RM_BINARY:
            if (dtype == NSX_U8) return pick<uint8_t, RM_BINARY>(conn, f);
            if (dtype == NSX_I32) return pick<int32_t, RM_BINARY>(conn, f);
            return false;
        case RM_FILL:
            if (dtype == NSX_U8) return pick_fill<uint8_t, RM_FILL>(conn, f);
            if (dtype == NSX_I32) return pick_fill<int32_t, RM_FILL>(conn, f);
            return false;
        case RM_FILLBIN:
            if (dtype == NSX_U8) return pick_fill<uint8_t, RM_FILLBIN>(conn, f);
            return false;
        default:
            return false;
    }
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

// Host-side state of one reconstruction between its launch and its check.
struct ReconRunImpl {
    ReconArgs A;
    RoundFn later;
    int grid;
    int next_round;   // first round not yet launched
    uint64_t key;     // round-count memory key
    bool drain = false;   // rounds 1 .. next_round - 1 ran in a persistent drain
    int mode = 0, dtype = 0, conn = 0;
    ReconBatch batch;     // to start again with per-round launches if the drain fails
};

int recon_tiles(int rows, int cols) {
    if (rows <= 0 || cols <= 0) return 0;
    return ((cols + TI - 1) / TI) * ((rows + TI - 1) / TI);
}

size_t recon_desc_bytes() { return sizeof(ImgDesc); }

size_t recon_list_bytes(int ntiles) { return 2 * sizeof(int32_t) * (size_t)NSEG * (size_t)ntiles; }

void recon_fill_desc(void* out, const View* marker, const View& mask, const View& state, int tile_base) {
    *reinterpret_cast<ImgDesc*>(out) = make_desc(marker, mask, state, tile_base);
}

void ReconRun::release() {
    delete impl;
    impl = nullptr;
}

// Records the rounds a problem needed.  A smaller count lowers the prediction
// by one round per call only: an over-prediction costs empty launches of a few
// microseconds, an under-prediction a host round trip and a re-run.
static void note_hint(nsx_ctx* ctx, uint64_t key, int64_t rounds_with_work) {
    ctx->round_hint.note(key, rounds_with_work);
}

nsx_status recon_launch(nsx_ctx* ctx, int mode, int dtype, int conn, const ReconBatch& B, hipStream_t s,
                        ReconRun* run, int32_t* pending_dev) {
    run->release();
    if (B.ntiles <= 0) return NSX_OK;
    RoundFns fns;
    if (!select_fns(mode, dtype, conn, &fns))
        return ctx->fail(NSX_ERR_UNSUPPORTED, "reconstruction mode/dtype not instantiated by the reference");
    const RoundFn first = fns.first, later = fns.later;
    ReconRunImpl* R = new ReconRunImpl();
    R->mode = mode;
    R->dtype = dtype;
    R->conn = conn;
    R->batch = B;
    ReconArgs& A = R->A;
    A = ReconArgs{};
    A.nimg = B.nimg;
    A.ntiles = B.ntiles;
    A.xcd_split = B.xcd_split >= 0 ? (B.xcd_split ? 1 : 0) : (B.nimg == 1 ? 1 : 0);
    if (B.nimg == 1) A.img0 = *reinterpret_cast<const ImgDesc*>(B.host_desc);
    A.imgs = reinterpret_cast<const ImgDesc*>(B.dev_desc);
    A.tile_img = B.tile_img;
    void *stamp, *lists, *counts;
    nsx_status st = ctx->ws(B.ws_base + 0, sizeof(uint32_t) * (size_t)A.ntiles, &stamp);
    if (st == NSX_OK) st = ctx->ws(B.ws_base + 1, recon_list_bytes(A.ntiles), &lists);
    if (st == NSX_OK) st = ctx->ws(B.ws_base + 2, sizeof(int32_t) * (3 * NSEG * CSTRIDE + 64), &counts);
    if (st != NSX_OK) { delete R; return st; }
    A.stamp = (uint32_t*)stamp;
    A.list0 = (int32_t*)lists;
    A.list1 = (int32_t*)lists + (size_t)NSEG * A.ntiles;
    A.counts = (int32_t*)counts;
    A.stats = B.stats ? B.stats : ctx->d_stats;
    // the round log of the stats stage (0 or 1) these statistics belong to
    A.round_log = A.stats == ctx->d_stats       ? ctx->d_round_log
                  : A.stats == ctx->d_stats + 8 ? ctx->d_round_log + nsx_ctx::kRoundLog
                                                : nullptr;
    const int cum_idx = mode * 3 + dtype;
    A.cum = (unsigned long long*)(ctx->d_stats + nsx_ctx::kCumWord + cum_idx);
    // four-wave hole-filling tiles run one pair more: their pairs are shorter
    // (4096^2 chain: 36 fill rounds instead of 41, 0.8 % faster)
    const bool quad = (mode == RM_FILL || mode == RM_FILLBIN) && conn == 4;
    const int pairs = B.visit_sweeps > 0 ? B.visit_sweeps : ctx->visit_sweeps;
    A.visit_sweeps = std::max(1, std::min(MAX_IT, pairs + (quad ? 1 : 0)));
    R->later = later;
    // u8 tiles run 3 waves/SIMD (12 per CU), 32-bit tiles 2
    R->grid = std::max(1, std::min(A.ntiles, (dtype == NSX_U8 ? 12 : 8) * ctx->cu_count));
    // Round predictions are shared by problems within ~19% of each other in tile
    // count (quarter-octave buckets), so mixed request sizes still hit.
    R->key = RoundHints::key(A.visit_sweeps, A.ntiles, mode, dtype, conn);
    run->impl = R;

    // Hole filling starts from the minimum everywhere inside the frame, so in
    // round 0 only the tiles that touch the frame can change: the others are
    // written with the initial state by one fill kernel and join when activated.
    // A gray fill of a large single image instead starts with image-wide row and
    // column scans that carry the frame's value across the open parts; its
    // round 0 then visits every tile from the scanned state.
    const ImgDesc& D0 = A.img0;
    const bool scan_start = mode == RM_FILL && B.nimg == 1 && ctx->fill_scans > 0 &&
                            (dtype == NSX_U8 || dtype == NSX_I32) && D0.cols <= SCAN_MAX_COLS &&
                            std::min(D0.ntx, D0.nty) >= SCAN_MIN_TILES;
    const bool border_start = (mode == RM_FILL || mode == RM_FILLBIN) && B.nimg == 1 && !scan_start;
    uint32_t* segf = nullptr;
    if (scan_start) {
        void* p = nullptr;
        nsx_status sst = ctx->ws(WS_SCAN_SEG, sizeof(uint32_t) * scan_segf_words(D0.rows, D0.cols), &p);
        if (sst != NSX_OK) { run->release(); return sst; }
        segf = (uint32_t*)p;
    }
    A.first_list = border_start ? 1 : 0;
    StateFill F{};
    int grid_reset = std::max(1, std::min(1024, (A.ntiles + 255) / 256));
    if (border_start) {
        // the stored state of the initial key: invert(min) = max for both fill modes
        F = StateFill{D0.state, D0.ps, D0.rows, D0.cols, dtype == NSX_U8 ? 1 : 4,
                      dtype == NSX_U8 ? 255u : (uint32_t)INT_MAX};
        grid_reset = std::max(grid_reset, std::min(4 * ctx->cu_count, D0.rows));
    }
    hipLaunchKernelGGL(recon_reset_kernel, dim3(grid_reset), dim3(256), 0, s, A.stamp, A.ntiles, A.counts, A.stats,
                       A.list0, border_start ? D0.ntx : 0, D0.nty, F, scan_start ? 1 : 0, A.round_log);
    if (scan_start) {
        const int hs = ctx->tic(TK_RECON_SCAN, s);
        if (dtype == NSX_U8) run_scans<uint8_t, RM_FILL>(D0, ctx->fill_scans, segf, s);
        else run_scans<int32_t, RM_FILL>(D0, ctx->fill_scans, segf, s);
        ctx->toc(hs, s);
        const int h0 = ctx->tic(TK_RECON_ROUND, s);
        later(A, 0, R->grid, s);
        ctx->toc(h0, s);
    } else {
        const int h0 = ctx->tic(TK_RECON_FIRST, s);
        first(A, 0, border_start ? std::min(R->grid, border_tiles(D0.ntx, D0.nty)) : A.ntiles, s);
        ctx->first_tiles[cum_idx] += border_start ? border_tiles(D0.ntx, D0.nty) : A.ntiles;
        ctx->toc(h0, s);
    }
    const bool drain = (B.drain > 0 || (B.drain == 0 && ctx->drain > 0)) && fns.drain;
    if (drain) {
        // every later round in one persistent launch, up to the cap; arrivals
        // at its barriers stay below 2^31
        Drain P{};
        P.ctl = reinterpret_cast<uint32_t*>(A.counts) + CTL_OFF;
        P.max_round = B.drain_cap > 0 ? B.drain_cap : (1 << 24);
        P.max_round = std::min<int64_t>(P.max_round, (int64_t)INT32_MAX / (12 * (int64_t)ctx->cu_count) - 1);
        P.stall = DRAIN_STALL_TICKS;
        const int h = ctx->tic(TK_RECON_DRAIN, s);
        const int g = fns.drain(A, P, ctx->cu_count, B.drain_occ, s);
        ctx->toc(h, s);
        if (g > 0) {
            R->drain = true;
            R->next_round = P.max_round;
        }
    }
    if (!R->drain) {
        // Rounds the same-sized problem needed last time, launched without a host check.
        const int hint = ctx->round_hint.launch(R->key);
        for (int r = 1; r <= hint; ++r) {
            const int h = ctx->tic(TK_RECON_ROUND, s);
            later(A, r, R->grid, s);
            ctx->toc(h, s);
        }
        R->next_round = hint + 1;
    }
    NSX_HIP_TRY(ctx, hipGetLastError());
    if (pending_dev)   // a kernel rather than a copy node, so the sequence can be graph-captured
        hipLaunchKernelGGL(sum_counts_kernel, dim3(1), dim3(64), 0, s, pending_dev, &A.counts[cnt_idx(R->next_round, 0)],
                           R->drain ? reinterpret_cast<const uint32_t*>(A.counts) + CTL_OFF : nullptr);
    NSX_HIP_TRY(ctx, hipGetLastError());
    return NSX_OK;
}

nsx_status recon_complete(nsx_ctx* ctx, ReconRun* run, hipStream_t s, bool* extended, int max_rounds,
                          bool* gave_up) {
    if (extended) *extended = false;
    if (gave_up) *gave_up = false;
    ReconRunImpl* R = run->impl;
    if (!R) return NSX_OK;
    if (R->drain) {
        // -1: the drain gave up (a block never arrived); start again with
        // per-round launches.  Otherwise every counter is zero (converged) or
        // the rounds go on from the drain's cap as below.
        hipLaunchKernelGGL(sum_counts_kernel, dim3(1), dim3(64), 0, s, &R->A.counts[CNT_SUM],
                           &R->A.counts[cnt_idx(R->next_round, 0)],
                           reinterpret_cast<const uint32_t*>(R->A.counts) + CTL_OFF);
        NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned, &R->A.counts[CNT_SUM], sizeof(int32_t), hipMemcpyDeviceToHost, s));
        NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
        if (ctx->h_pinned[0] < 0) {
            ++ctx->drain_failures;
            ReconBatch B = R->batch;
            B.drain = -1;
            const int mode = R->mode, dtype = R->dtype, conn = R->conn;
            NSX_TRY(recon_launch(ctx, mode, dtype, conn, B, s, run, nullptr));
            R = run->impl;
            if (!R) return NSX_OK;
        }
    }
    const ReconArgs& A = R->A;
    int batch = 2;
    while (true) {
        hipLaunchKernelGGL(sum_counts_kernel, dim3(1), dim3(64), 0, s, &A.counts[CNT_SUM],
                           &A.counts[cnt_idx(R->next_round, 0)], (const uint32_t*)nullptr);
        NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned, &A.counts[CNT_SUM], sizeof(int32_t), hipMemcpyDeviceToHost, s));
        NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
        if (ctx->h_pinned[0] == 0) break;
        if (extended) *extended = true;
        if (max_rounds > 0 && R->next_round > max_rounds) {
            // the caller finishes by another method; the next call of this
            // size goes there directly
            note_hint(ctx, R->key, (int64_t)R->next_round);
            run->release();
            if (gave_up) *gave_up = true;
            return NSX_OK;
        }
        if (R->next_round > (1 << 24)) return ctx->fail(NSX_ERR_NOT_CONVERGED, "reconstruction did not converge");
        for (int j = 0; j < batch; ++j, ++R->next_round) {
            const int h = ctx->tic(TK_RECON_ROUND, s);
            R->later(A, R->next_round, R->grid, s);
            ctx->toc(h, s);
        }
        NSX_HIP_TRY(ctx, hipGetLastError());
        batch = std::min(batch * 2, 64);
    }
    // remember how many rounds this size needed (the last launched round had work
    // or was the first empty one); read the device round counter for precision
    int64_t st[2];
    NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned + 32, A.stats, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
    std::memcpy(st, ctx->h_pinned + 32, sizeof(st));
    note_hint(ctx, R->key, st[0]);
    run->release();
    return NSX_OK;
}

void recon_status_gather(const ReconRun* rr, const ReconRun* rf, const int32_t* ncomp, const int64_t* rstats,
                         const int64_t* fstats, int32_t* out, hipStream_t s) {
    const ReconRunImpl* a = rr->impl;
    const ReconRunImpl* b = rf->impl;
    hipLaunchKernelGGL(chain_status_kernel, dim3(1), dim3(64), 0, s, out, ncomp,
                       &a->A.counts[cnt_idx(a->next_round, 0)], &b->A.counts[cnt_idx(b->next_round, 0)], rstats, fstats,
                       a->drain ? reinterpret_cast<const uint32_t*>(a->A.counts) + CTL_OFF : nullptr,
                       b->drain ? reinterpret_cast<const uint32_t*>(b->A.counts) + CTL_OFF : nullptr);
}

void recon_note_rounds(nsx_ctx* ctx, ReconRun* run, int64_t rounds_with_work) {
    if (run->impl) note_hint(ctx, run->impl->key, rounds_with_work);
    run->release();
}

int recon_predicted_rounds(nsx_ctx* ctx, int mode, int dtype, int conn, int ntiles, int visit_sweeps) {
    const bool quad = (mode == RM_FILL || mode == RM_FILLBIN) && conn == 4;
    const int pairs = std::max(1, std::min(MAX_IT, (visit_sweeps > 0 ? visit_sweeps : ctx->visit_sweeps) + (quad ? 1 : 0)));
    return ctx->round_hint.predicted(RoundHints::key(pairs, ntiles, mode, dtype, conn));
}

// u8 reconstructions that take more propagation rounds than this are finished
// through the bit planes under NSX_ENGINE_AUTO (a round costs a tile visit's
// latency, ~10-25 us; the eight bit-plane passes a fixed ~0.8 ms at 4096^2).
constexpr int AUTO_TREE_ROUNDS = 256;

// A u8 reconstruction with a fixed number of launches: the component tree
// under NSX_ENGINE_TREE (*ok false when one of its bounds was hit), the bit
// planes otherwise.
static nsx_status recon_tree(nsx_ctx* ctx, int mode, const View* marker, const View& mask, const View& state,
                             int conn, hipStream_t s, bool* ok) {
    *ok = false;
    if (ctx->engine != NSX_ENGINE_TREE) {
        NSX_TRY(recon_bitplane_u8(ctx, mode, marker, mask, state, conn, s));
        *ok = true;
        return NSX_OK;
    }
    uint32_t* err = nullptr;
    NSX_TRY(mtree_launch(ctx, mode, marker, mask, state, conn, s, &err));
    if (!err) return NSX_OK;
    NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned + 56, err, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
    *ok = ctx->h_pinned[56] == 0;
    return NSX_OK;
}

nsx_status recon_run(nsx_ctx* ctx, int mode, const View* marker, const View& mask, const View& state,
                     int conn, hipStream_t s, LevelLabelReq* req, int64_t* stats) {
    if (mask.rows <= 0 || mask.cols <= 0) return NSX_OK;
    if (mode == RM_RECON && mask.dtype == NSX_U8 && marker && ctx->levelset_levels > 0) {
        bool used = false;
        NSX_TRY(recon_levelset_u8(ctx, *marker, mask, state, conn == 8 ? 8 : 4, ctx->levelset_levels, &used, s, req));
        if (used) return NSX_OK;
    }
    conn = conn == 8 ? 8 : 4;
    ImgDesc D = make_desc(marker, mask, state, 0);
    const bool tree_ok = (mode == RM_RECON || mode == RM_FILL) && mask.dtype == NSX_U8;
    const bool tree_first = tree_ok && (ctx->engine == NSX_ENGINE_TREE || ctx->engine == NSX_ENGINE_PLANES ||
                                        (ctx->engine == NSX_ENGINE_AUTO &&
                                         recon_predicted_rounds(ctx, mode, mask.dtype, conn, D.ntx * D.nty, 0) >
                                             AUTO_TREE_ROUNDS));
    if (tree_first) {
        // bit planes, or the component tree (propagation only when one of its
        // bounds was hit)
        bool ok = false;
        NSX_TRY(recon_tree(ctx, mode, marker, mask, state, conn, s, &ok));
        if (ok) return NSX_OK;
    }
    ReconBatch B{};
    B.nimg = 1;
    B.ntiles = D.ntx * D.nty;
    B.host_desc = &D;
    B.ws_base = WS_RECON_STAMP;
    B.stats = stats;
    ReconRun run;
    const int cap = tree_ok && !tree_first && ctx->engine == NSX_ENGINE_AUTO ? AUTO_TREE_ROUNDS : 0;
    if (cap) B.drain_cap = cap + 1;
    NSX_TRY(recon_launch(ctx, mode, mask.dtype, conn, B, s, &run, nullptr));
    bool gave_up = false;
    nsx_status st = recon_complete(ctx, &run, s, nullptr, cap, &gave_up);
    run.release();
    if (st != NSX_OK || !gave_up) return st;
    // many rounds: the content carries values far; the bit planes' cost does
    // not depend on that (they overwrite the rounds' partial state)
    bool ok = false;
    NSX_TRY(recon_tree(ctx, mode, marker, mask, state, conn, s, &ok));
    if (ok) return NSX_OK;
    NSX_TRY(recon_launch(ctx, mode, mask.dtype, conn, B, s, &run, nullptr));
    st = recon_complete(ctx, &run, s, nullptr);
    run.release();
    return st;
}

// ---- the allocation-free reconstruction ---------------------------------------
// nsx_imreconstruct_noalloc keeps its tile work lists, stamps, round counters and
// statistics in one file-scope device array instead of context workspace, so a
// call allocates no device memory.  Every call on a device shares that device's
// copy of the array; overlapping calls on one device race.  The array's device
// address is looked up once per device, with the device current, when the
// first context on it is created; host code uses only that address.
namespace {

constexpr int NOALLOC_MAX_TILES = 16384;
constexpr size_t NA_STAMP = 0;                                                    // u32 [tiles]
constexpr size_t NA_LISTS = NA_STAMP + sizeof(uint32_t) * NOALLOC_MAX_TILES;      // i32 [2][NSEG][tiles]
constexpr size_t NA_COUNTS = NA_LISTS + 2 * sizeof(int32_t) * NSEG * (size_t)NOALLOC_MAX_TILES;
constexpr size_t NA_STATS = NA_COUNTS + ((sizeof(int32_t) * (3 * NSEG * CSTRIDE + 64) + 255) & ~(size_t)255);
constexpr size_t NA_BYTES = NA_STATS + 256;                                       // int64 [4] + spare

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[NA_MAX_DEVICES] = {};

}  // namespace

size_t recon_noalloc_scratch_bytes() { return NA_BYTES; }
int recon_noalloc_max_tiles() { return NOALLOC_MAX_TILES; }

nsx_status recon_noalloc_init(int device) {
    if (device < 0 || device >= NA_MAX_DEVICES) return NSX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return NSX_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return NSX_ERR_HIP;
    }
    g_scratch_ptr[device] = ptr;
    return NSX_OK;
}

nsx_status recon_run_noalloc(nsx_ctx* ctx, const View& marker, const View& mask, const View& state, int conn,
                             hipStream_t s) {
    if (mask.rows <= 0 || mask.cols <= 0) return NSX_OK;
    char* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (ctx->device >= 0 && ctx->device < NA_MAX_DEVICES) base = (char*)g_scratch_ptr[ctx->device];
    }
    if (!base) return ctx->fail(NSX_ERR_HIP, "allocation-free scratch array has no device address on this device");
    ImgDesc D = make_desc(&marker, mask, state, 0);
    const int ntiles = D.ntx * D.nty;
    if ((int64_t)D.ntx * D.nty > NOALLOC_MAX_TILES)
        return ctx->fail(NSX_ERR_INVALID, "image too large for nsx_imreconstruct_noalloc: at most 16384 tiles of "
                                          "62 x 62 pixels (e.g. 7936 x 7936)");
    RoundFns fns;
    if (!select_fns(RM_RECON, mask.dtype, conn, &fns))
        return ctx->fail(NSX_ERR_UNSUPPORTED, "reconstruction dtype not supported");
    const RoundFn first = fns.first, later = fns.later;
    ReconArgs A{};
    A.img0 = D;
    A.nimg = 1;
    A.ntiles = ntiles;
    A.stamp = (uint32_t*)(base + NA_STAMP);
    A.list0 = (int32_t*)(base + NA_LISTS);
    A.list1 = A.list0 + (size_t)NSEG * ntiles;
    A.counts = (int32_t*)(base + NA_COUNTS);
    A.stats = (int64_t*)(base + NA_STATS);
    A.visit_sweeps = std::max(1, std::min(MAX_IT, ctx->visit_sweeps));
    A.first_list = 0;
    A.xcd_split = 1;
    const int grid = std::max(1, std::min(ntiles, (mask.dtype == NSX_U8 ? 12 : 8) * ctx->cu_count));
    // the stamps, counters and statistics start cleared on every call (a kernel,
    // not a memset, so the sequence can also be captured)
    hipLaunchKernelGGL(recon_reset_kernel, dim3(std::max(1, std::min(1024, (ntiles + 255) / 256))), dim3(256), 0, s,
                       A.stamp, A.ntiles, A.counts, A.stats, A.list0, 0, 0, StateFill{}, 0, (int32_t*)nullptr);
    first(A, 0, ntiles, s);
    // the round count of the last same-sized reconstruction on this context
    // (read only: the prediction table is not extended here)
    const uint64_t key = RoundHints::key(A.visit_sweeps, ntiles, RM_RECON, mask.dtype, conn);
    const bool known = ctx->round_hint.has(key);
    int next = 1;
    const int hint = ctx->round_hint.launch(key);
    for (; next <= hint; ++next) later(A, next, grid, s);
    NSX_HIP_TRY(ctx, hipGetLastError());
    int batch = 2;
    while (true) {
        hipLaunchKernelGGL(sum_counts_kernel, dim3(1), dim3(64), 0, s, &A.counts[CNT_SUM], &A.counts[cnt_idx(next, 0)],
                           (const uint32_t*)nullptr);
        NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned, &A.counts[CNT_SUM], sizeof(int32_t), hipMemcpyDeviceToHost, s));
        NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
        if (ctx->h_pinned[0] == 0) break;
        if (next > (1 << 24)) return ctx->fail(NSX_ERR_NOT_CONVERGED, "reconstruction did not converge");
        for (int j = 0; j < batch; ++j, ++next) later(A, next, grid, s);
        NSX_HIP_TRY(ctx, hipGetLastError());
        batch = std::min(batch * 2, 64);
    }
    if (known) {
        NSX_HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned + 32, A.stats, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        NSX_HIP_TRY(ctx, hipStreamSynchronize(s));
        int64_t rounds;
        std::memcpy(&rounds, ctx->h_pinned + 32, sizeof(rounds));
        ctx->round_hint.note_existing(key, rounds);
    }
    return NSX_OK;
}

}  // namespace nsx
