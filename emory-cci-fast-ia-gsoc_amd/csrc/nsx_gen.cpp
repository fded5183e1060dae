// nsx_gen.cpp -- seeded synthetic image generators for the morphology / CCL
// workloads (host code, part of libnscale_hip.so).
//
// The generators are frozen: tests, fixtures and benchmarks on any machine
// reproduce the same bytes from the same seed.
//
//   RNG      SplitMix64; uniform integer in [lo, hi] = lo + next() % (hi-lo+1).
//   disks    for each disk, in order: centre x in [0,cols-1], centre y in
//            [0,rows-1], radius in [5,25], intensity in [50,220]; pixels with
//            dx^2+dy^2 <= r^2 take max(current, intensity) over a 0
//            background.  Then every pixel, in raster order, adds noise in
//            [-8,8] and is clamped to [0,255].
//   marker   saturating mask - 30 (u8); the int32 form scales both by 257.
//   maze     perfect maze on a 6-px cell pitch, carved by an iterative
//            depth-first backtracker from cell (0,0): a cell's 3x3 core and the
//            3-px passages to carved neighbours are 200, everything else 20.
//            The marker is 0 except a 5x5 block of 200 at the entrance (0,0).
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    int uniform(int lo, int hi) { return lo + (int)(next() % (uint64_t)(hi - lo + 1)); }
};

}  // namespace

extern "C" {

uint64_t nsx_splitmix64_next(uint64_t* state) {
    SplitMix64 g(*state);
    uint64_t v = g.next();
    *state = g.s;
    return v;
}

// Disks + noise mask; `pitch` is in elements (>= cols).  Returns 0 on success.
int nsx_gen_disks_u8(uint8_t* mask, int rows, int cols, long pitch, int ndisks, uint64_t seed) {
    if (!mask || rows <= 0 || cols <= 0 || pitch < cols || ndisks < 0) return 1;
    for (int y = 0; y < rows; ++y) std::memset(mask + (size_t)y * pitch, 0, (size_t)cols);
    SplitMix64 g(seed);
    for (int d = 0; d < ndisks; ++d) {
        const int cx = g.uniform(0, cols - 1);
        const int cy = g.uniform(0, rows - 1);
        const int r = g.uniform(5, 25);
        const int v = g.uniform(50, 220);
        const int y0 = std::max(0, cy - r), y1 = std::min(rows - 1, cy + r);
        const int x0 = std::max(0, cx - r), x1 = std::min(cols - 1, cx + r);
        for (int y = y0; y <= y1; ++y) {
            uint8_t* row = mask + (size_t)y * pitch;
            const int dy = y - cy;
            for (int x = x0; x <= x1; ++x) {
                const int dx = x - cx;
                if (dx * dx + dy * dy <= r * r && row[x] < v) row[x] = (uint8_t)v;
            }
        }
    }
    for (int y = 0; y < rows; ++y) {
        uint8_t* row = mask + (size_t)y * pitch;
        for (int x = 0; x < cols; ++x) {
            int v = (int)row[x] + g.uniform(-8, 8);
            row[x] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
        }
    }
    return 0;
}

// marker = max(mask - sub, 0), elementwise over n contiguous bytes.
void nsx_gen_marker_u8(const uint8_t* mask, uint8_t* marker, long n, int sub) {
    for (long i = 0; i < n; ++i) {
        int v = (int)mask[i] - sub;
        marker[i] = (uint8_t)(v < 0 ? 0 : v);
    }
}

// int32 form of a u8 image: v * scale (used with scale 257 -> [0, 65535]).
void nsx_gen_scale_i32(const uint8_t* in, int32_t* out, long n, int scale) {
    for (long i = 0; i < n; ++i) out[i] = (int32_t)in[i] * scale;
}

// float32 form of a u8 image: v / 255.0f (exact order-isomorphism of u8).
void nsx_gen_scale_f32(const uint8_t* in, float* out, long n) {
    for (long i = 0; i < n; ++i) out[i] = (float)in[i] / 255.0f;
}

// Maze mask and marker (contiguous rows x cols).  Returns 0 on success.
int nsx_gen_maze_u8(uint8_t* mask, uint8_t* marker, int rows, int cols, uint64_t seed) {
    if (!mask || !marker || rows < 6 || cols < 6) return 1;
    const size_t n = (size_t)rows * cols;
    std::memset(mask, 20, n);
    std::memset(marker, 0, n);
    const int cw = cols / 6, ch = rows / 6;
    const size_t ncell = (size_t)cw * ch;
    std::vector<uint8_t> visited(ncell, 0);
    std::vector<int32_t> stack;
    stack.reserve(1024);
    SplitMix64 g(seed);
    auto carve_core = [&](int cy, int cx) {
        for (int y = 0; y < 3; ++y) std::memset(mask + (size_t)(6 * cy + y) * cols + 6 * cx, 200, 3);
    };
    auto carve_right = [&](int cy, int cx) {   // passage (cy,cx) -> (cy,cx+1)
        for (int y = 0; y < 3; ++y) std::memset(mask + (size_t)(6 * cy + y) * cols + 6 * cx + 3, 200, 3);
    };
    auto carve_down = [&](int cy, int cx) {    // passage (cy,cx) -> (cy+1,cx)
        for (int y = 3; y < 6; ++y) std::memset(mask + (size_t)(6 * cy + y) * cols + 6 * cx, 200, 3);
    };
    visited[0] = 1;
    carve_core(0, 0);
    stack.push_back(0);
    while (!stack.empty()) {
        const int c = stack.back();
        const int cy = c / cw, cx = c % cw;
        int cand[4], nc = 0;   // 0 up, 1 right, 2 down, 3 left
        if (cy > 0 && !visited[(size_t)(cy - 1) * cw + cx]) cand[nc++] = 0;
        if (cx + 1 < cw && !visited[(size_t)cy * cw + cx + 1]) cand[nc++] = 1;
        if (cy + 1 < ch && !visited[(size_t)(cy + 1) * cw + cx]) cand[nc++] = 2;
        if (cx > 0 && !visited[(size_t)cy * cw + cx - 1]) cand[nc++] = 3;
        if (nc == 0) { stack.pop_back(); continue; }
        const int dir = cand[g.next() % (uint64_t)nc];
        int ny = cy, nx = cx;
        switch (dir) {
            case 0: ny = cy - 1; carve_down(ny, nx); break;
            case 1: nx = cx + 1; carve_right(cy, cx); break;
            case 2: ny = cy + 1; carve_down(cy, cx); break;
            default: nx = cx - 1; carve_right(ny, nx); break;
        }
        const int nidx = ny * cw + nx;
        visited[nidx] = 1;
        carve_core(ny, nx);
        stack.push_back(nidx);
    }
    for (int y = 0; y < 5 && y < rows; ++y) std::memset(marker + (size_t)y * cols, 200, 5);
    return 0;
}

}  // extern "C"
