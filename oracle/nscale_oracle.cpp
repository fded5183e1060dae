// nscale_oracle.cpp -- CPU restatement of the reference's morphology / CCL path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is compiled into
// oracle/liboracle.so and loaded by tests/, by __graft_entry__.smoke() as the
// checker, and by bench.py's cpu_baseline leg.  The product library never links,
// loads or calls it, and nothing here is a fallback for the GPU path.
//
// Every function restates, loop for loop, the reference CPU code it cites
// (paths relative to fastIA-master/).  OpenCV's Mat is replaced by plain
// row-major buffers; copyMakeBorder(BORDER_CONSTANT, v) becomes an explicit
// (rows+2)x(cols+2) padded copy, and the returned ROI becomes a crop.
//
// Pinning (what checks each function, see tests/test_oracle.py):
//   or_imreconstruct_*      brute-force geodesic dilation to the fixed point on
//                           tiny images, the 32x32 flood-fill fixture of
//                           test/segment/cuda/globalQueueTest.cu:446-583,
//                           order-isomorphism u8 -> i32/f32, closed forms.
//   or_imreconstruct_binary brute-force BFS flood fill + marker-kept quirk.
//   or_fill_holes           MATLAB imfill 'holes' closed form (complement of
//                           the border-reachable set / reconstruction from frame).
//   or_imfill_u8            image | brute-force reconstruction of the seeded
//                           complement (gray), image | BFS from the seeds
//                           through the complement (binary).
//   or_bwselect_u8          the components of {image != 0} holding a seed
//                           (scipy ndimage.label), with the image's values.
//   or_imhmin_*             invert(brute-force reconstruction of
//                           (invert(image) - h) under invert(image), 0 frame).
//   or_local_maxima/minima  candidates from the brute-force reconstruction,
//                           minus the equal-valued candidate plateaus that touch
//                           a higher neighbour (scipy); the flood follows
//                           OpenCV's documented floodFill (fixed range, mask
//                           only, mask value 1): parity pinned to that document.
//   or_bounding_box         the reference's own ConnComponents::boundingBox
//                           compiled into oracle/_ref (exact).
//   float invert's zero     parity unpinned: -(+0.0) is taken as +0.0 (see
//                           invert_one(float)); values are equal either way.
//   or_bwlabel / relabel /  the reference's own ConnComponents.cpp compiled
//   or_bwareaopen           unmodified into oracle/_ref (exact), scipy
//                           ndimage.label partitions, and the canonical-numbering
//                           rule.
//   i32 invert at INT_MIN   parity unpinned (OpenCV 2.x unavailable): both
//                           readings are restated (wrap, the default, and
//                           saturate); each one's int32 imfillHoles closed form
//                           pins it (constant INT_MIN under wrap; the textbook
//                           fill, or constant INT_MAX for binary, under
//                           saturation).  See invert_i32.
#include <cstdint>
#include <cstring>
#include <climits>
#include <cfloat>
#include <limits>
#include <cmath>
#include <deque>
#include <queue>
#include <vector>
#include <algorithm>
#include <unordered_map>

namespace {

// cci::common::type::min<T> (src/common/TypeUtils.h:33-41): integer min, or
// -max for floating point.
template <typename T> inline T type_min() {
    if constexpr (std::numeric_limits<T>::is_integer) return std::numeric_limits<T>::min();
    else return -std::numeric_limits<T>::max();
}

// copyMakeBorder(src, dst, 1,1,1,1, BORDER_CONSTANT, v)
template <typename T>
std::vector<T> pad1(const T* src, int rows, int cols, T v) {
    const int W = cols + 2, H = rows + 2;
    std::vector<T> dst((size_t)W * H, v);
    for (int y = 0; y < rows; ++y)
        std::memcpy(&dst[(size_t)(y + 1) * W + 1], src + (size_t)y * cols, sizeof(T) * cols);
    return dst;
}

template <typename T>
void crop1(const std::vector<T>& src, int rows, int cols, T* dst) {
    const int W = cols + 2;
    for (int y = 0; y < rows; ++y)
        std::memcpy(dst + (size_t)y * cols, &src[(size_t)(y + 1) * W + 1], sizeof(T) * cols);
}

// propagate<T> (MorphologicOperations.cpp:428-439)
template <typename T>
inline void propagate(std::queue<int>& xQ, std::queue<int>& yQ, int x, int y,
                      const T* iPtr, T* oPtr, const T& pval) {
    T qval = oPtr[x];
    T ival = iPtr[x];
    if ((qval < pval) && (ival != qval)) {
        oPtr[x] = std::min(pval, ival);
        xQ.push(x);
        yQ.push(y);
    }
}

// imreconstruct<T> (MorphologicOperations.cpp:665-838): Vincent's hybrid
// raster / anti-raster scan followed by a FIFO drain, on a 0-framed copy.
// Operates on padded buffers; returns the padded output.
template <typename T>
void imreconstruct_padded(std::vector<T>& output, const std::vector<T>& input,
                          int prow, int pcol, int connectivity) {
    const int W = pcol;
    T pval, preval;
    int xminus, xplus, yminus, yplus;
    const int maxx = pcol - 1;
    const int maxy = prow - 1;
    std::queue<int> xQ, yQ;
    T *oPtr, *oPtrMinus, *oPtrPlus;
    const T *iPtr, *iPtrPlus, *iPtrMinus;
    auto orow = [&](int y) { return output.data() + (size_t)y * W; };
    auto irow = [&](int y) { return input.data() + (size_t)y * W; };

    // raster scan (:691-713)
    for (int y = 1; y < maxy; ++y) {
        oPtr = orow(y);
        oPtrMinus = orow(y - 1);
        iPtr = irow(y);
        preval = oPtr[0];
        for (int x = 1; x < maxx; ++x) {
            xminus = x - 1;
            xplus = x + 1;
            pval = oPtr[x];
            pval = std::max(pval, std::max(preval, oPtrMinus[x]));
            if (connectivity == 8) pval = std::max(pval, std::max(oPtrMinus[xplus], oPtrMinus[xminus]));
            preval = std::min(pval, iPtr[x]);
            oPtr[x] = preval;
        }
    }

    // anti-raster scan with seed capture (:715-763)
    for (int y = maxy - 1; y > 0; --y) {
        oPtr = orow(y);
        oPtrPlus = orow(y + 1);
        oPtrMinus = orow(y - 1);
        iPtr = irow(y);
        iPtrPlus = irow(y + 1);
        preval = oPtr[maxx];
        for (int x = maxx - 1; x > 0; --x) {
            xminus = x - 1;
            xplus = x + 1;
            pval = oPtr[x];
            pval = std::max(pval, std::max(preval, oPtrPlus[x]));
            if (connectivity == 8) pval = std::max(pval, std::max(oPtrPlus[xplus], oPtrPlus[xminus]));
            preval = std::min(pval, iPtr[x]);
            oPtr[x] = preval;

            pval = oPtr[x];
            if ((oPtr[xplus] < std::min(pval, iPtr[xplus])) ||
                (oPtrPlus[x] < std::min(pval, iPtrPlus[x]))) {
                xQ.push(x);
                yQ.push(y);
                continue;
            }
            if (connectivity == 8) {
                if ((oPtrPlus[xplus] < std::min(pval, iPtrPlus[xplus])) ||
                    (oPtrPlus[xminus] < std::min(pval, iPtrPlus[xminus]))) {
                    xQ.push(x);
                    yQ.push(y);
                    continue;
                }
            }
        }
    }

    // FIFO drain (:772-828)
    int x, y;
    while (!xQ.empty()) {
        x = xQ.front();
        y = yQ.front();
        xQ.pop();
        yQ.pop();
        xminus = x - 1;
        xplus = x + 1;
        yminus = y - 1;
        yplus = y + 1;
        oPtr = orow(y);
        oPtrPlus = orow(yplus);
        oPtrMinus = orow(yminus);
        iPtr = irow(y);
        iPtrPlus = irow(yplus);
        iPtrMinus = irow(yminus);
        pval = oPtr[x];
        if (y > 0) propagate<T>(xQ, yQ, x, yminus, iPtrMinus, oPtrMinus, pval);
        if (y < maxy) propagate<T>(xQ, yQ, x, yplus, iPtrPlus, oPtrPlus, pval);
        if (x > 0) propagate<T>(xQ, yQ, xminus, y, iPtr, oPtr, pval);
        if (x < maxx) propagate<T>(xQ, yQ, xplus, y, iPtr, oPtr, pval);
        if (connectivity == 8) {
            if (y > 0) {
                if (x > 0) propagate<T>(xQ, yQ, xminus, yminus, iPtrMinus, oPtrMinus, pval);
                if (x < maxx) propagate<T>(xQ, yQ, xplus, yminus, iPtrMinus, oPtrMinus, pval);
            }
            if (y < maxy) {
                if (x > 0) propagate<T>(xQ, yQ, xminus, yplus, iPtrPlus, oPtrPlus, pval);
                if (x < maxx) propagate<T>(xQ, yQ, xplus, yplus, iPtrPlus, oPtrPlus, pval);
            }
        }
    }
}

template <typename T>
void imreconstruct(const T* seeds, const T* image, T* out, int rows, int cols, int connectivity) {
    std::vector<T> output = pad1<T>(seeds, rows, cols, T(0));   // :671-672
    std::vector<T> input = pad1<T>(image, rows, cols, T(0));    // :673-674
    imreconstruct_padded<T>(output, input, rows + 2, cols + 2, connectivity);
    crop1<T>(output, rows, cols, out);                          // :836
}

// propagateBinary<T> (MorphologicOperations.cpp:1665-1672)
template <typename T>
inline void propagateBinary(std::queue<int>& xQ, std::queue<int>& yQ, int x, int y,
                            const T* iPtr, T* oPtr, const T& foreground) {
    if ((oPtr[x] == 0) && (iPtr[x] != 0)) {
        oPtr[x] = foreground;
        xQ.push(x);
        yQ.push(y);
    }
}

// imreconstructBinary<T> (MorphologicOperations.cpp:1688-1829), padded form.
template <typename T>
void imreconstructBinary_padded(std::vector<T>& output, const std::vector<T>& input,
                                int prow, int pcol, int connectivity) {
    const int W = pcol;
    T pval, ival;
    int xminus, xplus, yminus, yplus;
    const int maxx = pcol - 1;
    const int maxy = prow - 1;
    std::queue<int> xQ, yQ;
    T *oPtr, *oPtrPlus, *oPtrMinus;
    const T *iPtr, *iPtrPlus, *iPtrMinus;
    auto orow = [&](int y) { return output.data() + (size_t)y * W; };
    auto irow = [&](int y) { return input.data() + (size_t)y * W; };

    // contour pixels (:1712-1756)
    for (int y = 1; y < maxy; ++y) {
        oPtr = orow(y);
        oPtrPlus = orow(y + 1);
        oPtrMinus = orow(y - 1);
        iPtr = irow(y);
        for (int x = 1; x < maxx; ++x) {
            pval = oPtr[x];
            ival = iPtr[x];
            if (pval != 0 && ival != 0) {
                xminus = x - 1;
                xplus = x + 1;
                if ((oPtrMinus[x] == 0) || (oPtrPlus[x] == 0) || (oPtr[xplus] == 0) || (oPtr[xminus] == 0)) {
                    xQ.push(x);
                    yQ.push(y);
                    continue;
                }
                if (connectivity == 8) {
                    if ((oPtrMinus[xminus] == 0) || (oPtrMinus[xplus] == 0) ||
                        (oPtrPlus[xminus] == 0) || (oPtrPlus[xplus] == 0)) {
                        xQ.push(x);
                        yQ.push(y);
                        continue;
                    }
                }
            }
        }
    }

    // BFS (:1763-1823)
    const T outval = std::numeric_limits<T>::max();
    int x, y;
    while (!xQ.empty()) {
        x = xQ.front();
        y = yQ.front();
        xQ.pop();
        yQ.pop();
        xminus = x - 1;
        yminus = y - 1;
        yplus = y + 1;
        xplus = x + 1;
        oPtr = orow(y);
        oPtrMinus = orow(y - 1);
        oPtrPlus = orow(y + 1);
        iPtr = irow(y);
        iPtrMinus = irow(y - 1);
        iPtrPlus = irow(y + 1);
        if (y > 0) propagateBinary<T>(xQ, yQ, x, yminus, iPtrMinus, oPtrMinus, outval);
        if (y < maxy) propagateBinary<T>(xQ, yQ, x, yplus, iPtrPlus, oPtrPlus, outval);
        if (x > 0) propagateBinary<T>(xQ, yQ, xminus, y, iPtr, oPtr, outval);
        if (x < maxx) propagateBinary<T>(xQ, yQ, xplus, y, iPtr, oPtr, outval);
        if (connectivity == 8) {
            if (y > 0) {
                if (x > 0) propagateBinary<T>(xQ, yQ, xminus, yminus, iPtrMinus, oPtrMinus, outval);
                if (x < maxx) propagateBinary<T>(xQ, yQ, xplus, yminus, iPtrMinus, oPtrMinus, outval);
            }
            if (y < maxy) {
                if (x > 0) propagateBinary<T>(xQ, yQ, xminus, yplus, iPtrPlus, oPtrPlus, outval);
                if (x < maxx) propagateBinary<T>(xQ, yQ, xplus, yplus, iPtrPlus, oPtrPlus, outval);
            }
        }
    }
}

template <typename T>
void imreconstructBinary(const T* seeds, const T* image, T* out, int rows, int cols, int connectivity) {
    std::vector<T> output = pad1<T>(seeds, rows, cols, T(0));   // :1694
    std::vector<T> input = pad1<T>(image, rows, cols, T(0));    // :1696
    imreconstructBinary_padded<T>(output, input, rows + 2, cols + 2, connectivity);
    crop1<T>(output, rows, cols, out);                          // :1827
}

// PixelOperations::invert<T> (PixelOperations.cpp:17-39).
//   unsigned: max - x ;  float: -x ;  signed int: bitwise_not(x) + 1.
// The signed case goes through OpenCV Mat+Scalar on CV_32S, whose behaviour at
// x == INT_MIN cannot be run here (OpenCV is absent).  The reference's own GPU
// kernel computes ~x + 1 in int and wraps (src/segment/cuda/pixel-ops.cu:
// 39-48, instantiated "for imfillholes" at PixelOperationsGPU.cpp:301), and
// OpenCV documents no saturation for CV_32S results, so wrapping is the
// default; saturation (INT_MIN -> INT_MAX) is the alternative reading.  All
// other inputs are exact two's-complement negation either way.
thread_local bool g_i32_saturate = false;   // set per call by the exported functions
inline int32_t invert_i32(int32_t v, bool saturate) {
    if (v == INT_MIN) return saturate ? INT_MAX : INT_MIN;
    return -v;
}
inline uint8_t invert_one(uint8_t v) { return (uint8_t)(255 - v); }
inline int32_t invert_one(int32_t v) { return invert_i32(v, g_i32_saturate); }
// OpenCV's unary minus on a Mat is a scale by -1 plus a shift of 0, which
// turns -0.0 into +0.0; 0.0f - v gives the same bits (OpenCV is absent here, so
// the sign of a zero result is unpinned; the value is exact).
inline float invert_one(float v) { return 0.0f - v; }

template <typename T>
void invert_vec(std::vector<T>& v) {
    for (auto& e : v) e = invert_one(e);
}

// imfillHoles<T> (MorphologicOperations.cpp:1859-1926).
template <typename T>
void imfillHoles(const T* image, T* out, int rows, int cols, bool binary, int connectivity) {
    const T mn = type_min<T>();                                  // :1889
    const T mx = std::numeric_limits<T>::max();                  // :1890
    std::vector<T> mask = pad1<T>(image, rows, cols, mn);         // :1894-1895
    std::vector<T> marker2((size_t)rows * cols, mn);             // :1898
    std::vector<T> marker = pad1<T>(marker2.data(), rows, cols, mx);  // :1900
    invert_vec(mask);                                            // :1903
    // imreconstruct / imreconstructBinary pad again with a 0 frame.
    const int R = rows + 2, C = cols + 2;
    std::vector<T> output;
    output.resize((size_t)R * C);
    if (binary) imreconstructBinary<T>(marker.data(), mask.data(), output.data(), R, C, connectivity);
    else imreconstruct<T>(marker.data(), mask.data(), output.data(), R, C, connectivity);
    invert_vec(output);                                          // :1923
    crop1<T>(output, rows, cols, out);                           // :1925
}

// imfill<T> with seeds (MorphologicOperations.cpp:1833-1857, instantiated for
// uchar at :2516): mask = invert(image) (:1848); marker = mask where seeds != 0,
// else 0 (:1850-1852, Mat::copyTo with the seeds as copy mask); marker =
// imreconstructBinary or imreconstruct(marker, mask, conn) (:1854-1855);
// result = image | marker (:1857).
template <typename T>
void imfill(const T* image, const uint8_t* seeds, T* out, int rows, int cols, bool binary, int connectivity) {
    const size_t n = (size_t)rows * cols;
    std::vector<T> mask(image, image + n);
    invert_vec(mask);
    std::vector<T> marker(n, T(0));
    for (size_t i = 0; i < n; ++i)
        if (seeds[i] != 0) marker[i] = mask[i];
    std::vector<T> rec(n);
    if (binary) imreconstructBinary<T>(marker.data(), mask.data(), rec.data(), rows, cols, connectivity);
    else imreconstruct<T>(marker.data(), mask.data(), rec.data(), rows, cols, connectivity);
    for (size_t i = 0; i < n; ++i) out[i] = (T)(image[i] | rec[i]);
}

// bwselect<T> (MorphologicOperations.cpp:1929-1951, instantiated for uchar at
// :2522): marker = binaryImage where seeds != 0, else 0 (:1945-1946);
// imreconstructBinary(marker, binaryImage, conn) (:1948); & binaryImage (:1950).
template <typename T>
void bwselect(const T* image, const uint8_t* seeds, T* out, int rows, int cols, int connectivity) {
    const size_t n = (size_t)rows * cols;
    std::vector<T> marker(n, T(0));
    for (size_t i = 0; i < n; ++i)
        if (seeds[i] != 0) marker[i] = image[i];
    std::vector<T> rec(n);
    imreconstructBinary<T>(marker.data(), image, rec.data(), rows, cols, connectivity);
    for (size_t i = 0; i < n; ++i) out[i] = (T)(rec[i] & image[i]);
}

// imhmin<T> (MorphologicOperations.cpp:2260-2282, instantiated for uchar and
// float at :2527-2528): mask = invert(image) (:2273); marker = mask - h
// (:2274, OpenCV Mat - scalar: saturating for uchar); output =
// invert(imreconstruct(marker, mask, conn)) (:2279-2280).
inline uint8_t sub_h(uint8_t v, uint8_t h) { return v > h ? (uint8_t)(v - h) : (uint8_t)0; }
inline float sub_h(float v, float h) { return v - h; }
template <typename T>
void imhmin(const T* image, T h, T* out, int rows, int cols, int connectivity) {
    const size_t n = (size_t)rows * cols;
    std::vector<T> mask(image, image + n);
    invert_vec(mask);
    std::vector<T> marker(n);
    for (size_t i = 0; i < n; ++i) marker[i] = sub_h(mask[i], h);
    imreconstruct<T>(marker.data(), mask.data(), out, rows, cols, connectivity);
    for (size_t i = 0; i < n; ++i) out[i] = invert_one(out[i]);
}

// localMaxima<T> (MorphologicOperations.cpp:2374-2446; uchar and float,
// :2529-2532).  marker = image - 1 (uchar saturating) (:2379); candidates =
// !(marker < imreconstruct(marker, image)) (:2380-2381: 0 marks a candidate);
// padded with 255 (mask) and type min (image) (:2392-2395); every candidate
// with a strictly higher neighbour (4, then the diagonals for 8) floods its
// equal-valued region of candidates in the mask (:2425-2440) -- cv::floodFill
// with FLOODFILL_FIXED_RANGE | FLOODFILL_MASK_ONLY, zero differences: pixels
// equal to the seed, connected through zero mask pixels, get mask value 1;
// result = 255 where the mask stayed 0 (:2446).
template <typename T>
void localMaxima(const T* image, uint8_t* out, int rows, int cols, int connectivity) {
    const size_t n = (size_t)rows * cols;
    std::vector<T> marker(n), rec(n);
    for (size_t i = 0; i < n; ++i) marker[i] = sub_h(image[i], T(1));
    imreconstruct<T>(marker.data(), image, rec.data(), rows, cols, connectivity);
    const int R = rows + 2, C = cols + 2;
    std::vector<uint8_t> mask((size_t)R * C, 255);
    std::vector<T> in((size_t)R * C, type_min<T>());
    for (int y = 0; y < rows; ++y)
        for (int x = 0; x < cols; ++x) {
            const size_t i = (size_t)y * cols + x;
            mask[(size_t)(y + 1) * C + x + 1] = marker[i] < rec[i] ? 255 : 0;
            in[(size_t)(y + 1) * C + x + 1] = image[i];
        }
    static const int dy[8] = {-1, 1, 0, 0, -1, -1, 1, 1}, dx[8] = {0, 0, -1, 1, -1, 1, -1, 1};
    const int nn = connectivity == 8 ? 8 : 4;
    std::vector<int> stack;
    for (int y = 1; y < R - 1; ++y)
        for (int x = 1; x < C - 1; ++x) {
            if (mask[(size_t)y * C + x] > 0) continue;
            const T v = in[(size_t)y * C + x];
            bool higher = false;
            for (int k = 0; k < nn && !higher; ++k) higher = v < in[(size_t)(y + dy[k]) * C + x + dx[k]];
            if (!higher) continue;
            // floodFill from (x, y): equal value, through zero mask pixels
            stack.assign(1, y * C + x);
            mask[(size_t)y * C + x] = 1;
            while (!stack.empty()) {
                const int p = stack.back();
                stack.pop_back();
                const int py = p / C, px = p % C;
                for (int k = 0; k < nn; ++k) {
                    const int q = (py + dy[k]) * C + px + dx[k];
                    if (mask[q] == 0 && in[q] == v) {
                        mask[q] = 1;
                        stack.push_back(q);
                    }
                }
            }
        }
    for (int y = 0; y < rows; ++y)
        for (int x = 0; x < cols; ++x) out[(size_t)y * cols + x] = mask[(size_t)(y + 1) * C + x + 1] == 0 ? 255 : 0;
}

// ---- ConnComponents (src/segment/ConnComponents.cpp) ----------------------
// find/flatten are written iteratively: the recursive originals (:24-32,
// :137-147) compress every node on the path to the root, which the two-pass
// loops below reproduce exactly, without deep recursion on very large images.
inline int cc_find(int* label, int x) {
    int r = x;
    while (label[r] != r) r = label[r];
    while (label[x] != r) { int n = label[x]; label[x] = r; x = n; }
    return r;
}

inline void cc_merge(int* label, int x, int y) {   // :34-43
    x = cc_find(label, x);
    y = cc_find(label, y);
    if (label[x] < label[y]) label[y] = x;
    else if (label[x] > label[y]) label[x] = y;
}

inline int cc_flatten(int* label, int x, int bgval) {   // :137-147
    if (label[x] == bgval) return bgval;
    int r = x;
    while (label[r] != r) {
        r = label[r];
        if (r == bgval) break;   // not reachable for well-formed forests
    }
    while (x != r && label[x] != r) { int n = label[x]; label[x] = r; x = n; }
    return r;
}

void cc_label(const uint8_t* img, int w, int h, int* label, int bgval, int connectivity) {  // :52-86
    const long length = (long)w * h;
    for (long i = 0; i < length; ++i) label[i] = (int)i;
    long i = -1, imw = 0;
    uint8_t p;
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            ++i;
            p = img[i];
            if (p == 0) {
                label[i] = bgval;
            } else {
                imw = i - w;
                if (x > 0 && p == img[i - 1]) cc_merge(label, (int)i, (int)(i - 1));
                if (y > 0 && p == img[imw]) cc_merge(label, (int)i, (int)imw);
                if (connectivity == 8) {
                    if (y > 0 && x > 0 && p == img[imw - 1]) cc_merge(label, (int)i, (int)(imw - 1));
                    if (y > 0 && x < w - 1 && p == img[imw + 1]) cc_merge(label, (int)i, (int)(imw + 1));
                }
            }
        }
    }
    for (long k = 0; k < length; ++k) label[k] = cc_flatten(label, (int)k, bgval);
}

// relabel (:89-118): roots (label[i]==i) numbered 1.. in raster order, bg -> 0.
int cc_relabel(int w, int h, int* label, int bgval) {
    const long length = (long)w * h;
    std::unordered_map<int, int> labelmap;
    int j = 1;
    labelmap[bgval] = 0;
    for (long i = 0; i < length; ++i) {
        if (label[i] == i) {
            labelmap[(int)i] = j;
            ++j;
        }
    }
    for (long i = 0; i < length; ++i) label[i] = labelmap[label[i]];
    return j - 1;
}

// areaThresholdLabeled (:205-239): keep labels whose pixel count is in
// [lower, upper); everything else becomes bgval.  Returns the kept count.
int cc_area_threshold_labeled(const int* label, int w, int h, int* n_label, int bgval, int lower, int upper) {
    const long length = (long)w * h;
    typedef std::unordered_map<int, int> AreaMap;
    AreaMap areas, thresholdedareas;
    for (long i = 0; i < length; ++i) {
        if (label[i] != bgval) areas[label[i]] += 1;
        n_label[i] = bgval;
    }
    int j = 0;
    for (AreaMap::iterator it = areas.begin(); it != areas.end(); ++it) {
        if (it->second >= lower && it->second < upper) {
            thresholdedareas[it->first] = 1;
            ++j;
        } else {
            thresholdedareas[it->first] = 0;
        }
    }
    areas.clear();
    for (long i = 0; i < length; ++i) {
        if (label[i] != bgval && thresholdedareas[label[i]] == 1) n_label[i] = label[i];
    }
    return j;
}

}  // namespace

extern "C" {

int or_imreconstruct_u8(const uint8_t* marker, const uint8_t* mask, uint8_t* out, int rows, int cols, int conn) {
    imreconstruct<uint8_t>(marker, mask, out, rows, cols, conn);
    return 0;
}
int or_imreconstruct_i32(const int32_t* marker, const int32_t* mask, int32_t* out, int rows, int cols, int conn) {
    imreconstruct<int32_t>(marker, mask, out, rows, cols, conn);
    return 0;
}
int or_imreconstruct_f32(const float* marker, const float* mask, float* out, int rows, int cols, int conn) {
    imreconstruct<float>(marker, mask, out, rows, cols, conn);
    return 0;
}
int or_imreconstruct_binary_u8(const uint8_t* marker, const uint8_t* mask, uint8_t* out, int rows, int cols, int conn) {
    imreconstructBinary<uint8_t>(marker, mask, out, rows, cols, conn);
    return 0;
}
int or_imreconstruct_binary_i32(const int32_t* marker, const int32_t* mask, int32_t* out, int rows, int cols, int conn) {
    imreconstructBinary<int32_t>(marker, mask, out, rows, cols, conn);
    return 0;
}
int or_fill_holes_u8(const uint8_t* in, uint8_t* out, int rows, int cols, int binary, int conn) {
    imfillHoles<uint8_t>(in, out, rows, cols, binary != 0, conn);
    return 0;
}
int or_fill_holes_i32(const int32_t* in, int32_t* out, int rows, int cols, int binary, int conn, int saturate) {
    g_i32_saturate = saturate != 0;
    imfillHoles<int32_t>(in, out, rows, cols, binary != 0, conn);
    g_i32_saturate = false;
    return 0;
}
int or_imfill_u8(const uint8_t* in, const uint8_t* seeds, uint8_t* out, int rows, int cols, int binary, int conn) {
    imfill<uint8_t>(in, seeds, out, rows, cols, binary != 0, conn);
    return 0;
}
int or_bwselect_u8(const uint8_t* in, const uint8_t* seeds, uint8_t* out, int rows, int cols, int conn) {
    bwselect<uint8_t>(in, seeds, out, rows, cols, conn);
    return 0;
}
// ConnComponents::boundingBox (ConnComponents.cpp:241-284): for each label !=
// bgval, in order of first appearance in raster order, its min x, max x, min y,
// max y; a label of -1 or 0 that is not bgval ends the scan (:251-252).
// out: labels, min x, max x, min y, max y (count each), at most capacity ints.
int or_bounding_box(const int32_t* label, int w, int h, int bgval, int32_t* out, int capacity) {
    const long length = (long)w * h;
    std::unordered_map<int32_t, int> index;
    std::vector<int32_t> lab, mnx, mxx, mny, mxy;
    for (long i = 0; i < length; ++i) {
        const int32_t v = label[i];
        if (v == bgval) continue;
        if (v == -1 || v == 0) break;
        const int x = (int)(i % w), y = (int)(i / w);
        auto it = index.find(v);
        if (it == index.end()) {
            index[v] = (int)lab.size();
            lab.push_back(v); mnx.push_back(x); mxx.push_back(x); mny.push_back(y); mxy.push_back(y);
        } else {
            const int j = it->second;
            mnx[j] = std::min(mnx[j], x); mxx[j] = std::max(mxx[j], x);
            mny[j] = std::min(mny[j], y); mxy[j] = std::max(mxy[j], y);
        }
    }
    const int n = (int)lab.size();
    const std::vector<int32_t>* cols[5] = {&lab, &mnx, &mxx, &mny, &mxy};
    for (int c = 0; c < 5; ++c)
        for (int j = 0; j < n && c * n + j < capacity; ++j) out[c * n + j] = (*cols[c])[j];
    return n;
}

// distanceTransform (MorphologicOperations.cpp:197-275).  Every pixel starts at
// the index rows*cols*3 (:214), a point (0, 3*rows) below the image; a background
// pixel holds its own index (:218) and enters a FIFO queue, in raster order,
// when any 8-neighbour is nonzero (checkDistNeighbors8, :33-88).  A popped pixel
// offers its index to its 8 neighbours in the order up-left, up-right, up,
// down-left, down-right, down, left, right (propagateDist8, :122-193); a
// neighbour takes it when the distance to it is strictly below the distance to
// its own (propagateDist, :90-119: sqrt in double, compared after rounding to
// float) and is queued again.  out_nn gets the indices; out_dist (optional) the
// float distances to them (:254-266).
int or_distance_transform(const uint8_t* mask, int rows, int cols, int32_t* out_nn, float* out_dist) {
    const int64_t n = (int64_t)rows * cols;
    const int32_t far = (int32_t)(n * 3);
    std::deque<int64_t> q;
    for (int y = 0; y < rows; ++y)
        for (int x = 0; x < cols; ++x) {
            const int64_t i = (int64_t)y * cols + x;
            out_nn[i] = far;
            if (mask[i] != 0) continue;
            out_nn[i] = (int32_t)i;
            bool front = false;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int qy = y + dy, qx = x + dx;
                    if ((dy || dx) && qy >= 0 && qy < rows && qx >= 0 && qx < cols && mask[(int64_t)qy * cols + qx])
                        front = true;
                }
            if (front) q.push_back(i);
        }
    auto dist = [&](int x, int y, int32_t label) {
        const int lx = label % cols, ly = label / cols;
        const double ddx = (double)(x - lx), ddy = (double)(y - ly);
        return (float)std::sqrt(ddx * ddx + ddy * ddy);
    };
    static const int order[8][2] = {{-1, -1}, {1, -1}, {0, -1}, {-1, 1}, {1, 1}, {0, 1}, {-1, 0}, {1, 0}};
    while (!q.empty()) {
        const int64_t i = q.front();
        q.pop_front();
        const int y = (int)(i / cols), x = (int)(i % cols);
        const int32_t cur = out_nn[i];
        for (const auto& o : order) {
            const int tx = x + o[0], ty = y + o[1];
            if (tx < 0 || tx >= cols || ty < 0 || ty >= rows) continue;
            const int64_t t = (int64_t)ty * cols + tx;
            if (dist(tx, ty, cur) < dist(tx, ty, out_nn[t])) {
                out_nn[t] = cur;
                q.push_back(t);
            }
        }
    }
    if (out_dist)
        for (int64_t i = 0; i < n; ++i) out_dist[i] = dist((int)(i % cols), (int)(i / cols), out_nn[i]);
    return 0;
}

int or_local_maxima_u8(const uint8_t* in, uint8_t* out, int rows, int cols, int conn) {
    localMaxima<uint8_t>(in, out, rows, cols, conn);
    return 0;
}
int or_local_maxima_f32(const float* in, uint8_t* out, int rows, int cols, int conn) {
    localMaxima<float>(in, out, rows, cols, conn);
    return 0;
}
// localMinima<T> (:2449-2456): localMaxima of invert(image).
int or_local_minima_u8(const uint8_t* in, uint8_t* out, int rows, int cols, int conn) {
    std::vector<uint8_t> c(in, in + (size_t)rows * cols);
    invert_vec(c);
    localMaxima<uint8_t>(c.data(), out, rows, cols, conn);
    return 0;
}
int or_local_minima_f32(const float* in, uint8_t* out, int rows, int cols, int conn) {
    std::vector<float> c(in, in + (size_t)rows * cols);
    invert_vec(c);
    localMaxima<float>(c.data(), out, rows, cols, conn);
    return 0;
}

int or_imhmin_u8(const uint8_t* in, int h, uint8_t* out, int rows, int cols, int conn) {
    imhmin<uint8_t>(in, (uint8_t)h, out, rows, cols, conn);
    return 0;
}
int or_imhmin_f32(const float* in, float h, float* out, int rows, int cols, int conn) {
    imhmin<float>(in, h, out, rows, cols, conn);
    return 0;
}
void or_invert_i32(const int32_t* in, int32_t* out, long n, int saturate) {
    for (long i = 0; i < n; ++i) out[i] = invert_i32(in[i], saturate != 0);
}

// bwlabel2 (MorphologicOperations.cpp:2012-2035).  relab=0 returns root-index
// labels with background -1; relab=1 returns 1..N in raster order of the roots,
// background 0.  Returns N (relab=1) or 0.
int or_bwlabel2(const uint8_t* bin, int32_t* out, int rows, int cols, int conn, int relab) {
    cc_label(bin, cols, rows, out, -1, conn);
    int j = 0;
    if (relab) j = cc_relabel(cols, rows, out, -1);
    return j;
}

// bwareaopen2 (MorphologicOperations.cpp:2199-2232).  labeled=0: input is u8,
// labelled with union-find first; labeled=1: input is int32 root labels (bg -1).
// flatten=0 writes int32 labels (bg -1) to out_i32; flatten=1 writes u8 255/0
// to out_u8.  Returns count (number of kept components).
int or_bwareaopen2(const void* in, int labeled, int flatten, int min_size, int max_size, int conn,
                   int rows, int cols, int32_t* out_i32, uint8_t* out_u8) {
    const long n = (long)rows * cols;
    std::vector<int> output(n > 0 ? n : 1);
    int count;
    if (!labeled) {
        std::vector<int> temp(n > 0 ? n : 1);
        cc_label((const uint8_t*)in, cols, rows, temp.data(), -1, conn);
        count = cc_area_threshold_labeled(temp.data(), cols, rows, output.data(), -1, min_size, max_size);
    } else {
        count = cc_area_threshold_labeled((const int*)in, cols, rows, output.data(), -1, min_size, max_size);
    }
    if (flatten) {
        for (long i = 0; i < n; ++i) out_u8[i] = output[i] > -1 ? 255 : 0;
    } else {
        std::memcpy(out_i32, output.data(), sizeof(int) * n);
    }
    return count;
}

// ConnComponents::relabel on an int32 label image with background bgval.
int or_relabel(int32_t* labels, int rows, int cols, int bgval) {
    return cc_relabel(cols, rows, labels, bgval);
}

// Threshold glue: OpenCV `Mat > k` on u8 (255 where v > k, else 0).
void or_threshold_gt_u8(const uint8_t* in, uint8_t* out, long n, int k) {
    for (long i = 0; i < n; ++i) out[i] = in[i] > k ? 255 : 0;
}

// The nuclei-segmentation chain pinned by HistologicalEntities.cpp:195,219,246,567
// in the order the reconstruct + CCL benchmark runs it:
//   r = imreconstruct<u8>(marker, mask, recon_conn)
//   f = imfillHoles<u8>(r, binary=false, fill_conn)
//   b = f > thresh
//   l = bwareaopen2(b, labeled=false, flatten=false, min, max, ccl_conn)
//   l = relabel(l)   (canonical raster order, background 0)
// Writes r, f and l; returns the component count after the area filter.
int or_chain_u8(const uint8_t* marker, const uint8_t* mask, int rows, int cols,
                int recon_conn, int fill_conn, int thresh, int ccl_conn, int min_size, int max_size,
                uint8_t* recon_out, uint8_t* fill_out, int32_t* labels_out) {
    const long n = (long)rows * cols;
    imreconstruct<uint8_t>(marker, mask, recon_out, rows, cols, recon_conn);
    imfillHoles<uint8_t>(recon_out, fill_out, rows, cols, false, fill_conn);
    std::vector<uint8_t> b(n > 0 ? n : 1);
    for (long i = 0; i < n; ++i) b[i] = fill_out[i] > thresh ? 255 : 0;
    or_bwareaopen2(b.data(), 0, 0, min_size, max_size, ccl_conn, rows, cols, labels_out, nullptr);
    return cc_relabel(cols, rows, labels_out, -1);
}

}  // extern "C"
