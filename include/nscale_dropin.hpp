// nscale_dropin.hpp -- header-only C++ drop-in for the nscale:: reconstruction
// and connected-component functions, over the C ABI of nscale_hip.h.
//
// Same names, template parameters, arguments and results as the reference
// declarations in src/segment/MorphologicOperations.h (paths relative to
// fastIA-master/):
//     imreconstruct<T>(seeds, image, connectivity)            :21-22
//     imreconstructBinary<T>(seeds, binaryImage, connectivity) :43-44
//     imfill<T>(image, seeds, binary, connectivity)            :48
//     imfillHoles<T>(image, binary, connectivity)              :50-51
//     bwselect<T>(binaryImage, seeds, connectivity)            :56
//     imhmin<T>(image, h, connectivity)                        :78
//     localMaxima<T>(image, connectivity)                      :84
//     localMinima<T>(image, connectivity)                      :86
//     distanceTransform(mask, calcDist)                        :92
//     bwlabel2(binaryImage, connectivity, relab)               :59
//     bwareaopen2(binaryImage, labeled, flatten, minSize, maxSize, connectivity, count)  :74
// At a call site, include this header instead of MorphologicOperations.h for
// these functions and link libnscale_hip.so.
//
// Behaviour: synchronous like the reference CPU functions.  Inputs are host
// Mats (single channel, any step); results are newly allocated continuous Mats.
// A failed CV_Assert of the reference (wrong type or channel count) and any
// device error throw std::runtime_error with the library's message.  Each
// calling thread gets its own GPU context (device 0) on first use.
//
// OpenCV's <opencv2/core/core.hpp> is included unless the including file
// defines NSCALE_DROPIN_CV_PROVIDED after providing a cv::Mat with the members
// used here: data, step (convertible to size_t), rows, cols, type(),
// channels(), and the constructors Mat(rows, cols, type), Mat_<int>(rows, cols)
// and Mat_<unsigned char>(rows, cols).
#ifndef NSCALE_DROPIN_HPP_
#define NSCALE_DROPIN_HPP_

#ifndef NSCALE_DROPIN_CV_PROVIDED
#include <opencv2/core/core.hpp>
#endif

#include <cstdint>
#include <stdexcept>
#include <string>

#include "nscale_hip.h"

namespace nscale {
namespace nsx_detail {

struct Ctx {
    nsx_ctx* h = nullptr;
    Ctx() {
        if (nsx_ctx_create(&h, 0, 0) != NSX_OK) h = nullptr;
    }
    ~Ctx() { nsx_ctx_destroy(h); }
};

inline nsx_ctx* ctx() {
    thread_local Ctx c;
    if (!c.h) throw std::runtime_error("nscale_hip: no GPU context (no ROCm device visible)");
    return c.h;
}

inline void check(nsx_status st, nsx_ctx* c) {
    if (st != NSX_OK) throw std::runtime_error(std::string("nscale_hip: ") + nsx_last_error(c));
}

template <typename T> struct Type;
template <> struct Type<unsigned char> { static constexpr int cv = CV_8U, nsx = NSX_U8; };
template <> struct Type<int> { static constexpr int cv = CV_32S, nsx = NSX_I32; };
template <> struct Type<float> { static constexpr int cv = CV_32F, nsx = NSX_F32; };

inline nsx_img host(const cv::Mat& m, int dtype) {
    nsx_img i;
    i.data = (void*)m.data;
    i.pitch_bytes = (int64_t)(size_t)m.step;
    i.rows = m.rows;
    i.cols = m.cols;
    i.dtype = dtype;
    i.mem = NSX_MEM_HOST;
    return i;
}

template <typename T>
inline void require(const cv::Mat& m, const char* what) {
    if (m.channels() != 1) throw std::runtime_error(std::string("nscale_hip: ") + what + " must have one channel");
    if (m.type() != Type<T>::cv) throw std::runtime_error(std::string("nscale_hip: ") + what + " has the wrong type");
}

}  // namespace nsx_detail

template <typename T>
cv::Mat imreconstruct(const cv::Mat& seeds, const cv::Mat& image, int connectivity) {
    using namespace nsx_detail;
    require<T>(seeds, "seeds");
    require<T>(image, "image");
    cv::Mat out(image.rows, image.cols, Type<T>::cv);
    nsx_img a = host(seeds, Type<T>::nsx), m = host(image, Type<T>::nsx), o = host(out, Type<T>::nsx);
    nsx_ctx* c = ctx();
    check(nsx_imreconstruct(c, &a, &m, &o, connectivity, nullptr), c);
    return out;
}

template <typename T>
cv::Mat imreconstructBinary(const cv::Mat& seeds, const cv::Mat& binaryImage, int connectivity) {
    using namespace nsx_detail;
    require<T>(seeds, "seeds");
    require<T>(binaryImage, "binaryImage");
    cv::Mat out(binaryImage.rows, binaryImage.cols, Type<T>::cv);
    nsx_img a = host(seeds, Type<T>::nsx), m = host(binaryImage, Type<T>::nsx), o = host(out, Type<T>::nsx);
    nsx_ctx* c = ctx();
    check(nsx_imreconstruct_binary(c, &a, &m, &o, connectivity, nullptr), c);
    return out;
}

template <typename T>
cv::Mat imfillHoles(const cv::Mat& image, bool binary, int connectivity) {
    using namespace nsx_detail;
    require<T>(image, "image");
    cv::Mat out(image.rows, image.cols, Type<T>::cv);
    nsx_img i = host(image, Type<T>::nsx), o = host(out, Type<T>::nsx);
    nsx_ctx* c = ctx();
    check(nsx_fill_holes(c, &i, &o, binary ? 1 : 0, connectivity, nullptr), c);
    return out;
}

// imfill and bwselect are instantiated for uchar only in the reference
// (MorphologicOperations.cpp:2516,2522); other T throw.
template <typename T>
cv::Mat imfill(const cv::Mat& image, const cv::Mat& seeds, bool binary, int connectivity) {
    using namespace nsx_detail;
    require<T>(image, "image");
    require<unsigned char>(seeds, "seeds");
    cv::Mat out(image.rows, image.cols, Type<T>::cv);
    nsx_img i = host(image, Type<T>::nsx), s = host(seeds, NSX_U8), o = host(out, Type<T>::nsx);
    nsx_ctx* c = ctx();
    check(nsx_imfill(c, &i, &s, &o, binary ? 1 : 0, connectivity, nullptr), c);
    return out;
}

template <typename T>
cv::Mat bwselect(const cv::Mat& binaryImage, const cv::Mat& seeds, int connectivity) {
    using namespace nsx_detail;
    require<T>(binaryImage, "binaryImage");
    require<unsigned char>(seeds, "seeds");
    cv::Mat out(binaryImage.rows, binaryImage.cols, Type<T>::cv);
    nsx_img b = host(binaryImage, Type<T>::nsx), s = host(seeds, NSX_U8), o = host(out, Type<T>::nsx);
    nsx_ctx* c = ctx();
    check(nsx_bwselect(c, &b, &s, &o, connectivity, nullptr), c);
    return out;
}

template <typename T>
cv::Mat imhmin(const cv::Mat& image, T h, int connectivity) {
    using namespace nsx_detail;
    require<T>(image, "image");
    cv::Mat out(image.rows, image.cols, Type<T>::cv);
    nsx_img i = host(image, Type<T>::nsx), o = host(out, Type<T>::nsx);
    nsx_ctx* c = ctx();
    check(nsx_imhmin(c, &i, (double)h, &o, connectivity, nullptr), c);
    return out;
}

template <typename T>
cv::Mat_<unsigned char> localMaxima(const cv::Mat& image, int connectivity) {
    using namespace nsx_detail;
    require<T>(image, "image");
    cv::Mat_<unsigned char> out(image.rows, image.cols);
    nsx_img i = host(image, Type<T>::nsx), o = host(out, NSX_U8);
    nsx_ctx* c = ctx();
    check(nsx_local_maxima(c, &i, &o, connectivity, nullptr), c);
    return out;
}

template <typename T>
cv::Mat_<unsigned char> localMinima(const cv::Mat& image, int connectivity) {
    using namespace nsx_detail;
    require<T>(image, "image");
    cv::Mat_<unsigned char> out(image.rows, image.cols);
    nsx_img i = host(image, Type<T>::nsx), o = host(out, NSX_U8);
    nsx_ctx* c = ctx();
    check(nsx_local_minima(c, &i, &o, connectivity, nullptr), c);
    return out;
}

inline cv::Mat distanceTransform(const cv::Mat& mask, bool calcDist = true) {
    using namespace nsx_detail;
    require<unsigned char>(mask, "mask");
    cv::Mat out(mask.rows, mask.cols, calcDist ? CV_32F : CV_32S);
    nsx_img m = host(mask, NSX_U8), o = host(out, calcDist ? NSX_F32 : NSX_I32);
    nsx_ctx* c = ctx();
    check(nsx_distance_transform(c, &m, &o, calcDist ? 1 : 0, nullptr), c);
    return out;
}

inline cv::Mat_<int> bwlabel2(const cv::Mat& binaryImage, int connectivity, bool relab) {
    using namespace nsx_detail;
    require<unsigned char>(binaryImage, "binaryImage");
    cv::Mat_<int> out(binaryImage.rows, binaryImage.cols);
    nsx_img b = host(binaryImage, NSX_U8), o = host(out, NSX_I32);
    nsx_ctx* c = ctx();
    check(nsx_bwlabel(c, &b, &o, connectivity, relab ? 1 : 0, nullptr, nullptr), c);
    return out;
}

inline cv::Mat bwareaopen2(const cv::Mat& binaryImage, bool labeled, bool flatten, int minSize, int maxSize,
                           int connectivity, int& count) {
    using namespace nsx_detail;
    if (labeled) require<int>(binaryImage, "binaryImage");
    else require<unsigned char>(binaryImage, "binaryImage");
    cv::Mat out(binaryImage.rows, binaryImage.cols, flatten ? CV_8U : CV_32S);
    nsx_img b = host(binaryImage, labeled ? NSX_I32 : NSX_U8), o = host(out, flatten ? NSX_U8 : NSX_I32);
    nsx_ctx* c = ctx();
    int32_t n = 0;
    check(nsx_bwareaopen(c, &b, labeled ? 1 : 0, flatten ? 1 : 0, minSize, maxSize, connectivity, &o, &n, nullptr), c);
    count = n;
    return out;
}

}  // namespace nscale

#endif  // NSCALE_DROPIN_HPP_
