// mini_cv.hpp -- the subset of OpenCV 2.x's cv::Mat that include/nscale_dropin.hpp
// uses (data, step, rows, cols, type(), channels(), Mat(rows, cols, type),
// Mat_<int>(rows, cols)), so the drop-in header compiles and runs where OpenCV
// is not installed.  Single-channel, owning, continuous matrices only.
#ifndef MINI_CV_HPP_
#define MINI_CV_HPP_

#include <cstddef>
#include <memory>
#include <vector>

#define CV_8U 0
#define CV_32S 4
#define CV_32F 5

namespace cv {

class Mat {
public:
    Mat() = default;
    Mat(int r, int c, int t) : rows(r), cols(c), step((size_t)c * esize(t)), type_(t),
                               buf_(std::make_shared<std::vector<unsigned char>>((size_t)r * c * esize(t))) {
        data = buf_->data();
    }
    int type() const { return type_; }
    int channels() const { return 1; }
    unsigned char* data = nullptr;
    int rows = 0, cols = 0;
    size_t step = 0;

    template <typename T> T& at(int y, int x) { return reinterpret_cast<T*>(data + (size_t)y * step)[x]; }
    template <typename T> const T& at(int y, int x) const {
        return reinterpret_cast<const T*>(data + (size_t)y * step)[x];
    }

private:
    static size_t esize(int t) { return t == CV_8U ? 1 : 4; }
    int type_ = CV_8U;
    std::shared_ptr<std::vector<unsigned char>> buf_;
};

template <typename T> class Mat_;
template <> class Mat_<int> : public Mat {
public:
    Mat_(int r, int c) : Mat(r, c, CV_32S) {}
};
template <> class Mat_<unsigned char> : public Mat {
public:
    Mat_(int r, int c) : Mat(r, c, CV_8U) {}
};

}  // namespace cv

#define NSCALE_DROPIN_CV_PROVIDED
#endif  // MINI_CV_HPP_
