// ref_driver.cpp -- TEST INFRASTRUCTURE ONLY.
// A C-callable shim around the reference's own, unmodified
// src/segment/ConnComponents.cpp, which oracle/Makefile compiles from where it
// lies under the reference checkout into oracle/_ref/libccref.so.  Tests use it
// to pin the oracle's union-find CCL, relabel and area filter exactly.
#include "ConnComponents.h"

#include <cstdlib>

extern "C" {

void ref_label(unsigned char* img, int w, int h, int* label, int bgval, int conn) {
    nscale::ConnComponents cc;
    cc.label(img, w, h, label, bgval, conn);
}

int ref_relabel(int w, int h, int* label, int bgval) {
    nscale::ConnComponents cc;
    return cc.relabel(w, h, label, bgval);
}

int ref_area_threshold_labeled(const int* label, int w, int h, int* n_label, int bgval, int lower, int upper) {
    nscale::ConnComponents cc;
    return cc.areaThresholdLabeled(label, w, h, n_label, bgval, lower, upper);
}

int ref_area_threshold(unsigned char* img, int w, int h, int* label, int bgval, int lower, int upper, int conn) {
    nscale::ConnComponents cc;
    return cc.areaThreshold(img, w, h, label, bgval, lower, upper, conn);
}

// boundingBox returns a malloc'ed 5 x count array: copied into out (at most
// `capacity` ints) and freed.  Returns count.
int ref_bounding_box(const int* label, int w, int h, int bgval, int* out, int capacity) {
    nscale::ConnComponents cc;
    int n = 0;
    int* b = cc.boundingBox(w, h, label, bgval, n);
    const int m = 5 * n < capacity ? 5 * n : capacity;
    for (int i = 0; i < m; ++i) out[i] = b[i];
    std::free(b);
    return n;
}

}
