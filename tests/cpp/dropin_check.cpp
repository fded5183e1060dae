// Runs every nscale:: drop-in function of include/nscale_dropin.hpp on a fixed
// synthetic image and writes the results as raw files into the directory given
// as argv[1] (the test compares them with the oracle).  With no GPU the first
// call throws and the program prints "no-gpu" and exits 3.
#include <cstdio>
#include <string>

#include "mini_cv.hpp"
#include "nscale_dropin.hpp"

static void save(const std::string& path, const cv::Mat& m, size_t esize) {
    FILE* f = std::fopen(path.c_str(), "wb");
    for (int y = 0; y < m.rows; ++y) std::fwrite(m.data + (size_t)y * m.step, esize, (size_t)m.cols, f);
    std::fclose(f);
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    const int R = 61, C = 83;
    // mask(y, x) = (x * 37 + y * 91 + ((x * y) >> 3)) & 255; marker = saturating mask - 30
    // seeds(y, x) = 1 where (13 x + 7 y) % 97 == 0
    cv::Mat mask(R, C, CV_8U), marker(R, C, CV_8U), bin(R, C, CV_8U), seeds(R, C, CV_8U);
    for (int y = 0; y < R; ++y)
        for (int x = 0; x < C; ++x) {
            const int v = (x * 37 + y * 91 + ((x * y) >> 3)) & 255;
            mask.at<unsigned char>(y, x) = (unsigned char)v;
            marker.at<unsigned char>(y, x) = (unsigned char)(v > 30 ? v - 30 : 0);
            bin.at<unsigned char>(y, x) = v > 128 ? 255 : 0;
            seeds.at<unsigned char>(y, x) = (x * 13 + y * 7) % 97 == 0 ? 1 : 0;
        }
    try {
        save(dir + "/recon8.bin", nscale::imreconstruct<unsigned char>(marker, mask, 8), 1);
        save(dir + "/recon4.bin", nscale::imreconstruct<unsigned char>(marker, mask, 4), 1);
        save(dir + "/reconbin8.bin", nscale::imreconstructBinary<unsigned char>(marker, bin, 8), 1);
        save(dir + "/fill4.bin", nscale::imfillHoles<unsigned char>(mask, false, 4), 1);
        save(dir + "/fillbin8.bin", nscale::imfillHoles<unsigned char>(bin, true, 8), 1);
        save(dir + "/imfill4.bin", nscale::imfill<unsigned char>(mask, seeds, false, 4), 1);
        save(dir + "/imfillbin8.bin", nscale::imfill<unsigned char>(bin, seeds, true, 8), 1);
        save(dir + "/bwselect8.bin", nscale::bwselect<unsigned char>(bin, seeds, 8), 1);
        save(dir + "/hmin8.bin", nscale::imhmin<unsigned char>(mask, (unsigned char)20, 8), 1);
        save(dir + "/lmax8.bin", nscale::localMaxima<unsigned char>(mask, 8), 1);
        save(dir + "/lmin4.bin", nscale::localMinima<unsigned char>(mask, 4), 1);
        save(dir + "/dist.bin", nscale::distanceTransform(bin), 4);
        save(dir + "/distnn.bin", nscale::distanceTransform(bin, false), 4);
        save(dir + "/label8.bin", nscale::bwlabel2(bin, 8, true), 4);
        int count = 0;
        save(dir + "/area8.bin", nscale::bwareaopen2(bin, false, false, 3, 40, 8, count), 4);
        std::printf("count %d\n", count);
        try {
            cv::Mat wrong(R, C, CV_32S);
            nscale::imreconstruct<unsigned char>(wrong, mask, 8);
            std::printf("type check missing\n");
            return 1;
        } catch (const std::runtime_error&) {
        }
    } catch (const std::runtime_error& e) {
        std::printf("no-gpu %s\n", e.what());
        return 3;
    }
    std::printf("ok\n");
    return 0;
}
