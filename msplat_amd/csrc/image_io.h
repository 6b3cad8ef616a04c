// image_io.h — image decoding and the float image operations of the dataset loaders
// (reference core/src/loaders/image_io.cpp): PNG (zlib inflate + the five scanline
// filters, every colour type and bit depth, Adam7), baseline JPEG (Huffman, 8-bit, any
// 1x1 / 2x1 / 2x2 chroma sampling, restart intervals), the area (box-filter) resize and
// the Brown-Conrady undistortion.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace msplat_hip {

// Float RGB in [0, 1], row-major [h][w][3].
struct HostImage {
    int width = 0, height = 0;
    std::vector<float> rgb;
    bool empty() const { return width <= 0 || height <= 0; }
};

// Decodes a PNG or JPEG file (by its signature) to 8-bit RGB. Grey images are replicated to
// RGB; an alpha channel is composited over black (what drawing into an opaque RGB bitmap
// initialised to zero gives); 16-bit samples are rounded to 8 bits. Throws on anything
// else (progressive or arithmetic-coded JPEG, CMYK, corrupt data).
std::vector<uint8_t> decode_image_file(const std::string& path, int& width, int& height);
std::vector<uint8_t> decode_png(const uint8_t* data, size_t n, int& width, int& height);
std::vector<uint8_t> decode_jpeg(const uint8_t* data, size_t n, int& width, int& height);

// imreadRGB (image_io.cpp:13-53): the decoded bytes / 255.
HostImage read_image_rgb(const std::string& path);

// resizeArea (image_io.cpp:113-162): every output pixel is the area-weighted mean of the
// source pixels its footprint covers.
HostImage resize_area(const HostImage& src, int dst_w, int dst_h);

// undistortImage (image_io.cpp:164-345): remaps a distorted image to the pinhole camera with
// the same intrinsics (bilinear, edge-clamped) and crops to the largest rectangle inside
// the undistorted image of the source border; returns the cropped principal point.
struct Undistorted {
    HostImage image;
    float fx, fy, cx, cy;
};
Undistorted undistort_image(const HostImage& src, float fx, float fy, float cx, float cy, float k1, float k2, float p1,
                            float p2, float k3);

// uint8 ground truth from float RGB (input_data.cpp:66-69): round(v * 255) clamped to [0, 255].
std::vector<uint8_t> quantize_rgb(const HostImage& img);

}  // namespace msplat_hip
