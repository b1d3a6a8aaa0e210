// Minimal .npy reader/writer for the rsf CLI (format versions 1.0-3.0,
// little-endian, C order).  Replaces tira::volume<float>::load_npy/save_npy
// (CUDA_RSF/main.cpp:145,149,316).
#pragma once

#include <cstdint>
#include <istream>
#include <string>
#include <vector>

namespace rsf {

struct NpyArray {
    std::string descr;             // e.g. "|u1", "<f4"
    std::vector<int64_t> shape;    // C order
    std::vector<uint8_t> data;     // raw little-endian payload
    int64_t count() const {
        int64_t n = 1;
        for (auto s : shape) n *= s;
        return n;
    }
};

// Returns an empty string on success, else a message.
std::string npy_load(const std::string& path, NpyArray* out);
// Parses the header of an open .npy stream (descr and shape of *out; data left
// empty) and gives the byte offset of the payload.
std::string npy_read_header(std::istream& f, const std::string& path, NpyArray* out, int64_t* data_offset);
size_t npy_dtype_size(const std::string& descr);
// The bytes of a version-1.0 header for a C-ordered little-endian fp32 array.
std::string npy_header_f32(const std::vector<int64_t>& shape);
// Converts n elements of dtype descr at p to float32.
std::string npy_convert_f32(const std::string& descr, const uint8_t* p, int64_t n, float* out);
std::string npy_save_f32(const std::string& path, const float* data, const std::vector<int64_t>& shape);
// Converts any supported numeric dtype to float32.
std::string npy_to_f32(const NpyArray& a, std::vector<float>* out);

}  // namespace rsf
