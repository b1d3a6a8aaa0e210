#include "npy.hpp"

#include <cstring>
#include <fstream>
#include <sstream>

namespace rsf {
namespace {

std::string dict_value(const std::string& hdr, const std::string& key) {
    const std::string k1 = "'" + key + "'";
    size_t p = hdr.find(k1);
    if (p == std::string::npos) return "";
    p = hdr.find(':', p + k1.size());
    if (p == std::string::npos) return "";
    ++p;
    while (p < hdr.size() && hdr[p] == ' ') ++p;
    if (p >= hdr.size()) return "";
    if (hdr[p] == '\'') {
        size_t e = hdr.find('\'', p + 1);
        return e == std::string::npos ? "" : hdr.substr(p, e - p + 1);
    }
    if (hdr[p] == '(') {
        size_t e = hdr.find(')', p);
        return e == std::string::npos ? "" : hdr.substr(p, e - p + 1);
    }
    size_t e = hdr.find_first_of(",}", p);
    return hdr.substr(p, e == std::string::npos ? std::string::npos : e - p);
}

size_t dtype_size(const std::string& d) {
    if (d.size() < 3) return 0;
    return (size_t)std::atoi(d.c_str() + 2);
}

}  // namespace

std::string npy_load(const std::string& path, NpyArray* out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return "cannot open '" + path + "'";
    char magic[8];
    if (!f.read(magic, 8) || std::memcmp(magic, "\x93NUMPY", 6) != 0) return "'" + path + "' is not a .npy file";
    const int major = (unsigned char)magic[6];
    uint32_t hlen = 0;
    if (major == 1) {
        unsigned char b[2];
        if (!f.read((char*)b, 2)) return "truncated .npy header in '" + path + "'";
        hlen = b[0] | (b[1] << 8);
    } else if (major == 2 || major == 3) {
        unsigned char b[4];
        if (!f.read((char*)b, 4)) return "truncated .npy header in '" + path + "'";
        hlen = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24);
    } else {
        return "unsupported .npy version " + std::to_string(major) + " in '" + path + "'";
    }
    if (hlen > (1u << 20)) return "oversized .npy header in '" + path + "'";
    std::string hdr(hlen, '\0');
    if (!f.read(&hdr[0], hlen)) return "truncated .npy header in '" + path + "'";
    std::string descr = dict_value(hdr, "descr");
    std::string fortran = dict_value(hdr, "fortran_order");
    std::string shape = dict_value(hdr, "shape");
    if (descr.size() < 2 || shape.empty()) return "malformed .npy header in '" + path + "'";
    descr = descr.substr(1, descr.size() - 2);
    if (fortran.find("True") != std::string::npos)
        return "'" + path + "' is Fortran-ordered; save it in C order (np.ascontiguousarray)";
    static const char* ok[] = {"|u1", "<u1", "|i1", "<i1", "|b1", "<u2", "<i2", "<u4", "<i4", "<f4", "<f8"};
    bool good = false;
    for (auto* d : ok) good |= descr == d;
    if (!good) return "unsupported dtype '" + descr + "' in '" + path + "'";
    out->descr = descr;
    out->shape.clear();
    std::string s = shape.substr(1, shape.size() - 2);
    std::stringstream ss(s);
    std::string tok;
    while (std::getline(ss, tok, ',')) {
        size_t a = tok.find_first_not_of(" ");
        if (a == std::string::npos) continue;
        long long v = std::atoll(tok.c_str() + a);
        if (v < 0) return "negative dimension in '" + path + "'";
        out->shape.push_back(v);
    }
    const size_t es = dtype_size(descr);
    const int64_t n = out->count();
    if (n < 0 || (uint64_t)n > (1ull << 40)) return "array too large in '" + path + "'";
    out->data.resize((size_t)n * es);
    if (n > 0 && !f.read((char*)out->data.data(), (std::streamsize)out->data.size()))
        return "truncated .npy payload in '" + path + "' (expected " + std::to_string(out->data.size()) + " bytes)";
    return "";
}

std::string npy_save_f32(const std::string& path, const float* data, const std::vector<int64_t>& shape) {
    std::ostringstream h;
    h << "{'descr': '<f4', 'fortran_order': False, 'shape': (";
    int64_t n = 1;
    for (size_t k = 0; k < shape.size(); ++k) {
        h << shape[k] << (shape.size() == 1 || k + 1 < shape.size() ? ", " : "");
        n *= shape[k];
    }
    h << "), }";
    std::string hs = h.str();
    const size_t total = 10 + hs.size() + 1;
    hs += std::string((64 - total % 64) % 64, ' ');
    hs += '\n';
    std::ofstream f(path, std::ios::binary);
    if (!f) return "cannot write '" + path + "'";
    f.write("\x93NUMPY\x01\x00", 8);
    const uint16_t hl = (uint16_t)hs.size();
    const unsigned char b[2] = {(unsigned char)(hl & 0xff), (unsigned char)(hl >> 8)};
    f.write((const char*)b, 2);
    f.write(hs.data(), (std::streamsize)hs.size());
    f.write((const char*)data, (std::streamsize)(n * sizeof(float)));
    if (!f) return "write failed for '" + path + "'";
    return "";
}

std::string npy_to_f32(const NpyArray& a, std::vector<float>* out) {
    const int64_t n = a.count();
    out->resize((size_t)n);
    const uint8_t* p = a.data.data();
    const std::string& d = a.descr;
    for (int64_t i = 0; i < n; ++i) {
        float v;
        if (d == "|u1" || d == "<u1" || d == "|b1") v = (float)p[i];
        else if (d == "|i1" || d == "<i1") v = (float)(int8_t)p[i];
        else if (d == "<u2") { uint16_t x; std::memcpy(&x, p + 2 * i, 2); v = (float)x; }
        else if (d == "<i2") { int16_t x; std::memcpy(&x, p + 2 * i, 2); v = (float)x; }
        else if (d == "<u4") { uint32_t x; std::memcpy(&x, p + 4 * i, 4); v = (float)x; }
        else if (d == "<i4") { int32_t x; std::memcpy(&x, p + 4 * i, 4); v = (float)x; }
        else if (d == "<f4") { std::memcpy(&v, p + 4 * i, 4); }
        else if (d == "<f8") { double x; std::memcpy(&x, p + 8 * i, 8); v = (float)x; }
        else return "unsupported dtype '" + d + "'";
        (*out)[(size_t)i] = v;
    }
    return "";
}

}  // namespace rsf
