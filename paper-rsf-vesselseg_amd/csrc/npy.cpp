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

}  // namespace

size_t npy_dtype_size(const std::string& d) {
    if (d.size() < 3) return 0;
    return (size_t)std::atoi(d.c_str() + 2);
}

std::string npy_read_header(std::istream& f, const std::string& path, NpyArray* out, int64_t* data_offset) {
    char magic[8];
    if (!f.read(magic, 8) || std::memcmp(magic, "\x93NUMPY", 6) != 0) return "'" + path + "' is not a .npy file";
    const int major = (unsigned char)magic[6];
    uint32_t hlen = 0;
    int64_t pre = 10;
    if (major == 1) {
        unsigned char b[2];
        if (!f.read((char*)b, 2)) return "truncated .npy header in '" + path + "'";
        hlen = b[0] | (b[1] << 8);
    } else if (major == 2 || major == 3) {
        unsigned char b[4];
        if (!f.read((char*)b, 4)) return "truncated .npy header in '" + path + "'";
        hlen = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24);
        pre = 12;
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
    const int64_t n = out->count();
    if (n < 0 || (uint64_t)n > (1ull << 40)) return "array too large in '" + path + "'";
    *data_offset = pre + (int64_t)hlen;
    return "";
}

std::string npy_load(const std::string& path, NpyArray* out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return "cannot open '" + path + "'";
    int64_t data_offset = 0;
    std::string err = npy_read_header(f, path, out, &data_offset);
    if (!err.empty()) return err;
    const size_t es = npy_dtype_size(out->descr);
    const int64_t n = out->count();
    out->data.resize((size_t)n * es);
    if (n > 0 && !f.read((char*)out->data.data(), (std::streamsize)out->data.size()))
        return "truncated .npy payload in '" + path + "' (expected " + std::to_string(out->data.size()) + " bytes)";
    return "";
}

std::string npy_header_f32(const std::vector<int64_t>& shape) {
    std::ostringstream h;
    h << "{'descr': '<f4', 'fortran_order': False, 'shape': (";
    for (size_t k = 0; k < shape.size(); ++k) h << shape[k] << (shape.size() == 1 || k + 1 < shape.size() ? ", " : "");
    h << "), }";
    std::string hs = h.str();
    const size_t total = 10 + hs.size() + 1;
    hs += std::string((64 - total % 64) % 64, ' ');
    hs += '\n';
    const uint16_t hl = (uint16_t)hs.size();
    std::string out("\x93NUMPY\x01\x00", 8);
    out += (char)(hl & 0xff);
    out += (char)(hl >> 8);
    return out + hs;
}

std::string npy_save_f32(const std::string& path, const float* data, const std::vector<int64_t>& shape) {
    int64_t n = 1;
    for (auto v : shape) n *= v;
    const std::string hs = npy_header_f32(shape);
    std::ofstream f(path, std::ios::binary);
    if (!f) return "cannot write '" + path + "'";
    f.write(hs.data(), (std::streamsize)hs.size());
    f.write((const char*)data, (std::streamsize)(n * sizeof(float)));
    if (!f) return "write failed for '" + path + "'";
    return "";
}

std::string npy_to_f32(const NpyArray& a, std::vector<float>* out) {
    out->resize((size_t)a.count());
    return npy_convert_f32(a.descr, a.data.data(), a.count(), out->data());
}

std::string npy_convert_f32(const std::string& d, const uint8_t* p, int64_t n, float* out) {
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
        out[i] = v;
    }
    return "";
}

}  // namespace rsf
