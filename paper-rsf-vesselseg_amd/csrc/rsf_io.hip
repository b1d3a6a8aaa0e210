// Streaming .npy I/O between files and device memory, and the file-to-file run
// of the rsf executable.  The reference loads whole volumes into host memory
// (tira::volume::load_npy, CUDA_RSF/main.cpp:145,149) and writes phi back the
// same way (main.cpp:316); gigavoxel volumes (0.8-3 GiB) pay for a pageable
// copy and a second full host buffer.  Here a file is read in chunks with
// pread straight into one of two pinned staging buffers while the other one is
// copied to the device on the context's stream (and the reverse for saving),
// so disk/page-cache reads overlap the PCIe transfers and no whole-volume host
// buffer exists.
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "npy.hpp"
#include "rsf_internal.hpp"

namespace rsf {
namespace {

constexpr size_t kChunk = size_t(32) << 20;  // bytes per staging buffer

struct Fd {
    int fd = -1;
    ~Fd() {
        if (fd >= 0) ::close(fd);
    }
};

struct Pinned {
    void* p = nullptr;
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
};

struct Events {
    hipEvent_t e[2] = {nullptr, nullptr};
    ~Events() {
        for (auto x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

// Declared after the staging buffers, so that on every return path the copies
// in flight finish before the buffers are freed.
struct StreamSync {
    hipStream_t s;
    ~StreamSync() { (void)hipStreamSynchronize(s); }
};

int header_of(const char* path, NpyArray* a, int64_t* data_offset) {
    if (!path) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL path");
    std::ifstream f(path, std::ios::binary);
    if (!f) return set_error(RSF_ERR_IO, "cannot open '%s'", path);
    const std::string err = npy_read_header(f, path, a, data_offset);
    if (!err.empty()) return set_error(RSF_ERR_IO, "%s", err.c_str());
    return RSF_OK;
}

bool is_u8(const std::string& d) { return d == "|u1" || d == "<u1" || d == "|b1"; }

// pread of exactly n bytes at offset off
int read_full(int fd, void* dst, size_t n, int64_t off, const char* path) {
    char* p = (char*)dst;
    while (n > 0) {
        const ssize_t r = ::pread(fd, p, n, (off_t)off);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return set_error(RSF_ERR_IO, "truncated .npy payload in '%s'", path);
        p += r;
        n -= (size_t)r;
        off += r;
    }
    return RSF_OK;
}

int write_full(int fd, const void* src, size_t n, int64_t off, const char* path) {
    const char* p = (const char*)src;
    while (n > 0) {
        const ssize_t r = ::pwrite(fd, p, n, (off_t)off);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return set_error(RSF_ERR_IO, "write failed for '%s'", path);
        p += r;
        n -= (size_t)r;
        off += r;
    }
    return RSF_OK;
}

}  // namespace

// Streams a 3-D .npy file into device memory: uint8 as is (dst_dtype U8, the
// file must be uint8/bool), anything else converted to float32 (dst_dtype F32).
int npy_load_device(hipStream_t s, const char* path, void* dst, int dst_dtype, int64_t dst_bytes,
                    std::vector<int64_t>* shape) {
    NpyArray a;
    int64_t off = 0;
    RSF_TRY(header_of(path, &a, &off));
    if (a.shape.size() != 3) return set_error(RSF_ERR_INVALID_ARGUMENT, "'%s' is not a 3-D array (Z, Y, X)", path);
    if (dst_dtype == RSF_DTYPE_U8 && !is_u8(a.descr))
        return set_error(RSF_ERR_INVALID_ARGUMENT, "'%s' holds %s, not uint8", path, a.descr.c_str());
    const int64_t n = a.count();
    const size_t es = npy_dtype_size(a.descr), os = dst_dtype == RSF_DTYPE_U8 ? 1 : 4;
    if (dst_bytes < n * (int64_t)os) return set_error(RSF_ERR_INVALID_ARGUMENT, "destination too small for '%s'", path);
    *shape = a.shape;
    Fd fd;
    fd.fd = ::open(path, O_RDONLY);
    if (fd.fd < 0) return set_error(RSF_ERR_IO, "cannot open '%s'", path);
    struct stat st;
    if (::fstat(fd.fd, &st) != 0 || st.st_size < off + n * (int64_t)es)
        return set_error(RSF_ERR_IO, "truncated .npy payload in '%s' (expected %lld bytes)", path,
                         (long long)(n * (int64_t)es));
    const bool raw = (os == es) && (os == 1 || a.descr == "<f4");  // no conversion: pread into pinned
    const int64_t per = (int64_t)(kChunk / std::max(es, os));          // elements per chunk
    Pinned stage[2];
    for (auto& b : stage) RSF_HIP(hipHostMalloc(&b.p, kChunk, hipHostMallocDefault));
    std::vector<uint8_t> tmp(raw ? 0 : (size_t)per * es);
    Events ev;
    for (auto& e : ev.e) RSF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    StreamSync guard{s};
    bool used[2] = {false, false};
    for (int64_t i = 0, k = 0; i < n; i += per, ++k) {
        const int b = (int)(k & 1);
        const int64_t m = std::min(per, n - i);
        if (used[b]) RSF_HIP(hipEventSynchronize(ev.e[b]));  // the copy out of this buffer is done
        if (raw) {
            RSF_TRY(read_full(fd.fd, stage[b].p, (size_t)(m * (int64_t)es), off + i * (int64_t)es, path));
        } else {
            RSF_TRY(read_full(fd.fd, tmp.data(), (size_t)(m * (int64_t)es), off + i * (int64_t)es, path));
            const std::string err = npy_convert_f32(a.descr, tmp.data(), m, (float*)stage[b].p);
            if (!err.empty()) return set_error(RSF_ERR_IO, "%s", err.c_str());
        }
        RSF_HIP(hipMemcpyAsync((char*)dst + i * (int64_t)os, stage[b].p, (size_t)(m * (int64_t)os),
                               hipMemcpyHostToDevice, s));
        RSF_HIP(hipEventRecord(ev.e[b], s));
        used[b] = true;
    }
    RSF_HIP(hipStreamSynchronize(s));
    return RSF_OK;
}

// Streams n float32 values from device memory into a .npy file of the given shape.
int npy_save_device(hipStream_t s, const char* path, const float* src, const std::vector<int64_t>& shape) {
    if (!path) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL path");
    int64_t n = 1;
    for (auto v : shape) n *= v;
    const std::string hdr = npy_header_f32(shape);
    Fd fd;
    fd.fd = ::open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fd.fd < 0) return set_error(RSF_ERR_IO, "cannot write '%s'", path);
    RSF_TRY(write_full(fd.fd, hdr.data(), hdr.size(), 0, path));
    const int64_t off = (int64_t)hdr.size(), per = (int64_t)(kChunk / 4);
    Pinned stage[2];
    for (auto& b : stage) RSF_HIP(hipHostMalloc(&b.p, kChunk, hipHostMallocDefault));
    Events ev;
    for (auto& e : ev.e) RSF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    StreamSync guard{s};
    // copy chunk k+1 to the host while chunk k is written to the file
    auto issue = [&](int64_t k) -> int {
        const int64_t i = k * per;
        if (i >= n) return RSF_OK;
        const int b = (int)(k & 1);
        RSF_HIP(hipMemcpyAsync(stage[b].p, src + i, (size_t)(std::min(per, n - i) * 4), hipMemcpyDeviceToHost, s));
        RSF_HIP(hipEventRecord(ev.e[b], s));
        return RSF_OK;
    };
    RSF_TRY(issue(0));
    for (int64_t k = 0; k * per < n; ++k) {
        const int b = (int)(k & 1);
        const int64_t i = k * per, m = std::min(per, n - i);
        RSF_HIP(hipEventSynchronize(ev.e[b]));
        RSF_TRY(issue(k + 1));
        RSF_TRY(write_full(fd.fd, stage[b].p, (size_t)(m * 4), off + i * 4, path));
    }
    if (::close(fd.fd) != 0) {
        fd.fd = -1;
        return set_error(RSF_ERR_IO, "write failed for '%s'", path);
    }
    fd.fd = -1;
    return RSF_OK;
}

}  // namespace rsf

using namespace rsf;

extern "C" {

int rsf_npy_info(const char* path, int64_t* shape, int* ndim, int* u8) {
    NpyArray a;
    int64_t off = 0;
    RSF_TRY(header_of(path, &a, &off));
    struct stat st;
    if (::stat(path, &st) != 0 || st.st_size < off + a.count() * (int64_t)npy_dtype_size(a.descr))
        return set_error(RSF_ERR_IO, "truncated .npy payload in '%s' (expected %lld bytes)", path,
                         (long long)(a.count() * (int64_t)npy_dtype_size(a.descr)));
    if (ndim) *ndim = (int)a.shape.size();
    if (shape)
        for (size_t k = 0; k < a.shape.size() && k < 8; ++k) shape[k] = a.shape[k];
    if (u8) *u8 = is_u8(a.descr) ? 1 : 0;
    return RSF_OK;
}

int rsf_npy_load_device(rsf_context* c, const char* path, void* dst, int dst_dtype, int64_t dst_bytes) {
    hipStream_t s;
    RSF_TRY(context_stream(c, &s));
    if (!dst) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL destination");
    if (dst_dtype != RSF_DTYPE_U8 && dst_dtype != RSF_DTYPE_F32) return set_error(RSF_ERR_INVALID_ARGUMENT, "dtype");
    std::vector<int64_t> shape;
    return npy_load_device(s, path, dst, dst_dtype, dst_bytes, &shape);
}

int rsf_npy_save_device(rsf_context* c, const char* path, const float* src, int64_t w, int64_t h, int64_t l) {
    hipStream_t s;
    RSF_TRY(context_stream(c, &s));
    if (!src || w <= 0 || h <= 0 || l <= 0) return set_error(RSF_ERR_INVALID_ARGUMENT, "bad source");
    return npy_save_device(s, path, src, {l, h, w});
}

int rsf_run_npy(rsf_context* c, const char* image_path, const char* seed_path, int seed_is_phi0,
                const char* out_path, const rsf_params* p, float* iter_ms) {
    hipStream_t s;
    RSF_TRY(context_stream(c, &s));
    if (!p) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL params");
    NpyArray ai, as;
    int64_t o1 = 0, o2 = 0;
    RSF_TRY(header_of(image_path, &ai, &o1));
    RSF_TRY(header_of(seed_path, &as, &o2));
    if (ai.shape.size() != 3) return set_error(RSF_ERR_INVALID_ARGUMENT, "'%s' is not a 3-D array (Z, Y, X)", image_path);
    if (as.shape != ai.shape) return set_error(RSF_ERR_INVALID_ARGUMENT, "initial level set and image shapes differ");
    const int64_t l = ai.shape[0], h = ai.shape[1], w = ai.shape[2], n = w * h * l;
    if (w <= 0 || h <= 0 || l <= 0) return set_error(RSF_ERR_INVALID_ARGUMENT, "'%s' has an empty dimension", image_path);
    const int dtype = is_u8(ai.descr) ? RSF_DTYPE_U8 : RSF_DTYPE_F32;
    DevBuf img, phi;
    RSF_TRY(img.ensure((size_t)n * (dtype == RSF_DTYPE_U8 ? 1 : 4)));
    RSF_TRY(phi.ensure((size_t)n * 4));
    std::vector<int64_t> shape;
    RSF_TRY(npy_load_device(s, image_path, img.p, dtype, (int64_t)img.bytes, &shape));
    if (seed_is_phi0) {
        RSF_TRY(npy_load_device(s, seed_path, phi.p, RSF_DTYPE_F32, (int64_t)phi.bytes, &shape));
    } else {
        // phi0 = sdf(seed - 13), main.cpp:174-183
        DevBuf seed;
        const int sd = is_u8(as.descr) ? RSF_DTYPE_U8 : RSF_DTYPE_F32;
        RSF_TRY(seed.ensure((size_t)n * (sd == RSF_DTYPE_U8 ? 1 : 4)));
        RSF_TRY(npy_load_device(s, seed_path, seed.p, sd, (int64_t)seed.bytes, &shape));
        RSF_TRY(rsf_sdf_from_seed_device(c, seed.p, sd, phi.as<float>(), w, h, l, 13.0f));
    }
    RSF_TRY(rsf_run_device(c, img.p, dtype, phi.as<float>(), w, h, l, p, iter_ms));
    return npy_save_device(s, out_path, phi.as<float>(), ai.shape);
}

}  // extern "C"
