// phi0 = sdf(seed - shift) on the device: CUDA_RSF/main.cpp:174-183 subtracts
// 13 from the seed volume and takes its signed distance field (TIRA's sdf(),
// a dependency the reference does not ship).  The field is the exact
// Euclidean one, positive inside {seed - shift > 0}, with the zero level
// half-way between voxel centres: inside voxels get dist(nearest outside
// voxel) - 1/2, outside voxels -(dist(nearest inside voxel) - 1/2), the
// distances capped at the volume diagonal.  The host form is rsf_sdf_from_seed;
// both give the same bits.
//
// Separable squared EDT with one signed field for both transforms: G(v) is
// -(squared distance to the nearest outside voxel) for an inside voxel and
// +(squared distance to the nearest inside voxel) for an outside one, over the
// axes processed so far (never 0, +-kInf when the line holds one kind only).
//   pass x: one thread per row, forward and backward scans;
//   pass y, pass z: one thread per voxel, G'(c) = min over c' of
//     (|G(c')| if G(c') has v's sign else 0) + (c - c')^2,
//   scanned outward from c and stopped once (c - c')^2 >= the best so far.
// A wave holds 64 consecutive x, so every scan step is one coalesced load.
// Below a volume diagonal of 4096 voxels every squared distance is an integer
// below 2^24 (or kInf), float arithmetic on them is exact, and the result is
// the host form's bit for bit.  Larger volumes carry the squared distances in
// double (exact below 2^53) and round the final distance once to float.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "rsf_internal.hpp"

namespace rsf {
namespace {

// "no such voxel on this line"; kInf + r^2 rounds back to kInf
template <typename D>
__device__ __forceinline__ D sdf_inf() { return sizeof(D) == 4 ? (D)1e20f : (D)1e300; }
template <typename D>
__device__ __forceinline__ D dmin(D a, D b) {
    if constexpr (sizeof(D) == 4) return fminf(a, b);
    else return fmin(a, b);
}

template <typename T>
__device__ __forceinline__ bool is_inside(const T* seed, int64_t i, float shift) {
    return ((float)seed[i] - shift) > 0.0f;
}

// Pass x.  flags[z] collects bit 0 (plane z has an inside voxel) and bit 1
// (an outside voxel).
template <typename T, typename D>
__global__ void k_sdf_rows(const T* __restrict__ seed, float shift, D* __restrict__ G, int64_t w, int64_t h,
                           int64_t rows, int* __restrict__ flags) {
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    const int64_t base = row * w;
    int64_t last_in = -1, last_out = -1;
    for (int64_t x = 0; x < w; ++x) {
        const bool ins = is_inside(seed, base + x, shift);
        const int64_t other = ins ? last_out : last_in;
        G[base + x] = other >= 0 ? (D)(x - other) : (D)-1;
        if (ins) last_in = x;
        else last_out = x;
    }
    const int bits = (last_in >= 0 ? 1 : 0) | (last_out >= 0 ? 2 : 0);
    int64_t next_in = -1, next_out = -1;
    for (int64_t x = w - 1; x >= 0; --x) {
        const bool ins = is_inside(seed, base + x, shift);
        const int64_t other = ins ? next_out : next_in;
        D d = G[base + x];
        if (other >= 0 && (d < (D)0 || (D)(other - x) < d)) d = (D)(other - x);
        const D q = d < (D)0 ? sdf_inf<D>() : d * d;
        G[base + x] = ins ? -q : q;
        if (ins) next_in = x;
        else next_out = x;
    }
    const int64_t z = row / h;
    if ((flags[z] & bits) != bits) atomicOr(&flags[z], bits);
}

// Pass y (FINAL = false, stride w, out = G') or pass z (FINAL = true, stride
// w*h, out = phi).  plane_flags (pass y): a plane without both kinds of voxel
// keeps +-kInf.  scan = false: the volume lacks one kind, every line stays kInf.
template <bool FINAL, typename D>
__global__ void k_sdf_pass(const D* __restrict__ in, std::conditional_t<FINAL, float, D>* __restrict__ out, int64_t w,
                           int64_t h, int64_t n, int64_t stride, int64_t len, const int* __restrict__ plane_flags,
                           bool scan, D cap) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const D g = in[v];
        const bool neg = g < (D)0;
        D best = g < (D)0 ? -g : g;
        const int64_t c = FINAL ? v / (w * h) : (v / w) % h;
        const bool go = scan && (FINAL || plane_flags[v / (w * h)] == 3);
        if (go) {
            for (int64_t r = 1; r < len; ++r) {
                const D rr = (D)(r * r);
                if (rr >= best) break;
                const bool lo = c - r >= 0, hi = c + r < len;
                if (!lo && !hi) break;
                if (lo) {
                    const D o = in[v - r * stride];
                    best = dmin(best, (o < (D)0) == neg ? (o < (D)0 ? -o : o) + rr : rr);
                }
                if (hi) {
                    const D o = in[v + r * stride];
                    best = dmin(best, (o < (D)0) == neg ? (o < (D)0 ? -o : o) + rr : rr);
                }
            }
        }
        if constexpr (FINAL) {
            if constexpr (sizeof(D) == 4) {
                const float a = fminf(sqrtf(best), cap) - 0.5f;
                out[v] = neg ? a : -a;
            } else {
                const double a = fmin(sqrt(best), cap) - 0.5;
                out[v] = (float)(neg ? a : -a);
            }
        } else {
            out[v] = neg ? -best : best;
        }
    }
}

unsigned grid_for(int64_t n, int64_t cap) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, cap)); }

}  // namespace
}  // namespace rsf

using namespace rsf;

namespace rsf {
namespace {

// The three passes with squared distances of type D; G0 (n elements of D) may
// alias phi when D is float.
template <typename D>
int sdf_passes(hipStream_t s, const void* seed, int seed_dtype, float shift, float* phi, D* G0, D* G1, int64_t w,
               int64_t h, int64_t l, D cap, int* flags) {
    const int64_t n = w * h * l, rows = h * l;
    RSF_HIP(hipMemsetAsync(flags, 0, (size_t)l * sizeof(int), s));
    if (seed_dtype == RSF_DTYPE_U8)
        hipLaunchKernelGGL((k_sdf_rows<uint8_t, D>), dim3(grid_for(rows, INT32_MAX)), dim3(256), 0, s,
                           (const uint8_t*)seed, shift, G0, w, h, rows, flags);
    else
        hipLaunchKernelGGL((k_sdf_rows<float, D>), dim3(grid_for(rows, INT32_MAX)), dim3(256), 0, s,
                           (const float*)seed, shift, G0, w, h, rows, flags);
    RSF_HIP(hipGetLastError());
    std::vector<int> pf((size_t)l);
    RSF_HIP(hipMemcpyAsync(pf.data(), flags, (size_t)l * sizeof(int), hipMemcpyDeviceToHost, s));
    RSF_HIP(hipStreamSynchronize(s));
    int all = 0;
    for (int f : pf) all |= f;
    const bool scan = all == 3;  // otherwise no distance is finite
    const unsigned g = grid_for(n, 1 << 20);
    hipLaunchKernelGGL((k_sdf_pass<false, D>), dim3(g), dim3(256), 0, s, G0, G1, w, h, n, w, h, flags, scan, cap);
    RSF_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_sdf_pass<true, D>), dim3(g), dim3(256), 0, s, G1, phi, w, h, n, w * h, l, flags, scan, cap);
    RSF_HIP(hipGetLastError());
    RSF_HIP(hipStreamSynchronize(s));
    return RSF_OK;
}

}  // namespace
}  // namespace rsf

extern "C" int rsf_sdf_from_seed_device(rsf_context* c, const void* seed, int seed_dtype, float* phi, int64_t w,
                                        int64_t h, int64_t l, float shift) {
    hipStream_t s;
    RSF_TRY(context_stream(c, &s));
    if (!seed || !phi) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (seed_dtype != RSF_DTYPE_U8 && seed_dtype != RSF_DTYPE_F32)
        return set_error(RSF_ERR_INVALID_ARGUMENT, "seed dtype must be RSF_DTYPE_U8 or RSF_DTYPE_F32");
    if (w <= 0 || h <= 0 || l <= 0 || w * h * l >= (int64_t(1) << 40))
        return set_error(RSF_ERR_INVALID_ARGUMENT, "bad dims");
    const int64_t n = w * h * l;
    DevBuf flags;
    RSF_TRY(flags.ensure((size_t)l * sizeof(int)));
    // the volume diagonal, evaluated as the host form does
    const float diag2 = (float)w * (float)w + (float)h * (float)h + (float)l * (float)l;
    if (diag2 < 16777216.0f) {
        DevBuf tmp;
        RSF_TRY(tmp.ensure((size_t)n * sizeof(float)));
        return sdf_passes<float>(s, seed, seed_dtype, shift, phi, phi, tmp.as<float>(), w, h, l, std::sqrt(diag2),
                                 flags.as<int>());
    }
    DevBuf g0, g1;
    RSF_TRY(g0.ensure((size_t)n * sizeof(double)));
    RSF_TRY(g1.ensure((size_t)n * sizeof(double)));
    const double cap = std::sqrt((double)w * (double)w + (double)h * (double)h + (double)l * (double)l);
    return sdf_passes<double>(s, seed, seed_dtype, shift, phi, g0.as<double>(), g1.as<double>(), w, h, l, cap,
                              flags.as<int>());
}

extern "C" int rsf_sdf_from_seed_host(rsf_context* c, const float* seed, float* phi, int64_t w, int64_t h, int64_t l,
                                      float shift) {
    hipStream_t s;
    RSF_TRY(context_stream(c, &s));
    if (!seed || !phi) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (w <= 0 || h <= 0 || l <= 0 || w * h * l >= (int64_t(1) << 40))
        return set_error(RSF_ERR_INVALID_ARGUMENT, "bad dims");
    const size_t bytes = (size_t)(w * h * l) * sizeof(float);
    DevBuf ds, dp;
    RSF_TRY(ds.ensure(bytes));
    RSF_TRY(dp.ensure(bytes));
    RSF_HIP(hipMemcpyAsync(ds.p, seed, bytes, hipMemcpyHostToDevice, s));
    RSF_TRY(rsf_sdf_from_seed_device(c, ds.p, RSF_DTYPE_F32, dp.as<float>(), w, h, l, shift));
    RSF_HIP(hipMemcpyAsync(phi, dp.p, bytes, hipMemcpyDeviceToHost, s));
    RSF_HIP(hipStreamSynchronize(s));
    return RSF_OK;
}
