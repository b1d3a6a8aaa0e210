// Per-op kernels: one-voxel-per-thread restatements of each reference kernel
// (CUDA_RSF/kernel.cu, cited per kernel), without FMA contraction, plus the
// C-ABI mirrors of the reference's host wrappers.  These are the building
// blocks of the unfused device pipeline and of the per-function parity tests;
// the fused sweeps (rsf_sweep.hip) are the fast path.
#include <mutex>

#include "rsf_internal.hpp"

namespace rsf {
namespace {

constexpr int kThreads = 256;

inline dim3 grid_for(int64_t n) {
    int64_t b = (n + kThreads - 1) / kThreads;
    if (b > (1 << 20)) b = 1 << 20;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

#define GRID_STRIDE(i, n) \
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

#pragma clang fp contract(off)

__global__ void k_heaviside(float* out, const float* in, int64_t n, float eps) {   // kernel.cu:248
    GRID_STRIDE(i, n) out[i] = heav_literal(in[i], eps);
}
__global__ void k_deri(float* out, const float* in, int64_t n, float eps) {        // kernel.cu:265
    GRID_STRIDE(i, n) out[i] = dirac_literal(in[i], eps);
}
__global__ void k_mul(float* out, const float* a, const float* b, int64_t n) {      // kernel.cu:210
    GRID_STRIDE(i, n) out[i] = a[i] * b[i];
}
__global__ void k_inverse(float* out, const float* a, int64_t n) {                   // kernel.cu:282
    GRID_STRIDE(i, n) out[i] = 1.0f - a[i];
}
__global__ void k_plus(float* out, const float* a, const float* b, const float* c, int64_t n) {  // :228
    GRID_STRIDE(i, n) out[i] = a[i] + b[i] + c[i];
}
__global__ void k_div(float* out, const float* a, const float* b, int64_t n) {       // kernel.cu:191
    GRID_STRIDE(i, n) out[i] = a[i] / b[i];
}
__global__ void k_div_grad(float* out, const float* g, const float* M, int64_t n) {  // kernel.cu:173
    GRID_STRIDE(i, n) out[i] = (float)((double)g[i] / ((double)M[i] + 1e-6));
}
__global__ void k_gradmag(float* M, const float* gx, const float* gy, const float* gz, int64_t n) {  // :914
    GRID_STRIDE(i, n) M[i] = sqrtf(gx[i] * gx[i] + gy[i] * gy[i] + gz[i] * gz[i]);
}
__global__ void k_force(float* F, const float* Eo, const float* Ei, float l1, float l2, int64_t n) {
    GRID_STRIDE(i, n) F[i] = (l1 == l2) ? l1 * (Eo[i] - Ei[i]) : (l1 * Eo[i] - l2 * Ei[i]);
}
__global__ void k_to_float(float* out, const void* img, int u8, int64_t n) {
    GRID_STRIDE(i, n) out[i] = u8 ? (float)((const uint8_t*)img)[i] : ((const float*)img)[i];
}

// kernel.cu:57-125: dx and dz difference in double, dy in float; x = 0 reads
// the previous linear element, linear index 0 reads itself.
__global__ void k_gradient(float* gx, float* gy, float* gz, const float* f, int w, int h, int l) {
    const int64_t plane = (int64_t)w * h, n = plane * l;
    GRID_STRIDE(i, n) {
        const int x = (int)(i % w);
        const int y = (int)((i / w) % h);
        const int z = (int)(i / plane);
        if (gx) {
            const int64_t r = (x + 1 >= w) ? i - x + (w - 1) : i + 1;
            const int64_t lft = (i == 0) ? 0 : i - 1;
            gx[i] = (float)(((double)f[r] - (double)f[lft]) / 2.0);
        }
        if (gy) {
            int yl = y - 1, yr = y + 1;
            if (yl < 0) { yl = 0; yr = 1; }
            else if (yr >= h) { yr = h - 1; yl = yr - 1; }
            gy[i] = (f[i + (int64_t)(yr - y) * w] - f[i + (int64_t)(yl - y) * w]) / 2.0f;
        }
        if (gz) {
            const int zl = (z > 0) ? z - 1 : 0, zr = (z < l - 1) ? z + 1 : l - 1;
            gz[i] = (float)(((double)f[i + (int64_t)(zr - z) * plane] -
                             (double)f[i + (int64_t)(zl - z) * plane]) / 2.0);
        }
    }
}

// kernel.cu:1000-1049: valid correlation along one axis, k = 0..P-1 in order.
__global__ void k_conv_axis(float* out, const float* in, int w, int h, int l, int axis, Taps t, int P) {
    const int ow = axis == 0 ? w - P + 1 : w;
    const int oh = axis == 1 ? h - P + 1 : h;
    const int ol = axis == 2 ? l - P + 1 : l;
    const int64_t n = (int64_t)ow * oh * ol;
    const int64_t stride = axis == 0 ? 1 : (axis == 1 ? (int64_t)w : (int64_t)w * h);
    GRID_STRIDE(o, n) {
        const int x = (int)(o % ow);
        const int y = (int)((o / ow) % oh);
        const int z = (int)(o / ((int64_t)ow * oh));
        const float* src = in + ((int64_t)z * h + y) * w + x;
        float c = 0.0f;
        for (int k = 0; k < P; ++k) c = fmaf(src[k * stride], t.g[k], c);
        out[o] = c;
    }
}

// kernel.cu:300-325.
__global__ void k_pad(float* out, const float* in, int vw, int vh, int vl, int pad) {
    const int pw = vw + 2 * pad, ph = vh + 2 * pad, pl = vl + 2 * pad;
    const int64_t n = (int64_t)pw * ph * pl;
    GRID_STRIDE(o, n) {
        const int j = (int)(o % pw);
        const int i = (int)((o / pw) % ph);
        const int k = (int)(o / ((int64_t)pw * ph));
        if (j < pad || j >= vw + pad || i < pad || i >= vh + pad || k < pad || k >= vl + pad)
            out[o] = 30.0f;
        else
            out[o] = in[((int64_t)(k - pad) * vh + (i - pad)) * vw + (j - pad)];
    }
}

// kernel.cu:25-55: window half-width s, double weight 1/(2s+1)^2, the running
// float sums rounded through a double add at every neighbour.
__global__ void k_energies(float* Eo, float* Ei, const float* fo, const float* fi, const float* I,
                           int w, int h, int l, int s) {
    const int64_t plane = (int64_t)w * h, n = plane * l;
    const double wgt = 1.0 / ((double)(2 * s + 1) * (double)(2 * s + 1));
    GRID_STRIDE(c, n) {
        const int j = (int)(c % w);
        const int i = (int)((c / w) % h);
        const int p = (int)(c / plane);
        const float fc = fo[c], gc = fi[c];
        float s1 = 0.f, s2 = 0.f;
        for (int u = -s; u <= s; ++u)
            for (int v = -s; v <= s; ++v)
                for (int q = -s; q <= s; ++q) {
                    const int ii = i + u, jj = j + v, pp = p + q;
                    if (ii < 0 || ii >= h || jj < 0 || jj >= w || pp < 0 || pp >= l) continue;
                    const float In = I[(int64_t)pp * plane + (int64_t)ii * w + jj];
                    const float d1 = In - fc, d2 = In - gc;
                    s1 = (float)((double)s1 + wgt * (double)d1 * (double)d1);
                    s2 = (float)((double)s2 + wgt * (double)d2 * (double)d2);
                }
        Eo[c] = s1;
        Ei[c] = s2;
    }
}

// kernel.cu:365-387, left-associative.
__global__ void k_final(float* out, const float* phi, const float* dh, const float* Eo, const float* Ei,
                        const float* kap, const float* L, float l1, float l2, float dt, float v,
                        float mu, int64_t n) {
    GRID_STRIDE(i, n) {
        const float d = dh[i], k = kap[i];
        const float F = (l1 == l2) ? l1 * (Eo[i] - Ei[i]) : (l1 * Eo[i] - l2 * Ei[i]);
        const float t1 = (d * F) * dt;
        const float t2 = ((v * dt) * d) * k;
        const float t3 = mu * (dt * (L[i] - k));
        out[i] = ((phi[i] - t1) + t2) + t3;
    }
}

#pragma clang fp contract(on)

#define LAUNCH(kernel, n, s, ...)                                                   \
    do {                                                                            \
        if ((n) <= 0) return RSF_OK;                                                \
        hipLaunchKernelGGL(kernel, grid_for(n), dim3(kThreads), 0, s, __VA_ARGS__); \
        RSF_HIP(hipGetLastError());                                                 \
        return RSF_OK;                                                              \
    } while (0)

}  // namespace

namespace ops {
int heaviside(float* out, const float* in, int64_t n, float eps, hipStream_t s) { LAUNCH(k_heaviside, n, s, out, in, n, eps); }
int deri_heaviside(float* out, const float* in, int64_t n, float eps, hipStream_t s) { LAUNCH(k_deri, n, s, out, in, n, eps); }
int multiplication(float* out, const float* a, const float* b, int64_t n, hipStream_t s) { LAUNCH(k_mul, n, s, out, a, b, n); }
int inverse(float* out, const float* a, int64_t n, hipStream_t s) { LAUNCH(k_inverse, n, s, out, a, n); }
int plus(float* out, const float* a, const float* b, const float* c, int64_t n, hipStream_t s) { LAUNCH(k_plus, n, s, out, a, b, c, n); }
int division(float* out, const float* a, const float* b, int64_t n, hipStream_t s) { LAUNCH(k_div, n, s, out, a, b, n); }
int division_grad(float* out, const float* g, const float* M, int64_t n, hipStream_t s) { LAUNCH(k_div_grad, n, s, out, g, M, n); }
int gradientmagnitude(float* M, const float* gx, const float* gy, const float* gz, int64_t n, hipStream_t s) {
    LAUNCH(k_gradmag, n, s, M, gx, gy, gz, n);
}
int gradient(float* gx, float* gy, float* gz, const float* f, int w, int h, int l, hipStream_t s) {
    const int64_t n = (int64_t)w * h * l;
    LAUNCH(k_gradient, n, s, gx, gy, gz, f, w, h, l);
}
int conv_axis(float* out, const float* in, int w, int h, int l, int axis, const Taps& t, int P, hipStream_t s) {
    const int64_t n = (int64_t)(axis == 0 ? w - P + 1 : w) * (axis == 1 ? h - P + 1 : h) *
                      (axis == 2 ? l - P + 1 : l);
    LAUNCH(k_conv_axis, n, s, out, in, w, h, l, axis, t, P);
}
int pad_border(float* out, const float* in, int vw, int vh, int vl, int pad, hipStream_t s) {
    const int64_t n = (int64_t)(vw + 2 * pad) * (vh + 2 * pad) * (vl + 2 * pad);
    LAUNCH(k_pad, n, s, out, in, vw, vh, vl, pad);
}
int energies(float* Eo, float* Ei, const float* fo, const float* fi, const float* I, int w, int h, int l,
             int sw, hipStream_t s) {
    const int64_t n = (int64_t)w * h * l;
    LAUNCH(k_energies, n, s, Eo, Ei, fo, fi, I, w, h, l, sw);
}
int final_phi(float* out, const float* phi, const float* dh, const float* Eo, const float* Ei, const float* kap,
              const float* L, float l1, float l2, float dt, float nu, float mu, int64_t n, hipStream_t s) {
    LAUNCH(k_final, n, s, out, phi, dh, Eo, Ei, kap, L, l1, l2, dt, nu, mu, n);
}
int force(float* F, const float* Eo, const float* Ei, float l1, float l2, int64_t n, hipStream_t s) {
    LAUNCH(k_force, n, s, F, Eo, Ei, l1, l2, n);
}
int to_float(float* out, const void* img, int u8, int64_t n, hipStream_t s) { LAUNCH(k_to_float, n, s, out, img, u8, n); }

int blur_pad(float* out, const float* in, float* tmp1, float* tmp2, int w, int h, int l, const Taps& t, int P,
             int m, hipStream_t s) {
    const int ow = w - P + 1, oh = h - P + 1, ol = l - P + 1;
    RSF_TRY(conv_axis(tmp1, in, w, h, l, 0, t, P, s));
    RSF_TRY(conv_axis(tmp2, tmp1, ow, h, l, 1, t, P, s));
    RSF_TRY(conv_axis(tmp1, tmp2, ow, oh, l, 2, t, P, s));
    return pad_border(out, tmp1, ow, oh, ol, m, s);
}
}  // namespace ops

// ------------------------------------------------ reference-wrapper mirrors
namespace {

// Process-wide state of the mirrors: device 0 and a private stream, with
// device buffers kept between calls (the reference allocates and frees on
// every call, kernel.cu:140-168 and the like).
struct MirrorState {
    std::mutex mu;
    hipStream_t stream = nullptr;
    DevBuf buf[8];
    int init() {
        if (stream) return RSF_OK;
        RSF_HIP(hipSetDevice(0));
        RSF_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        return RSF_OK;
    }
};
MirrorState& mirror() {
    static MirrorState* st = new MirrorState();  // never destroyed: no teardown-order issues
    return *st;
}

int check_dims3(int w, int h, int l) {
    if (w <= 0 || h <= 0 || l <= 0) return set_error(RSF_ERR_INVALID_ARGUMENT, "dims must be positive (%d,%d,%d)", w, h, l);
    return RSF_OK;
}

// Copies each host input in, runs fn on device pointers d[] = inputs, then
// outputs, then `nscratch` scratch buffers of `scratch_elems` floats, copies
// the outputs back and waits.
template <class Fn>
int mirror_call(std::initializer_list<std::pair<const float*, int64_t>> ins,
                std::initializer_list<std::pair<float*, int64_t>> outs, int nscratch,
                int64_t scratch_elems, Fn fn) {
    MirrorState& st = mirror();
    std::lock_guard<std::mutex> lk(st.mu);
    RSF_TRY(st.init());
    float* d[8];
    int k = 0;
    if ((int)(ins.size() + outs.size()) + nscratch > 8)
        return set_error(RSF_ERR_INVALID_ARGUMENT, "too many buffers");
    for (auto& in : ins) {
        if (!in.first) return set_error(RSF_ERR_INVALID_ARGUMENT, "null input pointer");
        RSF_TRY(st.buf[k].ensure((size_t)in.second * sizeof(float)));
        d[k] = st.buf[k].as<float>();
        RSF_HIP(hipMemcpyAsync(d[k], in.first, (size_t)in.second * sizeof(float), hipMemcpyHostToDevice, st.stream));
        ++k;
    }
    const int first_out = k;
    for (auto& out : outs) {
        if (!out.first) return set_error(RSF_ERR_INVALID_ARGUMENT, "null output pointer");
        RSF_TRY(st.buf[k].ensure((size_t)out.second * sizeof(float)));
        d[k] = st.buf[k].as<float>();
        ++k;
    }
    for (int q = 0; q < nscratch; ++q, ++k) {
        RSF_TRY(st.buf[k].ensure((size_t)scratch_elems * sizeof(float)));
        d[k] = st.buf[k].as<float>();
    }
    RSF_TRY(fn(d, st.stream));
    k = first_out;
    for (auto& out : outs) {
        RSF_HIP(hipMemcpyAsync(out.first, d[k], (size_t)out.second * sizeof(float), hipMemcpyDeviceToHost, st.stream));
        ++k;
    }
    RSF_HIP(hipStreamSynchronize(st.stream));
    return RSF_OK;
}

}  // namespace
}  // namespace rsf

using namespace rsf;

extern "C" {

int rsf_heaviside_cal(float* output, float* input, int img_w, int img_h, int img_l) {
    RSF_TRY(check_dims3(img_w, img_h, img_l));
    const int64_t n = (int64_t)img_w * img_h * img_l;
    return mirror_call({{input, n}}, {{output, n}}, 0, 0,
                       [&](float** d, hipStream_t s) { return ops::heaviside(d[1], d[0], n, 0.2f, s); });
}

int rsf_deri_heaviside_cal(float* output, float* input, int img_w, int img_h, int img_l) {
    RSF_TRY(check_dims3(img_w, img_h, img_l));
    const int64_t n = (int64_t)img_w * img_h * img_l;
    return mirror_call({{input, n}}, {{output, n}}, 0, 0,
                       [&](float** d, hipStream_t s) { return ops::deri_heaviside(d[1], d[0], n, 0.2f, s); });
}

int rsf_inverse_cal(float* output, float* input, int img_w, int img_h, int img_l) {
    RSF_TRY(check_dims3(img_w, img_h, img_l));
    const int64_t n = (int64_t)img_w * img_h * img_l;
    return mirror_call({{input, n}}, {{output, n}}, 0, 0,
                       [&](float** d, hipStream_t s) { return ops::inverse(d[1], d[0], n, s); });
}

int rsf_multiplication_cal(float* output, float* input_1, float* input_2, int img_w, int img_h, int img_l) {
    RSF_TRY(check_dims3(img_w, img_h, img_l));
    const int64_t n = (int64_t)img_w * img_h * img_l;
    return mirror_call({{input_1, n}, {input_2, n}}, {{output, n}}, 0, 0,
                       [&](float** d, hipStream_t s) { return ops::multiplication(d[2], d[0], d[1], n, s); });
}

int rsf_division_cal(float* output, float* input_1, float* input_2, int img_w, int img_h, int img_l) {
    RSF_TRY(check_dims3(img_w, img_h, img_l));
    const int64_t n = (int64_t)img_w * img_h * img_l;
    return mirror_call({{input_1, n}, {input_2, n}}, {{output, n}}, 0, 0,
                       [&](float** d, hipStream_t s) { return ops::division(d[2], d[0], d[1], n, s); });
}

int rsf_gradient_cal(float* input, float* output_x, float* output_y, float* output_z, int img_w, int img_h,
                     int img_l) {
    RSF_TRY(check_dims3(img_w, img_h, img_l));
    if (img_w < 2 || img_h < 2 || img_l < 2) return set_error(RSF_ERR_INVALID_ARGUMENT, "gradient needs dims >= 2");
    const int64_t n = (int64_t)img_w * img_h * img_l;
    return mirror_call({{input, n}}, {{output_x, n}, {output_y, n}, {output_z, n}}, 0, 0, [&](float** d, hipStream_t s) {
        return ops::gradient(d[1], d[2], d[3], d[0], img_w, img_h, img_l, s);
    });
}

int rsf_gradientmagnitude_cal(float* output, float* input_x, float* input_y, float* input_z, int img_w,
                              int img_h, int img_l) {
    RSF_TRY(check_dims3(img_w, img_h, img_l));
    const int64_t n = (int64_t)img_w * img_h * img_l;
    return mirror_call({{input_x, n}, {input_y, n}, {input_z, n}}, {{output, n}}, 0, 0, [&](float** d, hipStream_t s) {
        return ops::gradientmagnitude(d[3], d[0], d[1], d[2], n, s);
    });
}

int rsf_division_nDIV(float* input_x, float* input_y, float* input_z, float* gradientmagnitude, float* output_x,
                      float* output_y, float* output_z, int img_w, int img_h, int img_l) {
    RSF_TRY(check_dims3(img_w, img_h, img_l));
    const int64_t n = (int64_t)img_w * img_h * img_l;
    return mirror_call({{input_x, n}, {input_y, n}, {input_z, n}, {gradientmagnitude, n}},
                       {{output_x, n}, {output_y, n}, {output_z, n}}, 0, 0, [&](float** d, hipStream_t s) {
                           RSF_TRY(ops::division_grad(d[4], d[0], d[3], n, s));
                           RSF_TRY(ops::division_grad(d[5], d[1], d[3], n, s));
                           return ops::division_grad(d[6], d[2], d[3], n, s);
                       });
}

int rsf_gradient_cal_3_input(float* input_x, float* input_y, float* input_z, float* output_x, float* output_y,
                             float* output_z, int img_w, int img_h, int img_l) {
    RSF_TRY(check_dims3(img_w, img_h, img_l));
    if (img_w < 2 || img_h < 2 || img_l < 2) return set_error(RSF_ERR_INVALID_ARGUMENT, "gradient needs dims >= 2");
    const int64_t n = (int64_t)img_w * img_h * img_l;
    return mirror_call({{input_x, n}, {input_y, n}, {input_z, n}}, {{output_x, n}, {output_y, n}, {output_z, n}}, 0, 0,
                       [&](float** d, hipStream_t s) {
                           RSF_TRY(ops::gradient(d[3], nullptr, nullptr, d[0], img_w, img_h, img_l, s));
                           RSF_TRY(ops::gradient(nullptr, d[4], nullptr, d[1], img_w, img_h, img_l, s));
                           return ops::gradient(nullptr, nullptr, d[5], d[2], img_w, img_h, img_l, s);
                       });
}

int rsf_normalized_div_plus(float* normalize_DIV, float* input_x, float* input_y, float* input_z, int img_w,
                            int img_h, int img_l) {
    RSF_TRY(check_dims3(img_w, img_h, img_l));
    const int64_t n = (int64_t)img_w * img_h * img_l;
    return mirror_call({{input_x, n}, {input_y, n}, {input_z, n}}, {{normalize_DIV, n}}, 0, 0,
                       [&](float** d, hipStream_t s) { return ops::plus(d[3], d[0], d[1], d[2], n, s); });
}

int rsf_DIV_cal(float* DIV, float* input_x, float* input_y, float* input_z, int img_w, int img_h, int img_l) {
    RSF_TRY(check_dims3(img_w, img_h, img_l));
    if (img_w < 2 || img_h < 2 || img_l < 2) return set_error(RSF_ERR_INVALID_ARGUMENT, "gradient needs dims >= 2");
    const int64_t n = (int64_t)img_w * img_h * img_l;
    return mirror_call({{input_x, n}, {input_y, n}, {input_z, n}}, {{DIV, n}}, 3, n,
                       [&](float** d, hipStream_t s) {
                           float *tx = d[4], *ty = d[5], *tz = d[6];
                           RSF_TRY(ops::gradient(tx, nullptr, nullptr, d[0], img_w, img_h, img_l, s));
                           RSF_TRY(ops::gradient(nullptr, ty, nullptr, d[1], img_w, img_h, img_l, s));
                           RSF_TRY(ops::gradient(nullptr, nullptr, tz, d[2], img_w, img_h, img_l, s));
                           return ops::plus(d[3], tx, ty, tz, n, s);
                       });
}

int rsf_adddevice_convolution_seperable(float* output, float* in_img, int img_w, int img_h, int img_l,
                                        float sigma, float* gkernel, unsigned int k_size) {
    (void)sigma;  // unused by the reference too (kernel.cu:1051)
    RSF_TRY(check_dims3(img_w, img_h, img_l));
    if (!gkernel || k_size == 0 || (int)k_size > kMaxTaps)
        return set_error(RSF_ERR_INVALID_ARGUMENT, "k_size must be in 1..%d", kMaxTaps);
    const int P = (int)k_size;
    if (img_w < P || img_h < P || img_l < P) return set_error(RSF_ERR_INVALID_ARGUMENT, "dims smaller than the kernel");
    Taps t{};
    for (int k = 0; k < P; ++k) t.g[k] = gkernel[k];
    const int64_t n = (int64_t)img_w * img_h * img_l;
    const int ow = img_w - P + 1, oh = img_h - P + 1, ol = img_l - P + 1;
    const int64_t nout = (int64_t)ow * oh * ol;
    return mirror_call({{in_img, n}}, {{output, nout}}, 2, (int64_t)ow * img_h * img_l,
                       [&](float** d, hipStream_t s) {
                           RSF_TRY(ops::conv_axis(d[2], d[0], img_w, img_h, img_l, 0, t, P, s));
                           RSF_TRY(ops::conv_axis(d[3], d[2], ow, img_h, img_l, 1, t, P, s));
                           return ops::conv_axis(d[1], d[3], ow, oh, img_l, 2, t, P, s);
                       });
}

int rsf_pad_border_cal(float* output, float* input, int img_w, int img_h, int img_l, int pad_w) {
    RSF_TRY(check_dims3(img_w, img_h, img_l));
    if (pad_w < 0) return set_error(RSF_ERR_INVALID_ARGUMENT, "pad_w must be >= 0");
    const int64_t n = (int64_t)img_w * img_h * img_l;
    const int64_t nout = (int64_t)(img_w + 2 * pad_w) * (img_h + 2 * pad_w) * (img_l + 2 * pad_w);
    return mirror_call({{input, n}}, {{output, nout}}, 0, 0, [&](float** d, hipStream_t s) {
        return ops::pad_border(d[1], d[0], img_w, img_h, img_l, pad_w, s);
    });
}

int rsf_adddevice(float* input, float* fout, float* fin, float* Eo, float* Ei, int img_w, int img_h, int img_l,
                  float sigma) {
    RSF_TRY(check_dims3(img_w, img_h, img_l));
    const int64_t n = (int64_t)img_w * img_h * img_l;
    const int s_int = (int)sigma;  // float -> int parameter conversion, kernel.cu:1211
    if (s_int < 0) return set_error(RSF_ERR_INVALID_ARGUMENT, "sigma2 must be >= 0");
    return mirror_call({{input, n}, {fout, n}, {fin, n}}, {{Eo, n}, {Ei, n}}, 0, 0, [&](float** d, hipStream_t s) {
        return ops::energies(d[3], d[4], d[1], d[2], d[0], img_w, img_h, img_l, s_int, s);
    });
}

int rsf_Final_PHI_cal(float* output, float* phi, float* deri_heav, float* Eout, float* Ein, float* normalized_DIV,
                      float* DIV, float e, float dt, float v, float meu, int img_w, int img_h, int img_l) {
    RSF_TRY(check_dims3(img_w, img_h, img_l));
    const int64_t n = (int64_t)img_w * img_h * img_l;
    return mirror_call({{phi, n}, {deri_heav, n}, {Eout, n}, {Ein, n}, {normalized_DIV, n}, {DIV, n}},
                       {{output, n}}, 0, 0, [&](float** d, hipStream_t s) {
                           return ops::final_phi(d[6], d[0], d[1], d[2], d[3], d[4], d[5], e, e, dt, v, meu, n, s);
                       });
}

}  // extern "C"
