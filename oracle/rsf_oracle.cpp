/*
 * Oracle: CPU restatement of the reference RSF iteration -- TEST INFRASTRUCTURE.
 *
 * Only the test suite, the smoke check and the benchmark's CPU-baseline leg may
 * load this file's library, as the checker or as the timed CPU baseline; the
 * product path (paper-rsf-vesselseg_amd/) never links or calls it.
 *
 * Function -> reference source it restates -> what pins it (tests/):
 *   taps_impl            main.cpp:121-166          closed form 1/sqrt(u*pi)*exp(-(k-miu)^2/u),
 *                                                  P/miu/u integer rules (test_oracle_taps)
 *   heaviside            kernel.cu:248-263,494-524 H(0)=0.5, H(+-10) known answers, odd symmetry
 *   deri_heaviside       kernel.cu:265-280,558-588 delta(0)=1.59154928, delta(10) known answer,
 *                                                  derivative of H by finite differences
 *   multiplication/inverse/plus/division
 *                        kernel.cu:191-245,282-298 exact IEEE identities on random data
 *   division_grad        kernel.cu:173-189         FP64 divide-then-round vs numpy
 *   gradientmagnitude    kernel.cu:914-933         numpy restatement (FP64)
 *   gradient             kernel.cu:57-171          linear ramp closed forms incl. the x = 0
 *                                                  wrap into the previous row; independent
 *                                                  numpy restatement in linear-index form
 *   conv_separable       kernel.cu:1000-1121       brute-force 3-D correlation with the
 *                                                  outer-product kernel (FP64); constant
 *                                                  input -> c*(sum g)^3; delta -> g x g x g
 *   pad_border           kernel.cu:300-362         band == 30.0f, centre copy (numpy)
 *   energies             kernel.cu:25-55,1178-1236 s = 0 pointwise square; s > 0 brute force
 *   final_phi            kernel.cu:365-387         scalar closed form
 *   step (FP32)          main.cpp:222-300          independent numpy FP64 restatement of the
 *                                                  whole iteration; constant image -> zero
 *                                                  fitting force; sphere SDF curvature ~ -2/r;
 *                                                  committed golden fixtures
 *   step_f64             same                      agrees with the FP32 step to FP32 rounding
 *   step_adjoint_f64     derivative of step_f64    FP64 central finite differences; transpose
 *                                                  identities <Df,u> = <f,D^T u>
 *   det_atan             (atanf of kernel.cu:258)  numpy arctan: relative error < 4e-16, and
 *                                                  the float rounding equals numpy's on dense
 *                                                  samples
 * Parity unpinned (no reference output exists for them): the value the
 * reference reads for voxel (0,0,0) at index -1 (defined here as the clamp,
 * i.e. voxel 0 itself), nvcc's choice of FMA contraction outside the
 * correlation loops, and CUDA's atanf/sqrtf/expf last-ulp results (the
 * oracle uses the correctly rounded atanf, IEEE sqrtf and glibc expf).
 */
#include "rsf_oracle.h"

#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = 3.14159265358979323846;  // kernel.cu:14

inline int64_t nvox(int w, int h, int l) { return (int64_t)w * h * l; }
inline int64_t at(int x, int y, int z, int w, int h) { return ((int64_t)z * h + y) * w + x; }

/* main.cpp:152-166 (tap count, centre, unnormalised samples) and
 * main.cpp:264,273 (pad width).  FP32 throughout, as in the reference. */
int taps_impl(float sigma, float* g, int cap, int* P_out, int* m_out) {
    if (!(sigma > 0.0f)) return -1;
    float s7 = 7 * sigma;
    unsigned int p = (unsigned int)s7;             // main.cpp:152
    int k_size = (int)s7;                          // main.cpp:264
    if (p % 2 == 0) return -1;                     // even floor(7 sigma): buffers mis-sized
    if ((int)p > cap) return -1;
    float miu = (float)(p / 2);                    // main.cpp:155, integer division
    int u = (int)(2 * sigma * sigma);              // main.cpp:164
    if (u <= 0) return -1;
    for (int xi = 0; xi < (int)p; xi++) {          // main.cpp:163-166
        float d = (float)xi - miu;
        float ex = (-d) * d / (float)u;
        g[xi] = 1.0f / sqrtf((float)u * (float)kPi) * expf(ex);
    }
    *P_out = (int)p;
    *m_out = (k_size - 1) / 2;                     // main.cpp:273 (int division)
    return 0;
}

/* ---------------------------------------------------------------- FP32 ops */

/* The float arctangent.  CUDA's atanf (kernel.cu:258) is a vendor routine whose
 * last-ulp results no output of the reference pins; the oracle takes the
 * correctly rounded atanf instead: an arctangent in double precision from IEEE
 * +, -, *, / only (the rational approximation of S. L. Moshier's Cephes atan.c
 * after its range reduction: x > tan(3pi/8): pi/2 - atan(1/x); x > 0.66:
 * pi/4 + atan((x-1)/(x+1))), rounded once to float.  Built without
 * contraction, so any IEEE machine evaluating this sequence gets these bits.
 * Pinned against numpy's arctan (tests/test_oracle.py). */
double det_atan(double x) {
    const double ax = x < 0 ? -x : x;
    double y, v;
    int flag;
    if (ax > 2.41421356237309504880) {          /* tan(3 pi / 8) */
        y = 1.57079632679489661923; flag = 1; v = -(1.0 / ax);
    } else if (ax <= 0.66) {
        y = 0.0; flag = 0; v = ax;
    } else {
        y = 7.85398163397448309616e-1; flag = 2; v = (ax - 1.0) / (ax + 1.0);
    }
    const double z = v * v;
    const double pz = (((-8.750608600031904122785e-1 * z - 1.615753718733365076637e1) * z
                        - 7.500855792314704667340e1) * z - 1.228866684490136173410e2) * z
                      - 6.485021904942025371773e1;
    const double qz = ((((z + 2.485846490142306297962e1) * z + 1.650270098316988542046e2) * z
                        + 4.328810604912902668951e2) * z + 4.853903996359136964868e2) * z
                      + 1.945506571482613964425e2;
    double r = v * (z * pz / qz) + v;
    if (flag == 2) r = r + 0.5 * 6.123233995736765886130e-17;
    else if (flag == 1) r = r + 6.123233995736765886130e-17;
    y = y + r;
    return x < 0 ? -y : y;
}
inline float det_atanf(float x) { return (float)det_atan((double)x); }

void heaviside_f(float* out, const float* in, int64_t n, float eps) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)   // kernel.cu:258-259: double 2/pi, float atan
        out[i] = (float)(0.5 * (1 + (2 / kPi) * (double)det_atanf(in[i] / eps)));
}

void deri_heaviside_f(float* out, const float* in, int64_t n, float eps) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)   // kernel.cu:275-276: float quotient, double 1/pi
        out[i] = (float)((1 / kPi) * (double)(eps / ((eps * eps) + (in[i] * in[i]))));
}

/* kernel.cu:57-125.  dx: `size_t j_left = j - 1` wraps at x = 0, so the left
 * operand is the previous linear element; at linear index 0 it is defined
 * here as element 0.  dx and dz difference in double, dy in float. */
void gradient_f(float* gx, float* gy, float* gz, const float* f, int w, int h, int l) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < l; z++)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                int64_t i = at(x, y, z, w, h);
                /* dx: j_left = j-1 in size_t, so the left operand is always the
                 * previous linear element (kernel.cu:64-67, the `< 0` test is
                 * dead); j_right clamps to w-1 (:68-71). */
                int64_t il = i - 1;
                if (il < 0) il = 0;                              // (0,0,0): clamp
                int64_t ir = (x + 1 >= w) ? at(w - 1, y, z, w, h) : i + 1;
                gx[i] = (float)(((double)f[ir] - (double)f[il]) / 2.0);
                /* dy */
                int yl = y - 1, yr = y + 1;
                if (yl < 0) { yl = 0; yr = 1; }
                else if (yr >= h) { yr = h - 1; yl = yr - 1; }
                gy[i] = (f[at(x, yr, z, w, h)] - f[at(x, yl, z, w, h)]) / 2.0f;
                /* dz */
                int zl = (z > 0) ? z - 1 : 0;
                int zr = (z < l - 1) ? z + 1 : l - 1;
                gz[i] = (float)(((double)f[at(x, y, zr, w, h)] - (double)f[at(x, y, zl, w, h)]) / 2.0);
            }
}

/* kernel.cu:1000-1049 launched as in :1051-1121.  Valid correlation along x,
 * then y, then z; each output accumulates k = 0..P-1 in order with fmaf. */
void conv_separable_f(float* out, const float* in, int w, int h, int l, const float* g, int P) {
    int ow = w - P + 1, oh = h - P + 1, ol = l - P + 1;
    std::vector<float> tx((size_t)ow * h * l), ty((size_t)ow * oh * l);
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < l; z++)
        for (int y = 0; y < h; y++) {
            const float* row = in + at(0, y, z, w, h);
            float* o = tx.data() + ((int64_t)z * h + y) * ow;
            for (int x = 0; x < ow; x++) {
                float c = 0;
                for (int k = 0; k < P; k++) c = fmaf(row[x + k], g[k], c);
                o[x] = c;
            }
        }
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < l; z++)
        for (int y = 0; y < oh; y++) {
            float* o = ty.data() + ((int64_t)z * oh + y) * ow;
            for (int x = 0; x < ow; x++) {
                float c = 0;
                for (int k = 0; k < P; k++)
                    c = fmaf(tx[((int64_t)z * h + y + k) * ow + x], g[k], c);
                o[x] = c;
            }
        }
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < ol; z++)
        for (int y = 0; y < oh; y++) {
            float* o = out + ((int64_t)z * oh + y) * ow;
            for (int x = 0; x < ow; x++) {
                float c = 0;
                for (int k = 0; k < P; k++)
                    c = fmaf(ty[((int64_t)(z + k) * oh + y) * ow + x], g[k], c);
                o[x] = c;
            }
        }
}

/* kernel.cu:300-325: border voxels get the constant 30, the rest a copy. */
template <class T>
void pad_border_t(T* out, const T* in, int vw, int vh, int vl, int pad) {
    int pw = vw + 2 * pad, ph = vh + 2 * pad, pl = vl + 2 * pad;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < pl; k++)
        for (int i = 0; i < ph; i++)
            for (int j = 0; j < pw; j++) {
                int64_t o = ((int64_t)k * ph + i) * pw + j;
                if (j < pad || j >= vw + pad || i < pad || i >= vh + pad || k < pad || k >= vl + pad)
                    out[o] = (T)30;
                else
                    out[o] = in[((int64_t)(k - pad) * vh + (i - pad)) * vw + (j - pad)];
            }
}

/* kernel.cu:25-55.  The weight 1/pow(2s+1,2) is a double; each step rounds the
 * running float sum through a double add. */
void energies_f(float* Eo, float* Ei, const float* fo, const float* fi, const float* I,
                int w, int h, int l, int s) {
    const double wgt = 1 / std::pow((double)((2 * s) + 1), 2);
#pragma omp parallel for collapse(2) schedule(static)
    for (int p = 0; p < l; p++)
        for (int i = 0; i < h; i++)
            for (int j = 0; j < w; j++) {
                int64_t c = at(j, i, p, w, h);
                float s1 = 0, s2 = 0;
                for (int u = -s; u <= s; u++)
                    for (int v = -s; v <= s; v++)
                        for (int ww = -s; ww <= s; ww++) {
                            int ii = i + u, jj = j + v, pp = p + ww;
                            if (ii < 0 || ii >= h || jj < 0 || jj >= w || pp < 0 || pp >= l) continue;
                            float In = I[at(jj, ii, pp, w, h)];
                            float d1 = In - fo[c], d2 = In - fi[c];
                            s1 = (float)(s1 + wgt * d1 * d1);
                            s2 = (float)(s2 + wgt * d2 * d2);
                        }
                Eo[c] = s1;
                Ei[c] = s2;
            }
}

/* kernel.cu:384, left-associative, no contraction.  lambda1 == lambda2 keeps
 * the reference's e*(Eo-Ei); otherwise lambda1*Eo - lambda2*Ei. */
inline float final_phi_one(float phi, float dh, float Eo, float Ei, float kap, float L,
                           float l1, float l2, float dt, float v, float mu) {
    float F = (l1 == l2) ? l1 * (Eo - Ei) : (l1 * Eo - l2 * Ei);
    float t1 = (dh * F) * dt;
    float t2 = ((v * dt) * dh) * kap;
    float t3 = mu * (dt * (L - kap));
    return ((phi - t1) + t2) + t3;
}

struct Work32 {
    std::vector<float> H, A, Hc, B, gx, gy, gz, M, nx, ny, nz, dx, dy, dz, kap, L, blur, pad1, pad2,
        fo, fi, Eo, Ei, dh;
    void alloc(int64_t n, int64_t nv) {
        for (auto* v : {&H, &A, &Hc, &B, &gx, &gy, &gz, &M, &nx, &ny, &nz, &dx, &dy, &dz, &kap, &L,
                        &pad1, &pad2, &fo, &fi, &Eo, &Ei, &dh})
            v->resize(n);
        blur.resize(nv);
    }
};

int check_dims(int w, int h, int l, int P) {
    if (w < P || h < P || l < P || w < 2 || h < 2 || l < 2) return -2;
    return 0;
}

/* main.cpp:234-300, one iteration, FP32.  Writes the new phi to phi_out
 * (may alias phi) and keeps intermediates in wk. */
int step32(const float* phi, const float* I, int w, int h, int l, const oracle_params* p,
           float* phi_out, Work32& wk) {
    float g[256];
    int P, m;
    if (taps_impl(p->sigma, g, 256, &P, &m)) return -1;
    if (check_dims(w, h, l, P)) return -2;
    int64_t n = nvox(w, h, l);
    int vw = w - P + 1, vh = h - P + 1, vl = l - P + 1;
    wk.alloc(n, (int64_t)vw * vh * vl);

    heaviside_f(wk.H.data(), phi, n, p->epsilon);                                // :234
    oracle_multiplication(wk.A.data(), I, wk.H.data(), n);                         // :238
    oracle_inverse(wk.Hc.data(), wk.H.data(), n);                                  // :242
    oracle_multiplication(wk.B.data(), I, wk.Hc.data(), n);                        // :246
    gradient_f(wk.gx.data(), wk.gy.data(), wk.gz.data(), phi, w, h, l);           // :250
    oracle_gradientmagnitude(wk.M.data(), wk.gx.data(), wk.gy.data(), wk.gz.data(), n);  // :252
    oracle_division_grad(wk.nx.data(), wk.gx.data(), wk.M.data(), n);              // :254
    oracle_division_grad(wk.ny.data(), wk.gy.data(), wk.M.data(), n);
    oracle_division_grad(wk.nz.data(), wk.gz.data(), wk.M.data(), n);
    {   /* :256-258  kappa = dx(nx) + dy(ny) + dz(nz) */
        std::vector<float> t1(n), t2(n);
        gradient_f(wk.dx.data(), t1.data(), t2.data(), wk.nx.data(), w, h, l);
        gradient_f(t1.data(), wk.dy.data(), t2.data(), wk.ny.data(), w, h, l);
        gradient_f(t1.data(), t2.data(), wk.dz.data(), wk.nz.data(), w, h, l);
        oracle_plus(wk.kap.data(), wk.dx.data(), wk.dy.data(), wk.dz.data(), n);
        /* :260  L = dx(gx) + dy(gy) + dz(gz) */
        gradient_f(wk.dx.data(), t1.data(), t2.data(), wk.gx.data(), w, h, l);
        gradient_f(t1.data(), wk.dy.data(), t2.data(), wk.gy.data(), w, h, l);
        gradient_f(t1.data(), t2.data(), wk.dz.data(), wk.gz.data(), w, h, l);
        oracle_plus(wk.L.data(), wk.dx.data(), wk.dy.data(), wk.dz.data(), n);
    }
    conv_separable_f(wk.blur.data(), wk.A.data(), w, h, l, g, P);                 // :268
    pad_border_t(wk.pad1.data(), wk.blur.data(), vw, vh, vl, m);                  // :274
    conv_separable_f(wk.blur.data(), wk.H.data(), w, h, l, g, P);                 // :270
    pad_border_t(wk.pad2.data(), wk.blur.data(), vw, vh, vl, m);                  // :275
    oracle_division(wk.fo.data(), wk.pad1.data(), wk.pad2.data(), n);              // :277
    conv_separable_f(wk.blur.data(), wk.B.data(), w, h, l, g, P);                 // :282
    pad_border_t(wk.pad1.data(), wk.blur.data(), vw, vh, vl, m);                  // :286
    conv_separable_f(wk.blur.data(), wk.Hc.data(), w, h, l, g, P);                // :283
    pad_border_t(wk.pad2.data(), wk.blur.data(), vw, vh, vl, m);                  // :287
    oracle_division(wk.fi.data(), wk.pad1.data(), wk.pad2.data(), n);              // :289
    energies_f(wk.Eo.data(), wk.Ei.data(), wk.fo.data(), wk.fi.data(), I, w, h, l,
               (int)p->sigma2);                                                    // :293
    deri_heaviside_f(wk.dh.data(), phi, n, p->epsilon);                           // :296
    oracle_final_phi(phi_out, phi, wk.dh.data(), wk.Eo.data(), wk.Ei.data(), wk.kap.data(),
                     wk.L.data(), p->lambda1, p->lambda2, p->dt, p->nu, p->mu, n);   // :300
    return 0;
}

/* ---------------------------------------------------------------- FP64 twin */

/* Difference operator of kernel.cu:57-125 in linear-index form (FP64):
 * D f(i) = (f(R(i)) - f(L(i))) / 2 with, along x, R = i+1 (i at x = w-1) and
 * L = i-1 (0 at i = 0); along y and z, the clamped one-sided rules. */
void diff_d(double* out, const double* f, int w, int h, int l, int axis) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < l; z++)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                int64_t i = at(x, y, z, w, h), R, Lq;
                if (axis == 0) {
                    R = (x == w - 1) ? i : i + 1;
                    Lq = (i == 0) ? 0 : i - 1;
                } else if (axis == 1) {
                    int yl = y - 1, yr = y + 1;
                    if (yl < 0) { yl = 0; yr = 1; } else if (yr >= h) { yr = h - 1; yl = yr - 1; }
                    R = at(x, yr, z, w, h); Lq = at(x, yl, z, w, h);
                } else {
                    int zl = z > 0 ? z - 1 : 0, zr = z < l - 1 ? z + 1 : l - 1;
                    R = at(x, y, zr, w, h); Lq = at(x, y, zl, w, h);
                }
                out[i] = (f[R] - f[Lq]) / 2.0;
            }
}

/* Transpose of diff_d, as a gather: (D^T u)(j) = (sum_{i:R(i)=j} u(i) -
 * sum_{i:L(i)=j} u(i)) / 2. */
void diff_t_d(double* out, const double* u, int w, int h, int l, int axis) {
    int64_t N = nvox(w, h, l);
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < l; z++)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                int64_t j = at(x, y, z, w, h);
                double plus = 0, minus = 0;
                if (axis == 0) {
                    if (x != 0) plus += u[j - 1];
                    if (x == w - 1) plus += u[j];
                    if (j + 1 < N) minus += u[j + 1];
                    if (j == 0) minus += u[0];
                } else if (axis == 1) {
                    int64_t s = w;
                    if (y >= 1) plus += u[j - s];
                    if (y == h - 1) plus += u[j];
                    if (y == 0) minus += u[j];
                    if (y <= h - 2) minus += u[j + s];
                } else {
                    int64_t s = (int64_t)w * h;
                    if (z >= 1) plus += u[j - s];
                    if (z == l - 1) plus += u[j];
                    if (z == 0) minus += u[j];
                    if (z <= l - 2) minus += u[j + s];
                }
                out[j] = (plus - minus) / 2.0;
            }
}

/* Valid correlation along one axis (FP64). */
void corr_axis_d(double* out, const double* in, int w, int h, int l, int axis, const double* g, int P) {
    int ow = w, oh = h, ol = l;
    if (axis == 0) ow = w - P + 1; else if (axis == 1) oh = h - P + 1; else ol = l - P + 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < ol; z++)
        for (int y = 0; y < oh; y++)
            for (int x = 0; x < ow; x++) {
                double c = 0;
                for (int k = 0; k < P; k++) {
                    int xx = x + (axis == 0) * k, yy = y + (axis == 1) * k, zz = z + (axis == 2) * k;
                    c += in[at(xx, yy, zz, w, h)] * g[k];
                }
                out[at(x, y, z, ow, oh)] = c;
            }
}

/* Transpose of corr_axis_d: out has the input dims (w,h,l); in the output dims. */
void corr_axis_t_d(double* out, const double* in, int w, int h, int l, int axis, const double* g, int P) {
    int ow = w, oh = h, ol = l;
    if (axis == 0) ow = w - P + 1; else if (axis == 1) oh = h - P + 1; else ol = l - P + 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < l; z++)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                double c = 0;
                for (int k = 0; k < P; k++) {
                    int xx = x - (axis == 0) * k, yy = y - (axis == 1) * k, zz = z - (axis == 2) * k;
                    if (xx < 0 || yy < 0 || zz < 0 || xx >= ow || yy >= oh || zz >= ol) continue;
                    c += in[at(xx, yy, zz, ow, oh)] * g[k];
                }
                out[at(x, y, z, w, h)] = c;
            }
}

void blur_valid_d(double* out, const double* in, int w, int h, int l, const double* g, int P) {
    int ow = w - P + 1, oh = h - P + 1;
    std::vector<double> a((size_t)ow * h * l), b((size_t)ow * oh * l);
    corr_axis_d(a.data(), in, w, h, l, 0, g, P);
    corr_axis_d(b.data(), a.data(), ow, h, l, 1, g, P);
    corr_axis_d(out, b.data(), ow, oh, l, 2, g, P);
}

void blur_valid_t_d(double* out, const double* in, int w, int h, int l, const double* g, int P) {
    int ow = w - P + 1, oh = h - P + 1;
    std::vector<double> a((size_t)ow * h * l), b((size_t)ow * oh * l);
    corr_axis_t_d(b.data(), in, ow, oh, l, 2, g, P);
    corr_axis_t_d(a.data(), b.data(), ow, h, l, 1, g, P);
    corr_axis_t_d(out, a.data(), w, h, l, 0, g, P);
}

struct Fwd64 {
    int w, h, l, P, m, s;
    int64_t n;
    double g[256];
    std::vector<double> H, dh, gx, gy, gz, M, nx, ny, nz, kap, L, PA, PH, PB, PHc, fo, fi, Eo, Ei, F;
    std::vector<unsigned char> valid;
};

inline bool in_valid(int x, int y, int z, const Fwd64& f) {
    int vw = f.w - f.P + 1, vh = f.h - f.P + 1, vl = f.l - f.P + 1;
    return x >= f.m && x < vw + f.m && y >= f.m && y < vh + f.m && z >= f.m && z < vl + f.m;
}

/* Forward of one step in FP64, keeping what the adjoint needs. */
int forward64(const double* phi, const double* I, int w, int h, int l, const oracle_params* p,
              double* phi_out, Fwd64& f) {
    float g32[256];
    int P, m;
    if (taps_impl(p->sigma, g32, 256, &P, &m)) return -1;
    if (check_dims(w, h, l, P)) return -2;
    f.w = w; f.h = h; f.l = l; f.P = P; f.m = m; f.s = (int)p->sigma2;
    for (int k = 0; k < P; k++) f.g[k] = (double)g32[k];
    int64_t n = f.n = nvox(w, h, l);
    for (auto* v : {&f.H, &f.dh, &f.gx, &f.gy, &f.gz, &f.M, &f.nx, &f.ny, &f.nz, &f.kap, &f.L,
                    &f.PA, &f.PH, &f.PB, &f.PHc, &f.fo, &f.fi, &f.Eo, &f.Ei, &f.F})
        v->assign(n, 0.0);
    const double eps = p->epsilon;
    std::vector<double> A(n), B(n), Hc(n), t(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        f.H[i] = 0.5 * (1 + (2 / kPi) * std::atan(phi[i] / eps));
        f.dh[i] = (1 / kPi) * (eps / (eps * eps + phi[i] * phi[i]));
        A[i] = I[i] * f.H[i];
        Hc[i] = 1 - f.H[i];
        B[i] = I[i] * Hc[i];
    }
    diff_d(f.gx.data(), phi, w, h, l, 0);
    diff_d(f.gy.data(), phi, w, h, l, 1);
    diff_d(f.gz.data(), phi, w, h, l, 2);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        f.M[i] = std::sqrt(f.gx[i] * f.gx[i] + f.gy[i] * f.gy[i] + f.gz[i] * f.gz[i]);
        double r = 1.0 / (f.M[i] + 1e-6);
        f.nx[i] = f.gx[i] * r; f.ny[i] = f.gy[i] * r; f.nz[i] = f.gz[i] * r;
    }
    diff_d(t.data(), f.nx.data(), w, h, l, 0);
    for (int64_t i = 0; i < n; i++) f.kap[i] = t[i];
    diff_d(t.data(), f.ny.data(), w, h, l, 1);
    for (int64_t i = 0; i < n; i++) f.kap[i] += t[i];
    diff_d(t.data(), f.nz.data(), w, h, l, 2);
    for (int64_t i = 0; i < n; i++) f.kap[i] += t[i];
    diff_d(t.data(), f.gx.data(), w, h, l, 0);
    for (int64_t i = 0; i < n; i++) f.L[i] = t[i];
    diff_d(t.data(), f.gy.data(), w, h, l, 1);
    for (int64_t i = 0; i < n; i++) f.L[i] += t[i];
    diff_d(t.data(), f.gz.data(), w, h, l, 2);
    for (int64_t i = 0; i < n; i++) f.L[i] += t[i];

    int vw = w - P + 1, vh = h - P + 1, vl = l - P + 1;
    std::vector<double> bl((size_t)vw * vh * vl);
    auto blur_pad = [&](std::vector<double>& dst, const std::vector<double>& src) {
        blur_valid_d(bl.data(), src.data(), w, h, l, f.g, P);
        pad_border_t(dst.data(), bl.data(), vw, vh, vl, m);
    };
    blur_pad(f.PA, A); blur_pad(f.PH, f.H); blur_pad(f.PB, B); blur_pad(f.PHc, Hc);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) { f.fo[i] = f.PA[i] / f.PH[i]; f.fi[i] = f.PB[i] / f.PHc[i]; }
    const int s = f.s;
    const double wgt = 1.0 / ((2.0 * s + 1) * (2.0 * s + 1));
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < l; z++)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                int64_t c = at(x, y, z, w, h);
                double s1 = 0, s2 = 0;
                for (int u = -s; u <= s; u++)
                    for (int v = -s; v <= s; v++)
                        for (int ww = -s; ww <= s; ww++) {
                            int yy = y + u, xx = x + v, zz = z + ww;
                            if (yy < 0 || yy >= h || xx < 0 || xx >= w || zz < 0 || zz >= l) continue;
                            double In = I[at(xx, yy, zz, w, h)];
                            s1 += wgt * (In - f.fo[c]) * (In - f.fo[c]);
                            s2 += wgt * (In - f.fi[c]) * (In - f.fi[c]);
                        }
                f.Eo[c] = s1; f.Ei[c] = s2;
            }
    const double l1 = p->lambda1, l2 = p->lambda2, dt = p->dt, nu = p->nu, mu = p->mu;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        f.F[i] = l1 * f.Eo[i] - l2 * f.Ei[i];
        phi_out[i] = phi[i] - f.dh[i] * f.F[i] * dt + nu * dt * f.dh[i] * f.kap[i] +
                     mu * dt * (f.L[i] - f.kap[i]);
    }
    return 0;
}

}  // namespace

extern "C" {

int oracle_taps(float sigma, float* g, int cap, int* P, int* m) { return taps_impl(sigma, g, cap, P, m); }

void oracle_heaviside(float* out, const float* phi, int64_t n, float eps) { heaviside_f(out, phi, n, eps); }
void oracle_deri_heaviside(float* out, const float* phi, int64_t n, float eps) { deri_heaviside_f(out, phi, n, eps); }

void oracle_multiplication(float* out, const float* a, const float* b, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) out[i] = a[i] * (b[i]);      // kernel.cu:222
}
void oracle_inverse(float* out, const float* a, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) out[i] = 1 - a[i];           // kernel.cu:295
}
void oracle_plus(float* out, const float* a, const float* b, const float* c, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) out[i] = a[i] + b[i] + c[i]; // kernel.cu:242
}
void oracle_division(float* out, const float* a, const float* b, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) out[i] = a[i] / (b[i]);      // kernel.cu:203
}
void oracle_division_grad(float* out, const float* g, const float* M, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) out[i] = (float)(g[i] / (M[i] + 1e-6));   // kernel.cu:185
}
void oracle_atan_f64(double* out, const double* x, int64_t n) {
    for (int64_t i = 0; i < n; i++) out[i] = det_atan(x[i]);
}
void oracle_gradientmagnitude(float* M, const float* gx, const float* gy, const float* gz, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)                                 // kernel.cu:928
        M[i] = sqrtf(gx[i] * gx[i] + gy[i] * gy[i] + gz[i] * gz[i]);
}
void oracle_gradient(float* gx, float* gy, float* gz, const float* f, int w, int h, int l) {
    gradient_f(gx, gy, gz, f, w, h, l);
}
void oracle_conv_separable(float* out, const float* in, int w, int h, int l, const float* g, int P) {
    conv_separable_f(out, in, w, h, l, g, P);
}
void oracle_pad_border(float* out, const float* in, int vw, int vh, int vl, int pad) {
    pad_border_t(out, in, vw, vh, vl, pad);
}
void oracle_energies(float* Eo, float* Ei, const float* fo, const float* fi, const float* I,
                     int w, int h, int l, int s) {
    energies_f(Eo, Ei, fo, fi, I, w, h, l, s);
}
void oracle_final_phi(float* out, const float* phi, const float* dh, const float* Eo,
                      const float* Ei, const float* kappa, const float* L, float lambda1,
                      float lambda2, float dt, float nu, float mu, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        out[i] = final_phi_one(phi[i], dh[i], Eo[i], Ei[i], kappa[i], L[i], lambda1, lambda2, dt, nu, mu);
}

int oracle_step(float* phi, const float* I, int w, int h, int l, const oracle_params* p) {
    Work32 wk;
    return step32(phi, I, w, h, l, p, phi, wk);
}

int oracle_run(float* phi, const float* I, int w, int h, int l, const oracle_params* p, int T,
               int threads) {
#ifdef _OPENMP
    int saved = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
#endif
    Work32 wk;
    int rc = 0;
    for (int t = 0; t < T && rc == 0; t++) rc = step32(phi, I, w, h, l, p, phi, wk);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(saved);
#endif
    return rc;
}

int oracle_step_debug(const float* phi, const float* I, int w, int h, int l, const oracle_params* p,
                      float* phi_out, float* H, float* F, float* kappa, float* L, float* fo, float* fi) {
    Work32 wk;
    int64_t n = nvox(w, h, l);
    std::vector<float> out(n);
    int rc = step32(phi, I, w, h, l, p, out.data(), wk);
    if (rc) return rc;
    auto cp = [n](float* dst, const std::vector<float>& src) {
        if (dst) std::memcpy(dst, src.data(), n * sizeof(float));
    };
    cp(phi_out, out); cp(H, wk.H); cp(kappa, wk.kap); cp(L, wk.L); cp(fo, wk.fo); cp(fi, wk.fi);
    if (F) {
        for (int64_t i = 0; i < n; i++)
            F[i] = (p->lambda1 == p->lambda2) ? p->lambda1 * (wk.Eo[i] - wk.Ei[i])
                                              : (p->lambda1 * wk.Eo[i] - p->lambda2 * wk.Ei[i]);
    }
    return 0;
}

int oracle_step_f64(double* phi, const double* I, int w, int h, int l, const oracle_params* p) {
    Fwd64 f;
    std::vector<double> out(nvox(w, h, l));
    int rc = forward64(phi, I, w, h, l, p, out.data(), f);
    if (rc) return rc;
    std::memcpy(phi, out.data(), out.size() * sizeof(double));
    return 0;
}

/* Adjoint of forward64.  With u = dLoss/dphi':
 *   phi' = phi - dt*delta*F + nu*dt*delta*kappa + mu*dt*(L - kappa)
 * so  dphi += u + delta'(phi)*(-dt*F + nu*dt*kappa)*u,
 *     Fbar = -dt*delta*u, kbar = (nu*dt*delta - mu*dt)*u, Lbar = mu*dt*u,
 * F = l1*Eo - l2*Ei with Eo = sum w*(I(nb) - fo)^2 (valid region only; the
 * band has fo = fi = 1 and carries no gradient), fo = PA/PH, fi = PB/PHc,
 * A = I*H, B = I*(1-H), Hc = 1-H, H' = delta; kappa = sum_a D_a n_a with
 * n = g/(|g|+1e-6), L = sum_a D_a g_a, g = D phi. */
int oracle_step_adjoint_f64(const double* phi, const double* I, int w, int h, int l,
                            const oracle_params* p, const double* gout, double* gin,
                            double* gparams) {
    Fwd64 f;
    int64_t n = nvox(w, h, l);
    std::vector<double> out(n);
    int rc = forward64(phi, I, w, h, l, p, out.data(), f);
    if (rc) return rc;
    const double eps = p->epsilon, l1 = p->lambda1, l2 = p->lambda2, dt = p->dt, nu = p->nu,
                 mu = p->mu;
    std::vector<double> kbar(n), Lbar(n), QH(n, 0.0), QA(n, 0.0);
    double gl1 = 0, gl2 = 0, gnu = 0, gmu = 0;
    const int s = f.s;
    const double wgt = 1.0 / ((2.0 * s + 1) * (2.0 * s + 1));
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : gl1, gl2, gnu, gmu)
    for (int z = 0; z < l; z++)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                int64_t i = at(x, y, z, w, h);
                double u = gout[i];
                double d = f.dh[i];
                double ph = phi[i];
                double ddh = -(1 / kPi) * eps * 2 * ph / ((eps * eps + ph * ph) * (eps * eps + ph * ph));
                gin[i] = u + ddh * (-dt * f.F[i] + nu * dt * f.kap[i]) * u;
                double Fbar = -dt * d * u;
                kbar[i] = (nu * dt * d - mu * dt) * u;
                Lbar[i] = mu * dt * u;
                gl1 += f.Eo[i] * Fbar;
                gl2 += -f.Ei[i] * Fbar;
                gnu += dt * d * f.kap[i] * u;
                gmu += dt * (f.L[i] - f.kap[i]) * u;
                if (!in_valid(x, y, z, f)) continue;
                double Eob = l1 * Fbar, Eib = -l2 * Fbar;
                double sum_o = 0, sum_i = 0;   // d(E)/d(f) = -2 w sum (I(nb) - f)
                for (int uu = -s; uu <= s; uu++)
                    for (int v = -s; v <= s; v++)
                        for (int ww = -s; ww <= s; ww++) {
                            int yy = y + uu, xx = x + v, zz = z + ww;
                            if (yy < 0 || yy >= h || xx < 0 || xx >= w || zz < 0 || zz >= l) continue;
                            double In = I[at(xx, yy, zz, w, h)];
                            sum_o += In - f.fo[i];
                            sum_i += In - f.fi[i];
                        }
                double fob = -2 * wgt * sum_o * Eob, fib = -2 * wgt * sum_i * Eib;
                double PAb = fob / f.PH[i], PHb = -fob * f.fo[i] / f.PH[i];
                double PBb = fib / f.PHc[i], PHcb = -fib * f.fi[i] / f.PHc[i];
                QH[i] = PHb - PHcb;   // H-bar = blur^T(PHb - PHcb) + I * blur^T(PAb - PBb)
                QA[i] = PAb - PBb;
            }
    /* Transposed valid blur of the two combined channels. */
    int P = f.P, m = f.m, vw = w - P + 1, vh = h - P + 1, vl = l - P + 1;
    std::vector<double> vq((size_t)vw * vh * vl), bt(n);
    auto unpad = [&](const std::vector<double>& src) {
        for (int z = 0; z < vl; z++)
            for (int y = 0; y < vh; y++)
                for (int x = 0; x < vw; x++)
                    vq[at(x, y, z, vw, vh)] = src[at(x + m, y + m, z + m, w, h)];
    };
    unpad(QH);
    blur_valid_t_d(bt.data(), vq.data(), w, h, l, f.g, P);
    std::vector<double> Hbar(bt);
    unpad(QA);
    blur_valid_t_d(bt.data(), vq.data(), w, h, l, f.g, P);
    for (int64_t i = 0; i < n; i++) {
        Hbar[i] += I[i] * bt[i];
        gin[i] += f.dh[i] * Hbar[i];
    }
    /* Curvature and Laplacian. */
    std::vector<double> nb[3], gb[3], t(n);
    for (int a = 0; a < 3; a++) {
        nb[a].resize(n);
        gb[a].resize(n);
        diff_t_d(nb[a].data(), kbar.data(), w, h, l, a);
        diff_t_d(gb[a].data(), Lbar.data(), w, h, l, a);
    }
    for (int64_t i = 0; i < n; i++) {
        double M = f.M[i], g3[3] = {f.gx[i], f.gy[i], f.gz[i]};
        double r = 1.0 / (M + 1e-6);
        double dot = nb[0][i] * g3[0] + nb[1][i] * g3[1] + nb[2][i] * g3[2];
        double c = (M > 0) ? dot / (M * (M + 1e-6) * (M + 1e-6)) : 0.0;
        for (int a = 0; a < 3; a++) gb[a][i] += nb[a][i] * r - g3[a] * c;
    }
    for (int a = 0; a < 3; a++) {
        diff_t_d(t.data(), gb[a].data(), w, h, l, a);
        for (int64_t i = 0; i < n; i++) gin[i] += t[i];
    }
    if (gparams) {
        gparams[0] += gl1; gparams[1] += gl2; gparams[2] += gnu; gparams[3] += gmu;
    }
    return 0;
}

void oracle_diff_f64(double* out, const double* f, int w, int h, int l, int axis) {
    diff_d(out, f, w, h, l, axis);
}
void oracle_diff_transpose_f64(double* out, const double* u, int w, int h, int l, int axis) {
    diff_t_d(out, u, w, h, l, axis);
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
