/*
 * Oracle: CPU restatement of the reference RSF level-set iteration.
 *
 * TEST INFRASTRUCTURE ONLY.  Only the test suite, the smoke check and the
 * CPU-baseline leg of the benchmark may load this library, and only as the
 * checker or as the timed CPU baseline -- never as part of the product path.
 *
 * Each function restates one reference function (CUDA_RSF/kernel.cu and
 * CUDA_RSF/main.cpp, cited per function) on host memory, unfused and in the
 * reference's own operation order, FP32 with the reference's FP64 promotions
 * (double literals such as 2/pi, 1/pi and 1e-6).  The valid 1-D correlations
 * accumulate with fmaf, the contraction nvcc applies to `conv += a*b` by
 * default.  A second instantiation of the whole step runs in FP64 and an FP64
 * adjoint of one step backs the gradient checks.
 *
 * Volumes are X-fastest: voxel (x, y, z) lives at ((z*h)+y)*w+x.
 *
 * Parity status of each function is listed in the header comment of
 * rsf_oracle.cpp.
 */
#ifndef RSF_ORACLE_H
#define RSF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_params {
    float sigma;      /* Gaussian sigma of the fitting blur (--sigma)           */
    float lambda1;    /* weight of the outside energy (--e in the reference)    */
    float lambda2;    /* weight of the inside energy  (--e in the reference)    */
    float nu;         /* curvature weight (--v)                                 */
    float mu;         /* regulariser weight (--meu)                             */
    float dt;         /* time step (--dt)                                       */
    float epsilon;    /* Heaviside width (hard-coded 0.2 in the reference)      */
    float sigma2;     /* energy window half-width, truncated to int (--sigma2)  */
} oracle_params;

/* Taps of main.cpp:152-166.  Returns 0, or -1 when floor(7*sigma) is even
 * (the reference then mis-sizes its blurred buffers, main.cpp:121,191 vs
 * kernel.cu:1067) or when cap is too small. */
int oracle_taps(float sigma, float* g, int cap, int* P, int* m);

/* The arctangent behind the oracle's float atanf: double precision from IEEE
 * basic operations only (Cephes rational approximation). */
void oracle_atan_f64(double* out, const double* x, int64_t n);

/* Element-wise kernels, n voxels. */
void oracle_heaviside(float* out, const float* phi, int64_t n, float eps);
void oracle_deri_heaviside(float* out, const float* phi, int64_t n, float eps);
void oracle_multiplication(float* out, const float* a, const float* b, int64_t n);
void oracle_inverse(float* out, const float* a, int64_t n);
void oracle_plus(float* out, const float* a, const float* b, const float* c, int64_t n);
void oracle_division(float* out, const float* a, const float* b, int64_t n);
void oracle_division_grad(float* out, const float* g, const float* M, int64_t n);
void oracle_gradientmagnitude(float* M, const float* gx, const float* gy, const float* gz, int64_t n);

/* Central differences with the reference edge rules (x wraps to the previous
 * linear element at x = 0; voxel (0,0,0) is defined as the clamp). */
void oracle_gradient(float* gx, float* gy, float* gz, const float* f, int w, int h, int l);

/* Valid separable correlation x -> y -> z; out has (w-P+1)(h-P+1)(l-P+1) voxels. */
void oracle_conv_separable(float* out, const float* in, int w, int h, int l, const float* g, int P);

/* Centre a (vw,vh,vl) block at offset pad inside a border of 30.0f. */
void oracle_pad_border(float* out, const float* in, int vw, int vh, int vl, int pad);

/* Fitting energies with window half-width s (s = 0: pointwise). */
void oracle_energies(float* Eo, float* Ei, const float* fo, const float* fi, const float* I,
                     int w, int h, int l, int s);

/* Explicit-Euler update of kernel.cu:365-387 (lambda1 == lambda2 == e keeps
 * the reference's e*(Eo-Ei) rounding). */
void oracle_final_phi(float* out, const float* phi, const float* dh, const float* Eo,
                      const float* Ei, const float* kappa, const float* L,
                      float lambda1, float lambda2, float dt, float nu, float mu, int64_t n);

/* One full iteration of main.cpp:234-300, in place on phi (FP32).  I is the
 * image as float.  Returns 0 or a negative error. */
int oracle_step(float* phi, const float* I, int w, int h, int l, const oracle_params* p);

/* Same as oracle_step, iterated T times; threads <= 0 keeps the OpenMP default. */
int oracle_run(float* phi, const float* I, int w, int h, int l, const oracle_params* p,
               int T, int threads);

/* Intermediates of one FP32 step (any pointer may be NULL): Heaviside H, force
 * F = lambda*(Eo-Ei), curvature kappa, Laplacian L, local fits fo/fi. */
int oracle_step_debug(const float* phi, const float* I, int w, int h, int l,
                      const oracle_params* p, float* phi_out, float* H, float* F,
                      float* kappa, float* L, float* fo, float* fi);

/* FP64 twin of one step (same formulas without the FP32 roundings). */
int oracle_step_f64(double* phi, const double* I, int w, int h, int l, const oracle_params* p);

/* FP64 adjoint of one step.  Given phi (state before the step) and gout =
 * dLoss/dphi', writes gin = dLoss/dphi and accumulates dLoss/d(lambda1,
 * lambda2, nu, mu) into gparams[0..3]. */
int oracle_step_adjoint_f64(const double* phi, const double* I, int w, int h, int l,
                            const oracle_params* p, const double* gout, double* gin,
                            double* gparams);

/* Transposed difference operators (FP64), for adjoint tests: out = D_a^T u. */
void oracle_diff_transpose_f64(double* out, const double* u, int w, int h, int l, int axis);
void oracle_diff_f64(double* out, const double* f, int w, int h, int l, int axis);

int oracle_max_threads(void);

#ifdef __cplusplus
}
#endif

#endif
