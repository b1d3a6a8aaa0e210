// Host-side internals shared by the librsf translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <initializer_list>
#include <string>
#include <vector>

#include "rsf.h"
#include "rsf_device.hpp"

namespace rsf {

// ------------------------------------------------------------ error plumbing
int set_error(int status, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define RSF_HIP(expr)                                                         \
    do {                                                                      \
        hipError_t rsf_e_ = (expr);                                           \
        if (rsf_e_ != hipSuccess) return ::rsf::hip_fail(rsf_e_, #expr, __FILE__, __LINE__); \
    } while (0)

#define RSF_TRY(expr)                   \
    do {                                \
        int rsf_s_ = (expr);            \
        if (rsf_s_ != RSF_OK) return rsf_s_; \
    } while (0)

// Validates a context, makes its device current and returns its stream.
int context_stream(rsf_context* c, hipStream_t* s);

// Grow-only device allocation, or (bind) a view of memory someone else owns.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool owned = true;
    int ensure(size_t n);
    void bind(void* q, size_t n) {
        release();
        p = q;
        bytes = n;
        owned = false;
    }
    void release();
    ~DevBuf() { release(); }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Derived quantities of a parameter set.
struct Derived {
    int P = 0, m = 0;
    Taps taps{};
    float halfS3 = 0.f;   // (sum g)^3 / 2: blur of the constant 1/2
};
int derive(const rsf_params* p, Derived* d);
int check_params(const rsf_params* p);

// ------------------------------------------------------ per-op kernel launches
// Literal mirrors of the reference kernels (rsf_ops.hip).
namespace ops {
int heaviside(float* out, const float* in, int64_t n, float eps, hipStream_t s);
int deri_heaviside(float* out, const float* in, int64_t n, float eps, hipStream_t s);
int multiplication(float* out, const float* a, const float* b, int64_t n, hipStream_t s);
int inverse(float* out, const float* a, int64_t n, hipStream_t s);
int plus(float* out, const float* a, const float* b, const float* c, int64_t n, hipStream_t s);
int division(float* out, const float* a, const float* b, int64_t n, hipStream_t s);
int division_grad(float* out, const float* g, const float* M, int64_t n, hipStream_t s);
int gradientmagnitude(float* M, const float* gx, const float* gy, const float* gz, int64_t n,
                      hipStream_t s);
// any of gx/gy/gz may be null
int gradient(float* gx, float* gy, float* gz, const float* f, int w, int h, int l, hipStream_t s);
// valid correlation along one axis (0=x,1=y,2=z) of a (w,h,l) input
int conv_axis(float* out, const float* in, int w, int h, int l, int axis, const Taps& t, int P,
              hipStream_t s);
int pad_border(float* out, const float* in, int vw, int vh, int vl, int pad, hipStream_t s);
int energies(float* Eo, float* Ei, const float* fo, const float* fi, const float* I, int w, int h,
             int l, int sw, hipStream_t s);
int final_phi(float* out, const float* phi, const float* dh, const float* Eo, const float* Ei,
              const float* kap, const float* L, float l1, float l2, float dt, float nu, float mu,
              int64_t n, hipStream_t s);
// F = lambda*(Eo-Ei) or lambda1*Eo - lambda2*Ei
int force(float* F, const float* Eo, const float* Ei, float l1, float l2, int64_t n, hipStream_t s);
// image (u8 or f32) -> float
int to_float(float* out, const void* img, int u8, int64_t n, hipStream_t s);
// valid blur x->y->z of a block then padded to full size (pad value 30)
int blur_pad(float* out, const float* in, float* tmp1, float* tmp2, int w, int h, int l,
             const Taps& t, int P, int m, hipStream_t s);
}  // namespace ops

// --------------------------------------------------------- fused sweeps
struct SweepPlan;  // tile lists and kernel choice of one problem (rsf_sweep.hip)

struct SweepArgs {
    const float* tr;      // H (literal) or G (fast) of the current phi
    const void* img;
    int img_u8;
    const float* blurI;   // fast mode: padded blur of the image
    float* out0;          // F (out_fits 0), else fo
    float* out1;          // fi (out_fits 1, 2)
    float* out2;          // blur(H) (out_fits 2)
    float* out3;          // blur(1-H) (out_fits 2)
    int out_fits;         // 0: force, 1: local fits, 2: fits and blurred Heavisides
    const float* phi;     // sweep 2 input
    const float* F;       // sweep 2 force
    float* phi_out;
    float* tr_out;
};

// Kernel variants compiled into the library.
struct Sweep1Variant {
    int P;
    int literal;
    int tx, ty, nt;
    const void* fn;        // fused kernel, or the xy kernel of a split variant
    const char* name;
    const void* fn2;       // split variants: the z kernel (else null)
    int tx2, ty2;          // z-kernel columns per work item
    int nt2;               // z-kernel threads per workgroup (sweep modes)
    int64_t max_plane;     // largest plane (w*h) its z kernel addresses, 0: any
    const void* fn2s[10];  // z kernel by 2 * output mode + (image is uint8)
    const void* fn_img[2]; // split variants: the xy kernel by (image is uint8), else null
    const void* fn_blur[2];// split fast variants: the image-only xy kernel by (image is uint8), else null
    int strip = 0;         // strip xy kernel: waves per workgroup, each a plane of the work item's chunk
};
struct Sweep2Variant {
    int literal;
    int tx, ty;
    int vec;               // consecutive x voxels per thread of the interior kernel
    const void* fn;        // whole-block kernel, or the interior kernel when shell_fn is set
    const void* shell_fn;  // boundary-shell kernel of a split variant (else null)
    const void* scalar_fn; // vector variants: interior kernel of blocks with unaligned rows
    const char* name;
    const void* fix_fn;    // literal vector variant: exact recompute of the items the interior flagged
    int fullx = 0;         // the interior kernel covers whole rows (x in [0, w)) of aligned blocks
};
const std::vector<Sweep1Variant>& sweep1_variants();
const std::vector<Sweep2Variant>& sweep2_variants();

struct TilePlan {
    int s1 = -1, s2 = -1;        // variant indices
    int zc1 = 0, zc2 = 0;        // z-chunk of each sweep
    int grid1 = 0, grid2 = 0;    // persistent grid sizes
    int grid3 = 0, grid4 = 0, grid5 = 0;
    // sweep 1 (fused, or xy kernel); sweep 2 interior; sweep 2 shell; sweep 2 scalar
    // interior of unaligned blocks; sweep 1 z kernel of a split variant
    std::vector<Tile> tiles1, tiles2, tiles3, tiles4, tiles5;
    DevBuf d_tiles1, d_tiles2, d_tiles3, d_tiles4, d_tiles5, d_blocks;
    DevBuf d_flags2;             // sweep 2 interior: per item and wave, fast forms out of range
    DevBuf xy;                   // split sweep 1: the two xy-blurred channels
    int64_t extent = 0;          // voxels spanned by the packed buffers
    std::vector<BlockDesc> blocks;
    int64_t total_voxels = 0;
};

int plan_tiles(TilePlan& tp, const std::vector<BlockDesc>& blocks, int s1, int s2, int zc1,
               int zc2, int num_cus);
// plan_tiles in two halves: the tile lists and grids on the host, then the
// device copies.  plan_upload with an arena places every device buffer of the
// plan (tiles, flags, block table, split-sweep xy channels) inside it, aligned
// to 256 bytes, and copies on stream s; plan_arena_bytes(tp) is what it needs.
int plan_tiles_host(TilePlan& tp, const std::vector<BlockDesc>& blocks, int s1, int s2, int zc1, int zc2,
                    int num_cus);
size_t plan_arena_bytes(const TilePlan& tp);
int plan_upload(TilePlan& tp, unsigned char* arena, size_t arena_bytes, hipStream_t s);
int launch_sweep1(const TilePlan& tp, const SweepArgs& a, const Derived& d, const rsf_params* p,
                  hipStream_t s);
int launch_sweep2(const TilePlan& tp, const SweepArgs& a, const rsf_params* p, hipStream_t s);
// Fast mode, once per run: out = the image blurred x -> y -> z (valid
// correlation, kernel.cu:1000-1049) and padded with 30 (kernel.cu:300-325),
// through the split sweep-1 kernels of the plan (its image-only xy kernel, or
// for variants without one the xy kernel on T = 1 written into tr, which is then
// scratch).  RSF_ERR_UNSUPPORTED for a fused or literal plan.
int launch_blur_image(const TilePlan& tp, const void* img, int u8, float* tr, float* out, const Derived& d,
                      hipStream_t s);
// tr = H(phi) (literal) or G(phi) (fast) on n voxels
int launch_transform(float* tr, const float* phi, int64_t n, int literal, float eps, hipStream_t s);
int default_sweep1(int P, int literal);
int default_sweep2(int literal);

// --------------------------------------------------------------- adjoint
// Kernels of the one-step vector-Jacobian product (rsf_adjoint.hip); each
// works on one block (pointers already offset to the block).
namespace adj {
int pointwise_q(float* QH, float* QA, const float* phi, const float* gout, const void* img, int u8,
                const float* fo, const float* fi, const float* PH, const float* PHc, int w, int h, int l,
                int m, const rsf_params* p, hipStream_t s);
int blur_transpose(float* out, const float* in, float* t1, float* t2, int w, int h, int l,
                   const Derived& d, hipStream_t s);
// Both co-states at once (Bh <- blur^T QH, Ba <- blur^T QA); t1, t2 are scratch.
int blur_transpose2(float* Bh, float* Ba, const float* QH, const float* QA, float* t1, float* t2, int w,
                    int h, int l, const Derived& d, hipStream_t s);
// pointwise_q then blur_transpose2, the co-states formed inside the x pass where
// the tap count and the row length allow (QH, QA then untouched).
//
// q_blur_transpose2, stencil_g and finish also run nb blocks of one shape in one
// launch each: bt is then the device block table (pointers are the buffers' starts,
// each kernel adds its block's offset) and batch_ok(P, w, n, img, u8, {buffers})
// must hold for the buffers and for every block's first voxel (4-voxel aligned).
// The arithmetic and the partial rows (nb * finish_rows(n), block after block) are
// those of the per-block calls.
int q_blur_transpose2(float* Bh, float* Ba, float* QH, float* QA, float* t1, float* t2, const float* phi,
                      const float* gout, const void* img, int u8, const float* fo, const float* fi, const float* PH,
                      const float* PHc, int w, int h, int l, const Derived& d, const rsf_params* p, hipStream_t s,
                      const BlockDesc* bt = nullptr, int nb = 1);
bool batch_ok(int P, int w, int64_t n, const void* img, int u8, std::initializer_list<const void*> ptrs);
// whether stencil_g and finish take their four-voxel kernels for a block (rows a
// multiple of 4, every buffer 16-byte aligned, a uint8 image 4-byte aligned)
bool quads_ok(int w, int64_t n, const void* img, int u8, std::initializer_list<const void*> ptrs);
// The forward step's normal n = g/(|g|+1e-6) and gradient g at every voxel, which
// stencil_g computes and finish reads at a voxel's neighbours (its curvature and
// Laplacian).  g may be null where both take their four-voxel kernels (quads_ok):
// that finishing kernel forms the gradients from phi.
struct FwdStencil {
    float* n[3];
    float* g[3];
};
int stencil_g(float* Gx, float* Gy, float* Gz, FwdStencil fwd, const float* phi, const float* gout, int w,
              int h, int l, const rsf_params* p, hipStream_t s, const BlockDesc* bt = nullptr, int nb = 1);
int finish_rows(int64_t n);
int finish(float* gin, float* partials, const float* phi, const float* gout, const float* Gx,
           const float* Gy, const float* Gz, FwdStencil fwd, const float* Bh, const float* Ba, const void* img,
           int u8, const float* fo, const float* fi, int w, int h, int l, const rsf_params* p, hipStream_t s,
           const BlockDesc* bt = nullptr, int nb = 1);
int reduce_params(float* gparams, const float* partials, int rows, hipStream_t s);
}  // namespace adj

}  // namespace rsf
