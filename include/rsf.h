/*
 * rsf.h -- C ABI of the MI355X-native RSF level-set engine (librsf.so).
 *
 * Drop-in for the RSF evolution of STIM-Lab/paper-rsf-vesselseg (CUDA_RSF/).
 * Plain C types only: pointers, sizes, a POD parameter struct and int status
 * codes.  No C++ exceptions cross this boundary; on failure a function returns
 * a non-zero rsf_status and rsf_last_error() describes it (per host thread).
 *
 * Volumes are X-fastest: voxel (x, y, z) of a w x h x l volume is element
 * ((z*h)+y)*w+x -- the layout of the reference's tira::volume and of a C-order
 * numpy array of shape (l, h, w) (CUDA_RSF/kernel.cu:1010).
 *
 * Memory: functions named *_device take device pointers and work on the
 * context's stream; the caller owns inputs and outputs, the context owns its
 * workspace.  Functions named *_host and the reference-mirroring *_cal entry
 * points take host pointers and return when the result is in host memory.
 */
#ifndef RSF_H
#define RSF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RSF_ABI_VERSION 1

typedef enum rsf_status {
    RSF_OK = 0,
    RSF_ERR_INVALID_ARGUMENT = 1, /* bad pointer, dims, sigma, dtype ...          */
    RSF_ERR_UNSUPPORTED = 2,      /* valid request this build does not implement  */
    RSF_ERR_OUT_OF_MEMORY = 3,    /* device or host allocation failed             */
    RSF_ERR_DEVICE = 4,           /* HIP runtime or kernel launch error           */
    RSF_ERR_IO = 5,               /* file or .npy format error                    */
    RSF_ERR_NONFINITE = 6         /* NaN/Inf detected in phi                      */
} rsf_status;

typedef enum rsf_dtype { RSF_DTYPE_U8 = 0, RSF_DTYPE_F32 = 1 } rsf_dtype;

typedef enum rsf_mode {
    /* Two blurred channels per iteration: with G = atan(phi/eps)/pi, H = 1/2+G and
     * 1-H = 1/2-G, so blur(I*H), blur(H), blur(I*(1-H)), blur(1-H) all follow from
     * blur(I*G), blur(G) and the once-computed blur(I).  About 1.7x faster; one
     * step within 2e-5 of the reference's arithmetic, not bit-identical. */
    RSF_MODE_FAST = 0,
    /* Four blurred channels I*H, H, I*(1-H), 1-H in the reference's own order
     * and rounding (CUDA_RSF/main.cpp:234-289): bit-identical to the oracle.
     * The default. */
    RSF_MODE_LITERAL = 1
} rsf_mode;

/* Parameters of one RSF run.  Replaces the reference's CLI globals
 * (CUDA_RSF/main.cpp:43-51, defaults at :91-97).  The reference has a single
 * fitting weight e; here lambda1 (outside, phi > 0) and lambda2 (inside) --
 * equal values reproduce e*(Eo-Ei) exactly (kernel.cu:384).  epsilon is
 * hard-coded 0.2 in the reference (kernel.cu:258,275). */
typedef struct rsf_params {
    float sigma;        /* blur sigma; floor(7*sigma) must be odd (main.cpp:152-155)   */
    float lambda1;      /* weight of Eo                                                   */
    float lambda2;      /* weight of Ei                                                   */
    float nu;           /* curvature weight (--v)                                         */
    float mu;           /* regulariser weight (--meu)                                     */
    float dt;           /* time step (--dt)                                               */
    float epsilon;      /* Heaviside / Dirac width                                        */
    float sigma2;       /* energy window half-width, truncated to int (--sigma2)          */
    int32_t iterations; /* number of explicit Euler steps (--t)                           */
    int32_t mode;       /* rsf_mode                                                       */
} rsf_params;

typedef struct rsf_context rsf_context;

/* One sub-volume of a batched run: `offset` is its first voxel inside the
 * packed image / phi buffers, w/h/l its dims (each block is X-fastest and
 * contiguous).  Replaces one `rsf3d` process per block (README.md:56). */
typedef struct rsf_block {
    int64_t offset;
    int64_t w, h, l;
} rsf_block;

/* ------------------------------------------------------------ housekeeping */
int rsf_abi_version(void);
const char* rsf_last_error(void);
const char* rsf_status_string(int status);
/* Reference defaults (main.cpp:91-97) plus epsilon = 0.2, mode LITERAL. */
void rsf_params_default(rsf_params* p);
/* Gaussian taps and pad width as built at main.cpp:152-166,264,273.  Returns
 * RSF_ERR_INVALID_ARGUMENT when floor(7*sigma) is even (the reference then
 * writes its blurred buffers with a different shape than it reads them). */
int rsf_taps(float sigma, float* taps, int capacity, int* ntaps, int* pad);

/* A context binds one device and one stream (NULL = a stream the context
 * creates).  Not re-entrant: one host thread at a time per context. */
int rsf_context_create(int device, void* hip_stream, rsf_context** out);
int rsf_context_destroy(rsf_context* ctx);
int rsf_context_set_stream(rsf_context* ctx, void* hip_stream);
/* The caller promises that the image at `image` does not change until
 * rsf_context_release_image (or another hold).  Calls on this context with that
 * image then reuse its preparation (float copy and, in FAST mode, its blur)
 * while the block layout, taps and mode stay the same, instead of recomputing
 * it on every call -- for loops of rsf_step_blocks_device /
 * rsf_step_adjoint_device over one image.  No counterpart in the reference. */
int rsf_context_hold_image(rsf_context* ctx, const void* image);
int rsf_context_release_image(rsf_context* ctx);
/* Bytes of device workspace a plain run (rsf_run_device) of this size
 * allocates with the default kernels -- an upper bound for tuned kernel
 * choices -- not counting the tile lists (well under 1 MB). */
int rsf_workspace_size(const rsf_params* p, int64_t w, int64_t h, int64_t l, int64_t* bytes);

/* ---------------------------------------------------------- fused engine */
/* One iteration phi_out = step(phi_in) -- CUDA_RSF/main.cpp:234-300.
 * phi_in and phi_out must not overlap. */
int rsf_step_device(rsf_context* ctx, const void* image, int image_dtype, const float* phi_in,
                    float* phi_out, int64_t w, int64_t h, int64_t l, const rsf_params* p);
/* rsf_step_device without any device allocation -- no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Its workspace is a
 * static device array of the library (68 MiB per device the library runs on,
 * held until the process ends): the transform and force of at most 2^23
 * voxels (e.g. 203^3) plus the tile plan.  Larger volumes return
 * RSF_ERR_INVALID_ARGUMENT.  LITERAL mode with sigma2 = 0 only (else
 * RSF_ERR_UNSUPPORTED); same bits as rsf_step_device.  Every call on one
 * device shares that device's array: calls that overlap in time on one device
 * (two host threads, or two contexts' streams) race and give wrong results. */
int rsf_step_device_noalloc(rsf_context* ctx, const void* image, int image_dtype, const float* phi_in,
                            float* phi_out, int64_t w, int64_t h, int64_t l, const rsf_params* p);
/* p->iterations steps in place on phi -- the loop of CUDA_RSF/main.cpp:222-314,
 * captured in a hipGraph.  iter_ms (NULL or p->iterations floats) receives
 * the per-iteration device time, the figure main.cpp:309 prints.  After the
 * loop phi is checked: NaN or inf voxels return RSF_ERR_NONFINITE (the
 * reference carries on silently), with phi left as computed.  The same check
 * ends rsf_run_batched_device. */
int rsf_run_device(rsf_context* ctx, const void* image, int image_dtype, float* phi, int64_t w,
                   int64_t h, int64_t l, const rsf_params* p, float* iter_ms);
/* Host-pointer form of rsf_run_device (used by the rsf CLI). */
int rsf_run_host(rsf_context* ctx, const void* image, int image_dtype, float* phi, int64_t w,
                 int64_t h, int64_t l, const rsf_params* p, float* iter_ms);
/* Device time of the most recent rsf_run_device / rsf_run_batched_device
 * iteration loop, in milliseconds (hipEvent around the graph launch). */
int rsf_last_loop_ms(rsf_context* ctx, float* ms);

/* Many uneven sub-volumes in one persistent launch per sweep per iteration,
 * the iteration loop captured in a hipGraph.  Block-local boundary rules
 * (pad band, x-wrap) apply inside each block, as one rsf3d per block would. */
int rsf_run_batched_device(rsf_context* ctx, const rsf_block* blocks, int nblocks,
                           const void* image, int image_dtype, float* phi,
                           const rsf_params* p);
/* The same blocks run one after another through the single-volume engine
 * (the "plain per-block" schedule the batched mode is measured against). */
int rsf_run_blocks_sequential_device(rsf_context* ctx, const rsf_block* blocks, int nblocks,
                                     const void* image, int image_dtype, float* phi,
                                     const rsf_params* p);

/* Runs p->iterations iterations without a graph, with hipEvents around every
 * sweep launch on the context stream (phi is advanced in place), and returns
 * the mean device time of one sweep-1 and one sweep-2 launch in ms. */
int rsf_profile_sweeps(rsf_context* ctx, const rsf_block* blocks, int nblocks, const void* image,
                       int image_dtype, float* phi, const rsf_params* p, float* ms_sweep1,
                       float* ms_sweep2);

/* Time the compiled tile shapes for this problem on this device and keep the
 * fastest for later runs in this process (and in $RSF_TUNE_CACHE if set).
 * choice (nullable) receives a short description of the winner. */
int rsf_autotune(rsf_context* ctx, const rsf_block* blocks, int nblocks, int image_dtype,
                 const rsf_params* p, char* choice, int choice_len);

/* Kernel variants compiled into the library: the candidates rsf_autotune
 * times and RSF_SWEEP1 / RSF_SWEEP2 (= "<name>:<z chunk>") force.  sweep is 1
 * (region statistics) or 2 (stencil update).  rsf_variant_info writes the
 * name (NUL-terminated, truncated to name_len), the tap count it is built
 * for (0 for sweep 2: every tap count) and whether it runs the LITERAL mode.
 * No counterpart in the reference (it has one kernel per operation). */
int rsf_variant_count(int sweep, int* count);
int rsf_variant_info(int sweep, int index, char* name, int name_len, int* taps, int* literal);

/* ------------------------------------------------------- block pipeline
 * The gigavoxel workflow of README.md:6-14,56-59: cut a volume into
 * overlapping blocks, evolve every block, blend the blocks' phi back over the
 * overlaps with linear ramps (interpolating_volumes.py:10-40, the reference's
 * two halves along y generalised to every axis as a tensor product).
 *
 * Grid along each axis of n voxels with block size B and overlap k: B >= n
 * gives one block; otherwise blocks start at j*(B-k) with size B, the last one
 * clipped to n, and a last block shorter than max(min_size, 2k) is folded into
 * the one before it.  B must exceed 2k.  Blocks are numbered x-fastest
 * ((jz*Jy + jy)*Jx + jx) and packed back to back in that order.
 * blocks/origins NULL = size query (nblocks, packed_voxels only); origins
 * receives (x0, y0, z0) per block. */
int rsf_block_grid(int64_t w, int64_t h, int64_t l, int64_t bw, int64_t bh, int64_t bl, int64_t overlap,
                   int64_t min_size, rsf_block* blocks, int64_t* origins, int capacity, int* nblocks,
                   int64_t* packed_voxels);
/* Copies each block (origins[3i..3i+2], dims blocks[i].w/h/l) out of the
 * X-fastest volume into packed + blocks[i].offset.  dtype: RSF_DTYPE_U8 or
 * RSF_DTYPE_F32 (element size of both buffers).  Any block list that fits the
 * volume is accepted (the blocks need not form a grid). */
int rsf_disassemble_device(rsf_context* ctx, const void* volume, int dtype, int64_t w, int64_t h, int64_t l,
                           const rsf_block* blocks, const int64_t* origins, int nblocks, void* packed);
/* Inverse of rsf_disassemble_device for the grid rsf_block_grid(w..min_size)
 * describes: volume[v] = sum over the blocks covering v of
 * (wz*wy)*wx * phi_block, in double, rounded once to float.  Along an axis a
 * voxel i voxels into the overlap of blocks j-1 and j has weights 1-alpha_i and
 * alpha_i, alpha = numpy.linspace(0, 1, k); elsewhere weight 1.  For two
 * blocks along y this is float32 of interpolate_volumes(A, B, k) bit for bit. */
int rsf_reassemble_device(rsf_context* ctx, const float* packed, int64_t w, int64_t h, int64_t l, int64_t bw,
                          int64_t bh, int64_t bl, int64_t overlap, int64_t min_size, float* volume);
/* The whole pipeline on device buffers: grid (min_size = the tap count),
 * disassemble image and phi, one batched run of p->iterations over all
 * blocks, reassemble into phi (in place).  Replaces disassemble.py +
 * rsfbatch.py + reassemble.py (README.md:56-59). */
int rsf_run_volume_blocks_device(rsf_context* ctx, const void* image, int image_dtype, float* phi, int64_t w,
                                 int64_t h, int64_t l, int64_t bw, int64_t bh, int64_t bl, int64_t overlap,
                                 const rsf_params* p);
/* The seed producer's vote (Unet/run_model.py:29-35,86-91): pred holds
 * nz*ny*nx tile outputs of tz*ty*tx floats, tile t = (i*ny + j)*nx + k starting
 * at (x0[k], y0[j], z0[i]) (starts ascending per axis, tiles inside the volume).
 * Every voxel gets the mean of its covering tiles' outputs, summed in float64 in
 * the order of t, into prob, and prob >= 0.5 into mask (0 where no tile covers). */
int rsf_tile_vote_device(rsf_context* ctx, const float* pred, int64_t tz, int64_t ty, int64_t tx,
                         const int64_t* z0, int nz, const int64_t* y0, int ny, const int64_t* x0, int nx,
                         int64_t w, int64_t h, int64_t l, double* prob, uint8_t* mask);
/* Host-pointer forms of rsf_run_batched_device and rsf_run_volume_blocks_device
 * (used by the rsf CLI's --batch and --block modes). */
int rsf_run_batched_host(rsf_context* ctx, const rsf_block* blocks, int nblocks, const void* image,
                         int image_dtype, float* phi, const rsf_params* p);
int rsf_run_volume_blocks_host(rsf_context* ctx, const void* image, int image_dtype, float* phi, int64_t w,
                               int64_t h, int64_t l, int64_t bw, int64_t bh, int64_t bl, int64_t overlap,
                               const rsf_params* p);

/* ------------------------------------------------------ streaming .npy I/O
 * The reference reads and writes whole volumes through host memory
 * (tira::volume load_npy / save_npy, main.cpp:145,149,316).  These stream a
 * file in 32 MiB chunks through two pinned staging buffers, the file reads
 * overlapping the PCIe copies on the context's stream. */
/* Header of a .npy file: ndim, shape (up to 8 dims, C order) and whether it
 * holds uint8/bool. */
int rsf_npy_info(const char* path, int64_t* shape, int* ndim, int* is_u8);
/* A 3-D .npy file into device memory: dst_dtype RSF_DTYPE_U8 copies a uint8
 * file as is; RSF_DTYPE_F32 converts any supported dtype to float32.
 * dst_bytes is the destination's capacity. */
int rsf_npy_load_device(rsf_context* ctx, const char* path, void* dst, int dst_dtype, int64_t dst_bytes);
/* A float32 (l, h, w) device volume into a .npy file. */
int rsf_npy_save_device(rsf_context* ctx, const char* path, const float* src, int64_t w, int64_t h, int64_t l);
/* The whole rsf3d run file to file (main.cpp:143-316): image and seed streamed
 * in, phi0 = sdf(seed - 13) on the device (or the seed file taken as phi0),
 * p->iterations steps, phi streamed out.  iter_ms as in rsf_run_device. */
int rsf_run_npy(rsf_context* ctx, const char* image_path, const char* seed_path, int seed_is_phi0,
                const char* out_path, const rsf_params* p, float* iter_ms);

/* ---------------------------------------------------------------- adjoint */
/* Vector-Jacobian product of one step: given phi (state before the step) and
 * grad_out = dLoss/dphi', writes grad_in = dLoss/dphi and ADDS
 * dLoss/d(lambda1, lambda2, nu, mu) into grad_params[0..3] (device floats).
 * Blocks as in rsf_run_batched_device (nblocks = 1 for a single volume). */
int rsf_step_adjoint_device(rsf_context* ctx, const rsf_block* blocks, int nblocks,
                            const void* image, int image_dtype, const float* phi,
                            const float* grad_out, float* grad_in, float* grad_params,
                            const rsf_params* p);
/* Forward step over blocks (for checkpointed recomputation).  phi_in and
 * phi_out (each spanning the packed blocks) must not overlap. */
int rsf_step_blocks_device(rsf_context* ctx, const rsf_block* blocks, int nblocks,
                           const void* image, int image_dtype, const float* phi_in,
                           float* phi_out, const rsf_params* p);

/* ------------------------------------------------- synthetic data and phi0 */
/* Tube phantom of the benchmark configurations: `tubes` polylines of 5..10
 * segments (length 10..40, radius U(rmin,rmax)) at 200 on a background of 30,
 * a linear x ramp of +-bias, Gaussian noise, clipped and rounded to uint8.
 * mask (nullable) receives the 0/1 tube mask. */
int rsf_synth_tubes(uint8_t* image, uint8_t* mask, int64_t w, int64_t h, int64_t l, int tubes,
                    float rmin, float rmax, float bias, float noise_sigma, uint64_t seed);
/* phi0 of the benchmark configurations: Gaussian blur (sigma = blur_sigma,
 * clamped edges) of the image, threshold `> level`, then the signed Euclidean
 * distance, negative inside, zero level half-way between voxel centres. */
int rsf_initial_phi(const uint8_t* image, float* phi, int64_t w, int64_t h, int64_t l,
                    float blur_sigma, float level);
/* phi0 = sdf(seed - shift) of CUDA_RSF/main.cpp:174-183: the voxels with
 * seed - shift > 0 form the region; phi is the signed Euclidean distance to
 * its boundary, POSITIVE inside that region (the sign of seed - shift). */
int rsf_sdf_from_seed(const float* seed, float* phi, int64_t w, int64_t h, int64_t l, float shift);
/* rsf_sdf_from_seed on device buffers (seed uint8 or float32, by seed_dtype):
 * an exact separable squared EDT, x rows by forward/backward scans, y and z by
 * outward scans.  Below a volume diagonal of 4096 voxels the squared distances
 * are exact in float and the result is the host form's bit for bit; larger
 * volumes carry them in double (8 + 8 bytes of workspace per voxel) and round
 * each distance once to float. */
int rsf_sdf_from_seed_device(rsf_context* ctx, const void* seed, int seed_dtype, float* phi, int64_t w, int64_t h,
                             int64_t l, float shift);
/* rsf_sdf_from_seed_device on host buffers (float32 seed): upload, compute on
 * the device, download (main.cpp:174-183, the rsf CLI's block modes). */
int rsf_sdf_from_seed_host(rsf_context* ctx, const float* seed, float* phi, int64_t w, int64_t h, int64_t l,
                           float shift);

/* ------------------------------------------ mirrors of the reference wrappers
 * Same names (with an rsf_ prefix), argument order and meaning as the host
 * functions of CUDA_RSF/kernel.cu declared at CUDA_RSF/main.cpp:56-71: host
 * float buffers, int dims, synchronous.  Unlike the reference they return a
 * status instead of printing and continuing, and keep device buffers cached
 * between calls instead of allocating and freeing on every call. */
int rsf_heaviside_cal(float* output, float* input, int img_w, int img_h, int img_l);        /* kernel.cu:494  */
int rsf_deri_heaviside_cal(float* output, float* input, int img_w, int img_h, int img_l);   /* kernel.cu:558  */
int rsf_inverse_cal(float* output, float* input, int img_w, int img_h, int img_l);          /* kernel.cu:526  */
int rsf_multiplication_cal(float* output, float* input_1, float* input_2, int img_w, int img_h,
                           int img_l);                                                      /* kernel.cu:590  */
int rsf_division_cal(float* output, float* input_1, float* input_2, int img_w, int img_h,
                     int img_l);                                                            /* kernel.cu:453  */
int rsf_gradient_cal(float* input, float* output_x, float* output_y, float* output_z, int img_w,
                     int img_h, int img_l);                                                 /* kernel.cu:127  */
int rsf_gradientmagnitude_cal(float* output, float* input_x, float* input_y, float* input_z,
                              int img_w, int img_h, int img_l);                             /* kernel.cu:935  */
int rsf_division_nDIV(float* input_x, float* input_y, float* input_z, float* gradientmagnitude,
                      float* output_x, float* output_y, float* output_z, int img_w, int img_h,
                      int img_l);                                                           /* kernel.cu:687  */
int rsf_gradient_cal_3_input(float* input_x, float* input_y, float* input_z, float* output_x,
                             float* output_y, float* output_z, int img_w, int img_h,
                             int img_l);                                                    /* kernel.cu:858  */
int rsf_normalized_div_plus(float* normalize_DIV, float* input_x, float* input_y, float* input_z,
                            int img_w, int img_h, int img_l);                               /* kernel.cu:744  */
int rsf_DIV_cal(float* DIV, float* input_x, float* input_y, float* input_z, int img_w, int img_h,
                int img_l);                                                                 /* kernel.cu:630  */
int rsf_adddevice_convolution_seperable(float* output, float* in_img, int img_w, int img_h,
                                        int img_l, float sigma, float* gkernel,
                                        unsigned int k_size);                               /* kernel.cu:1051 */
int rsf_pad_border_cal(float* output, float* input, int img_w, int img_h, int img_l, int pad_w); /* kernel.cu:331 */
int rsf_adddevice(float* input, float* fout, float* fin, float* Eo, float* Ei, int img_w,
                  int img_h, int img_l, float sigma);                                       /* kernel.cu:1178 */
int rsf_Final_PHI_cal(float* output, float* phi, float* deri_heav, float* Eout, float* Ein,
                      float* normalized_DIV, float* DIV, float e, float dt, float v, float meu,
                      int img_w, int img_h, int img_l);                                     /* kernel.cu:390  */

/* The reference iteration run op by op on the device with the kernels behind
 * the mirrors above (no fusion, FP32-literal): used for tap counts without a
 * fused specialisation and as an in-library cross-check of the fused sweeps. */
int rsf_step_unfused_device(rsf_context* ctx, const float* image, const float* phi_in,
                            float* phi_out, int64_t w, int64_t h, int64_t l, const rsf_params* p);

#ifdef __cplusplus
}
#endif

#endif /* RSF_H */
