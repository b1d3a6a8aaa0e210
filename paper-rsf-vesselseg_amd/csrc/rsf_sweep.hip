// Fused sweeps of one RSF iteration (gfx950).
//
// Sweep 1 (region statistics) replaces CUDA_RSF/main.cpp:234-246,264-293:
// per input z-plane a workgroup stages its (TX+2m) x (TY+2m) tile of the
// transformed level set and the image in LDS, runs the x pass (register
// blocked, 4 outputs per item) and the y pass out of LDS, and pushes the
// plane into a register ring of P partial z sums per column and channel
// (z-marching: each input plane is loaded once per tile).  When a z output is
// complete it forms the local fits fo/fi (constant 1 in the pad band) and the
// force F = lambda*(Eo - Ei), and writes F -- 4 bytes per voxel.
//
// Sweep 2 (stencil update) replaces main.cpp:250-260,296-300: gradients with
// the reference edge rules and x-wrap, |grad|, normals, curvature, Laplacian,
// Dirac and the explicit Euler update; it also writes the transformed level
// set (H or G) that the next sweep 1 reads.
//
// Both sweeps walk a tile list (persistent workgroups), so the same kernels
// serve one volume or many uneven blocks.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <cstdio>

#include "rsf_internal.hpp"

namespace rsf {
namespace {
// Cache policy of the buffer operations whose data no other workgroup of the
// launch reads again: non-temporal (bit 1 of the builtins' aux operand on
// gfx950).  It marks the z kernel's reads of the xy intermediate and its F
// stores, the xy kernel's interleaved stores, and sweep 2's F reads and
// phi'/H(phi') stores, so they do not displace the halo lines that neighbouring
// tiles share in L2 (the window and stencil reads keep the default policy).
// Same bits; at 512^3 the literal loop went from 2.58 to 2.47 ms.
constexpr int kNT = 2;

// WPE: minimum waves per SIMD the register allocation must allow (0 = none),
// so that several workgroups share a CU and hide each other's barrier waits.
// DB: double-buffered LDS stages, one barrier per plane.
template <int P, bool LIT, int TX, int TY, int NT, int WPE, bool DB>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1))) void sweep1_kernel(
    const float* __restrict__ tr, const void* __restrict__ img, int img_u8,
    const float* __restrict__ blurI, float* __restrict__ out0, float* __restrict__ out1,
    float* __restrict__ out2, float* __restrict__ out3, int out_mode, const BlockDesc* __restrict__ blocks,
    const Tile* __restrict__ tiles, int ntiles, int zc, Taps taps, float halfS3, float l1, float l2,
    int same_lambda) {
    constexpr int CH = LIT ? 4 : 2;
    constexpr int M = (P - 1) / 2;
    constexpr int AL = (M + 3) & ~3;        // x halo rounded up to whole float4 groups
    constexpr int SH = AL - M;              // aligned-window columns left of the needed ones
    constexpr int RW = TX + 2 * M;          // LDS row: x0-M .. x0+TX+M-1
    constexpr int RH = TY + 2 * M;
    constexpr int RWP = (RW + 3) & ~3;
    constexpr int QW = (TX + 2 * AL) / 4;   // float4 groups per aligned row
    constexpr int NQ = RH * QW;             // float4 groups per plane
    constexpr int PF = (NQ + NT - 1) / NT;  // groups per thread
    constexpr int COLS = (TX * TY) / NT;
    static_assert(COLS * NT == TX * TY, "tile must be a multiple of the block");
    static_assert(TX % 4 == 0 && P <= 35, "unsupported tile");
    constexpr int RX = 4;
    constexpr int XG = TX / RX;
    constexpr int NC = COLS * CH;
    constexpr int WIN = (RX + P - 1 + 3) & ~3;

    constexpr int NB = DB ? 2 : 1;
    __shared__ __align__(16) float sIn[NB][CH][RH][RWP];
    // x-pass output rows padded by 4 floats: the y pass reads rows j and j+1 of a
    // column in one ds_read2, which would otherwise hit the same bank
    constexpr int TXP = TX + 4;
    __shared__ __align__(16) float sX[NB][CH][RH][TXP];

    const int tid = threadIdx.x;
    const int tx = tid % TX;
    const int ty0 = (tid / TX) * COLS;
    const uint8_t* img8 = (const uint8_t*)img;
    const float* imgf = (const float*)img;

    for (int item = blockIdx.x; item < ntiles; item += gridDim.x) {
        const Tile tl = tiles[item];
        if (tl.b < 0) continue;
        const BlockDesc bd = blocks[tl.b];
        const int w = bd.w, h = bd.h, l = bd.l;
        const int64_t off = bd.off;
        const int64_t plane = (int64_t)w * h;
        const int x0 = tl.x0, y0 = tl.y0, z0 = tl.z0;
        const int zend = min(z0 + zc, l);
        const int nplanes = (zend - z0) + P - 1;
        // float4 path: rows 16-byte aligned and the aligned window inside the row
        const bool vec = ((off | (int64_t)w) & 3) == 0 && x0 - AL >= 0 && x0 + TX + AL <= w &&
                         ((uintptr_t)tr & 15) == 0 &&
                         (img_u8 ? ((uintptr_t)img8 & 3) == 0 : ((uintptr_t)imgf & 15) == 0);

        // prefetch registers: tr and image of this thread's float4 groups
        float4 pt[PF], pi[PF];
        auto fetch = [&](int t) {
            const int zi = min(max(z0 - M + t, 0), l - 1);
            const int64_t zoff = off + (int64_t)zi * plane;
#pragma unroll
            for (int k = 0; k < PF; ++k) {
                const int e = tid + k * NT;
                if (NQ % NT != 0 && e >= NQ) break;
                const int r = e / QW, q = e - r * QW;
                const int gy = min(max(y0 - M + r, 0), h - 1);
                const int gx = x0 - AL + 4 * q;
                const int64_t row = zoff + (int64_t)gy * w;
                if (vec) {
                    pt[k] = *reinterpret_cast<const float4*>(tr + row + gx);
                    if (img_u8) {
                        const uint32_t b4 = *reinterpret_cast<const uint32_t*>(img8 + row + gx);
                        pi[k].x = __uint_as_float(b4);
                    } else {
                        pi[k] = *reinterpret_cast<const float4*>(imgf + row + gx);
                    }
                } else {
                    float tv[4], iv[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int64_t idx = row + min(max(gx + j, 0), w - 1);
                        tv[j] = tr[idx];
                        iv[j] = img_u8 ? (float)img8[idx] : imgf[idx];
                    }
                    pt[k] = make_float4(tv[0], tv[1], tv[2], tv[3]);
                    pi[k] = make_float4(iv[0], iv[1], iv[2], iv[3]);
                }
            }
        };
        auto store = [&](int buf) {
#pragma unroll
            for (int k = 0; k < PF; ++k) {
                const int e = tid + k * NT;
                if (NQ % NT != 0 && e >= NQ) break;
                const int r = e / QW, q = e - r * QW;
                float iv[4];
                if (vec && img_u8) {
                    const uint32_t b4 = __float_as_uint(pi[k].x);
                    iv[0] = (float)(b4 & 0xff);
                    iv[1] = (float)((b4 >> 8) & 0xff);
                    iv[2] = (float)((b4 >> 16) & 0xff);
                    iv[3] = (float)(b4 >> 24);
                } else {
                    iv[0] = pi[k].x; iv[1] = pi[k].y; iv[2] = pi[k].z; iv[3] = pi[k].w;
                }
                const float tv[4] = {pt[k].x, pt[k].y, pt[k].z, pt[k].w};
                float cv[CH][4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    cv[0][j] = iv[j] * tv[j];          // I*H (main.cpp:238) or I*G
                    cv[1][j] = tv[j];                  // H or G
                    if constexpr (LIT) {
                        const float hc = 1.0f - tv[j];  // kernel.cu:295
                        cv[2][j] = iv[j] * hc;          // I*(1-H) (main.cpp:246)
                        cv[3][j] = hc;
                    }
                }
                const int c0 = 4 * q - SH;             // LDS column of the group's first element
#pragma unroll
                for (int ch = 0; ch < CH; ++ch) {
                    if constexpr (SH % 2 == 0) {
                        if (c0 >= 0 && c0 + 1 < RW)
                            *reinterpret_cast<float2*>(&sIn[buf][ch][r][c0]) = make_float2(cv[ch][0], cv[ch][1]);
                        if (c0 + 2 >= 0 && c0 + 3 < RW)
                            *reinterpret_cast<float2*>(&sIn[buf][ch][r][c0 + 2]) = make_float2(cv[ch][2], cv[ch][3]);
                    } else {
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            if (c0 + j >= 0 && c0 + j < RW) sIn[buf][ch][r][c0 + j] = cv[ch][j];
                    }
                }
            }
        };

        // one 1-D array per ring channel (a 2-D array indexed by a run-time slot
        // would not be kept in registers)
        static_assert(NC <= 8, "ring channels");
        float acc0[P], acc1[P], acc2[P], acc3[P], acc4[P], acc5[P], acc6[P], acc7[P];
        auto ring = [&](int n) -> float (&)[P] {
            return n == 0 ? acc0 : n == 1 ? acc1 : n == 2 ? acc2 : n == 3 ? acc3
                 : n == 4 ? acc4 : n == 5 ? acc5 : n == 6 ? acc6 : acc7;
        };
#pragma unroll
        for (int n = 0; n < NC; ++n)
#pragma unroll
            for (int k = 0; k < P; ++k) ring(n)[k] = 0.f;

        // emission operands (image, precomputed blur) of the output plane of plane t
        float eI[COLS], eB[COLS];
        auto load_emit = [&](int t) {
            const int zo = z0 + t - (P - 1);
#pragma unroll
            for (int cc = 0; cc < COLS; ++cc) {
                const int x = min(x0 + tx, w - 1), y = min(y0 + ty0 + cc, h - 1);
                const int64_t i = off + (int64_t)zo * plane + (int64_t)y * w + x;
                eI[cc] = img_u8 ? (float)img8[i] : imgf[i];
                eB[cc] = LIT ? 0.f : blurI[i];
            }
        };
        // ---- x pass of one plane, RX consecutive outputs per work item
        auto xpass = [&](int buf) {
            for (int e = tid; e < CH * RH * XG; e += NT) {
                const int ch = e / (RH * XG);
                const int rem = e - ch * (RH * XG);
                const int r = rem / XG;
                const int xb = (rem - r * XG) * RX;
                float win[WIN];
                const float4* src = reinterpret_cast<const float4*>(&sIn[buf][ch][r][xb]);
#pragma unroll
                for (int q = 0; q < WIN / 4; ++q) {
                    const float4 v4 = src[q];
                    win[4 * q + 0] = v4.x;
                    win[4 * q + 1] = v4.y;
                    win[4 * q + 2] = v4.z;
                    win[4 * q + 3] = v4.w;
                }
                float o4[RX];
#pragma unroll
                for (int o = 0; o < RX; ++o) {
                    float a = win[o] * taps.g[0];
#pragma unroll
                    for (int k = 1; k < P; ++k) a = fmaf(win[o + k], taps.g[k], a);
                    o4[o] = a;
                }
                *reinterpret_cast<float4*>(&sX[buf][ch][r][xb]) = make_float4(o4[0], o4[1], o4[2], o4[3]);
            }
        };
        // ---- y pass of plane t for this thread's COLS rows, z ring, emission
        auto yzemit = [&](int buf, int t) {
            float v[NC];
#pragma unroll
            for (int ch = 0; ch < CH; ++ch) {
                float col[COLS + P - 1];
#pragma unroll
                for (int j = 0; j < COLS + P - 1; ++j) col[j] = sX[buf][ch][ty0 + j][tx];
#pragma unroll
                for (int cc = 0; cc < COLS; ++cc) {
                    float a = col[cc] * taps.g[0];
#pragma unroll
                    for (int k = 1; k < P; ++k) a = fmaf(col[cc + k], taps.g[k], a);
                    v[cc * CH + ch] = a;
                }
            }
            // ---- z ring.  ring(n)[j] collects the output whose first plane is = j
            // (mod P); plane t (phase s = t mod P) adds tap g[(s - j) mod P] to it,
            // i.e. the rotated row taps.g[P-1-s .. 2P-2-s] (the taps are symmetric),
            // and completes accumulator (s+1) mod P, which restarts from zero.  Each
            // output thus sums g[0]*x0 + ... + g[P-1]*x(P-1) in order, the summation
            // order of Convolution__on_Z (kernel.cu:1044-1045).  The accumulators stay
            // in fixed registers; only the finished one is read by index (uniform).
            const int s = t % P;
            const float* trow = &taps.g[P - 1 - s];
#pragma unroll
            for (int n = 0; n < NC; ++n)
#pragma unroll
                for (int j = 0; j < P; ++j) ring(n)[j] = fmaf(v[n], trow[j], ring(n)[j]);
            const int js = s == P - 1 ? 0 : s + 1;
            float done[NC];
#pragma unroll
            for (int n = 0; n < NC; ++n) {
                done[n] = ring(n)[js];
                ring(n)[js] = 0.f;
            }
            if (t < P - 1) return;
            // ---- emission of output plane zo
            const int zo = z0 + t - (P - 1);
            const bool zvalid = zo >= M && zo < l - M;
#pragma unroll
            for (int cc = 0; cc < COLS; ++cc) {
                const int x = x0 + tx, y = y0 + ty0 + cc;
                if (x >= w || y >= h) continue;
                const int64_t i = off + (int64_t)zo * plane + (int64_t)y * w + x;
                const float Iv = eI[cc];
                float fo = 1.0f, fi = 1.0f;  // pad band: 30/30 (kernel.cu:313, main.cpp:277)
                float bH = 30.0f, bHc = 30.0f;
                if (zvalid && x >= M && x < w - M && y >= M && y < h - M) {
                    if constexpr (LIT) {
                        bH = done[cc * CH + 1];
                        bHc = done[cc * CH + 3];
                        fo = done[cc * CH + 0] / bH;
                        fi = done[cc * CH + 2] / bHc;
                    } else {
                        fast_fits(eB[cc], done[cc * CH + 0], done[cc * CH + 1], halfS3, fo, fi, bH, bHc);
                    }
                }
                if (out_mode == 0) {
                    out0[i] = force_of(Iv, fo, fi, l1, l2, same_lambda);
                } else {
                    out0[i] = fo;
                    out1[i] = fi;
                    if (out_mode == 2) {
                        out2[i] = bH;
                        out3[i] = bHc;
                    }
                }
            }
        };

        fetch(0);
        if constexpr (!DB) {
            // two barriers per plane: LDS input -> x pass -> y pass
            for (int t = 0; t < nplanes; ++t) {
                store(0);
                if (t + 1 < nplanes) fetch(t + 1);
                if (t >= P - 1) load_emit(t);
                __syncthreads();
                xpass(0);
                __syncthreads();
                yzemit(0, t);
            }
        } else {
            // one barrier per plane: the x pass of plane t and the y pass, z ring and
            // emission of plane t-1 share a barrier interval on alternating buffers
            for (int t = 0; t <= nplanes; ++t) {
                if (t < nplanes) {
                    store(t & 1);
                    if (t + 1 < nplanes) fetch(t + 1);
                }
                if (t >= P) load_emit(t - 1);
                __syncthreads();
                if (t < nplanes) xpass(t & 1);
                if (t >= 1) yzemit((t - 1) & 1, t - 1);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Sweep 1 in two kernels.  The fused kernel above keeps a z ring of CH x P
// accumulators per voxel in registers next to its LDS passes, which caps its
// occupancy.  Here the x and y passes run alone (xy kernel, LDS only) and write
// the CH xy-blurred channels to a scratch volume; the z kernel marches each
// (x, y) column through z with the register ring and emits F.  Same operations
// in the same order as the fused kernel: identical results.  Channels: fast
// mode I*G and G; literal mode I*H, H, I*(1-H), 1-H (main.cpp:238-246).

// 16- and 8-byte buffer-load results
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// Channel values of one voxel from the transform t = T(phi) and the image i.
template <bool LIT>
__device__ __forceinline__ void sweep1_channels(float i, float t, float (&c)[LIT ? 4 : 2]) {
    c[0] = i * t;                 // I*H (main.cpp:238) or I*G
    c[1] = t;                     // H or G
    if constexpr (LIT) {
        const float hc = 1.0f - t;  // kernel.cu:295
        c[2] = i * hc;              // I*(1-H) (main.cpp:246)
        c[3] = hc;
    }
}

// xy kernel: tiles of TX x TY outputs inside the valid region [M, w-M) x
// [M, h-M) of a block, zc planes per work item.  Per plane the (TY+2M) x
// (TX+2M) window of the channels goes to LDS, the x pass writes (TY+2M) x TX
// outputs to LDS (RX = 4 per work item), the y pass R outputs per work item
// from R+P-1 LDS reads.  The work items of both passes are exact multiples of
// the thread count.  Channel ch of the output goes to xy + ch * stride.
//
// IMG: the image dtype fixed at compile time (1 uint8, 0 float32; -1 takes the
// img_u8 argument), so that a uint8 image's prefetch keeps one packed word per
// four voxels in registers.
// ONLY (fast mode): the xy blur of the image alone, into channel 0 (tr is not
// read): the first half of the padded image blur the fast fits need.
// IL (literal mode): the output interleaved per voxel, channels 0..3 of voxel v
// at xy[4v .. 4v+3]; each y-pass item blurs one channel pair (0,1) or (2,3) of a
// column and stores it as one 8-byte word, and the two halves of a wave take the
// two pairs of the same columns, so a wave's stores cover whole 16-byte voxels.
// (The fast kernel has no registers to spare for it: its variants spilled.)
template <int P, bool LIT, int TX, int TY, int NT, int R, int WPE, int HI = 0, int RX = 4, int UX = 1,
          int UY = 1, int IMG = -1, bool ONLY = false, bool IL = false>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void sweep1_xy_kernel(
    const float* __restrict__ tr, const void* __restrict__ img, int img_u8_arg, float* __restrict__ xy,
    int64_t stride, const BlockDesc* __restrict__ blocks, const Tile* __restrict__ tiles, int ntiles, int zc,
    Taps taps) {
    const int img_u8 = IMG >= 0 ? IMG : img_u8_arg;
    // HI (literal mode): LDS holds H and I, and each x-pass item forms a pair
    // of channels from them -- (I*H, H) or (I*(1-H), 1-H) -- before blurring
    // (HI = 1), or all four (HI = 2: one window read and one 1-H per element
    // for four channels), which halves the LDS input tile.
    static_assert(!HI || LIT, "HI is a literal-mode layout");
    static_assert(!ONLY || !LIT, "the image-only blur is a fast-mode pass");
    static_assert(!IL || (LIT && TX == 32 && !ONLY), "the interleaved layout: literal, 32-column tiles");
    constexpr int CH = LIT ? 4 : 2;
    constexpr int CIN = ONLY ? 1 : (HI ? 2 : CH);   // LDS input arrays (the image-only blur: one)
    constexpr int M = (P - 1) / 2;
    constexpr int RH = TY + 2 * M, RW = TX + 2 * M;
    constexpr int RWP = (RW + 3) & ~3;
    constexpr int TXP = TX + 4;
    constexpr int QW = RWP / 4;            // float4 groups per LDS input row
    constexpr int NQ = RH * QW;
    constexpr int PF = (NQ + NT - 1) / NT;
    constexpr int XG = TX / RX;  // RX: x-pass outputs per work item (4 or 8)
    static_assert(RX == 4 || RX == 8, "x-pass outputs per item");
    // UX, UY: unroll of the x- and y-pass item loops.  The LDS footprint caps the
    // kernel at two workgroups per CU, so registers are free to hold the next
    // item's loads in flight while the current item computes.
    // x- and y-pass work items; the image-only blur has one channel, and its
    // loops stop at the item count (whole idle waves in the last round)
    constexpr int NXI = ONLY || HI == 2 ? RH * XG : (HI ? 2 : CH) * RH * XG;
    constexpr int NYI = ONLY ? TX * (TY / R) : (IL ? CH / 2 : CH) * TX * (TY / R);
    constexpr int NXR = (NXI + NT - 1) / NT, NYR = (NYI + NT - 1) / NT;
    constexpr int WIN = (RX + P - 1 + 3) & ~3;
    static_assert(ONLY || HI == 2 || (NXI % NT == 0 && NYI % NT == 0), "work items must be multiples of the threads");
    static_assert(TY % R == 0, "y-pass rows");
    __shared__ __align__(16) float sIn[CIN][RH][RWP];
    // x-pass output: channel planes CS floats apart, CS = TX (mod 64), so the
    // groups of TX lanes of a wave in the y pass (same rows, neighbouring
    // channels) read disjoint LDS banks; in the interleaved y pass the two
    // halves are two channels apart: CS = 16 (mod 32)
    constexpr int CS0 = RH * TXP;
    constexpr int CS = IL ? CS0 + (((16 - CS0 % 32) % 32) + 32) % 32
                          : CS0 + (((TX % 64 - CS0 % 64) % 64) + 64) % 64;
    __shared__ __align__(16) float sX[(ONLY ? 1 : CH) * CS];
    const int tid = threadIdx.x;
    const uint8_t* img8 = (const uint8_t*)img;
    const float* imgf = (const float*)img;

    for (int item = blockIdx.x; item < ntiles; item += gridDim.x) {
        const Tile tl = tiles[item];
        if (tl.b < 0) continue;
        const BlockDesc bd = blocks[tl.b];
        const int w = bd.w, h = bd.h, l = bd.l;
        const int64_t off = bd.off, plane = (int64_t)w * h;
        const int x0 = tl.x0, y0 = tl.y0, z0 = tl.z0;
        const int zend = min(z0 + zc, l);
        // four-channel literal tiles past the valid region's last row (e.g. 492
        // rows in tiles of 60) skip the window rows and y-pass rows they do not
        // need (in the fast kernel the extra branch cost more than it saved)
        const int ny = min(TY, h - M - y0), rxv = ny + 2 * M;
        const bool vec = ((off | (int64_t)w) & 3) == 0 && ((uintptr_t)tr & 15) == 0 &&
                         (img_u8 ? ((uintptr_t)img8 & 3) == 0 : ((uintptr_t)imgf & 15) == 0);
        float4 pt[PF], pi[PF];
        // 16-byte aligned blocks: each prefetch round's element group as an index
        // inside a plane (the window's rows clamped to the block, its last group
        // to w-4: columns past w-1 only feed outputs that are not stored; the last
        // round's spare threads repeat the last group, and store() skips them),
        // loaded through buffer resources based at the plane
        int eidx[PF];
        if (vec) {
#pragma unroll
            for (int k = 0; k < PF; ++k) {
                const int e = min(tid + k * NT, NQ - 1);
                const int r = e / QW, q = e - r * QW;
                eidx[k] = min(y0 - M + r, h - 1) * w + min(x0 - M + 4 * q, w - 4);
            }
        }
        auto fetch = [&](int z) {
            const int64_t zoff = off + (int64_t)z * plane;
            if (vec) {
                const __amdgpu_buffer_rsrc_t rT =
                    __builtin_amdgcn_make_buffer_rsrc((void*)(tr + zoff), 0, (int)0xffffffff, 0x00020000);
                const __amdgpu_buffer_rsrc_t rI = __builtin_amdgcn_make_buffer_rsrc(
                    img_u8 ? (void*)(img8 + zoff) : (void*)(imgf + zoff), 0, (int)0xffffffff, 0x00020000);
#pragma unroll
                for (int k = 0; k < PF; ++k) {
                    const uint32_t vo = 4u * (uint32_t)eidx[k];
                    if constexpr (!ONLY) {
                        const u32x4 b = __builtin_amdgcn_raw_buffer_load_b128(rT, (int)vo, 0, 0);
                        pt[k] = make_float4(__uint_as_float(b[0]), __uint_as_float(b[1]), __uint_as_float(b[2]),
                                            __uint_as_float(b[3]));
                    }
                    if (img_u8) {
                        pi[k].x = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rI, (int)(vo >> 2), 0, 0));
                    } else {
                        const u32x4 b = __builtin_amdgcn_raw_buffer_load_b128(rI, (int)vo, 0, 0);
                        pi[k] = make_float4(__uint_as_float(b[0]), __uint_as_float(b[1]), __uint_as_float(b[2]),
                                            __uint_as_float(b[3]));
                    }
                }
                return;
            }
#pragma unroll
            for (int k = 0; k < PF; ++k) {
                const int e = tid + k * NT;
                if (NQ % NT != 0 && e >= NQ) break;
                const int r = e / QW, q = e - r * QW;
                const int gy = min(y0 - M + r, h - 1);
                const int gx = x0 - M + 4 * q;
                const int64_t row = zoff + (int64_t)gy * w;
                if (vec && gx + 3 < w) {
                    if constexpr (!ONLY) pt[k] = *reinterpret_cast<const float4*>(tr + row + gx);
                    if (img_u8) {
                        pi[k].x = __uint_as_float(*reinterpret_cast<const uint32_t*>(img8 + row + gx));
                    } else {
                        pi[k] = *reinterpret_cast<const float4*>(imgf + row + gx);
                    }
                } else {
                    float tv[4], iv[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int64_t idx = row + min(gx + j, w - 1);
                        tv[j] = ONLY ? 0.f : tr[idx];
                        iv[j] = img_u8 ? (float)img8[idx] : imgf[idx];
                    }
                    pt[k] = make_float4(tv[0], tv[1], tv[2], tv[3]);
                    pi[k] = make_float4(iv[0], iv[1], iv[2], iv[3]);
                    if (img_u8) pi[k].x = __uint_as_float((uint32_t)iv[0] | ((uint32_t)iv[1] << 8) |
                                                          ((uint32_t)iv[2] << 16) | ((uint32_t)iv[3] << 24));
                }
            }
        };
        auto store = [&]() {
#pragma unroll
            for (int k = 0; k < PF; ++k) {
                const int e = tid + k * NT;
                if (NQ % NT != 0 && e >= NQ) break;
                const int r = e / QW, q = e - r * QW;
                float iv[4];
                if (img_u8) {
                    const uint32_t b4 = __float_as_uint(pi[k].x);
                    iv[0] = (float)(b4 & 0xff);
                    iv[1] = (float)((b4 >> 8) & 0xff);
                    iv[2] = (float)((b4 >> 16) & 0xff);
                    iv[3] = (float)(b4 >> 24);
                } else {
                    iv[0] = pi[k].x; iv[1] = pi[k].y; iv[2] = pi[k].z; iv[3] = pi[k].w;
                }
                if constexpr (ONLY) {
                    *reinterpret_cast<float4*>(&sIn[0][r][4 * q]) = make_float4(iv[0], iv[1], iv[2], iv[3]);
                } else if constexpr (HI) {
                    *reinterpret_cast<float4*>(&sIn[0][r][4 * q]) = pt[k];
                    *reinterpret_cast<float4*>(&sIn[1][r][4 * q]) = make_float4(iv[0], iv[1], iv[2], iv[3]);
                } else {
                    const float tv[4] = {pt[k].x, pt[k].y, pt[k].z, pt[k].w};
                    float cv[4][CH];
#pragma unroll
                    for (int j = 0; j < 4; ++j) sweep1_channels<LIT>(iv[j], tv[j], cv[j]);
#pragma unroll
                    for (int ch = 0; ch < CH; ++ch)
                        *reinterpret_cast<float4*>(&sIn[ch][r][4 * q]) =
                            make_float4(cv[0][ch], cv[1][ch], cv[2][ch], cv[3][ch]);
                }
            }
        };
        fetch(z0);
        for (int z = z0; z < zend; ++z) {
            store();
            if (z + 1 < zend) fetch(z + 1);
            __syncthreads();
            // ---- x pass (one item at a time: registers, not latency, are the limit)
            if constexpr (HI == 2) {
                // items that would fill only part of a last round (e.g. 640 items on
                // 256 threads) run that round as channel-pair halves over all the
                // threads, so that every wave issues the same x-pass work
                constexpr int NF = NXI / NT, REM = NXI % NT;
                constexpr bool SPLIT = REM > 0 && 2 * REM == NT;
#pragma unroll UX
                for (int k = 0; k < (SPLIT ? NF : NXR); ++k) {
                    const int e = tid + k * NT;
                    if (!SPLIT && NXI % NT != 0 && e >= NXI) break;
                    const int r = e / XG;
                    if (r >= rxv) continue;
                    const int xb = (e - r * XG) * RX;
                    float cw[4][WIN];
                    const float4* sh = reinterpret_cast<const float4*>(&sIn[0][r][xb]);
                    const float4* si = reinterpret_cast<const float4*>(&sIn[1][r][xb]);
#pragma unroll
                    for (int q = 0; q < WIN / 4; ++q) {
                        const float4 h4 = sh[q], i4 = si[q];
                        const float hv[4] = {h4.x, h4.y, h4.z, h4.w}, iv[4] = {i4.x, i4.y, i4.z, i4.w};
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            // the channel values of sweep1_channels
                            const float hc = 1.0f - hv[j];
                            cw[0][4 * q + j] = iv[j] * hv[j];
                            cw[1][4 * q + j] = hv[j];
                            cw[2][4 * q + j] = iv[j] * hc;
                            cw[3][4 * q + j] = hc;
                        }
                    }
#pragma unroll
                    for (int ch = 0; ch < 4; ++ch) {
                        float oc[RX];
#pragma unroll
                        for (int o = 0; o < RX; ++o) {
                            float a = cw[ch][o] * taps.g[0];
#pragma unroll
                            for (int kk = 1; kk < P; ++kk) a = fmaf(cw[ch][o + kk], taps.g[kk], a);
                            oc[o] = a;
                        }
#pragma unroll
                        for (int o = 0; o < RX; o += 4)
                            *reinterpret_cast<float4*>(&sX[ch * CS + r * TXP + xb + o]) =
                                make_float4(oc[o], oc[o + 1], oc[o + 2], oc[o + 3]);
                    }
                }
                if constexpr (SPLIT) {
                    // pair 0: (I*H, H), pair 1: (I*(1-H), 1-H), the values of sweep1_channels
                    auto half = [&](auto PR_) __attribute__((always_inline)) {
                        constexpr int pr = decltype(PR_)::value;
                        const int e = NF * NT + tid % REM;
                        const int r = e / XG;
                        if (r >= rxv) return;
                        const int xb = (e - r * XG) * RX;
                        float wa[WIN], wb[WIN];
                        const float4* sh = reinterpret_cast<const float4*>(&sIn[0][r][xb]);
                        const float4* si = reinterpret_cast<const float4*>(&sIn[1][r][xb]);
#pragma unroll
                        for (int q = 0; q < WIN / 4; ++q) {
                            const float4 h4 = sh[q], i4 = si[q];
                            const float hv[4] = {h4.x, h4.y, h4.z, h4.w}, iv[4] = {i4.x, i4.y, i4.z, i4.w};
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                const float t = pr == 0 ? hv[j] : 1.0f - hv[j];
                                wa[4 * q + j] = iv[j] * t;
                                wb[4 * q + j] = t;
                            }
                        }
                        float oa[RX], ob[RX];
#pragma unroll
                        for (int o = 0; o < RX; ++o) {
                            float a = wa[o] * taps.g[0], b = wb[o] * taps.g[0];
#pragma unroll
                            for (int kk = 1; kk < P; ++kk) {
                                a = fmaf(wa[o + kk], taps.g[kk], a);
                                b = fmaf(wb[o + kk], taps.g[kk], b);
                            }
                            oa[o] = a;
                            ob[o] = b;
                        }
#pragma unroll
                        for (int o = 0; o < RX; o += 4) {
                            *reinterpret_cast<float4*>(&sX[(2 * pr) * CS + r * TXP + xb + o]) =
                                make_float4(oa[o], oa[o + 1], oa[o + 2], oa[o + 3]);
                            *reinterpret_cast<float4*>(&sX[(2 * pr + 1) * CS + r * TXP + xb + o]) =
                                make_float4(ob[o], ob[o + 1], ob[o + 2], ob[o + 3]);
                        }
                    };
                    // REM is a multiple of the wave size: the choice is wave-uniform
                    if (tid < REM) half(std::integral_constant<int, 0>{});
                    else half(std::integral_constant<int, 1>{});
                }
            } else if constexpr (HI) {
#pragma unroll UX
                for (int k = 0; k < NXI / NT; ++k) {
                    const int e = tid + k * NT;
                    const int pr = e / (RH * XG);
                    const int rem = e - pr * (RH * XG);
                    const int r = rem / XG;
                    const int xb = (rem - r * XG) * RX;
                    float wa[WIN], wb[WIN];
                    const float4* sh = reinterpret_cast<const float4*>(&sIn[0][r][xb]);
                    const float4* si = reinterpret_cast<const float4*>(&sIn[1][r][xb]);
#pragma unroll
                    for (int q = 0; q < WIN / 4; ++q) {
                        const float4 h4 = sh[q], i4 = si[q];
                        const float hv[4] = {h4.x, h4.y, h4.z, h4.w}, iv[4] = {i4.x, i4.y, i4.z, i4.w};
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            // the channel values of sweep1_channels: (I*H, H) or (I*(1-H), 1-H)
                            const float t = pr == 0 ? hv[j] : 1.0f - hv[j];
                            wa[4 * q + j] = iv[j] * t;
                            wb[4 * q + j] = t;
                        }
                    }
                    float oa[RX], ob[RX];
#pragma unroll
                    for (int o = 0; o < RX; ++o) {
                        float a = wa[o] * taps.g[0], b = wb[o] * taps.g[0];
#pragma unroll
                        for (int kk = 1; kk < P; ++kk) {
                            a = fmaf(wa[o + kk], taps.g[kk], a);
                            b = fmaf(wb[o + kk], taps.g[kk], b);
                        }
                        oa[o] = a;
                        ob[o] = b;
                    }
#pragma unroll
                    for (int o = 0; o < RX; o += 4) {
                        *reinterpret_cast<float4*>(&sX[(2 * pr) * CS + r * TXP + xb + o]) =
                            make_float4(oa[o], oa[o + 1], oa[o + 2], oa[o + 3]);
                        *reinterpret_cast<float4*>(&sX[(2 * pr + 1) * CS + r * TXP + xb + o]) =
                            make_float4(ob[o], ob[o + 1], ob[o + 2], ob[o + 3]);
                    }
                }
            } else
#pragma unroll UX
            for (int k = 0; k < NXR; ++k) {
                const int e = tid + k * NT;
                if (NXI % NT != 0 && e >= NXI) break;
                const int ch = e / (RH * XG);
                const int rem = e - ch * (RH * XG);
                const int r = rem / XG;
                const int xb = (rem - r * XG) * RX;
                float win[WIN];
                const float4* src = reinterpret_cast<const float4*>(&sIn[ch][r][xb]);
#pragma unroll
                for (int q = 0; q < WIN / 4; ++q) {
                    const float4 v4 = src[q];
                    win[4 * q + 0] = v4.x;
                    win[4 * q + 1] = v4.y;
                    win[4 * q + 2] = v4.z;
                    win[4 * q + 3] = v4.w;
                }
                float o4[RX];
#pragma unroll
                for (int o = 0; o < RX; ++o) {
                    float a = win[o] * taps.g[0];
#pragma unroll
                    for (int kk = 1; kk < P; ++kk) a = fmaf(win[o + kk], taps.g[kk], a);
                    o4[o] = a;
                }
#pragma unroll
                for (int o = 0; o < RX; o += 4)
                    *reinterpret_cast<float4*>(&sX[ch * CS + r * TXP + xb + o]) =
                        make_float4(o4[o], o4[o + 1], o4[o + 2], o4[o + 3]);
            }
            __syncthreads();
            // ---- y pass: R outputs of one column, rows r0 .. r0+R-1
#pragma unroll UY
            for (int k = 0; k < NYR; ++k) {
                // item e: column cx = e % TX, then channel, then row group, so that
                // with TX = 32 the two halves of a wave take two channels of the same rows
                const int e = tid + k * NT;
                if (NYI % NT != 0 && e >= NYI) break;
                const int cx = e % TX, t = e / TX;
                if constexpr (IL) {
                    // a wave's two halves take the two pairs of the same 32 columns
                    // and rows: the row group is wave-uniform
                    const int pr = t & 1, r0 = __builtin_amdgcn_readfirstlane((t >> 1) * R);
                    if (r0 >= ny) continue;
                    float ca[R + P - 1], cb[R + P - 1];
#pragma unroll
                    for (int j = 0; j < R + P - 1; ++j) {
                        ca[j] = sX[(2 * pr) * CS + (r0 + j) * TXP + cx];
                        cb[j] = sX[(2 * pr + 1) * CS + (r0 + j) * TXP + cx];
                    }
                    const int gx = x0 + cx;
                    const bool xin = gx < w - M;
                    // the plane's interleaved channels through a buffer resource: the
                    // column offset per lane, the row offset in an SGPR (16 w h < 2^32
                    // on this layout)
                    const __amdgpu_buffer_rsrc_t rX = __builtin_amdgcn_make_buffer_rsrc(
                        (void*)(xy + 4 * (off + (int64_t)z * plane)), 0, (int)0xffffffff, 0x00020000);
                    const int vo = 16 * gx + 8 * pr;
#pragma unroll
                    for (int o = 0; o < R; ++o) {
                        float a = ca[o] * taps.g[0], b = cb[o] * taps.g[0];
#pragma unroll
                        for (int kk = 1; kk < P; ++kk) {
                            a = fmaf(ca[o + kk], taps.g[kk], a);
                            b = fmaf(cb[o + kk], taps.g[kk], b);
                        }
                        const int gy = y0 + r0 + o;
                        if (gy < h - M && xin) {
                            const u32x2 v2 = {__float_as_uint(a), __float_as_uint(b)};
                            __builtin_amdgcn_raw_buffer_store_b64(v2, rX, vo, (int)(16u * (uint32_t)gy * (uint32_t)w), kNT);
                        }
                    }
                    continue;
                }
                const int ch = ONLY ? 0 : t % CH, r0 = (ONLY ? t : t / CH) * R;
                float col[R + P - 1];
#pragma unroll
                for (int j = 0; j < R + P - 1; ++j) col[j] = sX[ch * CS + (r0 + j) * TXP + cx];
                const int gx = x0 + cx;
                float* dst = xy + ch * stride + off + (int64_t)z * plane + gx;
#pragma unroll
                for (int o = 0; o < R; ++o) {
                    float a = col[o] * taps.g[0];
#pragma unroll
                    for (int kk = 1; kk < P; ++kk) a = fmaf(col[o + kk], taps.g[kk], a);
                    const int gy = y0 + r0 + o;
                    if (gx < w - M && gy < h - M) dst[(int64_t)gy * w] = a;
                }
            }
        }
        __syncthreads();
    }
}

// xy kernel over column strips: each wave blurs one strip of SW = (64 / CH) NCOL
// output columns of one plane along x and y, marching the plane's rows.
// Per input row the wave stages the channel values (sweep1_channels) of the
// strip's SW + P - 1 elements in LDS; lane (channel ch, column group cg) runs the
// x pass of its NCOL columns from a register window and adds the outputs to a
// register ring of P accumulators per column, the z kernel's ring along y:
// accumulator j holds the output whose first row is j (mod P), row t adds tap
// g[(t - j) mod P] to it -- starting it with the product at t = j -- and
// completes accumulator (t + 1) mod P, which sums g[0] x_0 + ... + g[P-1] x_(P-1)
// in order, as the tiled y pass does.  Every x-pass row is computed once (a
// tile of TY rows recomputes P - 1 halo rows) and the waves run without
// barriers, each in its own LDS.  Output row t - M completes at input row t.
// Literal mode writes the interleaved layout (the four channels of a voxel in 16
// bytes), transposed through LDS so that each lane stores one voxel; fast mode
// the two channel planes, each lane its NCOL columns.  Work
// items: (block, strip at column x0, first plane z0), waves taking planes z0 +
// wave, z0 + wave + NW, ... below z0 + zc.  D rows of loads stay in flight.
template <int P, bool LIT, int NCOL, int NW, int D, int WPE, int IMG>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(WPE))) void sweep1_xy_strip_kernel(
    const float* __restrict__ tr, const void* __restrict__ img, int img_u8_arg, float* __restrict__ xy,
    int64_t stride, const BlockDesc* __restrict__ blocks, const Tile* __restrict__ tiles, int ntiles, int zc,
    Taps taps) {
    constexpr int CH = LIT ? 4 : 2;
    constexpr int M = (P - 1) / 2;
    constexpr int NCG = 64 / CH;              // column groups per wave
    constexpr int SW = NCG * NCOL;            // output columns per strip
    constexpr int SE = SW + P - 1;            // input elements per strip row
    constexpr int NE = (SE + 63) / 64;        // staged elements per lane
    constexpr int NWIN = NCOL + P - 1;        // a lane's x-pass window
    constexpr int VW = NCOL % 4 == 0 ? 4 : 2; // LDS read width of the window
    constexpr int NV = (NWIN + VW - 1) / VW;
    constexpr int SEP = 64 * NE;              // staged row length (every lane stages NE elements)
    static_assert(NCOL % VW == 0 && (SW <= 64 || !LIT) && (NCG - 1) * NCOL + NV * VW <= SEP, "strip shape");
    static_assert(P % D == 0 && NCOL <= M, "P a multiple of the load queue depth, NCOL at most M");
    typedef float fvec __attribute__((ext_vector_type(NCOL), aligned(4)));   // a lane's columns (fast mode)
    const int img_u8 = IMG >= 0 ? IMG : img_u8_arg;
    __shared__ __align__(16) float sRow[NW][CH][SEP];
    // the transposition: voxel v's channels at 4 v + 4 (v / 16), so that the writes of
    // one column offset o by all 64 lanes (voxels 4 g + o, 4 channels) fall in 64
    // distinct banks (unpadded, lanes g and g + 4 collide)
    __shared__ __align__(16) float sT[LIT ? NW : 1][LIT ? SW * CH + 4 * (SW / 16) : 1];
    auto tpos = [](int v) { return 4 * v + 4 * (v >> 4); };
    // the wave index in an SGPR: the planes, and with them the buffer resources, are wave-uniform
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ch = lane / NCG, cg = lane % NCG;
    const uint8_t* img8 = (const uint8_t*)img;
    const float* imgf = (const float*)img;
    auto rsrc = [](const void* q) {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(q), 0, (int)0xffffffff, 0x00020000);
    };
    for (int item = blockIdx.x; item < ntiles; item += gridDim.x) {
        const Tile tl = tiles[item];
        if (tl.b < 0) continue;
        const BlockDesc bd = blocks[tl.b];
        const int w = bd.w, h = bd.h, l = bd.l;
        const int64_t off = bd.off, plane = (int64_t)w * h;
        const int x0 = tl.x0;
        const int zend = min(tl.z0 + zc, l);
        // staged elements: lane + 64 e of the strip, columns x0 - M + ... clamped to
        // the block (columns past w - 1, and elements past the strip's SE, feed only
        // outputs that are not stored; every lane stages NE elements, without branches)
        int gxe[NE];
#pragma unroll
        for (int e = 0; e < NE; ++e) gxe[e] = min(x0 - M + lane + 64 * e, w - 1);
        // the voxel this lane stores; lanes past the strip repeat its last voxel (same
        // address and value), and columns past w - 1 store at w - 1: columns from
        // w - M on are pad band, whose intermediate the z kernel reads into sums it
        // discards, so the ring's stores need no branch
        const int gxo = min(x0 + min(lane, SW - 1), w - 1);
        // fast mode: the lane's first column, groups past w - NCOL at w - NCOL (all
        // their columns are pad band, M >= NCOL)
        const int gxc = min(x0 + cg * NCOL, w - NCOL);
        for (int z = tl.z0 + wv; z < zend; z += NW) {
            const int64_t zoff = off + (int64_t)z * plane;
            const __amdgpu_buffer_rsrc_t rT = rsrc(tr + zoff);
            const __amdgpu_buffer_rsrc_t rI = rsrc(img_u8 ? (const void*)(img8 + zoff) : (const void*)(imgf + zoff));
            const __amdgpu_buffer_rsrc_t rX = rsrc(xy + 4 * zoff);
            // row t of the strip: H and I of the lane's staged elements
            // (raw bits: a conversion at the load would wait for it)
            uint32_t qh[D][NE], qi[D][NE];
            auto load_row = [&](int k, int t) {
                const uint32_t ro = (uint32_t)t * (uint32_t)w;   // element offset of the row
#pragma unroll
                for (int e = 0; e < NE; ++e) {
                    qh[k][e] = __builtin_amdgcn_raw_buffer_load_b32(rT, 4 * gxe[e], (int)(4u * ro), 0);
                    if (img_u8)
                        qi[k][e] = __builtin_amdgcn_raw_buffer_load_b8(rI, gxe[e], (int)ro, 0);
                    else
                        qi[k][e] = __builtin_amdgcn_raw_buffer_load_b32(rI, 4 * gxe[e], (int)(4u * ro), 0);
                }
            };
            float acc[NCOL][P];
#pragma unroll
            for (int o = 0; o < NCOL; ++o)
#pragma unroll
                for (int j = 0; j < P; ++j) acc[o][j] = 0.f;
            // groups of P rows, every step run: rows past h - 1 load row h - 1 and
            // store at row h - 1 (output rows from h - M on are pad band, whose
            // intermediate the z kernel reads into sums it discards); without
            // early exits the unrolled group keeps the ring in fixed registers
            auto group = [&](int t0, auto first_) __attribute__((always_inline)) {
                constexpr bool FIRST = decltype(first_)::value;
#pragma unroll
                for (int s = 0; s < P; ++s) {
                    const int t = t0 + s;
                    const int k = s % D;
                    // stage row t's channel values
#pragma unroll
                    for (int e = 0; e < NE; ++e) {
                        float cv[CH];
                        const float iv = img_u8 ? (float)qi[k][e] : __uint_as_float(qi[k][e]);
                        sweep1_channels<LIT>(iv, __uint_as_float(qh[k][e]), cv);
#pragma unroll
                        for (int c = 0; c < CH; ++c) sRow[wv][c][lane + 64 * e] = cv[c];
                    }
                    load_row(k, min(t + D, h - 1));
                    __builtin_amdgcn_wave_barrier();
                    // x pass of the lane's NCOL columns, each a sequential chain over
                    // the taps (g[0] x_0 first).  The window streams through in reads of
                    // VW elements, the next read in flight while the current one is
                    // used; scheduling barriers keep the compiler from hoisting all the
                    // reads (their registers would spill the ring)
                    float xo[NCOL];
                    const float* src = &sRow[wv][ch][cg * NCOL];
                    auto rd = [&](int q, float (&v)[VW]) {
                        if constexpr (VW == 4) {
                            const float4 f = *reinterpret_cast<const float4*>(src + 4 * q);
                            v[0] = f.x; v[1] = f.y; v[2] = f.z; v[3] = f.w;
                        } else {
                            const float2 f = *reinterpret_cast<const float2*>(src + 2 * q);
                            v[0] = f.x; v[1] = f.y;
                        }
                    };
                    float vc[VW], vn[VW];
                    rd(0, vc);
#pragma unroll
                    for (int q = 0; q < NV; ++q) {
                        if (q + 1 < NV) rd(q + 1, vn);
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int jj = 0; jj < VW; ++jj) {
                            const int el = VW * q + jj;
#pragma unroll
                            for (int o = 0; o < NCOL; ++o) {
                                const int kk = el - o;
                                if (kk == 0) xo[o] = vc[jj] * taps.g[0];
                                else if (kk > 0 && kk < P) xo[o] = fmaf(vc[jj], taps.g[kk], xo[o]);
                            }
                        }
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int jj = 0; jj < VW; ++jj) vc[jj] = vn[jj];
                    }
                    // y ring: accumulator s starts with its first product, the others add
                    // tap (s - j) mod P; accumulator (s + 1) mod P completes
#pragma unroll
                    for (int j = 0; j < P; ++j) {
#pragma unroll
                        for (int o = 0; o < NCOL; ++o)
                            acc[o][j] = j == s ? xo[o] * taps.g[0] : fmaf(xo[o], taps.g[(s - j + P) % P], acc[o][j]);
                    }
                    if (!FIRST || s >= P - 1) {
                        const int js = (s + 1) % P;
                        const uint32_t yo = (uint32_t)min(t - M, h - 1);
                        if constexpr (LIT) {
#pragma unroll
                            for (int o = 0; o < NCOL; ++o) sT[wv][tpos(cg * NCOL + o) + ch] = acc[o][js];
                            __builtin_amdgcn_wave_barrier();
                            const float4 vv = *reinterpret_cast<const float4*>(&sT[wv][tpos(min(lane, SW - 1))]);
                            const u32x4 b = {__float_as_uint(vv.x), __float_as_uint(vv.y), __float_as_uint(vv.z),
                                             __float_as_uint(vv.w)};
                            __builtin_amdgcn_raw_buffer_store_b128(b, rX, 16 * gxo, (int)(16u * yo * (uint32_t)w),
                                                                   kNT);
                        } else {
                            fvec v;
#pragma unroll
                            for (int o = 0; o < NCOL; ++o) v[o] = acc[o][js];
                            *reinterpret_cast<fvec*>(xy + ch * stride + zoff + (int64_t)yo * w + gxc) = v;
                        }
                    }
                }
            };
            for (int k = 0; k < D; ++k) load_row(k, min(k, h - 1));
            group(0, std::true_type{});
            for (int t0 = P; t0 < h; t0 += P) group(t0, std::false_type{});
        }
    }
}

// FAST xy kernel on the matrix cores.  The x pass of 16 rows x 16 outputs is
// the product of the 16 x 36 input window with the 36 x 16 band matrix
// T[m][j] = g[m - j] (0 <= m - j < P), nine v_mfma_f32_16x16x4_f32; the y pass
// of 16 outputs x 16 columns is T' (16 x 36) times 36 x-pass rows.  The sum
// over m is taken in a permuted order -- lane group k = lane/16 of slice s
// holds m = 9k + s -- so that a lane's nine window values lie next to each
// other in LDS, and the band operand of both passes is the same nine
// registers per lane.  The x-pass output goes to LDS column-major (a lane's
// four result rows are one 16-byte write, and the y pass reads nine
// consecutive rows).  The matrix cores round their sums in their own order:
// this is FAST mode's arithmetic, not the literal sequential correlation.
// Tiles: TX x 28 outputs, the window 48 rows = three 16-row blocks; the y
// pass covers the 28 rows with blocks at rows 0 and 12 (rows 12..15 from the
// first).
template <int P, int TX, int NT, int WPE, int IMG>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void sweep1_xy_mfma_kernel(
    const float* __restrict__ tr, const void* __restrict__ img, int img_u8_arg, float* __restrict__ xy,
    int64_t stride, const BlockDesc* __restrict__ blocks, const Tile* __restrict__ tiles, int ntiles, int zc,
    Taps taps) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    constexpr int TY = 28, M = (P - 1) / 2;
    // LDS strides = 2 (mod 32) words: the 16 row (or column) lanes of a read fall on
    // even banks, offset by 9k -- odd for k = 1, 3 -- so each 32-lane half is free
    // of bank conflicts (hence 8-byte LDS writes: rows are not 16-byte aligned)
    constexpr int RH = TY + P - 1, RW = TX + P - 1, RWP = ((RW + 29) / 32) * 32 + 2;
    constexpr int KW = 16 + P - 1, NS = KW / 4;      // window of 16 outputs, k slices
    constexpr int NRB = RH / 16, NCB = TX / 16;       // x-pass row blocks, column blocks
    constexpr int NBX = 2 * NRB * NCB, NBY = 2 * 2 * NCB;
    constexpr int NWV = NT / 64;
    constexpr int QW = RW / 4, NQ = RH * QW, PF = (NQ + NT - 1) / NT;
    constexpr int RXT = ((RH + 29) / 32) * 32 + 2;    // column of the x-pass output
    static_assert(KW % 4 == 0 && RH % 16 == 0 && TX % 16 == 0 && RW % 4 == 0 && NT % 64 == 0, "mfma tile");
    static_assert(TY - 16 + (NS - 1) + 9 * 3 < RH, "y pass reads inside the x-pass rows");
    __shared__ __align__(16) float sIn[2][RH][RWP];
    __shared__ __align__(16) float sXT[2][TX][RXT];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int img_u8 = IMG >= 0 ? IMG : img_u8_arg;
    const uint8_t* img8 = (const uint8_t*)img;
    const float* imgf = (const float*)img;
    // the band operand: slice s of lane (li, lk) is T[9 lk + s][li] (x pass, B) and
    // T'[li][9 lk + s] (y pass, A), both g[9 lk + s - li] or 0
    float tb[NS];
#pragma unroll
    for (int s2 = 0; s2 < NS; ++s2) {
        const int d = NS * lk + s2 - li;
        float v = 0.f;
#pragma unroll
        for (int k = 0; k < P; ++k) v = d == k ? taps.g[k] : v;
        tb[s2] = v;
    }
    for (int item = blockIdx.x; item < ntiles; item += gridDim.x) {
        const Tile tl = tiles[item];
        if (tl.b < 0) continue;
        const BlockDesc bd = blocks[tl.b];
        const int w = bd.w, h = bd.h, l = bd.l;
        const int64_t off = bd.off, plane = (int64_t)w * h;
        const int x0 = tl.x0, y0 = tl.y0, z0 = tl.z0;
        const int zend = min(z0 + zc, l);
        const bool vec = ((off | (int64_t)w) & 3) == 0 && ((uintptr_t)tr & 15) == 0 &&
                         (img_u8 ? ((uintptr_t)img8 & 3) == 0 : ((uintptr_t)imgf & 15) == 0);
        float4 pt[PF], pi[PF];
        auto fetch = [&](int z) {
            const int64_t zoff = off + (int64_t)z * plane;
#pragma unroll
            for (int k = 0; k < PF; ++k) {
                const int e = tid + k * NT;
                if (NQ % NT != 0 && e >= NQ) break;
                const int r = e / QW, q = e - r * QW;
                const int gy = min(y0 - M + r, h - 1);
                const int gx = x0 - M + 4 * q;
                const int64_t row = zoff + (int64_t)gy * w;
                if (vec && gx + 3 < w) {
                    pt[k] = *reinterpret_cast<const float4*>(tr + row + gx);
                    if (img_u8) pi[k].x = __uint_as_float(*reinterpret_cast<const uint32_t*>(img8 + row + gx));
                    else pi[k] = *reinterpret_cast<const float4*>(imgf + row + gx);
                } else {
                    float tv[4], iv[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int64_t idx = row + min(gx + j, w - 1);
                        tv[j] = tr[idx];
                        iv[j] = img_u8 ? (float)img8[idx] : imgf[idx];
                    }
                    pt[k] = make_float4(tv[0], tv[1], tv[2], tv[3]);
                    pi[k] = make_float4(iv[0], iv[1], iv[2], iv[3]);
                    if (img_u8) pi[k].x = __uint_as_float((uint32_t)iv[0] | ((uint32_t)iv[1] << 8) |
                                                          ((uint32_t)iv[2] << 16) | ((uint32_t)iv[3] << 24));
                }
            }
        };
        // channels I*G and G of the window into LDS
        auto store = [&]() {
#pragma unroll
            for (int k = 0; k < PF; ++k) {
                const int e = tid + k * NT;
                if (NQ % NT != 0 && e >= NQ) break;
                const int r = e / QW, q = e - r * QW;
                float iv[4];
                if (img_u8) {
                    const uint32_t b4 = __float_as_uint(pi[k].x);
                    iv[0] = (float)(b4 & 0xff);
                    iv[1] = (float)((b4 >> 8) & 0xff);
                    iv[2] = (float)((b4 >> 16) & 0xff);
                    iv[3] = (float)(b4 >> 24);
                } else {
                    iv[0] = pi[k].x; iv[1] = pi[k].y; iv[2] = pi[k].z; iv[3] = pi[k].w;
                }
                const float tv[4] = {pt[k].x, pt[k].y, pt[k].z, pt[k].w};
                float2* d0 = reinterpret_cast<float2*>(&sIn[0][r][4 * q]);
                float2* d1 = reinterpret_cast<float2*>(&sIn[1][r][4 * q]);
                d0[0] = make_float2(iv[0] * tv[0], iv[1] * tv[1]);
                d0[1] = make_float2(iv[2] * tv[2], iv[3] * tv[3]);
                d1[0] = make_float2(tv[0], tv[1]);
                d1[1] = make_float2(tv[2], tv[3]);
            }
        };
        fetch(z0);
        for (int z = z0; z < zend; ++z) {
            store();
            if (z + 1 < zend) fetch(z + 1);
            __syncthreads();
            // ---- x pass: block b = (channel, row block, column block), two at a time
#pragma unroll
            for (int it = 0; it < (NBX + 2 * NWV - 1) / (2 * NWV); ++it) {
                const int b = wv + it * 2 * NWV;
                if (b >= NBX) break;
                const int b2 = b + NWV;
                const bool two = b2 < NBX;
                const int ch = b / (NRB * NCB), rb = (b / NCB) % NRB, cb = b % NCB;
                const int ch2 = b2 / (NRB * NCB), rb2 = (b2 / NCB) % NRB, cb2 = b2 % NCB;
                const float* a1 = &sIn[ch][16 * rb + li][16 * cb + NS * lk];
                const float* a2 = &sIn[two ? ch2 : ch][16 * (two ? rb2 : rb) + li][16 * (two ? cb2 : cb) + NS * lk];
                f4 d1 = {0.f, 0.f, 0.f, 0.f}, d2 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int s2 = 0; s2 < NS; ++s2) {
                    d1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[s2], tb[s2], d1, 0, 0, 0);
                    d2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a2[s2], tb[s2], d2, 0, 0, 0);
                }
                float2* o1 = reinterpret_cast<float2*>(&sXT[ch][16 * cb + li][16 * rb + 4 * lk]);
                o1[0] = make_float2(d1[0], d1[1]);
                o1[1] = make_float2(d1[2], d1[3]);
                if (two) {
                    float2* o2 = reinterpret_cast<float2*>(&sXT[ch2][16 * cb2 + li][16 * rb2 + 4 * lk]);
                    o2[0] = make_float2(d2[0], d2[1]);
                    o2[1] = make_float2(d2[2], d2[3]);
                }
            }
            __syncthreads();
            // ---- y pass: block b = (channel, row block at 0 or TY-16, column block)
#pragma unroll
            for (int it = 0; it < (NBY + 2 * NWV - 1) / (2 * NWV); ++it) {
                const int b = wv + it * 2 * NWV;
                if (b >= NBY) break;
                const int b2 = b + NWV;
                const bool two = b2 < NBY;
                const int ch = b / (2 * NCB), yb = (b / NCB) & 1, cb = b % NCB;
                const int ch2 = b2 / (2 * NCB), yb2 = (b2 / NCB) & 1, cb2 = b2 % NCB;
                const int r01 = yb ? TY - 16 : 0, r02 = yb2 ? TY - 16 : 0;
                const float* s1 = &sXT[ch][16 * cb + li][r01 + NS * lk];
                const float* s2p = &sXT[two ? ch2 : ch][16 * (two ? cb2 : cb) + li][(two ? r02 : r01) + NS * lk];
                f4 d1 = {0.f, 0.f, 0.f, 0.f}, d2 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int s2 = 0; s2 < NS; ++s2) {
                    d1 = __builtin_amdgcn_mfma_f32_16x16x4f32(tb[s2], s1[s2], d1, 0, 0, 0);
                    d2 = __builtin_amdgcn_mfma_f32_16x16x4f32(tb[s2], s2p[s2], d2, 0, 0, 0);
                }
                auto put = [&](const f4& d, int c, int r0, int yb_, int cbb) {
                    const int gx = x0 + 16 * cbb + li;
                    float* dst = xy + c * stride + off + (int64_t)z * plane + gx;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int orow = r0 + 4 * lk + r;   // output row within the tile
                        const int gy = y0 + orow;
                        // the second block's rows below TY-16+4 repeat the first block's
                        if ((yb_ == 0 || 4 * lk + r >= 16 - r0) && gx < w - M && gy < h - M)
                            dst[(int64_t)gy * w] = d[r];
                    }
                };
                put(d1, ch, r01, yb, cb);
                if (two) put(d2, ch2, r02, yb2, cb2);
            }
        }
        __syncthreads();
    }
}

// z kernel: one thread per (x, y) column of a block, TXC x TYC consecutive
// columns per work item, marching all l planes.  Columns inside the valid xy region run the z
// ring over the xy-blurred channels; every voxel outside the valid region is
// in the pad band (fits 30/30 = 1).  U8: image dtype; MODE 0 writes F with
// lambda1 == lambda2, 3 F with separate lambdas, 1 the fits, 2 the fits and
// the blurred Heavisides, 4 the padded blur of channel 0 (the lambda form
// compiled in keeps the plane loop one basic block).  All accesses are buffer
// operations: the column's offset in a VGPR, the plane's offset in an SGPR
// (blocks stay below 2^32 bytes per buffer).
// IL (literal): the intermediate interleaved per voxel (xy[4v .. 4v+3]), one
// 16-byte load per plane; its buffer resource is rebased at the first plane of
// every group of P planes, so the 32-bit offsets span P + D planes (the host
// keeps 16 (P + D) w h below 2^32).
template <int P, bool LIT, int TXC, int TYC, int D, bool U8, int MODE, bool IL = false>
__global__ __launch_bounds__(TXC * TYC) __attribute__((amdgpu_waves_per_eu(4))) void sweep1_z_kernel(
    const float* __restrict__ xy, int64_t stride, const void* __restrict__ img,
    const float* __restrict__ blurI, float* __restrict__ out0, float* __restrict__ out1,
    float* __restrict__ out2, float* __restrict__ out3, const BlockDesc* __restrict__ blocks,
    const Tile* __restrict__ tiles, int ntiles, Taps taps, float halfS3, float l1, float l2, int same_lambda) {
    constexpr int CH = LIT ? 4 : 2;
    constexpr int CHL = MODE == 4 ? 1 : CH;   // channels read and blurred
    constexpr int M = (P - 1) / 2;
    constexpr int ESZ = U8 ? 1 : 4;
    static_assert(P % D == 0, "P must be a multiple of the load queue depth");
    static_assert(!IL || (LIT && MODE != 4), "the interleaved layout is literal");
    (void)same_lambda;   // the host picks MODE 0 or 3 from it
    for (int item = blockIdx.x; item < ntiles; item += gridDim.x) {
        const Tile tl = tiles[item];
        if (tl.b < 0) continue;
        const BlockDesc bd = blocks[tl.b];
        const int w = bd.w, h = bd.h, l = bd.l;
        // a work item is TXC*TYC consecutive columns of the block in (y, x) order
        // (whole waves stay busy on narrow blocks); no state crosses columns
        const int plane = w * h;
        const int c = tl.x0 + (int)threadIdx.x;
        if (c >= plane) continue;
        const int y = c / w, x = c - y * w;
        auto rsrc = [](const void* q) {
            return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(q), 0, (int)0xffffffff, 0x00020000);
        };
        __amdgpu_buffer_rsrc_t rC[CH];
#pragma unroll
        for (int ch = 0; ch < CH; ++ch) rC[ch] = rsrc(xy + ch * stride + bd.off);
        const __amdgpu_buffer_rsrc_t rL = rsrc(LIT ? out0 + bd.off : blurI + bd.off);
        const __amdgpu_buffer_rsrc_t rI = rsrc((const char*)img + bd.off * ESZ);
        const __amdgpu_buffer_rsrc_t r0 = rsrc(out0 + bd.off), r1 = rsrc(out1 + bd.off);
        const __amdgpu_buffer_rsrc_t r2 = rsrc(out2 + bd.off), r3 = rsrc(out3 + bd.off);
        const int vc = 4 * c, vci = ESZ * c;     // byte offsets of the column (VGPR)
        // plane byte offsets (SGPR) in unsigned 32-bit arithmetic: a block spans
        // fewer than 2^32 bytes per buffer, but 4 * z * plane can exceed 2^31
        auto so4 = [&](int z) { return (int)(4u * (uint32_t)z * (uint32_t)plane); };
        auto ldf = [&](__amdgpu_buffer_rsrc_t r, int z) {
            return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vc, so4(z), kNT));
        };
        auto ldi = [&](int z) {
            if constexpr (U8) return (float)__builtin_amdgcn_raw_buffer_load_b8(rI, vci, (int)((uint32_t)z * (uint32_t)plane), 0);
            else return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rI, vci, so4(z), 0));
        };
        auto st = [&](__amdgpu_buffer_rsrc_t r, int z, float v) {
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, vc, so4(z), kNT);
        };
        // interleaved layout: the four channels of plane (base + dz), r based at plane base
        auto ld4 = [&](float (&q)[CHL], __amdgpu_buffer_rsrc_t r, int dz) {
            const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)(16u * (uint32_t)c),
                                                                 (int)(16u * (uint32_t)dz * (uint32_t)plane), kNT);
            // each element through uint32_t (a bit_cast of the vector element itself
            // compiles to a one-dword load copied to every element)
            const uint32_t u[4] = {v[0], v[1], v[2], v[3]};
#pragma unroll
            for (int ch = 0; ch < CHL; ++ch) q[ch] = __builtin_bit_cast(float, u[ch]);
        };
        auto ilbase = [&](int z) { return rsrc(xy + 4 * (bd.off + (int64_t)z * plane)); };
        // ring outputs d[ch] (valid) or the pad band; bI: image blur (fast mode)
        auto store_out = [&](int z, float Iv, bool valid, float bI, const float (&d)[CH]) {
            if constexpr (MODE == 4) {
                // the padded blur of channel 0 itself (pad value 30, kernel.cu:313)
                st(r0, z, valid ? d[0] : 30.0f);
                return;
            }
            // the ring's fits, or the pad band's (kernel.cu:313, main.cpp:277), by
            // selection: a branch on the lane's validity inside the unrolled ring
            // would make the compiler copy the ring's registers around it
            float fo, fi, bH, bHc;
            if constexpr (LIT) {
                bH = d[1];
                bHc = d[3];
                fo = d[0] / bH;     // kernel.cu:203
                fi = d[2] / bHc;
            } else {
                fast_fits(bI, d[0], d[1], halfS3, fo, fi, bH, bHc);
            }
            fo = valid ? fo : 1.0f;
            fi = valid ? fi : 1.0f;
            bH = valid ? bH : 30.0f;
            bHc = valid ? bHc : 30.0f;
            if constexpr (MODE == 0 || MODE == 3) {
                st(r0, z, force_of(Iv, fo, fi, l1, l2, MODE == 0));
            } else {
                st(r0, z, fo);
                st(r1, z, fi);
                if constexpr (MODE == 2) {
                    st(r2, z, bH);
                    st(r3, z, bHc);
                }
            }
        };
        // pad-band planes [za, zb) of the column, in batches whose image loads are
        // issued before their stores (the compiler keeps a load after a store to
        // another buffer in order: plane by plane, every plane would wait a round trip)
        auto band = [&](int za, int zb) {
            constexpr int BB = 16;
            const float zero[CH] = {};
            for (int z = za; z < zb; z += BB) {
                float iv[BB];
#pragma unroll
                for (int k = 0; k < BB; ++k) iv[k] = MODE != 4 && z + k < zb ? ldi(z + k) : 0.f;
#pragma unroll
                for (int k = 0; k < BB; ++k)
                    if (z + k < zb) store_out(z + k, iv[k], false, 0.f, zero);
            }
        };
        // columns outside the valid xy region are pad band on every plane: a wave
        // of such columns only writes the band; a wave with valid columns takes its
        // band columns through the ring (their loads are in bounds, their sums unused)
        const bool xyvalid = x >= M && x < w - M && y >= M && y < h - M;
        if (__ballot(xyvalid) == 0) {
            band(0, l);
            continue;
        }
        band(0, M);
        float acc[CHL][P];
#pragma unroll
        for (int ch = 0; ch < CHL; ++ch)
#pragma unroll
            for (int j = 0; j < P; ++j) acc[ch][j] = 0.f;
        // D planes of loads in flight: the channels of plane t and the image (and,
        // fast mode, the image blur) of the plane it completes, o = t - (P-1) + M.
        // The plane loop is unrolled by P (a multiple of D): ring and queue slots
        // are fixed.
        float qc[D][CHL], qi[D], qB[D];
        if constexpr (MODE == 4)
#pragma unroll
            for (int k = 0; k < D; ++k) qi[k] = qB[k] = 0.f;
        auto group = [&](int t0, auto guard_) __attribute__((always_inline)) {
            constexpr bool GUARD = decltype(guard_)::value;
            __amdgpu_buffer_rsrc_t rG;
            if constexpr (IL) rG = ilbase(t0);
#pragma unroll
            for (int s = 0; s < P; ++s) {
                const int t = t0 + s;
                if (GUARD && t >= l) break;
                const int k = s % D;
                float v[CH];
#pragma unroll
                for (int ch = 0; ch < CHL; ++ch) v[ch] = qc[k][ch];
                const float vi = qi[k], vB = qB[k];
                const int tn = t + D;
                if (!GUARD || tn < l) {
                    if constexpr (IL) {
                        ld4(qc[k], rG, s + D);
                    } else {
#pragma unroll
                        for (int ch = 0; ch < CHL; ++ch) qc[k][ch] = ldf(rC[ch], tn);
                    }
                    if (!GUARD || tn >= P - 1) {
                        // blur mode (4) reads neither the image nor an image blur
                        if constexpr (MODE != 4) qi[k] = ldi(tn - (P - 1) + M);
                        if constexpr (!LIT && MODE != 4) qB[k] = ldf(rL, tn - (P - 1) + M);
                    }
                }
                // plane t adds tap g[(s - j) mod P] to accumulator j and completes
                // accumulator (s + 1) mod P (see the fused kernel's ring)
#pragma unroll
                for (int j = 0; j < P; ++j) {
                    const float g = taps.g[(s - j + P) % P];
#pragma unroll
                    for (int ch = 0; ch < CHL; ++ch) acc[ch][j] = fmaf(v[ch], g, acc[ch][j]);
                }
                const int js = (s + 1) % P;
                float d[CH] = {};
#pragma unroll
                for (int ch = 0; ch < CHL; ++ch) {
                    d[ch] = acc[ch][js];
                    acc[ch][js] = 0.f;
                }
                if (!GUARD || t >= P - 1) store_out(t - (P - 1) + M, vi, xyvalid, vB, d);
            }
        };
        // queue fill: planes 0 .. D-1 (their completed planes exist only from P-1 on)
#pragma unroll
        for (int k = 0; k < D; ++k) {
            if (k < l) {
                if constexpr (IL) {
                    ld4(qc[k], ilbase(0), k);
                } else {
#pragma unroll
                    for (int ch = 0; ch < CHL; ++ch) qc[k][ch] = ldf(rC[ch], k);
                }
            }
        }
        group(0, std::true_type{});
        int t0 = P;
        for (; t0 + P + D <= l; t0 += P) group(t0, std::false_type{});
        for (; t0 < l; t0 += P) group(t0, std::true_type{});
        band(l - M, l);
    }
}

// Sweep 2 kernels: no implicit contraction.  Each kernel variant (and each
// code path inside one, e.g. a chunk's first plane) then rounds every voxel's
// update the same way; the fused multiply-adds they want are written out
// (mag2, half_sum3, phi_update, dirac_fast).
#pragma clang fp contract(off)

// One voxel of sweep 2 with the full reference edge rules (boundary shell).
// The shell runs beside interior kernels that address a block through 32-bit
// buffer offsets, so its blocks hold fewer than 2^30 voxels: the indices are
// 32-bit and the loads go through buffer resources.  Literal normals take the
// shortened forms of the interior kernels and, for a |grad|^2 outside their
// range, the exact ones (the same bits either way).
template <bool LIT>
__device__ __forceinline__ void sweep2_voxel(const float* __restrict__ f, const float* __restrict__ F,
                                             float* __restrict__ phi_out, float* __restrict__ tr_out, int x,
                                             int y, int z, int w, int h, int l, float dt, float nu, float mu,
                                             float eps, float inv_eps) {
    const int P = w * h;
    const int i = (z * h + y) * w + x;
    const __amdgpu_buffer_rsrc_t rf = __builtin_amdgcn_make_buffer_rsrc((void*)f, 0, (int)0xffffffff, 0x00020000);
    auto ld = [&](int k) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rf, (int)(4u * (uint32_t)k), 0, 0));
    };
    // D f at point q (linear index qi) with the edge rules of kernel.cu:57-125
    auto grad = [&](int qx, int qy, int qz, int qi) {
        const int xr = qx == w - 1 ? qi : qi + 1, xl = qi == 0 ? 0 : qi - 1;
        const int yr = qy == h - 1 ? qi : qi + w, yl = qy == 0 ? qi : qi - w;
        const int zr = qz == l - 1 ? qi : qi + P, zl = qz == 0 ? qi : qi - P;
        Grad g;
        g.x = (ld(xr) - ld(xl)) * 0.5f;
        g.y = (ld(yr) - ld(yl)) * 0.5f;
        g.z = (ld(zr) - ld(zl)) * 0.5f;
        return g;
    };
    int qlx, qly, qlz;
    left_x(x, y, z, w, h, qlx, qly, qlz);
    const int rxx = (x == w - 1) ? x : x + 1;
    const int ryy = (y == 0) ? 1 : (y == h - 1 ? y : y + 1);
    const int lyy = (y == 0) ? 0 : (y == h - 1 ? h - 2 : y - 1);
    const int rzz = (z == 0) ? 1 : (z == l - 1 ? z : z + 1);
    const int lzz = (z == 0) ? 0 : (z == l - 1 ? l - 2 : z - 1);
    const int row = i - x;
    const Grad gRx = grad(rxx, y, z, row + rxx);
    const Grad gLx = grad(qlx, qly, qlz, (qlz * h + qly) * w + qlx);
    const Grad gRy = grad(x, ryy, z, i + (ryy - y) * w);
    const Grad gLy = grad(x, lyy, z, i + (lyy - y) * w);
    const Grad gRz = grad(x, y, rzz, i + (rzz - z) * P);
    const Grad gLz = grad(x, y, lzz, i + (lzz - z) * P);
    float nRx, nLx, nRy, nLy, nRz, nLz;
    if constexpr (LIT) {
        auto nrm = [](const Grad& g, float c) {
            const float m2 = mag2<LIT>(g.x, g.y, g.z);
            float n[1];
            if (normal_outside(m2)) normal_literal_exact<1>({c}, m2, n);
            else normal_literal_fast<1>({c}, m2, n);
            return n[0];
        };
        nRx = nrm(gRx, gRx.x); nLx = nrm(gLx, gLx.x);
        nRy = nrm(gRy, gRy.y); nLy = nrm(gLy, gLy.y);
        nRz = nrm(gRz, gRz.z); nLz = nrm(gLz, gLz.z);
    } else {
        auto nrm = [](const Grad& g, float c) {
            const float M = __builtin_amdgcn_sqrtf(mag2<false>(g.x, g.y, g.z));
            return c * __builtin_amdgcn_rcpf(M + 1e-6f);
        };
        nRx = nrm(gRx, gRx.x); nLx = nrm(gLx, gLx.x);
        nRy = nrm(gRy, gRy.y); nLy = nrm(gLy, gLy.y);
        nRz = nrm(gRz, gRz.z); nLz = nrm(gLz, gLz.z);
    }
    const float kap = half_sum3(nRx - nLx, nRy - nLy, nRz - nLz);
    const float lap = half_sum3(gRx.x - gLx.x, gRy.y - gLy.y, gRz.z - gLz.z);
    const float p0 = ld(i);
    const float d = LIT ? dirac_literal(p0, eps) : dirac_fast(p0, eps);
    const float p1 = phi_update<LIT>(p0, d, F[i], kap, lap, dt, nu * dt, mu);
    phi_out[i] = p1;
    tr_out[i] = LIT ? heav_literal_r(p1, eps, inv_eps) : heav_centred(p1, inv_eps);
}

// Sweep 2, boundary shell: voxels within 2 of a face of their block, and,
// for blocks whose interior runs on the vector kernel, the columns that kernel
// leaves out (x = 2, 3 and x >= 4 + 4*((w-8)/4)); the shared-interior kernel
// covers whole rows, and its shell is the planes and rows only.  Work items (block, first
// shell index, count, vector flag) enumerate, per block, the 4 planes z in
// {0,1,l-2,l-1}, then the rows y in {0,1,h-2,h-1} of the inner planes, then the
// boundary columns of the inner rows.
template <bool LIT>
__global__ __launch_bounds__(256) void sweep2_shell_kernel(
    const float* __restrict__ phi_all, const float* __restrict__ F, float* __restrict__ phi_out,
    float* __restrict__ tr_out, const BlockDesc* __restrict__ blocks, const Tile* __restrict__ items,
    int nitems, float dt, float nu, float mu, float eps, float inv_eps) {
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const Tile t = items[it];
        if (t.b < 0 || (int)threadIdx.x >= t.y0) continue;
        const BlockDesc bd = blocks[t.b];
        const int w = bd.w, h = bd.h, l = bd.l;
        // boundary columns of the inner rows: x < cl and x >= w - cr (z0: 0 scalar
        // interior, 1 vector interior, 2 an interior over whole rows: no columns)
        const int cl = t.z0 == 2 ? 0 : (t.z0 ? 4 : 2);
        const int cr = t.z0 == 2 ? 0 : (t.z0 ? w - 4 - 4 * ((w - 8) >> 2) : 2), cols = cl + cr;
        int64_t k = (int64_t)t.x0 + threadIdx.x;
        const int64_t cA = 4 * (int64_t)w * h, cB = 4 * (int64_t)w * (l - 4);
        int x, y, z;
        if (k < cA) {
            const int pz = (int)(k / ((int64_t)w * h));
            const int64_t r = k - (int64_t)pz * w * h;
            z = pz < 2 ? pz : l - 4 + pz;
            y = (int)(r / w);
            x = (int)(r - (int64_t)y * w);
        } else if (k < cA + cB) {
            k -= cA;
            z = 2 + (int)(k / (4 * (int64_t)w));
            const int r = (int)(k - (int64_t)(z - 2) * 4 * w);
            const int yy = r / w;
            y = yy < 2 ? yy : h - 4 + yy;
            x = r - yy * w;
        } else {
            k -= cA + cB;
            const int64_t per = (int64_t)cols * (h - 4);
            z = 2 + (int)(k / per);
            const int r = (int)(k - (int64_t)(z - 2) * per);
            y = 2 + r / cols;
            const int xx = r % cols;
            x = xx < cl ? xx : w - cols + xx;
        }
        sweep2_voxel<LIT>(phi_all + bd.off, F + bd.off, phi_out + bd.off, tr_out + bd.off, x, y, z, w, h, l, dt, nu,
                          mu, eps, inv_eps);
    }
}

// Sweep 2, interior: tiles cover x in [2,w-2), y in [2,h-2), z in [2,l-2) of
// each block; each thread marches one column through z.  The radius-2 stencil
// is read through a buffer resource whose base sits 2P+2W+2 elements before
// the block, so every tap has a non-negative byte offset: voffset (4i) + an
// SGPR for the y/z part + an immediate for the x part.
template <bool LIT, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void sweep2_inner_kernel(
    const float* __restrict__ phi_all, const float* __restrict__ F, float* __restrict__ phi_out,
    float* __restrict__ tr_out, const BlockDesc* __restrict__ blocks, const Tile* __restrict__ tiles,
    int ntiles, int zc, float dt, float nu, float mu, float eps, float inv_eps) {
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    for (int item = blockIdx.x; item < ntiles; item += gridDim.x) {
        const Tile tl = tiles[item];
        if (tl.b < 0) continue;
        const BlockDesc bd = blocks[tl.b];
        const int w = bd.w, h = bd.h, l = bd.l;
        const int x = tl.x0 + tx, y = tl.y0 + ty;
        if (x >= w - 2 || y >= h - 2) continue;
        const int64_t W = w, P = (int64_t)w * h;
        const int64_t shift = 2 * P + 2 * W + 2;
        const float* f = phi_all + bd.off;
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc((void*)(f - shift), 0, (int)0xffffffff, 0x00020000);
        const __amdgpu_buffer_rsrc_t rF =
            __builtin_amdgcn_make_buffer_rsrc((void*)(F + bd.off), 0, (int)0xffffffff, 0x00020000);
        const __amdgpu_buffer_rsrc_t rO =
            __builtin_amdgcn_make_buffer_rsrc((void*)(phi_out + bd.off), 0, (int)0xffffffff, 0x00020000);
        const __amdgpu_buffer_rsrc_t rT =
            __builtin_amdgcn_make_buffer_rsrc((void*)(tr_out + bd.off), 0, (int)0xffffffff, 0x00020000);
        auto S = [&](int ky, int kz) { return (int)(uint32_t)(((2 + ky) * W + (2 + kz) * P) * 4); };
        const int s00 = S(0, 0), sp10 = S(1, 0), sm10 = S(-1, 0), sp20 = S(2, 0), sm20 = S(-2, 0);
        const int s0p1 = S(0, 1), s0m1 = S(0, -1), s0p2 = S(0, 2), s0m2 = S(0, -2);
        const int spp = S(1, 1), spm = S(1, -1), smp = S(-1, 1), smm = S(-1, -1);
        const int zend = min(tl.z0 + zc, l - 2);
        const uint32_t vstep = (uint32_t)(P * 4);
        // taps: 0 c, 1 x+2, 2 x-2, 3 y+2, 4 y-2, 5 z+2, 6 z-2, 7 x+y+, 8 x+y-, 9 x-y+, 10 x-y-,
        // 11 x+z+, 12 x+z-, 13 x-z+, 14 x-z-, 15 y+z+, 16 y+z-, 17 y-z+, 18 y-z-, 19 F
        auto load = [&](float (&r)[20], uint32_t v) {
            auto L = [&](int kx, int so) {
                return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)(v + 4 * (kx + 2)), so, 0));
            };
            r[0] = L(0, s00);
            r[1] = L(2, s00); r[2] = L(-2, s00);
            r[3] = L(0, sp20); r[4] = L(0, sm20);
            r[5] = L(0, s0p2); r[6] = L(0, s0m2);
            r[7] = L(1, sp10); r[8] = L(1, sm10); r[9] = L(-1, sp10); r[10] = L(-1, sm10);
            r[11] = L(1, s0p1); r[12] = L(1, s0m1); r[13] = L(-1, s0p1); r[14] = L(-1, s0m1);
            r[15] = L(0, spp); r[16] = L(0, spm); r[17] = L(0, smp); r[18] = L(0, smm);
            r[19] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rF, (int)v, 0, kNT));
        };
        auto compute = [&](const float (&r)[20], uint32_t v) {
            const float c0 = r[0];
            // gradients at the six D-neighbours (interior rules)
            const Grad gRx{(r[1] - c0) * 0.5f, (r[7] - r[8]) * 0.5f, (r[11] - r[12]) * 0.5f};
            const Grad gLx{(c0 - r[2]) * 0.5f, (r[9] - r[10]) * 0.5f, (r[13] - r[14]) * 0.5f};
            const Grad gRy{(r[7] - r[9]) * 0.5f, (r[3] - c0) * 0.5f, (r[15] - r[16]) * 0.5f};
            const Grad gLy{(r[8] - r[10]) * 0.5f, (c0 - r[4]) * 0.5f, (r[17] - r[18]) * 0.5f};
            const Grad gRz{(r[11] - r[13]) * 0.5f, (r[15] - r[17]) * 0.5f, (r[5] - c0) * 0.5f};
            const Grad gLz{(r[12] - r[14]) * 0.5f, (r[16] - r[18]) * 0.5f, (c0 - r[6]) * 0.5f};
            float nRx, nLx, nRy, nLy, nRz, nLz;
            if constexpr (LIT) {
                auto nrm = [](const Grad& g, float cc) {
                    const float M = sqrtf(mag2<LIT>(g.x, g.y, g.z));
                    return (float)((double)cc / ((double)M + 1e-6));
                };
                nRx = nrm(gRx, gRx.x); nLx = nrm(gLx, gLx.x);
                nRy = nrm(gRy, gRy.y); nLy = nrm(gLy, gLy.y);
                nRz = nrm(gRz, gRz.z); nLz = nrm(gLz, gLz.z);
            } else {
                auto nrm = [](const Grad& g, float cc) {
                    const float M = __builtin_amdgcn_sqrtf(mag2<false>(g.x, g.y, g.z));
                    return cc * __builtin_amdgcn_rcpf(M + 1e-6f);
                };
                nRx = nrm(gRx, gRx.x); nLx = nrm(gLx, gLx.x);
                nRy = nrm(gRy, gRy.y); nLy = nrm(gLy, gLy.y);
                nRz = nrm(gRz, gRz.z); nLz = nrm(gLz, gLz.z);
            }
            const float kap = half_sum3(nRx - nLx, nRy - nLy, nRz - nLz);
            const float lap = half_sum3(gRx.x - gLx.x, gRy.y - gLy.y, gRz.z - gLz.z);
            const float d = LIT ? dirac_literal(c0, eps) : dirac_fast(c0, eps);
            const float p1 = phi_update<LIT>(c0, d, r[19], kap, lap, dt, nu * dt, mu);
            const float t1 = LIT ? heav_literal_r(p1, eps, inv_eps) : heav_centred(p1, inv_eps);
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, p1), rO, (int)v, 0, kNT);
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, t1), rT, (int)v, 0, kNT);
        };
        // two register sets: the next plane's loads are in flight while this one computes
        float A[20], B[20];
        int z = tl.z0;
        uint32_t v = (uint32_t)((((int64_t)tl.z0 * h + y) * W + x) * 4);
        if (z < zend) load(A, v);
        while (z < zend) {
            if (z + 1 < zend) load(B, v + vstep);
            compute(A, v);
            ++z;
            v += vstep;
            if (z >= zend) break;
            if (z + 1 < zend) load(A, v + vstep);
            compute(B, v);
            ++z;
            v += vstep;
        }
    }
}

// Sweep 2, vector interior.  Each thread owns 4 consecutive x voxels
// (x0..x0+3, x0 = 4 + 4q) of one row y and marches them through z.  Rows are
// read with 8- and 16-byte buffer loads (rows of a block must start on a
// 16-byte boundary: offset and w multiples of 4), and a 4-deep register ring
// along z keeps
//   C: phi(y, z-1 .. z+2) at x0-2 .. x0+5,
//   U, D: phi(y+1, .), phi(y-1, .) at x0-1 .. x0+4 for z-1 .. z+1,
//   the gradient at (x0..x0+3, y, z') computed once, at step z'-1 (as the
//   z+1 neighbour), and used again at z' (as x neighbours) and z'+1 (as the
//   z-1 neighbour).
// Per step a thread loads one new C row, one new U and D row, the rows
// y+-2 and F (12 loads for 4 voxels) and evaluates 14 gradients (3.5 per
// voxel).  Interior rules only: x in [4, 4+4*nq), y and z in [2, n-2).

// Normal components at one point.  Literal path: EX selects the exact forms;
// otherwise the fast forms run and `bad` records whether this lane's inputs
// lie outside their range (the caller then recomputes with EX).
template <bool LIT, bool EX>
__device__ __forceinline__ float norm_comp(float c, float gx, float gy, float gz, bool& bad) {
    if constexpr (LIT) {
        const float m2 = mag2<true>(gx, gy, gz);
        float n[1];
        if constexpr (EX) {
            normal_literal_exact<1>({c}, m2, n);
        } else {
            bad |= normal_outside(m2);
            normal_literal_fast<1>({c}, m2, n);
        }
        return n[0];
    } else {
        const float M = __builtin_amdgcn_sqrtf(mag2<false>(gx, gy, gz));
        return c * __builtin_amdgcn_rcpf(M + 1e-6f);
    }
}
template <bool LIT, bool EX>
__device__ __forceinline__ void norm_comp2(float c1, float c2, float gx, float gy, float gz, float& n1, float& n2,
                                           bool& bad) {
    if constexpr (LIT) {
        const float m2 = mag2<true>(gx, gy, gz);
        float n[2];
        if constexpr (EX) {
            normal_literal_exact<2>({c1, c2}, m2, n);
        } else {
            bad |= normal_outside(m2);
            normal_literal_fast<2>({c1, c2}, m2, n);
        }
        n1 = n[0];
        n2 = n[1];
    } else {
        n1 = norm_comp<false, false>(c1, gx, gy, gz, bad);
        n2 = norm_comp<false, false>(c2, gx, gy, gz, bad);
    }
}

// FIX = 0: literal mode computes the normals with the fast forms
// (normal_literal_fast) and records, per work item and wave, whether any lane
// met a |grad|^2 outside their range (flags, NW = TX*TY/64 ints per item,
// written every launch); fast mode ignores flags.  FIX = 1 (literal): the
// exact forms, only for the waves flagged -- the fix-up launch after FIX = 0,
// which rewrites those waves' outputs (sweep 2 reads phi and writes phi_out,
// so a wave's part of an item can be computed again).
template <bool LIT, int TX, int TY, int WPE, int FIX = 0>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE))) void sweep2_vec_kernel(
    const float* __restrict__ phi_all, const float* __restrict__ F, float* __restrict__ phi_out,
    float* __restrict__ tr_out, const BlockDesc* __restrict__ blocks, const Tile* __restrict__ tiles,
    int ntiles, int zc, float dt, float nu, float mu, float eps, float inv_eps, int* __restrict__ flags) {
    constexpr int NW = TX * TY / 64;
    static_assert(TX * TY % 64 == 0, "whole waves");
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    for (int item = blockIdx.x; item < ntiles; item += gridDim.x) {
        const Tile tl = tiles[item];
        if (tl.b < 0) continue;
        if constexpr (LIT && FIX) {
            if (!flags[item * NW + threadIdx.x / 64]) continue;   // this wave's part only
        }
        const BlockDesc bd = blocks[tl.b];
        const int w = bd.w, h = bd.h, l = bd.l;
        const int xend = 4 + 4 * ((w - 8) >> 2);
        const int x0 = tl.x0 + 4 * tx, y = tl.y0 + ty;
        bool bad = false;   // literal, FIX = 0: some input of this lane outside the fast forms' range
        if (x0 < xend && y < h - 2) {
            const int zb = tl.z0, zend = min(tl.z0 + zc, l - 2);
            const int W = w, P = w * h;
            const int shift = 2 * P + 2 * W + 4;
            const float* f = phi_all + bd.off;
            const __amdgpu_buffer_rsrc_t rs =
                __builtin_amdgcn_make_buffer_rsrc((void*)(f - shift), 0, (int)0xffffffff, 0x00020000);
            const __amdgpu_buffer_rsrc_t rF =
                __builtin_amdgcn_make_buffer_rsrc((void*)(F + bd.off), 0, (int)0xffffffff, 0x00020000);
            const __amdgpu_buffer_rsrc_t rO =
                __builtin_amdgcn_make_buffer_rsrc((void*)(phi_out + bd.off), 0, (int)0xffffffff, 0x00020000);
            const __amdgpu_buffer_rsrc_t rT =
                __builtin_amdgcn_make_buffer_rsrc((void*)(tr_out + bd.off), 0, (int)0xffffffff, 0x00020000);
            // byte offset of row (y+dy, z+dz) minus 4 elements: voffset + 0..32 bytes then
            // addresses x0-4 .. x0+4, and every sum stays non-negative
            auto S = [&](int dy, int dz) { return (int)(4u * (uint32_t)(shift + dz * P + dy * W - 4)); };
            const int sC = S(0, 0), sU = S(1, 0), sD = S(-1, 0), sU2 = S(2, 0), sD2 = S(-2, 0);
            const uint32_t vstep = 4u * (uint32_t)P;
            auto u2f = [](uint32_t u) { return __builtin_bit_cast(float, u); };
            // centre-type row (x0-2 .. x0+5) at voxel byte offset v + row (dy, dz)
            auto load8 = [&](float (&r)[8], uint32_t v, int so) {
                const u32x2 a = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)(v + 8), so, 0);
                const u32x4 b = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(v + 16), so, 0);
                const u32x2 c = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)(v + 32), so, 0);
                r[0] = u2f(a[0]); r[1] = u2f(a[1]);
                r[2] = u2f(b[0]); r[3] = u2f(b[1]); r[4] = u2f(b[2]); r[5] = u2f(b[3]);
                r[6] = u2f(c[0]); r[7] = u2f(c[1]);
            };
            // neighbour-type row (x0-1 .. x0+4)
            auto load6 = [&](float (&r)[6], uint32_t v, int so) {
                r[0] = u2f(__builtin_amdgcn_raw_buffer_load_b32(rs, (int)(v + 12), so, 0));
                const u32x4 b = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(v + 16), so, 0);
                r[1] = u2f(b[0]); r[2] = u2f(b[1]); r[3] = u2f(b[2]); r[4] = u2f(b[3]);
                r[5] = u2f(__builtin_amdgcn_raw_buffer_load_b32(rs, (int)(v + 32), so, 0));
            };
            auto load4 = [&](float (&r)[4], uint32_t v, int so) {
                const u32x4 b = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(v + 16), so, 0);
                r[0] = u2f(b[0]); r[1] = u2f(b[1]); r[2] = u2f(b[2]); r[3] = u2f(b[3]);
            };
            auto loadF = [&](float (&r)[4], uint32_t v) {
                const u32x4 b = __builtin_amdgcn_raw_buffer_load_b128(rF, (int)v, 0, kNT);
                r[0] = u2f(b[0]); r[1] = u2f(b[1]); r[2] = u2f(b[2]); r[3] = u2f(b[3]);
            };
            float C[4][8], U[4][6], D[4][6];          // phi rows, ring slot = (z' - zb) & 3
            float Q2u[2][4], Q2d[2][4], FF[2][4];     // rows y+-2 and F of the current / next plane
            float NZ[4][4], GZ[4][4], NX[4][4], GX[4][4];  // gradient at (x0+i, y, z')
            const uint32_t v0 = 4u * (uint32_t)(zb * P + y * W + x0);
            // the gradient at (x0+i, y, z') from the rows of planes z'-1, z', z'+1
            auto zpoint = [&](auto EX_, bool& bad, int i, const float (&cm)[8], const float (&c0)[8],
                              const float (&cp)[8], const float (&u0)[6], const float (&d0)[6], float& nx, float& gx,
                              float& nz, float& gz) __attribute__((always_inline)) {
                gx = (c0[i + 3] - c0[i + 1]) * 0.5f;
                const float gy = (u0[i + 1] - d0[i + 1]) * 0.5f;
                gz = (cp[i + 2] - cm[i + 2]) * 0.5f;
                norm_comp2<LIT, decltype(EX_)::value>(gx, gz, gx, gy, gz, nx, nz, bad);
            };
            // ---- prologue: rows of planes zb-2 .. zb+2, gradients at zb-1 and zb
            {
                float Cmm[8];
                load8(Cmm, v0, S(0, -2));
                load8(C[3], v0, S(0, -1));
                load8(C[0], v0, sC);
                load8(C[1], v0, S(0, 1));
                load8(C[2], v0, S(0, 2));
                load6(U[3], v0, S(1, -1));
                load6(D[3], v0, S(-1, -1));
                load6(U[0], v0, sU);
                load6(D[0], v0, sD);
                load6(U[1], v0, S(1, 1));
                load6(D[1], v0, S(-1, 1));
                load4(Q2u[0], v0, sU2);
                load4(Q2d[0], v0, sD2);
                loadF(FF[0], v0);
                bool unused = false;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    float nx, gx;
                    zpoint(std::true_type{}, unused, i, Cmm, C[3], C[0], U[3], D[3], nx, gx, NZ[3][i], GZ[3][i]);
                    zpoint(std::true_type{}, unused, i, C[3], C[0], C[1], U[0], D[0], NX[0][i], GX[0][i], NZ[0][i],
                           GZ[0][i]);
                }
            }
            auto step = [&](auto S_, uint32_t v, int z) __attribute__((always_inline)) {
                constexpr int s = decltype(S_)::value;
                constexpr int km1 = (s + 3) & 3, k0 = s & 3, kp1 = (s + 1) & 3, kp2 = (s + 2) & 3;
                constexpr int q = s & 1, qn = (s + 1) & 1;
                float (&Cm)[8] = C[km1];
                float (&C0)[8] = C[k0];
                float (&Cp)[8] = C[kp1];
                float (&Cpp)[8] = C[kp2];
                // the two values of plane z-1 still needed, then that slot takes plane z+3
                const float cmL = Cm[1], cmR = Cm[6];
                if (z + 1 < zend) {
                    load8(C[km1], v + vstep, S(0, 2));
                    load6(U[kp2], v + vstep, S(1, 1));
                    load6(D[kp2], v + vstep, S(-1, 1));
                    load4(Q2u[qn], v + vstep, sU2);
                    load4(Q2d[qn], v + vstep, sD2);
                    loadF(FF[qn], v + vstep);
                }
                // The step's arithmetic: the exact normal forms in the fix-up launch,
                // else the fast ones, which flag a tiny or non-finite |grad|^2.
                float p1[4], t1[4];
                auto body = [&](auto EX_, bool& bad) __attribute__((always_inline)) {
                    constexpr bool EX = decltype(EX_)::value;
                    // gradients at (x0+i, y, z+1): the z+1 neighbours now, x neighbours next step
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        zpoint(EX_, bad, i, C0, Cp, Cpp, U[kp1], D[kp1], NX[kp1][i], GX[kp1][i], NZ[kp1][i],
                               GZ[kp1][i]);
                        if constexpr (LIT) __builtin_amdgcn_sched_barrier(0);   // one point's temporaries at a time
                    }
                    // x neighbours at the quad's ends: (x0-1, y, z) and (x0+4, y, z)
                    float nxL, gxL, nxR, gxR;
                    {
                        gxL = (C0[2] - C0[0]) * 0.5f;
                        const float gy = (U[k0][0] - D[k0][0]) * 0.5f, gz = (Cp[1] - cmL) * 0.5f;
                        nxL = norm_comp<LIT, EX>(gxL, gxL, gy, gz, bad);
                    }
                    if constexpr (LIT) __builtin_amdgcn_sched_barrier(0);
                    {
                        gxR = (C0[7] - C0[5]) * 0.5f;
                        const float gy = (U[k0][5] - D[k0][5]) * 0.5f, gz = (Cp[6] - cmR) * 0.5f;
                        nxR = norm_comp<LIT, EX>(gxR, gxR, gy, gz, bad);
                    }
                    const float nxs[6] = {nxL, NX[k0][0], NX[k0][1], NX[k0][2], NX[k0][3], nxR};
                    const float gxs[6] = {gxL, GX[k0][0], GX[k0][1], GX[k0][2], GX[k0][3], gxR};
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        // y neighbours (x0+i, y+-1, z)
                        const float gxu = (U[k0][i + 2] - U[k0][i]) * 0.5f;
                        const float gyu = (Q2u[q][i] - C0[i + 2]) * 0.5f;
                        const float gzu = (U[kp1][i + 1] - U[km1][i + 1]) * 0.5f;
                        const float nyu = norm_comp<LIT, EX>(gyu, gxu, gyu, gzu, bad);
                        const float gxd = (D[k0][i + 2] - D[k0][i]) * 0.5f;
                        const float gyd = (C0[i + 2] - Q2d[q][i]) * 0.5f;
                        const float gzd = (D[kp1][i + 1] - D[km1][i + 1]) * 0.5f;
                        const float nyd = norm_comp<LIT, EX>(gyd, gxd, gyd, gzd, bad);
                        const float kap = half_sum3(nxs[i + 2] - nxs[i], nyu - nyd, NZ[kp1][i] - NZ[km1][i]);
                        const float lap = half_sum3(gxs[i + 2] - gxs[i], gyu - gyd, GZ[kp1][i] - GZ[km1][i]);
                        const float c0 = C0[i + 2];
                        const float d = LIT ? dirac_literal(c0, eps) : dirac_fast(c0, eps);
                        p1[i] = phi_update<LIT>(c0, d, FF[q][i], kap, lap, dt, nu * dt, mu);
                        t1[i] = LIT ? heav_literal_r(p1[i], eps, inv_eps) : heav_centred(p1[i], inv_eps);
                        if constexpr (LIT) __builtin_amdgcn_sched_barrier(0);
                    }
                };
                body(std::integral_constant<bool, FIX != 0>{}, bad);
                u32x4 po, pt;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    po[i] = __builtin_bit_cast(uint32_t, p1[i]);
                    pt[i] = __builtin_bit_cast(uint32_t, t1[i]);
                }
                __builtin_amdgcn_raw_buffer_store_b128(po, rO, (int)v, 0, kNT);
                __builtin_amdgcn_raw_buffer_store_b128(pt, rT, (int)v, 0, kNT);
            };
            uint32_t v = v0;
            for (int z = zb; z < zend; z += 4) {
                step(std::integral_constant<int, 0>{}, v, z);
                if (z + 1 >= zend) break;
                step(std::integral_constant<int, 1>{}, v + vstep, z + 1);
                if (z + 2 >= zend) break;
                step(std::integral_constant<int, 2>{}, v + 2 * vstep, z + 2);
                if (z + 3 >= zend) break;
                step(std::integral_constant<int, 3>{}, v + 3 * vstep, z + 3);
                v += 4 * vstep;
            }
        }
        if constexpr (LIT && !FIX) {
            const bool wbad = __builtin_amdgcn_ballot_w64(bad) != 0;
            if (threadIdx.x % 64 == 0) flags[item * NW + threadIdx.x / 64] = wbad ? 1 : 0;
        }
    }
}

// All three normal components at one point (the shared-interior kernel
// hands each of them to a different neighbour).
template <bool LIT, bool EX>
__device__ __forceinline__ void norm_comp3(float gx, float gy, float gz, float& nx, float& ny, float& nz, bool& bad) {
    if constexpr (LIT) {
        const float m2 = mag2<true>(gx, gy, gz);
        float n[3];
        if constexpr (EX) {
            normal_literal_exact<3>({gx, gy, gz}, m2, n);
        } else {
            bad |= normal_outside(m2);
            normal_literal_fast<3>({gx, gy, gz}, m2, n);
        }
        nx = n[0];
        ny = n[1];
        nz = n[2];
    } else {
        nx = norm_comp<false, false>(gx, gx, gy, gz, bad);
        ny = norm_comp<false, false>(gy, gx, gy, gz, bad);
        nz = norm_comp<false, false>(gz, gx, gy, gz, bad);
    }
}

// Sweep 2, shared interior.  The quads and z-march of the vector kernel, but
// every point's gradient and normal is computed once, by the thread that owns
// the point, and handed to the neighbours that need it through LDS: (n_y, g_y)
// of the rows y+-1 and (n_x, g_x) of the quad ends x0-1 and x0+4 (double-
// buffered by plane parity, one barrier per plane).  The rows and columns just
// outside the work item's active region (y0-1, the row after the last active
// one, x0-1 of the first quad, the column after the last active quad) have no
// owner: all threads of the workgroup compute those points together.  About
// 1.3 normal evaluations per voxel instead of 3.5, and rings of 4-wide rows.
// It covers whole rows, x in [0, w) (y and z in [2, n-2)), of blocks whose rows
// start on 16-byte boundaries, so its boundary shell is the planes and rows
// within 2 of a face only.  The same operations on the same values as the
// other sweep-2 kernels: the same bits.  FIX as in the vector kernel, except that a flagged item is
// recomputed whole (its lanes read normals that other waves computed).
template <bool LIT, int TX, int TY, int WPE, int FIX = 0>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE))) void sweep2_share_kernel(
    const float* __restrict__ phi_all, const float* __restrict__ F, float* __restrict__ phi_out,
    float* __restrict__ tr_out, const BlockDesc* __restrict__ blocks, const Tile* __restrict__ tiles,
    int ntiles, int zc, float dt, float nu, float mu, float eps, float inv_eps, int* __restrict__ flags) {
    constexpr int NT = TX * TY, NW = NT / 64;
    static_assert(NT % 64 == 0, "whole waves");
    // by plane parity: n_y and g_y of rows y0-1 .. y0+TY at row index y - y0 + 1 (a quad
    // per entry); (n_x, g_x) of each quad's last point (sL) and first point (sR) at
    // index tx + 1, the left halo column at sL[.][.][0], the right one at sR[.][.][nq + 1]
    __shared__ float4 sNY[2][TY + 2][TX], sGY[2][TY + 2][TX];
    __shared__ float2 sL[2][TY][TX + 2], sR[2][TY][TX + 2];
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    for (int item = blockIdx.x; item < ntiles; item += gridDim.x) {
        const Tile tl = tiles[item];
        if (tl.b < 0) continue;
        if constexpr (LIT && FIX) {
            bool any = false;   // the whole item: uniform across the workgroup
            for (int k = 0; k < NW; ++k) any = any || flags[item * NW + k] != 0;
            if (!any) continue;
        }
        const BlockDesc bd = blocks[tl.b];
        const int w = bd.w, h = bd.h, l = bd.l;
        // whole rows: the quad at x = 0 reads its left neighbour as the previous
        // linear element, (w-1, y-1) -- the reference's x-wrap -- and the quad
        // ending at w-1 takes the one-sided rule there (kernel.cu:64-74)
        const int xe = min(tl.x0 + 4 * TX, w);        // end of the item's active columns
        const int yb = min(tl.y0 + TY, h - 2);        // end of its active rows
        const int nq = (xe - tl.x0) >> 2;             // active quads per row
        const int x0 = tl.x0 + 4 * tx, y = tl.y0 + ty;
        const bool active = tx < nq && y < yb;
        const bool qlast = x0 + 4 == w;               // the quad ends at x = w-1
        const int zb = tl.z0, zend = min(tl.z0 + zc, l - 2);
        const int W = w, P = w * h;
        const int shift = 2 * P + 2 * W + 4;
        const float* f = phi_all + bd.off;
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc((void*)(f - shift), 0, (int)0xffffffff, 0x00020000);
        const __amdgpu_buffer_rsrc_t rF =
            __builtin_amdgcn_make_buffer_rsrc((void*)(F + bd.off), 0, (int)0xffffffff, 0x00020000);
        const __amdgpu_buffer_rsrc_t rO =
            __builtin_amdgcn_make_buffer_rsrc((void*)(phi_out + bd.off), 0, (int)0xffffffff, 0x00020000);
        const __amdgpu_buffer_rsrc_t rT =
            __builtin_amdgcn_make_buffer_rsrc((void*)(tr_out + bd.off), 0, (int)0xffffffff, 0x00020000);
        // byte offset of row (y+dy, z+dz) minus 4 elements (see the vector kernel)
        auto S = [&](int dy, int dz) { return (int)(4u * (uint32_t)(shift + dz * P + dy * W - 4)); };
        const uint32_t vstep = 4u * (uint32_t)P;
        auto u2f = [](uint32_t u) { return __builtin_bit_cast(float, u); };
        // row x0-1 .. x0+4 and row x0 .. x0+3 at voxel byte offset v + row (dy, dz)
        auto load6 = [&](float (&r)[6], uint32_t v, int so) {
            r[0] = u2f(__builtin_amdgcn_raw_buffer_load_b32(rs, (int)(v + 12), so, 0));
            const u32x4 b = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(v + 16), so, 0);
            r[1] = u2f(b[0]); r[2] = u2f(b[1]); r[3] = u2f(b[2]); r[4] = u2f(b[3]);
            r[5] = u2f(__builtin_amdgcn_raw_buffer_load_b32(rs, (int)(v + 32), so, 0));
        };
        auto load4 = [&](float (&r)[4], uint32_t v, int so) {
            const u32x4 b = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(v + 16), so, 0);
            r[0] = u2f(b[0]); r[1] = u2f(b[1]); r[2] = u2f(b[2]); r[3] = u2f(b[3]);
        };
        auto loadF = [&](float (&r)[4], uint32_t v) {
            const u32x4 b = __builtin_amdgcn_raw_buffer_load_b128(rF, (int)v, 0, kNT);
            r[0] = u2f(b[0]); r[1] = u2f(b[1]); r[2] = u2f(b[2]); r[3] = u2f(b[3]);
        };
        // The point of a halo row or column at plane zp (all interior rules:
        // 1 <= y0-1, yb <= h-2, 3 <= x0-1, xe <= w-4, 2 <= zp <= l-3): its normal
        // component along the halo's axis and the gradient component, into sNY/sGY
        // (rows) or sL/sR (columns) of buffer b.
        auto halo = [&](auto EX_, bool& bad, int zp, int b) __attribute__((always_inline)) {
            constexpr bool EX = decltype(EX_)::value;
            for (int t = threadIdx.x; t < 8 * TX + 2 * TY; t += NT) {
                int x, yh;
                const bool row = t < 8 * TX;
                int side, idx;
                if (row) {
                    side = t >= 4 * TX;
                    idx = t - side * 4 * TX;
                    x = tl.x0 + idx;
                    yh = side ? yb : tl.y0 - 1;
                    if (x >= xe) continue;
                } else {
                    const int u = t - 8 * TX;
                    side = u >= TY;
                    idx = u - side * TY;
                    yh = tl.y0 + idx;
                    x = side ? xe : tl.x0 - 1;
                    // right of x = w-1 the reference takes the point itself (see the step)
                    if (yh >= yb || x >= w) continue;
                }
                const uint32_t vp = 4u * (uint32_t)(zp * P + yh * W + x) + 16u;
                auto ld = [&](int dx, int dy, int dz) {
                    return u2f(__builtin_amdgcn_raw_buffer_load_b32(rs, (int)(vp + 4 * dx), S(dy, dz), 0));
                };
                // x = w-1, and x = -1 (the previous row's last voxel), difference one-sided
                const bool xl = x == w - 1 || x == -1;
                const float gx = ((xl ? ld(0, 0, 0) : ld(1, 0, 0)) - ld(-1, 0, 0)) * 0.5f;
                const float gy = (ld(0, 1, 0) - ld(0, -1, 0)) * 0.5f;
                const float gz = (ld(0, 0, 1) - ld(0, 0, -1)) * 0.5f;
                if (row) {
                    const float ny = norm_comp<LIT, EX>(gy, gx, gy, gz, bad);
                    const int r = side ? yb - tl.y0 + 1 : 0;
                    reinterpret_cast<float*>(&sNY[b][r][idx >> 2])[idx & 3] = ny;
                    reinterpret_cast<float*>(&sGY[b][r][idx >> 2])[idx & 3] = gy;
                } else {
                    const float nx = norm_comp<LIT, EX>(gx, gx, gy, gz, bad);
                    if (side) sR[b][idx][nq + 1] = make_float2(nx, gx);
                    else sL[b][idx][0] = make_float2(nx, gx);
                }
            }
        };
        float C[4][6];                 // phi row y, x0-1 .. x0+4, ring by (plane - zb) & 3
        float U[2][4], D[2][4];        // rows y+1, y-1 at x0 .. x0+3, by (plane - zb) & 1
        float FF[2][4];                // F of the plane, by parity
        float NX[2][4], GX[2][4];      // own points: n_x, g_x by parity
        float NZ[4][4], GZ[4][4];      // own points: n_z, g_z, ring by (plane - zb) & 3
        // the own points (x0+i, y, p) from rows cm (p-1), c0 (p), cp (p+1), u, d (p): the z
        // components (and n_x, g_x) kept, n_y and g_y (and the ends' n_x, g_x) into LDS
        auto own = [&](auto EX_, bool& bad, const float (&cm)[6], const float (&c0)[6], const float (&cp)[6],
                       const float (&u)[4], const float (&d)[4], float (&nx)[4], float (&gx)[4], float (&nz)[4],
                       float (&gz)[4], int b, bool share) __attribute__((always_inline)) {
            float ny[4], gy[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                gx[i] = ((i == 3 && qlast ? c0[i + 1] : c0[i + 2]) - c0[i]) * 0.5f;
                gy[i] = (u[i] - d[i]) * 0.5f;
                gz[i] = (cp[i + 1] - cm[i + 1]) * 0.5f;
                norm_comp3<LIT, decltype(EX_)::value>(gx[i], gy[i], gz[i], nx[i], ny[i], nz[i], bad);
                if constexpr (LIT) __builtin_amdgcn_sched_barrier(0);   // one point's temporaries at a time
            }
            if (share && active) {
                sNY[b][ty + 1][tx] = make_float4(ny[0], ny[1], ny[2], ny[3]);
                sGY[b][ty + 1][tx] = make_float4(gy[0], gy[1], gy[2], gy[3]);
                sL[b][ty][tx + 1] = make_float2(nx[3], gx[3]);
                sR[b][ty][tx + 1] = make_float2(nx[0], gx[0]);
            }
        };
        // Every lane loads and computes (no divergent branch around the register
        // rings, which would make the compiler copy them): an inactive lane reads
        // the item's first voxel instead of its own, and only its stores, LDS
        // writes and flag are masked.
        const uint32_t v0 = 4u * (uint32_t)(zb * P + (active ? y : tl.y0) * W + (active ? x0 : tl.x0));
        bool bad = false;   // literal, FIX = 0: some point this lane computed is outside the fast forms' range
        // ---- prologue: rows of planes zb-2 .. zb+2, own points at zb-1 and zb, halo at zb
        {
            float Cmm[6], Um[4], Dm[4];
            load6(Cmm, v0, S(0, -2));
            load6(C[3], v0, S(0, -1));
            load6(C[0], v0, S(0, 0));
            load6(C[1], v0, S(0, 1));
            load6(C[2], v0, S(0, 2));
            load4(Um, v0, S(1, -1));
            load4(Dm, v0, S(-1, -1));
            load4(U[0], v0, S(1, 0));
            load4(D[0], v0, S(-1, 0));
            load4(U[1], v0, S(1, 1));
            load4(D[1], v0, S(-1, 1));
            loadF(FF[0], v0);
            bool unused = false;
            float nx[4], gx[4];
            own(std::true_type{}, unused, Cmm, C[3], C[0], Um, Dm, nx, gx, NZ[3], GZ[3], 0, false);
            own(std::true_type{}, unused, C[3], C[0], C[1], U[0], D[0], NX[0], GX[0], NZ[0], GZ[0], 0, true);
        }
        {
            bool unused = false;
            halo(std::true_type{}, unused, zb, 0);
        }
        __syncthreads();
        auto step = [&](auto S_, uint32_t v, int z) __attribute__((always_inline)) {
            constexpr int s = decltype(S_)::value;
            constexpr int km1 = (s + 3) & 3, k0 = s & 3, kp1 = (s + 1) & 3, kp2 = (s + 2) & 3;
            constexpr int q = s & 1, qn = (s + 1) & 1;
            // plane z+3 (slot of z-1), or z+2 at the chunk's last plane, whose
            // loads feed no later step (z+3 may lie past the block)
            load6(C[km1], v + vstep, S(0, z + 1 < zend ? 2 : 1));
            load4(U[q], v + vstep, S(1, 1));            // plane z+2 (slot of z)
            load4(D[q], v + vstep, S(-1, 1));
            loadF(FF[qn], v + vstep);                   // plane z+1
            constexpr bool EX = FIX != 0;
            {
                // own points at plane z+1: the z neighbours now, shared for the next plane
                own(std::integral_constant<bool, EX>{}, bad, C[k0], C[kp1], C[kp2], U[qn], D[qn], NX[qn], GX[qn],
                    NZ[kp1], GZ[kp1], qn, z + 1 < zend);
                // plane z of the neighbours, written at the previous plane (buffer q)
                const float4 nyu = sNY[q][ty + 2][tx], nyd = sNY[q][ty][tx];
                const float4 gyu = sGY[q][ty + 2][tx], gyd = sGY[q][ty][tx];
                const float2 L = sL[q][ty][tx], R = sR[q][ty][tx + 2];
                const float nyus[4] = {nyu.x, nyu.y, nyu.z, nyu.w}, nyds[4] = {nyd.x, nyd.y, nyd.z, nyd.w};
                const float gyus[4] = {gyu.x, gyu.y, gyu.z, gyu.w}, gyds[4] = {gyd.x, gyd.y, gyd.z, gyd.w};
                // at x = w-1 the right neighbour is the point itself
                const float nxs[6] = {L.x, NX[q][0], NX[q][1], NX[q][2], NX[q][3], qlast ? NX[q][3] : R.x};
                const float gxs[6] = {L.y, GX[q][0], GX[q][1], GX[q][2], GX[q][3], qlast ? GX[q][3] : R.y};
                u32x4 po, pt;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float kap = half_sum3(nxs[i + 2] - nxs[i], nyus[i] - nyds[i], NZ[kp1][i] - NZ[km1][i]);
                    const float lap = half_sum3(gxs[i + 2] - gxs[i], gyus[i] - gyds[i], GZ[kp1][i] - GZ[km1][i]);
                    const float c0 = C[k0][i + 1];
                    const float d = LIT ? dirac_literal(c0, eps) : dirac_fast(c0, eps);
                    const float p1 = phi_update<LIT>(c0, d, FF[q][i], kap, lap, dt, nu * dt, mu);
                    const float t1 = LIT ? heav_literal_r(p1, eps, inv_eps) : heav_centred(p1, inv_eps);
                    po[i] = __builtin_bit_cast(uint32_t, p1);
                    pt[i] = __builtin_bit_cast(uint32_t, t1);
                    if constexpr (LIT) __builtin_amdgcn_sched_barrier(0);
                }
                if (active) {
                    __builtin_amdgcn_raw_buffer_store_b128(po, rO, (int)v, 0, kNT);
                    __builtin_amdgcn_raw_buffer_store_b128(pt, rT, (int)v, 0, kNT);
                }
            }
            if (z + 1 < zend) halo(std::integral_constant<bool, EX>{}, bad, z + 1, qn);
            // this plane's reads of buffer q precede the next plane's writes to it
            __syncthreads();
        };
        uint32_t v = v0;
        for (int z = zb; z < zend; z += 4) {
            step(std::integral_constant<int, 0>{}, v, z);
            if (z + 1 >= zend) break;
            step(std::integral_constant<int, 1>{}, v + vstep, z + 1);
            if (z + 2 >= zend) break;
            step(std::integral_constant<int, 2>{}, v + 2 * vstep, z + 2);
            if (z + 3 >= zend) break;
            step(std::integral_constant<int, 3>{}, v + 3 * vstep, z + 3);
            v += 4 * vstep;
        }
        if constexpr (LIT && !FIX) {
            const bool wbad = __builtin_amdgcn_ballot_w64(bad) != 0;
            if (threadIdx.x % 64 == 0) flags[item * NW + threadIdx.x / 64] = wbad ? 1 : 0;
        }
    }
}

// Sweep 2, direct form: each thread marches one (x,y) column of its tile
// through z.  Gradients at the six D-neighbours come from L1/L2-cached loads.
template <bool LIT, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void sweep2_kernel(
    const float* __restrict__ phi_all, const float* __restrict__ F, float* __restrict__ phi_out,
    float* __restrict__ tr_out, const BlockDesc* __restrict__ blocks, const Tile* __restrict__ tiles,
    int ntiles, int zc, float dt, float nu, float mu, float eps, float inv_eps) {
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    for (int item = blockIdx.x; item < ntiles; item += gridDim.x) {
        const Tile tl = tiles[item];
        if (tl.b < 0) continue;
        const BlockDesc bd = blocks[tl.b];
        const int w = bd.w, h = bd.h, l = bd.l;
        const int x = tl.x0 + tx, y = tl.y0 + ty;
        if (x >= w || y >= h) continue;
        const float* f = phi_all + bd.off;
        const int zend = min(tl.z0 + zc, l);
        // D-neighbour coordinates that do not depend on z
        int lxx, lxy, lxz_off;  // left-x operand: (lxx, lxy, z + lxz_off)
        if (x > 0) { lxx = x - 1; lxy = y; lxz_off = 0; }
        else if (y > 0) { lxx = w - 1; lxy = y - 1; lxz_off = 0; }
        else { lxx = w - 1; lxy = h - 1; lxz_off = -1; }
        const int rxx = (x == w - 1) ? x : x + 1;
        const int ryy = (y == 0) ? 1 : (y == h - 1 ? y : y + 1);
        const int lyy = (y == 0) ? 0 : (y == h - 1 ? h - 2 : y - 1);
        const bool xy_inner = x >= 2 && x <= w - 3 && y >= 2 && y <= h - 3;
        const int64_t P = (int64_t)w * h;
        // Interior voxels read their radius-2 stencil through a buffer resource
        // whose base sits 2P+2W+2 elements before the block: every tap then has
        // a non-negative byte offset = voffset (4i) + SGPR (y, z part) + imm (x part).
        const int64_t shift = 2 * P + 2 * (int64_t)w + 2;
        const bool buf_ok = ((int64_t)l * P + 2 * shift) * 4 < 0xffffffffLL;
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc((void*)(f - shift), 0, (int)0xffffffff, 0x00020000);
        auto S = [&](int ky, int kz) { return (uint32_t)(((2 + ky) * (int64_t)w + (2 + kz) * P) * 4); };
        const uint32_t s00 = S(0, 0), sp10 = S(1, 0), sm10 = S(-1, 0), sp20 = S(2, 0), sm20 = S(-2, 0);
        const uint32_t s0p1 = S(0, 1), s0m1 = S(0, -1), s0p2 = S(0, 2), s0m2 = S(0, -2);
        const uint32_t spp = S(1, 1), spm = S(1, -1), smp = S(-1, 1), smm = S(-1, -1);
        for (int z = tl.z0; z < zend; ++z) {
            const int64_t i = ((int64_t)z * h + y) * w + x;
            if (buf_ok && xy_inner && z >= 2 && z <= l - 3) {
                const uint32_t v = (uint32_t)i * 4u;
                auto L = [&](int kx, uint32_t so) {
                    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)(v + 4 * (kx + 2)),
                                                                                           (int)so, 0));
                };
                const float c0 = L(0, s00);
                const float xp2 = L(2, s00), xm2 = L(-2, s00);
                const float yp2 = L(0, sp20), ym2 = L(0, sm20);
                const float zp2 = L(0, s0p2), zm2 = L(0, s0m2);
                const float xpyp = L(1, sp10), xpym = L(1, sm10), xmyp = L(-1, sp10), xmym = L(-1, sm10);
                const float xpzp = L(1, s0p1), xpzm = L(1, s0m1), xmzp = L(-1, s0p1), xmzm = L(-1, s0m1);
                const float ypzp = L(0, spp), ypzm = L(0, spm), ymzp = L(0, smp), ymzm = L(0, smm);
                // gradients at the six D-neighbours
                const Grad gRx{(xp2 - c0) * 0.5f, (xpyp - xpym) * 0.5f, (xpzp - xpzm) * 0.5f};
                const Grad gLx{(c0 - xm2) * 0.5f, (xmyp - xmym) * 0.5f, (xmzp - xmzm) * 0.5f};
                const Grad gRy{(xpyp - xmyp) * 0.5f, (yp2 - c0) * 0.5f, (ypzp - ypzm) * 0.5f};
                const Grad gLy{(xpym - xmym) * 0.5f, (c0 - ym2) * 0.5f, (ymzp - ymzm) * 0.5f};
                const Grad gRz{(xpzp - xmzp) * 0.5f, (ypzp - ymzp) * 0.5f, (zp2 - c0) * 0.5f};
                const Grad gLz{(xpzm - xmzm) * 0.5f, (ypzm - ymzm) * 0.5f, (c0 - zm2) * 0.5f};
                float nRx, nLx, nRy, nLy, nRz, nLz;
                if constexpr (LIT) {
                    auto nrm = [](const Grad& g, float cc) {
                        const float M = sqrtf(mag2<LIT>(g.x, g.y, g.z));
                        return (float)((double)cc / ((double)M + 1e-6));
                    };
                    nRx = nrm(gRx, gRx.x); nLx = nrm(gLx, gLx.x);
                    nRy = nrm(gRy, gRy.y); nLy = nrm(gLy, gLy.y);
                    nRz = nrm(gRz, gRz.z); nLz = nrm(gLz, gLz.z);
                } else {
                    auto nrm = [](const Grad& g, float cc) {
                        const float M = __builtin_amdgcn_sqrtf(mag2<false>(g.x, g.y, g.z));
                        return cc * __builtin_amdgcn_rcpf(M + 1e-6f);
                    };
                    nRx = nrm(gRx, gRx.x); nLx = nrm(gLx, gLx.x);
                    nRy = nrm(gRy, gRy.y); nLy = nrm(gLy, gLy.y);
                    nRz = nrm(gRz, gRz.z); nLz = nrm(gLz, gLz.z);
                }
                const float kap = half_sum3(nRx - nLx, nRy - nLy, nRz - nLz);
                const float lap = half_sum3(gRx.x - gLx.x, gRy.y - gLy.y, gRz.z - gLz.z);
                const float d = LIT ? dirac_literal(c0, eps) : dirac_fast(c0, eps);
                const float Fv = F[bd.off + i];
                const float p1 = phi_update<LIT>(c0, d, Fv, kap, lap, dt, nu * dt, mu);
                phi_out[bd.off + i] = p1;
                tr_out[bd.off + i] = LIT ? heav_literal_r(p1, eps, inv_eps) : heav_centred(p1, inv_eps);
                continue;
            }
            int qlx = lxx, qly = lxy, qlz = z + lxz_off;
            if (qlz < 0) { qlx = 0; qly = 0; qlz = 0; }  // voxel (0,0,0): clamp
            const int rzz = (z == 0) ? 1 : (z == l - 1 ? z : z + 1);
            const int lzz = (z == 0) ? 0 : (z == l - 1 ? l - 2 : z - 1);
            const Grad gRx = grad_at(f, rxx, y, z, w, h, l);
            const Grad gLx = grad_at(f, qlx, qly, qlz, w, h, l);
            const Grad gRy = grad_at(f, x, ryy, z, w, h, l);
            const Grad gLy = grad_at(f, x, lyy, z, w, h, l);
            const Grad gRz = grad_at(f, x, y, rzz, w, h, l);
            const Grad gLz = grad_at(f, x, y, lzz, w, h, l);
            float nRx, nLx, nRy, nLy, nRz, nLz;
            if constexpr (LIT) {
                auto nrm = [](const Grad& g, float c) {
                    const float M = sqrtf(mag2<LIT>(g.x, g.y, g.z));
                    return (float)((double)c / ((double)M + 1e-6));
                };
                nRx = nrm(gRx, gRx.x); nLx = nrm(gLx, gLx.x);
                nRy = nrm(gRy, gRy.y); nLy = nrm(gLy, gLy.y);
                nRz = nrm(gRz, gRz.z); nLz = nrm(gLz, gLz.z);
            } else {
                auto nrm = [](const Grad& g, float c) {
                    const float M = __builtin_amdgcn_sqrtf(mag2<false>(g.x, g.y, g.z));
                    return c * __builtin_amdgcn_rcpf(M + 1e-6f);
                };
                nRx = nrm(gRx, gRx.x); nLx = nrm(gLx, gLx.x);
                nRy = nrm(gRy, gRy.y); nLy = nrm(gLy, gLy.y);
                nRz = nrm(gRz, gRz.z); nLz = nrm(gLz, gLz.z);
            }
            const float kap = half_sum3(nRx - nLx, nRy - nLy, nRz - nLz);
            const float lap = half_sum3(gRx.x - gLx.x, gRy.y - gLy.y, gRz.z - gLz.z);
            const float p0 = f[i];
            const float d = LIT ? dirac_literal(p0, eps) : dirac_fast(p0, eps);
            const float Fv = F[bd.off + i];
            const float p1 = phi_update<LIT>(p0, d, Fv, kap, lap, dt, nu * dt, mu);
            phi_out[bd.off + i] = p1;
            tr_out[bd.off + i] = LIT ? heav_literal_r(p1, eps, inv_eps) : heav_centred(p1, inv_eps);
        }
    }
}

// Sweep 2, LDS z-march.  Per gradient plane s the workgroup keeps phi planes
// s-1, s, s+1 of its tile (+2 halo) in an LDS ring, computes g = D phi and
// n = g/(|g|+1e-6) once per voxel of the tile (+1 halo) into LDS, keeps the
// z components of the own column in registers, and then emits output plane
// s-1 from the gradient planes s-2, s-1, s.  Edge rules as the reference:
// one-sided halves at the y/z ends and at x = w-1; at x = 0 the left operand
// is the previous linear element, so the x = -1 halo column holds phi (and g,
// n) of that element.
template <bool LIT, int TX, int TY, int WPE>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE))) void sweep2_lds_kernel(
    const float* __restrict__ phi_all, const float* __restrict__ F, float* __restrict__ phi_out,
    float* __restrict__ tr_out, const BlockDesc* __restrict__ blocks, const Tile* __restrict__ tiles,
    int ntiles, int zc, float dt, float nu, float mu, float eps, float inv_eps) {
    constexpr int NT = TX * TY;
    constexpr int PW = TX + 4, PH = TY + 4;   // phi tile with 2-voxel halo
    constexpr int GW = TX + 2, GH = TY + 2;   // gradient tile with 1-voxel halo
    __shared__ float sp[3][PH][PW];
    __shared__ float sgx[2][GH][GW], sgy[2][GH][GW], snx[2][GH][GW], sny[2][GH][GW];
    __shared__ float snz[TY][TX], sgz[TY][TX];  // z components of the interior, current plane
    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;

    for (int item = blockIdx.x; item < ntiles; item += gridDim.x) {
        const Tile tl = tiles[item];
        if (tl.b < 0) continue;
        const BlockDesc bd = blocks[tl.b];
        const int w = bd.w, h = bd.h, l = bd.l;
        const int64_t plane = (int64_t)w * h;
        const float* f = phi_all + bd.off;
        const int x0 = tl.x0, y0 = tl.y0, z0 = tl.z0;
        const int zend = min(z0 + zc, l);
        const int s_begin = max(z0 - 1, 0), s_last = min(zend, l - 1);

        // phi plane q into ring slot q % 3 (clamped reads; x = -1 holds the wrap)
        auto load_plane = [&](int q) {
            float(*dst)[PW] = sp[q % 3];
            for (int e = tid; e < PH * PW; e += NT) {
                const int r = e / PW, c = e - r * PW;
                const int gy = min(max(y0 - 2 + r, 0), h - 1);
                const int gx = x0 - 2 + c;
                float v;
                if (gx == -1 && y0 - 2 + r >= 0 && y0 - 2 + r < h) {
                    // previous linear element of (0, gy, q)
                    const int64_t lin = (int64_t)q * plane + (int64_t)gy * w - 1;
                    v = f[lin < 0 ? 0 : lin];
                } else {
                    v = f[(int64_t)q * plane + (int64_t)gy * w + min(max(gx, 0), w - 1)];
                }
                dst[r][c] = v;
            }
        };
        float nz_m2 = 0.f, gz_m2 = 0.f, nz_m1 = 0.f, gz_m1 = 0.f, nz_0 = 0.f, gz_0 = 0.f;

        // emission of output plane z from gradient planes (buffer zb = z&1 holds plane z)
        auto emit = [&](int z, float nzR, float gzR, float nzL, float gzL) {
            const int x = x0 + tx, y = y0 + ty;
            if (x >= w || y >= h) return;
            const int b = z & 1;
            const int cx = tx + 1, cy = ty + 1;
            const int cxr = (x == w - 1) ? cx : cx + 1, cxl = cx - 1;
            const int cyr = (y == 0) ? cy + 1 : (y == h - 1 ? cy : cy + 1);
            const int cyl = (y == 0) ? cy : (y == h - 1 ? cy - 1 : cy - 1);
            const float kap = half_sum3(snx[b][cy][cxr] - snx[b][cy][cxl], sny[b][cyr][cx] - sny[b][cyl][cx], nzR - nzL);
            const float lap = half_sum3(sgx[b][cy][cxr] - sgx[b][cy][cxl], sgy[b][cyr][cx] - sgy[b][cyl][cx], gzR - gzL);
            const float p0 = sp[z % 3][ty + 2][tx + 2];
            const float d = LIT ? dirac_literal(p0, eps) : dirac_fast(p0, eps);
            const int64_t i = bd.off + (int64_t)z * plane + (int64_t)y * w + x;
            const float p1 = phi_update<LIT>(p0, d, F[i], kap, lap, dt, nu * dt, mu);
            phi_out[i] = p1;
            tr_out[i] = LIT ? heav_literal_r(p1, eps, inv_eps) : heav_centred(p1, inv_eps);
        };

        load_plane(s_begin);
        if (s_begin + 1 <= l - 1) load_plane(s_begin + 1);
        for (int s = s_begin; s <= s_last; ++s) {
            if (s == s_begin && s > 0) load_plane(s - 1);
            if (s + 1 <= l - 1 && s + 1 > s_begin + 1) load_plane(s + 1);
            __syncthreads();
            // ---- gradient plane s on the tile + 1 halo
            const int zr = (s == 0) ? 1 : (s == l - 1 ? s : s + 1);
            const int zl = (s == 0) ? 0 : (s == l - 1 ? l - 2 : s - 1);
            const int gb = s & 1;
            for (int e = tid; e < GH * GW; e += NT) {
                const int r = e / GW, c = e - r * GW;
                const int x = x0 - 1 + c, y = y0 - 1 + r;
                Grad g;
                if (x < 0 || x >= w || y < 0 || y >= h) {
                    if (x == -1 && y >= 0 && y < h) {
                        // D_x left operand of (0, y, s): the previous linear element
                        int qx, qy, qz;
                        left_x(0, y, s, w, h, qx, qy, qz);
                        g = grad_at(f, qx, qy, qz, w, h, l);
                    } else {
                        g.x = g.y = g.z = 0.f;
                    }
                } else {
                    const int pc = c + 1, pr = r + 1;  // phi tile coords of (x, y)
                    const float* row = sp[s % 3][pr];
                    const int pcr = (x == w - 1) ? pc : pc + 1;
                    const int prr = (y == 0) ? pr + 1 : (y == h - 1 ? pr : pr + 1);
                    const int prl = (y == 0) ? pr : (y == h - 1 ? pr - 1 : pr - 1);
                    g.x = (row[pcr] - row[pc - 1]) * 0.5f;
                    g.y = (sp[s % 3][prr][pc] - sp[s % 3][prl][pc]) * 0.5f;
                    g.z = (sp[zr % 3][pr][pc] - sp[zl % 3][pr][pc]) * 0.5f;
                }
                float r_;
                if constexpr (LIT) {
                    const float M = sqrtf(mag2<LIT>(g.x, g.y, g.z));
                    snx[gb][r][c] = (float)((double)g.x / ((double)M + 1e-6));
                    sny[gb][r][c] = (float)((double)g.y / ((double)M + 1e-6));
                    r_ = 0.f;
                    if (r >= 1 && r <= TY && c >= 1 && c <= TX) {
                        snz[r - 1][c - 1] = (float)((double)g.z / ((double)M + 1e-6));
                        sgz[r - 1][c - 1] = g.z;
                    }
                } else {
                    const float M = __builtin_amdgcn_sqrtf(mag2<false>(g.x, g.y, g.z));
                    r_ = __builtin_amdgcn_rcpf(M + 1e-6f);
                    snx[gb][r][c] = g.x * r_;
                    sny[gb][r][c] = g.y * r_;
                    if (r >= 1 && r <= TY && c >= 1 && c <= TX) {
                        snz[r - 1][c - 1] = g.z * r_;
                        sgz[r - 1][c - 1] = g.z;
                    }
                }
                sgx[gb][r][c] = g.x;
                sgy[gb][r][c] = g.y;
            }
            __syncthreads();
            nz_0 = snz[ty][tx];
            gz_0 = sgz[ty][tx];
            // ---- output plane s-1 (its D_z partners: plane s and plane s-2, or s-1 itself at z = 0)
            const int z = s - 1;
            if (z >= z0 && z < zend) {
                if (z == 0) emit(z, nz_0, gz_0, nz_m1, gz_m1);
                else emit(z, nz_0, gz_0, nz_m2, gz_m2);
            }
            // the last plane of the volume: D_z partners are itself and l-2
            if (s == l - 1 && s >= z0 && s < zend) emit(s, nz_0, gz_0, nz_m1, gz_m1);
            nz_m2 = nz_m1; gz_m2 = gz_m1;
            nz_m1 = nz_0; gz_m1 = gz_0;
            __syncthreads();
        }
    }
}

#pragma clang fp contract(fast)

// T(phi) over n voxels: four per thread (one 16-byte load and store) where both
// buffers are 16-byte aligned (V4), then the tail one per thread.
template <bool LIT, bool V4>
__global__ void transform_kernel(float* __restrict__ tr, const float* __restrict__ phi, int64_t n, float eps,
                                 float inv_eps) {
    auto T = [&](float p) { return LIT ? heav_literal_r(p, eps, inv_eps) : heav_centred(p, inv_eps); };
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t done = 0;
    if constexpr (V4) {
        const int64_t n4 = n >> 2;
        for (int64_t j = i; j < n4; j += step) {
            const float4 p = reinterpret_cast<const float4*>(phi)[j];
            reinterpret_cast<float4*>(tr)[j] = make_float4(T(p.x), T(p.y), T(p.z), T(p.w));
        }
        done = n4 << 2;
    }
    for (int64_t j = done + i; j < n; j += step) tr[j] = T(phi[j]);
}

template <int P, bool LIT, int TX, int TY, int NT, int WPE = 0, bool DB = false>
Sweep1Variant s1v(const char* name) {
    return Sweep1Variant{P, LIT ? 1 : 0, TX, TY, NT, (const void*)&sweep1_kernel<P, LIT, TX, TY, NT, WPE, DB>, name,
                         nullptr, 0, 0, {}};
}
// split sweep 1 (fast mode): xy kernel with TX x TY tiles, z kernel with TXC x TYC columns
template <bool LIT, int TX, int TY>
Sweep2Variant s2v(const char* name) {
    return Sweep2Variant{LIT ? 1 : 0, TX, TY, 1, (const void*)&sweep2_kernel<LIT, TX, TY>, nullptr, nullptr, name};
}
template <bool LIT, int TX, int TY, int WPE>
Sweep2Variant s2l(const char* name) {
    return Sweep2Variant{LIT ? 1 : 0, TX, TY, 1, (const void*)&sweep2_lds_kernel<LIT, TX, TY, WPE>, nullptr, nullptr,
                         name};
}
template <bool LIT, int TX, int TY>
Sweep2Variant s2s(const char* name) {
    return Sweep2Variant{LIT ? 1 : 0, TX, TY, 1, (const void*)&sweep2_inner_kernel<LIT, TX, TY>,
                         (const void*)&sweep2_shell_kernel<LIT>, nullptr, name};
}
// vector interior for blocks with 16-byte aligned rows, the 64x4 scalar interior for the others
// (literal: the fast-form interior plus its exact fix-up)
template <bool LIT, int TX, int TY, int WPE>
Sweep2Variant s2q(const char* name) {
    return Sweep2Variant{LIT ? 1 : 0, TX, TY, 4, (const void*)&sweep2_vec_kernel<LIT, TX, TY, WPE, 0>,
                         (const void*)&sweep2_shell_kernel<LIT>, (const void*)&sweep2_inner_kernel<LIT, 64, 4>,
                         name, LIT ? (const void*)&sweep2_vec_kernel<LIT, TX, TY, WPE, 1> : nullptr};
}

// shared interior (each point's normal once, neighbours through LDS) for blocks with 16-byte
// aligned rows, the 64x4 scalar interior for the others
template <bool LIT, int TX, int TY, int WPE>
Sweep2Variant s2h(const char* name) {
    return Sweep2Variant{LIT ? 1 : 0, TX, TY, 4, (const void*)&sweep2_share_kernel<LIT, TX, TY, WPE, 0>,
                         (const void*)&sweep2_shell_kernel<LIT>, (const void*)&sweep2_inner_kernel<LIT, 64, 4>,
                         name, LIT ? (const void*)&sweep2_share_kernel<LIT, TX, TY, WPE, 1> : nullptr, 1};
}

// split sweep 1, fast mode, xy kernel on the matrix cores (TX x 28 tiles)
template <int P, int TX, int NT, int WPE, int TXC, int TYC, int D>
Sweep1Variant s1m(const char* name);

// split sweep 1: xy kernel with TX x TY tiles, z kernel with TXC x TYC columns
template <int P, bool LIT, int TX, int TY, int NT, int TXC, int TYC, int D, bool IL = false>
Sweep1Variant s1split(const void* xy_fn, const char* name) {
    Sweep1Variant v{P, LIT ? 1 : 0, TX, TY, NT, xy_fn, name,
                    (const void*)&sweep1_z_kernel<P, LIT, TXC, TYC, D, true, 0, IL>, TXC, TYC, {}};
    v.nt2 = TXC * TYC;
    v.fn2s[0] = (const void*)&sweep1_z_kernel<P, LIT, TXC, TYC, D, false, 0, IL>;
    v.fn2s[1] = (const void*)&sweep1_z_kernel<P, LIT, TXC, TYC, D, true, 0, IL>;
    v.fn2s[2] = (const void*)&sweep1_z_kernel<P, LIT, TXC, TYC, D, false, 1, IL>;
    v.fn2s[3] = (const void*)&sweep1_z_kernel<P, LIT, TXC, TYC, D, true, 1, IL>;
    v.fn2s[4] = (const void*)&sweep1_z_kernel<P, LIT, TXC, TYC, D, false, 2, IL>;
    v.fn2s[5] = (const void*)&sweep1_z_kernel<P, LIT, TXC, TYC, D, true, 2, IL>;
    v.fn2s[6] = (const void*)&sweep1_z_kernel<P, LIT, TXC, TYC, D, false, 3, IL>;
    v.fn2s[7] = (const void*)&sweep1_z_kernel<P, LIT, TXC, TYC, D, true, 3, IL>;
    v.fn2s[8] = (const void*)&sweep1_z_kernel<P, LIT, TXC, TYC, D, false, 4>;
    v.fn2s[9] = (const void*)&sweep1_z_kernel<P, LIT, TXC, TYC, D, true, 4>;
    return v;
}

template <int P, bool LIT, int TX, int TY, int NT, int R, int WPE, int TXC, int TYC, int D, int HI = 0,
          int RX = 4, int UX = 1, int UY = 1, bool IL = false>
Sweep1Variant s1s(const char* name) {
    Sweep1Variant v = s1split<P, LIT, TX, TY, NT, TXC, TYC, D, IL>(
        (const void*)&sweep1_xy_kernel<P, LIT, TX, TY, NT, R, WPE, HI, RX, UX, UY, 1, false, IL>, name);
    // a float image's four-channel kernel on the interleaved layout spills (its
    // prefetch holds a float4 per four voxels): it forms channel pairs instead
    constexpr int HIF = IL && HI == 2 ? 1 : HI;
    v.fn_img[0] = (const void*)&sweep1_xy_kernel<P, LIT, TX, TY, NT, R, WPE, HIF, RX, UX, UY, 0, false, IL>;
    v.fn_img[1] = (const void*)&sweep1_xy_kernel<P, LIT, TX, TY, NT, R, WPE, HI, RX, UX, UY, 1, false, IL>;
    if constexpr (IL) v.max_plane = (int64_t)(0xffffffffu / (16u * (P + D)));
    if constexpr (!LIT) {
        v.fn_blur[0] = (const void*)&sweep1_xy_kernel<P, LIT, TX, TY, NT, R, WPE, HI, RX, UX, UY, 0, true>;
        v.fn_blur[1] = (const void*)&sweep1_xy_kernel<P, LIT, TX, TY, NT, R, WPE, HI, RX, UX, UY, 1, true>;
    }
    return v;
}
// split sweep 1: the strip xy kernel (NW waves per workgroup, one plane per wave
// at a time) and the z kernel (literal mode: on the interleaved intermediate)
template <int P, bool LIT, int NCOL, int NW, int D, int WPE, int TXC, int DZ>
Sweep1Variant s1strip(const char* name) {
    Sweep1Variant v = s1split<P, LIT, (LIT ? 16 : 32) * NCOL, 1 << 20, 64 * NW, TXC, 1, DZ, LIT>(
        (const void*)&sweep1_xy_strip_kernel<P, LIT, NCOL, NW, D, WPE, 1>, name);
    v.fn_img[0] = (const void*)&sweep1_xy_strip_kernel<P, LIT, NCOL, NW, D, WPE, 0>;
    v.fn_img[1] = (const void*)&sweep1_xy_strip_kernel<P, LIT, NCOL, NW, D, WPE, 1>;
    if constexpr (LIT) v.max_plane = (int64_t)(0xffffffffu / (16u * (P + DZ)));
    v.strip = NW;
    return v;
}
template <int P, int TX, int NT, int WPE, int TXC, int TYC, int D>
Sweep1Variant s1m(const char* name) {
    Sweep1Variant v = s1split<P, false, TX, 28, NT, TXC, TYC, D>(
        (const void*)&sweep1_xy_mfma_kernel<P, TX, NT, WPE, 1>, name);
    v.fn_img[0] = (const void*)&sweep1_xy_mfma_kernel<P, TX, NT, WPE, 0>;
    v.fn_img[1] = (const void*)&sweep1_xy_mfma_kernel<P, TX, NT, WPE, 1>;
    return v;
}
}  // namespace

const std::vector<Sweep1Variant>& sweep1_variants() {
    static const std::vector<Sweep1Variant> v = {
        // fast default: the strip xy kernel, 128-column strips, a plane per wave (xy
        // kernel at 512^3: 0.49 ms against 0.60 for the 64x44 tiles below)
        s1strip<21, false, 4, 4, 1, 4, 256, 3>("p21_fast_strip_128w4"),
        // tiled xy kernels on channel planes (tuner candidates)
        s1s<21, false, 64, 44, 512, 11, 4, 256, 1, 3>("p21_fast_split_64x44w4"),
        s1s<21, false, 64, 28, 512, 7, 4, 256, 1, 3>("p21_fast_split_64x28w4"),
        s1s<21, false, 32, 28, 256, 7, 3, 256, 1, 3>("p21_fast_split_32x28w3"),
        s1s<21, false, 64, 44, 512, 11, 4, 256, 1, 7>("p21_fast_split_64x44d7"),
        // x and y passes on the matrix cores: a tuner candidate, 1.56 ms per sweep 1 at
        // 512^3 against 1.36 for the default (MFMA at about a third of its rate)
        s1m<21, 64, 256, 2, 256, 1, 3>("p21_fast_mfma_64x28"),
        // the strip xy kernel in fast mode, 64-column strips (the first of them the fast
        // default for blocks narrower than three 128-column strips)
        s1strip<21, false, 2, 4, 3, 5, 256, 3>("p21_fast_strip_64w5"),
        s1strip<21, false, 2, 4, 3, 4, 256, 3>("p21_fast_strip_64w4"),
        // literal default: the strip xy kernel, 64-column strips, a plane per wave (xy
        // kernel at 512^3: 0.82 ms against 1.02 for the 32x60 tiles below), with the
        // interleaved z kernel on work items of 1024 consecutive columns, one 1024-thread
        // workgroup per CU (0.520 ms against 0.543 with 256 columns at 512^3: a
        // workgroup's reads of a plane span 16 KB instead of 4)
        s1strip<21, true, 4, 4, 1, 4, 1024, 3>("p21_lit_strip_64z1k"),
        // 256-column z work items and 3 rows of loads in flight: the literal default for
        // planes that do not give every CU a 1024-column work item (sweep 1 at 256^3:
        // 0.236 ms against 0.281 with one row in flight, 96^3 to 160^3 and 8 x 128^3
        // 3-7% faster)
        s1strip<21, true, 4, 4, 3, 4, 256, 3>("p21_lit_strip_64w4d3"),
        s1strip<21, true, 4, 4, 1, 4, 256, 3>("p21_lit_strip_64w4"),
        // 32x60 tiles, 15 outputs per y item, each x-pass item forming all four channels
        // from one H/I window, the intermediate interleaved per voxel (the default for
        // narrow blocks before the 32-column strips; sweep 1 at 512^3: 2.05 ms before the z
        // kernel's band rework; 2.20 with channel planes, 2.24 with channel pairs per
        // x-pass item)
        s1s<21, true, 32, 60, 256, 15, 2, 256, 1, 3, 2, 4, 1, 1, true>("p21_lit_il_32x60r15"),
        // the same on channel planes: the default where a plane is too large for the
        // interleaved z kernel's 32-bit offsets
        s1s<21, true, 32, 60, 256, 15, 2, 256, 1, 3, 2>("p21_lit_q4_32x60r15"),
        // 32x60 tiles, 15 outputs per y item, one channel pair per x-pass item (the xy
        // kernel specialised for the image dtype: 2.36 ms at 512^3, 15.0 at 1024x1024x768)
        s1s<21, true, 32, 60, 256, 15, 2, 256, 1, 3, true>("p21_lit_hi_32x60r15"),
        s1s<21, true, 32, 60, 256, 10, 2, 256, 1, 3, true>("p21_lit_hi_32x60r10"),
        s1s<21, true, 32, 60, 256, 30, 2, 256, 1, 3, true>("p21_lit_hi_32x60r30"),
        s1s<21, true, 16, 108, 256, 9, 2, 256, 1, 3, true>("p21_lit_hi_16x108r9"),
        s1s<21, true, 16, 108, 256, 9, 2, 256, 1, 3, true, 8>("p21_lit_hi_16x108x8"),
        s1s<21, true, 32, 44, 256, 11, 2, 256, 1, 3, true>("p21_lit_hi_32x44w2"),
        s1s<21, true, 32, 44, 256, 11, 2, 256, 1, 3, true, 8>("p21_lit_hi_32x44x8"),
        s1s<21, true, 32, 44, 256, 11, 2, 256, 1, 3, 2, 8>("p21_lit_q4_32x44x8"),
        // more strip kernels: 3 waves per SIMD, 32-column strips (the first 32-column one
        // the literal default for narrow blocks; xy kernel at 512^3: 0.83, 0.98 and 1.01 ms)
        s1strip<21, true, 4, 4, 3, 3, 256, 3>("p21_lit_strip_64w3"),
        s1strip<21, true, 2, 4, 3, 7, 256, 3>("p21_lit_strip_32w7"),
        s1strip<21, true, 2, 4, 1, 8, 256, 3>("p21_lit_strip_32w8"),
        // more tiles on the interleaved intermediate
        s1s<21, true, 32, 60, 256, 15, 2, 256, 1, 3, 1, 4, 1, 1, true>("p21_lit_ilhi_32x60r15"),
        s1s<21, true, 32, 28, 256, 7, 4, 256, 1, 3>("p21_lit_split_32x28w4"),

        s1s<21, true, 64, 28, 512, 7, 4, 256, 1, 3>("p21_lit_split_64x28w4"),
        s1v<21, false, 32, 16, 512, 4>("p21_fast_32x16w4"),
        s1v<21, false, 32, 32, 1024, 4>("p21_fast_32x32t1k"),
        s1v<21, false, 32, 32, 1024, 4, true>("p21_fast_32x32t1kdb"),
        s1v<21, true, 32, 16, 512>("p21_lit_32x16"),
        s1v<7, false, 32, 32, 512>("p7_fast_32x32"),
        s1v<7, true, 32, 16, 512>("p7_lit_32x16"),
        s1v<35, false, 32, 16, 512>("p35_fast_32x16"),
        s1v<35, true, 32, 16, 512>("p35_lit_32x16"),
        // every other odd tap count up to 35 (sigma up to 5.14): the fused kernel, untuned
        s1v<9, false, 32, 16, 512>("p9_fast_32x16"),
        s1v<9, true, 32, 16, 512>("p9_lit_32x16"),
        s1v<11, false, 32, 16, 512>("p11_fast_32x16"),
        s1v<11, true, 32, 16, 512>("p11_lit_32x16"),
        s1v<13, false, 32, 16, 512>("p13_fast_32x16"),
        s1v<13, true, 32, 16, 512>("p13_lit_32x16"),
        s1v<15, false, 32, 16, 512>("p15_fast_32x16"),
        s1v<15, true, 32, 16, 512>("p15_lit_32x16"),
        s1v<17, false, 32, 16, 512>("p17_fast_32x16"),
        s1v<17, true, 32, 16, 512>("p17_lit_32x16"),
        s1v<19, false, 32, 16, 512>("p19_fast_32x16"),
        s1v<19, true, 32, 16, 512>("p19_lit_32x16"),
        s1v<23, false, 32, 16, 512>("p23_fast_32x16"),
        s1v<23, true, 32, 16, 512>("p23_lit_32x16"),
        s1v<25, false, 32, 16, 512>("p25_fast_32x16"),
        s1v<25, true, 32, 16, 512>("p25_lit_32x16"),
        s1v<27, false, 32, 16, 512>("p27_fast_32x16"),
        s1v<27, true, 32, 16, 512>("p27_lit_32x16"),
        s1v<29, false, 32, 16, 512>("p29_fast_32x16"),
        s1v<29, true, 32, 16, 512>("p29_lit_32x16"),
        s1v<31, false, 32, 16, 512>("p31_fast_32x16"),
        s1v<31, true, 32, 16, 512>("p31_lit_32x16"),
        s1v<33, false, 32, 16, 512>("p33_fast_32x16"),
        s1v<33, true, 32, 16, 512>("p33_lit_32x16"),
    };
    return v;
}

const std::vector<Sweep2Variant>& sweep2_variants() {
    static const std::vector<Sweep2Variant> v = {
        // fast default (the first fast entry): the shared interior over whole rows at four
        // waves per SIMD (127 VGPRs), two 512-thread workgroups per CU: 0.51 ms per sweep 2
        // at 512^3 with its shell, against 0.55 on 32 x 8 quads and 0.70 for the vector
        // kernel on 64 x 4 quads
        s2h<false, 32, 16, 4>("fast_share_32x16w4"),
        s2h<false, 32, 8, 4>("fast_share_32x8w4"),
        s2q<false, 64, 4, 2>("fast_vec_64x4"),
        s2q<false, 32, 8, 2>("fast_vec_32x8"),
        s2q<false, 32, 4, 2>("fast_vec_32x4"),
        s2q<false, 16, 16, 2>("fast_vec_16x16"),
        s2q<false, 16, 8, 2>("fast_vec_16x8"),
        s2h<false, 32, 8, 2>("fast_share_32x8"),
        s2h<false, 16, 16, 4>("fast_share_16x16w4"),
        s2s<false, 64, 4>("fast_split_64x4"),
        s2s<false, 64, 2>("fast_split_64x2"),
        s2s<false, 32, 8>("fast_split_32x8"),
        s2v<false, 64, 4>("fast_64x4"),
        s2l<false, 32, 16, 4>("fast_lds_32x16"),
        // literal default (the first literal entry): the shared interior over whole rows
        // on 16 x 16 quads, three 256-thread workgroups per CU (159 VGPRs): 0.76 ms per
        // sweep 2 at 512^3 with its shell, against 0.81 for 32 x 16 quads (one 512-thread
        // workgroup per CU) and 1.12 for the vector kernel, whose normals are evaluated
        // 3.5 times per voxel
        s2h<true, 16, 16, 2>("lit_share_16x16"),
        s2h<true, 32, 16, 2>("lit_share_32x16"),
        s2q<true, 32, 8, 2>("lit_vec_32x8"),
        s2h<true, 32, 8, 2>("lit_share_32x8"),
        s2h<true, 32, 8, 3>("lit_share_32x8w3"),
        s2h<true, 16, 16, 3>("lit_share_16x16w3"),
        s2s<true, 64, 4>("lit_split_64x4"),
        s2v<true, 64, 4>("lit_64x4"),
    };
    return v;
}

int default_sweep1(int P, int literal) {
    const auto& v = sweep1_variants();
    for (size_t k = 0; k < v.size(); ++k)
        if (v[k].P == P && v[k].literal == literal) return (int)k;
    return -1;
}

int default_sweep2(int literal) {
    const auto& v = sweep2_variants();
    for (size_t k = 0; k < v.size(); ++k)
        if (v[k].literal == literal) return (int)k;
    return -1;
}

// Orders tiles so that the workgroups of one XCD (blockIdx = k mod 8 lands on
// XCD k) walk a contiguous spatial range and share halo lines in that XCD's L2.
static void xcd_order(std::vector<Tile>& t) {
    const size_t n = t.size();
    const size_t per = (n + 7) / 8;
    std::vector<Tile> out(per * 8, Tile{-1, 0, 0, 0});
    for (size_t s = 0; s < n; ++s) {
        const size_t xcd = s / per, j = s % per;
        out[j * 8 + xcd] = t[s];
    }
    t.swap(out);
}

int plan_tiles_host(TilePlan& tp, const std::vector<BlockDesc>& blocks, int s1, int s2, int zc1, int zc2,
                    int num_cus) {
    const auto& V1 = sweep1_variants();
    const auto& V2 = sweep2_variants();
    if (s1 < 0 || s1 >= (int)V1.size() || s2 < 0 || s2 >= (int)V2.size())
        return set_error(RSF_ERR_UNSUPPORTED, "no fused kernel variant for this problem");
    tp.s1 = s1;
    tp.s2 = s2;
    tp.zc1 = zc1;
    tp.zc2 = zc2;
    tp.blocks = blocks;
    tp.tiles1.clear();
    tp.tiles2.clear();
    tp.tiles3.clear();
    tp.tiles4.clear();
    tp.total_voxels = 0;
    const int tx1 = V1[s1].tx, ty1 = V1[s1].ty, tx2 = V2[s2].tx, ty2 = V2[s2].ty;
    tp.tiles5.clear();
    tp.extent = 0;
    const bool split1 = V1[s1].fn2 != nullptr;
    const int M1 = (V1[s1].P - 1) / 2;
    for (size_t b = 0; b < blocks.size(); ++b) {
        const BlockDesc& bd = blocks[b];
        tp.total_voxels += (int64_t)bd.w * bd.h * bd.l;
        tp.extent = std::max(tp.extent, bd.off + (int64_t)bd.w * bd.h * bd.l);
        if (split1) {
            // xy tiles over the valid region, z kernel columns over the whole block
            for (int z = 0; z < bd.l; z += zc1)
                for (int y = M1; y < bd.h - M1; y += ty1)
                    for (int x = M1; x < bd.w - M1; x += tx1) tp.tiles1.push_back(Tile{(int)b, x, y, z});
            const int64_t cols = (int64_t)bd.w * bd.h;
            for (int64_t c = 0; c < cols; c += V1[s1].tx2 * V1[s1].ty2)
                tp.tiles5.push_back(Tile{(int)b, (int)c, 0, 0});
        } else {
            for (int z = 0; z < bd.l; z += zc1)
                for (int y = 0; y < bd.h; y += ty1)
                    for (int x = 0; x < bd.w; x += tx1) tp.tiles1.push_back(Tile{(int)b, x, y, z});
        }
        if (V2[s2].shell_fn) {
            // vector interior when every row of the block starts on a 16-byte boundary
            const bool vec = V2[s2].vec == 4 && (bd.off & 3) == 0 && (bd.w & 3) == 0;
            int cols = 4;
            if (vec) {
                // the shared-interior kernel takes whole rows, the vector kernel x in [4, xend)
                const int xbeg = V2[s2].fullx ? 0 : 4;
                const int xend = V2[s2].fullx ? bd.w : 4 + 4 * ((bd.w - 8) >> 2);
                cols = V2[s2].fullx ? 0 : 4 + (bd.w - xend);
                for (int z = 2; z < bd.l - 2; z += zc2)
                    for (int y = 2; y < bd.h - 2; y += ty2)
                        for (int x = xbeg; x < xend; x += 4 * tx2) tp.tiles2.push_back(Tile{(int)b, x, y, z});
            } else {
                std::vector<Tile>& dst = V2[s2].vec == 4 ? tp.tiles4 : tp.tiles2;
                const int txs = V2[s2].vec == 4 ? 64 : tx2, tys = V2[s2].vec == 4 ? 4 : ty2;
                for (int z = 2; z < bd.l - 2; z += zc2)
                    for (int y = 2; y < bd.h - 2; y += tys)
                        for (int x = 2; x < bd.w - 2; x += txs) dst.push_back(Tile{(int)b, x, y, z});
            }
            const int64_t shell = 4 * (int64_t)bd.w * bd.h + 4 * (int64_t)bd.w * (bd.l - 4) +
                                  (int64_t)cols * (bd.h - 4) * (bd.l - 4);
            for (int64_t k = 0; k < shell; k += 256)
                tp.tiles3.push_back(
                    Tile{(int)b, (int)k, (int)std::min<int64_t>(256, shell - k), vec ? (V2[s2].fullx ? 2 : 1) : 0});
        } else {
            for (int z = 0; z < bd.l; z += zc2)
                for (int y = 0; y < bd.h; y += ty2)
                    for (int x = 0; x < bd.w; x += tx2) tp.tiles2.push_back(Tile{(int)b, x, y, z});
        }
    }
    xcd_order(tp.tiles1);
    xcd_order(tp.tiles2);
    xcd_order(tp.tiles4);
    // Persistent grids: a few workgroups per CU, a multiple of 8 (XCD count).
    auto grid_of = [&](size_t n, int per_cu) {
        size_t g = std::min(n, (size_t)num_cus * per_cu);
        g = std::max<size_t>(8, (g + 7) / 8 * 8);
        return (int)std::min(g, n);
    };
    tp.grid1 = grid_of(tp.tiles1.size(), 2);
    tp.grid2 = grid_of(tp.tiles2.size(), 8);
    tp.grid3 = tp.tiles3.empty() ? 0 : grid_of(tp.tiles3.size(), 8);
    tp.grid4 = tp.tiles4.empty() ? 0 : grid_of(tp.tiles4.size(), 8);
    if (split1) {
        tp.grid1 = tp.tiles1.empty() ? 0 : grid_of(tp.tiles1.size(), 4);
        tp.grid5 = tp.tiles5.empty() ? 0 : grid_of(tp.tiles5.size(), 8);
    }
    return RSF_OK;
}

namespace {
size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
size_t flags_bytes(const TilePlan& tp) {
    const Sweep2Variant& v = sweep2_variants()[tp.s2];
    return (v.fix_fn && !tp.tiles2.empty()) ? tp.tiles2.size() * (size_t)(v.tx * v.ty / 64) * sizeof(int) : 0;
}
size_t xy_bytes(const TilePlan& tp) {
    const Sweep1Variant& v = sweep1_variants()[tp.s1];
    return v.fn2 ? (size_t)tp.extent * (v.literal ? 4 : 2) * sizeof(float) : 0;
}
}  // namespace

size_t plan_arena_bytes(const TilePlan& tp) {
    return align256(tp.tiles1.size() * sizeof(Tile)) + align256(tp.tiles2.size() * sizeof(Tile)) +
           align256(tp.tiles3.size() * sizeof(Tile)) + align256(tp.tiles4.size() * sizeof(Tile)) +
           align256(tp.tiles5.size() * sizeof(Tile)) + align256(tp.blocks.size() * sizeof(BlockDesc)) +
           align256(flags_bytes(tp)) + align256(xy_bytes(tp));
}

int plan_upload(TilePlan& tp, unsigned char* arena, size_t arena_bytes, hipStream_t s) {
    const size_t nf = flags_bytes(tp), nxy = xy_bytes(tp);
    struct Part {
        DevBuf* buf;
        const void* src;
        size_t bytes;
    };
    const Part parts[] = {
        {&tp.d_tiles1, tp.tiles1.data(), tp.tiles1.size() * sizeof(Tile)},
        {&tp.d_tiles2, tp.tiles2.data(), tp.tiles2.size() * sizeof(Tile)},
        {&tp.d_tiles3, tp.tiles3.data(), tp.tiles3.size() * sizeof(Tile)},
        {&tp.d_tiles4, tp.tiles4.data(), tp.tiles4.size() * sizeof(Tile)},
        {&tp.d_tiles5, tp.tiles5.data(), tp.tiles5.size() * sizeof(Tile)},
        {&tp.d_blocks, tp.blocks.data(), tp.blocks.size() * sizeof(BlockDesc)},
        {&tp.d_flags2, nullptr, nf},
        {&tp.xy, nullptr, nxy},
    };
    if (arena) {
        // every buffer a view of the caller's arena; copies and clears on stream s
        if (plan_arena_bytes(tp) > arena_bytes)
            return set_error(RSF_ERR_INVALID_ARGUMENT, "tile plan needs %zu bytes of scratch, %zu available",
                             plan_arena_bytes(tp), arena_bytes);
        size_t off = 0;
        for (const Part& q : parts) {
            if (q.bytes == 0) {
                q.buf->release();
                continue;
            }
            q.buf->bind(arena + off, q.bytes);
            if (q.src) RSF_HIP(hipMemcpyAsync(q.buf->p, q.src, q.bytes, hipMemcpyHostToDevice, s));
            else if (q.buf == &tp.d_flags2) RSF_HIP(hipMemsetAsync(q.buf->p, 0, q.bytes, s));
            off += align256(q.bytes);
        }
        return RSF_OK;
    }
    // owned buffers (grow-only), synchronous copies
    for (const Part& q : parts) {
        if (q.bytes == 0) continue;
        RSF_TRY(q.buf->ensure(q.bytes));
        if (q.src) RSF_HIP(hipMemcpy(q.buf->p, q.src, q.bytes, hipMemcpyHostToDevice));
    }
    // the interior rewrites its flags on every launch; zero them for a clean start
    if (nf) RSF_HIP(hipMemset(tp.d_flags2.p, 0, nf));
    return RSF_OK;
}

int plan_tiles(TilePlan& tp, const std::vector<BlockDesc>& blocks, int s1, int s2, int zc1, int zc2,
               int num_cus) {
    RSF_TRY(plan_tiles_host(tp, blocks, s1, s2, zc1, zc2, num_cus));
    return plan_upload(tp, nullptr, 0, nullptr);
}

int launch_sweep1(const TilePlan& tp, const SweepArgs& a, const Derived& d, const rsf_params* p,
                  hipStream_t s) {
    const Sweep1Variant& v = sweep1_variants()[tp.s1];
    if (v.P != d.P) return set_error(RSF_ERR_INVALID_ARGUMENT, "sweep-1 variant built for P=%d, run has P=%d", v.P, d.P);
    const int ntiles = (int)tp.tiles1.size();
    const int same = (p->lambda1 == p->lambda2) ? 1 : 0;
    if (v.fn2) {
        if (a.img_u8 != 0 && a.img_u8 != 1) return set_error(RSF_ERR_INVALID_ARGUMENT, "image dtype");
        float* xy = tp.xy.as<float>();
        const int64_t stride = tp.extent;
        if (ntiles > 0) {
            void* a1[] = {(void*)&a.tr, (void*)&a.img, (void*)&a.img_u8, (void*)&xy, (void*)&stride,
                          (void*)&tp.d_blocks.p, (void*)&tp.d_tiles1.p, (void*)&ntiles, (void*)&tp.zc1,
                          (void*)&d.taps};
            const void* f = v.fn_img[a.img_u8] ? v.fn_img[a.img_u8] : v.fn;
            RSF_HIP(hipLaunchKernel(f, dim3(tp.grid1), dim3(v.nt), a1, 0, s));
        }
        const int n5 = (int)tp.tiles5.size();
        if (a.out_fits < 0 || a.out_fits > 2) return set_error(RSF_ERR_INVALID_ARGUMENT, "sweep-1 output mode");
        void* a2[] = {(void*)&xy,       (void*)&stride,   (void*)&a.img,    (void*)&a.blurI,
                      (void*)&a.out0,   (void*)&a.out1,   (void*)&a.out2,   (void*)&a.out3,
                      (void*)&tp.d_blocks.p, (void*)&tp.d_tiles5.p, (void*)&n5, (void*)&d.taps,
                      (void*)&d.halfS3, (void*)&p->lambda1, (void*)&p->lambda2, (void*)&same};
        const int zmode = a.out_fits == 0 && !same ? 3 : a.out_fits;
        RSF_HIP(hipLaunchKernel(v.fn2s[2 * zmode + a.img_u8], dim3(tp.grid5), dim3(v.nt2), a2, 0, s));
        return RSF_OK;
    }
    void* args[] = {(void*)&a.tr,       (void*)&a.img,     (void*)&a.img_u8, (void*)&a.blurI,
                    (void*)&a.out0,     (void*)&a.out1,    (void*)&a.out2,   (void*)&a.out3,
                    (void*)&a.out_fits, (void*)&tp.d_blocks.p,
                    (void*)&tp.d_tiles1.p, (void*)&ntiles, (void*)&tp.zc1,   (void*)&d.taps,
                    (void*)&d.halfS3,   (void*)&p->lambda1, (void*)&p->lambda2, (void*)&same};
    RSF_HIP(hipLaunchKernel(v.fn, dim3(tp.grid1), dim3(v.nt), args, 0, s));
    return RSF_OK;
}

int launch_blur_image(const TilePlan& tp, const void* img, int u8, float* tr, float* out, const Derived& d,
                      hipStream_t s) {
    const Sweep1Variant& v = sweep1_variants()[tp.s1];
    if (!v.fn2 || v.P != d.P || v.literal) return set_error(RSF_ERR_UNSUPPORTED, "image blur needs a split fast plan");
    const int ntiles = (int)tp.tiles1.size(), n5 = (int)tp.tiles5.size();
    float* xy = tp.xy.as<float>();
    const int64_t stride = tp.extent;
    const int zero = 0;
    const float fzero = 0.f;
    const float* null = nullptr;
    float* nullo = nullptr;
    if (ntiles > 0) {
        const void* fn = v.fn_blur[u8];
        if (!fn) {
            // a variant without the image-only xy kernel blurs (I*1, 1): channel 0 is the image
            RSF_HIP(hipMemsetD32Async((hipDeviceptr_t)tr, 0x3f800000, (size_t)tp.extent, s));
            fn = v.fn_img[u8] ? v.fn_img[u8] : v.fn;
        }
        void* a1[] = {(void*)&tr, (void*)&img, (void*)&u8, (void*)&xy, (void*)&stride, (void*)&tp.d_blocks.p,
                      (void*)&tp.d_tiles1.p, (void*)&ntiles, (void*)&tp.zc1, (void*)&d.taps};
        RSF_HIP(hipLaunchKernel(fn, dim3(tp.grid1), dim3(v.nt), a1, 0, s));
    }
    void* a2[] = {(void*)&xy,    (void*)&stride, (void*)&img,   (void*)&null,  (void*)&out,
                  (void*)&nullo, (void*)&nullo,  (void*)&nullo, (void*)&tp.d_blocks.p, (void*)&tp.d_tiles5.p,
                  (void*)&n5,    (void*)&d.taps, (void*)&fzero, (void*)&fzero, (void*)&fzero, (void*)&zero};
    RSF_HIP(hipLaunchKernel(v.fn2s[8 + u8], dim3(tp.grid5), dim3(v.tx2 * v.ty2), a2, 0, s));
    return RSF_OK;
}

// The reciprocal argument of the sweep-2 and transform kernels.  Fast mode: 1/eps.
// Literal mode: 1/eps when eps is a power of two whose reciprocal is a normal
// float (the kernels then multiply instead of dividing, with the same bits),
// else 0 (they divide).
static float eps_reciprocal(float eps, bool literal) {
    if (!literal) return 1.0f / eps;
    int e = 0;
    const float m = std::frexp(eps, &e);   // eps = m * 2^e, m in [0.5, 1)
    // eps = 2^(e-1): eps and 1/eps = 2^(1-e) both normal for e in [-125, 127]
    return (std::isfinite(eps) && m == 0.5f && e >= -125 && e <= 127) ? 1.0f / eps : 0.0f;
}

int launch_sweep2(const TilePlan& tp, const SweepArgs& a, const rsf_params* p, hipStream_t s) {
    const Sweep2Variant& v = sweep2_variants()[tp.s2];
    const int ntiles = (int)tp.tiles2.size();
    const float inv_eps = eps_reciprocal(p->epsilon, p->mode == RSF_MODE_LITERAL);
    void* args[] = {(void*)&a.phi,  (void*)&a.F,     (void*)&a.phi_out, (void*)&a.tr_out,
                    (void*)&tp.d_blocks.p, (void*)&tp.d_tiles2.p, (void*)&ntiles, (void*)&tp.zc2,
                    (void*)&p->dt,  (void*)&p->nu,   (void*)&p->mu,     (void*)&p->epsilon,
                    (void*)&inv_eps, (void*)&tp.d_flags2.p};   // (kernels without flags ignore the last)
    if (ntiles > 0) RSF_HIP(hipLaunchKernel(v.fn, dim3(tp.grid2), dim3(v.tx * v.ty), args, 0, s));
    if (ntiles > 0 && v.fix_fn) RSF_HIP(hipLaunchKernel(v.fix_fn, dim3(tp.grid2), dim3(v.tx * v.ty), args, 0, s));
    if (v.scalar_fn && tp.grid4 > 0) {
        const int n4 = (int)tp.tiles4.size();
        void* a4[] = {(void*)&a.phi,  (void*)&a.F,     (void*)&a.phi_out, (void*)&a.tr_out,
                      (void*)&tp.d_blocks.p, (void*)&tp.d_tiles4.p, (void*)&n4, (void*)&tp.zc2,
                      (void*)&p->dt,  (void*)&p->nu,   (void*)&p->mu,     (void*)&p->epsilon,
                      (void*)&inv_eps};
        RSF_HIP(hipLaunchKernel(v.scalar_fn, dim3(tp.grid4), dim3(256), a4, 0, s));
    }
    if (v.shell_fn && tp.grid3 > 0) {
        const int nitems = (int)tp.tiles3.size();
        void* sargs[] = {(void*)&a.phi,  (void*)&a.F,      (void*)&a.phi_out, (void*)&a.tr_out,
                         (void*)&tp.d_blocks.p, (void*)&tp.d_tiles3.p, (void*)&nitems, (void*)&p->dt,
                         (void*)&p->nu, (void*)&p->mu, (void*)&p->epsilon, (void*)&inv_eps};
        RSF_HIP(hipLaunchKernel(v.shell_fn, dim3(tp.grid3), dim3(256), sargs, 0, s));
    }
    return RSF_OK;
}

int launch_transform(float* tr, const float* phi, int64_t n, int literal, float eps, hipStream_t s) {
    if (n <= 0) return RSF_OK;
    const bool v4 = (((uintptr_t)tr | (uintptr_t)phi) & 15) == 0;
    const int64_t g = std::max<int64_t>(1, std::min<int64_t>((v4 ? n / 4 + 255 : n + 255) / 256, 1 << 20));
    const void* fn = literal ? (v4 ? (const void*)&transform_kernel<true, true> : (const void*)&transform_kernel<true, false>)
                             : (v4 ? (const void*)&transform_kernel<false, true> : (const void*)&transform_kernel<false, false>);
    const float inv_eps = eps_reciprocal(eps, literal != 0);
    void* args[] = {(void*)&tr, (void*)&phi, (void*)&n, (void*)&eps, (void*)&inv_eps};
    RSF_HIP(hipLaunchKernel(fn, dim3((unsigned)g), dim3(256), args, 0, s));
    RSF_HIP(hipGetLastError());
    return RSF_OK;
}

}  // namespace rsf
