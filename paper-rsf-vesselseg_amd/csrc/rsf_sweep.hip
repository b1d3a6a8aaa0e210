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
#include <cstdio>

#include "rsf_internal.hpp"

namespace rsf {
namespace {

template <int P, bool LIT, int TX, int TY, int NT>
__global__ __launch_bounds__(NT) void sweep1_kernel(
    const float* __restrict__ tr, const void* __restrict__ img, int img_u8,
    const float* __restrict__ blurI, float* __restrict__ out0, float* __restrict__ out1,
    float* __restrict__ out2, float* __restrict__ out3, int out_mode, const BlockDesc* __restrict__ blocks, const Tile* __restrict__ tiles,
    int ntiles, int zc, Taps taps, float halfS3, float l1, float l2, int same_lambda) {
    constexpr int CH = LIT ? 4 : 2;
    constexpr int M = (P - 1) / 2;
    constexpr int RW = TX + 2 * M;
    constexpr int RH = TY + 2 * M;
    constexpr int RWP = (RW + 3) & ~3;
    constexpr int COLS = (TX * TY) / NT;
    static_assert(COLS * NT == TX * TY, "tile must be a multiple of the block");
    static_assert(TX % 4 == 0, "TX must be a multiple of 4");
    constexpr int RX = 4;
    constexpr int XG = TX / RX;
    constexpr int NC = COLS * CH;
    constexpr int WIN = (RX + P - 1 + 3) & ~3;  // x-pass window rounded to float4

    __shared__ __align__(16) float sIn[CH][RH][RWP];
    __shared__ __align__(16) float sX[CH][RH][TX];

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
        const int x0 = tl.x0, y0 = tl.y0, z0 = tl.z0;
        const int zend = min(z0 + zc, l);
        const int nplanes = (zend - z0) + P - 1;

        float acc[NC][P];
#pragma unroll
        for (int n = 0; n < NC; ++n)
#pragma unroll
            for (int k = 0; k < P; ++k) acc[n][k] = 0.f;

        for (int t = 0; t < nplanes; ++t) {
            const int zi = min(max(z0 - M + t, 0), l - 1);
            const int64_t zoff = off + (int64_t)zi * h * w;
            // ---- stage A: transformed level set and image -> channels in LDS
            for (int e = tid; e < RH * RW; e += NT) {
                const int r = e / RW, c = e - r * RW;
                const int gx = min(max(x0 - M + c, 0), w - 1);
                const int gy = min(max(y0 - M + r, 0), h - 1);
                const int64_t idx = zoff + (int64_t)gy * w + gx;
                const float tv = tr[idx];
                const float iv = img_u8 ? (float)img8[idx] : imgf[idx];
                if constexpr (LIT) {
                    const float hc = 1.0f - tv;        // kernel.cu:295
                    sIn[0][r][c] = iv * tv;            // I*H      (main.cpp:238)
                    sIn[1][r][c] = tv;                 // H
                    sIn[2][r][c] = iv * hc;            // I*(1-H)  (main.cpp:246)
                    sIn[3][r][c] = hc;                 // 1-H
                } else {
                    sIn[0][r][c] = iv * tv;            // I*G
                    sIn[1][r][c] = tv;                 // G
                }
            }
            __syncthreads();
            // ---- stage B: x pass, RX consecutive outputs per work item
            for (int e = tid; e < CH * RH * XG; e += NT) {
                const int ch = e / (RH * XG);
                const int rem = e - ch * (RH * XG);
                const int r = rem / XG;
                const int xb = (rem - r * XG) * RX;
                float win[WIN];
                const float4* src = reinterpret_cast<const float4*>(&sIn[ch][r][xb]);
#pragma unroll
                for (int q = 0; q < WIN / 4; ++q) {
                    const float4 v = src[q];
                    win[4 * q + 0] = v.x;
                    win[4 * q + 1] = v.y;
                    win[4 * q + 2] = v.z;
                    win[4 * q + 3] = v.w;
                }
                float o4[RX];
#pragma unroll
                for (int o = 0; o < RX; ++o) {
                    float a = win[o] * taps.g[0];
#pragma unroll
                    for (int k = 1; k < P; ++k) a = fmaf(win[o + k], taps.g[k], a);
                    o4[o] = a;
                }
                *reinterpret_cast<float4*>(&sX[ch][r][xb]) = make_float4(o4[0], o4[1], o4[2], o4[3]);
            }
            __syncthreads();
            // ---- stage C: y pass for this thread's COLS rows
            float v[NC];
#pragma unroll
            for (int ch = 0; ch < CH; ++ch) {
                float col[COLS + P - 1];
#pragma unroll
                for (int j = 0; j < COLS + P - 1; ++j) col[j] = sX[ch][ty0 + j][tx];
#pragma unroll
                for (int cc = 0; cc < COLS; ++cc) {
                    float a = col[cc] * taps.g[0];
#pragma unroll
                    for (int k = 1; k < P; ++k) a = fmaf(col[cc + k], taps.g[k], a);
                    v[cc * CH + ch] = a;
                }
            }
            // ---- stage D: z ring
            float done[NC];
            zdispatch<P, NC>(t % P, acc, v, taps, done, std::make_integer_sequence<int, P>{});
            if (t < P - 1) continue;
            // ---- emission of output plane zo
            const int zo = z0 + t - (P - 1);
            const bool zvalid = zo >= M && zo < l - M;
#pragma unroll
            for (int cc = 0; cc < COLS; ++cc) {
                const int x = x0 + tx, y = y0 + ty0 + cc;
                if (x >= w || y >= h) continue;
                const int64_t i = off + ((int64_t)zo * h + y) * w + x;
                const float Iv = img_u8 ? (float)img8[i] : imgf[i];
                float fo = 1.0f, fi = 1.0f;  // pad band: 30/30 (kernel.cu:313, main.cpp:277)
                float bH = 30.0f, bHc = 30.0f;
                if (zvalid && x >= M && x < w - M && y >= M && y < h - M) {
                    if constexpr (LIT) {
                        bH = done[cc * CH + 1];
                        bHc = done[cc * CH + 3];
                        fo = done[cc * CH + 0] / bH;
                        fi = done[cc * CH + 2] / bHc;
                    } else {
                        const float bI = blurI[i];
                        const float bIG = done[cc * CH + 0], bG = done[cc * CH + 1];
                        bH = halfS3 + bG;
                        bHc = halfS3 - bG;
                        fo = fmaf(0.5f, bI, bIG) / bH;
                        fi = fmaf(0.5f, bI, -bIG) / bHc;
                    }
                }
                if (out_mode == 0) {
                    const float d1 = Iv - fo, d2 = Iv - fi;
                    const float Eo = d1 * d1, Ei = d2 * d2;  // sigma2 = 0 (kernel.cu:44-45)
                    out0[i] = same_lambda ? l1 * (Eo - Ei) : (l1 * Eo - l2 * Ei);
                } else {
                    out0[i] = fo;
                    out1[i] = fi;
                    if (out_mode == 2) {
                        out2[i] = bH;
                        out3[i] = bHc;
                    }
                }
            }
        }
        __syncthreads();
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
        for (int z = tl.z0; z < zend; ++z) {
            const int64_t i = ((int64_t)z * h + y) * w + x;
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
                    const float M = sqrtf(g.x * g.x + g.y * g.y + g.z * g.z);
                    return (float)((double)c / ((double)M + 1e-6));
                };
                nRx = nrm(gRx, gRx.x); nLx = nrm(gLx, gLx.x);
                nRy = nrm(gRy, gRy.y); nLy = nrm(gLy, gLy.y);
                nRz = nrm(gRz, gRz.z); nLz = nrm(gLz, gLz.z);
            } else {
                auto nrm = [](const Grad& g, float c) {
                    const float M = __builtin_amdgcn_sqrtf(g.x * g.x + g.y * g.y + g.z * g.z);
                    return c * __builtin_amdgcn_rcpf(M + 1e-6f);
                };
                nRx = nrm(gRx, gRx.x); nLx = nrm(gLx, gLx.x);
                nRy = nrm(gRy, gRy.y); nLy = nrm(gLy, gLy.y);
                nRz = nrm(gRz, gRz.z); nLz = nrm(gLz, gLz.z);
            }
            const float kap = ((nRx - nLx) * 0.5f + (nRy - nLy) * 0.5f) + (nRz - nLz) * 0.5f;
            const float lap = ((gRx.x - gLx.x) * 0.5f + (gRy.y - gLy.y) * 0.5f) + (gRz.z - gLz.z) * 0.5f;
            const float p0 = f[i];
            const float d = LIT ? dirac_literal(p0, eps) : dirac_fast(p0, eps);
            const float Fv = F[bd.off + i];
            const float p1 = ((p0 - (d * Fv) * dt) + ((nu * dt) * d) * kap) + mu * (dt * (lap - kap));
            phi_out[bd.off + i] = p1;
            tr_out[bd.off + i] = LIT ? heav_literal(p1, eps) : heav_centred(p1, inv_eps);
        }
    }
}

__global__ void transform_kernel(float* tr, const float* phi, int64_t n, int lit, float eps, float inv_eps) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        tr[i] = lit ? heav_literal(phi[i], eps) : heav_centred(phi[i], inv_eps);
}

template <int P, bool LIT, int TX, int TY, int NT>
Sweep1Variant s1v(const char* name) {
    return Sweep1Variant{P, LIT ? 1 : 0, TX, TY, NT, (const void*)&sweep1_kernel<P, LIT, TX, TY, NT>, name};
}
template <bool LIT, int TX, int TY>
Sweep2Variant s2v(const char* name) {
    return Sweep2Variant{LIT ? 1 : 0, TX, TY, (const void*)&sweep2_kernel<LIT, TX, TY>, name};
}

}  // namespace

const std::vector<Sweep1Variant>& sweep1_variants() {
    static const std::vector<Sweep1Variant> v = {
        s1v<21, false, 32, 32, 512>("p21_fast_32x32"),
        s1v<21, false, 64, 16, 512>("p21_fast_64x16"),
        s1v<21, false, 32, 16, 512>("p21_fast_32x16"),
        s1v<21, true, 32, 16, 512>("p21_lit_32x16"),
        s1v<7, false, 32, 32, 512>("p7_fast_32x32"),
        s1v<7, true, 32, 16, 512>("p7_lit_32x16"),
        s1v<35, false, 32, 16, 512>("p35_fast_32x16"),
        s1v<35, true, 32, 16, 512>("p35_lit_32x16"),
    };
    return v;
}

const std::vector<Sweep2Variant>& sweep2_variants() {
    static const std::vector<Sweep2Variant> v = {
        s2v<false, 64, 4>("fast_64x4"),
        s2v<false, 32, 8>("fast_32x8"),
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

int plan_tiles(TilePlan& tp, const std::vector<BlockDesc>& blocks, int s1, int s2, int zc1, int zc2,
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
    tp.total_voxels = 0;
    const int tx1 = V1[s1].tx, ty1 = V1[s1].ty, tx2 = V2[s2].tx, ty2 = V2[s2].ty;
    for (size_t b = 0; b < blocks.size(); ++b) {
        const BlockDesc& bd = blocks[b];
        tp.total_voxels += (int64_t)bd.w * bd.h * bd.l;
        for (int z = 0; z < bd.l; z += zc1)
            for (int y = 0; y < bd.h; y += ty1)
                for (int x = 0; x < bd.w; x += tx1) tp.tiles1.push_back(Tile{(int)b, x, y, z});
        for (int z = 0; z < bd.l; z += zc2)
            for (int y = 0; y < bd.h; y += ty2)
                for (int x = 0; x < bd.w; x += tx2) tp.tiles2.push_back(Tile{(int)b, x, y, z});
    }
    xcd_order(tp.tiles1);
    xcd_order(tp.tiles2);
    // Persistent grids: a few workgroups per CU, a multiple of 8 (XCD count).
    auto grid_of = [&](size_t n, int per_cu) {
        size_t g = std::min(n, (size_t)num_cus * per_cu);
        g = std::max<size_t>(8, (g + 7) / 8 * 8);
        return (int)std::min(g, n);
    };
    tp.grid1 = grid_of(tp.tiles1.size(), 2);
    tp.grid2 = grid_of(tp.tiles2.size(), 8);
    RSF_TRY(tp.d_tiles1.ensure(tp.tiles1.size() * sizeof(Tile)));
    RSF_TRY(tp.d_tiles2.ensure(tp.tiles2.size() * sizeof(Tile)));
    RSF_TRY(tp.d_blocks.ensure(blocks.size() * sizeof(BlockDesc)));
    RSF_HIP(hipMemcpy(tp.d_tiles1.p, tp.tiles1.data(), tp.tiles1.size() * sizeof(Tile), hipMemcpyHostToDevice));
    RSF_HIP(hipMemcpy(tp.d_tiles2.p, tp.tiles2.data(), tp.tiles2.size() * sizeof(Tile), hipMemcpyHostToDevice));
    RSF_HIP(hipMemcpy(tp.d_blocks.p, blocks.data(), blocks.size() * sizeof(BlockDesc), hipMemcpyHostToDevice));
    return RSF_OK;
}

int launch_sweep1(const TilePlan& tp, const SweepArgs& a, const Derived& d, const rsf_params* p,
                  hipStream_t s) {
    const Sweep1Variant& v = sweep1_variants()[tp.s1];
    if (v.P != d.P) return set_error(RSF_ERR_INVALID_ARGUMENT, "sweep-1 variant built for P=%d, run has P=%d", v.P, d.P);
    const int ntiles = (int)tp.tiles1.size();
    const int same = (p->lambda1 == p->lambda2) ? 1 : 0;
    void* args[] = {(void*)&a.tr,       (void*)&a.img,     (void*)&a.img_u8, (void*)&a.blurI,
                    (void*)&a.out0,     (void*)&a.out1,    (void*)&a.out2,   (void*)&a.out3,
                    (void*)&a.out_fits, (void*)&tp.d_blocks.p,
                    (void*)&tp.d_tiles1.p, (void*)&ntiles, (void*)&tp.zc1,   (void*)&d.taps,
                    (void*)&d.halfS3,   (void*)&p->lambda1, (void*)&p->lambda2, (void*)&same};
    RSF_HIP(hipLaunchKernel(v.fn, dim3(tp.grid1), dim3(v.nt), args, 0, s));
    return RSF_OK;
}

int launch_sweep2(const TilePlan& tp, const SweepArgs& a, const rsf_params* p, hipStream_t s) {
    const Sweep2Variant& v = sweep2_variants()[tp.s2];
    const int ntiles = (int)tp.tiles2.size();
    const float inv_eps = 1.0f / p->epsilon;
    void* args[] = {(void*)&a.phi,  (void*)&a.F,     (void*)&a.phi_out, (void*)&a.tr_out,
                    (void*)&tp.d_blocks.p, (void*)&tp.d_tiles2.p, (void*)&ntiles, (void*)&tp.zc2,
                    (void*)&p->dt,  (void*)&p->nu,   (void*)&p->mu,     (void*)&p->epsilon,
                    (void*)&inv_eps};
    RSF_HIP(hipLaunchKernel(v.fn, dim3(tp.grid2), dim3(v.tx * v.ty), args, 0, s));
    return RSF_OK;
}

int launch_transform(float* tr, const float* phi, int64_t n, int literal, float eps, hipStream_t s) {
    if (n <= 0) return RSF_OK;
    int64_t g = std::min<int64_t>((n + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(transform_kernel, dim3((unsigned)g), dim3(256), 0, s, tr, phi, n, literal, eps, 1.0f / eps);
    RSF_HIP(hipGetLastError());
    return RSF_OK;
}

}  // namespace rsf
