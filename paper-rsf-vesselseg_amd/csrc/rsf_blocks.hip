// Block pipeline of the reference's gigavoxel workflow (README.md:6-14,56-59):
// a volume is cut into overlapping blocks, every block is evolved on its own
// (one rsf3d process per block in the reference, one batched launch here), and
// the blocks' phi are blended back over the overlaps with linear ramps
// (interpolating_volumes.py:10-40, generalised from two halves along y to a
// tensor-product of ramps along every axis).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "rsf_internal.hpp"

namespace rsf {
namespace {

// Blocks of one axis: starts j*S (S = B - overlap) and size B, the last one
// clipped to the volume; a last block smaller than max(min_size, 2*overlap) is
// folded into the one before it.  Every coordinate is covered by one block or,
// inside an overlap, by two consecutive ones.
std::vector<std::pair<int64_t, int64_t>> axis_grid(int64_t n, int64_t B, int64_t k, int64_t mn) {
    std::vector<std::pair<int64_t, int64_t>> g;
    if (B >= n) {
        g.push_back({0, n});
        return g;
    }
    const int64_t S = B - k;
    for (int64_t s = 0;; s += S) {
        if (s + B >= n) {
            g.push_back({s, n - s});
            break;
        }
        g.push_back({s, B});
    }
    if (g.size() >= 2 && g.back().second < std::max(mn, 2 * k)) {
        g.pop_back();
        g.back().second = n - g.back().first;
    }
    return g;
}

struct AxisDesc {     // the grid of one axis as the kernels see it
    int64_t step;     // S = B - overlap (the starts are j * S)
    int nblk;         // blocks along the axis
};

struct GridDesc {
    AxisDesc ax[3];   // x, y, z
    int64_t overlap;
    double astep;     // numpy.linspace(0, 1, overlap) step
};

// Blocks covering coordinate c of one axis and their weights: block j with
// weight 1, or blocks j-1 and j with 1 - alpha_i and alpha_i in the overlap,
// alpha = numpy.linspace(0, 1, k) evaluated as numpy does (i * step, last 1).
__device__ __forceinline__ int axis_cover(int64_t c, const AxisDesc& a, int64_t k, double astep, int* j,
                                          double* wgt) {
#pragma clang fp contract(off)
    int jh = (int)min<int64_t>(c / a.step, a.nblk - 1);
    const int64_t i = c - (int64_t)jh * a.step;
    if (jh >= 1 && i < k) {
        const double alpha = (k > 1 && i == k - 1) ? 1.0 : (double)i * astep;
        j[0] = jh - 1;
        wgt[0] = 1.0 - alpha;
        j[1] = jh;
        wgt[1] = alpha;
        return 2;
    }
    j[0] = jh;
    wgt[0] = 1.0;
    return 1;
}

__global__ void k_disassemble(void* __restrict__ packed, const void* __restrict__ vol, int esz, int64_t w, int64_t h,
                              const rsf_block* __restrict__ blocks, const int64_t* __restrict__ origins, int nblocks,
                              int64_t total) {
    // one thread per packed voxel; the block is found by a binary search on the
    // (ascending, checked on the host) offsets
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < total;
         p += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = nblocks - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (blocks[mid].offset <= p) lo = mid;
            else hi = mid - 1;
        }
        const rsf_block b = blocks[lo];
        const int64_t q = p - b.offset;
        if (q >= b.w * b.h * b.l) continue;  // a gap between packed blocks
        const int64_t bx = q % b.w, by = (q / b.w) % b.h, bz = q / (b.w * b.h);
        const int64_t src = ((origins[3 * lo + 2] + bz) * h + (origins[3 * lo + 1] + by)) * w + origins[3 * lo] + bx;
        if (esz == 1) ((uint8_t*)packed)[p] = ((const uint8_t*)vol)[src];
        else ((float*)packed)[p] = ((const float*)vol)[src];
    }
}

__global__ void k_reassemble(float* __restrict__ vol, const float* __restrict__ packed, int64_t w, int64_t h,
                             int64_t l, const rsf_block* __restrict__ blocks, const int64_t* __restrict__ origins,
                             GridDesc g) {
#pragma clang fp contract(off)
    const int64_t n = w * h * l;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t x = v % w, y = (v / w) % h, z = v / (w * h);
        int jx[2], jy[2], jz[2];
        double wx[2], wy[2], wz[2];
        const int nx = axis_cover(x, g.ax[0], g.overlap, g.astep, jx, wx);
        const int ny = axis_cover(y, g.ax[1], g.overlap, g.astep, jy, wy);
        const int nz = axis_cover(z, g.ax[2], g.overlap, g.astep, jz, wz);
        double acc = 0.0;
        for (int a = 0; a < nz; ++a)
            for (int b = 0; b < ny; ++b)
                for (int c = 0; c < nx; ++c) {
                    const int bi = (jz[a] * g.ax[1].nblk + jy[b]) * g.ax[0].nblk + jx[c];
                    const rsf_block blk = blocks[bi];
                    const int64_t lx = x - origins[3 * bi], ly = y - origins[3 * bi + 1],
                                  lz = z - origins[3 * bi + 2];
                    const double f = (double)packed[blk.offset + (lz * blk.h + ly) * blk.w + lx];
                    acc = acc + ((wz[a] * wy[b]) * wx[c]) * f;
                }
        vol[v] = (float)acc;
    }
}

int grid_of(int64_t w, int64_t h, int64_t l, int64_t bw, int64_t bh, int64_t bl, int64_t overlap, int64_t min_size,
            std::vector<rsf_block>* blocks, std::vector<int64_t>* origins, GridDesc* gd, int64_t* total) {
    if (w <= 0 || h <= 0 || l <= 0) return set_error(RSF_ERR_INVALID_ARGUMENT, "volume dims must be positive");
    if (overlap < 0) return set_error(RSF_ERR_INVALID_ARGUMENT, "overlap must be >= 0");
    const int64_t B[3] = {bw, bh, bl}, N[3] = {w, h, l};
    std::vector<std::pair<int64_t, int64_t>> g[3];
    for (int a = 0; a < 3; ++a) {
        if (B[a] < N[a] && B[a] < 2 * overlap + 1)
            return set_error(RSF_ERR_INVALID_ARGUMENT, "block size %lld must exceed twice the overlap %lld",
                             (long long)B[a], (long long)overlap);
        if (B[a] <= 0) return set_error(RSF_ERR_INVALID_ARGUMENT, "block dims must be positive");
        g[a] = axis_grid(N[a], B[a], overlap, min_size);
        gd->ax[a].step = B[a] >= N[a] ? N[a] : B[a] - overlap;
        gd->ax[a].nblk = (int)g[a].size();
    }
    if ((int64_t)g[0].size() * (int64_t)g[1].size() * (int64_t)g[2].size() > (int64_t(1) << 24))
        return set_error(RSF_ERR_INVALID_ARGUMENT, "more than 2^24 blocks: choose larger blocks");
    gd->overlap = overlap;
    gd->astep = overlap > 1 ? 1.0 / (double)(overlap - 1) : 0.0;
    blocks->clear();
    origins->clear();
    int64_t off = 0;
    for (auto& zz : g[2])
        for (auto& yy : g[1])
            for (auto& xx : g[0]) {
                blocks->push_back(rsf_block{off, xx.second, yy.second, zz.second});
                origins->push_back(xx.first);
                origins->push_back(yy.first);
                origins->push_back(zz.first);
                off += xx.second * yy.second * zz.second;
            }
    *total = off;
    return RSF_OK;
}

// Device copies of a grid's tables.
struct DevGrid {
    DevBuf blocks, origins;
    int upload(const std::vector<rsf_block>& b, const std::vector<int64_t>& o, hipStream_t s) {
        RSF_TRY(blocks.ensure(b.size() * sizeof(rsf_block)));
        RSF_TRY(origins.ensure(o.size() * sizeof(int64_t)));
        RSF_HIP(hipMemcpyAsync(blocks.p, b.data(), b.size() * sizeof(rsf_block), hipMemcpyHostToDevice, s));
        RSF_HIP(hipMemcpyAsync(origins.p, o.data(), o.size() * sizeof(int64_t), hipMemcpyHostToDevice, s));
        return RSF_OK;
    }
};

unsigned grid1d(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 1 << 20)); }

int check_tables(const rsf_block* blocks, const int64_t* origins, int nblocks, int64_t w, int64_t h, int64_t l) {
    if (!blocks || !origins || nblocks <= 0) return set_error(RSF_ERR_INVALID_ARGUMENT, "no blocks");
    for (int i = 0; i < nblocks; ++i) {
        const rsf_block& b = blocks[i];
        const int64_t* o = origins + 3 * i;
        if (b.w <= 0 || b.h <= 0 || b.l <= 0 || o[0] < 0 || o[1] < 0 || o[2] < 0 || o[0] + b.w > w ||
            o[1] + b.h > h || o[2] + b.l > l)
            return set_error(RSF_ERR_INVALID_ARGUMENT, "block %d does not fit the volume", i);
        if (b.offset < 0 || (i > 0 && b.offset < blocks[i - 1].offset + blocks[i - 1].w * blocks[i - 1].h *
                                                                         blocks[i - 1].l))
            return set_error(RSF_ERR_INVALID_ARGUMENT,
                             "block %d: packed offsets must be ascending and the blocks disjoint", i);
    }
    return RSF_OK;
}

}  // namespace
}  // namespace rsf

using namespace rsf;

extern "C" {

int rsf_block_grid(int64_t w, int64_t h, int64_t l, int64_t bw, int64_t bh, int64_t bl, int64_t overlap,
                   int64_t min_size, rsf_block* blocks, int64_t* origins, int capacity, int* nblocks,
                   int64_t* packed_voxels) {
    if (!nblocks) return set_error(RSF_ERR_INVALID_ARGUMENT, "nblocks is NULL");
    std::vector<rsf_block> b;
    std::vector<int64_t> o;
    GridDesc gd{};
    int64_t total = 0;
    RSF_TRY(grid_of(w, h, l, bw, bh, bl, overlap, min_size, &b, &o, &gd, &total));
    *nblocks = (int)b.size();
    if (packed_voxels) *packed_voxels = total;
    if (!blocks || !origins) return RSF_OK;  // size query
    if (capacity < (int)b.size())
        return set_error(RSF_ERR_INVALID_ARGUMENT, "capacity %d < %zu blocks", capacity, b.size());
    std::copy(b.begin(), b.end(), blocks);
    std::copy(o.begin(), o.end(), origins);
    return RSF_OK;
}

int rsf_disassemble_device(rsf_context* c, const void* volume, int dtype, int64_t w, int64_t h, int64_t l,
                           const rsf_block* blocks, const int64_t* origins, int nblocks, void* packed) {
    hipStream_t s;
    RSF_TRY(context_stream(c, &s));
    if (!volume || !packed) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (dtype != RSF_DTYPE_U8 && dtype != RSF_DTYPE_F32) return set_error(RSF_ERR_INVALID_ARGUMENT, "dtype");
    RSF_TRY(check_tables(blocks, origins, nblocks, w, h, l));
    std::vector<rsf_block> b(blocks, blocks + nblocks);
    std::vector<int64_t> o(origins, origins + 3 * (size_t)nblocks);
    int64_t total = 0;
    for (auto& x : b) total = std::max(total, x.offset + x.w * x.h * x.l);
    DevGrid dg;
    RSF_TRY(dg.upload(b, o, s));
    hipLaunchKernelGGL(k_disassemble, dim3(grid1d(total)), dim3(256), 0, s, packed, volume,
                       dtype == RSF_DTYPE_U8 ? 1 : 4, w, h, dg.blocks.as<rsf_block>(), dg.origins.as<int64_t>(),
                       nblocks, total);
    RSF_HIP(hipGetLastError());
    RSF_HIP(hipStreamSynchronize(s));
    return RSF_OK;
}

int rsf_reassemble_device(rsf_context* c, const float* packed, int64_t w, int64_t h, int64_t l, int64_t bw,
                          int64_t bh, int64_t bl, int64_t overlap, int64_t min_size, float* volume) {
    hipStream_t s;
    RSF_TRY(context_stream(c, &s));
    if (!volume || !packed) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL buffer");
    std::vector<rsf_block> b;
    std::vector<int64_t> o;
    GridDesc gd{};
    int64_t total = 0;
    RSF_TRY(grid_of(w, h, l, bw, bh, bl, overlap, min_size, &b, &o, &gd, &total));
    DevGrid dg;
    RSF_TRY(dg.upload(b, o, s));
    hipLaunchKernelGGL(k_reassemble, dim3(grid1d(w * h * l)), dim3(256), 0, s, volume, packed, w, h, l,
                       dg.blocks.as<rsf_block>(), dg.origins.as<int64_t>(), gd);
    RSF_HIP(hipGetLastError());
    RSF_HIP(hipStreamSynchronize(s));
    return RSF_OK;
}

int rsf_run_volume_blocks_device(rsf_context* c, const void* image, int image_dtype, float* phi, int64_t w,
                                 int64_t h, int64_t l, int64_t bw, int64_t bh, int64_t bl, int64_t overlap,
                                 const rsf_params* p) {
    hipStream_t s;
    RSF_TRY(context_stream(c, &s));
    if (!image || !phi || !p) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL argument");
    int P = 0, m = 0;
    float taps[256];
    RSF_TRY(rsf_taps(p->sigma, taps, 256, &P, &m));
    std::vector<rsf_block> b;
    std::vector<int64_t> o;
    GridDesc gd{};
    int64_t total = 0;
    RSF_TRY(grid_of(w, h, l, bw, bh, bl, overlap, P, &b, &o, &gd, &total));
    DevBuf pimg, pphi;
    RSF_TRY(pimg.ensure((size_t)total * (image_dtype == RSF_DTYPE_U8 ? 1 : 4)));
    RSF_TRY(pphi.ensure((size_t)total * sizeof(float)));
    RSF_TRY(rsf_disassemble_device(c, image, image_dtype, w, h, l, b.data(), o.data(), (int)b.size(), pimg.p));
    RSF_TRY(rsf_disassemble_device(c, phi, RSF_DTYPE_F32, w, h, l, b.data(), o.data(), (int)b.size(), pphi.p));
    RSF_TRY(rsf_run_batched_device(c, b.data(), (int)b.size(), pimg.p, image_dtype, pphi.as<float>(), p));
    RSF_TRY(rsf_reassemble_device(c, pphi.as<float>(), w, h, l, bw, bh, bl, overlap, P, phi));
    return RSF_OK;
}

int rsf_run_batched_host(rsf_context* c, const rsf_block* blocks, int nblocks, const void* image, int image_dtype,
                         float* phi, const rsf_params* p) {
    hipStream_t s;
    RSF_TRY(context_stream(c, &s));
    if (!image || !phi || !blocks || nblocks <= 0) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL argument");
    int64_t total = 0;
    for (int i = 0; i < nblocks; ++i) total = std::max(total, blocks[i].offset + blocks[i].w * blocks[i].h * blocks[i].l);
    const size_t ib = (size_t)total * (image_dtype == RSF_DTYPE_U8 ? 1 : 4);
    DevBuf dimg, dphi;
    RSF_TRY(dimg.ensure(ib));
    RSF_TRY(dphi.ensure((size_t)total * sizeof(float)));
    RSF_HIP(hipMemcpyAsync(dimg.p, image, ib, hipMemcpyHostToDevice, s));
    RSF_HIP(hipMemcpyAsync(dphi.p, phi, (size_t)total * sizeof(float), hipMemcpyHostToDevice, s));
    RSF_TRY(rsf_run_batched_device(c, blocks, nblocks, dimg.p, image_dtype, dphi.as<float>(), p));
    RSF_HIP(hipMemcpyAsync(phi, dphi.p, (size_t)total * sizeof(float), hipMemcpyDeviceToHost, s));
    RSF_HIP(hipStreamSynchronize(s));
    return RSF_OK;
}

int rsf_run_volume_blocks_host(rsf_context* c, const void* image, int image_dtype, float* phi, int64_t w, int64_t h,
                               int64_t l, int64_t bw, int64_t bh, int64_t bl, int64_t overlap, const rsf_params* p) {
    hipStream_t s;
    RSF_TRY(context_stream(c, &s));
    if (!image || !phi) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (w <= 0 || h <= 0 || l <= 0) return set_error(RSF_ERR_INVALID_ARGUMENT, "dims must be positive");
    const int64_t n = w * h * l;
    const size_t ib = (size_t)n * (image_dtype == RSF_DTYPE_U8 ? 1 : 4);
    DevBuf dimg, dphi;
    RSF_TRY(dimg.ensure(ib));
    RSF_TRY(dphi.ensure((size_t)n * sizeof(float)));
    RSF_HIP(hipMemcpyAsync(dimg.p, image, ib, hipMemcpyHostToDevice, s));
    RSF_HIP(hipMemcpyAsync(dphi.p, phi, (size_t)n * sizeof(float), hipMemcpyHostToDevice, s));
    RSF_TRY(rsf_run_volume_blocks_device(c, dimg.p, image_dtype, dphi.as<float>(), w, h, l, bw, bh, bl, overlap, p));
    RSF_HIP(hipMemcpyAsync(phi, dphi.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
    RSF_HIP(hipStreamSynchronize(s));
    return RSF_OK;
}

}  // extern "C"

// ------------------------------------------------------------ tile votes
// The vote of the seed producer (Unet/run_model.py:29-35,86-91): every voxel
// averages the outputs of the tiles covering it, the sum taken in float64 in the
// tiles' loop order (i along z, then j along y, then k along x), then >= 0.5.
namespace rsf {
namespace {

__global__ void k_tile_vote(const float* __restrict__ pred, int tz, int ty, int tx, int ny, int nx,
                            const int2* __restrict__ cz, const int2* __restrict__ cy, const int2* __restrict__ cx,
                            const int64_t* __restrict__ z0, const int64_t* __restrict__ y0,
                            const int64_t* __restrict__ x0, int64_t w, int64_t h, int64_t n,
                            double* __restrict__ prob, uint8_t* __restrict__ mask) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t x = v % w, y = (v / w) % h, z = v / (w * h);
        const int2 rz = cz[z], ry = cy[y], rx = cx[x];  // covering tile ranges [lo, hi)
        double acc = 0.0;
        int cnt = 0;
        for (int i = rz.x; i < rz.y; ++i)
            for (int j = ry.x; j < ry.y; ++j)
                for (int k = rx.x; k < rx.y; ++k) {
                    const int64_t t = ((int64_t)i * ny + j) * nx + k;
                    acc = acc + (double)pred[((t * tz + (z - z0[i])) * ty + (y - y0[j])) * tx + (x - x0[k])];
                    ++cnt;
                }
        const double p = cnt > 0 ? acc / (double)cnt : 0.0;
        prob[v] = p;
        mask[v] = p >= 0.5 ? 1 : 0;
    }
}

// [lo, hi) of the tiles (non-decreasing starts, one size) covering each coordinate
int cover_ranges(const int64_t* s, int nt, int64_t size, int64_t n, std::vector<int2>* out) {
    out->assign((size_t)n, int2{0, 0});
    for (int t = 0; t < nt; ++t) {
        if (s[t] < 0 || s[t] + size > n || (t > 0 && s[t] < s[t - 1]))
            return set_error(RSF_ERR_INVALID_ARGUMENT, "tile starts must be ascending and inside the volume");
    }
    for (int64_t c = 0; c < n; ++c) {
        int lo = 0;
        while (lo < nt && s[lo] + size <= c) ++lo;
        int hi = lo;
        while (hi < nt && s[hi] <= c) ++hi;
        (*out)[(size_t)c] = int2{lo, hi};
    }
    return RSF_OK;
}

}  // namespace
}  // namespace rsf

extern "C" int rsf_tile_vote_device(rsf_context* c, const float* pred, int64_t tz, int64_t ty, int64_t tx,
                                    const int64_t* z0, int nz, const int64_t* y0, int ny, const int64_t* x0, int nx,
                                    int64_t w, int64_t h, int64_t l, double* prob, uint8_t* mask) {
    hipStream_t s;
    RSF_TRY(context_stream(c, &s));
    if (!pred || !z0 || !y0 || !x0 || !prob || !mask) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL argument");
    if (tz <= 0 || ty <= 0 || tx <= 0 || nz <= 0 || ny <= 0 || nx <= 0 || w <= 0 || h <= 0 || l <= 0)
        return set_error(RSF_ERR_INVALID_ARGUMENT, "tile and volume dims must be positive");
    std::vector<int2> rz, ry, rx;
    RSF_TRY(cover_ranges(z0, nz, tz, l, &rz));
    RSF_TRY(cover_ranges(y0, ny, ty, h, &ry));
    RSF_TRY(cover_ranges(x0, nx, tx, w, &rx));
    DevBuf d_rz, d_ry, d_rx, d_z0, d_y0, d_x0;
    auto up = [&](DevBuf& d, const void* src, size_t bytes) -> int {
        RSF_TRY(d.ensure(bytes));
        RSF_HIP(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, s));
        return RSF_OK;
    };
    RSF_TRY(up(d_rz, rz.data(), rz.size() * sizeof(int2)));
    RSF_TRY(up(d_ry, ry.data(), ry.size() * sizeof(int2)));
    RSF_TRY(up(d_rx, rx.data(), rx.size() * sizeof(int2)));
    RSF_TRY(up(d_z0, z0, (size_t)nz * sizeof(int64_t)));
    RSF_TRY(up(d_y0, y0, (size_t)ny * sizeof(int64_t)));
    RSF_TRY(up(d_x0, x0, (size_t)nx * sizeof(int64_t)));
    const int64_t n = w * h * l;
    hipLaunchKernelGGL(k_tile_vote, dim3(grid1d(n)), dim3(256), 0, s, pred, (int)tz, (int)ty, (int)tx, ny, nx,
                       d_rz.as<int2>(), d_ry.as<int2>(), d_rx.as<int2>(), d_z0.as<int64_t>(), d_y0.as<int64_t>(),
                       d_x0.as<int64_t>(), w, h, n, prob, mask);
    RSF_HIP(hipGetLastError());
    RSF_HIP(hipStreamSynchronize(s));
    return RSF_OK;
}
