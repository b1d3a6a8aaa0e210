// Synthetic tube phantoms and initial level sets (host, OpenMP).
//
// rsf_synth_tubes / rsf_initial_phi build the benchmark inputs: polyline tubes
// at 200 on a background of 30, a linear x ramp, Gaussian noise; phi0 is the
// signed Euclidean distance to a thresholded Gaussian blur of the volume.
// rsf_sdf_from_seed is phi0 = sdf(seed - 13) of CUDA_RSF/main.cpp:174-183.
// Distances are exact: the separable squared-distance transform of
// Felzenszwalb & Huttenlocher ("Distance Transforms of Sampled Functions",
// Theory of Computing 8, 2012), one lower-envelope pass per axis.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "rsf.h"

namespace rsf {
int set_error(int status, const char* fmt, ...);
}

namespace {

inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    double u01() { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        double u1 = u01(), u2 = u01();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

// Counter-based standard normal for voxel i (independent of thread schedule).
inline double voxel_normal(uint64_t seed, uint64_t i) {
    uint64_t s = seed ^ (i * 0xD1B54A32D192ED03ull);
    Rng r(splitmix64(s));
    return r.normal();
}

struct Seg {
    double a[3], b[3];
    double r;
};

bool dims_ok(int64_t w, int64_t h, int64_t l) { return w > 0 && h > 0 && l > 0 && w * h * l < (int64_t(1) << 40); }

// 1-D squared distance transform of f (length n) into d: the lower envelope
// of the parabolas (q - p)^2 + f(p), with v/z scratch of n and n+1 entries.
void dt1d(const float* f, float* d, int n, int* v, double* z) {
    const double inf = std::numeric_limits<double>::infinity();
    auto inter = [&](int q, int p) {
        return (((double)f[q] + (double)q * q) - ((double)f[p] + (double)p * p)) / (2.0 * (q - p));
    };
    int k = 0;
    v[0] = 0;
    z[0] = -inf;
    z[1] = inf;
    for (int q = 1; q < n; ++q) {
        double s = inter(q, v[k]);
        while (s <= z[k]) {
            --k;
            s = inter(q, v[k]);
        }
        ++k;
        v[k] = q;
        z[k] = s;
        z[k + 1] = inf;
    }
    k = 0;
    for (int q = 0; q < n; ++q) {
        while (z[k + 1] < (double)q) ++k;
        const double dq = (double)(q - v[k]);
        d[q] = (float)(dq * dq + (double)f[v[k]]);
    }
}

// In-place squared EDT of g (0 at feature voxels, "infinity" elsewhere).
void edt3(std::vector<float>& g, int w, int h, int l) {
    const int nmax = std::max(w, std::max(h, l));
    auto pass = [&](int axis) {
        const int n = axis == 0 ? w : (axis == 1 ? h : l);
        const int64_t stride = axis == 0 ? 1 : (axis == 1 ? (int64_t)w : (int64_t)w * h);
        const int64_t lines = (int64_t)w * h * l / n;
#pragma omp parallel
        {
            std::vector<float> f(nmax), d(nmax);
            std::vector<double> z(nmax + 1);
            std::vector<int> v(nmax);
#pragma omp for schedule(static)
            for (int64_t li = 0; li < lines; ++li) {
                int64_t base;
                if (axis == 0) base = li * w;
                else if (axis == 1) base = (li / w) * (int64_t)w * h + (li % w);
                else base = li;
                for (int q = 0; q < n; ++q) f[q] = g[base + q * stride];
                dt1d(f.data(), d.data(), n, v.data(), z.data());
                for (int q = 0; q < n; ++q) g[base + q * stride] = d[q];
            }
        }
    };
    pass(0);
    pass(1);
    pass(2);
}

// Signed distance with the zero level half-way between voxel centres:
// outside voxels get +(distance to the nearest inside voxel) - 1/2, inside
// voxels -(distance to the nearest outside voxel) + 1/2; multiplied by `sign`.
void signed_distance(const std::vector<uint8_t>& inside, float* phi, int w, int h, int l, float sign) {
    const int64_t n = (int64_t)w * h * l;
    const float big = 1e20f;
    const float cap = std::sqrt((float)w * w + (float)h * h + (float)l * l);
    std::vector<float> din(n), dout(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        din[i] = inside[i] ? big : 0.0f;   // distance to nearest outside voxel
        dout[i] = inside[i] ? 0.0f : big;  // distance to nearest inside voxel
    }
    edt3(din, w, h, l);
    edt3(dout, w, h, l);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        float v;
        if (inside[i]) v = -(std::min(std::sqrt(din[i]), cap) - 0.5f);
        else v = std::min(std::sqrt(dout[i]), cap) - 0.5f;
        phi[i] = sign * v;
    }
}

}  // namespace

extern "C" {

int rsf_synth_tubes(uint8_t* image, uint8_t* mask_out, int64_t w, int64_t h, int64_t l, int tubes, float rmin,
                    float rmax, float bias, float noise_sigma, uint64_t seed) {
    if (!image) return rsf::set_error(RSF_ERR_INVALID_ARGUMENT, "image is NULL");
    if (!dims_ok(w, h, l)) return rsf::set_error(RSF_ERR_INVALID_ARGUMENT, "bad dims");
    if (tubes < 0 || !(rmin > 0) || rmax < rmin || noise_sigma < 0)
        return rsf::set_error(RSF_ERR_INVALID_ARGUMENT, "bad tube parameters");
    const int64_t n = w * h * l;
    std::vector<Seg> segs;
    const double dim[3] = {(double)w, (double)h, (double)l};
    for (int t = 0; t < tubes; ++t) {
        uint64_t s0 = seed * 0x9E3779B97F4A7C15ull + (uint64_t)t * 0x632BE59BD9B4E019ull + 1;
        Rng rng(splitmix64(s0));
        const int nseg = 5 + (int)(rng.u01() * 6.0);
        const double r = rmin + rng.u01() * (rmax - rmin);
        double p[3];
        for (int a = 0; a < 3; ++a) p[a] = rng.u01() * (dim[a] - 1);
        for (int s = 0; s < nseg; ++s) {
            double d[3], nrm = 0;
            for (int a = 0; a < 3; ++a) { d[a] = rng.normal(); nrm += d[a] * d[a]; }
            nrm = std::sqrt(std::max(nrm, 1e-12));
            const double len = 10.0 + rng.u01() * 30.0;
            Seg sg;
            for (int a = 0; a < 3; ++a) {
                sg.a[a] = p[a];
                sg.b[a] = std::min(std::max(p[a] + len * d[a] / nrm, 0.0), dim[a] - 1);
                p[a] = sg.b[a];
            }
            sg.r = r;
            segs.push_back(sg);
        }
    }
    std::vector<uint8_t> mask(n, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t z = 0; z < l; ++z) {
        for (const Seg& sg : segs) {
            const double zlo = std::min(sg.a[2], sg.b[2]) - sg.r, zhi = std::max(sg.a[2], sg.b[2]) + sg.r;
            if (z < zlo || z > zhi) continue;
            const int64_t ylo = std::max<int64_t>(0, (int64_t)std::floor(std::min(sg.a[1], sg.b[1]) - sg.r));
            const int64_t yhi = std::min<int64_t>(h - 1, (int64_t)std::ceil(std::max(sg.a[1], sg.b[1]) + sg.r));
            const int64_t xlo = std::max<int64_t>(0, (int64_t)std::floor(std::min(sg.a[0], sg.b[0]) - sg.r));
            const int64_t xhi = std::min<int64_t>(w - 1, (int64_t)std::ceil(std::max(sg.a[0], sg.b[0]) + sg.r));
            double ab[3], ab2 = 0;
            for (int a = 0; a < 3; ++a) { ab[a] = sg.b[a] - sg.a[a]; ab2 += ab[a] * ab[a]; }
            for (int64_t y = ylo; y <= yhi; ++y)
                for (int64_t x = xlo; x <= xhi; ++x) {
                    const double q[3] = {(double)x, (double)y, (double)z};
                    double t = 0;
                    if (ab2 > 0) {
                        for (int a = 0; a < 3; ++a) t += (q[a] - sg.a[a]) * ab[a];
                        t = std::min(std::max(t / ab2, 0.0), 1.0);
                    }
                    double d2 = 0;
                    for (int a = 0; a < 3; ++a) {
                        const double e = q[a] - (sg.a[a] + t * ab[a]);
                        d2 += e * e;
                    }
                    if (d2 <= sg.r * sg.r) mask[(z * h + y) * w + x] = 1;
                }
        }
    }
    const uint64_t nseed = seed ^ 0xA0761D6478BD642Full;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const int64_t x = i % w;
        double v = mask[i] ? 200.0 : 30.0;
        v += (w > 1) ? bias * (2.0 * (double)x / (double)(w - 1) - 1.0) : 0.0;
        if (noise_sigma > 0) v += noise_sigma * voxel_normal(nseed, (uint64_t)i);
        v = std::min(std::max(std::nearbyint(v), 0.0), 255.0);
        image[i] = (uint8_t)v;
        if (mask_out) mask_out[i] = mask[i];
    }
    return RSF_OK;
}

int rsf_initial_phi(const uint8_t* image, float* phi, int64_t w, int64_t h, int64_t l, float blur_sigma,
                    float level) {
    if (!image || !phi) return rsf::set_error(RSF_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (!dims_ok(w, h, l)) return rsf::set_error(RSF_ERR_INVALID_ARGUMENT, "bad dims");
    if (!(blur_sigma >= 0)) return rsf::set_error(RSF_ERR_INVALID_ARGUMENT, "blur_sigma must be >= 0");
    const int64_t n = w * h * l;
    std::vector<float> a(n), b(n);
    for (int64_t i = 0; i < n; ++i) a[i] = image[i];
    if (blur_sigma > 0) {
        const int R = std::max(1, (int)std::ceil(3.0 * blur_sigma));
        std::vector<float> g(2 * R + 1);
        double s = 0;
        for (int k = -R; k <= R; ++k) s += (g[k + R] = (float)std::exp(-0.5 * k * k / (blur_sigma * blur_sigma)));
        for (auto& x : g) x = (float)(x / s);
        const int64_t dims[3] = {w, h, l};
        const int64_t strides[3] = {1, w, w * h};
        for (int axis = 0; axis < 3; ++axis) {
            const int64_t D = dims[axis], st = strides[axis];
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; ++i) {
                const int64_t c = (i / st) % D;
                float acc = 0;
                for (int k = -R; k <= R; ++k) {
                    const int64_t cc = std::min(std::max(c + k, (int64_t)0), D - 1);
                    acc += g[k + R] * a[i + (cc - c) * st];
                }
                b[i] = acc;
            }
            a.swap(b);
        }
    }
    std::vector<uint8_t> inside(n);
    bool any = false, all = true;
    for (int64_t i = 0; i < n; ++i) {
        inside[i] = a[i] > level;
        any |= inside[i] != 0;
        all &= inside[i] != 0;
    }
    signed_distance(inside, phi, (int)w, (int)h, (int)l, 1.0f);
    (void)any;
    (void)all;
    return RSF_OK;
}

int rsf_sdf_from_seed(const float* seed, float* phi, int64_t w, int64_t h, int64_t l, float shift) {
    if (!seed || !phi) return rsf::set_error(RSF_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (!dims_ok(w, h, l)) return rsf::set_error(RSF_ERR_INVALID_ARGUMENT, "bad dims");
    const int64_t n = w * h * l;
    std::vector<uint8_t> inside(n);
    for (int64_t i = 0; i < n; ++i) inside[i] = (seed[i] - shift) > 0.0f;
    // positive inside the region {seed - shift > 0}
    signed_distance(inside, phi, (int)w, (int)h, (int)l, -1.0f);
    return RSF_OK;
}

}  // extern "C"
