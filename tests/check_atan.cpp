// Exhaustive check of the literal path's float arctangent (csrc/rsf_atan.hpp,
// host build) against the oracle's det_atan (oracle/rsf_oracle.cpp) rounded to
// float: every non-negative float bit pattern x (every STRIDE-th one with a
// stride argument) and -x for each.  Prints one line:
//   checked mismatches max_ulp
// max_ulp is the largest distance, in units in the last place of the oracle's
// double, between atan_poly(x) and det_atan(x) over finite nonzero results.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "rsf_atan.hpp"
#include "rsf_oracle.h"

static float f_of(uint32_t b) { float f; std::memcpy(&f, &b, 4); return f; }
static uint32_t b_of(float f) { uint32_t b; std::memcpy(&b, &f, 4); return b; }
static bool same(float a, float b) { return (std::isnan(a) && std::isnan(b)) || b_of(a) == b_of(b); }

int main(int argc, char** argv) {
    const uint64_t stride = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1;
    const uint64_t total = (1ull << 31) / stride;
    const uint64_t chunk = 1 << 16;
    long long mism = 0, checked = 0;
    double max_ulp = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : mism, checked) reduction(max : max_ulp)
    for (uint64_t c0 = 0; c0 < total; c0 += chunk) {
        const uint64_t n = c0 + chunk <= total ? chunk : total - c0;
        std::vector<double> xs(2 * n), ys(2 * n);
        for (uint64_t i = 0; i < n; ++i) {
            xs[i] = (double)f_of((uint32_t)((c0 + i) * stride));
            xs[n + i] = -xs[i];
        }
        oracle_atan_f64(ys.data(), xs.data(), (int64_t)(2 * n));
        for (uint64_t i = 0; i < 2 * n; ++i) {
            const float x = (float)xs[i];
            const float got = rsf::atanf_literal(x), ref = (float)ys[i];
            if (!same(got, ref)) {
                if (mism < 8) std::fprintf(stderr, "mismatch x=%a got %a ref %a\n", (double)x, (double)got, (double)ref);
                ++mism;
            }
            if (i < n && std::isfinite(ys[i]) && ys[i] != 0) {
                const double ulp = std::ldexp(1.0, std::ilogb(ys[i]) - 52);
                const double e = std::fabs(rsf::atan_poly(xs[i]) - ys[i]) / ulp;
                if (e > max_ulp) max_ulp = e;
            }
            ++checked;
        }
    }
    std::printf("%lld %lld %.3f\n", checked, mism, max_ulp);
    return mism != 0;
}
