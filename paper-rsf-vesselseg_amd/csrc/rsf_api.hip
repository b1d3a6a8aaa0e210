// Host engine behind the C ABI: contexts, workspace, the graph-captured
// iteration loop (CUDA_RSF/main.cpp:222-314), batched and per-block runs, the
// autotuner and the unfused device pipeline.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <initializer_list>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "rsf_internal.hpp"

namespace rsf {

// ------------------------------------------------------------ error plumbing
static thread_local std::string g_last_error;

int set_error(int status, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return status;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
    int st = (e == hipErrorOutOfMemory) ? RSF_ERR_OUT_OF_MEMORY : RSF_ERR_DEVICE;
    return set_error(st, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

int DevBuf::ensure(size_t n) {
    if (n <= bytes && p) return RSF_OK;
    release();
    if (n == 0) return RSF_OK;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
        p = nullptr;
        bytes = 0;
        return set_error(RSF_ERR_OUT_OF_MEMORY, "hipMalloc(%zu bytes): %s", n, hipGetErrorString(e));
    }
    bytes = n;
    return RSF_OK;
}

void DevBuf::release() {
    if (p && owned) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    owned = true;
}

// --------------------------------------------------------------- parameters
static int taps32(float sigma, float* g, int cap, int* P_out, int* m_out) {
    // main.cpp:152-166 (taps), :264,273 (pad).
    if (!(sigma > 0.0f) || !std::isfinite(sigma))
        return set_error(RSF_ERR_INVALID_ARGUMENT, "sigma must be positive, got %g", sigma);
    const float s7 = 7 * sigma;
    const unsigned int p = (unsigned int)s7;
    const int k_size = (int)s7;
    if (p % 2 == 0)
        return set_error(RSF_ERR_INVALID_ARGUMENT,
                         "floor(7*sigma) = %u is even: the reference writes its blurred buffers with "
                         "%u taps but sizes them for %u (main.cpp:121,152-155,191); choose sigma with "
                         "an odd floor(7*sigma)", p, p + 1, p);
    if ((int)p > cap) return set_error(RSF_ERR_UNSUPPORTED, "sigma %g needs %u taps (max %d)", sigma, p, cap);
    const float miu = (float)(p / 2);
    const int u = (int)(2 * sigma * sigma);
    if (u <= 0) return set_error(RSF_ERR_INVALID_ARGUMENT, "sigma %g gives a zero tap width (2 sigma^2 < 1)", sigma);
    for (int xi = 0; xi < (int)p; xi++) {
        const float d = (float)xi - miu;
        const float ex = (-d) * d / (float)u;
        g[xi] = 1.0f / sqrtf((float)u * (float)kPi) * expf(ex);
    }
    *P_out = (int)p;
    *m_out = (k_size - 1) / 2;
    return RSF_OK;
}

int check_params(const rsf_params* p) {
    if (!p) return set_error(RSF_ERR_INVALID_ARGUMENT, "params is NULL");
    const float v[] = {p->sigma, p->lambda1, p->lambda2, p->nu, p->mu, p->dt, p->epsilon, p->sigma2};
    for (float x : v)
        if (!std::isfinite(x)) return set_error(RSF_ERR_INVALID_ARGUMENT, "parameters must be finite");
    if (!(p->epsilon > 0)) return set_error(RSF_ERR_INVALID_ARGUMENT, "epsilon must be > 0");
    if (p->iterations < 0) return set_error(RSF_ERR_INVALID_ARGUMENT, "iterations must be >= 0");
    if (p->mode != RSF_MODE_FAST && p->mode != RSF_MODE_LITERAL)
        return set_error(RSF_ERR_INVALID_ARGUMENT, "unknown mode %d", p->mode);
    if ((int)p->sigma2 < 0) return set_error(RSF_ERR_INVALID_ARGUMENT, "sigma2 must be >= 0");
    return RSF_OK;
}

int derive(const rsf_params* p, Derived* d) {
    RSF_TRY(check_params(p));
    RSF_TRY(taps32(p->sigma, d->taps.g, kMaxTaps, &d->P, &d->m));
    for (int k = d->P; k < 2 * d->P - 1; ++k) d->taps.g[k] = d->taps.g[k - d->P];
    double s = 0;
    for (int k = 0; k < d->P; ++k) s += d->taps.g[k];
    d->halfS3 = (float)(0.5 * s * s * s);
    return RSF_OK;
}

}  // namespace rsf

using namespace rsf;

// ------------------------------------------------------------------- context
namespace {

struct GraphSet {
    hipGraphExec_t main = nullptr, rem = nullptr;
    int n_main = 0, n_launch = 0, n_rem = 0;
    void reset() {
        if (main) (void)hipGraphExecDestroy(main);
        if (rem) (void)hipGraphExecDestroy(rem);
        main = rem = nullptr;
        n_main = n_launch = n_rem = 0;
    }
};

struct Runner {
    std::string key;
    TilePlan plan;
    GraphSet graphs;
    std::string graph_key;
    uint64_t ws_epoch = 0;
    ~Runner() { graphs.reset(); }
};

struct TuneChoice {
    std::string s1, s2;
    int zc1 = 0, zc2 = 0;
};

}  // namespace

struct rsf_context {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int num_cus = 256;
    // workspace (grow-only); ws_epoch changes whenever a pointer changes
    DevBuf phi2, tr, F, blurI, imgf, tmp1, tmp2, fo, fi, Eo, Ei;
    DevBuf unf[20];
    DevBuf adj[14];       // adjoint: fo, fi, PH, PHc, QH, QA, Bh, Ba, Gx, Gy, Gz, t1, t2, partials
    DevBuf h_img, h_phi;  // staging of rsf_run_host
    uint64_t ws_epoch = 1;
    std::list<std::unique_ptr<Runner>> runners;
    std::map<std::string, TuneChoice> tune;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> iter_events;
    float last_ms = 0.f;
    // rsf_context_hold_image: an image the caller keeps unchanged, and the key of
    // the preparation (float copy, blur) last computed for it
    const void* held_image = nullptr;
    std::string held_key;
    DevBuf nonfinite;  // count of NaN/inf voxels found after a run
    unsigned char* scratch = nullptr;  // this device's copy of the allocation-free step's scratch
};

// ------------------------------------------------- allocation-free step scratch
// rsf_step_device_noalloc keeps its workspace here instead of allocating it:
// the transform T(phi) and the force F of up to kNoAllocMaxVoxels voxels, then
// the tile plan (tile lists, sweep-2 flags, block table).  Each device the
// library runs on holds its own copy for the life of the process.  In host code
// `scratch` only names the symbol; its device address comes from
// hipGetSymbolAddress, once per device (scratch_address).
namespace {
constexpr int64_t kNoAllocMaxVoxels = int64_t(1) << 23;
constexpr size_t kNoAllocPlanBytes = size_t(4) << 20;
constexpr size_t kNoAllocBytes = 2 * (size_t)kNoAllocMaxVoxels * sizeof(float) + kNoAllocPlanBytes;
}  // namespace
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch_addr;   // device -> address of its scratch

// The device address of `scratch` on the current device `device` (looked up once).
int scratch_address(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_addr.find(device);
    if (it != g_scratch_addr.end()) {
        *out = it->second;
        return RSF_OK;
    }
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    g_scratch_addr[device] = static_cast<unsigned char*>(ptr);
    *out = static_cast<unsigned char*>(ptr);
    return RSF_OK;
}

int ensure_ws(rsf_context* c, DevBuf& b, size_t bytes) {
    if (b.p && b.bytes >= bytes) return RSF_OK;
    RSF_TRY(b.ensure(bytes));
    c->ws_epoch++;
    return RSF_OK;
}

std::string blocks_key(const std::vector<BlockDesc>& b) {
    std::ostringstream os;
    for (auto& x : b) os << x.off << ':' << x.w << 'x' << x.h << 'x' << x.l << ';';
    return os.str();
}

std::string tune_key(const std::vector<BlockDesc>& b, int P, int literal) {
    std::vector<std::string> dims;
    for (auto& x : b) {
        std::ostringstream os;
        os << x.w << 'x' << x.h << 'x' << x.l;
        dims.push_back(os.str());
    }
    std::sort(dims.begin(), dims.end());
    std::ostringstream os;
    os << "P" << P << (literal ? "L" : "F") << '|';
    for (auto& d : dims) os << d << ',';
    return os.str();
}

int find_variant1(const std::string& name) {
    const auto& v = sweep1_variants();
    for (size_t k = 0; k < v.size(); ++k)
        if (name == v[k].name) return (int)k;
    return -1;
}
int find_variant2(const std::string& name) {
    const auto& v = sweep2_variants();
    for (size_t k = 0; k < v.size(); ++k)
        if (name == v[k].name) return (int)k;
    return -1;
}

void load_tune_file(rsf_context* c) {
    const char* path = std::getenv("RSF_TUNE_CACHE");
    if (!path) return;
    std::ifstream in(path);
    std::string key, s1, s2;
    int zc1, zc2;
    while (in >> key >> s1 >> zc1 >> s2 >> zc2) c->tune[key] = TuneChoice{s1, s2, zc1, zc2};
}

void save_tune_file(rsf_context* c) {
    const char* path = std::getenv("RSF_TUNE_CACHE");
    if (!path) return;
    std::ofstream out(path);
    for (auto& kv : c->tune)
        out << kv.first << ' ' << kv.second.s1 << ' ' << kv.second.zc1 << ' ' << kv.second.s2 << ' '
            << kv.second.zc2 << '\n';
}

int max_dim_l(const std::vector<BlockDesc>& b) {
    int m = 0;
    for (auto& x : b) m = std::max(m, x.l);
    return m;
}

// Chooses kernel variants and z-chunks for a problem: tuned if known, else defaults.
void pick_variants(rsf_context* c, const std::vector<BlockDesc>& blocks, const Derived& d, int literal,
                   bool aligned, int* s1, int* s2, int* zc1, int* zc2) {
    *s1 = default_sweep1(d.P, literal);
    *s2 = default_sweep2(literal);
    // a z kernel of wide work items (1024 columns, one workgroup per CU) is for
    // planes that give every CU one: smaller planes take the first variant whose z
    // work items are 256 columns
    if (*s1 >= 0 && sweep1_variants()[*s1].fn2) {
        const int64_t span = (int64_t)sweep1_variants()[*s1].tx2 * sweep1_variants()[*s1].ty2;
        bool small_plane = false;
        for (const auto& b : blocks) small_plane = small_plane || (int64_t)b.w * b.h < span * c->num_cus;
        for (size_t k = 0; span > 256 && small_plane && k < sweep1_variants().size(); ++k) {
            const auto& v = sweep1_variants()[k];
            if (v.P == d.P && v.literal == literal && v.fn2 && (int64_t)v.tx2 * v.ty2 <= 256) {
                *s1 = (int)k;
                break;
            }
        }
    }
    // a strip xy kernel wastes the lanes past a block's valid columns and runs one
    // wave per strip and plane: where some block is narrower than three strips, the
    // first strip variant of half the width (same z work items) takes over if it
    // wastes 15 % fewer lanes, or if the wider strips give under 8 waves per CU
    // (measured: 96^3 literal 0.092 -> 0.082 ms, 160^3 0.145 -> 0.130; fast 128^3
    // 0.115 -> 0.098, 8 x 128^3 0.169 -> 0.151; 8 x 128^3 literal keeps the 64-column
    // strips, 0.222 against 0.235)
    if (*s1 >= 0 && sweep1_variants()[*s1].strip > 0) {
        const Sweep1Variant& cur = sweep1_variants()[*s1];
        bool narrow = false;
        for (const auto& b : blocks) narrow = narrow || b.w - 2 * d.m < 3 * cur.tx;
        for (size_t k = 0; narrow && k < sweep1_variants().size(); ++k) {
            const auto& v = sweep1_variants()[k];
            if (v.P != d.P || v.literal != literal || v.strip == 0 || 2 * v.tx != cur.tx ||
                v.tx2 * v.ty2 != cur.tx2 * cur.ty2)
                continue;
            // valid columns over computed columns, planes weighted by their rows
            auto waste = [&](int tx, double* eff, int64_t* waves) {
                double valid = 0, computed = 0;
                *waves = 0;
                for (const auto& b : blocks) {
                    const int vw = std::max(b.w - 2 * d.m, 1), strips = (vw + tx - 1) / tx;
                    valid += (double)vw * b.h * b.l;
                    computed += (double)strips * tx * b.h * b.l;
                    *waves += (int64_t)strips * b.l;
                }
                *eff = valid / computed;
            };
            double e_cur, e_alt;
            int64_t w_cur, w_alt;
            waste(cur.tx, &e_cur, &w_cur);
            waste(v.tx, &e_alt, &w_alt);
            if (e_alt >= 1.15 * e_cur || w_cur < 8 * (int64_t)c->num_cus) *s1 = (int)k;
            break;
        }
    }
    const int lmax = max_dim_l(blocks);
    int64_t vox = 0;
    for (const auto& b : blocks) vox += (int64_t)b.w * b.h * b.l;
    // Longest z-chunks that still give every CU a few tiles: a sweep-1 chunk
    // re-reads P-1 halo planes, a sweep-2 chunk restarts its register pipeline.
    auto chunk = [&](int tx, int ty, int per_cu, int zmin) {
        int64_t cols = 0;
        for (const auto& b : blocks) cols += (int64_t)((b.w + tx - 1) / tx) * ((b.h + ty - 1) / ty);
        int zc = lmax;
        while (zc > zmin && cols * ((lmax + zc - 1) / zc) < (int64_t)per_cu * c->num_cus) zc = (zc + 1) / 2;
        return std::max(zc, 1);
    };
    // no compiled variant for this P (or mode): the caller takes the unfused path
    *zc1 = *s1 >= 0 ? chunk(sweep1_variants()[*s1].tx, sweep1_variants()[*s1].ty, 2, 4 * d.P) : std::min(lmax, 128);
    // split sweep 1: the xy kernel has no z state, short chunks balance its work items;
    // from 2^28 voxels on, longer chunks measured faster (32x60 literal tiles at
    // 1024x1024x768: 15.0 ms with 128 planes, 15.4 with 64, 16.4 with 32; at 512^3
    // 32 planes win); a strip xy kernel takes one plane per wave of a work item
    auto split_zc1 = [&](const Sweep1Variant& v) {
        return v.strip > 0 ? std::min(lmax, v.strip) : std::min(lmax, vox >= (int64_t(1) << 28) ? 128 : 32);
    };
    if (*s1 >= 0 && sweep1_variants()[*s1].fn2) *zc1 = split_zc1(sweep1_variants()[*s1]);
    // sweep 2: chunks of at least 16 planes, down to 4 where 16 would leave CUs without
    // a work item (a single 128^3 block: LITERAL 0.050 -> 0.030 ms, FAST 0.034 -> 0.020)
    if (*s2 >= 0) {
        const int tx2 = sweep2_variants()[*s2].tx * sweep2_variants()[*s2].vec, ty2 = sweep2_variants()[*s2].ty;
        int64_t cols = 0;
        for (const auto& b : blocks) cols += (int64_t)((b.w + tx2 - 1) / tx2) * ((b.h + ty2 - 1) / ty2);
        const int zmin = cols * ((lmax + 15) / 16) < (int64_t)c->num_cus ? 4 : 16;
        *zc2 = std::min(chunk(tx2, ty2, 4, zmin), 64);
    } else {
        *zc2 = std::min(lmax, 64);
    }
    // explicit overrides: RSF_SWEEP1 / RSF_SWEEP2 = "<variant name>:<z chunk>"
    for (int which = 1; which <= 2; ++which) {
        const char* env = std::getenv(which == 1 ? "RSF_SWEEP1" : "RSF_SWEEP2");
        if (!env) continue;
        std::string v(env);
        const size_t colon = v.find(':');
        const std::string name = v.substr(0, colon);
        const int zc = colon == std::string::npos ? 0 : std::atoi(v.c_str() + colon + 1);
        if (which == 1) {
            const int a = find_variant1(name);
            if (a >= 0 && sweep1_variants()[a].P == d.P && sweep1_variants()[a].literal == literal) {
                *s1 = a;
                if (zc > 0) *zc1 = zc;
                else if (sweep1_variants()[a].fn2) *zc1 = split_zc1(sweep1_variants()[a]);
            }
        } else {
            const int b = find_variant2(name);
            if (b >= 0 && sweep2_variants()[b].literal == literal) {
                *s2 = b;
                if (zc > 0) *zc2 = zc;
            }
        }
    }
    if (!std::getenv("RSF_SWEEP1") && !std::getenv("RSF_SWEEP2")) {
        auto it = c->tune.find(tune_key(blocks, d.P, literal));
        if (it != c->tune.end()) {
            const int a = find_variant1(it->second.s1), b = find_variant2(it->second.s2);
            if (a >= 0 && sweep1_variants()[a].P == d.P && sweep1_variants()[a].literal == literal) {
                *s1 = a;
                *zc1 = it->second.zc1;
            }
            if (b >= 0 && sweep2_variants()[b].literal == literal) {
                *s2 = b;
                *zc2 = it->second.zc2;
            }
        }
    }
    // a z kernel over the interleaved intermediate addresses P + D planes of it
    // through 32-bit offsets: larger planes take the first variant without that limit
    if (*s1 >= 0 && sweep1_variants()[*s1].max_plane > 0) {
        int64_t pmax = 0;
        for (const auto& b : blocks) pmax = std::max<int64_t>(pmax, (int64_t)b.w * b.h);
        if (pmax > sweep1_variants()[*s1].max_plane) {
            for (size_t k = 0; k < sweep1_variants().size(); ++k) {
                const auto& v = sweep1_variants()[k];
                if (v.P == d.P && v.literal == literal && v.max_plane == 0) {
                    *s1 = (int)k;
                    break;
                }
            }
        }
    }
    // The split and vector sweep-2 kernels address a block through 32-bit buffer
    // offsets; blocks of 2^30 voxels and more take the direct kernel.
    bool small = true;
    for (const auto& b : blocks)
        small = small && 4.0 * ((double)b.w * b.h * b.l + 4.0 * b.w * b.h + 4.0 * b.w + 64.0) < 4294967295.0;
    if (!small && *s2 >= 0 && sweep2_variants()[*s2].shell_fn) *s2 = find_variant2(literal ? "lit_64x4" : "fast_64x4");
    // likewise the z kernel of the split sweep 1: take the fused kernel
    if (!small && *s1 >= 0 && sweep1_variants()[*s1].fn2) {
        for (size_t k = 0; k < sweep1_variants().size(); ++k) {
            const auto& v = sweep1_variants()[k];
            if (!v.fn2 && v.P == d.P && v.literal == literal) {
                *s1 = (int)k;
                *zc1 = std::min(lmax, 512);
                break;
            }
        }
    }
    // the vector interior reads 16-byte groups: caller buffers must start on a 16-byte boundary
    if (!aligned && *s2 >= 0 && sweep2_variants()[*s2].vec == 4)
        *s2 = find_variant2(literal ? "lit_split_64x4" : "fast_split_64x4");
}

// true when the float ranges [a, a+n) and [b, b+n) share an element
bool ranges_overlap(const float* a, const float* b, int64_t n) {
    const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b, bytes = (uintptr_t)n * sizeof(float);
    return x < y + bytes && y < x + bytes;
}

// true when every pointer is 16-byte aligned (null pointers are ignored)
bool aligned16(std::initializer_list<const void*> ptrs) {
    for (const void* q : ptrs)
        if (q && ((uintptr_t)q & 15) != 0) return false;
    return true;
}

Runner* get_runner(rsf_context* c, const std::vector<BlockDesc>& blocks, int s1, int s2, int zc1, int zc2,
                   int* status) {
    std::ostringstream os;
    os << blocks_key(blocks) << '|' << s1 << '|' << s2 << '|' << zc1 << '|' << zc2;
    const std::string key = os.str();
    for (auto it = c->runners.begin(); it != c->runners.end(); ++it) {
        if ((*it)->key == key) {
            c->runners.splice(c->runners.begin(), c->runners, it);
            *status = RSF_OK;
            return c->runners.front().get();
        }
    }
    auto r = std::make_unique<Runner>();
    r->key = key;
    *status = plan_tiles(r->plan, blocks, s1, s2, zc1, zc2, c->num_cus);
    if (*status != RSF_OK) return nullptr;
    c->runners.push_front(std::move(r));
    while (c->runners.size() > 96) c->runners.pop_back();
    return c->runners.front().get();
}

int validate_blocks(const rsf_block* blocks, int nblocks, const Derived& d, std::vector<BlockDesc>* out,
                    int64_t* total) {
    if (!blocks || nblocks <= 0) return set_error(RSF_ERR_INVALID_ARGUMENT, "no blocks");
    out->clear();
    *total = 0;
    int64_t end = 0;
    for (int b = 0; b < nblocks; ++b) {
        const rsf_block& x = blocks[b];
        if (x.w < d.P || x.h < d.P || x.l < d.P)
            return set_error(RSF_ERR_INVALID_ARGUMENT,
                             "block %d dims (%lld,%lld,%lld) are smaller than the %d-tap blur", b,
                             (long long)x.w, (long long)x.h, (long long)x.l, d.P);
        if (x.w > (1 << 30) || x.h > (1 << 30) || x.l > (1 << 30) || x.offset < 0)
            return set_error(RSF_ERR_INVALID_ARGUMENT, "block %d out of range", b);
        out->push_back(BlockDesc{x.offset, (int)x.w, (int)x.h, (int)x.l, 0});
        *total += x.w * x.h * x.l;
        end = std::max(end, x.offset + x.w * x.h * x.l);
    }
    // blocks must not overlap
    std::vector<std::pair<int64_t, int64_t>> iv;
    for (auto& x : *out) iv.push_back({x.off, x.off + (int64_t)x.w * x.h * x.l});
    std::sort(iv.begin(), iv.end());
    for (size_t k = 1; k < iv.size(); ++k)
        if (iv[k].first < iv[k - 1].second) return set_error(RSF_ERR_INVALID_ARGUMENT, "blocks overlap");
    *total = std::max(*total, end);  // buffer extent in voxels
    return RSF_OK;
}

struct Prepared {
    Derived d;
    std::vector<BlockDesc> blocks;
    int64_t extent = 0;  // voxels spanned by the packed buffers
    int literal = 0, sw = 0;
    Runner* runner = nullptr;
    bool fused = true;
};

int prepare(rsf_context* c, const rsf_block* blocks, int nblocks, const rsf_params* p, bool aligned, Prepared* pr) {
    RSF_TRY(derive(p, &pr->d));
    RSF_TRY(validate_blocks(blocks, nblocks, pr->d, &pr->blocks, &pr->extent));
    pr->literal = p->mode == RSF_MODE_LITERAL;
    pr->sw = (int)p->sigma2;
    int s1, s2, zc1, zc2;
    pick_variants(c, pr->blocks, pr->d, pr->literal, aligned, &s1, &s2, &zc1, &zc2);
    pr->fused = s1 >= 0 && s2 >= 0;
    if (!pr->fused) return RSF_OK;
    int st = RSF_OK;
    pr->runner = get_runner(c, pr->blocks, s1, s2, zc1, zc2, &st);
    return st;
}

// true when the plan's sweep 1 is the split xy/z pair (which can also blur the image)
bool split_plan(const Prepared& pr) {
    return pr.fused && sweep1_variants()[pr.runner->plan.s1].fn2 != nullptr;
}

int ensure_workspace(rsf_context* c, const Prepared& pr) {
    const size_t n = (size_t)pr.extent * sizeof(float);
    RSF_TRY(ensure_ws(c, c->phi2, n));
    RSF_TRY(ensure_ws(c, c->tr, n));
    RSF_TRY(ensure_ws(c, c->F, n));
    if (!pr.literal) RSF_TRY(ensure_ws(c, c->blurI, n));
    if (pr.sw > 0) {
        RSF_TRY(ensure_ws(c, c->fo, n));
        RSF_TRY(ensure_ws(c, c->fi, n));
        RSF_TRY(ensure_ws(c, c->Eo, n));
        RSF_TRY(ensure_ws(c, c->Ei, n));
    }
    if (!pr.literal || pr.sw > 0) RSF_TRY(ensure_ws(c, c->imgf, n));
    if (!pr.literal && !split_plan(pr)) {
        // pass temporaries of the per-op image blur (a split plan blurs through its own kernels)
        size_t tmax = 0;
        for (auto& b : pr.blocks) tmax = std::max(tmax, (size_t)(b.w - pr.d.P + 1) * b.h * b.l);
        RSF_TRY(ensure_ws(c, c->tmp1, tmax * sizeof(float)));
        RSF_TRY(ensure_ws(c, c->tmp2, tmax * sizeof(float)));
    }
    return RSF_OK;
}

// The float copy and image blur of the workspace are about to be overwritten:
// whatever preparation the held image had there is gone.
void invalidate_held(rsf_context* c) { c->held_key.clear(); }

// Once per run: image as float and, in fast mode, the padded blur of the image
// (c->tr is scratch meanwhile).
int prepare_image_only(rsf_context* c, const Prepared& pr, const void* image, int u8) {
    if (pr.literal && pr.sw == 0) return RSF_OK;
    // a held image keeps its preparation while the layout, taps, mode and workspace match
    std::string key;
    if (image != c->held_image) {
        invalidate_held(c);
    } else {
        std::ostringstream os;
        os << blocks_key(pr.blocks) << '|' << u8 << '|' << pr.literal << '|' << pr.sw << '|' << c->ws_epoch << '|';
        for (int k = 0; k < pr.d.P; ++k) {
            uint32_t bits;
            std::memcpy(&bits, &pr.d.taps.g[k], 4);
            os << std::hex << bits << ',';
        }
        key = os.str();
        if (key == c->held_key) return RSF_OK;
        c->held_key.clear();
    }
    const bool split = split_plan(pr);
    if (pr.sw > 0 || (!pr.literal && !split))
        RSF_TRY(ops::to_float(c->imgf.as<float>(), image, u8, pr.extent, c->stream));
    if (!pr.literal && split) {
        // blur(I) through the plan's own xy and z kernels
        RSF_TRY(launch_blur_image(pr.runner->plan, image, u8, c->tr.as<float>(), c->blurI.as<float>(), pr.d,
                                  c->stream));
    } else if (!pr.literal) {
        for (auto& b : pr.blocks)
            RSF_TRY(ops::blur_pad(c->blurI.as<float>() + b.off, c->imgf.as<float>() + b.off, c->tmp1.as<float>(),
                                  c->tmp2.as<float>(), b.w, b.h, b.l, pr.d.taps, pr.d.P, pr.d.m, c->stream));
    }
    if (!key.empty()) c->held_key = key;
    return RSF_OK;
}

// Once per run: the image's preparation, then the transform T(phi) into c->tr.
int prepare_image(rsf_context* c, const Prepared& pr, const void* image, int u8, const float* phi, float eps) {
    RSF_TRY(prepare_image_only(c, pr, image, u8));
    return launch_transform(c->tr.as<float>(), phi, pr.extent, pr.literal, eps, c->stream);
}

// One sigma2 = 0 iteration on explicit buffers: tr holds T(phi_in) and gets
// T(phi_out); F is the force; blurI the image blur (fast mode only).
int enqueue_step_plain(const TilePlan& tp, const Derived& d, const void* image, int u8, const float* blurI,
                       float* tr, float* F, const float* phi_in, float* phi_out, const rsf_params* p,
                       hipStream_t s) {
    SweepArgs a{};
    a.tr = tr;
    a.img = image;
    a.img_u8 = u8;
    a.blurI = blurI;
    a.out0 = F;
    a.out_fits = 0;
    RSF_TRY(launch_sweep1(tp, a, d, p, s));
    a.phi = phi_in;
    a.F = F;
    a.phi_out = phi_out;
    a.tr_out = tr;
    return launch_sweep2(tp, a, p, s);
}

// One iteration phi_in -> phi_out over the prepared blocks (tr must hold the
// transform of phi_in; it is overwritten with the transform of phi_out).
int enqueue_step(rsf_context* c, const Prepared& pr, const void* image, int u8, const float* phi_in,
                 float* phi_out, const rsf_params* p) {
    const TilePlan& tp = pr.runner->plan;
    if (pr.sw == 0)
        return enqueue_step_plain(tp, pr.d, image, u8, c->blurI.as<float>(), c->tr.as<float>(), c->F.as<float>(),
                                  phi_in, phi_out, p, c->stream);
    SweepArgs a{};
    a.tr = c->tr.as<float>();
    a.img = image;
    a.img_u8 = u8;
    a.blurI = c->blurI.as<float>();
    {
        a.out0 = c->fo.as<float>();
        a.out1 = c->fi.as<float>();
        a.out_fits = 1;
        RSF_TRY(launch_sweep1(tp, a, pr.d, p, c->stream));
        for (auto& b : pr.blocks)
            RSF_TRY(ops::energies(c->Eo.as<float>() + b.off, c->Ei.as<float>() + b.off, c->fo.as<float>() + b.off,
                                  c->fi.as<float>() + b.off, c->imgf.as<float>() + b.off, b.w, b.h, b.l, pr.sw,
                                  c->stream));
        RSF_TRY(ops::force(c->F.as<float>(), c->Eo.as<float>(), c->Ei.as<float>(), p->lambda1, p->lambda2,
                           pr.extent, c->stream));
    }
    a.phi = phi_in;
    a.F = c->F.as<float>();
    a.phi_out = phi_out;
    a.tr_out = c->tr.as<float>();
    return launch_sweep2(tp, a, p, c->stream);
}

// Captures n iterations ping-ponging between A (input) and B; when n is odd
// and copy_back is set, a device copy brings the result back into A.
int capture(rsf_context* c, const Prepared& pr, const void* image, int u8, float* A0, float* B0,
            const rsf_params* p, int n, bool copy_back, hipGraphExec_t* out) {
    hipGraph_t g = nullptr;
    RSF_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int st = RSF_OK;
    float* A = A0;
    float* B = B0;
    for (int t = 0; t < n && st == RSF_OK; ++t) {
        st = enqueue_step(c, pr, image, u8, A, B, p);
        std::swap(A, B);
    }
    if (st == RSF_OK && copy_back && A != A0)
        st = (hipMemcpyAsync(A0, A, (size_t)pr.extent * sizeof(float), hipMemcpyDeviceToDevice, c->stream) ==
              hipSuccess) ? RSF_OK : set_error(RSF_ERR_DEVICE, "hipMemcpyAsync in capture failed");
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (st != RSF_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
    return RSF_OK;
}

// Voxels of phi (inside the blocks) whose exponent bits are all ones: NaN or +-inf.
__global__ void k_nonfinite(const float* __restrict__ phi, const BlockDesc* __restrict__ blocks, int nblocks,
                            unsigned long long* __restrict__ count) {
    unsigned long long bad = 0;
    auto nf = [](float v) { return (__float_as_uint(v) & 0x7f800000u) == 0x7f800000u ? 1u : 0u; };
    for (int b = blockIdx.y; b < nblocks; b += gridDim.y) {
        const BlockDesc bd = blocks[b];
        const int64_t n = (int64_t)bd.w * bd.h * bd.l;
        const float* p = phi + bd.off;
        const int64_t step = (int64_t)gridDim.x * blockDim.x;
        int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        int64_t done = 0;
        if (((uintptr_t)p & 15) == 0) {
            // 16-byte loads, four in flight per thread
            const float4* p4 = reinterpret_cast<const float4*>(p);
            const int64_t n4 = n >> 2;
            int64_t j = i;
            for (; j + 3 * step < n4; j += 4 * step) {
                float4 v[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) v[k] = p4[j + k * step];
#pragma unroll
                for (int k = 0; k < 4; ++k) bad += nf(v[k].x) + nf(v[k].y) + nf(v[k].z) + nf(v[k].w);
            }
            for (; j < n4; j += step) {
                const float4 v = p4[j];
                bad += nf(v.x) + nf(v.y) + nf(v.z) + nf(v.w);
            }
            done = n4 << 2;
        }
        for (i += done; i < n; i += step) bad += nf(p[i]);
    }
    if (bad) atomicAdd(count, bad);
}

// After a run: RSF_ERR_NONFINITE when phi holds NaN or inf (phi is left as computed).
int check_finite(rsf_context* c, const Prepared& pr, const float* phi) {
    RSF_TRY(c->nonfinite.ensure(sizeof(unsigned long long)));  // not a graph buffer: no epoch change
    RSF_HIP(hipMemsetAsync(c->nonfinite.p, 0, sizeof(unsigned long long), c->stream));
    int64_t nmax = 1;
    for (auto& b : pr.blocks) nmax = std::max(nmax, (int64_t)b.w * b.h * b.l);
    const dim3 grid((unsigned)std::min<int64_t>((nmax + 255) / 256, 2048),
                    (unsigned)std::min<size_t>(pr.blocks.size(), 65535));
    hipLaunchKernelGGL(k_nonfinite, grid, dim3(256), 0, c->stream, phi, pr.runner->plan.d_blocks.as<BlockDesc>(),
                       (int)pr.blocks.size(), c->nonfinite.as<unsigned long long>());
    RSF_HIP(hipGetLastError());
    unsigned long long bad = 0;
    RSF_HIP(hipMemcpyAsync(&bad, c->nonfinite.p, sizeof(bad), hipMemcpyDeviceToHost, c->stream));
    RSF_HIP(hipStreamSynchronize(c->stream));
    if (bad) return set_error(RSF_ERR_NONFINITE, "%llu voxels of phi are NaN or infinite after the run", bad);
    return RSF_OK;
}

int run_prepared(rsf_context* c, const Prepared& pr, const void* image, int u8, float* phi, const rsf_params* p,
                 float* iter_ms) {
    RSF_TRY(ensure_workspace(c, pr));
    RSF_TRY(prepare_image(c, pr, image, u8, phi, p->epsilon));
    const int T = p->iterations;
    if (T == 0) return RSF_OK;
    Runner* r = pr.runner;
    // The captured kernels hold the parameters by value: the key carries their
    // exact bit patterns (values that print alike must not share a graph).
    std::ostringstream gk;
    gk << image << '|' << u8 << '|' << (const void*)phi << '|' << std::hex;
    for (float v : {p->sigma, p->lambda1, p->lambda2, p->nu, p->mu, p->dt, p->epsilon, p->sigma2}) {
        uint32_t bits;
        std::memcpy(&bits, &v, 4);
        gk << bits << '|';
    }
    gk << std::dec << p->mode << '|' << T << '|' << (iter_ms ? 1 : 0);
    if (r->graph_key != gk.str() || r->ws_epoch != c->ws_epoch) {
        r->graphs.reset();
        r->graph_key.clear();
        float* B = c->phi2.as<float>();
        if (iter_ms) {
            // one graph per direction, launched alternately, events in between
            RSF_TRY(capture(c, pr, image, u8, phi, B, p, 1, false, &r->graphs.main));
            if (T > 1) RSF_TRY(capture(c, pr, image, u8, B, phi, p, 1, false, &r->graphs.rem));
            r->graphs.n_main = 1;
            r->graphs.n_launch = T;
        } else {
            const int chunk = std::min(T, 64) & ~1;  // even: each launch ends in phi
            if (chunk > 0) RSF_TRY(capture(c, pr, image, u8, phi, B, p, chunk, true, &r->graphs.main));
            r->graphs.n_main = chunk;
            r->graphs.n_launch = chunk > 0 ? T / chunk : 0;
            r->graphs.n_rem = chunk > 0 ? T % chunk : T;
            if (r->graphs.n_rem)
                RSF_TRY(capture(c, pr, image, u8, phi, B, p, r->graphs.n_rem, true, &r->graphs.rem));
        }
        r->graph_key = gk.str();
        r->ws_epoch = c->ws_epoch;
    }
    if (iter_ms) {
        if ((int)c->iter_events.size() < T + 1) {
            for (int k = (int)c->iter_events.size(); k < T + 1; ++k) {
                hipEvent_t e;
                RSF_HIP(hipEventCreate(&e));
                c->iter_events.push_back(e);
            }
        }
        RSF_HIP(hipEventRecord(c->iter_events[0], c->stream));
        for (int t = 0; t < T; ++t) {
            RSF_HIP(hipGraphLaunch((t & 1) ? r->graphs.rem : r->graphs.main, c->stream));
            RSF_HIP(hipEventRecord(c->iter_events[t + 1], c->stream));
        }
        if (T & 1)
            RSF_HIP(hipMemcpyAsync(phi, c->phi2.p, (size_t)pr.extent * sizeof(float), hipMemcpyDeviceToDevice,
                                   c->stream));
        RSF_HIP(hipStreamSynchronize(c->stream));
        for (int t = 0; t < T; ++t) RSF_HIP(hipEventElapsedTime(&iter_ms[t], c->iter_events[t], c->iter_events[t + 1]));
        RSF_HIP(hipEventElapsedTime(&c->last_ms, c->iter_events[0], c->iter_events[T]));
        return check_finite(c, pr, phi);
    }
    RSF_HIP(hipEventRecord(c->ev0, c->stream));
    for (int k = 0; k < r->graphs.n_launch; ++k) RSF_HIP(hipGraphLaunch(r->graphs.main, c->stream));
    if (r->graphs.rem) RSF_HIP(hipGraphLaunch(r->graphs.rem, c->stream));
    RSF_HIP(hipEventRecord(c->ev1, c->stream));
    RSF_HIP(hipEventSynchronize(c->ev1));
    RSF_HIP(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
    return check_finite(c, pr, phi);
}

int check_ctx(rsf_context* c) {
    if (!c) return set_error(RSF_ERR_INVALID_ARGUMENT, "context is NULL");
    RSF_HIP(hipSetDevice(c->device));
    return RSF_OK;
}

int check_dtype(int dtype) {
    if (dtype != RSF_DTYPE_U8 && dtype != RSF_DTYPE_F32)
        return set_error(RSF_ERR_INVALID_ARGUMENT, "image dtype must be RSF_DTYPE_U8 or RSF_DTYPE_F32");
    return RSF_OK;
}

int unfused_step(rsf_context* c, const float* I, const float* phi, float* out, int w, int h, int l,
                 const rsf_params* p, const Derived& d);

}  // namespace

int rsf::context_stream(rsf_context* c, hipStream_t* s) {
    RSF_TRY(check_ctx(c));
    *s = c->stream;
    return RSF_OK;
}

extern "C" {

int rsf_abi_version(void) { return RSF_ABI_VERSION; }

const char* rsf_last_error(void) { return g_last_error.c_str(); }

const char* rsf_status_string(int s) {
    switch (s) {
        case RSF_OK: return "ok";
        case RSF_ERR_INVALID_ARGUMENT: return "invalid argument";
        case RSF_ERR_UNSUPPORTED: return "unsupported";
        case RSF_ERR_OUT_OF_MEMORY: return "out of memory";
        case RSF_ERR_DEVICE: return "device error";
        case RSF_ERR_IO: return "i/o error";
        case RSF_ERR_NONFINITE: return "non-finite values";
        default: return "unknown status";
    }
}

void rsf_params_default(rsf_params* p) {
    if (!p) return;
    p->sigma = 5.0f;      // main.cpp:91
    p->iterations = 30;   // :92
    p->dt = 0.1f;         // :93
    p->mu = 1.0f;         // :94
    p->lambda1 = 0.1f;    // :95 (e)
    p->lambda2 = 0.1f;
    p->nu = 0.0001f;      // :96
    p->sigma2 = 0.0f;     // :97
    p->epsilon = 0.2f;    // kernel.cu:258
    p->mode = RSF_MODE_LITERAL;  // the reference's arithmetic; RSF_MODE_FAST is opt-in
}

int rsf_taps(float sigma, float* taps, int capacity, int* ntaps, int* pad) {
    if (!taps || !ntaps || !pad) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL output");
    return taps32(sigma, taps, capacity, ntaps, pad);
}

int rsf_context_create(int device, void* hip_stream, rsf_context** out) {
    if (!out) return set_error(RSF_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    int n = 0;
    RSF_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return set_error(RSF_ERR_INVALID_ARGUMENT, "device %d of %d", device, n);
    RSF_HIP(hipSetDevice(device));
    auto c = std::make_unique<rsf_context>();
    c->device = device;
    if (hip_stream) {
        c->stream = (hipStream_t)hip_stream;
    } else {
        RSF_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own_stream = true;
    }
    hipDeviceProp_t prop;
    RSF_HIP(hipGetDeviceProperties(&prop, device));
    c->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    RSF_HIP(hipEventCreate(&c->ev0));
    RSF_HIP(hipEventCreate(&c->ev1));
    RSF_TRY(scratch_address(device, &c->scratch));   // with `device` current
    load_tune_file(c.get());
    *out = c.release();
    return RSF_OK;
}

int rsf_context_destroy(rsf_context* c) {
    if (!c) return RSF_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->runners.clear();
    for (auto e : c->iter_events) (void)hipEventDestroy(e);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return RSF_OK;
}

int rsf_context_set_stream(rsf_context* c, void* hip_stream) {
    RSF_TRY(check_ctx(c));
    hipStream_t s = hip_stream ? (hipStream_t)hip_stream : nullptr;
    if (!s) return set_error(RSF_ERR_INVALID_ARGUMENT, "stream is NULL");
    if (s == c->stream) return RSF_OK;
    RSF_HIP(hipStreamSynchronize(c->stream));
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    c->stream = s;
    c->own_stream = false;
    return RSF_OK;
}

int rsf_context_hold_image(rsf_context* c, const void* image) {
    RSF_TRY(check_ctx(c));
    c->held_image = image;
    c->held_key.clear();
    return RSF_OK;
}

int rsf_context_release_image(rsf_context* c) {
    RSF_TRY(check_ctx(c));
    c->held_image = nullptr;
    c->held_key.clear();
    return RSF_OK;
}

int rsf_workspace_size(const rsf_params* p, int64_t w, int64_t h, int64_t l, int64_t* bytes) {
    if (!bytes) return set_error(RSF_ERR_INVALID_ARGUMENT, "bytes is NULL");
    Derived d;
    RSF_TRY(derive(p, &d));
    if (w < d.P || h < d.P || l < d.P) return set_error(RSF_ERR_INVALID_ARGUMENT, "dims smaller than the blur");
    const int64_t n = w * h * l;
    const int literal = p->mode == RSF_MODE_LITERAL;
    const int s1 = default_sweep1(d.P, literal);
    if (s1 < 0) {
        // no fused kernel for this tap count: the unfused pipeline's image copy,
        // phi ping-pong and 20 intermediate volumes
        *bytes = 22 * n * (int64_t)sizeof(float);
        return RSF_OK;
    }
    int64_t vols = 3;  // phi ping-pong, transform, force
    if (!literal) vols += 2;                       // blur(I), image as float
    if ((int)p->sigma2 > 0) vols += 4 + (literal ? 1 : 0);   // fo, fi, Eo, Ei (+ image as float)
    // the split sweep 1 holds its xy-blurred channels (4 literal, 2 fast); blocks of
    // 2^30 voxels and more run the fused kernel, which has none (pick_variants)
    const bool small = 4.0 * ((double)n + 4.0 * w * h + 4.0 * w + 64.0) < 4294967295.0;
    bool split = false;
    for (const auto& v : sweep1_variants()) split = split || (v.fn2 && v.P == d.P && v.literal == literal);
    if (small && split) vols += literal ? 4 : 2;
    // fast mode without a split sweep 1: the two pass temporaries of the per-op image
    // blur (fewer bytes than the split kernels' two xy channels)
    const int64_t tmp = (literal || (small && split)) ? 0 : 2 * (w - d.P + 1) * h * l;
    *bytes = (vols * n + tmp) * (int64_t)sizeof(float);
    return RSF_OK;
}

int rsf_variant_count(int sweep, int* count) {
    if (!count) return set_error(RSF_ERR_INVALID_ARGUMENT, "count is NULL");
    if (sweep == 1) *count = (int)sweep1_variants().size();
    else if (sweep == 2) *count = (int)sweep2_variants().size();
    else return set_error(RSF_ERR_INVALID_ARGUMENT, "sweep must be 1 or 2, got %d", sweep);
    return RSF_OK;
}

int rsf_variant_info(int sweep, int index, char* name, int name_len, int* taps, int* literal) {
    int n = 0;
    RSF_TRY(rsf_variant_count(sweep, &n));
    if (index < 0 || index >= n) return set_error(RSF_ERR_INVALID_ARGUMENT, "variant %d of %d", index, n);
    const char* nm = sweep == 1 ? sweep1_variants()[index].name : sweep2_variants()[index].name;
    if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", nm);
    if (taps) *taps = sweep == 1 ? sweep1_variants()[index].P : 0;
    if (literal) *literal = sweep == 1 ? sweep1_variants()[index].literal : sweep2_variants()[index].literal;
    return RSF_OK;
}

int rsf_last_loop_ms(rsf_context* c, float* ms) {
    if (!c || !ms) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL argument");
    *ms = c->last_ms;
    return RSF_OK;
}

int rsf_step_device(rsf_context* c, const void* image, int image_dtype, const float* phi_in, float* phi_out,
                    int64_t w, int64_t h, int64_t l, const rsf_params* p) {
    RSF_TRY(check_ctx(c));
    RSF_TRY(check_dtype(image_dtype));
    if (!image || !phi_in || !phi_out) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (phi_in == phi_out) return set_error(RSF_ERR_INVALID_ARGUMENT, "phi_in and phi_out must not alias");
    rsf_block b{0, w, h, l};
    return rsf_step_blocks_device(c, &b, 1, image, image_dtype, phi_in, phi_out, p);
}

int rsf_step_device_noalloc(rsf_context* c, const void* image, int image_dtype, const float* phi_in,
                            float* phi_out, int64_t w, int64_t h, int64_t l, const rsf_params* p) {
    RSF_TRY(check_ctx(c));
    RSF_TRY(check_dtype(image_dtype));
    if (!image || !phi_in || !phi_out) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (!c->scratch) return set_error(RSF_ERR_DEVICE, "no scratch address for device %d", c->device);
    Derived d;
    RSF_TRY(derive(p, &d));
    if (p->mode != RSF_MODE_LITERAL || (int)p->sigma2 != 0)
        return set_error(RSF_ERR_UNSUPPORTED, "the allocation-free step runs LITERAL mode with sigma2 = 0");
    if (w < d.P || h < d.P || l < d.P)
        return set_error(RSF_ERR_INVALID_ARGUMENT, "dims (%lld,%lld,%lld) are smaller than the %d-tap blur",
                         (long long)w, (long long)h, (long long)l, d.P);
    if (w > kNoAllocMaxVoxels || h > kNoAllocMaxVoxels || l > kNoAllocMaxVoxels || w * h * l > kNoAllocMaxVoxels)
        return set_error(RSF_ERR_INVALID_ARGUMENT,
                         "%lld x %lld x %lld voxels exceed the allocation-free step's scratch of %lld voxels",
                         (long long)w, (long long)h, (long long)l, (long long)kNoAllocMaxVoxels);
    const int64_t n = w * h * l;
    if (ranges_overlap(phi_in, phi_out, n)) return set_error(RSF_ERR_INVALID_ARGUMENT, "phi_in and phi_out overlap");
    const std::vector<BlockDesc> blocks = {BlockDesc{0, (int)w, (int)h, (int)l, 0}};
    int s1, s2, zc1, zc2;
    pick_variants(c, blocks, d, 1, aligned16({phi_in, phi_out}), &s1, &s2, &zc1, &zc2);
    if (s1 >= 0 && sweep1_variants()[s1].fn2) {
        // the fused sweep 1 keeps its xy-blurred channels on chip: no intermediate volume
        s1 = -1;
        for (size_t k = 0; k < sweep1_variants().size(); ++k) {
            const auto& v = sweep1_variants()[k];
            if (!v.fn2 && v.P == d.P && v.literal) {
                s1 = (int)k;
                break;
            }
        }
        if (s1 >= 0) {
            const auto& v = sweep1_variants()[s1];
            const int64_t cols = ((w + v.tx - 1) / v.tx) * ((h + v.ty - 1) / v.ty);
            zc1 = (int)l;
            while (zc1 > 4 * d.P && cols * ((l + zc1 - 1) / zc1) < 2 * (int64_t)c->num_cus) zc1 = (zc1 + 1) / 2;
        }
    }
    if (s1 < 0 || s2 < 0)
        return set_error(RSF_ERR_UNSUPPORTED, "tap count %d has no fused kernel for the allocation-free step", d.P);
    TilePlan tp;
    RSF_TRY(plan_tiles_host(tp, blocks, s1, s2, zc1, zc2, c->num_cus));
    // scratch layout: T(phi) | F | tile plan, each on a 256-byte boundary
    const size_t vol = ((size_t)n * sizeof(float) + 255) & ~size_t(255);
    float* tr = reinterpret_cast<float*>(c->scratch);
    float* F = reinterpret_cast<float*>(c->scratch + vol);
    unsigned char* arena = c->scratch + 2 * vol;
    const size_t arena_bytes = kNoAllocBytes - 2 * vol;
    const hipStream_t s = c->stream;
    RSF_TRY(plan_upload(tp, arena, arena_bytes, s));
    RSF_TRY(launch_transform(tr, phi_in, n, 1, p->epsilon, s));
    RSF_TRY(enqueue_step_plain(tp, d, image, image_dtype == RSF_DTYPE_U8, nullptr, tr, F, phi_in, phi_out, p, s));
    // also keeps the plan's host tile lists alive until their copies have run
    RSF_HIP(hipStreamSynchronize(s));
    return RSF_OK;
}

int rsf_step_blocks_device(rsf_context* c, const rsf_block* blocks, int nblocks, const void* image,
                           int image_dtype, const float* phi_in, float* phi_out, const rsf_params* p) {
    RSF_TRY(check_ctx(c));
    RSF_TRY(check_dtype(image_dtype));
    if (!image || !phi_in || !phi_out) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL buffer");
    Prepared pr;
    RSF_TRY(prepare(c, blocks, nblocks, p, aligned16({phi_in, phi_out}), &pr));
    if (ranges_overlap(phi_in, phi_out, pr.extent))
        return set_error(RSF_ERR_INVALID_ARGUMENT, "phi_in and phi_out overlap: sweep 2 reads neighbours of "
                                                   "voxels other workgroups are writing");
    const int u8 = image_dtype == RSF_DTYPE_U8;
    if (!pr.fused) {
        if (nblocks != 1) return set_error(RSF_ERR_UNSUPPORTED, "tap count %d has no fused kernel", pr.d.P);
        const BlockDesc& b = pr.blocks[0];
        RSF_TRY(ensure_ws(c, c->imgf, (size_t)pr.extent * sizeof(float)));
        invalidate_held(c);
        RSF_TRY(ops::to_float(c->imgf.as<float>(), image, u8, pr.extent, c->stream));
        RSF_TRY(unfused_step(c, c->imgf.as<float>() + b.off, phi_in + b.off, phi_out + b.off, b.w, b.h, b.l, p, pr.d));
        RSF_HIP(hipStreamSynchronize(c->stream));
        return RSF_OK;
    }
    RSF_TRY(ensure_workspace(c, pr));
    RSF_TRY(prepare_image(c, pr, image, u8, phi_in, p->epsilon));
    RSF_TRY(enqueue_step(c, pr, image, u8, phi_in, phi_out, p));
    RSF_HIP(hipStreamSynchronize(c->stream));
    return RSF_OK;
}

int rsf_run_device(rsf_context* c, const void* image, int image_dtype, float* phi, int64_t w, int64_t h,
                   int64_t l, const rsf_params* p, float* iter_ms) {
    RSF_TRY(check_ctx(c));
    RSF_TRY(check_dtype(image_dtype));
    if (!image || !phi) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL buffer");
    rsf_block b{0, w, h, l};
    Prepared pr;
    RSF_TRY(prepare(c, &b, 1, p, aligned16({phi}), &pr));
    const int u8 = image_dtype == RSF_DTYPE_U8;
    if (!pr.fused) {
        // unfused pipeline for tap counts without a fused specialisation
        RSF_TRY(ensure_ws(c, c->imgf, (size_t)pr.extent * sizeof(float)));
        RSF_TRY(ensure_ws(c, c->phi2, (size_t)pr.extent * sizeof(float)));
        invalidate_held(c);
        RSF_TRY(ops::to_float(c->imgf.as<float>(), image, u8, pr.extent, c->stream));
        float* A = phi;
        float* B = c->phi2.as<float>();
        RSF_HIP(hipEventRecord(c->ev0, c->stream));
        for (int t = 0; t < p->iterations; ++t) {
            if (iter_ms) RSF_HIP(hipEventRecord(c->ev0, c->stream));
            RSF_TRY(unfused_step(c, c->imgf.as<float>(), A, B, (int)w, (int)h, (int)l, p, pr.d));
            std::swap(A, B);
            if (iter_ms) {
                RSF_HIP(hipEventRecord(c->ev1, c->stream));
                RSF_HIP(hipEventSynchronize(c->ev1));
                RSF_HIP(hipEventElapsedTime(&iter_ms[t], c->ev0, c->ev1));
            }
        }
        if (A != phi) RSF_HIP(hipMemcpyAsync(phi, A, (size_t)pr.extent * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        RSF_HIP(hipStreamSynchronize(c->stream));
        return RSF_OK;
    }
    return run_prepared(c, pr, image, u8, phi, p, iter_ms);
}

int rsf_run_host(rsf_context* c, const void* image, int image_dtype, float* phi, int64_t w, int64_t h, int64_t l,
                 const rsf_params* p, float* iter_ms) {
    RSF_TRY(check_ctx(c));
    RSF_TRY(check_dtype(image_dtype));
    if (!image || !phi) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (w <= 0 || h <= 0 || l <= 0) return set_error(RSF_ERR_INVALID_ARGUMENT, "dims must be positive");
    const int64_t n = w * h * l;
    const size_t ib = (size_t)n * (image_dtype == RSF_DTYPE_U8 ? 1 : 4);
    RSF_TRY(c->h_img.ensure(ib));
    RSF_TRY(c->h_phi.ensure((size_t)n * sizeof(float)));
    RSF_HIP(hipMemcpyAsync(c->h_img.p, image, ib, hipMemcpyHostToDevice, c->stream));
    RSF_HIP(hipMemcpyAsync(c->h_phi.p, phi, (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    RSF_TRY(rsf_run_device(c, c->h_img.p, image_dtype, c->h_phi.as<float>(), w, h, l, p, iter_ms));
    RSF_HIP(hipMemcpyAsync(phi, c->h_phi.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    RSF_HIP(hipStreamSynchronize(c->stream));
    return RSF_OK;
}

int rsf_run_batched_device(rsf_context* c, const rsf_block* blocks, int nblocks, const void* image, int image_dtype,
                           float* phi, const rsf_params* p) {
    RSF_TRY(check_ctx(c));
    RSF_TRY(check_dtype(image_dtype));
    if (!image || !phi) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL buffer");
    Prepared pr;
    RSF_TRY(prepare(c, blocks, nblocks, p, aligned16({phi}), &pr));
    if (!pr.fused) return set_error(RSF_ERR_UNSUPPORTED, "tap count %d has no fused kernel for batched runs", pr.d.P);
    return run_prepared(c, pr, image, image_dtype == RSF_DTYPE_U8, phi, p, nullptr);
}

int rsf_profile_sweeps(rsf_context* c, const rsf_block* blocks, int nblocks, const void* image, int image_dtype,
                       float* phi, const rsf_params* p, float* ms_sweep1, float* ms_sweep2) {
    RSF_TRY(check_ctx(c));
    RSF_TRY(check_dtype(image_dtype));
    if (!image || !phi || !ms_sweep1 || !ms_sweep2) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL argument");
    Prepared pr;
    RSF_TRY(prepare(c, blocks, nblocks, p, aligned16({phi}), &pr));
    if (!pr.fused || pr.sw != 0) return set_error(RSF_ERR_UNSUPPORTED, "profiling covers the fused sigma2 = 0 path");
    const int u8 = image_dtype == RSF_DTYPE_U8;
    RSF_TRY(ensure_workspace(c, pr));
    RSF_TRY(prepare_image(c, pr, image, u8, phi, p->epsilon));
    const int T = std::max(1, (int)p->iterations);
    std::vector<hipEvent_t> ev(3 * T);
    for (auto& e : ev) RSF_HIP(hipEventCreate(&e));
    const TilePlan& tp = pr.runner->plan;
    float* A = phi;
    float* B = c->phi2.as<float>();
    int st = RSF_OK;
    for (int t = 0; t < T && st == RSF_OK; ++t) {
        SweepArgs a{};
        a.tr = c->tr.as<float>();
        a.img = image;
        a.img_u8 = u8;
        a.blurI = c->blurI.as<float>();
        a.out0 = c->F.as<float>();
        a.phi = A;
        a.F = c->F.as<float>();
        a.phi_out = B;
        a.tr_out = c->tr.as<float>();
        (void)hipEventRecord(ev[3 * t], c->stream);
        st = launch_sweep1(tp, a, pr.d, p, c->stream);
        (void)hipEventRecord(ev[3 * t + 1], c->stream);
        if (st == RSF_OK) st = launch_sweep2(tp, a, p, c->stream);
        (void)hipEventRecord(ev[3 * t + 2], c->stream);
        std::swap(A, B);
    }
    if (st == RSF_OK && A != phi)
        st = hipMemcpyAsync(phi, A, (size_t)pr.extent * sizeof(float), hipMemcpyDeviceToDevice, c->stream) == hipSuccess
                 ? RSF_OK : set_error(RSF_ERR_DEVICE, "copy back failed");
    (void)hipStreamSynchronize(c->stream);
    double s1 = 0, s2 = 0;
    for (int t = 0; t < T && st == RSF_OK; ++t) {
        float a = 0, b = 0;
        (void)hipEventElapsedTime(&a, ev[3 * t], ev[3 * t + 1]);
        (void)hipEventElapsedTime(&b, ev[3 * t + 1], ev[3 * t + 2]);
        s1 += a;
        s2 += b;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    RSF_TRY(st);
    *ms_sweep1 = (float)(s1 / T);
    *ms_sweep2 = (float)(s2 / T);
    return RSF_OK;
}

int rsf_run_blocks_sequential_device(rsf_context* c, const rsf_block* blocks, int nblocks, const void* image,
                                     int image_dtype, float* phi, const rsf_params* p) {
    RSF_TRY(check_ctx(c));
    RSF_TRY(check_dtype(image_dtype));
    if (!blocks || nblocks <= 0) return set_error(RSF_ERR_INVALID_ARGUMENT, "no blocks");
    const size_t esz = image_dtype == RSF_DTYPE_U8 ? 1 : 4;
    float total = 0.f;
    for (int b = 0; b < nblocks; ++b) {
        const rsf_block& x = blocks[b];
        RSF_TRY(rsf_run_device(c, (const char*)image + x.offset * esz, image_dtype, phi + x.offset, x.w, x.h, x.l, p,
                               nullptr));
        total += c->last_ms;
    }
    c->last_ms = total;
    return RSF_OK;
}

int rsf_autotune(rsf_context* c, const rsf_block* blocks, int nblocks, int image_dtype, const rsf_params* p,
                 char* choice, int choice_len) {
    RSF_TRY(check_ctx(c));
    RSF_TRY(check_dtype(image_dtype));
    Derived d;
    RSF_TRY(derive(p, &d));
    std::vector<BlockDesc> bl;
    int64_t extent = 0;
    RSF_TRY(validate_blocks(blocks, nblocks, d, &bl, &extent));
    const int literal = p->mode == RSF_MODE_LITERAL;
    const size_t n = (size_t)extent;
    // synthetic operands: a ramp level set and a constant image
    DevBuf img, phi, out, tr, F, fo, fi, bI;
    RSF_TRY(img.ensure(n * (image_dtype == RSF_DTYPE_U8 ? 1 : 4)));
    RSF_TRY(phi.ensure(n * 4));
    RSF_TRY(out.ensure(n * 4));
    RSF_TRY(tr.ensure(n * 4));
    RSF_TRY(F.ensure(n * 4));
    RSF_TRY(bI.ensure(n * 4));
    RSF_HIP(hipMemsetAsync(img.p, 0x40, img.bytes, c->stream));
    RSF_HIP(hipMemsetAsync(phi.p, 0, phi.bytes, c->stream));
    RSF_HIP(hipMemsetAsync(tr.p, 0, tr.bytes, c->stream));
    RSF_HIP(hipMemsetAsync(bI.p, 0, bI.bytes, c->stream));
    const int lmax = max_dim_l(bl);
    auto time_it = [&](auto launch) -> float {
        for (int k = 0; k < 2; ++k)
            if (launch() != RSF_OK) return 1e30f;
        if (hipEventRecord(c->ev0, c->stream) != hipSuccess) return 1e30f;
        for (int k = 0; k < 5; ++k)
            if (launch() != RSF_OK) return 1e30f;
        (void)hipEventRecord(c->ev1, c->stream);
        (void)hipEventSynchronize(c->ev1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
        return ms / 5;
    };
    SweepArgs a{};
    a.tr = tr.as<float>();
    a.img = img.p;
    a.img_u8 = image_dtype == RSF_DTYPE_U8;
    a.blurI = bI.as<float>();
    a.out0 = F.as<float>();
    a.phi = phi.as<float>();
    a.F = F.as<float>();
    a.phi_out = out.as<float>();
    a.tr_out = tr.as<float>();
    const auto& V1 = sweep1_variants();
    const auto& V2 = sweep2_variants();
    float best1 = 1e30f, best2 = 1e30f;
    TuneChoice tc;
    for (size_t k = 0; k < V1.size(); ++k) {
        if (V1[k].P != d.P || V1[k].literal != literal) continue;
        // strip xy kernels: a plane or two per wave of a work item
        const std::vector<int> zcs = V1[k].strip > 0 ? std::vector<int>{V1[k].strip, 2 * V1[k].strip}
                                                     : std::vector<int>{32, 64, 128, 256, lmax};
        for (int zc : zcs) {
            if (zc > lmax) continue;
            TilePlan tp;
            if (plan_tiles(tp, bl, (int)k, default_sweep2(literal), zc, std::min(lmax, 64), c->num_cus) != RSF_OK)
                continue;
            const float ms = time_it([&] { return launch_sweep1(tp, a, d, p, c->stream); });
            if (ms < best1) { best1 = ms; tc.s1 = V1[k].name; tc.zc1 = zc; }
        }
    }
    std::vector<int> zc2s = {8, 16, 32, 64, lmax};
    for (size_t k = 0; k < V2.size(); ++k) {
        if (V2[k].literal != literal) continue;
        for (int zc : zc2s) {
            if (zc > lmax) continue;
            const int s1 = default_sweep1(d.P, literal);
            if (s1 < 0) break;
            TilePlan tp;
            if (plan_tiles(tp, bl, s1, (int)k, std::min(lmax, 128), zc, c->num_cus) != RSF_OK) continue;
            const float ms = time_it([&] { return launch_sweep2(tp, a, p, c->stream); });
            if (ms < best2) { best2 = ms; tc.s2 = V2[k].name; tc.zc2 = zc; }
        }
    }
    RSF_HIP(hipStreamSynchronize(c->stream));
    if (tc.s1.empty() || tc.s2.empty()) return set_error(RSF_ERR_UNSUPPORTED, "no fused variant for P=%d", d.P);
    c->tune[tune_key(bl, d.P, literal)] = tc;
    save_tune_file(c);
    if (choice && choice_len > 0) {
        snprintf(choice, (size_t)choice_len, "sweep1=%s zc=%d (%.3f ms) sweep2=%s zc=%d (%.3f ms)", tc.s1.c_str(),
                 tc.zc1, best1, tc.s2.c_str(), tc.zc2, best2);
    }
    return RSF_OK;
}

int rsf_step_unfused_device(rsf_context* c, const float* image, const float* phi_in, float* phi_out, int64_t w,
                            int64_t h, int64_t l, const rsf_params* p) {
    RSF_TRY(check_ctx(c));
    if (!image || !phi_in || !phi_out) return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL buffer");
    Derived d;
    RSF_TRY(derive(p, &d));
    if (w < d.P || h < d.P || l < d.P) return set_error(RSF_ERR_INVALID_ARGUMENT, "dims smaller than the blur");
    RSF_TRY(unfused_step(c, image, phi_in, phi_out, (int)w, (int)h, (int)l, p, d));
    RSF_HIP(hipStreamSynchronize(c->stream));
    return RSF_OK;
}

int rsf_step_adjoint_device(rsf_context* c, const rsf_block* blocks, int nblocks, const void* image, int image_dtype,
                            const float* phi, const float* grad_out, float* grad_in, float* grad_params,
                            const rsf_params* p) {
    RSF_TRY(check_ctx(c));
    RSF_TRY(check_dtype(image_dtype));
    if (!image || !phi || !grad_out || !grad_in || !grad_params)
        return set_error(RSF_ERR_INVALID_ARGUMENT, "NULL buffer");
    Prepared pr;
    RSF_TRY(prepare(c, blocks, nblocks, p, aligned16({phi}), &pr));
    if (ranges_overlap(grad_in, grad_out, pr.extent) || ranges_overlap(grad_in, phi, pr.extent))
        return set_error(RSF_ERR_INVALID_ARGUMENT, "grad_in must not overlap grad_out or phi");
    if (!pr.fused) return set_error(RSF_ERR_UNSUPPORTED, "tap count %d has no fused kernel", pr.d.P);
    if (pr.sw != 0) return set_error(RSF_ERR_UNSUPPORTED, "the device adjoint covers sigma2 = 0");
    const int u8 = image_dtype == RSF_DTYPE_U8;
    RSF_TRY(ensure_workspace(c, pr));
    const size_t nb = (size_t)pr.extent * sizeof(float);
    for (int k = 0; k < 13; ++k) RSF_TRY(ensure_ws(c, c->adj[k], nb));
    int rows = 0;
    for (auto& b : pr.blocks) rows += adj::finish_rows((int64_t)b.w * b.h * b.l);
    RSF_TRY(ensure_ws(c, c->adj[13], (size_t)rows * 4 * sizeof(float)));
    float* A[14];
    for (int k = 0; k < 14; ++k) A[k] = c->adj[k].as<float>();
    float *fo = A[0], *fi = A[1], *PH = A[2], *PHc = A[3], *QH = A[4], *QA = A[5], *Bh = A[6], *Ba = A[7];
    float *Gx = A[8], *Gy = A[9], *Gz = A[10], *t1 = A[11], *t2 = A[12], *parts = A[13];
    hipStream_t s = c->stream;
    RSF_TRY(prepare_image(c, pr, image, u8, phi, p->epsilon));
    SweepArgs a{};
    a.tr = c->tr.as<float>();
    a.img = image;
    a.img_u8 = u8;
    a.blurI = c->blurI.as<float>();
    a.out0 = fo;
    a.out1 = fi;
    a.out2 = PH;
    a.out3 = PHc;
    a.out_fits = 2;
    RSF_TRY(launch_sweep1(pr.runner->plan, a, pr.d, p, s));
    const size_t esz = u8 ? 1 : 4;
    int row = 0;
    // blocks of one shape, each starting on a 4-voxel boundary: every kernel once
    // for all of them (the same bits as block by block)
    const BlockDesc& b0 = pr.blocks[0];
    const int64_t n0 = (int64_t)b0.w * b0.h * b0.l;
    bool batch = pr.blocks.size() > 1;
    for (auto& b : pr.blocks) batch = batch && b.w == b0.w && b.h == b0.h && b.l == b0.l && b.off % 4 == 0;
    if (batch && adj::batch_ok(pr.d.P, b0.w, n0, image, u8,
                               {fo, fi, PH, PHc, QH, QA, Bh, Ba, Gx, Gy, Gz, t1, t2, phi, grad_out, grad_in})) {
        const BlockDesc* bt = pr.runner->plan.d_blocks.as<BlockDesc>();
        const int nb = (int)pr.blocks.size();
        RSF_TRY(adj::q_blur_transpose2(Bh, Ba, QH, QA, t1, t2, phi, grad_out, image, u8, fo, fi, PH, PHc, b0.w, b0.h,
                                       b0.l, pr.d, p, s, bt, nb));
        // the quad finishing kernel forms the forward gradients from phi: none stored
        const adj::FwdStencil fwd{{PH, PHc, QH}, {nullptr, nullptr, nullptr}};
        RSF_TRY(adj::stencil_g(Gx, Gy, Gz, fwd, phi, grad_out, b0.w, b0.h, b0.l, p, s, bt, nb));
        RSF_TRY(adj::finish(grad_in, parts, phi, grad_out, Gx, Gy, Gz, fwd, Bh, Ba, image, u8, fo, fi, b0.w, b0.h,
                            b0.l, p, s, bt, nb));
        RSF_TRY(adj::reduce_params(grad_params, parts, nb * adj::finish_rows(n0), s));
        RSF_HIP(hipStreamSynchronize(s));
        return RSF_OK;
    }
    for (auto& b : pr.blocks) {
        const int64_t o = b.off;
        const void* im = (const char*)image + o * esz;
        RSF_TRY(adj::q_blur_transpose2(Bh + o, Ba + o, QH + o, QA + o, t1 + o, t2 + o, phi + o, grad_out + o, im, u8,
                                       fo + o, fi + o, PH + o, PHc + o, b.w, b.h, b.l, pr.d, p, s));
        // the forward normals and gradients into the buffers free by now (PH, PHc:
        // read by the point-wise step; QH, QA, t1, t2: by the transposed blur); the
        // gradients only for the one-voxel kernels, where some buffer is not aligned
        // for the quad ones
        const bool quads = adj::quads_ok(b.w, (int64_t)b.w * b.h * b.l, im, u8,
                                         {Gx + o, Gy + o, Gz + o, PH + o, PHc + o, QH + o, phi + o, grad_out + o,
                                          grad_in + o, Bh + o, Ba + o, fo + o, fi + o});
        const adj::FwdStencil fwd = quads ? adj::FwdStencil{{PH + o, PHc + o, QH + o}, {nullptr, nullptr, nullptr}}
                                          : adj::FwdStencil{{PH + o, PHc + o, QH + o}, {QA + o, t1 + o, t2 + o}};
        RSF_TRY(adj::stencil_g(Gx + o, Gy + o, Gz + o, fwd, phi + o, grad_out + o, b.w, b.h, b.l, p, s));
        RSF_TRY(adj::finish(grad_in + o, parts + 4 * row, phi + o, grad_out + o, Gx + o, Gy + o, Gz + o, fwd, Bh + o,
                            Ba + o, im, u8, fo + o, fi + o, b.w, b.h, b.l, p, s));
        row += adj::finish_rows((int64_t)b.w * b.h * b.l);
    }
    RSF_TRY(adj::reduce_params(grad_params, parts, row, s));
    RSF_HIP(hipStreamSynchronize(s));
    return RSF_OK;
}

}  // extern "C"

namespace {

// main.cpp:234-300 op by op, FP32-literal, on the device (kernels of rsf_ops.hip).
int unfused_step(rsf_context* c, const float* I, const float* phi, float* out, int w, int h, int l,
                 const rsf_params* p, const Derived& d) {
    const int64_t n = (int64_t)w * h * l;
    const size_t nb = (size_t)n * sizeof(float);
    for (int k = 0; k < 20; ++k) RSF_TRY(ensure_ws(c, c->unf[k], nb));
    float* U[20];
    for (int k = 0; k < 20; ++k) U[k] = c->unf[k].as<float>();
    float *H = U[0], *A = U[1], *Hc = U[2], *B = U[3], *gx = U[4], *gy = U[5], *gz = U[6], *M = U[7];
    float *nx = U[8], *ny = U[9], *nz = U[10], *tx = U[11], *ty = U[12], *tz = U[13], *kap = U[14], *L = U[15];
    float *p1 = U[16], *p2 = U[17], *fo = U[18], *fi = U[19];
    hipStream_t s = c->stream;
    const int P = d.P, m = d.m;
    RSF_TRY(ops::heaviside(H, phi, n, p->epsilon, s));
    RSF_TRY(ops::multiplication(A, I, H, n, s));
    RSF_TRY(ops::inverse(Hc, H, n, s));
    RSF_TRY(ops::multiplication(B, I, Hc, n, s));
    RSF_TRY(ops::gradient(gx, gy, gz, phi, w, h, l, s));
    RSF_TRY(ops::gradientmagnitude(M, gx, gy, gz, n, s));
    RSF_TRY(ops::division_grad(nx, gx, M, n, s));
    RSF_TRY(ops::division_grad(ny, gy, M, n, s));
    RSF_TRY(ops::division_grad(nz, gz, M, n, s));
    RSF_TRY(ops::gradient(tx, nullptr, nullptr, nx, w, h, l, s));
    RSF_TRY(ops::gradient(nullptr, ty, nullptr, ny, w, h, l, s));
    RSF_TRY(ops::gradient(nullptr, nullptr, tz, nz, w, h, l, s));
    RSF_TRY(ops::plus(kap, tx, ty, tz, n, s));
    RSF_TRY(ops::gradient(tx, nullptr, nullptr, gx, w, h, l, s));
    RSF_TRY(ops::gradient(nullptr, ty, nullptr, gy, w, h, l, s));
    RSF_TRY(ops::gradient(nullptr, nullptr, tz, gz, w, h, l, s));
    RSF_TRY(ops::plus(L, tx, ty, tz, n, s));
    // blurs: reuse gx/gy (no longer needed) as pass temporaries, M/nx as the blurred outputs
    RSF_TRY(ops::blur_pad(p1, A, gx, gy, w, h, l, d.taps, P, m, s));
    RSF_TRY(ops::blur_pad(p2, H, gx, gy, w, h, l, d.taps, P, m, s));
    RSF_TRY(ops::division(fo, p1, p2, n, s));
    RSF_TRY(ops::blur_pad(p1, B, gx, gy, w, h, l, d.taps, P, m, s));
    RSF_TRY(ops::blur_pad(p2, Hc, gx, gy, w, h, l, d.taps, P, m, s));
    RSF_TRY(ops::division(fi, p1, p2, n, s));
    float *Eo = A, *Ei = B, *dh = Hc;
    RSF_TRY(ops::energies(Eo, Ei, fo, fi, I, w, h, l, (int)p->sigma2, s));
    RSF_TRY(ops::deri_heaviside(dh, phi, n, p->epsilon, s));
    return ops::final_phi(out, phi, dh, Eo, Ei, kap, L, p->lambda1, p->lambda2, p->dt, p->nu, p->mu, n, s);
}

}  // namespace
