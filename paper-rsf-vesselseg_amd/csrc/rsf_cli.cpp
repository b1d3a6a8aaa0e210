// rsf -- drop-in for the reference's rsf3d executable (CUDA_RSF/main.cpp:73-319).
//
// Same long options and defaults (main.cpp:89-102, Boost style: "--name value"
// or "--name=value", unambiguous prefixes accepted, no short options), same
// inputs (--binary seed .npy, --image .npy) and output (fp32 phi .npy of the
// image's shape, --ouput), same per-iteration stdout line (main.cpp:309) and
// the same exit code 1 for --help.  Additions: --output (alias of --ouput),
// --lambda1/--lambda2 (default --e), --epsilon (default 0.2, the reference's
// hard-coded value), --phi0 (take fp32 phi directly instead of sdf(seed-13)),
// --mode fast|literal, --device, and two block modes that replace the
// reference's per-block process loop (README.md:56-59):
//   --block N|W,H,L [--overlap K]  the volume is cut into overlapping blocks,
//       all blocks evolve in one batched launch per sweep, and phi is blended
//       back with linear ramps over the overlaps (interpolating_volumes.py);
//   --batch manifest  one "image seed output" line per block file (paths
//       relative to the manifest's folder): every block evolves in one batched
//       launch and its phi is written to its output, as one rsf3d per line would.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "npy.hpp"
#include "rsf.h"
#include "rsf_cli.h"

namespace {

struct Opt {
    const char* name;
    bool takes_value;
    const char* help;
    const char* def;
};

const Opt kOpts[] = {
    {"help", false, "produce help message", nullptr},
    {"sigma", true, "blur kernel for the fitting term (in pixels)", "5"},
    {"t", true, "total number of time steps", "30"},
    {"dt", true, "time step size", "0.1"},
    {"meu", true, "weight for the regularization term", "1"},
    {"e", true, "weight for the fitting term", "0.1"},
    {"v", true, "weight for the smoothing term", "0.0001"},
    {"sigma2", true, "weight for sigma2", "0"},
    {"binary", true, "binary seed surface (.npy)", nullptr},
    {"image", true, "input image filename (.npy)", nullptr},
    {"ouput", true, "output level set (.npy)", "out.npy"},
    {"output", true, "alias of --ouput", nullptr},
    {"debug", false, "output all debugging details", nullptr},
    {"lambda1", true, "weight of the outside fitting energy (default: --e)", nullptr},
    {"lambda2", true, "weight of the inside fitting energy (default: --e)", nullptr},
    {"epsilon", true, "Heaviside/Dirac width", "0.2"},
    {"phi0", true, "initial level set (.npy, used instead of sdf(binary - 13))", nullptr},
    {"mode", true, "literal (the reference's arithmetic, bit for bit) | fast (two blurred channels)", "literal"},
    {"device", true, "HIP device index", "0"},
    {"block", true, "run as overlapping blocks of this size: N or W,H,L", nullptr},
    {"overlap", true, "overlap of neighbouring blocks with --block", "32"},
    {"batch", true, "manifest of 'image seed output' lines, run in one batched launch", nullptr},
};
constexpr int kNumOpts = sizeof(kOpts) / sizeof(kOpts[0]);

void print_help() {
    std::cout << "Allowed options:\n";
    for (const Opt& o : kOpts) {
        std::string head = std::string("  --") + o.name;
        if (o.takes_value) head += o.def ? std::string(" arg (=") + o.def + ")" : " arg";
        if (head.size() < 40) head += std::string(40 - head.size(), ' ');
        std::cout << head << ' ' << o.help << '\n';
    }
    std::cout << std::endl;
}

int find_opt(const std::string& name, std::string* err) {
    for (int k = 0; k < kNumOpts; ++k)
        if (name == kOpts[k].name) return k;
    int hit = -1, hits = 0;
    for (int k = 0; k < kNumOpts; ++k)
        if (std::strncmp(kOpts[k].name, name.c_str(), name.size()) == 0) { hit = k; ++hits; }
    if (hits == 1) return hit;
    *err = hits ? "option '--" + name + "' is ambiguous" : "unrecognised option '--" + name + "'";
    return -1;
}

bool parse_float(const std::string& s, float* v) {
    char* end = nullptr;
    *v = std::strtof(s.c_str(), &end);
    return end && *end == '\0' && !s.empty();
}

int fail(const std::string& msg) {
    std::cerr << "rsf: " << msg << std::endl;
    return 1;
}

// Image as the library takes it: uint8 kept as is, anything else as fp32.
struct Image {
    rsf::NpyArray arr;
    std::vector<float> f32;
    bool u8 = false;
};

std::string load_image(const std::string& path, Image* im) {
    std::string err = rsf::npy_load(path, &im->arr);
    if (!err.empty()) return err;
    if (im->arr.shape.size() != 3) return path + ": the image must be a 3-D array (Z, Y, X)";
    im->u8 = im->arr.descr == "|u1" || im->arr.descr == "<u1";
    if (!im->u8) return rsf::npy_to_f32(im->arr, &im->f32);
    return "";
}

// The seed (or, with --phi0, the initial level set) as float32, shape-checked.
std::string load_seed(const std::string& path, const std::vector<int64_t>& shape, std::vector<float>* seed) {
    rsf::NpyArray s;
    std::string err = rsf::npy_load(path, &s);
    if (!err.empty()) return err;
    if (s.shape != shape) return path + ": initial level set and image shapes differ";
    return rsf::npy_to_f32(s, seed);
}

// phi0 = sdf(seed - 13) (main.cpp:174-183) on the device, in place; with --phi0
// the array is phi0 itself.
std::string make_phi0(rsf_context* ctx, bool is_phi0, const std::vector<int64_t>& shape, std::vector<float>* phi) {
    if (is_phi0) return "";
    std::vector<float> seed(phi->size());
    seed.swap(*phi);
    if (rsf_sdf_from_seed_host(ctx, seed.data(), phi->data(), shape[2], shape[1], shape[0], 13.0f) != RSF_OK)
        return rsf_last_error();
    return "";
}

// Destroys the context on every exit path.
struct ContextGuard {
    rsf_context* c = nullptr;
    ~ContextGuard() {
        if (c) rsf_context_destroy(c);
    }
};

void print_iterations(rsf_context* ctx, int T) {
    float ms = 0.f;
    rsf_last_loop_ms(ctx, &ms);
    for (int k = 0; k < T; ++k) std::cout << "It takes " << ms / T << " ms to calculate " << std::endl;
}

bool parse_dims(const std::string& s, int64_t d[3]) {
    std::vector<int64_t> v;
    std::stringstream ss(s);
    std::string tok;
    while (std::getline(ss, tok, ',')) {
        char* end = nullptr;
        const long long x = std::strtoll(tok.c_str(), &end, 10);
        if (tok.empty() || *end || x <= 0) return false;
        v.push_back(x);
    }
    if (v.size() == 1) v = {v[0], v[0], v[0]};
    if (v.size() != 3) return false;
    std::copy(v.begin(), v.end(), d);
    return true;
}

// --batch: every manifest line is one block (an rsf3d call of the reference's
// batch driver); all blocks are packed back to back and run together.
int run_batch(const std::string& manifest, bool is_phi0, const rsf_params& p, int device) {
    std::ifstream f(manifest);
    if (!f) return fail("cannot open manifest " + manifest);
    const size_t slash = manifest.find_last_of('/');
    const std::string dir = slash == std::string::npos ? "" : manifest.substr(0, slash + 1);
    auto resolve = [&](const std::string& q) { return q.empty() || q[0] == '/' ? q : dir + q; };
    struct Item {
        Image im;
        std::vector<float> phi;
        std::string out;
    };
    std::vector<Item> items;
    std::string line;
    int lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        std::stringstream ss(line);
        std::string a, b, c, extra;
        if (!(ss >> a) || a[0] == '#') continue;
        if (!(ss >> b >> c) || (ss >> extra))
            return fail(manifest + ":" + std::to_string(lineno) + ": expected 'image seed output'");
        items.emplace_back();
        Item& it = items.back();
        std::string err = load_image(resolve(a), &it.im);
        if (err.empty()) err = load_seed(resolve(b), it.im.arr.shape, &it.phi);
        if (!err.empty()) return fail(err);
        it.out = resolve(c);
    }
    if (items.empty()) return fail("manifest " + manifest + " lists no blocks");
    ContextGuard g;
    if (rsf_context_create(device, nullptr, &g.c) != RSF_OK) return fail(rsf_last_error());
    for (Item& it : items) {
        const std::string err = make_phi0(g.c, is_phi0, it.im.arr.shape, &it.phi);
        if (!err.empty()) return fail(err);
    }
    const bool u8 = std::all_of(items.begin(), items.end(), [](const Item& it) { return it.im.u8; });
    std::vector<rsf_block> blocks;
    int64_t total = 0;
    for (const Item& it : items) {
        const auto& sh = it.im.arr.shape;
        blocks.push_back(rsf_block{total, sh[2], sh[1], sh[0]});
        total += sh[0] * sh[1] * sh[2];
    }
    std::vector<uint8_t> img8;
    std::vector<float> imgf, phi((size_t)total);
    if (u8) img8.resize((size_t)total);
    else imgf.resize((size_t)total);
    for (size_t i = 0; i < items.size(); ++i) {
        const Item& it = items[i];
        const size_t off = (size_t)blocks[i].offset, n = it.phi.size();
        std::copy(it.phi.begin(), it.phi.end(), phi.begin() + off);
        if (u8) {
            std::copy(it.im.arr.data.begin(), it.im.arr.data.begin() + n, img8.begin() + off);
        } else if (it.im.u8) {
            for (size_t k = 0; k < n; ++k) imgf[off + k] = (float)it.im.arr.data[k];
        } else {
            std::copy(it.im.f32.begin(), it.im.f32.end(), imgf.begin() + off);
        }
    }
    const int st = rsf_run_batched_host(g.c, blocks.data(), (int)blocks.size(),
                                        u8 ? (const void*)img8.data() : (const void*)imgf.data(),
                                        u8 ? RSF_DTYPE_U8 : RSF_DTYPE_F32, phi.data(), &p);
    if (st == RSF_OK) print_iterations(g.c, p.iterations);
    if (st != RSF_OK) return fail(std::string(rsf_status_string(st)) + ": " + rsf_last_error());
    for (size_t i = 0; i < items.size(); ++i) {
        const std::string err = rsf::npy_save_f32(items[i].out, phi.data() + blocks[i].offset, items[i].im.arr.shape);
        if (!err.empty()) return fail(err);
    }
    return 0;
}

}  // namespace

namespace {

int cli(int argc, char** argv) {
    std::vector<std::string> val(kNumOpts);
    std::vector<bool> seen(kNumOpts, false);
    for (int a = 1; a < argc; ++a) {
        std::string tok = argv[a];
        if (tok.size() < 3 || tok[0] != '-' || tok[1] != '-') return fail("unexpected argument '" + tok + "'");
        std::string name = tok.substr(2), value;
        bool has_eq = false;
        const size_t eq = name.find('=');
        if (eq != std::string::npos) { value = name.substr(eq + 1); name = name.substr(0, eq); has_eq = true; }
        std::string err;
        const int k = find_opt(name, &err);
        if (k < 0) return fail(err);
        if (kOpts[k].takes_value && !has_eq) {
            if (a + 1 >= argc) return fail("option '--" + name + "' requires an argument");
            value = argv[++a];
        } else if (!kOpts[k].takes_value && has_eq) {
            return fail("option '--" + name + "' takes no argument");
        }
        if (seen[k]) return fail("option '--" + std::string(kOpts[k].name) + "' given more than once");
        seen[k] = true;
        val[k] = value;
    }
    auto idx = [](const char* n) {
        for (int k = 0; k < kNumOpts; ++k)
            if (std::strcmp(kOpts[k].name, n) == 0) return k;
        return -1;
    };
    if (seen[idx("help")]) {
        print_help();
        return 1;  // main.cpp:109-112
    }
    rsf_params p;
    rsf_params_default(&p);
    auto getf = [&](const char* n, float* dst) -> bool {
        const int k = idx(n);
        if (!seen[k]) return true;
        return parse_float(val[k], dst);
    };
    float e = 0.1f, t = 30.f;
    if (!getf("sigma", &p.sigma) || !getf("dt", &p.dt) || !getf("meu", &p.mu) || !getf("e", &e) ||
        !getf("v", &p.nu) || !getf("sigma2", &p.sigma2) || !getf("epsilon", &p.epsilon) || !getf("t", &t))
        return fail("invalid numeric option value");
    if (t != (float)(int)t || t < 0) return fail("--t must be a non-negative integer");
    p.iterations = (int)t;
    p.lambda1 = p.lambda2 = e;
    if (!getf("lambda1", &p.lambda1) || !getf("lambda2", &p.lambda2)) return fail("invalid lambda value");
    if (seen[idx("mode")]) {
        const std::string& m = val[idx("mode")];
        if (m == "fast") p.mode = RSF_MODE_FAST;
        else if (m == "literal") p.mode = RSF_MODE_LITERAL;
        else return fail("--mode must be fast or literal");
    }
    int device = 0;
    if (seen[idx("device")]) device = std::atoi(val[idx("device")].c_str());
    const bool use_phi0 = seen[idx("phi0")];
    if (seen[idx("batch")]) {
        if (seen[idx("image")] || seen[idx("binary")] || seen[idx("block")])
            return fail("--batch takes its images and seeds from the manifest");
        return run_batch(val[idx("batch")], use_phi0, p, device);
    }
    std::string out_path = "out.npy";
    if (seen[idx("ouput")]) out_path = val[idx("ouput")];
    if (seen[idx("output")]) out_path = val[idx("output")];
    if (!seen[idx("image")]) return fail("--image is required");
    if (!seen[idx("binary")] && !seen[idx("phi0")]) return fail("--binary (or --phi0) is required");

    if (!seen[idx("block")]) {
        // plain run, file to file on the device: image and seed streamed in, phi0 on the
        // device, phi streamed out; the headers are checked before a device is touched
        const std::string ip = val[idx("image")], sp = val[idx(use_phi0 ? "phi0" : "binary")];
        int64_t is[8], ss[8];
        int ind = 0, snd = 0, u8 = 0;
        if (rsf_npy_info(ip.c_str(), is, &ind, &u8) != RSF_OK) return fail(rsf_last_error());
        if (ind != 3) return fail("--image must be a 3-D array (Z, Y, X)");
        if (rsf_npy_info(sp.c_str(), ss, &snd, &u8) != RSF_OK) return fail(rsf_last_error());
        if (snd != 3 || ss[0] != is[0] || ss[1] != is[1] || ss[2] != is[2])
            return fail("initial level set and image shapes differ");
        rsf_context* ctx = nullptr;
        if (rsf_context_create(device, nullptr, &ctx) != RSF_OK) return fail(rsf_last_error());
        std::vector<float> ms((size_t)std::max(p.iterations, 1));
        const int st = rsf_run_npy(ctx, ip.c_str(), sp.c_str(), use_phi0 ? 1 : 0, out_path.c_str(), &p, ms.data());
        rsf_context_destroy(ctx);
        if (st != RSF_OK) return fail(std::string(rsf_status_string(st)) + ": " + rsf_last_error());
        for (int k = 0; k < p.iterations; ++k) std::cout << "It takes " << ms[k] << " ms to calculate " << std::endl;
        return 0;
    }
    // --block: the whole volume cut into overlapping blocks, run together, blended
    int64_t bs[3];
    if (!parse_dims(val[idx("block")], bs)) return fail("--block must be N or W,H,L (positive integers)");
    float ov = 32.f;
    if (seen[idx("overlap")] && (!parse_float(val[idx("overlap")], &ov) || ov < 0 || ov != (float)(int)ov))
        return fail("--overlap must be a non-negative integer");
    Image im;
    std::string err = load_image(val[idx("image")], &im);
    if (!err.empty()) return fail(err);
    const std::vector<int64_t>& shape = im.arr.shape;
    const int64_t l = shape[0], h = shape[1], w = shape[2];
    std::vector<float> phi;
    err = load_seed(val[idx(use_phi0 ? "phi0" : "binary")], shape, &phi);
    if (!err.empty()) return fail(err);
    ContextGuard g;
    if (rsf_context_create(device, nullptr, &g.c) != RSF_OK) return fail(rsf_last_error());
    err = make_phi0(g.c, use_phi0, shape, &phi);
    if (!err.empty()) return fail(err);
    const void* image_ptr = im.u8 ? (const void*)im.arr.data.data() : (const void*)im.f32.data();
    const int dtype = im.u8 ? RSF_DTYPE_U8 : RSF_DTYPE_F32;
    const int st = rsf_run_volume_blocks_host(g.c, image_ptr, dtype, phi.data(), w, h, l, bs[0], bs[1], bs[2],
                                              (int64_t)ov, &p);
    if (st == RSF_OK) print_iterations(g.c, p.iterations);
    if (st != RSF_OK) return fail(std::string(rsf_status_string(st)) + ": " + rsf_last_error());
    err = rsf::npy_save_f32(out_path, phi.data(), shape);
    return err.empty() ? 0 : fail(err);
}

}  // namespace

// The whole command line as a C entry point: the executable's main, and a way to
// run the same code in a process that already uses the GPU (no child process).
extern "C" int rsf_cli_main(int argc, char** argv) {
    const int rc = cli(argc, argv);
    std::cout.flush();
    std::cerr.flush();
    return rc;
}

#ifndef RSF_CLI_NO_MAIN
int main(int argc, char** argv) { return rsf_cli_main(argc, argv); }
#endif
