// rsf -- drop-in for the reference's rsf3d executable (CUDA_RSF/main.cpp:73-319).
//
// Same long options and defaults (main.cpp:89-102, Boost style: "--name value"
// or "--name=value", unambiguous prefixes accepted, no short options), same
// inputs (--binary seed .npy, --image .npy) and output (fp32 phi .npy of the
// image's shape, --ouput), same per-iteration stdout line (main.cpp:309) and
// the same exit code 1 for --help.  Additions: --output (alias of --ouput),
// --lambda1/--lambda2 (default --e), --epsilon (default 0.2, the reference's
// hard-coded value), --phi0 (take fp32 phi directly instead of sdf(seed-13)),
// --mode fast|literal, --device.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "npy.hpp"
#include "rsf.h"

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
    {"mode", true, "fast | literal (four blurred channels in reference order)", "fast"},
    {"device", true, "HIP device index", "0"},
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

}  // namespace

int main(int argc, char** argv) {
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
    std::string out_path = "out.npy";
    if (seen[idx("ouput")]) out_path = val[idx("ouput")];
    if (seen[idx("output")]) out_path = val[idx("output")];
    if (!seen[idx("image")]) return fail("--image is required");
    if (!seen[idx("binary")] && !seen[idx("phi0")]) return fail("--binary (or --phi0) is required");

    rsf::NpyArray img;
    std::string err = rsf::npy_load(val[idx("image")], &img);
    if (!err.empty()) return fail(err);
    if (img.shape.size() != 3) return fail("--image must be a 3-D array (Z, Y, X)");
    const int64_t l = img.shape[0], h = img.shape[1], w = img.shape[2];
    const int64_t n = w * h * l;

    std::vector<float> phi;
    const bool use_phi0 = seen[idx("phi0")];
    {
        rsf::NpyArray s;
        err = rsf::npy_load(val[idx(use_phi0 ? "phi0" : "binary")], &s);
        if (!err.empty()) return fail(err);
        if (s.shape != img.shape) return fail("initial level set and image shapes differ");
        std::vector<float> seed;
        err = rsf::npy_to_f32(s, &seed);
        if (!err.empty()) return fail(err);
        if (use_phi0) {
            phi.swap(seed);
        } else {
            phi.resize((size_t)n);
            if (rsf_sdf_from_seed(seed.data(), phi.data(), w, h, l, 13.0f) != RSF_OK)  // main.cpp:174-183
                return fail(rsf_last_error());
        }
    }
    const void* image_ptr;
    int dtype;
    std::vector<float> imgf;
    if (img.descr == "|u1" || img.descr == "<u1") {
        image_ptr = img.data.data();
        dtype = RSF_DTYPE_U8;
    } else {
        err = rsf::npy_to_f32(img, &imgf);
        if (!err.empty()) return fail(err);
        image_ptr = imgf.data();
        dtype = RSF_DTYPE_F32;
    }
    rsf_context* ctx = nullptr;
    if (rsf_context_create(device, nullptr, &ctx) != RSF_OK) return fail(rsf_last_error());
    std::vector<float> ms((size_t)std::max(p.iterations, 1));
    int st = rsf_run_host(ctx, image_ptr, dtype, phi.data(), w, h, l, &p, ms.data());
    rsf_context_destroy(ctx);
    if (st != RSF_OK) return fail(std::string(rsf_status_string(st)) + ": " + rsf_last_error());
    for (int k = 0; k < p.iterations; ++k) std::cout << "It takes " << ms[k] << " ms to calculate " << std::endl;
    err = rsf::npy_save_f32(out_path, phi.data(), img.shape);
    if (!err.empty()) return fail(err);
    return 0;
}
