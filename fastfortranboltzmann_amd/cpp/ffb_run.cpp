// ffb_run -- C++ driver of the C ABI (include/ffb.h): one lattice, one run,
// one JSON line.  It stands where the reference's main program stands
// (source/layouts/xyd/main.f90:79-262: allocate, initial condition, the step
// loop, the MLUPS report of main.f90:324-326) for callers that link the
// library from C or C++ instead of Fortran or Python.  A failing call prints
// ffb_last_error() and exits with status 1, as the reference's `error stop`.
//
//   ffb_run --nx 4096 --ny 4096 --mode sl --precision fp32 --re 1000 --u 0.1
//           --steps 2000 --warmup 100 [--macros out.bin]
//
// --macros writes rho, u_x, u_y after the last step (n_y x n_x each, x fastest,
// FP64 for --precision fp64 and FP32 otherwise) one after the other.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ffb.h"

namespace {

void usage(FILE *out) {
    std::fprintf(out,
                 "usage: ffb_run [--nx N] [--ny N] [--mode sw|cf|pf|sl|bf] [--precision "
                 "fp32|fp64|fp16dev|bf16dev|fp32dev|e3m12dev]\n"
                 "               [--omega W | --re R] [--u U] [--force F] [--rho-in R] [--rho-out R]\n"
                 "               [--scale S] [--rounding stochastic|nearest] [--steps K] [--warmup W]\n"
                 "               [--device D] [--macros FILE]\n"
                 "Runs one lattice through the C ABI and prints one JSON line (kernel time, MLUPS,\n"
                 "FP64 sums of mass and momentum, largest speed).  --u is the lid, wall or wave\n"
                 "speed of the mode; --re sets omega from nu = u * n_y / re (settings.f90:170-172);\n"
                 "without --scale the deviation forms take ffb_default_scale_log2 at that speed.\n");
}

[[noreturn]] void die(const char *what) {
    std::fprintf(stderr, "error: %s: %s\n", what, ffb_last_error());
    std::exit(1);
}

void check(int rc, const char *what) {
    if (rc != FFB_OK) die(what);
}

[[noreturn]] void bad_arg(const std::string &msg) {
    std::fprintf(stderr, "ffb_run: %s\n", msg.c_str());
    usage(stderr);
    std::exit(2);
}

int mode_of(const std::string &s) {
    if (s == "sw" || s == "1") return FFB_SHEAR_WAVE;
    if (s == "cf" || s == "2") return FFB_COUETTE_FLOW;
    if (s == "pf" || s == "3") return FFB_POISEUILLE_FLOW;
    if (s == "sl" || s == "4") return FFB_SLIDING_LID;
    if (s == "bf" || s == "5") return FFB_BODY_FORCE_CHANNEL;
    bad_arg("unknown mode " + s);
}

int precision_of(const std::string &s) {
    if (s == "fp32") return FFB_PREC_FP32;
    if (s == "fp64") return FFB_PREC_FP64;
    if (s == "fp16dev") return FFB_PREC_FP16_DEV;
    if (s == "bf16dev") return FFB_PREC_BF16_DEV;
    if (s == "fp32dev") return FFB_PREC_FP32_DEV;
    if (s == "e3m12dev") return FFB_PREC_E3M12_DEV;
    bad_arg("unknown precision " + s);
}

}  // namespace

int main(int argc, char **argv) {
    ffb_config_t cfg;
    ffb_config_default(&cfg);
    cfg.n_x = 256;
    cfg.n_y = 256;
    std::string precision = "fp32", macros_path;
    double re = 0.0, u = -1.0;
    bool omega_given = false, scale_given = false;
    long long steps = 1000, warmup = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            usage(stdout);
            return 0;
        }
        if (i + 1 >= argc) bad_arg("missing value for " + a);
        const std::string v = argv[++i];
        char *end = nullptr;
        auto num = [&](const char *what) {
            const double x = std::strtod(v.c_str(), &end);
            if (end == v.c_str() || *end) bad_arg(std::string("not a number for ") + what + ": " + v);
            return x;
        };
        if (a == "--nx") cfg.n_x = (int32_t)num("--nx");
        else if (a == "--ny") cfg.n_y = (int32_t)num("--ny");
        else if (a == "--mode") cfg.sim_mode = mode_of(v);
        else if (a == "--precision") precision = v;
        else if (a == "--omega") cfg.omega = num("--omega"), omega_given = true;
        else if (a == "--re") re = num("--re");
        else if (a == "--u") u = num("--u");
        else if (a == "--force") cfg.body_force = num("--force");
        else if (a == "--rho-in") cfg.rho_in = num("--rho-in");
        else if (a == "--rho-out") cfg.rho_out = num("--rho-out");
        else if (a == "--scale") cfg.storage_scale_log2 = (int32_t)num("--scale"), scale_given = true;
        else if (a == "--rounding") {
            if (v == "stochastic") cfg.storage_rounding = FFB_ROUND_STOCHASTIC;
            else if (v == "nearest") cfg.storage_rounding = FFB_ROUND_NEAREST;
            else bad_arg("unknown rounding " + v);
        } else if (a == "--steps") steps = (long long)num("--steps");
        else if (a == "--warmup") warmup = (long long)num("--warmup");
        else if (a == "--device") cfg.device = (int32_t)num("--device");
        else if (a == "--macros") macros_path = v;
        else bad_arg("unknown option " + a);
    }
    cfg.precision = precision_of(precision);
    if (u >= 0.0) {
        cfg.u_lid = u;
        cfg.u_wall = u;
        cfg.u_max = u;
    }
    const double speed = cfg.sim_mode == FFB_COUETTE_FLOW ? cfg.u_wall
                         : cfg.sim_mode == FFB_SHEAR_WAVE ? cfg.u_max
                                                          : cfg.u_lid;
    if (re > 0.0) {
        if (omega_given) bad_arg("give --omega or --re, not both");
        cfg.omega = 1.0 / (3.0 * speed * cfg.n_y / re + 0.5);
    }
    if (!scale_given) {  // the library's default for the form at the mode's speed
        cfg.storage_scale_log2 = ffb_default_scale_log2(cfg.precision, speed < 0 ? -speed : speed);
        if (cfg.storage_scale_log2 == INT32_MIN) bad_arg("no default --scale for this precision and speed");
    }
    if (steps < 1 || warmup < 0 || steps > INT32_MAX || warmup > INT32_MAX) bad_arg("bad step counts");

    ffb_lattice_t *lat = nullptr;
    check(ffb_create(&cfg, &lat), "ffb_create");
    check(ffb_init(lat), "ffb_init");
    if (warmup > 0) check(ffb_step(lat, (int32_t)warmup, 0), "ffb_step (warm-up)");
    check(ffb_sync(lat), "ffb_sync");
    check(ffb_reset_timer(lat), "ffb_reset_timer");
    check(ffb_step(lat, (int32_t)steps, 1), "ffb_step");
    check(ffb_sync(lat), "ffb_sync");
    double ms = 0.0;
    check(ffb_elapsed_ms(lat, &ms), "ffb_elapsed_ms");
    double mass = 0.0, mom_x = 0.0, mom_y = 0.0;
    check(ffb_moments(lat, &mass, &mom_x, &mom_y), "ffb_moments");

    const size_t cells = (size_t)cfg.n_x * (size_t)cfg.n_y;
    const bool fp64 = cfg.precision == FFB_PREC_FP64;
    const size_t elem = fp64 ? 8 : 4;
    std::vector<unsigned char> fields(3 * cells * elem);
    check(ffb_read_macros(lat, fields.data(), fields.data() + cells * elem, fields.data() + 2 * cells * elem),
          "ffb_read_macros");
    double u_top = 0.0;
    for (size_t k = 0; k < cells; ++k) {
        const double ux = fp64 ? ((const double *)fields.data())[cells + k] : ((const float *)fields.data())[cells + k];
        const double uy =
            fp64 ? ((const double *)fields.data())[2 * cells + k] : ((const float *)fields.data())[2 * cells + k];
        const double s2 = ux * ux + uy * uy;
        if (s2 > u_top) u_top = s2;
    }
    if (!macros_path.empty()) {
        FILE *f = std::fopen(macros_path.c_str(), "wb");
        const bool ok = f && std::fwrite(fields.data(), 1, fields.size(), f) == fields.size();
        if (f && std::fclose(f) != 0) f = nullptr;
        if (!ok || !f) {
            std::fprintf(stderr, "error: cannot write %s\n", macros_path.c_str());
            return 1;
        }
    }
    check(ffb_destroy(lat), "ffb_destroy");

    const double mlups = (double)cells * (double)steps / (ms * 1e-3) / 1e6;
    std::printf("{\"driver\": \"ffb_run\", \"n_x\": %d, \"n_y\": %d, \"sim_mode\": %d, \"precision\": \"%s\", "
                "\"omega\": %.17g, \"storage_scale_log2\": %d, \"steps\": %lld, \"warmup\": %lld, \"kernel_ms\": %.6f, \"ms_per_step\": %.6f, "
                "\"mlups\": %.1f, \"mass\": %.17g, \"mom_x\": %.17g, \"mom_y\": %.17g, \"u_max\": %.17g}\n",
                cfg.n_x, cfg.n_y, cfg.sim_mode, precision.c_str(), cfg.omega, cfg.storage_scale_log2, steps, warmup, ms, ms / (double)steps,
                mlups, mass, mom_x, mom_y, std::sqrt(u_top));
    return 0;
}
