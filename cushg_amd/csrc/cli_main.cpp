// cli_main.cpp -- the `shg` program: drop-in for the reference executable.
//
//   shg Power[W] waist[um] T[C] save_only_last save_temperature [--key=value ...]
//
// The five positional arguments are those of src/cuSHG.cu:46-58 (a
// comma-separated list in any of the first three runs a batched sweep, see
// shg_cli_main).  The reference fixes grid, crystal and THERMAL at compile
// time (src/cuSHG.cu:11-15, src/headers/Crystal.h:26-52,
// src/headers/Solver.h:32); here they are options:
//   --nx= --ny= --nz= --nz-heat= --lx= --ly= --lz= (um)
//   --crystal=spplt|ppln  --thermal=0|1  --precision=fp32|fp32-storage|fp32-split|fp64|fp64-fused
//   --heat=jacobi|direct  --heat-bc=open|oven  --scheme=reference|strang  --max-outer=
//   --deff= (um/V) --period= (um) --alpha-p= --alpha-s= (1/um) --beta-s= (um/W)
//   --rho-p= --rho-s= (rad) --waist-y=UM (elliptical pump; default circular)
//   --out=DIR (default: current directory)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <string>
#include <vector>

#include "../../include/shg.h"

// value of a named choice: false (with a message) for an unknown name
static bool pick(const char* v, std::initializer_list<const char*> names, std::initializer_list<int> values,
                 const char* option, int32_t& out) {
    auto n = names.begin();
    for (int x : values) {
        if (std::strcmp(v, *n++) == 0) {
            out = x;
            return true;
        }
    }
    std::fprintf(stderr, "shg: %s=%s is not one of", option, v);
    for (const char* nm : names) std::fprintf(stderr, " %s", nm);
    std::fprintf(stderr, "\n");
    return false;
}

static bool opt(const char* a, const char* key, const char** val) {
    const size_t n = std::strlen(key);
    if (std::strncmp(a, key, n) == 0 && a[n] == '=') {
        *val = a + n + 1;
        return true;
    }
    return false;
}

int main(int argc, char** argv) {
    shg_params p;
    shg_params_default(&p);
    std::vector<const char*> pos{argv[0]};
    std::string out = ".";
    for (int i = 1; i < argc; i++) {
        const char* a = argv[i];
        const char* v = nullptr;
        if (std::strncmp(a, "--", 2) != 0) {
            pos.push_back(a);
        } else if (opt(a, "--nx", &v)) {
            p.nx = std::atoi(v);
        } else if (opt(a, "--ny", &v)) {
            p.ny = std::atoi(v);
        } else if (opt(a, "--nz", &v)) {
            p.nz = std::atoi(v);
        } else if (opt(a, "--nz-heat", &v)) {
            p.nz_heat = std::atoi(v);
        } else if (opt(a, "--lx", &v)) {
            p.lx_um = std::atof(v);
        } else if (opt(a, "--ly", &v)) {
            p.ly_um = std::atof(v);
        } else if (opt(a, "--lz", &v)) {
            p.lz_um = std::atof(v);
        } else if (opt(a, "--crystal", &v)) {
            const int c = std::strcmp(v, "ppln") == 0 ? SHG_CRYSTAL_MGO_PPLN
                          : std::strcmp(v, "spplt") == 0 ? SHG_CRYSTAL_MGO_SPPLT : -1;
            if (shg_params_set_crystal(&p, c) != SHG_OK) {
                std::fprintf(stderr, "shg: %s\n", shg_last_error());
                return 2;
            }
        } else if (opt(a, "--thermal", &v)) {
            p.thermal = std::atoi(v);
        } else if (opt(a, "--precision", &v)) {
            if (!pick(v, {"fp32", "fp32-storage", "fp64", "fp64-fused", "fp32-split"},
                      {SHG_PREC_FP32_MIXED, SHG_PREC_FP32_STORAGE, SHG_PREC_FP64_PLAIN, SHG_PREC_FP64_FUSED,
                       SHG_PREC_FP32_SPLIT},
                      "--precision", p.precision))
                return 2;
        } else if (opt(a, "--heat", &v)) {
            if (!pick(v, {"jacobi", "direct"}, {SHG_HEAT_JACOBI_REFERENCE, SHG_HEAT_DIRECT}, "--heat", p.heat_solver))
                return 2;
        } else if (opt(a, "--heat-bc", &v)) {
            if (!pick(v, {"open", "oven"}, {SHG_HEAT_BC_OPEN_TOP, SHG_HEAT_BC_OVEN_SURROUNDED}, "--heat-bc", p.heat_bc))
                return 2;
        } else if (opt(a, "--scheme", &v)) {
            if (!pick(v, {"reference", "strang"}, {SHG_SCHEME_REFERENCE, SHG_SCHEME_STRANG}, "--scheme", p.scheme))
                return 2;
        } else if (opt(a, "--max-outer", &v)) {
            p.max_outer = std::atoi(v);
        } else if (opt(a, "--deff", &v)) {
            p.deff_um_per_V = std::atof(v);
        } else if (opt(a, "--period", &v)) {
            p.Lambda_um = std::atof(v);
        } else if (opt(a, "--alpha-p", &v)) {
            p.alpha_p = std::atof(v);
        } else if (opt(a, "--alpha-s", &v)) {
            p.alpha_s = std::atof(v);
        } else if (opt(a, "--beta-s", &v)) {
            p.beta_s = std::atof(v);
        } else if (opt(a, "--rho-p", &v)) {
            p.rho_p = std::atof(v);
        } else if (opt(a, "--rho-s", &v)) {
            p.rho_s = std::atof(v);
        } else if (opt(a, "--waist-y", &v)) {
            p.waist_y_um = std::atof(v);
        } else if (opt(a, "--out", &v)) {
            out = v;
        } else {
            std::fprintf(stderr, "shg: unknown option %s\n", a);
            return 2;
        }
    }
    const int rc = shg_cli_main((int)pos.size(), pos.data(), &p, out.c_str());
    if (rc != SHG_OK) {
        std::fprintf(stderr, "shg: error %d: %s\n", rc, shg_last_error());
        return 1;
    }
    return 0;
}
