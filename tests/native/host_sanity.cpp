// host_sanity.cpp -- drives the host code of the HIP library through the C
// ABI for the AddressSanitizer / UndefinedBehaviorSanitizer build of its
// host side (python cushg_amd/build.py --sanitize).  Test infrastructure.
//
// Without a GPU: parameter defaults and the crystal database, validation of
// every rejected parameter, Sellmeier and dk, handle creation failing with
// SHG_EDEVICE, the CLI's argument checks.  With a GPU (argv[1] == "gpu"):
// also small runs of every form, THERMAL with both solvers and boundary
// sets, a batch of uneven instances with failing members, the per-slice
// trace, the adjoint, plans, graph replay and the CLI writing its .dat
// files into argv[2].  Exit status 0 when every call returned what it
// should.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "shg.h"

static int fails = 0;
#define EXPECT(c, what)                                            \
    do {                                                           \
        if (!(c)) {                                                \
            std::fprintf(stderr, "host_sanity: FAILED %s (%s)\n", what, \
                         shg_last_error());                        \
            fails++;                                               \
        }                                                          \
    } while (0)

static shg_params small(int n, int nz) {
    shg_params p;
    shg_params_default(&p);
    p.nx = p.ny = n;
    p.nz = nz;
    p.lx_um = p.ly_um = 800.0;
    p.power_W = 5.0;
    p.waist_um = 40.0;
    return p;
}

static void cpu_checks() {
    shg_params p;
    shg_params_default(&p);
    EXPECT(shg_validate(&p) == SHG_OK, "defaults validate");
    EXPECT(shg_params_set_crystal(&p, SHG_CRYSTAL_MGO_PPLN) == SHG_OK, "set PPLN");
    EXPECT(shg_params_set_crystal(&p, 9) == SHG_EINVAL, "unknown crystal");
    EXPECT(shg_params_set_crystal(nullptr, 0) == SHG_EINVAL, "null params");
    EXPECT(std::isfinite(shg_refractive_index(SHG_CRYSTAL_MGO_PPLN, 1.064, 40.0)), "n PPLN");
    EXPECT(std::isfinite(shg_delta_k(&p, 40.0)), "dk");
    EXPECT(shg_validate(nullptr) == SHG_EINVAL, "validate null");
    struct Bad {
        void (*f)(shg_params&);
    };
    const Bad bad[] = {
        {[](shg_params& q) { q.nx = 100; }},      {[](shg_params& q) { q.ny = 1024; }},
        {[](shg_params& q) { q.nz = 1; }},        {[](shg_params& q) { q.thermal = 1, q.nz_heat = 1; }},
        {[](shg_params& q) { q.thermal = 1, q.nz = 10, q.nz_heat = 20; }},
        {[](shg_params& q) { q.lz_um = 0; }},     {[](shg_params& q) { q.waist_um = -1; }},
        {[](shg_params& q) { q.power_W = NAN; }}, {[](shg_params& q) { q.crystal = 5; }},
        {[](shg_params& q) { q.scheme = 9; }},    {[](shg_params& q) { q.precision = 7; }},
        {[](shg_params& q) { q.thermal = 1, q.heat_solver = 4; }},
        {[](shg_params& q) { q.thermal = 1, q.heat_bc = 5; }},
        {[](shg_params& q) { q.Lambda_um = 0; }}, {[](shg_params& q) { q.waist_y_um = -2; }},
        {[](shg_params& q) { q.lambda_p_um = 9; }}, {[](shg_params& q) { q.T_C = INFINITY; }},
    };
    for (const Bad& b : bad) {
        shg_params q;
        shg_params_default(&q);
        b.f(q);
        EXPECT(shg_validate(&q) == SHG_EINVAL, "validate rejects");
        EXPECT(std::strlen(shg_last_error()) > 0, "message set");
    }
    for (int n = 0; n < 600; n++) (void)shg_fft_size_supported(n);
    const char* argv_short[] = {"shg", "30"};
    EXPECT(shg_cli_main(2, argv_short, nullptr, ".") == SHG_EINVAL, "cli: too few arguments");
}

static void gpu_checks(const char* out_dir) {
    shg_handle* h = nullptr;
    if (shg_handle_create(0, &h) != SHG_OK) {
        EXPECT(false, "handle create");
        return;
    }
    const int forms[] = {SHG_PREC_FP32_SPLIT, SHG_PREC_FP32_STORAGE, SHG_PREC_FP32_MIXED, SHG_PREC_FP64_FUSED,
                         SHG_PREC_FP64_PLAIN};
    for (int f : forms) {
        shg_params p = small(64, 12);
        p.precision = f;
        std::vector<double> P(2 * 64 * 64), S(2 * 64 * 64);
        shg_result r{};
        r.pump_out = P.data();
        r.signal_out = S.data();
        EXPECT(shg_run(h, &p, &r) == SHG_OK && std::isfinite(r.eta), "run");
    }
    for (int solver : {SHG_HEAT_DIRECT, SHG_HEAT_JACOBI_REFERENCE})
        for (int bc : {SHG_HEAT_BC_OPEN_TOP, SHG_HEAT_BC_OVEN_SURROUNDED}) {
            shg_params p = small(64, 16);
            p.thermal = 1;
            p.nz_heat = 4;
            p.heat_solver = solver;
            p.heat_bc = bc;
            p.alpha_p = 2e-6;
            p.alpha_s = 2e-5;
            p.check_every = 100;
            p.precision = SHG_PREC_FP32_SPLIT;
            std::vector<double> T(64 * 64 * 4);
            shg_result r{};
            r.temperature_out = T.data();
            EXPECT(shg_run(h, &p, &r) == SHG_OK && std::isfinite(r.peak_dT_K), "THERMAL run");
            p.precision = SHG_PREC_FP64_PLAIN;
            EXPECT(shg_run(h, &p, &r) == SHG_OK, "THERMAL plain run");
        }
    {
        std::vector<shg_params> ps;
        const int grids[] = {64, 128, 192, 256, 64, 384};
        for (int i = 0; i < 6; i++) {
            shg_params p = small(grids[i], 8 + 3 * i);
            p.precision = i % 2 ? SHG_PREC_FP32_STORAGE : SHG_PREC_FP32_SPLIT;
            ps.push_back(p);
        }
        ps[4].nx = 100;  // invalid member
        std::vector<shg_result> rs(ps.size());
        EXPECT(shg_run_batch(h, ps.data(), (int)ps.size(), rs.data()) == SHG_EINVAL, "batch first failure");
        EXPECT(rs[4].status == SHG_EINVAL && rs[0].status == SHG_OK && rs[5].status == SHG_OK, "batch statuses");
        EXPECT(shg_set_graph_steps(h, 4) == SHG_OK, "graph steps");
        EXPECT(shg_run_batch(h, ps.data(), (int)ps.size(), rs.data()) == SHG_EINVAL, "batch replayed");
        EXPECT(shg_set_graph_steps(h, -1) == SHG_OK, "graph default");
    }
    {
        shg_params p = small(64, 10);
        p.precision = SHG_PREC_FP32_SPLIT;
        shg_result r{};
        int calls = 0;
        auto fn = [](int32_t, int32_t, const double* a, const double*, void* u) -> int {
            ++*static_cast<int*>(u);
            return std::isfinite(a[0]) ? 0 : 1;
        };
        EXPECT(shg_run_traced(h, &p, 3, fn, &calls, &r) == SHG_OK && calls == 4, "trace");
    }
    {
        shg_params p = small(64, 24);
        p.precision = SHG_PREC_FP32_STORAGE;
        std::vector<double> pol(24, 0.01), g(3 + 24);
        p.poling_phase = pol.data();
        shg_result r{};
        EXPECT(shg_adjoint(h, &p, 5, g.data(), &r) == SHG_OK && std::isfinite(g[0]), "adjoint");
    }
    {
        int32_t rows = 0, cols = 0;
        EXPECT(shg_set_plan(h, 128, SHG_PREC_FP32_SPLIT, 2, 4) == SHG_OK, "set plan");
        EXPECT(shg_get_plan(h, 128, SHG_PREC_FP32_SPLIT, &rows, &cols) == SHG_OK && rows == 2 && cols == 4,
               "get plan");
        EXPECT(shg_autotune(h, 64, SHG_PREC_FP32_SPLIT, 2, 8) == SHG_OK, "autotune");
        const char* names[64];
        double ms[64];
        int32_t nl[64];
        EXPECT(shg_handle_set_profiling(h, 1) == SHG_OK, "profiling on");
        shg_params p = small(64, 6);
        shg_result r{};
        EXPECT(shg_run(h, &p, &r) == SHG_OK, "profiled run");
        EXPECT(shg_last_kernel_times(h, names, ms, nl, 64) > 0, "kernel times");
        EXPECT(shg_handle_set_profiling(h, 0) == SHG_OK, "profiling off");
    }
    {
        shg_params base = small(64, 8);
        base.precision = SHG_PREC_FP32_SPLIT;
        const char* argv[] = {"shg", "2.5", "35", "55.66", "1", "0"};
        EXPECT(shg_cli_main(6, argv, &base, out_dir) == SHG_OK, "cli");
    }
    EXPECT(shg_handle_destroy(h) == SHG_OK, "destroy");
}

int main(int argc, char** argv) {
    cpu_checks();
    if (argc > 1 && std::strcmp(argv[1], "gpu") == 0) {
        gpu_checks(argc > 2 ? argv[2] : ".");
    } else {
        shg_handle* h = nullptr;
        const int rc = shg_handle_create(0, &h);
        EXPECT(rc == SHG_EDEVICE && h == nullptr, "no GPU: handle create fails");
    }
    std::printf("host_sanity: %s (%d failures)\n", fails ? "FAILED" : "ok", fails);
    return fails ? 1 : 0;
}
