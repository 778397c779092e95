/*
 * sanity_main.c -- drives the FP64 oracle through every code path on small
 * grids for the AddressSanitizer / UndefinedBehaviorSanitizer build
 * (`make -C oracle sanitize`).  TEST INFRASTRUCTURE ONLY.
 *
 * Paths: plain march (reference and Strang schemes), walk-off, elliptical
 * pump, poling phases, THERMAL with the reference Jacobi loop and with the
 * direct solver on both boundary sets, a capped Jacobi loop (ENOCONV),
 * the test-only heat-source march, the FFT at radix-2/3/odd lengths and
 * the tridiagonal eigen-solver.  Exit status 0 when every call returned
 * what it should and every result is finite.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_shg.h"

static int fails = 0;
#define EXPECT(c, what)                                   \
    do {                                                  \
        if (!(c)) {                                       \
            fprintf(stderr, "sanity: FAILED %s\n", what); \
            fails++;                                      \
        }                                                 \
    } while (0)

/* reference values of src/headers/Crystal.h:19-52 and src/cuSHG.cu:11-58 */
static void base(shg_params* p, int n, int nz) {
    memset(p, 0, sizeof(*p));
    p->nx = p->ny = n;
    p->nz = nz;
    p->lx_um = p->ly_um = 800.0;
    p->lz_um = 30000.0;
    p->crystal = SHG_CRYSTAL_MGO_SPPLT;
    p->lambda_p_um = 1.064;
    p->deff_um_per_V = 11e-6;
    p->Lambda_um = 7.97;
    p->alpha_p = 0.17e-6;
    p->alpha_s = 1.57e-6;
    p->beta_s = 5e-5;
    p->k_th = 8.4e-6;
    p->h_conv = 10e-12;
    p->therm_exp_a = 2.2e-6;
    p->therm_exp_b = -5.9e-9;
    p->power_W = 5.0;
    p->waist_um = 40.0;
    p->focal_um = 0.5 * p->lz_um;
    p->T_C = p->T_peltier_C = 55.66;
    p->T_inf_C = 25.0;
    p->max_outer = 20;
    p->check_every = 50;
    p->max_sweeps = 2000000;
    p->jacobi_tol = 5e-4;
    p->heat_tol_K = 1e-8;
}

static void run_ok(const shg_params* p, const char* what, int want) {
    const size_t n2 = (size_t)p->nx * p->ny;
    const int nzh = p->nz_heat > 0 ? p->nz_heat : p->nz;
    double* pump = malloc(sizeof(double) * 2 * n2);
    double* sig = malloc(sizeof(double) * 2 * n2);
    double* T = p->thermal ? malloc(sizeof(double) * n2 * nzh) : NULL;
    shg_result r;
    memset(&r, 0, sizeof r);
    r.pump_out = pump;
    r.signal_out = sig;
    r.temperature_out = T;
    const int rc = orc_run(p, &r);
    EXPECT(rc == want, what);
    if (rc == SHG_OK) {
        EXPECT(isfinite(r.eta) && r.P_pump_in_W > 0, what);
        double s = 0;
        for (size_t i = 0; i < 2 * n2; i++) s += pump[i] + sig[i];
        EXPECT(isfinite(s), what);
    }
    printf("sanity: %-40s rc=%d eta=%.6e outer=%d sweeps=%lld\n", what, rc, r.eta, r.outer_iters,
           (long long)r.jacobi_sweeps);
    free(pump);
    free(sig);
    free(T);
}

int main(void) {
    orc_set_threads(2);
    shg_params p;

    base(&p, 16, 8);
    run_ok(&p, "plain, reference scheme", SHG_OK);
    p.scheme = SHG_SCHEME_STRANG;
    run_ok(&p, "plain, Strang scheme", SHG_OK);

    base(&p, 24, 10);
    p.nx = 12;
    p.rho_p = 0.01;
    p.rho_s = 0.02;
    p.waist_y_um = 55.0;
    double ph[10];
    for (int i = 0; i < 10; i++) ph[i] = 0.1 * i - 0.3;
    p.poling_phase = ph;
    run_ok(&p, "walk-off, elliptical pump, poling", SHG_OK);

    base(&p, 8, 6);
    p.thermal = 1;
    p.nz_heat = 3;
    p.heat_solver = SHG_HEAT_JACOBI_REFERENCE;
    p.T_peltier_C = 40.0;
    run_ok(&p, "THERMAL, reference Jacobi loop", SHG_OK);
    p.max_sweeps = 3;
    run_ok(&p, "THERMAL, Jacobi capped (ENOCONV)", SHG_ENOCONV);

    base(&p, 12, 9);
    p.thermal = 1;
    p.nz_heat = 4;
    p.heat_solver = SHG_HEAT_DIRECT;
    p.alpha_p = 2e-6;
    p.alpha_s = 2e-5;
    run_ok(&p, "THERMAL, direct, open top", SHG_OK);
    p.heat_bc = SHG_HEAT_BC_OVEN_SURROUNDED;
    run_ok(&p, "THERMAL, direct, oven surrounded", SHG_OK);
    p.crystal = SHG_CRYSTAL_MGO_PPLN;
    p.Lambda_um = 6.97;
    run_ok(&p, "THERMAL, direct, PPLN", SHG_OK);
    p.nz_heat = 20;
    run_ok(&p, "THERMAL, nz_heat > nz (EINVAL)", SHG_EINVAL);

    /* heat source march on a given T */
    base(&p, 8, 5);
    p.thermal = 1;
    p.nz_heat = 2;
    {
        double T[8 * 8 * 2], Q[8 * 8 * 2], pw[5 * 2];
        for (int i = 0; i < 128; i++) T[i] = 40.0 + 0.01 * i;
        EXPECT(orc_march_heat(&p, T, Q, pw) == SHG_OK, "march_heat");
    }

    /* FFT lengths: powers of two, radix 3, odd and prime factors up to 64;
     * a prime factor above 64 is refused */
    const int lens[] = {1, 2, 3, 7, 12, 30, 61, 64, 192, 97};
    for (size_t k = 0; k < sizeof lens / sizeof lens[0]; k++) {
        const int n = lens[k];
        const int want = n == 97 ? -1 : 0;
        double* x = malloc(sizeof(double) * 2 * n * 3);
        double* y = malloc(sizeof(double) * 2 * n);
        for (int i = 0; i < 2 * n * 3; i++) x[i] = sin(0.37 * i + n);
        EXPECT(orc_fft1d(x, n, 3, 0) == want, "fft1d");
        EXPECT(orc_fft1d(x, n, 3, 1) == want, "ifft1d");
        orc_dft1d_naive(x, y, n, 0);
        free(x);
        free(y);
    }
    {
        double x[2 * 12 * 6];
        for (int i = 0; i < 2 * 72; i++) x[i] = cos(0.11 * i);
        EXPECT(orc_fft2d(x, 12, 6, 0) == 0, "fft2d");
    }
    /* tridiagonal eigen-solver, n = 1 and a longer one */
    {
        double d[9], o[9], lam[9], V[81];
        for (int i = 0; i < 9; i++) {
            d[i] = 2.0 + 0.1 * i;
            o[i] = -1.0;
        }
        EXPECT(orc_tridiag_eigen(1, d, o, lam, V) == 0, "eigen n=1");
        EXPECT(orc_tridiag_eigen(9, d, o, lam, V) == 0, "eigen n=9");
    }
    printf("sanity: %s (%d failures)\n", fails ? "FAILED" : "ok", fails);
    return fails ? 1 : 0;
}
