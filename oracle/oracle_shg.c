/*
 * oracle_shg.c -- FP64 CPU restatement of cuSHG's split-step Fourier
 * propagation loop (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed
 * CPU baseline.  It is loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg, never by the product library.
 *
 * It restates, in double precision, the files the reference's hot path
 * lives in (paths relative to the cuSHG repository root):
 *   src/headers/Crystal.h:19-70      constants, grid steps, Sellmeier n(L,T)
 *   src/headers/PhaseMatching.h:12-153  dk(T), the running dk integral
 *   src/headers/Efields.h:15-241     pump, propagators, fftshift, FFT scale
 *   src/headers/Solver.h:17-390      dAdz, RK4, diffraction, SSFM, CWES, run
 *   src/headers/Tfield.h:13-61,93-201,243-402  Jacobi sweep, Q, ovens
 *   src/headers/Operators.h:13-192   complex arithmetic (C99 complex here)
 *   src/headers/Common.h:88-90       IDX layout (x fastest)
 * cuFFT is replaced by a mixed-radix Stockham FFT (forward sign -1,
 * unnormalised, as cufftExecC2C); Thrust's reduce by plain sums.
 *
 * Deliberate differences from the reference, all toward its evident intent:
 *  - everything is FP64 (the reference is FP32, src/cuSHG.cu:7-8);
 *  - z_s = s*dz exactly (the reference accumulates z in FP32,
 *    src/headers/Solver.h:303);
 *  - the SH field starts at 0 (the reference adds clock-seeded N(0,1e-20)
 *    noise, src/headers/Efields.h:144-168);
 *  - the running dk integral is an O(1) update per slice (the reference
 *    re-sums slices 0..s, src/headers/PhaseMatching.h:27-44: same value);
 *  - the heat grid may have nz_heat != nz slices: optical slice s uses heat
 *    slice floor(s*nz_heat/nz), and Q of a heat slice is the mean over the
 *    optical slices mapped to it (identical to the reference at
 *    nz_heat == nz);
 *  - aux_sum of the Jacobi stagnation test starts as NaN (the reference
 *    reads it uninitialised, src/headers/Solver.h:335,349).
 */
#include "oracle_shg.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* src/headers/Crystal.h:19-21: C [um/ps], EPS0 [W ps / (V^2 um)] */
static const double ORC_PI = 3.14159265358979323846;
static const double ORC_C = 299792458.0 * 1e6 / 1e12;
static const double ORC_EPS0 = 8.8541878128e-12 * 1e12 / 1e6;

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* material                                                            */
/* ------------------------------------------------------------------ */

/* MgO:sPPLT extraordinary index, Bruner et al. -- src/headers/Crystal.h:54-70 */
static double n_spplt(double L, double T) {
    const double A = 4.502483, B = 0.007294, C = 0.185087, D = -0.02357;
    const double E = 0.073423, F = 0.199595, G = 0.001, H = 7.99724;
    const double tk = T + 273.15;
    const double b = 3.483933e-8 * tk * tk;
    const double c = 1.607839e-8 * tk * tk;
    const double L2 = L * L;
    return sqrt(A + (B + b) / (L2 - (C + c) * (C + c)) + E / (L2 - F * F) +
                G / (L2 - H * H) + D * L2);
}

/* 5% MgO:CLN extraordinary index, Gayer et al., Appl. Phys. B 91, 343
 * (2008).  No reference counterpart (README.md:52 names PPLN, Crystal.h
 * holds only sPPLT): literature values, parity unpinned. */
static double n_ppln(double L, double T) {
    const double a1 = 5.756, a2 = 0.0983, a3 = 0.2020, a4 = 189.32;
    const double a5 = 12.52, a6 = 1.32e-2;
    const double b1 = 2.860e-6, b2 = 4.700e-8, b3 = 6.113e-8, b4 = 1.516e-4;
    const double f = (T - 24.5) * (T + 570.82);
    const double L2 = L * L;
    const double g = a3 + b3 * f;
    return sqrt(a1 + b1 * f + (a2 + b2 * f) / (L2 - g * g) +
                (a4 + b4 * f) / (L2 - a5 * a5) - a6 * L2);
}

double orc_refractive_index(int32_t crystal, double lambda_um, double T_C) {
    return crystal == SHG_CRYSTAL_MGO_PPLN ? n_ppln(lambda_um, T_C)
                                           : n_spplt(lambda_um, T_C);
}

/* dk(T) = 2 pi (2 n_p/l_p - n_s/l_s + 1/(Lambda (1 + a dT + b dT^2))),
 * dT = T - 25 -- src/headers/PhaseMatching.h:59,80 */
double orc_delta_k(const shg_params* p, double T) {
    const double lp = p->lambda_p_um, ls = 0.5 * lp;
    const double np = orc_refractive_index(p->crystal, lp, T);
    const double ns = orc_refractive_index(p->crystal, ls, T);
    const double d = T - 25.0;
    return 2.0 * ORC_PI *
           (2.0 * np / lp - ns / ls +
            1.0 / (p->Lambda_um * (1.0 + p->therm_exp_a * d + p->therm_exp_b * d * d)));
}

/* ------------------------------------------------------------------ */
/* FFT: mixed-radix Stockham autosort, FP64                            */
/* ------------------------------------------------------------------ */

typedef struct {
    int n, nrad;
    int rad[40];
    cplx* tw; /* tw[m] = exp(-2 pi i m / n) */
} fft_plan;

static int fft_plan_init(fft_plan* pl, int n) {
    memset(pl, 0, sizeof(*pl));
    if (n < 1) return -1;
    pl->n = n;
    int m = n;
    while (m % 8 == 0 && m > 8) { pl->rad[pl->nrad++] = 8; m /= 8; }
    while (m % 4 == 0) { pl->rad[pl->nrad++] = 4; m /= 4; }
    while (m % 2 == 0) { pl->rad[pl->nrad++] = 2; m /= 2; }
    for (int f = 3; m > 1; f += 2) {
        while (m % f == 0) {
            if (f > 64) return -1;
            pl->rad[pl->nrad++] = f;
            m /= f;
        }
    }
    if (m == 8) pl->rad[pl->nrad++] = 8; /* n == 8 exactly */
    pl->tw = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    if (!pl->tw) return -1;
    for (int k = 0; k < n; k++) {
        /* reduce to |angle| <= pi before sin/cos */
        int kk = k <= n / 2 ? k : k - n;
        double a = -2.0 * ORC_PI * (double)kk / (double)n;
        pl->tw[k] = cos(a) + I * sin(a);
    }
    return 0;
}

static void fft_plan_free(fft_plan* pl) { free(pl->tw); pl->tw = NULL; }

/* one Stockham pass: radix R, span L (product of earlier radices) */
static void stockham_pass(const cplx* in, cplx* out, const fft_plan* pl, int R,
                          int L, int inverse) {
    const int n = pl->n, m = n / R, twmul = n / (L * R), rstep = n / R;
    cplx x[64], y[64];
    for (int t = 0; t < m; t++) {
        const int k = t % L;
        for (int j = 0; j < R; j++) {
            cplx w = pl->tw[(j * k * twmul) % n];
            if (inverse) w = conj(w);
            x[j] = in[t + j * m] * w;
        }
        if (R == 2) {
            y[0] = x[0] + x[1];
            y[1] = x[0] - x[1];
        } else if (R == 4) {
            cplx a = x[0] + x[2], b = x[0] - x[2];
            cplx c = x[1] + x[3], d = (x[1] - x[3]) * (inverse ? I : -I);
            y[0] = a + c; y[2] = a - c; y[1] = b + d; y[3] = b - d;
        } else {
            for (int q = 0; q < R; q++) {
                cplx acc = 0;
                for (int j = 0; j < R; j++) {
                    cplx w = pl->tw[((j * q) % R) * rstep];
                    if (inverse) w = conj(w);
                    acc += x[j] * w;
                }
                y[q] = acc;
            }
        }
        const int base = (t - k) * R + k;
        for (int q = 0; q < R; q++) out[base + q * L] = y[q];
    }
}

/* in-place 1D FFT of a contiguous vector using scratch of n */
static void fft_exec(const fft_plan* pl, cplx* data, cplx* scratch, int inverse) {
    cplx* a = data;
    cplx* b = scratch;
    int L = 1;
    for (int i = 0; i < pl->nrad; i++) {
        stockham_pass(a, b, pl, pl->rad[i], L, inverse);
        L *= pl->rad[i];
        cplx* t = a; a = b; b = t;
    }
    if (a != data) memcpy(data, a, sizeof(cplx) * (size_t)pl->n);
}

int orc_fft1d(double* data, int32_t n, int32_t howmany, int32_t inverse) {
    fft_plan pl;
    if (fft_plan_init(&pl, n)) return -1;
    cplx* d = (cplx*)data;
#pragma omp parallel
    {
        cplx* scr = (cplx*)malloc(sizeof(cplx) * (size_t)n);
#pragma omp for schedule(static)
        for (int b = 0; b < howmany; b++) fft_exec(&pl, d + (size_t)b * n, scr, inverse);
        free(scr);
    }
    fft_plan_free(&pl);
    return 0;
}

/* 2D transform of data[y][x]: rows (length nx) then columns (length ny) */
static void fft2d_planned(const fft_plan* px, const fft_plan* py, cplx* d, int nx,
                          int ny, int inverse) {
    const int CB = 8;
#pragma omp parallel
    {
        cplx* scr = (cplx*)malloc(sizeof(cplx) * (size_t)(nx > ny ? nx : ny) * (CB + 1));
        cplx* col = scr + (nx > ny ? nx : ny);
#pragma omp for schedule(static)
        for (int y = 0; y < ny; y++) fft_exec(px, d + (size_t)y * nx, scr, inverse);
#pragma omp for schedule(static)
        for (int x0 = 0; x0 < nx; x0 += CB) {
            const int nb = nx - x0 < CB ? nx - x0 : CB;
            for (int y = 0; y < ny; y++)
                for (int c = 0; c < nb; c++) col[c * ny + y] = d[(size_t)y * nx + x0 + c];
            for (int c = 0; c < nb; c++) fft_exec(py, col + c * ny, scr, inverse);
            for (int y = 0; y < ny; y++)
                for (int c = 0; c < nb; c++) d[(size_t)y * nx + x0 + c] = col[c * ny + y];
        }
        free(scr);
    }
}

int orc_fft2d(double* data, int32_t nx, int32_t ny, int32_t inverse) {
    fft_plan px, py;
    if (fft_plan_init(&px, nx)) return -1;
    if (fft_plan_init(&py, ny)) { fft_plan_free(&px); return -1; }
    fft2d_planned(&px, &py, (cplx*)data, nx, ny, inverse);
    fft_plan_free(&px);
    fft_plan_free(&py);
    return 0;
}

void orc_dft1d_naive(const double* in, double* out, int32_t n, int32_t inverse) {
    const cplx* x = (const cplx*)in;
    cplx* y = (cplx*)out;
    const double sg = inverse ? 1.0 : -1.0;
    for (int k = 0; k < n; k++) {
        long double re = 0, im = 0;
        for (int j = 0; j < n; j++) {
            long m = ((long)j * k) % n;
            long double a = sg * 2.0L * 3.14159265358979323846264338327950288L * m / n;
            long double c = cosl(a), s = sinl(a);
            re += creal(x[j]) * c - cimag(x[j]) * s;
            im += creal(x[j]) * s + cimag(x[j]) * c;
        }
        y[k] = (double)re + I * (double)im;
    }
}

/* ------------------------------------------------------------------ */
/* grid helpers                                                        */
/* ------------------------------------------------------------------ */

typedef struct {
    int nx, ny, nz, nzh;
    double dx, dy, dz, dzh;
    double lp, ls;
} grid_t;

static void grid_init(const shg_params* p, grid_t* g) {
    g->nx = p->nx;
    g->ny = p->ny;
    g->nz = p->nz;
    g->nzh = p->nz_heat > 0 ? p->nz_heat : p->nz;
    /* src/headers/Crystal.h:34-36 */
    g->dx = p->lx_um / (p->nx - 1);
    g->dy = p->ly_um / (p->ny - 1);
    g->dz = p->lz_um / (p->nz - 1);
    g->dzh = p->lz_um / (g->nzh - 1);
    /* src/headers/Crystal.h:26-27 */
    g->lp = p->lambda_p_um;
    g->ls = 0.5 * p->lambda_p_um;
}

/* ------------------------------------------------------------------ */
/* setup: pump, propagators                                            */
/* ------------------------------------------------------------------ */

/* Gaussian pump focused at focal_um -- src/headers/Efields.h:84-119 */
void orc_pump_init(const shg_params* p, double Tc, double* pump) {
    grid_t g;
    grid_init(p, &g);
    cplx* A = (cplx*)pump;
    const double np = orc_refractive_index(p->crystal, g.lp, Tc);
    const double uX = 0.5 * g.nx, uY = 0.5 * g.ny;
    const double w2 = p->waist_um * p->waist_um;
    const double zR = ORC_PI * np * w2 / g.lp;
    if (p->waist_y_um > 0) {
        /* elliptical variant, commented out at src/headers/Efields.h:113-115:
         * A = Ap0 / (sqrt(MX) sqrt(MY)) exp(-x^2/(wX^2 MX) - y^2/(wY^2 MY)),
         * with Ap0 normalised by wX wY (the reference's line keeps wX^2) */
        const double wy2 = p->waist_y_um * p->waist_y_um;
        const double zRY = ORC_PI * np * wy2 / g.lp;
        const double Ap0 = sqrt(4.0 * p->power_W / (ORC_EPS0 * ORC_C * ORC_PI * np * p->waist_um * p->waist_y_um));
        const cplx MX = 1.0 - I * (p->focal_um / zR), MY = 1.0 - I * (p->focal_um / zRY);
        const cplx amp = Ap0 / (csqrt(MX) * csqrt(MY));
        for (int y = 0; y < g.ny; y++)
            for (int x = 0; x < g.nx; x++) {
                const double rx = (x - uX) * g.dx, ry = (y - uY) * g.dy;
                A[(size_t)y * g.nx + x] = amp * cexp(-rx * rx / (w2 * MX) - ry * ry / (wy2 * MY));
            }
        return;
    }
    const double Ap0 = sqrt(4.0 * p->power_W / (ORC_EPS0 * ORC_C * ORC_PI * np * w2));
    const cplx M = 1.0 - I * (p->focal_um / zR);
    for (int y = 0; y < g.ny; y++)
        for (int x = 0; x < g.nx; x++) {
            const double rx = (x - uX) * g.dx, ry = (y - uY) * g.dy;
            A[(size_t)y * g.nx + x] = (Ap0 / M) * cexp((-rx * rx - ry * ry) / (w2 * M));
        }
}

/* Paraxial propagators in natural FFT order: beamPropagator followed by
 * fftShift2D -- src/headers/Efields.h:122-141 and :32-81 (even n). */
void orc_propagators(const shg_params* p, double Tc, double* prop_p, double* prop_s) {
    grid_t g;
    grid_init(p, &g);
    cplx* Pp = (cplx*)prop_p;
    cplx* Ps = (cplx*)prop_s;
    const double kp = 2 * ORC_PI * orc_refractive_index(p->crystal, g.lp, Tc) / g.lp;
    const double ks = 2 * ORC_PI * orc_refractive_index(p->crystal, g.ls, Tc) / g.ls;
    const double dfX = 1.0 / g.dx / g.nx, dfY = 1.0 / g.dy / g.ny;
    const double uX = 0.5 * g.nx, uY = 0.5 * g.ny;
    for (int y = 0; y < g.ny; y++)
        for (int x = 0; x < g.nx; x++) {
            /* centred index (i) of natural index (x) after the swap of halves */
            const int ic = (x + g.nx / 2) % g.nx, jc = (y + g.ny / 2) % g.ny;
            const double fx2 = dfX * dfX * (ic - uX) * (ic - uX);
            const double fy2 = dfY * dfY * (jc - uY) * (jc - uY);
            const double ap = -g.dz * (2 * ORC_PI * ORC_PI / kp * (fx2 + fy2) +
                                       2 * ORC_PI * dfX * tan(p->rho_p) * (ic - uX));
            const double as = -g.dz * (2 * ORC_PI * ORC_PI / ks * (fx2 + fy2) +
                                       2 * ORC_PI * dfX * tan(p->rho_s) * (ic - uX));
            Pp[(size_t)y * g.nx + x] = cos(ap) + I * sin(ap);
            Ps[(size_t)y * g.nx + x] = cos(as) + I * sin(as);
        }
}

double orc_power(const shg_params* p, const double* Ad, double n) {
    grid_t g;
    grid_init(p, &g);
    const cplx* A = (const cplx*)Ad;
    double s = 0;
    const size_t N = (size_t)g.nx * g.ny;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (size_t i = 0; i < N; i++) s += creal(A[i]) * creal(A[i]) + cimag(A[i]) * cimag(A[i]);
    return 0.5 * ORC_EPS0 * ORC_C * n * s * g.dx * g.dy;
}

/* ------------------------------------------------------------------ */
/* nonlinear step                                                      */
/* ------------------------------------------------------------------ */

/* scheme constants of one RK4 call (src/headers/Solver.h:231-257) */
typedef struct {
    double c1, c2, c3; /* stage input A + c*k */
    double w;          /* final A += w (k1 + 2k2 + 2k3 + k4) */
    double kmul;       /* coupling multiplier */
    double off[3];     /* z offsets in units of dz: stage 1, stages 2-3, 4 */
} rk_call;

static void scheme_calls(const shg_params* p, double dz, rk_call c[2]) {
    if (p->scheme == SHG_SCHEME_STRANG) {
        const double h = 0.5 * dz;
        for (int i = 0; i < 2; i++) {
            c[i].c1 = 0.5 * h; c[i].c2 = 0.5 * h; c[i].c3 = h;
            c[i].w = h / 6.0; c[i].kmul = 2.0;
            c[i].off[0] = 0.5 * i; c[i].off[1] = 0.5 * i + 0.25; c[i].off[2] = 0.5 * i + 0.5;
        }
    } else {
        /* Q1: stage inputs without dz; Q2: offsets 0.5*{0,0.5,0.5,1};
         * Q3/Q4: full dz, both calls at the same z */
        for (int i = 0; i < 2; i++) {
            c[i].c1 = 0.5; c[i].c2 = 0.5; c[i].c3 = 1.0;
            c[i].w = dz / 6.0; c[i].kmul = 1.0;
            c[i].off[0] = 0.0; c[i].off[1] = 0.25; c[i].off[2] = 0.5;
        }
    }
}

/* dA/dz -- src/headers/Solver.h:33-55 (THERMAL) and :57-80 */
static inline void dadz(cplx P, cplx S, cplx e, double kp, double ks, double ap,
                        double as, double bfac, cplx* dP, cplx* dS) {
    *dP = I * kp * S * conj(P) * conj(e) - 0.5 * ap * P;
    const double s2 = creal(S) * creal(S) + cimag(S) * cimag(S);
    *dS = I * ks * P * P * e - 0.5 * (as + bfac * s2) * S;
}

static inline void rk4_point(cplx* P, cplx* S, const cplx e[3], double kp, double ks,
                             double ap, double as, double bfac, const rk_call* c) {
    cplx k1p, k1s, k2p, k2s, k3p, k3s, k4p, k4s;
    const cplx p0 = *P, s0 = *S;
    dadz(p0, s0, e[0], kp, ks, ap, as, bfac, &k1p, &k1s);
    dadz(p0 + c->c1 * k1p, s0 + c->c1 * k1s, e[1], kp, ks, ap, as, bfac, &k2p, &k2s);
    dadz(p0 + c->c2 * k2p, s0 + c->c2 * k2s, e[1], kp, ks, ap, as, bfac, &k3p, &k3s);
    dadz(p0 + c->c3 * k3p, s0 + c->c3 * k3s, e[2], kp, ks, ap, as, bfac, &k4p, &k4s);
    *P = p0 + (k1p + 2.0 * k2p + 2.0 * k3p + k4p) * c->w;
    *S = s0 + (k1s + 2.0 * k2s + 2.0 * k3s + k4s) * c->w;
}

/* coupling constants: kappa_p = pi dQ/(n_p l_p), kappa_s = 1/2 pi dQ/(n_s l_s),
 * 2PA factor beta_s * 1/2 c eps0 n_s -- src/headers/Solver.h:72,74 */
static inline void couplings(const shg_params* p, const grid_t* g, double T, double kmul,
                             double* kp, double* ks, double* bfac, double* np_,
                             double* ns_) {
    const double dQ = 2.0 * p->deff_um_per_V / ORC_PI;
    const double np = orc_refractive_index(p->crystal, g->lp, T);
    const double ns = orc_refractive_index(p->crystal, g->ls, T);
    *kp = kmul * ORC_PI * dQ / (np * g->lp);
    *ks = kmul * 0.5 * ORC_PI * dQ / (ns * g->ls);
    *bfac = p->beta_s * 0.5 * ORC_C * ORC_EPS0 * ns;
    if (np_) *np_ = np;
    if (ns_) *ns_ = ns;
}

/* One RK4 call over the grid, non-THERMAL: n at the scalar T_C and
 * phase (z_s + off*dz) dk + poling[s] -- src/headers/Solver.h:231-257 */
static void rk4_const(const shg_params* p, const grid_t* g, int s, double dk,
                      const rk_call* c, cplx* P, cplx* S) {
    double kp, ks, bfac;
    couplings(p, g, p->T_C, c->kmul, &kp, &ks, &bfac, NULL, NULL);
    const double pol = p->poling_phase ? p->poling_phase[s] : 0.0;
    cplx e[3];
    for (int i = 0; i < 3; i++) {
        const double ph = ((double)s * g->dz + c->off[i] * g->dz) * dk + pol;
        e[i] = cos(ph) + I * sin(ph);
    }
    const size_t N = (size_t)g->nx * g->ny;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; i++)
        rk4_point(&P[i], &S[i], e, kp, ks, p->alpha_p, p->alpha_s, bfac, c);
}

/* One RK4 call, THERMAL: n(T[x,y,h]) per point and phase
 * DKint (s + off) dz / (s + 1) -- src/headers/Solver.h:33-55,198-229 */
static void rk4_thermal(const shg_params* p, const grid_t* g, int s, const double* Th,
                        const double* dkint, const rk_call* c, cplx* P, cplx* S) {
    const double pol = p->poling_phase ? p->poling_phase[s] : 0.0;
    const size_t N = (size_t)g->nx * g->ny;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; i++) {
        double kp, ks, bfac;
        couplings(p, g, Th[i], c->kmul, &kp, &ks, &bfac, NULL, NULL);
        cplx e[3];
        for (int j = 0; j < 3; j++) {
            const double ph = dkint[i] * ((double)s + c->off[j]) * g->dz / ((double)s + 1.0) + pol;
            e[j] = cos(ph) + I * sin(ph);
        }
        rk4_point(&P[i], &S[i], e, kp, ks, p->alpha_p, p->alpha_s, bfac, c);
    }
}

/* A <- IFFT2(Prop . FFT2(A)) / (nx ny) -- src/headers/Solver.h:153-186 */
static void diffraction(const grid_t* g, const fft_plan* px, const fft_plan* py,
                        cplx* A, const cplx* prop) {
    const size_t N = (size_t)g->nx * g->ny;
    fft2d_planned(px, py, A, g->nx, g->ny, 0);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; i++) A[i] *= prop[i];
    fft2d_planned(px, py, A, g->nx, g->ny, 1);
    const double sc = 1.0 / (double)N;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; i++) A[i] *= sc;
}

void orc_ssfm_step(const shg_params* p, int32_t s, double* Pd, double* Sd) {
    grid_t g;
    grid_init(p, &g);
    fft_plan px, py;
    fft_plan_init(&px, g.nx);
    fft_plan_init(&py, g.ny);
    const size_t N = (size_t)g.nx * g.ny;
    cplx* pp = (cplx*)malloc(sizeof(cplx) * N * 2);
    orc_propagators(p, p->T_C, (double*)pp, (double*)(pp + N));
    rk_call c[2];
    scheme_calls(p, g.dz, c);
    const double dk = orc_delta_k(p, p->T_C);
    rk4_const(p, &g, s, dk, &c[0], (cplx*)Pd, (cplx*)Sd);
    diffraction(&g, &px, &py, (cplx*)Pd, pp);
    diffraction(&g, &px, &py, (cplx*)Sd, pp + N);
    rk4_const(p, &g, s, dk, &c[1], (cplx*)Pd, (cplx*)Sd);
    free(pp);
    fft_plan_free(&px);
    fft_plan_free(&py);
}

/* ------------------------------------------------------------------ */
/* heat                                                                */
/* ------------------------------------------------------------------ */

#define I3(x, y, z) (((size_t)(z) * g.ny + (y)) * g.nx + (x))

/* One Jacobi sweep Tf = F(Ti).  Open-top boundary set (the reference's):
 * src/headers/Tfield.h:93-201, the y = 0 plane is the Dirichlet oven and
 * is copied through (the reference never writes it).  Oven-surrounded set
 * (SHG_HEAT_BC_OVEN_SURROUNDED): src/headers/Tfield.h:205-240, the four
 * lateral planes are the Dirichlet oven, inner nodes as above, the z = 0 and
 * z = lz faces convective without Q. */
void orc_jacobi_sweep(const shg_params* p, int32_t nzh, const double* Ti, double* Tf,
                      const double* Q) {
    grid_t g;
    grid_init(p, &g);
    g.nz = nzh;
    const double dx = g.dx, dy = g.dy, dz = p->lz_um / (nzh - 1);
    const double kt = p->k_th, hc = p->h_conv, Tinf = p->T_inf_C;
    /* src/headers/Tfield.h:99-106 */
    const double Bix = hc * dx / kt, Biy = hc * dy / kt, Biz = hc * dz / kt;
    const double ax = dx * dx / (2 * dy * dy), ay = dy * dy / (2 * dx * dx),
                 az = dz * dz / (2 * dx * dx);
    const double bx = dx * dx / (2 * dz * dz), by = dy * dy / (2 * dz * dz),
                 bz = dz * dz / (2 * dy * dy);
    const double av = dz / (2 * dx), bv = dx / (2 * dz), cv = dx * dz / (4 * dy * dy),
                 V = 1 / (av + bv + 2 * cv + 0.5 * (Bix + Biz));
    const double ah = dz / (2 * dy), bh = dy / (2 * dz), ch = dy * dz / (4 * dx * dx),
                 H = 1 / (ah + bh + 2 * ch + 0.5 * (Biy + Biz));
    const double ad = dx / (2 * dy), bd = dy / (2 * dx), cd = dx * dy / (4 * dz * dz),
                 D = 1 / (ad + bd + 2 * cd + 0.5 * (Bix + Biy));
    const double AV = dy * dz / dx, BV = dx * dz / dy, CV = dx * dy / dz,
                 BiD = Bix * dy + Biy * dz + Biz * dx, GV = 1 / (AV + BV + CV + BiD);
    const double L2 = 1 / (2 / (dx * dx) + 2 / (dy * dy) + 2 / (dz * dz));
    const int NX = g.nx, NY = g.ny, NZ = nzh;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < NZ; z++)
        for (int y = 0; y < NY; y++)
            for (int x = 0; x < NX; x++) {
                const size_t i = I3(x, y, z);
                if (p->heat_bc == SHG_HEAT_BC_OVEN_SURROUNDED) {
                    if (x == 0 || x == NX - 1 || y == 0 || y == NY - 1) { Tf[i] = Ti[i]; continue; }
                    if (z > 0 && z < NZ - 1) { /* inner points, :217-219 */
                        Tf[i] = L2 * ((Ti[I3(x + 1, y, z)] + Ti[I3(x - 1, y, z)]) / (dx * dx) +
                                      (Ti[I3(x, y + 1, z)] + Ti[I3(x, y - 1, z)]) / (dy * dy) +
                                      (Ti[I3(x, y, z + 1)] + Ti[I3(x, y, z - 1)]) / (dz * dz) +
                                      Q[i] / kt);
                    } else { /* back/front faces, :224-232 */
                        const int zn = z == 0 ? z + 1 : z - 1;
                        Tf[i] = (Ti[I3(x, y, zn)] + az * (Ti[I3(x + 1, y, z)] + Ti[I3(x - 1, y, z)]) +
                                 bz * (Ti[I3(x, y + 1, z)] + Ti[I3(x, y - 1, z)]) + Biz * Tinf) /
                                (1 + 2 * az + 2 * bz + Biz);
                    }
                    continue;
                }
                if (y == 0) { Tf[i] = Ti[i]; continue; }
                const int xi = x > 0 && x < NX - 1, yi = y < NY - 1, zi = z > 0 && z < NZ - 1;
                const int xe = !xi, ze = !zi; /* on an x / z boundary */
                const int xn = x == 0 ? x + 1 : x - 1; /* inward neighbour */
                const int zn = z == 0 ? z + 1 : z - 1;
                double v;
                if (xi && yi && zi) { /* inner points, :113-115 */
                    v = L2 * ((Ti[I3(x + 1, y, z)] + Ti[I3(x - 1, y, z)]) / (dx * dx) +
                              (Ti[I3(x, y + 1, z)] + Ti[I3(x, y - 1, z)]) / (dy * dy) +
                              (Ti[I3(x, y, z + 1)] + Ti[I3(x, y, z - 1)]) / (dz * dz) +
                              Q[i] / kt);
                } else if (xe && yi && zi) { /* left/right faces, :119-125 */
                    v = (Ti[I3(xn, y, z)] + ax * (Ti[I3(x, y + 1, z)] + Ti[I3(x, y - 1, z)]) +
                         bx * (Ti[I3(x, y, z + 1)] + Ti[I3(x, y, z - 1)]) + Bix * Tinf) /
                        (1 + 2 * ax + 2 * bx + Bix);
                } else if (xi && !yi && zi) { /* top face, :127-129 */
                    v = (Ti[I3(x, y - 1, z)] + ay * (Ti[I3(x + 1, y, z)] + Ti[I3(x - 1, y, z)]) +
                         by * (Ti[I3(x, y, z + 1)] + Ti[I3(x, y, z - 1)]) + Biy * Tinf) /
                        (1 + 2 * ay + 2 * by + Biy);
                } else if (xi && yi && ze) { /* back/front faces, :131-137 */
                    v = (Ti[I3(x, y, zn)] + az * (Ti[I3(x + 1, y, z)] + Ti[I3(x - 1, y, z)]) +
                         bz * (Ti[I3(x, y + 1, z)] + Ti[I3(x, y - 1, z)]) + Biz * Tinf) /
                        (1 + 2 * az + 2 * bz + Biz);
                } else if (xe && yi && ze) { /* vertical sides, :143-157 */
                    v = (av * Ti[I3(xn, y, z)] + bv * Ti[I3(x, y, zn)] +
                         cv * (Ti[I3(x, y + 1, z)] + Ti[I3(x, y - 1, z)]) +
                         0.5 * Tinf * (Bix + Biz)) * V;
                } else if (xi && !yi && ze) { /* horizontal sides, :161-167 */
                    v = (ah * Ti[I3(x, y - 1, z)] + bh * Ti[I3(x, y, zn)] +
                         ch * (Ti[I3(x - 1, y, z)] + Ti[I3(x + 1, y, z)]) +
                         0.5 * Tinf * (Biy + Biz)) * H;
                } else if (xe && !yi && zi) { /* depth sides, :171-177 */
                    v = (ad * Ti[I3(x, y - 1, z)] + bd * Ti[I3(xn, y, z)] +
                         cd * (Ti[I3(x, y, z + 1)] + Ti[I3(x, y, z - 1)]) +
                         0.5 * Tinf * (Bix + Biy)) * D;
                } else { /* vertices, :181-196 */
                    v = (AV * Ti[I3(xn, y, z)] + BV * Ti[I3(x, y - 1, z)] +
                         CV * Ti[I3(x, y, zn)] + Tinf * BiD) * GV;
                }
                Tf[i] = v;
            }
}

/* Symmetric tridiagonal eigenproblem by the implicit QL method with
 * Wilkinson shifts (Golub & Van Loan, Matrix Computations, 8.3).
 * diag[n], off[n-1] (off[i] couples i, i+1).  evecs[i*n + k] = component i
 * of eigenvector k.  Returns 0 or -1 (no convergence). */
int orc_tridiag_eigen(int32_t n, const double* diag, const double* off, double* evals,
                      double* evecs) {
    double* d = evals;
    double* e = (double*)calloc((size_t)n + 1, sizeof(double));
    for (int i = 0; i < n; i++) d[i] = diag[i];
    for (int i = 0; i + 1 < n; i++) e[i] = off[i];
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) evecs[(size_t)i * n + k] = i == k ? 1.0 : 0.0;
    for (int l = 0; l < n; l++) {
        int iter = 0, m;
        for (;;) {
            for (m = l; m < n - 1; m++) {
                const double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) + dd == dd) break;
            }
            if (m == l) break;
            if (++iter > 200) { free(e); return -1; }
            /* shift from the trailing 2x2 of the active block at l */
            double gg = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hypot(gg, 1.0);
            gg = d[m] - d[l] + e[l] / (gg + (gg >= 0 ? fabs(r) : -fabs(r)));
            double s = 1.0, c = 1.0, pp = 0.0;
            int i;
            int deflated = 0;
            for (i = m - 1; i >= l; i--) {
                double f = s * e[i], b = c * e[i];
                r = hypot(f, gg);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= pp;
                    e[m] = 0.0;
                    deflated = 1;
                    break;
                }
                s = f / r;
                c = gg / r;
                gg = d[i + 1] - pp;
                r = (d[i] - gg) * s + 2.0 * c * b;
                pp = s * r;
                d[i + 1] = gg + pp;
                gg = c * r - b;
                for (int k = 0; k < n; k++) {
                    double* zr = evecs + (size_t)k * n;
                    f = zr[i + 1];
                    zr[i + 1] = s * zr[i] + c * f;
                    zr[i] = c * zr[i] - s * f;
                }
            }
            if (deflated) continue;
            d[l] -= pp;
            e[l] = gg;
            e[m] = 0.0;
        }
    }
    free(e);
    return 0;
}

/* 1D axis of the tensor-product finite-volume form of the Jacobi system */
typedef struct {
    int n;
    double* w;    /* control-volume width */
    double* lam;  /* eigenvalues of W^-1/2 K W^-1/2 */
    double* q;    /* eigenvectors q[i*n + k] */
    double* qt;   /* transpose */
} heat_axis;

/* K = conductances (1/h between neighbours, + h_conv/k_th at Robin ends);
 * robin_lo/hi select the Robin ends; dirichlet_lo/hi couple the end node to
 * a fixed value one spacing beyond it (a removed unknown). */
static int heat_axis_init(heat_axis* a, int n, double h, int robin_lo, int robin_hi,
                          int dirichlet_lo, int dirichlet_hi, double hk) {
    a->n = n;
    a->w = (double*)malloc(sizeof(double) * n);
    a->lam = (double*)malloc(sizeof(double) * n);
    a->q = (double*)malloc(sizeof(double) * (size_t)n * n);
    a->qt = (double*)malloc(sizeof(double) * (size_t)n * n);
    double* dg = (double*)malloc(sizeof(double) * n);
    double* of = (double*)malloc(sizeof(double) * (n > 1 ? n - 1 : 1));
    for (int i = 0; i < n; i++) {
        const int lo_end = i == 0 && !dirichlet_lo, hi_end = i == n - 1 && !dirichlet_hi;
        a->w[i] = (lo_end || hi_end) ? 0.5 * h : h;
        double kd = 0;
        if (i > 0 || dirichlet_lo) kd += 1.0 / h;
        if (i < n - 1 || dirichlet_hi) kd += 1.0 / h;
        if (i == 0 && robin_lo) kd += hk;
        if (i == n - 1 && robin_hi) kd += hk;
        dg[i] = kd / a->w[i];
    }
    for (int i = 0; i + 1 < n; i++) of[i] = -1.0 / h / sqrt(a->w[i] * a->w[i + 1]);
    int rc = orc_tridiag_eigen(n, dg, of, a->lam, a->q);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) a->qt[(size_t)k * n + i] = a->q[(size_t)i * n + k];
    free(dg);
    free(of);
    return rc;
}

static void heat_axis_free(heat_axis* a) {
    free(a->w); free(a->lam); free(a->q); free(a->qt);
}

/* out[..i..] = sum_j M[j*n + i] in[..j..] along one axis of u[nz][ny][nx] */
static void apply_axis(const double* M, int axis, int nx, int ny, int nz, const double* in,
                       double* out) {
    if (axis == 0) {
#pragma omp parallel for schedule(static)
        for (long r = 0; r < (long)ny * nz; r++) {
            const double* a = in + (size_t)r * nx;
            double* o = out + (size_t)r * nx;
            for (int i = 0; i < nx; i++) o[i] = 0;
            for (int j = 0; j < nx; j++) {
                const double aj = a[j];
                const double* mr = M + (size_t)j * nx;
                for (int i = 0; i < nx; i++) o[i] += mr[i] * aj;
            }
        }
    } else if (axis == 1) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int z = 0; z < nz; z++)
            for (int i = 0; i < ny; i++) {
                double* o = out + ((size_t)z * ny + i) * nx;
                for (int x = 0; x < nx; x++) o[x] = 0;
                for (int j = 0; j < ny; j++) {
                    const double m = M[(size_t)j * ny + i];
                    const double* a = in + ((size_t)z * ny + j) * nx;
                    for (int x = 0; x < nx; x++) o[x] += m * a[x];
                }
            }
    } else {
        const size_t plane = (size_t)nx * ny;
#pragma omp parallel for schedule(static)
        for (long c0 = 0; c0 < (long)plane; c0 += 256) {
            const size_t cn = plane - c0 < 256 ? plane - c0 : 256;
            for (int i = 0; i < nz; i++) {
                double* o = out + (size_t)i * plane + c0;
                for (size_t c = 0; c < cn; c++) o[c] = 0;
                for (int j = 0; j < nz; j++) {
                    const double m = M[(size_t)j * nz + i];
                    const double* a = in + (size_t)j * plane + c0;
                    for (size_t c = 0; c < cn; c++) o[c] += m * a[c];
                }
            }
        }
    }
}

/* Exact solution of the fixed point of orc_jacobi_sweep (the discrete
 * steady state) by fast diagonalisation of its tensor-product
 * finite-volume form  A = Kx(x)Wy(x)Wz + Wx(x)Ky(x)Wz + Wx(x)Wy(x)Kz. */
int orc_heat_solve_direct(const shg_params* p, int32_t nzh, const double* Q, double* T) {
    grid_t g;
    grid_init(p, &g);
    g.nz = nzh;
    const int NX = g.nx, NY = g.ny, NZ = nzh;
    const int oven = p->heat_bc == SHG_HEAT_BC_OVEN_SURROUNDED;
    /* unknowns: x in [X0, X0+NXu), y in [1, 1+NYu), all z */
    const int X0 = oven ? 1 : 0, NXu = oven ? NX - 2 : NX, NYu = oven ? NY - 2 : NY - 1;
    const double dx = g.dx, dy = g.dy, dz = p->lz_um / (nzh - 1);
    const double hk = p->h_conv / p->k_th, Tinf = p->T_inf_C, Tpel = p->T_peltier_C;
    heat_axis ax, ay, az;
    int rc = oven ? heat_axis_init(&ax, NXu, dx, 0, 0, 1, 1, hk) : heat_axis_init(&ax, NX, dx, 1, 1, 0, 0, hk);
    rc |= oven ? heat_axis_init(&ay, NYu, dy, 0, 0, 1, 1, hk) : heat_axis_init(&ay, NYu, dy, 0, 1, 1, 0, hk);
    rc |= heat_axis_init(&az, NZ, dz, 1, 1, 0, 0, hk);
    const size_t NU = (size_t)NXu * NYu * NZ;
    double* u = (double*)malloc(sizeof(double) * NU);
    double* v = (double*)malloc(sizeof(double) * NU);
    if (rc || !u || !v) { rc = -1; goto done; }
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < NZ; z++)
        for (int j = 0; j < NYu; j++)
            for (int xu = 0; xu < NXu; xu++) {
                const int y = j + 1, x = xu + X0;
                const double wx = ax.w[xu], wy = ay.w[j], wz = az.w[z];
                double b = 0;
                const int xi = x > 0 && x < NX - 1, yi = y < NY - 1, zi = z > 0 && z < NZ - 1;
                if (xi && yi && zi) b += Q[I3(x, y, z)] / p->k_th * wx * wy * wz; /* Q8 */
                if (z == 0) b += hk * Tinf * wx * wy;
                if (z == NZ - 1) b += hk * Tinf * wx * wy;
                if (j == 0) b += Tpel / dy * wx * wz;
                if (oven) {
                    if (j == NYu - 1) b += Tpel / dy * wx * wz;
                    if (xu == 0) b += Tpel / dx * wy * wz;
                    if (xu == NXu - 1) b += Tpel / dx * wy * wz;
                } else {
                    if (x == 0) b += hk * Tinf * wy * wz;
                    if (x == NX - 1) b += hk * Tinf * wy * wz;
                    if (y == NY - 1) b += hk * Tinf * wx * wz;
                }
                u[((size_t)z * NYu + j) * NXu + xu] = b / sqrt(wx * wy * wz);
            }
    apply_axis(ax.q, 0, NXu, NYu, NZ, u, v);
    apply_axis(ay.q, 1, NXu, NYu, NZ, v, u);
    apply_axis(az.q, 2, NXu, NYu, NZ, u, v);
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < NZ; z++)
        for (int j = 0; j < NYu; j++)
            for (int xu = 0; xu < NXu; xu++)
                v[((size_t)z * NYu + j) * NXu + xu] /= ax.lam[xu] + ay.lam[j] + az.lam[z];
    apply_axis(az.qt, 2, NXu, NYu, NZ, v, u);
    apply_axis(ay.qt, 1, NXu, NYu, NZ, u, v);
    apply_axis(ax.qt, 0, NXu, NYu, NZ, v, u);
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < NZ; z++)
        for (int y = 0; y < NY; y++)
            for (int x = 0; x < NX; x++) {
                const int j = y - 1, xu = x - X0;
                if (j < 0 || j >= NYu || xu < 0 || xu >= NXu) { T[I3(x, y, z)] = Tpel; continue; }
                T[I3(x, y, z)] = u[((size_t)z * NYu + j) * NXu + xu] /
                                 sqrt(ax.w[xu] * ay.w[j] * az.w[z]);
            }
done:
    free(u);
    free(v);
    heat_axis_free(&ax);
    heat_axis_free(&ay);
    heat_axis_free(&az);
    return rc;
}

static double diff_norm(const double* a, const double* b, size_t n) {
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (size_t i = 0; i < n; i++) s += (a[i] - b[i]) * (a[i] - b[i]);
    return sqrt(s);
}

/* Q = a_p I_p + a_s I_s + b_s I_s^2, I = 1/2 c eps0 n(T)|A|^2
 * -- src/headers/Tfield.h:260-276 */
static inline double heat_source(const shg_params* p, const grid_t* g, double T, cplx P,
                                 cplx S) {
    const double np = orc_refractive_index(p->crystal, g->lp, T);
    const double ns = orc_refractive_index(p->crystal, g->ls, T);
    const double Ip = 0.5 * ORC_C * ORC_EPS0 * np * (creal(P) * creal(P) + cimag(P) * cimag(P));
    const double Is = 0.5 * ORC_C * ORC_EPS0 * ns * (creal(S) * creal(S) + cimag(S) * cimag(S));
    return p->alpha_p * Ip + p->alpha_s * Is + p->beta_s * Is * Is;
}

/* ------------------------------------------------------------------ */
/* the run: Solver::run / Solver::CWES                                 */
/* ------------------------------------------------------------------ */

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* One THERMAL z-march on a given temperature field T[nzh][ny][nx]: the
 * heat source it deposits, Q[nzh][ny][nx] averaged over the optical slices
 * of each heat slice as in the outer loop below (src/headers/Solver.h:
 * 299-301, src/headers/Tfield.h:260-276), and the pump and SH powers after
 * every slice, power_after[2 s] and power_after[2 s + 1] (n at the centre
 * temperature, as the run's results).  Test infrastructure: pins Q against
 * the power the fields lose. */
int orc_march_heat(const shg_params* p, const double* T, double* Q, double* power_after) {
    if (p->nx < 4 || p->ny < 4 || p->nz < 2 || (p->nx & 1) || (p->ny & 1)) return SHG_EINVAL;
    grid_t g;
    grid_init(p, &g);
    if (g.nzh < 2) return SHG_EINVAL;
    if (p->thermal && g.nzh > g.nz) return SHG_EINVAL; /* a heat slab without an optical slice */
    const size_t N = (size_t)g.nx * g.ny, NV = N * g.nzh;
    fft_plan px, py;
    if (fft_plan_init(&px, g.nx) || fft_plan_init(&py, g.ny)) return SHG_EINVAL;
    cplx* P = (cplx*)calloc(N, sizeof(cplx));
    cplx* S = (cplx*)calloc(N, sizeof(cplx));
    cplx* prop = (cplx*)calloc(2 * N, sizeof(cplx));
    double* qsum = (double*)calloc(NV, sizeof(double));
    int* qcnt = (int*)calloc(g.nzh, sizeof(int));
    double* dk = (double*)malloc(sizeof(double) * NV);
    double* dkint = (double*)calloc(N, sizeof(double));
    rk_call calls[2];
    scheme_calls(p, g.dz, calls);
    for (size_t i = 0; i < NV; i++) dk[i] = orc_delta_k(p, T[i]);
    const double Tc = T[I3(g.nx / 2, g.ny / 2, 0)];
    orc_pump_init(p, Tc, (double*)P);
    orc_propagators(p, Tc, (double*)prop, (double*)(prop + N));
    for (int s = 0; s < g.nz; s++) {
        const int h = (int)(((long)s * g.nzh) / g.nz);
        const double* Th = T + (size_t)h * N;
        const double* dkh = dk + (size_t)h * N;
        for (size_t i = 0; i < N; i++) dkint[i] += dkh[i];
        rk4_thermal(p, &g, s, Th, dkint, &calls[0], P, S);
        diffraction(&g, &px, &py, P, prop);
        diffraction(&g, &px, &py, S, prop + N);
        rk4_thermal(p, &g, s, Th, dkint, &calls[1], P, S);
        double* qh = qsum + (size_t)h * N;
        for (size_t i = 0; i < N; i++) qh[i] += heat_source(p, &g, Th[i], P[i], S[i]);
        qcnt[h]++;
        if (power_after) {
            power_after[2 * s] = orc_power(p, (double*)P, orc_refractive_index(p->crystal, g.lp, Tc));
            power_after[2 * s + 1] = orc_power(p, (double*)S, orc_refractive_index(p->crystal, g.ls, Tc));
        }
    }
    for (int h = 0; h < g.nzh; h++)
        for (size_t i = 0; i < N; i++) Q[(size_t)h * N + i] = qcnt[h] ? qsum[(size_t)h * N + i] / qcnt[h] : 0.0;
    free(P); free(S); free(prop); free(qsum); free(qcnt); free(dk); free(dkint);
    fft_plan_free(&px);
    fft_plan_free(&py);
    return SHG_OK;
}

int orc_run(const shg_params* p, shg_result* r) {
    const double t0 = now_s();
    if (p->nx < 4 || p->ny < 4 || p->nz < 2 || (p->nx & 1) || (p->ny & 1)) return SHG_EINVAL;
    grid_t g;
    grid_init(p, &g);
    if (g.nzh < 2) return SHG_EINVAL;
    if (p->thermal && g.nzh > g.nz) return SHG_EINVAL; /* a heat slab without an optical slice */
    const size_t N = (size_t)g.nx * g.ny, NV = N * g.nzh;
    fft_plan px, py;
    if (fft_plan_init(&px, g.nx) || fft_plan_init(&py, g.ny)) return SHG_EINVAL;
    cplx* P = (cplx*)calloc(N, sizeof(cplx));
    cplx* S = (cplx*)calloc(N, sizeof(cplx));
    cplx* Po = (cplx*)calloc(N, sizeof(cplx));
    cplx* So = (cplx*)calloc(N, sizeof(cplx));
    cplx* prop = (cplx*)calloc(2 * N, sizeof(cplx));
    rk_call calls[2];
    scheme_calls(p, g.dz, calls);
    int rc = SHG_OK;
    long steps = 0, sweeps = 0;
    int outer = 0;
    double Tc = p->T_C, heat_res = 0;
    double* T = NULL;

    if (!p->thermal) {
        /* src/headers/Solver.h:376-383 */
        orc_pump_init(p, Tc, (double*)P);
        orc_propagators(p, Tc, (double*)prop, (double*)(prop + N));
        const double dk = orc_delta_k(p, Tc);
        for (int s = 0; s < g.nz; s++) {
            if (s == g.nz - 1) { /* Q6: saved before the last slice */
                memcpy(Po, P, sizeof(cplx) * N);
                memcpy(So, S, sizeof(cplx) * N);
            }
            rk4_const(p, &g, s, dk, &calls[0], P, S);
            diffraction(&g, &px, &py, P, prop);
            diffraction(&g, &px, &py, S, prop + N);
            rk4_const(p, &g, s, dk, &calls[1], P, S);
            steps++;
        }
        outer = 1;
    } else {
        /* src/headers/Solver.h:319-370 */
        T = (double*)malloc(sizeof(double) * NV);
        double* Tn = (double*)malloc(sizeof(double) * NV);
        double* Q = (double*)calloc(NV, sizeof(double));
        double* qsum = (double*)calloc(NV, sizeof(double));
        int* qcnt = (int*)calloc(g.nzh, sizeof(int));
        double* dk = (double*)malloc(sizeof(double) * NV);
        double* dkint = (double*)malloc(sizeof(double) * N);
        for (size_t i = 0; i < NV; i++) T[i] = p->T_C;             /* setTemperature */
        for (int z = 0; z < g.nzh; z++)                            /* setBottomOvens */
            for (int x = 0; x < g.nx; x++) T[I3(x, 0, z)] = p->T_peltier_C;
        if (p->heat_bc == SHG_HEAT_BC_OVEN_SURROUNDED)              /* setOvenSurrounded */
            for (int z = 0; z < g.nzh; z++)
                for (int y = 0; y < g.ny; y++)
                    for (int x = 0; x < g.nx; x++)
                        if (x == 0 || x == g.nx - 1 || y == g.ny - 1) T[I3(x, y, z)] = p->T_peltier_C;
        const int max_outer = p->max_outer > 0 ? p->max_outer : 20;
        const int chk = p->check_every > 0 ? p->check_every : 5000;
        for (int gi = 0; gi < max_outer; gi++) {
            long counter = 0;
            if (p->heat_solver == SHG_HEAT_JACOBI_REFERENCE) {
                /* src/headers/Solver.h:332-356 */
                double red = INFINITY, aux = NAN;
                while (red >= p->jacobi_tol) {
                    orc_jacobi_sweep(p, g.nzh, T, Tn, Q);
                    if (counter % chk == 0) {
                        if (counter > 0) aux = red;
                        red = diff_norm(Tn, T, NV);
                        if (red == aux) red = 0.5 * p->jacobi_tol;
                    }
                    memcpy(T, Tn, sizeof(double) * NV);
                    counter++;
                    if (p->max_sweeps > 0 && counter > p->max_sweeps) { rc = SHG_ENOCONV; break; }
                }
                heat_res = red;
            } else {
                orc_jacobi_sweep(p, g.nzh, T, Tn, Q);
                const double r0 = diff_norm(Tn, T, NV);
                if (r0 < p->jacobi_tol) {
                    memcpy(T, Tn, sizeof(double) * NV);
                    counter = 1;
                    heat_res = r0;
                } else {
                    if (orc_heat_solve_direct(p, g.nzh, Q, T)) { rc = SHG_ENOCONV; break; }
                    orc_jacobi_sweep(p, g.nzh, T, Tn, Q);
                    heat_res = diff_norm(Tn, T, NV);
                }
            }
            if (rc) break;
            sweeps += counter;
            /* setDKFromTemperature, src/headers/PhaseMatching.h:47-65 */
#pragma omp parallel for schedule(static)
            for (size_t i = 0; i < NV; i++) dk[i] = orc_delta_k(p, T[i]);
            Tc = T[I3(g.nx / 2, g.ny / 2, 0)];
            orc_pump_init(p, Tc, (double*)P);
            memset(S, 0, sizeof(cplx) * N);
            orc_propagators(p, Tc, (double*)prop, (double*)(prop + N));
            memset(dkint, 0, sizeof(double) * N);
            memset(qsum, 0, sizeof(double) * NV);
            memset(qcnt, 0, sizeof(int) * g.nzh);
            for (int s = 0; s < g.nz; s++) {
                if (s == g.nz - 1) {
                    memcpy(Po, P, sizeof(cplx) * N);
                    memcpy(So, S, sizeof(cplx) * N);
                }
                const int h = (int)(((long)s * g.nzh) / g.nz);
                const double* Th = T + (size_t)h * N;
                const double* dkh = dk + (size_t)h * N;
                for (size_t i = 0; i < N; i++) dkint[i] += dkh[i];
                rk4_thermal(p, &g, s, Th, dkint, &calls[0], P, S);
                diffraction(&g, &px, &py, P, prop);
                diffraction(&g, &px, &py, S, prop + N);
                rk4_thermal(p, &g, s, Th, dkint, &calls[1], P, S);
                /* Tfield::upDateQ after the slice, src/headers/Solver.h:299-301 */
                double* qh = qsum + (size_t)h * N;
#pragma omp parallel for schedule(static)
                for (size_t i = 0; i < N; i++) qh[i] += heat_source(p, &g, Th[i], P[i], S[i]);
                qcnt[h]++;
                steps++;
            }
            for (int h = 0; h < g.nzh; h++) {
                if (!qcnt[h]) continue; /* unreachable: nz_heat <= nz */
                for (size_t i = 0; i < N; i++)
                    Q[(size_t)h * N + i] = qsum[(size_t)h * N + i] / qcnt[h];
            }
            outer = gi + 1;
            if (counter == 1 && gi != 0) break; /* src/headers/Solver.h:367 */
        }
        free(Tn); free(Q); free(qsum); free(qcnt); free(dk); free(dkint);
    }

    if (r) {
        const double np = orc_refractive_index(p->crystal, g.lp, Tc);
        const double ns = orc_refractive_index(p->crystal, g.ls, Tc);
        r->P_sh_W = orc_power(p, (double*)So, ns);
        r->P_pump_out_W = orc_power(p, (double*)Po, np);
        {
            cplx* P0 = (cplx*)malloc(sizeof(cplx) * N);
            orc_pump_init(p, Tc, (double*)P0);
            r->P_pump_in_W = orc_power(p, (double*)P0, np);
            free(P0);
        }
        r->eta = r->P_sh_W / p->power_W;
        r->steps_done = steps;
        r->jacobi_sweeps = sweeps;
        r->outer_iters = outer;
        r->heat_residual_K = heat_res;
        r->peak_dT_K = 0;
        r->T_max_C = Tc;
        if (T) {
            double mx = -INFINITY;
            for (size_t i = 0; i < NV; i++) mx = T[i] > mx ? T[i] : mx;
            r->T_max_C = mx;
            r->peak_dT_K = mx - p->T_peltier_C;
            if (r->temperature_out) memcpy(r->temperature_out, T, sizeof(double) * NV);
        }
        if (r->pump_out) memcpy(r->pump_out, Po, sizeof(cplx) * N);
        if (r->signal_out) memcpy(r->signal_out, So, sizeof(cplx) * N);
        r->status = rc;
        r->wall_s = now_s() - t0;
    }
    free(T);
    free(P); free(S); free(Po); free(So); free(prop);
    fft_plan_free(&px);
    fft_plan_free(&py);
    return rc;
}
