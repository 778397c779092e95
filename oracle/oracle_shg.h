/*
 * oracle_shg.h -- FP64 CPU restatement of cuSHG's SSFM propagation loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this library, and only as the checker
 * (or the timed CPU baseline).  The product library never links it.
 *
 * Parameters and results use the public structs of include/shg.h.
 * Complex arrays are interleaved (re, im) doubles, index y*nx + x.
 */
#ifndef ORACLE_SHG_H
#define ORACLE_SHG_H

#include <stdint.h>
#include "../include/shg.h"

#ifdef __cplusplus
extern "C" {
#endif

void orc_set_threads(int n);
int orc_get_threads(void);

/* material */
double orc_refractive_index(int32_t crystal, double lambda_um, double T_C);
double orc_delta_k(const shg_params* p, double T_C);

/* FFT: unnormalised, forward sign -1, inverse +1 (cuFFT convention) */
int orc_fft1d(double* data, int32_t n, int32_t howmany, int32_t inverse);
int orc_fft2d(double* data, int32_t nx, int32_t ny, int32_t inverse);
void orc_dft1d_naive(const double* in, double* out, int32_t n, int32_t inverse);

/* setup kernels */
void orc_pump_init(const shg_params* p, double T_center, double* pump);
void orc_propagators(const shg_params* p, double T_center, double* prop_p,
                     double* prop_s);

/* One SSFM slice on (P, S) for a non-THERMAL run at slice index s. */
void orc_ssfm_step(const shg_params* p, int32_t s, double* P, double* S);

/* Power of a field: 1/2 eps0 c n sum|A|^2 dx dy */
double orc_power(const shg_params* p, const double* A, double n);

/* heat */
void orc_jacobi_sweep(const shg_params* p, int32_t nzh, const double* Ti,
                      double* Tf, const double* Q);
int orc_heat_solve_direct(const shg_params* p, int32_t nzh, const double* Q,
                          double* T);
int orc_tridiag_eigen(int32_t n, const double* diag, const double* off,
                      double* evals, double* evecs);

/* full run = Solver::run */
/* one THERMAL march on a given T: heat source per heat slice and the pump
 * and SH powers after every slice */
int orc_march_heat(const shg_params* p, const double* T, double* Q, double* power_after);
int orc_run(const shg_params* p, shg_result* r);

#ifdef __cplusplus
}
#endif
#endif
