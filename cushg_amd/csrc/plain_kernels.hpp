// plain_kernels.hpp -- launchers of the plain FP64 form (kernels_plain.hip).
#pragma once

#include "common.hpp"

namespace shg {

// scalars of one dA/dz stage
struct PlainStage {
    int n2, step, h, crystal;
    double inc;                // z offset of the stage in units of dz
    double dz, dk0, pol;       // non-THERMAL: phase (step + inc) dz dk0 + pol
    double kp, ks, bfac;       // non-THERMAL couplings
    double kp0, ks0, bf0;      // THERMAL couplings / n
    double alpha_p, alpha_s, beta_s, hce, lp, ls;
    const double* T;           // THERMAL: FP64 temperature [nzh][ny][nx]
    const double* dkbase;
    const double* dk;
};

hipError_t launch_dadz(hipStream_t st, bool thermal, const PlainStage& a, const double2* P, const double2* S,
                       double2* dP, double2* dS);
hipError_t launch_lincomb(hipStream_t st, int n2, double c, const double2* P, const double2* S, const double2* kP,
                          const double2* kS, double2* aP, double2* aS);
hipError_t launch_rk4_final(hipStream_t st, int n2, double w, double2* P, double2* S, double2* const* k);
hipError_t launch_cmul_inplace(hipStream_t st, int n2, double2* A, const double2* B);
hipError_t launch_scale(hipStream_t st, int n2, double s, double2* A);
hipError_t launch_setq(hipStream_t st, int n2, const PlainStage& a, const double2* P, const double2* S, float* Q,
                       bool first, double wgt);
hipError_t launch_fft_rows64(hipStream_t st, double2* A, int nx, int ny, const double2* tw, bool inverse);
hipError_t launch_fft_cols64(hipStream_t st, double2* A, int nx, int ny, const double2* tw, bool inverse);

}  // namespace shg
