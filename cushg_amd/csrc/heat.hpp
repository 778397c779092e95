// heat.hpp -- THERMAL: steady-state heat solve on the GPU and the optics
// inputs derived from the temperature (host orchestration).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/shg.h"

struct shg_handle;

namespace shg {

struct HeatState {
    int nx = 0, ny = 0, nzh = 0;
    double dx = 0, dy = 0, dz = 0;
    double* T = nullptr;       // FP64 temperature [nzh][ny][nx]
    double* Tn = nullptr;      // FP64 scratch
    float2* nn = nullptr;      // (n_p(T), n_s(T)) per voxel, read by the optics
    float* Q = nullptr;        // heat source [nzh][ny][nx]
    double* dk = nullptr;      // dk(T) [nzh][ny][nx]
    double* dkbase = nullptr;  // running-integral base per heat slice
    double* pol = nullptr;     // poling phases [nz] or null
    double* red = nullptr;     // reduction slots
    double* part = nullptr;    // per-block partials
    int nblocks = 0;
    void* solver = nullptr;    // direct-solver tables (opaque)
    double rho = 0.0;          // observed contraction of one defect-correction step
    double rho_max = 0.0;      // largest contraction seen on this state's solves (0: none yet)
    bool converged_in_one = false;
};

struct HeatSolveInfo {
    long long sweeps = 0;
    double residual = 0;
    bool converged_in_one = false;
    int refinements = 0;
};

size_t heat_bytes(const shg_params* p);
int heat_alloc(shg_handle* h, const shg_params* p, HeatState& hs);
int heat_init(shg_handle* h, const shg_params* p, HeatState& hs);
int heat_solve(shg_handle* h, const shg_params* p, HeatState& hs, HeatSolveInfo& info);
int heat_prepare_optics(shg_handle* h, const shg_params* p, HeatState& hs, double* Tc);
// Adjoint of the steady state (DIRECT solver tables): nu = A^-1 mu on the
// unknown nodes (0 on the oven nodes) by one FP64 fast-diagonalisation
// solve, w = dr/dQ nu (the heat-source cotangent, 0 where Q is not used) and
// *tp (device) = the derivative of <mu, T> with the oven temperature.  mu,
// nu, w: [nzh][ny][nx]; u64, v64: nxu * nyu * nzh doubles of work space.
int heat_adjoint_solve(shg_handle* h, const shg_params* p, HeatState& hs, const double* mu, double* nu, double* w,
                       double* u64, double* v64, double* tp);
int heat_finish(shg_handle* h, const shg_params* p, HeatState& hs, shg_result* r);
void heat_release(HeatState& hs);

}  // namespace shg
