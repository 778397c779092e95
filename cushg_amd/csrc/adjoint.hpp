// adjoint.hpp -- launch interface of the adjoint passes (kernels_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace shg {

// Parameters of one adjoint row / column launch (single instance).
struct AdjArgs {
    const MixedInst* I;     // the instance descriptor (constants, phase table, propagator factors)
    void* lamP;             // adjoint state, pump / SH ([ny][nx], storage type)
    void* lamS;
    const void* fwdP;       // forward state read by this pass (storage type)
    const void* fwdS;
    double* part;           // per-workgroup partial sums of this launch
    int s;                  // slice (row pass: boundary s; column pass: slice s)
    int first;              // row pass: apply R1_s^T (the adjoint state is in (kx, y) and is IFx'd first)
    int second;             // row pass: apply R2_{s-1}^T and Fx the result
    int fwd_spatial;        // row pass: fwd is the spatial initial state X_0 (s == 0)
    double dfx, dfy;        // frequency steps 1/(n d)
};

// THERMAL adjoint row pass: the heat-source cotangent and the per-voxel
// accumulators, indexed like the heat volume [nzh][ny][nx]:
//   gnp, gns  dL/dn_p, dL/dn_s of the voxel (couplings, two-photon term and
//             heat source through n(T));
//   gA, gC    sum over the slab's slices s of c_s (s - s0(h) + 1) and of
//             c_s, with c_s = dL/dDKint_s, the gradient of the running dk
//             integral DKint_s = sum_{j<=s} dk[h(j)] (s0(h): the slab's first
//             optical slice), from which dL/ddk[h] follows by a suffix sum.
struct AdjThermal {
    const double* w;  // dL/dQ per voxel (or null: the loss has no heat-source term)
    double *gnp, *gns, *gA, *gC;
    double seed;      // at boundary nz - 1: lam_S += seed S (the gradient of eta at the saved state)
};

// THERMAL adjoint row pass at boundary s (grid and workgroups as
// launch_adj_row): 2 partial sums per workgroup, dL/dpoling of R1_s and of
// R2_{s-1}
hipError_t launch_adj_row_th(bool f64, int nx, int ny, hipStream_t st, const AdjArgs& a, const AdjThermal& t,
                             const double2* tw);

// vector kernels of the THERMAL adjoint's accelerated fixed-point iteration
// (n doubles): out = sum_j c[j] x[j] over nx inputs (nx <= 6; out may be
// one of the inputs), and the partial dot products <a, b> per workgroup
// (adj_dot_workgroups() of them, summed by launch_reduce)
hipError_t launch_adj_lincomb(hipStream_t st, double* out, long n, int nx, const double* c, const double* const* x);
hipError_t launch_adj_dot(hipStream_t st, const double* a, const double* b, long n, double* part);
int adj_dot_workgroups();

// adjoint row pass: 12 partial sums per workgroup (R1_s: dL/dphi_0..2,
// dL/dkp, dL/dks, dL/dbfac; then the same for R2_{s-1})
hipError_t launch_adj_row(bool f64, int nx, int ny, hipStream_t st, const AdjArgs& a, const double2* tw);
int adj_row_workgroups(int nx, int ny);
// adjoint column pass: 2 partial sums per workgroup (pump, SH)
hipError_t launch_adj_col(bool f64, int nx, int ny, hipStream_t st, const AdjArgs& a, const double2* tw);
int adj_col_workgroups(int nx);
hipError_t launch_adj_seed(bool f64, void* lp, void* ls, const void* sout, long n, double coef, hipStream_t st);
// initial-field gradient: 3 partial sums per workgroup (waist, focus, n_p)
hipError_t launch_adj_pump(bool f64, const void* lam, int nx, int ny, double dx, double dy, double w0, double f,
                           double n, double zR, double Ap0, double* part, hipStream_t st);
int adj_pump_workgroups();
// out[r][v] = sum over nb workgroup partials of row r
hipError_t launch_adj_reduce(const double* part, int rows, int nb, int nv, double* out, hipStream_t st);

}  // namespace shg
