// heat_kernels.hpp -- coefficient structs and launchers of kernels_heat.hip.
#pragma once

#include "common.hpp"

namespace shg {

// Constants of the Jacobi sweep (src/headers/Tfield.h:99-106) and the
// unknown block of the steady-state problem: x in [x0, x0 + nxu),
// y in [1, 1 + nyu), every z (the other nodes are Dirichlet oven nodes).
struct HeatCoef {
    int nx, ny, nz;
    int oven;      // SHG_HEAT_BC_OVEN_SURROUNDED
    int x0, nxu, nyu;
    double kth, Tinf;
    double idx2, idy2, idz2, L2, ikth;
    double Bix, Biy, Biz;
    double ax, ay, az, bx, by, bz;
    double Fx, Fy, Fz;  // 1/(1 + 2a + 2b + Bi) of the x, top and z faces
    double av, bv, cv, V, ah, bh, ch, H, ad, bd, cd, D;
    double AV, BV, CV, BiD, GV;
};

// 1D factors of the tensor-product finite-volume form
//   A = Kx (x) Wy (x) Wz + Wx (x) Ky (x) Wz + Wx (x) Wy (x) Kz
// (kx/ky/kz: diagonals of K; wx/wy/wz: control-volume widths; lx/ly/lz:
// eigenvalues of W^-1/2 K W^-1/2, even then odd ones on a folded axis),
// over the unknown block of HeatCoef.
struct HeatAxes {
    const double *kx, *ky, *kz;
    const double *wx, *wy, *wz;
    const double *lx, *ly, *lz;
};

// One eigenbasis transform on the matrix cores (kernels_heat_gemm.hip):
// C[i][j] = sum_k R(i, k) S(k, j) over k < n, C row-major with row stride
// ldc, `batch` independent blocks (strides su, sc of the data, sq of the
// split eigenbasis);
//   kind A (kindB = 0): R = Q_op [i][k] (i < n), S(k, j) = u[k ldu + j] (j < nbig)
//   kind B (kindB = 1): R(i, k) = u[i ldu + k] (i < nbig), S(k, j) = Q_op [j][k] (j < n)
// qh, ql: the eigenbasis in the direction asked for, split into FP16 hi and
// lo parts of 2^qe q, [row][k] with rows padded to a multiple of 128 and k
// (row stride qpad) to a multiple of 32; umax: max|u| as float bits; cmax
// (optional): receives max|C| as float bits (cleared before), the
// per-workgroup maxima staged in cpart.
struct HeatGemm {
    const float* u;
    float* c;
    const _Float16* qh;
    const _Float16* ql;
    const unsigned* umax;
    unsigned* cmax;
    unsigned* cpart;
    long ldu, ldc, su, sc, sq, nbig;
    int n, qpad, qe, kindB;
};
hipError_t launch_heat_gemm(hipStream_t st, const HeatGemm& g, int batch);
// the split operand of launch_heat_gemm from the FP64 eigenvectors q (nb
// blocks of n x n, component j of vector i at q[j n + i]): Q_op[m][k] =
// q[k n + m] forward, q[m n + k] inverse, times 2^qe; nb blocks of rpad x
// kpad halves each in hi and lo.
hipError_t launch_split_basis(hipStream_t st, const double* q, int n, int nb, bool inv, int qe, _Float16* hi,
                              _Float16* lo, int rpad, int kpad);
// FP64 transform C[i][j] = sum_k r[i ldr + k] s[k lds + j] (i < I, j < J,
// k < n), `batch` blocks with strides sr, ss, sc.
struct HeatGemm64 {
    const double* r;
    const double* s;
    double* c;
    long ldr, lds, ldc, sr, ss, sc, I, J;
    int n;
};
hipError_t launch_heat_gemm64(hipStream_t st, const HeatGemm64& g, int batch);
// max |a| over n floats into *out (float bits; *out 0 before), per-workgroup
// maxima staged in part (1024 entries)
hipError_t launch_absmax(hipStream_t st, const float* a, long n, unsigned* part, unsigned* out);
hipError_t launch_zero_u32(hipStream_t st, unsigned* p, int n);

HeatCoef heat_coef(const shg_params* p, int nzh);
int heat_grid_blocks(const HeatCoef& c);

hipError_t launch_heat_sweep(hipStream_t st, const HeatCoef& c, const double* Ti, double* Tf, const float* Q,
                             double* part_sq, bool want_norm);
template <typename R>
hipError_t launch_heat_resid(hipStream_t st, const HeatCoef& c, const HeatAxes& a, const double* T, const float* Q,
                             R* u, double* part_sq);
// part, slot (optional, FP32 only): max |v| after the division, as float
// bits, into *slot (per-workgroup maxima staged in part, at least
// heat_max_partials entries; *slot must hold 0 or a smaller value before)
template <typename R>
hipError_t launch_fd_divide(hipStream_t st, R* v, const HeatAxes& a, int nxu, int nyu, int nz,
                            unsigned* part = nullptr, unsigned* slot = nullptr);
// even/odd fold of the x and/or z axis of u (in -> out; not in place);
// part, slot as for launch_fd_divide, of the output
template <typename R>
hipError_t launch_fd_fold(hipStream_t st, const R* in, R* out, int nxu, int nyu, int nz, bool fx, bool fz,
                          unsigned* part = nullptr, unsigned* slot = nullptr);
// entries of the part buffer that the max reductions of the heat solve's
// FP32 transforms need at most on an nxu x nyu x nz unknown block
long heat_max_partials(int nxu, int nyu, int nz);
long heat_gemm_max_blocks(int nxu, int nyu, int nz);
// *slot = max(*slot, max of part[0..n))
hipError_t launch_max_parts(hipStream_t st, const unsigned* part, long n, unsigned* slot);
// T += W^-1/2 u, u folded along x / z as launch_fd_fold leaves it; writes
// *nblocks partial maxima (at most heat_grid_blocks(c))
template <typename R>
hipError_t launch_fd_update(hipStream_t st, const HeatCoef& c, const HeatAxes& a, const R* u, double* T,
                            double* part_max, bool fx, bool fz, int* nblocks);
hipError_t launch_reduce(hipStream_t st, const double* part, int n, bool is_max, double* out);
hipError_t launch_heat_init(hipStream_t st, double* T, float* Q, int nx, int ny, int nz, double Tc, double Tpel,
                            bool oven);
// n(T), dk(T) and the running-integral base per voxel, one pass over T
hipError_t launch_heat_dk(hipStream_t st, const double* T, double* dk, float2* nn, double* base, int n2, int nzh,
                          int nz, int crystal, double lp, double Lambda, double ea, double eb);
hipError_t launch_max(hipStream_t st, const double* a, long n, double* part, int nblocks);
// adjoint of the steady state: u = W^-1/2 mu on the unknown block; after
// nu = A^-1 mu, the heat-source cotangent w and the partial sums of the oven
// temperature's derivative (at most heat_grid_blocks(c)); dL/dT per voxel
// from the adjoint row passes' accumulators (into gnp), plus tc at the
// centre voxel of slab 0
template <typename R>
hipError_t launch_adj_heat_in(hipStream_t st, const HeatCoef& c, const HeatAxes& a, const double* mu, R* u);
hipError_t launch_adj_heat_out(hipStream_t st, const HeatCoef& c, const HeatAxes& a, const double* mu,
                               const double* nu, double* w, double* part);
hipError_t launch_adj_heat_grad(hipStream_t st, const double* T, double* gnp, const double* gns, const double* gA,
                                const double* gC, int nx, int ny, int nzh, int nz, int crystal, double lp,
                                double Lambda, double ea, double eb, double tc);

}  // namespace shg
