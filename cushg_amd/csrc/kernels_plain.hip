// kernels_plain.hip -- the plain FP64 form: the reference's op sequence
// (src/headers/Solver.h:17-264) as separate FP64 kernels, FFTs included.
#include "fft.hpp"
#include "plain_kernels.hpp"

namespace shg {

// dA/dz (src/headers/Solver.h:33-80) for one RK4 stage, FP64.  Non-THERMAL:
// n at the scalar T_C, phase (z_s + inc dz) dk; THERMAL: n(T[x,y,h]) and
// phase DKint (s + inc) dz/(s + 1) with DKint = dkbase[h] + s dk[h].
template <bool THERMAL>
__global__ void k_dadz(PlainStage a, const double2* __restrict__ P, const double2* __restrict__ S,
                       double2* __restrict__ dP, double2* __restrict__ dS) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n2) return;
    double kp = a.kp, ks = a.ks, bfac = a.bfac, ph;
    if (THERMAL) {
        const size_t v = (size_t)a.h * a.n2 + i;
        const double T = a.T[v];
        const double np = refr_index<double>(a.crystal, a.lp, T);
        const double ns = refr_index<double>(a.crystal, a.ls, T);
        kp = a.kp0 / np;
        ks = a.ks0 / ns;
        bfac = a.bf0 * ns;
        const double dki = a.dkbase[v] + (double)a.step * a.dk[v];
        ph = dki * ((double)a.step + a.inc) * a.dz / ((double)a.step + 1.0) + a.pol;
    } else {
        ph = ((double)a.step + a.inc) * a.dz * a.dk0 + a.pol;
    }
    double sn, cs;
    sincos(ph, &sn, &cs);
    const double2 e = make_double2(cs, sn);
    const double2 p = P[i], s = S[i];
    const double2 spc = cmul(cmul(s, conjc(p)), conjc(e));
    dP[i] = make_double2(-kp * spc.y - 0.5 * a.alpha_p * p.x, kp * spc.x - 0.5 * a.alpha_p * p.y);
    const double2 ppe = cmul(cmul(p, p), e);
    const double g = 0.5 * (a.alpha_s + bfac * abs2(s));
    dS[i] = make_double2(-ks * ppe.y - g * s.x, ks * ppe.x - g * s.y);
}

// aux = A + c k  (linealCombination, src/headers/Solver.h:100-113)
__global__ void k_lincomb(int n2, double c, const double2* __restrict__ P, const double2* __restrict__ S,
                          const double2* __restrict__ kP, const double2* __restrict__ kS, double2* __restrict__ aP,
                          double2* __restrict__ aS) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n2) return;
    aP[i] = P[i] + c * kP[i];
    aS[i] = S[i] + c * kS[i];
}

// A += (k1 + 2k2 + 2k3 + k4) w  (kernelRK4, src/headers/Solver.h:83-97)
__global__ void k_rk4_final(int n2, double w, double2* __restrict__ P, double2* __restrict__ S,
                            const double2* __restrict__ k1p, const double2* __restrict__ k1s,
                            const double2* __restrict__ k2p, const double2* __restrict__ k2s,
                            const double2* __restrict__ k3p, const double2* __restrict__ k3s,
                            const double2* __restrict__ k4p, const double2* __restrict__ k4s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n2) return;
    P[i] = P[i] + w * (k1p[i] + 2.0 * k2p[i] + 2.0 * k3p[i] + k4p[i]);
    S[i] = S[i] + w * (k1s[i] + 2.0 * k2s[i] + 2.0 * k3s[i] + k4s[i]);
}

// C = A . B (kernelComplexProduct, src/headers/Solver.h:17-30), in place on A
__global__ void k_cmul_inplace(int n2, double2* __restrict__ A, const double2* __restrict__ B) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n2) A[i] = cmul(A[i], B[i]);
}

// A /= nx ny (cuFFTscale, src/headers/Efields.h:15-29)
__global__ void k_scale(int n2, double s, double2* __restrict__ A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n2) A[i] = s * A[i];
}

// Q of slice s (kernelSetQ, src/headers/Tfield.h:260-276); mean over the
// optical slices of a heat slice
__global__ void k_setq(int n2, PlainStage a, const double2* __restrict__ P, const double2* __restrict__ S,
                       float* __restrict__ Q, int first, double wgt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n2) return;
    const size_t v = (size_t)a.h * n2 + i;
    const double T = a.T[v];
    const double np = refr_index<double>(a.crystal, a.lp, T);
    const double ns = refr_index<double>(a.crystal, a.ls, T);
    const double Ip = a.hce * np * abs2(P[i]);
    const double Is = a.hce * ns * abs2(S[i]);
    const double q = (a.alpha_p * Ip + a.alpha_s * Is + a.beta_s * Is * Is) * wgt;
    Q[v] = first ? (float)q : (float)((double)Q[v] + q);
}

// in-place FFT of every row (length NX), FP64
template <int NX, bool INV>
__global__ __launch_bounds__(NX / 8 * 2) void k_fft_rows(double2* __restrict__ A, int ny,
                                                        const double2* __restrict__ tw) {
    constexpr int TPR = NX / 8;
    constexpr int LD = FftShape<NX>::LDS;
    __shared__ double2 lds[2 * LD];
    const int y = blockIdx.x * 2 + threadIdx.y;
    if (blockIdx.x * 2 >= ny) return;
    const int tp = threadIdx.x;
    double2 v[8];
    const size_t base = (size_t)y * NX + tp;
#pragma unroll
    for (int q = 0; q < 8; q++) v[q] = A[base + q * TPR];
    fft_dist<double2, NX, INV>(v, lds + threadIdx.y * LD, tw, tp);
#pragma unroll
    for (int q = 0; q < 8; q++) A[base + q * TPR] = v[q];
}

// in-place FFT of every column (length NY), FP64, 4 columns per workgroup
template <int NY, bool INV>
__global__ __launch_bounds__(NY / 8 * 4) void k_fft_cols(double2* __restrict__ A, int nx,
                                                        const double2* __restrict__ tw) {
    constexpr int TPC = NY / 8;
    constexpr int LD = FftShape<NY>::LDS;
    __shared__ double2 lds[4 * LD];
    const int x0 = blockIdx.x * 4;
    if (x0 >= nx) return;
    const int c = threadIdx.x, tp = threadIdx.y, x = x0 + c;
    double2 v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) v[q] = A[(size_t)(tp + q * TPC) * nx + x];
    fft_dist<double2, NY, INV>(v, lds + c * LD, tw, tp);
#pragma unroll
    for (int q = 0; q < 8; q++) A[(size_t)(tp + q * TPC) * nx + x] = v[q];
}

// ------------------------------------------------------------------
static inline dim3 g1(int n) { return dim3((n + 255) / 256); }

hipError_t launch_dadz(hipStream_t st, bool thermal, const PlainStage& a, const double2* P, const double2* S,
                       double2* dP, double2* dS) {
    if (thermal)
        hipLaunchKernelGGL(k_dadz<true>, g1(a.n2), dim3(256), 0, st, a, P, S, dP, dS);
    else
        hipLaunchKernelGGL(k_dadz<false>, g1(a.n2), dim3(256), 0, st, a, P, S, dP, dS);
    return hipGetLastError();
}
hipError_t launch_lincomb(hipStream_t st, int n2, double c, const double2* P, const double2* S, const double2* kP,
                          const double2* kS, double2* aP, double2* aS) {
    hipLaunchKernelGGL(k_lincomb, g1(n2), dim3(256), 0, st, n2, c, P, S, kP, kS, aP, aS);
    return hipGetLastError();
}
hipError_t launch_rk4_final(hipStream_t st, int n2, double w, double2* P, double2* S, double2* const* k) {
    hipLaunchKernelGGL(k_rk4_final, g1(n2), dim3(256), 0, st, n2, w, P, S, k[0], k[1], k[2], k[3], k[4], k[5], k[6],
                       k[7]);
    return hipGetLastError();
}
hipError_t launch_cmul_inplace(hipStream_t st, int n2, double2* A, const double2* B) {
    hipLaunchKernelGGL(k_cmul_inplace, g1(n2), dim3(256), 0, st, n2, A, B);
    return hipGetLastError();
}
hipError_t launch_scale(hipStream_t st, int n2, double s, double2* A) {
    hipLaunchKernelGGL(k_scale, g1(n2), dim3(256), 0, st, n2, s, A);
    return hipGetLastError();
}
hipError_t launch_setq(hipStream_t st, int n2, const PlainStage& a, const double2* P, const double2* S, float* Q,
                       bool first, double wgt) {
    hipLaunchKernelGGL(k_setq, g1(n2), dim3(256), 0, st, n2, a, P, S, Q, first ? 1 : 0, wgt);
    return hipGetLastError();
}

template <bool INV>
static hipError_t fft_rows_inv(hipStream_t st, double2* A, int nx, int ny, const double2* tw) {
    dim3 grid((ny + 1) / 2);
#define RC(N)                                                                                     \
    case N:                                                                                       \
        hipLaunchKernelGGL((k_fft_rows<N, INV>), grid, dim3(N / 8, 2), 0, st, A, ny, tw); \
        break;
    switch (nx) {
        RC(64) RC(128) RC(192) RC(256) RC(384) RC(512)
        default: return hipErrorInvalidValue;
    }
#undef RC
    return hipGetLastError();
}
template <bool INV>
static hipError_t fft_cols_inv(hipStream_t st, double2* A, int nx, int ny, const double2* tw) {
    dim3 grid((nx + 3) / 4);
#define CC(N)                                                                                     \
    case N:                                                                                       \
        hipLaunchKernelGGL((k_fft_cols<N, INV>), grid, dim3(4, N / 8), 0, st, A, nx, tw); \
        break;
    switch (ny) {
        CC(64) CC(128) CC(192) CC(256) CC(384) CC(512)
        default: return hipErrorInvalidValue;
    }
#undef CC
    return hipGetLastError();
}

hipError_t launch_fft_rows64(hipStream_t st, double2* A, int nx, int ny, const double2* tw, bool inverse) {
    return inverse ? fft_rows_inv<true>(st, A, nx, ny, tw) : fft_rows_inv<false>(st, A, nx, ny, tw);
}
hipError_t launch_fft_cols64(hipStream_t st, double2* A, int nx, int ny, const double2* tw, bool inverse) {
    return inverse ? fft_cols_inv<true>(st, A, nx, ny, tw) : fft_cols_inv<false>(st, A, nx, ny, tw);
}

}  // namespace shg
