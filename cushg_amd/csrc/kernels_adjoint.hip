// kernels_adjoint.hip -- reverse-mode (adjoint) passes of the fused
// non-THERMAL march, FP64 arithmetic, fields stored as S (float2 or
// double2).  No reference counterpart: the reference computes no gradients.
//
// Forward slice s on the spatial state X_s (notation of the fused passes):
//   U_s = R1_s(X_s);  V_s = Fx U_s;  C_s = IFy(Pi . Fy V_s);
//   Y_s = IFx C_s;    X_{s+1} = R2_s(Y_s)
// (R: RK4 calls, F/IF: unnormalised forward/inverse DFT along x or y, Pi:
// propagator with 1/(nx ny)).  For a real scalar L and the real inner
// product Re<a, b>, the gradients (Wirtinger form, grad = dL/dRe + i dL/dIm)
// run backwards through
//   lam_Y = R2_s^T lam_X(s+1);  lam_C = Fx lam_Y;
//   lam_V = IFy(conj(Pi) . Fy lam_C);  lam_U = IFx lam_V;  lam_X = R1_s^T lam_U
// (IF^T = F, F^T = IF for the unnormalised transforms), grouped like the
// forward passes into an adjoint row pass (IFx -> R1_s^T -> R2_{s-1}^T -> Fx)
// and an adjoint column pass (Fy -> conj(Pi) -> IFy).  The row pass rebuilds
// Y_{s-1} = IFx C_{s-1} and X_s = R2_{s-1}(Y_{s-1}) from the stored
// column-pass output C_{s-1}.
#include "adjoint.hpp"
#include "adjoint_math.hpp"
#include "mixed_passes.hpp"

namespace shg {

// block sum of NV values per thread into part[blockIdx.x * NV + v]
template <int NV>
__device__ __forceinline__ void block_sums(double (&v)[NV], double* part, double* red) {
    const int t = threadIdx.x + threadIdx.y * blockDim.x;
    const int nt = blockDim.x * blockDim.y;
    const int nw = (nt + 63) / 64;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double a = v[k];
        for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
        if ((t & 63) == 0) red[(t >> 6) * NV + k] = a;
    }
    __syncthreads();
    if (t < NV) {
        double a = 0.0;
        for (int w = 0; w < nw; w++) a += red[w * NV + t];
        part[(size_t)(blockIdx.x + gridDim.x * blockIdx.y) * NV + t] = a;
    }
}

// Adjoint row pass at boundary s: ROWS rows per workgroup of NX * ROWS / 2
// threads, grid (ny / ROWS).  The transforms run on groups of NX/8 threads
// holding 8 points each (group g: row g / 4, field g % 4 = lam_P, lam_S,
// Y_P, Y_S: every thread is in one); the pointwise reverse RK4 work, which
// dominates, runs two points per thread from LDS, so that a single
// 256 x 256 instance still spreads over every SIMD.
// Partial sums per workgroup (12):
//   [0..5]  call R1_s:     dL/dphi_0..2, dL/dkp, dL/dks, dL/dbfac
//   [6..11] call R2_{s-1}: the same
template <int NX, int ROWS, typename S>
__global__ __launch_bounds__(NX * ROWS / 2) void k_adj_row(AdjArgs a, const double2* __restrict__ tw) {
    using C = double2;
    constexpr int TPR = NX / 8;
    constexpr int NT = NX * ROWS / 2;  // threads: the 4 ROWS transform groups, two points each in the pointwise work
    constexpr int LD = FftShape<NX>::LDS;
    __shared__ C lds[4 * ROWS * LD];
    __shared__ double red[16 * 12];
    const MixedInst& I = *a.I;
    const int t = threadIdx.x;
    const int gi = t / TPR, tp = t % TPR;
    const S* fP = static_cast<const S*>(a.fwdP);
    const S* fS = static_cast<const S*>(a.fwdS);
    S* LP = static_cast<S*>(a.lamP);
    S* LS = static_cast<S*>(a.lamS);
    const size_t row0 = (size_t)blockIdx.x * ROWS * NX;
    // 1. the four rows per image row into LDS, spatial
    {
        const bool act = true;  // NT = 4 ROWS TPR: every thread is in a transform group
        const int f = gi & 3;
        const S* src = f == 0 ? LP : f == 1 ? LS : f == 2 ? fP : fS;
        const size_t base = row0 + (size_t)(gi >> 2) * NX + tp;
        C v[8];
        if (act) {
#pragma unroll
            for (int q = 0; q < 8; q++) v[q] = from_store<C>(src[base + q * TPR]);
        }
        C* buf = lds + (act ? gi : 0) * LD;
        if (a.first || !a.fwd_spatial)
            fft_pow<C, NX, true, true>(v, buf, tw, tp, act && (f < 2 ? a.first != 0 : !a.fwd_spatial));
        if (act) {
#pragma unroll
            for (int q = 0; q < 8; q++) buf[fpad(tp + q * TPR)] = v[q];
        }
    }
    __syncthreads();
    // 2. the pointwise reverse work, two points per thread
    const RkConst<double> k{I.c1, I.c2, I.c3, I.w, I.ap_h, I.as_h};
    const int s = a.s;
    C e1[3], e2[3];
    {
        const double2* t1 = I.phase + (size_t)(s * 2 + 0) * 3;            // R1_s
        const double2* t2 = I.phase + (size_t)(max(s - 1, 0) * 2 + 1) * 3;  // R2_{s-1}
#pragma unroll
        for (int j = 0; j < 3; j++) {
            e1[j] = t1[j];
            e2[j] = t2[j];
        }
    }
    double acc1[6] = {0, 0, 0, 0, 0, 0}, acc2[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 1
    for (int pt = t; pt < NX * ROWS; pt += NT) {
        C* rb = lds + (pt / NX) * 4 * LD;
        const int xi = fpad(pt % NX);
        const C Yp = rb[2 * LD + xi], Ys = rb[3 * LD + xi];
        C Xp = Yp, Xs = Ys;
        if (!a.fwd_spatial) rk4_point(Xp, Xs, e2, I.kp, I.ks, I.bfac, k);  // X_s = R2_{s-1}(Y_{s-1})
        C lp = rb[xi], ls = rb[LD + xi];
        if (a.first) rk4_adj(lp, ls, Xp, Xs, e1, I.kp, I.ks, I.bfac, k, acc1);
        if (a.second) rk4_adj(lp, ls, Yp, Ys, e2, I.kp, I.ks, I.bfac, k, acc2);
        rb[xi] = lp;
        rb[LD + xi] = ls;
    }
    __syncthreads();
    // 3. lam_P, lam_S back to row-spectral form and out
    {
        const bool act = gi < 2 * ROWS;
        const int r = gi >> 1, f = gi & 1;
        C* buf = lds + (act ? r * 4 + f : 0) * LD;
        C v[8];
        if (act) {
#pragma unroll
            for (int q = 0; q < 8; q++) v[q] = buf[fpad(tp + q * TPR)];
        }
        if (a.second) {
            __syncthreads();  // every group has read its buffer before the transform overwrites it
            fft_pow<C, NX, false, false>(v, buf, tw, tp, act);
        }
        if (act) {
            S* dst = f == 0 ? LP : LS;
            const uint32_t key = store_key(4, f, s);
            const size_t base = row0 + (size_t)r * NX + tp;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const size_t i = base + q * TPR;
                dst[i] = to_store<S>(v[q], i, key);
            }
        }
    }
    double v[12];
#pragma unroll
    for (int j = 0; j < 6; j++) {
        v[j] = acc1[j];
        v[6 + j] = acc2[j];
    }
    block_sums<12>(v, a.part, red);
}

// Adjoint column pass of slice s: COLS columns of one field per workgroup
// (blockIdx.y: 0 pump, 1 SH); grid (nx/COLS, 2).  Partial sums per
// workgroup, [pump, SH], of  sum_m Re(conj(Fy lam_C)_m i (Fy C_s)_m)
// (fx_m^2 + fy_m^2): the propagator-wavenumber gradient up to the factor
// dz 2 pi^2 / (kw^2 ny).
template <int NY, int COLS, typename S>
__global__ __launch_bounds__(NY / 8 * COLS) void k_adj_col(AdjArgs a, const double2* __restrict__ tw) {
    using C = double2;
    constexpr int TPC = NY / 8;
    constexpr int LD = FftShape<NY>::LDS;
    __shared__ C lds[COLS * LD];
    __shared__ double red[16 * 2];
    const MixedInst& I = *a.I;
    const int nx = I.nx;
    const int c = threadIdx.x, tp = threadIdx.y;
    // the K column groups sharing a 128-byte row segment on one XCD (ids 8
    // apart share an XCD and its L2), as in the forward column pass
    constexpr int K = COLS * 8 >= 128 ? 1 : 128 / (COLS * 8);
    int bx = blockIdx.x;
    if (K > 1 && gridDim.x % (8 * K) == 0) {
        const int lo = bx & 7, hi = (bx >> 3) % K, chunk = bx / (8 * K);
        bx = chunk * 8 * K + lo * K + hi;
    }
    const int x = bx * COLS + c;
    const bool sh = blockIdx.y != 0;
    S* L = static_cast<S*>(sh ? a.lamS : a.lamP);
    const S* F = static_cast<const S*>(sh ? a.fwdS : a.fwdP);
    const double2* ex = sh ? I.exS : I.exP;
    const double2* ey = sh ? I.eyS : I.eyP;
    C* my = lds + c * LD;
    C v[8], w[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        v[q] = from_store<C>(L[(size_t)(tp + q * TPC) * nx + x]);
        w[q] = from_store<C>(F[(size_t)(tp + q * TPC) * nx + x]);
    }
    // Fy w, then Fy v -> gradient term and * conj(Prop) -> IFy
    fft_pow<C, NY, false, true>(w, my, tw, tp);  // (its end barrier: v's transforms write next)
    const int ic = (x + nx / 2) % nx;
    const double fx = (ic - 0.5 * nx) * a.dfx;
    const double exc = ex[x].x, exs = -ex[x].y;  // conj(ex)
    double gk = 0.0;
    fft_fwd_op_inv<C, NY>(v, my, tw, tp, [&](C (&b)[8]) {
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int ky = tp + q * TPC;
            const int jc = (ky + NY / 2) % NY;
            const double fy = (jc - 0.5 * NY) * a.dfy;
            // Re(conj(v) i w) = v.x * (-w.y) + v.y * w.x
            gk += (-b[q].x * w[q].y + b[q].y * w[q].x) * (fx * fx + fy * fy);
            const double2 pr = cmul(mk<double2>(exc, exs), conjc(ey[ky]));
            b[q] = cmul(b[q], pr);
        }
    });
    const uint32_t key = store_key(5, sh ? 1 : 0, a.s);
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const size_t i = (size_t)(tp + q * TPC) * nx + x;
        L[i] = to_store<S>(v[q], i, key);
    }
    double vv[2] = {sh ? 0.0 : gk, sh ? gk : 0.0};
    block_sums<2>(vv, a.part, red);
}

// Seed: lam_P = 0, lam_S = coef * S_out (the gradient of
// eta = c_s sum |S|^2 / P with respect to the saved SH field).
template <typename S>
__global__ void k_adj_seed(S* __restrict__ lp, S* __restrict__ ls, const S* __restrict__ sout, long n, double coef) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)blockDim.x * gridDim.x) {
        const S v = sout[i];
        S z;
        z.x = 0;
        z.y = 0;
        lp[i] = z;
        S o;
        o.x = coef * v.x;
        o.y = coef * v.y;
        ls[i] = o;
    }
}

// Initial-field gradients: with lam the gradient at the pump's initial
// field A0 (src/headers/Efields.h:84-119),  sum Re(conj(lam) A0 D)  for
// D = d ln A0 / d waist, d ln A0 / d focus, d ln A0 / d n_p, where
//   A0 = (Ap0 / M) exp(-r^2 / (w0^2 M)),  M = 1 - i f / zR,
//   zR = pi n w0^2 / l,  Ap0 ~ (n w0^2)^(-1/2).
// grid-stride over rows, block (256); partials [block][3].
template <typename S>
__global__ __launch_bounds__(256) void k_adj_pump(const S* __restrict__ lam, int nx, int ny, double dx, double dy,
                                                  double w0, double f, double n, double zR, double Ap0,
                                                  double* __restrict__ part) {
    __shared__ double red[4 * 3];
    const double2 M = make_double2(1.0, -f / zR);
    const double2 Mi = make_double2(M.x / abs2(M), -M.y / abs2(M));  // 1/M
    const double2 Mi2 = cmul(Mi, Mi);
    const double w2 = w0 * w0;
    // dM/dw0 = -2 (M - 1) / w0 ; dM/df = -i / zR ; dM/dn = -(M - 1) / n
    const double2 Mm1 = make_double2(M.x - 1.0, M.y);
    const double2 dMw = make_double2(-2.0 * Mm1.x / w0, -2.0 * Mm1.y / w0);
    const double2 dMf = make_double2(0.0, -1.0 / zR);
    const double2 dMn = make_double2(-Mm1.x / n, -Mm1.y / n);
    double acc[3] = {0, 0, 0};
    const long N = (long)nx * ny;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < N; i += 256L * gridDim.x) {
        const int x = (int)(i % nx), y = (int)(i / nx);
        const double rx = (x - 0.5 * nx) * dx, ry = (y - 0.5 * ny) * dy;
        const double r2 = rx * rx + ry * ry;
        // A0 = Ap0 Mi exp(-r2 Mi / w2)
        const double2 ex = make_double2(-r2 * Mi.x / w2, -r2 * Mi.y / w2);
        const double m = exp(ex.x);
        const double2 A0 = cmul(make_double2(Ap0 * Mi.x, Ap0 * Mi.y), make_double2(m * cos(ex.y), m * sin(ex.y)));
        // d ln A0 = d ln Ap0 - dM / M + r2 [ d(w2) / (w2^2 M) + dM / (w2 M^2) ]
        const double2 dMw_M = cmul(dMw, Mi), dMf_M = cmul(dMf, Mi), dMn_M = cmul(dMn, Mi);
        const double2 dMw_M2 = cmul(dMw, Mi2), dMf_M2 = cmul(dMf, Mi2), dMn_M2 = cmul(dMn, Mi2);
        const double2 Dw = make_double2(-1.0 / w0 - dMw_M.x + r2 * (2.0 * w0 * Mi.x / (w2 * w2) + dMw_M2.x / w2),
                                        -dMw_M.y + r2 * (2.0 * w0 * Mi.y / (w2 * w2) + dMw_M2.y / w2));
        const double2 Df = make_double2(-dMf_M.x + r2 * dMf_M2.x / w2, -dMf_M.y + r2 * dMf_M2.y / w2);
        const double2 Dn = make_double2(-0.5 / n - dMn_M.x + r2 * dMn_M2.x / w2, -dMn_M.y + r2 * dMn_M2.y / w2);
        const S lv = lam[i];
        const double2 l = make_double2(lv.x, lv.y);
        const double2 aw = cmul(A0, Dw), af = cmul(A0, Df), an = cmul(A0, Dn);
        acc[0] += l.x * aw.x + l.y * aw.y;
        acc[1] += l.x * af.x + l.y * af.y;
        acc[2] += l.x * an.x + l.y * an.y;
    }
    block_sums<3>(acc, part, red);
}

// out[r * nv + v] = sum_b part[(r * nb + b) * nv + v] for every row r
// (one workgroup per row; deterministic order)
__global__ __launch_bounds__(256) void k_adj_reduce(const double* __restrict__ part, int nb, int nv,
                                                    double* __restrict__ out) {
    __shared__ double red[256];
    const double* pr = part + (size_t)blockIdx.x * nb * nv;
    for (int v = 0; v < nv; v++) {
        double a = 0.0;
        for (int b = threadIdx.x; b < nb; b += 256) a += pr[(size_t)b * nv + v];
        red[threadIdx.x] = a;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[(size_t)blockIdx.x * nv + v] = red[0];
        __syncthreads();
    }
}

// ------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------
template <int NX> constexpr int adj_rows() { return NX >= 256 ? 1 : 256 / NX; }  // ~256 threads per workgroup

template <typename S>
static hipError_t adj_row_n(int nx, dim3 grid, hipStream_t st, const AdjArgs& a, const double2* tw) {
    switch (nx) {
#define SHG_ADJROW(N)                                                                                 \
    case N:                                                                                           \
        hipLaunchKernelGGL((k_adj_row<N, adj_rows<N>(), S>), grid, dim3(N * adj_rows<N>() / 2), 0, st, a, tw);  \
        return hipGetLastError();
        SHG_ADJROW(64)
        SHG_ADJROW(128)
        SHG_ADJROW(192)
        SHG_ADJROW(256)
        SHG_ADJROW(384)
        SHG_ADJROW(512)
#undef SHG_ADJROW
    }
    return hipErrorInvalidValue;
}

int adj_row_workgroups(int nx, int ny) {
    int rows = 1;
    switch (nx) {
        case 64: rows = adj_rows<64>(); break;
        case 128: rows = adj_rows<128>(); break;
        case 192: rows = adj_rows<192>(); break;
        case 256: rows = adj_rows<256>(); break;
        case 384: rows = adj_rows<384>(); break;
        case 512: rows = adj_rows<512>(); break;
    }
    return (ny + rows - 1) / rows;
}

hipError_t launch_adj_row(bool f64, int nx, int ny, hipStream_t st, const AdjArgs& a, const double2* tw) {
    const dim3 grid(adj_row_workgroups(nx, ny));
    return f64 ? adj_row_n<double2>(nx, grid, st, a, tw) : adj_row_n<float2>(nx, grid, st, a, tw);
}

// columns per workgroup: the adjoint runs one instance, so the grid
// (nx / COLS x 2 fields) is what spreads it over the CUs; four columns
// (32-byte row segments) where that still gives 256 workgroups, else two
static int adj_cols(int nx) { return nx >= 512 ? 4 : 2; }

template <typename S>
static hipError_t adj_col_n(int ny, int nx, hipStream_t st, const AdjArgs& a, const double2* tw) {
    const int cols = adj_cols(nx);
    const dim3 grid(nx / cols, 2);
    switch (ny) {
#define SHG_ADJCOL(N)                                                                                 \
    case N:                                                                                           \
        if (cols == 4)                                                                                \
            hipLaunchKernelGGL((k_adj_col<N, 4, S>), grid, dim3(4, N / 8), 0, st, a, tw);            \
        else                                                                                          \
            hipLaunchKernelGGL((k_adj_col<N, 2, S>), grid, dim3(2, N / 8), 0, st, a, tw);            \
        return hipGetLastError();
        SHG_ADJCOL(64)
        SHG_ADJCOL(128)
        SHG_ADJCOL(192)
        SHG_ADJCOL(256)
        SHG_ADJCOL(384)
        SHG_ADJCOL(512)
#undef SHG_ADJCOL
    }
    return hipErrorInvalidValue;
}

int adj_col_workgroups(int nx) { return nx / adj_cols(nx) * 2; }

hipError_t launch_adj_col(bool f64, int nx, int ny, hipStream_t st, const AdjArgs& a, const double2* tw) {
    return f64 ? adj_col_n<double2>(ny, nx, st, a, tw) : adj_col_n<float2>(ny, nx, st, a, tw);
}

hipError_t launch_adj_seed(bool f64, void* lp, void* ls, const void* sout, long n, double coef, hipStream_t st) {
    if (f64)
        hipLaunchKernelGGL(k_adj_seed<double2>, dim3(1024), dim3(256), 0, st, static_cast<double2*>(lp),
                           static_cast<double2*>(ls), static_cast<const double2*>(sout), n, coef);
    else
        hipLaunchKernelGGL(k_adj_seed<float2>, dim3(1024), dim3(256), 0, st, static_cast<float2*>(lp),
                           static_cast<float2*>(ls), static_cast<const float2*>(sout), n, coef);
    return hipGetLastError();
}

int adj_pump_workgroups() { return 512; }

hipError_t launch_adj_pump(bool f64, const void* lam, int nx, int ny, double dx, double dy, double w0, double f,
                           double n, double zR, double Ap0, double* part, hipStream_t st) {
    if (f64)
        hipLaunchKernelGGL(k_adj_pump<double2>, dim3(adj_pump_workgroups()), dim3(256), 0, st,
                           static_cast<const double2*>(lam), nx, ny, dx, dy, w0, f, n, zR, Ap0, part);
    else
        hipLaunchKernelGGL(k_adj_pump<float2>, dim3(adj_pump_workgroups()), dim3(256), 0, st,
                           static_cast<const float2*>(lam), nx, ny, dx, dy, w0, f, n, zR, Ap0, part);
    return hipGetLastError();
}

hipError_t launch_adj_reduce(const double* part, int rows, int nb, int nv, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_reduce, dim3(rows), dim3(256), 0, st, part, nb, nv, out);
    return hipGetLastError();
}

}  // namespace shg
