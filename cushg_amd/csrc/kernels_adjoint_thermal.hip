// kernels_adjoint_thermal.hip -- reverse-mode row pass of the fused THERMAL
// march (FP64 arithmetic, fields stored as S).  No reference counterpart.
//
// The forward THERMAL row pass at boundary s (src/headers/Solver.h:33-55,
// 198-229, 299-301 restated by NlRow in mixed_passes.hpp) takes the state
// Y_{s-1} = IFx C_{s-1} through
//   X_s = R2_{s-1}(Y_{s-1})     couplings and phases of heat slab h(s-1)
//   q_{s-1}(X_s) -> Q[h(s-1)]   heat source, mean over the slab's slices
//   [X_{nz-1} is the saved output state]
//   U_s = R1_s(X_s)             couplings and phases of heat slab h(s)
// with, per voxel v of slab h, kp = kp0 / n_p(T_v), ks = ks0 / n_s(T_v),
// bfac = bf0 n_s(T_v) and the phases
//   phi_j = DKint_s (s + off_j) dz / (s + 1) + poling[s],
//   DKint_s = dkbase[v] + s dk[v] = sum_{j <= s} dk[h(j)].
// The reverse for a loss L = eta + sum_v w_v Q_v runs
//   lam_X = R1_s^T lam_U  (+ the eta seed at s = nz - 1)
//         + grad_X <w, Q>(q_{s-1})
//   lam_Y = R2_{s-1}^T lam_X
// and accumulates, per voxel, dL/dn_p, dL/dn_s and the two sums from which
// dL/ddk[h] follows (AdjThermal), plus dL/dpoling per slice.
#include "adjoint.hpp"
#include "adjoint_math.hpp"
#include "mixed_passes.hpp"

namespace shg {

// block sum of NV values per thread into part[block * NV + v]
template <int NV>
__device__ __forceinline__ void block_sums_th(double (&v)[NV], double* part, double* red) {
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
        part[(size_t)blockIdx.x * NV + t] = a;
    }
}

// per-point constants of one RK4 call of slice sl on heat voxel v
struct ThCall {
    double2 e[3];
    double kp, ks, bf, np, ns, dki, zs;
};
// the heat-grid values of voxel v that a call reads, loaded through the
// global address space (issued together, ahead of their use)
struct ThVox {
    float2 n;
    double dkb, dk;
};
__device__ __forceinline__ ThVox th_vox(const MixedInst& I, size_t v) {
    return ThVox{ld_global(I.nn + v), ld_global(I.dkbase + v), ld_global(I.dk + v)};
}
__device__ __forceinline__ ThCall th_call(const MixedInst& I, const ThVox& x, int sl, const double (&off)[3]) {
    ThCall c;
    const float2 n = x.n;
    c.np = n.x;
    c.ns = n.y;
    c.kp = I.kp0 * recip<double>(n.x);
    c.ks = I.ks0 * recip<double>(n.y);
    c.bf = I.bf0 * (double)n.y;
    c.dki = x.dkb + (double)sl * x.dk;
    c.zs = I.dz / ((double)sl + 1.0);
    const double pol = I.poling ? I.poling[sl] : 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) c.e[j] = expi_fast64(c.dki * ((double)sl + off[j]) * c.zs + pol);
    return c;
}

// Adjoint THERMAL row pass at boundary s: ROWS rows per workgroup of
// NX * ROWS / 2 threads, grid (ny / ROWS): the 4 ROWS transforms (lam_P,
// lam_S, Y_P, Y_S of each row) on groups of NX/8 threads, so that every
// wave takes part in them, and the pointwise reverse work two points per
// thread.
//   first       apply R1_s^T (lam is in (kx, y) and is IFx'd first)
//   second      apply the heat-source gradient of q_{s-1} and R2_{s-1}^T,
//               and Fx the result
//   fwd_spatial fwd is the spatial initial state X_0 (s == 0)
template <int NX, int ROWS, typename S>
__global__ __launch_bounds__(NX * ROWS / 2) void k_adj_row_th(AdjArgs a, AdjThermal th, const double2* __restrict__ tw) {
    using C = double2;
    constexpr int TPR = NX / 8;
    constexpr int NT = NX * ROWS / 2;  // threads: the 4 ROWS transform groups, two points each in the pointwise work
    constexpr int LD = FftShape<NX>::LDS;
    __shared__ C lds[4 * ROWS * LD];
    __shared__ double red[16 * 2];
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
    const int s = a.s, nz = I.nz, nzh = I.nzh;
    double pol1 = 0.0, pol2 = 0.0;
#pragma unroll 1
    for (int pt = t; pt < NX * ROWS; pt += NT) {
        const int r = pt / NX, x = pt % NX;
        const int y = blockIdx.x * ROWS + r;
        C* rb = lds + r * 4 * LD;
        const int xi = fpad(x);
        const size_t plane = (size_t)I.nx * I.ny, pix = (size_t)y * I.nx + x;
        const bool c1 = a.first != 0, c2 = a.second != 0;
        const int h0 = heat_slice(min(s, nz - 1), nz, nzh), h1 = heat_slice(max(s - 1, 0), nz, nzh);
        const size_t v0 = (size_t)h0 * plane + pix, v1 = (size_t)h1 * plane + pix;
        const double off0[3] = {I.off[0][0], I.off[0][1], I.off[0][2]};
        const double off1[3] = {I.off[1][0], I.off[1][1], I.off[1][2]};
        // the point's heat-grid loads, all in flight before any is used (both
        // voxels are valid rows whether or not their call runs)
        const ThVox x1 = th_vox(I, v1), x0 = th_vox(I, v0);
        const double w1 = c2 && th.w ? ld_global(th.w + v1) : 0.0;
        const C Yp = rb[2 * LD + xi], Ys = rb[3 * LD + xi];
        C Xp = Yp, Xs = Ys;
        ThCall k2{};
        if (c2 || !a.fwd_spatial) {
            k2 = th_call(I, x1, s - 1, off1);
            rk4_point(Xp, Xs, k2.e, k2.kp, k2.ks, k2.bf, k);  // X_s = R2_{s-1}(Y_{s-1})
        }
        C lp = rb[xi], ls = rb[LD + xi];
        double acc1[6] = {0, 0, 0, 0, 0, 0}, acc2[6] = {0, 0, 0, 0, 0, 0};
        double gnp0 = 0, gns0 = 0, gA0 = 0, gC0 = 0, gnp1 = 0, gns1 = 0, gA1 = 0, gC1 = 0;
        if (c1) {
            const ThCall k1 = th_call(I, x0, s, off0);
            rk4_adj(lp, ls, Xp, Xs, k1.e, k1.kp, k1.ks, k1.bf, k, acc1);
            double cs = 0.0;
#pragma unroll
            for (int j = 0; j < 3; j++) cs += acc1[j] * ((double)s + off0[j]) * k1.zs;
            const long long s0 = ((long long)h0 * nz + nzh - 1) / nzh;
            gA0 = cs * (double)(s - s0 + 1);
            gC0 = cs;
            gnp0 = -acc1[3] * k1.kp / k1.np;
            gns0 = -acc1[4] * k1.ks / k1.ns + acc1[5] * I.bf0;
            pol1 += acc1[0] + acc1[1] + acc1[2];
        }
        // the per-voxel accumulators' current values, loaded ahead of the
        // second call's reverse step (each point of a launch belongs to one
        // thread, so the read-modify-write has no other writer)
        const bool same = c1 && c2 && h0 == h1;
        double q0[4] = {0, 0, 0, 0}, q1[4] = {0, 0, 0, 0};
        if (c1) {
            q0[0] = ld_global(th.gnp + v0);
            q0[1] = ld_global(th.gns + v0);
            q0[2] = ld_global(th.gA + v0);
            q0[3] = ld_global(th.gC + v0);
        }
        if (c2 && !same) {
            q1[0] = ld_global(th.gnp + v1);
            q1[1] = ld_global(th.gns + v1);
            q1[2] = ld_global(th.gA + v1);
            q1[3] = ld_global(th.gC + v1);
        }
        if (th.seed != 0.0 && s == nz - 1) ls = axpy(th.seed, Xs, ls);
        if (c2) {
            if (th.w) {
                // heat source of slice s-1 at X_s: q = a_p I_p + a_s I_s + b_s I_s^2,
                // I = hce n |A|^2, averaged over the slab's slices
                const double wq = w1 / (double)heat_count(h1, nz, nzh);
                const double ap2 = abs2(Xp), as2 = abs2(Xs);
                const double Is = I.hce * k2.ns * as2;
                const double dqp = I.alpha_p * I.hce * k2.np;                         // dq / d|P|^2
                const double dqs = I.alpha_s * I.hce * k2.ns + 2.0 * I.beta_s * Is * I.hce * k2.ns;  // dq / d|S|^2
                lp = axpy(2.0 * wq * dqp, Xp, lp);
                ls = axpy(2.0 * wq * dqs, Xs, ls);
                gnp1 += wq * I.alpha_p * I.hce * ap2;
                gns1 += wq * (I.alpha_s * I.hce * as2 + 2.0 * I.beta_s * Is * I.hce * as2);
            }
            rk4_adj(lp, ls, Yp, Ys, k2.e, k2.kp, k2.ks, k2.bf, k, acc2);
            double cs = 0.0;
#pragma unroll
            for (int j = 0; j < 3; j++) cs += acc2[j] * ((double)(s - 1) + off1[j]) * k2.zs;
            const long long s0 = ((long long)h1 * nz + nzh - 1) / nzh;
            gA1 += cs * (double)(s - 1 - s0 + 1);
            gC1 += cs;
            gnp1 += -acc2[3] * k2.kp / k2.np;
            gns1 += -acc2[4] * k2.ks / k2.ns + acc2[5] * I.bf0;
            pol2 += acc2[0] + acc2[1] + acc2[2];
        }
        // per-voxel accumulation
        if (same) {
            st_global(th.gnp + v0, q0[0] + (gnp0 + gnp1));
            st_global(th.gns + v0, q0[1] + (gns0 + gns1));
            st_global(th.gA + v0, q0[2] + (gA0 + gA1));
            st_global(th.gC + v0, q0[3] + (gC0 + gC1));
        } else {
            if (c1) {
                st_global(th.gnp + v0, q0[0] + gnp0);
                st_global(th.gns + v0, q0[1] + gns0);
                st_global(th.gA + v0, q0[2] + gA0);
                st_global(th.gC + v0, q0[3] + gC0);
            }
            if (c2) {
                st_global(th.gnp + v1, q1[0] + gnp1);
                st_global(th.gns + v1, q1[1] + gns1);
                st_global(th.gA + v1, q1[2] + gA1);
                st_global(th.gC + v1, q1[3] + gC1);
            }
        }
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
    double vv[2] = {pol1, pol2};
    block_sums_th<2>(vv, a.part, red);
}

struct LinComb {
    double c[6];
    const double* x[6];
    int n;
};
__global__ __launch_bounds__(256) void k_adj_lincomb(double* out, long len, LinComb lc) {
    for (long i = blockIdx.x * 256L + threadIdx.x; i < len; i += 256L * gridDim.x) {
        double a = 0.0;
        for (int j = 0; j < lc.n; j++) a += lc.c[j] * lc.x[j][i];
        out[i] = a;
    }
}

constexpr int kAdjDotBlocks = 1024;
__global__ __launch_bounds__(256) void k_adj_dot(const double* __restrict__ a, const double* __restrict__ b, long n,
                                                 double* __restrict__ part) {
    __shared__ double red[4];
    double acc = 0.0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += 256L * gridDim.x) acc += a[i] * b[i];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

hipError_t launch_adj_lincomb(hipStream_t st, double* out, long n, int nx, const double* c, const double* const* x) {
    if (nx < 1 || nx > 6) return hipErrorInvalidValue;
    LinComb lc{};
    lc.n = nx;
    for (int j = 0; j < nx; j++) {
        lc.c[j] = c[j];
        lc.x[j] = x[j];
    }
    hipLaunchKernelGGL(k_adj_lincomb, dim3(2048), dim3(256), 0, st, out, n, lc);
    return hipGetLastError();
}

hipError_t launch_adj_dot(hipStream_t st, const double* a, const double* b, long n, double* part) {
    hipLaunchKernelGGL(k_adj_dot, dim3(kAdjDotBlocks), dim3(256), 0, st, a, b, n, part);
    return hipGetLastError();
}

int adj_dot_workgroups() { return kAdjDotBlocks; }

template <int NX> constexpr int adj_rows_th() { return NX >= 256 ? 1 : 256 / NX; }

template <typename S>
static hipError_t adj_row_th_n(int nx, dim3 grid, hipStream_t st, const AdjArgs& a, const AdjThermal& t,
                               const double2* tw) {
    switch (nx) {
#define SHG_ADJROWTH(N)                                                                                   \
    case N:                                                                                               \
        hipLaunchKernelGGL((k_adj_row_th<N, adj_rows_th<N>(), S>), grid, dim3(N * adj_rows_th<N>() / 2), 0, st, a, t, \
                           tw);                                                                          \
        return hipGetLastError();
        SHG_ADJROWTH(64)
        SHG_ADJROWTH(128)
        SHG_ADJROWTH(192)
        SHG_ADJROWTH(256)
        SHG_ADJROWTH(384)
        SHG_ADJROWTH(512)
#undef SHG_ADJROWTH
    }
    return hipErrorInvalidValue;
}

hipError_t launch_adj_row_th(bool f64, int nx, int ny, hipStream_t st, const AdjArgs& a, const AdjThermal& t,
                             const double2* tw) {
    const dim3 grid(adj_row_workgroups(nx, ny));
    return f64 ? adj_row_th_n<double2>(nx, grid, st, a, t, tw) : adj_row_th_n<float2>(nx, grid, st, a, t, tw);
}

}  // namespace shg
