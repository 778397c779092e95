// kernels_mixed.hip -- FP32-mixed fused SSFM passes for gfx950.
//
// One SSFM slice of the reference (src/headers/Solver.h:267-273) is
//   A <- R(A, z_s);  A <- IFFT2(Prop . FFT2(A)) / (nx ny);  A <- R(A, z_s)
// with R the RK4 call of src/headers/Solver.h:189-264.  Because the 2D FFT
// is separable and R is pointwise, the march is regrouped into two HBM
// passes per slice, each reading and writing both fields once:
//   row pass  (A): row IFFT (end of slice s-1) -> R_end(s-1) -> R_start(s)
//                  -> row FFT (start of slice s)
//   column pass (B): column FFT -> * Prop -> column IFFT
// so a slice costs 2 passes x 2 fields x (8 B read + 8 B write) = 64 B per
// grid point.  The 1/(nx ny) scale of cuFFTscale (src/headers/Efields.h:15-29)
// is folded into the propagator table.
//
// Precision: fields, FFTs and RK4 arithmetic in FP32; the QPM phase dk*z is
// formed in FP64 and reduced to a quarter turn before the FP32 sin/cos
// series (expi_fast).
#include "fft.hpp"
#include "kernels.hpp"

namespace shg {

// dA/dz of src/headers/Solver.h:57-80 (and :33-55 for THERMAL), FP32:
//   dP = i kp S conj(P) conj(e) - alpha_p/2 P
//   dS = i ks P^2 e - (alpha_s/2 + bfac/2 |S|^2) S
__device__ __forceinline__ void dadz(float2 P, float2 S, float2 e, float kp, float ks, float aph,
                                     float ash, float bfac, float2& dP, float2& dS) {
    const float2 spc = cmul(cmul(S, conjc(P)), conjc(e));
    dP = make_float2(-kp * spc.y - aph * P.x, kp * spc.x - aph * P.y);
    const float2 ppe = cmul(cmul(P, P), e);
    const float g = ash + 0.5f * bfac * abs2(S);
    dS = make_float2(-ks * ppe.y - g * S.x, ks * ppe.x - g * S.y);
}

// RK4 call of src/headers/Solver.h:231-257 on one point; stage inputs
// A + c*k with c = (c1, c2, c3), z offsets e[0], e[1] (stages 2 and 3), e[2].
__device__ __forceinline__ void rk4_point(float2& P, float2& S, const float2 (&e)[3], float kp,
                                          float ks, float aph, float ash, float bfac, float c1,
                                          float c2, float c3, float w) {
    float2 k1p, k1s, k2p, k2s, k3p, k3s, k4p, k4s;
    dadz(P, S, e[0], kp, ks, aph, ash, bfac, k1p, k1s);
    dadz(P + c1 * k1p, S + c1 * k1s, e[1], kp, ks, aph, ash, bfac, k2p, k2s);
    dadz(P + c2 * k2p, S + c2 * k2s, e[1], kp, ks, aph, ash, bfac, k3p, k3s);
    dadz(P + c3 * k3p, S + c3 * k3s, e[2], kp, ks, aph, ash, bfac, k4p, k4s);
    P = P + w * (k1p + 2.0f * k2p + 2.0f * k3p + k4p);
    S = S + w * (k1s + 2.0f * k2s + 2.0f * k3s + k4s);
}

__device__ __forceinline__ int heat_slice(int s, int nz, int nzh) {
    return (int)(((long long)s * nzh) / nz);
}
// number of optical slices mapped onto heat slice h: ceil((h+1)nz/nzh) - ceil(h nz/nzh)
__device__ __forceinline__ int heat_count(int h, int nz, int nzh) {
    const long long a = ((long long)(h + 1) * nz + nzh - 1) / nzh;
    const long long b = ((long long)h * nz + nzh - 1) / nzh;
    return (int)(a - b);
}

// Heat-voxel values a THERMAL R call reads at its 8 points of a row: the
// refractive indices (n_p, n_s)(T) and the running-dk base and slope.
struct VoxRow {
    float2 n[8];
    double dkb[8], dk[8];
};

template <int NX>
__device__ __forceinline__ void load_vox(VoxRow& v, const MixedInst& I, int h, int y, int tp) {
    constexpr int TPR = NX / 8;
    const size_t rowoff = (size_t)h * I.nx * I.ny + (size_t)y * I.nx;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int x = tp + q * TPR;
        v.n[q] = I.nn[rowoff + x];
        v.dkb[q] = I.dkbase[rowoff + x];
        v.dk[q] = I.dk[rowoff + x];
    }
}

// R call `call` (0 = first half of the slice, 1 = second half) of slice
// `step` on the 8 points x = tp + q*TPR of row y.  With THERMAL the values
// of the slice's heat voxels come in `vox` (heat slice h of `step`) and the
// heat source of the slice is deposited after the second call.
template <int NX, bool THERMAL>
__device__ __forceinline__ void nl_row(float2 (&p)[8], float2 (&s)[8], const MixedInst& I,
                                       int step, int call, int y, int tp, const VoxRow& vox, int h,
                                       bool deposit_q) {
    constexpr int TPR = NX / 8;
    const float c1 = I.c1, c2 = I.c2, c3 = I.c3, w = I.w;
    const float aph = I.ap_h, ash = I.as_h;
    if constexpr (!THERMAL) {
        const float2* ph = I.phase + (size_t)(step * 2 + call) * 3;
        const float2 e[3] = {ph[0], ph[1], ph[2]};
        const float kp = I.kp, ks = I.ks, bfac = I.bfac;
#pragma unroll
        for (int q = 0; q < 8; q++) rk4_point(p[q], s[q], e, kp, ks, aph, ash, bfac, c1, c2, c3, w);
    } else {
        const double pol = I.poling ? I.poling[step] : 0.0;
        const double zsc = I.dz / ((double)step + 1.0);
        const double o0 = (double)step + I.off[call][0];
        float* Qr = I.Q + (size_t)h * I.nx * I.ny + (size_t)y * I.nx;
        const bool first = step == 0 || heat_slice(step - 1, I.nz, I.nzh) != h;
        const float qw = 1.0f / (float)heat_count(h, I.nz, I.nzh);
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int x = tp + q * TPR;
            const float np = vox.n[q].x, ns = vox.n[q].y;  // n_p(T), n_s(T) of this voxel
            // 1/n by the hardware reciprocal refined with one Newton step
            float rp = __builtin_amdgcn_rcpf(np), rs = __builtin_amdgcn_rcpf(ns);
            rp = rp * __builtin_fmaf(-np, rp, 2.0f);
            rs = rs * __builtin_fmaf(-ns, rs, 2.0f);
            const float kp = I.kp0 * rp, ks = I.ks0 * rs, bfac = I.bf0 * ns;
            // running dk integral (src/headers/PhaseMatching.h:27-44) and the
            // phase DKint (s + off) dz / (s + 1) of src/headers/Solver.h:48,50;
            // the three offsets are 0.25 dz apart in both schemes, so
            // e(off + 1/4) = e(off) exp(i DKint dz / (4 (s + 1))).
            const double dki = vox.dkb[q] + (double)step * vox.dk[q];
            const float2 e0 = expi_fast(dki * o0 * zsc + pol);
            const float2 ed = expi_fast(0.25 * dki * zsc);
            const float2 e1 = cmul(e0, ed);
            const float2 e[3] = {e0, e1, cmul(e1, ed)};
            rk4_point(p[q], s[q], e, kp, ks, aph, ash, bfac, c1, c2, c3, w);
            if (deposit_q) {
                // Q = a_p I_p + a_s I_s + b_s I_s^2 -- src/headers/Tfield.h:260-276
                const float Ip = I.hce * np * abs2(p[q]);
                const float Is = I.hce * ns * abs2(s[q]);
                const float qv = (I.alpha_p * Ip + I.alpha_s * Is + I.beta_s * Is * Is) * qw;
                Qr[x] = first ? qv : Qr[x] + qv;
            }
        }
    }
}

// Row pass at global step g of every instance in the launch group.
//   g == 0       : R_start(0) -> row FFT
//   0 < g < nz   : row IFFT -> R_end(g-1) [-> save state if g-1 == nz-2] -> R_start(g) -> row FFT
//   g == nz      : row IFFT -> R_end(nz-1) -> spatial fields stored
// grid (ceil(max_ny/ROWS), 1, n_instances), block (NX/8, ROWS)
template <int NX, bool THERMAL, int ROWS>
__global__ __launch_bounds__(NX / 8 * ROWS) void k_rowpass(const MixedInst* __restrict__ insts, int g,
                                                           const double2* __restrict__ tw) {
    constexpr int TPR = NX / 8;
    constexpr int LD = FftShape<NX>::LDS;
    __shared__ float2 lds[2 * ROWS * LD];
    const MixedInst& I = insts[blockIdx.z];
    const int nz = I.nz;
    if (g > nz) return;
    const int y = blockIdx.x * ROWS + threadIdx.y;
    if (blockIdx.x * ROWS >= I.ny) return;
    const int tp = threadIdx.x;
    float2* my = lds + threadIdx.y * LD;  // pump buffer; SH buffer ROWS*LD further
    const size_t base = (size_t)y * NX + tp;
    float2 f[2][8];  // f[0] pump, f[1] SH: transformed together (shared barriers)
    float2(&p)[8] = f[0];
    float2(&s)[8] = f[1];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        p[q] = I.P[base + q * TPR];
        s[q] = I.S[base + q * TPR];
    }
    // heat-voxel values of the slab in registers, loaded once when both R
    // calls of this launch fall in the same heat slice
    VoxRow vox;
    int hv = -1;
    // per-slice trace: the state before the SSFM step of slice g
    const bool trace = I.trP && g <= nz - 1 && g % I.trace_every == 0;
    if (trace && g == 0) {
#pragma unroll
        for (int q = 0; q < 8; q++) {
            I.trP[base + q * TPR] = p[q];
            I.trS[base + q * TPR] = s[q];
        }
    }
    if (g > 0) {
        fft_multi<float2, NX, true, 2>(f, my, ROWS * LD, tw, tp);
        if constexpr (THERMAL) {
            hv = heat_slice(g - 1, nz, I.nzh);
            load_vox<NX>(vox, I, hv, y, tp);
        }
        nl_row<NX, THERMAL>(p, s, I, g - 1, 1, y, tp, vox, hv, THERMAL);
        if (g - 1 == nz - 2) {
#pragma unroll
            for (int q = 0; q < 8; q++) {
                I.outP[base + q * TPR] = p[q];
                I.outS[base + q * TPR] = s[q];
            }
        }
        if (trace) {
#pragma unroll
            for (int q = 0; q < 8; q++) {
                I.trP[base + q * TPR] = p[q];
                I.trS[base + q * TPR] = s[q];
            }
        }
    }
    if (g < nz) {
        int h = -1;
        if constexpr (THERMAL) {
            h = heat_slice(g, nz, I.nzh);
            if (h != hv) load_vox<NX>(vox, I, h, y, tp);
        }
        nl_row<NX, THERMAL>(p, s, I, g, 0, y, tp, vox, h, false);
        fft_multi<float2, NX, false, 2>(f, my, ROWS * LD, tw, tp);
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
        I.P[base + q * TPR] = p[q];
        I.S[base + q * TPR] = s[q];
    }
}

// Column pass of step g: column FFT -> * Prop -> column IFFT on COLS
// adjacent columns of one field (blockIdx.y: 0 pump, 1 SH).
// grid (ceil(max_nx/COLS), 2, n_instances), block (COLS, NY/8)
template <int NY, int COLS>
__global__ __launch_bounds__(NY / 8 * COLS) void k_colpass(const MixedInst* __restrict__ insts, int g,
                                                           const double2* __restrict__ tw) {
    constexpr int TPC = NY / 8;
    constexpr int LD = FftShape<NY>::LDS;
    __shared__ float2 lds[COLS * LD];
    const MixedInst& I = insts[blockIdx.z];
    if (g >= I.nz) return;
    const int nx = I.nx;
    // A workgroup reads COLS*8 bytes of each 128-byte row segment; the K
    // column groups sharing a segment are placed on one XCD (workgroups are
    // dealt to the 8 XCDs round-robin by linear id, so ids 8 apart share an
    // XCD and its L2), so the segment is fetched from HBM once, not K times.
    constexpr int K = COLS * 8 >= 128 ? 1 : 128 / (COLS * 8);
    int bx = blockIdx.x;
    if (K > 1 && gridDim.x % (8 * K) == 0) {
        const int lo = bx & 7, hi = (bx >> 3) % K, chunk = bx / (8 * K);
        bx = chunk * 8 * K + lo * K + hi;
    }
    const int x0 = bx * COLS;
    if (x0 >= nx) return;
    const int c = threadIdx.x, tp = threadIdx.y;
    const int x = x0 + c;
    const bool sh = blockIdx.y != 0;
    float2* F = sh ? I.S : I.P;
    const double2* ex = sh ? I.exS : I.exP;
    const double2* ey = sh ? I.eyS : I.eyP;
    float2* my = lds + c * LD;
    float2 v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) v[q] = F[(size_t)(tp + q * TPC) * nx + x];
    fft_dist<float2, NY, false>(v, my, tw, tp);
    // Prop = ex[x] ey[y], formed and applied in FP64 and rounded once, so
    // the propagator carries no fixed FP32 representation error per mode.
    const double2 exv = ex[x];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const double2 pr = cmul(exv, ey[tp + q * TPC]);
        const double vr = v[q].x, vi = v[q].y;
        v[q] = make_float2((float)(vr * pr.x - vi * pr.y), (float)(vr * pr.y + vi * pr.x));
    }
    fft_dist<float2, NY, true>(v, my, tw, tp);
#pragma unroll
    for (int q = 0; q < 8; q++) F[(size_t)(tp + q * TPC) * nx + x] = v[q];
}

// ------------------------------------------------------------------
// dispatch
// ------------------------------------------------------------------

// Built-in tile plans (measured on MI355X: rows per workgroup of the row
// pass, columns per workgroup of the column pass); shg_autotune refines them.
int rowpass_rows_default(int nx) {
    switch (nx) {
        case 64: return 8;
        case 128: return 4;
        case 192: return 2;
        case 256: return 2;
        case 384: return 1;
        case 512: return 1;
    }
    return 0;
}
bool rowpass_rows_supported(int nx, int rows) {
    return (rows == 1 || rows == 2 || rows == 4 || rows == 8 || rows == 16) && nx / 8 * rows <= 1024;
}
int colpass_cols_default(int ny) {
    switch (ny) {
        case 64: return 8;
        case 128: return 8;
        case 192: return 8;
        case 256: return 8;
        case 384: return 8;
        case 512: return 8;
    }
    return 0;
}
bool colpass_cols_supported(int ny, int cols) {
    return (cols == 2 || cols == 4 || cols == 8 || cols == 16) && ny / 8 * cols <= 1024;
}

template <int NX, bool TH>
static hipError_t rowpass_rows(int rows, dim3 grid, hipStream_t st, const MixedInst* insts, int g,
                               const double2* tw) {
    dim3 block(NX / 8, rows);
    switch (rows) {
        case 1: hipLaunchKernelGGL((k_rowpass<NX, TH, 1>), grid, block, 0, st, insts, g, tw); break;
        case 2: hipLaunchKernelGGL((k_rowpass<NX, TH, 2>), grid, block, 0, st, insts, g, tw); break;
        case 4: hipLaunchKernelGGL((k_rowpass<NX, TH, 4>), grid, block, 0, st, insts, g, tw); break;
        case 8:
            if constexpr (NX / 8 * 8 <= 1024) {
                hipLaunchKernelGGL((k_rowpass<NX, TH, 8>), grid, block, 0, st, insts, g, tw);
                break;
            }
            return hipErrorInvalidValue;
        case 16:
            if constexpr (NX / 8 * 16 <= 1024) {
                hipLaunchKernelGGL((k_rowpass<NX, TH, 16>), grid, block, 0, st, insts, g, tw);
                break;
            }
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_rowpass(int nx, bool thermal, int rows, dim3 grid, hipStream_t st,
                          const MixedInst* insts, int g, const double2* tw) {
#define ROWCASE(N)                                                                 \
    case N:                                                                        \
        return thermal ? rowpass_rows<N, true>(rows, grid, st, insts, g, tw)       \
                       : rowpass_rows<N, false>(rows, grid, st, insts, g, tw);
    switch (nx) {
        ROWCASE(64)
        ROWCASE(128)
        ROWCASE(192)
        ROWCASE(256)
        ROWCASE(384)
        ROWCASE(512)
    }
#undef ROWCASE
    return hipErrorInvalidValue;
}

template <int NY>
static hipError_t colpass_cols(int cols, dim3 grid, hipStream_t st, const MixedInst* insts, int g,
                               const double2* tw) {
    dim3 block(cols, NY / 8);
    switch (cols) {
        case 2: hipLaunchKernelGGL((k_colpass<NY, 2>), grid, block, 0, st, insts, g, tw); break;
        case 4: hipLaunchKernelGGL((k_colpass<NY, 4>), grid, block, 0, st, insts, g, tw); break;
        case 8: hipLaunchKernelGGL((k_colpass<NY, 8>), grid, block, 0, st, insts, g, tw); break;
        case 16:
            if constexpr (NY / 8 * 16 <= 1024) {
                hipLaunchKernelGGL((k_colpass<NY, 16>), grid, block, 0, st, insts, g, tw);
                break;
            }
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_colpass(int ny, int cols, dim3 grid, hipStream_t st, const MixedInst* insts,
                          int g, const double2* tw) {
    switch (ny) {
        case 64: return colpass_cols<64>(cols, grid, st, insts, g, tw);
        case 128: return colpass_cols<128>(cols, grid, st, insts, g, tw);
        case 192: return colpass_cols<192>(cols, grid, st, insts, g, tw);
        case 256: return colpass_cols<256>(cols, grid, st, insts, g, tw);
        case 384: return colpass_cols<384>(cols, grid, st, insts, g, tw);
        case 512: return colpass_cols<512>(cols, grid, st, insts, g, tw);
    }
    return hipErrorInvalidValue;
}

}  // namespace shg
