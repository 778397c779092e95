// kernels_heat_gemm.hip -- the eigenbasis transforms of the direct heat
// solver (one dense n x n matrix applied along one axis of the 3D unknown
// block) on gfx950 matrix cores, FP32 data.
//
// Split-FP16 products: each FP32 operand x, scaled by a power of two s so
// that max|x s| lies in [2^13, 2^14), is held as hi = fp16(x s) and
// lo = fp16(x s - hi) (22 significant bits together); a product is
//   a b ~ (a_hi b_hi + a_hi b_lo + a_lo b_hi) / (s_a s_b)
// with the three terms accumulated in FP32 by v_mfma_f32_32x32x16_f16
// (the dropped lo * lo term is ~2^-22 of the product).  The absolute error
// of an output is then ~2^-22 of max|a| max|b| K -- the error of an FP32
// product accumulation -- at the FP16 matrix rate (2.5 PFLOP/s dense on
// MI355X against 157 TFLOP/s for FP32 matrices; three products per term).
// The data's scale comes from the device slot `umax` (max|u| over the whole
// input, written by the kernel that produced it); the eigenbasis is split
// once per geometry (k_split_basis, from the FP64 eigenvectors, scale
// 2^qe with max|q 2^qe| in [2^13, 2^14)).
//
// The product C[i][j] = sum_k R(i, k) S(k, j), C row-major (j contiguous):
//   kind A (transform along y or z):  R = Q_op [m][k], S(k, j) = u[k ldu + j]
//   kind B (transform along x):       R(i, k) = u[i ldu + k], S(k, j) = Q_op [j][k]
// Q_op is the pre-split eigenbasis in the direction asked for, rows padded
// to a multiple of 128 and k to a multiple of 32 with zeros.
// Workgroup tile 128 x 128 (four wavefronts of 64 x 64, four 32 x 32
// accumulator blocks each), K in chunks of 32 through LDS, both operands
// stored in LDS with k contiguous so that an MFMA operand (8 consecutive k
// of one row or column) is one 16-byte read.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "heat_kernels.hpp"

namespace shg {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kTI = 128, kTJ = 128, kTK = 32, kLDK = kTK + 8;

// scale of an operand whose largest magnitude is mx: x * 2^(14 - e), mx = f 2^e, f in [0.5, 1)
__device__ __forceinline__ int scale_exp(float mx) {
    int e = 0;
    (void)frexpf(mx, &e);
    return 14 - e;
}

typedef float f4 __attribute__((ext_vector_type(4)));

// Each thread stages two 8-wide k groups of each operand per K chunk:
// item e = t + 256 s (s = 0, 1).  Eigenbasis: row e/4, k group e%4.  Data:
// kind B row e/4, k group e%4 (8 consecutive floats of one row); kind A
// column e%128, k group e/128 (8 rows of one column: each load instruction
// reads 64 consecutive floats of one row).  The next chunk's operands are
// loaded into registers while the current chunk's products run.
template <bool KINDB>
struct Stage {
    h8 qh[2], ql[2];
    float d[2][8];
};

template <bool KINDB>
__device__ __forceinline__ void stage_load(const HeatGemm& g, const float* __restrict__ u,
                                           const _Float16* __restrict__ qh, const _Float16* __restrict__ ql, long i0,
                                           long j0, long I, long J, int k0, int t, bool vec, Stage<KINDB>& st) {
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const int e = t + 256 * s;
        {
            const long r0 = KINDB ? j0 : i0;
            const size_t src = (size_t)(r0 + (e >> 2)) * g.qpad + k0 + (e & 3) * 8;
            st.qh[s] = *reinterpret_cast<const h8*>(&qh[src]);
            st.ql[s] = *reinterpret_cast<const h8*>(&ql[src]);
        }
        if constexpr (KINDB) {
            const long gi = i0 + (e >> 2);
            const int gk = k0 + (e & 3) * 8;
            const float* p = u + gi * g.ldu + gk;
            if (vec && gi < I && gk + 8 <= g.n) {
                const f4 a = *reinterpret_cast<const f4*>(p);
                const f4 b = *reinterpret_cast<const f4*>(p + 4);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    st.d[s][r] = a[r];
                    st.d[s][4 + r] = b[r];
                }
            } else {
#pragma unroll
                for (int r = 0; r < 8; r++) st.d[s][r] = (gi < I && gk + r < g.n) ? p[r] : 0.0f;
            }
        } else {
            const long gj = j0 + (e & 127);
            const int gk = k0 + (e >> 7) * 8;
            const float* p = u + (long)gk * g.ldu + gj;
            if (gj < J && gk + 8 <= g.n) {
#pragma unroll
                for (int r = 0; r < 8; r++) st.d[s][r] = p[r * g.ldu];
            } else {
#pragma unroll
                for (int r = 0; r < 8; r++) st.d[s][r] = (gj < J && gk + r < g.n) ? p[r * g.ldu] : 0.0f;
            }
        }
    }
}

template <bool KINDB>
__device__ __forceinline__ void stage_store(const Stage<KINDB>& st, float su, int t, _Float16* Rh, _Float16* Rl,
                                            _Float16* Sh, _Float16* Sl) {
    _Float16* Qh = KINDB ? Sh : Rh;
    _Float16* Ql = KINDB ? Sl : Rl;
    _Float16* Dh = KINDB ? Rh : Sh;
    _Float16* Dl = KINDB ? Rl : Sl;
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const int e = t + 256 * s;
        const int qo = (e >> 2) * kLDK + (e & 3) * 8;
        *reinterpret_cast<h8*>(&Qh[qo]) = st.qh[s];
        *reinterpret_cast<h8*>(&Ql[qo]) = st.ql[s];
        const int dof = KINDB ? (e >> 2) * kLDK + (e & 3) * 8 : (e & 127) * kLDK + (e >> 7) * 8;
        h8 hi, lo;
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const float x = st.d[s][r] * su;
            const _Float16 h = (_Float16)x;
            hi[r] = h;
            lo[r] = (_Float16)(x - (float)h);
        }
        *reinterpret_cast<h8*>(&Dh[dof]) = hi;
        *reinterpret_cast<h8*>(&Dl[dof]) = lo;
    }
}

// kind B fits three waves per SIMD without spilling; kind A does not
template <bool KINDB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KINDB ? 3 : 1))) void k_heat_gemm(HeatGemm g) {
    __shared__ __align__(16) _Float16 Rh[kTI * kLDK], Rl[kTI * kLDK], Sh[kTJ * kLDK], Sl[kTJ * kLDK];
    __shared__ float wmax[4];
    const int b = blockIdx.z;
    const float* u = g.u + (size_t)b * g.su;
    float* c = g.c + (size_t)b * g.sc;
    const _Float16* qh = g.qh + (size_t)b * g.sq;
    const _Float16* ql = g.ql + (size_t)b * g.sq;
    const long I = KINDB ? g.nbig : g.n, J = KINDB ? g.n : g.nbig;
    // Tile order (grid x: data tiles, y: eigenbasis tiles).  The gy tiles
    // that share a data tile are placed 8 apart in the dispatch order: the
    // same XCD (workgroups are dealt to the 8 XCDs round-robin) at about the
    // same time, so the data tile comes from HBM once and from that XCD's L2
    // for the others.  A bijection of the grid: groups of 8 data tiles, then
    // the remaining gx % 8 in plain order.
    int dt = blockIdx.x, et = blockIdx.y;
    {
        const int gx = gridDim.x, gy = gridDim.y;
        if (gy > 1) {
            const int L = blockIdx.x + gx * blockIdx.y;
            const int full = gx / 8 * 8, head = full * gy;
            if (L < head) {
                const int chunk = L / (8 * gy), r = L - chunk * 8 * gy;
                dt = chunk * 8 + (r & 7);
                et = r >> 3;
            } else {
                const int l = L - head, rem = gx - full;
                dt = full + l % rem;
                et = l / rem;
            }
        }
    }
    const long i0 = (long)(KINDB ? dt : et) * kTI;
    const long j0 = (long)(KINDB ? et : dt) * kTJ;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wi = w >> 1, wj = w & 1;
    const float mx = __uint_as_float(*g.umax);
    const int eu = mx > 0.0f ? scale_exp(mx) : 0;
    const float su = ldexpf(1.0f, eu);
    // 16-byte loads of 8 data floats: rows and the batch offset 4-aligned
    const bool vec = KINDB && (g.ldu & 3) == 0 && (g.su & 3) == 0 && ((size_t)g.u & 15) == 0;
    f32x16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int d = 0; d < 2; d++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[a][d][r] = 0.0f;
    Stage<KINDB> st;
    stage_load<KINDB>(g, u, qh, ql, i0, j0, I, J, 0, t, vec, st);
    for (int k0 = 0; k0 < g.n; k0 += kTK) {
        stage_store<KINDB>(st, su, t, Rh, Rl, Sh, Sl);
        __syncthreads();
        if (k0 + kTK < g.n) stage_load<KINDB>(g, u, qh, ql, i0, j0, I, J, k0 + kTK, t, vec, st);
#pragma unroll
        for (int ks = 0; ks < kTK; ks += 16) {
            const int kk = ks + 8 * (lane >> 5);
            h8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int a = 0; a < 2; a++) {
                const int row = wi * 64 + a * 32 + (lane & 31);
                ah[a] = *reinterpret_cast<const h8*>(&Rh[row * kLDK + kk]);
                al[a] = *reinterpret_cast<const h8*>(&Rl[row * kLDK + kk]);
            }
#pragma unroll
            for (int d = 0; d < 2; d++) {
                const int col = wj * 64 + d * 32 + (lane & 31);
                bh[d] = *reinterpret_cast<const h8*>(&Sh[col * kLDK + kk]);
                bl[d] = *reinterpret_cast<const h8*>(&Sl[col * kLDK + kk]);
            }
            // hi hi, then hi lo, then lo hi over the four accumulators (no
            // product waits on the one just issued)
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int d = 0; d < 2; d++)
                    acc[a][d] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[a], bh[d], acc[a][d], 0, 0, 0);
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int d = 0; d < 2; d++)
                    acc[a][d] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[a], bl[d], acc[a][d], 0, 0, 0);
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int d = 0; d < 2; d++)
                    acc[a][d] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[a], bh[d], acc[a][d], 0, 0, 0);
        }
        __syncthreads();
    }
    // C = acc / (2^qe s_u); accumulator (a, d, r) of lane l is
    // C[8 (r/4) + 4 (l/32) + r%4][l%32] of its 32 x 32 block
    const float so = ldexpf(1.0f, -g.qe - eu);
    float m = 0.0f;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int d = 0; d < 2; d++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const long gi = i0 + wi * 64 + a * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const long gj = j0 + wj * 64 + d * 32 + (lane & 31);
                if (gi < I && gj < J) {
                    const float v = acc[a][d][r] * so;
                    c[gi * g.ldc + gj] = v;
                    m = fmaxf(m, fabsf(v));
                }
            }
    if (g.cmax) {
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
        if (lane == 0) wmax[w] = m;
        __syncthreads();
        if (t == 0)
            g.cpart[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] =
                __float_as_uint(fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3])));
    }
}

// FP64 transforms: C = R S with R [I][n] and S [n][J] both row-major (the
// eigenvectors q or their transpose take one of the two places), on the
// FP64 matrix cores (v_mfma_f64_16x16x4_f64).  Workgroup tile 64 x 64 (four
// wavefronts of 32 x 32, four 16 x 16 accumulator blocks each), K in
// chunks of 16 through LDS, the next chunk loaded into registers while the
// current one's products run.  MFMA operands: lane l holds A[l%16][l/16]
// and B[l/16][l%16]; accumulator r of lane l is C[4 r + l/16][l%16].
typedef double f64x4 __attribute__((ext_vector_type(4)));
constexpr int kDT = 64, kDK = 16;

__global__ __launch_bounds__(256) void k_heat_gemm64(HeatGemm64 g) {
    __shared__ double Rs[kDT][kDK + 1];
    __shared__ double Ss[kDK][kDT + 1];
    const int b = blockIdx.z;
    const double* R = g.r + (size_t)b * g.sr;
    const double* S = g.s + (size_t)b * g.ss;
    double* C = g.c + (size_t)b * g.sc;
    const long i0 = (long)blockIdx.y * kDT, j0 = (long)blockIdx.x * kDT;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wi = w >> 1, wj = w & 1;
    // staging: R rows t/4, k (t%4)*4 + q; S row k t/16, columns (t%16)*4 + q
    const long gi = i0 + (t >> 2), gj = j0 + (t & 15) * 4;
    const int rk = (t & 3) * 4, sk = t >> 4;
    double rr[4], sv[4];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int k = k0 + rk + q;
            rr[q] = (gi < g.I && k < g.n) ? R[gi * g.ldr + k] : 0.0;
            const int ks = k0 + sk;
            sv[q] = (ks < g.n && gj + q < g.J) ? S[(long)ks * g.lds + gj + q] : 0.0;
        }
    };
    f64x4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int r = 0; r < 4; r++) acc[x][y][r] = 0.0;
    load(0);
    for (int k0 = 0; k0 < g.n; k0 += kDK) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            Rs[t >> 2][rk + q] = rr[q];
            Ss[sk][(t & 15) * 4 + q] = sv[q];
        }
        __syncthreads();
        if (k0 + kDK < g.n) load(k0 + kDK);
#pragma unroll
        for (int ks = 0; ks < kDK; ks += 4) {
            const int kk = ks + (lane >> 4);
            double a[2], bb[2];
#pragma unroll
            for (int x = 0; x < 2; x++) a[x] = Rs[wi * 32 + x * 16 + (lane & 15)][kk];
#pragma unroll
            for (int y = 0; y < 2; y++) bb[y] = Ss[kk][wj * 32 + y * 16 + (lane & 15)];
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], bb[y], acc[x][y], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const long ci = i0 + wi * 32 + x * 16 + 4 * r + (lane >> 4);
                const long cj = j0 + wj * 32 + y * 16 + (lane & 15);
                if (ci < g.I && cj < g.J) C[ci * g.ldc + cj] = acc[x][y][r];
            }
}

hipError_t launch_heat_gemm64(hipStream_t st, const HeatGemm64& g, int batch) {
    const long ti = (g.I + kDT - 1) / kDT, tj = (g.J + kDT - 1) / kDT;
    hipLaunchKernelGGL(k_heat_gemm64, dim3((unsigned)tj, (unsigned)ti, batch), dim3(256), 0, st, g);
    return hipGetLastError();
}

// max |a| over n floats: per-workgroup maxima into part (float bits)
__global__ __launch_bounds__(256) void k_absmax(const float* __restrict__ a, long n, unsigned* __restrict__ part) {
    __shared__ float wm[4];
    float m = 0.0f;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += 256L * gridDim.x) m = fmaxf(m, fabsf(a[i]));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = __float_as_uint(fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3])));
}

// *slot = max(*slot, part[0..n)): a few workgroups, one atomic each
__global__ __launch_bounds__(256) void k_max_parts(const unsigned* __restrict__ part, long n,
                                                   unsigned* __restrict__ slot) {
    __shared__ unsigned wm[4];
    unsigned m = 0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += 256L * gridDim.x) m = max(m, part[i]);
    for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off));
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(slot, max(max(wm[0], wm[1]), max(wm[2], wm[3])));
}

// the split eigenbasis of one axis in one direction: block b, row m, k of
// Q_op (Q_op[m][k] = q[k n + m] forward, q[m n + k] inverse; q row-major
// per block, component j of vector i at q[j n + i]), zero beyond n.
__global__ __launch_bounds__(256) void k_split_basis(const double* __restrict__ q, int n, int nb, int inv, int qe,
                                                     _Float16* __restrict__ hi, _Float16* __restrict__ lo, int rpad,
                                                     int kpad) {
    const long tot = (long)nb * rpad * kpad;
    const double s = ldexp(1.0, qe);
    for (long e = blockIdx.x * 256L + threadIdx.x; e < tot; e += 256L * gridDim.x) {
        const long b = e / ((long)rpad * kpad), r = e % ((long)rpad * kpad);
        const int m = (int)(r / kpad), k = (int)(r % kpad);
        double x = 0.0;
        if (m < n && k < n) x = s * q[(size_t)b * n * n + (inv ? (size_t)m * n + k : (size_t)k * n + m)];
        const _Float16 h = (_Float16)(float)x;
        hi[e] = h;
        lo[e] = (_Float16)(float)(x - (double)(float)h);
    }
}

__global__ void k_zero_u32(unsigned* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

hipError_t launch_max_parts(hipStream_t st, const unsigned* part, long n, unsigned* slot) {
    const long nb = std::min(64L, (n + 255) / 256);
    hipLaunchKernelGGL(k_max_parts, dim3((unsigned)std::max(1L, nb)), dim3(256), 0, st, part, n, slot);
    return hipGetLastError();
}

static long gemm_tiles(long rows, long cols) { return ((rows + kTI - 1) / kTI) * ((cols + kTJ - 1) / kTJ); }

long heat_gemm_max_blocks(int nxu, int nyu, int nz) {
    // unfolded lengths bound the folded ones (two blocks of half the length)
    return std::max({gemm_tiles((long)nyu * nz, nxu) * 2, gemm_tiles(nyu, nxu) * nz,
                     gemm_tiles(nz, (long)nxu * nyu) * 2});
}

hipError_t launch_heat_gemm(hipStream_t st, const HeatGemm& g, int batch) {
    const long ti = ((g.kindB ? g.nbig : g.n) + kTI - 1) / kTI;
    const long tj = ((g.kindB ? g.n : g.nbig) + kTJ - 1) / kTJ;
    if (g.kindB)
        hipLaunchKernelGGL(k_heat_gemm<true>, dim3((unsigned)ti, (unsigned)tj, batch), dim3(256), 0, st, g);
    else
        hipLaunchKernelGGL(k_heat_gemm<false>, dim3((unsigned)tj, (unsigned)ti, batch), dim3(256), 0, st, g);
    if (g.cmax) return launch_max_parts(st, g.cpart, ti * tj * batch, g.cmax);
    return hipGetLastError();
}

hipError_t launch_absmax(hipStream_t st, const float* a, long n, unsigned* part, unsigned* out) {
    hipLaunchKernelGGL(k_absmax, dim3(1024), dim3(256), 0, st, a, n, part);
    return launch_max_parts(st, part, 1024, out);
}

hipError_t launch_split_basis(hipStream_t st, const double* q, int n, int nb, bool inv, int qe, _Float16* hi,
                              _Float16* lo, int rpad, int kpad) {
    hipLaunchKernelGGL(k_split_basis, dim3(256), dim3(256), 0, st, q, n, nb, inv ? 1 : 0, qe, hi, lo, rpad, kpad);
    return hipGetLastError();
}

hipError_t launch_zero_u32(hipStream_t st, unsigned* p, int n) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, p, n);
    return hipGetLastError();
}

}  // namespace shg
