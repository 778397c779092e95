// kernels_heat.hip -- THERMAL kernels: the reference's Jacobi sweep (as a
// sweep and as an FP64 residual), the pointwise stages of the
// fast-diagonalisation solve, dk(T) and the running dk integral base.
//
// The stencil kernels march along z inside each thread (threads cover the
// x-y plane, x fastest), keeping T at z-1, z, z+1 in registers: every
// voxel's temperature is read from memory once per sweep (plus its x/y
// neighbours, which hit the caches), as in the reference's in-thread
// z-loop (src/headers/Tfield.h:111).  The z extent is cut into chunks
// (grid.z) so that a column's dependent loads are spread over several
// threads: one chunk re-reads only the plane below it.
#include <algorithm>

#include "heat_kernels.hpp"

namespace shg {

struct Nb {
    double c, xm, xp, ym, yp, zm, zp;
};

// One application of the Jacobi map at a non-Dirichlet node (x, y, z),
// given the values of its neighbours (those outside the crystal are not
// used).  Open top: src/headers/Tfield.h:93-201 (y = 0 is the oven plane).
// Oven-surrounded: src/headers/Tfield.h:205-240 (the four lateral planes
// are the oven; inner nodes, and convective z faces without Q).
__device__ __forceinline__ double jacobi_vals(const HeatCoef& c, int x, int y, int z, const Nb& t, float q) {
    if (c.oven) {
        if (z > 0 && z < c.nz - 1)  // inner points, :217-219
            return c.L2 * ((t.xp + t.xm) * c.idx2 + (t.yp + t.ym) * c.idy2 + (t.zp + t.zm) * c.idz2 +
                           (double)q * c.ikth);
        const double zn = z == 0 ? t.zp : t.zm;  // back/front faces, :224-232
        return (zn + c.az * (t.xp + t.xm) + c.bz * (t.yp + t.ym) + c.Biz * c.Tinf) * c.Fz;
    }
    const bool xi = x > 0 && x < c.nx - 1, yi = y < c.ny - 1, zi = z > 0 && z < c.nz - 1;
    const bool xe = !xi, ze = !zi;
    const double xn = x == 0 ? t.xp : t.xm;  // inward x neighbour
    const double zn = z == 0 ? t.zp : t.zm;  // inward z neighbour
    const double Tinf = c.Tinf;
    if (xi && yi && zi)  // inner points, :113-115
        return c.L2 * ((t.xp + t.xm) * c.idx2 + (t.yp + t.ym) * c.idy2 + (t.zp + t.zm) * c.idz2 + (double)q * c.ikth);
    if (xe && yi && zi)  // left/right faces, :119-125
        return (xn + c.ax * (t.yp + t.ym) + c.bx * (t.zp + t.zm) + c.Bix * Tinf) * c.Fx;
    if (xi && !yi && zi)  // top face, :127-129
        return (t.ym + c.ay * (t.xp + t.xm) + c.by * (t.zp + t.zm) + c.Biy * Tinf) * c.Fy;
    if (xi && yi && ze)  // back/front faces, :131-137
        return (zn + c.az * (t.xp + t.xm) + c.bz * (t.yp + t.ym) + c.Biz * Tinf) * c.Fz;
    if (xe && yi && ze)  // vertical sides, :143-157
        return (c.av * xn + c.bv * zn + c.cv * (t.yp + t.ym) + 0.5 * Tinf * (c.Bix + c.Biz)) * c.V;
    if (xi && !yi && ze)  // horizontal sides, :161-167
        return (c.ah * t.ym + c.bh * zn + c.ch * (t.xm + t.xp) + 0.5 * Tinf * (c.Biy + c.Biz)) * c.H;
    if (xe && !yi && zi)  // depth sides, :171-177
        return (c.ad * t.ym + c.bd * xn + c.cd * (t.zp + t.zm) + 0.5 * Tinf * (c.Bix + c.Biy)) * c.D;
    // vertices, :181-196
    return (c.AV * xn + c.BV * t.ym + c.CV * zn + Tinf * c.BiD) * c.GV;
}

// z range [z0, z1) of the chunk blockIdx.z
struct ZRange {
    int z0, z1;
};
__device__ __forceinline__ ZRange zchunk(const HeatCoef& c) {
    const int per = (c.nz + gridDim.z - 1) / gridDim.z;
    const int z0 = blockIdx.z * per;
    return {z0, min(c.nz, z0 + per)};
}

// Walks the z range [z0, z1) of the column of a non-Dirichlet node (x, y),
// calling f(z, index, F(T), F(T) - T).  The loads of the next node (its x
// and y neighbours and q, and T two planes up) are issued before this
// node's arithmetic, so that a node waits for no memory latency of its own.
// Neighbours outside the crystal, which jacobi_vals does not use, are read
// at the node itself (a valid address, no branch).
template <typename Fn>
__device__ __forceinline__ void jacobi_column(const HeatCoef& c, const double* __restrict__ T,
                                              const float* __restrict__ Q, int x, int y, ZRange zr, Fn&& f) {
    const size_t sy = (size_t)c.nx, sz = (size_t)c.nx * c.ny;
    size_t i = (size_t)zr.z0 * sz + (size_t)y * sy + x;
    if (zr.z0 >= zr.z1) return;
    const size_t dxm = x > 0 ? 1 : 0, dxp = x < c.nx - 1 ? 1 : 0, dyp = y < c.ny - 1 ? sy : 0;
    struct Plane {
        double xm, xp, ym, yp;
        float q;
    };
    auto plane = [&](size_t j) {
        return Plane{T[j - dxm], T[j + dxp], T[j - sy], T[j + dyp], Q[j]};
    };
    double cur = T[i];
    double nxt = T[zr.z0 + 1 < c.nz ? i + sz : i];
    double prv = T[zr.z0 > 0 ? i - sz : i];
    Plane pc = plane(i);
    for (int z = zr.z0; z < zr.z1; z++, i += sz) {
        const Plane pn = plane(z + 1 < zr.z1 ? i + sz : i);
        const double up2 = T[z + 2 < c.nz ? i + 2 * sz : i];
        const Nb t{cur, pc.xm, pc.xp, pc.ym, pc.yp, prv, nxt};
        const double v = jacobi_vals(c, x, y, z, t, pc.q);
        f(z, i, v, v - cur);
        prv = cur;
        cur = nxt;
        nxt = up2;
        pc = pn;
    }
}

// block (64, 4) reductions with wave shuffles
__device__ __forceinline__ double block_reduce(double v, bool is_max, double* red) {
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off);
        v = is_max ? fmax(v, o) : v + o;
    }
    const int t = threadIdx.x + threadIdx.y * blockDim.x;
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (t == 0) {
        r = red[0];
        const int nw = (blockDim.x * blockDim.y + 63) / 64;
        for (int w = 1; w < nw; w++) r = is_max ? fmax(r, red[w]) : r + red[w];
    }
    return r;
}

// Jacobi sweep Tf = F(Ti) (Tfield::upDate); with `want_norm`, per-block
// partial sums of (Tf - Ti)^2 (Tfield::checkConvergence, :366-376).
// block (64, 4), grid (nx/64, ny/4)
__global__ __launch_bounds__(256) void k_heat_sweep(HeatCoef c, const double* __restrict__ Ti,
                                                    double* __restrict__ Tf, const float* __restrict__ Q,
                                                    double* __restrict__ part_sq, int want_norm) {
    __shared__ double red[4];
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    const ZRange zr = zchunk(c);
    double d2 = 0.0;
    if (x < c.nx && y < c.ny) {
        const bool fixed = y == 0 || (c.oven && (x == 0 || x == c.nx - 1 || y == c.ny - 1));
        if (fixed) {  // oven nodes are copied through
            for (int z = zr.z0; z < zr.z1; z++) {
                const size_t i = ((size_t)z * c.ny + y) * c.nx + x;
                Tf[i] = Ti[i];
            }
        } else {
            jacobi_column(c, Ti, Q, x, y, zr, [&](int, size_t i, double v, double d) {
                Tf[i] = v;
                d2 += d * d;
            });
        }
    }
    if (want_norm) {
        const double s = block_reduce(d2, false, red);
        if (threadIdx.x == 0 && threadIdx.y == 0)
            part_sq[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = s;
    }
}

// FP64 residual of the steady state: delta = F(T) - T (the reference's
// one-sweep change) and the finite-volume residual r = A_ii * delta, written
// scaled by W^-1/2 into the unknown layout u[z][y-1][x-x0] that the
// fast-diagonalisation transforms consume; partial sums of delta^2.
template <typename R>
__global__ __launch_bounds__(256) void k_heat_resid(HeatCoef c, HeatAxes a, const double* __restrict__ T,
                                                    const float* __restrict__ Q, R* __restrict__ u,
                                                    double* __restrict__ part_sq) {
    __shared__ double red[4];
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    const ZRange zr = zchunk(c);
    double d2 = 0.0;
    const int xu = x - c.x0, j = y - 1;
    if (xu >= 0 && xu < c.nxu && j >= 0 && j < c.nyu) {
        const double wx = a.wx[xu], wy = a.wy[j], kx = a.kx[xu], ky = a.ky[j];
        const size_t nyu = (size_t)c.nyu, nxu = (size_t)c.nxu;
        jacobi_column(c, T, Q, x, y, zr, [&](int z, size_t, double, double d) {
            d2 += d * d;
            const double wz = a.wz[z];
            const double aii = kx * wy * wz + wx * ky * wz + wx * wy * a.kz[z];
            u[((size_t)z * nyu + j) * nxu + xu] = (R)(aii * d * rsqrt(wx * wy * wz));
        });
    }
    const double s = block_reduce(d2, false, red);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        part_sq[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = s;
}

// spectral division by lambda_x + lambda_y + lambda_z; grid (ceil(nxu/256),
// nyu, z chunks), each thread walking the z range of its chunk.  FP32 data
// are multiplied by the FP32 reciprocal of the FP64 sum (within an ulp or two
// of the division, the precision of the FP32 transforms around it); FP64
// data are divided.
// workgroup max of m (256 threads, every thread calling), as float bits,
// into part[linear workgroup id] (no atomics: a single-address atomic per
// workgroup serialises ~1e5 workgroups)
__device__ __forceinline__ void block_max_part(float m, unsigned* part) {
    __shared__ float wm[4];
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    if ((t & 63) == 0) wm[t >> 6] = m;
    __syncthreads();
    if (t == 0)
        part[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] =
            __float_as_uint(fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3])));
}

template <typename R>
__global__ __launch_bounds__(256) void k_fd_divide(R* __restrict__ v, const double* __restrict__ lx,
                                                   const double* __restrict__ ly, const double* __restrict__ lz,
                                                   int nxu, int nyu, int nz, unsigned* __restrict__ part) {
    const int x = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    float m = 0.0f;
    if (x < nxu) {
        const int per = (nz + gridDim.z - 1) / gridDim.z, z0 = blockIdx.z * per, z1 = min(nz, z0 + per);
        const double lxy = lx[x] + ly[j];
        for (int z = z0; z < z1; z++) {
            const size_t i = ((size_t)z * nyu + j) * nxu + x;
            if constexpr (sizeof(R) == 4) {
                const float r = v[i] * __builtin_amdgcn_rcpf((float)(lxy + lz[z]));
                v[i] = r;
                m = fmaxf(m, fabsf(r));
            } else {
                v[i] = v[i] / (lxy + lz[z]);
            }
        }
    }
    if (part) block_max_part(m, part);
}

// Even/odd fold of the symmetric axes of u[nz][nyu][nxu]: along a folded
// axis of even length n = 2m, entry i < m becomes (a_i + a_{n-1-i})/sqrt2 and
// entry m + i becomes (a_i - a_{n-1-i})/sqrt2.  The fold is orthogonal; in it
// the axis operator is block diagonal (even and odd eigenvectors), so each
// transform along that axis is two m x m products instead of one n x n.
// fx / fz: fold x / z.  grid (ceil(hx/64), ceil(nyu/4), hz), block (64, 4),
// with hx, hz the folded half lengths (the full length on an axis not folded).
constexpr double kRsqrt2 = 0.70710678118654752440;

template <typename R>
__global__ __launch_bounds__(256) void k_fd_fold(const R* __restrict__ in, R* __restrict__ out, int nxu, int nyu,
                                                 int nz, int fx, int fz, unsigned* __restrict__ part) {
    const int hx = fx ? nxu / 2 : nxu, hz = fz ? nz / 2 : nz;
    const int xu = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    float m = 0.0f;
    if (xu < hx && j < nyu && z < hz) {
        const int xm = nxu - 1 - xu, zm = nz - 1 - z;
        auto at = [&](int zi, int xi) -> size_t { return ((size_t)zi * nyu + j) * nxu + xi; };
        auto put = [&](size_t i, double v) {
            const R r = (R)v;
            out[i] = r;
            m = fmaxf(m, fabsf((float)r));
        };
        if (fx && fz) {
            const double a = in[at(z, xu)], b = in[at(z, xm)], c = in[at(zm, xu)], d = in[at(zm, xm)];
            const double ez = (a + b) * kRsqrt2, oz = (a - b) * kRsqrt2, ezm = (c + d) * kRsqrt2,
                         ozm = (c - d) * kRsqrt2;
            put(at(z, xu), (ez + ezm) * kRsqrt2);
            put(at(hz + z, xu), (ez - ezm) * kRsqrt2);
            put(at(z, hx + xu), (oz + ozm) * kRsqrt2);
            put(at(hz + z, hx + xu), (oz - ozm) * kRsqrt2);
        } else if (fx) {
            const double a = in[at(z, xu)], b = in[at(z, xm)];
            put(at(z, xu), (a + b) * kRsqrt2);
            put(at(z, hx + xu), (a - b) * kRsqrt2);
        } else if (fz) {
            const double a = in[at(z, xu)], c = in[at(zm, xu)];
            put(at(z, xu), (a + c) * kRsqrt2);
            put(at(hz + z, xu), (a - c) * kRsqrt2);
        } else {
            put(at(z, xu), in[at(z, xu)]);
        }
    }
    if (part) block_max_part(m, part);
}

// T += W^-1/2 u (the correction), u unfolded first along the folded axes
// (the inverse of k_fd_fold); partial max |correction|.  A thread owns the
// up to four nodes (xu | nxu-1-xu) x (z | nz-1-z) that one folded entry set
// maps to.  grid: x blocks over the half length hx, y over the node rows, z
// chunks over hz.
template <typename R>
__global__ __launch_bounds__(256) void k_fd_update(HeatCoef c, HeatAxes a, const R* __restrict__ u,
                                                   double* __restrict__ T, double* __restrict__ part_max, int fx,
                                                   int fz) {
    __shared__ double red[4];
    const int nxu = c.nxu, nyu = c.nyu, nz = c.nz;
    const int hx = fx ? nxu / 2 : nxu, hz = fz ? nz / 2 : nz;
    const int xu = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, j = y - 1;
    const int per = (hz + gridDim.z - 1) / gridDim.z, z0 = blockIdx.z * per, z1 = min(hz, z0 + per);
    double m = 0.0;
    if (xu < hx && j >= 0 && j < nyu) {
        const int xm = nxu - 1 - xu;
        const double wy = a.wy[j];
        auto at = [&](int zi, int xi) -> size_t { return ((size_t)zi * nyu + j) * nxu + xi; };
        auto apply = [&](int zi, int xi, double val) {
            const double corr = val * rsqrt(a.wx[xi] * wy * a.wz[zi]);
            T[((size_t)zi * c.ny + y) * c.nx + xi + c.x0] += corr;
            m = fmax(m, fabs(corr));
        };
        for (int z = z0; z < z1; z++) {
            const int zm = nz - 1 - z;
            if (fx && fz) {
                const double ee = u[at(z, xu)], eo = u[at(hz + z, xu)], oe = u[at(z, hx + xu)],
                             oo = u[at(hz + z, hx + xu)];
                const double az = (ee + eo) * kRsqrt2, azm = (ee - eo) * kRsqrt2, bz = (oe + oo) * kRsqrt2,
                             bzm = (oe - oo) * kRsqrt2;
                apply(z, xu, (az + bz) * kRsqrt2);
                apply(z, xm, (az - bz) * kRsqrt2);
                apply(zm, xu, (azm + bzm) * kRsqrt2);
                apply(zm, xm, (azm - bzm) * kRsqrt2);
            } else if (fx) {
                const double e = u[at(z, xu)], o = u[at(z, hx + xu)];
                apply(z, xu, (e + o) * kRsqrt2);
                apply(z, xm, (e - o) * kRsqrt2);
            } else if (fz) {
                const double e = u[at(z, xu)], o = u[at(hz + z, xu)];
                apply(z, xu, (e + o) * kRsqrt2);
                apply(zm, xu, (e - o) * kRsqrt2);
            } else {
                apply(z, xu, (double)u[at(z, xu)]);
            }
        }
    }
    const double s = block_reduce(m, true, red);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        part_max[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_reduce(const double* __restrict__ part, int n, int op,
                                                double* __restrict__ out) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc = op ? fmax(acc, part[i]) : acc + part[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = op ? fmax(red[threadIdx.x], red[threadIdx.x + s])
                                                   : red[threadIdx.x] + red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// T = T_C everywhere, the oven nodes = T_peltier (Tfield::setTemperature /
// setBottomOvens, src/headers/Tfield.h:28-61; setOvenSurrounded, :63-90,
// :333-348); Q = 0 (setInitialQ)
__global__ void k_heat_init(double* __restrict__ T, float* __restrict__ Q, int nx, int ny, int nz, double Tc,
                            double Tpel, int oven) {
    const long n = (long)nx * ny * nz;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)blockDim.x * gridDim.x) {
        const int x = (int)(i % nx);
        const int y = (int)((i / nx) % ny);
        const bool fixed = y == 0 || (oven && (x == 0 || x == nx - 1 || y == ny - 1));
        T[i] = fixed ? Tpel : Tc;
        Q[i] = 0.0f;
    }
}

// dk(T) per voxel in FP64 (PhaseMatching::setDKFromTemperature,
// src/headers/PhaseMatching.h:47-65) and (n_p(T), n_s(T)) per voxel for the
// optics (the n(lambda, T) of src/headers/Solver.h:48,50 and
// src/headers/Tfield.h:270), from one FP64 Sellmeier pair per voxel
// Per voxel of a heat column: n_p(T), n_s(T) and dk(T) (src/headers/PhaseMatching.h:59,80),
// and the base of the running dk integral per heat slice: the optical
// slices s in [s0(h), s0(h) + cnt(h)) see  DKint_s = sum_{j<=s} dk[h(j)]
// = dkbase[h] + s dk[h]  (src/headers/PhaseMatching.h:27-44 summed in O(1)).
// One thread per (x, y) column walks the heat slices, so dk is written and
// summed in one pass.
__global__ __launch_bounds__(256) void k_heat_dk(const double* __restrict__ T, double* __restrict__ dk,
                                                 float2* __restrict__ nn, double* __restrict__ base, int n2, int nzh,
                                                 int nz, int crystal, double lp, double Lambda, double ea, double eb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n2) return;
    const double ls = 0.5 * lp;
    double cum = 0.0;
    // T of the next two slices in flight during this slice's arithmetic
    // (past the last slice, the last one is read again: no branch)
    double t1 = T[i], t2 = T[(size_t)min(1, nzh - 1) * n2 + i];
    for (int h = 0; h < nzh; h++) {
        const size_t v = (size_t)h * n2 + i;
        const double t = t1;
        t1 = t2;
        t2 = T[(size_t)min(h + 2, nzh - 1) * n2 + i];
        const double np = refr_index<double>(crystal, lp, t);
        const double ns = refr_index<double>(crystal, ls, t);
        const double d = t - 25.0;
        const double k = 2.0 * kPi * (2.0 * np / lp - ns / ls + 1.0 / (Lambda * (1.0 + ea * d + eb * d * d)));
        dk[v] = k;
        nn[v] = make_float2((float)np, (float)ns);
        const long long s0 = ((long long)h * nz + nzh - 1) / nzh;
        const long long s1 = ((long long)(h + 1) * nz + nzh - 1) / nzh;
        base[v] = cum - (double)(s0 - 1) * k;
        cum += (double)(s1 - s0) * k;
    }
}

__global__ __launch_bounds__(256) void k_max(const double* __restrict__ a, long n, double* __restrict__ part) {
    __shared__ double red[256];
    double m = -1e300;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += 256L * gridDim.x) m = fmax(m, a[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// ------------------------------------------------------------------
// adjoint of the steady state (THERMAL gradients)
// ------------------------------------------------------------------
__device__ __forceinline__ bool heat_fixed(const HeatCoef& c, int x, int y) {
    return y == 0 || (c.oven && (x == 0 || x == c.nx - 1 || y == c.ny - 1));
}

// u = W^-1/2 mu on the unknown block, in the layout k_heat_resid writes:
// the right-hand side of the transposed (= the same, symmetric) solve.
template <typename R>
__global__ __launch_bounds__(256) void k_adj_heat_in(HeatCoef c, HeatAxes a, const double* __restrict__ mu,
                                                     R* __restrict__ u) {
    const int xu = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    if (xu >= c.nxu || j >= c.nyu) return;
    const int x = xu + c.x0, y = j + 1;
    const ZRange zr = zchunk(c);
    const double wxy = a.wx[xu] * a.wy[j];
    for (int z = zr.z0; z < zr.z1; z++)
        u[((size_t)z * c.nyu + j) * c.nxu + xu] =
            (R)(mu[((size_t)z * c.ny + y) * c.nx + x] * rsqrt(wxy * a.wz[z]));
}

// After nu = A^-1 mu (A the finite-volume operator of the steady state):
//   w_v = dr_v/dQ_v nu_v = A_vv L2 / k nu_v at the inner nodes that take Q
//   (the boundary nodes omit it, src/headers/Tfield.h:119-196), 0 elsewhere;
//   partial sums of  sum_{oven nodes} mu_v + sum_{unknowns} nu_v dr_v/dT_pel,
//   the derivative of <mu, T> with the oven temperature (dr_v/dT_pel: the
//   coupling of node v to its oven neighbours, A_vv (F(T) - T) at T = 1 on
//   the oven nodes, 0 elsewhere, without Q and T_inf).
__global__ __launch_bounds__(256) void k_adj_heat_out(HeatCoef c, HeatAxes a, const double* __restrict__ mu,
                                                      const double* __restrict__ nu, double* __restrict__ w,
                                                      double* __restrict__ part) {
    __shared__ double red[4];
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    const ZRange zr = zchunk(c);
    double acc = 0.0;
    if (x < c.nx && y < c.ny) {
        const bool fixed = heat_fixed(c, x, y);
        const int xu = x - c.x0, j = y - 1;
        HeatCoef c0 = c;
        c0.Tinf = 0.0;
        Nb t;
        t.c = 0.0;
        t.xm = x > 0 && heat_fixed(c, x - 1, y) ? 1.0 : 0.0;
        t.xp = x < c.nx - 1 && heat_fixed(c, x + 1, y) ? 1.0 : 0.0;
        t.ym = y > 0 && heat_fixed(c, x, y - 1) ? 1.0 : 0.0;
        t.yp = y < c.ny - 1 && heat_fixed(c, x, y + 1) ? 1.0 : 0.0;
        t.zm = 0.0;
        t.zp = 0.0;
        const bool inner_xy = x > 0 && x < c.nx - 1 && y > 0 && y < c.ny - 1;
        for (int z = zr.z0; z < zr.z1; z++) {
            const size_t i = ((size_t)z * c.ny + y) * c.nx + x;
            if (fixed) {
                acc += mu[i];
                w[i] = 0.0;
                continue;
            }
            const double wz = a.wz[z];
            const double aii = a.kx[xu] * a.wy[j] * wz + a.wx[xu] * a.ky[j] * wz + a.wx[xu] * a.wy[j] * a.kz[z];
            const bool takes_q = inner_xy && z > 0 && z < c.nz - 1;
            w[i] = takes_q ? aii * c.L2 * c.ikth * nu[i] : 0.0;
            acc += nu[i] * aii * jacobi_vals(c0, x, y, z, t, 0.0f);
        }
    }
    const double s = block_reduce(acc, false, red);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        part[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = s;
}

// dL/dT per voxel from the adjoint row passes' accumulators (one thread per
// (x, y) column, walking the heat slabs from the last):
//   dL/ddk[h] = gA[h] + cnt(h) sum_{h' > h} gC[h']
//   mu[h] = gnp[h] n_p'(T) + gns[h] n_s'(T) + dL/ddk[h] dk'(T)  (into gnp),
// n' and dk' by central differences of the FP64 closed forms (step 1e-3 K);
// plus `tc` at the voxel whose temperature sets the pump, the propagators and
// the power normalisation (slab 0, grid centre).
__global__ __launch_bounds__(256) void k_adj_heat_grad(const double* __restrict__ T, double* __restrict__ gnp,
                                                       const double* __restrict__ gns, const double* __restrict__ gA,
                                                       const double* __restrict__ gC, int nx, int ny, int nzh, int nz,
                                                       int crystal, double lp, double Lambda, double ea, double eb,
                                                       double tc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n2 = nx * ny;
    if (i >= n2) return;
    const double ls = 0.5 * lp, hT = 1e-3;
    double suffix = 0.0;
    for (int h = nzh - 1; h >= 0; h--) {
        const size_t v = (size_t)h * n2 + i;
        const double t = T[v];
        const double npd = (refr_index<double>(crystal, lp, t + hT) - refr_index<double>(crystal, lp, t - hT)) / (2 * hT);
        const double nsd = (refr_index<double>(crystal, ls, t + hT) - refr_index<double>(crystal, ls, t - hT)) / (2 * hT);
        const double dkd = (delta_k(crystal, lp, Lambda, ea, eb, t + hT) - delta_k(crystal, lp, Lambda, ea, eb, t - hT)) /
                           (2 * hT);
        const long long s0 = ((long long)h * nz + nzh - 1) / nzh;
        const long long s1 = ((long long)(h + 1) * nz + nzh - 1) / nzh;
        const double ddk = gA[v] + (double)(s1 - s0) * suffix;
        suffix += gC[v];
        double m = gnp[v] * npd + gns[v] * nsd + ddk * dkd;
        if (h == 0 && i == (ny / 2) * nx + nx / 2) m += tc;
        gnp[v] = m;
    }
}

// ------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------
// x-y plane tiles of 64 x 4 columns, and z chunks of at least 48 slices (at
// most 8 chunks)
static dim3 plane_grid(const HeatCoef& c) {
    const int zc = std::max(1, std::min(8, c.nz / 48));
    return dim3((c.nx + 63) / 64, (c.ny + 3) / 4, zc);
}
int heat_grid_blocks(const HeatCoef& c) {
    const dim3 g = plane_grid(c);
    return (int)(g.x * g.y * g.z);
}

hipError_t launch_heat_sweep(hipStream_t st, const HeatCoef& c, const double* Ti, double* Tf, const float* Q,
                             double* part_sq, bool want_norm) {
    hipLaunchKernelGGL(k_heat_sweep, plane_grid(c), dim3(64, 4), 0, st, c, Ti, Tf, Q, part_sq, want_norm ? 1 : 0);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_heat_resid(hipStream_t st, const HeatCoef& c, const HeatAxes& a, const double* T, const float* Q,
                             R* u, double* part_sq) {
    hipLaunchKernelGGL((k_heat_resid<R>), plane_grid(c), dim3(64, 4), 0, st, c, a, T, Q, u, part_sq);
    return hipGetLastError();
}
template hipError_t launch_heat_resid<float>(hipStream_t, const HeatCoef&, const HeatAxes&, const double*,
                                             const float*, float*, double*);
template hipError_t launch_heat_resid<double>(hipStream_t, const HeatCoef&, const HeatAxes&, const double*,
                                              const float*, double*, double*);

static dim3 divide_grid(int nxu, int nyu, int nz) {
    return dim3((nxu + 255) / 256, nyu, std::max(1, std::min(nz, 16)));
}
static dim3 fold_grid(int nxu, int nyu, int nz, bool fx, bool fz) {
    const int hx = fx ? nxu / 2 : nxu, hz = fz ? nz / 2 : nz;
    return dim3((hx + 63) / 64, (nyu + 3) / 4, hz);
}
static long blocks(dim3 g) { return (long)g.x * g.y * g.z; }

long heat_max_partials(int nxu, int nyu, int nz) {
    return std::max({blocks(divide_grid(nxu, nyu, nz)), blocks(fold_grid(nxu, nyu, nz, false, false)),
                     heat_gemm_max_blocks(nxu, nyu, nz), 1024L});
}

template <typename R>
hipError_t launch_fd_divide(hipStream_t st, R* v, const HeatAxes& a, int nxu, int nyu, int nz, unsigned* part,
                            unsigned* slot) {
    const dim3 grid = divide_grid(nxu, nyu, nz);
    hipLaunchKernelGGL((k_fd_divide<R>), grid, dim3(256), 0, st, v, a.lx, a.ly, a.lz, nxu, nyu, nz, part);
    if (part) return launch_max_parts(st, part, blocks(grid), slot);
    return hipGetLastError();
}
template hipError_t launch_fd_divide<float>(hipStream_t, float*, const HeatAxes&, int, int, int, unsigned*,
                                            unsigned*);
template hipError_t launch_fd_divide<double>(hipStream_t, double*, const HeatAxes&, int, int, int, unsigned*,
                                             unsigned*);

template <typename R>
hipError_t launch_fd_fold(hipStream_t st, const R* in, R* out, int nxu, int nyu, int nz, bool fx, bool fz,
                          unsigned* part, unsigned* slot) {
    const dim3 grid = fold_grid(nxu, nyu, nz, fx, fz);
    hipLaunchKernelGGL((k_fd_fold<R>), grid, dim3(64, 4), 0, st, in, out, nxu, nyu, nz, fx ? 1 : 0, fz ? 1 : 0, part);
    if (part) return launch_max_parts(st, part, blocks(grid), slot);
    return hipGetLastError();
}
template hipError_t launch_fd_fold<float>(hipStream_t, const float*, float*, int, int, int, bool, bool, unsigned*,
                                          unsigned*);
template hipError_t launch_fd_fold<double>(hipStream_t, const double*, double*, int, int, int, bool, bool,
                                           unsigned*, unsigned*);

// The grid keeps the block count of plane_grid (the size of the partials
// array) at most: the z chunks take the x blocks a fold saves.
template <typename R>
hipError_t launch_fd_update(hipStream_t st, const HeatCoef& c, const HeatAxes& a, const R* u, double* T,
                            double* part_max, bool fx, bool fz, int* nblocks) {
    const dim3 pg = plane_grid(c);
    const int hx = fx ? c.nxu / 2 : c.nxu, hz = fz ? c.nz / 2 : c.nz;
    const int bx = (hx + 63) / 64;
    const int zc = std::max(1, std::min(hz, (int)(pg.x * pg.z) / bx));
    const dim3 grid(bx, pg.y, zc);
    *nblocks = (int)(grid.x * grid.y * grid.z);
    hipLaunchKernelGGL((k_fd_update<R>), grid, dim3(64, 4), 0, st, c, a, u, T, part_max, fx ? 1 : 0, fz ? 1 : 0);
    return hipGetLastError();
}
template hipError_t launch_fd_update<float>(hipStream_t, const HeatCoef&, const HeatAxes&, const float*, double*,
                                            double*, bool, bool, int*);
template hipError_t launch_fd_update<double>(hipStream_t, const HeatCoef&, const HeatAxes&, const double*, double*,
                                             double*, bool, bool, int*);

hipError_t launch_reduce(hipStream_t st, const double* part, int n, bool is_max, double* out) {
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, st, part, n, is_max ? 1 : 0, out);
    return hipGetLastError();
}

hipError_t launch_heat_init(hipStream_t st, double* T, float* Q, int nx, int ny, int nz, double Tc, double Tpel,
                            bool oven) {
    hipLaunchKernelGGL(k_heat_init, dim3(2048), dim3(256), 0, st, T, Q, nx, ny, nz, Tc, Tpel, oven ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_heat_dk(hipStream_t st, const double* T, double* dk, float2* nn, double* base, int n2, int nzh,
                          int nz, int crystal, double lp, double Lambda, double ea, double eb) {
    hipLaunchKernelGGL(k_heat_dk, dim3((n2 + 255) / 256), dim3(256), 0, st, T, dk, nn, base, n2, nzh, nz, crystal, lp,
                       Lambda, ea, eb);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_adj_heat_in(hipStream_t st, const HeatCoef& c, const HeatAxes& a, const double* mu, R* u) {
    const dim3 pg = plane_grid(c);
    hipLaunchKernelGGL((k_adj_heat_in<R>), dim3((c.nxu + 63) / 64, (c.nyu + 3) / 4, pg.z), dim3(64, 4), 0, st, c, a,
                       mu, u);
    return hipGetLastError();
}
template hipError_t launch_adj_heat_in<float>(hipStream_t, const HeatCoef&, const HeatAxes&, const double*, float*);
template hipError_t launch_adj_heat_in<double>(hipStream_t, const HeatCoef&, const HeatAxes&, const double*,
                                               double*);

hipError_t launch_adj_heat_out(hipStream_t st, const HeatCoef& c, const HeatAxes& a, const double* mu,
                               const double* nu, double* w, double* part) {
    hipLaunchKernelGGL(k_adj_heat_out, plane_grid(c), dim3(64, 4), 0, st, c, a, mu, nu, w, part);
    return hipGetLastError();
}

hipError_t launch_adj_heat_grad(hipStream_t st, const double* T, double* gnp, const double* gns, const double* gA,
                                const double* gC, int nx, int ny, int nzh, int nz, int crystal, double lp,
                                double Lambda, double ea, double eb, double tc) {
    const int n2 = nx * ny;
    hipLaunchKernelGGL(k_adj_heat_grad, dim3((n2 + 255) / 256), dim3(256), 0, st, T, gnp, gns, gA, gC, nx, ny, nzh, nz,
                       crystal, lp, Lambda, ea, eb, tc);
    return hipGetLastError();
}

hipError_t launch_max(hipStream_t st, const double* a, long n, double* part, int nblocks) {
    hipLaunchKernelGGL(k_max, dim3(nblocks), dim3(256), 0, st, a, n, part);
    return hipGetLastError();
}

}  // namespace shg
