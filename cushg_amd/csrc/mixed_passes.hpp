// mixed_passes.hpp -- device code of the fused SSFM passes (gfx950),
// templated on the storage type S of the fields in device memory and the
// arithmetic type C in registers.
//
// One SSFM slice of the reference (src/headers/Solver.h:267-273) is
//   A <- R(A, z_s);  A <- IFFT2(Prop . FFT2(A)) / (nx ny);  A <- R(A, z_s)
// with R the RK4 call of src/headers/Solver.h:189-264.  Because the 2D FFT
// is separable and R is pointwise, the march is regrouped into two HBM
// passes per slice, each reading and writing both fields once:
//   row pass  (A): row IFFT (end of slice s-1) -> R_end(s-1) -> R_start(s)
//                  -> row FFT (start of slice s)
//   column pass (B): column FFT -> * Prop -> column IFFT
// so a slice costs 2 passes x 2 fields x (read + write) = 4 sizeof(S) bytes
// per grid point.  The 1/(nx ny) scale of cuFFTscale
// (src/headers/Efields.h:15-29) is folded into the propagator table.
//
// Forms (S, C):
//   (float2, float2)   SHG_PREC_FP32_MIXED: FFT butterflies, RK4 and the
//                      phase factor in FP32; twiddle and propagator
//                      products in FP64, rounded once
//   (float2, double2)  SHG_PREC_FP32_STORAGE: everything between the FP32
//                      load and the FP32 store in FP64
//   (double2, double2) SHG_PREC_FP64_FUSED: FP64 fields and arithmetic
// SHG_PREC_FP32_SPLIT runs the (float2, float2) row pass and the
// (float2, double2) column pass.
// The QPM phase dk*z is FP64 in all and reduced to a quarter turn before
// the sin/cos series.
#pragma once

#include "fft.hpp"
#include "kernels.hpp"

namespace shg {

template <typename C> using re_t = decltype(C{}.x);

template <typename C, typename S> __device__ __forceinline__ C from_store(S v) { return mk<C>(v.x, v.y); }

// 32-bit integer hash (Wellons' "lowbias32" finaliser): a counter-based
// random stream keyed by element index, step and store site.
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// FP64 -> FP32 with stochastic rounding: 29 random bits are added below the
// 23-bit FP32 mantissa of the FP64 pattern and the value is truncated, so
// it rounds up in magnitude with probability equal to the dropped fraction
// of an FP32 ulp.  Round-to-nearest would be biased for a component whose
// change per slice is below half an ulp (it would be stored unchanged slice
// after slice); stochastic rounding is unbiased for every value.
__device__ __forceinline__ float sr_f32(double v, uint32_t rnd) {
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    b += (unsigned long long)(rnd >> 3);
    b &= ~((1ull << 29) - 1);
    return (float)__longlong_as_double((long long)b);
}

// Store keys: site (0 row-pass state, 1 column-pass state, 2 saved output,
// 3 trace), field (0 pump, 1 SH) and step g.
__device__ __forceinline__ uint32_t store_key(int site, int field, int g) {
    return mix32(0x9e3779b9U * (uint32_t)(g + 1) ^ (uint32_t)(site << 28 | field << 27));
}
// Store of an arithmetic value C in the storage type S: exact when the two
// are the same; FP64 -> FP32 by stochastic rounding with one hash per
// complex value (16 random bits per component, placed at the top of the 29
// dropped bits: the 13 bits below them truncate, a bias of at most 2^-16
// ulp toward zero per store).
template <typename S, typename C> struct Store {
    static __device__ __forceinline__ S to(C v, size_t, uint32_t) { return v; }
};
template <> struct Store<float2, double2> {
    static __device__ __forceinline__ float2 to(double2 v, size_t idx, uint32_t key) {
        const uint32_t r = mix32(key ^ (uint32_t)idx);
        return make_float2(sr_f32(v.x, r & 0xffff0000U), sr_f32(v.y, r << 16));
    }
};
template <typename S, typename C> __device__ __forceinline__ S to_store(C v, size_t idx, uint32_t key) {
    return Store<S, C>::to(v, idx, key);
}

template <typename C> __device__ __forceinline__ C expi(double ph);
template <> __device__ __forceinline__ float2 expi<float2>(double ph) { return expi_fast(ph); }
template <> __device__ __forceinline__ double2 expi<double2>(double ph) { return expi_fast64(ph); }

// 1/n for an FP32 index n: hardware reciprocal, refined by Newton steps to
// the precision of R
template <typename R> __device__ __forceinline__ R recip(float n);
template <> __device__ __forceinline__ float recip<float>(float n) {
    const float r = __builtin_amdgcn_rcpf(n);
    return r * __builtin_fmaf(-n, r, 2.0f);
}
template <> __device__ __forceinline__ double recip<double>(float n) {
    const double nd = n;
    double r = __builtin_amdgcn_rcpf(n);
    r = r * (2.0 - nd * r);
    return r * (2.0 - nd * r);
}

// dA/dz of src/headers/Solver.h:57-80 (and :33-55 for THERMAL):
//   dP = i kp S conj(P) conj(e) - alpha_p/2 P
//   dS = i ks P^2 e - (alpha_s/2 + bfac/2 |S|^2) S
// Both coupling terms go through u = P e: S conj(P) conj(e) = S conj(u) and
// P^2 e = P u (three complex products instead of four).
template <typename C, typename R = re_t<C>>
__device__ __forceinline__ void dadz(C P, C S, C e, R kp, R ks, R aph, R ash, R bfac, C& dP, C& dS) {
    const C u = cmul(P, e);
    const C spc = cmul(S, conjc(u));
    dP = mk<C>(-kp * spc.y - aph * P.x, kp * spc.x - aph * P.y);
    const C ppe = cmul(P, u);
    const R g = ash + R(0.5) * bfac * abs2(S);
    dS = mk<C>(-ks * ppe.y - g * S.x, ks * ppe.x - g * S.y);
}

// Scheme constants of one R call in the arithmetic type.
template <typename R> struct RkConst {
    R c1, c2, c3, w, aph, ash;
};

// RK4 call of src/headers/Solver.h:231-257 on one point; stage inputs
// A + c*k with c = (c1, c2, c3), z offsets e[0], e[1] (stages 2 and 3), e[2].
template <typename C, typename R = re_t<C>>
__device__ __forceinline__ void rk4_point(C& P, C& S, const C (&e)[3], R kp, R ks, R bfac, const RkConst<R>& k) {
    // A += w (k1 + 2 k2 + 2 k3 + k4), the slope sum accumulated as each
    // stage ends so that only one stage's slopes are live at a time
    C kp_, ks_, sp, ss;
    dadz(P, S, e[0], kp, ks, k.aph, k.ash, bfac, kp_, ks_);
    sp = kp_;
    ss = ks_;
    dadz(axpy(k.c1, kp_, P), axpy(k.c1, ks_, S), e[1], kp, ks, k.aph, k.ash, bfac, kp_, ks_);
    sp = axpy(R(2), kp_, sp);
    ss = axpy(R(2), ks_, ss);
    dadz(axpy(k.c2, kp_, P), axpy(k.c2, ks_, S), e[1], kp, ks, k.aph, k.ash, bfac, kp_, ks_);
    sp = axpy(R(2), kp_, sp);
    ss = axpy(R(2), ks_, ss);
    dadz(axpy(k.c3, kp_, P), axpy(k.c3, ks_, S), e[2], kp, ks, k.aph, k.ash, bfac, kp_, ks_);
    P = axpy(k.w, sp + kp_, P);
    S = axpy(k.w, ss + ks_, S);
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

// Nonlinear work of the row pass at global step g on the points of row y:
// R call 2 of slice g-1 (if g > 0; with THERMAL the heat source of slice
// g-1 is deposited after it; the state is then saved and traced as asked),
// then R call 1 of slice g (if g < nz).  Both calls run point by point, so
// a point's heat-voxel values are loaded once and the second call's phase
// factor is the first's times a small-angle factor exp(i (phi_g - phi_{g-1})).
// The constructor holds what is uniform over the row; point() does one
// point (column x, element index i of the field).
template <bool THERMAL, typename S, typename C> struct NlRow {
    using R = re_t<C>;
    const MixedInst& I;
    int g, nz;
    bool c1, c0, save, trace;
    RkConst<R> k;
    uint32_t kop, kos;
    // non-THERMAL: the two calls' phase factors and couplings
    C e1[3], e0[3];
    R kp, ks, bfac;
    // THERMAL: heat-voxel rows of the two calls and the phase scales
    size_t row1 = 0, row0 = 0;
    int h1 = 0, h0 = 0;
    double zsc1 = 0, zsc0 = 0, o1 = 0, o0 = 0, pol1 = 0, pol0 = 0;
    bool first = false;
    R qw = 0, kp0 = 0, ks0 = 0, bf0 = 0, hce = 0;

    __device__ __forceinline__ NlRow(const MixedInst& I_, int g_, int y, bool trace_) : I(I_), g(g_), nz(I_.nz) {
        c1 = g > 0;
        c0 = g < nz;
        save = c1 && g - 1 == nz - 2;
        trace = trace_;
        k = RkConst<R>{(R)I.c1, (R)I.c2, (R)I.c3, (R)I.w, (R)I.ap_h, (R)I.as_h};
        // saved output and trace share a key: a traced slice nz-1 is that state, bit for bit
        kop = store_key(2, 0, g);
        kos = store_key(2, 1, g);
        if constexpr (!THERMAL) {
            // phase table: exp(i phi) of (slice, call, offset), FP64
            const double2* t1 = I.phase + (size_t)(max(g - 1, 0) * 2 + 1) * 3;
            const double2* t0 = I.phase + (size_t)(min(g, nz - 1) * 2) * 3;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                e1[j] = c1 ? mk<C>((R)t1[j].x, (R)t1[j].y) : C{};
                e0[j] = c0 ? mk<C>((R)t0[j].x, (R)t0[j].y) : C{};
            }
            kp = (R)I.kp;
            ks = (R)I.ks;
            bfac = (R)I.bfac;
        } else {
            const size_t plane = (size_t)I.nx * I.ny;
            // heat slices of the two calls (clamped to a valid slice when a call does not run)
            h1 = heat_slice(c1 ? g - 1 : g, nz, I.nzh);
            h0 = heat_slice(c0 ? g : g - 1, nz, I.nzh);
            row1 = (size_t)h1 * plane + (size_t)y * I.nx;
            row0 = (size_t)h0 * plane + (size_t)y * I.nx;
            // phase DKint (s + off) dz / (s + 1) of src/headers/Solver.h:48,50,
            // with DKint the running dk integral (src/headers/PhaseMatching.h:27-44)
            zsc1 = I.dz / (double)g;
            zsc0 = I.dz / ((double)g + 1.0);
            o1 = (double)(g - 1) + I.off[1][0];
            o0 = (double)g + I.off[0][0];
            pol1 = c1 && I.poling ? I.poling[g - 1] : 0.0;
            pol0 = c0 && I.poling ? I.poling[g] : 0.0;
            first = g - 1 == 0 || (c1 && heat_slice(g - 2, nz, I.nzh) != h1);
            qw = c1 ? R(1) / (R)heat_count(h1, nz, I.nzh) : R(0);
            kp0 = (R)I.kp0;
            ks0 = (R)I.ks0;
            bf0 = (R)I.bf0;
            hce = (R)I.hce;
        }
    }

    __device__ __forceinline__ void emit(const C& p, const C& s, size_t i) const {
        if (save) {
            static_cast<S*>(I.outP)[i] = to_store<S>(p, i, kop);
            static_cast<S*>(I.outS)[i] = to_store<S>(s, i, kos);
        }
        if (trace) {
            static_cast<S*>(I.trP)[i] = to_store<S>(p, i, kop);
            static_cast<S*>(I.trS)[i] = to_store<S>(s, i, kos);
        }
    }

    // THERMAL, FP32 arithmetic: the heat-grid loads of a row's next point
    // are issued before this point's arithmetic (nl_ahead), through the
    // global address space, so that they stay in flight while it runs (the
    // vector-memory counter completes them in order).  FP64 arithmetic has
    // no registers to spare for them: each point loads its own, in place.
    static constexpr bool kAhead = THERMAL && sizeof(C) == 8;
    // What a point reads from the heat grid for column x: the voxel of the
    // first call that runs, and the heat source it adds to.
    struct Vox {
        float2 n;       // n_p(T), n_s(T)
        double dkb, dkv;
        float q;        // heat source before this slice's deposit (unused on a slab's first slice)
    };
    // Q is read unconditionally (its value is unused when !c1 or on a
    // slab's first slice): a load on one side of a branch would make the
    // wait at the join wait for every load in flight.
    __device__ __forceinline__ Vox load(int x) const {
        Vox o{};
        if constexpr (THERMAL) {
            const size_t v = (c1 ? row1 : row0) + x;
            o.n = ld_global(I.nn + v);
            o.dkb = ld_global(I.dkbase + v);
            o.dkv = ld_global(I.dk + v);
            o.q = ld_global(I.Q + row1 + x);
        }
        return o;
    }
    // load(x) of the state of step g, row y, without constructing it: both
    // voxel rows are that of heat slice h(max(g - 1, 0)) (at g = 0 the
    // first call does not run and row0 = row1)
    static __device__ __forceinline__ Vox load_at(const MixedInst& I, int g, int y, int x) {
        Vox o{};
        if constexpr (THERMAL) {
            const size_t v = (size_t)heat_slice(max(g - 1, 0), I.nz, I.nzh) * ((size_t)I.nx * I.ny) +
                             (size_t)y * I.nx + x;
            o.n = ld_global(I.nn + v);
            o.dkb = ld_global(I.dkbase + v);
            o.dkv = ld_global(I.dk + v);
            o.q = ld_global(I.Q + v);
        }
        return o;
    }

    // one point, its heat-grid values loaded in place
    __device__ __forceinline__ void point(C& p, C& s, int x, size_t i) const { point_t<false>(p, s, x, i, nullptr); }
    // one point, its heat-grid values loaded ahead by load(x)
    __device__ __forceinline__ void point(C& p, C& s, int x, size_t i, const Vox& vx) const {
        point_t<true>(p, s, x, i, &vx);
    }

    template <bool AHEAD>
    __device__ __forceinline__ void point_t(C& p, C& s, int x, size_t i, const Vox* vx) const {
        if constexpr (!THERMAL) {
            if (c1) {
                rk4_point(p, s, e1, kp, ks, bfac, k);
                emit(p, s, i);
            }
            if (c0) rk4_point(p, s, e0, kp, ks, bfac, k);
        } else {
            const size_t v = (c1 ? row1 : row0) + x;
            float2 n;  // n_p(T), n_s(T) of the voxel
            double dkb, dkv;
            if constexpr (AHEAD) {
                n = vx->n;
                dkb = vx->dkb;
                dkv = vx->dkv;
            } else {
                n = I.nn[v];
                dkb = I.dkbase[v];
                dkv = I.dk[v];
            }
            R np = n.x, ns = n.y;
            R kpv = kp0 * recip<R>(n.x), ksv = ks0 * recip<R>(n.y), bfv = bf0 * ns;
            C e0v = C{};
            double ph = 0.0;
            if (c1) {
                const double dki = dkb + (double)(g - 1) * dkv;
                ph = dki * o1 * zsc1 + pol1;
                // the three offsets are 0.25 dz apart in both schemes:
                // e(off + 1/4) = e(off) exp(i DKint dz / (4 (s + 1)))
                e0v = expi<C>(ph);
                const C ed = expi_small<C>(0.25 * dki * zsc1);
                const C e1v = cmul(e0v, ed);
                const C e[3] = {e0v, e1v, cmul(e1v, ed)};
                rk4_point(p, s, e, kpv, ksv, bfv, k);
                // Q = a_p I_p + a_s I_s + b_s I_s^2 -- src/headers/Tfield.h:260-276
                const R Ip = hce * np * abs2(p);
                const R Is = hce * ns * abs2(s);
                const R qv = ((R)I.alpha_p * Ip + (R)I.alpha_s * Is + (R)I.beta_s * Is * Is) * qw;
                float* Qp = I.Q + row1 + x;
                if constexpr (AHEAD) st_global(Qp, first ? (float)qv : (float)((R)vx->q + qv));
                else *Qp = first ? (float)qv : (float)((R)*Qp + qv);
                emit(p, s, i);
            }
            if (c0) {
                if (c1 && h0 != h1) {  // the call falls in the next heat slab
                    if constexpr (AHEAD) {
                        n = ld_global(I.nn + row0 + x);
                        dkb = ld_global(I.dkbase + row0 + x);
                        dkv = ld_global(I.dk + row0 + x);
                    } else {
                        n = I.nn[row0 + x];
                        dkb = I.dkbase[row0 + x];
                        dkv = I.dk[row0 + x];
                    }
                    np = n.x;
                    ns = n.y;
                    kpv = kp0 * recip<R>(n.x);
                    ksv = ks0 * recip<R>(n.y);
                    bfv = bf0 * ns;
                }
                const double dki = dkb + (double)g * dkv;
                const double ph0 = dki * o0 * zsc0 + pol0;
                e0v = c1 ? cmul(e0v, expi_small<C>(ph0 - ph)) : expi<C>(ph0);
                const C ed = expi_small<C>(0.25 * dki * zsc0);
                const C e1v = cmul(e0v, ed);
                const C e[3] = {e0v, e1v, cmul(e1v, ed)};
                rk4_point(p, s, e, kpv, ksv, bfv, k);
            }
        }
    }
};

// The nonlinear work of a thread's 8 points x = tp + q*TPR of one row.
// Loads ahead (Nl::kAhead): cur holds the first point's heat-grid values,
// and each point's loads are issued before the previous point's arithmetic.
template <int NX, typename Nl, typename C>
__device__ __forceinline__ void nl_ahead(C (&p)[8], C (&s)[8], const Nl& nl, int tp, size_t base,
                                         typename Nl::Vox cur) {
    constexpr int TPR = NX / 8;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        typename Nl::Vox nxt{};
        if (q < 7) nxt = nl.load(tp + (q + 1) * TPR);
        nl.point(p[q], s[q], tp + q * TPR, base + q * TPR, cur);
        cur = nxt;
    }
}
// Loads in place.
template <int NX, bool THERMAL, typename S, typename C>
__device__ __forceinline__ void nl_inplace(C (&p)[8], C (&s)[8], const MixedInst& I, int g, int y, int tp,
                                           size_t base, bool trace) {
    constexpr int TPR = NX / 8;
    const NlRow<THERMAL, S, C> nl(I, g, y, trace);
#pragma unroll
    for (int q = 0; q < 8; q++) nl.point(p[q], s[q], tp + q * TPR, base + q * TPR);
}

// Row pass at global step g of every instance in the launch group.
//   g == 0       : R_start(0) -> row FFT
//   0 < g < nz   : row IFFT -> R_end(g-1) [-> save state if g-1 == nz-2] -> R_start(g) -> row FFT
//   g == nz      : row IFFT -> R_end(nz-1) -> spatial fields stored
// grid (ceil(max_ny/ROWS), 1, n_instances), block (NX/8, ROWS)
// Both fields of a row are transformed together (shared barriers, two LDS
// buffers per row) in FP32; in FP64 one after the other through one buffer,
// which halves the LDS footprint (occupancy is LDS-bound there).
template <typename C> struct RowLds { static constexpr int BUFS = sizeof(C) == 16 ? 1 : 2; };

template <typename C, int NX, bool INV>
__device__ __forceinline__ void fft_pair(C (&f)[2][8], C* my, int ld, const double2* __restrict__ tw, int tp) {
    if constexpr (RowLds<C>::BUFS == 1) {
        fft_dist<C, NX, INV>(f[0], my, tw, tp);
        fft_dist<C, NX, INV>(f[1], my, tw, tp);
    } else {
        fft_multi<C, NX, INV, 2>(f, my, ld, tw, tp);
    }
}

// FP64 arithmetic: ask for 3 waves per SIMD (<= 168 VGPRs)
template <typename C> struct RowWaves { static constexpr int W = sizeof(C) == 16 ? 3 : 1; };

template <int NX, bool THERMAL, int ROWS, typename S, typename C>
__global__ __launch_bounds__(NX / 8 * ROWS) __attribute__((amdgpu_waves_per_eu(RowWaves<C>::W)))
void k_rowpass(const MixedInst* __restrict__ insts, Step gs, const double2* __restrict__ tw) {
    constexpr int TPR = NX / 8;
    constexpr int LD = FftShape<NX>::LDS;
    __shared__ C lds[RowLds<C>::BUFS * ROWS * LD];
    const MixedInst& I = insts[blockIdx.z];
    // the descriptor fields and the step are read before any check, so that
    // the field loads wait for one round of scalar loads, not one per check
    const int nz = I.nz, ny = I.ny;
    void* const Pv = I.P;
    void* const Sv = I.S;
    const int g = gs.get();
    asm volatile("" ::"s"(Pv), "s"(Sv));  // keeps the pointer loads ahead of the checks
    if (g > nz) return;
    const int y = blockIdx.x * ROWS + threadIdx.y;
    if (blockIdx.x * ROWS >= ny) return;
    const int tp = threadIdx.x;
    C* my = lds + threadIdx.y * LD;  // pump buffer; with two buffers the SH one is ROWS*LD further
    const size_t base = (size_t)y * NX + tp;
    C f[2][8];  // f[0] pump, f[1] SH: transformed together (shared barriers)
    C(&p)[8] = f[0];
    C(&s)[8] = f[1];
    S* const P = static_cast<S*>(Pv);
    S* const Sf = static_cast<S*>(Sv);
#pragma unroll
    for (int q = 0; q < 8; q++) {
        p[q] = from_store<C>(P[base + q * TPR]);
        s[q] = from_store<C>(Sf[base + q * TPR]);
    }
    // per-slice trace: the state before the SSFM step of slice g
    const bool trace = I.trP && g <= nz - 1 && g % I.trace_every == 0;
    if (trace && g == 0) {
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const size_t i = base + q * TPR;
            static_cast<S*>(I.trP)[i] = to_store<S>(p[q], i, store_key(3, 0, g));
            static_cast<S*>(I.trS)[i] = to_store<S>(s[q], i, store_key(3, 1, g));
        }
    }
    using Nl = NlRow<THERMAL, S, C>;
    // Loads ahead from 192 points up (below, their registers would cost a
    // wave per SIMD).  The first point's loads are issued after the row
    // IFFT: issued before it, they were waited for with the IFFT's twiddle
    // loads (vector-memory loads complete in order), which made the pass
    // slower (44.6 against 41.2 us per 16-instance 512x512 launch).
    if constexpr (Nl::kAhead && NX >= 192) {
        const Nl nl(I, g, y, trace);
        if (g > 0) fft_pair<C, NX, true>(f, my, ROWS * LD, tw, tp);
        nl_ahead<NX>(p, s, nl, tp, base, nl.load(tp));
    } else {
        if (g > 0) fft_pair<C, NX, true>(f, my, ROWS * LD, tw, tp);
        nl_inplace<NX, THERMAL, S, C>(p, s, I, g, y, tp, base, trace);
    }
    if (g < nz) fft_pair<C, NX, false>(f, my, ROWS * LD, tw, tp);
    const uint32_t kp = store_key(0, 0, g), ks = store_key(0, 1, g);
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const size_t i = base + q * TPR;
        P[i] = to_store<S>(p[q], i, kp);
        Sf[i] = to_store<S>(s[q], i, ks);
    }
}

// Wide row pass: the same step as k_rowpass with NX threads per row, for
// launches too small to fill the chip with NX/8 threads per row (a single
// grid of 256 x 256 or less, or a small batch).  The transforms run on
// groups of NX/8 threads holding 8 points each (group 2r + f: row r, field
// f; the other groups only join the barriers), the nonlinear work runs one
// point per thread from LDS, which cuts the serial work of a thread by 8.
// Bit-identical to k_rowpass (same transforms, same per-point arithmetic,
// same stores).  WROWS = max(1, 256 / NX) rows per workgroup.
template <int NX> constexpr int wide_rows() { return NX >= 256 ? 1 : 256 / NX; }

template <int NX, bool THERMAL, typename S, typename C>
__global__ __launch_bounds__(NX * wide_rows<NX>()) void k_rowpass_wide(const MixedInst* __restrict__ insts, Step gs,
                                                                        const double2* __restrict__ tw) {
    constexpr int ROWS = wide_rows<NX>();
    constexpr int TPR = NX / 8;
    constexpr int LD = FftShape<NX>::LDS;
    __shared__ C lds[2 * ROWS * LD];
    const MixedInst& I = insts[blockIdx.z];
    // the descriptor fields and the step are read before any check, so that
    // the field loads wait for one round of scalar loads, not one per check
    const int nz = I.nz, ny = I.ny;
    void* const Pv = I.P;
    void* const Sv = I.S;
    const int g = gs.get();
    asm volatile("" ::"s"(Pv), "s"(Sv));  // keeps the pointer loads ahead of the checks
    if (g > nz) return;
    if (blockIdx.x * ROWS >= ny) return;
    const int t = threadIdx.x, gi = t / TPR, tp = t % TPR;
    const size_t row0 = (size_t)blockIdx.x * ROWS * NX;
    S* const P = static_cast<S*>(Pv);
    S* const Sf = static_cast<S*>(Sv);
    // per-slice trace: the state before the SSFM step of slice g
    const bool trace = I.trP && g <= nz - 1 && g % I.trace_every == 0;
    const int r = t / NX, x = t % NX;
    using Nl = NlRow<THERMAL, S, C>;
    // THERMAL from 384 (FP64 arithmetic) or 512 points (FP32): the point's
    // heat-grid loads are issued after the field loads and are in flight
    // during the row IFFT (at smaller sizes the registers they hold there
    // would cost a workgroup per CU)
    constexpr bool early = THERMAL && NX >= (sizeof(C) == 16 ? 384 : 512);
    typename Nl::Vox vx{};
    // 1. rows in, spatial (row IFFT unless g == 0), into LDS
    {
        const bool act = gi < 2 * ROWS;
        const int f = gi & 1;
        const size_t base = row0 + (size_t)(gi >> 1) * NX + tp;
        S* const F = f ? Sf : P;
        C v[8];
        if (act) {
#pragma unroll
            for (int q = 0; q < 8; q++) v[q] = from_store<C>(F[base + q * TPR]);
        }
        if constexpr (early) vx = Nl::load_at(I, g, blockIdx.x * ROWS + r, x);
        if (act) {
            if (trace && g == 0) {
                S* const T = static_cast<S*>(f ? I.trS : I.trP);
                const uint32_t key = store_key(3, f, g);
#pragma unroll
                for (int q = 0; q < 8; q++) T[base + q * TPR] = to_store<S>(v[q], base + q * TPR, key);
            }
        }
        C* buf = lds + (act ? gi : 0) * LD;
        if (g > 0) fft_dist<C, NX, true>(v, buf, tw, tp, act);
        if (act) {
#pragma unroll
            for (int q = 0; q < 8; q++) buf[fpad(tp + q * TPR)] = v[q];
        }
    }
    __syncthreads();
    // 2. one point per thread
    {
        C* rb = lds + 2 * r * LD;
        const Nl nl(I, g, blockIdx.x * ROWS + r, trace);
        C p = rb[fpad(x)], sv = rb[LD + fpad(x)];
        if constexpr (early) nl.point(p, sv, x, row0 + (size_t)r * NX + x, vx);
        else nl.point(p, sv, x, row0 + (size_t)r * NX + x);
        rb[fpad(x)] = p;
        rb[LD + fpad(x)] = sv;
    }
    __syncthreads();
    // 3. row FFT (unless g == nz) and out
    {
        const bool act = gi < 2 * ROWS;
        const int f = gi & 1;
        C* buf = lds + (act ? gi : 0) * LD;
        C v[8];
        if (act) {
#pragma unroll
            for (int q = 0; q < 8; q++) v[q] = buf[fpad(tp + q * TPR)];
        }
        if (g < nz) {
            __syncthreads();  // every group has read its buffer before the transform overwrites it
            fft_dist<C, NX, false>(v, buf, tw, tp, act);
        }
        if (act) {
            S* const F = f ? Sf : P;
            const uint32_t key = store_key(0, f, g);
            const size_t base = row0 + (size_t)(gi >> 1) * NX + tp;
#pragma unroll
            for (int q = 0; q < 8; q++) F[base + q * TPR] = to_store<S>(v[q], base + q * TPR, key);
        }
    }
}

// Column pass of step g: column FFT -> * Prop -> column IFFT on COLS
// adjacent columns of one field (blockIdx.y: 0 pump, 1 SH), through
// fft_fwd_op_inv (four workgroup barriers for the pair of transforms).
// grid (ceil(max_nx/COLS), 2, n_instances), block (COLS, NY/8)
// FP64 arithmetic: keep 4 waves per SIMD (<= 128 VGPRs), the most the
// LDS footprint of the FP64 exchange allows anyway
template <typename C> struct ColWaves { static constexpr int W = sizeof(C) == 16 ? 4 : 1; };

template <int NY, int COLS, typename S, typename C>
__global__ __launch_bounds__(NY / 8 * COLS) __attribute__((amdgpu_waves_per_eu(ColWaves<C>::W)))
void k_colpass(const MixedInst* __restrict__ insts, Step gs, const double2* __restrict__ tw) {
    using R = re_t<C>;
    constexpr int TPC = NY / 8;
    constexpr int LD = FftShape<NY>::LDS;
    __shared__ C lds[COLS * LD];
    const MixedInst& I = insts[blockIdx.z];
    // the descriptor fields and the step are read before any check, so that
    // the field loads wait for one round of scalar loads, not one per check
    const bool sh = blockIdx.y != 0;
    const int nz = I.nz, nx = I.nx;
    void* const Fv = sh ? I.S : I.P;
    const double2* const ex = sh ? I.exS : I.exP;
    const double2* const ey = sh ? I.eyS : I.eyP;
    const int g = gs.get();
    asm volatile("" ::"s"(Fv), "s"(ex), "s"(ey));  // keeps the pointer loads ahead of the checks
    if (g < 0 || g >= nz) return;  // step -1: a lagging launch before the first step
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
    S* F = static_cast<S*>(Fv);
    C* my = lds + c * LD;
    C v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) v[q] = from_store<C>(F[(size_t)(tp + q * TPC) * nx + x]);
    // Prop = ex[x] ey[y], formed and applied in FP64 and rounded once to the
    // arithmetic type, so it carries no fixed representation error per mode.
    const double2 exv = ex[x];
    fft_fwd_op_inv<C, NY>(v, my, tw, tp, [&](C (&a)[8]) {
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const double2 pr = cmul(exv, ey[tp + q * TPC]);
            const double vr = a[q].x, vi = a[q].y;
            a[q] = mk<C>((R)(vr * pr.x - vi * pr.y), (R)(vr * pr.y + vi * pr.x));
        }
    });
    const uint32_t key = store_key(1, sh ? 1 : 0, g);
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const size_t i = (size_t)(tp + q * TPC) * nx + x;
        F[i] = to_store<S>(v[q], i, key);
    }
    // slice history for the adjoint's recomputation: the column-pass output
    // of the recorded slices, [slice - hist_g0][field][ny][nx] (the same
    // stored values: the rounding is a function of value, index and key)
    if (I.hist && g >= I.hist_g0 && g < I.hist_g0 + I.hist_n) {
        S* H = static_cast<S*>(I.hist) + ((size_t)(g - I.hist_g0) * 2 + (sh ? 1 : 0)) * ((size_t)nx * I.ny);
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const size_t i = (size_t)(tp + q * TPC) * nx + x;
            H[i] = to_store<S>(v[q], i, key);
        }
    }
}


// ------------------------------------------------------------------
// dispatch over transform length and tile plan, per (storage, arithmetic)
// pair; each pair is instantiated in its own translation unit
// ------------------------------------------------------------------
template <int NX, bool TH, typename S, typename C>
static hipError_t rowpass_rows(int rows, dim3 grid, hipStream_t st, const MixedInst* insts, Step g,
                               const double2* tw) {
    dim3 block(NX / 8, rows);
    switch (rows) {
        case kRowsWide:
            pass_launch((k_rowpass_wide<NX, TH, S, C>), grid, dim3(NX * wide_rows<NX>()), st, insts, g, tw);
            break;
        case 1: pass_launch((k_rowpass<NX, TH, 1, S, C>), grid, block, st, insts, g, tw); break;
        case 2: pass_launch((k_rowpass<NX, TH, 2, S, C>), grid, block, st, insts, g, tw); break;
        case 4: pass_launch((k_rowpass<NX, TH, 4, S, C>), grid, block, st, insts, g, tw); break;
        case 8:
            if constexpr (NX / 8 * 8 <= 1024) {
                pass_launch((k_rowpass<NX, TH, 8, S, C>), grid, block, st, insts, g, tw);
                break;
            }
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename S, typename C>
static hipError_t rowpass_dispatch(int nx, bool thermal, int rows, dim3 grid, hipStream_t st,
                                   const MixedInst* insts, Step g, const double2* tw) {
#define SHG_ROWCASE(N)                                                                         \
    case N:                                                                                    \
        return thermal ? rowpass_rows<N, true, S, C>(rows, grid, st, insts, g, tw)             \
                       : rowpass_rows<N, false, S, C>(rows, grid, st, insts, g, tw);
    switch (nx) {
        SHG_ROWCASE(64)
        SHG_ROWCASE(128)
        SHG_ROWCASE(192)
        SHG_ROWCASE(256)
        SHG_ROWCASE(384)
        SHG_ROWCASE(512)
    }
#undef SHG_ROWCASE
    return hipErrorInvalidValue;
}

template <int NY, typename S, typename C>
static hipError_t colpass_cols(int cols, dim3 grid, hipStream_t st, const MixedInst* insts, Step g,
                               const double2* tw) {
    dim3 block(cols, NY / 8);
    switch (cols) {
        case 2: pass_launch((k_colpass<NY, 2, S, C>), grid, block, st, insts, g, tw); break;
        case 4: pass_launch((k_colpass<NY, 4, S, C>), grid, block, st, insts, g, tw); break;
        case 8: pass_launch((k_colpass<NY, 8, S, C>), grid, block, st, insts, g, tw); break;
        case 16:
            if constexpr (NY / 8 * 16 <= 1024 && sizeof(C) * 16 * FftShape<NY>::LDS <= 160 * 1024) {
                pass_launch((k_colpass<NY, 16, S, C>), grid, block, st, insts, g, tw);
                break;
            }
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename S, typename C>
static hipError_t colpass_dispatch(int ny, int cols, dim3 grid, hipStream_t st, const MixedInst* insts, Step g,
                                   const double2* tw) {
    switch (ny) {
        case 64: return colpass_cols<64, S, C>(cols, grid, st, insts, g, tw);
        case 128: return colpass_cols<128, S, C>(cols, grid, st, insts, g, tw);
        case 192: return colpass_cols<192, S, C>(cols, grid, st, insts, g, tw);
        case 256: return colpass_cols<256, S, C>(cols, grid, st, insts, g, tw);
        case 384: return colpass_cols<384, S, C>(cols, grid, st, insts, g, tw);
        case 512: return colpass_cols<512, S, C>(cols, grid, st, insts, g, tw);
    }
    return hipErrorInvalidValue;
}

}  // namespace shg
