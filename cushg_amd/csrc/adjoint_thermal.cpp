// adjoint_thermal.cpp -- shg_adjoint for THERMAL runs: the gradient of eta
// at the converged thermo-optic steady state.  No reference counterpart (the
// reference computes no gradients); the forward it differentiates is the
// reference's outer loop (src/headers/Solver.h:321-373): heat solve from the
// last march's heat source, Tfield::upDate (src/headers/Tfield.h:93-201),
// the march on the resulting T with the heat source deposited slice by slice
// (src/headers/Tfield.h:260-276), until a heat solve needs one sweep.
//
// At the converged state T* = H(Q(T*), T_oven) (H: the linear steady-state
// solve, Q: the heat source of the march on T) the implicit-function adjoint
// is
//   mu = d eta/dT + (dQ/dT)^T w,   w = B^T A^-1 mu   (B = dr/dQ, diagonal),
//   d eta/d theta = d eta/d theta|_T + (dQ/d theta)^T w + mu^T dT/dT_oven,
// solved as the fixed point w = F(w) of the affine map F: one reverse sweep
// of the march for the loss L = eta + <w, Q> (giving mu and the theta terms
// together), then one FP64 steady-state solve with the symmetric operator A
// (F(w) = B^T A^-1 mu(w)).  The first sweep (w = 0) is the gradient at
// frozen T.  The iterates are accelerated by Anderson mixing over the last
// five steps (Walker & Ni, SIAM J. Numer. Anal. 49, 1715 (2011)): the next w
// is F(w_k) - sum_i g_i (dW_i + dF_i), g the least-squares fit of the
// residual f_k = F(w_k) - w_k by the differences dF_i of earlier residuals
// (dW_i of earlier iterates).  The iteration stops when the gradient's
// error -- its change between sweeps, or, once the changes contract by r <
// 1/2 a sweep, the geometric-tail estimate r / (1 - r) times the last change
// -- is below 1e-5 of its size (FP32-field forms, whose rounding leaves
// ~3e-7) or 1e-10 (FP64 forms), or after 2 max_outer sweeps.
//
// The march is the FP32-field / FP64-arithmetic fused march (FP64 fields
// for the FP64 forms).  Its last forward pass keeps every slice's
// column-pass output when that fits in a quarter of the free device memory
// (the sweeps then only read it); otherwise it is checkpointed every K
// slices and recomputed segment by segment on the handle's second stream
// while the adjoint of the next segment runs on the first.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <utility>
#include <vector>

#include "adjoint.hpp"
#include "common.hpp"
#include "heat.hpp"
#include "heat_kernels.hpp"
#include "kernels.hpp"
#include "march_mixed.hpp"
#include "runtime.hpp"

namespace shg {

namespace {

double dn_dT_th(int crystal, double lam, double T) {
    const double hT = 1e-3;
    return (refr_index<double>(crystal, lam, T + hT) - refr_index<double>(crystal, lam, T - hT)) / (2 * hT);
}

}  // namespace

int run_adjoint_thermal(shg_handle* h, const shg_params* p, int ck, double* d_eta, shg_result* r) {
    if (p->heat_solver != SHG_HEAT_DIRECT) {
        set_error("shg_adjoint: THERMAL gradients need the DIRECT heat solver (its exact steady state)");
        return SHG_EUNSUPPORTED;
    }
    if (p->waist_y_um > 0) {
        set_error("shg_adjoint: the gradient is taken for the circular pump (waist_y_um = 0)");
        return SHG_EUNSUPPORTED;
    }
    const auto t0 = std::chrono::steady_clock::now();
    shg_result local{};
    shg_result* res = r ? r : &local;
    MixedHost m;
    m.p = p;
    m.r = res;
    m.nx = p->nx;
    m.ny = p->ny;
    m.nz = p->nz;
    m.nzh = p->nz_heat > 0 ? p->nz_heat : p->nz;
    m.N2 = (size_t)m.nx * m.ny;
    m.dx = p->lx_um / (m.nx - 1);
    m.dy = p->ly_um / (m.ny - 1);
    m.dz = p->lz_um / (m.nz - 1);
    m.thermal = true;
    m.Tc = p->T_C;
    const bool f64 = p->precision == SHG_PREC_FP64_PLAIN || p->precision == SHG_PREC_FP64_FUSED;
    const Arith arith = f64 ? Arith::F64 : Arith::S32;
    m.arith = arith;
    const size_t fb = arith_field_bytes(arith);
    const int nz = m.nz, nzh = m.nzh;
    // every slice reaches the loss: eta through the saved state after nz-1
    // slices, the heat source through all nz (the last, unsaved slice too)
    // checkpoint interval: as asked; by default the whole march in one
    // segment when its history (every slice's column-pass output) takes at
    // most a quarter of the free device memory -- the sweeps then read it
    // and recompute nothing -- else sqrt(nz) slices
    const size_t two_b = 2 * fb * m.N2;
    int K = ck;
    if (K <= 0) {
        size_t fr = 0, tot = 0;
        const bool whole = hipMemGetInfo(&fr, &tot) == hipSuccess && two_b * (size_t)(nz + 1) <= fr / 4;
        K = whole ? nz : (int)std::ceil(std::sqrt((double)nz));
    }
    K = std::max(1, std::min(K, nz));
    const int ncp = (nz + K - 1) / K;
    const int nbr = adj_row_workgroups(m.nx, m.ny), nbc = adj_col_workgroups(m.nx), nbp = adj_pump_workgroups();
    const size_t NV = m.N2 * nzh;
    shg_params pc = *p;  // the heat solve's parameter copy (precision decides the FP32/FP64 tables)
    const HeatCoef hc = heat_coef(p, nzh);
    const size_t nu = (size_t)hc.nxu * hc.nyu * nzh;

    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t two = 2 * fb * m.N2;
    constexpr int kMix = 5;  // Anderson history
    const int ndot = adj_dot_workgroups();
    const int nhb = ncp > 1 ? 2 : 1;  // history buffers (two alternate between segments)
    size_t total = mixed_bytes(m) + al(sizeof(MixedInst) * (ncp + 1)) + al(two * ncp) + nhb * al(two * K) + al(two) +
                   al(sizeof(double) * ((size_t)(nz + 1) * nbr * 2 + (size_t)nz * nbc * 2 + (size_t)nbp * 3)) +
                   al(sizeof(double) * ((size_t)(nz + 1) * 2 + (size_t)nz * 2 + 3)) +
                   (9 + 2 * kMix) * al(sizeof(double) * NV) + 2 * al(sizeof(double) * nu) + al(64) +
                   al(sizeof(double) * ndot * 8) + 16384;
    if (int rc = h->arena.reserve(total)) return rc;
    Arena& ar = h->arena;
    m.P = ar.carve<char>(fb * m.N2);
    m.S = ar.carve<char>(fb * m.N2);
    m.outP = ar.carve<char>(fb * m.N2);
    m.outS = ar.carve<char>(fb * m.N2);
    m.exP = ar.carve<double2>(m.nx);
    m.exS = ar.carve<double2>(m.nx);
    m.eyP = ar.carve<double2>(m.ny);
    m.eyS = ar.carve<double2>(m.ny);
    m.phase = ar.carve<double2>((size_t)nz * 6);
    m.slots = ar.carve<double>(4);
    MixedInst* D = ar.carve<MixedInst>((size_t)ncp + 1);
    void* cp = ar.carve<char>(two * ncp);
    void* hist[2] = {ar.carve<char>(two * K), nullptr};
    hist[1] = nhb > 1 ? ar.carve<char>(two * K) : hist[0];
    void* lam = ar.carve<char>(two);
    double* partR = ar.carve<double>((size_t)(nz + 1) * nbr * 2);
    double* partC = ar.carve<double>((size_t)nz * nbc * 2);
    double* partP = ar.carve<double>((size_t)nbp * 3);
    double* G = ar.carve<double>((size_t)(nz + 1) * 2);
    double* Gpsi = ar.carve<double>((size_t)nz * 2);
    double* Gpump = ar.carve<double>(3);
    double* gnp = ar.carve<double>(NV);  // dL/dn_p, then dL/dT (mu)
    double* gns = ar.carve<double>(NV);  // dL/dn_s, then nu = A^-1 mu
    double* gA = ar.carve<double>(NV);
    double* gC = ar.carve<double>(NV);
    // the fixed-point iterate w (read by the sweeps), F(w), the residuals
    // and the Anderson history
    double* w = ar.carve<double>(NV);
    double* wprev = ar.carve<double>(NV);
    double* Fw = ar.carve<double>(NV);
    double* fcur = ar.carve<double>(NV);
    double* fprev = ar.carve<double>(NV);
    double* dW[kMix];
    double* dF[kMix];
    for (int i = 0; i < kMix; i++) {
        dW[i] = ar.carve<double>(NV);
        dF[i] = ar.carve<double>(NV);
    }
    double* u64 = ar.carve<double>(nu);
    double* v64 = ar.carve<double>(nu);
    double* tp = ar.carve<double>(8);
    double* dpart = ar.carve<double>((size_t)ndot * 8);
    if (!dpart || !tp || !D || !cp || !hist[0] || !hist[1] || !lam || !m.slots || !dF[kMix - 1]) {
        set_error("arena exhausted (THERMAL adjoint)");
        return SHG_ENOMEM;
    }
    if (int rc = heat_alloc(h, &pc, m.heat)) return rc;
    struct Release {
        HeatState& hs;
        ~Release() { heat_release(hs); }
    } release{m.heat};
    const double2* twx = twiddles64(h, m.nx);
    const double2* twy = twiddles64(h, m.ny);
    if (!twx || !twy) {
        set_error("twiddle table allocation failed");
        return SHG_ENOMEM;
    }
    MixedInst& dsc = m.desc;
    dsc.P = m.P;
    dsc.S = m.S;
    dsc.outP = m.outP;
    dsc.outS = m.outS;
    dsc.exP = m.exP;
    dsc.exS = m.exS;
    dsc.eyP = m.eyP;
    dsc.eyS = m.eyS;
    dsc.phase = m.phase;
    dsc.trace_every = 1;
    dsc.dkbase = m.heat.dkbase;
    dsc.dk = m.heat.dk;
    dsc.Q = m.heat.Q;
    dsc.poling = m.heat.pol;
    const int rows = plan_rows(h, m.nx, arith, m.ny, true);
    const int cols = plan_cols(h, m.ny, arith, m.nx);
    hipStream_t s1 = h->stream, s2 = h->stream2;
    struct Drain {
        hipStream_t s;
        ~Drain() { (void)hipStreamSynchronize(s); }
    } drain{s2};
    auto row = [&](int g, const MixedInst* di, hipStream_t st) -> int {
        const int per = rowpass_rows_per_wg(m.nx, rows);
        ProfScope ps(h, rowpass_name(m.nx, true, arith), st);
        SHG_HIP(launch_rowpass(m.nx, true, arith, rows, dim3((m.ny + per - 1) / per, 1, 1), st, di, g, twx));
        return SHG_OK;
    };
    auto col = [&](int g, const MixedInst* di, hipStream_t st) -> int {
        ProfScope ps(h, colpass_name(m.ny, arith), st);
        SHG_HIP(launch_colpass(m.ny, arith, cols, dim3((m.nx + cols - 1) / cols, 2, 1), st, di, g, twy));
        return SHG_OK;
    };
    auto field = [&](void* base, size_t slot, int f) -> void* {
        return static_cast<char*>(base) + (slot * 2 + f) * fb * m.N2;
    };
    auto copy2 = [&](void* dP, void* dS, const void* sP, const void* sS, hipStream_t st) -> int {
        SHG_HIP(hipMemcpyAsync(dP, sP, fb * m.N2, hipMemcpyDeviceToDevice, st));
        SHG_HIP(hipMemcpyAsync(dS, sS, fb * m.N2, hipMemcpyDeviceToDevice, st));
        return SHG_OK;
    };
    auto seg = [&](int k) { return std::make_pair(k * K, std::min(k * K + K, nz)); };

    // 1. the forward outer loop; the last pass marches with checkpoints
    if (int rc = heat_init(h, &pc, m.heat)) return rc;
    const int max_outer = p->max_outer > 0 ? p->max_outer : 20;
    long long sweeps = 0;
    int outer = 0;
    double heat_res = 0;
    for (int go = 0; go < max_outer; go++) {
        HeatSolveInfo info{};
        if (int rc = heat_solve(h, &pc, m.heat, info)) return rc;
        sweeps += info.sweeps;
        heat_res = info.residual;
        if (int rc = heat_prepare_optics(h, &pc, m.heat, &m.Tc)) return rc;
        dsc.nn = m.heat.nn;
        if (int rc = setup_fields(h, m)) return rc;  // tables (into dsc), pump, zero SH, P_in
        // src/headers/Solver.h:367: the pass after a one-sweep heat solve is the last
        const bool last = (info.converged_in_one && go != 0) || go == max_outer - 1;
        std::vector<MixedInst> descs((size_t)ncp + 1, dsc);
        if (last)
            for (int k = -1; k < ncp; k++) {
                const int kk = k < 0 ? ncp - 1 : k;
                MixedInst& q = descs[(size_t)(k + 1)];
                q.hist = hist[kk % 2];
                q.hist_g0 = seg(kk).first;
                q.hist_n = seg(kk).second - seg(kk).first;
            }
        SHG_HIP(hipMemcpyAsync(D, descs.data(), sizeof(MixedInst) * descs.size(), hipMemcpyHostToDevice, s1));
        for (int g = 0; g <= nz; g++) {
            if (last && g % K == 0 && g / K < ncp)
                if (int rc = copy2(field(cp, g / K, 0), field(cp, g / K, 1), m.P, m.S, s1)) return rc;
            if (int rc = row(g, D, s1)) return rc;
            if (g < nz)
                if (int rc = col(g, D, s1)) return rc;
        }
        SHG_HIP(hipStreamSynchronize(s1));  // descs (host) outlive the copy
        outer = go + 1;
        if (last) break;
    }
    if (int rc = field_sum_sq(h, arith, m.outP, (long)m.N2, m.slots + 1)) return rc;
    if (int rc = field_sum_sq(h, arith, m.outS, (long)m.N2, m.slots + 2)) return rc;
    double sl[3];
    SHG_HIP(hipMemcpyAsync(sl, m.slots, sizeof sl, hipMemcpyDeviceToHost, s1));
    SHG_HIP(hipStreamSynchronize(s1));
    const double lp = p->lambda_p_um, ls = 0.5 * lp, Tc = m.Tc;
    const double np = refr_index<double>(p->crystal, lp, Tc), ns = refr_index<double>(p->crystal, ls, Tc);
    const double cpow = 0.5 * kEps0 * kC * m.dx * m.dy;
    res->P_pump_in_W = cpow * np * sl[0];
    res->P_pump_out_W = cpow * np * sl[1];
    res->P_sh_W = cpow * ns * sl[2];
    res->eta = p->power_W > 0 ? res->P_sh_W / p->power_W : 0.0;
    res->steps_done = (long long)nz * outer;
    res->jacobi_sweeps = sweeps;
    res->outer_iters = outer;
    res->heat_residual_K = heat_res;
    if (!std::isfinite(res->eta)) {
        set_error("non-finite field power");
        return SHG_ENAN;
    }
    if (int rc = heat_finish(h, &pc, m.heat, res)) return rc;

    // 2. fixed-point iteration of the adjoint
    const double seed = p->power_W > 0 ? 2.0 * cpow * ns / p->power_W : 0.0;
    const double npT = dn_dT_th(p->crystal, lp, Tc), nsT = dn_dT_th(p->crystal, ls, Tc);
    const double kwp = 2 * kPi * np / lp, kws = 2 * kPi * ns / ls;
    const double w2 = p->waist_um * p->waist_um;
    const double zR = kPi * np * w2 / lp;
    const double Ap0 = std::sqrt(4.0 * p->power_W / (kEps0 * kC * kPi * np * w2));
    void* lamP = field(lam, 0, 0);
    void* lamS = field(lam, 0, 1);
    AdjArgs aa{};
    aa.I = D;
    aa.lamP = lamP;
    aa.lamS = lamS;
    aa.dfx = 1.0 / (m.dx * m.nx);
    aa.dfy = 1.0 / (m.dy * m.ny);
    hipEvent_t ev_rec[2] = {h->ev[1], h->ev[2]}, ev_adj[2] = {h->ev[3], h->ev[4]}, ev_start = h->ev[0];
    std::vector<double> grad(3 + (size_t)nz, 0.0), prev(3 + (size_t)nz, 0.0);
    std::vector<double> g((size_t)(nz + 1) * 2), gpsi((size_t)nz * 2), gp(3);
    const int max_it = 2 * max_outer;
    const double tol = f64 ? 1e-10 : 1e-5;
    auto dot = [&](const double* a, const double* b, double* out) -> int {
        SHG_HIP(launch_adj_dot(s1, a, b, (long)NV, dpart));
        SHG_HIP(launch_reduce(s1, dpart, ndot, false, tp + 1));
        SHG_HIP(hipMemcpyAsync(out, tp + 1, sizeof(double), hipMemcpyDeviceToHost, s1));
        SHG_HIP(hipStreamSynchronize(s1));
        return SHG_OK;
    };
    auto comb = [&](double* out, std::initializer_list<std::pair<double, const double*>> terms) -> int {
        double c[6];
        const double* x[6];
        int n = 0;
        for (auto& t : terms) {
            c[n] = t.first;
            x[n] = t.second;
            n++;
        }
        SHG_HIP(launch_adj_lincomb(s1, out, (long)NV, n, c, x));
        return SHG_OK;
    };
    SHG_HIP(hipMemsetAsync(w, 0, sizeof(double) * NV, s1));
    int nhist = 0, hslot = 0;
    double Gram[kMix][kMix] = {};  // dF_i . dF_j of the live history slots
    double dprev = 0;           // the gradient's change in the previous sweep
    int it = 0;
    for (;; it++) {
        AdjThermal th{};
        th.w = it == 0 ? nullptr : w;
        th.gnp = gnp;
        th.gns = gns;
        th.gA = gA;
        th.gC = gC;
        SHG_HIP(hipMemsetAsync(gnp, 0, sizeof(double) * NV, s1));
        SHG_HIP(hipMemsetAsync(gns, 0, sizeof(double) * NV, s1));
        SHG_HIP(hipMemsetAsync(gA, 0, sizeof(double) * NV, s1));
        SHG_HIP(hipMemsetAsync(gC, 0, sizeof(double) * NV, s1));
        SHG_HIP(hipMemsetAsync(partR, 0, sizeof(double) * (size_t)(nz + 1) * nbr * 2, s1));
        SHG_HIP(hipMemsetAsync(partC, 0, sizeof(double) * (size_t)nz * nbc * 2, s1));
        SHG_HIP(hipMemsetAsync(lam, 0, two, s1));
        if (it > 0 && ncp >= 3) {
            // the last segment's history was overwritten by the earlier
            // iteration's recomputations of its parity: recompute it
            const int k = ncp - 1, a = seg(k).first, b = seg(k).second;
            if (int rc = copy2(m.P, m.S, field(cp, k, 0), field(cp, k, 1), s1)) return rc;
            for (int gg = a; gg < b; gg++) {
                if (int rc = row(gg, D + 1 + k, s1)) return rc;
                if (int rc = col(gg, D + 1 + k, s1)) return rc;
            }
        }
        SHG_HIP(hipEventRecord(ev_start, s1));
        SHG_HIP(hipStreamWaitEvent(s2, ev_start, 0));
        for (int k = ncp - 1; k >= 0; k--) {
            if (k >= 1) {
                const int j = k - 1, a = seg(j).first, b = seg(j).second;
                if (j + 2 <= ncp - 1) SHG_HIP(hipStreamWaitEvent(s2, ev_adj[j % 2], 0));
                if (int rc = copy2(m.P, m.S, field(cp, j, 0), field(cp, j, 1), s2)) return rc;
                for (int gg = a; gg < b; gg++) {
                    if (int rc = row(gg, D + 1 + j, s2)) return rc;
                    if (int rc = col(gg, D + 1 + j, s2)) return rc;
                }
                SHG_HIP(hipEventRecord(ev_rec[j % 2], s2));
            }
            if (k != ncp - 1) SHG_HIP(hipStreamWaitEvent(s1, ev_rec[k % 2], 0));
            const int a = seg(k).first, b = seg(k).second;
            void* H = hist[k % 2];
            auto fwd = [&](int s, int f) { return s < a ? field(cp, k, f) : field(H, (size_t)(s - a), f); };
            for (int s = b - 1; s >= a; s--) {
                if (s == nz - 1) {
                    // boundary nz: the heat source of the last (unsaved) slice, then R2_{nz-1}^T
                    aa.s = nz;
                    aa.first = 0;
                    aa.second = 1;
                    aa.fwd_spatial = 0;
                    aa.fwdP = fwd(s, 0);
                    aa.fwdS = fwd(s, 1);
                    aa.part = partR + (size_t)nz * nbr * 2;
                    th.seed = 0.0;
                    ProfScope ps(h, "adjoint_rowpass_thermal");
                    SHG_HIP(launch_adj_row_th(f64, m.nx, m.ny, s1, aa, th, twx));
                }
                {
                    aa.s = s;
                    aa.fwdP = fwd(s, 0);
                    aa.fwdS = fwd(s, 1);
                    aa.part = partC + (size_t)s * nbc * 2;
                    ProfScope ps(h, "adjoint_colpass");
                    SHG_HIP(launch_adj_col(f64, m.nx, m.ny, s1, aa, twy));
                }
                {
                    aa.s = s;
                    aa.first = 1;
                    aa.second = s >= 1;
                    aa.fwd_spatial = s == 0;
                    aa.fwdP = fwd(s - 1, 0);
                    aa.fwdS = fwd(s - 1, 1);
                    aa.part = partR + (size_t)s * nbr * 2;
                    th.seed = s == nz - 1 ? seed : 0.0;
                    ProfScope ps(h, "adjoint_rowpass_thermal");
                    SHG_HIP(launch_adj_row_th(f64, m.nx, m.ny, s1, aa, th, twx));
                }
            }
            SHG_HIP(hipEventRecord(ev_adj[k % 2], s1));
        }
        SHG_HIP(launch_adj_pump(f64, lamP, m.nx, m.ny, m.dx, m.dy, p->waist_um, p->focal_um, np, zR, Ap0, partP, s1));
        SHG_HIP(launch_adj_reduce(partP, 1, nbp, 3, Gpump, s1));
        SHG_HIP(launch_adj_reduce(partR, nz + 1, nbr, 2, G, s1));
        SHG_HIP(launch_adj_reduce(partC, nz, nbc, 2, Gpsi, s1));
        SHG_HIP(hipMemcpyAsync(g.data(), G, sizeof(double) * g.size(), hipMemcpyDeviceToHost, s1));
        SHG_HIP(hipMemcpyAsync(gpsi.data(), Gpsi, sizeof(double) * gpsi.size(), hipMemcpyDeviceToHost, s1));
        SHG_HIP(hipMemcpyAsync(gp.data(), Gpump, sizeof(double) * 3, hipMemcpyDeviceToHost, s1));
        SHG_HIP(hipStreamSynchronize(s1));
        // the temperature at the grid centre of slab 0 sets the pump's n_p,
        // the propagator wavenumbers and the n_s of the power normalisation
        double spsi_p = 0, spsi_s = 0;
        for (int s = 0; s < nz; s++) {
            spsi_p += gpsi[(size_t)s * 2];
            spsi_s += gpsi[(size_t)s * 2 + 1];
        }
        const double dkwp = m.dz * 2 * kPi * kPi / (kwp * kwp * m.ny) * spsi_p;
        const double dkws = m.dz * 2 * kPi * kPi / (kws * kws * m.ny) * spsi_s;
        const double tc = gp[2] * npT + dkwp * (2 * kPi * npT / lp) + dkws * (2 * kPi * nsT / ls) + res->eta * nsT / ns;
        SHG_HIP(launch_adj_heat_grad(s1, m.heat.T, gnp, gns, gA, gC, m.nx, m.ny, nzh, nz, p->crystal, lp, p->Lambda_um,
                                     p->therm_exp_a, p->therm_exp_b, tc));
        // nu = A^-1 mu into gns, F(w) = B^T nu, and d<mu, T>/dT_oven
        if (int rc = heat_adjoint_solve(h, &pc, m.heat, gnp, gns, Fw, u64, v64, tp)) return rc;
        double tpv = 0;
        SHG_HIP(hipMemcpyAsync(&tpv, tp, sizeof(double), hipMemcpyDeviceToHost, s1));
        SHG_HIP(hipStreamSynchronize(s1));
        grad[0] = gp[0];
        grad[1] = gp[1];
        grad[2] = tpv;
        for (int s = 0; s < nz; s++) grad[3 + (size_t)s] = g[(size_t)s * 2] + g[(size_t)(s + 1) * 2 + 1];
        double dmax = 0, gmax = 0;
        for (size_t i = 0; i < grad.size(); i++) {
            dmax = std::max(dmax, std::fabs(grad[i] - prev[i]));
            gmax = std::max(gmax, std::fabs(grad[i]));
        }
        prev = grad;
        if (!std::isfinite(gmax)) {
            set_error("non-finite gradient");
            return SHG_ENAN;
        }
        // stop when the change is below tol, or when the error left, estimated
        // from the observed contraction r of the last two changes as
        // r / (1 - r) times the last change, is (r < 1/2, from the third sweep)
        const double r = it >= 2 && dprev > 0 ? dmax / dprev : 1.0;
        const bool settled = r < 0.5 && r / (1.0 - r) * dmax <= tol * gmax;
        if ((it > 0 && (dmax <= tol * gmax || settled)) || it + 1 >= max_it) break;
        dprev = dmax;
        // Anderson step: f = F(w) - w; history of (w - w_prev, f - f_prev)
        if (int rc = comb(fcur, {{1.0, Fw}, {-1.0, w}})) return rc;
        if (it > 0) {
            if (int rc = comb(dF[hslot], {{1.0, fcur}, {-1.0, fprev}})) return rc;
            if (int rc = comb(dW[hslot], {{1.0, w}, {-1.0, wprev}})) return rc;
            nhist = std::min(nhist + 1, kMix);
            for (int j = 0; j < nhist; j++) {  // the new slot's row of the Gram matrix
                if (int rc = dot(dF[hslot], dF[j], &Gram[hslot][j])) return rc;
                Gram[j][hslot] = Gram[hslot][j];
            }
            hslot = (hslot + 1) % kMix;
        }
        std::swap(fcur, fprev);
        std::swap(w, wprev);  // w is rewritten below
        if (nhist == 0) {
            SHG_HIP(hipMemcpyAsync(w, Fw, sizeof(double) * NV, hipMemcpyDeviceToDevice, s1));
            continue;
        }
        // least squares  min || f - sum_i g_i dF_i ||  by the normal equations
        // (Cholesky, with a relative ridge against a nearly dependent history)
        double Gm[kMix][kMix], rhs[kMix], gam[kMix];
        for (int i = 0; i < nhist; i++) {
            if (int rc = dot(dF[i], fprev, &rhs[i])) return rc;
            for (int j = 0; j < nhist; j++) Gm[i][j] = Gram[i][j];
        }
        double tr = 0;
        for (int i = 0; i < nhist; i++) tr += Gm[i][i];
        for (int i = 0; i < nhist; i++) Gm[i][i] += 1e-12 * tr;
        double L[kMix][kMix] = {};
        bool ok = tr > 0;
        for (int i = 0; i < nhist && ok; i++)
            for (int j = 0; j <= i; j++) {
                double a = Gm[i][j];
                for (int q = 0; q < j; q++) a -= L[i][q] * L[j][q];
                if (i == j) {
                    if (!(a > 0)) {
                        ok = false;
                        break;
                    }
                    L[i][i] = std::sqrt(a);
                } else {
                    L[i][j] = a / L[j][j];
                }
            }
        if (!ok) {  // no usable history: a plain step
            SHG_HIP(hipMemcpyAsync(w, Fw, sizeof(double) * NV, hipMemcpyDeviceToDevice, s1));
            continue;
        }
        double y[kMix];
        for (int i = 0; i < nhist; i++) {
            double a = rhs[i];
            for (int q = 0; q < i; q++) a -= L[i][q] * y[q];
            y[i] = a / L[i][i];
        }
        for (int i = nhist - 1; i >= 0; i--) {
            double a = y[i];
            for (int q = i + 1; q < nhist; q++) a -= L[q][i] * gam[q];
            gam[i] = a / L[i][i];
        }
        // w = F(w_k) - sum_i g_i (dW_i + dF_i), two history terms per launch
        if (int rc = comb(w, {{1.0, Fw}})) return rc;
        for (int i = 0; i < nhist; i++)
            if (int rc = comb(w, {{1.0, w}, {-gam[i], dW[i]}, {-gam[i], dF[i]}})) return rc;
    }
    for (size_t i = 0; i < grad.size(); i++) d_eta[i] = grad[i];
    res->adjoint_sweeps = it + 1;
    res->status = SHG_OK;
    res->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return SHG_OK;
}

}  // namespace shg
