// adjoint.cpp -- shg_adjoint: gradient of eta by the discrete adjoint of the
// fused non-THERMAL march, with checkpointed recomputation.  No reference
// counterpart: the reference computes no gradients.
//
//   1. forward march (the fused passes), keeping the state before every
//      K-th slice (checkpoint k: the state before slice kK);
//   2. seed: the gradient of eta = c n_s sum|S_out|^2 dx dy / P with respect
//      to the saved SH field S_out (the state after nz-1 slices);
//   3. for the segments [kK, min(kK + K, nz - 1)) from the last to the first:
//      restore checkpoint k, recompute the segment's slices keeping each
//      column-pass output, then run the adjoint column and row passes of
//      those slices backwards;
//   4. the adjoint state left at slice 0 is the gradient at the initial
//      pump, which gives d/dwaist, d/dfocus and the pump's n_p(T) term.
// The per-slice sums of the row passes (phase, coupling derivatives) and
// column passes (propagator wavenumber) give the poling-phase gradient and,
// with the temperature derivatives of n_p, n_s and dk, d eta / dT.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "adjoint.hpp"
#include "common.hpp"
#include "kernels.hpp"
#include "march_mixed.hpp"
#include "runtime.hpp"

namespace shg {

namespace {

struct AdjRun {
    shg_handle* h;
    MixedHost m;
    Arith arith;
    bool f64;
    size_t fb;                   // bytes per complex field value
    MixedInst* d = nullptr;      // device descriptors: [0] forward, [1 + k] recomputation of segment k
    const double2* twx = nullptr;
    const double2* twy = nullptr;
    int rows = 0, cols = 0;

    void* field(void* base, size_t slot, int f) const {
        return static_cast<char*>(base) + (slot * 2 + f) * fb * m.N2;
    }

    int row(int g, const MixedInst* di, hipStream_t st) {
        const int per = rowpass_rows_per_wg(m.nx, rows);
        dim3 grid((m.ny + per - 1) / per, 1, 1);
        ProfScope ps(h, rowpass_name(m.nx, false, arith), st);
        SHG_HIP(launch_rowpass(m.nx, false, arith, rows, grid, st, di, g, twx));
        return SHG_OK;
    }
    int col(int g, const MixedInst* di, hipStream_t st) {
        dim3 grid((m.nx + cols - 1) / cols, 2, 1);
        ProfScope ps(h, colpass_name(m.ny, arith), st);
        SHG_HIP(launch_colpass(m.ny, arith, cols, grid, st, di, g, twy));
        return SHG_OK;
    }
    int copy_fields(void* dstP, void* dstS, const void* srcP, const void* srcS, hipStream_t st) {
        SHG_HIP(hipMemcpyAsync(dstP, srcP, fb * m.N2, hipMemcpyDeviceToDevice, st));
        SHG_HIP(hipMemcpyAsync(dstS, srcS, fb * m.N2, hipMemcpyDeviceToDevice, st));
        return SHG_OK;
    }
};

double dn_dT(int crystal, double lam, double T) {
    const double hT = 1e-3;  // central difference of the closed form, truncation < 1e-11 relative
    return (refr_index<double>(crystal, lam, T + hT) - refr_index<double>(crystal, lam, T - hT)) / (2 * hT);
}

}  // namespace

int run_adjoint_thermal(shg_handle* h, const shg_params* p, int ck, double* d_eta, shg_result* r);

int run_adjoint(shg_handle* h, const shg_params* p, int ck, double* d_eta, shg_result* r) {
    if (p->thermal) return run_adjoint_thermal(h, p, ck, d_eta, r);
    if (p->waist_y_um > 0) {
        set_error("shg_adjoint: the gradient is taken for the circular pump (waist_y_um = 0)");
        return SHG_EUNSUPPORTED;
    }
    const auto t0 = std::chrono::steady_clock::now();
    shg_result local{};
    shg_result* res = r ? r : &local;
    AdjRun A{};
    A.h = h;
    MixedHost& m = A.m;
    m.p = p;
    m.r = res;
    m.nx = p->nx;
    m.ny = p->ny;
    m.nz = p->nz;
    m.nzh = m.nz;
    m.N2 = (size_t)m.nx * m.ny;
    m.dx = p->lx_um / (m.nx - 1);
    m.dy = p->ly_um / (m.ny - 1);
    m.dz = p->lz_um / (m.nz - 1);
    m.thermal = false;
    m.Tc = p->T_C;
    // FP64 arithmetic in both: FP64 fields for the FP64 forms, FP32 otherwise
    A.f64 = p->precision == SHG_PREC_FP64_PLAIN || p->precision == SHG_PREC_FP64_FUSED;
    A.arith = A.f64 ? Arith::F64 : Arith::S32;
    m.arith = A.arith;
    A.fb = arith_field_bytes(A.arith);
    const int nz = m.nz;
    const int nsl = nz - 1;  // slices 0..nz-2 reach the output
    int K = ck > 0 ? ck : (int)std::ceil(std::sqrt((double)nsl));
    K = std::max(1, std::min(K, nsl));
    const int ncp = (nsl + K - 1) / K;
    const int nbr = adj_row_workgroups(m.nx, m.ny), nbc = adj_col_workgroups(m.nx), nbp = adj_pump_workgroups();
    A.rows = plan_rows(h, m.nx, A.arith, m.ny);
    A.cols = plan_cols(h, m.ny, A.arith, m.nx);

    // segment k covers slices [kK, min(kK + K, nsl)); its recomputation
    // keeps the column-pass output C_s of each of its slices in one of two
    // history buffers (segment k uses buffer k % 2), so that segment k - 1
    // is recomputed on a second stream while the adjoint of segment k runs
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t two = 2 * A.fb * m.N2;
    size_t total = mixed_bytes(m) + al(sizeof(MixedInst) * (ncp + 1)) + al(two * ncp) + 2 * al(two * K) +
                   al(two) + al(sizeof(double) * ((size_t)nz * nbr * 12 + (size_t)nz * nbc * 2 + (size_t)nbp * 3)) +
                   al(sizeof(double) * ((size_t)nz * 12 + (size_t)nz * 2 + 3)) + 8192;
    if (int rc = h->arena.reserve(total)) return rc;
    Arena& ar = h->arena;
    m.P = ar.carve<char>(A.fb * m.N2);
    m.S = ar.carve<char>(A.fb * m.N2);
    m.outP = ar.carve<char>(A.fb * m.N2);
    m.outS = ar.carve<char>(A.fb * m.N2);
    m.exP = ar.carve<double2>(m.nx);
    m.exS = ar.carve<double2>(m.nx);
    m.eyP = ar.carve<double2>(m.ny);
    m.eyS = ar.carve<double2>(m.ny);
    m.phase = ar.carve<double2>((size_t)nz * 6);
    m.slots = ar.carve<double>(4);
    A.d = ar.carve<MixedInst>((size_t)ncp + 1);
    void* cp = ar.carve<char>(two * ncp);
    void* hist[2] = {ar.carve<char>(two * K), ar.carve<char>(two * K)};
    void* lam = ar.carve<char>(two);
    double* partR = ar.carve<double>((size_t)nz * nbr * 12);
    double* partC = ar.carve<double>((size_t)nz * nbc * 2);
    double* partP = ar.carve<double>((size_t)nbp * 3);
    double* G = ar.carve<double>((size_t)nz * 12);
    double* Gpsi = ar.carve<double>((size_t)nz * 2);
    double* Gpump = ar.carve<double>(3);
    if (!Gpump || !A.d || !cp || !hist[0] || !hist[1] || !lam) {
        set_error("arena exhausted (adjoint)");
        return SHG_ENOMEM;
    }
    A.twx = twiddles64(h, m.nx);
    A.twy = twiddles64(h, m.ny);
    if (!A.twx || !A.twy) {
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
    if (int rc = setup_fields(h, m)) return rc;  // tables (into desc), pump, zero SH, P_in
    // descriptors: the forward records the last segment's history itself;
    // descriptor 1 + k recomputes segment k
    auto seg = [&](int k) { return std::make_pair(k * K, std::min(k * K + K, nsl)); };
    std::vector<MixedInst> descs((size_t)ncp + 1, dsc);
    for (int k = -1; k < ncp; k++) {
        const int kk = k < 0 ? ncp - 1 : k;
        MixedInst& q = descs[(size_t)(k + 1)];
        q.hist = hist[kk % 2];
        q.hist_g0 = seg(kk).first;
        q.hist_n = seg(kk).second - seg(kk).first;
    }
    SHG_HIP(hipMemcpyAsync(A.d, descs.data(), sizeof(MixedInst) * descs.size(), hipMemcpyHostToDevice, h->stream));
    SHG_HIP(hipMemsetAsync(partR, 0, sizeof(double) * (size_t)nz * nbr * 12, h->stream));
    SHG_HIP(hipMemsetAsync(partC, 0, sizeof(double) * (size_t)nz * nbc * 2, h->stream));
    hipStream_t s1 = h->stream, s2 = h->stream2;
    // whatever path leaves this function, the second stream has finished
    // its work before the arena it reads is handed to another call
    struct Drain {
        hipStream_t s;
        ~Drain() { (void)hipStreamSynchronize(s); }
    } drain{s2};
    hipEvent_t ev_fwd = h->ev[0], ev_rec[2] = {h->ev[1], h->ev[2]}, ev_adj[2] = {h->ev[3], h->ev[4]};

    // 1. forward march with checkpoints (state before slices 0, K, 2K, ...)
    for (int g = 0; g < nz; g++) {
        if (g % K == 0 && g / K < ncp)
            if (int rc = A.copy_fields(A.field(cp, g / K, 0), A.field(cp, g / K, 1), m.P, m.S, s1)) return rc;
        if (int rc = A.row(g, A.d, s1)) return rc;  // row pass nz-1 saves the output state
        if (g == nz - 1) break;
        if (int rc = A.col(g, A.d, s1)) return rc;
    }
    if (int rc = field_sum_sq(h, A.arith, m.outP, (long)m.N2, m.slots + 1)) return rc;
    if (int rc = field_sum_sq(h, A.arith, m.outS, (long)m.N2, m.slots + 2)) return rc;
    double sl[3];
    SHG_HIP(hipMemcpyAsync(sl, m.slots, sizeof sl, hipMemcpyDeviceToHost, s1));
    SHG_HIP(hipEventRecord(ev_fwd, s1));
    SHG_HIP(hipStreamWaitEvent(s2, ev_fwd, 0));
    const double lp = p->lambda_p_um, ls = 0.5 * lp, Tc = p->T_C;
    const double np = refr_index<double>(p->crystal, lp, Tc), ns = refr_index<double>(p->crystal, ls, Tc);
    const double cpow = 0.5 * kEps0 * kC * m.dx * m.dy;

    // 2. seed: d eta / d S_out = 2 c n_s dx dy / P * S_out
    void* lamP = A.field(lam, 0, 0);
    void* lamS = A.field(lam, 0, 1);
    const double coef = p->power_W > 0 ? 2.0 * cpow * ns / p->power_W : 0.0;
    SHG_HIP(launch_adj_seed(A.f64, lamP, lamS, m.outS, (long)m.N2, coef, s1));

    // 3. segments backwards; the recomputation of segment k - 1 (stream 2)
    // overlaps the adjoint of segment k (stream 1)
    AdjArgs aa{};
    aa.I = A.d;
    aa.lamP = lamP;
    aa.lamS = lamS;
    aa.dfx = 1.0 / (m.dx * m.nx);
    aa.dfy = 1.0 / (m.dy * m.ny);
    for (int k = ncp - 1; k >= 0; k--) {
        if (k >= 1) {
            const int j = k - 1, a = seg(j).first, b = seg(j).second;
            // history buffer j % 2 was last read by the adjoint of segment j + 2
            if (j + 2 <= ncp - 1) SHG_HIP(hipStreamWaitEvent(s2, ev_adj[j % 2], 0));
            if (int rc = A.copy_fields(m.P, m.S, A.field(cp, j, 0), A.field(cp, j, 1), s2)) return rc;
            for (int g = a; g < b; g++) {
                if (int rc = A.row(g, A.d + 1 + j, s2)) return rc;
                if (int rc = A.col(g, A.d + 1 + j, s2)) return rc;
            }
            SHG_HIP(hipEventRecord(ev_rec[j % 2], s2));
        }
        if (k != ncp - 1) SHG_HIP(hipStreamWaitEvent(s1, ev_rec[k % 2], 0));
        const int a = seg(k).first, b = seg(k).second;
        void* H = hist[k % 2];
        // C_s of slice s; C_{a-1} (or the initial spatial state at a = 0) is the checkpoint
        auto fwd = [&](int s, int f) { return s < a ? A.field(cp, k, f) : A.field(H, (size_t)(s - a), f); };
        for (int s = b - 1; s >= a; s--) {
            if (s == nsl - 1) {
                // adjoint of R2_{nz-2} from the seed (spatial): boundary nz-1, second half only
                aa.s = nz - 1;
                aa.first = 0;
                aa.second = 1;
                aa.fwd_spatial = 0;
                aa.fwdP = fwd(s, 0);
                aa.fwdS = fwd(s, 1);
                aa.part = partR + (size_t)(nz - 1) * nbr * 12;
                ProfScope ps(h, "adjoint_rowpass");
                SHG_HIP(launch_adj_row(A.f64, m.nx, m.ny, s1, aa, A.twx));
            }
            {
                aa.s = s;
                aa.fwdP = fwd(s, 0);
                aa.fwdS = fwd(s, 1);
                aa.part = partC + (size_t)s * nbc * 2;
                ProfScope ps(h, "adjoint_colpass");
                SHG_HIP(launch_adj_col(A.f64, m.nx, m.ny, s1, aa, A.twy));
            }
            {
                aa.s = s;
                aa.first = 1;
                aa.second = s >= 1;
                aa.fwd_spatial = s == 0;
                aa.fwdP = fwd(s - 1, 0);  // C_{s-1}, or X_0 at s = 0
                aa.fwdS = fwd(s - 1, 1);
                aa.part = partR + (size_t)s * nbr * 12;
                ProfScope ps(h, "adjoint_rowpass");
                SHG_HIP(launch_adj_row(A.f64, m.nx, m.ny, s1, aa, A.twx));
            }
        }
        SHG_HIP(hipEventRecord(ev_adj[k % 2], s1));
    }
    // 4. initial-field gradients and the per-slice sums
    const double w2 = p->waist_um * p->waist_um;
    const double zR = kPi * np * w2 / lp;
    const double Ap0 = std::sqrt(4.0 * p->power_W / (kEps0 * kC * kPi * np * w2));
    SHG_HIP(launch_adj_pump(A.f64, lamP, m.nx, m.ny, m.dx, m.dy, p->waist_um, p->focal_um, np, zR, Ap0, partP,
                            h->stream));
    SHG_HIP(launch_adj_reduce(partP, 1, nbp, 3, Gpump, h->stream));
    SHG_HIP(launch_adj_reduce(partR, nz, nbr, 12, G, h->stream));
    SHG_HIP(launch_adj_reduce(partC, nz, nbc, 2, Gpsi, h->stream));
    std::vector<double> g((size_t)nz * 12), gpsi((size_t)nz * 2), gp(3);
    SHG_HIP(hipMemcpyAsync(g.data(), G, sizeof(double) * g.size(), hipMemcpyDeviceToHost, h->stream));
    SHG_HIP(hipMemcpyAsync(gpsi.data(), Gpsi, sizeof(double) * gpsi.size(), hipMemcpyDeviceToHost, h->stream));
    SHG_HIP(hipMemcpyAsync(gp.data(), Gpump, sizeof(double) * 3, hipMemcpyDeviceToHost, h->stream));
    SHG_HIP(hipStreamSynchronize(h->stream));
    res->P_pump_in_W = cpow * np * sl[0];
    res->P_pump_out_W = cpow * np * sl[1];
    res->P_sh_W = cpow * ns * sl[2];
    res->eta = p->power_W > 0 ? res->P_sh_W / p->power_W : 0.0;
    res->steps_done = nz;
    res->outer_iters = 1;
    res->adjoint_sweeps = 1;
    res->status = SHG_OK;
    if (!std::isfinite(res->eta)) {
        set_error("non-finite field power");
        return SHG_ENAN;
    }

    // assembly: slice s gets R1_s from row-pass launch s ([0..5]) and R2_s
    // from launch s+1 ([6..11])
    double ddk = 0, dkp = 0, dks = 0, dbf = 0;
    for (int s = 0; s < nz; s++) d_eta[3 + s] = 0.0;
    for (int s = 0; s < nsl; s++) {
        const double* c1 = &g[(size_t)s * 12];
        const double* c2 = &g[(size_t)(s + 1) * 12 + 6];
        double dp = 0;
        for (int j = 0; j < 3; j++) {
            dp += c1[j] + c2[j];
            ddk += ((double)s + dsc.off[0][j]) * m.dz * c1[j] + ((double)s + dsc.off[1][j]) * m.dz * c2[j];
        }
        d_eta[3 + s] = dp;  // d phi / d poling[s] = 1 for all six phase factors of slice s
        dkp += c1[3] + c2[3];
        dks += c1[4] + c2[4];
        dbf += c1[5] + c2[5];
    }
    double spsi_p = 0, spsi_s = 0;
    for (int s = 0; s < nsl; s++) {
        spsi_p += gpsi[(size_t)s * 2];
        spsi_s += gpsi[(size_t)s * 2 + 1];
    }
    const double kwp = 2 * kPi * np / lp, kws = 2 * kPi * ns / ls;
    const double dkwp = m.dz * 2 * kPi * kPi / (kwp * kwp * m.ny) * spsi_p;
    const double dkws = m.dz * 2 * kPi * kPi / (kws * kws * m.ny) * spsi_s;
    d_eta[0] = gp[0];
    d_eta[1] = gp[1];
    // temperature: dk(T), the couplings kp ~ 1/n_p, ks ~ 1/n_s, bfac ~ n_s,
    // the propagator wavenumbers 2 pi n / lambda, the pump's n_p and the
    // n_s of the power normalisation
    const double npT = dn_dT(p->crystal, lp, Tc), nsT = dn_dT(p->crystal, ls, Tc);
    const double hT = 1e-3;
    const double dkT = (delta_k(p->crystal, lp, p->Lambda_um, p->therm_exp_a, p->therm_exp_b, Tc + hT) -
                        delta_k(p->crystal, lp, p->Lambda_um, p->therm_exp_a, p->therm_exp_b, Tc - hT)) /
                       (2 * hT);
    d_eta[2] = ddk * dkT + dkp * (-dsc.kp * npT / np) + dks * (-dsc.ks * nsT / ns) + dbf * (dsc.bfac * nsT / ns) +
               dkwp * (2 * kPi * npT / lp) + dkws * (2 * kPi * nsT / ls) + gp[2] * npT + res->eta * nsT / ns;
    for (int i = 0; i < 3 + nz; i++)
        if (!std::isfinite(d_eta[i])) {
            set_error("non-finite gradient");
            return SHG_ENAN;
        }
    res->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return SHG_OK;
}

}  // namespace shg
