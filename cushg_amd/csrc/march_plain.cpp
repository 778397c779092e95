// march_plain.cpp -- the plain FP64 GPU form: one instance, the reference's
// op sequence (Solver::SSFM = solverRK4, diffraction, solverRK4 with
// separate kernels per stage, src/headers/Solver.h:153-308), FP64 throughout,
// no fusion and no graph.  It is the denominator of the FP32-mixed speedup.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "heat.hpp"
#include "kernels.hpp"
#include "plain_kernels.hpp"
#include "runtime.hpp"

namespace shg {

namespace {

struct Plain {
    const shg_params* p;
    int nx, ny, nz, nzh, n2;
    double dx, dy, dz;
    double2 *P, *S, *outP, *outS, *aP, *aS, *propP, *propS;
    double2* k[8];  // k1p k1s k2p k2s k3p k3s k4p k4s
    double* slots;
    const double2 *twx, *twy;
    HeatState heat;
    bool thermal;
};

// propagators in natural FFT order (src/headers/Efields.h:122-141 followed
// by fftShift2D, :68-81), FP64, unscaled
void host_props(const Plain& m, double Tc, std::vector<double2>& pp, std::vector<double2>& ps) {
    const shg_params* p = m.p;
    const double lp = p->lambda_p_um, ls = 0.5 * lp;
    const double kwp = 2 * kPi * refr_index<double>(p->crystal, lp, Tc) / lp;
    const double kws = 2 * kPi * refr_index<double>(p->crystal, ls, Tc) / ls;
    const double dfX = 1.0 / m.dx / m.nx, dfY = 1.0 / m.dy / m.ny;
    pp.resize(m.n2);
    ps.resize(m.n2);
    for (int y = 0; y < m.ny; y++) {
        const double fy = (((y + m.ny / 2) % m.ny) - 0.5 * m.ny) * dfY;
        for (int x = 0; x < m.nx; x++) {
            const double fx = (((x + m.nx / 2) % m.nx) - 0.5 * m.nx) * dfX;
            const double ap = -m.dz * (2 * kPi * kPi / kwp * (fx * fx + fy * fy) + 2 * kPi * fx * std::tan(p->rho_p));
            const double as = -m.dz * (2 * kPi * kPi / kws * (fx * fx + fy * fy) + 2 * kPi * fx * std::tan(p->rho_s));
            pp[(size_t)y * m.nx + x] = make_double2(std::cos(ap), std::sin(ap));
            ps[(size_t)y * m.nx + x] = make_double2(std::cos(as), std::sin(as));
        }
    }
}

int diffraction(shg_handle* h, Plain& m, double2* A, const double2* prop) {
    {
        ProfScope ps(h, "plain_fft_rows");
        SHG_HIP(launch_fft_rows64(h->stream, A, m.nx, m.ny, m.twx, false));
    }
    {
        ProfScope ps(h, "plain_fft_cols");
        SHG_HIP(launch_fft_cols64(h->stream, A, m.nx, m.ny, m.twy, false));
    }
    {
        ProfScope ps(h, "plain_prop");
        SHG_HIP(launch_cmul_inplace(h->stream, m.n2, A, prop));
    }
    {
        ProfScope ps(h, "plain_fft_rows");
        SHG_HIP(launch_fft_rows64(h->stream, A, m.nx, m.ny, m.twx, true));
    }
    {
        ProfScope ps(h, "plain_fft_cols");
        SHG_HIP(launch_fft_cols64(h->stream, A, m.nx, m.ny, m.twy, true));
    }
    {
        ProfScope ps(h, "plain_scale");
        SHG_HIP(launch_scale(h->stream, m.n2, 1.0 / ((double)m.nx * m.ny), A));
    }
    return SHG_OK;
}

// Solver::solverRK4 (src/headers/Solver.h:189-264)
int rk4(shg_handle* h, Plain& m, PlainStage st, const double c[3], double w, const double off[3]) {
    const double inc[4] = {off[0], off[1], off[1], off[2]};
    const double2* inP = m.P;
    const double2* inS = m.S;
    for (int i = 0; i < 4; i++) {
        st.inc = inc[i];
        {
            ProfScope ps(h, "plain_dadz");
            SHG_HIP(launch_dadz(h->stream, m.thermal, st, inP, inS, m.k[2 * i], m.k[2 * i + 1]));
        }
        if (i < 3) {
            ProfScope ps(h, "plain_lincomb");
            SHG_HIP(launch_lincomb(h->stream, m.n2, c[i], m.P, m.S, m.k[2 * i], m.k[2 * i + 1], m.aP, m.aS));
            inP = m.aP;
            inS = m.aS;
        }
    }
    ProfScope ps(h, "plain_rk4");
    SHG_HIP(launch_rk4_final(h->stream, m.n2, w, m.P, m.S, m.k));
    return SHG_OK;
}

}  // namespace

int run_plain(shg_handle* h, const shg_params* p, shg_result* r, SliceTrace* trace) {
    const auto t0 = std::chrono::steady_clock::now();
    Plain m{};
    m.p = p;
    m.nx = p->nx;
    m.ny = p->ny;
    m.nz = p->nz;
    m.nzh = p->nz_heat > 0 ? p->nz_heat : p->nz;
    m.n2 = m.nx * m.ny;
    m.dx = p->lx_um / (m.nx - 1);
    m.dy = p->ly_um / (m.ny - 1);
    m.dz = p->lz_um / (m.nz - 1);
    m.thermal = p->thermal != 0;
    const size_t f = ((sizeof(double2) * m.n2 + 255) & ~size_t(255));
    size_t total = 16 * f + 4096 + (m.thermal ? heat_bytes(p) : 0);
    if (int rc = h->arena.reserve(total)) return rc;
    Arena& A = h->arena;
    double2** bufs[] = {&m.P, &m.S, &m.outP, &m.outS, &m.aP, &m.aS, &m.propP, &m.propS};
    for (auto b : bufs) *b = A.carve<double2>(m.n2);
    for (int i = 0; i < 8; i++) m.k[i] = A.carve<double2>(m.n2);
    m.slots = A.carve<double>(4);
    m.twx = twiddles64(h, m.nx);
    m.twy = twiddles64(h, m.ny);
    if (!m.slots || !m.k[7] || !m.twx || !m.twy) {
        set_error("plain form: allocation failed");
        return SHG_ENOMEM;
    }
    struct HeatGuard {
        HeatState* hs;
        ~HeatGuard() { heat_release(*hs); }
    } guard{&m.heat};
    if (m.thermal) {
        if (int rc = heat_alloc(h, p, m.heat)) return rc;
        if (int rc = heat_init(h, p, m.heat)) return rc;
    }

    // scheme constants (src/headers/Solver.h:235-256, or Strang)
    double c[3] = {0.5, 0.5, 1.0}, w = m.dz / 6.0, kmul = 1.0;
    double off[2][3] = {{0.0, 0.25, 0.5}, {0.0, 0.25, 0.5}};
    if (p->scheme == SHG_SCHEME_STRANG) {
        const double hh = 0.5 * m.dz;
        c[0] = c[1] = 0.5 * hh;
        c[2] = hh;
        w = hh / 6.0;
        kmul = 2.0;
        off[1][0] = 0.5;
        off[1][1] = 0.75;
        off[1][2] = 1.0;
    }
    const double lp = p->lambda_p_um, ls = 0.5 * lp;
    const double dQ = 2.0 * p->deff_um_per_V / kPi;
    const int max_outer = m.thermal ? (p->max_outer > 0 ? p->max_outer : 20) : 1;
    long long steps = 0, sweeps = 0;
    int outer = 0;
    double Tc = p->T_C, heat_res = 0;
    std::vector<double2> pp, ps;
    for (int go = 0; go < max_outer; go++) {
        HeatSolveInfo info{};
        if (m.thermal) {
            if (int rc = heat_solve(h, p, m.heat, info)) return rc;
            sweeps += info.sweeps;
            heat_res = info.residual;
            if (int rc = heat_prepare_optics(h, p, m.heat, &Tc)) return rc;
        }
        host_props(m, Tc, pp, ps);
        SHG_HIP(hipMemcpy(m.propP, pp.data(), sizeof(double2) * m.n2, hipMemcpyHostToDevice));
        SHG_HIP(hipMemcpy(m.propS, ps.data(), sizeof(double2) * m.n2, hipMemcpyHostToDevice));
        const double np = refr_index<double>(p->crystal, lp, Tc);
        const double ns = refr_index<double>(p->crystal, ls, Tc);
        SHG_HIP(launch_pump_init<double2>(h->stream, m.P, m.nx, m.ny, m.dx, m.dy, pump_coef(p, np)));
        SHG_HIP(hipMemsetAsync(m.S, 0, sizeof(double2) * m.n2, h->stream));
        if (int rc = sum_sq64(h, m.P, m.n2, m.slots + 0)) return rc;

        PlainStage st{};
        st.n2 = m.n2;
        st.crystal = p->crystal;
        st.dz = m.dz;
        st.dk0 = delta_k(p->crystal, lp, p->Lambda_um, p->therm_exp_a, p->therm_exp_b, Tc);
        st.kp = kmul * kPi * dQ / (np * lp);
        st.ks = kmul * 0.5 * kPi * dQ / (ns * ls);
        st.bfac = p->beta_s * 0.5 * kC * kEps0 * ns;
        st.kp0 = kmul * kPi * dQ / lp;
        st.ks0 = kmul * 0.5 * kPi * dQ / ls;
        st.bf0 = p->beta_s * 0.5 * kC * kEps0;
        st.alpha_p = p->alpha_p;
        st.alpha_s = p->alpha_s;
        st.beta_s = p->beta_s;
        st.hce = 0.5 * kC * kEps0;
        st.lp = lp;
        st.ls = ls;
        if (m.thermal) {
            st.T = m.heat.T;
            st.dkbase = m.heat.dkbase;
            st.dk = m.heat.dk;
        }
        // Solver::CWES (src/headers/Solver.h:276-308)
        for (int s = 0; s < m.nz; s++) {
            if (s == m.nz - 1) {  // saved before the last slice
                SHG_HIP(hipMemcpyAsync(m.outP, m.P, sizeof(double2) * m.n2, hipMemcpyDeviceToDevice, h->stream));
                SHG_HIP(hipMemcpyAsync(m.outS, m.S, sizeof(double2) * m.n2, hipMemcpyDeviceToDevice, h->stream));
            }
            if (trace && trace->wants(s, m.nz)) {
                trace->pump.resize(2 * m.n2);
                trace->signal.resize(2 * m.n2);
                SHG_HIP(hipMemcpyAsync(trace->pump.data(), m.P, sizeof(double2) * m.n2, hipMemcpyDeviceToHost,
                                       h->stream));
                SHG_HIP(hipMemcpyAsync(trace->signal.data(), m.S, sizeof(double2) * m.n2, hipMemcpyDeviceToHost,
                                       h->stream));
                SHG_HIP(hipStreamSynchronize(h->stream));
                if (int rc = trace->deliver(s, go)) return rc;
            }
            st.step = s;
            st.h = (int)(((long long)s * m.nzh) / m.nz);
            st.pol = p->poling_phase ? p->poling_phase[s] : 0.0;
            if (int rc = rk4(h, m, st, c, w, off[0])) return rc;
            if (int rc = diffraction(h, m, m.P, m.propP)) return rc;
            if (int rc = diffraction(h, m, m.S, m.propS)) return rc;
            if (int rc = rk4(h, m, st, c, w, off[1])) return rc;
            if (m.thermal) {
                const int hs = st.h;
                const bool first = s == 0 || (int)(((long long)(s - 1) * m.nzh) / m.nz) != hs;
                const long long a = ((long long)(hs + 1) * m.nz + m.nzh - 1) / m.nzh;
                const long long b = ((long long)hs * m.nz + m.nzh - 1) / m.nzh;
                ProfScope pq(h, "plain_setq");
                SHG_HIP(launch_setq(h->stream, m.n2, st, m.P, m.S, m.heat.Q, first, 1.0 / (double)(a - b)));
            }
            steps++;
        }
        if (int rc = sum_sq64(h, m.outP, m.n2, m.slots + 1)) return rc;
        if (int rc = sum_sq64(h, m.outS, m.n2, m.slots + 2)) return rc;
        outer = go + 1;
        if (m.thermal && info.converged_in_one && go != 0) break;
    }
    double sl[3];
    SHG_HIP(hipStreamSynchronize(h->stream));
    SHG_HIP(hipMemcpy(sl, m.slots, sizeof(sl), hipMemcpyDeviceToHost));
    const double np = refr_index<double>(p->crystal, lp, Tc);
    const double ns = refr_index<double>(p->crystal, ls, Tc);
    const double cc = 0.5 * kEps0 * kC * m.dx * m.dy;
    r->P_pump_in_W = cc * np * sl[0];
    r->P_pump_out_W = cc * np * sl[1];
    r->P_sh_W = cc * ns * sl[2];
    r->eta = p->power_W > 0 ? r->P_sh_W / p->power_W : 0.0;
    r->steps_done = steps;
    r->jacobi_sweeps = sweeps;
    r->outer_iters = outer;
    r->heat_residual_K = heat_res;
    r->peak_dT_K = 0;
    r->T_max_C = Tc;
    if (!std::isfinite(sl[0]) || !std::isfinite(sl[1]) || !std::isfinite(sl[2])) {
        set_error("non-finite field power");
        return SHG_ENAN;
    }
    if (r->pump_out) SHG_HIP(hipMemcpy(r->pump_out, m.outP, sizeof(double2) * m.n2, hipMemcpyDeviceToHost));
    if (r->signal_out) SHG_HIP(hipMemcpy(r->signal_out, m.outS, sizeof(double2) * m.n2, hipMemcpyDeviceToHost));
    if (m.thermal)
        if (int rc = heat_finish(h, p, m.heat, r)) return rc;
    r->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return SHG_OK;
}

}  // namespace shg
