// march_mixed.cpp -- host orchestration of the fused forms (FP32 fields
// with FP32 or FP64 arithmetic, FP64 fields): instance set-up, the batched
// z-march (Solver::CWES, src/headers/Solver.h:276-308) and, with THERMAL,
// the thermo-optic outer loop (src/headers/Solver.h:321-373).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "heat.hpp"
#include "kernels.hpp"
#include "march_mixed.hpp"
#include "runtime.hpp"

namespace shg {

Arith arith_of(int precision) {
    switch (precision) {
        case SHG_PREC_FP32_STORAGE: return Arith::S32;
        case SHG_PREC_FP64_FUSED: return Arith::F64;
        case SHG_PREC_FP32_SPLIT: return Arith::R32;
        default: return Arith::F32;
    }
}

// profile names: the pass and the arithmetic it runs in
static const char* arith_suffix(Arith a) {
    return a == Arith::F32 ? "" : a == Arith::S32 ? "_f64" : "_fp64fields";
}
const char* rowpass_name(int nx, bool th, Arith a) {
    static std::map<std::tuple<int, bool, int>, std::string> names;
    a = row_arith(a);
    auto& s = names[{nx, th, (int)a}];
    if (s.empty()) s = std::string("rowpass_") + std::to_string(nx) + (th ? "_thermal" : "") + arith_suffix(a);
    return s.c_str();
}
const char* colpass_name(int ny, Arith a) {
    static std::map<std::pair<int, int>, std::string> names;
    a = col_arith(a);
    auto& s = names[{ny, (int)a}];
    if (s.empty()) s = "colpass_" + std::to_string(ny) + arith_suffix(a);
    return s.c_str();
}

int plan_rows(shg_handle* h, int nx, Arith a, long lines, bool thermal) {
    if (thermal) {
        auto th = h->rows_plan_th.find({nx, (int)a});
        if (th != h->rows_plan_th.end() && th->second > 0) return th->second;
    }
    auto it = h->rows_plan.find({nx, (int)a});
    return it != h->rows_plan.end() && it->second > 0 ? it->second : rowpass_rows_auto(nx, a, lines, thermal);
}
int plan_cols(shg_handle* h, int ny, Arith a, long lines) {
    auto it = h->cols_plan.find({ny, (int)a});
    return it != h->cols_plan.end() && it->second > 0 ? it->second : colpass_cols_auto(ny, a, lines);
}

int field_sum_sq(shg_handle* h, Arith a, const void* A, long n, double* slot) {
    return a == Arith::F64 ? sum_sq64(h, static_cast<const double2*>(A), n, slot)
                           : sum_sq(h, static_cast<const float2*>(A), n, slot);
}

int field_to_host(shg_handle* h, Arith a, const void* A, size_t n, double* out) {
    if (a == Arith::F64) {
        SHG_HIP(hipMemcpyAsync(out, A, sizeof(double2) * n, hipMemcpyDeviceToHost, h->stream));
        SHG_HIP(hipStreamSynchronize(h->stream));
        return SHG_OK;
    }
    std::vector<float2> buf(n);
    SHG_HIP(hipMemcpyAsync(buf.data(), A, sizeof(float2) * n, hipMemcpyDeviceToHost, h->stream));
    SHG_HIP(hipStreamSynchronize(h->stream));
    for (size_t i = 0; i < n; i++) {
        out[2 * i] = buf[i].x;
        out[2 * i + 1] = buf[i].y;
    }
    return SHG_OK;
}

// Host-side FP64 tables of one instance at centre temperature Tc:
// separable propagator (src/headers/Efields.h:122-141 with the fftshift of
// :68-81 folded into the index map, and the 1/(nx ny) of cuFFTscale folded
// into the x factor), per-slice QPM phases, couplings and RK4 constants.
void host_tables(MixedHost& m, std::vector<double2>& ex_p, std::vector<double2>& ex_s,
                        std::vector<double2>& ey_p, std::vector<double2>& ey_s,
                        std::vector<double2>& phase) {
    ex_p.resize(m.nx);
    ex_s.resize(m.nx);
    ey_p.resize(m.ny);
    ey_s.resize(m.ny);
    phase.resize(m.thermal ? 0 : (size_t)m.nz * 6);
    host_tables_into(m, ex_p.data(), ex_s.data(), ey_p.data(), ey_s.data(), m.thermal ? nullptr : phase.data());
}

void host_tables_into(MixedHost& m, double2* ex_p, double2* ex_s, double2* ey_p, double2* ey_s, double2* phase) {
    const shg_params* p = m.p;
    const int nx = m.nx, ny = m.ny, nz = m.nz;
    const double lp = p->lambda_p_um, ls = 0.5 * lp;
    const double np = refr_index<double>(p->crystal, lp, m.Tc);
    const double ns = refr_index<double>(p->crystal, ls, m.Tc);
    const double kwp = 2 * kPi * np / lp, kws = 2 * kPi * ns / ls;
    const double dfX = 1.0 / m.dx / nx, dfY = 1.0 / m.dy / ny;
    const double scale = 1.0 / ((double)nx * ny);
    for (int x = 0; x < nx; x++) {
        const int ic = (x + nx / 2) % nx;
        const double fx = (ic - 0.5 * nx) * dfX;
        const double ap = -m.dz * (2 * kPi * kPi / kwp * fx * fx + 2 * kPi * fx * std::tan(p->rho_p));
        const double as = -m.dz * (2 * kPi * kPi / kws * fx * fx + 2 * kPi * fx * std::tan(p->rho_s));
        ex_p[x] = make_double2(scale * std::cos(ap), scale * std::sin(ap));
        ex_s[x] = make_double2(scale * std::cos(as), scale * std::sin(as));
    }
    for (int y = 0; y < ny; y++) {
        const int jc = (y + ny / 2) % ny;
        const double fy = (jc - 0.5 * ny) * dfY;
        const double ap = -m.dz * (2 * kPi * kPi / kwp * fy * fy);
        const double as = -m.dz * (2 * kPi * kPi / kws * fy * fy);
        ey_p[y] = make_double2(std::cos(ap), std::sin(ap));
        ey_s[y] = make_double2(std::cos(as), std::sin(as));
    }
    // scheme constants (src/headers/Solver.h:235-256 or textbook Strang)
    MixedInst& d = m.desc;
    double kmul = 1.0;
    double off[2][3] = {{0.0, 0.25, 0.5}, {0.0, 0.25, 0.5}};
    if (p->scheme == SHG_SCHEME_STRANG) {
        const double hh = 0.5 * m.dz;
        d.c1 = 0.5 * hh;
        d.c2 = 0.5 * hh;
        d.c3 = hh;
        d.w = hh / 6.0;
        kmul = 2.0;
        off[1][0] = 0.5;
        off[1][1] = 0.75;
        off[1][2] = 1.0;
    } else {
        d.c1 = 0.5;
        d.c2 = 0.5;
        d.c3 = 1.0;
        d.w = m.dz / 6.0;
    }
    for (int c = 0; c < 2; c++)
        for (int i = 0; i < 3; i++) d.off[c][i] = off[c][i];
    const double dQ = 2.0 * p->deff_um_per_V / kPi;
    d.kp = kmul * kPi * dQ / (np * lp);
    d.ks = kmul * 0.5 * kPi * dQ / (ns * ls);
    d.bfac = p->beta_s * 0.5 * kC * kEps0 * ns;
    d.kp0 = kmul * kPi * dQ / lp;
    d.ks0 = kmul * 0.5 * kPi * dQ / ls;
    d.bf0 = p->beta_s * 0.5 * kC * kEps0;
    d.ap_h = 0.5 * p->alpha_p;
    d.as_h = 0.5 * p->alpha_s;
    d.hce = 0.5 * kC * kEps0;
    d.alpha_p = p->alpha_p;
    d.alpha_s = p->alpha_s;
    d.beta_s = p->beta_s;
    d.crystal = p->crystal;
    d.dz = m.dz;
    d.nx = nx;
    d.ny = ny;
    d.nz = nz;
    d.nzh = m.nzh;
    // non-THERMAL phase table: exp(i ((s + off) dz dk + poling[s])), FP64
    if (!m.thermal) {
        const double dk = delta_k(p->crystal, lp, p->Lambda_um, p->therm_exp_a, p->therm_exp_b, m.Tc);
        for (int s = 0; s < nz; s++) {
            const double pol = p->poling_phase ? p->poling_phase[s] : 0.0;
            for (int c = 0; c < 2; c++)
                for (int i = 0; i < 3; i++) {
                    const double ph = ((double)s + off[c][i]) * m.dz * dk + pol;
                    phase[(size_t)(s * 2 + c) * 3 + i] = make_double2(std::cos(ph), std::sin(ph));
                }
        }
    }
}

// Upload tables, initialise the pump (src/headers/Efields.h:84-119) and a
// zero SH field, and reduce the initial pump power.
int setup_fields(shg_handle* h, MixedHost& m) {
    std::vector<double2> exp_, exs, eyp, eys;
    std::vector<double2> ph;
    host_tables(m, exp_, exs, eyp, eys, ph);
    SHG_HIP(hipMemcpyAsync(m.exP, exp_.data(), sizeof(double2) * m.nx, hipMemcpyHostToDevice, h->stream));
    SHG_HIP(hipMemcpyAsync(m.exS, exs.data(), sizeof(double2) * m.nx, hipMemcpyHostToDevice, h->stream));
    SHG_HIP(hipMemcpyAsync(m.eyP, eyp.data(), sizeof(double2) * m.ny, hipMemcpyHostToDevice, h->stream));
    SHG_HIP(hipMemcpyAsync(m.eyS, eys.data(), sizeof(double2) * m.ny, hipMemcpyHostToDevice, h->stream));
    if (!ph.empty())
        SHG_HIP(hipMemcpyAsync(m.phase, ph.data(), sizeof(double2) * ph.size(), hipMemcpyHostToDevice, h->stream));
    // the host vectors die at return: make the copies complete first
    SHG_HIP(hipStreamSynchronize(h->stream));
    const shg_params* p = m.p;
    const PumpCoef pc = pump_coef(p, refr_index<double>(p->crystal, p->lambda_p_um, m.Tc));
    {
        ProfScope ps(h, "pump_init");
        if (m.arith == Arith::F64)
            SHG_HIP(launch_pump_init<double2>(h->stream, static_cast<double2*>(m.P), m.nx, m.ny, m.dx, m.dy, pc));
        else
            SHG_HIP(launch_pump_init<float2>(h->stream, static_cast<float2*>(m.P), m.nx, m.ny, m.dx, m.dy, pc));
    }
    SHG_HIP(hipMemsetAsync(m.S, 0, arith_field_bytes(m.arith) * m.N2, h->stream));
    return field_sum_sq(h, m.arith, m.P, (long)m.N2, m.slots + 0);
}

// Copies the traced state of slice g (written by row pass g) to the host
// as FP64 and hands it to the trace callback.
static int deliver_slice(shg_handle* h, MixedHost& m, int g) {
    SliceTrace& t = *m.trace;
    t.pump.resize(2 * m.N2);
    t.signal.resize(2 * m.N2);
    if (int rc = field_to_host(h, m.arith, m.trP, m.N2, t.pump.data())) return rc;
    if (int rc = field_to_host(h, m.arith, m.trS, m.N2, t.signal.data())) return rc;
    return t.deliver(g, m.pass);
}

// Instances sharing one launch of a pass: same transform length (and
// THERMAL flag for the row pass) and form, descriptors at d, sorted by nz
// descending.
struct Group {
    int key_n;
    bool thermal;
    Arith arith;
    int half = 0;  // interleave half: 1 runs its column pass half a step behind
    std::vector<MixedHost*> m;
    MixedInst* d = nullptr;
    const double2* tw = nullptr;
};
using RowGroups = std::map<std::tuple<int, bool, int, int>, Group>;
using ColGroups = std::map<std::tuple<int, int, int>, Group>;

// Interleaved halves.  The row pass is bound by vector-ALU issue and the
// column pass by memory and LDS latency, so the two overlap well when they
// run side by side.  Within one instance they cannot (each needs the
// other's whole output), so a large class of like instances (same grid,
// THERMAL flag and form) is split into two halves on two streams, and half
// 1 runs its column pass of step g-1 beside half 0's row pass of step g:
//   half 0:  R(g)   C(g)   R(g+1) C(g+1) ...
//   half 1:  C(g-1) R(g)   C(g)   R(g+1) ...
// Every instance still sees its own passes in order, and every pass of an
// instance computes the same values in any launch that carries it, so the
// results are those of one launch group.  A class is split when each half
// holds at least kHalfMinPoints grid points; traced and profiled marches
// (one stream) are not split.
static constexpr size_t kHalfMinPoints = size_t(1) << 20;

static void assign_halves(std::vector<MixedHost*>& act, bool allowed) {
    std::map<std::tuple<int, int, bool, int>, std::vector<MixedHost*>> cls;
    for (auto* m : act) {
        m->half = 0;
        cls[{m->nx, m->ny, m->thermal, (int)m->arith}].push_back(m);
    }
    if (!allowed) return;
    for (auto& kv : cls) {
        auto& v = kv.second;
        size_t pts = 0;
        for (auto* m : v) pts += m->N2;
        if (v.size() < 2 || pts < 2 * kHalfMinPoints) continue;
        // alternate down the nz order, so both halves march about as long
        std::stable_sort(v.begin(), v.end(), [](const MixedHost* a, const MixedHost* b) { return a->nz > b->nz; });
        for (size_t i = 0; i < v.size(); i++) v[i]->half = (int)(i & 1);
    }
}

// Launch groups that share no instance, directly or through a chain of
// groups, are independent: each such component runs on its own stream (up
// to four: the handle's two and two more created on first use), forked
// from and joined back to the handle's stream.
struct Lanes {
    hipStream_t streams[4] = {};
    std::map<const Group*, int> lane;
    int n = 1;
};
static int march_lanes(shg_handle* h, RowGroups& rowg, ColGroups& colg, bool single, Lanes& L) {
    std::vector<const Group*> gs;
    std::map<const Group*, int> idx;
    for (auto& kv : rowg) idx[&kv.second] = (int)gs.size(), gs.push_back(&kv.second);
    for (auto& kv : colg) idx[&kv.second] = (int)gs.size(), gs.push_back(&kv.second);
    std::vector<int> par(gs.size());
    for (size_t i = 0; i < par.size(); i++) par[i] = (int)i;
    std::function<int(int)> find = [&](int i) { return par[i] == i ? i : par[i] = find(par[i]); };
    for (auto& kv : rowg)
        for (auto* m : kv.second.m) {
            const int a = find(idx[&kv.second]);
            const int b = find(idx[&colg.at({m->ny, (int)m->arith, m->half})]);
            par[a] = b;
        }
    std::map<int, int> comp;
    for (size_t i = 0; i < gs.size(); i++) {
        const int r = find((int)i);
        if (!comp.count(r)) comp[r] = (int)comp.size();
    }
    L.n = single ? 1 : std::min<int>((int)comp.size(), 4);
    for (size_t i = 0; i < gs.size(); i++) L.lane[gs[i]] = comp[find((int)i)] % L.n;
    L.streams[0] = h->stream;
    L.streams[1] = h->stream2;
    if (L.n > 1) {
        for (int l = 2; l < L.n; l++) {
            if (!h->lanes[l - 2]) SHG_HIP(hipStreamCreateWithFlags(&h->lanes[l - 2], hipStreamNonBlocking));
            L.streams[l] = h->lanes[l - 2];
        }
        if (!h->fork_ev) SHG_HIP(hipEventCreateWithFlags(&h->fork_ev, hipEventDisableTiming));
        for (auto& e : h->join_ev)
            if (!e) SHG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    return SHG_OK;
}

// Graph replay of the march: the launches of `chunk` consecutive steps
// (every group's row pass, then every group's column pass, per step) and
// a counter advance are captured once into a graph whose kernels read the
// step index from the handle's device counter; the graph is replayed
// ceil((G + 1) / chunk) times.  Every launch covers all instances of its
// group (an instance past its last step returns at once), so the captured
// shapes do not change along the march and the graph is cached on the
// handle, keyed by those shapes and pointers, for later marches.
static int march_graph(shg_handle* h, RowGroups& rowg, ColGroups& colg, int G, int C) {
    if (!h->step_counter) SHG_HIP(hipMalloc(&h->step_counter, 256));
    Lanes L;
    if (int rc = march_lanes(h, rowg, colg, false, L)) return rc;
    struct Launch {
        bool row;
        const Group* gr;
        int plan;
        dim3 grid;
        int lane;
    };
    std::vector<Launch> ls;
    std::vector<uint64_t> key{(uint64_t)C, (uint64_t)(uintptr_t)h->step_counter, (uint64_t)L.n};
    auto add = [&](bool row, const Group& gr) {
        int maxn = 0;
        for (auto* m : gr.m) maxn = std::max(maxn, row ? m->ny : m->nx);
        long lines = 0;
        for (auto* m : gr.m) lines += row ? m->ny : m->nx;
        const int plan = row ? plan_rows(h, gr.key_n, gr.arith, lines, gr.thermal)
                             : plan_cols(h, gr.key_n, gr.arith, lines);
        const int per = row ? rowpass_rows_per_wg(gr.key_n, plan) : plan;
        const dim3 grid((maxn + per - 1) / per, row ? 1 : 2, (unsigned)gr.m.size());
        const int ln = L.lane.at(&gr);
        ls.push_back({row, &gr, plan, grid, ln});
        for (uint64_t v : {(uint64_t)row, (uint64_t)gr.key_n, (uint64_t)gr.thermal, (uint64_t)gr.arith,
                           (uint64_t)plan, (uint64_t)grid.x, (uint64_t)grid.z, (uint64_t)(uintptr_t)gr.d,
                           (uint64_t)(uintptr_t)gr.tw, (uint64_t)ln, (uint64_t)gr.half})
            key.push_back(v);
    };
    for (auto& kv : rowg) add(true, kv.second);
    for (auto& kv : colg) add(false, kv.second);
    hipGraphExec_t ex = nullptr;
    auto it = h->graphs.find(key);
    if (it != h->graphs.end()) {
        ex = it->second;
    } else {
        if (h->graphs.size() >= 16) {
            for (auto& kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
            h->graphs.clear();
        }
        // independent groups become parallel branches of the graph (fork
        // from and join into the capturing stream); the counter advance
        // follows every branch
        SHG_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        if (L.n > 1) {
            e = hipEventRecord(h->fork_ev, h->stream);
            for (int l = 1; l < L.n && e == hipSuccess; l++) e = hipStreamWaitEvent(L.streams[l], h->fork_ev, 0);
        }
        // per step: half 0's row passes, half 1's column passes of the
        // previous step (step -1 returns at once), half 0's column passes,
        // half 1's row passes
        const struct { bool row; int half, lag; } order[4] = {{true, 0, 0}, {false, 1, 1}, {false, 0, 0},
                                                              {true, 1, 0}};
        for (int c = 0; c < C && e == hipSuccess; c++)
            for (const auto& o : order)
                for (const Launch& l : ls) {
                    if (l.row != o.row || l.gr->half != o.half) continue;
                    const Step st(h->step_counter, c - o.lag);
                    hipStream_t s = L.streams[l.lane];
                    e = l.row ? launch_rowpass(l.gr->key_n, l.gr->thermal, l.gr->arith, l.plan, l.grid, s, l.gr->d,
                                               st, l.gr->tw)
                              : launch_colpass(l.gr->key_n, l.gr->arith, l.plan, l.grid, s, l.gr->d, st, l.gr->tw);
                    if (e != hipSuccess) break;
                }
        for (int l = 1; l < L.n && e == hipSuccess; l++) {
            e = hipEventRecord(h->join_ev[l - 1], L.streams[l]);
            if (e == hipSuccess) e = hipStreamWaitEvent(h->stream, h->join_ev[l - 1], 0);
        }
        if (e == hipSuccess) e = launch_step_advance(h->stream, h->step_counter, C);
        hipGraph_t graph = nullptr;
        const hipError_t e2 = hipStreamEndCapture(h->stream, &graph);
        if (e == hipSuccess) e = e2;
        if (e == hipSuccess) e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            set_error("march graph capture: %s", hipGetErrorString(e));
            return SHG_EDEVICE;
        }
        h->graphs[key] = ex;
    }
    SHG_HIP(launch_step_set(h->stream, h->step_counter, 0));
    for (int r = 0; r < (G + C) / C; r++) SHG_HIP(hipGraphLaunch(ex, h->stream));
    return SHG_OK;
}

// The batched z-march: launches A(0) B(0) A(1) ... B(G-1) A(G), G = max nz.
static int march(shg_handle* h, std::vector<MixedHost*>& act, MixedInst* pool, size_t pool_cap) {
    size_t pool_used = 0;
    RowGroups rowg;
    ColGroups colg;
    const bool traced = std::any_of(act.begin(), act.end(), [](const MixedHost* m) { return m->trace != nullptr; });
    assign_halves(act, !traced && !h->profiling);
    for (auto* m : act) {
        auto& rg = rowg[{m->nx, m->thermal, (int)m->arith, m->half}];
        rg.key_n = m->nx;
        rg.thermal = m->thermal;
        rg.arith = m->arith;
        rg.half = m->half;
        rg.m.push_back(m);
        auto& cg = colg[{m->ny, (int)m->arith, m->half}];
        cg.key_n = m->ny;
        cg.thermal = false;
        cg.arith = m->arith;
        cg.half = m->half;
        cg.m.push_back(m);
    }
    auto by_nz = [](const MixedHost* a, const MixedHost* b) { return a->nz > b->nz; };
    std::vector<std::vector<MixedInst>> hostdesc;
    auto upload = [&](Group& g) -> int {
        std::sort(g.m.begin(), g.m.end(), by_nz);
        hostdesc.emplace_back();
        auto& v = hostdesc.back();
        for (auto* m : g.m) v.push_back(m->desc);
        g.d = pool_used + v.size() <= pool_cap ? pool + pool_used : nullptr;
        pool_used += v.size();
        if (!g.d) {
            set_error("arena exhausted (descriptors)");
            return SHG_ENOMEM;
        }
        SHG_HIP(hipMemcpyAsync(g.d, v.data(), sizeof(MixedInst) * v.size(), hipMemcpyHostToDevice, h->stream));
        g.tw = twiddles64(h, g.key_n);
        if (!g.tw) {
            set_error("twiddle table allocation failed");
            return SHG_ENOMEM;
        }
        return SHG_OK;
    };
    for (auto& kv : rowg)
        if (int rc = upload(kv.second)) return rc;
    for (auto& kv : colg)
        if (int rc = upload(kv.second)) return rc;
    int G = 0;
    for (auto* m : act) G = std::max(G, m->nz);
    Lanes L;
    if (int rc = march_lanes(h, rowg, colg, traced || h->profiling, L)) return rc;
    // graph replay: as set, or automatically (32 steps per graph) when the
    // march has independent lanes, where the replayed fork/join measured
    // faster than direct launches
    const int chunk = h->graph_chunk >= 0 ? h->graph_chunk : (L.n > 1 ? 32 : 0);
    if (chunk > 0 && !h->profiling && !traced) {
        if (int rc = march_graph(h, rowg, colg, G, chunk)) return rc;
        SHG_HIP(hipStreamSynchronize(h->stream));
        for (auto* m : act) m->steps += m->nz;
        return SHG_OK;
    }
    hipStream_t* streams = L.streams;
    auto& lane = L.lane;
    const int nlanes = L.n;
    if (nlanes > 1) {
        SHG_HIP(hipEventRecord(h->fork_ev, h->stream));
        for (int l = 1; l < nlanes; l++) SHG_HIP(hipStreamWaitEvent(streams[l], h->fork_ev, 0));
    }
    auto row_pass = [&](Group& gr, int g) -> int {
        hipStream_t st = streams[lane[&gr]];
        int cnt = 0, maxny = 0;
        long lines = 0;
        for (auto* m : gr.m) {
            if (m->nz < g) break;
            cnt++;
            maxny = std::max(maxny, m->ny);
            lines += m->ny;
        }
        if (!cnt) return SHG_OK;
        const int rows = plan_rows(h, gr.key_n, gr.arith, lines, gr.thermal);
        const int per = rowpass_rows_per_wg(gr.key_n, rows);
        dim3 grid((maxny + per - 1) / per, 1, cnt);
        ProfScope ps(h, rowpass_name(gr.key_n, gr.thermal, gr.arith), st, true);
        SHG_HIP(launch_rowpass(gr.key_n, gr.thermal, gr.arith, rows, grid, st, gr.d, g, gr.tw));
        return SHG_OK;
    };
    auto col_pass = [&](Group& gr, int g) -> int {
        hipStream_t st = streams[lane[&gr]];
        int cnt = 0, maxnx = 0;
        long lines = 0;
        for (auto* m : gr.m) {
            if (m->nz <= g) break;
            cnt++;
            maxnx = std::max(maxnx, m->nx);
            lines += m->nx;
        }
        if (!cnt) return SHG_OK;
        const int cols = plan_cols(h, gr.key_n, gr.arith, lines);
        dim3 grid((maxnx + cols - 1) / cols, 2, cnt);
        ProfScope ps(h, colpass_name(gr.key_n, gr.arith), st, true);
        SHG_HIP(launch_colpass(gr.key_n, gr.arith, cols, grid, st, gr.d, g, gr.tw));
        return SHG_OK;
    };
    // per step: half 0's row passes, half 1's column passes of the previous
    // step, half 0's column passes, half 1's row passes (assign_halves)
    for (int g = 0; g <= G; g++) {
        for (auto& kv : rowg)
            if (kv.second.half == 0)
                if (int rc = row_pass(kv.second, g)) return rc;
        for (auto* m : act)
            if (m->trace && m->trace->wants(g, m->nz))
                if (int rc = deliver_slice(h, *m, g)) return rc;
        if (g > 0)
            for (auto& kv : colg)
                if (kv.second.half == 1)
                    if (int rc = col_pass(kv.second, g - 1)) return rc;
        if (g < G)
            for (auto& kv : colg)
                if (kv.second.half == 0)
                    if (int rc = col_pass(kv.second, g)) return rc;
        for (auto& kv : rowg)
            if (kv.second.half == 1)
                if (int rc = row_pass(kv.second, g)) return rc;
    }
    for (int l = 1; l < nlanes; l++) {
        SHG_HIP(hipEventRecord(h->join_ev[l - 1], streams[l]));
        SHG_HIP(hipStreamWaitEvent(h->stream, h->join_ev[l - 1], 0));
    }
    // descriptors' host copies must outlive the async copies
    SHG_HIP(hipStreamSynchronize(h->stream));
    for (auto* m : act) m->steps += m->nz;
    return SHG_OK;
}

size_t mixed_bytes(const MixedHost& m) {
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t fb = arith_field_bytes(m.arith);
    size_t b = 4 * al(fb * m.N2) + 2 * al(sizeof(double2) * m.nx) +
               2 * al(sizeof(double2) * m.ny) + al(sizeof(double2) * (size_t)m.nz * 6) + al(sizeof(double) * 4) +
               2 * al(sizeof(MixedInst));
    if (m.thermal) b += heat_bytes(m.p);
    if (m.trace) b += 2 * al(fb * m.N2);
    return b;
}

static int run_mixed_impl(shg_handle* h, const std::vector<const shg_params*>& ps,
                          const std::vector<shg_result*>& rs, std::vector<MixedHost>& ms);

int run_mixed(shg_handle* h, const std::vector<const shg_params*>& ps, const std::vector<shg_result*>& rs,
              SliceTrace* trace) {
    std::vector<MixedHost> ms(ps.size());
    if (trace) {
        if (ps.size() != 1) {
            set_error("a per-slice trace needs a single run");
            return SHG_EINVAL;
        }
        ms[0].trace = trace;
    }
    const int rc = run_mixed_impl(h, ps, rs, ms);
    for (auto& m : ms) heat_release(m.heat);
    // a failure that ends the whole batch leaves the instances it did not
    // finish without results: each of them reports that failure
    if (rc != SHG_OK)
        for (size_t i = 0; i < ms.size(); i++)
            if (!ms[i].finalised && rs[i]->status == SHG_OK) rs[i]->status = rc;
    return rc;
}

static int run_mixed_impl(shg_handle* h, const std::vector<const shg_params*>& ps,
                          const std::vector<shg_result*>& rs, std::vector<MixedHost>& ms) {
    const auto t0 = std::chrono::steady_clock::now();
    const int n = (int)ps.size();
    size_t total = 0;
    for (int i = 0; i < n; i++) {
        MixedHost& m = ms[i];
        m.p = ps[i];
        m.r = rs[i];
        m.nx = m.p->nx;
        m.ny = m.p->ny;
        m.nz = m.p->nz;
        m.nzh = m.p->nz_heat > 0 ? m.p->nz_heat : m.p->nz;
        m.N2 = (size_t)m.nx * m.ny;
        m.dx = m.p->lx_um / (m.nx - 1);
        m.dy = m.p->ly_um / (m.ny - 1);
        m.dz = m.p->lz_um / (m.nz - 1);
        m.thermal = m.p->thermal != 0;
        m.arith = arith_of(m.p->precision);
        m.Tc = m.p->T_C;
        total += mixed_bytes(m);
    }
    total += 2 * sizeof(MixedInst) * (size_t)n + 4096;
    if (int rc = h->arena.reserve(total)) return rc;
    for (auto& m : ms) {
        const size_t fb = arith_field_bytes(m.arith);
        m.P = h->arena.carve<char>(fb * m.N2);
        m.S = h->arena.carve<char>(fb * m.N2);
        m.outP = h->arena.carve<char>(fb * m.N2);
        m.outS = h->arena.carve<char>(fb * m.N2);
        m.exP = h->arena.carve<double2>(m.nx);
        m.exS = h->arena.carve<double2>(m.nx);
        m.eyP = h->arena.carve<double2>(m.ny);
        m.eyS = h->arena.carve<double2>(m.ny);
        m.phase = h->arena.carve<double2>((size_t)m.nz * 6);
        if (m.trace) {
            m.trP = h->arena.carve<char>(fb * m.N2);
            m.trS = h->arena.carve<char>(fb * m.N2);
        }
        m.slots = h->arena.carve<double>(4);
        if (!m.slots) {
            set_error("arena exhausted");
            return SHG_ENOMEM;
        }
        if (m.thermal) {
            if (int rc = heat_alloc(h, m.p, m.heat)) return rc;
        }
        MixedInst& d = m.desc;
        d.P = m.P;
        d.S = m.S;
        d.outP = m.outP;
        d.outS = m.outS;
        d.trP = m.trP;
        d.trS = m.trS;
        d.trace_every = m.trace ? m.trace->every : 1;
        d.exP = m.exP;
        d.exS = m.exS;
        d.eyP = m.eyP;
        d.eyS = m.eyS;
        d.phase = m.phase;
        if (m.thermal) {
            d.nn = m.heat.nn;
            d.dkbase = m.heat.dkbase;
            d.dk = m.heat.dk;
            d.Q = m.heat.Q;
            d.poling = m.heat.pol;
        }
    }

    MixedInst* pool = h->arena.carve<MixedInst>(2 * (size_t)n);
    if (!pool) {
        set_error("arena exhausted (descriptor pool)");
        return SHG_ENOMEM;
    }
    std::vector<MixedHost*> act;
    for (auto& m : ms)
        if (!m.thermal) {
            if (int rc = setup_fields(h, m)) return rc;
            act.push_back(&m);
        }
    // non-THERMAL instances march once; THERMAL ones iterate the outer loop
    std::vector<MixedHost*> th;
    for (auto& m : ms)
        if (m.thermal) {
            if (int rc = heat_init(h, m.p, m.heat)) return rc;
            th.push_back(&m);
        }
    int max_outer = 0;
    for (auto* m : th) max_outer = std::max(max_outer, m->p->max_outer > 0 ? m->p->max_outer : 20);
    for (int go = 0; go <= max_outer; go++) {
        // heat solve + optics set-up of the THERMAL instances still running
        for (auto* m : th) {
            if (!m->running) continue;
            if (go >= (m->p->max_outer > 0 ? m->p->max_outer : 20)) {
                m->running = false;
                continue;
            }
            HeatSolveInfo info{};
            int rc = heat_solve(h, m->p, m->heat, info);
            if (rc) {
                m->r->status = rc;
                m->running = false;
                continue;
            }
            m->sweeps += info.sweeps;
            m->heat_res = info.residual;
            m->heat.converged_in_one = info.converged_in_one;
            rc = heat_prepare_optics(h, m->p, m->heat, &m->Tc);
            if (rc) return rc;
            m->desc.nn = m->heat.nn;
            m->pass = go;
            if ((rc = setup_fields(h, *m))) return rc;
            act.push_back(m);
        }
        if (act.empty()) break;
        if (int rc = march(h, act, pool, 2 * (size_t)n)) return rc;
        for (auto* m : act) {
            if (!m->thermal) continue;
            m->outer = go + 1;
            // src/headers/Solver.h:367: stop after a pass whose Jacobi loop took one sweep
            if (m->heat.converged_in_one && go != 0) m->running = false;
        }
        // the saved output state of this pass is kept in outP/outS; reduce now
        for (auto* m : act) {
            if (int rc = field_sum_sq(h, m->arith, m->outP, (long)m->N2, m->slots + 1)) return rc;
            if (int rc = field_sum_sq(h, m->arith, m->outS, (long)m->N2, m->slots + 2)) return rc;
        }
        act.clear();
        if (th.empty()) break;
        bool any = false;
        for (auto* m : th) any |= m->running;
        if (!any) break;
    }

    // results
    SHG_HIP(hipStreamSynchronize(h->stream));
    for (auto& m : ms) {
        if (m.r->status != SHG_OK) continue;
        double sl[4];
        SHG_HIP(hipMemcpy(sl, m.slots, sizeof(double) * 3, hipMemcpyDeviceToHost));
        const shg_params* p = m.p;
        const double np = refr_index<double>(p->crystal, p->lambda_p_um, m.Tc);
        const double ns = refr_index<double>(p->crystal, 0.5 * p->lambda_p_um, m.Tc);
        const double c = 0.5 * kEps0 * kC * m.dx * m.dy;
        shg_result* r = m.r;
        r->P_pump_in_W = c * np * sl[0];
        r->P_pump_out_W = c * np * sl[1];
        r->P_sh_W = c * ns * sl[2];
        r->eta = p->power_W > 0 ? r->P_sh_W / p->power_W : 0.0;
        r->steps_done = m.steps;
        r->jacobi_sweeps = m.sweeps;
        r->outer_iters = m.thermal ? m.outer : 1;
        r->heat_residual_K = m.heat_res;
        r->peak_dT_K = 0;
        r->T_max_C = m.Tc;
        if (!std::isfinite(sl[0]) || !std::isfinite(sl[1]) || !std::isfinite(sl[2])) {
            set_error("non-finite field power");
            r->status = SHG_ENAN;
            continue;
        }
        if (r->pump_out)
            if (int rc = field_to_host(h, m.arith, m.outP, m.N2, r->pump_out)) return rc;
        if (r->signal_out)
            if (int rc = field_to_host(h, m.arith, m.outS, m.N2, r->signal_out)) return rc;
        if (m.thermal) {
            if (int rc = heat_finish(h, p, m.heat, r)) return rc;
        }
        r->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        m.finalised = true;
    }
    return SHG_OK;
}

}  // namespace shg
