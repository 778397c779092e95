// plans.cpp -- launch plans of the fused passes (shg_set_plan, shg_get_plan),
// the autotuner that chooses them (shg_autotune) and the graph-replay
// setting (shg_set_graph_steps).
//
// A plan is the tile shape of a pass for one transform length and form:
// rows per workgroup of the row pass (1, 2, 4, 8, or the wide form with one
// point per thread) and columns per workgroup of the column pass (2, 4, 8,
// 16).  The radix sequence of the transform is
// fixed by its length ({8, n/64, 8}).  The autotuner times every supported
// candidate on a synthetic batched march of the caller's shape, with HIP
// events around each launch, and keeps the fastest on the handle; the row
// pass is timed a second time on a THERMAL march, whose voxel loads can
// shift its best tile, and that choice is kept for THERMAL launches.
#include <algorithm>
#include <cfloat>
#include <vector>

#include "heat.hpp"
#include "kernels.hpp"
#include "march_mixed.hpp"
#include "runtime.hpp"
using namespace shg;

namespace {

bool fused_form(int precision) {
    return precision == SHG_PREC_FP32_MIXED || precision == SHG_PREC_FP32_STORAGE || precision == SHG_PREC_FP64_FUSED ||
           precision == SHG_PREC_FP32_SPLIT;
}

// mean launch time [ms] of kernel class `name` over one march of `ps`, or
// a negative value if the march failed
double time_pass(shg_handle* h, const std::vector<shg_params>& ps, const char* name) {
    std::vector<const shg_params*> pp;
    std::vector<shg_result> rs(ps.size());
    std::vector<shg_result*> rp;
    for (size_t i = 0; i < ps.size(); i++) {
        pp.push_back(&ps[i]);
        rp.push_back(&rs[i]);
    }
    prof_reset(h);
    h->profiling = true;
    const int rc = run_mixed(h, pp, rp, nullptr);
    prof_collect(h);
    h->profiling = false;
    if (rc != SHG_OK) return -1.0;
    auto it = h->ktimes.find(name);
    if (it == h->ktimes.end() || it->second.launches == 0) return -1.0;
    return it->second.ms / it->second.launches;
}

}  // namespace

extern "C" {

int shg_set_plan(shg_handle* h, int32_t n, int32_t precision, int32_t r, int32_t c) {
    if (!h) return SHG_EINVAL;
    if (!shg_fft_size_supported(n) || !fused_form(precision)) {
        set_error("shg_set_plan: n=%d precision=%d not a fused-pass plan", n, precision);
        return SHG_EINVAL;
    }
    if ((r && !rowpass_rows_supported(n, r)) || (c && !colpass_cols_supported(n, c))) {
        set_error("shg_set_plan: rows=%d cols=%d not supported at n=%d", r, c, n);
        return SHG_EINVAL;
    }
    const int a = (int)arith_of(precision);
    if (r) {
        h->rows_plan[{n, a}] = r;
        h->rows_plan_th.erase({n, a});  // a set plan holds for THERMAL launches too
    }
    if (c) h->cols_plan[{n, a}] = c;
    return SHG_OK;
}

int shg_get_plan(shg_handle* h, int32_t n, int32_t precision, int32_t* rows, int32_t* cols) {
    if (!h || !rows || !cols) return SHG_EINVAL;
    if (!shg_fft_size_supported(n) || !fused_form(precision)) {
        set_error("shg_get_plan: n=%d precision=%d not a fused-pass plan", n, precision);
        return SHG_EINVAL;
    }
    // built-in plans as for one n x n instance
    *rows = plan_rows(h, n, arith_of(precision), n);
    *cols = plan_cols(h, n, arith_of(precision), n);
    return SHG_OK;
}

int shg_autotune(shg_handle* h, int32_t n, int32_t precision, int32_t batch, int32_t nz) {
    if (!h) return SHG_EINVAL;
    if (!shg_fft_size_supported(n) || !fused_form(precision)) {
        set_error("shg_autotune: n=%d precision=%d not a fused-pass plan", n, precision);
        return SHG_EINVAL;
    }
    if (batch < 1 || batch > 1024 || nz < 2 || nz > 100000) {
        set_error("shg_autotune: batch=%d (1..1024) nz=%d (2..100000)", batch, nz);
        return SHG_EINVAL;
    }
    SHG_HIP(hipSetDevice(h->device));
    const Arith a = arith_of(precision);
    shg_params base;
    shg_params_default(&base);
    base.nx = base.ny = n;
    base.nz = nz;
    base.thermal = 0;
    base.precision = precision;
    base.lx_um = base.ly_um = 8.0 * base.waist_um;
    std::vector<shg_params> ps(batch, base);
    const bool prof = h->profiling;
    const std::pair<int, int> key{n, (int)a};
    const auto rows_had = h->rows_plan.find(key) != h->rows_plan.end() ? h->rows_plan[key] : 0;
    const auto cols_had = h->cols_plan.find(key) != h->cols_plan.end() ? h->cols_plan[key] : 0;
    const auto th_had = h->rows_plan_th.find(key) != h->rows_plan_th.end() ? h->rows_plan_th[key] : 0;
    auto restore = [&](int rc) {
        prof_reset(h);
        h->profiling = prof;
        if (rc != SHG_OK) {
            if (rows_had) h->rows_plan[key] = rows_had; else h->rows_plan.erase(key);
            if (cols_had) h->cols_plan[key] = cols_had; else h->cols_plan.erase(key);
            if (th_had) h->rows_plan_th[key] = th_had; else h->rows_plan_th.erase(key);
        }
        return rc;
    };
    // warm-up march (arena, twiddles, code objects)
    if (time_pass(h, ps, rowpass_name(n, false, a)) < 0) return restore(SHG_EDEVICE);
    double best = DBL_MAX;
    int best_rows = 0, best_cols = 0;
    for (int r : {1, 2, 4, 8, (int)SHG_PLAN_ROWS_WIDE}) {
        if (!rowpass_rows_supported(n, r)) continue;
        h->rows_plan[key] = r;
        const double t = time_pass(h, ps, rowpass_name(n, false, a));
        if (t > 0 && t < best) {
            best = t;
            best_rows = r;
        }
    }
    if (!best_rows) {
        set_error("shg_autotune: no row-pass plan ran at n=%d", n);
        return restore(SHG_EDEVICE);
    }
    h->rows_plan[key] = best_rows;
    // the THERMAL row pass: one thermo-optic pass with a 4-slice heat grid
    // and the direct heat solver (the heat solve itself is not timed)
    {
        std::vector<shg_params> pt(ps);
        for (auto& q : pt) {
            q.thermal = 1;
            q.nz_heat = std::min(4, nz);
            q.heat_solver = SHG_HEAT_DIRECT;
            q.max_outer = 1;
        }
        const char* name = rowpass_name(n, true, a);
        h->rows_plan_th.erase(key);
        if (time_pass(h, pt, name) < 0) return restore(SHG_EDEVICE);  // warm-up (heat tables)
        double bt = DBL_MAX;
        int br = 0;
        for (int r : {1, 2, 4, 8, (int)SHG_PLAN_ROWS_WIDE}) {
            if (!rowpass_rows_supported(n, r)) continue;
            h->rows_plan_th[key] = r;
            const double t = time_pass(h, pt, name);
            if (t > 0 && t < bt) {
                bt = t;
                br = r;
            }
        }
        if (!br) {
            set_error("shg_autotune: no THERMAL row-pass plan ran at n=%d", n);
            return restore(SHG_EDEVICE);
        }
        h->rows_plan_th[key] = br;
    }
    best = DBL_MAX;
    for (int c : {2, 4, 8, 16}) {
        if (!colpass_cols_supported(n, c)) continue;
        h->cols_plan[key] = c;
        const double t = time_pass(h, ps, colpass_name(n, a));
        if (t > 0 && t < best) {
            best = t;
            best_cols = c;
        }
    }
    if (!best_cols) {
        set_error("shg_autotune: no column-pass plan ran at n=%d", n);
        return restore(SHG_EDEVICE);
    }
    h->cols_plan[key] = best_cols;
    clear_error();
    return restore(SHG_OK);
}

int shg_set_graph_steps(shg_handle* h, int32_t steps) {
    if (!h) return SHG_EINVAL;
    if (steps < -1 || steps > 4096) {
        set_error("shg_set_graph_steps: steps=%d (-1..4096)", steps);
        return SHG_EINVAL;
    }
    h->graph_chunk = steps;
    return SHG_OK;
}

}  // extern "C"
