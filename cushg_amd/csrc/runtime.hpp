// runtime.hpp -- handle, device arena, error and timing helpers (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/shg.h"

namespace shg {

void set_error(const char* fmt, ...);
void clear_error();

#define SHG_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            ::shg::set_error("%s:%d: %s -> %s", __FILE__, __LINE__, #call,             \
                             hipGetErrorString(e_));                                   \
            return e_ == hipErrorOutOfMemory ? SHG_ENOMEM : SHG_EDEVICE;               \
        }                                                                              \
    } while (0)

// Bump allocator over one device allocation, grown on demand and reused
// across calls on the same handle (no per-call hipMalloc once warm).
struct Arena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
    int reserve(size_t bytes);  // ensure capacity; invalidates earlier carves
    void reset() { used = 0; }
    template <typename T> T* carve(size_t count) {
        size_t b = (count * sizeof(T) + 255) & ~size_t(255);
        if (used + b > cap) return nullptr;
        T* r = reinterpret_cast<T*>(base + used);
        used += b;
        return r;
    }
};

// Per-slice trace of a single run (shg_run_traced): the callback, the
// slice stride and host staging for the two fields.
struct SliceTrace {
    int every = 1;
    shg_slice_fn fn = nullptr;
    void* user = nullptr;
    std::vector<double> pump, signal;
    // hands slice `slice` of pass `pass` (already in pump/signal) to fn
    int deliver(int slice, int pass) {
        if (fn(slice, pass, pump.data(), signal.data(), user) != 0) {
            set_error("run cancelled by the slice callback at slice %d of pass %d", slice, pass);
            return SHG_EINVAL;
        }
        return SHG_OK;
    }
    bool wants(int slice, int nz) const { return slice <= nz - 1 && slice % every == 0; }
};

struct KernelTime {
    double ms = 0;
    int launches = 0;
};

}  // namespace shg

namespace shg {
struct HeatGeom;
}

struct shg_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream (default priority) for work that overlaps the main
    // stream's (the adjoint's segment recomputation), and ordering events
    hipStream_t stream2 = nullptr;
    hipEvent_t ev[6] = {};
    // two more streams (created on first use) so that independent launch
    // groups of a batched march run concurrently on up to four streams,
    // with a fork event and per-stream join events
    hipStream_t lanes[2] = {};
    hipEvent_t fork_ev = nullptr, join_ev[3] = {};
    shg::Arena arena;
    std::map<int, double2*> tw64;
    std::map<std::string, shg::HeatGeom*> heat_cache;  // direct-solver tables per geometry
    double* d_scalars = nullptr;  // reduction slots
    double* d_partials = nullptr;
    int n_partials = 0;
    // tuned launch plans: rows per workgroup (row pass, by (nx, form)),
    // columns per workgroup (column pass, by (ny, form)); absent = built-in
    // default.  rows_plan_th: the autotuner's row plan for THERMAL row
    // passes, ahead of rows_plan for those launches
    std::map<std::pair<int, int>, int> rows_plan, cols_plan, rows_plan_th;
    // graph replay of the fused march: steps per captured graph (0: direct
    // launches; negative: automatic -- graphs of 32 steps for marches with
    // more than one independent launch lane), the device step counter and
    // the cached executable graphs keyed by the launch structure they were
    // captured from
    int graph_chunk = -1;
    int* step_counter = nullptr;
    std::map<std::vector<uint64_t>, hipGraphExec_t> graphs;
    // profiling
    bool profiling = false;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, shg::KernelTime> ktimes;
    std::vector<std::string> knames;  // storage for names handed out
};

namespace shg {

// Profiling wrapper: brackets a launch with events on the handle's stream.
// Times what runs on stream s while it lives, under `name` (when the handle
// profiles).  With one_launch, the scope wraps exactly one fused-pass launch
// and its events are recorded by that launch's own dispatch packet
// (pass_launch), not as separate packets around it.
struct ProfScope {
    shg_handle* h;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t st;
    bool one = false;
    ProfScope(shg_handle* h_, const char* n, hipStream_t s = nullptr, bool one_launch = false);
    ~ProfScope();
};
void prof_collect(shg_handle* h);
void prof_reset(shg_handle* h);

// the zero-allocation run's device scratch: looks up its address on
// `device` (the current device) once; called by shg_handle_create
int noalloc_attach(int device);

const double2* twiddles64(shg_handle* h, int n);
// the values of that table, tw[m] = exp(-2 pi i m / n), into host memory
void fill_twiddles64(double2* t, int n);

int sum_sq(shg_handle* h, const float2* A, long n, double* slot);
int sum_sq64(shg_handle* h, const double2* A, long n, double* slot);

}  // namespace shg
