// noalloc.hip -- shg_run_noalloc: a single non-THERMAL run of a fused form
// that allocates no device memory.  Every device buffer of the run (fields,
// saved output, propagator and phase tables, FFT twiddles, the instance
// descriptor, reduction partials and slots, output staging) is carved from
// one file-scope device array, `scratch`, of kNoallocScratchBytes.  Host
// tables live in per-thread static arrays.  The kernels are the batched
// march's own (k_rowpass / k_colpass, launched directly with the same plans
// as a one-instance shg_run), so the results are those of shg_run bit for
// bit.
//
// Each GPU holds its own copy of `scratch` until the process ends.  Its
// device address is looked up once per device with hipGetSymbolAddress
// (noalloc_attach, called by shg_handle_create with the device current);
// host code uses that address, never the array's name.  Calls on one GPU
// share the array and each leaves its contents behind: two calls that
// overlap on the same GPU race.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <mutex>

#include "common.hpp"
#include "kernels.hpp"
#include "march_mixed.hpp"
#include "runtime.hpp"

namespace shg {

static constexpr size_t kNoallocScratchBytes = size_t(SHG_NOALLOC_SCRATCH_MIB) << 20;
static constexpr int kNoallocMaxNz = SHG_NOALLOC_MAX_NZ;
static constexpr int kMaxDevices = 64;

static __device__ __align__(256) char scratch[kNoallocScratchBytes];

static std::mutex g_scratch_mu;
static char* g_scratch_ptr[kMaxDevices] = {};

// Looks up this device's address of `scratch` (the current device must be
// `device`); idempotent per device.
int noalloc_attach(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("device %d beyond the %d devices the scratch table holds", device, kMaxDevices);
        return SHG_EINVAL;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return SHG_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        set_error("hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        return SHG_EDEVICE;
    }
    g_scratch_ptr[device] = static_cast<char*>(ptr);
    return SHG_OK;
}

__global__ void k_zero_f32x2(float2* a, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        a[i] = make_float2(0.f, 0.f);
}
__global__ void k_zero_f64x2(double2* a, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        a[i] = make_double2(0.0, 0.0);
}
__global__ void k_widen_f32x2(const float2* __restrict__ in, double2* __restrict__ out, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = make_double2(in[i].x, in[i].y);
}

// bump allocation over the scratch array (256-byte aligned pieces)
struct ScratchCarve {
    char* base;
    size_t used = 0;
    bool over = false;
    template <typename T> T* take(size_t count) {
        const size_t b = (count * sizeof(T) + 255) & ~size_t(255);
        if (used + b > kNoallocScratchBytes) {
            over = true;
            return nullptr;
        }
        T* r = reinterpret_cast<T*>(base + used);
        used += b;
        return r;
    }
};

// per-thread host staging of the FP64 tables (no heap)
struct NoallocHost {
    double2 exp_[512], exs[512], eyp[512], eys[512];
    double2 twx[512], twy[512];
    double2 phase[(size_t)kNoallocMaxNz * 6];
    double slots[4];
};
static thread_local NoallocHost t_host;

static int sum_sq_scratch(hipStream_t st, Arith a, const void* A, long n, double* partials, double* slot) {
    const int nb = 256;
    if (a == Arith::F64)
        SHG_HIP(launch_sumsq_f64(st, static_cast<const double2*>(A), n, partials, nb));
    else
        SHG_HIP(launch_sumsq_f32(st, static_cast<const float2*>(A), n, partials, nb));
    SHG_HIP(launch_sum_partials(st, partials, nb, slot));
    return SHG_OK;
}

// field (FP32 or FP64 complex) -> caller's interleaved FP64 host array;
// FP32 fields are widened on the device into `stage` first
static int field_out(hipStream_t st, Arith a, const void* A, double2* stage, size_t n, double* out) {
    const void* src = A;
    if (a != Arith::F64) {
        k_widen_f32x2<<<256, 256, 0, st>>>(static_cast<const float2*>(A), stage, (long)n);
        SHG_HIP(hipGetLastError());
        src = stage;
    }
    SHG_HIP(hipMemcpyAsync(out, src, sizeof(double2) * n, hipMemcpyDeviceToHost, st));
    return SHG_OK;
}

static int run_noalloc(shg_handle* h, const shg_params* p, shg_result* r) {
    const auto t0 = std::chrono::steady_clock::now();
    char* base = g_scratch_ptr[h->device];
    if (!base) {
        set_error("device scratch of device %d not attached", h->device);
        return SHG_EDEVICE;
    }
    hipStream_t st = h->stream;
    MixedHost m;
    m.p = p;
    m.r = r;
    m.nx = p->nx;
    m.ny = p->ny;
    m.nz = p->nz;
    m.nzh = p->nz;
    m.N2 = (size_t)m.nx * m.ny;
    m.dx = p->lx_um / (m.nx - 1);
    m.dy = p->ly_um / (m.ny - 1);
    m.dz = p->lz_um / (m.nz - 1);
    m.thermal = false;
    m.arith = arith_of(p->precision);
    m.Tc = p->T_C;
    const size_t fb = arith_field_bytes(m.arith);
    ScratchCarve sc{base};
    // P and S adjacent: at the end they are free and hold the FP64 output staging
    char* ps = sc.take<char>(2 * fb * m.N2);
    m.P = ps;
    m.S = ps + fb * m.N2;
    double2* stage = reinterpret_cast<double2*>(ps);
    m.outP = sc.take<char>(fb * m.N2);
    m.outS = sc.take<char>(fb * m.N2);
    m.exP = sc.take<double2>(m.nx);
    m.exS = sc.take<double2>(m.nx);
    m.eyP = sc.take<double2>(m.ny);
    m.eyS = sc.take<double2>(m.ny);
    m.phase = sc.take<double2>((size_t)m.nz * 6);
    double2* twx = sc.take<double2>(m.nx);
    double2* twy = sc.take<double2>(m.ny);
    MixedInst* d = sc.take<MixedInst>(1);
    double* partials = sc.take<double>(256);
    m.slots = sc.take<double>(4);
    if (sc.over || 2 * fb * m.N2 < sizeof(double2) * m.N2) {
        set_error("shg_run_noalloc: %dx%d x %d slices in this form needs more than the %zu-byte scratch", m.nx, m.ny,
                  m.nz, kNoallocScratchBytes);
        return SHG_EINVAL;
    }
    NoallocHost& hb = t_host;
    host_tables_into(m, hb.exp_, hb.exs, hb.eyp, hb.eys, hb.phase);
    fill_twiddles64(hb.twx, m.nx);
    fill_twiddles64(hb.twy, m.ny);
    MixedInst& dd = m.desc;
    dd.P = m.P;
    dd.S = m.S;
    dd.outP = m.outP;
    dd.outS = m.outS;
    dd.trace_every = 1;
    dd.exP = m.exP;
    dd.exS = m.exS;
    dd.eyP = m.eyP;
    dd.eyS = m.eyS;
    dd.phase = m.phase;
    SHG_HIP(hipMemcpyAsync(m.exP, hb.exp_, sizeof(double2) * m.nx, hipMemcpyHostToDevice, st));
    SHG_HIP(hipMemcpyAsync(m.exS, hb.exs, sizeof(double2) * m.nx, hipMemcpyHostToDevice, st));
    SHG_HIP(hipMemcpyAsync(m.eyP, hb.eyp, sizeof(double2) * m.ny, hipMemcpyHostToDevice, st));
    SHG_HIP(hipMemcpyAsync(m.eyS, hb.eys, sizeof(double2) * m.ny, hipMemcpyHostToDevice, st));
    SHG_HIP(hipMemcpyAsync(m.phase, hb.phase, sizeof(double2) * m.nz * 6, hipMemcpyHostToDevice, st));
    SHG_HIP(hipMemcpyAsync(twx, hb.twx, sizeof(double2) * m.nx, hipMemcpyHostToDevice, st));
    SHG_HIP(hipMemcpyAsync(twy, hb.twy, sizeof(double2) * m.ny, hipMemcpyHostToDevice, st));
    SHG_HIP(hipMemcpyAsync(d, &dd, sizeof(MixedInst), hipMemcpyHostToDevice, st));
    // the host sources are reused by the next call on this thread and dd
    // lives on this stack frame: the copies complete before anything else
    SHG_HIP(hipStreamSynchronize(st));

    // pump (src/headers/Efields.h:84-119), a zero SH field, P_in
    const PumpCoef pc = pump_coef(p, refr_index<double>(p->crystal, p->lambda_p_um, m.Tc));
    if (m.arith == Arith::F64) {
        SHG_HIP(launch_pump_init<double2>(st, static_cast<double2*>(m.P), m.nx, m.ny, m.dx, m.dy, pc));
        k_zero_f64x2<<<256, 256, 0, st>>>(static_cast<double2*>(m.S), (long)m.N2);
    } else {
        SHG_HIP(launch_pump_init<float2>(st, static_cast<float2*>(m.P), m.nx, m.ny, m.dx, m.dy, pc));
        k_zero_f32x2<<<256, 256, 0, st>>>(static_cast<float2*>(m.S), (long)m.N2);
    }
    SHG_HIP(hipGetLastError());
    if (int rc = sum_sq_scratch(st, m.arith, m.P, (long)m.N2, partials, m.slots + 0)) return rc;

    // the z-march (Solver::CWES, src/headers/Solver.h:276-308): row pass g,
    // column pass g, ..., row pass nz -- the launches a one-instance shg_run
    // makes, with the same plans
    const int rows = plan_rows(h, m.nx, m.arith, m.ny, false);
    const int per = rowpass_rows_per_wg(m.nx, rows);
    const dim3 rgrid((m.ny + per - 1) / per, 1, 1);
    const int cols = plan_cols(h, m.ny, m.arith, m.nx);
    const dim3 cgrid((m.nx + cols - 1) / cols, 2, 1);
    for (int g = 0; g <= m.nz; g++) {
        SHG_HIP(launch_rowpass(m.nx, false, m.arith, rows, rgrid, st, d, g, twx));
        if (g < m.nz) SHG_HIP(launch_colpass(m.ny, m.arith, cols, cgrid, st, d, g, twy));
    }
    if (int rc = sum_sq_scratch(st, m.arith, m.outP, (long)m.N2, partials, m.slots + 1)) return rc;
    if (int rc = sum_sq_scratch(st, m.arith, m.outS, (long)m.N2, partials, m.slots + 2)) return rc;
    SHG_HIP(hipMemcpyAsync(hb.slots, m.slots, sizeof(double) * 3, hipMemcpyDeviceToHost, st));
    SHG_HIP(hipStreamSynchronize(st));
    const double* sl = hb.slots;
    if (!std::isfinite(sl[0]) || !std::isfinite(sl[1]) || !std::isfinite(sl[2])) {
        set_error("non-finite field power");
        r->status = SHG_ENAN;
        return SHG_ENAN;
    }
    if (r->pump_out) {
        if (int rc = field_out(st, m.arith, m.outP, stage, m.N2, r->pump_out)) return rc;
        SHG_HIP(hipStreamSynchronize(st));
    }
    if (r->signal_out) {
        if (int rc = field_out(st, m.arith, m.outS, stage, m.N2, r->signal_out)) return rc;
        SHG_HIP(hipStreamSynchronize(st));
    }
    const double np = refr_index<double>(p->crystal, p->lambda_p_um, m.Tc);
    const double ns = refr_index<double>(p->crystal, 0.5 * p->lambda_p_um, m.Tc);
    const double c = 0.5 * kEps0 * kC * m.dx * m.dy;
    r->P_pump_in_W = c * np * sl[0];
    r->P_pump_out_W = c * np * sl[1];
    r->P_sh_W = c * ns * sl[2];
    r->eta = p->power_W > 0 ? r->P_sh_W / p->power_W : 0.0;
    r->steps_done = m.nz;
    r->jacobi_sweeps = 0;
    r->outer_iters = 1;
    r->heat_residual_K = 0;
    r->peak_dT_K = 0;
    r->T_max_C = m.Tc;
    r->status = SHG_OK;
    r->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return SHG_OK;
}

}  // namespace shg

extern "C" int shg_run_noalloc(shg_handle* h, const shg_params* p, shg_result* r) {
    using namespace shg;
    if (!h || !p || !r) {
        set_error("shg_run_noalloc: null argument");
        return SHG_EINVAL;
    }
    clear_error();
    r->status = shg_validate(p);
    if (r->status != SHG_OK) return r->status;
    int rc = SHG_OK;
    if (p->thermal) {
        set_error("shg_run_noalloc: THERMAL runs need the heat volumes of shg_run");
        rc = SHG_EUNSUPPORTED;
    } else if (p->precision == SHG_PREC_FP64_PLAIN) {
        set_error("shg_run_noalloc: the FP64 plain form is not served (fused forms only)");
        rc = SHG_EUNSUPPORTED;
    } else if (p->nz > kNoallocMaxNz) {
        set_error("shg_run_noalloc: nz=%d exceeds the maximum %d", p->nz, kNoallocMaxNz);
        rc = SHG_EINVAL;
    }
    if (rc == SHG_OK) {
        const hipError_t e = hipSetDevice(h->device);
        if (e != hipSuccess) {
            set_error("hipSetDevice(%d): %s", h->device, hipGetErrorString(e));
            rc = SHG_EDEVICE;
        } else {
            rc = run_noalloc(h, p, r);
        }
    }
    r->status = rc;
    return rc;
}
