// runtime.cpp -- error state, arena, profiling, twiddle tables, reductions.
#include "runtime.hpp"

#include <cmath>
#include <cstring>

#include "kernels.hpp"

namespace shg {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}
void clear_error() { g_err.clear(); }
const char* last_error() { return g_err.c_str(); }

int Arena::reserve(size_t bytes) {
    if (bytes <= cap) {
        used = 0;
        return SHG_OK;
    }
    if (base) {
        (void)hipDeviceSynchronize();
        (void)hipFree(base);
        base = nullptr;
        cap = 0;
    }
    size_t want = bytes + bytes / 8 + (1 << 20);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&base), want);
    if (e != hipSuccess) {
        base = nullptr;
        // the failed allocation is also the thread's last HIP error: clear it,
        // or the next launch check of a later call would report it again
        (void)hipGetLastError();
        set_error("arena: hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
        return SHG_ENOMEM;
    }
    cap = want;
    used = 0;
    return SHG_OK;
}

LaunchEvents& launch_events() {
    thread_local LaunchEvents e;
    return e;
}

ProfScope::ProfScope(shg_handle* h_, const char* n, hipStream_t s, bool one_launch)
    : h(h_), name(n), st(s ? s : h_->stream), one(one_launch) {
    if (!h->profiling) return;
    if (h->event_pool.size() < 2) {
        for (int i = 0; i < 64; i++) {
            hipEvent_t e;
            if (hipEventCreate(&e) == hipSuccess) h->event_pool.push_back(e);
        }
    }
    if (h->event_pool.size() < 2) return;
    a = h->event_pool.back();
    h->event_pool.pop_back();
    b = h->event_pool.back();
    h->event_pool.pop_back();
    if (one)
        launch_events() = LaunchEvents{a, b, false};
    else
        (void)hipEventRecord(a, st);
}

ProfScope::~ProfScope() {
    if (!a) return;
    if (one) {
        LaunchEvents& le = launch_events();
        const bool used = le.used;
        le = LaunchEvents{};
        if (!used) {  // no launch took the events: an empty interval
            (void)hipEventRecord(a, st);
            (void)hipEventRecord(b, st);
        }
    } else {
        (void)hipEventRecord(b, st);
    }
    h->pending.push_back({name, {a, b}});
    if (h->pending.size() > 4096) prof_collect(h);
}

void prof_collect(shg_handle* h) {
    if (h->pending.empty()) return;
    (void)hipStreamSynchronize(h->stream);
    if (h->stream2) (void)hipStreamSynchronize(h->stream2);
    for (auto s : h->lanes)
        if (s) (void)hipStreamSynchronize(s);
    for (auto& pe : h->pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
            auto& kt = h->ktimes[pe.first];
            kt.ms += ms;
            kt.launches += 1;
        }
        h->event_pool.push_back(pe.second.first);
        h->event_pool.push_back(pe.second.second);
    }
    h->pending.clear();
}

void prof_reset(shg_handle* h) {
    prof_collect(h);
    h->ktimes.clear();
}

static void tw_value(int n, int k, double& c, double& s) {
    const int kk = k <= n / 2 ? k : k - n;
    const double a = -2.0 * kPi * (double)kk / (double)n;
    c = std::cos(a);
    s = std::sin(a);
}
static void tw_set(double2& t, double c, double s) {
    t.x = c;
    t.y = s;
}

void fill_twiddles64(double2* t, int n) {
    for (int k = 0; k < n; k++) {
        double c, s;
        tw_value(n, k, c, s);
        tw_set(t[k], c, s);
    }
}

// tw[m] = exp(-2 pi i m / n) in FP64
template <typename C>
static C* make_twiddles(int n) {
    std::vector<C> t(n);
    fill_twiddles64(t.data(), n);
    C* d = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&d), sizeof(C) * n) != hipSuccess) return nullptr;
    if (hipMemcpy(d, t.data(), sizeof(C) * n, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    return d;
}

const double2* twiddles64(shg_handle* h, int n) {
    auto it = h->tw64.find(n);
    if (it != h->tw64.end()) return it->second;
    double2* d = make_twiddles<double2>(n);
    if (d) h->tw64[n] = d;
    return d;
}

static int ensure_partials(shg_handle* h, int n) {
    if (h->n_partials >= n) return SHG_OK;
    if (h->d_partials) (void)hipFree(h->d_partials);
    h->d_partials = nullptr;
    SHG_HIP(hipMalloc(reinterpret_cast<void**>(&h->d_partials), sizeof(double) * n));
    h->n_partials = n;
    return SHG_OK;
}

int sum_sq(shg_handle* h, const float2* A, long n, double* slot) {
    const int nb = 256;
    int rc = ensure_partials(h, nb);
    if (rc) return rc;
    ProfScope ps(h, "reduce_sumsq");
    SHG_HIP(launch_sumsq_f32(h->stream, A, n, h->d_partials, nb));
    SHG_HIP(launch_sum_partials(h->stream, h->d_partials, nb, slot));
    return SHG_OK;
}

int sum_sq64(shg_handle* h, const double2* A, long n, double* slot) {
    const int nb = 256;
    int rc = ensure_partials(h, nb);
    if (rc) return rc;
    ProfScope ps(h, "reduce_sumsq");
    SHG_HIP(launch_sumsq_f64(h->stream, A, n, h->d_partials, nb));
    SHG_HIP(launch_sum_partials(h->stream, h->d_partials, nb, slot));
    return SHG_OK;
}

}  // namespace shg
