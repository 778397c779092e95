// kernels_util.hip -- set-up and reduction kernels shared by both forms.
#include "kernels.hpp"

namespace shg {

// Gaussian pump focused at f (src/headers/Efields.h:84-119), evaluated in
// FP64 and stored in the field precision:
//   A = (Ap0/M) exp(-(rx^2 + ry^2) / (w0^2 M)),  M = 1 - i f/zR,
//   rx = (x - nx/2) dx,  ry = (y - ny/2) dy
template <typename C>
__global__ void k_pump_init(C* __restrict__ P, int nx, int ny, double dx, double dy, PumpCoef c) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= nx || y >= ny) return;
    const double rx = (x - 0.5 * nx) * dx, ry = (y - 0.5 * ny) * dy;
    // amp * exp(-rx^2 cx - ry^2 cy), cx = 1/(wx^2 Mx), cy = 1/(wy^2 My)
    const double er = -(rx * rx) * c.cx.x - (ry * ry) * c.cy.x;
    const double ei = -(rx * rx) * c.cx.y - (ry * ry) * c.cy.y;
    const double mag = exp(er);
    double sn, cs;
    sincos(ei, &sn, &cs);
    const double ar = mag * cs, ai = mag * sn;
    C v;
    v.x = (decltype(v.x))(c.amp.x * ar - c.amp.y * ai);
    v.y = (decltype(v.y))(c.amp.x * ai + c.amp.y * ar);
    P[(size_t)y * nx + x] = v;
}

template <typename C>
hipError_t launch_pump_init(hipStream_t st, C* P, int nx, int ny, double dx, double dy, const PumpCoef& c) {
    dim3 block(64, 4);
    dim3 grid((nx + 63) / 64, (ny + 3) / 4);
    hipLaunchKernelGGL((k_pump_init<C>), grid, block, 0, st, P, nx, ny, dx, dy, c);
    return hipGetLastError();
}
template hipError_t launch_pump_init<float2>(hipStream_t, float2*, int, int, double, double, const PumpCoef&);
template hipError_t launch_pump_init<double2>(hipStream_t, double2*, int, int, double, double, const PumpCoef&);

// FP64 partial sums of |A|^2 (block-level tree reduction, deterministic)
template <typename C>
__global__ __launch_bounds__(256) void k_sumsq(const C* __restrict__ A, long n, double* __restrict__ partials) {
    __shared__ double red[256];
    double acc = 0.0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += 256L * gridDim.x) {
        const C v = A[i];
        acc += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

hipError_t launch_sumsq_f32(hipStream_t st, const float2* A, long n, double* partials, int nblocks) {
    hipLaunchKernelGGL((k_sumsq<float2>), dim3(nblocks), dim3(256), 0, st, A, n, partials);
    return hipGetLastError();
}
hipError_t launch_sumsq_f64(hipStream_t st, const double2* A, long n, double* partials, int nblocks) {
    hipLaunchKernelGGL((k_sumsq<double2>), dim3(nblocks), dim3(256), 0, st, A, n, partials);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ partials, int n,
                                                      double* __restrict__ out) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc += partials[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

hipError_t launch_sum_partials(hipStream_t st, const double* partials, int n, double* out) {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, partials, n, out);
    return hipGetLastError();
}

__global__ void k_f64_to_f32(const double* __restrict__ in, float* __restrict__ out, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)blockDim.x * gridDim.x)
        out[i] = (float)in[i];
}

hipError_t launch_f64_to_f32(hipStream_t st, const double* in, float* out, long n) {
    long nb = (n + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(k_f64_to_f32, dim3((unsigned)nb), dim3(256), 0, st, in, out, n);
    return hipGetLastError();
}

// step counter of graph-replayed marches: set, and advance by a chunk
// (one thread; a fill kernel rather than a memset so that it may be captured)
__global__ void k_step_set(int* c, int v) {
    if (threadIdx.x == 0) *c = v;
}
__global__ void k_step_advance(int* c, int by) {
    if (threadIdx.x == 0) *c += by;
}

hipError_t launch_step_set(hipStream_t st, int* counter, int value) {
    hipLaunchKernelGGL(k_step_set, dim3(1), dim3(64), 0, st, counter, value);
    return hipGetLastError();
}

hipError_t launch_step_advance(hipStream_t st, int* counter, int by) {
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, st, counter, by);
    return hipGetLastError();
}

}  // namespace shg
