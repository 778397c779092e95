// fft_api.hip -- shg_fft2d: the library's 2D DFT on its own.  The fused
// passes transform rows inside the row pass and columns inside the column
// pass (the replacement for cufftExecC2C in the diffraction step,
// src/headers/Solver.h:159-183); this entry point runs the same transform
// code (fft_dist, the same twiddle tables, the same arithmetic types) over a
// caller's array, so the transform can be checked against golden DFT
// vectors directly rather than only through whole marches.
//
// Rows: one workgroup of N/8 threads per row.  Columns: four adjacent
// columns per workgroup, N/8 threads per column.  Unnormalised, like cuFFT:
// forward exp(-2 pi i jk/N), inverse exp(+2 pi i jk/N).
#include <hip/hip_runtime.h>

#include <vector>

#include "common.hpp"
#include "fft.hpp"
#include "runtime.hpp"

namespace shg {

template <int N, typename C, bool INV>
__global__ __launch_bounds__(N / 8) void k_fft_rows(C* __restrict__ d, const double2* __restrict__ tw) {
    constexpr int TPR = N / 8;
    __shared__ C lds[FftShape<N>::LDS];
    const int tp = threadIdx.x;
    const size_t base = (size_t)blockIdx.x * N + tp;
    C v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) v[q] = d[base + q * TPR];
    fft_dist<C, N, INV>(v, lds, tw, tp);
#pragma unroll
    for (int q = 0; q < 8; q++) d[base + q * TPR] = v[q];
}

constexpr int kFftCols = 4;

template <int N, typename C, bool INV>
__global__ __launch_bounds__(N / 8 * kFftCols) void k_fft_cols(C* __restrict__ d, int nx,
                                                               const double2* __restrict__ tw) {
    constexpr int TPC = N / 8;
    constexpr int LD = FftShape<N>::LDS;
    __shared__ C lds[kFftCols * LD];
    const int c = threadIdx.x, tp = threadIdx.y;
    const int x = blockIdx.x * kFftCols + c;
    C v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) v[q] = d[(size_t)(tp + q * TPC) * nx + x];
    fft_dist<C, N, INV>(v, lds + c * LD, tw, tp);
#pragma unroll
    for (int q = 0; q < 8; q++) d[(size_t)(tp + q * TPC) * nx + x] = v[q];
}

template <int N, typename C>
static hipError_t fft_rows_n(hipStream_t st, C* d, int rows, bool inv, const double2* tw) {
    if (inv)
        hipLaunchKernelGGL((k_fft_rows<N, C, true>), dim3(rows), dim3(N / 8), 0, st, d, tw);
    else
        hipLaunchKernelGGL((k_fft_rows<N, C, false>), dim3(rows), dim3(N / 8), 0, st, d, tw);
    return hipGetLastError();
}

template <int N, typename C>
static hipError_t fft_cols_n(hipStream_t st, C* d, int nx, bool inv, const double2* tw) {
    const dim3 grid(nx / kFftCols), block(kFftCols, N / 8);
    if (inv)
        hipLaunchKernelGGL((k_fft_cols<N, C, true>), grid, block, 0, st, d, nx, tw);
    else
        hipLaunchKernelGGL((k_fft_cols<N, C, false>), grid, block, 0, st, d, nx, tw);
    return hipGetLastError();
}

template <typename C>
static hipError_t fft_rows(hipStream_t st, C* d, int nx, int rows, bool inv, const double2* tw) {
    switch (nx) {
        case 64: return fft_rows_n<64, C>(st, d, rows, inv, tw);
        case 128: return fft_rows_n<128, C>(st, d, rows, inv, tw);
        case 192: return fft_rows_n<192, C>(st, d, rows, inv, tw);
        case 256: return fft_rows_n<256, C>(st, d, rows, inv, tw);
        case 384: return fft_rows_n<384, C>(st, d, rows, inv, tw);
        case 512: return fft_rows_n<512, C>(st, d, rows, inv, tw);
    }
    return hipErrorInvalidValue;
}

template <typename C>
static hipError_t fft_cols(hipStream_t st, C* d, int nx, int ny, bool inv, const double2* tw) {
    switch (ny) {
        case 64: return fft_cols_n<64, C>(st, d, nx, inv, tw);
        case 128: return fft_cols_n<128, C>(st, d, nx, inv, tw);
        case 192: return fft_cols_n<192, C>(st, d, nx, inv, tw);
        case 256: return fft_cols_n<256, C>(st, d, nx, inv, tw);
        case 384: return fft_cols_n<384, C>(st, d, nx, inv, tw);
        case 512: return fft_cols_n<512, C>(st, d, nx, inv, tw);
    }
    return hipErrorInvalidValue;
}

template <typename C>
static int fft2d(shg_handle* h, int nx, int ny, bool inv, const double* in, double* out) {
    using R = decltype(C{}.x);
    const size_t n = (size_t)nx * ny;
    std::vector<C> host(n);
    for (size_t i = 0; i < n; i++) host[i] = mk<C>((R)in[2 * i], (R)in[2 * i + 1]);
    const double2* twx = twiddles64(h, nx);
    const double2* twy = twiddles64(h, ny);
    if (!twx || !twy) {
        set_error("twiddle table allocation failed");
        return SHG_ENOMEM;
    }
    C* d = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&d), sizeof(C) * n) != hipSuccess) {
        (void)hipGetLastError();
        set_error("device allocation of %zu bytes failed", sizeof(C) * n);
        return SHG_ENOMEM;
    }
    struct Free {
        C* p;
        ~Free() { (void)hipFree(p); }
    } release{d};
    SHG_HIP(hipMemcpyAsync(d, host.data(), sizeof(C) * n, hipMemcpyHostToDevice, h->stream));
    SHG_HIP(fft_rows<C>(h->stream, d, nx, ny, inv, twx));
    SHG_HIP(fft_cols<C>(h->stream, d, nx, ny, inv, twy));
    SHG_HIP(hipMemcpyAsync(host.data(), d, sizeof(C) * n, hipMemcpyDeviceToHost, h->stream));
    SHG_HIP(hipStreamSynchronize(h->stream));
    for (size_t i = 0; i < n; i++) {
        out[2 * i] = host[i].x;
        out[2 * i + 1] = host[i].y;
    }
    return SHG_OK;
}

}  // namespace shg

using namespace shg;

extern "C" SHG_API int shg_fft2d(shg_handle* h, int32_t nx, int32_t ny, int32_t precision, int32_t inverse,
                                 const double* in, double* out) {
    if (!h || !in || !out) {
        set_error("shg_fft2d: null handle or buffer");
        return SHG_EINVAL;
    }
    if (!shg_fft_size_supported(nx) || !shg_fft_size_supported(ny)) {
        set_error("shg_fft2d: unsupported size %d x %d", (int)nx, (int)ny);
        return SHG_EINVAL;
    }
    SHG_HIP(hipSetDevice(h->device));
    switch (precision) {
        case SHG_PREC_FP64_PLAIN:
        case SHG_PREC_FP64_FUSED: return fft2d<double2>(h, nx, ny, inverse != 0, in, out);
        case SHG_PREC_FP32_MIXED:
        case SHG_PREC_FP32_STORAGE:
        case SHG_PREC_FP32_SPLIT: return fft2d<float2>(h, nx, ny, inverse != 0, in, out);
    }
    set_error("shg_fft2d: unknown precision %d", (int)precision);
    return SHG_EINVAL;
}
