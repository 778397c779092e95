// stubs.cpp -- launch plans (shg_set_plan) and the autotuner entry point.
#include "heat.hpp"
#include "kernels.hpp"
#include "march_mixed.hpp"
#include "runtime.hpp"
using namespace shg;
extern "C" {
int shg_set_plan(shg_handle* h, int32_t n, int32_t precision, int32_t r, int32_t c) {
    if (!h) return SHG_EINVAL;
    if (!shg_fft_size_supported(n) || (precision != SHG_PREC_FP32_MIXED && precision != SHG_PREC_FP32_STORAGE &&
                                       precision != SHG_PREC_FP64_FUSED)) {
        set_error("shg_set_plan: n=%d precision=%d not a fused-pass plan", n, precision);
        return SHG_EINVAL;
    }
    if ((r && !rowpass_rows_supported(n, r)) || (c && !colpass_cols_supported(n, c))) {
        set_error("shg_set_plan: rows=%d cols=%d not supported at n=%d", r, c, n);
        return SHG_EINVAL;
    }
    const int a = (int)arith_of(precision);
    if (r) h->rows_plan[{n, a}] = r;
    if (c) h->cols_plan[{n, a}] = c;
    return SHG_OK;
}
int shg_autotune(shg_handle*, int32_t, int32_t, int32_t, int32_t) { return SHG_EUNSUPPORTED; }
}
