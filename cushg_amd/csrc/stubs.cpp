// temporary: THERMAL, FP64-plain and adjoint paths not built yet
#include "heat.hpp"
#include "kernels.hpp"
#include "runtime.hpp"
namespace shg {
int run_adjoint(shg_handle*, const shg_params*, int, double*, shg_result*) { set_error("adjoint not built"); return SHG_EUNSUPPORTED; }
}
using namespace shg;
extern "C" {
int shg_set_plan(shg_handle* h, int32_t n, int32_t precision, int32_t r, int32_t c) {
    if (!h) return SHG_EINVAL;
    if (!shg_fft_size_supported(n) || (precision != SHG_PREC_FP32_MIXED && precision != SHG_PREC_FP32_STORAGE)) {
        set_error("shg_set_plan: n=%d precision=%d not a fused-pass plan", n, precision);
        return SHG_EINVAL;
    }
    if ((r && !rowpass_rows_supported(n, r)) || (c && !colpass_cols_supported(n, c))) {
        set_error("shg_set_plan: rows=%d cols=%d not supported at n=%d", r, c, n);
        return SHG_EINVAL;
    }
    const bool f64 = precision == SHG_PREC_FP32_STORAGE;
    if (r) h->rows_plan[{n, f64}] = r;
    if (c) h->cols_plan[{n, f64}] = c;
    return SHG_OK;
}
int shg_autotune(shg_handle*, int32_t, int32_t, int32_t, int32_t) { return SHG_EUNSUPPORTED; }
}
