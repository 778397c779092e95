// temporary: THERMAL, FP64-plain and adjoint paths not built yet
#include "heat.hpp"
#include "runtime.hpp"
namespace shg {
int run_adjoint(shg_handle*, const shg_params*, int, double*, shg_result*) { set_error("adjoint not built"); return SHG_EUNSUPPORTED; }
}
extern "C" {
int shg_set_plan(shg_handle* h, int32_t n, int32_t r, int32_t c) {
    if (!h) return SHG_EINVAL;
    if (r) h->rows_plan[n] = r;
    if (c) h->cols_plan[n] = c;
    return SHG_OK;
}
int shg_autotune(shg_handle*, int32_t, int32_t, int32_t) { return SHG_EUNSUPPORTED; }
}
