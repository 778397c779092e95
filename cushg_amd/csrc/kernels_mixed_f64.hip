// kernels_mixed_f64.hip -- fused passes with FP64 fields, FP64 arithmetic.
#include "mixed_passes.hpp"

namespace shg {

hipError_t rowpass_f64(int nx, bool thermal, int rows, dim3 grid, hipStream_t st, const MixedInst* insts, Step g,
                        const double2* tw) {
    return rowpass_dispatch<double2, double2>(nx, thermal, rows, grid, st, insts, g, tw);
}

hipError_t colpass_f64(int ny, int cols, dim3 grid, hipStream_t st, const MixedInst* insts, Step g,
                        const double2* tw) {
    return colpass_dispatch<double2, double2>(ny, cols, grid, st, insts, g, tw);
}

}  // namespace shg
