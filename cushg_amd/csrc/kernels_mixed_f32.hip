// kernels_mixed_f32.hip -- fused passes with FP32 fields, FP32 arithmetic.
#include "mixed_passes.hpp"

namespace shg {

hipError_t rowpass_f32(int nx, bool thermal, int rows, dim3 grid, hipStream_t st, const MixedInst* insts, Step g,
                        const double2* tw) {
    return rowpass_dispatch<float2, float2>(nx, thermal, rows, grid, st, insts, g, tw);
}

hipError_t colpass_f32(int ny, int cols, dim3 grid, hipStream_t st, const MixedInst* insts, Step g,
                        const double2* tw) {
    return colpass_dispatch<float2, float2>(ny, cols, grid, st, insts, g, tw);
}

}  // namespace shg
