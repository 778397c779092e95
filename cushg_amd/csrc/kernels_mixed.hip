// kernels_mixed.hip -- launch front of the fused SSFM passes: the form
// switch (storage and arithmetic types; the passes themselves are in
// mixed_passes.hpp, instantiated per form in kernels_mixed_{f32,s32,f64}.hip)
// and the built-in tile plans.
#include "kernels.hpp"

namespace shg {

hipError_t rowpass_f32(int, bool, int, dim3, hipStream_t, const MixedInst*, int, const double2*);
hipError_t rowpass_s32(int, bool, int, dim3, hipStream_t, const MixedInst*, int, const double2*);
hipError_t rowpass_f64(int, bool, int, dim3, hipStream_t, const MixedInst*, int, const double2*);
hipError_t colpass_f32(int, int, dim3, hipStream_t, const MixedInst*, int, const double2*);
hipError_t colpass_s32(int, int, dim3, hipStream_t, const MixedInst*, int, const double2*);
hipError_t colpass_f64(int, int, dim3, hipStream_t, const MixedInst*, int, const double2*);

// Built-in tile plans (measured on MI355X: rows per workgroup of the row
// pass, columns per workgroup of the column pass); shg_autotune refines them.
int rowpass_rows_default(int nx, Arith a) {
    const bool fp64 = a != Arith::F32;
    switch (nx) {
        case 64: return fp64 ? 4 : 8;
        case 128: return fp64 ? 2 : 4;
        case 192: return 2;
        case 256: return fp64 ? 1 : 2;
        case 384: return 1;
        case 512: return 1;
    }
    return 0;
}
bool rowpass_rows_supported(int nx, int rows) {
    return (rows == 1 || rows == 2 || rows == 4 || rows == 8) && nx / 8 * rows <= 1024;
}
int colpass_cols_default(int ny, Arith a) {
    (void)ny;
    (void)a;
    return 8;
}
bool colpass_cols_supported(int ny, int cols) {
    return (cols == 4 || cols == 8 || cols == 16) && ny / 8 * cols <= 1024;
}

hipError_t launch_rowpass(int nx, bool thermal, Arith a, int rows, dim3 grid, hipStream_t st,
                          const MixedInst* insts, int g, const double2* tw) {
    switch (a) {
        case Arith::F32: return rowpass_f32(nx, thermal, rows, grid, st, insts, g, tw);
        case Arith::S32: return rowpass_s32(nx, thermal, rows, grid, st, insts, g, tw);
        case Arith::F64: return rowpass_f64(nx, thermal, rows, grid, st, insts, g, tw);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_colpass(int ny, Arith a, int cols, dim3 grid, hipStream_t st, const MixedInst* insts, int g,
                          const double2* tw) {
    switch (a) {
        case Arith::F32: return colpass_f32(ny, cols, grid, st, insts, g, tw);
        case Arith::S32: return colpass_s32(ny, cols, grid, st, insts, g, tw);
        case Arith::F64: return colpass_f64(ny, cols, grid, st, insts, g, tw);
    }
    return hipErrorInvalidValue;
}

}  // namespace shg
