// kernels_mixed.hip -- launch front of the fused SSFM passes: the form
// switch (storage and arithmetic types; the passes themselves are in
// mixed_passes.hpp, instantiated per form in kernels_mixed_{f32,s32,f64}.hip)
// and the built-in tile plans.
#include "kernels.hpp"

namespace shg {

hipError_t rowpass_f32(int, bool, int, dim3, hipStream_t, const MixedInst*, Step, const double2*);
hipError_t rowpass_s32(int, bool, int, dim3, hipStream_t, const MixedInst*, Step, const double2*);
hipError_t rowpass_f64(int, bool, int, dim3, hipStream_t, const MixedInst*, Step, const double2*);
hipError_t colpass_f32(int, int, dim3, hipStream_t, const MixedInst*, Step, const double2*);
hipError_t colpass_s32(int, int, dim3, hipStream_t, const MixedInst*, Step, const double2*);
hipError_t colpass_f64(int, int, dim3, hipStream_t, const MixedInst*, Step, const double2*);

// Built-in tile plans (measured on MI355X: rows per workgroup of the row
// pass, columns per workgroup of the column pass); shg_autotune refines them.
// Rows up to 256 points take the wide form: measured faster for one
// instance and for batches of 16 (256 x 256 THERMAL FP64 arithmetic: 40 us
// against 58 us per launch); 384 and 512 keep NX/8 threads per row, which
// is faster once a batch fills the chip (512 x 512 x 16: 108 us against
// 160 us).  The row plan depends on the row length only, never on the
// batch, so an instance gives the same bits alone and in any batch.
int rowpass_rows_default(int nx, Arith a) {
    (void)a;
    switch (nx) {
        case 64:
        case 128:
        case 192:
        case 256: return kRowsWide;
        case 384: return 1;
        case 512: return 1;
    }
    return 0;
}
bool rowpass_rows_supported(int nx, int rows) {
    if (rows == kRowsWide) return nx >= 64 && nx <= 512;
    return (rows == 1 || rows == 2 || rows == 4 || rows == 8) && nx / 8 * rows <= 1024;
}
int rowpass_rows_per_wg(int nx, int rows) {
    if (rows == kRowsWide) return nx >= 256 ? 1 : 256 / nx;
    return rows;
}
int colpass_cols_default(int ny, Arith a) {
    (void)ny;
    (void)a;
    return 8;
}
// Launch-size-aware column plan: the column pass narrows its tiles until
// the launch has at least 512 workgroups (every column plan gives the same
// bits, so this may depend on the batch).
int colpass_cols_auto(int ny, Arith a, long lines) {
    int c = colpass_cols_default(ny, a);
    while (lines > 0 && c > 2 && lines * 2 / c < 512) c /= 2;
    return c;
}
bool colpass_cols_supported(int ny, int cols) {
    return (cols == 2 || cols == 4 || cols == 8 || cols == 16) && ny / 8 * cols <= 1024;
}

hipError_t launch_rowpass(int nx, bool thermal, Arith a, int rows, dim3 grid, hipStream_t st,
                          const MixedInst* insts, Step g, const double2* tw) {
    switch (a) {
        case Arith::F32: return rowpass_f32(nx, thermal, rows, grid, st, insts, g, tw);
        case Arith::S32: return rowpass_s32(nx, thermal, rows, grid, st, insts, g, tw);
        case Arith::F64: return rowpass_f64(nx, thermal, rows, grid, st, insts, g, tw);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_colpass(int ny, Arith a, int cols, dim3 grid, hipStream_t st, const MixedInst* insts, Step g,
                          const double2* tw) {
    switch (a) {
        case Arith::F32: return colpass_f32(ny, cols, grid, st, insts, g, tw);
        case Arith::S32: return colpass_s32(ny, cols, grid, st, insts, g, tw);
        case Arith::F64: return colpass_f64(ny, cols, grid, st, insts, g, tw);
    }
    return hipErrorInvalidValue;
}

}  // namespace shg
