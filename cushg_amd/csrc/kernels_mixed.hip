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

// Built-in tile plans: rows per workgroup of the row pass and columns per
// workgroup of the column pass, from shg_autotune runs on MI355X over
// batches of 1, 4 and 16 instances per size (they agree with the tuned
// plans at most sizes; shg_autotune refines them for a given workload).
// shg_autotune times the optics-only row pass; the THERMAL row pass at 512
// in FP32 arithmetic, with its heat-voxel loads, runs faster at one row per
// workgroup (2357 against 2526 ms per run of a 16 x 512 x 512 x 4000 THERMAL
// batch).
int rowpass_rows_default(int nx, Arith a, bool thermal) {
    const bool fp64 = row_arith(a) != Arith::F32;
    switch (nx) {
        case 64: return fp64 ? 4 : 8;
        case 128: return fp64 ? 2 : 4;
        case 192: return 2;
        case 256: return 2;
        case 384: return 4;
        case 512: return fp64 || thermal ? 1 : 2;
    }
    return 0;
}
// A launch whose NX/8-thread rows make fewer than 1024 waves (4 per CU)
// leaves the chip underfilled and takes the wide form (8x the threads, one
// point each in the nonlinear step): one 512 x 512 THERMAL instance runs
// at 16 us against 24 us per launch, while a 16 x 512 x 512 batch keeps
// NX/8 threads per row (108 us against 160 us).  Both forms give the same
// bits (FMAs are contracted per source expression only), so the choice may
// follow the launch size.
int rowpass_rows_auto(int nx, Arith a, long lines, bool thermal) {
    if (lines > 0 && lines * (nx / 8) < 1024L * 64) return kRowsWide;
    return rowpass_rows_default(nx, a, thermal);
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
    const bool fp64 = col_arith(a) != Arith::F32;
    switch (ny) {
        case 64: return 8;
        case 128: return 4;
        case 192: return 16;
        case 256: return fp64 ? 16 : 8;
        case 384: return fp64 ? 16 : 4;
        case 512: return 8;
    }
    return 8;
}
// Launch-size-aware column plan: the column pass narrows its tiles until
// the launch has at least 256 workgroups (every column plan gives the same
// bits, so this may depend on the batch).
int colpass_cols_auto(int ny, Arith a, long lines) {
    int c = colpass_cols_default(ny, a);
    while (lines > 0 && c > 2 && lines * 2 / c < 256) c /= 2;
    return c;
}
bool colpass_cols_supported(int ny, int cols) {
    return (cols == 2 || cols == 4 || cols == 8 || cols == 16) && ny / 8 * cols <= 1024;
}

hipError_t launch_rowpass(int nx, bool thermal, Arith a, int rows, dim3 grid, hipStream_t st,
                          const MixedInst* insts, Step g, const double2* tw) {
    switch (row_arith(a)) {
        case Arith::F32: return rowpass_f32(nx, thermal, rows, grid, st, insts, g, tw);
        case Arith::S32: return rowpass_s32(nx, thermal, rows, grid, st, insts, g, tw);
        case Arith::F64: return rowpass_f64(nx, thermal, rows, grid, st, insts, g, tw);
        default: break;
    }
    return hipErrorInvalidValue;
}

hipError_t launch_colpass(int ny, Arith a, int cols, dim3 grid, hipStream_t st, const MixedInst* insts, Step g,
                          const double2* tw) {
    switch (col_arith(a)) {
        case Arith::F32: return colpass_f32(ny, cols, grid, st, insts, g, tw);
        case Arith::S32: return colpass_s32(ny, cols, grid, st, insts, g, tw);
        case Arith::F64: return colpass_f64(ny, cols, grid, st, insts, g, tw);
        default: break;
    }
    return hipErrorInvalidValue;
}

}  // namespace shg
