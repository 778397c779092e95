// kernels.hpp -- device-side instance descriptors and launch entry points.
#pragma once

#include "common.hpp"

namespace shg {

// Storage / arithmetic pair of the fused passes.
enum class Arith {
    F32,  // FP32 fields, FP32 arithmetic          (SHG_PREC_FP32_MIXED)
    S32,  // FP32 fields, FP64 arithmetic          (SHG_PREC_FP32_STORAGE)
    F64,  // FP64 fields, FP64 arithmetic          (SHG_PREC_FP64_FUSED)
};
inline size_t arith_field_bytes(Arith a) { return a == Arith::F64 ? 16 : 8; }

// One instance of a fused march, as the kernels see it.  Lives in device
// memory (an array per launch group, indexed by blockIdx.z).  Field
// pointers point to float2 or double2 per the launch's form.  Constants
// are FP64; the FP32-arithmetic kernels round them once at use.
struct MixedInst {
    void* P;                // pump field [ny][nx]
    void* S;                // SH field   [ny][nx]
    void* outP;             // state after nz-1 slices (the reference's Pump_out)
    void* outS;
    void* trP;              // per-slice trace: state before slice g's SSFM step (or null)
    void* trS;
    void* hist;             // column-pass outputs of slices [hist_g0, hist_g0 + hist_n) (or null)
    const double2* exP;     // separable propagator, x factor [nx] (includes 1/(nx ny))
    const double2* exS;
    const double2* eyP;     // y factor [ny]
    const double2* eyS;
    const double2* phase;   // non-THERMAL: exp(i phi) [nz][2 calls][3 offsets]
    const float2* nn;       // THERMAL: (n_p(T), n_s(T)) per heat voxel [nzh][ny][nx]
    const double* dkbase;   // THERMAL: running dk integral = dkbase[h] + s*dk[h]
    const double* dk;       // THERMAL: dk(T) [nzh][ny][nx]
    float* Q;               // THERMAL: heat source [nzh][ny][nx]
    const double* poling;   // THERMAL: per-slice poling phase [nz] or null
    double dz;
    double kp, ks, bfac;    // non-THERMAL couplings at T_C
    double kp0, ks0, bf0;   // THERMAL: kp = kp0/n_p(T), ks = ks0/n_s(T), bfac = bf0 n_s(T)
    double ap_h, as_h;      // alpha/2
    double c1, c2, c3, w;   // RK4 stage coefficients and final weight
    double off[2][3];       // z offsets (units of dz) per call
    double hce;             // c eps0 / 2
    double alpha_p, alpha_s, beta_s;
    int nx, ny, nz, nzh;
    int crystal;
    int trace_every;
    int hist_g0, hist_n;
};

// fused passes (kernels_mixed*.hip)
hipError_t launch_rowpass(int nx, bool thermal, Arith a, int rows_pw, dim3 grid, hipStream_t st,
                          const MixedInst* insts, int g, const double2* tw);
hipError_t launch_colpass(int ny, Arith a, int cols_pw, dim3 grid, hipStream_t st, const MixedInst* insts,
                          int g, const double2* tw);
int rowpass_rows_default(int nx, Arith a);
int colpass_cols_default(int ny, Arith a);
bool colpass_cols_supported(int ny, int cols);
bool rowpass_rows_supported(int nx, int rows);

// utility kernels (kernels_util.hip)
template <typename C>
hipError_t launch_pump_init(hipStream_t st, C* P, int nx, int ny, double dx, double dy, double w2,
                            double Ap0, double Mre, double Mim);
hipError_t launch_sumsq_f32(hipStream_t st, const float2* A, long n, double* partials, int nblocks);
hipError_t launch_sumsq_f64(hipStream_t st, const double2* A, long n, double* partials, int nblocks);
hipError_t launch_sum_partials(hipStream_t st, const double* partials, int n, double* out);
hipError_t launch_f64_to_f32(hipStream_t st, const double* in, float* out, long n);

}  // namespace shg
