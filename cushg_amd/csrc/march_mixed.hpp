// march_mixed.hpp -- host-side instance state of the fused forms, shared
// by the forward march (march_mixed.cpp) and the adjoint (adjoint.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "heat.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace shg {

struct MixedHost {
    const shg_params* p = nullptr;
    shg_result* r = nullptr;
    int nx = 0, ny = 0, nz = 0, nzh = 0;
    size_t N2 = 0;
    double dx = 0, dy = 0, dz = 0;
    double Tc = 0;  // temperature at the grid centre of heat slice 0
    Arith arith = Arith::F32;
    void *P = nullptr, *S = nullptr, *outP = nullptr, *outS = nullptr;  // float2 or double2 per arith
    SliceTrace* trace = nullptr;  // per-slice trace (single runs)
    void *trP = nullptr, *trS = nullptr;
    int pass = 0;                 // thermo-optic pass being marched
    double2 *exP = nullptr, *exS = nullptr, *eyP = nullptr, *eyS = nullptr;
    double2* phase = nullptr;
    double* pol = nullptr;
    double* slots = nullptr;  // [0] P_in, [1] P_out, [2] P_sh
    HeatState heat;
    MixedInst desc{};
    bool thermal = false;
    bool running = true;  // THERMAL: still in the outer loop
    bool finalised = false;  // results written to *r (the last step of a run)
    int half = 0;         // interleave half of the batched march (0 or 1)
    int outer = 0;
    long long steps = 0, sweeps = 0;
    double heat_res = 0;
};

Arith arith_of(int precision);
int run_mixed(shg_handle* h, const std::vector<const shg_params*>& ps, const std::vector<shg_result*>& rs,
              SliceTrace* trace);
const char* rowpass_name(int nx, bool th, Arith a);
const char* colpass_name(int ny, Arith a);
// plans: the handle's set or tuned plan, else the built-in one for a launch
// covering `lines` rows (row pass) or columns (column pass) in total (0: the
// static default)
int plan_rows(shg_handle* h, int nx, Arith a, long lines = 0, bool thermal = false);
int plan_cols(shg_handle* h, int ny, Arith a, long lines = 0);
// sum of |A|^2 into a device slot, for either storage type
int field_sum_sq(shg_handle* h, Arith a, const void* A, long n, double* slot);
// device field -> interleaved FP64 host array (synchronous)
int field_to_host(shg_handle* h, Arith a, const void* A, size_t n, double* out);
// FP64 host tables (propagator factors, phase table, couplings) into m.desc
void host_tables(MixedHost& m, std::vector<double2>& ex_p, std::vector<double2>& ex_s,
                 std::vector<double2>& ey_p, std::vector<double2>& ey_s, std::vector<double2>& phase);
// the same tables into caller-provided arrays (nx, nx, ny, ny and, without
// THERMAL, 6 nz entries; phase may be null with THERMAL)
void host_tables_into(MixedHost& m, double2* ex_p, double2* ex_s, double2* ey_p, double2* ey_s, double2* phase);
// uploads the tables, initialises the pump and a zero SH, reduces P_in into slots[0]
int setup_fields(shg_handle* h, MixedHost& m);
size_t mixed_bytes(const MixedHost& m);

}  // namespace shg
