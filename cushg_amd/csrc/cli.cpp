// cli.cpp -- shg_cli_main: the reference program as a library call.
//
// Keeps the reference's command line (src/cuSHG.cu:46-58): five positional
// arguments Power[W] waist[um] T[C] save_only_last save_temperature, the
// focus at lz/2, T_inf = 25 C and the bottom oven at T (src/cuSHG.cu:48-53),
// and its output files in their text format (src/headers/Files.h:69-90,
// 169-188): Pump_out_{r,i}.dat and Signal_out_{r,i}.dat (save_only_last),
// Pump_<s>_{r,i}.dat / Signal_<s>_{r,i}.dat for every slice otherwise
// (src/headers/Solver.h:282-294), and Tfinal_<z>.dat per heat slice
// (src/headers/Solver.h:372-373).
//
// A comma-separated list in any of the first three arguments runs the
// Cartesian product of the lists (power outermost, then waist, then T, the
// loop order of the sweep script src/cuSHG.sh:58-94) as one batched run and
// writes each point into its own folder named as that script names it.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "runtime.hpp"

namespace shg {
namespace {

std::vector<std::string> split_list(const char* a) {
    std::vector<std::string> out;
    std::string cur;
    for (const char* c = a; *c; c++) {
        if (*c == ',') {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
        } else {
            cur += *c;
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

bool parse_double(const std::string& s, double& v) {
    char* end = nullptr;
    errno = 0;
    v = std::strtod(s.c_str(), &end);
    return errno == 0 && end && *end == '\0' && std::isfinite(v);
}

int make_dir(const std::string& d) {
    if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) {
        set_error("cannot create directory %s: %s", d.c_str(), std::strerror(errno));
        return SHG_EINVAL;
    }
    return SHG_OK;
}

// saveMatrixComplex of src/headers/Files.h:69-90: <name>_r.dat and
// <name>_i.dat, ny lines of nx tab-terminated values, 20 significant digits
int save_matrix_complex(const std::string& dir, const std::string& name, const double* A, int nx, int ny) {
    for (int part = 0; part < 2; part++) {
        const std::string fn = dir + "/" + name + (part == 0 ? "_r.dat" : "_i.dat");
        std::ofstream f(fn);
        if (!f) {
            set_error("cannot write %s", fn.c_str());
            return SHG_EINVAL;
        }
        f << std::setprecision(20);
        for (int y = 0; y < ny; y++) {
            for (int x = 0; x < nx; x++) f << A[2 * ((size_t)y * nx + x) + part] << "\t";
            f << "\n";
        }
    }
    return SHG_OK;
}

// saveTensorReal of src/headers/Files.h:169-188: <name>_<z>.dat per slice,
// default stream precision
int save_tensor_real(const std::string& dir, const std::string& name, const double* T, int nx, int ny, int nz) {
    for (int z = 0; z < nz; z++) {
        const std::string fn = dir + "/" + name + "_" + std::to_string(z) + ".dat";
        std::ofstream f(fn);
        if (!f) {
            set_error("cannot write %s", fn.c_str());
            return SHG_EINVAL;
        }
        for (int y = 0; y < ny; y++) {
            for (int x = 0; x < nx; x++) f << T[((size_t)z * ny + y) * nx + x] << "\t";
            f << "\n";
        }
    }
    return SHG_OK;
}

struct SliceWriter {
    std::string dir;
    int nx, ny;
};

int write_slice(int32_t slice, int32_t /*pass*/, const double* pump, const double* signal, void* user) {
    auto* w = static_cast<SliceWriter*>(user);
    const std::string s = std::to_string(slice);
    if (save_matrix_complex(w->dir, "Pump_" + s, pump, w->nx, w->ny)) return 1;
    if (save_matrix_complex(w->dir, "Signal_" + s, signal, w->nx, w->ny)) return 1;
    return 0;
}

void print_params(const shg_params& p) {
    std::printf("\n\nPump power =  %g W\n", p.power_W);
    std::printf("Beam waist = %g um\n", p.waist_um);
    std::printf("Focal distance = %g um\n", p.focal_um);
    std::printf("Phase-matching temperature =  %g C\n", p.T_C);
    std::printf("Environment temperature =  %g C\n", p.T_inf_C);
    std::printf("Peltier 1 temperature =  %g C\n", p.T_peltier_C);
    std::printf("Crystal: %s, grid %d x %d x %d (heat z-slices %d), window %g x %g x %g um, THERMAL %s, %s\n",
                p.crystal == SHG_CRYSTAL_MGO_PPLN ? "MgO:PPLN" : "MgO:sPPLT", p.nx, p.ny, p.nz,
                p.nz_heat > 0 ? p.nz_heat : p.nz, p.lx_um, p.ly_um, p.lz_um, p.thermal ? "on" : "off",
                p.precision == SHG_PREC_FP64_PLAIN      ? "FP64 plain"
                : p.precision == SHG_PREC_FP64_FUSED    ? "FP64 fused"
                : p.precision == SHG_PREC_FP32_STORAGE  ? "FP32 storage"
                : p.precision == SHG_PREC_FP32_SPLIT    ? "FP32 split"
                                                        : "FP32 mixed");
}

void print_result(const shg_result& r, bool thermal) {
    std::printf("SH efficiency eta = %.10g  (P_SH = %.10g W, pump out %.10g W)\n", r.eta, r.P_sh_W, r.P_pump_out_W);
    if (thermal)
        std::printf("Peak temperature rise = %.8g K (T_max %.8g C), %d thermo-optic passes, %lld Jacobi sweeps\n",
                    r.peak_dT_K, r.T_max_C, r.outer_iters, (long long)r.jacobi_sweeps);
    std::printf("%lld SSFM steps in %.3f s\n", (long long)r.steps_done, r.wall_s);
}

}  // namespace
}  // namespace shg

using namespace shg;

extern "C" int shg_cli_main(int argc, const char* const* argv, const shg_params* base, const char* out_dir) {
    clear_error();
    if (argc < 6 || !argv) {
        set_error("usage: shg Power[W] waist[um] T[C] save_only_last save_temperature "
                  "(comma-separated lists in the first three run a sweep)");
        return SHG_EINVAL;
    }
    shg_params p0;
    if (base) {
        p0 = *base;
    } else {
        shg_params_default(&p0);
    }
    const std::string out = out_dir && *out_dir ? out_dir : ".";
    const std::vector<std::string> Ps = split_list(argv[1]), Ws = split_list(argv[2]), Ts = split_list(argv[3]);
    if (Ps.empty() || Ws.empty() || Ts.empty()) {
        set_error("empty Power, waist or T argument");
        return SHG_EINVAL;
    }
    const bool save_only_last = std::atoi(argv[4]) != 0;
    const bool save_temperature = std::atoi(argv[5]) != 0;

    struct Point {
        shg_params p;
        std::string dir;
        shg_result r;
        std::vector<double> pump, signal, T;
    };
    std::vector<Point> pts;
    const bool sweep = Ps.size() * Ws.size() * Ts.size() > 1;
    const char* tag = p0.crystal == SHG_CRYSTAL_MGO_PPLN ? "PPLN" : "sPPLT";
    for (const auto& a : Ps)
        for (const auto& b : Ws)
            for (const auto& c : Ts) {
                Point pt;
                pt.p = p0;
                double P, W, T;
                if (!parse_double(a, P) || !parse_double(b, W) || !parse_double(c, T)) {
                    set_error("not a number in '%s' '%s' '%s'", a.c_str(), b.c_str(), c.c_str());
                    return SHG_EINVAL;
                }
                pt.p.power_W = P;
                pt.p.waist_um = W;
                pt.p.T_C = T;
                pt.p.T_peltier_C = T;                 // src/cuSHG.cu:53
                pt.p.T_inf_C = 25.0;                  // src/cuSHG.cu:52
                pt.p.focal_um = 0.5 * pt.p.lz_um;     // src/cuSHG.cu:48
                pt.p.poling_phase = p0.poling_phase;  // a grating profile from the caller, if any
                pt.dir = sweep ? out + "/" + tag + "_P_" + a + "_W_T1_" + c + "_WAIST_" + b : out;
                if (int rc = shg_validate(&pt.p)) return rc;
                pts.push_back(pt);
            }
    if (int rc = make_dir(out)) return rc;
    std::printf("#######---SHG efficiency calculator---#######\n\n");
    std::fflush(stdout);

    shg_handle* h = nullptr;
    if (int rc = shg_handle_create(0, &h)) return rc;
    struct HandleGuard {
        shg_handle* h;
        ~HandleGuard() { shg_handle_destroy(h); }
    } guard{h};

    for (auto& pt : pts) {
        std::memset(&pt.r, 0, sizeof(pt.r));
        const size_t n2 = (size_t)pt.p.nx * pt.p.ny;
        const int nzh = pt.p.nz_heat > 0 ? pt.p.nz_heat : pt.p.nz;
        if (save_only_last) {
            pt.pump.resize(2 * n2);
            pt.signal.resize(2 * n2);
            pt.r.pump_out = pt.pump.data();
            pt.r.signal_out = pt.signal.data();
        }
        if (save_temperature && pt.p.thermal) {
            pt.T.resize(n2 * nzh);
            pt.r.temperature_out = pt.T.data();
        }
        if (sweep)
            if (int rc = make_dir(pt.dir)) return rc;
    }
    if (save_only_last) {
        // one batched run over every point
        std::vector<shg_params> pv;
        std::vector<shg_result> rv;
        for (auto& pt : pts) {
            pv.push_back(pt.p);
            rv.push_back(pt.r);
        }
        const int rc = shg_run_batch(h, pv.data(), (int32_t)pv.size(), rv.data());
        for (size_t i = 0; i < pts.size(); i++) pts[i].r = rv[i];
        if (rc) return rc;
    } else {
        for (auto& pt : pts) {
            SliceWriter w{pt.dir, pt.p.nx, pt.p.ny};
            if (int rc = shg_run_traced(h, &pt.p, 1, write_slice, &w, &pt.r)) return rc;
        }
    }
    for (auto& pt : pts) {
        print_params(pt.p);
        print_result(pt.r, pt.p.thermal != 0);
        if (save_only_last) {
            if (int rc = save_matrix_complex(pt.dir, "Pump_out", pt.pump.data(), pt.p.nx, pt.p.ny)) return rc;
            if (int rc = save_matrix_complex(pt.dir, "Signal_out", pt.signal.data(), pt.p.nx, pt.p.ny)) return rc;
        }
        if (!pt.T.empty()) {
            const int nzh = pt.p.nz_heat > 0 ? pt.p.nz_heat : pt.p.nz;
            if (int rc = save_tensor_real(pt.dir, "Tfinal", pt.T.data(), pt.p.nx, pt.p.ny, nzh)) return rc;
        }
        std::fflush(stdout);
    }
    return SHG_OK;
}
