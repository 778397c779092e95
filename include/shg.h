/*
 * shg.h -- C ABI of the MI355X split-step Fourier SHG propagation library.
 *
 * The reference (cuSHG) has no library ABI: its boundary is the program
 * main(argc, argv) with five positional arguments (src/cuSHG.cu:27,46-58),
 * the method
 *     void Solver::run(real_t Power, real_t waist, real_t focalpoint,
 *                      real_t Temp, real_t T_inf, real_t Tpeltier,
 *                      bool save_only_last, bool save_temperature)
 * (src/headers/Solver.h:149,311), compile-time grid/crystal/THERMAL settings
 * (src/cuSHG.cu:11-15, src/headers/Crystal.h:26-52, src/headers/Solver.h:32)
 * and text files as outputs (src/headers/Files.h:69-90,169-188).
 *
 * This header replaces that boundary with plain C: every compile-time
 * constant of the reference becomes a field of shg_params, Solver::run
 * becomes shg_run(), and the output files become caller-owned buffers in
 * shg_result.  No C++ types, no exceptions and no torch types cross it.
 * All functions return SHG_OK (0) or a negative shg_status code; none of
 * them calls exit() (the reference's CHECK macro does,
 * src/headers/Common.h:10-20).
 *
 * Units follow the reference (src/headers/Crystal.h:19-52): micrometres,
 * picoseconds, watts, volts, degrees Celsius.
 *
 * Layout of every 2D field buffer: index = y*nx + x, x fastest
 * (the reference's IDX, src/headers/Common.h:88-90).  Complex values are
 * interleaved (re, im) doubles.  3D temperature buffers: z*nx*ny + y*nx + x.
 */
#ifndef SHG_H
#define SHG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHG_ABI_VERSION 1

#if defined(__GNUC__) || defined(__clang__)
#define SHG_API __attribute__((visibility("default")))
#else
#define SHG_API
#endif

typedef enum shg_status {
    SHG_OK = 0,
    SHG_EINVAL = -1,       /* invalid parameter (message via shg_last_error) */
    SHG_ENOMEM = -2,       /* device or host allocation failed */
    SHG_EDEVICE = -3,      /* HIP runtime / kernel error */
    SHG_ENOCONV = -4,      /* heat solver did not reach its tolerance */
    SHG_ENAN = -5,         /* non-finite value detected in the fields */
    SHG_EUNSUPPORTED = -6  /* valid request this build cannot serve */
} shg_status;

/* Crystal database.  MGO_SPPLT is the reference's only crystal
 * (src/headers/Crystal.h:54-70); MGO_PPLN is named by the reference's
 * README.md:52 but absent from its sources (literature Sellmeier). */
enum { SHG_CRYSTAL_MGO_SPPLT = 0, SHG_CRYSTAL_MGO_PPLN = 1 };

/* Split-step scheme.  REFERENCE reproduces src/headers/Solver.h:189-273
 * exactly (RK4 stage inputs A + c*k without dz, z offsets {0,dz/4,dz/4,dz/2},
 * two full-dz nonlinear calls per slice).  STRANG is the textbook
 * NL(dz/2)-D(dz)-NL(dz/2) with standard RK4 (used for closed-form checks). */
enum { SHG_SCHEME_REFERENCE = 0, SHG_SCHEME_STRANG = 1 };

/* Arithmetic form on the GPU.
 *   FP32_MIXED:   FP32 fields, FFT butterflies and RK4; FP64 phase (dk*z),
 *                 twiddle and propagator products, reductions and heat
 *                 residual.
 *   FP64_PLAIN:   FP64 everywhere, unfused kernels in the reference's op
 *                 order (one instance at a time).
 *   FP32_STORAGE: the fused, batched march of FP32_MIXED with the same FP32
 *                 fields in device memory (same HBM bytes), but every
 *                 operation between a load and a store -- FFTs, propagator,
 *                 phase, RK4 -- in FP64 registers.
 *   FP64_FUSED:   the fused, batched march with FP64 fields and FP64
 *                 arithmetic (twice the HBM bytes of the FP32-field forms).
 *   FP32_SPLIT:   FP32 fields; the row pass (row transforms, RK4, heat
 *                 source) in the FP32 arithmetic of FP32_MIXED, the column
 *                 pass (column transforms, propagator) in the FP64
 *                 arithmetic of FP32_STORAGE.  The column transforms carry
 *                 most of FP32_MIXED's error, the row pass most of the
 *                 FP64 arithmetic's cost. */
enum {
    SHG_PREC_FP32_MIXED = 0,
    SHG_PREC_FP64_PLAIN = 1,
    SHG_PREC_FP32_STORAGE = 2,
    SHG_PREC_FP64_FUSED = 3,
    SHG_PREC_FP32_SPLIT = 4
};

/* Heat solver for THERMAL runs.  JACOBI_REFERENCE: the reference's loop
 * (src/headers/Solver.h:332-356): Jacobi sweeps, convergence tested every
 * check_every sweeps as ||T_f - T_i||_2 < jacobi_tol.  DIRECT: exact solve
 * of the same discrete steady state (the fixed point of that Jacobi map)
 * by fast diagonalisation, refined until the update is below heat_tol_K. */
enum { SHG_HEAT_JACOBI_REFERENCE = 0, SHG_HEAT_DIRECT = 1 };

/* Thermal boundary set.  OPEN_TOP is the reference's (kernelHeatEquationTopOpen,
 * src/headers/Tfield.h:93-201): the y = 0 plane held at T_peltier, all other
 * faces convective to T_inf.  OVEN_SURROUNDED is the reference's unused
 * variant (kernelSetOvenSurrounded / kernelHeatEquationOvenSurrounded,
 * src/headers/Tfield.h:64-90,205-240,333-348): the four lateral planes
 * x = 0, x = lx, y = 0, y = ly held at T_peltier, the faces z = 0 and
 * z = lz convective. */
enum { SHG_HEAT_BC_OPEN_TOP = 0, SHG_HEAT_BC_OVEN_SURROUNDED = 1 };

typedef struct shg_params {
    /* grid (src/cuSHG.cu:11-15, src/headers/Crystal.h:31-36) */
    int32_t nx, ny;          /* transverse points; even; supported sizes below */
    int32_t nz;              /* z slices = SSFM steps per march (reference NZ) */
    int32_t nz_heat;         /* heat-grid slices; 0 means nz (reference) */
    double lx_um, ly_um;     /* transverse window; dx = lx/(nx-1) */
    double lz_um;            /* crystal length; dz = lz/(nz-1) */

    /* crystal (src/headers/Crystal.h:26-52) */
    int32_t crystal;         /* SHG_CRYSTAL_* */
    int32_t scheme;          /* SHG_SCHEME_* */
    double lambda_p_um;      /* pump wavelength; SH is lambda_p/2 */
    double deff_um_per_V;    /* d_eff; the QPM coefficient is dQ = 2 d_eff/pi */
    double Lambda_um;        /* QPM grating period at 25 C */
    double alpha_p, alpha_s; /* linear absorption [1/um] */
    double beta_s;           /* SH two-photon absorption [um/W] */
    double rho_p, rho_s;     /* walk-off angles [rad] */
    double k_th;             /* thermal conductivity [W/(um K)] */
    double h_conv;           /* heat-transfer coefficient [W/(um^2 K)] */
    double therm_exp_a;      /* grating expansion 1 + a(T-25) + b(T-25)^2 */
    double therm_exp_b;

    /* beam and oven (src/cuSHG.cu:46-53) */
    double power_W;          /* argv[1] */
    double waist_um;         /* argv[2] */
    double focal_um;         /* focus position; reference: 0.5*lz */
    double T_C;              /* argv[3]: crystal / phase-matching temperature */
    double T_inf_C;          /* environment temperature; reference: 25 */
    double T_peltier_C;      /* bottom oven; reference: argv[3] */

    /* THERMAL (src/headers/Solver.h:321-373) */
    int32_t thermal;         /* 0/1: the reference's -DTHERMAL */
    int32_t precision;       /* SHG_PREC_*; default SHG_PREC_FP32_STORAGE */
    int32_t heat_solver;     /* SHG_HEAT_* */
    int32_t max_outer;       /* outer thermo-optic passes; reference 20 */
    int32_t check_every;     /* Jacobi check period; reference 5000 */
    int32_t heat_bc;         /* SHG_HEAT_BC_*; reference: OPEN_TOP */
    int64_t max_sweeps;      /* Jacobi safety cap (SHG_ENOCONV beyond) */
    double jacobi_tol;       /* ||dT||_2 per sweep [K]; reference 5e-4 */
    double heat_tol_K;       /* DIRECT refinement tolerance [K] */

    /* extensions without a reference counterpart */
    double waist_y_um;       /* elliptical pump: waist along y (waist_um is the
                                x waist); 0 = circular, the reference's beam.
                                The variant commented out at
                                src/headers/Efields.h:113-115, normalised to
                                power_W with sqrt(wx wy) */
    const double* poling_phase; /* nullable; nz entries [rad], added to the
                                   QPM phase of slice s (pump -, SH +) */
} shg_params;

typedef struct shg_result {
    double eta;              /* P_SH(out) / power_W */
    double P_sh_W;           /* SH power of the saved output slice */
    double P_pump_out_W;     /* pump power of the saved output slice */
    double P_pump_in_W;      /* discrete power of the initial pump field */
    double peak_dT_K;        /* max(T) - T_peltier (THERMAL) */
    double T_max_C;          /* max(T) (THERMAL) */
    double heat_residual_K;  /* last one-sweep change / refinement update */
    double wall_s;           /* wall time of the call */
    int64_t steps_done;      /* SSFM steps executed (all outer passes) */
    int64_t jacobi_sweeps;   /* Jacobi sweeps executed (all passes) */
    int32_t outer_iters;     /* thermo-optic passes executed */
    int32_t status;          /* shg_status of this instance */

    /* optional caller-owned outputs (NULL to skip).  The fields are the
     * state the reference saves as Pump_out/Signal_out: after nz-1 SSFM
     * steps of the last pass (src/headers/Solver.h:282-287). */
    double* pump_out;        /* 2*nx*ny doubles, interleaved */
    double* signal_out;      /* 2*nx*ny doubles, interleaved */
    double* temperature_out; /* nx*ny*nz_heat doubles (THERMAL) */

    /* set by shg_adjoint only: reverse sweeps of the march it ran (1 without
     * THERMAL; the fixed-point sweeps at the thermo-optic steady state with) */
    int32_t adjoint_sweeps;
    int32_t reserved0;
} shg_result;

/* Supported transverse sizes of the GPU FFT: 64, 128, 192, 256, 384, 512
 * (nx == ny not required, each must be one of these). */
SHG_API int shg_fft_size_supported(int32_t n);

/* Fill *p with the reference defaults: src/headers/Crystal.h:19-52 (MgO:sPPLT,
 * lx=2000, ly=1000, lz=30000), src/cuSHG.cu:11-13 (128x128x100) and
 * src/cuSHG.cu:46-53 / src/cuSHG.sh:43-45 (30 W, 29 um, 55.66 C,
 * focus lz/2, T_inf 25, oven = T).  Replaces the compile-time constants. */
SHG_API void shg_params_default(shg_params* p);

/* Apply the crystal database entry `crystal` (Sellmeier id, d_eff, period,
 * absorption, conductivity, expansion) to *p. */
SHG_API int shg_params_set_crystal(shg_params* p, int32_t crystal);

/* Refractive index n(lambda, T) of the crystal (extraordinary), FP64.
 * Replaces n(L, T) of src/headers/Crystal.h:54-70. */
SHG_API double shg_refractive_index(int32_t crystal, double lambda_um, double T_C);

/* Phase mismatch dk(T) [1/um], FP64: src/headers/PhaseMatching.h:59,80. */
SHG_API double shg_delta_k(const shg_params* p, double T_C);

/* Validate parameters; SHG_OK or SHG_EINVAL (with message). */
SHG_API int shg_validate(const shg_params* p);

/* Message of the last error on this thread ("" if none). */
SHG_API const char* shg_last_error(void);

/* ---------------------------------------------------------------------
 * Handles: own device buffers, FFT plans, captured graphs and the HIP
 * stream they run on.  One handle per host thread; a handle is not
 * thread-safe.  Replaces the RAII classes Efields/Tfield/PhaseMatching/
 * Solver (src/headers/Efields.h:171-196, src/headers/Tfield.h:279-298,
 * src/headers/PhaseMatching.h:88-103, src/headers/Solver.h:116-143).
 * ------------------------------------------------------------------- */
typedef struct shg_handle shg_handle;

SHG_API int shg_handle_create(int32_t device, shg_handle** out);
SHG_API int shg_handle_destroy(shg_handle* h);

/* One full run = Solver::run (src/headers/Solver.h:311-390): setup, the
 * optional thermo-optic outer loop, the z-march (Solver::CWES,
 * src/headers/Solver.h:276-308).  Result scalars and the optional output
 * buffers are filled in *r. */
SHG_API int shg_run(shg_handle* h, const shg_params* p, shg_result* r);

/* Per-slice trace callback: the state before the SSFM step of slice
 * `slice` of thermo-optic pass `pass` (0-based), i.e. what the reference
 * writes as Pump_<s>/Signal_<s> when save_only_last is false
 * (src/headers/Solver.h:289-294).  pump/signal: 2*nx*ny interleaved
 * doubles, valid during the call only.  A non-zero return cancels the run
 * (shg_run_traced then returns SHG_EINVAL with that message). */
typedef int (*shg_slice_fn)(int32_t slice, int32_t pass, const double* pump, const double* signal,
                            void* user);
/* shg_run with the per-slice trace: fn is called for slices 0, every,
 * 2*every, ... <= nz-1 of every pass (the reference overwrites its slice
 * files on each thermo-optic pass, so the last pass's calls are the
 * reference's files).  every <= 0 means 1.  Each traced slice waits for the
 * GPU and copies the two fields to the host: a debugging path, not for
 * timing. */
SHG_API int shg_run_traced(shg_handle* h, const shg_params* p, int32_t every, shg_slice_fn fn, void* user,
                           shg_result* r);
/* n independent runs in one batched execution (replaces the one process
 * per point sweep of src/cuSHG.sh:58-94).  Instances may differ in every
 * parameter, including grid size and nz.  r[i].status holds each
 * instance's own status; the return value is the first failure or SHG_OK.
 * Like instances (same grid, THERMAL flag and precision) of at least 2^21
 * grid points in total march as two halves on two streams, one half's
 * column pass beside the other's row pass; every instance's result is
 * bit-identical to its shg_run result either way. */
SHG_API int shg_run_batch(shg_handle* h, const shg_params* p, int32_t n, shg_result* r);

/* Gradient of eta by a hand-written adjoint of the discrete SSFM/RK4
 * march with checkpointed recomputation.
 *   d_eta[0] = d eta / d waist_um
 *   d_eta[1] = d eta / d focal_um
 *   d_eta[2] = d eta / d T, with T = T_C = T_peltier_C moved together (the
 *              reference's argv[3], src/cuSHG.cu:46-53): the crystal
 *              temperature without THERMAL, the oven temperature with it
 *   d_eta[3 + s] = d eta / d poling_phase[s], s = 0..nz-1
 * d_eta must hold 3 + nz doubles.  checkpoint_every <= 0 selects a default
 * (with THERMAL: the whole last march when its history fits in a quarter of
 * the free device memory).
 * THERMAL runs (SHG_HEAT_DIRECT only) are differentiated at the converged
 * thermo-optic steady state of the reference's outer loop
 * (src/headers/Solver.h:321-373): the implicit-function adjoint through the
 * heat solve, the heat source and the temperature dependence of n and dk,
 * solved by Anderson-accelerated fixed-point sweeps.
 * Arithmetic: FP64 for the FP64 forms; every FP32-field form is
 * differentiated with FP32 fields and FP64 arithmetic (the FP32_STORAGE
 * march), so they give the same result.
 * No reference counterpart (the reference computes no gradients). */
SHG_API int shg_adjoint(shg_handle* h, const shg_params* p, int32_t checkpoint_every,
                double* d_eta, shg_result* r);

/* Drop-in for the reference program: same five positional arguments as
 * src/cuSHG.cu:46-58 (power, waist, T, save_only_last, save_temperature),
 * runs with *base (or defaults if NULL) and writes Pump_out_r.dat,
 * Pump_out_i.dat, Signal_out_r.dat, Signal_out_i.dat (and Tfinal_<z>.dat)
 * into out_dir in the text format of src/headers/Files.h:69-90,169-188. */
SHG_API int shg_cli_main(int argc, const char* const* argv, const shg_params* base,
                 const char* out_dir);

/* Timing of the last shg_run/shg_run_batch on this handle, per kernel
 * class, measured with HIP events on the handle's stream (milliseconds).
 * names/ms hold up to cap entries; returns the number of entries. */
SHG_API int shg_last_kernel_times(shg_handle* h, const char** names, double* ms,
                          int32_t* launches, int32_t cap);

/* Enable (1) or disable (0) per-launch HIP-event timing on the handle's
 * stream; read the totals with shg_last_kernel_times.  Resets the totals. */
SHG_API int shg_handle_set_profiling(shg_handle* h, int32_t on);

/* Row-pass plan value of the wide form: n threads per row, one point each
 * in the nonlinear step, max(1, 256/n) rows per workgroup -- for launches
 * too small to fill the GPU with n/8 threads per row. */
#define SHG_PLAN_ROWS_WIDE 16

/* Launch plan of the fused passes for transform length n and precision
 * (SHG_PREC_FP32_MIXED, _FP32_STORAGE, _FP64_FUSED or _FP32_SPLIT): rows per workgroup of
 * the row pass (1, 2, 4, 8 with n/8 threads per row holding 8 points each,
 * or SHG_PLAN_ROWS_WIDE) and columns per workgroup of the column pass
 * (2, 4, 8, 16); 0 keeps the current choice.  A row plan set here holds
 * for THERMAL and non-THERMAL launches alike.  Replaces the reference's
 * fixed 16x16 blocks (src/headers/Solver.h:119-135) and per-slice cuFFT
 * plans (src/headers/Solver.h:159,183). */
SHG_API int shg_set_plan(shg_handle* h, int32_t n, int32_t precision, int32_t rows_per_wg, int32_t cols_per_wg);

/* The plan the handle uses for (n, precision): tuned or set, else the
 * built-in plan for one non-THERMAL n x n instance (built-in plans follow
 * the launch size and, for the row pass, THERMAL: every plan gives the same
 * results bit for bit). */
SHG_API int shg_get_plan(shg_handle* h, int32_t n, int32_t precision, int32_t* rows_per_wg, int32_t* cols_per_wg);

/* Autotuner: time every supported plan for transform length n on a
 * synthetic non-THERMAL march of `batch` n x n instances with nz slices
 * (HIP events around each launch) and keep the fastest row and column
 * plans on the handle.  The row pass is timed again on a THERMAL march of
 * the same shape (one thermo-optic pass, 4 heat slices) and that choice is
 * kept for THERMAL launches (shg_get_plan reports the non-THERMAL one).
 * No reference counterpart. */
SHG_API int shg_autotune(shg_handle* h, int32_t n, int32_t precision, int32_t batch, int32_t nz);

/* Graph replay of the fused march: `steps` > 0 captures the launches of
 * that many consecutive z-steps (every launch group's row and column pass;
 * independent groups as parallel branches) once into a hipGraph, cached on
 * the handle, and replays it along the march; 0 launches every kernel
 * directly; -1 (the default) replays graphs of 32 steps for batches that
 * march on several streams (launch groups that share no instance, or the
 * two halves of a large class of like instances, see shg_run_batch) and
 * launches directly otherwise.
 * Runs with a per-slice trace or with profiling on always launch directly.  Replaces the reference's
 * per-slice host loop with its synchronisations (src/headers/Solver.h:276-308). */
SHG_API int shg_set_graph_steps(shg_handle* h, int32_t steps);

/* The library's 2D DFT on a caller's array: ny rows of nx complex points
 * (interleaved doubles, x fastest), unnormalised, forward exp(-2 pi i jk/N)
 * (inverse != 0: exp(+2 pi i jk/N)).  The same transform code, twiddle
 * tables and arithmetic as the fused passes that replace cufftExecC2C
 * (src/headers/Solver.h:159-183): FP64 for SHG_PREC_FP64_PLAIN/_FUSED; FP32
 * points with FP64 twiddle products for the FP32 forms (inputs rounded to
 * FP32).  nx and ny as shg_fft_size_supported; allocates its device buffer
 * (SHG_ENOMEM on failure). */
SHG_API int shg_fft2d(shg_handle* h, int32_t nx, int32_t ny, int32_t precision, int32_t inverse, const double* in,
                      double* out);

/* ---------------------------------------------------------------------
 * Zero-allocation run.  No counterpart in the reference: this entry point
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 *
 * shg_run_noalloc is shg_run for one non-THERMAL instance of a fused form
 * (SHG_PREC_FP32_SPLIT, _FP32_STORAGE, _FP32_MIXED, _FP64_FUSED) that calls
 * no hipMalloc, hipFree or other allocating routine: every device buffer of
 * the run (fields, saved output, propagator and phase tables, FFT twiddles,
 * the instance descriptor, reduction slots, output staging) lives in a
 * fixed device scratch array of SHG_NOALLOC_SCRATCH_MIB MiB that the
 * library holds on each GPU it runs on until the process ends (attached by
 * shg_handle_create).  The handle must exist (create it at set-up); the
 * run uses the handle's stream and plans and gives shg_run's results bit
 * for bit.  Limits: nz <= SHG_NOALLOC_MAX_NZ (SHG_EINVAL beyond), THERMAL
 * and SHG_PREC_FP64_PLAIN return SHG_EUNSUPPORTED.  All calls on one GPU
 * share that GPU's scratch, which keeps the last call's contents: calls
 * that overlap on the same GPU (from several threads or handles) race. */
#define SHG_NOALLOC_SCRATCH_MIB 20
#define SHG_NOALLOC_MAX_NZ 16384
SHG_API int shg_run_noalloc(shg_handle* h, const shg_params* p, shg_result* r);

#ifdef __cplusplus
}
#endif

#endif /* SHG_H */
