// heat.cpp -- THERMAL host orchestration: the steady heat solve on the GPU
// and the optics inputs derived from T.
//
// Two solvers reach the same discrete steady state, the fixed point of the
// reference's Jacobi map (src/headers/Tfield.h:93-201):
//   JACOBI_REFERENCE  the reference loop itself (src/headers/Solver.h:332-356):
//                     FP64 sweeps, ||T_f - T_i||_2 tested every check_every.
//   DIRECT            defect correction  T <- T + M^-1 r,  r = A_ii (F(T) - T)
//                     the FP64 residual of that same map, M^-1 the exact
//                     inverse of its tensor-product finite-volume form by
//                     fast diagonalisation (three axis transforms = plain
//                     GEMMs, FP32 in the mixed form, FP64 in the plain form),
//                     repeated until max|update| < heat_tol_K.
#include <hipblaslt/hipblaslt.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "heat.hpp"
#include "heat_kernels.hpp"
#include "runtime.hpp"

namespace shg {

HeatCoef heat_coef(const shg_params* p, int nzh) {
    HeatCoef c{};
    c.nx = p->nx;
    c.ny = p->ny;
    c.nz = nzh;
    c.oven = p->heat_bc == SHG_HEAT_BC_OVEN_SURROUNDED;
    c.x0 = c.oven ? 1 : 0;
    c.nxu = c.oven ? p->nx - 2 : p->nx;
    c.nyu = c.oven ? p->ny - 2 : p->ny - 1;
    const double dx = p->lx_um / (p->nx - 1), dy = p->ly_um / (p->ny - 1), dz = p->lz_um / (nzh - 1);
    const double kt = p->k_th, hc = p->h_conv;
    c.kth = kt;
    c.Tinf = p->T_inf_C;
    c.idx2 = 1.0 / (dx * dx);
    c.idy2 = 1.0 / (dy * dy);
    c.idz2 = 1.0 / (dz * dz);
    c.L2 = 1.0 / (2 / (dx * dx) + 2 / (dy * dy) + 2 / (dz * dz));
    c.Bix = hc * dx / kt;
    c.Biy = hc * dy / kt;
    c.Biz = hc * dz / kt;
    c.ax = dx * dx / (2 * dy * dy);
    c.ay = dy * dy / (2 * dx * dx);
    c.az = dz * dz / (2 * dx * dx);
    c.bx = dx * dx / (2 * dz * dz);
    c.by = dy * dy / (2 * dz * dz);
    c.bz = dz * dz / (2 * dy * dy);
    c.ikth = 1.0 / kt;
    c.Fx = 1.0 / (1 + 2 * c.ax + 2 * c.bx + c.Bix);
    c.Fy = 1.0 / (1 + 2 * c.ay + 2 * c.by + c.Biy);
    c.Fz = 1.0 / (1 + 2 * c.az + 2 * c.bz + c.Biz);
    c.av = dz / (2 * dx);
    c.bv = dx / (2 * dz);
    c.cv = dx * dz / (4 * dy * dy);
    c.V = 1 / (c.av + c.bv + 2 * c.cv + 0.5 * (c.Bix + c.Biz));
    c.ah = dz / (2 * dy);
    c.bh = dy / (2 * dz);
    c.ch = dy * dz / (4 * dx * dx);
    c.H = 1 / (c.ah + c.bh + 2 * c.ch + 0.5 * (c.Biy + c.Biz));
    c.ad = dx / (2 * dy);
    c.bd = dy / (2 * dx);
    c.cd = dx * dy / (4 * dz * dz);
    c.D = 1 / (c.ad + c.bd + 2 * c.cd + 0.5 * (c.Bix + c.Biy));
    c.AV = dy * dz / dx;
    c.BV = dx * dz / dy;
    c.CV = dx * dy / dz;
    c.BiD = c.Bix * dy + c.Biy * dz + c.Biz * dx;
    c.GV = 1 / (c.AV + c.BV + c.CV + c.BiD);
    return c;
}

// Eigen-decomposition of a symmetric tridiagonal matrix (diagonal d, first
// off-diagonal e) by implicit QL sweeps with a Wilkinson shift
// (Golub & Van Loan, Matrix Computations, sec. 8.3).  On return d holds the
// eigenvalues and z[j*n + k] component j of eigenvector k.
static bool sym_tridiag_eig(int n, std::vector<double>& d, std::vector<double> e, std::vector<double>& z) {
    z.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) z[(size_t)i * n + i] = 1.0;
    e.resize(n + 1, 0.0);
    e[n - 1] = 0.0;
    for (int l = 0; l < n; l++) {
        for (int sweep = 0;; sweep++) {
            int m = l;
            while (m < n - 1) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) + dd == dd) break;
                m++;
            }
            if (m == l) break;
            if (sweep > 300) return false;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            bool early = false;
            for (int i = m - 1; i >= l; i--) {
                double f = s * e[i];
                const double b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    early = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                for (int k = 0; k < n; k++) {
                    double* zk = &z[(size_t)k * n];
                    f = zk[i + 1];
                    zk[i + 1] = s * zk[i] + c * f;
                    zk[i] = c * zk[i] - s * f;
                }
            }
            if (early) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    return true;
}

// One axis of the finite-volume form (values divided by k_th): node widths
// w, stiffness diagonal k (1/h per neighbour link, + h_conv/k_th at Robin
// ends), eigenpairs of the symmetric W^-1/2 K W^-1/2.
//
// Folded axis (both ends of the same kind, even length n = 2m): the matrix
// is then persymmetric, its eigenvectors are [a; Ja]/sqrt2 (even) or
// [a; -Ja]/sqrt2 (odd, J the reversal), and a runs over the eigenvectors of
// the m x m leading block with the coupling o_{m-1} across the middle added
// to (even) or subtracted from (odd) its last diagonal entry.  q then holds
// the even basis (m x m) followed by the odd one, lam the even eigenvalues
// followed by the odd ones.
struct AxisHost {
    int n = 0;
    bool fold = false;
    std::vector<double> k, w, lam, q;
};
// An end marked Dirichlet couples its node to a fixed value one spacing
// beyond it (a removed unknown, full-width control volume).
static bool axis_build(int n, double h, bool robin_lo, bool robin_hi, bool dirichlet_lo, bool dirichlet_hi, double hk,
                       bool foldable, AxisHost& a) {
    a.n = n;
    a.fold = foldable && n >= 4 && n % 2 == 0 && robin_lo == robin_hi && dirichlet_lo == dirichlet_hi;
    a.k.resize(n);
    a.w.resize(n);
    std::vector<double> diag(n), off(std::max(n - 1, 1), 0.0);
    for (int i = 0; i < n; i++) {
        const bool lo_end = i == 0 && !dirichlet_lo, hi_end = i == n - 1 && !dirichlet_hi;
        a.w[i] = (lo_end || hi_end) ? 0.5 * h : h;
        double kd = 0;
        if (i > 0 || dirichlet_lo) kd += 1.0 / h;
        if (i < n - 1 || dirichlet_hi) kd += 1.0 / h;
        if (i == 0 && robin_lo) kd += hk;
        if (i == n - 1 && robin_hi) kd += hk;
        a.k[i] = kd;
        diag[i] = kd / a.w[i];
    }
    for (int i = 0; i + 1 < n; i++) off[i] = -1.0 / h / std::sqrt(a.w[i] * a.w[i + 1]);
    if (!a.fold) {
        a.lam = diag;
        return sym_tridiag_eig(n, a.lam, off, a.q);
    }
    const int m = n / 2;
    a.lam.clear();
    a.q.clear();
    for (int sign : {1, -1}) {
        std::vector<double> d(diag.begin(), diag.begin() + m), e(off.begin(), off.begin() + (m - 1)), z;
        d[m - 1] += sign * off[m - 1];
        if (!sym_tridiag_eig(m, d, e, z)) return false;
        a.lam.insert(a.lam.end(), d.begin(), d.end());
        a.q.insert(a.q.end(), z.begin(), z.end());
    }
    return true;
}

// Device-side tables of the direct solver (carved from the run arena).
struct DirectTables {
    HeatAxes ax{};
    const void* q[3] = {nullptr, nullptr, nullptr};  // eigenvector matrices (R precision)
    bool fold[3] = {false, false, false};           // axis transformed in its even/odd fold
    void* u = nullptr;
    void* v = nullptr;
    bool fp64 = false;
};

// Per-geometry tables of the direct solver (axis widths, stiffness
// diagonals, eigenpairs in FP64 and FP32), built once per geometry and kept
// on the handle: the host eigen-decompositions are O(n^3).
struct HeatGeom {
    HeatAxes ax{};
    bool fold[3] = {};
    void* mem[9] = {};
    double* q64[3] = {};
    float* q32[3] = {};
};

static std::string geom_key(const shg_params* p, const HeatState& hs) {
    char b[256];
    snprintf(b, sizeof b, "%d/%d/%d/%.17g/%.17g/%.17g/%.17g/%d", hs.nx, hs.ny, hs.nzh, hs.dx, hs.dy, hs.dz,
             p->h_conv / p->k_th, p->heat_bc);
    return b;
}

// On failure returns nullptr with *rc = SHG_ENOCONV (eigen-solver) or
// SHG_ENOMEM (device tables).
static HeatGeom* heat_geom(shg_handle* h, const shg_params* p, const HeatState& hs, int* rc) {
    *rc = SHG_OK;
    const std::string key = geom_key(p, hs);
    auto it = h->heat_cache.find(key);
    if (it != h->heat_cache.end()) return it->second;
    // open top: x (Robin both ends), y' = y-1 (Dirichlet below, Robin top), z (Robin both);
    // oven-surrounded: x' = x-1 and y' = y-1 (Dirichlet both ends), z (Robin both).
    // x and z are folded when their length is even (y is transformed whole).
    const double hk = p->h_conv / p->k_th;
    const bool oven = p->heat_bc == SHG_HEAT_BC_OVEN_SURROUNDED;
    AxisHost axes[3];
    const bool ok = (oven ? axis_build(hs.nx - 2, hs.dx, false, false, true, true, hk, true, axes[0])
                          : axis_build(hs.nx, hs.dx, true, true, false, false, hk, true, axes[0])) &&
                    (oven ? axis_build(hs.ny - 2, hs.dy, false, false, true, true, hk, false, axes[1])
                          : axis_build(hs.ny - 1, hs.dy, false, true, true, false, hk, false, axes[1])) &&
                    axis_build(hs.nzh, hs.dz, true, true, false, false, hk, true, axes[2]);
    if (!ok) {
        set_error("heat: tridiagonal eigen-solver did not converge");
        *rc = SHG_ENOCONV;
        return nullptr;
    }
    HeatGeom* g = new HeatGeom();
    const double** kp[3] = {&g->ax.kx, &g->ax.ky, &g->ax.kz};
    const double** wp[3] = {&g->ax.wx, &g->ax.wy, &g->ax.wz};
    const double** lp[3] = {&g->ax.lx, &g->ax.ly, &g->ax.lz};
    bool fail = false;
    for (int i = 0; i < 3 && !fail; i++) {
        const int n = axes[i].n;
        const size_t nq = axes[i].q.size();  // n*n, or 2 (n/2)^2 folded
        g->fold[i] = axes[i].fold;
        double* d = nullptr;
        fail |= hipMalloc(reinterpret_cast<void**>(&d), 8 * 3 * (size_t)n) != hipSuccess;
        if (fail) break;
        g->mem[3 * i] = d;
        fail |= hipMemcpy(d, axes[i].k.data(), 8 * n, hipMemcpyHostToDevice) != hipSuccess;
        fail |= hipMemcpy(d + n, axes[i].w.data(), 8 * n, hipMemcpyHostToDevice) != hipSuccess;
        fail |= hipMemcpy(d + 2 * n, axes[i].lam.data(), 8 * n, hipMemcpyHostToDevice) != hipSuccess;
        *kp[i] = d;
        *wp[i] = d + n;
        *lp[i] = d + 2 * n;
        double* q64 = nullptr;
        float* q32 = nullptr;
        fail |= hipMalloc(reinterpret_cast<void**>(&q64), 8 * nq) != hipSuccess;
        if (fail) break;
        g->mem[3 * i + 1] = q64;
        fail |= hipMalloc(reinterpret_cast<void**>(&q32), 4 * nq) != hipSuccess;
        if (fail) break;
        g->mem[3 * i + 2] = q32;
        std::vector<float> qf(axes[i].q.begin(), axes[i].q.end());
        fail |= hipMemcpy(q64, axes[i].q.data(), 8 * nq, hipMemcpyHostToDevice) != hipSuccess;
        fail |= hipMemcpy(q32, qf.data(), 4 * nq, hipMemcpyHostToDevice) != hipSuccess;
        g->q64[i] = q64;
        g->q32[i] = q32;
    }
    if (fail) {
        (void)hipGetLastError();  // not left for a later launch check to report
        for (void* m : g->mem)
            if (m) (void)hipFree(m);
        delete g;
        set_error("heat: device allocation of solver tables failed");
        *rc = SHG_ENOMEM;
        return nullptr;
    }
    h->heat_cache[key] = g;
    return g;
}

struct LtPlan;
static void lt_plan_free(LtPlan* p);

void heat_cache_free(shg_handle* h) {
    for (auto& kv : h->lt_plans) lt_plan_free(static_cast<LtPlan*>(kv.second));
    h->lt_plans.clear();
    if (h->blaslt) hipblasLtDestroy((hipblasLtHandle_t)h->blaslt);
    h->blaslt = nullptr;
    for (auto& kv : h->heat_cache) {
        for (void* m : kv.second->mem)
            if (m) (void)hipFree(m);
        delete kv.second;
    }
    h->heat_cache.clear();
}

static size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

size_t heat_bytes(const shg_params* p) {
    const size_t nzh = p->nz_heat > 0 ? p->nz_heat : p->nz;
    const size_t n3 = (size_t)p->nx * p->ny * nzh;
    const HeatCoef c = heat_coef(p, (int)nzh);
    const size_t nu = (size_t)c.nxu * c.nyu * nzh;
    const size_t rs = (p->precision == SHG_PREC_FP64_PLAIN || p->precision == SHG_PREC_FP64_FUSED) ? 8 : 4;
    size_t b = 4 * al256(8 * n3) + al256(8 * n3) + al256(4 * n3);  // T, Tn, dk, dkbase; nn, Q
    b += 2 * al256(rs * nu);                           // u, v
    const size_t nb = (size_t)((p->nx + 63) / 64) * ((p->ny + 3) / 4) * nzh;
    b += al256(8 * nb) + al256(8 * 8);                 // partials, slots
    b += al256(8 * (size_t)p->nz);                     // poling
    b += 4096;
    return b;
}

static rocblas_handle blas(shg_handle* h) {
    if (!h->blas) {
        rocblas_handle bh = nullptr;
        if (rocblas_create_handle(&bh) != rocblas_status_success) return nullptr;
        rocblas_set_stream(bh, h->stream);
        h->blas = bh;
    }
    return (rocblas_handle)h->blas;
}

int heat_alloc(shg_handle* h, const shg_params* p, HeatState& hs) {
    hs.nx = p->nx;
    hs.ny = p->ny;
    hs.nzh = p->nz_heat > 0 ? p->nz_heat : p->nz;
    hs.dx = p->lx_um / (p->nx - 1);
    hs.dy = p->ly_um / (p->ny - 1);
    hs.dz = p->lz_um / (hs.nzh - 1);
    const size_t n3 = (size_t)hs.nx * hs.ny * hs.nzh;
    const HeatCoef c = heat_coef(p, hs.nzh);
    const size_t nu = (size_t)c.nxu * c.nyu * hs.nzh;
    const bool f64 = p->precision == SHG_PREC_FP64_PLAIN || p->precision == SHG_PREC_FP64_FUSED;
    Arena& A = h->arena;
    hs.T = A.carve<double>(n3);
    hs.Tn = A.carve<double>(n3);
    hs.dk = A.carve<double>(n3);
    hs.dkbase = A.carve<double>(n3);
    hs.nn = A.carve<float2>(n3);
    hs.Q = A.carve<float>(n3);
    hs.nblocks = heat_grid_blocks(c);
    hs.part = A.carve<double>(hs.nblocks);
    hs.red = A.carve<double>(8);
    hs.pol = nullptr;
    if (p->poling_phase) {
        hs.pol = A.carve<double>(p->nz);
        if (hs.pol)
            SHG_HIP(hipMemcpyAsync(hs.pol, p->poling_phase, sizeof(double) * p->nz, hipMemcpyHostToDevice, h->stream));
    }
    DirectTables* dt = new DirectTables();
    hs.solver = dt;
    dt->fp64 = f64;
    dt->u = f64 ? (void*)A.carve<double>(nu) : (void*)A.carve<float>(nu);
    dt->v = f64 ? (void*)A.carve<double>(nu) : (void*)A.carve<float>(nu);
    if (!hs.T || !hs.Q || !hs.red || !dt->u || !dt->v) {
        set_error("arena exhausted (heat buffers)");
        return SHG_ENOMEM;
    }
    int grc = SHG_OK;
    HeatGeom* g = heat_geom(h, p, hs, &grc);
    if (!g) return grc;
    dt->ax = g->ax;
    for (int i = 0; i < 3; i++) {
        dt->fold[i] = g->fold[i];
        if (f64) {
            if (!g->q64[i]) return SHG_ENOMEM;
            dt->q[i] = g->q64[i];
        } else {
            if (!g->q32[i]) return SHG_ENOMEM;
            dt->q[i] = g->q32[i];
        }
    }
    return SHG_OK;
}

void heat_release(HeatState& hs) {
    delete static_cast<DirectTables*>(hs.solver);
    hs.solver = nullptr;
}

int heat_init(shg_handle* h, const shg_params* p, HeatState& hs) {
    ProfScope ps(h, "heat_init");
    SHG_HIP(launch_heat_init(h->stream, hs.T, hs.Q, hs.nx, hs.ny, hs.nzh, p->T_C, p->T_peltier_C,
                             p->heat_bc == SHG_HEAT_BC_OVEN_SURROUNDED));
    return SHG_OK;
}

static int read_scalar(shg_handle* h, const double* d, double* out) {
    SHG_HIP(hipMemcpyAsync(out, d, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    SHG_HIP(hipStreamSynchronize(h->stream));
    return SHG_OK;
}

// out[..i..] = sum_j M[j][i] in[..j..] along one axis of u[nz][nyu][nx]
// (M = Q forward, Q^T inverse; Q row-major, Q[j*n+i] = component j of vector
// i).  On a folded axis (x or z) the data are in the even/odd fold and Q is
// the even block followed by the odd one: two products of half the length,
// one batched call.
template <typename R>
static rocblas_status gemm_sb(rocblas_handle bh, rocblas_operation ta, rocblas_operation tb, int m, int n, int k,
                              const R* a, int lda, rocblas_stride sa, const R* b, int ldb, rocblas_stride sb, R* c,
                              int ldc, rocblas_stride sc, int batch) {
    const R one = 1, zero = 0;
    if (batch == 1) {
        if constexpr (sizeof(R) == 4)
            return rocblas_sgemm(bh, ta, tb, m, n, k, &one, a, lda, b, ldb, &zero, c, ldc);
        else
            return rocblas_dgemm(bh, ta, tb, m, n, k, &one, a, lda, b, ldb, &zero, c, ldc);
    }
    if constexpr (sizeof(R) == 4)
        return rocblas_sgemm_strided_batched(bh, ta, tb, m, n, k, &one, a, lda, sa, b, ldb, sb, &zero, c, ldc, sc,
                                             batch);
    else
        return rocblas_dgemm_strided_batched(bh, ta, tb, m, n, k, &one, a, lda, sa, b, ldb, sb, &zero, c, ldc, sc,
                                             batch);
}

// The z transform in FP32 goes through hipBLASLt: for its shape (M = nx nyu,
// N = K = nz or nz/2) the algorithm hipBLASLt's heuristic ranks first runs
// about twice as fast as rocBLAS's choice (365 against 714 us folded at
// 512 x 511 x 400, measured on MI355X); rocBLAS stays faster for the x and y
// transforms.  One descriptor set and algorithm per shape, cached on the
// handle; no workspace.  A shape hipBLASLt has no algorithm for stays on
// rocBLAS.
struct LtPlan {
    bool ok = false;
    hipblasLtMatmulDesc_t d = nullptr;
    hipblasLtMatrixLayout_t la = nullptr, lb = nullptr, lc = nullptr;
    hipblasLtMatmulAlgo_t algo{};
};
static void lt_plan_free(LtPlan* p) {
    if (!p) return;
    if (p->la) hipblasLtMatrixLayoutDestroy(p->la);
    if (p->lb) hipblasLtMatrixLayoutDestroy(p->lb);
    if (p->lc) hipblasLtMatrixLayoutDestroy(p->lc);
    if (p->d) hipblasLtMatmulDescDestroy(p->d);
    delete p;
}
// C[M x N] = A[M x K] op(B), column-major, `batch` strided products
static LtPlan* lt_plan(shg_handle* h, bool tb, int M, int N, int K, long sa, long sb, long sc, int batch) {
    char key[160];
    snprintf(key, sizeof key, "%d/%d/%d/%d/%ld/%ld/%ld/%d", (int)tb, M, N, K, sa, sb, sc, batch);
    auto it = h->lt_plans.find(key);
    if (it != h->lt_plans.end()) return static_cast<LtPlan*>(it->second);
    LtPlan* p = new LtPlan();
    h->lt_plans[key] = p;
    if (!h->blaslt) {
        hipblasLtHandle_t lt = nullptr;
        if (hipblasLtCreate(&lt) != HIPBLAS_STATUS_SUCCESS) return p;
        h->blaslt = lt;
    }
    const hipblasOperation_t opn = HIPBLAS_OP_N, opb = tb ? HIPBLAS_OP_T : HIPBLAS_OP_N;
    bool ok = hipblasLtMatmulDescCreate(&p->d, HIPBLAS_COMPUTE_32F, HIP_R_32F) == HIPBLAS_STATUS_SUCCESS &&
              hipblasLtMatmulDescSetAttribute(p->d, HIPBLASLT_MATMUL_DESC_TRANSA, &opn, sizeof opn) ==
                  HIPBLAS_STATUS_SUCCESS &&
              hipblasLtMatmulDescSetAttribute(p->d, HIPBLASLT_MATMUL_DESC_TRANSB, &opb, sizeof opb) ==
                  HIPBLAS_STATUS_SUCCESS &&
              hipblasLtMatrixLayoutCreate(&p->la, HIP_R_32F, M, K, M) == HIPBLAS_STATUS_SUCCESS &&
              hipblasLtMatrixLayoutCreate(&p->lb, HIP_R_32F, tb ? N : K, tb ? K : N, tb ? N : K) ==
                  HIPBLAS_STATUS_SUCCESS &&
              hipblasLtMatrixLayoutCreate(&p->lc, HIP_R_32F, M, N, M) == HIPBLAS_STATUS_SUCCESS;
    const int32_t bc = batch;
    const int64_t so[3] = {sa, sb, sc};
    hipblasLtMatrixLayout_t ls[3] = {p->la, p->lb, p->lc};
    for (int i = 0; i < 3 && ok; i++)
        ok = hipblasLtMatrixLayoutSetAttribute(ls[i], HIPBLASLT_MATRIX_LAYOUT_BATCH_COUNT, &bc, sizeof bc) ==
                 HIPBLAS_STATUS_SUCCESS &&
             hipblasLtMatrixLayoutSetAttribute(ls[i], HIPBLASLT_MATRIX_LAYOUT_STRIDED_BATCH_OFFSET, &so[i],
                                               sizeof so[i]) == HIPBLAS_STATUS_SUCCESS;
    if (ok) {
        hipblasLtMatmulPreference_t pref = nullptr;
        const uint64_t ws = 0;
        hipblasLtMatmulHeuristicResult_t res{};
        int n = 0;
        ok = hipblasLtMatmulPreferenceCreate(&pref) == HIPBLAS_STATUS_SUCCESS &&
             hipblasLtMatmulPreferenceSetAttribute(pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES, &ws, sizeof ws) ==
                 HIPBLAS_STATUS_SUCCESS &&
             hipblasLtMatmulAlgoGetHeuristic((hipblasLtHandle_t)h->blaslt, p->d, p->la, p->lb, p->lc, p->lc, pref, 1,
                                             &res, &n) == HIPBLAS_STATUS_SUCCESS &&
             n > 0;
        if (pref) hipblasLtMatmulPreferenceDestroy(pref);
        if (ok) p->algo = res.algo;
    }
    p->ok = ok;
    return p;
}

template <typename R>
static int axis_gemm(shg_handle* h, int axis, bool fwd, bool fold, const R* q, const R* in, R* out, int nx, int nyu,
                     int nz) {
    rocblas_handle bh = blas(h);
    if (!bh) {
        set_error("rocBLAS handle creation failed");
        return SHG_EDEVICE;
    }
    rocblas_status st;
    const rocblas_operation N_ = rocblas_operation_none, T_ = rocblas_operation_transpose;
    const rocblas_stride plane = (rocblas_stride)nx * nyu;
    if (axis == 0) {
        const int n = fold ? nx / 2 : nx;  // per block
        st = gemm_sb<R>(bh, fwd ? N_ : T_, N_, n, nyu * nz, n, q, n, (rocblas_stride)n * n, in, nx, n, out, nx, n,
                        fold ? 2 : 1);
    } else if (axis == 1) {
        st = gemm_sb<R>(bh, N_, fwd ? T_ : N_, nx, nyu, nyu, in, nx, plane, q, nyu, 0, out, nx, plane, nz);
    } else {
        const int n = fold ? nz / 2 : nz;
        if constexpr (sizeof(R) == 4) {
            LtPlan* lp = lt_plan(h, fwd, nx * nyu, n, n, plane * n, (long)n * n, plane * n, fold ? 2 : 1);
            if (lp->ok) {
                const float one = 1.0f, zero = 0.0f;
                const hipblasStatus_t ls =
                    hipblasLtMatmul((hipblasLtHandle_t)h->blaslt, lp->d, &one, in, lp->la, q, lp->lb, &zero, out,
                                    lp->lc, out, lp->lc, &lp->algo, nullptr, 0, h->stream);
                if (ls != HIPBLAS_STATUS_SUCCESS) {
                    set_error("hipBLASLt matmul failed (%d) on axis 2", (int)ls);
                    return SHG_EDEVICE;
                }
                return SHG_OK;
            }
        }
        st = gemm_sb<R>(bh, N_, fwd ? T_ : N_, nx * nyu, n, n, in, nx * nyu, plane * n, q, n, (rocblas_stride)n * n,
                        out, nx * nyu, plane * n, fold ? 2 : 1);
    }
    if (st != rocblas_status_success) {
        set_error("rocBLAS gemm failed (%d) on axis %d", (int)st, axis);
        return SHG_EDEVICE;
    }
    return SHG_OK;
}

// residual pass: u = W^-1/2 A_ii (F(T) - T); optionally returns ||F(T) - T||_2^2
template <typename R>
static int resid_pass(shg_handle* h, const HeatCoef& c, HeatState& hs, DirectTables& dt, double* r_sq) {
    {
        ProfScope ps(h, "heat_resid");
        SHG_HIP(launch_heat_resid<R>(h->stream, c, dt.ax, hs.T, hs.Q, static_cast<R*>(dt.u), hs.part));
    }
    if (!r_sq) return SHG_OK;
    SHG_HIP(launch_reduce(h->stream, hs.part, hs.nblocks, false, hs.red + 0));
    return read_scalar(h, hs.red + 0, r_sq);
}

// correction pass: T += W^-1/2 S^-1 u (u from the residual pass); returns max|update|
template <typename R>
static int correct_pass(shg_handle* h, const HeatCoef& c, HeatState& hs, DirectTables& dt, double* max_corr) {
    const int nx = c.nxu, nyu = c.nyu, nz = hs.nzh;
    const bool fx = dt.fold[0], fz = dt.fold[2];
    R* a = static_cast<R*>(dt.u);
    R* b = static_cast<R*>(dt.v);
    const R* qx = static_cast<const R*>(dt.q[0]);
    const R* qy = static_cast<const R*>(dt.q[1]);
    const R* qz = static_cast<const R*>(dt.q[2]);
    {
        ProfScope ps(h, "heat_fd_gemm");
        if (fx || fz) {
            SHG_HIP(launch_fd_fold<R>(h->stream, a, b, nx, nyu, nz, fx, fz));
            std::swap(a, b);
        }
        if (int rc = axis_gemm<R>(h, 0, true, fx, qx, a, b, nx, nyu, nz)) return rc;
        if (int rc = axis_gemm<R>(h, 1, true, false, qy, b, a, nx, nyu, nz)) return rc;
        if (int rc = axis_gemm<R>(h, 2, true, fz, qz, a, b, nx, nyu, nz)) return rc;
        SHG_HIP(launch_fd_divide<R>(h->stream, b, dt.ax, nx, nyu, nz));
        if (int rc = axis_gemm<R>(h, 2, false, fz, qz, b, a, nx, nyu, nz)) return rc;
        if (int rc = axis_gemm<R>(h, 1, false, false, qy, a, b, nx, nyu, nz)) return rc;
        if (int rc = axis_gemm<R>(h, 0, false, fx, qx, b, a, nx, nyu, nz)) return rc;
    }
    {
        ProfScope ps(h, "heat_update");
        int nb = 0;
        SHG_HIP(launch_fd_update<R>(h->stream, c, dt.ax, a, hs.T, hs.part, fx, fz, &nb));
        SHG_HIP(launch_reduce(h->stream, hs.part, nb, true, hs.red + 1));
    }
    return read_scalar(h, hs.red + 1, max_corr);
}

template <typename R>
static int direct_solve(shg_handle* h, const shg_params* p, const HeatCoef& c, HeatState& hs, DirectTables& dt,
                        HeatSolveInfo& info) {
    // the reference's first check: the change of one sweep from the current T
    // (its residual pass also feeds the first correction)
    double r0sq = 0;
    if (int rc = resid_pass<R>(h, c, hs, dt, &r0sq)) return rc;
    if (std::sqrt(r0sq) < p->jacobi_tol) {
        ProfScope ps(h, "heat_jacobi_sweep");
        SHG_HIP(launch_heat_sweep(h->stream, c, hs.T, hs.Tn, hs.Q, hs.part, false));
        std::swap(hs.T, hs.Tn);
        info.sweeps = 1;
        info.residual = std::sqrt(r0sq);
        info.converged_in_one = true;
        return SHG_OK;
    }
    // defect correction until the error left is below tol (each step
    // contracts the error by ~1e-6 with FP32 transforms, ~1e-14 with FP64).
    // The error after an update of size mc is about rho * mc; once this
    // state's solves have measured rho, the loop stops when ten times that
    // estimate is below tol, one correction before an update itself falls
    // below tol (the criterion used until rho is known).
    const double tol = p->heat_tol_K > 0 ? p->heat_tol_K : 1e-8;
    double mc = INFINITY, prev = INFINITY;
    int it = 0;
    for (;; it++) {
        if (it > 0)
            if (int rc = resid_pass<R>(h, c, hs, dt, nullptr)) return rc;
        if (int rc = correct_pass<R>(h, c, hs, dt, &mc)) return rc;
        if (!std::isfinite(mc)) {
            set_error("heat: non-finite correction");
            return SHG_ENAN;
        }
        if (it > 0 && prev > 0 && std::isfinite(prev)) {
            hs.rho = std::max(mc / prev, 1e-16);
            hs.rho_max = std::max(hs.rho_max, hs.rho);
        }
        prev = mc;
        if (mc < tol) break;
        const double est = 10.0 * hs.rho_max;
        if (hs.rho_max > 0 && est < 0.1 && est * mc < tol) break;
        if (it >= 40) {
            set_error("heat: defect correction did not reach %g K (last update %g K)", tol, mc);
            return SHG_ENOCONV;
        }
    }
    info.sweeps = 0;
    info.residual = mc;
    info.converged_in_one = false;
    info.refinements = it + 1;
    return SHG_OK;
}

int heat_solve(shg_handle* h, const shg_params* p, HeatState& hs, HeatSolveInfo& info) {
    const HeatCoef c = heat_coef(p, hs.nzh);
    info = HeatSolveInfo{};
    if (p->heat_solver == SHG_HEAT_JACOBI_REFERENCE) {
        // src/headers/Solver.h:332-356 (norm ||T_f - T_i||_2 instead of
        // sqrt(sum)/TSIZE against 5e-4/TSIZE: the same test)
        const long long chk = p->check_every > 0 ? p->check_every : 5000;
        double red = INFINITY, aux = NAN;
        long long counter = 0;
        while (red >= p->jacobi_tol) {
            const bool check = counter % chk == 0;
            {
                ProfScope ps(h, "heat_jacobi_sweep");
                SHG_HIP(launch_heat_sweep(h->stream, c, hs.T, hs.Tn, hs.Q, hs.part, check));
            }
            if (check) {
                SHG_HIP(launch_reduce(h->stream, hs.part, hs.nblocks, false, hs.red + 0));
                double s2 = 0;
                if (int rc = read_scalar(h, hs.red + 0, &s2)) return rc;
                if (counter > 0) aux = red;
                red = std::sqrt(s2);
                if (red == aux) red = 0.5 * p->jacobi_tol;
            }
            std::swap(hs.T, hs.Tn);
            counter++;
            if (p->max_sweeps > 0 && counter > p->max_sweeps) {
                set_error("Jacobi loop exceeded max_sweeps=%lld", (long long)p->max_sweeps);
                return SHG_ENOCONV;
            }
        }
        info.sweeps = counter;
        info.residual = red;
        info.converged_in_one = counter == 1;
        return SHG_OK;
    }
    DirectTables& dt = *static_cast<DirectTables*>(hs.solver);
    return dt.fp64 ? direct_solve<double>(h, p, c, hs, dt, info) : direct_solve<float>(h, p, c, hs, dt, info);
}

int heat_adjoint_solve(shg_handle* h, const shg_params* p, HeatState& hs, const double* mu, double* nu, double* w,
                       double* u64, double* v64, double* tp) {
    const HeatCoef c = heat_coef(p, hs.nzh);
    int grc = SHG_OK;
    HeatGeom* g = heat_geom(h, p, hs, &grc);
    if (!g) return grc;
    if (!g->q64[0] || !g->q64[1] || !g->q64[2]) {
        set_error("heat adjoint: FP64 eigenbases missing");
        return SHG_ENOMEM;
    }
    const int nx = c.nxu, nyu = c.nyu, nz = hs.nzh;
    const bool fx = g->fold[0], fz = g->fold[2];
    double* a = u64;
    double* b = v64;
    ProfScope ps(h, "heat_adjoint_solve");
    SHG_HIP(launch_adj_heat_in(h->stream, c, g->ax, mu, a));
    if (fx || fz) {
        SHG_HIP(launch_fd_fold<double>(h->stream, a, b, nx, nyu, nz, fx, fz));
        std::swap(a, b);
    }
    // nu = W^-1/2 Q S^-1 Q^T W^-1/2 mu: the exact inverse, in FP64 (the
    // operator is symmetric, so the transposed solve is this one)
    if (int rc = axis_gemm<double>(h, 0, true, fx, g->q64[0], a, b, nx, nyu, nz)) return rc;
    if (int rc = axis_gemm<double>(h, 1, true, false, g->q64[1], b, a, nx, nyu, nz)) return rc;
    if (int rc = axis_gemm<double>(h, 2, true, fz, g->q64[2], a, b, nx, nyu, nz)) return rc;
    SHG_HIP(launch_fd_divide<double>(h->stream, b, g->ax, nx, nyu, nz));
    if (int rc = axis_gemm<double>(h, 2, false, fz, g->q64[2], b, a, nx, nyu, nz)) return rc;
    if (int rc = axis_gemm<double>(h, 1, false, false, g->q64[1], a, b, nx, nyu, nz)) return rc;
    if (int rc = axis_gemm<double>(h, 0, false, fx, g->q64[0], b, a, nx, nyu, nz)) return rc;
    SHG_HIP(hipMemsetAsync(nu, 0, sizeof(double) * (size_t)c.nx * c.ny * nz, h->stream));
    int nb = 0;
    SHG_HIP(launch_fd_update<double>(h->stream, c, g->ax, a, nu, hs.part, fx, fz, &nb));
    SHG_HIP(launch_adj_heat_out(h->stream, c, g->ax, mu, nu, w, hs.part));
    SHG_HIP(launch_reduce(h->stream, hs.part, heat_grid_blocks(c), false, tp));
    return SHG_OK;
}

int heat_prepare_optics(shg_handle* h, const shg_params* p, HeatState& hs, double* Tc) {
    {
        ProfScope ps(h, "heat_dk");
        SHG_HIP(launch_heat_dk(h->stream, hs.T, hs.dk, hs.nn, hs.dkbase, hs.nx * hs.ny, hs.nzh, p->nz, p->crystal,
                               p->lambda_p_um, p->Lambda_um, p->therm_exp_a, p->therm_exp_b));
    }
    // temperature at the grid centre of heat slice 0 (src/headers/Efields.h:88,124)
    const size_t ic = (size_t)(hs.ny / 2) * hs.nx + hs.nx / 2;
    return read_scalar(h, hs.T + ic, Tc);
}

int heat_finish(shg_handle* h, const shg_params* p, HeatState& hs, shg_result* r) {
    const long n3 = (long)hs.nx * hs.ny * hs.nzh;
    SHG_HIP(launch_max(h->stream, hs.T, n3, hs.part, std::min(hs.nblocks, 1024)));
    SHG_HIP(launch_reduce(h->stream, hs.part, std::min(hs.nblocks, 1024), true, hs.red + 2));
    double mx = 0;
    if (int rc = read_scalar(h, hs.red + 2, &mx)) return rc;
    r->T_max_C = mx;
    r->peak_dT_K = mx - p->T_peltier_C;
    if (r->temperature_out)
        SHG_HIP(hipMemcpy(r->temperature_out, hs.T, sizeof(double) * n3, hipMemcpyDeviceToHost));
    return SHG_OK;
}

}  // namespace shg
