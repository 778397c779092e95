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
    const double* q[3] = {nullptr, nullptr, nullptr};   // FP64 eigenvector matrices
    const double* qt[3] = {nullptr, nullptr, nullptr};  // and their (per-block) transposes
    bool fold[3] = {false, false, false};           // axis transformed in its even/odd fold
    void* u = nullptr;
    void* v = nullptr;
    bool fp64 = false;
    // FP32 transforms on the matrix cores: split eigenbases (per axis: hi
    // and lo of the forward operator, then of the inverse, nb blocks of
    // rpad x kpad halves each), their scale exponents, and the max|u| slots
    const _Float16* qs[3] = {nullptr, nullptr, nullptr};
    int qe[3] = {}, rpad[3] = {}, kpad[3] = {};
    unsigned* amax = nullptr;
    unsigned* apart = nullptr;  // per-workgroup maxima of those reductions
};

// Per-geometry tables of the direct solver (axis widths, stiffness
// diagonals, eigenpairs in FP64 and FP32), built once per geometry and kept
// on the handle: the host eigen-decompositions are O(n^3).
struct HeatGeom {
    HeatAxes ax{};
    bool fold[3] = {};
    void* mem[12] = {};
    double* q64[3] = {};
    double* q64t[3] = {};
    _Float16* qs[3] = {};
    int qe[3] = {}, rpad[3] = {}, kpad[3] = {};
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
        const int nb = axes[i].fold ? 2 : 1, nblk = n / nb;
        double* q64 = nullptr;
        double* q64t = nullptr;
        fail |= hipMalloc(reinterpret_cast<void**>(&q64), 8 * nq) != hipSuccess;
        if (fail) break;
        g->mem[3 * i + 1] = q64;
        fail |= hipMalloc(reinterpret_cast<void**>(&q64t), 8 * nq) != hipSuccess;
        if (fail) break;
        g->mem[3 * i + 2] = q64t;
        std::vector<double> qt(nq);
        for (int bl = 0; bl < nb; bl++)
            for (int r = 0; r < nblk; r++)
                for (int cc = 0; cc < nblk; cc++)
                    qt[(size_t)bl * nblk * nblk + (size_t)r * nblk + cc] =
                        axes[i].q[(size_t)bl * nblk * nblk + (size_t)cc * nblk + r];
        fail |= hipMemcpy(q64, axes[i].q.data(), 8 * nq, hipMemcpyHostToDevice) != hipSuccess;
        fail |= hipMemcpy(q64t, qt.data(), 8 * nq, hipMemcpyHostToDevice) != hipSuccess;
        g->q64[i] = q64;
        g->q64t[i] = q64t;
        // the split operators of the FP32 (matrix-core FP16) transforms
        const int rpad = (nblk + 127) / 128 * 128, kpad = (nblk + 31) / 32 * 32;
        double qm = 0;
        for (double v : axes[i].q) qm = std::max(qm, std::fabs(v));
        int e = 0;
        (void)std::frexp(qm > 0 ? qm : 1.0, &e);
        const size_t blk = (size_t)nb * rpad * kpad;
        _Float16* qs = nullptr;
        fail |= hipMalloc(reinterpret_cast<void**>(&qs), 2 * 4 * blk) != hipSuccess;
        if (fail) break;
        g->mem[9 + i] = qs;
        g->qs[i] = qs;
        g->qe[i] = 14 - e;
        g->rpad[i] = rpad;
        g->kpad[i] = kpad;
        fail |= launch_split_basis(nullptr, q64, nblk, nb, false, g->qe[i], qs, qs + blk, rpad, kpad) != hipSuccess;
        fail |= launch_split_basis(nullptr, q64, nblk, nb, true, g->qe[i], qs + 2 * blk, qs + 3 * blk, rpad, kpad) !=
                hipSuccess;
    }
    if (!fail) fail |= hipDeviceSynchronize() != hipSuccess;
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

void heat_cache_free(shg_handle* h) {
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
    b += al256(8 * nb) + al256(8 * 8) + al256(4 * 8);  // partials, slots, max|u| slots
    b += al256(4 * (size_t)heat_max_partials(c.nxu, c.nyu, (int)nzh));  // their per-workgroup maxima
    b += al256(8 * (size_t)p->nz);                     // poling
    b += 4096;
    return b;
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
    dt->amax = A.carve<unsigned>(8);
    dt->apart = A.carve<unsigned>((size_t)heat_max_partials(c.nxu, c.nyu, hs.nzh));
    if (!hs.T || !hs.Q || !hs.red || !dt->u || !dt->v || !dt->amax || !dt->apart) {
        set_error("arena exhausted (heat buffers)");
        return SHG_ENOMEM;
    }
    int grc = SHG_OK;
    HeatGeom* g = heat_geom(h, p, hs, &grc);
    if (!g) return grc;
    dt->ax = g->ax;
    for (int i = 0; i < 3; i++) {
        dt->fold[i] = g->fold[i];
        dt->qs[i] = g->qs[i];
        dt->qe[i] = g->qe[i];
        dt->rpad[i] = g->rpad[i];
        dt->kpad[i] = g->kpad[i];
        dt->q[i] = g->q64[i];
        dt->qt[i] = g->q64t[i];
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
// i, qt its transpose).  On a folded axis (x or z) the data are in the
// even/odd fold and Q is the even block followed by the odd one: two
// products of half the length, one batched launch.  FP64, as R S products
// of row-major operands (kernels_heat_gemm.hip):
//   x: out[row][i] = sum_k in[row][k] Q_op[i][k]   (S = q forward, qt inverse)
//   y, z: out[i][col] = sum_k Q_op[i][k] in[k][col] (R = qt forward, q inverse)
static int axis_gemm64(shg_handle* h, int axis, bool fwd, bool fold, const double* q, const double* qt,
                       const double* in, double* out, int nx, int nyu, int nz) {
    const long plane = (long)nx * nyu;
    HeatGemm64 g{};
    int batch = 1;
    if (axis == 0) {
        const int n = fold ? nx / 2 : nx;
        g.r = in;
        g.s = fwd ? q : qt;
        g.c = out;
        g.ldr = g.ldc = nx;
        g.lds = n;
        g.sr = g.sc = n;
        g.ss = (long)n * n;
        g.I = (long)nyu * nz;
        g.J = n;
        g.n = n;
        batch = fold ? 2 : 1;
    } else if (axis == 1) {
        g.r = fwd ? qt : q;
        g.s = in;
        g.c = out;
        g.ldr = nyu;
        g.lds = g.ldc = nx;
        g.sr = 0;
        g.ss = g.sc = plane;
        g.I = nyu;
        g.J = nx;
        g.n = nyu;
        batch = nz;
    } else {
        const int n = fold ? nz / 2 : nz;
        g.r = fwd ? qt : q;
        g.s = in;
        g.c = out;
        g.ldr = n;
        g.lds = g.ldc = plane;
        g.sr = (long)n * n;
        g.ss = g.sc = plane * n;
        g.I = n;
        g.J = plane;
        g.n = n;
        batch = fold ? 2 : 1;
    }
    SHG_HIP(launch_heat_gemm64(h->stream, g, batch));
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

// One FP32 eigenbasis transform along `axis` of u[nz][nyu][nx] on the
// matrix cores (split-FP16 products, kernels_heat_gemm.hip): the same
// products as axis_gemm, with max|in| read from umax and max|out| written
// to cmax (when given) for the next transform's scaling.
static int axis_mfma(shg_handle* h, int axis, bool fwd, const DirectTables& dt, const float* in, float* out, int nx,
                     int nyu, int nz, const unsigned* umax, unsigned* cmax) {
    const bool fold = dt.fold[axis];
    const int nb = fold ? 2 : 1;
    const long blk = (long)dt.rpad[axis] * dt.kpad[axis];
    const _Float16* base = dt.qs[axis] + (fwd ? 0 : 2) * nb * blk;
    const long plane = (long)nx * nyu;
    HeatGemm g{};
    g.u = in;
    g.c = out;
    g.qh = base;
    g.ql = base + nb * blk;
    g.umax = umax;
    g.cmax = cmax;
    g.cpart = dt.apart;
    g.qpad = dt.kpad[axis];
    g.qe = dt.qe[axis];
    int batch = nb;
    if (axis == 0) {  // C[row][m] = sum_k u[row][k] Q_op[m][k], rows = (z, y)
        g.kindB = 1;
        g.n = fold ? nx / 2 : nx;
        g.nbig = (long)nyu * nz;
        g.ldu = g.ldc = nx;
        g.su = g.sc = g.n;
        g.sq = blk;
    } else if (axis == 1) {  // per z plane: C[m][x] = sum_k Q_op[m][k] u[k][x]
        g.kindB = 0;
        g.n = nyu;
        g.nbig = nx;
        g.ldu = g.ldc = nx;
        g.su = g.sc = plane;
        g.sq = 0;
        batch = nz;
    } else {  // C[m][(y, x)] = sum_k Q_op[m][k] u[k][(y, x)]
        g.kindB = 0;
        g.n = fold ? nz / 2 : nz;
        g.nbig = plane;
        g.ldu = g.ldc = plane;
        g.su = g.sc = plane * g.n;
        g.sq = blk;
    }
    if (g.n > dt.rpad[axis] || g.n > dt.kpad[axis]) {
        set_error("heat: transform length %d exceeds its split eigenbasis (%d x %d)", g.n, dt.rpad[axis],
                  dt.kpad[axis]);
        return SHG_EDEVICE;
    }
    SHG_HIP(launch_heat_gemm(h->stream, g, batch));
    return SHG_OK;
}

// a <- S^-1 a in FP32 on the matrix cores (fold, the six transforms around
// the spectral division), leaving the result folded as launch_fd_update
// expects; b is scratch.  Max|.| slots: 0 the folded input, 1-2 after the x
// and y transforms, 3 after the division, 4-5 after the inverse z and y
// transforms.
static int inverse_fp32(shg_handle* h, const DirectTables& dt, float*& a, float*& b, int nx, int nyu, int nz) {
    const bool fx = dt.fold[0], fz = dt.fold[2];
    unsigned* m = dt.amax;
    SHG_HIP(launch_zero_u32(h->stream, m, 8));
    if (fx || fz) {
        SHG_HIP(launch_fd_fold<float>(h->stream, a, b, nx, nyu, nz, fx, fz, dt.apart, m + 0));
        std::swap(a, b);
    } else {
        SHG_HIP(launch_absmax(h->stream, a, (long)nx * nyu * nz, dt.apart, m + 0));
    }
    if (int rc = axis_mfma(h, 0, true, dt, a, b, nx, nyu, nz, m + 0, m + 1)) return rc;
    if (int rc = axis_mfma(h, 1, true, dt, b, a, nx, nyu, nz, m + 1, m + 2)) return rc;
    if (int rc = axis_mfma(h, 2, true, dt, a, b, nx, nyu, nz, m + 2, nullptr)) return rc;
    SHG_HIP(launch_fd_divide<float>(h->stream, b, dt.ax, nx, nyu, nz, dt.apart, m + 3));
    if (int rc = axis_mfma(h, 2, false, dt, b, a, nx, nyu, nz, m + 3, m + 4)) return rc;
    if (int rc = axis_mfma(h, 1, false, dt, a, b, nx, nyu, nz, m + 4, m + 5)) return rc;
    if (int rc = axis_mfma(h, 0, false, dt, b, a, nx, nyu, nz, m + 5, nullptr)) return rc;
    return SHG_OK;
}

// correction pass: T += W^-1/2 S^-1 u (u from the residual pass); returns max|update|
template <typename R>
static int correct_pass(shg_handle* h, const HeatCoef& c, HeatState& hs, DirectTables& dt, double* max_corr) {
    const int nx = c.nxu, nyu = c.nyu, nz = hs.nzh;
    const bool fx = dt.fold[0], fz = dt.fold[2];
    R* a = static_cast<R*>(dt.u);
    R* b = static_cast<R*>(dt.v);
    if constexpr (sizeof(R) == 4) {
        ProfScope ps(h, "heat_fd_gemm");
        if (int rc = inverse_fp32(h, dt, a, b, nx, nyu, nz)) return rc;
    } else {
        ProfScope ps(h, "heat_fd_gemm");
        if (fx || fz) {
            SHG_HIP(launch_fd_fold<R>(h->stream, a, b, nx, nyu, nz, fx, fz));
            std::swap(a, b);
        }
        if (int rc = axis_gemm64(h, 0, true, fx, dt.q[0], dt.qt[0], a, b, nx, nyu, nz)) return rc;
        if (int rc = axis_gemm64(h, 1, true, false, dt.q[1], dt.qt[1], b, a, nx, nyu, nz)) return rc;
        if (int rc = axis_gemm64(h, 2, true, fz, dt.q[2], dt.qt[2], a, b, nx, nyu, nz)) return rc;
        SHG_HIP(launch_fd_divide<R>(h->stream, b, dt.ax, nx, nyu, nz));
        if (int rc = axis_gemm64(h, 2, false, fz, dt.q[2], dt.qt[2], b, a, nx, nyu, nz)) return rc;
        if (int rc = axis_gemm64(h, 1, false, false, dt.q[1], dt.qt[1], a, b, nx, nyu, nz)) return rc;
        if (int rc = axis_gemm64(h, 0, false, fx, dt.q[0], dt.qt[0], b, a, nx, nyu, nz)) return rc;
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
    if (!g->q64[0] || !g->q64[1] || !g->q64[2] || !g->q64t[0] || !g->q64t[1] || !g->q64t[2]) {
        set_error("heat adjoint: FP64 eigenbases missing");
        return SHG_ENOMEM;
    }
    const int nx = c.nxu, nyu = c.nyu, nz = hs.nzh;
    const bool fx = g->fold[0], fz = g->fold[2];
    ProfScope ps(h, "heat_adjoint_solve");
    const DirectTables* dt = static_cast<const DirectTables*>(hs.solver);
    if (dt && !dt->fp64) {
        // FP32-field forms: one FP32 solve (relative error ~1e-6, below the
        // forms' own rounding of the gradient) on the matrix cores
        float* a = reinterpret_cast<float*>(u64);
        float* b = reinterpret_cast<float*>(v64);
        SHG_HIP(launch_adj_heat_in<float>(h->stream, c, g->ax, mu, a));
        if (int rc = inverse_fp32(h, *dt, a, b, nx, nyu, nz)) return rc;
        SHG_HIP(hipMemsetAsync(nu, 0, sizeof(double) * (size_t)c.nx * c.ny * nz, h->stream));
        int nb = 0;
        SHG_HIP(launch_fd_update<float>(h->stream, c, g->ax, a, nu, hs.part, fx, fz, &nb));
        SHG_HIP(launch_adj_heat_out(h->stream, c, g->ax, mu, nu, w, hs.part));
        SHG_HIP(launch_reduce(h->stream, hs.part, heat_grid_blocks(c), false, tp));
        return SHG_OK;
    }
    double* a = u64;
    double* b = v64;
    SHG_HIP(launch_adj_heat_in<double>(h->stream, c, g->ax, mu, a));
    if (fx || fz) {
        SHG_HIP(launch_fd_fold<double>(h->stream, a, b, nx, nyu, nz, fx, fz));
        std::swap(a, b);
    }
    // nu = W^-1/2 Q S^-1 Q^T W^-1/2 mu: the exact inverse, in FP64 (the
    // operator is symmetric, so the transposed solve is this one)
    if (int rc = axis_gemm64(h, 0, true, fx, g->q64[0], g->q64t[0], a, b, nx, nyu, nz)) return rc;
    if (int rc = axis_gemm64(h, 1, true, false, g->q64[1], g->q64t[1], b, a, nx, nyu, nz)) return rc;
    if (int rc = axis_gemm64(h, 2, true, fz, g->q64[2], g->q64t[2], a, b, nx, nyu, nz)) return rc;
    SHG_HIP(launch_fd_divide<double>(h->stream, b, g->ax, nx, nyu, nz));
    if (int rc = axis_gemm64(h, 2, false, fz, g->q64[2], g->q64t[2], b, a, nx, nyu, nz)) return rc;
    if (int rc = axis_gemm64(h, 1, false, false, g->q64[1], g->q64t[1], a, b, nx, nyu, nz)) return rc;
    if (int rc = axis_gemm64(h, 0, false, fx, g->q64[0], g->q64t[0], b, a, nx, nyu, nz)) return rc;
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
