// adjoint_math.hpp -- pointwise reverse-mode pieces shared by the adjoint
// row passes (kernels_adjoint.hip, kernels_adjoint_thermal.hip): the
// transposes of dA/dz and of one RK4 call, FP64.
#pragma once

#include "mixed_passes.hpp"

namespace shg {

// Reverse of dA/dz (src/headers/Solver.h:57-80) for gradients GP, GS of
// dP, dS:  returns grad_P, grad_S and grad_e, and accumulates dL/dkp,
// dL/dks, dL/dbfac.
//   dP = i kp S P* e* - aph P        dS = i ks P^2 e - (ash + bfac |S|^2 / 2) S
__device__ __forceinline__ void dadz_adj(double2 P, double2 S, double2 e, double kp, double ks, double aph,
                                         double ash, double bfac, double2 GP, double2 GS, double2& gP,
                                         double2& gS, double2& ge, double& gkp, double& gks, double& gb) {
    const double2 ec = conjc(e), Pc = conjc(P);
    const double2 GPc = conjc(GP), GSc = conjc(GS);
    // grad_P = -aph GP + i kp S e* conj(GP) - 2 i ks P* e* GS
    const double2 t1 = cmul(cmul(S, ec), GPc);
    const double2 t2 = cmul(cmul(Pc, ec), GS);
    gP = mk<double2>(-aph * GP.x - kp * t1.y + 2.0 * ks * t2.y, -aph * GP.y + kp * t1.x - 2.0 * ks * t2.x);
    // grad_S = -i kp P e GP - (ash + bfac |S|^2) GS - bfac/2 S^2 conj(GS)
    const double2 t3 = cmul(cmul(P, e), GP);
    const double s2 = abs2(S);
    const double2 t4 = cmul(cmul(S, S), GSc);
    const double g = ash + bfac * s2;
    gS = mk<double2>(kp * t3.y - g * GS.x - 0.5 * bfac * t4.x, -kp * t3.x - g * GS.y - 0.5 * bfac * t4.y);
    // grad_e = i kp S P* conj(GP) - i ks P*^2 GS
    const double2 t5 = cmul(cmul(S, Pc), GPc);
    const double2 t6 = cmul(cmul(Pc, Pc), GS);
    ge = mk<double2>(-kp * t5.y + ks * t6.y, kp * t5.x - ks * t6.x);
    // dL/dkp = Re(conj(GP) i S P* e*), dL/dks = Re(conj(GS) i P^2 e), dL/dbfac = Re(conj(GS)(-|S|^2 S / 2))
    const double2 u = cmul(cmul(S, Pc), ec);
    gkp += GP.y * u.x - GP.x * u.y;
    const double2 w = cmul(cmul(P, P), e);
    gks += GS.y * w.x - GS.x * w.y;
    gb += -0.5 * s2 * (GS.x * S.x + GS.y * S.y);
}

// Reverse of one RK4 call (src/headers/Solver.h:231-257) at forward input
// (P, S): lam <- J^T lam.  acc[0..2] += dL/dphi_j of the three phase factors
// (e_j = exp(i phi_j): dL/dphi_j = Re(conj(grad_e_j) i e_j)), acc[3..5] +=
// dL/dkp, dL/dks, dL/dbfac.
__device__ __forceinline__ void rk4_adj(double2& lP, double2& lS, double2 P, double2 S, const double2 (&e)[3],
                                        double kp, double ks, double bfac, const RkConst<double>& k,
                                        double (&acc)[6]) {
    double2 k1p, k1s, k2p, k2s, k3p, k3s, k4p, k4s;
    dadz(P, S, e[0], kp, ks, k.aph, k.ash, bfac, k1p, k1s);
    const double2 u2p = axpy(k.c1, k1p, P), u2s = axpy(k.c1, k1s, S);
    dadz(u2p, u2s, e[1], kp, ks, k.aph, k.ash, bfac, k2p, k2s);
    const double2 u3p = axpy(k.c2, k2p, P), u3s = axpy(k.c2, k2s, S);
    dadz(u3p, u3s, e[1], kp, ks, k.aph, k.ash, bfac, k3p, k3s);
    const double2 u4p = axpy(k.c3, k3p, P), u4s = axpy(k.c3, k3s, S);
    (void)k4p;
    (void)k4s;
    double2 aP = lP, aS = lS;
    double2 a4p = k.w * lP, a4s = k.w * lS;
    double2 a3p = (2.0 * k.w) * lP, a3s = (2.0 * k.w) * lS;
    double2 a2p = a3p, a2s = a3s;
    double2 a1p = a4p, a1s = a4s;
    double2 gP, gS, ge;
    double2 ge1 = mk<double2>(0.0, 0.0);
    dadz_adj(u4p, u4s, e[2], kp, ks, k.aph, k.ash, bfac, a4p, a4s, gP, gS, ge, acc[3], acc[4], acc[5]);
    aP = aP + gP;
    aS = aS + gS;
    a3p = axpy(k.c3, gP, a3p);
    a3s = axpy(k.c3, gS, a3s);
    const double2 ge2 = ge;
    dadz_adj(u3p, u3s, e[1], kp, ks, k.aph, k.ash, bfac, a3p, a3s, gP, gS, ge, acc[3], acc[4], acc[5]);
    aP = aP + gP;
    aS = aS + gS;
    a2p = axpy(k.c2, gP, a2p);
    a2s = axpy(k.c2, gS, a2s);
    ge1 = ge1 + ge;
    dadz_adj(u2p, u2s, e[1], kp, ks, k.aph, k.ash, bfac, a2p, a2s, gP, gS, ge, acc[3], acc[4], acc[5]);
    aP = aP + gP;
    aS = aS + gS;
    a1p = axpy(k.c1, gP, a1p);
    a1s = axpy(k.c1, gS, a1s);
    ge1 = ge1 + ge;
    dadz_adj(P, S, e[0], kp, ks, k.aph, k.ash, bfac, a1p, a1s, gP, gS, ge, acc[3], acc[4], acc[5]);
    aP = aP + gP;
    aS = aS + gS;
    const double2 ge0 = ge;
    // dL/dphi = Re(conj(g) i e) = g.x * (-e.y) + g.y * e.x
    acc[0] += -ge0.x * e[0].y + ge0.y * e[0].x;
    acc[1] += -ge1.x * e[1].y + ge1.y * e[1].x;
    acc[2] += -ge2.x * e[2].y + ge2.y * e[2].x;
    lP = aP;
    lS = aS;
}

}  // namespace shg
