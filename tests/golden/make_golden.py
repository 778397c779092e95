"""Writes the golden fixtures under tests/golden/ (run in the build container).

The reference ships no fixtures, known-answer tests or golden vectors for
this path (its repository has no tests at all), so every fixture here comes
from an independent source:
  fft_vectors.npz   numpy.fft (pocketfft) outputs for seeded complex inputs
                    of every GPU transform length, 1D and 2D.
  heat_fv.npz       steady states of the open-top and oven-surrounded heat problems assembled
                    independently as a finite-volume sparse matrix (face
                    conductances k*A/d, convective faces h*A, Dirichlet oven planes)
                    and solved with scipy.sparse.linalg.spsolve.
  sellmeier.json    n(lambda, T) and dk(T) evaluated with Python's decimal
                    module at 40 significant digits from the published
                    coefficients (Bruner et al. for MgO:sPPLT, as in
                    src/headers/Crystal.h:54-70).
  sellmeier_ppln.json
                    the same for 5 mol% MgO-doped congruent LiNbO3
                    (extraordinary ray): Gayer, Sacks, Galun, Arie,
                    Appl. Phys. B 91, 343-348 (2008), Table 1, typed here
                    from the paper; the grating expansion is Jundt's
                    (Opt. Lett. 22, 1553 (1997)).  The reference has no
                    PPLN entry (its README.md:52 names the crystal).
Usage: python tests/golden/make_golden.py [fft|heat|sellmeier|ppln ...]
"""
from __future__ import annotations

import json
import os
import sys
from decimal import Decimal, getcontext

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))


def fft_vectors():
    rng = np.random.default_rng(20260101)
    out = {}
    for n in (64, 128, 192, 256, 384, 512):
        x = (rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))).astype(np.complex128)
        out[f"x{n}"] = x
        out[f"X{n}"] = np.fft.fft(x, axis=-1)
        out[f"I{n}"] = np.fft.ifft(x, axis=-1) * n  # unnormalised inverse
    for (ny, nx) in ((64, 64), (128, 192)):
        x = (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))).astype(np.complex128)
        out[f"x2d_{ny}_{nx}"] = x
        out[f"X2d_{ny}_{nx}"] = np.fft.fft2(x)
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **out)


def fv_heat_solution(nx, ny, nz, lx, ly, lz, k, h, T_inf, T_pel, Q, bc="open"):
    """Finite-volume steady heat equation on the node grid of the reference
    (dx = lx/(nx-1) ...), each node owning the cell [x-dx/2, x+dx/2] clipped
    to the crystal, Q applied to interior nodes only.
      bc="open": faces x=0, x=lx, y=ly, z=0, z=lz convective (h), the y = 0
                 plane held at T_pel (the reference's open-top oven);
      bc="oven": the four lateral planes x=0, x=lx, y=0, y=ly held at T_pel,
                 faces z=0, z=lz convective (the oven-surrounded variant)."""
    dx, dy, dz = lx / (nx - 1), ly / (ny - 1), lz / (nz - 1)

    def width(i, n, d):
        return d / 2 if i in (0, n - 1) else d

    def fixed(x, y):
        if bc == "oven":
            return x in (0, nx - 1) or y in (0, ny - 1)
        return y == 0

    idx = -np.ones((nz, ny, nx), dtype=np.int64)
    cnt = 0
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                if not fixed(x, y):
                    idx[z, y, x] = cnt
                    cnt += 1
    rows, cols, vals = [], [], []
    b = np.zeros(cnt)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                if fixed(x, y):
                    continue
                i = idx[z, y, x]
                wx, wy, wz = width(x, nx, dx), width(y, ny, dy), width(z, nz, dz)
                diag = 0.0
                # conductive links
                for (xx, yy, zz, area, d) in (
                    (x - 1, y, z, wy * wz, dx), (x + 1, y, z, wy * wz, dx),
                    (x, y - 1, z, wx * wz, dy), (x, y + 1, z, wx * wz, dy),
                    (x, y, z - 1, wx * wy, dz), (x, y, z + 1, wx * wy, dz),
                ):
                    if not (0 <= xx < nx and 0 <= yy < ny and 0 <= zz < nz):
                        continue
                    g = k * area / d
                    diag += g
                    if fixed(xx, yy):
                        b[i] += g * T_pel
                    else:
                        rows.append(i)
                        cols.append(idx[zz, yy, xx])
                        vals.append(-g)
                # convective faces
                conv = 0.0
                if bc == "open":
                    if x == 0 or x == nx - 1:
                        conv += h * wy * wz
                    if y == ny - 1:
                        conv += h * wx * wz
                if z == 0 or z == nz - 1:
                    conv += h * wx * wy
                diag += conv
                b[i] += conv * T_inf
                interior = 0 < x < nx - 1 and 0 < y < ny - 1 and 0 < z < nz - 1
                if interior:
                    b[i] += Q[z, y, x] * wx * wy * wz
                rows.append(i)
                cols.append(i)
                vals.append(diag)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(cnt, cnt))
    t = spla.spsolve(A.tocsc(), b)
    T = np.full((nz, ny, nx), T_pel)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                if not fixed(x, y):
                    T[z, y, x] = t[idx[z, y, x]]
    return T


def heat_fv():
    rng = np.random.default_rng(7)
    cases = {}
    specs = [
        # nx, ny, nz, lx, ly, lz, k, h, T_inf, T_pel, Qscale
        (8, 8, 6, 800.0, 800.0, 30000.0, 8.4e-6, 10e-12, 25.0, 40.0, 1e-9),
        (6, 5, 4, 2000.0, 1000.0, 30000.0, 8.4e-6, 10e-12, 25.0, 55.66, 5e-9),
        (10, 7, 5, 800.0, 600.0, 1000.0, 4.6e-6, 5e-8, 20.0, 35.0, 2e-10),
    ]
    for ci, (nx, ny, nz, lx, ly, lz, k, h, Ti, Tp, qs) in enumerate(specs):
        Q = rng.random((nz, ny, nx)) * qs
        T = fv_heat_solution(nx, ny, nz, lx, ly, lz, k, h, Ti, Tp, Q)
        cases[f"spec{ci}"] = np.array([nx, ny, nz, lx, ly, lz, k, h, Ti, Tp], dtype=np.float64)
        cases[f"Q{ci}"] = Q
        cases[f"T{ci}"] = T
    # the oven-surrounded variant on the same three geometries (own sources)
    for ci, (nx, ny, nz, lx, ly, lz, k, h, Ti, Tp, qs) in enumerate(specs):
        Q = rng.random((nz, ny, nx)) * qs
        T = fv_heat_solution(nx, ny, nz, lx, ly, lz, k, h, Ti, Tp, Q, bc="oven")
        cases[f"ospec{ci}"] = np.array([nx, ny, nz, lx, ly, lz, k, h, Ti, Tp], dtype=np.float64)
        cases[f"oQ{ci}"] = Q
        cases[f"oT{ci}"] = T
    np.savez_compressed(os.path.join(HERE, "heat_fv.npz"), **cases)


def sellmeier():
    getcontext().prec = 40
    D = Decimal

    def n_spplt(L, T):
        A, B, Cc, Dd = D("4.502483"), D("0.007294"), D("0.185087"), D("-0.02357")
        E, F, G, H = D("0.073423"), D("0.199595"), D("0.001"), D("7.99724")
        tk = D(T) + D("273.15")
        b = D("3.483933e-8") * tk * tk
        c = D("1.607839e-8") * tk * tk
        L = D(L)
        L2 = L * L
        return (A + (B + b) / (L2 - (Cc + c) ** 2) + E / (L2 - F * F) + G / (L2 - H * H) + Dd * L2).sqrt()

    pi = D("3.141592653589793238462643383279502884197")

    def dk(T, Lambda="7.97", a="2.2e-6", b="-5.9e-9"):
        lp, ls = D("1.064"), D("0.532")
        d = D(T) - D(25)
        return 2 * pi * (2 * n_spplt(lp, T) / lp - n_spplt(ls, T) / ls
                         + 1 / (D(Lambda) * (1 + D(a) * d + D(b) * d * d)))

    rows = []
    for T in ("25", "40", "55.66", "60"):
        rows.append({"T": float(T), "n_p": float(n_spplt("1.064", T)), "n_s": float(n_spplt("0.532", T)),
                     "dk": float(dk(T))})
    with open(os.path.join(HERE, "sellmeier.json"), "w") as f:
        json.dump({"crystal": "MgO:sPPLT", "Lambda_um": 7.97, "rows": rows}, f, indent=1)


def sellmeier_ppln():
    getcontext().prec = 40
    D = Decimal
    # Gayer et al. 2008, Table 1, 5% MgO:CLN, n_e:
    #   n_e^2 = a1 + b1 f + (a2 + b2 f)/(lambda^2 - (a3 + b3 f)^2)
    #           + (a4 + b4 f)/(lambda^2 - a5^2) - a6 lambda^2,
    #   f = (t - 24.5 C)(t + 570.82), lambda in um, t in C
    a = {1: D("5.756"), 2: D("0.0983"), 3: D("0.2020"), 4: D("189.32"), 5: D("12.52"), 6: D("1.32e-2")}
    b = {1: D("2.860e-6"), 2: D("4.700e-8"), 3: D("6.113e-8"), 4: D("1.516e-4")}

    def n_e(lam, t):
        t, lam = D(t), D(lam)
        f = (t - D("24.5")) * (t + D("570.82"))
        l2 = lam * lam
        pole_uv = a[3] + b[3] * f
        n2 = (a[1] + b[1] * f + (a[2] + b[2] * f) / (l2 - pole_uv * pole_uv)
              + (a[4] + b[4] * f) / (l2 - a[5] * a[5]) - a[6] * l2)
        return n2.sqrt()

    pi = D("3.141592653589793238462643383279502884197")
    # first-order grating of the crystal database entry and Jundt's thermal
    # expansion of LiNbO3: L(T) = L(25) (1 + 1.54e-5 dT + 5.3e-9 dT^2)
    Lam, ea, eb = D("6.97"), D("1.54e-5"), D("5.3e-9")

    def dk(t):
        lp, ls = D("1.064"), D("0.532")
        d = D(t) - D(25)
        return 2 * pi * (2 * n_e(lp, t) / lp - n_e(ls, t) / ls + 1 / (Lam * (1 + ea * d + eb * d * d)))

    rows = []
    for t in ("20", "25", "40", "60", "100", "150"):
        rows.append({"T": float(t), "n_p": float(n_e("1.064", t)), "n_s": float(n_e("0.532", t)),
                     "dk": float(dk(t))})
    # other wavelengths of the fit's range (0.5-4 um) at 25 and 80 C
    extra = []
    for lam in ("0.5", "0.6328", "0.8", "1.55", "2.1", "3.0"):
        for t in ("25", "80"):
            extra.append({"lambda_um": float(lam), "T": float(t), "n": float(n_e(lam, t))})
    with open(os.path.join(HERE, "sellmeier_ppln.json"), "w") as fh:
        json.dump({"crystal": "MgO:PPLN (5% MgO:CLN, Gayer et al. 2008)", "Lambda_um": float(Lam),
                   "therm_exp_a": float(ea), "therm_exp_b": float(eb), "rows": rows, "index": extra},
                  fh, indent=1)


if __name__ == "__main__":
    parts = sys.argv[1:] or ["fft", "heat", "sellmeier", "ppln"]
    if "fft" in parts:
        fft_vectors()
    if "heat" in parts:
        heat_fv()
    if "sellmeier" in parts:
        sellmeier()
    if "ppln" in parts:
        sellmeier_ppln()
    print("fixtures written to", HERE)
