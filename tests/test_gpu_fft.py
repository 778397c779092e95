"""The GPU FFT itself (shg_fft2d: the transform code of the fused passes,
the replacement for cufftExecC2C at src/headers/Solver.h:159-183) against
the committed numpy golden vectors (tests/golden/fft_vectors.npz), in FP64
and in the FP32 forms' arithmetic (FP32 points, FP64 twiddle products).

1D vectors are checked through the 2D transform: a field that is zero
except for one row (column) r0 = x transforms to X (the 1D DFT of x) in
every row (column), times exp(-2 pi i k r0 / N) of the other axis -- with
r0 = 0 exactly X everywhere.  Tolerances, relative to the largest output:
FP64 1e-13, FP32 3e-6 (FP32 rounding of the inputs and of each of the
log8 N passes' outputs).
"""
import os

import numpy as np
import pytest

from cushg_amd import abi, shg

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "fft_vectors.npz")
SIZES = [64, 128, 192, 256, 384, 512]
FORMS = [pytest.param(abi.PREC_FP64_FUSED, 1e-13, id="fp64"),
         pytest.param(abi.PREC_FP32_SPLIT, 3e-6, id="fp32")]


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("prec,tol", FORMS)
@pytest.mark.parametrize("shape", [(64, 64), (128, 192)])
def test_fft2d_golden(handle, gold, shape, prec, tol):
    ny, nx = shape
    x = gold[f"x2d_{ny}_{nx}"]
    X = handle.fft2d(x, prec)
    err = rel(X, gold[f"X2d_{ny}_{nx}"])
    print(f"fft2d {ny}x{nx} precision {prec}: rel err {err:.2e}", flush=True)
    assert err < tol


@pytest.mark.parametrize("prec,tol", FORMS)
@pytest.mark.parametrize("n", SIZES)
def test_row_and_column_transforms_golden(handle, gold, n, prec, tol):
    for inv, key in ((False, "X"), (True, "I")):
        for v in range(2):
            x, ref = gold[f"x{n}"][v], gold[f"{key}{n}"][v]
            rows = np.zeros((64, n), complex)
            rows[0] = x
            cols = np.zeros((n, 64), complex)
            cols[:, 0] = x
            R = handle.fft2d(rows, prec, inverse=inv)
            Cc = handle.fft2d(cols, prec, inverse=inv)
            er = max(rel(R[k], ref) for k in range(64))
            ec = max(rel(Cc[:, k], ref) for k in range(64))
            assert er < tol, (n, inv, v, er)
            assert ec < tol, (n, inv, v, ec)


@pytest.mark.parametrize("prec,tol", FORMS)
def test_shifted_impulse_row(handle, gold, prec, tol):
    """A row r0 != 0 carries the phase exp(-2 pi i ky r0 / ny) of the column
    transform (catches a row/column or sign swap)."""
    n, ny, r0 = 256, 128, 37
    x = gold[f"x{n}"][0]
    a = np.zeros((ny, n), complex)
    a[r0] = x
    X = handle.fft2d(a, prec)
    ky = np.arange(ny)
    ref = np.exp(-2j * np.pi * ky * r0 / ny)[:, None] * gold[f"X{n}"][0][None, :]
    assert rel(X, ref) < tol


@pytest.mark.parametrize("prec,tol", FORMS)
@pytest.mark.parametrize("shape", [(512, 512), (384, 192), (64, 512)])
def test_round_trip_and_parseval(handle, shape, prec, tol):
    ny, nx = shape
    rng = np.random.default_rng(nx + ny)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    X = handle.fft2d(x, prec)
    # Parseval: sum |X|^2 = N sum |x|^2
    p = np.sum(np.abs(X) ** 2) / (nx * ny * np.sum(np.abs(x) ** 2))
    assert abs(p - 1) < 10 * tol
    y = handle.fft2d(X, prec, inverse=True) / (nx * ny)
    assert rel(y, x) < 2 * tol
    # and against numpy's transform of the same input
    assert rel(X, np.fft.fft2(x)) < tol


def test_fft2d_rejects_unsupported_sizes(handle):
    with pytest.raises(shg.ShgError) as e:
        handle.fft2d(np.zeros((64, 100), complex))
    assert e.value.status == abi.SHG_EINVAL
    with pytest.raises(shg.ShgError) as e:
        handle.fft2d(np.zeros((64, 64), complex), precision=99)
    assert e.value.status == abi.SHG_EINVAL
