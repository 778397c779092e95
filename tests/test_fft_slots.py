"""The LDS slot scheme of the column pass's transform pair (fft_fwd_op_inv
and fft_pow in cushg_amd/csrc/fft.hpp), restated on the CPU for every
supported length: within each pass every slot is written once, the passes
read exactly what the previous pass wrote, the barrier-free hand-overs are
thread-local (a middle-pass butterfly rewrites only the slots it read; the
inverse transform's first pass writes only the slots its thread read in the
forward transform's last pass), and the pair computes
IFFT(FFT(x) * prop) * N.  Replaces cufftExecC2C of the diffraction step
(src/headers/Solver.h:159-183)."""
import numpy as np
import pytest


def _dft(x, inv):
    n = len(x)
    k = np.arange(n)
    return np.exp((1 if inv else -1) * 2j * np.pi * np.outer(k, k) / n) @ x


@pytest.mark.parametrize("N", [64, 128, 192, 256, 384, 512])
def test_transform_pair_slots(N):
    tpr, m = N // 8, N // 64
    per = (64 + tpr - 1) // tpr
    rng = np.random.default_rng(N)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    prop = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    W = lambda e, inv: np.exp((1 if inv else -1) * 2j * np.pi * e / N)
    v = [np.array([x[tp + q * tpr] for q in range(8)]) for tp in range(tpr)]
    lds = np.full(N, np.nan + 0j)
    read3 = {}
    for inv in (False, True):
        # pass 1
        written = set()
        for tp in range(tpr):
            w, u = tp & 7, tp >> 3
            y = _dft(v[tp], inv)
            slots = [8 * q + w + 64 * u if inv else 8 * tp + q for q in range(8)]
            if inv:  # thread-local hand-over from the forward transform's last pass
                assert set(slots) == read3[tp]
            for q, s in enumerate(slots):
                assert 0 <= s < N and s not in written
                written.add(s)
                lds[s] = y[q]
        assert len(written) == N
        # middle pass, in place
        if m > 1:
            written = set()
            for tp in range(tpr):
                for i in range(per):
                    b = tp + i * tpr
                    if b >= 64:
                        continue
                    k = b & 7
                    s0 = 8 * k + (b >> 3) if inv else b
                    slots = [s0 + 64 * j for j in range(m)]
                    xs = np.array([lds[s] * (W(8 * j * k, inv) if j else 1) for j, s in enumerate(slots)])
                    assert not np.isnan(xs).any()
                    ys = _dft(xs, inv)
                    for q, s in enumerate(slots):
                        assert s not in written
                        written.add(s)
                        lds[s] = ys[q]
            assert len(written) == N
        # last pass
        seen = set()
        for tp in range(tpr):
            w, u = tp & 7, tp >> 3
            slots = [8 * w + j + 64 * u if inv else 8 * j + w + 64 * u for j in range(8)]
            assert not seen & set(slots)
            seen |= set(slots)
            read3[tp] = set(slots)
            a = np.array([lds[s] * (W(j * tp, inv) if j else 1) for j, s in enumerate(slots)])
            v[tp] = _dft(a, inv)
        assert len(seen) == N
        if not inv:
            got = np.array([v[i % tpr][i // tpr] for i in range(N)])
            assert np.allclose(got, np.fft.fft(x), atol=1e-9 * N)
            for tp in range(tpr):
                for q in range(8):
                    v[tp][q] *= prop[tp + q * tpr]
    got = np.array([v[i % tpr][i // tpr] for i in range(N)])
    assert np.allclose(got, np.fft.ifft(np.fft.fft(x) * prop) * N, atol=1e-8 * N)
