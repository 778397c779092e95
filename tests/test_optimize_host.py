"""Host logic of the eta-maximisation step (cushg_amd.optimize), on the CPU
with stand-in handles: every instance's gradient is taken exactly once
whatever the number of handles, each handle is used by one thread at a
time, results keep the instances' order, and a failing instance's error
reaches the caller."""
import threading
import time

import numpy as np
import pytest

from cushg_amd import abi, optimize


class FakeResult:
    def __init__(self, eta):
        self.eta = eta


class FakeHandle:
    def __init__(self, fail_on=None):
        self.fail_on = fail_on
        self.calls = []
        self.busy = threading.Lock()

    def adjoint(self, p, checkpoint_every=0):
        assert self.busy.acquire(blocking=False), "a handle was used by two threads at once"
        try:
            time.sleep(0.002)
            if self.fail_on is not None and p.waist_um == self.fail_on:
                raise RuntimeError("instance failed")
            self.calls.append(p.waist_um)
            g = np.zeros(3 + p.nz)
            g[:3] = (p.waist_um, 2 * p.waist_um, 3 * p.waist_um)
            return FakeResult(p.waist_um * 1e-3), g
        finally:
            self.busy.release()


def params(n):
    out = []
    for i in range(n):
        p = abi.default_params()
        p.waist_um = 10.0 + i
        out.append(p)
    return out


@pytest.mark.parametrize("nh", [1, 3])
def test_gradients_once_each_in_order(nh):
    hs = [FakeHandle() for _ in range(nh)]
    ps = params(11)
    out = optimize.gradients(hs, ps)
    assert sorted(w for h in hs for w in h.calls) == [p.waist_um for p in ps]
    for p, (eta, g) in zip(ps, out):
        assert eta == p.waist_um * 1e-3
        assert np.array_equal(g, [p.waist_um, 2 * p.waist_um, 3 * p.waist_um])


def test_gradients_error_reaches_caller():
    hs = [FakeHandle(fail_on=14.0), FakeHandle(fail_on=14.0)]
    with pytest.raises(RuntimeError, match="instance failed"):
        optimize.gradients(hs, params(8))
