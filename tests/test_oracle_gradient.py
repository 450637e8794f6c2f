"""The oracle's cost gradients (the adjoint of the exit-wave model, one slice or three): equal
to PyNX's ML gradients at one slice (whose runs the pynx_near fixtures pin), and, for every
exit of the reference's three-slice fixtures and both costs, the directional derivative
Re Σ conj(G)·d against five-point central differences of the cost at 1e-7."""

import numpy as np
import pytest

from oracle.lsqml import AMPLITUDE, POISSON, cost_grads
from test_gpu_rpie import oracle, small_scan
from test_oracle_multislice import state as multislice_state

pytestmark = []


def _slope(f, h):
    return (8.0 * (f(h) - f(-h)) - (f(2.0 * h) - f(-2.0 * h))) / (12.0 * h)


def test_one_slice_equals_pynx_ml_gradient():
    d = small_scan(n=32, nprobe=2, npos=12)
    o = oracle(d)
    v, go, gp = cost_grads(o.optics, o.views_all, o.iobs, o.addr.nmode, o.obj, o.probe)
    want_o, want_p = o.grads(True, True)
    assert np.abs(go - want_o).max() <= 1e-12 * np.abs(want_o).max()
    assert np.abs(gp - want_p).max() <= 1e-12 * np.abs(want_p).max()
    n = o.nb_obs
    assert abs(v / n - o.llk().poisson) <= 1e-12 * abs(v / n)


@pytest.mark.parametrize("case", range(3))
@pytest.mark.parametrize("cost", [POISSON, AMPLITUDE])
def test_multislice_gradients_match_central_differences(case, cost):
    addr, objects, probe, cases = multislice_state()
    optics, iobs, _, _ = cases[case]
    views = addr.views()[:6 * addr.nmode]
    _, go, gp = cost_grads(optics, views, iobs, addr.nmode, objects, probe, cost)
    rng = np.random.default_rng(17 + case)
    for i in range(3):
        d = (rng.uniform(-0.5, 0.5, objects[i].shape)
             + 1j * rng.uniform(-0.5, 0.5, objects[i].shape)) * np.abs(objects[i]).max()

        def f(t, i=i, d=d):
            os = list(objects)
            os[i] = objects[i] + t * d
            return cost_grads(optics, views, iobs, addr.nmode, os, probe, cost)[0]

        want = _slope(f, 1e-4)
        got = float((np.conj(go[i]) * d).real.sum())
        assert abs(got - want) <= 1e-7 * abs(want), (i, got, want)
    d = (rng.uniform(-0.5, 0.5, probe.shape)
         + 1j * rng.uniform(-0.5, 0.5, probe.shape)) * np.abs(probe).max()
    want = _slope(lambda t: cost_grads(optics, views, iobs, addr.nmode, objects, probe + t * d,
                                       cost)[0], 1e-4)
    got = float((np.conj(gp) * d).real.sum())
    assert abs(got - want) <= 1e-7 * abs(want)
