"""The engine's hand-written adjoint on the GPU (ptyrs_engine_grad) and the
torch.autograd.Function over it: the cost and its gradients against the oracle's (pinned by
PyNX's ML gradient and by central differences, tests/test_oracle_gradient.py) for one slice
in the far field and the angular spectrum and for the reference's three-slice fixtures; the
GPU cost's own central differences against its gradients (≤ 1e-5 relative, complex128), and
torch's backward through the Function."""

import numpy as np
import pytest

import ptyrs_amd as pa
from oracle.lsqml import AMPLITUDE, POISSON, cost_grads
from test_gpu_multislice import start as ms_start
from test_gpu_rpie import gpu, oracle, small_scan
from test_oracle_multislice import state as multislice_state

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-9, "c64": 1e-4}


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("cost", ["poisson", "amplitude"])
def test_far_field_gradient_matches_oracle(dtype, cost):
    d = small_scan(n=64, nprobe=2, npos=30)
    g, o = gpu(d, dtype), oracle(d)
    v, go, gp = g.grad(cost)
    wv, wo, wp = cost_grads(o.optics, o.views_all, o.iobs, o.addr.nmode, o.obj, o.probe,
                            POISSON if cost == "poisson" else AMPLITUDE)
    assert abs(v - wv) <= TOL[dtype] * abs(wv)
    assert rel(go[0], wo) < TOL[dtype] * 10
    assert rel(gp, wp) < TOL[dtype] * 10


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("case", range(3))
def test_multislice_gradient_matches_oracle(dtype, case):
    addr, objects, probe, cases = multislice_state()
    optics, iobs, _, _ = cases[case]
    g, _, _ = ms_start(case, dtype)
    v, go, gp = g.grad("poisson")
    wv, wo, wp = cost_grads(optics, addr.views(), iobs, addr.nmode, objects, probe)
    tol = {"c128": 1e-9, "c64": 1e-3}[dtype]
    assert abs(v - wv) <= tol * abs(wv)
    for s in range(3):
        assert rel(go[s], wo[s]) < tol * 10, s
    assert rel(gp, wp) < tol * 10


def _slope(f, h):
    return (8.0 * (f(h) - f(-h)) - (f(2.0 * h) - f(-2.0 * h))) / (12.0 * h)


@pytest.mark.parametrize("case", [0, 1])
def test_gpu_gradient_matches_its_central_differences(case):
    """Re Σ conj(G)·d against five-point central differences of the GPU cost along random
    directions of every slice and of the probe (complex128), 1e-5 relative."""
    g, objects, probe = ms_start(case, "c128")
    _, go, gp = g.grad("poisson")
    rng = np.random.default_rng(23)

    def cost(objs, pr):
        g.set_object(objs)
        g.set_probe(pr)
        return g.grad("poisson")[0]

    for s in range(3):
        d = np.zeros_like(objects)
        d[s] = (rng.uniform(-0.5, 0.5, objects[s].shape)
                + 1j * rng.uniform(-0.5, 0.5, objects[s].shape)) * np.abs(objects[s]).max()
        want = _slope(lambda t: cost(objects + t * d, probe), 1e-4)
        got = float((np.conj(go) * d).real.sum())
        assert abs(got - want) <= 1e-5 * abs(want), (s, got, want)
    d = (rng.uniform(-0.5, 0.5, probe.shape)
         + 1j * rng.uniform(-0.5, 0.5, probe.shape)) * np.abs(probe).max()
    want = _slope(lambda t: cost(objects, probe + t * d), 1e-4)
    got = float((np.conj(gp) * d).real.sum())
    assert abs(got - want) <= 1e-5 * abs(want)


def test_autograd_function_backward():
    torch = pytest.importorskip("torch")
    from ptyrs_amd.autograd import cost

    g, objects, probe = ms_start(0, "c128")
    v, go, gp = g.grad("poisson")
    ot = torch.tensor(objects, dtype=torch.complex128, requires_grad=True)
    pt = torch.tensor(probe, dtype=torch.complex128, requires_grad=True)
    loss = cost(g, ot, pt)
    assert abs(loss.item() - v) <= 1e-12 * abs(v)
    (3.0 * loss).backward()
    assert np.allclose(ot.grad.numpy(), 3.0 * go, rtol=0, atol=1e-12 * np.abs(go).max())
    assert np.allclose(pt.grad.numpy(), 3.0 * gp, rtol=0, atol=1e-12 * np.abs(gp).max())
    # A descent step along −grad lowers the cost (torch's convention: grad = 2·∂L/∂z*).
    with torch.no_grad():
        step = 1e-3 * abs(v) / float((ot.grad.abs() ** 2).sum())
        lower = cost(g, ot - step * ot.grad / 3.0, pt).item()
    assert lower < v


def test_gradient_refuses_positions():
    d = small_scan(n=32, nprobe=1, npos=6)
    g = gpu(d, "c64")
    g.set_features(position=np.zeros((6, 2)))
    with pytest.raises(ValueError):
        g.grad()
    with pytest.raises(ValueError):
        g.set_features()
        g.grad("gaussian")
