"""The C++ FP64 rPIE of the CPU baseline (oracle/cpu_rpie.cpp) against the numpy oracle's
`Ptycho.rpie` (pinned to the reference's rPIE fixtures at 1e-12, tests/test_oracle_rpie.py):
one batch and whole epochs, 1–4 modes, masked pixels, fixed arrays, and 1 or 4 threads giving
the same bits."""

import numpy as np
import pytest

from oracle import cpu
from oracle.ptycho import far_field_ptycho
from ptyrs_amd import sim


@pytest.fixture(scope="module", autouse=True)
def built():
    cpu.build()


def scan(n, nmode, npos, seed=3, masked=True):
    pixel = sim.pixel_size(n)
    y, x = sim.fermat_positions(npos, n / 8, pixel)
    ext = int(np.ceil(np.abs(np.r_[y - y.mean(), x - x.mean()] / pixel).max())) + 2
    side = (2 * ext + n + 8 + 3) // 4 * 4
    obj = sim.random_object(side, side, seed)
    probe = sim.hermite_gauss_probe(n, powers=[0.7, 0.15, 0.1, 0.05][:nmode])
    cy, cx = sim.corners(y, x, pixel, obj.shape[-2:], n)
    I = sim.intensities(obj, probe, cy, cx, sim.quad_phase(n))
    s = 1e5 / I.sum(axis=(1, 2)).mean()
    counts = np.random.default_rng(seed).poisson(I * s).astype(np.float32)
    if masked:
        counts[:, 5, :] = -1
        counts[:, :, n // 3] = -1
    start = np.exp(1j * np.random.default_rng(seed + 1).uniform(-1.0, 1.0, obj.shape))
    o = far_field_ptycho(np.where(counts < 0, -1.0, counts).astype(np.float64), y, x, start,
                         probe * np.sqrt(s) * 0.9, sim.WAVELENGTH, sim.DISTANCE,
                         sim.DETECTOR_PIXEL)
    c = cpu.CpuRpie(counts, np.c_[o.addr.oy, o.addr.ox], o.obj, o.probe, o.quad)
    return o, c


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n,nmode", [(32, 1), (64, 2), (32, 4)])
def test_batch_matches_oracle(n, nmode):
    o, c = scan(n, nmode, 24)
    o0, p0 = o.obj.copy(), o.probe.copy()
    frames = [1, 2, 5, 8, 13, 21]
    o.rpie_batch(frames, 0.9, 0.8, 0.15, 0.05)
    c.batch(frames, 0.9, 0.8, 0.15, 0.05)
    assert rel(c.obj - o0[0], o.obj[0] - o0[0]) < 1e-12
    assert rel(c.probe - p0, o.probe - p0) < 1e-12


def test_epochs_match_oracle_and_threads_agree():
    o, c = scan(64, 2, 40)
    p0 = o.probe.copy()
    c4 = cpu.CpuRpie(c.counts, c.corners, c.obj.copy(), c.probe.copy(), c.quad, threads=4)
    o.rpie(3, batch_size=16)
    c.rpie(3, batch_size=16)
    c4.rpie(3, batch_size=16)
    assert rel(c.obj, o.obj[0]) < 1e-11
    assert rel(c.probe - p0, o.probe - p0) < 1e-11
    assert np.array_equal(c.obj, c4.obj) and np.array_equal(c.probe, c4.probe)


def test_fixed_arrays():
    o, c = scan(32, 2, 12, masked=False)
    p0, ob0 = c.probe.copy(), c.obj.copy()
    c.batch(range(12), obj_step=None)
    o.rpie_batch(list(range(12)), None, 1.0, 0.1, 0.1)
    assert np.array_equal(c.obj, ob0)
    assert rel(c.probe - p0, o.probe - p0) < 1e-12
    c2 = cpu.CpuRpie(c.counts, c.corners, ob0, p0, c.quad)
    c2.batch(range(12), probe_step=None)
    assert np.array_equal(c2.probe, p0)


def test_refuses_bad_shapes():
    o, c = scan(32, 1, 6, masked=False)
    bad = cpu.CpuRpie(c.counts, c.corners + 10_000, c.obj, c.probe, c.quad)
    with pytest.raises(ValueError):
        bad.batch([0, 1])
