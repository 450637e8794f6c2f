"""The GPU's own 2D FFT (every frame side it is built for) through the far-field model
intensities Σ_m |U(P_m·O_win·q)|² of white-noise probes and a rough object, against the FP64
oracle's: complex64 to max|ΔI|/max I ≤ 2e-5 (twice the reference's FP32 amplitude bound
max|ΔΨ|/max|Ψ| ≤ 1e-5 of cuFFT against an FP64 DFT, ptyrs/tests/fft_accuracy.rs:61-84),
complex128 to 1e-12."""

import numpy as np
import pytest

import ptyrs_amd as pa
from oracle.base import to_natural
from oracle.ptycho import _mode_sum, far_field_ptycho
from ptyrs_amd import sim

pytestmark = pytest.mark.gpu


def noise_scan(n, nmode=2, npos=6, seed=17):
    rng = np.random.default_rng(seed)
    pixel = sim.pixel_size(n)
    y, x = sim.fermat_positions(npos, max(n / 8, 2.0), pixel)
    side = (n + 2 * int(np.ceil(np.abs(np.r_[y, x]).max() / pixel)) + 12 + 3) // 4 * 4
    obj = sim.random_object(side, side, seed, amp=(0.5, 1.0), phase=(-3.0, 3.0), sigma=0.7)
    probe = rng.standard_normal((nmode, n, n)) + 1j * rng.standard_normal((nmode, n, n))
    counts = np.zeros((npos, n, n), dtype=np.float32)
    return counts, y, x, obj, probe


@pytest.mark.parametrize("n", [16, 24, 32, 48, 64, 96, 128, 256, 512])
@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_fft_accuracy_through_model_intensities(n, dtype):
    counts, y, x, obj, probe = noise_scan(n)
    g = pa.GpuPtycho(pa.Scan(counts, y, x, sim.WAVELENGTH, sim.DISTANCE, sim.DETECTOR_PIXEL),
                     obj, probe, dtype=dtype)
    got = g.intensities()
    o = far_field_ptycho(counts.astype(np.float64), y, x, obj, probe, sim.WAVELENGTH,
                         sim.DISTANCE, sim.DETECTOR_PIXEL)
    assert np.array_equal(g.corners, np.c_[o.addr.oy, o.addr.ox])
    want = to_natural(_mode_sum(o.optics.exit_forward(o.calc_psi()), o.addr.nmode))
    err = float(np.abs(got - want).max() / np.abs(want).max())
    assert err <= {"c64": 2e-5, "c128": 1e-12}[dtype], err
