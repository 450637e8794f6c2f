"""The oracle's propagators against pty-chi's f64 outputs that the reference saved
(ptyrs/tests/engine_propagate.rs:42-105), and its FFT against a brute-force DFT
(ptyrs/src/engine/fft.rs:105-126)."""

import math

import numpy as np

from conftest import table
from oracle.base import shift_index, to_natural, to_shifted
from oracle.optics import ANGULAR_SPECTRUM, FOURIER, FRESNEL, Propagator, Wavefield, fft2, ifft2


def _rel(a, b):
    return 0.0 if a == b else abs(a - b) / abs(b)


def _wavefield(r):
    return Wavefield([int(r[0]), int(r[1])], r[2], [r[3], r[4]], r[5])


def test_free_space_selection_matches_ptychi():
    for r in table("propagator_select_f64.txt"):
        w = _wavefield(r)
        p = Propagator.free_space(w)
        assert p.kind == int(r[6]), r
        assert _rel(w.fresnel_number(), r[7]) <= 1e-15
        assert _rel(p.intensity_scale(), r[8]) <= 1e-15


def test_propagators_match_ptychi():
    fields = table("propagator_fields_f64.txt")
    for c in table("propagator_cases_f64.txt"):
        w = _wavefield(c[2:8])
        p = Propagator.angular_spectrum(w) if c[1] == 1.0 else Propagator.free_space(w)
        assert p.kind == int(c[8])
        k = p.intensity_scale()
        assert _rel(k, c[9]) <= 1e-15
        ny, nx = w.shape
        rows = np.array([r for r in fields if r[0] == c[0]])
        assert len(rows) == ny * nx
        at = lambda col: (rows[:, col] + 1j * rows[:, col + 1])
        j = np.arange(ny * nx)
        sj = shift_index(j, ny, nx)
        if p.kind == FOURIER:
            det_i, sign = j, np.where(((j // nx) + (j % nx)) % 2 == 0, 1.0, -1.0)
        else:
            det_i, sign = sj, np.ones(ny * nx)
        x = at(3)[sj].reshape(ny, nx)
        y = (at(5)[det_i] * sign).reshape(ny, nx)
        got_f = p.forward(x[None])[0].ravel()
        got_b = p.adjoint(y[None])[0].ravel()
        want_f = at(7)[det_i] * sign / math.sqrt(k)
        want_b = at(9)[sj] * math.sqrt(k)
        for got, want in ((got_f, want_f), (got_b, want_b)):
            d = np.abs(got - want).max() / np.abs(want).max()
            assert d <= 1e-14, (c[0], d)


def test_fft_matches_direct_dft_and_inverts():
    rng = np.random.default_rng(1)
    ny, nx = 4, 6
    x = rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))
    ky, kx = np.mgrid[:ny, :nx]
    want = np.empty((ny, nx), dtype=complex)
    for a in range(ny):
        for b in range(nx):
            ph = -2 * np.pi * (a * ky / ny + b * kx / nx)
            want[a, b] = (x * np.exp(1j * ph)).sum() / math.sqrt(ny * nx)
    assert np.abs(fft2(x) - want).max() < 1e-12
    assert np.abs(ifft2(fft2(x)) - x).max() < 1e-13


def test_propagators_are_unitary_and_pass_the_dot_test():
    rng = np.random.default_rng(3)
    for w, kind in [
        (Wavefield([12, 16], 1e-10, [50e-9, 50e-9], math.inf), FOURIER),
        (Wavefield([12, 16], 1e-10, [50e-9, 50e-9], 1e-4), ANGULAR_SPECTRUM),
        (Wavefield([12, 16], 1e-10, [60e-9, 45e-9], -2e-4), ANGULAR_SPECTRUM),
        (Wavefield([12, 16], 1e-10, [50e-9, 50e-9], 1.25e-3), FRESNEL),
        (Wavefield([12, 16], 1e-10, [40e-9, 50e-9], -1.6e-3), FRESNEL),
    ]:
        p = Propagator.free_space(w)
        assert p.kind == kind
        x = rng.standard_normal((3, 12, 16)) + 1j * rng.standard_normal((3, 12, 16))
        y = rng.standard_normal((3, 12, 16)) + 1j * rng.standard_normal((3, 12, 16))
        lhs = np.vdot(y, p.forward(x))
        rhs = np.vdot(p.adjoint(y), x)
        assert abs(lhs - rhs) / abs(lhs) < 1e-13
        fx = p.forward(x)
        assert abs((np.abs(fx) ** 2).sum() / (np.abs(x) ** 2).sum() - 1) < 1e-13
        assert np.abs(p.adjoint(fx) - x).max() < 1e-13


def test_shift_layout_is_its_own_inverse():
    a = np.arange(24.0).reshape(4, 6)
    assert shift_index(0, 4, 6) == 2 * 6 + 3
    assert shift_index(shift_index(7, 4, 6), 4, 6) == 7
    assert to_shifted(a).ravel()[0] == a.ravel()[2 * 6 + 3]
    assert np.array_equal(to_natural(to_shifted(a)), a)
