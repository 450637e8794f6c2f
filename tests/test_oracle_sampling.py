"""The oracle's FullObject sampling (oracle/sampling.py) against closed forms, restating the
reference's unit tests (ptyrs/src/engine/sampling.rs:263-436): the windows of a band-limited
object are its trigonometric interpolant and the interpolant's derivatives, past the plane's
edge too; `place` is the adjoint of `windows`; `diag` is the diagonal of the normal operator
at fractional and whole-pixel offsets (where it is `place_norm`); at zero offsets the placed
|P|² is the PATCH illumination; the AUTO rule."""

import math

import numpy as np

from oracle.base import AddrTable
from oracle.lsqml import illumination
from oracle.optics import ANGULAR_SPECTRUM, FOURIER, freq
from oracle.sampling import Sampler, auto_sampling, full_object_illumination

# Odd rows: index 5 of 11 is +5/11, which one term of each mode reads. Even columns: index 5 of
# 10 is the Nyquist, −1/2, which mode 1's k_x = −5 term reads as exp(−iπx).
OBJ = (11, 10)
WINDOW = (6, 5)
# Views (frame, mode, probe, obj, oy, ox): two probe modes at one point of object mode 0, then
# object mode 1 past the plane's corner and at a whole pixel.
VIEWS = [(0, 0, 0, 0, 1, 2), (0, 0, 1, 0, 1, 2), (0, 0, 0, 1, 8, 7), (0, 0, 1, 1, 3, 0)]
OFFSETS = [[0.3, -0.2], [0.3, -0.2], [-0.45, 0.35], [0.0, 0.0]]
TERMS = [
    [(1.0 + 0.0j, (0.0, 0.0)), (0.3 - 0.2j, (1.0, -2.0)), (-0.25 + 0.4j, (-5.0, 4.0)),
     (0.1 + 0.15j, (3.0, 1.0))],
    [(0.7 + 0.2j, (0.0, 1.0)), (-0.2 + 0.35j, (-2.0, 3.0)), (0.45 + 0.05j, (5.0, -4.0)),
     (-0.3 + 0.25j, (3.0, -5.0))],
]


def value(mode, t, axes=()):
    """Object mode `mode`'s polynomial at t, differentiated over t's components `axes`."""
    out = 0j
    for c, k in TERMS[mode]:
        w = [2 * math.pi * k[a] / OBJ[a] for a in (0, 1)]
        d = 1.0 + 0j
        for a in axes:
            d *= 1j * w[a]
        out += c * d * np.exp(1j * (w[0] * t[0] + w[1] * t[1]))
    return out


def noise(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def test_fftfreq_is_the_references_bin_convention():
    for n in (10, 11):
        assert np.array_equal(np.fft.fftfreq(n), freq(np.arange(n), n))


def test_windows_are_the_interpolant_and_its_derivatives():
    yy, xx = np.meshgrid(np.arange(OBJ[0]), np.arange(OBJ[1]), indexing="ij")
    obj = np.stack([value(m, (yy, xx)) for m in (0, 1)])
    sm = Sampler(OBJ, WINDOW, VIEWS, OFFSETS)
    uy, ux = np.meshgrid(np.arange(WINDOW[0]), np.arange(WINDOW[1]), indexing="ij")
    for axes in [(), (0,), (1,), (0, 1), (0, 0), (1, 1)]:
        got = sm.windows(obj, axes)
        want = np.stack([value(v[3], (v[4] + d[0] + uy, v[5] + d[1] + ux), axes)
                         for v, d in zip(VIEWS, OFFSETS)])
        assert np.abs(got - want).max() < 1e-13 * np.abs(want).max(), axes


def test_place_is_the_adjoint_of_windows():
    sm = Sampler(OBJ, WINDOW, VIEWS, OFFSETS)
    obj = noise((2,) + OBJ, 1)
    patches = noise((4,) + WINDOW, 2)
    win = sm.windows(obj)
    lhs = np.vdot(win, patches)
    rhs = np.vdot(obj, sm.place(patches, 2))
    # the dot test's bound, relative to the product of the norms (the random sum cancels)
    assert abs(lhs - rhs) < 1e-14 * np.linalg.norm(win) * np.linalg.norm(patches)


def test_diag_is_the_diagonal_of_the_normal_operator():
    power = np.abs(noise((4,) + WINDOW, 4)) ** 2
    whole = np.round(OFFSETS)
    for offsets in (OFFSETS, whole):
        sm = Sampler(OBJ, WINDOW, VIEWS, offsets)
        got = sm.diag(power, 2)
        want = np.zeros((2,) + OBJ)
        unit = np.zeros((2,) + OBJ, dtype=np.complex128)
        for i in np.ndindex(*want.shape):
            unit[i] = 1.0
            want[i] = (power * np.abs(sm.windows(unit)) ** 2).sum()
            unit[i] = 0.0
        scale = want.max()
        assert np.abs(got - want).max() < 1e-14 * scale
        if offsets is whole:
            assert np.abs(got - sm.place_norm(power, 2)).max() < 1e-14 * scale


def test_full_object_illumination_without_offsets_is_the_patch_one():
    addr = AddrTable(np.array([0.0, 2.0, -3.0]), np.array([1.0, -2.0, 2.0]), [1.0, 1.0],
                     [8, 6], [14, 12], 2, 2)
    probe = noise((2, 8, 6), 3)
    views = addr.views()
    want = illumination(views, (2, 14, 12), probe)
    got = full_object_illumination(views, (2, 14, 12), probe, np.zeros((3, 2)))
    assert np.abs(got - want).max() < 1e-14 * want.max()


def test_auto_takes_full_object_for_a_near_field_probe_over_most_of_the_object():
    assert auto_sampling(FOURIER, (9, 7), (8, 6)) == "patch"
    assert auto_sampling(ANGULAR_SPECTRUM, (9, 7), (8, 6)) == "full_object"
    # 5·8 = 4·10 and 5·6 < 4·8
    assert auto_sampling(ANGULAR_SPECTRUM, (10, 7), (8, 6)) == "patch"
    assert auto_sampling(ANGULAR_SPECTRUM, (9, 8), (8, 6)) == "patch"
