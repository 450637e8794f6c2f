"""pty-chi's FULL_OBJECT position model (`Sampling::FullObject`): the probe stays put and each
view reads its object plane at its corner plus its frame's offset by Fourier interpolation of
the whole plane, `patch[u] = DFT⁻¹(DFT(O)·exp(+2πi f·(corner + d)))[u]` over the window's
pixels `u` (numpy `fftfreq` bins `f` of the object plane; a sample past the plane's edge
wraps). Its adjoint puts a patch at the plane's origin, applies the conjugate phase in the
spectrum and sums over the views before one inverse transform per plane.

Test infrastructure (see oracle/__init__.py). Follows ptyrs/src/engine/sampling.rs:1-261.
"""

import numpy as np

from .optics import FOURIER


def auto_sampling(propagator_kind, obj_shape, frame):
    """pty-chi's AUTO rule: FullObject in the near field when 5·frame > 4·object on both axes,
    else PATCH. ptyrs/src/engine/sampling.rs:38-49"""
    large = all(5 * frame[a] > 4 * obj_shape[a] for a in range(2))
    return "full_object" if propagator_kind != FOURIER and large else "patch"


def _phase(shape, t, axes=()):
    """exp(+2πi f·t)·Π_a 2πi f_a over the plane's spectrum. ptyrs/src/engine/sampling.rs:112-120"""
    fy = np.fft.fftfreq(shape[0])[:, None]
    fx = np.fft.fftfreq(shape[1])[None, :]
    e = np.exp(2j * np.pi * (fy * t[0] + fx * t[1]))
    for a in axes:
        e = e * (2j * np.pi * (fy if a == 0 else fx))
    return e


def square_spectrum(t, n):
    """The spectrum over n points of |h_t[−v]|², h_t the one-axis kernel that samples at t: at
    index m, (1 − m/n)·exp(−2πi t m/n) + (m/n)·exp(−2πi t (m − n)/n).
    ptyrs/src/engine/sampling.rs:217-230"""
    a = np.arange(n) / n
    return (1.0 - a) * np.exp(-2j * np.pi * t * a) + a * np.exp(-2j * np.pi * t * (a - 1.0))


class Sampler:
    """The sampling of a set of views: view k reads object plane `views[k]`'s at its corner
    plus `offsets[k]` (one offset per view). ptyrs/src/engine/sampling.rs:79-110"""

    def __init__(self, obj_shape, window, views, offsets):
        self.shape = tuple(int(s) for s in obj_shape)
        self.window = tuple(int(s) for s in window)
        offsets = np.asarray(offsets, dtype=np.float64).reshape(-1, 2)
        if len(offsets) != len(views):
            raise ValueError("one offset per view")
        self.points = [(int(v[3]), (float(v[4]) + d[0], float(v[5]) + d[1]))
                       for v, d in zip(views, offsets)]

    def _groups(self):
        """Runs of consecutive views at one point (a frame's modes on one object mode), which
        share one transform. ptyrs/src/engine/sampling.rs:130-133"""
        out, start = [], 0
        for k in range(1, len(self.points) + 1):
            if k == len(self.points) or self.points[k] != self.points[start]:
                out.append((start, k))
                start = k
        return out

    def windows(self, obj, axes=()):
        """Every view's window of `obj` ([nobj, nyo, nxo]) at its position, or with `axes`
        (0 for y, 1 for x, repeats for higher orders) its derivative over the position.
        ptyrs/src/engine/sampling.rs:122-145"""
        py, px = self.window
        spectra = np.fft.fft2(obj)
        out = np.empty((len(self.points), py, px), dtype=np.complex128)
        for a, b in self._groups():
            o, t = self.points[a]
            plane = np.fft.ifft2(spectra[o] * _phase(self.shape, t, axes))
            out[a:b] = plane[:py, :px]
        return out

    def _convolve(self, patches, nobj, kernel):
        """Per view point, `patches` summed at the plane's origin and circularly convolved by
        the kernel whose spectrum `kernel(t)` gives, summed onto `nobj` object planes.
        ptyrs/src/engine/sampling.rs:154-187"""
        py, px = self.window
        patches = np.asarray(patches).reshape(len(self.points), py, px)
        spec = np.zeros((nobj,) + self.shape, dtype=np.complex128)
        for a, b in self._groups():
            o, t = self.points[a]
            plane = np.zeros(self.shape, dtype=np.complex128)
            for k in range(a, b):
                plane[:py, :px] += patches[k]
            spec[o] += np.fft.fft2(plane) * kernel(t)
        return np.fft.ifft2(spec)

    def place(self, patches, nobj):
        """The adjoint of `windows` without derivative: `patches` ([nview, py, px], natural
        layout) placed onto `nobj` object planes. ptyrs/src/engine/sampling.rs:147-152"""
        return self._convolve(patches, nobj, lambda t: np.conj(_phase(self.shape, t)))

    def place_norm(self, patches, nobj):
        """`place` of the real `patches`, its real part floored at 0: a placed sum of squared
        moduli. ptyrs/src/engine/sampling.rs:189-194"""
        return np.maximum(self.place(np.asarray(patches, dtype=np.complex128), nobj).real, 0.0)

    def diag(self, patches, nobj):
        """The diagonal of AᴴA, A the views' windows each weighted by a W with |W|² the real
        `patches`, its real part floored at 0. ptyrs/src/engine/sampling.rs:196-209"""
        ny, nx = self.shape

        def kernel(t):
            return square_spectrum(t[0], ny)[:, None] * square_spectrum(t[1], nx)[None, :]
        return np.maximum(self._convolve(np.asarray(patches, dtype=np.complex128), nobj,
                                         kernel).real, 0.0)


def view_offsets(views, offsets):
    """Each view's offset: its frame's row of `offsets` ([nframe, 2])."""
    return np.array([offsets[v[0]] for v in views], dtype=np.float64).reshape(-1, 2)


def full_object_illumination(views, obj_shape, probe, offsets):
    """pty-chi's illumination map under FullObject: each view's |P|² placed at its corner plus
    its frame's offset by the sampling's adjoint, real part floored at 0.
    ptyrs/src/engine/sampling.rs:51-75"""
    sm = Sampler(obj_shape[-2:], probe.shape[-2:], views, view_offsets(views, offsets))
    p2 = probe.real ** 2 + probe.imag ** 2
    return sm.place_norm(np.stack([p2[v[2]] for v in views]), obj_shape[0])
