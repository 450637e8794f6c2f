"""Exit transforms of the oracle: the unitary 2D FFT, the far-field quadratic phase and the
free-space propagators (Fourier, angular spectrum, Fresnel), all in the reference's
FFT-shifted layout.

Test infrastructure (see oracle/__init__.py). The reference's CPU FFT is rustfft 6.4.1
(ptyrs/src/engine/fft.rs:1-4), not vendored; any accurate FP64 FFT reproduces its results at
the 1e-12 level its fixtures are checked at, so numpy's pocketfft stands in for it here.
"""

import math

import numpy as np

from .base import to_natural, to_shifted


def fft2(planes):
    """Unitary forward transform of every `[ny, nx]` plane. ptyrs/src/engine/fft.rs:6-88"""
    return np.fft.fft2(planes, norm="ortho")


def ifft2(planes):
    return np.fft.ifft2(planes, norm="ortho")


def freq(k, n):
    """numpy `fftfreq` of index `k`. ptyrs/src/engine/features.rs:337-344"""
    k = np.asarray(k)
    return np.where(k < (n + 1) // 2, k, k - n) * (1.0 / n)


def quad_phase(wavelength, distance, pixel, frame):
    """Far-field `q[y, x] = exp(iπ/(λz)·(dx²px² + dy²py²))`, natural layout.
    ptyrs/src/engine/ops.rs:11-33"""
    ny, nx = frame
    py, px = pixel
    f = math.pi / (wavelength * distance)
    dy = (np.arange(ny, dtype=np.float64) - (ny // 2))[:, None]
    dx = (np.arange(nx, dtype=np.float64) - (nx // 2))[None, :]
    ph = f * (dx * dx * px * px + dy * dy * py * py)
    return np.cos(ph) + 1j * np.sin(ph)


class Wavefield:
    """One free-space propagation in metres. ptyrs/src/engine/propagate.rs:218-299"""

    def __init__(self, shape, wavelength, pixel, distance):
        if any(n == 0 or n % 2 for n in shape):
            raise ValueError(f"plane shape {shape} must be even and non-zero")
        for v in (wavelength, pixel[0], pixel[1]):
            if not (math.isfinite(v) and v > 0):
                raise ValueError("wavelength and pixel must be finite and > 0")
        if math.isnan(distance) or distance == 0.0 or distance == -math.inf:
            raise ValueError("distance must be non-zero and not -inf")
        self.shape = list(shape)
        self.wavelength = wavelength
        self.pixel = list(pixel)
        self.distance = distance

    def pixel_width_wlu(self):
        return self.pixel[1] / self.wavelength

    def aspect_ratio(self):
        return self.pixel[1] / self.pixel[0]

    def distance_wlu(self):
        return self.distance / self.wavelength

    def fresnel_number(self):
        w = self.pixel_width_wlu()
        return w * w / self.distance_wlu()

    def prefers_fresnel(self):
        ny, nx = self.shape
        d = abs(self.distance_wlu())
        w = self.pixel_width_wlu()
        h = w / self.aspect_ratio()
        return w * w / d < 1.0 / nx or h * h / d < 1.0 / ny


FOURIER, ANGULAR_SPECTRUM, FRESNEL = 0, 1, 2


class Propagator:
    """A unitary propagator over shifted planes. ptyrs/src/engine/propagate.rs:301-503"""

    def __init__(self, kind, shape, tf=None, a=None, b=None, reverse=False, scale=1.0,
                 plane_quad=(0.0, 0.0)):
        self.kind, self.shape = kind, list(shape)
        self.tf, self.a, self.b = tf, a, b
        self.reverse, self.scale, self.plane_quad = reverse, scale, plane_quad

    @staticmethod
    def free_space(w):
        """pty-chi `build_free_space_propagator`. ptyrs/src/engine/propagate.rs:356-366"""
        if math.isinf(w.distance):
            return Propagator(FOURIER, w.shape, scale=float(w.shape[0] * w.shape[1]))
        if w.prefers_fresnel():
            return Propagator._fresnel(w)
        return Propagator._angular_spectrum(w)

    @staticmethod
    def angular_spectrum(w):
        if math.isinf(w.distance):
            raise ValueError("angular spectrum needs a finite distance")
        return Propagator._angular_spectrum(w)

    @staticmethod
    def _angular_spectrum(w):
        """`T = exp(2πi z √(1 − r))` inside the propagating band, else 0, on the unshifted
        fftfreq grid. ptyrs/src/engine/propagate.rs:505-531"""
        ny, nx = w.shape
        ar = w.aspect_ratio()
        width = w.pixel_width_wlu()
        i2piz = 2.0 * math.pi * w.distance_wlu()
        fy = freq(np.arange(ny), ny)[:, None]
        fx = freq(np.arange(nx), nx)[None, :]
        ratio = (fx * fx + (ar * fy) * (ar * fy)) / (width * width)
        inside = ratio < 1.0
        ph = i2piz * np.sqrt(np.where(inside, 1.0 - ratio, 0.0))
        tf = np.where(inside, np.cos(ph) + 1j * np.sin(ph), 0.0)
        return Propagator(ANGULAR_SPECTRUM, w.shape, tf=tf, scale=1.0)

    @staticmethod
    def _fresnel(w):
        """`x ↦ a·Fft2(b·x)` (Fft2⁻¹ for z < 0). ptyrs/src/engine/propagate.rs:533-566"""
        ny, nx = w.shape
        fr = w.fresnel_number()
        ar = w.aspect_ratio()
        c1 = complex(math.cos(2 * math.pi * w.distance_wlu()),
                     math.sin(2 * math.pi * w.distance_wlu()))
        c1c0 = c1 * complex(0.0, -math.copysign(1.0, fr) if fr != 0 else 0.0)
        n, m = float(nx), float(ny)
        yy = (np.arange(ny, dtype=np.float64) - (ny // 2))[:, None]
        xx = (np.arange(nx, dtype=np.float64) - (nx // 2))[None, :]
        s2 = (xx / n) * (xx / n) + (ar * yy / m) * (ar * yy / m)
        pha = s2 * math.pi / fr
        a_nat = c1c0 * (np.cos(pha) + 1j * np.sin(pha))
        sb = xx * xx + (yy / ar) * (yy / ar)
        phb = math.pi * fr * sb
        b_nat = np.cos(phb) + 1j * np.sin(phb)
        c0 = fr / ar
        return Propagator(FRESNEL, w.shape, a=to_shifted(a_nat), b=to_shifted(b_nat),
                          reverse=w.distance < 0.0, scale=(ny * nx) * (c0 * c0),
                          plane_quad=(math.pi * fr / (ar * ar), math.pi * fr))

    def intensity_scale(self):
        return self.scale

    def after_plane(self, planes):
        """Forward of planes that already carry the Fresnel `B`.
        ptyrs/src/engine/propagate.rs:380-398"""
        if self.kind == FOURIER:
            return fft2(planes)
        if self.kind == ANGULAR_SPECTRUM:
            return ifft2(fft2(planes) * self.tf)
        out = ifft2(planes) if self.reverse else fft2(planes)
        return out * self.a

    def before_plane(self, planes):
        """Adjoint of `after_plane`. ptyrs/src/engine/propagate.rs:410-428"""
        if self.kind == FOURIER:
            return ifft2(planes)
        if self.kind == ANGULAR_SPECTRUM:
            return ifft2(fft2(planes) * np.conj(self.tf))
        x = planes * np.conj(self.a)
        return fft2(x) if self.reverse else ifft2(x)

    def forward(self, planes):
        if self.kind == FRESNEL:
            planes = planes * self.b
        return self.after_plane(planes)

    def adjoint(self, planes):
        out = self.before_plane(planes)
        if self.kind == FRESNEL:
            out = out * np.conj(self.b)
        return out

    def adjoint_inverts(self):
        if self.kind == ANGULAR_SPECTRUM:
            return not np.any(self.tf == 0)
        return True


class Optics:
    """The exit-wave model `psi = E(P·O·q)`: plane phase `quad` (natural layout), the exit
    propagator, the object pixel, and in-object gaps for multislice.
    ptyrs/src/engine/propagate.rs:24-186"""

    def __init__(self, pixel, quad, propagator, gaps=(), plane_quad=(0.0, 0.0)):
        self.pixel = list(pixel)
        self.quad = quad
        self.propagator = propagator
        self.gaps = list(gaps)
        # quad = exp(i·(c[0]·Y² + c[1]·X²)) over the centred pixel (propagate.rs:29-31)
        self.plane_quad = tuple(plane_quad)

    @staticmethod
    def far_field(wavelength, distance, detector_pixel, frame):
        """PyNX far field: `quad_phase` and Fourier. ptyrs/src/engine/propagate.rs:52-62"""
        from .base import pixel_size_object
        pixel = pixel_size_object(wavelength, distance, detector_pixel, frame)
        q = quad_phase(wavelength, distance, pixel, frame)
        lz = wavelength * distance
        return Optics(pixel, q, Propagator(FOURIER, frame, scale=float(frame[0] * frame[1])),
                      plane_quad=(math.pi * pixel[0] ** 2 / lz, math.pi * pixel[1] ** 2 / lz))

    @staticmethod
    def free_space(w):
        """pty-chi planar model; the plane phase is the Fresnel `B`.
        ptyrs/src/engine/propagate.rs:66-83"""
        p = Propagator.free_space(w)
        ny, nx = w.shape
        quad = np.ones((ny, nx), dtype=np.complex128)
        plane_quad = (0.0, 0.0)
        if p.kind == FRESNEL:
            quad = to_natural(p.b)
            plane_quad = p.plane_quad
        return Optics(w.pixel, quad, p, plane_quad=plane_quad)

    def with_slices(self, gaps):
        return Optics(self.pixel, self.quad, self.propagator,
                      [Propagator.angular_spectrum(w) for w in gaps], self.plane_quad)

    def nslice(self):
        return len(self.gaps) + 1

    def exit_forward(self, planes):
        return self.propagator.after_plane(planes)

    def exit_adjoint(self, planes):
        return self.propagator.before_plane(planes)
