"""Seeded synthetic scans for the benchmark configurations: probes, objects, scan positions and
Poisson counts. Data preparation only; no reconstruction runs here.

The far-field intensities are |DFT(fftshift(P·O·q))|² summed over modes, written in detector
order, the model of ptyrs/src/engine/ops.rs:42-154 (quad_phase, obj_probe_to_psi, llk).
"""

import math

import numpy as np

WAVELENGTH, DISTANCE, DETECTOR_PIXEL = 1.5e-10, 1.0, 75e-6


def pixel_size(n, wavelength=WAVELENGTH, distance=DISTANCE, detector_pixel=DETECTOR_PIXEL):
    """Object pixel λz/(N·dpix) (ptyrs/src/base/geometry.rs:118-126)."""
    return wavelength * distance / (n * detector_pixel)


def quad_phase(n, wavelength=WAVELENGTH, distance=DISTANCE, detector_pixel=DETECTOR_PIXEL):
    px = pixel_size(n, wavelength, distance, detector_pixel)
    d = np.arange(n, dtype=np.float64) - n // 2
    ph = math.pi / (wavelength * distance) * (d[None, :] ** 2 + d[:, None] ** 2) * px * px
    return np.exp(1j * ph)


def _hermite(k, t):
    """Physicists' Hermite polynomial H_k(t), k <= 3."""
    return [np.ones_like(t), 2 * t, 4 * t * t - 2, 8 * t ** 3 - 12 * t][k]


def hermite_gauss_probe(n, powers=(0.7, 0.15, 0.1, 0.05), sigma=None, curvature=1.5):
    """Incoherent Hermite–Gauss modes (0,0), (1,0), (0,1), (1,1), … on one Gaussian envelope
    of width `sigma` (default n/12) with a defocus phase `curvature`·r²/σ² (the construction of
    ptyrs/src/base/sim.rs:93-125), mode k carrying `powers[k]` of a total of 1 (sum of
    |P|² over the plane)."""
    orders = [(0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (0, 2), (2, 1), (1, 2), (3, 0),
              (0, 3)][:len(powers)]
    sigma = sigma or n / 12
    v, u = (np.mgrid[:n, :n] - n / 2) / sigma
    r2 = u * u + v * v
    modes = []
    for k, (a, b) in enumerate(orders):
        m = _hermite(a, u) * _hermite(b, v) * np.exp(-r2 / 2) * np.exp(1j * (curvature * r2 + 0.3 * k))
        modes.append(m * math.sqrt(powers[k]) / np.linalg.norm(m))
    return np.stack(modes)


def gaussian_probe(n, sigma=10.0, defocus=0.02):
    """One mode: Gaussian amplitude (σ px) times a quadratic phase `defocus`·r² rad/px²."""
    v, u = np.mgrid[:n, :n] - n / 2
    r2 = u * u + v * v
    p = np.exp(-r2 / (2 * sigma * sigma)) * np.exp(1j * defocus * r2)
    return (p / np.linalg.norm(p))[None]


def random_object(nyo, nxo, seed, amp=(0.9, 1.0), phase=(-0.5, 0.5), sigma=4.0):
    """Amplitude uniform in `amp`; phase Gaussian-filtered (σ px, periodic) white noise scaled to
    the range `phase`."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(amp[0], amp[1], (nyo, nxo))
    w = rng.standard_normal((nyo, nxo))
    fy = np.fft.fftfreq(nyo)[:, None]
    fx = np.fft.fftfreq(nxo)[None, :]
    g = np.exp(-2 * (math.pi * sigma) ** 2 * (fy * fy + fx * fx))
    s = np.fft.ifft2(np.fft.fft2(w) * g).real
    s = (s - s.min()) / (s.max() - s.min())
    ph = phase[0] + (phase[1] - phase[0]) * s
    return (a * np.exp(1j * ph))[None]


def raster_positions(ny, nx, step, jitter, seed, pixel):
    """`ny`×`nx` raster of `step` px with uniform ±`jitter` px, integer pixels, centred; metres."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[:ny, :nx]
    y = (yy.ravel() - (ny - 1) / 2) * step + rng.integers(-jitter, jitter + 1, ny * nx)
    x = (xx.ravel() - (nx - 1) / 2) * step + rng.integers(-jitter, jitter + 1, ny * nx)
    return np.round(y) * pixel, np.round(x) * pixel


def fermat_positions(npos, spacing, pixel):
    """Fermat spiral of `npos` points `spacing` px apart, metres."""
    k = np.arange(npos)
    r = spacing * np.sqrt(k)
    th = k * math.pi * (3 - math.sqrt(5))
    return r * np.sin(th) * pixel, r * np.cos(th) * pixel


def corners(y, x, pixel, obj_shape, n):
    """View corners as the engine builds them for integer-pixel positions (mean-centred)."""
    cy = np.round((obj_shape[0] - n) // 2 + (y - y.mean()) / pixel).astype(np.int64)
    cx = np.round((obj_shape[1] - n) // 2 + (x - x.mean()) / pixel).astype(np.int64)
    return cy, cx


def intensities(obj, probe, cy, cx, quad, chunk=64):
    """Model intensities Σ_m |DFT_unitary(fftshift(P_m·O_win·q))|², detector order, float64."""
    n = probe.shape[-1]
    out = np.empty((len(cy), n, n))
    for s in range(0, len(cy), chunk):
        idx = range(s, min(len(cy), s + chunk))
        win = np.stack([obj[0, cy[i]:cy[i] + n, cx[i]:cx[i] + n] for i in idx])
        psi = probe[None] * win[:, None] * quad
        far = np.fft.fft2(np.fft.ifftshift(psi, axes=(-2, -1)), norm="ortho")
        out[s:s + len(idx)] = np.fft.fftshift((np.abs(far) ** 2).sum(axis=1), axes=(-2, -1))
    return out


def scan(config, seed=20261202):
    """The synthetic scan of a benchmark configuration: dict of counts (float32, detector order),
    y, x (metres), truth object and probe, start object and probe, geometry."""
    if config == "headline":
        n, obj_side, npos_side, step, jitter, photons = 256, 1024, 32, 24, 2, 1e7
        probe = hermite_gauss_probe(n)
    elif config == "small":
        n, obj_side, npos_side, step, jitter, photons = 64, 256, 10, 8, 1, 1e6
        probe = gaussian_probe(n, 10.0, 0.02)
    else:
        raise ValueError(config)
    pixel = pixel_size(n)
    y, x = raster_positions(npos_side, npos_side, step, jitter, seed, pixel)
    obj = random_object(obj_side, obj_side, seed + 1)
    cy, cx = corners(y, x, pixel, obj.shape[-2:], n)
    q = quad_phase(n)
    I = intensities(obj, probe, cy, cx, q)
    scale = photons / I.sum(axis=(1, 2)).mean()
    rng = np.random.default_rng(seed + 2)
    counts = rng.poisson(I * scale).astype(np.float32)
    truth_probe = probe * math.sqrt(scale)
    # Random-phase start (a flat start is an unstable point of the iteration).
    start_obj = np.exp(1j * np.random.default_rng(seed + 3).uniform(-1.0, 1.0, obj.shape))
    start_probe = truth_probe.copy()
    return dict(counts=counts, y=y, x=x, obj=obj, probe=truth_probe, start_obj=start_obj,
                start_probe=start_probe, wavelength=WAVELENGTH, distance=DISTANCE,
                detector_pixel=DETECTOR_PIXEL, n=n)
