"""pty-chi's position model under its PATCH shift operator and LSQML's position step, as the
reference's CPU engine restates them: the probe (with the plane phase it carries) is moved by
each frame's sub-pixel offset in Fourier space on its own periodic grid, the window stays on
whole object pixels; the offsets are refined by a gradient direction per frame and an SGD step
clipped by the batch's mean absolute deviation.

Test infrastructure (see oracle/__init__.py). Follows ptyrs/src/engine/features.rs:334-521 and
ptyrs/src/engine/propagate.rs:131-163.
"""

import math

import numpy as np

from .optics import freq

POSITION_STEP = 0.3     # ptyrs/src/engine/features.rs:448-455
POSITION_LIMIT = 0.1
EPS_POSITION = 1e-6


def _freqs(shape):
    ny, nx = shape
    return freq(np.arange(ny), ny)[:, None], freq(np.arange(nx), nx)[None, :]


def ramp(shape, offset):
    """`exp(−2πi f·offset)` on the unshifted fftfreq grid. ptyrs/src/engine/features.rs:362-368"""
    fy, fx = _freqs(shape)
    return np.exp(-2j * math.pi * (fy * offset[0] + fx * offset[1]))


def move_planes(planes, quad, offset):
    """Natural-layout planes moved by `offset` ([dy, dx] pixels) as pty-chi LSQML moves its
    probe: `conj(q)·DFT⁻¹(DFT(q·p)·exp(−2πi f·offset))` per plane, `quad` None for none.
    ptyrs/src/engine/features.rs:386-429"""
    q = np.ones(planes.shape[-2:], dtype=np.complex128) if quad is None else quad
    spec = np.fft.fft2(planes * q) * ramp(planes.shape[-2:], offset)
    return np.fft.ifft2(spec) * np.conj(q)


def moved_quad(optics, offset):
    """The last slice's plane phase displaced by a frame's offset, `q(r − d)`:
    `q·exp(i Σ_a c_a·d_a·(d_a − 2 r_a))` over the centred pixel r (multislice only).
    ptyrs/src/engine/propagate.rs:134-163"""
    ny, nx = optics.quad.shape
    c = optics.plane_quad
    ry = (np.arange(ny) - ny // 2)[:, None].astype(np.float64)
    rx = (np.arange(nx) - nx // 2)[None, :].astype(np.float64)
    ph = c[0] * offset[0] * (offset[0] - 2.0 * ry) + c[1] * offset[1] * (offset[1] - 2.0 * rx)
    return optics.quad * (np.cos(ph) + 1j * np.sin(ph))


def gradient_sums(patch, w, chi):
    """Per axis [y, x]: `num = Σ Re(conj(∂O·w)·chi)`, `den = Σ |∂O·w|²`, ∂O the Fourier
    derivative of the natural-layout window taken as periodic; w and chi natural layout.
    ptyrs/src/engine/features.rs:457-497"""
    spec = np.fft.fft2(patch)
    fy, fx = _freqs(patch.shape)
    num, den = [0.0, 0.0], [0.0, 0.0]
    for axis, f in enumerate((fy, fx)):
        g = np.fft.ifft2(spec * (2j * math.pi * f))
        u = g * w
        num[axis] = float((np.conj(u) * chi).real.sum())
        den[axis] = float((u.real ** 2 + u.imag ** 2).sum())
    return num, den


def derivative_sums(gy, gx, w, chi):
    """`gradient_sums` of given window derivatives [∂O/∂y, ∂O/∂x] (under FullObject those of
    the sampled window over the position); all natural layout.
    ptyrs/src/engine/features.rs:474-492"""
    num, den = [0.0, 0.0], [0.0, 0.0]
    for axis, g in enumerate((gy, gx)):
        u = g * w
        num[axis] = float((np.conj(u) * chi).real.sum())
        den[axis] = float((u.real ** 2 + u.imag ** 2).sum())
    return num, den


def gradient_directions(sums, eps):
    """`num / (den + max(max_batch den, eps))` per frame and axis.
    ptyrs/src/engine/features.rs:499-507"""
    mx = [max([eps] + [s[1][a] for s in sums]) for a in (0, 1)]
    return [[s[0][a] / (s[1][a] + mx[a]) for a in (0, 1)] for s in sums]


def batch_position_step(d, frames, delta):
    """pty-chi's SGD position step: 0.3·delta at the batch's frames, 0 elsewhere, each axis
    clipped to min(10·mean|dx − mean dx|, 0.1) over every frame; `d` [nframe, 2] in place.
    ptyrs/src/engine/features.rs:509-521"""
    dx = np.zeros_like(d)
    for f, v in zip(frames, delta):
        dx[f] = [POSITION_STEP * v[0], POSITION_STEP * v[1]]
    n = float(len(d))
    for axis in (0, 1):
        mean = dx[:, axis].sum() / n
        mad = np.abs(dx[:, axis] - mean).sum() / n
        limit = min(10.0 * mad, POSITION_LIMIT)
        d[:, axis] += np.clip(dx[:, axis], -limit, limit)


def position_step(d, sums):
    """AP's, DM's and ML's position step from every frame's [num, den] (`sums` [nframe, 2, 2]):
    per axis 0.3·num/(den + max den), clipped to ±0.1, less its mean over the frames; `d`
    [nframe, 2] in place. ptyrs/src/engine/features.rs:1026-1047"""
    for axis in (0, 1):
        mx = max(0.0, float(sums[:, 1, axis].max()))
        den = sums[:, 1, axis] + mx
        with np.errstate(divide="ignore", invalid="ignore"):
            step = np.where(den > 0.0, np.clip(POSITION_STEP * sums[:, 0, axis] / den,
                                               -POSITION_LIMIT, POSITION_LIMIT), 0.0)
        d[:, axis] += step - step.sum() / len(step)
