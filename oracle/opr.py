"""pty-chi's orthogonal probe relaxation (OPR, Odstrčil et al. 2016) of probe mode 0: with k
eigenmodes U_i, frame f sees P_f0 = c_f·P_0 + Σ_i w_fi·U_i and P_fm = P_m for m ≥ 1; c_f is 1
unless modelled. LSQML gives every frame of a batch its own probe planes (`unique`) and, after
each step, moves the eigenmodes and the batch's weights (`step`).

Test infrastructure (see oracle/__init__.py). Follows ptyrs/src/engine/opr.rs.
"""

import math

import numpy as np

from .base import to_natural
from .lsqml import AMPLITUDE
from .ptycho import obj_probe_to_psi


class ProbeVariation:
    """k eigenmodes [k, ny, nx] of probe mode 0, their weights for every frame (frame-major,
    w_fi at f·k + i) and, when modelled, the main mode's weight c_f of every frame; each of
    the weights and c refined or held. ptyrs/src/engine/opr.rs:19-90"""

    def __init__(self, eigenmodes, weights, refine=True, intensity=None,
                 refine_intensity=True):
        self.eigenmodes = np.array(eigenmodes, dtype=np.complex128)
        self.k = self.eigenmodes.shape[0]
        self.weights = np.array(weights, dtype=np.float64).ravel()
        if self.k == 0 or self.weights.size % self.k:
            raise ValueError("weights must be [nframe, k] for k >= 1 eigenmodes")
        self.refine = refine
        self.intensity = None if intensity is None else np.array(intensity, dtype=np.float64)
        if self.intensity is not None and self.intensity.size != self.weights.size // self.k:
            raise ValueError("one main-mode weight per frame")
        self.refine_intensity = refine_intensity

    def copy(self):
        return ProbeVariation(self.eigenmodes.copy(), self.weights.copy(), self.refine,
                              None if self.intensity is None else self.intensity.copy(),
                              self.refine_intensity)

    def unique(self, views, probe):
        """`views` (whole frames, in order) with probe plane b·np + m for the b-th frame and
        probe mode m, and those planes: c_f·P_0 + w_f0·U_0 + w_f1·U_1 + … in that order, then
        P_1 … ptyrs/src/engine/opr.rs:92-128"""
        np_ = probe.shape[0]
        frames, out = [], []
        for v in views:
            if not frames or frames[-1] != v[0]:
                frames.append(v[0])
            out.append((v[0], v[1], (len(frames) - 1) * np_ + v[2], v[3], v[4], v[5]))
        planes = []
        for f in frames:
            c = 1.0 if self.intensity is None else self.intensity[f]
            p0 = c * probe[0]
            for i in range(self.k):
                p0 = p0 + self.weights[f * self.k + i] * self.eigenmodes[i]
            planes.append(p0)
            planes.extend(probe[1:])
        return out, np.stack(planes)

    def step(self, frames, probe0, moved, d, vv, np_):
        """pty-chi OPRModeWeights.update_opr_probe_modes_and_weights on the batch `frames` from
        d (plane b·np: conj(q·V_b)·χ_b0), vv (|V_b|²) and `moved` (d taken back through the
        frames' probe moves), probe0 the shared probe's mode 0; the main-mode weights' SGD step
        first. ptyrs/src/engine/opr.rs:130-237"""
        if self.intensity is not None and self.refine_intensity:
            pp = probe0.real ** 2 + probe0.imag ** 2
            for b, f in enumerate(frames):
                den = float((pp * vv[b * np_]).sum())
                self.intensity[f] += 0.1 * _dot(probe0, d[b * np_]) / den
        k, count = self.k, len(frames)
        if k == 0 or count < 2:
            return
        n = d[0].size
        nn, cnt = float(n), float(count)
        mean = sum(moved[b * np_] for b in range(count)) / cnt
        r = [moved[b * np_] - mean for b in range(count)]
        a_u = min(0.1, cnt / (self.weights.size // k))
        for i in range(k):
            w = [self.weights[f * k + i] for f in frames]
            w2 = sum(x * x for x in w)
            u = self.eigenmodes[i]
            e = np.zeros_like(u)
            for rb, wb in zip(r, w):
                p = (_dot(rb, u) / (nn * nn) + wb) / w2
                e = e + rb * p
            e = e / cnt
            u = u + (a_u * math.sqrt(nn) / _mnorm(e)) * e
            u = u * (math.sqrt(nn) / _mnorm(u))
            if self.refine:
                den = [float((np.abs(u) ** 2 * vv[b * np_]).sum()) for b in range(count)]
                reg = 0.1 * sum(den) / cnt
                for b, f in enumerate(frames):
                    self.weights[f * k + i] = w[b] + _dot(u, d[b * np_]) / (den[b] + reg)
            if i + 1 < k:
                uu = float((np.abs(u) ** 2).sum())
                r = [rb - ((rb * np.conj(u)).sum() / uu) * u for rb in r]
            self.eigenmodes[i] = u


def _dot(a, b):
    """Σ Re(conj(a)·b). ptyrs/src/engine/opr.rs:240-243"""
    return float((np.conj(a) * b).real.sum())


def _mnorm(x):
    """√(mean |x|²). ptyrs/src/engine/opr.rs:245-248"""
    return math.sqrt(float((np.abs(x) ** 2).sum()) / x.size)


def opr_weight_grad(cost, views, iobs, quad, obj, probe, var, nmode, exit_forward,
                    exit_adjoint):
    """∂cost/∂w_bi for every frame b of `views` (whole frames, in order) and eigenmode i,
    frame-major: ⟨U_i, g_b0⟩ with g the gradient over each frame's own probe planes, the
    probe back-projection of F⁻¹(2·slope·ψ), slope the cost's derivative over icalc
    (Poisson 1 − iobs/icalc, amplitude 1 − √(iobs/icalc), icalc floored at 1e-12).
    ptyrs/src/engine/opr.rs:250-290, ptyrs/src/engine/features.rs:325-331"""
    np_ = probe.shape[0]
    uviews, planes = var.unique(views, probe)
    ny, nx = probe.shape[-2:]
    psi = exit_forward(obj_probe_to_psi(None, uviews, obj, planes, quad))
    frames = []
    for v in uviews:
        if not frames or frames[-1] != v[0]:
            frames.append(v[0])
    nb = len(frames)
    blk = psi.reshape(nb, nmode, ny, nx)
    icalc = np.maximum((blk.real ** 2 + blk.imag ** 2).sum(axis=1), 1e-12)
    io = np.stack([iobs[f] for f in frames])
    masked = np.isnan(io)
    if cost == AMPLITUDE:
        slope = 1.0 - np.sqrt(np.where(masked, 0.0, io) / icalc)
    else:
        slope = 1.0 - np.where(masked, 0.0, io) / icalc
    s = np.where(masked, 0.0, 2.0 * slope)
    g_psi = exit_adjoint((blk * s[:, None]).reshape(psi.shape))
    g = np.zeros_like(planes)
    for k_, (_, _, p, o, oy, ox) in enumerate(uviews):
        g[p] += to_natural(g_psi[k_]) * np.conj(quad * obj[o, oy:oy + ny, ox:ox + nx])
    k = var.k
    return [_dot(var.eigenmodes[bi % k], g[(bi // k) * np_]) for bi in range(nb * k)]
