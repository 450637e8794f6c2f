"""pty-chi's BH (the bilinear-Hessian method of Carlsson, Nikitin and Wendt) on the amplitude
cost, one object slice, far- or near-field exit, no features: a batch's gradient and bilinear
Hessian at one state (`BhBatch`) and the epochs (`Bh`): steepest descent on each batch, or with
one batch of every frame conjugate gradient across epochs, each step's length from the Hessian.

Test infrastructure (see oracle/__init__.py). Layout as oracle/ptycho.py: detector planes
shifted, objects and probes natural. Follows ptyrs/src/engine/bh.rs.
"""

import numpy as np

from .base import random_batches, to_natural, to_shifted

# pty-chi's offset on the modelled amplitude in the residual and the Hessian.
EPS = 1e-8


def frames_of(views):
    """The frames of `views` (whole frames, in order) and the modes per frame."""
    nmode = 1
    while nmode < len(views) and views[nmode][0] == views[0][0]:
        nmode += 1
    if len(views) % nmode:
        raise ValueError("views must be whole frames")
    return [views[k][0] for k in range(0, len(views), nmode)], nmode


class BhBatch:
    """The amplitude cost F = Σ_frames Σ_unmasked (A − √iobs)², A = √(Σ_m |U ψ_m|²), ψ_k =
    P_k·O_k[win]·q, U the exit transform, of a batch of views (whole frames in order) at one
    state, and its derivatives. Stores U ψ, A + EPS, √iobs (NaN where masked) and td = U ψ·(1 −
    √iobs/(A + EPS)) (0 where masked), half the gradient of F over U ψ.
    ptyrs/src/engine/bh.rs:130-272"""

    def __init__(self, views, iobs, optics, obj, probe):
        self.views, self.optics = views, optics
        self.obj, self.probe = obj, probe
        self.frames, self.nmode = frames_of(views)
        self.psi = optics.exit_forward(self._waves(obj, probe))
        nb = len(self.frames)
        blk = self.psi.reshape((nb, self.nmode) + self.psi.shape[1:])
        self.amp = np.sqrt((blk.real ** 2 + blk.imag ** 2).sum(axis=1)) + EPS
        self.root = np.sqrt(np.stack([iobs[f] for f in self.frames]))
        w = np.where(np.isnan(self.root), 0.0, 1.0 - self.root / self.amp)
        self.td = (blk * w[:, None]).reshape(self.psi.shape)

    def _waves(self, obj, probe, other=None):
        """Shifted exit waves P_k·O_k[win]·q, or with `other` = (obj2, probe2) the bilinear
        term P2_k·O_k[win] + P_k·O2_k[win] (times q)."""
        ny, nx = self.probe.shape[-2:]
        q = self.optics.quad
        out = np.empty((len(self.views), ny, nx), dtype=np.complex128)
        for k, (_, _, p, o, oy, ox) in enumerate(self.views):
            w = obj[o, oy:oy + ny, ox:ox + nx]
            if other is None:
                out[k] = to_shifted(probe[p] * w * q)
            else:
                w2 = other[0][o, oy:oy + ny, ox:ox + nx]
                out[k] = to_shifted((other[1][p] * w + probe[p] * w2) * q)
        return out

    def grads(self):
        """(2·∂F/∂O*, 2·∂F/∂P*): U^H(2·td) back-projected, the object's onto the windows
        through conj(P·q), the probe's summed over the views through conj(O[win]·q).
        ptyrs/src/engine/bh.rs:274-343"""
        chi = self.optics.exit_adjoint(2.0 * self.td)
        ny, nx = self.probe.shape[-2:]
        q = self.optics.quad
        g_obj = np.zeros_like(self.obj)
        g_probe = np.zeros_like(self.probe)
        for k, (_, _, p, o, oy, ox) in enumerate(self.views):
            c = to_natural(chi[k])
            g_obj[o, oy:oy + ny, ox:ox + nx] += c * np.conj(self.probe[p] * q)
            g_probe[p] += c * np.conj(self.obj[o, oy:oy + ny, ox:ox + nx] * q)
        return g_obj, g_probe

    def hessian(self, u, v):
        """The second derivative of F along the displacements u = (u_O, u_P) and v, pty-chi's
        bilinear Hessian with A + EPS in place of A: Σ 2·Re(conj(td)·L_uv) + 2·Σ_unmasked
        [(1 − d/A)·Σ_m Re(conj(L_u)·L_v) + d·c_u·c_v/A³], L_u = U(u_O·P + O·u_P)·q, L_uv =
        U(u_O·v_P + v_O·u_P)·q, c_u = Σ_m Re(conj(U ψ)·L_u), d = √iobs.
        ptyrs/src/engine/bh.rs:497-577"""
        fwd = self.optics.exit_forward
        lu = fwd(self._waves(self.obj, self.probe, u))
        lv = fwd(self._waves(self.obj, self.probe, v))
        luv = fwd(self._waves(u[0], u[1], v))
        h = float(2.0 * (np.conj(self.td) * luv).real.sum())
        nb = len(self.frames)
        sh = (nb, self.nmode) + self.psi.shape[1:]
        lu, lv, psi = lu.reshape(sh), lv.reshape(sh), self.psi.reshape(sh)
        uv = (np.conj(lu) * lv).real.sum(axis=1)
        cu = (np.conj(psi) * lu).real.sum(axis=1)
        cv = (np.conj(psi) * lv).real.sum(axis=1)
        keep = ~np.isnan(self.root)
        d, a = self.root[keep], self.amp[keep]
        # pty-chi's d/A³ (kept for its step lengths).
        h += float((2.0 * ((1.0 - d / a) * uv[keep] + d * cu[keep] * cv[keep] / a ** 3)).sum())
        return h


def dot(a, b):
    """Σ Re(conj(a)·b) over (object, probe) pairs. ptyrs/src/engine/bh.rs:580-583"""
    return float(sum((np.conj(x) * y).real.sum() for x, y in zip(a, b)))


class Bh:
    """BH across epochs: batch size, probe ρ, seed and, with one batch of every frame, the
    previous direction. ptyrs/src/engine/bh.rs:585-799"""

    def __init__(self, batch_size, probe_rho=0.01, seed=7):
        if batch_size < 1:
            raise ValueError("BH batch_size must be >= 1")
        self.batch_size, self.probe_rho, self.seed = batch_size, probe_rho, seed
        self.epochs = 0
        self.prev = None

    def epoch(self, addr, iobs, optics, obj, probe):
        """One epoch over random_batches(nframe, size, seed, epoch): per batch g = grads, Pg =
        g with the probe part times ρ² = probe_rho²·(the exit's intensity scale), the
        direction D = −Pg, or with one batch of every frame after the first epoch D = −Pg +
        β·D′, β = H(Pg, D′)/H(D′, D′); α = −⟨g, D⟩/H(D, D)·count/size; O += α·D_O, P += α·D_P.
        Returns (obj, probe, steps), steps (frames, β, α, top, bottom) per batch."""
        nframe, nmode = addr.nframe, addr.nmode
        size = min(self.batch_size, nframe)
        full = size == nframe
        rho2 = self.probe_rho ** 2 * optics.propagator.intensity_scale()
        views_all = addr.views()
        epoch = self.epochs
        self.epochs += 1
        steps = []
        obj, probe = obj.copy(), probe.copy()
        for frames in random_batches(nframe, size, self.seed, epoch):
            views = [v for f in frames for v in views_all[f * nmode:(f + 1) * nmode]]
            b = BhBatch(views, iobs, optics, obj, probe)
            g = b.grads()
            pg = (g[0], g[1] * rho2)
            prev, self.prev = self.prev, None
            if prev is not None and full:
                beta = b.hessian(pg, prev) / b.hessian(prev, prev)
                d = (beta * prev[0] - pg[0], beta * prev[1] - pg[1])
            else:
                beta, d = 0.0, (-pg[0], -pg[1])
            top = -dot(g, d)
            bottom = b.hessian(d, d)
            alpha = top / bottom * len(frames) / size
            obj = obj + alpha * d[0]
            probe = probe + alpha * d[1]
            if full:
                self.prev = d
            steps.append((frames, beta, alpha, top, bottom))
        return obj, probe, steps
