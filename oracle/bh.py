"""pty-chi's BH (the bilinear-Hessian method of Carlsson, Nikitin and Wendt) on the amplitude
cost, far- or near-field exit, one object slice with or without position offsets (PATCH, or
FullObject: the object sampled at corner plus offset) or
several slices (`BhSlices`, the object variable every slice, [nslice, nobj, nyo, nxo]): a
batch's gradient and bilinear Hessian at one state (`BhBatch`) and the epochs (`Bh`): steepest
descent on each batch, or with one batch of every frame conjugate gradient across epochs, each
step's length from the Hessian. Refined offsets join the variables (their gradient scaled by
position_rho², their step 0.3·α·D_d).

Test infrastructure (see oracle/__init__.py). Layout as oracle/ptycho.py: detector planes
shifted, objects and probes natural. Follows ptyrs/src/engine/bh.rs.
"""

import math

import numpy as np

from .base import random_batches, to_natural, to_shifted
from .positions import POSITION_STEP, _freqs, ramp
from .sampling import Sampler

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
    W_k·O_k[win], U the exit transform, of a batch of views (whole frames in order) at one
    state, and its derivatives. W_k is the probe mode times the plane phase, P·q, or under PATCH
    offsets (one per frame of the batch, [dy, dx] object pixels) that product moved by its
    frame's offset, W_b = DFT⁻¹(DFT(q·P)·exp(−2πi f·d_b)) (the copy conj(q)·W_b times q).
    Stores U ψ, A + EPS, √iobs (NaN where masked) and td = U ψ·(1 − √iobs/(A + EPS)) (0 where
    masked), half the gradient of F over U ψ. ptyrs/src/engine/bh.rs:130-272"""

    def __init__(self, views, iobs, optics, obj, probe, offsets=None, refine=False,
                 sampling="patch"):
        self.views, self.optics = views, optics
        self.obj, self.probe = obj, probe
        self.frames, self.nmode = frames_of(views)
        self.shape = probe.shape[-2:]
        offsets = None if offsets is None else np.asarray(offsets, dtype=np.float64)
        if offsets is not None and offsets.shape != (len(self.frames), 2):
            raise ValueError("one offset per frame of the batch")
        self.refine = refine and offsets is not None
        # FullObject: every view reads the object sampled at its corner plus its frame's offset
        # (ptyrs/src/engine/bh.rs:96-107); the probe terms stay.
        self.sm = None
        if offsets is not None and sampling == "full_object":
            self.sm = Sampler(obj.shape[-2:], self.shape, views,
                              [offsets[k // self.nmode] for k in range(len(views))])
            offsets = None
        self.offsets = offsets
        if self.offsets is not None:
            self.ramps = [ramp(self.shape, d) for d in self.offsets]
            fy, fx = _freqs(self.shape)
            self.f = (-2j * math.pi * fy, -2j * math.pi * fx)   # −2πi f_a
            self.spec = np.fft.fft2(probe * optics.quad)
        self.w = self._moved(probe)
        self.win = self._reads(obj)
        self.psi = optics.exit_forward(self._exit(obj, self.w))
        nb = len(self.frames)
        blk = self.psi.reshape((nb, self.nmode) + self.psi.shape[1:])
        self.amp = np.sqrt((blk.real ** 2 + blk.imag ** 2).sum(axis=1)) + EPS
        self.root = np.sqrt(np.stack([iobs[f] for f in self.frames]))
        w = np.where(np.isnan(self.root), 0.0, 1.0 - self.root / self.amp)
        self.td = (blk * w[:, None]).reshape(self.psi.shape)

    def _slot(self, k):
        return k // self.nmode

    def _moved(self, probe, terms=None):
        """Per view: probe mode times q, moved by its frame's offset, or with `terms` = [(spec,
        coefficient per frame)] the plane DFT⁻¹(Σ spec·c_b·exp(−2πi f·d_b)) (spec a per-mode
        spectrum of q·X, c_b a per-frame multiplier of the frequency grid)."""
        q = self.optics.quad
        out = np.empty((len(self.views),) + tuple(self.shape), dtype=np.complex128)
        if self.offsets is None:
            for k, v in enumerate(self.views):
                out[k] = probe[v[2]] * q
            return out
        spec = np.fft.fft2(probe * q) if terms is None else None
        for k, v in enumerate(self.views):
            b = self._slot(k)
            if terms is None:
                s = spec[v[2]]
            else:
                s = sum(t[v[2]] * c(b) for t, c in terms)
            out[k] = np.fft.ifft2(s * self.ramps[b])
        return out

    def _fd(self, d, b):
        """Σ_a d[b][a]·(−2πi f_a): the first derivative's factor along a frame's offset change."""
        return d[b][0] * self.f[0] + d[b][1] * self.f[1]

    def _reads(self, x, axes=()):
        """What the views read of object-shaped x: the windows at their corners, under
        FullObject sampled at corner plus offset (with `axes` its derivative over the offset)."""
        if self.sm is not None:
            return self.sm.windows(x, axes)
        ny, nx = self.shape
        return np.stack([x[o, oy:oy + ny, ox:ox + nx] for (_, _, _, o, oy, ox) in self.views])

    def _exit(self, obj, w, obj2=None, w2=None):
        """Shifted exit-plane waves W_k·O_k[win], or with obj2, w2 the bilinear term W_k·O2_k +
        W2_k·O_k."""
        win = self._reads(obj)
        if obj2 is None:
            return np.stack([to_shifted(p) for p in w * win])
        return np.stack([to_shifted(p) for p in w * self._reads(obj2) + w2 * win])

    def grads(self):
        """(2·∂F/∂O*, 2·∂F/∂P*, ∂F/∂d): χ = Uᴴ(2·td) back-projected, the object's onto the
        windows through conj(W), the probe's through each frame's copy, conj(q)·DFT⁻¹(DFT(conj(O)·
        χ)·exp(+2πi f·d_b)) summed over the views, and with refined offsets frame b's Re Σ_views
        conj(χ)·O·∂_a W_b (∂_a W_b = DFT⁻¹(DFT(q·P)·(−2πi f_a)·exp(−2πi f·d_b))); d is None
        otherwise. ptyrs/src/engine/bh.rs:274-343"""
        chi = self.optics.exit_adjoint(2.0 * self.td)
        ny, nx = self.shape
        q = self.optics.quad
        g_obj = np.zeros_like(self.obj)
        g_probe = np.zeros_like(self.probe)
        g_d = np.zeros((len(self.frames), 2)) if self.refine else None
        if self.sm is not None:
            # the reads' gradient placed by the sampling's adjoint; the offsets' through the
            # reads' derivatives (ptyrs/src/engine/bh.rs:296-308)
            c = np.stack([to_natural(x) for x in chi])
            g_win = c * np.conj(self.w)
            g_obj = self.sm.place(g_win, self.obj.shape[0])
            for k, v in enumerate(self.views):
                g_probe[v[2]] += c[k] * np.conj(self.win[k] * q)
            if self.refine:
                for a in range(2):
                    dw = self.sm.windows(self.obj, (a,))
                    for k in range(len(self.views)):
                        g_d[k // self.nmode, a] += float((np.conj(g_win[k]) * dw[k]).real.sum())
            return g_obj, g_probe, g_d
        for k, (_, _, p, o, oy, ox) in enumerate(self.views):
            c = to_natural(chi[k])
            win = self.obj[o, oy:oy + ny, ox:ox + nx]
            g_obj[o, oy:oy + ny, ox:ox + nx] += c * np.conj(self.w[k])
            if self.offsets is None:
                g_probe[p] += c * np.conj(win * q)
                continue
            b = self._slot(k)
            s = np.fft.fft2(np.conj(win) * c)
            g_probe[p] += np.conj(q) * np.fft.ifft2(s * np.conj(self.ramps[b]))
            if self.refine:
                for a in range(2):
                    dw = np.fft.ifft2(self.spec[p] * self.f[a] * self.ramps[b])
                    g_d[b, a] += float((np.conj(c) * win * dw).real.sum())
        return g_obj, g_probe, g_d

    def _lift(self, u):
        """Per view the moved waves' first-order change along u: W(u_P) + Σ_a u_da·∂_a W."""
        if self.offsets is None:
            return self._moved(u[1])
        su = np.fft.fft2(u[1] * self.optics.quad)
        terms = [(su, lambda b: 1.0)]
        if self.refine:
            terms.append((self.spec, lambda b: self._fd(u[2], b)))
        return self._moved(None, terms)

    def _second(self, u, v):
        """Per view the moved waves' second-order change along u and v: Σ_a (v_da·∂_a W(u_P) +
        u_da·∂_a W(v_P)) + Σ_ac u_da·v_dc·∂_a ∂_c W; None without refined offsets."""
        if not self.refine:
            return None
        q = self.optics.quad
        su, sv = np.fft.fft2(u[1] * q), np.fft.fft2(v[1] * q)
        return self._moved(None, [(su, lambda b: self._fd(v[2], b)),
                                  (sv, lambda b: self._fd(u[2], b)),
                                  (self.spec, lambda b: self._fd(u[2], b) * self._fd(v[2], b))])

    def _fo_lift(self, u):
        """FullObject: the reads' first-order change S(u_O) + Σ_a u_da·∂_a S(O)."""
        r = self._reads(u[0])
        if self.refine:
            for a in range(2):
                d = self.sm.windows(self.obj, (a,))
                r = r + np.stack([u[2][k // self.nmode][a] * d[k]
                                  for k in range(len(self.views))])
        return r

    def _fo_second(self, u, v):
        """FullObject: the reads' second-order change Σ_a (v_da·∂_a S(u_O) + u_da·∂_a S(v_O)) +
        Σ_ac u_da·v_dc·∂_a ∂_c S(O); None with held offsets."""
        if not self.refine:
            return None
        nv = len(self.views)
        cu = [u[2][k // self.nmode] for k in range(nv)]
        cv = [v[2][k // self.nmode] for k in range(nv)]
        out = np.zeros((nv,) + tuple(self.shape), dtype=np.complex128)
        for a in range(2):
            du, dv = self.sm.windows(u[0], (a,)), self.sm.windows(v[0], (a,))
            for k in range(nv):
                out[k] += cv[k][a] * du[k] + cu[k][a] * dv[k]
            for c in range(2):
                dd = self.sm.windows(self.obj, (a, c))
                for k in range(nv):
                    out[k] += cu[k][a] * cv[k][c] * dd[k]
        return out

    def _hessian_fo(self, u, v):
        """The Hessian's exit-plane fields under FullObject: L_u = W·u′ + u_P·q·S(O), L_uv =
        u′·v_P·q + v′·u_P·q + W·o″ (shifted)."""
        q = self.optics.quad
        wu = np.stack([u[1][x[2]] * q for x in self.views])
        wv = np.stack([v[1][x[2]] * q for x in self.views])
        ru, rv = self._fo_lift(u), self._fo_lift(v)
        lu = self.w * ru + wu * self.win
        lv = self.w * rv + wv * self.win
        luv = ru * wv + rv * wu
        o2 = self._fo_second(u, v)
        if o2 is not None:
            luv = luv + self.w * o2
        return [np.stack([to_shifted(p) for p in x]) for x in (lu, lv, luv)]

    def hessian(self, u, v):
        """The second derivative of F along the displacements u = (u_O, u_P[, u_d]) and v,
        pty-chi's bilinear Hessian with A + EPS in place of A: Σ 2·Re(conj(td)·L_uv) +
        2·Σ_unmasked [(1 − d/A)·Σ_m Re(conj(L_u)·L_v) + d·c_u·c_v/A³], L_u = U(u_O·W + O·W_u),
        W_u the waves' first-order change, L_uv = U(u_O·W_v + v_O·W_u + O·W_uv), W_uv their
        second-order change, c_u = Σ_m Re(conj(U ψ)·L_u), d = √iobs.
        ptyrs/src/engine/bh.rs:437-577"""
        fwd = self.optics.exit_forward
        if self.sm is not None:
            lu, lv, luv = [fwd(x) for x in self._hessian_fo(u, v)]
        else:
            wu, wv = self._lift(u), self._lift(v)
            lu = fwd(self._exit(self.obj, self.w, u[0], wu))
            lv = fwd(self._exit(self.obj, self.w, v[0], wv))
            ny, nx = self.shape
            luv = np.empty((len(self.views), ny, nx), dtype=np.complex128)
            w2 = self._second(u, v)
            for k, (_, _, _, o, oy, ox) in enumerate(self.views):
                t = (u[0][o, oy:oy + ny, ox:ox + nx] * wv[k] +
                     v[0][o, oy:oy + ny, ox:ox + nx] * wu[k])
                if w2 is not None:
                    t = t + self.obj[o, oy:oy + ny, ox:ox + nx] * w2[k]
                luv[k] = to_shifted(t)
            luv = fwd(luv)
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


class BhSlices:
    """BhBatch through several object slices: ψ the exit waves of SliceWaves, the object
    variable every slice ([nslice, nobj, nyo, nxo]). The gradient back-propagates χ =
    Uᴴ(2·td) through the gaps; the Hessian's L_u, L_v and L_uv are the tangent and the mixed
    second derivative of the slice model along u and v. With FullObject `offsets` (one per frame
    of the batch) every slice is read sampled at corner plus offset: the slices' gradients are
    placed by the sampling's adjoint and a direction's slice is read through the same windows;
    refined (`refine`), the offsets join the variable: their gradient sums every slice's reads'
    gradient against the reads' derivatives, and the reads' first- and second-order changes
    take the offsets' terms (BhBatch._fo_lift and _fo_second, slice by slice). ptyrs/src/engine/bh.rs:130-577, ptyrs/src/engine/multislice.rs:318-383, the
    sampled slices of ptyrs/src/engine/gpu/bh.rs:266-316"""

    def __init__(self, views, iobs, optics, objects, probe, offsets=None, refine=False):
        from .multislice import SliceWaves
        self.views, self.optics = views, optics
        self.objects, self.probe = objects, probe
        self.frames, self.nmode = frames_of(views)
        self.refine = refine and offsets is not None
        self.sm = None
        if offsets is not None:
            offsets = np.asarray(offsets, dtype=np.float64)
            self.sm = Sampler(objects[0].shape[-2:], probe.shape[-2:], views,
                              [offsets[k // self.nmode] for k in range(len(views))])
        self.sw = SliceWaves(optics, views, list(objects), probe, sampler=self.sm)
        self.psi = optics.exit_forward(self.sw.psi)
        nb = len(self.frames)
        blk = self.psi.reshape((nb, self.nmode) + self.psi.shape[1:])
        self.amp = np.sqrt((blk.real ** 2 + blk.imag ** 2).sum(axis=1)) + EPS
        self.root = np.sqrt(np.stack([iobs[f] for f in self.frames]))
        w = np.where(np.isnan(self.root), 0.0, 1.0 - self.root / self.amp)
        self.td = (blk * w[:, None]).reshape(self.psi.shape)

    def grads(self):
        """(2·∂F/∂O_s* for every slice, 2·∂F/∂P*, None): χ back through the slices, each slice's
        conj(q_s·w_s)·χ_s placed, the probe's conj(q_0·O_0[win])·χ_0 summed over the views."""
        chi = self.optics.exit_adjoint(2.0 * self.td)
        back = self.sw.back_propagate(chi)
        g_obj = np.stack([self.sw.obj_adjoint(s, back[s]) for s in range(len(self.objects))])
        np_ = self.probe.shape[0]
        per = self.sw.probe_adjoint(back[0], np_)
        g_probe = per.reshape(-1, np_, *self.probe.shape[1:]).sum(axis=0)
        g_d = None
        if self.refine:
            sw = self.sw
            g_d = np.zeros((len(self.frames), 2))
            for s in range(len(self.objects)):
                for a in range(2):
                    dw = self.sm.windows(self.objects[s], (a,))
                    for k in range(len(self.views)):
                        gw = to_natural(back[s][k]) * np.conj(sw.quad(s, k) * sw.incident[s][k])
                        g_d[k // self.nmode, a] += float((np.conj(gw) * dw[k]).real.sum())
        return g_obj, g_probe, g_d

    def _reads_lift(self, s, u):
        """Slice s's reads' first-order change along u: its windows of u_O,s, plus with refined
        offsets Σ_a u_da·∂_a S(O_s)."""
        r = self.sw.windows(u[0][s])
        if self.refine:
            for a in range(2):
                d = self.sm.windows(self.objects[s], (a,))
                r = r + np.stack([u[2][k // self.nmode][a] * d[k]
                                  for k in range(len(self.views))])
        return r

    def _reads_second(self, s, u, v):
        """Slice s's reads' second-order change: Σ_a (v_da·∂_a S(u_O,s) + u_da·∂_a S(v_O,s)) +
        Σ_ac u_da·v_dc·∂_a ∂_c S(O_s); None without refined offsets."""
        if not self.refine:
            return None
        nv = len(self.views)
        cu = [u[2][k // self.nmode] for k in range(nv)]
        cv = [v[2][k // self.nmode] for k in range(nv)]
        out = np.zeros((nv,) + tuple(self.probe.shape[-2:]), dtype=np.complex128)
        for a in range(2):
            du, dv = self.sm.windows(u[0][s], (a,)), self.sm.windows(v[0][s], (a,))
            for k in range(nv):
                out[k] += cv[k][a] * du[k] + cu[k][a] * dv[k]
            for c in range(2):
                dd = self.sm.windows(self.objects[s], (a, c))
                for k in range(nv):
                    out[k] += cu[k][a] * cv[k][c] * dd[k]
        return out

    def _tangents(self, u, v):
        """[L_u, L_v, L_uv] in the exit plane before U (natural layout): per slice s the changes
        of the wave leaving it, q_s·(O_s·w_u,s + u_s·in_s), and the mixed one
        q_s·(O_s·w_uv,s + u_s·w_v,s + v_s·w_u,s), carried through the gaps."""
        sw = self.sw
        wu = np.stack([u[1][v_[2]] for v_ in self.views])
        wv = np.stack([v[1][v_[2]] for v_ in self.views])
        wuv = np.zeros_like(wu)
        last = len(self.objects) - 1
        for s in range(last + 1):
            o = sw.objects[s]   # what the views read (under FullObject their windows)
            us, vs = self._reads_lift(s, u), self._reads_lift(s, v)
            ou = sw._times_window(s, o, wu) + sw._times_window(s, us, sw.incident[s])
            ov = sw._times_window(s, o, wv) + sw._times_window(s, vs, sw.incident[s])
            ouv = (sw._times_window(s, o, wuv) + sw._times_window(s, us, wv)
                   + sw._times_window(s, vs, wu))
            o2 = self._reads_second(s, u, v)
            if o2 is not None:
                ouv = ouv + sw._times_window(s, o2, sw.incident[s])
            if s < last:
                gap = self.optics.gaps[s]
                wu, wv, wuv = gap.forward(ou), gap.forward(ov), gap.forward(ouv)
        return ou, ov, ouv

    def hessian(self, u, v):
        """BhBatch.hessian through the slices."""
        fwd = self.optics.exit_forward
        lu, lv, luv = [fwd(np.stack([to_shifted(p) for p in x])) for x in self._tangents(u, v)]
        h = float(2.0 * (np.conj(self.td) * luv).real.sum())
        nb = len(self.frames)
        sh = (nb, self.nmode) + self.psi.shape[1:]
        lu, lv, psi = lu.reshape(sh), lv.reshape(sh), self.psi.reshape(sh)
        uv = (np.conj(lu) * lv).real.sum(axis=1)
        cu = (np.conj(psi) * lu).real.sum(axis=1)
        cv = (np.conj(psi) * lv).real.sum(axis=1)
        keep = ~np.isnan(self.root)
        d, a = self.root[keep], self.amp[keep]
        h += float((2.0 * ((1.0 - d / a) * uv[keep] + d * cu[keep] * cv[keep] / a ** 3)).sum())
        return h


def dot(a, b):
    """Σ Re(conj(a)·b) over the object and probe parts, plus Σ a_d·b_d over the offsets' part
    where both have one. ptyrs/src/engine/bh.rs:79-90, 580-583"""
    s = float(sum((np.conj(x) * y).real.sum() for x, y in zip(a[:2], b[:2])))
    if len(a) > 2 and len(b) > 2 and a[2] is not None and b[2] is not None:
        s += float((np.asarray(a[2]) * np.asarray(b[2])).sum())
    return s


class Bh:
    """BH across epochs: batch size, probe ρ, position ρ, seed, the PATCH offsets of every frame
    (refined or held) and, with one batch of every frame, the previous direction (its offsets'
    part by frame). ptyrs/src/engine/bh.rs:585-799"""

    def __init__(self, batch_size, probe_rho=0.01, seed=7, position_rho=0.1, offsets=None,
                 refine=True, sampling="patch"):
        if batch_size < 1:
            raise ValueError("BH batch_size must be >= 1")
        self.batch_size, self.probe_rho, self.seed = batch_size, probe_rho, seed
        self.position_rho = position_rho
        self.offsets = None if offsets is None else np.array(offsets, dtype=np.float64)
        self.refine = refine and self.offsets is not None
        self.sampling = sampling
        self.epochs = 0
        self.prev = None

    def epoch(self, addr, iobs, optics, obj, probe):
        """One epoch over random_batches(nframe, size, seed, epoch): per batch g = grads, Pg =
        g with the probe part times ρ² = probe_rho²·(the exit's intensity scale) and the
        offsets' part times position_rho², the direction D = −Pg, or with one batch of every
        frame after the first epoch D = −Pg + β·D′, β = H(Pg, D′)/H(D′, D′); α =
        −⟨g, D⟩/H(D, D)·count/size; O += α·D_O, P += α·D_P and refined offsets += 0.3·α·D_d.
        Returns (obj, probe, steps), steps (frames, β, α, top, bottom) per batch; the offsets
        are in self.offsets."""
        nframe, nmode = addr.nframe, addr.nmode
        size = min(self.batch_size, nframe)
        full = size == nframe
        rho2 = self.probe_rho ** 2 * optics.propagator.intensity_scale()
        pos2 = self.position_rho ** 2
        views_all = addr.views()
        epoch = self.epochs
        self.epochs += 1
        steps = []
        obj, probe = obj.copy(), probe.copy()
        for frames in random_batches(nframe, size, self.seed, epoch):
            views = [v for f in frames for v in views_all[f * nmode:(f + 1) * nmode]]
            offs = None if self.offsets is None else self.offsets[frames]
            if optics.nslice() > 1:
                if offs is not None and self.sampling != "full_object":
                    raise ValueError("BH through several slices models FullObject offsets "
                                     "only here")
                b = BhSlices(views, iobs, optics, obj, probe, offs, self.refine)
            else:
                b = BhBatch(views, iobs, optics, obj, probe, offs, self.refine, self.sampling)
            g = b.grads()
            pg = (g[0], g[1] * rho2, None if g[2] is None else g[2] * pos2)
            prev, self.prev = self.prev, None
            if prev is not None and full:
                pd = None if prev[2] is None else prev[2][frames]
                pv = (prev[0], prev[1], pd)
                beta = b.hessian(pg, pv) / b.hessian(pv, pv)
                d = (beta * pv[0] - pg[0], beta * pv[1] - pg[1],
                     None if pd is None else beta * pd - pg[2])
            else:
                beta = 0.0
                d = (-pg[0], -pg[1], None if pg[2] is None else -pg[2])
            top = -dot(g, d)
            bottom = b.hessian(d, d)
            alpha = top / bottom * len(frames) / size
            obj = obj + alpha * d[0]
            probe = probe + alpha * d[1]
            by_frame = None
            if self.refine:
                by_frame = np.zeros((nframe, 2))
                by_frame[frames] = d[2]
                self.offsets[frames] += POSITION_STEP * alpha * d[2]
            if full:
                self.prev = (d[0], d[1], by_frame)
            steps.append((frames, beta, alpha, top, bottom))
        return obj, probe, steps
