"""PyNX operators and the AP, DM, RAAR, ML and rPIE loops of the reference's CPU engine in
FP64. Of the features only the position offsets are modelled: under PATCH sampling (the probe
moved by each frame's sub-pixel offset in Fourier space) and under FullObject sampling (each
view's window Fourier-interpolated at its corner plus the offset, `sampling.py`; AP's, DM's
and RAAR's object by the PCG solve `solve_obj`) in the LLK and every algorithm. No scale or
background.

Test infrastructure (see oracle/__init__.py). Layout: exit waves `psi[k]` shifted, `k` in
table order; objects `[nobj, nyo, nxo]`, probes `[nprobe, ny, nx]`, natural layout.
"""

import math

import numpy as np

from .base import AddrTable, observed, pynx_preprocess, random_batches, to_natural, to_shifted
from .optics import Optics, fft2, freq
from .positions import (EPS_POSITION, batch_position_step, derivative_sums, gradient_directions,
                        gradient_sums, move_planes, position_step, ramp)
from .sampling import Sampler, auto_sampling, view_offsets

AP, DM, RAAR, RPIE, ML, BH, LSQML = "ap", "dm", "raar", "rpie", "ml", "bh", "lsqml"


def obj_probe_to_psi(addr, views, obj, probe, quad):
    """`psi[k] = fftshift(P[probe_k]·O[obj_k, win_k]·q)`. ptyrs/src/engine/ops.rs:42-71"""
    ny, nx = probe.shape[-2:]
    out = np.empty((len(views), ny, nx), dtype=np.complex128)
    for k, (_, _, p, o, oy, ox) in enumerate(views):
        out[k] = to_shifted(probe[p] * obj[o, oy:oy + ny, ox:ox + nx] * quad)
    return out


def _mode_sum(psi, nmode):
    n = psi.shape[0] // nmode
    a = psi.reshape(n, nmode, *psi.shape[1:])
    return (a.real ** 2 + a.imag ** 2).sum(axis=1)


def apply_amplitude(iobs, nmode, psi):
    """`psi *= sqrt(iobs)/max(1e-12, sqrt(Σ_m|psi|²))` at unmasked pixels.
    ptyrs/src/engine/ops.rs:73-107"""
    calc = np.sqrt(_mode_sum(psi, nmode))
    r = np.sqrt(iobs) / np.maximum(calc, 1e-12)
    r = np.where(np.isnan(iobs), 1.0, r)
    n = psi.shape[0] // nmode
    return (psi.reshape(n, nmode, *psi.shape[1:]) * r[:, None]).reshape(psi.shape)


class Llk:
    """Poisson, Gaussian, Euclidian sums and `nb_photons_calc`. ptyrs/src/engine/ops.rs:255-277"""

    def __init__(self, poisson=0.0, gaussian=0.0, euclidian=0.0, nb_photons_calc=0.0):
        self.poisson, self.gaussian = poisson, gaussian
        self.euclidian, self.nb_photons_calc = euclidian, nb_photons_calc

    def normalised(self, nb_obs):
        n = float(nb_obs)
        return Llk(self.poisson / n, self.gaussian / n, self.euclidian / n, self.nb_photons_calc)

    def as_dict(self):
        return dict(poisson=self.poisson, gaussian=self.gaussian, euclidian=self.euclidian,
                    nb_photons_calc=self.nb_photons_calc)


def llk(iobs, nmode, psi):
    """LLK sums of detector-plane psi; masked pixels excluded. ptyrs/src/engine/ops.rs:109-154"""
    icalc = _mode_sum(psi, nmode)
    keep = ~np.isnan(iobs)
    io, ic = iobs[keep], icalc[keep]
    pos = io > 0.0
    safe = np.where(pos, io, 1.0)
    pois = np.where(pos, ic - io + io * np.log(safe / np.where(pos, ic, 1.0)), ic)
    return Llk(float(pois.sum()), float(((io - ic) ** 2 / (io + 1.0)).sum()),
               float((4.0 * (np.sqrt(io) - np.sqrt(ic)) ** 2).sum()), float(ic.sum()))


def poisson_grad(iobs, nmode, psi):
    """`psi *= 2(1 − iobs/max(1e-12, icalc))`, masked pixels 0. ptyrs/src/engine/ops.rs:156-195"""
    icalc = _mode_sum(psi, nmode)
    r = 2.0 * (1.0 - iobs / np.maximum(icalc, 1e-12))
    r = np.where(np.isnan(iobs), 0.0, r)
    n = psi.shape[0] // nmode
    return (psi.reshape(n, nmode, *psi.shape[1:]) * r[:, None]).reshape(psi.shape)


def cg_poisson_gamma(iobs, nmode, po, a, b):
    """Numerator and denominator of the Newton step along a direction.
    ptyrs/src/engine/ops.rs:197-253"""
    n = po.shape[0] // nmode
    sh = (n, nmode) + po.shape[1:]
    po, a, b = po.reshape(sh), a.reshape(sh), b.reshape(sh)
    calc = (po.real ** 2 + po.imag ** 2).sum(axis=1)
    r = (np.conj(po) * a).real.sum(axis=1)
    s = (a.real ** 2 + a.imag ** 2 + 2.0 * (np.conj(po) * b).real).sum(axis=1)
    keep = ~np.isnan(iobs)
    calc = np.maximum(calc[keep], 1e-32)
    r, s, io = r[keep], s[keep], iobs[keep]
    d = 1.0 - io / calc
    return float((-2.0 * r * d).sum()), float((2.0 * s * d + io * (2.0 * r / calc) ** 2).sum())


def accumulate_obj(views, obj_shape, probe, quad, psi):
    """Object back-projection over views in order. ptyrs/src/engine/ops.rs:279-310"""
    ny, nx = probe.shape[-2:]
    num = np.zeros(obj_shape, dtype=np.complex128)
    norm = np.zeros(obj_shape, dtype=np.float64)
    for k, (_, _, p, o, oy, ox) in enumerate(views):
        w = to_natural(psi[k])
        num[o, oy:oy + ny, ox:ox + nx] += w * np.conj(quad * probe[p])
        norm[o, oy:oy + ny, ox:ox + nx] += probe[p].real ** 2 + probe[p].imag ** 2
    return num, norm


def accumulate_probe(views, obj, probe_shape, quad, psi):
    """Probe back-projection over views in order. ptyrs/src/engine/ops.rs:312-343"""
    ny, nx = probe_shape[-2:]
    num = np.zeros(probe_shape, dtype=np.complex128)
    norm = np.zeros(probe_shape, dtype=np.float64)
    for k, (_, _, p, o, oy, ox) in enumerate(views):
        w = to_natural(psi[k])
        ob = obj[o, oy:oy + ny, ox:ox + nx]
        num[p] += w * np.conj(quad * ob)
        norm[p] += ob.real ** 2 + ob.imag ** 2
    return num, norm


def blend(old, num, norm, inertia):
    """`(num + reg·old)/max(norm + reg, 1e-12)`, `reg = inertia·max(norm)`.
    ptyrs/src/engine/ops.rs:367-381"""
    reg = max(float(norm.max()), 0.0) * inertia
    return (num + old * reg) / np.maximum(norm + reg, 1e-12)


def psi_to_obj_probe(views, quad, psi, obj, probe, obj_inertia, probe_inertia):
    """Psi2ObjProbe, both sums from the arrays on entry. ptyrs/src/engine/ops.rs:345-365"""
    obj_acc = None
    if obj_inertia is not None:
        obj_acc = accumulate_obj(views, obj.shape, probe, quad, psi)
    if probe_inertia is not None:
        num, norm = accumulate_probe(views, obj, probe.shape, quad, psi)
        probe = blend(probe, num, norm, probe_inertia)
    if obj_inertia is not None:
        obj = blend(obj, obj_acc[0], obj_acc[1], obj_inertia)
    return obj, probe


OBJ_PCG = 10      # ptyrs/src/engine/sampling.rs:214
OBJ_HELD = 1e-5   # ptyrs/src/engine/sampling.rs:221


def solve_obj(sm, w, psi_n, obj, inertia):
    """Under FullObject the object of min ‖psi − A·O'‖² + reg·‖O' − O‖², A the views' waves
    `w` (natural) times their reads through the sampler `sm`: OBJ_PCG preconditioned
    conjugate-gradient steps from `obj`, preconditioner D + reg with D = sm.diag(|w|²) and
    reg = inertia·max D, pixels with D + reg ≤ OBJ_HELD·max D held (a zero share of every
    step); a step along no positive curvature, or after a residual of no positive norm, is 0.
    ptyrs/src/engine/features.rs:726-793"""
    nobj = obj.shape[0]
    d = sm.diag(w.real ** 2 + w.imag ** 2, nobj)
    mx = float(d.max())
    reg = mx * inertia
    pre = np.where(d + reg > OBJ_HELD * mx, d + reg, np.inf)

    def normal(x):
        return sm.place(np.conj(w) * w * sm.windows(x), nobj)

    r = sm.place(np.conj(w) * psi_n, nobj) - normal(obj)
    rz = float((np.abs(r) ** 2 / pre).sum())
    p = r / pre
    x = obj.copy()
    for step in range(OBJ_PCG):
        ap = normal(p) + p * reg
        pap = float((np.conj(p) * ap).real.sum())
        alpha = rz / pap if pap > 0.0 else 0.0
        x = x + p * alpha
        r = r - ap * alpha
        if step + 1 < OBJ_PCG:
            nxt = float((np.abs(r) ** 2 / pre).sum())
            beta = nxt / rz if rz > 0.0 else 0.0
            p = r / pre + p * beta
            rz = nxt
    return x


def llk_due(calc_llk, ic, ncycle):
    """ptyrs/src/engine/ptycho.rs:43-46"""
    return calc_llk > 0 and (ic % calc_llk == 0 or ic + 1 == ncycle)


def zero_phase_ramp(obj, probe):
    """PyNX ZeroPhaseRamp. ptyrs/src/engine/ptycho.rs:593-603, :1132-1181"""
    ny, nx = probe.shape[-2:]
    far = fft2(to_shifted(probe))
    v_nat = to_natural((far.real ** 2 + far.imag ** 2).sum(axis=0))
    w = v_nat.sum()
    yy, xx = np.mgrid[:ny, :nx]
    dy = (v_nat * yy).sum() / w - (ny // 2)
    dx = (v_nat * xx).sum() / w - (nx // 2)

    def ramp(length, half, scale, d, sign):
        return sign * 2.0 * math.pi * d * (np.arange(length) - half) / scale

    def apply(a, ry, rx):
        ph = ry[:, None] + rx[None, :]
        return a * (np.cos(ph) + 1j * np.sin(ph))

    probe = apply(probe, ramp(ny, ny // 2, ny, dy, -1.0), ramp(nx, nx // 2, nx, dx, -1.0))
    nyo, nxo = obj.shape[-2:]
    obj = apply(obj, ramp(nyo, nyo // 2, ny, dy, 1.0), ramp(nxo, nxo // 2, nx, dx, 1.0))
    return obj, probe


class Ptycho:
    """Reconstruction state: object, probe, carried exit waves, LLK history.
    ptyrs/src/engine/ptycho.rs:188-285 (no features)"""

    def __init__(self, counts, y, x, obj, probe, optics):
        probe = np.array(probe, dtype=np.complex128)
        obj = np.array(obj, dtype=np.complex128)
        frame = list(np.asarray(counts).shape[-2:])
        if list(probe.shape[-2:]) != frame:
            raise ValueError("probe plane differs from frame")
        if optics.nslice() != 1:
            raise ValueError("Ptycho runs one object slice")
        if not optics.propagator.adjoint_inverts():
            raise ValueError("Ptycho needs a unitary propagator")
        py_, px_ = pynx_preprocess(y, x, optics.pixel)
        self.addr = AddrTable(py_, px_, optics.pixel, frame, obj.shape[-2:],
                              probe.shape[0], obj.shape[0])
        self.optics = optics
        self.iobs, self.nb_obs = observed(counts)
        self.obj, self.probe = obj, probe
        self.views_all = self.addr.views()
        self.psi = np.zeros((len(self.views_all),) + tuple(frame), dtype=np.complex128)
        self.cycle = 0
        self.history = []
        self.offsets = None   # Features::position: [nframe, 2] object pixels
        self.refine = False
        # Sampling::auto (ptyrs/src/engine/ptycho.rs:271): "patch" or "full_object"
        self.sampling = auto_sampling(optics.propagator.kind, obj.shape[-2:], frame)

    def set_sampling(self, sampling):
        """How the offsets enter the exit waves. ptyrs/src/engine/ptycho.rs:289-295"""
        if sampling not in ("patch", "full_object"):
            raise ValueError(sampling)
        self.sampling = sampling

    def _sampler(self, views):
        """The FullObject sampling of `views` at their frames' offsets, None under PATCH."""
        if self.offsets is None or self.sampling != "full_object":
            return None
        return Sampler(self.obj.shape[-2:], self.probe.shape[-2:], views,
                       view_offsets(views, self.offsets))

    def _patch_only(self, what):
        if self.offsets is not None and self.sampling == "full_object":
            raise ValueError(what + " under FullObject sampling is not modelled")

    def set_position(self, offsets, refine=True):
        """Features::position (ptyrs/src/engine/features.rs:304-330): per-frame [dy, dx]
        offsets in object pixels, refined by rPIE's steps unless held; None for none."""
        self.offsets = None if offsets is None else np.array(offsets, dtype=np.float64)
        self.refine = bool(refine) and offsets is not None

    def moved_waves(self, views):
        """Under the offsets, every view's wave W = DFT⁻¹(DFT(P·q)·exp(−2πi f·d)): its probe
        mode times the plane phase translated by its frame's offset on the plane's periodic
        grid, natural layout (`Shifted` under PATCH, ptyrs/src/engine/features.rs:545-640);
        under FullObject the unmoved P·q (features.rs:609-615)."""
        shape = self.probe.shape[-2:]
        if self.sampling == "full_object":
            return np.stack([self.probe[v[2]] * self.quad for v in views])
        spec = np.fft.fft2(self.probe * self.quad)
        return np.stack([np.fft.ifft2(spec[v[2]] * ramp(shape, self.offsets[v[0]]))
                         for v in views])

    @property
    def quad(self):
        return self.optics.quad

    def calc_psi(self):
        if self.offsets is not None:
            return self._moved_psi(self.views_all)[1]
        return obj_probe_to_psi(self.addr, self.views_all, self.obj, self.probe, self.quad)

    def _moved_psi(self, views):
        """The moved waves W and the exit waves fftshift(W·O_win) of `views` (under FullObject
        the sampled windows)."""
        ny, nx = self.probe.shape[-2:]
        w = self.moved_waves(views)
        sm = self._sampler(views)
        if sm is not None:
            win = sm.windows(self.obj)
        else:
            win = np.stack([self.obj[o, oy:oy + ny, ox:ox + nx]
                            for (_, _, _, o, oy, ox) in views])
        return w, to_shifted(w * win)

    def llk(self):
        """ptyrs/src/engine/ptycho.rs:452-466"""
        psi = self.optics.exit_forward(self.calc_psi())
        return llk(self.iobs, self.addr.nmode, psi).normalised(self.nb_obs)

    def _record(self, algorithm, value):
        self.history.append((self.cycle, algorithm, value))

    def grad_psi(self):
        """ptyrs/src/engine/ptycho.rs:547-561"""
        psi = self.optics.exit_forward(self.calc_psi())
        psi = poisson_grad(self.iobs, self.addr.nmode, psi)
        return self.optics.exit_adjoint(psi)

    def grads(self, obj, probe):
        """2·∂L/∂O* and 2·∂L/∂P*; under the offsets the back-projections through the moved
        waves (`Shifted::accumulate_obj` / `accumulate_probe`,
        ptyrs/src/engine/features.rs:865-958). ptyrs/src/engine/ptycho.rs:492-513"""
        g = self.grad_psi()
        sm = self._sampler(self.views_all)
        if sm is not None:
            return self._sampled_grads(sm, g, obj, probe)
        if self.offsets is not None:
            return self._moved_grads(g, obj, probe)
        go = (accumulate_obj(self.views_all, self.obj.shape, self.probe, self.quad, g)[0]
              if obj else np.zeros_like(self.obj))
        gp = (accumulate_probe(self.views_all, self.obj, self.probe.shape, self.quad, g)[0]
              if probe else np.zeros_like(self.probe))
        return go, gp

    def _sampled_grads(self, sm, g, obj, probe):
        """Under FullObject: the object's Σ g·conj(W) on each view's read placed by the
        sampling's adjoint, the probe's Σ conj(S(O)·q)·g (W = P·q unmoved).
        ptyrs/src/engine/features.rs:865-958"""
        views = self.views_all
        gn = to_natural(g)
        q = self.quad
        w = np.stack([self.probe[v[2]] * q for v in views])
        go = sm.place(gn * np.conj(w), self.obj.shape[0]) if obj else np.zeros_like(self.obj)
        gp = np.zeros_like(self.probe)
        if probe:
            win = sm.windows(self.obj)
            for k, v in enumerate(views):
                gp[v[2]] += gn[k] * np.conj(win[k] * q)
        return go, gp

    def _moved_grads(self, g, obj, probe):
        """Object: Σ g·conj(W) on the windows. Probe: per frame Y = Σ conj(O_win)·g over its
        views of each mode, moved back by the frame's offset in Fourier space and summed,
        times conj(q)."""
        ny, nx = self.probe.shape[-2:]
        views = self.views_all
        w = self.moved_waves(views)
        gn = to_natural(g)
        go = np.zeros_like(self.obj)
        gp = np.zeros_like(self.probe)
        if obj:
            for k, (_, _, _, o, oy, ox) in enumerate(views):
                go[o, oy:oy + ny, ox:ox + nx] += gn[k] * np.conj(w[k])
        if probe:
            spec = np.zeros_like(self.probe)
            nmode = self.addr.nmode
            for f in range(self.addr.nframe):
                y = np.zeros_like(self.probe)
                for k in range(f * nmode, (f + 1) * nmode):
                    _, _, pm, o, oy, ox = views[k]
                    y[pm] += np.conj(self.obj[o, oy:oy + ny, ox:ox + nx]) * gn[k]
                spec += np.fft.fft2(y) * ramp((ny, nx), -self.offsets[f])
            gp = np.conj(self.quad) * np.fft.ifft2(spec)
        return go, gp

    def cg_gamma(self, d_obj, d_probe):
        """ptyrs/src/engine/ptycho.rs:515-545; under the offsets the probe and d_probe terms
        are their moved waves."""
        v, q = self.views_all, self.quad
        sm = self._sampler(v)
        if sm is not None:
            w = np.stack([self.probe[x[2]] * q for x in v])
            dw = np.stack([d_probe[x[2]] * q for x in v])
            win, dwin = sm.windows(self.obj), sm.windows(d_obj)
            po, a, b = (to_shifted(w * win), to_shifted(w * dwin + dw * win),
                        to_shifted(dw * dwin))
        elif self.offsets is not None:
            ny, nx = self.probe.shape[-2:]
            w = self.moved_waves(v)
            keep = self.probe
            self.probe = d_probe
            dw = self.moved_waves(v)
            self.probe = keep
            win = lambda x: np.stack([x[o, oy:oy + ny, ox:ox + nx] for (_, _, _, o, oy, ox) in v])
            po = to_shifted(w * win(self.obj))
            a = to_shifted(w * win(d_obj) + dw * win(self.obj))
            b = to_shifted(dw * win(d_obj))
        else:
            po = obj_probe_to_psi(self.addr, v, self.obj, self.probe, q)
            a = (obj_probe_to_psi(self.addr, v, d_obj, self.probe, q)
                 + obj_probe_to_psi(self.addr, v, self.obj, d_probe, q))
            b = obj_probe_to_psi(self.addr, v, d_obj, d_probe, q)
        po, a, b = (self.optics.exit_forward(x) for x in (po, a, b))
        return cg_poisson_gamma(self.iobs, self.addr.nmode, po, a, b)

    def projected_position_sums(self):
        """Per frame [num, den] of the position step from the magnitude projection of the
        current estimate's exit waves (`position_sums` of them).
        ptyrs/src/engine/ptycho.rs:468-479"""
        x = self.optics.exit_forward(self._moved_psi(self.views_all)[1])
        x = apply_amplitude(self.iobs, self.addr.nmode, x)
        return self.position_sums(self.optics.exit_adjoint(x))

    def position_sums(self, x):
        """Per frame [num, den] of the position step of new exit waves x (shifted layout):
        num_a = Re Σ conj(x − W·O)·O·∂_a W, den_a = Σ |O·∂_a W|² over the frame's views,
        ∂_a W = DFT⁻¹(DFT(P·q)·exp(−2πi f·d)·(−2πi f_a)), W and O those of the current arrays.
        ptyrs/src/engine/features.rs:960-1025"""
        ny, nx = self.probe.shape[-2:]
        views = self.views_all
        w = self.moved_waves(views)
        x = to_natural(x)
        sm = self._sampler(views)
        if sm is not None:
            # FullObject: ∂psi = W·∂_a S(O), the sampled windows' derivatives
            # (ptyrs/src/engine/features.rs:978-985)
            win = sm.windows(self.obj)
            dwin = [sm.windows(self.obj, (a,)) for a in range(2)]
            sums = np.zeros((self.addr.nframe, 2, 2))
            for k, v in enumerate(views):
                c = x[k] - w[k] * win[k]
                for axis in range(2):
                    dpsi = w[k] * dwin[axis][k]
                    sums[v[0], 0, axis] += float((np.conj(c) * dpsi).real.sum())
                    sums[v[0], 1, axis] += float((dpsi.real ** 2 + dpsi.imag ** 2).sum())
            return sums
        spec = np.fft.fft2(self.probe * self.quad)
        fy = freq(np.arange(ny), ny)[:, None] * np.ones((1, nx))
        fx = np.ones((ny, 1)) * freq(np.arange(nx), nx)[None, :]
        sums = np.zeros((self.addr.nframe, 2, 2))
        for k, (f, _, pm, o, oy, ox) in enumerate(views):
            win = self.obj[o, oy:oy + ny, ox:ox + nx]
            r = spec[pm] * ramp((ny, nx), self.offsets[f])
            c = x[k] - w[k] * win
            for axis, fa in enumerate((fy, fx)):
                dpsi = win * np.fft.ifft2(r * (-2j * math.pi * fa))
                sums[f, 0, axis] += float((np.conj(c) * dpsi).real.sum())
                sums[f, 1, axis] += float((dpsi.real ** 2 + dpsi.imag ** 2).sum())
        return sums

    def scale_obj_probe(self):
        """PyNX ScaleObjProbe(absolute=True). ptyrs/src/engine/ptycho.rs:571-581"""
        calc = self.llk().nb_photons_calc
        total = float(np.nansum(self.iobs))
        self.split_scale(math.sqrt(total / calc))

    def split_scale(self, s):
        """ptyrs/src/engine/ptycho.rs:583-591"""
        os_ = float(np.abs(self.obj).sum())
        ps = float(np.abs(self.probe).sum())
        fp, fo = math.sqrt(os_ / ps * s), math.sqrt(ps / os_ * s)
        self.probe = self.probe * fp
        self.obj = self.obj * fo

    def zero_phase_ramp(self):
        self.obj, self.probe = zero_phase_ramp(self.obj, self.probe)

    def _update_from_psi(self, obj_inertia, probe_inertia):
        """psi_to_obj_probe of self.psi, through the moved waves under the offsets.
        ptyrs/src/engine/ptycho.rs:387-410"""
        if self.offsets is not None and self.sampling == "full_object":
            self._sampled_psi_to_obj_probe(obj_inertia, probe_inertia)
        elif self.offsets is not None:
            self._moved_psi_to_obj_probe(obj_inertia, probe_inertia)
        else:
            self.obj, self.probe = psi_to_obj_probe(self.views_all, self.quad, self.psi,
                                                    self.obj, self.probe, obj_inertia,
                                                    probe_inertia)

    def _sampled_psi_to_obj_probe(self, obj_inertia, probe_inertia):
        """psi_to_obj_probe under FullObject, both from the arrays on entry: the probe's first
        conjugate-gradient step through the sampled windows S(O) (the waves are not moved, so
        it is `blend` to rounding), then the object solve (`solve_obj`): OBJ_PCG = 10
        conjugate-gradient steps on (AᴴA + reg)·O' = Aᴴ·psi + reg·O from O, A the views' waves
        times their reads, preconditioned by D + reg (D the exact diagonal of AᴴA,
        Sampler.diag; reg = inertia·max D), pixels with D + reg ≤ 1e-5·max D held.
        ptyrs/src/engine/features.rs:653-793"""
        views = self.views_all
        sm = self._sampler(views)
        q = self.quad
        psi_n = to_natural(self.psi)
        w = np.stack([self.probe[v[2]] * q for v in views])
        win = sm.windows(self.obj)
        new_probe = self.probe
        if probe_inertia is not None:
            resid = psi_n - w * win
            num = np.zeros_like(self.probe)
            norm = np.zeros(self.probe.shape)
            for k, v in enumerate(views):
                num[v[2]] += np.conj(win[k] * q) * resid[k]
                norm[v[2]] += win[k].real ** 2 + win[k].imag ** 2
            reg = float(norm.max()) * probe_inertia
            d = np.maximum(norm + reg, 1e-12)
            slope = [float((np.abs(num[m]) ** 2 / d[m]).sum()) for m in range(num.shape[0])]
            step = num / d
            curv = [reg * float((np.abs(step[m]) ** 2).sum()) for m in range(num.shape[0])]
            for k, v in enumerate(views):
                curv[v[2]] += float((np.abs(step[v[2]] * q * win[k]) ** 2).sum())
            new_probe = self.probe.copy()
            for m in range(num.shape[0]):
                if slope[m] > 0.0:
                    new_probe[m] = new_probe[m] + step[m] * (slope[m] / curv[m])
        if obj_inertia is not None:
            self.obj = solve_obj(sm, w, psi_n, self.obj, obj_inertia)
        self.probe = new_probe

    def _moved_psi_to_obj_probe(self, obj_inertia, probe_inertia):
        """psi_to_obj_probe under PATCH offsets, both from the arrays on entry: the probe takes
        the first conjugate-gradient step of min ‖psi − A·P'‖² + reg·‖P' − P‖² preconditioned by
        Σ|O|² + reg (A the moved-wave model, reg = inertia·max Σ|O|²): s = g/D with g the
        residual back-projected through the moves, P += (<g, s>/(‖A·s‖² + reg·‖s‖²))·s per mode
        where <g, s> > 0; the object blends Σ psi·conj(W) and Σ|W|² with the waves of the probe
        on entry. ptyrs/src/engine/features.rs:653-720, :851-958"""
        ny, nx = self.probe.shape[-2:]
        views = self.views_all
        psi_n = to_natural(self.psi)
        w = self.moved_waves(views)
        win = np.stack([self.obj[o, oy:oy + ny, ox:ox + nx] for (_, _, _, o, oy, ox) in views])
        new_probe = self.probe
        if probe_inertia is not None:
            resid = psi_n - w * win
            nmode = self.addr.nmode
            spec = np.zeros_like(self.probe)
            norm = np.zeros(self.probe.shape)
            for f in range(self.addr.nframe):
                y = np.zeros_like(self.probe)
                for k in range(f * nmode, (f + 1) * nmode):
                    pm = views[k][2]
                    y[pm] += np.conj(win[k]) * resid[k]
                    norm[pm] += win[k].real ** 2 + win[k].imag ** 2
                spec += np.fft.fft2(y) * ramp((ny, nx), -self.offsets[f])
            num = np.conj(self.quad) * np.fft.ifft2(spec)
            reg = float(norm.max()) * probe_inertia
            d = np.maximum(norm + reg, 1e-12)
            slope = [float((np.abs(num[m]) ** 2 / d[m]).sum()) for m in range(num.shape[0])]
            step = num / d
            curv = [reg * float((np.abs(step[m]) ** 2).sum()) for m in range(num.shape[0])]
            keep = self.probe
            self.probe = step
            ws = self.moved_waves(views)
            self.probe = keep
            for k, v in enumerate(views):
                curv[v[2]] += float((np.abs(ws[k] * win[k]) ** 2).sum())
            new_probe = self.probe.copy()
            for m in range(num.shape[0]):
                if slope[m] > 0.0:
                    new_probe[m] = new_probe[m] + step[m] * (slope[m] / curv[m])
        if obj_inertia is not None:
            num = np.zeros_like(self.obj)
            norm = np.zeros(self.obj.shape)
            for k, (_, _, _, o, oy, ox) in enumerate(views):
                num[o, oy:oy + ny, ox:ox + nx] += psi_n[k] * np.conj(w[k])
                norm[o, oy:oy + ny, ox:ox + nx] += w[k].real ** 2 + w[k].imag ** 2
            self.obj = blend(self.obj, num, norm, obj_inertia)
        self.probe = new_probe

    def ap(self, ncycle, obj_inertia=0.01, probe_inertia=None, calc_llk=0,
           zero_phase_ramp=True):
        """PyNX AP; under refined offsets the position sums of the projection, then the
        update, then the position step. ptyrs/src/engine/ptycho.rs:605-641"""
        nmode = self.addr.nmode
        for ic in range(ncycle):
            self.psi = self.optics.exit_forward(self.calc_psi())
            rec = None
            if llk_due(calc_llk, ic, ncycle):
                rec = llk(self.iobs, nmode, self.psi).normalised(self.nb_obs)
            self.psi = apply_amplitude(self.iobs, nmode, self.psi)
            self.psi = self.optics.exit_adjoint(self.psi)
            sums = self.position_sums(self.psi) if self.refine else None
            self._update_from_psi(obj_inertia, probe_inertia)
            if sums is not None:
                position_step(self.offsets, sums)
            if rec is not None:
                self._record(AP, rec)
            self.cycle += 1
        if zero_phase_ramp:
            self.zero_phase_ramp()

    def _update_obj_probe(self, inertia, loops):
        """DM's and RAAR's update: the position sums while refined, `loops` times an object-only
        then a probe-only psi_to_obj_probe, then the position step.
        ptyrs/src/engine/ptycho.rs:733-755"""
        sums = self.position_sums(self.psi) if self.refine else None
        for _ in range(loops):
            for o, p in ((inertia[0], None), (None, inertia[1])):
                if o is not None or p is not None:
                    self._update_from_psi(o, p)
        if sums is not None:
            position_step(self.offsets, sums)

    def dm(self, ncycle, obj_inertia=0.01, probe_inertia=0.001, alpha=0.02, loop_obj_probe=1,
           calc_llk=0, zero_phase_ramp=True):
        """PyNX DM. ptyrs/src/engine/ptycho.rs:643-691"""
        nmode, a = self.addr.nmode, alpha
        refl = None
        for ic in range(ncycle):
            rec = self.llk() if llk_due(calc_llk, ic, ncycle) else None
            po = self.calc_psi()
            if ic == 0:
                self.psi = po.copy()
            else:
                self.psi = po * (2.0 - a) - self.psi * (1.0 - a)
                refl = self.psi.copy()
            psi = self.optics.exit_forward(self.psi)
            psi = apply_amplitude(self.iobs, nmode, psi)
            self.psi = self.optics.exit_adjoint(psi)
            if ic > 0:
                self.psi = self.psi + (po - refl) * (1.0 - a)
            self._update_obj_probe((obj_inertia, probe_inertia), loop_obj_probe)
            if rec is not None:
                self._record(DM, rec)
                self.split_scale(1.0)
            self.cycle += 1
        if zero_phase_ramp:
            self.zero_phase_ramp()

    def raar(self, ncycle, obj_inertia=0.01, probe_inertia=0.001, beta=0.9, loop_obj_probe=1,
             calc_llk=0, zero_phase_ramp=True):
        """PyNX RAAR. ptyrs/src/engine/ptycho.rs:693-731"""
        nmode, h = self.addr.nmode, 0.5 * beta
        iterate = None
        for ic in range(ncycle):
            rec = self.llk() if llk_due(calc_llk, ic, ncycle) else None
            po = self.calc_psi()
            if ic == 0:
                self.psi = po.copy()
            else:
                self.psi = self.psi * (1.0 - h) + (po * 2.0 - iterate) * h
            iterate = self.psi.copy()
            psi = self.optics.exit_forward(self.psi)
            psi = apply_amplitude(self.iobs, nmode, psi)
            self.psi = self.optics.exit_adjoint(psi)
            self._update_obj_probe((obj_inertia, probe_inertia), loop_obj_probe)
            if rec is not None:
                self._record(RAAR, rec)
                self.split_scale(1.0)
            self.cycle += 1
        if zero_phase_ramp:
            self.zero_phase_ramp()

    def ml(self, ncycle, update_object=True, update_probe=False, calc_llk=0,
           zero_phase_ramp=True):
        """PyNX ML Poisson conjugate gradient. ptyrs/src/engine/ptycho.rs:757-842"""
        self.scale_obj_probe()
        last = None
        d = [np.zeros_like(self.obj), np.zeros_like(self.probe)]
        for ic in range(ncycle):
            rec = self.llk() if llk_due(calc_llk, ic, ncycle) else None
            sums = self.projected_position_sums() if self.refine else None
            g = self.grads(update_object, update_probe)
            beta = 0.0
            if last is not None:
                num = sum(float((np.conj(gi) * (gi - hi)).real.sum()) for gi, hi in zip(g, last))
                den = sum(float((hi.real ** 2 + hi.imag ** 2).sum()) for hi in last)
                b = num / den if den != 0 else math.nan
                beta = b if (math.isfinite(b) and b >= 0.0) else 0.0
            d = [di * beta - gi for di, gi in zip(d, g)]
            num, den = self.cg_gamma(d[0], d[1])
            gamma = num / den if den != 0 else math.inf
            if math.isfinite(gamma):
                self.obj = self.obj + d[0] * gamma
                self.probe = self.probe + d[1] * gamma
            else:
                d = [-gi for gi in g]
            last = g
            if sums is not None:
                position_step(self.offsets, sums)
            if rec is not None:
                self._record(ML, rec)
            self.cycle += 1
        if zero_phase_ramp:
            self.zero_phase_ramp()

    def rpie(self, nepoch, obj_step=1.0, probe_step=1.0, obj_alpha=0.1, probe_alpha=0.1,
             batch_size=100, calc_llk=0, seed=7):
        """rPIE after pty-chi with the overlap-normalised object step.
        ptyrs/src/engine/ptycho.rs:844-898"""
        if self.addr.nobj != 1:
            raise ValueError("rPIE needs exactly one object mode")
        if batch_size <= 0:
            raise ValueError("rPIE batch_size must be >= 1")
        for ic in range(nepoch):
            for frames in random_batches(self.addr.nframe, batch_size, seed, self.cycle):
                self.rpie_batch(frames, obj_step, probe_step, obj_alpha, probe_alpha)
            if llk_due(calc_llk, ic, nepoch):
                self._record(RPIE, self.llk())
            self.cycle += 1

    def rpie_delta(self, frames):
        """The batch's `psi` and `Δ = IFT(ψ') − ψ` (shifted layout).
        ptyrs/src/engine/ptycho.rs:900-945"""
        nmode = self.addr.nmode
        views = [v for f in frames for v in self.views_all[f * nmode:(f + 1) * nmode]]
        if self.offsets is not None:
            psi = self._moved_psi(views)[1]
        else:
            psi = obj_probe_to_psi(self.addr, views, self.obj, self.probe, self.quad)
        delta = self.optics.exit_forward(psi)
        io = self.iobs[frames]
        calc = np.sqrt(_mode_sum(delta, nmode))
        r = np.sqrt(io + 1e-7) / (calc + 1e-7)
        r = np.where(np.isnan(io), 1.0, r)
        nb = len(frames)
        delta = (delta.reshape(nb, nmode, *delta.shape[1:]) * r[:, None]).reshape(delta.shape)
        delta = self.optics.exit_adjoint(delta) - psi
        return views, psi, delta

    def rpie_batch(self, frames, obj_step, probe_step, obj_alpha, probe_alpha):
        """One rPIE batch from the batch-start arrays. ptyrs/src/engine/ptycho.rs:900-1086, :1089-1130"""
        if self.offsets is not None and self.sampling == "full_object":
            return self._rpie_batch_sampled(frames, obj_step, probe_step, obj_alpha, probe_alpha)
        if self.offsets is not None:
            return self._rpie_batch_moved(frames, obj_step, probe_step, obj_alpha, probe_alpha)
        ny, nx = self.probe.shape[-2:]
        nmode = self.addr.nmode
        nprobe = self.probe.shape[0]
        nbatch = len(frames)
        views, psi, delta = self.rpie_delta(frames)
        delta_nat = to_natural(delta)
        w_nat = self.probe * self.quad  # W_m = P_m q, natural
        w2_nat = self.probe.real ** 2 + self.probe.imag ** 2
        d_obj = None
        if obj_step is not None:
            nyo, nxo = self.obj.shape[-2:]
            ov = np.zeros((nprobe, nyo, nxo))
            for (_, _, p, _, oy, ox) in views:
                ov[p, oy:oy + ny, ox:ox + nx] += w2_nat[p]
            dens = np.stack([ov[p, oy:oy + ny, ox:ox + nx] for (_, _, p, _, oy, ox) in views])
            pmax = float(dens.reshape(nbatch, nmode, ny, nx).sum(axis=1).max())
            a = obj_alpha
            d_obj = np.zeros_like(self.obj)
            for k, (_, _, p, _, oy, ox) in enumerate(views):
                step = np.conj(w_nat[p]) / ((1.0 - a) * dens[k] + a * pmax) * delta_nat[k]
                d_obj[0, oy:oy + ny, ox:ox + nx] += step
        d_probe = None
        if probe_step is not None:
            a = probe_alpha
            d_probe = np.zeros_like(self.probe)
            for b in range(nbatch):
                fv = views[b * nmode:(b + 1) * nmode]
                oy, ox = fv[0][4], fv[0][5]
                win = self.obj[0, oy:oy + ny, ox:ox + nx]
                w2 = win.real ** 2 + win.imag ** 2
                omax = float(w2.max())
                wgt = np.conj(win * self.quad) / ((1.0 - a) * w2 + a * omax)
                for m, v in enumerate(fv):
                    d_probe[v[2]] += wgt * delta_nat[b * nmode + m]
        if obj_step is not None:
            self.obj = self.obj + d_obj * obj_step
        if probe_step is not None:
            self.probe = self.probe + d_probe * (probe_step / nbatch)


    def _rpie_batch_moved(self, frames, obj_step, probe_step, obj_alpha, probe_alpha):
        """One rPIE batch under the offsets (PATCH), from the batch-start object, probe and
        offsets: the views' waves W are the moved probe terms; the object step divides
        conj(W)·Δ by the batch overlap of |W|² per probe mode; each frame's probe direction
        conj(q·O_win)·Δ/((1 − α)|O_win|² + α·max) is moved back by the frame's offset before it
        joins the batch mean; with `refine` the frames' position directions (gradient sums of
        the window's Fourier derivative against mode 0's W and Δ) take the clipped SGD step.
        ptyrs/src/engine/ptycho.rs:900-1086"""
        ny, nx = self.probe.shape[-2:]
        nmode = self.addr.nmode
        nprobe = self.probe.shape[0]
        nbatch = len(frames)
        views, psi, delta = self.rpie_delta(frames)
        w, _ = self._moved_psi(views)
        delta_nat = to_natural(delta)
        d_obj = None
        if obj_step is not None:
            nyo, nxo = self.obj.shape[-2:]
            w2 = w.real ** 2 + w.imag ** 2
            ov = np.zeros((nprobe, nyo, nxo))
            for k, (_, _, p, _, oy, ox) in enumerate(views):
                ov[p, oy:oy + ny, ox:ox + nx] += w2[k]
            dens = np.stack([ov[p, oy:oy + ny, ox:ox + nx] for (_, _, p, _, oy, ox) in views])
            pmax = float(dens.reshape(nbatch, nmode, ny, nx).sum(axis=1).max())
            a = obj_alpha
            d_obj = np.zeros_like(self.obj)
            for k, (_, _, p, _, oy, ox) in enumerate(views):
                d_obj[0, oy:oy + ny, ox:ox + nx] += (np.conj(w[k]) / ((1.0 - a) * dens[k] + a * pmax)
                                                     * delta_nat[k])
        d_probe = None
        if probe_step is not None:
            a = probe_alpha
            d_probe = np.zeros_like(self.probe)
            for b in range(nbatch):
                fv = views[b * nmode:(b + 1) * nmode]
                oy, ox = fv[0][4], fv[0][5]
                win = self.obj[0, oy:oy + ny, ox:ox + nx]
                w2 = win.real ** 2 + win.imag ** 2
                wgt = np.conj(win * self.quad) / ((1.0 - a) * w2 + a * float(w2.max()))
                frame = np.zeros_like(self.probe)
                for m, v in enumerate(fv):
                    frame[v[2]] += wgt * delta_nat[b * nmode + m]
                d_probe += move_planes(frame, self.quad, -self.offsets[fv[0][0]])
        d_pos = None
        if self.refine:
            sums = []
            for b in range(nbatch):
                k = b * nmode
                _, _, _, o, oy, ox = views[k]
                sums.append(gradient_sums(self.obj[o, oy:oy + ny, ox:ox + nx], w[k],
                                          delta_nat[k]))
            d_pos = gradient_directions(sums, EPS_POSITION
                                        * self.optics.propagator.intensity_scale())
        if obj_step is not None:
            self.obj = self.obj + d_obj * obj_step
        if probe_step is not None:
            self.probe = self.probe + d_probe * (probe_step / nbatch)
        if d_pos is not None:
            batch_position_step(self.offsets, frames, d_pos)


    def _rpie_batch_sampled(self, frames, obj_step, probe_step, obj_alpha, probe_alpha):
        """One rPIE batch under FullObject sampling, from the batch-start arrays: the waves are
        the unmoved P·q and each view reads its window sampled at its corner plus its frame's
        offset; the object step's terms conj(W)·Δ/((1 − α)D_m + α·max) are placed by the
        sampling's adjoint, D_m the batch's |W|² of probe mode m placed and read back through
        the windows (real part floored at 0); the probe direction reads the sampled window
        and is not moved; the position sums take the sampled window's derivatives over the
        position. ptyrs/src/engine/ptycho.rs:900-1086, :1089-1130"""
        ny, nx = self.probe.shape[-2:]
        nmode = self.addr.nmode
        nprobe = self.probe.shape[0]
        nbatch = len(frames)
        views, psi, delta = self.rpie_delta(frames)
        sm = self._sampler(views)
        w = self.moved_waves(views)
        delta_nat = to_natural(delta)
        d_obj = None
        if obj_step is not None:
            w2 = w.real ** 2 + w.imag ** 2
            dens = np.zeros((len(views), ny, nx))
            for m in range(nprobe):
                ks = [k for k, v in enumerate(views) if v[2] == m]
                if not ks:
                    continue
                patches = np.zeros((len(views), ny, nx), dtype=np.complex128)
                patches[ks] = w2[ks]
                read = sm.windows(sm.place(patches, 1))
                dens[ks] = np.maximum(read[ks].real, 0.0)
            pmax = float(dens.reshape(nbatch, nmode, ny, nx).sum(axis=1).max())
            a = obj_alpha
            d_obj = sm.place(np.conj(w) / ((1.0 - a) * dens + a * pmax) * delta_nat, 1)
        d_probe = None
        if probe_step is not None:
            a = probe_alpha
            windows = sm.windows(self.obj)
            d_probe = np.zeros_like(self.probe)
            for b in range(nbatch):
                fv = views[b * nmode:(b + 1) * nmode]
                win = windows[b * nmode]
                w2 = win.real ** 2 + win.imag ** 2
                wgt = np.conj(win * self.quad) / ((1.0 - a) * w2 + a * float(w2.max()))
                for m, v in enumerate(fv):
                    d_probe[v[2]] += wgt * delta_nat[b * nmode + m]
        d_pos = None
        if self.refine:
            gy, gx = sm.windows(self.obj, (0,)), sm.windows(self.obj, (1,))
            sums = [derivative_sums(gy[b * nmode], gx[b * nmode], w[b * nmode],
                                    delta_nat[b * nmode]) for b in range(nbatch)]
            d_pos = gradient_directions(sums, EPS_POSITION
                                        * self.optics.propagator.intensity_scale())
        if obj_step is not None:
            self.obj = self.obj + d_obj * obj_step
        if probe_step is not None:
            self.probe = self.probe + d_probe * (probe_step / nbatch)
        if d_pos is not None:
            batch_position_step(self.offsets, frames, d_pos)


def far_field_ptycho(counts, y, x, obj, probe, wavelength, distance, detector_pixel):
    """`Ptycho::new`: the far field of the scan geometry. ptyrs/src/engine/ptycho.rs:218-221"""
    frame = list(np.asarray(counts).shape[-2:])
    optics = Optics.far_field(wavelength, distance, detector_pixel, frame)
    return Ptycho(counts, y, x, obj, probe, optics)
