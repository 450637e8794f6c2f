"""pty-chi LSQML as the reference's CPU engine restates it, one object slice or several
(`SliceWaves`), with position offsets (refined or held) under the PATCH or FullObject model
and no OPR:
one batch's directions (`LsqmlBatch`), the per-frame step sizes, the step and the epochs with
the preconditioner schedule. An object is one `[nobj, nyo, nxo]` array, or a list of them, one
per slice of the optics.

Test infrastructure (see oracle/__init__.py). Follows ptyrs/src/engine/lsqml.rs.
"""

import math

import numpy as np

from .base import random_batches, to_natural, to_shifted
from .multislice import SliceWaves
from .positions import (EPS_POSITION, batch_position_step, derivative_sums, gradient_directions,
                        gradient_sums, move_planes, moved_quad)
from .sampling import Sampler, auto_sampling, full_object_illumination, view_offsets
from .ptycho import obj_probe_to_psi, llk

EPS_AMPLITUDE = 1e-9   # ptyrs/src/engine/lsqml.rs:25-31
EPS_POISSON = 1e-6
EPS_XI = 1e-5
LAMBDA0 = 1.2e-7

AMPLITUDE, POISSON = "amplitude", "poisson"


def illumination(views, obj_shape, probe):
    """Σ_views |P[probe_k]|² on each view's window. ptyrs/src/engine/lsqml.rs:372-388"""
    ny, nx = probe.shape[-2:]
    out = np.zeros(obj_shape)
    p2 = probe.real ** 2 + probe.imag ** 2
    for (_, _, p, o, oy, ox) in views:
        out[o, oy:oy + ny, ox:ox + nx] += p2[p]
    return out


def trim_mean(x):
    """MATLAB trimmean(x, 10). ptyrs/src/engine/lsqml.rs:390-398"""
    s = sorted(x)
    k = int(math.ceil(len(s) * 0.1 / 2.0 - 0.5))
    kept = s[k:len(s) - k]
    return sum(kept) / len(kept)


def frame_mean(d, np_):
    """Mean over frames of planes b·np + m per mode. ptyrs/src/engine/lsqml.rs:353-370"""
    count = d.shape[0] // np_
    out = np.zeros((np_,) + d.shape[1:], dtype=d.dtype)
    for b in range(count):
        for m in range(np_):
            out[m] += d[b * np_ + m]
    return out / count


def as_slices(obj):
    """An object as its list of slices."""
    return list(obj) if isinstance(obj, (list, tuple)) else [obj]


def unique_waves(run, views, optics, objects, probe, var=None):
    """The slice waves of `views` (whole frames) at each frame's own probe planes (`var`'s
    OPR planes, or the shared probe's): plane b·np + m of the b-th frame, moved by its offset
    under PATCH (q = the first slice's plane phase), and past one slice the last slice's phase
    displaced with it. ptyrs/src/engine/lsqml.rs:483-536, ptyrs/src/engine/opr.rs:88-124"""
    nmode = 0
    while nmode < len(views) and views[nmode][0] == views[0][0]:
        nmode += 1
    frames = [views[k][0] for k in range(0, len(views), nmode)]
    np_ = probe.shape[0]
    if var is not None:
        uviews, planes = var.unique(views, probe)
    else:
        uviews = [(f, m, (k // nmode) * np_ + p, o, oy, ox)
                  for k, (f, m, p, o, oy, ox) in enumerate(views)]
        planes = None
    if run.offsets is None:
        if planes is None:
            return frames, nmode, None, SliceWaves(optics, views, objects, probe)
        return frames, nmode, None, SliceWaves(optics, uviews, objects, planes)
    if run.sampling_of(optics, objects[0], probe) == "full_object":
        sm = Sampler(objects[0].shape[-2:], probe.shape[-2:], views,
                     view_offsets(views, run.offsets))
        if planes is None:
            return frames, nmode, None, SliceWaves(optics, views, objects, probe, sampler=sm)
        return frames, nmode, None, SliceWaves(optics, uviews, objects, planes, sampler=sm)
    q0 = optics.quad if len(objects) == 1 else None
    moves = [run.offsets[f].copy() for f in frames]
    if planes is None:
        planes = np.concatenate([move_planes(probe, q0, d) for d in moves])
    else:
        planes = np.concatenate([move_planes(planes[b * np_:(b + 1) * np_], q0, d)
                                 for b, d in enumerate(moves)])
    quads = None
    if len(objects) > 1:
        quads = np.stack([moved_quad(optics, moves[k // nmode]) for k in range(len(views))])
    return frames, nmode, moves, SliceWaves(optics, uviews, objects, planes, quads)


class LsqmlBatch:
    """One batch's directions at one state. ptyrs/src/engine/lsqml.rs:97-181"""

    def __init__(self, run, views, iobs, optics, obj, probe, var=None):
        cost = run.cost
        ny, nx = probe.shape[-2:]
        self.objects, self.probe, self.optics = as_slices(obj), probe, optics
        self.frames, nmode, self.moves, self.waves = unique_waves(run, views, optics,
                                                                  self.objects, probe, var)
        self.views, self.nmode = views, nmode
        psi = self.waves.psi
        self.psi = psi
        far = optics.exit_forward(psi)
        nb = len(self.frames)
        fa = far.reshape(nb, nmode, ny, nx)
        icalc = (fa.real ** 2 + fa.imag ** 2).sum(axis=1)
        io = np.stack([iobs[f] for f in self.frames])
        if cost == POISSON:
            a_far = [run.a_far_of(f) for f in self.frames]
            for _ in range(2):
                prev = sum(a_far) / len(a_far)
                new = []
                for k in range(nb):
                    yt, yp = io[k], icalc[k]
                    keep = ~np.isnan(yt)
                    yt, yp = yt[keep], yp[keep]
                    xi = 1.0 - yt / (yp + EPS_XI)
                    num = float((xi * (yp - yt / (1.0 - prev * xi))).sum())
                    den = float((xi * xi * yp).sum())
                    ratio = num / den if den > 0.0 else 0.0
                    new.append(min(max(0.5 * prev + 0.5 * ratio, 0.0), 1.0))
                a_far = new
        else:
            a_far = [0.5] * nb
        self.a_far = a_far
        masked = np.isnan(io)
        if cost == AMPLITUDE:
            c = 2.0 * (1.0 - np.sqrt(np.maximum(np.where(masked, 0.0, io), 0.0))
                       / (np.sqrt(icalc) + EPS_AMPLITUDE))
        else:
            c = 1.0 - np.where(masked, 0.0, io) / (icalc + EPS_POISSON)
        c = np.where(masked, 0.0, c)
        scale = 1.0 - np.asarray(a_far)[:, None, None] * c
        chi = (fa * scale[:, None]).reshape(far.shape)
        # chi_s per slice, the exit's on the last (ptyrs/src/engine/lsqml.rs:166-169)
        self.chis = self.waves.back_propagate(optics.exit_adjoint(chi) - psi)
        self.chi = self.chis[0]

    def obj_directions(self):
        """pty-chi `delta_o_comb` per slice. ptyrs/src/engine/lsqml.rs:191-198"""
        return [self.waves.obj_adjoint(s, c) for s, c in enumerate(self.chis)]

    def obj_direction(self):
        """The first slice's object direction."""
        return self.obj_directions()[0]

    def frame_probe_directions(self):
        """Plane b·np + m: conj(q·O_view)·chi_m on slice 0. ptyrs/src/engine/lsqml.rs:200-205"""
        return self.waves.probe_adjoint(self.chis[0], self.probe.shape[0])

    def frame_probe_norms(self):
        """Plane b·np + m: |O_view|² on slice 0 (accumulate_probe's norm).
        ptyrs/src/engine/ops.rs:312-343"""
        return self.waves.probe_norms(self.probe.shape[0])

    def adjoint_move(self, d):
        """Frame b's planes of `d` moved back by −offset (the adjoint of the probe move).
        ptyrs/src/engine/lsqml.rs:207-222"""
        if self.moves is None:
            return d
        np_ = self.probe.shape[0]
        q0 = self.optics.quad if len(self.objects) == 1 else None
        out = d.copy()
        for b, off in enumerate(self.moves):
            out[b * np_:(b + 1) * np_] = move_planes(d[b * np_:(b + 1) * np_], q0, -off)
        return out

    def probe_direction(self):
        return frame_mean(self.adjoint_move(self.frame_probe_directions()), self.probe.shape[0])

    def position_directions(self):
        """Per frame [Δy, Δx] on slice S//2 from its first view (probe and object mode 0):
        gradient sums of the Fourier derivative of the object window against w·q and chi_s.
        ptyrs/src/engine/lsqml.rs:230-270"""
        s = len(self.objects) // 2
        ny, nx = self.probe.shape[-2:]
        sm = self.waves.sampler
        if sm is not None:
            gy, gx = sm.windows(self.objects[s], (0,)), sm.windows(self.objects[s], (1,))
        sums = []
        for b in range(len(self.frames)):
            k = b * self.nmode
            _, _, _, o, oy, ox = self.views[k]
            wq = self.waves.incident[s][k] * self.waves.quad(s, k)
            chi = to_natural(self.chis[s][k])
            if sm is not None:
                sums.append(derivative_sums(gy[k], gx[k], wq, chi))
                continue
            patch = self.objects[s][o, oy:oy + ny, ox:ox + nx]
            sums.append(gradient_sums(patch, wq, chi))
        eps = EPS_POSITION * self.optics.propagator.intensity_scale()
        return gradient_directions(sums, eps)

    def step_sizes(self, d_obj, d_probe):
        """Per slice and frame a_obj, per frame a_probe: slice 0 jointly with the probe by the
        2×2 system, every other slice alone (Eq. 23b). ptyrs/src/engine/lsqml.rs:272-350"""
        d_obj = as_slices(d_obj)
        ny, nx = self.probe.shape[-2:]
        n = ny * nx
        # the shared views on what they read (under FullObject the sampled windows)
        read = [(f, m, p, o, oy, ox) for (f, m, p, _, _, _), (_, _, _, o, oy, ox)
                in zip(self.views, self.waves.views)]
        dpo = obj_probe_to_psi(None, read, self.waves.objects[0], d_probe,
                               self.waves.quad(0, 0))
        dop = self.waves.obj_tangent(0, d_obj[0])
        lam = LAMBDA0 * self.optics.propagator.intensity_scale() / n
        nb = len(self.frames)
        sh = (nb, self.nmode * n)
        u, v, c = dop.reshape(sh), dpo.reshape(sh), self.chis[0].reshape(sh)
        a12 = u * np.conj(v)
        sums = np.stack([(u.real ** 2 + u.imag ** 2 + lam).sum(axis=1),
                         (v.real ** 2 + v.imag ** 2 + lam).sum(axis=1),
                         a12.real.sum(axis=1), a12.imag.sum(axis=1),
                         (np.conj(u) * c).real.sum(axis=1),
                         (np.conj(v) * c).real.sum(axis=1)], axis=1)
        reg11, reg22 = 0.1 * sums[:, 0].mean(), 0.1 * sums[:, 1].mean()
        scale = float(len(self.objects) * self.nmode)
        a_obj, a_probe = [], []
        for s in sums:
            a11, a22 = s[0] + reg11, s[1] + reg22
            det = a11 * a22 - (s[2] ** 2 + s[3] ** 2)
            x0 = (a22 * s[4] - s[2] * s[5]) / det
            x1 = (a11 * s[5] - s[2] * s[4]) / det
            a_obj.append(0.9 * max(x0, 0.0) / scale)
            a_probe.append(0.9 * max(x1, 0.0))
        a_objs = [a_obj]
        for s in range(1, len(self.objects)):
            u = self.waves.obj_tangent(s, d_obj[s]).reshape(sh)
            c = self.chis[s].reshape(sh)
            num = (np.conj(u) * c).real.sum(axis=1)
            den = (u.real ** 2 + u.imag ** 2).sum(axis=1)
            a_objs.append([max(0.5 * a / b * 0.9 / scale, 0.0) for a, b in zip(num, den)])
        if len(self.objects) == 1:
            return a_obj, a_probe, sums
        return a_objs, a_probe, sums


class Lsqml:
    """LSQML state across epochs. ptyrs/src/engine/lsqml.rs:411-714"""

    def __init__(self, batch_size, cost=POISSON, seed=7, offsets=None, refine=False,
                 sampling=None):
        """`offsets` ([nframe, 2] object pixels) are held, or refined after every step with
        `refine` (Features::position, ptyrs/src/engine/features.rs:304-330), modelled as
        `sampling` ("patch" or "full_object"; None: Sampling::auto at the first use)."""
        if batch_size <= 0:
            raise ValueError("LSQML batch_size must be >= 1")
        if cost not in (AMPLITUDE, POISSON):
            raise ValueError(cost)
        self.batch_size, self.cost, self.seed = batch_size, cost, seed
        self.epochs = 0
        self.preconditioner = None
        self.a_far = {}
        self.offsets = None if offsets is None else np.array(offsets, dtype=np.float64)
        self.refine = refine and offsets is not None
        self.sampling = sampling

    def sampling_of(self, optics, obj, probe):
        """The run's sampling, Sampling::auto of the optics, object and probe fixed at the first
        use if none was set. ptyrs/src/engine/lsqml.rs:450-458"""
        if self.sampling is None:
            self.sampling = auto_sampling(optics.propagator.kind, obj.shape[-2:],
                                          probe.shape[-2:])
        return self.sampling

    def a_far_of(self, f):
        return self.a_far.get(f, 0.5)

    def step(self, views, iobs, optics, obj, probe, var=None):
        """`run_minibatch`: directions, step sizes, updates, and `var`'s OPR step (its
        eigenmodes and the batch's weights) from the state on entry.
        ptyrs/src/engine/lsqml.rs:571-643. Returns the object in the form given (one array, or
        a list of slices)."""
        pre = self.preconditioner
        floor = (0.1 * pre.max()) ** 2
        b = LsqmlBatch(self, views, iobs, optics, obj, probe, var)
        d_obj = [d / np.sqrt(pre * pre + floor) for d in b.obj_directions()]
        np_ = probe.shape[0]
        d = b.frame_probe_directions()
        moved = b.adjoint_move(d)
        d_probe = frame_mean(moved, np_)
        a_obj, a_probe, _ = b.step_sizes(d_obj, d_probe)
        d_pos = b.position_directions() if self.refine else None
        for f, a in zip(b.frames, b.a_far):
            self.a_far[f] = a
        if var is not None:
            var.step(b.frames, probe[0], moved, d, b.frame_probe_norms(), np_)
        per_slice = a_obj if len(b.objects) > 1 else [a_obj]
        new = [o + trim_mean(a) * d for o, a, d in zip(as_slices(obj), per_slice, d_obj)]
        probe = probe + (sum(a_probe) / len(a_probe)) * d_probe
        if d_pos is not None:
            batch_position_step(self.offsets, b.frames, d_pos)
        out = new if isinstance(obj, (list, tuple)) else new[0]
        return out, probe, dict(frames=b.frames, a_far=b.a_far, a_obj=a_obj, a_probe=a_probe)

    def epoch(self, addr, iobs, optics, obj, probe, var=None):
        """One epoch with pty-chi's preconditioner schedule. ptyrs/src/engine/lsqml.rs:650-714"""
        nframe, nmode = addr.nframe, addr.nmode
        views_all = addr.views()
        if self.epochs < 10 or self.epochs % 10 == 0:
            first = as_slices(obj)[0]
            if (self.offsets is not None
                    and self.sampling_of(optics, first, probe) == "full_object"):
                illum = full_object_illumination(views_all, first.shape, probe, self.offsets)
            else:
                illum = illumination(views_all, first.shape, probe)
            pre = self.preconditioner
            if pre is None:
                n = probe.shape[-1] * probe.shape[-2]
                k = optics.propagator.intensity_scale()
                power = max(float(np.nansum(iobs[f])) for f in range(nframe))
                pre = illum * (power / (4.0 * n * k))
            self.preconditioner = (pre + illum) / 2.0
        epoch = self.epochs
        self.epochs += 1
        size = min(self.batch_size, nframe)
        steps = []
        for frames in random_batches(nframe, size, self.seed, epoch):
            views = [v for f in frames for v in views_all[f * nmode:(f + 1) * nmode]]
            obj, probe, st = self.step(views, iobs, optics, obj, probe, var)
            steps.append(st)
        return obj, probe, steps

    def llk(self, ptycho, obj, probe, var=None):
        """Normalised LLK of obj, probe over every frame. ptyrs/src/engine/lsqml.rs:538-555"""
        return slice_llk(ptycho.optics, ptycho.views_all, ptycho.iobs, ptycho.addr.nmode,
                         ptycho.nb_obs, obj, probe, self, var)


def slice_llk(optics, views, iobs, nmode, nb_obs, obj, probe, run=None, var=None):
    """Normalised LLK of the slice model's exit waves over `views`, at `run`'s offsets and
    `var`'s per-frame probes if given. ptyrs/src/engine/lsqml.rs:538-555"""
    if (run is not None and run.offsets is not None) or var is not None:
        if run is None:
            run = Lsqml(1)
        waves = unique_waves(run, views, optics, as_slices(obj), probe, var)[3]
    else:
        waves = SliceWaves(optics, views, as_slices(obj), probe)
    psi = optics.exit_forward(waves.psi)
    return llk(iobs, nmode, psi).normalised(nb_obs)


def cost_grads(optics, views, iobs, nmode, obj, probe, cost=POISSON):
    """The cost over `views` and its gradients 2·∂L/∂O_s* (every slice) and 2·∂L/∂P*: L the
    Poisson LLK sum (`POISSON`) or Σ(√icalc − √iobs)² (`AMPLITUDE`) over unmasked pixels; the
    exit transform's and the slices' adjoints carry the detector-plane gradient back.
    ptyrs/src/engine/ptycho.rs:492-561, ptyrs/src/engine/features.rs:310-330,
    ptyrs/src/engine/multislice.rs:266-442"""
    objects = as_slices(obj)
    np_ = probe.shape[0]
    waves = SliceWaves(optics, views, objects, probe)
    far = optics.exit_forward(waves.psi)
    nb = far.shape[0] // nmode
    fa = far.reshape(nb, nmode, *far.shape[1:])
    icalc = (fa.real ** 2 + fa.imag ** 2).sum(axis=1)
    io = np.stack([iobs[v[0]] for v in views[::nmode]])
    sums = llk(io, nmode, far)
    masked = np.isnan(io)
    ic = np.maximum(icalc, 1e-12)
    if cost == POISSON:
        value = sums.poisson
        r = 2.0 * (1.0 - np.where(masked, 0.0, io) / ic)
    else:
        value = sums.euclidian / 4.0
        r = 2.0 * (1.0 - np.sqrt(np.where(masked, 0.0, io) / ic))
    r = np.where(masked, 0.0, r)
    chis = waves.back_propagate(optics.exit_adjoint((fa * r[:, None]).reshape(far.shape)))
    og = [waves.obj_adjoint(s, c) for s, c in enumerate(chis)]
    d = waves.probe_adjoint(chis[0], np_)
    pg = d.reshape(-1, np_, *d.shape[1:]).sum(axis=0)
    return value, (og if isinstance(obj, (list, tuple)) else og[0]), pg
