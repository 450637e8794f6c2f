"""pty-chi's multislice exit wave over the oracle's arrays: the wave entering slice 0 is the
view's probe plane; slice s multiplies the wave entering it by its object window, giving psi_s,
which reaches slice s + 1 through gap s (an angular-spectrum propagator); the last slice's psi,
times the plane phase q, is the exit wave. One slice is `obj_probe_to_psi` itself. Under
PATCH a frame's position offset enters through its moved probe planes (the caller's) and, past
one slice, the last slice's plane phase displaced with them; under FullObject (`sampler`) every
slice is read through each view's window sampled at its corner plus its frame's offset, and
the maps of an object change take that sampling, or its adjoint, around the same model.

Test infrastructure (see oracle/__init__.py). Follows ptyrs/src/engine/multislice.rs:35-430.
"""

import numpy as np

from .base import to_natural, to_shifted


class SliceWaves:
    """The waves of `views` through the slices of `optics` at `objects` (one `[nobj, nyo, nxo]`
    array per slice) and `probe`. ptyrs/src/engine/multislice.rs:70-171"""

    def __init__(self, optics, views, objects, probe, view_quads=None, sampler=None):
        if len(objects) != optics.nslice():
            raise ValueError("one object per slice of the optics")
        self.sampler = sampler
        self.obj_shape = objects[0].shape
        if sampler is not None:
            # what the views read: one sampled window per view, at corner (0, 0)
            objects = [sampler.windows(o) for o in objects]
            views = [(f, m, p, k, 0, 0) for k, (f, m, p, _, _, _) in enumerate(views)]
        self.optics, self.views, self.objects = optics, list(views), list(objects)
        # per view, the last slice's phase displaced by its frame's offset (more than one slice)
        self.view_quads = view_quads if len(objects) > 1 else None
        ny, nx = probe.shape[-2:]
        self.shape = (ny, nx)
        self.probe_shape = probe.shape
        # incident[s][k]: the wave entering slice s at view k, natural layout
        w = np.stack([probe[v[2]] for v in self.views])
        self.incident = [w]
        for s, gap in enumerate(optics.gaps):
            w = gap.forward(self._times_window(s, self.objects[s], w))
            self.incident.append(w)
        last = len(self.objects) - 1
        nat = self._times_window(last, self.objects[last], self.incident[last])
        self.psi = np.stack([to_shifted(p) for p in nat])

    def quad(self, s, k=0):
        """The plane phase slice s applies to view k: q on the last slice (displaced per frame
        where offsets move it), 1 elsewhere. ptyrs/src/engine/multislice.rs:187-241"""
        if s + 1 == len(self.objects):
            return self.optics.quad if self.view_quads is None else self.view_quads[k]
        return np.ones(self.shape, dtype=np.complex128)

    def _times_window(self, s, obj, w):
        """`q_s·O[win_k]·w_k` per view, natural layout. ptyrs/src/engine/multislice.rs:173-185"""
        ny, nx = self.shape
        out = np.empty_like(w)
        for k, (_, _, _, o, oy, ox) in enumerate(self.views):
            out[k] = w[k] * obj[o, oy:oy + ny, ox:ox + nx] * self.quad(s, k)
        return out

    def back_propagate(self, chi):
        """`chi` (shifted planes, a change of the exit waves) back through the slices: slice
        S − 1 keeps it, slice s takes `gap_sᴴ(conj(q_{s+1}·O_{s+1}[win])·chi_{s+1})`.
        ptyrs/src/engine/multislice.rs:266-291"""
        ny, nx = self.shape
        out = [chi]
        for s in range(len(self.objects) - 2, -1, -1):
            up = out[-1]
            obj = self.objects[s + 1]
            g = np.empty((len(self.views), ny, nx), dtype=np.complex128)
            for k, (_, _, _, o, oy, ox) in enumerate(self.views):
                q = self.quad(s + 1, k)
                g[k] = to_natural(up[k]) * np.conj(q * obj[o, oy:oy + ny, ox:ox + nx])
            g = self.optics.gaps[s].adjoint(g)
            out.append(np.stack([to_shifted(p) for p in g]))
        out.reverse()
        return out

    def windows(self, d):
        """A change `d` of a slice as the views read it: `d` itself, under FullObject its
        sampled windows. ptyrs/src/engine/multislice.rs:301-306"""
        return d if self.sampler is None else self.sampler.windows(d)

    def obj_adjoint(self, s, chi_s):
        """`Σ_views conj(q_s·in_s)·chi_s` scattered onto slice s (under FullObject onto the
        views' windows, then placed by the sampling's adjoint).
        ptyrs/src/engine/multislice.rs:402-430"""
        ny, nx = self.shape
        out = np.zeros(self.objects[s].shape, dtype=np.complex128)
        for k, (_, _, _, o, oy, ox) in enumerate(self.views):
            out[o, oy:oy + ny, ox:ox + nx] += to_natural(chi_s[k]) * np.conj(
                self.quad(s, k) * self.incident[s][k])
        if self.sampler is not None:
            return self.sampler.place(out, self.obj_shape[0])
        return out

    def obj_tangent(self, s, d):
        """The change of psi_s along `d` on slice s with the wave entering it held (shifted
        planes). ptyrs/src/engine/multislice.rs:385-397"""
        nat = self._times_window(s, self.windows(d), self.incident[s])
        return np.stack([to_shifted(p) for p in nat])

    def probe_adjoint(self, chi_0, np_):
        """Plane b·np + m: `conj(q_0·O_0[win])·chi_0` of the b-th frame and probe mode m
        (a view's probe index taken modulo `np_`, the shared probe's modes).
        ptyrs/src/engine/multislice.rs:432-442, lsqml.rs:200-205"""
        ny, nx = self.shape
        nmode = 0
        while nmode < len(self.views) and self.views[nmode][0] == self.views[0][0]:
            nmode += 1
        nb = len(self.views) // nmode
        out = np.zeros((nb * np_, ny, nx), dtype=np.complex128)
        obj = self.objects[0]
        for k, (_, _, p, o, oy, ox) in enumerate(self.views):
            b = k // nmode
            q = self.quad(0, k)
            out[b * np_ + p % np_] += np.conj(q * obj[o, oy:oy + ny, ox:ox + nx]) * to_natural(
                chi_0[k])
        return out

    def probe_norms(self, np_):
        """Plane b·np + m: |O_0[win]|² of the b-th frame's view of probe mode m.
        ptyrs/src/engine/ops.rs:312-343"""
        ny, nx = self.shape
        nmode = 0
        while nmode < len(self.views) and self.views[nmode][0] == self.views[0][0]:
            nmode += 1
        nb = len(self.views) // nmode
        out = np.zeros((nb * np_, ny, nx))
        obj = self.objects[0]
        for k, (_, _, p, o, oy, ox) in enumerate(self.views):
            w = obj[o, oy:oy + ny, ox:ox + nx]
            out[(k // nmode) * np_ + p % np_] += w.real ** 2 + w.imag ** 2
        return out

    def tangent(self, d_objects, d_probe):
        """The change of the exit waves along `d_objects` (one per slice) and `d_probe`: the
        linearisation of the slice model. ptyrs/src/engine/multislice.rs:359-383"""
        w = np.stack([d_probe[v[2]] for v in self.views])
        for s, (o, d) in enumerate(zip(self.objects, d_objects)):
            out = (self._times_window(s, o, w)
                   + self._times_window(s, self.windows(d), self.incident[s]))
            if s < len(self.optics.gaps):
                w = self.optics.gaps[s].forward(out)
        return np.stack([to_shifted(p) for p in out])
