"""The oracle's multislice model and LSQML against the reference's self-contained three-slice
fixtures (ptyrs/tests/engine_lsqml_multislice.rs): pty-chi's forward intensities at 1e-14, the
batch directions against central differences of the weighted Poisson cost at 1e-7, the
tangent against differences of the exit waves, one step on frames 0..8 and the 30-epoch
reconstructions at 1e-12, every case (exit 0 angular spectrum, 1 Fresnel transform, 2 far
field, whose q multiplies the last slice)."""

import numpy as np
import pytest

from conftest import max_rel, table
from oracle.base import AddrTable, observed, pixel_size_object, to_natural
from oracle.lsqml import POISSON, LsqmlBatch, Lsqml, illumination, slice_llk
from oracle.multislice import SliceWaves
from oracle.optics import Optics, Propagator, Wavefield

SHAPE, OBJ, NPROBE, NSLICE, BATCH = (24, 24), (40, 40), 2, 3, 8


def state():
    """ptyrs/tests/engine_lsqml_multislice.rs:66-145"""
    corners = np.array(table("lsqml_multislice_corners_f64.txt"))
    at = [corners[:, 1 + a] - (OBJ[a] - SHAPE[a]) // 2 for a in (0, 1)]
    addr = AddrTable(at[0], at[1], [1.0, 1.0], SHAPE, OBJ, NPROBE, 1)
    assert np.array_equal(addr.oy, corners[:, 1]) and np.array_equal(addr.ox, corners[:, 2])
    rows = np.array(table("lsqml_multislice_obj_f64.txt"))
    objects = [(rows[rows[:, 0] == s][:, 3] + 1j * rows[rows[:, 0] == s][:, 4]).reshape(1, *OBJ)
               for s in range(NSLICE)]
    rows = np.array(table("lsqml_multislice_probe_f64.txt"))
    probe = (rows[:, 3] + 1j * rows[:, 4]).reshape(NPROBE, *SHAPE)
    counts = table("lsqml_multislice_counts_f64.txt")
    cases = []
    for c in table("lsqml_multislice_cases_f64.txt"):
        if c[4] == 0.0:
            lz = c[1] * c[3]
            dpix = lz / (SHAPE[0] * c[2])
            optics = Optics.far_field(c[1], c[3], dpix, list(SHAPE))
            pixel = pixel_size_object(c[1], c[3], dpix, list(SHAPE))
        else:
            pixel = [c[2], c[2]]
            w = Wavefield(SHAPE, c[1], pixel, c[3])
            optics = Optics.free_space(w)
            assert Propagator.free_space(w).kind == int(c[4])
        optics = optics.with_slices([Wavefield(SHAPE, c[1], pixel, d) for d in c[7:]])
        assert optics.nslice() == NSLICE
        assert abs(optics.propagator.intensity_scale() - c[5]) <= 1e-15 * c[5]
        vals = np.array([r[2:] for r in counts if r[0] == c[0]]).reshape(len(corners), *SHAPE)
        iobs, nb_obs = observed(vals)
        cases.append((optics, iobs, nb_obs, c[6]))
    return addr, objects, probe, cases


def poisson_llk(addr, case, objects, probe):
    optics, iobs, nb_obs, _ = case
    return slice_llk(optics, addr.views(), iobs, addr.nmode, nb_obs, objects, probe).poisson


def test_forward_matches_ptychi():
    addr, objects, probe, cases = state()
    nmode = addr.nmode
    for r in table("lsqml_multislice_icalc_f64.txt"):
        optics = cases[int(r[0])][0]
        f = int(r[1])
        views = addr.views()[f * nmode:(f + 1) * nmode]
        far = optics.exit_forward(SliceWaves(optics, views, objects, probe).psi)
        got = to_natural((far.real ** 2 + far.imag ** 2).sum(axis=0))
        assert max_rel(got.ravel(), r[2:]) < 1e-14, (r[0], f)


def _direction(like, rng):
    scale = np.abs(like).max()
    return (rng.uniform(-0.5, 0.5, like.shape) + 1j * rng.uniform(-0.5, 0.5, like.shape)) * scale


def _slope(f, h):
    return (8.0 * (f(h) - f(-h)) - (f(2.0 * h) - f(-2.0 * h))) / (12.0 * h)


def test_directions_match_central_differences():
    """−2·obj_directions per slice and −2·BATCH·probe_direction against five-point central
    differences of Σ_f a_far_f Σ (icalc + 1e-6 − iobs·ln(icalc + 1e-6)) over frames 0..8.
    ptyrs/tests/engine_lsqml_multislice.rs:168-285"""
    addr, objects, probe, cases = state()
    views = addr.views()[:BATCH * addr.nmode]
    run = Lsqml(BATCH, POISSON, 0)
    rng = np.random.default_rng(5)
    for k, (optics, iobs, _, _) in enumerate(cases):
        b = LsqmlBatch(run, views, iobs, optics, objects, probe)
        w = np.asarray(b.a_far)

        def cost(os, pr):
            far = optics.exit_forward(SliceWaves(optics, views, os, pr).psi)
            ic = (far.real ** 2 + far.imag ** 2).reshape(BATCH, addr.nmode, *SHAPE).sum(axis=1)
            io = iobs[:BATCH]
            keep = ~np.isnan(io)
            t = np.where(keep, ic + 1e-6 - np.where(keep, io, 0.0) * np.log(ic + 1e-6), 0.0)
            return float((w * t.sum(axis=(1, 2))).sum())

        for i, g in enumerate(b.obj_directions()):
            d = _direction(objects[i], rng)

            def at(t, i=i, d=d):
                os = list(objects)
                os[i] = objects[i] + t * d
                return cost(os, probe)

            want = _slope(at, 1e-3)
            got = float((np.conj(-2.0 * g) * d).real.sum())
            assert abs(got - want) / abs(want) < 1e-7, (k, i)
        d = _direction(probe, rng)
        want = _slope(lambda t: cost(objects, probe + t * d), 1e-3)
        got = float((np.conj(-2.0 * BATCH * b.probe_direction()) * d).real.sum())
        assert abs(got - want) / abs(want) < 1e-7, (k, "probe")


def test_tangent_is_the_derivative_of_the_exit_waves():
    addr, objects, probe, cases = state()
    views = addr.views()[:6 * addr.nmode]
    rng = np.random.default_rng(9)
    d_obj = [_direction(o, rng) for o in objects]
    d_probe = _direction(probe, rng)
    h = 1e-5
    for optics, _, _, _ in cases:
        def psi(t):
            os = [o + t * d for o, d in zip(objects, d_obj)]
            return SliceWaves(optics, views, os, probe + t * d_probe).psi

        fd = (psi(h) - psi(-h)) / (2.0 * h)
        tan = SliceWaves(optics, views, objects, probe).tangent(d_obj, d_probe)
        assert np.abs(fd - tan).max() / np.abs(tan).max() < 1e-8


def _check_step(st, rows, col):
    assert st["frames"] == [int(r[col]) for r in rows]
    assert max_rel(st["a_far"], [r[col + 1] for r in rows]) < 1e-13
    for s in range(NSLICE):
        assert max_rel(st["a_obj"][s], [r[col + 2 + s] for r in rows]) < 1e-12, s
    assert max_rel(st["a_probe"], [r[col + 2 + NSLICE] for r in rows]) < 1e-12


def _check_changes(case, objects, probe, objects0, probe0, obj_file, probe_file):
    flat = lambda a: np.stack([a.real.ravel(), a.imag.ravel()], axis=1).ravel()
    got_o = np.concatenate([flat(o - o0) for o, o0 in zip(objects, objects0)])
    for got, file in ((got_o, obj_file), (flat(probe - probe0), probe_file)):
        rows = np.array([r for r in table(file) if r[0] == case])
        assert max_rel(got, rows[:, 4:6].ravel()) < 1e-12, file


@pytest.mark.parametrize("case", range(3))
def test_step_matches_ptychi(case):
    """ptyrs/tests/engine_lsqml_multislice.rs:340-372"""
    addr, objects, probe, cases = state()
    optics, iobs, _, _ = cases[case]
    run = Lsqml(BATCH, POISSON, 0)
    run.preconditioner = illumination(addr.views(), objects[0].shape, probe)
    views = addr.views()[:BATCH * addr.nmode]
    new_o, new_p, st = run.step(views, iobs, optics, objects, probe)
    rows = [r for r in table("lsqml_multislice_step_steps_f64.txt") if r[0] == case]
    _check_step(st, rows, 1)
    _check_changes(case, new_o, new_p, objects, probe, "lsqml_multislice_step_obj_f64.txt",
                   "lsqml_multislice_step_probe_f64.txt")


@pytest.mark.parametrize("case", range(3))
def test_reconstruction_matches_ptychi(case):
    """30 epochs of batches of 8 (seed 7): every step, the LLK after every epoch, the changes.
    ptyrs/tests/engine_lsqml_multislice.rs:478-541"""
    addr, objects, probe, cases = state()
    c = cases[case]
    want = [r[2] for r in table("lsqml_multislice_llk_f64.txt") if r[0] == case]
    rows = [r for r in table("lsqml_multislice_epochs_steps_f64.txt") if r[0] == case]
    run = Lsqml(BATCH, POISSON, 7)
    os, pr = list(objects), probe
    llks = [poisson_llk(addr, c, os, pr)]
    nxt = 0
    for e in range(len(want) - 1):
        os, pr, steps = run.epoch(addr, c[1], c[0], os, pr)
        for st in steps:
            sub = rows[nxt:nxt + len(st["frames"])]
            nxt += len(st["frames"])
            assert all(r[1] == e for r in sub)
            _check_step(st, sub, 2)
        llks.append(poisson_llk(addr, c, os, pr))
    assert nxt == len(rows)
    assert max_rel(llks, want) < 1e-12
    reached = lambda t: next((i for i, v in enumerate(t) if v <= c[3]), None)
    assert reached(llks) == reached(want) and reached(llks) is not None
    _check_changes(case, os, pr, objects, probe, "lsqml_multislice_epochs_obj_f64.txt",
                   "lsqml_multislice_epochs_probe_f64.txt")
