"""The oracle's position model (pty-chi PATCH) and LSQML position step against the reference's
fixtures: one Poisson LSQML step on frames 0..8 from the truth object with every frame moved by
its offset and the offsets refined, near field (ptyrs/tests/engine_lsqml_near.rs:362-450, cases
0 and 2) and multislice (ptyrs/tests/engine_lsqml_multislice.rs:380-470, case 0, the exit with
no plane phase), each under PATCH and under FULL_OBJECT (the `_full` files; the preconditioner
is the sampling's illumination at the start offsets): the step sizes, the object and probe
changes at 1e-12 and the offset changes at 1e-11. Plus the closed forms of the probe move."""

import numpy as np
import pytest

from conftest import max_rel, table
from oracle.lsqml import POISSON, Lsqml, illumination
from oracle.positions import batch_position_step, move_planes, ramp
from oracle.sampling import full_object_illumination
from test_oracle_lsqml import state as near_state
from test_oracle_multislice import state as multislice_state

BATCH = 8


def _flat(a):
    return np.stack([a.real.ravel(), a.imag.ravel()], axis=1).ravel()


def _run(addr, objects, probe, case, offsets_file, sampling="patch"):
    optics, iobs = case[0], case[1]
    rows = [r for r in table(offsets_file) if r[0] == case[4]]
    d0 = np.array([[r[2], r[3]] for r in rows])
    want = np.array([[r[4], r[5]] for r in rows])
    obj0 = objects if isinstance(objects, list) else objects
    run = Lsqml(BATCH, POISSON, 0, offsets=d0, refine=True, sampling=sampling)
    first = objects[0] if isinstance(objects, list) else objects
    if sampling == "full_object":
        run.preconditioner = full_object_illumination(addr.views(), first.shape, probe, d0)
    else:
        run.preconditioner = illumination(addr.views(), first.shape, probe)
    views = addr.views()[:BATCH * addr.nmode]
    new_o, new_p, st = run.step(views, iobs, optics, obj0, probe)
    return new_o, new_p, st, run.offsets - d0, want - d0


@pytest.mark.parametrize("sampling,tag", [("patch", "pos"), ("full_object", "pos_full")])
@pytest.mark.parametrize("case", [0, 2])
def test_near_field_position_step_matches_ptychi(case, sampling, tag):
    addr, _, probe, cases = near_state()
    start = np.array(table("lsqml_near_pos_start_f64.txt"))
    obj = (start[:, 2] + 1j * start[:, 3]).reshape(1, 40, 40)
    c = cases[case] + (case,)
    new_o, new_p, st, got_d, want_d = _run(addr, obj, probe, c,
                                           f"lsqml_near_{tag}_offsets_f64.txt", sampling)
    rows = [r for r in table(f"lsqml_near_{tag}_steps_f64.txt") if r[0] == case]
    assert st["frames"] == [int(r[1]) for r in rows]
    assert max_rel(st["a_far"], [r[2] for r in rows]) < 1e-13
    assert max_rel(st["a_obj"], [r[3] for r in rows]) < 1e-12
    assert max_rel(st["a_probe"], [r[4] for r in rows]) < 1e-12
    for got, file, lead in ((new_o - obj, f"lsqml_near_{tag}_obj_f64.txt", 3),
                            (new_p - probe, f"lsqml_near_{tag}_probe_f64.txt", 4)):
        r = np.array([t for t in table(file) if t[0] == case])
        assert max_rel(_flat(got), r[:, lead:lead + 2].ravel()) < 1e-12, file
    assert max_rel(got_d.ravel(), want_d.ravel()) < 1e-11


@pytest.mark.parametrize("sampling,tag", [("patch", "pos"), ("full_object", "pos_full")])
def test_multislice_position_step_matches_ptychi(sampling, tag):
    addr, _, probe, cases = multislice_state()
    start = np.array(table("lsqml_multislice_pos_start_f64.txt"))
    objects = [(start[start[:, 0] == s][:, 3] + 1j * start[start[:, 0] == s][:, 4])
               .reshape(1, 40, 40) for s in range(3)]
    case = 0
    assert np.all(cases[case][0].quad == 1.0)
    c = cases[case] + (case,)
    new_o, new_p, st, got_d, want_d = _run(addr, objects, probe, c,
                                           f"lsqml_multislice_{tag}_offsets_f64.txt", sampling)
    rows = [r for r in table(f"lsqml_multislice_{tag}_steps_f64.txt") if r[0] == case]
    assert st["frames"] == [int(r[1]) for r in rows]
    assert max_rel(st["a_far"], [r[2] for r in rows]) < 1e-13
    for s in range(3):
        assert max_rel(st["a_obj"][s], [r[3 + s] for r in rows]) < 1e-12, s
    assert max_rel(st["a_probe"], [r[6] for r in rows]) < 1e-12
    got = np.concatenate([_flat(o - o0) for o, o0 in zip(new_o, objects)])
    r = np.array([t for t in table(f"lsqml_multislice_{tag}_obj_f64.txt") if t[0] == case])
    assert max_rel(got, r[:, 4:6].ravel()) < 1e-12
    r = np.array([t for t in table(f"lsqml_multislice_{tag}_probe_f64.txt") if t[0] == case])
    assert max_rel(_flat(new_p - probe), r[:, 4:6].ravel()) < 1e-12
    assert max_rel(got_d.ravel(), want_d.ravel()) < 1e-11


def test_probe_move_closed_forms():
    """A whole-pixel move is a cyclic roll; the move by −d is the adjoint (and inverse) of the
    move by d; the ramp is the DFT shift theorem's."""
    rng = np.random.default_rng(4)
    p = rng.normal(size=(2, 12, 10)) + 1j * rng.normal(size=(2, 12, 10))
    q = np.exp(1j * rng.uniform(-1, 1, (12, 10)))
    assert np.abs(move_planes(p, None, [2.0, -3.0]) - np.roll(p, (2, -3), axis=(1, 2))).max() < 1e-12
    d = [0.37, -0.81]
    u, v = move_planes(p, q, d), rng.normal(size=p.shape) + 1j * rng.normal(size=p.shape)
    lhs = np.vdot(v, u)
    rhs = np.vdot(move_planes(v, q, [-d[0], -d[1]]), p)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    assert np.abs(move_planes(u, q, [-d[0], -d[1]]) - p).max() < 1e-12
    assert abs(ramp((12, 10), d)[1, 0] - np.exp(-2j * np.pi * d[0] / 12)) < 1e-15


def test_position_step_clips_by_the_mean_absolute_deviation():
    d = np.zeros((5, 2))
    batch_position_step(d, [0, 1], [[1.0, 0.01], [-1.0, 0.02]])
    # dx = 0.3·delta; axis 0: mean 0, MAD 0.12 → limit min(1.2, 0.1) = 0.1
    assert np.allclose(d[:, 0], [0.1, -0.1, 0, 0, 0])
    # axis 1: 0.003, 0.006; mean 0.0018, MAD (0.0012 + 0.0042 + 3·0.0018)/5 = 0.00216
    assert np.allclose(d[:, 1], [0.003, 0.006, 0, 0, 0])
