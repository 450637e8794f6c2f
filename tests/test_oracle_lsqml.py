"""The oracle's LSQML (Poisson) against the reference's self-contained near-field fixtures
(ptyrs/tests/engine_lsqml_near.rs): one step on frames 0..8 with the illumination as
preconditioner, and the 30-epoch reconstructions; cases angular spectrum, Fresnel, angular
spectrum with an evanescent band, Fresnel back. Step sizes at 1e-13/1e-12, changes and LLK
traces at 1e-12."""

import numpy as np
import pytest

from conftest import max_rel, table
from oracle.base import AddrTable, observed
from oracle.lsqml import POISSON, Lsqml, illumination
from oracle.optics import Optics, Propagator, Wavefield
from oracle.ptycho import llk, obj_probe_to_psi

SHAPE, OBJ, NPROBE, BATCH = (24, 24), (40, 40), 2, 8


def state():
    corners = np.array(table("lsqml_near_corners_f64.txt"))
    at = [corners[:, 1 + a] - (OBJ[a] - SHAPE[a]) // 2 for a in (0, 1)]
    addr = AddrTable(at[0], at[1], [1.0, 1.0], SHAPE, OBJ, NPROBE, 1)
    assert np.array_equal(addr.oy, corners[:, 1]) and np.array_equal(addr.ox, corners[:, 2])
    obj = np.array([r[2] + 1j * r[3] for r in table("lsqml_near_obj_f64.txt")]).reshape(1, *OBJ)
    probe = np.array([r[3] + 1j * r[4] for r in table("lsqml_near_probe_f64.txt")])
    probe = probe.reshape(NPROBE, *SHAPE)
    counts = table("lsqml_near_counts_f64.txt")
    cases = []
    for c in table("lsqml_near_cases_f64.txt"):
        w = Wavefield(SHAPE, c[1], [c[2], c[2]], c[3])
        optics = Optics.free_space(w)
        assert Propagator.free_space(w).kind == int(c[4])
        vals = np.array([r[2:] for r in counts if r[0] == c[0]]).reshape(len(corners), *SHAPE)
        iobs, nb_obs = observed(vals)
        cases.append((optics, iobs, nb_obs, c[6]))
    return addr, obj, probe, cases


def poisson_llk(addr, case, obj, probe):
    optics, iobs, nb_obs, _ = case
    psi = optics.exit_forward(obj_probe_to_psi(None, addr.views(), obj, probe, optics.quad))
    return llk(iobs, addr.nmode, psi).normalised(nb_obs).poisson


def check_step(st, rows, col):
    assert st["frames"] == [int(r[col]) for r in rows]
    for name, c, tol in (("a_far", col + 1, 1e-13), ("a_obj", col + 2, 1e-12),
                         ("a_probe", col + 3, 1e-12)):
        assert max_rel(st[name], [r[c] for r in rows]) < tol, name


def check_changes(case, obj, probe, obj0, probe0, obj_file, probe_file):
    for got, file, lead in ((obj - obj0, obj_file, 3), (probe - probe0, probe_file, 4)):
        rows = np.array([r for r in table(file) if r[0] == case])
        g = np.stack([got.real.ravel(), got.imag.ravel()], axis=1).ravel()
        assert max_rel(g, rows[:, lead:lead + 2].ravel()) < 1e-12, file


@pytest.mark.parametrize("case", range(4))
def test_step_matches_ptychi(case):
    addr, obj, probe, cases = state()
    optics, iobs, _, _ = cases[case]
    run = Lsqml(BATCH, POISSON, 0)
    run.preconditioner = illumination(addr.views(), obj.shape, probe)
    views = addr.views()[:BATCH * addr.nmode]
    o2, p2, st = run.step(views, iobs, optics, obj, probe)
    rows = [r for r in table("lsqml_near_step_steps_f64.txt") if r[0] == case]
    check_step(st, rows, 1)
    check_changes(case, o2, p2, obj, probe, "lsqml_near_step_obj_f64.txt",
                  "lsqml_near_step_probe_f64.txt")


@pytest.mark.parametrize("case", range(4))
def test_reconstruction_matches_ptychi(case):
    addr, obj0, probe0, cases = state()
    optics, iobs, _, truth = cases[case]
    want = [r[2] for r in table("lsqml_near_llk_f64.txt") if r[0] == case]
    rows = [r for r in table("lsqml_near_epochs_steps_f64.txt") if r[0] == case]
    run = Lsqml(BATCH, POISSON, 7)
    obj, probe = obj0, probe0
    llks = [poisson_llk(addr, cases[case], obj, probe)]
    nxt = 0
    for e in range(len(want) - 1):
        obj, probe, steps = run.epoch(addr, iobs, optics, obj, probe)
        for st in steps:
            part = rows[nxt:nxt + len(st["frames"])]
            nxt += len(st["frames"])
            assert all(r[1] == e for r in part)
            check_step(st, part, 2)
        llks.append(poisson_llk(addr, cases[case], obj, probe))
    assert nxt == len(rows)
    assert max_rel(llks, want) < 1e-12
    reached = lambda t: next((i for i, v in enumerate(t) if v <= truth), None)
    assert reached(llks) == reached(want)
    check_changes(case, obj, probe, obj0, probe0, "lsqml_near_epochs_obj_f64.txt",
                  "lsqml_near_epochs_probe_f64.txt")
