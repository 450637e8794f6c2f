"""The oracle's AP, DM, RAAR and ML (Poisson CG) against PyNX promoted to f64 in its near-field
model, the reference's self-contained `pynx_near_*` fixtures (ptyrs/tests/engine_pynx_near.rs):
LLK columns of every record, final object and probe; exact-kernel runs at 5e-12, paraxial ones
at 3e-5 (the reference's bounds)."""

import numpy as np
import pytest

from conftest import max_rel, table, worst_of
from oracle.optics import ANGULAR_SPECTRUM, Optics, Propagator, Wavefield
from oracle.ptycho import Ptycho

SHAPE, OBJ, NPROBE, AP_BEFORE_ML = (24, 24), (40, 40), 2, 10


def start():
    c = table("pynx_near_case_f64.txt")[0]
    corners = np.array(table("pynx_near_corners_f64.txt"))
    pixel = c[1]
    at = [(corners[:, 1 + a] - (OBJ[a] - SHAPE[a]) // 2) * pixel for a in (0, 1)]
    counts = np.array([r[1:] for r in table("pynx_near_counts_f64.txt")])
    counts = counts.reshape(len(corners), *SHAPE)
    w = Wavefield(SHAPE, c[0], [pixel, pixel], c[2])
    assert Propagator.free_space(w).kind == ANGULAR_SPECTRUM
    obj = np.array([r[2] + 1j * r[3] for r in table("pynx_near_obj_f64.txt")]).reshape(1, *OBJ)
    probe = np.array([r[3] + 1j * r[4] for r in table("pynx_near_probe_f64.txt")])
    p = Ptycho(counts, at[0], at[1], obj, probe.reshape(NPROBE, *SHAPE), Optics.free_space(w))
    assert np.array_equal(p.addr.oy, corners[:, 1]) and np.array_equal(p.addr.ox, corners[:, 2])
    return c, p


def columns(l):
    return [l.poisson, l.gaussian, l.euclidian, l.nb_photons_calc]


def engine_run(algorithm, cycles):
    """ptyrs/tests/engine_pynx_near.rs:103-152"""
    _, p = start()
    ap = dict(obj_inertia=0.01, probe_inertia=0.001, calc_llk=1, zero_phase_ramp=False)
    if algorithm == 0:
        p.ap(cycles, **ap)
    elif algorithm == 1:
        p.dm(cycles, calc_llk=1, zero_phase_ramp=False)
    elif algorithm == 2:
        p.raar(cycles, calc_llk=1, zero_phase_ramp=False)
    else:
        p.ap(AP_BEFORE_ML, **dict(ap, calc_llk=0))
        p.ml(cycles, update_probe=True, calc_llk=1, zero_phase_ramp=False)
    llk = [columns(h[2]) for h in p.history]
    if algorithm > 0:
        llk = llk[1:] + [columns(p.llk())]
    assert len(llk) == cycles
    return llk, p


def unit(a):
    a = a.ravel()
    s = np.abs(a).sum()
    return np.stack([a.real / s, a.imag / s], axis=1).ravel()


@pytest.mark.parametrize("run", range(8))
def test_ap_dm_raar_ml_match_pynx(run):
    r = table("pynx_near_runs_f64.txt")[run]
    exact, a, cycles = r[1] == 1.0, int(r[2]), int(r[3])
    llk, p = engine_run(a, cycles)
    want = np.array([t[2:6] for t in table("pynx_near_llk_f64.txt") if t[0] == r[0]])
    got = np.array(llk)
    llk_err = worst_of(worst_of(np.abs(got[:, k] - want[:, k]) / np.abs(want[:, k])) for k in range(4))
    wo = np.array([t[3] + 1j * t[4] for t in table("pynx_near_run_obj_f64.txt") if t[0] == r[0]])
    wp = np.array([t[4] + 1j * t[5] for t in table("pynx_near_run_probe_f64.txt") if t[0] == r[0]])
    if a in (1, 2):
        errs = [max_rel(unit(p.obj), unit(wo)), max_rel(unit(p.probe), unit(wp))]
    else:
        flat = lambda v: np.stack([v.ravel().real, v.ravel().imag], axis=1).ravel()
        errs = [max_rel(flat(p.obj), flat(wo)), max_rel(flat(p.probe), flat(wp))]
    bound = 5e-12 if exact else 3e-5
    for name, e in zip(("LLK", "object", "probe"), [llk_err] + errs):
        assert e < bound, (name, e)
