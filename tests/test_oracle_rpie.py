"""The oracle's rPIE against the reference's self-contained near-field fixtures
(ptyrs/tests/engine_rpie_near.rs:102-169): three propagators (angular spectrum, Fresnel
forward and back), batches of 1, 4 and 16 frames, 30 epochs; the LLK trace and the object
and probe changes at 1e-12."""

import numpy as np
import pytest

from conftest import max_rel, table
from oracle.optics import Optics, Propagator, Wavefield
from oracle.ptycho import Ptycho

SHAPE, OBJ, NPROBE = (24, 24), (40, 40), 2


def start(c):
    """Case row `case wavelength pixel distance kind scale truth_llk` at the reference's start
    (ptyrs/tests/engine_rpie_near.rs:55-100)."""
    corners = np.array(table("rpie_near_corners_f64.txt"))
    pixel = c[2]
    at = [(corners[:, 1 + a] - (OBJ[a] - SHAPE[a]) // 2) * pixel for a in (0, 1)]
    counts = np.array([r[2:] for r in table("rpie_near_counts_f64.txt") if r[0] == c[0]])
    counts = counts.reshape(len(corners), *SHAPE)
    w = Wavefield(SHAPE, c[1], [pixel, pixel], c[3])
    assert Propagator.free_space(w).kind == int(c[4])
    obj = np.array([r[2] + 1j * r[3] for r in table("rpie_near_obj_f64.txt")]).reshape(1, *OBJ)
    probe = np.array([r[3] + 1j * r[4] for r in table("rpie_near_probe_f64.txt")])
    probe = probe.reshape(NPROBE, *SHAPE)
    p = Ptycho(counts, at[0], at[1], obj, probe, Optics.free_space(w))
    assert np.array_equal(p.addr.oy, corners[:, 1]) and np.array_equal(p.addr.ox, corners[:, 2])
    return p


def _runs():
    return table("rpie_near_runs_f64.txt")


@pytest.mark.parametrize("run", range(9))
def test_rpie_matches_ptychi(run):
    cases = table("rpie_near_cases_f64.txt")
    trace = table("rpie_near_llk_f64.txt")
    r = _runs()[run]
    c = cases[int(r[1])]
    p = start(c)
    o0, p0 = p.obj.copy(), p.probe.copy()
    llks = [p.llk().poisson]
    p.rpie(int(r[3]), obj_step=r[4], probe_step=r[5], obj_alpha=r[6], probe_alpha=r[7],
           batch_size=int(r[2]), calc_llk=1)
    llks += [h[2].poisson for h in p.history]
    want = [t[2] for t in trace if t[0] == r[0]]
    assert max_rel(llks, want) < 1e-12
    reached = lambda t: next((i for i, v in enumerate(t) if v <= c[6]), None)
    assert reached(llks) == reached(want)
    assert llks[-1] < 0.1 * llks[0]
    for name, lead, got in (("rpie_near_run_obj_f64.txt", 3, p.obj - o0),
                            ("rpie_near_run_probe_f64.txt", 4, p.probe - p0)):
        rows = np.array([t for t in table(name) if t[0] == r[0]])
        want = rows[:, lead:lead + 2].ravel()
        g = np.stack([got.real.ravel(), got.imag.ravel()], axis=1).ravel()
        assert max_rel(g, want) < 1e-12, name
