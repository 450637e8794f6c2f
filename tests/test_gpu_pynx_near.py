"""AP, DM, RAAR and ML on the GPU through the C ABI under an angular-spectrum exit, against the
reference's PyNX near-field fixtures (ptyrs/tests/engine_pynx_near.rs): the four LLK columns of
every record, the final object and probe. Exact-kernel runs hold c128 to FP64 rounding over 30
cycles; paraxial runs to the reference's 3e-5 model bound; c64 to 1e-3."""

import numpy as np
import pytest

import ptyrs_amd as pa
from conftest import max_rel, table, worst_of

pytestmark = pytest.mark.gpu

SHAPE, OBJ, NPROBE, AP_BEFORE_ML = (24, 24), (40, 40), 2, 10
COLUMNS = ("poisson", "gaussian", "euclidian", "nb_photons_calc")


def start(dtype):
    c = table("pynx_near_case_f64.txt")[0]
    corners = np.array(table("pynx_near_corners_f64.txt"))
    pixel = c[1]
    at = [(corners[:, 1 + a] - (OBJ[a] - SHAPE[a]) // 2) * pixel for a in (0, 1)]
    counts = np.array([r[1:] for r in table("pynx_near_counts_f64.txt")])
    counts = counts.reshape(len(corners), *SHAPE).astype(np.float32)
    w = pa.Wavefield(list(SHAPE), c[0], [pixel, pixel], c[2])
    obj = np.array([r[2] + 1j * r[3] for r in table("pynx_near_obj_f64.txt")]).reshape(1, *OBJ)
    probe = np.array([r[3] + 1j * r[4] for r in table("pynx_near_probe_f64.txt")])
    scan = pa.Scan(counts, at[0], at[1], c[0], 1.0, 75e-6)
    g = pa.GpuPtycho(scan, obj, probe.reshape(NPROBE, *SHAPE), optics=pa.Optics.free_space(w),
                     dtype=dtype)
    assert np.array_equal(g.corners, corners[:, 1:3])
    return g


def engine_run(dtype, algorithm, cycles):
    """ptyrs/tests/engine_pynx_near.rs:103-152"""
    g = start(dtype)
    ap = dict(obj_inertia=0.01, probe_inertia=0.001, calc_llk=1, zero_phase_ramp=False)
    if algorithm == 0:
        g.ap(cycles, **ap)
    elif algorithm == 1:
        g.dm(cycles, calc_llk=1, zero_phase_ramp=False)
    elif algorithm == 2:
        g.raar(cycles, calc_llk=1, zero_phase_ramp=False)
    else:
        g.ap(AP_BEFORE_ML, **dict(ap, calc_llk=0))
        g.ml(cycles, update_probe=True, calc_llk=1, zero_phase_ramp=False)
    h = g.history
    llk = np.stack([h[k] for k in COLUMNS], axis=1)
    if algorithm > 0:
        last = g.llk()
        llk = np.concatenate([llk[1:], [[last[k] for k in COLUMNS]]])
    assert len(llk) == cycles
    return llk, g


def unit(a):
    a = a.ravel()
    s = np.abs(a).sum()
    return np.stack([a.real / s, a.imag / s], axis=1).ravel()


def flat(a):
    return np.stack([a.ravel().real, a.ravel().imag], axis=1).ravel()


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("run", range(8))
def test_ap_dm_raar_ml_match_pynx_near_field(dtype, run):
    r = table("pynx_near_runs_f64.txt")[run]
    exact, a, cycles = r[1] == 1.0, int(r[2]), int(r[3])
    got, g = engine_run(dtype, a, cycles)
    want = np.array([t[2:6] for t in table("pynx_near_llk_f64.txt") if t[0] == r[0]])
    llk_err = worst_of(worst_of(np.abs(got[:, k] - want[:, k]) / np.abs(want[:, k]))
                       for k in range(4))
    wo = np.array([t[3] + 1j * t[4] for t in table("pynx_near_run_obj_f64.txt") if t[0] == r[0]])
    wp = np.array([t[4] + 1j * t[5] for t in table("pynx_near_run_probe_f64.txt") if t[0] == r[0]])
    # DM and RAAR leave the object-probe scale free: compare shapes at unit L1 norm.
    norm = unit if a in (1, 2) else flat
    errs = [max_rel(norm(g.object), norm(wo)), max_rel(norm(g.probe), norm(wp))]
    if dtype == "c64":
        bound = 1e-3
    else:
        bound = 1e-9 if exact else 3e-5
    for name, e in zip(("LLK", "object", "probe"), [llk_err] + errs):
        assert e < bound, (name, e)
