"""Multislice LSQML on the GPU through the C ABI against the reference's three-slice fixtures
(ptyrs/tests/engine_lsqml_multislice.rs:478-541): 30 epochs of Poisson LSQML in batches of 8
(seed 7) under an angular-spectrum exit (case 0), a Fresnel one (case 1) and the far field
(case 2); the LLK after every epoch, every step of the last epoch (a_far, a_obj of every slice,
a_probe) and the changes of every slice and the probe. c128 to FP64 rounding, c64 to 1e-3."""

import numpy as np
import pytest

import ptyrs_amd as pa
from conftest import max_rel, table
from test_gpu_rpie import rel_l2

pytestmark = pytest.mark.gpu

SHAPE, OBJ, NPROBE, NSLICE, BATCH = (24, 24), (40, 40), 2, 3, 8


def start(case, dtype):
    corners = np.array(table("lsqml_multislice_corners_f64.txt"))
    rows = np.array(table("lsqml_multislice_obj_f64.txt"))
    objects = np.stack([(rows[rows[:, 0] == s][:, 3] + 1j * rows[rows[:, 0] == s][:, 4])
                        .reshape(1, *OBJ) for s in range(NSLICE)])
    rows = np.array(table("lsqml_multislice_probe_f64.txt"))
    probe = (rows[:, 3] + 1j * rows[:, 4]).reshape(NPROBE, *SHAPE)
    c = table("lsqml_multislice_cases_f64.txt")[case]
    counts = np.array([r[2:] for r in table("lsqml_multislice_counts_f64.txt") if r[0] == case])
    counts = counts.reshape(len(corners), *SHAPE).astype(np.float32)
    nf = len(corners)
    if c[4] == 0.0:
        dpix = c[1] * c[3] / (SHAPE[0] * c[2])
        optics = pa.Optics.far_field(c[1], c[3], dpix, list(SHAPE))
        pixel = [c[1] * c[3] / (SHAPE[0] * dpix)] * 2
    else:
        pixel = [c[2], c[2]]
        optics = pa.Optics.free_space(pa.Wavefield(list(SHAPE), c[1], pixel, c[3]))
    optics = optics.with_slices([pa.Wavefield(list(SHAPE), c[1], pixel, d) for d in c[7:]])
    scan = pa.Scan(counts, np.zeros(nf), np.zeros(nf), c[1], c[3], 75e-6,
                   corners=corners[:, 1:3].astype(np.int64))
    g = pa.GpuPtycho(scan, objects, probe, optics=optics, dtype=dtype)
    return g, objects, probe


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("case", range(3))
def test_multislice_lsqml_matches_reference_fixtures(dtype, case):
    g, objects, probe = start(case, dtype)
    assert g.objects.shape == (NSLICE, 1, *OBJ)
    want = [r[2] for r in table("lsqml_multislice_llk_f64.txt") if r[0] == case]
    nepoch = len(want) - 1
    llks = [g.llk()["poisson"]]
    g.lsqml(nepoch, 1, BATCH, "poisson", 7)
    llks += list(g.history["poisson"])
    tol = {"c128": 1e-9, "c64": 1e-3}[dtype]
    assert max_rel(llks, want) < tol
    rows = [r for r in table("lsqml_multislice_epochs_steps_f64.txt")
            if r[0] == case and r[1] == nepoch - 1]
    steps = g.lsqml_steps()
    assert [int(f) for s in steps for f in s["frames"]] == [int(r[2]) for r in rows]
    got_far = np.concatenate([s["a_far"] for s in steps])
    assert max_rel(got_far, [r[3] for r in rows]) < tol * 10
    a_obj = np.concatenate([s["a_obj"] for s in steps], axis=1)
    for s in range(NSLICE):
        assert max_rel(a_obj[s], [r[4 + s] for r in rows]) < tol * 10, s
    got_p = np.concatenate([s["a_probe"] for s in steps])
    assert max_rel(got_p, [r[4 + NSLICE] for r in rows]) < tol * 10
    r = np.array([t for t in table("lsqml_multislice_epochs_obj_f64.txt") if t[0] == case])
    assert rel_l2((g.objects - objects).ravel(), r[:, 4] + 1j * r[:, 5]) < tol * 10
    r = np.array([t for t in table("lsqml_multislice_epochs_probe_f64.txt") if t[0] == case])
    assert rel_l2((g.probe - probe).ravel(), r[:, 4] + 1j * r[:, 5]) < tol * 10


def test_multislice_runs_lsqml_and_llk_only():
    """The reference's GPU engine runs a multislice object through LSQML, BH and the LLK pass
    alone (ptyrs/src/engine/gpu/device.rs:3100-3105)."""
    g, objects, _ = start(0, "c64")
    for name in ("rpie", "ap", "dm", "raar", "ml"):
        with pytest.raises(ValueError, match="slice"):
            getattr(g, name)(1)
    assert np.array_equal(g.objects, objects.astype(np.complex64).astype(np.complex128))
    with pytest.raises(ValueError):
        g.set_object(objects[0])


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_multislice_forward_matches_ptychi(dtype):
    """The GPU model's detector intensities against pty-chi's forward-model output `y_pred` at
    the start, frames 0..4 of every case (ptyrs/tests/engine_lsqml_multislice.rs:146-182)."""
    tol = {"c128": 1e-12, "c64": 1e-5}[dtype]
    for case in range(3):
        g, _, _ = start(case, dtype)
        got = g.intensities()
        for r in table("lsqml_multislice_icalc_f64.txt"):
            if r[0] != case:
                continue
            want = np.array(r[2:]).reshape(24, 24)
            assert max_rel(got[int(r[1])].ravel(), want.ravel()) < tol, (case, r[1])
