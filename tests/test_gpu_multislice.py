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


def deep_state(nslice=8, n=64, npos=40, seed=61):
    """An eight-slice far-field scan with 2 probe modes and Poisson counts of its own model."""
    from ptyrs_amd import sim
    pixel = 10e-9
    lam = 1.2398e-10
    z = pixel * n * 75e-6 / lam
    probe = sim.hermite_gauss_probe(n, powers=(0.8, 0.2)) * 100.0
    y, x = sim.fermat_positions(npos, 6.0, pixel)
    phi = np.angle(sim.random_object(192, 192, seed, amp=(1.0, 1.0), phase=(0.0, 0.2)))
    objects = np.stack([np.exp(1j * np.roll(phi, (5 * s, 7 * s), axis=(-2, -1))
                               - 0.05 * np.roll(phi, (5 * s, 7 * s), axis=(-2, -1)))
                        for s in range(nslice)])
    optics = pa.Optics.far_field(lam, z, 75e-6, [n, n]).with_slices(
        [pa.Wavefield([n, n], lam, [pixel, pixel], 1e-6)] * (nslice - 1))
    zero = np.zeros((npos, n, n), dtype=np.float32)
    g = pa.GpuPtycho(pa.Scan(zero, y, x, lam, z, 75e-6), objects, probe, optics=optics,
                     dtype="c128")
    counts = np.random.default_rng(seed).poisson(g.intensities()).astype(np.float32)
    start = objects * np.exp(1j * np.random.default_rng(seed + 1).uniform(-0.1, 0.1,
                                                                         objects.shape))
    return pa.Scan(counts, y, x, lam, z, 75e-6), start, probe * 0.9, optics


@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_checkpointed_waves_are_bit_identical(dtype):
    """Keeping every 2nd or 3rd slice's incident wave and recomputing the rest in the adjoint
    gives the same bits as keeping them all: the gradient, and LSQML with refined positions on
    eight slices (and the three-slice fixture case), with fewer wave bytes held."""
    scan, start, probe, optics = deep_state()
    d0 = np.random.default_rng(5).normal(0.0, 1.0, (scan.counts.shape[0], 2))
    out, held = {}, {}
    for every in (0, 2, 3):
        g = pa.GpuPtycho(scan, start, probe, optics=optics, dtype=dtype)
        g.set_checkpoint(every)
        v, go, gp = g.grad("poisson")
        held[every] = g.checkpoint
        g.set_features(position=d0)
        g.lsqml(3, 1, 16, "poisson", 7)
        out[every] = (v, go, gp, g.objects, g.probe, g.features["position"],
                      g.history["poisson"])
    for every in (2, 3):
        for a, b in zip(out[every], out[0]):
            assert np.array_equal(a, b), every
        assert held[every][0] == every
    # Eight slices: 7 incident wave sets kept without checkpoints, 4 with every 2 or 3.
    assert held[2][1] * 7 == held[0][1] * 4 and held[3][1] == held[2][1]


@pytest.mark.parametrize("every", [2, 3])
def test_checkpointed_fixture_case_is_bit_identical(every):
    runs = []
    for k in (0, every):
        g, _, _ = start(0, "c64")
        g.set_checkpoint(k)
        g.lsqml(5, 1, BATCH, "poisson", 7)
        runs.append((g.objects, g.probe, g.history["poisson"]))
    for a, b in zip(*runs):
        assert np.array_equal(a, b)
