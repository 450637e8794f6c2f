"""rPIE on the GPU through the C ABI against the FP64 oracle (same seeded inputs), and against
the reference's own near-field fixtures. c128 runs must match to FP64 rounding; c64 runs to
the reference's GPU-vs-CPU bound, 1e-4 relative (ptyrs/tests/engine_gpu.rs:128-164)."""

import math

import numpy as np
import pytest

import ptyrs_amd as pa
from conftest import max_rel, table
from oracle.optics import Optics, Wavefield
from oracle.ptycho import Ptycho, far_field_ptycho
from ptyrs_amd import sim

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-9, "c64": 1e-4}


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def small_scan(n=64, nprobe=2, npos=30, photons=2e5, seed=3, masked=True):
    """A far-field scan: Fermat positions n/8 px apart, a random object, Hermite–Gauss modes,
    Poisson counts, rows and columns masked; start from a flat object."""
    pixel = sim.pixel_size(n)
    y, x = sim.fermat_positions(npos, n / 8, pixel)
    ext = int(np.ceil(np.abs(np.r_[y - y.mean(), x - x.mean()] / pixel).max())) + 2
    side = (2 * ext + n + 8 + 3) // 4 * 4
    obj = sim.random_object(side, side, seed)
    probe = sim.hermite_gauss_probe(n, powers=([0.8, 0.2, 0.05, 0.02] + [0.01] * 6)[:nprobe])
    cy, cx = sim.corners(y, x, pixel, obj.shape[-2:], n)
    I = sim.intensities(obj, probe, cy, cx, sim.quad_phase(n))
    s = photons / I.sum(axis=(1, 2)).mean()
    counts = np.random.default_rng(seed).poisson(I * s).astype(np.float32)
    if masked:
        counts[:, 3, :] = -1
        counts[:, :, n // 3] = -1
    # A random-phase start, as the reference's scenario (a flat start is an unstable point of
    # the iteration: a 1e-7 change of it moves one epoch's result by percents).
    start = np.exp(1j * np.random.default_rng(seed + 1).uniform(-1.0, 1.0, obj.shape))
    return dict(counts=counts, y=y, x=x, obj=start, probe=probe * math.sqrt(s) * 0.9)


def gpu(d, dtype, **kw):
    scan = pa.Scan(d["counts"], d["y"], d["x"], sim.WAVELENGTH, sim.DISTANCE, sim.DETECTOR_PIXEL)
    return pa.GpuPtycho(scan, d["obj"], d["probe"], dtype=dtype, **kw)


def oracle(d):
    c = np.where(d["counts"] < 0, -1.0, d["counts"]).astype(np.float64)
    return far_field_ptycho(c, d["y"], d["x"], d["obj"], d["probe"], sim.WAVELENGTH,
                            sim.DISTANCE, sim.DETECTOR_PIXEL)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n,nprobe", [(64, 1), (64, 2), (32, 3), (128, 4)])
def test_llk_matches_oracle(dtype, n, nprobe):
    d = small_scan(n=n, nprobe=nprobe, npos=12)
    g, o = gpu(d, dtype), oracle(d)
    assert np.array_equal(g.corners, np.c_[o.addr.oy, o.addr.ox])
    got, want = g.llk(), o.llk()
    for k in ("poisson", "gaussian", "euclidian", "nb_photons_calc"):
        assert abs(got[k] - want.as_dict()[k]) <= TOL[dtype] * abs(want.as_dict()[k]), k


@pytest.mark.parametrize("n", [16, 24, 48, 96, 256, 512])
def test_every_frame_side_matches_oracle_llk(n):
    d = small_scan(n=n, nprobe=2, npos=6)
    g, o = gpu(d, "c128"), oracle(d)
    got, want = g.llk(), o.llk()
    assert abs(got["poisson"] - want.poisson) <= 1e-9 * abs(want.poisson)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("batch", [1, 7, 30])
def test_rpie_far_field_matches_oracle(dtype, batch):
    d = small_scan(n=64, nprobe=2, npos=30)
    g, o = gpu(d, dtype), oracle(d)
    o0, p0 = o.obj.copy(), o.probe.copy()
    kw = dict(batch_size=batch, calc_llk=1, obj_step=0.9, probe_step=0.8, obj_alpha=0.15,
              probe_alpha=0.05, seed=11)
    g.rpie(4, **kw)
    o.rpie(4, **kw)
    assert g.cycle == o.cycle == 4
    h = g.history
    want = [r[2].poisson for r in o.history]
    assert list(h["cycle"]) == [r[0] for r in o.history]
    assert max_rel(h["poisson"], want) < TOL[dtype]
    assert rel_l2(g.object - o0, o.obj - o0) < TOL[dtype] * 10
    assert rel_l2(g.probe - p0, o.probe - p0) < TOL[dtype] * 10


def test_rpie_fixed_arrays_and_split_calls():
    d = small_scan(n=64, nprobe=2, npos=20)
    g, o = gpu(d, "c128"), oracle(d)
    p0 = o.probe.copy()
    g.rpie(2, probe_step=None, batch_size=6)
    g.rpie(1, probe_step=None, batch_size=6)
    o.rpie(3, probe_step=None, batch_size=6)
    assert np.array_equal(g.probe, p0)
    assert rel_l2(g.object, o.obj) < 1e-9
    o1 = g.object
    g.rpie(1, obj_step=None, batch_size=6)
    o.rpie(1, obj_step=None, batch_size=6)
    assert np.array_equal(g.object, o1)
    assert rel_l2(g.probe, o.probe) < 1e-9


def _near_start(c):
    shape, obj_shape, nprobe = (24, 24), (40, 40), 2
    corners = np.array(table("rpie_near_corners_f64.txt"))
    pixel = c[2]
    at = [(corners[:, 1 + a] - (obj_shape[a] - shape[a]) // 2) * pixel for a in (0, 1)]
    counts = np.array([r[2:] for r in table("rpie_near_counts_f64.txt") if r[0] == c[0]])
    counts = counts.reshape(len(corners), *shape)
    obj = np.array([r[2] + 1j * r[3] for r in table("rpie_near_obj_f64.txt")]).reshape(1, *obj_shape)
    probe = np.array([r[3] + 1j * r[4] for r in table("rpie_near_probe_f64.txt")])
    return counts, at, obj, probe.reshape(nprobe, *shape), corners


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("run", range(9))
def test_rpie_near_field_matches_reference_fixtures(dtype, run):
    """The reference's rPIE under an angular-spectrum exit (case 0) and a Fresnel one forward
    (case 1) and back (case 2), batches 1, 4 and 16, 30 epochs
    (ptyrs/tests/engine_rpie_near.rs:102-169)."""
    r = table("rpie_near_runs_f64.txt")[run]
    c = table("rpie_near_cases_f64.txt")[int(r[1])]
    counts, at, obj, probe, corners = _near_start(c)
    w = pa.Wavefield([24, 24], c[1], [c[2], c[2]], c[3])
    scan = pa.Scan(counts.astype(np.float32), at[0], at[1], c[1], 1.0, 75e-6)
    g = pa.GpuPtycho(scan, obj, probe, optics=pa.Optics.free_space(w), dtype=dtype)
    assert np.array_equal(g.corners, corners[:, 1:3])
    llks = [g.llk()["poisson"]]
    g.rpie(int(r[3]), obj_step=r[4], probe_step=r[5], obj_alpha=r[6], probe_alpha=r[7],
           batch_size=int(r[2]), calc_llk=1)
    llks += list(g.history["poisson"])
    want = [t[2] for t in table("rpie_near_llk_f64.txt") if t[0] == r[0]]
    tol = {"c128": 1e-9, "c64": 1e-3}[dtype]
    assert max_rel(llks, want) < tol
    for name, lead, got in (("rpie_near_run_obj_f64.txt", 3, g.object - obj),
                            ("rpie_near_run_probe_f64.txt", 4, g.probe - probe)):
        rows = np.array([t for t in table(name) if t[0] == r[0]])
        want = rows[:, lead] + 1j * rows[:, lead + 1]
        assert rel_l2(got.ravel(), want) < tol * 10, name


def test_banded_angular_spectrum_runs_lsqml_and_llk_only():
    """A transfer function with an evanescent band is not inverted by its adjoint: rPIE, AP, DM,
    RAAR and ML refuse it, LSQML and the LLK pass run it (ptyrs/src/engine/gpu/device.rs:
    3087-3105). The LSQML fixtures' case 2 has such a band."""
    c = table("lsqml_near_cases_f64.txt")[2]
    corners = np.array(table("lsqml_near_corners_f64.txt"))
    counts = np.array([r[2:] for r in table("lsqml_near_counts_f64.txt") if r[0] == 2])
    counts = counts.reshape(len(corners), 24, 24).astype(np.float32)
    obj = np.array([r[2] + 1j * r[3] for r in table("lsqml_near_obj_f64.txt")]).reshape(1, 40, 40)
    probe = np.array([r[3] + 1j * r[4] for r in table("lsqml_near_probe_f64.txt")])
    nf = len(corners)
    scan = pa.Scan(counts, np.zeros(nf), np.zeros(nf), c[1], 1.0, 75e-6,
                   corners=corners[:, 1:3].astype(np.int64))
    w = pa.Wavefield([24, 24], c[1], [c[2], c[2]], c[3])
    g = pa.GpuPtycho(scan, obj, probe.reshape(-1, 24, 24), optics=pa.Optics.free_space(w))
    for name in ("rpie", "ap", "dm", "raar", "ml"):
        with pytest.raises(ValueError, match="unitary"):
            getattr(g, name)(1)
    assert np.array_equal(g.object, obj)
    assert np.isfinite(g.llk()["poisson"])
    g.lsqml(1, 1, 8, "poisson", 7)
    assert len(g.history["poisson"]) == 1


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("nprobe", [1, 3, 4])
@pytest.mark.parametrize("n", [16, 24, 32, 48, 64, 96, 128, 256, 512])
def test_rpie_every_side_and_mode_count_matches_oracle(n, nprobe, dtype):
    """Two rPIE epochs in batches of 4 at every frame side and 1, 3 or 4 modes: the launch
    shapes of every (side, modes, precision) the passes instantiate."""
    d = small_scan(n=n, nprobe=nprobe, npos=6, masked=False)
    g, o = gpu(d, dtype), oracle(d)
    o0, p0 = o.obj.copy(), o.probe.copy()
    g.rpie(2, batch_size=4, calc_llk=1)
    o.rpie(2, batch_size=4, calc_llk=1)
    assert max_rel(g.history["poisson"], [r[2].poisson for r in o.history]) < TOL[dtype]
    assert rel_l2(g.object - o0, o.obj - o0) < TOL[dtype] * 10
    assert rel_l2(g.probe - p0, o.probe - p0) < TOL[dtype] * 10
