"""LSQML with position offsets (pty-chi PATCH) on the GPU through the C ABI against the FP64
oracle, whose position model and step the reference's fixtures pin
(tests/test_oracle_positions.py): several epochs with the offsets refined, from the fixtures'
truth object with every frame moved by its offset, near field (angular spectrum; angular
spectrum with an evanescent band), three slices, and the far field; the LLK after every epoch,
the object, probe and offsets at the end. c128 to FP64 rounding, c64 to 1e-3. Plus the feature
API: held offsets stay, the algorithms that do not model them refuse them. AP, DM and RAAR
under offsets, refined or held, against the oracle."""

import numpy as np
import pytest

import ptyrs_amd as pa
from conftest import max_rel, table
from oracle.lsqml import POISSON, Lsqml, slice_llk
from test_gpu_rpie import gpu, oracle, rel_l2, small_scan
from test_oracle_lsqml import state as near_state
from test_oracle_multislice import state as multislice_state

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-9, "c64": 1e-3}


def run_oracle(addr, optics, iobs, nb_obs, obj, probe, d0, nepoch, batch, seed, sampling=None):
    run = Lsqml(batch, POISSON, seed, offsets=d0, refine=True, sampling=sampling)
    llks = []
    for _ in range(nepoch):
        obj, probe, _ = run.epoch(addr, iobs, optics, obj, probe)
        llks.append(slice_llk(optics, addr.views(), iobs, addr.nmode, nb_obs, obj, probe,
                              run).poisson)
    return obj, probe, run.offsets, llks


def near_gpu(case, dtype, obj):
    corners = np.array(table("lsqml_near_corners_f64.txt"))
    c = table("lsqml_near_cases_f64.txt")[case]
    counts = np.array([r[2:] for r in table("lsqml_near_counts_f64.txt") if r[0] == case])
    counts = counts.reshape(len(corners), 24, 24).astype(np.float32)
    probe = np.array([r[3] + 1j * r[4] for r in table("lsqml_near_probe_f64.txt")])
    nf = len(corners)
    scan = pa.Scan(counts, np.zeros(nf), np.zeros(nf), c[1], 1.0, 75e-6,
                   corners=corners[:, 1:3].astype(np.int64))
    w = pa.Wavefield([24, 24], c[1], [c[2], c[2]], c[3])
    return pa.GpuPtycho(scan, obj, probe.reshape(2, 24, 24), optics=pa.Optics.free_space(w),
                        dtype=dtype)


def offsets_of(file, case):
    return np.array([[r[2], r[3]] for r in table(file) if r[0] == case])


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("case", [0, 2])
def test_near_field_positions_match_oracle(dtype, case):
    addr, _, probe, cases = near_state()
    start = np.array(table("lsqml_near_pos_start_f64.txt"))
    obj = (start[:, 2] + 1j * start[:, 3]).reshape(1, 40, 40)
    d0 = offsets_of("lsqml_near_pos_offsets_f64.txt", case)
    optics, iobs, nb_obs, _ = cases[case]
    o_obj, o_probe, o_d, o_llk = run_oracle(addr, optics, iobs, nb_obs, obj, probe, d0, 4, 8, 7)
    g = near_gpu(case, dtype, obj)
    if g.sampling != "patch":
        g.set_sampling("patch")
    g.set_features(position=d0)
    g.lsqml(4, 1, 8, "poisson", 7)
    tol = TOL[dtype]
    assert max_rel(g.history["poisson"], o_llk) < tol
    assert rel_l2(g.object, o_obj) < tol * 10
    assert rel_l2(g.probe, o_probe) < tol * 10
    got = g.features["position"]
    assert max_rel((got - d0).ravel(), (o_d - d0).ravel()) < tol * 100


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_multislice_positions_match_oracle(dtype):
    addr, objects, probe, cases = multislice_state()
    start = np.array(table("lsqml_multislice_pos_start_f64.txt"))
    objects = [(start[start[:, 0] == s][:, 3] + 1j * start[start[:, 0] == s][:, 4])
               .reshape(1, 40, 40) for s in range(3)]
    d0 = offsets_of("lsqml_multislice_pos_offsets_f64.txt", 0)
    optics, iobs, nb_obs, _ = cases[0]
    o_obj, o_probe, o_d, o_llk = run_oracle(addr, optics, iobs, nb_obs, list(objects), probe,
                                            d0, 3, 8, 7)
    from test_gpu_multislice import start as ms_start
    g, _, _ = ms_start(0, dtype)
    g.set_object(np.stack(objects))
    g.set_sampling("patch")
    g.set_features(position=d0)
    g.lsqml(3, 1, 8, "poisson", 7)
    tol = TOL[dtype]
    assert max_rel(g.history["poisson"], o_llk) < tol
    assert rel_l2(g.objects, np.stack(o_obj)) < tol * 10
    assert rel_l2(g.probe, o_probe) < tol * 10
    got = g.features["position"]
    assert max_rel((got - d0).ravel(), (o_d - d0).ravel()) < tol * 100


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_multislice_full_object_positions_match_oracle(dtype):
    # Three slices under FULL_OBJECT (every slice read through the batch's sampled windows, the
    # slice directions placed by the sampling's adjoint, the probe unmoved): the reference's
    # lsqml_multislice_pos_full state and offsets, three epochs of batches of 8, refined.
    addr, objects, probe, cases = multislice_state()
    start = np.array(table("lsqml_multislice_pos_start_f64.txt"))
    objects = [(start[start[:, 0] == s][:, 3] + 1j * start[start[:, 0] == s][:, 4])
               .reshape(1, 40, 40) for s in range(3)]
    d0 = offsets_of("lsqml_multislice_pos_full_offsets_f64.txt", 0)
    optics, iobs, nb_obs, _ = cases[0]
    o_obj, o_probe, o_d, o_llk = run_oracle(addr, optics, iobs, nb_obs, list(objects), probe,
                                            d0, 3, 8, 7, sampling="full_object")
    from test_gpu_multislice import start as ms_start
    g, _, _ = ms_start(0, dtype)
    g.set_object(np.stack(objects))
    g.set_sampling("full_object")
    g.set_features(position=d0)
    g.lsqml(3, 1, 8, "poisson", 7)
    tol = TOL[dtype]
    print("multislice full object", dtype, max_rel(g.history["poisson"], o_llk),
          rel_l2(g.objects, np.stack(o_obj)), rel_l2(g.probe, o_probe))
    assert max_rel(g.history["poisson"], o_llk) < tol
    assert rel_l2(g.objects, np.stack(o_obj)) < tol * 10
    assert rel_l2(g.probe, o_probe) < tol * 10
    got = g.features["position"]
    assert max_rel((got - d0).ravel(), (o_d - d0).ravel()) < tol * 100


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_far_field_positions_match_oracle(dtype):
    d = small_scan(n=32, nprobe=2, npos=20)
    o = oracle(d)
    rng = np.random.default_rng(11)
    d0 = rng.uniform(-0.6, 0.6, (20, 2))
    o_obj, o_probe, o_d, o_llk = run_oracle(o.addr, o.optics, o.iobs, o.nb_obs, o.obj, o.probe,
                                            d0, 3, 7, 5)
    g = gpu(d, dtype)
    assert g.sampling == "patch"
    g.set_features(position=d0)
    g.lsqml(3, 1, 7, "poisson", 5)
    tol = TOL[dtype]
    assert max_rel(g.history["poisson"], o_llk) < tol
    assert rel_l2(g.object, o_obj) < tol * 10
    assert rel_l2(g.probe, o_probe) < tol * 10
    got = g.features["position"]
    assert max_rel((got - d0).ravel(), (o_d - d0).ravel()) < tol * 100


def test_position_feature_api():
    d = small_scan(n=32, nprobe=1, npos=12)
    g = gpu(d, "c64")
    assert g.features["position"] is None
    d0 = np.linspace(-0.4, 0.4, 24).reshape(12, 2)
    g.set_features(position=d0, hold=["position"])
    f = g.features
    assert np.array_equal(f["position"], d0) and f["hold"] == ["position"]
    g.lsqml(1, 1, 6, "poisson", 1)
    assert np.array_equal(g.features["position"], d0)
    g.rpie(1, batch_size=5)  # held offsets: rPIE models them and leaves them
    assert np.array_equal(g.features["position"], d0)
    g.ml(1)  # ML models held offsets too
    assert np.array_equal(g.features["position"], d0)
    for name in ("ap", "dm", "raar"):   # held offsets: the projections model them
        getattr(g, name)(1)
        assert np.array_equal(g.features["position"], d0)
    g.set_sampling("full_object")   # held offsets under FullObject: every algorithm runs
    g.rpie(1, batch_size=5)
    g.lsqml(1, 1, 6, "poisson", 1)
    g.ml(1)
    for name in ("ap", "dm", "raar"):
        getattr(g, name)(1)
    assert np.array_equal(g.features["position"], d0)
    g.set_sampling("patch")
    with pytest.raises(ValueError):
        g.set_features(position=d0[:5])
    with pytest.raises(ValueError):
        g.set_features(scale=np.ones(12))
    g.set_features()
    assert g.features["position"] is None
    g.rpie(1)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("nprobe,batch,refine", [(1, 7, True), (3, 30, True), (2, 5, False)])
def test_rpie_positions_match_oracle(dtype, nprobe, batch, refine):
    # rPIE under PATCH offsets (ptyrs/src/engine/ptycho.rs:900-1086 with Features::position):
    # far field, masked pixels, 1-3 modes, refined or held, two epochs with LLK records.
    d = small_scan(n=64, nprobe=nprobe, npos=30, seed=11)
    d0 = np.random.default_rng(12).uniform(-0.4, 0.4, (30, 2))
    o = oracle(d)
    o.set_position(d0, refine=refine)
    kw = dict(batch_size=batch, obj_alpha=0.2, probe_alpha=0.3, calc_llk=1, seed=5)
    o.rpie(2, **kw)
    g = gpu(d, dtype)
    g.set_features(position=d0, hold=() if refine else ("position",))
    g.rpie(2, **kw)
    tol = TOL[dtype]
    want = [r[2].poisson for r in o.history]
    print("rpie positions", dtype, nprobe, max_rel(g.history["poisson"], want),
          rel_l2(g.object - d["obj"], o.obj - d["obj"]), rel_l2(g.probe, o.probe))
    assert max_rel(g.history["poisson"], want) < tol
    assert rel_l2(g.object - d["obj"], o.obj - d["obj"]) < tol * 10
    assert rel_l2(g.probe, o.probe) < tol * 10
    got = g.features["position"]
    if refine:
        assert max_rel((got - d0).ravel(), (o.offsets - d0).ravel()) < tol * 100
    else:
        assert np.array_equal(got, d0)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("nprobe,refine,up", [(1, True, True), (3, True, False), (2, False, True)])
def test_ml_positions_match_oracle(dtype, nprobe, refine, up):
    # ML under PATCH offsets (ptyrs/src/engine/ptycho.rs:757-842 with Features::position):
    # far field, masked pixels, 1-3 modes, held or refined, three cycles with LLK records.
    d = small_scan(n=64, nprobe=nprobe, npos=30, seed=13)
    d0 = np.random.default_rng(14).uniform(-0.4, 0.4, (30, 2))
    o = oracle(d)
    o.set_position(d0, refine=refine)
    kw = dict(update_object=True, update_probe=up, calc_llk=1)
    o.ml(3, **kw)
    g = gpu(d, dtype)
    g.set_features(position=d0, hold=() if refine else ("position",))
    g.ml(3, **kw)
    tol = TOL[dtype]
    want = [r[2].poisson for r in o.history]
    print("ml positions", dtype, nprobe, max_rel(g.history["poisson"], want),
          rel_l2(g.object, o.obj), rel_l2(g.probe, o.probe))
    assert max_rel(g.history["poisson"], want) < tol
    assert rel_l2(g.object, o.obj) < tol * 10
    assert rel_l2(g.probe, o.probe) < tol * 10
    got = g.features["position"]
    if refine:
        assert max_rel((got - d0).ravel(), (o.offsets - d0).ravel()) < tol * 100
    else:
        assert np.array_equal(got, d0)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("name,nprobe,refine,kw", [
    ("ap", 1, True, dict(probe_inertia=0.01)), ("ap", 3, False, dict(probe_inertia=0.01)),
    ("ap", 2, True, {}), ("dm", 2, True, {}), ("dm", 1, False, dict(loop_obj_probe=2)),
    ("raar", 2, True, {}), ("raar", 3, True, dict(loop_obj_probe=2))])
def test_projections_with_positions_match_oracle(dtype, name, nprobe, refine, kw):
    # AP, DM and RAAR under PATCH offsets (ptyrs/src/engine/ptycho.rs:605-755 with
    # Shifted::psi_to_obj_probe): far field, masked pixels, zero-mean offsets within 0.4 px,
    # refined or held, three cycles with LLK records (the oracle pins the moved waves and the
    # position sums by the reference's fixtures and the projections by the zero-offset runs).
    d = small_scan(n=32, nprobe=nprobe, npos=25, seed=31)
    d0 = np.random.default_rng(32).uniform(-0.4, 0.4, (25, 2))
    d0 -= d0.mean(axis=0)
    o = oracle(d)
    o.set_position(d0, refine=refine)
    params = dict(calc_llk=1, **kw)
    getattr(o, name)(3, **params)
    g = gpu(d, dtype)
    g.set_features(position=d0, hold=() if refine else ("position",))
    getattr(g, name)(3, **params)
    tol = TOL[dtype]
    want = [r[2].poisson for r in o.history]
    print(name, nprobe, refine, kw, dtype, max_rel(g.history["poisson"], want),
          rel_l2(g.object, o.obj), rel_l2(g.probe, o.probe))
    assert max_rel(g.history["poisson"], want) < tol
    assert rel_l2(g.object, o.obj) < tol * 10
    assert rel_l2(g.probe, o.probe) < tol * 10
    got = g.features["position"]
    if refine:
        assert max_rel((got - d0).ravel(), (o.offsets - d0).ravel()) < tol * 100
    else:
        assert np.array_equal(got, d0)
