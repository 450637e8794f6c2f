"""rPIE, ML, LSQML, the LLK and the model intensities under FullObject sampling on the GPU
(pty-chi's FULL_OBJECT shift operator: each view's window Fourier-interpolated from the whole
object plane at its corner plus its frame's offset, ptyrs/src/engine/sampling.rs): rPIE against
the reference's own fixtures on its seeded far-field scenario (`rpie_full_object_*`,
ptyrs/tests/engine_rpie_pos.rs:178-226, one of them on an odd-sided 75 x 79 plane); rPIE and
LSQML against the FP64 oracle (whose FullObject step the reference's `rpie_full_object_*`,
`lsqml_near_pos_full_*` and `lsqml_multislice_pos_full_*` fixtures pin) on far- and near-field
scans (masked pixels, 1-3 modes, held and refined offsets, both LSQML costs). complex128 to
FP64 rounding; complex64 to the reference's GPU-vs-CPU bound on LLK records
(ptyrs/tests/engine_gpu.rs:128-164), 1e-3 on the fixtures' changes and LSQML's runs."""

import numpy as np
import pytest

import ptyrs_amd as pa
from conftest import max_rel, table
from oracle import scenario as sc
from oracle.base import to_natural
from oracle.lsqml import POISSON, Lsqml, slice_llk
from oracle.optics import Optics, Wavefield
from oracle.ptycho import Ptycho, _mode_sum
from test_gpu_positions import TOL as PTOL, near_gpu, offsets_of
from test_gpu_rpie import TOL, _near_start, gpu, oracle, rel_l2, small_scan
from test_gpu_scenario import RPIE_POS, start
from test_oracle_lsqml import state as near_state

pytestmark = pytest.mark.gpu


def scenario_engine(dtype, odd=False, **kw):
    counts, y, x, obj, probe, _ = start(2)
    if odd:
        obj = obj[:, 1:, 1:].copy()
    g = pa.GpuPtycho(pa.Scan(counts, y, x, *sc.GEOMETRY), obj, probe, dtype=dtype, **kw)
    g.set_sampling("full_object")
    return g, obj, probe


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("case", ["step", "epochs", "odd"])
def test_rpie_full_object_matches_ptychi(dtype, case):
    params, nepoch = RPIE_POS["step" if case == "odd" else case]
    rows = np.array(table(f"rpie_full_object_{case}_offsets_f64.txt"))
    off0 = rows[:, 1:3]
    g, obj0, probe0 = scenario_engine(dtype, odd=case == "odd")
    assert g.sampling == "full_object"
    g.set_features(position=off0)
    g.rpie(nepoch, **params)
    tol = {"c128": 1e-9, "c64": 1e-3}[dtype]
    for got, file, lead in ((g.object - obj0, f"rpie_full_object_{case}_obj_f64.txt", 2),
                            (g.probe - probe0, f"rpie_full_object_{case}_probe_f64.txt", 3)):
        r = np.array(table(file))
        gg = np.stack([got.real.ravel(), got.imag.ravel()], axis=1).ravel()
        err = max_rel(gg, r[:, lead:lead + 2].ravel())
        print("rpie full object", case, dtype, file, err)
        assert err < tol, file
    err = max_rel((g.features["position"] - off0).ravel(), (rows[:, 3:5] - off0).ravel())
    print("rpie full object offsets", case, dtype, err)
    assert err < tol
    assert len(g.history["cycle"]) == nepoch


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n,nprobe,batch,refine", [(64, 1, 7, True), (32, 3, 30, True),
                                                   (64, 2, 4, False)])
def test_rpie_full_object_matches_oracle(dtype, n, nprobe, batch, refine):
    d = small_scan(n=n, nprobe=nprobe, npos=30, seed=13)
    d0 = np.random.default_rng(14).uniform(-0.45, 0.45, (30, 2))
    o = oracle(d)
    o.set_sampling("full_object")
    o.set_position(d0, refine=refine)
    kw = dict(batch_size=batch, obj_alpha=0.2, probe_alpha=0.3, calc_llk=1, seed=5)
    o.rpie(2, **kw)
    g = gpu(d, dtype)
    g.set_sampling("full_object")
    g.set_features(position=d0, hold=() if refine else ("position",))
    g.rpie(2, **kw)
    tol = TOL[dtype]
    want = [r[2].poisson for r in o.history]
    print("rpie full object vs oracle", dtype, n, nprobe, d["obj"].shape,
          max_rel(g.history["poisson"], want), rel_l2(g.object - d["obj"], o.obj - d["obj"]),
          rel_l2(g.probe, o.probe))
    assert max_rel(g.history["poisson"], want) < tol
    assert rel_l2(g.object - d["obj"], o.obj - d["obj"]) < tol * 10
    assert rel_l2(g.probe, o.probe) < tol * 10
    got = g.features["position"]
    if refine:
        assert max_rel((got - d0).ravel(), (o.offsets - d0).ravel()) < tol * 100
    else:
        assert np.array_equal(got, d0)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("case", [0, 1, 2])
def test_rpie_full_object_near_field_matches_oracle(dtype, case):
    # The reference's near-field fixture state (angular spectrum, Fresnel forward and back)
    # from its object with a random phase added (at the fixture's own object, the truth, the
    # position gradients vanish and the offsets' change is rounding: a 1e-13 change of the
    # start moves it by 3e-6 to 3e-5 in the oracle), refined offsets, two epochs in batches
    # of 4.
    c = table("rpie_near_cases_f64.txt")[case]
    counts, at, obj, probe, _ = _near_start(c)
    obj = obj * np.exp(0.2j * np.random.default_rng(16).uniform(-1, 1, obj.shape))
    d0 = np.random.default_rng(15).uniform(-0.4, 0.4, (len(counts), 2))
    w = Wavefield([24, 24], c[1], [c[2], c[2]], c[3])
    o = Ptycho(counts.astype(np.float64), at[0], at[1], obj, probe, Optics.free_space(w))
    o.set_sampling("full_object")
    o.set_position(d0, refine=True)
    kw = dict(batch_size=4, obj_alpha=0.3, probe_alpha=0.2, calc_llk=1, seed=3)
    o.rpie(2, **kw)
    pw = pa.Wavefield([24, 24], c[1], [c[2], c[2]], c[3])
    scan = pa.Scan(counts.astype(np.float32), at[0], at[1], c[1], 1.0, 75e-6)
    g = pa.GpuPtycho(scan, obj, probe, optics=pa.Optics.free_space(pw), dtype=dtype)
    g.set_sampling("full_object")
    g.set_features(position=d0)
    g.rpie(2, **kw)
    tol = TOL[dtype]
    want = [r[2].poisson for r in o.history]
    assert max_rel(g.history["poisson"], want) < tol
    assert rel_l2(g.object - obj, o.obj - obj) < tol * 10
    assert max_rel((g.features["position"] - d0).ravel(), (o.offsets - d0).ravel()) < tol * 100


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_full_object_llk_matches_oracle(dtype):
    d = small_scan(n=64, nprobe=2, npos=30, seed=16)
    d0 = np.random.default_rng(17).uniform(-0.5, 0.5, (30, 2))
    o = oracle(d)
    o.set_sampling("full_object")
    o.set_position(d0, refine=False)
    want = o.llk()
    g = gpu(d, dtype)
    g.set_sampling("full_object")
    g.set_features(position=d0, hold=("position",))
    got = g.llk()
    for k in ("poisson", "gaussian", "euclidian"):
        assert abs(got[k] - getattr(want, k)) <= TOL[dtype] * abs(getattr(want, k)), k
    # Zero offsets under FullObject read the windows at their corners: the LLK without
    # offsets, up to the transforms' rounding.
    g.set_features(position=np.zeros((30, 2)), hold=("position",))
    a = g.llk()["poisson"]
    g.set_features(position=None)
    b = g.llk()["poisson"]
    assert abs(a - b) <= TOL[dtype] * abs(b)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_full_object_graph_replays_bitwise_like_eager(dtype):
    params, _ = RPIE_POS["epochs"]
    rows = np.array(table("rpie_full_object_epochs_offsets_f64.txt"))
    out = []
    for graph in (True, False):
        g, _, _ = scenario_engine(dtype, graph=graph)
        g.set_features(position=rows[:, 1:3])
        g.rpie(2, **params)
        out.append((g.object, g.probe, g.features["position"], g.history["poisson"]))
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_full_object_refusals():
    d = small_scan(n=32, nprobe=1, npos=10, seed=18)
    g = gpu(d, "c64")
    assert g.sampling == "patch"   # far field: Sampling::auto takes PATCH
    g.set_sampling("full_object")
    g.set_features(position=np.zeros((10, 2)))
    with pytest.raises(ValueError):
        g.set_sampling("other")
    g.set_features(position=None)
    g.ml(1)   # without offsets the sampling reads nothing


def lsqml_oracle(addr, optics, iobs, nb_obs, obj, probe, d0, nepoch, batch, seed):
    run = Lsqml(batch, POISSON, seed, offsets=d0, refine=True, sampling="full_object")
    llks = []
    for _ in range(nepoch):
        obj, probe, _ = run.epoch(addr, iobs, optics, obj, probe)
        llks.append(slice_llk(optics, addr.views(), iobs, addr.nmode, nb_obs, obj, probe,
                              run).poisson)
    return obj, probe, run.offsets, llks


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("case", [0, 2])
def test_lsqml_full_object_near_field_matches_oracle(dtype, case):
    # LSQML under FULL_OBJECT on the near-field LSQML fixture state (the oracle's step is
    # pinned by `lsqml_near_pos_full_*`): four epochs of batches of 8, offsets refined.
    addr, _, probe, cases = near_state()
    start_ = np.array(table("lsqml_near_pos_start_f64.txt"))
    obj = (start_[:, 2] + 1j * start_[:, 3]).reshape(1, 40, 40)
    d0 = offsets_of("lsqml_near_pos_full_offsets_f64.txt", case)
    optics, iobs, nb_obs, _ = cases[case]
    o_obj, o_probe, o_d, o_llk = lsqml_oracle(addr, optics, iobs, nb_obs, obj, probe, d0, 4, 8,
                                              7)
    g = near_gpu(case, dtype, obj)
    g.set_sampling("full_object")
    g.set_features(position=d0)
    g.lsqml(4, 1, 8, "poisson", 7)
    tol = PTOL[dtype]
    print("lsqml full object near", case, dtype, max_rel(g.history["poisson"], o_llk),
          rel_l2(g.object, o_obj), max_rel((g.features["position"] - d0).ravel(),
                                           (o_d - d0).ravel()))
    assert max_rel(g.history["poisson"], o_llk) < tol
    assert rel_l2(g.object, o_obj) < tol * 10
    assert rel_l2(g.probe, o_probe) < tol * 10
    assert max_rel((g.features["position"] - d0).ravel(), (o_d - d0).ravel()) < tol * 100


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("cost", ["poisson", "amplitude"])
def test_lsqml_full_object_far_field_matches_oracle(dtype, cost):
    d = small_scan(n=32, nprobe=2, npos=20, seed=19)
    o = oracle(d)
    d0 = np.random.default_rng(20).uniform(-0.6, 0.6, (20, 2))
    run = Lsqml(7, cost, 5, offsets=d0, refine=True, sampling="full_object")
    obj, probe, llks = o.obj, o.probe, []
    for _ in range(3):
        obj, probe, _ = run.epoch(o.addr, o.iobs, o.optics, obj, probe)
        llks.append(slice_llk(o.optics, o.addr.views(), o.iobs, o.addr.nmode, o.nb_obs, obj,
                              probe, run).poisson)
    g = gpu(d, dtype)
    g.set_sampling("full_object")
    g.set_features(position=d0)
    g.lsqml(3, 1, 7, cost, 5)
    tol = PTOL[dtype]
    assert max_rel(g.history["poisson"], llks) < tol
    assert rel_l2(g.object, obj) < tol * 10
    assert rel_l2(g.probe, probe) < tol * 10
    assert max_rel((g.features["position"] - d0).ravel(), (run.offsets - d0).ravel()) < tol * 100



@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_full_object_intensities_match_oracle(dtype):
    # The model's detector intensities at sampled windows (pty-chi's y_pred).
    d = small_scan(n=32, nprobe=2, npos=20, seed=21)
    d0 = np.random.default_rng(22).uniform(-0.6, 0.6, (20, 2))
    o = oracle(d)
    o.set_sampling("full_object")
    o.set_position(d0, refine=False)
    want = to_natural(_mode_sum(o.optics.exit_forward(o.calc_psi()), o.addr.nmode))
    g = gpu(d, dtype)
    g.set_sampling("full_object")
    g.set_features(position=d0, hold=("position",))
    err = float(np.abs(g.intensities() - want).max() / np.abs(want).max())
    assert err <= {"c64": 2e-5, "c128": 1e-12}[dtype], err


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("nprobe,refine,up", [(1, True, True), (3, True, False), (2, False, True)])
def test_ml_full_object_matches_oracle(dtype, nprobe, refine, up):
    # PyNX ML under FULL_OBJECT (the oracle's: the reads' gradient placed by the sampling's
    # adjoint, γ through the direction's sampled windows, the position sums through the
    # windows' derivatives): far field, masked pixels, offsets within 0.45 px, three cycles.
    d = small_scan(n=32, nprobe=nprobe, npos=25, seed=19)
    d0 = np.random.default_rng(20).uniform(-0.45, 0.45, (25, 2))
    d0 -= d0.mean(axis=0)
    o = oracle(d)
    o.set_sampling("full_object")
    o.set_position(d0, refine=refine)
    o.ml(3, update_object=True, update_probe=up, calc_llk=1)
    g = gpu(d, dtype)
    g.set_sampling("full_object")
    g.set_features(position=d0, hold=() if refine else ("position",))
    g.ml(3, update_object=True, update_probe=up, calc_llk=1)
    tol = PTOL[dtype]
    want = [r[2].poisson for r in o.history]
    print("ml full object", dtype, nprobe, refine, up, max_rel(g.history["poisson"], want),
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
    ("ap", 1, True, dict(probe_inertia=0.01)), ("ap", 2, False, {}),
    ("dm", 2, True, {}), ("dm", 1, False, dict(loop_obj_probe=2)),
    ("raar", 3, True, {}), ("raar", 1, True, dict(loop_obj_probe=2))])
def test_projections_full_object_match_oracle(dtype, name, nprobe, refine, kw):
    # AP, DM and RAAR under FULL_OBJECT (the oracle's: the probe's conjugate-gradient step
    # through the sampled windows, the object's preconditioned CG solve with the exact diagonal
    # and held pixels, the position sums through the windows' derivatives): far field, masked
    # pixels, zero-mean offsets within 0.45 px, refined or held, three cycles with LLK records.
    d = small_scan(n=32, nprobe=nprobe, npos=25, seed=33)
    d0 = np.random.default_rng(34).uniform(-0.45, 0.45, (25, 2))
    d0 -= d0.mean(axis=0)
    o = oracle(d)
    o.set_sampling("full_object")
    o.set_position(d0, refine=refine)
    params = dict(calc_llk=1, **kw)
    getattr(o, name)(3, **params)
    g = gpu(d, dtype)
    g.set_sampling("full_object")
    g.set_features(position=d0, hold=() if refine else ("position",))
    getattr(g, name)(3, **params)
    tol = PTOL[dtype]
    want = [r[2].poisson for r in o.history]
    print(name, "full object", nprobe, refine, kw, dtype, max_rel(g.history["poisson"], want),
          rel_l2(g.object, o.obj), rel_l2(g.probe, o.probe))
    assert max_rel(g.history["poisson"], want) < tol
    assert rel_l2(g.object, o.obj) < tol * 10
    assert rel_l2(g.probe, o.probe) < tol * 10
    got = g.features["position"]
    if refine:
        assert max_rel((got - d0).ravel(), (o.offsets - d0).ravel()) < tol * 100
    else:
        assert np.array_equal(got, d0)
