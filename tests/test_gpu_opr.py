"""LSQML with orthogonal probe relaxation on the GPU (ptyrs_engine_lsqml_opr) against the FP64
oracle, whose OPR step the reference's fixtures pin (tests/test_oracle_opr.py): the reference's
seeded far-field scenario with its fixtures' eigenmodes and weights (one eigenmode; three; three
with the main mode's weights), both costs, weights and intensity refined or held, two epochs:
the LLK records, object, probe, eigenmodes, weights and main-mode weights. Also three slices
(the angular-spectrum multislice fixture state), refined PATCH offsets (the reference's
lsqml_pos_opr state) and refined offsets under FullObject sampling. complex128 to FP64 rounding (1e-9), complex64
to 1e-3 of the changes."""

import numpy as np
import pytest

import ptyrs_amd as pa
from conftest import max_rel, table
from oracle import scenario as sc
from oracle.lsqml import AMPLITUDE, POISSON, Lsqml, slice_llk
from oracle.opr import ProbeVariation
from test_gpu_rpie import rel_l2
from test_gpu_scenario import start
from test_oracle_opr import Ref

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-9, "c64": 1e-3}


def oracle_run(var, cost, nepoch, batch, seed):
    from test_oracle_opr import state
    p = state()
    run = Lsqml(batch, cost, seed)
    obj, probe, llks = p.obj, p.probe, []
    for _ in range(nepoch):
        obj, probe, _ = run.epoch(p.addr, p.iobs, p.optics, obj, probe, var)
        llks.append(slice_llk(p.optics, p.views_all, p.iobs, p.addr.nmode, p.nb_obs, obj, probe,
                              None, var).poisson)
    return obj, probe, llks


def engine(dtype):
    counts, y, x, obj, probe, _ = start(2)
    return pa.GpuPtycho(pa.Scan(counts, y, x, *sc.GEOMETRY), obj, probe, dtype=dtype)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("name,cost,hold", [
    ("opr", POISSON, ()), ("lsqml_opr", AMPLITUDE, ()), ("lsqml_opr_int", POISSON, ()),
    ("lsqml_opr_int", AMPLITUDE, ("weights",)), ("lsqml_opr_int", POISSON, ("intensity",))])
def test_opr_epochs_match_oracle(dtype, name, cost, hold):
    r = Ref(name)
    var = r.variation(refine="weights" not in hold, refine_c="intensity" not in hold)
    start_var = var.copy()
    o_obj, o_probe, o_llk = oracle_run(var, cost, 2, 12, 5)
    g = engine(dtype)
    obj0, probe0 = g.object, g.probe
    out = g.lsqml(2, 1, 12, cost, 5,
                  opr=(start_var.eigenmodes, start_var.weights.reshape(-1, r.k),
                       start_var.intensity), hold=hold)
    eig, w, inten = out
    tol = TOL[dtype]
    print("opr", name, cost, hold, dtype, max_rel(g.history["poisson"], o_llk),
          rel_l2(g.object - obj0, o_obj - obj0), rel_l2(eig, var.eigenmodes))
    assert max_rel(g.history["poisson"], o_llk) < tol
    assert rel_l2(g.object - obj0, o_obj - obj0) < tol * 10
    assert rel_l2(g.probe - probe0, o_probe - probe0) < tol * 10
    assert rel_l2(eig, var.eigenmodes) < tol * 10
    w0 = start_var.weights.reshape(-1, r.k)
    if "weights" in hold:
        assert np.array_equal(w, w0)
    else:
        assert max_rel((w - w0).ravel(), (var.weights.reshape(-1, r.k) - w0).ravel()) < tol * 10
    if start_var.intensity is None:
        assert inten is None
    elif "intensity" in hold:
        assert np.array_equal(inten, start_var.intensity)
    else:
        assert max_rel(inten - start_var.intensity, var.intensity - start_var.intensity) < tol * 10


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_opr_multislice_matches_oracle(dtype):
    # Three slices (the angular-spectrum multislice fixture state), two eigenmodes.
    from test_gpu_multislice import start as ms_start
    from test_oracle_multislice import state as multislice_state
    addr, objects, probe, cases = multislice_state()
    optics, iobs, nb_obs, _ = cases[0]
    rng = np.random.default_rng(30)
    n = probe.shape[-1]
    eig = 0.05 * (rng.normal(size=(2, n, n)) + 1j * rng.normal(size=(2, n, n)))
    w0 = rng.uniform(-0.5, 0.5, (addr.nframe, 2))
    var = ProbeVariation(eig, w0.ravel())
    run = Lsqml(8, POISSON, 7)
    obj, pr, llks = list(objects), probe, []
    for _ in range(2):
        obj, pr, _ = run.epoch(addr, iobs, optics, obj, pr, var)
        llks.append(slice_llk(optics, addr.views(), iobs, addr.nmode, nb_obs, obj, pr, None,
                              var).poisson)
    g, _, _ = ms_start(0, dtype)
    g.set_object(np.stack(objects))
    geig, gw, _ = g.lsqml(2, 1, 8, "poisson", 7, opr=(eig, w0, None))
    tol = TOL[dtype]
    assert max_rel(g.history["poisson"], llks) < tol
    assert rel_l2(g.objects, np.stack(obj)) < tol * 10
    assert rel_l2(geig, var.eigenmodes) < tol * 10
    assert max_rel((gw - w0).ravel(), (var.weights.reshape(-1, 2) - w0).ravel()) < tol * 10


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_opr_multislice_full_object_matches_oracle(dtype):
    # Three slices under FULL_OBJECT with two eigenmodes and refined offsets (the reference runs
    # this combination): each frame's own probe planes, unmoved, through every slice's sampled
    # windows. The reference's lsqml_multislice_pos_full state and offsets: its slices carry
    # structure, so the position steps are well conditioned (on the featureless slices of the
    # plain multislice state a 1e-15 change of the start moves the refined offsets by 1e-7).
    from test_gpu_multislice import start as ms_start
    from test_gpu_positions import offsets_of
    from test_oracle_multislice import state as multislice_state
    addr, _, probe, cases = multislice_state()
    start = np.array(table("lsqml_multislice_pos_start_f64.txt"))
    objects = [(start[start[:, 0] == s][:, 3] + 1j * start[start[:, 0] == s][:, 4])
               .reshape(1, 40, 40) for s in range(3)]
    d0 = offsets_of("lsqml_multislice_pos_full_offsets_f64.txt", 0)
    optics, iobs, nb_obs, _ = cases[0]
    rng = np.random.default_rng(31)
    n = probe.shape[-1]
    eig = 0.05 * (rng.normal(size=(2, n, n)) + 1j * rng.normal(size=(2, n, n)))
    w0 = rng.uniform(-0.5, 0.5, (addr.nframe, 2))
    var = ProbeVariation(eig, w0.ravel())
    run = Lsqml(8, POISSON, 7, offsets=d0, refine=True, sampling="full_object")
    obj, pr, llks = list(objects), probe, []
    for _ in range(2):
        obj, pr, _ = run.epoch(addr, iobs, optics, obj, pr, var)
        llks.append(slice_llk(optics, addr.views(), iobs, addr.nmode, nb_obs, obj, pr, run,
                              var).poisson)
    g, _, _ = ms_start(0, dtype)
    g.set_object(np.stack(objects))
    g.set_sampling("full_object")
    g.set_features(position=d0)
    geig, gw, _ = g.lsqml(2, 1, 8, "poisson", 7, opr=(eig, w0, None))
    tol = TOL[dtype]
    print("opr multislice full object", dtype, max_rel(g.history["poisson"], llks),
          rel_l2(g.objects, np.stack(obj)), rel_l2(geig, var.eigenmodes))
    assert max_rel(g.history["poisson"], llks) < tol
    assert rel_l2(g.objects, np.stack(obj)) < tol * 10
    assert rel_l2(geig, var.eigenmodes) < tol * 10
    assert max_rel((gw - w0).ravel(), (var.weights.reshape(-1, 2) - w0).ravel()) < tol * 10
    got = g.features["position"]
    assert max_rel((got - d0).ravel(), (run.offsets - d0).ravel()) < tol * 100


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("cost", [AMPLITUDE, POISSON])
def test_opr_with_positions_matches_oracle(dtype, cost):
    # OPR with refined PATCH offsets: the reference's lsqml_pos_opr state (three eigenmodes,
    # its weights and offsets), two epochs of batches of 10.
    from test_oracle_opr import state
    p = state()
    eig = np.array(table("lsqml_pos_opr_eigen_f64.txt"))
    wt = np.array(table("lsqml_pos_opr_weights_f64.txt"))
    offs = np.array(table("lsqml_pos_opr_offsets_f64.txt"))[:, 1:3]
    u0 = (eig[:, 3] + 1j * eig[:, 4]).reshape(3, 32, 32)
    w0 = wt[:, 1:4].copy()
    var = ProbeVariation(u0, w0.ravel())
    run = Lsqml(10, cost, 3, offsets=offs, refine=True, sampling="patch")
    obj, probe, llks = p.obj, p.probe, []
    for _ in range(2):
        obj, probe, _ = run.epoch(p.addr, p.iobs, p.optics, obj, probe, var)
        llks.append(slice_llk(p.optics, p.views_all, p.iobs, p.addr.nmode, p.nb_obs, obj, probe,
                              run, var).poisson)
    g = engine(dtype)
    obj0 = g.object
    g.set_features(position=offs)
    geig, gw, _ = g.lsqml(2, 1, 10, cost, 3, opr=(u0, w0, None))
    tol = TOL[dtype]
    print("opr pos", dtype, cost, max_rel(g.history["poisson"], llks), rel_l2(geig, var.eigenmodes))
    assert max_rel(g.history["poisson"], llks) < tol
    assert rel_l2(g.object - obj0, obj - obj0) < tol * 10
    assert rel_l2(geig, var.eigenmodes) < tol * 10
    assert max_rel((gw - w0).ravel(), (var.weights.reshape(-1, 3) - w0).ravel()) < tol * 10
    got = g.features["position"]
    assert max_rel((got - offs).ravel(), (run.offsets - offs).ravel()) < tol * 100


def test_opr_refusals():
    g = engine("c64")
    k = 2
    eig = np.zeros((k, 32, 32), dtype=np.complex128)
    w = np.zeros((40, k))
    with pytest.raises(ValueError):
        g.lsqml(1, 1, 8, "poisson", 1, opr=(eig, w[:10], None))
    with pytest.raises(ValueError):
        g.lsqml(1, 1, 8, "poisson", 1, opr=(eig[:, :16], w, None))
    with pytest.raises(ValueError):
        g.lsqml(1, 1, 8, "poisson", 1, opr=(eig, w, None), hold=("intensity",))
    with pytest.raises(TypeError):
        g.lsqml(1, 1, 8, "poisson", 1, variation=None)
    assert g.cycle == 0


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("name,cost,hold", [
    ("lsqml_opr", POISSON, ()), ("lsqml_opr_int", AMPLITUDE, ()),
    ("lsqml_opr_int", POISSON, ("weights",))])
def test_opr_full_object_matches_oracle(dtype, name, cost, hold):
    # OPR under FULL_OBJECT with refined offsets (the reference runs LSQML's OPR under both
    # samplings): each frame's own probe planes, unmoved, against its sampled window; the
    # frames' directions and window powers from the windows. The seeded scenario with the
    # reference's eigenmodes and weights, zero-mean offsets within 0.4 px, two epochs.
    from test_oracle_opr import state
    p = state()
    r = Ref(name)
    var = r.variation(refine="weights" not in hold, refine_c=True)
    start_var = var.copy()
    d0 = np.random.default_rng(35).uniform(-0.4, 0.4, (len(p.iobs), 2))
    d0 -= d0.mean(axis=0)
    run = Lsqml(12, cost, 5, offsets=d0, refine=True, sampling="full_object")
    obj, probe, llks = p.obj, p.probe, []
    for _ in range(2):
        obj, probe, _ = run.epoch(p.addr, p.iobs, p.optics, obj, probe, var)
        llks.append(slice_llk(p.optics, p.views_all, p.iobs, p.addr.nmode, p.nb_obs, obj, probe,
                              run, var).poisson)
    g = engine(dtype)
    obj0, probe0 = g.object, g.probe
    g.set_sampling("full_object")
    g.set_features(position=d0)
    eig, w, inten = g.lsqml(2, 1, 12, cost, 5,
                            opr=(start_var.eigenmodes, start_var.weights.reshape(-1, r.k),
                                 start_var.intensity), hold=hold)
    tol = TOL[dtype]
    print("opr full object", name, cost, hold, dtype, max_rel(g.history["poisson"], llks),
          rel_l2(g.object - obj0, obj - obj0), rel_l2(eig, var.eigenmodes))
    assert max_rel(g.history["poisson"], llks) < tol
    assert rel_l2(g.object - obj0, obj - obj0) < tol * 10
    assert rel_l2(g.probe - probe0, probe - probe0) < tol * 10
    assert rel_l2(eig, var.eigenmodes) < tol * 10
    w0 = start_var.weights.reshape(-1, r.k)
    if "weights" in hold:
        assert np.array_equal(w, w0)
    else:
        assert max_rel((w - w0).ravel(), (var.weights.reshape(-1, r.k) - w0).ravel()) < tol * 10
    if start_var.intensity is not None:
        assert max_rel(inten - start_var.intensity, var.intensity - start_var.intensity) < tol * 10
    got = g.features["position"]
    assert max_rel((got - d0).ravel(), (run.offsets - d0).ravel()) < tol * 100
