"""BH (pty-chi's bilinear-Hessian method on the amplitude cost) on the GPU through the C ABI
(ptyrs_engine_bh) against the reference's fixtures and the FP64 oracle, whose BH the fixtures
pin (tests/test_oracle_bh.py): the seeded far-field scenario with one batch of all 40 frames
(conjugate gradient from the second epoch) and with batches of 16 (bh_{full,mini}_*: every
step's β, α, ⟨g, D⟩ and H(D, D), the object and probe changes), the near-field cases (angular
spectrum, Fresnel forward and back, an evanescent band) against the oracle, LLK records, graph
replays against eager launches, refined PATCH position offsets against the reference's
bh_pos_{full,mini}_* and under FullObject sampling against pty-chi's bh_full_object_*, held
ones against the oracle, and the refusals. complex128 to FP64
rounding (1e-9), complex64 to 1e-3 of the steps (1e-2 of the changes)."""

import numpy as np
import pytest

import ptyrs_amd as pa
from conftest import max_rel, table
from oracle.bh import Bh
from test_gpu_rpie import rel_l2
from test_gpu_scenario import engine

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-9, "c64": 1e-3}


def check_steps(got, rows, tol, lead=0):
    """`got` (bh_steps dicts) against fixture rows `step epoch count beta alpha top bottom
    frames..` (after `lead` leading columns)."""
    assert len(got) == len(rows)
    for s, r in zip(got, rows):
        r = r[lead:]
        assert s["frames"] == [int(f) for f in r[7:]]
        if r[3] == 0.0:
            assert s["beta"] == 0.0
        else:
            assert abs(s["beta"] - r[3]) < tol * 10 * abs(r[3]), (s["beta"], r[3])
        for name, want in (("alpha", r[4]), ("top", r[5]), ("bottom", r[6])):
            assert abs(s[name] - want) < tol * abs(want), (name, s[name], want)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("batch_size,name", [(40, "bh_full"), (16, "bh_mini")])
def test_far_field_matches_fixtures(dtype, batch_size, name):
    # ptyrs/tests/engine_bh.rs:164-243 (probe ρ 0.01, seed 7, two epochs): each call is a
    # fresh BH, so epoch 0's steps come from a one-epoch call and epoch 1's from a two-epoch one.
    rows = table(f"{name}_steps_f64.txt")
    tol = TOL[dtype]
    g = engine(2, dtype)
    g.bh(1, 0, batch_size, 0.01, 7)
    check_steps(g.bh_steps(), [r for r in rows if r[1] == 0], tol)
    g = engine(2, dtype)
    obj0, probe0 = g.object, g.probe
    g.bh(2, 1, batch_size, 0.01, 7)
    check_steps(g.bh_steps(), [r for r in rows if r[1] == 1], tol)
    assert list(g.history["algorithm"]) == ["bh", "bh"]
    assert list(g.history["cycle"]) == [0, 1]
    obj_rows = np.array(table(f"{name}_obj_f64.txt"))
    probe_rows = np.array(table(f"{name}_probe_f64.txt"))
    want_o = (obj_rows[:, 2] + obj_rows[:, 4]) + 1j * (obj_rows[:, 3] + obj_rows[:, 5])
    do = (g.object - obj0)[0][obj_rows[:, 0].astype(int), obj_rows[:, 1].astype(int)]
    idx = probe_rows[:, :3].astype(int)
    want_p = (probe_rows[:, 3] + probe_rows[:, 5]) + 1j * (probe_rows[:, 4] + probe_rows[:, 6])
    dp = (g.probe - probe0)[idx[:, 0], idx[:, 1], idx[:, 2]]
    print(name, dtype, rel_l2(do, want_o), rel_l2(dp, want_p))
    assert rel_l2(do, want_o) < tol * 10
    assert rel_l2(dp, want_p) < tol * 10


def near(case, dtype, graph=True):
    from test_gpu_lsqml import _near_state
    corners, obj, probe, counts = _near_state()
    c = table("lsqml_near_cases_f64.txt")[case]
    vals = np.array([r[2:] for r in counts if r[0] == case]).reshape(len(corners), 24, 24)
    nf = len(corners)
    scan = pa.Scan(vals.astype(np.float32), np.zeros(nf), np.zeros(nf), c[1], 1.0, 75e-6,
                   corners=corners[:, 1:3].astype(np.int64))
    w = pa.Wavefield([24, 24], c[1], [c[2], c[2]], c[3])
    return pa.GpuPtycho(scan, obj, probe, optics=pa.Optics.free_space(w), dtype=dtype,
                        graph=graph)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("case", [0, 1, 2, 3])
@pytest.mark.parametrize("method,batch_size", [(0, 16), (1, 6)])
def test_near_field_matches_fixtures(dtype, case, method, batch_size):
    # pty-chi's planar near field (ptyrs/tests/engine_bh_near.rs:103-176): probe ρ 1, two
    # epochs, the last epoch's steps and the changes over the run.
    g = near(case, dtype)
    obj0, probe0 = g.object, g.probe
    g.bh(2, 0, batch_size, 1.0, 7)
    tol = TOL[dtype]
    rows = [r for r in table("bh_near_steps_f64.txt")
            if r[0] == case and r[1] == method and r[3] == 1]
    check_steps(g.bh_steps(), rows, tol, lead=2)
    for now, start, name, lead in ((g.object, obj0, "bh_near_obj_f64.txt", 4),
                                   (g.probe, probe0, "bh_near_probe_f64.txt", 5)):
        r = np.array([t for t in table(name) if t[0] == case and t[1] == method])
        want = r[:, lead] + 1j * r[:, lead + 1]
        print(case, method, dtype, name, rel_l2((now - start).ravel(), want))
        assert rel_l2((now - start).ravel(), want) < tol * 10


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_more_epochs_match_oracle(dtype):
    # Four conjugate-gradient epochs and three of batches of 7 (a ragged last batch of 5) on the
    # seeded scenario with one probe mode, against the oracle, with the LLK after each epoch.
    from oracle import scenario as sc
    from oracle.ptycho import llk, obj_probe_to_psi
    for batch in (40, 7):
        p, _ = sc.parity_start(nprobe=1)
        bh = Bh(batch, probe_rho=0.02, seed=3)
        obj, probe, want = p.obj, p.probe, []
        nepoch = 4 if batch == 40 else 3
        for _ in range(nepoch):
            obj, probe, steps = bh.epoch(p.addr, p.iobs, p.optics, obj, probe)
            psi = p.optics.exit_forward(obj_probe_to_psi(None, p.views_all, obj, probe, p.quad))
            want.append(llk(p.iobs, p.addr.nmode, psi).normalised(p.nb_obs).poisson)
        g = engine(1, dtype)
        g.bh(nepoch, 1, batch, 0.02, 3)
        tol = TOL[dtype]
        assert max_rel(g.history["poisson"], want) < tol
        got = g.bh_steps()
        assert [s["frames"] for s in got] == [list(s[0]) for s in steps]
        assert max_rel([s["alpha"] for s in got], [s[2] for s in steps]) < tol
        assert rel_l2(g.object, obj) < tol * 10
        assert rel_l2(g.probe, probe) < tol * 10


@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_graph_replays_match_eager(dtype):
    out = []
    for graph in (True, False):
        g = near(0, dtype, graph=graph)
        g.bh(3, 1, 16, 1.0, 7)
        g.bh(2, 0, 6, 1.0, 9)
        g.bh(2, 1, 16, 1.0, 7)
        out.append((g.object, g.probe, g.history["poisson"], g.bh_steps()))
    a, b = out
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])
    assert a[3] == b[3]


def test_refusals():
    g = engine(2, "c64")
    with pytest.raises(ValueError, match="batch_size"):
        g.bh(1, 0, 0, 0.01, 7)
    with pytest.raises(ValueError):
        g.bh(1, 0, 16, float("nan"), 7)
    assert g.cycle == 0 and g.bh_steps() == []
    g.bh(0, 0, 16, 0.01, 7)
    assert g.bh_steps() == []


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("batch_size,name,sampling", [
    (40, "bh_pos_full", "patch"), (16, "bh_pos_mini", "patch"),
    (40, "bh_full_object_full", "full_object"), (16, "bh_full_object_mini", "full_object")])
def test_offsets_match_fixtures(dtype, batch_size, name, sampling):
    # Refined offsets under PATCH (the reference's model) and FullObject sampling (pty-chi's
    # FULL_OBJECT BH) (ptyrs/tests/engine_bh_pos.rs:61-184): probe ρ 0.01, position ρ 2, seed 7,
    # two epochs from the fixture's offsets: each epoch's steps, the object, probe and offset
    # changes over the run.
    rows = np.array(table(f"{name}_offsets_f64.txt"))
    d0 = rows[:, 1:3].copy()
    steps = table(f"{name}_steps_f64.txt")
    tol = TOL[dtype]
    g = engine(2, dtype)
    g.set_features(position=d0)
    g.set_sampling(sampling)
    g.bh(1, 0, batch_size, 0.01, 7, position_rho=2.0)
    check_steps(g.bh_steps(), [r for r in steps if r[1] == 0], tol)
    g = engine(2, dtype)
    g.set_features(position=d0)
    g.set_sampling(sampling)
    obj0, probe0 = g.object, g.probe
    g.bh(2, 1, batch_size, 0.01, 7, position_rho=2.0)
    check_steps(g.bh_steps(), [r for r in steps if r[1] == 1], tol)
    for now, start, fname, lead in ((g.object, obj0, f"{name}_obj_f64.txt", 2),
                                    (g.probe, probe0, f"{name}_probe_f64.txt", 3)):
        r = np.array(table(fname))
        want = r[:, lead] + 1j * r[:, lead + 1]
        print(name, dtype, fname, rel_l2((now - start).ravel(), want))
        assert rel_l2((now - start).ravel(), want) < tol * 10
    got = g.features["position"] - d0
    want = rows[:, 3:5] - d0
    print(name, dtype, "offsets", np.abs(got - want).max() / np.abs(want).max())
    assert np.abs(got - want).max() < tol * 10 * np.abs(want).max()


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("batch_size", [40, 7])
def test_held_offsets_match_oracle(dtype, batch_size):
    # Held PATCH offsets (the moved probe copies alone): three epochs against the oracle, one
    # probe mode, offsets within 0.5 px; the offsets stay.
    from oracle import scenario as sc
    p, _ = sc.parity_start(nprobe=1)
    d0 = np.random.default_rng(41).uniform(-0.5, 0.5, (p.addr.nframe, 2))
    bh = Bh(batch_size, probe_rho=0.02, seed=5, offsets=d0, refine=False)
    obj, probe = p.obj, p.probe
    for _ in range(3):
        obj, probe, steps = bh.epoch(p.addr, p.iobs, p.optics, obj, probe)
    g = engine(1, dtype)
    g.set_features(position=d0, hold=("position",))
    g.bh(3, 0, batch_size, 0.02, 5)
    tol = TOL[dtype]
    got = g.bh_steps()
    assert [s["frames"] for s in got] == [list(s[0]) for s in steps]
    assert max_rel([s["alpha"] for s in got], [s[2] for s in steps]) < tol
    assert max_rel([s["bottom"] for s in got], [s[4] for s in steps]) < tol
    assert rel_l2(g.object, obj) < tol * 10
    assert rel_l2(g.probe, probe) < tol * 10
    assert np.array_equal(g.features["position"], d0)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("case", [0, 1, 2])
@pytest.mark.parametrize("method,batch_size", [(0, 16), (1, 6)])
def test_multislice_matches_fixtures(dtype, case, method, batch_size):
    # Three slices with unequal gaps under the angular-spectrum, Fresnel and far-field exits
    # (ptyrs/tests/engine_bh_multislice.rs:114-185): probe ρ 1, three epochs: the last epoch's
    # steps, every slice's and the probe's change over the run.
    from test_gpu_multislice import start as ms_start
    g, objects, probe0 = ms_start(case, dtype)
    g.bh(3, 1, batch_size, 1.0, 7)
    tol = TOL[dtype]
    rows = [r for r in table("bh_multislice_steps_f64.txt")
            if r[0] == case and r[1] == method and r[3] == 2]
    got = g.bh_steps()
    assert len(got) == len(rows)
    for s, r in zip(got, rows):
        assert s["frames"] == [int(f) for f in r[10:]]
        if r[5] != 0.0:
            assert abs(s["beta"] - r[5]) < tol * 10 * abs(r[5]), (s["beta"], r[5])
        for name, want in (("alpha", r[6]), ("top", r[7]), ("bottom", r[8])):
            assert abs(s[name] - want) < tol * abs(want), (name, s[name], want)
    for now, start, name in ((g.objects, objects, "bh_multislice_obj_f64.txt"),
                             (g.probe, probe0, "bh_multislice_probe_f64.txt")):
        r = np.array([t for t in table(name) if t[0] == case and t[1] == method])
        want = r[:, 5] + 1j * r[:, 6]
        print(case, method, dtype, name, rel_l2((now - start).ravel(), want))
        assert rel_l2((now - start).ravel(), want) < tol * 10
    assert list(g.history["algorithm"]) == ["bh"] * 3


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("case,batch_size", [(0, 16), (1, 6), (2, 40)])
def test_multislice_held_full_object_offsets_match_oracle(dtype, case, batch_size):
    # Three slices read through their sampled windows at held FullObject offsets (the oracle's
    # BhSlices with a sampler: the slices' gradients placed by the sampling's adjoint, the
    # directions' slices read through the same windows), all three exits, steepest descent and
    # (one batch of every frame) conjugate gradient, three epochs.
    from test_gpu_multislice import start as ms_start
    from test_oracle_multislice import state as ms_state
    addr, objects, probe0, cases = ms_state()
    optics, iobs, nb_obs = cases[case][0], cases[case][1], cases[case][2]
    d0 = np.random.default_rng(36).uniform(-0.45, 0.45, (addr.nframe, 2))
    bh = Bh(batch_size, probe_rho=1.0, seed=7, offsets=d0, refine=False,
            sampling="full_object")
    obj, probe, got = np.stack(objects), probe0, []
    for _ in range(3):
        obj, probe, s = bh.epoch(addr, iobs, optics, obj, probe)
        got += s
    g, _, _ = ms_start(case, dtype)
    g.set_sampling("full_object")
    g.set_features(position=d0, hold=("position",))
    g.bh(3, 1, batch_size, 1.0, 7)
    tol = TOL[dtype]
    want = got[-len(g.bh_steps()):]
    for s, (frames, beta, alpha, top, bottom) in zip(g.bh_steps(), want):
        assert s["frames"] == list(frames)
        for name, w in (("alpha", alpha), ("top", top), ("bottom", bottom)):
            assert abs(s[name] - w) < tol * 10 * abs(w), (name, s[name], w)
    print("bh multislice full object", case, batch_size, dtype,
          rel_l2(g.objects - np.stack(objects), obj - np.stack(objects)),
          rel_l2(g.probe - probe0, probe - probe0))
    assert rel_l2(g.objects - np.stack(objects), obj - np.stack(objects)) < tol * 10
    assert rel_l2(g.probe - probe0, probe - probe0) < tol * 10
    assert np.array_equal(g.features["position"], d0)
    # PATCH offsets through several slices are refused
    g.set_sampling("patch")
    g.set_features(position=d0, hold=("position",))
    with pytest.raises(ValueError, match="several object slices"):
        g.bh(1, 0, batch_size, 1.0, 7)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("case,batch_size", [(0, 16), (1, 6), (2, 40)])
def test_multislice_refined_full_object_offsets_match_oracle(dtype, case, batch_size):
    # Three slices at refined FullObject offsets (the oracle's BhSlices with refine: the
    # offsets' gradient over every slice's reads, the reads' offset terms slice by slice, the
    # step 0.3·α·D_d), all three exits, steepest descent and conjugate gradient, three epochs.
    from test_gpu_multislice import start as ms_start
    from test_oracle_multislice import state as ms_state
    addr, objects, probe0, cases = ms_state()
    optics, iobs = cases[case][0], cases[case][1]
    d0 = np.random.default_rng(37).uniform(-0.45, 0.45, (addr.nframe, 2))
    bh = Bh(batch_size, probe_rho=1.0, seed=7, offsets=d0, refine=True, sampling="full_object")
    obj, probe, got = np.stack(objects), probe0, []
    for _ in range(3):
        obj, probe, s = bh.epoch(addr, iobs, optics, obj, probe)
        got += s
    g, _, _ = ms_start(case, dtype)
    g.set_sampling("full_object")
    g.set_features(position=d0)
    g.bh(3, 1, batch_size, 1.0, 7)
    tol = TOL[dtype]
    want = got[-len(g.bh_steps()):]
    for s, (frames, beta, alpha, top, bottom) in zip(g.bh_steps(), want):
        assert s["frames"] == list(frames)
        for name, w in (("alpha", alpha), ("top", top), ("bottom", bottom)):
            assert abs(s[name] - w) < tol * 10 * abs(w), (name, s[name], w)
    dd = g.features["position"] - d0
    print("bh multislice refined full object", case, batch_size, dtype,
          rel_l2(g.objects - np.stack(objects), obj - np.stack(objects)),
          rel_l2(g.probe - probe0, probe - probe0), max_rel(dd.ravel(), (bh.offsets - d0).ravel()))
    assert rel_l2(g.objects - np.stack(objects), obj - np.stack(objects)) < tol * 10
    assert rel_l2(g.probe - probe0, probe - probe0) < tol * 10
    assert max_rel(dd.ravel(), (bh.offsets - d0).ravel()) < tol * 100
