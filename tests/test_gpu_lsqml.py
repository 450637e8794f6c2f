"""LSQML on the GPU through the C ABI against the FP64 oracle (same seeded inputs) and against
the reference's own near-field LSQML fixtures (ptyrs/tests/engine_lsqml_near.rs:455-513). c128
must match to FP64 rounding; c64 to the reference's GPU-vs-CPU bound (ptyrs/tests/
engine_gpu_lsqml.rs)."""

import numpy as np
import pytest

import ptyrs_amd as pa
from conftest import max_rel, table
from oracle.lsqml import Lsqml
from test_gpu_rpie import gpu, oracle, rel_l2, small_scan

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-9, "c64": 1e-4}


def run_oracle(o, nepoch, batch, cost, seed):
    run = Lsqml(batch, cost, seed)
    obj, probe = o.obj, o.probe
    llks, steps = [], None
    for _ in range(nepoch):
        obj, probe, steps = run.epoch(o.addr, o.iobs, o.optics, obj, probe)
        llks.append(run.llk(o, obj, probe).poisson)
    return obj, probe, llks, steps


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("cost", ["poisson", "amplitude"])
@pytest.mark.parametrize("batch", [7, 30])
def test_lsqml_far_field_matches_oracle(dtype, cost, batch):
    d = small_scan(n=64, nprobe=2, npos=30)
    g, o = gpu(d, dtype), oracle(d)
    obj, probe, llks, steps = run_oracle(o, 3, batch, cost, 5)
    g.lsqml(3, 1, batch, cost, 5)
    assert g.cycle == 3
    assert max_rel(g.history["poisson"], llks) < TOL[dtype]
    assert set(g.history["algorithm"]) == {"lsqml"}
    gs = g.lsqml_steps()
    assert [list(s["frames"]) for s in gs] == [s["frames"] for s in steps]
    for name in ("a_far", "a_probe"):
        got = np.concatenate([s[name] for s in gs])
        want = np.concatenate([s[name] for s in steps])
        assert max_rel(got, want) < TOL[dtype] * 100, name
    got = np.concatenate([s["a_obj"][0] for s in gs])
    assert max_rel(got, np.concatenate([s["a_obj"] for s in steps])) < TOL[dtype] * 100
    tol = TOL[dtype] * (10 if dtype == "c128" else 30)
    assert rel_l2(g.object - d["obj"], obj - d["obj"]) < tol
    assert rel_l2(g.probe - d["probe"], probe - d["probe"]) < tol


def test_lsqml_preconditioner_schedule_and_repeat_calls():
    """Eleven epochs cross the schedule (epochs 0–9 and every tenth); a second call starts a
    fresh LSQML from the state the first left."""
    d = small_scan(n=32, nprobe=1, npos=20)
    g, o = gpu(d, "c128"), oracle(d)
    obj, probe, llks, _ = run_oracle(o, 11, 6, "poisson", 2)
    g.lsqml(11, 1, 6, "poisson", 2)
    # FP64 rounding grows over eleven epochs; 1e-8 still pins every step.
    assert max_rel(g.history["poisson"], llks) < 1e-8
    o.obj, o.probe = obj, probe
    obj2, probe2, llks2, _ = run_oracle(o, 2, 6, "poisson", 2)
    g.lsqml(2, 1, 6, "poisson", 2)
    assert g.cycle == 13
    # A fresh LSQML's first large steps amplify the ~1e-8 state difference after eleven epochs;
    # a state that was not reset (a_far, preconditioner) would differ at the percent level.
    assert max_rel(g.history["poisson"][11:], llks2) < 1e-6
    assert rel_l2(g.object, obj2) < 1e-6


def _near_state():
    shape, obj_shape, nprobe = (24, 24), (40, 40), 2
    corners = np.array(table("lsqml_near_corners_f64.txt"))
    obj = np.array([r[2] + 1j * r[3] for r in table("lsqml_near_obj_f64.txt")]).reshape(1, *obj_shape)
    probe = np.array([r[3] + 1j * r[4] for r in table("lsqml_near_probe_f64.txt")])
    counts = table("lsqml_near_counts_f64.txt")
    return corners, obj, probe.reshape(nprobe, *shape), counts


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("case", [0, 1, 2, 3])
def test_lsqml_near_field_matches_reference_fixtures(dtype, case):
    """Poisson LSQML, batches of 8 (seed 7), under an angular-spectrum exit (case 0; case 2 with
    an evanescent band the transfer function zeroes) and a Fresnel one forward (case 1) and back
    (case 3): every step of the last epoch, the LLK after every epoch and the changes over the
    run."""
    corners, obj, probe, counts = _near_state()
    c = table("lsqml_near_cases_f64.txt")[case]
    vals = np.array([r[2:] for r in counts if r[0] == case]).reshape(len(corners), 24, 24)
    nf = len(corners)
    scan = pa.Scan(vals.astype(np.float32), np.zeros(nf), np.zeros(nf), c[1], 1.0, 75e-6,
                   corners=corners[:, 1:3].astype(np.int64))
    w = pa.Wavefield([24, 24], c[1], [c[2], c[2]], c[3])
    g = pa.GpuPtycho(scan, obj, probe, optics=pa.Optics.free_space(w), dtype=dtype)
    want = [r[2] for r in table("lsqml_near_llk_f64.txt") if r[0] == case]
    nepoch = len(want) - 1
    llk0 = g.llk()["poisson"]
    g.lsqml(nepoch, 1, 8, "poisson", 7)
    llks = [llk0] + list(g.history["poisson"])
    tol = {"c128": 1e-9, "c64": 1e-3}[dtype]
    assert max_rel(llks, want) < tol
    rows = [r for r in table("lsqml_near_epochs_steps_f64.txt") if r[0] == case and r[1] == nepoch - 1]
    steps = g.lsqml_steps()
    assert [int(f) for s in steps for f in s["frames"]] == [int(r[2]) for r in rows]
    for name, col in (("a_far", 3), ("a_probe", 5)):
        got = np.concatenate([s[name] for s in steps])
        assert max_rel(got, [r[col] for r in rows]) < tol * 10, name
    for got, name, lead in ((g.object - obj, "lsqml_near_epochs_obj_f64.txt", 3),
                            (g.probe - probe, "lsqml_near_epochs_probe_f64.txt", 4)):
        r = np.array([t for t in table(name) if t[0] == case])
        assert rel_l2(got.ravel(), r[:, lead] + 1j * r[:, lead + 1]) < tol * 10, name


def test_lsqml_rejects_bad_params():
    d = small_scan(n=32, nprobe=1, npos=6)
    g = gpu(d, "c64")
    with pytest.raises(ValueError):
        g.lsqml(1, 0, 0, "poisson", 1)
    with pytest.raises(ValueError):
        g.lsqml(1, 0, 4, "gaussian", 1)
    with pytest.raises(TypeError):
        g.lsqml(1, 0, 4, "poisson", 1, variation=None)
    assert g.lsqml(1, 0, 4, "poisson", 1, opr=None) is None   # no probe variation
