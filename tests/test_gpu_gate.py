"""The C ABI's contract on the GPU: u16 and float counts give the same bits, the abort gate
(ptyrs/src/engine/gpu/gate.rs:14-71) driven by the fault hook, and the GpuOptions that change
no result (llk_slots, stack_size, fused)."""

import numpy as np
import pytest

import ptyrs_amd as pa
from ptyrs_amd import base as pb
from ptyrs_amd import sim
from test_gpu_rpie import small_scan

pytestmark = pytest.mark.gpu

GEOM = (sim.WAVELENGTH, sim.DISTANCE, sim.DETECTOR_PIXEL)


def _u16_pair(d):
    """The same counts as uint16 (65535 masked) and float32 (-1 masked), all below 65535."""
    c = np.minimum(d["counts"], 65534)
    u = np.where(c < 0, 65535, c).astype(np.uint16)
    f = np.where(u == 65535, -1.0, u.astype(np.float64)).astype(np.float32)
    return u, f


@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_u16_counts_run_bit_identical_to_float_counts(dtype):
    # counts.rs:3-7: below MAX_COUNT the float path (negative = masked) is the u16 path
    # (65535 = masked) bit for bit, through rPIE, LSQML and AP and their LLK records.
    d = small_scan(n=64, nprobe=2, npos=30)
    u, f = _u16_pair(d)
    assert (u == 65535).any()
    out = []
    for counts in (u, f):
        g = pa.GpuPtycho(pa.Scan(counts, d["y"], d["x"], *GEOM), d["obj"], d["probe"],
                         dtype=dtype)
        g.rpie(2, batch_size=7, calc_llk=1)
        g.lsqml(2, 1, 10, "poisson", 3)
        g.ap(2, calc_llk=1)
        out.append((g.object, g.probe, g.history["poisson"], g.llk()["poisson"]))
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)


def _gate_engine():
    d = small_scan(n=32, nprobe=2, npos=12)
    return pa.GpuPtycho(pa.Scan(d["counts"], d["y"], d["x"], *GEOM), d["obj"], d["probe"],
                        dtype="c128", graph=False), d


def test_a_failure_before_any_launch_changes_nothing_and_aborts():
    g, d = _gate_engine()
    g.rpie(1, batch_size=5)
    before = (g.object, g.probe, g.cycle)
    try:
        pb.arm_fault("launch", 0)
        with pytest.raises(RuntimeError, match="injected fault"):
            g.rpie(1, batch_size=5)
        assert pb.fault_fired()
    finally:
        pb.arm_fault(None)
    # Downloads run on the aborted engine and see the state before the call.
    np.testing.assert_array_equal(g.object, before[0])
    np.testing.assert_array_equal(g.probe, before[1])
    assert g.cycle == before[2]
    # Every mutating call is refused, whatever its arguments; reads still run.
    for call in (lambda: g.rpie(1, batch_size=5), lambda: g.ap(1), lambda: g.llk(),
                 lambda: g.set_object(before[0]), lambda: g.rpie(1, batch_size=0),
                 lambda: g.lsqml(1, 0, 5, "poisson", 1)):
        with pytest.raises(RuntimeError, match="aborted"):
            call()
    np.testing.assert_array_equal(g.object, before[0])
    assert len(g.history["poisson"]) == 0


@pytest.mark.parametrize("algorithm", ["rpie", "ap", "dm", "ml", "lsqml"])
def test_a_failure_part_way_aborts_every_algorithm(algorithm):
    g, d = _gate_engine()
    calls = {"rpie": lambda: g.rpie(2, batch_size=5), "ap": lambda: g.ap(3),
             "dm": lambda: g.dm(3), "ml": lambda: g.ml(3),
             "lsqml": lambda: g.lsqml(2, 0, 5, "poisson", 1)}
    try:
        pb.arm_fault("launch", 7)
        with pytest.raises(RuntimeError, match="injected fault"):
            calls[algorithm]()
    finally:
        pb.arm_fault(None)
    with pytest.raises(RuntimeError, match="aborted"):
        calls[algorithm]()
    # The part-changed state downloads; a new engine on the same inputs runs.
    assert np.all(np.isfinite(g.object)) and np.all(np.isfinite(g.probe))
    g2 = pa.GpuPtycho(pa.Scan(d["counts"], d["y"], d["x"], *GEOM), d["obj"], d["probe"],
                      dtype="c128")
    g2.rpie(1, batch_size=5)


def test_a_failed_capture_aborts_and_a_new_engine_replays():
    d = small_scan(n=32, nprobe=1, npos=12)
    g = pa.GpuPtycho(pa.Scan(d["counts"], d["y"], d["x"], *GEOM), d["obj"], d["probe"])
    try:
        pb.arm_fault("capture", 0)
        with pytest.raises(RuntimeError, match="injected fault"):
            g.rpie(1, batch_size=5)
    finally:
        pb.arm_fault(None)
    with pytest.raises(RuntimeError, match="aborted"):
        g.rpie(1, batch_size=5)
    g2 = pa.GpuPtycho(pa.Scan(d["counts"], d["y"], d["x"], *GEOM), d["obj"], d["probe"])
    g3 = pa.GpuPtycho(pa.Scan(d["counts"], d["y"], d["x"], *GEOM), d["obj"], d["probe"],
                      graph=False)
    g2.rpie(2, batch_size=5)
    g3.rpie(2, batch_size=5)
    np.testing.assert_array_equal(g2.object, g3.object)


@pytest.mark.parametrize("slots", [1, 3, 64])
def test_llk_slots_change_no_record(slots):
    # GpuOptions::llk_slots (ptyrs/src/engine/gpu/mod.rs:80-84): how many records the device
    # holds before the host drains them, not what they are.
    d = small_scan(n=32, nprobe=2, npos=12)
    hs = []
    for kw in ({}, {"llk_slots": slots}):
        g = pa.GpuPtycho(pa.Scan(d["counts"], d["y"], d["x"], *GEOM), d["obj"], d["probe"],
                         dtype="c128", **kw)
        g.rpie(9, batch_size=5, calc_llk=1)
        g.ap(5, calc_llk=2)
        g.lsqml(4, 1, 5, "poisson", 1)
        h = g.history
        hs.append((h["cycle"], h["poisson"], h["gaussian"]))
    for a, b in zip(*hs):
        np.testing.assert_array_equal(a, b)
    assert len(hs[0][0]) == 9 + 3 + 4


@pytest.mark.parametrize("stack", [1, 5, 64])
def test_stack_size_cuts_change_results_only_by_rounding(stack):
    # GpuOptions::stack_size: frames per stack of AP, DM, RAAR and ML; the sums keep view order.
    d = small_scan(n=32, nprobe=2, npos=23)
    out = []
    for kw in ({}, {"stack_size": stack, "fused": False}):
        g = pa.GpuPtycho(pa.Scan(d["counts"], d["y"], d["x"], *GEOM), d["obj"], d["probe"],
                         dtype="c128", **kw)
        g.ap(3, calc_llk=1)
        g.dm(3, calc_llk=1)
        g.ml(3, calc_llk=1)
        out.append((g.object, g.probe, g.history["poisson"]))
    for a, b in zip(*out):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-10 * np.abs(b).max())
