"""Live ingestion on the GPU (ptyrs_engine_live_reset / _push / _frames; LiveAp::push and
epoch, ptyrs/src/live/recon.rs:1-25, :184-266): the frames of a planned scan land in order and
each epoch is the engine's AP cycle over the frames landed so far with the probe fixed, as the
reference's live AP covers "exactly the frames landed when it was queued". Checked against
engines built on the landed frames alone (the same views, by explicit corners) from the same
state: object, probe, LLK records and model intensities, complex128 to 1e-12 and complex64 to
1e-5; AP itself is pinned against the oracle in tests/test_gpu_projection.py."""

import numpy as np
import pytest

import ptyrs_amd as pa
from conftest import max_rel
from ptyrs_amd import sim
from test_gpu_rpie import rel_l2, small_scan

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 1e-5}


def _scan(d, corners, k):
    return pa.Scan(d["counts"][:k], d["y"][:k], d["x"][:k], sim.WAVELENGTH, sim.DISTANCE,
                   sim.DETECTOR_PIXEL, corners=corners[:k])


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_live_ap_epochs_are_the_engines_ap_over_the_landed_frames(dtype):
    n, npos = 32, 40
    d = small_scan(n=n, nprobe=2, npos=npos, seed=41)
    cy, cx = sim.corners(d["y"], d["x"], sim.pixel_size(n), d["obj"].shape[-2:], n)
    corners = np.stack([cy, cx], axis=1)
    live = pa.GpuPtycho(_scan(d, corners, npos), d["obj"], d["probe"], dtype=dtype)
    assert live.live_frames == (npos, npos)
    live.live_reset()
    assert live.live_frames == (0, npos)
    before = live.object
    live.ap(1, zero_phase_ramp=False)   # nothing landed: an epoch changes nothing
    assert np.array_equal(live.object, before)
    params = dict(calc_llk=1, zero_phase_ramp=False)
    obj = d["obj"]
    tol = TOL[dtype]
    counts16 = np.where(d["counts"] < 0, 65535, d["counts"]).astype(np.uint16)
    for k0, k1 in ((0, 16), (16, 29), (29, 40)):
        # uint16 counts for one landing, float for the others: the same conversion as the scan's
        live.live_push(counts16[k0:k1] if k0 == 16 else d["counts"][k0:k1])
        assert live.live_frames == (k1, npos)
        before = len(live.history["poisson"])
        live.ap(2, **params)
        ref = pa.GpuPtycho(_scan(d, corners, k1), obj, d["probe"], dtype=dtype)
        ref.ap(2, **params)
        print("live", dtype, k1, rel_l2(live.object, ref.object),
              max_rel(live.history["poisson"][before:], ref.history["poisson"]))
        assert rel_l2(live.object, ref.object) < tol
        assert np.array_equal(live.probe, ref.probe)   # AP's probe stays fixed
        assert max_rel(live.history["poisson"][before:], ref.history["poisson"]) < tol
        obj = ref.object
    assert rel_l2(live.intensities(), ref.intensities()) < tol
    with pytest.raises(ValueError):
        live.live_push(d["counts"][:1])   # past the planned frames
    live.live_reset()
    with pytest.raises(ValueError):
        live.live_push(np.concatenate([d["counts"], d["counts"][:1]]))
    assert live.live_frames == (0, npos)
