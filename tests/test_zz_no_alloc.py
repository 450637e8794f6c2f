"""ptyrs_engine_llk_noalloc, the LLK entry point that allocates nothing (its work buffers live
in one static device array per GPU), against ptyrs_engine_llk through the C ABI: the same bits
on two calls in a row, the second on a smaller scan, with stacks cut by the array's size."""

import numpy as np
import pytest

import ptyrs_amd as pa
from ptyrs_amd import sim
from test_gpu_rpie import small_scan

pytestmark = pytest.mark.gpu

GEOM = (sim.WAVELENGTH, sim.DISTANCE, sim.DETECTOR_PIXEL)


def engine(dtype, **kw):
    d = small_scan(**kw)
    return pa.GpuPtycho(pa.Scan(d["counts"], d["y"], d["x"], *GEOM), d["obj"], d["probe"],
                        dtype=dtype)


@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_noalloc_llk_equals_llk_on_two_calls_the_second_smaller(dtype):
    # 256² × 4 modes × 40 frames: 2 MiB (c64) or 4 MiB (c128) of exit waves per frame, so the
    # 64 MiB array cuts the frames into two or three stacks; then a 32² single-mode scan of 7
    # frames on the array the first call left behind.
    big = engine(dtype, n=256, nprobe=4, npos=40, seed=3)
    small = engine(dtype, n=32, nprobe=1, npos=7, seed=4)
    for g in (big, small):
        want = g.llk()
        got = g.llk_noalloc()
        assert got == want, (got, want)
    # After the state changes, and the larger scan again after the smaller one.
    big.rpie(1, batch_size=9)
    small.rpie(1, batch_size=3)
    for g in (big, small):
        assert g.llk_noalloc() == g.llk()


def test_noalloc_llk_refuses_what_it_does_not_run():
    g = engine("c64", n=32, nprobe=1, npos=7, seed=5)
    g.set_features(position=np.zeros((7, 2)))
    with pytest.raises(ValueError, match="allocation-free"):
        g.llk_noalloc()
    g.set_features(position=None)
    g.profile(True)
    with pytest.raises(ValueError, match="profiling"):
        g.llk_noalloc()
    g.profile(False)
    assert g.llk_noalloc() == g.llk()
