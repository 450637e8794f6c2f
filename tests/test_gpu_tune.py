"""The launch-shape tuner on the GPU (ptyrs_amd.tune.tune_engine, ptyrs_engine_set_launch): it
picks one of its candidates on a copy of the engine and sets it, leaving the engine's state as
it was; every launch shape gives the default's results to rounding (the pass C chunks only
reorder the probe direction's partial sums; the stacks only cut the frames), and a captured
epoch is not replayed under another launch shape."""

import numpy as np
import pytest

from ptyrs_amd import tune
from test_gpu_rpie import gpu, small_scan

pytestmark = pytest.mark.gpu


def test_rpie_and_ap_tuning_keep_results(tmp_path, monkeypatch):
    monkeypatch.setattr(tune, "CACHE", tmp_path / "tune.json")
    d = small_scan(n=64, nprobe=2, npos=60, seed=23)
    g = gpu(d, "c128")
    o0, p0 = g.object, g.probe
    r = tune.tune_engine(g, "rpie", {"batch_size": 30})
    assert r["value"] in tune.ENGINE_CANDIDATES["rpie"][1] and not r["cached"]
    assert g.launch["pass_c_blocks"] == r["value"]
    assert g.cycle == 0 and np.array_equal(g.object, o0) and np.array_equal(g.probe, p0)
    ref = gpu(d, "c128")
    for e in (g, ref):
        e.rpie(3, batch_size=30, calc_llk=1)
        e.ap(2, calc_llk=1)
    assert np.abs(g.object - ref.object).max() < 1e-12 * np.abs(ref.object).max()
    assert np.abs(g.probe - ref.probe).max() < 1e-12 * np.abs(ref.probe).max()
    r = tune.tune_engine(g, "ap", {"calc_llk": 0})
    assert g.launch["stack_size"] == r["value"]
    assert tune.tune_engine(g, "ap", {"calc_llk": 0})["cached"]


@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_changing_the_launch_recaptures_the_epoch(dtype):
    d = small_scan(n=64, nprobe=3, npos=40, seed=24)
    out = []
    for graph in (True, False):
        g = gpu(d, dtype, graph=graph)
        g.rpie(1, batch_size=20)
        g.set_launch(pass_c_blocks=256, chi_blocks=256, stack_size=8)
        g.rpie(1, batch_size=20)
        g.lsqml(1, 0, 20, "poisson", 3)
        g.dm(1)
        g.set_launch(pass_c_blocks=0, chi_blocks=0, stack_size=0)
        assert g.launch == dict(stack_size=0, pass_c_blocks=768, chi_blocks=1024)
        g.rpie(1, batch_size=20)
        out.append((g.object, g.probe))
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_launch_is_checked():
    g = gpu(small_scan(n=32, nprobe=1, npos=8), "c64")
    with pytest.raises(ValueError):
        g.set_launch(pass_c_blocks=9000)
    with pytest.raises(ValueError):
        g.set_launch(stack_size=70000)
    with pytest.raises(TypeError):
        g.set_launch(tiles=4)
    assert g.launch == dict(stack_size=0, pass_c_blocks=768, chi_blocks=1024)
