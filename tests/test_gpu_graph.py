"""Captured epochs (GpuOptions graph=True, ptyrs/src/engine/gpu/mod.rs:61-108) against eager
launches (graph=False): rPIE and LSQML replayed over several epochs, with the eager LLK passes
and order copies between replays, across calls (the same graph reused) and after a parameter
change (recaptured). The kernels are the same, so the results must be bit-identical."""

import numpy as np
import pytest

from test_gpu_rpie import gpu, small_scan

pytestmark = pytest.mark.gpu


def engines(d, dtype="c64"):
    return [gpu(d, dtype, graph=g) for g in (True, False)]


def same(a, b):
    assert np.array_equal(a.object, b.object)
    assert np.array_equal(a.probe, b.probe)
    assert np.array_equal(a.history["poisson"], b.history["poisson"])


@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_rpie_graph_replays_match_eager(dtype):
    d = small_scan(n=64, nprobe=2, npos=30)
    g, e = engines(d, dtype)
    for eng in (g, e):
        eng.rpie(4, batch_size=7, calc_llk=1)
        eng.rpie(3, batch_size=7, calc_llk=2)
        eng.rpie(2, batch_size=7, obj_alpha=0.2, calc_llk=1)
        eng.rpie(2, batch_size=11, probe_step=None)
    same(g, e)


@pytest.mark.parametrize("cost", ["poisson", "amplitude"])
def test_lsqml_graph_replays_match_eager(cost):
    d = small_scan(n=64, nprobe=2, npos=30)
    g, e = engines(d)
    for eng in (g, e):
        eng.lsqml(12, 2, 7, cost, 3)
        eng.lsqml(3, 1, 9, cost, 3)
    same(g, e)


@pytest.mark.parametrize("name,kw", [("ap", dict(probe_inertia=0.001, calc_llk=1)),
                                     ("dm", dict(calc_llk=2)),
                                     ("raar", dict(calc_llk=1, loop_obj_probe=2))])
def test_projection_graph_replays_match_eager(name, kw):
    d = small_scan(n=64, nprobe=2, npos=30)
    g, e = engines(d)
    for eng in (g, e):
        getattr(eng, name)(5, **kw)
        getattr(eng, name)(3, **kw)
    same(g, e)


def test_ml_graph_replays_match_eager():
    d = small_scan(n=64, nprobe=2, npos=30)
    g, e = engines(d)
    for eng in (g, e):
        eng.ap(3, zero_phase_ramp=False)
        eng.ml(5, update_probe=True, calc_llk=1)
        eng.ml(3, calc_llk=2)
    same(g, e)
