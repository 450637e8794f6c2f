"""Several independent scans in one call (ptyrs_group_rpie): each engine's rPIE epochs run on
its own stream beside the others and give the same bits as the engine run alone; parameters
are checked for every engine before any changes."""

import numpy as np
import pytest

import ptyrs_amd as pa
from ptyrs_amd import sim
from test_gpu_rpie import small_scan

pytestmark = pytest.mark.gpu

SCANS = [dict(n=64, nprobe=1, npos=30, seed=3), dict(n=128, nprobe=2, npos=20, seed=4),
         dict(n=32, nprobe=3, npos=41, seed=5), dict(n=64, nprobe=4, npos=12, seed=6)]
PARAMS = [dict(batch_size=7, calc_llk=1), dict(batch_size=20, calc_llk=2, obj_alpha=0.2),
          dict(batch_size=100, calc_llk=1, probe_step=None), dict(batch_size=3, seed=11)]


def engines(dtypes, **options):
    out = []
    for s, dt in zip(SCANS, dtypes):
        d = small_scan(**s)
        scan = pa.Scan(d["counts"], d["y"], d["x"], sim.WAVELENGTH, sim.DISTANCE,
                       sim.DETECTOR_PIXEL)
        out.append(pa.GpuPtycho(scan, d["obj"], d["probe"], dtype=dt, **options))
    return out


@pytest.mark.parametrize("graph,one_graph", [(True, False), (False, False), (True, True)])
def test_group_matches_engines_run_alone(graph, one_graph):
    dtypes = ["c64", "c128", "c64", "c128"]
    alone = engines(dtypes)
    for g, p in zip(alone, PARAMS):
        g.rpie(3, **p)
    group = engines(dtypes, graph=graph)
    pa.rpie_group(group, 3, PARAMS, one_graph=one_graph)
    for a, b in zip(alone, group):
        assert a.cycle == b.cycle == 3
        assert np.array_equal(a.object, b.object)
        assert np.array_equal(a.probe, b.probe)
        ha, hb = a.history, b.history
        assert np.array_equal(ha["cycle"], hb["cycle"])
        assert np.array_equal(ha["poisson"], hb["poisson"])


def test_group_checks_every_engine_first():
    group = engines(["c64"] * 4)
    before = [(g.object, g.probe) for g in group]
    bad = [dict(p) for p in PARAMS]
    bad[2]["batch_size"] = 0
    with pytest.raises(ValueError):
        pa.rpie_group(group, 2, bad)
    for g, (o, p) in zip(group, before):
        assert g.cycle == 0
        assert np.array_equal(g.object, o) and np.array_equal(g.probe, p)
    with pytest.raises(ValueError):
        pa.rpie_group([group[0], group[0]], 1)
    pa.rpie_group(group, 1, {"batch_size": 10})
    assert [g.cycle for g in group] == [1] * 4


def test_tuned_waves_give_the_same_bits(tmp_path, monkeypatch):
    from ptyrs_amd import tune
    monkeypatch.setattr(tune, "CACHE", tmp_path / "tune.json")
    dtypes = ["c64"] * 4
    a, b = engines(dtypes), engines(dtypes)
    start = [(g.object, g.probe) for g in a]
    r = tune.tune_group(a, PARAMS, widths=(1, 2, 4))
    assert r["width"] in (1, 2, 4) and r["one_graph"] in (False, True) and not r["cached"]
    assert len(r["seconds"]) == 6
    # Tuning times copies: the engines are where they were, on a miss as on a hit.
    for g, (o, p) in zip(a, start):
        assert g.cycle == 0 and len(g.history["cycle"]) == 0
        assert np.array_equal(g.object, o) and np.array_equal(g.probe, p)
    assert tune.tune_group(a, PARAMS)["cached"]
    tune.run_waves(a, 3, PARAMS, r["width"], r["one_graph"])
    for g, p in zip(b, PARAMS):
        g.rpie(3, **p)
    for g, h in zip(a, b):
        assert g.cycle == h.cycle == 3
        assert np.array_equal(g.object, h.object) and np.array_equal(g.probe, h.probe)


def test_one_graph_group_with_an_eager_engine_runs_it_eagerly():
    # An engine built with graph=False (or profiling) is not captured: the group falls back to
    # the per-stream launch, with the bits of each engine run alone.
    a = engines(["c64", "c64"])
    b = engines(["c64", "c64"])
    eager = engines(["c64"], graph=False)[0]
    alone = engines(["c64"], graph=False)[0]
    pa.rpie_group(a + [eager], 2, PARAMS[:2] + [PARAMS[2]], one_graph=True)
    for g, p in zip(b + [alone], PARAMS[:3]):
        g.rpie(2, **p)
    for g, h in zip(a + [eager], b + [alone]):
        assert np.array_equal(g.object, h.object) and np.array_equal(g.probe, h.probe)
    a[0].profile(True)
    pa.rpie_group(a, 1, PARAMS[:2], one_graph=True)
    assert a[0].kernel_stats()["pass_a"][0] > 0
