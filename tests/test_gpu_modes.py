"""More than four views per frame (probe modes × object modes, up to the engine's eight): every
algorithm against the FP64 oracle on the same seeded far-field inputs, at sides whose launch
shapes differ (pass B's per-mode thread groups fill a 1024-thread workgroup at eight modes;
rPIE's pass C sums eight modes through one term buffer at 512² in complex128). Nine views are
refused. The reference takes any count by cutting the modes into chunks (ptyrs/src/gpu/
fused.rs:78-82)."""

import numpy as np
import pytest

import ptyrs_amd as pa
from conftest import max_rel
from ptyrs_amd import sim
from test_gpu_lsqml import run_oracle
from test_gpu_rpie import gpu, oracle, rel_l2, small_scan

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-9, "c64": 1e-4}


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n,nprobe", [(32, 5), (64, 8), (256, 6), (512, 8)])
def test_rpie_with_five_to_eight_modes_matches_oracle(n, nprobe, dtype):
    d = small_scan(n=n, nprobe=nprobe, npos=6, masked=False)
    g, o = gpu(d, dtype), oracle(d)
    o0, p0 = o.obj.copy(), o.probe.copy()
    g.rpie(2, batch_size=4, calc_llk=1)
    o.rpie(2, batch_size=4, calc_llk=1)
    assert max_rel(g.history["poisson"], [r[2].poisson for r in o.history]) < TOL[dtype]
    assert rel_l2(g.object - o0, o.obj - o0) < TOL[dtype] * 10
    assert rel_l2(g.probe - p0, o.probe - p0) < TOL[dtype] * 10


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("name,kw", [("ap", dict(probe_inertia=0.001)), ("dm", {}),
                                     ("raar", {}), ("ml", dict(update_probe=True))])
def test_projections_and_ml_with_eight_modes_match_oracle(name, kw, dtype):
    d = small_scan(n=64, nprobe=8, npos=8)
    g, o = gpu(d, dtype), oracle(d)
    getattr(g, name)(2, calc_llk=1, **kw)
    getattr(o, name)(2, calc_llk=1, **kw)
    assert max_rel(g.history["poisson"], [r[2].poisson for r in o.history]) < TOL[dtype] * 10
    tol = TOL[dtype] * (10 if dtype == "c128" else 30)
    assert rel_l2(g.object, o.obj) < tol
    assert rel_l2(g.probe, o.probe) < tol


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_lsqml_and_llk_with_seven_modes_match_oracle(dtype):
    d = small_scan(n=96, nprobe=7, npos=8)
    g, o = gpu(d, dtype), oracle(d)
    assert abs(g.llk()["poisson"] - o.llk().poisson) < TOL[dtype] * o.llk().poisson
    obj, probe, llks, _ = run_oracle(o, 2, 4, "poisson", 5)
    g.lsqml(2, 1, 4, "poisson", 5)
    assert max_rel(g.history["poisson"], llks) < TOL[dtype] * 10
    tol = TOL[dtype] * (10 if dtype == "c128" else 30)
    assert rel_l2(g.object - d["obj"], obj - d["obj"]) < tol
    assert rel_l2(g.probe - d["probe"], probe - d["probe"]) < tol


def test_nine_views_per_frame_are_refused():
    d = small_scan(n=32, nprobe=3, npos=6)
    obj = np.stack([d["obj"][0]] * 3)  # three object modes × three probe modes
    with pytest.raises(ValueError, match="at most 8 modes"):
        pa.GpuPtycho(pa.Scan(d["counts"], d["y"], d["x"], sim.WAVELENGTH, sim.DISTANCE,
                             sim.DETECTOR_PIXEL), obj, d["probe"])
