"""Every algorithm at the frame sides and mode counts whose launch shapes differ most (the
smallest side, the sides with a factor 3, the largest side; one mode and four), in both
precisions, against the FP64 oracle on the same seeded far-field inputs: AP, DM, RAAR and ML
one LLK record per cycle, LSQML the LLK after each epoch, with the final object and probe."""

import pytest

from conftest import max_rel
from test_gpu_lsqml import run_oracle
from test_gpu_rpie import gpu, oracle, rel_l2, small_scan

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-9, "c64": 1e-4}
SIDES = [16, 24, 96, 512]


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("nprobe", [1, 4])
@pytest.mark.parametrize("n", SIDES)
@pytest.mark.parametrize("name,kw", [("ap", dict(probe_inertia=0.001)), ("dm", {}),
                                     ("raar", {}), ("ml", dict(update_probe=True))])
def test_projections_and_ml_at_every_shape(name, kw, n, nprobe, dtype):
    d = small_scan(n=n, nprobe=nprobe, npos=6, masked=False)
    g, o = gpu(d, dtype), oracle(d)
    getattr(g, name)(2, calc_llk=1, **kw)
    getattr(o, name)(2, calc_llk=1, **kw)
    assert max_rel(g.history["poisson"], [r[2].poisson for r in o.history]) < TOL[dtype] * 10
    tol = TOL[dtype] * (10 if dtype == "c128" else 30)
    assert rel_l2(g.object, o.obj) < tol
    assert rel_l2(g.probe, o.probe) < tol


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("nprobe", [1, 4])
@pytest.mark.parametrize("n", SIDES)
def test_lsqml_at_every_shape(n, nprobe, dtype):
    d = small_scan(n=n, nprobe=nprobe, npos=6, masked=False)
    g, o = gpu(d, dtype), oracle(d)
    obj, probe, llks, _ = run_oracle(o, 2, 4, "poisson", 5)
    g.lsqml(2, 1, 4, "poisson", 5)
    assert max_rel(g.history["poisson"], llks) < TOL[dtype] * 10
    tol = TOL[dtype] * (10 if dtype == "c128" else 30)
    assert rel_l2(g.object - d["obj"], obj - d["obj"]) < tol
    assert rel_l2(g.probe - d["probe"], probe - d["probe"]) < tol
