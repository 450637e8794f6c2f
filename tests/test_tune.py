"""The width tuner's host logic (no device): size-class keys, the largest-first order, the
waves it cuts and its cache, with stand-in engines and a stand-in group call."""

import types

import pytest

from ptyrs_amd import tune


def fake(n, modes, frames, dtype="c64", clones=None):
    g = types.SimpleNamespace(_probe_shape=(modes, n, n), nframe=frames, dtype=dtype,
                              synchronize=lambda: None)
    def clone():
        c = fake(n, modes, frames, dtype)
        if clones is not None:
            clones.append(c)
        return c
    g._clone = clone
    return g


def test_size_class_rounds_positions_and_ignores_order():
    a = [fake(128, 2, 300), fake(256, 1, 1000)]
    b = [fake(256, 1, 900), fake(128, 2, 257)]
    assert tune.size_class(a) == tune.size_class(b)
    assert tune.size_class(a) != tune.size_class([fake(128, 2, 300, "c128"), fake(256, 1, 1000)])


def test_waves_are_largest_first_and_cover_every_engine(monkeypatch):
    calls = []
    monkeypatch.setattr(tune, "rpie_group", lambda gs, nepoch, ps, one: calls.append((gs, ps)))
    es = [fake(128, 1, 200), fake(256, 4, 100), fake(128, 2, 3000), fake(256, 1, 50)]
    ps = [{"batch_size": k} for k in range(4)]
    tune.run_waves(es, 1, ps, width=3)
    assert [len(g) for g, _ in calls] == [3, 1]
    flat = [g for gs, _ in calls for g in gs]
    assert sorted(map(id, flat)) == sorted(map(id, es))
    sizes = [g.nframe * g._probe_shape[-1] ** 2 * g._probe_shape[0] for g in flat]
    assert sizes == sorted(sizes, reverse=True)
    # Each engine keeps its own parameters.
    for gs, pp in calls:
        for g, p in zip(gs, pp):
            assert p == ps[es.index(g)]


def test_tuner_caches_its_choice(tmp_path, monkeypatch):
    monkeypatch.setattr(tune, "CACHE", tmp_path / "t.json")
    ran = []
    monkeypatch.setattr(tune, "rpie_group", lambda gs, nepoch, ps, one: ran.extend(gs))
    clones = []
    es = [fake(128, 1, 200, clones=clones), fake(256, 4, 100, clones=clones)]
    r = tune.tune_group(es, widths=(1, 2))
    # Only copies of the engines run epochs.
    assert ran and all(any(g is c for c in clones) for g in ran)
    assert not any(g is e for g in ran for e in es)
    assert r["width"] in (1, 2) and not r["cached"]
    again = tune.tune_group(es, widths=(1, 2))
    assert again["cached"] and again["width"] == r["width"]


def fake_engine(times, algorithm_calls, nframe=300):
    """A stand-in engine whose copies run `algorithm` in `times[launch value]` seconds."""
    state = {"launch": dict(stack_size=0, pass_c_blocks=768, chi_blocks=1024)}
    g = types.SimpleNamespace(_probe_shape=(4, 256, 256), nframe=nframe, dtype="c64",
                              nslice=1, object=types.SimpleNamespace(shape=(1, 1000, 1000)),
                              synchronize=lambda: None)
    g.launch = state["launch"]

    def set_launch(**kw):
        state["launch"].update(kw)
    g.set_launch = set_launch

    def clone():
        c = fake_engine(times, algorithm_calls, nframe)
        c.copy = True
        return c
    g._clone = clone

    def run(*args, **kw):
        algorithm_calls.append(dict(state["launch"]))
        import time as _t
        key = state["launch"]["pass_c_blocks"]
        _t.sleep(times.get(key, 0.0))
    g.rpie = run
    return g


def test_engine_tuner_times_copies_and_caches(tmp_path, monkeypatch):
    monkeypatch.setattr(tune, "CACHE", tmp_path / "t.json")
    calls = []
    g = fake_engine({512: 0.03, 768: 0.001, 1024: 0.02}, calls)
    r = tune.tune_engine(g, "rpie", {"batch_size": 100}, values=(512, 768, 1024))
    assert r["name"] == "pass_c_blocks" and r["value"] == 768 and not r["cached"]
    assert len(calls) == 6   # one untimed and one timed epoch per candidate, on the copy
    assert g.launch["pass_c_blocks"] == 768
    assert tune.engine_class(g, "rpie") == "rpie:256x4:s1:c64:1024x1024:512"
    again = tune.tune_engine(g, "rpie", {"batch_size": 100})
    assert again["cached"] and again["value"] == 768 and len(calls) == 6
    with pytest.raises(ValueError):
        tune.tune_engine(g, "bh")
