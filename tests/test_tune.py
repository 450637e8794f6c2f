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
