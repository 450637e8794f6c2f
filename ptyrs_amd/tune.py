"""Launch-shape autotuning.

`tune_engine` picks, per size class, the launch shapes of one engine that change results only
by rounding: frames per stack of AP, DM, RAAR and ML (`stack_size`), and the workgroups per
launch that rPIE's (`pass_c_blocks`) and LSQML's (`chi_blocks`) pass C cut a batch into chunks
for. It times one epoch (or cycle) of a copy of the engine for each candidate after one
untimed epoch (which captures the candidate's graph), keeps the fastest, caches it per size
class — the algorithm, frame side, modes, slices, dtype, object plane and frames rounded to
powers of two — and sets it on the engine.

Several scans reconstructed at once:

`rpie_group` (ptyrs_group_rpie) runs every engine's rPIE epochs side by side, each on its own
stream, or (ptyrs_group_rpie_graph) in one graph per epoch. How many scans share the device at
a time, and which of the two launches, are choices that leave every result unchanged (each
engine's epochs are bit-identical however the scans are grouped), so they can be tuned on the
reconstruction's own epochs: `tune_group` times one epoch of every scan for each candidate
(scans taken largest first, in waves of `width` scans), keeps the fastest, and caches it per
size class — the multiset of (frame side, modes, dtype) and positions rounded to a power of
two — in a JSON file.
"""

import json
import os
import pathlib
import time

from . import rpie_group

CACHE = pathlib.Path(os.environ.get("PTYRS_TUNE_CACHE",
                                    pathlib.Path.home() / ".cache" / "ptyrs_amd_tune.json"))
WIDTHS = (1, 2, 4, 8, 16)


def size_class(engines):
    """The tuning key of a set of engines."""
    parts = []
    for g in engines:
        n = g._probe_shape[-1]
        pos = 1 << max(int(g.nframe) - 1, 0).bit_length()
        parts.append(f"{n}x{g._probe_shape[0]}:{g.dtype}:{pos}")
    return ",".join(sorted(parts))


def _order(engines):
    """Largest scans first (frames × plane × modes), so waves start with the longest work."""
    return sorted(engines, key=lambda g: -g.nframe * g._probe_shape[-1] ** 2 * g._probe_shape[0])


def run_waves(engines, nepoch, params=None, width=None, one_graph=False):
    """`nepoch` rPIE epochs of every engine, `width` scans at a time (all at once if None),
    each wave's epochs on the engines' streams or in one graph per epoch."""
    engines = list(engines)
    if params is None or isinstance(params, dict):
        params = [params or {}] * len(engines)
    pairs = list(zip(engines, params))
    order = _order(engines)
    pairs.sort(key=lambda gp: order.index(gp[0]))
    width = width or len(pairs)
    for i in range(0, len(pairs), width):
        wave = pairs[i:i + width]
        rpie_group([g for g, _ in wave], nepoch, [p for _, p in wave], one_graph)


def _load():
    try:
        return json.loads(CACHE.read_text())
    except (OSError, ValueError):
        return {}


def tune_group(engines, params=None, widths=WIDTHS, cache=True, modes=(False, True)):
    """The fastest launch for these engines: a width and whether each wave's epochs go on the
    engines' streams (one_graph False) or in one graph per epoch. The candidates are timed on
    copies of the engines (same scan, options, object and probe), one rPIE epoch of every copy
    each, so the caller's engines are left as they were, cache hit or miss. Returns {"width",
    "one_graph", "seconds": {"width/mode": s}, "key", "cached"}."""
    engines = list(engines)
    key = size_class(engines)
    if cache:
        hit = _load().get(key)
        if hit:
            return dict(hit, key=key, cached=True)
    copies = [g._clone() for g in engines]
    times = {}
    for w in widths:
        if w > len(copies) and w != widths[0]:
            continue
        for one in modes:
            for g in copies:
                g.synchronize()
            t = time.perf_counter()
            run_waves(copies, 1, params, w, one)
            times[(w, one)] = time.perf_counter() - t
    del copies
    best = min(times, key=times.get)
    out = {"width": best[0], "one_graph": best[1],
           "seconds": {f"{w}/{'graph' if one else 'streams'}": v for (w, one), v in times.items()}}
    if cache:
        data = _load()
        data[key] = out
        try:
            CACHE.parent.mkdir(parents=True, exist_ok=True)
            CACHE.write_text(json.dumps(data, indent=1))
        except OSError:
            pass
    return dict(out, key=key, cached=False)


# Candidates per algorithm: the launch key and its values.
ENGINE_CANDIDATES = {
    "rpie": ("pass_c_blocks", (512, 768, 1024, 1536)),
    "lsqml": ("chi_blocks", (512, 1024, 2048)),
    "ap": ("stack_size", (32, 64, 128, 256)),
    "dm": ("stack_size", (32, 64, 128, 256)),
    "raar": ("stack_size", (32, 64, 128, 256)),
    "ml": ("stack_size", (32, 64, 128, 256)),
}


def _pow2(x):
    return 1 << max(int(x) - 1, 0).bit_length()


def engine_class(g, algorithm):
    """The tuning key of one engine and algorithm."""
    m, ny, nx = g._probe_shape
    obj = g.objects.shape[-2:] if g.nslice > 1 else g.object.shape[-2:]
    return (f"{algorithm}:{nx}x{m}:s{g.nslice}:{g.dtype}:{_pow2(obj[0])}x{_pow2(obj[1])}:"
            f"{_pow2(g.nframe)}")


def _run(g, algorithm, params):
    p = dict(params or {})
    if algorithm == "lsqml":
        g.lsqml(1, 0, p.get("batch_size", 100), p.get("cost", "poisson"), p.get("seed", 7))
    else:
        getattr(g, algorithm)(1, **p)


def tune_engine(g, algorithm, params=None, values=None, cache=True, apply=True):
    """The fastest launch shape for `algorithm` (one of ENGINE_CANDIDATES) on engine `g` with
    `params` (the algorithm's keywords; LSQML's `batch_size`, `cost`, `seed`): times one epoch
    of a copy of `g` per candidate value (`values`, else the defaults), keeps the fastest,
    caches it per `engine_class` and, with `apply`, sets it on `g`. The copy starts from `g`'s
    state, so `g` is left as it was. Returns {"key", "name", "value", "seconds", "cached"}."""
    if algorithm not in ENGINE_CANDIDATES:
        raise ValueError(f"algorithm must be one of {sorted(ENGINE_CANDIDATES)}")
    name, default = ENGINE_CANDIDATES[algorithm]
    key = engine_class(g, algorithm)
    out = None
    if cache:
        hit = _load().get(key)
        if hit:
            out = dict(hit, key=key, cached=True)
    if out is None:
        copy = g._clone()
        times = {}
        for v in (values or default):
            copy.set_launch(**{name: v})
            _run(copy, algorithm, params)  # captures the candidate's graph
            copy.synchronize()
            t = time.perf_counter()
            _run(copy, algorithm, params)
            copy.synchronize()
            times[v] = time.perf_counter() - t
        del copy
        best = min(times, key=times.get)
        out = {"name": name, "value": best, "seconds": {str(v): t for v, t in times.items()}}
        if cache:
            data = _load()
            data[key] = out
            try:
                CACHE.parent.mkdir(parents=True, exist_ok=True)
                CACHE.write_text(json.dumps(data, indent=1))
            except OSError:
                pass
        out = dict(out, key=key, cached=False)
    if apply:
        g.set_launch(**{out["name"]: out["value"]})
    return out
