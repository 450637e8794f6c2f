"""Shared test setup: the `gpu` marker, the repo root on sys.path, fixture tables."""

import gzip
import os
import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def table(name):
    """Rows of whitespace-separated floats of a reference fixture (`inf` parses)."""
    with gzip.open(GOLDEN / (name + ".gz"), "rt") as f:
        rows = [[float(v) for v in line.split()] for line in f if line.strip()]
    return rows


def worst_of(values):
    """Largest value, NaN anywhere fails (the reference's rule, ptyrs/tests/common/worst.rs:1-8)."""
    v = np.asarray(list(values), dtype=np.float64)
    if v.size == 0:
        return 0.0
    if np.isnan(v).any():
        return float("inf")
    return float(v.max())


def max_rel(a, b):
    """`max |a − b| / max |b|` (ptyrs/tests/engine_rpie_near.rs:25-30)."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, (a.shape, b.shape)
    return worst_of(np.abs(a - b).ravel()) / np.abs(b).max()
