"""The C++ FP64 far-field rPIE (`cpu_rpie.cpp`): the CPU baseline timed beside the GPU.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): `bench.py`'s `cpu_baseline` times it and
tests/test_oracle_cpu.py checks it against the numpy oracle's `Ptycho.rpie`
(ptyrs/src/engine/ptycho.rs:844-1086). `build()` compiles it with g++ and OpenMP (`library()` builds
it when it is missing); the hot loops carry an FMA clone chosen at load time, so one build runs
on any x86-64 host.
"""

import ctypes
import pathlib
import subprocess

import numpy as np

from .base import random_batches

HERE = pathlib.Path(__file__).resolve().parent
SRC = HERE / "cpu_rpie.cpp"
LIB = HERE / "libcpu_rpie.so"

_lib = None


def build(verbose=False):
    cmd = ["g++", "-O3", "-fopenmp", "-shared", "-fPIC", "-o", str(LIB), str(SRC)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def library():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = ctypes.CDLL(str(LIB))
        f = lib.cpu_rpie_batch
        P = ctypes.c_void_p
        f.argtypes = [ctypes.c_int] * 4 + [P] * 6 + [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                     ctypes.c_int] + [ctypes.c_double] * 3 + [ctypes.c_int]
        f.restype = ctypes.c_int
        _lib = lib
    return _lib


class CpuRpie:
    """Far-field rPIE state for the C++ batch: counts `[nframe, n, n]` (detector order,
    negative = masked), corners `[nframe, 2]`, object `[1, nyo, nxo]`, probe `[nmode, n, n]`,
    far-field phase `quad` `[n, n]`; `threads` OpenMP threads (0: the runtime's default)."""

    def __init__(self, counts, corners, obj, probe, quad, threads=1):
        self.counts = np.ascontiguousarray(counts, dtype=np.float32)
        self.corners = np.ascontiguousarray(corners, dtype=np.int64)
        o = np.asarray(obj)
        if o.ndim == 3:
            if o.shape[0] != 1:
                raise ValueError("rPIE needs exactly one object mode")
            o = o[0]
        # Own copies: the batch updates them in place.
        self.obj = np.array(o, dtype=np.complex128, order="C", copy=True)
        self.probe = np.array(probe, dtype=np.complex128, order="C", copy=True)
        self.quad = np.ascontiguousarray(quad, dtype=np.complex128)
        self.threads = int(threads)
        self.cycle = 0

    def batch(self, frames, obj_step=1.0, probe_step=1.0, obj_alpha=0.1, probe_alpha=0.1):
        fr = np.ascontiguousarray(frames, dtype=np.int64)
        nmode, n = self.probe.shape[0], self.probe.shape[-1]
        nyo, nxo = self.obj.shape
        ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        rc = library().cpu_rpie_batch(
            n, nmode, nyo, nxo, ptr(self.obj), ptr(self.probe), ptr(self.quad), ptr(self.counts),
            ptr(self.corners), ptr(fr), len(fr), int(obj_step is not None),
            float(obj_step or 0.0), int(probe_step is not None), float(probe_step or 0.0),
            float(obj_alpha), float(probe_alpha), self.threads)
        if rc != 0:
            raise ValueError("cpu_rpie_batch refused its arguments")

    def rpie(self, nepoch, obj_step=1.0, probe_step=1.0, obj_alpha=0.1, probe_alpha=0.1,
             batch_size=100, seed=7):
        """`nepoch` epochs over SplitMix64 batches (ptyrs/src/engine/batch.rs:14-38)."""
        for _ in range(nepoch):
            for frames in random_batches(len(self.counts), batch_size, seed, self.cycle):
                self.batch(frames, obj_step, probe_step, obj_alpha, probe_alpha)
            self.cycle += 1
