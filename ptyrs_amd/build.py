"""Builds libptyrs_hip.so for gfx950 in-tree: one object per supported frame side (compiled in
parallel) plus the C ABI, linked with hipcc.

    python -m ptyrs_amd.build [--jobs J]
"""

import argparse
import concurrent.futures as cf
import os
import pathlib
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "_build"
LIB = HERE / "libptyrs_hip.so"
SIDES = (16, 24, 32, 48, 64, 96, 128, 256, 512)
ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # Packed FP32 instructions off (v_pk_*_f32 have given wrong results on these boxes).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _compile(src, obj, defines=()):
    cmd = [HIPCC, *FLAGS, *defines, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name} {' '.join(defines)}:\n{r.stderr[-4000:]}")
    return obj


def build(jobs=None, verbose=False, lib=LIB, out=OUT, defines=(), sides=SIDES):
    """Compiles every unit (with extra `defines`) into `out` and links `lib`."""
    out = pathlib.Path(out)
    lib = pathlib.Path(lib)
    out.mkdir(parents=True, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    units = [(CSRC / "capi.hip", out / "capi.o", tuple(defines))]
    units += [(CSRC / "engine_inst.hip", out / f"engine_{n}.o", (f"-DPTYRS_N={n}", *defines))
              for n in sides]
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda u: _compile(*u), units))
    tmp = lib.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, lib)
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    build(a.jobs, verbose=True)
    sys.exit(0)
