"""FP64 CPU oracle: a numpy restatement of ptyrs's CPU engine for the per-position
reconstruction step.

TEST INFRASTRUCTURE ONLY. Nothing in the product (`ptyrs_amd`, `include/`) imports, calls,
links or executes anything under `oracle/`. Only `tests/`, `__graft_entry__.smoke()` and
`bench.py`'s `cpu_baseline` leg use it, and there only as the checker (or, for the CPU
baseline, as the thing timed beside the GPU), never as a code path of the HIP engine.

Every function cites the reference file:line it follows. Layout conventions are the
reference's (`ptyrs/src/engine/mod.rs:6-19`): exit waves `[nview, ny, nx]` stored FFT-shifted,
view `k = frame * nmode + mode`, unitary 2D FFT, every object/probe sum in view order.

Pinning (what checks each function; see tests/test_oracle_*.py):
- base.py: reference unit-test values (`base/addr.rs:155-214`, `base/positions.rs:114-143`,
  `engine/obs.rs:79-107`), random_batches against `refs/f64/random_batches.py`'s published
  algorithm plus permutation invariants.
- optics.py: `propagator_{select,cases,fields}_f64.txt` at 1e-14; FFT against a brute-force
  DFT; unitarity and adjoint dot tests.
- ptycho.py: ops hand values (`engine/ops.rs:461-502`), dot tests, the consistent-pair
  invariant; rPIE against `rpie_near_*_f64.txt` (LLK trace, object and probe change) at 1e-12.
"""
