"""FP64 CPU oracle: a numpy restatement of ptyrs's CPU engine for the per-position
reconstruction step.

TEST INFRASTRUCTURE ONLY. Nothing in the product (`ptyrs_amd`, `include/`) imports, calls,
links or executes anything under `oracle/`. Only `tests/`, `__graft_entry__.smoke()` and
`bench.py`'s `cpu_baseline` leg use it, and there only as the checker (or, for the CPU
baseline, as the thing timed beside the GPU), never as a code path of the HIP engine.

Every function cites the reference file:line it follows. Layout conventions are the
reference's (`ptyrs/src/engine/mod.rs:6-19`): exit waves `[nview, ny, nx]` stored FFT-shifted,
view `k = frame * nmode + mode`, unitary 2D FFT, every object/probe sum in view order.

What pins each module (the test files named run on the CPU):
- base.py — tests/test_base_pins.py restates the reference's unit-test values:
  `view_corner` half-even ties and edge rejection, `AddrTable` order, mode mapping and
  rejection, `required_obj_shape` (`base/addr.rs:151-214`); the three `pynx_preprocess`
  cases, including the shared sub-pixel offset (`base/positions.rs:102-143`); `observed`'s
  shift, masking and totals, and the u16 MASKED branch (`engine/obs.rs:73-107`,
  `base/counts.rs:140-198`). `random_batches` is pinned by the `rpie_near_*` runs below
  (their batch order decides every traced value).
- optics.py — tests/test_oracle_optics.py: `propagator_{select,cases,fields}_f64.txt` at
  1e-14, `fft2` against a direct DFT, unitarity and adjoint dot tests; `quad_phase` by the
  `engine/ops.rs:409-419` values in tests/test_base_pins.py.
- ptycho.py — tests/test_base_pins.py (`engine/ops.rs:383-502` hand values for
  `apply_amplitude` and `llk`, adjoint dot tests of the back-projections, the consistent-pair
  invariant); tests/test_oracle_rpie.py (`rpie_near_*`, 9 runs × 30 epochs at 1e-12);
  tests/test_oracle_pynx.py (AP, DM, RAAR, ML against `pynx_near_*`);
  tests/test_oracle_scenario.py (`dm_pynx`, `raar_pynx`, `rpie_ptychi`, `ml_grad`; rPIE and
  ML under refined PATCH offsets: `rpie_pos_*` at 1e-12, `ml_pos_*` at 1e-13; rPIE under
  FULL_OBJECT: `rpie_full_object_{step,epochs,odd}_*` at 1e-12; ML, AP, DM and RAAR under
  FULL_OBJECT: zero offsets reduce them to the plain runs, ML's gradient against central
  differences, `solve_obj` descends its quadratic and, run long, reaches its minimiser (the
  quadratic's slope along random directions vanishes there)).
- lsqml.py — tests/test_oracle_lsqml.py (`lsqml_near_*`), tests/test_oracle_scenario.py
  (`lsqml_{gaussian,poisson}_*`, `lsqml_epochs_*`), tests/test_oracle_gradient.py
  (`cost_grads` against PyNX's ML gradient and five-point central differences).
- opr.py — tests/test_oracle_opr.py restates `ptyrs/tests/engine_opr.rs` on the seeded
  scenario: per-frame probe costs (`opr_values`, `lsqml_opr_values`, `lsqml_opr_int_values` at
  1e-13), weight gradients against the fixtures (1e-13) and central differences (1e-8), the OPR
  step (one eigenmode) and LSQML steps with three eigenmodes, with and without the main mode's
  weights (steps, changes, eigenmodes, weights at 1e-12/1e-13), a one-frame batch, held
  weights; LSQML with PATCH offsets and OPR (`lsqml_pos_opr_*`), and the far-field
  `lsqml_pos_*` step and three epochs; under FULL_OBJECT at zero offsets the OPR step
  reproduces the `lsqml_opr_*` and `lsqml_opr_int_*` fixtures at 1e-11.
- bh.py — tests/test_oracle_bh.py restates `ptyrs/tests/engine_bh.rs` and
  `engine_bh_near.rs`: two epochs on the seeded far-field scenario with one batch of all frames
  (conjugate gradient) and with batches of 16 (`bh_full_*`, `bh_mini_*`: β, α, ⟨g, D⟩, H(D, D)
  at 1e-13, object changes at 1e-13, probe at 5e-12), the four near-field cases with both batch
  sizes (`bh_near_*`); three slices with unequal gaps under the three exits (`bh_multislice_*`,
  steps at 2e-13, every slice's and the probe's change at 2e-12); refined offsets under PATCH (`bh_pos_full_*`, `bh_pos_mini_*`) and FullObject sampling
  (`bh_full_object_full_*`, `bh_full_object_mini_*`, pty-chi's) at 3e-12; the
  gradient against central differences of the cost, the Hessian against central differences of
  the gradient (with and without offsets; through three slices at held and refined
  FULL_OBJECT offsets, the offsets' part alone), its symmetry and bilinearity, zero offsets
  reducing to the unmoved batch.
- multislice.py — tests/test_oracle_multislice.py (`lsqml_multislice_*`: forward
  intensities, one step, 30 epochs; directions and tangent against central differences).
- positions.py — tests/test_oracle_positions.py (`lsqml_near_pos_*`,
  `lsqml_multislice_pos_*`; closed forms: a whole-pixel move is a roll, the MAD clip by hand);
  `position_step` through the `ml_pos_*` run; `derivative_sums` through the FULL_OBJECT runs.
- sampling.py — tests/test_oracle_sampling.py restates `engine/sampling.rs:263-436`: windows of
  a trigonometric polynomial are its interpolant and derivatives (past the plane's edge, odd
  and even sides) at 1e-13, `place` is the adjoint of `windows`, `diag` the brute-force
  diagonal of the normal operator, the zero-offset illumination the PATCH one, the AUTO rule;
  then the FULL_OBJECT fixtures: `rpie_full_object_*` (tests/test_oracle_scenario.py) and
  LSQML's `lsqml_near_pos_full_*` (cases 0 and 2) and `lsqml_multislice_pos_full_*` at 1e-12
  (tests/test_oracle_positions.py).
- scenario.py — tests/test_oracle_scenario.py: the ChaCha block function against RFC 8439's
  A.1 keystream (20 rounds) and the reference's seeded scenario, whose Poisson LLK sum
  equals `ml_grad_values_f64.txt`'s at 9e-16 (every count is the reference's); Poisson
  sampler moments.
- cpu.py / cpu_rpie.cpp (the CPU baseline) — tests/test_oracle_cpu.py: one batch at 1e-12
  and three epochs at 1e-11 against ptycho.py's rPIE, any thread count bit-identical.
"""
