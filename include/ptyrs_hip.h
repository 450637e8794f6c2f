/* ptyrs_hip: the per-position ptychographic reconstruction step on an AMD Instinct MI355X
 * (gfx950), behind a C ABI.
 *
 * It replaces the device layer of ptyrs's GPU engine: `crate::gpu` (checked kernel launches,
 * ptyrs/src/gpu/mod.rs:1-492), `crate::fft` (cuFFT plans, ptyrs/src/fft/mod.rs:114+) and the
 * cuda-core/cuda-async runtime under `ptyrs::engine::GpuPtycho`
 * (ptyrs/src/engine/gpu/mod.rs:597-912). Each entry point below names the reference interface
 * it stands for. The engine owns every device buffer; one engine is driven from one thread.
 *
 * Arrays are row-major. Complex arrays are interleaved (re, im) doubles, as numpy complex128.
 * Counts are [nframe, ny, nx] in detector order (zero frequency at [ny/2, nx/2]).
 *
 * Status codes: PTYRS_OK; PTYRS_EINVAL (bad input; nothing changed, as
 * EngineError::Core/Invalid); PTYRS_EDEVICE (a device call failed part-way; the engine is then
 * aborted, ptyrs/src/engine/gpu/gate.rs:14-71); PTYRS_EABORTED (refused: an earlier call
 * failed part-way; downloads still run, as EngineError::Aborted, ptyrs/src/engine/mod.rs:81-85).
 * ptyrs_last_error() gives the message of the last failure on this thread.
 */
#ifndef PTYRS_HIP_H
#define PTYRS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PTYRS_OK = 0,
  PTYRS_EINVAL = 1,
  PTYRS_EDEVICE = 2,
  PTYRS_EABORTED = 3,
};

/* Exit transform kinds (ptyrs/src/engine/propagate.rs:301-366). */
enum { PTYRS_FOURIER = 0, PTYRS_ANGULAR_SPECTRUM = 1, PTYRS_FRESNEL = 2 };

/* Element types of the engine's arithmetic. */
enum { PTYRS_C64 = 0, PTYRS_C128 = 1 };

/* Sampling, how position offsets enter the exit waves (ptyrs/src/engine/sampling.rs:32-36). */
enum { PTYRS_PATCH = 0, PTYRS_FULL_OBJECT = 1 };

/* Algorithms of LLK records (ptyrs/src/engine/ptycho.rs:22-31). */
enum { PTYRS_AP = 0, PTYRS_DM = 1, PTYRS_RAAR = 2, PTYRS_RPIE = 3, PTYRS_ML = 4, PTYRS_BH = 5,
       PTYRS_LSQML = 6 };

/* A far-field scan, as ptyrs::base::Scan (ptyrs/src/base/scan.rs, ptyrs-py/src/lib.rs:433-463).
 * Exactly one of counts_u16 (65535 = masked, ptyrs/src/base/counts.rs:3-7) and counts_f32
 * (a negative value is masked; values above 65534 are kept, which the u16 path refuses) is
 * non-null. y, x are the sample positions in metres. corners, if non-null, gives the view
 * corners [nframe][2] (oy, ox) directly, an address table as ptyrs::base::AddrTable holds it
 * (ptyrs/src/base/addr.rs:20-81), and y, x are then not read. */
typedef struct {
  size_t nframe, ny, nx;
  const uint16_t* counts_u16;
  const float* counts_f32;
  const double* y;
  const double* x;
  double wavelength, distance, detector_pixel;
  const int64_t* corners;
} ptyrs_scan;

/* The exit-wave model (ptyrs::engine::Optics, ptyrs/src/engine/propagate.rs:24-186):
 * free_space = 0 takes the far field of the scan geometry (Optics::far_field); free_space = 1
 * takes pty-chi's planar model of one Wavefield (Optics::free_space) of wavelength, object
 * pixel [pixel_y, pixel_x] and signed distance (INFINITY for the far field).
 * ngap > 0 makes it multislice (Optics::with_slices, ptyrs/src/engine/propagate.rs:84-98):
 * gaps[4·s ..] = {wavelength, pixel_y, pixel_x, distance} of the angular-spectrum gap from
 * slice s to s + 1; the object then holds ngap + 1 slices. */
typedef struct {
  int free_space;
  double wavelength, pixel_y, pixel_x, distance;
  size_t ngap;
  const double* gaps;
} ptyrs_optics;

/* GpuOptions (ptyrs/src/engine/gpu/mod.rs:61-108) plus the element type. stack_size: frames
 * per stack for AP, DM, RAAR, ML and the LLK pass (0 lets the engine choose, the reference's
 * None); a stack, and an rPIE or LSQML batch, of more than 65535 views is refused. graph:
 * replay captured epochs (same bits either way). llk_slots >= 1: LLK records the device holds;
 * the host waits for them only when that many are held or at the end of a call. fused: this
 * engine's passes are always the fused ones, so 0 and 1 run the same kernels (the reference's
 * M1 and M2 launch paths agree to rounding). psi_f16: DM's and RAAR's carried exit waves (and
 * RAAR's iterate) held as one pair of IEEE halves per value, times a power of two fixed from
 * the counts (ptyrs/src/gpu/psi.rs:40-50), half the bytes of complex64; needs fused and
 * dtype c64, else PTYRS_EINVAL. */
typedef struct {
  size_t stack_size;
  int graph;
  size_t llk_slots;
  int fused;
  int psi_f16;
  int dtype;
} ptyrs_gpu_options;

/* RpieParams (ptyrs/src/engine/ptycho.rs:145-186); a step with has_* = 0 leaves that array
 * fixed. Defaults: steps 1.0, alphas 0.1, batch_size 100, calc_llk 0, seed 7. */
typedef struct {
  int has_obj_step;
  double obj_step;
  int has_probe_step;
  double probe_step;
  double obj_alpha, probe_alpha;
  size_t batch_size;
  size_t calc_llk;
  uint64_t seed;
} ptyrs_rpie_params;

/* ApParams, DmParams, RaarParams (ptyrs/src/engine/ptycho.rs:48-121); an inertia with has_* = 0
 * leaves that array fixed. Defaults: AP obj_inertia 0.01, no probe update; DM and RAAR
 * inertias 0.01 / 0.001, alpha 0.02, beta 0.9, loop_obj_probe 1; calc_llk 0;
 * zero_phase_ramp 1. */
typedef struct {
  int has_obj_inertia;
  double obj_inertia;
  int has_probe_inertia;
  double probe_inertia;
  size_t calc_llk;
  int zero_phase_ramp;
} ptyrs_ap_params;

typedef struct {
  int has_obj_inertia;
  double obj_inertia;
  int has_probe_inertia;
  double probe_inertia;
  double alpha;
  size_t loop_obj_probe;
  size_t calc_llk;
  int zero_phase_ramp;
} ptyrs_dm_params;

typedef struct {
  int has_obj_inertia;
  double obj_inertia;
  int has_probe_inertia;
  double probe_inertia;
  double beta;
  size_t loop_obj_probe;
  size_t calc_llk;
  int zero_phase_ramp;
} ptyrs_raar_params;

/* MlParams (ptyrs/src/engine/ptycho.rs:123-143). Defaults: update_object 1, update_probe 0,
 * calc_llk 0, zero_phase_ramp 1. */
typedef struct {
  int update_object;
  int update_probe;
  size_t calc_llk;
  int zero_phase_ramp;
} ptyrs_ml_params;

/* LsqmlParams (ptyrs/src/engine/lsqml.rs:34-53) and Cost (ptyrs/src/engine/features.rs). */
enum { PTYRS_AMPLITUDE = 0, PTYRS_POISSON = 1 };
typedef struct {
  size_t batch_size;
  int cost;
  uint64_t seed;
} ptyrs_lsqml_params;

/* One frame's LSQML step sizes of the last epoch (LsqmlStep, ptyrs/src/engine/lsqml.rs:55-66),
 * one object slice: the batch index within the epoch, the frame, a_far, a_obj and a_probe. */
typedef struct {
  size_t batch;
  size_t frame;
  double a_far, a_obj, a_probe;
} ptyrs_lsqml_step;

/* One LLK record (LlkRecord, ptyrs/src/engine/ptycho.rs:33-41), normalised by nb_obs
 * (Llk::normalised, ptyrs/src/engine/ops.rs:268-276). */
typedef struct {
  size_t cycle;
  int algorithm;
  double poisson, gaussian, euclidian, nb_photons_calc;
} ptyrs_llk_record;

typedef struct ptyrs_engine ptyrs_engine;

/* GpuPtycho::with_slices (ptyrs/src/engine/gpu/mod.rs:619-760): object [nslice, nobj, nyo, nxo]
 * (nslice = optics->ngap + 1, else 1) and probe [nprobe, ny, nx] complex128; positions go
 * through PyNX preprocessing and become view corners. optics may be NULL (far field of the
 * scan). device is the HIP device ordinal. More than one slice runs LSQML and the LLK pass
 * only, of one object mode (the others fail with PTYRS_EINVAL).
 * An angular-spectrum exit whose transfer function drops an evanescent band is not unitary:
 * rPIE, AP, DM, RAAR and ML then fail with PTYRS_EINVAL, and LSQML and the LLK pass run
 * (ptyrs/src/engine/gpu/device.rs:3087-3105). */
int ptyrs_engine_create(const ptyrs_scan* scan, const ptyrs_optics* optics, const double* obj,
                        size_t nobj, size_t nyo, size_t nxo, const double* probe, size_t nprobe,
                        const ptyrs_gpu_options* options, int device, ptyrs_engine** out);

/* GpuPtycho::rpie (ptyrs/src/engine/gpu/rpie.rs:31-66): nepoch epochs of rPIE. */
int ptyrs_engine_rpie(ptyrs_engine* e, const ptyrs_rpie_params* params, size_t nepoch);

/* rPIE on several independent scans at once: `nepoch` epochs of every engine (engines[i] with
 * params[i]), each as ptyrs_engine_rpie would run it (same results, same LLK records), the
 * engines' epochs enqueued round-robin, each a graph replay on its engine's own stream, so that
 * scans of uneven sizes share the device instead of running one after another. Returns when
 * all are done. Every engine's parameters are checked first: PTYRS_EINVAL changes none. A
 * device failure aborts every engine of the group. No counterpart in the reference, which runs
 * one scan per GpuPtycho call (ptyrs/src/engine/gpu/rpie.rs:31-62). */
int ptyrs_group_rpie(ptyrs_engine* const* engines, size_t count,
                     const ptyrs_rpie_params* params, size_t nepoch);
/* The same with every engine's epoch captured in one graph, a branch per engine, launched once
 * per epoch (same results). When any engine of the group was built with graph = 0 or has
 * profiling on, the group runs as ptyrs_group_rpie instead (each engine eagerly or from its
 * own graphs). No counterpart in the reference. */
int ptyrs_group_rpie_graph(ptyrs_engine* const* engines, size_t count,
                           const ptyrs_rpie_params* params, size_t nepoch);

/* GpuPtycho::ap, ::dm, ::raar (ptyrs/src/engine/gpu/ap.rs:28, dm.rs:32, raar.rs:17): ncycle
 * cycles of PyNX AP, DM or RAAR, the exit waves of DM and RAAR carried across its cycles. */
int ptyrs_engine_ap(ptyrs_engine* e, const ptyrs_ap_params* params, size_t ncycle);
int ptyrs_engine_dm(ptyrs_engine* e, const ptyrs_dm_params* params, size_t ncycle);
int ptyrs_engine_raar(ptyrs_engine* e, const ptyrs_raar_params* params, size_t ncycle);

/* GpuPtycho::ml (ptyrs/src/engine/gpu/ml.rs:38-56): PyNX ScaleObjProbe(absolute=True), then
 * ncycle cycles of Poisson conjugate gradient (Polak–Ribière β, one Newton step γ). */
int ptyrs_engine_ml(ptyrs_engine* e, const ptyrs_ml_params* params, size_t ncycle);

/* GpuPtycho::lsqml (ptyrs/src/engine/gpu/lsqml.rs:75-120) without a probe variation or
 * features: nepoch epochs from a fresh LSQML, an LLK record every calc_llk epochs and after
 * the last. */
int ptyrs_engine_lsqml(ptyrs_engine* e, const ptyrs_lsqml_params* params, size_t nepoch,
                       size_t calc_llk);

/* ProbeVariation (ptyrs/src/engine/opr.rs:19-90): k >= 1 eigenmodes of probe mode 0
 * ([k][ny][nx] complex128, interleaved re/im), their weights for every frame ([nframe][k]) and,
 * unless intensity is NULL, the main mode's weight c_f of every frame ([nframe]); frame f sees
 * c_f·P_0 + Σ_i w_fi·U_i as its probe mode 0. refine_weights / refine_intensity: refined by each
 * step, else held (the eigenmodes move either way). */
typedef struct {
  size_t k;
  double* eigenmodes;
  double* weights;
  double* intensity;
  int refine_weights;
  int refine_intensity;
} ptyrs_opr;

/* GpuPtycho::lsqml with a ProbeVariation (ptyrs/src/engine/gpu/lsqml.rs:75-120): LSQML epochs
 * that model pty-chi's orthogonal probe relaxation of probe mode 0; with nepoch >= 1 the
 * eigenmodes, weights and intensity of `var` are overwritten with the values the epochs leave.
 * Runs with one slice or several, with or without position offsets (refined or held) under
 * PATCH or FullObject sampling. */
int ptyrs_engine_lsqml_opr(ptyrs_engine* e, const ptyrs_lsqml_params* params, size_t nepoch,
                           size_t calc_llk, ptyrs_opr* var);

/* BhParams (ptyrs/src/engine/bh.rs:28-58): frames per batch (capped at the scan's; one batch of
 * every frame runs conjugate gradient across the epochs of a call, smaller batches steepest
 * descent), probe_rho (the probe's gradient is scaled by probe_rho²·k, k the exit's intensity
 * scale: ny·nx in the far field), position_rho (refined position offsets' gradient is scaled by
 * position_rho², their step is 0.3·α·D_d), seed of every epoch's random_batches. */
typedef struct {
  size_t batch_size;
  double probe_rho;
  double position_rho;
  uint64_t seed;
} ptyrs_bh_params;

/* One BH batch's step of the last epoch (BhStep, ptyrs/src/engine/bh.rs:60-72): its `count`
 * frames are entries first .. first + count of ptyrs_engine_bh_frames' order; β (0 in steepest
 * descent and the first epoch), α (after the factor count/batch_size), top = −⟨g, D⟩ and
 * bottom = H(D, D). */
typedef struct {
  size_t first, count;
  double beta, alpha, top, bottom;
} ptyrs_bh_step;

/* GpuPtycho::bh (ptyrs/src/engine/gpu/bh.rs:33-68): nepoch epochs of pty-chi's BH on the
 * amplitude cost from a fresh BH (epoch indices 0.. of the call), an LLK record every calc_llk
 * epochs and after the last. One object mode; one slice with or without position offsets
 * (refined or held) under PATCH or FullObject sampling, or several slices (the object variable
 * every slice) without offsets or with FullObject offsets (refined or held); PATCH offsets
 * through several slices give PTYRS_EINVAL. */
int ptyrs_engine_bh(ptyrs_engine* e, const ptyrs_bh_params* params, size_t nepoch,
                    size_t calc_llk);

/* GpuPtycho::bh_steps (ptyrs/src/engine/gpu/bh.rs:70-72): the last BH epoch's steps, *n of them
 * (none before a BH call); at most cap are copied to out (which may be NULL). */
int ptyrs_engine_bh_steps(ptyrs_engine* e, ptyrs_bh_step* out, size_t cap, size_t* n);
/* The last BH epoch's frames, batch after batch (len = nframe after a BH call, else 0). */
int ptyrs_engine_bh_frames(ptyrs_engine* e, size_t* out, size_t len);

/* GpuPtycho::lsqml_steps (ptyrs/src/engine/gpu/lsqml.rs:122-127): the last LSQML epoch's steps,
 * one per frame in batch order (a_obj of slice 0); copies up to cap and sets *n to the number
 * held. */
int ptyrs_engine_lsqml_steps(ptyrs_engine* e, ptyrs_lsqml_step* out, size_t cap, size_t* n);

/* LsqmlStep::a_obj of every slice (ptyrs/src/engine/lsqml.rs:55-66): out[i·nslice + s] is a_obj
 * of slice s at step i of ptyrs_engine_lsqml_steps; len must be n·nslice. */
int ptyrs_engine_lsqml_slice_steps(ptyrs_engine* e, double* out, size_t len);

/* Launch shapes that change results only by rounding, no counterpart in the reference (whose
 * GpuOptions::stack_size and StackRule cover the first): stack_size frames per stack of AP, DM,
 * RAAR, ML and the LLK pass (0: the engine's default), and the workgroups per launch rPIE's
 * (pass_c_blocks) and LSQML's (chi_blocks) pass C cut a batch into chunks for (0: the
 * defaults, 768 and 1024; at most 8192). For ptyrs_amd.tune, which times them per size class. */
typedef struct {
  size_t stack_size;
  int pass_c_blocks;
  int chi_blocks;
} ptyrs_launch;
int ptyrs_engine_set_launch(ptyrs_engine* e, const ptyrs_launch* launch);
int ptyrs_engine_launch(ptyrs_engine* e, ptyrs_launch* out);

/* GpuPtycho::set_features with Features::position only (ptyrs/src/engine/gpu/mod.rs:846-857,
 * ptyrs/src/engine/features.rs:304-330): offsets [nframe][2] (dy, dx) in object pixels,
 * len = 2·nframe, refined by every rPIE batch, LSQML step or AP, DM, RAAR or ML cycle when
 * refine != 0, else held; offsets = NULL stops modelling them. Offsets are kept in FP64 on the
 * device (BH refines them too). Under PATCH sampling every algorithm and the LLK pass run with
 * offsets (AP, DM and RAAR through Shifted::psi_to_obj_probe,
 * ptyrs/src/engine/features.rs:653-720: the probe's first conjugate-gradient step, DM and RAAR
 * with complex waves only); under FullObject sampling every algorithm (AP, DM and RAAR through
 * the FullObject psi_to_obj_probe, ptyrs/src/engine/features.rs:653-793: the probe's first
 * conjugate-gradient step through the sampled windows, the object's preconditioned CG solve),
 * the LLK pass and the model intensities, with one object mode (LSQML, the LLK pass and the
 * model intensities also through several slices, every slice read through its sampled
 * windows). */
int ptyrs_engine_set_position(ptyrs_engine* e, const double* offsets, size_t len, int refine);

/* GpuPtycho::features' position (ptyrs/src/engine/gpu/mod.rs:823-825): *state = 0 (none),
 * 1 (held) or 2 (refined); the offsets are copied to out (len = 2·nframe) if out is non-null. */
int ptyrs_engine_position(ptyrs_engine* e, double* out, size_t len, int* state);

/* GpuPtycho::sampling / set_sampling (ptyrs/src/engine/gpu/mod.rs:827-844): PTYRS_PATCH or
 * PTYRS_FULL_OBJECT, Sampling::auto of the optics, object and frame until set. Under
 * PTYRS_FULL_OBJECT each view's window is Fourier-interpolated from the whole object plane at
 * its corner plus its offset (object sides up to 4096 pixels in complex64, 2048 in
 * complex128); the algorithms that do not run it fail with PTYRS_EINVAL while offsets are set. */
int ptyrs_engine_sampling(ptyrs_engine* e, int* sampling);
int ptyrs_engine_set_sampling(ptyrs_engine* e, int sampling);

/* Multislice wave checkpointing for the adjoint (LSQML and ptyrs_engine_grad). The reference
 * keeps the wave entering every slice of a stack (SliceWaves::new,
 * ptyrs/src/engine/multislice.rs:86-171); every = 0 or 1 does the same. every = k >= 2 keeps
 * the waves entering slices k, 2k, … and recomputes the others from those one segment at a
 * time on the way back, so a stack holds (nslice − 1)/k + k − 1 wave sets instead of
 * nslice − 1, for one more forward pass through the non-checkpoint slices; results are
 * bit-identical. At most 64. */
int ptyrs_engine_set_checkpoint(ptyrs_engine* e, size_t every);
/* The interval set, and the device bytes the kept incident waves hold now (0 before the first
 * multislice call after a change). */
int ptyrs_engine_checkpoint(ptyrs_engine* e, size_t* every, size_t* wave_bytes);

/* The cost of the current state over every frame and its gradients, the adjoint of the
 * exit-wave model (object window -> probe multiply -> slices -> exit transform -> cost), for a
 * torch.autograd.Function: *value = the Poisson LLK sum Σ(icalc − iobs + iobs·ln(iobs/icalc))
 * (cost PTYRS_POISSON; its gradients are those of the Poisson NLL) or Σ(√icalc − √iobs)²
 * (PTYRS_AMPLITUDE) over unmasked pixels; obj_grad = 2·∂L/∂O* of every slice (len as the
 * object's), probe_grad = 2·∂L/∂P* (len as the probe's), complex128 interleaved — for a real L
 * of complex z, 2·∂L/∂z* is torch's gradient. Follows Ptycho::obj_grad / probe_grad and the
 * multislice adjoint (ptyrs/src/engine/ptycho.rs:492-561, multislice.rs:266-442,
 * features.rs:310-330). No position offsets. */
int ptyrs_engine_grad(ptyrs_engine* e, int cost, double* value, double* obj_grad,
                      size_t obj_len, double* probe_grad, size_t probe_len);

/* The model's detector intensities Σ_m |U ψ_m|² of every frame at the current state, detector
 * order, len = nframe·ny·nx: what the counts are compared with, pty-chi's forward-model output
 * (ptyrs/tests/engine_lsqml_multislice.rs:146-182; the reference keeps them on the device). */
int ptyrs_engine_intensities(ptyrs_engine* e, double* out, size_t len);

/* Live ingestion (LiveAp::push and epoch, ptyrs/src/live/recon.rs:184-266; the frames' ring,
 * ptyrs/src/live/ring.rs:116-175): the engine's scan is the planned one (its positions or
 * corners fix every frame's view up front); frames land in scan order and every algorithm,
 * the LLK and the model intensities run over the frames landed so far (the reference's live
 * AP is ptyrs_engine_ap with update_probe = 0 on the landed frames).
 * ptyrs_engine_live_reset leaves no frame landed (the engine holds every planned frame's slot,
 * so landing is a copy, never a reallocation). ptyrs_engine_live_push lands the next nframes
 * planned frames: counts [nframes][ny][nx] in detector order, exactly one of counts_u16
 * (65535 masked) and counts_f32 (negative masked); PTYRS_EINVAL past the planned frames.
 * ptyrs_engine_live_frames reports the landed and planned counts (either pointer may be
 * null). An engine starts with every planned frame landed. */
int ptyrs_engine_live_reset(ptyrs_engine* e);
int ptyrs_engine_live_push(ptyrs_engine* e, const uint16_t* counts_u16, const float* counts_f32,
                           size_t nframes);
int ptyrs_engine_live_frames(ptyrs_engine* e, size_t* landed, size_t* planned);

/* GpuPtycho::llk (ptyrs/src/engine/gpu/mod.rs:853-868): normalised LLK of the current state. */
int ptyrs_engine_llk(ptyrs_engine* e, ptyrs_llk_record* out);
/* The LLK of the current state, as ptyrs_engine_llk gives it (the same bits), for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). No counterpart in the reference. The call allocates nothing: its
 * exit-wave stacks, per-workgroup sums and result live in one static device array of 64 MiB
 * per GPU, held until the process ends. A scan
 * whose per-frame sums (32 B per frame and pass-B workgroup) and one frame's exit waves do not
 * fit returns PTYRS_EINVAL, as do several slices, position offsets and profiling. Every call
 * on a GPU shares that GPU's array: calls that overlap on one GPU (engines on several host
 * threads) race and must be serialised by the caller. */
int ptyrs_engine_llk_noalloc(ptyrs_engine* e, ptyrs_llk_record* out);

/* GpuPtycho::objects / probe (ptyrs/src/engine/gpu/mod.rs:762-815): complex128 downloads of
 * len doubles (2·nslice·nobj·nyo·nxo, every slice; 2·nprobe·ny·nx). Allowed on an aborted
 * engine. */
int ptyrs_engine_object(ptyrs_engine* e, double* out, size_t len);
int ptyrs_engine_probe(ptyrs_engine* e, double* out, size_t len);

/* GpuPtycho::set_object (ptyrs/src/engine/gpu/mod.rs:816-840) and its probe counterpart. */
int ptyrs_engine_set_object(ptyrs_engine* e, const double* obj, size_t len);
int ptyrs_engine_set_probe(ptyrs_engine* e, const double* probe, size_t len);

/* GpuPtycho::history (ptyrs/src/engine/gpu/mod.rs:842-846): copies up to cap records and sets
 * *n to the number held. */
int ptyrs_engine_history(ptyrs_engine* e, ptyrs_llk_record* out, size_t cap, size_t* n);

/* GpuPtycho::cycle (ptyrs/src/engine/gpu/mod.rs:848-851). */
int ptyrs_engine_cycle(ptyrs_engine* e, size_t* cycle);

/* The view corners [nframe, 2] (oy, ox) the engine built (AddrTable, ptyrs/src/base/addr.rs). */
int ptyrs_engine_corners(ptyrs_engine* e, int64_t* out, size_t len);

/* GpuPtycho::synchronize (ptyrs/src/engine/gpu/mod.rs:870-881). */
int ptyrs_engine_synchronize(ptyrs_engine* e);

/* Kernel timing with HIP events on the engine's stream, for measurement: enable = 1 starts
 * recording every launch, 0 stops. ptyrs_engine_kernel_stats reports, for kernel name `name`
 * (as listed by ptyrs_kernel_names), the launches recorded and their summed duration in ms.
 * No reference counterpart (the reference times its kernels with external profilers). */
int ptyrs_engine_profile(ptyrs_engine* e, int enable);
int ptyrs_engine_kernel_stats(ptyrs_engine* e, const char* name, size_t* launches, double* total_ms);
const char* ptyrs_kernel_names(void);

void ptyrs_engine_destroy(ptyrs_engine* e);

/* Host-side scan preparation, the steps ptyrs_engine_create runs before it touches a device;
 * none of these calls a device.
 *
 * Positions::pynx_preprocess along one axis (ptyrs/src/base/positions.rs:51-100): v[n] in
 * metres, pixel the object pixel along that axis, out[n]; PTYRS_EINVAL where a position or a
 * preprocessed one is not finite with |pos| <= 1e10. */
int ptyrs_pynx_preprocess(const double* v, size_t n, double pixel, double* out);
/* view_corner (ptyrs/src/base/addr.rs:118-127): the window's corner along one axis,
 * (obj − frame)/2 + shift rounded half to even, or -1 when the window leaves the object. */
int ptyrs_view_corner(size_t obj, size_t frame, double shift, int64_t* corner);
/* required_obj_shape (ptyrs/src/base/addr.rs:129-149) of positions y, x (metres, as given;
 * the reference applies it to preprocessed positions): out[0], out[1] = ny, nx. */
int ptyrs_required_obj_shape(const double* y, const double* x, size_t n, double pixel_y,
                             double pixel_x, size_t frame_y, size_t frame_x, size_t margin,
                             size_t* out);
/* AddrTable::new over the preprocessed positions (ptyrs/src/base/addr.rs:35-81), as
 * ptyrs_engine_create builds it: the view corners [nframe][2] (oy, ox) of `scan` under
 * `optics` (NULL: far field) in an object of nyo × nxo; len = 2·nframe. */
int ptyrs_scan_corners(const ptyrs_scan* scan, const ptyrs_optics* optics, size_t nyo,
                       size_t nxo, int64_t* out, size_t len);
/* Observed::from_counts (ptyrs/src/engine/obs.rs:13-35): the counts as the device reads them,
 * [nframe][ny][nx] float in the exit transform's DFT index order (for the far field the
 * reference's fftshifted layout), -1 where masked; nb_obs and the unmasked total. out may be
 * NULL (then len is not read). */
int ptyrs_scan_observed(const ptyrs_scan* scan, const ptyrs_optics* optics, float* out,
                        size_t len, size_t* nb_obs, double* total);
/* quad_phase (ptyrs/src/engine/ops.rs:11-33): the far-field quadratic phase of a geometry,
 * [ny][nx] complex interleaved in the natural (centred) layout, len = 2·ny·nx. */
int ptyrs_quad_phase(double wavelength, double distance, double detector_pixel, size_t ny,
                     size_t nx, double* out, size_t len);

/* Fault injection for tests of the abort gate, process-wide, as the reference's test-only
 * fault::arm (ptyrs/src/fault.rs:62-114): the exit `exit` fails at its hit after `pass` others,
 * once, as a device failure (PTYRS_EDEVICE) raised before anything is enqueued, so every
 * kernel already launched runs to its end. PTYRS_FAULT_LAUNCH is each kernel launch of an
 * engine call, PTYRS_FAULT_CAPTURE each graph capture before it begins. exit = -1 disarms.
 * ptyrs_test_fault_fired tells whether the armed exit has failed. */
enum { PTYRS_FAULT_LAUNCH = 0, PTYRS_FAULT_CAPTURE = 1 };
int ptyrs_test_arm_fault(int exit, size_t pass);
int ptyrs_test_fault_fired(int* fired);

/* Message of the last failed call on this thread; returns its length. */
size_t ptyrs_last_error(char* buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* PTYRS_HIP_H */
