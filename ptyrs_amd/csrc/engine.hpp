// The engine: device state of one scan and the algorithm loops that drive the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "host.hpp"
#include "kernels.hpp"
#include "bh.hpp"
#include "opr.hpp"
#include "sampling.hpp"

namespace ptyrs {

#define PTYRS_HIP(expr)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw DeviceError(std::string(#expr) + ": " + hipGetErrorString(e_));               \
  } while (0)

// Fault injection for tests of the abort gate (ptyrs_test_arm_fault): the armed exit fails
// at its hit after `pass` others, once, by throwing before anything is enqueued, so every
// kernel already launched runs to its end. Process-wide, as the reference's test hook
// (ptyrs/src/fault.rs:62-114); disarmed, an exit costs one relaxed atomic load.
struct FaultArm {
  std::atomic<bool> armed{false};
  std::mutex m;
  int exit = -1;
  size_t pass = 0;
  bool fired = false;
};
inline FaultArm& fault_arm() {
  static FaultArm a;
  return a;
}
inline void fault_at(int exit) {
  FaultArm& a = fault_arm();
  if (!a.armed.load(std::memory_order_relaxed)) return;
  std::lock_guard<std::mutex> g(a.m);
  if (a.exit != exit || a.fired) return;
  if (a.pass > 0) {
    --a.pass;
    return;
  }
  a.fired = true;
  throw DeviceError("injected fault at exit " + std::to_string(exit));
}

// Kernel names for timing; the index is the `kind` of a timed launch.
enum Kind { K_PASS_A = 0, K_PASS_B, K_PASS_C, K_BIN, K_OVERLAP, K_OBJ_UPDATE, K_PROBE, K_LLK,
            K_OTHER, K_OBJ_DIR, K_DOTS, K_HESS_A, K_HESS_B, K_COUNT };
inline const char* kind_names() {
  return "pass_a,pass_b,pass_c,bin_tiles,obj_overlap,obj_update,probe_update,llk,other,"
         "obj_dir,dots,hess_a,hess_b";
}

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  // Bumped on every (re)allocation: captured graphs hold device addresses.
  static inline std::atomic<uint64_t> generation{0};
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void ensure(size_t want) {
    if (want <= bytes) return;
    if (p) PTYRS_HIP(hipFree(p));
    p = nullptr;
    bytes = 0;
    PTYRS_HIP(hipMalloc(&p, want));
    bytes = want;
    ++generation;
  }
  // Frees the allocation (after the stream has drained); the next ensure allocates anew.
  void release() {
    if (p) PTYRS_HIP(hipFree(p));
    p = nullptr;
    bytes = 0;
    ++generation;
  }
  template <typename U>
  U* as() const { return static_cast<U*>(p); }
};

struct Record {
  size_t cycle;
  int algorithm;
  double llk[4];
};

// Everything of a scan the host prepares once (ptyrs/src/engine/gpu/mod.rs:635-760).
struct Prepared {
  size_t nframe = 0, ny = 0, nx = 0;
  std::vector<float> counts;  // DFT index order, negative = masked
  std::vector<int> corner;    // [nframe][2] oy, ox
  size_t nb_obs = 0;
  double total = 0;
  double max_frame_total = 0;  // max over frames of the unmasked counts' sum
  bool graph = true;           // GpuOptions::graph: replay captured epochs
  size_t stack_size = 0;       // GpuOptions::stack_size: frames per stack, 0 = the default
  size_t llk_slots = 8;        // GpuOptions::llk_slots: LLK records the device holds
  bool psi_f16 = false;        // GpuOptions::psi_f16: DM's and RAAR's carried waves as f16 pairs
  ExitModel exit;
  // Multislice: object slices and, per gap, its angular-spectrum transfer function.
  size_t nslice = 1;
  std::vector<std::vector<double>> gaps;
};

// AP, DM and RAAR in one form (ptyrs/src/engine/ptycho.rs:605-755).
struct ProjParams {
  int algorithm;  // PTYRS_AP, PTYRS_DM, PTYRS_RAAR
  bool upd_obj, upd_probe;
  double obj_inertia, probe_inertia;
  double mix;     // DM alpha, RAAR beta
  size_t loops, calc_llk;
  bool zero_phase_ramp;
};

// One LSQML slot's step sizes (LsqmlStep, ptyrs/src/engine/lsqml.rs:55-66), single slice.
struct StepRecord {
  int batch;
  int frame;
  double a_far, a_obj, a_probe;
  std::vector<double> a_objs;  // every slice's a_obj, slice 0's first
};

// One BH batch's step (BhStep, ptyrs/src/engine/bh.rs:60-72): its frames (`count` of them from
// `first` in the epoch's order), β, α, −⟨g, D⟩ and H(D, D).
struct BhRecord {
  size_t first, count;
  double beta, alpha, top, bottom;
};

// A captured epoch (GpuOptions::graph, ptyrs/src/engine/gpu/mod.rs:110-179): one graph per
// epoch kind, keyed by what its launches bake in (shapes, parameters, buffer generation).
struct EpochGraph {
  hipGraphExec_t exec = nullptr;
  std::vector<uint64_t> key;
  ~EpochGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
};

class EngineBase {
 public:
  virtual ~EngineBase() {
    if (stream_) {
      (void)hipStreamSynchronize(stream_);
      (void)hipStreamDestroy(stream_);
    }
  }
  virtual void rpie(const ptyrs_rpie_params& p, size_t nepoch) = 0;
  virtual void rpie_check(const ptyrs_rpie_params& p) = 0;
  virtual void rpie_begin(const ptyrs_rpie_params& p, size_t nepoch) = 0;
  virtual bool rpie_epoch() = 0;  // enqueues the next epoch; false once none is left
  virtual void rpie_end() = 0;
  virtual std::vector<uint64_t> rpie_epoch_order() = 0;
  virtual void rpie_epoch_body() = 0;
  virtual bool rpie_epoch_after() = 0;
  virtual bool rpie_pending() const = 0;  // an epoch of the call in progress is left
  hipStream_t stream() const { return stream_; }
  virtual void lsqml(const ptyrs_lsqml_params& p, size_t nepoch, size_t calc_llk) = 0;
  virtual void lsqml_opr(const ptyrs_lsqml_params& p, size_t nepoch, size_t calc_llk,
                         ptyrs_opr& var) = 0;
  virtual void bh(const ptyrs_bh_params& p, size_t nepoch, size_t calc_llk) = 0;
  // Launch shapes that change results only by rounding (the autotuner's knobs): frames per
  // stack of AP, DM, RAAR, ML and the LLK pass (0: the default), and the workgroups per launch
  // that rPIE's and LSQML's pass C cut a batch into chunks for (0: the default).
  virtual void set_launch(size_t stack_size, int pass_c_blocks, int chi_blocks) = 0;
  virtual void get_launch(size_t* stack_size, int* pass_c_blocks, int* chi_blocks) const = 0;
  virtual void project(const ProjParams& p, size_t ncycle) = 0;
  virtual void ml(const ptyrs_ml_params& p, size_t ncycle) = 0;
  virtual void llk(double out[4]) = 0;
  // The LLK with its exit-wave stack, per-workgroup sums and result in `scratch` (`bytes`
  // long, device memory the caller owns) instead of the engine's workspace: no allocation.
  virtual void llk_noalloc(void* scratch, size_t bytes, double out[4]) = 0;
  // Bytes of scratch llk_noalloc needs for a stack of `frames` frames.
  virtual size_t llk_scratch_bytes(size_t frames) const = 0;
  virtual void grad(int cost, double* value, double* obj_grad, double* probe_grad) = 0;
  virtual void intensities(double* out) = 0;
  virtual void get_object(double* out) = 0;
  virtual void get_probe(double* out) = 0;
  virtual void set_object(const double* in) = 0;
  virtual void set_probe(const double* in) = 0;
  // Features::position (ptyrs/src/engine/features.rs:304-330): offsets [nframe][2] in object
  // pixels, or null for none; refine or hold.
  virtual void set_position(const double* offsets, bool refine) = 0;
  virtual int get_position(double* out) = 0;  // 0 none, 1 held, 2 refined
  int sampling = 0;                            // PTYRS_PATCH or PTYRS_FULL_OBJECT
  // Multislice wave checkpointing: 0 or 1 keeps the wave entering every slice for the
  // adjoint; k >= 2 keeps every k-th and recomputes the rest one segment at a time.
  size_t checkpoint = 0;
  virtual void set_checkpoint(size_t every) = 0;
  virtual size_t wave_bytes() const = 0;  // device bytes holding the kept incident waves
  // Live ingestion (ptyrs/src/live/recon.rs:184-266, ring.rs): the scan's planned frames land
  // in order; every algorithm runs over the frames landed so far. live_reset leaves none
  // landed; live_push lands the next `nframes` frames, `counts` in the device's index order
  // (the counts' conversion of the scan), with their unmasked pixel count, total and largest
  // frame total.
  virtual void live_reset() = 0;
  virtual void live_push(const float* counts, size_t nframes, size_t nb_obs, double total,
                         double max_frame_total) = 0;
  size_t landed() const { return nframe_; }
  size_t planned() const { return corners.size() / 2; }
  bool counts_reverse() const { return reverse_counts_; }
  bool counts_dft_plane() const { return dft_plane_counts_; }

  void synchronize() { PTYRS_HIP(hipStreamSynchronize(stream_)); }

  size_t object_len() const { return 2 * nslice_ * nobj_ * nyo_ * nxo_; }
  size_t nslice() const { return nslice_; }
  size_t probe_len() const { return 2 * nprobe_ * ny_ * nx_; }
  size_t nprobe() const { return nprobe_; }
  size_t frame_len() const { return nframe_ * ny_ * nx_; }
  size_t frame_side() const { return ny_; }

  std::vector<Record> history;
  std::vector<StepRecord> lsqml_steps;  // the last LSQML epoch's
  std::vector<BhRecord> bh_steps;       // the last BH epoch's
  std::vector<int> bh_order;            // the last BH epoch's frames, batch after batch
  size_t cycle = 0;
  bool aborted = false;
  std::vector<int> corners;

  void profile(bool on) {
    profiling_ = on;
    if (on) {
      for (auto& s : stats_) s = {0, 0.0};
    }
  }
  int device() const { return device_; }
  void kernel_stats(int kind, size_t* launches, double* ms) {
    flush_timing();
    *launches = stats_[kind].first;
    *ms = stats_[kind].second;
  }

 protected:
  int device_ = 0;
  hipStream_t stream_ = nullptr;
  size_t nframe_ = 0, ny_ = 0, nx_ = 0, nprobe_ = 0, nobj_ = 0, nyo_ = 0, nxo_ = 0;
  size_t nslice_ = 1;
  size_t nb_obs_ = 0;
  bool reverse_counts_ = false, dft_plane_counts_ = true;  // the counts' index map

  bool profiling_ = false;
  bool graph_ = true;

 public:
  // Whether this engine's epochs run as graph replays (graph option on, profiling off).
  bool captures() const { return graph_ && !profiling_; }

 protected:
  EpochGraph rpie_graph_, lsqml_graph_, proj_graph_, ml_graph_;
  virtual uint64_t launch_key() const = 0;

  // Replays `g` if its key matches, else captures `body` (the launches of one epoch on the
  // engine's stream, no allocation, memset or host wait inside) into it and launches it.
  // Profiling runs `body` eagerly, its launches timed one by one.
  template <typename F>
  void epoch(EpochGraph& g, std::vector<uint64_t> key, F&& body) {
    if (!graph_ || profiling_) {
      body();
      return;
    }
    key.push_back(DevBuf::generation.load());
    key.push_back(uint64_t(launch_key()));
    key.push_back(uint64_t(nframe_));  // frames landed (live ingestion)
    if (g.exec == nullptr || g.key != key) {
      if (g.exec) {
        PTYRS_HIP(hipGraphExecDestroy(g.exec));
        g.exec = nullptr;
      }
      hipGraph_t graph = nullptr;
      fault_at(PTYRS_FAULT_CAPTURE);
      PTYRS_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
      try {
        body();
      } catch (...) {
        hipGraph_t dropped = nullptr;
        (void)hipStreamEndCapture(stream_, &dropped);
        if (dropped) (void)hipGraphDestroy(dropped);
        throw;
      }
      PTYRS_HIP(hipStreamEndCapture(stream_, &graph));
      const hipError_t inst = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (inst != hipSuccess) {
        g.exec = nullptr;
        throw DeviceError(std::string("hipGraphInstantiate: ") + hipGetErrorString(inst));
      }
      g.key = std::move(key);
    }
    PTYRS_HIP(hipGraphLaunch(g.exec, stream_));
  }
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending_;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pool_;
  std::pair<size_t, double> stats_[K_COUNT] = {};

  // Runs `launch` on the engine's stream, timed with events when profiling.
  template <typename F>
  void run(int kind, F&& launch) {
    fault_at(PTYRS_FAULT_LAUNCH);
    if (!profiling_) {
      launch();
      PTYRS_HIP(hipGetLastError());
      return;
    }
    std::pair<hipEvent_t, hipEvent_t> ev;
    if (!pool_.empty()) {
      ev = pool_.back();
      pool_.pop_back();
    } else {
      PTYRS_HIP(hipEventCreate(&ev.first));
      PTYRS_HIP(hipEventCreate(&ev.second));
    }
    PTYRS_HIP(hipEventRecord(ev.first, stream_));
    launch();
    PTYRS_HIP(hipGetLastError());
    PTYRS_HIP(hipEventRecord(ev.second, stream_));
    pending_.push_back({kind, ev});
  }
  void flush_timing() {
    if (pending_.empty()) return;
    PTYRS_HIP(hipStreamSynchronize(stream_));
    for (auto& p : pending_) {
      float ms = 0;
      PTYRS_HIP(hipEventElapsedTime(&ms, p.second.first, p.second.second));
      stats_[p.first].first += 1;
      stats_[p.first].second += ms;
      pool_.push_back(p.second);
    }
    pending_.clear();
  }
};

template <typename T, int N>
class Engine : public EngineBase {
  using S = RowShape<T, N>;    // pass A
  using SR = RowShapeC<T, N>;  // pass C
  using SC = ColShape<T, N>;   // column passes
  using SB = BShape<T, N>;     // pass B
  static constexpr int n = N * N;

 public:
  Engine(const Prepared& pre, const double* obj, size_t nobj, size_t nyo, size_t nxo,
         const double* probe, size_t nprobe, int device) {
    device_ = device;
    PTYRS_HIP(hipSetDevice(device));
    PTYRS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    nframe_ = pre.nframe;
    ny_ = nx_ = N;
    nprobe_ = nprobe;
    nobj_ = nobj;
    nyo_ = nyo;
    nxo_ = nxo;
    nb_obs_ = pre.nb_obs;
    corners = pre.corner;
    scale_ = pre.exit.scale;
    max_frame_total_ = pre.max_frame_total;
    total_ = pre.total;
    M_ = int(nprobe * nobj);
    if (M_ > kMaxModes) throw Invalid(modes_error());
    nslice_ = pre.nslice;
    graph_ = pre.graph;
    stack_size_ = pre.stack_size;
    hist_slots_ = std::max<size_t>(pre.llk_slots, 1);
    psi_f16_ = pre.psi_f16;
    half_scale_ = half_scale(pre.max_frame_total);
    reverse_ = reverse_counts_ = pre.exit.reverse;
    dft_plane_ = dft_plane_counts_ = pre.exit.dft_plane;
    plane_quad_[0] = pre.exit.plane_quad[0];
    plane_quad_[1] = pre.exit.plane_quad[1];
    // Sampling::auto (ptyrs/src/engine/sampling.rs:38-49): FullObject in the near field when
    // 5·frame > 4·object on both axes.
    sampling = (pre.exit.kind != PTYRS_FOURIER && 5 * N > 4 * int(nyo) && 5 * N > 4 * int(nxo))
                   ? PTYRS_FULL_OBJECT
                   : PTYRS_PATCH;

    O_.ensure(sizeof(cplx<T>) * nslice_ * nobj * nyo * nxo);
    P_.ensure(sizeof(cplx<T>) * nprobe * n);
    Wt_.ensure(sizeof(cplx<T>) * nprobe * n);
    q_.ensure(sizeof(cplx<T>) * n);
    tw_.ensure(sizeof(cplx<T>) * N);
    counts_.ensure(sizeof(float) * std::max<size_t>(1, nframe_ * n));
    corner_.ensure(sizeof(int2) * std::max<size_t>(1, nframe_));
    iota_.ensure(sizeof(int) * std::max<size_t>(1, nframe_));
    hist_.ensure(sizeof(double) * 4 * hist_slots_);

    upload_complex(O_, obj, nslice_ * nobj * nyo * nxo);
    if (nslice_ > 1) {
      gaptf_.ensure(sizeof(cplx<T>) * (nslice_ - 1) * n);
      std::vector<double> all;
      for (auto& g : pre.gaps) all.insert(all.end(), g.begin(), g.end());
      upload_complex(gaptf_, all.data(), (nslice_ - 1) * n);
      std::vector<double> one(2 * n, 0.0);
      for (int i = 0; i < n; ++i) one[2 * i] = 1.0;
      ones_.ensure(sizeof(cplx<T>) * n);
      upload_complex(ones_, one.data(), n);
    }
    upload_complex(P_, probe, nprobe * n);
    upload_complex(q_, pre.exit.quad.data(), n);
    if (pre.exit.kind == PTYRS_ANGULAR_SPECTRUM) {
      angular_ = true;
      banded_ = pre.exit.banded;
      tf_.ensure(sizeof(cplx<T>) * n);
      upload_complex(tf_, pre.exit.tf.data(), n);
    }
    std::vector<double> tw(2 * N);
    for (int t = 0; t < N; ++t) {
      // exp(−2πi t/N) from the exact angle of the reduced fraction.
      const long double a = -2.0L * 3.14159265358979323846264338327950288L * t / N;
      tw[2 * t] = double(cosl(a));
      tw[2 * t + 1] = double(sinl(a));
    }
    upload_complex(tw_, tw.data(), N);
    if (nframe_) {
      PTYRS_HIP(hipMemcpy(counts_.p, pre.counts.data(), sizeof(float) * nframe_ * n,
                          hipMemcpyHostToDevice));
      std::vector<int2> c(nframe_);
      std::vector<int> io(nframe_);
      for (size_t f = 0; f < nframe_; ++f) {
        c[f] = int2{pre.corner[2 * f], pre.corner[2 * f + 1]};
        io[f] = int(f);
      }
      PTYRS_HIP(hipMemcpy(corner_.p, c.data(), sizeof(int2) * nframe_, hipMemcpyHostToDevice));
      PTYRS_HIP(hipMemcpy(iota_.p, io.data(), sizeof(int) * nframe_, hipMemcpyHostToDevice));
    }
    ntx_ = int((nxo + TILE_X - 1) / TILE_X);
    nty_ = int((nyo + TILE_Y - 1) / TILE_Y);
  }

  void get_object(double* out) override {
    download_complex(O_, out, nslice_ * nobj_ * nyo_ * nxo_);
  }
  void get_probe(double* out) override { download_complex(P_, out, nprobe_ * n); }
  void set_object(const double* in) override {
    upload_complex(O_, in, nslice_ * nobj_ * nyo_ * nxo_);
  }
  void set_probe(const double* in) override { upload_complex(P_, in, nprobe_ * n); }

  void llk(double out[4]) override {
    ensure_workspace(std::min<size_t>(nframe_, kDefaultStack), 1);
    if (full_object()) fo_ensure(cap_, true);
    probe_terms();
    llk_pass(0);
    PTYRS_HIP(hipMemcpyAsync(out, hist_.p, sizeof(double) * 4, hipMemcpyDeviceToHost, stream_));
    synchronize();
    flush_timing();
    // Llk::normalised (ptyrs/src/engine/ops.rs:268-276): nb_photons_calc stays a total.
    for (int i = 0; i < 3; ++i) out[i] /= double(nb_obs_);
  }

  // The rPIE call in progress (between rpie_begin and rpie_end).
  struct RpieRun {
    ptyrs_rpie_params p{};
    size_t nepoch = 0, ic = 0;
    std::vector<std::vector<std::pair<int, int>>> spans;  // per epoch: (offset, length)
    std::vector<uint64_t> base;                           // the epoch graph's key prefix
    std::vector<std::pair<size_t, int>> due;              // LLK records: (cycle, slot)
    std::vector<int> order;                               // every epoch's batches, uploaded
  } run_;

  void rpie(const ptyrs_rpie_params& p, size_t nepoch) override {
    rpie_check(p);
    rpie_begin(p, nepoch);
    while (rpie_epoch()) {
    }
    rpie_end();
  }

  void rpie_check(const ptyrs_rpie_params& p) override {
    if (nslice_ > 1)
      throw Invalid("rPIE runs one object slice; only LSQML and the LLK pass run several");
    unitary("rPIE");
    if (nobj_ != 1) throw Invalid("rPIE needs exactly one object mode");
    if (p.batch_size == 0) throw Invalid("rPIE batch_size must be >= 1");
    if (M_ > kMaxModes) throw Invalid(modes_error());
  }

  // An rPIE call in three steps, so that several engines can interleave their epochs on their
  // own streams: begin uploads every epoch's batches (the host never waits inside the call),
  // each rpie_epoch() enqueues the next epoch (a graph replay), end drains the LLK records
  // and waits for the stream.
  void rpie_begin(const ptyrs_rpie_params& p, size_t nepoch) override {
    run_ = RpieRun{};
    run_.p = p;
    run_.nepoch = nepoch;
    if (nepoch == 0 || nframe_ == 0) {
      cycle += nepoch;
      run_.nepoch = 0;
      return;
    }
    const size_t cap = std::min(p.batch_size, nframe_);
    ensure_workspace(cap, nepoch);
    if (full_object()) fo_ensure(cap_);
    if (pos_on_ && !full_object()) {
      // Under offsets pass C leaves one probe direction per frame, moved back and averaged.
      dP_.ensure(sizeof(cplx<T>) * cap_ * M_ * n);
      dPm_.ensure(sizeof(cplx<T>) * M_ * n);
    }
    std::vector<int> order(nepoch * nframe_);
    run_.spans.assign(nepoch, {});
    for (size_t ic = 0; ic < nepoch; ++ic) {
      auto batches = random_batches(nframe_, p.batch_size, p.seed, uint64_t(cycle + ic));
      size_t off = 0;
      for (auto& b : batches) {
        std::copy(b.begin(), b.end(), order.begin() + ic * nframe_ + off);
        run_.spans[ic].push_back({int(off), int(b.size())});
        off += b.size();
      }
    }
    PTYRS_HIP(hipMemcpyAsync(order_.p, order.data(), sizeof(int) * order.size(),
                             hipMemcpyHostToDevice, stream_));
    // The host copy must outlive the asynchronous upload.
    run_.order = std::move(order);
    probe_terms();
    // |W_m|² of the starting probe terms: each batch's object and probe update keeps them
    // current for the next batch's overlaps.
    if (p.has_obj_step) run(K_BIN, [&] { norm2_planes(); });
    // Every epoch has the same batch lengths; its order is copied into the captured order
    // buffer ahead of the replay.
    run_.base = {uint64_t(p.has_obj_step), uint64_t(p.has_probe_step), bits(p.obj_step),
                 bits(p.probe_step), bits(p.obj_alpha), bits(p.probe_alpha), cap_,
                 uint64_t(pos_on_), uint64_t(pos_refine_), uint64_t(sampling)};
  }

  bool rpie_epoch() override {
    if (run_.ic >= run_.nepoch) return false;
    std::vector<uint64_t> key = rpie_epoch_order();
    epoch(rpie_graph_, key, [&] { rpie_epoch_body(); });
    return rpie_epoch_after();
  }

  // The three parts of an epoch, for a caller that captures several engines' epochs in one
  // graph: the epoch's order copied into the captured order buffer (returns the graph key of
  // the body), the body's launches, and the LLK record and counters after it.
  std::vector<uint64_t> rpie_epoch_order() override {
    const size_t ic = run_.ic;
    PTYRS_HIP(hipMemcpyAsync(eord_.p, order_.as<int>() + ic * nframe_, sizeof(int) * nframe_,
                             hipMemcpyDeviceToDevice, stream_));
    std::vector<uint64_t> key = run_.base;
    for (auto& s : run_.spans[ic]) key.push_back(uint64_t(s.second));
    key.push_back(DevBuf::generation.load());
    return key;
  }
  void rpie_epoch_body() override {
    const int* ord = eord_.as<int>();
    for (auto& s : run_.spans[run_.ic]) rpie_batch(run_.p, ord + s.first, s.second);
  }
  bool rpie_epoch_after() override {
    const size_t ic = run_.ic;
    const ptyrs_rpie_params& p = run_.p;
    const bool rec = p.calc_llk > 0 && (ic % p.calc_llk == 0 || ic + 1 == run_.nepoch);
    if (rec) {
      if (run_.due.size() == hist_slots_) drain(run_.due, PTYRS_RPIE);
      const int slot = int(run_.due.size());
      llk_pass(slot);
      run_.due.push_back({cycle, slot});
    }
    ++cycle;
    ++run_.ic;
    return run_.ic < run_.nepoch;
  }

  bool rpie_pending() const override { return run_.ic < run_.nepoch; }

  void rpie_end() override {
    drain(run_.due, PTYRS_RPIE);
    synchronize();
    flush_timing();
    run_ = RpieRun{};
  }

  // GpuPtycho::lsqml (ptyrs/src/engine/gpu/lsqml.rs:75-120) without OPR or positions: nepoch
  // epochs from a fresh LSQML (preconditioner, a_far = 1/2), batches random_batches(nframe,
  // min(batch, nframe), seed, epoch) with the epoch counted from 0 in every call.
  void lsqml(const ptyrs_lsqml_params& p, size_t nepoch, size_t calc_llk) override {
    if (nobj_ != 1) throw Invalid("the GPU LSQML runs one object mode");
    if (p.batch_size == 0) throw Invalid("LSQML batch_size must be >= 1");
    if (p.cost != PTYRS_AMPLITUDE && p.cost != PTYRS_POISSON)
      throw Invalid("cost must be amplitude or poisson");
    if (M_ > kMaxModes) throw Invalid(modes_error());
    if (nepoch == 0 || nframe_ == 0) {
      cycle += nepoch;
      return;
    }
    const size_t size = std::min(p.batch_size, nframe_);
    ensure_workspace(size, nepoch);
    ensure_lsqml(size);
    if (full_object()) fo_ensure(cap_, true);
    if (opr_on_) opr_ensure(cap_);
    std::vector<int> order(nepoch * nframe_);
    std::vector<std::vector<std::pair<int, int>>> spans(nepoch);
    for (size_t ic = 0; ic < nepoch; ++ic) {
      auto batches = random_batches(nframe_, size, p.seed, uint64_t(ic));
      size_t off = 0;
      for (auto& b : batches) {
        std::copy(b.begin(), b.end(), order.begin() + ic * nframe_ + off);
        spans[ic].push_back({int(off), int(b.size())});
        off += b.size();
      }
    }
    PTYRS_HIP(hipMemcpyAsync(order_.p, order.data(), sizeof(int) * order.size(),
                             hipMemcpyHostToDevice, stream_));
    {
      std::vector<T> half(nframe_, T(0.5));
      PTYRS_HIP(hipMemcpyAsync(afar_all_.p, half.data(), sizeof(T) * nframe_,
                               hipMemcpyHostToDevice, stream_));
      PTYRS_HIP(hipStreamSynchronize(stream_));
    }
    probe_terms();
    std::vector<std::pair<size_t, int>> due;
    for (size_t ic = 0; ic < nepoch; ++ic) {
      if (ic < 10 || ic % 10 == 0) precondition(ic == 0);
      PTYRS_HIP(hipMemcpyAsync(eord_.p, order_.as<int>() + ic * nframe_, sizeof(int) * nframe_,
                               hipMemcpyDeviceToDevice, stream_));
      const int* ord = eord_.as<int>();
      if (full_object() && nslice_ == 1) {
        for (auto& s : spans[ic]) lsqml_batch_sampled(p, ord + s.first, s.second, s.first);
      } else if (nslice_ > 1 || pos_on_ || opr_on_) {
        for (auto& s : spans[ic]) lsqml_batch_ms(p, ord + s.first, s.second, s.first);
      } else {
        std::vector<uint64_t> key = {uint64_t(p.cost), cap_};
        for (auto& s : spans[ic]) key.push_back(uint64_t(s.second));
        epoch(lsqml_graph_, key, [&] {
          for (auto& s : spans[ic]) lsqml_batch(p, ord + s.first, s.second, s.first);
        });
      }
      const bool rec = calc_llk > 0 && (ic % calc_llk == 0 || ic + 1 == nepoch);
      if (rec) {
        if (due.size() == hist_slots_) drain(due, PTYRS_LSQML);
        const int slot = int(due.size());
        llk_pass(slot);
        due.push_back({cycle, slot});
      }
      ++cycle;
    }
    drain(due, PTYRS_LSQML);
    // The last epoch's steps.
    const size_t ls = 2 + nslice_;
    std::vector<double> lg(ls * nframe_);
    PTYRS_HIP(hipMemcpyAsync(lg.data(), log_.p, sizeof(double) * ls * nframe_,
                             hipMemcpyDeviceToHost, stream_));
    synchronize();
    flush_timing();
    lsqml_steps.clear();
    const int* last = order.data() + (nepoch - 1) * nframe_;
    int bi = 0;
    for (auto& s : spans[nepoch - 1]) {
      for (int k = 0; k < s.second; ++k) {
        const size_t slot = size_t(s.first + k);
        const double* r = lg.data() + ls * slot;
        lsqml_steps.push_back(StepRecord{bi, last[slot], r[0], r[1], r[ls - 1],
                                         std::vector<double>(r + 1, r + 1 + nslice_)});
      }
      ++bi;
    }
  }

  // pty-chi's BH on the amplitude cost (GpuPtycho::bh, ptyrs/src/engine/gpu/bh.rs:33-68;
  // Bh::epoch, ptyrs/src/engine/bh.rs:684-799): `nepoch` epochs from a fresh BH over
  // random_batches(nframe, batch_size, seed, epoch); with one batch of every frame, conjugate
  // gradient from the second epoch of the call. An LLK record after the epochs `calc_llk` asks.
  void bh(const ptyrs_bh_params& p, size_t nepoch, size_t calc_llk) override {
    if (p.batch_size == 0) throw Invalid("BH batch_size must be >= 1");
    if (!std::isfinite(p.probe_rho) || !std::isfinite(p.position_rho))
      throw Invalid("BH probe_rho and position_rho must be finite");
    if (nslice_ > 1 && pos_on_ && !full_object())
      throw Invalid("BH through several object slices runs without position offsets or with "
                    "offsets under FullObject sampling on this GPU engine");
    if (full_object() && nobj_ != 1)
      throw Invalid("FullObject sampling with several object modes is not run on this GPU "
                    "engine");
    if (nobj_ != 1) throw Invalid("the GPU BH runs one object mode");
    if (M_ > kMaxModes) throw Invalid(modes_error());
    if (nepoch == 0 || nframe_ == 0) {
      cycle += nepoch;
      return;
    }
    const size_t size = std::min(p.batch_size, nframe_);
    const bool full = size == nframe_;
    ensure_workspace(size, nepoch);
    ensure_bh(size);
    std::vector<int> order(nepoch * nframe_);
    std::vector<std::vector<std::pair<int, int>>> spans(nepoch);
    for (size_t ic = 0; ic < nepoch; ++ic) {
      auto batches = random_batches(nframe_, size, p.seed, uint64_t(ic));
      size_t off = 0;
      for (auto& b : batches) {
        std::copy(b.begin(), b.end(), order.begin() + ic * nframe_ + off);
        spans[ic].push_back({int(off), int(b.size())});
        off += b.size();
      }
    }
    PTYRS_HIP(hipMemcpyAsync(order_.p, order.data(), sizeof(int) * order.size(),
                             hipMemcpyHostToDevice, stream_));
    probe_terms();
    const double rho2 = p.probe_rho * p.probe_rho * scale_;
    const double pos2 = p.position_rho * p.position_rho;
    if (full_object() && nslice_ > 1) {
      fo_ensure(cap_, true);
      ensure_bh_sampled(size);
    } else if (full_object()) {
      fo_ensure(cap_);
      ensure_bh_sampled(size);
    } else if (pos_on_) {
      ensure_bh_moved(size, full);
    }
    std::vector<std::pair<size_t, int>> due;
    for (size_t ic = 0; ic < nepoch; ++ic) {
      PTYRS_HIP(hipMemcpyAsync(eord_.p, order_.as<int>() + ic * nframe_, sizeof(int) * nframe_,
                               hipMemcpyDeviceToDevice, stream_));
      const int* ord = eord_.as<int>();
      const bool cg = full && ic > 0;
      if (nslice_ > 1) {
        // Through several slices the batches launch eagerly, as LSQML's.
        int k = 0;
        for (auto& s : spans[ic])
          bh_batch_ms(rho2, ord + s.first, s.second, double(s.second) / double(size), cg, full,
                      k++, pos2);
      } else if (full_object()) {
        int k = 0;
        for (auto& s : spans[ic])
          bh_batch_sampled(rho2, pos2, ord + s.first, s.second,
                           double(s.second) / double(size), cg, full, k++);
      } else if (pos_on_) {
        // Under offsets the batches launch eagerly, as the other position paths.
        int k = 0;
        for (auto& s : spans[ic])
          bh_batch_moved(rho2, pos2, ord + s.first, s.second, double(s.second) / double(size),
                         cg, full, k++);
      } else {
        std::vector<uint64_t> key = {uint64_t(cg), uint64_t(full), cap_, bits_of(rho2)};
        for (auto& s : spans[ic]) key.push_back(uint64_t(s.second));
        epoch(bh_graph_, key, [&] {
          int k = 0;
          for (auto& s : spans[ic])
            bh_batch(rho2, ord + s.first, s.second, double(s.second) / double(size), cg, full,
                     k++);
        });
      }
      const bool rec = calc_llk > 0 && (ic % calc_llk == 0 || ic + 1 == nepoch);
      if (rec) {
        if (due.size() == hist_slots_) drain(due, PTYRS_BH);
        const int slot = int(due.size());
        llk_pass(slot);
        due.push_back({cycle, slot});
      }
      ++cycle;
    }
    drain(due, PTYRS_BH);
    const auto& last = spans[nepoch - 1];
    std::vector<double> lg(4 * last.size());
    PTYRS_HIP(hipMemcpyAsync(lg.data(), bh_log_.p, sizeof(double) * lg.size(),
                             hipMemcpyDeviceToHost, stream_));
    synchronize();
    flush_timing();
    bh_steps.clear();
    for (size_t k = 0; k < last.size(); ++k)
      bh_steps.push_back(BhRecord{size_t(last[k].first), size_t(last[k].second), lg[4 * k],
                                  lg[4 * k + 1], lg[4 * k + 2], lg[4 * k + 3]});
    bh_order.assign(order.begin() + (nepoch - 1) * nframe_, order.end());
  }

  // PyNX AP, DM, RAAR (ptyrs/src/engine/ptycho.rs:605-755), stacks of the whole scan; DM and
  // RAAR carry their exit waves across the cycles of a call.
  void project(const ProjParams& p, size_t ncycle) override {
    const char* name = p.algorithm == PTYRS_AP ? "AP" : p.algorithm == PTYRS_DM ? "DM" : "RAAR";
    if (nslice_ > 1)
      throw Invalid(std::string(name) +
                    " runs one object slice; only LSQML and the LLK pass run several");
    unitary(name);
    if (pos_on_ && psi_f16_ && p.algorithm != PTYRS_AP)
      throw Invalid(std::string(name) +
                    " with position offsets does not carry f16 waves on this GPU engine");
    if (nobj_ != 1) throw Invalid("the GPU projection algorithms run one object mode");
    if (M_ > kMaxModes) throw Invalid(modes_error());
    if (p.algorithm != PTYRS_AP && p.loops == 0)
      throw Invalid("loop_obj_probe must be >= 1");
    if (pos_on_ && ncycle > 0 && nframe_ > 0) {
      if (full_object())
        project_sampled(p, ncycle);
      else
        project_moved(p, ncycle);
      return;
    }
    if (ncycle == 0 || nframe_ == 0) {
      cycle += ncycle;
      if (p.zero_phase_ramp && ncycle > 0) {
        device_zero_phase_ramp();
        synchronize();
      }
      return;
    }
    const size_t stack = std::min<size_t>(nframe_, stack_size_ ? stack_size_ : kDefaultStack);
    ensure_workspace(stack, 1);
    const size_t nb_all = nframe_;
    // psi_f16 holds DM's and RAAR's carried waves as f16 pairs (AP's stay complex).
    const bool half = psi_f16_ && p.algorithm != PTYRS_AP;
    DevBuf& cbuf = half ? carried16_ : carried_;
    DevBuf& ibuf = half ? iter16_ : iter_;
    const size_t wbytes = half ? sizeof(uint32_t) : sizeof(cplx<T>);
    cbuf.ensure(wbytes * nb_all * M_ * n);
    if (p.algorithm == PTYRS_RAAR) ibuf.ensure(wbytes * nb_all * M_ * n);
    const int nt = ntiles();
    list_all_.ensure(sizeof(int) * size_t(nt) * nb_all);
    count_all_.ensure(sizeof(int) * size_t(nt));
    const int nstack = int((nb_all + stack - 1) / stack);
    const int chunk = chunk_of(int(stack));
    const int nchunk = int((stack + chunk - 1) / chunk);
    pnum_part_.ensure(sizeof(cplx<T>) * size_t(nstack) * nchunk * M_ * n);
    pden_part_.ensure(sizeof(T) * size_t(nstack) * nchunk * n);
    pnum_.ensure(sizeof(cplx<T>) * M_ * n);
    pden_.ensure(sizeof(T) * n);
    pmaxpart_.ensure(sizeof(T) * ((n + THREADS - 1) / THREADS));
    pmax2_.ensure(sizeof(T) * 2);
    abs_part_.ensure(sizeof(double) * 2 * kAbsBlocks);
    probe_terms();
    std::vector<std::pair<size_t, int>> due;
    for (size_t ic = 0; ic < ncycle; ++ic) {
      const bool rec = p.calc_llk > 0 && (ic % p.calc_llk == 0 || ic + 1 == ncycle);
      if (rec) {
        if (due.size() == hist_slots_) drain(due, p.algorithm);
        const int slot = int(due.size());
        llk_pass(slot);
        due.push_back({cycle, slot});
      }
      // Every call's first cycle starts from ψ = PO (ptyrs/src/engine/ptycho.rs:663-668).
      const bool first = ic == 0;
      const std::vector<uint64_t> key = {uint64_t(p.algorithm), uint64_t(first), bits(p.mix),
                                         uint64_t(p.loops), uint64_t(p.upd_obj),
                                         uint64_t(p.upd_probe), bits(p.obj_inertia),
                                         bits(p.probe_inertia), stack, cap_};
      epoch(proj_graph_, key, [&] {
      with_carried(half, [&](auto cw, auto iw) {
      constexpr bool H = decltype(cw)::half;
      MixArgs<T, H> mix{MIX_NONE, cw, iw, T(0), T(0)};
      if (p.algorithm == PTYRS_DM && !first) {
        mix.kind = MIX_DM;
        mix.c0 = T(2.0 - p.mix);
        mix.c1 = T(1.0 - p.mix);
      } else if (p.algorithm == PTYRS_RAAR) {
        mix.kind = first ? MIX_RAAR0 : MIX_RAAR;
        mix.c0 = T(1.0 - 0.5 * p.mix);
        mix.c1 = T(0.5 * p.mix);
      }
      const bool ap_probe = p.algorithm == PTYRS_AP && p.upd_probe;
      for (int s = 0; s < nstack; ++s) {
        const int f0 = int(s * stack);
        const int nb = std::min<int>(int(stack), int(nframe_) - f0);
        const int* ord = iota_.as<int>() + f0;
        pass_a(ord, nb, false, mix);
        pass_b<B_AMPLITUDE>(ord, nb);
        const int ck = chunk_of(nb);
        const int nck = (nb + ck - 1) / ck;
        const bool dm = p.algorithm == PTYRS_DM && !first;
        run(K_PASS_C, [&] {
          hipLaunchKernelGGL((k_pass_c_proj<T, N, H>), dim3(SR::NBLK, M_, nck), dim3(SR::NT), 0,
                             stream_, psi_.as<cplx<T>>(), cw, O_.as<cplx<T>>(), int(nxo_),
                             Wt_.as<cplx<T>>(), q_.as<cplx<T>>(), corner_.as<int2>(), ord, nb,
                             M_, ck, int(dm), T(2.0 - p.mix), T(1.0 - p.mix), int(ap_probe),
                             pnum_part_.as<cplx<T>>() + size_t(s) * nchunk * M_ * n,
                             pden_part_.as<T>() + size_t(s) * nchunk * n, tw_.as<cplx<T>>());
        });
        if (ap_probe && nck < nchunk) zero_partials(s, nck, nchunk);
      }
      const int tot_chunks = nstack * nchunk;
      if (p.algorithm == PTYRS_AP) {
        if (p.upd_obj) object_blend(p.obj_inertia, cw);
        if (p.upd_probe) probe_blend(tot_chunks, p.probe_inertia);
      } else {
        for (size_t l = 0; l < p.loops; ++l) {
          if (p.upd_obj) object_blend(p.obj_inertia, cw);
          if (p.upd_probe) {
            probe_sums_all(stack, nstack, nchunk, cw);
            probe_blend(tot_chunks, p.probe_inertia);
          }
        }
      }
      });
      });
      if (rec && p.algorithm != PTYRS_AP) balance(1.0);
      ++cycle;
    }
    drain(due, p.algorithm);
    if (p.zero_phase_ramp) device_zero_phase_ramp();
    synchronize();
    flush_timing();
  }

  // AP, DM and RAAR under PATCH offsets (ptyrs/src/engine/ptycho.rs:605-755 with
  // Shifted::psi_to_obj_probe, ptyrs/src/engine/features.rs:653-720, 851-1047): every frame's
  // moved wave W at the cycle's probe (wall_) takes the probe terms' place in pass A and in DM's
  // PO of pass C; the position sums of the new carried waves against the arrays the cycle read;
  // the object blends Σ conj(W)·ψ over Σ|W|²; the probe takes the first conjugate-gradient step
  // of its least-squares problem (probe_cg_step); then the position step. AP updates both from
  // the arrays on entry (the probe first, the object with the waves of the probe on entry); DM
  // and RAAR loop object-only then probe-only updates, the probe seeing the new object.
  void project_moved(const ProjParams& p, size_t ncycle) {
    // The offsets' derivative planes (two per view) fill one launch's grid y.
    const size_t stack = std::min<size_t>({nframe_, stack_size_ ? stack_size_ : kDefaultStack,
                                           size_t(32767 / M_)});
    ensure_workspace(stack, 1);
    const int xb = aux_blocks();
    const size_t vplane = size_t(M_) * n;
    carried_.ensure(sizeof(cplx<T>) * nframe_ * vplane);
    if (p.algorithm == PTYRS_RAAR) iter_.ensure(sizeof(cplx<T>) * nframe_ * vplane);
    wall_.ensure(sizeof(cplx<T>) * nframe_ * vplane);
    if (pos_refine_) {
      dwpos_.ensure(sizeof(cplx<T>) * 2 * stack * vplane);
      pos_all_.ensure(sizeof(double) * 4 * nframe_ * xb);
      pstep_.ensure(sizeof(double) * 4 * nframe_);
    }
    const int nt = ntiles();
    list_all_.ensure(sizeof(int) * size_t(nt) * nframe_);
    count_all_.ensure(sizeof(int) * size_t(nt));
    for (DevBuf* b : {&gP_, &dPc_, &dWt_}) b->ensure(sizeof(cplx<T>) * nprobe_ * n);
    pden_.ensure(sizeof(T) * n);
    pmax2_.ensure(sizeof(T) * 2);
    cg_part_.ensure(sizeof(double) * 2 * size_t(M_) * xb);
    vpow_.ensure(sizeof(double) * nframe_ * vplane / n * xb);
    alpha_.ensure(sizeof(double) * M_);
    abs_part_.ensure(sizeof(double) * 2 * kAbsBlocks);
    const int nstack = int((nframe_ + stack - 1) / stack);
    probe_terms();
    cplx<T>* wall = wall_.as<cplx<T>>();
    cplx<T>* carried = carried_.as<cplx<T>>();
    std::vector<std::pair<size_t, int>> due;
    for (size_t ic = 0; ic < ncycle; ++ic) {
      const bool rec = p.calc_llk > 0 && (ic % p.calc_llk == 0 || ic + 1 == ncycle);
      if (rec) {
        if (due.size() == hist_slots_) drain(due, p.algorithm);
        const int slot = int(due.size());
        llk_pass(slot);
        due.push_back({cycle, slot});
      }
      const bool first = ic == 0;
      MixArgs<T> mix{MIX_NONE, Carried<T, false>{carried},
                     Carried<T, false>{iter_.as<cplx<T>>()}, T(0), T(0)};
      if (p.algorithm == PTYRS_DM && !first) {
        mix.kind = MIX_DM;
        mix.c0 = T(2.0 - p.mix);
        mix.c1 = T(1.0 - p.mix);
      } else if (p.algorithm == PTYRS_RAAR) {
        mix.kind = first ? MIX_RAAR0 : MIX_RAAR;
        mix.c0 = T(1.0 - 0.5 * p.mix);
        mix.c1 = T(0.5 * p.mix);
      }
      const bool dm = p.algorithm == PTYRS_DM && !first;
      moved_waves_all(stack, nstack);
      for (int s = 0; s < nstack; ++s) {
        const int f0 = int(s * stack);
        const int nb = std::min<int>(int(stack), int(nframe_) - f0);
        const int* ord = iota_.as<int>() + f0;
        const cplx<T>* ws = wall + size_t(f0) * vplane;
        const int K = nb * M_;
        pass_a(ord, nb, false, mix, ws);
        pass_b<B_AMPLITUDE>(ord, nb);
        const int ck = chunk_of(nb);
        const int nck = (nb + ck - 1) / ck;
        run(K_PASS_C, [&] {
          hipLaunchKernelGGL((k_pass_c_proj<T, N, false>), dim3(SR::NBLK, M_, nck), dim3(SR::NT),
                             0, stream_, psi_.as<cplx<T>>(), Carried<T, false>{carried},
                             O_.as<cplx<T>>(), int(nxo_), ws, q_.as<cplx<T>>(),
                             corner_.as<int2>(), ord, nb, M_, ck, int(dm), T(2.0 - p.mix),
                             T(1.0 - p.mix), 0, (cplx<T>*)nullptr, (T*)nullptr,
                             tw_.as<cplx<T>>(), 1);
        });
        if (pos_refine_) {
          // The position sums of the new carried waves (spec_ holds the probe's spectra).
          run(K_OTHER, [&] {
            hipLaunchKernelGGL(k_ramp_deriv_planes<T>, dim3((n + 255) / 256, K), dim3(256), 0,
                               stream_, dwpos_.as<cplx<T>>(),
                               (const cplx<T>*)spec_.as<cplx<T>>(),
                               (const double2*)offs_.as<double2>(), ord, N, M_, 1.0 / double(n));
            hipLaunchKernelGGL((k_cols<T, N, true>), dim3(SC::NBLK, 2 * K), dim3(SC::NT), 0,
                               stream_, dwpos_.as<cplx<T>>(), tw_.as<cplx<T>>());
            hipLaunchKernelGGL((k_rows<T, N, true>), dim3(S::NBLK, 2 * K), dim3(S::NT), 0,
                               stream_, dwpos_.as<cplx<T>>(), tw_.as<cplx<T>>());
            hipLaunchKernelGGL(k_pos_sums_proj<T>, dim3(xb, nb), dim3(THREADS), 0, stream_,
                               (const cplx<T>*)(carried + size_t(f0) * vplane), ws,
                               (const cplx<T>*)O_.as<cplx<T>>(), int(nxo_), corner_.as<int2>(),
                               ord, (const cplx<T>*)dwpos_.as<cplx<T>>(), N, M_,
                               pos_all_.as<double>() + size_t(f0) * xb * 4);
          });
        }
      }
      if (p.algorithm == PTYRS_AP) {
        if (p.upd_probe) probe_cg_step(stack, nstack, p.probe_inertia);
        if (p.upd_obj) object_blend_moved(p.obj_inertia);
      } else {
        bool stale = false;  // wall_ is of an earlier probe
        for (size_t l = 0; l < p.loops; ++l) {
          if (p.upd_obj) {
            if (stale) moved_waves_all(stack, nstack);
            stale = false;
            object_blend_moved(p.obj_inertia);
          }
          if (p.upd_probe) {
            if (stale) moved_waves_all(stack, nstack);
            probe_cg_step(stack, nstack, p.probe_inertia);
            stale = true;
          }
        }
      }
      if (pos_refine_) {
        run(K_OTHER, [&] {
          hipLaunchKernelGGL(k_pos_step_all, dim3(1), dim3(1024), 0, stream_,
                             (const double*)pos_all_.as<double>(), xb, int(nframe_),
                             offs_.as<double2>(), pstep_.as<double>());
        });
      }
      if (rec && p.algorithm != PTYRS_AP) balance(1.0);
      ++cycle;
    }
    drain(due, p.algorithm);
    if (p.zero_phase_ramp) device_zero_phase_ramp();
    synchronize();
    flush_timing();
  }

  // Every frame's moved wave W = DFT⁻¹(DFT(P·q)·exp(−2πi f·d)) at the current probe into
  // wall_ ([frame·M + m][n]); spec_ keeps the probe terms' spectra.
  void moved_waves_all(size_t stack, int nstack) {
    run(K_PROBE, [&] { probe_spectra(); });
    for (int s = 0; s < nstack; ++s) {
      const int f0 = int(s * stack);
      const int nb = std::min<int>(int(stack), int(nframe_) - f0);
      moved_probes(iota_.as<int>() + f0, nb, false,
                   wall_.as<cplx<T>>() + size_t(f0) * M_ * n, nullptr, true);
    }
  }

  // The object's blend through the moved waves: Σ conj(W)·ψ over Σ|W|² (+ reg·O).
  void object_blend_moved(double inertia) {
    const int nt = ntiles();
    const cplx<T>* wall = wall_.as<cplx<T>>();
    run(K_BIN, [&] {
      hipLaunchKernelGGL(k_bin_tiles<T>, dim3(nt), dim3(THREADS), 0, stream_,
                         corner_.as<int2>(), iota_.as<int>(), int(nframe_), N, ntx_,
                         int(nframe_), list_all_.as<int>(), count_all_.as<int>(),
                         (const T*)nullptr, M_, int(nyo_), int(nxo_), D_.as<T>(),
                         tilemax_.as<T>(), wall);
    });
    run(K_OVERLAP, [&] {
      hipLaunchKernelGGL(k_max_reduce<T>, dim3(1), dim3(MAX_THREADS), 0, stream_,
                         tilemax_.as<T>(), nt, pmax_.as<T>());
    });
    run(K_OBJ_UPDATE, [&] {
      hipLaunchKernelGGL((k_obj_blend<T, false>), dim3(nt), dim3(THREADS), 0, stream_,
                         O_.as<cplx<T>>(), wall, D_.as<T>(),
                         Carried<T, false>{carried_.as<cplx<T>>()}, N, M_, corner_.as<int2>(),
                         iota_.as<int>(), list_all_.as<int>(), count_all_.as<int>(),
                         int(nframe_), int(nyo_), int(nxo_), ntx_, pmax_.as<T>(), T(inertia), 1);
    });
  }

  // The probe's first conjugate-gradient step of min ‖ψ − A·P'‖² + reg·‖P' − P‖² under the
  // moves (ptyrs/src/engine/features.rs:673-711): g = Σ_f the adjoint move of conj(O)·(ψ − W·O)
  // times conj(q), D = Σ|O|² + reg (reg = inertia·max Σ|O|²), s = g/D, and per mode
  // P += (<g, s> / (‖A·s‖² + reg·‖s‖²))·s where <g, s> > 0; A·s the moved waves of s·q times the
  // windows. wall_ holds the current probe's waves.
  void probe_cg_step(size_t stack, int nstack, double inertia) {
    const int xb = aux_blocks();
    const size_t vplane = size_t(M_) * n;
    const int np = int(nprobe_ * n);
    cplx<T>* g = gspec_.as<cplx<T>>();
    for (int s = 0; s < nstack; ++s) {
      const int f0 = int(s * stack);
      const int nb = std::min<int>(int(stack), int(nframe_) - f0);
      const int* ord = iota_.as<int>() + f0;
      const int K = nb * M_;
      run(K_PROBE, [&] {
        hipLaunchKernelGGL(k_resid_dir<T>, dim3((n + 255) / 256, K), dim3(256), 0, stream_, g,
                           (const cplx<T>*)carried_.as<cplx<T>>(),
                           (const cplx<T>*)(wall_.as<cplx<T>>() + size_t(f0) * vplane),
                           (const cplx<T>*)O_.as<cplx<T>>(), int(nxo_), corner_.as<int2>(), ord,
                           N, M_);
        hipLaunchKernelGGL((k_rows<T, N, false>), dim3(S::NBLK, K), dim3(S::NT), 0, stream_, g,
                           tw_.as<cplx<T>>());
        hipLaunchKernelGGL((k_cols<T, N, false>), dim3(SC::NBLK, K), dim3(SC::NT), 0, stream_,
                           g, tw_.as<cplx<T>>());
        hipLaunchKernelGGL(k_unramp_mean<T>, dim3((n + 255) / 256, M_), dim3(256), 0, stream_,
                           gP_.as<cplx<T>>(), (const cplx<T>*)g,
                           (const double2*)offs_.as<double2>(), ord, nb, N, M_,
                           1.0 / double(n), int(s > 0));
      });
    }
    run(K_PROBE, [&] {
      hipLaunchKernelGGL((k_cols<T, N, true>), dim3(SC::NBLK, M_), dim3(SC::NT), 0, stream_,
                         gP_.as<cplx<T>>(), tw_.as<cplx<T>>());
      hipLaunchKernelGGL((k_rows<T, N, true>), dim3(S::NBLK, M_), dim3(S::NT), 0, stream_,
                         gP_.as<cplx<T>>(), tw_.as<cplx<T>>());
      hipLaunchKernelGGL((k_times_q<T, true>), dim3((n + 255) / 256, M_), dim3(256), 0, stream_,
                         gP_.as<cplx<T>>(), (const cplx<T>*)q_.as<cplx<T>>(), n);
      hipLaunchKernelGGL(k_win_power<T>, dim3((n + 255) / 256), dim3(256), 0, stream_,
                         pden_.as<T>(), (const cplx<T>*)O_.as<cplx<T>>(), int(nxo_),
                         corner_.as<int2>(), iota_.as<int>(), int(nframe_), N);
      hipLaunchKernelGGL(k_max_reduce<T>, dim3(1), dim3(MAX_THREADS), 0, stream_,
                         (const T*)pden_.as<T>(), n, pmax2_.as<T>());
      hipLaunchKernelGGL(k_cg_probe_sums<T>, dim3(xb, M_), dim3(THREADS), 0, stream_,
                         (const cplx<T>*)gP_.as<cplx<T>>(), (const T*)pden_.as<T>(),
                         (const T*)pmax2_.as<T>(), T(inertia), n, dPc_.as<cplx<T>>(),
                         cg_part_.as<double>());
      // A·s: the moved waves of s·q against the windows.
      hipLaunchKernelGGL(k_probe_terms<T>, dim3((np + 255) / 256), dim3(256), 0, stream_,
                         dPc_.as<cplx<T>>(), q_.as<cplx<T>>(), dWt_.as<cplx<T>>(), n, np);
      probe_spectra(dWt_.as<cplx<T>>());
    });
    for (int s = 0; s < nstack; ++s) {
      const int f0 = int(s * stack);
      const int nb = std::min<int>(int(stack), int(nframe_) - f0);
      const int* ord = iota_.as<int>() + f0;
      moved_probes(ord, nb, false, inc0_.as<cplx<T>>(), nullptr, true);
      run(K_PROBE, [&] {
        hipLaunchKernelGGL(k_view_power<T>, dim3(xb, nb * M_), dim3(THREADS), 0, stream_,
                           (const cplx<T>*)inc0_.as<cplx<T>>(), (const cplx<T>*)O_.as<cplx<T>>(),
                           int(nxo_), corner_.as<int2>(), ord, N, M_, size_t(f0) * M_,
                           vpow_.as<double>());
      });
    }
    run(K_PROBE, [&] {
      hipLaunchKernelGGL(k_cg_probe_alpha, dim3(1), dim3(1), 0, stream_,
                         (const double*)cg_part_.as<double>(), xb,
                         (const double*)vpow_.as<double>(), xb, int(nframe_) * M_, M_,
                         alpha_.as<double>());
      hipLaunchKernelGGL(k_probe_axpy<T>, dim3((np + 255) / 256), dim3(256), 0, stream_,
                         P_.as<cplx<T>>(), (const cplx<T>*)dPc_.as<cplx<T>>(),
                         (const double*)alpha_.as<double>(), n, np);
    });
    probe_terms();
  }

  // AP, DM and RAAR under FullObject offsets (ptyrs/src/engine/ptycho.rs:605-755 with the
  // FullObject psi_to_obj_probe, ptyrs/src/engine/features.rs:653-793): the probe stays put and
  // every frame reads its window sampled at its corner plus offset (fo_win_, read by passes A
  // and C at the tall corners); the position sums of the new carried waves through the
  // windows' derivatives; the probe takes the first conjugate-gradient step of its
  // least-squares problem through the reads (probe_step_sampled); the object solves its
  // least-squares problem by preconditioned conjugate gradients (object_solve_sampled). AP
  // updates both from the arrays on entry; DM and RAAR loop object-only then probe-only
  // updates, the probe seeing the new object.
  void project_sampled(const ProjParams& p, size_t ncycle) {
    const size_t stack = std::min<size_t>(nframe_, stack_size_ ? stack_size_ : kDefaultStack);
    ensure_workspace(stack, 1);
    fo_ensure(stack);
    const int xb = aux_blocks();
    const size_t vplane = size_t(M_) * n, no = nyo_ * nxo_;
    carried_.ensure(sizeof(cplx<T>) * nframe_ * vplane);
    if (p.algorithm == PTYRS_RAAR) iter_.ensure(sizeof(cplx<T>) * nframe_ * vplane);
    if (pos_refine_) {
      fo_dd_.ensure(sizeof(cplx<T>) * 2 * stack * n);
      pos_all_.ensure(sizeof(double) * 4 * nframe_ * xb);
      pstep_.ensure(sizeof(double) * 4 * nframe_);
    }
    for (DevBuf* b : {&gP_, &dPc_}) b->ensure(sizeof(cplx<T>) * nprobe_ * n);
    for (DevBuf* b : {&gO_, &hO_, &dOc_}) b->ensure(sizeof(cplx<T>) * no);
    fo_p2_.ensure(sizeof(cplx<T>) * n);
    fo_pcg_d_.ensure(sizeof(T) * no);
    fo_pcg_pinv_.ensure(sizeof(T) * no);
    fo_pcg_sc_.ensure(sizeof(double) * 4);
    pden_.ensure(sizeof(T) * n);
    pmax2_.ensure(sizeof(T) * 2);
    cg_part_.ensure(sizeof(double) * std::max<size_t>(2 * size_t(M_) * xb, kAbsBlocks));
    vpow_.ensure(sizeof(double) * M_ * xb);
    alpha_.ensure(sizeof(double) * M_);
    abs_part_.ensure(sizeof(double) * 2 * kAbsBlocks);
    const int nstack = int((nframe_ + stack - 1) / stack);
    probe_terms();
    cplx<T>* carried = carried_.as<cplx<T>>();
    cplx<T>* win = fo_win_.as<cplx<T>>();
    std::vector<std::pair<size_t, int>> due;
    for (size_t ic = 0; ic < ncycle; ++ic) {
      const bool rec = p.calc_llk > 0 && (ic % p.calc_llk == 0 || ic + 1 == ncycle);
      if (rec) {
        if (due.size() == hist_slots_) drain(due, p.algorithm);
        const int slot = int(due.size());
        llk_pass(slot);
        due.push_back({cycle, slot});
      }
      const bool first = ic == 0;
      MixArgs<T> mix{MIX_NONE, Carried<T, false>{carried},
                     Carried<T, false>{iter_.as<cplx<T>>()}, T(0), T(0)};
      if (p.algorithm == PTYRS_DM && !first) {
        mix.kind = MIX_DM;
        mix.c0 = T(2.0 - p.mix);
        mix.c1 = T(1.0 - p.mix);
      } else if (p.algorithm == PTYRS_RAAR) {
        mix.kind = first ? MIX_RAAR0 : MIX_RAAR;
        mix.c0 = T(1.0 - 0.5 * p.mix);
        mix.c1 = T(0.5 * p.mix);
      }
      const bool dm = p.algorithm == PTYRS_DM && !first;
      run(K_BIN, [&] { fo_object_spectrum(O_.as<cplx<T>>()); });
      for (int s = 0; s < nstack; ++s) {
        const int f0 = int(s * stack);
        const int nb = std::min<int>(int(stack), int(nframe_) - f0);
        const int* ord = iota_.as<int>() + f0;
        run(K_BIN, [&] {
          fo_points(ord, nb);
          fo_object_windows(nb);
        });
        pass_a(ord, nb, false, mix, nullptr, win, N, fo_tall_.as<int2>());
        pass_b<B_AMPLITUDE>(ord, nb);
        const int ck = chunk_of(nb);
        const int nck = (nb + ck - 1) / ck;
        run(K_PASS_C, [&] {
          hipLaunchKernelGGL((k_pass_c_proj<T, N, false>), dim3(SR::NBLK, M_, nck), dim3(SR::NT),
                             0, stream_, psi_.as<cplx<T>>(), Carried<T, false>{carried},
                             (const cplx<T>*)win, N, Wt_.as<cplx<T>>(), q_.as<cplx<T>>(),
                             fo_tall_.as<int2>(), ord, nb, M_, ck, int(dm), T(2.0 - p.mix),
                             T(1.0 - p.mix), 0, (cplx<T>*)nullptr, (T*)nullptr,
                             tw_.as<cplx<T>>(), 0);
        });
        if (pos_refine_) {
          // The position sums of the new carried waves (fo_spec_ holds the object's spectrum).
          run(K_OTHER, [&] {
            fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 1, fo_dd_.p, 2ll * n);
            fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 2, fo_dd_.as<cplx<T>>() + n,
                                 2ll * n);
            hipLaunchKernelGGL(k_pos_sums_fo<T>, dim3(xb, nb), dim3(THREADS), 0, stream_,
                               (const cplx<T>*)(carried + size_t(f0) * vplane),
                               (const cplx<T>*)Wt_.as<cplx<T>>(), (const cplx<T>*)win,
                               (const cplx<T>*)fo_dd_.as<cplx<T>>(), n, M_,
                               pos_all_.as<double>() + size_t(f0) * xb * 4);
          });
        }
      }
      if (p.algorithm == PTYRS_AP) {
        // Both from the arrays on entry: the probe's step is applied after the object's solve,
        // which reads the probe on entry.
        if (p.upd_probe) probe_step_sampled(stack, nstack, p.probe_inertia);
        if (p.upd_obj) object_solve_sampled(stack, nstack, p.obj_inertia);
        if (p.upd_probe) probe_apply();
      } else {
        bool stale = false;  // fo_spec_ is not the object's spectrum
        for (size_t l = 0; l < p.loops; ++l) {
          if (p.upd_obj) {
            object_solve_sampled(stack, nstack, p.obj_inertia);
            stale = true;
          }
          if (p.upd_probe) {
            if (stale) run(K_BIN, [&] { fo_object_spectrum(O_.as<cplx<T>>()); });
            stale = false;
            probe_step_sampled(stack, nstack, p.probe_inertia);
            probe_apply();
          }
        }
      }
      if (pos_refine_) {
        run(K_OTHER, [&] {
          hipLaunchKernelGGL(k_pos_step_all, dim3(1), dim3(1024), 0, stream_,
                             (const double*)pos_all_.as<double>(), xb, int(nframe_),
                             offs_.as<double2>(), pstep_.as<double>());
        });
      }
      if (rec && p.algorithm != PTYRS_AP) balance(1.0);
      ++cycle;
    }
    drain(due, p.algorithm);
    if (p.zero_phase_ramp) device_zero_phase_ramp();
    synchronize();
    flush_timing();
  }

  // The probe's first conjugate-gradient step of min ‖ψ − A·P'‖² + reg·‖P' − P‖² under
  // FullObject (ptyrs/src/engine/features.rs:673-711), the reads S of every frame not moving
  // with the probe: g_m = Σ_f conj(S_f·q)·(ψ_{f,m} − W_m·S_f), D = Σ_f |S_f|² + reg (reg =
  // inertia·max Σ_f |S_f|²), s = g/D and α_m = <g_m, s_m> / (Σ|s_m·q|²·Σ_f |S_f|² + reg·‖s_m‖²)
  // where <g_m, s_m> > 0, into dPc_ and alpha_ (probe_apply adds them). fo_spec_ holds the
  // object's spectrum; Wt_ the probe's terms.
  void probe_step_sampled(size_t stack, int nstack, double inertia) {
    const int xb = aux_blocks();
    for (int s = 0; s < nstack; ++s) {
      const int f0 = int(s * stack);
      const int nb = std::min<int>(int(stack), int(nframe_) - f0);
      const int* ord = iota_.as<int>() + f0;
      run(K_PROBE, [&] {
        fo_points(ord, nb);
        fo_object_windows(nb);
        hipLaunchKernelGGL(k_fo_probe_num<T>, dim3((n + 255) / 256, M_), dim3(256), 0, stream_,
                           gP_.as<cplx<T>>(), pden_.as<T>(), (const cplx<T>*)fo_win_.as<cplx<T>>(),
                           (const cplx<T>*)q_.as<cplx<T>>(), (const cplx<T>*)Wt_.as<cplx<T>>(),
                           (const cplx<T>*)carried_.as<cplx<T>>(), ord, nb, n, M_, int(s > 0));
      });
    }
    run(K_PROBE, [&] {
      hipLaunchKernelGGL(k_max_reduce<T>, dim3(1), dim3(MAX_THREADS), 0, stream_,
                         (const T*)pden_.as<T>(), n, pmax2_.as<T>());
      hipLaunchKernelGGL(k_cg_probe_sums<T>, dim3(xb, M_), dim3(THREADS), 0, stream_,
                         (const cplx<T>*)gP_.as<cplx<T>>(), (const T*)pden_.as<T>(),
                         (const T*)pmax2_.as<T>(), T(inertia), n, dPc_.as<cplx<T>>(),
                         cg_part_.as<double>());
      hipLaunchKernelGGL(k_fo_probe_vpow<T>, dim3(xb, M_), dim3(THREADS), 0, stream_,
                         vpow_.as<double>(), (const cplx<T>*)dPc_.as<cplx<T>>(),
                         (const cplx<T>*)q_.as<cplx<T>>(), (const T*)pden_.as<T>(), n);
      hipLaunchKernelGGL(k_cg_probe_alpha, dim3(1), dim3(1), 0, stream_,
                         (const double*)cg_part_.as<double>(), xb,
                         (const double*)vpow_.as<double>(), xb, M_, M_, alpha_.as<double>());
    });
  }
  // P += α_m·s_m (the step probe_step_sampled left in dPc_ and alpha_), then the probe terms.
  void probe_apply() {
    const int np = int(nprobe_ * n);
    run(K_PROBE, [&] {
      hipLaunchKernelGGL(k_probe_axpy<T>, dim3((np + 255) / 256), dim3(256), 0, stream_,
                         P_.as<cplx<T>>(), (const cplx<T>*)dPc_.as<cplx<T>>(),
                         (const double*)alpha_.as<double>(), n, np);
    });
    probe_terms();
  }

  // out = AᴴA·x under FullObject: x's windows at every frame's point times Σ_m |W_m|² (fo_p2_),
  // placed back by the sampling's adjoint. Overwrites fo_spec_.
  void fo_normal(const cplx<T>* x, cplx<T>* out, size_t stack, int nstack) {
    fo_object_spectrum(x);
    for (int s = 0; s < nstack; ++s) {
      const int f0 = int(s * stack);
      const int nb = std::min<int>(int(stack), int(nframe_) - f0);
      fo_points(iota_.as<int>() + f0, nb);
      fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 0, fo_win_.p, (long long)n);
      hipLaunchKernelGGL(k_fo_times_power<T>, dim3((n + 255) / 256, nb), dim3(256), 0, stream_,
                         fo_win_.as<cplx<T>>(), (const cplx<T>*)fo_p2_.as<cplx<T>>(), n);
      fo_place_spectrum(fo_win_.as<cplx<T>>(), (long long)n, nb, s > 0);
    }
    fo_add_placed<AS_PLAIN>(out, 1.0);
  }

  // The object of min ‖ψ − A·O'‖² + reg·‖O' − O‖² under FullObject
  // (ptyrs/src/engine/features.rs:726-793), A the frames' waves W_m = P_m·q times their reads:
  // kObjPcg preconditioned conjugate-gradient steps on (AᴴA + reg)·O' = Aᴴψ + reg·O from O, the
  // preconditioner D + reg with D the exact diagonal of AᴴA (Σ_m |W_m|² placed with the squared
  // sampling kernels, real part floored at 0) and reg = inertia·max D; pixels with D + reg ≤
  // 1e-5·max D are held. Every sum in FP64; the scalars stay on the device. Wt_ holds the probe
  // terms, carried_ the waves ψ; O_ is updated in place; fo_spec_ is overwritten.
  void object_solve_sampled(size_t stack, int nstack, double inertia) {
    const long long no = (long long)(nyo_ * nxo_), tp = (long long)N * nxo_;
    const size_t vplane = size_t(M_) * n;
    const int grid = int(std::min<long long>(kAbsBlocks, (no + THREADS - 1) / THREADS));
    const unsigned gno = unsigned((no + 255) / 256);
    cplx<T>* r = gO_.as<cplx<T>>();
    cplx<T>* pd = hO_.as<cplx<T>>();
    cplx<T>* ap = dOc_.as<cplx<T>>();
    T* D = fo_pcg_d_.as<T>();
    T* pinv = fo_pcg_pinv_.as<T>();
    double* sc = fo_pcg_sc_.as<double>();
    double* part = cg_part_.as<double>();
    run(K_OBJ_UPDATE, [&] {
      // D: the spectrum of Σ_m |W_m|² at the plane's origin times Σ_f K_f.
      hipLaunchKernelGGL(k_fo_power<T>, dim3((n + 255) / 256), dim3(256), 0, stream_,
                         fo_p2_.as<cplx<T>>(), (const cplx<T>*)Wt_.as<cplx<T>>(), M_, n);
      axis_pass<AL_PLAIN, AS_PLAIN>(0, false, fo_p2_.as<cplx<T>>(), fo_tmp_.p,
                                    rows_geom(n, N, tp, int(nxo_), N, N, int(nxo_)), 1);
      axis_pass<AL_PLAIN, AS_PLAIN>(1, false, fo_tmp_.as<cplx<T>>(), fo_planes_.p,
                                    cols_geom(tp, N, no, int(nyo_)), 1);
      for (int s = 0; s < nstack; ++s) {
        const int f0 = int(s * stack);
        const int nb = std::min<int>(int(stack), int(nframe_) - f0);
        fo_points(iota_.as<int>() + f0, nb);
        hipLaunchKernelGGL(k_fo_diag_acc<T>, dim3(gno), dim3(256), 0, stream_,
                           fo_acc_.as<cplx<T>>(), (const cplx<T>*)fo_planes_.as<cplx<T>>(),
                           (const double2*)fo_pts_.as<double2>(), nb, int(nyo_), int(nxo_),
                           int(s > 0));
      }
      fo_add_placed<AS_REAL_FLOOR>(D, 1.0);
      hipLaunchKernelGGL(k_max_reduce<T>, dim3(1), dim3(MAX_THREADS), 0, stream_, (const T*)D,
                         int(no), pmax2_.as<T>() + 1);
      hipLaunchKernelGGL(k_pcg_pinv<T>, dim3(gno), dim3(256), 0, stream_, pinv, (const T*)D,
                         (const T*)(pmax2_.as<T>() + 1), inertia, no, sc);
      // The right-hand side Aᴴψ into r.
      for (int s = 0; s < nstack; ++s) {
        const int f0 = int(s * stack);
        const int nb = std::min<int>(int(stack), int(nframe_) - f0);
        fo_points(iota_.as<int>() + f0, nb);
        hipLaunchKernelGGL(k_fo_chi_terms<T>, dim3((n + 255) / 256, nb), dim3(256), 0, stream_,
                           fo_dwin_.as<cplx<T>>(), (const cplx<T>*)Wt_.as<cplx<T>>(),
                           (const cplx<T>*)(carried_.as<cplx<T>>() + size_t(f0) * vplane), M_,
                           n);
        fo_place_spectrum(fo_dwin_.as<cplx<T>>(), (long long)n, nb, s > 0);
      }
      fo_add_placed<AS_PLAIN>(r, 1.0);
      // r = Aᴴψ − AᴴA·O (the regulariser's terms cancel at O), p = r/(D + reg).
      fo_normal(O_.as<cplx<T>>(), ap, stack, nstack);
      hipLaunchKernelGGL(k_pcg_start<T>, dim3(grid), dim3(THREADS), 0, stream_, r,
                         (const cplx<T>*)ap, (const T*)pinv, pd, no, part);
      hipLaunchKernelGGL(k_pcg_scalars, dim3(1), dim3(THREADS), 0, stream_,
                         (const double*)part, grid, sc, 0);
      for (int it = 0; it < kObjPcg; ++it) {
        fo_normal(pd, ap, stack, nstack);
        hipLaunchKernelGGL(k_pcg_ap<T>, dim3(grid), dim3(THREADS), 0, stream_, ap,
                           (const cplx<T>*)pd, (const double*)sc, no, part);
        hipLaunchKernelGGL(k_pcg_scalars, dim3(1), dim3(THREADS), 0, stream_,
                           (const double*)part, grid, sc, 1);
        hipLaunchKernelGGL(k_pcg_step<T>, dim3(grid), dim3(THREADS), 0, stream_,
                           O_.as<cplx<T>>(), r, (const cplx<T>*)pd, (const cplx<T>*)ap,
                           (const T*)pinv, (const double*)sc, no, part);
        if (it + 1 < kObjPcg) {
          hipLaunchKernelGGL(k_pcg_scalars, dim3(1), dim3(THREADS), 0, stream_,
                             (const double*)part, grid, sc, 2);
          hipLaunchKernelGGL(k_pcg_dir<T>, dim3(gno), dim3(256), 0, stream_, pd,
                             (const cplx<T>*)r, (const T*)pinv, (const double*)sc, no);
        }
      }
    });
  }

  // PyNX ML, Poisson conjugate gradient (ptyrs/src/engine/ptycho.rs:757-842): scale_obj_probe,
  // then per cycle the gradients, Polak–Ribière β, d = β·d − g, the Newton step γ along d and
  // O += γ·d_O, P += γ·d_P (or d = −g where γ is not finite).
  void ml(const ptyrs_ml_params& p, size_t ncycle) override {
    if (nslice_ > 1)
      throw Invalid("ML runs one object slice; only LSQML and the LLK pass run several");
    unitary("ML");
    if (nobj_ != 1) throw Invalid("the GPU ML runs one object mode");
    if (M_ > kMaxModes) throw Invalid(modes_error());
    if (nframe_ == 0) {
      cycle += ncycle;
      return;
    }
    if (full_object()) {
      ml_sampled(p, ncycle);
      return;
    }
    if (pos_on_) {
      ml_moved(p, ncycle);
      return;
    }
    const bool uo = p.update_object != 0, up = p.update_probe != 0;
    const size_t stack = std::min<size_t>(nframe_, stack_size_ ? stack_size_ : kDefaultStack);
    ensure_workspace(stack, 1);
    const size_t no = nobj_ * nyo_ * nxo_, np = nprobe_ * n;
    const int nt = ntiles();
    carried_.ensure(sizeof(cplx<T>) * nframe_ * M_ * n);
    psi3_.ensure(sizeof(cplx<T>) * 3 * stack * M_ * n);
    for (DevBuf* b : {&gO_, &hO_, &dOc_}) b->ensure(sizeof(cplx<T>) * no);
    for (DevBuf* b : {&gP_, &hP_, &dPc_, &dWt_}) b->ensure(sizeof(cplx<T>) * np);
    list_all_.ensure(sizeof(int) * size_t(nt) * nframe_);
    count_all_.ensure(sizeof(int) * size_t(nt));
    const int nstack = int((nframe_ + stack - 1) / stack);
    const int chunk = chunk_of(int(stack));
    const int nchunk = int((stack + chunk - 1) / chunk);
    pnum_part_.ensure(sizeof(cplx<T>) * size_t(nstack) * nchunk * M_ * n);
    pden_part_.ensure(sizeof(T) * size_t(nstack) * nchunk * n);
    pnum_.ensure(sizeof(cplx<T>) * M_ * n);
    pden_.ensure(sizeof(T) * n);
    pmaxpart_.ensure(sizeof(T) * ((n + THREADS - 1) / THREADS));
    abs_part_.ensure(sizeof(double) * 2 * kAbsBlocks);
    cg_part_.ensure(sizeof(double) * 2 * kAbsBlocks);
    cg_sums_.ensure(sizeof(double) * 2);
    gam_part_.ensure(sizeof(double) * 2 * nframe_ * SC::NBLK);
    gam_.ensure(sizeof(double) * 2);
    PTYRS_HIP(hipMemsetAsync(dOc_.p, 0, sizeof(cplx<T>) * no, stream_));
    PTYRS_HIP(hipMemsetAsync(dPc_.p, 0, sizeof(cplx<T>) * np, stream_));
    PTYRS_HIP(hipMemsetAsync(gO_.p, 0, sizeof(cplx<T>) * no, stream_));
    PTYRS_HIP(hipMemsetAsync(gP_.p, 0, sizeof(cplx<T>) * np, stream_));
    probe_terms();
    // ScaleObjProbe(absolute=True) from the model photons of the start.
    llk_pass(0);
    balance(1.0, hist_.as<double>() + 3);
    if (uo) {
      run(K_BIN, [&] {
        hipLaunchKernelGGL(k_bin_tiles<T>, dim3(nt), dim3(THREADS), 0, stream_,
                           corner_.as<int2>(), iota_.as<int>(), int(nframe_), N, ntx_,
                           int(nframe_), list_all_.as<int>(), count_all_.as<int>(),
                           (const T*)nullptr, M_, int(nyo_), int(nxo_), (T*)nullptr,
                           tilemax_.as<T>(),
                           (const cplx<T>*)nullptr);
      });
    }
    const int grid_o = int(std::min<size_t>(kAbsBlocks, (no + THREADS - 1) / THREADS));
    const int grid_p = int(std::min<size_t>(kAbsBlocks, (np + THREADS - 1) / THREADS));
    std::vector<std::pair<size_t, int>> due;
    for (size_t ic = 0; ic < ncycle; ++ic) {
      const bool rec = p.calc_llk > 0 && (ic % p.calc_llk == 0 || ic + 1 == ncycle);
      if (rec) {
        if (due.size() == hist_slots_) drain(due, PTYRS_ML);
        const int slot = int(due.size());
        llk_pass(slot);
        due.push_back({cycle, slot});
      }
      const std::vector<uint64_t> key = {uint64_t(ic == 0), uint64_t(uo), uint64_t(up), stack,
                                         cap_};
      epoch(ml_graph_, key, [&] {
      // Gradients: 2·∂L/∂ψ* back-projected (ptyrs/src/engine/ptycho.rs:492-561).
      for (int s = 0; s < nstack; ++s) {
        const int f0 = int(s * stack);
        const int nb = std::min<int>(int(stack), int(nframe_) - f0);
        const int* ord = iota_.as<int>() + f0;
        pass_a(ord, nb, false);
        pass_b<B_GRAD>(ord, nb);
        const int ck = chunk_of(nb);
        const int nck = (nb + ck - 1) / ck;
        run(K_PASS_C, [&] {
          hipLaunchKernelGGL((k_pass_c_proj<T, N, false>), dim3(SR::NBLK, M_, nck),
                             dim3(SR::NT), 0, stream_, psi_.as<cplx<T>>(),
                             Carried<T, false>{carried_.as<cplx<T>>()}, O_.as<cplx<T>>(),
                             int(nxo_), Wt_.as<cplx<T>>(), q_.as<cplx<T>>(),
                             corner_.as<int2>(), ord, nb, M_, ck, 0, T(0), T(0), int(up),
                             pnum_part_.as<cplx<T>>() + size_t(s) * nchunk * M_ * n,
                             pden_part_.as<T>() + size_t(s) * nchunk * n, tw_.as<cplx<T>>());
        });
        if (up && nck < nchunk) zero_partials(s, nck, nchunk);
      }
      if (uo) {
        run(K_OBJ_UPDATE, [&] {
          hipLaunchKernelGGL(k_lsqml_obj_dir<T>, dim3(nt), dim3(THREADS), 0, stream_,
                             gO_.as<cplx<T>>(), Wt_.as<cplx<T>>(), carried_.as<cplx<T>>(), N,
                             M_, corner_.as<int2>(), iota_.as<int>(), list_all_.as<int>(),
                             count_all_.as<int>(), int(nframe_), int(nyo_), int(nxo_), ntx_,
                             (const T*)nullptr, (const T*)nullptr);
        });
      }
      if (up) {
        const int nblk = (n + THREADS - 1) / THREADS;
        run(K_PROBE, [&] {
          hipLaunchKernelGGL(k_probe_reduce<T>, dim3(nblk), dim3(THREADS), 0, stream_,
                             pnum_part_.as<cplx<T>>(), pden_part_.as<T>(), nstack * nchunk, n,
                             M_, gP_.as<cplx<T>>(), pden_.as<T>(), pmaxpart_.as<T>());
        });
      }
      // Polak–Ribière β over both arrays, then the direction; h keeps this cycle's g.
      run(K_OTHER, [&] {
        if (ic > 0) {
          zero(cg_sums_.as<double>(), 2);
          if (uo) {
            hipLaunchKernelGGL(k_cg_dots<T>, dim3(grid_o), dim3(THREADS), 0, stream_,
                               gO_.as<cplx<T>>(), hO_.as<cplx<T>>(), no, cg_part_.as<double>());
            hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(THREADS), 0, stream_,
                               cg_part_.as<double>(), grid_o, cg_sums_.as<double>(), 1);
          }
          if (up) {
            hipLaunchKernelGGL(k_cg_dots<T>, dim3(grid_p), dim3(THREADS), 0, stream_,
                               gP_.as<cplx<T>>(), hP_.as<cplx<T>>(), np, cg_part_.as<double>());
            hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(THREADS), 0, stream_,
                               cg_part_.as<double>(), grid_p, cg_sums_.as<double>(), 1);
          }
        }
        if (uo)
          hipLaunchKernelGGL(k_cg_direction<T>, dim3((no + 255) / 256), dim3(256), 0, stream_,
                             dOc_.as<cplx<T>>(), hO_.as<cplx<T>>(), gO_.as<cplx<T>>(), no,
                             cg_sums_.as<double>(), int(ic == 0));
        if (up)
          hipLaunchKernelGGL(k_cg_direction<T>, dim3((np + 255) / 256), dim3(256), 0, stream_,
                             dPc_.as<cplx<T>>(), hP_.as<cplx<T>>(), gP_.as<cplx<T>>(), np,
                             cg_sums_.as<double>(), int(ic == 0));
        hipLaunchKernelGGL(k_probe_terms<T>, dim3((np + 255) / 256), dim3(256), 0, stream_,
                           dPc_.as<cplx<T>>(), q_.as<cplx<T>>(), dWt_.as<cplx<T>>(), n, int(np));
      });
      // Newton step γ = num/den along d (ptyrs/src/engine/ops.rs:197-253).
      const size_t field = stack * M_ * n;
      for (int s = 0; s < nstack; ++s) {
        const int f0 = int(s * stack);
        const int nb = std::min<int>(int(stack), int(nframe_) - f0);
        const int* ord = iota_.as<int>() + f0;
        run(K_PASS_A, [&] {
          hipLaunchKernelGGL((k_pass_a3<T, N>), dim3(S::NBLK, nb * M_), dim3(S::NT), 0, stream_,
                             O_.as<cplx<T>>(), dOc_.as<cplx<T>>(), int(nxo_), Wt_.as<cplx<T>>(),
                             dWt_.as<cplx<T>>(), corner_.as<int2>(), ord, M_, field,
                             psi3_.as<cplx<T>>(), tw_.as<cplx<T>>());
        });
        run(K_PASS_B, [&] {
          if (angular_) pass_t<false>(psi3_.as<cplx<T>>(), nb * M_, 3, field);
          const T norm = T(1) / (angular_ ? T(n) * T(n) : T(n));
          auto go = [&](auto col, auto row) {
            hipLaunchKernelGGL(angular_ ? row : col, dim3(SC::NBLK, nb), dim3(SC::NT), 0,
                               stream_, psi3_.as<cplx<T>>(), field, counts_.as<float>(), ord,
                               gam_part_.as<double>() + size_t(f0) * SC::NBLK * 2,
                               tw_.as<cplx<T>>(), norm);
          };
          with_modes([&](auto mm) {
            constexpr int MM = decltype(mm)::value;
            go(k_pass_b_gamma<T, N, MM, false>, k_pass_b_gamma<T, N, MM, true>);
          });
        });
      }
      run(K_OTHER, [&] {
        hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(THREADS), 0, stream_,
                           gam_part_.as<double>(), int(nframe_ * SC::NBLK), gam_.as<double>(), 0);
        if (uo)
          hipLaunchKernelGGL(k_cg_step<T>, dim3((no + 255) / 256), dim3(256), 0, stream_,
                             O_.as<cplx<T>>(), dOc_.as<cplx<T>>(), gO_.as<cplx<T>>(), no,
                             gam_.as<double>());
        if (up)
          hipLaunchKernelGGL(k_cg_step<T>, dim3((np + 255) / 256), dim3(256), 0, stream_,
                             P_.as<cplx<T>>(), dPc_.as<cplx<T>>(), gP_.as<cplx<T>>(), np,
                             gam_.as<double>());
        hipLaunchKernelGGL(k_probe_terms<T>, dim3((np + 255) / 256), dim3(256), 0, stream_,
                           P_.as<cplx<T>>(), q_.as<cplx<T>>(), Wt_.as<cplx<T>>(), n, int(np));
      });
      });
      ++cycle;
    }
    drain(due, PTYRS_ML);
    if (p.zero_phase_ramp) device_zero_phase_ramp();
    synchronize();
    flush_timing();
  }

  // PyNX ML under position offsets (PATCH; ptyrs/src/engine/ptycho.rs:757-842 with
  // Features::position, the reference GPU's ml under its position model,
  // ptyrs/src/engine/gpu/ml.rs:86-160): every view's wave is the probe terms moved by its
  // frame's offset (kept for the whole scan in wall_); the object gradient sums conj(W)·g over
  // the views, the probe gradient moves each frame's conj(O)·g back in Fourier space (its
  // spectra summed stack by stack); γ's a and b take the moved waves of dP·q; with refinement
  // the position sums come from the magnitude projection of the cycle's starting estimate and
  // AP's, DM's and ML's step follows the update (ptyrs/src/engine/features.rs:960-1047).
  DevBuf wall_, dwpos_, pos_all_, pstep_, pden1_, vpow_, alpha_;
  void ml_moved(const ptyrs_ml_params& p, size_t ncycle) {
    const bool uo = p.update_object != 0, up = p.update_probe != 0;
    // The offsets' derivative planes (two per view) fill one launch's grid y.
    const size_t stack = std::min<size_t>({nframe_, stack_size_ ? stack_size_ : kDefaultStack,
                                           size_t(32767 / M_)});
    ensure_workspace(stack, 1);
    const size_t no = nobj_ * nyo_ * nxo_, np = nprobe_ * n;
    const int nt = ntiles();
    const int xb = aux_blocks();
    carried_.ensure(sizeof(cplx<T>) * nframe_ * M_ * n);
    wall_.ensure(sizeof(cplx<T>) * nframe_ * M_ * n);
    pden1_.ensure(sizeof(T) * stack * n);
    if (pos_refine_) {
      dwpos_.ensure(sizeof(cplx<T>) * 2 * stack * M_ * n);
      pos_all_.ensure(sizeof(double) * 4 * nframe_ * xb);
      pstep_.ensure(sizeof(double) * 4 * nframe_);
    }
    psi3_.ensure(sizeof(cplx<T>) * 3 * stack * M_ * n);
    for (DevBuf* b : {&gO_, &hO_, &dOc_}) b->ensure(sizeof(cplx<T>) * no);
    for (DevBuf* b : {&gP_, &hP_, &dPc_, &dWt_}) b->ensure(sizeof(cplx<T>) * np);
    list_all_.ensure(sizeof(int) * size_t(nt) * nframe_);
    count_all_.ensure(sizeof(int) * size_t(nt));
    abs_part_.ensure(sizeof(double) * 2 * kAbsBlocks);
    cg_part_.ensure(sizeof(double) * 2 * kAbsBlocks);
    cg_sums_.ensure(sizeof(double) * 2);
    gam_part_.ensure(sizeof(double) * 2 * nframe_ * SC::NBLK);
    gam_.ensure(sizeof(double) * 2);
    const int nstack = int((nframe_ + stack - 1) / stack);
    PTYRS_HIP(hipMemsetAsync(dOc_.p, 0, sizeof(cplx<T>) * no, stream_));
    PTYRS_HIP(hipMemsetAsync(dPc_.p, 0, sizeof(cplx<T>) * np, stream_));
    PTYRS_HIP(hipMemsetAsync(gO_.p, 0, sizeof(cplx<T>) * no, stream_));
    PTYRS_HIP(hipMemsetAsync(gP_.p, 0, sizeof(cplx<T>) * np, stream_));
    probe_terms();
    // ScaleObjProbe(absolute=True) from the model photons of the start.
    llk_pass(0);
    balance(1.0, hist_.as<double>() + 3);
    if (uo) {
      run(K_BIN, [&] {
        hipLaunchKernelGGL(k_bin_tiles<T>, dim3(nt), dim3(THREADS), 0, stream_,
                           corner_.as<int2>(), iota_.as<int>(), int(nframe_), N, ntx_,
                           int(nframe_), list_all_.as<int>(), count_all_.as<int>(),
                           (const T*)nullptr, M_, int(nyo_), int(nxo_), (T*)nullptr,
                           tilemax_.as<T>(), (const cplx<T>*)nullptr);
      });
    }
    const int grid_o = int(std::min<size_t>(kAbsBlocks, (no + THREADS - 1) / THREADS));
    const int grid_p = int(std::min<size_t>(kAbsBlocks, (np + THREADS - 1) / THREADS));
    cplx<T>* wall = wall_.as<cplx<T>>();
    cplx<T>* g = gspec_.as<cplx<T>>();
    const size_t vplane = size_t(M_) * n;  // one frame's views
    std::vector<std::pair<size_t, int>> due;
    for (size_t ic = 0; ic < ncycle; ++ic) {
      const bool rec = p.calc_llk > 0 && (ic % p.calc_llk == 0 || ic + 1 == ncycle);
      if (rec) {
        if (due.size() == hist_slots_) drain(due, PTYRS_ML);
        const int slot = int(due.size());
        llk_pass(slot);
        due.push_back({cycle, slot});
      }
      const std::vector<uint64_t> key = {uint64_t(ic == 0), uint64_t(uo), uint64_t(up), stack,
                                         cap_, 1, uint64_t(pos_refine_)};
      epoch(ml_graph_, key, [&] {
      // Every view's moved wave at this cycle's probe.
      run(K_PROBE, [&] { probe_spectra(); });
      for (int s = 0; s < nstack; ++s) {
        const int f0 = int(s * stack);
        const int nb = std::min<int>(int(stack), int(nframe_) - f0);
        moved_probes(iota_.as<int>() + f0, nb, false, wall + size_t(f0) * vplane, nullptr, true);
      }
      // Position sums of the magnitude projection of the starting estimate.
      if (pos_refine_) {
        for (int s = 0; s < nstack; ++s) {
          const int f0 = int(s * stack);
          const int nb = std::min<int>(int(stack), int(nframe_) - f0);
          const int* ord = iota_.as<int>() + f0;
          const cplx<T>* ws = wall + size_t(f0) * vplane;
          const int K = nb * M_;
          pass_a(ord, nb, false, MixArgs<T>{MIX_NONE, {}, {}, T(0), T(0)}, ws);
          pass_b<B_AMPLITUDE>(ord, nb);
          run(K_OTHER, [&] {
            rows(true, K);
            hipLaunchKernelGGL(k_ramp_deriv_planes<T>, dim3((n + 255) / 256, K), dim3(256), 0,
                               stream_, dwpos_.as<cplx<T>>(),
                               (const cplx<T>*)spec_.as<cplx<T>>(),
                               (const double2*)offs_.as<double2>(), ord, N, M_,
                               1.0 / double(n));
            hipLaunchKernelGGL((k_cols<T, N, true>), dim3(SC::NBLK, 2 * K), dim3(SC::NT), 0,
                               stream_, dwpos_.as<cplx<T>>(), tw_.as<cplx<T>>());
            hipLaunchKernelGGL((k_rows<T, N, true>), dim3(S::NBLK, 2 * K), dim3(S::NT), 0,
                               stream_, dwpos_.as<cplx<T>>(), tw_.as<cplx<T>>());
            hipLaunchKernelGGL(k_pos_sums_proj<T>, dim3(xb, nb), dim3(THREADS), 0, stream_,
                               (const cplx<T>*)psi_.as<cplx<T>>(), ws,
                               (const cplx<T>*)O_.as<cplx<T>>(), int(nxo_), corner_.as<int2>(),
                               ord, (const cplx<T>*)dwpos_.as<cplx<T>>(), N, M_,
                               pos_all_.as<double>() + size_t(f0) * xb * 4);
          });
        }
      }
      // Gradients: 2·∂L/∂ψ* through the moved waves.
      for (int s = 0; s < nstack; ++s) {
        const int f0 = int(s * stack);
        const int nb = std::min<int>(int(stack), int(nframe_) - f0);
        const int* ord = iota_.as<int>() + f0;
        const cplx<T>* ws = wall + size_t(f0) * vplane;
        const int K = nb * M_;
        pass_a(ord, nb, false, MixArgs<T>{MIX_NONE, {}, {}, T(0), T(0)}, ws);
        pass_b<B_GRAD>(ord, nb);
        // One slot per chunk: each frame's conj(q·O)·g of its own.
        run(K_PASS_C, [&] {
          hipLaunchKernelGGL((k_pass_c_proj<T, N, false>), dim3(SR::NBLK, M_, nb),
                             dim3(SR::NT), 0, stream_, psi_.as<cplx<T>>(),
                             Carried<T, false>{carried_.as<cplx<T>>()}, O_.as<cplx<T>>(),
                             int(nxo_), Wt_.as<cplx<T>>(), q_.as<cplx<T>>(),
                             corner_.as<int2>(), ord, nb, M_, 1, 0, T(0), T(0), int(up), g,
                             pden1_.as<T>(), tw_.as<cplx<T>>());
        });
        if (up) {
          run(K_PROBE, [&] {
            hipLaunchKernelGGL((k_times_q<T, false>), dim3((n + 255) / 256, K), dim3(256), 0,
                               stream_, g, (const cplx<T>*)q_.as<cplx<T>>(), n);
            hipLaunchKernelGGL((k_rows<T, N, false>), dim3(S::NBLK, K), dim3(S::NT), 0, stream_,
                               g, tw_.as<cplx<T>>());
            hipLaunchKernelGGL((k_cols<T, N, false>), dim3(SC::NBLK, K), dim3(SC::NT), 0,
                               stream_, g, tw_.as<cplx<T>>());
            hipLaunchKernelGGL(k_unramp_mean<T>, dim3((n + 255) / 256, M_), dim3(256), 0,
                               stream_, gP_.as<cplx<T>>(), (const cplx<T>*)g,
                               (const double2*)offs_.as<double2>(), ord, nb, N, M_,
                               1.0 / double(n), int(s > 0));
          });
        }
      }
      if (up) {
        run(K_PROBE, [&] {
          hipLaunchKernelGGL((k_cols<T, N, true>), dim3(SC::NBLK, M_), dim3(SC::NT), 0, stream_,
                             gP_.as<cplx<T>>(), tw_.as<cplx<T>>());
          hipLaunchKernelGGL((k_rows<T, N, true>), dim3(S::NBLK, M_), dim3(S::NT), 0, stream_,
                             gP_.as<cplx<T>>(), tw_.as<cplx<T>>());
          hipLaunchKernelGGL((k_times_q<T, true>), dim3((n + 255) / 256, M_), dim3(256), 0,
                             stream_, gP_.as<cplx<T>>(), (const cplx<T>*)q_.as<cplx<T>>(), n);
        });
      }
      if (uo) {
        run(K_OBJ_UPDATE, [&] {
          hipLaunchKernelGGL(k_lsqml_obj_dir<T>, dim3(nt), dim3(THREADS), 0, stream_,
                             gO_.as<cplx<T>>(), (const cplx<T>*)wall, carried_.as<cplx<T>>(),
                             N, M_, corner_.as<int2>(), iota_.as<int>(), list_all_.as<int>(),
                             count_all_.as<int>(), int(nframe_), int(nyo_), int(nxo_), ntx_,
                             (const T*)nullptr, (const T*)nullptr, 1, (const cplx<T>*)nullptr,
                             0, 0);
        });
      }
      // Polak–Ribière β over both arrays, then the direction; h keeps this cycle's g.
      run(K_OTHER, [&] {
        if (ic > 0) {
          zero(cg_sums_.as<double>(), 2);
          if (uo) {
            hipLaunchKernelGGL(k_cg_dots<T>, dim3(grid_o), dim3(THREADS), 0, stream_,
                               gO_.as<cplx<T>>(), hO_.as<cplx<T>>(), no, cg_part_.as<double>());
            hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(THREADS), 0, stream_,
                               cg_part_.as<double>(), grid_o, cg_sums_.as<double>(), 1);
          }
          if (up) {
            hipLaunchKernelGGL(k_cg_dots<T>, dim3(grid_p), dim3(THREADS), 0, stream_,
                               gP_.as<cplx<T>>(), hP_.as<cplx<T>>(), np, cg_part_.as<double>());
            hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(THREADS), 0, stream_,
                               cg_part_.as<double>(), grid_p, cg_sums_.as<double>(), 1);
          }
        }
        if (uo)
          hipLaunchKernelGGL(k_cg_direction<T>, dim3((no + 255) / 256), dim3(256), 0, stream_,
                             dOc_.as<cplx<T>>(), hO_.as<cplx<T>>(), gO_.as<cplx<T>>(), no,
                             cg_sums_.as<double>(), int(ic == 0));
        if (up)
          hipLaunchKernelGGL(k_cg_direction<T>, dim3((np + 255) / 256), dim3(256), 0, stream_,
                             dPc_.as<cplx<T>>(), hP_.as<cplx<T>>(), gP_.as<cplx<T>>(), np,
                             cg_sums_.as<double>(), int(ic == 0));
        hipLaunchKernelGGL(k_probe_terms<T>, dim3((np + 255) / 256), dim3(256), 0, stream_,
                           dPc_.as<cplx<T>>(), q_.as<cplx<T>>(), dWt_.as<cplx<T>>(), n, int(np));
      });
      // Newton step γ along d with the moved waves of P·q and of dP·q.
      run(K_PROBE, [&] { probe_spectra(dWt_.as<cplx<T>>()); });
      const size_t field = stack * M_ * n;
      for (int s = 0; s < nstack; ++s) {
        const int f0 = int(s * stack);
        const int nb = std::min<int>(int(stack), int(nframe_) - f0);
        const int* ord = iota_.as<int>() + f0;
        moved_probes(ord, nb, false, inc0_.as<cplx<T>>(), nullptr, true);
        run(K_PASS_A, [&] {
          hipLaunchKernelGGL((k_pass_a3<T, N>), dim3(S::NBLK, nb * M_), dim3(S::NT), 0, stream_,
                             O_.as<cplx<T>>(), dOc_.as<cplx<T>>(), int(nxo_),
                             (const cplx<T>*)(wall + size_t(f0) * vplane),
                             (const cplx<T>*)inc0_.as<cplx<T>>(), corner_.as<int2>(), ord, M_,
                             field, psi3_.as<cplx<T>>(), tw_.as<cplx<T>>(), 1);
        });
        run(K_PASS_B, [&] {
          if (angular_) pass_t<false>(psi3_.as<cplx<T>>(), nb * M_, 3, field);
          const T norm = T(1) / (angular_ ? T(n) * T(n) : T(n));
          auto go = [&](auto col, auto row) {
            hipLaunchKernelGGL(angular_ ? row : col, dim3(SC::NBLK, nb), dim3(SC::NT), 0,
                               stream_, psi3_.as<cplx<T>>(), field, counts_.as<float>(), ord,
                               gam_part_.as<double>() + size_t(f0) * SC::NBLK * 2,
                               tw_.as<cplx<T>>(), norm);
          };
          with_modes([&](auto mm) {
            constexpr int MM = decltype(mm)::value;
            go(k_pass_b_gamma<T, N, MM, false>, k_pass_b_gamma<T, N, MM, true>);
          });
        });
      }
      run(K_OTHER, [&] {
        hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(THREADS), 0, stream_,
                           gam_part_.as<double>(), int(nframe_ * SC::NBLK), gam_.as<double>(), 0);
        if (uo)
          hipLaunchKernelGGL(k_cg_step<T>, dim3((no + 255) / 256), dim3(256), 0, stream_,
                             O_.as<cplx<T>>(), dOc_.as<cplx<T>>(), gO_.as<cplx<T>>(), no,
                             gam_.as<double>());
        if (up)
          hipLaunchKernelGGL(k_cg_step<T>, dim3((np + 255) / 256), dim3(256), 0, stream_,
                             P_.as<cplx<T>>(), dPc_.as<cplx<T>>(), gP_.as<cplx<T>>(), np,
                             gam_.as<double>());
        hipLaunchKernelGGL(k_probe_terms<T>, dim3((np + 255) / 256), dim3(256), 0, stream_,
                           P_.as<cplx<T>>(), q_.as<cplx<T>>(), Wt_.as<cplx<T>>(), n, int(np));
        if (pos_refine_)
          hipLaunchKernelGGL(k_pos_step_all, dim3(1), dim3(1024), 0, stream_,
                             (const double*)pos_all_.as<double>(), xb, int(nframe_),
                             offs_.as<double2>(), pstep_.as<double>());
      });
      });
      ++cycle;
    }
    drain(due, PTYRS_ML);
    if (p.zero_phase_ramp) device_zero_phase_ramp();
    synchronize();
    flush_timing();
  }

  // PyNX ML under FullObject sampling (ptyrs/src/engine/ptycho.rs:757-842 with Shifted's
  // sampler; the reference GPU's ml under its position model, ptyrs/src/engine/gpu/ml.rs:1-160):
  // the probe terms are not moved; every frame reads its object window sampled at its corner
  // plus offset (the tall windows of a stack, from the object's spectrum once per cycle); the
  // object gradient places each slot's Σ conj(W)·g by the sampling's adjoint, the probe
  // gradient sums conj(S·q)·g; γ's a and b read the direction's sampled windows; with
  // refinement the position sums take the windows' derivatives against W.
  DevBuf fo_dd_, fo_dspec_;
  void ml_sampled(const ptyrs_ml_params& p, size_t ncycle) {
    const bool uo = p.update_object != 0, up = p.update_probe != 0;
    const size_t stack = std::min<size_t>(nframe_, stack_size_ ? stack_size_ : kDefaultStack);
    ensure_workspace(stack, 1);
    fo_ensure(stack);
    const size_t no = nyo_ * nxo_, np = nprobe_ * n;
    const int xb = aux_blocks();
    if (pos_refine_) {
      fo_dd_.ensure(sizeof(cplx<T>) * 2 * stack * n);
      pos_all_.ensure(sizeof(double) * 4 * nframe_ * xb);
      pstep_.ensure(sizeof(double) * 4 * nframe_);
    }
    fo_dspec_.ensure(sizeof(cplx<T>) * no);
    psi3_.ensure(sizeof(cplx<T>) * 3 * stack * M_ * n);
    for (DevBuf* b : {&gO_, &hO_, &dOc_}) b->ensure(sizeof(cplx<T>) * no);
    for (DevBuf* b : {&gP_, &hP_, &dPc_, &dWt_}) b->ensure(sizeof(cplx<T>) * np);
    cg_part_.ensure(sizeof(double) * 2 * kAbsBlocks);
    cg_sums_.ensure(sizeof(double) * 2);
    gam_part_.ensure(sizeof(double) * 2 * nframe_ * SC::NBLK);
    gam_.ensure(sizeof(double) * 2);
    abs_part_.ensure(sizeof(double) * 2 * kAbsBlocks);
    const int nstack = int((nframe_ + stack - 1) / stack);
    PTYRS_HIP(hipMemsetAsync(dOc_.p, 0, sizeof(cplx<T>) * no, stream_));
    PTYRS_HIP(hipMemsetAsync(dPc_.p, 0, sizeof(cplx<T>) * np, stream_));
    PTYRS_HIP(hipMemsetAsync(gO_.p, 0, sizeof(cplx<T>) * no, stream_));
    PTYRS_HIP(hipMemsetAsync(gP_.p, 0, sizeof(cplx<T>) * np, stream_));
    probe_terms();
    // ScaleObjProbe(absolute=True) from the model photons of the start.
    llk_pass(0);
    balance(1.0, hist_.as<double>() + 3);
    const int grid_o = int(std::min<size_t>(kAbsBlocks, (no + THREADS - 1) / THREADS));
    const int grid_p = int(std::min<size_t>(kAbsBlocks, (np + THREADS - 1) / THREADS));
    const MixArgs<T> none{MIX_NONE, {}, {}, T(0), T(0)};
    cplx<T>* win = fo_win_.as<cplx<T>>();
    std::vector<std::pair<size_t, int>> due;
    for (size_t ic = 0; ic < ncycle; ++ic) {
      const bool rec = p.calc_llk > 0 && (ic % p.calc_llk == 0 || ic + 1 == ncycle);
      if (rec) {
        if (due.size() == hist_slots_) drain(due, PTYRS_ML);
        const int slot = int(due.size());
        llk_pass(slot);
        due.push_back({cycle, slot});
      }
      run(K_BIN, [&] { fo_object_spectrum(O_.as<cplx<T>>()); });
      // Position sums of the magnitude projection of the starting estimate.
      if (pos_refine_) {
        for (int s = 0; s < nstack; ++s) {
          const int f0 = int(s * stack);
          const int nb = std::min<int>(int(stack), int(nframe_) - f0);
          const int* ord = iota_.as<int>() + f0;
          const int K = nb * M_;
          run(K_BIN, [&] {
            fo_points(ord, nb);
            fo_object_windows(nb);
          });
          pass_a(ord, nb, false, none, nullptr, win, N, fo_tall_.as<int2>());
          pass_b<B_AMPLITUDE>(ord, nb);
          run(K_OTHER, [&] {
            rows(true, K);
            fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 1, fo_dd_.p, 2ll * n);
            fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 2, fo_dd_.as<cplx<T>>() + n,
                                 2ll * n);
            hipLaunchKernelGGL(k_pos_sums_fo<T>, dim3(xb, nb), dim3(THREADS), 0, stream_,
                               (const cplx<T>*)psi_.as<cplx<T>>(),
                               (const cplx<T>*)Wt_.as<cplx<T>>(), (const cplx<T>*)win,
                               (const cplx<T>*)fo_dd_.as<cplx<T>>(), n, M_,
                               pos_all_.as<double>() + size_t(f0) * xb * 4);
          });
        }
      }
      // Gradients: 2·∂L/∂ψ* through the sampled windows.
      for (int s = 0; s < nstack; ++s) {
        const int f0 = int(s * stack);
        const int nb = std::min<int>(int(stack), int(nframe_) - f0);
        const int* ord = iota_.as<int>() + f0;
        const int K = nb * M_;
        run(K_BIN, [&] {
          fo_points(ord, nb);
          fo_object_windows(nb);
        });
        pass_a(ord, nb, false, none, nullptr, win, N, fo_tall_.as<int2>());
        pass_b<B_GRAD>(ord, nb);
        const cplx<T>* g = psi_.as<cplx<T>>();
        run(K_PASS_C, [&] { rows(true, K); });
        if (up)
          run(K_PROBE, [&] {
            hipLaunchKernelGGL(k_ms_probe_dir<T>, dim3((n + 255) / 256, M_), dim3(256), 0,
                               stream_, gP_.as<cplx<T>>(), (const cplx<T>*)win, N,
                               (const cplx<T>*)q_.as<cplx<T>>(), g, fo_tall_.as<int2>(), ord, nb,
                               N, M_, T(1), int(s > 0));
          });
        if (uo)
          run(K_OBJ_UPDATE, [&] {
            hipLaunchKernelGGL(k_fo_chi_terms<T>, dim3((n + 255) / 256, nb), dim3(256), 0,
                               stream_, fo_dwin_.as<cplx<T>>(), (const cplx<T>*)Wt_.as<cplx<T>>(),
                               g, M_, n);
            fo_place_spectrum(fo_dwin_.as<cplx<T>>(), (long long)n, nb, s > 0);
          });
      }
      if (uo) run(K_OBJ_UPDATE, [&] { fo_add_placed<AS_PLAIN>(gO_.p, 1.0); });
      // Polak–Ribière β over both arrays, then the direction; h keeps this cycle's g.
      run(K_OTHER, [&] {
        if (ic > 0) {
          zero(cg_sums_.as<double>(), 2);
          if (uo) {
            hipLaunchKernelGGL(k_cg_dots<T>, dim3(grid_o), dim3(THREADS), 0, stream_,
                               gO_.as<cplx<T>>(), hO_.as<cplx<T>>(), no, cg_part_.as<double>());
            hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(THREADS), 0, stream_,
                               cg_part_.as<double>(), grid_o, cg_sums_.as<double>(), 1);
          }
          if (up) {
            hipLaunchKernelGGL(k_cg_dots<T>, dim3(grid_p), dim3(THREADS), 0, stream_,
                               gP_.as<cplx<T>>(), hP_.as<cplx<T>>(), np, cg_part_.as<double>());
            hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(THREADS), 0, stream_,
                               cg_part_.as<double>(), grid_p, cg_sums_.as<double>(), 1);
          }
        }
        if (uo)
          hipLaunchKernelGGL(k_cg_direction<T>, dim3((no + 255) / 256), dim3(256), 0, stream_,
                             dOc_.as<cplx<T>>(), hO_.as<cplx<T>>(), gO_.as<cplx<T>>(), no,
                             cg_sums_.as<double>(), int(ic == 0));
        if (up)
          hipLaunchKernelGGL(k_cg_direction<T>, dim3((np + 255) / 256), dim3(256), 0, stream_,
                             dPc_.as<cplx<T>>(), hP_.as<cplx<T>>(), gP_.as<cplx<T>>(), np,
                             cg_sums_.as<double>(), int(ic == 0));
        hipLaunchKernelGGL(k_probe_terms<T>, dim3((np + 255) / 256), dim3(256), 0, stream_,
                           dPc_.as<cplx<T>>(), q_.as<cplx<T>>(), dWt_.as<cplx<T>>(), n, int(np));
      });
      // Newton step γ along d: the three fields read S(O) and S(dO).
      run(K_OTHER, [&] { fo_spectrum_to(dOc_.as<cplx<T>>(), fo_dspec_.as<cplx<T>>()); });
      const size_t field = stack * M_ * n;
      for (int s = 0; s < nstack; ++s) {
        const int f0 = int(s * stack);
        const int nb = std::min<int>(int(stack), int(nframe_) - f0);
        const int* ord = iota_.as<int>() + f0;
        run(K_BIN, [&] {
          fo_points(ord, nb);
          fo_object_windows(nb);
          fo_windows<AS_PLAIN>(fo_dspec_.as<cplx<T>>(), nb, 0, fo_dwin_.p, (long long)n);
        });
        run(K_PASS_A, [&] {
          hipLaunchKernelGGL((k_pass_a3<T, N>), dim3(S::NBLK, nb * M_), dim3(S::NT), 0, stream_,
                             (const cplx<T>*)win, (const cplx<T>*)fo_dwin_.as<cplx<T>>(), N,
                             (const cplx<T>*)Wt_.as<cplx<T>>(),
                             (const cplx<T>*)dWt_.as<cplx<T>>(), fo_tall_.as<int2>(), ord, M_,
                             field, psi3_.as<cplx<T>>(), tw_.as<cplx<T>>(), 0);
        });
        run(K_PASS_B, [&] {
          if (angular_) pass_t<false>(psi3_.as<cplx<T>>(), nb * M_, 3, field);
          const T norm = T(1) / (angular_ ? T(n) * T(n) : T(n));
          auto go = [&](auto col, auto row) {
            hipLaunchKernelGGL(angular_ ? row : col, dim3(SC::NBLK, nb), dim3(SC::NT), 0,
                               stream_, psi3_.as<cplx<T>>(), field, counts_.as<float>(), ord,
                               gam_part_.as<double>() + size_t(f0) * SC::NBLK * 2,
                               tw_.as<cplx<T>>(), norm);
          };
          with_modes([&](auto mm) {
            constexpr int MM = decltype(mm)::value;
            go(k_pass_b_gamma<T, N, MM, false>, k_pass_b_gamma<T, N, MM, true>);
          });
        });
      }
      run(K_OTHER, [&] {
        hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(THREADS), 0, stream_,
                           gam_part_.as<double>(), int(nframe_ * SC::NBLK), gam_.as<double>(), 0);
        if (uo)
          hipLaunchKernelGGL(k_cg_step<T>, dim3((no + 255) / 256), dim3(256), 0, stream_,
                             O_.as<cplx<T>>(), dOc_.as<cplx<T>>(), gO_.as<cplx<T>>(), no,
                             gam_.as<double>());
        if (up)
          hipLaunchKernelGGL(k_cg_step<T>, dim3((np + 255) / 256), dim3(256), 0, stream_,
                             P_.as<cplx<T>>(), dPc_.as<cplx<T>>(), gP_.as<cplx<T>>(), np,
                             gam_.as<double>());
        hipLaunchKernelGGL(k_probe_terms<T>, dim3((np + 255) / 256), dim3(256), 0, stream_,
                           P_.as<cplx<T>>(), q_.as<cplx<T>>(), Wt_.as<cplx<T>>(), n, int(np));
        if (pos_refine_)
          hipLaunchKernelGGL(k_pos_step_all, dim3(1), dim3(1024), 0, stream_,
                             (const double*)pos_all_.as<double>(), xb, int(nframe_),
                             offs_.as<double2>(), pstep_.as<double>());
      });
      ++cycle;
    }
    drain(due, PTYRS_ML);
    if (p.zero_phase_ramp) device_zero_phase_ramp();
    synchronize();
    flush_timing();
  }

 private:
  static constexpr size_t kDefaultStack = 128;
  // Views per frame (probe modes × object modes) the passes take: each is compiled per count.
  static constexpr int kMaxModes = 8;
  static std::string modes_error() {
    return "the GPU engine runs at most " + std::to_string(kMaxModes) +
           " modes (probe modes × object modes) per frame";
  }
  // Calls f with the views per frame as a compile-time constant.
  template <typename F>
  void with_modes(F&& f) {
    switch (M_) {
      case 1: f(std::integral_constant<int, 1>{}); break;
      case 2: f(std::integral_constant<int, 2>{}); break;
      case 3: f(std::integral_constant<int, 3>{}); break;
      case 4: f(std::integral_constant<int, 4>{}); break;
      case 5: f(std::integral_constant<int, 5>{}); break;
      case 6: f(std::integral_constant<int, 6>{}); break;
      case 7: f(std::integral_constant<int, 7>{}); break;
      case 8: f(std::integral_constant<int, 8>{}); break;
      default: throw Invalid(modes_error());
    }
  }
  int M_ = 1;
  double scale_ = 1.0, max_frame_total_ = 0.0, total_ = 0.0;
  // Angular-spectrum exit: the transfer function T on the unshifted grid, and whether it drops
  // an evanescent band (then its adjoint does not invert it).
  bool angular_ = false, banded_ = false;
  DevBuf tf_;

  // The algorithms that take the exit adjoint for its inverse refuse a banded transfer
  // function; LSQML and the LLK pass run it (ptyrs/src/engine/gpu/device.rs:3087-3105).
  void unitary(const char* what) const {
    if (banded_)
      throw Invalid(std::string(what) +
                    " needs a unitary propagator; this one drops an evanescent band: only "
                    "LSQML and the LLK pass run it");
  }
  // ... and one object slice (ptyrs/src/engine/gpu/device.rs:3100-3105).
  void one_unitary(const char* what) const {
    if (nslice_ > 1)
      throw Invalid(std::string(what) +
                    " runs one object slice; only LSQML and the LLK pass run several");
    if (pos_on_)
      throw Invalid(std::string(what) +
                    " with position offsets is not run on this GPU engine; LSQML and the LLK "
                    "pass model them");
    unitary(what);
  }

  // ---- Orthogonal probe relaxation (ptyrs/src/engine/opr.rs) ---------------------------------
  // During an LSQML call with a ProbeVariation: k eigenmodes of probe mode 0 (opr_U_, [k][n]),
  // the weights of every frame (opr_w_, [nframe·k] FP64) and the main mode's weights (opr_c_,
  // [nframe] FP64) when modelled; the stacks read each frame's own planes (stack_waves).
  bool opr_on_ = false, opr_c_on_ = false, opr_refine_w_ = false, opr_refine_c_ = false;
  int opr_k_ = 0;
  DevBuf opr_U_, opr_w_, opr_c_, opr_d_, opr_m_, opr_vv_, opr_r_, opr_e_, opr_part_, opr_pe_,
      opr_pu_, opr_puu_, opr_sc_;
  static constexpr int kOprBlocks = 64;
  int opr_blocks() const {
    return std::max(1, std::min(kOprBlocks, (n + kOprThreads - 1) / kOprThreads));
  }

  void opr_ensure(size_t cap) {
    inc0_.ensure(sizeof(cplx<T>) * cap * M_ * n);
    opr_d_.ensure(sizeof(cplx<T>) * cap * n);
    opr_m_.ensure(sizeof(cplx<T>) * cap * n);
    opr_vv_.ensure(sizeof(T) * cap * n);
    opr_r_.ensure(sizeof(cplx<T>) * cap * n);
    opr_e_.ensure(sizeof(cplx<T>) * n);
    opr_part_.ensure(sizeof(double) * 2 * cap * aux_blocks());
    opr_pe_.ensure(sizeof(double) * kOprBlocks);
    opr_pu_.ensure(sizeof(double) * kOprBlocks);
    opr_puu_.ensure(sizeof(double) * kOprBlocks);
    opr_sc_.ensure(sizeof(double) * cap);
  }

  void lsqml_opr(const ptyrs_lsqml_params& p, size_t nepoch, size_t calc_llk,
                 ptyrs_opr& var) override {
    if (var.k == 0 || !var.eigenmodes || !var.weights)
      throw Invalid("OPR needs k >= 1 eigenmodes and their weights");
    if (var.k > 64) throw Invalid("OPR takes at most 64 eigenmodes");
    const size_t k = var.k;
    for (size_t i = 0; i < 2 * k * n; ++i)
      if (!std::isfinite(var.eigenmodes[i])) throw Invalid("eigenmodes must be finite");
    for (size_t i = 0; i < k * nframe_; ++i)
      if (!std::isfinite(var.weights[i])) throw Invalid("OPR weights must be finite");
    for (size_t i = 0; var.intensity && i < nframe_; ++i)
      if (!std::isfinite(var.intensity[i])) throw Invalid("main-mode weights must be finite");
    opr_U_.ensure(sizeof(cplx<T>) * k * n);
    upload_complex(opr_U_, var.eigenmodes, k * n);
    opr_w_.ensure(sizeof(double) * std::max<size_t>(1, k * nframe_));
    opr_c_.ensure(sizeof(double) * std::max<size_t>(1, nframe_));
    if (nframe_) {
      PTYRS_HIP(hipMemcpy(opr_w_.p, var.weights, sizeof(double) * k * nframe_,
                          hipMemcpyHostToDevice));
      if (var.intensity)
        PTYRS_HIP(hipMemcpy(opr_c_.p, var.intensity, sizeof(double) * nframe_,
                            hipMemcpyHostToDevice));
    }
    opr_k_ = int(k);
    opr_c_on_ = var.intensity != nullptr;
    opr_refine_w_ = var.refine_weights != 0;
    opr_refine_c_ = var.refine_intensity != 0;
    opr_on_ = true;
    try {
      lsqml(p, nepoch, calc_llk);
    } catch (...) {
      opr_on_ = false;
      throw;
    }
    opr_on_ = false;
    if (nepoch == 0 || nframe_ == 0) return;
    download_complex(opr_U_, var.eigenmodes, k * n);
    PTYRS_HIP(hipMemcpy(var.weights, opr_w_.p, sizeof(double) * k * nframe_,
                        hipMemcpyDeviceToHost));
    if (var.intensity)
      PTYRS_HIP(hipMemcpy(var.intensity, opr_c_.p, sizeof(double) * nframe_,
                          hipMemcpyDeviceToHost));
  }

  // OPRModeWeights.update_opr_probe_modes_and_weights on a batch (ptyrs/src/engine/opr.rs:
  // 130-237): from slice 0's χ (psi_ after the back-propagation) and the batch-start object and
  // probe, each slot's mode-0 direction d_b = conj(q·O[win])·χ_b0 and |O[win]|²; the main-mode
  // weights' step; then, with two slots or more, per eigenmode in order its move along the
  // residues' principal direction, its renormalisation, its weights' step and the residues'
  // projection off it.
  // Under FullObject the windows are the batch's sampled reads (fo_win_ at the tall corners)
  // and the directions are not moved.
  void opr_step(const int* ord, int nb) {
    const int xb = aux_blocks(), eb = opr_blocks();
    const cplx<T>* q0 = nslice_ == 1 ? q_.as<cplx<T>>() : nullptr;
    const int k = opr_k_;
    const bool fo = full_object();
    run(K_PROBE, [&] {
      hipLaunchKernelGGL(k_opr_frame_dir<T>, dim3((n + 255) / 256, nb), dim3(256), 0, stream_,
                         opr_d_.as<cplx<T>>(), opr_vv_.as<T>(),
                         fo ? (const cplx<T>*)fo_win_.as<cplx<T>>()
                            : (const cplx<T>*)slice_obj(0),
                         fo ? N : int(nxo_), q0, (const cplx<T>*)psi_.as<cplx<T>>(),
                         fo ? fo_tall_.as<int2>() : corner_.as<int2>(), ord, N, M_);
      if (opr_c_on_ && opr_refine_c_) {
        hipLaunchKernelGGL((k_opr_sums<T, 0>), dim3(xb, nb), dim3(kOprThreads), 0, stream_,
                           (const cplx<T>*)P_.as<cplx<T>>(), (const cplx<T>*)opr_d_.as<cplx<T>>(),
                           (const T*)opr_vv_.as<T>(), n, opr_part_.as<double>());
        hipLaunchKernelGGL(k_opr_intensity, dim3(1), dim3(1), 0, stream_,
                           (const double*)opr_part_.as<double>(), xb, nb, ord,
                           opr_c_.as<double>());
      }
      if (nb < 2) return;
      // The residues of the directions taken back through the frames' moves (the directions
      // themselves without offsets).
      const cplx<T>* moved =
          pos_on_ && !fo ? opr_m_.as<cplx<T>>() : opr_d_.as<cplx<T>>();
      hipLaunchKernelGGL(k_opr_residues<T>, dim3((n + 255) / 256), dim3(256), 0, stream_,
                         opr_r_.as<cplx<T>>(), moved, nb, n);
      const double a_u = std::min(0.1, double(nb) / double(nframe_));
      for (int i = 0; i < k; ++i) {
        cplx<T>* U = opr_U_.as<cplx<T>>() + size_t(i) * n;
        hipLaunchKernelGGL((k_opr_sums<T, 1>), dim3(xb, nb), dim3(kOprThreads), 0, stream_,
                           (const cplx<T>*)U, (const cplx<T>*)opr_r_.as<cplx<T>>(),
                           (const T*)nullptr, n, opr_part_.as<double>());
        hipLaunchKernelGGL(k_opr_p, dim3(1), dim3(1), 0, stream_,
                           (const double*)opr_part_.as<double>(), xb, nb, ord,
                           (const double*)opr_w_.as<double>(), k, i, double(n),
                           opr_sc_.as<double>());
        hipLaunchKernelGGL(k_opr_e<T>, dim3(eb), dim3(kOprThreads), 0, stream_,
                           opr_e_.as<cplx<T>>(), (const cplx<T>*)opr_r_.as<cplx<T>>(),
                           (const double*)opr_sc_.as<double>(), nb, n, opr_pe_.as<double>());
        hipLaunchKernelGGL(k_opr_axpy_norm<T>, dim3(eb), dim3(kOprThreads), 0, stream_, U,
                           (const cplx<T>*)opr_e_.as<cplx<T>>(),
                           (const double*)opr_pe_.as<double>(), eb, a_u, n,
                           opr_pu_.as<double>());
        hipLaunchKernelGGL(k_opr_normalise<T>, dim3(eb), dim3(kOprThreads), 0, stream_, U,
                           (const double*)opr_pu_.as<double>(), eb, n, opr_puu_.as<double>());
        if (opr_refine_w_) {
          hipLaunchKernelGGL((k_opr_sums<T, 0>), dim3(xb, nb), dim3(kOprThreads), 0, stream_,
                             (const cplx<T>*)U, (const cplx<T>*)opr_d_.as<cplx<T>>(),
                             (const T*)opr_vv_.as<T>(), n, opr_part_.as<double>());
          hipLaunchKernelGGL(k_opr_weights, dim3(1), dim3(1), 0, stream_,
                             (const double*)opr_part_.as<double>(), xb, nb, ord,
                             opr_w_.as<double>(), k, i);
        }
        if (i + 1 < k) {
          hipLaunchKernelGGL((k_opr_sums<T, 2>), dim3(xb, nb), dim3(kOprThreads), 0, stream_,
                             (const cplx<T>*)U, (const cplx<T>*)opr_r_.as<cplx<T>>(),
                             (const T*)nullptr, n, opr_part_.as<double>());
          hipLaunchKernelGGL(k_opr_project_out<T>, dim3((n + 255) / 256, nb), dim3(256), 0,
                             stream_, opr_r_.as<cplx<T>>(), (const cplx<T>*)U,
                             (const double*)opr_part_.as<double>(), xb,
                             (const double*)opr_puu_.as<double>(), eb, n);
        }
      }
    });
  }

  // ---- FullObject sampling (ptyrs/src/engine/sampling.rs:79-209) -----------------------------
  // Under offsets and FullObject the probe stays put and each slot reads its window sampled
  // at its frame's corner plus offset: the windows of a batch are laid out as one tall plane
  // ([slot·N + y][x], fo_win_), read by the stack passes at the corners fo_tall_ ((slot·N, 0)
  // per frame); the overlaps D_m are the same tall layout (fo_D_, [m][cap·N][N]). The object's
  // transforms run at its own sides (AxisPlan, any length up to kAxisMax; fo_plan_[0] rows of
  // nxo, [1] columns of nyo).
  AxisPlan fo_plan_[2]{};
  DevBuf fo_tw_[2], fo_spec_, fo_tmp_, fo_planes_, fo_win_, fo_dwin_, fo_F_, fo_acc_, fo_E_,
      fo_pts_, fo_tall_, fo_D_, fo_part_, fo_p2_, fo_pcg_d_, fo_pcg_pinv_, fo_pcg_sc_;
  size_t fo_cap_ = 0;
  // The object solve's conjugate-gradient steps under FullObject (ptyrs/src/engine/sampling.rs:214).
  static constexpr int kObjPcg = 10;
  static constexpr int kFoBlocks = 256;
  bool full_object() const { return pos_on_ && sampling == PTYRS_FULL_OBJECT; }

  static AxisPlan axis_plan(int L) {
    AxisPlan pl{};
    pl.L = L;
    int x = L;
    auto push = [&](int r) {
      if (pl.nst >= kAxisMaxStages) throw Invalid("object side has too many prime factors");
      pl.radix[pl.nst++] = r;
    };
    while (x % 4 == 0) push(4), x /= 4;
    while (x % 2 == 0) push(2), x /= 2;
    for (int r = 3; x > 1; r += 2)
      while (x % r == 0) push(r), x /= r;
    return pl;
  }
  // The longest object side the FullObject transforms take: one line of it (two complex
  // buffers) in 64 KiB of LDS.
  static constexpr size_t fo_max_side() { return sizeof(T) == 4 ? kAxisMax : kAxisMax / 2; }

  // `slices`: the caller runs several object slices (LSQML, the LLK and the model intensities).
  void fo_ensure(size_t cap, bool slices = false) {
    if (nslice_ != 1 && !slices)
      throw Invalid("FullObject sampling with several object slices runs LSQML, the LLK and "
                    "the model intensities on this GPU engine");
    if (nobj_ != 1)
      throw Invalid("FullObject sampling with several object modes is not run on this GPU "
                    "engine");
    if (nyo_ < size_t(N) || nxo_ < size_t(N))
      throw Invalid("FullObject sampling needs an object plane at least the frame's size");
    if (nyo_ > fo_max_side() || nxo_ > fo_max_side())
      throw Invalid("FullObject sampling takes object sides up to " +
                    std::to_string(fo_max_side()) + " pixels on this GPU engine");
    if (fo_plan_[0].L == 0) {
      fo_plan_[0] = axis_plan(int(nxo_));
      fo_plan_[1] = axis_plan(int(nyo_));
      for (int a = 0; a < 2; ++a) {
        const int L = fo_plan_[a].L;
        std::vector<double> tw(2 * size_t(L));
        for (int t = 0; t < L; ++t) {
          const long double ang = -2.0L * 3.14159265358979323846264338327950288L * t / L;
          tw[2 * t] = double(cosl(ang));
          tw[2 * t + 1] = double(sinl(ang));
        }
        fo_tw_[a].ensure(sizeof(cplx<T>) * L);
        upload_complex(fo_tw_[a], tw.data(), size_t(L));
      }
    }
    cap = std::max<size_t>(cap, 1);
    const size_t no = nyo_ * nxo_;
    fo_spec_.ensure(sizeof(cplx<T>) * no);
    fo_acc_.ensure(sizeof(cplx<T>) * no);
    fo_F_.ensure(sizeof(cplx<T>) * M_ * no);
    fo_tmp_.ensure(sizeof(cplx<T>) * std::max<size_t>(cap, M_) * N * nxo_);
    fo_planes_.ensure(sizeof(cplx<T>) * cap * no);
    fo_win_.ensure(sizeof(cplx<T>) * cap * n);
    fo_dwin_.ensure(sizeof(cplx<T>) * cap * n);
    fo_E_.ensure(sizeof(double2) * no);
    fo_D_.ensure(sizeof(T) * M_ * cap * n);
    fo_pts_.ensure(sizeof(double2) * cap);
    fo_tall_.ensure(sizeof(int2) * std::max<size_t>(1, nframe_));
    fo_part_.ensure(sizeof(T) * kFoBlocks);
    fo_cap_ = std::max(fo_cap_, cap);
    if (nslice_ > 1) fo_swin_.ensure(sizeof(cplx<T>) * (nslice_ - 1) * fo_cap_ * n);
  }
  // Several slices under FullObject: every slice's windows of the batch (slice 0's in fo_win_,
  // slice s's in fo_swin_ plane s − 1, tall layout at the corners fo_tall_), read by the
  // multislice passes in place of the slices at the frames' corners.
  DevBuf fo_swin_;
  bool ms_fo() const { return full_object() && nslice_ > 1; }
  cplx<T>* fo_swin(size_t s) {
    return s == 0 ? fo_win_.as<cplx<T>>() : fo_swin_.as<cplx<T>>() + (s - 1) * fo_cap_ * n;
  }
  void fo_slice_windows(const int* ord, int nb) {
    fo_points(ord, nb);
    for (size_t sl = 0; sl < nslice_; ++sl) {
      fo_object_spectrum(slice_obj(sl));
      fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 0, fo_swin(sl), (long long)n);
    }
  }
  // A slice direction as the batch reads it: its sampled windows into fo_dwin_.
  void fo_dir_windows(const cplx<T>* dir, int nb) {
    fo_object_spectrum(dir);
    fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 0, fo_dwin_.p, (long long)n);
  }
  // The position sums on slice s under FullObject: its windows' derivatives over the position
  // against the wave entering it (w, per slot with w_slot), its phase q and χ (psi_).
  void fo_position_sums(size_t s, const cplx<T>* w, int w_slot, const cplx<T>* q, int q_frame,
                        const int* ord, int nb) {
    (void)ord;
    cplx<T>* grad = pgrad_.as<cplx<T>>() + cap_ * n;
    run(K_OTHER, [&] {
      fo_object_spectrum(slice_obj(s));
      fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 1, grad, 2ll * n);
      fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 2, grad + n, 2ll * n);
      hipLaunchKernelGGL(k_pos_sums<T>, dim3(aux_blocks(), nb), dim3(THREADS), 0, stream_,
                         (const cplx<T>*)grad, w, q, q_frame, (const cplx<T>*)psi_.as<cplx<T>>(),
                         N, M_, pos_part_.as<double>(), w_slot);
    });
  }

  // One pass of object-side line transforms along `axis` (0: rows of nxo, 1: columns of nyo).
  template <int LOAD, int STORE>
  void axis_pass(int axis, bool inverse, const cplx<T>* in, void* out, const AxisGeom& g,
                 int nplanes, PhaseArgs ph = PhaseArgs{nullptr, 0, 0, 0, 0}, T scale = T(1)) {
    const AxisPlan& pl = fo_plan_[axis];
    const int L = pl.L;
    const int G = std::max(1, std::min(16, 1024 / L));
    const int Lp = G > 1 ? L + 1 : L;
    const size_t lds = 4 * size_t(G) * Lp * sizeof(T);
    hipLaunchKernelGGL((k_axis_fft<T, LOAD, STORE>), dim3((g.nlines + G - 1) / G, nplanes),
                       dim3(kAxisThreads), lds, stream_, in, out, fo_tw_[axis].as<cplx<T>>(), pl,
                       int(inverse), g, G, ph, scale);
  }
  AxisGeom rows_geom(long long in_plane, int n_in, long long out_plane, int n_out, int nlines,
                     int in_line, int out_line) const {
    return AxisGeom{nlines, in_plane, in_line, 1, n_in, out_plane, out_line, 1, n_out};
  }
  AxisGeom cols_geom(long long in_plane, int n_in, long long out_plane, int n_out) const {
    return AxisGeom{int(nxo_), in_plane, 1, int(nxo_), n_in, out_plane, 1, int(nxo_), n_out};
  }

  // The slots' points (corner + offset of each frame of `ord`) and the tall corners.
  void fo_points(const int* ord, int nb) {
    hipLaunchKernelGGL(k_fo_points, dim3((nb + 255) / 256), dim3(256), 0, stream_,
                       corner_.as<int2>(), (const double2*)offs_.as<double2>(), ord, nb, N,
                       fo_pts_.as<double2>(), fo_tall_.as<int2>());
  }
  // The unnormalised spectrum of object plane `obj` into fo_spec_.
  void fo_object_spectrum(const cplx<T>* obj) {
    const long long no = (long long)(nyo_ * nxo_);
    axis_pass<AL_PLAIN, AS_PLAIN>(
        0, false, obj, fo_spec_.p,
        rows_geom(no, int(nxo_), no, int(nxo_), int(nyo_), int(nxo_), int(nxo_)), 1);
    axis_pass<AL_PLAIN, AS_PLAIN>(1, false, fo_spec_.as<cplx<T>>(), fo_spec_.p,
                                  cols_geom(no, int(nyo_), no, int(nyo_)), 1);
  }
  // Every slot's N × N window of the plane whose spectrum is `spec`, sampled at its point
  // (with `deriv` 1 or 2 its derivative over the position along y or x): plane b of `out`
  // at out + b·out_plane, real part floored at 0 with STORE = AS_REAL_FLOOR.
  template <int STORE>
  void fo_windows(const cplx<T>* spec, int nb, int deriv, void* out, long long out_plane) {
    const long long no = (long long)(nyo_ * nxo_), tp = (long long)N * nxo_;
    axis_pass<AL_PHASE, AS_PLAIN>(
        1, true, spec, fo_tmp_.p, cols_geom(0, int(nyo_), tp, N), nb,
        PhaseArgs{(const double2*)fo_pts_.as<double2>(), int(nyo_), int(nxo_), deriv, 1});
    (void)no;
    axis_pass<AL_PLAIN, STORE>(0, true, fo_tmp_.as<cplx<T>>(), out,
                               rows_geom(tp, int(nxo_), out_plane, N, N, int(nxo_), N), nb,
                               PhaseArgs{nullptr, 0, 0, 0, 0},
                               T(1.0 / (double(nyo_) * double(nxo_))));
  }
  // The slots' N × N patches (plane b at patches + b·in_plane, natural layout) placed by the
  // sampling's adjoint: the spectrum Σ_b conj(exp(+2πi f·t_b))·DFT(patch_b at the origin) into
  // fo_acc_ (added to what it holds with `accumulate`).
  void fo_place_spectrum(const cplx<T>* patches, long long in_plane, int nb, bool accumulate) {
    const long long no = (long long)(nyo_ * nxo_), tp = (long long)N * nxo_;
    axis_pass<AL_PLAIN, AS_PLAIN>(0, false, patches, fo_tmp_.p,
                                  rows_geom(in_plane, N, tp, int(nxo_), N, N, int(nxo_)), nb);
    axis_pass<AL_PLAIN, AS_PLAIN>(1, false, fo_tmp_.as<cplx<T>>(), fo_planes_.p,
                                  cols_geom(tp, N, no, int(nyo_)), nb);
    hipLaunchKernelGGL(k_fo_place_acc<T>, dim3(unsigned((no + 255) / 256)), dim3(256), 0,
                       stream_, fo_acc_.as<cplx<T>>(), (const cplx<T>*)fo_planes_.as<cplx<T>>(),
                       no, (const double2*)fo_pts_.as<double2>(), nb, int(nyo_), int(nxo_),
                       int(accumulate));
  }
  // out (+)= scale·DFT⁻¹(fo_acc_)/(nyo·nxo) with STORE AS_ADD (AS_PLAIN: =; AS_REAL_FLOOR:
  // the real part floored at 0 into a real plane); fo_acc_ is overwritten.
  template <int STORE = AS_ADD>
  void fo_add_placed(void* out, double scale) {
    const long long no = (long long)(nyo_ * nxo_);
    axis_pass<AL_PLAIN, AS_PLAIN>(
        0, true, fo_acc_.as<cplx<T>>(), fo_acc_.p,
        rows_geom(no, int(nxo_), no, int(nxo_), int(nyo_), int(nxo_), int(nxo_)), 1);
    axis_pass<AL_PLAIN, STORE>(1, true, fo_acc_.as<cplx<T>>(), out,
                               cols_geom(no, int(nyo_), no, int(nyo_)), 1,
                               PhaseArgs{nullptr, 0, 0, 0, 0},
                               T(scale / (double(nyo_) * double(nxo_))));
  }
  // pty-chi's illumination map under FullObject (ptyrs/src/engine/sampling.rs:51-75): every
  // frame's Σ_m |P_m|² placed at its corner plus offset by the sampling's adjoint, real part
  // floored at 0, into illum_.
  void fo_illumination() {
    const long long no = (long long)(nyo_ * nxo_), tp = (long long)N * nxo_;
    axis_pass<AL_ABS2, AS_PLAIN>(0, false, P_.as<cplx<T>>(), fo_tmp_.p,
                                 rows_geom(n, N, tp, int(nxo_), N, N, int(nxo_)), M_);
    axis_pass<AL_PLAIN, AS_PLAIN>(1, false, fo_tmp_.as<cplx<T>>(), fo_F_.p,
                                  cols_geom(tp, N, no, int(nyo_)), M_);
    const int step = int(fo_cap_);
    for (int f0 = 0; f0 < int(nframe_); f0 += step) {
      const int nb = std::min(step, int(nframe_) - f0);
      fo_points(iota_.as<int>() + f0, nb);
      hipLaunchKernelGGL(k_fo_phase_acc, dim3(unsigned((no + 255) / 256)), dim3(256), 0, stream_,
                         fo_E_.as<double2>(), (const double2*)fo_pts_.as<double2>(), nb,
                         int(nyo_), int(nxo_), int(f0 > 0));
    }
    hipLaunchKernelGGL(k_fo_times_sum<T>, dim3(unsigned((no + 255) / 256)), dim3(256), 0,
                       stream_, fo_acc_.as<cplx<T>>(), (const cplx<T>*)fo_F_.as<cplx<T>>(),
                       (const double2*)fo_E_.as<double2>(), M_, no);
    fo_add_placed<AS_REAL_FLOOR>(illum_.p, 1.0);
  }
  // The batch's sampled object windows (tall, fo_win_) from the object's spectrum in fo_spec_.
  void fo_object_windows(int nb) {
    fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 0, fo_win_.p, (long long)n);
  }
  // rPIE's overlaps under FullObject (ptyrs/src/engine/ptycho.rs:1089-1130): per mode m the
  // batch's |W_m|² placed at every slot's point by the sampling's adjoint and read back through
  // each slot's window, real part floored at 0, into fo_D_; their max over the slots and
  // pixels of Σ_m into pmax_. The placed plane's transform back and forth is left out (the
  // read takes the placed spectrum directly).
  void fo_overlaps(int nb) {
    const long long no = (long long)(nyo_ * nxo_), tp = (long long)N * nxo_;
    axis_pass<AL_ABS2, AS_PLAIN>(0, false, Wt_.as<cplx<T>>(), fo_tmp_.p,
                                 rows_geom(n, N, tp, int(nxo_), N, N, int(nxo_)), M_);
    axis_pass<AL_PLAIN, AS_PLAIN>(1, false, fo_tmp_.as<cplx<T>>(), fo_F_.p,
                                  cols_geom(tp, N, no, int(nyo_)), M_);
    hipLaunchKernelGGL(k_fo_place_same<T>, dim3(unsigned((no + 255) / 256)), dim3(256), 0,
                       stream_, fo_F_.as<cplx<T>>(), M_, (const double2*)fo_pts_.as<double2>(),
                       nb, int(nyo_), int(nxo_));
    for (int m = 0; m < M_; ++m)
      fo_windows<AS_REAL_FLOOR>(fo_F_.as<cplx<T>>() + m * no, nb, 0,
                                fo_D_.as<T>() + size_t(m) * fo_cap_ * n, (long long)n);
    const long long count = (long long)nb * n;
    const int blocks = int(std::min<long long>(kFoBlocks, (count + 255) / 256));
    hipLaunchKernelGGL(k_fo_dens_max<T>, dim3(blocks), dim3(256), 0, stream_,
                       (const T*)fo_D_.as<T>(), (long long)(fo_cap_ * n), M_, count,
                       fo_part_.as<T>());
    hipLaunchKernelGGL(k_max_reduce<T>, dim3(1), dim3(MAX_THREADS), 0, stream_,
                       (const T*)fo_part_.as<T>(), blocks, pmax_.as<T>());
  }

  // Position offsets (PATCH): device [nframe] (dy, dx) in FP64; the moved probe planes of a
  // stack, the probe spectra, the moved last-slice phases, the window derivative planes.
  bool pos_on_ = false, pos_refine_ = false;
  bool reverse_ = false, dft_plane_ = true;  // the counts' index map (capi.hip prepare)
  double plane_quad_[2] = {0, 0};
  DevBuf offs_, inc0_, spec_, qv_, pgrad_, gspec_, pos_part_, pos_sums_;

  void set_position(const double* offsets, bool refine) override {
    if (offsets == nullptr) {
      pos_on_ = pos_refine_ = false;
      return;
    }
    // One offset per planned frame (the frames not landed yet keep theirs for when they land).
    const size_t nall = planned();
    for (size_t i = 0; i < 2 * nall; ++i)
      if (!std::isfinite(offsets[i])) throw Invalid("position offsets must be finite");
    offs_.ensure(sizeof(double) * 2 * std::max<size_t>(1, nall));
    if (nall)
      PTYRS_HIP(hipMemcpy(offs_.p, offsets, sizeof(double) * 2 * nall, hipMemcpyHostToDevice));
    pos_on_ = true;
    pos_refine_ = refine;
  }
  // Live ingestion: the counts buffer holds every planned frame's slot from the start, so a
  // landing frame is one copy into its slot (through pinned staging), no reallocation; the
  // epochs' graphs are keyed by the landed count and captured again when it changes.
  DevBuf stage_dev_;
  void* stage_host_ = nullptr;
  size_t stage_bytes_ = 0;
  void live_reset() override {
    synchronize();
    nframe_ = 0;
    nb_obs_ = 0;
    total_ = 0.0;
    max_frame_total_ = 0.0;
  }
  void live_push(const float* counts, size_t nframes, size_t nb_obs, double total,
                 double max_frame_total) override {
    if (nframes == 0) return;
    if (nframe_ + nframes > planned())
      throw Invalid("live_push lands " + std::to_string(nframes) + " frames past the " +
                    std::to_string(planned() - nframe_) + " planned frames left");
    const size_t bytes = sizeof(float) * nframes * n;
    synchronize();
    if (bytes > stage_bytes_) {
      if (stage_host_) PTYRS_HIP(hipHostFree(stage_host_));
      stage_host_ = nullptr;
      stage_bytes_ = 0;
      PTYRS_HIP(hipHostMalloc(&stage_host_, bytes, hipHostMallocDefault));
      stage_bytes_ = bytes;
    }
    std::memcpy(stage_host_, counts, bytes);
    PTYRS_HIP(hipMemcpyAsync(counts_.as<float>() + nframe_ * n, stage_host_, bytes,
                             hipMemcpyHostToDevice, stream_));
    synchronize();
    nframe_ += nframes;
    nb_obs_ += nb_obs;
    total_ += total;
    max_frame_total_ = std::max(max_frame_total_, max_frame_total);
  }
  ~Engine() override {
    if (stage_host_) (void)hipHostFree(stage_host_);
  }

  int get_position(double* out) override {
    if (!pos_on_) return 0;
    if (out && planned()) {
      PTYRS_HIP(hipStreamSynchronize(stream_));
      PTYRS_HIP(hipMemcpy(out, offs_.p, sizeof(double) * 2 * planned(),
                          hipMemcpyDeviceToHost));
    }
    return pos_refine_ ? 2 : 1;
  }

  // The probe planes of a stack's views moved by their frames' offsets as pty-chi PATCH moves
  // them, conj(q0)·DFT⁻¹(DFT(q0·P_m)·exp(−2πi f·d_b)), q0 the first slice's plane phase
  // (ptyrs/src/engine/features.rs:386-429, lsqml.rs:483-536), into inc0_.
  // With conj_q false (one slice) the planes keep q: W = DFT⁻¹(DFT(P·q)·exp(−2πi f·d_b)),
  // the views' waves.
  // `src` (M planes, nullable) replaces the probe terms (the ML direction's dP·q); `out`
  // (nullable) replaces inc0_; with `spectra` the spectra of `src` are already in spec_.
  void moved_probes(const int* ord, int nb, bool conj_q = true, cplx<T>* out = nullptr,
                    const cplx<T>* src = nullptr, bool spectra = false) {
    const int K = nb * M_;
    if (out == nullptr) out = inc0_.as<cplx<T>>();
    run(K_PROBE, [&] {
      if (!spectra) probe_spectra(src);
      hipLaunchKernelGGL(k_ramp_planes<T>, dim3((n + 255) / 256, K), dim3(256), 0, stream_, out,
                         (const cplx<T>*)spec_.as<cplx<T>>(),
                         (const double2*)offs_.as<double2>(), ord, N, M_, 1.0 / double(n));
      hipLaunchKernelGGL((k_cols<T, N, true>), dim3(SC::NBLK, K), dim3(SC::NT), 0, stream_, out,
                         tw_.as<cplx<T>>());
      hipLaunchKernelGGL((k_rows<T, N, true>), dim3(S::NBLK, K), dim3(S::NT), 0, stream_, out,
                         tw_.as<cplx<T>>());
      if (nslice_ == 1 && conj_q)
        hipLaunchKernelGGL((k_times_q<T, true>), dim3((n + 255) / 256, K), dim3(256), 0, stream_,
                           out, (const cplx<T>*)q_.as<cplx<T>>(), n);
    });
  }
  // The unnormalised spectra of the probe terms (or of `src`, M planes) into spec_.
  void probe_spectra(const cplx<T>* src = nullptr) {
    const size_t pn = size_t(M_) * n;
    const void* from = src ? (const void*)src : (nslice_ == 1 ? Wt_.p : P_.p);
    PTYRS_HIP(hipMemcpyAsync(spec_.p, from, sizeof(cplx<T>) * pn, hipMemcpyDeviceToDevice,
                             stream_));
    hipLaunchKernelGGL((k_rows<T, N, false>), dim3(S::NBLK, M_), dim3(S::NT), 0, stream_,
                       spec_.as<cplx<T>>(), tw_.as<cplx<T>>());
    hipLaunchKernelGGL((k_cols<T, N, false>), dim3(SC::NBLK, M_), dim3(SC::NT), 0, stream_,
                       spec_.as<cplx<T>>(), tw_.as<cplx<T>>());
  }

  // The wave entering slice 0 (P per mode, or the moved planes per view) and the last slice's
  // plane phase (q, or past one slice with offsets and a plane phase, q moved per frame).
  struct Waves {
    const cplx<T>* w0;
    int w0_view;
    const cplx<T>* qlast;
    int q_frame;
  };
  // Under FullObject with OPR the slots' waves W_b = P_b·q from each frame's own probe planes
  // (unmoved), plane slot·M + m of inc0_ (ptyrs/src/engine/lsqml.rs:483-536 with the sampled
  // windows).
  const cplx<T>* opr_waves_sampled(const int* ord, int nb) {
    const int K = nb * M_;
    cplx<T>* pl = inc0_.as<cplx<T>>();
    run(K_PROBE, [&] {
      hipLaunchKernelGGL(k_opr_planes<T>, dim3((n + 255) / 256, K), dim3(256), 0, stream_, pl,
                         (const cplx<T>*)P_.as<cplx<T>>(), (const cplx<T>*)opr_U_.as<cplx<T>>(),
                         (const double*)opr_w_.as<double>(),
                         opr_c_on_ ? (const double*)opr_c_.as<double>() : nullptr, ord, opr_k_,
                         M_, n);
      hipLaunchKernelGGL((k_times_q<T, false>), dim3((n + 255) / 256, K), dim3(256), 0, stream_,
                         pl, (const cplx<T>*)q_.as<cplx<T>>(), n);
    });
    return pl;
  }

  Waves stack_waves(const int* ord, int nb) {
    Waves w{P_.as<cplx<T>>(), 0, q_.as<cplx<T>>(), 0};
    // Under FullObject the probe does not move (the windows do).
    const bool moving = pos_on_ && !full_object();
    if (opr_on_) {
      // Each frame's own probe planes (OPR), moved by its offset with offsets:
      // conj(q0)·DFT⁻¹(DFT(q0·P_b)·exp(−2πi f·d_b)) (ptyrs/src/engine/lsqml.rs:507-520).
      const int K = nb * M_;
      cplx<T>* pl = inc0_.as<cplx<T>>();
      const cplx<T>* q0 = nslice_ == 1 ? q_.as<cplx<T>>() : nullptr;
      run(K_PROBE, [&] {
        hipLaunchKernelGGL(k_opr_planes<T>, dim3((n + 255) / 256, K), dim3(256), 0, stream_, pl,
                           (const cplx<T>*)P_.as<cplx<T>>(),
                           (const cplx<T>*)opr_U_.as<cplx<T>>(),
                           (const double*)opr_w_.as<double>(),
                           opr_c_on_ ? (const double*)opr_c_.as<double>() : nullptr, ord,
                           opr_k_, M_, n);
        if (moving) {
          if (q0)
            hipLaunchKernelGGL((k_times_q<T, false>), dim3((n + 255) / 256, K), dim3(256), 0,
                               stream_, pl, q0, n);
          hipLaunchKernelGGL((k_rows<T, N, false>), dim3(S::NBLK, K), dim3(S::NT), 0, stream_,
                             pl, tw_.as<cplx<T>>());
          hipLaunchKernelGGL((k_cols<T, N, false>), dim3(SC::NBLK, K), dim3(SC::NT), 0,
                             stream_, pl, tw_.as<cplx<T>>());
          hipLaunchKernelGGL(k_ramp_planes<T>, dim3((n + 255) / 256, K), dim3(256), 0, stream_,
                             pl, (const cplx<T>*)pl, (const double2*)offs_.as<double2>(), ord, N,
                             M_, 1.0 / double(n), 1);
          hipLaunchKernelGGL((k_cols<T, N, true>), dim3(SC::NBLK, K), dim3(SC::NT), 0, stream_,
                             pl, tw_.as<cplx<T>>());
          hipLaunchKernelGGL((k_rows<T, N, true>), dim3(S::NBLK, K), dim3(S::NT), 0, stream_,
                             pl, tw_.as<cplx<T>>());
          if (q0)
            hipLaunchKernelGGL((k_times_q<T, true>), dim3((n + 255) / 256, K), dim3(256), 0,
                               stream_, pl, q0, n);
        }
      });
      w.w0 = pl;
      w.w0_view = 1;
      if (!moving) return w;
    } else {
      if (!moving) return w;
      moved_probes(ord, nb);
    }
    w.w0 = inc0_.as<cplx<T>>();
    w.w0_view = 1;
    if (nslice_ > 1 && (plane_quad_[0] != 0.0 || plane_quad_[1] != 0.0)) {
      run(K_PROBE, [&] {
        hipLaunchKernelGGL(k_moved_quad<T>, dim3((n + 255) / 256, nb), dim3(256), 0, stream_,
                           qv_.as<cplx<T>>(), (const cplx<T>*)q_.as<cplx<T>>(),
                           (const double2*)offs_.as<double2>(), ord, N, plane_quad_[0],
                           plane_quad_[1]);
      });
      w.qlast = qv_.as<cplx<T>>();
      w.q_frame = 1;
    }
    return w;
  }
  void ensure_positions(size_t cap) {
    if (!pos_on_) return;
    inc0_.ensure(sizeof(cplx<T>) * cap * M_ * n);
    spec_.ensure(sizeof(cplx<T>) * M_ * n);
    qv_.ensure(sizeof(cplx<T>) * cap * n);
    gspec_.ensure(sizeof(cplx<T>) * cap * M_ * n);
    pgrad_.ensure(sizeof(cplx<T>) * 3 * cap * n);
    pos_part_.ensure(sizeof(double) * 4 * cap * aux_blocks());
    pos_sums_.ensure(sizeof(double) * 4 * cap);
  }

  // Multislice state: the gaps' transfer functions, a plane of ones, the incident waves of
  // slices 1.. of a stack ([slot][cap·M][n], natural layout). Without checkpointing slice s
  // has slot s − 1. With checkpoints every k slices, slices that are multiples of k keep slots
  // 0 .. (S−1)/k − 1 and the others share k − 1 segment slots (slice s in slot
  // (S−1)/k + s mod k − 1), which hold one segment [c+1, c+k) at a time: the forward pass
  // leaves the top segment there, and the adjoint recomputes each lower one from its
  // checkpoint c when it reaches it (`seg_` = the c held).
  DevBuf gaptf_, ones_, inc_;
  size_t seg_ = 0;
  cplx<T>* slice_obj(size_t s) { return O_.as<cplx<T>>() + s * nobj_ * nyo_ * nxo_; }
  cplx<T>* slice_dir(size_t s) { return dO_.as<cplx<T>>() + s * nyo_ * nxo_; }
  // The interval in use: at most nslice − 1 (a longer one would only add segment slots).
  size_t every() const {
    return checkpoint > 1 && nslice_ > 2 ? std::min(checkpoint, nslice_ - 1) : 1;
  }
  void set_checkpoint(size_t k) override {
    synchronize();
    checkpoint = k;
    inc_.release();  // reallocated at the new slot count by the next call
  }
  size_t wave_bytes() const override { return inc_.bytes; }
  size_t inc_slots() const {
    if (nslice_ < 2) return 0;
    const size_t k = every();
    return k == 1 ? nslice_ - 1 : (nslice_ - 1) / k + k - 1;
  }
  size_t inc_slot(size_t s) const {
    const size_t k = every();
    if (k == 1) return s - 1;
    return s % k == 0 ? s / k - 1 : (nslice_ - 1) / k + s % k - 1;
  }
  cplx<T>* incident(size_t s) { return inc_.as<cplx<T>>() + inc_slot(s) * cap_ * M_ * n; }

  // The incident waves of the segment holding slice s, recomputed from its checkpoint if the
  // segment slots hold another (the adjoint's descent through the slices).
  void segment_for(size_t s, const Waves& w, const int* ord, int nb) {
    const size_t k = every();
    if (k == 1 || s % k == 0 || seg_ == s - s % k) return;
    const size_t c = s - s % k, top = std::min(c + k - 1, nslice_ - 1);
    run(K_PASS_A, [&] {
      for (size_t t = c; t < top; ++t) {
        auto in = wave(t, w);
        window_rows<MS_MUL>(incident(t + 1), t, in.first, in.second, nullptr, 0, ord, nb);
        gap_tail(incident(t + 1), nb * M_, t, false);
      }
    });
    seg_ = c;
  }

  // An angular-spectrum gap's propagation of `nplane` planes in place, or its adjoint.
  void gap(cplx<T>* planes, int nplane, size_t g, bool adjoint) {
    hipLaunchKernelGGL((k_rows<T, N, false>), dim3(S::NBLK, nplane), dim3(S::NT), 0, stream_,
                       planes, tw_.as<cplx<T>>());
    gap_tail(planes, nplane, g, adjoint);
  }
  // The gap after its forward row DFTs (done by k_rows_ms with the window product before it).
  void gap_tail(cplx<T>* planes, int nplane, size_t g, bool adjoint) {
    const cplx<T>* tf = gaptf_.as<cplx<T>>() + g * n;
    const T inv_n = T(1) / T(n);
    if (adjoint)
      hipLaunchKernelGGL((k_gap_cols<T, N, true>), dim3(SC::NBLK, nplane), dim3(SC::NT), 0,
                         stream_, planes, tf, tw_.as<cplx<T>>(), inv_n);
    else
      hipLaunchKernelGGL((k_gap_cols<T, N, false>), dim3(SC::NBLK, nplane), dim3(SC::NT), 0,
                         stream_, planes, tf, tw_.as<cplx<T>>(), inv_n);
    hipLaunchKernelGGL((k_rows<T, N, true>), dim3(S::NBLK, nplane), dim3(S::NT), 0, stream_,
                       planes, tw_.as<cplx<T>>());
  }

  // window<OP> followed by the forward row DFTs of its output, in one pass.
  template <int OP>
  void window_rows(cplx<T>* out, size_t s, const cplx<T>* w, int w_view, const cplx<T>* q,
                   int q_frame, const int* ord, int nb) {
    const bool fo = ms_fo();
    hipLaunchKernelGGL((k_rows_ms<T, N, OP>), dim3(S::NBLK, nb * M_), dim3(S::NT), 0, stream_,
                       out, (const cplx<T>*)(fo ? fo_swin(s) : slice_obj(s)),
                       fo ? N : int(nxo_), w, w_view, q,
                       fo ? fo_tall_.as<int2>() : corner_.as<int2>(), ord, M_, q_frame,
                       tw_.as<cplx<T>>());
  }

  template <int OP>
  void window(cplx<T>* out, size_t s, const cplx<T>* w, int w_view, const cplx<T>* q,
              int q_frame, const int* ord, int nb) {
    const bool fo = ms_fo();
    hipLaunchKernelGGL((k_ms_window<T, OP>), dim3((n + 255) / 256, nb * M_), dim3(256), 0,
                       stream_, out, fo ? fo_swin(s) : slice_obj(s), fo ? N : int(nxo_), w,
                       w_view, q, fo ? fo_tall_.as<int2>() : corner_.as<int2>(), ord, N, M_,
                       q_frame);
  }

  // The wave entering slice s: slice 0's from `w`, the others' kept per view.
  std::pair<const cplx<T>*, int> wave(size_t s, const Waves& w) {
    if (s == 0) return {w.w0, w.w0_view};
    return {incident(s), 1};
  }

  // The exit waves of a stack through the slices (ptyrs/src/engine/multislice.rs:86-171):
  // the incident waves of slices 1.. kept, psi = q·O_last[win]·w_last, natural layout.
  void ms_forward(const int* ord, int nb, const Waves& w) {
    const size_t S = nslice_;
    run(K_PASS_A, [&] {
      for (size_t s = 0; s + 1 < S; ++s) {
        auto in = wave(s, w);
        window_rows<MS_MUL>(incident(s + 1), s, in.first, in.second, nullptr, 0, ord, nb);
        gap_tail(incident(s + 1), nb * M_, s, false);
      }
      auto in = wave(S - 1, w);
      window<MS_MUL>(psi_.as<cplx<T>>(), S - 1, in.first, in.second, w.qlast, w.q_frame, ord,
                     nb);
    });
    seg_ = (S - 1) / every() * every();
  }

  void rows(bool inverse, int nplane) {
    if (inverse)
      hipLaunchKernelGGL((k_rows<T, N, true>), dim3(S::NBLK, nplane), dim3(S::NT), 0, stream_,
                         psi_.as<cplx<T>>(), tw_.as<cplx<T>>());
    else
      hipLaunchKernelGGL((k_rows<T, N, false>), dim3(S::NBLK, nplane), dim3(S::NT), 0, stream_,
                         psi_.as<cplx<T>>(), tw_.as<cplx<T>>());
  }

  // The angular spectrum's column passes over `nplane` planes of `nfield` fields `field` apart.
  template <bool INV>
  void pass_t(cplx<T>* psi, int nplane, int nfield = 1, size_t field = 0) {
    hipLaunchKernelGGL((k_pass_t<T, N, INV>), dim3(SC::NBLK, nplane, nfield), dim3(SC::NT), 0,
                       stream_, psi, field, tf_.as<cplx<T>>(), tw_.as<cplx<T>>());
  }
  DevBuf psi3_, gO_, hO_, dOc_, gP_, hP_, dPc_, dWt_, cg_part_, cg_sums_, gam_part_, gam_;
  // GpuOptions::psi_f16: DM's and RAAR's carried waves and RAAR's iterate as f16 pairs times
  // half_scale_ (complex64 engines only).
  bool psi_f16_ = false;
  float half_scale_ = 1.f;
  DevBuf carried16_, iter16_;
  // The scale of f16 carried waves for frames whose counts sum to `max_frame_sum` at most: the
  // largest power of two s with s²·2^ceil(log2 sum) ≤ 2^30, i.e. s·√sum ≤ 2^15, half the
  // largest finite half (a projected frame's components are within √sum; the factor two is
  // for DM's and RAAR's combinations). ptyrs/src/gpu/psi.rs:40-50
  static float half_scale(double max_frame_sum) {
    const uint64_t sum = uint64_t(std::ceil(std::max(max_frame_sum, 1.0)));
    int log2 = 0;
    while (log2 < 64 && (uint64_t(1) << log2) < sum) ++log2;  // ceil(log2(sum)), 0 for 1
    const int k = 30 - log2;
    const int e = k >= 0 ? k / 2 : -((-k + 1) / 2);  // floor(k / 2)
    return std::ldexp(1.0f, e);
  }
  // Calls f with the carried waves (and RAAR's iterate) in the engine's storage form.
  template <typename F>
  void with_carried(bool half, F&& f) {
    if constexpr (sizeof(T) == 4) {
      if (half) {
        const float us = 1.0f / half_scale_;
        f(Carried<T, true>{carried16_.as<uint32_t>(), half_scale_, us},
          Carried<T, true>{iter16_.as<uint32_t>(), half_scale_, us});
        return;
      }
    }
    f(Carried<T, false>{carried_.as<cplx<T>>()}, Carried<T, false>{iter_.as<cplx<T>>()});
  }
  DevBuf carried_, iter_, list_all_, count_all_, pnum_part_, pden_part_, pnum_, pden_,
      pmaxpart_, pmax2_, abs_part_;
  DevBuf dO_, pre_, illum_, afar_all_, afar_tmp_, prev_, icalc_, xi_part_, xi_part2_, dots_part_,
      sums_, log_, step_, pre_part_, floor_, dPm_;
  int ntx_ = 0, nty_ = 0;
  size_t cap_ = 0, nchunk_cap_ = 0, order_cap_ = 0;
  DevBuf O_, P_, Wt_, W2_, q_, tw_, counts_, corner_, iota_, hist_, eord_;

  static uint64_t bits(double v) {
    uint64_t u;
    std::memcpy(&u, &v, sizeof u);
    return u;
  }
  DevBuf psi_, omax_, dP_, list_, lcount_, tilemax_, pmax_, order_, llk_part_, D_;

  void upload_complex(DevBuf& d, const double* src, size_t count) {
    if (stream_) PTYRS_HIP(hipStreamSynchronize(stream_));
    std::vector<cplx<T>> h(count);
    for (size_t i = 0; i < count; ++i) h[i] = cplx<T>{T(src[2 * i]), T(src[2 * i + 1])};
    PTYRS_HIP(hipMemcpy(d.p, h.data(), sizeof(cplx<T>) * count, hipMemcpyHostToDevice));
  }
  void download_complex(const DevBuf& d, double* out, size_t count) {
    std::vector<cplx<T>> h(count);
    PTYRS_HIP(hipStreamSynchronize(stream_));
    PTYRS_HIP(hipMemcpy(h.data(), d.p, sizeof(cplx<T>) * count, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < count; ++i) {
      out[2 * i] = double(h[i].x);
      out[2 * i + 1] = double(h[i].y);
    }
  }

  static constexpr size_t kIllumChunk = 1024;
  // rPIE's pass C cuts a batch into chunks of slots so that its grid has about this many
  // workgroups (each chunk leaves one probe-direction partial per mode). At 256² × 4 modes pass
  // C holds 150 VGPRs and 53 KB of LDS, three workgroups per CU: 768 is one round, 73 µs per
  // launch against 87 at 512 and 85 at 1024.
  static constexpr int kPassCBlocks = 768;
  size_t illum_chunk() const { return std::max<size_t>(1, std::min(nframe_, kIllumChunk)); }

  void ensure_lsqml(size_t cap) {
    const size_t no = nyo_ * nxo_;
    list_all_.ensure(sizeof(int) * size_t(ntiles()) * illum_chunk());
    count_all_.ensure(sizeof(int) * size_t(ntiles()));
    dO_.ensure(sizeof(cplx<T>) * nslice_ * no);
    pre_.ensure(sizeof(T) * no);
    illum_.ensure(sizeof(T) * no);
    afar_all_.ensure(sizeof(T) * std::max<size_t>(1, nframe_));
    afar_tmp_.ensure(sizeof(T) * cap);
    prev_.ensure(sizeof(T) * 2);
    icalc_.ensure(sizeof(T) * cap * n);
    xi_part_.ensure(sizeof(double) * 2 * cap * SB::NBLK);
    xi_part2_.ensure(sizeof(double) * 2 * cap * aux_blocks());
    dots_part_.ensure(sizeof(double) * 6 * cap * aux_blocks());
    sums_.ensure(sizeof(double) * 6 * cap);
    log_.ensure(sizeof(double) * (2 + nslice_) * std::max<size_t>(1, nframe_));
    step_.ensure(sizeof(double) * (2 + nslice_));
    pre_part_.ensure(sizeof(T) * kPreBlocks);
    floor_.ensure(sizeof(T) * 1);
    dPm_.ensure(sizeof(cplx<T>) * nprobe_ * n);
  }

  // BH: U·ψ of a batch's views, L_u and L_v (two fields of a batch), the gradient and the
  // previous direction of the object and the probe, the partial sums, the step, the epoch's
  // per-batch log {β, α, top, bottom}.
  DevBuf bh_keep_, bh_l_, bh_gO_, bh_dO_, bh_gP_, bh_dP_, bh_dots_, bh_pa_, bh_pb_, bh_sum_,
      bh_step_, bh_log_;
  EpochGraph bh_graph_;
  static constexpr int kBhDotBlocks = 256;
  static constexpr int kBhSumBlocks = 128;  // first-level blocks of each Hessian partial array
  void ensure_bh(size_t cap) {
    const size_t no = nslice_ * nyo_ * nxo_, np = nprobe_ * n, views = cap_ * M_;
    bh_keep_.ensure(sizeof(cplx<T>) * views * n);
    bh_l_.ensure(sizeof(cplx<T>) * 2 * views * n);
    for (DevBuf* b : {&bh_gO_, &bh_dO_}) b->ensure(sizeof(cplx<T>) * no);
    for (DevBuf* b : {&bh_gP_, &bh_dP_}) b->ensure(sizeof(cplx<T>) * np);
    bh_dots_.ensure(sizeof(double) * 2 * kBhDotBlocks);
    bh_pa_.ensure(sizeof(double) * 3 * views * HessShape<T, N, 2>::NBLK);
    bh_pb_.ensure(sizeof(double) * 3 * cap_ * SC::NBLK);
    bh_sum_.ensure(sizeof(double) * 3 * 2 * kBhSumBlocks);
    bh_step_.ensure(sizeof(double) * 2);
    bh_log_.ensure(sizeof(double) * 4 * ((nframe_ + cap - 1) / cap));
    if (nslice_ > 1) {
      if (2 * views > 65535)
        throw Invalid("BH through several slices runs batches of at most 32767 views");
      bh_chi_.ensure(sizeof(cplx<T>) * nslice_ * views * n);
      bh_t_.ensure(sizeof(cplx<T>) * 2 * views * n);
      bh_ps_.ensure(sizeof(double) * 3 * nslice_ * views * ((n + 255) / 256));
    }
  }
  // BH through several slices: χ_s of every slice, the tangent chains' waves, their partials.
  DevBuf bh_chi_, bh_t_, bh_ps_;
  // BH under PATCH offsets: the batch's moved waves, the spectra of conj(O)·r per view, the
  // spectra of q·g_P and q·D′_P, the wave changes (NP planes per view), the offsets' gradient
  // per slot and the previous direction's offsets by frame.
  DevBuf bh_w_, bh_gg_, bh_su_, bh_sv_, bh_pl_, bh_gd_, bh_ddf_;
  void ensure_bh_moved(size_t cap, bool full) {
    const size_t views = cap_ * M_;
    for (DevBuf* b : {&bh_w_, &bh_gg_}) b->ensure(sizeof(cplx<T>) * views * n);
    for (DevBuf* b : {&bh_su_, &bh_sv_}) b->ensure(sizeof(cplx<T>) * M_ * n);
    bh_pl_.ensure(sizeof(cplx<T>) * views * (full ? 5 : 2) * n);
    bh_gd_.ensure(sizeof(double) * 2 * cap);
    bh_ddf_.ensure(sizeof(double) * 2 * std::max<size_t>(1, nframe_));
  }
  // The unnormalised spectra of q·src (M planes) into dst.
  void q_spectra(const cplx<T>* src, cplx<T>* dst) {
    const int np = int(M_) * n;
    hipLaunchKernelGGL(k_probe_terms<T>, dim3((np + 255) / 256), dim3(256), 0, stream_, src,
                       q_.as<cplx<T>>(), dst, n, np);
    hipLaunchKernelGGL((k_rows<T, N, false>), dim3(S::NBLK, M_), dim3(S::NT), 0, stream_, dst,
                       tw_.as<cplx<T>>());
    hipLaunchKernelGGL((k_cols<T, N, false>), dim3(SC::NBLK, M_), dim3(SC::NT), 0, stream_, dst,
                       tw_.as<cplx<T>>());
  }
  // Inverse transforms of `planes` planes from p, at most 65535 per launch (grid y).
  void inverse_planes(cplx<T>* p, size_t planes) {
    for (size_t p0 = 0; p0 < planes; p0 += 65535) {
      const unsigned k = unsigned(std::min<size_t>(65535, planes - p0));
      hipLaunchKernelGGL((k_cols<T, N, true>), dim3(SC::NBLK, k), dim3(SC::NT), 0, stream_,
                         p + p0 * n, tw_.as<cplx<T>>());
      hipLaunchKernelGGL((k_rows<T, N, true>), dim3(S::NBLK, k), dim3(S::NT), 0, stream_,
                         p + p0 * n, tw_.as<cplx<T>>());
    }
  }

  // One BH batch under PATCH offsets (ptyrs/src/engine/bh.rs:130-577; the reference GPU's
  // bh_batch with its moved probes, ptyrs/src/engine/gpu/bh.rs:426-600): the views' waves are
  // the frames' moved probe terms W_b; the probe gradient moves each frame's conj(O)·r back in
  // Fourier space; refined offsets get their gradient by Parseval against DFT(q·P), their
  // first and second derivatives enter the wave changes along u and v, and their step follows
  // the update.
  void bh_batch_moved(double rho2, double pos2, const int* ord, int nb, double frac, bool cg,
                      bool keep, int slot) {
    const int nt = ntiles();
    const size_t no = nyo_ * nxo_, np = nprobe_ * n;
    const int total = int(np), K = nb * M_;
    const bool refine = pos_refine_;
    const double2* offs = offs_.as<double2>();
    run(K_BIN, [&] {
      hipLaunchKernelGGL(k_bin_tiles<T>, dim3(nt), dim3(THREADS), 0, stream_,
                         corner_.as<int2>(), ord, nb, N, ntx_, int(cap_), list_.as<int>(),
                         lcount_.as<int>(), (const T*)nullptr, M_, int(nyo_), int(nxo_),
                         (T*)nullptr, tilemax_.as<T>(), (const cplx<T>*)nullptr);
    });
    cplx<T>* w = bh_w_.as<cplx<T>>();
    moved_probes(ord, nb, false, w);  // W_b = DFT⁻¹(DFT(q·P)·e_b); spec_ = DFT(q·P)
    pass_a(ord, nb, false, MixArgs<T>{MIX_NONE, {}, {}, T(0), T(0)}, w);
    BArgs<T> ba{nullptr, nullptr, nullptr, nullptr};
    ba.keep = bh_keep_.as<cplx<T>>();
    pass_b<B_BH>(ord, nb, ba);
    const int chunk = chunk_of(nb);
    const int nchunk = (nb + chunk - 1) / chunk;
    run(K_PASS_C, [&] {
      hipLaunchKernelGGL((k_pass_c_chi<T, N>), dim3(SR::NBLK, M_, nchunk), dim3(SR::NT), 0,
                         stream_, psi_.as<cplx<T>>(), O_.as<cplx<T>>(), int(nxo_),
                         Wt_.as<cplx<T>>(), q_.as<cplx<T>>(), corner_.as<int2>(), ord, nb, M_,
                         chunk, 0, dP_.as<cplx<T>>(), tw_.as<cplx<T>>(), 0);
    });
    run(K_OBJ_DIR, [&] {
      hipLaunchKernelGGL(k_lsqml_obj_dir<T>, dim3(nt), dim3(THREADS), 0, stream_,
                         bh_gO_.as<cplx<T>>(), (const cplx<T>*)w, psi_.as<cplx<T>>(), N, M_,
                         corner_.as<int2>(), ord, list_.as<int>(), lcount_.as<int>(), int(cap_),
                         int(nyo_), int(nxo_), ntx_, (const T*)nullptr, (const T*)nullptr, 1);
    });
    cplx<T>* gg = bh_gg_.as<cplx<T>>();
    run(K_PROBE, [&] {
      hipLaunchKernelGGL(k_bh_conj_o<T>, dim3((n + 255) / 256, K), dim3(256), 0, stream_, gg,
                         (const cplx<T>*)psi_.as<cplx<T>>(), (const cplx<T>*)O_.as<cplx<T>>(),
                         int(nxo_), (const int2*)corner_.as<int2>(), ord, N, M_);
      hipLaunchKernelGGL((k_rows<T, N, false>), dim3(S::NBLK, K), dim3(S::NT), 0, stream_, gg,
                         tw_.as<cplx<T>>());
      hipLaunchKernelGGL((k_cols<T, N, false>), dim3(SC::NBLK, K), dim3(SC::NT), 0, stream_, gg,
                         tw_.as<cplx<T>>());
      // g_P = conj(q)·Σ_b DFT⁻¹(G_b·exp(+2πi f·d_b))
      hipLaunchKernelGGL(k_unramp_mean<T>, dim3((n + 255) / 256, M_), dim3(256), 0, stream_,
                         bh_gP_.as<cplx<T>>(), (const cplx<T>*)gg, offs, ord, nb, N, M_,
                         1.0 / double(n), 0);
      inverse_planes(bh_gP_.as<cplx<T>>(), M_);
      hipLaunchKernelGGL((k_times_q<T, true>), dim3((n + 255) / 256, M_), dim3(256), 0, stream_,
                         bh_gP_.as<cplx<T>>(), (const cplx<T>*)q_.as<cplx<T>>(), n);
      if (refine)
        hipLaunchKernelGGL(k_bh_pos_grad<T>, dim3(nb), dim3(THREADS), 0, stream_,
                           (const cplx<T>*)gg, (const cplx<T>*)spec_.as<cplx<T>>(), offs, ord,
                           N, M_, bh_gd_.as<double>());
    });
    const int ndots =
        int(std::min<size_t>(kBhDotBlocks, (no + np + THREADS - 1) / THREADS));
    run(K_DOTS, [&] {
      hipLaunchKernelGGL(k_bh_dots<T>, dim3(ndots), dim3(THREADS), 0, stream_,
                         (const cplx<T>*)bh_gO_.as<cplx<T>>(),
                         (const cplx<T>*)bh_dO_.as<cplx<T>>(), no,
                         (const cplx<T>*)bh_gP_.as<cplx<T>>(),
                         (const cplx<T>*)bh_dP_.as<cplx<T>>(), np, rho2, int(cg),
                         bh_dots_.as<double>());
    });
    const int NP = cg ? 5 : 2;
    cplx<T>* pl = bh_pl_.as<cplx<T>>();
    run(K_PROBE, [&] {
      q_spectra(bh_gP_.as<cplx<T>>(), bh_su_.as<cplx<T>>());
      if (cg) q_spectra(bh_dP_.as<cplx<T>>(), bh_sv_.as<cplx<T>>());
      auto mp = cg ? k_bh_moved_planes<T, 5> : k_bh_moved_planes<T, 2>;
      hipLaunchKernelGGL(mp, dim3((n + 255) / 256, K), dim3(256), 0, stream_, pl,
                         (const cplx<T>*)spec_.as<cplx<T>>(),
                         (const cplx<T>*)bh_su_.as<cplx<T>>(), rho2,
                         (const cplx<T>*)bh_sv_.as<cplx<T>>(), offs, ord,
                         refine ? (const double*)bh_gd_.as<double>() : nullptr, pos2,
                         refine && cg ? (const double*)bh_ddf_.as<double>() : nullptr, N, M_,
                         1.0 / double(n));
      inverse_planes(pl, size_t(K) * NP);
    });
    const T norm = T(1) / (angular_ ? T(n) * T(n) : T(n));
    using HA = HessShape<T, N, 2>;
    run(K_HESS_A, [&] {
      auto ha = cg ? k_bh_hess_a_moved<T, N, 2> : k_bh_hess_a_moved<T, N, 1>;
      hipLaunchKernelGGL(ha, dim3(HA::NBLK, K), dim3(HA::NT), 0, stream_,
                         (const cplx<T>*)O_.as<cplx<T>>(),
                         (const cplx<T>*)bh_gO_.as<cplx<T>>(),
                         (const cplx<T>*)bh_dO_.as<cplx<T>>(), int(nxo_), (const cplx<T>*)w,
                         (const cplx<T>*)pl, (const int2*)corner_.as<int2>(), ord, M_,
                         (const cplx<T>*)psi_.as<cplx<T>>(), bh_l_.as<cplx<T>>(),
                         bh_pa_.as<double>(), (const cplx<T>*)tw_.as<cplx<T>>());
    });
    bh_hess_b_and_step(ord, nb, cg, norm, ndots, frac, slot,
                       refine ? (const double*)bh_gd_.as<double>() : nullptr, refine ? nb : 0,
                       pos2);
    run(K_OBJ_UPDATE, [&] {
      const size_t all = no + np;
      hipLaunchKernelGGL(k_bh_update<T>, dim3(unsigned((all + 255) / 256)), dim3(256), 0,
                         stream_, O_.as<cplx<T>>(), bh_dO_.as<cplx<T>>(),
                         (const cplx<T>*)bh_gO_.as<cplx<T>>(), no, P_.as<cplx<T>>(),
                         bh_dP_.as<cplx<T>>(), (const cplx<T>*)bh_gP_.as<cplx<T>>(), np, rho2,
                         (const double*)bh_step_.as<double>(), int(cg), int(keep));
      if (refine)
        hipLaunchKernelGGL(k_bh_update_pos, dim3((nb + 255) / 256), dim3(256), 0, stream_,
                           offs_.as<double2>(), bh_ddf_.as<double>(),
                           (const double*)bh_gd_.as<double>(), ord, nb, pos2,
                           (const double*)bh_step_.as<double>(), int(cg), int(keep));
    });
    run(K_PROBE, [&] {
      hipLaunchKernelGGL(k_probe_terms<T>, dim3((total + 255) / 256), dim3(256), 0, stream_,
                         P_.as<cplx<T>>(), q_.as<cplx<T>>(), Wt_.as<cplx<T>>(), n, total);
    });
  }

  // One BH batch through several object slices (ptyrs/src/engine/bh.rs:130-577 with
  // SliceWaves, ptyrs/src/engine/multislice.rs:266-442): the exit waves through the slices,
  // the gradient carried back through the gaps (each slice's χ_s kept), then the tangent
  // chains along u = (g_O, ρ²·g_P) and v = D′ up through the slices, whose second-order terms
  // meet χ_s slice by slice and whose exit changes L_u, L_v go through the detector-plane pass.
  void bh_batch_ms(double rho2, const int* ord, int nb, double frac, bool cg, bool keep,
                   int slot, double pos2 = 0.0) {
    const int nt = ntiles();
    const size_t S = nslice_, no1 = nyo_ * nxo_, no = S * no1, np = nprobe_ * n;
    const int total = int(np), K = nb * M_, NF = cg ? 2 : 1;
    // FullObject offsets: every slice read through the batch's sampled windows; refined, the
    // offsets' gradient sums every slice's reads' gradient against the reads' derivatives and
    // the reads' changes take the offsets' terms (k_bh_fo_reads per slice).
    const bool fo = ms_fo();
    const bool refine = fo && pos_refine_;
    run(K_BIN, [&] {
      if (fo) {
        fo_slice_windows(ord, nb);
        return;
      }
      hipLaunchKernelGGL(k_bin_tiles<T>, dim3(nt), dim3(THREADS), 0, stream_,
                         corner_.as<int2>(), ord, nb, N, ntx_, int(cap_), list_.as<int>(),
                         lcount_.as<int>(), (const T*)nullptr, M_, int(nyo_), int(nxo_),
                         (T*)nullptr, tilemax_.as<T>(), (const cplx<T>*)nullptr);
    });
    const Waves wv = stack_waves(ord, nb);
    ms_forward(ord, nb, wv);
    run(K_PASS_A, [&] { rows(false, K); });
    BArgs<T> ba{nullptr, nullptr, nullptr, nullptr};
    ba.keep = bh_keep_.as<cplx<T>>();
    pass_b<B_BH>(ord, nb, ba);
    cplx<T>* chi = psi_.as<cplx<T>>();
    run(K_PASS_C, [&] { rows(true, K); });  // r = Uᴴ(2·td) at the last slice's output
    const size_t vn = size_t(K) * n;
    for (size_t s = S; s-- > 0;) {
      segment_for(s, wv, ord, nb);
      const auto in = wave(s, wv);
      const cplx<T>* qs = s + 1 == S ? wv.qlast : nullptr;
      PTYRS_HIP(hipMemcpyAsync(bh_chi_.as<cplx<T>>() + s * vn, chi, sizeof(cplx<T>) * vn,
                               hipMemcpyDeviceToDevice, stream_));
      run(K_OBJ_DIR, [&] {
        if (fo) {
          // Σ_m conj(q_s·in_s)·χ_s per slot, placed by the sampling's adjoint.
          hipLaunchKernelGGL(k_fo_chi_terms<T>, dim3((n + 255) / 256, nb), dim3(256), 0,
                             stream_, fo_dwin_.as<cplx<T>>(), in.first, (const cplx<T>*)chi, M_,
                             n, in.second, qs);
          if (refine) {
            cplx<T>* fd = bh_fd_.as<cplx<T>>();
            fo_object_spectrum(slice_obj(s));
            fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 1, fd, 2ll * n);
            fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 2, fd + n, 2ll * n);
            hipLaunchKernelGGL(k_bh_win_dot<T>, dim3(nb), dim3(THREADS), 0, stream_,
                               (const cplx<T>*)fo_dwin_.as<cplx<T>>(), (const cplx<T>*)fd, n,
                               bh_gd_.as<double>(), int(s + 1 < S));
          }
          fo_place_spectrum(fo_dwin_.as<cplx<T>>(), (long long)n, nb, false);
          fo_add_placed<AS_PLAIN>(bh_gO_.as<cplx<T>>() + s * no1, 1.0);
          return;
        }
        hipLaunchKernelGGL(k_lsqml_obj_dir<T>, dim3(nt), dim3(THREADS), 0, stream_,
                           bh_gO_.as<cplx<T>>() + s * no1, in.first, (const cplx<T>*)chi, N, M_,
                           corner_.as<int2>(), ord, list_.as<int>(), lcount_.as<int>(), int(cap_),
                           int(nyo_), int(nxo_), ntx_, (const T*)nullptr, (const T*)nullptr,
                           in.second, qs, 0);
      });
      if (s == 0) break;
      run(K_PASS_C, [&] {
        window_rows<MS_ADJ>(chi, s, nullptr, 0, qs, 0, ord, nb);
        gap_tail(chi, K, s - 1, true);
      });
    }
    run(K_PROBE, [&] {
      hipLaunchKernelGGL(k_ms_probe_dir<T>, dim3((n + 255) / 256, M_), dim3(256), 0, stream_,
                         bh_gP_.as<cplx<T>>(),
                         fo ? (const cplx<T>*)fo_swin(0) : (const cplx<T>*)slice_obj(0),
                         fo ? N : int(nxo_), (const cplx<T>*)nullptr, (const cplx<T>*)chi,
                         fo ? fo_tall_.as<int2>() : corner_.as<int2>(), ord, nb, N, M_, T(1), 0);
    });
    const int ndots =
        int(std::min<size_t>(kBhDotBlocks, (no + np + THREADS - 1) / THREADS));
    run(K_DOTS, [&] {
      hipLaunchKernelGGL(k_bh_dots<T>, dim3(ndots), dim3(THREADS), 0, stream_,
                         (const cplx<T>*)bh_gO_.as<cplx<T>>(),
                         (const cplx<T>*)bh_dO_.as<cplx<T>>(), no,
                         (const cplx<T>*)bh_gP_.as<cplx<T>>(),
                         (const cplx<T>*)bh_dP_.as<cplx<T>>(), np, rho2, int(cg),
                         bh_dots_.as<double>());
    });
    const int nblk = (n + 255) / 256;
    for (size_t s = 0; s < S; ++s) {
      if (s > 0) segment_for(s, wv, ord, nb);
      const auto in = wave(s, wv);
      const bool last = s + 1 == S;
      cplx<T>* out = last ? bh_l_.as<cplx<T>>() : bh_t_.as<cplx<T>>();
      run(K_HESS_A, [&] {
        // Under FullObject the directions' slices as the slots read them: their sampled
        // windows (tall, bh_fd_: u then v).
        const cplx<T>* uo = bh_gO_.as<cplx<T>>() + s * no1;
        const cplx<T>* vo = bh_dO_.as<cplx<T>>() + s * no1;
        const cplx<T>* rd = nullptr;
        if (refine) {
          // The reads' changes of slice s: the nine windows k_bh_fo_reads takes, from the
          // spectra of O_s, u_O,s and v_O,s.
          cplx<T>* fw = bh_fw_.as<cplx<T>>();
          cplx<T>* fd = bh_fd_.as<cplx<T>>();
          cplx<T>* su = bh_fs_.as<cplx<T>>();
          cplx<T>* sv = su + no1;
          const long long st9 = 9ll * n;
          fo_object_spectrum(slice_obj(s));
          fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 1, fd, 2ll * n);
          fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 2, fd + n, 2ll * n);
          for (int d = 3; d <= 5; ++d)
            fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, d, fw + (3 + d) * n, st9);
          fo_spectrum_to(uo, su);
          fo_windows<AS_PLAIN>(su, nb, 0, fw, st9);
          fo_windows<AS_PLAIN>(su, nb, 1, fw + 2 * n, st9);
          fo_windows<AS_PLAIN>(su, nb, 2, fw + 3 * n, st9);
          if (cg) {
            fo_spectrum_to(vo, sv);
            fo_windows<AS_PLAIN>(sv, nb, 0, fw + n, st9);
            fo_windows<AS_PLAIN>(sv, nb, 1, fw + 4 * n, st9);
            fo_windows<AS_PLAIN>(sv, nb, 2, fw + 5 * n, st9);
          }
          auto kr = cg ? k_bh_fo_reads<T, 2> : k_bh_fo_reads<T, 1>;
          hipLaunchKernelGGL(kr, dim3((n + 255) / 256, nb), dim3(256), 0, stream_,
                             bh_fr_.as<cplx<T>>(), (const cplx<T>*)fw, (const cplx<T>*)fd,
                             (const double*)bh_gd_.as<double>(), pos2,
                             (const double*)bh_ddf_.as<double>(), ord, n);
          rd = bh_fr_.as<cplx<T>>();
        } else if (fo) {
          cplx<T>* fw = bh_fd_.as<cplx<T>>();
          fo_object_spectrum(uo);
          fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 0, fw, (long long)n);
          if (cg) {
            fo_object_spectrum(vo);
            fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 0, fw + cap_ * n, (long long)n);
          }
          uo = fw;
          vo = fw + cap_ * n;
        }
        auto tk = cg ? k_bh_slice_tangent<T, 2> : k_bh_slice_tangent<T, 1>;
        hipLaunchKernelGGL(tk, dim3(nblk, K), dim3(256), 0, stream_, out,
                           (const cplx<T>*)bh_t_.as<cplx<T>>(),
                           (const cplx<T>*)bh_gP_.as<cplx<T>>(), T(rho2),
                           (const cplx<T>*)bh_dP_.as<cplx<T>>(), int(s == 0), in.first,
                           in.second, fo ? (const cplx<T>*)fo_swin(s) : (const cplx<T>*)slice_obj(s),
                           uo, vo, fo ? N : int(nxo_),
                           last ? wv.qlast : (const cplx<T>*)nullptr,
                           (const cplx<T>*)(bh_chi_.as<cplx<T>>() + s * vn),
                           fo ? (const int2*)fo_tall_.as<int2>() : (const int2*)corner_.as<int2>(),
                           ord, N, M_, bh_ps_.as<double>() + s * size_t(K) * nblk * 3, rd);
        if (!last) gap(bh_t_.as<cplx<T>>(), K * NF, s, false);
      });
    }
    run(K_HESS_A, [&] {
      hipLaunchKernelGGL((k_rows<T, N, false>), dim3(S::NBLK, K * NF), dim3(S::NT), 0, stream_,
                         bh_l_.as<cplx<T>>(), tw_.as<cplx<T>>());
    });
    const T norm = T(1) / (angular_ ? T(n) * T(n) : T(n));
    bh_hess_b_and_step(ord, nb, cg, norm, ndots, frac, slot,
                       refine ? (const double*)bh_gd_.as<double>() : nullptr, refine ? nb : 0,
                       pos2, bh_ps_.as<double>(), int(S) * K * nblk);
    run(K_OBJ_UPDATE, [&] {
      const size_t all = no + np;
      hipLaunchKernelGGL(k_bh_update<T>, dim3(unsigned((all + 255) / 256)), dim3(256), 0,
                         stream_, O_.as<cplx<T>>(), bh_dO_.as<cplx<T>>(),
                         (const cplx<T>*)bh_gO_.as<cplx<T>>(), no, P_.as<cplx<T>>(),
                         bh_dP_.as<cplx<T>>(), (const cplx<T>*)bh_gP_.as<cplx<T>>(), np, rho2,
                         (const double*)bh_step_.as<double>(), int(cg), int(keep));
      if (refine)
        hipLaunchKernelGGL(k_bh_update_pos, dim3((nb + 255) / 256), dim3(256), 0, stream_,
                           offs_.as<double2>(), bh_ddf_.as<double>(),
                           (const double*)bh_gd_.as<double>(), ord, nb, pos2,
                           (const double*)bh_step_.as<double>(), int(cg), int(keep));
    });
    run(K_PROBE, [&] {
      hipLaunchKernelGGL(k_probe_terms<T>, dim3((total + 255) / 256), dim3(256), 0, stream_,
                         P_.as<cplx<T>>(), q_.as<cplx<T>>(), Wt_.as<cplx<T>>(), n, total);
    });
  }

  // BH under FullObject: the spectra of u_O and v_O, the reads' derivatives ∂_a S(O) per slot,
  // the nine further reads per slot (k_bh_fo_reads' inputs) and the reads' changes.
  DevBuf bh_fs_, bh_fd_, bh_fw_, bh_fr_;
  void ensure_bh_sampled(size_t cap) {
    bh_fs_.ensure(sizeof(cplx<T>) * 2 * nyo_ * nxo_);
    bh_fd_.ensure(sizeof(cplx<T>) * 2 * cap_ * n);
    bh_fw_.ensure(sizeof(cplx<T>) * 9 * cap_ * n);
    bh_fr_.ensure(sizeof(cplx<T>) * 5 * cap_ * n);
    bh_gd_.ensure(sizeof(double) * 2 * cap);
    bh_ddf_.ensure(sizeof(double) * 2 * std::max<size_t>(1, nframe_));
  }
  // The unnormalised spectrum of object plane `obj` into dst.
  void fo_spectrum_to(const cplx<T>* obj, cplx<T>* dst) {
    const long long no = (long long)(nyo_ * nxo_);
    axis_pass<AL_PLAIN, AS_PLAIN>(
        0, false, obj, dst,
        rows_geom(no, int(nxo_), no, int(nxo_), int(nyo_), int(nxo_), int(nxo_)), 1);
    axis_pass<AL_PLAIN, AS_PLAIN>(1, false, dst, dst, cols_geom(no, int(nyo_), no, int(nyo_)), 1);
  }

  // One BH batch under FullObject sampling (ptyrs/src/engine/bh.rs:96-107, :274-577; the
  // reference GPU's bh_sampled_batch, ptyrs/src/engine/gpu/bh.rs:613-760): every slot reads
  // S(O) (the tall windows fo_win_), the probe terms are not moved; the reads' gradient per
  // slot is placed onto the object by the sampling's adjoint and, with refined offsets, dotted
  // with the reads' derivatives; the Hessian's reads change by u′, v′ and o″ (k_bh_fo_reads)
  // from the windows of the spectra of O, u_O and v_O and their derivatives.
  void bh_batch_sampled(double rho2, double pos2, const int* ord, int nb, double frac, bool cg,
                        bool keep, int slot) {
    const size_t no = nyo_ * nxo_, np = nprobe_ * n;
    const int total = int(np), K = nb * M_;
    const bool refine = pos_refine_;
    run(K_BIN, [&] {
      fo_points(ord, nb);
      fo_object_spectrum(O_.as<cplx<T>>());
      fo_object_windows(nb);
    });
    pass_a(ord, nb, false, MixArgs<T>{MIX_NONE, {}, {}, T(0), T(0)}, nullptr,
           fo_win_.as<cplx<T>>(), N, fo_tall_.as<int2>());
    BArgs<T> ba{nullptr, nullptr, nullptr, nullptr};
    ba.keep = bh_keep_.as<cplx<T>>();
    pass_b<B_BH>(ord, nb, ba);
    const int chunk = chunk_of(nb);
    const int nchunk = (nb + chunk - 1) / chunk;
    run(K_PASS_C, [&] {
      hipLaunchKernelGGL((k_pass_c_chi<T, N>), dim3(SR::NBLK, M_, nchunk), dim3(SR::NT), 0,
                         stream_, psi_.as<cplx<T>>(), fo_win_.as<cplx<T>>(), N,
                         Wt_.as<cplx<T>>(), q_.as<cplx<T>>(), fo_tall_.as<int2>(), ord, nb, M_,
                         chunk, 1, dP_.as<cplx<T>>(), tw_.as<cplx<T>>(), 0);
    });
    const cplx<T>* r = psi_.as<cplx<T>>();
    cplx<T>* gwin = fo_dwin_.as<cplx<T>>();
    cplx<T>* fd = bh_fd_.as<cplx<T>>();
    run(K_OBJ_DIR, [&] {
      hipLaunchKernelGGL(k_fo_chi_terms<T>, dim3((n + 255) / 256, nb), dim3(256), 0, stream_,
                         gwin, (const cplx<T>*)Wt_.as<cplx<T>>(), r, M_, n);
      if (refine) {
        fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 1, fd, 2ll * n);
        fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 2, fd + n, 2ll * n);
        hipLaunchKernelGGL(k_bh_win_dot<T>, dim3(nb), dim3(THREADS), 0, stream_,
                           (const cplx<T>*)gwin, (const cplx<T>*)fd, n, bh_gd_.as<double>());
      }
      fo_place_spectrum(gwin, (long long)n, nb, false);
      fo_add_placed<AS_PLAIN>(bh_gO_.p, 1.0);
    });
    run(K_PROBE, [&] {
      hipLaunchKernelGGL(k_probe_mean<T>, dim3((total + 255) / 256), dim3(256), 0, stream_,
                         bh_gP_.as<cplx<T>>(), dP_.as<cplx<T>>(), nchunk, total, 1);
    });
    const int ndots =
        int(std::min<size_t>(kBhDotBlocks, (no + np + THREADS - 1) / THREADS));
    run(K_DOTS, [&] {
      hipLaunchKernelGGL(k_bh_dots<T>, dim3(ndots), dim3(THREADS), 0, stream_,
                         (const cplx<T>*)bh_gO_.as<cplx<T>>(),
                         (const cplx<T>*)bh_dO_.as<cplx<T>>(), no,
                         (const cplx<T>*)bh_gP_.as<cplx<T>>(),
                         (const cplx<T>*)bh_dP_.as<cplx<T>>(), np, rho2, int(cg),
                         bh_dots_.as<double>());
    });
    // The reads along u and v: windows of the spectra of u_O, v_O and O.
    cplx<T>* fw = bh_fw_.as<cplx<T>>();
    cplx<T>* su = bh_fs_.as<cplx<T>>();
    cplx<T>* sv = su + no;
    const long long st9 = 9ll * n;
    run(K_HESS_A, [&] {
      fo_spectrum_to(bh_gO_.as<cplx<T>>(), su);
      fo_windows<AS_PLAIN>(su, nb, 0, fw, st9);
      if (cg) {
        fo_spectrum_to(bh_dO_.as<cplx<T>>(), sv);
        fo_windows<AS_PLAIN>(sv, nb, 0, fw + n, st9);
      }
      if (refine) {
        fo_windows<AS_PLAIN>(su, nb, 1, fw + 2 * n, st9);
        fo_windows<AS_PLAIN>(su, nb, 2, fw + 3 * n, st9);
        if (cg) {
          fo_windows<AS_PLAIN>(sv, nb, 1, fw + 4 * n, st9);
          fo_windows<AS_PLAIN>(sv, nb, 2, fw + 5 * n, st9);
        }
        for (int d = 3; d <= 5; ++d)
          fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, d, fw + (3 + d) * n, st9);
      }
      auto kr = cg ? k_bh_fo_reads<T, 2> : k_bh_fo_reads<T, 1>;
      hipLaunchKernelGGL(kr, dim3((n + 255) / 256, nb), dim3(256), 0, stream_,
                         bh_fr_.as<cplx<T>>(), (const cplx<T>*)fw, (const cplx<T>*)fd,
                         refine ? (const double*)bh_gd_.as<double>() : nullptr, pos2,
                         (const double*)bh_ddf_.as<double>(), ord, n);
      using HA = HessShape<T, N, 2>;
      auto ha = cg ? k_bh_hess_a_fo<T, N, 2> : k_bh_hess_a_fo<T, N, 1>;
      hipLaunchKernelGGL(ha, dim3(HA::NBLK, K), dim3(HA::NT), 0, stream_,
                         (const cplx<T>*)fo_win_.as<cplx<T>>(),
                         (const cplx<T>*)bh_fr_.as<cplx<T>>(), (const cplx<T>*)Wt_.as<cplx<T>>(),
                         (const cplx<T>*)bh_gP_.as<cplx<T>>(), T(rho2),
                         (const cplx<T>*)bh_dP_.as<cplx<T>>(), (const cplx<T>*)q_.as<cplx<T>>(),
                         M_, r, bh_l_.as<cplx<T>>(), bh_pa_.as<double>(),
                         (const cplx<T>*)tw_.as<cplx<T>>());
    });
    const T norm = T(1) / (angular_ ? T(n) * T(n) : T(n));
    bh_hess_b_and_step(ord, nb, cg, norm, ndots, frac, slot,
                       refine ? (const double*)bh_gd_.as<double>() : nullptr, refine ? nb : 0,
                       pos2);
    run(K_OBJ_UPDATE, [&] {
      const size_t all = no + np;
      hipLaunchKernelGGL(k_bh_update<T>, dim3(unsigned((all + 255) / 256)), dim3(256), 0,
                         stream_, O_.as<cplx<T>>(), bh_dO_.as<cplx<T>>(),
                         (const cplx<T>*)bh_gO_.as<cplx<T>>(), no, P_.as<cplx<T>>(),
                         bh_dP_.as<cplx<T>>(), (const cplx<T>*)bh_gP_.as<cplx<T>>(), np, rho2,
                         (const double*)bh_step_.as<double>(), int(cg), int(keep));
      if (refine)
        hipLaunchKernelGGL(k_bh_update_pos, dim3((nb + 255) / 256), dim3(256), 0, stream_,
                           offs_.as<double2>(), bh_ddf_.as<double>(),
                           (const double*)bh_gd_.as<double>(), ord, nb, pos2,
                           (const double*)bh_step_.as<double>(), int(cg), int(keep));
    });
    run(K_PROBE, [&] {
      hipLaunchKernelGGL(k_probe_terms<T>, dim3((total + 255) / 256), dim3(256), 0, stream_,
                         P_.as<cplx<T>>(), q_.as<cplx<T>>(), Wt_.as<cplx<T>>(), n, total);
    });
  }

  // The Hessian's detector-plane pass and the batch's step (every BH path).
  void bh_hess_b_and_step(const int* ord, int nb, bool cg, T norm, int ndots, double frac,
                          int slot, const double* gd, int npos, double pos2,
                          const double* pa = nullptr, int npa = 0) {
    run(K_HESS_B, [&] {
      // A view's fields are adjacent planes: the column passes see nb·M·NF planes, at most
      // 65535 per launch (grid y).
      if (angular_) {
        const int planes = nb * M_ * (cg ? 2 : 1);
        for (int p0 = 0; p0 < planes; p0 += 65535)
          pass_t<false>(bh_l_.as<cplx<T>>() + size_t(p0) * n, std::min(65535, planes - p0));
      }
      with_modes([&](auto mm) {
        constexpr int MM = decltype(mm)::value;
        auto hb = angular_ ? (cg ? k_bh_hess_b<T, N, MM, true, 2> : k_bh_hess_b<T, N, MM, true, 1>)
                           : (cg ? k_bh_hess_b<T, N, MM, false, 2>
                                 : k_bh_hess_b<T, N, MM, false, 1>);
        hipLaunchKernelGGL(hb, dim3(SC::NBLK, nb), dim3(SC::NT), 0, stream_,
                           (const cplx<T>*)bh_l_.as<cplx<T>>(),
                           (const cplx<T>*)bh_keep_.as<cplx<T>>(), counts_.as<float>(), ord,
                           bh_pb_.as<double>(), (const cplx<T>*)tw_.as<cplx<T>>(), norm);
      });
    });
    run(K_OTHER, [&] {
      // Two-level sums of the Hessian partials: at most kBhSumBlocks rows each, then the step.
      const int cnt[2] = {pa ? npa : nb * M_ * int(HessShape<T, N, 2>::NBLK),
                          nb * int(SC::NBLK)};
      const double* in[2] = {pa ? pa : bh_pa_.as<double>(), bh_pb_.as<double>()};
      int g[2];
      for (int i = 0; i < 2; ++i) {
        g[i] = std::min(kBhSumBlocks, (cnt[i] + THREADS - 1) / THREADS);
        const int per = (cnt[i] + g[i] - 1) / g[i];
        hipLaunchKernelGGL(k_sum_rows3, dim3(g[i]), dim3(THREADS), 0, stream_, in[i], cnt[i],
                           per, bh_sum_.as<double>() + 3 * size_t(i) * kBhSumBlocks);
      }
      hipLaunchKernelGGL(k_bh_step, dim3(1), dim3(THREADS), 0, stream_,
                         (const double*)bh_dots_.as<double>(), ndots,
                         (const double*)bh_sum_.as<double>(), g[0],
                         (const double*)bh_sum_.as<double>() + 3 * kBhSumBlocks, g[1], int(cg),
                         frac, bh_step_.as<double>(), bh_log_.as<double>() + 4 * size_t(slot),
                         gd, (const double*)bh_ddf_.as<double>(), ord, npos, pos2);
    });
  }

  static uint64_t bits_of(double v) {
    uint64_t u;
    std::memcpy(&u, &v, sizeof u);
    return u;
  }

  // One BH batch (Bh::epoch's loop body, ptyrs/src/engine/bh.rs:720-797): the gradient g at
  // the batch-start arrays, the Hessian sums along u = (g_O, ρ²·g_P) and, with cg, v = D′,
  // the step and the update; `keep` stores D as the next D′; log slot `slot`.
  void bh_batch(double rho2, const int* ord, int nb, double frac, bool cg, bool keep,
                int slot) {
    const int nt = ntiles();
    const size_t no = nyo_ * nxo_, np = nprobe_ * n;
    const int total = int(np);
    run(K_BIN, [&] {
      hipLaunchKernelGGL(k_bin_tiles<T>, dim3(nt), dim3(THREADS), 0, stream_,
                         corner_.as<int2>(), ord, nb, N, ntx_, int(cap_), list_.as<int>(),
                         lcount_.as<int>(), (const T*)nullptr, M_, int(nyo_), int(nxo_),
                         (T*)nullptr, tilemax_.as<T>(), (const cplx<T>*)nullptr);
    });
    pass_a(ord, nb, false);
    BArgs<T> ba{nullptr, nullptr, nullptr, nullptr};
    ba.keep = bh_keep_.as<cplx<T>>();
    pass_b<B_BH>(ord, nb, ba);
    const int chunk = chunk_of(nb);
    const int nchunk = (nb + chunk - 1) / chunk;
    run(K_PASS_C, [&] {
      hipLaunchKernelGGL((k_pass_c_chi<T, N>), dim3(SR::NBLK, M_, nchunk), dim3(SR::NT), 0,
                         stream_, psi_.as<cplx<T>>(), O_.as<cplx<T>>(), int(nxo_),
                         Wt_.as<cplx<T>>(), q_.as<cplx<T>>(), corner_.as<int2>(), ord, nb, M_,
                         chunk, 1, dP_.as<cplx<T>>(), tw_.as<cplx<T>>(), 0);
    });
    run(K_OBJ_DIR, [&] {
      hipLaunchKernelGGL(k_lsqml_obj_dir<T>, dim3(nt), dim3(THREADS), 0, stream_,
                         bh_gO_.as<cplx<T>>(), Wt_.as<cplx<T>>(), psi_.as<cplx<T>>(), N, M_,
                         corner_.as<int2>(), ord, list_.as<int>(), lcount_.as<int>(), int(cap_),
                         int(nyo_), int(nxo_), ntx_, (const T*)nullptr, (const T*)nullptr);
    });
    run(K_PROBE, [&] {
      hipLaunchKernelGGL(k_probe_mean<T>, dim3((total + 255) / 256), dim3(256), 0, stream_,
                         bh_gP_.as<cplx<T>>(), dP_.as<cplx<T>>(), nchunk, total, 1);
    });
    const int ndots =
        int(std::min<size_t>(kBhDotBlocks, (no + np + THREADS - 1) / THREADS));
    run(K_DOTS, [&] {
      hipLaunchKernelGGL(k_bh_dots<T>, dim3(ndots), dim3(THREADS), 0, stream_,
                         (const cplx<T>*)bh_gO_.as<cplx<T>>(),
                         (const cplx<T>*)bh_dO_.as<cplx<T>>(), no,
                         (const cplx<T>*)bh_gP_.as<cplx<T>>(),
                         (const cplx<T>*)bh_dP_.as<cplx<T>>(), np, rho2, int(cg),
                         bh_dots_.as<double>());
    });
    const T norm = T(1) / (angular_ ? T(n) * T(n) : T(n));
    run(K_HESS_A, [&] {
      using HA = HessShape<T, N, 2>;
      auto ha = cg ? k_bh_hess_a<T, N, 2> : k_bh_hess_a<T, N, 1>;
      hipLaunchKernelGGL(ha, dim3(HA::NBLK, nb * M_), dim3(HA::NT), 0, stream_,
                         (const cplx<T>*)O_.as<cplx<T>>(),
                         (const cplx<T>*)bh_gO_.as<cplx<T>>(),
                         (const cplx<T>*)bh_dO_.as<cplx<T>>(), int(nxo_),
                         (const cplx<T>*)Wt_.as<cplx<T>>(),
                         (const cplx<T>*)bh_gP_.as<cplx<T>>(), T(rho2),
                         (const cplx<T>*)bh_dP_.as<cplx<T>>(), (const cplx<T>*)q_.as<cplx<T>>(),
                         (const int2*)corner_.as<int2>(), ord, M_,
                         (const cplx<T>*)psi_.as<cplx<T>>(), bh_l_.as<cplx<T>>(),
                         bh_pa_.as<double>(), (const cplx<T>*)tw_.as<cplx<T>>());
    });
    bh_hess_b_and_step(ord, nb, cg, norm, ndots, frac, slot, nullptr, 0, 0.0);
    run(K_OBJ_UPDATE, [&] {
      const size_t all = no + np;
      hipLaunchKernelGGL(k_bh_update<T>, dim3(unsigned((all + 255) / 256)), dim3(256), 0,
                         stream_, O_.as<cplx<T>>(), bh_dO_.as<cplx<T>>(),
                         (const cplx<T>*)bh_gO_.as<cplx<T>>(), no, P_.as<cplx<T>>(),
                         bh_dP_.as<cplx<T>>(), (const cplx<T>*)bh_gP_.as<cplx<T>>(), np, rho2,
                         (const double*)bh_step_.as<double>(), int(cg), int(keep));
    });
    run(K_PROBE, [&] {
      hipLaunchKernelGGL(k_probe_terms<T>, dim3((total + 255) / 256), dim3(256), 0, stream_,
                         P_.as<cplx<T>>(), q_.as<cplx<T>>(), Wt_.as<cplx<T>>(), n, total);
    });
  }

  static constexpr int kPreBlocks = 1024;
  // Workgroups per frame of the per-frame reductions.
  static int aux_blocks() { return std::max(1, n / 4096); }

  // Preconditioner update at an epoch's start (ptyrs/src/engine/lsqml.rs:671-699):
  // illumination of every view at the current probe, then pre = (pre + illum)/2, the first
  // from illum·max_f Σ iobs_f/(4·n·k); floor = (0.1·max pre)².
  void precondition(bool first) {
    const size_t no = nyo_ * nxo_;
    const int nt = ntiles();
    // Up to kIllumChunk frames' views at once: one tile list and one illumination pass each.
    const int step = int(illum_chunk());
    if (full_object())
      run(K_OTHER, [&] { fo_illumination(); });
    else
      PTYRS_HIP(hipMemsetAsync(illum_.p, 0, sizeof(T) * no, stream_));
    for (int f0 = 0; f0 < int(nframe_) && !full_object(); f0 += step) {
      const int nb = std::min(step, int(nframe_) - f0);
      const int* ord = iota_.as<int>() + f0;
      run(K_BIN, [&] {
        hipLaunchKernelGGL(k_bin_tiles<T>, dim3(nt), dim3(THREADS), 0, stream_,
                           corner_.as<int2>(), ord, nb, N, ntx_, step, list_all_.as<int>(),
                           count_all_.as<int>(), (const T*)nullptr, M_, int(nyo_), int(nxo_),
                           (T*)nullptr, tilemax_.as<T>(),
                           (const cplx<T>*)nullptr);
      });
      run(K_OTHER, [&] {
        hipLaunchKernelGGL(k_illumination<T>, dim3(nt), dim3(THREADS), 0, stream_,
                           illum_.as<T>(), P_.as<cplx<T>>(), N, M_, corner_.as<int2>(), ord,
                           list_all_.as<int>(), count_all_.as<int>(), step, int(nyo_),
                           int(nxo_), ntx_);
      });
    }
    const double init = max_frame_total_ / (4.0 * double(n) * scale_);
    run(K_OTHER, [&] {
      hipLaunchKernelGGL(k_pre_update<T>, dim3(kPreBlocks), dim3(THREADS), 0, stream_,
                         pre_.as<T>(), illum_.as<T>(), no, int(first), init, pre_part_.as<T>());
      hipLaunchKernelGGL(k_pre_floor<T>, dim3(1), dim3(THREADS), 0, stream_, pre_part_.as<T>(),
                         kPreBlocks, floor_.as<T>());
    });
  }

  // One LSQML batch (ptyrs/src/engine/lsqml.rs:571-643), single slice, no features.
  void lsqml_batch(const ptyrs_lsqml_params& p, const int* ord, int nb, int off) {
    const int nt = ntiles();
    const bool gauss = p.cost == PTYRS_AMPLITUDE;
    run(K_BIN, [&] {
      hipLaunchKernelGGL(k_bin_tiles<T>, dim3(nt), dim3(THREADS), 0, stream_,
                         corner_.as<int2>(), ord, nb, N, ntx_, int(cap_), list_.as<int>(),
                         lcount_.as<int>(), (const T*)nullptr, M_, int(nyo_), int(nxo_),
                         (T*)nullptr, tilemax_.as<T>(),
                           (const cplx<T>*)nullptr);
    });
    pass_a(ord, nb, false);
    BArgs<T> ba{afar_tmp_.as<T>(), prev_.as<T>(), xi_part_.as<double>(), icalc_.as<T>()};
    if (gauss) {
      pass_b<B_GAUSS>(ord, nb, ba);
    } else {
      const int xb = aux_blocks();
      run(K_OTHER, [&] {
        hipLaunchKernelGGL(k_lsqml_prev<T>, dim3(1), dim3(THREADS), 0, stream_,
                           afar_all_.as<T>(), ord, nb, prev_.as<T>());
      });
      pass_b<B_ICALC>(ord, nb, ba);
      run(K_OTHER, [&] {
        hipLaunchKernelGGL(k_lsqml_xi<T>, dim3(1), dim3(THREADS), 0, stream_,
                           xi_part_.as<double>(), int(SB::NBLK), nb, prev_.as<T>(),
                           afar_tmp_.as<T>(), prev_.as<T>() + 1);
        hipLaunchKernelGGL(k_lsqml_xi_sums<T>, dim3(xb, nb), dim3(THREADS), 0, stream_,
                           icalc_.as<T>(), counts_.as<float>(), ord, n, prev_.as<T>() + 1,
                           xi_part2_.as<double>());
        hipLaunchKernelGGL(k_lsqml_xi<T>, dim3(1), dim3(THREADS), 0, stream_,
                           xi_part2_.as<double>(), xb, nb, prev_.as<T>() + 1, afar_tmp_.as<T>(),
                           (T*)nullptr);
      });
      pass_b<B_POISSON>(ord, nb, ba);
    }
    run(K_OTHER, [&] {
      hipLaunchKernelGGL(k_lsqml_log_afar<T>, dim3((nb + 255) / 256), dim3(256), 0, stream_,
                         afar_tmp_.as<T>(), ord, nb, afar_all_.as<T>(),
                         log_.as<double>() + 3 * size_t(off), int(gauss), 3);
    });
    const int chunk = chunk_of(nb);
    const int nchunk = (nb + chunk - 1) / chunk;
    run(K_PASS_C, [&] {
      hipLaunchKernelGGL((k_pass_c_chi<T, N>), dim3(SR::NBLK, M_, nchunk), dim3(SR::NT), 0,
                         stream_, psi_.as<cplx<T>>(), O_.as<cplx<T>>(), int(nxo_),
                         Wt_.as<cplx<T>>(), q_.as<cplx<T>>(), corner_.as<int2>(), ord, nb, M_,
                         chunk, 1, dP_.as<cplx<T>>(), tw_.as<cplx<T>>());
    });
    run(K_OBJ_DIR, [&] {
      hipLaunchKernelGGL(k_lsqml_obj_dir<T>, dim3(nt), dim3(THREADS), 0, stream_,
                         dO_.as<cplx<T>>(), Wt_.as<cplx<T>>(), psi_.as<cplx<T>>(), N, M_,
                         corner_.as<int2>(), ord, list_.as<int>(), lcount_.as<int>(), int(cap_),
                         int(nyo_), int(nxo_), ntx_, pre_.as<T>(), floor_.as<T>());
    });
    const int total = int(nprobe_ * n);
    const int xb = aux_blocks();
    // λ0 = 1.2e-7·k/n per pixel and mode (ptyrs/src/engine/lsqml.rs:31-33), summed over both.
    const double lam = 1.2e-7 * scale_ * double(M_);
    run(K_PROBE, [&] {
      hipLaunchKernelGGL(k_probe_mean<T>, dim3((total + 255) / 256), dim3(256), 0, stream_,
                         dPm_.as<cplx<T>>(), dP_.as<cplx<T>>(), nchunk, total, nb);
    });
    run(K_DOTS, [&] {
      hipLaunchKernelGGL(k_lsqml_dots<T>, dim3(xb, nb), dim3(THREADS), 0, stream_,
                         dO_.as<cplx<T>>(), O_.as<cplx<T>>(), int(nxo_), Wt_.as<cplx<T>>(),
                         dPm_.as<cplx<T>>(), q_.as<cplx<T>>(), psi_.as<cplx<T>>(), N, M_,
                         corner_.as<int2>(), ord, dots_part_.as<double>());
      hipLaunchKernelGGL(k_lsqml_step<T>, dim3(1), dim3(1024), 0, stream_,
                         dots_part_.as<double>(), xb, nb, lam, M_, sums_.as<double>(),
                         log_.as<double>() + 3 * size_t(off), step_.as<double>(), 3);
    });
    const size_t no = nyo_ * nxo_;
    run(K_OBJ_UPDATE, [&] {
      hipLaunchKernelGGL(k_lsqml_apply_obj<T>, dim3((no + 255) / 256), dim3(256), 0, stream_,
                         O_.as<cplx<T>>(), dO_.as<cplx<T>>(), step_.as<double>(), no);
    });
    run(K_PROBE, [&] {
      hipLaunchKernelGGL(k_lsqml_apply_probe<T>, dim3((total + 255) / 256), dim3(256), 0,
                         stream_, P_.as<cplx<T>>(), Wt_.as<cplx<T>>(), q_.as<cplx<T>>(),
                         dPm_.as<cplx<T>>(), step_.as<double>(), n, total);
    });
  }

  // The cost of the current state over every frame and its gradients 2·∂L/∂O_s* (every slice)
  // and 2·∂L/∂P*: the adjoint of the exit-wave model (window, probe, slices, plane phase, exit
  // transform), L the Poisson LLK sum Σ(icalc − iobs + iobs·ln(iobs/icalc)) (icalc where
  // iobs = 0) or the amplitude cost Σ(√icalc − √iobs)² over unmasked pixels.
  // ptyrs/src/engine/ptycho.rs:492-561, ptyrs/src/engine/features.rs:310-330,
  // ptyrs/src/engine/multislice.rs:266-442
  void grad(int cost, double* value, double* og, double* pg) override {
    if (nobj_ != 1) throw Invalid("the GPU gradient runs one object mode");
    if (M_ > kMaxModes) throw Invalid(modes_error());
    if (pos_on_) throw Invalid("the GPU gradient models no position offsets");
    if (cost != PTYRS_POISSON && cost != PTYRS_AMPLITUDE)
      throw Invalid("cost must be amplitude or poisson");
    const size_t stack = std::min<size_t>(std::max<size_t>(nframe_, 1), kDefaultStack);
    ensure_workspace(stack, 1);
    ensure_lsqml(stack);
    probe_terms();
    const size_t S = nslice_, no = nyo_ * nxo_;
    double v[4] = {0, 0, 0, 0};
    if (nframe_) {
      llk_pass(0);
      PTYRS_HIP(hipMemcpyAsync(v, hist_.p, sizeof(double) * 4, hipMemcpyDeviceToHost, stream_));
    } else {
      PTYRS_HIP(hipMemsetAsync(dO_.p, 0, sizeof(cplx<T>) * S * no, stream_));
      PTYRS_HIP(hipMemsetAsync(dPm_.p, 0, sizeof(cplx<T>) * nprobe_ * n, stream_));
    }
    const int nt = ntiles();
    const Waves wv{P_.as<cplx<T>>(), 0, q_.as<cplx<T>>(), 0};
    const cplx<T>* q0 = S == 1 ? q_.as<cplx<T>>() : nullptr;
    cplx<T>* chi = psi_.as<cplx<T>>();
    for (size_t f0 = 0; f0 < nframe_; f0 += stack) {
      const int nb = int(std::min(stack, nframe_ - f0));
      const int* ord = iota_.as<int>() + f0;
      const int K = nb * M_;
      const int acc = f0 > 0;
      run(K_BIN, [&] {
        hipLaunchKernelGGL(k_bin_tiles<T>, dim3(nt), dim3(THREADS), 0, stream_,
                           corner_.as<int2>(), ord, nb, N, ntx_, int(cap_), list_.as<int>(),
                           lcount_.as<int>(), (const T*)nullptr, M_, int(nyo_), int(nxo_),
                           (T*)nullptr, tilemax_.as<T>(),
                           (const cplx<T>*)nullptr);
      });
      ms_forward(ord, nb, wv);
      run(K_PASS_A, [&] { rows(false, K); });
      if (cost == PTYRS_POISSON)
        pass_b<B_GRAD>(ord, nb);
      else
        pass_b<B_GRAD_AMP>(ord, nb);
      run(K_PASS_C, [&] { rows(true, K); });
      for (size_t s = S; s-- > 0;) {
        segment_for(s, wv, ord, nb);
        const auto in = wave(s, wv);
        const cplx<T>* qs = s + 1 == S ? q_.as<cplx<T>>() : nullptr;
        run(K_OBJ_UPDATE, [&] {
          hipLaunchKernelGGL(k_lsqml_obj_dir<T>, dim3(nt), dim3(THREADS), 0, stream_,
                             slice_dir(s), in.first, (const cplx<T>*)chi, N, M_,
                             corner_.as<int2>(), ord, list_.as<int>(), lcount_.as<int>(),
                             int(cap_), int(nyo_), int(nxo_), ntx_, (const T*)nullptr,
                             (const T*)nullptr, in.second, qs, 0, acc);
        });
        if (s == 0) break;
        run(K_PASS_C, [&] {
          window_rows<MS_ADJ>(chi, s, nullptr, 0, qs, 0, ord, nb);
          gap_tail(chi, K, s - 1, true);
        });
      }
      run(K_PROBE, [&] {
        hipLaunchKernelGGL(k_ms_probe_dir<T>, dim3((n + 255) / 256, M_), dim3(256), 0, stream_,
                           dPm_.as<cplx<T>>(), (const cplx<T>*)slice_obj(0), int(nxo_), q0,
                           (const cplx<T>*)chi, corner_.as<int2>(), ord, nb, N, M_, T(1), acc);
      });
    }
    download_complex(dO_, og, S * no);
    download_complex(dPm_, pg, nprobe_ * n);
    flush_timing();
    *value = cost == PTYRS_POISSON ? v[0] : v[2] / 4.0;
  }

  // The model's detector intensities Σ_m |U ψ_m|² of every frame at the current state (what
  // the counts are compared with), detector order — pty-chi's forward-model output `y_pred`
  // (ptyrs/tests/engine_lsqml_multislice.rs:146-182).
  void intensities(double* out) override {
    if (M_ > kMaxModes) throw Invalid(modes_error());
    if (nframe_ == 0) return;
    const size_t stack = std::min<size_t>(nframe_, kDefaultStack);
    ensure_workspace(stack, 1);
    ensure_lsqml(stack);
    const bool fo = full_object();
    if (fo) fo_ensure(cap_, true);
    probe_terms();
    if (fo && nslice_ == 1) run(K_OTHER, [&] { fo_object_spectrum(O_.as<cplx<T>>()); });
    zero(prev_.as<T>(), 2);
    std::vector<T> ic(stack * n);
    for (size_t f0 = 0; f0 < nframe_; f0 += stack) {
      const int nb = int(std::min(stack, nframe_ - f0));
      const int* ord = iota_.as<int>() + f0;
      if (fo && nslice_ == 1) {
        run(K_OTHER, [&] {
          fo_points(ord, nb);
          fo_object_windows(nb);
        });
        pass_a(ord, nb, false, MixArgs<T>{MIX_NONE, {}, {}, T(0), T(0)}, nullptr,
               fo_win_.as<cplx<T>>(), N, fo_tall_.as<int2>());
      } else if (nslice_ > 1 || pos_on_) {
        if (fo) run(K_OTHER, [&] { fo_slice_windows(ord, nb); });
        ms_forward(ord, nb, stack_waves(ord, nb));
        run(K_PASS_A, [&] { rows(false, nb * M_); });
      } else {
        pass_a(ord, nb, false);
      }
      BArgs<T> ba{afar_tmp_.as<T>(), prev_.as<T>(), xi_part_.as<double>(), icalc_.as<T>()};
      pass_b<B_ICALC>(ord, nb, ba);
      PTYRS_HIP(hipMemcpyAsync(ic.data(), icalc_.p, sizeof(T) * nb * n, hipMemcpyDeviceToHost,
                               stream_));
      PTYRS_HIP(hipStreamSynchronize(stream_));
      const int h = dft_plane_ ? N / 2 : 0;
      for (int b = 0; b < nb; ++b)
        for (int ky = 0; ky < N; ++ky)
          for (int kx = 0; kx < N; ++kx) {
            const int dy = reverse_ ? (N - ky) % N : ky, dx = reverse_ ? (N - kx) % N : kx;
            out[(f0 + b) * size_t(n) + size_t((dy + h) % N) * N + (dx + h) % N] =
                double(ic[size_t(b) * n + size_t(ky) * N + kx]);
          }
    }
    flush_timing();
  }

  // One LSQML batch of the general model (ptyrs/src/engine/lsqml.rs:571-643 over
  // ptyrs/src/engine/multislice.rs): any number of slices, the probe planes moved by the
  // frames' offsets under PATCH. The exit waves through the slices, the detector operator as
  // for one slice, χ back through the slices with each slice's object direction and step
  // (Eq. 23b) on the way, slice 0 jointly with the probe (Eq. 22), the position directions on
  // slice S/2, then every update from the state on entry and the position step. One object
  // mode; unfused (one kernel per stage).
  // Under FullObject (several slices) every slice is read through the batch's sampled windows
  // (fo_slice_windows; ptyrs/src/engine/multislice.rs:70-171 with the sampler): the slice
  // directions' terms are placed by the sampling's adjoint, their step sums read the
  // directions' windows, the position sums the windows' derivatives, and the probe does not
  // move.
  void lsqml_batch_ms(const ptyrs_lsqml_params& p, const int* ord, int nb, int off) {
    const int nt = ntiles();
    const bool gauss = p.cost == PTYRS_AMPLITUDE;
    const size_t S = nslice_;
    const int ls = int(2 + S);
    const int K = nb * M_;
    const bool fo = ms_fo();
    const long long nol = (long long)(nyo_ * nxo_);
    run(K_BIN, [&] {
      if (fo) {
        fo_slice_windows(ord, nb);
        return;
      }
      hipLaunchKernelGGL(k_bin_tiles<T>, dim3(nt), dim3(THREADS), 0, stream_,
                         corner_.as<int2>(), ord, nb, N, ntx_, int(cap_), list_.as<int>(),
                         lcount_.as<int>(), (const T*)nullptr, M_, int(nyo_), int(nxo_),
                         (T*)nullptr, tilemax_.as<T>(),
                           (const cplx<T>*)nullptr);
    });
    const Waves wv = stack_waves(ord, nb);
    ms_forward(ord, nb, wv);
    run(K_PASS_A, [&] { rows(false, K); });
    BArgs<T> ba{afar_tmp_.as<T>(), prev_.as<T>(), xi_part_.as<double>(), icalc_.as<T>()};
    const int xb = aux_blocks();
    if (gauss) {
      pass_b<B_GAUSS>(ord, nb, ba);
    } else {
      run(K_OTHER, [&] {
        hipLaunchKernelGGL(k_lsqml_prev<T>, dim3(1), dim3(THREADS), 0, stream_,
                           afar_all_.as<T>(), ord, nb, prev_.as<T>());
      });
      pass_b<B_ICALC>(ord, nb, ba);
      run(K_OTHER, [&] {
        hipLaunchKernelGGL(k_lsqml_xi<T>, dim3(1), dim3(THREADS), 0, stream_,
                           xi_part_.as<double>(), int(SB::NBLK), nb, prev_.as<T>(),
                           afar_tmp_.as<T>(), prev_.as<T>() + 1);
        hipLaunchKernelGGL(k_lsqml_xi_sums<T>, dim3(xb, nb), dim3(THREADS), 0, stream_,
                           icalc_.as<T>(), counts_.as<float>(), ord, n, prev_.as<T>() + 1,
                           xi_part2_.as<double>());
        hipLaunchKernelGGL(k_lsqml_xi<T>, dim3(1), dim3(THREADS), 0, stream_,
                           xi_part2_.as<double>(), xb, nb, prev_.as<T>() + 1, afar_tmp_.as<T>(),
                           (T*)nullptr);
      });
      pass_b<B_POISSON>(ord, nb, ba);
    }
    double* log = log_.as<double>() + size_t(ls) * off;
    double* step = step_.as<double>();
    run(K_OTHER, [&] {
      hipLaunchKernelGGL(k_lsqml_log_afar<T>, dim3((nb + 255) / 256), dim3(256), 0, stream_,
                         afar_tmp_.as<T>(), ord, nb, afar_all_.as<T>(), log, int(gauss), ls);
    });
    cplx<T>* chi = psi_.as<cplx<T>>();
    run(K_PASS_C, [&] {
      rows(true, K);
      auto in = wave(S - 1, wv);
      window<MS_SUB>(chi, S - 1, in.first, in.second, wv.qlast, wv.q_frame, ord, nb);
    });
    const size_t sp = S / 2;  // pty-chi's slice_for_correction
    for (size_t s = S; s-- > 0;) {
      segment_for(s, wv, ord, nb);
      const auto in = wave(s, wv);
      const cplx<T>* qs = s + 1 == S ? wv.qlast : nullptr;
      const int qf = s + 1 == S ? wv.q_frame : 0;
      run(K_OBJ_UPDATE, [&] {
        if (fo) {
          // Σ_m conj(q_s·in_s)·χ_s on each window, placed, then preconditioned.
          hipLaunchKernelGGL(k_fo_chi_terms<T>, dim3((n + 255) / 256, nb), dim3(256), 0,
                             stream_, fo_dwin_.as<cplx<T>>(), in.first, (const cplx<T>*)chi, M_,
                             n, in.second, qs);
          fo_place_spectrum(fo_dwin_.as<cplx<T>>(), (long long)n, nb, false);
          fo_add_placed<AS_PLAIN>(slice_dir(s), 1.0);
          hipLaunchKernelGGL(k_fo_precondition<T>, dim3(unsigned((nol + 255) / 256)), dim3(256),
                             0, stream_, slice_dir(s), (const T*)pre_.as<T>(),
                             (const T*)floor_.as<T>(), nol);
          return;
        }
        hipLaunchKernelGGL(k_lsqml_obj_dir<T>, dim3(nt), dim3(THREADS), 0, stream_,
                           slice_dir(s), in.first, (const cplx<T>*)chi, N, M_, corner_.as<int2>(),
                           ord, list_.as<int>(), lcount_.as<int>(), int(cap_), int(nyo_),
                           int(nxo_), ntx_, pre_.as<T>(), floor_.as<T>(), in.second, qs, qf);
      });
      if (pos_refine_ && s == sp) {
        if (fo)
          fo_position_sums(s, in.first, in.second, qs, qf, ord, nb);
        else
          position_sums(s, in.first, qs, qf, ord, nb);
      }
      if (s == 0) break;
      run(K_OTHER, [&] {
        if (fo) fo_dir_windows(slice_dir(s), nb);
        hipLaunchKernelGGL(k_ms_dots<T>, dim3(xb, nb), dim3(THREADS), 0, stream_,
                           fo ? (const cplx<T>*)fo_dwin_.as<cplx<T>>()
                              : (const cplx<T>*)slice_dir(s),
                           fo ? N : int(nxo_), in.first, qs, (const cplx<T>*)chi, N, M_,
                           fo ? fo_tall_.as<int2>() : corner_.as<int2>(), ord,
                           xi_part2_.as<double>(), qf);
        hipLaunchKernelGGL(k_ms_slice_step, dim3(1), dim3(1024), 0, stream_,
                           xi_part2_.as<double>(), xb, nb, int(S) * M_, log, ls, int(1 + s),
                           step + 1 + s);
      });
      run(K_PASS_C, [&] {
        window_rows<MS_ADJ>(chi, s, nullptr, 0, qs, qf, ord, nb);
        gap_tail(chi, K, s - 1, true);
      });
    }
    // Slice 0: the probe direction (moved back per frame with offsets), then the joint step.
    const cplx<T>* q0 = S == 1 ? q_.as<cplx<T>>() : nullptr;
    if (pos_on_ && !fo) {
      cplx<T>* g = gspec_.as<cplx<T>>();
      run(K_PROBE, [&] {
        PTYRS_HIP(hipMemcpyAsync(g, chi, sizeof(cplx<T>) * size_t(K) * n,
                                 hipMemcpyDeviceToDevice, stream_));
        window<MS_ADJ>(g, 0, nullptr, 0, nullptr, 0, ord, nb);  // conj(O_0)·χ_0 = q0·g
        hipLaunchKernelGGL((k_rows<T, N, false>), dim3(S::NBLK, K), dim3(S::NT), 0, stream_, g,
                           tw_.as<cplx<T>>());
        hipLaunchKernelGGL((k_cols<T, N, false>), dim3(SC::NBLK, K), dim3(SC::NT), 0, stream_, g,
                           tw_.as<cplx<T>>());
        hipLaunchKernelGGL(k_unramp_mean<T>, dim3((n + 255) / 256, M_), dim3(256), 0, stream_,
                           dPm_.as<cplx<T>>(), (const cplx<T>*)g,
                           (const double2*)offs_.as<double2>(), ord, nb, N, M_,
                           1.0 / (double(n) * double(nb)));
        hipLaunchKernelGGL((k_cols<T, N, true>), dim3(SC::NBLK, M_), dim3(SC::NT), 0, stream_,
                           dPm_.as<cplx<T>>(), tw_.as<cplx<T>>());
        hipLaunchKernelGGL((k_rows<T, N, true>), dim3(S::NBLK, M_), dim3(S::NT), 0, stream_,
                           dPm_.as<cplx<T>>(), tw_.as<cplx<T>>());
        if (q0)
          hipLaunchKernelGGL((k_times_q<T, true>), dim3((n + 255) / 256, M_), dim3(256), 0,
                             stream_, dPm_.as<cplx<T>>(), q0, n);
        if (opr_on_) {
          // Each frame's mode-0 direction moved back, for OPR's residues.
          cplx<T>* mv = opr_m_.as<cplx<T>>();
          hipLaunchKernelGGL(k_unramp_planes<T>, dim3((n + 255) / 256, nb), dim3(256), 0,
                             stream_, mv, (const cplx<T>*)g, (const double2*)offs_.as<double2>(),
                             ord, N, M_, 0, 1.0 / double(n));
          hipLaunchKernelGGL((k_cols<T, N, true>), dim3(SC::NBLK, nb), dim3(SC::NT), 0, stream_,
                             mv, tw_.as<cplx<T>>());
          hipLaunchKernelGGL((k_rows<T, N, true>), dim3(S::NBLK, nb), dim3(S::NT), 0, stream_,
                             mv, tw_.as<cplx<T>>());
          if (q0)
            hipLaunchKernelGGL((k_times_q<T, true>), dim3((n + 255) / 256, nb), dim3(256), 0,
                               stream_, mv, q0, n);
        }
      });
    } else {
      run(K_PROBE, [&] {
        hipLaunchKernelGGL(k_ms_probe_dir<T>, dim3((n + 255) / 256, M_), dim3(256), 0, stream_,
                           dPm_.as<cplx<T>>(),
                           fo ? (const cplx<T>*)fo_swin(0) : (const cplx<T>*)slice_obj(0),
                           fo ? N : int(nxo_), q0, (const cplx<T>*)chi,
                           fo ? fo_tall_.as<int2>() : corner_.as<int2>(), ord, nb, N, M_,
                           T(1) / T(nb), 0);
      });
    }
    const int total = int(nprobe_ * n);
    const double lam = 1.2e-7 * scale_ * double(M_);
    run(K_OTHER, [&] {
      if (fo) fo_dir_windows(slice_dir(0), nb);
      hipLaunchKernelGGL(k_lsqml_dots<T>, dim3(xb, nb), dim3(THREADS), 0, stream_,
                         fo ? (const cplx<T>*)fo_dwin_.as<cplx<T>>()
                            : (const cplx<T>*)slice_dir(0),
                         fo ? (const cplx<T>*)fo_swin(0) : (const cplx<T>*)slice_obj(0),
                         fo ? N : int(nxo_), wv.w0, (const cplx<T>*)dPm_.as<cplx<T>>(),
                         q0 ? q0 : (const cplx<T>*)ones_.as<cplx<T>>(), (const cplx<T>*)chi, N,
                         M_, fo ? fo_tall_.as<int2>() : corner_.as<int2>(), ord,
                         dots_part_.as<double>(), wv.w0_view, q0);
      hipLaunchKernelGGL(k_lsqml_step<T>, dim3(1), dim3(1024), 0, stream_,
                         dots_part_.as<double>(), xb, nb, lam, int(S) * M_, sums_.as<double>(),
                         log, step, ls);
    });
    // OPR's eigenmodes and weights from slice 0's χ and the state on entry.
    if (opr_on_) opr_step(ord, nb);
    const size_t no = nyo_ * nxo_;
    run(K_OBJ_UPDATE, [&] {
      for (size_t s = 0; s < S; ++s)
        hipLaunchKernelGGL(k_lsqml_apply_obj<T>, dim3((no + 255) / 256), dim3(256), 0, stream_,
                           slice_obj(s), (const cplx<T>*)slice_dir(s),
                           (const double*)(s == 0 ? step : step + 1 + s), no);
    });
    run(K_PROBE, [&] {
      hipLaunchKernelGGL(k_lsqml_apply_probe<T>, dim3((total + 255) / 256), dim3(256), 0,
                         stream_, P_.as<cplx<T>>(), Wt_.as<cplx<T>>(), q_.as<cplx<T>>(),
                         dPm_.as<cplx<T>>(), (const double*)step, n, total);
    });
    if (pos_refine_) {
      const double eps = 1e-6 * scale_;  // EPS_POSITION·k (ptyrs/src/engine/features.rs:455)
      run(K_OTHER, [&] {
        hipLaunchKernelGGL(k_pos_step, dim3(1), dim3(1024), 0, stream_,
                           (const double*)pos_part_.as<double>(), xb, nb, int(nframe_), eps, ord,
                           offs_.as<double2>(), pos_sums_.as<double>());
      });
    }
  }

  // The position sums of the batch's frames on slice s (ptyrs/src/engine/lsqml.rs:230-270):
  // the windows' Fourier derivatives against the wave entering the slice times its phase and χ_s.
  void position_sums(size_t s, const cplx<T>* w, const cplx<T>* q, int q_frame, const int* ord,
                     int nb, const cplx<T>* chi = nullptr) {
    if (chi == nullptr) chi = psi_.as<cplx<T>>();
    cplx<T>* spec = pgrad_.as<cplx<T>>();
    cplx<T>* grad = spec + cap_ * n;
    run(K_OTHER, [&] {
      hipLaunchKernelGGL(k_patch<T>, dim3((n + 255) / 256, nb), dim3(256), 0, stream_, spec,
                         (const cplx<T>*)slice_obj(s), int(nxo_), corner_.as<int2>(), ord, N);
      hipLaunchKernelGGL((k_rows<T, N, false>), dim3(S::NBLK, nb), dim3(S::NT), 0, stream_, spec,
                         tw_.as<cplx<T>>());
      hipLaunchKernelGGL((k_cols<T, N, false>), dim3(SC::NBLK, nb), dim3(SC::NT), 0, stream_,
                         spec, tw_.as<cplx<T>>());
      hipLaunchKernelGGL(k_deriv<T>, dim3((n + 255) / 256, nb), dim3(256), 0, stream_, grad,
                         (const cplx<T>*)spec, N, 1.0 / double(n));
      hipLaunchKernelGGL((k_cols<T, N, true>), dim3(SC::NBLK, 2 * nb), dim3(SC::NT), 0, stream_,
                         grad, tw_.as<cplx<T>>());
      hipLaunchKernelGGL((k_rows<T, N, true>), dim3(S::NBLK, 2 * nb), dim3(S::NT), 0, stream_,
                         grad, tw_.as<cplx<T>>());
      hipLaunchKernelGGL(k_pos_sums<T>, dim3(aux_blocks(), nb), dim3(THREADS), 0, stream_,
                         (const cplx<T>*)grad, w, q, q_frame, chi, N, M_,
                         pos_part_.as<double>(), 1);
    });
  }

  static constexpr int kAbsBlocks = 256;
  size_t stack_size_ = 0;
  // LLK records the device holds before the host drains them (GpuOptions::llk_slots).
  size_t hist_slots_ = 8;

  // Chunks of a short last stack hold no partials: zero them so every stack's slab sums.
  void zero_partials(int s, int from, int nchunk) {
    const size_t cn = size_t(M_) * n;
    zero(pnum_part_.as<cplx<T>>() + (size_t(s) * nchunk + from) * cn, (nchunk - from) * cn);
    zero(pden_part_.as<T>() + (size_t(s) * nchunk + from) * n, (nchunk - from) * size_t(n));
  }

  // count elements at p set to 0 by a fill kernel (capturable, unlike a memset).
  template <typename U>
  void zero(U* p, size_t count) {
    const size_t blocks = std::min<size_t>(1024, (count + 255) / 256);
    hipLaunchKernelGGL(k_zero<U>, dim3(unsigned(std::max<size_t>(blocks, 1))), dim3(256), 0,
                       stream_, p, count);
  }

  // Object update of AP/DM/RAAR: tile lists over every frame, norm = Σ_m D_m, its max, blend.
  template <bool H>
  void object_blend(double inertia, Carried<T, H> cw) {
    const int nt = ntiles();
    run(K_BIN, [&] {
      norm2_planes();
      hipLaunchKernelGGL(k_bin_tiles<T>, dim3(nt), dim3(THREADS), 0, stream_,
                         corner_.as<int2>(), iota_.as<int>(), int(nframe_), N, ntx_,
                         int(nframe_), list_all_.as<int>(), count_all_.as<int>(),
                         (const T*)W2_.as<T>(), M_, int(nyo_), int(nxo_), D_.as<T>(),
                         tilemax_.as<T>(),
                           (const cplx<T>*)nullptr);
    });
    run(K_OVERLAP, [&] {
      hipLaunchKernelGGL(k_max_reduce<T>, dim3(1), dim3(MAX_THREADS), 0, stream_,
                         tilemax_.as<T>(), nt, pmax_.as<T>());
    });
    run(K_OBJ_UPDATE, [&] {
      hipLaunchKernelGGL((k_obj_blend<T, H>), dim3(nt), dim3(THREADS), 0, stream_,
                         O_.as<cplx<T>>(), Wt_.as<cplx<T>>(), D_.as<T>(), cw, N, M_,
                         corner_.as<int2>(), iota_.as<int>(),
                         list_all_.as<int>(), count_all_.as<int>(), int(nframe_), int(nyo_),
                         int(nxo_), ntx_, pmax_.as<T>(), T(inertia));
    });
  }

  // Probe sums of the carried waves against the current object (DM, RAAR).
  template <bool H>
  void probe_sums_all(size_t stack, int nstack, int nchunk, Carried<T, H> cw) {
    for (int s = 0; s < nstack; ++s) {
      const int f0 = int(s * stack);
      const int nb = std::min<int>(int(stack), int(nframe_) - f0);
      const int ck = chunk_of(nb);
      const int nck = (nb + ck - 1) / ck;
      run(K_PROBE, [&] {
        hipLaunchKernelGGL((k_probe_sums<T, H>), dim3((n + THREADS - 1) / THREADS, M_, nck),
                           dim3(THREADS), 0, stream_, cw, O_.as<cplx<T>>(),
                           int(nxo_), q_.as<cplx<T>>(), corner_.as<int2>(),
                           iota_.as<int>() + f0, nb, N, M_, ck,
                           pnum_part_.as<cplx<T>>() + size_t(s) * nchunk * M_ * n,
                           pden_part_.as<T>() + size_t(s) * nchunk * n);
      });
      if (nck < nchunk) zero_partials(s, nck, nchunk);
    }
  }

  void probe_blend(int tot_chunks, double inertia) {
    const int nblk = (n + THREADS - 1) / THREADS;
    const int total = int(nprobe_ * n);
    run(K_PROBE, [&] {
      hipLaunchKernelGGL(k_probe_reduce<T>, dim3(nblk), dim3(THREADS), 0, stream_,
                         pnum_part_.as<cplx<T>>(), pden_part_.as<T>(), tot_chunks, n, M_,
                         pnum_.as<cplx<T>>(), pden_.as<T>(), pmaxpart_.as<T>());
      hipLaunchKernelGGL(k_max_reduce<T>, dim3(1), dim3(MAX_THREADS), 0, stream_,
                         pmaxpart_.as<T>(), nblk, pmax2_.as<T>());
      hipLaunchKernelGGL(k_probe_blend<T>, dim3((total + 255) / 256), dim3(256), 0, stream_,
                         P_.as<cplx<T>>(), Wt_.as<cplx<T>>(), q_.as<cplx<T>>(),
                         pnum_.as<cplx<T>>(), pden_.as<T>(), pmax2_.as<T>(), T(inertia), n,
                         total);
    });
  }

  // split_scale(s) (ptyrs/src/engine/ptycho.rs:583-591); with `calc` (device, the model photons
  // over unmasked pixels) s = √(Σ iobs / calc), scale_obj_probe (:172-182).
  void balance(double s, const double* calc = nullptr) {
    run(K_OTHER, [&] {
      hipLaunchKernelGGL(k_abs_sums<T>, dim3(kAbsBlocks), dim3(THREADS), 0, stream_,
                         O_.as<cplx<T>>(), nobj_ * nyo_ * nxo_, P_.as<cplx<T>>(), nprobe_ * n,
                         abs_part_.as<double>());
      hipLaunchKernelGGL(k_balance<T>, dim3(kAbsBlocks), dim3(THREADS), 0, stream_,
                         O_.as<cplx<T>>(), nobj_ * nyo_ * nxo_, P_.as<cplx<T>>(),
                         Wt_.as<cplx<T>>(), q_.as<cplx<T>>(), n, nprobe_ * n,
                         abs_part_.as<double>(), kAbsBlocks, s, calc, total_);
    });
  }

  // zero_phase_ramp (ptyrs/src/engine/ptycho.rs:593-603, 1134-1181) on the device: the
  // probe's spectrum (row and column DFTs of a copy), its centre of mass in FP64, then opposite
  // linear phase ramps on the probe and the object, enqueued on the engine's stream.
  DevBuf zspec_, zpart_, zd_;
  static constexpr int kRampBlocks = 64;
  void device_zero_phase_ramp() {
    const size_t np = nprobe_ * n;
    zspec_.ensure(sizeof(cplx<T>) * np);
    zpart_.ensure(sizeof(double) * 3 * kRampBlocks);
    zd_.ensure(sizeof(double) * 2);
    run(K_OTHER, [&] {
      PTYRS_HIP(hipMemcpyAsync(zspec_.p, P_.p, sizeof(cplx<T>) * np, hipMemcpyDeviceToDevice,
                               stream_));
      hipLaunchKernelGGL((k_rows<T, N, false>), dim3(S::NBLK, nprobe_), dim3(S::NT), 0, stream_,
                         zspec_.as<cplx<T>>(), tw_.as<cplx<T>>());
      hipLaunchKernelGGL((k_cols<T, N, false>), dim3(SC::NBLK, nprobe_), dim3(SC::NT), 0,
                         stream_, zspec_.as<cplx<T>>(), tw_.as<cplx<T>>());
      hipLaunchKernelGGL(k_spectrum_moments<T>, dim3(kRampBlocks), dim3(THREADS), 0, stream_,
                         (const cplx<T>*)zspec_.as<cplx<T>>(), N, int(nprobe_),
                         zpart_.as<double>());
      hipLaunchKernelGGL(k_ramp_offset, dim3(1), dim3(THREADS), 0, stream_,
                         (const double*)zpart_.as<double>(), kRampBlocks, N, zd_.as<double>());
      hipLaunchKernelGGL(k_apply_ramp<T>, dim3(int((np + 255) / 256)), dim3(256), 0, stream_,
                         P_.as<cplx<T>>(), int(nprobe_), N, N, (const double*)zd_.as<double>(),
                         -1.0, N);
      const size_t no = nobj_ * nyo_ * nxo_;
      hipLaunchKernelGGL(k_apply_ramp<T>, dim3(int(std::min<size_t>((no + 255) / 256, 8192))),
                         dim3(256), 0, stream_, O_.as<cplx<T>>(), int(nobj_), int(nyo_),
                         int(nxo_), (const double*)zd_.as<double>(), 1.0, N);
    });
  }

  int ntiles() const { return ntx_ * nty_; }

  // Pass C chunking: enough workgroups to fill the chip, few partial probe sums.
  int chunk_of(int nb) const {
    const int per = SR::NBLK * M_;
    int nchunk = std::max(1, std::min(nb, (chi_blocks_ + per - 1) / per));
    return (nb + nchunk - 1) / nchunk;
  }
  // The same for rPIE's mode-summing pass C, whose workgroups hold G rows of every mode.
  int chunk_of_m(int nb) const {
    const int per = N / rows_per_group(N, SR::L, int(M_));
    int nchunk = std::max(1, std::min(nb, (pass_c_blocks_ + per - 1) / per));
    return (nb + nchunk - 1) / nchunk;
  }
  int pass_c_blocks_ = kPassCBlocks, chi_blocks_ = kChiBlocks;
  static constexpr int kChiBlocks = 1024;
  static constexpr int kMaxLaunchBlocks = 8192;
  uint64_t launch_key() const override {
    return (uint64_t(pass_c_blocks_) << 32) | uint64_t(chi_blocks_);
  }
  void set_launch(size_t stack_size, int pass_c_blocks, int chi_blocks) override {
    if (pass_c_blocks < 0 || pass_c_blocks > kMaxLaunchBlocks || chi_blocks < 0 ||
        chi_blocks > kMaxLaunchBlocks)
      throw Invalid("pass C workgroup targets must be in 0.." + std::to_string(kMaxLaunchBlocks));
    if (stack_size * size_t(M_) > 65535)
      throw Invalid("a stack of " + std::to_string(stack_size) + " frames of " +
                    std::to_string(M_) + " modes holds more than 65535 views");
    stack_size_ = stack_size;
    pass_c_blocks_ = pass_c_blocks ? pass_c_blocks : kPassCBlocks;
    chi_blocks_ = chi_blocks ? chi_blocks : kChiBlocks;
  }
  void get_launch(size_t* stack_size, int* pass_c_blocks, int* chi_blocks) const override {
    if (stack_size) *stack_size = stack_size_;
    if (pass_c_blocks) *pass_c_blocks = pass_c_blocks_;
    if (chi_blocks) *chi_blocks = chi_blocks_;
  }

  // Scratch layout of llk_noalloc: the per-workgroup sums (4 doubles per frame and pass-B
  // workgroup), the 4 result sums, then the stack, each part 256-byte aligned.
  static size_t align256(size_t b) { return (b + 255) / 256 * 256; }
  size_t llk_part_bytes() const { return align256(sizeof(double) * 4 * nframe_ * SB::NBLK); }
  size_t llk_scratch_bytes(size_t frames) const override {
    return llk_part_bytes() + 256 + sizeof(cplx<T>) * frames * M_ * n;
  }

  void llk_noalloc(void* scratch, size_t bytes, double out[4]) override {
    if (nslice_ > 1 || pos_on_)
      throw Invalid("the allocation-free LLK runs one slice without position offsets");
    if (profiling_)
      throw Invalid("the allocation-free LLK does not run with profiling on (it creates events)");
    const size_t head = llk_part_bytes() + 256;
    const size_t per_frame = sizeof(cplx<T>) * M_ * n;
    if (bytes < head + per_frame)
      throw Invalid("the allocation-free LLK needs " + std::to_string(llk_scratch_bytes(1)) +
                    " bytes of scratch for this scan (" + std::to_string(nframe_) +
                    " frames); it has " + std::to_string(bytes));
    // Stacks as long as the scratch holds, at most 65535 views per launch.
    const size_t frames = std::min<size_t>({(bytes - head) / per_frame,
                                            std::max<size_t>(nframe_, 1),
                                            size_t(65535) / size_t(M_)});
    char* base = static_cast<char*>(scratch);
    double* result = reinterpret_cast<double*>(base + llk_part_bytes());
    struct Restore {
      Engine* e;
      ~Restore() {
        e->stack_override_ = nullptr;
        e->part_override_ = nullptr;
      }
    } restore{this};
    part_override_ = reinterpret_cast<double*>(base);
    stack_override_ = reinterpret_cast<cplx<T>*>(base + head);
    probe_terms();
    llk_pass_to(result, frames);
    PTYRS_HIP(hipMemcpyAsync(out, result, sizeof(double) * 4, hipMemcpyDeviceToHost, stream_));
    PTYRS_HIP(hipStreamSynchronize(stream_));
    // Llk::normalised (ptyrs/src/engine/ops.rs:268-276): nb_photons_calc stays a total.
    for (int i = 0; i < 3; ++i) out[i] /= double(nb_obs_);
  }

  void ensure_workspace(size_t cap, size_t nepoch) {
    cap = std::max<size_t>(cap, 1);
    // A stack's views fill one launch's grid y (the reference's limit, GpuOptions::stack_size).
    if (cap * size_t(M_) > 65535)
      throw Invalid("a stack of " + std::to_string(cap) + " frames of " + std::to_string(M_) +
                    " modes holds more than 65535 views");
    psi_.ensure(sizeof(cplx<T>) * cap * M_ * n);
    omax_.ensure(sizeof(T) * cap * S::NBLK);
    // chunk_of and chunk_of_m never cut more chunks than this, whatever the batch length.
    const size_t per =
        std::min(size_t(SR::NBLK) * M_, size_t(N / rows_per_group(N, SR::L, int(M_))));
    const size_t nchunk =
        std::min(cap, (size_t(std::max(pass_c_blocks_, chi_blocks_)) + per - 1) / per);
    dP_.ensure(sizeof(cplx<T>) * std::max<size_t>(nchunk, 1) * M_ * n);
    list_.ensure(sizeof(int) * size_t(ntiles()) * cap);
    lcount_.ensure(sizeof(int) * size_t(ntiles()));
    tilemax_.ensure(sizeof(T) * size_t(ntiles()));
    D_.ensure(sizeof(T) * size_t(M_) * nyo_ * nxo_);
    W2_.ensure(sizeof(T) * size_t(M_) * n);
    pmax_.ensure(sizeof(T) * 1);
    order_.ensure(sizeof(int) * std::max<size_t>(1, nepoch * nframe_));
    eord_.ensure(sizeof(int) * std::max<size_t>(1, nframe_));
    llk_part_.ensure(sizeof(double) * 4 * nframe_ * SB::NBLK);
    cap_ = std::max(cap_, cap);
    if (nslice_ > 1) inc_.ensure(sizeof(cplx<T>) * inc_slots() * cap_ * M_ * n);
    ensure_positions(cap_);
  }

  // |W_m|² of the current probe terms into W2_, for the per-mode overlaps.
  void norm2_planes() {
    const int total = int(M_) * n;
    hipLaunchKernelGGL(k_norm2_planes<T>, dim3((total + 255) / 256), dim3(256), 0, stream_,
                       (const cplx<T>*)Wt_.as<cplx<T>>(), W2_.as<T>(), total);
  }

  void probe_terms() {
    const int total = int(nprobe_ * n);
    run(K_OTHER, [&] {
      hipLaunchKernelGGL(k_probe_terms<T>, dim3((total + 255) / 256), dim3(256), 0, stream_,
                         P_.as<cplx<T>>(), q_.as<cplx<T>>(), Wt_.as<cplx<T>>(), n, total);
    });
  }

  // Pass A; `waves` (nullable) gives every view its own wave (plane view of it) instead of
  // the probe terms Wt.
  template <bool H = false>
  // `obj`, `obj_nxo` and `corners` (nullable) replace the object plane, its width and the
  // frames' corners (FullObject's tall windows).
  void pass_a(const int* ord, int nb, bool omax,
              MixArgs<T, H> mix = MixArgs<T, H>{MIX_NONE, {}, {}, T(0), T(0)},
              const cplx<T>* waves = nullptr, const cplx<T>* obj = nullptr, int obj_nxo = 0,
              const int2* corners = nullptr) {
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(S::NBLK, nb * M_), dim3(S::NT), 0, stream_,
                         obj ? obj : O_.as<cplx<T>>(), obj ? obj_nxo : int(nxo_),
                         waves ? waves : Wt_.as<cplx<T>>(),
                         corners ? corners : corner_.as<int2>(), ord, M_, stack(),
                         omax ? omax_.as<T>() : nullptr, tw_.as<cplx<T>>(), mix,
                         int(waves != nullptr));
    };
    run(K_PASS_A, [&] {
      switch (mix.kind) {
        case MIX_DM: go(k_pass_a<T, N, MIX_DM, H>); break;
        case MIX_RAAR: go(k_pass_a<T, N, MIX_RAAR, H>); break;
        case MIX_RAAR0: go(k_pass_a<T, N, MIX_RAAR0, H>); break;
        default: go(k_pass_a<T, N, MIX_NONE, H>); break;
      }
    });
  }

  template <int MODE>
  void pass_b(const int* ord, int nb, BArgs<T> ba = BArgs<T>{nullptr, nullptr, nullptr, nullptr}) {
    // Far field and Fresnel: the column DFTs around the detector operator, 1/n for the round
    // trip. Angular spectrum: columns, T, columns; then the row DFTs around the detector
    // operator, 1/n² for two round trips; then the inverse column passes with conj T
    // (nothing to carry back for B_LLK and B_ICALC). B_POISSON runs on the stack B_ICALC left
    // behind, whose forward column passes are done.
    constexpr bool back = MODE != B_LLK && MODE != B_ICALC;
    constexpr bool fwd = MODE != B_POISSON;
    run(MODE == B_LLK ? K_LLK : K_PASS_B, [&] {
      if (angular_ && fwd) pass_t<false>(stack(), nb * M_);
      const T norm = T(1) / (angular_ ? T(n) * T(n) : T(n));
      auto args = [&](auto mm) {
        constexpr int MM = decltype(mm)::value;
        auto col = k_pass_b<T, N, MM, MODE, false>;
        auto row = k_pass_b<T, N, MM, MODE, true>;
        hipLaunchKernelGGL(angular_ ? row : col, dim3(SB::NBLK, nb),
                           dim3(pass_b_threads<T, N, MM>()), 0, stream_, stack(),
                           counts_.as<float>(), ord, llk_parts(), tw_.as<cplx<T>>(),
                           ba, norm);
      };
      with_modes(args);
      if (angular_ && back) pass_t<true>(stack(), nb * M_);
    });
  }

  // One forward pass of every frame and the four LLK sums into history slot `slot`.
  void llk_pass(int slot) { llk_pass_to(hist_.as<double>() + 4 * slot, cap_); }

  // The same, `step_frames` frames per stack, the sums into `result`. Every frame's
  // per-workgroup sums are kept and reduced in one fixed order, so the stack length changes
  // no bit.
  void llk_pass_to(double* result, size_t step_frames) {
    const int step = int(std::max<size_t>(step_frames, 1));
    const bool fo = full_object();
    if (fo && nslice_ == 1) run(K_OTHER, [&] { fo_object_spectrum(O_.as<cplx<T>>()); });
    for (int f0 = 0; f0 < int(nframe_); f0 += step) {
      const int nb = std::min(step, int(nframe_) - f0);
      const int* ord = iota_.as<int>() + f0;
      if (fo && nslice_ == 1) {
        // FullObject: the frames' sampled windows, read at their tall corners.
        run(K_OTHER, [&] {
          fo_points(ord, nb);
          fo_object_windows(nb);
        });
        pass_a(ord, nb, false, MixArgs<T>{MIX_NONE, {}, {}, T(0), T(0)},
               opr_on_ ? opr_waves_sampled(ord, nb) : nullptr, fo_win_.as<cplx<T>>(), N,
               fo_tall_.as<int2>());
      } else if (nslice_ > 1 || pos_on_ || opr_on_) {
        if (fo) run(K_OTHER, [&] { fo_slice_windows(ord, nb); });
        ms_forward(ord, nb, stack_waves(ord, nb));
        run(K_PASS_A, [&] { rows(false, nb * M_); });
      } else {
        pass_a(ord, nb, false);
      }
      pass_b<B_LLK>(ord, nb);
    }
    run(K_LLK, [&] {
      hipLaunchKernelGGL(k_llk_reduce, dim3(1), dim3(THREADS), 0, stream_, llk_parts(),
                         int(nframe_ * SB::NBLK), result);
    });
  }

  // The exit-wave stack and the LLK's per-workgroup sums: the engine's workspace, or the
  // caller's scratch while llk_noalloc runs.
  cplx<T>* stack_override_ = nullptr;
  double* part_override_ = nullptr;
  cplx<T>* stack() const { return stack_override_ ? stack_override_ : psi_.as<cplx<T>>(); }
  double* llk_parts() const { return part_override_ ? part_override_ : llk_part_.as<double>(); }

  void drain(std::vector<std::pair<size_t, int>>& due, int algorithm) {
    if (due.empty()) return;
    std::vector<double> h(4 * hist_slots_);
    PTYRS_HIP(hipMemcpyAsync(h.data(), hist_.p, sizeof(double) * 4 * hist_slots_,
                             hipMemcpyDeviceToHost, stream_));
    PTYRS_HIP(hipStreamSynchronize(stream_));
    const double nb = double(nb_obs_);
    for (auto& d : due) {
      const double* v = h.data() + 4 * d.second;
      history.push_back(Record{d.first, algorithm, {v[0] / nb, v[1] / nb, v[2] / nb, v[3]}});
    }
    due.clear();
  }

  // One rPIE batch (ptyrs/src/engine/ptycho.rs:900-1086) from the batch-start object and probe.
  void rpie_batch(const ptyrs_rpie_params& p, const int* ord, int nb) {
    if (full_object()) {
      rpie_batch_sampled(p, ord, nb);
      return;
    }
    if (pos_on_) {
      rpie_batch_moved(p, ord, nb);
      return;
    }
    const bool upd_obj = p.has_obj_step != 0, upd_probe = p.has_probe_step != 0;
    const int nt = ntiles();
    if (upd_obj) {
      // W2_ holds |W_m|² of the batch-start probe terms (rpie_begin, then each batch's
      // k_rpie_updates; without probe updates the probe does not change).
      run(K_BIN, [&] {
        hipLaunchKernelGGL(k_bin_tiles<T>, dim3(nt), dim3(THREADS), 0, stream_,
                           corner_.as<int2>(), ord, nb, N, ntx_, int(cap_), list_.as<int>(),
                           lcount_.as<int>(), (const T*)W2_.as<T>(), M_, int(nyo_), int(nxo_),
                           D_.as<T>(), tilemax_.as<T>(),
                           (const cplx<T>*)nullptr);
      });
      run(K_OVERLAP, [&] {
        hipLaunchKernelGGL(k_max_reduce<T>, dim3(1), dim3(MAX_THREADS), 0, stream_,
                           tilemax_.as<T>(), nt, pmax_.as<T>());
      });
    }
    pass_a(ord, nb, upd_probe);
    pass_b<B_RPIE>(ord, nb);
    const int chunk = chunk_of_m(nb);
    const int nchunk = (nb + chunk - 1) / chunk;
    run(K_PASS_C, [&] {
      auto launch = [&](auto mm) {
        constexpr int MM = decltype(mm)::value;
        constexpr int G = rows_per_group(N, SR::L, MM);
        hipLaunchKernelGGL((k_pass_c_rpie_sum<T, N, MM>), dim3(N / G, nchunk),
                           dim3(pass_c_sum_threads<T, N, MM>()), 0, stream_, psi_.as<cplx<T>>(), O_.as<cplx<T>>(), int(nxo_),
                           Wt_.as<cplx<T>>(), q_.as<cplx<T>>(), corner_.as<int2>(), ord, nb,
                           omax_.as<T>(), int(S::NBLK), chunk, T(p.probe_alpha),
                           int(upd_probe), dP_.as<cplx<T>>(), tw_.as<cplx<T>>(), int(upd_obj),
                           (const T*)D_.as<T>(), int(nyo_), (const T*)pmax_.as<T>(),
                           T(p.obj_alpha), 0, (cplx<T>*)nullptr);
      };
      with_modes(launch);
    });
    if (upd_obj && upd_probe) {
      const int total = int(nprobe_ * n);
      run(K_OBJ_UPDATE, [&] {
        hipLaunchKernelGGL(k_rpie_updates<T>, dim3(nt + (total + THREADS - 1) / THREADS),
                           dim3(THREADS), 0, stream_, O_.as<cplx<T>>(),
                           (const cplx<T>*)psi_.as<cplx<T>>(), N, M_, corner_.as<int2>(), ord,
                           list_.as<int>(), lcount_.as<int>(), int(cap_), int(nyo_), int(nxo_),
                           ntx_, T(p.obj_step), nt, P_.as<cplx<T>>(), Wt_.as<cplx<T>>(),
                           q_.as<cplx<T>>(), dP_.as<cplx<T>>(), nchunk, total,
                           T(p.probe_step / nb), W2_.as<T>());
      });
      return;
    }
    if (upd_obj) {
      run(K_OBJ_UPDATE, [&] {
        hipLaunchKernelGGL(k_obj_update_rpie_sum<T>, dim3(nt), dim3(THREADS), 0, stream_,
                           O_.as<cplx<T>>(), (const cplx<T>*)psi_.as<cplx<T>>(), N, M_,
                           corner_.as<int2>(), ord, list_.as<int>(), lcount_.as<int>(),
                           int(cap_), int(nyo_), int(nxo_), ntx_, T(p.obj_step));
      });
    }
    if (upd_probe) {
      const int total = int(nprobe_ * n);
      run(K_PROBE, [&] {
        hipLaunchKernelGGL(k_probe_apply<T>, dim3((total + 255) / 256), dim3(256), 0, stream_,
                           P_.as<cplx<T>>(), Wt_.as<cplx<T>>(), q_.as<cplx<T>>(),
                           dP_.as<cplx<T>>(), nchunk, n, total, T(p.probe_step / nb));
      });
    }
  }

  // One rPIE batch under position offsets (PATCH; ptyrs/src/engine/ptycho.rs:900-1086 with
  // Features::position, the reference GPU's rpie under its position model,
  // ptyrs/src/engine/gpu/rpie.rs:69-133): every view's wave W = DFT⁻¹(DFT(P·q)·exp(−2πi f·d))
  // (its frame's offset d) takes the place of the probe terms in the overlaps, pass A and pass
  // C; pass C runs one slot per chunk, so its probe-direction partials are the frames' own,
  // which are moved back by their offsets in Fourier space (conj(q)·DFT⁻¹(DFT(q·g)·exp(+2πi
  // f·d))) and averaged; with refinement, mode 0's W and Δ give the position sums on the
  // batch-start object, then the MAD-clipped SGD step (after the probe's move back, which
  // reads the batch-start offsets).
  void rpie_batch_moved(const ptyrs_rpie_params& p, const int* ord, int nb) {
    const bool upd_obj = p.has_obj_step != 0, upd_probe = p.has_probe_step != 0;
    const int nt = ntiles();
    const int K = nb * M_;
    moved_probes(ord, nb, false);
    const cplx<T>* W = inc0_.as<cplx<T>>();
    if (upd_obj) {
      run(K_BIN, [&] {
        hipLaunchKernelGGL(k_bin_tiles<T>, dim3(nt), dim3(THREADS), 0, stream_,
                           corner_.as<int2>(), ord, nb, N, ntx_, int(cap_), list_.as<int>(),
                           lcount_.as<int>(), (const T*)nullptr, M_, int(nyo_), int(nxo_),
                           D_.as<T>(), tilemax_.as<T>(), W);
      });
      run(K_OVERLAP, [&] {
        hipLaunchKernelGGL(k_max_reduce<T>, dim3(1), dim3(MAX_THREADS), 0, stream_,
                           tilemax_.as<T>(), nt, pmax_.as<T>());
      });
    }
    pass_a(ord, nb, upd_probe, MixArgs<T>{MIX_NONE, {}, {}, T(0), T(0)}, W);
    pass_b<B_RPIE>(ord, nb);
    cplx<T>* chi0 = pos_refine_ ? gspec_.as<cplx<T>>() : nullptr;
    run(K_PASS_C, [&] {
      auto launch = [&](auto mm) {
        constexpr int MM = decltype(mm)::value;
        constexpr int G = rows_per_group(N, SR::L, MM);
        hipLaunchKernelGGL((k_pass_c_rpie_sum<T, N, MM>), dim3(N / G, nb),
                           dim3(pass_c_sum_threads<T, N, MM>()), 0, stream_, psi_.as<cplx<T>>(),
                           O_.as<cplx<T>>(), int(nxo_), W, q_.as<cplx<T>>(), corner_.as<int2>(),
                           ord, nb, omax_.as<T>(), int(S::NBLK), 1, T(p.probe_alpha),
                           int(upd_probe), dP_.as<cplx<T>>(), tw_.as<cplx<T>>(), int(upd_obj),
                           (const T*)D_.as<T>(), int(nyo_), (const T*)pmax_.as<T>(),
                           T(p.obj_alpha), 1, chi0);
      };
      with_modes(launch);
    });
    if (pos_refine_) position_sums(0, W, nullptr, 0, ord, nb, chi0);
    if (upd_obj) {
      run(K_OBJ_UPDATE, [&] {
        hipLaunchKernelGGL(k_obj_update_rpie_sum<T>, dim3(nt), dim3(THREADS), 0, stream_,
                           O_.as<cplx<T>>(), (const cplx<T>*)psi_.as<cplx<T>>(), N, M_,
                           corner_.as<int2>(), ord, list_.as<int>(), lcount_.as<int>(),
                           int(cap_), int(nyo_), int(nxo_), ntx_, T(p.obj_step));
      });
    }
    if (upd_probe) {
      cplx<T>* g = dP_.as<cplx<T>>();
      const int total = int(nprobe_ * n);
      run(K_PROBE, [&] {
        hipLaunchKernelGGL((k_times_q<T, false>), dim3((n + 255) / 256, K), dim3(256), 0,
                           stream_, g, (const cplx<T>*)q_.as<cplx<T>>(), n);
        hipLaunchKernelGGL((k_rows<T, N, false>), dim3(S::NBLK, K), dim3(S::NT), 0, stream_, g,
                           tw_.as<cplx<T>>());
        hipLaunchKernelGGL((k_cols<T, N, false>), dim3(SC::NBLK, K), dim3(SC::NT), 0, stream_, g,
                           tw_.as<cplx<T>>());
        hipLaunchKernelGGL(k_unramp_mean<T>, dim3((n + 255) / 256, M_), dim3(256), 0, stream_,
                           dPm_.as<cplx<T>>(), (const cplx<T>*)g,
                           (const double2*)offs_.as<double2>(), ord, nb, N, M_,
                           1.0 / (double(n) * double(nb)));
        hipLaunchKernelGGL((k_cols<T, N, true>), dim3(SC::NBLK, M_), dim3(SC::NT), 0, stream_,
                           dPm_.as<cplx<T>>(), tw_.as<cplx<T>>());
        hipLaunchKernelGGL((k_rows<T, N, true>), dim3(S::NBLK, M_), dim3(S::NT), 0, stream_,
                           dPm_.as<cplx<T>>(), tw_.as<cplx<T>>());
        hipLaunchKernelGGL((k_times_q<T, true>), dim3((n + 255) / 256, M_), dim3(256), 0,
                           stream_, dPm_.as<cplx<T>>(), (const cplx<T>*)q_.as<cplx<T>>(), n);
        hipLaunchKernelGGL(k_probe_apply<T>, dim3((total + 255) / 256), dim3(256), 0, stream_,
                           P_.as<cplx<T>>(), Wt_.as<cplx<T>>(), q_.as<cplx<T>>(),
                           (const cplx<T>*)dPm_.as<cplx<T>>(), 1, n, total, T(p.probe_step));
      });
    }
    if (pos_refine_) {
      const double eps = 1e-6 * scale_;  // EPS_POSITION·k (ptyrs/src/engine/features.rs:455)
      run(K_OTHER, [&] {
        hipLaunchKernelGGL(k_pos_step, dim3(1), dim3(1024), 0, stream_,
                           (const double*)pos_part_.as<double>(), aux_blocks(), nb, int(nframe_),
                           eps, ord, offs_.as<double2>(), pos_sums_.as<double>());
      });
    }
  }

  // One LSQML batch under offsets and FullObject sampling, one slice (ptyrs/src/engine/
  // lsqml.rs:97-350 with SliceWaves' sampler, ptyrs/src/engine/multislice.rs:84-171, :385-442):
  // the probe stays put and every slot reads its sampled window (the tall fo_win_); χ as without
  // offsets; the object direction Σ_m conj(W_m)·χ_m on each window placed onto the object by the
  // sampling's adjoint, then preconditioned; the probe direction read against the windows and
  // not moved; the step sizes from the direction's sampled windows; the position directions
  // from the windows' derivatives over the position against W = P·q and mode 0's χ.
  void lsqml_batch_sampled(const ptyrs_lsqml_params& p, const int* ord, int nb, int off) {
    const bool gauss = p.cost == PTYRS_AMPLITUDE;
    const long long no = (long long)(nyo_ * nxo_);
    run(K_BIN, [&] {
      fo_points(ord, nb);
      fo_object_spectrum(O_.as<cplx<T>>());
      fo_object_windows(nb);
    });
    // The waves: the shared probe terms, or with OPR each slot's own (W_view).
    const cplx<T>* W = opr_on_ ? opr_waves_sampled(ord, nb) : Wt_.as<cplx<T>>();
    const int wview = opr_on_ ? 1 : 0;
    pass_a(ord, nb, false, MixArgs<T>{MIX_NONE, {}, {}, T(0), T(0)}, opr_on_ ? W : nullptr,
           fo_win_.as<cplx<T>>(), N, fo_tall_.as<int2>());
    BArgs<T> ba{afar_tmp_.as<T>(), prev_.as<T>(), xi_part_.as<double>(), icalc_.as<T>()};
    const int xb = aux_blocks();
    if (gauss) {
      pass_b<B_GAUSS>(ord, nb, ba);
    } else {
      run(K_OTHER, [&] {
        hipLaunchKernelGGL(k_lsqml_prev<T>, dim3(1), dim3(THREADS), 0, stream_,
                           afar_all_.as<T>(), ord, nb, prev_.as<T>());
      });
      pass_b<B_ICALC>(ord, nb, ba);
      run(K_OTHER, [&] {
        hipLaunchKernelGGL(k_lsqml_xi<T>, dim3(1), dim3(THREADS), 0, stream_,
                           xi_part_.as<double>(), int(SB::NBLK), nb, prev_.as<T>(),
                           afar_tmp_.as<T>(), prev_.as<T>() + 1);
        hipLaunchKernelGGL(k_lsqml_xi_sums<T>, dim3(xb, nb), dim3(THREADS), 0, stream_,
                           icalc_.as<T>(), counts_.as<float>(), ord, n, prev_.as<T>() + 1,
                           xi_part2_.as<double>());
        hipLaunchKernelGGL(k_lsqml_xi<T>, dim3(1), dim3(THREADS), 0, stream_,
                           xi_part2_.as<double>(), xb, nb, prev_.as<T>() + 1, afar_tmp_.as<T>(),
                           (T*)nullptr);
      });
      pass_b<B_POISSON>(ord, nb, ba);
    }
    run(K_OTHER, [&] {
      hipLaunchKernelGGL(k_lsqml_log_afar<T>, dim3((nb + 255) / 256), dim3(256), 0, stream_,
                         afar_tmp_.as<T>(), ord, nb, afar_all_.as<T>(),
                         log_.as<double>() + 3 * size_t(off), int(gauss), 3);
    });
    const int chunk = chunk_of(nb);
    const int nchunk = (nb + chunk - 1) / chunk;
    run(K_PASS_C, [&] {
      hipLaunchKernelGGL((k_pass_c_chi<T, N>), dim3(SR::NBLK, M_, nchunk), dim3(SR::NT), 0,
                         stream_, psi_.as<cplx<T>>(), fo_win_.as<cplx<T>>(), N, W,
                         q_.as<cplx<T>>(), fo_tall_.as<int2>(), ord, nb, M_, chunk, 1,
                         dP_.as<cplx<T>>(), tw_.as<cplx<T>>(), 1, wview);
    });
    const cplx<T>* chi = psi_.as<cplx<T>>();
    if (pos_refine_) {
      // The windows' derivatives from the batch-start object's spectrum (still in fo_spec_).
      cplx<T>* grad = pgrad_.as<cplx<T>>() + cap_ * n;
      run(K_OTHER, [&] {
        fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 1, grad, 2ll * n);
        fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 2, grad + n, 2ll * n);
        hipLaunchKernelGGL(k_pos_sums<T>, dim3(xb, nb), dim3(THREADS), 0, stream_,
                           (const cplx<T>*)grad, W, (const cplx<T>*)nullptr, 0, chi, N, M_,
                           pos_part_.as<double>(), wview);
      });
    }
    run(K_OBJ_DIR, [&] {
      hipLaunchKernelGGL(k_fo_chi_terms<T>, dim3((n + 255) / 256, nb), dim3(256), 0, stream_,
                         fo_dwin_.as<cplx<T>>(), W, chi, M_, n, wview);
      fo_place_spectrum(fo_dwin_.as<cplx<T>>(), (long long)n, nb, false);
      fo_add_placed<AS_PLAIN>(dO_.p, 1.0);
      hipLaunchKernelGGL(k_fo_precondition<T>, dim3(unsigned((no + 255) / 256)), dim3(256), 0,
                         stream_, dO_.as<cplx<T>>(), (const T*)pre_.as<T>(),
                         (const T*)floor_.as<T>(), no);
    });
    const int total = int(nprobe_ * n);
    // λ0 = 1.2e-7·k/n per pixel and mode (ptyrs/src/engine/lsqml.rs:31-33), summed over both.
    const double lam = 1.2e-7 * scale_ * double(M_);
    run(K_PROBE, [&] {
      hipLaunchKernelGGL(k_probe_mean<T>, dim3((total + 255) / 256), dim3(256), 0, stream_,
                         dPm_.as<cplx<T>>(), dP_.as<cplx<T>>(), nchunk, total, nb);
    });
    run(K_DOTS, [&] {
      // The direction as the slots read it: its sampled windows (the object's spectrum is no
      // longer needed).
      fo_object_spectrum(dO_.as<cplx<T>>());
      fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 0, fo_dwin_.p, (long long)n);
      hipLaunchKernelGGL(k_lsqml_dots<T>, dim3(xb, nb), dim3(THREADS), 0, stream_,
                         (const cplx<T>*)fo_dwin_.as<cplx<T>>(),
                         (const cplx<T>*)fo_win_.as<cplx<T>>(), N, W, dPm_.as<cplx<T>>(),
                         q_.as<cplx<T>>(), chi, N, M_, fo_tall_.as<int2>(), ord,
                         dots_part_.as<double>(), wview);
      hipLaunchKernelGGL(k_lsqml_step<T>, dim3(1), dim3(1024), 0, stream_,
                         dots_part_.as<double>(), xb, nb, lam, M_, sums_.as<double>(),
                         log_.as<double>() + 3 * size_t(off), step_.as<double>(), 3);
    });
    // OPR's eigenmodes and weights from χ and the batch-start windows and probe.
    if (opr_on_) opr_step(ord, nb);
    run(K_OBJ_UPDATE, [&] {
      hipLaunchKernelGGL(k_lsqml_apply_obj<T>, dim3(unsigned((no + 255) / 256)), dim3(256), 0,
                         stream_, O_.as<cplx<T>>(), dO_.as<cplx<T>>(), step_.as<double>(),
                         size_t(no));
    });
    run(K_PROBE, [&] {
      hipLaunchKernelGGL(k_lsqml_apply_probe<T>, dim3((total + 255) / 256), dim3(256), 0,
                         stream_, P_.as<cplx<T>>(), Wt_.as<cplx<T>>(), q_.as<cplx<T>>(),
                         dPm_.as<cplx<T>>(), step_.as<double>(), n, total);
    });
    if (pos_refine_) {
      const double eps = 1e-6 * scale_;  // EPS_POSITION·k (ptyrs/src/engine/features.rs:455)
      run(K_OTHER, [&] {
        hipLaunchKernelGGL(k_pos_step, dim3(1), dim3(1024), 0, stream_,
                           (const double*)pos_part_.as<double>(), xb, nb, int(nframe_), eps, ord,
                           offs_.as<double2>(), pos_sums_.as<double>());
      });
    }
  }

  // One rPIE batch under offsets and FullObject sampling (ptyrs/src/engine/ptycho.rs:900-1086
  // with Shifted::sampler, the reference GPU's rpie under its position model,
  // ptyrs/src/engine/gpu/rpie.rs:69-133): the waves are the unmoved probe terms; every slot
  // reads its window sampled at its frame's corner plus offset (the tall windows fo_win_);
  // the overlaps D_m are each mode's |W_m|² placed at every slot's point and read back through
  // the windows (fo_overlaps); pass C sums the modes' terms per slot as without offsets, and
  // the summed terms are placed onto the object by the sampling's adjoint; the probe
  // direction is not moved; with refined offsets the position sums take the sampled window's
  // derivatives over the position against mode 0's W and Δ.
  void rpie_batch_sampled(const ptyrs_rpie_params& p, const int* ord, int nb) {
    const bool upd_obj = p.has_obj_step != 0, upd_probe = p.has_probe_step != 0;
    run(K_BIN, [&] {
      fo_points(ord, nb);
      fo_object_spectrum(O_.as<cplx<T>>());
      fo_object_windows(nb);
    });
    if (upd_obj) run(K_OVERLAP, [&] { fo_overlaps(nb); });
    pass_a(ord, nb, upd_probe, MixArgs<T>{MIX_NONE, {}, {}, T(0), T(0)}, nullptr,
           fo_win_.as<cplx<T>>(), N, fo_tall_.as<int2>());
    pass_b<B_RPIE>(ord, nb);
    const int chunk = chunk_of_m(nb);
    const int nchunk = (nb + chunk - 1) / chunk;
    cplx<T>* chi0 = pos_refine_ ? gspec_.as<cplx<T>>() : nullptr;
    run(K_PASS_C, [&] {
      auto launch = [&](auto mm) {
        constexpr int MM = decltype(mm)::value;
        constexpr int G = rows_per_group(N, SR::L, MM);
        hipLaunchKernelGGL((k_pass_c_rpie_sum<T, N, MM>), dim3(N / G, nchunk),
                           dim3(pass_c_sum_threads<T, N, MM>()), 0, stream_, psi_.as<cplx<T>>(),
                           fo_win_.as<cplx<T>>(), N, Wt_.as<cplx<T>>(), q_.as<cplx<T>>(),
                           fo_tall_.as<int2>(), ord, nb, omax_.as<T>(), int(S::NBLK), chunk,
                           T(p.probe_alpha), int(upd_probe), dP_.as<cplx<T>>(),
                           tw_.as<cplx<T>>(), int(upd_obj), (const T*)fo_D_.as<T>(),
                           int(fo_cap_ * N), (const T*)pmax_.as<T>(), T(p.obj_alpha), 0, chi0);
      };
      with_modes(launch);
    });
    if (pos_refine_) {
      cplx<T>* grad = pgrad_.as<cplx<T>>() + cap_ * n;
      run(K_OTHER, [&] {
        fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 1, grad, 2ll * n);
        fo_windows<AS_PLAIN>(fo_spec_.as<cplx<T>>(), nb, 2, grad + n, 2ll * n);
        hipLaunchKernelGGL(k_pos_sums<T>, dim3(aux_blocks(), nb), dim3(THREADS), 0, stream_,
                           (const cplx<T>*)grad, (const cplx<T>*)Wt_.as<cplx<T>>(),
                           (const cplx<T>*)nullptr, 0, (const cplx<T>*)chi0, N, M_,
                           pos_part_.as<double>(), 0);
      });
    }
    if (upd_obj) {
      // The summed terms sit in each slot's mode-0 plane of the stack.
      run(K_OBJ_UPDATE, [&] {
        fo_place_spectrum(psi_.as<cplx<T>>(), (long long)M_ * n, nb, false);
        fo_add_placed(O_.as<cplx<T>>(), p.obj_step);
      });
    }
    if (upd_probe) {
      const int total = int(nprobe_ * n);
      run(K_PROBE, [&] {
        hipLaunchKernelGGL(k_probe_apply<T>, dim3((total + 255) / 256), dim3(256), 0, stream_,
                           P_.as<cplx<T>>(), Wt_.as<cplx<T>>(), q_.as<cplx<T>>(),
                           dP_.as<cplx<T>>(), nchunk, n, total, T(p.probe_step / nb));
      });
    }
    if (pos_refine_) {
      const double eps = 1e-6 * scale_;  // EPS_POSITION·k (ptyrs/src/engine/features.rs:455)
      run(K_OTHER, [&] {
        hipLaunchKernelGGL(k_pos_step, dim3(1), dim3(1024), 0, stream_,
                           (const double*)pos_part_.as<double>(), aux_blocks(), nb, int(nframe_),
                           eps, ord, offs_.as<double2>(), pos_sums_.as<double>());
      });
    }
  }
};

EngineBase* make_engine(int dtype, const Prepared& pre, const double* obj, size_t nobj,
                        size_t nyo, size_t nxo, const double* probe, size_t nprobe, int device);

}  // namespace ptyrs
