// The C ABI of include/ptyrs_hip.h: scan preparation on the host, engine construction for the
// frame side, status codes, and the abort gate.
#include <cstring>
#include <string>

#include "engine.hpp"

namespace ptyrs {

// One instantiation per supported frame side (engine_inst.hip, built once per side).
#define PTYRS_SIDES(X) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(256) X(512)
#define PTYRS_DECL(n)                                                                       \
  EngineBase* make_engine_##n(int dtype, const Prepared& pre, const double* obj, size_t nobj, \
                              size_t nyo, size_t nxo, const double* probe, size_t nprobe,     \
                              int device);
PTYRS_SIDES(PTYRS_DECL)

EngineBase* make_engine(int dtype, const Prepared& pre, const double* obj, size_t nobj,
                        size_t nyo, size_t nxo, const double* probe, size_t nprobe, int device) {
#define PTYRS_CASE(n) \
  case n: return make_engine_##n(dtype, pre, obj, nobj, nyo, nxo, probe, nprobe, device);
  switch (pre.ny) {
    PTYRS_SIDES(PTYRS_CASE)
    default:
      throw Invalid("frame side " + std::to_string(pre.ny) +
                    " not supported on the GPU (16, 24, 32, 48, 64, 96, 128, 256, 512)");
  }
#undef PTYRS_CASE
}

thread_local std::string last_error;

// The exit model of a scan under `optics` (NULL: the far field of the scan geometry).
// Follows ptyrs/src/engine/propagate.rs:24-98 and ptyrs/src/base/geometry.rs:90-126.
ExitModel prepare_exit(const ptyrs_scan& s, const ptyrs_optics* optics) {
  if (s.ny != s.nx) throw Invalid("frames must be square");
  if (s.ny == 0 || s.ny % 2) throw Invalid("frame sides must be even and non-zero");
  const int N = int(s.ny);
  if (optics == nullptr || !optics->free_space) {
    if (!(s.wavelength > 0 && s.distance > 0 && s.detector_pixel > 0) ||
        !std::isfinite(s.wavelength * s.distance / s.detector_pixel))
      throw Invalid("wavelength, distance and detector pixel must be finite and > 0");
    return far_field(N, N, s.wavelength, s.distance, s.detector_pixel);
  }
  return free_space(N, N, optics->wavelength, optics->pixel_y, optics->pixel_x,
                    optics->distance);
}

// View corners [nframe][2] (oy, ox) in an object of nyo × nxo: the scan's corners as given,
// or its positions through PyNX preprocessing and view_corner. Follows
// ptyrs/src/base/positions.rs:51-100 and ptyrs/src/base/addr.rs:35-81.
std::vector<int> prepare_corners(const ptyrs_scan& s, const ExitModel& exit, size_t nyo,
                                 size_t nxo) {
  const int N = int(s.ny);
  if (nyo < s.ny || nxo < s.nx) throw Invalid("object smaller than frame");
  if (s.nframe && s.corners == nullptr && (s.y == nullptr || s.x == nullptr))
    throw Invalid("positions missing");
  std::vector<int> corner(2 * s.nframe);
  if (s.corners) {
    for (size_t f = 0; f < s.nframe; ++f) {
      const int64_t oy = s.corners[2 * f], ox = s.corners[2 * f + 1];
      if (oy < 0 || ox < 0 || oy + N > int64_t(nyo) || ox + N > int64_t(nxo))
        throw Invalid("corner of frame " + std::to_string(f) + " leaves the object");
      corner[2 * f] = int(oy);
      corner[2 * f + 1] = int(ox);
    }
    return corner;
  }
  std::vector<double> y(s.y, s.y + s.nframe), x(s.x, s.x + s.nframe);
  for (size_t f = 0; f < s.nframe; ++f)
    if (!(std::isfinite(y[f]) && std::isfinite(x[f]) && std::fabs(y[f]) <= 1e10 &&
          std::fabs(x[f]) <= 1e10))
      throw Invalid("positions must be finite with |pos| <= 1e10");
  y = pynx_axis(y, exit.pixel[0]);
  x = pynx_axis(x, exit.pixel[1]);
  for (size_t f = 0; f < s.nframe; ++f) {
    const long oy = view_corner(long(nyo), N, y[f] / exit.pixel[0]);
    const long ox = view_corner(long(nxo), N, x[f] / exit.pixel[1]);
    if (oy < 0 || ox < 0)
      throw Invalid("frame " + std::to_string(f) +
                    " leaves the object; size it with required_obj_shape");
    corner[2 * f] = int(oy);
    corner[2 * f + 1] = int(ox);
  }
  return corner;
}

// Counts to the device's DFT index order (the shifted observed intensities of
// ptyrs/src/engine/obs.rs:13-35 for the far field), with nb_obs, the unmasked total and the
// largest frame total. u16: 65535 is masked (ptyrs/src/base/counts.rs:3-7); float: a negative
// value is masked, every other must be finite.
void prepare_counts(const ptyrs_scan& s, const ExitModel& exit, Prepared& p) {
  if ((s.counts_u16 == nullptr) == (s.counts_f32 == nullptr))
    throw Invalid("give exactly one of counts_u16 and counts_f32");
  const int N = int(s.ny);
  const size_t n = size_t(N) * N;
  p.counts.resize(s.nframe * n);
  p.nb_obs = 0;
  p.total = 0;
  p.max_frame_total = 0;
  for (size_t f = 0; f < s.nframe; ++f) {
    double frame_total = 0;
    for (int ky = 0; ky < N; ++ky)
      for (int kx = 0; kx < N; ++kx) {
        // DFT index (ky, kx) sees detector pixel k; the inverse transform flips it (the angular
        // spectrum's two forward transforms land the field at −k).
        int dy = ky, dx = kx;
        if (exit.reverse) {
          dy = (N - ky) % N;
          dx = (N - kx) % N;
        }
        // A DFT plane's index to detector order, (d + N/2) mod N; a spatial plane's is its own.
        const int h = exit.dft_plane ? N / 2 : 0;
        const size_t src = f * n + size_t((dy + h) % N) * N + (dx + h) % N;
        float v;
        if (s.counts_u16) {
          const uint16_t c = s.counts_u16[src];
          v = c == 65535 ? -1.0f : float(c);
        } else {
          const float c = s.counts_f32[src];
          if (!std::isfinite(c)) throw Invalid("counts must be finite");
          v = c < 0 ? -1.0f : c;
        }
        if (v >= 0) {
          ++p.nb_obs;
          p.total += v;
          frame_total += v;
        }
        p.counts[f * n + size_t(ky) * N + kx] = v;
      }
    p.max_frame_total = std::max(p.max_frame_total, frame_total);
  }
}

// A scan, its optics and the object's shape as the device takes them. Follows
// ptyrs/src/engine/gpu/mod.rs:635-760 (shape checks, pynx_preprocess, AddrTable).
Prepared prepare(const ptyrs_scan& s, const ptyrs_optics* optics, size_t nprobe, size_t nobj,
                 size_t nyo, size_t nxo) {
  if (nprobe == 0 || nobj == 0) throw Invalid("nprobe and nobj must be >= 1");
  Prepared p;
  p.exit = prepare_exit(s, optics);
  p.nframe = s.nframe;
  p.ny = s.ny;
  p.nx = s.nx;
  const int N = int(s.ny);
  if (optics && optics->ngap) {
    // Optics::with_slices: every gap an angular-spectrum propagator over the frame
    // (ptyrs/src/engine/propagate.rs:84-98).
    if (!optics->gaps) throw Invalid("ngap > 0 needs gaps");
    p.nslice = optics->ngap + 1;
    for (size_t g = 0; g < optics->ngap; ++g) {
      const double* w = optics->gaps + 4 * g;
      if (!(std::isfinite(w[0]) && w[0] > 0 && std::isfinite(w[1]) && w[1] > 0 &&
            std::isfinite(w[2]) && w[2] > 0))
        throw Invalid("gap wavelength and pixel must be finite and > 0");
      if (!std::isfinite(w[3]) || w[3] == 0.0)
        throw Invalid("gap distance must be finite and non-zero");
      p.gaps.push_back(angular_spectrum_tf(N, N, w[0], w[1], w[2], w[3], nullptr));
    }
    if (nobj != 1) throw Invalid("a multislice object has one mode");
  }
  p.corner = prepare_corners(s, p.exit, nyo, nxo);
  prepare_counts(s, p.exit, p);
  return p;
}

}  // namespace ptyrs

using namespace ptyrs;

struct ptyrs_engine {
  std::unique_ptr<EngineBase> e;
};

namespace {

template <typename F>
int guarded(ptyrs_engine* h, bool mutating, F&& f) {
  try {
    if (h == nullptr || !h->e) throw Invalid("null engine");
    if (mutating && h->e->aborted)
      throw Aborted("aborted: an earlier call failed part-way; rebuild the engine");
    f(*h->e);
    return PTYRS_OK;
  } catch (const Invalid& x) {
    last_error = x.what();
    return PTYRS_EINVAL;
  } catch (const Aborted& x) {
    last_error = x.what();
    return PTYRS_EABORTED;
  } catch (const std::exception& x) {
    last_error = x.what();
    if (h && h->e && mutating) h->e->aborted = true;
    return PTYRS_EDEVICE;
  }
}

}  // namespace

namespace {

template <typename F>
int host_call(F&& f) {
  try {
    f();
    return PTYRS_OK;
  } catch (const Invalid& x) {
    last_error = x.what();
    return PTYRS_EINVAL;
  } catch (const std::exception& x) {
    last_error = x.what();
    return PTYRS_EDEVICE;
  }
}

}  // namespace

// Scratch of ptyrs_engine_llk_noalloc: one array per GPU, held from the library's load on that
// GPU until the process ends. Its device address is taken once per device with that device
// current (hipGetSymbolAddress) and only that address is passed to kernels: on the host the
// array's name is a stand-in, not a device address.
constexpr size_t kLlkScratchBytes = size_t(64) << 20;
__device__ __align__(256) static unsigned char scratch[kLlkScratchBytes];

namespace {

void* scratch_on(int device) {
  static std::mutex m;
  static std::vector<void*> ptrs;
  std::lock_guard<std::mutex> g(m);
  if (device < 0) throw Invalid("bad device");
  if (size_t(device) < ptrs.size() && ptrs[device]) return ptrs[device];
  int prev = 0;
  PTYRS_HIP(hipGetDevice(&prev));
  PTYRS_HIP(hipSetDevice(device));
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  (void)hipSetDevice(prev);
  if (e != hipSuccess || p == nullptr)
    throw DeviceError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  if (ptrs.size() <= size_t(device)) ptrs.resize(device + 1, nullptr);
  ptrs[device] = p;
  return p;
}

}  // namespace

extern "C" {

int ptyrs_engine_create(const ptyrs_scan* scan, const ptyrs_optics* optics, const double* obj,
                        size_t nobj, size_t nyo, size_t nxo, const double* probe, size_t nprobe,
                        const ptyrs_gpu_options* options, int device, ptyrs_engine** out) {
  try {
    if (!scan || !obj || !probe || !out) throw Invalid("null argument");
    const int dtype = options ? options->dtype : PTYRS_C64;
    if (dtype != PTYRS_C64 && dtype != PTYRS_C128) throw Invalid("dtype must be c64 or c128");
    if (options) {
      if (options->llk_slots == 0) throw Invalid("llk_slots must be >= 1");
      if (options->psi_f16 && !options->fused) throw Invalid("psi_f16 needs fused passes");
      if (options->psi_f16 && dtype != PTYRS_C64)
        throw Invalid("psi_f16 holds the f16 form of complex64 waves: dtype must be c64");
    }
    Prepared pre = prepare(*scan, optics, nprobe, nobj, nyo, nxo);
    pre.graph = options ? options->graph != 0 : true;
    pre.stack_size = options ? options->stack_size : 0;
    pre.llk_slots = options ? options->llk_slots : 8;
    pre.psi_f16 = options ? options->psi_f16 != 0 : false;
    if (pre.stack_size && std::min(pre.stack_size, pre.nframe) * nprobe * nobj > 65535)
      throw Invalid("a stack of " + std::to_string(pre.stack_size) +
                    " frames holds more than 65535 views");
    auto h = std::make_unique<ptyrs_engine>();
    h->e.reset(make_engine(dtype, pre, obj, nobj, nyo, nxo, probe, nprobe, device));
    *out = h.release();
    return PTYRS_OK;
  } catch (const Invalid& x) {
    last_error = x.what();
    return PTYRS_EINVAL;
  } catch (const std::exception& x) {
    last_error = x.what();
    return PTYRS_EDEVICE;
  }
}

int ptyrs_engine_rpie(ptyrs_engine* h, const ptyrs_rpie_params* p, size_t nepoch) {
  if (!p) {
    last_error = "null params";
    return PTYRS_EINVAL;
  }
  return guarded(h, true, [&](EngineBase& e) { e.rpie(*p, nepoch); });
}

namespace {

// Every engine's next epoch captured in one graph, a branch per engine on its own stream
// (forked from and joined back to a group stream), replayed while the engines' keys hold.
void group_epochs_one_graph(std::vector<EngineBase*>& es) {
  hipStream_t gs = nullptr;
  PTYRS_HIP(hipStreamCreateWithFlags(&gs, hipStreamNonBlocking));
  std::vector<hipEvent_t> ev(2 * es.size() + 2, nullptr);
  for (auto& e : ev) PTYRS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  hipEvent_t fork = ev[0], done = ev[1];
  hipGraphExec_t exec = nullptr;
  std::vector<uint64_t> held;
  auto cleanup = [&] {
    if (exec) (void)hipGraphExecDestroy(exec);
    (void)hipStreamSynchronize(gs);
    for (auto& e : ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(gs);
  };
  try {
    for (;;) {
      std::vector<size_t> act;
      for (size_t k = 0; k < es.size(); ++k)
        if (es[k]->rpie_pending()) act.push_back(k);
      if (act.empty()) break;
      std::vector<uint64_t> key;
      for (size_t k : act) {
        const auto kk = es[k]->rpie_epoch_order();
        key.push_back(uint64_t(k));
        key.push_back(kk.size());
        key.insert(key.end(), kk.begin(), kk.end());
        PTYRS_HIP(hipEventRecord(ev[2 + k], es[k]->stream()));
        PTYRS_HIP(hipStreamWaitEvent(gs, ev[2 + k], 0));
      }
      if (exec == nullptr || key != held) {
        if (exec) PTYRS_HIP(hipGraphExecDestroy(exec));
        exec = nullptr;
        hipGraph_t graph = nullptr;
        PTYRS_HIP(hipStreamBeginCapture(gs, hipStreamCaptureModeThreadLocal));
        try {
          PTYRS_HIP(hipEventRecord(fork, gs));
          for (size_t k : act) {
            PTYRS_HIP(hipStreamWaitEvent(es[k]->stream(), fork, 0));
            es[k]->rpie_epoch_body();
            PTYRS_HIP(hipEventRecord(ev[2 + es.size() + k], es[k]->stream()));
            PTYRS_HIP(hipStreamWaitEvent(gs, ev[2 + es.size() + k], 0));
          }
        } catch (...) {
          hipGraph_t dropped = nullptr;
          (void)hipStreamEndCapture(gs, &dropped);
          if (dropped) (void)hipGraphDestroy(dropped);
          throw;
        }
        PTYRS_HIP(hipStreamEndCapture(gs, &graph));
        const hipError_t inst = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (inst != hipSuccess) {
          exec = nullptr;
          throw DeviceError(std::string("hipGraphInstantiate: ") + hipGetErrorString(inst));
        }
        held = key;
      }
      PTYRS_HIP(hipGraphLaunch(exec, gs));
      PTYRS_HIP(hipEventRecord(done, gs));
      for (size_t k : act) {
        PTYRS_HIP(hipStreamWaitEvent(es[k]->stream(), done, 0));
        es[k]->rpie_epoch_after();
      }
    }
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
}

}  // namespace

namespace {

int group_rpie(ptyrs_engine* const* engines, size_t count, const ptyrs_rpie_params* params,
               size_t nepoch, bool one_graph) {
  try {
    if (!engines || !params) throw Invalid("null argument");
    std::vector<EngineBase*> es;
    for (size_t i = 0; i < count; ++i) {
      ptyrs_engine* h = engines[i];
      if (h == nullptr || !h->e) throw Invalid("null engine");
      if (h->e->aborted)
        throw Aborted("aborted: an earlier call failed part-way; rebuild the engine");
      for (EngineBase* o : es)
        if (o == h->e.get()) throw Invalid("an engine appears twice in the group");
      es.push_back(h->e.get());
    }
    // Every engine's parameters are checked before any engine changes.
    for (size_t i = 0; i < count; ++i) es[i]->rpie_check(params[i]);
    size_t i = 0;
    try {
      for (; i < count; ++i) es[i]->rpie_begin(params[i], nepoch);
      // One graph only when every engine captures its epochs: an engine built with graph off,
      // or profiling its launches, runs eagerly, as ptyrs_engine_rpie would run it.
      bool all_capture = true;
      for (EngineBase* e : es) all_capture = all_capture && e->captures();
      if (one_graph && all_capture) {
        group_epochs_one_graph(es);
      } else {
        // Epoch by epoch, round-robin: each engine's epoch is one replay on its own stream, so
        // the engines' epochs run side by side on the device.
        bool more = true;
        while (more) {
          more = false;
          for (EngineBase* e : es) more = e->rpie_epoch() || more;
        }
      }
      for (EngineBase* e : es) e->rpie_end();
    } catch (const Invalid&) {
      throw;
    } catch (const std::exception&) {
      for (size_t k = 0; k < count; ++k) es[k]->aborted = true;
      throw;
    }
    return PTYRS_OK;
  } catch (const Invalid& x) {
    last_error = x.what();
    return PTYRS_EINVAL;
  } catch (const Aborted& x) {
    last_error = x.what();
    return PTYRS_EABORTED;
  } catch (const std::exception& x) {
    last_error = x.what();
    return PTYRS_EDEVICE;
  }
}

}  // namespace

int ptyrs_group_rpie(ptyrs_engine* const* engines, size_t count,
                     const ptyrs_rpie_params* params, size_t nepoch) {
  return group_rpie(engines, count, params, nepoch, false);
}

int ptyrs_group_rpie_graph(ptyrs_engine* const* engines, size_t count,
                           const ptyrs_rpie_params* params, size_t nepoch) {
  return group_rpie(engines, count, params, nepoch, true);
}

namespace {
ProjParams proj(int algorithm, int ho, double oi, int hp, double pi, double mix, size_t loops,
                size_t calc_llk, int zpr) {
  return ProjParams{algorithm, ho != 0, hp != 0, oi, pi, mix, loops, calc_llk, zpr != 0};
}
}  // namespace

int ptyrs_engine_ap(ptyrs_engine* h, const ptyrs_ap_params* p, size_t ncycle) {
  if (!p) {
    last_error = "null params";
    return PTYRS_EINVAL;
  }
  return guarded(h, true, [&](EngineBase& e) {
    e.project(proj(PTYRS_AP, p->has_obj_inertia, p->obj_inertia, p->has_probe_inertia,
                   p->probe_inertia, 0.0, 1, p->calc_llk, p->zero_phase_ramp),
              ncycle);
  });
}

int ptyrs_engine_dm(ptyrs_engine* h, const ptyrs_dm_params* p, size_t ncycle) {
  if (!p) {
    last_error = "null params";
    return PTYRS_EINVAL;
  }
  return guarded(h, true, [&](EngineBase& e) {
    e.project(proj(PTYRS_DM, p->has_obj_inertia, p->obj_inertia, p->has_probe_inertia,
                   p->probe_inertia, p->alpha, p->loop_obj_probe, p->calc_llk,
                   p->zero_phase_ramp),
              ncycle);
  });
}

int ptyrs_engine_raar(ptyrs_engine* h, const ptyrs_raar_params* p, size_t ncycle) {
  if (!p) {
    last_error = "null params";
    return PTYRS_EINVAL;
  }
  return guarded(h, true, [&](EngineBase& e) {
    e.project(proj(PTYRS_RAAR, p->has_obj_inertia, p->obj_inertia, p->has_probe_inertia,
                   p->probe_inertia, p->beta, p->loop_obj_probe, p->calc_llk,
                   p->zero_phase_ramp),
              ncycle);
  });
}

int ptyrs_engine_ml(ptyrs_engine* h, const ptyrs_ml_params* p, size_t ncycle) {
  if (!p) {
    last_error = "null params";
    return PTYRS_EINVAL;
  }
  return guarded(h, true, [&](EngineBase& e) { e.ml(*p, ncycle); });
}

int ptyrs_engine_lsqml(ptyrs_engine* h, const ptyrs_lsqml_params* p, size_t nepoch,
                       size_t calc_llk) {
  if (!p) {
    last_error = "null params";
    return PTYRS_EINVAL;
  }
  return guarded(h, true, [&](EngineBase& e) { e.lsqml(*p, nepoch, calc_llk); });
}

int ptyrs_engine_lsqml_opr(ptyrs_engine* h, const ptyrs_lsqml_params* p, size_t nepoch,
                           size_t calc_llk, ptyrs_opr* var) {
  if (!p || !var) {
    last_error = "null params or variation";
    return PTYRS_EINVAL;
  }
  return guarded(h, true, [&](EngineBase& e) { e.lsqml_opr(*p, nepoch, calc_llk, *var); });
}

int ptyrs_engine_bh(ptyrs_engine* h, const ptyrs_bh_params* p, size_t nepoch, size_t calc_llk) {
  if (!p) {
    last_error = "null params";
    return PTYRS_EINVAL;
  }
  return guarded(h, true, [&](EngineBase& e) { e.bh(*p, nepoch, calc_llk); });
}

int ptyrs_engine_bh_steps(ptyrs_engine* h, ptyrs_bh_step* out, size_t cap, size_t* n) {
  return guarded(h, false, [&](EngineBase& e) {
    if (n) *n = e.bh_steps.size();
    for (size_t i = 0; i < e.bh_steps.size() && i < cap && out; ++i) {
      const BhRecord& r = e.bh_steps[i];
      out[i] = ptyrs_bh_step{r.first, r.count, r.beta, r.alpha, r.top, r.bottom};
    }
  });
}

int ptyrs_engine_bh_frames(ptyrs_engine* h, size_t* out, size_t len) {
  return guarded(h, false, [&](EngineBase& e) {
    if (len != e.bh_order.size() || (len && !out)) throw Invalid("buffer has the wrong length");
    for (size_t i = 0; i < len; ++i) out[i] = size_t(e.bh_order[i]);
  });
}

int ptyrs_engine_lsqml_steps(ptyrs_engine* h, ptyrs_lsqml_step* out, size_t cap, size_t* n) {
  return guarded(h, false, [&](EngineBase& e) {
    if (n) *n = e.lsqml_steps.size();
    for (size_t i = 0; i < e.lsqml_steps.size() && i < cap && out; ++i) {
      const StepRecord& r = e.lsqml_steps[i];
      out[i] = ptyrs_lsqml_step{size_t(r.batch), size_t(r.frame), r.a_far, r.a_obj, r.a_probe};
    }
  });
}

int ptyrs_engine_lsqml_slice_steps(ptyrs_engine* h, double* out, size_t len) {
  return guarded(h, false, [&](EngineBase& e) {
    const size_t S = e.nslice();
    if (!out || len != e.lsqml_steps.size() * S) throw Invalid("buffer has the wrong length");
    for (size_t i = 0; i < e.lsqml_steps.size(); ++i)
      for (size_t s = 0; s < S; ++s) out[i * S + s] = e.lsqml_steps[i].a_objs[s];
  });
}

int ptyrs_engine_set_launch(ptyrs_engine* h, const ptyrs_launch* l) {
  if (!l) {
    last_error = "null launch";
    return PTYRS_EINVAL;
  }
  return guarded(h, true, [&](EngineBase& e) {
    e.set_launch(l->stack_size, l->pass_c_blocks, l->chi_blocks);
  });
}

int ptyrs_engine_launch(ptyrs_engine* h, ptyrs_launch* out) {
  return guarded(h, false, [&](EngineBase& e) {
    if (out) e.get_launch(&out->stack_size, &out->pass_c_blocks, &out->chi_blocks);
  });
}

int ptyrs_engine_set_position(ptyrs_engine* h, const double* offsets, size_t len, int refine) {
  return guarded(h, true, [&](EngineBase& e) {
    if (offsets && len != e.corners.size())  // 2·nframe
      throw Invalid("position offsets must be [nframe, 2]");
    e.set_position(offsets, refine != 0);
  });
}

int ptyrs_engine_position(ptyrs_engine* h, double* out, size_t len, int* state) {
  return guarded(h, false, [&](EngineBase& e) {
    if (out && len != e.corners.size()) throw Invalid("position buffer has the wrong length");
    const int s = e.get_position(out);
    if (state) *state = s;
  });
}

int ptyrs_engine_sampling(ptyrs_engine* h, int* sampling) {
  return guarded(h, false, [&](EngineBase& e) {
    if (sampling) *sampling = e.sampling;
  });
}

int ptyrs_engine_set_sampling(ptyrs_engine* h, int sampling) {
  return guarded(h, true, [&](EngineBase& e) {
    if (sampling != PTYRS_PATCH && sampling != PTYRS_FULL_OBJECT)
      throw Invalid("sampling must be PATCH or FULL_OBJECT");
    e.sampling = sampling;
  });
}

int ptyrs_engine_set_checkpoint(ptyrs_engine* h, size_t every) {
  return guarded(h, true, [&](EngineBase& e) {
    if (every > 64) throw Invalid("checkpoint interval must be at most 64 slices");
    e.set_checkpoint(every);
  });
}

int ptyrs_engine_checkpoint(ptyrs_engine* h, size_t* every, size_t* wave_bytes) {
  return guarded(h, false, [&](EngineBase& e) {
    if (every) *every = e.checkpoint;
    if (wave_bytes) *wave_bytes = e.wave_bytes();
  });
}

int ptyrs_engine_grad(ptyrs_engine* h, int cost, double* value, double* obj_grad,
                     size_t obj_len, double* probe_grad, size_t probe_len) {
  return guarded(h, true, [&](EngineBase& e) {
    if (!value || !obj_grad || !probe_grad) throw Invalid("null output");
    if (obj_len != e.object_len() || probe_len != e.probe_len())
      throw Invalid("gradient buffer has the wrong length");
    e.grad(cost, value, obj_grad, probe_grad);
  });
}

int ptyrs_engine_intensities(ptyrs_engine* h, double* out, size_t len) {
  return guarded(h, true, [&](EngineBase& e) {
    if (!out || len != e.frame_len())
      throw Invalid("intensity buffer has the wrong length");
    e.intensities(out);
  });
}

int ptyrs_engine_live_reset(ptyrs_engine* h) {
  return guarded(h, true, [&](EngineBase& e) { e.live_reset(); });
}

int ptyrs_engine_live_push(ptyrs_engine* h, const uint16_t* counts_u16, const float* counts_f32,
                           size_t nframes) {
  return guarded(h, true, [&](EngineBase& e) {
    if (nframes > e.planned() - e.landed())
      throw Invalid("live_push lands more frames than are planned");
    // The counts' conversion of the scan (prepare_counts) on the landing frames alone.
    ptyrs_scan s{};
    s.nframe = nframes;
    s.ny = s.nx = e.frame_side();
    s.counts_u16 = counts_u16;
    s.counts_f32 = counts_f32;
    ExitModel exit;
    exit.reverse = e.counts_reverse();
    exit.dft_plane = e.counts_dft_plane();
    Prepared p;
    prepare_counts(s, exit, p);
    e.live_push(p.counts.data(), nframes, p.nb_obs, p.total, p.max_frame_total);
  });
}

int ptyrs_engine_live_frames(ptyrs_engine* h, size_t* landed, size_t* planned) {
  return guarded(h, false, [&](EngineBase& e) {
    if (landed) *landed = e.landed();
    if (planned) *planned = e.planned();
  });
}

int ptyrs_engine_llk(ptyrs_engine* h, ptyrs_llk_record* out) {
  return guarded(h, true, [&](EngineBase& e) {
    if (!out) throw Invalid("null output");
    double v[4];
    e.llk(v);
    out->cycle = e.cycle;
    out->algorithm = -1;
    out->poisson = v[0];
    out->gaussian = v[1];
    out->euclidian = v[2];
    out->nb_photons_calc = v[3];
  });
}

int ptyrs_engine_llk_noalloc(ptyrs_engine* h, ptyrs_llk_record* out) {
  return guarded(h, true, [&](EngineBase& e) {
    if (!out) throw Invalid("null output");
    void* s = scratch_on(e.device());
    double v[4];
    e.llk_noalloc(s, kLlkScratchBytes, v);
    out->cycle = e.cycle;
    out->algorithm = -1;
    out->poisson = v[0];
    out->gaussian = v[1];
    out->euclidian = v[2];
    out->nb_photons_calc = v[3];
  });
}

int ptyrs_engine_object(ptyrs_engine* h, double* out, size_t len) {
  return guarded(h, false, [&](EngineBase& e) {
    if (!out || len != e.object_len()) throw Invalid("object buffer has the wrong length");
    e.get_object(out);
  });
}

int ptyrs_engine_probe(ptyrs_engine* h, double* out, size_t len) {
  return guarded(h, false, [&](EngineBase& e) {
    if (!out || len != e.probe_len()) throw Invalid("probe buffer has the wrong length");
    e.get_probe(out);
  });
}

int ptyrs_engine_set_object(ptyrs_engine* h, const double* in, size_t len) {
  return guarded(h, true, [&](EngineBase& e) {
    if (!in || len != e.object_len()) throw Invalid("object buffer has the wrong length");
    e.set_object(in);
  });
}

int ptyrs_engine_set_probe(ptyrs_engine* h, const double* in, size_t len) {
  return guarded(h, true, [&](EngineBase& e) {
    if (!in || len != e.probe_len()) throw Invalid("probe buffer has the wrong length");
    e.set_probe(in);
  });
}

int ptyrs_engine_history(ptyrs_engine* h, ptyrs_llk_record* out, size_t cap, size_t* n) {
  return guarded(h, false, [&](EngineBase& e) {
    if (n) *n = e.history.size();
    for (size_t i = 0; i < e.history.size() && i < cap && out; ++i) {
      const Record& r = e.history[i];
      out[i] = ptyrs_llk_record{r.cycle, r.algorithm, r.llk[0], r.llk[1], r.llk[2], r.llk[3]};
    }
  });
}

int ptyrs_engine_cycle(ptyrs_engine* h, size_t* cycle) {
  return guarded(h, false, [&](EngineBase& e) {
    if (cycle) *cycle = e.cycle;
  });
}

int ptyrs_engine_corners(ptyrs_engine* h, int64_t* out, size_t len) {
  return guarded(h, false, [&](EngineBase& e) {
    if (!out || len != e.corners.size()) throw Invalid("corner buffer has the wrong length");
    for (size_t i = 0; i < len; ++i) out[i] = e.corners[i];
  });
}

int ptyrs_engine_synchronize(ptyrs_engine* h) {
  return guarded(h, false, [&](EngineBase& e) { e.synchronize(); });
}

int ptyrs_engine_profile(ptyrs_engine* h, int enable) {
  return guarded(h, false, [&](EngineBase& e) { e.profile(enable != 0); });
}

int ptyrs_engine_kernel_stats(ptyrs_engine* h, const char* name, size_t* launches,
                              double* total_ms) {
  return guarded(h, false, [&](EngineBase& e) {
    if (!name || !launches || !total_ms) throw Invalid("null argument");
    std::string names = kind_names();
    int kind = 0;
    size_t start = 0;
    while (true) {
      size_t end = names.find(',', start);
      std::string k = names.substr(start, end == std::string::npos ? std::string::npos : end - start);
      if (k == name) break;
      if (end == std::string::npos) throw Invalid(std::string("unknown kernel name ") + name);
      start = end + 1;
      ++kind;
    }
    e.kernel_stats(kind, launches, total_ms);
  });
}

const char* ptyrs_kernel_names(void) { return kind_names(); }

int ptyrs_test_arm_fault(int exit, size_t pass) {
  return host_call([&] {
    if (exit != -1 && exit != PTYRS_FAULT_LAUNCH && exit != PTYRS_FAULT_CAPTURE)
      throw Invalid("unknown fault exit");
    FaultArm& a = fault_arm();
    std::lock_guard<std::mutex> g(a.m);
    a.exit = exit;
    a.pass = pass;
    a.fired = false;
    a.armed.store(exit != -1);
  });
}

int ptyrs_test_fault_fired(int* fired) {
  return host_call([&] {
    if (!fired) throw Invalid("null argument");
    FaultArm& a = fault_arm();
    std::lock_guard<std::mutex> g(a.m);
    *fired = a.fired ? 1 : 0;
  });
}

int ptyrs_pynx_preprocess(const double* v, size_t n, double pixel, double* out) {
  return host_call([&] {
    if (n && (!v || !out)) throw Invalid("null argument");
    for (size_t i = 0; i < n; ++i)
      if (!(std::isfinite(v[i]) && std::fabs(v[i]) <= 1e10))
        throw Invalid("positions must be finite with |pos| <= 1e10");
    const std::vector<double> r = pynx_axis(std::vector<double>(v, v + n), pixel);
    for (size_t i = 0; i < n; ++i)
      if (!(std::isfinite(r[i]) && std::fabs(r[i]) <= 1e10))
        throw Invalid("preprocessed positions must be finite with |pos| <= 1e10");
    std::copy(r.begin(), r.end(), out);
  });
}

int ptyrs_view_corner(size_t obj, size_t frame, double shift, int64_t* corner) {
  return host_call([&] {
    if (!corner) throw Invalid("null argument");
    *corner = view_corner(long(obj), long(frame), shift);
  });
}

int ptyrs_required_obj_shape(const double* y, const double* x, size_t n, double pixel_y,
                             double pixel_x, size_t frame_y, size_t frame_x, size_t margin,
                             size_t* out) {
  return host_call([&] {
    if (!out || (n && (!y || !x))) throw Invalid("null argument");
    out[0] = required_obj_axis(y, n, pixel_y, frame_y, margin);
    out[1] = required_obj_axis(x, n, pixel_x, frame_x, margin);
  });
}

int ptyrs_scan_corners(const ptyrs_scan* scan, const ptyrs_optics* optics, size_t nyo,
                       size_t nxo, int64_t* out, size_t len) {
  return host_call([&] {
    if (!scan || !out) throw Invalid("null argument");
    if (len != 2 * scan->nframe) throw Invalid("corner buffer has the wrong length");
    const ExitModel exit = prepare_exit(*scan, optics);
    const std::vector<int> c = prepare_corners(*scan, exit, nyo, nxo);
    for (size_t i = 0; i < len; ++i) out[i] = c[i];
  });
}

int ptyrs_scan_observed(const ptyrs_scan* scan, const ptyrs_optics* optics, float* out,
                        size_t len, size_t* nb_obs, double* total) {
  return host_call([&] {
    if (!scan) throw Invalid("null argument");
    if (out && len != scan->nframe * scan->ny * scan->nx)
      throw Invalid("observed buffer has the wrong length");
    Prepared p;
    const ExitModel exit = prepare_exit(*scan, optics);
    prepare_counts(*scan, exit, p);
    if (out) std::copy(p.counts.begin(), p.counts.end(), out);
    if (nb_obs) *nb_obs = p.nb_obs;
    if (total) *total = p.total;
  });
}

int ptyrs_quad_phase(double wavelength, double distance, double detector_pixel, size_t ny,
                     size_t nx, double* out, size_t len) {
  return host_call([&] {
    if (!out) throw Invalid("null argument");
    if (ny == 0 || nx == 0 || ny % 2 || nx % 2) throw Invalid("frame sides must be even");
    if (len != 2 * ny * nx) throw Invalid("phase buffer has the wrong length");
    if (!(wavelength > 0 && distance > 0 && detector_pixel > 0) ||
        !std::isfinite(wavelength * distance / detector_pixel))
      throw Invalid("wavelength, distance and detector pixel must be finite and > 0");
    const ExitModel m = far_field(int(ny), int(nx), wavelength, distance, detector_pixel);
    std::copy(m.quad.begin(), m.quad.end(), out);
  });
}

void ptyrs_engine_destroy(ptyrs_engine* h) { delete h; }

size_t ptyrs_last_error(char* buf, size_t len) {
  if (buf && len) {
    size_t k = std::min(len - 1, last_error.size());
    std::memcpy(buf, last_error.data(), k);
    buf[k] = 0;
  }
  return last_error.size();
}

}  // extern "C"
