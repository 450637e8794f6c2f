// The engine for one frame side, PTYRS_N, in both element types. Built once per supported
// side so that the sides compile in parallel.
#include "engine.hpp"

#ifndef PTYRS_N
#error "define PTYRS_N, the frame side"
#endif

#define PTYRS_CAT2(a, b) a##b
#define PTYRS_CAT(a, b) PTYRS_CAT2(a, b)

namespace ptyrs {

EngineBase* PTYRS_CAT(make_engine_, PTYRS_N)(int dtype, const Prepared& pre, const double* obj,
                                             size_t nobj, size_t nyo, size_t nxo,
                                             const double* probe, size_t nprobe, int device) {
  if (dtype == PTYRS_C64)
    return new Engine<float, PTYRS_N>(pre, obj, nobj, nyo, nxo, probe, nprobe, device);
  return new Engine<double, PTYRS_N>(pre, obj, nobj, nyo, nxo, probe, nprobe, device);
}

}  // namespace ptyrs
