// Host-side preparation of a scan for the device: position preprocessing, view corners, the
// exit-wave plane phase, the detector index map of the exit transform, and seeded batches.
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ptyrs_hip.h"

namespace ptyrs {

struct Invalid : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct Aborted : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline double rem_euclid(double a, double b) {
  double r = std::fmod(a, b);
  return r < 0 ? r + std::fabs(b) : r;
}

// PyNX position preprocessing along one axis: centre on the mean, then, if every position
// shares one sub-pixel offset (within 0.02 px), remove it and snap to whole pixels.
// Follows ptyrs/src/base/positions.rs:76-100.
inline std::vector<double> pynx_axis(const std::vector<double>& v, double px) {
  std::vector<double> pos(v.size());
  if (v.empty()) return pos;
  const double n = double(v.size());
  double sum = 0;
  for (double a : v) sum += a;
  const double mean = sum / n;
  for (size_t i = 0; i < v.size(); ++i) pos[i] = v[i] - mean;
  std::vector<double> frac(v.size());
  bool near = true;
  for (size_t i = 0; i < v.size(); ++i) {
    frac[i] = rem_euclid(pos[i] / px, 1.0);
    if (!(std::fabs(std::fmin(frac[i], 1.0 - frac[i])) < 0.1)) near = false;
  }
  if (near)
    for (size_t i = 0; i < v.size(); ++i) frac[i] = rem_euclid(pos[i] + 0.5, 1.0) - 0.5;
  double fs = 0;
  for (double d : frac) fs += d;
  const double fmean = fs / n;
  bool shared = true;
  for (double d : frac)
    if (!(std::fabs(d - fmean) <= 0.02 + 1e-5 * std::fabs(fmean))) shared = false;
  if (shared)
    for (double& p : pos) p = std::nearbyint((p - fmean * px) / px) * px;
  return pos;
}

// Corner of the window along one axis, (obj − frame)/2 + shift rounded half to even; -1 when
// the window leaves the object. Follows ptyrs/src/base/addr.rs:118-127.
inline long view_corner(long obj, long frame, double shift) {
  if (obj < frame || !std::isfinite(shift)) return -1;
  const long span = obj - frame;
  const double c = std::nearbyint(double(span / 2) + shift);  // FE_TONEAREST: ties to even
  if (c < 0.0 || c > double(span)) return -1;
  return long(c);
}

// PyNX calc_obj_shape(..., margin, multiple=4) along one axis of preprocessed positions v
// (metres, pixel p): 2·max(|ceil(v/p)| + 1) + frame + margin, rounded up to a multiple of 4.
// Follows ptyrs/src/base/addr.rs:129-149.
inline size_t required_obj_axis(const double* v, size_t count, double p, size_t frame,
                                size_t margin) {
  double ext = 0.0;
  for (size_t i = 0; i < count; ++i) ext = std::fmax(ext, std::fabs(std::ceil(v[i] / p)) + 1.0);
  const size_t len = size_t(2.0 * ext) + frame + margin;
  return (len + 3) / 4 * 4;
}

// SplitMix64 output function and the per-epoch batches of ptyrs/src/engine/batch.rs:7-38:
// Fisher–Yates from the end with draws from seed ^ mix(epoch), chunks sorted ascending.
inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

inline std::vector<std::vector<int>> random_batches(size_t nframe, size_t size, uint64_t seed,
                                                    uint64_t epoch) {
  if (size == 0) throw Invalid("batch size must be positive");
  uint64_t state = seed ^ mix64(epoch);
  std::vector<int> order(nframe);
  for (size_t i = 0; i < nframe; ++i) order[i] = int(i);
  for (size_t i = nframe; i-- > 1;) {
    state += 0x9e3779b97f4a7c15ULL;
    const uint64_t r = mix64(state);
    const size_t j = size_t(((unsigned __int128)r * (unsigned __int128)(i + 1)) >> 64);
    std::swap(order[i], order[j]);
  }
  std::vector<std::vector<int>> out;
  for (size_t s = 0; s < nframe; s += size) {
    std::vector<int> c(order.begin() + s, order.begin() + std::min(nframe, s + size));
    std::sort(c.begin(), c.end());
    out.push_back(std::move(c));
  }
  return out;
}

// The exit transform seen by the device: a plane phase q (natural layout) applied before one
// unitary DFT, and for each DFT index the detector pixel it lands on.
struct ExitModel {
  std::vector<double> quad;  // interleaved re, im; n values
  bool reverse = false;      // the detector sees the inverse DFT (Fresnel with z < 0)
  // The detector plane is a DFT of the exit wave (far field, Fresnel: its frame is stored
  // fftshifted against the DFT index), or a plane in the exit wave's own layout (angular
  // spectrum: the frame's natural index is the field's).
  bool dft_plane = true;
  double pixel[2] = {0, 0};  // object-plane pixel [py, px], metres
  int kind = 0;              // PTYRS_FOURIER / PTYRS_FRESNEL / PTYRS_ANGULAR_SPECTRUM
  // pty-chi's intensity scale k of the propagator (ptyrs/src/engine/propagate.rs:373-385):
  // the engine's probe is pty-chi's times √k.
  double scale = 1.0;
  // Angular spectrum: the transfer function on the unshifted fftfreq grid (interleaved re, im)
  // and whether it has an evanescent band (zeros), where its adjoint does not invert it.
  std::vector<double> tf;
  bool banded = false;
  // The plane phase as exp(i·(c[0]·Y² + c[1]·X²)) over the centred pixel, c = plane_quad
  // (ptyrs/src/engine/propagate.rs:29-31); 0 where there is none.
  double plane_quad[2] = {0, 0};
};

inline void set_phase(std::vector<double>& q, size_t i, double ph) {
  q[2 * i] = std::cos(ph);
  q[2 * i + 1] = std::sin(ph);
}

// Far field of the scan geometry: q = exp(iπ/(λz)·(dx²px² + dy²py²)), pixel λz/(N·dpix).
// Follows ptyrs/src/engine/ops.rs:11-33 and ptyrs/src/base/geometry.rs:118-126.
inline ExitModel far_field(int ny, int nx, double wavelength, double distance, double dpix) {
  ExitModel m;
  m.kind = PTYRS_FOURIER;
  m.scale = double(ny) * nx;
  const double lz = wavelength * distance;
  m.pixel[0] = lz / (ny * dpix);
  m.pixel[1] = lz / (nx * dpix);
  const double f = M_PI / lz;
  m.plane_quad[0] = f * m.pixel[0] * m.pixel[0];
  m.plane_quad[1] = f * m.pixel[1] * m.pixel[1];
  m.quad.resize(2 * size_t(ny) * nx);
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      const double dy = double(y) - double(ny / 2), dx = double(x) - double(nx / 2);
      set_phase(m.quad, size_t(y) * nx + x,
                f * (dx * dx * m.pixel[1] * m.pixel[1] + dy * dy * m.pixel[0] * m.pixel[0]));
    }
  return m;
}

// The angular-spectrum transfer function T = exp(2πi z √(1 − r)) for
// r = (fx² + (a fy)²)/px² < 1, else 0 (the evanescent band, which sets *banded), lengths in
// wavelengths, on the unshifted fftfreq grid, interleaved re, im.
// ptyrs/src/engine/propagate.rs:505-531
inline std::vector<double> angular_spectrum_tf(int ny, int nx, double wavelength, double py,
                                               double px, double distance, bool* banded) {
  const double w = px / wavelength, ar = px / py, d = distance / wavelength;
  std::vector<double> tf(2 * size_t(ny) * nx, 0.0);
  const double i2piz = 2.0 * M_PI * d;
  auto freq = [](int k, int len) {
    const double kk = k < (len + 1) / 2 ? double(k) : double(k) - double(len);
    return kk * (1.0 / double(len));
  };
  bool band = false;
  for (int y = 0; y < ny; ++y) {
    const double fy = freq(y, ny);
    for (int x = 0; x < nx; ++x) {
      const double fx = freq(x, nx);
      const double ratio = (fx * fx + (ar * fy) * (ar * fy)) / (w * w);
      if (ratio < 1.0) {
        set_phase(tf, size_t(y) * nx + x, i2piz * std::sqrt(1.0 - ratio));
      } else {
        band = true;
      }
    }
  }
  if (banded) *banded = band;
  return tf;
}

// pty-chi's free-space model (ptyrs/src/engine/propagate.rs:66-83, 270-298, 356-366,
// 533-566): Fourier for an infinite distance, the Fresnel transform x ↦ a·DFT(b·x) where its
// output pitch is finer than the source pitch, else angular spectrum. Under the Fresnel
// transform the plane phase is b and the unit-modulus a drops out of every detector
// intensity and projection.
inline ExitModel free_space(int ny, int nx, double wavelength, double py, double px,
                            double distance) {
  if (ny <= 0 || nx <= 0 || ny % 2 || nx % 2) throw Invalid("plane shape must be even and non-zero");
  if (!(std::isfinite(wavelength) && wavelength > 0 && std::isfinite(py) && py > 0 &&
        std::isfinite(px) && px > 0))
    throw Invalid("wavelength and pixel must be finite and > 0");
  if (std::isnan(distance) || distance == 0.0 || distance == -INFINITY)
    throw Invalid("distance must be non-zero and not -inf");
  ExitModel m;
  m.pixel[0] = py;
  m.pixel[1] = px;
  m.quad.assign(2 * size_t(ny) * nx, 0.0);
  for (size_t i = 0; i < size_t(ny) * nx; ++i) m.quad[2 * i] = 1.0;
  if (std::isinf(distance)) {
    m.kind = PTYRS_FOURIER;
    m.scale = double(ny) * nx;
    return m;
  }
  const double w = px / wavelength, ar = px / py, d = distance / wavelength;
  const double h = w / ar;
  const bool fresnel = w * w / std::fabs(d) < 1.0 / nx || h * h / std::fabs(d) < 1.0 / ny;
  if (!fresnel) {
    // The detector sees IDFT(T·DFT(ψ)), which at index k is DFT(T·DFT(ψ)) at −k over n: the
    // counts are stored index-reversed.
    m.kind = PTYRS_ANGULAR_SPECTRUM;
    m.reverse = true;
    m.dft_plane = false;
    m.scale = 1.0;
    m.tf = angular_spectrum_tf(ny, nx, wavelength, py, px, distance, &m.banded);
    return m;
  }
  m.kind = PTYRS_FRESNEL;
  const double fr = w * w / d;
  m.reverse = distance < 0.0;
  m.scale = double(ny) * nx * (fr / ar) * (fr / ar);
  m.plane_quad[0] = M_PI * fr / (ar * ar);
  m.plane_quad[1] = M_PI * fr;
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      const double xx = double(x) - double(nx / 2), yy = double(y) - double(ny / 2);
      set_phase(m.quad, size_t(y) * nx + x, M_PI * fr * (xx * xx + (yy / ar) * (yy / ar)));
    }
  return m;
}

}  // namespace ptyrs
