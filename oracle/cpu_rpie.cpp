// FP64 far-field rPIE batch on the CPU: a C++ restatement of the oracle's `Ptycho.rpie_batch`
// (oracle/ptycho.py), which follows ptyrs/src/engine/ptycho.rs:900-1086 (batch) and :518-546
// (modulus and Δ), with the far-field exit of ptyrs/src/engine/propagate.rs:52-62.
//
// TEST INFRASTRUCTURE ONLY: the CPU baseline timed beside the GPU, and checked against the
// numpy oracle by tests/test_oracle_cpu.py. Nothing in the product links or loads it.
//
// Planes are split into real and imaginary arrays (row-major, natural layout). The unitary 2D
// DFT of the shifted layout is computed in the natural layout: for even sides the half-plane
// roll before and after the transform multiplies the spectrum by (−1)^(ky+kx), which leaves
// |Ψ| unchanged and cancels between the forward and the inverse transform, so
//     Δ = IDFT(r ⊙ DFT(ψ))/(N·N) − ψ,  r = √(I+1e-7)/(√(Σ_m|DFT ψ_m|²/(N·N)) + 1e-7),
// with I the counts indexed in FFT order (the detector's centre at index 0) and r = 1 where
// a count is masked (negative). The transform is radix-2 over whole rows: each butterfly
// combines two rows with one twiddle, so its inner loop runs over contiguous columns; the x
// axis runs the same stages on the transposed plane, and the spectrum stays transposed through
// the pointwise modulus step until the inverse transform transposes it back.
//
// Threads (OpenMP): the exit waves of a batch's frames in parallel; the object sum per object
// row and the probe sum per probe row, each pixel summing its terms in view order, so the
// result does not depend on the thread count.

#include <omp.h>

#include <cmath>
#include <cstdint>
#include <algorithm>
#include <cstring>
#include <vector>

#define HOT __attribute__((target_clones("fma", "default")))

namespace {

struct Plan {
  int n = 0, logn = 0;
  std::vector<double> wr, wi;  // twiddles exp(-2πi j/(2h)) at [h + j], h = 1, 2, 4, …, n/2
  std::vector<int> rev;
  explicit Plan(int n_) : n(n_), wr(n_), wi(n_), rev(n_) {
    while ((1 << logn) < n) ++logn;
    for (int h = 1; h < n; h *= 2)
      for (int j = 0; j < h; ++j) {
        const double a = -M_PI * j / h;
        wr[h + j] = std::cos(a);
        wi[h + j] = std::sin(a);
      }
    for (int i = 0; i < n; ++i) {
      int r = 0;
      for (int b = 0; b < logn; ++b) r |= ((i >> b) & 1) << (logn - 1 - b);
      rev[i] = r;
    }
  }
};

// Unnormalised DFT along y of every column of an n×n plane (sign −1 forward, +1 inverse).
HOT void dft_rows_of_columns(const Plan& p, double* re, double* im, bool inverse) {
  const int n = p.n;
  for (int i = 0; i < n; ++i) {
    const int r = p.rev[i];
    if (i < r) {
      double* a = re + (size_t)i * n; double* b = re + (size_t)r * n;
      double* c = im + (size_t)i * n; double* d = im + (size_t)r * n;
      for (int x = 0; x < n; ++x) {
        double t = a[x]; a[x] = b[x]; b[x] = t;
        t = c[x]; c[x] = d[x]; d[x] = t;
      }
    }
  }
  const double s = inverse ? -1.0 : 1.0;
  for (int h = 1; h < n; h *= 2)
    for (int base = 0; base < n; base += 2 * h)
      for (int j = 0; j < h; ++j) {
        const double wr = p.wr[h + j], wi = s * p.wi[h + j];
        double* ar = re + (size_t)(base + j) * n;
        double* ai = im + (size_t)(base + j) * n;
        double* br = re + (size_t)(base + j + h) * n;
        double* bi = im + (size_t)(base + j + h) * n;
        for (int x = 0; x < n; ++x) {
          const double tr = wr * br[x] - wi * bi[x];
          const double ti = wr * bi[x] + wi * br[x];
          br[x] = ar[x] - tr;
          bi[x] = ai[x] - ti;
          ar[x] += tr;
          ai[x] += ti;
        }
      }
}

// dst = srcᵀ (n×n), contiguous writes.
HOT void transpose(const double* src, double* dst, int n) {
  constexpr int B = 16;
  for (int x0 = 0; x0 < n; x0 += B)
    for (int y0 = 0; y0 < n; y0 += B)
      for (int x = x0; x < x0 + B && x < n; ++x)
        for (int y = y0; y < y0 + B && y < n; ++y) dst[(size_t)x * n + y] = src[(size_t)y * n + x];
}

// Forward 2D DFT of the natural plane (re, im), which it overwrites, into (tr, ti) in
// transposed order: tr[kx·n + ky]. The modulus step is pointwise, so the spectrum is never
// transposed back.
void dft2_forward_t(const Plan& p, double* re, double* im, double* tr, double* ti) {
  dft_rows_of_columns(p, re, im, false);
  transpose(re, tr, p.n);
  transpose(im, ti, p.n);
  dft_rows_of_columns(p, tr, ti, false);
}

// Inverse of the above: the transposed spectrum (tr, ti), which it overwrites, to the natural
// plane (re, im).
void dft2_inverse_t(const Plan& p, double* tr, double* ti, double* re, double* im) {
  dft_rows_of_columns(p, tr, ti, true);
  transpose(tr, re, p.n);
  transpose(ti, im, p.n);
  dft_rows_of_columns(p, re, im, true);
}

// Buffers kept between calls, so that a batch touches no fresh pages (grown, never shrunk).
struct Workspace {
  std::vector<double> d_re, d_im, ov;
  std::vector<std::vector<double>> per_thread;
};
Workspace& workspace() {
  static Workspace w;
  return w;
}

void grow(std::vector<double>& v, size_t n) {
  if (v.size() < n) v.resize(n);
}

}  // namespace

extern "C" {

// One rPIE batch of `nb` frames (`frames`, ascending) from the batch-start object and probe;
// `nmode` probe modes, one object mode. Complex arrays are interleaved doubles: object
// [nyo][nxo], probe [nmode][n][n], far-field phase `quad` [n][n]. Counts are float32
// [nframe][n][n] in detector order (negative = masked); corners [nframe][2] (y, x). A step
// flag of 0 leaves that array unchanged. Returns 0, or −1 for a side that is not a power of
// two or a window outside the object.
int cpu_rpie_batch(int n, int nmode, int nyo, int nxo, double* obj, double* probe,
                   const double* quad, const float* counts, const int64_t* corners,
                   const int64_t* frames, int nb, int has_obj, double obj_step, int has_probe,
                   double probe_step, double obj_alpha, double probe_alpha, int threads) {
  if (n < 2 || (n & (n - 1)) || nmode < 1 || nb < 1) return -1;
  for (int b = 0; b < nb; ++b) {
    const int64_t oy = corners[2 * frames[b]], ox = corners[2 * frames[b] + 1];
    if (oy < 0 || ox < 0 || oy + n > nyo || ox + n > nxo) return -1;
  }
  if (threads > 0) omp_set_num_threads(threads);
  const Plan plan(n);
  const size_t nn = (size_t)n * n;
  const double inv_nn = 1.0 / (double)nn;
  // Batch-start copies the updates read.
  std::vector<double> o0(obj, obj + 2 * (size_t)nyo * nxo);
  std::vector<double> p0(probe, probe + 2 * nmode * nn);
  // W = P·q and |P|², natural layout.
  std::vector<double> w_re(nmode * nn), w_im(nmode * nn), w2(nmode * nn);
  for (size_t i = 0; i < nmode * nn; ++i) {
    const double pr = p0[2 * i], pi = p0[2 * i + 1];
    const double qr = quad[2 * (i % nn)], qi = quad[2 * (i % nn) + 1];
    w_re[i] = pr * qr - pi * qi;
    w_im[i] = pr * qi + pi * qr;
    w2[i] = pr * pr + pi * pi;
  }
  // Δ of every view (split planes) and the per-frame max |O_win|².
  Workspace& ws = workspace();
  grow(ws.d_re, (size_t)nb * nmode * nn);
  grow(ws.d_im, (size_t)nb * nmode * nn);
  double* const d_re = ws.d_re.data();
  double* const d_im = ws.d_im.data();
  std::vector<double> omax(nb);
  ws.per_thread.resize(std::max(ws.per_thread.size(), (size_t)omp_get_max_threads()));
#pragma omp parallel
  {
    std::vector<double>& buf = ws.per_thread[omp_get_thread_num()];
    grow(buf, (2 * nmode + 3) * nn);
    double* const sr = buf.data();
    double* const si = sr + nmode * nn;
    double* const pr = si + nmode * nn;
    double* const pi = pr + nn;
    double* const ratio = pi + nn;
#pragma omp for schedule(dynamic)
    for (int b = 0; b < nb; ++b) {
      const int64_t f = frames[b], oy = corners[2 * f], ox = corners[2 * f + 1];
      double mx = 0.0;
      for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
          const size_t o = 2 * ((size_t)(oy + y) * nxo + (ox + x));
          const double a = o0[o] * o0[o] + o0[o + 1] * o0[o + 1];
          mx = a > mx ? a : mx;
        }
      omax[b] = mx;
      double* dr = d_re + (size_t)b * nmode * nn;
      double* di = d_im + (size_t)b * nmode * nn;
      for (int m = 0; m < nmode; ++m) {
        for (int y = 0; y < n; ++y)
          for (int x = 0; x < n; ++x) {
            const size_t i = (size_t)y * n + x, o = 2 * ((size_t)(oy + y) * nxo + (ox + x));
            const double ar = w_re[m * nn + i], ai = w_im[m * nn + i];
            pr[i] = ar * o0[o] - ai * o0[o + 1];
            pi[i] = ar * o0[o + 1] + ai * o0[o];
          }
        // ψ kept for Δ.
        std::memcpy(dr + m * nn, pr, nn * sizeof(double));
        std::memcpy(di + m * nn, pi, nn * sizeof(double));
        dft2_forward_t(plan, pr, pi, sr + m * nn, si + m * nn);
      }
      // Spectra in transposed order: index kx·n + ky.
      const float* cnt = counts + (size_t)f * nn;
      const int h = n / 2;
      for (int kx = 0; kx < n; ++kx)
        for (int ky = 0; ky < n; ++ky) {
          const size_t i = (size_t)kx * n + ky;
          double s = 0.0;
          for (int m = 0; m < nmode; ++m) s += sr[m * nn + i] * sr[m * nn + i] + si[m * nn + i] * si[m * nn + i];
          const double c = cnt[(size_t)((ky + h) % n) * n + (kx + h) % n];
          ratio[i] = c < 0.0 ? 1.0 : std::sqrt(c + 1e-7) / (std::sqrt(s * inv_nn) + 1e-7);
        }
      for (int m = 0; m < nmode; ++m) {
        double* xr = sr + m * nn;
        double* xi = si + m * nn;
        for (size_t i = 0; i < nn; ++i) {
          xr[i] *= ratio[i];
          xi[i] *= ratio[i];
        }
        dft2_inverse_t(plan, xr, xi, pr, pi);
        double* er = dr + m * nn;
        double* ei = di + m * nn;
        for (size_t i = 0; i < nn; ++i) {
          er[i] = pr[i] * inv_nn - er[i];
          ei[i] = pi[i] * inv_nn - ei[i];
        }
      }
    }
  }
  if (has_obj) {
    // Overlap of |W_m|² over the batch's views per probe mode, and max over frames and pixels
    // of Σ_m of it read back at each view.
    grow(ws.ov, (size_t)nmode * nyo * nxo);
    double* const ov = ws.ov.data();
    std::fill(ov, ov + (size_t)nmode * nyo * nxo, 0.0);
#pragma omp parallel for schedule(static)
    for (int yo = 0; yo < nyo; ++yo)
      for (int b = 0; b < nb; ++b) {
        const int64_t f = frames[b], oy = corners[2 * f], ox = corners[2 * f + 1];
        if (yo < oy || yo >= oy + n) continue;
        for (int m = 0; m < nmode; ++m) {
          double* row = ov + ((size_t)m * nyo + yo) * nxo + ox;
          const double* src = w2.data() + m * nn + (size_t)(yo - oy) * n;
          for (int x = 0; x < n; ++x) row[x] += src[x];
        }
      }
    double pmax = 0.0;
#pragma omp parallel for reduction(max : pmax) schedule(static)
    for (int b = 0; b < nb; ++b) {
      const int64_t f = frames[b], oy = corners[2 * f], ox = corners[2 * f + 1];
      for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
          double s = 0.0;
          for (int m = 0; m < nmode; ++m) s += ov[((size_t)m * nyo + oy + y) * nxo + ox + x];
          pmax = s > pmax ? s : pmax;
        }
    }
    const double a = obj_alpha;
#pragma omp parallel
    {
      std::vector<double> acc(2 * (size_t)nxo);
#pragma omp for schedule(static)
      for (int yo = 0; yo < nyo; ++yo) {
        bool touched = false;
        for (int b = 0; b < nb; ++b) {
          const int64_t f = frames[b], oy = corners[2 * f], ox = corners[2 * f + 1];
          if (yo < oy || yo >= oy + n) continue;
          if (!touched) std::fill(acc.begin(), acc.end(), 0.0);
          touched = true;
          const int y = (int)(yo - oy);
          for (int m = 0; m < nmode; ++m) {
            const size_t k = ((size_t)b * nmode + m) * nn + (size_t)y * n;
            const double* ovr = ov + ((size_t)m * nyo + yo) * nxo + ox;
            double* ar = acc.data() + 2 * ox;
            for (int x = 0; x < n; ++x) {
              const size_t i = (size_t)m * nn + (size_t)y * n + x;
              const double den = (1.0 - a) * ovr[x] + a * pmax;
              const double cr = w_re[i] / den, ci = -w_im[i] / den;  // conj(W)/den
              const double er = d_re[k + x], ei = d_im[k + x];
              ar[2 * x] += cr * er - ci * ei;
              ar[2 * x + 1] += cr * ei + ci * er;
            }
          }
        }
        if (!touched) continue;
        double* orow = obj + 2 * (size_t)yo * nxo;
        for (int x = 0; x < 2 * nxo; ++x) orow[x] = o0[2 * (size_t)yo * nxo + x] + acc[x] * obj_step;
      }
    }
  }
  if (has_probe) {
    const double a = probe_alpha, scale = probe_step / nb;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < n; ++y) {
      std::vector<double> acc(2 * (size_t)nmode * n, 0.0);
      for (int b = 0; b < nb; ++b) {
        const int64_t f = frames[b], oy = corners[2 * f], ox = corners[2 * f + 1];
        for (int x = 0; x < n; ++x) {
          const size_t o = 2 * ((size_t)(oy + y) * nxo + ox + x), i = (size_t)y * n + x;
          const double orr = o0[o], oi = o0[o + 1];
          const double qr = quad[2 * i], qi = quad[2 * i + 1];
          const double wr = orr * qr - oi * qi, wi = orr * qi + oi * qr;  // O·q
          const double den = (1.0 - a) * (orr * orr + oi * oi) + a * omax[b];
          const double gr = wr / den, gi = -wi / den;  // conj(O·q)/den
          for (int m = 0; m < nmode; ++m) {
            const size_t k = ((size_t)b * nmode + m) * nn + i;
            acc[2 * ((size_t)m * n + x)] += gr * d_re[k] - gi * d_im[k];
            acc[2 * ((size_t)m * n + x) + 1] += gr * d_im[k] + gi * d_re[k];
          }
        }
      }
      for (int m = 0; m < nmode; ++m)
        for (int x = 0; x < n; ++x) {
          const size_t i = 2 * ((size_t)m * nn + (size_t)y * n + x);
          probe[i] = p0[i] + scale * acc[2 * ((size_t)m * n + x)];
          probe[i + 1] = p0[i + 1] + scale * acc[2 * ((size_t)m * n + x) + 1];
        }
    }
  }
  return 0;
}

}  // extern "C"
