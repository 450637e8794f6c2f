// FullObject sampling on the device (pty-chi's FULL_OBJECT shift operator; ptyrs/src/engine/
// sampling.rs:79-209): a view reads its object plane at its corner plus its frame's offset t by
// Fourier interpolation of the whole plane, patch = crop(DFT⁻¹(DFT(O)·exp(+2πi f·t))); the
// adjoint puts a patch at the plane's origin, transforms it and sums conj(exp(+2πi f·t)) times
// it over the views in the spectrum, then one inverse transform.
//
// The object plane has any side (its prime factors decide the cost, not whether it runs), so
// its transforms are a runtime-length Stockham autosort over lines held in LDS: the plan's
// stages are radices 4, 2, then the odd primes; stage s with radix R and Ns (the product of
// the earlier radices) forms output o from R inputs as
//   out[o] = Σ_b in[j + b·L/R]·w^(b·((j mod Ns) + a·Ns)·L/(Ns·R)),
//   a = (o/Ns) mod R, j = (o/(Ns·R))·Ns + o mod Ns, w = exp(−2πi/L)
// (the butterfly of Govindaraju et al., SC 2008, one output per thread, so any radix runs the
// same code). Inverse transforms conjugate in and out; none is normalised.
#pragma once

#include "cplx.hpp"

namespace ptyrs {

constexpr int kAxisMaxStages = 24;
constexpr int kAxisMax = 4096;   // longest object side the transforms take
constexpr int kAxisThreads = 256;

struct AxisPlan {
  int L, nst;
  int radix[kAxisMaxStages];
};

// Where a pass reads and writes: element i of line l of plane p at
// base + p·plane + l·line + i·elem; n_in elements are read (the rest of the L are zero), the
// first n_out written.
struct AxisGeom {
  int nlines;
  long long in_plane;
  int in_line, in_elem, n_in;
  long long out_plane;
  int out_line, out_elem, n_out;
};

enum AxisLoad { AL_PLAIN = 0, AL_PHASE = 1, AL_ABS2 = 2 };
enum AxisStore { AS_PLAIN = 0, AS_REAL_FLOOR = 1, AS_ADD = 2 };

// AL_PHASE: the loaded spectrum value times exp(+2πi(f_y·t_y + f_x·t_x)) of plane p's point
// pts[p] (and by 2πi·f_a for a derivative along axis a: deriv 1 for y, 2 for x; by
// (2πi)²·f_a·f_c for a second derivative: 3 yy, 4 yx, 5 xx), its bins (jy, jx) numpy's fftfreq
// of an ny × nx plane; `cols` says the lines are columns.
struct PhaseArgs {
  const double2* pts;
  int ny, nx, deriv, cols;
};

__device__ __forceinline__ double fftfreq(int j, int n) {
  return double(j < (n + 1) / 2 ? j : j - n) / double(n);
}

// exp(+2πi(fy·t.x + fx·t.y)) in FP64, the argument reduced to [−1/2, 1/2] turns first.
__device__ __forceinline__ double2 sample_phase(double2 t, int jy, int jx, int ny, int nx) {
  double a = fftfreq(jy, ny) * t.x + fftfreq(jx, nx) * t.y;
  a -= rint(a);
  double s, c;
  sincospi(2.0 * a, &s, &c);
  return make_double2(c, s);
}

// One pass of line transforms over planes. Grid (ceil(nlines/G), nplanes), kAxisThreads
// threads, dynamic LDS 4·G·Lp·sizeof(T) (Lp = L + 1 when G > 1, so that a column load's lanes
// fall on different banks).
template <typename T, int LOAD, int STORE>
__global__ __launch_bounds__(kAxisThreads) void k_axis_fft(const cplx<T>* __restrict__ in,
                                                           void* __restrict__ out,
                                                           const cplx<T>* __restrict__ tw,
                                                           AxisPlan plan, int inverse,
                                                           AxisGeom g, int G, PhaseArgs ph,
                                                           T scale) {
  extern __shared__ unsigned char axis_lds[];
  const int L = plan.L, Lp = G > 1 ? L + 1 : L;
  T* re0 = reinterpret_cast<T*>(axis_lds);
  T* im0 = re0 + G * Lp;
  T* re1 = im0 + G * Lp;
  T* im1 = re1 + G * Lp;
  const int p = blockIdx.y;
  const int l0 = blockIdx.x * G;
  const int gl = g.nlines - l0 < G ? g.nlines - l0 : G;
  const int tot = G * L;
  // Lanes run along a line where its elements are adjacent in memory, else across lines.
  const bool along_in = g.in_elem == 1;
  for (int idx = threadIdx.x; idx < tot; idx += kAxisThreads) {
    const int gi = along_in ? idx / L : idx % G;
    const int i = along_in ? idx % L : idx / G;
    cplx<T> v{T(0), T(0)};
    if (gi < gl && i < g.n_in) {
      const int line = l0 + gi;
      v = ld(in + p * g.in_plane + (long long)line * g.in_line + (long long)i * g.in_elem);
      if constexpr (LOAD == AL_ABS2) {
        v = cplx<T>{norm2(v), T(0)};
      } else if constexpr (LOAD == AL_PHASE) {
        const int jy = ph.cols ? i : line, jx = ph.cols ? line : i;
        double2 e = sample_phase(ph.pts[p], jy, jx, ph.ny, ph.nx);
        if (ph.deriv) {
          const double fy = fftfreq(jy, ph.ny), fx = fftfreq(jx, ph.nx);
          if (ph.deriv <= 2) {
            const double w = 2.0 * M_PI * (ph.deriv == 1 ? fy : fx);
            e = make_double2(-e.y * w, e.x * w);  // times i·w
          } else {
            // second derivatives (3: yy, 4: yx, 5: xx): times (2πi)²·f_a·f_c
            const double w = -4.0 * M_PI * M_PI *
                             (ph.deriv == 3 ? fy * fy : ph.deriv == 4 ? fy * fx : fx * fx);
            e = make_double2(e.x * w, e.y * w);
          }
        }
        v = cplx<T>{T(double(v.x) * e.x - double(v.y) * e.y),
                    T(double(v.x) * e.y + double(v.y) * e.x)};
      }
    }
    if (inverse) v.y = -v.y;
    re0[gi * Lp + i] = v.x;
    im0[gi * Lp + i] = v.y;
  }
  __syncthreads();
  T *sre = re0, *sim = im0, *dre = re1, *dim = im1;
  int Ns = 1;
  for (int s = 0; s < plan.nst; ++s) {
    const int R = plan.radix[s];
    const int LR = L / R, NR = Ns * R, stride = L / NR;
    for (int idx = threadIdx.x; idx < tot; idx += kAxisThreads) {
      const int gi = idx / L, o = idx % L;
      const int a = (o / Ns) % R, j = (o / NR) * Ns + o % Ns;
      const int base = ((j % Ns) + a * Ns) * stride;  // < L
      const T* xr = sre + gi * Lp + j;
      const T* xi = sim + gi * Lp + j;
      T sr = T(0), si = T(0);
      int t = 0;
      for (int b = 0; b < R; ++b) {
        const cplx<T> w = ld(tw + t);
        const T ar = xr[b * LR], ai = xi[b * LR];
        sr += ar * w.x - ai * w.y;
        si += ar * w.y + ai * w.x;
        t += base;
        if (t >= L) t -= L;
      }
      dre[gi * Lp + o] = sr;
      dim[gi * Lp + o] = si;
    }
    __syncthreads();
    T* x = sre; sre = dre; dre = x;
    x = sim; sim = dim; dim = x;
    Ns *= R;
  }
  const bool along_out = g.out_elem == 1;
  const int tout = G * g.n_out;
  for (int idx = threadIdx.x; idx < tout; idx += kAxisThreads) {
    const int gi = along_out ? idx / g.n_out : idx % G;
    const int i = along_out ? idx % g.n_out : idx / G;
    if (gi >= gl) continue;
    cplx<T> v{sre[gi * Lp + i], sim[gi * Lp + i]};
    if (inverse) v.y = -v.y;
    v = v * scale;
    const long long at =
        p * g.out_plane + (long long)(l0 + gi) * g.out_line + (long long)i * g.out_elem;
    if constexpr (STORE == AS_REAL_FLOOR) {
      T* o = reinterpret_cast<T*>(out) + at;
      *o = v.x > T(0) ? v.x : T(0);
    } else if constexpr (STORE == AS_ADD) {
      cplx<T>* o = reinterpret_cast<cplx<T>*>(out) + at;
      st(o, ld(o) + v);
    } else {
      st(reinterpret_cast<cplx<T>*>(out) + at, v);
    }
  }
}

// The batch's sampling points and the tall layout its windows take: slot b's point is the
// corner of its frame plus the frame's offset (t = (oy + dy, ox + dx)), and its window sits at
// rows b·N.. of a [nb·N][N] plane, so the stack passes read it at corner (b·N, 0).
__global__ static void k_fo_points(const int2* __restrict__ corner,
                                   const double2* __restrict__ offs,
                                   const int* __restrict__ order, int nb, int N,
                                   double2* __restrict__ pts, int2* __restrict__ tall) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int f = order[b];
  const int2 c = corner[f];
  const double2 d = offs[f];
  pts[b] = make_double2(double(c.x) + d.x, double(c.y) + d.y);
  tall[f] = make_int2(b * N, 0);
}

// acc[j] (+)= Σ_b X[b][j]·conj(exp(+2πi f_j·t_b)) over the slots in order: the spectrum of the
// patches placed by the sampling's adjoint (X the patches' transforms at the plane's origin).
template <typename T>
__global__ void k_fo_place_acc(cplx<T>* __restrict__ acc, const cplx<T>* __restrict__ X,
                               long long xplane, const double2* __restrict__ pts, int nb, int ny,
                               int nx, int accumulate) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ny * nx) return;
  const int jy = j / nx, jx = j % nx;
  double sr = 0, si = 0;
  if (accumulate) {
    const cplx<T> a = ld(acc + j);
    sr = double(a.x);
    si = double(a.y);
  }
  for (int b = 0; b < nb; ++b) {
    const cplx<T> x = ld(X + b * xplane + j);
    const double2 e = sample_phase(pts[b], jy, jx, ny, nx);
    sr += double(x.x) * e.x + double(x.y) * e.y;
    si += double(x.y) * e.x - double(x.x) * e.y;
  }
  st(acc + j, cplx<T>{T(sr), T(si)});
}

// S_m = F_m · Σ_b conj(exp(+2πi f·t_b)) for every mode plane m of F (in place): the spectrum of
// one mode's |W|², the same at every slot, placed at every slot's point.
template <typename T>
__global__ void k_fo_place_same(cplx<T>* __restrict__ F, int nplane,
                                const double2* __restrict__ pts, int nb, int ny, int nx) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ny * nx) return;
  const int jy = j / nx, jx = j % nx;
  double er = 0, ei = 0;
  for (int b = 0; b < nb; ++b) {
    const double2 e = sample_phase(pts[b], jy, jx, ny, nx);
    er += e.x;
    ei -= e.y;
  }
  const size_t n = size_t(ny) * nx;
  for (int m = 0; m < nplane; ++m) {
    const cplx<T> f = ld(F + m * n + j);
    st(F + m * n + j, cplx<T>{T(double(f.x) * er - double(f.y) * ei),
                              T(double(f.x) * ei + double(f.y) * er)});
  }
}

// E[j] (+)= Σ_b conj(exp(+2πi f_j·t_b)) over the slots in order (FP64): the spectrum of a
// unit patch at the origin placed at every slot's point, per transform of the patch.
__global__ static void k_fo_phase_acc(double2* __restrict__ E, const double2* __restrict__ pts,
                                      int nb, int ny, int nx, int accumulate) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ny * nx) return;
  const int jy = j / nx, jx = j % nx;
  double er = 0, ei = 0;
  if (accumulate) {
    er = E[j].x;
    ei = E[j].y;
  }
  for (int b = 0; b < nb; ++b) {
    const double2 e = sample_phase(pts[b], jy, jx, ny, nx);
    er += e.x;
    ei -= e.y;
  }
  E[j] = make_double2(er, ei);
}

// acc[j] = (Σ_m F_m[j])·E[j]: every frame's summed mode patches placed at its point.
template <typename T>
__global__ void k_fo_times_sum(cplx<T>* __restrict__ acc, const cplx<T>* __restrict__ F,
                               const double2* __restrict__ E, int M, long long no) {
  const long long j = blockIdx.x * 256ll + threadIdx.x;
  if (j >= no) return;
  double fr = 0, fi = 0;
  for (int m = 0; m < M; ++m) {
    const cplx<T> f = ld(F + m * no + j);
    fr += double(f.x);
    fi += double(f.y);
  }
  const double2 e = E[j];
  st(acc + j, cplx<T>{T(fr * e.x - fi * e.y), T(fr * e.y + fi * e.x)});
}

// Per slot the modes' object-direction terms on its window: out[b][j] = Σ_m conj(W_m[j]·q[j])·
// χ_bm[j] (W = P·q and q null, or a slice's incident waves and its plane phase), natural layout;
// with w_view every slot's modes have their own W (planes slot·M + m).
template <typename T>
__global__ void k_fo_chi_terms(cplx<T>* __restrict__ out, const cplx<T>* __restrict__ W,
                               const cplx<T>* __restrict__ chi, int M, int n, int w_view = 0,
                               const cplx<T>* __restrict__ q = nullptr) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int b = blockIdx.y;
  const cplx<T>* wb = w_view ? W + (size_t)b * M * n : W;
  const cplx<T> qj = q ? ld(q + j) : cplx<T>{T(1), T(0)};
  cplx<T> s{T(0), T(0)};
  for (int m = 0; m < M; ++m)
    s = s + cmulc(ld(wb + (size_t)m * n + j) * qj, ld(chi + ((size_t)b * M + m) * n + j));
  st(out + (size_t)b * n + j, s);
}

// Per slot and workgroup the position sums of AP's, DM's and ML's step under FullObject
// (ptyrs/src/engine/features.rs:960-1025 with `derivatives` of a sampled window): over the
// slot's views (every mode) num_a = Re Σ conj(x − W·S)·W·∂_a S and den_a = Σ |W·∂_a S|², x the
// projected exit waves ([slot·M + m][n], natural), W = P·q ([m][n]), S the slot's read
// ([slot][n]) and ∂_a S its derivatives ([slot·2 + a][n]). part[slot][blk] = {num_y, num_x,
// den_y, den_x}. Grid (nblk, nb).
template <typename T>
__global__ __launch_bounds__(THREADS) void k_pos_sums_fo(
    const cplx<T>* __restrict__ x, const cplx<T>* __restrict__ W, const cplx<T>* __restrict__ win,
    const cplx<T>* __restrict__ dwin, int n, int M, double* __restrict__ part) {
  __shared__ double red[THREADS];
  const int b = blockIdx.y;
  double s[4] = {0, 0, 0, 0};
  for (int j = blockIdx.x * THREADS + threadIdx.x; j < n; j += gridDim.x * THREADS) {
    const cplx<T> o = ld(win + (size_t)b * n + j);
    const cplx<T> dy = ld(dwin + ((size_t)b * 2) * n + j), dx = ld(dwin + ((size_t)b * 2 + 1) * n + j);
    for (int m = 0; m < M; ++m) {
      const cplx<T> w = ld(W + (size_t)m * n + j);
      const cplx<T> r = ld(x + ((size_t)b * M + m) * n + j) - w * o;
      const cplx<T> uy = w * dy, ux = w * dx;
      s[0] += double(cmulc(r, uy).x);
      s[1] += double(cmulc(r, ux).x);
      s[2] += double(norm2(uy));
      s[3] += double(norm2(ux));
    }
  }
  for (int i = 0; i < 4; ++i) {
    const double t = block_reduce<THREADS>(s[i], red, AddOp());
    if (threadIdx.x == 0) part[((size_t)b * gridDim.x + blockIdx.x) * 4 + i] = t;
  }
}

// ---- AP, DM and RAAR under FullObject: the object's PCG solve (ptyrs/src/engine/features.rs:
// 726-793) and the probe's step through the sampled windows --------------------------------------
// One axis of the diagonal's kernel, the spectrum of |h_t[−v]|² (h_t the one-axis kernel that
// samples at t) at index m of n: (1 − m/n)·exp(−2πi t m/n) + (m/n)·exp(−2πi t (m − n)/n).
// ptyrs/src/engine/sampling.rs:217-230
__device__ __forceinline__ double2 square_spectrum(double t, int m, int n) {
  const double a = double(m) / double(n);
  double s0, c0, s1, c1;
  double u = t * a;
  u -= rint(u);
  sincospi(-2.0 * u, &s0, &c0);
  double v = t * (a - 1.0);
  v -= rint(v);
  sincospi(-2.0 * v, &s1, &c1);
  return make_double2((1.0 - a) * c0 + a * c1, (1.0 - a) * s0 + a * s1);
}

// acc[j] (+)= X[j]·Σ_b K_b[j], K_b the diagonal's kernel at slot b's point (the product of the
// two axes' square_spectrum): the spectrum of the one patch X (the same at every slot) placed
// for AᴴA's exact diagonal (Sampler::diag, ptyrs/src/engine/sampling.rs:196-209).
template <typename T>
__global__ void k_fo_diag_acc(cplx<T>* __restrict__ acc, const cplx<T>* __restrict__ X,
                              const double2* __restrict__ pts, int nb, int ny, int nx,
                              int accumulate) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ny * nx) return;
  const int jy = j / nx, jx = j % nx;
  double kr = 0, ki = 0;
  for (int b = 0; b < nb; ++b) {
    const double2 a = square_spectrum(pts[b].x, jy, ny), c = square_spectrum(pts[b].y, jx, nx);
    kr += a.x * c.x - a.y * c.y;
    ki += a.x * c.y + a.y * c.x;
  }
  const cplx<T> x = ld(X + j);
  double sr = double(x.x) * kr - double(x.y) * ki, si = double(x.x) * ki + double(x.y) * kr;
  if (accumulate) {
    const cplx<T> a = ld(acc + j);
    sr += double(a.x);
    si += double(a.y);
  }
  st(acc + j, cplx<T>{T(sr), T(si)});
}

// out = Σ_m |W_m|² (a real patch as complex, zero imaginary part). Grid (ceil(n/256)).
template <typename T>
__global__ void k_fo_power(cplx<T>* __restrict__ out, const cplx<T>* __restrict__ W, int M, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  T s = 0;
  for (int m = 0; m < M; ++m) s += norm2(ld(W + (size_t)m * n + j));
  st(out + j, cplx<T>{s, T(0)});
}

// planes[b][j] *= p[j].x over nb planes (the reads times Σ_m |W_m|²). Grid (ceil(n/256), nb).
template <typename T>
__global__ void k_fo_times_power(cplx<T>* __restrict__ planes, const cplx<T>* __restrict__ p,
                                 int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  cplx<T>* v = planes + (size_t)blockIdx.y * n + j;
  st(v, ld(v) * ld(p + j).x);
}

// The probe step's sums through the reads (slots in order, Kahan): num_m (+)= Σ_b conj(S_b·q)·
// (ψ_{f_b, m} − W_m·S_b) and, from the mode-0 blocks, norm (+)= Σ_b |S_b|²; ψ the carried waves by
// frame ([frame·M + m][n]), S the stack's reads ([slot][n]). Grid (ceil(n/256), M).
template <typename T>
__global__ void k_fo_probe_num(cplx<T>* __restrict__ num, T* __restrict__ norm,
                               const cplx<T>* __restrict__ win, const cplx<T>* __restrict__ q,
                               const cplx<T>* __restrict__ W, const cplx<T>* __restrict__ psi,
                               const int* __restrict__ order, int nb, int n, int M,
                               int accumulate) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int m = blockIdx.y;
  const cplx<T> qj = ld(q + j), w = ld(W + (size_t)m * n + j);
  kahan<T> sr, si, sn;
  for (int b = 0; b < nb; ++b) {
    const cplx<T> o = ld(win + (size_t)b * n + j);
    const cplx<T> r = ld(psi + ((size_t)order[b] * M + m) * n + j) - w * o;
    const cplx<T> g = cmulc(o * qj, r);
    sr.add(g.x);
    si.add(g.y);
    if (m == 0) sn.add(norm2(o));
  }
  cplx<T> v{sr.s, si.s};
  if (accumulate) v = v + ld(num + (size_t)m * n + j);
  st(num + (size_t)m * n + j, v);
  if (m == 0) norm[j] = accumulate ? norm[j] + sn.s : sn.s;
}

// Per mode and workgroup Σ |s_m·q|²·norm (FP64): ‖A·s‖² of the probe step (the reads do not
// move with the probe) into part[m·gridDim.x + blk]. Grid (nblk, M).
template <typename T>
__global__ __launch_bounds__(THREADS) void k_fo_probe_vpow(double* __restrict__ part,
                                                           const cplx<T>* __restrict__ s,
                                                           const cplx<T>* __restrict__ q,
                                                           const T* __restrict__ norm, int n) {
  __shared__ double red[THREADS];
  const int m = blockIdx.y;
  double a = 0;
  for (int j = blockIdx.x * THREADS + threadIdx.x; j < n; j += gridDim.x * THREADS)
    a += double(norm2(ld(s + (size_t)m * n + j) * ld(q + j))) * double(norm[j]);
  const double t = block_reduce<THREADS>(a, red, AddOp());
  if (threadIdx.x == 0) part[(size_t)m * gridDim.x + blockIdx.x] = t;
}

// The PCG's vector steps over the object (no values), FP64 partial sums per workgroup into
// part[blk]; sc = {α, β, rz, reg} on the device. pinv = 1/(D + reg), 0 where held.
template <typename T>
__global__ void k_pcg_pinv(T* __restrict__ pinv, const T* __restrict__ D,
                           const T* __restrict__ dmax, double inertia, long long no,
                           double* __restrict__ sc) {
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  const double mx = double(dmax[0]), reg = mx * inertia;
  if (i == 0) sc[3] = reg;
  if (i >= no) return;
  const double v = double(D[i]) + reg;
  pinv[i] = v > 1e-5 * mx ? T(1.0 / v) : T(0);
}
// r −= ax (r holding the right-hand side); p = r·pinv; Σ |r|²·pinv.
template <typename T>
__global__ __launch_bounds__(THREADS) void k_pcg_start(cplx<T>* __restrict__ r,
                                                       const cplx<T>* __restrict__ ax,
                                                       const T* __restrict__ pinv,
                                                       cplx<T>* __restrict__ p, long long no,
                                                       double* __restrict__ part) {
  __shared__ double red[THREADS];
  double a = 0;
  for (long long i = blockIdx.x * (long long)THREADS + threadIdx.x; i < no;
       i += (long long)gridDim.x * THREADS) {
    const cplx<T> v = ld(r + i) - ld(ax + i);
    st(r + i, v);
    st(p + i, v * pinv[i]);
    a += double(norm2(v)) * double(pinv[i]);
  }
  a = block_reduce<THREADS>(a, red, AddOp());
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}
// ap += reg·p; Re Σ conj(p)·ap.
template <typename T>
__global__ __launch_bounds__(THREADS) void k_pcg_ap(cplx<T>* __restrict__ ap,
                                                    const cplx<T>* __restrict__ p,
                                                    const double* __restrict__ sc, long long no,
                                                    double* __restrict__ part) {
  __shared__ double red[THREADS];
  const T reg = T(sc[3]);
  double a = 0;
  for (long long i = blockIdx.x * (long long)THREADS + threadIdx.x; i < no;
       i += (long long)gridDim.x * THREADS) {
    const cplx<T> pv = ld(p + i), v = ld(ap + i) + pv * reg;
    st(ap + i, v);
    a += double(cmulc(pv, v).x);
  }
  a = block_reduce<THREADS>(a, red, AddOp());
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}
// x += α·p; r −= α·ap; Σ |r|²·pinv.
template <typename T>
__global__ __launch_bounds__(THREADS) void k_pcg_step(cplx<T>* __restrict__ x,
                                                      cplx<T>* __restrict__ r,
                                                      const cplx<T>* __restrict__ p,
                                                      const cplx<T>* __restrict__ ap,
                                                      const T* __restrict__ pinv,
                                                      const double* __restrict__ sc,
                                                      long long no, double* __restrict__ part) {
  __shared__ double red[THREADS];
  const double alpha = sc[0];
  double a = 0;
  for (long long i = blockIdx.x * (long long)THREADS + threadIdx.x; i < no;
       i += (long long)gridDim.x * THREADS) {
    const cplx<T> xv = ld(x + i), pv = ld(p + i), rv = ld(r + i), av = ld(ap + i);
    st(x + i, cplx<T>{T(double(xv.x) + alpha * pv.x), T(double(xv.y) + alpha * pv.y)});
    const cplx<T> nr{T(double(rv.x) - alpha * av.x), T(double(rv.y) - alpha * av.y)};
    st(r + i, nr);
    a += double(norm2(nr)) * double(pinv[i]);
  }
  a = block_reduce<THREADS>(a, red, AddOp());
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}
// p = r·pinv + β·p.
template <typename T>
__global__ void k_pcg_dir(cplx<T>* __restrict__ p, const cplx<T>* __restrict__ r,
                          const T* __restrict__ pinv, const double* __restrict__ sc,
                          long long no) {
  const long long i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= no) return;
  const double beta = sc[1];
  const cplx<T> rv = ld(r + i) * pinv[i], pv = ld(p + i);
  st(p + i, cplx<T>{T(rv.x + beta * pv.x), T(rv.y + beta * pv.y)});
}
// One workgroup: the sum of part[0..nblk) in a fixed order, then mode 0: rz = sum; mode 1: α =
// rz/sum where sum > 0, else 0; mode 2: β = sum/rz where rz > 0, else 0, and rz = sum.
__global__ __launch_bounds__(THREADS) static void k_pcg_scalars(const double* __restrict__ part,
                                                                int nblk, double* __restrict__ sc,
                                                                int mode) {
  __shared__ double red[THREADS];
  double a = 0;
  for (int i = threadIdx.x; i < nblk; i += THREADS) a += part[i];
  a = block_reduce<THREADS>(a, red, AddOp());
  if (threadIdx.x != 0) return;
  if (mode == 0) {
    sc[2] = a;
  } else if (mode == 1) {
    sc[0] = a > 0.0 ? sc[2] / a : 0.0;
  } else {
    sc[1] = sc[2] > 0.0 ? a / sc[2] : 0.0;
    sc[2] = a;
  }
}

// dO /= √(pre² + floor) (LSQML's preconditioned object direction).
template <typename T>
__global__ void k_fo_precondition(cplx<T>* __restrict__ dO, const T* __restrict__ pre,
                                  const T* __restrict__ floor, long long no) {
  const long long p = blockIdx.x * 256ll + threadIdx.x;
  if (p >= no) return;
  const T pr = pre[p];
  const T inv = T(1) / sqrt(pr * pr + floor[0]);
  st(dO + p, ld(dO + p) * inv);
}

// Per workgroup the max over its pixels of Σ_m D_m on the slots' windows: D is [M][rows][N]
// real with slot b's window at rows b·N (stride `mstride` between modes). Grid (blocks).
template <typename T>
__global__ __launch_bounds__(256) void k_fo_dens_max(const T* __restrict__ D, long long mstride,
                                                     int M, long long count,
                                                     T* __restrict__ part) {
  __shared__ T red[256];
  T mx = T(0);
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < count; i += gridDim.x * 256ll) {
    T s = T(0);
    for (int m = 0; m < M; ++m) s += D[m * mstride + i];
    mx = s > mx ? s : mx;
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] = red[threadIdx.x + k] > red[threadIdx.x]
                                                ? red[threadIdx.x + k]
                                                : red[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

}  // namespace ptyrs
