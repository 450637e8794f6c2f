// pty-chi's orthogonal probe relaxation in LSQML (ptyrs/src/engine/opr.rs): with k eigenmodes
// U_i of probe mode 0, frame f sees P_f0 = c_f·P_0 + Σ_i w_fi·U_i and P_fm = P_m (m ≥ 1). The
// eigenmodes are kept on the device in the engine's element type, the weights w_fi and c_f in
// FP64; every sum of the step is an FP64 reduction in a fixed order (per slot, the workgroups'
// partials in order), so results do not depend on the launch.
#pragma once

#include "cplx.hpp"

namespace ptyrs {

constexpr int kOprThreads = 256;

// Per-view probe planes of a stack: plane b·M + m is c_f·P_0 + Σ_i w_fi·U_i for m = 0 (f the
// frame of slot b; the sum in eigenmode order, FP64) and P_m otherwise. Grid (ceil(n/256), nb·M).
// ptyrs/src/engine/opr.rs:92-128
template <typename T>
__global__ void k_opr_planes(cplx<T>* __restrict__ out, const cplx<T>* __restrict__ P,
                             const cplx<T>* __restrict__ U, const double* __restrict__ w,
                             const double* __restrict__ c, const int* __restrict__ order, int k,
                             int M, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int v = blockIdx.y, b = v / M, m = v % M;
  if (m != 0) {
    st(out + (size_t)v * n + j, ld(P + (size_t)m * n + j));
    return;
  }
  const int f = order[b];
  const double cf = c ? c[f] : 1.0;
  const cplx<T> p = ld(P + j);
  double re = cf * double(p.x), im = cf * double(p.y);
  for (int i = 0; i < k; ++i) {
    const double wi = w[(size_t)f * k + i];
    const cplx<T> u = ld(U + (size_t)i * n + j);
    re += wi * double(u.x);
    im += wi * double(u.y);
  }
  st(out + (size_t)v * n + j, cplx<T>{T(re), T(im)});
}

// Slot b's mode-0 probe direction and window power on slice 0: d_b = conj(q·O[win_b])·χ_b0,
// vv_b = |O[win_b]|² (accumulate_probe's num and norm, ptyrs/src/engine/ops.rs:312-343), q
// nullable. Grid (ceil(n/256), nb).
template <typename T>
__global__ void k_opr_frame_dir(cplx<T>* __restrict__ d, T* __restrict__ vv,
                                const cplx<T>* __restrict__ O, int nxo,
                                const cplx<T>* __restrict__ q, const cplx<T>* __restrict__ chi,
                                const int2* __restrict__ corner, const int* __restrict__ order,
                                int N, int M) {
  const int n = N * N;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int b = blockIdx.y, y = j / N, x = j % N;
  const int2 c = corner[order[b]];
  const cplx<T> o = ld(O + (size_t)(c.x + y) * nxo + c.y + x);
  const cplx<T> oq = q ? o * ld(q + j) : o;
  st(d + (size_t)b * n + j, cmulc(oq, ld(chi + (size_t)b * M * n + j)));
  vv[(size_t)b * n + j] = norm2(o);
}

// Per slot and workgroup, FP64 sums over the workgroup's pixels (grid (nblk, nb)):
//   mode 0: Re Σ conj(a)·x_b, Σ |a|²·vv_b          (a = P_0 or an eigenmode, x_b = d_b)
//   mode 1: Re Σ conj(x_b)·a                         (x_b = r_b, a = U_i)
//   mode 2: Σ x_b·conj(a) (re, im)                   (x_b = r_b, a = u)
// into part[(b·nblk + blk)·2 + {0, 1}].
template <typename T, int MODE>
__global__ __launch_bounds__(kOprThreads) void k_opr_sums(const cplx<T>* __restrict__ a,
                                                          const cplx<T>* __restrict__ x,
                                                          const T* __restrict__ vv, int n,
                                                          double* __restrict__ part) {
  __shared__ double red[kOprThreads];
  const int b = blockIdx.y;
  double s0 = 0, s1 = 0;
  for (int j = blockIdx.x * kOprThreads + threadIdx.x; j < n; j += gridDim.x * kOprThreads) {
    const cplx<T> av = ld(a + j), xv = ld(x + (size_t)b * n + j);
    if constexpr (MODE == 0) {
      s0 += double(av.x) * double(xv.x) + double(av.y) * double(xv.y);
      s1 += (double(av.x) * double(av.x) + double(av.y) * double(av.y)) *
            double(vv[(size_t)b * n + j]);
    } else if constexpr (MODE == 1) {
      s0 += double(xv.x) * double(av.x) + double(xv.y) * double(av.y);
    } else {
      s0 += double(xv.x) * double(av.x) + double(xv.y) * double(av.y);
      s1 += double(xv.y) * double(av.x) - double(xv.x) * double(av.y);
    }
  }
  for (int i = 0; i < 2; ++i) {
    red[threadIdx.x] = i == 0 ? s0 : s1;
    __syncthreads();
    for (int s = kOprThreads / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[((size_t)b * gridDim.x + blockIdx.x) * 2 + i] = red[0];
    __syncthreads();
  }
}

// The sum of slot b's partials, in workgroup order.
__device__ __forceinline__ double2 opr_slot_sum(const double* part, int nblk, int b) {
  double s0 = 0, s1 = 0;
  for (int i = 0; i < nblk; ++i) {
    s0 += part[((size_t)b * nblk + i) * 2];
    s1 += part[((size_t)b * nblk + i) * 2 + 1];
  }
  return make_double2(s0, s1);
}

// One thread: the main-mode weights' SGD step c_f += 0.1·Re Σ conj(P_0)·d_b / Σ |P_0|²·vv_b.
// ptyrs/src/engine/opr.rs:171-177
__global__ static void k_opr_intensity(const double* __restrict__ part, int nblk, int nb,
                                       const int* __restrict__ order, double* __restrict__ c) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int b = 0; b < nb; ++b) {
    const double2 s = opr_slot_sum(part, nblk, b);
    c[order[b]] += 0.1 * s.x / s.y;
  }
}

// The batch mean of the slots' directions and their residues r_b = d_b − mean (FP64 sums).
// Grid (ceil(n/256)). ptyrs/src/engine/opr.rs:184-192
template <typename T>
__global__ void k_opr_residues(cplx<T>* __restrict__ r, const cplx<T>* __restrict__ d, int nb,
                               int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double mr = 0, mi = 0;
  for (int b = 0; b < nb; ++b) {
    const cplx<T> v = ld(d + (size_t)b * n + j);
    mr += double(v.x);
    mi += double(v.y);
  }
  mr /= double(nb);
  mi /= double(nb);
  for (int b = 0; b < nb; ++b) {
    const cplx<T> v = ld(d + (size_t)b * n + j);
    st(r + (size_t)b * n + j, cplx<T>{T(double(v.x) - mr), T(double(v.y) - mi)});
  }
}

// One thread: p_b = (Re Σ conj(r_b)·U_i / n² + w_bi) / Σ_b w_bi² into sc[b].
// ptyrs/src/engine/opr.rs:195-203
__global__ static void k_opr_p(const double* __restrict__ part, int nblk, int nb,
                               const int* __restrict__ order, const double* __restrict__ w,
                               int k, int i, double nn, double* __restrict__ sc) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double w2 = 0;
  for (int b = 0; b < nb; ++b) {
    const double wb = w[(size_t)order[b] * k + i];
    w2 += wb * wb;
  }
  for (int b = 0; b < nb; ++b) {
    const double wb = w[(size_t)order[b] * k + i];
    sc[b] = (opr_slot_sum(part, nblk, b).x / (nn * nn) + wb) / w2;
  }
}

// e = Σ_b r_b·p_b / nb per pixel (FP64), and per workgroup Σ |e|² into part[blk].
// Grid (nblk). ptyrs/src/engine/opr.rs:199-205
template <typename T>
__global__ __launch_bounds__(kOprThreads) void k_opr_e(cplx<T>* __restrict__ e,
                                                       const cplx<T>* __restrict__ r,
                                                       const double* __restrict__ p, int nb,
                                                       int n, double* __restrict__ part) {
  __shared__ double red[kOprThreads];
  double s = 0;
  for (int j = blockIdx.x * kOprThreads + threadIdx.x; j < n; j += gridDim.x * kOprThreads) {
    double er = 0, ei = 0;
    for (int b = 0; b < nb; ++b) {
      const cplx<T> v = ld(r + (size_t)b * n + j);
      er += double(v.x) * p[b];
      ei += double(v.y) * p[b];
    }
    er /= double(nb);
    ei /= double(nb);
    const cplx<T> et{T(er), T(ei)};
    st(e + j, et);
    s += double(et.x) * double(et.x) + double(et.y) * double(et.y);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = kOprThreads / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// u = U + s·x with s = scale·√n / √(Σ part / n) (the partials of Σ|x|², in order), in place in
// U, and per workgroup Σ |u|² into out_part[blk]. Grid (nblk). ptyrs/src/engine/opr.rs:205-209
template <typename T>
__global__ __launch_bounds__(kOprThreads) void k_opr_axpy_norm(
    cplx<T>* __restrict__ U, const cplx<T>* __restrict__ x, const double* __restrict__ part,
    int nparts, double scale, int n, double* __restrict__ out_part) {
  __shared__ double red[kOprThreads];
  double tot = 0;
  for (int i = 0; i < nparts; ++i) tot += part[i];
  const double s = scale * sqrt(double(n)) / sqrt(tot / double(n));
  double acc = 0;
  for (int j = blockIdx.x * kOprThreads + threadIdx.x; j < n; j += gridDim.x * kOprThreads) {
    const cplx<T> u = ld(U + j), xv = ld(x + j);
    const cplx<T> v{T(double(u.x) + s * double(xv.x)), T(double(u.y) + s * double(xv.y))};
    st(U + j, v);
    acc += double(v.x) * double(v.x) + double(v.y) * double(v.y);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int k = kOprThreads / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out_part[blockIdx.x] = red[0];
}

// U *= √n / √(Σ part / n) (the partials of Σ|U|², in order), and per workgroup Σ |U|² after
// into out_part[blk]. Grid (nblk). ptyrs/src/engine/opr.rs:208-209
template <typename T>
__global__ __launch_bounds__(kOprThreads) void k_opr_normalise(cplx<T>* __restrict__ U,
                                                               const double* __restrict__ part,
                                                               int nparts, int n,
                                                               double* __restrict__ out_part) {
  __shared__ double red[kOprThreads];
  double tot = 0;
  for (int i = 0; i < nparts; ++i) tot += part[i];
  const double s = sqrt(double(n)) / sqrt(tot / double(n));
  double acc = 0;
  for (int j = blockIdx.x * kOprThreads + threadIdx.x; j < n; j += gridDim.x * kOprThreads) {
    const cplx<T> u = ld(U + j);
    const cplx<T> v{T(double(u.x) * s), T(double(u.y) * s)};
    st(U + j, v);
    acc += double(v.x) * double(v.x) + double(v.y) * double(v.y);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int k = kOprThreads / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out_part[blockIdx.x] = red[0];
}

// One thread: the eigenmode's weights, w_bi += Re Σ conj(u)·d_b / (Σ |u|²·vv_b + 0.1·mean_b of
// those). ptyrs/src/engine/opr.rs:210-220
__global__ static void k_opr_weights(const double* __restrict__ part, int nblk, int nb,
                                     const int* __restrict__ order, double* __restrict__ w,
                                     int k, int i) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double reg = 0;
  for (int b = 0; b < nb; ++b) reg += opr_slot_sum(part, nblk, b).y;
  reg = 0.1 * reg / double(nb);
  for (int b = 0; b < nb; ++b) {
    const double2 s = opr_slot_sum(part, nblk, b);
    w[(size_t)order[b] * k + i] += s.x / (s.y + reg);
  }
}

// r_b −= (Σ r_b·conj(u) / Σ |u|²)·u, the slot sums from part (mode 2 of k_opr_sums), Σ|u|² from
// uu_part (nuu partials). Grid (ceil(n/256), nb). ptyrs/src/engine/opr.rs:221-227
template <typename T>
__global__ void k_opr_project_out(cplx<T>* __restrict__ r, const cplx<T>* __restrict__ u,
                                  const double* __restrict__ part, int nblk,
                                  const double* __restrict__ uu_part, int nuu, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int b = blockIdx.y;
  double uu = 0;
  for (int i = 0; i < nuu; ++i) uu += uu_part[i];
  const double2 c = opr_slot_sum(part, nblk, b);
  const double cr = c.x / uu, ci = c.y / uu;
  const cplx<T> uv = ld(u + j), rv = ld(r + (size_t)b * n + j);
  st(r + (size_t)b * n + j,
     cplx<T>{T(double(rv.x) - (cr * double(uv.x) - ci * double(uv.y))),
             T(double(rv.y) - (cr * double(uv.y) + ci * double(uv.x)))});
}

}  // namespace ptyrs
