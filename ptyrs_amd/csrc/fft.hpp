// Line FFTs for the stack passes: one line of N complex values spread over TL = N/E threads of
// a workgroup, E values per thread in registers, as a mixed-radix Stockham autosort
// (Govindaraju et al., "High Performance Discrete Fourier Transforms on Graphics Processors",
// SC 2008: butterfly j of a radix-R stage reads x[j + r·N/R], twiddles by
// exp(∓2πi·(j mod Ns)·r/(Ns·R)), and writes to (j/Ns)·Ns·R + j mod Ns + r·Ns).
//
// Register layout before and after the transform: thread jl of the line holds element
// e in [0, E) at line index jl + TL·e. Every stage radix divides E, so each thread runs E/R
// butterflies per stage and the layout is the same at both ends whatever the plan; stages
// exchange through LDS (one padded region per line), or through registers when TL = 1.
//
// Forward transforms use exp(−2πi·k/N) without normalisation; inverse ones conjugate in and out.
// Twiddles exp(−2πi·t/N), t < N, come from a table computed on the host in double precision.
#pragma once

#include "cplx.hpp"
#include "lds_layout.hpp"

namespace ptyrs {

constexpr int pick_radix(int rem, int e) {
  const int rs[6] = {16, 8, 4, 2, 3, 5};
  for (int i = 0; i < 6; ++i)
    if (rs[i] <= e && e % rs[i] == 0 && rem % rs[i] == 0) return rs[i];
  return 0;
}
constexpr int stage_radix(int n, int e, int s) {
  int rem = n;
  for (int i = 0; i < s; ++i) rem /= pick_radix(rem, e);
  return rem == 1 ? 1 : pick_radix(rem, e);
}
constexpr int stage_ns(int n, int e, int s) {
  int ns = 1;
  for (int i = 0; i < s; ++i) ns *= stage_radix(n, e, i);
  return ns;
}
constexpr int n_stages(int n, int e) {
  int rem = n, s = 0;
  while (rem > 1) {
    int r = pick_radix(rem, e);
    if (r == 0) return -1;
    rem /= r;
    ++s;
  }
  return s;
}
// Placement of element i of a line among the line's LDS words, (i + (i >> SH)·MU)·ST: padded
// so that the exchanges' strides spread over the banks, ST apart when the workgroup's lines are
// interleaved (lds_layout.hpp picks SH, MU and ST per line shape).
template <int SH, int MU, int ST>
struct LdsAt {
  static constexpr int at(int i) { return (i + (i >> SH) * MU) * ST; }
};
using LdsPlain = LdsAt<4, 1, 1>;

// ---- small DFTs in registers, forward direction --------------------------------------------
template <typename T>
__device__ __forceinline__ void dft2(cplx<T>* a) {
  cplx<T> t = a[0];
  a[0] = t + a[1];
  a[1] = t - a[1];
}

template <typename T>
__device__ __forceinline__ void dft3(cplx<T>* a) {
  const T s3 = T(0.86602540378443864676);
  cplx<T> t1 = a[1] + a[2];
  cplx<T> t2 = a[0] - t1 * T(0.5);
  cplx<T> d = a[1] - a[2];
  cplx<T> t3 = {d.y * s3, -d.x * s3};  // -i·sin(2π/3)·d
  a[0] = a[0] + t1;
  a[1] = t2 + t3;
  a[2] = t2 - t3;
}

template <typename T>
__device__ __forceinline__ void dft4(cplx<T>& a0, cplx<T>& a1, cplx<T>& a2, cplx<T>& a3) {
  cplx<T> s02 = a0 + a2, d02 = a0 - a2, s13 = a1 + a3, d13 = mul_mi(a1 - a3);
  a0 = s02 + s13;
  a2 = s02 - s13;
  a1 = d02 + d13;
  a3 = d02 - d13;
}

template <typename T>
__device__ __forceinline__ void dft4(cplx<T>* a) { dft4(a[0], a[1], a[2], a[3]); }

template <typename T>
__device__ __forceinline__ void dft5(cplx<T>* a) {
  // y_k = Σ a_j exp(−2πi jk/5), by the pairs (1,4), (2,3).
  const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
  const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);
  cplx<T> p1 = a[1] + a[4], m1 = a[1] - a[4], p2 = a[2] + a[3], m2 = a[2] - a[3];
  cplx<T> r1 = a[0] + p1 * c1 + p2 * c2;
  cplx<T> r2 = a[0] + p1 * c2 + p2 * c1;
  cplx<T> i1 = m1 * s1 + m2 * s2;   // times -i below
  cplx<T> i2 = m1 * s2 - m2 * s1;
  a[0] = a[0] + p1 + p2;
  a[1] = r1 + mul_mi(i1);
  a[4] = r1 - mul_mi(i1);
  a[2] = r2 + mul_mi(i2);
  a[3] = r2 - mul_mi(i2);
}

template <typename T>
__device__ __forceinline__ void dft8(cplx<T>* a) {
  const T h = T(0.70710678118654752440);
  cplx<T> b[4], c[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    b[k] = a[k] + a[k + 4];
    c[k] = a[k] - a[k + 4];
  }
  // c[k] *= exp(-2πi k/8)
  c[1] = cplx<T>{(c[1].x + c[1].y) * h, (c[1].y - c[1].x) * h};
  c[2] = mul_mi(c[2]);
  c[3] = cplx<T>{(c[3].y - c[3].x) * h, -(c[3].x + c[3].y) * h};
  dft4(b);
  dft4(c);
  a[0] = b[0]; a[2] = b[1]; a[4] = b[2]; a[6] = b[3];
  a[1] = c[0]; a[3] = c[1]; a[5] = c[2]; a[7] = c[3];
}

template <typename T>
__device__ __forceinline__ void dft16(cplx<T>* a) {
  const T c8 = T(0.92387953251128675613), s8 = T(0.38268343236508977173);
  const T h = T(0.70710678118654752440);
  cplx<T> b[4][4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    b[k][0] = a[k]; b[k][1] = a[k + 4]; b[k][2] = a[k + 8]; b[k][3] = a[k + 12];
    dft4(b[k]);
  }
  // b[k][p] *= exp(-2πi kp/16) for kp in {1,2,3,2,4,6,3,6,9}
  const cplx<T> w1 = {c8, -s8}, w2 = {h, -h}, w3 = {s8, -c8};
  const cplx<T> w6 = {-h, -h}, w9 = {-c8, s8};
  b[1][1] = b[1][1] * w1; b[1][2] = b[1][2] * w2; b[1][3] = b[1][3] * w3;
  b[2][1] = b[2][1] * w2; b[2][2] = mul_mi(b[2][2]); b[2][3] = b[2][3] * w6;
  b[3][1] = b[3][1] * w3; b[3][2] = b[3][2] * w6; b[3][3] = b[3][3] * w9;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    cplx<T> t0 = b[0][p], t1 = b[1][p], t2 = b[2][p], t3 = b[3][p];
    dft4(t0, t1, t2, t3);
    a[p] = t0; a[p + 4] = t1; a[p + 8] = t2; a[p + 12] = t3;
  }
}

template <typename T, int R>
__device__ __forceinline__ void dft(cplx<T>* a) {
  if constexpr (R == 2) dft2(a);
  else if constexpr (R == 3) dft3(a);
  else if constexpr (R == 4) dft4(a);
  else if constexpr (R == 5) dft5(a);
  else if constexpr (R == 8) dft8(a);
  else if constexpr (R == 16) dft16(a);
}

// ---- per-thread twiddles ----------------------------------------------------------------------
// A thread that transforms several lines at one lane position jl uses the same twiddles in
// every one: they can be read from the LDS table once into registers. Stage S (NS > 1, lines
// over several threads) needs U·(R − 1) of them, at offset tw_cache_off(N, E, S).
constexpr int tw_cache_off(int n, int e, int s) {
  int off = 0;
  for (int i = 0; i < s; ++i) {
    const int r = stage_radix(n, e, i);
    if (n / e > 1 && stage_ns(n, e, i) > 1) off += (e / r) * (r - 1);
  }
  return off;
}
constexpr int tw_cache_size(int n, int e) {
  const int c = tw_cache_off(n, e, n_stages(n, e));
  return c > 0 ? c : 1;
}

template <typename T, int N, int E, int S = 0>
__device__ __forceinline__ void load_tw_cache(cplx<T>* c, int jl, const cplx<T>* tw) {
  if constexpr (S < n_stages(N, E)) {
    constexpr int R = stage_radix(N, E, S), NS = stage_ns(N, E, S), TL = N / E, U = E / R;
    constexpr int OFF = tw_cache_off(N, E, S);
    if constexpr (TL > 1 && NS > 1) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = (jl + TL * u) % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) c[OFF + u * (R - 1) + r - 1] = tw[k * r * (N / (NS * R))];
      }
    }
    load_tw_cache<T, N, E, S + 1>(c, jl, tw);
  }
}

// ---- one Stockham stage ----------------------------------------------------------------------
// Stage S of the plan for (N, E): butterflies jb = jl + TL·u, u < E/R. `re`/`im` point at this
// line's element 0 in LDS, its elements placed by A (unused when TL == 1 or on the last stage).
// CACHED: the twiddles come from `twc` (load_tw_cache for this jl) in registers.
template <typename T, int N, int E, int S, typename A, bool CACHED>
__device__ __forceinline__ void fft_stage(cplx<T> (&v)[E], int jl, T* re, T* im,
                                          const cplx<T>* __restrict__ tw,
                                          const cplx<T>* twc) {
  constexpr int R = stage_radix(N, E, S);
  constexpr int NS = stage_ns(N, E, S);
  constexpr int TL = N / E;
  constexpr int U = E / R;
  constexpr bool LAST = (NS * R == N);
  if constexpr (!LAST && TL == 1) {
    // One thread per line: the exchange is a register permutation known at compile time.
    cplx<T> t[N];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      cplx<T> a[R];
#pragma unroll
      for (int r = 0; r < R; ++r) a[r] = v[u + U * r];
      if constexpr (NS > 1) {
        const int k = u % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) a[r] = a[r] * tw[k * r * (N / (NS * R))];
      }
      dft<T, R>(a);
#pragma unroll
      for (int r = 0; r < R; ++r) t[(u / NS) * NS * R + u % NS + r * NS] = a[r];
    }
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = t[e];
  } else {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      cplx<T> a[R];
#pragma unroll
      for (int r = 0; r < R; ++r) a[r] = v[u + U * r];
      const int jb = jl + TL * u;
      if constexpr (NS > 1) {
        constexpr int OFF = tw_cache_off(N, E, S);
        const int k = jb % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          if constexpr (CACHED && TL > 1) a[r] = a[r] * twc[OFF + u * (R - 1) + r - 1];
          else a[r] = a[r] * tw[k * r * (N / (NS * R))];
        }
      }
      dft<T, R>(a);
      if constexpr (LAST) {
        // The last stage writes to index jb + r·N/R, the element it read: in place.
#pragma unroll
        for (int r = 0; r < R; ++r) v[u + U * r] = a[r];
      } else {
        const int base = (jb / NS) * NS * R + jb % NS;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int idx = A::at(base + r * NS);
          re[idx] = a[r].x;
          im[idx] = a[r].y;
        }
      }
    }
    if constexpr (!LAST) {
      __syncthreads();
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int idx = A::at(jl + TL * e);
        v[e] = cplx<T>{re[idx], im[idx]};
      }
      __syncthreads();
    }
  }
}

template <typename T, int N, int E, int S, typename A, bool CACHED>
__device__ __forceinline__ void fft_stages(cplx<T> (&v)[E], int jl, T* re, T* im,
                                           const cplx<T>* __restrict__ tw, const cplx<T>* twc) {
  if constexpr (S < n_stages(N, E)) {
    fft_stage<T, N, E, S, A, CACHED>(v, jl, re, im, tw, twc);
    fft_stages<T, N, E, S + 1, A, CACHED>(v, jl, re, im, tw, twc);
  }
}

// Forward (INV = false) or inverse unnormalised DFT of one line, its LDS elements placed by A,
// its twiddles from `tw` (the LDS table) or, CACHED, from `twc` (load_tw_cache for this lane).
// All threads of the workgroup must call it together (it synchronises when the line spans
// several threads).
template <typename T, int N, int E, bool INV, typename A = LdsPlain, bool CACHED = false>
__device__ __forceinline__ void line_fft(cplx<T> (&v)[E], int jl, T* re, T* im,
                                         const cplx<T>* __restrict__ tw,
                                         const cplx<T>* twc = nullptr) {
  static_assert(n_stages(N, E) > 0, "no radix plan for this line length");
  if constexpr (INV) {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e].y = -v[e].y;
  }
  fft_stages<T, N, E, 0, A, CACHED>(v, jl, re, im, tw, twc);
  if constexpr (INV) {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e].y = -v[e].y;
  }
}

}  // namespace ptyrs
