// Device kernels of the per-position reconstruction step.
//
// Data layout in HBM (n = N·N pixels per plane, M modes):
//   object O        [nobj][nyo][nxo]   complex, row-major
//   probe P, Wt     [M][N][N]          complex, natural (centred) layout; Wt = P·q
//   plane phase q   [N][N]             complex, natural layout
//   counts          [nframe][N][N]     float, detector plane in DFT index order (zero frequency
//                                      at [0,0]); a negative value is a masked pixel
//   psi stack       [nb][M][N][N]      complex; exit waves in natural layout, after pass A with
//                                      rows transformed, after pass B in the detector plane
//                                      transformed back along columns, after pass C the update
//                                      term (Δ, χ) in natural layout again
//   corner          [nframe]           int2 {oy, ox}, the view's top-left object pixel
//   order           [nb]               frame index of each batch slot
//
// Exit waves are kept in the natural layout, not FFT-shifted: the DFT of the natural plane is
// the DFT of the shifted plane times (−1)^(ky+kx), a unit-modulus factor that every modulus
// projection, LLK and gradient scale leaves alone and the inverse transform takes back out.
#pragma once

#include "cplx.hpp"
#include "fft.hpp"

namespace ptyrs {

constexpr int THREADS = 256;

// Elements per thread for a line of n: `pref` where its radix plan exists and the line fits in
// THREADS_MAX threads, else the first divisor of n that has a plan.
constexpr int choose_e(int n, int pref) {
  if (n % pref == 0 && n_stages(n, pref) > 0 && n / pref <= 1024) return pref;
  const int cands[12] = {8, 16, 12, 24, 4, 6, 10, 20, 32, 40, 48, 2};
  for (int i = 0; i < 12; ++i) {
    int e = cands[i];
    if (n % e == 0 && n_stages(n, e) > 0 && n / e <= 256) return e;
  }
  return 0;
}

// Lines per workgroup: the largest divisor of N with at most maxl lines and maxt threads.
constexpr int lines_per_block(int n, int tl, int maxt, int maxl) {
  int best = 1;
  for (int l = 1; l <= n && l <= maxl; ++l)
    if (n % l == 0 && l * tl <= maxt) best = l;
  return best;
}

// A workgroup of L lines of N values, TL = N/E threads per line.
template <int N, int E_, int MAXT, int MAXL>
struct Lines {
  static constexpr int E = E_;
  static_assert(E > 0, "unsupported frame side");
  static constexpr int TL = N / E;                            // threads per line
  static constexpr int L = lines_per_block(N, TL, MAXT, MAXL);
  static constexpr int NT = TL * L;                           // threads per workgroup
  static constexpr int LD = lds_line(N);
  static constexpr int NBLK = N / L;                          // workgroups per plane
  static constexpr int LDS_WORDS = TL > 1 ? L * LD : 1;
};

// Row passes hold one mode per workgroup: 16 values per thread (float), 8 (double).
template <typename T, int N>
using RowShape = Lines<N, choose_e(N, sizeof(T) == 4 ? 16 : 8), THREADS, N>;
// Pass C loops over frames with its probe terms held: 8 values per thread (float), 4 (double).
template <typename T, int N>
using RowShapeC = Lines<N, choose_e(N, sizeof(T) == 4 ? 8 : 4), THREADS, N>;
// The column pass holds every mode of a 128-byte stripe of columns (16 float or 8 double
// columns): 8 values per thread (float), 4 (double), so that four modes stay in registers at
// several waves per SIMD.
template <typename T, int N>
using ColShape = Lines<N, choose_e(N, sizeof(T) == 4 ? 8 : 4), 512, sizeof(T) == 4 ? 16 : 8>;

// ---- block reductions --------------------------------------------------------------------------
template <int NT, typename V, typename Op>
__device__ __forceinline__ V block_reduce(V v, V* red, Op op) {
  // Fixed tree: deterministic for a fixed block size.
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
#pragma unroll
  for (int s = 1; s < NT; s <<= 1) {
    if ((tid % (2 * s)) == 0 && tid + s < NT) red[tid] = op(red[tid], red[tid + s]);
    __syncthreads();
  }
  V r = red[0];
  __syncthreads();
  return r;
}

struct MaxOp {
  template <typename V>
  __device__ V operator()(V a, V b) const { return a > b ? a : b; }
};
struct AddOp {
  template <typename V>
  __device__ V operator()(V a, V b) const { return a + b; }
};

// ---- probe terms ------------------------------------------------------------------------------
// Wt[m][i] = P[m][i]·q[i].
template <typename T>
__global__ void k_probe_terms(const cplx<T>* __restrict__ P, const cplx<T>* __restrict__ q,
                              cplx<T>* __restrict__ Wt, int n, int total) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) st(Wt + i, ld(P + i) * ld(q + i % n));
}

// ---- pass A: gather O·P·q and transform rows ---------------------------------------------------
// Grid (NBLK, nb·M). View (b, m) = (blockIdx.y / M, blockIdx.y % M). With `omax_part`, mode-0
// workgroups also write max |O|² over their rows of the window.
template <typename T, int N>
__global__ __launch_bounds__((RowShape<T, N>::NT)) void k_pass_a(
    const cplx<T>* __restrict__ O, int nxo, const cplx<T>* __restrict__ Wt,
    const int2* __restrict__ corner, const int* __restrict__ order, int M,
    cplx<T>* __restrict__ psi, T* __restrict__ omax_part, const cplx<T>* __restrict__ tw) {
  using S = RowShape<T, N>;
  constexpr int E = S::E, TL = S::TL;
  __shared__ T lre[S::LDS_WORDS], lim[S::LDS_WORDS];
  __shared__ T red[S::NT];
  const int tid = threadIdx.x, l = tid / TL, jl = tid % TL;
  const int y = blockIdx.x * S::L + l;
  const int b = blockIdx.y / M, m = blockIdx.y % M;
  const int2 c = corner[order[b]];
  const cplx<T>* orow = O + (size_t)(c.x + y) * nxo + c.y;
  const cplx<T>* wrow = Wt + ((size_t)m * N + y) * N;
  cplx<T> v[E];
  T mx = 0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int x = jl + TL * e;
    cplx<T> o = ld(orow + x);
    v[e] = ld(wrow + x) * o;
    T o2 = norm2(o);
    mx = o2 > mx ? o2 : mx;
  }
  line_fft<T, N, E, false>(v, jl, lre + l * S::LD, lim + l * S::LD, tw);
  cplx<T>* prow = psi + ((size_t)(b * M + m) * N + y) * N;
#pragma unroll
  for (int e = 0; e < E; ++e) st(prow + jl + TL * e, v[e]);
  if (omax_part != nullptr && m == 0) {
    T r = block_reduce<S::NT>(mx, red, MaxOp());
    if (tid == 0) omax_part[(size_t)b * S::NBLK + blockIdx.x] = r;
  }
}

// ---- pass B: column transforms, the detector-plane operator, inverse column transforms --------
enum PassB { B_RPIE = 0, B_LLK = 1, B_AMPLITUDE = 2 };

// Grid (NBLK, nb). Each workgroup holds L columns of every mode of one frame. B_RPIE scales
// the field by the pty-chi modulus sqrt(iobs + 1e-7)/(sqrt(icalc) + 1e-7) (masked pixels keep
// their value), with the 1/n of the unnormalised round trip folded in; B_LLK writes the four
// LLK sums of the workgroup's pixels (FP64) to llk_part[frame][blockIdx.x][4] and stores
// nothing.
template <typename T, int N, int M, int MODE>
__global__ __launch_bounds__((ColShape<T, N>::NT)) void k_pass_b(
    cplx<T>* __restrict__ psi, const float* __restrict__ counts, const int* __restrict__ order,
    double* __restrict__ llk_part, const cplx<T>* __restrict__ tw) {
  using S = ColShape<T, N>;
  constexpr int E = S::E, TL = S::TL, CG = S::L;
  constexpr int n = N * N;
  __shared__ T lre[S::LDS_WORDS], lim[S::LDS_WORDS];
  const int tid = threadIdx.x, col = tid % CG, jl = tid / CG;
  const int x = blockIdx.x * CG + col;
  const int b = blockIdx.y;
  const int f = order[b];
  T* re = lre + col * S::LD;
  T* im = lim + col * S::LD;
  cplx<T> v[M][E];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const cplx<T>* plane = psi + (size_t)(b * M + m) * n;
#pragma unroll
    for (int e = 0; e < E; ++e) v[m][e] = ld(plane + (jl + TL * e) * N + x);
    line_fft<T, N, E, false>(v[m], jl, re, im, tw);
  }
  const float* cnt = counts + (size_t)f * n;
  const T inv_n = T(1) / T(n);
  double acc[4] = {0, 0, 0, 0};
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const float io = cnt[(jl + TL * e) * N + x];
    T I = 0;
#pragma unroll
    for (int m = 0; m < M; ++m) I += norm2(v[m][e]);
    if constexpr (MODE == B_RPIE) {
      T r = inv_n;
      if (io >= 0.f) r = sqrt(T(io) + T(1e-7)) / (sqrt(I * inv_n) + T(1e-7)) * inv_n;
#pragma unroll
      for (int m = 0; m < M; ++m) v[m][e] = v[m][e] * r;
    } else if constexpr (MODE == B_AMPLITUDE) {
      // PyNX ApplyAmplitude: sqrt(iobs)/max(1e-12, sqrt(icalc)); masked pixels unchanged.
      T r = inv_n;
      if (io >= 0.f) {
        T calc = sqrt(I * inv_n);
        r = sqrt(T(io)) / (calc > T(1e-12) ? calc : T(1e-12)) * inv_n;
      }
#pragma unroll
      for (int m = 0; m < M; ++m) v[m][e] = v[m][e] * r;
    } else {
      if (io >= 0.f) {
        const double ic = double(I) * (1.0 / double(n));
        const double ob = double(io);
        acc[0] += ob > 0.0 ? ic - ob + ob * log(ob / ic) : ic;
        acc[1] += (ob - ic) * (ob - ic) / (ob + 1.0);
        const double d = sqrt(ob) - sqrt(ic);
        acc[2] += 4.0 * d * d;
        acc[3] += ic;
      }
    }
  }
  if constexpr (MODE == B_LLK) {
    __shared__ double red[S::NT];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double s = block_reduce<S::NT>(acc[i], red, AddOp());
      if (tid == 0) llk_part[((size_t)f * S::NBLK + blockIdx.x) * 4 + i] = s;
    }
  } else {
#pragma unroll
    for (int m = 0; m < M; ++m) {
      line_fft<T, N, E, true>(v[m], jl, re, im, tw);
      cplx<T>* plane = psi + (size_t)(b * M + m) * n;
#pragma unroll
      for (int e = 0; e < E; ++e) st(plane + (jl + TL * e) * N + x, v[m][e]);
    }
  }
}

// ---- pass C for rPIE: inverse rows, Δ = ψ' − W·O, probe direction ------------------------------
// Grid (NBLK, M, nchunk). Each workgroup owns rows of mode m and runs over the batch slots of
// its chunk in order: it transforms the rows back, stores Δ over the projected field, and adds
// conj(O·q)·Δ / ((1 − α)|O|² + α·max_window |O|²) into a per-chunk probe direction. The window
// max comes from pass A's `na` partial maxima per slot.
template <typename T, int N>
__global__ __launch_bounds__((RowShapeC<T, N>::NT)) void k_pass_c_rpie(
    cplx<T>* __restrict__ psi, const cplx<T>* __restrict__ O, int nxo,
    const cplx<T>* __restrict__ Wt, const cplx<T>* __restrict__ q,
    const int2* __restrict__ corner, const int* __restrict__ order, int nb, int M,
    const T* __restrict__ omax_part, int na, int chunk, T alpha, int probe_dir,
    cplx<T>* __restrict__ dP_part, const cplx<T>* __restrict__ tw) {
  using S = RowShapeC<T, N>;
  constexpr int E = S::E, TL = S::TL;
  constexpr int n = N * N;
  __shared__ T lre[S::LDS_WORDS], lim[S::LDS_WORDS];
  const int tid = threadIdx.x, l = tid / TL, jl = tid % TL;
  const int y = blockIdx.x * S::L + l;
  const int m = blockIdx.y, ch = blockIdx.z;
  const int b0 = ch * chunk;
  const int b1 = b0 + chunk < nb ? b0 + chunk : nb;
  cplx<T> w[E], qq[E], acc[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int x = jl + TL * e;
    w[e] = ld(Wt + (size_t)m * n + (size_t)y * N + x);
    qq[e] = ld(q + (size_t)y * N + x);
    acc[e] = cplx<T>{0, 0};
  }
  const T inv = T(1) - alpha;
  for (int b = b0; b < b1; ++b) {
    const int2 c = corner[order[b]];
    cplx<T>* prow = psi + ((size_t)(b * M + m) * N + y) * N;
    cplx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = ld(prow + jl + TL * e);
    T omax = 0;
    for (int i = 0; i < na; ++i) {
      T t = omax_part[(size_t)b * na + i];
      omax = t > omax ? t : omax;
    }
    line_fft<T, N, E, true>(v, jl, lre + l * S::LD, lim + l * S::LD, tw);
    const cplx<T>* orow = O + (size_t)(c.x + y) * nxo + c.y;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int x = jl + TL * e;
      const cplx<T> o = ld(orow + x);
      const cplx<T> d = v[e] - w[e] * o;
      st(prow + x, d);
      if (probe_dir) {
        const T den = inv * norm2(o) + alpha * omax;
        acc[e] = acc[e] + cmulc(o * qq[e], d) * (T(1) / den);
      }
    }
  }
  if (probe_dir) {
    cplx<T>* out = dP_part + ((size_t)ch * M + m) * n + (size_t)y * N;
#pragma unroll
    for (int e = 0; e < E; ++e) st(out + jl + TL * e, acc[e]);
  }
}

// P += s·Σ_chunks dP_part, then Wt = P·q.
template <typename T>
__global__ void k_probe_apply(cplx<T>* __restrict__ P, cplx<T>* __restrict__ Wt,
                              const cplx<T>* __restrict__ q, const cplx<T>* __restrict__ dP_part,
                              int nchunk, int n, int total, T s) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  kahan<T> sr, si;
  for (int c = 0; c < nchunk; ++c) {
    cplx<T> d = ld(dP_part + (size_t)c * total + i);
    sr.add(d.x);
    si.add(d.y);
  }
  cplx<T> p = ld(P + i);
  p = p + cplx<T>{sr.s, si.s} * s;
  st(P + i, p);
  st(Wt + i, p * ld(q + i % n));
}

// ---- object tiles --------------------------------------------------------------------------------
constexpr int TILE_X = 64, TILE_Y = 4;

// One workgroup per object tile: (1) the batch slots whose N×N window meets the tile, in slot
// order, compacted with wave ballots into list[tile][·] and count[tile]; (2) with `P`, the max
// over the tile's pixels of the batch overlap Σ_m Σ_views |P_m|² (rPIE's regulariser max).
template <typename T>
__global__ __launch_bounds__(THREADS) void k_bin_tiles(
    const int2* __restrict__ corner, const int* __restrict__ order, int nb, int N, int ntx,
    int cap, int* __restrict__ list, int* __restrict__ count, const cplx<T>* __restrict__ P,
    int M, int nyo, int nxo, T* __restrict__ tilemax) {
  __shared__ int wave_n[THREADS / 64];
  __shared__ int base_s;
  __shared__ T red[THREADS];
  const int t = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ty0 = (t / ntx) * TILE_Y, tx0 = (t % ntx) * TILE_X;
  int* out = list + (size_t)t * cap;
  if (tid == 0) base_s = 0;
  __syncthreads();
  for (int s = 0; s < nb; s += THREADS) {
    const int b = s + tid;
    bool hit = false;
    if (b < nb) {
      const int2 c = corner[order[b]];
      hit = c.x < ty0 + TILE_Y && c.x + N > ty0 && c.y < tx0 + TILE_X && c.y + N > tx0;
    }
    const unsigned long long mask = __ballot(hit);
    const int before = __popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) wave_n[wave] = __popcll(mask);
    __syncthreads();
    int off = base_s;
    for (int w = 0; w < wave; ++w) off += wave_n[w];
    if (hit) out[off + before] = b;
    __syncthreads();
    if (tid == 0) {
      int tot = 0;
      for (int w = 0; w < THREADS / 64; ++w) tot += wave_n[w];
      base_s += tot;
    }
    __syncthreads();
  }
  const int nl = base_s;
  if (tid == 0) count[t] = nl;
  if (P == nullptr) return;
  __shared__ int2 cs[THREADS];
  const int Y = ty0 + tid / TILE_X, X = tx0 + tid % TILE_X;
  const bool inside = Y < nyo && X < nxo;
  const int n = N * N;
  T sum = 0;
  for (int s = 0; s < nl; s += THREADS) {
    const int k = nl - s < THREADS ? nl - s : THREADS;
    if (tid < k) cs[tid] = corner[order[out[s + tid]]];
    __syncthreads();
    if (inside) {
      for (int i = 0; i < k; ++i) {
        const int y = Y - cs[i].x, x = X - cs[i].y;
        if (y >= 0 && y < N && x >= 0 && x < N) {
          for (int m = 0; m < M; ++m) sum += norm2(ld(P + (size_t)m * n + y * N + x));
        }
      }
    }
    __syncthreads();
  }
  T r = block_reduce<THREADS>(sum, red, MaxOp());
  if (tid == 0) tilemax[t] = r;
}

template <typename T>
__global__ __launch_bounds__(THREADS) void k_max_reduce(const T* __restrict__ v, int count,
                                                        T* __restrict__ out) {
  __shared__ T red[THREADS];
  T m = 0;
  for (int i = threadIdx.x; i < count; i += THREADS) m = v[i] > m ? v[i] : m;
  T r = block_reduce<THREADS>(m, red, MaxOp());
  if (threadIdx.x == 0) out[0] = r;
}

// rPIE object step, pass 2: every object pixel owned by one thread adds, in slot order,
// O += step · Σ_m Σ_views conj(W_m)·Δ_m / ((1 − α)·D_m + α·pmax) with D_m the batch overlap
// of |P_m|² at the pixel, the sums Kahan-compensated. No atomics; the result does not depend on
// the launch shape.
template <typename T>
__global__ __launch_bounds__(THREADS) void k_obj_update_rpie(
    cplx<T>* __restrict__ O, const cplx<T>* __restrict__ Wt, const cplx<T>* __restrict__ P,
    const cplx<T>* __restrict__ delta, int N, int M, const int2* __restrict__ corner,
    const int* __restrict__ order, const int* __restrict__ list, const int* __restrict__ count,
    int cap, int nyo, int nxo, int ntx, const T* __restrict__ pmax, T alpha, T step) {
  constexpr int MM = 4;
  __shared__ int3 cs[THREADS];
  const int t = blockIdx.x, tid = threadIdx.x;
  const int Y = (t / ntx) * TILE_Y + tid / TILE_X;
  const int X = (t % ntx) * TILE_X + tid % TILE_X;
  const int nl = count[t];
  if (nl == 0) return;
  const bool inside = Y < nyo && X < nxo;
  const int n = N * N;
  kahan<T> nr[MM], ni[MM], dd[MM];
  bool covered = false;
  const int* lst = list + (size_t)t * cap;
  for (int s = 0; s < nl; s += THREADS) {
   const int k = nl - s < THREADS ? nl - s : THREADS;
   if (tid < k) {
     const int b = lst[s + tid];
     const int2 c = corner[order[b]];
     cs[tid] = int3{b, c.x, c.y};
   }
   __syncthreads();
   for (int i = 0; i < k && inside; ++i) {
    const int b = cs[i].x;
    const int y = Y - cs[i].y, x = X - cs[i].z;
    if (y >= 0 && y < N && x >= 0 && x < N) {
      const int j = y * N + x;
      const cplx<T>* dv = delta + (size_t)b * M * n + j;
      cplx<T> dl[MM];
#pragma unroll
      for (int m = 0; m < MM; ++m)
        if (m < M) dl[m] = ld(dv + (size_t)m * n);
#pragma unroll
      for (int m = 0; m < MM; ++m) {
        if (m < M) {
          const cplx<T> g = cmulc(ld(Wt + (size_t)m * n + j), dl[m]);
          nr[m].add(g.x);
          ni[m].add(g.y);
          dd[m].add(norm2(ld(P + (size_t)m * n + j)));
        }
      }
      covered = true;
    }
   }
   __syncthreads();
  }
  if (!covered) return;
  const T reg = alpha * pmax[0];
  cplx<T> upd = {0, 0};
#pragma unroll
  for (int m = 0; m < MM; ++m)
    if (m < M) upd = upd + cplx<T>{nr[m].s, ni[m].s} * (T(1) / ((T(1) - alpha) * dd[m].s + reg));
  cplx<T>* o = O + (size_t)Y * nxo + X;
  st(o, ld(o) + upd * step);
}

// ---- LLK sums -------------------------------------------------------------------------------------
// Sums llk_part[count][4] in a fixed order into out[4].
static __global__ void k_llk_reduce(const double* __restrict__ part, int count, double* __restrict__ out) {
  __shared__ double red[THREADS];
  for (int k = 0; k < 4; ++k) {
    double s = 0;
    for (int i = threadIdx.x; i < count; i += THREADS) s += part[(size_t)i * 4 + k];
    double r = block_reduce<THREADS>(s, red, AddOp());
    if (threadIdx.x == 0) out[k] = r;
  }
}

}  // namespace ptyrs
