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

#include <type_traits>

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

// A workgroup of L lines of N values, TL = N/E threads per line, elements of WB bytes.
template <int N, int E_, int MAXT, int MAXL, int WB>
struct Lines {
  static constexpr int E = E_;
  static_assert(E > 0, "unsupported frame side");
  static constexpr int TL = N / E;                            // threads per line
  static constexpr int L = lines_per_block(N, TL, MAXT, MAXL);
  static constexpr int NT = TL * L;                           // threads per workgroup
  static constexpr int NBLK = N / L;                          // workgroups per plane
  // LDS placement of the lines' exchanges for lanes along a line (row passes) or across the
  // lines (column passes): element placement and the offset of line `line`'s element 0.
  static constexpr LdsLay ROWL = lds_layout(N, E, L, true, WB);
  static constexpr LdsLay COLL = lds_layout(N, E, L, false, WB);
  using AtRow = LdsAt<ROWL.shift, ROWL.mul, ROWL.inter ? L : 1>;
  using AtCol = LdsAt<COLL.shift, COLL.mul, COLL.inter ? L : 1>;
  static constexpr int base_row(int line) { return ROWL.inter ? line : line * ROWL.ld; }
  static constexpr int base_col(int line) { return COLL.inter ? line : line * COLL.ld; }
  static constexpr int LDS_WORDS =
      TL > 1 ? (ROWL.words > COLL.words ? ROWL.words : COLL.words) : 1;
  // The last element of the last line lies inside the arrays under both placements.
  static_assert(TL == 1 || (AtRow::at(N - 1) + base_row(L - 1) < LDS_WORDS &&
                            AtCol::at(N - 1) + base_col(L - 1) < LDS_WORDS),
                "LDS layout exceeds its arrays");
};

// Row passes hold one mode per workgroup: 16 values per thread (float), 8 (double).
template <typename T, int N>
using RowShape = Lines<N, choose_e(N, sizeof(T) == 4 ? 16 : 8), THREADS, N, sizeof(T)>;
// Pass C loops over frames with its probe terms held: 8 values per thread (float), 4 (double).
template <typename T, int N>
using RowShapeC = Lines<N, choose_e(N, sizeof(T) == 4 ? 8 : 4), THREADS, N, sizeof(T)>;
// The column pass holds every mode of a 64-byte stripe of columns (8 float or 4 double
// columns): 8 values per thread (float), 4 (double), so that four modes stay in registers with
// two workgroups per CU; neighbouring stripes share their 128-byte lines through L2
// (block_xy).
template <typename T, int N>
using ColShape = Lines<N, choose_e(N, sizeof(T) == 4 ? 8 : 4), 512, sizeof(T) == 4 ? 8 : 4,
                       sizeof(T)>;

// ---- workgroup placement --------------------------------------------------------------------------
// Workgroups are dealt round-robin to the 8 XCDs (each with its own L2): dispatch slot L runs on
// XCD L % 8. The column passes renumber a 2D grid whose size divides by 8 so that consecutive
// logical blocks (neighbouring column stripes of a plane, which share 128-byte lines) run on one
// XCD at about the same time and share its L2.
__device__ __forceinline__ int2 block_xy() {
  const int total = gridDim.x * gridDim.y;
  const int L = blockIdx.y * gridDim.x + blockIdx.x;
  if (total % 8 == 0) {
    const int id = (L % 8) * (total / 8) + L / 8;
    return int2{id % int(gridDim.x), id / int(gridDim.x)};
  }
  return int2{int(blockIdx.x), int(blockIdx.y)};
}

// The object tile of a tile kernel (1D grid of tiles, row-major), in dispatch order (tiles
// dealt to the XCDs in contiguous bands measured slower).
__device__ __forceinline__ int tile_id() { return int(blockIdx.x); }

// A thread-group index tid / GROUP: when groups are whole waves it is the same in every lane
// of a wave, and is read from the first lane so the compiler keeps it (and the plane addresses
// and branches that follow from it) in scalar registers.
template <int GROUP>
__device__ __forceinline__ int wave_uniform(int v) {
  if constexpr (GROUP % 64 == 0) return __builtin_amdgcn_readfirstlane(v);
  else return v;
}

// ---- block reductions --------------------------------------------------------------------------
// Deterministic for a fixed block size: whole waves reduce by a butterfly of lane exchanges
// and then the waves' results are combined in wave order (one barrier); other block sizes
// reduce by a fixed tree in LDS. `red` holds NT values.
template <int NT, typename V, typename Op>
__device__ __forceinline__ V block_reduce(V v, V* red, Op op) {
  const int tid = threadIdx.x;
  if constexpr (NT % 64 == 0) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off, 64));
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    V r = red[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) r = op(r, red[w]);
    __syncthreads();
    return r;
  } else {
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
}

struct MaxOp {
  template <typename V>
  __device__ V operator()(V a, V b) const { return a > b ? a : b; }
};
struct AddOp {
  template <typename V>
  __device__ V operator()(V a, V b) const { return a + b; }
};

// The DFT twiddle table exp(−2πi·t/N), t < N, copied to LDS once per workgroup: every butterfly
// of every line then reads it from LDS instead of global memory.
template <typename T, int N, int NT>
__device__ __forceinline__ void stage_twiddles(cplx<T>* stw, const cplx<T>* __restrict__ tw) {
  for (int i = threadIdx.x; i < N; i += NT) stw[i] = ld(tw + i);
  __syncthreads();
}

// ---- probe terms ------------------------------------------------------------------------------
// Wt[m][i] = P[m][i]·q[i].
template <typename T>
__global__ void k_probe_terms(const cplx<T>* __restrict__ P, const cplx<T>* __restrict__ q,
                              cplx<T>* __restrict__ Wt, int n, int total) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) st(Wt + i, ld(P + i) * ld(q + i % n));
}

// p[0 .. count) = 0: a fill kernel rather than a memset node, so that it can sit in a captured
// epoch.
template <typename U>
__global__ void k_zero(U* __restrict__ p, size_t count) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = U{};
}

// W2[i] = |Wt[i]|² over the M planes of the probe terms (the per-mode overlap's summands).
template <typename T>
__global__ void k_norm2_planes(const cplx<T>* __restrict__ Wt, T* __restrict__ W2, int total) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) W2[i] = norm2(ld(Wt + i));
}

// ---- pass A: gather O·P·q and transform rows ---------------------------------------------------
// Grid (NBLK, nb·M). View (b, m) = (blockIdx.y / M, blockIdx.y % M). With `omax_part`, mode-0
// workgroups also write max |O|² over their rows of the window.
// The exit wave pass A transforms: PO itself, or DM's and RAAR's mix of PO with the waves
// they carry across cycles (ptyrs/src/engine/ptycho.rs:643-731).
enum Mix { MIX_NONE = 0, MIX_DM = 1, MIX_RAAR = 2, MIX_RAAR0 = 3 };

// Exit waves carried across cycles, value i of a [nframe][M][N][N] array: complex T, or with
// H (GpuOptions::psi_f16) one pair of IEEE halves per value, the real part in the low half,
// holding the value times `scale`, a power of two whose inverse `unscale` is exact, so the
// format adds f16 rounding and nothing else. A component beyond the f16 range is stored as
// the largest finite half of its sign; a NaN stays one. ptyrs-kernels/src/pixel.rs:360-442
template <typename T, bool H>
struct Carried {
  static constexpr bool half = false;
  cplx<T>* p;
  __device__ cplx<T> get(size_t i) const { return ld(p + i); }
  __device__ void put(size_t i, cplx<T> v) const { st(p + i, v); }
};
template <typename T>
struct Carried<T, true> {
  static constexpr bool half = true;
  uint32_t* p;
  float scale, unscale;
  __device__ cplx<T> get(size_t i) const {
    const uint32_t w = p[i];
    const _Float16 re = __builtin_bit_cast(_Float16, uint16_t(w & 0xffffu));
    const _Float16 im = __builtin_bit_cast(_Float16, uint16_t(w >> 16));
    return cplx<T>{T(float(re) * unscale), T(float(im) * unscale)};
  }
  __device__ void put(size_t i, cplx<T> v) const {
    // Comparisons keep a NaN (fminf/fmaxf would return the bound).
    auto sat = [](float x) { return x > 65504.f ? 65504.f : (x < -65504.f ? -65504.f : x); };
    const _Float16 re = _Float16(sat(float(v.x) * scale));
    const _Float16 im = _Float16(sat(float(v.y) * scale));
    p[i] = uint32_t(__builtin_bit_cast(uint16_t, re)) |
           (uint32_t(__builtin_bit_cast(uint16_t, im)) << 16);
  }
};

template <typename T, bool H = false>
struct MixArgs {
  int kind;
  Carried<T, H> carried;  // DM's and RAAR's ψ
  Carried<T, H> iterate;  // RAAR's iterate, rewritten with the mix
  T c0, c1;               // DM: (2 − a, 1 − a); RAAR: (1 − β/2, β/2)
};

template <typename T, int N, int MIX, bool H = false>
__global__ __launch_bounds__((RowShape<T, N>::NT)) void k_pass_a(
    const cplx<T>* __restrict__ O, int nxo, const cplx<T>* __restrict__ Wt,
    const int2* __restrict__ corner, const int* __restrict__ order, int M,
    cplx<T>* __restrict__ psi, T* __restrict__ omax_part, const cplx<T>* __restrict__ tw,
    MixArgs<T, H> mix, int w_view) {
  using S = RowShape<T, N>;
  constexpr int E = S::E, TL = S::TL;
  __shared__ T lre[S::LDS_WORDS], lim[S::LDS_WORDS];
  __shared__ T red[S::NT];
  __shared__ cplx<T> stw[N];
  const int tid = threadIdx.x, l = tid / TL, jl = tid % TL;
  const int y = blockIdx.x * S::L + l;
  const int b = blockIdx.y / M, m = blockIdx.y % M;
  stage_twiddles<T, N, S::NT>(stw, tw);
  const int f = order[b];
  const int2 c = corner[f];
  const cplx<T>* orow = O + (size_t)(c.x + y) * nxo + c.y;
  const cplx<T>* wrow = Wt + ((size_t)(w_view ? int(blockIdx.y) : m) * N + y) * N;
  const size_t crow = (((size_t)f * M + m) * N + y) * N;
  cplx<T> v[E];
  T mx = 0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int x = jl + TL * e;
    cplx<T> o = ld(orow + x);
    v[e] = ld(wrow + x) * o;
    if constexpr (MIX == MIX_DM) {
      v[e] = v[e] * mix.c0 - mix.carried.get(crow + x) * mix.c1;
    } else if constexpr (MIX == MIX_RAAR) {
      const cplx<T> it = mix.iterate.get(crow + x);
      v[e] = mix.carried.get(crow + x) * mix.c0 + (v[e] * T(2) - it) * mix.c1;
      mix.iterate.put(crow + x, v[e]);
    } else if constexpr (MIX == MIX_RAAR0) {
      mix.iterate.put(crow + x, v[e]);  // RAAR's first cycle: ψ = PO, iterate = ψ
    }
    T o2 = norm2(o);
    mx = o2 > mx ? o2 : mx;
  }
  line_fft<T, N, E, false, typename S::AtRow>(v, jl, lre + S::base_row(l), lim + S::base_row(l),
                                              stw);
  cplx<T>* prow = psi + ((size_t)(b * M + m) * N + y) * N;
#pragma unroll
  for (int e = 0; e < E; ++e) st(prow + jl + TL * e, v[e]);
  if (omax_part != nullptr && m == 0) {
    T r = block_reduce<S::NT>(mx, red, MaxOp());
    if (tid == 0) omax_part[(size_t)b * S::NBLK + blockIdx.x] = r;
  }
}

// ---- pass B: column transforms, the detector-plane operator, inverse column transforms --------
enum PassB { B_RPIE = 0, B_LLK = 1, B_AMPLITUDE = 2, B_GAUSS = 3, B_ICALC = 4, B_POISSON = 5,
             B_GRAD = 6, B_GRAD_AMP = 7, B_BH = 8 };

// Per-slot values some pass-B modes read or write.
template <typename T>
struct BArgs {
  const T* a_far;   // B_POISSON: LSQML far-field step per batch slot
  const T* prev;    // B_ICALC: batch mean of the slots' last a_far
  double* part;     // B_ICALC: per-workgroup (num, den) of the Poisson xi solver
  T* icalc;         // B_ICALC: model intensity per slot and pixel, DFT index order
  cplx<T>* keep = nullptr;  // B_BH: the detector-plane field U·ψ per view ([slot·M + m][n])
};

// Pass B's lines: per mode, a stripe of CG columns (or rows) of N values, E per thread (16
// float / 8 double: one LDS exchange fewer than the column passes' 8 / 4). Every mode of the
// stripe has its own group of threads and its own exchange region.
template <typename T>
constexpr int pass_b_lines(int n) {
  return (sizeof(T) == 4 ? 8 : 4) / (n >= 512 ? 2 : 1);
}
template <typename T, int N>
using BShape = Lines<N, choose_e(N, sizeof(T) == 4 ? 16 : 8), 1024, pass_b_lines<T>(N), sizeof(T)>;
template <typename T, int N, int M>
constexpr int pass_b_threads() {
  return M * BShape<T, N>::NT;
}

// Grid (NBLK, nb) of BShape. Mode group m of each workgroup transforms mode m of the stripe;
// the modes' |ψ_m|² meet in LDS, where the group owning a pixel sums them in mode order and
// leaves the pixel's detector-plane factor for every group. B_RPIE
// scales the field by the pty-chi modulus sqrt(iobs + 1e-7)/(sqrt(icalc) + 1e-7) (masked pixels
// keep their value), with the 1/n of the unnormalised round trip folded in; B_LLK writes the
// four LLK sums of the workgroup's pixels (FP64) to llk_part[frame][blockIdx.x][4] and stores
// nothing.
// B_BH keeps U·ψ (the field times √norm) in ba.keep and scales the field by 2(1 − √iobs/(A +
// 1e-8))·norm, A = √(icalc), 0 where masked: Uᴴ of it is BH's 2·∂F/∂(U ψ)* taken back
// (ptyrs/src/engine/bh.rs:233-260).
// B_GAUSS and B_POISSON scale the field by (1 − a_far·c)/n with LSQML's residual c (Gaussian
// 2(1 − √iobs/(√icalc + 1e-9)) with a_far = 1/2, Poisson 1 − iobs/(icalc + 1e-6)), 0 where
// masked (ptyrs/src/engine/lsqml.rs:97-170). B_ICALC stores icalc and the first pass of the
// Poisson a_far solver's sums and stores no field.
// ROW = false: lines are columns (the far field's column pass); ROW = true: lines are rows (the
// angular spectrum's second row pass). `norm` is 1/n (one unnormalised DFT round trip) or 1/n²
// (two).
template <typename T, int N, int M, int MODE, bool ROW>
__global__ __launch_bounds__((pass_b_threads<T, N, M>())) void k_pass_b(
    cplx<T>* __restrict__ psi, const float* __restrict__ counts, const int* __restrict__ order,
    double* __restrict__ llk_part, const cplx<T>* __restrict__ tw, BArgs<T> ba, T norm) {
  using S = BShape<T, N>;
  constexpr int E = S::E, TL = S::TL, CG = S::L;
  constexpr int NT = pass_b_threads<T, N, M>();
  // Each group's region holds its exchanges and, after the forward transform, its intensities.
  constexpr int W = S::LDS_WORDS > CG * N ? S::LDS_WORDS : CG * N;
  constexpr int n = N * N;
  __shared__ T lre[M * W], lim[M * W];
  __shared__ cplx<T> stw[N];
  // Columns: lanes run along the stripe of lines; rows: along each line.
  const int tid = threadIdx.x;
  const int m = wave_uniform<S::NT>(tid / S::NT), tg = tid % S::NT;
  const int line = ROW ? tg / TL : tg % CG, jl = ROW ? tg % TL : tg / CG;
  using At = std::conditional_t<ROW, typename S::AtRow, typename S::AtCol>;
  const int2 bxy = block_xy();
  const int x = bxy.x * CG + line;
  const int b = bxy.y;
  const int f = order[b];
  const int lb = ROW ? S::base_row(line) : S::base_col(line);
  T* re = lre + m * W + lb;
  T* im = lim + m * W + lb;
  auto at = [&](int e) { return ROW ? x * N + jl + TL * e : (jl + TL * e) * N + x; };
  // Element e of the lane is at index i0 + e·step of the plane: one offset VGPR (buffer
  // accesses), the step in SGPRs.
  const uint32_t i0 = uint32_t(ROW ? x * N + jl : jl * N + x);
  constexpr uint32_t step = ROW ? TL : TL * N;
  constexpr uint32_t cb = sizeof(cplx<T>);
  // Pixel e of the lane is owned by group e mod M: that group reads its count, forms its
  // detector-plane operator once and shares it through LDS.
  const Buf cnt = buf(counts + (size_t)f * n, n * 4u);
  float cv[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    if (e % M != m) continue;
    cv[e] = bld1(cnt, i0 * 4u, e * step * 4u);
  }
  const Buf plane = buf(psi + (size_t)(b * M + m) * n, n * cb);
  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = bld<T>(plane, i0 * cb, e * step * cb);
  stage_twiddles<T, N, NT>(stw, tw);
  // The forward and inverse transforms share the lane's twiddles, held in registers for float
  // (for double they would cost a wave per SIMD).
  constexpr bool TC = sizeof(T) == 4;
  cplx<T> twc[TC ? tw_cache_size(N, E) : 1];
  if constexpr (TC) load_tw_cache<T, N, E>(twc, jl, stw);
  line_fft<T, N, E, false, At, TC>(v, jl, re, im, stw, twc);
  if constexpr (MODE == B_BH) {
    const T c = sqrt(norm);
    cplx<T>* kp = ba.keep + (size_t)(b * M + m) * n;
#pragma unroll
    for (int e = 0; e < E; ++e) st(kp + at(e), v[e] * c);
  }
  // The transform's last exchange has been read by every thread: the regions take the
  // intensities, pixel p = jl + TL·e of line `line` of group k at lre[k·W + p·CG + line], and
  // the owners' operators at lim[p·CG + line].
  auto key = [&](int e) { return (jl + TL * e) * CG + line; };
#pragma unroll
  for (int e = 0; e < E; ++e) lre[m * W + key(e)] = norm2(v[e]);
  __syncthreads();
  const T inv_n = norm;
  double acc[4] = {0, 0, 0, 0};
  T afar = 0;
  double prev = 0;
  if constexpr (MODE == B_POISSON) afar = ba.a_far[b];
  if constexpr (MODE == B_ICALC) prev = double(ba.prev[0]);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    if (e % M != m) continue;
    const float io = cv[e];
    T I = 0;
#pragma unroll
    for (int k = 0; k < M; ++k) I += lre[k * W + key(e)];
    T r = 0;
    if constexpr (MODE == B_RPIE) {
      r = inv_n;
      if (io >= 0.f) r = sqrt(T(io) + T(1e-7)) / (sqrt(I * inv_n) + T(1e-7)) * inv_n;
    } else if constexpr (MODE == B_AMPLITUDE) {
      // PyNX ApplyAmplitude: sqrt(iobs)/max(1e-12, sqrt(icalc)); masked pixels unchanged.
      r = inv_n;
      if (io >= 0.f) {
        T calc = sqrt(I * inv_n);
        r = sqrt(T(io)) / (calc > T(1e-12) ? calc : T(1e-12)) * inv_n;
      }
    } else if constexpr (MODE == B_GRAD) {
      // PyNX GradPoissonFourier: 2(1 − iobs/max(icalc, 1e-12)), 0 where masked
      // (ptyrs/src/engine/ops.rs:156-195).
      if (io >= 0.f) {
        const T ic = I * inv_n;
        r = T(2) * (T(1) - T(io) / (ic > T(1e-12) ? ic : T(1e-12))) * inv_n;
      }
    } else if constexpr (MODE == B_GRAD_AMP) {
      // The amplitude cost (√icalc − √iobs)²: 2(1 − √(iobs/max(icalc, 1e-12))), 0 where masked
      // (ptyrs/src/engine/features.rs:322-330).
      if (io >= 0.f) {
        const T ic = I * inv_n;
        r = T(2) * (T(1) - sqrt(T(io) / (ic > T(1e-12) ? ic : T(1e-12)))) * inv_n;
      }
    } else if constexpr (MODE == B_BH) {
      if (io >= 0.f) {
        const T A = sqrt(I * inv_n) + T(1e-8);
        r = T(2) * (T(1) - sqrt(T(io)) / A) * inv_n;
      }
    } else if constexpr (MODE == B_GAUSS || MODE == B_POISSON) {
      const T ic = I * inv_n;
      T c = 0;
      if (io >= 0.f) {
        if constexpr (MODE == B_GAUSS) c = T(2) * (T(1) - sqrt(T(io)) / (sqrt(ic) + T(1e-9)));
        else c = T(1) - T(io) / (ic + T(1e-6));
      }
      r = (T(1) - (MODE == B_GAUSS ? T(0.5) : afar) * c) * inv_n;
    } else if constexpr (MODE == B_ICALC) {
      const T ic = I * inv_n;
      ba.icalc[(size_t)b * n + at(e)] = ic;
      if (io >= 0.f) {
        const double yp = double(ic), yt = double(io);
        const double xi = 1.0 - yt / (yp + 1e-5);
        acc[0] += xi * (yp - yt / (1.0 - prev * xi));
        acc[1] += xi * xi * yp;
      }
    } else {
      if (io >= 0.f) {
        const double ic = double(I) * double(norm);
        const double ob = double(io);
        acc[0] += ob > 0.0 ? ic - ob + ob * log(ob / ic) : ic;
        acc[1] += (ob - ic) * (ob - ic) / (ob + 1.0);
        const double d = sqrt(ob) - sqrt(ic);
        acc[2] += 4.0 * d * d;
        acc[3] += ic;
      }
    }
    if constexpr (MODE != B_LLK && MODE != B_ICALC) lim[key(e)] = r;
  }
  if constexpr (MODE == B_LLK) {
    __shared__ double red[NT];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double s = block_reduce<NT>(acc[i], red, AddOp());
      if (tid == 0) llk_part[((size_t)f * S::NBLK + bxy.x) * 4 + i] = s;
    }
  } else if constexpr (MODE == B_ICALC) {
    __shared__ double red[NT];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      double s = block_reduce<NT>(acc[i], red, AddOp());
      if (tid == 0) ba.part[((size_t)b * S::NBLK + bxy.x) * 2 + i] = s;
    }
  } else {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = v[e] * lim[key(e)];
    // Every group has read the operators before the inverse transform's exchange.
    __syncthreads();
    line_fft<T, N, E, true, At, TC>(v, jl, re, im, stw, twc);
#pragma unroll
    for (int e = 0; e < E; ++e) bst<T>(plane, i0 * cb, e * step * cb, v[e]);
  }
}

// ---- the angular spectrum's column passes -------------------------------------------------------------
// Per plane and stripe of columns: DFT along columns, times T, DFT again (INV: the inverse DFTs
// and conj T), unnormalised. Grid (NBLK, nplanes, nfield), planes of field z at psi + z·field;
// linear, one stripe of one plane per workgroup. ptyrs/src/engine/propagate.rs:380-428
template <typename T, int N, bool INV>
__global__ __launch_bounds__((ColShape<T, N>::NT)) void k_pass_t(cplx<T>* __restrict__ psi,
                                                                  size_t field,
                                                                  const cplx<T>* __restrict__ tf,
                                                                  const cplx<T>* __restrict__ tw) {
  using S = ColShape<T, N>;
  constexpr int E = S::E, TL = S::TL, CG = S::L;
  __shared__ T lre[S::LDS_WORDS], lim[S::LDS_WORDS];
  __shared__ cplx<T> stw[N];
  const int tid = threadIdx.x, col = tid % CG, jl = tid / CG;
  const int2 bxy = block_xy();
  const int x = bxy.x * CG + col;
  cplx<T>* plane = psi + blockIdx.z * field + (size_t)bxy.y * N * N;
  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = ld(plane + (jl + TL * e) * N + x);
  stage_twiddles<T, N, S::NT>(stw, tw);
  T* re = lre + S::base_col(col);
  T* im = lim + S::base_col(col);
  line_fft<T, N, E, INV, typename S::AtCol>(v, jl, re, im, stw);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const cplx<T> t = ld(tf + (jl + TL * e) * N + x);
    v[e] = INV ? v[e] * conj(t) : v[e] * t;
  }
  line_fft<T, N, E, INV, typename S::AtCol>(v, jl, re, im, stw);
#pragma unroll
  for (int e = 0; e < E; ++e) st(plane + (jl + TL * e) * N + x, v[e]);
}

// ---- pass C: inverse rows, Δ = ψ' − W·O, probe direction ----------------------------------------
// LSQML's pass C. Grid (NBLK, M, nchunk). Each workgroup owns rows of mode m and runs over the
// batch slots of its chunk in order: it transforms the rows back, stores χ = ψ' − W·O over the
// projected field (sub_wo = 0: χ = ψ', BH's back-projected residual), and adds conj(O·q)·χ into
// a per-chunk probe direction. With w_view every slot's mode has its own wave W (plane slot·M + m
// of Wt, a frame's own probe).
// ptyrs/src/engine/lsqml.rs:97-228
template <typename T, int N>
__global__ __launch_bounds__((RowShapeC<T, N>::NT)) void k_pass_c_chi(
    cplx<T>* __restrict__ psi, const cplx<T>* __restrict__ O, int nxo,
    const cplx<T>* __restrict__ Wt, const cplx<T>* __restrict__ q,
    const int2* __restrict__ corner, const int* __restrict__ order, int nb, int M, int chunk,
    int probe_dir, cplx<T>* __restrict__ dP_part, const cplx<T>* __restrict__ tw,
    int sub_wo = 1, int w_view = 0) {
  using S = RowShapeC<T, N>;
  constexpr int E = S::E, TL = S::TL;
  constexpr int n = N * N;
  __shared__ T lre[S::LDS_WORDS], lim[S::LDS_WORDS];
  __shared__ cplx<T> stw[N];
  const int tid = threadIdx.x, l = tid / TL, jl = tid % TL;
  const int y = blockIdx.x * S::L + l;
  const int m = blockIdx.y, ch = blockIdx.z;
  const int b0 = ch * chunk;
  const int b1 = b0 + chunk < nb ? b0 + chunk : nb;
  stage_twiddles<T, N, S::NT>(stw, tw);
  cplx<T> w[E], qq[E], acc[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int x = jl + TL * e;
    w[e] = ld(Wt + (size_t)m * n + (size_t)y * N + x);
    qq[e] = ld(q + (size_t)y * N + x);
    acc[e] = cplx<T>{0, 0};
  }
  // The next slot's rows are loaded before this slot's transform.
  cplx<T> nxt[E], onx[E];
  if (b0 < b1) {
    const cplx<T>* p0 = psi + ((size_t)(b0 * M + m) * N + y) * N;
    const int2 c0 = corner[order[b0]];
    const cplx<T>* o0 = O + (size_t)(c0.x + y) * nxo + c0.y;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      nxt[e] = ld(p0 + jl + TL * e);
      onx[e] = ld(o0 + jl + TL * e);
    }
  }
  for (int b = b0; b < b1; ++b) {
    cplx<T>* prow = psi + ((size_t)(b * M + m) * N + y) * N;
    cplx<T> v[E], ow[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      v[e] = nxt[e];
      ow[e] = onx[e];
    }
    if (b + 1 < b1) {
      const cplx<T>* pn = psi + ((size_t)((b + 1) * M + m) * N + y) * N;
      const int2 cn = corner[order[b + 1]];
      const cplx<T>* on = O + (size_t)(cn.x + y) * nxo + cn.y;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        nxt[e] = ld(pn + jl + TL * e);
        onx[e] = ld(on + jl + TL * e);
      }
    }
    line_fft<T, N, E, true, typename S::AtRow>(v, jl, lre + S::base_row(l),
                                               lim + S::base_row(l), stw);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int x = jl + TL * e;
      const cplx<T> o = ow[e];
      const cplx<T> we = w_view ? ld(Wt + ((size_t)b * M + m) * n + (size_t)y * N + x) : w[e];
      const cplx<T> d = sub_wo ? v[e] - we * o : v[e];
      st(prow + x, d);
      if (probe_dir) acc[e] = acc[e] + cmulc(o * qq[e], d);
    }
  }
  if (probe_dir) {
    cplx<T>* out = dP_part + ((size_t)ch * M + m) * n + (size_t)y * N;
#pragma unroll
    for (int e = 0; e < E; ++e) st(out + jl + TL * e, acc[e]);
  }
}

// Rows per mode group of rPIE's mode-summing pass C: the largest divisor of N with the M groups
// within pass C's lines per workgroup.
constexpr int rows_per_group(int n, int l, int m) {
  int best = 1;
  for (int d = 1; d * m <= l; ++d)
    if (n % d == 0) best = d;
  return best;
}

// rPIE's pass C with the modes summed. Grid (N/G, nchunk) of M·G·TL threads: mode group m
// (G rows of TL threads each, pass C's line shape) owns rows of mode m, and the groups run over
// the batch slots of their chunk together. Per slot each group transforms its rows back, forms
// Δ_m = ψ'_m − W_m·O, and puts conj(W_m)·Δ_m/((1 − α_o)·D_m + α_o·pmax) in LDS — D_m the batch
// overlap of |W_m|² at the window pixel, so the object update only sums over views. One slot
// later (after the next transform's barriers) the groups add those terms over the modes, in
// mode order, each group for every M-th of its elements, into the slot's mode-0 plane; the
// terms alternate between two LDS buffers. With `probe_dir` each group adds conj(O·q)·Δ_m
// divided by (1 − α_p)|O|² + α_p·max_window |O|² into its mode's per-chunk probe direction.
// Under position offsets (`w_view`, one slot per chunk) W is the slot's own moved wave, plane
// slot·M + m of Wt, and with `chi0` mode 0's Δ is also stored ([slot·M][n], the position sums'
// χ). ptyrs/src/engine/ptycho.rs:900-1086
// Threads of k_pass_c_rpie_sum: M groups of G rows of pass C's line width.
template <typename T, int N, int M>
constexpr int pass_c_sum_threads() {
  return M * rows_per_group(N, RowShapeC<T, N>::L, M) * RowShapeC<T, N>::TL;
}

template <typename T, int N, int M>
__global__ __launch_bounds__((pass_c_sum_threads<T, N, M>())) void k_pass_c_rpie_sum(
    cplx<T>* __restrict__ psi, const cplx<T>* __restrict__ O, int nxo,
    const cplx<T>* __restrict__ Wt, const cplx<T>* __restrict__ q,
    const int2* __restrict__ corner, const int* __restrict__ order, int nb,
    const T* __restrict__ omax_part, int na, int chunk, T probe_alpha, int probe_dir,
    cplx<T>* __restrict__ dP_part, const cplx<T>* __restrict__ tw, int upd_obj,
    const T* __restrict__ D, int nyo, const T* __restrict__ pmax, T obj_alpha, int w_view,
    cplx<T>* __restrict__ chi0) {
  using S = RowShapeC<T, N>;
  constexpr int E = S::E, TL = S::TL;
  constexpr int G = rows_per_group(N, S::L, M);
  constexpr int n = N * N;
  constexpr int TW = M * G * N;  // terms per buffer
  // The M·G lines exchange through pass C's LDS placement when they fit its L lines, else
  // (M > L, one row per group) through one plain padded region per line.
  constexpr bool FITS = M * G <= S::L;
  constexpr int PLD = N + (N >> 4) + 1;
  using At = std::conditional_t<FITS, typename S::AtRow, LdsPlain>;
  constexpr int LDW = FITS ? S::LDS_WORDS : M * G * PLD;
  __shared__ T lre[LDW], lim[LDW];
  __shared__ cplx<T> stw[N];
  // Two term buffers (slot b in buffer b mod 2) where they fit in LDS, else one, with a
  // barrier between a slot's sums and the next slot's terms.
  constexpr bool DB = 2 * LDW * sizeof(T) + N * sizeof(cplx<T>) + 2 * TW * sizeof(cplx<T>) <=
                      160 * 1024;
  constexpr int NBUF = DB ? 2 : 1;
  __shared__ cplx<T> terms[NBUF * TW];
  const int tid = threadIdx.x;
  const int m = wave_uniform<G * TL>(tid / (G * TL)), l = (tid / TL) % G, jl = tid % TL;
  const int line = m * G + l;
  const int y = blockIdx.x * G + l;
  const int ch = blockIdx.y;
  const int b0 = ch * chunk;
  const int b1 = b0 + chunk < nb ? b0 + chunk : nb;
  stage_twiddles<T, N, G * M * TL>(stw, tw);
  // Every slot's transform uses the lane's twiddles: held in registers.
  cplx<T> twc[tw_cache_size(N, E)];
  load_tw_cache<T, N, E>(twc, jl, stw);
  cplx<T> w[E], acc[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int x = jl + TL * e;
    w[e] = ld(Wt + (size_t)(w_view ? b0 * M + m : m) * n + (size_t)y * N + x);
    acc[e] = cplx<T>{0, 0};
  }
  const T inv_p = T(1) - probe_alpha, inv_o = T(1) - obj_alpha;
  const T reg = upd_obj ? obj_alpha * pmax[0] : T(0);
  const size_t no = (size_t)nyo * nxo;
  // Slot `sb`'s sum over the modes of this thread's share of the row, from terms buffer `buf`.
  auto store_sum = [&](int sb, int buf) {
    cplx<T>* out = psi + ((size_t)(sb * M) * N + y) * N + jl;
    const cplx<T>* t = terms + buf * TW + l * N + jl;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      if (e % M != m) continue;
      cplx<T> sum = t[TL * e];
#pragma unroll
      for (int k = 1; k < M; ++k) sum = sum + t[k * G * N + TL * e];
      st(out + TL * e, sum);
    }
  };
  // Each slot's rows (field, object window, overlap) are loaded as the slot starts: held for
  // the next slot, they would cost a workgroup per CU (two instead of three), which hides the
  // loads' latency better than the prefetch did. The probe direction omits conj(q) until the
  // store (q is the same at every slot).
  for (int b = b0; b < b1; ++b) {
    cplx<T> v[E], ow[E];
    T dw[E];
    {
      const cplx<T>* pn = psi + ((size_t)(b * M + m) * N + y) * N;
      const int2 cn = corner[order[b]];
      const size_t rn = (size_t)(cn.x + y) * nxo + cn.y;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        v[e] = ld(pn + jl + TL * e);
        ow[e] = ld(O + rn + jl + TL * e);
        dw[e] = upd_obj ? D[m * no + rn + jl + TL * e] : T(0);
      }
    }
    T omax = 0;
    if (probe_dir) {
      for (int i = 0; i < na; ++i) {
        const T t = omax_part[(size_t)b * na + i];
        omax = t > omax ? t : omax;
      }
    }
    // The transform's barriers also order the previous slot's term writes before the reads
    // below, and this slot's reads of buffer (b − 1) before the writes of slot b + 1.
    const int lb = FITS ? S::base_row(line) : line * PLD;
    line_fft<T, N, E, true, At, true>(v, jl, lre + lb, lim + lb, stw, twc);
    if (upd_obj && b > b0) {
      store_sum(b - 1, DB ? (b - 1) & 1 : 0);
      if constexpr (!DB) __syncthreads();
    }
    cplx<T>* trow = terms + (DB ? b & 1 : 0) * TW + line * N + jl;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const cplx<T> o = ow[e];
      const cplx<T> d = v[e] - w[e] * o;
      if (chi0 != nullptr && m == 0) st(chi0 + ((size_t)b * M * N + y) * N + jl + TL * e, d);
      if (upd_obj) trow[TL * e] = cmulc(w[e], d) * rcp(inv_o * dw[e] + reg);
      if (probe_dir)
        acc[e] = acc[e] + cmulc(o, d) * rcp(inv_p * norm2(o) + probe_alpha * omax);
    }
  }
  if (upd_obj && b1 > b0) {
    __syncthreads();
    store_sum(b1 - 1, DB ? (b1 - 1) & 1 : 0);
  }
  if (probe_dir) {
    cplx<T>* out = dP_part + ((size_t)ch * M + m) * n + (size_t)y * N;
#pragma unroll
    for (int e = 0; e < E; ++e)
      st(out + jl + TL * e, cmulc(ld(q + (size_t)y * N + jl + TL * e), acc[e]));
  }
}

// P += s·Σ_chunks dP_part, then Wt = P·q (and, with W2, |Wt|²), at element i.
template <typename T>
__device__ __forceinline__ void probe_apply_at(int i, cplx<T>* __restrict__ P,
                                               cplx<T>* __restrict__ Wt,
                                               const cplx<T>* __restrict__ q,
                                               const cplx<T>* __restrict__ dP_part, int nchunk,
                                               int n, int total, T s, T* __restrict__ W2) {
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
  const cplx<T> w = p * ld(q + i % n);
  st(Wt + i, w);
  if (W2) W2[i] = norm2(w);
}

template <typename T>
__global__ void k_probe_apply(cplx<T>* __restrict__ P, cplx<T>* __restrict__ Wt,
                              const cplx<T>* __restrict__ q, const cplx<T>* __restrict__ dP_part,
                              int nchunk, int n, int total, T s) {
  probe_apply_at(int(blockIdx.x * blockDim.x + threadIdx.x), P, Wt, q, dP_part, nchunk, n, total,
                 s, (T*)nullptr);
}

// ---- object tiles --------------------------------------------------------------------------------
// Object tiles of 64×4 pixels, one pixel per thread (TROWS rows per thread in general).
constexpr int TILE_X = 64, TILE_Y = 4, TROWS = TILE_Y / (THREADS / TILE_X);

// One workgroup per object tile: (1) the batch slots whose N×N window meets the tile, in slot
// order, compacted with wave ballots into list[tile][·] and count[tile]; (2) with `W2` (|W_m|²
// per mode, W = P·q) or `Wv` (every view's own wave W, plane slot·M + m, under position
// offsets), each pixel's batch overlap D_m = Σ_views |W_m|² per mode into D[m][pixel] and the
// tile max of Σ_m D_m (rPIE's regulariser max).
template <typename T>
__global__ __launch_bounds__(THREADS) void k_bin_tiles(
    const int2* __restrict__ corner, const int* __restrict__ order, int nb, int N, int ntx,
    int cap, int* __restrict__ list, int* __restrict__ count, const T* __restrict__ W2,
    int M, int nyo, int nxo, T* __restrict__ D, T* __restrict__ tilemax,
    const cplx<T>* __restrict__ Wv) {
  __shared__ int wave_n[THREADS / 64];
  __shared__ int base_s;
  __shared__ T red[THREADS];
  // The corners (and batch slots) of the tile's first THREADS views, kept as they are found.
  __shared__ int2 cs[THREADS];
  __shared__ int bs[THREADS];
  const int t = tile_id(), tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ty0 = (t / ntx) * TILE_Y, tx0 = (t % ntx) * TILE_X;
  int* out = list + (size_t)t * cap;
  if (tid == 0) base_s = 0;
  __syncthreads();
  for (int s = 0; s < nb; s += THREADS) {
    const int b = s + tid;
    bool hit = false;
    int2 c = {0, 0};
    if (b < nb) {
      c = corner[order[b]];
      hit = c.x < ty0 + TILE_Y && c.x + N > ty0 && c.y < tx0 + TILE_X && c.y + N > tx0;
    }
    const unsigned long long mask = __ballot(hit);
    const int before = __popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) wave_n[wave] = __popcll(mask);
    __syncthreads();
    int off = base_s;
    for (int w = 0; w < wave; ++w) off += wave_n[w];
    if (hit) {
      out[off + before] = b;
      if (off + before < THREADS) {
        cs[off + before] = c;
        bs[off + before] = b;
      }
    }
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
  if (W2 == nullptr && Wv == nullptr) return;
  // Modes in chunks of MM, each chunk's overlaps in registers; Σ_m D_m in mode order.
  constexpr int MM = 4;
  const int X = tx0 + tid % TILE_X, Y0 = ty0 + tid / TILE_X;
  constexpr int RS = THREADS / TILE_X;  // row stride between a thread's pixels
  const int n = N * N;
  const size_t no = (size_t)nyo * nxo;
  T tot[TROWS] = {};
  for (int m0 = 0; m0 < M; m0 += MM) {
  const int mc = M - m0 < MM ? M - m0 : MM;
  T dm[TROWS][MM] = {};
  for (int s = 0; s < nl; s += THREADS) {
    const int k = nl - s < THREADS ? nl - s : THREADS;
    if (s > 0 || (m0 > 0 && nl > THREADS)) {
      if (tid < k) {
        bs[tid] = out[s + tid];
        cs[tid] = corner[order[bs[tid]]];
      }
      __syncthreads();
    }
    if (X < nxo) {
      // SU slots per step, added in slot order. One slot at a time in this form is fastest
      // (256² × 4 modes: 2 slots took 8 % longer, 4 slots 14 %, 8 slots 30 %, and a plain
      // loop with the bounds test as a branch 31 %).
      constexpr int SU = 1;
      for (int i0 = 0; i0 < k; i0 += SU) {
#pragma unroll
        for (int r = 0; r < TROWS; ++r) {
          T w2[SU][MM];
          bool ok[SU];
#pragma unroll
          for (int u = 0; u < SU; ++u) {
            const int i = i0 + u < k ? i0 + u : k - 1;  // cs holds k entries
            const int x = X - cs[i].y, y = Y0 + RS * r - cs[i].x;
            ok[u] = i0 + u < k && x >= 0 && x < N && y >= 0 && y < N;
            if (ok[u]) {
#pragma unroll
              for (int m = 0; m < MM; ++m)
                if (m < mc)
                  w2[u][m] = Wv ? norm2(Wv[((size_t)bs[i] * M + m0 + m) * n + y * N + x])
                                : W2[(size_t)(m0 + m) * n + y * N + x];
            }
          }
#pragma unroll
          for (int u = 0; u < SU; ++u)
            if (ok[u]) {
#pragma unroll
              for (int m = 0; m < MM; ++m)
                if (m < mc) dm[r][m] += w2[u][m];
            }
        }
      }
    }
    __syncthreads();
  }
  if (X < nxo && nl > 0) {
#pragma unroll
    for (int r = 0; r < TROWS; ++r) {
      const int Y = Y0 + RS * r;
      if (Y >= nyo) continue;
      const size_t p = (size_t)Y * nxo + X;
      T s = 0;
#pragma unroll
      for (int m = 0; m < MM; ++m)
        if (m < mc) {
          D[(m0 + m) * no + p] = dm[r][m];
          s += dm[r][m];
        }
      tot[r] += s;
    }
  }
  }
  T sum = 0;
  if (X < nxo && nl > 0) {
#pragma unroll
    for (int r = 0; r < TROWS; ++r)
      if (Y0 + RS * r < nyo) sum = tot[r] > sum ? tot[r] : sum;
  }
  T r = block_reduce<THREADS>(sum, red, MaxOp());
  if (tid == 0) tilemax[t] = r;
}

// The max of `count` non-negative values by one workgroup of MAX_THREADS.
constexpr int MAX_THREADS = 1024;
template <typename T>
__global__ __launch_bounds__(MAX_THREADS) void k_max_reduce(const T* __restrict__ v, int count,
                                                            T* __restrict__ out) {
  __shared__ T red[MAX_THREADS];
  T m = 0;
  for (int i = threadIdx.x; i < count; i += MAX_THREADS) m = v[i] > m ? v[i] : m;
  T r = block_reduce<MAX_THREADS>(m, red, MaxOp());
  if (threadIdx.x == 0) out[0] = r;
}

// rPIE's object update from k_pass_c_rpie_sum's per-slot sums (stored in each slot's mode-0
// plane, every mode's term already divided by its denominator): every object pixel owned by one
// thread adds, in slot order and Kahan-compensated, the terms of the views over it, then
// O += step·Σ. No atomics; the result does not depend on the launch shape.
template <typename T>
__device__ __forceinline__ void obj_update_rpie_tile(
    int t, cplx<T>* __restrict__ O, const cplx<T>* __restrict__ wsum, int N, int M,
    const int2* __restrict__ corner, const int* __restrict__ order, const int* __restrict__ list,
    const int* __restrict__ count, int cap, int nyo, int nxo, int ntx, T step) {
  constexpr int RS = THREADS / TILE_X;
  __shared__ int3 cs[THREADS];
  const int tid = threadIdx.x;
  const int X = (t % ntx) * TILE_X + tid % TILE_X, Y0 = (t / ntx) * TILE_Y + tid / TILE_X;
  const int nl = count[t];
  if (nl == 0) return;
  const int n = N * N;
  kahan<T> nr[TROWS], ni[TROWS];
  bool covered[TROWS] = {};
  const int* lst = list + (size_t)t * cap;
  for (int s = 0; s < nl; s += THREADS) {
    const int k = nl - s < THREADS ? nl - s : THREADS;
    if (tid < k) {
      const int b = lst[s + tid];
      const int2 c = corner[order[b]];
      cs[tid] = int3{b, c.x, c.y};
    }
    __syncthreads();
    // SU slots per step: their loads are issued together, then added in slot order.
    constexpr int SU = 4;
    for (int i0 = 0; i0 < k && X < nxo; i0 += SU) {
#pragma unroll
      for (int r = 0; r < TROWS; ++r) {
        cplx<T> g[SU];
        bool ok[SU];
#pragma unroll
        for (int u = 0; u < SU; ++u) {
          const int i = i0 + u < k ? i0 + u : k - 1;  // cs holds k entries
          const int x = X - cs[i].z, y = Y0 + RS * r - cs[i].y;
          ok[u] = i0 + u < k && x >= 0 && x < N && y >= 0 && y < N;
          if (ok[u]) g[u] = ld(wsum + (size_t)cs[i].x * M * n + y * N + x);
        }
#pragma unroll
        for (int u = 0; u < SU; ++u)
          if (ok[u]) {
            nr[r].add(g[u].x);
            ni[r].add(g[u].y);
            covered[r] = true;
          }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < TROWS; ++r) {
    const int Y = Y0 + RS * r;
    if (!covered[r] || Y >= nyo) continue;
    cplx<T>* o = O + (size_t)Y * nxo + X;
    st(o, ld(o) + cplx<T>{nr[r].s, ni[r].s} * step);
  }
}

template <typename T>
__global__ __launch_bounds__(THREADS) void k_obj_update_rpie_sum(
    cplx<T>* __restrict__ O, const cplx<T>* __restrict__ wsum, int N, int M,
    const int2* __restrict__ corner, const int* __restrict__ order, const int* __restrict__ list,
    const int* __restrict__ count, int cap, int nyo, int nxo, int ntx, T step) {
  obj_update_rpie_tile(tile_id(), O, wsum, N, M, corner, order, list, count, cap, nyo, nxo, ntx,
                       step);
}

// The object update's ntile workgroups and then the probe update's (THREADS elements each) in
// one launch: the two read and write disjoint buffers. The probe update also leaves |W_m|² of
// the new probe terms in W2 for the next batch's overlaps.
template <typename T>
__global__ __launch_bounds__(THREADS) void k_rpie_updates(
    cplx<T>* __restrict__ O, const cplx<T>* __restrict__ wsum, int N, int M,
    const int2* __restrict__ corner, const int* __restrict__ order, const int* __restrict__ list,
    const int* __restrict__ count, int cap, int nyo, int nxo, int ntx, T step, int ntile,
    cplx<T>* __restrict__ P, cplx<T>* __restrict__ Wt, const cplx<T>* __restrict__ q,
    const cplx<T>* __restrict__ dP_part, int nchunk, int total, T pstep, T* __restrict__ W2) {
  if (int(blockIdx.x) < ntile)
    obj_update_rpie_tile(int(blockIdx.x), O, wsum, N, M, corner, order, list, count, cap, nyo,
                         nxo, ntx, step);
  else
    probe_apply_at(int((blockIdx.x - ntile) * THREADS + threadIdx.x), P, Wt, q, dP_part, nchunk,
                   N * N, total, pstep, W2);
}

// ---- LSQML ------------------------------------------------------------------------------------------
// Batch mean of the slots' last a_far (the Poisson solver's first `prev`).
template <typename T>
__global__ void k_lsqml_prev(const T* __restrict__ a_far_all, const int* __restrict__ order, int nb,
                             T* __restrict__ prev) {
  __shared__ double red[THREADS];
  double s = 0;
  for (int b = threadIdx.x; b < nb; b += THREADS) s += double(a_far_all[order[b]]);
  double r = block_reduce<THREADS>(s, red, AddOp());
  if (threadIdx.x == 0) prev[0] = T(r / nb);
}

// One pass of the Poisson a_far solver per slot: a = clip(prev/2 + ratio/2, 0, 1) with ratio
// num/den of the slot's sums (0 where den ≤ 0); writes a[b] and, with `prev_out`, their batch
// mean. ptyrs/src/engine/lsqml.rs:137-155
template <typename T>
__global__ void k_lsqml_xi(const double* __restrict__ part, int nblk, int nb,
                           const T* __restrict__ prev, T* __restrict__ a, T* __restrict__ prev_out) {
  __shared__ double red[THREADS];
  double mine = 0;
  for (int b = threadIdx.x; b < nb; b += THREADS) {
    double num = 0, den = 0;
    for (int i = 0; i < nblk; ++i) {
      num += part[((size_t)b * nblk + i) * 2];
      den += part[((size_t)b * nblk + i) * 2 + 1];
    }
    const double ratio = den > 0.0 ? num / den : 0.0;
    double v = 0.5 * double(prev[0]) + 0.5 * ratio;
    v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
    a[b] = T(v);
    mine += v;
  }
  double r = block_reduce<THREADS>(mine, red, AddOp());
  if (threadIdx.x == 0 && prev_out) prev_out[0] = T(r / nb);
}

// Second pass sums of the Poisson solver from stored icalc: grid (nblk, nb), per workgroup
// (num, den) with `prev` the first pass's batch mean.
template <typename T>
__global__ __launch_bounds__(THREADS) void k_lsqml_xi_sums(
    const T* __restrict__ icalc, const float* __restrict__ counts, const int* __restrict__ order,
    int n, const T* __restrict__ prev, double* __restrict__ part) {
  __shared__ double red[THREADS];
  const int b = blockIdx.y, f = order[b];
  const double pv = double(prev[0]);
  double num = 0, den = 0;
  for (int j = blockIdx.x * THREADS + threadIdx.x; j < n; j += gridDim.x * THREADS) {
    const float io = counts[(size_t)f * n + j];
    if (io >= 0.f) {
      const double yp = double(icalc[(size_t)b * n + j]), yt = double(io);
      const double xi = 1.0 - yt / (yp + 1e-5);
      num += xi * (yp - yt / (1.0 - pv * xi));
      den += xi * xi * yp;
    }
  }
  double rn = block_reduce<THREADS>(num, red, AddOp());
  double rd = block_reduce<THREADS>(den, red, AddOp());
  if (threadIdx.x == 0) {
    part[((size_t)b * gridDim.x + blockIdx.x) * 2] = rn;
    part[((size_t)b * gridDim.x + blockIdx.x) * 2 + 1] = rd;
  }
}

// LSQML object direction (and, with pre = null, the object back-projection numerator of ML's
// gradient): every object pixel owned by one thread sums, in slot order,
// Σ_m Σ_views conj(W·q)·χ_m (Kahan-compensated) and divides by √(pre² + floor); 0 where no view
// covers it. W is W_m ([M][n]) or, with per_view, the slot's own plane ([slot·M + m][n], a
// multislice slice's incident waves or a moved probe); q (nullable) a plane phase, per slot
// ([slot][n]) with q_frame. With `accumulate` the result is added to dO.
// ptyrs/src/engine/lsqml.rs:183-188, :587-598; ptyrs/src/engine/ops.rs:279-310;
// ptyrs/src/engine/multislice.rs:402-430
template <typename T>
__global__ __launch_bounds__(THREADS) void k_lsqml_obj_dir(
    cplx<T>* __restrict__ dO, const cplx<T>* __restrict__ Wt, const cplx<T>* __restrict__ chi,
    int N, int M, const int2* __restrict__ corner, const int* __restrict__ order,
    const int* __restrict__ list, const int* __restrict__ count, int cap, int nyo, int nxo,
    int ntx, const T* __restrict__ pre, const T* __restrict__ floor, int per_view = 0,
    const cplx<T>* __restrict__ q = nullptr, int q_frame = 0, int accumulate = 0) {
  constexpr int RS = THREADS / TILE_X;
  __shared__ int3 cs[THREADS];
  const int t = tile_id(), tid = threadIdx.x;
  const int X = (t % ntx) * TILE_X + tid % TILE_X, Y0 = (t / ntx) * TILE_Y + tid / TILE_X;
  const int nl = count[t];
  const int n = N * N;
  kahan<T> sr[TROWS], si[TROWS];
  const int* lst = list + (size_t)t * cap;
  for (int s = 0; s < nl; s += THREADS) {
    const int k = nl - s < THREADS ? nl - s : THREADS;
    if (tid < k) {
      const int b = lst[s + tid];
      const int2 c = corner[order[b]];
      cs[tid] = int3{b, c.x, c.y};
    }
    __syncthreads();
    for (int i = 0; i < k && X < nxo; ++i) {
      const int x = X - cs[i].z;
      if (x < 0 || x >= N) continue;
      const cplx<T>* cv = chi + (size_t)cs[i].x * M * n;
      const cplx<T>* wv = per_view ? Wt + (size_t)cs[i].x * M * n : Wt;
      const cplx<T>* qv = q && q_frame ? q + (size_t)cs[i].x * n : q;
#pragma unroll
      for (int r = 0; r < TROWS; ++r) {
        const int y = Y0 + RS * r - cs[i].y;
        if (y < 0 || y >= N) continue;
        const int j = y * N + x;
        for (int m = 0; m < M; ++m) {
          cplx<T> w = ld(wv + (size_t)m * n + j);
          if (qv) w = w * ld(qv + j);
          const cplx<T> g = cmulc(w, ld(cv + (size_t)m * n + j));
          sr[r].add(g.x);
          si[r].add(g.y);
        }
      }
    }
    __syncthreads();
  }
  if (X >= nxo) return;
  const T fl = pre ? floor[0] : T(0);
#pragma unroll
  for (int r = 0; r < TROWS; ++r) {
    const int Y = Y0 + RS * r;
    if (Y >= nyo) continue;
    const size_t p = (size_t)Y * nxo + X;
    T inv = T(1);
    if (pre) {
      const T pr = pre[p];
      inv = T(1) / sqrt(pr * pr + fl);
    }
    cplx<T> v = cplx<T>{sr[r].s * inv, si[r].s * inv};
    if (accumulate) v = v + ld(dO + p);
    st(dO + p, v);
  }
}

// dP = Σ_chunks partials / nb (the frame mean of the probe directions).
template <typename T>
__global__ void k_probe_mean(cplx<T>* __restrict__ dP, const cplx<T>* __restrict__ dP_part,
                             int nchunk, int total, int nb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  kahan<T> sr, si;
  for (int c = 0; c < nchunk; ++c) {
    cplx<T> d = ld(dP_part + (size_t)c * total + i);
    sr.add(d.x);
    si.add(d.y);
  }
  st(dP + i, cplx<T>{sr.s, si.s} * (T(1) / T(nb)));
}

// Per slot and workgroup the six sums of the step-size system (FP64):
// Σ|dop|², Σ|dpo|², Re Σ dop·conj(dpo), Im Σ dop·conj(dpo), Re Σ conj(dop)·χ, Re Σ conj(dpo)·χ
// over modes and the workgroup's pixels, dop = dO_window·W·qu, dpo = O_window·dP_m·q, W = W_m or
// with per_view the slot's own plane (a moved probe), qu nullable (1).
// Grid (nblk, nb). ptyrs/src/engine/lsqml.rs:300-318
template <typename T>
__global__ __launch_bounds__(THREADS) void k_lsqml_dots(
    const cplx<T>* __restrict__ dO, const cplx<T>* __restrict__ O, int nxo,
    const cplx<T>* __restrict__ Wt, const cplx<T>* __restrict__ dP, const cplx<T>* __restrict__ q,
    const cplx<T>* __restrict__ chi, int N, int M, const int2* __restrict__ corner,
    const int* __restrict__ order, double* __restrict__ part, int per_view = 0,
    const cplx<T>* __restrict__ qu = nullptr) {
  __shared__ double red[THREADS];
  const int b = blockIdx.y;
  const int2 c = corner[order[b]];
  const int n = N * N;
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int j = blockIdx.x * THREADS + threadIdx.x; j < n; j += gridDim.x * THREADS) {
    const int y = j / N, x = j % N;
    const size_t p = (size_t)(c.x + y) * nxo + c.y + x;
    const cplx<T> ow = ld(O + p), qj = ld(q + j);
    cplx<T> dow = ld(dO + p);
    if (qu) dow = dow * ld(qu + j);
    const cplx<T>* wv = per_view ? Wt + (size_t)b * M * n : Wt;
    for (int m = 0; m < M; ++m) {
      const cplx<T> u = dow * ld(wv + (size_t)m * n + j);
      const cplx<T> v = ow * ld(dP + (size_t)m * n + j) * qj;
      const cplx<T> ch = ld(chi + ((size_t)b * M + m) * n + j);
      const cplx<T> a12 = cmulc(v, u);  // u·conj(v)
      s[0] += double(norm2(u));
      s[1] += double(norm2(v));
      s[2] += double(a12.x);
      s[3] += double(a12.y);
      s[4] += double(cmulc(u, ch).x);
      s[5] += double(cmulc(v, ch).x);
    }
  }
  for (int i = 0; i < 6; ++i) {
    double r = block_reduce<THREADS>(s[i], red, AddOp());
    if (threadIdx.x == 0) part[((size_t)b * gridDim.x + blockIdx.x) * 6 + i] = r;
  }
}

// One workgroup: per slot the 2×2 step-size system in FP64 (ptyrs/src/engine/lsqml.rs:319-335),
// a_obj = 0.9·max(x0, 0)/div (div = slices × modes) and a_probe = 0.9·max(x1, 0); then the
// object step, the trimmed mean of a_obj (MATLAB trimmean 10 %, ranks with ties broken by slot),
// and the probe step, the mean of a_probe, into step[0], step[1]. Per slot, log[b·ls + 1] gets
// a_obj and log[b·ls + ls − 1] a_probe (log[b·ls] is a_far, written before).
template <typename T>
__global__ __launch_bounds__(1024) void k_lsqml_step(const double* __restrict__ part, int nblk,
                                                     int nb, double lam, int div,
                                                     double* __restrict__ sums,
                                                     double* __restrict__ log,
                                                     double* __restrict__ step, int ls = 3) {
  __shared__ double red[1024];
  const int tid = threadIdx.x;
  double m0 = 0, m1 = 0;
  for (int b = tid; b < nb; b += 1024) {
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < nblk; ++i)
      for (int k = 0; k < 6; ++k) s[k] += part[((size_t)b * nblk + i) * 6 + k];
    s[0] += lam;
    s[1] += lam;
    for (int k = 0; k < 6; ++k) sums[(size_t)b * 6 + k] = s[k];
    m0 += s[0];
    m1 += s[1];
  }
  const double reg11 = 0.1 * block_reduce<1024>(m0, red, AddOp()) / nb;
  const double reg22 = 0.1 * block_reduce<1024>(m1, red, AddOp()) / nb;
  double pmean = 0;
  for (int b = tid; b < nb; b += 1024) {
    const double* s = sums + (size_t)b * 6;
    const double a11 = s[0] + reg11, a22 = s[1] + reg22;
    const double det = a11 * a22 - (s[2] * s[2] + s[3] * s[3]);
    const double x0 = (a22 * s[4] - s[2] * s[5]) / det;
    const double x1 = (a11 * s[5] - s[2] * s[4]) / det;
    const double ao = 0.9 * (x0 > 0.0 ? x0 : 0.0) / double(div);
    const double ap = 0.9 * (x1 > 0.0 ? x1 : 0.0);
    log[(size_t)b * ls + 1] = ao;
    log[(size_t)b * ls + ls - 1] = ap;
    pmean += ap;
  }
  const double psum = block_reduce<1024>(pmean, red, AddOp());
  __syncthreads();
  // Trimmed mean: slot b's rank among a_obj (ties by slot), kept ranks [k, nb − k).
  const int kcut = int(ceil(double(nb) * 0.1 / 2.0 - 0.5));
  double kept = 0;
  for (int b = tid; b < nb; b += 1024) {
    const double v = log[(size_t)b * ls + 1];
    int rank = 0;
    for (int o = 0; o < nb; ++o) {
      const double w = log[(size_t)o * ls + 1];
      rank += (w < v) || (w == v && o < b);
    }
    if (rank >= kcut && rank < nb - kcut) kept += v;
  }
  const double ksum = block_reduce<1024>(kept, red, AddOp());
  if (tid == 0) {
    step[0] = ksum / double(nb - 2 * kcut);
    step[1] = psum / double(nb);
  }
}

// O += step[0]·dO over the object.
template <typename T>
__global__ void k_lsqml_apply_obj(cplx<T>* __restrict__ O, const cplx<T>* __restrict__ dO,
                                  const double* __restrict__ step, size_t total) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  st(O + i, ld(O + i) + ld(dO + i) * T(step[0]));
}

// P += step[1]·dP, then Wt = P·q.
template <typename T>
__global__ void k_lsqml_apply_probe(cplx<T>* __restrict__ P, cplx<T>* __restrict__ Wt,
                                    const cplx<T>* __restrict__ q, const cplx<T>* __restrict__ dP,
                                    const double* __restrict__ step, int n, int total) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  cplx<T> p = ld(P + i) + ld(dP + i) * T(step[1]);
  st(P + i, p);
  st(Wt + i, p * ld(q + i % n));
}

// Illumination Σ_m Σ_views |P_m|² over a tile list's views, added into illum (each pixel owned
// by one thread).
template <typename T>
__global__ __launch_bounds__(THREADS) void k_illumination(
    T* __restrict__ illum, const cplx<T>* __restrict__ P, int N, int M,
    const int2* __restrict__ corner, const int* __restrict__ order, const int* __restrict__ list,
    const int* __restrict__ count, int cap, int nyo, int nxo, int ntx) {
  constexpr int RS = THREADS / TILE_X;
  __shared__ int2 cs[THREADS];
  const int t = tile_id(), tid = threadIdx.x;
  const int X = (t % ntx) * TILE_X + tid % TILE_X, Y0 = (t / ntx) * TILE_Y + tid / TILE_X;
  const int nl = count[t];
  const int n = N * N;
  T sum[TROWS] = {};
  for (int s = 0; s < nl; s += THREADS) {
    const int k = nl - s < THREADS ? nl - s : THREADS;
    if (tid < k) cs[tid] = corner[order[list[(size_t)t * cap + s + tid]]];
    __syncthreads();
    // SU views per step: their loads issued together, added in view order.
    constexpr int SU = 8;
    for (int i0 = 0; i0 < k && X < nxo; i0 += SU) {
#pragma unroll
      for (int r = 0; r < TROWS; ++r) {
        T a[SU];
#pragma unroll
        for (int u = 0; u < SU; ++u) {
          const int i = i0 + u < k ? i0 + u : k - 1;  // cs holds k entries
          const int x = X - cs[i].y, y = Y0 + RS * r - cs[i].x;
          a[u] = 0;
          if (i0 + u < k && x >= 0 && x < N && y >= 0 && y < N)
            for (int m = 0; m < M; ++m) a[u] += norm2(ld(P + (size_t)m * n + y * N + x));
        }
#pragma unroll
        for (int u = 0; u < SU; ++u) sum[r] += a[u];
      }
    }
    __syncthreads();
  }
  if (X >= nxo || nl == 0) return;
#pragma unroll
  for (int r = 0; r < TROWS; ++r) {
    const int Y = Y0 + RS * r;
    if (Y < nyo) illum[(size_t)Y * nxo + X] += sum[r];
  }
}

// pre = (pre + illum)/2, or on the first update (illum·init + illum)/2; pre's max goes to
// tilemax-style partials for the floor (0.1·max pre)². ptyrs/src/engine/lsqml.rs:671-699
template <typename T>
__global__ __launch_bounds__(THREADS) void k_pre_update(T* __restrict__ pre, const T* __restrict__ illum,
                                                        size_t total, int first, double init,
                                                        T* __restrict__ part) {
  __shared__ T red[THREADS];
  T mx = 0;
  for (size_t i = (size_t)blockIdx.x * THREADS + threadIdx.x; i < total;
       i += (size_t)gridDim.x * THREADS) {
    const T il = illum[i];
    const T p = first ? T(double(il) * init) : pre[i];
    const T v = (p + il) / T(2);
    pre[i] = v;
    mx = v > mx ? v : mx;
  }
  T r = block_reduce<THREADS>(mx, red, MaxOp());
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// floor = (0.1·max)² from partial maxima.
template <typename T>
__global__ __launch_bounds__(THREADS) void k_pre_floor(const T* __restrict__ part, int count,
                                                       T* __restrict__ floor) {
  __shared__ T red[THREADS];
  T m = 0;
  for (int i = threadIdx.x; i < count; i += THREADS) m = part[i] > m ? part[i] : m;
  T r = block_reduce<THREADS>(m, red, MaxOp());
  if (threadIdx.x == 0) floor[0] = (T(0.1) * r) * (T(0.1) * r);
}

// The slot's a_far into the step log and the per-frame store.
template <typename T>
__global__ void k_lsqml_log_afar(const T* __restrict__ a, const int* __restrict__ order, int nb,
                                 T* __restrict__ a_far_all, double* __restrict__ log, int gauss,
                                 int ls = 3) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const T v = gauss ? T(0.5) : a[b];
  a_far_all[order[b]] = v;
  log[(size_t)b * ls] = double(v);
}

// ---- zero phase ramp (ptyrs/src/engine/ptycho.rs:1134-1181) ----------------------------------
// The probe spectrum's centre of mass: per workgroup the FP64 sums of v, v·y and v·x over the
// DFT indices k of the `M` planes `spec` (each the unnormalised DFT of a probe mode), v the
// modes' |spec|² summed at k and (y, x) the fftshifted index of k. The DFT of the fftshifted
// probe is this one times (−1)^(ky+kx) on even sides, so the moduli are the same.
template <typename T>
__global__ __launch_bounds__(THREADS) void k_spectrum_moments(const cplx<T>* __restrict__ spec,
                                                              int N, int M,
                                                              double* __restrict__ part) {
  const int n = N * N;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * THREADS + threadIdx.x; i < n; i += gridDim.x * THREADS) {
    double v = 0.0;
    for (int m = 0; m < M; ++m) v += double(norm2(spec[(size_t)m * n + i]));
    const int y = (i / N + N / 2) % N, x = (i % N + N / 2) % N;
    acc[0] += v;
    acc[1] += v * double(y);
    acc[2] += v * double(x);
  }
  __shared__ double red[THREADS];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double r = block_reduce<THREADS>(acc[k], red, AddOp());
    if (threadIdx.x == 0) part[blockIdx.x * 3 + k] = r;
  }
}

// The centre of mass's offset from the plane's centre, (dy, dx) in pixels, from `nb`
// workgroups' sums (added in workgroup order per thread, then across the workgroup).
__global__ __launch_bounds__(THREADS) static void k_ramp_offset(const double* __restrict__ part,
                                                                int nb, int N,
                                                                double* __restrict__ d) {
  double acc[3] = {0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nb; b += THREADS)
    for (int k = 0; k < 3; ++k) acc[k] += part[b * 3 + k];
  __shared__ double red[THREADS];
  double s[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) s[k] = block_reduce<THREADS>(acc[k], red, AddOp());
  if (threadIdx.x == 0) {
    d[0] = s[1] / s[0] - double(N / 2);
    d[1] = s[2] / s[0] - double(N / 2);
  }
}

// `planes` planes of py × px multiplied by exp(i·(sy·(y − py/2) + sx·(x − px/2))), the phase
// in FP64 with s = sgn·2π·d/N (the probe takes sgn = −1, the object +1).
template <typename T>
__global__ __launch_bounds__(256) void k_apply_ramp(cplx<T>* __restrict__ a, int planes, int py,
                                                    int px, const double* __restrict__ d,
                                                    double sgn, int N) {
  const size_t total = size_t(planes) * py * px;
  const double sy = sgn * 2.0 * M_PI * d[0] / double(N);
  const double sx = sgn * 2.0 * M_PI * d[1] / double(N);
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < total;
       i += size_t(gridDim.x) * 256) {
    const size_t r = i % (size_t(py) * px);
    const int y = int(r / px), x = int(r % px);
    const double ph = sy * (double(y) - double(py / 2)) + sx * (double(x) - double(px / 2));
    double sn, cs;
    sincos(ph, &sn, &cs);
    const double re = double(a[i].x), im = double(a[i].y);
    a[i] = cplx<T>{T(re * cs - im * sn), T(re * sn + im * cs)};
  }
}

// ---- multislice (ptyrs/src/engine/multislice.rs) ----------------------------------------------------
// Row DFTs of `nplane` planes in place (INV: inverse), unnormalised. Grid (NBLK, nplane).
template <typename T, int N, bool INV>
__global__ __launch_bounds__((RowShape<T, N>::NT)) void k_rows(cplx<T>* __restrict__ psi,
                                                                const cplx<T>* __restrict__ tw) {
  using S = RowShape<T, N>;
  constexpr int E = S::E, TL = S::TL;
  __shared__ T lre[S::LDS_WORDS], lim[S::LDS_WORDS];
  __shared__ cplx<T> stw[N];
  const int tid = threadIdx.x, line = tid / TL, jl = tid % TL;
  const int y = blockIdx.x * S::L + line;
  cplx<T>* row = psi + (size_t)blockIdx.y * N * N + (size_t)y * N;
  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = ld(row + jl + TL * e);
  stage_twiddles<T, N, S::NT>(stw, tw);
  line_fft<T, N, E, INV, typename S::AtRow>(v, jl, lre + S::base_row(line),
                                            lim + S::base_row(line), stw);
#pragma unroll
  for (int e = 0; e < E; ++e) st(row + jl + TL * e, v[e]);
}

// An angular-spectrum gap's column passes over row-transformed planes: forward column DFT,
// times T (CONJ: conj T, the adjoint), inverse column DFT, times `scale` (1/n for the unitary
// round trip). Grid (NBLK, nplane). ptyrs/src/engine/propagate.rs:380-428
template <typename T, int N, bool CONJ>
__global__ __launch_bounds__((ColShape<T, N>::NT)) void k_gap_cols(cplx<T>* __restrict__ psi,
                                                                    const cplx<T>* __restrict__ tf,
                                                                    const cplx<T>* __restrict__ tw,
                                                                    T scale) {
  using S = ColShape<T, N>;
  constexpr int E = S::E, TL = S::TL, CG = S::L;
  __shared__ T lre[S::LDS_WORDS], lim[S::LDS_WORDS];
  __shared__ cplx<T> stw[N];
  const int tid = threadIdx.x, col = tid % CG, jl = tid / CG;
  const int2 bxy = block_xy();
  const int x = bxy.x * CG + col;
  cplx<T>* plane = psi + (size_t)bxy.y * N * N;
  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = ld(plane + (jl + TL * e) * N + x);
  stage_twiddles<T, N, S::NT>(stw, tw);
  T* re = lre + S::base_col(col);
  T* im = lim + S::base_col(col);
  line_fft<T, N, E, false, typename S::AtCol>(v, jl, re, im, stw);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const cplx<T> t = ld(tf + (jl + TL * e) * N + x);
    v[e] = (CONJ ? v[e] * conj(t) : v[e] * t) * scale;
  }
  line_fft<T, N, E, true, typename S::AtCol>(v, jl, re, im, stw);
#pragma unroll
  for (int e = 0; e < E; ++e) st(plane + (jl + TL * e) * N + x, v[e]);
}

// A slice's window product over the views of a stack (natural layout, one plane per slot and
// mode): MS_MUL out = O_s[win]·w·q, MS_SUB out −= O_s[win]·w·q, MS_ADJ out = conj(O_s[win]·q)·out.
// w is P_m (w_view = 0) or the slot's own plane (w_view = 1); q is nullable (1), per slot
// ([slot][n]) with q_frame.
// Grid (ceil(n/256), nb·M). ptyrs/src/engine/multislice.rs:173-185, :266-291
enum MsOp { MS_MUL = 0, MS_SUB, MS_ADJ };
template <typename T, int OP>
__global__ void k_ms_window(cplx<T>* __restrict__ out, const cplx<T>* __restrict__ O, int nxo,
                            const cplx<T>* __restrict__ w, int w_view,
                            const cplx<T>* __restrict__ q, const int2* __restrict__ corner,
                            const int* __restrict__ order, int N, int M, int q_frame) {
  const int n = N * N;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int k = blockIdx.y, b = k / M, m = k % M;
  const int2 c = corner[order[b]];
  const int y = j / N, x = j % N;
  cplx<T> o = ld(O + (size_t)(c.x + y) * nxo + c.y + x);
  if (q) o = o * ld(q + (q_frame ? (size_t)b * n : 0) + j);
  cplx<T>* p = out + (size_t)k * n + j;
  if (OP == MS_ADJ) {
    st(p, cmulc(o, ld(p)));
  } else {
    const cplx<T> v = o * ld(w + (size_t)(w_view ? k : m) * n + j);
    st(p, OP == MS_SUB ? ld(p) - v : v);
  }
}

// The window product of k_ms_window (MS_MUL: out = O_s[win]·w·q; MS_ADJ: out = conj(O_s[win]·q)
// ·out, in place) fused with the forward row DFT of its result, the first pass of the gap that
// follows: one pass over the planes instead of two. Grid (NBLK, nb·M).
template <typename T, int N, int OP>
__global__ __launch_bounds__((RowShape<T, N>::NT)) void k_rows_ms(
    cplx<T>* __restrict__ out, const cplx<T>* __restrict__ O, int nxo,
    const cplx<T>* __restrict__ w, int w_view, const cplx<T>* __restrict__ q,
    const int2* __restrict__ corner, const int* __restrict__ order, int M, int q_frame,
    const cplx<T>* __restrict__ tw) {
  using S = RowShape<T, N>;
  constexpr int E = S::E, TL = S::TL;
  constexpr int n = N * N;
  __shared__ T lre[S::LDS_WORDS], lim[S::LDS_WORDS];
  __shared__ cplx<T> stw[N];
  const int tid = threadIdx.x, line = tid / TL, jl = tid % TL;
  const int y = blockIdx.x * S::L + line;
  const int k = blockIdx.y, b = k / M, m = k % M;
  const int2 c = corner[order[b]];
  const cplx<T>* orow = O + (size_t)(c.x + y) * nxo + c.y;
  const cplx<T>* qrow = q ? q + (q_frame ? (size_t)b * n : 0) + (size_t)y * N : nullptr;
  cplx<T>* row = out + (size_t)k * n + (size_t)y * N;
  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int x = jl + TL * e;
    cplx<T> o = ld(orow + x);
    if (qrow) o = o * ld(qrow + x);
    if constexpr (OP == MS_ADJ)
      v[e] = cmulc(o, ld(row + x));
    else
      v[e] = o * ld(w + (size_t)(w_view ? k : m) * n + (size_t)y * N + x);
  }
  stage_twiddles<T, N, S::NT>(stw, tw);
  line_fft<T, N, E, false, typename S::AtRow>(v, jl, lre + S::base_row(line),
                                              lim + S::base_row(line), stw);
#pragma unroll
  for (int e = 0; e < E; ++e) st(row + jl + TL * e, v[e]);
}

// The probe direction of slice 0: dP_m = scale·Σ_b conj(O_0[win_b]·q)·χ_0[b·M + m], summed in
// slot order (Kahan); scale 1/nb for LSQML's frame mean, 1 for a gradient's sum, which
// `accumulate` adds to dP. Grid (ceil(n/256), M). ptyrs/src/engine/lsqml.rs:200-229
template <typename T>
__global__ void k_ms_probe_dir(cplx<T>* __restrict__ dP, const cplx<T>* __restrict__ O, int nxo,
                               const cplx<T>* __restrict__ q, const cplx<T>* __restrict__ chi,
                               const int2* __restrict__ corner, const int* __restrict__ order,
                               int nb, int N, int M, T scale, int accumulate) {
  const int n = N * N;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int m = blockIdx.y, y = j / N, x = j % N;
  const cplx<T> qj = q ? ld(q + j) : cplx<T>{T(1), T(0)};
  kahan<T> sr, si;
  for (int b = 0; b < nb; ++b) {
    const int2 c = corner[order[b]];
    const cplx<T> o = ld(O + (size_t)(c.x + y) * nxo + c.y + x) * qj;
    const cplx<T> g = cmulc(o, ld(chi + ((size_t)b * M + m) * n + j));
    sr.add(g.x);
    si.add(g.y);
  }
  cplx<T> v = cplx<T>{sr.s, si.s} * scale;
  if (accumulate) v = v + ld(dP + (size_t)m * n + j);
  st(dP + (size_t)m * n + j, v);
}

// Per slot and workgroup of a slice past the first: Re Σ conj(dop)·χ_s and Σ|dop|² over modes and
// the workgroup's pixels, dop = dO_s[win]·w_s·q. Grid (nblk, nb). ptyrs/src/engine/lsqml.rs:336-349
template <typename T>
__global__ __launch_bounds__(THREADS) void k_ms_dots(
    const cplx<T>* __restrict__ dO, int nxo, const cplx<T>* __restrict__ w,
    const cplx<T>* __restrict__ q, const cplx<T>* __restrict__ chi, int N, int M,
    const int2* __restrict__ corner, const int* __restrict__ order, double* __restrict__ part,
    int q_frame) {
  __shared__ double red[THREADS];
  const int b = blockIdx.y;
  const int2 c = corner[order[b]];
  const int n = N * N;
  double s[2] = {0, 0};
  for (int j = blockIdx.x * THREADS + threadIdx.x; j < n; j += gridDim.x * THREADS) {
    const int y = j / N, x = j % N;
    cplx<T> d = ld(dO + (size_t)(c.x + y) * nxo + c.y + x);
    if (q) d = d * ld(q + (q_frame ? (size_t)b * n : 0) + j);
    for (int m = 0; m < M; ++m) {
      const size_t i = ((size_t)b * M + m) * n + j;
      const cplx<T> u = d * ld(w + i);
      s[0] += double(cmulc(u, ld(chi + i)).x);
      s[1] += double(norm2(u));
    }
  }
  for (int i = 0; i < 2; ++i) {
    double r = block_reduce<THREADS>(s[i], red, AddOp());
    if (threadIdx.x == 0) part[((size_t)b * gridDim.x + blockIdx.x) * 2 + i] = r;
  }
}

// One workgroup: a slice past the first takes a_obj = max(0.5·num/den·0.9/div, 0) per slot
// (Eq. 23b), log[b·ls + col], and its step, the trimmed mean, into step[0].
// ptyrs/src/engine/lsqml.rs:336-349, :418-426
__global__ __launch_bounds__(1024) static void k_ms_slice_step(const double* __restrict__ part,
                                                               int nblk, int nb, int div,
                                                               double* __restrict__ log, int ls,
                                                               int col,
                                                               double* __restrict__ step) {
  __shared__ double red[1024];
  const int tid = threadIdx.x;
  for (int b = tid; b < nb; b += 1024) {
    double num = 0, den = 0;
    for (int i = 0; i < nblk; ++i) {
      num += part[((size_t)b * nblk + i) * 2];
      den += part[((size_t)b * nblk + i) * 2 + 1];
    }
    const double a = 0.5 * num / den * 0.9 / double(div);
    log[(size_t)b * ls + col] = a > 0.0 ? a : 0.0;
  }
  __syncthreads();
  const int kcut = int(ceil(double(nb) * 0.1 / 2.0 - 0.5));
  double kept = 0;
  for (int b = tid; b < nb; b += 1024) {
    const double v = log[(size_t)b * ls + col];
    int rank = 0;
    for (int o = 0; o < nb; ++o) {
      const double w = log[(size_t)o * ls + col];
      rank += (w < v) || (w == v && o < b);
    }
    if (rank >= kcut && rank < nb - kcut) kept += v;
  }
  const double ksum = block_reduce<1024>(kept, red, AddOp());
  if (tid == 0) step[0] = ksum / double(nb - 2 * kcut);
}

// ---- positions (pty-chi PATCH; ptyrs/src/engine/features.rs:334-521) ---------------------------------
// Column DFTs of `nplane` planes in place (INV: inverse), unnormalised. Grid (NBLK, nplane).
template <typename T, int N, bool INV>
__global__ __launch_bounds__((ColShape<T, N>::NT)) void k_cols(cplx<T>* __restrict__ psi,
                                                                const cplx<T>* __restrict__ tw) {
  using S = ColShape<T, N>;
  constexpr int E = S::E, TL = S::TL, CG = S::L;
  __shared__ T lre[S::LDS_WORDS], lim[S::LDS_WORDS];
  __shared__ cplx<T> stw[N];
  const int tid = threadIdx.x, col = tid % CG, jl = tid / CG;
  const int2 bxy = block_xy();
  const int x = bxy.x * CG + col;
  cplx<T>* plane = psi + (size_t)bxy.y * N * N;
  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = ld(plane + (jl + TL * e) * N + x);
  stage_twiddles<T, N, S::NT>(stw, tw);
  line_fft<T, N, E, INV, typename S::AtCol>(v, jl, lre + S::base_col(col),
                                            lim + S::base_col(col), stw);
#pragma unroll
  for (int e = 0; e < E; ++e) st(plane + (jl + TL * e) * N + x, v[e]);
}

// fftfreq of DFT index k on an axis of length len.
__device__ __forceinline__ double dft_freq(int k, int len) {
  return double(k < (len + 1) / 2 ? k : k - len) * (1.0 / double(len));
}

// Spectra moved by the slots' offsets: out[b·M + m] = spec[m]·exp(−2πi f·d_b)·scale (the probe
// move's translation phase, ptyrs/src/engine/features.rs:362-368), d_b = offs[order[b]] in
// object pixels [dy, dx]; with per_view every view's own spectrum spec[b·M + m] (out may be
// spec). Grid (ceil(n/256), nb·M).
template <typename T>
__global__ void k_ramp_planes(cplx<T>* out, const cplx<T>* spec, const double2* __restrict__ offs,
                              const int* __restrict__ order, int N, int M, double scale,
                              int per_view = 0) {
  const int n = N * N;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int k = blockIdx.y, b = k / M, m = per_view ? k : k % M;
  const double2 d = offs[order[b]];
  const double ph = -2.0 * M_PI * (dft_freq(j / N, N) * d.x + dft_freq(j % N, N) * d.y);
  double s, c;
  sincos(ph, &s, &c);
  const cplx<T> v = ld(spec + (size_t)m * n + j);
  st(out + (size_t)k * n + j,
     cplx<T>{T((double(v.x) * c - double(v.y) * s) * scale),
             T((double(v.x) * s + double(v.y) * c) * scale)});
}

// The frame mean of spectra moved back: out[m] = scale·Σ_b G[b·M + m]·exp(+2πi f·d_b), slots in
// order (Kahan). Grid (ceil(n/256), M). ptyrs/src/engine/lsqml.rs:207-229
// With `accumulate` the result is added to out (a sum over stacks, in stack order).
template <typename T>
__global__ void k_unramp_mean(cplx<T>* __restrict__ out, const cplx<T>* __restrict__ G,
                              const double2* __restrict__ offs, const int* __restrict__ order,
                              int nb, int N, int M, double scale, int accumulate = 0) {
  const int n = N * N;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int m = blockIdx.y;
  const double fy = dft_freq(j / N, N), fx = dft_freq(j % N, N);
  kahan<T> sr, si;
  for (int b = 0; b < nb; ++b) {
    const double2 d = offs[order[b]];
    double s, c;
    sincos(2.0 * M_PI * (fy * d.x + fx * d.y), &s, &c);
    const cplx<T> v = ld(G + ((size_t)b * M + m) * n + j);
    sr.add(T(double(v.x) * c - double(v.y) * s));
    si.add(T(double(v.x) * s + double(v.y) * c));
  }
  cplx<T> r{T(double(sr.s) * scale), T(double(si.s) * scale)};
  if (accumulate) r = r + ld(out + (size_t)m * n + j);
  st(out + (size_t)m * n + j, r);
}

// Each slot's spectrum of mode m moved back: out[b] = scale·G[b·M + m]·exp(+2πi f·d_b) (the
// adjoint move of a frame's direction, ptyrs/src/engine/lsqml.rs:207-222). Grid (ceil(n/256),
// nb).
template <typename T>
__global__ void k_unramp_planes(cplx<T>* __restrict__ out, const cplx<T>* __restrict__ G,
                                const double2* __restrict__ offs, const int* __restrict__ order,
                                int N, int M, int m, double scale) {
  const int n = N * N;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int b = blockIdx.y;
  const double2 d = offs[order[b]];
  double s, c;
  sincos(2.0 * M_PI * (dft_freq(j / N, N) * d.x + dft_freq(j % N, N) * d.y), &s, &c);
  const cplx<T> v = ld(G + ((size_t)b * M + m) * n + j);
  st(out + (size_t)b * n + j, cplx<T>{T((double(v.x) * c - double(v.y) * s) * scale),
                                      T((double(v.x) * s + double(v.y) * c) * scale)});
}

// The moved waves' derivatives over the offset: out[2k + a] = spec[m]·exp(−2πi f·d_b)·(−2πi
// f_a)·scale for view k = b·M + m and axis a (0: dy, 1: dx), d_b = offs[order[b]]; their
// inverse transforms are ∂W/∂d_a (ptyrs/src/engine/features.rs:370-384, 960-985). Grid
// (ceil(n/256), nb·M).
template <typename T>
__global__ void k_ramp_deriv_planes(cplx<T>* __restrict__ out, const cplx<T>* __restrict__ spec,
                                    const double2* __restrict__ offs,
                                    const int* __restrict__ order, int N, int M, double scale) {
  const int n = N * N;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int k = blockIdx.y, b = k / M, m = k % M;
  const double2 d = offs[order[b]];
  const double fy = dft_freq(j / N, N), fx = dft_freq(j % N, N);
  double s, c;
  sincos(-2.0 * M_PI * (fy * d.x + fx * d.y), &s, &c);
  const cplx<T> v = ld(spec + (size_t)m * n + j);
  const double re = double(v.x) * c - double(v.y) * s, im = double(v.x) * s + double(v.y) * c;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const double g = -2.0 * M_PI * (a == 0 ? fy : fx) * scale;  // times −2πi f_a: (re, im)·i·g
    st(out + ((size_t)2 * k + a) * n + j, cplx<T>{T(-im * g), T(re * g)});
  }
}

// Per slot and workgroup the position sums of AP's, DM's and ML's step over the slot's views
// (every mode): num_a = Re Σ conj(x − W·O)·O·∂_a W and den_a = Σ |O·∂_a W|², x the projected
// exit waves, W the moved waves ([slot·M + m][n]), ∂_a W from k_ramp_deriv_planes
// ([2(slot·M + m) + a][n]), all natural layout. part[slot][blk] = {num_y, num_x, den_y, den_x}.
// Grid (nblk, nb). ptyrs/src/engine/features.rs:987-1025
template <typename T>
__global__ __launch_bounds__(THREADS) void k_pos_sums_proj(
    const cplx<T>* __restrict__ x, const cplx<T>* __restrict__ W, const cplx<T>* __restrict__ O,
    int nxo, const int2* __restrict__ corner, const int* __restrict__ order,
    const cplx<T>* __restrict__ dW, int N, int M, double* __restrict__ part) {
  __shared__ double red[THREADS];
  const int b = blockIdx.y;
  const int n = N * N;
  const int2 c = corner[order[b]];
  double s[4] = {0, 0, 0, 0};
  for (int j = blockIdx.x * THREADS + threadIdx.x; j < n; j += gridDim.x * THREADS) {
    const cplx<T> o = ld(O + (size_t)(c.x + j / N) * nxo + c.y + j % N);
    for (int m = 0; m < M; ++m) {
      const size_t k = (size_t)b * M + m;
      const cplx<T> r = ld(x + k * n + j) - ld(W + k * n + j) * o;
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const cplx<T> u = o * ld(dW + (2 * k + a) * n + j);
        s[a] += double(cmulc(r, u).x);
        s[2 + a] += double(norm2(u));
      }
    }
  }
  for (int i = 0; i < 4; ++i) {
    const double r = block_reduce<THREADS>(s[i], red, AddOp());
    if (threadIdx.x == 0) part[((size_t)b * gridDim.x + blockIdx.x) * 4 + i] = r;
  }
}

// One workgroup: AP's, DM's and ML's position step over every frame from their sums (`part`,
// nblk per frame, frames in order): per axis 0.3·num/(den + max_frames den) clipped to ±0.1
// (0 where that denominator is 0), less its mean over the frames, added to the offsets (FP64).
// `steps` holds 4·nframe values of scratch. ptyrs/src/engine/features.rs:1026-1047
__global__ __launch_bounds__(1024) static void k_pos_step_all(const double* __restrict__ part,
                                                              int nblk, int nframe,
                                                              double2* __restrict__ offs,
                                                              double* __restrict__ steps) {
  __shared__ double red[1024];
  const int tid = threadIdx.x;
  double mden[2] = {0.0, 0.0};
  for (int f = tid; f < nframe; f += 1024) {
    double v[4] = {0, 0, 0, 0};
    for (int i = 0; i < nblk; ++i)
      for (int a = 0; a < 4; ++a) v[a] += part[((size_t)f * nblk + i) * 4 + a];
    // num in steps[f][a], den in steps[nframe + f][a] until the max is known.
    for (int a = 0; a < 2; ++a) {
      steps[(size_t)f * 2 + a] = v[a];
      steps[((size_t)nframe + f) * 2 + a] = v[2 + a];
      mden[a] = v[2 + a] > mden[a] ? v[2 + a] : mden[a];
    }
  }
  double mx[2];
  for (int a = 0; a < 2; ++a) mx[a] = block_reduce<1024>(mden[a], red, MaxOp());
  double sum[2] = {0, 0};
  for (int f = tid; f < nframe; f += 1024)
    for (int a = 0; a < 2; ++a) {
      const double den = steps[((size_t)nframe + f) * 2 + a] + mx[a];
      double st_ = 0.0;
      if (den > 0.0) {
        st_ = 0.3 * steps[(size_t)f * 2 + a] / den;
        st_ = st_ < -0.1 ? -0.1 : (st_ > 0.1 ? 0.1 : st_);
      }
      steps[(size_t)f * 2 + a] = st_;
      sum[a] += st_;
    }
  double mean[2];
  for (int a = 0; a < 2; ++a) mean[a] = block_reduce<1024>(sum[a], red, AddOp()) / nframe;
  for (int f = tid; f < nframe; f += 1024) {
    double2 d = offs[f];
    d.x += steps[(size_t)f * 2] - mean[0];
    d.y += steps[(size_t)f * 2 + 1] - mean[1];
    offs[f] = d;
  }
}

// planes[k] *= q (CONJ: conj q) over `nplane` planes. Grid (ceil(n/256), nplane).
template <typename T, bool CONJ>
__global__ void k_times_q(cplx<T>* __restrict__ planes, const cplx<T>* __restrict__ q, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  cplx<T>* p = planes + (size_t)blockIdx.y * n + j;
  const cplx<T> qj = ld(q + j);
  st(p, CONJ ? cmulc(qj, ld(p)) : ld(p) * qj);
}

// The last slice's plane phase displaced by each slot's offset, q(r − d) =
// q·exp(i Σ_a c_a·d_a·(d_a − 2 r_a)) over the centred pixel r. Grid (ceil(n/256), nb).
// ptyrs/src/engine/propagate.rs:134-163
template <typename T>
__global__ void k_moved_quad(cplx<T>* __restrict__ out, const cplx<T>* __restrict__ q,
                             const double2* __restrict__ offs, const int* __restrict__ order,
                             int N, double cy, double cx) {
  const int n = N * N;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int b = blockIdx.y;
  const double2 d = offs[order[b]];
  const double ry = double(j / N - N / 2), rx = double(j % N - N / 2);
  const double ph = cy * d.x * (d.x - 2.0 * ry) + cx * d.y * (d.y - 2.0 * rx);
  double s, c;
  sincos(ph, &s, &c);
  st(out + (size_t)b * n + j, ld(q + j) * cplx<T>{T(c), T(s)});
}

// The object windows of the slots, one plane each. Grid (ceil(n/256), nb).
template <typename T>
__global__ void k_patch(cplx<T>* __restrict__ out, const cplx<T>* __restrict__ O, int nxo,
                        const int2* __restrict__ corner, const int* __restrict__ order, int N) {
  const int n = N * N;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int b = blockIdx.y;
  const int2 c = corner[order[b]];
  st(out + (size_t)b * n + j, ld(O + (size_t)(c.x + j / N) * nxo + c.y + j % N));
}

// The window's Fourier derivatives: out[2b + a] = spec[b]·(2πi f_a)·scale. Grid (ceil(n/256), nb).
template <typename T>
__global__ void k_deriv(cplx<T>* __restrict__ out, const cplx<T>* __restrict__ spec, int N,
                        double scale) {
  const int n = N * N;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int b = blockIdx.y;
  const cplx<T> v = ld(spec + (size_t)b * n + j);
  const double f[2] = {dft_freq(j / N, N), dft_freq(j % N, N)};
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const double w = 2.0 * M_PI * f[a] * scale;  // times i
    st(out + ((size_t)2 * b + a) * n + j, cplx<T>{T(-double(v.y) * w), T(double(v.x) * w)});
  }
}

// Per slot and workgroup the position sums of its first view (mode 0):
// num_a = Re Σ conj(g_a·w·q)·χ, den_a = Σ |g_a·w·q|² over the workgroup's pixels, g_a the window's
// derivative along axis a; w per slot and mode ([slot·M + m][n]) with w_slot, else mode 0's
// plane for every slot; q nullable, per slot with q_frame. part[b][blk] = {num_y, num_x, den_y,
// den_x}. Grid (nblk, nb). ptyrs/src/engine/features.rs:457-497
template <typename T>
__global__ __launch_bounds__(THREADS) void k_pos_sums(
    const cplx<T>* __restrict__ grad, const cplx<T>* __restrict__ w, const cplx<T>* __restrict__ q,
    int q_frame, const cplx<T>* __restrict__ chi, int N, int M, double* __restrict__ part,
    int w_slot) {
  __shared__ double red[THREADS];
  const int b = blockIdx.y;
  const int n = N * N;
  double s[4] = {0, 0, 0, 0};
  const size_t k = (size_t)b * M * n;
  const size_t kw = w_slot ? k : 0;
  for (int j = blockIdx.x * THREADS + threadIdx.x; j < n; j += gridDim.x * THREADS) {
    cplx<T> wq = ld(w + kw + j);
    if (q) wq = wq * ld(q + (q_frame ? (size_t)b * n : 0) + j);
    const cplx<T> c = ld(chi + k + j);
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const cplx<T> u = ld(grad + ((size_t)2 * b + a) * n + j) * wq;
      s[a] += double(cmulc(u, c).x);
      s[2 + a] += double(norm2(u));
    }
  }
  for (int i = 0; i < 4; ++i) {
    double r = block_reduce<THREADS>(s[i], red, AddOp());
    if (threadIdx.x == 0) part[((size_t)b * gridDim.x + blockIdx.x) * 4 + i] = r;
  }
}

// One workgroup: the batch's position directions num/(den + max(max_batch den, eps)) and
// pty-chi's SGD step 0.3·Δ clipped per axis to min(10·mean|dx − mean dx|, 0.1) over all nframe
// frames (dx = 0 off the batch), added to the batch's offsets (FP64).
// ptyrs/src/engine/features.rs:499-521
__global__ __launch_bounds__(1024) static void k_pos_step(const double* __restrict__ part,
                                                          int nblk, int nb, int nframe,
                                                          double eps,
                                                          const int* __restrict__ order,
                                                          double2* __restrict__ offs,
                                                          double* __restrict__ sums) {
  __shared__ double red[1024];
  const int tid = threadIdx.x;
  double mden[2] = {eps, eps};
  for (int b = tid; b < nb; b += 1024) {
    double s[4] = {0, 0, 0, 0};
    for (int i = 0; i < nblk; ++i)
      for (int a = 0; a < 4; ++a) s[a] += part[((size_t)b * nblk + i) * 4 + a];
    for (int a = 0; a < 4; ++a) sums[(size_t)b * 4 + a] = s[a];
    for (int a = 0; a < 2; ++a) mden[a] = s[2 + a] > mden[a] ? s[2 + a] : mden[a];
  }
  double mx[2];
  for (int a = 0; a < 2; ++a) mx[a] = block_reduce<1024>(mden[a], red, MaxOp());
  // dx per slot, its sum over the batch (the mean over all frames has the others at 0)
  double dsum[2] = {0, 0};
  for (int b = tid; b < nb; b += 1024)
    for (int a = 0; a < 2; ++a) {
      const double dx = 0.3 * sums[(size_t)b * 4 + a] / (sums[(size_t)b * 4 + 2 + a] + mx[a]);
      sums[(size_t)b * 4 + a] = dx;
      dsum[a] += dx;
    }
  double mean[2];
  for (int a = 0; a < 2; ++a) mean[a] = block_reduce<1024>(dsum[a], red, AddOp()) / nframe;
  double dev[2] = {0, 0};
  for (int b = tid; b < nb; b += 1024)
    for (int a = 0; a < 2; ++a) dev[a] += fabs(sums[(size_t)b * 4 + a] - mean[a]);
  double limit[2];
  for (int a = 0; a < 2; ++a) {
    const double mad = (block_reduce<1024>(dev[a], red, AddOp()) +
                        double(nframe - nb) * fabs(mean[a])) / nframe;
    limit[a] = 10.0 * mad < 0.1 ? 10.0 * mad : 0.1;
  }
  for (int b = tid; b < nb; b += 1024) {
    double2 d = offs[order[b]];
    double v[2];
    for (int a = 0; a < 2; ++a) {
      const double x = sums[(size_t)b * 4 + a];
      v[a] = x < -limit[a] ? -limit[a] : (x > limit[a] ? limit[a] : x);
    }
    d.x += v[0];
    d.y += v[1];
    offs[order[b]] = d;
  }
}

// ---- AP, DM, RAAR ----------------------------------------------------------------------------------
// Pass C of the projection algorithms. Grid (NBLK, M, nchunk); each workgroup runs over the
// slots of its chunk: inverse row transform of the projected field Π, then the new carried wave
// ψ' = Π (AP, RAAR) or Π + (1 − a)·(PO − R) with R = (2 − a)·PO − (1 − a)·ψ (DM), written to
// carried[frame]. With `probe_sums` (AP: both updates read the old arrays) it adds
// conj(q·O)·ψ' and |O|² of the slot's window into per-chunk partials. With w_view every slot's
// mode has its own wave W (plane slot·M + m of Wt, a moved probe) in PO.
template <typename T, int N, bool H = false>
__global__ __launch_bounds__((RowShapeC<T, N>::NT)) void k_pass_c_proj(
    const cplx<T>* __restrict__ psi, Carried<T, H> carried, const cplx<T>* __restrict__ O,
    int nxo, const cplx<T>* __restrict__ Wt, const cplx<T>* __restrict__ q,
    const int2* __restrict__ corner, const int* __restrict__ order, int nb, int M, int chunk,
    int dm, T c0, T c1, int probe_sums, cplx<T>* __restrict__ num_part, T* __restrict__ den_part,
    const cplx<T>* __restrict__ tw, int w_view = 0) {
  using S = RowShapeC<T, N>;
  constexpr int E = S::E, TL = S::TL;
  constexpr int n = N * N;
  __shared__ T lre[S::LDS_WORDS], lim[S::LDS_WORDS];
  __shared__ cplx<T> stw[N];
  const int tid = threadIdx.x, l = tid / TL, jl = tid % TL;
  const int y = blockIdx.x * S::L + l;
  const int m = blockIdx.y, ch = blockIdx.z;
  const int b0 = ch * chunk;
  const int b1 = b0 + chunk < nb ? b0 + chunk : nb;
  stage_twiddles<T, N, S::NT>(stw, tw);
  cplx<T> w[E], qq[E], acc[E];
  T den[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int x = jl + TL * e;
    w[e] = ld(Wt + (size_t)m * n + (size_t)y * N + x);
    qq[e] = ld(q + (size_t)y * N + x);
    acc[e] = cplx<T>{0, 0};
    den[e] = 0;
  }
  for (int b = b0; b < b1; ++b) {
    const int f = order[b];
    const int2 c = corner[f];
    const cplx<T>* prow = psi + ((size_t)(b * M + m) * N + y) * N;
    cplx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = ld(prow + jl + TL * e);
    line_fft<T, N, E, true, typename S::AtRow>(v, jl, lre + S::base_row(l),
                                               lim + S::base_row(l), stw);
    const cplx<T>* orow = O + (size_t)(c.x + y) * nxo + c.y;
    const size_t crow = (((size_t)f * M + m) * N + y) * N;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int x = jl + TL * e;
      const cplx<T> o = ld(orow + x);
      cplx<T> p = v[e];
      if (dm) {
        const cplx<T> we =
            w_view ? ld(Wt + ((size_t)b * M + m) * n + (size_t)y * N + x) : w[e];
        const cplx<T> po = we * o;
        const cplx<T> r = po * c0 - carried.get(crow + x) * c1;
        p = p + (po - r) * c1;
      }
      carried.put(crow + x, p);
      if (probe_sums) {
        acc[e] = acc[e] + cmulc(o * qq[e], p);
        den[e] += norm2(o);
      }
    }
  }
  if (probe_sums) {
    const size_t base = ((size_t)ch * M + m) * n + (size_t)y * N;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      st(num_part + base + jl + TL * e, acc[e]);
      if (m == 0) den_part[(size_t)ch * n + (size_t)y * N + jl + TL * e] = den[e];
    }
  }
}

// Probe sums of carried waves against the current object (DM, RAAR: the probe update sees the
// new object): grid (N·N/THREADS, M, nchunk), per-chunk partials of conj(q·O)·ψ and |O|².
template <typename T, bool H = false>
__global__ __launch_bounds__(THREADS) void k_probe_sums(
    Carried<T, H> carried, const cplx<T>* __restrict__ O, int nxo,
    const cplx<T>* __restrict__ q, const int2* __restrict__ corner, const int* __restrict__ order,
    int nb, int N, int M, int chunk, cplx<T>* __restrict__ num_part, T* __restrict__ den_part) {
  const int n = N * N;
  const int i = blockIdx.x * THREADS + threadIdx.x;
  const int m = blockIdx.y, ch = blockIdx.z;
  if (i >= n) return;
  const int y = i / N, x = i % N;
  const int b0 = ch * chunk, b1 = b0 + chunk < nb ? b0 + chunk : nb;
  const cplx<T> qq = ld(q + i);
  cplx<T> acc = {0, 0};
  T den = 0;
  for (int b = b0; b < b1; ++b) {
    const int f = order[b];
    const int2 c = corner[f];
    const cplx<T> o = ld(O + (size_t)(c.x + y) * nxo + c.y + x);
    acc = acc + cmulc(o * qq, carried.get(((size_t)f * M + m) * n + i));
    den += norm2(o);
  }
  st(num_part + ((size_t)ch * M + m) * n + i, acc);
  if (m == 0) den_part[(size_t)ch * n + i] = den;
}

// Sums the probe partials over chunks (Kahan) into num[M][n] and den[n], and writes the per
// workgroup max of den for the regulariser.
template <typename T>
__global__ __launch_bounds__(THREADS) void k_probe_reduce(
    const cplx<T>* __restrict__ num_part, const T* __restrict__ den_part, int nchunk, int n, int M,
    cplx<T>* __restrict__ num, T* __restrict__ den, T* __restrict__ maxpart) {
  __shared__ T red[THREADS];
  const int i = blockIdx.x * THREADS + threadIdx.x;
  T mx = 0;
  if (i < n) {
    kahan<T> d;
    for (int c = 0; c < nchunk; ++c) d.add(den_part[(size_t)c * n + i]);
    den[i] = d.s;
    mx = d.s;
    for (int m = 0; m < M; ++m) {
      kahan<T> sr, si;
      for (int c = 0; c < nchunk; ++c) {
        const cplx<T> v = ld(num_part + ((size_t)c * M + m) * n + i);
        sr.add(v.x);
        si.add(v.y);
      }
      st(num + (size_t)m * n + i, cplx<T>{sr.s, si.s});
    }
  }
  T r = block_reduce<THREADS>(mx, red, MaxOp());
  if (threadIdx.x == 0) maxpart[blockIdx.x] = r;
}

// PyNX blend of the probe: P = (num + reg·P)/max(den + reg, 1e-12), reg = inertia·max(den);
// then Wt = P·q. ptyrs/src/engine/ops.rs:367-381
template <typename T>
__global__ void k_probe_blend(cplx<T>* __restrict__ P, cplx<T>* __restrict__ Wt,
                              const cplx<T>* __restrict__ q, const cplx<T>* __restrict__ num,
                              const T* __restrict__ den, const T* __restrict__ dmax, T inertia,
                              int n, int total) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const T reg = dmax[0] * inertia;
  const T d = den[i % n] + reg;
  const cplx<T> p = (ld(num + i) + ld(P + i) * reg) * (T(1) / (d > T(1e-12) ? d : T(1e-12)));
  st(P + i, p);
  st(Wt + i, p * ld(q + i % n));
}

// PyNX blend of the object from tile-owned sums over every view: every pixel owned by one thread
// sums Σ_m Σ_views conj(W_m)·ψ_m (Kahan) in view order, and O = (num + reg·O)/max(norm + reg,
// 1e-12) with norm = Σ_m D_m from k_bin_tiles and reg = inertia·max(norm). With w_view W is the
// frame's own wave (plane frame·M + m of Wt, a moved probe; the order is every frame's).
template <typename T, bool H = false>
__global__ __launch_bounds__(THREADS) void k_obj_blend(
    cplx<T>* __restrict__ O, const cplx<T>* __restrict__ Wt, const T* __restrict__ D,
    Carried<T, H> carried, int N, int M, const int2* __restrict__ corner,
    const int* __restrict__ order, const int* __restrict__ list, const int* __restrict__ count,
    int cap, int nyo, int nxo, int ntx, const T* __restrict__ nmax, T inertia, int w_view = 0) {
  __shared__ int3 cs[THREADS];
  const int t = tile_id(), tid = threadIdx.x;
  const int X = (t % ntx) * TILE_X + tid % TILE_X, Y = (t / ntx) * TILE_Y + tid / TILE_X;
  const int nl = count[t];
  const bool inside = X < nxo && Y < nyo;
  const int n = N * N;
  kahan<T> sr, si;
  const int* lst = list + (size_t)t * cap;
  for (int s = 0; s < nl; s += THREADS) {
    const int k = nl - s < THREADS ? nl - s : THREADS;
    if (tid < k) {
      const int f = order[lst[s + tid]];
      const int2 c = corner[f];
      cs[tid] = int3{f, c.x, c.y};
    }
    __syncthreads();
    for (int i = 0; i < k && inside; ++i) {
      const int y = Y - cs[i].y, x = X - cs[i].z;
      if (y >= 0 && y < N && x >= 0 && x < N) {
        const int j = y * N + x;
        const size_t cv = (size_t)cs[i].x * M * n + j;
        const cplx<T>* wv = w_view ? Wt + (size_t)cs[i].x * M * n : Wt;
        for (int m = 0; m < M; ++m) {
          const cplx<T> g = cmulc(ld(wv + (size_t)m * n + j), carried.get(cv + (size_t)m * n));
          sr.add(g.x);
          si.add(g.y);
        }
      }
    }
    __syncthreads();
  }
  if (!inside) return;
  const size_t no = (size_t)nyo * nxo, p = (size_t)Y * nxo + X;
  T norm = 0;
  if (nl > 0)
    for (int m = 0; m < M; ++m) norm += D[m * no + p];
  const T reg = nmax[0] * inertia;
  const T d = norm + reg;
  const cplx<T> o = (cplx<T>{sr.s, si.s} + ld(O + p) * reg) * (T(1) / (d > T(1e-12) ? d : T(1e-12)));
  st(O + p, o);
}

// ---- The probe's conjugate-gradient step under moved waves (AP, DM, RAAR with PATCH offsets;
// ptyrs/src/engine/features.rs:653-720) ---------------------------------------------------------
// Per view of a stack the probe-space residual term g_v = conj(O[win])·(ψ_v − W_v·O[win]), ψ the
// frame's carried wave, W its moved wave (stack-local plane). Grid (ceil(n/256), nb·M).
template <typename T>
__global__ void k_resid_dir(cplx<T>* __restrict__ g, const cplx<T>* __restrict__ carried,
                            const cplx<T>* __restrict__ W, const cplx<T>* __restrict__ O, int nxo,
                            const int2* __restrict__ corner, const int* __restrict__ order,
                            int N, int M) {
  const int n = N * N;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int v = blockIdx.y, b = v / M, m = v % M;
  const int f = order[b];
  const int2 c = corner[f];
  const cplx<T> o = ld(O + (size_t)(c.x + j / N) * nxo + c.y + j % N);
  const cplx<T> r = ld(carried + ((size_t)f * M + m) * n + j) - ld(W + (size_t)v * n + j) * o;
  st(g + (size_t)v * n + j, cmulc(o, r));
}

// Σ_f |O[win_f]|² per window pixel over the frames `order[0..nb)` (FP64), the probe sums' norm
// (the same for every mode). Grid (ceil(n/256)).
template <typename T>
__global__ void k_win_power(T* __restrict__ out, const cplx<T>* __restrict__ O, int nxo,
                            const int2* __restrict__ corner, const int* __restrict__ order,
                            int nb, int N) {
  const int n = N * N;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int y = j / N, x = j % N;
  double s = 0;
  for (int b = 0; b < nb; ++b) {
    const int2 c = corner[order[b]];
    const cplx<T> o = ld(O + (size_t)(c.x + y) * nxo + c.y + x);
    s += double(o.x) * double(o.x) + double(o.y) * double(o.y);
  }
  out[j] = T(s);
}

// Per mode and workgroup: the preconditioned step s = num/max(norm + reg, 1e-12) (into step),
// and the partials of Σ |num|²/d (slope) and reg·Σ |s|² into part[(m·gridDim.x + blk)·2],
// reg = inertia·max norm. Grid (nblk, M).
template <typename T>
__global__ __launch_bounds__(THREADS) void k_cg_probe_sums(
    const cplx<T>* __restrict__ num, const T* __restrict__ norm, const T* __restrict__ nmax,
    T inertia, int n, cplx<T>* __restrict__ step, double* __restrict__ part) {
  __shared__ double red[THREADS];
  const int m = blockIdx.y;
  const T reg = nmax[0] * inertia;
  double a = 0, c = 0;
  for (int j = blockIdx.x * THREADS + threadIdx.x; j < n; j += gridDim.x * THREADS) {
    const T dv = norm[j] + reg;
    const T d = dv > T(1e-12) ? dv : T(1e-12);
    const cplx<T> g = ld(num + (size_t)m * n + j);
    const cplx<T> sv = g * (T(1) / d);
    st(step + (size_t)m * n + j, sv);
    a += double(norm2(g)) / double(d);
    c += double(reg) * double(norm2(sv));
  }
  for (int i = 0; i < 2; ++i) {
    double r = block_reduce<THREADS>(i == 0 ? a : c, red, AddOp());
    if (threadIdx.x == 0) part[((size_t)m * gridDim.x + blockIdx.x) * 2 + i] = r;
  }
}

// Per view of a stack and workgroup Σ |W_v·O[win]|² (FP64) into part[(v0 + v)·nblk + blk], v0
// the stack's first view. Grid (nblk, nb·M).
template <typename T>
__global__ __launch_bounds__(THREADS) void k_view_power(const cplx<T>* __restrict__ W,
                                                        const cplx<T>* __restrict__ O, int nxo,
                                                        const int2* __restrict__ corner,
                                                        const int* __restrict__ order, int N,
                                                        int M, size_t v0,
                                                        double* __restrict__ part) {
  __shared__ double red[THREADS];
  const int v = blockIdx.y, b = v / M;
  const int n = N * N;
  const int2 c = corner[order[b]];
  double s = 0;
  for (int j = blockIdx.x * THREADS + threadIdx.x; j < n; j += gridDim.x * THREADS) {
    const cplx<T> o = ld(O + (size_t)(c.x + j / N) * nxo + c.y + j % N);
    s += double(norm2(ld(W + (size_t)v * n + j) * o));
  }
  const double r = block_reduce<THREADS>(s, red, AddOp());
  if (threadIdx.x == 0) part[(v0 + v) * gridDim.x + blockIdx.x] = r;
}

// One thread: per mode α_m = slope_m / (reg·Σ|s_m|² + Σ_views of m ‖W_s·O‖²) where the slope is
// positive, else 0 (the views summed in view order). ptyrs/src/engine/features.rs:686-716
__global__ static void k_cg_probe_alpha(const double* __restrict__ part, int nblk,
                                        const double* __restrict__ vpart, int vblk, int nview,
                                        int M, double* __restrict__ alpha) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int m = 0; m < M; ++m) {
    double slope = 0, curv = 0;
    for (int i = 0; i < nblk; ++i) {
      slope += part[((size_t)m * nblk + i) * 2];
      curv += part[((size_t)m * nblk + i) * 2 + 1];
    }
    for (int v = m; v < nview; v += M)
      for (int i = 0; i < vblk; ++i) curv += vpart[(size_t)v * vblk + i];
    alpha[m] = slope > 0.0 ? slope / curv : 0.0;
  }
}

// P_m += α_m·s_m. Grid (ceil(total/256)).
template <typename T>
__global__ void k_probe_axpy(cplx<T>* __restrict__ P, const cplx<T>* __restrict__ step,
                             const double* __restrict__ alpha, int n, int total) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const double a = alpha[i / n];
  const cplx<T> s = ld(step + i), p = ld(P + i);
  st(P + i, cplx<T>{T(double(p.x) + a * double(s.x)), T(double(p.y) + a * double(s.y))});
}

// Σ|O| and Σ|P| in FP64: per-workgroup partials into part[2·blocks]. The host-side factors of
// split_scale (ptyrs/src/engine/ptycho.rs:583-591) are then applied by k_balance.
template <typename T>
__global__ __launch_bounds__(THREADS) void k_abs_sums(const cplx<T>* __restrict__ O, size_t no,
                                                      const cplx<T>* __restrict__ P, size_t np,
                                                      double* __restrict__ part) {
  __shared__ double red[THREADS];
  double so = 0, sp = 0;
  for (size_t i = (size_t)blockIdx.x * THREADS + threadIdx.x; i < no; i += (size_t)gridDim.x * THREADS) {
    const cplx<T> v = ld(O + i);
    so += sqrt(double(v.x) * v.x + double(v.y) * v.y);
  }
  for (size_t i = (size_t)blockIdx.x * THREADS + threadIdx.x; i < np; i += (size_t)gridDim.x * THREADS) {
    const cplx<T> v = ld(P + i);
    sp += sqrt(double(v.x) * v.x + double(v.y) * v.y);
  }
  double a = block_reduce<THREADS>(so, red, AddOp());
  double b = block_reduce<THREADS>(sp, red, AddOp());
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

// P *= sqrt(Σ|O|/Σ|P|·s), O *= sqrt(Σ|P|/Σ|O|·s), Wt = P·q; sums from k_abs_sums' partials.
template <typename T>
__global__ __launch_bounds__(THREADS) void k_balance(cplx<T>* __restrict__ O, size_t no,
                                                     cplx<T>* __restrict__ P, cplx<T>* __restrict__ Wt,
                                                     const cplx<T>* __restrict__ q, int n, size_t np,
                                                     const double* __restrict__ part, int nparts,
                                                     double s, const double* __restrict__ calc,
                                                     double total) {
  __shared__ double fac[2];
  if (threadIdx.x == 0) {
    // With `calc` (the model photons over unmasked pixels), s = √(total/calc): PyNX
    // ScaleObjProbe(absolute=True), ptyrs/src/engine/ptycho.rs:571-581.
    if (calc) s = sqrt(total / calc[0]);
    double so = 0, sp = 0;
    for (int i = 0; i < nparts; ++i) {
      so += part[2 * i];
      sp += part[2 * i + 1];
    }
    fac[0] = sqrt(sp / so * s);  // object
    fac[1] = sqrt(so / sp * s);  // probe
  }
  __syncthreads();
  const T fo = T(fac[0]), fp = T(fac[1]);
  for (size_t i = (size_t)blockIdx.x * THREADS + threadIdx.x; i < no; i += (size_t)gridDim.x * THREADS)
    st(O + i, ld(O + i) * fo);
  for (size_t i = (size_t)blockIdx.x * THREADS + threadIdx.x; i < np; i += (size_t)gridDim.x * THREADS) {
    const cplx<T> p = ld(P + i) * fp;
    st(P + i, p);
    st(Wt + i, p * ld(q + i % n));
  }
}

// ---- ML (PyNX Poisson conjugate gradient) ------------------------------------------------------------
// Pass A of the Newton step: per view the rows of po = W·O, a = W·dO + dW·O and b = dW·dO
// (dW = dP·q; with `w_view` W and dW are every view's own moved waves, plane slot·M + m),
// each row-transformed, into psi3[field][slot][mode][N][N].
template <typename T, int N>
__global__ __launch_bounds__((RowShape<T, N>::NT)) void k_pass_a3(
    const cplx<T>* __restrict__ O, const cplx<T>* __restrict__ dO, int nxo,
    const cplx<T>* __restrict__ Wt, const cplx<T>* __restrict__ dWt,
    const int2* __restrict__ corner, const int* __restrict__ order, int M, size_t field,
    cplx<T>* __restrict__ psi3, const cplx<T>* __restrict__ tw, int w_view = 0) {
  using S = RowShape<T, N>;
  constexpr int E = S::E, TL = S::TL;
  __shared__ T lre[S::LDS_WORDS], lim[S::LDS_WORDS];
  __shared__ cplx<T> stw[N];
  const int tid = threadIdx.x, l = tid / TL, jl = tid % TL;
  const int y = blockIdx.x * S::L + l;
  const int b = blockIdx.y / M, m = blockIdx.y % M;
  stage_twiddles<T, N, S::NT>(stw, tw);
  const int2 c = corner[order[b]];
  const size_t op = (size_t)(c.x + y) * nxo + c.y;
  const size_t wp = ((size_t)(w_view ? int(blockIdx.y) : m) * N + y) * N;
  cplx<T> po[E], a[E], bb[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int x = jl + TL * e;
    const cplx<T> o = ld(O + op + x), d = ld(dO + op + x);
    const cplx<T> w = ld(Wt + wp + x), dw = ld(dWt + wp + x);
    po[e] = w * o;
    a[e] = w * d + dw * o;
    bb[e] = dw * d;
  }
  T* re = lre + S::base_row(l);
  T* im = lim + S::base_row(l);
  line_fft<T, N, E, false, typename S::AtRow>(po, jl, re, im, stw);
  line_fft<T, N, E, false, typename S::AtRow>(a, jl, re, im, stw);
  line_fft<T, N, E, false, typename S::AtRow>(bb, jl, re, im, stw);
  const size_t row = ((size_t)(b * M + m) * N + y) * N;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    st(psi3 + row + jl + TL * e, po[e]);
    st(psi3 + field + row + jl + TL * e, a[e]);
    st(psi3 + 2 * field + row + jl + TL * e, bb[e]);
  }
}

// Pass B of the Newton step: column transforms of po, a, b per mode, then over unmasked pixels
// with calc = max(Σ|po|², 1e-32), r = Σ Re(conj(po)·a), s = Σ |a|² + 2·Re(conj(po)·b),
// d = 1 − iobs/calc: num −= 2·r·d, den += 2·s·d + iobs·(2r/calc)² (FP64, per workgroup into
// part[slot][blockIdx.x][2]). ptyrs/src/engine/ops.rs:197-253
template <typename T, int N, int M, bool ROW>
__global__ __launch_bounds__((ColShape<T, N>::NT)) void k_pass_b_gamma(
    const cplx<T>* __restrict__ psi3, size_t field, const float* __restrict__ counts,
    const int* __restrict__ order, double* __restrict__ part, const cplx<T>* __restrict__ tw,
    T norm) {
  using S = ColShape<T, N>;
  constexpr int E = S::E, TL = S::TL, CG = S::L;
  constexpr int n = N * N;
  __shared__ T lre[S::LDS_WORDS], lim[S::LDS_WORDS];
  __shared__ cplx<T> stw[N];
  __shared__ double red[S::NT];
  const int tid = threadIdx.x;
  const int line = ROW ? tid / TL : tid % CG, jl = ROW ? tid % TL : tid / CG;
  using At = std::conditional_t<ROW, typename S::AtRow, typename S::AtCol>;
  const int2 bxy = block_xy();
  const int x = bxy.x * CG + line;
  const int b = bxy.y;
  const int f = order[b];
  T* re = lre + (ROW ? S::base_row(line) : S::base_col(line));
  T* im = lim + (ROW ? S::base_row(line) : S::base_col(line));
  auto at = [&](int e) { return ROW ? x * N + jl + TL * e : (jl + TL * e) * N + x; };
  stage_twiddles<T, N, S::NT>(stw, tw);
  T calc[E], r[E], s[E];
#pragma unroll
  for (int e = 0; e < E; ++e) calc[e] = r[e] = s[e] = 0;
  for (int m = 0; m < M; ++m) {
    const size_t plane = (size_t)(b * M + m) * n;
    cplx<T> po[E], a[E], bb[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const size_t i = plane + at(e);
      po[e] = ld(psi3 + i);
      a[e] = ld(psi3 + field + i);
      bb[e] = ld(psi3 + 2 * field + i);
    }
    line_fft<T, N, E, false, At>(po, jl, re, im, stw);
    line_fft<T, N, E, false, At>(a, jl, re, im, stw);
    line_fft<T, N, E, false, At>(bb, jl, re, im, stw);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      calc[e] += norm2(po[e]);
      r[e] += cmulc(po[e], a[e]).x;
      s[e] += norm2(a[e]) + T(2) * cmulc(po[e], bb[e]).x;
    }
  }
  const double inv_n = double(norm);
  double num = 0, den = 0;
  const float* cnt = counts + (size_t)f * n;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const float io = cnt[at(e)];
    if (io >= 0.f) {
      double cl = double(calc[e]) * inv_n;
      cl = cl > 1e-32 ? cl : 1e-32;
      const double rr = double(r[e]) * inv_n, ss = double(s[e]) * inv_n, ob = double(io);
      const double d = 1.0 - ob / cl;
      num -= 2.0 * rr * d;
      const double t = 2.0 * rr / cl;
      den += 2.0 * ss * d + ob * t * t;
    }
  }
  const double rn = block_reduce<S::NT>(num, red, AddOp());
  const double rd = block_reduce<S::NT>(den, red, AddOp());
  if (tid == 0) {
    part[((size_t)b * gridDim.x + bxy.x) * 2] = rn;
    part[((size_t)b * gridDim.x + bxy.x) * 2 + 1] = rd;
  }
}

// Sums of pairs in a fixed order: out[0] = Σ part[2i], out[1] = Σ part[2i + 1] (add = 1 adds to
// out instead of writing it).
static __global__ void k_sum_pairs(const double* __restrict__ part, int count, double* __restrict__ out,
                            int add) {
  __shared__ double red[THREADS];
  double a = 0, b = 0;
  for (int i = threadIdx.x; i < count; i += THREADS) {
    a += part[2 * (size_t)i];
    b += part[2 * (size_t)i + 1];
  }
  a = block_reduce<THREADS>(a, red, AddOp());
  b = block_reduce<THREADS>(b, red, AddOp());
  if (threadIdx.x == 0) {
    out[0] = add ? out[0] + a : a;
    out[1] = add ? out[1] + b : b;
  }
}

// Polak–Ribière sums over one array in FP64: part[2·block] = Σ Re(conj(g)·(g − h)),
// part[2·block + 1] = Σ |h|². ptyrs/src/engine/ptycho.rs:785-806
template <typename T>
__global__ __launch_bounds__(THREADS) void k_cg_dots(const cplx<T>* __restrict__ g,
                                                     const cplx<T>* __restrict__ h, size_t total,
                                                     double* __restrict__ part) {
  __shared__ double red[THREADS];
  double a = 0, b = 0;
  for (size_t i = (size_t)blockIdx.x * THREADS + threadIdx.x; i < total; i += (size_t)gridDim.x * THREADS) {
    const cplx<T> gv = ld(g + i), hv = ld(h + i);
    a += double(gv.x) * (double(gv.x) - hv.x) + double(gv.y) * (double(gv.y) - hv.y);
    b += double(hv.x) * hv.x + double(hv.y) * hv.y;
  }
  a = block_reduce<THREADS>(a, red, AddOp());
  b = block_reduce<THREADS>(b, red, AddOp());
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

// d = β·d − g with β = num/den from sums[0..1] (0 when negative, not finite, or first);
// then h = g.
template <typename T>
__global__ void k_cg_direction(cplx<T>* __restrict__ d, cplx<T>* __restrict__ h,
                               const cplx<T>* __restrict__ g, size_t total,
                               const double* __restrict__ sums, int first) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  double beta = 0;
  if (!first) {
    const double b = sums[0] / sums[1];
    beta = (isfinite(b) && b >= 0.0) ? b : 0.0;
  }
  const cplx<T> gv = ld(g + i);
  st(d + i, ld(d + i) * T(beta) - gv);
  st(h + i, gv);
}

// x += γ·d with γ = nd[0]/nd[1]; where γ is not finite, d = −g instead.
template <typename T>
__global__ void k_cg_step(cplx<T>* __restrict__ x, cplx<T>* __restrict__ d,
                          const cplx<T>* __restrict__ g, size_t total,
                          const double* __restrict__ nd) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const double gamma = nd[0] / nd[1];
  if (isfinite(gamma)) {
    st(x + i, ld(x + i) + ld(d + i) * T(gamma));
  } else {
    const cplx<T> gv = ld(g + i);
    st(d + i, cplx<T>{-gv.x, -gv.y});
  }
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
