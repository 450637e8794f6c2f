// pty-chi's BH (bilinear-Hessian method) on the amplitude cost: the kernels of a batch's
// Hessian sums, step and update (ptyrs/src/engine/bh.rs:130-799, the reference GPU's
// ptyrs/src/engine/gpu/bh.rs:399-600).
//
// A batch takes its gradient through the stack passes (pass B's B_BH mode keeps U·ψ and scales
// the field by 2(1 − √iobs/(A + EPS))/n; pass C keeps r = Uᴴ(2·td) per view and sums the probe
// gradient; the object gradient comes from the tile pass). With the displacement u = (g_O,
// ρ²·g_P) and, for conjugate gradient, v = D′ (the previous direction), the Hessian H(a, b) of
// the amplitude cost over the batch is
//   Σ Re(conj(r)·(a_O·b_W + b_O·a_W))                          (second order, exit plane)
//   + 2·Σ_unmasked [(1 − d/A)·Σ_m Re(conj(L_a)·L_b) + d·c_a·c_b/A³]   (detector plane)
// with a_W = a_P·q, L_a = U(W·a_O + a_W·O), c_a = Σ_m Re(conj(U ψ)·L_a), d = √iobs, A =
// √(Σ_m |U ψ|²) + EPS. The first term uses ⟨2·td, U x⟩ = ⟨Uᴴ(2·td), x⟩, so only L_u and L_v
// are transformed. H(u, u), H(u, v) and H(v, v) come from one pass each way; H(D, D) for D =
// β·v − u follows by bilinearity.
#pragma once

#include "kernels.hpp"

namespace ptyrs {

// Pass A of the Hessian sums. Per view row: L_u = W·u_O + su·u_P·q·O and (NF = 2) L_v = W·v_O +
// v_P·q·O, row-transformed into lf[(slot·M + m)·NF + field] (a view's fields side by side);
// the second-order sums over the row with
// r (psi, natural layout): Re Σ conj(r)·2·u_O·u_W, Re Σ conj(r)·(u_O·v_W + v_O·u_W), Re Σ
// conj(r)·2·v_O·v_W (FP64, per workgroup into part[(view·NBLK + blk)·3 + i]). Grid
// (NBLK, nb·M) of HessShape: pass C's line shape (half pass A's values per thread), which keeps
// a thread's loads in flight within its registers.
template <typename T, int N, int NF>
using HessShape = RowShapeC<T, N>;
template <typename T, int N, int NF>
__global__ __launch_bounds__((HessShape<T, N, NF>::NT)) void k_bh_hess_a(
    const cplx<T>* __restrict__ O, const cplx<T>* __restrict__ uO,
    const cplx<T>* __restrict__ vO, int nxo, const cplx<T>* __restrict__ Wt,
    const cplx<T>* __restrict__ uP, T su, const cplx<T>* __restrict__ vP,
    const cplx<T>* __restrict__ q, const int2* __restrict__ corner,
    const int* __restrict__ order, int M, const cplx<T>* __restrict__ r,
    cplx<T>* __restrict__ lf, double* __restrict__ part, const cplx<T>* __restrict__ tw) {
  using S = HessShape<T, N, NF>;
  constexpr int E = S::E, TL = S::TL;
  constexpr int n = N * N;
  __shared__ T lre[S::LDS_WORDS], lim[S::LDS_WORDS];
  __shared__ cplx<T> stw[N];
  __shared__ double red[S::NT];
  const int tid = threadIdx.x, l = tid / TL, jl = tid % TL;
  const int y = blockIdx.x * S::L + l;
  const int v = blockIdx.y, b = v / M, m = v % M;
  stage_twiddles<T, N, S::NT>(stw, tw);
  const int2 c = corner[order[b]];
  const size_t op = (size_t)(c.x + y) * nxo + c.y;
  const size_t pr = (size_t)m * n + (size_t)y * N;
  const size_t row = ((size_t)v * N + y) * N;
  const size_t lrow = ((size_t)v * NF * N + y) * N;
  cplx<T> lu[E], lv[E];
  double s[3] = {0, 0, 0};
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int x = jl + TL * e;
    const cplx<T> o = ld(O + op + x), w = ld(Wt + pr + x), qq = ld(q + (size_t)y * N + x);
    const cplx<T> rr = ld(r + row + x);
    const cplx<T> ao = ld(uO + op + x), aw = ld(uP + pr + x) * qq * su;
    lu[e] = w * ao + aw * o;
    s[0] += 2.0 * double(cmulc(rr, ao * aw).x);
    if constexpr (NF == 2) {
      const cplx<T> bo = ld(vO + op + x), bw = ld(vP + pr + x) * qq;
      lv[e] = w * bo + bw * o;
      s[1] += double(cmulc(rr, ao * bw + bo * aw).x);
      s[2] += 2.0 * double(cmulc(rr, bo * bw).x);
    }
  }
  T* re = lre + S::base_row(l);
  T* im = lim + S::base_row(l);
  line_fft<T, N, E, false, typename S::AtRow>(lu, jl, re, im, stw);
  if constexpr (NF == 2) line_fft<T, N, E, false, typename S::AtRow>(lv, jl, re, im, stw);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    st(lf + lrow + jl + TL * e, lu[e]);
    if constexpr (NF == 2) st(lf + lrow + n + jl + TL * e, lv[e]);
  }
  // (u, v) and (v, v) are 0 with NF = 1.
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double t = block_reduce<S::NT>(s[i], red, AddOp());
    if (tid == 0) part[((size_t)v * gridDim.x + blockIdx.x) * 3 + i] = t;
  }
}

// Pass B of the Hessian sums: per frame and stripe, the column transforms of L_u (and L_v) per
// mode (ROW: the angular spectrum's row transforms, after its column passes), then over the
// unmasked pixels with U ψ from `keep` (already in the detector plane, unitary units), lu = c·L_u
// (c² = `norm`, the unnormalised transforms' factor): A = √(Σ_m |U ψ|²) + EPS, d = √iobs, the
// detector-plane terms 2[(1 − d/A)·Σ_m Re(conj(lu)·lv) + d·cu·cv/A³] for (u, u), (u, v), (v, v)
// (FP64, per workgroup into part[(slot·gridDim.x + blk)·3 + i]). Grid (NBLK, nb) of ColShape.
template <typename T, int N, int M, bool ROW, int NF>
__global__ __launch_bounds__((ColShape<T, N>::NT)) void k_bh_hess_b(
    const cplx<T>* __restrict__ lf, const cplx<T>* __restrict__ keep,
    const float* __restrict__ counts, const int* __restrict__ order, double* __restrict__ part,
    const cplx<T>* __restrict__ tw, T norm) {
  using S = ColShape<T, N>;
  constexpr int E = S::E, TL = S::TL, CG = S::L;
  constexpr int n = N * N;
  constexpr double EPS = 1e-8;
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
  const T c = sqrt(norm);
  double I[E], uu[E], uv[E], vv[E], cu[E], cv[E];
#pragma unroll
  for (int e = 0; e < E; ++e) I[e] = uu[e] = uv[e] = vv[e] = cu[e] = cv[e] = 0;
  for (int m = 0; m < M; ++m) {
    const size_t plane = (size_t)(b * M + m) * n;
    const size_t lplane = plane * NF;
    cplx<T> a[E], bb[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      a[e] = ld(lf + lplane + at(e));
      if constexpr (NF == 2) bb[e] = ld(lf + lplane + n + at(e));
    }
    line_fft<T, N, E, false, At>(a, jl, re, im, stw);
    if constexpr (NF == 2) line_fft<T, N, E, false, At>(bb, jl, re, im, stw);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const cplx<T> p = ld(keep + plane + at(e));
      const cplx<T> la = a[e] * c;
      I[e] += double(norm2(p));
      uu[e] += double(norm2(la));
      cu[e] += double(cmulc(p, la).x);
      if constexpr (NF == 2) {
        const cplx<T> lb = bb[e] * c;
        uv[e] += double(cmulc(la, lb).x);
        vv[e] += double(norm2(lb));
        cv[e] += double(cmulc(p, lb).x);
      }
    }
  }
  double h[3] = {0, 0, 0};
  const float* cnt = counts + (size_t)f * n;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const float io = cnt[at(e)];
    if (io >= 0.f) {
      const double A = sqrt(I[e]) + EPS, d = sqrt(double(io));
      const double w = 1.0 - d / A, k3 = d / (A * A * A);
      h[0] += 2.0 * (w * uu[e] + k3 * cu[e] * cu[e]);
      if constexpr (NF == 2) {
        h[1] += 2.0 * (w * uv[e] + k3 * cu[e] * cv[e]);
        h[2] += 2.0 * (w * vv[e] + k3 * cv[e] * cv[e]);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double t = block_reduce<S::NT>(h[i], red, AddOp());
    if (tid == 0) part[((size_t)b * gridDim.x + bxy.x) * 3 + i] = t;
  }
}

// ---- BH under PATCH position offsets (ptyrs/src/engine/bh.rs:180-577) ---------------------------
// Frame b's probe copy is conj(q)·W_b with W_b = DFT⁻¹(DFT(q·P)·exp(−2πi f·d_b)), so ψ = W_b·O. In
// these "wave units" the displacement u = (u_O, u_P, u_d) changes W_b by A_u = DFT⁻¹((S_u +
// S_P·F_u)·e_b) and to second order (u, v) by B_uv = DFT⁻¹((S_u·F_v + S_v·F_u + S_P·F_u·F_v)·e_b),
// with S_X = DFT(q·X), F_u = −2πi Σ_a u_da·f_a and e_b = exp(−2πi f·d_b).

// g[v] = conj(O[win])·r[v] per view of a batch (slot-local, natural layout). Grid
// (ceil(n/256), nb·M).
template <typename T>
__global__ void k_bh_conj_o(cplx<T>* __restrict__ g, const cplx<T>* __restrict__ r,
                            const cplx<T>* __restrict__ O, int nxo, const int2* __restrict__ corner,
                            const int* __restrict__ order, int N, int M) {
  const int n = N * N;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int v = blockIdx.y;
  const int2 c = corner[order[v / M]];
  const cplx<T> o = ld(O + (size_t)(c.x + j / N) * nxo + c.y + j % N);
  st(g + (size_t)v * n + j, cmulc(o, ld(r + (size_t)v * n + j)));
}

// The offsets' gradient of each frame of a batch, Re Σ_views conj(r)·O·∂_a W_b, from G =
// DFT(conj(O)·r) (unnormalised, [slot·M + m]) by Parseval: Re Σ_m Σ_k conj(G)·S_P[m]·(−2πi f_a)·
// e_b / n into gd[b·2 + a] (FP64). One workgroup per frame. ptyrs/src/engine/bh.rs:290-317
template <typename T>
__global__ __launch_bounds__(THREADS) void k_bh_pos_grad(const cplx<T>* __restrict__ G,
                                                         const cplx<T>* __restrict__ spec,
                                                         const double2* __restrict__ offs,
                                                         const int* __restrict__ order, int N,
                                                         int M, double* __restrict__ gd) {
  __shared__ double red[THREADS];
  const int n = N * N, b = blockIdx.x;
  const double2 d = offs[order[b]];
  double acc[2] = {0, 0};
  for (int m = 0; m < M; ++m)
    for (int j = threadIdx.x; j < n; j += THREADS) {
      const double fy = dft_freq(j / N, N), fx = dft_freq(j % N, N);
      double sn, cs;
      sincos(-2.0 * M_PI * (fy * d.x + fx * d.y), &sn, &cs);
      const cplx<T> g = ld(G + ((size_t)b * M + m) * n + j), p = ld(spec + (size_t)m * n + j);
      // t = conj(g)·p·e_b; Re(t·(−2πi f_a)) = 2π f_a·Im(t)
      const double pr = double(p.x) * cs - double(p.y) * sn, pi = double(p.x) * sn + double(p.y) * cs;
      const double ti = double(g.x) * pi - double(g.y) * pr;
      acc[0] += 2.0 * M_PI * fy * ti;
      acc[1] += 2.0 * M_PI * fx * ti;
    }
  for (int a = 0; a < 2; ++a) {
    const double t = block_reduce<THREADS>(acc[a], red, AddOp());
    if (threadIdx.x == 0) gd[2 * (size_t)b + a] = t / double(n);
  }
}

// The spectra of a batch's wave changes along u and v, NP planes per view ([view·NP + p]):
// A_u, B_uu (NP = 2) or A_u, A_v, B_uu, B_uv, B_vv (NP = 5), each times `scale`; S_u is the
// spectrum of q·g_P and u_P = su·g_P, u_d = pos2·gd[slot] (gd null: 0), v_d = ddf[frame] (ddf
// null: 0). Grid (ceil(n/256), nb·M).
template <typename T, int NP>
__global__ void k_bh_moved_planes(cplx<T>* __restrict__ out, const cplx<T>* __restrict__ sp,
                                  const cplx<T>* __restrict__ su_spec, double su,
                                  const cplx<T>* __restrict__ sv_spec,
                                  const double2* __restrict__ offs, const int* __restrict__ order,
                                  const double* __restrict__ gd, double pos2,
                                  const double* __restrict__ ddf, int N, int M, double scale) {
  const int n = N * N;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int v = blockIdx.y, b = v / M, m = v % M, f = order[b];
  const double2 d = offs[f];
  const double fy = dft_freq(j / N, N), fx = dft_freq(j % N, N);
  double sn, cs;
  sincos(-2.0 * M_PI * (fy * d.x + fx * d.y), &sn, &cs);
  // F = i·φ, φ = −2π Σ_a d_a·f_a
  const double pu = gd ? -2.0 * M_PI * pos2 * (gd[2 * b] * fy + gd[2 * b + 1] * fx) : 0.0;
  const double pv = ddf ? -2.0 * M_PI * (ddf[2 * f] * fy + ddf[2 * f + 1] * fx) : 0.0;
  const size_t at = (size_t)m * n + j;
  const cplx<T> P = ld(sp + at), U = ld(su_spec + at);
  const double Pr = P.x, Pi = P.y, Ur = su * U.x, Ui = su * U.y;
  // x·e_b·scale for x = (re, im)
  auto put = [&](int p, double re, double im) {
    st(out + ((size_t)v * NP + p) * n + j,
       cplx<T>{T((re * cs - im * sn) * scale), T((re * sn + im * cs) * scale)});
  };
  // A_u = U + i·pu·P; B_uu = 2i·pu·U − pu²·P
  put(0, Ur - pu * Pi, Ui + pu * Pr);
  if constexpr (NP == 2) {
    put(1, -2.0 * pu * Ui - pu * pu * Pr, 2.0 * pu * Ur - pu * pu * Pi);
  } else {
    const cplx<T> V = ld(sv_spec + at);
    const double Vr = V.x, Vi = V.y;
    put(1, Vr - pv * Pi, Vi + pv * Pr);                                         // A_v
    put(2, -2.0 * pu * Ui - pu * pu * Pr, 2.0 * pu * Ur - pu * pu * Pi);        // B_uu
    // B_uv = i·pv·U + i·pu·V − pu·pv·P
    put(3, -pv * Ui - pu * Vi - pu * pv * Pr, pv * Ur + pu * Vr - pu * pv * Pi);
    put(4, -2.0 * pv * Vi - pv * pv * Pr, 2.0 * pv * Vr - pv * pv * Pi);        // B_vv
  }
}

// Pass A of the Hessian sums under offsets: as k_bh_hess_a with every view's own waves: W from
// w ([view][n]), the changes A_u, A_v and B_uu, B_uv, B_vv from pl ([view·NP + p][n], natural):
// L_u = W·u_O + A_u·O, L_v = W·v_O + A_v·O; second-order sums Re Σ conj(r)·(2·u_O·A_u + O·B_uu),
// Re Σ conj(r)·(u_O·A_v + v_O·A_u + O·B_uv), Re Σ conj(r)·(2·v_O·A_v + O·B_vv).
template <typename T, int N, int NF>
__global__ __launch_bounds__((HessShape<T, N, NF>::NT)) void k_bh_hess_a_moved(
    const cplx<T>* __restrict__ O, const cplx<T>* __restrict__ uO,
    const cplx<T>* __restrict__ vO, int nxo, const cplx<T>* __restrict__ w,
    const cplx<T>* __restrict__ pl, const int2* __restrict__ corner,
    const int* __restrict__ order, int M, const cplx<T>* __restrict__ r,
    cplx<T>* __restrict__ lf, double* __restrict__ part, const cplx<T>* __restrict__ tw) {
  using S = HessShape<T, N, NF>;
  constexpr int E = S::E, TL = S::TL;
  constexpr int n = N * N;
  constexpr int NP = NF == 2 ? 5 : 2;
  __shared__ T lre[S::LDS_WORDS], lim[S::LDS_WORDS];
  __shared__ cplx<T> stw[N];
  __shared__ double red[S::NT];
  const int tid = threadIdx.x, l = tid / TL, jl = tid % TL;
  const int y = blockIdx.x * S::L + l;
  const int v = blockIdx.y, b = v / M;
  stage_twiddles<T, N, S::NT>(stw, tw);
  const int2 c = corner[order[b]];
  const size_t op = (size_t)(c.x + y) * nxo + c.y;
  const size_t row = ((size_t)v * N + y) * N;
  const size_t prow = ((size_t)v * NP * N + y) * N;
  const size_t lrow = ((size_t)v * NF * N + y) * N;
  cplx<T> lu[E], lv[E];
  double s[3] = {0, 0, 0};
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int x = jl + TL * e;
    const cplx<T> o = ld(O + op + x), ww = ld(w + row + x), rr = ld(r + row + x);
    const cplx<T> ao = ld(uO + op + x), au = ld(pl + prow + x);
    lu[e] = ww * ao + au * o;
    if constexpr (NF == 1) {
      const cplx<T> buu = ld(pl + prow + n + x);
      s[0] += double(cmulc(rr, ao * au * T(2) + o * buu).x);
    } else {
      const cplx<T> av = ld(pl + prow + n + x), bo = ld(vO + op + x);
      const cplx<T> buu = ld(pl + prow + 2 * n + x), buv = ld(pl + prow + 3 * n + x),
                    bvv = ld(pl + prow + 4 * n + x);
      lv[e] = ww * bo + av * o;
      s[0] += double(cmulc(rr, ao * au * T(2) + o * buu).x);
      s[1] += double(cmulc(rr, ao * av + bo * au + o * buv).x);
      s[2] += double(cmulc(rr, bo * av * T(2) + o * bvv).x);
    }
  }
  T* re = lre + S::base_row(l);
  T* im = lim + S::base_row(l);
  line_fft<T, N, E, false, typename S::AtRow>(lu, jl, re, im, stw);
  if constexpr (NF == 2) line_fft<T, N, E, false, typename S::AtRow>(lv, jl, re, im, stw);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    st(lf + lrow + jl + TL * e, lu[e]);
    if constexpr (NF == 2) st(lf + lrow + n + jl + TL * e, lv[e]);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double t = block_reduce<S::NT>(s[i], red, AddOp());
    if (tid == 0) part[((size_t)v * gridDim.x + blockIdx.x) * 3 + i] = t;
  }
}

// ---- BH under FullObject sampling (ptyrs/src/engine/bh.rs:96-107, :380-577) -------------------
// Every slot reads S(O), the object sampled at its corner plus offset (one N × N plane per slot,
// natural layout); a displacement changes the reads by u′ = S(u_O) + Σ_a u_da·∂_a S(O) and, to
// second order along u and v, by o″ = Σ_a (v_da·∂_a S(u_O) + u_da·∂_a S(v_O)) +
// Σ_ac u_da·v_dc·∂_a ∂_c S(O).

// Per slot Re Σ conj(g)·d_a over the slot's plane for a = y, x: g the reads' gradient
// ([slot][n]), d the reads' derivatives ([slot·2 + a][n]); into gd[slot·2 + a] (FP64). One
// workgroup per slot.
template <typename T>
__global__ __launch_bounds__(THREADS) void k_bh_win_dot(const cplx<T>* __restrict__ g,
                                                        const cplx<T>* __restrict__ d, int n,
                                                        double* __restrict__ gd,
                                                        int accumulate = 0) {
  __shared__ double red[THREADS];
  const int b = blockIdx.x;
  double acc[2] = {0, 0};
  for (int j = threadIdx.x; j < n; j += THREADS) {
    const cplx<T> gv = ld(g + (size_t)b * n + j);
    for (int a = 0; a < 2; ++a) acc[a] += double(cmulc(gv, ld(d + ((size_t)b * 2 + a) * n + j)).x);
  }
  for (int a = 0; a < 2; ++a) {
    const double t = block_reduce<THREADS>(acc[a], red, AddOp());
    if (threadIdx.x == 0) gd[2 * (size_t)b + a] = accumulate ? gd[2 * (size_t)b + a] + t : t;
  }
}

// The reads' changes per slot: out[slot·NR + p] = u′ (p 0), v′ (1, NF = 2), o″_uu, o″_uv, o″_vv
// (NR = 5; NF = 1: u′, o″_uu, NR = 2), from w[slot·9 + p] = S(u_O), S(v_O), ∂_y S(u_O),
// ∂_x S(u_O), ∂_y S(v_O), ∂_x S(v_O), ∂_yy S(O), ∂_yx S(O), ∂_xx S(O) and d[slot·2 + a] =
// ∂_a S(O); u_d = pos2·gd[slot], v_d = ddf[frame] (gd null: held offsets, no offset terms).
// Grid (ceil(n/256), nb).
template <typename T, int NF>
__global__ void k_bh_fo_reads(cplx<T>* __restrict__ out, const cplx<T>* __restrict__ w,
                              const cplx<T>* __restrict__ d, const double* __restrict__ gd,
                              double pos2, const double* __restrict__ ddf,
                              const int* __restrict__ order, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int b = blockIdx.y;
  constexpr int NR = NF == 2 ? 5 : 2;
  auto at = [&](int p) { return ld(w + ((size_t)b * 9 + p) * n + j); };
  const cplx<T> su = at(0);
  cplx<T>* o = out + (size_t)b * NR * n + j;
  if (!gd) {
    st(o, su);
    if constexpr (NF == 2) {
      st(o + n, at(1));
      st(o + 2 * n, cplx<T>{T(0), T(0)});
      st(o + 3 * n, cplx<T>{T(0), T(0)});
      st(o + 4 * n, cplx<T>{T(0), T(0)});
    } else {
      st(o + n, cplx<T>{T(0), T(0)});
    }
    return;
  }
  const T u0 = T(pos2 * gd[2 * b]), u1 = T(pos2 * gd[2 * b + 1]);
  const cplx<T> dy = ld(d + ((size_t)b * 2) * n + j), dx = ld(d + ((size_t)b * 2 + 1) * n + j);
  const cplx<T> dyy = at(6), dyx = at(7), dxx = at(8);
  // Σ_ac a_a·c_c·∂_a∂_c S(O)
  auto dd = [&](T a0, T a1, T c0, T c1) { return dyy * (a0 * c0) + dyx * (a0 * c1 + a1 * c0) + dxx * (a1 * c1); };
  st(o, su + dy * u0 + dx * u1);
  const cplx<T> uy = at(2), ux = at(3);
  if constexpr (NF == 1) {
    st(o + n, (uy * u0 + ux * u1) * T(2) + dd(u0, u1, u0, u1));
  } else {
    const int f = order[b];
    const T v0 = T(ddf[2 * f]), v1 = T(ddf[2 * f + 1]);
    const cplx<T> vy = at(4), vx = at(5);
    st(o + n, at(1) + dy * v0 + dx * v1);
    st(o + 2 * n, (uy * u0 + ux * u1) * T(2) + dd(u0, u1, u0, u1));
    st(o + 3 * n, uy * v0 + ux * v1 + vy * u0 + vx * u1 + dd(u0, u1, v0, v1));
    st(o + 4 * n, (vy * v0 + vx * v1) * T(2) + dd(v0, v1, v0, v1));
  }
}

// Pass A of the Hessian sums under FullObject: as k_bh_hess_a with the reads per slot: S(O) from
// win ([slot][n]) and u′, v′, o″_uu, o″_uv, o″_vv from rd ([slot·NR + p][n]): L_u = W·u′ +
// u_W·S(O), L_v = W·v′ + v_W·S(O) (u_W = su·u_P·q, v_W = v_P·q); second-order sums
// Re Σ conj(r)·(2·u′·u_W + W·o″_uu), Re Σ conj(r)·(u′·v_W + v′·u_W + W·o″_uv),
// Re Σ conj(r)·(2·v′·v_W + W·o″_vv).
template <typename T, int N, int NF>
__global__ __launch_bounds__((HessShape<T, N, NF>::NT)) void k_bh_hess_a_fo(
    const cplx<T>* __restrict__ win, const cplx<T>* __restrict__ rd,
    const cplx<T>* __restrict__ Wt, const cplx<T>* __restrict__ uP, T su,
    const cplx<T>* __restrict__ vP, const cplx<T>* __restrict__ q, int M,
    const cplx<T>* __restrict__ r, cplx<T>* __restrict__ lf, double* __restrict__ part,
    const cplx<T>* __restrict__ tw) {
  using S = HessShape<T, N, NF>;
  constexpr int E = S::E, TL = S::TL;
  constexpr int n = N * N;
  constexpr int NR = NF == 2 ? 5 : 2;
  __shared__ T lre[S::LDS_WORDS], lim[S::LDS_WORDS];
  __shared__ cplx<T> stw[N];
  __shared__ double red[S::NT];
  const int tid = threadIdx.x, l = tid / TL, jl = tid % TL;
  const int y = blockIdx.x * S::L + l;
  const int v = blockIdx.y, b = v / M, m = v % M;
  stage_twiddles<T, N, S::NT>(stw, tw);
  const size_t wrow = ((size_t)b * N + y) * N;
  const size_t rrow = ((size_t)b * NR * N + y) * N;
  const size_t pr = (size_t)m * n + (size_t)y * N;
  const size_t row = ((size_t)v * N + y) * N;
  const size_t lrow = ((size_t)v * NF * N + y) * N;
  cplx<T> lu[E], lv[E];
  double s[3] = {0, 0, 0};
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int x = jl + TL * e;
    const cplx<T> o = ld(win + wrow + x), w = ld(Wt + pr + x), qq = ld(q + (size_t)y * N + x);
    const cplx<T> rr = ld(r + row + x);
    const cplx<T> ao = ld(rd + rrow + x), aw = ld(uP + pr + x) * qq * su;
    lu[e] = w * ao + aw * o;
    if constexpr (NF == 1) {
      const cplx<T> ouu = ld(rd + rrow + n + x);
      s[0] += double(cmulc(rr, ao * aw * T(2) + w * ouu).x);
    } else {
      const cplx<T> bo = ld(rd + rrow + n + x), bw = ld(vP + pr + x) * qq;
      const cplx<T> ouu = ld(rd + rrow + 2 * n + x), ouv = ld(rd + rrow + 3 * n + x),
                    ovv = ld(rd + rrow + 4 * n + x);
      lv[e] = w * bo + bw * o;
      s[0] += double(cmulc(rr, ao * aw * T(2) + w * ouu).x);
      s[1] += double(cmulc(rr, ao * bw + bo * aw + w * ouv).x);
      s[2] += double(cmulc(rr, bo * bw * T(2) + w * ovv).x);
    }
  }
  T* re = lre + S::base_row(l);
  T* im = lim + S::base_row(l);
  line_fft<T, N, E, false, typename S::AtRow>(lu, jl, re, im, stw);
  if constexpr (NF == 2) line_fft<T, N, E, false, typename S::AtRow>(lv, jl, re, im, stw);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    st(lf + lrow + jl + TL * e, lu[e]);
    if constexpr (NF == 2) st(lf + lrow + n + jl + TL * e, lv[e]);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double t = block_reduce<S::NT>(s[i], red, AddOp());
    if (tid == 0) part[((size_t)v * gridDim.x + blockIdx.x) * 3 + i] = t;
  }
}

// ---- BH through several object slices (ptyrs/src/engine/multislice.rs:318-383) ---------------
// One slice of the Hessian's tangent chains: per view the changes of the wave leaving slice s
// along u and (NF = 2) v, ou = q_s·(O_s·w_u + u_s·in) and ov = q_s·(O_s·w_v + v_s·in), written
// at out[(view·NF + f)·n] (over w itself, in place), and the slice's second-order terms against
// χ_s, the residual carried back to slice s's output: Re Σ conj(χ_s)·q_s·2·u_s·w_u,
// Re Σ conj(χ_s)·q_s·(u_s·w_v + v_s·w_u), Re Σ conj(χ_s)·q_s·2·v_s·w_v (FP64, per workgroup
// into part[(view·gridDim.x + blk)·3 + i]); ⟨r, J″(u, v)⟩ is their sum over the slices. At
// slice 0 (w_mode) w_u = su·u_P[m] and w_v = v_P[m]; `in` is P[m] there (in_view 0), else the
// views' incident waves. With `rd` (FullObject with refined offsets) the slot's reads' changes
// come from rd[slot·NR + p] (k_bh_fo_reads' layout: u′, v′, o″_uu, o″_uv, o″_vv; NF = 1: u′,
// o″_uu) in place of uO and vO, and the second-order terms add Re Σ conj(χ_s)·q_s·o″·in.
// Grid (ceil(n/256), nb·M).
template <typename T, int NF>
__global__ __launch_bounds__(256) void k_bh_slice_tangent(
    cplx<T>* out, const cplx<T>* w, const cplx<T>* __restrict__ uP, T su,
    const cplx<T>* __restrict__ vP, int w_mode, const cplx<T>* __restrict__ in, int in_view,
    const cplx<T>* __restrict__ O, const cplx<T>* __restrict__ uO,
    const cplx<T>* __restrict__ vO, int nxo, const cplx<T>* __restrict__ q,
    const cplx<T>* __restrict__ chi, const int2* __restrict__ corner,
    const int* __restrict__ order, int N, int M, double* __restrict__ part,
    const cplx<T>* __restrict__ rd = nullptr) {
  __shared__ double red[256];
  constexpr int NR = NF == 2 ? 5 : 2;
  const int n = N * N;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int v = blockIdx.y, b = v / M, m = v % M;
  double s[3] = {0, 0, 0};
  if (j < n) {
    const int2 c = corner[order[b]];
    const size_t at = (size_t)(c.x + j / N) * nxo + c.y + j % N;
    const cplx<T>* r = rd ? rd + (size_t)b * NR * n + j : nullptr;
    const cplx<T> o = ld(O + at), ao = r ? ld(r) : ld(uO + at);
    const cplx<T> qq = q ? ld(q + j) : cplx<T>{T(1), T(0)};
    const cplx<T> iw = ld(in + (size_t)(in_view ? v : m) * n + j);
    const cplx<T> cq = cmulc(qq, ld(chi + (size_t)v * n + j));  // conj(q)·χ
    const cplx<T> wu = w_mode ? ld(uP + (size_t)m * n + j) * su : ld(w + ((size_t)v * NF) * n + j);
    s[0] = 2.0 * double(cmulc(cq, ao * wu).x);
    if (r) s[0] += double(cmulc(cq, ld(r + (NF == 2 ? 2 : 1) * n) * iw).x);
    const cplx<T> ou = qq * (o * wu + ao * iw);
    if constexpr (NF == 2) {
      const cplx<T> bo = r ? ld(r + n) : ld(vO + at);
      const cplx<T> wv = w_mode ? ld(vP + (size_t)m * n + j) : ld(w + ((size_t)v * NF + 1) * n + j);
      s[1] = double(cmulc(cq, ao * wv + bo * wu).x);
      s[2] = 2.0 * double(cmulc(cq, bo * wv).x);
      if (r) {
        s[1] += double(cmulc(cq, ld(r + 3 * n) * iw).x);
        s[2] += double(cmulc(cq, ld(r + 4 * n) * iw).x);
      }
      st(out + ((size_t)v * NF + 1) * n + j, qq * (o * wv + bo * iw));
    }
    st(out + ((size_t)v * NF) * n + j, ou);
  }
  for (int i = 0; i < 3; ++i) {
    const double t = block_reduce<256>(s[i], red, AddOp());
    if (threadIdx.x == 0) part[((size_t)v * gridDim.x + blockIdx.x) * 3 + i] = t;
  }
}

// The refined offsets' step of a batch: D_d = β·D′_d[frame] (cg) − pos2·gd[slot]; offsets +=
// 0.3·α·D_d; with `keep` D′_d[frame] = D_d. ptyrs/src/engine/bh.rs:778-787
__global__ static void k_bh_update_pos(double2* __restrict__ offs, double* __restrict__ ddf,
                                       const double* __restrict__ gd,
                                       const int* __restrict__ order, int nb, double pos2,
                                       const double* __restrict__ step, int cg, int keep) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int f = order[b];
  const double alpha = step[0], beta = step[1];
  double dy = -pos2 * gd[2 * b], dx = -pos2 * gd[2 * b + 1];
  if (cg) {
    dy += beta * ddf[2 * f];
    dx += beta * ddf[2 * f + 1];
  }
  double2 o = offs[f];
  o.x += 0.3 * alpha * dy;
  o.y += 0.3 * alpha * dx;
  offs[f] = o;
  if (keep) {
    ddf[2 * f] = dy;
    ddf[2 * f + 1] = dx;
  }
}

// Per workgroup over the object (no values) and the probe (np): Σ|g_O|² + ρ²·Σ|g_P|² = ⟨g, Pg⟩
// and (cg) Σ Re(conj(g)·D′) = ⟨g, D′⟩ into part[blk·2 + {0, 1}] (FP64).
template <typename T>
__global__ __launch_bounds__(THREADS) void k_bh_dots(
    const cplx<T>* __restrict__ gO, const cplx<T>* __restrict__ dO, size_t no,
    const cplx<T>* __restrict__ gP, const cplx<T>* __restrict__ dP, size_t np, double rho2,
    int cg, double* __restrict__ part) {
  __shared__ double red[THREADS];
  double a = 0, b = 0;
  const size_t total = no + np;
  for (size_t i = (size_t)blockIdx.x * THREADS + threadIdx.x; i < total;
       i += (size_t)gridDim.x * THREADS) {
    const bool obj = i < no;
    const cplx<T> g = obj ? ld(gO + i) : ld(gP + (i - no));
    const double gg = double(g.x) * g.x + double(g.y) * g.y;
    a += obj ? gg : rho2 * gg;
    if (cg) {
      const cplx<T> d = obj ? ld(dO + i) : ld(dP + (i - no));
      b += double(g.x) * d.x + double(g.y) * d.y;
    }
  }
  a = block_reduce<THREADS>(a, red, AddOp());
  b = block_reduce<THREADS>(b, red, AddOp());
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

// Per workgroup the sums of rows [blk·per, blk·per + per) of a [count][3] array (FP64, a fixed
// order for a fixed count) into out[blk][3]: the first level of the Hessian sums' reduction.
__global__ __launch_bounds__(THREADS) static void k_sum_rows3(const double* __restrict__ in,
                                                              int count, int per,
                                                              double* __restrict__ out) {
  __shared__ double red[THREADS];
  const int i0 = blockIdx.x * per, i1 = i0 + per < count ? i0 + per : count;
  double s[3] = {0, 0, 0};
  for (int i = i0 + threadIdx.x; i < i1; i += THREADS)
    for (int k = 0; k < 3; ++k) s[k] += in[3 * (size_t)i + k];
  for (int k = 0; k < 3; ++k) {
    const double t = block_reduce<THREADS>(s[k], red, AddOp());
    if (threadIdx.x == 0) out[3 * (size_t)blockIdx.x + k] = t;
  }
}

// One workgroup: the batch's step from the partial sums, in a fixed order. H(u, u), H(u, v),
// H(v, v) = second-order (npa partials of 3) + detector-plane (npb partials of 3) terms; β =
// H(u, v)/H(v, v) with cg, else 0; top = ⟨g, Pg⟩ − β·⟨g, D′⟩ (= −⟨g, D⟩ for D = β·D′ − Pg);
// bottom = H(D, D) = β²·H(v, v) − 2β·H(u, v) + H(u, u); α = top/bottom·count/size. step =
// {α, β}; log[0..4) = {β, α, top, bottom}. ptyrs/src/engine/bh.rs:729-760
// Refined offsets (npos = the batch's frames, gd their gradient, ddf D′_d by frame) add
// pos2·Σ|g_d|² and Σ g_d·D′_d to the two dots.
__global__ __launch_bounds__(THREADS) static void k_bh_step(
    const double* __restrict__ dots, int ndots, const double* __restrict__ pa, int npa,
    const double* __restrict__ pb, int npb, int cg, double frac, double* __restrict__ step,
    double* __restrict__ log, const double* __restrict__ gd = nullptr,
    const double* __restrict__ ddf = nullptr, const int* __restrict__ order = nullptr,
    int npos = 0, double pos2 = 0.0) {
  __shared__ double red[THREADS];
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < ndots; i += THREADS) {
    s[0] += dots[2 * (size_t)i];
    s[1] += dots[2 * (size_t)i + 1];
  }
  for (int b = threadIdx.x; b < npos; b += THREADS) {
    const double gy = gd[2 * b], gx = gd[2 * b + 1];
    s[0] += pos2 * (gy * gy + gx * gx);
    if (cg) {
      const int f = order[b];
      s[1] += gy * ddf[2 * f] + gx * ddf[2 * f + 1];
    }
  }
  for (int i = threadIdx.x; i < npa; i += THREADS)
    for (int k = 0; k < 3; ++k) s[2 + k] += pa[3 * (size_t)i + k];
  for (int i = threadIdx.x; i < npb; i += THREADS)
    for (int k = 0; k < 3; ++k) s[5 + k] += pb[3 * (size_t)i + k];
  double t[8];
  for (int k = 0; k < 8; ++k) t[k] = block_reduce<THREADS>(s[k], red, AddOp());
  if (threadIdx.x != 0) return;
  const double huu = t[2] + t[5], huv = t[3] + t[6], hvv = t[4] + t[7];
  const double beta = cg ? huv / hvv : 0.0;
  const double top = t[0] - beta * t[1];
  const double bottom = cg ? beta * beta * hvv - 2.0 * beta * huv + huu : huu;
  const double alpha = top / bottom * frac;
  step[0] = alpha;
  step[1] = beta;
  log[0] = beta;
  log[1] = alpha;
  log[2] = top;
  log[3] = bottom;
}

// D = β·D′ − s·g over the object (s = 1) then the probe (s = ρ²); X += α·D; with `keep` D′ = D.
template <typename T>
__global__ void k_bh_update(cplx<T>* __restrict__ O, cplx<T>* __restrict__ dO,
                            const cplx<T>* __restrict__ gO, size_t no, cplx<T>* __restrict__ P,
                            cplx<T>* __restrict__ dP, const cplx<T>* __restrict__ gP, size_t np,
                            double rho2, const double* __restrict__ step, int cg, int keep) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= no + np) return;
  const double alpha = step[0], beta = step[1];
  const bool obj = i < no;
  const size_t j = obj ? i : i - no;
  cplx<T>* X = obj ? O : P;
  cplx<T>* D = obj ? dO : dP;
  const cplx<T> g = ld((obj ? gO : gP) + j);
  const double s = obj ? 1.0 : rho2;
  double dr = -s * double(g.x), di = -s * double(g.y);
  if (cg) {
    const cplx<T> d0 = ld(D + j);
    dr += beta * double(d0.x);
    di += beta * double(d0.y);
  }
  const cplx<T> x = ld(X + j);
  st(X + j, cplx<T>{T(double(x.x) + alpha * dr), T(double(x.y) + alpha * di)});
  if (keep) st(D + j, cplx<T>{T(dr), T(di)});
}

}  // namespace ptyrs
