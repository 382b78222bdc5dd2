// Adjoint (vector-Jacobian product) of the stabilised step, FP64.
//
// With real operators D_x, D_y (spectral derivative, Nyquist zeroed,
// D^T = -D), Lambda = IFFT K^2 FFT, P_den = IFFT (1/den) FFT and M =
// diag(M0 h(phi)), the step of core_shell.jl:72-76 is
//   c' = c + dt P_den (D_x M D_x + D_y M D_y) mu(c),  mu = f'(c) + 2 kappa Lambda c,
// so for gbar = dL/dc'
//   v = P_den gbar,  r_v = (D_x M D_x + D_y M D_y) v,
//   dL/dc = gbar + dt (f''(c) r_v + 2 kappa Lambda r_v),
// with per-step parameter terms
//   dL/dkappa += dt <r_v, 2 Lambda c> + dt <gbar, IFFT(-2Q K^4 dc^ / den)>,
//   dL/dP += -dt chi <r_v, g>,  dL/dchi += -dt P <r_v, g>,
// where dc^ = c^_{s+1} - c^_s (= dt rate^ / den of the forward step).
// Passes (gbar kept in the state's packed line layout, like S):
//   A1 column: gbar, S_{s+1} - S_s -> FFT_j -> V = gbar^/den, kappa-den term
//              -> Vx = i kdx V, Vy = i kdy V -> IFFT_j
//   A2 row:    the forward P3 (J = M D v per component)
//   A3 column: Jx, Jy, S_s -> FFT_j -> r_v^ = i kdx Jx^ + i kdy Jy^,
//              kappa-mu term -> IFFT_j: r_v (row spectra), gbar += dt 2 kappa Lambda r_v
//   A4 row:    r_v, S_s -> C2R -> w = dt f''(c) r_v, <r_v, g> -> R2C -> gbar += w
// Per-workgroup partials accumulate over the whole backward sweep and are
// reduced once at the end.
#pragma once
#include "csg_kernels.hpp"

namespace csg {

struct AdjArgs {
  cplx<double>* G;          // adjoint, packed line layout (A1 in, A3/A4 in-out)
  const cplx<double>* S0;   // state s
  const cplx<double>* S1;   // state s+1 (A1)
  cplx<double>* X0;         // A1 out Vx / A3 in Jx
  cplx<double>* X1;         // A1 out Vy / A3 in Jy
  cplx<double>* Xr;         // r_v row spectra (A3 out, A4 in)
  const double* phi;
  double* acc;              // per-workgroup accumulators (added to)
  double dt;
  DevParams p;
};

// Register tile of the adjoint passes: 4 elements per thread for power-of-two
// lines of 64 .. 1024 (the fastest tile of the forward passes there), the
// default otherwise.
constexpr int adj_E(int len) { return (len % 3 != 0 && len >= 64 && len <= 1024) ? 4 : fft_E(len); }

// Exchange layout of the adjoint FFTs (XM of fft_regs): two buffer sets, one
// ordering point per stage (ADJ_XM = 2), stepped down to one buffer per array
// (1) and to one shared buffer (0) while the buffers of `na` arrays over
// `lines_lp` complex doubles exceed ADJ_XM_LDS bytes.
constexpr int ADJ_XM = 2;
constexpr int ADJ_XM_LDS = 64 * 1024;
constexpr int adj_xm(int na, int lines_lp) {
  int xm = ADJ_XM;
  while (xm > 0 && (xm == 1 ? na : 2 * na) * lines_lp * 16 > ADJ_XM_LDS) --xm;
  return xm;
}

template <int N>
struct AdjColCfg {
  static constexpr int E = adj_E(N);
  static constexpr int TPL = N / E;
  static constexpr int C = col_lines<N, double, E>();
  static constexpr int LP = padded_len<double>(N);
  static constexpr int NT = C * TPL;
  static constexpr bool WAVE = TPL <= 64;
  static constexpr int GRID = (N / 2) / C;
  // up to three arrays per transform (adj_col3), at stride AS, second set at BS
  static constexpr int XM = adj_xm(3, C * LP);
  static constexpr int AS = XM == 0 ? 0 : C * LP, BS = 3 * C * LP;
  static constexpr int LDS = (XM == 0 ? 1 : XM == 1 ? 3 : 6) * C * LP;
};

// Stage the values at -ky of `src` through the line buffer (packed line 0).
template <int N, int E, int TPL, bool WAVE>
__device__ __forceinline__ void partner_of(const cplx<double> (&src)[E], cplx<double> (&dst)[E], int t,
                                           cplx<double>* line) {
  lds_sync<WAVE>();
#pragma unroll
  for (int j = 0; j < E; ++j) line[pad<double, N>(t + TPL * j)] = src[j];
  lds_sync<WAVE>();
#pragma unroll
  for (int j = 0; j < E; ++j) dst[j] = line[pad<double, N>(neg_index<N>(t + TPL * j))];
  lds_sync<WAVE>();
}

// Unpack the packed line 0: A part (k_x = 0) and B part (k_x = N/2) at ky
// from Z[ky] and Z[-ky].
__device__ __forceinline__ void unpack0(cplx<double> z, cplx<double> zn, cplx<double>& a, cplx<double>& b) {
  a = 0.5 * (z + conj(zn));
  b = 0.5 * mul_mi(z - conj(zn));
}
__device__ __forceinline__ double re_conj_mul(cplx<double> x, cplx<double> y) { return x.x * y.x + x.y * y.y; }

template <int N>
__global__ void __launch_bounds__((AdjColCfg<N>::NT))
adj_col1(AdjArgs a, const cplx<double>* __restrict__ tw_g) {
  using Cfg = AdjColCfg<N>;
  constexpr int E = Cfg::E, TPL = Cfg::TPL, LP = Cfg::LP, NT = Cfg::NT;
  constexpr bool WAVE = Cfg::WAVE;
  __shared__ cplx<double> lds[Cfg::LDS];
  __shared__ cplx<double> tw_s[tw_lds_len<N, double>()];
  __shared__ double scratch[8 * ((NT + 63) / 64)];
  const cplx<double>* tw = tw_in_lds<N, double>() ? tw_s : tw_g;
  const int l = threadIdx.x / TPL, t = threadIdx.x % TPL;
  const int kx = blockIdx.x * Cfg::C + l;
  cplx<double>* line = lds + l * LP;
  const DevParams& p = a.p;
  const double inv_nsp = p.inv_nsp, dt = a.dt;
  const double kN2 = kval<N>(N / 2, inv_nsp) * kval<N>(N / 2, inv_nsp);
  const double q2 = 2.0 * p.Q * p.kappa;

  cplx<double> v[2][E];
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const size_t o = lidx<N>(kx, t + TPL * j);
    v[0][j] = a.G[o];
    v[1][j] = a.S1[o] - a.S0[o];
  }
  if constexpr (tw_in_lds<N, double>()) load_twiddles<N, NT>(tw_s, tw_g);
  __syncthreads();
  fft_regs<double, -1, N, 2, N, WAVE, E, Cfg::XM, Cfg::AS, Cfg::BS>(v, t, line, tw);
  cplx<double> gn[E], dn[E];
  if (kx == 0) {
    partner_of<N, E, TPL, WAVE>(v[0], gn, t, line);
    partner_of<N, E, TPL, WAVE>(v[1], dn, t, line);
  }
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int ky = t + TPL * j;
    const double kyv = kval<N>(ky, inv_nsp), ky2 = kyv * kyv;
    const cplx<double> g = v[0][j], d = v[1][j];
    cplx<double> V;
    if (kx == 0) {
      const double k2a = ky2, k2b = kN2 + ky2;
      const double da = 1.0 + q2 * (k2a * k2a) * dt, db = 1.0 + q2 * (k2b * k2b) * dt;
      const double ia = 1.0 / da, ib = 1.0 / db;
      V = (0.5 * (ia + ib)) * g + (0.5 * (ia - ib)) * conj(gn[j]);
      cplx<double> gA, gB, dA, dB;
      unpack0(g, gn[j], gA, gB);
      unpack0(d, dn[j], dA, dB);
      acc += (-2.0 * p.Q * k2a * k2a / da) * re_conj_mul(gA, dA) + (-2.0 * p.Q * k2b * k2b / db) * re_conj_mul(gB, dB);
    } else {
      const double kxv = kval<N>(kx, inv_nsp);
      const double k2 = kxv * kxv + ky2;
      const double den = 1.0 + q2 * (k2 * k2) * dt;
      V = (1.0 / den) * g;
      acc += 2.0 * (-2.0 * p.Q * k2 * k2 / den) * re_conj_mul(g, d);
    }
    const double sx = dkval<N>(kx, inv_nsp) / N, sy = dkval<N>(ky, inv_nsp) / N;
    v[0][j] = mul_i(sx * V);
    v[1][j] = mul_i(sy * V);
  }
  pin_partial(acc);
  fft_regs<double, +1, N, 2, N, WAVE, E, Cfg::XM, Cfg::AS, Cfg::BS>(v, t, line, tw);
#pragma unroll
  for (int j = 0; j < E; ++j) {
    a.X0[lidx<N>(kx, t + TPL * j)] = v[0][j];
    a.X1[lidx<N>(kx, t + TPL * j)] = v[1][j];
  }
  double r[1] = {acc};
  const int ops[1] = {RED_SUM};
  block_reduce0<NT, 1>(r, ops, scratch);
  if (threadIdx.x == 0) a.acc[blockIdx.x] += r[0] * dt / ((double)N * N);
}

template <int N>
__global__ void __launch_bounds__((AdjColCfg<N>::NT))
adj_col3(AdjArgs a, const cplx<double>* __restrict__ tw_g) {
  using Cfg = AdjColCfg<N>;
  constexpr int E = Cfg::E, TPL = Cfg::TPL, LP = Cfg::LP, NT = Cfg::NT;
  constexpr bool WAVE = Cfg::WAVE;
  __shared__ cplx<double> lds[Cfg::LDS];
  __shared__ cplx<double> tw_s[tw_lds_len<N, double>()];
  __shared__ double scratch[8 * ((NT + 63) / 64)];
  const cplx<double>* tw = tw_in_lds<N, double>() ? tw_s : tw_g;
  const int l = threadIdx.x / TPL, t = threadIdx.x % TPL;
  const int kx = blockIdx.x * Cfg::C + l;
  cplx<double>* line = lds + l * LP;
  const DevParams& p = a.p;
  const double inv_nsp = p.inv_nsp, dt = a.dt;
  const double kN2 = kval<N>(N / 2, inv_nsp) * kval<N>(N / 2, inv_nsp);
  const double tk = 2.0 * p.kappa;

  cplx<double> v[3][E], gp[E];
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const size_t o = lidx<N>(kx, t + TPL * j);
    v[0][j] = a.X0[o];
    v[1][j] = a.X1[o];
    v[2][j] = a.S0[o];
    gp[j] = a.G[o];   // the adjoint line, updated at the end
  }
  if constexpr (tw_in_lds<N, double>()) load_twiddles<N, NT>(tw_s, tw_g);
  __syncthreads();
  fft_regs<double, -1, N, 3, N, WAVE, E, Cfg::XM, Cfg::AS, Cfg::BS>(v, t, line, tw);
  cplx<double> jyn[E], cn[E];
  if (kx == 0) {
    partner_of<N, E, TPL, WAVE>(v[1], jyn, t, line);
    partner_of<N, E, TPL, WAVE>(v[2], cn, t, line);
  }
  double acc = 0.0;
  cplx<double> w[2][E];
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int ky = t + TPL * j;
    const double kyv = kval<N>(ky, inv_nsp), ky2 = kyv * kyv;
    const double sy = dkval<N>(ky, inv_nsp);
    const cplx<double> ch = v[2][j];
    cplx<double> r, lap;
    if (kx == 0) {
      const int kn = neg_index<N>(ky);
      r = mul_i(sy * v[1][j]);
      const cplx<double> rn = mul_i(dkval<N>(kn, inv_nsp) * jyn[j]);
      const double k2a = ky2, k2b = kN2 + ky2;
      lap = (0.5 * tk * (k2a + k2b)) * r + (0.5 * tk * (k2a - k2b)) * conj(rn);
      cplx<double> rA, rB, cA, cB;
      unpack0(r, rn, rA, rB);
      unpack0(ch, cn[j], cA, cB);
      acc += k2a * re_conj_mul(rA, cA) + k2b * re_conj_mul(rB, cB);
    } else {
      const double kxv = kval<N>(kx, inv_nsp);
      const double k2 = kxv * kxv + ky2;
      r = mul_i(kxv * v[0][j]) + mul_i(sy * v[1][j]);
      lap = (tk * k2) * r;
      acc += 2.0 * k2 * re_conj_mul(r, ch);
    }
    w[0][j] = (1.0 / N) * r;
    w[1][j] = (dt / N) * lap;
  }
  pin_partial(acc);
  fft_regs<double, +1, N, 2, N, WAVE, E, Cfg::XM, Cfg::AS, Cfg::BS>(w, t, line, tw);
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const size_t o = lidx<N>(kx, t + TPL * j);
    a.Xr[o] = w[0][j];
    a.G[o] = gp[j] + w[1][j];
  }
  double rr[1] = {acc};
  const int ops[1] = {RED_SUM};
  block_reduce0<NT, 1>(rr, ops, scratch);
  if (threadIdx.x == 0) a.acc[blockIdx.x] += rr[0] * 2.0 * dt / ((double)N * N);
}

// Rows per workgroup of the adjoint row pass: the forward rule for two
// arrays, halved while the grid has fewer workgroups than MI355X has CUs
// (256) and the workgroup stays whole wavefronts (at 512^2 the rule alone
// gives 4 rows and 128 workgroups, half the chip idle).
template <int N, int E>
constexpr int adj_rows() {
  int b = row_rows<N, double, 2, E>();
  while (b > 1 && N / b < 256 && (b / 2) * (N / 2 / E) >= 64) b /= 2;
  return b;
}

template <int N>
struct AdjRowCfg {
  static constexpr int L = N / 2;
  static constexpr int E = adj_E(L);
  static constexpr int TPL = L / E;
  static constexpr int B = adj_rows<N, E>();
  static constexpr int LP = tile_stride<double>(L, B);
  static constexpr int NT = B * TPL;
  static constexpr bool WAVE = TPL <= 64;
  static constexpr int GRID = N / B;
  // two arrays per transform; the transposed tiles of both fill the first set
  static constexpr int XM = adj_xm(2, B * LP);
  static constexpr int AS = XM == 0 ? 0 : B * LP, BS = 2 * B * LP;
  static constexpr int LDS = (XM == 2 ? 4 : 2) * B * LP;
};

template <int N>
__global__ void __launch_bounds__((AdjRowCfg<N>::NT))
adj_row4(AdjArgs a, const cplx<double>* __restrict__ tw_g) {
  using Cfg = AdjRowCfg<N>;
  constexpr int L = Cfg::L, E = Cfg::E, TPL = Cfg::TPL, B = Cfg::B, LP = Cfg::LP, NT = Cfg::NT;
  constexpr bool WAVE = Cfg::WAVE;
  __shared__ cplx<double> lds[Cfg::LDS];
  __shared__ cplx<double> tw_s[tw_lds_len<N, double>()];
  __shared__ double scratch[8 * ((NT + 63) / 64)];
  const cplx<double>* tw = tw_in_lds<N, double>() ? tw_s : tw_g;
  const int tile = xcd_tile(blockIdx.x, Cfg::GRID);
  const int j0 = tile * B;
  const int b = threadIdx.x / TPL, t = threadIdx.x % TPL;
  cplx<double>* row = lds + b * LP;
  const size_t rbase = (size_t)(j0 + b) * N;
  const DevParams& p = a.p;

  // every global load first: transposed tiles of r_v and the state, the
  // adjoint tile updated at the end, phi, twiddles
  static_assert(B * L == NT * E, "each thread moves E elements of the tile");
  cplx<double> tr_x[E], tr_s[E], gp[E];
  double ph[E][2];
#pragma unroll
  for (int r = 0; r < E; ++r) {
    int bb, kx;
    tile_elem<N, B>(threadIdx.x + NT * r, bb, kx);
    const size_t o = lidx<N>(kx, j0 + bb);
    tr_x[r] = a.Xr[o];
    tr_s[r] = a.S0[o];
    gp[r] = a.G[o];
  }
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int k = t + TPL * j;
    ph[j][0] = a.phi[rbase + 2 * k];
    ph[j][1] = a.phi[rbase + 2 * k + 1];
  }
  if constexpr (tw_in_lds<N, double>()) load_twiddles<N, NT>(tw_s, tw_g);
#pragma unroll
  for (int r = 0; r < E; ++r) {
    int bb, kx;
    tile_elem<N, B>(threadIdx.x + NT * r, bb, kx);
    lds[bb * LP + pad<double, L>(kx)] = tr_x[r];
    lds[(B + bb) * LP + pad<double, L>(kx)] = tr_s[r];
  }
  __syncthreads();
  cplx<double> v[2][E];
  const double scale = 1.0 / N;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const cplx<double>* tr = lds + (q * B + b) * LP;
#pragma unroll
    for (int j = 0; j < E; ++j) {
      const int k = t + TPL * j;
      const cplx<double> xk = tr[pad<double, L>(k)];
      if (k == 0) {
        v[q][j] = mk(scale * (xk.x + xk.y), scale * (xk.x - xk.y));
      } else {
        const cplx<double> xm = tr[pad<double, L>(L - k)];
        v[q][j] = scale * ((xk + conj(xm)) + mul_i(cmul(xk - conj(xm), conj(tw[k]))));
      }
    }
  }
  lds_sync<WAVE>();
  fft_regs<double, +1, L, 2, N, WAVE, E, Cfg::XM, Cfg::AS, Cfg::BS>(v, t, row, tw);
  // w = dt f''(c) r_v (core_shell.jl:53-54 differentiated in c); <r_v, g>
  const double A = p.A, Bc = p.B, cb = p.cb, cc = p.cc, dt = a.dt;
  double srg = 0.0;
  cplx<double> w[1][E];
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const double rv[2] = {v[0][j].x, v[0][j].y};
    const double cv[2] = {v[1][j].x, v[1][j].y};
    double out[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const double q = ph[j][e], c = cv[e];
      const double h = q * q * q * (10.0 - 15.0 * q + 6.0 * q * q);
      const double g = q * q * (1.0 - q) * (1.0 - q);
      const double fpp = 2.0 * A * (1.0 - h) +
                         2.0 * Bc * h * ((2.0 * c - cb - cc) * (2.0 * c - cb - cc) + 2.0 * (c - cb) * (c - cc));
      out[e] = dt * fpp * rv[e];
      srg += rv[e] * g;
    }
    w[0][j] = mk(out[0], out[1]);
  }
  pin_partial(srg);
  fft_regs<double, -1, L, 1, N, WAVE, E, Cfg::XM, Cfg::AS, Cfg::BS>(w, t, row, tw);
#pragma unroll
  for (int j = 0; j < E; ++j) row[pad<double, L>(t + TPL * j)] = w[0][j];
  lds_sync<WAVE>();
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int k = t + TPL * j;
    const cplx<double> zk = w[0][j];
    if (k == 0) {
      w[0][j] = mk(zk.x + zk.y, zk.x - zk.y);
    } else {
      const cplx<double> zm = row[pad<double, L>(L - k)];
      w[0][j] = 0.5 * (zk + conj(zm)) + cmul(tw[k], 0.5 * mul_mi(zk - conj(zm)));
    }
  }
  lds_sync<WAVE>();
#pragma unroll
  for (int j = 0; j < E; ++j) row[pad<double, L>(t + TPL * j)] = w[0][j];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < E; ++r) {
    int bb, kx;
    tile_elem<N, B>(threadIdx.x + NT * r, bb, kx);
    a.G[lidx<N>(kx, j0 + bb)] = gp[r] + lds[bb * LP + pad<double, L>(kx)];
  }
  double rr[1] = {srg};
  const int ops[1] = {RED_SUM};
  block_reduce0<NT, 1>(rr, ops, scratch);
  if (threadIdx.x == 0) a.acc[blockIdx.x] += rr[0] * dt;
}

}  // namespace csg
