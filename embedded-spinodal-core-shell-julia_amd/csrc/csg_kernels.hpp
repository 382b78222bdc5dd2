// Pass kernels of the core-shell Cahn-Hilliard step (core_shell.jl:57-103).
//
// Data layout in HBM (N x N grid, L = N/2):
//   * real fields (phi, host I/O staging) in "row layout": x[j*N + i], the
//     Julia column-major matrix c[i+1, j+1]; a row is fixed j, contiguous i.
//   * spectra in "line layout": X[kx*N + j], kx in [0, L), j in [0, N), one
//     complex value per site pair.  X[kx][j] is the R2C transform along i of
//     row j (packed: line 0 holds (X(0,j), X(L,j)) as (re, im), both real),
//     optionally transformed along j as well inside a column pass.
// Every array is N*N reals: in-place row passes need no ping-pong.
//
// One step = four passes (each a kernel):
//   P1 row pass    S -> c (C2R along i) -> f'(c), bulk f, min/max -> F' (R2C)
//   P2 column pass S, F' -> FFT_j -> mu^ = F'^ + 2 kappa K^2 c^ -> Gx, Gy
//                  = i kd mu^ -> IFFT_j; gradient-energy partial
//   P3 row pass    Gx or Gy -> g -> J = M0 h(phi) g, dissipation -> R2C
//                  (one gradient component per workgroup)
//   P4 column pass Jx, Jy -> FFT_j -> inc = dt (i kdx Jx^ + i kdy Jy^)/den
//                  -> IFFT_j -> S += inc   (core_shell.jl:74-75, linear in c^);
//                  workgroup 0 also reduces the step's partials into the
//                  history row.
// Row passes read and write their tiles transposed (line layout), so column
// passes see contiguous lines.
#pragma once
#include "csg_fft.hpp"

namespace csg {

struct DevParams {
  double A, B, chi, P, ca, cb, cc, kappa, mobility, Q;
  double spacing;
  double inv_nsp;  // (1/spacing)/N, the fftfreq multiplier
};

constexpr double TWO_PI = 6.283185307179586;  // Julia's 2pi

// k_m = 2pi * fftfreq(N, 1/spacing)[m]   (core_shell.jl:18)
template <int N>
__device__ __forceinline__ double kval(int m, double inv_nsp) {
  const int ms = (m < N / 2) ? m : m - N;
  return TWO_PI * (static_cast<double>(ms) * inv_nsp);
}
// derivative wavenumber: Nyquist zeroed (core_shell.jl:19-20)
template <int N>
__device__ __forceinline__ double dkval(int m, double inv_nsp) {
  return (m == N / 2) ? 0.0 : kval<N>(m, inv_nsp);
}

enum RowMode { ROW_FWD = 0, ROW_INV = 1, ROW_P1 = 2, ROW_P3 = 3 };
enum ColMode { COL_P2 = 0, COL_MU = 1, COL_P4 = 2, COL_RATE = 3 };

// Reduction of the per-workgroup partials into one history row
// [t, mean c, F, min c, max c, dissipation] (core_shell.jl:92).
struct FinalArgs {
  const double* p1;   // 5 per P1 workgroup: sum f, sum of row sums, min, max, #non-finite
  const double* p2;   // 1 per P2 workgroup: gradient-energy sum
  const double* p3;   // 1 per P3 workgroup: dissipation sum
  int n1, n2, n3;
  double* hist;       // history buffer (6 per row); NULL: no history
  long long* ctr;     // [0] = row index of this step, [1] = first non-finite step (-1)
  long long every;    // rows s with s % every == 0 are kept, at hist + 6 (s / every); 0 or 1: all
  double dt;
  double spacing2;
  double inv_n2;      // 1 / N^2
  double kappa;
  // ensembles: instance q uses p1..p3 + q*s_part, hist + q*s_hist,
  // ctr + 2q and, when params is set, params[q].kappa / .spacing
  const struct DevParams* params;
  long long s_part, s_hist;
};

// Ensemble of same-size instances: instance q (blockIdx.y) finds its arrays
// at the base pointers plus q times these byte strides, and its parameters at
// params[q] (NULL: PassArgs::p for every instance).
struct InstanceStrides {
  const DevParams* params;
  long long in0, in1, out0, out1, phi;   // bytes
  long long part;                        // doubles
};

struct PassArgs {
  const void* in0;
  const void* in1;
  void* out0;
  void* out1;
  const void* phi;   // real rows, compute type
  double* part;      // per-workgroup partial records
  double dt;
  DevParams p;
  FinalArgs fin;     // P4 only
  InstanceStrides is;
  int ninst;         // instances (gridDim.y); 0 or 1: a single field
  const void* sin;   // P4: state read from here and the sum written to out0 (NULL: in place)
  int nt;            // cache hints of the step passes: NT_STORE | NT_LOAD | NT_PHI
};

__device__ __forceinline__ const void* offset_bytes(const void* p, long long q, long long s) {
  return p ? static_cast<const char*>(p) + q * s : p;
}
__device__ __forceinline__ void* offset_bytes(void* p, long long q, long long s) {
  return p ? static_cast<char*>(p) + q * s : p;
}

// FinalArgs of instance q.
__device__ __forceinline__ FinalArgs final_view(const FinalArgs& f0, int q) {
  FinalArgs f = f0;
  if (!f0.hist) return f;   // no history row: nothing of the record is used
  f.p1 = f0.p1 + q * f0.s_part;
  f.p2 = f0.p2 + q * f0.s_part;
  f.p3 = f0.p3 + q * f0.s_part;
  f.hist = f0.hist + q * f0.s_hist;
  f.ctr = f0.ctr + 2 * q;
  if (f0.params) {
    f.kappa = f0.params[q].kappa;
    f.spacing2 = f0.params[q].spacing * f0.params[q].spacing;
  }
  return f;
}

// PassArgs of instance q (the grid row of a batched launch).
__device__ __forceinline__ PassArgs instance_view(const PassArgs& a0, int q) {
  if (q == 0 && !a0.is.params) return a0;
  PassArgs a = a0;
  a.in0 = offset_bytes(a0.in0, q, a0.is.in0);
  a.in1 = offset_bytes(a0.in1, q, a0.is.in1);
  a.out0 = offset_bytes(a0.out0, q, a0.is.out0);
  a.out1 = offset_bytes(a0.out1, q, a0.is.out1);
  a.phi = offset_bytes(a0.phi, q, a0.is.phi);
  a.sin = offset_bytes(a0.sin, q, a0.is.out0);
  if (a0.part) a.part = a0.part + q * a0.is.part;
  if (a0.is.params) a.p = a0.is.params[q];
  a.fin = final_view(a0.fin, q);
  return a;
}

template <class T> __device__ __forceinline__ bool finite(T v) { return isfinite(v); }

// Materialise a pass's partial sums where they are computed.  The partials
// are only read by the end-of-pass reduction, so without this the compiler
// sinks the bulk-energy, min/max and dissipation terms below the forward FFT
// and the stores, keeping every site's intermediate values live across the
// transform (4096^2 FP64 P1: 256 VGPRs plus 94 AGPRs, one wave per SIMD).
// The empty asm has side effects, so it stays ahead of the FFT's barriers.
__device__ __forceinline__ void pin_partial(double& r) { asm volatile("" : "+v"(r)); }
template <int K>
__device__ __forceinline__ void pin_partials(double (&r)[K]) {
#pragma unroll
  for (int q = 0; q < K; ++q) pin_partial(r[q]);
}

// Line blocking of the line layout (a build option, off by default).  The L
// lines are stored in blocks of KB consecutive k_x, j-major inside a block:
// element (kx, j) sits at (kx / KB) N KB + j KB + kx % KB.  A row tile of b
// rows then moves b KB contiguous elements per block instead of b, and a
// column workgroup reads its line with a stride of KB elements.  Measured at
// 4096^2 with KB = 2 (-DCSG_LINE_BLOCK_MIN_N=4096): the row passes gain 6-8 %
// (FP64) and 4-6 % (mixed), the column passes lose 40-70 %, so KB = 1 (the
// plain layout kx N + j) is used for every grid.
#ifndef CSG_LINE_BLOCK_MIN_N
#define CSG_LINE_BLOCK_MIN_N (1 << 30)
#endif
#ifndef CSG_LINE_BLOCK
#define CSG_LINE_BLOCK 2
#endif
template <int N> constexpr int line_block() { return N >= CSG_LINE_BLOCK_MIN_N ? CSG_LINE_BLOCK : 1; }

template <int N, int KB = line_block<N>()> __host__ __device__ __forceinline__ size_t lidx(int kx, int j) {
  if constexpr (KB == 1) return (size_t)kx * N + j;
  else return (size_t)(kx / KB) * (N * KB) + (size_t)j * KB + (kx % KB);
}

// Element f of a row tile of B rows in its transposed (line-layout) order:
// lanes take consecutive k_x of a block, then consecutive rows, then blocks,
// so neighbouring lanes touch neighbouring addresses.  Sets the tile row bb
// and the k_x.
template <int N, int B, int KB = line_block<N>()>
__host__ __device__ __forceinline__ void tile_elem(int f, int& bb, int& kx) {
  const int r = f / KB;
  bb = r % B;
  kx = (r / B) * KB + f % KB;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Remap so that each XCD receives a contiguous range of tiles: row
// tiles j0 and j0+B share 128-byte lines of the line-layout arrays, which
// are then fetched once into that XCD's L2.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcd_tile(int b, int grid) {
  if (grid % NUM_XCD != 0) return b;
  return (b % NUM_XCD) * (grid / NUM_XCD) + b / NUM_XCD;
}

// The twiddle half table (N/2 entries) is staged in LDS when it takes at most
// 16 KiB; larger tables (N = 4096 in FP64) are read through the caches.
template <int N, class CT>
constexpr bool tw_in_lds() { return (N / 2) * 2 * sizeof(CT) <= 16384; }
template <int N, class CT>
constexpr int tw_lds_len() { return tw_in_lds<N, CT>() ? N / 2 : 1; }
// Schedules may have the FFT engine read its twiddles from per-stage tables
// in LDS (lanes read consecutive entries) where the half table is
// LDS-resident and the tables of the (LEN, E) plans fit 32 KiB.
template <int N, class CT, int LEN, int E>
constexpr bool stab_fits() { return tw_in_lds<N, CT>() && stab_len(LEN, E) * 2 * (int)sizeof(CT) <= 32 * 1024; }

// Cache hints (PassArgs::nt, chosen by the host from the device footprint):
//   NT_STORE  pass outputs stored non-temporally: not kept in this XCD's L2,
//             which the next pass (on other XCDs) would not use; pays while
//             the working set stays in the Infinity Cache;
//   NT_LOAD   the column passes' line inputs, read once, loaded
//             non-temporally; pays when the working set lives in HBM, where
//             it leaves L2 to the row passes' shared transposed lines;
//   NT_PHI    phi loaded non-temporally in the row passes; pays only when a
//             single field is far larger than the Infinity Cache, and is
//             honoured by the FP64 row passes of N >= PHI_HINT_MIN_N only.
enum CacheHint { NT_STORE = 1, NT_LOAD = 2, NT_PHI = 4 };
constexpr int PHI_HINT_MIN_N = 4096;

template <class T>
__device__ __forceinline__ void store_out(cplx<T>* p, cplx<T> v, int nt) {
  if (nt & NT_STORE) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    v2 w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<v2*>(p));
  } else {
    *p = v;
  }
}

template <class T>
__device__ __forceinline__ cplx<T> load_line(const cplx<T>* p, int nt) {
  if (nt & NT_LOAD) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 w = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return cplx<T>{w.x, w.y};
  }
  return *p;
}

__device__ __forceinline__ void neumaier(double& s, double& c, double v) {
  const double t = s + v;
  if (fabs(s) >= fabs(v)) c += (s - t) + v; else c += (v - t) + s;
  s = t;
}

// Reduce the partials of one step into its history row (one workgroup).
template <int NT>
__device__ void finalize_block(const FinalArgs& f, double* scratch) {
  double s[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0};
  double mn = INFINITY, mx = -INFINITY, bad = 0.0;
  for (int i = threadIdx.x; i < f.n1; i += NT) {
    const double* r = f.p1 + 5 * i;
    neumaier(s[0], c[0], r[0]);
    neumaier(s[1], c[1], r[1]);
    mn = red_combine<RED_MIN>(mn, r[2]);
    mx = red_combine<RED_MAX>(mx, r[3]);
    bad += r[4];
  }
  for (int i = threadIdx.x; i < f.n2; i += NT) neumaier(s[2], c[2], f.p2[i]);
  for (int i = threadIdx.x; i < f.n3; i += NT) neumaier(s[3], c[3], f.p3[i]);
  double v[7] = {s[0] + c[0], s[1] + c[1], s[2] + c[2], s[3] + c[3], mn, mx, bad};
  const int ops[7] = {RED_SUM, RED_SUM, RED_SUM, RED_SUM, RED_MIN, RED_MAX, RED_SUM};
  block_reduce<NT, 7>(v, ops, scratch);
  if (threadIdx.x == 0) {
    const long long row = f.ctr[0];
    const long long every = f.every > 1 ? f.every : 1;
    const bool nonfinite = v[6] > 0 || v[4] != v[4] || v[5] != v[5];
    if (row % every == 0) {
      double* h = f.hist + 6 * (row / every);
      h[0] = static_cast<double>(row) * f.dt;
      h[1] = v[1] * f.inv_n2;
      h[2] = (v[0] + f.kappa * v[2] * f.inv_n2) * f.spacing2;   // core_shell.jl:80-81
      h[3] = nonfinite ? NAN : v[4];
      h[4] = nonfinite ? NAN : v[5];
      h[5] = v[3] * f.spacing2;                                  // core_shell.jl:68
    }
    if (nonfinite && f.ctr[1] < 0) f.ctr[1] = row;
    f.ctr[0] = row + 1;
  }
}

// ---------------------------------------------------------------------------
// Row pass: B rows of L = N/2 packed complex per workgroup, TPL threads per
// row.  CT: compute type; TIN/TOUT: packed input/output element types; the
// real rows (FWD input, INV output) are double.
// Rows per workgroup: as many as fit an LDS budget (more rows make the
// transposed accesses to the line layout wider), up to 8 and 1024 threads.
#ifndef CSG_ROW_LDS
#define CSG_ROW_LDS (40 * 1024)
#endif
template <int N, class CT, int NARR, int E = fft_E(N / 2)>
constexpr int row_rows() {
  constexpr int L = N / 2, TPL = L / E;
  const int lp = NARR * padded_len<CT>(L) * 2 * (int)sizeof(CT);
  int b = 1;
  while (b < 8 && N % (b * 2) == 0 && (b * 2) * lp <= CSG_ROW_LDS && (b * 2) * TPL <= 1024) b *= 2;
  while (b * TPL < 64 && N % (b * 2) == 0) b *= 2;
  return b;
}

// Schedule of a pass kernel (0 = the default rule):
//   E   elements per thread (register tile; sets the radix plan),
//   W   rows (row pass) or lines (column pass) per workgroup,
//   G   row passes: transposed-tile loads issued in groups of G, each group
//       written to LDS before the next is issued (E = all at once),
//   PF  1: the pass's other inputs (phi in P1/P3, the state in P4) are loaded
//       with the first group of loads; 0: where they are used; 2: once the
//       first loads have arrived, before the first transform,
//   X   FFT exchange layout (fft_regs XM: 0 one array at a time, 1 all
//       arrays together, 2 all together with two buffer sets); -1 = 0,
//   T   FFT twiddles: 1 per-stage tables copied to LDS (lanes read
//       consecutive entries), 0 the half table (staged in LDS up to 16
//       KiB, else read through the caches), 2 the half table staged in LDS
//       whatever its size; -1 = 0,
//   S   P3 only: 1 gives each gradient component its own workgroups (phi is
//       then read once per component), 0 transforms both components of W
//       rows in one workgroup (phi read once), 2 is 1 with the two
//       components of the same rows on neighbouring workgroups.
template <int E_ = 0, int W_ = 0, int G_ = 0, int PF_ = 1, int X_ = -1, int T_ = -1, int S_ = 1>
struct Sched {
  static constexpr int E = E_, W = W_, G = G_, PF = PF_, X = X_, T = T_, S = S_;
};

template <class CT, class TIN, class TOUT, int N, int MODE, class SC = Sched<>>
struct RowCfg {
  static constexpr int NN = N;
  static constexpr int L = N / 2;
  static constexpr int E = SC::E ? SC::E : fft_E(L);
  static constexpr int TPL = L / E;
  static constexpr bool SPLIT = (MODE == ROW_P3) && SC::S != 0;
  static constexpr int NARR = (MODE == ROW_P3 && !SPLIT) ? 2 : 1;   // arrays per workgroup
  static constexpr int B = SC::W ? SC::W : row_rows<N, CT, NARR, E>();
  static constexpr int G = SC::G ? SC::G : E;
  static constexpr int PF = SC::PF;
  static constexpr int XM = SC::X >= 0 ? SC::X : 0;
  static constexpr int LP = tile_stride<CT>(L, B);   // row stride of the tile
  static constexpr int AS = B * LP, BS = NARR * B * LP;   // exchange strides
  static constexpr bool ST = SC::T == 1;
  static constexpr bool TWL = tw_in_lds<N, CT>() || SC::T == 2;   // half table in LDS
  static_assert(!ST || stab_fits<N, CT, L, E>(), "stage tables do not fit this schedule");
  static constexpr int LDS = (XM == 2 ? 2 : 1) * NARR * B * LP;
  static constexpr int NT = B * TPL;
  static constexpr bool WAVE = TPL <= 64;
  static constexpr int TILES = N / B;
  static constexpr int GRID = SPLIT ? 2 * TILES : TILES;
  static_assert(fft_E_valid(L, E) && E % G == 0 && N % B == 0, "row schedule");
  static_assert(NT <= 1024 && (NT % 64 == 0 || (NT < 64 && (NT & (NT - 1)) == 0)), "workgroup of whole wavefronts");
};

// LDS of one row-pass workgroup, carved from one byte array: the tile and
// exchange buffers, the twiddle half table, the stage tables and the
// reduction scratch.
template <class Cfg, class CT>
struct RowShared {
  static constexpr int NLDS = Cfg::LDS, NTW = Cfg::TWL ? Cfg::NN / 2 : 0,
                       NST = Cfg::ST ? stab_len(Cfg::L, Cfg::E) : 0, NSCR = 8 * ((Cfg::NT + 63) / 64);
  static constexpr int OFF_TW = NLDS * (int)sizeof(cplx<CT>);
  static constexpr int OFF_ST = OFF_TW + NTW * (int)sizeof(cplx<CT>);
  static constexpr int OFF_SCR = OFF_ST + NST * (int)sizeof(cplx<CT>);
  static constexpr int BYTES = OFF_SCR + NSCR * (int)sizeof(double);
};

// P1 pointwise work at a thread's E site pairs (c in v[j].x / .y, phi in
// ph[j][0 / 1]): v becomes f'(c) (bulk_derivative, core_shell.jl:53-54);
// the bulk energy sum, min c, max c and the non-finite count go to red[0],
// red[2], red[3], red[4].  Non-finite values are counted and kept out of
// min/max (branchless).
template <class CT, int E>
__device__ __forceinline__ void p1_sites(cplx<CT> (&v)[E], const CT (&ph)[E][2], const DevParams& p, double (&red)[5]) {
  const CT A = p.A, Bc = p.B, chi = p.chi, P = p.P, ca = p.ca, cb = p.cb, cc = p.cc;
  double fsum = 0.0, cmin = INFINITY, cmax = -INFINITY, nbad = 0.0;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const CT cin[2] = {v[j].x, v[j].y};
    const CT qq[2] = {ph[j][0], ph[j][1]};
    CT out[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const CT c = cin[e], q = qq[e];
      const CT h = q * q * q * (CT(10) - CT(15) * q + CT(6) * q * q);   // core_shell.jl:40
      const CT g = q * q * (CT(1) - q) * (CT(1) - q);                     // :41
      // bulk_derivative, core_shell.jl:53-54
      out[e] = CT(2) * A * (CT(1) - h) * (c - ca) +
               CT(2) * Bc * h * (c - cb) * (c - cc) * (CT(2) * c - cb - cc) - chi * P * g;
      // bulk_energy, core_shell.jl:47-48
      const CT fb = A * (CT(1) - h) * (c - ca) * (c - ca) +
                    Bc * h * (c - cb) * (c - cb) * (c - cc) * (c - cc) + P * (CT(1) - chi * c) * g;
      fsum += static_cast<double>(fb);
      const bool ok = finite(c);
      const double cd = static_cast<double>(c);
      cmin = fmin(cmin, ok ? cd : INFINITY);
      cmax = fmax(cmax, ok ? cd : -INFINITY);
      nbad += ok ? 0.0 : 1.0;
    }
    v[j] = mk(out[0], out[1]);
  }
  red[0] += fsum;
  red[2] = fmin(red[2], cmin);
  red[3] = fmax(red[3], cmax);
  red[4] += nbad;
  pin_partials(red);
}

// P3 pointwise work: J = M0 h(phi) g for each of the NARR gradient
// components in v (core_shell.jl:42, 66-67); the dissipation sum
// sum(M (gx^2 + gy^2)) (:68) goes to red[0].
template <class CT, int NARR, int E>
__device__ __forceinline__ void p3_sites(cplx<CT> (&v)[NARR][E], const CT (&ph)[E][2], const DevParams& p,
                                         double (&red)[5]) {
  const CT mob = p.mobility;
  double dsum = 0.0;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const CT q0 = ph[j][0], q1 = ph[j][1];
    const CT M0 = mob * (q0 * q0 * q0 * (CT(10) - CT(15) * q0 + CT(6) * q0 * q0));   // :42
    const CT M1 = mob * (q1 * q1 * q1 * (CT(10) - CT(15) * q1 + CT(6) * q1 * q1));
#pragma unroll
    for (int q = 0; q < NARR; ++q) {
      const cplx<CT> g = v[q][j];
      dsum += static_cast<double>(M0 * (g.x * g.x)) + static_cast<double>(M1 * (g.y * g.y));
      v[q][j] = mk(M0 * g.x, M1 * g.y);
    }
  }
  red[0] += dsum;
  pin_partials(red);
}

// R2C post-processing of one row after the forward FFT along i: X_k = E_k +
// W_N^k O_k from (Z_k, Z_{L-k}), slot 0 packing (X_0, X_L).  `tr` is the
// row's LDS line buffer (padded layout); both orderings points are included.
template <class CT, int L, int E, bool WAVE>
__device__ __forceinline__ void r2c_post(cplx<CT> (&v)[E], cplx<CT>* tr, int t, const cplx<CT>* tw) {
  constexpr int TPL = L / E;
#pragma unroll
  for (int j = 0; j < E; ++j) tr[pad<CT, L>(t + TPL * j)] = v[j];
  lds_sync<WAVE>();
  const CT half = static_cast<CT>(0.5);
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int k = t + TPL * j;
    const cplx<CT> zk = v[j];
    if (k == 0) {
      v[j] = mk(zk.x + zk.y, zk.x - zk.y);
    } else {
      const cplx<CT> zm = tr[pad<CT, L>(L - k)];
      const cplx<CT> ek = half * (zk + conj(zm));
      const cplx<CT> ok = half * mul_mi(zk - conj(zm));
      v[j] = ek + cmul(tw[k], ok);
    }
  }
  lds_sync<WAVE>();
}

// Body of a row pass for tile-workgroup `bx` of instance q, with its LDS
// regions (tile and exchange buffers, twiddle half table, stage tables,
// reduction scratch; RowShared sizes): the batched kernel below runs it for
// (blockIdx.x, blockIdx.y) on separate __shared__ arrays (distinct objects,
// so the compiler knows they do not alias), a persistent kernel for its work
// items on regions of one buffer.
template <class CT, class TIN, class TOUT, int N, int MODE, class SC>
__device__ __forceinline__ void row_body(const PassArgs& a0, int bx, int q, const cplx<CT>* __restrict__ tw_g,
                                         cplx<CT>* lds, cplx<CT>* tw_s, cplx<CT>* st_s, double* scratch) {
  using Cfg = RowCfg<CT, TIN, TOUT, N, MODE, SC>;
  const PassArgs a = instance_view(a0, q);
  constexpr int L = Cfg::L, E = Cfg::E, TPL = Cfg::TPL, B = Cfg::B, LP = Cfg::LP, NT = Cfg::NT;
  constexpr int NARR = Cfg::NARR;
  constexpr bool WAVE = Cfg::WAVE;
  constexpr int XM = Cfg::XM, AS = Cfg::AS, BS = Cfg::BS;
  constexpr bool ST = Cfg::ST;
  const cplx<CT>* tw = Cfg::TWL ? tw_s : tw_g;   // real-FFT pre/post twiddles
  const cplx<CT>* ftw = ST ? st_s : tw;                     // FFT engine twiddles
  const int b = threadIdx.x / TPL, t = threadIdx.x % TPL;
  cplx<CT>* row = lds + b * LP;                 // FFT exchange buffer of this row
  cplx<CT> v[NARR][E];
  // partials, reduced at the end: P1 [sum f, sum of row sums, min c, max c,
  // #non-finite]; P3 [dissipation]
  double red[5] = {0.0, 0.0, INFINITY, -INFINITY, 0.0};

  // Loads are issued before the first wait where the schedule says so (tile
  // group, phi, twiddles): one memory latency instead of one per phase.
  static_assert(B * L == NT * E, "each thread moves E elements of the tile");
  constexpr int G = Cfg::G;
  // phi: 1 loaded with the tile, 2 after the tile's barrier, 0 where used
  constexpr int PFM = (MODE == ROW_P1 || MODE == ROW_P3) ? Cfg::PF : 0;
  constexpr bool PF = PFM == 1;
  constexpr bool PHI_HINT = std::is_same<CT, double>::value && N >= PHI_HINT_MIN_N;
  // tile -> component (split P3: tiles [0, TILES) take Gx/Jx, [TILES, 2
  // TILES) take Gy/Jy) and first row
  // (S = 2: tiles 2t and 2t+1 take Gx and Gy of the same rows, so the XCD
  // that runs one also runs the other and the second phi read hits its L2)
  auto arr_of = [&](int tile) {
    if constexpr (Cfg::SPLIT && SC::S == 2) return tile & 1;
    else return (Cfg::SPLIT && tile >= Cfg::TILES) ? 1 : 0;
  };
  auto row0_of = [&](int tile) {
    if constexpr (Cfg::SPLIT && SC::S == 2) return (tile >> 1) * B;
    else return (tile - arr_of(tile) * Cfg::TILES) * B;
  };
  auto load_phi = [&](size_t rbase, CT (&ph)[E][2]) {
    const CT* phi = static_cast<const CT*>(a.phi);
    // The NT_PHI branch is compiled only where the hint can be set (FP64
    // fields of 4096^2 and up); elsewhere it would cost its control flow.
    if constexpr (PHI_HINT) {
      if (a.nt & NT_PHI) {
        typedef CT v2 __attribute__((ext_vector_type(2)));
#pragma unroll
        for (int j = 0; j < E; ++j) {
          const v2 w = __builtin_nontemporal_load(reinterpret_cast<const v2*>(phi + rbase + 2 * (t + TPL * j)));
          ph[j][0] = w.x;
          ph[j][1] = w.y;
        }
        return;
      }
    }
#pragma unroll
    for (int j = 0; j < E; ++j) {
      const int k = t + TPL * j;
      ph[j][0] = phi[rbase + 2 * k];
      ph[j][1] = phi[rbase + 2 * k + 1];
    }
  };
  // transposed tile loads from line layout: lane-consecutive rows (coalesced)
  auto tile_load = [&](int arr0, int j0, int g0, cplx<TIN>(&tl)[NARR][E]) {
#pragma unroll
    for (int q = 0; q < NARR; ++q) {
      const cplx<TIN>* X = static_cast<const cplx<TIN>*>((arr0 + q) == 0 ? a.in0 : a.in1);
#pragma unroll
      for (int r = 0; r < G; ++r) {
        int bb, kx;
        tile_elem<N, B>(threadIdx.x + NT * (g0 + r), bb, kx);
        tl[q][g0 + r] = X[lidx<N>(kx, j0 + bb)];
      }
    }
  };
  auto tile_to_lds = [&](int g0, const cplx<TIN>(&tl)[NARR][E]) {
#pragma unroll
    for (int q = 0; q < NARR; ++q)
#pragma unroll
      for (int r = 0; r < G; ++r) {
        int bb, kx;
        tile_elem<N, B>(threadIdx.x + NT * (g0 + r), bb, kx);
        lds[(q * B + bb) * LP + pad<CT, L>(kx)] = cvt<CT>(tl[q][g0 + r]);
      }
  };
  // mean composition from the FP64 k_x = 0 row sums of the state (P1)
  auto row_sum = [&](int j0) {
    return threadIdx.x < B ? static_cast<double>(static_cast<const cplx<TIN>*>(a.in0)[lidx<N>(0, j0 + threadIdx.x)].x)
                           : 0.0;
  };
  // C2R pre-processing fused with the read: Z_k = (X_k + conj X_{L-k}) +
  // i (X_k - conj X_{L-k}) W_N^{-k}, times 1/N; slot 0 holds (X_0, X_L).
  // Then the inverse FFT along i.
  auto c2r = [&]() {
    const CT scale = static_cast<CT>(1.0 / N);
#pragma unroll
    for (int q = 0; q < NARR; ++q) {
      const cplx<CT>* tr = lds + (q * B + b) * LP;
#pragma unroll
      for (int j = 0; j < E; ++j) {
        const int k = t + TPL * j;
        const cplx<CT> xk = tr[pad<CT, L>(k)];
        if (k == 0) {
          v[q][j] = mk(scale * (xk.x + xk.y), scale * (xk.x - xk.y));
        } else {
          const cplx<CT> xm = tr[pad<CT, L>(L - k)];
          v[q][j] = scale * ((xk + conj(xm)) + mul_i(cmul(xk - conj(xm), conj(tw[k]))));
        }
      }
    }
    lds_sync<WAVE>();
    fft_regs<CT, +1, L, NARR, N, WAVE, E, XM, AS, BS, ST>(v, t, row, ftw);
  };
  // pointwise work at this thread's sites (P1: f'(c); bulk energy, min,
  // max, non-finite partials.  P3: J = M0 h(phi) g; dissipation partial),
  // then R2C along i (forward FFT, X_k = E_k + W_N^k O_k from (Z_k, Z_{L-k}))
  // and the transposed store of the tile to line layout
  auto pointwise_r2c_store = [&](int arr0, int j0, const CT(&ph)[E][2]) {
    if constexpr (MODE == ROW_P1) p1_sites<CT, E>(v[0], ph, a.p, red);
    if constexpr (MODE == ROW_P3) p3_sites<CT, NARR, E>(v, ph, a.p, red);
    fft_regs<CT, -1, L, NARR, N, WAVE, E, XM, AS, BS, ST>(v, t, row, ftw);
    // Multi-row FP64 tiles do the R2C post-processing in the transposed read
    // (one LDS round trip less: 1024^2 step -0.5 us); FP32 tiles keep it
    // before the tile write, where their conflict-free tile strides apply to
    // every read (fused: +0.7 us).
    constexpr bool FUSED = (B > 1 || NARR > 1) && std::is_same<CT, double>::value;
    if constexpr (!FUSED) {
#pragma unroll
      for (int q = 0; q < NARR; ++q) r2c_post<CT, L, E, WAVE>(v[q], lds + (q * B + b) * LP, t, tw);
    }
    if constexpr (B == 1 && NARR == 1) {
      // one row: the thread's own elements k = t + TPL j are its stores
      cplx<TOUT>* X = static_cast<cplx<TOUT>*>(arr0 == 0 ? a.out0 : a.out1);
#pragma unroll
      for (int j = 0; j < E; ++j) store_out(&X[lidx<N>(t + TPL * j, j0)], cvt<TOUT>(v[0][j]), a.nt);
      return;
    }
    // the tile (Z, or X when not fused) to the tile buffer, then the
    // transposed read and store
#pragma unroll
    for (int q = 0; q < NARR; ++q) {
      cplx<CT>* tr = lds + (q * B + b) * LP;
#pragma unroll
      for (int j = 0; j < E; ++j) tr[pad<CT, L>(t + TPL * j)] = v[q][j];
    }
    __syncthreads();
    const CT half = static_cast<CT>(0.5);
#pragma unroll
    for (int q = 0; q < NARR; ++q) {
      cplx<TOUT>* X = static_cast<cplx<TOUT>*>((arr0 + q) == 0 ? a.out0 : a.out1);
#pragma unroll
      for (int r = 0; r < E; ++r) {
        int bb, kx;
        tile_elem<N, B>(threadIdx.x + NT * r, bb, kx);
        const cplx<CT>* tr = lds + (q * B + bb) * LP;
        const cplx<CT> zk = tr[pad<CT, L>(kx)];
        cplx<CT> xk = zk;
        if constexpr (FUSED) {
          if (kx == 0) {
            xk = mk(zk.x + zk.y, zk.x - zk.y);
          } else {
            const cplx<CT> zm = tr[pad<CT, L>(L - kx)];
            xk = half * (zk + conj(zm)) + cmul(tw[kx], half * mul_mi(zk - conj(zm)));
          }
        }
        store_out(&X[lidx<N>(kx, j0 + bb)], cvt<TOUT>(xk), a.nt);
      }
    }
  };

  {
    const int tile = xcd_tile(bx, Cfg::GRID);
    const int arr0 = arr_of(tile);
    const int j0 = row0_of(tile);
    const size_t rbase = (size_t)(j0 + b) * N;   // real-row offset
    CT ph[E][2];   // phi at this thread's sites (P1, P3)
    if constexpr (MODE == ROW_FWD) {
      const double* x = static_cast<const double*>(a.in0);
#pragma unroll
      for (int j = 0; j < E; ++j) {
        const int k = t + TPL * j;
        v[0][j] = mk(static_cast<CT>(x[rbase + 2 * k]), static_cast<CT>(x[rbase + 2 * k + 1]));
      }
      if constexpr (Cfg::TWL) load_twiddles<N, NT>(tw_s, tw_g);
      if constexpr (ST) load_stage_tables<N, L, E, NT>(st_s, tw_g);
      __syncthreads();
    } else {
      // in groups of G elements per thread, each written to LDS before the
      // next is issued
      cplx<TIN> tl[NARR][E];
#pragma unroll
      for (int g0 = 0; g0 < E; g0 += G) {
        tile_load(arr0, j0, g0, tl);
        if (g0 == 0) {
          if constexpr (MODE == ROW_P1) red[1] = row_sum(j0);
          if constexpr (PF) load_phi(rbase, ph);
          if constexpr (Cfg::TWL) load_twiddles<N, NT>(tw_s, tw_g);
          if constexpr (ST) load_stage_tables<N, L, E, NT>(st_s, tw_g);
        }
        tile_to_lds(g0, tl);
      }
      lds_barrier();   // the prefetched phi row stays in flight
      if constexpr (PFM == 2) load_phi(rbase, ph);
      c2r();
    }

    if constexpr (MODE == ROW_INV) {
      double* x = static_cast<double*>(a.out0);
#pragma unroll
      for (int j = 0; j < E; ++j) {
        const int k = t + TPL * j;
        x[rbase + 2 * k] = static_cast<double>(v[0][j].x);
        x[rbase + 2 * k + 1] = static_cast<double>(v[0][j].y);
      }
      return;
    }
    if constexpr ((MODE == ROW_P1 || MODE == ROW_P3) && PFM == 0) load_phi(rbase, ph);
    pointwise_r2c_store(arr0, j0, ph);
  }

  if constexpr (MODE == ROW_P1) {
    const int ops[5] = {RED_SUM, RED_SUM, RED_MIN, RED_MAX, RED_SUM};
    block_reduce0<NT, 5>(red, ops, scratch);
    if (threadIdx.x == 0) {
      double* rec = a.part + 5 * bx;
#pragma unroll
      for (int q = 0; q < 5; ++q) rec[q] = red[q];
    }
  }
  if constexpr (MODE == ROW_P3) {
    double r[1] = {red[0]};
    const int ops[1] = {RED_SUM};
    block_reduce0<NT, 1>(r, ops, scratch);
    if (threadIdx.x == 0) a.part[bx] = r[0];
  }
}

template <class CT, class TIN, class TOUT, int N, int MODE, class SC = Sched<>>
__global__ void __launch_bounds__((RowCfg<CT, TIN, TOUT, N, MODE, SC>::NT))
row_pass(PassArgs a0, const cplx<CT>* __restrict__ tw_g) {
  using Cfg = RowCfg<CT, TIN, TOUT, N, MODE, SC>;
  using Sh = RowShared<Cfg, CT>;
  __shared__ cplx<CT> lds[Sh::NLDS];
  __shared__ cplx<CT> tw_s[Sh::NTW > 0 ? Sh::NTW : 1];
  __shared__ cplx<CT> st_s[Sh::NST > 0 ? Sh::NST : 1];
  __shared__ double scratch[Sh::NSCR];
  row_body<CT, TIN, TOUT, N, MODE, SC>(a0, blockIdx.x, blockIdx.y, tw_g, lds, tw_s, st_s, scratch);
}

// ---------------------------------------------------------------------------
// Column pass: C lines kx0 .. kx0+C-1 of length N per workgroup, TPL threads
// per line; each thread group transforms both input arrays of its line.
template <int N, class CT, int E = fft_E(N)>
constexpr int col_lines() {
  constexpr int TPL = N / E;
  if (TPL > 64) return 1;   // lines spanning wavefronts: one per workgroup
  const int lp = padded_len<CT>(N) * 2 * (int)sizeof(CT);
  int c = 1;
  while ((N / 2) % (c * 2) == 0 && (c * 2) * lp <= 40 * 1024 && (c * 2) * TPL <= 256) c *= 2;
  while (c * TPL < 64 && (N / 2) % (c * 2) == 0) c *= 2;
  return c;
}

template <class CT, class TS, int N, int MODE, class SC = Sched<>>
struct ColCfg {
  static constexpr int NN = N;
  static constexpr int E = SC::E ? SC::E : fft_E(N);
  static constexpr int TPL = N / E;
  static constexpr int C = SC::W ? SC::W : col_lines<N, CT, E>();
  static constexpr int PF = SC::PF;
  static constexpr int XM = SC::X >= 0 ? SC::X : 0;
  static constexpr int LP = padded_len<CT>(N);
  // exchange buffers: arrays at stride AS, the second set (XM = 2) at BS
  static constexpr int AS = XM == 0 ? 0 : C * LP, BS = 2 * C * LP;
  static constexpr int LDS = (XM == 0 ? 1 : XM == 1 ? 2 : 4) * C * LP;
  static constexpr bool ST = SC::T == 1;
  static constexpr bool TWL = tw_in_lds<N, CT>() || SC::T == 2;   // half table in LDS
  static_assert(!ST || stab_fits<N, CT, N, E>(), "stage tables do not fit this schedule");
  static constexpr int NT = C * TPL;
  static constexpr bool WAVE = TPL <= 64;
  static constexpr int GRID = (N / 2) / C;   // line workgroups
  // P4 gets one more workgroup (index 0) that reduces the step's partials into
  // the history row alongside the line workgroups instead of after them
  static constexpr int LAUNCH = GRID + (MODE == COL_P4 ? 1 : 0);
  static_assert(fft_E_valid(N, E) && (N / 2) % C == 0, "column schedule");
  static_assert(WAVE || C == 1, "block-synchronised lines need one line per workgroup");
  static_assert(NT <= 1024 && (NT % 64 == 0 || (NT < 64 && (NT & (NT - 1)) == 0)), "workgroup of whole wavefronts");
};

// LDS of one column-pass workgroup: line exchange buffers, the stage tables
// or the half table, the reduction scratch.
template <class Cfg, class CT>
struct ColShared {
  static constexpr int NLDS = Cfg::LDS,
                       NTW = Cfg::ST ? stab_len(Cfg::NN, Cfg::E) : (Cfg::TWL ? Cfg::NN / 2 : 0),
                       NSCR = 8 * ((Cfg::NT + 63) / 64);
  static constexpr int OFF_TW = NLDS * (int)sizeof(cplx<CT>);
  static constexpr int OFF_SCR = OFF_TW + NTW * (int)sizeof(cplx<CT>);
  static constexpr int BYTES = OFF_SCR + NSCR * (int)sizeof(double);
};

// Body of a column pass for workgroup `bx` of instance q (P4: bx = 0 is the
// history-row reduction, bx - 1 the line group), with its LDS regions (line
// buffers, stage tables or half table, reduction scratch; ColShared sizes),
// for the persistent kernel.  The batched kernel col_pass includes the same
// statements directly: compiled through a separate inlined function, the
// 4096^2 FP64 column kernels needed more registers and spilled (P2 183 ->
// 233 us per launch).
template <class CT, class TS, int N, int MODE, class SC>
__device__ __forceinline__ void col_body(const PassArgs& a0, int bx, int q, const cplx<CT>* __restrict__ tw_g,
                                         cplx<CT>* lds, cplx<CT>* tw_s, double* scratch) {
#include "csg_col_body.inc"
}

template <class CT, class TS, int N, int MODE, class SC = Sched<>>
__global__ void __launch_bounds__((ColCfg<CT, TS, N, MODE, SC>::NT))
col_pass(PassArgs a0, const cplx<CT>* __restrict__ tw_g) {
  using Sh = ColShared<ColCfg<CT, TS, N, MODE, SC>, CT>;
  __shared__ cplx<CT> lds[Sh::NLDS];
  __shared__ cplx<CT> tw_s[Sh::NTW > 0 ? Sh::NTW : 1];
  __shared__ double scratch[Sh::NSCR];
  const int bx = static_cast<int>(blockIdx.x), q = static_cast<int>(blockIdx.y);
#include "csg_col_body.inc"
}

// ---------------------------------------------------------------------------
// Pointwise evaluations on real rows (geometry, bulk_energy, bulk_derivative;
// core_shell.jl:39-55), FP64, for the single-call API.
enum PointMode { PT_GEOMETRY = 0, PT_BULK_ENERGY = 1, PT_BULK_DERIVATIVE = 2 };

template <class TP, int MODE>
__global__ void __launch_bounds__(256)
pointwise_kernel(const double* __restrict__ c, const TP* __restrict__ phi, double* __restrict__ o0,
                 double* __restrict__ o1, double* __restrict__ o2, long long nn, DevParams p) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= nn) return;
  const double q = static_cast<double>(phi[i]);
  const double h = q * q * q * (10.0 - 15.0 * q + 6.0 * q * q);
  const double g = q * q * (1.0 - q) * (1.0 - q);
  if constexpr (MODE == PT_GEOMETRY) {
    o0[i] = h;
    o1[i] = g;
    o2[i] = p.mobility * h;
  } else {
    const double x = c[i];
    if constexpr (MODE == PT_BULK_ENERGY)
      o0[i] = p.A * (1.0 - h) * (x - p.ca) * (x - p.ca) +
              p.B * h * (x - p.cb) * (x - p.cb) * (x - p.cc) * (x - p.cc) + p.P * (1.0 - p.chi * x) * g;
    else
      o0[i] = 2.0 * p.A * (1.0 - h) * (x - p.ca) +
              2.0 * p.B * h * (x - p.cb) * (x - p.cc) * (2.0 * x - p.cb - p.cc) - p.chi * p.P * g;
  }
}

// Standalone history reduction (the last row of a run, single evaluations);
// one workgroup per instance.
template <int NT>
__global__ void __launch_bounds__(NT) finalize_kernel(FinalArgs f) {
  __shared__ double scratch[8 * ((NT + 63) / 64)];
  finalize_block<NT>(final_view(f, blockIdx.x), scratch);
}

}  // namespace csg
