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
//   P3 row pass    Gx, Gy -> gx, gy -> J = M0 h(phi) g, dissipation -> R2C
//   P4 column pass Jx, Jy -> FFT_j -> inc = dt (i kdx Jx^ + i kdy Jy^)/den
//                  -> IFFT_j -> S += inc   (core_shell.jl:74-75, linear in c^)
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

struct PassArgs {
  const void* in0;
  const void* in1;
  void* out0;
  void* out1;
  const void* phi;   // real rows, compute type
  double* part;      // per-workgroup partial records
  double dt;
  DevParams p;
};

// Workgroup shapes.  Row passes: B rows of L complex per array; column passes:
// C lines of N complex per array.  Threads = lines*len/8 (one radix-8
// butterfly per thread per stage).
template <int N, int ELEM, int NARR>
constexpr int row_B() {
  // LDS budget per workgroup ~64 KiB, at least 1 and at most N rows.
  int b = (64 * 1024) / (NARR * (N / 2 + 1) * 2 * ELEM);
  int p = 1;
  while (p * 2 <= b && p * 2 <= 8) p *= 2;
  if (p > N) p = N;
  // at least 64 threads where the grid allows it
  while (NARR * p * (N / 2) / 8 < 64 && p * 2 <= N) p *= 2;
  return p;
}
template <int N, int ELEM>
constexpr int col_C() {
  int c = (64 * 1024) / (2 * (N + 1) * 2 * ELEM);
  int p = 1;
  while (p * 2 <= c && p * 2 <= 8 && 2 * p * 2 * N / 8 <= 1024) p *= 2;
  if (p > N / 2) p = N / 2;
  while (2 * p * N / 8 < 64 && p * 2 <= N / 2) p *= 2;
  return p;
}

template <class T> __device__ __forceinline__ bool finite(T v) { return isfinite(v); }

// Block reduction helpers (deterministic tree over an LDS scratch of NT doubles).
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}
template <int NT, bool MAX>
__device__ __forceinline__ double block_minmax(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      double a = red[threadIdx.x], b = red[threadIdx.x + s];
      // NaN-propagating: a NaN anywhere poisons the result
      red[threadIdx.x] = (a != a || b != b) ? (a + b) : (MAX ? fmax(a, b) : fmin(a, b));
    }
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// Row pass.  CT: compute/LDS type; TIN/TOUT: element types of the packed
// input/output arrays; the real rows (FWD input, INV output) are double.
template <class CT, class TIN, class TOUT, int N, int MODE>
struct RowCfg {
  static constexpr int L = N / 2;
  static constexpr int NARR = (MODE == ROW_P3) ? 2 : 1;
  static constexpr int B = row_B<N, sizeof(CT), NARR>();
  static constexpr int PITCH = L + 1;
  static constexpr int NLINES = NARR * B;
  static constexpr int NT_FFT = NLINES * L / 8;
  static constexpr int NT = NT_FFT < 64 ? 64 : NT_FFT;
  static constexpr int GRID = N / B;
};

template <class CT, class TIN, class TOUT, int N, int MODE>
__global__ void __launch_bounds__((RowCfg<CT, TIN, TOUT, N, MODE>::NT))
row_pass(PassArgs a, const cplx<CT>* __restrict__ tw) {
  using Cfg = RowCfg<CT, TIN, TOUT, N, MODE>;
  constexpr int L = Cfg::L, B = Cfg::B, PITCH = Cfg::PITCH, NARR = Cfg::NARR;
  constexpr int NLINES = Cfg::NLINES, NT = Cfg::NT;
  __shared__ cplx<CT> lds[NLINES * PITCH];
  __shared__ double red[NT];
  const int j0 = blockIdx.x * B;

  if constexpr (MODE == ROW_FWD) {
    // real rows -> LDS (row b, as L complex = interleaved reals)
    const double* x = static_cast<const double*>(a.in0);
    for (int f = threadIdx.x; f < B * L; f += NT) {
      const int b = f / L, n = f - b * L;
      const double* r = x + (size_t)(j0 + b) * N + 2 * n;
      lds[b * PITCH + n] = mk(static_cast<CT>(r[0]), static_cast<CT>(r[1]));
    }
  } else {
    // transposed tile load from line layout: lane-consecutive b (coalesced)
#pragma unroll
    for (int arr = 0; arr < NARR; ++arr) {
      const cplx<TIN>* X = static_cast<const cplx<TIN>*>(arr == 0 ? a.in0 : a.in1);
      for (int f = threadIdx.x; f < B * L; f += NT) {
        const int b = f % B, kx = f / B;
        lds[(arr * B + b) * PITCH + kx] = cvt<CT>(X[(size_t)kx * N + j0 + b]);
      }
    }
  }
  __syncthreads();

  if constexpr (MODE != ROW_FWD) {
    // C2R along i: packed -> real row (interleaved in place), includes 1/N
    c2r_pre<CT, N, NLINES, PITCH, NT>(lds, tw, static_cast<CT>(1.0 / N));
    fft_lines<CT, +1, L, NLINES, PITCH, NT, 2>(lds, tw);
  }

  if constexpr (MODE == ROW_INV) {
    double* x = static_cast<double*>(a.out0);
    for (int f = threadIdx.x; f < B * L; f += NT) {
      const int b = f / L, n = f - b * L;
      const cplx<CT> v = lds[b * PITCH + n];
      double* r = x + (size_t)(j0 + b) * N + 2 * n;
      r[0] = static_cast<double>(v.x);
      r[1] = static_cast<double>(v.y);
    }
    return;
  }

  if constexpr (MODE == ROW_P1) {
    // pointwise: f'(c) in place; bulk energy, min, max, non-finite partials
    const CT* phi = static_cast<const CT*>(a.phi);
    const DevParams& p = a.p;
    const CT A = p.A, Bc = p.B, chi = p.chi, P = p.P, ca = p.ca, cb = p.cb, cc = p.cc;
    double fsum = 0.0, cmin = INFINITY, cmax = -INFINITY;
    int bad = 0;
    for (int f = threadIdx.x; f < B * L; f += NT) {
      const int b = f / L, n = f - b * L;
      cplx<CT> v = lds[b * PITCH + n];
      const CT* ph = phi + (size_t)(j0 + b) * N + 2 * n;
      CT out[2];
      CT cin[2] = {v.x, v.y};
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const CT c = cin[e], q = ph[e];
        const CT h = q * q * q * (CT(10) - CT(15) * q + CT(6) * q * q);   // core_shell.jl:40
        const CT g = q * q * (CT(1) - q) * (CT(1) - q);                     // :41
        // bulk_derivative, core_shell.jl:53-54
        out[e] = CT(2) * A * (CT(1) - h) * (c - ca) +
                 CT(2) * Bc * h * (c - cb) * (c - cc) * (CT(2) * c - cb - cc) - chi * P * g;
        // bulk_energy, core_shell.jl:47-48
        const CT fb = A * (CT(1) - h) * (c - ca) * (c - ca) +
                      Bc * h * (c - cb) * (c - cb) * (c - cc) * (c - cc) + P * (CT(1) - chi * c) * g;
        fsum += static_cast<double>(fb);
        const double cd = static_cast<double>(c);
        cmin = fmin(cmin, cd);
        cmax = fmax(cmax, cd);
        bad |= !finite(c);
      }
      lds[b * PITCH + n] = mk(out[0], out[1]);
    }
    // mean composition from the FP64 k_x = 0 row sums of the state
    double msum = 0.0;
    if (threadIdx.x < B) {
      const cplx<TIN>* S = static_cast<const cplx<TIN>*>(a.in0);
      msum = static_cast<double>(S[j0 + threadIdx.x].x);
    }
    __syncthreads();
    fsum = block_sum<NT>(fsum, red);
    msum = block_sum<NT>(msum, red);
    cmin = block_minmax<NT, false>(bad ? NAN : cmin, red);
    cmax = block_minmax<NT, true>(bad ? NAN : cmax, red);
    const double nbad = block_sum<NT>(bad ? 1.0 : 0.0, red);
    if (threadIdx.x == 0) {
      double* rec = a.part + 5 * blockIdx.x;
      rec[0] = fsum;
      rec[1] = msum;
      rec[2] = cmin;
      rec[3] = cmax;
      rec[4] = nbad;
    }
  }

  if constexpr (MODE == ROW_P3) {
    // J = M0 h(phi) g for both gradient components; dissipation partial
    const CT* phi = static_cast<const CT*>(a.phi);
    const CT mob = a.p.mobility;
    double dsum = 0.0;
    for (int f = threadIdx.x; f < B * L; f += NT) {
      const int b = f / L, n = f - b * L;
      cplx<CT> gx = lds[b * PITCH + n];
      cplx<CT> gy = lds[(B + b) * PITCH + n];
      const CT* ph = phi + (size_t)(j0 + b) * N + 2 * n;
      CT M[2];
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const CT q = ph[e];
        M[e] = mob * (q * q * q * (CT(10) - CT(15) * q + CT(6) * q * q));   // :42
      }
      // core_shell.jl:68 sum(M .* (gx.^2 .+ gy.^2))
      dsum += static_cast<double>(M[0] * (gx.x * gx.x + gy.x * gy.x)) +
              static_cast<double>(M[1] * (gx.y * gx.y + gy.y * gy.y));
      lds[b * PITCH + n] = mk(M[0] * gx.x, M[1] * gx.y);
      lds[(B + b) * PITCH + n] = mk(M[0] * gy.x, M[1] * gy.y);
    }
    __syncthreads();
    dsum = block_sum<NT>(dsum, red);
    if (threadIdx.x == 0) a.part[blockIdx.x] = dsum;
  }

  if constexpr (MODE == ROW_P1 || MODE == ROW_P3) {
    // (block_sum ended with a barrier)
  } else {
    __syncthreads();
  }

  // R2C along i and transposed store
  fft_lines<CT, -1, L, NLINES, PITCH, NT, 2>(lds, tw);
  r2c_post<CT, N, NLINES, PITCH, NT>(lds, tw);
#pragma unroll
  for (int arr = 0; arr < NARR; ++arr) {
    cplx<TOUT>* X = static_cast<cplx<TOUT>*>(arr == 0 ? a.out0 : a.out1);
    for (int f = threadIdx.x; f < B * L; f += NT) {
      const int b = f % B, kx = f / B;
      X[(size_t)kx * N + j0 + b] = cvt<TOUT>(lds[(arr * B + b) * PITCH + kx]);
    }
  }
}

// ---------------------------------------------------------------------------
// Column pass.  Lines kx0 .. kx0+C-1 of length N.
template <class CT, class TS, int N, int MODE>
struct ColCfg {
  static constexpr int C = col_C<N, sizeof(CT)>();
  static constexpr int PITCH = N;
  static constexpr int NLINES = 2 * C;
  static constexpr int NT_FFT = NLINES * N / 8;
  static constexpr int NT = NT_FFT < 64 ? 64 : NT_FFT;
  static constexpr int GRID = (N / 2) / C;
};

template <class CT, class TS, int N, int MODE>
__global__ void __launch_bounds__((ColCfg<CT, TS, N, MODE>::NT))
col_pass(PassArgs a, const cplx<CT>* __restrict__ tw) {
  using Cfg = ColCfg<CT, TS, N, MODE>;
  constexpr int C = Cfg::C, PITCH = Cfg::PITCH, NLINES = Cfg::NLINES, NT = Cfg::NT;
  constexpr int PER = (C * N + NT - 1) / NT;  // k-space elements per thread
  __shared__ cplx<CT> lds[NLINES * PITCH];
  __shared__ double red[NT];
  const int kx0 = blockIdx.x * C;
  const DevParams& p = a.p;
  const double inv_nsp = p.inv_nsp;

  // load: lines [0,C) <- in0, [C,2C) <- in1 (contiguous)
  {
    const size_t base = (size_t)kx0 * N;
    if constexpr (MODE == COL_P2 || MODE == COL_MU) {
      const cplx<TS>* S = static_cast<const cplx<TS>*>(a.in0);
      const cplx<CT>* F = static_cast<const cplx<CT>*>(a.in1);
      for (int f = threadIdx.x; f < C * N; f += NT) {
        lds[f] = cvt<CT>(S[base + f]);
        lds[C * PITCH + f] = F[base + f];
      }
    } else {
      const cplx<CT>* X0 = static_cast<const cplx<CT>*>(a.in0);
      const cplx<CT>* X1 = static_cast<const cplx<CT>*>(a.in1);
      for (int f = threadIdx.x; f < C * N; f += NT) {
        lds[f] = X0[base + f];
        lds[C * PITCH + f] = X1[base + f];
      }
    }
  }
  __syncthreads();
  fft_lines<CT, -1, N, NLINES, PITCH, NT, 1>(lds, tw);

  const double kN2 = kval<N>(N / 2, inv_nsp) * kval<N>(N / 2, inv_nsp);  // (pi/spacing)^2

  if constexpr (MODE == COL_P2 || MODE == COL_MU) {
    cplx<CT> o0[PER], o1[PER];
    double esum = 0.0;
    const double tk = 2.0 * p.kappa;
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int f = threadIdx.x + t * NT;
      if (C * N % NT == 0 || f < C * N) {
        const int l = f / N, ky = f - l * N;
        const int kx = kx0 + l;
        const cplx<CT> ch = lds[l * PITCH + ky];
        const cplx<CT> fh = lds[(C + l) * PITCH + ky];
        const double kyv = kval<N>(ky, inv_nsp);
        const double ky2 = kyv * kyv;
        cplx<CT> mu;
        if (kx == 0) {
          // packed line: A part is k_x = 0, B part is k_x = N/2 (Nyquist, k^2 kept)
          const cplx<CT> chn = lds[l * PITCH + ((N - ky) & (N - 1))];
          const CT sa = static_cast<CT>(tk * (ky2 + 0.5 * kN2));   // (a+b)/2
          const CT sd = static_cast<CT>(-0.5 * tk * kN2);           // (a-b)/2
          mu = fh + sa * ch + sd * conj(chn);
          // |A|^2 K_A^2 + |B|^2 K_B^2 summed over ky, written per element
          const double m2 = (double)ch.x * ch.x + (double)ch.y * ch.y;
          const double re = (double)ch.x * chn.x - (double)ch.y * chn.y;
          esum += (ky2 + 0.5 * kN2) * m2 - 0.5 * kN2 * re;
        } else {
          const double kxv = kval<N>(kx, inv_nsp);
          const double k2 = kxv * kxv + ky2;   // core_shell.jl:23
          mu = fh + static_cast<CT>(tk * k2) * ch;
          const double m2 = (double)ch.x * ch.x + (double)ch.y * ch.y;
          esum += 2.0 * k2 * m2;   // Hermitian weight of interior k_x
        }
        if constexpr (MODE == COL_MU) {
          o0[t] = static_cast<CT>(1.0 / N) * mu;
        } else {
          const CT sx = static_cast<CT>(dkval<N>(kx, inv_nsp) / N);   // 0 on line 0
          const CT sy = static_cast<CT>(dkval<N>(ky, inv_nsp) / N);
          o0[t] = mul_i(sx * mu);
          o1[t] = mul_i(sy * mu);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int f = threadIdx.x + t * NT;
      if (C * N % NT == 0 || f < C * N) {
        const int l = f / N, ky = f - l * N;
        lds[l * PITCH + ky] = o0[t];
        if constexpr (MODE == COL_P2) lds[(C + l) * PITCH + ky] = o1[t];
      }
    }
    __syncthreads();
    if constexpr (MODE == COL_P2) {
      fft_lines<CT, +1, N, NLINES, PITCH, NT, 1>(lds, tw);
    } else {
      fft_lines<CT, +1, N, C, PITCH, NT, 1>(lds, tw);
    }
    const size_t base = (size_t)kx0 * N;
    cplx<CT>* O0 = static_cast<cplx<CT>*>(a.out0);
    for (int f = threadIdx.x; f < C * N; f += NT) O0[base + f] = lds[f];
    if constexpr (MODE == COL_P2) {
      cplx<CT>* O1 = static_cast<cplx<CT>*>(a.out1);
      for (int f = threadIdx.x; f < C * N; f += NT) O1[base + f] = lds[C * PITCH + f];
      esum = block_sum<NT>(esum, red);
      if (threadIdx.x == 0) a.part[blockIdx.x] = esum;
    }
    return;
  }

  if constexpr (MODE == COL_P4 || MODE == COL_RATE) {
    const double dt = a.dt;
    const double qk2dt = 2.0 * p.Q * p.kappa;   // core_shell.jl:74  2p.Q*p.kappa*k2^2*dt
    cplx<CT> o0[PER], o1[PER];
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int f = threadIdx.x + t * NT;
      if (C * N % NT == 0 || f < C * N) {
        const int l = f / N, ky = f - l * N;
        const int kx = kx0 + l;
        const cplx<CT> jx = lds[l * PITCH + ky];
        const cplx<CT> jy = lds[(C + l) * PITCH + ky];
        const double kyv = kval<N>(ky, inv_nsp);
        const double ky2 = kyv * kyv;
        const CT sy = static_cast<CT>(dkval<N>(ky, inv_nsp));
        if (kx == 0) {
          const int kn = (N - ky) & (N - 1);
          const cplx<CT> jyn = lds[(C + l) * PITCH + kn];
          const CT syn = static_cast<CT>(dkval<N>(kn, inv_nsp));
          // rate on the packed line: i*kdy*Jy^ (k_dx = 0 for k_x = 0 and N/2)
          const cplx<CT> r = mul_i(sy * jy);
          const cplx<CT> rn = mul_i(syn * jyn);
          if constexpr (MODE == COL_RATE) {
            // unpack: A = (r + conj rn)/2 (k_x = 0), B = (r - conj rn)/(2i) (k_x = N/2)
            o0[t] = static_cast<CT>(0.5) * (r + conj(rn));
            o1[t] = static_cast<CT>(0.5) * mul_mi(r - conj(rn));
          } else {
            const double k2a = ky2, k2b = kN2 + ky2;
            const double ia = 1.0 / (1.0 + qk2dt * (k2a * k2a) * dt);
            const double ib = 1.0 / (1.0 + qk2dt * (k2b * k2b) * dt);
            const CT s = static_cast<CT>(dt / N);
            const CT ha = static_cast<CT>(0.5 * (ia + ib)), hd = static_cast<CT>(0.5 * (ia - ib));
            o0[t] = s * (ha * r + hd * conj(rn));
          }
        } else {
          const double kxv = kval<N>(kx, inv_nsp);
          const CT sx = static_cast<CT>(kxv);   // interior k_x: no Nyquist
          const cplx<CT> r = mul_i(sx * jx) + mul_i(sy * jy);
          if constexpr (MODE == COL_RATE) {
            o0[t] = r;
          } else {
            const double k2 = kxv * kxv + ky2;
            const double den = 1.0 + qk2dt * (k2 * k2) * dt;
            o0[t] = static_cast<CT>(dt / (den * N)) * r;
          }
        }
      }
    }
    if constexpr (MODE == COL_RATE) {
      // k-space output, k_x = 0 .. N/2 lines of N (double): line kx at kx*N
      cplx<double>* R = static_cast<cplx<double>*>(a.out0);
#pragma unroll
      for (int t = 0; t < PER; ++t) {
        const int f = threadIdx.x + t * NT;
        if (C * N % NT == 0 || f < C * N) {
          const int l = f / N, ky = f - l * N;
          const int kx = kx0 + l;
          R[(size_t)kx * N + ky] = cvt<double>(o0[t]);
          if (kx == 0) R[(size_t)(N / 2) * N + ky] = cvt<double>(o1[t]);
        }
      }
      return;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int f = threadIdx.x + t * NT;
      if (C * N % NT == 0 || f < C * N) lds[f] = o0[t];
    }
    __syncthreads();
    fft_lines<CT, +1, N, C, PITCH, NT, 1>(lds, tw);
    // S += IFFT_j(inc)
    cplx<TS>* S = static_cast<cplx<TS>*>(a.out0);
    const size_t base = (size_t)kx0 * N;
    for (int f = threadIdx.x; f < C * N; f += NT) {
      cplx<TS> s = S[base + f];
      const cplx<CT> d = lds[f];
      s.x += static_cast<TS>(d.x);
      s.y += static_cast<TS>(d.y);
      S[base + f] = s;
    }
  }
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

// ---------------------------------------------------------------------------
// Reduction of the per-workgroup partials into one history row
// [t, mean c, F, min c, max c, dissipation] (core_shell.jl:92), FP64 with
// Neumaier-compensated sums in a fixed order.
struct FinalArgs {
  const double* p1;   // 5 per row-pass workgroup
  const double* p2;   // 1 per column-pass workgroup
  const double* p3;   // 1 per row-pass workgroup
  int n1, n2, n3;
  double* hist;       // base of the history buffer (6 per row)
  long long* ctr;     // [0] = row index of this step, [1] = first non-finite step (-1)
  double dt;
  double spacing2;
  double inv_n2;      // 1 / N^2
  double kappa;
  long long step0;    // step number of row 0
};

__device__ __forceinline__ void neumaier(double& s, double& c, double v) {
  double t = s + v;
  if (fabs(s) >= fabs(v)) c += (s - t) + v; else c += (v - t) + s;
  s = t;
}

template <int NT>
__device__ void finalize_block(const FinalArgs& f, double* red) {
  // each thread sums a strided slice (fixed order), then a fixed tree
  double s[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0};
  double mn = INFINITY, mx = -INFINITY, bad = 0.0;
  for (int i = threadIdx.x; i < f.n1; i += NT) {
    const double* r = f.p1 + 5 * i;
    neumaier(s[0], c[0], r[0]);
    neumaier(s[1], c[1], r[1]);
    const double a = r[2], b = r[3];
    if (a != a || b != b) bad += 1.0; else { mn = fmin(mn, a); mx = fmax(mx, b); }
    bad += r[4];
  }
  for (int i = threadIdx.x; i < f.n2; i += NT) neumaier(s[2], c[2], f.p2[i]);
  for (int i = threadIdx.x; i < f.n3; i += NT) neumaier(s[3], c[3], f.p3[i]);
  double tot[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) tot[q] = block_sum<NT>(s[q], red) + block_sum<NT>(c[q], red);
  mn = block_minmax<NT, false>(mn, red);
  mx = block_minmax<NT, true>(mx, red);
  bad = block_sum<NT>(bad, red);
  if (threadIdx.x == 0) {
    const long long row = f.ctr[0];
    const long long step = f.step0 + row;
    double* h = f.hist + 6 * row;
    h[0] = static_cast<double>(step) * f.dt;
    h[1] = tot[1] * f.inv_n2;
    // energy, core_shell.jl:80-81
    h[2] = (tot[0] + f.kappa * tot[2] * f.inv_n2) * f.spacing2;
    h[3] = bad > 0 ? NAN : mn;
    h[4] = bad > 0 ? NAN : mx;
    h[5] = tot[3] * f.spacing2;
    if (bad > 0 && f.ctr[1] < 0) f.ctr[1] = step;
    f.ctr[0] = row + 1;
  }
}

template <int NT>
__global__ void __launch_bounds__(NT) finalize_kernel(FinalArgs f) {
  __shared__ double red[NT];
  finalize_block<NT>(f, red);
}

}  // namespace csg
