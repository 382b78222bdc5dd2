// LDS-resident mixed-radix Stockham FFT engine for gfx950.
//
// A workgroup owns NLINES independent lines of LEN complex values stored in
// LDS with a line pitch of PITCH elements.  Each stage reads R values per
// butterfly from LDS into registers, applies the stage twiddles and an
// in-register radix-R DFT, and writes the results back in autosorted
// (Stockham) order, so the output is in natural order and no bit reversal is
// needed.  Stage s with radix R and stride product P (product of the radices
// of the previous stages) maps butterfly i (0 <= i < LEN/R), k = i mod P, to
//   x_r = in[i + r*LEN/R],  x_r *= W_{P*R}^{r*k},  y = DFT_R(x),
//   out[(i-k)*R + k + q*P] = y_q.
// Radices are 8 while LEN allows, then 4 or 2; the workgroup is sized to
// NLINES*LEN/8 threads so a radix-8 stage is one butterfly per thread.
//
// Twiddles come from a table tw[m] = exp(-2*pi*i*m/NTAB) (host-computed in
// extended precision and rounded once); W_LEN^e = tw[e * (NTAB/LEN)].
// DIR = -1 is the forward (exp(-i...), unnormalised) transform, DIR = +1 the
// inverse without the 1/LEN factor.
#pragma once
#include <hip/hip_runtime.h>

namespace csg {

template <class T>
struct alignas(2 * sizeof(T)) cplx {
  T x, y;
};

template <class T> __device__ __forceinline__ cplx<T> mk(T x, T y) { return cplx<T>{x, y}; }
template <class T> __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T> __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T> __device__ __forceinline__ cplx<T> operator*(T s, cplx<T> a) { return {s * a.x, s * a.y}; }
template <class T> __device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class T> __device__ __forceinline__ cplx<T> conj(cplx<T> a) { return {a.x, -a.y}; }
// a * (i)  and  a * (-i)
template <class T> __device__ __forceinline__ cplx<T> mul_i(cplx<T> a) { return {-a.y, a.x}; }
template <class T> __device__ __forceinline__ cplx<T> mul_mi(cplx<T> a) { return {a.y, -a.x}; }
// a * (DIR * i): the radix-4 rotation W_4^{-DIR}... i.e. W_4 = exp(DIR*i*pi/2)
template <int DIR, class T> __device__ __forceinline__ cplx<T> rot(cplx<T> a) {
  if constexpr (DIR < 0) return mul_mi(a); else return mul_i(a);
}

template <class T> __device__ __forceinline__ cplx<T> cvt(cplx<double> a) {
  return {static_cast<T>(a.x), static_cast<T>(a.y)};
}
template <class T> __device__ __forceinline__ cplx<T> cvt(cplx<float> a) {
  return {static_cast<T>(a.x), static_cast<T>(a.y)};
}

// ---- in-register DFTs ------------------------------------------------------
template <int DIR, class T> __device__ __forceinline__ void dft2(cplx<T>* x) {
  cplx<T> a = x[0], b = x[1];
  x[0] = a + b;
  x[1] = a - b;
}

template <int DIR, class T> __device__ __forceinline__ void dft4(cplx<T>& x0, cplx<T>& x1, cplx<T>& x2, cplx<T>& x3) {
  cplx<T> t0 = x0 + x2, t1 = x0 - x2, t2 = x1 + x3, t3 = rot<DIR>(x1 - x3);
  x0 = t0 + t2;
  x1 = t1 + t3;
  x2 = t0 - t2;
  x3 = t1 - t3;
}

template <int DIR, class T> __device__ __forceinline__ void dft4(cplx<T>* x) { dft4<DIR>(x[0], x[1], x[2], x[3]); }

template <int DIR, class T> __device__ __forceinline__ void dft8(cplx<T>* x) {
  const T s = static_cast<T>(0.70710678118654752440084436210484903928);
  cplx<T> e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
  cplx<T> o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
  dft4<DIR>(e0, e1, e2, e3);
  dft4<DIR>(o0, o1, o2, o3);
  // o_q *= W_8^q with W_8 = exp(DIR * i * pi / 4)
  if constexpr (DIR < 0) {
    o1 = cplx<T>{s * (o1.x + o1.y), s * (o1.y - o1.x)};
    o3 = cplx<T>{s * (o3.y - o3.x), -s * (o3.x + o3.y)};
  } else {
    o1 = cplx<T>{s * (o1.x - o1.y), s * (o1.y + o1.x)};
    o3 = cplx<T>{-s * (o3.x + o3.y), s * (o3.x - o3.y)};
  }
  o2 = rot<DIR>(o2);
  x[0] = e0 + o0; x[4] = e0 - o0;
  x[1] = e1 + o1; x[5] = e1 - o1;
  x[2] = e2 + o2; x[6] = e2 - o2;
  x[3] = e3 + o3; x[7] = e3 - o3;
}

template <int R, int DIR, class T> __device__ __forceinline__ void dft(cplx<T>* x) {
  if constexpr (R == 2) dft2<DIR>(x);
  else if constexpr (R == 4) dft4<DIR>(x);
  else dft8<DIR>(x);
}

template <int REM> constexpr int pick_radix() { return (REM % 8 == 0) ? 8 : ((REM % 4 == 0) ? 4 : 2); }

// Twiddle W_{LEN}^{e} (direction DIR) from the table; TM = NTAB / LEN.
template <int DIR, int TM, class T>
__device__ __forceinline__ cplx<T> twiddle(const cplx<T>* __restrict__ tw, int e) {
  cplx<T> w = tw[e * TM];
  if constexpr (DIR > 0) w.y = -w.y;
  return w;
}

// One Stockham stage over all lines resident in LDS.
template <class T, int DIR, int LEN, int R, int P, int NLINES, int PITCH, int NT, int TM>
__device__ __forceinline__ void stockham_stage(cplx<T>* buf, const cplx<T>* __restrict__ tw) {
  constexpr int NB = LEN / R;
  constexpr int TOT = NB * NLINES;
  constexpr int PER = (TOT + NT - 1) / NT;
  cplx<T> v[PER][R];
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int f = threadIdx.x + t * NT;
    if (TOT % NT == 0 || f < TOT) {
      const int line = f / NB, i = f - line * NB;
      const cplx<T>* L = buf + line * PITCH;
#pragma unroll
      for (int r = 0; r < R; ++r) v[t][r] = L[i + r * NB];
      if constexpr (P > 1) {
        const int k = i % P;
        constexpr int STEP = LEN / (P * R);
#pragma unroll
        for (int r = 1; r < R; ++r) v[t][r] = cmul(v[t][r], twiddle<DIR, TM>(tw, r * k * STEP));
      }
      dft<R, DIR>(v[t]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int f = threadIdx.x + t * NT;
    if (TOT % NT == 0 || f < TOT) {
      const int line = f / NB, i = f - line * NB;
      const int k = i % P;
      cplx<T>* L = buf + line * PITCH;
      const int j = (i - k) * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) L[j + q * P] = v[t][q];
    }
  }
  __syncthreads();
}

template <class T, int DIR, int LEN, int P, int NLINES, int PITCH, int NT, int TM>
__device__ __forceinline__ void stockham_from(cplx<T>* buf, const cplx<T>* __restrict__ tw) {
  if constexpr (P < LEN) {
    constexpr int R = pick_radix<LEN / P>();
    stockham_stage<T, DIR, LEN, R, P, NLINES, PITCH, NT, TM>(buf, tw);
    stockham_from<T, DIR, LEN, P * R, NLINES, PITCH, NT, TM>(buf, tw);
  }
}

// Full transform of every line. Requires a barrier before the call (data in
// LDS complete); ends with a barrier.
template <class T, int DIR, int LEN, int NLINES, int PITCH, int NT, int TM>
__device__ __forceinline__ void fft_lines(cplx<T>* buf, const cplx<T>* __restrict__ tw) {
  stockham_from<T, DIR, LEN, 1, NLINES, PITCH, NT, TM>(buf, tw);
}

// ---- real <-> packed half spectrum (rows of N = 2L reals) ------------------
// Packed format of the R2C transform of a real row x[0..N): X[k] for
// 1 <= k < L as usual, slot 0 holds (X[0], X[L]) as (re, im) -- both are real.

// Forward post-processing: z = FFT_L(x as L complex) -> packed X (in place).
// Each (k, L-k) pair is owned by one thread; requires the FFT output complete.
template <class T, int N, int NLINES, int PITCH, int NT>
__device__ __forceinline__ void r2c_post(cplx<T>* buf, const cplx<T>* __restrict__ tw) {
  constexpr int L = N / 2;
  constexpr int NP = L / 2 + 1;  // pairs k = 0 .. L/2
  const T half = static_cast<T>(0.5);
  for (int f = threadIdx.x; f < NLINES * NP; f += NT) {
    const int line = f / NP, k = f - line * NP;
    cplx<T>* X = buf + line * PITCH;
    if (k == 0) {
      const cplx<T> z = X[0];
      X[0] = mk(z.x + z.y, z.x - z.y);
    } else {
      const int m = L - k;
      const cplx<T> zk = X[k], zm = X[m];
      // E_k = (Z_k + conj Z_m)/2, O_k = (Z_k - conj Z_m)/(2i)
      const cplx<T> ek = half * (zk + conj(zm));
      const cplx<T> ok = half * mul_mi(zk - conj(zm));
      const cplx<T> wk = tw[k];  // W_N^k
      const cplx<T> xk = ek + cmul(wk, ok);
      if (m != k) {
        const cplx<T> em = half * (zm + conj(zk));
        const cplx<T> om = half * mul_mi(zm - conj(zk));
        const cplx<T> wm = tw[m];
        X[m] = em + cmul(wm, om);
      }
      X[k] = xk;
    }
  }
  __syncthreads();
}

// Inverse pre-processing: packed X -> Z such that x (as L complex) =
// FFT^{-1}_L(Z) without normalisation; the overall 1/N is folded in via
// `scale`.  In place, pairs owned by one thread.
template <class T, int N, int NLINES, int PITCH, int NT>
__device__ __forceinline__ void c2r_pre(cplx<T>* buf, const cplx<T>* __restrict__ tw, T scale) {
  constexpr int L = N / 2;
  constexpr int NP = L / 2 + 1;
  for (int f = threadIdx.x; f < NLINES * NP; f += NT) {
    const int line = f / NP, k = f - line * NP;
    cplx<T>* X = buf + line * PITCH;
    if (k == 0) {
      const cplx<T> p = X[0];
      X[0] = mk(scale * (p.x + p.y), scale * (p.x - p.y));
    } else {
      const int m = L - k;
      const cplx<T> xk = X[k], xm = X[m];
      // Z_k = (X_k + conj X_m) + i (X_k - conj X_m) W_N^{-k}
      cplx<T> wk = tw[k];
      wk.y = -wk.y;
      const cplx<T> zk = scale * ((xk + conj(xm)) + mul_i(cmul(xk - conj(xm), wk)));
      if (m != k) {
        cplx<T> wm = tw[m];
        wm.y = -wm.y;
        X[m] = scale * ((xm + conj(xk)) + mul_i(cmul(xm - conj(xk), wm)));
      }
      X[k] = zk;
    }
  }
  __syncthreads();
}

}  // namespace csg
