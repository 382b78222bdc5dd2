// Register-resident mixed-radix Stockham FFT engine for gfx950.
//
// A line of LEN complex values is owned by TPL = LEN/E threads; thread t holds
// the E elements t + TPL*j (j = 0..E-1) in registers v[j] ("canonical"
// layout: lanes with consecutive t touch consecutive elements, so global
// loads and stores of whole lines coalesce).  A plan of radices R_0..R_{S-1}
// (each dividing E, product LEN) is applied as Stockham stages: stage s with
// radix R and stride product P = R_0...R_{s-1} maps butterfly
// i = t + TPL*m (m < E/R), k = i mod P, to
//   x_r = in[i + r*LEN/R] = v[m + r*E/R],  x_r *= W_{PR}^{rk},  y = DFT_R(x),
//   out[(i-k)*R + k + q*P] = y_q.
// Between stages the outputs go through a padded LDS line buffer and the next
// stage reads its canonical inputs back; the last stage has P = LEN/R, k = i,
// so its outputs land directly in canonical order in registers.  Inverse
// transforms use the reversed plan, so a forward transform's outputs are
// exactly the inputs of a following inverse transform (and vice versa):
// pointwise work between them needs no data movement.
//
// Twiddles W_NTAB^e come from a half table tw[m] = exp(-2*pi*i*m/NTAB),
// 0 <= m < NTAB/2 (W^e = -tw[e - NTAB/2] above), rounded once from extended
// precision on the host; a line of length LEN uses W_LEN^e = W_NTAB^{e*NTAB/LEN}.
// DIR = -1: forward, unnormalised; DIR = +1: inverse without 1/LEN.
//
// The index helpers and in-register DFTs are host-callable so the plan and
// stage indexing can be checked on the CPU.
#pragma once
#include <hip/hip_runtime.h>

#define CSG_HD __host__ __device__ __forceinline__

namespace csg {

template <class T>
struct alignas(2 * sizeof(T)) cplx {
  T x, y;
};

template <class T> CSG_HD cplx<T> mk(T x, T y) { return cplx<T>{x, y}; }
template <class T> CSG_HD cplx<T> operator+(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T> CSG_HD cplx<T> operator-(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T> CSG_HD cplx<T> operator*(T s, cplx<T> a) { return {s * a.x, s * a.y}; }
template <class T> CSG_HD cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class T> CSG_HD cplx<T> conj(cplx<T> a) { return {a.x, -a.y}; }
template <class T> CSG_HD cplx<T> mul_i(cplx<T> a) { return {-a.y, a.x}; }
template <class T> CSG_HD cplx<T> mul_mi(cplx<T> a) { return {a.y, -a.x}; }
// a * W_4 with W_4 = exp(DIR * i * pi / 2)
template <int DIR, class T> CSG_HD cplx<T> rot(cplx<T> a) {
  if constexpr (DIR < 0) return mul_mi(a); else return mul_i(a);
}
template <class T> CSG_HD cplx<T> cvt(cplx<double> a) { return {static_cast<T>(a.x), static_cast<T>(a.y)}; }
template <class T> CSG_HD cplx<T> cvt(cplx<float> a) { return {static_cast<T>(a.x), static_cast<T>(a.y)}; }

// ---- in-register DFTs (x[r * STRIDE], r < R) --------------------------------
template <int DIR, int ST, class T> CSG_HD void dft2(cplx<T>* x) {
  const cplx<T> a = x[0], b = x[ST];
  x[0] = a + b;
  x[ST] = a - b;
}

template <int DIR, class T> CSG_HD void dft4v(cplx<T>& x0, cplx<T>& x1, cplx<T>& x2, cplx<T>& x3) {
  const cplx<T> t0 = x0 + x2, t1 = x0 - x2, t2 = x1 + x3, t3 = rot<DIR>(x1 - x3);
  x0 = t0 + t2;
  x1 = t1 + t3;
  x2 = t0 - t2;
  x3 = t1 - t3;
}

template <int DIR, int ST, class T> CSG_HD void dft4(cplx<T>* x) { dft4v<DIR>(x[0], x[ST], x[2 * ST], x[3 * ST]); }

// multiply by W_8^q (q = 1, 3) for direction DIR
template <int DIR, class T> CSG_HD cplx<T> w8_1(cplx<T> o) {
  const T s = static_cast<T>(0.70710678118654752440084436210484903928);
  if constexpr (DIR < 0) return {s * (o.x + o.y), s * (o.y - o.x)};
  else return {s * (o.x - o.y), s * (o.y + o.x)};
}
template <int DIR, class T> CSG_HD cplx<T> w8_3(cplx<T> o) {
  const T s = static_cast<T>(0.70710678118654752440084436210484903928);
  if constexpr (DIR < 0) return {s * (o.y - o.x), -s * (o.x + o.y)};
  else return {-s * (o.x + o.y), s * (o.x - o.y)};
}

template <int DIR, int ST, class T> CSG_HD void dft8(cplx<T>* x) {
  cplx<T> e0 = x[0], e1 = x[2 * ST], e2 = x[4 * ST], e3 = x[6 * ST];
  cplx<T> o0 = x[ST], o1 = x[3 * ST], o2 = x[5 * ST], o3 = x[7 * ST];
  dft4v<DIR>(e0, e1, e2, e3);
  dft4v<DIR>(o0, o1, o2, o3);
  o1 = w8_1<DIR>(o1);
  o2 = rot<DIR>(o2);
  o3 = w8_3<DIR>(o3);
  x[0] = e0 + o0; x[4 * ST] = e0 - o0;
  x[ST] = e1 + o1; x[5 * ST] = e1 - o1;
  x[2 * ST] = e2 + o2; x[6 * ST] = e2 - o2;
  x[3 * ST] = e3 + o3; x[7 * ST] = e3 - o3;
}

// radix 16 as 4 x 4: x[n1 + 4 n2] -> X[k2 + 4 k1]... computed in place with
// the natural output order y[q] at x[q * ST].
template <int DIR, int ST, class T> CSG_HD void dft16(cplx<T>* x) {
  const T c1 = static_cast<T>(0.92387953251128675612818318939678828682);  // cos(pi/8)
  const T s1 = static_cast<T>(0.38268343236508977172845998403039886676);  // sin(pi/8)
  const T h = static_cast<T>(0.70710678118654752440084436210484903928);
  cplx<T> a[16];
#pragma unroll
  for (int n = 0; n < 16; ++n) a[n] = x[n * ST];
  // first pass: DFT4 over n2 for each n1 (elements n1, n1+4, n1+8, n1+12)
#pragma unroll
  for (int n1 = 0; n1 < 4; ++n1) dft4v<DIR>(a[n1], a[n1 + 4], a[n1 + 8], a[n1 + 12]);
  // twiddles W_16^{n1 * k2}: a[n1 + 4 k2] *= W16^(n1 k2)
  // W16^1 = (c1, DIR*s1), W16^2 = (h, DIR*h), W16^3 = (s1, DIR*c1), W16^4 = (0, DIR), ...
  const T d = static_cast<T>(DIR);
  const cplx<T> w1{c1, d * s1}, w2{h, d * h}, w3{s1, d * c1}, w6{-h, d * h}, w9{-c1, -d * s1};
  a[1 + 4] = cmul(a[1 + 4], w1);
  a[1 + 8] = cmul(a[1 + 8], w2);
  a[1 + 12] = cmul(a[1 + 12], w3);
  a[2 + 4] = cmul(a[2 + 4], w2);
  a[2 + 8] = rot<DIR>(a[2 + 8]);
  a[2 + 12] = cmul(a[2 + 12], w6);
  a[3 + 4] = cmul(a[3 + 4], w3);
  a[3 + 8] = cmul(a[3 + 8], w6);
  a[3 + 12] = cmul(a[3 + 12], w9);
  // second pass: DFT4 over n1 for each k2 -> X[k2 + 4 k1]
#pragma unroll
  for (int k2 = 0; k2 < 4; ++k2) dft4v<DIR>(a[4 * k2], a[4 * k2 + 1], a[4 * k2 + 2], a[4 * k2 + 3]);
#pragma unroll
  for (int k2 = 0; k2 < 4; ++k2)
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) x[(k2 + 4 * k1) * ST] = a[4 * k2 + k1];
}

// radix 3: W_3 = exp(DIR * 2 pi i / 3) = -1/2 + DIR i sqrt(3)/2
template <int DIR, int ST, class T> CSG_HD void dft3(cplx<T>* x) {
  const T s = static_cast<T>(DIR * 0.86602540378443864676372317075293618347);  // DIR sqrt(3)/2
  const cplx<T> a = x[0], b = x[ST], c = x[2 * ST];
  const cplx<T> t = b + c;
  const cplx<T> m = a - static_cast<T>(0.5) * t;      // a + (b + c) Re W_3
  const cplx<T> d = s * mul_i(b - c);                 // (b - c) i Im W_3
  x[0] = a + t;
  x[ST] = m + d;
  x[2 * ST] = m - d;
}

template <int R, int DIR, int ST, class T> CSG_HD void dft(cplx<T>* x) {
  if constexpr (R == 2) dft2<DIR, ST>(x);
  else if constexpr (R == 3) dft3<DIR, ST>(x);
  else if constexpr (R == 4) dft4<DIR, ST>(x);
  else if constexpr (R == 8) dft8<DIR, ST>(x);
  else dft16<DIR, ST>(x);
}

// ---- plans -----------------------------------------------------------------
// Elements per thread E (the register tile; the largest radix is the largest
// power of two <= 16 dividing E).  Default: 16 for LEN in {16, 128, 256}, 8
// otherwise (LEN <= 8: LEN); lengths 3 * 2^m use 12 and one radix-3 stage
// first.  Kernels may pick another E per pass (any divisor of LEN, a multiple
// of 3 for 3 * 2^m).
constexpr int fft_E(int len) {
  return (len % 3 == 0) ? (len < 12 ? len : 12)
         : len <= 8 ? len : ((len == 16 || len == 128 || len == 256) ? 16 : 8);
}
// A valid tile leaves a radix >= 2 for every stage: E >= 2 dividing LEN, and
// for 3 * 2^m lines a multiple of 3 that is even unless LEN = 3.
constexpr bool fft_E_valid(int len, int e) {
  return e >= 2 && e <= len && len % e == 0 && e <= 32 &&
         (len % 3 != 0 || (e % 3 == 0 && (len == 3 || e % 2 == 0)));
}
// Radix of forward stage s: 3 while 3 divides what remains, then the largest
// power of two <= 16 dividing both E and what remains.
constexpr int fft_radix(int len, int s, int e) {
  int rem = len;
  int r = 1;
  for (int k = 0; k <= s; ++k) {
    if (rem % 3 == 0) {
      r = 3;
    } else {
      r = 16;
      while (r > e || rem % r != 0 || e % r != 0) r >>= 1;
    }
    rem /= r;
  }
  return r;
}
constexpr int fft_nstages(int len, int e) {
  int n = 0, rem = len;
  while (rem > 1) {
    rem /= fft_radix(len, n, e);
    ++n;
  }
  return n;
}
// radix / stride product of stage s of the forward (INV = false) or inverse
// (reversed) plan
constexpr int plan_radix(int len, int s, bool inv, int e) {
  return inv ? fft_radix(len, fft_nstages(len, e) - 1 - s, e) : fft_radix(len, s, e);
}
constexpr int plan_pstride(int len, int s, bool inv, int e) {
  int p = 1;
  for (int k = 0; k < s; ++k) p *= plan_radix(len, k, inv, e);
  return p;
}

// LDS slot of element e of a line buffer of LEN elements (buffers hold
// padded_len<T>(LEN) slots).  Lines of up to SWZ_MAX_LEN elements: an XOR
// swizzle that permutes aligned blocks of 8 (16-byte elements: low 3 bits
// ^= bits 3..5) or 32 (8-byte elements: bits 0..4 ^= a mix of bits 4..7);
// longer lines: one pad slot per 128 bytes of line data.  Chosen from
// bank-conflict counts of ds_write/ds_read b128 and b64 on MI355X over the
// exchange patterns of the radix-4/8 stages and the consecutive reads, then
// by the step time (1024^2 FP64: 61 % fewer conflict cycles); the pad slots
// suit the radix-8/16 stages of long lines.
constexpr int SWZ_MAX_LEN = 1024;
template <class T> constexpr int pad_shift() { return sizeof(T) == 8 ? 3 : 4; }
template <class T, int LEN> CSG_HD int pad(int e) {
  if constexpr (LEN <= SWZ_MAX_LEN) {
    if constexpr (sizeof(T) == 8)
      return e ^ ((e >> 3) & 7);
    else
      return e ^ (((e >> 4) & 1) * 5) ^ (((e >> 5) & 1) * 10) ^ (((e >> 6) & 1) * 12) ^ (((e >> 7) & 1) * 16);
  } else {
    return e + (e >> pad_shift<T>());
  }
}
template <class T> constexpr int padded_len(int len) { return len + (len >> pad_shift<T>()) + 1; }

// Row stride (slots) of a tile of b rows of len elements whose transposed
// loads and stores put the b rows of one element in neighbouring lanes.  For
// 8-byte elements the rows' base slots are spread over the banks, stride =
// 16/b modulo 16, the conflict-free strides measured for b = 2..8 (1024^2
// mixed step -2.4 %).  16-byte elements keep padded_len: their conflict-free
// strides (8/b modulo 16) made the 1024^2 FP64 step no faster.  Rows longer
// than SWZ_MAX_LEN keep padded_len.
template <class T>
constexpr int tile_stride(int len, int b) {
  const int base = padded_len<T>(len);
  const int per = 16 / (b > 0 ? b : 1);
  if (sizeof(T) == 8 || len > SWZ_MAX_LEN || b < 2 || per < 1) return base;
  return base + ((per - base % 16) % 16 + 16) % 16;
}

// Index of -k on a periodic line of length N.
template <int N> CSG_HD int neg_index(int k) { return k == 0 ? 0 : N - k; }

// Stockham output position of value q of butterfly i at stride product P.
CSG_HD int stockham_out(int i, int q, int R, int P) {
  const int k = i % P;
  return (i - k) * R + k + q * P;
}

// ---- stage twiddle tables ----------------------------------------------------
// The twiddles of a plan laid out per stage so that the lanes of a stage read
// consecutive entries (no LDS bank conflicts): for stage s >= 1 with radix R
// and stride product P, entry (P - 1) + (r - 1) P + k holds
// W_LEN^(r k LEN / (P R)) with the forward sign, r in [1, R), k in [0, P).
// The offsets telescope to LEN - 1 entries; stage 0's (P = 1) are all 1 and
// go unused.
// The inverse plan (radices reversed) has its own table after the forward one
// unless both plans are the same.
constexpr bool plan_symmetric(int len, int e) {
  for (int s = 0; s < fft_nstages(len, e); ++s)
    if (plan_radix(len, s, false, e) != plan_radix(len, s, true, e)) return false;
  return true;
}
constexpr int stab_len(int len, int e) { return (len - 1) * (plan_symmetric(len, e) ? 1 : 2); }
constexpr int stab_inv_off(int len, int e) { return plan_symmetric(len, e) ? 0 : len - 1; }

// Exponent (units of W_LEN) of entry q of the table of plan INV.
template <int LEN, int E, bool INV, int S = 0>
CSG_HD int stab_exp(int q) {
  if constexpr (S >= fft_nstages(LEN, E)) {
    return 0;
  } else {
    constexpr int P = plan_pstride(LEN, S, INV, E), R = plan_radix(LEN, S, INV, E);
    if (q < P * R - 1) {
      const int l = q - (P - 1);
      return (1 + l / P) * (l % P) * (LEN / (P * R));
    }
    return stab_exp<LEN, E, INV, S + 1>(q);
  }
}

// The same for run-time (len, e): host code fills the tables with it.
constexpr int stab_exp_rt(int len, int e, bool inv, int q) {
  for (int s = 0; s < fft_nstages(len, e); ++s) {
    const int p = plan_pstride(len, s, inv, e), r = plan_radix(len, s, inv, e);
    if (q < p * r - 1) {
      const int l = q - (p - 1);
      return (1 + l / p) * (l % p) * (len / (p * r));
    }
  }
  return 0;
}

// Twiddle buffer of an N-point grid: W_N^m for m in [0, N), then the stage
// tables (forward plan, and the inverse plan's when it differs) of every
// (LEN, E) a kernel may use: LEN in {N, N/2} and every valid E of the
// stab_e_at list.  Kernels copy their tables from the buffer to LDS with
// contiguous loads.
constexpr int stab_e_count = 8;
constexpr int stab_e_at(int i) {
  constexpr int es[stab_e_count] = {2, 4, 6, 8, 12, 16, 24, 32};
  return es[i];
}
constexpr bool stab_e_ok(int len, int e) { return len >= 2 && fft_E_valid(len, e); }
// offset (in entries) of the (len, e) tables in the buffer of an n-point grid; -1 if none
constexpr int stab_offset(int n, int len, int e) {
  int off = n;
  for (int h = 0; h < 2; ++h) {
    const int l = h == 0 ? n : n / 2;
    for (int i = 0; i < stab_e_count; ++i) {
      const int x = stab_e_at(i);
      if (!stab_e_ok(l, x)) continue;
      if (l == len && x == e) return off;
      off += stab_len(l, x);
    }
  }
  return -1;
}
constexpr int twiddle_len(int n) {
  int off = n;
  for (int h = 0; h < 2; ++h) {
    const int l = h == 0 ? n : n / 2;
    for (int i = 0; i < stab_e_count; ++i)
      if (stab_e_ok(l, stab_e_at(i))) off += stab_len(l, stab_e_at(i));
  }
  return off;
}

// ---- device side -------------------------------------------------------------

// W_NTAB^e (0 <= e < NTAB) for direction DIR from the half table.
template <int DIR, int NTAB, class T>
CSG_HD cplx<T> tw_at(const cplx<T>* tw, int e) {
  constexpr int H = NTAB / 2;
  const bool hi = e >= H;
  cplx<T> w = tw[hi ? e - H : e];
  if (hi) w = cplx<T>{-w.x, -w.y};
  if constexpr (DIR > 0) w.y = -w.y;
  return w;
}

// Stage the half table in LDS: all loads issued, then all LDS stores.  Loads
// and stores are unconditional, with the index clamped into the table (lanes
// past its end store the last entry's own value again): a predicated store
// becomes a branch that the compiler fills with its load and a wait for every
// load issued before it, which serialises the kernel's first memory accesses.
template <int NTAB, int NT, class T>
__device__ __forceinline__ void load_twiddles(cplx<T>* lds_tw, const cplx<T>* __restrict__ g) {
  constexpr int H = NTAB / 2, K = (H + NT - 1) / NT;
  cplx<T> w[K];
#pragma unroll
  for (int r = 0; r < K; ++r) {
    const int m = static_cast<int>(threadIdx.x) + NT * r;
    w[r] = g[m < H ? m : H - 1];
  }
#pragma unroll
  for (int r = 0; r < K; ++r) {
    const int m = static_cast<int>(threadIdx.x) + NT * r;
    lds_tw[m < H ? m : H - 1] = w[r];
  }
}

// Copy the stage table(s) of the (LEN, E) plans of an N-point grid's
// twiddle buffer g (twiddle_len(N) entries) to LDS: all loads issued, then
// the LDS stores, both clamped and unconditional, as in load_twiddles.
template <int N, int LEN, int E, int NT, class T>
__device__ __forceinline__ void load_stage_tables(cplx<T>* st, const cplx<T>* __restrict__ g) {
  constexpr int Q = stab_len(LEN, E), K = (Q + NT - 1) / NT, OFF = stab_offset(N, LEN, E);
  static_assert(OFF >= N, "no stage table for this plan in the twiddle buffer");
  cplx<T> w[K];
#pragma unroll
  for (int r = 0; r < K; ++r) {
    const int q = static_cast<int>(threadIdx.x) + NT * r;
    w[r] = g[OFF + (q < Q ? q : Q - 1)];
  }
#pragma unroll
  for (int r = 0; r < K; ++r) {
    const int q = static_cast<int>(threadIdx.x) + NT * r;
    st[q < Q ? q : Q - 1] = w[r];
  }
}

// Workgroup barrier for LDS traffic only: the wave's LDS accesses complete
// (s_waitcnt lgkmcnt(0); the vmcnt and expcnt fields are left at their
// no-wait maxima), then s_barrier.  Wavefront-scope fences keep the compiler
// from moving memory accesses across it.  Unlike __syncthreads() it does not
// wait for the wave's outstanding global loads, so loads issued before an
// exchange (a prefetched phi row or state line) stay in flight through it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xC07F);
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Order LDS traffic between stages: lines that fit in one wavefront only need
// the wave's own accesses ordered; wider lines need the workgroup barrier.
template <bool WAVE>
__device__ __forceinline__ void lds_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    lds_barrier();
  }
}

// Butterflies of one stage for one array, in registers.  ST: `tw` is the
// stage table of this direction's plan; otherwise the half table of W_NTAB.
template <class T, int DIR, int LEN, int TPL, int E, int R, int P, int NTAB, bool ST = false>
CSG_HD void stage_compute(cplx<T> (&v)[E], int t, const cplx<T>* tw) {
  constexpr int NBT = E / R;  // butterflies per thread
  constexpr int TM = NTAB / LEN;
#pragma unroll
  for (int m = 0; m < NBT; ++m) {
    if constexpr (P > 1) {
      const int i = t + TPL * m;
      const int k = i % P;
      constexpr int STEP = LEN / (P * R);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cplx<T> w;
        if constexpr (ST) {
          w = tw[(P - 1) + (r - 1) * P + k];
          if constexpr (DIR > 0) w.y = -w.y;
        } else {
          w = tw_at<DIR, NTAB>(tw, r * k * STEP * TM);
        }
        v[m + r * NBT] = cmul(v[m + r * NBT], w);
      }
    }
    dft<R, DIR, NBT>(&v[m]);
  }
}

// Write a non-final stage's outputs to the line buffer, then read the next
// stage's canonical inputs.
template <class T, int LEN, int TPL, int E, int R, int P, bool WAVE>
__device__ __forceinline__ void stage_exchange(cplx<T> (&v)[E], int t, cplx<T>* line) {
  constexpr int NBT = E / R;
#pragma unroll
  for (int m = 0; m < NBT; ++m) {
    const int i = t + TPL * m;
#pragma unroll
    for (int q = 0; q < R; ++q) line[pad<T, LEN>(stockham_out(i, q, R, P))] = v[m + q * NBT];
  }
  lds_sync<WAVE>();
#pragma unroll
  for (int j = 0; j < E; ++j) v[j] = line[pad<T, LEN>(t + TPL * j)];
  // the next exchange overwrites what other lanes read here: without this
  // ordering point the compiler may hoist its stores above these loads
  lds_sync<WAVE>();
}

// Exchange layouts of a transform of NA arrays.  `line` points at array 0's
// buffer (padded_len<T>(LEN) elements); array a's is at line + a * AS.
//   XM = 0: the arrays are exchanged one after another (AS = 0: one buffer
//           for all); two ordering points per array and stage.
//   XM = 1: all arrays written, one ordering point, all read, one more.
//   XM = 2: as 1, but odd stages use a second set of buffers at line + BS, so
//           a stage's reads and the next stage's writes touch different
//           memory and need no ordering point between them; one closing
//           ordering point at the end of the transform frees the buffers.
template <class T, int LEN, int TPL, int E, int R, int P, bool WAVE, int NA, int XM, int AS, int BS, int S>
__device__ __forceinline__ void stage_exchange_all(cplx<T> (&v)[NA][E], int t, cplx<T>* line) {
  if constexpr (XM == 0) {
#pragma unroll
    for (int a = 0; a < NA; ++a) stage_exchange<T, LEN, TPL, E, R, P, WAVE>(v[a], t, line + a * AS);
  } else {
    constexpr int NBT = E / R;
    cplx<T>* buf = line + ((XM == 2 && (S & 1)) ? BS : 0);
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int m = 0; m < NBT; ++m) {
        const int i = t + TPL * m;
#pragma unroll
        for (int q = 0; q < R; ++q) buf[a * AS + pad<T, LEN>(stockham_out(i, q, R, P))] = v[a][m + q * NBT];
      }
    lds_sync<WAVE>();
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int j = 0; j < E; ++j) v[a][j] = buf[a * AS + pad<T, LEN>(t + TPL * j)];
    if constexpr (XM == 1) lds_sync<WAVE>();
  }
}

template <class T, int DIR, int LEN, int NA, int S, int NTAB, bool WAVE, int E, int XM, int AS, int BS, bool ST>
__device__ __forceinline__ void fft_stages(cplx<T> (&v)[NA][E], int t, cplx<T>* line, const cplx<T>* tw) {
  constexpr bool INV = DIR > 0;
  constexpr int NS = fft_nstages(LEN, E);
  if constexpr (S < NS) {
    constexpr int TPL = LEN / E;
    constexpr int R = plan_radix(LEN, S, INV, E), P = plan_pstride(LEN, S, INV, E);
#pragma unroll
    for (int a = 0; a < NA; ++a) stage_compute<T, DIR, LEN, TPL, E, R, P, NTAB, ST>(v[a], t, tw);
    if constexpr (S + 1 < NS) stage_exchange_all<T, LEN, TPL, E, R, P, WAVE, NA, XM, AS, BS, S>(v, t, line);
    fft_stages<T, DIR, LEN, NA, S + 1, NTAB, WAVE, E, XM, AS, BS, ST>(v, t, line, tw);
  } else if constexpr (XM == 2 && NS > 1) {
    lds_sync<WAVE>();
  }
}

// Transform NA lines held in canonical registers by this thread group (TPL =
// LEN/E threads, t = index within the group), using the group's LDS line
// buffers at `line` (layout XM, array stride AS, buffer-set stride BS; each
// buffer padded_len<T>(LEN) elements) for the exchanges.  Twiddles: ST, the
// stage tables of the (LEN, E) plans (load_stage_tables); otherwise the half
// table of W_NTAB.
template <class T, int DIR, int LEN, int NA, int NTAB, bool WAVE, int E = fft_E(LEN), int XM = 0, int AS = 0, int BS = 0,
          bool ST = false>
__device__ __forceinline__ void fft_regs(cplx<T> (&v)[NA][E], int t, cplx<T>* line, const cplx<T>* tw) {
  static_assert(fft_E_valid(LEN, E), "E must divide LEN (a multiple of 3 for 3 * 2^m lines)");
  static_assert(XM == 0 || NA == 1 || AS >= padded_len<T>(LEN), "merged exchanges need one buffer per array");
  static_assert(XM != 2 || BS >= (NA - 1) * AS + padded_len<T>(LEN), "second buffer set after the first");
  constexpr int IO = (ST && DIR > 0) ? stab_inv_off(LEN, E) : 0;   // folded: the plan search is not cheap
  const cplx<T>* t0 = tw + IO;
  fft_stages<T, DIR, LEN, NA, 0, NTAB, WAVE, E, XM, AS, BS, ST>(v, t, line, t0);
}

// ---- workgroup reductions (wave shuffles, one barrier) ---------------------
enum RedOp { RED_SUM = 0, RED_MIN = 1, RED_MAX = 2 };

template <int OP> __device__ __forceinline__ double red_combine(double a, double b) {
  if constexpr (OP == RED_SUM) return a + b;
  // NaN-propagating min/max
  else if constexpr (OP == RED_MIN) return (a != a || b != b) ? a + b : fmin(a, b);
  else return (a != a || b != b) ? a + b : fmax(a, b);
}

// Reduce over the W = min(NT, 64) active lanes of a wavefront (NT a power of
// two): partners lane ^ o stay below W, so lanes of a workgroup narrower than
// a wavefront never read from lanes that do not run.
template <int OP, int W> __device__ __forceinline__ double wave_reduce(double v) {
#pragma unroll
  for (int o = W / 2; o > 0; o >>= 1) v = red_combine<OP>(v, __shfl_xor(v, o, 64));
  return v;
}

// Reduce K values across the workgroup; every thread gets the results.
// `scratch` must hold K * (NT/64) doubles.  Starts and ends with a barrier.
template <int NT, int K>
__device__ __forceinline__ void block_reduce(double (&v)[K], const int (&ops)[K], double* scratch) {
  constexpr int NW = (NT + 63) / 64;
  constexpr int W = NT < 64 ? NT : 64;
  static_assert((W & (W - 1)) == 0 && (NT <= 64 || NT % 64 == 0), "whole wavefronts or one power-of-two part");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    if (ops[q] == RED_SUM) v[q] = wave_reduce<RED_SUM, W>(v[q]);
    else if (ops[q] == RED_MIN) v[q] = wave_reduce<RED_MIN, W>(v[q]);
    else v[q] = wave_reduce<RED_MAX, W>(v[q]);
  }
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < K; ++q) scratch[q * NW + wave] = v[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < K; ++q) {
    double a = scratch[q * NW];
    for (int w = 1; w < NW; ++w) {
      const double b = scratch[q * NW + w];
      a = (ops[q] == RED_SUM) ? a + b : (ops[q] == RED_MIN ? red_combine<RED_MIN>(a, b) : red_combine<RED_MAX>(a, b));
    }
    v[q] = a;
  }
  __syncthreads();
}

// Reduce K values across the workgroup into thread 0, in the same order as
// block_reduce (bit-identical results); the other threads' values are left
// unspecified.  One barrier (none for a single wavefront) and no barrier at
// the end: for the partials a pass writes once, when it is done with
// `scratch` (a caller that reuses the LDS afterwards synchronises first).
template <int NT, int K>
__device__ __forceinline__ void block_reduce0(double (&v)[K], const int (&ops)[K], double* scratch) {
  constexpr int NW = (NT + 63) / 64;
  constexpr int W = NT < 64 ? NT : 64;
  static_assert((W & (W - 1)) == 0 && (NT <= 64 || NT % 64 == 0), "whole wavefronts or one power-of-two part");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    if (ops[q] == RED_SUM) v[q] = wave_reduce<RED_SUM, W>(v[q]);
    else if (ops[q] == RED_MIN) v[q] = wave_reduce<RED_MIN, W>(v[q]);
    else v[q] = wave_reduce<RED_MAX, W>(v[q]);
  }
  if constexpr (NW > 1) {
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < K; ++q) scratch[q * NW + wave] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
      for (int q = 0; q < K; ++q) {
        double a = scratch[q * NW];
        for (int w = 1; w < NW; ++w) {
          const double b = scratch[q * NW + w];
          a = (ops[q] == RED_SUM) ? a + b : (ops[q] == RED_MIN ? red_combine<RED_MIN>(a, b) : red_combine<RED_MAX>(a, b));
        }
        v[q] = a;
      }
    }
  }
}

}  // namespace csg
