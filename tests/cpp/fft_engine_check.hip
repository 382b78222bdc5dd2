// Host-only check of the register Stockham engine (csg_fft.hpp): drives the
// engine's stage_compute / plan / padding / Stockham output indexing over
// emulated threads of one line and compares forward and inverse transforms
// with a naive DFT for every line length the kernels use (2^m and 3*2^m).
// Each transform runs twice, with twiddles from the half table and from the
// per-stage tables (entries placed by stab_exp, values from the half table):
// both must match the DFT and each other bit for bit.  No GPU needed.
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "../../embedded-spinodal-core-shell-julia_amd/csrc/csg_kernels.hpp"

using namespace csg;

template <int DIR, int LEN, int S, class T, int NTAB, int E, bool ST>
void run_stages(std::vector<std::vector<cplx<T>>>& v, const cplx<T>* tw, std::vector<cplx<T>>& line) {
  constexpr bool INV = DIR > 0;
  constexpr int NS = fft_nstages(LEN, E);
  if constexpr (S < NS) {
    constexpr int TPL = LEN / E;
    constexpr int R = plan_radix(LEN, S, INV, E), P = plan_pstride(LEN, S, INV, E);
    static_assert(E % R == 0, "radix must divide E");
    for (int t = 0; t < TPL; ++t) {
      cplx<T> r[E];
      for (int j = 0; j < E; ++j) r[j] = v[t][j];
      stage_compute<T, DIR, LEN, TPL, E, R, P, NTAB, ST>(r, t, tw);
      for (int j = 0; j < E; ++j) v[t][j] = r[j];
    }
    if constexpr (S + 1 < NS) {
      constexpr int NBT = E / R;
      for (int t = 0; t < TPL; ++t)
        for (int m = 0; m < NBT; ++m)
          for (int q = 0; q < R; ++q) line[pad<T, LEN>(stockham_out(t + TPL * m, q, R, P))] = v[t][m + q * NBT];
      for (int t = 0; t < TPL; ++t)
        for (int j = 0; j < E; ++j) v[t][j] = line[pad<T, LEN>(t + TPL * j)];
    }
    run_stages<DIR, LEN, S + 1, T, NTAB, E, ST>(v, tw, line);
  }
}

// NTAB: length of the twiddle table's full circle (LEN for column lines,
// 2 LEN for the half-length complex transforms of the real row passes).
template <int LEN, class T, int NTAB = LEN, int E = fft_E(LEN)>
int check(double tol) {
  static_assert(fft_E_valid(LEN, E), "E");
  constexpr int TPL = LEN / E;
  std::vector<cplx<T>> tw(NTAB / 2);
  for (int m = 0; m < NTAB / 2; ++m) {
    const long double a = -2.0L * 3.141592653589793238462643383279502884L * m / NTAB;
    tw[m] = cplx<T>{static_cast<T>(cosl(a)), static_cast<T>(sinl(a))};
  }
  std::vector<std::complex<double>> x(LEN);
  unsigned s = 12345u + LEN;
  for (auto& z : x) {
    s = s * 1664525u + 1013904223u;
    double a = (s >> 8) / 16777216.0 - 0.5;
    s = s * 1664525u + 1013904223u;
    double b = (s >> 8) / 16777216.0 - 0.5;
    z = {a, b};
  }
  // stage tables of the forward and inverse plans
  std::vector<cplx<T>> st(stab_len(LEN, E));
  for (int q = 0; q < (int)st.size(); ++q) {
    const int e = q < LEN - 1 ? stab_exp<LEN, E, false>(q) : stab_exp<LEN, E, true>(q - (LEN - 1));
    if (e < 0 || e >= LEN) { printf("LEN=%d E=%d entry %d exponent %d out of range\n", LEN, E, q, e); return 1; }
    // the run-time form the host uses to fill the twiddle buffer
    const int er = q < LEN - 1 ? stab_exp_rt(LEN, E, false, q) : stab_exp_rt(LEN, E, true, q - (LEN - 1));
    if (er != e) { printf("LEN=%d E=%d entry %d: stab_exp_rt %d != stab_exp %d\n", LEN, E, q, er, e); return 1; }
    st[q] = tw_at<-1, NTAB>(tw.data(), e * (NTAB / LEN));
  }
  int bad = 0;
  for (int dir : {-1, +1}) {
    auto load = [&]() {
      std::vector<std::vector<cplx<T>>> v(TPL, std::vector<cplx<T>>(E));
      for (int t = 0; t < TPL; ++t)
        for (int j = 0; j < E; ++j) {
          const auto z = x[t + TPL * j];
          v[t][j] = cplx<T>{static_cast<T>(z.real()), static_cast<T>(z.imag())};
        }
      return v;
    };
    std::vector<std::vector<cplx<T>>> v = load(), vs = load();
    std::vector<cplx<T>> line(padded_len<T>(LEN));
    if (dir < 0) {
      run_stages<-1, LEN, 0, T, NTAB, E, false>(v, tw.data(), line);
      run_stages<-1, LEN, 0, T, NTAB, E, true>(vs, st.data(), line);
    } else {
      run_stages<+1, LEN, 0, T, NTAB, E, false>(v, tw.data(), line);
      run_stages<+1, LEN, 0, T, NTAB, E, true>(vs, st.data() + stab_inv_off(LEN, E), line);
    }
    int diff = 0;
    for (int t = 0; t < TPL; ++t)
      for (int j = 0; j < E; ++j) diff += (v[t][j].x != vs[t][j].x) || (v[t][j].y != vs[t][j].y);
    if (diff) { printf("LEN=%d E=%d dir=%+d: stage-table transform differs in %d values\n", LEN, E, dir, diff); ++bad; }
    double err = 0, mag = 0;
    for (int k = 0; k < LEN; ++k) {
      std::complex<long double> acc = 0;
      for (int n = 0; n < LEN; ++n) {
        const long double ang = dir * 2.0L * 3.141592653589793238462643383279502884L * ((long long)n * k % LEN) / LEN;
        acc += std::complex<long double>(x[n].real(), x[n].imag()) * std::complex<long double>(cosl(ang), sinl(ang));
      }
      const cplx<T> g = v[k % TPL][k / TPL];
      err = std::fmax(err, std::abs(std::complex<double>((double)g.x - (double)acc.real(), (double)g.y - (double)acc.imag())));
      mag = std::fmax(mag, std::abs(std::complex<double>((double)acc.real(), (double)acc.imag())));
    }
    const bool ok = err <= tol * mag;
    printf("LEN=%5d NTAB=%5d E=%2d stages=%d dir=%+d %s rel err %.3e\n", LEN, NTAB, E, fft_nstages(LEN, E), dir,
           sizeof(T) == 8 ? "f64" : "f32", err / mag);
    if (!ok) { printf("  FAIL\n"); ++bad; }
  }
  return bad;
}

// The stage tables of every (LEN, E) of an n-point grid tile [n, twiddle_len(n))
// of the twiddle buffer without overlap.
int check_layout(int n) {
  std::vector<int> owner(twiddle_len(n), -1);
  for (int h = 0; h < 2; ++h) {
    const int len = h == 0 ? n : n / 2;
    for (int i = 0; i < stab_e_count; ++i) {
      const int e = stab_e_at(i);
      if (!stab_e_ok(len, e)) continue;
      const int off = stab_offset(n, len, e);
      if (off < n || off + stab_len(len, e) > twiddle_len(n)) { printf("n=%d len=%d e=%d: tables outside the buffer\n", n, len, e); return 1; }
      for (int q = 0; q < stab_len(len, e); ++q) {
        if (owner[off + q] != -1) { printf("n=%d len=%d e=%d: tables overlap\n", n, len, e); return 1; }
        owner[off + q] = len * 64 + e;
      }
    }
  }
  for (int q = n; q < twiddle_len(n); ++q)
    if (owner[q] == -1) { printf("n=%d: gap at %d\n", n, q); return 1; }
  return 0;
}

// The LDS slot map of a line buffer of LEN elements is injective into the
// buffer's padded_len slots, and a tile's row stride leaves every row its
// padded_len slots, for 16- and 8-byte elements and b = 1..16 rows.
template <class T, int LEN>
int check_slots() {
  std::vector<char> used(padded_len<T>(LEN), 0);
  for (int e = 0; e < LEN; ++e) {
    const int s = pad<T, LEN>(e);
    if (s < 0 || s >= padded_len<T>(LEN) || used[s]) { printf("LEN=%d %s: slot map not injective at %d\n", LEN, sizeof(T) == 8 ? "f64" : "f32", e); return 1; }
    used[s] = 1;
  }
  for (int b = 1; b <= 16; b *= 2)
    if (tile_stride<T>(LEN, b) < padded_len<T>(LEN)) { printf("LEN=%d b=%d: tile stride too short\n", LEN, b); return 1; }
  return 0;
}
template <int LEN>
int check_slots_both() { return check_slots<double, LEN>() + check_slots<float, LEN>(); }

// The line-layout map of a grid of N (blocks of KB lines) and the element
// order of row tiles of B rows: every tile element (row bb, line kx) is
// visited once, the tiles of a grid cover the N*L elements once, and the B*KB
// elements of one block and tile sit at consecutive addresses in lane order.
template <int N, int B, int KB>
int check_line_map() {
  constexpr int L = N / 2;
  std::vector<char> seen(static_cast<size_t>(N) * L, 0);
  for (int j0 = 0; j0 < N; j0 += B) {
    for (int f = 0; f < B * L; ++f) {
      int bb, kx;
      tile_elem<N, B, KB>(f, bb, kx);
      const size_t o = lidx<N, KB>(kx, j0 + bb);
      if (bb < 0 || bb >= B || kx < 0 || kx >= L || o >= seen.size() || seen[o]) {
        printf("N=%d B=%d KB=%d: line map not a bijection at j0=%d f=%d\n", N, B, KB, j0, f);
        return 1;
      }
      seen[o] = 1;
      if (f % (B * KB) != 0) {
        int pb, pk;
        tile_elem<N, B, KB>(f - 1, pb, pk);
        if (o != lidx<N, KB>(pk, j0 + pb) + 1) {
          printf("N=%d B=%d KB=%d: tile block not contiguous at f=%d\n", N, B, KB, f);
          return 1;
        }
      }
    }
  }
  if (KB == 1 && lidx<N, KB>(3, 5) != static_cast<size_t>(3) * N + 5) { printf("KB=1 is not kx N + j\n"); return 1; }
  return 0;
}

int main() {
  int bad = 0;
  bad += check_line_map<64, 1, 1>() + check_line_map<64, 4, 1>() + check_line_map<64, 1, 2>() +
         check_line_map<64, 2, 2>() + check_line_map<64, 4, 4>() + check_line_map<256, 8, 2>() +
         check_line_map<4096, 1, 2>() + check_line_map<4096, 2, 2>() + check_line_map<4096, 2, 1>();
  for (int n : {16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 48, 96, 192, 384, 768}) bad += check_layout(n);
  bad += check_slots_both<8>() + check_slots_both<16>() + check_slots_both<24>() + check_slots_both<32>() +
         check_slots_both<48>() + check_slots_both<64>() + check_slots_both<96>() + check_slots_both<128>() +
         check_slots_both<192>() + check_slots_both<256>() + check_slots_both<384>() + check_slots_both<512>() +
         check_slots_both<768>() + check_slots_both<1024>() + check_slots_both<2048>() + check_slots_both<4096>();
  bad += check<8, double>(1e-14);
  bad += check<16, double>(1e-14);
  bad += check<32, double>(1e-14);
  bad += check<64, double>(1e-14);
  bad += check<128, double>(1e-14);
  bad += check<256, double>(1e-14);
  bad += check<512, double>(1e-14);
  bad += check<1024, double>(1e-14);
  bad += check<2048, double>(1e-14);
  bad += check<4096, double>(1e-14);
  bad += check<24, double>(1e-14);
  bad += check<48, double>(1e-14);
  bad += check<96, double>(1e-14);
  bad += check<192, double>(1e-14);
  bad += check<384, double>(1e-14);
  bad += check<384, float>(2e-6);
  bad += check<768, double>(1e-14);
  bad += check<1024, float>(2e-6);
  bad += check<4096, float>(2e-6);
  // row lengths L = N/2 with the N-point table
  bad += check<8, double, 16>(1e-14);
  bad += check<512, double, 1024>(1e-14);
  bad += check<2048, double, 4096>(1e-14);
  bad += check<24, double, 48>(1e-14);
  bad += check<48, double, 96>(1e-14);
  bad += check<96, double, 192>(1e-14);
  bad += check<192, double, 384>(1e-14);
  bad += check<384, double, 768>(1e-14);
  bad += check<192, float, 384>(2e-6);
  // other register tiles (elements per thread) the pass schedules may use
  bad += check<512, double, 1024, 4>(1e-14);
  bad += check<1024, double, 1024, 4>(1e-14);
  bad += check<1024, double, 1024, 2>(1e-14);
  bad += check<512, double, 1024, 2>(1e-14);
  bad += check<1024, float, 1024, 2>(2e-6);
  bad += check<1024, double, 1024, 16>(1e-14);
  bad += check<2048, double, 4096, 16>(1e-14);
  bad += check<256, double, 512, 4>(1e-14);
  bad += check<128, double, 128, 8>(1e-14);
  bad += check<256, float, 256, 8>(2e-6);
  bad += check<192, double, 384, 12>(1e-14);
  bad += check<384, double, 384, 12>(1e-14);
  bad += check<384, double, 384, 6>(1e-14);
  bad += check<96, double, 192, 6>(1e-14);
  bad += check<384, double, 768, 24>(1e-14);
  bad += check<768, double, 768, 24>(1e-14);
  bad += check<768, float, 768, 24>(2e-6);
  printf(bad ? "FAILED %d\n" : "ALL OK\n", bad);
  return bad ? 1 : 0;
}
