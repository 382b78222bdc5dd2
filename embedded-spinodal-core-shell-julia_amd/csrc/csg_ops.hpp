// Per-(N, precision) kernel tables.  The instantiations live in separate
// translation units (csg_ops_*.hip) so they compile in parallel.
#pragma once
#include <type_traits>

#include "csg_adjoint.hpp"
#include "csg_sched.hpp"

namespace csg {

using LaunchFn = void (*)(const PassArgs&, const void* tw, hipStream_t);

constexpr int MAX_CANDS = 7;
constexpr int SCHED_FIELDS = 7;   // E, W, G, PF, X, T, S

// Compiled schedules of one step pass.
struct PassCands {
  LaunchFn fn[MAX_CANDS] = {};
  int grid[MAX_CANDS] = {};
  int sched[MAX_CANDS][SCHED_FIELDS] = {};   // resolved E, W, G, PF, X, T, S
  int n = 0;
};

// Kernel table for one (N, precision).
struct Ops {
  LaunchFn fwd;        // real rows (double) -> state S (line layout, double)
  LaunchFn inv_state;  // state S -> real rows (double)
  LaunchFn inv_bulk;   // bulk-precision line array -> real rows (double)
  LaunchFn p1, p2, p3, p4, mu, rate;   // p1..p4: the selected schedules
  int n_row1, n_col, n_row3;  // workgroups of the selected P1, P2, P3 (partials count)
  PassCands cand[4];          // P1..P4 candidates
  int sel[4] = {0, 0, 0, 0};
  int max_row1 = 0, max_col = 0, max_row3 = 0;   // partial buffers are sized for these
  // adjoint passes (FP64 contexts only)
  void (*adj1)(const AdjArgs&, const void* tw, hipStream_t) = nullptr;
  void (*adj3)(const AdjArgs&, const void* tw, hipStream_t) = nullptr;
  void (*adj4)(const AdjArgs&, const void* tw, hipStream_t) = nullptr;
  int n_adj_col = 0, n_adj_row = 0;
};

// Point p1..p4 and the partial counts at candidates sel[0..3]; false if an
// index is out of range (nothing changed then).
inline bool select_schedule(Ops& o, const int sel[4]) {
  for (int q = 0; q < 4; ++q)
    if (sel[q] < 0 || sel[q] >= o.cand[q].n) return false;
  for (int q = 0; q < 4; ++q) o.sel[q] = sel[q];
  o.p1 = o.cand[0].fn[sel[0]];
  o.p2 = o.cand[1].fn[sel[1]];
  o.p3 = o.cand[2].fn[sel[2]];
  o.p4 = o.cand[3].fn[sel[3]];
  o.n_row1 = o.cand[0].grid[sel[0]];
  o.n_col = o.cand[1].grid[sel[1]];
  o.n_row3 = o.cand[2].grid[sel[2]];
  return true;
}

// Launches cover a.ninst instances (grid y) of the same size.
inline unsigned instances(const PassArgs& a) { return a.ninst > 1 ? static_cast<unsigned>(a.ninst) : 1u; }

template <class CT, class TIN, class TOUT, int N, int MODE, class SC = Sched<>>
void launch_row(const PassArgs& a, const void* tw, hipStream_t s) {
  using Cfg = RowCfg<CT, TIN, TOUT, N, MODE, SC>;
  hipLaunchKernelGGL((row_pass<CT, TIN, TOUT, N, MODE, SC>), dim3(Cfg::GRID, instances(a)), dim3(Cfg::NT), 0, s, a,
                     static_cast<const cplx<CT>*>(tw));
}

template <class CT, class TS, int N, int MODE, class SC = Sched<>>
void launch_col(const PassArgs& a, const void* tw, hipStream_t s) {
  using Cfg = ColCfg<CT, TS, N, MODE, SC>;
  hipLaunchKernelGGL((col_pass<CT, TS, N, MODE, SC>), dim3(Cfg::LAUNCH, instances(a)), dim3(Cfg::NT), 0, s, a,
                     static_cast<const cplx<CT>*>(tw));
}

template <int N>
void launch_adj1(const AdjArgs& a, const void* tw, hipStream_t s) {
  hipLaunchKernelGGL(adj_col1<N>, dim3(AdjColCfg<N>::GRID), dim3(AdjColCfg<N>::NT), 0, s, a,
                     static_cast<const cplx<double>*>(tw));
}
template <int N>
void launch_adj3(const AdjArgs& a, const void* tw, hipStream_t s) {
  hipLaunchKernelGGL(adj_col3<N>, dim3(AdjColCfg<N>::GRID), dim3(AdjColCfg<N>::NT), 0, s, a,
                     static_cast<const cplx<double>*>(tw));
}
template <int N>
void launch_adj4(const AdjArgs& a, const void* tw, hipStream_t s) {
  hipLaunchKernelGGL(adj_row4<N>, dim3(AdjRowCfg<N>::GRID), dim3(AdjRowCfg<N>::NT), 0, s, a,
                     static_cast<const cplx<double>*>(tw));
}

template <int N, class T, int PASS, class SC>
void add_cand(PassCands& pc) {
  const int i = pc.n++;
  if constexpr (PASS == 1 || PASS == 3) {
    constexpr int MODE = PASS == 1 ? ROW_P1 : ROW_P3;
    using TIN = std::conditional_t<PASS == 1, double, T>;
    using Cfg = RowCfg<T, TIN, T, N, MODE, SC>;
    pc.fn[i] = launch_row<T, TIN, T, N, MODE, SC>;
    pc.grid[i] = Cfg::GRID;
    const int r[SCHED_FIELDS] = {Cfg::E, Cfg::B, Cfg::G, Cfg::PF, Cfg::XM, Cfg::ST ? 1 : (SC::T == 2 ? 2 : 0),
                                 PASS == 3 ? (Cfg::SPLIT ? SC::S : 0) : 0};
    for (int q = 0; q < SCHED_FIELDS; ++q) pc.sched[i][q] = r[q];
  } else {
    constexpr int MODE = PASS == 2 ? COL_P2 : COL_P4;
    using Cfg = ColCfg<T, double, N, MODE, SC>;
    pc.fn[i] = launch_col<T, double, N, MODE, SC>;
    pc.grid[i] = Cfg::GRID;
    const int r[SCHED_FIELDS] = {Cfg::E, Cfg::C, 0, Cfg::PF, Cfg::XM, Cfg::ST ? 1 : (SC::T == 2 ? 2 : 0), 0};
    for (int q = 0; q < SCHED_FIELDS; ++q) pc.sched[i][q] = r[q];
  }
}

template <int N, class T, int PASS, class... S>
void fill_cands(PassCands& pc, SchedList<S...>) {
  static_assert(sizeof...(S) >= 1 && sizeof...(S) <= MAX_CANDS, "candidate count");
  (add_cand<N, T, PASS, S>(pc), ...);
}

// T: bulk (compute) type; the state is always double.
template <int N, class T>
Ops make_ops() {
  constexpr bool F64 = std::is_same<T, double>::value;
  Ops o;
  o.fwd = launch_row<double, double, double, N, ROW_FWD>;
  o.inv_state = launch_row<double, double, double, N, ROW_INV>;
  o.inv_bulk = launch_row<T, T, double, N, ROW_INV>;
  o.mu = launch_col<T, double, N, COL_MU>;
  o.rate = launch_col<T, double, N, COL_RATE>;
  fill_cands<N, T, 1>(o.cand[0], typename StepCands<N, F64, 1>::type{});
  fill_cands<N, T, 2>(o.cand[1], typename StepCands<N, F64, 2>::type{});
  fill_cands<N, T, 3>(o.cand[2], typename StepCands<N, F64, 3>::type{});
  fill_cands<N, T, 4>(o.cand[3], typename StepCands<N, F64, 4>::type{});
  for (int i = 0; i < o.cand[0].n; ++i) o.max_row1 = o.cand[0].grid[i] > o.max_row1 ? o.cand[0].grid[i] : o.max_row1;
  for (int i = 0; i < o.cand[1].n; ++i) o.max_col = o.cand[1].grid[i] > o.max_col ? o.cand[1].grid[i] : o.max_col;
  for (int i = 0; i < o.cand[2].n; ++i) o.max_row3 = o.cand[2].grid[i] > o.max_row3 ? o.cand[2].grid[i] : o.max_row3;
  const int def[4] = {0, 0, 0, 0};
  select_schedule(o, def);
  if constexpr (F64) {
    o.adj1 = launch_adj1<N>;
    o.adj3 = launch_adj3<N>;
    o.adj4 = launch_adj4<N>;
    o.n_adj_col = AdjColCfg<N>::GRID;
    o.n_adj_row = AdjRowCfg<N>::GRID;
  }
  return o;
}

// Kernel table for grid n in FP64 (T = double) or mixed (T = float); false
// when n is not compiled in.  Split by size group so the units compile in
// parallel.
bool ops_fp64_small(int n, Ops* o);   // 16 .. 256
bool ops_fp64_mid(int n, Ops* o);     // 512, 1024
bool ops_fp64_big(int n, Ops* o);     // 2048, 4096
bool ops_fp64_pow3(int n, Ops* o);    // 48 .. 768
bool ops_mixed_small(int n, Ops* o);
bool ops_mixed_mid(int n, Ops* o);
bool ops_mixed_big(int n, Ops* o);
bool ops_mixed_pow3(int n, Ops* o);


}  // namespace csg
