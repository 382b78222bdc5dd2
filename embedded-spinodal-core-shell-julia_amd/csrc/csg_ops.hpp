// Per-(N, precision) kernel tables.  The instantiations live in separate
// translation units (csg_ops_*.hip) so they compile in parallel.
#pragma once
#include <type_traits>

#include "csg_adjoint.hpp"

namespace csg {

using LaunchFn = void (*)(const PassArgs&, const void* tw, hipStream_t);

// Kernel table for one (N, precision).
struct Ops {
  LaunchFn fwd;        // real rows (double) -> state S (line layout, double)
  LaunchFn inv_state;  // state S -> real rows (double)
  LaunchFn inv_bulk;   // bulk-precision line array -> real rows (double)
  LaunchFn p1, p2, p3, p4, mu, rate;
  int n_row1, n_col, n_row3;  // workgroups of P1, P2, P3 (partials count)
  // adjoint passes (FP64 contexts only)
  void (*adj1)(const AdjArgs&, const void* tw, hipStream_t) = nullptr;
  void (*adj3)(const AdjArgs&, const void* tw, hipStream_t) = nullptr;
  void (*adj4)(const AdjArgs&, const void* tw, hipStream_t) = nullptr;
  int n_adj_col = 0, n_adj_row = 0;
};

// Launches cover a.ninst instances (grid y) of the same size.
inline unsigned instances(const PassArgs& a) { return a.ninst > 1 ? static_cast<unsigned>(a.ninst) : 1u; }

template <class CT, class TIN, class TOUT, int N, int MODE>
void launch_row(const PassArgs& a, const void* tw, hipStream_t s) {
  using Cfg = RowCfg<CT, TIN, TOUT, N, MODE>;
  hipLaunchKernelGGL((row_pass<CT, TIN, TOUT, N, MODE>), dim3(Cfg::GRID, instances(a)), dim3(Cfg::NT), 0, s, a,
                     static_cast<const cplx<CT>*>(tw));
}

template <class CT, class TS, int N, int MODE>
void launch_col(const PassArgs& a, const void* tw, hipStream_t s) {
  using Cfg = ColCfg<CT, TS, N, MODE>;
  hipLaunchKernelGGL((col_pass<CT, TS, N, MODE>), dim3(Cfg::GRID, instances(a)), dim3(Cfg::NT), 0, s, a,
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

// T: bulk (compute) type; the state is always double.
template <int N, class T>
Ops make_ops() {
  Ops o;
  o.fwd = launch_row<double, double, double, N, ROW_FWD>;
  o.inv_state = launch_row<double, double, double, N, ROW_INV>;
  o.inv_bulk = launch_row<T, T, double, N, ROW_INV>;
  o.p1 = launch_row<T, double, T, N, ROW_P1>;
  o.p3 = launch_row<T, T, T, N, ROW_P3>;
  o.p2 = launch_col<T, double, N, COL_P2>;
  o.mu = launch_col<T, double, N, COL_MU>;
  o.p4 = launch_col<T, double, N, COL_P4>;
  o.rate = launch_col<T, double, N, COL_RATE>;
  o.n_row1 = RowCfg<T, double, T, N, ROW_P1>::GRID;
  o.n_row3 = RowCfg<T, T, T, N, ROW_P3>::GRID;
  o.n_col = ColCfg<T, double, N, COL_P2>::GRID;
  if constexpr (std::is_same<T, double>::value) {
    o.adj1 = launch_adj1<N>;
    o.adj3 = launch_adj3<N>;
    o.adj4 = launch_adj4<N>;
    o.n_adj_col = AdjColCfg<N>::GRID;
    o.n_adj_row = AdjRowCfg<N>::GRID;
  }
  return o;
}

// Kernel table for grid n in FP64 (T = double) or mixed (T = float); false
// when n is not compiled in.
bool ops_fp64_pow2(int n, Ops* o);
bool ops_fp64_pow3(int n, Ops* o);
bool ops_mixed_pow2(int n, Ops* o);
bool ops_mixed_pow3(int n, Ops* o);


}  // namespace csg
