// Candidate schedules of the four step passes (P1 row, P2 column, P3 row,
// P4 column) per grid size and precision, fastest first.
//
// Each entry is Sched<E, W, G, PF> (elements per thread, rows or lines per
// workgroup, tile-load group size with 0 = all at once, prefetch): the
// three fastest distinct (E, W) choices of an exhaustive sweep over valid
// schedules, each pass timed alone on MI355X over 40 back-to-back launches
// on a 0.3 N particle field (the times are in the comments).  The first entry
// is the default; the runtime tuner (csg_autotune) times the candidates inside
// the captured step loop and keeps the fastest combination.  Sizes without an
// entry use the single rule-based schedule Sched<>.
#pragma once
#include "csg_kernels.hpp"

namespace csg {

template <class... S>
struct SchedList {};

template <int N, bool F64, int PASS>
struct StepCands {
  using type = SchedList<Sched<>>;
};

template <> struct StepCands<128, true, 1> { using type = SchedList<Sched<4, 4, 0, 0>, Sched<4, 8, 1, 0>, Sched<4, 16, 1, 0>>; };  // 4.7 / 5.0 / 5.9 us
template <> struct StepCands<128, true, 2> { using type = SchedList<Sched<4, 2, 0, 1>, Sched<4, 4, 0, 1>, Sched<8, 4, 0, 1>>; };  // 4.0 / 4.4 / 5.3 us
template <> struct StepCands<128, true, 3> { using type = SchedList<Sched<4, 4, 0, 0>, Sched<4, 8, 1, 0>, Sched<4, 16, 1, 0>>; };  // 3.7 / 3.9 / 4.5 us
template <> struct StepCands<128, true, 4> { using type = SchedList<Sched<4, 2, 0, 1>, Sched<4, 4, 0, 0>, Sched<4, 8, 0, 0>>; };  // 4.1 / 4.3 / 5.0 us
template <> struct StepCands<128, false, 1> { using type = SchedList<Sched<4, 4, 0, 0>, Sched<4, 8, 0, 0>, Sched<4, 16, 0, 0>>; };  // 4.3 / 4.6 / 5.3 us
template <> struct StepCands<128, false, 2> { using type = SchedList<Sched<4, 2, 0, 1>, Sched<4, 4, 0, 1>, Sched<4, 8, 0, 1>>; };  // 3.6 / 4.0 / 4.5 us
template <> struct StepCands<128, false, 3> { using type = SchedList<Sched<4, 4, 0, 1>, Sched<4, 8, 0, 0>, Sched<4, 16, 1, 0>>; };  // 3.3 / 3.5 / 3.9 us
template <> struct StepCands<128, false, 4> { using type = SchedList<Sched<4, 2, 0, 0>, Sched<4, 4, 0, 0>, Sched<4, 8, 0, 0>>; };  // 3.7 / 3.9 / 4.3 us
template <> struct StepCands<192, true, 1> { using type = SchedList<Sched<6, 4, 0, 1>, Sched<6, 8, 1, 0>, Sched<6, 16, 1, 0>>; };  // 7.4 / 7.4 / 8.8 us
template <> struct StepCands<192, true, 2> { using type = SchedList<Sched<6, 2, 0, 1>, Sched<6, 4, 0, 1>, Sched<12, 4, 0, 1>>; };  // 7.3 / 7.8 / 9.1 us
template <> struct StepCands<192, true, 3> { using type = SchedList<Sched<6, 8, 1, 0>, Sched<6, 4, 1, 0>, Sched<6, 16, 1, 0>>; };  // 6.2 / 6.2 / 7.3 us
template <> struct StepCands<192, true, 4> { using type = SchedList<Sched<6, 2, 0, 0>, Sched<6, 4, 0, 0>, Sched<6, 8, 0, 0>>; };  // 6.7 / 6.9 / 8.7 us
template <> struct StepCands<192, false, 1> { using type = SchedList<Sched<6, 4, 1, 0>, Sched<6, 8, 1, 0>, Sched<6, 16, 1, 0>>; };  // 6.4 / 6.7 / 7.4 us
template <> struct StepCands<192, false, 2> { using type = SchedList<Sched<6, 2, 0, 1>, Sched<6, 4, 0, 1>, Sched<12, 4, 0, 1>>; };  // 5.9 / 6.4 / 7.5 us
template <> struct StepCands<192, false, 3> { using type = SchedList<Sched<6, 4, 0, 0>, Sched<6, 8, 0, 0>, Sched<6, 16, 1, 0>>; };  // 5.1 / 5.3 / 5.9 us
template <> struct StepCands<192, false, 4> { using type = SchedList<Sched<6, 2, 0, 0>, Sched<6, 4, 0, 0>, Sched<6, 8, 0, 0>>; };  // 5.6 / 6.1 / 7.0 us
template <> struct StepCands<256, true, 1> { using type = SchedList<Sched<4, 4, 1, 0>, Sched<4, 2, 0, 0>, Sched<4, 8, 0, 0>>; };  // 5.5 / 5.7 / 6.4 us
template <> struct StepCands<256, true, 2> { using type = SchedList<Sched<4, 2, 0, 1>, Sched<4, 1, 0, 1>, Sched<4, 4, 0, 1>>; };  // 4.6 / 4.6 / 5.8 us
template <> struct StepCands<256, true, 3> { using type = SchedList<Sched<4, 4, 1, 0>, Sched<4, 2, 0, 0>, Sched<4, 8, 1, 0>>; };  // 4.4 / 4.8 / 5.1 us
template <> struct StepCands<256, true, 4> { using type = SchedList<Sched<4, 2, 0, 1>, Sched<4, 1, 0, 0>, Sched<4, 4, 0, 1>>; };  // 4.3 / 4.4 / 5.2 us
template <> struct StepCands<256, false, 1> { using type = SchedList<Sched<4, 2, 0, 1>, Sched<4, 4, 1, 0>, Sched<4, 8, 1, 0>>; };  // 4.9 / 5.0 / 5.6 us
template <> struct StepCands<256, false, 2> { using type = SchedList<Sched<4, 1, 0, 1>, Sched<4, 2, 0, 1>, Sched<4, 4, 0, 1>>; };  // 3.8 / 4.2 / 4.9 us
template <> struct StepCands<256, false, 3> { using type = SchedList<Sched<4, 2, 0, 1>, Sched<4, 4, 1, 0>, Sched<4, 8, 1, 0>>; };  // 3.9 / 4.0 / 4.2 us
template <> struct StepCands<256, false, 4> { using type = SchedList<Sched<4, 1, 0, 0>, Sched<4, 2, 0, 1>, Sched<4, 4, 0, 0>>; };  // 3.7 / 3.9 / 4.5 us
template <> struct StepCands<384, true, 1> { using type = SchedList<Sched<6, 2, 0, 1>, Sched<6, 4, 1, 1>, Sched<6, 8, 1, 0>>; };  // 8.6 / 8.7 / 9.7 us
template <> struct StepCands<384, true, 2> { using type = SchedList<Sched<6, 1, 0, 1>, Sched<6, 2, 0, 1>, Sched<12, 2, 0, 1>>; };  // 8.6 / 9.2 / 10.7 us
template <> struct StepCands<384, true, 3> { using type = SchedList<Sched<6, 4, 1, 1>, Sched<6, 8, 1, 0>, Sched<6, 2, 0, 1>>; };  // 7.6 / 8.3 / 8.4 us
template <> struct StepCands<384, true, 4> { using type = SchedList<Sched<6, 1, 0, 0>, Sched<6, 2, 0, 0>, Sched<6, 4, 0, 0>>; };  // 7.6 / 8.1 / 9.8 us
template <> struct StepCands<384, false, 1> { using type = SchedList<Sched<6, 2, 0, 0>, Sched<6, 4, 1, 1>, Sched<6, 8, 1, 0>>; };  // 7.4 / 7.4 / 8.2 us
template <> struct StepCands<384, false, 2> { using type = SchedList<Sched<6, 1, 0, 1>, Sched<6, 2, 0, 1>, Sched<12, 2, 0, 1>>; };  // 6.6 / 7.2 / 8.6 us
template <> struct StepCands<384, false, 3> { using type = SchedList<Sched<6, 4, 0, 1>, Sched<6, 2, 0, 1>, Sched<6, 8, 1, 0>>; };  // 6.1 / 6.2 / 6.5 us
template <> struct StepCands<384, false, 4> { using type = SchedList<Sched<6, 1, 0, 1>, Sched<6, 2, 0, 0>, Sched<6, 4, 0, 0>>; };  // 6.1 / 6.6 / 7.7 us
template <> struct StepCands<512, true, 1> { using type = SchedList<Sched<4, 2, 0, 0>, Sched<4, 4, 1, 0>, Sched<4, 1, 0, 0>>; };  // 6.7 / 6.9 / 7.3 us
template <> struct StepCands<512, true, 2> { using type = SchedList<Sched<8, 2, 0, 1>, Sched<8, 1, 0, 1>, Sched<4, 1, 0, 1>>; };  // 7.2 / 7.2 / 7.2 us
template <> struct StepCands<512, true, 3> { using type = SchedList<Sched<4, 4, 1, 0>, Sched<8, 4, 0, 0>, Sched<8, 8, 1, 0>>; };  // 6.1 / 7.2 / 7.4 us
template <> struct StepCands<512, true, 4> { using type = SchedList<Sched<4, 1, 0, 0>, Sched<8, 2, 0, 0>, Sched<8, 1, 0, 0>>; };  // 6.5 / 6.7 / 6.7 us
template <> struct StepCands<512, false, 1> { using type = SchedList<Sched<4, 2, 0, 0>, Sched<4, 1, 0, 0>, Sched<4, 4, 1, 0>>; };  // 5.5 / 5.5 / 5.9 us
template <> struct StepCands<512, false, 2> { using type = SchedList<Sched<8, 1, 0, 1>, Sched<8, 2, 0, 1>, Sched<4, 1, 0, 1>>; };  // 5.4 / 5.7 / 5.9 us
template <> struct StepCands<512, false, 3> { using type = SchedList<Sched<4, 4, 1, 0>, Sched<4, 2, 1, 0>, Sched<4, 8, 1, 0>>; };  // 4.9 / 5.3 / 5.4 us
template <> struct StepCands<512, false, 4> { using type = SchedList<Sched<8, 1, 0, 1>, Sched<4, 1, 0, 0>, Sched<8, 2, 0, 0>>; };  // 5.2 / 5.3 / 5.5 us
template <> struct StepCands<768, true, 1> { using type = SchedList<Sched<6, 4, 1, 0>, Sched<6, 2, 1, 0>, Sched<6, 1, 1, 0>>; };  // 12.3 / 12.9 / 13.2 us
template <> struct StepCands<768, true, 2> { using type = SchedList<Sched<12, 1, 0, 1>, Sched<6, 1, 0, 1>, Sched<12, 4, 0, 1>>; };  // 14.5 / 17.3 / 24.0 us
template <> struct StepCands<768, true, 3> { using type = SchedList<Sched<6, 8, 1, 0>, Sched<6, 2, 1, 0>, Sched<12, 2, 1, 0>>; };  // 15.5 / 15.7 / 15.7 us
template <> struct StepCands<768, true, 4> { using type = SchedList<Sched<12, 1, 0, 0>, Sched<6, 1, 0, 0>, Sched<12, 4, 0, 1>>; };  // 12.5 / 14.8 / 20.8 us
template <> struct StepCands<768, false, 1> { using type = SchedList<Sched<6, 4, 1, 0>, Sched<6, 1, 0, 1>, Sched<6, 2, 1, 0>>; };  // 9.3 / 9.5 / 10.3 us
template <> struct StepCands<768, false, 2> { using type = SchedList<Sched<12, 1, 0, 1>, Sched<12, 2, 0, 1>, Sched<6, 1, 0, 1>>; };  // 10.3 / 10.4 / 12.2 us
template <> struct StepCands<768, false, 3> { using type = SchedList<Sched<6, 8, 1, 0>, Sched<12, 8, 1, 0>, Sched<6, 2, 1, 0>>; };  // 9.8 / 10.0 / 10.3 us
template <> struct StepCands<768, false, 4> { using type = SchedList<Sched<12, 2, 0, 0>, Sched<12, 1, 0, 0>, Sched<6, 1, 0, 0>>; };  // 9.4 / 9.6 / 10.3 us
template <> struct StepCands<1024, true, 1> { using type = SchedList<Sched<8, 2, 1, 1>, Sched<4, 4, 1, 0>, Sched<4, 2, 1, 0>>; };  // 13.8 / 13.9 / 14.6 us
template <> struct StepCands<1024, true, 2> { using type = SchedList<Sched<4, 1, 0, 1>, Sched<16, 2, 0, 1>, Sched<8, 1, 0, 1>>; };  // 15.2 / 15.3 / 15.6 us
template <> struct StepCands<1024, true, 3> { using type = SchedList<Sched<8, 8, 1, 0>, Sched<8, 2, 1, 0>, Sched<16, 4, 1, 0>>; };  // 15.5 / 18.1 / 19.4 us
template <> struct StepCands<1024, true, 4> { using type = SchedList<Sched<4, 1, 0, 0>, Sched<8, 1, 0, 0>, Sched<16, 2, 0, 0>>; };  // 12.7 / 13.0 / 13.2 us
template <> struct StepCands<1024, false, 1> { using type = SchedList<Sched<8, 4, 1, 1>, Sched<8, 2, 0, 0>, Sched<4, 2, 1, 0>>; };  // 8.8 / 9.4 / 9.8 us
template <> struct StepCands<1024, false, 2> { using type = SchedList<Sched<4, 1, 0, 1>, Sched<8, 1, 0, 1>, Sched<16, 1, 0, 1>>; };  // 9.1 / 9.6 / 10.7 us
template <> struct StepCands<1024, false, 3> { using type = SchedList<Sched<8, 8, 1, 0>, Sched<8, 4, 1, 0>, Sched<16, 8, 1, 0>>; };  // 8.9 / 9.2 / 10.8 us
template <> struct StepCands<1024, false, 4> { using type = SchedList<Sched<4, 1, 0, 0>, Sched<8, 1, 0, 0>, Sched<16, 2, 0, 0>>; };  // 8.2 / 9.0 / 10.4 us
template <> struct StepCands<2048, true, 1> { using type = SchedList<Sched<4, 2, 1, 0>, Sched<4, 1, 0, 0>, Sched<8, 2, 1, 0>>; };  // 57.4 / 57.8 / 59.0 us
template <> struct StepCands<2048, true, 2> { using type = SchedList<Sched<8, 1, 0, 1>, Sched<4, 1, 0, 1>, Sched<16, 1, 0, 1>>; };  // 48.0 / 51.2 / 55.8 us
template <> struct StepCands<2048, true, 3> { using type = SchedList<Sched<8, 2, 1, 1>, Sched<4, 2, 1, 0>, Sched<8, 4, 1, 0>>; };  // 62.3 / 69.2 / 69.8 us
template <> struct StepCands<2048, true, 4> { using type = SchedList<Sched<4, 1, 0, 0>, Sched<8, 1, 0, 0>, Sched<16, 1, 0, 0>>; };  // 42.6 / 43.0 / 49.7 us
template <> struct StepCands<2048, false, 1> { using type = SchedList<Sched<8, 2, 1, 0>, Sched<4, 1, 0, 0>, Sched<8, 4, 1, 0>>; };  // 28.8 / 31.1 / 31.2 us
template <> struct StepCands<2048, false, 2> { using type = SchedList<Sched<16, 1, 0, 1>, Sched<8, 1, 0, 1>, Sched<4, 1, 0, 1>>; };  // 24.1 / 24.6 / 31.3 us
template <> struct StepCands<2048, false, 3> { using type = SchedList<Sched<16, 4, 1, 0>, Sched<16, 8, 1, 0>, Sched<8, 2, 1, 0>>; };  // 30.4 / 32.0 / 33.2 us
template <> struct StepCands<2048, false, 4> { using type = SchedList<Sched<8, 1, 0, 0>, Sched<16, 1, 0, 0>, Sched<4, 1, 0, 0>>; };  // 22.2 / 24.2 / 28.8 us
template <> struct StepCands<4096, true, 1> { using type = SchedList<Sched<8, 1, 0, 1>, Sched<4, 1, 0, 1>, Sched<8, 2, 1, 1>>; };  // 291.7 / 323.4 / 339.6 us
template <> struct StepCands<4096, true, 2> { using type = SchedList<Sched<16, 1, 0, 1>, Sched<8, 1, 0, 1>, Sched<4, 1, 0, 1>>; };  // 213.0 / 219.1 / 294.8 us
template <> struct StepCands<4096, true, 3> { using type = SchedList<Sched<8, 2, 0, 1>, Sched<16, 2, 0, 1>, Sched<4, 2, 1, 1>>; };  // 425.5 / 444.2 / 471.5 us
template <> struct StepCands<4096, true, 4> { using type = SchedList<Sched<16, 1, 0, 1>, Sched<8, 1, 0, 0>, Sched<4, 1, 0, 0>>; };  // 187.3 / 199.0 / 254.5 us
template <> struct StepCands<4096, false, 1> { using type = SchedList<Sched<8, 2, 0, 0>, Sched<16, 2, 1, 0>, Sched<8, 1, 1, 1>>; };  // 126.5 / 146.1 / 146.1 us
template <> struct StepCands<4096, false, 2> { using type = SchedList<Sched<16, 1, 0, 1>, Sched<8, 1, 0, 1>, Sched<4, 1, 0, 1>>; };  // 92.9 / 134.8 / 162.8 us
template <> struct StepCands<4096, false, 3> { using type = SchedList<Sched<16, 4, 1, 0>, Sched<8, 2, 1, 1>, Sched<16, 2, 1, 1>>; };  // 167.6 / 169.9 / 172.4 us
template <> struct StepCands<4096, false, 4> { using type = SchedList<Sched<16, 1, 0, 0>, Sched<8, 1, 0, 0>, Sched<4, 1, 0, 0>>; };  // 82.7 / 109.7 / 139.3 us

}  // namespace csg
