// Candidate schedules of the four step passes (P1 row, P2 column, P3 row,
// P4 column) per grid size and precision, fastest first.
//
// Each entry is Sched<E, W, G, PF> (elements per thread, rows or lines per
// workgroup, tile-load group size with 0 = all at once, prefetch): the
// three fastest distinct (E, W) choices of an exhaustive sweep over valid
// schedules, each pass timed alone on MI355X over 40 back-to-back launches
// on a 0.3 N particle field (the times are in the comments).  For the
// ensemble sizes 128 .. 512 a fourth entry is the fastest schedule of the same
// sweep run over 64 instances per launch, when it is not among the three.  The
// first entry is the default; the runtime tuner (csg_autotune,
// csg_batch_autotune) times the candidates inside the captured step loop and
// keeps the fastest combination.  Sizes without an entry use the single
// rule-based schedule Sched<>.
#pragma once
#include "csg_kernels.hpp"

namespace csg {

template <class... S>
struct SchedList {};

template <int N, bool F64, int PASS>
struct StepCands {
  using type = SchedList<Sched<>>;
};

template <> struct StepCands<128, true, 1> { using type = SchedList<Sched<4, 4, 0, 1>, Sched<4, 8, 1, 0>, Sched<4, 16, 1, 0>, Sched<8, 16, 1, 0>>; };  // 4.8 / 5.0 / 5.9; ensemble (64 instances per launch): 8.8 us
template <> struct StepCands<128, true, 2> { using type = SchedList<Sched<4, 2, 0, 1>, Sched<4, 4, 0, 1>, Sched<8, 4, 0, 1>>; };  // 4.0 / 4.4 / 5.4 us
template <> struct StepCands<128, true, 3> { using type = SchedList<Sched<4, 4, 0, 0>, Sched<4, 8, 1, 0>, Sched<4, 16, 1, 0>, Sched<8, 8, 1, 0>>; };  // 3.7 / 4.0 / 4.5; ensemble (64 instances per launch): 10.9 us
template <> struct StepCands<128, true, 4> { using type = SchedList<Sched<4, 2, 0, 0>, Sched<4, 4, 0, 0>, Sched<4, 8, 0, 0>>; };  // 4.0 / 4.3 / 4.9 us
template <> struct StepCands<128, false, 1> { using type = SchedList<Sched<4, 4, 0, 0>, Sched<4, 8, 0, 0>, Sched<4, 16, 1, 0>, Sched<4, 8, 1, 0>>; };  // 4.3 / 4.7 / 5.4; ensemble (64 instances per launch): 6.9 us
template <> struct StepCands<128, false, 2> { using type = SchedList<Sched<4, 2, 0, 1>, Sched<4, 4, 0, 1>, Sched<4, 8, 0, 1>>; };  // 3.7 / 4.1 / 4.5 us
template <> struct StepCands<128, false, 3> { using type = SchedList<Sched<4, 4, 0, 0>, Sched<4, 8, 1, 0>, Sched<4, 16, 1, 0>, Sched<8, 16, 1, 0>>; };  // 3.4 / 3.6 / 4.0; ensemble (64 instances per launch): 6.9 us
template <> struct StepCands<128, false, 4> { using type = SchedList<Sched<4, 2, 0, 0>, Sched<4, 4, 0, 0>, Sched<4, 8, 0, 0>>; };  // 3.8 / 3.9 / 4.3 us
template <> struct StepCands<192, true, 1> { using type = SchedList<Sched<6, 4, 1, 1>, Sched<6, 8, 1, 0>, Sched<6, 16, 1, 0>>; };  // 7.0 / 7.5 / 8.9 us
template <> struct StepCands<192, true, 2> { using type = SchedList<Sched<6, 2, 0, 1>, Sched<6, 4, 0, 1>, Sched<12, 4, 0, 1>>; };  // 7.0 / 7.9 / 8.9 us
template <> struct StepCands<192, true, 3> { using type = SchedList<Sched<6, 4, 0, 1>, Sched<6, 8, 0, 0>, Sched<6, 16, 1, 0>>; };  // 5.8 / 6.1 / 7.3 us
template <> struct StepCands<192, true, 4> { using type = SchedList<Sched<6, 2, 0, 0>, Sched<6, 4, 0, 0>, Sched<6, 8, 0, 0>>; };  // 6.5 / 6.9 / 8.8 us
template <> struct StepCands<192, false, 1> { using type = SchedList<Sched<6, 4, 1, 0>, Sched<6, 8, 1, 0>, Sched<6, 16, 1, 0>>; };  // 6.3 / 6.7 / 7.5 us
template <> struct StepCands<192, false, 2> { using type = SchedList<Sched<6, 2, 0, 1>, Sched<6, 4, 0, 1>, Sched<12, 4, 0, 1>>; };  // 5.9 / 6.4 / 7.6 us
template <> struct StepCands<192, false, 3> { using type = SchedList<Sched<6, 4, 1, 0>, Sched<6, 8, 0, 0>, Sched<6, 16, 1, 0>>; };  // 5.0 / 5.3 / 5.9 us
template <> struct StepCands<192, false, 4> { using type = SchedList<Sched<6, 2, 0, 0>, Sched<6, 4, 0, 0>, Sched<6, 8, 0, 0>>; };  // 5.7 / 6.1 / 7.0 us
template <> struct StepCands<256, true, 1> { using type = SchedList<Sched<4, 2, 0, 0>, Sched<4, 4, 0, 0>, Sched<4, 8, 0, 0>>; };  // 5.3 / 5.5 / 6.4 us
template <> struct StepCands<256, true, 2> { using type = SchedList<Sched<4, 1, 0, 1>, Sched<4, 2, 0, 1>, Sched<8, 2, 0, 1>, Sched<8, 4, 0, 1>>; };  // 4.3 / 4.6 / 5.7; ensemble (64 instances per launch): 25.3 us
template <> struct StepCands<256, true, 3> { using type = SchedList<Sched<4, 2, 1, 0>, Sched<4, 4, 1, 0>, Sched<4, 8, 1, 0>, Sched<8, 16, 0, 0>>; };  // 4.4 / 4.4 / 5.1; ensemble (64 instances per launch): 34.2 us
template <> struct StepCands<256, true, 4> { using type = SchedList<Sched<4, 1, 0, 0>, Sched<4, 2, 0, 1>, Sched<4, 4, 0, 0>, Sched<8, 8, 0, 0>>; };  // 4.0 / 4.4 / 5.1; ensemble (64 instances per launch): 22.8 us
template <> struct StepCands<256, false, 1> { using type = SchedList<Sched<4, 2, 0, 0>, Sched<4, 4, 1, 0>, Sched<4, 8, 1, 0>, Sched<4, 4, 0, 0>>; };  // 4.8 / 5.0 / 5.6; ensemble (64 instances per launch): 20.4 us
template <> struct StepCands<256, false, 2> { using type = SchedList<Sched<4, 1, 0, 1>, Sched<4, 2, 0, 1>, Sched<4, 4, 0, 1>, Sched<8, 4, 0, 1>>; };  // 3.8 / 4.2 / 4.9; ensemble (64 instances per launch): 17.4 us
template <> struct StepCands<256, false, 3> { using type = SchedList<Sched<4, 2, 1, 1>, Sched<4, 4, 0, 0>, Sched<4, 8, 1, 0>, Sched<8, 8, 1, 0>>; };  // 3.8 / 4.0 / 4.2; ensemble (64 instances per launch): 20.1 us
template <> struct StepCands<256, false, 4> { using type = SchedList<Sched<4, 1, 0, 0>, Sched<4, 2, 0, 0>, Sched<4, 4, 0, 0>>; };  // 3.7 / 4.0 / 4.5 us
template <> struct StepCands<384, true, 1> { using type = SchedList<Sched<6, 2, 1, 1>, Sched<6, 4, 1, 0>, Sched<6, 8, 1, 0>, Sched<6, 4, 0, 0>>; };  // 8.0 / 8.2 / 9.7; ensemble (64 instances per launch): 79.0 us
template <> struct StepCands<384, true, 2> { using type = SchedList<Sched<6, 1, 0, 1>, Sched<6, 2, 0, 1>, Sched<12, 2, 0, 1>, Sched<12, 8, 0, 1>>; };  // 8.1 / 8.9 / 10.3; ensemble (64 instances per launch): 76.6 us
template <> struct StepCands<384, true, 3> { using type = SchedList<Sched<6, 4, 1, 0>, Sched<6, 2, 1, 1>, Sched<6, 8, 1, 0>, Sched<12, 4, 1, 0>>; };  // 6.9 / 7.8 / 8.2; ensemble (64 instances per launch): 107.4 us
template <> struct StepCands<384, true, 4> { using type = SchedList<Sched<6, 1, 0, 1>, Sched<6, 2, 0, 0>, Sched<6, 4, 0, 0>, Sched<12, 4, 0, 0>>; };  // 7.2 / 7.7 / 9.8; ensemble (64 instances per launch): 62.8 us
template <> struct StepCands<384, false, 1> { using type = SchedList<Sched<6, 4, 1, 0>, Sched<6, 2, 0, 0>, Sched<6, 8, 1, 0>>; };  // 7.2 / 7.3 / 8.2 us
template <> struct StepCands<384, false, 2> { using type = SchedList<Sched<6, 1, 0, 1>, Sched<6, 2, 0, 1>, Sched<12, 2, 0, 1>, Sched<12, 8, 0, 1>>; };  // 6.7 / 7.1 / 8.7; ensemble (64 instances per launch): 44.9 us
template <> struct StepCands<384, false, 3> { using type = SchedList<Sched<6, 4, 1, 0>, Sched<6, 2, 1, 1>, Sched<6, 8, 1, 0>, Sched<12, 4, 0, 0>>; };  // 6.0 / 6.2 / 6.5; ensemble (64 instances per launch): 49.7 us
template <> struct StepCands<384, false, 4> { using type = SchedList<Sched<6, 1, 0, 1>, Sched<6, 2, 0, 0>, Sched<6, 4, 0, 0>, Sched<12, 2, 0, 0>>; };  // 6.1 / 6.6 / 7.8; ensemble (64 instances per launch): 41.3 us
template <> struct StepCands<512, true, 1> { using type = SchedList<Sched<4, 2, 1, 0>, Sched<4, 1, 1, 0>, Sched<4, 4, 1, 0>, Sched<4, 4, 0, 1>>; };  // 6.2 / 6.5 / 6.9; ensemble (64 instances per launch): 131.2 us
template <> struct StepCands<512, true, 2> { using type = SchedList<Sched<8, 1, 0, 1>, Sched<4, 1, 0, 1>, Sched<8, 2, 0, 1>, Sched<8, 4, 0, 1>>; };  // 6.7 / 6.8 / 6.9; ensemble (64 instances per launch): 108.5 us
template <> struct StepCands<512, true, 3> { using type = SchedList<Sched<4, 4, 1, 0>, Sched<8, 4, 1, 0>, Sched<4, 2, 1, 0>, Sched<8, 8, 0, 0>>; };  // 6.0 / 6.6 / 6.8; ensemble (64 instances per launch): 148.7 us
template <> struct StepCands<512, true, 4> { using type = SchedList<Sched<4, 1, 0, 0>, Sched<8, 1, 0, 0>, Sched<8, 2, 0, 1>, Sched<8, 4, 0, 1>>; };  // 6.0 / 6.1 / 6.3; ensemble (64 instances per launch): 92.3 us
template <> struct StepCands<512, false, 1> { using type = SchedList<Sched<4, 2, 0, 0>, Sched<4, 1, 1, 0>, Sched<4, 4, 1, 0>, Sched<8, 2, 0, 1>>; };  // 5.5 / 5.6 / 5.9; ensemble (64 instances per launch): 66.4 us
template <> struct StepCands<512, false, 2> { using type = SchedList<Sched<8, 1, 0, 1>, Sched<8, 2, 0, 1>, Sched<4, 1, 0, 1>>; };  // 5.5 / 5.8 / 5.9 us
template <> struct StepCands<512, false, 3> { using type = SchedList<Sched<4, 4, 1, 0>, Sched<4, 2, 1, 0>, Sched<4, 8, 1, 0>, Sched<8, 16, 0, 1>>; };  // 4.9 / 5.3 / 5.4; ensemble (64 instances per launch): 81.8 us
template <> struct StepCands<512, false, 4> { using type = SchedList<Sched<4, 1, 0, 0>, Sched<8, 1, 0, 1>, Sched<8, 2, 0, 0>>; };  // 5.3 / 5.4 / 5.6 us
template <> struct StepCands<768, true, 1> { using type = SchedList<Sched<6, 4, 1, 0>, Sched<6, 1, 1, 0>, Sched<6, 2, 1, 0>>; };  // 11.6 / 11.9 / 12.2 us
template <> struct StepCands<768, true, 2> { using type = SchedList<Sched<12, 2, 0, 1>, Sched<12, 1, 0, 1>, Sched<6, 1, 0, 1>>; };  // 13.1 / 13.4 / 16.4 us
template <> struct StepCands<768, true, 3> { using type = SchedList<Sched<12, 8, 1, 0>, Sched<12, 2, 1, 1>, Sched<6, 2, 1, 0>>; };  // 13.7 / 14.7 / 14.8 us
template <> struct StepCands<768, true, 4> { using type = SchedList<Sched<12, 2, 0, 0>, Sched<12, 1, 0, 0>, Sched<6, 1, 0, 0>>; };  // 11.2 / 11.6 / 13.6 us
template <> struct StepCands<768, false, 1> { using type = SchedList<Sched<6, 4, 0, 0>, Sched<6, 1, 1, 1>, Sched<6, 2, 1, 0>>; };  // 9.2 / 9.6 / 10.2 us
template <> struct StepCands<768, false, 2> { using type = SchedList<Sched<12, 2, 0, 1>, Sched<12, 1, 0, 1>, Sched<6, 1, 0, 1>>; };  // 10.5 / 10.5 / 12.4 us
template <> struct StepCands<768, false, 3> { using type = SchedList<Sched<6, 8, 1, 0>, Sched<12, 8, 1, 0>, Sched<6, 2, 1, 0>>; };  // 9.8 / 9.9 / 10.2 us
template <> struct StepCands<768, false, 4> { using type = SchedList<Sched<12, 2, 0, 0>, Sched<12, 1, 0, 0>, Sched<6, 1, 0, 0>>; };  // 9.5 / 9.7 / 10.6 us
template <> struct StepCands<1024, true, 1> { using type = SchedList<Sched<8, 4, 1, 0>, Sched<8, 2, 1, 0>, Sched<4, 2, 1, 0>>; };  // 11.6 / 12.8 / 13.6 us
template <> struct StepCands<1024, true, 2> { using type = SchedList<Sched<4, 1, 0, 1>, Sched<8, 1, 0, 1>, Sched<16, 2, 0, 1>>; };  // 13.8 / 14.4 / 14.8 us
template <> struct StepCands<1024, true, 3> { using type = SchedList<Sched<8, 8, 1, 0>, Sched<8, 4, 1, 0>, Sched<8, 2, 1, 0>>; };  // 15.5 / 15.5 / 16.6 us
template <> struct StepCands<1024, true, 4> { using type = SchedList<Sched<4, 1, 0, 0>, Sched<8, 1, 0, 0>, Sched<16, 2, 0, 0>>; };  // 11.4 / 11.8 / 12.6 us
template <> struct StepCands<1024, false, 1> { using type = SchedList<Sched<8, 4, 1, 0>, Sched<8, 2, 0, 0>, Sched<4, 2, 1, 0>>; };  // 8.7 / 9.4 / 9.8 us
template <> struct StepCands<1024, false, 2> { using type = SchedList<Sched<4, 1, 0, 1>, Sched<8, 1, 0, 1>, Sched<16, 2, 0, 1>>; };  // 9.1 / 9.6 / 10.8 us
template <> struct StepCands<1024, false, 3> { using type = SchedList<Sched<8, 8, 1, 0>, Sched<8, 4, 1, 0>, Sched<16, 8, 1, 0>>; };  // 8.8 / 9.1 / 10.6 us
template <> struct StepCands<1024, false, 4> { using type = SchedList<Sched<4, 1, 0, 0>, Sched<8, 1, 0, 0>, Sched<16, 2, 0, 0>>; };  // 8.2 / 9.1 / 10.5 us
template <> struct StepCands<2048, true, 1> { using type = SchedList<Sched<4, 1, 1, 0>, Sched<8, 2, 1, 0>, Sched<4, 2, 1, 0>>; };  // 53.3 / 56.9 / 57.4 us
template <> struct StepCands<2048, true, 2> { using type = SchedList<Sched<8, 1, 0, 1>, Sched<16, 1, 0, 1>, Sched<4, 1, 0, 1>>; };  // 44.5 / 51.5 / 51.6 us
template <> struct StepCands<2048, true, 3> { using type = SchedList<Sched<8, 2, 1, 1>, Sched<16, 4, 1, 0>, Sched<4, 1, 1, 0>>; };  // 60.9 / 66.6 / 69.1 us
template <> struct StepCands<2048, true, 4> { using type = SchedList<Sched<8, 1, 0, 0>, Sched<4, 1, 0, 0>, Sched<16, 1, 0, 0>>; };  // 38.1 / 43.9 / 45.6 us
template <> struct StepCands<2048, false, 1> { using type = SchedList<Sched<8, 2, 1, 0>, Sched<4, 1, 0, 1>, Sched<8, 1, 0, 0>>; };  // 28.6 / 30.5 / 30.8 us
template <> struct StepCands<2048, false, 2> { using type = SchedList<Sched<16, 1, 0, 1>, Sched<8, 1, 0, 1>, Sched<4, 1, 0, 1>>; };  // 24.7 / 24.9 / 32.4 us
template <> struct StepCands<2048, false, 3> { using type = SchedList<Sched<16, 4, 1, 0>, Sched<16, 8, 1, 0>, Sched<8, 2, 1, 0>>; };  // 30.9 / 32.6 / 33.9 us
template <> struct StepCands<2048, false, 4> { using type = SchedList<Sched<8, 1, 0, 0>, Sched<16, 1, 0, 0>, Sched<4, 1, 0, 0>>; };  // 22.9 / 24.3 / 29.3 us
template <> struct StepCands<4096, true, 1> { using type = SchedList<Sched<8, 1, 0, 1>, Sched<4, 1, 0, 1>, Sched<8, 2, 1, 1>>; };  // 292.1 / 322.5 / 338.2 us
template <> struct StepCands<4096, true, 2> { using type = SchedList<Sched<16, 1, 0, 1>, Sched<8, 1, 0, 1>, Sched<4, 1, 0, 1>>; };  // 211.3 / 220.1 / 295.7 us
template <> struct StepCands<4096, true, 3> { using type = SchedList<Sched<8, 2, 0, 1>, Sched<16, 2, 0, 1>, Sched<4, 2, 1, 1>>; };  // 425.5 / 444.7 / 471.8 us
template <> struct StepCands<4096, true, 4> { using type = SchedList<Sched<16, 1, 0, 1>, Sched<8, 1, 0, 0>, Sched<4, 1, 0, 0>>; };  // 190.3 / 197.9 / 253.2 us
template <> struct StepCands<4096, false, 1> { using type = SchedList<Sched<8, 2, 0, 0>, Sched<8, 1, 0, 1>, Sched<16, 2, 0, 0>>; };  // 125.8 / 140.5 / 145.3 us
template <> struct StepCands<4096, false, 2> { using type = SchedList<Sched<16, 1, 0, 1>, Sched<8, 1, 0, 1>, Sched<4, 1, 0, 1>>; };  // 96.3 / 134.6 / 161.1 us
template <> struct StepCands<4096, false, 3> { using type = SchedList<Sched<16, 4, 1, 1>, Sched<8, 2, 1, 1>, Sched<16, 2, 1, 0>>; };  // 170.0 / 172.0 / 185.1 us
template <> struct StepCands<4096, false, 4> { using type = SchedList<Sched<16, 1, 0, 0>, Sched<8, 1, 0, 0>, Sched<4, 1, 0, 0>>; };  // 84.9 / 112.4 / 139.9 us

}  // namespace csg
