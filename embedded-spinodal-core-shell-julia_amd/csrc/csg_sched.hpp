// Candidate schedules of the four step passes (P1 row, P2 column, P3 row,
// P4 column) per grid size and precision, fastest first.
//
// Each entry is Sched<E, W, G, PF, X, T> (elements per thread, rows or lines
// per workgroup, tile-load group size with 0 = all at once, prefetch (0 where
// used, 1 with the first loads, 2 once they have arrived), FFT
// exchange layout, twiddle source with 1 = per-stage tables, 2 = the half
// table staged in LDS although it exceeds 16 KiB).  The (E, W, G,
// PF) choices are the three fastest distinct (E, W) of an exhaustive sweep
// over valid schedules, each pass timed alone on MI355X over 40 back-to-back
// launches on a 0.3 N particle field, plus, for the ensemble sizes 128 .. 512,
// the fastest of the same sweep run over 64 instances per launch when it is
// not among them.  Each choice then carries the fastest (X, T) of a second
// sweep over those, and the list is ordered by that sweep's times (in the
// comments).  For the ensemble sizes, the fastest (E, W, G, PF, X, T) of the
// second sweep run over 64 instances per launch is appended when it is not
// already listed.  Where the captured loop ranks them differently (marked),
// the loop's order wins.  The first entry is the default; the runtime tuner
// (csg_autotune, csg_batch_autotune) times the candidates inside the captured
// step loop and keeps the fastest combination.  Sizes without an entry use the
// single rule-based schedule Sched<>.  After the twiddle staging became
// branch-free, the 256 .. 4096 lists were swept again in the captured loop
// (one pass varied at a time, the others at their defaults, with prefetch
// mode 2 added) and reordered where that sweep and an alternating A/B of the
// whole step put another candidate first; those comments give in-loop times.
#pragma once
#include "csg_kernels.hpp"

namespace csg {

template <class... S>
struct SchedList {};

template <int N, bool F64, int PASS>
struct StepCands {
  using type = SchedList<Sched<>>;
};

template <> struct StepCands<128, true, 1> { using type = SchedList<Sched<4, 4, 0, 1, 2, 0>, Sched<4, 8, 1, 0, 2, 0>, Sched<4, 16, 1, 0, 0, 0>, Sched<8, 16, 1, 0, 0, 0>, Sched<8, 16, 1, 0, 2, 0>, Sched<8, 8, 1, 0, 0, 0>>; };  // 4.7 / 5.0 / 5.8 / 6.6; 64 instances per launch: 8.7 us; appended after the partial sums were pinned ahead of the forward FFT: picked by csg_batch_autotune on the 256-particle ensemble (step 510 -> 497 us mixed, 811 -> 790 us FP64 with all appended entries)
template <> struct StepCands<128, true, 2> { using type = SchedList<Sched<4, 2, 0, 1, 0, 0>, Sched<4, 4, 0, 1, 0, 0>, Sched<8, 4, 0, 1, 0, 0>, Sched<4, 4, 0, 1, 0, 1>>; };  // 4.0 / 4.4 / 5.3; 64 instances per launch: 9.8 us
template <> struct StepCands<128, true, 3> { using type = SchedList<Sched<4, 4, 0, 0, 2, 0>, Sched<4, 8, 1, 0, 2, 0>, Sched<4, 16, 1, 0, 2, 0>, Sched<8, 8, 1, 0, 0, 0>, Sched<8, 8, 1, 0, 2, 0>>; };  // 3.6 / 3.9 / 4.5 / 4.6; 64 instances per launch: 11.0 us
template <> struct StepCands<128, true, 4> { using type = SchedList<Sched<4, 2, 0, 0, 0, 0>, Sched<4, 4, 0, 0, 2, 0>, Sched<4, 8, 0, 0, 2, 1>, Sched<4, 2, 0, 0, 2, 0>>; };  // 3.9 / 4.2 / 4.8; 64 instances per launch: 7.5 us
template <> struct StepCands<128, false, 1> { using type = SchedList<Sched<4, 4, 0, 0, 0, 0>, Sched<4, 8, 0, 0, 0, 0>, Sched<4, 8, 1, 0, 0, 0>, Sched<4, 16, 1, 0, 2, 0>>; };  // 4.3 / 4.6 / 4.7 / 5.3 us
template <> struct StepCands<128, false, 2> { using type = SchedList<Sched<4, 2, 0, 1, 0, 0>, Sched<4, 4, 0, 1, 0, 0>, Sched<4, 8, 0, 1, 0, 0>, Sched<4, 4, 0, 1, 0, 1>>; };  // 3.7 / 3.9 / 4.5; 64 instances per launch: 6.8 us
template <> struct StepCands<128, false, 3> { using type = SchedList<Sched<4, 4, 0, 0, 2, 0>, Sched<4, 8, 1, 0, 2, 0>, Sched<4, 16, 1, 0, 2, 0>, Sched<8, 16, 1, 0, 0, 0>, Sched<8, 16, 1, 0, 2, 0>>; };  // 3.3 / 3.5 / 3.9 / 4.3; 64 instances per launch: 6.9 us
template <> struct StepCands<128, false, 4> { using type = SchedList<Sched<4, 2, 0, 0, 1, 0>, Sched<4, 4, 0, 0, 0, 0>, Sched<4, 8, 0, 0, 2, 1>, Sched<4, 2, 0, 0, 0, 0>>; };  // 3.7 / 3.8 / 4.3; 64 instances per launch: 6.5 us
template <> struct StepCands<192, true, 1> { using type = SchedList<Sched<6, 4, 1, 1, 0, 0>, Sched<6, 8, 1, 0, 2, 0>, Sched<6, 16, 1, 0, 2, 0>>; };  // 7.0 / 7.4 / 8.6 us
template <> struct StepCands<192, true, 2> { using type = SchedList<Sched<6, 2, 0, 1, 0, 0>, Sched<6, 4, 0, 1, 2, 0>, Sched<12, 4, 0, 1, 0, 0>>; };  // 7.0 / 7.7 / 8.9 us
template <> struct StepCands<192, true, 3> { using type = SchedList<Sched<6, 4, 0, 1, 0, 0>, Sched<6, 8, 0, 0, 2, 0>, Sched<6, 16, 1, 0, 2, 0>>; };  // 5.8 / 6.0 / 7.0 us
template <> struct StepCands<192, true, 4> { using type = SchedList<Sched<6, 2, 0, 0, 0, 0>, Sched<6, 4, 0, 0, 0, 0>, Sched<6, 8, 0, 0, 1, 1>>; };  // 6.4 / 6.8 / 8.4 us
template <> struct StepCands<192, false, 1> { using type = SchedList<Sched<6, 4, 1, 0, 0, 0>, Sched<6, 8, 1, 0, 0, 0>, Sched<6, 16, 1, 0, 2, 0>>; };  // 6.2 / 6.6 / 7.2 us
template <> struct StepCands<192, false, 2> { using type = SchedList<Sched<6, 2, 0, 1, 0, 0>, Sched<6, 4, 0, 1, 0, 0>, Sched<12, 4, 0, 1, 0, 0>>; };  // 5.9 / 6.4 / 7.6 us
template <> struct StepCands<192, false, 3> { using type = SchedList<Sched<6, 4, 1, 0, 0, 0>, Sched<6, 8, 0, 0, 0, 0>, Sched<6, 16, 1, 0, 2, 0>>; };  // 5.0 / 5.3 / 5.8 us
template <> struct StepCands<192, false, 4> { using type = SchedList<Sched<6, 2, 0, 0, 0, 0>, Sched<6, 4, 0, 0, 0, 0>, Sched<6, 8, 0, 0, 1, 0>>; };  // 5.7 / 6.0 / 6.9 us
template <> struct StepCands<256, true, 1> { using type = SchedList<Sched<4, 2, 0, 0, 0, 0>, Sched<4, 4, 0, 0, 0, 0>, Sched<4, 8, 0, 0, 2, 0>, Sched<4, 8, 0, 0, 2, 1>>; };  // 5.3 / 5.5 / 6.3; 64 instances per launch: 33.1 us
template <> struct StepCands<256, true, 2> { using type = SchedList<Sched<4, 1, 0, 1, 2, 0>, Sched<4, 2, 0, 1, 0, 0>, Sched<8, 2, 0, 1, 0, 0>, Sched<8, 4, 0, 1, 0, 0>>; };  // 4.3 / 4.6 / 5.7 / 6.0 us
template <> struct StepCands<256, true, 3> { using type = SchedList<Sched<4, 2, 1, 0, 0, 0>, Sched<4, 4, 1, 0, 2, 0>, Sched<4, 8, 1, 0, 2, 0>, Sched<8, 16, 0, 0, 2, 1>, Sched<4, 8, 1, 0, 0, 1>>; };  // 4.3 / 4.4 / 5.0 / 6.6; 64 instances per launch: 35.2 us
template <> struct StepCands<256, true, 4> { using type = SchedList<Sched<4, 1, 0, 0, 2, 0>, Sched<4, 2, 0, 1, 2, 1>, Sched<4, 4, 0, 0, 1, 1>, Sched<8, 8, 0, 0, 0, 0>>; };  // 4.0 / 4.3 / 4.9 / 6.8 us; the single-field loop ranks the first two the other way (5.58 vs 6.39 us, step 20.0 -> 19.3 us), but over 68 instances per launch the second takes 65.1 vs 47.5 us, so the first stays the default of both
template <> struct StepCands<256, false, 1> { using type = SchedList<Sched<4, 2, 0, 0, 2, 0>, Sched<4, 4, 1, 0, 0, 0>, Sched<4, 4, 0, 0, 2, 0>, Sched<4, 8, 1, 0, 0, 0>>; };  // 4.8 / 5.0 / 5.0 / 5.5 us
template <> struct StepCands<256, false, 2> { using type = SchedList<Sched<4, 1, 0, 1, 1, 0>, Sched<4, 2, 0, 1, 0, 0>, Sched<4, 4, 0, 1, 2, 0>, Sched<8, 4, 0, 1, 0, 0>, Sched<8, 4, 0, 1, 0, 1>>; };  // 3.8 / 4.2 / 4.8 / 5.4; 64 instances per launch: 17.3 us
template <> struct StepCands<256, false, 3> { using type = SchedList<Sched<4, 2, 1, 1, 0, 0>, Sched<4, 4, 0, 0, 0, 0>, Sched<4, 8, 1, 0, 2, 0>, Sched<8, 8, 1, 0, 2, 0>>; };  // 3.8 / 4.0 / 4.1 / 5.0 us
template <> struct StepCands<256, false, 4> { using type = SchedList<Sched<4, 2, 0, 0, 0, 0>, Sched<4, 1, 0, 0, 0, 0>, Sched<4, 4, 0, 0, 1, 0>, Sched<4, 2, 0, 0, 0, 1>>; };  // 3.8 / 3.6 / 4.3 alone; 64 instances per launch: 15.5 us; the loop ranks the first two the other way (marked): 5.66 vs 6.35 us, step 18.5 -> 18.0 us
template <> struct StepCands<384, true, 1> { using type = SchedList<Sched<6, 2, 1, 1, 0, 0>, Sched<6, 4, 1, 0, 2, 0>, Sched<6, 4, 0, 0, 0, 0>, Sched<6, 8, 1, 0, 2, 0>, Sched<6, 8, 1, 0, 2, 1>, Sched<12, 4, 1, 2, 0, 0>>; };  // 8.0 / 8.1 / 8.2 / 9.4; 64 instances per launch: 74.3 us; appended after the partial sums were pinned ahead of the forward FFT: picked by csg_batch_autotune on the 256-particle ensemble (step 510 -> 497 us mixed, 811 -> 790 us FP64 with all appended entries)
template <> struct StepCands<384, true, 2> { using type = SchedList<Sched<6, 1, 0, 1, 0, 0>, Sched<6, 2, 0, 1, 2, 0>, Sched<12, 2, 0, 1, 0, 0>, Sched<12, 8, 0, 1, 1, 0>, Sched<12, 8, 0, 1, 0, 0>>; };  // 8.1 / 8.9 / 10.3 / 15.2; 64 instances per launch: 80.1 us
template <> struct StepCands<384, true, 3> { using type = SchedList<Sched<6, 4, 1, 0, 2, 0>, Sched<6, 2, 1, 1, 2, 0>, Sched<6, 8, 1, 0, 2, 0>, Sched<12, 4, 1, 0, 2, 0>, Sched<12, 4, 1, 0, 0, 0>>; };  // 6.9 / 7.4 / 8.0 / 8.3; 64 instances per launch: 104.2 us
template <> struct StepCands<384, true, 4> { using type = SchedList<Sched<6, 1, 0, 1, 2, 0>, Sched<6, 2, 0, 0, 2, 0>, Sched<6, 4, 0, 0, 2, 0>, Sched<12, 4, 0, 0, 0, 0>>; };  // 7.1 / 7.7 / 9.6 / 10.4 us
template <> struct StepCands<384, false, 1> { using type = SchedList<Sched<6, 2, 0, 0, 2, 0>, Sched<6, 4, 1, 0, 0, 0>, Sched<6, 8, 1, 0, 2, 0>, Sched<6, 8, 1, 0, 2, 1>, Sched<12, 8, 1, 0, 2, 0>, Sched<12, 4, 1, 0, 0, 0>>; };  // 7.1 / 7.2 / 7.9; 64 instances per launch: 46.6 us; appended after the partial sums were pinned ahead of the forward FFT: picked by csg_batch_autotune on the 256-particle ensemble (step 510 -> 497 us mixed, 811 -> 790 us FP64 with all appended entries)
template <> struct StepCands<384, false, 2> { using type = SchedList<Sched<6, 1, 0, 1, 0, 0>, Sched<6, 2, 0, 1, 0, 0>, Sched<12, 2, 0, 1, 0, 0>, Sched<12, 8, 0, 1, 1, 0>>; };  // 6.6 / 7.2 / 8.7 / 11.4 us
template <> struct StepCands<384, false, 3> { using type = SchedList<Sched<6, 4, 1, 0, 0, 0>, Sched<6, 2, 1, 1, 0, 0>, Sched<6, 8, 1, 0, 2, 0>, Sched<12, 4, 0, 0, 0, 0>, Sched<12, 8, 0, 0, 0, 0>>; };  // 6.0 / 6.2 / 6.4 / 6.9 us; appended after the partial sums were pinned ahead of the forward FFT: the one csg_batch_autotune picked on the 256-particle ensemble (mixed step 493-498 -> 489-492 us, alternating A/B)
template <> struct StepCands<384, false, 4> { using type = SchedList<Sched<6, 1, 0, 1, 0, 0>, Sched<6, 2, 0, 0, 0, 0>, Sched<6, 4, 0, 0, 1, 0>, Sched<12, 2, 0, 0, 0, 0>>; };  // 6.1 / 6.6 / 7.6 / 8.6 us
template <> struct StepCands<512, true, 1> { using type = SchedList<Sched<4, 1, 0, 0, 2, 1>, Sched<4, 2, 1, 0, 2, 0>, Sched<4, 1, 1, 0, 2, 0>, Sched<4, 4, 0, 1, 2, 1>, Sched<4, 4, 0, 1, 0, 1>, Sched<8, 8, 1, 0, 0, 0>>; };  // in the loop after the branch-free twiddle staging: 27.3 / 28.5 us per step for the first two; earlier 6.2 / 6.4 / 6.8 us alone; 64 instances per launch: 128.6 us; appended after the partial sums were pinned ahead of the forward FFT: picked by csg_batch_autotune on the 256-particle ensemble (step 510 -> 497 us mixed, 811 -> 790 us FP64 with all appended entries)
template <> struct StepCands<512, true, 2> { using type = SchedList<Sched<4, 1, 0, 1, 2, 0>, Sched<8, 1, 0, 1, 2, 0>, Sched<8, 2, 0, 1, 0, 0>, Sched<8, 4, 0, 1, 2, 1>, Sched<8, 4, 0, 1, 1, 0>>; };  // 6.4 / 6.6 / 6.8 / 8.2; 64 instances per launch: 108.2 us
template <> struct StepCands<512, true, 3> { using type = SchedList<Sched<4, 4, 1, 0, 2, 1>, Sched<4, 2, 1, 0, 2, 1>, Sched<8, 4, 1, 0, 2, 0>, Sched<8, 8, 0, 0, 0, 0>>; };  // 5.9 / 6.6 / 6.6 / 7.4 us
template <> struct StepCands<512, true, 4> { using type = SchedList<Sched<4, 1, 0, 0, 2, 0>, Sched<8, 1, 0, 0, 0, 0>, Sched<8, 2, 0, 1, 2, 0>, Sched<8, 4, 0, 1, 1, 1>, Sched<8, 2, 0, 1, 0, 0>>; };  // 5.7 / 6.1 / 6.2 / 7.2; 64 instances per launch: 94.8 us
template <> struct StepCands<512, false, 1> { using type = SchedList<Sched<4, 2, 0, 0, 2, 0>, Sched<4, 1, 1, 0, 0, 0>, Sched<4, 4, 1, 0, 2, 0>, Sched<8, 2, 0, 1, 0, 0>, Sched<8, 2, 0, 1, 2, 0>, Sched<8, 8, 1, 2, 2, 0>, Sched<8, 4, 1, 2, 0, 1>>; };  // 5.5 / 5.5 / 5.9 / 6.7; 64 instances per launch: 66.2 us; appended after the partial sums were pinned ahead of the forward FFT: picked by csg_batch_autotune on the 256-particle ensemble (step 510 -> 497 us mixed, 811 -> 790 us FP64 with all appended entries)
template <> struct StepCands<512, false, 2> { using type = SchedList<Sched<4, 1, 0, 1, 2, 0>, Sched<8, 1, 0, 1, 0, 0>, Sched<8, 2, 0, 1, 1, 0>, Sched<8, 1, 0, 1, 1, 0>>; };  // 5.3 / 5.5 / 5.7; 64 instances per launch: 59.2 us
template <> struct StepCands<512, false, 3> { using type = SchedList<Sched<4, 4, 1, 0, 2, 0>, Sched<4, 2, 1, 0, 2, 0>, Sched<4, 8, 1, 0, 0, 0>, Sched<8, 16, 0, 1, 0, 1>, Sched<8, 16, 0, 1, 0, 0>, Sched<8, 8, 0, 0, 0, 0>>; };  // 4.9 / 5.2 / 5.4 / 7.2; 64 instances per launch: 77.3 us; appended after the partial sums were pinned ahead of the forward FFT: the one csg_batch_autotune picked on the 256-particle ensemble (mixed step 493-498 -> 489-492 us, alternating A/B)
template <> struct StepCands<512, false, 4> { using type = SchedList<Sched<4, 1, 0, 0, 2, 0>, Sched<8, 1, 0, 1, 0, 0>, Sched<8, 2, 0, 0, 0, 0>>; };  // 4.8 / 5.1 / 5.5 us
template <> struct StepCands<768, true, 1> { using type = SchedList<Sched<6, 4, 1, 0, 2, 1>, Sched<6, 1, 1, 0, 2, 0>, Sched<6, 2, 1, 0, 2, 0>>; };  // 11.3 / 11.8 / 12.1 us
template <> struct StepCands<768, true, 2> { using type = SchedList<Sched<12, 2, 0, 1, 0, 0>, Sched<12, 1, 0, 1, 0, 0>, Sched<6, 1, 0, 1, 2, 0>>; };  // 12.9 / 13.4 / 14.6 us
template <> struct StepCands<768, true, 3> { using type = SchedList<Sched<12, 8, 1, 0, 2, 0>, Sched<12, 2, 1, 1, 2, 0>, Sched<6, 2, 1, 0, 2, 0>>; };  // 13.2 / 14.0 / 14.4 us
template <> struct StepCands<768, true, 4> { using type = SchedList<Sched<12, 2, 0, 0, 0, 0>, Sched<12, 1, 0, 0, 0, 0>, Sched<6, 1, 0, 0, 2, 0>>; };  // 11.1 / 11.4 / 12.0 us
template <> struct StepCands<768, false, 1> { using type = SchedList<Sched<6, 4, 0, 0, 2, 1>, Sched<6, 1, 1, 1, 0, 0>, Sched<6, 2, 1, 0, 2, 0>>; };  // 9.2 / 9.4 / 9.9 us
template <> struct StepCands<768, false, 2> { using type = SchedList<Sched<12, 2, 0, 1, 0, 0>, Sched<12, 1, 0, 1, 0, 0>, Sched<6, 1, 0, 1, 2, 0>>; };  // 10.4 / 10.8 / 11.5 us
template <> struct StepCands<768, false, 3> { using type = SchedList<Sched<6, 8, 1, 0, 2, 1>, Sched<12, 8, 1, 0, 2, 0>, Sched<6, 2, 1, 0, 2, 1>>; };  // 9.5 / 9.6 / 9.8 us
template <> struct StepCands<768, false, 4> { using type = SchedList<Sched<12, 2, 0, 0, 0, 0>, Sched<12, 1, 0, 0, 0, 0>, Sched<6, 1, 0, 0, 2, 0>>; };  // 9.5 / 9.7 / 9.8 us
// 1024^2 FP64 P1: ranked by the captured loop's step time, where two rows
// of two waves each beat four one-wave rows (50.2 vs 50.8 us per step)
// although they lose alone (isolated 11.2 us for Sched<8, 4, ...>); with
// the LDS-only exchange barrier, issuing the whole tile at once (G = E)
// is faster again: in the loop 13.4 vs 13.9 us, step 50.2 vs 51.0 us
template <> struct StepCands<1024, true, 1> { using type = SchedList<Sched<8, 4, 1, 2, 2, 1>, Sched<8, 4, 1, 2, 0, 1>, Sched<8, 4, 1, 0, 0, 1>, Sched<8, 4, 1, 2, 0, 0>, Sched<4, 2, 1, 0, 2, 1>>; };  // after the partial sums were pinned ahead of the forward FFT, in the loop (one pass varied, whole step): 45.75 / 45.81 / 46.12 / 46.23 / 46.81 us
template <> struct StepCands<1024, true, 2> { using type = SchedList<Sched<4, 1, 0, 1, 0, 1>, Sched<8, 1, 0, 1, 1, 1>, Sched<16, 2, 0, 1, 0, 0>>; };  // 13.1 / 14.1 / 14.8 us
template <> struct StepCands<1024, true, 3> { using type = SchedList<Sched<8, 8, 1, 0, 0, 1>, Sched<8, 4, 1, 0, 0, 1>, Sched<8, 2, 1, 0, 0, 0>, Sched<8, 4, 1, 0, 0, 1, 0>>; };  // 14.6 / 14.9 / 16.6 us; in the loop 14.0 / 15.5 / 19.9, and 17.1 for the last (both components, phi read once)
template <> struct StepCands<1024, true, 4> { using type = SchedList<Sched<4, 1, 0, 2, 0, 1>, Sched<4, 1, 0, 0, 0, 1>, Sched<8, 1, 0, 0, 1, 1>, Sched<16, 2, 0, 0, 0, 0>>; };  // in the loop 12.4 / 12.6 us (whole step 48.0 / 48.1); alone 10.4 / 11.1 / 12.4 us (the last three)
template <> struct StepCands<1024, false, 1> { using type = SchedList<Sched<8, 4, 1, 2, 0, 1>, Sched<8, 4, 1, 0, 0, 0>, Sched<4, 2, 0, 0, 2, 1>, Sched<8, 4, 1, 1, 0, 1>, Sched<8, 2, 0, 0, 0, 0>>; };  // in the loop: 37.1 / 37.4 / 37.6 / 37.5 / 38.3 us per step; with P4 candidate 0 against both old defaults, alternating: 36.1 vs 36.9 us
template <> struct StepCands<1024, false, 2> { using type = SchedList<Sched<4, 1, 0, 1, 2, 1>, Sched<8, 1, 0, 1, 2, 0>, Sched<16, 2, 0, 1, 0, 0>>; };  // 8.6 / 9.4 / 10.7 us
template <> struct StepCands<1024, false, 3> { using type = SchedList<Sched<8, 8, 0, 0, 2, 1>, Sched<8, 8, 1, 0, 2, 1>, Sched<8, 4, 1, 0, 2, 1>, Sched<16, 8, 1, 0, 0, 0>, Sched<4, 4, 1, 0, 2, 1, 0>>; };  // in the loop 11.3 / 11.5 / 11.9 / 13.9 us, and 14.3 for the last (both components, phi read once); the first two equal after the branch-free twiddle staging
template <> struct StepCands<1024, false, 4> { using type = SchedList<Sched<4, 1, 0, 2, 2, 1>, Sched<4, 1, 0, 0, 2, 1>, Sched<4, 1, 0, 0, 0, 1>, Sched<8, 1, 0, 0, 2, 0>, Sched<16, 2, 0, 0, 0, 0>>; };  // in the loop 9.6 / 9.9 / 10.1 / - / 13.8 us
template <> struct StepCands<2048, true, 1> { using type = SchedList<Sched<8, 2, 0, 0, 0, 0>, Sched<4, 1, 0, 2, 0, 0>, Sched<8, 1, 0, 0, 0, 0>, Sched<8, 1, 0, 2, 0, 0>, Sched<4, 1, 1, 0, 2, 1>, Sched<4, 2, 1, 0, 2, 1>, Sched<8, 2, 1, 0, 2, 0>>; };  // re-swept in the captured loop after the partial sums were pinned ahead of the forward FFT: 36.8 / 43.2 / 44.0 / 45.5 / 46.9 / 50.1 / 52.9 us (step 174.6 -> 163.6 us)
template <> struct StepCands<2048, true, 2> { using type = SchedList<Sched<8, 1, 0, 1, 0, 0>, Sched<16, 1, 0, 1, 0, 0>, Sched<4, 1, 0, 1, 0, 0>>; };  // 44.8 / 51.4 / 52.9 us
template <> struct StepCands<2048, true, 3> { using type = SchedList<Sched<8, 2, 1, 1, 0, 0>, Sched<16, 4, 1, 0, 0, 0>, Sched<4, 1, 1, 0, 0, 1>>; };  // 60.0 / 67.0 / 72.3 us; in the loop after the pinning 53.2 us for the first, and the 4096^2 winners (E = 16 or 8, two rows, G = 0 or 1, PF = 2) no faster (52.8 - 61.7 us)
template <> struct StepCands<2048, true, 4> { using type = SchedList<Sched<8, 1, 0, 0, 0, 0>, Sched<4, 1, 0, 0, 0, 0>, Sched<16, 1, 0, 0, 0, 0>>; };  // 38.4 / 44.6 / 45.7 us
template <> struct StepCands<2048, false, 1> { using type = SchedList<Sched<8, 4, 0, 0, 0, 0>, Sched<8, 2, 0, 0, 0, 0>, Sched<16, 2, 0, 0, 0, 0>, Sched<4, 1, 0, 1, 2, 1>, Sched<8, 2, 1, 0, 2, 1>, Sched<8, 1, 0, 0, 2, 0>>; };  // re-swept in the captured loop after the partial sums were pinned ahead of the forward FFT: 26.3 / 27.2 / 28.0 / 30.5 / 32.3 / 33.9 us (step 109.0 -> 103.5 us)
template <> struct StepCands<2048, false, 2> { using type = SchedList<Sched<16, 1, 0, 1, 0, 0>, Sched<8, 1, 0, 1, 0, 0>, Sched<4, 1, 0, 1, 0, 1>>; };  // 24.6 / 25.0 / 29.4 us
template <> struct StepCands<2048, false, 3> { using type = SchedList<Sched<16, 4, 1, 0, 0, 0>, Sched<16, 8, 1, 0, 0, 0>, Sched<8, 2, 1, 0, 0, 0>>; };  // 30.8 / 32.5 / 34.0 us; in the loop after the pinning 32.1 us for the first, the 4096^2 split variant Sched<16, 4, 1, 0, 0, 1, 2> 34.4 us
template <> struct StepCands<2048, false, 4> { using type = SchedList<Sched<8, 1, 0, 0, 0, 0>, Sched<16, 1, 0, 0, 0, 0>, Sched<4, 1, 0, 0, 1, 1>>; };  // 23.1 / 24.5 / 27.5 us
// 4096^2 FP64: the 32 KiB half table of W_4096 staged in LDS (T = 2) instead
// of read through the caches in every FFT stage, with the (G, PF, X) that
// then time best; candidates with and without it ranked by their in-loop
// event times (A/B on one box)
template <> struct StepCands<4096, true, 1> { using type = SchedList<Sched<8, 1, 0, 2, 0, 2>, Sched<8, 2, 0, 0, 0, 2>, Sched<8, 1, 0, 1, 0, 2>, Sched<4, 1, 0, 2, 0, 2>, Sched<8, 1, 0, 0, 0, 2>>; };  // after the partial sums were pinned ahead of the forward FFT (two workgroups per CU with one exchange buffer), in the loop: 172.4 / 176.7 / 186.8 / 190.5 / 192.7 us (the earlier default <8, 1, 0, 0, 2, 2>: 231.8)
template <> struct StepCands<4096, true, 2> { using type = SchedList<Sched<8, 1, 0, 1, 0, 2>, Sched<8, 1, 0, 1, 1, 0>, Sched<16, 1, 0, 1, 1, 0>, Sched<4, 1, 0, 1, 0, 0>>; };  // in the loop 181.2 / 183.7 / 195.5 / 284.9 us
template <> struct StepCands<4096, true, 3> { using type = SchedList<Sched<16, 2, 0, 2, 0, 2>, Sched<16, 2, 1, 2, 0, 2>, Sched<8, 2, 1, 2, 0, 2>, Sched<8, 2, 1, 1, 0, 2>, Sched<8, 2, 0, 2, 0, 2>, Sched<8, 2, 0, 1, 2, 0>>; };  // E = 16 tried after the partial sums were pinned ahead of the forward FFT: in the loop 281.0 / 282.4 us against 295.8 us for the earlier default (step 797.7 -> 786.1 us); then 305 / 316 / 302 / 349 us as measured before
template <> struct StepCands<4096, true, 4> { using type = SchedList<Sched<8, 1, 0, 2, 0, 2>, Sched<8, 1, 0, 0, 0, 2>, Sched<8, 1, 0, 1, 0, 2>, Sched<16, 1, 0, 1, 1, 0>, Sched<8, 1, 0, 0, 0, 0>>; };  // in the loop 159 / 166 / 170 / 188 / 185 us
template <> struct StepCands<4096, false, 1> { using type = SchedList<Sched<16, 2, 0, 0, 0, 0>, Sched<8, 2, 0, 0, 0, 0>, Sched<8, 4, 0, 0, 0, 0>, Sched<8, 2, 0, 0, 2, 1>, Sched<8, 2, 0, 0, 0, 1>>; };  // after the partial sums were pinned ahead of the forward FFT, in the loop: 105.5 / 112.1 / 118.4 / 127.3 / 130.5 us
template <> struct StepCands<4096, false, 2> { using type = SchedList<Sched<16, 1, 0, 1, 0, 1>, Sched<8, 1, 0, 1, 1, 1>, Sched<4, 1, 0, 1, 1, 1>>; };  // 96.1 / 125.8 / 148.8 us
template <> struct StepCands<4096, false, 3> { using type = SchedList<Sched<16, 4, 1, 0, 0, 1, 2>, Sched<16, 4, 1, 0, 0, 1>, Sched<16, 4, 1, 1, 0, 1>, Sched<8, 2, 1, 1, 0, 0>, Sched<16, 2, 1, 0, 0, 0>>; };  // the first: both components of the same rows on neighbouring workgroups, step 408.5 -> 405.9 us (alternating A/B); then in the loop 132.3 / 137.3 / 159.0 / 174.3 us
template <> struct StepCands<4096, false, 4> { using type = SchedList<Sched<8, 1, 0, 0, 0, 1>, Sched<16, 1, 0, 0, 0, 0>, Sched<4, 1, 0, 0, 1, 1>>; };  // 80.0 / 84.8 / 121.9 us

}  // namespace csg
