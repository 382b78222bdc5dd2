// Persistent ensemble kernel: every pass of every member of an ensemble, for
// several steps, in ONE launch (an alternative to one batched launch per
// size class and pass).
//
// Work items are the workgroups of the batched passes: for member q of size
// class c, pass p has items[c][p] items (P4's item 0 is the history-row
// reduction).  Per pass, the items of all members are laid out in one
// flattened list (class, member, item) and cut into G contiguous ranges of
// equal estimated cost, one per workgroup; a workgroup walks its ranges step
// by step, pass by pass.  There is no grid-wide barrier: each member has a
// completion counter, an item of (step s, pass p) of member q waits until
// q's counter has counted all of q's items before (s, p), and a workgroup adds
// to a member's counter once per range segment of that member.  A member that
// finishes its passes early runs ahead to its next step.
//
// Deadlock freedom: all G workgroups are resident at once (G = occupancy x
// CUs, checked on the host), and every wait is for items that precede the
// waiting item in every workgroup's order, so the earliest unfinished
// (step, pass) always progresses.  Every wait is also bounded in wall time:
// on expiry the workgroup raises the abort flag, every workgroup leaves at
// its next wait, and the host reports the stall.
//
// Memory ordering: an item's stores are released at agent scope (written
// back from this XCD's L2) before the counter add; a waiting workgroup
// acquires at agent scope after it observes the count.
#pragma once
#include "csg_kernels.hpp"

namespace csg {

constexpr int PERSIST_NT = 256;            // threads of every work item
constexpr int PERSIST_MAX_CLASSES = 4;     // size classes of one launch
constexpr int PERSIST_SIZES[PERSIST_MAX_CLASSES] = {128, 256, 384, 512};

// Schedules of the work items (all 256 threads): rows per workgroup and
// lines per workgroup chosen so every pass has PERSIST_NT threads; the other
// knobs follow the batched candidates where an (E, W) pair matches, and the
// column passes exchange through one LDS buffer set so that the kernel's LDS
// (the largest item's) leaves room for several workgroups per CU.
template <int N, bool F64, int PASS> struct PSched;
template <> struct PSched<128, true, 1> { using type = Sched<4, 16, 1, 0, 0, 0>; };
template <> struct PSched<128, true, 2> { using type = Sched<4, 8, 0, 1, 0, 0>; };
template <> struct PSched<128, true, 3> { using type = Sched<4, 16, 1, 0, 2, 0>; };
template <> struct PSched<128, true, 4> { using type = Sched<4, 8, 0, 0, 2, 1>; };
template <> struct PSched<128, false, 1> { using type = Sched<4, 16, 1, 0, 2, 0>; };
template <> struct PSched<128, false, 2> { using type = Sched<4, 8, 0, 1, 0, 0>; };
template <> struct PSched<128, false, 3> { using type = Sched<4, 16, 1, 0, 2, 0>; };
template <> struct PSched<128, false, 4> { using type = Sched<4, 8, 0, 0, 2, 1>; };
template <> struct PSched<256, true, 1> { using type = Sched<4, 8, 0, 0, 2, 0>; };
template <> struct PSched<256, true, 2> { using type = Sched<4, 4, 0, 1, 0, 0>; };
template <> struct PSched<256, true, 3> { using type = Sched<4, 8, 1, 0, 2, 0>; };
template <> struct PSched<256, true, 4> { using type = Sched<4, 4, 0, 0, 1, 1>; };
template <> struct PSched<256, false, 1> { using type = Sched<4, 8, 1, 0, 0, 0>; };
template <> struct PSched<256, false, 2> { using type = Sched<4, 4, 0, 1, 2, 0>; };
template <> struct PSched<256, false, 3> { using type = Sched<4, 8, 1, 0, 2, 0>; };
template <> struct PSched<256, false, 4> { using type = Sched<4, 4, 0, 0, 1, 0>; };
template <> struct PSched<384, true, 1> { using type = Sched<6, 8, 1, 0, 2, 0>; };
template <> struct PSched<384, true, 2> { using type = Sched<6, 4, 0, 1, 0, 0>; };
template <> struct PSched<384, true, 3> { using type = Sched<6, 8, 1, 0, 2, 0>; };
template <> struct PSched<384, true, 4> { using type = Sched<6, 4, 0, 0, 0, 0>; };
template <> struct PSched<384, false, 1> { using type = Sched<6, 8, 1, 0, 2, 0>; };
template <> struct PSched<384, false, 2> { using type = Sched<6, 4, 0, 1, 0, 0>; };
template <> struct PSched<384, false, 3> { using type = Sched<6, 8, 1, 0, 2, 0>; };
template <> struct PSched<384, false, 4> { using type = Sched<6, 4, 0, 0, 1, 0>; };
template <> struct PSched<512, true, 1> { using type = Sched<4, 4, 0, 1, 2, 1>; };
template <> struct PSched<512, true, 2> { using type = Sched<8, 4, 0, 1, 0, 1>; };
template <> struct PSched<512, true, 3> { using type = Sched<4, 4, 1, 0, 2, 1>; };
template <> struct PSched<512, true, 4> { using type = Sched<8, 4, 0, 1, 0, 1>; };
template <> struct PSched<512, false, 1> { using type = Sched<4, 4, 1, 0, 2, 0>; };
template <> struct PSched<512, false, 2> { using type = Sched<8, 4, 0, 1, 0, 0>; };
template <> struct PSched<512, false, 3> { using type = Sched<4, 4, 1, 0, 2, 0>; };
template <> struct PSched<512, false, 4> { using type = Sched<8, 4, 0, 0, 0, 0>; };

template <class T, int N, int PASS>
using PRowCfg = RowCfg<T, std::conditional_t<PASS == 1, double, T>, T, N, PASS == 1 ? ROW_P1 : ROW_P3,
                       typename PSched<N, std::is_same<T, double>::value, PASS>::type>;
template <class T, int N, int PASS>
using PColCfg = ColCfg<T, double, N, PASS == 2 ? COL_P2 : COL_P4,
                       typename PSched<N, std::is_same<T, double>::value, PASS>::type>;

// Work items per member of each pass (the batched launch widths).
template <class T, int N>
constexpr int persist_items(int pass) {
  return pass == 1 ? PRowCfg<T, N, 1>::GRID
       : pass == 2 ? PColCfg<T, N, 2>::LAUNCH
       : pass == 3 ? PRowCfg<T, N, 3>::GRID
                   : PColCfg<T, N, 4>::LAUNCH;
}

template <class T, int N>
constexpr int persist_lds_bytes() {
  constexpr int a = RowShared<PRowCfg<T, N, 1>, T>::BYTES, b = ColShared<PColCfg<T, N, 2>, T>::BYTES;
  constexpr int c = RowShared<PRowCfg<T, N, 3>, T>::BYTES, d = ColShared<PColCfg<T, N, 4>, T>::BYTES;
  constexpr int ab = a > b ? a : b, cd = c > d ? c : d;
  return ab > cd ? ab : cd;
}
template <class T>
constexpr int persist_smem_bytes() {
  constexpr int a = persist_lds_bytes<T, 128>(), b = persist_lds_bytes<T, 256>();
  constexpr int c = persist_lds_bytes<T, 384>(), d = persist_lds_bytes<T, 512>();
  constexpr int ab = a > b ? a : b, cd = c > d ? c : d;
  return ab > cd ? ab : cd;
}

struct PersistClass {
  int n;            // grid size
  int ninst;        // members
  int inst_base;    // counter index of member 0
  int items[4];     // work items per member per pass
  int per_step;     // their sum
  PassArgs args[4]; // pass arguments covering all members of the class (slab strides)
  const void* tw;   // bulk-precision twiddles of the class
};

struct PersistPlan {
  PersistClass cls[PERSIST_MAX_CLASSES];
  int ncls;
  long long off[4][PERSIST_MAX_CLASSES + 1];   // flattened index of class c's first item in pass p
};

struct PersistArgs {
  const PersistPlan* plan;   // device
  const long long* ranges;   // device, [4][G + 1]: workgroup w takes [ranges[p][w], ranges[p][w + 1])
  unsigned* ctr;             // device, one completed-item counter per member, zero at launch start
  int* flag;                 // device abort flag (non-zero: leave)
  int steps;                 // loop iterations of this launch
  int last_advances;         // 0: the last iteration writes its history row only
  long long spin_ticks;      // wait limit in s_memrealtime ticks (100 MHz)
};

// Wait until *c >= target (thread 0 polls, the workgroup follows); false if
// the abort flag is raised or the wait outlasts the limit (which raises it).
__device__ __forceinline__ bool persist_wait(const unsigned* c, unsigned target, int* flag, long long ticks,
                                             int* ok_s) {
  if (threadIdx.x == 0) {
    int ok = 1;
    if (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      const long long t0 = static_cast<long long>(__builtin_amdgcn_s_memrealtime());
      while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
          ok = 0;
          break;
        }
        if (static_cast<long long>(__builtin_amdgcn_s_memrealtime()) - t0 > ticks) {
          __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    *ok_s = ok;
  }
  __syncthreads();
  const bool ok = *ok_s != 0;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return ok;
}

// Publish this workgroup's stores and count n finished items of a member.
__device__ __forceinline__ void persist_signal(unsigned* c, unsigned n) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(c, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// The plan is read through a constant-address-space pointer: it does not
// change during a launch, so its (uniform) fields load into scalar registers
// even though the kernel stores to global memory.
typedef const __attribute__((address_space(4))) PersistClass ConstClass;
typedef const __attribute__((address_space(4))) PersistPlan ConstPlan;

template <class T, int N>
__device__ __forceinline__ void persist_item(ConstClass* k, int pass, int bx, int q, bool advance, char* smem) {
  constexpr bool F64 = std::is_same<T, double>::value;
  const cplx<T>* tw = static_cast<const cplx<T>*>(k->tw);
  const PassArgs* args = (const PassArgs*)(k->args);
  // the item's LDS regions, carved from the kernel's one buffer
  auto row = [&](auto cfg, auto fn) {
    using Sh = RowShared<decltype(cfg), T>;
    fn(reinterpret_cast<cplx<T>*>(smem), reinterpret_cast<cplx<T>*>(smem + Sh::OFF_TW),
       reinterpret_cast<cplx<T>*>(smem + Sh::OFF_ST), reinterpret_cast<double*>(smem + Sh::OFF_SCR));
  };
  auto col = [&](auto cfg, auto fn) {
    using Sh = ColShared<decltype(cfg), T>;
    fn(reinterpret_cast<cplx<T>*>(smem), reinterpret_cast<cplx<T>*>(smem + Sh::OFF_TW),
       reinterpret_cast<double*>(smem + Sh::OFF_SCR));
  };
  if (pass == 0) {
    row(PRowCfg<T, N, 1>{}, [&](cplx<T>* l, cplx<T>* t, cplx<T>* st, double* sc) {
      row_body<T, double, T, N, ROW_P1, typename PSched<N, F64, 1>::type>(args[0], bx, q, tw, l, t, st, sc);
    });
  } else if (pass == 1) {
    col(PColCfg<T, N, 2>{}, [&](cplx<T>* l, cplx<T>* t, double* sc) {
      col_body<T, double, N, COL_P2, typename PSched<N, F64, 2>::type>(args[1], bx, q, tw, l, t, sc);
    });
  } else if (pass == 2) {
    row(PRowCfg<T, N, 3>{}, [&](cplx<T>* l, cplx<T>* t, cplx<T>* st, double* sc) {
      row_body<T, T, T, N, ROW_P3, typename PSched<N, F64, 3>::type>(args[2], bx, q, tw, l, t, st, sc);
    });
  } else if (advance || bx == 0) {
    // the last iteration of a run writes its history row without a step
    col(PColCfg<T, N, 4>{}, [&](cplx<T>* l, cplx<T>* t, double* sc) {
      col_body<T, double, N, COL_P4, typename PSched<N, F64, 4>::type>(args[3], bx, q, tw, l, t, sc);
    });
  }
}

// One workgroup per CU: the kernel holds every pass body of every size class
// and spills at any register budget; with the whole register file per wave
// it spills half as much (FP64 287 vs 593 VGPRs).  On the 256-member
// ensemble the mixed form runs 21-24 % faster than with two workgroups per
// CU; FP64 is within run-to-run spread.
template <class T>
__global__ void __launch_bounds__(PERSIST_NT, 1) persist_kernel(PersistArgs pa) {
  static_assert(PRowCfg<T, 128, 1>::NT == PERSIST_NT && PColCfg<T, 128, 2>::NT == PERSIST_NT &&
                PRowCfg<T, 128, 3>::NT == PERSIST_NT && PColCfg<T, 128, 4>::NT == PERSIST_NT &&
                PRowCfg<T, 256, 1>::NT == PERSIST_NT && PColCfg<T, 256, 2>::NT == PERSIST_NT &&
                PRowCfg<T, 256, 3>::NT == PERSIST_NT && PColCfg<T, 256, 4>::NT == PERSIST_NT &&
                PRowCfg<T, 384, 1>::NT == PERSIST_NT && PColCfg<T, 384, 2>::NT == PERSIST_NT &&
                PRowCfg<T, 384, 3>::NT == PERSIST_NT && PColCfg<T, 384, 4>::NT == PERSIST_NT &&
                PRowCfg<T, 512, 1>::NT == PERSIST_NT && PColCfg<T, 512, 2>::NT == PERSIST_NT &&
                PRowCfg<T, 512, 3>::NT == PERSIST_NT && PColCfg<T, 512, 4>::NT == PERSIST_NT,
                "every work item runs on the kernel's 256 threads");
  __shared__ __attribute__((aligned(16))) char smem[persist_smem_bytes<T>()];
  __shared__ int ok_s;
  ConstPlan* plan = (ConstPlan*)pa.plan;
  const int w = static_cast<int>(blockIdx.x), G = static_cast<int>(gridDim.x);
  for (int s = 0; s < pa.steps; ++s) {
    const bool advance = pa.last_advances || s + 1 < pa.steps;
    for (int p = 0; p < 4; ++p) {
      const long long g0 = pa.ranges[p * (G + 1) + w], g1 = pa.ranges[p * (G + 1) + w + 1];
      int cur = -1;        // member whose items this segment counts
      unsigned done = 0;
      int c = 0;
      for (long long g = g0; g < g1; ++g) {
        while (g >= plan->off[p][c + 1]) ++c;
        ConstClass* k = &plan->cls[c];
        const long long r = g - plan->off[p][c];
        const int q = static_cast<int>(r / k->items[p]);
        const int bx = static_cast<int>(r - static_cast<long long>(q) * k->items[p]);
        const int inst = k->inst_base + q;
        if (inst != cur) {
          if (cur >= 0) persist_signal(pa.ctr + cur, done);
          int before = 0;
          for (int pp = 0; pp < p; ++pp) before += k->items[pp];
          const unsigned target = static_cast<unsigned>(s) * static_cast<unsigned>(k->per_step) +
                                  static_cast<unsigned>(before);
          if (!persist_wait(pa.ctr + inst, target, pa.flag, pa.spin_ticks, &ok_s)) return;
          cur = inst;
          done = 0;
        }
        switch (k->n) {
          case 128: persist_item<T, 128>(k, p, bx, q, advance, smem); break;
          case 256: persist_item<T, 256>(k, p, bx, q, advance, smem); break;
          case 384: persist_item<T, 384>(k, p, bx, q, advance, smem); break;
          default: persist_item<T, 512>(k, p, bx, q, advance, smem); break;
        }
        __syncthreads();   // the next item reuses the LDS
        ++done;
      }
      if (cur >= 0) persist_signal(pa.ctr + cur, done);
    }
  }
}

// Host side (csg_persist.hip).
bool persist_size_supported(int n);
// Work items per member of each pass for grid n (F64 or mixed).
void persist_items_of(int n, bool f64, int items[4]);
int persist_blocks_per_cu(bool f64, int* blocks);
hipError_t persist_launch(bool f64, const PersistArgs& a, int G, int nctr, hipStream_t s);

}  // namespace csg
