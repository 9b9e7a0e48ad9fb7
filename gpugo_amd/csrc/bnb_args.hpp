// bnb_args.hpp -- layouts shared by libgpugo.so (hipcc) and the hipRTC
// objective modules.  Plain structs, no pointers to host memory.
#pragma once
#ifndef GPUGO_BNB_ARGS_HPP
#define GPUGO_BNB_ARGS_HPP

#ifndef GPUGO_H
// status values of include/gpugo.h, for code compiled without it (modules)
enum gpugo_status {
  GPUGO_ST_RUNNING = 0,
  GPUGO_ST_CERTIFIED = 1,
  GPUGO_ST_X_CONVERGED = 2,
  GPUGO_ST_MAX_ITERS = 3,
  GPUGO_ST_EXHAUSTED = 4,
  GPUGO_ST_OVERFLOW = 5
};
#endif

// A batch slot that would outgrow its box queue or parent arena in the next
// step stops in this state instead (nothing consumed, the queue compacted),
// so the host can move it into a slot of a larger tier and resume it there.
// Internal to batches: never reported as a result.
#define GPUGO_ST_GROW 6

// layout version of the structs below; modules report it and the library
// refuses a module built against another layout
#define GPUGO_MODULE_ABI_V 10

namespace gpugo {

// Per-problem state kept in device memory for the whole solve.  One
// branch-and-bound step is: select (k_select), split+bound (objective
// module), compact (k_compact).  The reference keeps the same state as host
// NumPy arrays and scalars (1-Ackley/ackley50.py:971-1006).
struct DevState {
  unsigned long long ub_key;     // incumbent upper bound, order-preserving key (atomicMin)
  double ub_snap;                // upper bound at the start of the step (pruning)
  double lb_global;              // smallest live lower bound at the last select
  double out_min_lb;             // smallest lower bound among output boxes
  double width;                  // ub - lb rounded up, set when the step stops
  long long q_count;             // queue entries incl. this step's appends
  long long compact_count;       // live entries after this step's compaction
  long long parents_used;        // parent records written
  long long first_parent;        // first parent record split in this step
  long long iterations;          // steps that split boxes
  long long bounded;             // children bounded
  long long pruned_value;
  long long pruned_kkt;
  long long n_out;               // output boxes (x-width rule)
  long long queue_peak;
  long long compactions;         // compaction passes run (selects the live buffer)
  unsigned long long lbmin_key;  // smallest / largest live lower bound after the
  unsigned long long lbmax_key;  // last compaction (order keys)
  int n_split;                   // boxes split in this step
  int status;                    // gpugo_status
  int overflow;                  // 1: queue append, 2: parent arena, 4: output arena
  int compact_final;             // 1: the stop step's compaction still has to run
  int fresh;                     // batch slots: 1 until the root step ran
  int cblocks;                   // batch slots: compaction workgroups finished this step
  // one box per step with stale entries left in place between compactions:
  long long dirty_lo;            // chunk minima from this queue index up are stale
  long long consumed;            // queue index the last select consumed, or -1
};

// Arguments of the module kernel gpugo_bound_{exact,fast}.
struct BoundArgs {
  const double *plo, *phi;       // parent arena, [max_parents][d]
  const double *pw;              // (hi - lo) / wid per parent coordinate
  const int *pwin;               // first split coordinate of each parent
  const double *params;          // objective coefficients
  const double *dom_lo, *dom_hi; // initial box (monotonicity boundary rule)
  double *q_lb;                  // queue, structure of arrays
  unsigned int *q_par, *q_pos;
  DevState *st;
  long long q_cap;
  int d, wid, c, n_samples;
  unsigned int W;                // wid^c children per parent
  int flags;                     // GPUGO_FLAG_* of the problem (+ FLAG_F32_PRUNE)
  double inv_wid, half_inv_wid;  // exact digit extraction: q = floor(r*inv + half)
  const void *params32;          // coefficients as outward FP32 intervals (iv<float>)
  const double *prest;           // windowed modules: rest sums per parent record, [p][2][nrest] intervals
  int nrest, pad1;
  // windowed modules: term tables of the records split in this step (record
  // first_parent + b at [b][tab_n] / [b][tab_n32]), written by the step's
  // prep kernel; tab_n == 0: no tables, terms evaluated per child
  void *ptab;                    // iv<double>
  void *ptab32;                  // iv<float> (FP32 pruning pass)
  int tab_n, tab_n32;
  // windowed modules at large d: each rest sum is split over rsplit
  // workgroups (chunks of the terms); their partials, [b][pair][rsplit],
  // are folded in chunk order by the step's prep fold kernel (rsplit 1: one
  // workgroup per sum, no fold)
  void *rpart;                   // iv<double>
  int rsplit, pad3;
};

// BoundArgs.flags bit of the FP32 pruning pass (problem precision F32)
#define GPUGO_FLAG_F32_PRUNE 2
#ifndef GPUGO_FLAG_SPLIT
#define GPUGO_FLAG_SPLIT 4
#endif

// Arguments of the module kernel gpugo_eval_{exact,fast} (parity hook).
struct EvalArgs {
  const double *params;
  const double *lo, *hi;         // [n][d]
  const double *dom_lo, *dom_hi;
  double *flo, *fhi, *glo, *ghi, *ub;
  int *kkt;
  long long n;
  int d, n_samples;
};

// ---- batches: many independent problems, one slot each -------------------
// Per-slot problem description (written by the host when a problem is
// admitted).  Pools are indexed by slot with fixed strides.
struct SlotInfo {
  int active;                    // 1 while the slot holds a problem
  int family;                    // objective family index in the batch module
  int d, c, wid, k_select, n_samples, flags;
  unsigned W;                    // wid^c
  int pad0;
  long long max_iters;
  double tol_f, inv_wid, half_inv_wid;
};

// Arguments of the batch module kernel gpugo_batch_bound_{exact,fast}.
// The step's children of all slots form one list of wavefront-sized work
// items: slot s owns waves [wave_off[s], wave_off[s+1]) (exclusive scan of
// ceil(n_children / 64), written by the plan kernel), so a wavefront always
// works on one slot and the family switch never diverges.
struct BatchBoundArgs {
  const SlotInfo *info;
  DevState *st;                  // [slots]
  const double *plo, *phi, *pw;  // [slots][pcap][dmax]
  const int *pwin;               // [slots][pcap]
  const double *params;          // [slots][pmax]
  const double *dom_lo, *dom_hi; // [slots][dmax]
  double *q_lb;                  // [slots][2][qcap]
  unsigned *q_par, *q_pos;
  const long long *wave_off;     // [slots + 1]
  long long pcap, qcap;
  int slots, pmax;
};

// Admission record of a batch slot, followed in memory by params[pmax],
// lo[dmax], hi[dmax] and pw[dmax] (the root parent record).
struct AdmitHead {
  int slot, pad;
  SlotInfo info;
  DevState st;
};

// Terms j in [0, nt) of a loop whose term j uses the coordinates j + k,
// k in [klo, khi], and that meet the split window {(win + t) mod d : t < c}
// (c <= d): at most two ranges [lo, hi) (the window may wrap past d),
// merged when they overlap.  Returns the number of ranges.  The windowed
// code lists use the same ranges for the rest sums of a parent record
// (terms outside them) and for its children (terms inside them).
__host__ __device__ inline int win_ranges(int win, int c, int d, int klo, int khi, int nt, int *lo, int *hi) {
  int n = 0;
  const int e0 = win + c < d ? win + c : d;  // coordinates [win, e0) and [0, e1)
  const int e1 = win + c - d;
  int l = win - khi, h = e0 - klo;
  l = l < 0 ? 0 : l;
  h = h > nt ? nt : h;
  if (l < h) {
    lo[n] = l;
    hi[n] = h;
    ++n;
  }
  if (e1 > 0) {
    int l2 = -khi, h2 = e1 - klo;
    l2 = l2 < 0 ? 0 : l2;
    h2 = h2 > nt ? nt : h2;
    if (l2 < h2) {
      if (n == 1 && l2 <= hi[0] && lo[0] <= h2) {
        lo[0] = l2 < lo[0] ? l2 : lo[0];
        hi[0] = h2 > hi[0] ? h2 : hi[0];
      } else {
        lo[n] = l2;
        hi[n] = h2;
        ++n;
      }
    }
  }
  return n;
}

__host__ __device__ inline bool in_ranges(int j, int n, const int *lo, const int *hi) {
  return (n > 0 && j >= lo[0] && j < hi[0]) || (n > 1 && j >= lo[1] && j < hi[1]);
}

__host__ __device__ inline unsigned long long order_key(double x) {
  unsigned long long b;
  __builtin_memcpy(&b, &x, 8);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}
__host__ __device__ inline double key_value(unsigned long long k) {
  unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffULL) : ~k;
  double x;
  __builtin_memcpy(&x, &b, 8);
  return x;
}

}  // namespace gpugo

#endif  // GPUGO_BNB_ARGS_HPP
