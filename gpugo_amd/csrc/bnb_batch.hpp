// bnb_batch.hpp -- split+bound kernel of a batch module: many independent
// problems (slots) of several objective families and dimensions d <= DMAX
// in one launch.  Included (as text, by gpugo_amd.emit) after ia.hpp,
// bnb_args.hpp and the variable-dimension code lists of each family with the
// dispatchers fam_value / fam_kkt / fam_grad.  In modules built with
// GPUGO_BATCH_MV, slots whose flags have GPUGO_FLAG_MEAN_VALUE use the
// mean-value bound (it multiplies the hipRTC time several times).
//
// A wavefront works on one slot at a time, so the slot's family and
// dimension are uniform and the family switch does not diverge.  Per slot
// the work is the same as the single-problem kernel (bnb_kernel.hpp) and the
// reference kernels sampling + ruleOutGPU (GPUGO_1.1/1-Ackley/ackley50.py:
// 602-766), on the slot's own queue, parent records and incumbent.
#pragma once


namespace gpugo {

template <int R>
__device__ __forceinline__ double batch_sample_ub(int f, const ChildBox &x, const double *__restrict__ P, int d,
                                                  int S) {
  double best = __builtin_huge_val();
  for (int t = 0; t < S; ++t) {
    const DiagPoint<ChildBox> xs{x, t, S};
    const iv<double> fv = fam_value<R>(f, xs, P, d);
    if (fv.hi < best) best = fv.hi;
  }
  return best;
}

__device__ __forceinline__ ChildBox batch_child(const BatchBoundArgs &a, int slot, const SlotInfo &I, long long p,
                                                unsigned pos) {
  const long long rec = ((long long)slot * a.pcap + p);
  return child_box(a.plo + rec * GPUGO_DMAX, a.phi + rec * GPUGO_DMAX, a.pw + rec * GPUGO_DMAX, a.pwin[rec], I.d, I.c,
                   (unsigned)I.wid, I.inv_wid, I.half_inv_wid, pos);
}

__device__ __forceinline__ long long wave_uniform(long long v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((unsigned long long)v >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}

// Persistent split+bound over the work list of all slots.  Each wavefront
// takes wave items gw, gw + (waves in the grid), ...; item gw belongs to the
// slot s with wave_off[s] <= gw < wave_off[s+1] (binary search, starting
// from the previous item's slot since gw increases).  Survivors are appended
// to the slot's queue with one atomic per wavefront.
template <int R>
__device__ __forceinline__ void batch_bound_impl(const BatchBoundArgs &a) {
  const int lane = threadIdx.x & 63;
  const long long grid_waves = (long long)gridDim.x * (blockDim.x >> 6);
  const long long total = a.wave_off[a.slots];
  int slot = 0;
  for (long long gw = wave_uniform((long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)); gw < total;
       gw += grid_waves) {
    int lo = slot, hi = a.slots;  // wave_off[lo] <= gw < wave_off[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (a.wave_off[mid] <= gw) lo = mid;
      else hi = mid;
    }
    slot = lo;
    const SlotInfo &I = a.info[slot];
    DevState *st = a.st + slot;
    const long long n_child = (long long)st->n_split * (long long)I.W;
    const long long first = st->first_parent;
    const double ub0 = st->ub_snap;
    const int cur = (int)(st->compactions & 1);
    const long long qoff = ((long long)slot * 2 + cur) * a.qcap;
    const double *__restrict__ P = a.params + (long long)slot * a.pmax;
    const double *__restrict__ dlo = a.dom_lo + (long long)slot * GPUGO_DMAX;
    const double *__restrict__ dhi = a.dom_hi + (long long)slot * GPUGO_DMAX;
    const int d = I.d, f = I.family;
    const long long t = (gw - a.wave_off[slot]) * 64 + lane;
    long long pl = 0;
    unsigned pos = 0;
    double lb = __builtin_huge_val(), su = __builtin_huge_val();
    bool keep = false, pv = false, pk = false;
    if (t < n_child) {
      pl = (long long)__builtin_fma((double)t, 1.0 / (double)I.W, 0.5 / (double)I.W);
      pos = (unsigned)(t - pl * (long long)I.W);
      const ChildBox x = batch_child(a, slot, I, first + pl, pos);
      const iv<double> fv = fam_value<R>(f, x, P, d);
      lb = fv.lo;
      if (!(lb == lb)) lb = -__builtin_huge_val();
      if (lb <= ub0) {
#if GPUGO_BATCH_MV
        if (I.flags & 1) {
          // mean-value form, as bound_mv in the single-problem kernel
          const DiagPoint<ChildBox> xc{x, 0, 1};
          const iv<double> s0 = fam_value<R>(f, xc, P, d);
          su = I.n_samples == 1 ? s0.hi : batch_sample_ub<R>(f, x, P, d, I.n_samples);
          MVSink<R, ChildBox> g{x, dlo, dhi, d, s0, false};
          fam_grad<R>(f, x, P, g, d);
          if (g.acc.lo > lb) lb = g.acc.lo;
          if (lb <= ub0) {
            keep = !g.prune;
            pk = !keep;
          } else {
            pv = true;
          }
        } else
#endif
        {
          // as in the single-problem kernel: monotonicity test, then the
          // sample bound of the survivor of the value test
          keep = !fam_kkt<R>(f, x, P, dlo, dhi, d);
          pk = !keep;
          su = batch_sample_ub<R>(f, x, P, d, I.n_samples);
        }
      } else {
        pv = true;
      }
    }
    const unsigned long long mask = __ballot(keep);
    const unsigned pre =
        __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
    double m = su;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double t2 = __shfl_xor(m, o, 64);
      m = t2 < m ? t2 : m;
    }
    const unsigned long long pvm = __ballot(pv), pkm = __ballot(pk);
    unsigned long long base = 0;
    if (lane == 0) {
      const unsigned tot = (unsigned)__popcll(mask);
      if (tot) {
        base = atomicAdd((unsigned long long *)&st->q_count, (unsigned long long)tot);
        if (base + tot > (unsigned long long)a.qcap) atomicOr(&st->overflow, 1);
      }
      if (m < __builtin_huge_val()) {
        const unsigned long long k = order_key(m);
        if (k < *(volatile unsigned long long *)&st->ub_key) atomicMin(&st->ub_key, k);
      }
      if (pvm) atomicAdd((unsigned long long *)&st->pruned_value, (unsigned long long)__popcll(pvm));
      if (pkm) atomicAdd((unsigned long long *)&st->pruned_kkt, (unsigned long long)__popcll(pkm));
    }
    base = (unsigned long long)__shfl((long long)base, 0, 64);
    if (keep) {
      const unsigned long long o = base + pre;
      if (o < (unsigned long long)a.qcap) {
        a.q_lb[qoff + o] = lb;
        a.q_par[qoff + o] = (unsigned)(first + pl);
        a.q_pos[qoff + o] = pos;
      }
    }
  }
}

}  // namespace gpugo

#if GPUGO_WITH_EXACT
extern "C" __global__ __launch_bounds__(256) void gpugo_batch_bound_exact(gpugo::BatchBoundArgs a) {
  gpugo::batch_bound_impl<gpugo::ROUND_EXACT>(a);
}
#endif
#if GPUGO_WITH_FAST
extern "C" __global__ __launch_bounds__(256) void gpugo_batch_bound_fast(gpugo::BatchBoundArgs a) {
  gpugo::batch_bound_impl<gpugo::ROUND_FAST>(a);
}
#endif
// out: [DMAX, NFAM, ABI, rounding variants, PMAX, nparams of family 0..NFAM-1]
extern "C" __global__ void gpugo_batch_info(int *out) {
  out[0] = GPUGO_DMAX;
  out[1] = GPUGO_NFAM;
  out[2] = GPUGO_MODULE_ABI_V;
  out[3] = (GPUGO_WITH_EXACT ? 1 : 0) | (GPUGO_WITH_FAST ? 2 : 0) | (GPUGO_BATCH_MV ? 4 : 0);
  out[4] = GPUGO_PMAX;
  const int np[] = {GPUGO_FAM_NPARAMS};
  // the library reads 64 ints; it rejects modules with more than 59 families
  for (int i = 0; i < GPUGO_NFAM && i < 59; ++i) out[5 + i] = np[i];
}
// declared coefficient signs, GPUGO_PMAX entries per family
extern "C" __global__ void gpugo_batch_param_signs(signed char *out) {
  const signed char s[] = {GPUGO_FAM_PSIGN};
  for (int i = threadIdx.x; i < GPUGO_NFAM * GPUGO_PMAX; i += blockDim.x) out[i] = s[i];
}
