// bnb_batch.hpp -- split+bound kernel of a batch module: many independent
// problems (slots) of several objective families and dimensions d <= DMAX
// in one launch.  Included (as text, by gpugo_amd.emit) after ia.hpp,
// bnb_args.hpp and the variable-dimension code lists of each family with the
// dispatchers fam_value / fam_kkt / fam_grad.
//
// Block b works on slot b / vblocks: the slot's family and dimension are
// uniform within a block, so the family switch does not diverge.  Per slot
// the work is the same as the single-problem kernel (bnb_kernel.hpp) and the
// reference kernels sampling + ruleOutGPU (GPUGO_1.1/1-Ackley/ackley50.py:
// 602-766), on the slot's own queue, parent records and incumbent.
#pragma once

#ifndef GPUGO_MODULE_ABI_V
#define GPUGO_MODULE_ABI_V 1
#endif

namespace gpugo {

template <int R>
__device__ __forceinline__ double batch_sample_ub(int f, const iv<double> *x, const double *__restrict__ P, int d,
                                                  int S) {
  double best = __builtin_huge_val();
  for (int t = 0; t < S; ++t) {
    iv<double> xs[GPUGO_DMAX];
#pragma unroll
    for (int j = 0; j < GPUGO_DMAX; ++j) {
      if (j < d) {
        double w = x[j].hi - x[j].lo;
        double h = (S == 1) ? w * 0.5 : w / (double)(S + 1);
        double p = __builtin_fma(h, (double)(t + 1), x[j].lo);
        xs[j] = mk<double>(p, p);
      } else {
        xs[j] = mk<double>(0.0, 0.0);
      }
    }
    iv<double> fv = fam_value<R>(f, xs, P, d);
    if (fv.hi < best) best = fv.hi;
  }
  return best;
}

// decode with a runtime dimension d <= DMAX (see decode_child)
__device__ __forceinline__ void batch_decode(const BatchBoundArgs &a, int slot, const SlotInfo &I, long long p,
                                             unsigned pos, iv<double> *x) {
  const long long rec = ((long long)slot * a.pcap + p);
  const double *__restrict__ plo = a.plo + rec * GPUGO_DMAX;
  const double *__restrict__ phi = a.phi + rec * GPUGO_DMAX;
  const double *__restrict__ pw = a.pw + rec * GPUGO_DMAX;
  const int win = a.pwin[rec];
  const int d = I.d;
  const unsigned wid = (unsigned)I.wid;
  const int bpd = 32 - __builtin_clz(wid - 1u);
  unsigned long long packed = 0;
  if ((wid & (wid - 1u)) == 0u) {
    packed = pos;
  } else {
    unsigned r = pos;
    for (int k = 0; k < I.c; ++k) {
      unsigned q = (unsigned)__builtin_fma((double)r, I.inv_wid, I.half_inv_wid);
      unsigned dg = r - q * wid;
      packed |= (unsigned long long)dg << (k * bpd);
      r = q;
    }
  }
  const unsigned long long mask = (1ull << bpd) - 1ull;
#pragma unroll
  for (int j = 0; j < GPUGO_DMAX; ++j) {
    if (j < d) {
      const double L = plo[j], U = phi[j];
      int k = j - win;
      if (k < 0) k += d;
      if (k < I.c) {
        const unsigned dg = (unsigned)((packed >> (k * bpd)) & mask);
        const double w = pw[j];
        const double lo = __builtin_fma(w, (double)dg, L);
        const double hi = (dg + 1u == wid) ? U : __builtin_fma(w, (double)(dg + 1u), L);
        x[j] = mk<double>(lo, hi);
      } else {
        x[j] = mk<double>(L, U);
      }
    } else {
      x[j] = mk<double>(0.0, 0.0);
    }
  }
}

template <int R>
__device__ __forceinline__ void batch_bound_impl(const BatchBoundArgs &a) {
  __shared__ unsigned s_wcnt[4];
  __shared__ unsigned long long s_base;
  __shared__ unsigned long long s_ubmin;
  __shared__ unsigned s_pv, s_pk;
  const int slot = blockIdx.x / a.vblocks, vb = blockIdx.x % a.vblocks;
  if (slot >= a.slots) return;
  const SlotInfo I = a.info[slot];
  DevState *st = a.st + slot;
  if (!I.active || st->status != 0 || st->fresh) return;
  const long long n_child = (long long)st->n_split * (long long)I.W;
  if (n_child == 0) return;
  const long long first = st->first_parent;
  const double ub0 = st->ub_snap;
  const int cur = (int)(st->compactions & 1);
  const long long qoff = ((long long)slot * 2 + cur) * a.qcap;
  double *__restrict__ q_lb = a.q_lb + qoff;
  unsigned *__restrict__ q_par = a.q_par + qoff;
  unsigned *__restrict__ q_pos = a.q_pos + qoff;
  const double *__restrict__ P = a.params + (long long)slot * a.pmax;
  const double *__restrict__ dlo = a.dom_lo + (long long)slot * GPUGO_DMAX;
  const double *__restrict__ dhi = a.dom_hi + (long long)slot * GPUGO_DMAX;
  const int d = I.d, f = I.family;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const double invW = 1.0 / (double)I.W, halfInvW = 0.5 / (double)I.W;
  if (threadIdx.x == 0) {
    s_ubmin = order_key(__builtin_huge_val());
    s_pv = 0;
    s_pk = 0;
  }
  __syncthreads();
  for (long long base = (long long)vb * 256; base < n_child; base += (long long)a.vblocks * 256) {
    const long long t = base + threadIdx.x;
    const bool active = t < n_child;
    long long pl = 0;
    unsigned pos = 0;
    double lb = __builtin_huge_val(), su = __builtin_huge_val();
    bool keep = false, pv = false, pk = false;
    if (active) {
      pl = (long long)__builtin_fma((double)t, invW, halfInvW);
      pos = (unsigned)(t - pl * (long long)I.W);
      iv<double> x[GPUGO_DMAX];
      batch_decode(a, slot, I, first + pl, pos, x);
      iv<double> fv = fam_value<R>(f, x, P, d);
      lb = fv.lo;
      if (!(lb == lb)) lb = -__builtin_huge_val();
      if (lb <= ub0) {
        keep = !fam_kkt<R>(f, x, P, dlo, dhi, d);
        pk = !keep;
        su = batch_sample_ub<R>(f, x, P, d, I.n_samples);
      } else {
        pv = true;
      }
    }
    const unsigned long long mask = __ballot(keep);
    const unsigned pre =
        __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
    if (lane == 0) s_wcnt[w] = (unsigned)__popcll(mask);
    double m = su;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      double t2 = __shfl_xor(m, o, 64);
      m = t2 < m ? t2 : m;
    }
    const unsigned long long pvm = __ballot(pv), pkm = __ballot(pk);
    if (lane == 0) {
      if (m < __builtin_huge_val()) atomicMin(&s_ubmin, order_key(m));
      if (pvm) atomicAdd(&s_pv, (unsigned)__popcll(pvm));
      if (pkm) atomicAdd(&s_pk, (unsigned)__popcll(pkm));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const unsigned tot = s_wcnt[0] + s_wcnt[1] + s_wcnt[2] + s_wcnt[3];
      unsigned long long b = 0;
      if (tot) {
        b = atomicAdd((unsigned long long *)&st->q_count, (unsigned long long)tot);
        if (b + tot > (unsigned long long)a.qcap) atomicOr(&st->overflow, 1);
      }
      s_base = b;
    }
    __syncthreads();
    if (keep) {
      unsigned off = pre;
      for (int q = 0; q < w; ++q) off += s_wcnt[q];
      const unsigned long long o = s_base + off;
      if (o < (unsigned long long)a.qcap) {
        q_lb[o] = lb;
        q_par[o] = (unsigned)(first + pl);
        q_pos[o] = pos;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (s_ubmin < *(volatile unsigned long long *)&st->ub_key) atomicMin(&st->ub_key, s_ubmin);
    if (s_pv) atomicAdd((unsigned long long *)&st->pruned_value, (unsigned long long)s_pv);
    if (s_pk) atomicAdd((unsigned long long *)&st->pruned_kkt, (unsigned long long)s_pk);
    if (vb == 0) {
      st->iterations += 1;
      st->bounded += n_child;
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
  out[3] = (GPUGO_WITH_EXACT ? 1 : 0) | (GPUGO_WITH_FAST ? 2 : 0);
  out[4] = GPUGO_PMAX;
  const int np[] = {GPUGO_FAM_NPARAMS};
  for (int i = 0; i < GPUGO_NFAM; ++i) out[5 + i] = np[i];
}
// declared coefficient signs, GPUGO_PMAX entries per family
extern "C" __global__ void gpugo_batch_param_signs(signed char *out) {
  const signed char s[] = {GPUGO_FAM_PSIGN};
  for (int i = threadIdx.x; i < GPUGO_NFAM * GPUGO_PMAX; i += blockDim.x) out[i] = s[i];
}
