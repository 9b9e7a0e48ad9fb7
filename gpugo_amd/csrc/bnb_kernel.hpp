// bnb_kernel.hpp -- kernels of an objective module.  Included (as text, by
// gpugo_amd.emit) after ia.hpp, bnb_args.hpp and the generated code lists,
// which define GPUGO_D, GPUGO_NPARAMS and
//   obj_value<R>(x, P, d)          interval value          (<fn>Value_IA)
//   obj_grad<R>(x, P, g, d)        interval gradient       (<fn>Derivative_IA)
//   obj_kkt<R>(x, P, dlo, dhi, d)  monotonicity prune test (<fn>Derivative_IA)
// where x[j] is the j-th coordinate interval (an array or an accessor) and
// d the runtime dimension (= GPUGO_D; objectives built with per-coordinate
// guards branch on it, which keeps the code lists' live ranges short)
// Reference kernels replaced: sampling and ruleOutGPU
// (GPUGO_1.1/1-Ackley/ackley50.py:602-766).
#pragma once

#ifndef GPUGO_WITH_F32
#define GPUGO_WITH_F32 0
#endif
#ifndef GPUGO_WITH_DIGITS15
#define GPUGO_WITH_DIGITS15 0
#endif
#ifndef GPUGO_SPLIT_G
#define GPUGO_SPLIT_G 0
#endif

// dimensions above this decode child coordinates on use instead of into an
// array (the array would keep 4 d registers live through the code lists)
#ifndef GPUGO_LAZY_D
#define GPUGO_LAZY_D 16
#endif

namespace gpugo {

// Rigorous upper bound of f at n_samples points of the box diagonal
// (DiagPoint); S = 1 is the midpoint.
template <int R, class XS>
__device__ __forceinline__ double sample_ub(const XS &x, const double *__restrict__ P, int S, int d) {
  double best = __builtin_huge_val();
  for (int t = 0; t < S; ++t) {
    const DiagPoint<XS> xs{x, t, S};
    const iv<double> f = obj_value<R>(xs, P, d);
    if (f.hi < best) best = f.hi;
  }
  return best;
}

// FP32 pruning pass: the value code list in FP32 intervals; a child whose
// FP32 lower bound exceeds the incumbent is ruled out without any FP64
// work (rigorous: the FP32 enclosure contains f over the child).  Nothing
// can be ruled out before there is an incumbent.
template <int R, class XS>
__device__ __forceinline__ bool f32_prunes(const BoundArgs &a, const XS &x, double ub0) {
#if GPUGO_WITH_F32
  if (!(a.flags & GPUGO_FLAG_F32_PRUNE) || !(ub0 < __builtin_huge_val())) return false;
  const F32Box<XS> xf{x};
  const iv<float> f = obj_value32<R>(xf, (const iv<float> *)a.params32, a.d);
  return (double)f.lo > ub0;
#else
  return false;
#endif
}

// value test, then for survivors the monotonicity test and the sample bound
// (the gradient reuses the value code list's temporaries, which then die
// before the sample evaluation: fewer live registers)
template <int R, class XS>
__device__ __forceinline__ void bound_plain(const BoundArgs &a, const XS &x, double ub0, double &lb, double &su,
                                            bool &keep, bool &pv, bool &pk) {
  const iv<double> f = obj_value<R>(x, a.params, a.d);
  lb = f.lo;
  if (!(lb == lb)) lb = -__builtin_huge_val();
  if (lb <= ub0) {
    // a child pruned by value cannot improve the incumbent: its sample
    // bound is >= its lower bound > ub0; so only survivors are sampled
    keep = !obj_kkt<R>(x, a.params, a.dom_lo, a.dom_hi, a.d);
    pk = !keep;
    su = sample_ub<R>(x, a.params, a.n_samples, a.d);
  } else {
    pv = true;
  }
}

// value test, then for survivors the mean-value form
// f(c) + sum_j G_j (X_j - c_j) (c the midpoint, G the gradient enclosure the
// monotonicity test needs anyway), the monotonicity test and the sample bound.
template <int R, class XS>
__device__ __forceinline__ void bound_mv(const BoundArgs &a, const XS &x, double ub0, double &lb, double &su,
                                         bool &keep, bool &pv, bool &pk) {
  const iv<double> f = obj_value<R>(x, a.params, a.d);
  lb = f.lo;
  if (!(lb == lb)) lb = -__builtin_huge_val();
  if (lb <= ub0) {
    const DiagPoint<XS> xc{x, 0, 1};
    const iv<double> s0 = obj_value<R>(xc, a.params, a.d);
    su = a.n_samples == 1 ? s0.hi : sample_ub<R>(x, a.params, a.n_samples, a.d);
    MVSink<R, XS> g{x, a.dom_lo, a.dom_hi, a.d, s0, false};
    obj_grad<R>(x, a.params, g, a.d);
    if (g.acc.lo > lb) lb = g.acc.lo;
    if (lb <= ub0) {
      keep = !g.prune;
      pk = !keep;
    } else {
      pv = true;
    }
  } else {
    pv = true;
  }
}

// Child `pos` of parent record `p`: window coordinates (win + k) mod d for
// k < c take base-wid digit k of pos (least significant first); child
// bounds are lo + w*digit and lo + w*(digit+1) with one fma each, w the
// parent's (hi-lo)/wid, and the top child ends exactly at hi
// (1-Ackley/ackley50.py:565-593).
__device__ __forceinline__ void decode_child(const BoundArgs &a, long long p, unsigned pos, iv<double> *x) {
  const double *__restrict__ plo = a.plo + p * GPUGO_D;
  const double *__restrict__ phi = a.phi + p * GPUGO_D;
  const double *__restrict__ pw = a.pw + p * GPUGO_D;
  const int win = a.pwin[p];
  const unsigned wid = (unsigned)a.wid;
  // digits packed at bpd bits each (wid^c < 2^32 keeps c*bpd <= 64)
  const int bpd = 32 - __builtin_clz(wid - 1u);
  unsigned long long packed = 0;
  if ((wid & (wid - 1u)) == 0u) {
    packed = pos;
  } else {
    unsigned r = pos;
    for (int k = 0; k < a.c; ++k) {
      unsigned q = (unsigned)__builtin_fma((double)r, a.inv_wid, a.half_inv_wid);
      unsigned dg = r - q * wid;
      packed |= (unsigned long long)dg << (k * bpd);
      r = q;
    }
  }
  const unsigned long long mask = (1ull << bpd) - 1ull;
#pragma unroll
  for (int j = 0; j < GPUGO_D; ++j) {
    const double L = plo[j], U = phi[j];
    int k = j - win;
    if (k < 0) k += GPUGO_D;
    if (k < a.c) {
      const unsigned dg = (unsigned)((packed >> (k * bpd)) & mask);
      const double w = pw[j];
      const double lo = __builtin_fma(w, (double)dg, L);
      const double hi = (dg + 1u == wid) ? U : __builtin_fma(w, (double)(dg + 1u), L);
      x[j] = mk<double>(lo, hi);
    } else {
      x[j] = mk<double>(L, U);
    }
  }
}

template <int R, bool MV>
__device__ __forceinline__ void bound_impl(const BoundArgs &a) {
  __shared__ unsigned s_wcnt[4];
  __shared__ unsigned long long s_base;
  __shared__ unsigned long long s_ubmin;
  __shared__ unsigned s_pv, s_pk;
  DevState *st = a.st;
  if (st->status != 0) return;
  const long long n_child = (long long)st->n_split * (long long)a.W;
  if (n_child == 0) return;
  const long long first = st->first_parent;
  const double ub0 = st->ub_snap;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const double invW = 1.0 / (double)a.W, halfInvW = 0.5 / (double)a.W;
  if (threadIdx.x == 0) {
    s_ubmin = order_key(__builtin_huge_val());
    s_pv = 0;
    s_pk = 0;
  }
  __syncthreads();
  for (long long base = (long long)blockIdx.x * 256; base < n_child; base += (long long)gridDim.x * 256) {
    const long long t = base + threadIdx.x;
    const bool active = t < n_child;
    long long pl = 0;
    unsigned pos = 0;
    double lb = __builtin_huge_val(), su = __builtin_huge_val();
    bool keep = false, pv = false, pk = false;
    if (active) {
      pl = (long long)__builtin_fma((double)t, invW, halfInvW);
      pos = (unsigned)(t - pl * (long long)a.W);
      if (MV || GPUGO_D > GPUGO_LAZY_D) {
        // coordinates decoded where the code lists use them
        const long long p = first + pl;
        const ChildBox x = child_box(a.plo + p * GPUGO_D, a.phi + p * GPUGO_D, a.pw + p * GPUGO_D, a.pwin[p],
                                     GPUGO_D, a.c, (unsigned)a.wid, a.inv_wid, a.half_inv_wid, pos);
        if (f32_prunes<R>(a, x, ub0)) pv = true;
        else if (MV) bound_mv<R>(a, x, ub0, lb, su, keep, pv, pk);
        else bound_plain<R>(a, x, ub0, lb, su, keep, pv, pk);
      } else {
        iv<double> x[GPUGO_D];
        decode_child(a, first + pl, pos, x);
        if (f32_prunes<R>(a, (const iv<double> *)x, ub0)) pv = true;
        else bound_plain<R>(a, (const iv<double> *)x, ub0, lb, su, keep, pv, pk);
      }
    }
    // order-free append of survivors: one global atomic per block
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
        if (b + tot > (unsigned long long)a.q_cap) atomicOr(&st->overflow, 1);
      }
      s_base = b;
    }
    __syncthreads();
    if (keep) {
      unsigned off = pre;
      for (int q = 0; q < w; ++q) off += s_wcnt[q];
      const unsigned long long o = s_base + off;
      if (o < (unsigned long long)a.q_cap) {
        a.q_lb[o] = lb;
        a.q_par[o] = (unsigned)(first + pl);
        a.q_pos[o] = pos;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (s_ubmin < *(volatile unsigned long long *)&st->ub_key) atomicMin(&st->ub_key, s_ubmin);
    if (s_pv) atomicAdd((unsigned long long *)&st->pruned_value, (unsigned long long)s_pv);
    if (s_pk) atomicAdd((unsigned long long *)&st->pruned_kkt, (unsigned long long)s_pk);
    if (blockIdx.x == 0) {
      st->iterations += 1;
      st->bounded += n_child;
    }
  }
}

#if GPUGO_SPLIT_G > 1
// Term-parallel split+bound: GPUGO_SPLIT_G consecutive lanes share one
// child; the code lists' template loops run over j = lane, lane + G, ...
// and reduce over the group, so a step's children occupy G times more
// lanes and each lane runs about 1/G of the terms.  The same tests as
// bound_impl (value, then for survivors the monotonicity test and the
// sample bound, or the mean-value form); lane 0 of a group reports the box.
template <int R, int G>
__device__ __forceinline__ double sample_ub_g(const ChildBox &x, const double *__restrict__ P, int S, int d, int lg) {
  double best = __builtin_huge_val();
  for (int t = 0; t < S; ++t) {
    const DiagPoint<ChildBox> xs{x, t, S};
    const iv<double> f = obj_value_g<R, G>(xs, P, d, lg);
    if (f.hi < best) best = f.hi;
  }
  return best;
}

// blk / nblk: this workgroup and the number sharing the pass (any block
// size up to 1024 threads)
template <int R, bool MV>
__device__ __forceinline__ void bound_phase_g(const BoundArgs &a, int blk, int nblk) {
  constexpr int G = GPUGO_SPLIT_G;
  typedef IA<double, R> A;
  __shared__ unsigned s_wcnt[16];
  __shared__ unsigned long long s_base;
  __shared__ unsigned long long s_ubmin;
  __shared__ unsigned s_pv, s_pk;
  DevState *st = a.st;
  if (st->status != 0) return;
  const long long n_child = (long long)st->n_split * (long long)a.W;
  if (n_child == 0) return;
  const long long first = st->first_parent;
  const double ub0 = st->ub_snap;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, lg = threadIdx.x & (G - 1);
  const double invW = 1.0 / (double)a.W, halfInvW = 0.5 / (double)a.W;
  if (threadIdx.x == 0) {
    s_ubmin = order_key(__builtin_huge_val());
    s_pv = 0;
    s_pk = 0;
  }
  __syncthreads();
  const int nw = (int)(blockDim.x >> 6), bsz = (int)blockDim.x;
  for (long long base = (long long)blk * bsz; base < n_child * G; base += (long long)nblk * bsz) {
    const long long ci = (base + threadIdx.x) / G;
    const bool active = ci < n_child;
    long long pl = 0;
    unsigned pos = 0;
    double lb = __builtin_huge_val(), su = __builtin_huge_val();
    bool keep = false, pv = false, pk = false;
    if (active) {
      pl = (long long)__builtin_fma((double)ci, invW, halfInvW);
      pos = (unsigned)(ci - pl * (long long)a.W);
      const long long p = first + pl;
      const ChildBox x = child_box(a.plo + p * GPUGO_D, a.phi + p * GPUGO_D, a.pw + p * GPUGO_D, a.pwin[p],
                                   GPUGO_D, a.c, (unsigned)a.wid, a.inv_wid, a.half_inv_wid, pos);
      const iv<double> f = obj_value_g<R, G>(x, a.params, a.d, lg);
      lb = f.lo;
      if (!(lb == lb)) lb = -__builtin_huge_val();
      if (lb <= ub0) {
        if (MV) {
          const DiagPoint<ChildBox> xc{x, 0, 1};
          const iv<double> s0 = obj_value_g<R, G>(xc, a.params, a.d, lg);
          su = a.n_samples == 1 ? s0.hi : sample_ub_g<R, G>(x, a.params, a.n_samples, a.d, lg);
          MVSink<R, ChildBox> g{x, a.dom_lo, a.dom_hi, a.d, mk<double>(0.0, 0.0), false};
          obj_grad_g<R, G>(x, a.params, g, a.d, lg);
          const iv<double> s = A::add(s0, group_sum<R, G>(g.acc));
          const bool prune = group_any<G>(g.prune);
          if (s.lo > lb) lb = s.lo;
          if (lb <= ub0) {
            keep = !prune;
            pk = !keep;
          } else {
            pv = true;
          }
        } else {
          keep = !obj_kkt_g<R, G>(x, a.params, a.dom_lo, a.dom_hi, a.d, lg);
          pk = !keep;
          su = sample_ub_g<R, G>(x, a.params, a.n_samples, a.d, lg);
        }
      } else {
        pv = true;
      }
      if (lg != 0) keep = pv = pk = false;
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
      unsigned tot = 0;
      for (int q = 0; q < nw; ++q) tot += s_wcnt[q];
      unsigned long long b = 0;
      if (tot) {
        b = atomicAdd((unsigned long long *)&st->q_count, (unsigned long long)tot);
        if (b + tot > (unsigned long long)a.q_cap) atomicOr(&st->overflow, 1);
      }
      s_base = b;
    }
    __syncthreads();
    if (keep) {
      unsigned off = pre;
      for (int q = 0; q < w; ++q) off += s_wcnt[q];
      const unsigned long long o = s_base + off;
      if (o < (unsigned long long)a.q_cap) {
        a.q_lb[o] = lb;
        a.q_par[o] = (unsigned)(first + pl);
        a.q_pos[o] = pos;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (s_ubmin < *(volatile unsigned long long *)&st->ub_key) atomicMin(&st->ub_key, s_ubmin);
    if (s_pv) atomicAdd((unsigned long long *)&st->pruned_value, (unsigned long long)s_pv);
    if (s_pk) atomicAdd((unsigned long long *)&st->pruned_kkt, (unsigned long long)s_pk);
    if (blk == 0) {
      st->iterations += 1;
      st->bounded += n_child;
    }
  }
}
#endif

#if GPUGO_SPLIT_G > 1
// Persistent step loop: one cooperative launch of resident 1024-thread
// workgroups runs the steps back to back (select on workgroup 0, then the
// term-parallel split+bound and the compaction on all), with grid barriers
// instead of kernel boundaries: no launch gaps between the step's phases
// and no host polling.  Every exit condition is read after a barrier from
// state no workgroup writes until the next barrier, so all workgroups leave
// at the same step.
template <int R, bool MV>
__device__ __forceinline__ void persist_impl(const PersistArgs &p) {
  const int blk = blockIdx.x, nblk = p.nblocks;
  DevState *st = p.ba.st;
  for (long long k = 0; k < p.max_steps; ++k) {
    const long long step = p.step0 + k;
    const int cur = (int)(step & 1);
    if (step > 0 && blk == 0) {
      SelectArgs sa = p.sa;
      sa.q_lb = p.q_lb[cur];
      sa.q_par = p.q_par[cur];
      sa.q_pos = p.q_pos[cur];
      select_body(sa);
    }
    if (!grid_barrier(p.bar, nblk)) {
      if (threadIdx.x == 0) atomicOr(p.err, 1);
      return;
    }
    const int status = __hip_atomic_load(&st->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int cfinal = __hip_atomic_load(&st->compact_final, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (status != GPUGO_ST_RUNNING && !cfinal) return;
    if (status == GPUGO_ST_RUNNING) {
      BoundArgs ba = p.ba;
      ba.q_lb = p.q_lb[cur];
      ba.q_par = p.q_par[cur];
      ba.q_pos = p.q_pos[cur];
      bound_phase_g<R, MV>(ba, blk, nblk);
    }
    if (!grid_barrier(p.bar, nblk)) {
      if (threadIdx.x == 0) atomicOr(p.err, 1);
      return;
    }
    CompactArgs ca;
    ca.st = st;
    ca.lb_in = p.q_lb[cur];
    ca.par_in = p.q_par[cur];
    ca.pos_in = p.q_pos[cur];
    ca.lb_out = p.q_lb[cur ^ 1];
    ca.par_out = p.q_par[cur ^ 1];
    ca.pos_out = p.q_pos[cur ^ 1];
    ca.cap = p.cap;
    ca.poll = nullptr;
    ca.tie_xor = 0u;
    ca.wmin = nullptr;
    ca.cmin = nullptr;
    ca.chunk_log2 = 0;
    ca.pad_c = 0;
    ca.band = p.sa.band;
    compact_body(ca, blk, nblk);
    if (!grid_barrier(p.bar, nblk)) {
      if (threadIdx.x == 0) atomicOr(p.err, 1);
      return;
    }
  }
}
#endif

template <int R>
__device__ __forceinline__ void eval_impl(const EvalArgs &a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  iv<double> x[GPUGO_D];
#pragma unroll
  for (int j = 0; j < GPUGO_D; ++j) x[j] = mk<double>(a.lo[i * GPUGO_D + j], a.hi[i * GPUGO_D + j]);
  if (a.flo) {
    iv<double> f = obj_value<R>(x, a.params, a.d);
    a.flo[i] = f.lo;
    a.fhi[i] = f.hi;
  }
  if (a.glo) {
    iv<double> g[GPUGO_D];
    obj_grad<R>(x, a.params, g, a.d);
#pragma unroll
    for (int j = 0; j < GPUGO_D; ++j) {
      a.glo[i * GPUGO_D + j] = g[j].lo;
      a.ghi[i * GPUGO_D + j] = g[j].hi;
    }
  }
  if (a.kkt) a.kkt[i] = obj_kkt<R>(x, a.params, a.dom_lo, a.dom_hi, a.d) ? 1 : 0;
  if (a.ub) a.ub[i] = sample_ub<R>(x, a.params, a.n_samples, a.d);
}

#if GPUGO_WITH_F32
// FP32 sample pass of the first step of an FP32-prune solve, before
// split+bound (the reference's order: sampling, 1-Ackley/ackley50.py:602-684,
// then rule-out against the lowered incumbent, :686-766).  The first step
// has no incumbent yet, so the FP32 value test could rule nothing out
// there.  The FP32 value code list at each child's S diagonal points, each
// coordinate rounded outward (the enclosure contains f at the FP64 point: a
// rigorous upper bound of the minimum); their minimum lowers the incumbent
// and, by the last workgroup (ticket in st->cblocks, which it leaves at 0),
// the step's pruning bound ub_snap.
template <int R>
__device__ __forceinline__ void sample32_impl(const BoundArgs &a) {
  __shared__ unsigned long long s_ubmin;
  DevState *st = a.st;
  if (st->status != 0) return;
  const long long n_child = (long long)st->n_split * (long long)a.W;
  const long long first = st->first_parent;
  const double invW = 1.0 / (double)a.W, halfInvW = 0.5 / (double)a.W;
  const int lane = threadIdx.x & 63;
  if (threadIdx.x == 0) s_ubmin = order_key(__builtin_huge_val());
  __syncthreads();
  double m = __builtin_huge_val();
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n_child;
       t += (long long)gridDim.x * blockDim.x) {
    const long long pl = (long long)__builtin_fma((double)t, invW, halfInvW);
    const unsigned pos = (unsigned)(t - pl * (long long)a.W);
    const long long p = first + pl;
    const ChildBox x = child_box(a.plo + p * GPUGO_D, a.phi + p * GPUGO_D, a.pw + p * GPUGO_D, a.pwin[p], GPUGO_D,
                                 a.c, (unsigned)a.wid, a.inv_wid, a.half_inv_wid, pos);
    for (int ts = 0; ts < a.n_samples; ++ts) {
      const DiagPoint<ChildBox> xs{x, ts, a.n_samples};
      const F32Box<DiagPoint<ChildBox>> xf{xs};
      const double u = (double)obj_value32<R>(xf, (const iv<float> *)a.params32, a.d).hi;
      m = u < m ? u : m;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double t2 = __shfl_xor(m, o, 64);
    m = t2 < m ? t2 : m;
  }
  if (lane == 0 && m < __builtin_huge_val()) atomicMin(&s_ubmin, order_key(m));
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_ubmin < order_key(__builtin_huge_val())) atomicMin(&st->ub_key, s_ubmin);
    __threadfence();
    if (atomicAdd(&st->cblocks, 1) == (int)gridDim.x - 1) {
      __threadfence();
      const double ub = key_value(__hip_atomic_load(&st->ub_key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      if (ub < st->ub_snap) st->ub_snap = ub;
      st->cblocks = 0;
    }
  }
}
#endif

#if GPUGO_WITH_F32
// FP32 value enclosure of explicit boxes (parity hook of the pruning pass);
// a.ub carries the coefficient intervals (iv<float>[n_params])
template <int R>
__device__ __forceinline__ void eval32_impl(const EvalArgs &a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  iv<double> x[GPUGO_D];
#pragma unroll
  for (int j = 0; j < GPUGO_D; ++j) x[j] = mk<double>(a.lo[i * GPUGO_D + j], a.hi[i * GPUGO_D + j]);
  const F32Box<iv<double>[GPUGO_D]> xf{x};
  const iv<float> f = obj_value32<R>(xf, (const iv<float> *)a.ub, a.d);
  a.flo[i] = (double)f.lo;
  a.fhi[i] = (double)f.hi;
}
#endif

}  // namespace gpugo

// Rounding variants instantiated in this module (the generator defines
// GPUGO_WITH_EXACT / GPUGO_WITH_FAST); the library looks the kernels up by
// name and reports a missing variant as an error.
#if GPUGO_WITH_EXACT
extern "C" __global__ __launch_bounds__(256) void gpugo_bound_exact(gpugo::BoundArgs a) {
  gpugo::bound_impl<gpugo::ROUND_EXACT, false>(a);
}
extern "C" __global__ __launch_bounds__(256) void gpugo_bound_exact_mv(gpugo::BoundArgs a) {
  gpugo::bound_impl<gpugo::ROUND_EXACT, true>(a);
}
extern "C" __global__ __launch_bounds__(64) void gpugo_eval_exact(gpugo::EvalArgs a) {
  gpugo::eval_impl<gpugo::ROUND_EXACT>(a);
}
#endif
#if GPUGO_SPLIT_G > 1 && GPUGO_WITH_EXACT
extern "C" __global__ __launch_bounds__(1024) void gpugo_persist_exact_g(gpugo::PersistArgs p) {
  gpugo::persist_impl<gpugo::ROUND_EXACT, false>(p);
}
extern "C" __global__ __launch_bounds__(1024) void gpugo_persist_exact_mv_g(gpugo::PersistArgs p) {
  gpugo::persist_impl<gpugo::ROUND_EXACT, true>(p);
}
extern "C" __global__ __launch_bounds__(256) void gpugo_bound_exact_g(gpugo::BoundArgs a) {
  gpugo::bound_phase_g<gpugo::ROUND_EXACT, false>(a, blockIdx.x, gridDim.x);
}
extern "C" __global__ __launch_bounds__(256) void gpugo_bound_exact_mv_g(gpugo::BoundArgs a) {
  gpugo::bound_phase_g<gpugo::ROUND_EXACT, true>(a, blockIdx.x, gridDim.x);
}
#endif
#if GPUGO_SPLIT_G > 1 && GPUGO_WITH_FAST
extern "C" __global__ __launch_bounds__(1024) void gpugo_persist_fast_g(gpugo::PersistArgs p) {
  gpugo::persist_impl<gpugo::ROUND_FAST, false>(p);
}
extern "C" __global__ __launch_bounds__(1024) void gpugo_persist_fast_mv_g(gpugo::PersistArgs p) {
  gpugo::persist_impl<gpugo::ROUND_FAST, true>(p);
}
extern "C" __global__ __launch_bounds__(256) void gpugo_bound_fast_g(gpugo::BoundArgs a) {
  gpugo::bound_phase_g<gpugo::ROUND_FAST, false>(a, blockIdx.x, gridDim.x);
}
extern "C" __global__ __launch_bounds__(256) void gpugo_bound_fast_mv_g(gpugo::BoundArgs a) {
  gpugo::bound_phase_g<gpugo::ROUND_FAST, true>(a, blockIdx.x, gridDim.x);
}
#endif
#if GPUGO_WITH_F32 && GPUGO_WITH_EXACT
extern "C" __global__ __launch_bounds__(64) void gpugo_eval32_exact(gpugo::EvalArgs a) {
  gpugo::eval32_impl<gpugo::ROUND_EXACT>(a);
}
extern "C" __global__ __launch_bounds__(256) void gpugo_sample32_exact(gpugo::BoundArgs a) {
  gpugo::sample32_impl<gpugo::ROUND_EXACT>(a);
}
#endif
#if GPUGO_WITH_F32 && GPUGO_WITH_FAST
extern "C" __global__ __launch_bounds__(64) void gpugo_eval32_fast(gpugo::EvalArgs a) {
  gpugo::eval32_impl<gpugo::ROUND_FAST>(a);
}
extern "C" __global__ __launch_bounds__(256) void gpugo_sample32_fast(gpugo::BoundArgs a) {
  gpugo::sample32_impl<gpugo::ROUND_FAST>(a);
}
#endif
#if GPUGO_WITH_DIGITS15
// the IA of the paper's results: 15-significant-digit outward rounding
extern "C" __global__ __launch_bounds__(256) void gpugo_bound_digits15(gpugo::BoundArgs a) {
  gpugo::bound_impl<gpugo::ROUND_DIGITS15, false>(a);
}
extern "C" __global__ __launch_bounds__(256) void gpugo_bound_digits15_mv(gpugo::BoundArgs a) {
  gpugo::bound_impl<gpugo::ROUND_DIGITS15, true>(a);
}
extern "C" __global__ __launch_bounds__(64) void gpugo_eval_digits15(gpugo::EvalArgs a) {
  gpugo::eval_impl<gpugo::ROUND_DIGITS15>(a);
}
#endif
#if GPUGO_WITH_FAST
extern "C" __global__ __launch_bounds__(256) void gpugo_bound_fast(gpugo::BoundArgs a) {
  gpugo::bound_impl<gpugo::ROUND_FAST, false>(a);
}
extern "C" __global__ __launch_bounds__(256) void gpugo_bound_fast_mv(gpugo::BoundArgs a) {
  gpugo::bound_impl<gpugo::ROUND_FAST, true>(a);
}
extern "C" __global__ __launch_bounds__(64) void gpugo_eval_fast(gpugo::EvalArgs a) {
  gpugo::eval_impl<gpugo::ROUND_FAST>(a);
}
#endif
// declared coefficient signs (+1: >= 0, -1: <= 0, 0: any); the library
// rejects problems that violate them because the code lists assume them
extern "C" __global__ void gpugo_param_signs(signed char *out) {
  const signed char s[] = {GPUGO_PSIGN};
  for (int i = threadIdx.x; i < GPUGO_NPARAMS; i += blockDim.x) out[i] = s[i];
}
extern "C" __global__ void gpugo_info(int *out) {
  out[0] = GPUGO_D;
  out[1] = GPUGO_NPARAMS;
  out[2] = GPUGO_MODULE_ABI_V;
  out[3] = (GPUGO_WITH_EXACT ? 1 : 0) | (GPUGO_WITH_FAST ? 2 : 0) | (GPUGO_WITH_DIGITS15 ? 4 : 0);
  out[4] = GPUGO_WITH_F32;
  out[5] = GPUGO_SPLIT_G;
}
