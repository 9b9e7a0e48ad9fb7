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

#define GPUGO_MODULE_ABI_V 1
// dimensions above this decode child coordinates on use instead of into an
// array (the array would keep 4 d registers live through the code lists)
#ifndef GPUGO_LAZY_D
#define GPUGO_LAZY_D 16
#endif

namespace gpugo {

// Rigorous upper bound of f at n_samples points of the box diagonal,
// x_j = lo_j + (hi_j - lo_j)/(S+1) * (t+1) (1-Ackley/ackley50.py:288-290);
// S = 1 is the midpoint.
template <int R>
__device__ __forceinline__ double sample_ub(const iv<double> *x, const double *__restrict__ P, int S, int d) {
  double best = __builtin_huge_val();
  for (int t = 0; t < S; ++t) {
    iv<double> xs[GPUGO_D];
#pragma unroll
    for (int j = 0; j < GPUGO_D; ++j) {
      double w = x[j].hi - x[j].lo;
      double h = (S == 1) ? w * 0.5 : w / (double)(S + 1);
      double p = __builtin_fma(h, (double)(t + 1), x[j].lo);
      xs[j] = mk<double>(p, p);
    }
    iv<double> f = obj_value<R>(xs, P, d);
    if (f.hi < best) best = f.hi;
  }
  return best;
}

// The same bound on a lazily decoded child (coordinate accessor).
template <int R>
__device__ __forceinline__ double sample_ub(const ChildBox &x, const double *__restrict__ P, int S, int d) {
  double best = __builtin_huge_val();
  for (int t = 0; t < S; ++t) {
    const DiagPoint xs{x, t, S};
    const iv<double> f = obj_value<R>(xs, P, d);
    if (f.hi < best) best = f.hi;
  }
  return best;
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

// Monotonicity test on a computed gradient enclosure
// (1-Ackley/ackley50.py:458-477).
__device__ __forceinline__ bool kkt_from_grad(const iv<double> *x, const iv<double> *g,
                                              const double *__restrict__ dlo, const double *__restrict__ dhi) {
  bool prune = false;
#pragma unroll
  for (int j = 0; j < GPUGO_D; ++j)
    prune |= (g[j].lo > 0.0 && x[j].lo != dlo[j]) || (g[j].hi < 0.0 && x[j].hi != dhi[j]);
  return prune;
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
      if (!MV && GPUGO_D > GPUGO_LAZY_D) {
        // large d: coordinates decoded where the code lists use them
        const long long p = first + pl;
        const ChildBox x = child_box(a.plo + p * GPUGO_D, a.phi + p * GPUGO_D, a.pw + p * GPUGO_D, a.pwin[p],
                                     GPUGO_D, a.c, (unsigned)a.wid, a.inv_wid, a.half_inv_wid, pos);
        bound_plain<R>(a, x, ub0, lb, su, keep, pv, pk);
      } else if (!MV) {
        iv<double> x[GPUGO_D];
        decode_child(a, first + pl, pos, x);
        bound_plain<R>(a, (const iv<double> *)x, ub0, lb, su, keep, pv, pk);
      } else {
        iv<double> x[GPUGO_D];
        decode_child(a, first + pl, pos, x);
        iv<double> f = obj_value<R>((const iv<double> *)x, a.params, a.d);
        lb = f.lo;
        if (!(lb == lb)) lb = -__builtin_huge_val();
        if (lb <= ub0) {
          // mean-value form f(c) + sum_j G_j (X_j - c_j) from the gradient
          // the monotonicity test needs anyway; c is the S = 1 sample point
          typedef IA<double, R> A;
          iv<double> g[GPUGO_D], xc[GPUGO_D];
          obj_grad<R>((const iv<double> *)x, a.params, g, a.d);
#pragma unroll
          for (int j = 0; j < GPUGO_D; ++j) {
            double c = __builtin_fma((x[j].hi - x[j].lo) * 0.5, 1.0, x[j].lo);
            xc[j] = mk<double>(c, c);
          }
          iv<double> s = obj_value<R>((const iv<double> *)xc, a.params, a.d);
          su = a.n_samples == 1 ? s.hi : sample_ub<R>((const iv<double> *)x, a.params, a.n_samples, a.d);
#pragma unroll
          for (int j = 0; j < GPUGO_D; ++j) {
            iv<double> dx = mk<double>(A::RR::sub_dn(x[j].lo, xc[j].lo), A::RR::sub_up(x[j].hi, xc[j].lo));
            s = A::add(s, A::mul(g[j], dx));
          }
          if (s.lo > lb) lb = s.lo;
          if (lb <= ub0) {
            keep = !kkt_from_grad(x, g, a.dom_lo, a.dom_hi);
            pk = !keep;
          } else {
            pv = true;
          }
        } else {
          pv = true;
        }
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
  out[3] = (GPUGO_WITH_EXACT ? 1 : 0) | (GPUGO_WITH_FAST ? 2 : 0);
}
