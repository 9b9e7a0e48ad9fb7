// bnb_window.hpp -- kernels of a windowed objective module (large d).
// Included (as text, by gpugo_amd.emit.window_module_source) after ia.hpp,
// bnb_args.hpp, bnb_box.hpp, the windowed code lists (obj_rest,
// obj_value_w, obj_kkt_w), the loop-form code lists (obj_value_g,
// obj_grad_g, obj_kkt_g) and bnb_step.hpp.
//
// A step splits a parent record in the c coordinates of its window
// (win + t) mod d, t < c (1-Ackley/ackley50.py:482-594); every child
// shares the parent's other d - c coordinates.  So each template sum of the
// objective is enclosed as
//   rest (the terms that do not touch the window, over the parent record,
//         summed once per parent by gpugo_prep_*)
//   + the terms that do (over the child)
// and a child costs O(c) terms instead of O(d).  The reference evaluates all
// d coordinates for every child (6-Levy/levy10000.py:160-232).  Rest sums
// are also kept at the parent's S diagonal sample points, for the samples
// of the incumbent (the reference's numSamples diagonal points, :235-380).
//
// Reference kernels replaced: sampling and ruleOutGPU
// (GPUGO_1.1/6-Levy/levy10000.py:686-850).
#pragma once

#ifndef GPUGO_WITH_F32
#define GPUGO_WITH_F32 0
#endif

namespace gpugo {

// FP32 pruning pass of the windowed form (precision F32): the child's value
// in FP32 intervals (coordinates rounded outward, the parent's FP64 rest
// sums rounded outward); a child whose FP32 lower bound exceeds the
// incumbent is ruled out without FP64 work.  Rigorous: the FP32 enclosure
// contains f over the child.
template <int R>
__device__ __forceinline__ bool f32_prunes_w(const BoundArgs &a, const ChildBox &x, const iv<double> *rest, int win,
                                             double ub0) {
#if GPUGO_WITH_F32
  if (!(a.flags & GPUGO_FLAG_F32_PRUNE) || !(ub0 < __builtin_huge_val())) return false;
  const F32Box<ChildBox> xf{x};
  const iv<float> f = obj_value_w32<R>(xf, (const iv<float> *)a.params32, a.d, rest, win, a.c);
  return (double)f.lo > ub0;
#else
  return false;
#endif
}

#if GPUGO_TAB
// Entries [blk * 256, blk * 256 + 256) of the term tables of record
// first_parent + b (layout tab_layout): each thread evaluates one term, or
// one gradient component, over the TabBox of its digit combination, at the
// record's box or at one of its diagonal sample points, in FP64 and, for
// the FP32 pass, in FP32.  Entries of terms the window does not reach, or
// of digit combinations no child has, are not written: no child reads them.
template <int R>
__device__ __forceinline__ void tab_build(const BoundArgs &a, int b, int blk) {
  if (a.tab_n == 0) return;
  const int e = blk * (int)blockDim.x + (int)threadIdx.x;
  const int S = a.n_samples, d = a.d, c = a.c;
  const unsigned wid = (unsigned)a.wid;
  const TabLayout L = tab_layout(c, wid, S, d);
  if (e >= L.total) return;
  const long long p = a.st->first_parent + b;
  const int win = a.pwin[p];
  const int NT = L.off[GPUGO_NREST];
  int s, q, pt, w, klo, nr, rlo[2], rhi[2];
  unsigned combo;
#if GPUGO_TAB_PARTS
  if (e >= L.poff) {
    // parts table: part k of component j over the record's box (r < d), or
    // over the TabBox of a window component's digit combination
    iv<double> *T = (iv<double> *)a.ptab + (long long)b * a.tab_n;
    const int k = (e - L.poff) / L.pstride, r = (e - L.poff) - k * L.pstride;
    if (r < d) {
      const RowBox xb{a.plo + p * GPUGO_D, a.phi + p * GPUGO_D};
      T[e] = wgpart<R>(k, xb, a.params, d, r);
      return;
    }
    q = (r - d) / (int)L.gE;
    combo = (unsigned)(r - d) - (unsigned)q * L.gE;
    nr = win_ranges(win, c, d, GPUGO_TAB_GKLO, GPUGO_TAB_GKLO + GPUGO_TAB_GW - 1, d, rlo, rhi);
    const int n0 = nr > 0 ? rhi[0] - rlo[0] : 0, nq = n0 + (nr > 1 ? rhi[1] - rlo[1] : 0);
    if (q >= nq) return;
    const int j = q < n0 ? rlo[0] + q : rlo[1] + (q - n0);
    unsigned rr = combo;
    for (int t = 0; t < GPUGO_TAB_GW; ++t) {
      const int i = j + GPUGO_TAB_GKLO + t;
      int kk = i - win;
      if (kk < 0) kk += d;
      if (!(i >= 0 && i < d && kk < c) && (rr % wid) != 0u) return;
      rr /= wid;
    }
    const TabBox tb{a.plo + p * GPUGO_D, a.phi + p * GPUGO_D, a.pw + p * GPUGO_D, win, d, c, j + GPUGO_TAB_GKLO, wid,
                    combo};
    T[e] = wgpart<R>(k, tb, a.params, d, j);
    return;
  }
#endif
  if (e < L.goff) {
    pt = e / NT;
    const int e2 = e - pt * NT;
    s = 0;
    while (s + 1 < GPUGO_NREST && e2 >= L.off[s + 1]) ++s;
    q = (e2 - L.off[s]) / (int)L.E[s];
    combo = (unsigned)(e2 - L.off[s]) - (unsigned)q * L.E[s];
    w = tab_w(s);
    klo = tab_klo(s);
    nr = wranges_any(s, win, c, d, rlo, rhi);
  } else {
    pt = 0;
    s = -1;
    const int e2 = e - L.goff;
    q = e2 / (int)L.gE;
    combo = (unsigned)e2 - (unsigned)q * L.gE;
    w = GPUGO_TAB_GW;
    klo = GPUGO_TAB_GKLO;
    nr = win_ranges(win, c, d, GPUGO_TAB_GKLO, GPUGO_TAB_GKLO + GPUGO_TAB_GW - 1, d, rlo, rhi);
  }
  // term (component) j: the q-th of the window ranges, in the order the
  // code lists walk them
  const int n0 = nr > 0 ? rhi[0] - rlo[0] : 0, nq = n0 + (nr > 1 ? rhi[1] - rlo[1] : 0);
  if (q >= nq) return;
  const int j = q < n0 ? rlo[0] + q : rlo[1] + (q - n0);
  unsigned r = combo;
  for (int k = 0; k < w; ++k) {
    const int i = j + klo + k;
    int kk = i - win;
    if (kk < 0) kk += d;
    if (!(i >= 0 && i < d && kk < c) && (r % wid) != 0u) return;
    r /= wid;
  }
  const TabBox tb{a.plo + p * GPUGO_D, a.phi + p * GPUGO_D, a.pw + p * GPUGO_D, win, d, c, j + klo, wid, combo};
  iv<double> *T = (iv<double> *)a.ptab + (long long)b * a.tab_n;
  if (s < 0) {
#if GPUGO_TAB_GRAD
    T[e] = wgrad_t<R>(tb, a.params, d, j);
#endif
    return;
  }
  if (pt == 0) {
    T[e] = wterm_any<R>(s, tb, a.params, d, j);
  } else {
    const DiagPoint<TabBox> xs{tb, pt - 1, S};
    T[e] = wterm_any<R>(s, xs, a.params, d, j);
  }
#if GPUGO_WITH_F32
  if (a.tab_n32 > 0) {
    iv<float> *T32 = (iv<float> *)a.ptab32 + (long long)b * a.tab_n32;
    const iv<float> *P32 = (const iv<float> *)a.params32;
    if (pt == 0) {
      T32[e] = wterm32_any<R>(s, F32Box<TabBox>{tb}, P32, d, j);
    } else {
      const DiagPoint<TabBox> xs{tb, pt - 1, S};
      T32[e] = wterm32_any<R>(s, F32Box<DiagPoint<TabBox>>{xs}, P32, d, j);
    }
  }
#endif
}
#endif

#if !GPUGO_TAB
// objectives whose stencils are too wide for term tables
struct TabLayout {};
__device__ __forceinline__ TabLayout tab_layout(int, unsigned, int, int) { return TabLayout{}; }
#endif

// The tests of one child (record first_parent + b) from the term tables
// when the launch has them (tab_n > 0), else from the terms themselves;
// both give the same bits.  pt: 0 the child, 1 + t its diagonal sample t;
// rest_pt: the record's rest sums at that point.
template <int R>
__device__ __forceinline__ iv<double> value_wx(const BoundArgs &a, const ChildBox &x, long long b, int pt,
                                               const iv<double> *rest_pt, int win, const TabLayout &L) {
#if GPUGO_TAB
  if (a.tab_n > 0) {
    const iv<double> *T = (const iv<double> *)a.ptab + b * a.tab_n + (long long)pt * L.off[GPUGO_NREST];
    if (pt == 0) return obj_value_wt<R>(x, x, T, L, a.params, a.d, rest_pt, win, a.c);
    const DiagPoint<ChildBox> xs{x, pt - 1, a.n_samples};
    return obj_value_wt<R>(xs, x, T, L, a.params, a.d, rest_pt, win, a.c);
  }
#endif
  if (pt == 0) return obj_value_w<R>(x, a.params, a.d, rest_pt, win, a.c);
  const DiagPoint<ChildBox> xs{x, pt - 1, a.n_samples};
  return obj_value_w<R>(xs, a.params, a.d, rest_pt, win, a.c);
}

template <int R>
__device__ __forceinline__ bool kkt_wx(const BoundArgs &a, const ChildBox &x, long long b, const iv<double> *rest,
                                       int win, const TabLayout &L) {
#if GPUGO_TAB && GPUGO_TAB_GRAD
  if (a.tab_n > 0)
    return obj_kkt_wt<R>(x, (const iv<double> *)a.ptab + b * a.tab_n + L.goff, L, a.dom_lo, a.dom_hi, a.d, win, a.c);
#elif GPUGO_TAB && GPUGO_TAB_PARTS
  if (a.tab_n > 0)
    return obj_kkt_wl<R>(x, (const iv<double> *)a.ptab + b * a.tab_n, L, a.params, a.dom_lo, a.dom_hi, a.d, rest,
                         win, a.c);
#elif GPUGO_TAB
  if (a.tab_n > 0)
    return obj_kkt_wr<R>(x, (const iv<double> *)a.ptab + b * a.tab_n, L, a.params, a.dom_lo, a.dom_hi, a.d, rest,
                         win, a.c);
#endif
  return obj_kkt_w<R>(x, a.params, a.dom_lo, a.dom_hi, a.d, rest, win, a.c);
}

template <int R>
__device__ __forceinline__ bool f32_prunes_wx(const BoundArgs &a, const ChildBox &x, long long b,
                                              const iv<double> *rest, int win, double ub0, const TabLayout &L) {
#if GPUGO_WITH_F32 && GPUGO_TAB
  if (a.tab_n32 > 0) {
    if (!(a.flags & GPUGO_FLAG_F32_PRUNE) || !(ub0 < __builtin_huge_val())) return false;
    const F32Box<ChildBox> xf{x};
    const iv<float> f = obj_value_wt32<R>(xf, x, (const iv<float> *)a.ptab32 + b * a.tab_n32, L,
                                          (const iv<float> *)a.params32, a.d, rest, win, a.c);
    return (double)f.lo > ub0;
  }
#endif
  return f32_prunes_w<R>(a, x, rest, win, ub0);
}

// Rest sums (and products) of the parent records split in this step: one
// workgroup per record, threads over the terms, a butterfly in each
// wavefront and a fixed-order fold over the wavefronts (deterministic).  Per record
// (1 + S) * NREST intervals: [0, NREST) over the box, then NREST at each of
// its S diagonal sample points (DiagPoint t = 0 .. S-1).
template <int R>
__device__ __forceinline__ void prep_impl(const BoundArgs &a) {
  typedef IA<double, R> A;
  __shared__ iv<double> part[16];
  const DevState *st = a.st;
  if (st->status != 0) return;
  // grid (pairs + table blocks, parents): workgroup (q, b) with q < pairs
  // sums template s = q mod NREST at point k = q / NREST (0: the parent
  // box, k > 0: diagonal sample k - 1); the workgroups after them fill the
  // record's term tables
  const int b = blockIdx.y;
  if (b >= st->n_split) return;
  const int S = a.n_samples;
  const int npairs = (1 + S) * GPUGO_NREST, rs = a.rsplit;
  if ((int)blockIdx.x >= npairs * rs) {
#if GPUGO_TAB
    tab_build<R>(a, b, (int)blockIdx.x - npairs * rs);
#endif
    return;
  }
  // pair (k, s) and chunk ch of its terms: j in [ch * CH, (ch + 1) * CH)
  const int pair = (int)blockIdx.x / rs, ch = (int)blockIdx.x - pair * rs;
  const int k = pair / GPUGO_NREST, s = pair - k * GPUGO_NREST;
  const int CH = (a.d + 2 + rs - 1) / rs;  // every template runs over fewer than d + 2 terms
  const long long p = st->first_parent + b;
  const int win = a.pwin[p];
  const RowBox x{a.plo + p * GPUGO_D, a.phi + p * GPUGO_D};
  iv<double> *out = (iv<double> *)(a.prest) + p * (long long)((1 + S) * GPUGO_NREST);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nw = blockDim.x >> 6;
  const int j0 = ch * CH, j1 = j0 + CH;
  iv<double> v = k == 0 ? obj_rest<R>(s, x, a.params, a.d, win, a.c, j0 + tid, (int)blockDim.x, j1)
                        : obj_rest<R>(s, DiagPoint<RowBox>{x, k - 1, S}, a.params, a.d, win, a.c, j0 + tid,
                                      (int)blockDim.x, j1);
  const bool prod = rest_is_prod(s);
  v = prod ? group_prod<R, 64>(v) : group_sum<R, 64>(v);
  if (lane == 0) part[w] = v;
  __syncthreads();
  if (tid == 0) {
    iv<double> t = part[0];
    for (int i = 1; i < nw; ++i) t = prod ? A::mul(t, part[i]) : A::add(t, part[i]);
    if (rs == 1) out[k * GPUGO_NREST + s] = t;
    else ((iv<double> *)a.rpart)[((long long)b * npairs + pair) * rs + ch] = t;
  }
}

// The rest sums split over several workgroups (rsplit > 1): fold each
// pair's partials in chunk order (deterministic), one thread per pair.
template <int R>
__device__ __forceinline__ void prep_fold_impl(const BoundArgs &a) {
  typedef IA<double, R> A;
  const DevState *st = a.st;
  if (st->status != 0) return;
  const int b = blockIdx.y;
  if (b >= st->n_split) return;
  const int S = a.n_samples, npairs = (1 + S) * GPUGO_NREST, rs = a.rsplit;
  const int pair = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (pair >= npairs) return;
  const int s = pair % GPUGO_NREST;
  const bool prod = rest_is_prod(s);
  const iv<double> *pp = (const iv<double> *)a.rpart + ((long long)b * npairs + pair) * rs;
  iv<double> t = pp[0];
  for (int i = 1; i < rs; ++i) t = prod ? A::mul(t, pp[i]) : A::add(t, pp[i]);
  const long long p = st->first_parent + b;
  ((iv<double> *)(a.prest) + p * (long long)npairs)[pair] = t;
}

// Wavefronts where at most this many children reach an FP64 test evaluate
// them one at a time with all 64 lanes (obj_value_wp / obj_kkt_wp): one
// child's FP64 tests on one lane would otherwise be the launch's critical
// path (a lone lane runs its whole dependent chain with nothing to overlap).
#ifndef GPUGO_COOP_MAX
#define GPUGO_COOP_MAX 8
#endif

// The FP64 tests of child (p, pos) by the whole wavefront: value (unless
// lb is given), then, if lb <= ub0, the monotonicity test and the S
// diagonal samples.  Every lane returns the same results.
template <int R>
__device__ __forceinline__ void coop_tests(const BoundArgs &a, long long p, unsigned pos, double ub0, bool value,
                                           int lane, double &lb, bool &keep, bool &pv, bool &pk, double &su) {
  const int win = a.pwin[p];
  const ChildBox x = child_box(a.plo + p * GPUGO_D, a.phi + p * GPUGO_D, a.pw + p * GPUGO_D, win, GPUGO_D, a.c,
                               (unsigned)a.wid, a.inv_wid, a.half_inv_wid, pos);
  const int S = a.n_samples;
  const iv<double> *rest = (const iv<double> *)(a.prest) + p * (long long)((1 + S) * GPUGO_NREST);
  if (value) {
    lb = obj_value_wp<R, 64>(x, a.params, a.d, rest, win, a.c, lane).lo;
    if (!(lb == lb)) lb = -__builtin_huge_val();
  }
  keep = false;
  pv = false;
  pk = false;
  su = __builtin_huge_val();
  if (lb <= ub0) {
    keep = !obj_kkt_wp<R, 64>(x, a.params, a.dom_lo, a.dom_hi, a.d, rest, win, a.c, lane);
    pk = !keep;
    for (int ts = 0; ts < S; ++ts) {
      const DiagPoint<ChildBox> xs{x, ts, S};
      const double u = obj_value_wp<R, 64>(xs, a.params, a.d, rest + (1 + ts) * GPUGO_NREST, win, a.c, lane).hi;
      su = u < su ? u : su;
    }
  } else {
    pv = true;
  }
}

// Split+bound with windowed code lists: value test, then for survivors the
// monotonicity test and the diagonal samples; survivors are appended to the
// queue (one atomic per workgroup), the incumbent is lowered by one atomic
// per workgroup (same protocol as bound_impl in bnb_kernel.hpp).  A
// wavefront with few children left for an FP64 stage (after the FP32 test,
// or after the FP64 value test) runs that stage cooperatively, one child at
// a time (coop_tests).
template <int R>
__device__ __forceinline__ void bound_w_impl(const BoundArgs &a) {
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
  const TabLayout TL = tab_layout(a.c, (unsigned)a.wid, a.n_samples, a.d);
  if (threadIdx.x == 0) {
    s_ubmin = order_key(__builtin_huge_val());
    s_pv = 0;
    s_pk = 0;
  }
  __syncthreads();
  for (long long base = (long long)blockIdx.x * 256; base < n_child; base += (long long)gridDim.x * 256) {
    const long long t = base + threadIdx.x;
    long long pl = 0;
    unsigned pos = 0;
    double lb = __builtin_huge_val(), su = __builtin_huge_val();
    bool keep = false, pv = false, pk = false;
    bool need = false;  // the child reaches the FP64 value test
    if (t < n_child) {
      pl = (long long)__builtin_fma((double)t, invW, halfInvW);
      pos = (unsigned)(t - pl * (long long)a.W);
      const long long p = first + pl;
      const int win = a.pwin[p];
      const ChildBox x = child_box(a.plo + p * GPUGO_D, a.phi + p * GPUGO_D, a.pw + p * GPUGO_D, win, GPUGO_D, a.c,
                                   (unsigned)a.wid, a.inv_wid, a.half_inv_wid, pos);
      const iv<double> *rest = (const iv<double> *)(a.prest) + p * (long long)((1 + a.n_samples) * GPUGO_NREST);
      if (f32_prunes_wx<R>(a, x, pl, rest, win, ub0, TL)) pv = true;
      else need = true;
    }
    // the wavefront's children in each FP64 stage: few -> all lanes on each
    const unsigned long long need_m = __ballot(need);
    bool stage2 = false;  // value computed on the child's own lane; tests next
    if (need_m != 0ull && __popcll(need_m) <= GPUGO_COOP_MAX) {
      unsigned long long m = need_m;
      while (m != 0ull) {
        const int src = __builtin_ctzll(m);
        m &= m - 1ull;
        const long long ps = first + (long long)__shfl((int)pl, src, 64);
        const unsigned poss = (unsigned)__shfl((int)pos, src, 64);
        double l2 = 0.0, s2;
        bool k2, v2, q2;
        coop_tests<R>(a, ps, poss, ub0, true, lane, l2, k2, v2, q2, s2);
        if (lane == src) {
          lb = l2;
          keep = k2;
          pv = v2;
          pk = q2;
          su = s2;
        }
      }
    } else if (need) {
      const long long p = first + pl;
      const int win = a.pwin[p];
      const ChildBox x = child_box(a.plo + p * GPUGO_D, a.phi + p * GPUGO_D, a.pw + p * GPUGO_D, win, GPUGO_D, a.c,
                                   (unsigned)a.wid, a.inv_wid, a.half_inv_wid, pos);
      const iv<double> *rest = (const iv<double> *)(a.prest) + p * (long long)((1 + a.n_samples) * GPUGO_NREST);
      const iv<double> f = value_wx<R>(a, x, pl, 0, rest, win, TL);
      lb = f.lo;
      if (!(lb == lb)) lb = -__builtin_huge_val();
      if (lb <= ub0) stage2 = true;
      else pv = true;
    }
    const unsigned long long s2_m = __ballot(stage2);
    if (s2_m != 0ull && __popcll(s2_m) <= GPUGO_COOP_MAX) {
      unsigned long long m = s2_m;
      while (m != 0ull) {
        const int src = __builtin_ctzll(m);
        m &= m - 1ull;
        const long long ps = first + (long long)__shfl((int)pl, src, 64);
        const unsigned poss = (unsigned)__shfl((int)pos, src, 64);
        double l2 = __shfl(lb, src, 64), s2;
        bool k2, v2, q2;
        coop_tests<R>(a, ps, poss, ub0, false, lane, l2, k2, v2, q2, s2);
        if (lane == src) {
          keep = k2;
          pv = v2;
          pk = q2;
          su = s2;
        }
      }
    } else if (stage2) {
      const long long p = first + pl;
      const int win = a.pwin[p];
      const ChildBox x = child_box(a.plo + p * GPUGO_D, a.phi + p * GPUGO_D, a.pw + p * GPUGO_D, win, GPUGO_D, a.c,
                                   (unsigned)a.wid, a.inv_wid, a.half_inv_wid, pos);
      const int S = a.n_samples;
      const iv<double> *rest = (const iv<double> *)(a.prest) + p * (long long)((1 + S) * GPUGO_NREST);
      keep = !kkt_wx<R>(a, x, pl, rest, win, TL);
      pk = !keep;
      // S points on the child's diagonal: window coordinates from the
      // child, the others from the parent's same sample point (its rest)
      for (int ts = 0; ts < S; ++ts) {
        const double u = value_wx<R>(a, x, pl, 1 + ts, rest + (1 + ts) * GPUGO_NREST, win, TL).hi;
        su = u < su ? u : su;
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

// Gradient sink writing component j of box i to explicit arrays.
struct RowSink {
  double *glo, *ghi;
  struct Ref {
    double *lo, *hi;
    __device__ __forceinline__ void operator=(const iv<double> &g) const {
      *lo = g.lo;
      *hi = g.hi;
    }
  };
  __device__ __forceinline__ Ref operator[](int j) const { return Ref{glo + j, ghi + j}; }
};

// Explicit boxes (parity hook): the loop-form code lists over all d
// coordinates (one lane per box), the same functions the windowed lists
// split into rest + window terms.
template <int R>
__device__ __forceinline__ void eval_w_impl(const EvalArgs &a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const RowBox x{a.lo + i * GPUGO_D, a.hi + i * GPUGO_D};
  if (a.flo) {
    const iv<double> f = obj_value_g<R, 1>(x, a.params, a.d, 0);
    a.flo[i] = f.lo;
    a.fhi[i] = f.hi;
  }
  if (a.glo) obj_grad_g<R, 1>(x, a.params, RowSink{a.glo + i * GPUGO_D, a.ghi + i * GPUGO_D}, a.d, 0);
  if (a.kkt) a.kkt[i] = obj_kkt_g<R, 1>(x, a.params, a.dom_lo, a.dom_hi, a.d, 0) ? 1 : 0;
  if (a.ub) {
    double best = __builtin_huge_val();
    for (int t = 0; t < a.n_samples; ++t) {
      const DiagPoint<RowBox> xs{x, t, a.n_samples};
      const iv<double> f = obj_value_g<R, 1>(xs, a.params, a.d, 0);
      if (f.hi < best) best = f.hi;
    }
    a.ub[i] = best;
  }
}

#if GPUGO_WITH_F32
// FP32 sample pass of the first step of an FP32-prune solve (windowed
// form): the FP32 value at each child's S diagonal points, the window's
// coordinates from the child and the other terms from the parent's rest sums
// at the same sample point; their minimum lowers the incumbent and, by the
// last workgroup (ticket in st->cblocks, left at 0), the step's pruning
// bound ub_snap.  Same protocol as the unrolled modules' sample pass.
template <int R>
__device__ __forceinline__ void sample32_w_impl(const BoundArgs &a) {
  __shared__ unsigned long long s_ubmin;
  DevState *st = a.st;
  if (st->status != 0) return;
  const long long n_child = (long long)st->n_split * (long long)a.W;
  const long long first = st->first_parent;
  const double invW = 1.0 / (double)a.W, halfInvW = 0.5 / (double)a.W;
  const int lane = threadIdx.x & 63, S = a.n_samples;
  const TabLayout TL = tab_layout(a.c, (unsigned)a.wid, S, a.d);
  if (threadIdx.x == 0) s_ubmin = order_key(__builtin_huge_val());
  __syncthreads();
  double m = __builtin_huge_val();
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n_child;
       t += (long long)gridDim.x * blockDim.x) {
    const long long pl = (long long)__builtin_fma((double)t, invW, halfInvW);
    const unsigned pos = (unsigned)(t - pl * (long long)a.W);
    const long long p = first + pl;
    const int win = a.pwin[p];
    const ChildBox x = child_box(a.plo + p * GPUGO_D, a.phi + p * GPUGO_D, a.pw + p * GPUGO_D, win, GPUGO_D, a.c,
                                 (unsigned)a.wid, a.inv_wid, a.half_inv_wid, pos);
    const iv<double> *rest = (const iv<double> *)(a.prest) + p * (long long)((1 + S) * GPUGO_NREST);
    for (int ts = 0; ts < S; ++ts) {
      double u;
#if GPUGO_TAB
      if (a.tab_n32 > 0) {
        const DiagPoint<ChildBox> xs{x, ts, S};
        const F32Box<DiagPoint<ChildBox>> xf{xs};
        u = (double)obj_value_wt32<R>(xf, x,
                                      (const iv<float> *)a.ptab32 + pl * a.tab_n32 + (1 + ts) * TL.off[GPUGO_NREST],
                                      TL, (const iv<float> *)a.params32, a.d, rest + (1 + ts) * GPUGO_NREST, win, a.c)
                .hi;
      } else
#endif
      {
        const DiagPoint<ChildBox> xs{x, ts, S};
        const F32Box<DiagPoint<ChildBox>> xf{xs};
        u = (double)obj_value_w32<R>(xf, (const iv<float> *)a.params32, a.d, rest + (1 + ts) * GPUGO_NREST, win, a.c)
                .hi;
      }
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

// FP32 value enclosure of explicit boxes (parity hook of the windowed
// pruning pass): the windowed FP32 list with the window covering every
// coordinate (c = d) and empty rest sums, so every term is evaluated over
// the box; a.ub carries the coefficient intervals (iv<float>[n_params]).
template <int R>
__device__ __forceinline__ void eval32_w_impl(const EvalArgs &a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const RowBox x{a.lo + i * GPUGO_D, a.hi + i * GPUGO_D};
  const F32Box<RowBox> xf{x};
  iv<double> rest[GPUGO_NREST];
  for (int s = 0; s < GPUGO_NREST; ++s) rest[s] = rest_is_prod(s) ? mk<double>(1.0, 1.0) : mk<double>(0.0, 0.0);
  const iv<float> f = obj_value_w32<R>(xf, (const iv<float> *)a.ub, a.d, rest, 0, a.d);
  a.flo[i] = (double)f.lo;
  a.fhi[i] = (double)f.hi;
}
#endif

}  // namespace gpugo

#if GPUGO_WITH_F32 && GPUGO_WITH_EXACT
extern "C" __global__ __launch_bounds__(256) void gpugo_sample32_exact(gpugo::BoundArgs a) {
  gpugo::sample32_w_impl<gpugo::ROUND_EXACT>(a);
}
extern "C" __global__ __launch_bounds__(64) void gpugo_eval32_exact(gpugo::EvalArgs a) {
  gpugo::eval32_w_impl<gpugo::ROUND_EXACT>(a);
}
#endif
#if GPUGO_WITH_F32 && GPUGO_WITH_FAST
extern "C" __global__ __launch_bounds__(256) void gpugo_sample32_fast(gpugo::BoundArgs a) {
  gpugo::sample32_w_impl<gpugo::ROUND_FAST>(a);
}
extern "C" __global__ __launch_bounds__(64) void gpugo_eval32_fast(gpugo::EvalArgs a) {
  gpugo::eval32_w_impl<gpugo::ROUND_FAST>(a);
}
#endif
#if GPUGO_WITH_EXACT
extern "C" __global__ __launch_bounds__(256) void gpugo_bound_exact(gpugo::BoundArgs a) {
  gpugo::bound_w_impl<gpugo::ROUND_EXACT>(a);
}
extern "C" __global__ __launch_bounds__(256) void gpugo_prep_exact(gpugo::BoundArgs a) {
  gpugo::prep_impl<gpugo::ROUND_EXACT>(a);
}
extern "C" __global__ __launch_bounds__(64) void gpugo_prep_fold_exact(gpugo::BoundArgs a) {
  gpugo::prep_fold_impl<gpugo::ROUND_EXACT>(a);
}
extern "C" __global__ __launch_bounds__(64) void gpugo_eval_exact(gpugo::EvalArgs a) {
  gpugo::eval_w_impl<gpugo::ROUND_EXACT>(a);
}
#endif
#if GPUGO_WITH_DIGITS15
// the IA of the paper's results (15 significant digits, never narrower than EXACT)
extern "C" __global__ __launch_bounds__(256) void gpugo_bound_digits15(gpugo::BoundArgs a) {
  gpugo::bound_w_impl<gpugo::ROUND_DIGITS15>(a);
}
extern "C" __global__ __launch_bounds__(256) void gpugo_prep_digits15(gpugo::BoundArgs a) {
  gpugo::prep_impl<gpugo::ROUND_DIGITS15>(a);
}
extern "C" __global__ __launch_bounds__(64) void gpugo_prep_fold_digits15(gpugo::BoundArgs a) {
  gpugo::prep_fold_impl<gpugo::ROUND_DIGITS15>(a);
}
extern "C" __global__ __launch_bounds__(64) void gpugo_eval_digits15(gpugo::EvalArgs a) {
  gpugo::eval_w_impl<gpugo::ROUND_DIGITS15>(a);
}
#endif
#if GPUGO_WITH_FAST
extern "C" __global__ __launch_bounds__(256) void gpugo_bound_fast(gpugo::BoundArgs a) {
  gpugo::bound_w_impl<gpugo::ROUND_FAST>(a);
}
extern "C" __global__ __launch_bounds__(256) void gpugo_prep_fast(gpugo::BoundArgs a) {
  gpugo::prep_impl<gpugo::ROUND_FAST>(a);
}
extern "C" __global__ __launch_bounds__(64) void gpugo_prep_fold_fast(gpugo::BoundArgs a) {
  gpugo::prep_fold_impl<gpugo::ROUND_FAST>(a);
}
extern "C" __global__ __launch_bounds__(64) void gpugo_eval_fast(gpugo::EvalArgs a) {
  gpugo::eval_w_impl<gpugo::ROUND_FAST>(a);
}
#endif
extern "C" __global__ void gpugo_param_signs(signed char *out) {
  const signed char s[] = {GPUGO_PSIGN};
  for (int i = threadIdx.x; i < GPUGO_NPARAMS; i += blockDim.x) out[i] = s[i];
}
// out: [D, NPARAMS, ABI, rounding variants, f32, split G, rest slots (windowed)]
extern "C" __global__ void gpugo_info(int *out) {
  out[0] = GPUGO_D;
  out[1] = GPUGO_NPARAMS;
  out[2] = GPUGO_MODULE_ABI_V;
  out[3] = (GPUGO_WITH_EXACT ? 1 : 0) | (GPUGO_WITH_FAST ? 2 : 0) | (GPUGO_WITH_DIGITS15 ? 4 : 0);
  out[4] = GPUGO_WITH_F32;
  out[5] = 0;
  out[6] = GPUGO_NREST;
  // term tables: out[7] bit 0 tables, bit 1 gradient table, bit 2 parts
  // table; out[8 + s] the stencil width of slot s; out[16] the gradient's;
  // out[17] the gradient parts (the host sizes the tables with tab_layout's
  // formula)
#if GPUGO_TAB
  out[7] = 1 | (GPUGO_TAB_GRAD ? 2 : 0) | (GPUGO_TAB_PARTS ? 4 : 0);
  for (int s = 0; s < GPUGO_NREST && s < 8; ++s) out[8 + s] = gpugo::tab_w(s);
  out[16] = GPUGO_TAB_GW;
  out[17] = GPUGO_TAB_PARTS;
#else
  out[7] = 0;
#endif
}
