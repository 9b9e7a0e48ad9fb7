// bnb_box.hpp -- children of a parent record as coordinate accessors, for
// the code lists of objective and batch modules (x[j] is an interval).
// Included (as text, by gpugo_amd.emit) after ia.hpp and bnb_args.hpp.
#pragma once

namespace gpugo {

// Child `pos` of a parent record as a coordinate accessor: x[j]
// decodes coordinate j when it is used (same arithmetic as decode_child in
// bnb_kernel.hpp: window coordinates (win + k) mod d for k < c take base-wid
// digit k of pos, bounds L + w*digit by one fma, the top child ends at U),
// so no array of d intervals stays live across the code lists.
struct ChildBox {
  const double *__restrict__ plo;
  const double *__restrict__ phi;
  const double *__restrict__ pw;
  unsigned long long packed, mask;
  int win, d, c, bpd;
  unsigned wid;
  __device__ __forceinline__ iv<double> operator[](int j) const {
    // Above d = 16 the empty asm hides packed/win from the optimiser, so
    // the digit shifts and window tests are recomputed at each use instead
    // of being kept live for all d coordinates, and the parent loads stay
    // here.  Small fixed-d modules let the compiler reuse the loads (the
    // registers are there: measured 3% faster on the d = 10 chain solve).
    unsigned long long packed = this->packed;
    int win = this->win;
#if !defined(GPUGO_ACCESSOR_D) || GPUGO_ACCESSOR_D > 16
    asm volatile("" : "+v"(packed), "+v"(win)::"memory");
#endif
    const double L = plo[j], U = phi[j];
    int k = j - win;
    if (k < 0) k += d;
    if (k < c) {
      const unsigned dg = (unsigned)((packed >> (k * bpd)) & mask);
      const double w = pw[j];
      const double lo = __builtin_fma(w, (double)dg, L);
      const double hi = (dg + 1u == wid) ? U : __builtin_fma(w, (double)(dg + 1u), L);
      return mk<double>(lo, hi);
    }
    return mk<double>(L, U);
  }
};

// Sample point t of S on the diagonal of a box (an array or accessor),
// x_j = lo_j + (hi_j - lo_j)/(S+1) * (t+1) (1-Ackley/ackley50.py:288-290);
// t = 0, S = 1 is the midpoint.
template <class XS>
struct DiagPoint {
  const XS &b;
  int t, S;
  __device__ __forceinline__ iv<double> operator[](int j) const {
    const iv<double> x = b[j];
    const double w = x.hi - x.lo;
    const double h = (S == 1) ? w * 0.5 : w / (double)(S + 1);
    const double p = __builtin_fma(h, (double)(t + 1), x.lo);
    return mk<double>(p, p);
  }
};

// A box (array or accessor of FP64 intervals) seen as FP32 intervals,
// each coordinate rounded outward.
template <class XS>
struct F32Box {
  const XS &x;
  __device__ __forceinline__ iv<float> operator[](int j) const {
    const iv<double> v = x[j];
    return mk<float>(to_dn<float>(v.lo), to_up<float>(v.hi));
  }
};

// Gradient sink for the mean-value bound: obj_grad's g[j] = G_j folds
// s += G_j * (X_j - c_j) (c the midpoint, j in increasing order) and the
// monotonicity test on the same G_j (1-Ackley/ackley50.py:458-477), so no
// array of d gradient intervals is kept.  Components j >= d are ignored.
template <int R, class XS>
struct MVSink {
  typedef IA<double, R> A;
  const XS &x;
  const double *__restrict__ dlo;
  const double *__restrict__ dhi;
  int d;
  iv<double> acc;
  bool prune;
  struct Ref {
    MVSink *s;
    int j;
    __device__ __forceinline__ void operator=(const iv<double> &g) const { s->put(j, g); }
  };
  __device__ __forceinline__ Ref operator[](int j) { return Ref{this, j}; }
  __device__ __forceinline__ void put(int j, const iv<double> &g) {
    if (j >= d) return;
    const iv<double> xj = x[j];
    const double c = __builtin_fma((xj.hi - xj.lo) * 0.5, 1.0, xj.lo);
    const iv<double> dx = mk<double>(A::RR::sub_dn(xj.lo, c), A::RR::sub_up(xj.hi, c));
    acc = A::add(acc, A::mul(g, dx));
    prune |= (g.lo > 0.0 && xj.lo != dlo[j]) || (g.hi < 0.0 && xj.hi != dhi[j]);
  }
};

// Reductions over groups of G consecutive lanes (G a power of two <= 64)
// that share one box in the term-parallel kernels.  The butterfly leaves
// the same value on every lane of the group (+ is commutative).
template <int R, int G>
__device__ __forceinline__ iv<double> group_sum(iv<double> v) {
  typedef IA<double, R> A;
#pragma unroll
  for (int o = 1; o < G; o <<= 1) {
    const double lo = __shfl_xor(v.lo, o, 64), hi = __shfl_xor(v.hi, o, 64);
    v = A::add(v, mk<double>(lo, hi));
  }
  return v;
}

template <int G>
__device__ __forceinline__ bool group_any(bool p) {
  const unsigned long long m = __ballot(p);
  const int base = (threadIdx.x & 63) & ~(G - 1);
  const unsigned long long gm = (G >= 64) ? ~0ull : ((1ull << G) - 1ull);
  return ((m >> base) & gm) != 0ull;
}

// A box given by its bounds arrays (a parent record, or a row of explicit boxes).
struct RowBox {
  const double *__restrict__ lo;
  const double *__restrict__ hi;
  __device__ __forceinline__ iv<double> operator[](int j) const { return mk<double>(lo[j], hi[j]); }
};

// Term tables of the windowed modules.  A term whose coordinates are
// base, base + 1, ..., base + W - 1 takes, over a child, one value per
// combination of the digits of those coordinates that lie in the split
// window (the others are the parent's).  Entry index = sum of digit_k *
// wid^k, with digit 0 for a coordinate outside the window or outside [0, d).
template <int W>
__device__ __forceinline__ unsigned tab_idx(const ChildBox &x, int base) {
  unsigned idx = 0, m = 1;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const int i = base + k;
    int kk = i - x.win;
    if (kk < 0) kk += x.d;
    if (i >= 0 && i < x.d && kk < x.c) idx += (unsigned)((x.packed >> (kk * x.bpd)) & x.mask) * m;
    m *= x.wid;
  }
  return idx;
}

// The box of one table entry as a coordinate accessor: coordinate j of the
// stencil (base <= j) in the window takes digit (combo / wid^(j - base)) mod
// wid, with ChildBox's arithmetic (so a term over it is bit-identical to
// the same term over every child with those digits); the others are the
// parent's.
struct TabBox {
  const double *__restrict__ plo;
  const double *__restrict__ phi;
  const double *__restrict__ pw;
  int win, d, c, base;
  unsigned wid, combo;
  __device__ __forceinline__ iv<double> operator[](int j) const {
    const double L = plo[j], U = phi[j];
    int k = j - win;
    if (k < 0) k += d;
    if (k < c) {
      unsigned r = combo;
      for (int t = base; t < j; ++t) r /= wid;
      const unsigned dg = r % wid;
      const double w = pw[j];
      const double lo = __builtin_fma(w, (double)dg, L);
      const double hi = (dg + 1u == wid) ? U : __builtin_fma(w, (double)(dg + 1u), L);
      return mk<double>(lo, hi);
    }
    return mk<double>(L, U);
  }
};

__device__ __forceinline__ ChildBox child_box(const double *plo, const double *phi, const double *pw, int win, int d,
                                              int c, unsigned wid, double inv_wid, double half_inv_wid, unsigned pos) {
  ChildBox x;
  x.plo = plo;
  x.phi = phi;
  x.pw = pw;
  x.win = win;
  x.d = d;
  x.c = c;
  x.wid = wid;
  // digits packed at bpd bits each (wid^c < 2^32 keeps c*bpd <= 64);
  // q = floor(r / wid) exactly as fma(r, 1/wid, 1/(2 wid)) for r < 2^32
  x.bpd = 32 - __builtin_clz(wid - 1u);
  x.mask = (1ull << x.bpd) - 1ull;
  unsigned long long packed = 0;
  if ((wid & (wid - 1u)) == 0u) {
    packed = pos;
  } else {
    unsigned r = pos;
    for (int k = 0; k < c; ++k) {
      const unsigned q = (unsigned)__builtin_fma((double)r, inv_wid, half_inv_wid);
      const unsigned dg = r - q * wid;
      packed |= (unsigned long long)dg << (k * x.bpd);
      r = q;
    }
  }
  x.packed = packed;
  return x;
}

}  // namespace gpugo
