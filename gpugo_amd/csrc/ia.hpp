// ia.hpp -- interval arithmetic device library for gfx950 (FP64 and FP32).
//
// Compiled both by hipcc (libgpugo.so) and, as text, into every hipRTC
// objective module, so it uses only compiler builtins and device math.
// Build with -ffp-contract=off: every fused multiply-add below is written
// explicitly and no other a*b+c may be contracted.
//
// Replaces the device functions of the reference interval module
// ("Updated Interval Arithmetic Operations/IA.py":57-570).  CDNA has no
// per-instruction rounding mode, so directed rounding is built from
// round-to-nearest with one of three policies:
//
//   EXACT  error-free transformations: TwoSum (Knuth) for +/-, the fma
//          residual for * / and sqrt.  The sign of the exact error decides
//          whether the nearest result is already the round-down / round-up
//          result or is one ulp off; the ulp step is done on the bit pattern.
//          For finite operands this equals IEEE directed rounding bit for
//          bit, except for nonzero products below 2^-968 in magnitude whose
//          residual may underflow: those are stepped outward unconditionally
//          (one ulp wider, still rigorous).
//   FAST   round-to-nearest followed by one outward step
//          x +/- fma(|x|, u(1+2u), eta), u the unit roundoff and eta the
//          smallest subnormal, which is >= succ(x) / <= pred(x) for every x
//          (Rump, Zimmermann, Boldo & Melquiond, "Computing predecessor and
//          successor in rounding to nearest", BIT 49 (2009)).  At most one
//          ulp wider than EXACT per operation.
//
//   DIGITS15  (FP64 only) the semantics of the IA behind the paper's results
//          (1-Ackley/IAAckley.py:106-141, outRound with n = 15): every
//          result is widened to the enclosing decimal of 15 significant
//          digits.  Each bound is the EXACT bound, lowered (raised) further
//          to floor (ceil) of the nearest result at 15 significant digits,
//          so it is never narrower than EXACT and stays rigorous where the
//          decimal scaling itself rounds.  Magnitudes below 1e-292 widen to
//          0 (lower bound of a positive value) or to -/+1e-292.
//
// exp/sin/cos use the device library's round-to-nearest functions and are
// widened outward by W_TRANS ulps on the bit pattern, following the shape of
// IA.py:363-570 (nearest value, then integer steps); the bound of the
// device functions is checked against mpmath in the tests.  sin/cos find
// interior extrema from rigorous bounds of x/pi, not from IA.py:429-444's
// nudged floor division.
#pragma once

#ifndef GPUGO_IA_HPP
#define GPUGO_IA_HPP

namespace gpugo {

enum { ROUND_EXACT = 0, ROUND_FAST = 1, ROUND_DIGITS15 = 2 };

template <typename T> struct fp;

template <> struct fp<double> {
  typedef long long I;
  static constexpr double PHI = 0x1.0000000000001p-53;  // u(1+2u), u = 2^-53
  static constexpr double ETA = 0x1p-1074;
  static constexpr double TINY = 0x1p-968;
  static constexpr double MAXV = 0x1.fffffffffffffp+1023;
  static constexpr I INF_BITS = 0x7ff0000000000000LL;
  static constexpr double INV_PI_LO = 0x1.45f306dc9c882p-2;
  static constexpr double INV_PI_HI = 0x1.45f306dc9c883p-2;
  static constexpr double TWO_PI_LO = 0x1.921fb54442d18p+2;
  static constexpr double TRIG_BIG = 0x1p40;
  static constexpr double TRIG_NARROW = 0x1p-40;
  static constexpr int W_TRANS = 2;
  __device__ static __forceinline__ I bits(double x) { return __double_as_longlong(x); }
  __device__ static __forceinline__ double from(I b) { return __longlong_as_double(b); }
  __device__ static __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }
  __device__ static __forceinline__ double abs_(double a) { return __builtin_fabs(a); }
  __device__ static __forceinline__ double sqrt_(double a) { return __builtin_sqrt(a); }
  __device__ static __forceinline__ double floor_(double a) { return __builtin_floor(a); }
  __device__ static __forceinline__ double ceil_(double a) { return __builtin_ceil(a); }
  __device__ static __forceinline__ double exp_(double a) { return ::exp(a); }
  __device__ static __forceinline__ double sin_(double a) { return ::sin(a); }
  __device__ static __forceinline__ double cos_(double a) { return ::cos(a); }
};

template <> struct fp<float> {
  typedef int I;
  static constexpr float PHI = 0x1.000002p-24f;   // u(1+2u), u = 2^-24
  static constexpr float ETA = 0x1p-126f;         // smallest normal: safe if denormals flush
  static constexpr float TINY = 0x1p-100f;
  static constexpr float MAXV = 0x1.fffffep+127f;
  static constexpr I INF_BITS = 0x7f800000;
  static constexpr float INV_PI_LO = 0x1.45f306p-2f;
  static constexpr float INV_PI_HI = 0x1.45f308p-2f;
  static constexpr float TWO_PI_LO = 0x1.921fb4p+2f;
  static constexpr float TRIG_BIG = 0x1p20f;
  static constexpr float TRIG_NARROW = 0x1p-16f;
  static constexpr int W_TRANS = 3;
  __device__ static __forceinline__ I bits(float x) { return __float_as_int(x); }
  __device__ static __forceinline__ float from(I b) { return __int_as_float(b); }
  __device__ static __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
  __device__ static __forceinline__ float abs_(float a) { return __builtin_fabsf(a); }
  __device__ static __forceinline__ float sqrt_(float a) { return __builtin_sqrtf(a); }
  __device__ static __forceinline__ float floor_(float a) { return __builtin_floorf(a); }
  __device__ static __forceinline__ float ceil_(float a) { return __builtin_ceilf(a); }
  __device__ static __forceinline__ float exp_(float a) { return ::expf(a); }
  __device__ static __forceinline__ float sin_(float a) { return ::sinf(a); }
  __device__ static __forceinline__ float cos_(float a) { return ::cosf(a); }
};

// exact neighbours on the bit pattern; +inf and NaN stay, -0 treated as +0
template <typename T> __device__ __forceinline__ T succ(T x) {
  typedef typename fp<T>::I I;
  I b = fp<T>::bits(x + T(0));
  I n = (b >= 0) ? ((b < fp<T>::INF_BITS) ? b + 1 : b) : b - 1;
  return fp<T>::from(n);
}
template <typename T> __device__ __forceinline__ T pred(T x) { return -succ<T>(-x); }

template <typename T> __device__ __forceinline__ T min_(T a, T b) { return a < b ? a : b; }
template <typename T> __device__ __forceinline__ T max_(T a, T b) { return a > b ? a : b; }

// ---------------------------------------------------------------------------
// Directed rounding primitives: dn = toward -inf, up = toward +inf.
// Down results are clamped to <= MAXV and up results to >= -MAXV so a finite
// exact result that overflows is reported as the largest finite value, as
// IEEE directed rounding does.
// ---------------------------------------------------------------------------
template <typename T, int R> struct rnd;

template <typename T> struct rnd<T, ROUND_EXACT> {
  __device__ static __forceinline__ T two_sum_err(T a, T b, T s) {
    T bb = s - a;
    return (a - (s - bb)) + (b - bb);
  }
  __device__ static __forceinline__ T add_dn(T a, T b) {
    T s = a + b, e = two_sum_err(a, b, s);
    return min_<T>(e < T(0) ? pred<T>(s) : s, fp<T>::MAXV);
  }
  __device__ static __forceinline__ T add_up(T a, T b) {
    T s = a + b, e = two_sum_err(a, b, s);
    return max_<T>(e > T(0) ? succ<T>(s) : s, -fp<T>::MAXV);
  }
  __device__ static __forceinline__ T sub_dn(T a, T b) { return add_dn(a, -b); }
  __device__ static __forceinline__ T sub_up(T a, T b) { return add_up(a, -b); }
  __device__ static __forceinline__ bool tiny_inexact(T a, T b, T p) {
    return fp<T>::abs_(p) < fp<T>::TINY && a != T(0) && b != T(0);
  }
  __device__ static __forceinline__ T mul_dn(T a, T b) {
    T p = a * b, e = fp<T>::fma_(a, b, -p);
    bool step = (e < T(0)) || tiny_inexact(a, b, p);
    return min_<T>(step ? pred<T>(p) : p, fp<T>::MAXV);
  }
  __device__ static __forceinline__ T mul_up(T a, T b) {
    T p = a * b, e = fp<T>::fma_(a, b, -p);
    bool step = (e > T(0)) || tiny_inexact(a, b, p);
    return max_<T>(step ? succ<T>(p) : p, -fp<T>::MAXV);
  }
  // a/b - q = r/b with r = a - q*b exact
  __device__ static __forceinline__ T div_dn(T a, T b) {
    T q = a / b, r = fp<T>::fma_(-q, b, a);
    bool below = (b > T(0)) ? (r < T(0)) : (r > T(0));
    bool step = below || (fp<T>::abs_(q) < fp<T>::TINY && a != T(0));
    return min_<T>(step ? pred<T>(q) : q, fp<T>::MAXV);
  }
  __device__ static __forceinline__ T div_up(T a, T b) {
    T q = a / b, r = fp<T>::fma_(-q, b, a);
    bool above = (b > T(0)) ? (r > T(0)) : (r < T(0));
    bool step = above || (fp<T>::abs_(q) < fp<T>::TINY && a != T(0));
    return max_<T>(step ? succ<T>(q) : q, -fp<T>::MAXV);
  }
  __device__ static __forceinline__ T sqrt_dn(T a) {
    T r = fp<T>::sqrt_(a), e = fp<T>::fma_(-r, r, a);
    return e < T(0) ? pred<T>(r) : r;
  }
  __device__ static __forceinline__ T sqrt_up(T a) {
    T r = fp<T>::sqrt_(a), e = fp<T>::fma_(-r, r, a);
    return e > T(0) ? succ<T>(r) : r;
  }
  // outward step of a nearest result that carries no error information
  __device__ static __forceinline__ T step_dn(T x) { return pred<T>(x); }
  __device__ static __forceinline__ T step_up(T x) { return succ<T>(x); }
};

template <typename T> struct rnd<T, ROUND_FAST> {
  __device__ static __forceinline__ T step_dn(T x) {
    T c = min_<T>(x, fp<T>::MAXV);
    return c - fp<T>::fma_(fp<T>::abs_(c), fp<T>::PHI, fp<T>::ETA);
  }
  __device__ static __forceinline__ T step_up(T x) {
    T c = max_<T>(x, -fp<T>::MAXV);
    return c + fp<T>::fma_(fp<T>::abs_(c), fp<T>::PHI, fp<T>::ETA);
  }
  __device__ static __forceinline__ T add_dn(T a, T b) { return step_dn(a + b); }
  __device__ static __forceinline__ T add_up(T a, T b) { return step_up(a + b); }
  __device__ static __forceinline__ T sub_dn(T a, T b) { return step_dn(a - b); }
  __device__ static __forceinline__ T sub_up(T a, T b) { return step_up(a - b); }
  __device__ static __forceinline__ T mul_dn(T a, T b) { return step_dn(a * b); }
  __device__ static __forceinline__ T mul_up(T a, T b) { return step_up(a * b); }
  __device__ static __forceinline__ T div_dn(T a, T b) { return step_dn(a / b); }
  __device__ static __forceinline__ T div_up(T a, T b) { return step_up(a / b); }
  __device__ static __forceinline__ T sqrt_dn(T a) {
    T r = fp<T>::sqrt_(a), e = fp<T>::fma_(-r, r, a);
    return e < T(0) ? pred<T>(r) : r;
  }
  __device__ static __forceinline__ T sqrt_up(T a) {
    T r = fp<T>::sqrt_(a), e = fp<T>::fma_(-r, r, a);
    return e > T(0) ? succ<T>(r) : r;
  }
};

// ---------------------------------------------------------------------------
// Intervals.
// ---------------------------------------------------------------------------
// 15-significant-digit outward rounding of a double (DIGITS15): the largest
// (smallest) decimal of 15 significant digits <= x (>= x), computed as
// floor(x * 10^(14-e)) / 10^(14-e), e = floor(log10 |x|), the shape of
// 1-Ackley/IAAckley.py:106-141; callers take the min (max) with a rigorous
// bound, which absorbs the rounding of this scaling.
__device__ __forceinline__ double digits15(double x, bool up) {
  if (!(x == x) || x == 0.0 || __builtin_isinf(x)) return x;
  const double ax = __builtin_fabs(x);
  if (ax < 1e-292) return (x > 0.0) == up ? (up ? 1e-292 : -1e-292) : 0.0;
  const int e = (int)__builtin_floor(::log10(ax));
  // 10^k exactly for 0 <= k <= 22 (the reference's integer power), else pow
  const int k = 14 - e;
  double sc = 1.0;
  if (k >= 0 && k <= 22) {
    for (int i = 0; i < k; ++i) sc *= 10.0;
  } else {
    sc = ::pow(10.0, (double)k);
  }
  const double q = ax * sc;
  // magnitude rounded away from 0 when the bound moves away from 0
  const bool away = (x > 0.0) == up;
  const double m = (away ? __builtin_ceil(q) : __builtin_floor(q)) / sc;
  return x > 0.0 ? m : -m;
}

template <typename T> struct rnd<T, ROUND_DIGITS15> {
  typedef rnd<T, ROUND_EXACT> X;
  __device__ static __forceinline__ T dn(T x, T r) { return min_<T>(digits15(x, false), r); }
  __device__ static __forceinline__ T up(T x, T r) { return max_<T>(digits15(x, true), r); }
  __device__ static __forceinline__ T add_dn(T a, T b) { return dn(a + b, X::add_dn(a, b)); }
  __device__ static __forceinline__ T add_up(T a, T b) { return up(a + b, X::add_up(a, b)); }
  __device__ static __forceinline__ T sub_dn(T a, T b) { return dn(a - b, X::sub_dn(a, b)); }
  __device__ static __forceinline__ T sub_up(T a, T b) { return up(a - b, X::sub_up(a, b)); }
  __device__ static __forceinline__ T mul_dn(T a, T b) { return dn(a * b, X::mul_dn(a, b)); }
  __device__ static __forceinline__ T mul_up(T a, T b) { return up(a * b, X::mul_up(a, b)); }
  __device__ static __forceinline__ T div_dn(T a, T b) { return dn(a / b, X::div_dn(a, b)); }
  __device__ static __forceinline__ T div_up(T a, T b) { return up(a / b, X::div_up(a, b)); }
  __device__ static __forceinline__ T sqrt_dn(T a) { return dn(__builtin_sqrt(a), X::sqrt_dn(a)); }
  __device__ static __forceinline__ T sqrt_up(T a) { return up(__builtin_sqrt(a), X::sqrt_up(a)); }
  // outward rounding of an already rigorous bound (transcendental results)
  __device__ static __forceinline__ T out_dn(T r) { return dn(r, r); }
  __device__ static __forceinline__ T out_up(T r) { return up(r, r); }
};

template <typename T> struct iv { T lo, hi; };

// rigorous bounds of transcendental functions pass through the policy's
// out_dn/out_up where it has them (DIGITS15), unchanged otherwise
template <class RR, class = void> struct has_out { static constexpr bool value = false; };
template <class RR> struct has_out<RR, decltype((void)RR::out_dn(0.0), void())> {
  static constexpr bool value = true;
};
template <class RR, typename T> __device__ __forceinline__ T out_dn(T x) {
  if constexpr (has_out<RR>::value) return RR::out_dn(x);
  else return x;
}
template <class RR, typename T> __device__ __forceinline__ T out_up(T x) {
  if constexpr (has_out<RR>::value) return RR::out_up(x);
  else return x;
}

template <typename T> __device__ __forceinline__ iv<T> mk(T lo, T hi) { iv<T> r; r.lo = lo; r.hi = hi; return r; }
template <typename T> __device__ __forceinline__ iv<T> point(T x) { return mk<T>(x, x); }
template <typename T> __device__ __forceinline__ iv<T> entire() {
  return mk<T>(-__builtin_huge_val(), __builtin_huge_val());
}

// outward conversion of a double to T (exact for double)
template <typename T> __device__ __forceinline__ T to_dn(double x);
template <typename T> __device__ __forceinline__ T to_up(double x);
template <> __device__ __forceinline__ double to_dn<double>(double x) { return x; }
template <> __device__ __forceinline__ double to_up<double>(double x) { return x; }
template <> __device__ __forceinline__ float to_dn<float>(double x) {
  float f = (float)x;
  return ((double)f > x) ? pred<float>(f) : f;
}
template <> __device__ __forceinline__ float to_up<float>(double x) {
  float f = (float)x;
  return ((double)f < x) ? succ<float>(f) : f;
}

template <typename T, int R> struct IA {
  typedef rnd<T, R> RR;
  typedef iv<T> V;

  __device__ static __forceinline__ V add(V a, V b) { return mk<T>(RR::add_dn(a.lo, b.lo), RR::add_up(a.hi, b.hi)); }
  __device__ static __forceinline__ V sub(V a, V b) { return mk<T>(RR::sub_dn(a.lo, b.hi), RR::sub_up(a.hi, b.lo)); }
  __device__ static __forceinline__ V neg(V a) { return mk<T>(-a.hi, -a.lo); }
  __device__ static __forceinline__ V addc(V a, T s) { return mk<T>(RR::add_dn(a.lo, s), RR::add_up(a.hi, s)); }
  __device__ static __forceinline__ V rsubc(T s, V a) { return mk<T>(RR::sub_dn(s, a.hi), RR::sub_up(s, a.lo)); }

  // sign-aware scalar product (IA.py:306-333 assumes s > 0)
  __device__ static __forceinline__ V scale(V a, T s) {
    T x = s >= T(0) ? a.lo : a.hi;
    T y = s >= T(0) ? a.hi : a.lo;
    V r = mk<T>(RR::mul_dn(x, s), RR::mul_up(y, s));
    return s == T(0) ? mk<T>(T(0), T(0)) : r;
  }
  // known positive / negative scalar (emitter uses these for literals)
  __device__ static __forceinline__ V scale_pos(V a, T s) { return mk<T>(RR::mul_dn(a.lo, s), RR::mul_up(a.hi, s)); }
  __device__ static __forceinline__ V scale_neg(V a, T s) { return mk<T>(RR::mul_dn(a.hi, s), RR::mul_up(a.lo, s)); }

  __device__ static __forceinline__ T zmul_dn(T a, T b) { return (a == T(0) || b == T(0)) ? T(0) : RR::mul_dn(a, b); }
  __device__ static __forceinline__ T zmul_up(T a, T b) { return (a == T(0) || b == T(0)) ? T(0) : RR::mul_up(a, b); }

  // product with an interval known to be >= 0 (a coefficient declared
  // non-negative and rounded outward to T); 0 * inf counts as 0
  __device__ static __forceinline__ V mulp(V a, V p) {
    return mk<T>(zmul_dn(a.lo, a.lo >= T(0) ? p.lo : p.hi), zmul_up(a.hi, a.hi >= T(0) ? p.hi : p.lo));
  }

  __device__ static __forceinline__ V mul(V a, V b) {
    T l = min_<T>(min_<T>(zmul_dn(a.lo, b.lo), zmul_dn(a.lo, b.hi)), min_<T>(zmul_dn(a.hi, b.lo), zmul_dn(a.hi, b.hi)));
    T h = max_<T>(max_<T>(zmul_up(a.lo, b.lo), zmul_up(a.lo, b.hi)), max_<T>(zmul_up(a.hi, b.lo), zmul_up(a.hi, b.hi)));
    return mk<T>(l, h);
  }

  // extended division (textbook; IA.py:154-207 special cases by the mathematics)
  __device__ static V div(V a, V b) {
    const T inf = __builtin_huge_val();
    if (b.lo > T(0) || b.hi < T(0)) {
      T l = min_<T>(min_<T>(RR::div_dn(a.lo, b.lo), RR::div_dn(a.lo, b.hi)), min_<T>(RR::div_dn(a.hi, b.lo), RR::div_dn(a.hi, b.hi)));
      T h = max_<T>(max_<T>(RR::div_up(a.lo, b.lo), RR::div_up(a.lo, b.hi)), max_<T>(RR::div_up(a.hi, b.lo), RR::div_up(a.hi, b.hi)));
      return mk<T>(l, h);
    }
    if (b.lo == T(0) && b.hi > T(0)) {
      if (a.lo >= T(0)) return mk<T>(RR::div_dn(a.lo, b.hi), inf);
      if (a.hi <= T(0)) return mk<T>(-inf, RR::div_up(a.hi, b.hi));
      return entire<T>();
    }
    if (b.hi == T(0) && b.lo < T(0)) {
      if (a.lo >= T(0)) return mk<T>(-inf, RR::div_up(a.lo, b.lo));
      if (a.hi <= T(0)) return mk<T>(RR::div_dn(a.hi, b.lo), inf);
      return entire<T>();
    }
    return entire<T>();
  }

  __device__ static __forceinline__ V divc(V a, T s) {
    T x = s > T(0) ? a.lo : a.hi;
    T y = s > T(0) ? a.hi : a.lo;
    return mk<T>(RR::div_dn(x, s), RR::div_up(y, s));
  }
  __device__ static __forceinline__ V divc_pos(V a, T s) { return mk<T>(RR::div_dn(a.lo, s), RR::div_up(a.hi, s)); }

  __device__ static __forceinline__ V sqr(V a) {
    T al = a.lo, ah = a.hi;
    bool pos = al >= T(0), negv = ah <= T(0);
    T m = pos ? al : (negv ? -ah : T(0));          // mignitude
    T M = max_<T>(-al, ah);                         // magnitude
    T l = (pos || negv) ? RR::mul_dn(m, m) : T(0);
    return mk<T>(l, RR::mul_up(M, M));
  }

  __device__ static __forceinline__ T pow_dn(T x, int n) {
    int top = 31 - __builtin_clz((unsigned)n);
    T r = x;
    for (int b = top - 1; b >= 0; --b) { r = RR::mul_dn(r, r); if ((n >> b) & 1) r = RR::mul_dn(r, x); }
    return r;
  }
  __device__ static __forceinline__ T pow_up(T x, int n) {
    int top = 31 - __builtin_clz((unsigned)n);
    T r = x;
    for (int b = top - 1; b >= 0; --b) { r = RR::mul_up(r, r); if ((n >> b) & 1) r = RR::mul_up(r, x); }
    return r;
  }
  __device__ static V pown(V a, int n) {
    if (n <= 0) return mk<T>(T(1), T(1));
    if (n == 1) return a;
    if ((n & 1) == 0) {
      if (a.lo >= T(0)) return mk<T>(pow_dn(a.lo, n), pow_up(a.hi, n));
      if (a.hi <= T(0)) return mk<T>(pow_dn(-a.hi, n), pow_up(-a.lo, n));
      return mk<T>(T(0), pow_up(max_<T>(-a.lo, a.hi), n));
    }
    T l = a.lo >= T(0) ? pow_dn(a.lo, n) : -pow_up(-a.lo, n);
    T h = a.hi >= T(0) ? pow_up(a.hi, n) : -pow_dn(-a.hi, n);
    return mk<T>(l, h);
  }

  __device__ static __forceinline__ V sqrt(V a) {
    const T nan = __builtin_nan("");
    if (a.hi < T(0)) return mk<T>(nan, nan);
    T l = a.lo <= T(0) ? T(0) : RR::sqrt_dn(a.lo);
    return mk<T>(l, RR::sqrt_up(a.hi));
  }

  __device__ static __forceinline__ T widen_dn(T x) {
#pragma unroll
    for (int i = 0; i < fp<T>::W_TRANS; ++i) x = pred<T>(x);
    return x;
  }
  __device__ static __forceinline__ T widen_up(T x) {
#pragma unroll
    for (int i = 0; i < fp<T>::W_TRANS; ++i) x = succ<T>(x);
    return x;
  }

  // out of line: keeps unrolled code lists (and hipRTC compile time) small
  __device__ __attribute__((noinline)) static V exp(V a) {
    T l = max_<T>(out_dn<RR>(widen_dn(fp<T>::exp_(a.lo))), T(0));
    return mk<T>(l, out_up<RR>(widen_up(fp<T>::exp_(a.hi))));
  }

  // shift 0: cos (extrema at m*pi), shift 1: sin (extrema at pi/2 + m*pi);
  // m even -> maximum +1, m odd -> minimum -1.
  template <int SHIFT> __device__ __attribute__((noinline)) static V trig(V a) {
    const T one = T(1);
    if (!(fp<T>::abs_(a.lo) <= fp<T>::TRIG_BIG && fp<T>::abs_(a.hi) <= fp<T>::TRIG_BIG))
      return mk<T>(-one, one);
    const T w = RR::sub_up(a.hi, a.lo);
    if (w <= fp<T>::TRIG_NARROW) {
      // narrow argument (sample points): |f(x) - f(a)| <= |x - a| <= w since
      // |sin'|, |cos'| <= 1, so one evaluation and a widening by w suffice
      T v = SHIFT ? fp<T>::sin_(a.lo) : fp<T>::cos_(a.lo);
      T l = max_<T>(RR::sub_dn(widen_dn(v), w), -one);
      T h = min_<T>(RR::add_up(widen_up(v), w), one);
      return mk<T>(l, h);
    }
    if (w >= fp<T>::TWO_PI_LO) return mk<T>(-one, one);
    T va = SHIFT ? fp<T>::sin_(a.lo) : fp<T>::cos_(a.lo);
    T vb = SHIFT ? fp<T>::sin_(a.hi) : fp<T>::cos_(a.hi);
    T l = max_<T>(widen_dn(min_<T>(va, vb)), -one);
    T h = min_<T>(widen_up(max_<T>(va, vb)), one);
    T ta = a.lo >= T(0) ? RR::mul_dn(a.lo, fp<T>::INV_PI_LO) : RR::mul_dn(a.lo, fp<T>::INV_PI_HI);
    T tb = a.hi >= T(0) ? RR::mul_up(a.hi, fp<T>::INV_PI_HI) : RR::mul_up(a.hi, fp<T>::INV_PI_LO);
    if (SHIFT) { ta = RR::sub_dn(ta, T(0.5)); tb = RR::sub_up(tb, T(0.5)); }
    T m0 = fp<T>::ceil_(ta), m1 = fp<T>::floor_(tb);
    if (m1 >= m0) {
      if (m1 > m0) return mk<T>(-one, one);           // both parities present
      long long mi = (long long)m0;
      if ((mi & 1) == 0) h = one; else l = -one;
    }
    return mk<T>(l, h);
  }
  __device__ static __forceinline__ V cos(V a) {
    const V r = trig<0>(a);
    return mk<T>(out_dn<RR>(r.lo), out_up<RR>(r.hi));
  }
  __device__ static __forceinline__ V sin(V a) {
    const V r = trig<1>(a);
    return mk<T>(out_dn<RR>(r.lo), out_up<RR>(r.hi));
  }
};

}  // namespace gpugo

#endif  // GPUGO_IA_HPP
