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
// exp/sin/cos endpoints come from this file's own FP64 evaluation (tr64
// below: Cody-Waite reduction with its rounding error bounded, Taylor
// polynomials with an a-priori error bound added outward), so they are
// rigorous by construction; FP32 endpoints are FP64 enclosures from shorter
// polynomials (error far below an FP32 ulp) rounded outward.
// IA.py:363-570 instead widens library values by fixed ulp counts.  sin/cos
// find interior extrema from rigorous bounds of x/pi, not from
// IA.py:429-444's nudged floor division.
#pragma once

#ifndef GPUGO_IA_HPP
#define GPUGO_IA_HPP

// exp and sin/cos of the interval type: out of line by default (unrolled code
// lists call them many times); a module that calls them a few times per box
// defines GPUGO_TRANS_INLINE 1 and saves the call's register saves
#if defined(GPUGO_TRANS_INLINE) && GPUGO_TRANS_INLINE
#define GPUGO_TRANS_ATTR __forceinline__
#else
#define GPUGO_TRANS_ATTR __attribute__((noinline))
#endif

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
  __device__ static __forceinline__ I bits(double x) { return __double_as_longlong(x); }
  __device__ static __forceinline__ double from(I b) { return __longlong_as_double(b); }
  __device__ static __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }
  __device__ static __forceinline__ double abs_(double a) { return __builtin_fabs(a); }
  __device__ static __forceinline__ double sqrt_(double a) { return __builtin_sqrt(a); }
  __device__ static __forceinline__ double floor_(double a) { return __builtin_floor(a); }
  __device__ static __forceinline__ double ceil_(double a) { return __builtin_ceil(a); }
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
  __device__ static __forceinline__ I bits(float x) { return __float_as_int(x); }
  __device__ static __forceinline__ float from(I b) { return __int_as_float(b); }
  __device__ static __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
  __device__ static __forceinline__ float abs_(float a) { return __builtin_fabsf(a); }
  __device__ static __forceinline__ float sqrt_(float a) { return __builtin_sqrtf(a); }
  __device__ static __forceinline__ float floor_(float a) { return __builtin_floorf(a); }
  __device__ static __forceinline__ float ceil_(float a) { return __builtin_ceilf(a); }
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
// 10^k for k in [-294, 309]: each decimal literal is converted with correct
// rounding, so 10^0 ... 10^22 are exact (the reference's integer powers) and
// the others are the nearest doubles (as the reference's 10.0 ** k).
#define GPUGO_P10_MIN (-294)
static __device__ const double gpugo_p10[] = {
    1e-294, 1e-293, 1e-292, 1e-291, 1e-290, 1e-289, 1e-288, 1e-287,
    1e-286, 1e-285, 1e-284, 1e-283, 1e-282, 1e-281, 1e-280, 1e-279,
    1e-278, 1e-277, 1e-276, 1e-275, 1e-274, 1e-273, 1e-272, 1e-271,
    1e-270, 1e-269, 1e-268, 1e-267, 1e-266, 1e-265, 1e-264, 1e-263,
    1e-262, 1e-261, 1e-260, 1e-259, 1e-258, 1e-257, 1e-256, 1e-255,
    1e-254, 1e-253, 1e-252, 1e-251, 1e-250, 1e-249, 1e-248, 1e-247,
    1e-246, 1e-245, 1e-244, 1e-243, 1e-242, 1e-241, 1e-240, 1e-239,
    1e-238, 1e-237, 1e-236, 1e-235, 1e-234, 1e-233, 1e-232, 1e-231,
    1e-230, 1e-229, 1e-228, 1e-227, 1e-226, 1e-225, 1e-224, 1e-223,
    1e-222, 1e-221, 1e-220, 1e-219, 1e-218, 1e-217, 1e-216, 1e-215,
    1e-214, 1e-213, 1e-212, 1e-211, 1e-210, 1e-209, 1e-208, 1e-207,
    1e-206, 1e-205, 1e-204, 1e-203, 1e-202, 1e-201, 1e-200, 1e-199,
    1e-198, 1e-197, 1e-196, 1e-195, 1e-194, 1e-193, 1e-192, 1e-191,
    1e-190, 1e-189, 1e-188, 1e-187, 1e-186, 1e-185, 1e-184, 1e-183,
    1e-182, 1e-181, 1e-180, 1e-179, 1e-178, 1e-177, 1e-176, 1e-175,
    1e-174, 1e-173, 1e-172, 1e-171, 1e-170, 1e-169, 1e-168, 1e-167,
    1e-166, 1e-165, 1e-164, 1e-163, 1e-162, 1e-161, 1e-160, 1e-159,
    1e-158, 1e-157, 1e-156, 1e-155, 1e-154, 1e-153, 1e-152, 1e-151,
    1e-150, 1e-149, 1e-148, 1e-147, 1e-146, 1e-145, 1e-144, 1e-143,
    1e-142, 1e-141, 1e-140, 1e-139, 1e-138, 1e-137, 1e-136, 1e-135,
    1e-134, 1e-133, 1e-132, 1e-131, 1e-130, 1e-129, 1e-128, 1e-127,
    1e-126, 1e-125, 1e-124, 1e-123, 1e-122, 1e-121, 1e-120, 1e-119,
    1e-118, 1e-117, 1e-116, 1e-115, 1e-114, 1e-113, 1e-112, 1e-111,
    1e-110, 1e-109, 1e-108, 1e-107, 1e-106, 1e-105, 1e-104, 1e-103,
    1e-102, 1e-101, 1e-100, 1e-99, 1e-98, 1e-97, 1e-96, 1e-95,
    1e-94, 1e-93, 1e-92, 1e-91, 1e-90, 1e-89, 1e-88, 1e-87,
    1e-86, 1e-85, 1e-84, 1e-83, 1e-82, 1e-81, 1e-80, 1e-79,
    1e-78, 1e-77, 1e-76, 1e-75, 1e-74, 1e-73, 1e-72, 1e-71,
    1e-70, 1e-69, 1e-68, 1e-67, 1e-66, 1e-65, 1e-64, 1e-63,
    1e-62, 1e-61, 1e-60, 1e-59, 1e-58, 1e-57, 1e-56, 1e-55,
    1e-54, 1e-53, 1e-52, 1e-51, 1e-50, 1e-49, 1e-48, 1e-47,
    1e-46, 1e-45, 1e-44, 1e-43, 1e-42, 1e-41, 1e-40, 1e-39,
    1e-38, 1e-37, 1e-36, 1e-35, 1e-34, 1e-33, 1e-32, 1e-31,
    1e-30, 1e-29, 1e-28, 1e-27, 1e-26, 1e-25, 1e-24, 1e-23,
    1e-22, 1e-21, 1e-20, 1e-19, 1e-18, 1e-17, 1e-16, 1e-15,
    1e-14, 1e-13, 1e-12, 1e-11, 1e-10, 1e-9, 1e-8, 1e-7,
    1e-6, 1e-5, 1e-4, 1e-3, 1e-2, 1e-1, 1e0, 1e1,
    1e2, 1e3, 1e4, 1e5, 1e6, 1e7, 1e8, 1e9,
    1e10, 1e11, 1e12, 1e13, 1e14, 1e15, 1e16, 1e17,
    1e18, 1e19, 1e20, 1e21, 1e22, 1e23, 1e24, 1e25,
    1e26, 1e27, 1e28, 1e29, 1e30, 1e31, 1e32, 1e33,
    1e34, 1e35, 1e36, 1e37, 1e38, 1e39, 1e40, 1e41,
    1e42, 1e43, 1e44, 1e45, 1e46, 1e47, 1e48, 1e49,
    1e50, 1e51, 1e52, 1e53, 1e54, 1e55, 1e56, 1e57,
    1e58, 1e59, 1e60, 1e61, 1e62, 1e63, 1e64, 1e65,
    1e66, 1e67, 1e68, 1e69, 1e70, 1e71, 1e72, 1e73,
    1e74, 1e75, 1e76, 1e77, 1e78, 1e79, 1e80, 1e81,
    1e82, 1e83, 1e84, 1e85, 1e86, 1e87, 1e88, 1e89,
    1e90, 1e91, 1e92, 1e93, 1e94, 1e95, 1e96, 1e97,
    1e98, 1e99, 1e100, 1e101, 1e102, 1e103, 1e104, 1e105,
    1e106, 1e107, 1e108, 1e109, 1e110, 1e111, 1e112, 1e113,
    1e114, 1e115, 1e116, 1e117, 1e118, 1e119, 1e120, 1e121,
    1e122, 1e123, 1e124, 1e125, 1e126, 1e127, 1e128, 1e129,
    1e130, 1e131, 1e132, 1e133, 1e134, 1e135, 1e136, 1e137,
    1e138, 1e139, 1e140, 1e141, 1e142, 1e143, 1e144, 1e145,
    1e146, 1e147, 1e148, 1e149, 1e150, 1e151, 1e152, 1e153,
    1e154, 1e155, 1e156, 1e157, 1e158, 1e159, 1e160, 1e161,
    1e162, 1e163, 1e164, 1e165, 1e166, 1e167, 1e168, 1e169,
    1e170, 1e171, 1e172, 1e173, 1e174, 1e175, 1e176, 1e177,
    1e178, 1e179, 1e180, 1e181, 1e182, 1e183, 1e184, 1e185,
    1e186, 1e187, 1e188, 1e189, 1e190, 1e191, 1e192, 1e193,
    1e194, 1e195, 1e196, 1e197, 1e198, 1e199, 1e200, 1e201,
    1e202, 1e203, 1e204, 1e205, 1e206, 1e207, 1e208, 1e209,
    1e210, 1e211, 1e212, 1e213, 1e214, 1e215, 1e216, 1e217,
    1e218, 1e219, 1e220, 1e221, 1e222, 1e223, 1e224, 1e225,
    1e226, 1e227, 1e228, 1e229, 1e230, 1e231, 1e232, 1e233,
    1e234, 1e235, 1e236, 1e237, 1e238, 1e239, 1e240, 1e241,
    1e242, 1e243, 1e244, 1e245, 1e246, 1e247, 1e248, 1e249,
    1e250, 1e251, 1e252, 1e253, 1e254, 1e255, 1e256, 1e257,
    1e258, 1e259, 1e260, 1e261, 1e262, 1e263, 1e264, 1e265,
    1e266, 1e267, 1e268, 1e269, 1e270, 1e271, 1e272, 1e273,
    1e274, 1e275, 1e276, 1e277, 1e278, 1e279, 1e280, 1e281,
    1e282, 1e283, 1e284, 1e285, 1e286, 1e287, 1e288, 1e289,
    1e290, 1e291, 1e292, 1e293, 1e294, 1e295, 1e296, 1e297,
    1e298, 1e299, 1e300, 1e301, 1e302, 1e303, 1e304, 1e305,
    1e306, 1e307, 1e308, __builtin_huge_val(),
};

__device__ __forceinline__ double p10(int k) { return gpugo_p10[k - GPUGO_P10_MIN]; }

// The paper's outRound (n = 15): 10^(14 - e) scaling with e = floor(log10|x|).
// e comes from the binary exponent: with |x| in [2^(b-1), 2^b),
// floor((b - 1) log10 2) is floor(log10|x|) or one less, so one comparison
// with the table fixes it.
__device__ __forceinline__ double digits15(double x, bool up) {
  if (!(x == x) || x == 0.0 || __builtin_isinf(x)) return x;
  const double ax = __builtin_fabs(x);
  if (ax < 1e-292) return (x > 0.0) == up ? (up ? 1e-292 : -1e-292) : 0.0;
  int b;
  (void)__builtin_frexp(ax, &b);
  int e = (int)__builtin_floor((double)(b - 1) * 0.30102999566398120);
  if (ax >= p10(e + 1)) ++e;
  const double sc = p10(14 - e);
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

// ---------------------------------------------------------------------------
// Point enclosures of exp, sin and cos in FP64, rigorous by construction (no
// property of a device math library is assumed).
//
// Argument reduction (Cody & Waite, "Software Manual for the Elementary
// Functions", 1980): x = k c + r with c = ln 2 (exp) or pi/2 (sin, cos)
// split as c1 + c2 (+ c3); c1 and c2 carry at most 33 significant bits, so
// k c1 and k c2 are exact for |k| < 2^20, and the last part is bracketed by
// two doubles.  r is formed as an interval with the directed operations of
// the policy D, so it encloses x - k c for any k (outside |k| < 2^20 it is
// only wider).  For |k| < 2^20 (always for exp) the reduction is instead
// three (two) fma steps in round-to-nearest whose rounding errors (at most u
// times each intermediate) and the last part's bracket are bounded and added
// to the error term; beyond, it is an interval in EXACT directed rounding.
// k = rint(x / c) keeps |r| <= ln(2)/2 + 2^-40 (exp,
// |x| < 746) and |r| <= pi/4 + 2^-11 (sin, cos, |x| <= 2^40).
//
// Kernels: Taylor polynomials evaluated by Horner's rule (one fma per term)
// in round-to-nearest at r0 = lo(r).  Their error is bounded a priori and
// added outward:
//   truncation (Lagrange remainder): sin to r^21: |r|^23/23! < 2e-25;
//     cos to r^22: |r|^24/24! < 1e-26; exp to r^13: e^|r| |r|^14/14! < 1e-17;
//   rounding (Higham, Accuracy and Stability of Numerical Algorithms, 2nd
//     ed., 2002, section 5.1: Horner's rule on a degree-n polynomial with
//     one rounding per step returns a value within gamma_n sum_i |a_i| |x|^i
//     of the exact one, gamma_n = n u / (1 - n u), u = 2^-53), with one more
//     u for the rounding of each coefficient and of z = r^2 and the final
//     step: sin: the correction r z Q(z) has |z Q| < 0.12, so the error is
//     below (0.12 (gamma_9 + 4u) + u) |r| < 3u |r|; cos: |z C| < 0.34, error
//     below 0.34 (gamma_10 + 4u) + u < 6u; exp: sum |a_i||r|^i < e^0.35 < 1.42,
//     error below 1.42 (gamma_13 + 2u) < 21u.
//   The bounds added are EPS_SIN |r|, EPS_COS and EPS_EXP, each at least
//   three times these, plus the reduction width w (|sin'|, |cos'| <= 1,
//   (e^r)' < 1.5).
// ---------------------------------------------------------------------------
template <int D> struct tr64 {
  typedef rnd<double, D> RD;
  static constexpr double PIO2_1 = 0x1.921fb544p+0;
  static constexpr double PIO2_2 = 0x1.0b4611a6p-34;
  static constexpr double PIO2_3_LO = 0x1.3198a2e037073p-69;
  static constexpr double PIO2_3_HI = 0x1.3198a2e037074p-69;
  static constexpr double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
  static constexpr double LN2_1 = 0x1.62e42feep-1;
  static constexpr double LN2_2_LO = 0x1.a39ef35793c76p-33;
  static constexpr double LN2_2_HI = 0x1.a39ef35793c77p-33;
  static constexpr double LOG2E = 0x1.71547652b82fep+0;
  static constexpr double EPS_SIN = 1e-15, EPS_COS = 2e-15, EPS_EXP = 8e-15;

  // [a - k b] for an interval a and k b's enclosure [dn(k bl), up(k bh)]
  __device__ static __forceinline__ iv<double> sub_k(iv<double> a, double k, double bl, double bh) {
    const double pl = k >= 0.0 ? RD::mul_dn(k, bl) : RD::mul_dn(k, bh);
    const double ph = k >= 0.0 ? RD::mul_up(k, bh) : RD::mul_up(k, bl);
    return mk<double>(RD::sub_dn(a.lo, ph), RD::sub_up(a.hi, pl));
  }
  __device__ static __forceinline__ double sin_poly(double r) {
    const double z = r * r;
    double q = 1.957294106339126123085e-20;
    q = __builtin_fma(q, z, -8.220635246624329716956e-18);
    q = __builtin_fma(q, z, 2.811457254345520763199e-15);
    q = __builtin_fma(q, z, -7.647163731819816475901e-13);
    q = __builtin_fma(q, z, 1.605904383682161459939e-10);
    q = __builtin_fma(q, z, -2.505210838544171877505e-8);
    q = __builtin_fma(q, z, 2.755731922398589065256e-6);
    q = __builtin_fma(q, z, -1.984126984126984126984e-4);
    q = __builtin_fma(q, z, 8.333333333333333333333e-3);
    q = __builtin_fma(q, z, -1.666666666666666666667e-1);
    return __builtin_fma(r * z, q, r);
  }
  __device__ static __forceinline__ double cos_poly(double r) {
    const double z = r * r;
    double c = -8.896791392450573286749e-22;
    c = __builtin_fma(c, z, 4.110317623312164858478e-19);
    c = __builtin_fma(c, z, -1.561920696858622646222e-16);
    c = __builtin_fma(c, z, 4.779477332387385297438e-14);
    c = __builtin_fma(c, z, -1.147074559772972471385e-11);
    c = __builtin_fma(c, z, 2.087675698786809897921e-9);
    c = __builtin_fma(c, z, -2.755731922398589065256e-7);
    c = __builtin_fma(c, z, 2.480158730158730158730e-5);
    c = __builtin_fma(c, z, -1.388888888888888888889e-3);
    c = __builtin_fma(c, z, 4.166666666666666666667e-2);
    c = __builtin_fma(c, z, -0.5);
    return __builtin_fma(z, c, 1.0);
  }
  __device__ static __forceinline__ double exp_poly(double r) {
    double p = 1.605904383682161459939e-10;
    p = __builtin_fma(p, r, 2.087675698786809897921e-9);
    p = __builtin_fma(p, r, 2.505210838544171877505e-8);
    p = __builtin_fma(p, r, 2.755731922398589065256e-7);
    p = __builtin_fma(p, r, 2.755731922398589065256e-6);
    p = __builtin_fma(p, r, 2.480158730158730158730e-5);
    p = __builtin_fma(p, r, 1.984126984126984126984e-4);
    p = __builtin_fma(p, r, 1.388888888888888888889e-3);
    p = __builtin_fma(p, r, 8.333333333333333333333e-3);
    p = __builtin_fma(p, r, 4.166666666666666666667e-2);
    p = __builtin_fma(p, r, 1.666666666666666666667e-1);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_fma(p, r, 1.0);
  }

  // Kernels of the FP32 interval functions, whose endpoints are rounded
  // outward to FP32 (unit roundoff 2^-24): Taylor to r^9 (sin), r^10 (cos),
  // r^7 (exp).  Truncation: |r|^11/11! < 2.3e-9 |r|, |r|^12/12! < 1.2e-10,
  // e^|r| |r|^8/8! < 7.2e-9; the rounding of these short Horner loops is
  // below 30u.  The bounds added, EPS_*_LP, are at least three times these.
  static constexpr double EPS_SIN_LP = 7e-9, EPS_COS_LP = 4e-10, EPS_EXP_LP = 2.2e-8;
  __device__ static __forceinline__ double sin_poly_lp(double r) {
    const double z = r * r;
    double q = 2.755731922398589065256e-6;
    q = __builtin_fma(q, z, -1.984126984126984126984e-4);
    q = __builtin_fma(q, z, 8.333333333333333333333e-3);
    q = __builtin_fma(q, z, -1.666666666666666666667e-1);
    return __builtin_fma(r * z, q, r);
  }
  __device__ static __forceinline__ double cos_poly_lp(double r) {
    const double z = r * r;
    double c = -2.755731922398589065256e-7;
    c = __builtin_fma(c, z, 2.480158730158730158730e-5);
    c = __builtin_fma(c, z, -1.388888888888888888889e-3);
    c = __builtin_fma(c, z, 4.166666666666666666667e-2);
    c = __builtin_fma(c, z, -0.5);
    return __builtin_fma(z, c, 1.0);
  }
  __device__ static __forceinline__ double exp_poly_lp(double r) {
    double p = 1.984126984126984126984e-4;
    p = __builtin_fma(p, r, 1.388888888888888888889e-3);
    p = __builtin_fma(p, r, 8.333333333333333333333e-3);
    p = __builtin_fma(p, r, 4.166666666666666666667e-2);
    p = __builtin_fma(p, r, 1.666666666666666666667e-1);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_fma(p, r, 1.0);
  }

  typedef rnd<double, ROUND_FAST> RF;  // one outward step after a nearest result
  static constexpr double U = 0x1p-53;

  // sin(x) (shift 0) or cos(x) (shift 1) for |x| <= 2^40; LP: the short
  // kernels (for endpoints that are rounded to FP32)
  template <bool LP = false> __device__ static iv<double> sincos(double x, int shift) {
    const double k = __builtin_rint(x * TWO_OVER_PI);
    const double ak = __builtin_fabs(k);
    double r0, ered;
    if (ak < 0x1p20) {
      // three fma steps.  t1 = x - k c1 is exact: k c1 is a double (33-bit
      // c1, |k| < 2^20) and, for k != 0, k c1 / 2 <= |x| <= 2 k c1 (Sterbenz);
      // |t2 - (t1 - k c2)| <= u |t2| and |r - (t2 - k c3)| <= u |r| (one
      // rounding each), and c3, the low end of its bracket, is within 2^-121
      // of pi/2 - c1 - c2; so |r - (x - k pi/2)| <= u (|t2| + |r|) + |k| 2^-121
      const double c3 = PIO2_3_LO;
      const double t1 = __builtin_fma(-k, PIO2_1, x);
      const double t2 = __builtin_fma(-k, PIO2_2, t1);
      r0 = __builtin_fma(-k, c3, t2);
      ered = __builtin_fma(U, __builtin_fabs(t2) + __builtin_fabs(r0), ak * 0x1p-120);
    } else {
      // the same reduction as an interval in directed rounding
      iv<double> r = sub_k(point<double>(x), k, PIO2_1, PIO2_1);
      r = sub_k(r, k, PIO2_2, PIO2_2);
      r = sub_k(r, k, PIO2_3_LO, PIO2_3_HI);
      r0 = r.lo;
      ered = RD::sub_up(r.hi, r.lo);
    }
    const int q = ((int)((long long)k & 3) + shift) & 3;
    // q = 0: sin r, 1: cos r, 2: -sin r, 3: -cos r; |sin'|, |cos'| <= 1
    const bool use_sin = (q & 1) == 0;
    const double v = use_sin ? (LP ? sin_poly_lp(r0) : sin_poly(r0)) : (LP ? cos_poly_lp(r0) : cos_poly(r0));
    const double es = LP ? EPS_SIN_LP : EPS_SIN, ec = LP ? EPS_COS_LP : EPS_COS;
    const double e = (use_sin ? __builtin_fma(__builtin_fabs(r0), es, 0x1p-1074) : ec) + ered * (1.0 + 4 * U);
    double lo = RF::sub_dn(v, e), hi = RF::add_up(v, e);
    if (q & 2) {
      const double t = lo;
      lo = -hi;
      hi = -t;
    }
    return mk<double>(lo, hi);
  }

  template <bool LP = false> __device__ static iv<double> exp(double x) {
    if (!(x == x)) return mk<double>(x, x);
    if (x > 709.8) return mk<double>(fp<double>::MAXV, __builtin_huge_val());  // e^x > MAXV
    if (x < -745.2) return mk<double>(0.0, 0x1p-1074);                          // 0 < e^x < 2^-1074
    const double k = __builtin_rint(x * LOG2E);  // |k| <= 1076
    // two fma steps: |t - (x - k c1)| <= u |t|, |r - (t - k c2)| <= u |r| with
    // c2 the lower end of its bracket (one ulp, 2^-85, below its upper end)
    const double t = __builtin_fma(-k, LN2_1, x);
    const double r = __builtin_fma(-k, LN2_2_LO, t);
    const double ered = __builtin_fma(U, __builtin_fabs(t) + __builtin_fabs(r), __builtin_fabs(k) * 0x1p-84);
    const double v = LP ? exp_poly_lp(r) : exp_poly(r);
    // (e^r)' < 1.5 on |r| <= ln(2)/2 + 2^-40
    const double e = (LP ? EPS_EXP_LP : EPS_EXP) + 1.5 * ered * (1.0 + 4 * U);
    double lo = RF::sub_dn(v, e), hi = RF::add_up(v, e);
    // times 2^k: exact in the normal range; outward by one step otherwise
    const int ki = (int)k;
    double l2 = __builtin_ldexp(lo, ki), h2 = __builtin_ldexp(hi, ki);
    if (ki < -1020) {
      l2 = __builtin_fmax(pred<double>(l2), 0.0);
      h2 = succ<double>(h2);
    }
    if (__builtin_isinf(l2)) l2 = fp<double>::MAXV;
    return mk<double>(l2, h2);
  }
};

template <typename T, int R> struct IA {
  typedef rnd<T, R> RR;
  typedef iv<T> V;
  // reduction and error-bound arithmetic of the transcendental enclosures:
  // EXACT directed operations in every policy (the reduction's products and
  // differences are exact or nearly so; a FAST step would widen them by an
  // ulp of x)
  typedef tr64<ROUND_EXACT> TR;
  // FP32 intervals take the short kernels: their error is far below an FP32 ulp
  static constexpr bool LP = sizeof(T) == 4;

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

  // out of line: keeps unrolled code lists (and hipRTC compile time) small;
  // inline where a code list calls them a few times (GPUGO_TRANS_INLINE)
  __device__ GPUGO_TRANS_ATTR static V exp(V a) {
    const iv<double> el = TR::template exp<LP>((double)a.lo), eh = TR::template exp<LP>((double)a.hi);
    T l = max_<T>(out_dn<RR>(to_dn<T>(el.lo)), T(0));
    return mk<T>(l, out_up<RR>(to_up<T>(eh.hi)));
  }

  // shift 0: cos (extrema at m*pi), shift 1: sin (extrema at pi/2 + m*pi);
  // m even -> maximum +1, m odd -> minimum -1.
  template <int SHIFT> __device__ GPUGO_TRANS_ATTR static V trig(V a) {
    const T one = T(1);
    if (!(fp<T>::abs_(a.lo) <= fp<T>::TRIG_BIG && fp<T>::abs_(a.hi) <= fp<T>::TRIG_BIG))
      return mk<T>(-one, one);
    const T w = RR::sub_up(a.hi, a.lo);
    if (w <= fp<T>::TRIG_NARROW * max_<T>(one, fp<T>::abs_(a.lo))) {
      // narrow argument (sample points, a few ulps of a large argument):
      // |f(x) - f(a)| <= |x - a| <= w since |sin'|, |cos'| <= 1, so one
      // evaluation and a widening by w suffice
      const iv<double> v = TR::template sincos<LP>((double)a.lo, 1 - SHIFT);
      T l = max_<T>(RR::sub_dn(to_dn<T>(v.lo), w), -one);
      T h = min_<T>(RR::add_up(to_up<T>(v.hi), w), one);
      return mk<T>(l, h);
    }
    if (w >= fp<T>::TWO_PI_LO) return mk<T>(-one, one);
    const iv<double> va = TR::template sincos<LP>((double)a.lo, 1 - SHIFT),
                     vb = TR::template sincos<LP>((double)a.hi, 1 - SHIFT);
    T l = max_<T>(min_<T>(to_dn<T>(va.lo), to_dn<T>(vb.lo)), -one);
    T h = min_<T>(max_<T>(to_up<T>(va.hi), to_up<T>(vb.hi)), one);
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
