// Elementary functions used by the solver and by the registered systems, compiled
// from this one source for the gfx950 device pass and for the host (g++ / clang).
//
// Why the project owns these: the reference calls the CUDA device libm (`pow` in
// SourceCodes/SingleSystem_PerThread_ExplicitRungeKutta_ErrorControllers.cuh:51-53,
// `cos` in the user right-hand sides, e.g.
// Tutorials_SingleSystem_PerThread/T1_Reference/Reference_SystemDefinition.cuh:12).
// Bit-identical accepted-step and event counts between the GPU and the host oracle
// require that both sides evaluate the same function bit for bit, so every routine
// here uses only IEEE add/sub/mul/div, explicit fused multiply-add and round-to-
// nearest-even, all of which are correctly rounded on gfx950 and on x86-64.
// Compile every user of this header with -ffp-contract=off.
//
// Methods (all published, coefficients fitted by gen_math_coeffs.py beside this file):
//  * cos/sin: Cody-Waite reduction by pi/2 with a two-part constant applied with
//    fma (exact first part), then near-minimax polynomials on [-pi/4, pi/4]
//    in z = r^2 (Chebyshev fits; error below 2e-17 of the polynomial part).
//    The reduced argument's absolute error is below n * 1.6e-33 plus half an
//    ulp, and the rounded quotient keeps |r| within pi/4 + 0.1 while
//    |x| < 2^50, so results stay within a few 1e-16 of the true value there;
//    from 2^50 on the spacing of doubles is at least 1/4 radian (the phase of
//    the argument itself is uncertain) and cos/sin return NaN, so the solver
//    reports the instance as non-finite instead of integrating noise.
//  * exp: reduction by ln2 (two-part, fma), polynomial e^r = 1 + r + r^2 E(r),
//    scaling by 2^n through the exponent field.
//  * log: x = 2^k m, m in [sqrt(2)/2, sqrt(2)), f = m-1, s = f/(2+f),
//    log(1+f) = f - (f^2/2 - s (f^2/2 + R(s^2))), the 2 atanh(s) form of
//    Tang / Sun fdlibm, with R fitted here.
//  * pow(x, y) = exp(y log x) with log x carried in double-double, a few ulp
//    over the controller's x^(1/5) and x^(1/4) for x up to 1e30.
// Accuracy against the platform libm is asserted in tests/test_math.py.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MPGOS_HD __host__ __device__
#else
#define MPGOS_HD
#endif

#define MPGOS_INLINE inline __attribute__((always_inline))

namespace mpgos {
namespace math {

MPGOS_HD MPGOS_INLINE uint64_t bits(double x) { return __builtin_bit_cast(uint64_t, x); }
MPGOS_HD MPGOS_INLINE double from_bits(uint64_t u) { return __builtin_bit_cast(double, u); }
MPGOS_HD MPGOS_INLINE uint32_t bits(float x) { return __builtin_bit_cast(uint32_t, x); }
MPGOS_HD MPGOS_INLINE float from_bits32(uint32_t u) { return __builtin_bit_cast(float, u); }

MPGOS_HD MPGOS_INLINE double fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
MPGOS_HD MPGOS_INLINE float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
MPGOS_HD MPGOS_INLINE double rint(double x) { return __builtin_rint(x); }
MPGOS_HD MPGOS_INLINE float rint(float x) { return __builtin_rintf(x); }
MPGOS_HD MPGOS_INLINE double fabs(double x) { return __builtin_fabs(x); }
MPGOS_HD MPGOS_INLINE float fabs(float x) { return __builtin_fabsf(x); }
MPGOS_HD MPGOS_INLINE bool isfinite(double x) { return (bits(x) & 0x7ff0000000000000ull) != 0x7ff0000000000000ull; }
MPGOS_HD MPGOS_INLINE bool isfinite(float x) { return (bits(x) & 0x7f800000u) != 0x7f800000u; }
MPGOS_HD MPGOS_INLINE bool isnan(double x) { return x != x; }

// IEEE square root: correctly rounded on the host and on gfx950 (the device
// expansion of the square root refines the hardware estimate to the correctly
// rounded result; tests compare both bit for bit).
MPGOS_HD MPGOS_INLINE double sqrt(double x) { return __builtin_sqrt(x); }

// 2^n for integer n in [-1022, 1023] built from the exponent field.
MPGOS_HD MPGOS_INLINE double pow2i(int n) { return from_bits((uint64_t)(n + 1023) << 52); }

// ---------------------------------------------------------------- cos / sin
namespace detail {
// pi/2 = PIO2_HI + PIO2_LO (+ ~1e-33); PIO2_HI is the double nearest pi/2.
constexpr double PIO2_HI = 0x1.921fb54442d18p+0;
constexpr double PIO2_LO = 0x1.1a62633145c07p-54;
constexpr double TWO_OVER_PI = 0x1.45f306dc9c883p-1;

// cos(r) = 1 - z/2 + z^2 C(z), z = r^2, |r| <= pi/4 (+ rounding slack)
MPGOS_HD MPGOS_INLINE double cos_kernel(double r)
{
    const double z = r * r;
    double c = -0x1.907da367a37cbp-37;
    c = fma(c, z, 0x1.1eeb68e93b64cp-29);
    c = fma(c, z, -0x1.27e4fa17da09ep-22);
    c = fma(c, z, 0x1.a01a019f4eb01p-16);
    c = fma(c, z, -0x1.6c16c16c16967p-10);
    c = fma(c, z, 0x1.5555555555555p-5);
    // 1 - z/2 + z^2 c, with the large term kept separate until the end
    const double hz = 0.5 * z;
    const double w = 1.0 - hz;                    // exact for z <= 0.62 up to one rounding
    return w + (((1.0 - w) - hz) + z * z * c);    // recover the rounding error of w
}

// sin(r) = r + r z S(z)
MPGOS_HD MPGOS_INLINE double sin_kernel(double r)
{
    const double z = r * r;
    double s = 0x1.5e0b19f8b1451p-33;
    s = fma(s, z, -0x1.ae600b02b6262p-26);
    s = fma(s, z, 0x1.71de37968a100p-19);
    s = fma(s, z, -0x1.a01a019e83aaep-13);
    s = fma(s, z, 0x1.1111111110bb2p-7);
    s = fma(s, z, -0x1.5555555555555p-3);
    return fma(r * z, s, r);
}

// r = x - n pi/2 and the quadrant n mod 4, for |x| < 2^50.  n = rint(x 2/pi):
// adding 1.5 * 2^52 rounds y = x 2/pi to an integer (to nearest, ties to even,
// as rint) for |y| < 2^51 and leaves n mod 4 in the two lowest mantissa bits
// (the mantissa then holds 2^51 + n).
MPGOS_HD MPGOS_INLINE double reduce_pio2(double x, int& quadrant)
{
    constexpr double SHIFT = 0x1.8p52;
    const double y = x * TWO_OVER_PI;
    const double sh = y + SHIFT;
    const double n = sh - SHIFT;      // rint(y), exactly
    double r = fma(-n, PIO2_HI, x);   // exact: granularity of n*PIO2_HI and x is >= 2^-52 and |r| < 1
    r = fma(-n, PIO2_LO, r);
    quadrant = (int)(bits(sh) & 3u);
    return r;
}
} // namespace detail

// Arguments outside |x| < 2^50 (including +-inf and NaN) are replaced by a
// NaN before the reduction, which then flows through to the result, instead
// of returning early: the function is one block of arithmetic, which a
// compiler can evaluate once for equal arguments (RK4's two mid-step stages
// share their time, so a right-hand side cos(T) is computed once for both).
// The result's sign is flipped by the quadrant's bit in the sign position.
MPGOS_HD MPGOS_INLINE double cos(double x)
{
    const bool ok = fabs(x) < 0x1p50;   // false for NaN
    int q;
    const double r = detail::reduce_pio2(ok ? x : __builtin_nan(""), q);
    const double c = detail::cos_kernel(r);
    const double s = detail::sin_kernel(r);
    const double v = (q & 1) ? s : c;
    const uint64_t sign = (uint64_t)(((unsigned)q + 1u) & 2u) << 62;   // quadrants 1 and 2
    return from_bits(bits(v) ^ sign);
}

MPGOS_HD MPGOS_INLINE double sin(double x)
{
    const bool ok = fabs(x) < 0x1p50;
    int q;
    const double r = detail::reduce_pio2(ok ? x : __builtin_nan(""), q);
    const double c = detail::cos_kernel(r);
    const double s = detail::sin_kernel(r);
    const double v = (q & 1) ? c : s;
    const uint64_t sign = (uint64_t)((unsigned)q & 2u) << 62;           // quadrants 2 and 3
    return from_bits(bits(v) ^ sign);
}

// ---------------------------------------------------------------- exp
namespace detail {
constexpr double LOG2E = 0x1.71547652b82fep+0;
constexpr double LN2_HI = 0x1.62e42fefa39efp-1;    // double nearest ln 2
constexpr double LN2_LO = 0x1.abc9e3b39803fp-56;   // ln 2 - LN2_HI

// e^r = 1 + r + r^2 E(r), |r| <= ln2/2
MPGOS_HD MPGOS_INLINE double exp_kernel(double r)
{
    double e = 0x1.1f72fc730b4ffp-29;
    e = fma(e, r, 0x1.af4ddd84882fep-26);
    e = fma(e, r, 0x1.27e4db67b4303p-22);
    e = fma(e, r, 0x1.71de02375656cp-19);
    e = fma(e, r, 0x1.a01a01a6d7808p-16);
    e = fma(e, r, 0x1.a01a01abe62ddp-13);
    e = fma(e, r, 0x1.6c16c16c162d6p-10);
    e = fma(e, r, 0x1.11111111100dfp-7);
    e = fma(e, r, 0x1.5555555555556p-5);
    e = fma(e, r, 0x1.5555555555557p-3);
    e = fma(e, r, 0.5);
    return 1.0 + fma(r * r, e, r);
}
} // namespace detail

// The range checks select the result after the polynomial instead of branching
// around it: a wavefront then runs one instruction stream whatever its lanes'
// arguments (no divergent paths, no saved execution masks).  The value is the
// same as with early returns for every argument.
MPGOS_HD MPGOS_INLINE double exp(double x)
{
    constexpr double HI = 709.782712893384, LO = -745.1332191019412;
    const double xc = x >= LO ? (x <= HI ? x : HI) : LO;     // NaN -> LO, replaced below
    const double n = rint(xc * detail::LOG2E);
    double r = fma(-n, detail::LN2_HI, xc);   // exact (see reduce_pio2)
    r = fma(-n, detail::LN2_LO, r);
    const double p = detail::exp_kernel(r);
    const int k = (int)n;
    // two-step scaling keeps 2^k representable near the overflow / subnormal ends
    const int k1 = k / 2;
    double v = (p * pow2i(k1)) * pow2i(k - k1);
    v = x < LO ? 0.0 : v;
    v = x > HI ? from_bits(0x7ff0000000000000ull) : v;
    return x != x ? x : v;
}

// ---------------------------------------------------------------- log
namespace detail {
// 2 atanh(s) = 2 s + s R(z), z = s^2, R(z) = z * L(z), |s| <= (sqrt2-1)/(sqrt2+1)
MPGOS_HD MPGOS_INLINE double log_R(double z)
{
    double l = 0x1.2b584aae78a57p-3;
    l = fma(l, z, 0x1.39fe606542ddep-3);
    l = fma(l, z, 0x1.7462b4ab2ef6bp-3);
    l = fma(l, z, 0x1.c71c62e5800a1p-3);
    l = fma(l, z, 0x1.2492492df148dp-2);
    l = fma(l, z, 0x1.99999999952e2p-2);
    l = fma(l, z, 0x1.5555555555558p-1);
    return z * l;
}
} // namespace detail

MPGOS_HD MPGOS_INLINE double log(double x)
{
    if (x != x) return x;
    if (x < 0.0) return from_bits(0x7ff8000000000000ull);
    if (x == 0.0) return -from_bits(0x7ff0000000000000ull);
    if (!isfinite(x)) return x;
    int k = 0;
    if (x < 0x1p-1022) { x *= 0x1p54; k = -54; }       // subnormal input
    uint64_t u = bits(x);
    k += (int)((u >> 52) & 0x7ff) - 1023;
    double m = from_bits((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);   // [1, 2)
    if (m > 0x1.6a09e667f3bcdp+0) { m *= 0.5; k += 1; }                        // [sqrt2/2, sqrt2)
    const double f = m - 1.0;                 // exact (Sterbenz)
    const double s = f / (2.0 + f);
    const double hfsq = 0.5 * f * f;
    const double R = detail::log_R(s * s);
    const double dk = (double)k;
    const double lo = fma(dk, detail::LN2_LO, s * (hfsq + R));
    return fma(dk, detail::LN2_HI, f - (hfsq - lo));
}

namespace detail {
// log(x) as an unevaluated sum hi + lo (absolute error ~1e-16 * |lo|-scale),
// x > 0 finite.  ln 2 split so that k * LN2_HI_SPLIT is exact for |k| < 2^11.
constexpr double LN2_HI_SPLIT = 0x1.62e42feep-1;
constexpr double LN2_LO_SPLIT = 0x1.a39ef35793c76p-33;
MPGOS_HD MPGOS_INLINE void log_dd(double x, double& hi, double& lo)
{
    int k = 0;
    if (x < 0x1p-1022) { x *= 0x1p54; k = -54; }
    const uint64_t u = bits(x);
    k += (int)((u >> 52) & 0x7ff) - 1023;
    double m = from_bits((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 0x1.6a09e667f3bcdp+0) { m *= 0.5; k += 1; }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double hfsq = 0.5 * f * f;
    const double R = log_R(s * s);
    const double L = f - (hfsq - s * (hfsq + R));   // log(1 + f), |L| <= 0.35
    const double dk = (double)k;
    const double a = dk * LN2_HI_SPLIT;             // exact
    const double sum = a + L;                       // two-sum a + L
    const double bb = sum - a;
    const double err = (a - (sum - bb)) + (L - bb);
    hi = sum;
    lo = fma(dk, LN2_LO_SPLIT, err);
}
} // namespace detail

// x^y = exp(y log x) with log x carried in double-double and the product's
// rounding error folded back (exp(p + e) ~ exp(p) (1 + e)), so the result is
// within a few ulp over the controller's range (x up to 1e30, y = 1/5, 1/4).
// IEEE special cases for x in {0, inf, NaN}; negative x gives NaN unless y is
// an integer.  The general formula runs for every argument (on 1 where x is 0,
// inf or NaN) and the special cases then override it from the lowest to the
// highest precedence, so no lane branches.
MPGOS_HD MPGOS_INLINE double pow(double x, double y)
{
    constexpr double INF = __builtin_huge_val(), QNAN = __builtin_nan("");
    const double ax = fabs(x);
    const bool regular = ax > 0.0 && isfinite(ax);
    double lh, ll;
    detail::log_dd(regular ? ax : 1.0, lh, ll);
    const double ph = y * lh;
    const double pl = fma(y, lh, -ph) + y * ll;
    const double e = exp(ph);
    double r = (isfinite(e) && e != 0.0) ? fma(e, pl, e) : e;
    // negative base: integer exponents only, the sign from the exponent's parity
    const double half = 0.5 * y;
    const bool odd = rint(half) != half;
    if (x < 0.0) r = rint(y) != y ? QNAN : (odd ? -r : r);
    if (!isfinite(x)) r = x > 0.0 ? (y > 0.0 ? x : 0.0) : QNAN;
    if (x == 0.0) r = y > 0.0 ? 0.0 : INF;
    if (x == 1.0) r = 1.0;
    if (y == 0.0) r = 1.0;
    if (x != x || y != y) r = x + y;
    return r;
}

// pow for the error controllers' R^(1/5) and R^(1/4): R = min(tol/err) is
// >= 0 or NaN.  On the domain x >= 0, +inf or NaN and 0 < y <= 1/4 it returns
// pow(x, y)'s bits: |y log x| < 187, so exp neither overflows nor
// underflows and the scaling is one exact multiply; x = 1 gives exactly 1
// through the general formula.  Without pow's range selections and
// negative-base handling the controller's step is ~35 instructions shorter.
MPGOS_HD MPGOS_INLINE double pow_ctrl(double x, double y)
{
    constexpr double MAXF = 0x1.fffffffffffffp+1023;
    const bool regular = x > 0.0 && x <= MAXF;
    double lh, ll;
    detail::log_dd(regular ? x : 1.0, lh, ll);
    const double ph = y * lh;
    const double pl = fma(y, lh, -ph) + y * ll;
    const double n = rint(ph * detail::LOG2E);
    double r = fma(-n, detail::LN2_HI, ph);
    r = fma(-n, detail::LN2_LO, r);
    const double e = detail::exp_kernel(r) * pow2i((int)n);
    double v = fma(e, pl, e);
    v = x == 0.0 ? 0.0 : v;
    v = x > MAXF ? x : v;
    return x != x ? x + y : v;
}

// ---------------------------------------------------------------- controller roots
// The per-thread controllers' factor R^(1/5) (accepted step) and R^(1/4)
// (rejected step), R >= 0, from roots instead of exp(y log R):
//  * R^(1/4) = sqrt(sqrt(R)), two correctly rounded square roots (< 1 ulp);
//  * R^(1/5): R = 2^e m, e = 5q + r (0 <= r < 5), z = 2^r m in [1, 32);
//    w ~ z^(-1/5) starts from a quartic in m (m^(-1/5) on [1, 2], relative
//    error 2.1e-5) times a quartic in r through the five values 2^(-r/5),
//    then two Newton steps w += w (1 - z w^5) / 5 (error 3 eps^2 per step);
//    z^(1/5) = z w^4 corrected by its residual, scaled by 2^q: a few ulp.
// Only IEEE operations and integer arithmetic, so host and device agree bit
// for bit.  Subnormal R is scaled by 2^1000 first; 0 gives 0, +inf gives
// +inf, NaN gives NaN.
namespace detail {
MPGOS_HD MPGOS_INLINE double root5(double x)
{
    const bool tiny = x < 0x1p-1000;
    const double xs = tiny ? x * 0x1p1000 : x;
    const uint64_t b = bits(xs);
    const int e = (int)(b >> 52) - 1023;
    const double m = from_bits((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    const int q = (e + 1025) / 5 - 205;          // floor(e / 5)
    const int r = e - 5 * q;                     // 0 .. 4
    const double z = m * pow2i(r);
    double pm = 0x1.d2cadb0bb3296p-7;            // m^(-1/5), Chebyshev fit on [1, 2]
    pm = fma(pm, m, -0x1.ca9fab0a0d793p-4);
    pm = fma(pm, m, 0x1.714a36f55ea3cp-2);
    pm = fma(pm, m, -0x1.484ddb53b357dp-1);
    pm = fma(pm, m, 0x1.60d7515a3c0e5p+0);
    const double rd = (double)r;
    double pr = 0x1.88970ec6f3ab9p-17;           // 2^(-r/5) at r = 0 .. 4
    pr = fma(pr, rd, -0x1.c4b4e3b1a9125p-12);
    pr = fma(pr, rd, 0x1.3a4e88f4a3e7ap-7);
    pr = fma(pr, rd, -0x1.1be579c2b43dbp-3);
    pr = fma(pr, rd, 1.0);
    double w = pm * pr;
    for (int i = 0; i < 2; i++) {
        const double w2 = w * w;
        const double w5 = (w2 * w2) * w;
        w = fma(0.2 * w, fma(-z, w5, 1.0), w);
    }
    const double w2 = w * w, w4 = w2 * w2;
    const double y = z * w4;
    const double d = fma(-z, w4 * w, 1.0);      // w = z^(-1/5) (1 - d / 5): y = z^(1/5) (1 - 4 d / 5)
    double v = fma(0.8 * d, y, y) * pow2i(tiny ? q - 200 : q);
    v = x == 0.0 ? 0.0 : v;
    v = x > 0x1.fffffffffffffp+1023 ? x : v;
    return x != x ? x : v;
}
} // namespace detail

MPGOS_HD MPGOS_INLINE double ctrl_root(double x, bool fifth)
{
    if (fifth) return detail::root5(x);
    return sqrt(sqrt(x));
}

// ---------------------------------------------------------------- float versions
// Used by the FP32 stepping mode: computed through the double routines and
// rounded once, so float results are deterministic and within 1 ulp.
MPGOS_HD MPGOS_INLINE float cos(float x) { return (float)cos((double)x); }
MPGOS_HD MPGOS_INLINE float sin(float x) { return (float)sin((double)x); }
MPGOS_HD MPGOS_INLINE float exp(float x) { return (float)exp((double)x); }
MPGOS_HD MPGOS_INLINE float log(float x) { return (float)log((double)x); }
MPGOS_HD MPGOS_INLINE float pow(float x, float y) { return (float)pow((double)x, (double)y); }
MPGOS_HD MPGOS_INLINE float pow_ctrl(float x, float y) { return (float)pow_ctrl((double)x, (double)y); }
MPGOS_HD MPGOS_INLINE float ctrl_root(float x, bool fifth) { return (float)ctrl_root((double)x, fifth); }
MPGOS_HD MPGOS_INLINE float sqrt(float x) { return (float)sqrt((double)x); }   // exact: double has > 2 * 24 + 2 bits

} // namespace math
} // namespace mpgos
