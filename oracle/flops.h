// ORACLE -- test infrastructure only (see pt_oracle.h).
//
// Op-counting scalar: runs the oracle's own solver code and counts the
// algorithmic floating-point operations it performs (add/sub/mul/div = 1,
// fused multiply-add = 2, comparisons, fabs, fmin, fmax = 0).  The library's
// elementary functions are charged the operations their source performs
// (csrc/mpgos_math.h): cos or sin 38 (reduction 5, both polynomial kernels
// 19 + 14), exp 31, log 28, pow 70 (double-double log 32, product 5,
// exp 31, final fma 2), sqrt 1, the per-thread controller's root 44 (R^(1/5):
// start polynomials 17, two Newton steps 16, result and correction 11) or 2
// (R^(1/4) = sqrt(sqrt(R))).  Used for the
// roofline's algorithmic FLOPs per step.
#pragma once

#include <cmath>

namespace oracle {
namespace flops {

extern thread_local long long g_count;

struct F {
    double v;
    F() : v(0.0) {}
    F(double x) : v(x) {}
    explicit operator double() const { return v; }
};

inline F operator+(F a, F b) { g_count += 1; return F(a.v + b.v); }
inline F operator-(F a, F b) { g_count += 1; return F(a.v - b.v); }
inline F operator*(F a, F b) { g_count += 1; return F(a.v * b.v); }
inline F operator/(F a, F b) { g_count += 1; return F(a.v / b.v); }
inline F operator-(F a) { return F(-a.v); }
inline F& operator+=(F& a, F b) { a = a + b; return a; }
inline bool operator<(F a, F b) { return a.v < b.v; }
inline bool operator>(F a, F b) { return a.v > b.v; }
inline bool operator<=(F a, F b) { return a.v <= b.v; }
inline bool operator>=(F a, F b) { return a.v >= b.v; }
inline bool operator==(F a, F b) { return a.v == b.v; }
inline bool operator!=(F a, F b) { return a.v != b.v; }

inline F fma(F a, F b, F c) { g_count += 2; return F(std::fma(a.v, b.v, c.v)); }
inline F fabs(F a) { return F(std::fabs(a.v)); }
inline F fmin(F a, F b) { return F(std::fmin(a.v, b.v)); }
inline F fmax(F a, F b) { return F(std::fmax(a.v, b.v)); }
inline bool isfinite(F a) { return std::isfinite(a.v); }
inline F frozen(F a) { return a; }
F cos(F a);
F sin(F a);
F exp(F a);
F log(F a);
F pow(F a, F b);
F ctrl_root(F a, bool fifth);
F sqrt(F a);

} // namespace flops
} // namespace oracle
