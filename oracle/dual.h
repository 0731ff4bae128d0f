// ORACLE -- test infrastructure only (see pt_oracle.h).
//
// Forward-mode dual numbers with N directions: running the oracle's solver
// over Dual<N> gives the tangent-linear (forward-sensitivity) derivatives of
// the discrete map.  Step sizes are frozen (their derivative dropped after the
// controller and the event control), so the derivatives are those of the
// accepted-step sequence actually taken -- the map the discrete adjoint
// differentiates.
#pragma once

#include <cmath>

#include "mpgos_math.h"

namespace oracle {
namespace dual {

template <int N>
struct D {
    double v;
    double d[N];
    D() : v(0.0) { for (int i = 0; i < N; i++) d[i] = 0.0; }
    D(double x) : v(x) { for (int i = 0; i < N; i++) d[i] = 0.0; }
    explicit operator double() const { return v; }
};

template <int N> D<N> operator+(const D<N>& a, const D<N>& b) { D<N> r(a.v + b.v); for (int i = 0; i < N; i++) r.d[i] = a.d[i] + b.d[i]; return r; }
template <int N> D<N> operator-(const D<N>& a, const D<N>& b) { D<N> r(a.v - b.v); for (int i = 0; i < N; i++) r.d[i] = a.d[i] - b.d[i]; return r; }
template <int N> D<N> operator-(const D<N>& a) { D<N> r(-a.v); for (int i = 0; i < N; i++) r.d[i] = -a.d[i]; return r; }
template <int N> D<N> operator*(const D<N>& a, const D<N>& b) { D<N> r(a.v * b.v); for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
template <int N> D<N> operator/(const D<N>& a, const D<N>& b)
{
    D<N> r(a.v / b.v);
    for (int i = 0; i < N; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
    return r;
}
template <int N> D<N>& operator+=(D<N>& a, const D<N>& b) { a = a + b; return a; }
template <int N> bool operator<(const D<N>& a, const D<N>& b) { return a.v < b.v; }
template <int N> bool operator>(const D<N>& a, const D<N>& b) { return a.v > b.v; }
template <int N> bool operator<=(const D<N>& a, const D<N>& b) { return a.v <= b.v; }
template <int N> bool operator>=(const D<N>& a, const D<N>& b) { return a.v >= b.v; }
template <int N> bool operator==(const D<N>& a, const D<N>& b) { return a.v == b.v; }
template <int N> bool operator!=(const D<N>& a, const D<N>& b) { return a.v != b.v; }

template <int N> D<N> fma(const D<N>& a, const D<N>& b, const D<N>& c)
{
    D<N> r(std::fma(a.v, b.v, c.v));
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i] + c.d[i];
    return r;
}
template <int N> D<N> fabs(const D<N>& a) { return a.v < 0.0 ? -a : a; }
template <int N> D<N> fmin(const D<N>& a, const D<N>& b) { return (b.v < a.v || a.v != a.v) ? b : a; }
template <int N> D<N> fmax(const D<N>& a, const D<N>& b) { return (b.v > a.v || a.v != a.v) ? b : a; }
template <int N> bool isfinite(const D<N>& a) { return std::isfinite(a.v); }
template <int N> D<N> cos(const D<N>& a)
{
    D<N> r(mpgos::math::cos(a.v));
    const double s = -mpgos::math::sin(a.v);
    for (int i = 0; i < N; i++) r.d[i] = s * a.d[i];
    return r;
}
template <int N> D<N> sin(const D<N>& a)
{
    D<N> r(mpgos::math::sin(a.v));
    const double c = mpgos::math::cos(a.v);
    for (int i = 0; i < N; i++) r.d[i] = c * a.d[i];
    return r;
}
template <int N> D<N> sqrt(const D<N>& a)
{
    D<N> r(mpgos::math::sqrt(a.v));
    const double g = 0.5 / r.v;
    for (int i = 0; i < N; i++) r.d[i] = g * a.d[i];
    return r;
}
// only the controller calls pow; its result is frozen, so the value suffices
template <int N> D<N> pow(const D<N>& a, const D<N>& b) { return D<N>(mpgos::math::pow(a.v, b.v)); }
// the per-thread controller's root, frozen like pow
template <int N> D<N> ctrl_root(const D<N>& a, bool fifth) { return D<N>(mpgos::math::ctrl_root(a.v, fifth)); }
template <int N> D<N> frozen(const D<N>& a) { return D<N>(a.v); }

} // namespace dual

// Directions: one per state component, then one per control parameter.
template <int N> inline void seed_directions(dual::D<N>* x, int sd, dual::D<N>* cp, int ncp)
{
    for (int i = 0; i < sd && i < N; i++) x[i].d[i] = 1.0;
    for (int j = 0; j < ncp && sd + j < N; j++) cp[j].d[sd + j] = 1.0;
}

namespace dual {

} // namespace dual
} // namespace oracle
