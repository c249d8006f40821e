// ORACLE — test infrastructure only (CPU checker; never part of the product path).
//
// Restatement of toolbox/Rotations.jl and the small static-array algebra the
// beam element uses (src/Dots.jl ∘₁ on SMatrix/SVector). Matrices are stored
// column-major like Julia's SMatrix: m[i + 3*j] is M[i+1, j+1].
#pragma once
#include <cmath>
#include <limits>
#include "dual.hpp"

namespace orc {

template <class T> struct V3 { T v[3]; T& operator[](int i) { return v[i]; } const T& operator[](int i) const { return v[i]; } };
template <class T> struct M3 { T m[9]; T& operator()(int i, int j) { return m[i + 3 * j]; } const T& operator()(int i, int j) const { return m[i + 3 * j]; } };

template <class T> inline V3<T> vzero() { V3<T> r; for (int i = 0; i < 3; ++i) r[i] = T(0.); return r; }
template <class T> inline M3<T> mzero() { M3<T> r; for (int i = 0; i < 9; ++i) r.m[i] = T(0.); return r; }

template <class A, class B> inline auto vadd(const V3<A>& a, const V3<B>& b) {
    V3<decltype(a[0] + b[0])> r; for (int i = 0; i < 3; ++i) r[i] = a[i] + b[i]; return r; }
template <class A, class B> inline auto vsub(const V3<A>& a, const V3<B>& b) {
    V3<decltype(a[0] - b[0])> r; for (int i = 0; i < 3; ++i) r[i] = a[i] - b[i]; return r; }
template <class A, class S> inline auto vscale(const V3<A>& a, const S& s) {
    V3<decltype(a[0] * s)> r; for (int i = 0; i < 3; ++i) r[i] = a[i] * s; return r; }
template <class A, class S> inline auto vscale(const S& s, const V3<A>& a) {
    V3<decltype(s * a[0])> r; for (int i = 0; i < 3; ++i) r[i] = s * a[i]; return r; }
// M ∘₁ v
template <class A, class B> inline auto mv(const M3<A>& m, const V3<B>& v) {
    V3<decltype(m(0, 0) * v[0])> r;
    for (int i = 0; i < 3; ++i) r[i] = m(i, 0) * v[0] + m(i, 1) * v[1] + m(i, 2) * v[2];
    return r; }
// M' ∘₁ v
template <class A, class B> inline auto mtv(const M3<A>& m, const V3<B>& v) {
    V3<decltype(m(0, 0) * v[0])> r;
    for (int i = 0; i < 3; ++i) r[i] = m(0, i) * v[0] + m(1, i) * v[1] + m(2, i) * v[2];
    return r; }
// A ∘₁ B
template <class A, class B> inline auto mm(const M3<A>& a, const M3<B>& b) {
    M3<decltype(a(0, 0) * b(0, 0))> r;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j)
        r(i, j) = a(i, 0) * b(0, j) + a(i, 1) * b(1, j) + a(i, 2) * b(2, j);
    return r; }
// A' ∘₁ B
template <class A, class B> inline auto mtm(const M3<A>& a, const M3<B>& b) {
    M3<decltype(a(0, 0) * b(0, 0))> r;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j)
        r(i, j) = a(0, i) * b(0, j) + a(1, i) * b(1, j) + a(2, i) * b(2, j);
    return r; }
// A ∘₁ B'
template <class A, class B> inline auto mmt(const M3<A>& a, const M3<B>& b) {
    M3<decltype(a(0, 0) * b(0, 0))> r;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j)
        r(i, j) = a(i, 0) * b(j, 0) + a(i, 1) * b(j, 1) + a(i, 2) * b(j, 2);
    return r; }
template <class T> inline M3<T> madd(const M3<T>& a, const M3<T>& b) { M3<T> r; for (int i = 0; i < 9; ++i) r.m[i] = a.m[i] + b.m[i]; return r; }

// ------------------------------------------------------------- sinc1 family
// Rotations.jl:13-45. sinc1(x) = Base.sinc(x/π); Base.sinc switches to the
// series 1 - (πy)²/6 + (πy)⁴/120 for |y| < 1e-3, y = x/π.
// sinpi with Base's reduction to [-1/4,1/4] around multiples of 1/2, so that
// integer arguments give exact zeros (sinc1(π) == 0).
inline double sinpi_(double x_) {
    double x = std::fabs(x_);
    double n = std::nearbyint(2. * x);
    double rx = -0.5 * n + x;
    long q = (long)n & 3;
    double res;
    if (q == 0) res = std::sin(M_PI * rx);
    else if (q == 1) res = std::cos(M_PI * rx);
    else if (q == 2) res = 0. - std::sin(M_PI * rx);
    else res = 0. - std::cos(M_PI * rx);
    return std::signbit(x_) ? -res : res;
}
inline double sinc1_0(double x) {
    double y = x / M_PI;
    if (std::fabs(y) < 1e-3) {
        double y2 = y * y;
        return 1. + y2 * (-(M_PI * M_PI) / 6. + y2 * (M_PI * M_PI * M_PI * M_PI / 120.));
    }
    return sinpi_(y) / (M_PI * y);
}
inline double sinc1_1(double x) {
    if (std::fabs(x) > 1e-3) { double s = std::sin(x), c = std::cos(x); return c / x - s / (x * x); }
    double x2 = x * x; return x * (-1. / 3. + x2 / 30.);
}
inline double sinc1_2(double x) {
    if (std::fabs(x) > 1e-1) { double s = std::sin(x), c = std::cos(x); return -s / x - 2. * c / (x * x) + 2. * s / (x * x * x); }
    double x2 = x * x; return -1. / 3. + x2 * (1. / 10. + x2 * (-1. / 168. + x2 * (1. / 6480.)));
}
inline double sinc1_3(double x) {
    if (std::fabs(x) > 0.4) { double s = std::sin(x), c = std::cos(x);
        return -c / x + 3. * s / (x * x) + 6. * c / (x * x * x) - 6. * s / (x * x * x * x); }
    double x2 = x * x; return x * (1. / 5. + x2 * (-1. / 42. + x2 * (1. / 1080. + x2 * (-1. / 55440. + x2 * (1. / 4717440.)))));
}
inline double sinc1_4(double x) {
    double x2 = x * x;
    return 1. / 5. + x2 * (-1. / 14. + x2 * (1. / 216. + x2 * (-1. / 7920. + x2 * (1. / 524160. + x2 * (-1. / 54432000. +
           x2 * (1. / 54432000. + x2 * (-1. / 8143027200. + x2 * (1. / 1656387532800.))))))));
}
inline double sinc1_5(double x) { return x * std::numeric_limits<double>::quiet_NaN(); }

template <int K> inline double sinc1_k(double x) {
    if constexpr (K == 0) return sinc1_0(x);
    else if constexpr (K == 1) return sinc1_1(x);
    else if constexpr (K == 2) return sinc1_2(x);
    else if constexpr (K == 3) return sinc1_3(x);
    else if constexpr (K == 4) return sinc1_4(x);
    else return sinc1_5(x);
}
// DiffRule1 chain: d/dx sinc1^(K)(x) = sinc1^(K+1)(x)   (Rotations.jl:47-51)
template <int K, int N, class T> inline Dual<N, T> sinc1_k(const Dual<N, T>& a) {
    Dual<N, T> r; r.x = sinc1_k<K>(a.x);
    T d = sinc1_k<(K < 5 ? K + 1 : 5)>(a.x);
    for (int i = 0; i < N; ++i) r.dx[i] = d * a.dx[i];
    return r;
}
template <class T> inline T sinc1(const T& x) { return sinc1_k<0>(x); }

// scac(x) = sinc1(acos(x)), with a 5th order Taylor series for |x-1|<=1e-3 (Rotations.jl:61-68)
template <class T> inline T scac(const T& x) {
    T dx = x - 1.;
    if (std::fabs(VALUE(dx)) > 1e-3) return sinc1(dacos(x));
    return 1. + dx * (1. / 3. + dx * (-2. / 90. + dx * (0.0052911879917544626 + dx * (-0.0016229317117234072 + dx * (0.0005625)))));
}

template <class T> inline M3<T> spin(const V3<T>& v) {
    M3<T> m;
    m(0, 0) = T(0.); m(1, 0) = v[2]; m(2, 0) = -v[1];
    m(0, 1) = -v[2]; m(1, 1) = T(0.); m(2, 1) = v[0];
    m(0, 2) = v[1]; m(1, 2) = -v[0]; m(2, 2) = T(0.);
    return m;
}
template <class T> inline V3<T> spin_inv(const M3<T>& m) {
    V3<T> r;
    r[0] = (m(2, 1) - m(1, 2)) / 2.;
    r[1] = (m(0, 2) - m(2, 0)) / 2.;
    r[2] = (m(1, 0) - m(0, 1)) / 2.;
    return r;
}
template <class T> inline T trace(const M3<T>& m) { return m(0, 0) + m(1, 1) + m(2, 2); }

template <class T> inline T norm3(const V3<T>& v) {   // Rotations.jl:106-112
    T n = dsqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (VALUE(n) < 1e-14) n = T(0.);
    return n;
}
template <class T> inline M3<T> spin2(const M3<T>& S) {   // Rotations.jl:114-122
    T ab = S(1, 2) * S(2, 0), ca = S(0, 1) * S(1, 2), bc = S(2, 0) * S(0, 1);
    T ma2 = S(2, 1) * S(1, 2), mb2 = S(2, 0) * S(0, 2), mc2 = S(1, 0) * S(0, 1);
    M3<T> r;
    r(0, 0) = mc2 + mb2; r(1, 0) = ab; r(2, 0) = ca;
    r(0, 1) = ab; r(1, 1) = ma2 + mc2; r(2, 1) = bc;
    r(0, 2) = ca; r(1, 2) = bc; r(2, 2) = ma2 + mb2;
    return r;
}
template <class T> inline M3<T> Rodrigues(const V3<T>& v) {   // Rotations.jl:131-135
    M3<T> S = spin(v);
    T th = norm3(v);
    T a = sinc1(th);
    T b = dsq(sinc1(th / 2.)) / 2.;
    M3<T> S2 = spin2(S);
    M3<T> r;
    for (int i = 0; i < 9; ++i) r.m[i] = a * S.m[i] + b * S2.m[i];
    for (int i = 0; i < 3; ++i) r(i, i) = 1. + r(i, i);
    return r;
}
template <class T> inline V3<T> Rodrigues_inv(const M3<T>& m) {   // Rotations.jl:105
    V3<T> s = spin_inv(m);
    T d = scac((trace(m) - 1.) / 2.);
    V3<T> r; for (int i = 0; i < 3; ++i) r[i] = s[i] / d;
    return r;
}
// Rotations.jl:162-168
template <class T> inline V3<T> adjust(V3<T> u, V3<T> v) {
    T nu = dsqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    T nv = dsqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int i = 0; i < 3; ++i) { u[i] = u[i] / nu; v[i] = v[i] / nv; }
    T c = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
    V3<T> w; w[0] = u[1] * v[2] - u[2] * v[1]; w[1] = u[2] * v[0] - u[0] * v[2]; w[2] = u[0] * v[1] - u[1] * v[0];
    T s = norm3(w);
    T th = datan2(s, c);
    T sc = sinc1(th);
    V3<T> r; for (int i = 0; i < 3; ++i) r[i] = w[i] / sc;
    return r;
}

}  // namespace orc
