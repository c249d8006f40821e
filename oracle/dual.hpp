// ORACLE — test infrastructure only. Nothing in the product path may include,
// link or call this code; it is the CPU checker for the HIP kernels.
//
// Nested forward-mode dual numbers, restating the semantics of Muscade's
// ∂ℝ{P,N,R} (src/Adiff.jl:13-16, rules :197-304). The nesting level plays the
// role of the precedence P: an outer Dual is a newer (higher-precedence)
// variation than the Dual it contains. Mixed arithmetic between levels treats
// the lower-level operand as a constant of the higher level, as the
// promote/DiffRule2 machinery does (Adiff.jl:46-62,204-221).
#pragma once
#include <cmath>
#include <type_traits>

namespace orc {

template <int N, class T> struct Dual;

template <class T> struct level { static constexpr int v = 0; };
template <int N, class T> struct level<Dual<N, T>> { static constexpr int v = 1 + level<T>::v; };

template <class T> struct npart { static constexpr int v = 0; };
template <int N, class T> struct npart<Dual<N, T>> { static constexpr int v = N; };

// lift a lower-level value into type T (zero partials at every added level)
template <class T, class S> inline T lift(const S& s);

template <int N, class T> struct Dual {
    T x;
    T dx[N];
    Dual() : x() { for (int i = 0; i < N; ++i) dx[i] = T(); }
    Dual(double v) : x(lift<T>(v)) { for (int i = 0; i < N; ++i) dx[i] = T(); }
    template <class S, std::enable_if_t<(level<S>::v < level<Dual>::v) && (level<S>::v > 0), int> = 0>
    Dual(const S& s) : x(lift<T>(s)) { for (int i = 0; i < N; ++i) dx[i] = T(); }
};

template <class T, class S> struct lifter { static T go(const S& s) { return T(s); } };
template <class S> struct lifter<S, S> { static S go(const S& s) { return s; } };
template <class T, class S> inline T lift(const S& s) { return lifter<T, S>::go(s); }

inline double VALUE(double a) { return a; }
template <int N, class T> inline double VALUE(const Dual<N, T>& a) { return VALUE(a.x); }

// Integer power a^b with Muscade's rule (Adiff.jl:230):
//   a^b = b==0 ? zero(a) : ∂ℝ(a.x^b, a.dx*b*a.x^(b-1)).
// For a dual, a^0 is zero(a) (not one), so the partials of a^1 taken at the
// nesting level below are lost; for Float64, Base's x^0 = 1 and x^2 = x*x.
// This is reproduced as is: the reference's own golden for the 4th derivative
// of scac at 0 (test/TestRotations.jl:63) is the value this rule produces.
inline double ipow(double a, int b) { if (b == 0) return 1.; double r = a; for (int k = 1; k < b; ++k) r = r * a; return r; }
template <int N, class T> inline Dual<N, T> ipow(const Dual<N, T>& a, int b) {
    Dual<N, T> r;
    if (b == 0) return r;   // zero(a)
    r.x = ipow(a.x, b); T p = ipow(a.x, b - 1);
    for (int i = 0; i < N; ++i) r.dx[i] = a.dx[i] * double(b) * p;
    return r; }
template <class T> inline T dsq(const T& a) { return ipow(a, 2); }

// ---------------------------------------------------------------- same type
template <int N, class T> inline Dual<N, T> operator+(const Dual<N, T>& a, const Dual<N, T>& b) {
    Dual<N, T> r; r.x = a.x + b.x; for (int i = 0; i < N; ++i) r.dx[i] = a.dx[i] + b.dx[i]; return r; }
template <int N, class T> inline Dual<N, T> operator-(const Dual<N, T>& a, const Dual<N, T>& b) {
    Dual<N, T> r; r.x = a.x - b.x; for (int i = 0; i < N; ++i) r.dx[i] = a.dx[i] - b.dx[i]; return r; }
template <int N, class T> inline Dual<N, T> operator-(const Dual<N, T>& a) {
    Dual<N, T> r; r.x = -a.x; for (int i = 0; i < N; ++i) r.dx[i] = -a.dx[i]; return r; }
template <int N, class T> inline Dual<N, T> operator*(const Dual<N, T>& a, const Dual<N, T>& b) {
    Dual<N, T> r; r.x = a.x * b.x; for (int i = 0; i < N; ++i) r.dx[i] = a.dx[i] * b.x + a.x * b.dx[i]; return r; }
template <int N, class T> inline Dual<N, T> operator/(const Dual<N, T>& a, const Dual<N, T>& b) {
    // a.dx/b.x - a.x/b.x^2*b.dx  (Adiff.jl:228)
    Dual<N, T> r; r.x = a.x / b.x; T b2 = ipow(b.x, 2);
    for (int i = 0; i < N; ++i) r.dx[i] = a.dx[i] / b.x - a.x / b2 * b.dx[i]; return r; }

// ------------------------------------------------- mixed with a lower level
#define ORC_LOWER(S, ...) std::enable_if_t<(level<S>::v < level<__VA_ARGS__>::v), int> = 0
template <int N, class T, class S, ORC_LOWER(S, Dual<N, T>)>
inline Dual<N, T> operator+(const Dual<N, T>& a, const S& b) { Dual<N, T> r = a; r.x = a.x + b; return r; }
template <int N, class T, class S, ORC_LOWER(S, Dual<N, T>)>
inline Dual<N, T> operator+(const S& b, const Dual<N, T>& a) { Dual<N, T> r = a; r.x = b + a.x; return r; }
template <int N, class T, class S, ORC_LOWER(S, Dual<N, T>)>
inline Dual<N, T> operator-(const Dual<N, T>& a, const S& b) { Dual<N, T> r = a; r.x = a.x - b; return r; }
template <int N, class T, class S, ORC_LOWER(S, Dual<N, T>)>
inline Dual<N, T> operator-(const S& b, const Dual<N, T>& a) {
    Dual<N, T> r; r.x = b - a.x; for (int i = 0; i < N; ++i) r.dx[i] = -a.dx[i]; return r; }
template <int N, class T, class S, ORC_LOWER(S, Dual<N, T>)>
inline Dual<N, T> operator*(const Dual<N, T>& a, const S& b) {
    Dual<N, T> r; r.x = a.x * b; for (int i = 0; i < N; ++i) r.dx[i] = a.dx[i] * b; return r; }
template <int N, class T, class S, ORC_LOWER(S, Dual<N, T>)>
inline Dual<N, T> operator*(const S& b, const Dual<N, T>& a) {
    Dual<N, T> r; r.x = b * a.x; for (int i = 0; i < N; ++i) r.dx[i] = b * a.dx[i]; return r; }
template <int N, class T, class S, ORC_LOWER(S, Dual<N, T>)>
inline Dual<N, T> operator/(const Dual<N, T>& a, const S& b) {
    Dual<N, T> r; r.x = a.x / b; for (int i = 0; i < N; ++i) r.dx[i] = a.dx[i] / b; return r; }
template <int N, class T, class S, ORC_LOWER(S, Dual<N, T>)>
inline Dual<N, T> operator/(const S& a, const Dual<N, T>& b) {
    // -a/b.x^2*b.dx
    Dual<N, T> r; r.x = a / b.x; T b2 = ipow(b.x, 2);
    for (int i = 0; i < N; ++i) r.dx[i] = -a / b2 * b.dx[i]; return r; }

template <class A, class B> inline A& operator+=(A& a, const B& b) { a = a + b; return a; }
template <class A, class B> inline A& operator-=(A& a, const B& b) { a = a - b; return a; }
template <class A, class B> inline A& operator*=(A& a, const B& b) { a = a * b; return a; }
template <class A, class B> inline A& operator/=(A& a, const B& b) { a = a / b; return a; }

// comparisons act on VALUE (Adiff.jl:198-202)
template <class A, class B, std::enable_if_t<(level<A>::v + level<B>::v > 0), int> = 0>
inline bool operator<(const A& a, const B& b) { return VALUE(a) < VALUE(b); }
template <class A, class B, std::enable_if_t<(level<A>::v + level<B>::v > 0), int> = 0>
inline bool operator>(const A& a, const B& b) { return VALUE(a) > VALUE(b); }

// ------------------------------------------------------------- functions
// The derivative expressions follow the DiffRule1 forms of Adiff.jl:236-268.
inline double dsqrt(double a) { return std::sqrt(a); }
template <int N, class T> inline Dual<N, T> dsqrt(const Dual<N, T>& a) {
    Dual<N, T> r; r.x = dsqrt(a.x); T s = r.x;
    for (int i = 0; i < N; ++i) r.dx[i] = a.dx[i] / 2. / s; return r; }
inline double dsin(double a) { return std::sin(a); }
inline double dcos(double a) { return std::cos(a); }
template <int N, class T> inline Dual<N, T> dcos(const Dual<N, T>& a);
template <int N, class T> inline Dual<N, T> dsin(const Dual<N, T>& a) {
    Dual<N, T> r; r.x = dsin(a.x); T c = dcos(a.x);
    for (int i = 0; i < N; ++i) r.dx[i] = c * a.dx[i]; return r; }
template <int N, class T> inline Dual<N, T> dcos(const Dual<N, T>& a) {
    Dual<N, T> r; r.x = dcos(a.x); T s = dsin(a.x);
    for (int i = 0; i < N; ++i) r.dx[i] = -s * a.dx[i]; return r; }
inline double dacos(double a) { return std::acos(a); }
template <int N, class T> inline Dual<N, T> dacos(const Dual<N, T>& a) {
    Dual<N, T> r; r.x = dacos(a.x); T q = dsqrt(1. - ipow(a.x, 2));
    for (int i = 0; i < N; ++i) r.dx[i] = -a.dx[i] / q; return r; }
inline double dexp(double a) { return std::exp(a); }
template <int N, class T> inline Dual<N, T> dexp(const Dual<N, T>& a) {
    Dual<N, T> r; r.x = dexp(a.x); for (int i = 0; i < N; ++i) r.dx[i] = r.x * a.dx[i]; return r; }
inline double dexp10(double a) { return std::pow(10., a); }
template <int N, class T> inline Dual<N, T> dexp10(const Dual<N, T>& a) {
    Dual<N, T> r; r.x = dexp10(a.x); for (int i = 0; i < N; ++i) r.dx[i] = std::log(10.) * r.x * a.dx[i]; return r; }
inline double datan2(double a, double b) { return std::atan2(a, b); }
template <int N, class T> inline Dual<N, T> datan2(const Dual<N, T>& a, const Dual<N, T>& b) {
    // (a.dx*b.x-b.dx*a.x)/(a.x^2+b.x^2)  (Adiff.jl:223)
    Dual<N, T> r; r.x = datan2(a.x, b.x); T d = ipow(a.x, 2) + ipow(b.x, 2);
    for (int i = 0; i < N; ++i) r.dx[i] = (a.dx[i] * b.x - b.dx[i] * a.x) / d;
    return r;
}


// a variable with unit seed in partial i (0-based) — ∂ℝ{P,N}(x,i,s)
template <int N, class T> inline Dual<N, T> seed(const T& x, int i, double s = 1.) {
    Dual<N, T> r; r.x = x; r.dx[i] = lift<T>(s); return r; }

}  // namespace orc
