// Dynamic EulerBeam3D (X with time derivatives, OX = 1 or 2): time jets of the
// corotated kinematics and the inertia, added-mass, damping and roll-inertia loads.
//
// Reference: toolbox/BeamElement.jl:28-64 (resultants), :151-208 (residual,
// kinematics, corotated), toolbox/Rotations.jl:13-182, src/Taylor.jl:24-65
// (motion / motion⁻¹), src/SweepX.jl:45-74 (the solver's trial path).
//
// Muscade evaluates the element on nested forward-mode duals: two time levels
// (motion{P}) around the solver partials. Two things follow, and are reproduced:
//
//   * time derivatives are exact calculus through every formula, except that
//     the integer power rule of src/Adiff.jl:230 (a^0 → zero(a)) drops the
//     cross-level term of a square taken on a two-level time jet: for the
//     `sinc1(θ/2)^2` of Rodrigues (Rotations.jl:134) the second time
//     derivative is 2·s·s̈ instead of 2·s·s̈ + 2·ṡ². `jsq_muscade` below;
//   * the solver's tangent is the derivative along one trial direction ε:
//     X₀ + ε·e, X₁ + ε·a₁e, X₂ + ε·b₁e (:iter) or X₁ + ε·a, X₂ + ε·b (the δr
//     partial of :step).
//
// `Jet<NT,E>` carries q(t,ε) truncated at order NT-1 in t and 1 in ε:
// v[n] = ∂ⁿq/∂tⁿ and e[n] = ∂(∂ⁿq/∂tⁿ)/∂ε.  With E = false it is the plain time
// jet (value pass).
#pragma once
#include "beam_math.hpp"

namespace mh {

template <int NT, bool E> struct Jet {
    double v[NT];
    double e[E ? NT : 1];
};

template <int NT, bool E> MH_HD Jet<NT, E> jzero() {
    Jet<NT, E> r;
#pragma unroll
    for (int n = 0; n < NT; ++n) r.v[n] = 0.;
#pragma unroll
    for (int n = 0; n < (E ? NT : 1); ++n) r.e[n] = 0.;
    return r;
}
template <int NT, bool E> MH_HD Jet<NT, E> jconst(double c) { Jet<NT, E> r = jzero<NT, E>(); r.v[0] = c; return r; }
template <int NT, bool E> MH_HD Jet<NT, E> operator+(const Jet<NT, E>& a, const Jet<NT, E>& b) {
    Jet<NT, E> r;
#pragma unroll
    for (int n = 0; n < NT; ++n) r.v[n] = a.v[n] + b.v[n];
    if constexpr (E) {
#pragma unroll
        for (int n = 0; n < NT; ++n) r.e[n] = a.e[n] + b.e[n];
    } else r.e[0] = 0.;
    return r;
}
template <int NT, bool E> MH_HD Jet<NT, E> operator-(const Jet<NT, E>& a, const Jet<NT, E>& b) {
    Jet<NT, E> r;
#pragma unroll
    for (int n = 0; n < NT; ++n) r.v[n] = a.v[n] - b.v[n];
    if constexpr (E) {
#pragma unroll
        for (int n = 0; n < NT; ++n) r.e[n] = a.e[n] - b.e[n];
    } else r.e[0] = 0.;
    return r;
}
template <int NT, bool E> MH_HD Jet<NT, E> operator-(const Jet<NT, E>& a) {
    Jet<NT, E> r;
#pragma unroll
    for (int n = 0; n < NT; ++n) r.v[n] = -a.v[n];
    if constexpr (E) {
#pragma unroll
        for (int n = 0; n < NT; ++n) r.e[n] = -a.e[n];
    } else r.e[0] = 0.;
    return r;
}
template <int NT, bool E> MH_HD Jet<NT, E> operator*(double s, const Jet<NT, E>& a) {
    Jet<NT, E> r;
#pragma unroll
    for (int n = 0; n < NT; ++n) r.v[n] = s * a.v[n];
    if constexpr (E) {
#pragma unroll
        for (int n = 0; n < NT; ++n) r.e[n] = s * a.e[n];
    } else r.e[0] = 0.;
    return r;
}
template <int NT, bool E> MH_HD Jet<NT, E> operator+(const Jet<NT, E>& a, double c) { Jet<NT, E> r = a; r.v[0] = a.v[0] + c; return r; }
// Leibniz product of truncated jets.
template <int NT, bool E> MH_HD Jet<NT, E> operator*(const Jet<NT, E>& a, const Jet<NT, E>& b) {
    Jet<NT, E> r;
    r.v[0] = a.v[0] * b.v[0];
    if constexpr (NT > 1) r.v[1] = a.v[1] * b.v[0] + a.v[0] * b.v[1];
    if constexpr (NT > 2) r.v[2] = a.v[2] * b.v[0] + 2. * (a.v[1] * b.v[1]) + a.v[0] * b.v[2];
    if constexpr (E) {
        r.e[0] = a.e[0] * b.v[0] + a.v[0] * b.e[0];
        if constexpr (NT > 1) r.e[1] = (a.e[1] * b.v[0] + a.v[1] * b.e[0]) + (a.e[0] * b.v[1] + a.v[0] * b.e[1]);
        if constexpr (NT > 2)
            r.e[2] = (a.e[2] * b.v[0] + a.v[2] * b.e[0]) + 2. * (a.e[1] * b.v[1] + a.v[1] * b.e[1]) +
                     (a.e[0] * b.v[2] + a.v[0] * b.e[2]);
    } else r.e[0] = 0.;
    return r;
}
// Composition with a scalar function given f, f', f'', f''' at a.v[0] (chain rule
// to third total order: two in t, one in ε).
template <int NT, bool E> MH_HD Jet<NT, E> jfun(const Jet<NT, E>& a, const double* f) {
    Jet<NT, E> r;
    r.v[0] = f[0];
    if constexpr (NT > 1) r.v[1] = f[1] * a.v[1];
    if constexpr (NT > 2) r.v[2] = f[2] * (a.v[1] * a.v[1]) + f[1] * a.v[2];
    if constexpr (E) {
        r.e[0] = f[1] * a.e[0];
        if constexpr (NT > 1) r.e[1] = f[2] * (a.e[0] * a.v[1]) + f[1] * a.e[1];
        if constexpr (NT > 2)
            r.e[2] = f[3] * (a.e[0] * (a.v[1] * a.v[1])) + f[2] * (2. * (a.v[1] * a.e[1]) + a.e[0] * a.v[2]) + f[1] * a.e[2];
    } else r.e[0] = 0.;
    return r;
}
// x^2 as Muscade's integer power rule evaluates it on a time jet (src/Adiff.jl:230):
// exact to first order in t; at second order only 2·a·ä survives (see header).
template <int NT, bool E> MH_HD Jet<NT, E> jsq_muscade(const Jet<NT, E>& a) {
    Jet<NT, E> r;
    r.v[0] = a.v[0] * a.v[0];
    if constexpr (NT > 1) r.v[1] = 2. * (a.v[0] * a.v[1]);
    if constexpr (NT > 2) r.v[2] = 2. * (a.v[0] * a.v[2]);
    if constexpr (E) {
        r.e[0] = 2. * (a.v[0] * a.e[0]);
        if constexpr (NT > 1) r.e[1] = 2. * (a.e[0] * a.v[1] + a.v[0] * a.e[1]);
        if constexpr (NT > 2) r.e[2] = 2. * (a.e[0] * a.v[2] + a.v[0] * a.e[2]);
    } else r.e[0] = 0.;
    return r;
}

// ------------------------------------------------------------ scalar functions
// Derivatives of sinc1 with the branches of Rotations.jl:14-45.
MH_HD double sinc1_d1(double x) {
    if (fabs(x) > 1e-3) { double s = sin(x), c = cos(x); return c / x - s / (x * x); }
    double x2 = x * x;
    return x * (-1. / 3. + x2 / 30.);
}
MH_HD double sinc1_d2(double x) {
    if (fabs(x) > 1e-1) { double s = sin(x), c = cos(x); return -s / x - 2. * c / (x * x) + 2. * s / (x * x * x); }
    double x2 = x * x;
    return -1. / 3. + x2 * (1. / 10. + x2 * (-1. / 168. + x2 * (1. / 6480.)));
}
MH_HD double sinc1_d3(double x) {
    if (fabs(x) > 0.4) {
        double s = sin(x), c = cos(x);
        return -c / x + 3. * s / (x * x) + 6. * c / (x * x * x) - 6. * s / (x * x * x * x);
    }
    double x2 = x * x;
    return x * (1. / 5. + x2 * (-1. / 42. + x2 * (1. / 1080. + x2 * (-1. / 55440. + x2 * (1. / 4717440.)))));
}
template <int NT, bool E> MH_HD Jet<NT, E> jsinc1(const Jet<NT, E>& a) {
    const double x = a.v[0];
    double f[4];
    f[0] = sinc1(x);
    f[1] = sinc1_d1(x);
    f[2] = (NT + (E ? 1 : 0) > 2) ? sinc1_d2(x) : 0.;
    f[3] = (NT + (E ? 1 : 0) > 3) ? sinc1_d3(x) : 0.;
    return jfun(a, f);
}
template <int NT, bool E> MH_HD Jet<NT, E> jsqrt(const Jet<NT, E>& a) {
    const double x = a.v[0], s = sqrt(x);
    double f[4] = {s, 0.5 / s, -0.25 / (x * s), 0.375 / (x * x * s)};
    return jfun(a, f);
}
template <int NT, bool E> MH_HD Jet<NT, E> jrecip(const Jet<NT, E>& a) {
    const double x = a.v[0], r = 1. / x;
    double f[4] = {r, -(r * r), 2. * (r * r * r), -6. * (r * r * r * r)};
    return jfun(a, f);
}
template <int NT, bool E> MH_HD Jet<NT, E> jacos(const Jet<NT, E>& a) {
    const double x = a.v[0], q = 1. - x * x, sq = sqrt(q);
    double f[4] = {acos(x), -1. / sq, -x / (q * sq), -(1. + 2. * x * x) / (q * q * sq)};
    return jfun(a, f);
}
// scac(x) = sinc1(acos x); 5th-order series in x-1 for |x-1| <= 1e-3 (Rotations.jl:61-68)
template <int NT, bool E> MH_HD Jet<NT, E> jscac(const Jet<NT, E>& a) {
    const double d = a.v[0] - 1.;
    if (fabs(d) > 1e-3) return jsinc1(jacos(a));
    const double c1 = 1. / 3., c2 = -2. / 90., c3 = 0.0052911879917544626, c4 = -0.0016229317117234072, c5 = 0.0005625;
    double f[4];
    f[0] = 1. + d * (c1 + d * (c2 + d * (c3 + d * (c4 + d * c5))));
    f[1] = c1 + d * (2. * c2 + d * (3. * c3 + d * (4. * c4 + d * (5. * c5))));
    f[2] = 2. * c2 + d * (6. * c3 + d * (12. * c4 + d * (20. * c5)));
    f[3] = 6. * c3 + d * (24. * c4 + d * (60. * c5));
    return jfun(a, f);
}

// ------------------------------------------------------------ rotations on jets
template <int NT, bool E> struct JM3 { Jet<NT, E> a[9]; };   // column-major 3×3
template <int NT, bool E> struct JV3 { Jet<NT, E> a[3]; };

// Rodrigues(v) = I + sinc1(θ) S + (sinc1(θ/2)²/2) spin²(S), θ = norm3(v) (Rotations.jl:106-135):
// the coefficient jets a = sinc1(θ), b = sinc1(θ/2)²/2.
// XSQ = false: the square by Muscade's power rule on jets over solver partials (the
// assembly); XSQ = true: the exact square (jets over plain values, as getresult evaluates).
template <int NT, bool E, bool XSQ = false>
MH_HD void jrodrigues_coef(const JV3<NT, E>& v, Jet<NT, E>& a, Jet<NT, E>& b) {
    using J = Jet<NT, E>;
    const J& x = v.a[0]; const J& y = v.a[1]; const J& z = v.a[2];
    J ss = (x * x + y * y) + z * z;
    J th;
    if (sqrt(ss.v[0]) < 1e-14) th = jzero<NT, E>();
    else th = jsqrt(ss);
    a = jsinc1(th);
    J s = jsinc1(0.5 * th);
    if constexpr (XSQ) b = 0.5 * (s * s);
    else b = 0.5 * jsq_muscade(s);
}
// Column k of Rodrigues(v) from its coefficients, spin²(S) entrywise (Rotations.jl:114-122).
template <int K, int NT, bool E>
MH_HD JV3<NT, E> jrodrigues_col(const JV3<NT, E>& v, const Jet<NT, E>& a, const Jet<NT, E>& b) {
    const Jet<NT, E>& x = v.a[0]; const Jet<NT, E>& y = v.a[1]; const Jet<NT, E>& z = v.a[2];
    JV3<NT, E> r;
    if constexpr (K == 0) {
        r.a[0] = b * (-(z * z) + -(y * y)) + 1.;
        r.a[1] = a * z + b * (x * y);
        r.a[2] = -(a * y) + b * (z * x);
    } else if constexpr (K == 1) {
        r.a[0] = -(a * z) + b * (x * y);
        r.a[1] = b * (-(x * x) + -(z * z)) + 1.;
        r.a[2] = a * x + b * (y * z);
    } else {
        r.a[0] = a * y + b * (z * x);
        r.a[1] = -(a * x) + b * (y * z);
        r.a[2] = b * (-(x * x) + -(y * y)) + 1.;
    }
    return r;
}
template <int NT, bool E, bool XSQ = false> MH_HD JM3<NT, E> jrodrigues(const JV3<NT, E>& v) {
    Jet<NT, E> a, b;
    jrodrigues_coef<NT, E, XSQ>(v, a, b);
    JV3<NT, E> c0 = jrodrigues_col<0>(v, a, b), c1 = jrodrigues_col<1>(v, a, b), c2 = jrodrigues_col<2>(v, a, b);
    JM3<NT, E> r;
#pragma unroll
    for (int i = 0; i < 3; ++i) { r.a[i] = c0.a[i]; r.a[3 + i] = c1.a[i]; r.a[6 + i] = c2.a[i]; }
    return r;
}
// Rodrigues⁻¹(m) = spin⁻¹(m)/scac((tr m - 1)/2) (Rotations.jl:101)
template <int NT, bool E> MH_HD JV3<NT, E> jrodrigues_inv(const JM3<NT, E>& m) {
    using J = Jet<NT, E>;
    J tr = (m.a[0] + m.a[4]) + m.a[8];
    J rs = jrecip(jscac(0.5 * (tr + (-1.))));
    JV3<NT, E> r;
    r.a[0] = (0.5 * (m.a[5] - m.a[7])) * rs;
    r.a[1] = (0.5 * (m.a[6] - m.a[2])) * rs;
    r.a[2] = (0.5 * (m.a[1] - m.a[3])) * rs;
    return r;
}
template <int NT, bool E> MH_HD JM3<NT, E> jmul(const JM3<NT, E>& A, const JM3<NT, E>& B) {
    JM3<NT, E> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.a[i + 3 * j] = (A.a[i] * B.a[3 * j] + A.a[i + 3] * B.a[1 + 3 * j]) + A.a[i + 6] * B.a[2 + 3 * j];
    return r;
}
template <int NT, bool E> MH_HD JM3<NT, E> jmulABt(const JM3<NT, E>& A, const JM3<NT, E>& B) {
    JM3<NT, E> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.a[i + 3 * j] = (A.a[i] * B.a[j] + A.a[i + 3] * B.a[j + 3]) + A.a[i + 6] * B.a[j + 6];
    return r;
}
// A ∘₁ B with a constant B
template <int NT, bool E> MH_HD JM3<NT, E> jmul_const(const JM3<NT, E>& A, const m3& B) {
    JM3<NT, E> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.a[i + 3 * j] = (at(B, 0, j) * A.a[i] + at(B, 1, j) * A.a[i + 3]) + at(B, 2, j) * A.a[i + 6];
    return r;
}
template <int NT, bool E> MH_HD JV3<NT, E> jmulT(const JM3<NT, E>& A, const JV3<NT, E>& v) {   // Aᵀ v
    JV3<NT, E> r;
#pragma unroll
    for (int j = 0; j < 3; ++j) r.a[j] = (A.a[3 * j] * v.a[0] + A.a[3 * j + 1] * v.a[1]) + A.a[3 * j + 2] * v.a[2];
    return r;
}
template <int NT, bool E> MH_HD JV3<NT, E> jmulv(const JM3<NT, E>& A, const JV3<NT, E>& v) {   // A v
    JV3<NT, E> r;
#pragma unroll
    for (int i = 0; i < 3; ++i) r.a[i] = (A.a[i] * v.a[0] + A.a[i + 3] * v.a[1]) + A.a[i + 6] * v.a[2];
    return r;
}

// ------------------------------------------------------------ element dynamics
// Per element, what the loads need from the kinematics jets.
template <bool E> struct DynKin {
    Jet<1, E> r0[9];          // rₛₘ
    Jet<1, E> r1[9], r2[9];   // ṙₛₘ, r̈ₛₘ
    Jet<1, E> xd[4][3];       // ẋ at the Gauss points
    Jet<1, E> xdd[4][3];      // ẍ at the Gauss points
    double xe[4][3];          // ∂x/∂ε at the Gauss points (a column of the Jacobian of x w.r.t. X₀)
};
template <int NT, bool E> MH_HD Jet<1, E> jslice(const Jet<NT, E>& a, int n) {
    Jet<1, E> r;
    r.v[0] = n < NT ? a.v[n] : 0.;
    if constexpr (E) r.e[0] = n < NT ? a.e[n] : 0.;
    else r.e[0] = 0.;
    return r;
}

// The rotation chain (v₁,v₂) ↦ (rₛₘ, Δv) of BeamElement.jl:192-208 on time jets.
// Neither R₂ = Rodrigues(v₂) nor the relative rotation M = R₂R₁ᵀ is held whole: φ = Rodrigues⁻¹(M)
// reads only tr M = Σₖ R₁eₖ·R₂eₖ and the axial vector of M − Mᵀ, Σₖ R₁eₖ × R₂eₖ, so R₂ is
// formed a column at a time; rₛₘ = Rodrigues(Δv)·(R₁ rₘ) applies Rodrigues(Δv) to one column
// of R₁rₘ at a time, w ↦ w + a Δv×w + b (Δv (Δv·w) − |Δv|² w). This keeps the live jets
// of the chain to about 21 instead of the 27 of three whole matrices.
template <int NT, bool E, bool XSQ = false>
MH_HD void dyn_rotation_chain(const BeamGeo& G, const JV3<NT, E>& v1, const JV3<NT, E>& v2, JM3<NT, E>& rsm,
                              JV3<NT, E>& dv) {
    using J = Jet<NT, E>;
    JM3<NT, E> R1 = jrodrigues<NT, E, XSQ>(v1);
    J tr, ax[3];
    {
        J a2, b2;
        jrodrigues_coef<NT, E, XSQ>(v2, a2, b2);
        auto add = [&](const JV3<NT, E>& c2, int k, bool first) {
            const J* c1 = &R1.a[3 * k];
            J d = (c1[0] * c2.a[0] + c1[1] * c2.a[1]) + c1[2] * c2.a[2];
            J x = c1[1] * c2.a[2] - c1[2] * c2.a[1];
            J y = c1[2] * c2.a[0] - c1[0] * c2.a[2];
            J z = c1[0] * c2.a[1] - c1[1] * c2.a[0];
            if (first) { tr = d; ax[0] = x; ax[1] = y; ax[2] = z; }
            else { tr = tr + d; ax[0] = ax[0] + x; ax[1] = ax[1] + y; ax[2] = ax[2] + z; }
        };
        add(jrodrigues_col<0>(v2, a2, b2), 0, true);
        add(jrodrigues_col<1>(v2, a2, b2), 1, false);
        add(jrodrigues_col<2>(v2, a2, b2), 2, false);
    }
    // Rodrigues⁻¹(M) = spin⁻¹(M)/scac((tr M - 1)/2) (Rotations.jl:101), Δv = φ/2
    J rs = jrecip(jscac(0.5 * (tr + (-1.))));
#pragma unroll
    for (int c = 0; c < 3; ++c) dv.a[c] = 0.5 * ((0.5 * ax[c]) * rs);
    JM3<NT, E> R1m = jmul_const(R1, G.rm);
    J ah, bh;
    jrodrigues_coef<NT, E, XSQ>(dv, ah, bh);
    const J& x = dv.a[0]; const J& y = dv.a[1]; const J& z = dv.a[2];
    J n2 = (x * x + y * y) + z * z;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const J* w = &R1m.a[3 * j];
        J cx = y * w[2] - z * w[1], cy = z * w[0] - x * w[2], cz = x * w[1] - y * w[0];
        J s = (x * w[0] + y * w[1]) + z * w[2];
        rsm.a[3 * j] = (w[0] + ah * cx) + bh * (x * s - n2 * w[0]);
        rsm.a[3 * j + 1] = (w[1] + ah * cy) + bh * (y * s - n2 * w[1]);
        rsm.a[3 * j + 2] = (w[2] + ah * cz) + bh * (z * s - n2 * w[2]);
    }
}
// A time jet with no ε part as one whose ε part is zero.
template <int NT, bool E> MH_HD Jet<NT, E> jlift(const Jet<NT, false>& a) {
    Jet<NT, E> r = jzero<NT, E>();
#pragma unroll
    for (int n = 0; n < NT; ++n) r.v[n] = a.v[n];
    return r;
}

// Kinematics (BeamElement.jl:176-208) on time jets. Xj: the 12 element dofs as jets.
// ROTE = false: the ε seed is on a translation dof, so the rotation chain carries no ε part
// and is evaluated on plain time jets (about half the work of a direction).
template <int NT, bool E, bool ROTE = true, bool XSQ = false>
MH_HD void dyn_kinematics(const BeamGeo& G, const Jet<NT, E>* Xj, DynKin<E>& k) {
    using J = Jet<NT, E>;
    const double L = G.L;
    JV3<NT, E> u1, u2;
#pragma unroll
    for (int c = 0; c < 3; ++c) { u1.a[c] = Xj[c]; u2.a[c] = Xj[6 + c]; }
    JM3<NT, E> rsm;
    JV3<NT, E> dv;
    if constexpr (ROTE || !E) {
        JV3<NT, E> v1, v2;
#pragma unroll
        for (int c = 0; c < 3; ++c) { v1.a[c] = Xj[3 + c]; v2.a[c] = Xj[9 + c]; }
        dyn_rotation_chain<NT, E, XSQ>(G, v1, v2, rsm, dv);
    } else {
        JV3<NT, false> w1, w2;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int n = 0; n < NT; ++n) { w1.a[c].v[n] = Xj[3 + c].v[n]; w2.a[c].v[n] = Xj[9 + c].v[n]; }
#pragma unroll
        for (int c = 0; c < 3; ++c) { w1.a[c].e[0] = 0.; w2.a[c].e[0] = 0.; }
        JM3<NT, false> rf;
        JV3<NT, false> df;
        dyn_rotation_chain<NT, false, XSQ>(G, w1, w2, rf, df);
#pragma unroll
        for (int i = 0; i < 9; ++i) rsm.a[i] = jlift<NT, E>(rf.a[i]);
#pragma unroll
        for (int c = 0; c < 3; ++c) dv.a[c] = jlift<NT, E>(df.a[c]);
    }
    // uₗ₂ = rₛₘᵀ(u₂ + tgₘ/2 - cₛ) - tgₑ/2,  vₗ₂ = rₛₘᵀ Δv
    JV3<NT, E> dd, cs;
    const double tg[3] = {G.tgm.x, G.tgm.y, G.tgm.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        cs.a[c] = 0.5 * (u1.a[c] + u2.a[c]);
        dd.a[c] = (u2.a[c] + tg[c] * 0.5) - cs.a[c];
    }
    JV3<NT, E> ul = jmulT(rsm, dd), vl = jmulT(rsm, dv);
    ul.a[0] = ul.a[0] + (-(L * 0.5));
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const double z = G.zgp[g];
        const double ya = 2. * z, yu = -4. * (z * z * z) + 3. * z, yv = (z * z - 1. / 4.) * L;
        JV3<NT, E> p;
        p.a[0] = ya * ul.a[0] + L * z;
        p.a[1] = yu * ul.a[1] + yv * vl.a[2];
        p.a[2] = yu * ul.a[2] - yv * vl.a[1];
        JV3<NT, E> x = jmulv(rsm, p);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            J xc = x.a[c] + cs.a[c];
            k.xd[g][c] = jslice(xc, 1);
            k.xdd[g][c] = jslice(xc, 2);
            if constexpr (E) k.xe[g][c] = xc.e[0];
            else k.xe[g][c] = 0.;
        }
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) { k.r0[i] = jslice(rsm.a[i], 0); k.r1[i] = jslice(rsm.a[i], 1); k.r2[i] = jslice(rsm.a[i], 2); }
}

struct BeamDynSection {
    double mu, iota1, w;
    double Ca[3], Cl[3], Cq[3];
};

// External loads of the resultants (BeamElement.jl:28-64): per Gauss point the force
// fₑ (fw, inertia, added mass, linear and quadratic damping, minus U) and the
// element's roll-inertia couple mₑ (the same at every Gauss point).
template <bool E> MH_HD void dyn_loads(const DynKin<E>& k, const BeamDynSection& S, const double* U, Jet<1, E> (*fe)[3],
                                       Jet<1, E>* me) {
    using J = Jet<1, E>;
    const J* r0 = k.r0;
    for (int g = 0; g < 4; ++g) {
        J xl1[3], xl2[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {   // x ∘₁ r₀ = r₀ᵀx
            xl1[j] = (k.xd[g][0] * r0[3 * j] + k.xd[g][1] * r0[3 * j + 1]) + k.xd[g][2] * r0[3 * j + 2];
            xl2[j] = (k.xdd[g][0] * r0[3 * j] + k.xdd[g][1] * r0[3 * j + 1]) + k.xdd[g][2] * r0[3 * j + 2];
        }
        J fal[3], fll[3], fql[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            fal[j] = S.Ca[j] * xl2[j];
            fll[j] = S.Cl[j] * xl1[j];
            J q = S.Cq[j] * (xl1[j] * xl1[j]);
            fql[j] = xl1[j].v[0] < 0. ? -q : q;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            J fa = (r0[i] * fal[0] + r0[i + 3] * fal[1]) + r0[i + 6] * fal[2];
            J fl = (r0[i] * fll[0] + r0[i + 3] * fll[1]) + r0[i + 6] * fll[2];
            J fq = (r0[i] * fql[0] + r0[i + 3] * fql[1]) + r0[i + 6] * fql[2];
            J f = (((S.mu * k.xdd[g][i] + fa) + fl) + fq) + (i == 2 ? S.w : 0.);
            fe[g][i] = f + (-U[i]);
        }
    }
    // vᵢ₂ = spin⁻¹(ṙᵀṙ + r₀ᵀr̈) (Rotations.jl:177-182); only its first component is used
    J a21 = ((k.r1[6] * k.r1[3] + k.r1[7] * k.r1[4]) + k.r1[8] * k.r1[5]) +
            ((r0[6] * k.r2[3] + r0[7] * k.r2[4]) + r0[8] * k.r2[5]);
    J a12 = ((k.r1[3] * k.r1[6] + k.r1[4] * k.r1[7]) + k.r1[5] * k.r1[8]) +
            ((r0[3] * k.r2[6] + r0[4] * k.r2[7]) + r0[5] * k.r2[8]);
    J m1 = S.iota1 * (0.5 * (a21 - a12));
#pragma unroll
    for (int i = 0; i < 3; ++i) me[i] = r0[i] * m1;
}

// Left-Jacobian inverse direction of the rotation vector of rₛₘ: ∂vₛₘ/∂ε with
// vₛₘ = Rodrigues⁻¹(rₛₘ) (BeamElement.jl:203).
MH_HD v3 dyn_vsm_e(const Jet<1, true>* r0) {
    JM3<1, true> m;
#pragma unroll
    for (int i = 0; i < 9; ++i) m.a[i] = r0[i];
    JV3<1, true> v = jrodrigues_inv(m);
    return mk(v.a[0].e[0], v.a[1].e[0], v.a[2].e[0]);
}
// Rotation vector of a rotation matrix (value only).
MH_HD v3 rotvec(const m3& r) {
    double sc = scac((at(r, 0, 0) + at(r, 1, 1) + at(r, 2, 2) - 1.) / 2.);
    return mk((at(r, 2, 1) - at(r, 1, 2)) / 2. / sc, (at(r, 0, 2) - at(r, 2, 0)) / 2. / sc, (at(r, 1, 0) - at(r, 0, 1)) / 2. / sc);
}

// The element dofs of the solver's trial path as jets.
//   dir 0..11: δX along dof `dir`, with weights (w0·s, a₁s, b₁s), s = scale.X[dir]
//              (w0 = 1 for the solver; w0 = 0 isolates the X′ or X″ partials);
//   dir 12:    δr of :step, (0, a, b) with a = a₂X₁ + a₃X₂, b = b₂X₁ + b₃X₂ per dof;
//   dir 13:    X₀ + ε·lam (a directional derivative along lam, 12 values);
//   otherwise: no ε seed.
// X: ND rows of 12; c = (a₁,a₂,a₃,b₁,b₂,b₃).
template <int NT, bool E> MH_HD void dyn_seed(const double (*X)[12], const double* scale, const double* c, int dir,
                                              Jet<NT, E>* Xj, double w0 = 1., const double* lam = nullptr) {
#pragma unroll
    for (int i = 0; i < 12; ++i) {
        Jet<NT, E> q = jzero<NT, E>();
#pragma unroll
        for (int n = 0; n < NT; ++n) q.v[n] = X[n][i];
        if constexpr (E) {
            if (dir == 13) {
                q.e[0] = lam[i];
            } else if (dir == i) {
                q.e[0] = w0 * scale[i];
                if constexpr (NT > 1) q.e[1] = c[0] * scale[i];
                if constexpr (NT > 2) q.e[2] = c[3] * scale[i];
            } else if (dir == 12) {
                if constexpr (NT == 2) q.e[1] = c[1] * X[1][i];
                if constexpr (NT == 3) {
                    q.e[1] = c[1] * X[1][i] + c[2] * X[2][i];
                    q.e[2] = c[4] * X[1][i] + c[5] * X[2][i];
                }
            }
        }
        Xj[i] = q;
    }
}

// Value pass: the frozen loads of the dynamic element, eg[gp] = dL_gp·fₑ(gp) and
// mtot = Σ_gp dL_gp·mₑ, for beam_static<true>. X: ND rows of 12 element dofs.
// kb (optional, stride ks): the ε-free kinematics that dyn_translation_direction reads, the
// kDynBase numbers ∂ⁿrₛₘ(i,j)/∂tⁿ at n·9 + i + 3j (n < 3) and ẋ_gp,i at 27 + 3·gp + i.
constexpr int kDynBase = 39;
template <int ND> MH_HD void dyn_frozen_loads(const BeamGeo& G, const BeamDynSection& D, const double (*X)[12],
                                              const double* U, v3* eg, v3& mtot, double* kb = nullptr,
                                              int64_t ks = 1) {
    Jet<ND, false> Xj[12];
    dyn_seed<ND, false>(X, nullptr, nullptr, -1, Xj);
    DynKin<false> k;
    dyn_kinematics<ND, false>(G, Xj, k);
    if (kb) {
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            kb[q * ks] = k.r0[q].v[0];
            kb[(9 + q) * ks] = k.r1[q].v[0];
            kb[(18 + q) * ks] = k.r2[q].v[0];
        }
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int i = 0; i < 3; ++i) kb[(27 + 3 * g + i) * ks] = k.xd[g][i].v[0];
    }
    Jet<1, false> fe[4][3], me[3];
    dyn_loads<false>(k, D, U, fe, me);
    mtot = mk(0., 0., 0.);
    for (int g = 0; g < 4; ++g) {
        eg[g] = G.dL[g] * mk(fe[g][0].v[0], fe[g][1].v[0], fe[g][2].v[0]);
        mtot = mtot + G.dL[g] * mk(me[0].v[0], me[1].v[0], me[2].v[0]);
    }
}

// Requestables of the element (BeamElement.jl:28-64,151-174) at X (ND time levels), what
// getresult extracts (src/Output.jl:131-181). getresult evaluates the element on plain
// values, so the time jets carry no solver partials and every square is exact (no
// power-rule artefact). ☼ε and ♢κ = (vₗ₂[1], vₗ₂[3], −vₗ₂[2])·2/L
// with their ND time levels, then per Gauss point ☼x and ☼κgp (ND levels each) and the
// resultants' ☼fᵢ, ☼mᵢ, ☼fₑ (as returned by `resultants`, before the U-load is subtracted)
// and ☼mₑ. Layout: ε[ND], κ[ND][3], then per Gauss point x[ND][3], κgp[ND][3], fᵢ, mᵢ[3],
// fₑ[3], mₑ[3] (beam_nres(ND) numbers).
constexpr int beam_nres(int ND) { return 4 * ND + 4 * (6 * ND + 10); }
template <int ND> MH_HD void dyn_requestables(const BeamGeo& G, const BeamSection& S, const BeamDynSection& D,
                                              const double (*X)[12], double* __restrict__ out) {
    using J = Jet<ND, false>;
    J Xj[12];
    dyn_seed<ND, false>(X, nullptr, nullptr, -1, Xj);
    const double L = G.L;
    JV3<ND, false> u1, u2, v1, v2;
#pragma unroll
    for (int c = 0; c < 3; ++c) { u1.a[c] = Xj[c]; u2.a[c] = Xj[6 + c]; v1.a[c] = Xj[3 + c]; v2.a[c] = Xj[9 + c]; }
    JM3<ND, false> rsm;
    JV3<ND, false> dv;
    dyn_rotation_chain<ND, false, true>(G, v1, v2, rsm, dv);
    JV3<ND, false> dd, cs;
    const double tg[3] = {G.tgm.x, G.tgm.y, G.tgm.z};
    const double cm[3] = {G.cm.x, G.cm.y, G.cm.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        cs.a[c] = 0.5 * (u1.a[c] + u2.a[c]);
        dd.a[c] = (u2.a[c] + tg[c] * 0.5) - cs.a[c];
    }
    JV3<ND, false> ul = jmulT(rsm, dd), vl = jmulT(rsm, dv);
    ul.a[0] = ul.a[0] + (-(L * 0.5));
    // ε = √((uₗ₂[1]+L/2)² + uₗ₂[2]² + uₗ₂[3]²)·2/L − 1
    const J a0 = ul.a[0] + 0.5 * L;
    const J eps = (2. / L) * jsqrt((a0 * a0 + ul.a[1] * ul.a[1]) + ul.a[2] * ul.a[2]) + (-1.);
    int q = 0;
#pragma unroll
    for (int d = 0; d < ND; ++d) out[q++] = eps.v[d];
    const J km[3] = {(2. / L) * vl.a[0], (2. / L) * vl.a[2], (2. / L) * (-vl.a[1])};
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
        for (int c = 0; c < 3; ++c) out[q++] = km[c].v[d];
    // loads (values) from the time jets of the Gauss points
    DynKin<false> k;
    dyn_kinematics<ND, false, true, true>(G, Xj, k);
    Jet<1, false> fe[4][3], me[3];
    const double U0[3] = {0., 0., 0.};
    dyn_loads<false>(k, D, U0, fe, me);
    const double ka = 2. / L, kv = 2. / L;
    for (int g = 0; g < 4; ++g) {
        const double z = G.zgp[g];
        const double ya = 2. * z, yu = -4. * (z * z * z) + 3. * z, yv = (z * z - 1. / 4.) * L;
        const double ku = (-24. * z) / (L * L);
        JV3<ND, false> p;
        p.a[0] = ya * ul.a[0] + L * z;
        p.a[1] = yu * ul.a[1] + yv * vl.a[2];
        p.a[2] = yu * ul.a[2] - yv * vl.a[1];
        JV3<ND, false> x = jmulv(rsm, p);
        J kg[3] = {ka * vl.a[0], ku * ul.a[1] + kv * vl.a[2], ku * ul.a[2] - kv * vl.a[1]};
#pragma unroll
        for (int d = 0; d < ND; ++d)
#pragma unroll
            for (int c = 0; c < 3; ++c) out[q++] = d == 0 ? (x.a[c].v[0] + cs.a[c].v[0]) + cm[c] : x.a[c].v[d] + cs.a[c].v[d];
#pragma unroll
        for (int d = 0; d < ND; ++d)
#pragma unroll
            for (int c = 0; c < 3; ++c) out[q++] = kg[c].v[d];
        out[q++] = S.EA * eps.v[0];
        out[q++] = S.GJ * kg[0].v[0];
        out[q++] = S.EI3 * kg[1].v[0];
        out[q++] = S.EI2 * kg[2].v[0];
#pragma unroll
        for (int c = 0; c < 3; ++c) out[q++] = fe[g][c].v[0];
#pragma unroll
        for (int c = 0; c < 3; ++c) out[q++] = me[c].v[0];
    }
}

// The strain-gauge requestables εₐₓ = ε and κ = ♢κ (BeamElement.jl:155,172;
// StrainGaugeOnBeamElement.jl:73) on jets whose "time" level carries ∂/∂X_i and whose ε
// level carries ∂/∂X_j: with the seeds X_k = X₀_k + t·δ_ki + ε·δ_kj, v[0] is the value,
// v[1] = ∂/∂X_i, e[0] = ∂/∂X_j and e[1] = ∂²/∂X_i∂X_j (exact squares: second derivatives
// at the value level, as revariate{2} forms them, have no power-rule artefact).
MH_HD void gauge_jets(const BeamGeo& G, const double* X0, int i, int j, Jet<2, true>* q) {
    using J = Jet<2, true>;
    J Xj[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        Xj[k].v[0] = X0[k]; Xj[k].v[1] = k == i ? 1. : 0.;
        Xj[k].e[0] = k == j ? 1. : 0.; Xj[k].e[1] = 0.;
    }
    const double L = G.L;
    JV3<2, true> u1, u2, v1, v2;
#pragma unroll
    for (int c = 0; c < 3; ++c) { u1.a[c] = Xj[c]; u2.a[c] = Xj[6 + c]; v1.a[c] = Xj[3 + c]; v2.a[c] = Xj[9 + c]; }
    JM3<2, true> rsm;
    JV3<2, true> dv;
    dyn_rotation_chain<2, true, true>(G, v1, v2, rsm, dv);
    JV3<2, true> dd;
    const double tg[3] = {G.tgm.x, G.tgm.y, G.tgm.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const J cs = 0.5 * (u1.a[c] + u2.a[c]);
        dd.a[c] = (u2.a[c] + tg[c] * 0.5) - cs;
    }
    JV3<2, true> ul = jmulT(rsm, dd), vl = jmulT(rsm, dv);
    ul.a[0] = ul.a[0] + (-(L * 0.5));
    const J a0 = ul.a[0] + 0.5 * L;
    q[0] = (2. / L) * jsqrt((a0 * a0 + ul.a[1] * ul.a[1]) + ul.a[2] * ul.a[2]) + (-1.);
    q[1] = (2. / L) * vl.a[0];
    q[2] = (2. / L) * vl.a[2];
    q[3] = (2. / L) * (-vl.a[1]);
}

// One trial direction (see dyn_seed): the ε-derivatives of the loads (DF, Dm) and,
// for a δX direction, the column of the Jacobians of x_gp and vₛₘ w.r.t. X₀ (Jx, Jv).
struct DynDir { double DF[4][3]; double Dm[3]; double Jx[4][3]; double Jv[3]; };
// ROTE = false only for δX along a translation dof (0..2, 6..8): rₛₘ and vₛₘ then carry no ε.
template <int ND, bool ROTE = true>
MH_HD void dyn_direction(const BeamGeo& G, const BeamDynSection& D, const double (*X)[12], const double* U,
                         const double* scale, const double* c, int dir, DynDir& o, double w0 = 1.,
                         const double* lam = nullptr) {
    Jet<ND, true> Xj[12];
    dyn_seed<ND, true>(X, scale, c, dir, Xj, w0, lam);
    DynKin<true> k;
    dyn_kinematics<ND, true, ROTE>(G, Xj, k);
    Jet<1, true> fe[4][3], me[3];
    dyn_loads<true>(k, D, U, fe, me);
    for (int g = 0; g < 4; ++g)
        for (int q = 0; q < 3; ++q) { o.DF[g][q] = fe[g][q].e[0]; o.Jx[g][q] = k.xe[g][q]; }
    for (int q = 0; q < 3; ++q) o.Dm[q] = me[q].e[0];
    if constexpr (ROTE) {
        v3 jv = dyn_vsm_e(k.r0);
        o.Jv[0] = jv.x; o.Jv[1] = jv.y; o.Jv[2] = jv.z;
    } else {
        o.Jv[0] = 0.; o.Jv[1] = 0.; o.Jv[2] = 0.;
    }
}
// The δX direction along a translation dof (dir 0..2, 6..8) in closed form, from the
// kinematics jets k of any direction of the same element (only their ε-free parts are read).
// Along a translation rₛₘ and vₛₘ carry no ε, so with the seed σ = (w0·s, a₁s, b₁s) on dof
// component c of node 1 (η = −½) or node 2 (η = +½):
//   x_gp,ε = (η·M_gp(:,c) + ½·e_c) ⊛ σ,   M_gp = rₛₘ·diag(ya, yu, yu)·rₛₘᵀ   (Leibniz products
// of time jets), the loads vary through ẋ and ẍ alone, and the roll couple mₑ not at all:
//   DFᵢ = μ·ẍᵢ,ε + Σⱼ r₀(i,j)·(Caⱼ·(r₀ᵀẍ_ε)ⱼ + (Clⱼ + 2Cqⱼ|(r₀ᵀẋ)ⱼ|)·(r₀ᵀẋ_ε)ⱼ).
// The same numbers as dyn_direction<ND, false>(…, dir, …) up to rounding.
template <int ND>
MH_HD void dyn_translation_direction(const BeamGeo& G, const BeamDynSection& D, const double* kb, int64_t ks,
                                     const double* scale, const double* c, int dir, DynDir& o, double w0 = 1.) {
    const bool node2 = dir >= 6;
    const int cc = node2 ? dir - 6 : dir;
    const double eta = node2 ? 0.5 : -0.5;
    const double s = scale[dir];
    const double sg0 = w0 * s, sg1 = ND > 1 ? c[0] * s : 0., sg2 = ND > 2 ? c[3] * s : 0.;
    // R[n][i + 3j]: ∂ⁿrₛₘ(i,j)/∂tⁿ
    double R[3][9];
#pragma unroll
    for (int q = 0; q < 9; ++q) { R[0][q] = kb[q * ks]; R[1][q] = kb[(9 + q) * ks]; R[2][q] = kb[(18 + q) * ks]; }
    // Wᵢⱼ = rₛₘ(i,j) ⊛ rₛₘ(c,j): the gp-independent products of M_gp(i,c) = Σⱼ Yⱼ Wᵢⱼ
    // row cc of R by selects (an index known only at run time would put R in scratch memory)
    double Rc[3][3];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int j = 0; j < 3; ++j) Rc[m][j] = cc == 0 ? R[m][3 * j] : (cc == 1 ? R[m][3 * j + 1] : R[m][3 * j + 2]);
    double W[3][3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int a = i + 3 * j;
            W[i][j][0] = R[0][a] * Rc[0][j];
            W[i][j][1] = R[1][a] * Rc[0][j] + R[0][a] * Rc[1][j];
            W[i][j][2] = (R[2][a] * Rc[0][j] + 2. * (R[1][a] * Rc[1][j])) + R[0][a] * Rc[2][j];
        }
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const double z = G.zgp[g];
        const double Y[3] = {2. * z, -4. * (z * z * z) + 3. * z, -4. * (z * z * z) + 3. * z};
        double xd[3], xdd[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double m[3];
#pragma unroll
            for (int n = 0; n < 3; ++n) m[n] = eta * ((Y[0] * W[i][0][n] + Y[1] * W[i][1][n]) + Y[2] * W[i][2][n]);
            if (i == cc) m[0] = m[0] + 0.5;
            o.Jx[g][i] = m[0] * sg0;
            xd[i] = ND > 1 ? m[1] * sg0 + m[0] * sg1 : 0.;
            xdd[i] = ND > 2 ? (m[2] * sg0 + 2. * (m[1] * sg1)) + m[0] * sg2 : 0.;
        }
        double fl[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double l1 = (kb[(27 + 3 * g) * ks] * R[0][3 * j] + kb[(28 + 3 * g) * ks] * R[0][3 * j + 1]) +
                              kb[(29 + 3 * g) * ks] * R[0][3 * j + 2];
            const double l1e = (xd[0] * R[0][3 * j] + xd[1] * R[0][3 * j + 1]) + xd[2] * R[0][3 * j + 2];
            const double l2e = (xdd[0] * R[0][3 * j] + xdd[1] * R[0][3 * j + 1]) + xdd[2] * R[0][3 * j + 2];
            fl[j] = D.Ca[j] * l2e + (D.Cl[j] + 2. * (D.Cq[j] * fabs(l1))) * l1e;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
            o.DF[g][i] = D.mu * xdd[i] + ((R[0][i] * fl[0] + R[0][i + 3] * fl[1]) + R[0][i + 6] * fl[2]);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) { o.Dm[q] = 0.; o.Jv[q] = 0.; }
}
// Row i of the load-variation part of the tangent along direction d:
// Σ_gp dL_gp·(∂x_gp/∂X₀ᵢ · DF_gp + ∂vₛₘ/∂X₀ᵢ · Dm)   (Jx_i, Jv_i: column i of the Jacobians,
// from the δX direction i, so already × scale.X[i] — the scaling of row i of R)
MH_HD double dyn_row(const BeamGeo& G, const double (*Jx_i)[3], const double* Jv_i, const DynDir& d) {
    double s = 0.;
    for (int g = 0; g < 4; ++g) {
        double t = (Jx_i[g][0] * d.DF[g][0] + Jx_i[g][1] * d.DF[g][1]) + Jx_i[g][2] * d.DF[g][2];
        t = t + ((Jv_i[0] * d.Dm[0] + Jv_i[1] * d.Dm[1]) + Jv_i[2] * d.Dm[2]);
        s = s + t * G.dL[g];
    }
    return s;
}

// ---- adjoint (vector-Jacobian products) of the residual ------------------------
// For L = Λ·R the load part of ∂L/∂X⁽ᵈ⁾ₖ is Σ_gp dL (y_gp·∂fₑ/∂X⁽ᵈ⁾ₖ + y_v·∂mₑ/∂X⁽ᵈ⁾ₖ) with
// y_gp = J_x,gp Λ and y_v = J_v Λ, the derivatives of x_gp and vₛₘ along Λ.
struct DynY { double y[4][3]; double yv[3]; };
MH_HD void dyn_y(const BeamGeo& G, const double (*X)[12], const double* lam, DynY& o) {
    Jet<1, true> Xj[12];
    dyn_seed<1, true>(X, nullptr, nullptr, 13, Xj, 0., lam);
    DynKin<true> k;
    dyn_kinematics<1, true>(G, Xj, k);
    for (int g = 0; g < 4; ++g)
        for (int q = 0; q < 3; ++q) o.y[g][q] = k.xe[g][q];
    v3 jv = dyn_vsm_e(k.r0);
    o.yv[0] = jv.x; o.yv[1] = jv.y; o.yv[2] = jv.z;
}
MH_HD double dyn_ycontract(const BeamGeo& G, const DynY& y, const DynDir& d) {
    double s = 0.;
    for (int g = 0; g < 4; ++g) {
        double t = (y.y[g][0] * d.DF[g][0] + y.y[g][1] * d.DF[g][1]) + y.y[g][2] * d.DF[g][2];
        t = t + ((y.yv[0] * d.Dm[0] + y.yv[1] * d.Dm[1]) + y.yv[2] * d.Dm[2]);
        s = s + t * G.dL[g];
    }
    return s;
}
// ∂L/∂U (U-dofs: fₑ - U, per length): -Σ_gp dL y_gp
MH_HD v3 dyn_vjp_u(const BeamGeo& G, const DynY& y) {
    v3 s = mk(0., 0., 0.);
    for (int g = 0; g < 4; ++g) s = s + G.dL[g] * mk(y.y[g][0], y.y[g][1], y.y[g][2]);
    return mk(-s.x, -s.y, -s.z);
}
// Whole dynamic element on one thread (host checks): gX[der][k] = Σ_i Λᵢ ∂Rᵢ/∂X⁽ᵈᵉʳ⁾ₖ (unscaled)
// for der < ND, gU = Λ·∂R/∂U.
template <int ND> MH_HD int dyn_vjp_element(const BeamGeo& G, const BeamSection& S, const BeamDynSection& D,
                                            const double (*X)[12], const double* U, const double* lam, double (*gX)[12],
                                            double* gU) {
    v3 eg[4], mtot;
    dyn_frozen_loads<ND>(G, D, X, U, eg, mtot);
    BeamOut out;
    beam_static<true>(G, S, mk(0., 0., 0.), X[0], out, eg, mtot);
    DynY y;
    dyn_y(G, X, lam, y);
    const double one[12] = {1., 1., 1., 1., 1., 1., 1., 1., 1., 1., 1., 1.};
    int nan = out.nan;
    for (int der = 0; der < ND; ++der) {
        double c[6] = {der == 1 ? 1. : 0., 0., 0., der == 2 ? 1. : 0., 0., 0.};
        for (int k = 0; k < 12; ++k) {
            DynDir d;
            dyn_direction<ND>(G, D, X, U, one, c, k, d, der == 0 ? 1. : 0.);
            double g = dyn_ycontract(G, y, d);
            if (der == 0)
                for (int i = 0; i < 12; ++i) g = g + compact_entry(out.Kc, k, i) * lam[i];
            gX[der][k] = g;
            nan |= g != g;
        }
    }
    v3 gu = dyn_vjp_u(G, y);
    gU[0] = gu.x; gU[1] = gu.y; gU[2] = gu.z;
    return nan;
}

// Whole element on one thread (host checks): scaled R, K = ∂R/∂δX (12×12,
// column-major) and, for :step, dRr = ∂R/∂δr. c = Newmark (a₁,a₂,a₃,b₁,b₂,b₃).
template <int ND> MH_HD int dyn_element(const BeamGeo& G, const BeamSection& S, const BeamDynSection& D,
                                        const double (*X)[12], const double* U, const double* scale, const double* c,
                                        double* R, double* K, double* dRr) {
    v3 eg[4], mtot;
    dyn_frozen_loads<ND>(G, D, X, U, eg, mtot);
    BeamOut out;
    beam_static<true>(G, S, mk(0., 0., 0.), X[0], out, eg, mtot);
    DynDir dir[12];
    for (int k = 0; k < 12; ++k) dyn_direction<ND>(G, D, X, U, scale, c, k, dir[k]);
    for (int i = 0; i < 12; ++i) R[i] = out.R[i] * scale[i];
    for (int k = 0; k < 12; ++k)
        for (int i = 0; i < 12; ++i)
            K[i + 12 * k] = compact_entry(out.Kc, i, k) * (scale[i] * scale[k]) + dyn_row(G, dir[i].Jx, dir[i].Jv, dir[k]);
    if (dRr) {
        DynDir r;
        dyn_direction<ND>(G, D, X, U, scale, c, 12, r);
        for (int i = 0; i < 12; ++i) dRr[i] = dyn_row(G, dir[i].Jx, dir[i].Jv, r);
    }
    int nan = out.nan;
    for (int i = 0; i < 144; ++i) nan |= K[i] != K[i];
    return nan;
}

}  // namespace mh
