// Analytic FP64 residual and tangent of the static corotated EulerBeam3D.
//
// Reference: toolbox/BeamElement.jl:122-208 (constructor, residual, kinematics,
// corotated), toolbox/Rotations.jl:13-135 (sinc1, scac, Rodrigues, Rodrigues⁻¹).
// Muscade differentiates the element with nested forward-mode duals
// (src/Adiff.jl, src/Taylor.jl). This file computes the same R = ∂Π/∂X and
// K = ∂R/∂X in closed form instead:
//
//   * the static residual is the gradient of Π(X) = Σ_gp dL·(½EA ε² +
//     ½ mᵢ·κ + fₑ·x) (fₑ constant in the static case), so K is its Hessian;
//   * Π depends on the 12 element dofs through the rotation core
//     (v₁,v₂) ↦ (rₛₘ, Δv) and through the translations (d, cₛₘ);
//   * first derivatives of the rotation core use spatial spins
//     (δQ Qᵀ = spin(ω δp)); second derivatives use the symmetric second-order
//     spin W (δ²Q Qᵀ = spin(W) + ½{spin ω_a, spin ω_b}) with composition
//     rules for Exp, Log, products and transposes. The second-order part is
//     never stored: it is contracted with adjoint vectors in a reverse sweep
//     and accumulated into the 6×6 rotation block of K.
//
// Everything is `__host__ __device__` so the same code can be exercised on the
// host in unit harnesses; the product path runs it on the GPU only.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace mh {

#define MH_HD __host__ __device__ __forceinline__

struct v3 { double x, y, z; };
struct m3 { double a[9]; };   // column-major: a[i + 3*j] = M(i,j)

MH_HD v3 mk(double x, double y, double z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
MH_HD v3 operator+(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
MH_HD v3 operator-(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
MH_HD v3 operator*(double s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
MH_HD double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MH_HD v3 cross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
MH_HD double& at(m3& m, int i, int j) { return m.a[i + 3 * j]; }
MH_HD double at(const m3& m, int i, int j) { return m.a[i + 3 * j]; }
MH_HD v3 col(const m3& m, int j) { return mk(m.a[3 * j], m.a[3 * j + 1], m.a[3 * j + 2]); }
MH_HD v3 mul(const m3& m, v3 v) {
    return mk(at(m, 0, 0) * v.x + at(m, 0, 1) * v.y + at(m, 0, 2) * v.z,
              at(m, 1, 0) * v.x + at(m, 1, 1) * v.y + at(m, 1, 2) * v.z,
              at(m, 2, 0) * v.x + at(m, 2, 1) * v.y + at(m, 2, 2) * v.z);
}
MH_HD v3 mulT(const m3& m, v3 v) {   // Mᵀ v
    return mk(at(m, 0, 0) * v.x + at(m, 1, 0) * v.y + at(m, 2, 0) * v.z,
              at(m, 0, 1) * v.x + at(m, 1, 1) * v.y + at(m, 2, 1) * v.z,
              at(m, 0, 2) * v.x + at(m, 1, 2) * v.y + at(m, 2, 2) * v.z);
}
MH_HD m3 mul(const m3& A, const m3& B) {
    m3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.a[i + 3 * j] = at(A, i, 0) * at(B, 0, j) + at(A, i, 1) * at(B, 1, j) + at(A, i, 2) * at(B, 2, j);
    return r;
}
MH_HD m3 mulABt(const m3& A, const m3& B) {   // A Bᵀ
    m3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.a[i + 3 * j] = at(A, i, 0) * at(B, j, 0) + at(A, i, 1) * at(B, j, 1) + at(A, i, 2) * at(B, j, 2);
    return r;
}

// ------------------------------------------------------------ scalar kernels
// sinc1(x) = sin(x)/x with the small-argument series that Base.sinc uses
// (Rotations.jl:13): |x/π| < 1e-3 → 1 - x²/6 + x⁴/120.
MH_HD double sinpi_red(double y) {
    double x = fabs(y);
    double n = rint(2. * x);
    double rx = -0.5 * n + x;
    long q = (long)n & 3;
    double res;
    const double pi = 3.141592653589793;
    if (q == 0) res = sin(pi * rx);
    else if (q == 1) res = cos(pi * rx);
    else if (q == 2) res = 0. - sin(pi * rx);
    else res = 0. - cos(pi * rx);
    return __builtin_signbit(y) ? -res : res;
}
MH_HD double sinc1(double x) {
    const double pi = 3.141592653589793;
    double y = x / pi;
    if (fabs(y) < 1e-3) { double y2 = y * y; return 1. + y2 * (-(pi * pi) / 6. + y2 * (pi * pi * pi * pi / 120.)); }
    return sinpi_red(y) / (pi * y);
}
// scac(x) = sinc1(acos x), 5th-order series for |x-1| <= 1e-3 (Rotations.jl:61-68)
MH_HD double scac(double x) {
    double dx = x - 1.;
    if (fabs(dx) > 1e-3) return sinc1(acos(x));
    return 1. + dx * (1. / 3. + dx * (-2. / 90. + dx * (0.0052911879917544626 + dx * (-0.0016229317117234072 + dx * (0.0005625)))));
}

// Coefficients of the rotation-vector exponential and its derivatives, for θ = |v|:
//   a = sinθ/θ, b = (1-cosθ)/θ², c = (θ-sinθ)/θ³, β = b'(θ)/θ, γ = c'(θ)/θ,
//   e = (1 - a/(2b))/θ²  (coefficient of spin² in T⁻¹).
// a and b are evaluated like Rodrigues (Rotations.jl:131-135): a = sinc1(θ),
// b = sinc1(θ/2)²/2, with θ reset to 0 below 1e-14 like norm3 (Rotations.jl:106-112).
struct ExpCoef { double a, b, c, beta, gamma, e; };
MH_HD ExpCoef exp_coef(double th) {
    ExpCoef k;
    if (th < 1e-14) th = 0.;
    k.a = sinc1(th);
    double s = sinc1(th / 2.);
    k.b = s * s / 2.;
    double t2 = th * th;
    if (th < 0.1) {
        k.c = 1. / 6. + t2 * (-1. / 120. + t2 * (1. / 5040. + t2 * (-1. / 362880.)));
        k.beta = -1. / 12. + t2 * (1. / 180. + t2 * (-1. / 6720. + t2 * (1. / 453600.)));
        k.gamma = -1. / 60. + t2 * (1. / 1260. + t2 * (-1. / 60480. + t2 * (1. / 4989600.)));
        k.e = 1. / 12. + t2 * (1. / 720. + t2 * (1. / 30240. + t2 * (1. / 1209600.)));
    } else {
        k.c = (1. - k.a) / t2;
        k.beta = (k.a - 2. * k.b) / t2;
        k.gamma = (k.b - 3. * k.c) / t2;
        k.e = (1. - k.a / (2. * k.b)) / t2;
    }
    return k;
}
MH_HD double norm3(v3 v) {
    double n = sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    return n < 1e-14 ? 0. : n;
}
// Rodrigues(v) = I + a S + b S² (Rotations.jl:131-135, spin² per :114-122)
MH_HD m3 rodrigues(v3 v, const ExpCoef& k) {
    m3 r;
    double x = v.x, y = v.y, z = v.z;
    // S = spin(v); S² = v vᵀ - |v|² I
    double ab = x * y, ca = z * x, bc = y * z;
    double a2 = -x * x, b2 = -y * y, c2 = -z * z;
    at(r, 0, 0) = 1. + k.b * (c2 + b2);
    at(r, 1, 0) = k.a * z + k.b * ab;
    at(r, 2, 0) = -k.a * y + k.b * ca;
    at(r, 0, 1) = -k.a * z + k.b * ab;
    at(r, 1, 1) = 1. + k.b * (a2 + c2);
    at(r, 2, 1) = k.a * x + k.b * bc;
    at(r, 0, 2) = k.a * y + k.b * ca;
    at(r, 1, 2) = -k.a * x + k.b * bc;
    at(r, 2, 2) = 1. + k.b * (a2 + b2);
    return r;
}
// Left Jacobian T(v) = I + b S + c S²: δRod(v) = spin(T δv) Rod(v)
MH_HD m3 tangent_op(v3 v, const ExpCoef& k) {
    m3 r;
    double x = v.x, y = v.y, z = v.z;
    double xx = x * x, yy = y * y, zz = z * z;
    at(r, 0, 0) = 1. - k.c * (yy + zz);
    at(r, 1, 1) = 1. - k.c * (xx + zz);
    at(r, 2, 2) = 1. - k.c * (xx + yy);
    at(r, 0, 1) = -k.b * z + k.c * x * y;
    at(r, 1, 0) = k.b * z + k.c * x * y;
    at(r, 0, 2) = k.b * y + k.c * x * z;
    at(r, 2, 0) = -k.b * y + k.c * x * z;
    at(r, 1, 2) = -k.b * x + k.c * y * z;
    at(r, 2, 1) = k.b * x + k.c * y * z;
    return r;
}
// T(v)⁻¹ = I - ½ S + e S²
MH_HD m3 tangent_inv(v3 v, const ExpCoef& k) {
    m3 r;
    double x = v.x, y = v.y, z = v.z;
    double xx = x * x, yy = y * y, zz = z * z;
    at(r, 0, 0) = 1. - k.e * (yy + zz);
    at(r, 1, 1) = 1. - k.e * (xx + zz);
    at(r, 2, 2) = 1. - k.e * (xx + yy);
    at(r, 0, 1) = 0.5 * z + k.e * x * y;
    at(r, 1, 0) = -0.5 * z + k.e * x * y;
    at(r, 0, 2) = -0.5 * y + k.e * x * z;
    at(r, 2, 0) = 0.5 * y + k.e * x * z;
    at(r, 1, 2) = 0.5 * x + k.e * y * z;
    at(r, 2, 1) = -0.5 * x + k.e * y * z;
    return r;
}
// Symmetric 3×3 stored as (xx,yy,zz,xy,xz,yz).
struct s3 { double xx, yy, zz, xy, xz, yz; };
// sym part of the bilinear form (x,y) ↦ g·∂T(v)[x]y:
//   β sym(v (g×v)ᵀ) + γ sym(v (S²g)ᵀ) + c[(g·v)I - sym(g vᵀ)],  S²g = v(v·g) - g|v|²
MH_HD s3 dT_form(v3 v, v3 g, const ExpCoef& k) {
    v3 gv = cross(g, v);
    double vg = dot(v, g), vv = dot(v, v);
    v3 s2g = vg * v - vv * g;
    v3 p = k.beta * gv + k.gamma * s2g;   // B = v pᵀ + c[(g·v)I - g vᵀ]...
    s3 r;
    r.xx = v.x * p.x + k.c * (vg - g.x * v.x);
    r.yy = v.y * p.y + k.c * (vg - g.y * v.y);
    r.zz = v.z * p.z + k.c * (vg - g.z * v.z);
    r.xy = 0.5 * (v.x * p.y + p.x * v.y) - 0.5 * k.c * (g.x * v.y + v.x * g.y);
    r.xz = 0.5 * (v.x * p.z + p.x * v.z) - 0.5 * k.c * (g.x * v.z + v.x * g.z);
    r.yz = 0.5 * (v.y * p.z + p.y * v.z) - 0.5 * k.c * (g.y * v.z + v.y * g.z);
    return r;
}
MH_HD v3 smul(const s3& s, v3 x) {
    return mk(s.xx * x.x + s.xy * x.y + s.xz * x.z, s.xy * x.x + s.yy * x.y + s.yz * x.z, s.xz * x.x + s.yz * x.y + s.zz * x.z);
}

// ---------------------------------------------------------- element geometry
struct BeamGeo {
    v3 cm, tgm;
    m3 rm;        // columns t, n, b
    double L;
    double zgp[4], dL[4];
    int bad;      // orient2 nearly parallel to the element (BeamElement.jl:135)
};
MH_HD BeamGeo beam_geometry(v3 c1, v3 c2, v3 o2) {   // BeamElement.jl:122-148
    BeamGeo g;
    g.cm = mk((c1.x + c2.x) / 2., (c1.y + c2.y) / 2., (c1.z + c2.z) / 2.);
    g.tgm = c2 - c1;
    double L = sqrt(dot(g.tgm, g.tgm));
    v3 t = mk(g.tgm.x / L, g.tgm.y / L, g.tgm.z / L);
    double no = sqrt(dot(o2, o2));
    o2 = mk(o2.x / no, o2.y / no, o2.z / no);
    double d = dot(o2, t);
    v3 n = o2 - d * t;
    double nn = sqrt(dot(n, n));
    g.bad = !(nn > 1e-3);
    n = mk(n.x / nn, n.y / nn, n.z / nn);
    v3 b = cross(t, n);
    g.rm.a[0] = t.x; g.rm.a[1] = t.y; g.rm.a[2] = t.z;
    g.rm.a[3] = n.x; g.rm.a[4] = n.y; g.rm.a[5] = n.z;
    g.rm.a[6] = b.x; g.rm.a[7] = b.y; g.rm.a[8] = b.z;
    g.L = L;
    const double s30 = 5.477225575051661;      // √30
    const double s65 = 1.0954451150103321;     // √(6/5)
    g.dL[0] = L / 2. * (18. - s30) / 36.; g.dL[1] = L / 2. * (18. + s30) / 36.;
    g.dL[2] = g.dL[1]; g.dL[3] = g.dL[0];
    double z1 = 0.5 * sqrt(3. / 7. + 2. / 7. * s65), z2 = 0.5 * sqrt(3. / 7. - 2. / 7. * s65);
    g.zgp[0] = -z1; g.zgp[1] = -z2; g.zgp[2] = z2; g.zgp[3] = z1;
    return g;
}

struct BeamSection { double EA, EI2, EI3, GJ, w; };

// Output: R[12], and K as the upper triangle (incl. diagonal) in the order
// K[i][j], i<=j, row-major over the 78 entries; K is symmetric.
struct BeamOut { double R[12]; double K[78]; int nan; };
MH_HD int tri(int i, int j) { return i * 12 - (i * (i - 1)) / 2 + (j - i); }   // i <= j

// Static residual and tangent; fe = (0,0,w) - U is the external load per length.
MH_HD void beam_static(const BeamGeo& G, const BeamSection& S, v3 fe, const double* X, BeamOut& out) {
    const double L = G.L;
    v3 u1 = mk(X[0], X[1], X[2]), v1 = mk(X[3], X[4], X[5]);
    v3 u2 = mk(X[6], X[7], X[8]), v2 = mk(X[9], X[10], X[11]);
    // ---- rotation core ------------------------------------------------------
    ExpCoef k1 = exp_coef(norm3(v1)), k2 = exp_coef(norm3(v2));
    m3 R1 = rodrigues(v1, k1), R2 = rodrigues(v2, k2);
    m3 T1 = tangent_op(v1, k1), T2 = tangent_op(v2, k2);
    m3 M = mulABt(R2, R1);
    // φ = Rodrigues⁻¹(M) = spin⁻¹(M)/scac((tr M - 1)/2), Δv = φ/2
    double sc = scac((at(M, 0, 0) + at(M, 1, 1) + at(M, 2, 2) - 1.) / 2.);
    v3 phi = mk((at(M, 2, 1) - at(M, 1, 2)) / 2. / sc, (at(M, 0, 2) - at(M, 2, 0)) / 2. / sc, (at(M, 1, 0) - at(M, 0, 1)) / 2. / sc);
    v3 dv = 0.5 * phi;
    ExpCoef kp = exp_coef(norm3(phi)), kh = exp_coef(norm3(dv));
    m3 Tpi = tangent_inv(phi, kp);
    m3 H = rodrigues(dv, kh), TH = tangent_op(dv, kh);
    m3 HR1 = mul(H, R1);
    m3 rsm = mul(HR1, G.rm);
    // first-order spin Jacobians, 3×6 stored as 6 columns: [0..2] w.r.t. δv₁, [3..5] w.r.t. δv₂
    v3 wM[6], Jphi[6], JD[6], wH[6], wsm[6], HT1[3], MT1[3];
    for (int j = 0; j < 3; ++j) {
        v3 t1 = col(T1, j), t2 = col(T2, j);
        MT1[j] = mul(M, t1);
        HT1[j] = mul(H, t1);
        wM[j] = -1. * MT1[j];
        wM[3 + j] = t2;
    }
    for (int j = 0; j < 6; ++j) {
        Jphi[j] = mul(Tpi, wM[j]);
        JD[j] = 0.5 * Jphi[j];
        wH[j] = mul(TH, JD[j]);
        wsm[j] = wH[j];
    }
    for (int j = 0; j < 3; ++j) wsm[j] = wsm[j] + HT1[j];
    // ---- local deformations ---------------------------------------------------
    v3 cs = mk(0.5 * (u1.x + u2.x), 0.5 * (u1.y + u2.y), 0.5 * (u1.z + u2.z));
    v3 d = mk(u2.x + G.tgm.x * 0.5 - cs.x, u2.y + G.tgm.y * 0.5 - cs.y, u2.z + G.tgm.z * 0.5 - cs.z);
    v3 ul = mulT(rsm, d);
    ul.x = ul.x - L * 0.5;
    v3 vl = mulT(rsm, dv);
    // axial
    v3 w = mk(ul.x + L / 2., ul.y, ul.z);
    double ell = sqrt(w.x * w.x + w.y * w.y + w.z * w.z);
    double eps = ell * 2. / L - 1.;
    double Lsum = ((G.dL[0] + G.dL[1]) + G.dL[2]) + G.dL[3];
    // f_q (gradient of Π w.r.t. q = (ul, vl)) and D_q (Hessian of the internal part)
    double fq[6] = {0, 0, 0, 0, 0, 0};
    double Dq[21];   // upper triangle of 6×6, row-major
    for (int i = 0; i < 21; ++i) Dq[i] = 0.;
    auto dq = [&](int i, int j) -> double& { return Dq[i * 6 - (i * (i - 1)) / 2 + (j - i)]; };
    {
        double de = 2. / L / ell;                  // ∂ε/∂w = de·w
        double fa = S.EA * Lsum * eps;             // ∂Π_a/∂ε
        fq[0] += fa * de * w.x; fq[1] += fa * de * w.y; fq[2] += fa * de * w.z;
        double c1 = S.EA * Lsum * de * de;          // EA Lsum (∂ε)(∂ε)ᵀ
        double c2 = fa * de;                        // EA Lsum ε ∂²ε = fa·de·(I - w wᵀ/ℓ²)
        double il2 = 1. / (ell * ell);
        double ww[3] = {w.x, w.y, w.z};
        for (int i = 0; i < 3; ++i)
            for (int j = i; j < 3; ++j) dq(i, j) += c1 * ww[i] * ww[j] + c2 * ((i == j ? 1. : 0.) - ww[i] * ww[j] * il2);
    }
    const double ka = 2. / L, kv = 2. / L;
    for (int g = 0; g < 4; ++g) {
        double z = G.zgp[g];
        double ku = (-24. * z) / (L * L);
        double k1v = ka * vl.x, k2v = ku * ul.y + kv * vl.z, k3v = ku * ul.z - kv * vl.y;
        double dl = G.dL[g];
        // mᵢ = (GJ κ₁, EI₃ κ₂, EI₂ κ₃)
        double m1 = S.GJ * k1v, m2 = S.EI3 * k2v, m3v = S.EI2 * k3v;
        fq[3] += dl * m1 * ka;
        fq[1] += dl * m2 * ku; fq[5] += dl * m2 * kv;
        fq[2] += dl * m3v * ku; fq[4] += -dl * m3v * kv;
        dq(3, 3) += dl * S.GJ * ka * ka;
        dq(1, 1) += dl * S.EI3 * ku * ku; dq(1, 5) += dl * S.EI3 * ku * kv; dq(5, 5) += dl * S.EI3 * kv * kv;
        dq(2, 2) += dl * S.EI2 * ku * ku; dq(2, 4) += -dl * S.EI2 * ku * kv; dq(4, 4) += dl * S.EI2 * kv * kv;
    }
    // external load: Π_e = Σ_gp dL fₑ·(rₛₘ p_gp + cₛₘ), p_gp = tgₑ ζ + Y_gp q
    v3 gphi_e = mk(0, 0, 0);
    double Hee[6] = {0, 0, 0, 0, 0, 0};   // Σ [sym(e sᵀ) - (e·s) I] as s3 (xx,yy,zz,xy,xz,yz)
    double Ye[3][6];                       // Ỹ = Σ spin(e) rₛₘ Y_gp (3×6)
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 6; ++j) Ye[i][j] = 0.;
    double esum[3] = {0, 0, 0};
    for (int g = 0; g < 4; ++g) {
        double z = G.zgp[g];
        double ya = 2. * z, yu = -4. * (z * z * z) + 3. * z, yv = (z * z - 1. / 4.) * L;
        v3 y = mk(ya * ul.x, yu * ul.y + yv * vl.z, yu * ul.z - yv * vl.y);
        v3 p = mk(L * z + y.x, y.y, y.z);
        v3 s = mul(rsm, p);
        v3 e = G.dL[g] * fe;
        esum[0] += e.x; esum[1] += e.y; esum[2] += e.z;
        gphi_e = gphi_e + cross(s, e);
        double es = dot(e, s);
        Hee[0] += e.x * s.x - es; Hee[1] += e.y * s.y - es; Hee[2] += e.z * s.z - es;
        Hee[3] += 0.5 * (e.x * s.y + s.x * e.y); Hee[4] += 0.5 * (e.x * s.z + s.x * e.z); Hee[5] += 0.5 * (e.y * s.z + s.y * e.z);
        // f_q^ext = Yᵀ rₛₘᵀ e
        v3 re = mulT(rsm, e);
        fq[0] += ya * re.x; fq[1] += yu * re.y; fq[5] += yv * re.y; fq[2] += yu * re.z; fq[4] += -yv * re.z;
        // spin(e) rₛₘ Y: columns of rₛₘ Y are rₛₘ·(Y column)
        v3 r0 = col(rsm, 0), r1 = col(rsm, 1), r2 = col(rsm, 2);
        v3 yc[6] = {ya * r0, yu * r1, yu * r2, mk(0, 0, 0), -yv * r2, yv * r1};
        for (int j = 0; j < 6; ++j) { v3 c = cross(e, yc[j]); Ye[0][j] += c.x; Ye[1][j] += c.y; Ye[2][j] += c.z; }
    }
    // ---- gradients w.r.t. local coordinates z = (φ, Δ, d, c) -----------------
    v3 Fu = mul(rsm, mk(fq[0], fq[1], fq[2]));
    v3 Fv = mul(rsm, mk(fq[3], fq[4], fq[5]));
    v3 gphi = cross(Fu, d) + cross(Fv, dv) + gphi_e;
    v3 gD = Fv;
    // ---- residual R = Aᵀ g_z -----------------------------------------------------
    for (int i = 0; i < 3; ++i) {
        double gd = (&Fu.x)[i], gc = esum[i];
        out.R[i] = -0.5 * gd + 0.5 * gc;
        out.R[6 + i] = 0.5 * gd + 0.5 * gc;
    }
    for (int j = 0; j < 6; ++j) {
        double r = dot(wsm[j], gphi) + dot(JD[j], gD);
        out.R[j < 3 ? 3 + j : 6 + j] = r;
    }
    // ---- tangent --------------------------------------------------------------
    // Bq = ∂q/∂X (6×12): δul = rᵀ(δd + spin(d) φ), δvl = rᵀ(Δ + spin(Δv) φ)
    double Bq[6][12];
    for (int i = 0; i < 6; ++i) for (int j = 0; j < 12; ++j) Bq[i][j] = 0.;
    for (int i = 0; i < 3; ++i) {
        v3 ri = col(rsm, i);   // row i of rₛₘᵀ
        for (int k = 0; k < 3; ++k) { Bq[i][k] = -0.5 * (&ri.x)[k]; Bq[i][6 + k] = 0.5 * (&ri.x)[k]; }
        for (int j = 0; j < 6; ++j) {
            int cj = j < 3 ? 3 + j : 6 + j;
            Bq[i][cj] = dot(ri, cross(d, wsm[j]));
            Bq[3 + i][cj] = dot(ri, JD[j] + cross(dv, wsm[j]));
        }
    }
    // K1 = Bqᵀ Dq Bq
    double DB[6][12];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 12; ++j) {
            double s = 0.;
            for (int k = 0; k < 6; ++k) s += (i <= k ? dq(i, k) : dq(k, i)) * Bq[k][j];
            DB[i][j] = s;
        }
    for (int i = 0; i < 12; ++i)
        for (int j = i; j < 12; ++j) {
            double s = 0.;
            for (int k = 0; k < 6; ++k) s += Bq[k][i] * DB[k][j];
            out.K[tri(i, j)] = s;
        }
    // Hφφ' = sym(Fu dᵀ) - (Fu·d)I + sym(Fv Δvᵀ) - (Fv·Δv)I + Σ[sym(e sᵀ) - (e·s)I]
    double fud = dot(Fu, d), fvd = dot(Fv, dv);
    s3 Hpp;
    Hpp.xx = Fu.x * d.x - fud + Fv.x * dv.x - fvd + Hee[0];
    Hpp.yy = Fu.y * d.y - fud + Fv.y * dv.y - fvd + Hee[1];
    Hpp.zz = Fu.z * d.z - fud + Fv.z * dv.z - fvd + Hee[2];
    Hpp.xy = 0.5 * (Fu.x * d.y + d.x * Fu.y) + 0.5 * (Fv.x * dv.y + dv.x * Fv.y) + Hee[3];
    Hpp.xz = 0.5 * (Fu.x * d.z + d.x * Fu.z) + 0.5 * (Fv.x * dv.z + dv.x * Fv.z) + Hee[4];
    Hpp.yz = 0.5 * (Fu.y * d.z + d.y * Fu.z) + 0.5 * (Fv.y * dv.z + dv.y * Fv.z) + Hee[5];
    // N' = spin(Fv) AΔ + spin(Fu) Ad - Ỹ Bq   (3×12); Aφ = wsm on the rotation columns
    double Np[3][12];
    for (int j = 0; j < 12; ++j) {
        v3 c = mk(0, 0, 0);
        if (j < 3) c = cross(Fu, mk(j == 0 ? -0.5 : 0., j == 1 ? -0.5 : 0., j == 2 ? -0.5 : 0.));
        else if (j >= 6 && j < 9) c = cross(Fu, mk(j == 6 ? 0.5 : 0., j == 7 ? 0.5 : 0., j == 8 ? 0.5 : 0.));
        else { int r = j < 6 ? j - 3 : j - 6; c = cross(Fv, JD[r]); }
        for (int i = 0; i < 3; ++i) {
            double s = 0.;
            for (int k = 0; k < 6; ++k) s += Ye[i][k] * Bq[k][j];
            Np[i][j] = (&c.x)[i] - s;
        }
    }
    // K += Aφᵀ Hφφ' Aφ + Aφᵀ N' + N'ᵀ Aφ   (Aφ nonzero only on rotation columns)
    auto rot = [](int r) { return r < 3 ? 3 + r : 6 + r; };
    for (int a = 0; a < 6; ++a) {
        v3 hw = smul(Hpp, wsm[a]);
        for (int b = a; b < 6; ++b) {
            int i = rot(a), j = rot(b);
            out.K[tri(i, j)] += dot(wsm[b], hw);
        }
    }
    for (int a = 0; a < 6; ++a) {
        int i = rot(a);
        for (int j = 0; j < 12; ++j) {
            double t = wsm[a].x * Np[0][j] + wsm[a].y * Np[1][j] + wsm[a].z * Np[2][j];
            if (i <= j) out.K[tri(i, j)] += t;
            if (j <= i) out.K[tri(j, i)] += t;   // (N'ᵀAφ)[j][i]; the diagonal gets both
        }
    }
    // ---- second-order rotation chain, contracted with the adjoints --------------
    // μ = ½ T(φ)⁻ᵀ (gΔ + T_Hᵀ gφ);  λ₁ = Hᵀ gφ - Mᵀ μ;  λ₂ = μ
    v3 lamD = gD + mulT(TH, gphi);
    v3 mu = 0.5 * mulT(Tpi, lamD);
    v3 lam1 = mulT(H, gphi) - mulT(M, mu);
    double Gr[21];
    for (int i = 0; i < 21; ++i) Gr[i] = 0.;
    auto gr = [&](int i, int j) -> double& { return Gr[i * 6 - (i * (i - 1)) / 2 + (j - i)]; };
    {   // (i) λ₁·W₁, (ii) λ₂·W₂
        s3 B1 = dT_form(v1, lam1, k1), B2 = dT_form(v2, mu, k2);
        gr(0, 0) += B1.xx; gr(1, 1) += B1.yy; gr(2, 2) += B1.zz; gr(0, 1) += B1.xy; gr(0, 2) += B1.xz; gr(1, 2) += B1.yz;
        gr(3, 3) += B2.xx; gr(4, 4) += B2.yy; gr(5, 5) += B2.zz; gr(3, 4) += B2.xy; gr(3, 5) += B2.xz; gr(4, 5) += B2.yz;
    }
    {   // (iii) J_Δᵀ sym B(gφ; Δv) J_Δ  and  (iv) -J_φᵀ sym B(μ; φ) J_φ
        s3 BH = dT_form(dv, gphi, kh), BP = dT_form(phi, mu, kp);
        for (int a = 0; a < 6; ++a) {
            v3 h1 = smul(BH, JD[a]), h2 = smul(BP, Jphi[a]);
            for (int b = a; b < 6; ++b) gr(a, b) += dot(JD[b], h1) - dot(Jphi[b], h2);
        }
    }
    {   // (v) -½(ω_Hᵀ S_g Hω₁ - (Hω₁)ᵀ S_g ω_H), S_g = spin(gφ), Hω₁ = [H T₁, 0]
        //     as a bilinear form: -½(ω_H a · (gφ × Hω₁ b) - Hω₁ a · (gφ × ω_H b))
        // (vi) +½(ω₂ᵀ S_μ Mω₁ - (Mω₁)ᵀ S_μ ω₂), ω₂ = [0, T₂], Mω₁ = [M T₁, 0]
        v3 t2c[3] = {col(T2, 0), col(T2, 1), col(T2, 2)};
        for (int a = 0; a < 6; ++a)
            for (int b = a; b < 6; ++b) {
                v3 hw_a = a < 3 ? HT1[a] : mk(0, 0, 0), hw_b = b < 3 ? HT1[b] : mk(0, 0, 0);
                double t = -0.5 * (dot(wH[a], cross(gphi, hw_b)) - dot(hw_a, cross(gphi, wH[b])));
                v3 w2a = a >= 3 ? t2c[a - 3] : mk(0, 0, 0), w2b = b >= 3 ? t2c[b - 3] : mk(0, 0, 0);
                v3 mw_a = a < 3 ? MT1[a] : mk(0, 0, 0), mw_b = b < 3 ? MT1[b] : mk(0, 0, 0);
                t += 0.5 * (dot(w2a, cross(mu, mw_b)) - dot(mw_a, cross(mu, w2b)));
                gr(a, b) += t;
            }
    }
    for (int a = 0; a < 6; ++a)
        for (int b = a; b < 6; ++b) out.K[tri(rot(a), rot(b))] += gr(a, b);
    // NaN check like getresidual's hasnan (Assemble.jl:625)
    int nan = 0;
    for (int i = 0; i < 12; ++i) nan |= (out.R[i] != out.R[i]);
    for (int i = 0; i < 78; ++i) nan |= (out.K[i] != out.K[i]);
    out.nan = nan;
}

}  // namespace mh
