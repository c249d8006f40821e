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



// ------------------------------------------------------------ compact tangent
// The static tangent of the 2-node beam has an exact block structure: the
// translations enter Π only through d = ½(u₂-u₁) (internal terms) and
// c = ½(u₁+u₂) (linearly, external load), so with A = ∂²Π/∂d² (global frame),
// T = ∂²Π/∂d∂v (3×6 over the rotation dofs v = (v₁,v₂)) and S = ∂²Π/∂v² (6×6):
//     K[u₁,u₁] = K[u₂,u₂] = ¼A,  K[u₁,u₂] = K[u₂,u₁] = -¼A,
//     K[u₁,v]  = -½T,            K[u₂,v]  = +½T,          K[v,v] = S,
// and K is symmetric. These 45 numbers are the element output; the scatter
// multiplies by the exact powers of two ±1, ±½, ±¼.
// Slots: A at 0..5 (xx,yy,zz,xy,xz,yz), T at 6..23 (T[a][r] = 6 + 6a + r),
// S at 24..44 (upper triangle, row-major: 24 + sym6(r,s)).
constexpr int kCompact = 45;
MH_HD int sym3(int a, int b) {   // (xx,yy,zz,xy,xz,yz)
    if (a > b) { int t = a; a = b; b = t; }
    return a == b ? a : (a == 0 ? 2 + b : 5);
}
MH_HD int sym6(int r, int s) {   // upper triangle of 6×6, row-major
    if (r > s) { int t = r; r = s; s = t; }
    return r * 6 - (r * (r - 1)) / 2 + (s - r);
}
// Compact slot and exact multiplier of element entry K[i][j] (i,j = 0..11,
// element dof order t1,t2,t3,r1,r2,r3 of node 1 then node 2).
MH_HD int compact_slot(int i, int j, double* mult) {
    int ni = i / 6, a = i % 6, nj = j / 6, b = j % 6;
    bool ri = a >= 3, rj = b >= 3;
    if (!ri && !rj) { *mult = (ni == nj) ? 0.25 : -0.25; return sym3(a, b); }
    if (!ri && rj) { *mult = ni == 0 ? -0.5 : 0.5; return 6 + 6 * a + (3 * nj + b - 3); }
    if (ri && !rj) { *mult = nj == 0 ? -0.5 : 0.5; return 6 + 6 * b + (3 * ni + a - 3); }
    *mult = 1.;
    return 24 + sym6(3 * ni + a - 3, 3 * nj + b - 3);
}

struct BeamSection { double EA, EI2, EI3, GJ, w; };
struct BeamOut { double R[12]; double Kc[kCompact]; int nan; };

// Residual and compact tangent of Π = Π_int + Σ_gp e_gp·x_gp + m·vₛₘ with the loads
// e_gp and the couple m held fixed.
//   GEN = false: e_gp = dL_gp·fe with fe = (0,0,w) - U, m = 0 (the static element);
//   GEN = true:  e_gp = eg[gp] (already × dL_gp) and m = mtot = Σ_gp dL_gp·mₑ, the
//                frozen external loads of the dynamic element (inertia, added mass,
//                damping, roll inertia). vₛₘ = Rodrigues⁻¹(rₛₘ) (BeamElement.jl:203).
template <bool GEN = false>
MH_HD void beam_static(const BeamGeo& G, const BeamSection& S, v3 fe, const double* X, BeamOut& out,
                       const v3* eg = nullptr, v3 mtot = v3{0., 0., 0.}, double* hms = nullptr, int hst = 0) {
    const double L = G.L;
    v3 u1 = mk(X[0], X[1], X[2]), v1 = mk(X[3], X[4], X[5]);
    v3 u2 = mk(X[6], X[7], X[8]), v2 = mk(X[9], X[10], X[11]);
    // ---- rotation core, values --------------------------------------------------
    ExpCoef k1 = exp_coef(norm3(v1)), k2 = exp_coef(norm3(v2));
    m3 M, H, rsm;
    v3 phi, dv;
    ExpCoef kp, kh;
    {
        m3 R1 = rodrigues(v1, k1), R2 = rodrigues(v2, k2);
        M = mulABt(R2, R1);
        // φ = Rodrigues⁻¹(M) = spin⁻¹(M)/scac((tr M - 1)/2), Δv = φ/2
        double sc = scac((at(M, 0, 0) + at(M, 1, 1) + at(M, 2, 2) - 1.) / 2.);
        phi = mk((at(M, 2, 1) - at(M, 1, 2)) / 2. / sc, (at(M, 0, 2) - at(M, 2, 0)) / 2. / sc,
                 (at(M, 1, 0) - at(M, 0, 1)) / 2. / sc);
        dv = 0.5 * phi;
        kp = exp_coef(norm3(phi));
        kh = exp_coef(norm3(dv));
        H = rodrigues(dv, kh);
        rsm = mul(mul(H, R1), G.rm);
    }
    // ---- local deformations and the local gradient ---------------------------------
    v3 cs = mk(0.5 * (u1.x + u2.x), 0.5 * (u1.y + u2.y), 0.5 * (u1.z + u2.z));
    v3 d = mk(u2.x + G.tgm.x * 0.5 - cs.x, u2.y + G.tgm.y * 0.5 - cs.y, u2.z + G.tgm.z * 0.5 - cs.z);
    v3 ul = mulT(rsm, d);
    ul.x = ul.x - L * 0.5;
    v3 vl = mulT(rsm, dv);
    v3 w = mk(ul.x + L / 2., ul.y, ul.z);
    double ell = sqrt(w.x * w.x + w.y * w.y + w.z * w.z);
    double eps = ell * 2. / L - 1.;
    double Lsum = ((G.dL[0] + G.dL[1]) + G.dL[2]) + G.dL[3];
    double fq[6] = {0, 0, 0, 0, 0, 0};
    // D_q = ∂²Π_int/∂q² (q = (ul, vl)): dense uu block, u₂v₃ and u₃v₂ couplings, diagonal vv
    s3 Duu;
    double D15 = 0., D24 = 0., D33 = 0., D44 = 0., D55 = 0.;
    {
        double de = 2. / L / ell;
        double fa = S.EA * Lsum * eps;
        fq[0] += fa * de * w.x; fq[1] += fa * de * w.y; fq[2] += fa * de * w.z;
        double c1 = S.EA * Lsum * de * de, c2 = fa * de, il2 = 1. / (ell * ell);
        Duu.xx = c1 * w.x * w.x + c2 * (1. - w.x * w.x * il2);
        Duu.yy = c1 * w.y * w.y + c2 * (1. - w.y * w.y * il2);
        Duu.zz = c1 * w.z * w.z + c2 * (1. - w.z * w.z * il2);
        Duu.xy = c1 * w.x * w.y - c2 * w.x * w.y * il2;
        Duu.xz = c1 * w.x * w.z - c2 * w.x * w.z * il2;
        Duu.yz = c1 * w.y * w.z - c2 * w.y * w.z * il2;
    }
    const double ka = 2. / L, kv = 2. / L;
    for (int g = 0; g < 4; ++g) {
        double z = G.zgp[g];
        double ku = (-24. * z) / (L * L);
        double k1v = ka * vl.x, k2v = ku * ul.y + kv * vl.z, k3v = ku * ul.z - kv * vl.y;
        double dl = G.dL[g];
        double m1 = S.GJ * k1v, m2 = S.EI3 * k2v, m3v = S.EI2 * k3v;
        fq[3] += dl * m1 * ka;
        fq[1] += dl * m2 * ku; fq[5] += dl * m2 * kv;
        fq[2] += dl * m3v * ku; fq[4] += -dl * m3v * kv;
        D33 += dl * S.GJ * ka * ka;
        Duu.yy += dl * S.EI3 * ku * ku; D15 += dl * S.EI3 * ku * kv; D55 += dl * S.EI3 * kv * kv;
        Duu.zz += dl * S.EI2 * ku * ku; D24 += -dl * S.EI2 * ku * kv; D44 += dl * S.EI2 * kv * kv;
    }
    // external load Π_e = Σ_gp dL fₑ·(rₛₘ p_gp + cₛₘ), p_gp = tgₑ ζ + Y_gp q
    v3 gphi_e = mk(0, 0, 0), esum = mk(0, 0, 0);
    s3 Hee; Hee.xx = Hee.yy = Hee.zz = Hee.xy = Hee.xz = Hee.yz = 0.;
    v3 Ye[6];   // columns of Ỹ = Σ spin(e) rₛₘ Y_gp over q = (u1,u2,u3,v1,v2,v3); column v1 is zero
    for (int j = 0; j < 6; ++j) Ye[j] = mk(0, 0, 0);
    {
        v3 r0 = col(rsm, 0), r1 = col(rsm, 1), r2 = col(rsm, 2);
        for (int g = 0; g < 4; ++g) {
            double z = G.zgp[g];
            double ya = 2. * z, yu = -4. * (z * z * z) + 3. * z, yv = (z * z - 1. / 4.) * L;
            v3 y = mk(ya * ul.x, yu * ul.y + yv * vl.z, yu * ul.z - yv * vl.y);
            v3 s = mul(rsm, mk(L * z + y.x, y.y, y.z));
            v3 e = GEN ? eg[g] : G.dL[g] * fe;
            esum = esum + e;
            gphi_e = gphi_e + cross(s, e);
            double es = dot(e, s);
            Hee.xx += e.x * s.x - es; Hee.yy += e.y * s.y - es; Hee.zz += e.z * s.z - es;
            Hee.xy += 0.5 * (e.x * s.y + s.x * e.y); Hee.xz += 0.5 * (e.x * s.z + s.x * e.z); Hee.yz += 0.5 * (e.y * s.z + s.y * e.z);
            v3 re = mulT(rsm, e);
            fq[0] += ya * re.x; fq[1] += yu * re.y; fq[5] += yv * re.y; fq[2] += yu * re.z; fq[4] += -yv * re.z;
            Ye[0] = Ye[0] + cross(e, ya * r0);
            Ye[1] = Ye[1] + cross(e, yu * r1);
            Ye[2] = Ye[2] + cross(e, yu * r2);
            Ye[4] = Ye[4] + cross(e, -yv * r2);
            Ye[5] = Ye[5] + cross(e, yv * r1);
        }
    }
    v3 Fu = mul(rsm, mk(fq[0], fq[1], fq[2]));
    v3 Fv = mul(rsm, mk(fq[3], fq[4], fq[5]));
    v3 gphi = cross(Fu, d) + cross(Fv, dv) + gphi_e;
    // couple term m·Log(rₛₘ): gradient T(ψ)⁻ᵀm w.r.t. the spatial spin of rₛₘ, and the
    // second-order part -Jᵀ B J (J = T(ψ)⁻¹, B = dT_form(ψ, T⁻ᵀm)) of its Hessian
    s3 Hm; Hm.xx = Hm.yy = Hm.zz = Hm.xy = Hm.xz = Hm.yz = 0.;
    if (GEN) {
        v3 psi = mk(0, 0, 0);
        {
            double sc = scac((at(rsm, 0, 0) + at(rsm, 1, 1) + at(rsm, 2, 2) - 1.) / 2.);
            psi = mk((at(rsm, 2, 1) - at(rsm, 1, 2)) / 2. / sc, (at(rsm, 0, 2) - at(rsm, 2, 0)) / 2. / sc,
                     (at(rsm, 1, 0) - at(rsm, 0, 1)) / 2. / sc);
        }
        ExpCoef kpsi = exp_coef(norm3(psi));
        m3 Jp = tangent_inv(psi, kpsi);
        v3 mu_m = mulT(Jp, mtot);
        gphi = gphi + mu_m;
        s3 B = dT_form(psi, mu_m, kpsi);
        v3 J0 = col(Jp, 0), J1 = col(Jp, 1), J2 = col(Jp, 2);
        v3 BJ0 = smul(B, J0), BJ1 = smul(B, J1), BJ2 = smul(B, J2);
        Hm.xx = -dot(J0, BJ0); Hm.yy = -dot(J1, BJ1); Hm.zz = -dot(J2, BJ2);
        Hm.xy = -dot(J0, BJ1); Hm.xz = -dot(J0, BJ2); Hm.yz = -dot(J1, BJ2);
        // parked (LDS, stride hst) until Hφφ' is formed: fewer registers live through the
        // rotation Jacobians
        if (hms) { hms[0] = Hm.xx; hms[hst] = Hm.yy; hms[2 * hst] = Hm.zz; hms[3 * hst] = Hm.xy; hms[4 * hst] = Hm.xz; hms[5 * hst] = Hm.yz; }
    }
    v3 gD = Fv;
    // translational residual R = Aᵀ g_z on u₁, u₂
    out.R[0] = -0.5 * Fu.x + 0.5 * esum.x; out.R[1] = -0.5 * Fu.y + 0.5 * esum.y; out.R[2] = -0.5 * Fu.z + 0.5 * esum.z;
    out.R[6] = 0.5 * Fu.x + 0.5 * esum.x;  out.R[7] = 0.5 * Fu.y + 0.5 * esum.y;  out.R[8] = 0.5 * Fu.z + 0.5 * esum.z;
    // ---- rotation core, first-order spin Jacobians (6 columns: δv₁ then δv₂) ----------
    v3 JD[6], W[6], HT1[3], MT1[3], T2c[3];
    m3 TH = tangent_op(dv, kh);
    {
        m3 T1 = tangent_op(v1, k1), T2 = tangent_op(v2, k2), Tpi = tangent_inv(phi, kp);
        for (int j = 0; j < 3; ++j) {
            v3 t1 = col(T1, j);
            T2c[j] = col(T2, j);
            MT1[j] = mul(M, t1);
            HT1[j] = mul(H, t1);
            JD[j] = mul(Tpi, -0.5 * MT1[j]);
            JD[3 + j] = mul(Tpi, 0.5 * T2c[j]);
        }
        for (int j = 0; j < 6; ++j) W[j] = mul(TH, JD[j]);
        for (int j = 0; j < 3; ++j) W[j] = W[j] + HT1[j];
        // ---- second-order rotation chain contracted with the adjoints (into S) ------
        // μ = ½ T(φ)⁻ᵀ (gΔ + T_Hᵀ gφ);  λ₁ = Hᵀ gφ - Mᵀ μ;  λ₂ = μ
        v3 mu = 0.5 * mulT(Tpi, gD + mulT(TH, gphi));
        v3 lam1 = mulT(H, gphi) - mulT(M, mu);
        for (int i = 0; i < 21; ++i) out.Kc[24 + i] = 0.;
        double* Sm = out.Kc + 24;
        s3 B1 = dT_form(v1, lam1, k1), B2 = dT_form(v2, mu, k2);
        Sm[sym6(0, 0)] += B1.xx; Sm[sym6(1, 1)] += B1.yy; Sm[sym6(2, 2)] += B1.zz;
        Sm[sym6(0, 1)] += B1.xy; Sm[sym6(0, 2)] += B1.xz; Sm[sym6(1, 2)] += B1.yz;
        Sm[sym6(3, 3)] += B2.xx; Sm[sym6(4, 4)] += B2.yy; Sm[sym6(5, 5)] += B2.zz;
        Sm[sym6(3, 4)] += B2.xy; Sm[sym6(3, 5)] += B2.xz; Sm[sym6(4, 5)] += B2.yz;
        s3 BH = dT_form(dv, gphi, kh), BP = dT_form(phi, mu, kp);
        for (int a = 0; a < 6; ++a) {
            // J_φ = 2 J_Δ:  J_Δᵀ B_H J_Δ - J_φᵀ B_P J_φ = J_Δᵀ (B_H - 4 B_P) J_Δ
            v3 h = smul(BH, JD[a]) - 4. * smul(BP, JD[a]);
            v3 wHa = a < 3 ? W[a] - HT1[a] : W[a];
            for (int b = a; b < 6; ++b) {
                v3 wHb = b < 3 ? W[b] - HT1[b] : W[b];
                double t = dot(JD[b], h);
                v3 hwa = a < 3 ? HT1[a] : mk(0, 0, 0), hwb = b < 3 ? HT1[b] : mk(0, 0, 0);
                t += -0.5 * (dot(wHa, cross(gphi, hwb)) - dot(hwa, cross(gphi, wHb)));
                v3 w2a = a >= 3 ? T2c[a - 3] : mk(0, 0, 0), w2b = b >= 3 ? T2c[b - 3] : mk(0, 0, 0);
                v3 mwa = a < 3 ? MT1[a] : mk(0, 0, 0), mwb = b < 3 ? MT1[b] : mk(0, 0, 0);
                t += 0.5 * (dot(w2a, cross(mu, mwb)) - dot(mwa, cross(mu, w2b)));
                Sm[sym6(a, b)] += t;
            }
        }
    }
    // rotational residual
    for (int j = 0; j < 6; ++j) out.R[j < 3 ? 3 + j : 6 + j] = dot(W[j], gphi) + dot(JD[j], gD);
    // ---- A = rₛₘ D_uu rₛₘᵀ ----------------------------------------------------------------
    {
        v3 c0 = col(rsm, 0), c1 = col(rsm, 1), c2 = col(rsm, 2);
        // rows of rₛₘ D_uu: (rₛₘ D_uu)[i][k] = Σ_m rsm[i][m] D[m][k]
        v3 Dc0 = mk(Duu.xx, Duu.xy, Duu.xz), Dc1 = mk(Duu.xy, Duu.yy, Duu.yz), Dc2 = mk(Duu.xz, Duu.yz, Duu.zz);
        v3 P0 = mul(rsm, Dc0), P1 = mul(rsm, Dc1), P2 = mul(rsm, Dc2);   // columns of rₛₘ D_uu
        // A = (rₛₘ D_uu) rₛₘᵀ: A[i][j] = Σ_k P_k[i] rsm[j][k]
        auto Aij = [&](int i, int j) {
            return (&P0.x)[i] * at(rsm, j, 0) + (&P1.x)[i] * at(rsm, j, 1) + (&P2.x)[i] * at(rsm, j, 2);
        };
        out.Kc[0] = Aij(0, 0); out.Kc[1] = Aij(1, 1); out.Kc[2] = Aij(2, 2);
        out.Kc[3] = Aij(0, 1); out.Kc[4] = Aij(0, 2); out.Kc[5] = Aij(1, 2);
        (void)c0; (void)c1; (void)c2;
    }
    // ---- Br = ∂q/∂v (6×6), T, and the remaining rotation-rotation terms ---------------------
    {
        // Q3 = spin(F_u) - Ỹ_u rₛₘᵀ   (3×3),  T = rₛₘ(D_uu Br_u + D_uv Br_v) + Q3ᵀ W
        s3 Hpp;   // Hφφ' = sym(Fu dᵀ) - (Fu·d)I + sym(Fv Δvᵀ) - (Fv·Δv)I + Σ[sym(e sᵀ) - (e·s)I]
        double fud = dot(Fu, d), fvd = dot(Fv, dv);
        Hpp.xx = Fu.x * d.x - fud + Fv.x * dv.x - fvd + Hee.xx;
        Hpp.yy = Fu.y * d.y - fud + Fv.y * dv.y - fvd + Hee.yy;
        Hpp.zz = Fu.z * d.z - fud + Fv.z * dv.z - fvd + Hee.zz;
        Hpp.xy = 0.5 * (Fu.x * d.y + d.x * Fu.y) + 0.5 * (Fv.x * dv.y + dv.x * Fv.y) + Hee.xy;
        Hpp.xz = 0.5 * (Fu.x * d.z + d.x * Fu.z) + 0.5 * (Fv.x * dv.z + dv.x * Fv.z) + Hee.xz;
        Hpp.yz = 0.5 * (Fu.y * d.z + d.y * Fu.z) + 0.5 * (Fv.y * dv.z + dv.y * Fv.z) + Hee.yz;
        if (GEN && hms) { Hm.xx = hms[0]; Hm.yy = hms[hst]; Hm.zz = hms[2 * hst]; Hm.xy = hms[3 * hst]; Hm.xz = hms[4 * hst]; Hm.yz = hms[5 * hst]; }
        if (GEN) {
            Hpp.xx += Hm.xx; Hpp.yy += Hm.yy; Hpp.zz += Hm.zz;
            Hpp.xy += Hm.xy; Hpp.xz += Hm.xz; Hpp.yz += Hm.yz;
        }
        double* Sm = out.Kc + 24;
        v3 Bu[6], Bv[6], Np[6];
        for (int j = 0; j < 6; ++j) {
            Bu[j] = mulT(rsm, cross(d, W[j]));
            Bv[j] = mulT(rsm, JD[j] + cross(dv, W[j]));
            // N'_rot column j = spin(Fv) JD_j - Ỹ Br_j   (Ỹ column 3 (v1) is zero)
            v3 yb = (Bu[j].x * Ye[0] + Bu[j].y * Ye[1]) + Bu[j].z * Ye[2];
            yb = yb + Bv[j].y * Ye[4] + Bv[j].z * Ye[5];
            Np[j] = cross(Fv, JD[j]) - yb;
        }
        for (int a = 0; a < 6; ++a) {
            // D Br column a
            v3 DBu = smul(Duu, Bu[a]) + mk(0., D15 * Bv[a].z, D24 * Bv[a].y);
            v3 DBv = mk(D33 * Bv[a].x, D44 * Bv[a].y + D24 * Bu[a].z, D55 * Bv[a].z + D15 * Bu[a].y);
            v3 hw = smul(Hpp, W[a]);
            for (int b = a; b < 6; ++b) {
                double t = dot(Bu[b], DBu) + dot(Bv[b], DBv) + dot(W[b], hw) + dot(W[a], Np[b]) + dot(W[b], Np[a]);
                Sm[sym6(a, b)] += t;
            }
            // T column a: rₛₘ (D_uu Bu_a + D_uv Bv_a) + Q3ᵀ W_a
            v3 C = mul(rsm, DBu);
            // Q3ᵀ W_a = spin(Fu)ᵀ W_a - rₛₘ Ỹ_uᵀ W_a = W_a × Fu - rₛₘ (Ye_k · W_a)_k
            v3 yw = mk(dot(Ye[0], W[a]), dot(Ye[1], W[a]), dot(Ye[2], W[a]));
            v3 Tc = C + cross(W[a], Fu) - mul(rsm, yw);
            out.Kc[6 + 0 * 6 + a] = Tc.x; out.Kc[6 + 1 * 6 + a] = Tc.y; out.Kc[6 + 2 * 6 + a] = Tc.z;
        }
    }
    int nan = 0;
    // Muscade's residual always contracts mₑ with ∂vₛₘ/∂X (BeamElement.jl:170), with
    // vₛₘ = Rodrigues⁻¹(rₛₘ) = spin⁻¹(rₛₘ)/scac((tr rₛₘ - 1)/2): for a corotated frame at
    // a half turn, scac and acos′ there make it NaN even where mₑ = 0.
    if (!GEN) nan |= !((at(rsm, 0, 0) + at(rsm, 1, 1) + at(rsm, 2, 2) - 1.) / 2. > -1.);
    for (int i = 0; i < 12; ++i) nan |= (out.R[i] != out.R[i]);
    for (int i = 0; i < kCompact; ++i) nan |= (out.Kc[i] != out.Kc[i]);
    out.nan = nan;
}

// Full 12×12 entry from the compact form (host harnesses and checks).
MH_HD double compact_entry(const double* Kc, int i, int j) {
    double m;
    int s = compact_slot(i, j, &m);
    return m * Kc[s];
}

}  // namespace mh
