// ORACLE — test infrastructure only (CPU checker; never part of the product path).
//
// Restatement of the corotated EulerBeam3D of toolbox/BeamElement.jl with
// Muscade's own differentiation strategy:
//   * values and time derivatives by nested time-jets over the solver's duals
//     (motion/motion⁻¹, src/Taylor.jl:24-65),
//   * Jacobians of ε, κ, x and vₛₘ w.r.t. X₀ by re-variating X₀ to the
//     solver precedence P (revariate, Taylor.jl:143-151) and composing the
//     result back onto the solver's variation (composeJacobian/chainrule/
//     McLaurin, Taylor.jl:225-273,319-323).
// apply{:chainrule} inside `corotated` (BeamElement.jl:196-203) is evaluated
// directly over the same nested dual: chainrule is a performance device whose
// result equals direct evaluation through second order (Taylor.jl:271-295).
#pragma once
#include <cmath>
#include <type_traits>
#include "dual.hpp"
#include "rotations.hpp"

namespace orc {

struct BeamCrossSection {   // BeamElement.jl:6-25 (field order kept)
    double EA, EI2, EI3, GJ, mu, iota1, w, Ca1, Cl1, Cq1, Ca2, Cl2, Cq2, Ca3, Cl3, Cq3;
};

struct EulerBeam3D {   // BeamElement.jl:87-103
    double cm[3];
    double rm[9];    // column-major, columns t,n,b
    double zgp[4];
    double znod[2];
    double tgm[3];
    double tge[3];
    double ya[4], yu[4], yv[4], ka[4], ku[4], kv[4];
    double L;
    double dL[4];
    BeamCrossSection mat;
    int Udof;
};

// BeamElement.jl:122-148. Returns 0, or 1 when orient2 is nearly parallel to the element.
inline int make_beam(const double* c1, const double* c2, const double* orient2_in, const BeamCrossSection& mat, int Udof,
                     EulerBeam3D* o) {
    double tg[3], t[3], n[3], b[3], or2[3];
    for (int i = 0; i < 3; ++i) { o->cm[i] = (c1[i] + c2[i]) / 2.; tg[i] = c2[i] - c1[i]; o->tgm[i] = tg[i]; }
    double L = std::sqrt(tg[0] * tg[0] + tg[1] * tg[1] + tg[2] * tg[2]);
    for (int i = 0; i < 3; ++i) t[i] = tg[i] / L;
    double no = std::sqrt(orient2_in[0] * orient2_in[0] + orient2_in[1] * orient2_in[1] + orient2_in[2] * orient2_in[2]);
    for (int i = 0; i < 3; ++i) or2[i] = orient2_in[i] / no;
    double d = or2[0] * t[0] + or2[1] * t[1] + or2[2] * t[2];
    for (int i = 0; i < 3; ++i) n[i] = or2[i] - t[i] * d;
    double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (!(nn > 1e-3)) return 1;
    for (int i = 0; i < 3; ++i) n[i] /= nn;
    b[0] = t[1] * n[2] - t[2] * n[1]; b[1] = t[2] * n[0] - t[0] * n[2]; b[2] = t[0] * n[1] - t[1] * n[0];
    for (int i = 0; i < 3; ++i) { o->rm[i] = t[i]; o->rm[3 + i] = n[i]; o->rm[6 + i] = b[i]; }
    o->tge[0] = L; o->tge[1] = 0.; o->tge[2] = 0.;
    double s30 = std::sqrt(30.);
    o->dL[0] = L / 2. * (18. - s30) / 36.; o->dL[1] = L / 2. * (18. + s30) / 36.;
    o->dL[2] = L / 2. * (18. + s30) / 36.; o->dL[3] = L / 2. * (18. - s30) / 36.;
    double s65 = std::sqrt(6. / 5.);
    o->zgp[0] = -1. / 2. * std::sqrt(3. / 7. + 2. / 7. * s65); o->zgp[1] = -1. / 2. * std::sqrt(3. / 7. - 2. / 7. * s65);
    o->zgp[2] = +1. / 2. * std::sqrt(3. / 7. - 2. / 7. * s65); o->zgp[3] = +1. / 2. * std::sqrt(3. / 7. + 2. / 7. * s65);
    o->znod[0] = -1. / 2.; o->znod[1] = 1. / 2.;
    for (int g = 0; g < 4; ++g) {   // shape functions BeamElement.jl:74-79,146
        double z = o->zgp[g];
        o->ya[g] = 2. * z;
        o->yu[g] = -4. * (z * z * z) + 3. * z;
        o->yv[g] = (z * z - 1. / 4.) * L;
        o->ka[g] = 2. / L;
        o->ku[g] = (-24. * z) / (L * L);
        o->kv[g] = 2. / L;
    }
    o->L = L;
    o->mat = mat;
    o->Udof = Udof;
    return 0;
}

// corotated + kinematics (BeamElement.jl:176-208), generic in the scalar type.
template <class T> struct Kin {
    V3<T> vsm; M3<T> rsm; V3<T> ul2, vl2, csm;
    T eps;
    V3<T> kap[4], x[4];
};
template <class T> inline Kin<T> kinematics(const EulerBeam3D& o, const T* X0) {
    Kin<T> k;
    V3<T> ug1, vg1, ug2, vg2;
    for (int i = 0; i < 3; ++i) { ug1[i] = X0[i]; vg1[i] = X0[3 + i]; ug2[i] = X0[6 + i]; vg2[i] = X0[9 + i]; }
    M3<T> rs1 = Rodrigues(vg1);
    M3<T> rs2 = Rodrigues(vg2);
    V3<T> dv = Rodrigues_inv(mmt(rs2, rs1));
    for (int i = 0; i < 3; ++i) dv[i] = 0.5 * dv[i];
    M3<double> rm; for (int i = 0; i < 9; ++i) rm.m[i] = o.rm[i];
    k.rsm = mm(mm(Rodrigues(dv), rs1), rm);
    k.vsm = Rodrigues_inv(k.rsm);
    V3<T> cs; for (int i = 0; i < 3; ++i) cs[i] = 0.5 * (ug1[i] + ug2[i]);
    V3<T> arm; for (int i = 0; i < 3; ++i) arm[i] = ug2[i] + o.tgm[i] * o.znod[1] - cs[i];
    k.ul2 = mtv(k.rsm, arm);
    for (int i = 0; i < 3; ++i) k.ul2[i] = k.ul2[i] - o.tge[i] * o.znod[1];
    k.vl2 = mtv(k.rsm, dv);
    for (int i = 0; i < 3; ++i) k.csm[i] = cs[i] + o.cm[i];
    double L = o.L;
    k.eps = dsqrt(dsq(k.ul2[0] + L / 2.) + dsq(k.ul2[1]) + dsq(k.ul2[2])) * 2. / L - 1.;
    for (int g = 0; g < 4; ++g) {
        k.kap[g][0] = o.ka[g] * k.vl2[0];
        k.kap[g][1] = o.ku[g] * k.ul2[1] + o.kv[g] * k.vl2[2];
        k.kap[g][2] = o.ku[g] * k.ul2[2] - o.kv[g] * k.vl2[1];
        V3<T> y;
        y[0] = o.ya[g] * k.ul2[0];
        y[1] = o.yu[g] * k.ul2[1] + o.yv[g] * k.vl2[2];
        y[2] = o.yu[g] * k.ul2[2] - o.yv[g] * k.vl2[1];
        V3<T> p; for (int i = 0; i < 3; ++i) p[i] = o.tge[i] * o.zgp[g] + y[i];
        V3<T> q = mv(k.rsm, p);
        for (int i = 0; i < 3; ++i) k.x[g][i] = q[i] + k.csm[i];
    }
    return k;
}

// ---- time jets (motion / motion⁻¹) ----------------------------------------
template <class TX, int ND> struct Jet;
template <class TX> struct Jet<TX, 1> {
    using type = TX;
    static type make(const TX* a) { return a[0]; }
    static void unpack(const type& y, TX* out) { out[0] = y; }
};
template <class TX> struct Jet<TX, 2> {
    using type = Dual<1, TX>;
    static type make(const TX* a) { type r; r.x = a[0]; r.dx[0] = a[1]; return r; }
    static void unpack(const type& y, TX* out) { out[0] = y.x; out[1] = y.dx[0]; }
};
template <class TX> struct Jet<TX, 3> {
    using type = Dual<1, Dual<1, TX>>;
    static type make(const TX* a) {
        type r; r.x.x = a[0]; r.x.dx[0] = a[1]; r.dx[0].x = a[1]; r.dx[0].dx[0] = a[2]; return r; }
    static void unpack(const type& y, TX* out) { out[0] = y.x.x; out[1] = y.dx[0].x; out[2] = y.dx[0].dx[0]; }
};

// ---- re-variation of X₀ and composition of Jacobians ----------------------
// P = 1 (TX = double): first-order re-variation; Jacobians are doubles.
// P = 2 (TX carries solver partials): second-order; Jacobian entries become
// TX values whose partials hold ∂²y/∂X₀² contracted with the solver's δX.
template <class TX> struct Revar;
template <> struct Revar<double> {
    using HD = Dual<12, double>;
    static void make(const double* X0, HD* T) { for (int i = 0; i < 12; ++i) T[i] = seed<12, double>(X0[i], i); }
    static void jac(const HD& y, const double* /*X0*/, double* J) { for (int i = 0; i < 12; ++i) J[i] = y.dx[i]; }
};
template <int NP, class R> struct Revar<Dual<NP, R>> {
    using TX = Dual<NP, R>;
    using HD = Dual<12, Dual<12, double>>;
    static void make(const TX* X0, HD* T) {
        for (int i = 0; i < 12; ++i) {
            double v = VALUE(X0[i]);
            T[i].x = seed<12, double>(v, i);
            T[i].dx[i] = Dual<12, double>(1.);
        }
    }
    // McLaurin(∂{2,12}(y), X₀ - VALUE(X₀))  (Taylor.jl:229-243,271-274,320)
    static void jac(const HD& y, const TX* X0, TX* J) {
        TX dX[12];
        for (int k = 0; k < 12; ++k) dX[k] = X0[k] - VALUE(X0[k]);
        for (int i = 0; i < 12; ++i) {
            const Dual<12, double>& e = y.dx[i];
            TX s = e.dx[0] * dX[0];
            for (int k = 1; k < 12; ++k) s = s + e.dx[k] * dX[k];
            s = s / 1.;
            J[i] = e.x + s;
        }
    }
};

// Intrinsic rotation rates (Rotations.jl:177-182): returns vᵢ₁, vᵢ₂.
template <class TX, int ND> inline void intrinsic_rates(const M3<TX>* r, V3<TX>& vi1, V3<TX>& vi2) {
    vi1 = vzero<TX>(); vi2 = vzero<TX>();
    if constexpr (ND >= 2) vi1 = spin_inv(mtm(r[0], r[1]));
    if constexpr (ND >= 3) vi2 = spin_inv(madd(mtm(r[1], r[1]), mtm(r[0], r[2])));
}

// Optional requestables of the residual (used by the strain-gauge wrapper).
template <class TX> struct BeamReq { TX eps; V3<TX> kappa_mid; };
// All requestables of the residual (BeamElement.jl:28-64,151-174), plain values: ☼ε, ♢κ,
// and per Gauss point ☼x, ☼κgp with their ND time levels, and the resultants' ☼fᵢ, ☼mᵢ,
// ☼fₑ (before the U-load is subtracted, BeamElement.jl:161), ☼mₑ.
struct BeamRes {
    double eps[3], kmid[3][3];
    double x[4][3][3], kgp[4][3][3];   // [gp][der][comp]
    double fi[4], mi[4][3], fe[4][3], me[4][3];
};

// residual(::EulerBeam3D, X, U, A, t) — BeamElement.jl:151-174.
// X: ND vectors of 12; U: NDU vectors of 3 (only if o.Udof).
template <class TX, int ND, int NDU>
inline void beam_residual(const EulerBeam3D& o, const TX (*X)[12], const TX (*U)[3], TX* R, BeamReq<TX>* req = nullptr,
                          BeamRes* res = nullptr) {
    using J = Jet<TX, ND>;
    using JT = typename J::type;
    // direct kinematics over time-jets
    JT Xj[12];
    for (int i = 0; i < 12; ++i) { TX a[3]; for (int d = 0; d < ND; ++d) a[d] = X[d][i]; Xj[i] = J::make(a); }
    Kin<JT> kj = kinematics<JT>(o, Xj);
    TX eps[3], kap[4][3][3], xg[4][3][3];   // [gp][comp][der] ; eps[der]
    M3<TX> rsm[3];
    {
        TX tmp[3];
        J::unpack(kj.eps, tmp); for (int d = 0; d < ND; ++d) eps[d] = tmp[d];
        for (int g = 0; g < 4; ++g) for (int c = 0; c < 3; ++c) {
            J::unpack(kj.kap[g][c], tmp); for (int d = 0; d < ND; ++d) kap[g][c][d] = tmp[d];
            J::unpack(kj.x[g][c], tmp); for (int d = 0; d < ND; ++d) xg[g][c][d] = tmp[d];
        }
        for (int e = 0; e < 9; ++e) { J::unpack(kj.rsm.m[e], tmp); for (int d = 0; d < ND; ++d) rsm[d].m[e] = tmp[d]; }
    }
    V3<TX> vi1, vi2;
    intrinsic_rates<TX, ND>(rsm, vi1, vi2);
    // Jacobians w.r.t. X₀
    using RV = Revar<TX>;
    typename RV::HD T0[12];
    RV::make(X[0], T0);
    Kin<typename RV::HD> kt = kinematics<typename RV::HD>(o, T0);
    TX epsJ[12], kapJ[4][3][12], xJ[4][3][12], vsmJ[3][12];
    RV::jac(kt.eps, X[0], epsJ);
    for (int g = 0; g < 4; ++g) for (int c = 0; c < 3; ++c) { RV::jac(kt.kap[g][c], X[0], kapJ[g][c]); RV::jac(kt.x[g][c], X[0], xJ[g][c]); }
    for (int c = 0; c < 3; ++c) RV::jac(kt.vsm[c], X[0], vsmJ[c]);
    // resultants (BeamElement.jl:28-64)
    const BeamCrossSection& m = o.mat;
    const M3<TX>& r0 = rsm[0];
    TX Rsum[12];
    for (int g = 0; g < 4; ++g) {
        V3<TX> x1, x2;
        for (int c = 0; c < 3; ++c) { x1[c] = ND >= 2 ? xg[g][c][1] : TX(0.); x2[c] = ND >= 3 ? xg[g][c][2] : TX(0.); }
        V3<TX> xl1 = mtv(r0, x1), xl2 = mtv(r0, x2);
        V3<TX> fi = vscale(m.mu, x2);
        V3<TX> fal; fal[0] = m.Ca1 * xl2[0]; fal[1] = m.Ca2 * xl2[1]; fal[2] = m.Ca3 * xl2[2];
        V3<TX> fag = mv(r0, fal);
        V3<TX> fll; fll[0] = m.Cl1 * xl1[0]; fll[1] = m.Cl2 * xl1[1]; fll[2] = m.Cl3 * xl1[2];
        V3<TX> flg = mv(r0, fll);
        TX fq1 = m.Cq1 * dsq(xl1[0]); if (xl1[0] < 0.) fq1 = -fq1;
        TX fq2 = m.Cq2 * dsq(xl1[1]); if (xl1[1] < 0.) fq2 = -fq2;
        TX fq3 = m.Cq3 * dsq(xl1[2]); if (xl1[2] < 0.) fq3 = -fq3;
        V3<TX> fql; fql[0] = fq1; fql[1] = fq2; fql[2] = fq3;
        V3<TX> fqg = mv(r0, fql);
        V3<TX> fe;
        for (int c = 0; c < 3; ++c) fe[c] = fi[c] + fag[c] + flg[c] + fqg[c] + (c == 2 ? TX(m.w) : TX(0.));
        TX m1l = m.iota1 * vi2[0];
        V3<TX> me; for (int c = 0; c < 3; ++c) me[c] = r0(c, 0) * m1l;
        TX fint = m.EA * eps[0];
        V3<TX> mint;
        mint[0] = m.GJ * kap[g][0][0]; mint[1] = m.EI3 * kap[g][1][0]; mint[2] = m.EI2 * kap[g][2][0];
        if constexpr (std::is_same_v<TX, double>) {
            if (res) {
                res->fi[g] = fint;
                for (int c = 0; c < 3; ++c) { res->mi[g][c] = mint[c]; res->fe[g][c] = fe[c]; res->me[g][c] = me[c]; }
                for (int d = 0; d < ND; ++d)
                    for (int c = 0; c < 3; ++c) { res->x[g][d][c] = xg[g][c][d]; res->kgp[g][d][c] = kap[g][c][d]; }
            }
        }
        if (o.Udof && U) for (int c = 0; c < 3; ++c) fe[c] = fe[c] - U[0][c];
        for (int j = 0; j < 12; ++j) {
            TX a = fint * epsJ[j];
            TX b = mint[0] * kapJ[g][0][j] + mint[1] * kapJ[g][1][j] + mint[2] * kapJ[g][2][j];
            TX c = fe[0] * xJ[g][0][j] + fe[1] * xJ[g][1][j] + fe[2] * xJ[g][2][j];
            TX d = me[0] * vsmJ[0][j] + me[1] * vsmJ[1][j] + me[2] * vsmJ[2][j];
            TX Rg = (a + b + c + d) * o.dL[g];
            Rsum[j] = (g == 0) ? Rg : Rsum[j] + Rg;
        }
    }
    for (int j = 0; j < 12; ++j) R[j] = Rsum[j];
    if constexpr (std::is_same_v<TX, double>) {
        if (res) {
            TX tmp[3];
            V3<JT> kk; kk[0] = kj.vl2[0]; kk[1] = kj.vl2[2]; kk[2] = -kj.vl2[1];
            for (int d = 0; d < ND; ++d) res->eps[d] = eps[d];
            for (int c = 0; c < 3; ++c) {
                J::unpack(kk[c], tmp);
                for (int d = 0; d < ND; ++d) res->kmid[d][c] = tmp[d] * (2. / o.L);
            }
        }
    }
    if (req) {
        req->eps = eps[0];
        TX tmp[3];
        V3<JT> kk; kk[0] = kj.vl2[0]; kk[1] = kj.vl2[2]; kk[2] = -kj.vl2[1];
        for (int c = 0; c < 3; ++c) { J::unpack(kk[c], tmp); req->kappa_mid[c] = tmp[0] * (2. / o.L); }
    }
}

}  // namespace orc
