// ORACLE — test infrastructure only (CPU checker; never part of the product path).
//
// Restatement of toolbox/BarElement.jl (Bar3D, AxisymmetricBarCrossSection)
// with Muscade's differentiation strategy: the element dofs are time jets over
// the solver's duals (motion / motion⁻¹, src/Taylor.jl:24-65); the virtual
// strain ε∂X₀ is the explicit formula of BarElement.jl:180 on ∂0(δ), which
// carries the solver partials.
#pragma once
#include <cmath>
#include "beam.hpp"

namespace orc {

struct AxisymmetricBarCrossSection {   // BarElement.jl:36-48 (field order kept)
    double EA, mu, w, Cat, Clt, Cqt, Can, Cln, Cqn;
};

struct Bar3D {   // BarElement.jl:89-101
    double cm[3], tgm[3], tge[3];
    double L0, Ls;
    AxisymmetricBarCrossSection mat;
    double wgp[4], zgp[4], znod[2], psi1[4], psi2[4];
    int Udof;
};

// BarElement.jl:106-121; eps_s is ϵₛ (default eps() = 2^-52).
inline void make_bar(const double* c1, const double* c2, const AxisymmetricBarCrossSection& mat, double eps_s, int Udof,
                     Bar3D* o) {
    for (int i = 0; i < 3; ++i) { o->cm[i] = (c1[i] + c2[i]) / 2.; o->tgm[i] = c2[i] - c1[i]; }
    o->L0 = std::sqrt(o->tgm[0] * o->tgm[0] + o->tgm[1] * o->tgm[1] + o->tgm[2] * o->tgm[2]);
    o->Ls = (1. - eps_s) * o->L0;
    o->tge[0] = o->L0; o->tge[1] = 0.; o->tge[2] = 0.;
    const double L0 = o->L0, s30 = std::sqrt(30.);
    o->wgp[0] = L0 / 2. * (18. - s30) / 36.; o->wgp[1] = L0 / 2. * (18. + s30) / 36.;
    o->wgp[2] = L0 / 2. * (18. + s30) / 36.; o->wgp[3] = L0 / 2. * (18. - s30) / 36.;
    const double s65 = std::sqrt(6. / 5.);
    o->zgp[0] = -1. / 2. * std::sqrt(3. / 7. + 2. / 7. * s65); o->zgp[1] = -1. / 2. * std::sqrt(3. / 7. - 2. / 7. * s65);
    o->zgp[2] = +1. / 2. * std::sqrt(3. / 7. - 2. / 7. * s65); o->zgp[3] = +1. / 2. * std::sqrt(3. / 7. + 2. / 7. * s65);
    o->znod[0] = -1. / 2.; o->znod[1] = 1. / 2.;
    for (int g = 0; g < 4; ++g) { o->psi1[g] = -o->zgp[g] + 1. / 2.; o->psi2[g] = o->zgp[g] + 1. / 2.; }   // :75-76
    o->mat = mat;
    o->Udof = Udof;
}

// resultants(::AxisymmetricBarCrossSection, ε, x, u, t) — BarElement.jl:136-161.
// x: the Gauss point's value/velocity/acceleration (x[der]), dl = ∂0(δ).
template <class TX>
inline void bar_resultants(const AxisymmetricBarCrossSection& o, const TX& eps0, const V3<TX>* x, int ND, const V3<TX>& dl,
                           double t, TX& fi_out, V3<TX>& fe) {
    V3<TX> v = ND >= 2 ? x[1] : vzero<TX>();
    V3<TX> a = ND >= 3 ? x[2] : vzero<TX>();
    V3<TX> fi = vscale(o.mu, a);
    double ramp = (std::min(t, -5.) + 10.) / 5. * o.w;
    TX at = a[0] * dl[0] + a[1] * dl[1] + a[2] * dl[2];
    V3<TX> an; for (int i = 0; i < 3; ++i) an[i] = a[i] - at * dl[i];
    V3<TX> fa; fa[0] = o.Cat * at; fa[1] = o.Can * an[1]; fa[2] = o.Can * an[2];
    TX vt = v[0] * dl[0] + v[1] * dl[1] + v[2] * dl[2];
    V3<TX> vn; for (int i = 0; i < 3; ++i) vn[i] = v[i] - vt * dl[i];
    TX fqt = o.Cqt * dsq(vt); if (vt < 0.) fqt = -fqt;
    TX fqn2 = o.Cqn * dsq(vn[1]); if (vn[1] < 0.) fqn2 = -fqn2;
    TX fqn3 = o.Cqn * dsq(vn[2]); if (vn[2] < 0.) fqn3 = -fqn3;
    V3<TX> fd; fd[0] = o.Clt * vt + fqt; fd[1] = o.Cln * vn[1] + fqn2; fd[2] = o.Cln * vn[2] + fqn3;
    for (int i = 0; i < 3; ++i) fe[i] = ((fi[i] + (i == 2 ? TX(ramp) : TX(0.))) + fa[i]) + fd[i];
    fi_out = o.EA * eps0;
}

// residual(::Bar3D, X, U, A, t) — BarElement.jl:167-202. X: ND rows of 6; U: 3 (Udof only).
template <class TX, int ND>
inline void bar_residual(const Bar3D& o, const TX (*X)[12], const TX* U, double t, TX* R) {
    using J = Jet<TX, ND>;
    using JT = typename J::type;
    JT x_[6];
    for (int i = 0; i < 6; ++i) { TX a[3]; for (int d = 0; d < ND; ++d) a[d] = X[d][i]; x_[i] = J::make(a); }
    V3<JT> u1, u2, c, tg;
    for (int i = 0; i < 3; ++i) { u1[i] = x_[i]; u2[i] = x_[3 + i]; }
    for (int i = 0; i < 3; ++i) c[i] = o.cm[i] + 0.5 * (u1[i] + u2[i]);
    for (int i = 0; i < 3; ++i) tg[i] = o.tgm[i] + u2[i] - u1[i];
    JT L = dsqrt(dsq(tg[0]) + dsq(tg[1]) + dsq(tg[2]));
    V3<JT> dl_; for (int i = 0; i < 3; ++i) dl_[i] = tg[i] / L;
    JT eps_ = L / o.Ls - 1.;
    TX tmp[3];
    J::unpack(eps_, tmp);
    TX eps0 = tmp[0];
    V3<TX> dl0;
    for (int i = 0; i < 3; ++i) { J::unpack(dl_[i], tmp); dl0[i] = tmp[0]; }
    TX edX[6];
    for (int i = 0; i < 3; ++i) { edX[i] = (1. / o.L0) * (-dl0[i]); edX[3 + i] = (1. / o.L0) * dl0[i]; }
    TX Rs[6];
    for (int g = 0; g < 4; ++g) {
        double z = o.zgp[g];
        double p1 = -z + 1. / 2., p2 = z + 1. / 2.;
        V3<JT> xg; for (int i = 0; i < 3; ++i) xg[i] = c[i] + tg[i] * z;
        V3<TX> xd[3];
        for (int i = 0; i < 3; ++i) { J::unpack(xg[i], tmp); for (int d = 0; d < ND; ++d) xd[d][i] = tmp[d]; }
        TX fi;
        V3<TX> fe;
        bar_resultants(o.mat, eps0, xd, ND, dl0, t, fi, fe);
        if (o.Udof && U) for (int i = 0; i < 3; ++i) fe[i] = fe[i] - U[i];
        // fᵢ ∘₀ ε∂X₀ + fₑ ∘₁ x∂X₀, x∂X₀ = [ψ₁ I, ψ₂ I]
        for (int j = 0; j < 6; ++j) {
            TX xj = j < 3 ? p1 * fe[j] : p2 * fe[j - 3];
            TX Rg = (fi * edX[j] + xj) * o.wgp[g];
            Rs[j] = g == 0 ? Rg : Rs[j] + Rg;
        }
    }
    for (int j = 0; j < 6; ++j) R[j] = Rs[j];
}

}  // namespace orc
