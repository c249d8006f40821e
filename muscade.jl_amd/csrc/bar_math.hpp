// Analytic FP64 residual and tangent of Bar3D (toolbox/BarElement.jl:36-202).
//
// The bar's Gauss-point motion is linear in the element dofs, x_gp = c + tg·ζ,
// so its velocity and acceleration are exact linear interpolations of the
// nodal ones; only the unit direction δ = tg/L and the strain ε = L/Lₛ - 1 are
// nonlinear (and only their values enter the loads). The residual is
//     R = Σ_gp w_gp (EA ε · ε∂X₀ + x∂X₀ᵀ fₑ(δ, ẋ_gp, ẍ_gp, t) - x∂X₀ᵀ U),
// with the reference's virtual strain ε∂X₀ = (-δ, δ)/L₀ (BarElement.jl:180)
// and x∂X₀ = [ψ₁I ψ₂I]. The tangent along the solver's trial path
// (X₀ + εe, X₁ + a₁εe, X₂ + b₁εe) follows from ∂δ/∂u₂ = (I - δδᵀ)/L,
// ∂ε/∂u₂ = δ/Lₛ and the 3×3 load Jacobians ∂fₑ/∂δ, ∂fₑ/∂ẋ, ∂fₑ/∂ẍ.
#pragma once
#include "beam_math.hpp"

namespace mh {

struct BarSection { double EA, mu, w, Cat, Clt, Cqt, Can, Cln, Cqn; };   // AxisymmetricBarCrossSection order
struct BarGeo { v3 tgm; double L0, Ls; double wgp[4], zgp[4]; };

MH_HD BarGeo bar_geometry(v3 c1, v3 c2, double eps_s) {   // BarElement.jl:106-121
    BarGeo g;
    g.tgm = c2 - c1;
    g.L0 = sqrt(g.tgm.x * g.tgm.x + g.tgm.y * g.tgm.y + g.tgm.z * g.tgm.z);
    g.Ls = (1. - eps_s) * g.L0;
    const double s30 = 5.477225575051661, s65 = 1.0954451150103321;
    g.wgp[0] = g.L0 / 2. * (18. - s30) / 36.; g.wgp[1] = g.L0 / 2. * (18. + s30) / 36.;
    g.wgp[2] = g.wgp[1]; g.wgp[3] = g.wgp[0];
    const double z1 = 0.5 * sqrt(3. / 7. + 2. / 7. * s65), z2 = 0.5 * sqrt(3. / 7. - 2. / 7. * s65);
    g.zgp[0] = -z1; g.zgp[1] = -z2; g.zgp[2] = z2; g.zgp[3] = z1;
    return g;
}

MH_HD double comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// Load Jacobians of the bar's resultants (BarElement.jl:136-161) at (δ, v, a):
// fₑ = μa + (0,0,ramp) + (Caₜ aₜ, Caₙ aₙ₂, Caₙ aₙ₃) + (Clₜvₜ + Cqₜvₜ|vₜ|, Clₙvₙ₂ + Cqₙvₙ₂|vₙ₂|, …),
// aₜ = a·δ, aₙ = a - aₜδ, vₜ = v·δ, vₙ = v - vₜδ. Row i of each matrix is ∂fₑᵢ/∂(·).
struct BarLoad { v3 fe; m3 Dd, Dv, Da; };
MH_HD BarLoad bar_load(const BarSection& S, v3 dl, v3 v, v3 a, double ramp) {
    BarLoad o;
    const double a_t = (a.x * dl.x + a.y * dl.y) + a.z * dl.z;
    const v3 an = mk(a.x - a_t * dl.x, a.y - a_t * dl.y, a.z - a_t * dl.z);
    const double vt = (v.x * dl.x + v.y * dl.y) + v.z * dl.z;
    const v3 vn = mk(v.x - vt * dl.x, v.y - vt * dl.y, v.z - vt * dl.z);
    double fqt = S.Cqt * (vt * vt); if (vt < 0.) fqt = -fqt;
    double fq2 = S.Cqn * (vn.y * vn.y); if (vn.y < 0.) fq2 = -fq2;
    double fq3 = S.Cqn * (vn.z * vn.z); if (vn.z < 0.) fq3 = -fq3;
    const v3 fa = mk(S.Cat * a_t, S.Can * an.y, S.Can * an.z);
    const v3 fd = mk(S.Clt * vt + fqt, S.Cln * vn.y + fq2, S.Cln * vn.z + fq3);
    o.fe = mk(((S.mu * a.x + 0.) + fa.x) + fd.x, ((S.mu * a.y + 0.) + fa.y) + fd.y, ((S.mu * a.z + ramp) + fa.z) + fd.z);
    const double kt = S.Clt + 2. * S.Cqt * fabs(vt);
    const double k2 = S.Cln + 2. * S.Cqn * fabs(vn.y), k3 = S.Cln + 2. * S.Cqn * fabs(vn.z);
    for (int j = 0; j < 3; ++j) {
        const double dj = comp(dl, j), aj = comp(a, j), vj = comp(v, j);
        const double ey = j == 1 ? 1. : 0., ez = j == 2 ? 1. : 0.;
        // ∂/∂a
        at(o.Da, 0, j) = (j == 0 ? S.mu : 0.) + S.Cat * dj;
        at(o.Da, 1, j) = (j == 1 ? S.mu : 0.) + S.Can * (ey - dl.y * dj);
        at(o.Da, 2, j) = (j == 2 ? S.mu : 0.) + S.Can * (ez - dl.z * dj);
        // ∂/∂v
        at(o.Dv, 0, j) = kt * dj;
        at(o.Dv, 1, j) = k2 * (ey - dl.y * dj);
        at(o.Dv, 2, j) = k3 * (ez - dl.z * dj);
        // ∂/∂δ
        at(o.Dd, 0, j) = S.Cat * aj + kt * vj;
        at(o.Dd, 1, j) = -S.Can * (dl.y * aj + a_t * ey) - k2 * (dl.y * vj + vt * ey);
        at(o.Dd, 2, j) = -S.Can * (dl.z * aj + a_t * ez) - k3 * (dl.z * vj + vt * ez);
    }
    return o;
}

// Whole element. X: ND rows of 6 (t1,t2,t3 of node 1 then node 2); U: 3 (zeros without
// U-dofs); t: time; scale: scale.X (6); c: Newmark (a₁,a₂,a₃,b₁,b₂,b₃).
// Outputs: R (6, scaled), K (36, column-major, = ∂R/∂δX scaled like Muscade's δX and R),
// dRr (6, ∂R/∂δr of :step, scaled rows) when non-null. Returns 1 on NaN.
// K = w0·∂R/∂X₀ + a₁·∂R/∂X₁ + b₁·∂R/∂X₂; w0 = 0 isolates the X′ and X″ partials.
template <int ND>
MH_HD int bar_element(const BarGeo& G, const BarSection& S, const double (*X)[6], const double* U, double t,
                      const double* scale, const double* c, double* R, double* K, double* dRr, double w0 = 1.) {
    const v3 u1 = mk(X[0][0], X[0][1], X[0][2]), u2 = mk(X[0][3], X[0][4], X[0][5]);
    const v3 tg = mk((G.tgm.x + u2.x) - u1.x, (G.tgm.y + u2.y) - u1.y, (G.tgm.z + u2.z) - u1.z);
    const double L = sqrt((tg.x * tg.x + tg.y * tg.y) + tg.z * tg.z);
    const v3 dl = mk(tg.x / L, tg.y / L, tg.z / L);
    const double eps = L / G.Ls - 1.;
    const double fi = S.EA * eps;
    const double ramp = ((t < -5. ? t : -5.) + 10.) / 5. * S.w;
    double edX[6];
    for (int i = 0; i < 3; ++i) { edX[i] = (1. / G.L0) * (-comp(dl, i)); edX[3 + i] = (1. / G.L0) * comp(dl, i); }
    // P = I - δδᵀ; ∂δ/∂u₂ = P/L = -∂δ/∂u₁
    double P[3][3];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) P[i][j] = (i == j ? 1. : 0.) - comp(dl, i) * comp(dl, j);
    double Rs[6] = {0, 0, 0, 0, 0, 0};
    double Kl[6][6];   // Kl[j][k] = ∂R_j/∂(trial direction k), unscaled
    double Rr[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < 6; ++j) for (int k = 0; k < 6; ++k) Kl[j][k] = 0.;
    double av[6], bv[6];   // δr seeds: a = a₂x′ + a₃x″, b = b₂x′ + b₃x″
    for (int k = 0; k < 6; ++k) {
        av[k] = ND >= 2 ? c[1] * X[ND >= 2 ? 1 : 0][k] + (ND >= 3 ? c[2] * X[ND >= 3 ? 2 : 0][k] : 0.) : 0.;
        bv[k] = ND >= 3 ? c[4] * X[ND >= 2 ? 1 : 0][k] + c[5] * X[ND >= 3 ? 2 : 0][k] : 0.;
    }
    for (int g = 0; g < 4; ++g) {
        const double z = G.zgp[g], w = G.wgp[g];
        const double p1 = -z + 1. / 2., p2 = z + 1. / 2.;
        v3 v = mk(0, 0, 0), a = mk(0, 0, 0);
        if (ND >= 2) {
            const double* x1 = X[ND >= 2 ? 1 : 0];
            v = mk(0.5 * (x1[0] + x1[3]) + (x1[3] - x1[0]) * z, 0.5 * (x1[1] + x1[4]) + (x1[4] - x1[1]) * z,
                   0.5 * (x1[2] + x1[5]) + (x1[5] - x1[2]) * z);
        }
        if (ND >= 3) {
            const double* x2 = X[ND >= 3 ? 2 : 0];
            a = mk(0.5 * (x2[0] + x2[3]) + (x2[3] - x2[0]) * z, 0.5 * (x2[1] + x2[4]) + (x2[4] - x2[1]) * z,
                   0.5 * (x2[2] + x2[5]) + (x2[5] - x2[2]) * z);
        }
        BarLoad Ld = bar_load(S, dl, v, a, ramp);
        const v3 fe = mk(Ld.fe.x - U[0], Ld.fe.y - U[1], Ld.fe.z - U[2]);
        for (int j = 0; j < 6; ++j) {
            const double xj = j < 3 ? p1 * comp(fe, j) : p2 * comp(fe, j - 3);
            Rs[j] = Rs[j] + (fi * edX[j] + xj) * w;
        }
        // δr: ∂ẋ_gp/∂r, ∂ẍ_gp/∂r
        v3 vr = mk(0.5 * (av[0] + av[3]) + (av[3] - av[0]) * z, 0.5 * (av[1] + av[4]) + (av[4] - av[1]) * z,
                   0.5 * (av[2] + av[5]) + (av[5] - av[2]) * z);
        v3 ar = mk(0.5 * (bv[0] + bv[3]) + (bv[3] - bv[0]) * z, 0.5 * (bv[1] + bv[4]) + (bv[4] - bv[1]) * z,
                   0.5 * (bv[2] + bv[5]) + (bv[5] - bv[2]) * z);
        const v3 fr = mul(Ld.Dv, vr) + mul(Ld.Da, ar);
        for (int j = 0; j < 6; ++j) Rr[j] += (j < 3 ? p1 * comp(fr, j) : p2 * comp(fr, j - 3)) * w;
        for (int k = 0; k < 6; ++k) {
            const int q = k % 3;
            const double sgn = k < 3 ? -1. : 1.;
            const double pk = k < 3 ? p1 : p2;
            // ∂δ/∂X_k and ∂ε/∂X_k
            const v3 dd = mk(sgn * P[0][q] / L, sgn * P[1][q] / L, sgn * P[2][q] / L);
            const double deps = sgn * comp(dl, q) / G.Ls;
            // ∂fₑ/∂(direction k) = Dδ ∂δ/∂X₀ₖ + a₁ Dv ψₖ e_q + b₁ Da ψₖ e_q
            v3 dfe = w0 * mul(Ld.Dd, dd);
            if (ND >= 2) dfe = dfe + (c[0] * pk) * col(Ld.Dv, q);
            if (ND >= 3) dfe = dfe + (c[3] * pk) * col(Ld.Da, q);
            for (int j = 0; j < 6; ++j) {
                const int qj = j % 3;
                const double sj = j < 3 ? -1. : 1.;
                const double dedX = sj * sgn * P[qj][q] / (G.L0 * L);   // ∂(ε∂X₀)_j/∂X_k
                const double xj = j < 3 ? p1 * comp(dfe, j) : p2 * comp(dfe, j - 3);
                Kl[j][k] += (w0 * ((S.EA * deps) * edX[j] + fi * dedX) + xj) * w;
            }
        }
    }
    int nan = 0;
    for (int j = 0; j < 6; ++j) { R[j] = Rs[j] * scale[j]; nan |= R[j] != R[j]; }
    for (int k = 0; k < 6; ++k)
        for (int j = 0; j < 6; ++j) { K[j + 6 * k] = Kl[j][k] * (scale[j] * scale[k]); nan |= K[j + 6 * k] != K[j + 6 * k]; }
    if (dRr) for (int j = 0; j < 6; ++j) { dRr[j] = Rr[j] * scale[j]; nan |= dRr[j] != dRr[j]; }
    return nan;
}

// Λ·∂R/∂X⁽ᵈᵉʳ⁾ (der < ND) and Λ·∂R/∂U for one bar (unscaled).
template <int ND>
MH_HD int bar_vjp_element(const BarGeo& G, const BarSection& S, const double (*X)[6], const double* U, double t,
                          const double* lam, double (*gX)[6], double* gU) {
    const double one[6] = {1., 1., 1., 1., 1., 1.};
    double R[6], K[36];
    int nan = 0;
    for (int der = 0; der < ND; ++der) {
        const double c[6] = {der == 1 ? 1. : 0., 0., 0., der == 2 ? 1. : 0., 0., 0.};
        nan |= bar_element<ND>(G, S, X, U, t, one, c, R, K, nullptr, der == 0 ? 1. : 0.);
        for (int k = 0; k < 6; ++k) {
            double g = 0.;
            for (int i = 0; i < 6; ++i) g = g + lam[i] * K[i + 6 * k];
            // constant row indices (der is a loop counter: an index by it would put gX in scratch)
            if (der == 0) gX[0][k] = g;
            else if (ND > 1 && der == 1) gX[ND > 1 ? 1 : 0][k] = g;
            else if (ND > 2) gX[ND > 2 ? 2 : 0][k] = g;
        }
    }
    // ∂Rⱼ/∂U_q = -Σ_gp w ψⱼ δ(j mod 3, q)
    for (int q = 0; q < 3; ++q) {
        double s = 0.;
        for (int g = 0; g < 4; ++g) {
            const double z = G.zgp[g];
            s = s + ((-z + 1. / 2.) * lam[q] + (z + 1. / 2.) * lam[3 + q]) * G.wgp[g];
        }
        gU[q] = -s;
    }
    return nan;
}

}  // namespace mh
