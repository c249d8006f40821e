// ORACLE — test infrastructure only (CPU checker; never part of the product path).
//
// Point elements and the toy fixture elements of the reference tests, generic
// in the solver's dual type TX.
#pragma once
#include <cmath>
#include "dual.hpp"
#include "rotations.hpp"

namespace orc {

// DofConstraint{:X} with gap(x,t)=x[1] and mode `equal` — what Hold builds
// (src/BasicElements.jl:426-438,484-488). X = (x, λ):  R = (-g∂x·λ, -gap).
// g: the prescribed value of a DofConstraint{:X} with gap(x,t) = x[1] - g(t) (0 for Hold,
// whose gap is x[1], BasicElements.jl:484-488).
template <class TX> inline void hold_residual(const TX* X0, double g, TX* R) {
    // P = constants(∂0(X),t); x∂ = variate{P,1}(x); gap,g∂x = value_∂{P,1}(gap(x∂))
    TX x = X0[0], lam = X0[1];
    double gdx = 1.;
    R[0] = -gdx * lam;
    R[1] = -(x - g);
}

// DofLoad (BasicElements.jl:281-284): R = [-F], F = value(t) evaluated by the caller.
template <class TX> inline void dofload_residual(double F, TX* R) { R[0] = TX(-F); }

// Turbine (test/SomeElements.jl:15-26) with sea(t,x)=(1,0), sky(t,x)=(0,1)
// as used by test/TestAssemble.jl:9-11,47-49.
struct Turbine { double xm[2]; double z; double seadrag, skydrag; };
template <class TX> inline void turbine_residual(const Turbine& o, const TX* /*X0*/, const double* A, TX* R) {
    double sea[2] = {1., 0.}, sky[2] = {0., 1.};
    for (int i = 0; i < 2; ++i) R[i] = TX(-sea[i] * o.seadrag * (1. + A[0]) - sky[i] * o.skydrag * (1. + A[1]));
}

// AnchorLine (test/SomeElements.jl:54-94). Its Lagrangian is L = Λ[1:2]∘Fd + Λ[3]·m3,
// so the residual obtained through getresidual (∂L/∂Λ, src/Assemble.jl:661-676) is (Fd, m3).
struct AnchorLine { double xmtop[3]; double dxmtop[3]; double xmbot[2]; double L; double buoyancy; };
inline const double* anchor_poly() {
    static const double p[11] = {2.82040487827, -24.86027164695, 153.69500343165, -729.52107422849, 2458.11921356871,
                                 -5856.85610233072, 9769.49700812681, -11141.12651712473, 8260.66447746395, -3582.36704093187,
                                 687.83550335374};
    return p;
}
// AnchorLine's residual (Fd, m3) and its requestables Fh and cr (SomeElements.jl:77-93), with
// the A-dofs (ΔL, Δbuoyancy) of the same scalar type so that they can carry partials too.
template <class TX> inline void anchorline_eval(const AnchorLine& o, const TX* x, const TX* A, TX* R, TX* Fh_out,
                                                TX* cr_out) {
    TX L = o.L * (1. + A[0]), buoy = o.buoyancy * (1. + A[1]);
    TX Xtop[3] = {x[0] + o.xmtop[0], x[1] + o.xmtop[1], TX(0. + o.xmtop[2])};
    TX alpha = x[2];
    TX c = dcos(alpha), s = dsin(alpha);
    TX dXtop[3];
    dXtop[0] = c * o.dxmtop[0] + (-s) * o.dxmtop[1] + 0. * o.dxmtop[2];
    dXtop[1] = s * o.dxmtop[0] + c * o.dxmtop[1] + 0. * o.dxmtop[2];
    dXtop[2] = TX(0. * o.dxmtop[0] + 0. * o.dxmtop[1] + 1. * o.dxmtop[2]);
    TX ch[2] = {Xtop[0] + dXtop[0] - o.xmbot[0], Xtop[1] + dXtop[1] - o.xmbot[1]};
    TX xaf = dsqrt(ch[0] * ch[0] + ch[1] * ch[1]);
    const double* p = anchor_poly();
    TX arg = (L - xaf) / Xtop[2];
    TX hy = TX(0.);
    for (int i = 10; i >= 0; --i) hy = p[i] + arg * hy;
    TX cr = dexp10(hy) * Xtop[2];
    TX Fh = -cr * buoy;
    TX Fd[2] = {ch[0] / xaf * Fh, ch[1] / xaf * Fh};
    TX m3 = dXtop[0] * Fd[1] - dXtop[1] * Fd[0];
    R[0] = Fd[0]; R[1] = Fd[1]; R[2] = m3;
    if (Fh_out) *Fh_out = Fh;
    if (cr_out) *cr_out = cr;
}
template <class TX> inline void anchorline_residual(const AnchorLine& o, const TX* x, const double* A, TX* R) {
    const TX Ad[2] = {TX(A[0]), TX(A[1])};
    anchorline_eval<TX>(o, x, Ad, R, nullptr, nullptr);
}

}  // namespace orc
