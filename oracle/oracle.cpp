// ORACLE — test infrastructure only. Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load this library, and only as the checker
// (or as the timed CPU baseline). The product path never calls it.
//
// CPU restatement of Muscade.jl's element-evaluation-and-assembly path:
//   dof numbering          src/ModelDescription.jl:195-263   (orc_number_dofs)
//   asmvec!/asmmat!        src/Assemble.jl:340-357,373-448   (orc_model_prepare_sweepx)
//   assemble!/addin!       src/Assemble.jl:470-505,559-608, src/SweepX.jl:39-96 (orc_model_assemble_sweepx)
//   element residuals      toolbox/BeamElement.jl, toolbox/BarElement.jl, src/BasicElements.jl,
//                          test/SomeElements.jl
//   rotations              toolbox/Rotations.jl (orc_sinc1, orc_scac_d, orc_rodrigues ...)
//   requestables           toolbox/BeamElement.jl:28-64,151-174 (orc_beam_results), the strain
//                          gauges' εₐₓ, κ with second derivatives (orc_beam_gauge2), AnchorLine's
//                          Fh, cr (orc_anchor_eval, pinned by test/TestElementCost.jl:28-34,81-95)
// Parity pins: see the header of each test module in tests/ for the golden
// literal (reference file:line) each function is checked against.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <tuple>
#include <vector>
#include "bar.hpp"
#include "beam.hpp"
#include "elements.hpp"

using namespace orc;

enum Kind { K_BEAM = 0, K_HOLD = 1, K_DOFLOAD = 2, K_ZERO = 3, K_TURBINE = 4, K_ANCHOR = 5, K_BAR = 6 };

struct EType {
    int kind, nele, nX, nU, nA;
    int udof;
    std::vector<int64_t> iX, iU, iA;   // 1-based model dof indices, [iele*n + k]
    std::vector<double> scaleX;        // per element dof (dis.dis[ieletyp].scale.X)
    std::vector<double> scaleL, scaleU;   // scale.Λ, scale.U (AssemblyDirect's revariate)
    std::vector<EulerBeam3D> beams;
    std::vector<Bar3D> bars;
    std::vector<double> F;             // DofLoad values
    std::vector<Turbine> turbines;
    std::vector<AnchorLine> anchors;
    std::vector<int64_t> asm1, asm2;   // asm[1,ieletyp] (nX×nele), asm[2,ieletyp] (nX²×nele), 1-based, 0 = skip
};
// One class-pair pattern of AssemblyDirect (asmmat!(asm[α,β], asm[α], asm[β]), Assemble.jl:373-448).
struct DPattern {
    int64_t nr = 0, nc = 0, nnz = 0;
    std::vector<int64_t> colptr, rowval;
    std::vector<std::vector<int64_t>> asm2;   // per type: (nα·nβ)×nele, entry i + nα·j, 1-based
};
struct OModel {
    int64_t nX, nU, nA;
    int dOX = -1, dOU = 0, dIA = 0;    // AssemblyDirect{OX,OU,IA} prepared
    DPattern dpat[3][3];               // classes Λ, X, U
    double t = 0.;                     // state.time handed to time-dependent elements
    std::vector<EType> types;
    int64_t nnz = 0;
    std::vector<int64_t> colptr, rowval;
};

// --------------------------------------------------------------- elements
template <class TX, int ND>
static int elem_residual(const EType& ty, int e, const TX (*X)[12], const double* U, const double* A, double t, TX* R) {
    switch (ty.kind) {
    case K_BAR: {
        TX Uj[3];
        for (int c = 0; c < 3; ++c) Uj[c] = TX(ty.udof ? U[c] : 0.);
        bar_residual<TX, ND>(ty.bars[e], X, ty.udof ? Uj : nullptr, t, R);
        return 0;
    }
    case K_BEAM: {
        if (ty.udof) {
            TX Uj[1][3];
            for (int c = 0; c < 3; ++c) Uj[0][c] = TX(U[c]);
            beam_residual<TX, ND, 1>(ty.beams[e], X, Uj, R);
        } else {
            beam_residual<TX, ND, 1>(ty.beams[e], X, nullptr, R);
        }
        return 0;
    }
    case K_HOLD: hold_residual<TX>(X[0], ty.F.empty() ? 0. : ty.F[e], R); return 0;
    case K_DOFLOAD: dofload_residual<TX>(ty.F[e], R); return 0;
    case K_ZERO: for (int i = 0; i < ty.nX; ++i) R[i] = TX(0.); return 0;
    case K_TURBINE: turbine_residual<TX>(ty.turbines[e], X[0], A, R); return 0;
    case K_ANCHOR: anchorline_residual<TX>(ty.anchors[e], X[0], A, R); return 0;
    }
    return 1;
}

static bool has_nan(double v) { return std::isnan(v); }
template <int N, class T> static bool has_nan(const Dual<N, T>& v) {
    if (has_nan(v.x)) return true;
    for (int i = 0; i < N; ++i) if (has_nan(v.dx[i])) return true;
    return false;
}

extern "C" {

// ---------------------------------------------------------------- the power rule
// x^b (Adiff.jl:230) on a three-level nested dual x = W + W′ε₃, W = V + V′ε₂, W′ = Q + Q′ε₂,
// V = v + v′ε₁, V′ = p + p′ε₁, Q = q + q′ε₁, Q′ = r + r′ε₁ (ε₁ innermost): in and out are
// {v, v′, p, p′, q, q′, r, r′}. A probe of the rule for the tests.
void orc_ipow_probe(const double* in, int b, double* out) {
    using D1 = Dual<1, double>;
    using D2 = Dual<1, D1>;
    using D3 = Dual<1, D2>;
    D3 x;
    x.x.x.x = in[0]; x.x.x.dx[0] = in[1]; x.x.dx[0].x = in[2]; x.x.dx[0].dx[0] = in[3];
    x.dx[0].x.x = in[4]; x.dx[0].x.dx[0] = in[5]; x.dx[0].dx[0].x = in[6]; x.dx[0].dx[0].dx[0] = in[7];
    const D3 y = ipow(x, b);
    out[0] = y.x.x.x; out[1] = y.x.x.dx[0]; out[2] = y.x.dx[0].x; out[3] = y.x.dx[0].dx[0];
    out[4] = y.dx[0].x.x; out[5] = y.dx[0].x.dx[0]; out[6] = y.dx[0].dx[0].x; out[7] = y.dx[0].dx[0].dx[0];
}

// ---------------------------------------------------------------- rotations
double orc_sinc1_k(int k, double x) {
    switch (k) { case 0: return sinc1_0(x); case 1: return sinc1_1(x); case 2: return sinc1_2(x);
                 case 3: return sinc1_3(x); case 4: return sinc1_4(x); }
    return 0.;
}
// k-th derivative of scac at x by nested variation (test/TestRotations.jl:31-55)
double orc_scac_d(int k, double x) {
    if (k == 0) return scac(x);
    if (k == 1) { auto y = scac(seed<1, double>(x, 0)); return y.dx[0]; }
    if (k == 2) { using D1 = Dual<1, double>; Dual<1, D1> v; v.x = seed<1, double>(x, 0); v.dx[0] = D1(1.);
                  auto y = scac(v); return y.dx[0].dx[0]; }
    if (k == 3) { using D1 = Dual<1, double>; using D2 = Dual<1, D1>; Dual<1, D2> v;
                  v.x.x = seed<1, double>(x, 0); v.x.dx[0] = D1(1.); v.dx[0] = D2(1.);
                  auto y = scac(v); return y.dx[0].dx[0].dx[0]; }
    if (k == 4) { using D1 = Dual<1, double>; using D2 = Dual<1, D1>; using D3 = Dual<1, D2>; Dual<1, D3> v;
                  v.x.x.x = seed<1, double>(x, 0); v.x.x.dx[0] = D1(1.); v.x.dx[0] = D2(1.); v.dx[0] = D3(1.);
                  auto y = scac(v); return y.dx[0].dx[0].dx[0].dx[0]; }
    return 0.;
}
void orc_rodrigues(const double* v, double* M) {
    V3<double> a; for (int i = 0; i < 3; ++i) a[i] = v[i];
    M3<double> r = Rodrigues(a); for (int i = 0; i < 9; ++i) M[i] = r.m[i];
}
void orc_rodrigues_inv(const double* M, double* v) {
    M3<double> m; for (int i = 0; i < 9; ++i) m.m[i] = M[i];
    V3<double> r = Rodrigues_inv(m); for (int i = 0; i < 3; ++i) v[i] = r[i];
}
// w = Rodrigues⁻¹(Rodrigues(v)) and its Jacobian (column-major 3×3)
void orc_rodrigues_roundtrip(const double* v, double* w, double* J) {
    using D = Dual<3, double>;
    V3<D> a; for (int i = 0; i < 3; ++i) a[i] = seed<3, double>(v[i], i);
    V3<D> r = Rodrigues_inv(Rodrigues(a));
    for (int i = 0; i < 3; ++i) { w[i] = r[i].x; for (int j = 0; j < 3; ++j) J[i + 3 * j] = r[i].dx[j]; }
}
void orc_adjust(const double* u, const double* v, double* r) {
    V3<double> a, b; for (int i = 0; i < 3; ++i) { a[i] = u[i]; b[i] = v[i]; }
    V3<double> w = adjust(a, b); for (int i = 0; i < 3; ++i) r[i] = w[i];
}

// ---------------------------------------------------------------- beam
// Constructor (BeamElement.jl:122-148). out: cm(3) rm(9) zgp(4) znod(2) tgm(3) tge(3)
// ya yu yv ka ku kv (6×4) L(1) dL(4) = 53 doubles. Returns 1 if orient2 is nearly parallel.
int orc_beam_construct(const double* c1, const double* c2, const double* orient2, double* out) {
    BeamCrossSection mat{}; EulerBeam3D o;
    if (make_beam(c1, c2, orient2, mat, 0, &o)) return 1;
    double* p = out;
    for (int i = 0; i < 3; ++i) *p++ = o.cm[i];
    for (int i = 0; i < 9; ++i) *p++ = o.rm[i];
    for (int i = 0; i < 4; ++i) *p++ = o.zgp[i];
    for (int i = 0; i < 2; ++i) *p++ = o.znod[i];
    for (int i = 0; i < 3; ++i) *p++ = o.tgm[i];
    for (int i = 0; i < 3; ++i) *p++ = o.tge[i];
    const double* sh[6] = {o.ya, o.yu, o.yv, o.ka, o.ku, o.kv};
    for (int s = 0; s < 6; ++s) for (int i = 0; i < 4; ++i) *p++ = sh[s][i];
    *p++ = o.L;
    for (int i = 0; i < 4; ++i) *p++ = o.dL[i];
    return 0;
}

// Bar3D constructor (BarElement.jl:106-121). out: cm(3) tgm(3) L0 Ls wgp(4) zgp(4) znod(2) psi1(4) psi2(4) = 26.
void orc_bar_construct(const double* c1, const double* c2, double eps_s, double* out) {
    AxisymmetricBarCrossSection mat{}; Bar3D o;
    make_bar(c1, c2, mat, eps_s, 0, &o);
    double* p = out;
    for (int i = 0; i < 3; ++i) *p++ = o.cm[i];
    for (int i = 0; i < 3; ++i) *p++ = o.tgm[i];
    *p++ = o.L0; *p++ = o.Ls;
    for (int i = 0; i < 4; ++i) *p++ = o.wgp[i];
    for (int i = 0; i < 4; ++i) *p++ = o.zgp[i];
    for (int i = 0; i < 2; ++i) *p++ = o.znod[i];
    for (int i = 0; i < 4; ++i) *p++ = o.psi1[i];
    for (int i = 0; i < 4; ++i) *p++ = o.psi2[i];
}
// diffed_residual of one Bar3D (Diagnostic.jl:929-961): R (6), dR/dX[der] (ND × 6×6, column-major),
// dR/dU (6×3) if udof. X: ND×6. mat9 = AxisymmetricBarCrossSection fields.
int orc_bar_diffed_residual(const double* c1, const double* c2, const double* mat9, double eps_s, int udof, int ND,
                            const double* X, const double* U, double t, double* R, double* dRdX, double* dRdU) {
    AxisymmetricBarCrossSection mat; std::memcpy(&mat, mat9, sizeof(mat));
    Bar3D o;
    make_bar(c1, c2, mat, eps_s, udof, &o);
    constexpr int NP = 21;   // revariate{1}((;X,U)) over ND*6 X entries then 3 U entries
    using TX = Dual<NP, double>;
    TX Xd[3][12];
    for (int d = 0; d < ND; ++d) for (int i = 0; i < 6; ++i) Xd[d][i] = seed<NP, double>(X[d * 6 + i], d * 6 + i);
    TX Ud[3];
    for (int c = 0; c < 3; ++c) Ud[c] = udof ? seed<NP, double>(U[c], ND * 6 + c) : TX(0.);
    TX Rr[6];
    if (ND == 1) bar_residual<TX, 1>(o, Xd, udof ? Ud : nullptr, t, Rr);
    else if (ND == 2) bar_residual<TX, 2>(o, Xd, udof ? Ud : nullptr, t, Rr);
    else bar_residual<TX, 3>(o, Xd, udof ? Ud : nullptr, t, Rr);
    for (int i = 0; i < 6; ++i) {
        R[i] = Rr[i].x;
        for (int d = 0; d < ND; ++d) for (int j = 0; j < 6; ++j) dRdX[d * 36 + i + 6 * j] = Rr[i].dx[d * 6 + j];
        if (udof && dRdU) for (int c = 0; c < 3; ++c) dRdU[i + 6 * c] = Rr[i].dx[ND * 6 + c];
    }
    return 0;
}

// diffed_residual of one beam (src/Diagnostic.jl:929-961): R (12), and
// dR/dX[der] for der < ND (each 12×12 column-major), dR/dU (12×3) if udof.
// X: ND×12, U: 3 (Udof only). mat: 16 BeamCrossSection fields.
int orc_beam_diffed_residual(const double* c1, const double* c2, const double* orient2, const double* mat16, int udof,
                             int ND, const double* X, const double* U, double* R, double* dRdX, double* dRdU) {
    BeamCrossSection mat; std::memcpy(&mat, mat16, sizeof(mat));
    EulerBeam3D o;
    if (make_beam(c1, c2, orient2, mat, udof, &o)) return 1;
    // revariate{1}((;X,U,A)): partials over ND*12 X entries then U (Diagnostic.jl:951)
    constexpr int NP = 39;
    using TX = Dual<NP, double>;
    TX Xd[3][12];
    for (int d = 0; d < ND; ++d) for (int i = 0; i < 12; ++i) Xd[d][i] = seed<NP, double>(X[d * 12 + i], d * 12 + i);
    TX Rr[12];
    if (udof) {
        TX Ud[1][3]; for (int c = 0; c < 3; ++c) Ud[0][c] = seed<NP, double>(U[c], ND * 12 + c);
        if (ND == 1) beam_residual<TX, 1, 1>(o, Xd, Ud, Rr);
        else if (ND == 2) beam_residual<TX, 2, 1>(o, Xd, Ud, Rr);
        else beam_residual<TX, 3, 1>(o, Xd, Ud, Rr);
    } else {
        if (ND == 1) beam_residual<TX, 1, 1>(o, Xd, nullptr, Rr);
        else if (ND == 2) beam_residual<TX, 2, 1>(o, Xd, nullptr, Rr);
        else beam_residual<TX, 3, 1>(o, Xd, nullptr, Rr);
    }
    for (int i = 0; i < 12; ++i) {
        R[i] = Rr[i].x;
        for (int d = 0; d < ND; ++d) for (int j = 0; j < 12; ++j) dRdX[d * 144 + i + 12 * j] = Rr[i].dx[d * 12 + j];
        if (udof && dRdU) for (int c = 0; c < 3; ++c) dRdU[i + 12 * c] = Rr[i].dx[ND * 12 + c];
    }
    return 0;
}

// residual with plain-double X (P=1 path, BeamElement.jl:151-174), ND=1..3.
// Also returns the requestables ε and ♢κ (mid-element curvature).
int orc_beam_residual_plain(const double* c1, const double* c2, const double* orient2, const double* mat16, int ND,
                            const double* X, double* R, double* eps, double* kappa_mid) {
    BeamCrossSection mat; std::memcpy(&mat, mat16, sizeof(mat));
    EulerBeam3D o;
    if (make_beam(c1, c2, orient2, mat, 0, &o)) return 1;
    double Xd[3][12];
    for (int d = 0; d < ND; ++d) for (int i = 0; i < 12; ++i) Xd[d][i] = X[d * 12 + i];
    BeamReq<double> rq;
    if (ND == 1) beam_residual<double, 1, 1>(o, Xd, nullptr, R, &rq);
    else if (ND == 2) beam_residual<double, 2, 1>(o, Xd, nullptr, R, &rq);
    else beam_residual<double, 3, 1>(o, Xd, nullptr, R, &rq);
    if (eps) *eps = rq.eps;
    if (kappa_mid) for (int c = 0; c < 3; ++c) kappa_mid[c] = rq.kappa_mid[c];
    return 0;
}

// All requestables of EulerBeam3D at X (ND time levels) and U (3 or none), in the device's
// layout (include/muscade_hip.h, mh_beam_results): ε[ND], κ[ND][3], then per Gauss point
// x[ND][3], κgp[ND][3], fᵢ, mᵢ[3], fₑ[3], mₑ[3].
int orc_beam_results(const double* c1, const double* c2, const double* orient2, const double* mat16, int ND,
                     const double* X, const double* U, double* out) {
    BeamCrossSection mat; std::memcpy(&mat, mat16, sizeof(mat));
    EulerBeam3D o;
    if (make_beam(c1, c2, orient2, mat, U ? 1 : 0, &o)) return 1;
    double Xd[3][12], Ud[1][3] = {{0., 0., 0.}}, R[12];
    for (int d = 0; d < ND; ++d) for (int i = 0; i < 12; ++i) Xd[d][i] = X[d * 12 + i];
    if (U) for (int c = 0; c < 3; ++c) Ud[0][c] = U[c];
    BeamRes rs;
    if (ND == 1) beam_residual<double, 1, 1>(o, Xd, Ud, R, nullptr, &rs);
    else if (ND == 2) beam_residual<double, 2, 1>(o, Xd, Ud, R, nullptr, &rs);
    else beam_residual<double, 3, 1>(o, Xd, Ud, R, nullptr, &rs);
    int q = 0;
    for (int d = 0; d < ND; ++d) out[q++] = rs.eps[d];
    for (int d = 0; d < ND; ++d) for (int c = 0; c < 3; ++c) out[q++] = rs.kmid[d][c];
    for (int g = 0; g < 4; ++g) {
        for (int d = 0; d < ND; ++d) for (int c = 0; c < 3; ++c) out[q++] = rs.x[g][d][c];
        for (int d = 0; d < ND; ++d) for (int c = 0; c < 3; ++c) out[q++] = rs.kgp[g][d][c];
        out[q++] = rs.fi[g];
        for (int c = 0; c < 3; ++c) out[q++] = rs.mi[g][c];
        for (int c = 0; c < 3; ++c) out[q++] = rs.fe[g][c];
        for (int c = 0; c < 3; ++c) out[q++] = rs.me[g][c];
    }
    return 0;
}

// εₐₓ = ε and κ = ♢κ with their first and second derivatives w.r.t. X₀, over the
// restated kinematics on second-order nested duals (revariate{2}, Taylor.jl:143-151).
// out (364): values (4), gradients (4×12), Hessians (4×78: upper triangle, row-major,
// (0,0), (0,1) … (0,11), (1,1) …).
int orc_beam_gauge2(const double* c1, const double* c2, const double* orient2, const double* mat16, const double* X,
                    double* out) {
    BeamCrossSection mat; std::memcpy(&mat, mat16, sizeof(mat));
    EulerBeam3D o;
    if (make_beam(c1, c2, orient2, mat, 0, &o)) return 1;
    using T = Dual<12, Dual<12, double>>;
    T X0[12];
    for (int i = 0; i < 12; ++i) { X0[i].x = seed<12, double>(X[i], i); X0[i].dx[i] = Dual<12, double>(1.); }
    Kin<T> k = kinematics<T>(o, X0);
    const double f = 2. / o.L;
    T q[4] = {k.eps, k.vl2[0] * f, k.vl2[2] * f, k.vl2[1] * (-f)};
    for (int a = 0; a < 4; ++a) {
        out[a] = q[a].x.x;
        for (int i = 0; i < 12; ++i) out[4 + 12 * a + i] = q[a].dx[i].x;
        int pr = 0;
        for (int i = 0; i < 12; ++i)
            for (int j = i; j < 12; ++j) out[52 + 78 * a + pr++] = q[a].dx[i].dx[j];
    }
    return 0;
}

// AnchorLine (test/SomeElements.jl:54-94) at x (tx1, tx2, rx3) and A (ΔL, Δbuoyancy): R (3),
// its gradient (3×5, row-major, partials x then A), the requestables Fh and cr, Fh's gradient
// (5) and Hessian (5×5). p10 = xmtop(3), Δxmtop(3), xmbot(2), L, buoyancy.
int orc_anchor_eval(const double* p10, const double* x, const double* A, double* R, double* dR, double* Fh,
                    double* dFh, double* HFh, double* cr) {
    AnchorLine o;
    for (int i = 0; i < 3; ++i) { o.xmtop[i] = p10[i]; o.dxmtop[i] = p10[3 + i]; }
    o.xmbot[0] = p10[6]; o.xmbot[1] = p10[7]; o.L = p10[8]; o.buoyancy = p10[9];
    using T = Dual<5, Dual<5, double>>;
    T v[5];
    for (int i = 0; i < 5; ++i) {
        v[i].x = seed<5, double>(i < 3 ? x[i] : A[i - 3], i);
        v[i].dx[i] = Dual<5, double>(1.);
    }
    T Rr[3], fh, c;
    anchorline_eval<T>(o, v, v + 3, Rr, &fh, &c);
    for (int k = 0; k < 3; ++k) {
        R[k] = Rr[k].x.x;
        for (int i = 0; i < 5; ++i) dR[5 * k + i] = Rr[k].dx[i].x;
    }
    *Fh = fh.x.x;
    *cr = c.x.x;
    for (int i = 0; i < 5; ++i) {
        dFh[i] = fh.dx[i].x;
        for (int j = 0; j < 5; ++j) HFh[5 * i + j] = fh.dx[i].dx[j];
    }
    return 0;
}

// ε and ♢κ with their gradients w.r.t. X₀ (12): the strain-gauge requestables
// (toolbox/StrainGaugeOnBeamElement.jl:70-76) differentiated as by
// diffed_lagrangian{1} (Diagnostic.jl:851-911). Also R and the first-order
// residual partials K (12×12). out: eps, deps(12), kap(3), dkap(3×12 row-major by comp)
int orc_beam_gauge(const double* c1, const double* c2, const double* orient2, const double* mat16, const double* X,
                   double* R, double* K, double* eps, double* deps, double* kap, double* dkap) {
    BeamCrossSection mat; std::memcpy(&mat, mat16, sizeof(mat));
    EulerBeam3D o;
    if (make_beam(c1, c2, orient2, mat, 0, &o)) return 1;
    using TX = Dual<12, double>;
    TX Xd[1][12];
    for (int i = 0; i < 12; ++i) Xd[0][i] = seed<12, double>(X[i], i);
    TX Rr[12];
    BeamReq<TX> rq;
    beam_residual<TX, 1, 1>(o, Xd, nullptr, Rr, &rq);
    for (int i = 0; i < 12; ++i) { R[i] = Rr[i].x; for (int j = 0; j < 12; ++j) K[i + 12 * j] = Rr[i].dx[j]; }
    *eps = rq.eps.x;
    for (int j = 0; j < 12; ++j) deps[j] = rq.eps.dx[j];
    for (int c = 0; c < 3; ++c) { kap[c] = rq.kappa_mid[c].x; for (int j = 0; j < 12; ++j) dkap[c * 12 + j] = rq.kappa_mid[c].dx[j]; }
    return 0;
}

// ---------------------------------------------------------------- dof numbering
// addelement! (ModelDescription.jl:195-263) for a sequence of single-element calls.
// Inputs: nel elements, element e has type key tkey[e] (concrete-type identity),
// nodes nodes[noff[e] .. noff[e+1]); each type key k has a doflist given by
// dl_off[k]..dl_off[k+1] entries of (inod 1-based, class 0/1/2, field code).
// Outputs: ieletyp[e] (1-based, first-appearance order), iele[e] (1-based within type),
// dofid[doff[e]+k] = class-local 1-based dof number of eledof k (doff = prefix of doflist lengths),
// ndof[3] counts per class, and per-dof (class-local order) node and field:
// dofnode_c / doffield_c for c=X,U,A (arrays of capacity cap).
int orc_number_dofs(int nel, const int* tkey, const int* noff, const int* nodes, int ntkey, const int* dl_off,
                    const int* dl_inod, const int* dl_class, const int* dl_field, int* ieletyp, int* iele, int* dofid,
                    int64_t* ndof, int cap, int* dofnode_x, int* doffield_x, int* dofnode_u, int* doffield_u,
                    int* dofnode_a, int* doffield_a) {
    std::vector<int> type_of_key(ntkey, 0), count(1, 0);
    int ntype = 0;
    std::vector<std::pair<int, int>> doftyp;   // (class, field) in first-appearance order
    // node → list of (idoftyp, class, class-local id)
    std::vector<std::vector<std::tuple<int, int, int>>> nodedofs;
    int64_t n[3] = {0, 0, 0};
    int* dn[3] = {dofnode_x, dofnode_u, dofnode_a};
    int* df[3] = {doffield_x, doffield_u, doffield_a};
    int dpos = 0;
    for (int e = 0; e < nel; ++e) {
        int k = tkey[e];
        if (type_of_key[k] == 0) { type_of_key[k] = ++ntype; count.push_back(0); }
        int t = type_of_key[k];
        ieletyp[e] = t;
        iele[e] = ++count[t];
        for (int q = dl_off[k]; q < dl_off[k + 1]; ++q) {
            int nod = nodes[noff[e] + dl_inod[q] - 1];
            int cls = dl_class[q], fld = dl_field[q];
            int idt = -1;
            for (size_t s = 0; s < doftyp.size(); ++s) if (doftyp[s].first == cls && doftyp[s].second == fld) { idt = (int)s; break; }
            if (idt < 0) { doftyp.push_back({cls, fld}); idt = (int)doftyp.size() - 1; }
            if ((int)nodedofs.size() <= nod) nodedofs.resize(nod + 1);
            int id = -1;
            for (auto& d : nodedofs[nod]) if (std::get<0>(d) == idt) { id = std::get<2>(d); break; }
            if (id < 0) {
                id = (int)(++n[cls]);
                if (id > cap) return 2;
                nodedofs[nod].push_back({idt, cls, id});
                if (dn[cls]) { dn[cls][id - 1] = nod; df[cls][id - 1] = fld; }
            }
            dofid[dpos++] = id;
        }
    }
    for (int c = 0; c < 3; ++c) ndof[c] = n[c];
    return 0;
}

// ---------------------------------------------------------------- model
void* orc_model_new(int64_t nX, int64_t nU, int64_t nA) {
    OModel* m = new OModel; m->nX = nX; m->nU = nU; m->nA = nA; return m;
}
void orc_model_free(void* h) { delete (OModel*)h; }

// Append an element type (one Muscade eletype). idx*: 1-based model dofs, nele×n (element-major).
// params per kind:
//   K_BEAM    : nele × 25 = c1(3) c2(3) orient2(3) mat(16)
//   K_BAR     : nele × 16 = c1(3) c2(3) mat(9: EA μ w Caₜ Clₜ Cqₜ Caₙ Clₙ Cqₙ) ϵₛ
//   K_DOFLOAD : nele × 1  = F
//   K_HOLD    : nele × 1  = g, the prescribed value (gap x - g), or none (Hold: g = 0)
//   K_TURBINE : nele × 5  = xm(2) z seadrag skydrag
//   K_ANCHOR  : nele × 10 = xmtop(3) dxmtop(3) xmbot(2) L buoyancy
int orc_model_add_type(void* h, int kind, int nele, int nX, int nU, int nA, const int64_t* iX, const int64_t* iU,
                       const int64_t* iA, const double* scaleX, const double* params, int udof) {
    OModel* m = (OModel*)h;
    EType t; t.kind = kind; t.nele = nele; t.nX = nX; t.nU = nU; t.nA = nA; t.udof = udof;
    t.iX.assign(iX, iX + (size_t)nele * nX);
    if (nU) t.iU.assign(iU, iU + (size_t)nele * nU);
    if (nA) t.iA.assign(iA, iA + (size_t)nele * nA);
    t.scaleX.assign(nX, 1.);
    if (scaleX) t.scaleX.assign(scaleX, scaleX + nX);
    t.scaleL = t.scaleX;
    t.scaleU.assign(nU, 1.);
    if (kind == K_BEAM) {
        t.beams.resize(nele);
        for (int e = 0; e < nele; ++e) {
            const double* p = params + 25 * (size_t)e;
            BeamCrossSection mat; std::memcpy(&mat, p + 9, sizeof(mat));
            if (make_beam(p, p + 3, p + 6, mat, udof, &t.beams[e])) return 1;
        }
    } else if (kind == K_BAR) {
        t.bars.resize(nele);
        for (int e = 0; e < nele; ++e) {
            const double* p = params + 16 * (size_t)e;
            AxisymmetricBarCrossSection mat; std::memcpy(&mat, p + 6, sizeof(mat));
            make_bar(p, p + 3, mat, p[15], udof, &t.bars[e]);
        }
    } else if (kind == K_DOFLOAD || (kind == K_HOLD && params)) {   // loads; prescribed values of a held dof
        t.F.assign(params, params + nele);
    } else if (kind == K_TURBINE) {
        t.turbines.resize(nele);
        for (int e = 0; e < nele; ++e) { const double* p = params + 5 * (size_t)e;
            t.turbines[e] = Turbine{{p[0], p[1]}, p[2], p[3], p[4]}; }
    } else if (kind == K_ANCHOR) {
        t.anchors.resize(nele);
        for (int e = 0; e < nele; ++e) { const double* p = params + 10 * (size_t)e;
            AnchorLine a; for (int i = 0; i < 3; ++i) { a.xmtop[i] = p[i]; a.dxmtop[i] = p[3 + i]; }
            a.xmbot[0] = p[6]; a.xmbot[1] = p[7]; a.L = p[8]; a.buoyancy = p[9]; t.anchors[e] = a; }
    }
    m->types.push_back(std::move(t));
    return 0;
}
void orc_model_set_time(void* h, double t) { ((OModel*)h)->t = t; }
int orc_model_set_dofload(void* h, int itype, const double* F) {
    OModel* m = (OModel*)h; EType& t = m->types[itype];
    if (t.kind != K_DOFLOAD) return 1;
    std::copy(F, F + t.nele, t.F.begin());
    return 0;
}

// prepare(AssemblySweepX) (SweepX.jl:24-33): asmvec! over allXdofs, then asmmat!.
// Returns nnz; pairs are sorted lexicographically on (j,i) with ties kept in
// traversal order, as sortperm does (Assemble.jl:373-448).
int64_t orc_model_prepare_sweepx(void* h) {
    OModel* m = (OModel*)h;
    // asmvec!: with allXdofs, X[idof] = idof, so asm[1] = index.X (Assemble.jl:347-357)
    for (auto& t : m->types) t.asm1 = t.iX;
    // asmmat! step 1-2
    int64_t npair = 0;
    for (auto& t : m->types) for (int e = 0; e < t.nele; ++e) {
        int a = 0; for (int k = 0; k < t.nX; ++k) a += t.asm1[(size_t)e * t.nX + k] != 0;
        npair += (int64_t)a * a;
    }
    std::vector<std::pair<int64_t, int64_t>> A(npair);
    int64_t ip = 0;
    for (auto& t : m->types) for (int e = 0; e < t.nele; ++e)
        for (int j = 0; j < t.nX; ++j) for (int i = 0; i < t.nX; ++i) {
            int64_t ii = t.asm1[(size_t)e * t.nX + i], jj = t.asm1[(size_t)e * t.nX + j];
            if (ii != 0 && jj != 0) A[ip++] = {jj, ii};
        }
    // step 3: sortperm (stable)
    std::vector<int64_t> I(npair);
    std::iota(I.begin(), I.end(), 0);
    std::stable_sort(I.begin(), I.end(), [&](int64_t a, int64_t b) { return A[a] < A[b]; });
    // step 4
    int64_t nnz = 0;
    for (int64_t p = 0; p < npair; ++p) if (p == 0 || A[I[p]] != A[I[p - 1]]) ++nnz;
    std::vector<int64_t> K(npair);
    m->colptr.assign(m->nX + 1, 0);
    m->rowval.assign(nnz, 0);
    m->colptr[0] = 1;
    int64_t inz = 0, icol = 1;
    for (int64_t p = 0; p < npair; ++p) {
        if (p == 0 || A[I[p]] != A[I[p - 1]]) {
            ++inz;
            int64_t j = A[I[p]].first, i = A[I[p]].second;
            m->rowval[inz - 1] = i;
            while (j > icol) { ++icol; m->colptr[icol - 1] = inz; }
        }
        K[I[p]] = inz;
    }
    for (int64_t i = icol + 1; i <= m->nX + 1; ++i) m->colptr[i - 1] = nnz + 1;
    // step 5
    ip = 0;
    for (auto& t : m->types) {
        t.asm2.assign((size_t)t.nele * t.nX * t.nX, 0);
        for (int e = 0; e < t.nele; ++e)
            for (int j = 0; j < t.nX; ++j) for (int i = 0; i < t.nX; ++i) {
                int64_t ii = t.asm1[(size_t)e * t.nX + i], jj = t.asm1[(size_t)e * t.nX + j];
                if (ii != 0 && jj != 0) t.asm2[(size_t)e * t.nX * t.nX + i + t.nX * j] = K[ip++];
            }
    }
    m->nnz = nnz;
    return nnz;
}
void orc_model_get_pattern(void* h, int64_t* colptr, int64_t* rowval) {
    OModel* m = (OModel*)h;
    std::copy(m->colptr.begin(), m->colptr.end(), colptr);
    std::copy(m->rowval.begin(), m->rowval.end(), rowval);
}
void orc_model_get_asm(void* h, int itype, int64_t* asm1, int64_t* asm2) {
    OModel* m = (OModel*)h; EType& t = m->types[itype];
    if (asm1) std::copy(t.asm1.begin(), t.asm1.end(), asm1);
    if (asm2) std::copy(t.asm2.begin(), t.asm2.end(), asm2);
}

}  // extern "C"

// assemble!{mission}(AssemblySweepX{OX}) (Assemble.jl:470-505 + SweepX.jl:39-96).
// c = (a1,a2,a3,b1,b2,b3) Newmark coefficients; mission 0 = :step, 1 = :iter.
// X0..X2: model X vectors (X1,X2 used for OX>=1,2); U, A model vectors; t time.
// On a NaN the element is reported through err_type/err_ele (0-based) and 3 is returned.
template <int OX>
static int assemble_sweepx(OModel* m, int mission, const double* c, const double* const* Xs, const double* U,
                           const double* A, double* rhs, double* nzval, int* err_type, int* err_ele) {
    std::fill(rhs, rhs + m->nX, 0.);
    std::fill(nzval, nzval + m->nnz, 0.);
    constexpr int ND = OX + 1;
    const bool step = (mission == 0) && OX > 0;
    for (size_t it = 0; it < m->types.size(); ++it) {
        const EType& t = m->types[it];
        const int nX = t.nX;
        if (nX == 0) continue;   // SweepX.jl:39-44
        for (int e = 0; e < t.nele; ++e) {
            double x[3][12];
            for (int d = 0; d < ND; ++d) for (int k = 0; k < nX; ++k) x[d][k] = Xs[d][t.iX[(size_t)e * nX + k] - 1];
            double Ue[3] = {0, 0, 0}, Ae[4] = {0, 0, 0, 0};
            for (int k = 0; k < t.nU && k < 3; ++k) Ue[k] = U[t.iU[(size_t)e * t.nU + k] - 1];
            for (int k = 0; k < t.nA && k < 4; ++k) Ae[k] = A[t.iA[(size_t)e * t.nA + k] - 1];
            const int64_t* a1 = &t.asm1[(size_t)e * nX];
            const int64_t* a2 = &t.asm2[(size_t)e * nX * nX];
            if (step) {
                // :step — reδ{1}((;X=x,r=0.),(;X=scale.X,r=1.)): 13 partials (SweepX.jl:45-59)
                using TX = Dual<13, double>;
                TX Xd[3][12];
                for (int k = 0; k < nX; ++k) {
                    TX dX; dX.dx[k] = t.scaleX[k];
                    TX dr; dr.dx[12] = 1.;
                    (void)dr;
                    double a = 0., b = 0.;
                    if (OX >= 1) a = c[1] * x[1][k] + (OX >= 2 ? c[2] * x[2][k] : 0.);
                    if (OX >= 2) b = c[4] * x[1][k] + c[5] * x[2][k];
                    Xd[0][k] = x[0][k] + dX;
                    if (OX >= 1) { TX v = x[1][k] + c[0] * dX; v.dx[12] = a; Xd[1][k] = v; }
                    if (OX >= 2) { TX v = x[2][k] + c[3] * dX; v.dx[12] = b; Xd[2][k] = v; }
                }
                TX R[12];
                elem_residual<TX, ND>(t, e, Xd, Ue, Ae, m->t, R);
                for (int i = 0; i < nX; ++i) if (has_nan(R[i])) { *err_type = (int)it; *err_ele = e; return 3; }
                for (int i = 0; i < nX; ++i) R[i] = R[i] * t.scaleX[i];
                for (int i = 0; i < nX; ++i) if (a1[i]) rhs[a1[i] - 1] += R[i].x;            // add_value!
                for (int i = 0; i < nX; ++i) if (a1[i]) rhs[a1[i] - 1] -= R[i].dx[12];       // add_∂!{1,:minus} (Nx+1,)
                for (int i = 0; i < nX; ++i) for (int j = 0; j < nX; ++j) {
                    int64_t k = a2[i + nX * j]; if (k) nzval[k - 1] += R[i].dx[j]; }
            } else {
                using TX = Dual<12, double>;
                TX Xd[3][12];
                for (int k = 0; k < nX; ++k) {
                    TX dX; dX.dx[k] = t.scaleX[k];   // δ{1,Nx,𝕣}(scale.X)
                    Xd[0][k] = x[0][k] + dX;
                    if (OX >= 1) Xd[1][k] = x[1][k] + c[0] * dX;
                    if (OX >= 2) Xd[2][k] = x[2][k] + c[3] * dX;
                }
                TX R[12];
                elem_residual<TX, ND>(t, e, Xd, Ue, Ae, m->t, R);
                for (int i = 0; i < nX; ++i) if (has_nan(R[i])) { *err_type = (int)it; *err_ele = e; return 3; }
                for (int i = 0; i < nX; ++i) R[i] = R[i] * t.scaleX[i];
                for (int i = 0; i < nX; ++i) if (a1[i]) rhs[a1[i] - 1] += R[i].x;
                for (int i = 0; i < nX; ++i) for (int j = 0; j < nX; ++j) {
                    int64_t k = a2[i + nX * j]; if (k) nzval[k - 1] += R[i].dx[j]; }
            }
        }
    }
    return 0;
}
// The same static assemble!{:iter} (OX = 0) with the element evaluations spread over
// nthreads OpenMP threads (the all-cores CPU baseline): each element's scaled R and
// ∂R/∂δX go to its own slot, then add_value!/add_∂! run sequentially in (ieletyp, iele)
// order, so rhs and nzval equal the sequential assembly bit for bit.
static int assemble_sweepx0_omp(OModel* m, int nthreads, const double* X0, const double* U, const double* A,
                                double* rhs, double* nzval, int* err_type, int* err_ele) {
    std::fill(rhs, rhs + m->nX, 0.);
    std::fill(nzval, nzval + m->nnz, 0.);
    using TX = Dual<12, double>;
    for (size_t it = 0; it < m->types.size(); ++it) {
        const EType& t = m->types[it];
        const int nX = t.nX;
        if (nX == 0) continue;
        std::vector<double> Rv((size_t)t.nele * nX), Kv((size_t)t.nele * nX * nX);
        std::vector<int> bad(t.nele, 0);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
        for (int e = 0; e < t.nele; ++e) {
            double Ue[3] = {0, 0, 0}, Ae[4] = {0, 0, 0, 0};
            for (int k = 0; k < t.nU && k < 3; ++k) Ue[k] = U[t.iU[(size_t)e * t.nU + k] - 1];
            for (int k = 0; k < t.nA && k < 4; ++k) Ae[k] = A[t.iA[(size_t)e * t.nA + k] - 1];
            TX Xd[3][12];
            for (int k = 0; k < nX; ++k) {
                TX dX; dX.dx[k] = t.scaleX[k];
                Xd[0][k] = X0[t.iX[(size_t)e * nX + k] - 1] + dX;
            }
            TX R[12];
            elem_residual<TX, 1>(t, e, Xd, Ue, Ae, m->t, R);
            for (int i = 0; i < nX; ++i) bad[e] |= has_nan(R[i]);
            for (int i = 0; i < nX; ++i) {
                const TX r = R[i] * t.scaleX[i];
                Rv[(size_t)e * nX + i] = r.x;
                for (int j = 0; j < nX; ++j) Kv[(size_t)e * nX * nX + i + (size_t)nX * j] = r.dx[j];
            }
        }
        for (int e = 0; e < t.nele; ++e) {
            if (bad[e]) { *err_type = (int)it; *err_ele = e; return 3; }
            const int64_t* a1 = &t.asm1[(size_t)e * nX];
            const int64_t* a2 = &t.asm2[(size_t)e * nX * nX];
            for (int i = 0; i < nX; ++i) if (a1[i]) rhs[a1[i] - 1] += Rv[(size_t)e * nX + i];
            for (int i = 0; i < nX; ++i) for (int j = 0; j < nX; ++j) {
                int64_t k = a2[i + nX * j]; if (k) nzval[k - 1] += Kv[(size_t)e * nX * nX + i + (size_t)nX * j]; }
        }
    }
    return 0;
}

extern "C" {
int orc_model_assemble_sweepx0_omp(void* h, int nthreads, const double* X0, const double* U, const double* A,
                                   double* rhs, double* nzval, int* err_type, int* err_ele) {
    return assemble_sweepx0_omp((OModel*)h, nthreads, X0, U, A, rhs, nzval, err_type, err_ele);
}
int orc_model_assemble_sweepx(void* h, int OX, int mission, const double* c, const double* X0, const double* X1,
                              const double* X2, const double* U, const double* A, double* rhs, double* nzval,
                              int* err_type, int* err_ele) {
    OModel* m = (OModel*)h;
    const double* Xs[3] = {X0, X1, X2};
    if (OX == 0) return assemble_sweepx<0>(m, mission, c, Xs, U, A, rhs, nzval, err_type, err_ele);
    if (OX == 1) return assemble_sweepx<1>(m, mission, c, Xs, U, A, rhs, nzval, err_type, err_ele);
    if (OX == 2) return assemble_sweepx<2>(m, mission, c, Xs, U, A, rhs, nzval, err_type, err_ele);
    return 1;
}

}  // extern "C"

// ================================================================ AssemblyDirect (DirectXUA)
// The element-level maps of one class (asmvec!, Assemble.jl:340-357): with the all-dofs
// groups of DirectXUA.jl:23, Λ and X use index.X, U uses index.U.
static const std::vector<int64_t>& class_index(const EType& t, int cls) { return cls == 2 ? t.iU : t.iX; }
static int class_ndof(const EType& t, int cls) { return cls == 2 ? t.nU : t.nX; }

// asmmat! for (α,β): pairs (jmod, imod) in traversal order, stable sort, run-length
// numbering, asm2 per type (Assemble.jl:373-448).
static void build_dpattern(OModel* m, int a, int b, DPattern& P) {
    P.nr = a == 2 ? m->nU : m->nX;
    P.nc = b == 2 ? m->nU : m->nX;
    std::vector<std::pair<int64_t, int64_t>> A;
    for (auto& t : m->types) {
        const int na = class_ndof(t, a), nb = class_ndof(t, b);
        const auto& ia = class_index(t, a); const auto& ib = class_index(t, b);
        for (int e = 0; e < t.nele; ++e)
            for (int j = 0; j < nb; ++j) for (int i = 0; i < na; ++i)
                A.push_back({ib[(size_t)e * nb + j], ia[(size_t)e * na + i]});
    }
    const int64_t npair = (int64_t)A.size();
    std::vector<int64_t> I(npair);
    std::iota(I.begin(), I.end(), 0);
    std::stable_sort(I.begin(), I.end(), [&](int64_t x, int64_t y) { return A[x] < A[y]; });
    std::vector<int64_t> K(npair);
    P.colptr.assign(P.nc + 1, 0);
    P.rowval.clear();
    P.colptr[0] = 1;
    int64_t inz = 0, icol = 1;
    for (int64_t q = 0; q < npair; ++q) {
        if (q == 0 || A[I[q]] != A[I[q - 1]]) {
            ++inz;
            P.rowval.push_back(A[I[q]].second);
            while (A[I[q]].first > icol) { ++icol; P.colptr[icol - 1] = inz; }
        }
        K[I[q]] = inz;
    }
    for (int64_t i = icol + 1; i <= P.nc + 1; ++i) P.colptr[i - 1] = inz + 1;
    P.nnz = inz;
    P.asm2.assign(m->types.size(), {});
    int64_t ip = 0;
    for (size_t it = 0; it < m->types.size(); ++it) {
        auto& t = m->types[it];
        const int na = class_ndof(t, a), nb = class_ndof(t, b);
        P.asm2[it].assign((size_t)t.nele * na * nb, 0);
        for (int e = 0; e < t.nele; ++e)
            for (int j = 0; j < nb; ++j) for (int i = 0; i < na; ++i)
                P.asm2[it][(size_t)e * na * nb + i + (size_t)na * j] = K[ip++];
    }
}

// Derivative orders of a class in AssemblyDirect{OX,OU,IA} (nder, DirectXUA.jl:25)
static int dnder(const OModel* m, int cls) { return cls == 0 ? 1 : (cls == 1 ? m->dOX + 1 : m->dOU + 1); }
// Does L2[α,β] hold matrices (DirectXUA.jl:37-48; no independence flags)
static bool dhas(int a, int b) { return !(a == 0 && b == 0); }

// Offsets of the flat outputs: L1 = [Λ | X·(OX+1) | U·(OU+1)], L2 = pairs (α,β) in
// row-major class order, each nder(α)×nder(β) blocks (i-major) of nnz(α,β) values.
static int64_t l1_offset(const OModel* m, int cls) {
    return cls == 0 ? 0 : (cls == 1 ? m->nX : m->nX + (int64_t)(m->dOX + 1) * m->nX);
}
static int64_t l2_offset(const OModel* m, int a, int b) {
    int64_t o = 0;
    for (int x = 0; x < 3; ++x) for (int y = 0; y < 3; ++y) {
        if (x == a && y == b) return o;
        if (dhas(x, y)) o += (int64_t)dnder(m, x) * dnder(m, y) * m->dpat[x][y].nnz;
    }
    return o;
}

template <class TX, int ND>
static void direct_residual(const EType& t, int e, const TX (*X)[12], const TX* U, double time, TX* R) {
    if (t.kind == K_BEAM) {
        if (t.udof) { TX Uj[1][3]; for (int c = 0; c < 3; ++c) Uj[0][c] = U[c]; beam_residual<TX, ND, 1>(t.beams[e], X, Uj, R); }
        else beam_residual<TX, ND, 1>(t.beams[e], X, nullptr, R);
    } else if (t.kind == K_BAR) {
        bar_residual<TX, ND>(t.bars[e], X, t.udof ? U : nullptr, time, R);
    } else if (t.kind == K_HOLD) {
        hold_residual<TX>(X[0], t.F.empty() ? 0. : t.F[e], R);
    } else if (t.kind == K_DOFLOAD) {
        dofload_residual<TX>(t.F[e], R);
    }
}

// assemble!{mission}(::AssemblyDirect{OX,OU,IA}) for one step (DirectXUA.jl:85-150):
// beams and bars by the residual path (no_second_order, :85-120), Hold and DofLoad by
// the Lagrangian path L = Λ∘₁R with second-order partials (:121-171).
template <int OX>
static int assemble_direct(OModel* m, int mission, double dt, const double* Lam, const double* const* Xs,
                           const double* const* Us, double time, double* L1, double* L2) {
    constexpr int ND = OX + 1;
    const int OU = m->dOU;
    const int64_t nL1 = l1_offset(m, 2) + (int64_t)(OU + 1) * m->nU;
    std::fill(L1, L1 + nL1, 0.);
    const int64_t nL2 = l2_offset(m, 3, 0);
    std::fill(L2, L2 + nL2, 0.);
    const bool mats = mission == 1;
    for (size_t it = 0; it < m->types.size(); ++it) {
        const EType& t = m->types[it];
        const int Nx = t.nX, Nu = t.nU;
        if (t.kind == K_ZERO || t.kind == K_TURBINE || t.kind == K_ANCHOR) return 4;   // not restated here
        for (int e = 0; e < t.nele; ++e) {
            double x[3][12] = {}, u[3][3] = {}, lam[12] = {};
            for (int d = 0; d < ND; ++d) for (int k = 0; k < Nx; ++k) x[d][k] = Xs[d][t.iX[(size_t)e * Nx + k] - 1];
            for (int d = 0; d <= OU; ++d) for (int k = 0; k < Nu; ++k) u[d][k] = Us[d][t.iU[(size_t)e * Nu + k] - 1];
            for (int k = 0; k < Nx; ++k) lam[k] = Lam[t.iX[(size_t)e * Nx + k] - 1];
            const int64_t* aL = &t.iX[(size_t)e * Nx];
            const int64_t* aU = Nu ? &t.iU[(size_t)e * Nu] : nullptr;
            if (t.kind == K_BEAM || t.kind == K_BAR) {
                // residual path: revariate{1}((;X,U),(;X=scale.X,U=scale.U)) — partials X₀..X_OX, U₀..U_OU
                using TX = Dual<45, double>;
                TX Xd[3][12], Ud[3];
                for (int d = 0; d < ND; ++d) for (int k = 0; k < Nx; ++k) {
                    TX v(x[d][k]); v.dx[d * Nx + k] = t.scaleX[k]; Xd[d][k] = v; }
                for (int k = 0; k < Nu; ++k) { TX v(u[0][k]); v.dx[ND * Nx + k] = t.scaleU[k]; Ud[k] = v; }
                TX R[12];
                direct_residual<TX, ND>(t, e, Xd, Ud, time, R);
                for (int i = 0; i < Nx; ++i) if (has_nan(R[i])) return 3;
                for (int i = 0; i < Nx; ++i) L1[aL[i] - 1] += R[i].x * dt;
                if (!mats) continue;
                for (int b = 1; b <= 2; ++b) {
                    const int Nb = b == 1 ? Nx : Nu;
                    for (int j = 0; j < dnder(m, b); ++j) {
                        const int pbase = b == 1 ? j * Nx : ND * Nx + j * Nu;
                        double* Lab = L2 + l2_offset(m, 0, b) + (int64_t)j * m->dpat[0][b].nnz;
                        double* Lba = L2 + l2_offset(m, b, 0) + (int64_t)j * dnder(m, 0) * m->dpat[b][0].nnz;
                        const auto& A0b = m->dpat[0][b].asm2[it];
                        const auto& Ab0 = m->dpat[b][0].asm2[it];
                        for (int i = 0; i < Nx; ++i) for (int k = 0; k < Nb; ++k) {
                            const bool ok = b == 1 || j <= OU;
                            const double v = (ok && (b == 1 ? j < ND : j == 0)) ? R[i].dx[pbase + k] : 0.;
                            Lab[A0b[(size_t)e * Nx * Nb + i + (size_t)Nx * k] - 1] += v * dt;     // add_∂!{1}
                            Lba[Ab0[(size_t)e * Nb * Nx + k + (size_t)Nb * i] - 1] += v * dt;     // add_∂!{1,:plus,:transpose}
                        }
                    }
                }
            } else {
                // Lagrangian path: revariate{2}((;Λ,X,U),(;Λ=scale.Λ,X=scale.X,U=scale.U)), L = Λ∘₁R
                constexpr int NP = 12;
                using D1 = Dual<NP, double>;
                using TT = Dual<NP, D1>;
                auto var = [](double v, int p, double sc) { TT r; r.x = D1(v); r.x.dx[p] = sc; r.dx[p] = D1(sc); return r; };
                int p = 0;
                TT Lv[12], Xd[3][12], Ud[3];
                int pL = p; for (int k = 0; k < Nx; ++k) Lv[k] = var(lam[k], p++, t.scaleL[k]);
                int pX = p; for (int d = 0; d < ND; ++d) for (int k = 0; k < Nx; ++k) Xd[d][k] = var(x[d][k], p++, t.scaleX[k]);
                int pU = p; for (int d = 0; d <= OU; ++d) for (int k = 0; k < Nu; ++k) Ud[k] = var(u[d][k], p++, t.scaleU[k]);
                if (p > NP) return 5;
                TT R[12];
                direct_residual<TT, ND>(t, e, Xd, Ud, time, R);
                TT L = Lv[0] * R[0];
                for (int i = 1; i < Nx; ++i) L = L + Lv[i] * R[i];
                auto pstart = [&](int c, int d) { return c == 0 ? pL : (c == 1 ? pX + d * Nx : pU + d * Nu); };
                auto nloc = [&](int c) { return c == 2 ? Nu : Nx; };
                auto amap = [&](int c) { return c == 2 ? aU : aL; };
                for (int a = 0; a < 3; ++a)
                    for (int i = 0; i < dnder(m, a); ++i)
                        for (int k = 0; k < nloc(a); ++k)
                            L1[l1_offset(m, a) + (int64_t)i * (a == 2 ? m->nU : m->nX) + amap(a)[k] - 1] +=
                                L.dx[pstart(a, i) + k].x * dt;
                if (!mats) continue;
                for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) {
                    if (!dhas(a, b)) continue;
                    const auto& Aab = m->dpat[a][b].asm2[it];
                    const int na = nloc(a), nb = nloc(b);
                    for (int i = 0; i < dnder(m, a); ++i) for (int j = 0; j < dnder(m, b); ++j) {
                        double* Lab = L2 + l2_offset(m, a, b) + ((int64_t)i * dnder(m, b) + j) * m->dpat[a][b].nnz;
                        for (int k = 0; k < na; ++k) for (int l = 0; l < nb; ++l)
                            Lab[Aab[(size_t)e * na * nb + k + (size_t)na * l] - 1] += L.dx[pstart(a, i) + k].dx[pstart(b, j) + l] * dt;
                    }
                }
            }
        }
    }
    return 0;
}

extern "C" {
int orc_model_set_scales(void* h, int itype, const double* scaleL, const double* scaleU) {
    OModel* m = (OModel*)h; EType& t = m->types[itype];
    if (scaleL) t.scaleL.assign(scaleL, scaleL + t.nX);
    if (scaleU) t.scaleU.assign(scaleU, scaleU + t.nU);
    return 0;
}
// prepare(AssemblyDirect{OX,OU,IA}) (DirectXUA.jl:22-56); nnz of the 9 class pairs (Λ,X,U)².
int orc_model_prepare_direct(void* h, int OX, int OU, int IA, int64_t* nnz9) {
    OModel* m = (OModel*)h;
    if (m->nA) return 1;
    m->dOX = OX; m->dOU = OU; m->dIA = IA;
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) {
        build_dpattern(m, a, b, m->dpat[a][b]);
        if (nnz9) nnz9[3 * a + b] = m->dpat[a][b].nnz;
    }
    return 0;
}
void orc_model_direct_pattern(void* h, int a, int b, int64_t* colptr, int64_t* rowval) {
    OModel* m = (OModel*)h; const DPattern& P = m->dpat[a][b];
    std::copy(P.colptr.begin(), P.colptr.end(), colptr);
    std::copy(P.rowval.begin(), P.rowval.end(), rowval);
}
int64_t orc_model_direct_sizes(void* h, int64_t* nL1) {
    OModel* m = (OModel*)h;
    if (nL1) *nL1 = l1_offset(m, 2) + (int64_t)(m->dOU + 1) * m->nU;
    return l2_offset(m, 3, 0);
}
int orc_model_assemble_direct(void* h, int mission, double dt, const double* Lam, const double* X0, const double* X1,
                              const double* X2, const double* U0, const double* U1, const double* U2, double t,
                              double* L1, double* L2) {
    OModel* m = (OModel*)h;
    const double* Xs[3] = {X0, X1, X2};
    const double* Us[3] = {U0, U1, U2};
    if (m->dOX == 0) return assemble_direct<0>(m, mission, dt, Lam, Xs, Us, t, L1, L2);
    if (m->dOX == 1) return assemble_direct<1>(m, mission, dt, Lam, Xs, Us, t, L1, L2);
    if (m->dOX == 2) return assemble_direct<2>(m, mission, dt, Lam, Xs, Us, t, L1, L2);
    return 1;
}
}  // extern "C"
