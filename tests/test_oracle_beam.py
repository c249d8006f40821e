"""Oracle EulerBeam3D (oracle/beam.hpp) pinned to the reference's beam tests and to mathematics.

`approx` is Julia's `≈` on arrays: norm(a-b) <= rtol*max(norm(a),norm(b)), rtol=√eps
(optionally with atol), which is how the reference states these goldens.
"""
import math

import numpy as np
import pytest

from tests import _oracle as orc

RTOL = math.sqrt(np.finfo(float).eps)


def approx(a, b, atol=0.0, rtol=RTOL):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) <= max(atol, rtol * max(np.linalg.norm(a), np.linalg.norm(b)))


def mat16(EA, EI2, EI3, GJ, mu, iota1, w=0., Ca=(0, 0, 0), Cl=(0, 0, 0), Cq=(0, 0, 0)):
    return np.array([EA, EI2, EI3, GJ, mu, iota1, w, Ca[0], Cl[0], Cq[0], Ca[1], Cl[1], Cq[1], Ca[2], Cl[2], Cq[2]],
                    float)


def test_constructor_golden():   # TestBeamElement.jl:6-31
    L = 5.
    o = orc.beam_construct([0, 0, 0], [4, 3, 0], [0., 1., 0.])
    assert approx(o["cm"], [2.0, 1.5, 0.0])
    assert approx(o["rm"], [[0.8, -0.6, 0.0], [0.6, 0.8, -0.0], [0.0, 0.0, 1.0]])
    assert approx(o["zgp"], [-0.4305681557970263, -0.16999052179242816, 0.16999052179242816, 0.4305681557970263])
    assert approx(o["znod"], [-0.5, 0.5])
    assert approx(o["tgm"], [4.0, 3.0, 0.0])
    assert approx(o["tge"], [5.0, 0.0, 0.0])
    assert approx(o["ya"], [-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526])
    assert approx(o["yu"], [-0.972414176921822, -0.49032285223640754, 0.49032285223640754, 0.972414176921822])
    assert approx(o["yv"], np.array([-0.0646110632135477, -0.221103222500738, -0.221103222500738, -0.0646110632135477]) * L)
    assert approx(o["ka"], np.array([2.0, 2.0, 2.0, 2.0]) / L)
    assert approx(o["ku"], np.array([10.333635739128631, 4.079772523018276, -4.079772523018276, -10.333635739128631]) / L ** 2)
    assert approx(o["kv"], np.array([2.0, 2.0, 2.0, 2.0]) / L)
    assert approx(o["dL"], np.array([0.17392742256872692, 0.3260725774312731, 0.3260725774312731, 0.17392742256872692]) * L)


def test_constructor_rejects_parallel_orient2():   # BeamElement.jl:135
    with pytest.raises(ValueError):
        orc.beam_construct([0, 0, 0], [1, 0, 0], [1., 1e-5, 0.])


EA, EI2, EI3, GJ, L, MU, IOTA = 10., 3., 2., 4., 2., 1., 5.
C1, C2, O2 = [0, 0, 0], [L, 0, 0], [0., 1., 0.]


def resid(x, mat=None, X=None):
    m = mat16(EA, EI2, EI3, GJ, MU, IOTA) if mat is None else mat
    R, _, _ = orc.beam_residual_plain(C1, C2, O2, m, X if X is not None else [x])
    return R


def test_residual_closed_forms():   # TestBeamElement.jl:52-82
    t1n2 = 0.1
    x = np.zeros(12); x[6] = t1n2
    assert approx(resid(x), [-EA / L * t1n2, 0, 0, 0, 0, 0, EA / L * t1n2, 0, 0, 0, 0, 0])
    t2n2 = 0.01
    x = np.zeros(12); x[7] = t2n2
    ref = [0, -12 * EI3 / L ** 3 * t2n2, 0, 0, 0, -6 * EI3 / L ** 2 * t2n2, 0, 12 * EI3 / L ** 3 * t2n2, 0, 0, 0,
           -6 * EI3 / L ** 2 * t2n2]
    assert approx(resid(x), ref, atol=1e-2)
    r3n2 = 0.01
    x = np.zeros(12); x[11] = r3n2
    ref = [0, 6 * EI3 / L ** 2 * r3n2, 0, 0, 0, 2 * EI3 / L * r3n2, 0, -6 * EI3 / L ** 2 * r3n2, 0, 0, 0, 4 * EI3 / L * r3n2]
    assert approx(resid(x), ref, atol=1e-2)
    r1n2 = 0.1
    x = np.zeros(12); x[9] = r1n2
    assert approx(resid(x), [0, 0, 0, -GJ / L * r1n2, 0, 0, 0, 0, 0, GJ / L * r1n2, 0, 0])


@pytest.fixture(scope="module")
def KCM():
    R, Ks, _ = orc.beam_diffed_residual(C1, C2, O2, mat16(EA, EI2, EI3, GJ, MU, IOTA), [np.zeros(12)] * 3)
    return Ks


def test_stiffness_closed_forms(KCM):   # TestBeamElement.jl:107-172
    K = KCM[0]
    a = lambda v, r: approx(v, r)   # noqa: E731
    for (i, j), v in {(1, 1): EA / L, (7, 7): EA / L, (1, 7): -EA / L, (7, 1): -EA / L,
                      (2, 2): 12 * EI3 / L ** 3, (8, 8): 12 * EI3 / L ** 3, (3, 3): 12 * EI2 / L ** 3, (9, 9): 12 * EI2 / L ** 3,
                      (3, 9): -12 * EI2 / L ** 3, (9, 3): -12 * EI2 / L ** 3, (2, 8): -12 * EI3 / L ** 3, (8, 2): -12 * EI3 / L ** 3,
                      (8, 6): -6 * EI3 / L ** 2, (2, 12): 6 * EI3 / L ** 2, (9, 5): 6 * EI2 / L ** 2, (3, 11): -6 * EI2 / L ** 2,
                      (5, 5): 4 * EI2 / L, (11, 11): 4 * EI2 / L, (6, 6): 4 * EI3 / L, (12, 12): 4 * EI3 / L,
                      (5, 11): 2 * EI2 / L, (11, 5): 2 * EI2 / L, (6, 12): 2 * EI3 / L, (12, 6): 2 * EI3 / L,
                      (5, 9): 6 * EI2 / L ** 2, (11, 3): -6 * EI2 / L ** 2, (6, 8): -6 * EI3 / L ** 2, (12, 2): 6 * EI3 / L ** 2,
                      (4, 4): GJ / L, (10, 10): GJ / L, (4, 10): -GJ / L, (10, 4): -GJ / L}.items():
        assert a(K[i - 1, j - 1], v), (i, j, K[i - 1, j - 1], v)
    zero_rows = {1: [2, 3, 4, 5, 6, 8, 9, 10, 11, 12], 7: [2, 3, 4, 5, 6, 8, 9, 10, 11, 12],
                 4: [1, 2, 3, 5, 6, 7, 8, 9, 11, 12], 10: [1, 2, 3, 5, 6, 7, 8, 9, 11, 12],
                 2: [1, 3, 4, 5, 7, 9, 10, 11], 3: [1, 2, 4, 6, 7, 8, 10, 12], 8: [1, 3, 4, 5, 7, 9, 10, 11],
                 9: [1, 2, 4, 6, 7, 8, 10, 12], 5: [1, 2, 4, 6, 7, 8, 10, 12], 6: [1, 3, 4, 5, 7, 9, 10, 11],
                 11: [1, 2, 4, 6, 7, 8, 10, 12], 12: [1, 3, 4, 5, 7, 9, 10, 11]}
    for r, cols in zero_rows.items():
        assert np.linalg.norm(K[r - 1, np.array(cols) - 1]) <= 1e-12 * np.abs(K).max()


def test_mass_closed_forms(KCM):   # TestBeamElement.jl:175-239
    M = KCM[2]
    mu, i1 = MU, IOTA
    for (i, j), v in {(1, 1): mu * L / 3, (7, 7): mu * L / 3, (1, 7): mu * L / 6, (7, 1): mu * L / 6,
                      (8, 8): 156 * mu * L / 420, (2, 2): 156 * mu * L / 420, (3, 3): 156 * mu * L / 420, (9, 9): 156 * mu * L / 420,
                      (3, 9): 54 * mu * L / 420, (9, 3): 54 * mu * L / 420, (2, 8): 54 * mu * L / 420, (8, 2): 54 * mu * L / 420,
                      (8, 6): 13 * mu * L ** 2 / 420, (2, 12): -13 * mu * L ** 2 / 420, (9, 5): -13 * mu * L ** 2 / 420,
                      (3, 11): 13 * mu * L ** 2 / 420, (5, 5): 4 * mu * L ** 3 / 420, (11, 11): 4 * mu * L ** 3 / 420,
                      (6, 6): 4 * mu * L ** 3 / 420, (12, 12): 4 * mu * L ** 3 / 420, (5, 11): -3 * mu * L ** 3 / 420,
                      (11, 5): -3 * mu * L ** 3 / 420, (6, 12): -3 * mu * L ** 3 / 420, (12, 6): -3 * mu * L ** 3 / 420,
                      (5, 9): -13 * mu * L ** 2 / 420, (11, 3): 13 * mu * L ** 2 / 420, (6, 8): 13 * mu * L ** 2 / 420,
                      (12, 2): -13 * mu * L ** 2 / 420, (4, 4): i1 * L / 4, (10, 10): i1 * L / 4, (4, 10): i1 * L / 4,
                      (10, 4): i1 * L / 4}.items():
        assert approx(M[i - 1, j - 1], v), (i, j, M[i - 1, j - 1], v)


def test_weight_added_mass_damping():   # TestBeamElement.jl:242-326
    w = 10.
    m = mat16(EA, EI2, EI3, GJ, MU, IOTA, w=w)
    x = np.zeros(12); x[4] = w * L ** 3 / (24 * EI2); x[10] = -w * L ** 3 / (24 * EI2)
    assert approx(resid(x, m), [0, 0, w * L / 2, 0, 0, 0, 0, 0, w * L / 2, 0, 0, 0])
    Ca = (1., 2., 3.)
    m = mat16(EA, EI2, EI3, GJ, 1., IOTA, Ca=Ca)
    a1, a2, a3 = 4.0, 3.0, 2.0
    v0 = np.zeros(12)
    acc = np.zeros(12); acc[0] = acc[6] = a1
    assert approx(resid(None, m, [np.zeros(12), v0, acc]), [(1 + Ca[0]) * a1 * L / 2, 0, 0, 0, 0, 0, (1 + Ca[0]) * a1 * L / 2, 0, 0, 0, 0, 0])
    x = np.zeros(12); x[5] = -(1 + Ca[1]) * a2 * L ** 3 / (24 * EI3); x[11] = -x[5]
    acc = np.zeros(12); acc[1] = acc[7] = a2
    assert approx(resid(None, m, [x, v0, acc]), [0, (1 + Ca[1]) * a2 * L / 2, 0, 0, 0, 0, 0, (1 + Ca[1]) * a2 * L / 2, 0, 0, 0, 0])
    x = np.zeros(12); x[4] = (1 + Ca[2]) * a3 * L ** 3 / (24 * EI2); x[10] = -x[4]
    acc = np.zeros(12); acc[2] = acc[8] = a3
    assert approx(resid(None, m, [x, v0, acc]), [0, 0, (1 + Ca[2]) * a3 * L / 2, 0, 0, 0, 0, 0, (1 + Ca[2]) * a3 * L / 2, 0, 0, 0])
    Cl, Cq = (1., 2., 3.), (.1, .2, .3)
    m = mat16(EA, EI2, EI3, GJ, 1., IOTA, Cl=Cl, Cq=Cq)
    v1, v2, v3 = 1.0, 1.1, 0.1
    vel = np.zeros(12); vel[0] = vel[6] = v1
    f = (Cl[0] + Cq[0] * abs(v1)) * v1 * L / 2
    assert approx(resid(None, m, [np.zeros(12), vel, np.zeros(12)]), [f, 0, 0, 0, 0, 0, f, 0, 0, 0, 0, 0])
    f = (Cl[1] + Cq[1] * abs(v2)) * v2
    x = np.zeros(12); x[5] = -f * L ** 3 / (24 * EI3); x[11] = -x[5]
    vel = np.zeros(12); vel[1] = vel[7] = v2
    assert approx(resid(None, m, [x, vel, np.zeros(12)]), [0, f * L / 2, 0, 0, 0, 0, 0, f * L / 2, 0, 0, 0, 0])
    f = (Cl[2] + Cq[2] * abs(v3)) * v3
    x = np.zeros(12); x[4] = f * L ** 3 / (24 * EI2); x[10] = -x[4]
    vel = np.zeros(12); vel[2] = vel[8] = v3
    assert approx(resid(None, m, [x, vel, np.zeros(12)]), [0, 0, f * L / 2, 0, 0, 0, 0, 0, f * L / 2, 0, 0, 0])


def test_strain_gauge_golden():   # TestBeamElementStrainGauge.jl:5-56,90-111
    mat = mat16(10., 3., 3., 4., 1., 1.0)
    c1, c2, o2 = [0, 0, 0], [4, 0, 0], [0., 1., 0.]
    P = np.array([[0., .5, 0.], [0., 0, .5], [0., -.5, 0.], [0., 0, -.5], [0., .5, 0.]]).T
    D = np.array([[1., 0., 0.], [1., 0., 0.], [1., 0., 0.], [1., 0., 0.], [1 / math.sqrt(2), 0, 1 / math.sqrt(2)]]).T
    E = D[0] ** 2
    K1 = D[0] * (D[2] * P[1] - D[1] * P[2])
    K2 = -D[0] ** 2 * P[1]
    K3 = -D[0] ** 2 * P[2]
    assert approx(E, [1., 1., 1., 1., .5]) and approx(K1, [0, 0, 0, 0, .25])
    assert approx(K2, [-0.5, 0, 0.5, 0, -.25]) and approx(K3, [0, -0.5, 0, 0.5, 0])

    def gauges(X):
        R, K, eps, deps, kap, dkap = orc.beam_gauge(c1, c2, o2, mat, X)
        e = E * eps + K1 * kap[0] + K2 * kap[1] + K3 * kap[2]
        de = np.outer(E, deps) + np.outer(K1, dkap[0]) + np.outer(K2, dkap[1]) + np.outer(K3, dkap[2])
        return R, K, eps, kap, e, de
    _, _, eps, kap, e, _ = gauges(np.array([0, 0, 0, 0, .1, 0, 0, 0, 0, 0, -.1, 0.]))
    assert abs(eps) <= 1e-12 and approx(kap * 20, [0, 0, 1], atol=1e-12) and approx(e * 40, [0, -1, 0, 1, 0], atol=1e-12)
    _, _, eps, kap, e, _ = gauges(np.array([0, 0, 0, 0, 0, .1, 0, 0, 0, 0, 0, -.1]))
    assert abs(eps) <= 1e-12 and approx(kap * 20, [0, -1, 0], atol=1e-12) and approx(e * 40, [1, 0, -1, 0, .5], atol=1e-12)
    _, _, eps, kap, e, _ = gauges(np.array([0, 0, 0, 0, 0, 0, 0, 0, 0, 1., 0, 0]))
    assert abs(eps) <= 1e-12 and approx(kap, [.25, 0, 0], atol=1e-12) and approx(e, [0, 0, 0, 0, 0.0625], atol=1e-12)
    # ElementCost(StrainGauge) Lagrangian gradient: L = Λ·R + Δε·Δε/(2σ²), εₘ = (cos t,0,-cos t,0,cos t/2)·1e-3, t=0
    Lam = np.array([0, 0, 0, 0, .1, 0, 0, 0, 0, 0, -.1, 0])
    X = np.array([0, 0, 0, 0, .1, 0, 0, 0, 0, 0, -.1, 0.])
    R, K, eps, kap, e, de = gauges(X)
    sigma = 15e-6
    em = np.array([1., 0., -1., 0., .5]) * 0.001
    gradL = K.T @ Lam + de.T @ (e - em) / sigma ** 2
    gold = [277777.7777783019, 0.0, 1.1102230246251565e-16, 201441.02435832855, 2.7777777927777793e7,
            -1.2430497627256343e6, -277777.7777783019, 0.0, -1.1102230246251565e-16, -76336.75341947998,
            -2.7777777927777793e7, 1.2569502372743965e6]
    assert approx(gradL, gold)


@pytest.mark.parametrize("seed", range(6))
def test_tangent_matches_central_differences(seed):
    """K from the nested duals equals central differences of the plain residual (independent of the AD)."""
    rng = np.random.default_rng(seed)
    c1 = rng.normal(size=3)
    c2 = c1 + rng.normal(size=3) * 2
    o2 = rng.normal(size=3)
    m = mat16(*rng.uniform(1, 10, 4), 1., 1., w=rng.uniform(-1, 1))
    X = rng.normal(size=12) * 0.3
    R, Ks, _ = orc.beam_diffed_residual(c1, c2, o2, m, [X])
    h = 1e-6
    Kfd = np.zeros((12, 12))
    for j in range(12):
        dx = np.zeros(12); dx[j] = h
        Rp, _, _ = orc.beam_residual_plain(c1, c2, o2, m, [X + dx])
        Rm, _, _ = orc.beam_residual_plain(c1, c2, o2, m, [X - dx])
        Kfd[:, j] = (Rp - Rm) / (2 * h)
    assert np.abs(Ks[0] - Kfd).max() <= 1e-6 * np.abs(Ks[0]).max()
    # static K is the Hessian of a potential: symmetric
    assert np.abs(Ks[0] - Ks[0].T).max() <= 1e-13 * np.abs(Ks[0]).max()


def test_rigid_body_motion_is_stress_free():
    """A rigid rotation + translation of the element gives R = 0 (frame invariance)."""
    rng = np.random.default_rng(5)
    c1, c2 = np.array([0.3, -0.2, 0.1]), np.array([1.7, 0.4, -0.5])
    m = mat16(10., 3., 2., 4., 1., 1.)
    v = rng.normal(size=3) * 0.8
    M = np.zeros(9)
    orc.lib.orc_rodrigues(orc.dp(orc.f64(v)), orc.dp(M))
    Q = M.reshape(3, 3, order="F")
    u0 = rng.normal(size=3)
    cm = (c1 + c2) / 2
    u1 = u0 + Q @ (c1 - cm) - (c1 - cm)
    u2 = u0 + Q @ (c2 - cm) - (c2 - cm)
    X = np.concatenate([u1, v, u2, v])
    R, _, _ = orc.beam_residual_plain(c1, c2, [0., 0., 1.], m, [X])
    assert np.abs(R).max() <= 1e-13


# ---------------------------------------------------------------- the power rule of the time jets
def test_ipow_three_level_square_hand_derived():
    """Adiff.jl:230, a^b = b==0 ? zero(a) : ∂ℝ(a.x^b, a.dx·b·a.x^(b−1)), worked by hand on
    x = W + W′ε₃, W = V + V′ε₂, W′ = Q + Q′ε₂, V = v + v′ε₁, V′ = p + p′ε₁, Q = q + q′ε₁, Q′ = r + r′ε₁:
    x² = (W², W′·2·W¹) with W² = (V², V′·2·V¹), V¹ = (v, v′) (v is a Float64: v⁰ = 1), but
    W¹ = (V¹, V′·1·V⁰) = (V, 0) because V is a dual and V⁰ = zero(V). Hence
    x² = [v², 2vv′, 2pv, 2(pv′ + p′v), 2qv, 2(qv′ + q′v), 2rv, 2(rv′ + r′v)]: against the exact
    square, the ε₃ε₂ and ε₃ε₂ε₁ parts lack 2qp and 2(qp′ + q′p). Small integers keep every
    product exact."""
    v, v1, p, p1, q, q1, r, r1 = 3., -2., 5., 7., -4., 6., 2., -3.
    got = orc.ipow_probe([v, v1, p, p1, q, q1, r, r1], 2)
    muscade = [v * v, 2 * v * v1, 2 * p * v, 2 * (p * v1 + p1 * v), 2 * q * v, 2 * (q * v1 + q1 * v), 2 * r * v,
               2 * (r * v1 + r1 * v)]
    exact_last = [2 * (r * v + q * p), 2 * (r * v1 + r1 * v + q * p1 + q1 * p)]
    assert list(got) == muscade
    assert list(got[6:]) != exact_last
    # x¹ = (W¹, W′·1·W⁰) = ((v, v′), 0): the rule drops every outer part of a first power
    assert list(orc.ipow_probe([v, v1, p, p1, q, q1, r, r1], 1)) == [v, v1, 0., 0., 0., 0., 0., 0.]
    # x³ = (W³, W′·3·W²) with W² exact: the cube equals the exact product x·x·x
    x = np.array([v, v1, p, p1, q, q1, r, r1])

    def mul(a, b):   # exact product of two such three-level duals (multilinear in ε₁, ε₂, ε₃)
        A = {(i, j, k): a[4 * k + 2 * j + i] for i in (0, 1) for j in (0, 1) for k in (0, 1)}
        B = {(i, j, k): b[4 * k + 2 * j + i] for i in (0, 1) for j in (0, 1) for k in (0, 1)}
        c = np.zeros(8)
        for (i, j, k), av in A.items():
            for (i2, j2, k2), bv in B.items():
                if i + i2 <= 1 and j + j2 <= 1 and k + k2 <= 1:
                    c[4 * (k + k2) + 2 * (j + j2) + (i + i2)] += av * bv
        return c
    assert np.array_equal(orc.ipow_probe(x, 3), mul(mul(x, x), x))
    assert np.array_equal(orc.ipow_probe(x, 2)[:6], mul(x, x)[:6])


def _moving_state(rng, udof):
    c1 = rng.normal(size=3)
    c2 = c1 + np.array([2.0, 0.3, -0.2]) + 0.2 * rng.normal(size=3)
    o2 = np.array([0.1, 1.0, 0.2])
    mat = np.zeros(16)
    mat[:4] = rng.uniform(1., 10., size=4)
    mat[4] = rng.uniform(0.5, 2.)          # μ
    mat[5] = rng.uniform(0.1, 1.)          # ι₁
    mat[6] = rng.uniform(-2., 2.)          # w
    mat[7:16] = rng.uniform(0., 1., size=9)
    X = [rng.normal(size=12) * np.r_[[1e-2] * 3, [0.2] * 3, [1e-2] * 3, [0.2] * 3],
         rng.normal(size=12) * 0.5, rng.normal(size=12) * 0.5]
    U = rng.normal(size=3) if udof else None
    return c1, c2, o2, mat, X, U


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_dynamic_tangent_matches_central_differences(seed):
    """∂R/∂X, ∂R/∂X′, ∂R/∂X″ and ∂R/∂U of the dynamic beam (ND = 3: inertia, added mass, linear
    and quadratic drag, roll inertia, weight, a distributed U-load) at random moving states,
    against central differences of R. R is taken as the oracle evaluates it on the solver's duals,
    power rule included: the rule drops the same terms from the value and from its partials
    (zero(a) zeroes every level below), so the tangent is the exact derivative of that R and
    central differences pin it; the rule itself is pinned by the hand-derived test above."""
    rng = np.random.default_rng(100 + seed)
    c1, c2, o2, mat, X, U = _moving_state(rng, udof=True)
    R0, Ks, KU = orc.beam_diffed_residual(c1, c2, o2, mat, X, U)
    for d in range(3):
        h = 1e-6
        FD = np.zeros((12, 12))
        for k in range(12):
            Xp = [x.copy() for x in X]
            Xm = [x.copy() for x in X]
            Xp[d][k] += h
            Xm[d][k] -= h
            FD[:, k] = (orc.beam_diffed_residual(c1, c2, o2, mat, Xp, U)[0]
                        - orc.beam_diffed_residual(c1, c2, o2, mat, Xm, U)[0]) / (2 * h)
        scale = np.abs(Ks[d]).max()
        assert np.abs(FD - Ks[d]).max() <= 1e-6 * scale, (d, np.abs(FD - Ks[d]).max() / scale)
    FDU = np.zeros((12, 3))
    for c in range(3):
        Up, Um = U.copy(), U.copy()
        Up[c] += 1e-6
        Um[c] -= 1e-6
        FDU[:, c] = (orc.beam_diffed_residual(c1, c2, o2, mat, X, Up)[0]
                     - orc.beam_diffed_residual(c1, c2, o2, mat, X, Um)[0]) / 2e-6
    assert np.abs(FDU - KU).max() <= 1e-6 * np.abs(KU).max()
    assert np.abs(Ks[1]).max() > 0. and np.abs(Ks[2]).max() > 0.


def test_power_rule_changes_the_dynamic_residual_only():
    """Where the rule acts on the beam: the residual on the solver's duals (three levels) and the
    residual on plain X (two levels, where the dropped V⁰ is a Float64 power and equals 1) agree
    at rest in time (X′ = X″ = 0) and at constant rotation rate zero; with moving rotations they
    differ, the difference being the 2ṡ² term of sinc1(θ/2)² that the rule removes."""
    rng = np.random.default_rng(7)
    c1, c2, o2, mat, X, _ = _moving_state(rng, udof=False)
    still = [X[0], np.zeros(12), np.zeros(12)]
    Rd = orc.beam_diffed_residual(c1, c2, o2, mat, still)[0]
    Rp = orc.beam_residual_plain(c1, c2, o2, mat, still)[0]
    assert np.abs(Rd - Rp).max() <= 1e-12 * np.abs(Rp).max()
    Rd = orc.beam_diffed_residual(c1, c2, o2, mat, X)[0]
    Rp = orc.beam_residual_plain(c1, c2, o2, mat, X)[0]
    assert np.abs(Rd - Rp).max() > 1e-9 * np.abs(Rp).max()
