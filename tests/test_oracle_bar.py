"""Oracle Bar3D (oracle/bar.hpp) pinned to the reference's bar tests and to mathematics.

Goldens: test/TestBarElement.jl (constructor :17-26, tension :35-38, axial and
transverse stiffness :50-69, spurious stiffness :70-75, consistent mass :78-100,
weight :116-118, added mass :131-149, damping :163-178). `approx` is Julia's `≈`
(rtol √eps on the norm). Central differences pin the tangent at random states,
including velocities and accelerations.
"""
import math

import numpy as np
import pytest

from tests import _oracle as orc

RTOL = math.sqrt(np.finfo(float).eps)
EA, L0, MU = 10., 2., 1.


def approx(a, b, atol=0.0, rtol=RTOL):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) <= max(atol, rtol * max(np.linalg.norm(a), np.linalg.norm(b)))


def mat9(EA, mu, w=0., Cat=0., Clt=0., Cqt=0., Can=0., Cln=0., Cqn=0.):
    return np.array([EA, mu, w, Cat, Clt, Cqt, Can, Cln, Cqn], float)


C1, C2 = [0., 0., 0.], [L0, 0., 0.]


def test_constructor_golden():   # TestBarElement.jl:17-26
    o = orc.bar_construct(C1, C2)
    assert approx(o["cm"], [L0 / 2, 0., 0.]) and approx(o["tgm"], [L0, 0., 0.]) and approx(o["L0"], L0)
    assert approx(o["wgp"], [0.34785484513745385, 0.6521451548625462, 0.6521451548625462, 0.34785484513745385])
    assert approx(o["zgp"], [-0.4305681557970263, -0.16999052179242816, 0.16999052179242816, 0.4305681557970263])
    assert approx(o["znod"], [-0.5, 0.5])
    assert approx(o["psi1"], [0.9305681557970262, 0.6699905217924281, 0.33000947820757187, 0.06943184420297371])
    assert approx(o["psi2"], [0.06943184420297371, 0.33000947820757187, 0.6699905217924281, 0.9305681557970262])
    assert o["Ls"] == (1. - np.finfo(float).eps) * L0


def test_residual_tension():   # TestBarElement.jl:30-38
    dx = .1
    R, _, _ = orc.bar_diffed_residual(C1, C2, mat9(EA, MU), [[0, 0, 0, dx, 0, 0]])
    assert approx(R, [-EA / L0 * dx, 0, 0, EA / L0 * dx, 0, 0])


def test_stiffness_and_mass():   # TestBarElement.jl:40-100
    dx = .1
    X = [np.array([0, 0, 0, dx, 0, 0.]), np.zeros(6), np.zeros(6)]
    R, (K, C, M), _ = orc.bar_diffed_residual(C1, C2, mat9(EA, MU), X)
    for (i, j), v in {(1, 1): EA / L0, (4, 4): EA / L0, (1, 4): -EA / L0, (4, 1): -EA / L0}.items():
        assert approx(K[i - 1, j - 1], v)
    kt = (EA / L0) * dx / (L0 + dx)
    for (i, j), v in {(2, 2): kt, (3, 3): kt, (5, 5): kt, (6, 6): kt, (2, 5): -kt, (5, 2): -kt, (3, 6): -kt,
                      (6, 3): -kt}.items():
        assert approx(K[i - 1, j - 1], v)
    assert np.linalg.norm(K[np.ix_([0, 3], [1, 2, 4, 5])]) <= 1e-12 * EA
    assert np.linalg.norm(K[np.ix_([1, 4], [0, 2, 3, 5])]) <= 1e-12 * EA
    assert np.linalg.norm(K[np.ix_([2, 5], [0, 1, 3, 4])]) <= 1e-12 * EA
    for i in range(6):
        assert approx(M[i, i], MU * L0 / 3)
    for i, j in ((0, 3), (1, 4), (2, 5)):
        assert approx(M[i, j], MU * L0 / 6) and approx(M[j, i], MU * L0 / 6)
    for r, cols in {0: [1, 2, 4, 5], 3: [1, 2, 4, 5], 1: [0, 2, 3, 5], 4: [0, 2, 3, 5], 2: [0, 1, 3, 4], 5: [0, 1, 3, 4]}.items():
        assert np.linalg.norm(M[r, cols]) <= 1e-14
    assert np.abs(C).max() == 0.


def test_weight_ramp():   # TestBarElement.jl:103-118 and the ramp of BarElement.jl:144
    w = 10.
    for t, f in ((0., 1.), (-5., 1.), (-7.5, .5), (-10., 0.), (-20., -2.)):
        R, _, _ = orc.bar_diffed_residual(C1, C2, mat9(EA, MU, w=w), [np.zeros(6)], t=t)
        assert approx(R, [0, 0, f * w * L0 / 2, 0, 0, f * w * L0 / 2], atol=1e-14), t


def test_added_mass():   # TestBarElement.jl:120-149
    Cat, Can = 2., 3.
    m = mat9(EA, MU, Cat=Cat, Can=Can)
    z = np.zeros(6)
    for k, a in ((0, 4.), (1, 3.), (2, 2.)):
        acc = np.zeros(6); acc[k] = acc[k + 3] = a
        R, _, _ = orc.bar_diffed_residual(C1, C2, m, [z, z, acc])
        exp = np.zeros(6); exp[k] = exp[k + 3] = (MU + (Cat if k == 0 else Can)) * a * L0 / 2
        assert approx(R, exp)


def test_damping():   # TestBarElement.jl:151-178
    Clt, Cln, Cqt, Cqn = 1., 2., 3., 4.
    m = mat9(EA, MU, Clt=Clt, Cln=Cln, Cqt=Cqt, Cqn=Cqn)
    z = np.zeros(6)
    for k, v in ((0, 1.0), (1, 1.1), (2, 0.1)):
        vel = np.zeros(6); vel[k] = vel[k + 3] = v
        R, _, _ = orc.bar_diffed_residual(C1, C2, m, [z, vel, z])
        Cl, Cq = (Clt, Cqt) if k == 0 else (Cln, Cqn)
        exp = np.zeros(6); exp[k] = exp[k + 3] = (Cl + Cq * abs(v)) * v * L0 / 2
        assert approx(R, exp)


@pytest.mark.parametrize("seed", range(4))
def test_tangent_matches_central_differences(seed):
    """∂R/∂X₀, ∂R/∂X₁, ∂R/∂X₂ and ∂R/∂U of a tilted bar at a random state, against
    central differences of the residual (h = 1e-6, relative agreement ~1e-7)."""
    rng = np.random.default_rng(seed)
    c1 = rng.normal(size=3)
    c2 = c1 + rng.normal(size=3) * 2
    m = mat9(*rng.uniform(1., 10., 2), w=rng.uniform(-2, 2), Cat=.5, Clt=.7, Cqt=.3, Can=1.5, Cln=.9, Cqn=1.1)
    X = [rng.normal(size=6) * 0.2, rng.normal(size=6), rng.normal(size=6)]
    U = rng.normal(size=3)
    R, Ks, dU = orc.bar_diffed_residual(c1, c2, m, X, U=U, t=-3.)
    h = 1e-6
    for d in range(3):
        for j in range(6):
            Xp = [x.copy() for x in X]; Xm = [x.copy() for x in X]
            Xp[d][j] += h; Xm[d][j] -= h
            Rp, _, _ = orc.bar_diffed_residual(c1, c2, m, Xp, U=U, t=-3.)
            Rm, _, _ = orc.bar_diffed_residual(c1, c2, m, Xm, U=U, t=-3.)
            fd = (Rp - Rm) / (2 * h)
            assert np.abs(fd - Ks[d][:, j]).max() <= 1e-6 * max(1., np.abs(Ks[d]).max()), (d, j)
    for c in range(3):
        Up, Um = U.copy(), U.copy(); Up[c] += h; Um[c] -= h
        fd = (orc.bar_diffed_residual(c1, c2, m, X, U=Up, t=-3.)[0] - orc.bar_diffed_residual(c1, c2, m, X, U=Um, t=-3.)[0]) / (2 * h)
        assert np.abs(fd - dU[:, c]).max() <= 1e-8
