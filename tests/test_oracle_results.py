"""Oracle requestables of EulerBeam3D (orc_beam_results: ☼ε, ♢κ, and per Gauss point ☼x,
☼κgp, ☼fᵢ, ☼mᵢ, ☼fₑ, ☼mₑ of toolbox/BeamElement.jl:28-64,151-174), pinned by:
- ε and κ: the same values as the residual's pinned requestables (orc_beam_residual_plain),
  and the strain-gauge goldens of TestBeamElementStrainGauge.jl:35-56 through E, K1, K2, K3;
- closed forms: the element at rest, a rigid translation, a uniform stretch, a rigid
  translational motion (inertia, added mass, linear and quadratic drag in the element
  frame, weight) and a roll acceleration about the element axis (roll-inertia couple).
"""
import math

import numpy as np
import pytest

from tests import _oracle as orc
from tests.test_oracle_beam import mat16


def split(out, ND):
    """The layout of mh_beam_results as named arrays."""
    q = 0
    r = {"ε": out[q:q + ND]}; q += ND
    r["κ"] = out[q:q + 3 * ND].reshape(ND, 3); q += 3 * ND
    gp = []
    for g in range(4):
        d = {"x": out[q:q + 3 * ND].reshape(ND, 3)}; q += 3 * ND
        d["κgp"] = out[q:q + 3 * ND].reshape(ND, 3); q += 3 * ND
        d["fᵢ"] = out[q]; q += 1
        d["mᵢ"] = out[q:q + 3]; q += 3
        d["fₑ"] = out[q:q + 3]; q += 3
        d["mₑ"] = out[q:q + 3]; q += 3
        gp.append(d)
    assert q == out.size
    r["gp"] = gp
    return r


C1, C2, O2 = np.array([1., 2., 0.5]), np.array([4., 2.5, 1.3]), np.array([0., 0., 1.])
M = mat16(7., 3., 2., 4., 1.5, 0.7, w=2.5, Ca=(0.3, 1.1, 1.3), Cl=(0.2, 0.9, 0.8), Cq=(0.5, 1.7, 2.1))
ZGP = np.array([-0.5 * math.sqrt(3 / 7 + 2 / 7 * math.sqrt(6 / 5)), -0.5 * math.sqrt(3 / 7 - 2 / 7 * math.sqrt(6 / 5)),
                0.5 * math.sqrt(3 / 7 - 2 / 7 * math.sqrt(6 / 5)), 0.5 * math.sqrt(3 / 7 + 2 / 7 * math.sqrt(6 / 5))])


def frame():
    t = (C2 - C1) / np.linalg.norm(C2 - C1)
    n = O2 - t * (O2 @ t)
    n /= np.linalg.norm(n)
    return np.stack([t, n, np.cross(t, n)], axis=1), np.linalg.norm(C2 - C1)


@pytest.mark.parametrize("ND", [1, 2, 3])
def test_eps_kappa_equal_pinned_requestables(ND):
    rng = np.random.default_rng(ND)
    X = [rng.normal(size=12) * 0.2 for _ in range(ND)]
    r = split(orc.beam_results(C1, C2, O2, M, X), ND)
    _, eps, kap = orc.beam_residual_plain(C1, C2, O2, M, X)
    assert r["ε"][0] == eps and np.array_equal(r["κ"][0], kap)


def test_strain_gauge_goldens():   # TestBeamElementStrainGauge.jl:35-56
    mat = mat16(10., 3., 3., 4., 1., 1.0)
    c1, c2, o2 = [0, 0, 0], [4, 0, 0], [0., 1., 0.]
    E = np.array([1., 1., 1., 1., .5]); K1 = np.array([0, 0, 0, 0, .25])
    K2 = np.array([-0.5, 0, 0.5, 0, -.25]); K3 = np.array([0, -0.5, 0, 0.5, 0])
    cases = [([0, 0, 0, 0, .1, 0, 0, 0, 0, 0, -.1, 0], [0, 0, 1], 20, [0, -1, 0, 1, 0], 40),
             ([0, 0, 0, 0, 0, .1, 0, 0, 0, 0, 0, -.1], [0, -1, 0], 20, [1, 0, -1, 0, .5], 40),
             ([0, 0, 0, 0, 0, 0, 0, 0, 0, 1., 0, 0], [.25, 0, 0], 1, [0, 0, 0, 0, 0.0625], 1)]
    for X, kgold, kf, egold, ef in cases:
        r = split(orc.beam_results(c1, c2, o2, mat, [np.array(X, float)]), 1)
        eps, kap = r["ε"][0], r["κ"][0]
        e = E * eps + K1 * kap[0] + K2 * kap[1] + K3 * kap[2]
        assert abs(eps) <= 1e-12
        assert np.allclose(kap * kf, kgold, atol=1e-12) and np.allclose(e * ef, egold, atol=1e-12)


def test_at_rest():
    r = split(orc.beam_results(C1, C2, O2, M, [np.zeros(12)] * 3), 3)
    assert np.abs(r["ε"]).max() <= 1e-15 and np.abs(r["κ"]).max() <= 1e-15   # L from √ of the chord: round-off
    for g, d in enumerate(r["gp"]):
        assert np.allclose(d["x"][0], C1 + (ZGP[g] + 0.5) * (C2 - C1), rtol=0, atol=1e-14)
        assert np.all(d["x"][1:] == 0) and np.abs(d["κgp"]).max() <= 1e-15
        assert abs(d["fᵢ"]) <= 1e-14 and np.abs(d["mᵢ"]).max() <= 1e-14 and np.all(d["mₑ"] == 0)
        assert np.array_equal(d["fₑ"], [0., 0., M[6]])


def test_rigid_translation_and_stretch():
    u = np.array([0.3, -0.2, 0.7])
    X = np.concatenate([u, np.zeros(3), u, np.zeros(3)])
    r = split(orc.beam_results(C1, C2, O2, M, [X]), 1)
    assert abs(r["ε"][0]) <= 1e-15 and np.abs(r["κ"]).max() <= 1e-15
    for g, d in enumerate(r["gp"]):
        assert np.allclose(d["x"][0], C1 + (ZGP[g] + 0.5) * (C2 - C1) + u, rtol=0, atol=1e-14)
        assert np.abs(d["κgp"]).max() <= 1e-15
    rm, L = frame()
    δ = 0.05
    t = rm[:, 0]
    X = np.concatenate([-δ / 2 * t, np.zeros(3), δ / 2 * t, np.zeros(3)])
    r = split(orc.beam_results(C1, C2, O2, M, [X]), 1)
    assert abs(r["ε"][0] - δ / L) <= 1e-15
    for g, d in enumerate(r["gp"]):
        assert abs(d["fᵢ"] - M[0] * δ / L) <= 1e-14
        assert np.allclose(d["x"][0], (C1 + C2) / 2 + ZGP[g] * (L + δ) * t, rtol=0, atol=1e-14)
        assert np.abs(d["κgp"]).max() <= 1e-15


def test_rigid_motion_loads():
    """Translational velocity v and acceleration a of the whole element: ẋ = v, ẍ = a at every
    Gauss point and fₑ = μa + r(Ca∘rᵀa) + r(Cl∘rᵀv) + r(sign·Cq∘(rᵀv)²) + (0,0,w)."""
    rm, L = frame()
    v = np.array([0.4, -1.3, 0.6]); a = np.array([-0.7, 0.2, 1.1])
    X = [np.zeros(12), np.concatenate([v, np.zeros(3), v, np.zeros(3)]),
         np.concatenate([a, np.zeros(3), a, np.zeros(3)])]
    r = split(orc.beam_results(C1, C2, O2, M, X), 3)
    mu, w = M[4], M[6]
    Ca, Cl, Cq = M[[7, 10, 13]], M[[8, 11, 14]], M[[9, 12, 15]]
    vl, al = rm.T @ v, rm.T @ a
    fe = mu * a + rm @ (Ca * al) + rm @ (Cl * vl) + rm @ (np.sign(vl) * Cq * vl * vl) + np.array([0, 0, w])
    assert np.abs(r["ε"]).max() <= 1e-15
    for d in r["gp"]:
        assert np.allclose(d["x"][1], v, rtol=0, atol=1e-14) and np.allclose(d["x"][2], a, rtol=0, atol=1e-14)
        assert np.allclose(d["fₑ"], fe, rtol=1e-14, atol=1e-14)
        assert np.abs(d["mₑ"]).max() <= 1e-14


def test_roll_acceleration_couple():
    """Both nodes turned by 0.3 rad about the element axis t, with roll acceleration α about t:
    vᵢ₂ = (α, 0, 0) and mₑ = t·ι₁α at every Gauss point (BeamElement.jl:54-56)."""
    rm, L = frame()
    t = rm[:, 0]
    α = 0.9
    rot = np.concatenate([np.zeros(3), 0.3 * t, np.zeros(3), 0.3 * t])
    acc = np.concatenate([np.zeros(3), α * t, np.zeros(3), α * t])
    r = split(orc.beam_results(C1, C2, O2, M, [rot, np.zeros(12), acc]), 3)
    for d in r["gp"]:
        assert np.allclose(d["mₑ"], M[5] * α * t, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("seed", range(4))
def test_gauge_second_derivatives(seed):
    """orc_beam_gauge2: values and gradients equal the pinned first-order requestables
    (orc_beam_gauge, checked against the costedbeam ∇L golden), Hessians equal central
    differences of those gradients and are symmetric."""
    rng = np.random.default_rng(seed)
    X = rng.normal(size=12) * 0.2
    if seed == 0:
        X[[3, 4, 5, 9, 10, 11]] = 0.   # zero rotations (the series branches)
    v, g, H = orc.beam_gauge2(C1, C2, O2, M, X)
    _, _, eps, deps, kap, dkap = orc.beam_gauge(C1, C2, O2, M, X)
    assert np.allclose(v, [eps, *kap], rtol=1e-14, atol=1e-15)
    assert np.allclose(g, np.vstack([deps, dkap]), rtol=1e-13, atol=1e-14)
    assert np.abs(H - np.swapaxes(H, 1, 2)).max() == 0.
    h = 1e-6
    for i in range(12):
        e = np.zeros(12); e[i] = h
        _, _, _, dp, _, kp = orc.beam_gauge(C1, C2, O2, M, X + e)
        _, _, _, dm, _, km = orc.beam_gauge(C1, C2, O2, M, X - e)
        fd = (np.vstack([dp, kp]) - np.vstack([dm, km])) / (2 * h)
        assert np.allclose(H[:, i, :], fd, rtol=1e-6, atol=1e-8), i
