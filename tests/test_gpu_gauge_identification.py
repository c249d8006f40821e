"""The all-steps adjoint gradient of the gauge identification (AllStepsGradient: the steps'
beams as copies in one SweepX context, one assembly and one mh_vjp_lagrangian for all steps,
host Lagrangians for sensors, priors and Holds) on the device:

- for OX = 2, 1 and 0: L and ∂L/∂Λ, ∂L/∂X, ∂L/∂U, ∂L/∂A against the step-block route with the oracle's Jacobians
  (the oracle restates Muscade's forward-mode duals), 1e-10 of each class's largest entry, and
  against the same route with the device's step blocks (AssemblyDirect);
- against central differences of its own L (h = 1e-6 relative to the entry) at a few entries of
  each class, within 1e-7 of the class's largest entry plus the rounding floor of L.
"""
import numpy as np
import pytest

from tests.helpers import OracleAssemblyDirect, allsteps_lagrangian_blocks
from tests.test_gauge_identification import DT, NEL, NG, NSEG, NSTEP, problem, truth  # noqa: F401

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


@pytest.fixture(scope="module")
def point(problem):
    fwd, model, A, U, states, meas = problem
    rng = np.random.default_rng(13)
    nX = model.ndof["X"]
    X = np.array([s.X[0] for s in states]) + rng.normal(0., 1e-4, (NSTEP, nX))
    Λ = rng.normal(0., 1e-3, (NSTEP, nX))
    return (model, mc_initialize_dis(model), Λ, X, np.tile(0.9 * U, (NSTEP, 1)), A + rng.normal(0., 0.01, A.size),
            DT * np.arange(1, NSTEP + 1))


def mc_initialize_dis(model):
    import muscade_amd
    return muscade_amd.initialize(model).dis


@pytest.mark.parametrize("OX", [2, 1, 0])
def test_adjoint_route_matches_oracle_blocks(mc, point, OX):
    model, dis, Λ, X, U, A, time = point
    grad = mc.AllStepsGradient(OX, 0, 1, model, dis, NSTEP, DT)
    L, g = grad(Λ, X, U, A, time)
    for Assembly in (OracleAssemblyDirect, mc.AssemblyDirect):
        Lr, gr = allsteps_lagrangian_blocks(model, dis, Λ, X, U, A, time, DT, Assembly, OX=OX)
        assert abs(L - Lr) <= 1e-10 * max(abs(Lr), np.abs(Λ * gr["Λ"]).sum()), (Assembly, L, Lr)
        for c in ("Λ", "X", "U", "A"):
            scale = np.abs(gr[c]).max()
            assert np.abs(g[c] - gr[c]).max() <= 1e-10 * scale, (Assembly, c, np.abs(g[c] - gr[c]).max(), scale)


@pytest.mark.parametrize("OX", [2, 1])
def test_adjoint_route_difference_quotients(mc, point, OX):
    model, dis, Λ, X, U, A, time = point
    grad = mc.AllStepsGradient(OX, 0, 1, model, dis, NSTEP, DT)
    L, g = grad(Λ, X, U, A, time)
    for cls, k, i in [("X", 0, 30), ("X", 4, 40), ("X", NSTEP - 1, 51), ("Λ", 3, 31), ("U", 5, 7), ("A", None, 2)]:
        base = A[i] if cls == "A" else {"Λ": Λ, "X": X, "U": U}[cls][k, i]
        h = 1e-6 * max(1., abs(base))   # relative to the entry (loads are of order 100 N/m)
        fd = 0.
        for sgn in (1., -1.):
            v = {"Λ": Λ.copy(), "X": X.copy(), "U": U.copy(), "A": A.copy()}
            if cls == "A":
                v["A"][i] += sgn * h
            else:
                v[cls][k, i] += sgn * h
            fd += sgn * grad(v["Λ"], v["X"], v["U"], v["A"], time)[0]
        fd /= 2 * h
        got = g["A"][i] if cls == "A" else g[cls][k, i]
        # the difference quotient resolves no better than the rounding of L, about eps·|L|/h
        tol = 1e-7 * np.abs(g[cls]).max() + 8 * np.finfo(float).eps * abs(L) / h
        assert abs(got - fd) <= tol, (cls, k, i, got, fd, tol)
