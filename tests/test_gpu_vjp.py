"""GPU adjoint (vector-Jacobian products) of the residual Lagrangian L = Σ_e Λ_eᵀ R_e.

Checked against the oracle's forward-mode Jacobians (Λᵀ·∂R/∂X⁽ᵈ⁾, Λᵀ·∂R/∂U from
diffed_residual, i.e. Muscade's Adiff gradient restated) within 1e-12 normwise relative
error, and against central differences of the GPU-assembled residual.
"""
import numpy as np
import pytest

from tests import _oracle as orc
from tests.helpers import relerr
from tests.test_gpu_bar import random_bars, tower
from tests.test_gpu_dynamic import dyn_beams

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


def oracle_vjp(model, state, lam, OX, kinds):
    """Σ_e scatter(Λ_eᵀ ∂R_e/∂X⁽ᵈ⁾) and Λ_eᵀ ∂R_e/∂U from the oracle's element Jacobians."""
    nX, nU = model.ndof["X"], model.ndof["U"]
    gX = [np.zeros(nX) for _ in range(OX + 1)]
    gU = np.zeros(nU)
    xyz = model.coords()
    for it, E in enumerate(model.eletypes):
        d = state.dis.dis[it]
        nodes = E.cat_nodes()
        for e in range(d.X.shape[0]):
            ix = d.X[e] - 1
            X = [state.X[k][ix] for k in range(OX + 1)]
            U = state.U[0][d.U[e] - 1] if d.U.shape[1] else None
            le = lam[ix]
            if E.cls.kind in (1, 2):
                p = E.params[0] if len(E.params) == 1 else None
                mat = np.concatenate([q["mat"] for q in E.params])[e]
                o2 = np.concatenate([q["orient2"] for q in E.params])[e]
                R, Ks, dU = orc.beam_diffed_residual(xyz[nodes[e, 0], :3], xyz[nodes[e, 1], :3], o2, mat, X, U)
            elif E.cls.kind in (6, 7):
                mat = np.concatenate([q["mat"] for q in E.params])[e]
                eps = np.concatenate([q["eps_s"] for q in E.params])[e]
                R, Ks, dU = orc.bar_diffed_residual(xyz[nodes[e, 0], :3], xyz[nodes[e, 1], :3], mat, X, U=U,
                                                    t=state.time, eps_s=eps)
            elif E.cls.kind == 3:   # Hold: R = (-λ, -x)
                Ks = [np.array([[0., -1.], [-1., 0.]])] + [np.zeros((2, 2))] * OX
                dU = None
            else:
                continue
            for k in range(OX + 1):
                np.add.at(gX[k], ix, le @ Ks[k])
            if dU is not None and U is not None:
                np.add.at(gU, d.U[e] - 1, le @ dU)
    return gX, gU


def check(gX, gU, ref):
    rX, rU = ref
    for k in range(len(gX)):
        assert relerr(gX[k], rX[k]) <= TOL, k
    if rU.size:
        assert relerr(gU, rU) <= TOL


@pytest.mark.parametrize("udof", [False, True])
def test_vjp_static_beams(mc, udof):
    model, state, *_ = dyn_beams(mc, 200, seed=21, rotscale=0.5, udof=udof)
    lam = np.random.default_rng(1).normal(size=model.ndof["X"])
    out = mc.AssemblySweepX(0, model, state.dis)
    gX, gU, gA = out.vjp(state, lam, OX=0)
    check(gX, gU, oracle_vjp(model, state, lam, 0, None))


@pytest.mark.parametrize("OX", [1, 2])
@pytest.mark.parametrize("udof", [False, True])
@pytest.mark.parametrize("nbeam", [150, 7, 253])   # whole and partial tiles of 10 elements
def test_vjp_dynamic_beams(mc, OX, udof, nbeam):
    model, state, *_ = dyn_beams(mc, nbeam, seed=30 + OX, rotscale=0.4, udof=udof)
    lam = np.random.default_rng(2).normal(size=model.ndof["X"])
    out = mc.AssemblySweepX(OX, model, state.dis)
    gX, gU, _ = out.vjp(state, lam)
    check(gX, gU, oracle_vjp(model, state, lam, OX, None))


@pytest.mark.parametrize("OX", [0, 1, 2])
def test_vjp_bars(mc, OX):
    model, state, *_ = random_bars(mc, 200, seed=50 + OX, udof=OX == 2)
    lam = np.random.default_rng(3).normal(size=model.ndof["X"])
    out = mc.AssemblySweepX(OX, model, state.dis)
    gX, gU, _ = out.vjp(state, lam)
    check(gX, gU, oracle_vjp(model, state, lam, OX, None))


def test_vjp_tower_and_directional_differences(mc):
    """Mixed bars, beams, Holds, DofLoad: the adjoint against the oracle, and its contraction
    with random directions against central differences of Λ·R assembled on the GPU."""
    model = tower(mc, w=40.)
    state = mc.initialize(model)
    rng = np.random.default_rng(4)
    state.X[0][:] = rng.normal(size=model.ndof["X"]) * 1e-2
    state.time = -6.
    lam = rng.normal(size=model.ndof["X"])
    out = mc.AssemblySweepX(0, model, state.dis)
    gX, gU, _ = out.vjp(state, lam)
    check(gX, gU, oracle_vjp(model, state, lam, 0, None))
    h = 1e-6
    for _ in range(4):
        v = rng.normal(size=model.ndof["X"])
        sp, sm = state.copy(), state.copy()
        sp.X[0] = state.X[0] + h * v
        sm.X[0] = state.X[0] - h * v
        Lp = lam @ out.assemble(sp, "iter")[0]
        Lm = lam @ out.assemble(sm, "iter")[0]
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - gX[0] @ v) <= 1e-6 * max(1., abs(fd)), (fd, gX[0] @ v)


def test_vjp_is_deterministic(mc):
    model, state, *_ = dyn_beams(mc, 300, seed=8, rotscale=0.3)
    lam = np.random.default_rng(9).normal(size=model.ndof["X"])
    out = mc.AssemblySweepX(2, model, state.dis)
    a = [g.copy() for g in out.vjp(state, lam)[0]]
    b = out.vjp(state, lam)[0]
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_vjp_batched_equals_plain(mc):
    """Uneven cantilevers with U-loads shared by whole stretches of beams (long U segments):
    the gradients of the joined batch equal each model's own, bit for bit, and match the oracle."""
    from muscade_amd.workloads import cantilever_batch, cantilever_batch_specs
    specs = cantilever_batch_specs(6, 150, 700, seed=77)
    [(mj, sj, lj)] = cantilever_batch(specs)
    sep = cantilever_batch(specs, joined=False)
    gXj, gUj, _ = mc.AssemblySweepX(0, mj, sj.dis).vjp(sj, lj)
    xo = 0
    for k, (m, s, l) in enumerate(sep):
        gX, gU, _ = mc.AssemblySweepX(0, m, s.dis).vjp(s, l)
        nd = 6 * (specs[k]["ne"] + 1)
        assert np.array_equal(gXj[0][xo:xo + nd], gX[0][:nd])
        assert np.array_equal(gUj[24 * k:24 * k + 24], gU)
        xo += nd
        if k < 2:
            check(gX, gU, oracle_vjp(m, s, l, 0, None))
