"""GPU parity of Bar3D (toolbox/BarElement.jl): the HIP bar kernel against the oracle.

Bar (FP64): R, K + a₁C + b₁M and the :step rhs within 1e-12 normwise relative
error (max|gpu-oracle| / max|oracle|); CSC pattern bit-exact.
"""
import numpy as np
import pytest

from tests import _oracle as orc
from tests.helpers import OracleAssembly, oracle_model, relerr

pytestmark = pytest.mark.gpu
TOL = 1e-12
DT = 0.05


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


def random_bars(mc, n, seed, scale=0.2, udof=False):
    rng = np.random.default_rng(seed)
    model = mc.Model("bars")
    c1 = rng.normal(size=(n, 3))
    c2 = c1 + rng.normal(size=(n, 3)) * 2.
    xyz = np.empty((2 * n, 3))
    xyz[0::2], xyz[1::2] = c1, c2
    mc.addnode(model, xyz)
    mat = np.zeros((n, 9))
    mat[:, 0] = rng.uniform(1., 10., n)
    mat[:, 1] = rng.uniform(.5, 2., n)
    mat[:, 2] = rng.uniform(-2., 2., n)
    mat[:, 3:] = rng.uniform(0., 1., (n, 6))
    nodes = np.stack([np.arange(1, 2 * n, 2), np.arange(2, 2 * n + 1, 2)], axis=1)
    eps_s = rng.uniform(0., 1e-3, n)
    if udof:
        mc.addnode(model, rng.normal(size=(n, 3)))
        nodes = np.concatenate([nodes, np.arange(2 * n + 1, 3 * n + 1)[:, None]], axis=1)
        mc.addelement(model, mc.Bar3DU, nodes, mat=mat, eps_s=eps_s)
    else:
        mc.addelement(model, mc.Bar3D, nodes, mat=mat, eps_s=eps_s)
    state = mc.initialize(model, nXder=3)
    state.X[0][:] = rng.normal(size=model.ndof["X"]) * scale
    state.X[1][:] = rng.normal(size=model.ndof["X"])
    state.X[2][:] = rng.normal(size=model.ndof["X"])
    if udof:
        state.U[0][:] = rng.normal(size=model.ndof["U"])
    state.time = -7.
    return model, state, xyz, nodes, mat, eps_s


@pytest.mark.parametrize("OX", [0, 1, 2])
def test_bar_element_blocks(mc, OX):
    n = 300
    model, state, xyz, nodes, mat, eps_s = random_bars(mc, n, seed=40 + OX)
    out = mc.AssemblySweepX(OX, model, state.dis)
    out.c = mc.Newmarkβcoefficients(OX, DT, 0.25, 0.5)
    out.assemble(state, "iter")
    R, K = out.element_blocks(0)
    d = state.dis.dis[0]
    worst = 0.
    for e in range(n):
        X = [state.X[k][d.X[e] - 1] for k in range(OX + 1)]
        Ro, Ks, _ = orc.bar_diffed_residual(xyz[nodes[e, 0] - 1], xyz[nodes[e, 1] - 1], mat[e], X, t=state.time,
                                            eps_s=eps_s[e])
        Ko = Ks[0] + (out.c.a1 * Ks[1] if OX >= 1 else 0.) + (out.c.b1 * Ks[2] if OX == 2 else 0.)
        worst = max(worst, relerr(R[e], Ro), relerr(K[e], Ko))
    assert worst <= TOL, worst


@pytest.mark.parametrize("OX,mission,udof", [(0, "iter", False), (1, "step", False), (2, "iter", True),
                                             (2, "step", False), (2, "step", True)])
def test_bar_assembly_vs_oracle(mc, OX, mission, udof):
    model, state, *_ = random_bars(mc, 200, seed=3 + OX, udof=udof)
    out = mc.AssemblySweepX(OX, model, state.dis)
    out.c = mc.Newmarkβcoefficients(OX, DT, 0.25, 0.5)
    Lλ, nzval = out.assemble(state, mission)
    om = oracle_model(model, state.dis, t=state.time)
    r, rhs, nz, err = om.assemble(OX, 0 if mission == "step" else 1, out.c.array(), state.X[:OX + 1],
                                  state.U[0] if udof else None, None)
    assert r == 0, err
    assert np.array_equal(out.colptr, om.colptr) and np.array_equal(out.rowval, om.rowval)
    assert relerr(Lλ, rhs) <= TOL and relerr(nzval, nz) <= TOL


def tower(mc, nlev=6, w=0.):
    """A square lattice tower of Bar3D: 4 columns, rings and diagonals per level, base pinned,
    horizontal load on the top ring, plus two EulerBeam3D on the top (a mixed model)."""
    model = mc.Model("tower")
    xyz = []
    for k in range(nlev + 1):
        for (x, y) in ((0., 0.), (1., 0.), (1., 1.), (0., 1.)):
            xyz.append((x, y, 1.2 * k))
    mc.addnode(model, np.array(xyz))
    nid = lambda k, c: 1 + 4 * k + c   # noqa: E731
    bars = []
    for k in range(nlev):
        for c in range(4):
            bars.append((nid(k, c), nid(k + 1, c)))                 # column
            bars.append((nid(k + 1, c), nid(k + 1, (c + 1) % 4)))   # ring
            bars.append((nid(k, c), nid(k + 1, (c + 1) % 4)))       # diagonal
    mat = mc.AxisymmetricBarCrossSection(EA=2e8, mu=10., w=w)
    mc.addelement(model, mc.Bar3D, np.array(bars), mat=mat)
    top = [nid(nlev, c) for c in range(4)]
    sec = mc.BeamCrossSection(EA=2e8, EI2=1e5, EI3=1e5, GJ=8e4, mu=10., iota1=0.1)
    mc.addelement(model, mc.EulerBeam3D, np.array([[top[0], top[2]], [top[1], top[3]]]), mat=sec,
                  orient2=np.array([0., 0., 1.]))
    for f in ("t1", "t2", "t3"):
        mc.addelement(model, mc.Hold, np.array([[nid(0, c)] for c in range(4)]), field=f)
    for f in ("r1", "r2", "r3"):   # the beams' rotations at the top are held
        mc.addelement(model, mc.Hold, np.array([[n] for n in top]), field=f)

    def push(t, F=2e4):
        return F * t
    mc.addelement(model, mc.DofLoad, np.array([[n] for n in top]), field="t1", value=push)
    return model


def test_mixed_tower_assembly(mc):
    model = tower(mc, w=50.)
    state = mc.initialize(model, nXder=3)
    rng = np.random.default_rng(5)
    state.X[0][:] = rng.normal(size=model.ndof["X"]) * 1e-2
    state.X[1][:] = rng.normal(size=model.ndof["X"]) * 0.1
    state.X[2][:] = rng.normal(size=model.ndof["X"])
    state.time = -6.
    for OX, mission in ((0, "iter"), (2, "step"), (2, "iter")):
        out = mc.AssemblySweepX(OX, model, state.dis)
        out.c = mc.Newmarkβcoefficients(OX, DT, 0.25, 0.5)
        Lλ, nzval = out.assemble(state, mission)
        om = oracle_model(model, state.dis, t=state.time)
        r, rhs, nz, err = om.assemble(OX, 0 if mission == "step" else 1, out.c.array(), state.X[:OX + 1], None, None)
        assert r == 0, err
        assert np.array_equal(out.colptr, om.colptr) and np.array_equal(out.rowval, om.rowval)
        assert relerr(Lλ, rhs) <= TOL and relerr(nzval, nz) <= TOL, (OX, mission)


def test_sweepx0_tower_matches_oracle_solve(mc):
    """SweepX{0} on the mixed bar/beam tower under a ramped load: GPU assembly and the oracle's
    assembler give the same load path."""
    model = tower(mc)
    st0 = mc.initialize(model, time=0.)
    kw = dict(initialstate=st0, time=np.arange(1., 5.), maxΔx=1e-6, verbose=False, catcherror=False)
    gpu = mc.solve(mc.SweepX[0], **kw)
    ref = mc.solve(mc.SweepX[0], Assembly=OracleAssembly, **kw)
    nodes = np.arange(1, model.nnod + 1)
    for sg, so in zip(gpu, ref):
        for f in ("t1", "t2", "t3"):   # displacements, then the reactions (λ) separately
            ug, uo = mc.getdof(sg, field=f, nodID=nodes), mc.getdof(so, field=f, nodID=nodes)
            assert np.abs(ug - uo).max() <= 1e-9 * np.abs(uo).max() + 1e-15, f
        lam = model.getdofID("X", "λt1") - 1
        assert np.abs(sg.X[0][lam] - so.X[0][lam]).max() <= 1e-9 * np.abs(so.X[0][lam]).max()
    tip = mc.getdof(gpu[-1], field="t1", nodID=[25])[0]
    assert 0.1 < tip < 0.5
