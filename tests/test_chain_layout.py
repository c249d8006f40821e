"""Chain layout for the device linear solve (muscade_amd/chainsolve.py) and the block cyclic
reduction it feeds (chain_solve.hpp), checked on the CPU:

- a numpy restatement of the device algorithm (pack, λ-pair rows, cyclic reduction levels,
  back substitution, λ recovery) solves the oracle-assembled Lλx·Δx = Lλ of chain models
  as a backward-stable solver does: normwise backward error
  ‖Lλx·Δx − Lλ‖∞ / (‖Lλx‖∞‖Δx‖∞ + ‖Lλ‖∞) ≤ 1e-14, and agrees with SciPy's sparse LU (the
  reference's `Lλx\\Lλ`, src/SweepX.jl:205-209) to 1e-10 where Lλx is well conditioned.
  The dynamic chains are not (condition number ~1e19, from Hold rows and EA = 5e9 against
  the mass terms): there sparse LU and dense LU themselves differ by ~1e-8, and the
  backward error is the measure;
- the layout takes every position once, follows the paths node by node whatever the
  element order, and pairs each Hold's λ dof with its held dof;
- branched, looped and lattice models are refused.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import muscade_amd as mc
from muscade_amd.chainsolve import NotAChain, chain_layout
from muscade_amd.workloads import base_section, catenary_chains
from tests.helpers import OracleAssembly, relerr


def bcr_emulate(K, r, lay):
    """The device solve, level by level, in numpy (K: scipy CSC, r: rhs)."""
    B, nb = lay.B, lay.nblock
    D = np.zeros((nb, B, B)); L = np.zeros((nb, B, B)); U = np.zeros((nb, B, B)); b = np.zeros((nb, B))
    npair = lay.pair_hpos.size
    pc = np.zeros((npair, 4))
    for q in lay.dummy_pos - 1:
        D[q // B, q % B, q % B] = 1.
    Kc = K.tocoo()
    pos = lay.pos
    for i, j, v in zip(Kc.row, Kc.col, Kc.data):
        pr, pcl = pos[i], pos[j]
        if pr > 0 and pcl > 0:
            pr -= 1; pcl -= 1
            br, bc = pr // B, pcl // B
            M = {0: D, 1: L, -1: U}[br - bc]
            M[br, pr % B, pcl % B] = v
        elif pr < 0 and pcl < 0:
            assert pr == pcl
            pc[-pr - 1, 2] = v
        elif pcl < 0:
            pc[-pcl - 1, 1] = v
        else:
            pc[-pr - 1, 0] = v
    for d in range(K.shape[0]):
        if pos[d] > 0:
            b[(pos[d] - 1) // B, (pos[d] - 1) % B] = r[d]
        else:
            pc[-pos[d] - 1, 3] = r[d]
    saved = np.zeros((npair, 3 * B + 1))
    for p in range(npair):
        h = lay.pair_hpos[p] - 1
        k, i = h // B, h % B
        assert pc[p, 2] == 0.
        saved[p] = np.concatenate([L[k, i], D[k, i], U[k, i], [b[k, i]]])
        L[k, i] = 0.; U[k, i] = 0.; D[k, i] = 0.; D[k, i, i] = 1.
        b[k, i] = pc[p, 3] / pc[p, 0]
    s = 1
    while s < nb:
        el = np.arange(s, nb, 2 * s)
        for i in el:
            rhs = np.concatenate([L[i], U[i], b[i][:, None]], axis=1)
            sol = np.linalg.solve(D[i], rhs)
            L[i], U[i], b[i] = sol[:, :B], sol[:, B:2 * B], sol[:, 2 * B]
        for j in np.arange(0, nb, 2 * s):
            nL = np.zeros((B, B)); nU = np.zeros((B, B))
            if j >= s:
                l = j - s
                D[j] -= L[j] @ U[l]; b[j] -= L[j] @ b[l]; nL = -L[j] @ L[l]
            if j + s < nb:
                q = j + s
                D[j] -= U[j] @ L[q]; b[j] -= U[j] @ b[q]; nU = -U[j] @ U[q]
            L[j], U[j] = nL, nU
        s *= 2
    b[0] = np.linalg.solve(D[0], b[0])
    s //= 2
    while s >= 1:
        for i in np.arange(s, nb, 2 * s):
            x = b[i] - L[i] @ b[i - s]
            if i + s < nb:
                x -= U[i] @ b[i + s]
            b[i] = x
        s //= 2
    x = b.ravel()
    dx = np.zeros(K.shape[0])
    for d in range(K.shape[0]):
        if pos[d] > 0:
            dx[d] = x[pos[d] - 1]
        else:
            p = -pos[d] - 1
            h = lay.pair_hpos[p] - 1
            k = h // B
            rr = saved[p, 3 * B]
            if k > 0:
                rr -= saved[p, :B] @ x[(k - 1) * B:k * B]
            rr -= saved[p, B:2 * B] @ x[k * B:(k + 1) * B]
            if k + 1 < nb:
                rr -= saved[p, 2 * B:3 * B] @ x[(k + 1) * B:(k + 2) * B]
            dx[d] = rr / pc[p, 1]
    return dx


def backward_error(K, x, r):
    Ka = abs(K)
    return np.abs(K @ x - r).max() / (Ka.sum(axis=1).max() * np.abs(x).max() + np.abs(r).max())


def cantilever_model(nel=7, reverse=False, shuffle_seed=None):
    model = mc.Model("cl")
    x = np.linspace(0., 3., nel + 1)
    mc.addnode(model, np.stack([x, 0.1 * x * x, 0 * x], axis=1))
    con = np.stack([np.arange(1, nel + 1), np.arange(2, nel + 2)], axis=1)
    if reverse:
        con = con[::-1, ::-1].copy()
    if shuffle_seed is not None:
        con = con[np.random.default_rng(shuffle_seed).permutation(nel)]
    mc.addelement(model, mc.EulerBeam3D, con, mat=base_section(), orient2=np.array([0., 1., 0.]))
    for f in ("t1", "t2", "t3", "r1", "r2", "r3"):
        mc.addelement(model, mc.Hold, [1], field=f)
    mc.addelement(model, mc.DofLoad, [nel + 1], field="t3", value=lambda t: 1e3 * t)
    return model


def assembled(model, state, OX, mission="iter", Δt=0.05):
    out = OracleAssembly(OX, model, state.dis)
    if OX:
        out.c = mc.Newmarkβcoefficients(OX, Δt)
    r, _ = out.assemble(state, mission)
    return out.Lλx(), r.copy()


def random_state(model, OX, seed=0, scale=1e-3):
    st = mc.initialize(model, nXder=3, time=1.0)
    rng = np.random.default_rng(seed)
    for d in range(OX + 1):
        st.X[d][:] = rng.normal(0., scale, model.ndof["X"])
    return st


@pytest.mark.parametrize("reverse,shuffle", [(False, None), (True, None), (False, 3)])
def test_cantilever_static(reverse, shuffle):
    model = cantilever_model(reverse=reverse, shuffle_seed=shuffle)
    st = random_state(model, 0)
    K, r = assembled(model, st, 0)
    col = K.indptr.astype(np.int64) + 1
    row = K.indices.astype(np.int64) + 1
    lay = chain_layout(model, st.dis, col, row)
    assert lay.B == 6 and lay.nblock == 8 and lay.pair_hpos.size == 6 and lay.dummy_pos.size == 0
    # path order: node 1 (the held end) first, then along the beam
    assert list(lay.node_order) == list(range(8))
    taken = np.sort(lay.pos[lay.pos > 0])
    assert np.array_equal(taken, np.arange(1, 6 * 8 + 1))
    ref = spla.spsolve(K.tocsc(), r)
    dx = bcr_emulate(K, r, lay)
    assert relerr(dx, ref) <= 1e-10
    assert backward_error(K, dx, r) <= 1e-14


def test_catenary_chains_dynamic():
    model, st = catenary_chains(nchain=3, nmin=5, nmax=13, seed=11)
    st.time = 1.0
    K, r = assembled(model, st, 2, mission="step")
    lay = chain_layout(model, st.dis, K.indptr.astype(np.int64) + 1, K.indices.astype(np.int64) + 1)
    assert lay.pair_hpos.size == 9   # t1..t3 held at each anchor
    dx = bcr_emulate(K, r, lay)
    assert backward_error(K, dx, r) <= 1e-14
    assert backward_error(K, spla.spsolve(K.tocsc(), r), r) <= 1e-14


def bar_chain(n=9):
    model = mc.Model("bars")
    x = np.linspace(0., 8., n + 1)
    mc.addnode(model, np.stack([x, 0 * x, -0.05 * x * (8. - x)], axis=1))
    sec = mc.AxisymmetricBarCrossSection(EA=1e7, mu=10., w=50., Can=1., Cln=0.5, Cqn=5.)
    mc.addelement(model, mc.Bar3D, np.stack([np.arange(1, n + 1), np.arange(2, n + 2)], axis=1), mat=sec)
    for f in ("t1", "t2", "t3"):
        mc.addelement(model, mc.Hold, [1], field=f)
        mc.addelement(model, mc.Hold, [n + 1], field=f)
    return model


def test_bar_chain_three_dof_blocks():
    model = bar_chain()
    st = random_state(model, 2, scale=1e-2)
    K, r = assembled(model, st, 2, mission="step")
    lay = chain_layout(model, st.dis, K.indptr.astype(np.int64) + 1, K.indices.astype(np.int64) + 1)
    assert lay.B == 3 and lay.nblock == 10 and lay.pair_hpos.size == 6
    ref = spla.spsolve(K.tocsc(), r)
    dx = bcr_emulate(K, r, lay)
    assert backward_error(K, dx, r) <= 1e-14
    print("bar chain: rel. diff to sparse LU", relerr(dx, ref))


def test_refuses_non_chains():
    from tests.helpers import build_lattice, oracle_model

    def layout(m, dis):
        om = oracle_model(m, dis)
        return chain_layout(m, dis, np.asarray(om.colptr, np.int64), np.asarray(om.rowval, np.int64))

    model, st = build_lattice(3, 3, 2, seed=1)
    with pytest.raises(NotAChain, match="branch"):
        layout(model, st.dis)
    # a Y branch
    m = mc.Model("y")
    mc.addnode(m, np.array([[0., 0, 0], [1, 0, 0], [2, 1, 0], [2, -1, 0]]))
    mc.addelement(m, mc.EulerBeam3D, np.array([[1, 2], [2, 3], [2, 4]]), mat=base_section(),
                  orient2=np.array([0., 0., 1.]))
    with pytest.raises(NotAChain, match="branch"):
        layout(m, mc.initialize(m).dis)
    # a closed ring
    m = mc.Model("ring")
    a = np.linspace(0., 2 * np.pi, 7)[:-1]
    mc.addnode(m, np.stack([np.cos(a), np.sin(a), 0 * a], axis=1))
    mc.addelement(m, mc.EulerBeam3D, np.stack([np.arange(1, 7), np.roll(np.arange(1, 7), -1)], axis=1),
                  mat=base_section(), orient2=np.array([0., 0., 1.]))
    with pytest.raises(NotAChain, match="loop"):
        layout(m, mc.initialize(m).dis)
    # a dof held twice
    m = cantilever_model(nel=3)
    mc.addelement(m, mc.Hold, [1], field="t1", λfield="μt1")
    with pytest.raises(NotAChain, match="held by several"):
        layout(m, mc.initialize(m).dis)


def test_two_chains_and_a_free_node():
    """Chains in one model with nodes numbered out of path order, one ending in a bar whose
    far node has 3 dofs (padded to a 6-position block)."""
    m = mc.Model("two")
    x = np.linspace(0., 2., 5)
    mc.addnode(m, np.stack([x, 0 * x, 0 * x], axis=1))           # nodes 1-5: chain A
    mc.addnode(m, np.stack([x[::-1], 1 + 0 * x, 0 * x], axis=1))  # nodes 6-10: chain B, reversed
    mc.addnode(m, np.array([[2.5, 0., -0.2]]))                    # node 11: end of a bar on chain A
    mc.addelement(m, mc.EulerBeam3D, np.array([[6, 7], [3, 4], [1, 2], [8, 9], [2, 3], [10, 9], [4, 5], [7, 8]]),
                  mat=base_section(), orient2=np.array([0., 0., 1.]))
    for f in ("t1", "t2", "t3", "r1", "r2", "r3"):
        mc.addelement(m, mc.Hold, [1], field=f)
        mc.addelement(m, mc.Hold, [10], field=f)
    mc.addelement(m, mc.Bar3D, [5, 11], mat=mc.AxisymmetricBarCrossSection(EA=1e6, mu=5., w=10.))
    st = random_state(m, 2, seed=4)
    K, r = assembled(m, st, 2, mission="step")
    lay = chain_layout(m, st.dis, K.indptr.astype(np.int64) + 1, K.indices.astype(np.int64) + 1)
    assert list(lay.node_order) == [0, 1, 2, 3, 4, 10, 5, 6, 7, 8, 9]
    assert lay.B == 6 and lay.dummy_pos.size == 3   # node 11 has three X dofs
    dx = bcr_emulate(K, r, lay)
    assert backward_error(K, dx, r) <= 1e-14
