"""The oracle's AssemblyDirect restatement (oracle.cpp assemble_direct, DirectXUA.jl:22-171)
checked against independent pieces: the class-pair patterns against a dense brute force,
L2[Λ,X][1,d] and L2[X,Λ][d,1] against the element Jacobians of diffed_residual assembled
densely, and the Lagrangian path of Hold (L = Λ∘₁R) against its closed form.

The goldens of test/TestDirectXUA.jl:93-136 use fixture elements with user cost
functions (El1, Spring, SingleAcost, SingleUdof, SingleDofCost) that are not restated
here; for those values this restatement is pinned only through the pieces above."""
import numpy as np

from tests import _oracle as orc


def beam_model(seed, nbeam=6, OX=2, dt=0.1):
    rng = np.random.default_rng(seed)
    nnod = nbeam + 1
    xyz = np.cumsum(rng.uniform(0.5, 1.5, (nnod, 3)), axis=0)
    mat = np.array([5., 3., 4., 2., 1.2, 0.4, 0.5, .3, .2, .1, .6, .3, .2, .6, .3, .2])
    iX = np.array([[6 * e + k + 1 for k in range(12)] for e in range(nbeam)])
    nXb = 6 * nnod
    holds = np.array([[1, nXb + 1], [2, nXb + 2], [3, nXb + 3]])
    nX = nXb + 3
    om = orc.OracleModel(nX)
    params = np.array([np.concatenate([xyz[e], xyz[e + 1], [0., 0., 1.], mat]) for e in range(nbeam)])
    om.add_type(orc.K_BEAM, iX, None, None, None, params)
    om.add_type(orc.K_HOLD, holds)
    om.add_type(orc.K_DOFLOAD, np.array([[nXb - 2]]), params=np.array([3.0]))
    X = [rng.normal(size=nX) * s for s in (0.1, 0.5, 1.0)][:OX + 1]
    lam = rng.normal(size=nX)
    return om, iX, holds, xyz, params, X, lam, nX


def dense(cp, rv, v, nr, nc):
    M = np.zeros((nr, nc))
    for j in range(nc):
        for p in range(cp[j] - 1, cp[j + 1] - 1):
            M[rv[p] - 1, j] += v[p]
    return M


def test_direct_blocks_match_element_jacobians():
    OX, dt = 2, 0.1
    om, iX, holds, xyz, params, X, lam, nX = beam_model(3, OX=OX)
    nnz = om.prepare_direct(OX, 0, 0)
    r, L1, L2 = om.assemble_direct(1, dt, lam, X)
    assert r == 0
    # dense reference: Σ_e ∂R_e/∂X⁽ᵈ⁾ (beams) + Hold's ∂R/∂X for d = 0
    ref = [np.zeros((nX, nX)) for _ in range(OX + 1)]
    R = np.zeros(nX)
    for e in range(iX.shape[0]):
        ix = iX[e] - 1
        Re, Ks, _ = orc.beam_diffed_residual(params[e, :3], params[e, 3:6], params[e, 6:9], params[e, 9:],
                                             [x[ix] for x in X])
        R[ix] += Re * dt
        for d in range(OX + 1):
            ref[d][np.ix_(ix, ix)] += Ks[d] * dt
    for x, lm in holds - 1:
        ref[0][x, lm] += -dt
        ref[0][lm, x] += -dt
        R[x] += -X[0][lm] * dt
        R[lm] += -X[0][x] * dt
    R[(nX - 3) - 3] += -3.0 * dt          # DofLoad R = -F on its dof (1-based nX-5)
    cp, rv = om.dpattern[(0, 1)]
    nder = OX + 1
    # L2 layout: pairs (Λ,X),(Λ,U),(X,Λ),(X,X),(X,U),(U,Λ),(U,X),(U,U)
    n = nnz[0, 1]
    for d in range(nder):
        got = dense(cp, rv, L2[d * n:(d + 1) * n], nX, nX)
        assert np.abs(got - ref[d]).max() <= 1e-12 * np.abs(ref[d]).max()
    # (X,Λ) blocks follow (Λ,X) and (Λ,U) (empty: no U dofs)
    off = nder * n
    for d in range(nder):
        got = dense(cp, rv, L2[off + d * n:off + (d + 1) * n], nX, nX)
        assert np.abs(got - ref[d].T).max() <= 1e-12 * np.abs(ref[d]).max()
    # L1[Λ] = R·Δt (beams: residual path; Hold, DofLoad: Lagrangian path with scale.Λ = 1)
    assert np.abs(L1[:nX] - R).max() <= 1e-12 * np.abs(R).max()
    # L1[X][1] = Λᵀ∂R/∂X of the Holds (beams add nothing to L1[X] on the residual path)
    gX = np.zeros(nX)
    for x, lm in holds - 1:
        gX[x] += -lam[lm] * dt
        gX[lm] += -lam[x] * dt
    assert np.abs(L1[nX:2 * nX] - gX).max() <= 1e-14
    assert np.abs(L1[2 * nX:]).max() == 0.


def test_direct_patterns_bruteforce():
    om, iX, holds, *_ , nX = beam_model(5, OX=1)
    nnz = om.prepare_direct(1, 0, 0)
    M = np.zeros((nX, nX), bool)
    for dofs in list(iX) + list(holds) + [np.array([nX - 5])]:
        for i in dofs:
            for j in dofs:
                M[i - 1, j - 1] = True
    for a, b in ((0, 1), (1, 0), (1, 1)):
        cp, rv = om.dpattern[(a, b)]
        assert nnz[a, b] == M.sum()
        for j in range(nX):
            assert (rv[cp[j] - 1:cp[j + 1] - 1] - 1).tolist() == np.nonzero(M[:, j])[0].tolist()
