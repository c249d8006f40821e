"""Oracle assembly (oracle.cpp: asmvec!/asmmat!/assemble!) pinned to test/TestAssemble.jl.

The model is Turbine + AnchorLine on 3 nodes (TestAssemble.jl:43-50); the goldens are the
asm index maps, the CSC pattern and the assembled Lλ / Lλx at X = 0 (:118-158). These need
the restated fixture elements of test/SomeElements.jl (oracle/elements.hpp).
"""
import math

import numpy as np

from tests import _oracle as orc

RTOL = math.sqrt(np.finfo(float).eps)


def approx(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) <= RTOL * max(np.linalg.norm(a), np.linalg.norm(b))


def build():
    om = orc.OracleModel(3, 0, 4)
    # Turbine: X (tx1,tx2) = dofs 1,2; A = 1,2; xm=(0,0), z=100, seadrag=2, skydrag=3
    om.add_type(orc.K_TURBINE, np.array([[1, 2]]), None, np.array([[1, 2]]), None, np.array([[0., 0., 100., 2., 3.]]))
    # AnchorLine: X (tx1,tx2,rx3) = 1,2,3; A = 3,4; xmtop=(0,0,100), Δxmtop=(5,0,0), xmbot=(150,0), L=180, buoyancy=-1e3
    om.add_type(orc.K_ANCHOR, np.array([[1, 2, 3]]), None, np.array([[3, 4]]), None,
                np.array([[0., 0., 100., 5., 0., 0., 150., 0., 180., -1e3]]))
    nnz = om.prepare()
    return om, nnz


def test_asm_and_pattern_golden():   # TestAssemble.jl:118-127
    om, nnz = build()
    assert nnz == 9   # AnchorLine couples its three X-dofs: a dense 3×3 pattern
    a1, a2 = om.asm(0, 1, 2)
    assert a1.tolist() == [[1, 2]] and a2.tolist() == [[1, 2, 4, 5]]
    a1, a2 = om.asm(1, 1, 3)
    assert a1.tolist() == [[1, 2, 3]] and a2.tolist() == [[1, 2, 3, 4, 5, 6, 7, 8, 9]]
    assert om.colptr.tolist() == [1, 4, 7, 10] and om.rowval.tolist() == [1, 2, 3] * 3


def test_assembled_values_golden():   # TestAssemble.jl:152-158
    om, nnz = build()
    r, rhs, nz, _ = om.assemble(0, 1, np.zeros(6), [np.zeros(3)], None, np.zeros(4))
    assert r == 0
    assert approx(rhs, [-152130.71199858442, -3.0, 0.0])
    # the golden is the dense 3×3 of Lλx; scatter our CSC into it
    K = np.zeros((3, 3))
    for j in range(3):
        for p in range(om.colptr[j] - 1, om.colptr[j + 1] - 1):
            K[om.rowval[p] - 1, j] = nz[p]
    gold = np.array([10323.069597975566, 0.0, 0.0, 0.0, 1049.1635310247202, 5245.817655123601, 0.0, 5245.8176551236,
                     786872.6482685402]).reshape(3, 3, order="F")
    assert approx(K, gold)


def test_pattern_bruteforce_random():
    """The sortperm/run-length pattern equals a dense brute-force build on random small models,
    and asm[2] points at the entry of its (row, column) (Assemble.jl:373-448)."""
    rng = np.random.default_rng(0)
    for trial in range(20):
        nX = int(rng.integers(3, 15))
        om = orc.OracleModel(nX)
        types = []
        for t in range(int(rng.integers(1, 4))):
            ne, nd = int(rng.integers(1, 5)), int(rng.integers(1, 4))
            iX = np.stack([rng.choice(np.arange(1, nX + 1), nd, replace=False) for _ in range(ne)])
            om.add_type(orc.K_ZERO, iX)
            types.append(iX)
        nnz = om.prepare()
        dense = np.zeros((nX, nX), bool)
        for iX in types:
            for e in iX:
                for i in e:
                    for j in e:
                        dense[i - 1, j - 1] = True
        assert nnz == dense.sum()
        for j in range(nX):
            rows = om.rowval[om.colptr[j] - 1:om.colptr[j + 1] - 1] - 1
            assert rows.tolist() == np.nonzero(dense[:, j])[0].tolist()
        for it, iX in enumerate(types):
            ne, nd = iX.shape
            a1, a2 = om.asm(it, ne, nd)
            for e in range(ne):
                for j in range(nd):
                    for i in range(nd):
                        k = a2[e, i + nd * j] - 1
                        assert om.rowval[k] == iX[e, i]
                        assert om.colptr[iX[e, j] - 1] - 1 <= k < om.colptr[iX[e, j]] - 1


def test_parallel_assembly_equals_sequential():
    """The all-cores CPU baseline (element evaluations on OpenMP threads, then the ordered
    add_value!/add_∂!) gives the sequential assembly bit for bit, Holds included."""
    from tests.helpers import build_lattice, oracle_model
    model, state = build_lattice(4, 3, 3, seed=2)
    om = oracle_model(model, state.dis)
    r1, rhs1, nz1, _ = om.assemble(0, 1, np.zeros(6), [state.X[0]])
    r2, rhs2, nz2, _ = om.assemble0_parallel(4, state.X[0])
    assert r1 == 0 and r2 == 0
    assert np.array_equal(rhs1.view(np.uint64), rhs2.view(np.uint64))
    assert np.array_equal(nz1.view(np.uint64), nz2.view(np.uint64))
