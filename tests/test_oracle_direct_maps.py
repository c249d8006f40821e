"""The integer maps of DirectXUA's all-steps system: the oracle (oracle/direct_maps.py) against
the reference's goldens, and the product (muscade_amd.directxua.class_patterns, BigPattern,
sparser) against the oracle, bit for bit.

Goldens: test/TestDirectXUA.jl:109-136 (asm of the class arrays, Lvv.colptr[1:50],
Lvv.rowval[1:60], Lvasm, Lvdis, Lvvasm) on the model of TestDirectXUA.jl:23-44, and
test/TestSparseTools.jl:7-62 (block assembly, bigsparseasm, sparser!)."""
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from tests.fixture_elements import El1, Spring1

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import direct_maps as om  # noqa: E402


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


def fa(a):
    return a ** 2 * 1e-14


def fu(u, t):
    return u ** 2


def l1(tx1, t):
    return (tx1 - 0.1 * np.sin(t)) ** 2


def l2(tx1, t):
    return (tx1 - 0.1 * np.cos(t)) ** 2


def directxua_model(mc):
    """TestDirectXUA.jl:23-44."""
    model = mc.Model("TrueModel")
    n1 = mc.addnode(model, [0.])
    n2 = mc.addnode(model, [1.])
    n3 = mc.addnode(model, [])
    mc.addelement(model, El1, [n1], K=1., C=0.05, M=1.)
    mc.addelement(model, El1, [n2], K=0., C=0.0, M=1.)
    mc.addelement(model, Spring1, [n1, n2, n3], EA=1.1)
    mc.addelement(model, mc.SingleAcost, [n3], field="ΞL₀", cost=fa)
    mc.addelement(model, mc.SingleAcost, [n3], field="ΞEI", cost=fa)
    mc.addelement(model, mc.SingleAcost, [n1], field="ΞC", cost=fa)
    mc.addelement(model, mc.SingleAcost, [n1], field="ΞM", cost=fa)
    mc.addelement(model, mc.SingleAcost, [n2], field="ΞC", cost=fa)
    mc.addelement(model, mc.SingleAcost, [n2], field="ΞM", cost=fa)
    mc.addelement(model, mc.SingleUdof, [n1], Xfield="tx1", Ufield="utx1", cost=fu)
    mc.addelement(model, mc.SingleUdof, [n2], Xfield="tx1", Ufield="utx1", cost=fu)
    mc.addelement(model, mc.SingleDofCost, [n1], dofclass="X", field="tx1", cost=l1)
    mc.addelement(model, mc.SingleDofCost, [n2], dofclass="X", field="tx1", cost=l2)
    return model, mc.initialize(model)


def oracle_maps(model, dis, OX, OU, IA):
    index = {c: [[list(r) for r in getattr(d, c)] for d in dis.dis] for c in ("X", "U", "A")}
    return om.direct_maps(index, dict(model.ndof), OX, OU, IA)


def cols(matrix_rows):
    """A Julia matrix literal [a b; c d] (rows) as per-element columns."""
    m = np.array(matrix_rows, np.int64)
    return [list(c) for c in m.T]


def test_model_layout(mc):
    model, s0 = directxua_model(mc)
    assert model.ndof == {"X": 2, "U": 4, "A": 6}
    assert len(model.eletypes) == 9


def test_oracle_asm_goldens(mc):
    """TestDirectXUA.jl:109-120."""
    model, s0 = directxua_model(mc)
    asm, pat = oracle_maps(model, s0.dis, 2, 0, 1)
    assert asm[1][0] == cols([[1, 2]])
    assert asm[1][1] == cols([[1], [2]])
    assert asm[2][0] == cols([[1, 2]])
    assert asm[2][1] == cols([[1], [2]])
    assert asm[3][0] == cols([[1, 2]])
    assert asm[3][1] == [[]]
    assert asm[4][0] == cols([[1, 3], [2, 4]])
    assert asm[4][1] == cols([[5], [6]])
    assert asm[5][0] == cols([[1, 4]])
    assert asm[20][0] == cols([[1, 5], [2, 6], [3, 7], [4, 8]])


LVV_COLPTR_50 = [1, 13, 25, 43, 61, 68, 75, 80, 85, 97, 109, 133, 157, 164, 171, 176, 181, 193, 205, 235, 265, 272, 279,
                 284, 289, 301, 313, 343, 373, 380, 387, 392, 397, 409, 421, 445, 469, 476, 483, 488, 493, 505, 517, 535,
                 553, 560, 567, 572, 577, 597]
LVV_ROWVAL_60 = [3, 4, 5, 7, 11, 12, 19, 20, 49, 50, 53, 54, 3, 4, 6, 8, 11, 12, 19, 20, 51, 52, 53, 54, 1, 2, 3, 4, 5, 7, 9,
                 10, 11, 12, 13, 15, 19, 20, 49, 50, 53, 54, 1, 2, 3, 4, 6, 8, 9, 10, 11, 12, 14, 16, 19, 20, 51, 52, 53,
                 54]
LVASM = [1, 3, 5, 9, 11, 13, 17, 19, 21, 25, 27, 29, 33, 35, 37, 41, 43, 45, 49, 55]
LVVASM_COLPTR = [1, 6, 14, 20, 25, 36, 42, 47, 61, 67, 72, 86, 92, 97, 108, 114, 119, 127, 133, 152]


def oracle_big(model, dis, nstep=6, OX=2, OU=0, IA=1):
    asm, pat = oracle_maps(model, dis, OX, OU, IA)
    bcp, brv, blocks = om.makepattern(IA, [nstep], pat)
    return (bcp, brv) + om.sparse_prepare(bcp, brv, blocks, pat)


def test_oracle_preparebig_goldens(mc):
    """TestDirectXUA.jl:122-136: preparebig(IA=1, nstep=6, out) for AssemblyDirect{2,0,1}."""
    model, s0 = directxua_model(mc)
    bcp, brv, colptr, rowval, asm_igv, pgr, pgc = oracle_big(model, s0.dis)
    assert len(colptr) - 1 == 54
    assert colptr[:50] == LVV_COLPTR_50
    assert rowval[:60] == LVV_ROWVAL_60
    assert pgr == LVASM and pgc == LVASM
    assert bcp == LVVASM_COLPTR
    assert asm_igv[0] == [1, 2, 13, 14]
    assert asm_igv[3] == [7, 8, 19, 20]
    assert asm_igv[150] == [595, 596, 615, 616, 635, 636, 655, 656, 681, 682, 707, 708]


def test_oracle_sparsetools_goldens():
    """TestSparseTools.jl:7-62: block assembly into the big matrix, bigsparseasm, sparser!."""
    # pattern 1: a 3×3 block at block positions (1,2), (2,1), (2,2), (3,1), (3,2)
    rng = np.random.default_rng(0)
    blk = sp.csc_matrix((rng.standard_normal(5), ([0, 0, 1, 2, 2], [0, 1, 1, 1, 2])), shape=(3, 3))
    pat = {"B": (list(blk.indptr + 1), list(blk.indices + 1), (3, 3), 1, 1)}
    bcp, brv = [1, 3, 6], [2, 3, 1, 2, 3]
    colptr, rowval, asm_igv, pgr, pgc = om.sparse_prepare(bcp, brv, ["B"] * 5, pat)
    nz = np.zeros(len(rowval))
    for ibv in range(5):
        nz[np.array(asm_igv[ibv]) - 1] += blk.data
    big = sp.csc_matrix((nz, np.array(rowval) - 1, np.array(colptr) - 1), shape=(9, 6)).toarray()
    assert np.array_equal(big[3:6, 0:3], blk.toarray())
    assert np.array_equal(big[3:6, 3:6], blk.toarray())
    # pattern 2: a 4×4 block on the diagonal of a 3×3 pattern
    pat = {"B": ([1, 2, 4, 5, 7], [1, 1, 2, 3, 1, 4], (4, 4), 1, 1)}
    colptr, rowval, asm_igv, pgr, pgc = om.sparse_prepare([1, 2, 3, 4], [1, 2, 3], ["B"] * 3, pat)
    assert asm_igv == [[1, 2, 3, 4, 5, 6], [7, 8, 9, 10, 11, 12], [13, 14, 15, 16, 17, 18]]
    assert pgr == [1, 5, 9, 13] and pgc == [1, 5, 9, 13]
    # sparser!(s, i -> s.nzval[i] > 0.5) on the literal of TestSparseTools.jl:53-55
    i = [3, 7, 2, 3, 6, 2, 7, 9, 2, 6, 4, 5, 9, 3, 9, 1, 7, 8, 10, 4, 9, 7]
    j = [2, 2, 3, 3, 3, 4, 4, 4, 6, 6, 7, 7, 7, 8, 8, 9, 9, 9, 9, 10, 10, 11]
    v = [0., 1., 0., 0., 1., 1., 1., 0., 0., 0., 1., 0., 0., 0., 0., 1., 0., 0., 0., 0., 1., 0.]
    S = sp.csc_matrix((v, (np.array(i) - 1, np.array(j) - 1)), shape=(10, 11))
    cp, rv, nzv = om.sparser(list(S.indptr + 1), list(S.indices + 1), list(S.data), rtol=0.5)
    assert cp == [1, 1, 2, 3, 5, 5, 5, 6, 6, 7, 8, 8]
    assert rv == [7, 6, 2, 7, 4, 1, 9]
    assert nzv == [1.] * 7


def test_product_pattern_equals_oracle(mc):
    """The product's numpy pattern (class_patterns + BigPattern) is the oracle's, bit for bit, on
    the golden model and on models with several experiments, IA = 0, 1 and OX, OU = 0..2."""
    from muscade_amd.directxua import BigPattern, class_patterns
    model, s0 = directxua_model(mc)
    for (OX, OU, IA, nstep) in ((2, 0, 1, [6]), (0, 0, 1, [3]), (1, 1, 0, [7, 6]), (2, 2, 1, [6, 8]), (0, 0, 0, [1])):
        asm, pat = oracle_maps(model, s0.dis, OX, OU, IA)
        bcp, brv, blocks = om.makepattern(IA, nstep, pat)
        colptr, rowval, asm_igv, pgr, pgc = om.sparse_prepare(bcp, brv, blocks, pat)
        P = BigPattern(IA, nstep, class_patterns(model, s0.dis, OX, OU, IA))
        assert np.array_equal(P.colptr + 1, colptr), (OX, OU, IA, nstep)
        assert np.array_equal(P.rowval + 1, rowval), (OX, OU, IA, nstep)
        assert np.array_equal(P.pgr + 1, pgr), (OX, OU, IA, nstep)


def test_product_pattern_beams_equals_oracle(mc):
    """The same on a curved beam cantilever with U-loads and Holds (the load identification's
    kinds), DirectXUA{2,0,0} over 6 steps."""
    from muscade_amd.directxua import BigPattern, class_patterns
    from tests.test_gpu_directxua import beam_model
    model = beam_model(mc, inverse=True)
    s0 = mc.initialize(model, time=0.)
    for (OX, OU, IA) in ((2, 0, 0), (0, 0, 0)):
        asm, pat = oracle_maps(model, s0.dis, OX, OU, IA)
        bcp, brv, blocks = om.makepattern(IA, [6], pat)
        colptr, rowval, _, pgr, _ = om.sparse_prepare(bcp, brv, blocks, pat)
        P = BigPattern(IA, [6], class_patterns(model, s0.dis, OX, OU, IA))
        assert np.array_equal(P.colptr + 1, colptr)
        assert np.array_equal(P.rowval + 1, rowval)


def test_product_sparser_equals_oracle():
    from muscade_amd.directxua import sparser
    rng = np.random.default_rng(4)
    S = sp.random(30, 30, density=0.3, random_state=5, format="csc")
    S.data = np.where(rng.random(S.nnz) < 0.3, S.data * 1e-22, S.data)
    T = sparser(S, 1e-20)
    cp, rv, nz = om.sparser(list(S.indptr + 1), list(S.indices + 1), list(S.data), rtol=1e-20)
    assert np.array_equal(T.indptr + 1, cp) and np.array_equal(T.indices + 1, rv) and np.array_equal(T.data, nz)
