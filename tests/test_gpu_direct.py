"""GPU parity of one DirectXUA step, AssemblyDirect{OX,OU,IA} (src/DirectXUA.jl:18-171), against
the oracle's restatement (oracle.cpp assemble_direct): class-pair patterns bit-exact, L1 and
every L2 block within 1e-12 normwise relative error (per block; blocks that are zero in the
oracle must be exactly zero)."""
import numpy as np
import pytest

from tests.helpers import oracle_model
from tests.test_gpu_bar import random_bars
from tests.test_gpu_dynamic import dyn_beams

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


def add_supports(mc, model, nodes, rng):
    """Holds on a few X dofs and a DofLoad, appended to a beam or bar model."""
    for f in ("t1", "t2", "t3"):
        mc.addelement(model, mc.Hold, np.array([[n] for n in nodes[:3]]), field=f)

    def load(t, F=2.):
        return F * np.sin(t)
    mc.addelement(model, mc.DofLoad, np.array([[nodes[-1]]]), field="t2", value=load)


def pollute_device_memory(mc):
    """Fill and free device memory first, so a buffer the library fails to initialise shows up
    (freshly mapped memory is zero; reused memory is not)."""
    from tests.helpers import build_lattice
    m, st = build_lattice(12, 10, 6, seed=3, rot=5e-2, disp=1e-2)
    a = mc.AssemblySweepX(0, m, st.dis)
    a.assemble(st, "iter", fetch=False)
    del a


def compare(mc, model, state, lam, OX, OU, dt=0.05):
    pollute_device_memory(mc)
    out = mc.AssemblyDirect(OX, OU, 0, model, state.dis)
    L1, L2 = out.assemble(state, lam, "matrices", dt)
    # L1[X][1] comes only from Lagrangian-path elements (Hold, DofLoad): zero elsewhere, exactly
    touched = np.zeros(model.ndof["X"], bool)
    for it, E in enumerate(model.eletypes):
        if E.cls.kind in (3, 4):
            touched[state.dis.dis[it].X.ravel() - 1] = True
    assert np.all(L1["X"][0][~touched] == 0.)
    om = oracle_model(model, state.dis, t=state.time)
    om.prepare_direct(OX, OU, 0)
    Us = [state.U[0]] + [np.zeros(model.ndof["U"])] * OU if model.ndof["U"] else None
    r, o1, o2 = om.assemble_direct(1, dt, lam, state.X[:OX + 1], Us, t=state.time)
    assert r == 0, r
    # patterns: (Λ|X,Λ|X), (Λ|X,U), (U,Λ|X), (U,U)
    for w, (a, b) in enumerate(((1, 1), (1, 2), (2, 1), (2, 2))):
        cp, rv, _ = out.patterns[w]
        ocp, orv = om.dpattern[(a, b)]
        assert np.array_equal(cp, ocp) and np.array_equal(rv, orv), (a, b)
    g1, g2 = out.flat
    assert g1.size == o1.size and g2.size == o2.size
    # L1 per class segment, L2 per block
    nX, nU = model.ndof["X"], model.ndof["U"]
    segs = [(0, nX)] + [(nX * (1 + d), nX * (2 + d)) for d in range(OX + 1)] + \
           [(nX * (OX + 2) + nU * d, nX * (OX + 2) + nU * (d + 1)) for d in range(OU + 1)]
    for a, b in segs:
        ref, got = o1[a:b], g1[a:b]
        if np.abs(ref).max(initial=0.) == 0.:
            assert np.abs(got).max(initial=0.) == 0.
        else:
            assert np.abs(got - ref).max() <= TOL * np.abs(ref).max(), (a, b)
    nder = [1, OX + 1, OU + 1]
    nnz = om.dnnz
    off = 0
    for a in range(3):
        for b in range(3):
            if a == 0 and b == 0:
                continue
            n = nnz[a, b]
            for blk in range(nder[a] * nder[b]):
                ref, got = o2[off:off + n], g2[off:off + n]
                if np.abs(ref).max(initial=0.) == 0.:
                    assert np.abs(got).max(initial=0.) == 0., (a, b, blk)
                else:
                    assert np.abs(got - ref).max() <= TOL * np.abs(ref).max(), (a, b, blk)
                off += n
    return L1, L2


@pytest.mark.parametrize("OX", [0, 1, 2])
@pytest.mark.parametrize("udof", [False, True])
def test_direct_beams(mc, OX, udof):
    model, state, *_ = dyn_beams(mc, 60, seed=60 + OX, rotscale=0.3, udof=udof)
    model.locked = False
    add_supports(mc, model, [1, 2, 3, 5], np.random.default_rng(1))
    st = mc.initialize(model, nXder=3)
    for k in range(3):
        st.X[k][:state.X[k].size] = state.X[k]
    if udof:
        st.U[0][:] = state.U[0]
    st.time = 0.7
    lam = np.random.default_rng(OX).normal(size=model.ndof["X"])
    compare(mc, model, st, lam, OX, 0)


@pytest.mark.parametrize("OX", [0, 2])
def test_direct_bars(mc, OX):
    model, state, *_ = random_bars(mc, 80, seed=70 + OX, udof=True)
    model.locked = False
    add_supports(mc, model, [1, 2, 4, 6], np.random.default_rng(2))
    st = mc.initialize(model, nXder=3)
    for k in range(3):
        st.X[k][:state.X[k].size] = state.X[k]
    st.U[0][:] = state.U[0]
    st.time = -6.
    lam = np.random.default_rng(5).normal(size=model.ndof["X"])
    compare(mc, model, st, lam, OX, 0)


def test_direct_blocks_are_the_element_jacobians(mc):
    """L2[Λ,X][1,d] is Σ ∂R/∂X⁽ᵈ⁾·Δt assembled and L2[X,Λ][d,1] is its transpose."""
    model, state, *_ = dyn_beams(mc, 40, seed=81, rotscale=0.3)
    lam = np.random.default_rng(3).normal(size=model.ndof["X"])
    out = mc.AssemblyDirect(2, 0, 0, model, state.dis)
    L1, L2 = out.assemble(state, lam, "matrices", 0.1)
    for d in range(3):
        A, B = L2[("Λ", "X")][0][d], L2[("X", "Λ")][d][0]
        assert abs(A - B.T).max() == 0.
    sw = mc.AssemblySweepX(2, model, state.dis)
    sw.c = mc.Newmarkβcoefficients(2, 0.05, 0.25, 0.5)
    Lλ, _ = sw.assemble(state, "iter")
    K = sw.Lλx()
    comb = (L2[("Λ", "X")][0][0] + sw.c.a1 * L2[("Λ", "X")][0][1] + sw.c.b1 * L2[("Λ", "X")][0][2]) / 0.1
    assert abs(comb - K).max() <= 1e-12 * abs(K).max()
    assert np.abs(L1["Λ"][0] / 0.1 - Lλ).max() <= 1e-12 * np.abs(Lλ).max()
