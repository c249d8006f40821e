"""Element results on the device (mh_beam_results, muscade_amd.results.getresult) against the
oracle's requestables (orc_beam_results, pinned in tests/test_oracle_results.py): EulerBeam3D's
ε, κ and Gauss-point x, κgp, fᵢ, mᵢ, fₑ, mₑ with 1, 2 and 3 time levels, per column within
1e-12 of the column's largest value; StrainGaugeOnEulerBeam3D's εₐₓ, κ, ε against the goldens
of TestBeamElementStrainGauge.jl:35-56; request selection."""
import math

import numpy as np
import pytest

from tests import _oracle as orc
from tests.helpers import build_lattice

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


def oracle_raw(model, state, it, ND):
    E = model.eletypes[it]
    xyz = model.coords()
    nodes = E.cat_nodes()
    o2 = np.concatenate([p["orient2"] for p in E.params])
    mat = np.concatenate([p["mat"] for p in E.params])
    idx = state.dis.dis[it].X - 1
    U = state.dis.dis[it].U - 1 if E.cls.Udof else None
    out = np.zeros((E.nele, orc.beam_nres(ND)))
    for e in range(E.nele):
        X = [state.X[d][idx[e]] for d in range(ND)]
        u = state.U[0][U[e]] if U is not None else None
        out[e] = orc.beam_results(xyz[nodes[e, 0], :3], xyz[nodes[e, 1], :3], o2[e], mat[e], X, u)
    return out


def check_columns(a, b, tol=1e-12):
    scale = np.abs(b).max(axis=0)
    err = np.abs(a - b).max(axis=0)
    bad = err > tol * np.maximum(scale, 1e-300)
    assert not bad.any(), (np.nonzero(bad)[0][:10], err[bad][:10], scale[bad][:10])


@pytest.mark.parametrize("ND", [1, 2, 3])
def test_lattice_requestables(mc, ND):
    from muscade_amd.results import ElementResults
    print("lattice ND", ND, flush=True)
    model, st = build_lattice(3, 3, 2, seed=9, disp=1e-2, rot=0.2)
    st = st.with_ders(1, 3, 1)
    rng = np.random.default_rng(ND)
    for d in range(1, 3):
        st.X[d][:] = rng.normal(size=model.ndof["X"]) * 0.3
    er = ElementResults(model, st.dis)
    r = er.beam(st, 0, ND)
    nres = orc.beam_nres(ND)
    raw = np.concatenate([r["ε"], r["κ"].reshape(len(r["ε"]), -1)], axis=1)
    g = r["gp"]
    per = [np.concatenate([g["x"][:, k].reshape(-1, 3 * ND), g["κgp"][:, k].reshape(-1, 3 * ND),
                           g["resultants"]["fᵢ"][:, k:k + 1], g["resultants"]["mᵢ"][:, k],
                           g["resultants"]["fₑ"][:, k], g["resultants"]["mₑ"][:, k]], axis=1) for k in range(4)]
    raw = np.concatenate([raw] + per, axis=1)
    assert raw.shape[1] == nres
    check_columns(raw, oracle_raw(model, st, 0, ND))


def test_dynamic_chain_requestables(mc):
    from muscade_amd.results import ElementResults
    from muscade_amd.workloads import catenary_chains
    model, st = catenary_chains(nchain=3, nmin=20, nmax=60, seed=3, disp=1e-2, rot=1e-2, vel=0.5, acc=0.5)
    r = ElementResults(model, st.dis).beam(st, 0, 3)
    ref = oracle_raw(model, st, 0, 3)
    check_columns(np.concatenate([r["ε"], r["κ"].reshape(len(r["ε"]), -1)], axis=1), ref[:, :12])
    fe = r["gp"]["resultants"]["fₑ"]
    off = 12
    per = 6 * 3 + 10
    for k in range(4):
        check_columns(fe[:, k], ref[:, off + k * per + 18 + 4: off + k * per + 18 + 7])


def test_strain_gauge_goldens(mc):   # TestBeamElementStrainGauge.jl:35-56
    model = mc.Model("gauge")
    n1 = mc.addnode(model, [0., 0., 0.])
    n2 = mc.addnode(model, [4., 0., 0.])
    P = np.array([[0., .5, 0.], [0., 0, .5], [0., -.5, 0.], [0., 0, -.5], [0., .5, 0.]]).T
    D = np.array([[1., 0., 0.], [1., 0., 0.], [1., 0., 0.], [1., 0., 0.], [1 / math.sqrt(2), 0, 1 / math.sqrt(2)]]).T
    mat = mc.BeamCrossSection(EA=10., EI2=3., EI3=3., GJ=4., mu=1., iota1=1.0)
    mc.addelement(model, mc.StrainGaugeOnEulerBeam3D, [n1, n2], P=P, D=D,
                  elementkwargs=dict(mat=mat, orient2=np.array([0., 1., 0.])))
    st = mc.initialize(model, time=0.)
    req = mc.request("εₐₓ", "κ", "ε")
    cases = [([0, 0, 0, 0, .1, 0, 0, 0, 0, 0, -.1, 0], [0, 0, 1], 20, [0, -1, 0, 1, 0], 40),
             ([0, 0, 0, 0, 0, .1, 0, 0, 0, 0, 0, -.1], [0, -1, 0], 20, [1, 0, -1, 0, .5], 40),
             ([0, 0, 0, 0, 0, 0, 0, 0, 0, 1., 0, 0], [.25, 0, 0], 1, [0, 0, 0, 0, 0.0625], 1)]
    idx = st.dis.dis[0].X[0] - 1
    for X, kgold, kf, egold, ef in cases:
        st.X[0][:] = 0.
        st.X[0][idx] = X
        r = mc.getresult(st, req, 0)
        assert abs(r["εₐₓ"][0]) <= 1e-12
        assert np.allclose(r["κ"][0] * kf, kgold, atol=1e-12) and np.allclose(r["ε"][0] * ef, egold, atol=1e-12)
    r = mc.getresult(st, mc.request(eleres=mc.request("ε", gp=mc.request("κgp"))), 0)
    assert set(r) == {"eleres"} and set(r["eleres"]) == {"ε", "gp"} and set(r["eleres"]["gp"]) == {"κgp"}
    assert r["eleres"]["gp"]["κgp"].shape == (1, 4, 1, 3)
    with pytest.raises(mc.MuscadeError):
        mc.getresult(st, mc.request("σ"), 0)
