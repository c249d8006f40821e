"""Gauge-requestable derivatives on the device (mh_beam_gauge_derivs) and ElementCost over
StrainGaugeOnEulerBeam3D in solve(DirectXUA) on the device, against the oracle
(orc_beam_gauge2, pinned by central differences, and the oracle-backed runs of
tests/test_elementcost.py): values, gradients and Hessians per column within 1e-11 of the
column's largest value; the costed-beam ∇L golden (TestBeamElementStrainGauge.jl:90-111);
the load identification from gauge strains, device run = oracle run."""
import numpy as np
import pytest

from tests import _oracle as orc
from tests.test_elementcost import (GOLD, costed_beam_model, gauge_cantilever, lagrangian_gradient, state_of)
from tests.helpers import OracleAssemblyDirect

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


def test_gauge_derivatives_vs_oracle(mc):
    from muscade_amd.results import ElementResults
    from muscade_amd.workloads import base_section
    rng = np.random.default_rng(5)
    model = mc.Model("gauges")
    n = 40
    xyz = np.cumsum(rng.normal(size=(n + 1, 3)), axis=0)
    mc.addnode(model, xyz)
    P = np.array([[0., .1, 0.], [0., 0., .1]]).T
    D = np.array([[1., 0., 0.], [.6, .8, 0.]]).T
    mc.addelement(model, mc.StrainGaugeOnEulerBeam3D, np.stack([np.arange(1, n + 1), np.arange(2, n + 2)], axis=1),
                  P=P, D=D, elementkwargs=dict(mat=base_section(), orient2=np.array([0.3, 0.2, 1.])))
    st = mc.initialize(model, time=0.)
    st.X[0][:] = rng.normal(size=model.ndof["X"]) * 0.3
    st.X[0][model.getdofID("X", "r1", np.arange(1, 4)) - 1] = 0.   # a few zero rotations
    er = ElementResults(model, st.dis)
    v, g, H = er.gauge_derivatives(0, st)
    E = model.eletypes[0]
    o2 = np.concatenate([p["orient2"] for p in E.params])
    mat = np.concatenate([p["mat"] for p in E.params])
    idx = st.dis.dis[0].X - 1
    vo, go, Ho = np.zeros_like(v), np.zeros_like(g), np.zeros_like(H)
    for e in range(n):
        vo[e], go[e], Ho[e] = orc.beam_gauge2(xyz[e], xyz[e + 1], o2[e], mat[e], st.X[0][idx[e]])
    for a, b in ((v, vo), (g.reshape(n, -1), go.reshape(n, -1)), (H.reshape(n, -1), Ho.reshape(n, -1))):
        # columns that are zero up to round-off (∂ε/∂rotations at this state) get the array's scale
        scale = np.maximum(np.abs(b).max(axis=0), 1e-3 * np.abs(b).max())
        assert np.all(np.abs(a - b).max(axis=0) <= 1e-11 * scale)


def test_costed_beam_gradient_golden_device(mc):
    model = costed_beam_model()
    st, idx = state_of(model)
    gX, _ = lagrangian_gradient(model, st, mc.AssemblyDirect, lambda out: out.gauge_derivatives)
    g = gX[idx]
    assert np.linalg.norm(g - GOLD) <= 1.5e-8 * np.linalg.norm(GOLD)


def test_load_identification_device_equals_oracle(mc):
    from muscade_amd.results import ElementResults
    F = 3e4
    fwd = gauge_cantilever(F=F)
    s = mc.solve(mc.SweepX[0], initialstate=mc.initialize(fwd, time=0.), time=[1.], maxΔx=1e-9, verbose=False,
                 catcherror=False)[-1]
    r = mc.getresult(s, mc.request("ε"), 0)
    target = r["ε"][0]
    inv = gauge_cantilever(target=target)
    kw = dict(initialstate=[mc.initialize(inv, time=0.)], time=[np.array([1.])], maxΔx=1e-9, maxΔλ=1e-3,
              maxΔu=1e-6, maxiter=30, verbose=False, catcherror=False)
    dev = mc.solve(mc.DirectXUA[0, 0, 0], **kw)
    ora = mc.solve(mc.DirectXUA[0, 0, 0], Assembly=OracleAssemblyDirect, **kw)
    ud, uo = dev[0][0].U[0][0], ora[0][0].U[0][0]
    assert abs(ud - F) <= 1e-6 * F and abs(ud - uo) <= 1e-9 * F, (ud, uo, F)
