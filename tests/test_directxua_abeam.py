"""DirectXUA{0,0,1} with A-scaled beams (EulerBeam3DA: EI₂, EI₃ × 10^a per half of a cantilever)
identifies the stiffness parameters from deflections, on the CPU with the oracle's step
assembly (tests/helpers.py OracleAssemblyDirect, whose ∂R/∂a is ln 10 · R of the masked section
by the oracle's beam):

- the measured deflections are those of a forward SweepX{0} with the true parameters (the
  oracle's assembly with the scaled sections), and DirectXUA recovers the parameters to
  1e-3 of their size (the weak prior on A biases them slightly);
- ∂R/∂a of the oracle is the central difference of its residual in a (h = 1e-6), 1e-7;
- the A column reaches the solver: without ∂R/∂A the problem has no information on A and the
  identified A stays at the prior's minimum.
"""
import numpy as np
import pytest

import muscade_amd as mc
from tests import _oracle as orc
from tests.helpers import OracleAssembly, OracleAssemblyDirect

A_TRUE = np.array([0.08, -0.05])
TIMES = np.arange(1., 7.)


def abeam_model(nel=6, L=3., measurements=None, prior=1e-2):
    from muscade_amd.workloads import base_section
    model = mc.Model("abeam_identification")
    x = np.linspace(0., L, nel + 1)
    mc.addnode(model, np.stack([x, 0 * x, 0 * x], axis=1))
    mc.addnode(model, np.array([[L / 4, 1., 0.], [3 * L / 4, 1., 0.]]))
    seg = np.where(np.arange(nel) < nel // 2, nel + 2, nel + 3)
    nodes = np.stack([np.arange(1, nel + 1), np.arange(2, nel + 2), seg], axis=1)
    mc.addelement(model, mc.EulerBeam3DA, nodes, mat=base_section(), orient2=np.array([0., 1., 0.]))
    for f in ("t1", "t2", "t3", "r1", "r2", "r3"):
        mc.addelement(model, mc.Hold, [1], field=f)
    mc.addelement(model, mc.DofLoad, [nel + 1], field="t3", value=tip_load)
    if measurements is not None:
        for nod, xt in measurements.items():
            mc.addelement(model, mc.SingleDofCost, [nod], dofclass="X", field="t3",
                          cost=lambda x, t, xt=xt: 1e8 * (x - xt[int(round(t)) - 1]) ** 2)
        for nod in (nel + 2, nel + 3):
            mc.addelement(model, mc.SingleAcost, [nod], field="ΞEI", cost=lambda a: prior * a * a)
    return model


def tip_load(t):
    return 4e3 * t


def truth(nel=6):
    """t3 of the mid node and the tip at each step, SweepX{0} with A = A_TRUE."""
    model = abeam_model(nel)
    st = mc.initialize(model, time=0.)
    st.A[:] = A_TRUE
    out = mc.solve(mc.SweepX[0], initialstate=st, time=TIMES, maxΔx=1e-10, verbose=False, catcherror=False,
                   Assembly=OracleAssembly)
    mid, tip = nel // 2 + 1, nel + 1
    dofs = {n: model.getdofID("X", "t3", nodID=[n])[0] - 1 for n in (mid, tip)}
    return {n: np.array([s.X[0][d] for s in out]) for n, d in dofs.items()}


def identify(meas, Assembly=OracleAssemblyDirect, solve_kw=None, **kw):
    model = abeam_model(measurements=meas, **kw)
    st0 = mc.initialize(model, time=0.)
    skw = dict(maxΔx=1e-10, maxΔλ=1e-6, maxΔa=1e-9, maxiter=40, verbose=False, catcherror=False, Assembly=Assembly)
    skw.update(solve_kw or {})
    out = mc.solve(mc.DirectXUA[0, 0, 1], initialstate=[st0], time=[TIMES], **skw)
    return out[0][0].A.copy(), out


def test_oracle_dRdA_is_the_difference_quotient():
    model = abeam_model()
    st = mc.initialize(model, time=0.)
    rng = np.random.default_rng(3)
    st.X[0][:] = rng.normal(0., 1e-2, st.X[0].size)
    st.A[:] = [0.1, -0.2]
    oa = OracleAssemblyDirect(0, 0, 1, model, st.dis)
    oa._state = st
    D = oa.beam_dRdA(0)
    E, d = model.eletypes[0], st.dis.dis[0]
    xyz, nodes = model.coords(), E.cat_nodes()
    o2 = np.concatenate([p["orient2"] for p in E.params])
    base = np.concatenate([p["mat"] for p in E.params])
    h = 1e-6
    for e in (0, 4):
        R = []
        for s in (1., -1.):
            A = st.A.copy()
            A[d.A[e, 0] - 1] += s * h
            mat = E.cls.scaled_section(base, A[d.A[:, 0] - 1], E.params[0]["amask"])
            r, _, _ = orc.beam_diffed_residual(xyz[nodes[e, 0], :3], xyz[nodes[e, 1], :3], o2[e], mat[e],
                                               [st.X[0][d.X[e] - 1]], np.zeros(3))
            R.append(r)
        fd = (R[0] - R[1]) / (2 * h)
        assert np.abs(D[e] - fd).max() <= 1e-7 * np.abs(fd).max()


@pytest.fixture(scope="module")
def measurements():
    return truth()


def test_directxua_identifies_section_scaling(measurements):
    A, _ = identify(measurements)
    assert np.abs(A - A_TRUE).max() <= 1e-3 * np.abs(A_TRUE).max()


class _NoDRdA(OracleAssemblyDirect):
    def beam_dRdA(self, it):
        return np.zeros((self.model.eletypes[it].nele, 12))


def test_without_dRdA_nothing_is_identified(measurements):
    # Λ (of order 1e6 here) then stagnates at a few 1e-6: the tolerance on it is relaxed
    A, _ = identify(measurements, Assembly=_NoDRdA, solve_kw=dict(maxΔλ=1e-4))
    assert np.abs(A).max() <= 1e-12
