"""ElementConstraint over StrainGaugeOnEulerBeam3D on the device (src/BasicElements.jl:490-576):
the gauge-requestable derivatives from mh_beam_gauge_derivs in place of the oracle's, against the
oracle-backed runs of tests/test_elementconstraint.py:
- the host Lagrangian blocks (∂L/∂λ, ∂L/∂X, the (U, X) and (X, X) Hessian blocks) within 1e-10
  of the oracle provider's;
- solve(DirectXUA{0,0,0}) of the constrained load identification: the device run recovers the
  load and equals the oracle run (u to 1e-9·F, λ to 1e-6);
- getresult of λ, gap and eleres(ε) at the solution (gap = 0 up to the solver tolerance);
- SweepX{0} on the device: the constraint term contributes nothing."""
import numpy as np
import pytest

from tests.helpers import OracleAssemblyDirect
from tests.test_elementconstraint import (SOLVE_KW, constrained_beam_model, gap_at, gauge_cantilever_constrained,
                                          gauge_target, host_blocks, state_for)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


@pytest.mark.parametrize("exact", [False, True])
def test_host_blocks_device_vs_oracle(mc, exact):
    from muscade_amd.directxua import HostLagrangians
    model = constrained_beam_model(mc.equal, exact=exact)
    st = state_for(model, 1.3, seed=8)
    L1o, Ho = host_blocks(model, st)
    out = mc.AssemblyDirect(0, 0, 0, model, st.dis)
    host = HostLagrangians(model, st.dis, 0, 0, 0, gauge_provider=out.gauge_derivatives)
    L1, L2 = host.step(st)
    for c in ("X", "U"):
        a, b = L1[c][0], L1o[c][0]
        assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max(), c
    nX, nU = model.ndof["X"], model.ndof["U"]
    H = {k: np.zeros(v.shape) for k, v in Ho.items()}
    for (a, i, b, j, r, cc, v) in L2:
        if (a, b) in H and i == 0 and j == 0:
            np.add.at(H[(a, b)], (r, cc), v)
    for k in Ho:
        if np.abs(Ho[k]).max() > 0:
            assert np.abs(H[k] - Ho[k]).max() <= 1e-10 * np.abs(Ho[k]).max(), k
    assert nX == 12 and nU == 1


def test_constrained_identification_device_equals_oracle(mc):
    F = 3e4
    target = gauge_target(F)
    inv = gauge_cantilever_constrained(target=target)
    st0 = mc.initialize(inv, time=0.)
    dev = mc.solve(mc.DirectXUA[0, 0, 0], initialstate=[st0], **SOLVE_KW)
    ora = mc.solve(mc.DirectXUA[0, 0, 0], initialstate=[st0], Assembly=OracleAssemblyDirect, **SOLVE_KW)
    iu = inv.getdofID("U", "ut3")[0] - 1
    il = inv.getdofID("U", "λ")[0] - 1
    sd, so = dev[0][0], ora[0][0]
    assert abs(sd.U[0][iu] - F) <= 1e-6 * F
    assert abs(sd.U[0][iu] - so.U[0][iu]) <= 1e-9 * F
    assert abs(sd.U[0][il] - so.U[0][il]) <= 1e-6 * abs(so.U[0][il])
    r = mc.getresult(sd, mc.request("λ", "gap", eleres=mc.request("ε")), 0)
    assert r["λ"][0] == sd.U[0][il]
    assert abs(r["gap"][0]) <= 1e-6
    assert abs(r["eleres"]["ε"][0, 0] - target) <= 1e-9 * abs(target)


def test_getresult_gap_matches_oracle(mc):
    model = constrained_beam_model(mc.positive)
    st = state_for(model, 0.4, seed=11)
    r = mc.getresult(st, mc.request("λ", "gap"), 0)
    assert r["λ"][0] == 0.4
    g = gap_at(model, st)
    assert abs(r["gap"][0] - g) <= 1e-12 * max(abs(g), 1.)


def test_sweepx_device_ignores_the_constraint_term(mc):
    F = 3e4
    plain = gauge_cantilever_constrained(F=F)
    cons = gauge_cantilever_constrained(F=F, target=1e-4)
    kw = dict(time=[1.], maxΔx=1e-9, verbose=False, catcherror=False)
    sp_ = mc.solve(mc.SweepX[0], initialstate=mc.initialize(plain, time=0.), **kw)[-1]
    sc = mc.solve(mc.SweepX[0], initialstate=mc.initialize(cons, time=0.), **kw)[-1]
    assert np.array_equal(sc.X[0], sp_.X[0])


def test_constrained_identification_steps_batched(mc):
    """Three steps, loads F·t: all steps in one launch with the all-steps system on the device
    (batch_steps=True, big="device") equals the step-by-step run and recovers each load."""
    from tests.helpers import OracleAssembly
    F = 2e4
    times = np.array([1., 2., 3.])
    fwd = gauge_cantilever_constrained(F=F)
    fwd.eletypes[-1].params[0]["value"] = lambda t: F * t
    states = mc.solve(mc.SweepX[0], initialstate=mc.initialize(fwd, time=0.), time=list(times), maxΔx=1e-9,
                      verbose=False, catcherror=False, Assembly=OracleAssembly)
    targets = {}
    for t, s in zip(times, states):
        targets[float(t)] = mc.getresult(s, mc.request("ε"), 0)["ε"][0, 0]
    inv = gauge_cantilever_constrained(target=0.)
    inv.eletypes[0].params[0]["gap"] = lambda eleres, t: (eleres["ε"][0] - targets[float(t)]) * 1e4
    kw = dict(SOLVE_KW, time=[times])
    st0 = mc.initialize(inv, time=0.)
    a = mc.solve(mc.DirectXUA[0, 0, 0], initialstate=[st0], **kw)
    b = mc.solve(mc.DirectXUA[0, 0, 0], initialstate=[st0], batch_steps=True, big="device", **kw)
    iu = inv.getdofID("U", "ut3")[0] - 1
    for k, t in enumerate(times):
        ua, ub = a[0][k].U[0][iu], b[0][k].U[0][iu]
        assert abs(ua - F * t) <= 1e-6 * F * t, (t, ua)
        assert abs(ub - ua) <= 1e-9 * F * t, (t, ua, ub)
