"""The product's solver loops against the reference's solver goldens (CPU).

The reference's test elements (SdofOscillator, AdjustableSdofOscillator, Spring{D}, El1) are
restated as host residual elements (tests/fixture_elements.py). The SweepX loop runs with the
product's host-element evaluation scattered on the CPU (tests/helpers.HostSweepXAssembly); the GPU
variant of the same tests (tests/test_gpu_solver_goldens.py) runs it with the device assembler.
Comparisons use Julia's isapprox: ‖a − b‖ ≤ rtol·max(‖a‖, ‖b‖), rtol = √eps unless the
reference states one."""
import math

import numpy as np
import pytest

from tests.fixture_elements import AdjustableSdofOscillator, SdofOscillator
from tests.helpers import HostSweepXAssembly

RTOL = math.sqrt(np.finfo(float).eps)


def isapprox(a, b, rtol=RTOL):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.linalg.norm(a - b) <= rtol * max(np.linalg.norm(a), np.linalg.norm(b))


# TestSweepX2.jl:20-24: X, X′, X″ of the SDOF Newmark-β run at steps 1, 11, …, 91
SWEEPX2_X = [0.3653456491624315, 0.039495394592936224, -0.9800856952523974, 0.0204208015740059, 0.10202734361080085,
             -0.08876358375431506, 0.020279004568508258, 0.009410960025011333, -0.01164572216979256,
             0.0044937076208295505]
SWEEPX2_X1 = [0.8267282458121575, -0.5543584424772612, 0.1761431693326722, 0.17142199935793515, -0.07990836783415763,
              0.009615311853025504, 0.01797222270368847, -0.012951778973228531, 0.0032459856194525815,
              0.0015257230558188115]
SWEEPX2_X2 = [-0.8663587709392137, -0.033410494488486674, 0.1512397675677002, -0.08357963580025408,
              -0.012670847567464316, 0.025533223926927903, -0.0130068667401185, 0.0010595839803349055,
              0.0027793256172318247, -0.002010048120257385]
# TestSweepX1.jl:23-26 (rtol 1e-5): x and x′ of the first-order decay at t = 0.1 … 2
SWEEPX1_X = [0.857143, 0.612245, 0.437318, 0.31237, 0.223121, 0.159372, 0.113837, 0.0813124, 0.0580803, 0.0414859,
             0.0296328, 0.0211663, 0.0151188, 0.0107991, 0.00771366, 0.00550976, 0.00393554, 0.0028111, 0.00200793,
             0.00143424]
SWEEPX1_X1 = [-2.85714, -2.04082, -1.45773, -1.04123, -0.743738, -0.531241, -0.379458, -0.271041, -0.193601, -0.138286,
              -0.098776, -0.0705543, -0.0503959, -0.0359971, -0.0257122, -0.0183659, -0.0131185, -0.00937034,
              -0.0066931, -0.00478079]


def sweepx2_sdof(mc, Assembly, **kw):
    """TestSweepX2.jl:8-17."""
    model = mc.Model("TestModel")
    node = mc.addnode(model, [])
    mc.addelement(model, SdofOscillator, [node], K1=1., K2=.3, C1=1., C2=2., M1=3.)
    s0 = mc.initialize(model, time=0.)
    s0 = mc.setdof(s0, [1.], field="tx1", nodID=[node], order=1)
    T = 0.4 * np.arange(1, 101)
    states = mc.solve(mc.SweepX[2], initialstate=s0, time=T, verbose=False, catcherror=False, Assembly=Assembly, **kw)
    return [mc.getdof(states, field="tx1", nodID=[node], order=k)[0] for k in range(3)]


def sweepx1_decay(mc, Assembly, **kw):
    """TestSweepX1.jl:7-21."""
    model = mc.Model("TestModel")
    node = mc.addnode(model, [])
    mc.addelement(model, AdjustableSdofOscillator, [node], K=1., C=.3, M=0.)
    s0 = mc.initialize(model, time=0.)
    s0 = mc.setdof(s0, [1.], field="tx1", nodID=[node], order=0)
    t = np.round(np.arange(1, 21) * 0.1, 12)
    states = mc.solve(mc.SweepX[1], initialstate=s0, time=t, maxiter=5, verbose=False, catcherror=False,
                      Assembly=Assembly, **kw)
    return np.array([s.X[0][0] for s in states]), np.array([s.X[1][0] for s in states])


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


def check_sweepx2(X, X1, X2):
    assert isapprox(X[0::10], SWEEPX2_X)
    assert isapprox(X1[0::10], SWEEPX2_X1)
    assert isapprox(X2[0::10], SWEEPX2_X2)


def check_sweepx1(x, x1):
    assert isapprox(x, SWEEPX1_X, 1e-5)
    assert isapprox(x1, SWEEPX1_X1, 1e-5)


def test_sweepx2_sdof_oscillator_golden(mc):
    check_sweepx2(*sweepx2_sdof(mc, HostSweepXAssembly))


def test_sweepx1_exponential_decay_golden(mc):
    check_sweepx1(*sweepx1_decay(mc, HostSweepXAssembly))



# TestDirectXUA001.jl:33-37 (rtol 1e-4): X and A of the second step of the identification
DXUA001_X = [0.0154897, 0.0154897, 0.0, 0.0, 0.0, 0.0, -0.0100155, 1.55379e-5, 1.55379e-5, -0.0100155]
DXUA001_A = [-0.000195471, -0.0400374]


def load(t):
    return 0.1 * t


def position_meas(x, t):
    return 0.5 * ((x - 0.12 * t) / 0.01) ** 2


def acost(a):
    return 0.5 * (a / .1) ** 2


def directxua001(mc, SweepAssembly, DirectAssembly, times=None, scale=None, Λscale=None, sweep_times=(0.,), **kw):
    """TestDirectXUA001.jl:7-31: two springs sharing their A-dofs, loads, holds, position
    measurements and A priors; SweepX{0} at t = 0, then DirectXUA{0,0,1} over 0:0.1:1.
    scale, Λscale: setscale! before initialize (TestScale.jl:31)."""
    from tests.fixture_elements import Spring2
    model = mc.Model("TestModel")
    n1 = mc.addnode(model, [0., 0.])
    n2 = mc.addnode(model, [10., 0.])
    n3 = mc.addnode(model, [0., 10.])
    n4 = mc.addnode(model, [])
    mc.addelement(model, Spring2, [n1, n2, n4], EA=10.)
    mc.addelement(model, Spring2, [n1, n3, n4], EA=10.)
    mc.addelement(model, mc.DofLoad, [n1], field="tx1", value=load)
    mc.addelement(model, mc.DofLoad, [n1], field="tx2", value=load)
    mc.addelement(model, mc.Hold, [n2], field="tx1")
    mc.addelement(model, mc.Hold, [n2], field="tx2")
    mc.addelement(model, mc.Hold, [n3], field="tx1")
    mc.addelement(model, mc.Hold, [n3], field="tx2")
    mc.addelement(model, mc.SingleDofCost, [n1], dofclass="X", field="tx1", cost=position_meas)
    mc.addelement(model, mc.SingleDofCost, [n1], dofclass="X", field="tx2", cost=position_meas)
    mc.addelement(model, mc.SingleAcost, [n4], field="ΞL₀", cost=acost)
    mc.addelement(model, mc.SingleAcost, [n4], field="ΞEI", cost=acost)
    if scale is not None or Λscale is not None:
        mc.setscale(model, scale=scale, Λscale=Λscale)
    s0 = mc.initialize(model)
    stateX = mc.solve(mc.SweepX[0], initialstate=s0, time=list(sweep_times), verbose=False, catcherror=False,
                      Assembly=SweepAssembly)
    times = times or [np.array([k / 10 for k in range(11)])]
    return mc.solve(mc.DirectXUA[0, 0, 1], initialstate=[stateX[0]] * len(times), time=times, verbose=False,
                    maxiter=50, catcherror=False, Assembly=DirectAssembly, **kw)


def check_directxua001(out):
    st = out[0][1]
    assert isapprox(st.X[0], DXUA001_X, 1e-4)
    assert isapprox(st.A, DXUA001_A, 1e-4)
    assert st.A is out[0][0].A                     # TestDirectXUA001.jl:36: all steps share one A


def test_directxua001_solution_golden(mc):
    from tests.helpers import OracleAssemblyDirect
    check_directxua001(directxua001(mc, HostSweepXAssembly, OracleAssemblyDirect))


def test_scale_invariance(mc):
    """TestScale.jl:28-48: the same model with setscale!(X: tx1 = 1, tx2 = 10, rx3 = 2; A: three
    fields it does not have; Λscale = 2) gives the same SweepX{0} state and the same DirectXUA{0,0,1}
    X and A (≈, Julia's default rtol √eps). The reference runs SweepX over [0, 1] unscaled and
    [0] scaled and compares the first step."""
    from tests.helpers import OracleAssemblyDirect
    kw = dict(maxΔλ=.5, maxΔa=1e-4, maxΔx=1e-4)
    sc = {"X": {"tx1": 1., "tx2": 10., "rx3": 2.}, "A": {"Δseadrag": 3., "Δskydrag": 4., "ΔL": 5.}}
    out1 = directxua001(mc, HostSweepXAssembly, OracleAssemblyDirect, sweep_times=(0., 1.), **kw)
    out2 = directxua001(mc, HostSweepXAssembly, OracleAssemblyDirect, scale=sc, Λscale=2., **kw)
    rtol = np.sqrt(np.finfo(float).eps)
    s1, s2 = out1[0][0], out2[0][0]
    assert isapprox(s1.X[0], s2.X[0], rtol)
    assert isapprox(s1.A, s2.A, rtol)
    d1, d2 = out1[0][0].dis, out2[0][0].dis
    assert np.any(np.asarray(d2.scaleX) == 10.) and not np.any(np.asarray(d1.scaleX) == 10.)   # the scaling took


def test_directxua001_multiple_experiments_share_A(mc):
    """TestDirectXUA001.jl:50-54: two experiments, one A."""
    from tests.helpers import OracleAssemblyDirect
    out = directxua001(mc, HostSweepXAssembly, OracleAssemblyDirect,
                       times=[np.array([k / 10 for k in range(11)]), np.array([k / 10 for k in range(1, 11)])])
    assert out[0][1].A is out[1][2].A


def test_directxua_step_blocks_golden(mc):
    """TestDirectXUA.jl:92-107: the gradient and Hessian blocks of one step (the state at t = 6,
    all dofs zero) of AssemblyDirect{2,0,1} on the model of TestDirectXUA.jl:23-44, from the
    product's host Lagrangians (El1 and Spring{1} by the Lagrangian path, the costs, and the
    A priors, which the reference's step loop reaches too)."""
    from muscade_amd.directxua import HostLagrangians
    from tests.test_oracle_direct_maps import directxua_model
    model, s0 = directxua_model(mc)
    st = s0.with_ders(1, 3, 1)
    st.time, st.SP = 6., {"γ": 0., "iter": 1}
    L1, L2 = HostLagrangians(model, s0.dis, 2, 0, 1).step(st)
    n = {"Λ": 2, "X": 2, "U": 4, "A": 6}

    def block(a, i, b, j):
        M = np.zeros((n[a], n[b]))
        for (aa, ii, bb, jj, r, c, v) in L2:
            if (aa, ii, bb, jj) == (a, i, b, j):
                np.add.at(M, (r, c), v)
        return M

    assert isapprox(L1["Λ"][0], [0., 0.]) or np.abs(L1["Λ"][0]).max() == 0.
    assert isapprox(L1["X"][0], [0.055883099639785175, -0.1920340573300732])
    assert np.abs(L1["X"][1]).max() == 0. and np.abs(L1["X"][2]).max() == 0.
    assert np.abs(L1["U"][0]).max() == 0. and np.abs(L1["A"][0]).max() == 0.
    assert isapprox(block("X", 0, "X", 0), [[2., 0.], [0., 2.]])
    assert np.abs(block("X", 0, "X", 1)).max() == 0. and np.abs(block("X", 0, "X", 2)).max() == 0.
    assert isapprox(block("X", 0, "Λ", 0), [[2.1, -1.1], [-1.1, 1.1]])
    assert isapprox(block("X", 1, "Λ", 0), [[.05, 0.], [0., 0.]])
    assert isapprox(block("X", 2, "Λ", 0), [[1., 0.], [0., 1.]])
    assert isapprox(block("U", 0, "U", 0), np.diag([0., 0., 2., 2.]))
    assert np.abs(block("U", 0, "A", 0)).max() == 0.
    assert isapprox(block("A", 0, "A", 0), np.diag([2e-14] * 6))
    # TestDirectXUA.jl:73-84: assembleA! alone gives zero gradients and the priors' 2e-14 diagonal
    gA, pieces = HostLagrangians(model, s0.dis, 2, 0, 1).once(st)
    assert np.abs(gA).max() == 0.
    M = np.zeros((6, 6))
    for r, c, v in pieces:
        np.add.at(M, (r, c), v)
    assert isapprox(M, np.diag([2e-14] * 6))


def test_directxua_two_experiments_converge(mc):
    """TestDirectXUA.jl:88: DirectXUA{2,0,1} over two experiments (0:1:5 and 6:1:12) converges
    (the reference keeps no golden for this solution)."""
    from tests.helpers import OracleAssemblyDirect
    from tests.test_oracle_direct_maps import directxua_model
    model, s0 = directxua_model(mc)
    out = mc.solve(mc.DirectXUA[2, 0, 1], initialstate=[s0, s0], time=[np.arange(0., 6.), np.arange(6., 13.)],
                   maxiter=100, verbose=False, catcherror=False, Assembly=OracleAssemblyDirect)
    assert len(out) == 2 and len(out[0]) == 6 and len(out[1]) == 7
    assert np.all(np.isfinite(out[0][0].A))
