"""Prescribed motion (DofMotion: the DofConstraint{:X} with gap x − value(t), run on the device as
a Hold whose gap carries the host-evaluated value) on the catenary chains with their tops moved
horizontally:

- the assembled Lλ and Lλx at a moving state against the oracle's (its Hold with the same
  prescribed value), 1e-12 of the largest element force;
- SweepX{2} with the device chain solve against the same loop with the oracle's assembly and
  the host LU, and the tops on their prescribed paths at every step."""
import numpy as np
import pytest

from tests.helpers import OracleAssembly, oracle_model

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


def chains(mc, **kw):
    from muscade_amd.workloads import catenary_chains
    return catenary_chains(3, 30, 60, seed=7, motion=True, **kw)


def tops(mc, model):
    E = [E for E in model.eletypes if E.cls is mc.DofMotion][0]
    nodes = E.cat_nodes()[:, 0] + 1
    dofs = np.array([model.getdofID("X", "t1", nodID=[n])[0] for n in nodes])
    return dofs, [(p["args"]["a"], p["args"]["T"]) for p in E.params]


def test_motion_assembly_matches_oracle(mc):
    model, state = chains(mc, disp=1e-2, rot=1e-2, vel=0.1, acc=0.1)
    state.time = 1.7
    out = mc.AssemblySweepX(2, model, state.dis)
    out.c = mc.Newmarkβcoefficients(2, 0.05, 0.25, 0.5)
    Lλ, nz = out.assemble(state, "step")
    Re, _ = out.element_blocks(0)
    om = oracle_model(model, state.dis, t=state.time)
    r, rhs, nzo, err = om.assemble(2, 0, out.c.array(), state.X[:3], None, None)
    assert r == 0
    assert np.abs(Lλ - rhs).max() <= 1e-12 * np.abs(Re).max()
    assert np.abs(nz - nzo).max() <= 1e-12 * np.abs(nzo).max()
    # the prescribed rows hold −(x − a·sin(2πt/T))
    dofs, par = tops(mc, model)
    lam = np.array([model.getdofID("X", "λt1", nodID=[n])[0] for n in
                    [E for E in model.eletypes if E.cls is mc.DofMotion][0].cat_nodes()[:, 0] + 1])
    for d, l, (a, T) in zip(dofs, lam, par):
        assert Lλ[l - 1] == pytest.approx(-(state.X[0][d - 1] - a * np.sin(2 * np.pi * 1.7 / T)), rel=1e-14)


def test_motion_run_device_against_oracle_loop(mc):
    model, state = chains(mc, disp=0., rot=0., vel=0., acc=0.)
    state.time = 0.
    times = 0.05 * np.arange(1, 6)
    kw = dict(initialstate=state, time=times, maxΔx=1e-5, maxiter=20, verbose=False, catcherror=False)
    dev = mc.solve(mc.SweepX[2], linsolve="device", **kw)
    ref = mc.solve(mc.SweepX[2], Assembly=OracleAssembly, **kw)
    dofs, par = tops(mc, model)
    for s, r in zip(dev, ref):
        # the chain systems are ill-conditioned (the axial stiffness of long lines): the two
        # factorizations agree to about 1e-8, the accelerations (b₁ = 1/(βΔt²) times Δx) least
        for d, tol in zip(range(3), (1e-6, 1e-5, 1e-4)):
            assert np.abs(s.X[d] - r.X[d]).max() <= tol * np.abs(r.X[d]).max()
        for dof, (a, T) in zip(dofs, par):
            assert s.X[0][dof - 1] == pytest.approx(a * np.sin(2 * np.pi * s.time / T), abs=1e-8)
