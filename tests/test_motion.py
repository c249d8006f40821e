"""Prescribed motion (DofMotion, the DofConstraint{:X} with gap x − value(t)) on the CPU: the
product's SweepX{2} loop with the oracle's assembly moves each chain's top along its prescribed
path a·sin(2πt/T), and the oracle's constraint row is −(x − value(t))."""
import numpy as np
import pytest

from tests.helpers import OracleAssembly, oracle_model


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


def test_tops_follow_the_prescribed_path(mc):
    from muscade_amd.workloads import catenary_chains
    model, state = catenary_chains(2, 8, 12, seed=3, disp=0., rot=0., vel=0., acc=0., motion=True)
    state.time = 0.
    E = [E for E in model.eletypes if E.cls is mc.DofMotion][0]
    nodes = E.cat_nodes()[:, 0] + 1
    dofs = [model.getdofID("X", "t1", nodID=[n])[0] for n in nodes]
    lams = [model.getdofID("X", "λt1", nodID=[n])[0] for n in nodes]
    par = [(p["args"]["a"], p["args"]["T"]) for p in E.params]
    out = mc.solve(mc.SweepX[2], initialstate=state, time=0.05 * np.arange(1, 5), maxΔx=1e-5, maxiter=20,
                   verbose=False, catcherror=False, Assembly=OracleAssembly)
    for s in out:
        for d, (a, T) in zip(dofs, par):
            assert s.X[0][d - 1] == pytest.approx(a * np.sin(2 * np.pi * s.time / T), abs=1e-8)
    # the constraint row of the oracle at an arbitrary state
    st = state.copy()
    st.time = 0.9
    st.X[0][:] = np.random.default_rng(1).normal(0., 1e-2, model.ndof["X"])
    om = oracle_model(model, st.dis, t=st.time)
    r, rhs, nz, err = om.assemble(0, 1, np.zeros(6), [st.X[0]], None, None)
    for d, l, (a, T) in zip(dofs, lams, par):
        assert rhs[l - 1] == pytest.approx(-(st.X[0][d - 1] - a * np.sin(2 * np.pi * 0.9 / T)), rel=1e-15)
