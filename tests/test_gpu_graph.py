"""Captured HIP-graph assembly (mh_assemble_sweepx_graph) against the eager launches of
mh_assemble_sweepx: bitwise equal rhs and nzval on every replay, with the program's own
eager work and synchronisations between the replays, recaptured when the time or the
Newmark coefficients change (one graph per mission)."""
import numpy as np
import pytest

from tests.helpers import build_lattice

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def check_replays(mc, model, state, OX, missions, rng, times, scale=1e-2):
    eager = mc.AssemblySweepX(OX, model, state.dis)
    graph = mc.AssemblySweepX(OX, model, state.dis, graph=True)
    for o in (eager, graph):
        o.c = mc.Newmarkβcoefficients(OX, 0.05, 0.25, 0.5)
    for t in times:
        state.time = t
        for rep, mission in enumerate(missions):
            for k in range(OX + 1):   # new state: the graph reads it through the same buffers
                state.X[k][:] = state.X[k] + rng.normal(size=state.X[k].size) * scale
            Le, Ke = [a.copy() for a in eager.assemble(state, mission)]
            Lg, Kg = [a.copy() for a in graph.assemble(state, mission)]
            assert np.array_equal(bits(Le), bits(Lg)), (t, rep, mission)
            assert np.array_equal(bits(Ke), bits(Kg)), (t, rep, mission)
    return graph


def test_graph_static_lattice(mc):
    model, state = build_lattice(6, 5, 4, seed=14, rot=5e-2, disp=1e-2)
    rng = np.random.default_rng(1)
    g = check_replays(mc, model, state, 0, ["iter"] * 4, rng, [0.0, 1.0])
    assert g.graph_captures() == 2   # one per time value


def test_graph_dynamic_beams_and_bars(mc):
    from tests.test_gpu_dynamic import dyn_beams
    from tests.test_gpu_bar import tower
    rng = np.random.default_rng(2)
    model, state, *_ = dyn_beams(mc, 80, seed=91, rotscale=0.2)
    g = check_replays(mc, model, state, 2, ["step", "iter", "iter", "iter"], rng, [0.3, 0.35], scale=1e-3)
    assert g.graph_captures() == 4   # (:step, :iter) × two time steps
    model = tower(mc, nlev=3, w=50.)
    state = mc.initialize(model, nXder=3)
    for k in range(3):
        state.X[k][:] = rng.normal(size=model.ndof["X"]) * 1e-2
    check_replays(mc, model, state, 2, ["step", "iter", "iter"], rng, [-6.0, -5.95], scale=1e-3)


def test_graph_solver_matches_eager(mc):
    """SweepX{0} Newton loop: graph replays give the eager trajectory bit for bit."""
    from muscade_amd.workloads import cantilever
    model = cantilever(nel=30)
    st0 = mc.initialize(model)
    kw = dict(initialstate=st0, time=np.arange(1., 6.), maxΔx=1e-10, verbose=False, catcherror=False)
    a = mc.solve(mc.SweepX[0], graph=False, **kw)
    b = mc.solve(mc.SweepX[0], graph=True, **kw)
    for sa, sb in zip(a, b):
        assert np.array_equal(bits(sa.X[0]), bits(sb.X[0]))
