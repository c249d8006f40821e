"""Captured HIP-graph assembly (mh_assemble_sweepx_graph) against the eager launches of
mh_assemble_sweepx: bitwise equal rhs and nzval on every replay, with the program's own
eager work and synchronisations between the replays, recaptured when the Newmark coefficients
or, with bars (whose kernel takes the time), the time change (one graph per mission); and
whole device Newton iterations as graphs (mh_newton_iteration_graph): the same SweepX
trajectories bit for bit as the eager iterations, two captures for a run at constant Δt."""
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
    assert g.graph_captures() == 1   # no kernel of these element kinds takes the time


def test_graph_dynamic_beams_and_bars(mc):
    from tests.test_gpu_dynamic import dyn_beams
    from tests.test_gpu_bar import tower
    rng = np.random.default_rng(2)
    model, state, *_ = dyn_beams(mc, 80, seed=91, rotscale=0.2)
    g = check_replays(mc, model, state, 2, ["step", "iter", "iter", "iter"], rng, [0.3, 0.35], scale=1e-3)
    assert g.graph_captures() == 2   # :step and :iter; the beams' kernels do not take the time
    model = tower(mc, nlev=3, w=50.)
    state = mc.initialize(model, nXder=3)
    for k in range(3):
        state.X[k][:] = rng.normal(size=model.ndof["X"]) * 1e-2
    g = check_replays(mc, model, state, 2, ["step", "iter", "iter"], rng, [-6.0, -5.95], scale=1e-3)
    assert g.graph_captures() == 4   # (:step, :iter) × two time values: the bars' kernel takes the time


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


def test_newton_iteration_graph_matches_eager(mc):
    """SweepX{2} of hanging chains with the device solve: whole Newton iterations replayed from
    graphs give the trajectory of the eager iterations (and of graphed assemblies) bit for bit."""
    from muscade_amd.workloads import catenary_chains
    model, state = catenary_chains(4, 30, 60, seed=3, disp=0., rot=0., vel=0., acc=0., motion=True)
    state.time = 0.
    # uneven steps: the Newmark coefficients change from step to step
    steps = np.array([0.05, 0.05, 0.04, 0.06, 0.05, 0.03, 0.05, 0.05])
    kw = dict(initialstate=state, time=np.cumsum(steps), maxΔx=1e-6, maxiter=20, verbose=False,
              catcherror=False, linsolve="device")
    runs = {}
    for g in (False, "assembly", True):
        st = {}
        runs[g] = (mc.solve(mc.SweepX[2], graph=g, stats=st, **kw), st)
    ref = runs[False][0]
    for g in ("assembly", True):
        for sa, sb in zip(ref, runs[g][0]):
            for d in range(3):
                assert np.array_equal(bits(sa.X[d]), bits(sb.X[d])), (g, sa.time, d)
    assert runs[True][1]["iterations"] == runs[False][1]["iterations"]
    # the first and the later iterations of a step, each captured once for the run: the graphs
    # read the Newmark coefficients from device memory
    assert runs[True][1]["graph_captures"] == 2 and runs["assembly"][1]["graph_captures"] == 2


def test_newton_iteration_graph_bar_chain(mc):
    """A sagging bar chain (3-dof blocks) in SweepX{2} with the device solve: whole iterations
    replayed from graphs, the assembly writing the chain blocks directly, give the eager
    trajectory bit for bit. The bars' kernel takes the time, so each step captures its two
    graphs anew."""
    from tests.test_chain_layout import bar_chain
    model = bar_chain(n=11)
    st0 = mc.initialize(model, nXder=3, time=0.)
    kw = dict(initialstate=st0, time=0.02 * np.arange(1, 7), maxΔx=1e-8, maxiter=30, verbose=False,
              catcherror=False, linsolve="device")
    runs = {}
    for g in (False, True):
        st = {}
        runs[g] = (mc.solve(mc.SweepX[2], graph=g, stats=st, **kw), st)
    for sa, sb in zip(runs[False][0], runs[True][0]):
        for d in range(3):
            assert np.array_equal(bits(sa.X[d]), bits(sb.X[d])), (sa.time, d)
    assert np.abs(runs[True][0][-1].X[0]).max() > 1e-4   # the chain moved
    assert runs[True][1]["graph_captures"] == 2 * 6
