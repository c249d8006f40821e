"""GPU parity: HIP element kernels + device pattern + segmented scatter vs the CPU oracle.

Bar (all FP64): integer maps (colptr, rowval, asm[1], asm[2]) bit-exact;
residual and tangent within 1e-12 normwise relative error
(max|gpu-oracle| / max|oracle|) for relative nodal rotations below 2.5 rad
(the Log map of Rotations.jl:105 is singular at π; both methods lose digits
as that angle is approached).
"""
import numpy as np
import pytest

from tests import _oracle as orc
from tests.helpers import build_lattice, oracle_assemble_sweepx0, oracle_model, relerr

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


def random_beams(mc, n, seed, rotscale=0.3, w=True, bad_angle=2.5):
    """n independent beams (2 nodes each) with random geometry, sections and states."""
    rng = np.random.default_rng(seed)
    model = mc.Model("beams")
    c1 = rng.normal(size=(n, 3))
    c2 = c1 + rng.normal(size=(n, 3)) * 2.
    xyz = np.empty((2 * n, 3))
    xyz[0::2], xyz[1::2] = c1, c2
    mc.addnode(model, xyz)
    o2 = rng.normal(size=(n, 3))
    # keep orient2 away from the element axis
    t = (c2 - c1) / np.linalg.norm(c2 - c1, axis=1)[:, None]
    par = np.abs(np.sum(o2 * t, axis=1)) / np.linalg.norm(o2, axis=1) > 0.99
    o2[par] = np.cross(t[par], [0.3, 0.5, 0.7])
    mat = np.zeros((n, 16))
    mat[:, :4] = rng.uniform(1., 10., size=(n, 4))
    if w:
        mat[:, 6] = rng.uniform(-2., 2., size=n)
    nodes = np.stack([np.arange(1, 2 * n, 2), np.arange(2, 2 * n + 1, 2)], axis=1)
    mc.addelement(model, mc.EulerBeam3D, nodes, mat=mat, orient2=o2)
    state = mc.initialize(model)
    X = rng.normal(size=model.ndof["X"]) * rotscale
    d = state.dis.dis[0]
    # bound the relative rotation of each beam's nodes (see module docstring)
    for e in range(n):
        v1, v2 = X[d.X[e, 3:6] - 1], X[d.X[e, 9:12] - 1]
        if np.linalg.norm(v2 - v1) > bad_angle:
            X[d.X[e, 9:12] - 1] = v1 + (v2 - v1) * (bad_angle / np.linalg.norm(v2 - v1)) * 0.9
    state.X[0][:] = X
    return model, state, xyz, nodes, o2, mat


@pytest.mark.parametrize("rotscale", [0.0, 1e-3, 0.1, 0.6])
def test_beam_element_blocks(mc, rotscale):
    n = 300
    model, state, xyz, nodes, o2, mat = random_beams(mc, n, seed=int(rotscale * 1e4) + 3, rotscale=rotscale)
    out = mc.AssemblySweepX(0, model, state.dis)
    out.assemble(state, "iter")
    R, K = out.element_blocks(0)
    d = state.dis.dis[0]
    worst = 0.
    for e in range(n):
        x = state.X[0][d.X[e] - 1]
        Ro, Ks, _ = orc.beam_diffed_residual(xyz[nodes[e, 0] - 1], xyz[nodes[e, 1] - 1], o2[e], mat[e], [x])
        worst = max(worst, relerr(R[e], Ro), relerr(K[e], Ks[0]))
    assert worst <= TOL, worst


def test_beam_special_states(mc):
    """Exact zero rotations (norm3 reset), equal nodal rotations (Δv = 0), pure translations."""
    model, state, xyz, nodes, o2, mat = random_beams(mc, 60, seed=11, rotscale=0.2)
    d = state.dis.dis[0]
    X = state.X[0]
    for e in range(60):
        if e % 3 == 0:
            X[d.X[e, 3:6] - 1] = 0.
            X[d.X[e, 9:12] - 1] = 0.
        elif e % 3 == 1:
            X[d.X[e, 9:12] - 1] = X[d.X[e, 3:6] - 1]
        else:
            X[d.X[e, 3:6] - 1] = 0.
    out = mc.AssemblySweepX(0, model, state.dis)
    out.assemble(state, "iter")
    R, K = out.element_blocks(0)
    for e in range(60):
        x = X[d.X[e] - 1]
        Ro, Ks, _ = orc.beam_diffed_residual(xyz[nodes[e, 0] - 1], xyz[nodes[e, 1] - 1], o2[e], mat[e], [x])
        assert relerr(R[e], Ro) <= TOL and relerr(K[e], Ks[0]) <= TOL, e


def test_lattice_pattern_and_values(mc):
    model, state = build_lattice(5, 4, 3, seed=21)
    out = mc.AssemblySweepX(0, model, state.dis)
    Lλ, nzval = out.assemble(state, "iter")
    ref = oracle_assemble_sweepx0(model, state)
    assert np.array_equal(out.colptr, ref["colptr"])
    assert np.array_equal(out.rowval, ref["rowval"])
    om = ref["om"]
    for it in range(len(model.eletypes)):
        nele, nX = state.dis.dis[it].X.shape
        a1, a2 = out.asm(it)
        o1, o2 = om.asm(it, nele, nX)
        assert np.array_equal(a1, o1) and np.array_equal(a2, o2), it
    assert relerr(Lλ, ref["rhs"]) <= TOL
    assert relerr(nzval, ref["nzval"]) <= TOL


def test_lattice_larger_and_deterministic(mc):
    model, state = build_lattice(16, 12, 6, seed=33, rot=5e-2, disp=1e-2)
    out = mc.AssemblySweepX(0, model, state.dis)
    Lλ1, nz1 = [a.copy() for a in out.assemble(state, "iter")]
    Lλ2, nz2 = [a.copy() for a in out.assemble(state, "iter")]
    assert np.array_equal(nz1, nz2) and np.array_equal(Lλ1, Lλ2)   # deterministic, no atomics
    K = out.Lλx()
    assert (K != K.T).nnz == 0                                     # bitwise symmetric
    ref = oracle_assemble_sweepx0(model, state)
    assert np.array_equal(out.colptr, ref["colptr"]) and np.array_equal(out.rowval, ref["rowval"])
    assert relerr(Lλ1, ref["rhs"]) <= TOL
    assert relerr(nz1, ref["nzval"]) <= TOL


def test_scaled_dofs(mc):
    """setscale! factors enter as δX = δ(scale.X) and Lλ .* scale.X (SweepX.jl:90-95)."""
    from muscade_amd.workloads import space_frame
    import muscade_amd.model as mm
    model, state = space_frame(3, 3, 2, seed=4)
    model.locked = False
    mm.setscale(model, scale={"X": {"t1": 10., "r2": 0.5, "λt3": 3.}})
    model.locked = False
    state2 = mc.initialize(model)
    state2.X[0][:] = state.X[0]
    out = mc.AssemblySweepX(0, model, state2.dis)
    Lλ, nzval = out.assemble(state2, "iter")
    ref = oracle_assemble_sweepx0(model, state2)
    assert relerr(Lλ, ref["rhs"]) <= TOL and relerr(nzval, ref["nzval"]) <= TOL


def test_cantilever_hold_and_load(mc):
    from muscade_amd.workloads import cantilever
    model = cantilever(nel=20)
    state = mc.initialize(model)
    rng = np.random.default_rng(2)
    state.X[0][:] = rng.normal(size=model.ndof["X"]) * 1e-2
    state.time = 3.0
    out = mc.AssemblySweepX(0, model, state.dis)
    Lλ, nzval = out.assemble(state, "iter")
    ref = oracle_assemble_sweepx0(model, state)
    assert np.array_equal(out.colptr, ref["colptr"]) and np.array_equal(out.rowval, ref["rowval"])
    assert relerr(Lλ, ref["rhs"]) <= TOL and relerr(nzval, ref["nzval"]) <= TOL


def test_nan_reports_first_element(mc):
    """A NaN in an element's inputs raises with that element (Assemble.jl:625-631)."""
    model, state = build_lattice(3, 3, 2, seed=1)
    d = state.dis.dis[0]
    state.X[0][d.X[7, 4] - 1] = np.nan
    out = mc.AssemblySweepX(0, model, state.dis)
    with pytest.raises(mc.MuscadeError) as ei:
        out.assemble(state, "iter")
    first = min(e for e in range(d.X.shape[0]) if np.isnan(state.X[0][d.X[e] - 1]).any())
    assert ei.value.dbg["ieletyp"] == 1 and ei.value.dbg["iele"] == first + 1


def test_sweepx0_cantilever_equilibrium(mc):
    """SweepX{0} with GPU assembly converges to a state the oracle's residual calls equilibrium."""
    from muscade_amd.workloads import cantilever
    model = cantilever(nel=40)
    st0 = mc.initialize(model)
    states = mc.solve(mc.SweepX[0], initialstate=st0, time=np.arange(1., 11.), maxΔx=1e-10, verbose=False,
                      catcherror=False)
    assert all(s is not None for s in states)
    s = states[-1]
    ref = oracle_assemble_sweepx0(model, s)
    F = 5e4
    assert np.abs(ref["rhs"]).max() <= 1e-6 * F
    tip = mc.getdof(s, field="t3", nodID=[41])[0]
    assert 0.3 < tip < 10.   # large deflection under α = PL²/EI ≈ 2.38


def test_corotated_half_turn_is_nan(mc):
    """A beam along -x has a corotated frame at a half turn; Muscade's ∂vₛₘ/∂X is NaN there
    (Rotations.jl:101), so the residual is NaN and assembly raises, as in the oracle."""
    model = mc.Model("minus_x")
    mc.addnode(model, [[1., 0., 0.], [0., 0., 0.]])
    mat = mc.BeamCrossSection(1., 2., 3., 4., 0., 0.)
    mc.addelement(model, mc.EulerBeam3D, [[1, 2]], mat=mat, orient2=[0., 1., 0.])
    state = mc.initialize(model)
    Ro, _, _ = orc.beam_diffed_residual([1., 0., 0.], [0., 0., 0.], [0., 1., 0.], mat.values, [np.zeros(12)])
    assert np.isnan(Ro).all()
    out = mc.AssemblySweepX(0, model, state.dis)
    with pytest.raises(mc.MuscadeError) as ei:
        out.assemble(state, "iter")
    assert ei.value.dbg["iele"] == 1
