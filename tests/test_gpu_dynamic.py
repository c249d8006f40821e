"""GPU parity of the dynamic EulerBeam3D (SweepX{1}, SweepX{2}): the HIP value and
direction kernels against the CPU oracle's nested-dual restatement.

Bar (FP64): R, the assembled tangent K + a₁C + b₁M and the :step rhs
(R - C·a - M·b, src/SweepX.jl:45-59) within 1e-12 normwise relative error
(max|gpu-oracle| / max|oracle|); CSC pattern bit-exact. Domain: relative nodal
rotations below 2.5 rad and corotated-frame angle |Rodrigues⁻¹(rₛₘ)| below 2.8 rad
(the roll-inertia couple is contracted with ∂vₛₘ/∂X, singular at π: both codes
lose digits as the angle approaches π).
"""
import numpy as np
import pytest

from tests import _oracle as orc
from tests.helpers import oracle_model, relerr

pytestmark = pytest.mark.gpu
TOL = 1e-12
DT = 0.05


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


def _rotvec_angle(R):
    c = np.clip((np.trace(R) - 1.) / 2., -1., 1.)
    return np.arccos(c)


def _rodrigues(v):
    th = np.linalg.norm(v)
    S = np.array([[0., -v[2], v[1]], [v[2], 0., -v[0]], [-v[1], v[0], 0.]])
    if th < 1e-14:
        return np.eye(3) + S + 0.5 * S @ S
    return np.eye(3) + np.sin(th) / th * S + (1. - np.cos(th)) / th ** 2 * S @ S


def _rsm_angle(c1, c2, o2, v1, v2):
    rm = np.asarray(orc.beam_construct(c1, c2, o2)["rm"])
    R1, R2 = _rodrigues(v1), _rodrigues(v2)
    M = R2 @ R1.T
    th = _rotvec_angle(M)
    ax = np.array([M[2, 1] - M[1, 2], M[0, 2] - M[2, 0], M[1, 0] - M[0, 1]])
    phi = ax * (th / (2. * np.sin(th))) if th > 1e-12 else 0.5 * ax
    return _rotvec_angle(_rodrigues(0.5 * phi) @ R1 @ rm), np.linalg.norm(phi)


def dyn_beams(mc, n, seed, rotscale=0.3, vscale=0.5, ascale=1.0, udof=False):
    """n independent beams with random geometry, full BeamCrossSection (inertia, added mass,
    linear and quadratic damping, roll inertia, weight) and random X, X′, X″."""
    rng = np.random.default_rng(seed)
    model = mc.Model("dynbeams")
    c1 = rng.normal(size=(n, 3))
    dc = rng.normal(size=(n, 3)) * 2.
    dc[dc[:, 0] < 0.] *= -1.   # beams towards -x have an as-meshed frame near a half turn
    c2 = c1 + dc
    xyz = np.empty((2 * n, 3))
    xyz[0::2], xyz[1::2] = c1, c2
    mc.addnode(model, xyz)
    o2 = rng.normal(size=(n, 3))
    t = (c2 - c1) / np.linalg.norm(c2 - c1, axis=1)[:, None]
    par = np.abs(np.sum(o2 * t, axis=1)) / np.linalg.norm(o2, axis=1) > 0.99
    o2[par] = np.cross(t[par], [0.3, 0.5, 0.7])
    mat = np.zeros((n, 16))
    mat[:, :4] = rng.uniform(1., 10., size=(n, 4))
    mat[:, 4] = rng.uniform(0.5, 2., size=n)      # μ
    mat[:, 5] = rng.uniform(0.1, 1., size=n)      # ι₁
    mat[:, 6] = rng.uniform(-2., 2., size=n)      # w
    mat[:, 7:16] = rng.uniform(0., 1., size=(n, 9))
    nodes = np.stack([np.arange(1, 2 * n, 2), np.arange(2, 2 * n + 1, 2)], axis=1)
    if udof:
        mc.addnode(model, rng.normal(size=(n, 3)))
        nodes = np.concatenate([nodes, np.arange(2 * n + 1, 3 * n + 1)[:, None]], axis=1)
        mc.addelement(model, mc.EulerBeam3DU, nodes, mat=mat, orient2=o2)
    else:
        mc.addelement(model, mc.EulerBeam3D, nodes, mat=mat, orient2=o2)
    state = mc.initialize(model, nXder=3)
    d = state.dis.dis[0]
    X0 = rng.normal(size=model.ndof["X"]) * rotscale
    for e in range(n):
        for it in range(60):   # stay inside the well-conditioned domain (module docstring)
            v1, v2 = X0[d.X[e, 3:6] - 1], X0[d.X[e, 9:12] - 1]
            a_rsm, a_rel = _rsm_angle(c1[e], c2[e], o2[e], v1, v2)
            if a_rsm < 2.8 and a_rel < 2.5:
                break
            if it % 2 == 0:   # the as-meshed frame itself may be near a half turn: re-draw orient2
                o2[e] = np.cross(t[e], rng.normal(size=3))
            else:
                X0[d.X[e, 3:6] - 1] *= 0.7
                X0[d.X[e, 9:12] - 1] *= 0.7
        else:
            raise AssertionError("could not draw a well-conditioned beam")
    state.X[0][:] = X0
    model.eletypes[0].params[0]["orient2"][:] = o2
    state.X[1][:] = rng.normal(size=model.ndof["X"]) * vscale
    state.X[2][:] = rng.normal(size=model.ndof["X"]) * ascale
    if udof:
        state.U[0][:] = rng.normal(size=model.ndof["U"])
    return model, state, xyz, nodes, o2, mat


def _oracle_keff(OX, c, Ks):
    K = Ks[0] + c.a1 * Ks[1]
    if OX == 2:
        K = K + c.b1 * Ks[2]
    return K


@pytest.mark.parametrize("OX", [1, 2])
@pytest.mark.parametrize("rotscale", [0.0, 1e-3, 0.3, 0.8])
def test_dynamic_element_blocks(mc, OX, rotscale):
    """Per element: R and K + a₁C (+ b₁M) against diffed_residual of the oracle."""
    n = 200
    model, state, xyz, nodes, o2, mat = dyn_beams(mc, n, seed=100 + OX + int(rotscale * 1e3), rotscale=rotscale)
    out = mc.AssemblySweepX(OX, model, state.dis)
    out.c = mc.Newmarkβcoefficients(OX, DT, 0.25, 0.5)
    out.assemble(state, "iter")
    R, K = out.element_blocks(0)
    d = state.dis.dis[0]
    worst = 0.
    for e in range(n):
        X = [state.X[k][d.X[e] - 1] for k in range(OX + 1)]
        Ro, Ks, _ = orc.beam_diffed_residual(xyz[nodes[e, 0] - 1], xyz[nodes[e, 1] - 1], o2[e], mat[e], X)
        worst = max(worst, relerr(R[e], Ro), relerr(K[e], _oracle_keff(OX, out.c, Ks)))
    assert worst <= TOL, worst


@pytest.mark.parametrize("OX,mission,udof", [(1, "iter", False), (1, "step", False), (2, "iter", False),
                                             (2, "step", False), (2, "step", True)])
def test_dynamic_assembly_vs_oracle(mc, OX, mission, udof):
    """Assembled Lλ (with the δr term of :step) and Lλx against the oracle's sequential assembler."""
    model, state, *_ = dyn_beams(mc, 150, seed=7 + OX, rotscale=0.4, udof=udof)
    out = mc.AssemblySweepX(OX, model, state.dis)
    out.c = mc.Newmarkβcoefficients(OX, DT, 0.25, 0.5)
    Lλ, nzval = out.assemble(state, mission)
    om = oracle_model(model, state.dis)
    r, rhs, nz, err = om.assemble(OX, 0 if mission == "step" else 1, out.c.array(), state.X[:OX + 1],
                                  state.U[0] if udof else None, None)
    assert r == 0, err
    assert np.array_equal(out.colptr, om.colptr) and np.array_equal(out.rowval, om.rowval)
    assert relerr(Lλ, rhs) <= TOL
    assert relerr(nzval, nz) <= TOL


def test_dynamic_lattice_with_holds(mc):
    """Space frame with a full dynamic section, Holds on the base, OX = 2 :step then :iter."""
    from muscade_amd.workloads import space_frame
    model, st = space_frame(5, 4, 3, seed=12, disp=1e-2, rot=5e-2)
    for p in model.eletypes[0].params:
        m = p["mat"]
        m[:, 4] = 7.8e3 * 1e-2       # μ
        m[:, 5] = 7.8e3 * 2e-5       # ι₁
        m[:, 6] = 500.               # w
        m[:, 7:16] = [50., 10., 5., 80., 20., 10., 80., 20., 10.]
    state = st.copy().with_ders(1, 3, 1)
    rng = np.random.default_rng(3)
    state.X[1][:] = rng.normal(size=model.ndof["X"]) * 0.1
    state.X[2][:] = rng.normal(size=model.ndof["X"]) * 0.5
    out = mc.AssemblySweepX(2, model, state.dis)
    out.c = mc.Newmarkβcoefficients(2, DT, 0.25, 0.5)
    om = oracle_model(model, state.dis)
    for mission in ("step", "iter"):
        Lλ, nzval = out.assemble(state, mission)
        r, rhs, nz, err = om.assemble(2, 0 if mission == "step" else 1, out.c.array(), state.X[:3], None, None)
        assert r == 0, err
        assert relerr(Lλ, rhs) <= TOL and relerr(nzval, nz) <= TOL, mission


def test_dynamic_zero_motion_equals_static(mc):
    """With X′ = X″ = 0 and no loads depending on motion, R equals the static residual and
    K + a₁C + b₁M equals the static tangent plus b₁·M; M against the consistent-mass closed
    forms of test/TestBeamElement.jl:175-239 (unit beam along x)."""
    model = mc.Model("one")
    mc.addnode(model, [[0., 0., 0.], [1., 0., 0.]])
    EA, EI2, EI3, GJ, mu, iota = 10., 3., 4., 2., 2., 0.5
    mat = mc.BeamCrossSection(EA, EI2, EI3, GJ, mu, iota)
    mc.addelement(model, mc.EulerBeam3D, [[1, 2]], mat=mat, orient2=[0., 1., 0.])
    state = mc.initialize(model, nXder=3)
    out = mc.AssemblySweepX(2, model, state.dis)
    c = mc.Newmarkβcoefficients(2, DT, 0.25, 0.5)
    out.c = c
    out.assemble(state, "iter")
    R, K = out.element_blocks(0)
    out0 = mc.AssemblySweepX(0, model, state.dis)
    out0.assemble(state, "iter")
    R0, K0 = out0.element_blocks(0)
    assert np.abs(R[0]).max() == 0. and np.abs(R0[0]).max() == 0.
    M = (K[0] - K0[0]) / c.b1
    L = 1.
    assert abs(M[0, 0] - mu * L / 3) <= 1e-12 * mu
    assert abs(M[1, 1] - 156 * mu * L / 420) <= 1e-12 * mu
    assert abs(M[1, 7] - 54 * mu * L / 420) <= 1e-12 * mu
    assert abs(M[4, 4] - 4 * mu * L ** 3 / 420) <= 1e-12 * mu
    assert abs(M[3, 3] - iota * L / 4) <= 1e-12 * mu


def test_dynamic_nan_reports_element(mc):
    model, state, *_ = dyn_beams(mc, 30, seed=5)
    d = state.dis.dis[0]
    state.X[2][d.X[11, 2] - 1] = np.nan
    out = mc.AssemblySweepX(2, model, state.dis)
    out.c = mc.Newmarkβcoefficients(2, DT, 0.25, 0.5)
    with pytest.raises(mc.MuscadeError) as ei:
        out.assemble(state, "iter")
    first = min(e for e in range(30) if np.isnan(state.X[2][d.X[e] - 1]).any())
    assert ei.value.dbg["iele"] == first + 1


def test_prepared_order_is_checked(mc):
    model, state, *_ = dyn_beams(mc, 4, seed=9)
    out = mc.AssemblySweepX(2, model, state.dis)
    out.OX = 1
    with pytest.raises(mc.MuscadeError):
        out.assemble(state, "iter")


def test_sweepx2_cantilever_matches_oracle_solve(mc):
    """SweepX{2} (Newmark-β) on a cantilever with mass and roll inertia under a ramped tip
    load: the trajectory with GPU assembly equals the one with the oracle's assembler."""
    from muscade_amd.workloads import cantilever
    from tests.helpers import OracleAssembly
    model = cantilever(nel=12, F=2e4, nstep=10)
    for p in model.eletypes[0].params:
        p["mat"][:, 4] = 78.        # μ
        p["mat"][:, 5] = 0.16       # ι₁
        p["mat"][:, 8] = 5.         # Cl₁
        p["mat"][:, 12] = 2.        # Cq₂
    st0 = mc.initialize(model, nXder=3, time=0.)
    time = np.arange(1, 9) * 0.01
    kw = dict(initialstate=st0, time=time, maxΔx=1e-8, verbose=False, catcherror=False)
    gpu = mc.solve(mc.SweepX[2], **kw)
    ref = mc.solve(mc.SweepX[2], Assembly=OracleAssembly, **kw)
    scale = max(np.abs(s.X[0]).max() for s in ref)
    assert scale > 1e-4
    for sg, so in zip(gpu, ref):
        for k in range(3):
            assert np.abs(sg.X[k] - so.X[k]).max() <= 1e-9 * max(np.abs(so.X[k]).max(), 1e-30), k
