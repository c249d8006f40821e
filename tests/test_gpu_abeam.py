"""EulerBeam3DA (section stiffnesses scaled by 10^a, a an A-dof) on the device and the adjoint
gradients of a limit state (mh_group_set_adofs, mh_vjp_lagrangian's ∂L/∂A, mh_adjoint_solve):

- element R and K against the oracle's EulerBeam3D with the scaled section, 1e-12;
- ∂L/∂A against the oracle's ln 10 · Λᵀ R(scaled stiffnesses alone, no loads), 1e-12, and
  against central differences of Λ·R(A) assembled on the device (h = 1e-6), 1e-6;
- the limit-state gradient dg/dA, dg/dU of the lateral top displacement of a tensioned riser at
  equilibrium against Richardson-extrapolated central differences of re-solved equilibria,
  1e-4 (the equilibria are converged to about 1e-9 m); the adjoint solve's ∂L/∂X = −∂g/∂X;
  the batch's gradients against each model's own."""
import numpy as np
import pytest

from tests import _oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mc():
    import muscade_amd
    return muscade_amd


def batch(mc, ne=(20, 31, 45), joined=True):
    from muscade_amd.workloads import design_point_batch, design_point_specs
    specs = design_point_specs(len(ne), 1, 2)
    for s, n in zip(specs, ne):
        s["ne"] = n
    return design_point_batch(specs, joined=joined)


def random_state(model, state, seed):
    rng = np.random.default_rng(seed)
    st = state.copy()
    st.X[0][:] = rng.normal(0., 1e-2, model.ndof["X"])
    rd = np.concatenate([model.getdofID("X", f) for f in ("r1", "r2", "r3")]) - 1
    st.X[0][rd] = rng.normal(0., 0.1, rd.size)
    st.A[:] = rng.normal(0., 0.2, model.ndof["A"])
    return st


def oracle_elements(model, st, it, loads=True, masked_only=False, elements=None):
    E = model.eletypes[it]
    d = st.dis.dis[it]
    xyz = model.coords()
    nodes = E.cat_nodes()
    o2 = np.concatenate([p["orient2"] for p in E.params])
    mat = np.concatenate([p["mat"] for p in E.params])
    amask = E.params[0]["amask"]
    mat = E.cls.scaled_section(mat, st.A[d.A[:, 0] - 1], amask)
    if masked_only:
        mat = mat.copy()
        for k in range(4):
            if not amask >> k & 1:
                mat[:, k] = 0.
        mat[:, 4:] = 0.
    els = np.arange(E.nele) if elements is None else np.asarray(elements)
    R, K = np.zeros((els.size, 12)), np.zeros((els.size, 12, 12))
    for q, e in enumerate(els):
        U = st.U[0][d.U[e] - 1] if (loads and not masked_only) else np.zeros(3)
        r, Ks, _ = orc.beam_diffed_residual(xyz[nodes[e, 0], :3], xyz[nodes[e, 1], :3], o2[e], mat[e],
                                            [st.X[0][d.X[e] - 1]], U)
        R[q], K[q] = r, Ks[0]
    return R, K


def test_scaled_beam_element_blocks(mc):
    [(model, s0, _)] = batch(mc)
    st = random_state(model, s0, 1)
    out = mc.AssemblySweepX(0, model, st.dis)
    out.assemble(st, "iter")
    R, K = out.element_blocks(0)
    Ro, Ko = oracle_elements(model, st, 0)
    assert np.abs(R - Ro).max() <= 1e-12 * np.abs(Ro).max()
    assert np.abs(K - Ko).max() <= 1e-12 * np.abs(Ko).max()
    # a zero A gives the plain beam: the base section is kept apart from the scaled one
    st0 = st.copy()
    st0.A[:] = 0.
    out.assemble(st0, "iter")
    R0, _ = out.element_blocks(0)
    Rb, _ = oracle_elements(model, st0, 0)
    assert np.abs(R0 - Rb).max() <= 1e-12 * np.abs(Rb).max()


def test_scaled_beam_large_group(mc):
    """A group of 1.5M A-scaled beams: the base section copy and the scaling kernel run in
    stream order (a null-stream copy racing the scaling left later section fields at zero
    at this size); sampled elements, the last ones included, against the oracle."""
    from muscade_amd.workloads import design_point_batch
    [(model, s0, _)] = design_point_batch([dict(ne=15000, seed=9 + m) for m in range(100)])
    st = random_state(model, s0, 4)
    out = mc.AssemblySweepX(0, model, st.dis, graph=False)
    out.assemble(st, "iter", fetch=False)
    R, K = out.element_blocks(0)
    n = R.shape[0]
    sample = np.unique(np.concatenate([np.arange(0, n, n // 40), np.arange(n - 20, n)]))
    Ro, Ko = oracle_elements(model, st, 0, elements=sample)
    assert np.abs(R[sample] - Ro).max() <= 1e-12 * np.abs(Ro).max()
    assert np.abs(K[sample] - Ko).max() <= 1e-12 * np.abs(Ko).max()


def test_dLdA_against_oracle_and_differences(mc):
    [(model, s0, _)] = batch(mc)
    st = random_state(model, s0, 2)
    rng = np.random.default_rng(3)
    lam = rng.normal(size=model.ndof["X"])
    out = mc.AssemblySweepX(0, model, st.dis)
    gX, gU, gA = out.vjp(st, lam)
    d = st.dis.dis[0]
    Rm, _ = oracle_elements(model, st, 0, masked_only=True)
    want = np.zeros(model.ndof["A"])
    np.add.at(want, d.A[:, 0] - 1, np.log(10.) * np.sum(lam[d.X - 1] * Rm, axis=1))
    assert np.abs(gA - want).max() <= 1e-12 * np.abs(want).max()
    h = 1e-6
    for a in (0, 7, model.ndof["A"] - 1):
        sp, sm = st.copy(), st.copy()
        sp.A[a] += h
        sm.A[a] -= h
        Lp = lam @ (out.assemble(sp, "iter")[0].copy() / st.dis.scaleX)
        Lm = lam @ (out.assemble(sm, "iter")[0].copy() / st.dis.scaleX)
        assert (Lp - Lm) / (2 * h) == pytest.approx(gA[a], rel=1e-6, abs=1e-9 * np.abs(gA).max())


def equilibrium(mc, model, s0):
    from muscade_amd.workloads import design_point_equilibrium
    return design_point_equilibrium(s0, maxΔx=1e-9)


def richardson(g, h):
    """(4·D(h/2) − D(h))/3 of the central difference D(h) = (g(h) − g(−h))/(2h): error O(h⁴)."""
    def D(k):
        return (g(k) - g(-k)) / (2 * k)
    return (4 * D(h / 2) - D(h)) / 3


def test_limit_state_gradient_against_differences(mc):
    [(model, s0, tip)] = batch(mc)
    st = equilibrium(mc, model, s0)
    lsg = mc.LimitStateGradient(model, st.dis)
    dg = np.zeros(model.ndof["X"])
    dg[tip[1] - 1] = 1.                     # g = lateral top displacement of the second riser
    gA, gU, gX = lsg(st, dg)
    assert np.abs(gX + dg).max() <= 1e-10

    def g_of(field, k):
        def g(h):
            s = s0.copy()
            if field == "A":
                s.A[k] += h
            else:
                s.U[0][k] += h
            return equilibrium(mc, model, s).X[0][tip[1] - 1]
        return g
    for a in (16, 20, 31):                  # the second riser's A-dofs
        assert gA[a] == pytest.approx(richardson(g_of("A", a), 2e-2), rel=1e-4)
    for u in (24 + 0, 24 + 12):             # its lateral (t1) loads of segments 1 and 5
        assert gU[u] == pytest.approx(richardson(g_of("U", u), 2e-2 * abs(s0.U[0][u])), rel=1e-4)
    assert np.abs(gA[:16]).max() == 0. and np.abs(gA[32:]).max() == 0.


def test_limit_state_gradient_batched_equals_per_model(mc):
    [(model, s0, tip)] = batch(mc)
    st = equilibrium(mc, model, s0)
    dg = np.zeros(model.ndof["X"])
    dg[tip - 1] = 1.
    lsg = mc.LimitStateGradient(model, st.dis)
    gA, gU, _ = lsg(st, dg)
    # ∂g/∂X by its nonzeros gives the same right-hand side, so the same gradients bit for bit
    gAs, gUs, gXs = lsg(st, (tip, np.ones(tip.size)), upload=False, check=False)
    assert gXs is None and np.array_equal(gAs, gA) and np.array_equal(gUs, gU)
    for m, (mm, sm0, tm) in enumerate(batch(mc, joined=False)):
        sm = equilibrium(mc, mm, sm0)
        d1 = np.zeros(mm.ndof["X"])
        d1[tm - 1] = 1.
        a1, u1, _ = mc.LimitStateGradient(mm, sm.dis)(sm, d1)
        assert np.abs(gA[16 * m:16 * m + 16] - a1).max() <= 1e-9 * np.abs(a1).max()
        assert np.abs(gU[24 * m:24 * m + 24] - u1).max() <= 1e-9 * np.abs(u1).max()
