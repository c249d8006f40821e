"""DofConstraint (src/BasicElements.jl:349-460) and KKT (src/BasicElements.jl:289-307) on the host,
against the goldens of TestDofConstraints.jl:1-160: residual and its Jacobian for λclass X,
the Lagrangian's gradient and (pseudo-)Hessian for λclass U, in the modes equal, off and
positive, with γ = 1 and γ = 0, at a contact point (g = 0) and a gap point (g = 2.4)."""
import numpy as np
import pytest

import muscade_amd as mc
from muscade_amd.hostcost import Dual2

XCTC, XGAP = np.array([4., -3., 10.]), np.array([4., 3., 10.])


def build(λclass, mode):
    if λclass == "X":
        kw = dict(xinod=(1, 1), xfield=("t1", "t2"), gap=lambda x, t: .3 * x[0] + .4 * x[1])
    else:
        kw = dict(uinod=(1, 1), ufield=("t1", "t2"), gap=lambda x, u, a, t: .3 * u[0] + .4 * u[1])
    key, doflist, p = mc.DofConstraint.construct(None, None, λinod=1, λclass=λclass, λfield="λ", mode=mode, **kw)
    assert doflist == ((1, 1, 1), (λclass,) * 3, ("t1", "t2", "λ"))
    return p


CASES = [  # mode, point, γ, R, R∂
    (mc.equal, XCTC, 1., [-3, -4, 0], [[0, 0, -.3], [0, 0, -.4], [-.3, -.4, 0]]),
    (mc.equal, XGAP, 1., [-3, -4, -2.4], [[0, 0, -.3], [0, 0, -.4], [-.3, -.4, 0]]),
    (mc.off, XCTC, 1., [0, 0, -10], [[0, 0, 0], [0, 0, 0], [0, 0, -1]]),
    (mc.positive, XCTC, 1., [-3, -4, 1], [[0, 0, -.3], [0, 0, -.4], [-3, -4, 0]]),
    (mc.positive, XGAP, 1., [-3, -4, -23], [[0, 0, -.3], [0, 0, -.4], [-3, -4, -2.4]]),
    (mc.positive, XCTC, 0., [-3, -4, 0], [[0, 0, -.3], [0, 0, -.4], [-3, -4, 0]]),
    (mc.positive, XGAP, 0., [-3, -4, -24], [[0, 0, -.3], [0, 0, -.4], [-3, -4, -2.4]]),
]


@pytest.mark.parametrize("mode,X,γ,R,dR", CASES)
def test_X_class_residual(mode, X, γ, R, dR):   # TestDofConstraints.jl:24-86
    p = build("X", mode)
    Rv, K = mc.DofConstraint.residual_blocks(p, X[None, :], 0., γ)
    assert np.allclose(Rv[0], R, atol=1e-12) and np.allclose(K[0], dR, atol=1e-12)


@pytest.mark.parametrize("mode,X,γ,R,dR", CASES)
def test_U_class_lagrangian(mode, X, γ, R, dR):   # TestDofConstraints.jl:88-160
    p = build("U", mode)
    U = [[Dual2.seed(X[k:k + 1], k, 3) for k in range(3)]]
    L = mc.DofConstraint.lagrangian(p, [], [[]], U, [], 0., SP={"γ": γ})
    assert np.allclose(L.g[0], R, atol=1e-12) and np.allclose(L.h[0], dR, atol=1e-12)


def test_X_class_in_directxua_is_residual_path():
    """λclass X in DirectXUA: L = Λ·R gives L1[Λ] = R, L2[Λ,X] = ∂R/∂X, L1[X] = (∂R/∂X)ᵀΛ and no
    (X, X) block, for a nonlinear gap."""
    key, _, p = mc.DofConstraint.construct(None, None, λinod=1, λclass="X", λfield="λ", mode=mc.equal,
                                           xinod=(1, 1), xfield=("t1", "t2"),
                                           gap=lambda x, t: x[0] * x[0] + np.sin(x[1]))
    xv = np.array([0.3, -0.7, 2.])
    lam = np.array([0.5, -1.5, 0.25])
    n = 6
    Λ = [Dual2.seed(lam[k:k + 1], k, n) for k in range(3)]
    X = [[Dual2.seed(xv[k:k + 1], 3 + k, n) for k in range(3)]]
    L = mc.DofConstraint.lagrangian(p, Λ, X, [[]], [], 0.)
    R, K = mc.DofConstraint.residual_blocks(p, xv[None, :], 0.)
    assert np.allclose(L.g[0, :3], R[0]) and np.allclose(L.g[0, 3:], K[0].T @ lam)
    assert np.allclose(L.h[0, :3, 3:], K[0]) and np.allclose(L.h[0, 3:, :3], K[0].T)
    assert np.all(L.h[0, 3:, 3:] == 0.)


def test_prescribed_motion_is_not_a_hold():
    """A DofConstraint{:X} whose gap is not Hold's (the prescribed motion horizMove of
    examples/DynamicBeamAnalysis.jl:114-116, gap = x − (1 − e^min(t,0))·181 − xMotion(t)) is never
    routed to the Hold kernel: it is a host-evaluated group, and its λ-row residual is −gap(x, t)
    at the element's own t. Even a user gap equal to x[1] stays a host group: only Hold's own
    constructor gives the Hold kernel (its gap, BasicElements.jl:484-488)."""
    import math
    from muscade_amd import _lib
    import muscade_amd as mc

    def x_motion(t):
        return 0.5 * math.sin(0.7 * t)

    def horiz_move(x, t):
        return x[0] - ((1.0 - math.exp(min(t, 0.))) * 181.0 + x_motion(t))

    def same_as_hold(x, t):
        return x[0]

    model = mc.Model("prescribed")
    n = mc.addnode(model, [0., 0., 0.])
    mc.addelement(model, mc.DofConstraint, [n], xinod=(1,), xfield=("t1",), λinod=1, λclass="X", λfield="λt1",
                  gap=horiz_move, mode=mc.equal)
    mc.addelement(model, mc.DofConstraint, [n], xinod=(1,), xfield=("t3",), λinod=1, λclass="X", λfield="λt3",
                  gap=same_as_hold, mode=mc.equal)
    mc.addelement(model, mc.Hold, [n], field="t2")
    kinds = [E.cls.kind for E in model.eletypes]
    assert kinds == [_lib.MH_KIND_HOST, _lib.MH_KIND_HOST, _lib.MH_KIND_HOLD]
    E = model.eletypes[0]
    for t in (-3., 0.4, 2.):
        R, K = E.cls.residual_blocks(E.params[0], np.array([[0.25, 3.]]), t)
        assert R[0, 1] == pytest.approx(-horiz_move([0.25], t), rel=1e-15)
        assert R[0, 0] == pytest.approx(-3., rel=1e-15)            # −∂gap/∂x·λ
        assert np.allclose(K[0], [[0., -1.], [-1., 0.]])
