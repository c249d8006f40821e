"""Element types of the assembly path.

Each class mirrors one Muscade element: its `doflist` (src/ElementAPI.jl:79-99),
its constructor's validation, and the identity of its concrete type (which
decides the element-type grouping of ModelDescription.jl:203-208). The
residuals themselves are evaluated by the HIP kernels; `kind` names the kernel.
"""
import numpy as np

from . import _lib
from .hostcost import KKT, S, Dual2, call_cost, compose
from .model import MuscadeError

BEAM_FIELDS = ("EA", "EI₂", "EI₃", "GJ", "μ", "ι₁", "w", "Ca₁", "Cl₁", "Cq₁", "Ca₂", "Cl₂", "Cq₂", "Ca₃", "Cl₃", "Cq₃")


class BeamCrossSection:
    """BeamCrossSection(;EA,EI₂,EI₃,GJ,μ,ι₁,w=0,Ca₁..Cq₃=0) — toolbox/BeamElement.jl:6-25."""

    def __init__(self, EA, EI2, EI3, GJ, mu, iota1, w=0., Ca1=0., Cl1=0., Cq1=0., Ca2=0., Cl2=0., Cq2=0., Ca3=0.,
                 Cl3=0., Cq3=0.):
        self.values = np.array([EA, EI2, EI3, GJ, mu, iota1, w, Ca1, Cl1, Cq1, Ca2, Cl2, Cq2, Ca3, Cl3, Cq3], float)

    def __repr__(self):
        return "BeamCrossSection(" + ", ".join(f"{n}={v:g}" for n, v in zip(BEAM_FIELDS, self.values)) + ")"


def _per_element(value, nele, width, name):
    a = np.asarray(value, dtype=np.float64)
    if a.ndim == 1:
        if a.size != width:
            raise MuscadeError(f"{name} must have {width} entries")
        return np.broadcast_to(a, (nele, width)).copy()
    if a.shape != (nele, width):
        raise MuscadeError(f"{name} must be ({nele},{width}) for per-element values")
    return a.copy()


def check_orient(c1, c2, orient2):
    """The constructor's orient2 test, vectorized (BeamElement.jl:126-135)."""
    tg = c2 - c1
    L = np.linalg.norm(tg, axis=1)
    t = tg / L[:, None]
    o = orient2 / np.linalg.norm(orient2, axis=1)[:, None]
    n = o - t * np.sum(o * t, axis=1)[:, None]
    nn = np.linalg.norm(n, axis=1)
    bad = ~(nn > 1e-3)
    if bad.any():
        raise MuscadeError("Provide a 'orient' input that is not nearly parallel to the element")
    return L


class EulerBeam3D:
    """EulerBeam3D{Udof}(nod; mat, orient2=SVector(0,1,0)) — toolbox/BeamElement.jl:87-148.

    `mat` is a BeamCrossSection or an (nele,16) array of per-element sections;
    `orient2` a 3-vector or an (nele,3) array."""
    Udof = False
    kind = _lib.MH_KIND_BEAM

    @classmethod
    def doflist(cls):   # BeamElement.jl:109-118
        inod = (1,) * 6 + (2,) * 6
        field = ("t1", "t2", "t3", "r1", "r2", "r3") * 2
        dclass = ("X",) * 12
        if cls.Udof:
            inod += (3, 3, 3)
            field += ("t1", "t2", "t3")
            dclass += ("U",) * 3
        return inod, dclass, field

    @classmethod
    def construct(cls, model, nodes0, mat, orient2=(0., 1., 0.)):
        nele = nodes0.shape[0]
        m = mat.values if isinstance(mat, BeamCrossSection) else mat
        m = _per_element(m, nele, 16, "mat")
        o = _per_element(orient2, nele, 3, "orient2")
        xyz = model.coords()
        if nele:
            check_orient(xyz[nodes0[:, 0]], xyz[nodes0[:, 1]], o)
        return ("EulerBeam3D", "BeamCrossSection", cls.Udof), cls.doflist(), {"mat": m, "orient2": o}


class EulerBeam3DU(EulerBeam3D):
    """EulerBeam3D{true}: with 3 U-dofs on a third node, a distributed load per length (BeamElement.jl:114-118,167)."""
    Udof = True
    kind = _lib.MH_KIND_BEAM_U


STIFFNESS_FIELDS = ("EA", "EI₂", "EI₃", "GJ")


class EulerBeam3DA(EulerBeam3D):
    """An EulerBeam3D whose section stiffnesses `scaled` (a subset of EA, EI₂, EI₃, GJ) are
    multiplied by 10^a, a being one A-dof of field `afield` at the element's last node: the
    parameter-identification wrapper in the style of the reference's Spring, which scales its L₀
    and EA by exp10 of its A-dofs (test/SomeElements.jl:166-186). Elements of a segment share
    one A node, hence one parameter. The residual is EulerBeam3D's with the scaled section
    (toolbox/BeamElement.jl:151-174); R is linear in each stiffness, so
    ∂R/∂a = ln 10 · (R of the scaled stiffnesses alone, without loads). The device evaluates it
    like any beam (mh_group_set_adofs)."""
    Udof = False
    kind = _lib.MH_KIND_BEAM

    @classmethod
    def doflist(cls, afield="ΞEI"):
        inod, dclass, field = EulerBeam3D.doflist.__func__(cls)
        return inod + (max(inod) + 1,), dclass + ("A",), field + (afield,)

    @classmethod
    def construct(cls, model, nodes0, mat, orient2=(0., 1., 0.), scaled=("EI₂", "EI₃"), afield="ΞEI"):
        scaled = tuple(scaled)
        if not scaled or any(f not in STIFFNESS_FIELDS for f in scaled):
            raise MuscadeError(f"scaled must name section stiffnesses among {STIFFNESS_FIELDS}")
        key, _, params = EulerBeam3D.construct.__func__(cls, model, nodes0, mat, orient2)
        params["amask"] = sum(1 << STIFFNESS_FIELDS.index(f) for f in scaled)
        return key + ("A", params["amask"], afield), cls.doflist(afield), params

    @staticmethod
    def scaled_section(mat, a, amask):
        """The section with its scaled stiffnesses × 10^a (mat (nele,16), a (nele,))."""
        m = np.array(mat, dtype=float, copy=True)
        for k in range(4):
            if amask >> k & 1:
                m[:, k] = m[:, k] * 10. ** np.asarray(a, float)
        return m


class EulerBeam3DAU(EulerBeam3DA):
    """EulerBeam3DA with 3 U-dofs on a third node, a distributed load per length (EulerBeam3D{true});
    the A-dof is on the fourth node."""
    Udof = True
    kind = _lib.MH_KIND_BEAM_U


class StrainGaugeOnEulerBeam3D(EulerBeam3D):
    """StrainGaugeOnEulerBeam3D(nod; P, D, elementkwargs=(mat, orient2)) —
    toolbox/StrainGaugeOnBeamElement.jl:55-76. An EulerBeam3D (same dofs, same residual: the
    device assembles it as a beam) with strain gauges halfway along it. P (3×Ngauge) are the
    gauges' offsets from the axis (P[0,:] = 0), D (3×Ngauge) their directions. Its
    requestables (muscade_amd.results.getresult): εₐₓ, κ and the gauge strains
    ε = E·εₐₓ + K1·κ₁ + K2·κ₂ + K3·κ₃, with E, K1, K2, K3 from P and D as the reference's
    constructor computes them."""

    @classmethod
    def construct(cls, model, nodes0, P, D, elementkwargs):
        P = np.asarray(P, dtype=np.float64)
        D = np.asarray(D, dtype=np.float64)
        if P.ndim != 2 or P.shape[0] != 3 or D.shape != P.shape:
            raise MuscadeError("P and D must be 3×Ngauge")
        if not np.all(P[0] == 0.):
            raise MuscadeError("In arguments of StrainGaugeOnEulerBeam3D, P[1,:] must all be zero")
        key, doflist, params = EulerBeam3D.construct.__func__(cls, model, nodes0, **elementkwargs)
        E = D[0] ** 2
        K1 = D[0] * (D[2] * P[1] - D[1] * P[2])
        K2 = -D[0] ** 2 * P[1]
        K3 = -D[0] ** 2 * P[2]
        nele = nodes0.shape[0]
        params["gauges"] = np.broadcast_to(np.stack([E, K1, K2, K3], axis=1), (nele, P.shape[1], 4)).copy()
        params["P"], params["D"] = P, D
        return ("StrainGaugeOnEulerBeam3D", P.shape[1]) + key, doflist, params


class ElementCost(EulerBeam3D):
    """ElementCost(nod; req, cost, costargs=(), ElementType, elementkwargs) — src/BasicElements.jl:88-132,
    for ElementType = StrainGaugeOnEulerBeam3D: L = Λ·R(beam) + cost(eleres, t, costargs...),
    with eleres the requested gauge requestables (εₐₓ, κ, ε). The wrapped beam's residual is
    assembled on the device like any beam; the cost is evaluated on the host from the device's
    values, gradients and Hessians of εₐₓ and κ (mh_beam_gauge_derivs), composed with the
    second-order duals of the solver's partials. In SweepX the cost contributes nothing (it
    has no Λ).

    DirectXUA evaluates ElementCost with its accelerator (src/DirectXUA.jl:180-199): the
    wrapped element's requestables are revariated to first order only, so the cost's Hessian
    is Jᵀ·∂²cost·J (Gauss-Newton, without Σ ∂cost/∂ε·∂²ε/∂X²). That is the default here, so
    that the Newton iterates are the reference's; exact=True adds the second derivatives of
    the requestables (exact Newton; the same optimum, and the gradient is the same)."""
    host_cost = True

    @classmethod
    def construct(cls, model, nodes0, req, cost, ElementType, elementkwargs, costargs=(), exact=False):
        if ElementType is not StrainGaugeOnEulerBeam3D:
            raise MuscadeError("ElementCost is built for StrainGaugeOnEulerBeam3D elements (the device requestables)")
        key, doflist, params = ElementType.construct(model, nodes0, **elementkwargs)
        name = getattr(cost, "__name__", type(cost).__name__)
        params.update(req=dict(req), cost=cost, costargs=costargs, exact=bool(exact))
        return ("ElementCost", name, bool(exact)) + key, doflist, params

    @staticmethod
    def eleres(p, X):
        """The requested gauge requestables as second-order duals over the solver's partials;
        p["_gd"] = (values (m,4), gradients (m,4,12), Hessians (m,4,12,12)) of (εₐₓ, κ₁, κ₂, κ₃)
        w.r.t. the element's X-dofs at this state (the Hessians dropped unless p["exact"])."""
        v, g, H = p["_gd"]
        exact = p.get("exact", False)
        q = [compose(v[:, k], g[:, k], H[:, k] if exact else None, X[0][:12]) for k in range(4)]
        G = p["gauges"]                                   # m, ngauge, 4
        eps = [((G[:, k, 0] * q[0] + G[:, k, 1] * q[1]) + G[:, k, 2] * q[2]) + G[:, k, 3] * q[3]
               for k in range(G.shape[1])]
        full = {"εₐₓ": q[0], "κ": q[1:], "ε": eps}
        out = {}
        for name in p["req"]:
            if name not in full:
                raise MuscadeError(f"{name} is not a requestable of StrainGaugeOnEulerBeam3D")
            out[name] = full[name]
        return out

    @staticmethod
    def term(p, eleres, t):
        """cost(eleres, t, costargs...) (BasicElements.jl:128-129)."""
        return call_cost(p["cost"], (eleres, t), p["costargs"])

    @staticmethod
    def lagrangian(p, Λ, X, U, A, t):
        """The cost part (the wrapped beam's Λ·R is assembled on the device)."""
        return ElementCost.term(p, ElementCost.eleres(p, X), t)


class ElementConstraint(EulerBeam3D):
    """ElementConstraint(nod; λinod, λfield, req, gap, gargs=(), mode, ElementType, elementkwargs)
    — src/BasicElements.jl:490-576, for ElementType = StrainGaugeOnEulerBeam3D. A constraint
    gap(eleres, t, gargs...) ≥ 0 (or = 0) on the requested gauge requestables, with a Lagrange
    multiplier λ of class U at element node λinod, field λfield, appended to the beam's dofs:
    L = Λ·R(beam) − gap·λ (mode equal) | −KKT(λ, gap, γ) (positive) | −½λ² (off), γ from SP.
    As for ElementCost, the beam is assembled on the device and the constraint term on the host
    from the device's gauge values, gradients and Hessians. In SweepX the term contributes
    nothing (it has no Λ; λ is a U-dof).

    DirectXUA's accelerator for this element (src/DirectXUA.jl:204-241) revariates the
    requestables to first order, like ElementCost's, but also takes λ = ∂0(U)[Nu] undifferentiated,
    so that ∂L/∂λ = −gap, the constraint equation itself, is lost and the λ rows of the all-steps
    Hessian are zero. Here λ is differentiated as in the element's own lagrangian; the
    requestables are first order by default (the Hessian misses −λ·∂²gap/∂X², the same
    solution), and exact=True adds their second derivatives."""
    host_cost = True
    uses_SP = True

    @classmethod
    def construct(cls, model, nodes0, λinod, λfield, req, gap, mode, ElementType, elementkwargs, gargs=(),
                  exact=False):
        if ElementType is not StrainGaugeOnEulerBeam3D:
            raise MuscadeError("ElementConstraint is built for StrainGaugeOnEulerBeam3D elements (the device "
                               "requestables)")
        λinod = int(λinod)
        if not 1 <= λinod <= max(2, nodes0.shape[1]):
            raise MuscadeError("λinod must be an element node number")
        key, (inod, dclass, field), params = ElementType.construct(model, nodes0, **elementkwargs)
        name = getattr(gap, "__name__", type(gap).__name__)
        params.update(req=dict(req), gap=gap, gargs=tuple(gargs), mode=mode, exact=bool(exact))
        doflist = (inod + (λinod,), dclass + ("U",), field + (λfield,))
        key = ("ElementConstraint", name, getattr(mode, "__name__", str(mode)), λinod, λfield, bool(exact)) + key
        return key, doflist, params

    @staticmethod
    def term(p, eleres, lam, t, γ=0.):
        """−gap·λ | −KKT(λ, gap, γ) | −½λ² with gap = gap(eleres, t, gargs...)
        (BasicElements.jl:569-575)."""
        mode = p["mode"](t)
        if mode == "off":
            return -0.5 * (lam * lam)
        return DofConstraint._mode_lagrangian(mode, lam, call_cost(p["gap"], (eleres, t), p["gargs"]), γ)

    @staticmethod
    def lagrangian(p, Λ, X, U, A, t, SP=None):
        γ = float((SP or {}).get("γ", 0.))
        return ElementConstraint.term(p, ElementCost.eleres(p, X), U[0][-1], t, γ)


BAR_FIELDS = ("EA", "μ", "w", "Caₜ", "Clₜ", "Cqₜ", "Caₙ", "Clₙ", "Cqₙ")


class AxisymmetricBarCrossSection:
    """AxisymmetricBarCrossSection(;EA,μ,w=0,Caₜ=0,Clₜ=0,Cqₜ=0,Caₙ=0,Clₙ=0,Cqₙ=0) — toolbox/BarElement.jl:36-48."""

    def __init__(self, EA, mu, w=0., Cat=0., Clt=0., Cqt=0., Can=0., Cln=0., Cqn=0.):
        self.values = np.array([EA, mu, w, Cat, Clt, Cqt, Can, Cln, Cqn], float)

    def __repr__(self):
        return "AxisymmetricBarCrossSection(" + ", ".join(f"{n}={v:g}" for n, v in zip(BAR_FIELDS, self.values)) + ")"


class Bar3D:
    """Bar3D{Udof}(nod; mat, ϵₛ=eps()) — toolbox/BarElement.jl:89-121.

    `mat` is an AxisymmetricBarCrossSection or an (nele,9) array; `eps_s` (ϵₛ) makes the
    stress-free length (1-ϵₛ)·L₀, a scalar or one value per element."""
    Udof = False
    kind = _lib.MH_KIND_BAR

    @classmethod
    def doflist(cls):   # BarElement.jl:106-113
        inod = (1, 1, 1, 2, 2, 2)
        field = ("t1", "t2", "t3") * 2
        dclass = ("X",) * 6
        if cls.Udof:
            inod += (3, 3, 3)
            field += ("t1", "t2", "t3")
            dclass += ("U",) * 3
        return inod, dclass, field

    @classmethod
    def construct(cls, model, nodes0, mat, eps_s=float(np.finfo(float).eps)):
        nele = nodes0.shape[0]
        m = mat.values if isinstance(mat, AxisymmetricBarCrossSection) else mat
        m = _per_element(m, nele, 9, "mat")
        e = np.broadcast_to(np.asarray(eps_s, float), (nele,)).copy()
        return ("Bar3D", "AxisymmetricBarCrossSection", cls.Udof), cls.doflist(), {"mat": m, "eps_s": e}


class Bar3DU(Bar3D):
    """Bar3D{true}: with 3 U-dofs on a third node, a distributed load per length (BarElement.jl:110-113,192)."""
    Udof = True
    kind = _lib.MH_KIND_BAR_U


class Hold:
    """Hold(nod; field, λfield=Symbol(:λ,field)) — src/BasicElements.jl:484-488.
    A DofConstraint{:X} with gap(x)=x[1] and mode `equal` (BasicElements.jl:426-438)."""
    kind = _lib.MH_KIND_HOLD
    direct_host = True   # in DirectXUA: L = Λ∘₁R on the host (getlagrangian, src/Assemble.jl:707-736)

    @staticmethod
    def lagrangian(p, Λ, X, U, A, t):
        return Λ[0] * (-X[0][1]) + Λ[1] * (-X[0][0])

    @classmethod
    def construct(cls, model, nodes0, field, λfield=None):
        λf = λfield if λfield is not None else "λ" + field
        doflist = ((1, 1), ("X", "X"), (field, λf))
        return ("DofConstraint", "X", (field,), λf, "equal"), doflist, {}


class DofMotion(Hold):
    """DofMotion(nod; field, value, λfield=Symbol(:λ,field), args...): a prescribed motion of one
    X-dof, the DofConstraint{:X} with gap(x, t) = x[1] − value(t, args...) and mode equal
    (src/BasicElements.jl:406-438; examples/DynamicBeamAnalysis.jl:114-117 prescribes the top of
    a riser this way, with horizMove and vertMove). R = (−λ, −(x − value(t))), K = [[0,−1],[−1,0]]:
    the device runs it as a Hold whose gap carries the value, which the host evaluates at each
    time like DofLoad's (the Julia binding keeps a generic DofConstraint on the host kind)."""
    kind = _lib.MH_KIND_HOLD
    direct_host = True

    @staticmethod
    def lagrangian(p, Λ, X, U, A, t):
        return Λ[0] * (-X[0][1]) + Λ[1] * (-(X[0][0] - float(p["value"](t, **p["args"]))))

    @staticmethod
    def values(E, t):
        """value(t, args...) of every element of type E, in element order."""
        return np.concatenate([np.full(n.shape[0], float(p["value"](t, **p["args"])))
                               for n, p in zip(E.nodes, E.params)])

    @classmethod
    def construct(cls, model, nodes0, field, value, λfield=None, **args):
        λf = λfield if λfield is not None else "λ" + field
        name = getattr(value, "__name__", type(value).__name__)
        doflist = ((1, 1), ("X", "X"), (field, λf))
        return ("DofConstraint", "X", (field,), λf, "equal", "motion", name, tuple(sorted(args))), doflist, \
            {"value": value, "args": args}


def equal(t):
    """mode=equal (src/BasicElements.jl:322)."""
    return "equal"


def positive(t):
    """mode=positive (src/BasicElements.jl:333)."""
    return "positive"


def off(t):
    """mode=off (src/BasicElements.jl:311)."""
    return "off"


class DofConstraint:
    """DofConstraint(nod; xinod, xfield, uinod, ufield, ainod, afield, λinod, λclass, λfield, gap, gargs, mode)
    — src/BasicElements.jl:349-460. Gap g ≥ 0, multiplier λ ≥ 0; mode(t) ∈ equal, positive, off.
      λclass "X": residual R = (−∂g/∂x·λ, −g | −S(λ,g,γ) | −λ for off), gap(x, t, gargs...);
      λclass "U": L = −g·λ | −KKT(λ,g,γ) | −½λ², gap(x, u, a, t, gargs...), λ a U-dof;
      λclass "A": the same over A-dofs only, gap(a, gargs...), at every step (like any element
      with a Lagrangian; assembleA! takes Acost elements only).
    The gap is a user function, so the element is evaluated on the host: in SweepX as a
    host-evaluated group (R and ∂R/∂X handed to the device scatter, MH_KIND_HOST), in DirectXUA
    as a host Lagrangian (the X class by the residual path: R, ∂R/∂X, no X-X block, like the
    device kinds). γ comes from state.SP (default 0)."""
    kind = _lib.MH_KIND_HOST
    host = True
    uses_SP = True

    @classmethod
    def construct(cls, model, nodes0, λinod, λclass, λfield, gap, mode, xinod=(), xfield=(), uinod=(), ufield=(),
                  ainod=(), afield=(), gargs=()):
        xinod, xfield, uinod, ufield = tuple(xinod), tuple(xfield), tuple(uinod), tuple(ufield)
        ainod, afield = tuple(ainod), tuple(afield)
        Nx, Nu, Na = len(xinod), len(uinod), len(ainod)
        if λclass not in ("X", "U", "A"):
            raise MuscadeError("λclass must be :X, :U or :A")
        if λclass == "X" and (Nu > 0 or Na > 0):
            raise MuscadeError("Constraints with λclass=:X must have zero U-dofs and zero A-dofs")
        if λclass == "A" and (Nx > 0 or Nu > 0):
            raise MuscadeError("Constraints with λclass=:A must have zero X-dofs and zero U-dofs")
        inod = xinod + uinod + ainod + (λinod,)
        dclass = ("X",) * Nx + ("U",) * Nu + ("A",) * Na + (λclass,)
        field = xfield + ufield + afield + (λfield,)
        name = getattr(gap, "__name__", type(gap).__name__)
        key = ("DofConstraint", λclass, xinod, xfield, uinod, ufield, ainod, afield, λinod, λfield, name,
               getattr(mode, "__name__", str(mode)))
        return key, (inod, dclass, field), {"gap": gap, "gargs": tuple(gargs), "mode": mode, "λclass": λclass,
                                            "N": (Nx, Nu, Na)}

    @staticmethod
    def residual_blocks(p, Xv, t, γ=0.):
        """SweepX: R (m, nX) and ∂R/∂X (m, nX, nX) over the element's X-dofs, from their values
        Xv (m, nX). X class: x then λ; U class: structural zeros (its X-dofs carry no residual)."""
        mE, nXe = Xv.shape
        if p["λclass"] != "X":
            return np.zeros((mE, nXe)), np.zeros((mE, nXe, nXe))
        Nx = p["N"][0]
        seeds = [Dual2.seed(Xv[:, k], k, Nx) for k in range(Nx)]
        mode = p["mode"](t)
        g = p["gap"](seeds, t, *p["gargs"])
        if not isinstance(g, Dual2):
            g = Dual2(np.broadcast_to(np.asarray(g, float), (mE,)).copy(), np.zeros((mE, Nx)), np.zeros((mE, Nx, Nx)))
        lam = Xv[:, Nx]
        R = np.zeros((mE, Nx + 1))
        K = np.zeros((mE, Nx + 1, Nx + 1))
        if mode == "off":
            R[:, Nx] = -lam
            K[:, Nx, Nx] = -1.
            return R, K
        R[:, :Nx] = -g.g * lam[:, None]
        K[:, :Nx, :Nx] = -g.h * lam[:, None, None]
        K[:, :Nx, Nx] = -g.g
        if mode == "equal":
            R[:, Nx] = -g.v
            K[:, Nx, :Nx] = -g.g
        elif mode == "positive":
            R[:, Nx] = -S(lam, g.v, γ)
            K[:, Nx, :Nx] = -lam[:, None] * g.g
            K[:, Nx, Nx] = -g.v
        else:
            raise MuscadeError(f"unknown constraint mode {mode!r}")
        return R, K

    @staticmethod
    def lagrangian(p, Λ, X, U, A, t, SP=None):
        γ = float((SP or {}).get("γ", 0.))
        Nx, Nu, Na = p["N"]
        if p["λclass"] == "X":
            # residual path (src/DirectXUA.jl:85-120): L = Λ·R with R carrying ∂R/∂X but no
            # second derivatives, so L2 has the (Λ, X) blocks and no (X, X) block
            xs = X[0]
            vals = np.stack([v.v if isinstance(v, Dual2) else np.asarray(v, float) for v in xs], axis=1)
            R, K = DofConstraint.residual_blocks(p, vals, t, γ)
            L = 0.
            for k in range(Nx + 1):
                Rk = compose(R[:, k], K[:, k, :], None, xs)
                L = L + Λ[k] * Rk
            return L
        mode = p["mode"](t)
        if p["λclass"] == "A":
            # gap(a, gargs...), λ the last A-dof (BasicElements.jl:449-458), at every step's t
            return DofConstraint._mode_lagrangian(mode, A[Na], call_cost(p["gap"], (A[:Na],), p["gargs"]), γ)
        x = X[0][:Nx]
        u = U[0][:Nu]
        a = A[:Na] if Na else []
        lam = U[0][Nu]
        g = call_cost(p["gap"], (x, u, a, t), p["gargs"])
        return DofConstraint._mode_lagrangian(mode, lam, g, γ)

    @staticmethod
    def _mode_lagrangian(mode, lam, g, γ):
        if mode == "equal":
            return -(g * lam)
        if mode == "positive":
            return -KKT(lam, g, γ)
        if mode == "off":
            return -0.5 * (lam * lam)
        raise MuscadeError(f"unknown constraint mode {mode!r}")



class ResidualElement:
    """Base of element types whose residual is a host function: Muscade's element API
    (src/ElementAPI.jl:79-173: `doflist`, `residual`, `no_second_order`) for elements that have no
    device kernel, such as user and test elements. A subclass defines

      doflist(cls) -> (inod, class, field)    as Muscade.doflist (ElementAPI.jl:79-99);
      make(cls, model, nodes0, **kw) -> dict  the parameters of the elements of one addelement
                                              call (the element constructor E(nod; kwargs...),
                                              ModelDescription.jl:3-15, 190-199);
      type_key(cls, **kw) -> tuple            what makes a distinct concrete type (a Julia type
                                              parameter such as Spring{D}'s D; default none);
      residual(p, X, U, A, t) -> sequence     the Nx residual components, as Muscade.residual
                                              (ElementAPI.jl:144-173). X[d][k] is time derivative
                                              d (0..2) of the element's k-th X-dof, U[d][k]
                                              likewise, A[k] its k-th A-dof; each is a Dual2 or a
                                              float array over the elements of the call, and the
                                              derivatives the solver does not carry are zeros.

    no_second_order = False (ElementAPI.jl:110): DirectXUA differentiates L = Λ·R to second
    order (the Lagrangian path, DirectXUA.jl:152-171); True: the residual path (:85-120).
    In SweepX the host evaluates R and its derivative along the solver's trial path
    (SweepX.jl:39-96) and the device scatters them with everything else (MH_KIND_HOST)."""
    kind = _lib.MH_KIND_HOST
    host = True
    residual_host = True
    no_second_order = False

    @classmethod
    def doflist(cls):
        raise MuscadeError(f"method 'doflist' must be provided for elements of type '{cls.__name__}'")

    @classmethod
    def make(cls, model, nodes0, **kw):
        return dict(kw)

    @classmethod
    def type_key(cls, **kw):
        return ()

    @classmethod
    def construct(cls, model, nodes0, **kw):
        return (cls.__name__,) + tuple(cls.type_key(**kw)), cls.doflist(), cls.make(model, nodes0, **kw)

    @staticmethod
    def residual(p, X, U, A, t):
        raise MuscadeError("method 'residual' must be provided")

    @classmethod
    def counts(cls):
        _, dclass, _ = cls.doflist()
        return dclass.count("X"), dclass.count("U"), dclass.count("A")

    @classmethod
    def _residual(cls, p, X, U, A, t):
        m = np.shape(X[0][0].v if isinstance(X[0][0], Dual2) else X[0][0])[0] if X and X[0] else 0
        R = list(cls.residual(p, X, U, A, t))
        Nx = cls.counts()[0]
        if len(R) != Nx:
            raise MuscadeError(f"residual of {cls.__name__} returned {len(R)} components, expected {Nx}")
        return [r if isinstance(r, Dual2) else np.broadcast_to(np.asarray(r, float), (m,)) for r in R]

    @classmethod
    def sweepx_blocks(cls, p, Xv, Uv, Av, t, c, mission, OX):
        """R (m, Nx), ∂R/∂δX (m, Nx, Nx) and, on a :step of OX ≥ 1, ∂R/∂δr (m, Nx), unscaled, at the
        element values Xv[d] (m, Nx), Uv (m, Nu), Av (m, Na): the trial path of SweepX{OX}'s addin!
        (SweepX.jl:45-96) vx = x + δX, vx′ = x′ + a₁δX + a·δr, vx″ = x″ + b₁δX + b·δr, with
        a = a₂x′ + a₃x″, b = b₂x′ + b₃x″ on :step."""
        m, Nx = Xv[0].shape
        step = mission == "step" and OX > 0
        P = Nx + (1 if step else 0)
        zero = np.zeros(m)
        X = [[None] * Nx for _ in range(3)]
        x1 = Xv[1] if OX >= 1 else np.zeros((m, Nx))
        x2 = Xv[2] if OX >= 2 else np.zeros((m, Nx))
        a = (c.a2 * x1 + c.a3 * x2) if step else None
        b = (c.b2 * x1 + c.b3 * x2) if step else None
        for k in range(Nx):
            g0 = np.zeros((m, P)); g0[:, k] = 1.
            X[0][k] = Dual2(Xv[0][:, k].astype(float), g0, np.zeros((m, P, P)))
            if OX >= 1:
                g1 = np.zeros((m, P)); g1[:, k] = c.a1
                if step:
                    g1[:, Nx] = a[:, k]
                X[1][k] = Dual2(x1[:, k].astype(float), g1, np.zeros((m, P, P)))
            else:
                X[1][k] = zero
            if OX >= 2:
                g2 = np.zeros((m, P)); g2[:, k] = c.b1
                if step:
                    g2[:, Nx] = b[:, k]
                X[2][k] = Dual2(x2[:, k].astype(float), g2, np.zeros((m, P, P)))
            else:
                X[2][k] = zero
        U = [[Uv[:, k] for k in range(Uv.shape[1])], [zero] * Uv.shape[1], [zero] * Uv.shape[1]]
        A = [Av[:, k] for k in range(Av.shape[1])]
        R = cls._residual(p, X, U, A, t)
        Rv = np.stack([r.v if isinstance(r, Dual2) else r for r in R], axis=1)
        G = np.stack([r.g if isinstance(r, Dual2) else np.zeros((m, P)) for r in R], axis=1)   # m, Nx, P
        return Rv, G[:, :, :Nx].copy(), (G[:, :, Nx].copy() if step else None)

    @classmethod
    def lagrangian(cls, p, Λ, X, U, A, t):
        """L = Λ∘₁R (getlagrangian, src/Assemble.jl:707-736) over the solver's partials. With
        no_second_order the residual is first order (revariate{1}, DirectXUA.jl:96-102): its
        second derivatives are dropped, so L keeps only the Λ-β cross terms."""
        zero = 0. * (X[0][0].v if isinstance(X[0][0], Dual2) else np.asarray(X[0][0], float))
        Xp = [list(X[d]) if d < len(X) else [zero] * len(X[0]) for d in range(3)]
        Up = [list(U[d]) if d < len(U) else [zero] * len(U[0]) for d in range(3)] if U and len(U[0]) else [[], [], []]
        R = cls._residual(p, Xp, Up, list(A), t)
        if cls.no_second_order:
            R = [Dual2(r.v, r.g, None if r.h is None else np.zeros_like(r.h)) if isinstance(r, Dual2) else r
                 for r in R]
        L = Λ[0] * R[0]
        for k in range(1, len(R)):
            L = L + Λ[k] * R[k]
        return L


class DofLoad:
    """DofLoad(nod; field, value, args...) — src/BasicElements.jl:275-284. R = [-value(t, args...)].

    `value` is a host callable; it is evaluated on the host at each step and
    the values are handed to the device (no device kernel for user functions)."""
    kind = _lib.MH_KIND_DOFLOAD
    direct_host = True   # in DirectXUA: L = Λ·(−value(t)) on the host

    @staticmethod
    def lagrangian(p, Λ, X, U, A, t):
        return Λ[0] * (-float(p["value"](t, **p["args"])))

    @classmethod
    def construct(cls, model, nodes0, field, value, **args):
        name = getattr(value, "__name__", type(value).__name__)
        return ("DofLoad", field, name, tuple(sorted(args))), ((1,), ("X",), (field,)), {"value": value, "args": args}


class DofCost:
    """DofCost(nod; xinod, xfield, uinod, ufield, ainod, afield, cost, costargs) — src/BasicElements.jl:32-50.

    A Lagrangian-only element: in SweepX its residual ∂L/∂Λ is identically
    zero, but its X-dof pairs are part of the pattern (structural zeros). In DirectXUA
    its Lagrangian L = cost(X, U, A, t, costargs...) is evaluated on the host
    (`lagrangian`): X[d][k] is time derivative d of the element's k-th X-dof, likewise U;
    A[k] the k-th A-dof."""
    kind = _lib.MH_KIND_ZERO
    host = True

    @staticmethod
    def lagrangian(p, Λ, X, U, A, t):
        return call_cost(p["cost"], (X, U, A, t), p["costargs"])

    @classmethod
    def construct(cls, model, nodes0, cost, xinod=(), xfield=(), uinod=(), ufield=(), ainod=(), afield=(), costargs=()):
        inod = tuple(xinod) + tuple(uinod) + tuple(ainod)
        dclass = ("X",) * len(xinod) + ("U",) * len(uinod) + ("A",) * len(ainod)
        field = tuple(xfield) + tuple(ufield) + tuple(afield)
        name = getattr(cost, "__name__", type(cost).__name__)
        key = ("DofCost", tuple(xinod), tuple(xfield), tuple(uinod), tuple(ufield), tuple(ainod), tuple(afield), name)
        return key, (inod, dclass, field), {"cost": cost, "costargs": costargs}


class SingleDofCost:
    """SingleDofCost(nod; class, field, cost, derivative=0, costargs) — src/BasicElements.jl:198-208:
    a DofCost on one X- or U-dof, L = cost(∂ⁿx, t, costargs...) with n = derivative.
    `class` is a Python keyword: pass it as dofclass="X" (or **{"class": "X"})."""
    kind = _lib.MH_KIND_ZERO
    host = True

    @classmethod
    def construct(cls, model, nodes0, cost, field, derivative=0, costargs=(), **kw):
        dclass = kw.get("class", kw.get("dofclass", "X"))
        if dclass not in ("X", "U"):
            raise MuscadeError("'class' must be :X or :U")
        name = getattr(cost, "__name__", type(cost).__name__)
        key = ("DofCost", dclass, field, derivative, name)
        return key, ((1,), (dclass,), (field,)), {"cost": cost, "costargs": costargs, "derivative": int(derivative),
                                                   "class": dclass}

    @staticmethod
    def lagrangian(p, Λ, X, U, A, t):
        d = p["derivative"]
        v = X[d][0] if p["class"] == "X" else U[d][0]
        return call_cost(p["cost"], (v, t), p["costargs"])


class SingleUdof:
    """SingleUdof(nod; Xfield, Ufield, cost, costargs) — src/BasicElements.jl:239-248: a U-dof applied
    as a load on an X-dof of the same node, with L = cost(u, t, costargs...) − λ·u."""
    kind = _lib.MH_KIND_ZERO
    host = True

    @classmethod
    def construct(cls, model, nodes0, Xfield, Ufield, cost, costargs=()):
        name = getattr(cost, "__name__", type(cost).__name__)
        return ("SingleUdof", Xfield, Ufield, name), ((1, 1), ("X", "U"), (Xfield, Ufield)), \
            {"cost": cost, "costargs": costargs}

    @staticmethod
    def lagrangian(p, Λ, X, U, A, t):
        u = U[0][0]
        return call_cost(p["cost"], (u, t), p["costargs"]) - Λ[0] * u


class Acost:
    """Acost(nod; inod, field, cost, costargs) — src/BasicElements.jl:77-86: a cost cost(A, costargs...)
    on A-dofs. DirectXUA adds it once through assembleA! (src/Assemble.jl:507-519, unscaled partials),
    and, like the reference's per-step element loop, again at every step through the Lagrangian path
    (scaled partials; see HostLagrangians.step)."""
    kind = _lib.MH_KIND_ZERO
    host = True
    once_off = True

    @classmethod
    def construct(cls, model, nodes0, cost, inod=(), field=(), costargs=()):
        name = getattr(cost, "__name__", type(cost).__name__)
        return ("Acost", tuple(inod), tuple(field), name), (tuple(inod), ("A",) * len(inod), tuple(field)), \
            {"cost": cost, "costargs": costargs}

    @staticmethod
    def acost(p, A):
        return call_cost(p["cost"], (A,), p["costargs"])


class SingleAcost(Acost):
    """SingleAcost(nod; field, cost, costargs) — src/BasicElements.jl:163-167: Acost on one A-dof,
    cost(a, costargs...)."""

    @classmethod
    def construct(cls, model, nodes0, field, cost, costargs=()):
        name = getattr(cost, "__name__", type(cost).__name__)
        return ("Acost", (1,), (field,), "single", name), ((1,), ("A",), (field,)), {"cost": cost, "costargs": costargs}

    @staticmethod
    def acost(p, A):
        return call_cost(p["cost"], (A[0],), p["costargs"])
