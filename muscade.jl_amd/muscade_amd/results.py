"""Element results on the device: getresult of EulerBeam3D's requestables (src/Output.jl:131-181,
toolbox/BeamElement.jl:28-64,151-174) and of StrainGaugeOnEulerBeam3D
(toolbox/StrainGaugeOnBeamElement.jl:62-76), without host re-evaluation of the elements.

A request mirrors `@request`: `request("ε", gp=request("κgp", resultants=request("fᵢ", "mᵢ")))`
stands for `@request ε, gp(κgp, resultants(fᵢ, mᵢ))`. getresult returns nested dicts of numpy
arrays whose first axis is the element (eleres[iele] in the reference) and, under "gp", whose
second axis is the Gauss point (eleres[iele].gp[igp]). Quantities with time levels (ε, κ, x,
κgp) carry an axis of the state's X derivatives (ND) before their components.
"""
import numpy as np

from . import _lib
from ._lib import as_f64, dptr, lib
from .assemble import DeviceContext, upload_model
from .model import MuscadeError

BEAM_GP = ("x", "κgp", "resultants")
BEAM_RESULTANTS = ("fᵢ", "mᵢ", "fₑ", "mₑ")


def request(*names, **sub):
    """@request: leaf names, and sub-requests for nested requestables."""
    r = {n: None for n in names}
    r.update(sub)
    return r


def _nres(ND):
    return 4 * ND + 4 * (6 * ND + 10)


def _split(raw, ND, ngauge):
    """mh_beam_results' layout as named arrays (leading axis: element)."""
    n = raw.shape[0]
    q = 0
    out = {"ε": raw[:, q:q + ND]}; q += ND
    out["κ"] = raw[:, q:q + 3 * ND].reshape(n, ND, 3); q += 3 * ND
    per = 6 * ND + 10
    g = raw[:, q:q + 4 * per].reshape(n, 4, per); q += 4 * per
    gp = {"x": g[:, :, :3 * ND].reshape(n, 4, ND, 3), "κgp": g[:, :, 3 * ND:6 * ND].reshape(n, 4, ND, 3)}
    o = 6 * ND
    gp["resultants"] = {"fᵢ": g[:, :, o], "mᵢ": g[:, :, o + 1:o + 4], "fₑ": g[:, :, o + 4:o + 7],
                        "mₑ": g[:, :, o + 7:o + 10]}
    out["gp"] = gp
    out["gauges"] = raw[:, q:q + ngauge]
    return out


def _select(res, req, where):
    sel = {}
    for k, v in req.items():
        if k not in res:
            raise MuscadeError(f"{k} is not a requestable of {where} (requestables: {', '.join(res)})")
        sel[k] = res[k] if v is None else _select(res[k], v, f"{where}.{k}")
    return sel


class ElementResults:
    """The model's element groups on a device context, for result extraction."""

    def __init__(self, model, dis, device=0):
        self.model, self.dis = model, dis
        self.ctx = DeviceContext(device)
        self.groups = upload_model(self.ctx, model, dis)

    def beam(self, state, ieletyp, ND=None):
        """All requestables of beam type ieletyp (0-based) at `state` (X up to ND levels)."""
        E = self.model.eletypes[ieletyp]
        if E.cls.kind not in (_lib.MH_KIND_BEAM, _lib.MH_KIND_BEAM_U):
            raise MuscadeError(f"element type {ieletyp + 1} ({E.key[0]}) has no device requestables")
        ND = min(len(state.X), 3) if ND is None else ND
        h = self.ctx.h
        for d in range(ND):
            self.ctx.check(lib.mh_set_X(h, d, dptr(as_f64(state.X[d]))), "set X")
        if self.model.ndof["U"]:
            self.ctx.check(lib.mh_set_U(h, dptr(as_f64(state.U[0]))), "set U")
        gauges = None
        ng = 0
        if "gauges" in E.params[0]:
            gauges = as_f64(np.concatenate([p["gauges"] for p in E.params]))
            ng = gauges.shape[1]
        raw = np.zeros((E.nele, _nres(ND) + ng))
        self.ctx.check(lib.mh_beam_results(h, self.groups[ieletyp], ND, ng, dptr(gauges), dptr(raw)), "beam results")
        return _split(raw, ND, ng)

    def gauge_derivatives(self, ieletyp, state):
        return gauge_derivatives_on(self.ctx, self.groups[ieletyp], state, self.model.eletypes[ieletyp].nele)


def gauge_derivatives_on(ctx, group, state, nele, upload=True, hessians=True):
    """(values (nele,4), gradients (nele,4,12), Hessians (nele,4,12,12)) of εₐₓ, κ₁, κ₂, κ₃
    w.r.t. the element X-dofs of a beam group at state.X[0] (mh_beam_gauge_derivs);
    upload=False: at the X already on the device (state unused); hessians=False: None in
    their place (for first-order partials)."""
    if upload:
        ctx.check(lib.mh_set_X(ctx.h, 0, dptr(as_f64(state.X[0]))), "set X")
    raw = np.zeros((nele, 364))
    ctx.check(lib.mh_beam_gauge_derivs(ctx.h, group, 0, None, dptr(raw)), "gauge derivatives")
    if not hessians:
        return raw[:, :4], raw[:, 4:52].reshape(nele, 4, 12), None
    H = np.zeros((nele, 4, 12, 12))
    iu = np.triu_indices(12)
    h78 = raw[:, 52:].reshape(nele, 4, 78)
    H[:, :, iu[0], iu[1]] = h78
    H[:, :, iu[1], iu[0]] = h78
    return raw[:, :4], raw[:, 4:52].reshape(nele, 4, 12), H


def getresult(state, req, ieletyp, device=0, _cache={}):
    """eleres = getresult(state, req, eletype) for all elements of type ieletyp (0-based)
    (src/Output.jl:164-175). EulerBeam3D requestables: ε, κ, gp(x, κgp, resultants(fᵢ, mᵢ,
    fₑ, mₑ)); StrainGaugeOnEulerBeam3D: εₐₓ, κ, ε (the gauge strains) and eleres(...) with
    the beam's requestables; ElementConstraint over StrainGaugeOnEulerBeam3D: λ, gap and
    eleres(...) with the strain gauge's requestables."""
    model = state.model
    key = (id(model), device)
    er = _cache.get(key)
    if er is None or er.model is not model:
        er = _cache[key] = ElementResults(model, state.dis, device)
    r = er.beam(state, ieletyp)
    E = model.eletypes[ieletyp]
    if "gauges" in E.params[0]:
        res = {"εₐₓ": r["ε"][:, 0], "κ": r["κ"][:, 0], "ε": r["gauges"],
               "eleres": {k: r[k] for k in ("ε", "κ", "gp")}}
        if E.key[0] == "ElementConstraint":
            res = _constraint_results(state, ieletyp, res)
        return _select(res, req, E.key[0])
    return _select({k: r[k] for k in ("ε", "κ", "gp")}, req, E.key[0])


def _constraint_results(state, ieletyp, gauge):
    """ElementConstraint's requestables (src/BasicElements.jl:564-576): λ, gap (the gap function
    of the requested gauge requestables, evaluated on the host) and eleres(...), the target
    StrainGaugeOnEulerBeam3D's requestables."""
    from .hostcost import call_cost
    E = state.model.eletypes[ieletyp]
    lam = state.U[0][state.dis.dis[ieletyp].U[:, -1] - 1]
    gap = np.zeros(E.nele)
    e0 = 0
    for p in E.params:
        m = p["gauges"].shape[0]
        sl = slice(e0, e0 + m)
        full = {"εₐₓ": gauge["εₐₓ"][sl], "κ": [gauge["κ"][sl, k] for k in range(3)],
                "ε": [gauge["ε"][sl, k] for k in range(gauge["ε"].shape[1])]}
        eleres = {}
        for name in p["req"]:
            if name not in full:
                raise MuscadeError(f"{name} is not a requestable of StrainGaugeOnEulerBeam3D")
            eleres[name] = full[name]
        gap[sl] = call_cost(p["gap"], (eleres, state.time), p["gargs"])
        e0 += m
    return {"λ": lam, "gap": gap, "eleres": gauge}
