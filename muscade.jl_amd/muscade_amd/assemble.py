"""AssemblySweepX on the GPU: prepare / zero! / assemble! of src/SweepX.jl:17-96.

The host keeps Muscade's call structure; the element loop, the index maps and
the scatter run in libmuscade_hip.so (include/muscade_hip.h).
"""
import ctypes

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import as_f64, as_i64, dptr, i64ptr, lib
from .model import MuscadeError


class Newmarkβcoefficients:
    """Newmarkβcoefficients{OX}(Δt,β,γ) — src/SweepX.jl:3-15."""

    def __init__(self, OX, Δt=0., β=0.25, γ=0.5):
        self.OX, self.Δt = OX, Δt
        if OX == 0 or Δt == 0.:
            self.a1 = self.a2 = self.a3 = self.b1 = self.b2 = self.b3 = 0.
        elif OX == 1:
            self.a1, self.a2, self.a3, self.b1, self.b2, self.b3 = 1 / (γ * Δt), 1 / γ, 0., 0., 0., 0.
        else:
            self.a1, self.a2, self.a3 = γ / (β * Δt), γ / β, (γ / (2 * β) - 1) * Δt
            self.b1, self.b2, self.b3 = 1 / (β * Δt ** 2), 1 / (β * Δt), 1 / (2 * β)

    def array(self):
        return np.array([self.a1, self.a2, self.a3, self.b1, self.b2, self.b3])


class DeviceContext:
    """Owns one mh_ctx (one HIP stream and the device buffers of one model)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        r = lib.mh_ctx_create(device, ctypes.byref(h))
        if r != _lib.MH_OK:
            raise RuntimeError(f"mh_ctx_create failed: {lib.mh_strerror(r).decode()} (no usable GPU {device}?)")
        self.h = h

    def check(self, r, what):
        if r != _lib.MH_OK:
            msg = lib.mh_last_error(self.h).decode()
            raise MuscadeError(f"{what}: {lib.mh_strerror(r).decode()}: {msg}")

    def __del__(self):
        if getattr(self, "h", None):
            lib.mh_ctx_destroy(self.h)
            self.h = None


def upload_model(ctx, model, dis, skip=(), nA=None):
    """Describe the model to the device: one group per element type, in eletype order.
    Element types in `skip` (evaluated on the host) get no group (None); nA overrides the
    model's A-dof count (0 when no device kind uses A-dofs)."""
    xyz = as_f64(model.coords()[:, :3]) if model.nnod else np.zeros((0, 3))
    nA = model.ndof["A"] if nA is None else nA
    ctx.check(lib.mh_model_begin(ctx.h, model.ndof["X"], model.ndof["U"], nA, model.nnod,
                                 dptr(xyz) if model.nnod else None), "mh_model_begin")
    groups = []
    for it, E in enumerate(model.eletypes):
        if it in skip:
            groups.append(None)
            continue
        d = dis.dis[it]
        nele = d.X.shape[0]
        nX = d.X.shape[1]
        idx = as_i64(d.X)
        sX = as_f64(d.scaleX) if nX else None
        nodes = orient = mat = idxU = None
        nU = 0
        kind = E.cls.kind
        if kind in (_lib.MH_KIND_BEAM, _lib.MH_KIND_BEAM_U):
            nodes = as_i64(E.cat_nodes()[:, :2] + 1)
            orient = as_f64(np.concatenate([p["orient2"] for p in E.params]))
            mat = as_f64(np.concatenate([p["mat"] for p in E.params]))
        elif kind in (_lib.MH_KIND_BAR, _lib.MH_KIND_BAR_U):
            nodes = as_i64(E.cat_nodes()[:, :2] + 1)
            mat = as_f64(np.concatenate([np.concatenate([p["mat"], p["eps_s"][:, None]], axis=1) for p in E.params]))
        if kind in (_lib.MH_KIND_BEAM_U, _lib.MH_KIND_BAR_U):
            idxU = as_i64(d.U)
            nU = 3
        g = ctypes.c_int()
        ctx.check(lib.mh_group_add(ctx.h, kind, nele, nX, i64ptr(idx) if nX else None, dptr(sX), nU, i64ptr(idxU),
                                   i64ptr(nodes), dptr(orient), dptr(mat), ctypes.byref(g)), f"element type {it + 1}")
        amask = E.params[0].get("amask", 0) if E.params else 0
        if amask:   # EulerBeam3DA: stiffnesses scaled by 10^A
            if nA == 0:
                raise MuscadeError(f"element type {it + 1} ({E.key[0]}) is scaled by A-dofs, which this "
                                   "assembler does not carry")
            ctx.check(lib.mh_group_set_adofs(ctx.h, g.value, i64ptr(as_i64(d.A[:, :1])), int(amask)),
                      f"A-dofs of element type {it + 1}")
        groups.append(g.value)
    return groups


class AssemblySweepX:
    """AssemblySweepX{OX}: prepare(AssemblySweepX{OX},model,dis) + assemble!{mission} (SweepX.jl:17-96)."""

    def __init__(self, OX, model, dis, device=0, graph=False):
        """graph=True: each assemble! replays one captured HIP graph of the element kernels and
        scatters, one per mission (recaptured when the dof scales change, or the time when
        bars are present: their kernel takes it); results are bitwise those of the eager
        launches."""
        self.OX = OX
        self.graph = graph
        self.model, self.dis = model, dis
        self.ctx = DeviceContext(device)
        self.groups = upload_model(self.ctx, model, dis)
        nnz = ctypes.c_int64()
        self.ctx.check(lib.mh_prepare_sweepx(self.ctx.h, OX, ctypes.byref(nnz)), "prepare")
        self.nX = model.ndof["X"]
        self.nnz = nnz.value
        self.colptr = np.zeros(self.nX + 1, np.int64)
        self.rowval = np.zeros(self.nnz, np.int64)
        self.ctx.check(lib.mh_get_pattern(self.ctx.h, i64ptr(self.colptr), i64ptr(self.rowval)), "pattern")
        self.c = Newmarkβcoefficients(OX)
        self.Lλ = np.zeros(self.nX)
        self.nzval = np.zeros(self.nnz)
        # element types evaluated on the host (user functions): their blocks are handed over
        # before each assembly (mh_set_host_blocks)
        self.host_types = [it for it, E in enumerate(model.eletypes)
                           if E.cls.kind == _lib.MH_KIND_HOST and E.nele and dis.dis[it].X.shape[1]]
        self._indptr0 = self.colptr - 1
        self._indices0 = self.rowval - 1

    def asm(self, ieletyp):
        """asm[1,ieletyp], asm[2,ieletyp] as Muscade builds them (1-based, element-major)."""
        d = self.dis.dis[ieletyp]
        nele, nX = d.X.shape
        a1 = np.zeros((nele, nX), np.int64)
        a2 = np.zeros((nele, nX * nX), np.int64)
        self.ctx.check(lib.mh_get_asm(self.ctx.h, self.groups[ieletyp], i64ptr(a1), i64ptr(a2)), "asm")
        return a1, a2

    def set_state(self, state, t, X=True):
        """Upload the state (X=False: all but X, X′, X″, which stay as they are on the device)."""
        h = self.ctx.h
        self.ctx.check(lib.mh_set_time(h, float(t)), "set time")
        for ider in range(min(self.OX + 1, len(state.X)) if X else 0):
            self.ctx.check(lib.mh_set_X(h, ider, dptr(as_f64(state.X[ider]))), "set X")
        if self.model.ndof["U"]:
            self.ctx.check(lib.mh_set_U(h, dptr(as_f64(state.U[0]))), "set U")
        if self.model.ndof["A"]:
            self.ctx.check(lib.mh_set_A(h, dptr(as_f64(state.A))), "set A")
        for it, E in enumerate(self.model.eletypes):
            if E.cls.kind in (_lib.MH_KIND_DOFLOAD, _lib.MH_KIND_HOLD) and E.nele and "value" in E.params[0]:
                # DofLoad's value(t), DofMotion's prescribed value
                vals = np.concatenate([np.full(n.shape[0], float(p["value"](t, **p["args"])))
                                       for n, p in zip(E.nodes, E.params)])
                self.ctx.check(lib.mh_set_point_values(h, self.groups[it], dptr(as_f64(vals))), "point values")

    def assemble(self, state, mission, dbg=None, fetch=True, upload=True):
        """assemble!{:step|:iter}(out,asm,dis,model,state,Δt,dbg). mission: 'step' or 'iter'.
        upload=False assembles from the state already on the device (a device-resident
        Newton loop, muscade_amd.chainsolve)."""
        if upload:
            self.set_state(state, state.time)
        if self.host_types:
            self._host_blocks(state, upload, dbg, mission)
        err = _lib.mh_err()
        m = 0 if mission == "step" else 1
        fn = lib.mh_assemble_sweepx_graph if self.graph else lib.mh_assemble_sweepx
        r = fn(self.ctx.h, self.OX, m, dptr(self.c.array()), ctypes.byref(err))
        if r == _lib.MH_ERR_NAN:
            E = self.model.eletypes[err.igroup]
            raise MuscadeError(f"residual({E.key[0]},...) returned NaN in R, FB or derivatives",
                               dict(dbg or {}, ieletyp=err.igroup + 1, iele=err.iele + 1, t=state.time))
        self.ctx.check(r, "assemble")
        if fetch:
            self.ctx.check(lib.mh_get_rhs(self.ctx.h, dptr(self.Lλ)), "rhs")
            self.ctx.check(lib.mh_get_nzval(self.ctx.h, dptr(self.nzval)), "nzval")
        return self.Lλ, self.nzval

    def _host_blocks(self, state, upload, dbg, mission="iter"):
        """R and ∂R/∂δX (and ∂R/∂δr on :step) of the host-evaluated element types at the current
        state, along SweepX{OX}'s trial path (SweepX.jl:45-96)."""
        nd = self.OX + 1 if any(getattr(self.model.eletypes[it].cls, "residual_host", False)
                                for it in self.host_types) else 1
        if upload:
            Xs = [state.X[d] if d < len(state.X) else np.zeros(self.nX) for d in range(nd)]
        else:
            Xs = [np.zeros(self.nX) for _ in range(nd)]
            for d in range(nd):
                self.ctx.check(lib.mh_get_X(self.ctx.h, d, dptr(Xs[d])), "X")
        SP = getattr(state, "SP", None)
        γ = float(SP.get("γ", 0.)) if isinstance(SP, dict) else 0.   # default{:γ}(SP, 0.)
        for it in self.host_types:
            E = self.model.eletypes[it]
            d = self.dis.dis[it]
            Rs, Ks, dRs, e0 = [], [], [], 0
            for nodes, p in zip(E.nodes, E.params):
                m = nodes.shape[0]
                iX = d.X[e0:e0 + m] - 1
                if getattr(E.cls, "residual_host", False):
                    Uv = state.U[0][d.U[e0:e0 + m] - 1] if d.U.shape[1] else np.zeros((m, 0))
                    Av = state.A[d.A[e0:e0 + m] - 1] if d.A.shape[1] else np.zeros((m, 0))
                    R, K, dR = E.cls.sweepx_blocks(p, [x[iX] for x in Xs], Uv, Av, state.time, self.c, mission,
                                                   self.OX)
                else:
                    R, K = E.cls.residual_blocks(p, Xs[0][iX], state.time, γ)
                    dR = None
                Rs.append(R); Ks.append(K); dRs.append(dR)
                e0 += m
            R, K = np.concatenate(Rs), np.concatenate(Ks)
            bad = ~(np.isfinite(R).all(axis=1) & np.isfinite(K).reshape(K.shape[0], -1).all(axis=1))
            if bad.any():
                raise MuscadeError(f"residual({E.key[0]},...) returned NaN in R, FB or derivatives",
                                   dict(dbg or {}, ieletyp=it + 1, iele=int(np.nonzero(bad)[0][0]) + 1, t=state.time))
            self.ctx.check(lib.mh_set_host_blocks(self.ctx.h, self.groups[it], dptr(as_f64(R)),
                                                  dptr(as_f64(K.transpose(0, 2, 1)))), "host blocks")
            if any(x is not None for x in dRs):
                dR = np.concatenate([x if x is not None else np.zeros_like(r) for x, r in zip(dRs, Rs)])
                if not np.isfinite(dR).all():
                    raise MuscadeError(f"residual({E.key[0]},...) returned NaN in R, FB or derivatives",
                                       dict(dbg or {}, ieletyp=it + 1, t=state.time))
                self.ctx.check(lib.mh_set_host_dr(self.ctx.h, self.groups[it], dptr(as_f64(dR))), "host ∂R/∂δr")

    def graph_captures(self):
        n = ctypes.c_int64()
        self.ctx.check(lib.mh_graph_captures(self.ctx.h, ctypes.byref(n)), "graph captures")
        return n.value

    def vjp(self, state, Λ, OX=None, dbg=None, upload=True):
        """Gradients of L = Σ_e Λ_eᵀ R_e(X,U,A,t) (getlagrangian, Assemble.jl:707-736) at `state`:
        returns ([∂L/∂X, ∂L/∂X′, …] up to order OX, ∂L/∂U, ∂L/∂A), unscaled model coordinates.
        Computed by the device adjoint kernels (include/muscade_hip.h, mh_vjp_lagrangian).
        upload=False: at the state already on the device (the last assembly's)."""
        OX = self.OX if OX is None else OX
        if upload:
            self.set_state(state, state.time)
        self.ctx.check(lib.mh_set_lambda(self.ctx.h, dptr(as_f64(Λ))), "set Λ")
        err = _lib.mh_err()
        r = lib.mh_vjp_lagrangian(self.ctx.h, OX, ctypes.byref(err))
        if r == _lib.MH_ERR_NAN:
            E = self.model.eletypes[err.igroup]
            raise MuscadeError(f"adjoint of {E.key[0]} returned NaN",
                               dict(dbg or {}, ieletyp=err.igroup + 1, iele=err.iele + 1, t=state.time))
        self.ctx.check(r, "adjoint")
        gX = np.zeros((OX + 1) * self.nX)
        gU = np.zeros(self.model.ndof["U"])
        gA = np.zeros(self.model.ndof["A"])
        self.ctx.check(lib.mh_get_vjp(self.ctx.h, dptr(gX), dptr(gU) if gU.size else None, dptr(gA) if gA.size else None),
                       "adjoint results")
        return [gX[d * self.nX:(d + 1) * self.nX] for d in range(OX + 1)], gU, gA

    def Lλx(self):
        """out.Lλx as a scipy CSC matrix sharing the pattern."""
        return sp.csc_matrix((self.nzval.copy(), self._indices0, self._indptr0), shape=(self.nX, self.nX))

    def element_blocks(self, ieletyp):
        d = self.dis.dis[ieletyp]
        nele, nX = d.X.shape
        R = np.zeros((nele, nX))
        K = np.zeros((nele, nX * nX))
        self.ctx.check(lib.mh_get_element_blocks(self.ctx.h, self.groups[ieletyp], dptr(R), dptr(K)), "element blocks")
        return R, K.reshape(nele, nX, nX).transpose(0, 2, 1)   # K[e][i][j]

    def timings(self):
        ms = np.zeros(3)
        self.ctx.check(lib.mh_last_timings(self.ctx.h, dptr(ms)), "timings")
        return ms


CLASSES_DIRECT = ("Λ", "X", "U")


class AssemblyDirect:
    """AssemblyDirect{OX,OU,IA} for one DirectXUA step on the GPU: prepare (DirectXUA.jl:22-56)
    and assemble!{:vectors|:matrices} (DirectXUA.jl:85-171) over the classes Λ, X, U.

    After `assemble`, `L1[α][αder]` are vectors and `L2[(α,β)][αder][βder]` are scipy CSC
    matrices with the class-pair patterns, as out.L1/out.L2 of the reference (without the
    empty L2[Λ,Λ])."""

    def __init__(self, OX, OU, IA, model, dis, device=0, skip=(), nA=None):
        """skip: element types left to the host (user-cost elements, whose Lagrangians the
        DirectXUA solver evaluates itself); the device then assembles the others only."""
        self.OX, self.OU, self.IA = OX, OU, IA
        self.model, self.dis = model, dis
        self.ctx = DeviceContext(device)
        # A-scaled beams (EulerBeam3DA) on the device read A; otherwise the device needs no A-dofs
        self.ascaled = [it for it, E in enumerate(model.eletypes)
                        if it not in skip and E.params and E.params[0].get("amask")]
        if nA is None and skip and not self.ascaled:
            nA = 0
        self.groups = upload_model(self.ctx, model, dis, skip=skip, nA=nA)
        for it, g in enumerate(self.groups):
            d = dis.dis[it]
            if g is not None and d.X.shape[1]:
                sU = as_f64(d.scaleU) if d.U.shape[1] else None
                self.ctx.check(lib.mh_set_group_scales(self.ctx.h, g, dptr(as_f64(d.scaleΛ)), dptr(sU)), "scales")
        nnz = np.zeros(4, np.int64)
        self.ctx.check(lib.mh_prepare_direct(self.ctx.h, OX, OU, IA, i64ptr(nnz)), "prepare direct")
        self.nX, self.nU = model.ndof["X"], model.ndof["U"]
        shapes = [(self.nX, self.nX), (self.nX, self.nU), (self.nU, self.nX), (self.nU, self.nU)]
        self.patterns = []
        for w, (nr, nc) in enumerate(shapes):
            cp = np.zeros(nc + 1, np.int64)
            rv = np.zeros(nnz[w], np.int64)
            self.ctx.check(lib.mh_get_direct_pattern(self.ctx.h, w, i64ptr(cp), i64ptr(rv)), "pattern")
            self.patterns.append((cp, rv, (nr, nc)))
        self.nder = {"Λ": 1, "X": OX + 1, "U": OU + 1}

    def beam_dRdA(self, it):
        """∂R/∂a (nele × 12, unscaled) of the A-scaled beam type it at the state of the last
        `assemble` (mh_beam_dRdA)."""
        E = self.model.eletypes[it]
        out = np.zeros((E.nele, 12))
        r = lib.mh_beam_dRdA(self.ctx.h, self.groups[it], dptr(out))
        if r == _lib.MH_ERR_NAN:
            raise MuscadeError(f"residual({E.key[0]},...) returned NaN in ∂R/∂A", dict(ieletyp=it + 1))
        self.ctx.check(r, "dR/dA")
        return out

    def gauge_derivatives(self, it, state):
        """Values, gradients and Hessians of the gauge requestables of beam type it at
        state.X[0] (for ElementCost types; mh_beam_gauge_derivs on this context)."""
        from .results import gauge_derivatives_on
        return gauge_derivatives_on(self.ctx, self.groups[it], state, self.model.eletypes[it].nele)

    def _which(self, a, b):
        return (0 if a != "U" else 2) if b != "U" else (1 if a != "U" else 3)

    def assemble(self, state, Λ, mission="matrices", Δt=1.0, dbg=None):
        h = self.ctx.h
        self.ctx.check(lib.mh_set_time(h, float(state.time)), "time")
        for ider in range(self.OX + 1):
            self.ctx.check(lib.mh_set_X(h, ider, dptr(as_f64(state.X[ider]))), "X")
        if self.nU:
            self.ctx.check(lib.mh_set_U(h, dptr(as_f64(state.U[0]))), "U")
        self.ctx.check(lib.mh_set_lambda(h, dptr(as_f64(Λ))), "Λ")
        if self.ascaled:
            self.ctx.check(lib.mh_set_A(h, dptr(as_f64(state.A))), "A")
        for it, E in enumerate(self.model.eletypes):
            if E.cls.kind == _lib.MH_KIND_DOFLOAD and E.nele and self.groups[it] is not None:
                vals = np.concatenate([np.full(n.shape[0], float(p["value"](state.time, **p["args"])))
                                       for n, p in zip(E.nodes, E.params)])
                self.ctx.check(lib.mh_set_point_values(h, self.groups[it], dptr(as_f64(vals))), "DofLoad values")
        err = _lib.mh_err()
        r = lib.mh_assemble_direct(h, 1 if mission == "matrices" else 0, float(Δt), ctypes.byref(err))
        if r == _lib.MH_ERR_NAN:
            E = self.model.eletypes[err.igroup]
            raise MuscadeError(f"residual({E.key[0]},...) returned NaN",
                               dict(dbg or {}, ieletyp=err.igroup + 1, iele=err.iele + 1, t=state.time))
        self.ctx.check(r, "assemble direct")
        nnz = [p[1].size for p in self.patterns]
        nL1 = self.nX + (self.OX + 1) * self.nX + (self.OU + 1) * self.nU
        sizes = {}
        nL2 = 0
        for a in CLASSES_DIRECT:
            for b in CLASSES_DIRECT:
                if a == "Λ" and b == "Λ":
                    continue
                sizes[(a, b)] = nL2
                nL2 += self.nder[a] * self.nder[b] * nnz[self._which(a, b)]
        L1f, L2f = np.zeros(nL1), np.zeros(nL2)
        self.ctx.check(lib.mh_get_direct(h, dptr(L1f), dptr(L2f)), "direct results")
        self.flat = (L1f, L2f)
        L1 = {"Λ": [L1f[:self.nX]],
              "X": [L1f[self.nX * (1 + d):self.nX * (2 + d)] for d in range(self.OX + 1)],
              "U": [L1f[self.nX * (self.OX + 2) + self.nU * d:self.nX * (self.OX + 2) + self.nU * (d + 1)]
                    for d in range(self.OU + 1)]}
        L2 = {}
        for (a, b), off in sizes.items():
            cp, rv, shape = self.patterns[self._which(a, b)]
            n = rv.size
            blocks = []
            for i in range(self.nder[a]):
                row = []
                for j in range(self.nder[b]):
                    v = L2f[off + (i * self.nder[b] + j) * n: off + (i * self.nder[b] + j + 1) * n]
                    row.append(sp.csc_matrix((v.copy(), rv - 1, cp - 1), shape=shape))
                blocks.append(row)
            L2[(a, b)] = blocks
        return L1, L2
