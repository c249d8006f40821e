"""Shared test helpers: build models with the product's host API and hand the
same model to the CPU oracle (test infrastructure)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "muscade.jl_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

from tests import _oracle as orc  # noqa: E402


def model_api():
    """The host package without loading the HIP library (model description only)."""
    from muscade_amd import model, elements  # noqa: F401
    return model, elements


def build_lattice(nx, ny, nz, seed=7, disp=1e-3, rot=1e-2, sigma=0.2):
    from muscade_amd.workloads import space_frame
    return space_frame(nx, ny, nz, seed=seed, disp=disp, rot=rot, sigma=sigma)


KIND_TO_ORACLE = {1: orc.K_BEAM, 2: orc.K_BEAM, 3: orc.K_HOLD, 4: orc.K_DOFLOAD, 5: orc.K_ZERO, 6: orc.K_BAR, 7: orc.K_BAR}


def oracle_model(model, dis, t=0.0, skip=(), nA=None, A=None):
    """The same model for the oracle's sequential assembler (element types in `skip` left out).
    A: the A-dof values, which give A-scaled beams (EulerBeam3DA) their scaled sections (the
    oracle's beam has no A-dofs; it evaluates the scaled section it is given)."""
    om = orc.OracleModel(model.ndof["X"], model.ndof["U"], model.ndof["A"] if nA is None else nA)
    xyz = model.coords()
    io = -1
    for it, E in enumerate(model.eletypes):
        if it in skip:
            continue
        io += 1
        d = dis.dis[it]
        kind = KIND_TO_ORACLE[E.cls.kind]
        params = None
        if kind == orc.K_BEAM:
            nodes = E.cat_nodes()
            o2 = np.concatenate([p["orient2"] for p in E.params])
            mat = np.concatenate([p["mat"] for p in E.params])
            if E.params[0].get("amask") and A is not None:
                mat = E.cls.scaled_section(mat, np.asarray(A)[d.A[:, 0] - 1], E.params[0]["amask"])
            params = np.concatenate([xyz[nodes[:, 0], :3], xyz[nodes[:, 1], :3], o2, mat], axis=1)
        elif kind == orc.K_BAR:
            nodes = E.cat_nodes()
            mat = np.concatenate([p["mat"] for p in E.params])
            eps = np.concatenate([p["eps_s"] for p in E.params])[:, None]
            params = np.concatenate([xyz[nodes[:, 0], :3], xyz[nodes[:, 1], :3], mat, eps], axis=1)
        elif kind == orc.K_DOFLOAD or (kind == orc.K_HOLD and E.nele and "value" in E.params[0]):
            params = np.concatenate([np.full(n.shape[0], float(p["value"](t, **p["args"])))
                                     for n, p in zip(E.nodes, E.params)])
        om.add_type(kind, d.X, d.U if d.U.shape[1] else None, d.A if d.A.shape[1] else None, d.scaleX, params,
                    udof=int(getattr(E.cls, "Udof", False)))
        om.set_scales(io, d.scaleΛ, d.scaleU if d.U.shape[1] else None)
    om.prepare()
    om.set_time(t)
    return om


def oracle_assemble_sweepx0(model, state, mission=1):
    om = oracle_model(model, state.dis, t=state.time)
    r, rhs, nz, err = om.assemble(0, mission, np.zeros(6), [state.X[0]],
                                  state.U[0] if model.ndof["U"] else None, state.A if model.ndof["A"] else None)
    assert r == 0, (r, err)
    return {"rhs": rhs, "nzval": nz, "colptr": om.colptr, "rowval": om.rowval, "om": om}


def relerr(a, b):
    """Normwise relative error max|a-b| / max|b| (the parity metric of the tests)."""
    a, b = np.asarray(a), np.asarray(b)
    d = np.abs(b).max()
    return np.abs(a - b).max() / (d if d > 0 else 1.0)


class OracleAssembly:
    """The oracle's sequential assembler behind the interface of muscade_amd.AssemblySweepX,
    so that the host solver loop can be run against it (test infrastructure only)."""

    def __init__(self, OX, model, dis, device=0):
        import scipy.sparse as sp
        from muscade_amd.assemble import Newmarkβcoefficients
        self._sp = sp
        self.OX, self.model, self.dis = OX, model, dis
        self.nX = model.ndof["X"]
        self.c = Newmarkβcoefficients(OX)
        self.om = None
        self.Lλ = self.nzval = None

    def assemble(self, state, mission, dbg=None, fetch=True):
        self.om = oracle_model(self.model, self.dis, t=state.time, A=state.A)
        r, rhs, nz, err = self.om.assemble(self.OX, 0 if mission == "step" else 1, self.c.array(),
                                           state.X[:self.OX + 1], state.U[0] if self.model.ndof["U"] else None,
                                           state.A if self.model.ndof["A"] else None)
        assert r == 0, err
        self.Lλ, self.nzval = rhs, nz
        return rhs, nz

    def Lλx(self):
        om = self.om
        return self._sp.csc_matrix((self.nzval.copy(), om.rowval - 1, om.colptr - 1), shape=(self.nX, self.nX))


class OracleAssemblyDirect:
    """The oracle's AssemblyDirect behind the interface of muscade_amd.AssemblyDirect, so that the
    DirectXUA host solver can be run against it (test infrastructure only)."""

    def __init__(self, OX, OU, IA, model, dis, device=0, skip=()):
        self.OX, self.OU, self.IA, self.model, self.dis, self.skip = OX, OU, IA, model, dis, tuple(skip)
        self.nX, self.nU = model.ndof["X"], model.ndof["U"]
        self.nder = {"Λ": 1, "X": OX + 1, "U": OU + 1}
        self._state = None

    def beam_dRdA(self, it):
        """∂R/∂a of an A-scaled beam type at the last assembled state: ln 10 · R(X; the masked
        stiffnesses of the scaled section alone, no loads), element by element with the oracle's
        beam."""
        E, d, st = self.model.eletypes[it], self.dis.dis[it], self._state
        xyz = self.model.coords()
        nodes = E.cat_nodes()
        o2 = np.concatenate([p["orient2"] for p in E.params])
        amask = E.params[0]["amask"]
        mat = E.cls.scaled_section(np.concatenate([p["mat"] for p in E.params]), st.A[d.A[:, 0] - 1], amask).copy()
        for k in range(4):
            if not amask >> k & 1:
                mat[:, k] = 0.
        mat[:, 4:] = 0.
        out = np.zeros((E.nele, 12))
        for e in range(E.nele):
            r, _, _ = orc.beam_diffed_residual(xyz[nodes[e, 0], :3], xyz[nodes[e, 1], :3], o2[e], mat[e],
                                               [st.X[0][d.X[e] - 1]], np.zeros(3))
            out[e] = np.log(10.) * r
        return out

    def gauge_derivatives(self, it, state):
        """The oracle's (εₐₓ, κ) values, gradients and Hessians (orc_beam_gauge2) per element."""
        E = self.model.eletypes[it]
        xyz = self.model.coords()
        nodes = E.cat_nodes()
        o2 = np.concatenate([p["orient2"] for p in E.params])
        mat = np.concatenate([p["mat"] for p in E.params])
        idx = self.dis.dis[it].X - 1
        v, g, H = np.zeros((E.nele, 4)), np.zeros((E.nele, 4, 12)), np.zeros((E.nele, 4, 12, 12))
        for e in range(E.nele):
            v[e], g[e], H[e] = orc.beam_gauge2(xyz[nodes[e, 0], :3], xyz[nodes[e, 1], :3], o2[e], mat[e],
                                               state.X[0][idx[e]])
        return v, g, H

    def assemble(self, state, Λ, mission="matrices", Δt=1.0, dbg=None):
        import scipy.sparse as sp
        self._state = state
        om = oracle_model(self.model, self.dis, t=state.time, skip=self.skip, nA=0, A=state.A)
        om.prepare_direct(self.OX, self.OU, 0)
        U = [state.U[d] for d in range(self.OU + 1)] if self.nU else None
        r, L1f, L2f = om.assemble_direct(1 if mission == "matrices" else 0, Δt, Λ, state.X[:self.OX + 1], U,
                                         t=state.time)
        assert r == 0, r
        nX, nU, OX, OU = self.nX, self.nU, self.OX, self.OU
        L1 = {"Λ": [L1f[:nX]], "X": [L1f[nX * (1 + d):nX * (2 + d)] for d in range(OX + 1)],
              "U": [L1f[nX * (OX + 2) + nU * d:nX * (OX + 2) + nU * (d + 1)] for d in range(OU + 1)]}
        cls = ("Λ", "X", "U")
        L2, off = {}, 0
        for a in range(3):
            for b in range(3):
                if a == 0 and b == 0:
                    continue
                cp, rv = om.dpattern[(a, b)]
                n = rv.size
                shape = (nU if a == 2 else nX, nU if b == 2 else nX)
                blocks = []
                for i in range(self.nder[cls[a]]):
                    row = []
                    for j in range(self.nder[cls[b]]):
                        v = L2f[off:off + n]
                        off += n
                        row.append(sp.csc_matrix((v.copy(), rv - 1, cp - 1), shape=shape))
                    blocks.append(row)
                L2[(cls[a], cls[b])] = blocks
        return L1, L2


class HostSweepXAssembly:
    """AssemblySweepX{OX} for models whose elements are all host residual elements
    (muscade_amd.ResidualElement): the product's host evaluation of R, ∂R/∂δX and ∂R/∂δr along the
    trial path, scattered on the CPU with add_value!/add_∂! (src/Assemble.jl:559-608) in element
    order. It lets the product's SweepX loop run on a CPU against the reference's solver goldens
    (test infrastructure only)."""

    def __init__(self, OX, model, dis, device=0):
        import scipy.sparse as sp
        from muscade_amd.assemble import Newmarkβcoefficients
        self._sp = sp
        self.OX, self.model, self.dis = OX, model, dis
        self.nX = model.ndof["X"]
        self.c = Newmarkβcoefficients(OX)
        from muscade_amd import _lib
        self._point = {_lib.MH_KIND_HOLD: "hold", _lib.MH_KIND_DOFLOAD: "load", _lib.MH_KIND_ZERO: "zero"}
        for E in model.eletypes:
            if not getattr(E.cls, "residual_host", False) and E.cls.kind not in self._point:
                raise TypeError(f"{E.key[0]} is neither a host residual element nor Hold, DofLoad or a cost")
        self.Lλ = np.zeros(self.nX)
        self._K = None

    def assemble(self, state, mission, dbg=None, fetch=True):
        nX = self.nX
        rhs = np.zeros(nX)
        rows, cols, vals = [], [], []
        for it, E in enumerate(self.model.eletypes):
            d = self.dis.dis[it]
            s = d.scaleX
            e0 = 0
            point = self._point.get(E.cls.kind) if not getattr(E.cls, "residual_host", False) else None
            for nodes, p in zip(E.nodes, E.params):
                m = nodes.shape[0]
                iX = d.X[e0:e0 + m] - 1
                if point is not None:   # Hold (BasicElements.jl:426-438), DofLoad (:275-284), cost elements
                    e0 += m
                    if point == "zero" or iX.shape[1] == 0:
                        continue
                    if point == "hold":
                        x, lam = state.X[0][iX[:, 0]], state.X[0][iX[:, 1]]
                        g = float(p["value"](state.time, **p["args"])) if "value" in p else 0.
                        R = np.stack([-lam, -(x - g)], axis=1)
                        K = np.broadcast_to(np.array([[0., -1.], [-1., 0.]]), (m, 2, 2))
                    else:
                        R = np.full((m, 1), -float(p["value"](state.time, **p["args"])))
                        K = np.zeros((m, 1, 1))
                    for e in range(m):
                        np.add.at(rhs, iX[e], R[e] * s)
                        rows.append(np.repeat(iX[e], iX.shape[1]))
                        cols.append(np.tile(iX[e], iX.shape[1]))
                        vals.append((K[e] * s[:, None] * s[None, :]).ravel())
                    continue
                Xv = [state.X[k][iX] if k < len(state.X) else np.zeros(iX.shape) for k in range(self.OX + 1)]
                Uv = state.U[0][d.U[e0:e0 + m] - 1] if d.U.shape[1] else np.zeros((m, 0))
                Av = state.A[d.A[e0:e0 + m] - 1] if d.A.shape[1] else np.zeros((m, 0))
                R, K, dR = E.cls.sweepx_blocks(p, Xv, Uv, Av, state.time, self.c, mission, self.OX)
                for e in range(m):
                    np.add.at(rhs, iX[e], R[e] * s)
                    if dR is not None:
                        np.add.at(rhs, iX[e], -(dR[e] * s))
                    rows.append(np.repeat(iX[e], iX.shape[1]))
                    cols.append(np.tile(iX[e], iX.shape[1]))
                    vals.append((K[e] * s[:, None] * s[None, :]).ravel())
                e0 += m
        cat = (lambda a: np.concatenate(a)) if rows else (lambda a: np.zeros(0))
        self._K = self._sp.coo_matrix((cat(vals), (cat(rows).astype(int), cat(cols).astype(int))),
                                      shape=(nX, nX)).tocsc()
        self.Lλ = rhs
        return rhs, self._K.data

    def Lλx(self):
        return self._K.copy()


def allsteps_lagrangian_blocks(model, dis, Λ, X, U, A, time, dt, Assembly, OX=2, OU=0, IA=1):
    """The all-steps DirectXUA Lagrangian and its gradient by the step blocks of `Assembly`
    (AssemblyDirect or the oracle's): the route of solve(DirectXUA)'s Lv with Λgradient,
    L1 + (L2[Λ,·])ᵀΛ per step plus the A column ∂R/∂A, at unit dof scales. L = Σₖ Λₖ·Rₖ + the
    host Lagrangians. Returns (L, {"Λ","X","U" (nstep × n), "A"}), as AllStepsGradient."""
    from muscade_amd.directxua import HostLagrangians, _StepState, step_derivatives, step_pullback
    n, nX, nU, nA = len(time), model.ndof["X"], model.ndof["U"], model.ndof["A"]
    host = HostLagrangians(model, dis, OX, OU, IA)
    out = Assembly(OX, OU, IA, model, dis, skip=tuple(host.skip))
    if host.costed:
        host.gauge_provider = out.gauge_derivatives
    ascaled = [it for it, E in enumerate(model.eletypes) if it not in host.skip and E.params
               and E.params[0].get("amask")]
    Xd, Ud = step_derivatives(X, OX, dt), step_derivatives(U, OU, dt)
    gΛ = np.zeros((n, nX))
    gXd = [np.zeros((n, nX)) for _ in range(OX + 1)]
    gUd = [np.zeros((n, nU)) for _ in range(OU + 1)]
    gA = np.zeros(nA)
    L = 0.
    for k in range(n):
        st = _StepState(model, dis, float(time[k]), [Λ[k]], [x[k] for x in Xd], [u[k] for u in Ud], A,
                        dict(γ=0., iter=1))
        L1d, L2d = out.assemble(st, Λ[k], "matrices", 1.0)
        L1h, _ = host.step(st, Λgradient=True, pieces=False)
        L += float(Λ[k] @ L1d["Λ"][0]) + host.value
        gΛ[k] = L1d["Λ"][0] + L1h["Λ"][0]
        for d in range(OX + 1):
            gXd[d][k] = L1d["X"][d] + L2d[("Λ", "X")][0][d].T @ Λ[k] + L1h["X"][d]
        for d in range(OU + 1):
            if nU:
                gUd[d][k] = L1d["U"][d] + L2d[("Λ", "U")][0][d].T @ Λ[k] + L1h["U"][d]
        if IA:
            gA += L1h["A"][0]
            for it in ascaled:
                D = out.beam_dRdA(it)
                np.add.at(gA, dis.dis[it].A[:, 0] - 1, np.einsum("ei,ei->e", D, Λ[k][dis.dis[it].X - 1]))
    if IA and nA:
        g1, _ = host.once(_StepState(model, dis, float(time[0]), [Λ[0]], [X[0]], [U[0]], A, None))
        gA += g1
        L += host.once_value
    return L, {"Λ": gΛ, "X": step_pullback(gXd, dt), "U": step_pullback(gUd, dt), "A": gA}
