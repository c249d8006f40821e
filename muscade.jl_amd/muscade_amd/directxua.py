"""solve(DirectXUA{OX,OU,IA}; initialstate, time, maxiter, maxΔλ, maxΔx, maxΔu, maxΔa, saveiter) —
the all-steps Newton solver of optimisation FEM, src/DirectXUA.jl:245-513.

Each Newton iteration assembles, for every step of every experiment, the step's gradient
L1[α][αder] and Hessian blocks L2[α,β][αder,βder] over Λ, X, U (and A):

- the residual elements (beams, bars) on the GPU, through AssemblyDirect (residual path,
  src/DirectXUA.jl:85-120);
- the Lagrangian elements on the host with second-order duals (hostcost.Dual2): the
  user-cost elements (DofCost, SingleDofCost, SingleUdof; Acost once), which carry user
  Functors, and Hold and DofLoad, whose Lagrangians are a few products
  (src/DirectXUA.jl:121-171).

These are combined into one sparse system over all steps with the finite-difference kernels
of src/FiniteDifferences.jl for the time derivatives (assemblebig!, :322-357). The sparse LU
of that system stays on the host (UMFPACK `lu` in the reference, SuperLU here). The
increments are applied back through the same kernels (decrementbig!, :358-383).

The unknowns are laid out step after step, [Λ (nX) | X (nX) | U (nU)] per step, then A: the
block order of makepattern (:245-296). The big matrix has Muscade's pattern (BigPattern:
makepattern + SparseTools.prepare), bit for bit.

One deliberate difference. For residual elements (beams, bars: no_second_order), the
reference's step assembly adds L1[Λ] = R and the Λ-row/column Jacobian blocks, but not the
gradient Λᵀ∂R/∂(X,U) (src/DirectXUA.jl:104-119). Its Newton step then returns the whole
multiplier in ΔΛ instead of an increment, and the iterates' Λ keeps drifting. Here that
gradient is added from the same blocks, (L2[Λ,X])ᵀΛ and (L2[Λ,U])ᵀΛ, so Λ converges to the
multiplier. X and U converge to the same KKT point either way. The Hessian keeps the
no_second_order approximation (no Λᵀ∂²R). Hold and DofLoad are evaluated here as
Lagrangians on the host, so the device assembles residual elements only.
"""
import copy
import ctypes
import math
import time as _time

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import _lib
from ._lib import as_f64, dptr, i64ptr, lib
from .assemble import AssemblyDirect
from .hostcost import Dual2
from .model import MuscadeError

CLASSES = ("Λ", "X", "U", "A")

# [order][left | right | centre][points] — src/FiniteDifferences.jl:2-4
_FD = [[[(0, 1.)]],
       [[(0, -1.), (1, 1.)], [(-1, -1.), (0, 1.)], [(-1, -.5), (1, .5)]],
       [[(0, 1.), (1, -2.), (2, 1.)], [(-2, 1.), (-1, -2.), (0, 1.)], [(-1, 1.), (0, -2.), (1, 1.)]]]


def finitediff(order, n, s):
    """(Δs, w) of the finite-difference kernel of `order` at step s of n (1-based) —
    src/FiniteDifferences.jl:8-34."""
    if order > 0 and n < 6:
        raise MuscadeError("Number of steps must be ≥6")
    if order == 0:
        return _FD[0][0]
    return _FD[order][0] if s == 1 else (_FD[order][1] if s == n else _FD[order][2])


def class_patterns(model, dis, OX, OU, IA):
    """The class-pair patterns of prepare(AssemblyDirect{OX,OU,IA}) (DirectXUA.jl:22-56): asmmat!
    (Assemble.jl:373-448) over the dof groups allΛdofs, allXdofs, allUdofs, allAdofs and every
    element type. Returns {(α,β): (colptr, rowval (0-based), shape, nder α, nder β)}, with nder
    0×0 for (Λ,Λ), which is always zero."""
    idx = {"Λ": [d.X for d in dis.dis], "X": [d.X for d in dis.dis], "U": [d.U for d in dis.dis],
           "A": [d.A for d in dis.dis]}
    n = {"Λ": model.ndof["X"], "X": model.ndof["X"], "U": model.ndof["U"], "A": model.ndof["A"]}
    nder = {"Λ": 1, "X": OX + 1, "U": OU + 1, "A": IA}
    pats = {}
    for a in CLASSES:
        for b in CLASSES:
            keys = []
            if n[a] and n[b]:
                for ia, ib in zip(idx[a], idx[b]):
                    if ia.shape[1] and ib.shape[1] and ia.shape[0]:
                        keys.append(((ib[:, None, :] - 1) * n[a] + (ia[:, :, None] - 1)).ravel())
            u = np.unique(np.concatenate(keys)) if keys else np.zeros(0, np.int64)
            col, row = (u // n[a], u % n[a]) if n[a] else (u, u)
            colptr = np.searchsorted(col, np.arange(n[b] + 1)).astype(np.int64)
            na, nb = (0, 0) if a == b == "Λ" else (nder[a], nder[b])
            pats[(a, b)] = (colptr, row.astype(np.int64), (n[a], n[b]), na, nb)
    return pats


class BigPattern:
    """The sparsity of DirectXUA's all-steps matrix as Muscade lays it out: makepattern
    (DirectXUA.jl:245-296) decides which blocks (step, class) × (step, class) are structurally
    nonzero; SparseTools.prepare (SparseTools.jl:32-94) expands each block by its class-pair
    pattern. Block rows are the steps' Λ, X, U blocks of every experiment in turn, then the A
    block; `pgr` holds their first rows (0-based, the reference's Lvasm/Lvdis minus one).
    colptr/rowval are 0-based."""

    def __init__(self, IA, nstep, pats):
        cls3 = CLASSES[:3]
        blk = {}                                   # (block row, block column) → class pair, first appearance
        cum = 0
        for ns in nstep:
            for istep in range(1, ns + 1):
                for ia, a in enumerate(cls3):
                    for ib, b in enumerate(cls3):
                        _, _, _, na, nb = pats[(a, b)]
                        for ad in range(na):
                            for bd in range(nb):
                                for Δa, _ in finitediff(ad, ns, istep):
                                    for Δb, _ in finitediff(bd, ns, istep):
                                        blk.setdefault((cum + 3 * (istep + Δa - 1) + ia, cum + 3 * (istep + Δb - 1) + ib),
                                                       (a, b))
            cum += 3 * ns
        nblock = 3 * sum(nstep) + (1 if IA == 1 else 0)
        if IA == 1:
            Ab = nblock - 1
            blk.setdefault((Ab, Ab), ("A", "A"))
            cum = 0
            for ns in nstep:
                for istep in range(1, ns + 1):
                    for ia, a in enumerate(cls3):
                        if pats[("A", a)][3] > 0:
                            blk.setdefault((Ab, cum + 3 * (istep - 1) + ia), ("A", a))
                            blk.setdefault((cum + 3 * (istep - 1) + ia, Ab), (a, "A"))
                cum += 3 * ns
        size = {c: pats[(c, c)][2][0] for c in CLASSES}
        nl = np.full(nblock, -1, np.int64)
        for (r, c), (a, b) in blk.items():
            nl[r] = size[a]
            nl[c] = size[b]
        if (nl < 0).any():
            raise MuscadeError("invalid sparse-of-sparse pattern: a block row or column has only empty blocks")
        self.pgr = np.concatenate([[0], np.cumsum(nl)]).astype(np.int64)
        self.nv = int(self.pgr[-1])
        # SparseTools.prepare: global column by global column, the blocks of the block column in
        # increasing block row, each block's entries of that local column in its rowval order
        bycol = {}
        for (r, c), pair in blk.items():
            bycol.setdefault(c, []).append((r, pair))
        rows, cols = [], []
        for c in range(nblock):
            entries = sorted(bycol.get(c, []))
            lr, lc, rk = [], [], []
            for k, (r, pair) in enumerate(entries):
                cp, rv, _, _, _ = pats[pair]
                lc.append(np.repeat(np.arange(cp.size - 1), np.diff(cp)))
                lr.append(rv + self.pgr[r])
                rk.append(np.full(rv.size, k))
            if not lr:
                continue
            lr, lc, rk = np.concatenate(lr), np.concatenate(lc), np.concatenate(rk)
            o = np.lexsort((rk, lc))
            rows.append(lr[o])
            cols.append(lc[o] + self.pgr[c])
        self.rowval = np.concatenate(rows).astype(np.int64) if rows else np.zeros(0, np.int64)
        col = np.concatenate(cols).astype(np.int64) if cols else np.zeros(0, np.int64)
        self.colptr = np.searchsorted(col, np.arange(self.nv + 1)).astype(np.int64)
        self._keys = col * self.nv + self.rowval

    @property
    def nnz(self):
        return int(self.rowval.size)

    def locate(self, rows, cols):
        """nz numbers of the entries (rows, cols); every entry must be in the pattern."""
        k = np.asarray(cols, np.int64) * self.nv + np.asarray(rows, np.int64)
        pos = np.searchsorted(self._keys, k)
        if k.size and (pos.max() >= self._keys.size or not np.array_equal(self._keys[pos], k)):
            raise MuscadeError("a contribution falls outside the all-steps pattern")
        return pos

    def coords(self):
        """(rows, cols) of every nz, in nz order."""
        return self.rowval, np.repeat(np.arange(self.nv), np.diff(self.colptr))

    def matrix(self, rows, cols, vals):
        """The matrix with this pattern holding the sum of the contributions (rows, cols, vals),
        each nz summed in contribution order."""
        pos = self.locate(rows, cols)
        nz = np.bincount(pos, weights=vals, minlength=self.nnz) if pos.size else np.zeros(self.nnz)
        return sp.csc_matrix((nz, self.rowval.copy(), self.colptr.copy()), shape=(self.nv, self.nv))


def sparser(M, rtol=1e-20):
    """sparser!(cLvv, Lvv, rtol) (SparseTools.jl:192-196, called with 1e-20 by DirectXUA.jl:486):
    drop the stored entries below rtol·max|M|, before the LU."""
    M = M.copy()
    atol = rtol * np.abs(M.data).max() if M.nnz else 0.
    if atol > 0.:
        M.data[np.abs(M.data) < atol] = 0.
        M.eliminate_zeros()
    return M


class HostLagrangians:
    """Gradient and Hessian of the host-evaluated elements' Lagrangians, in the layout of
    DirectXUA_lagrangian_addition! (src/DirectXUA.jl:121-150): partials Λ | X⁽⁰..OX⁾ | U⁽⁰..OU⁾
    | A (IA = 1), each scaled like revariate with the element type's dof scales."""

    def __init__(self, model, dis, OX, OU, IA, gauge_provider=None):
        """gauge_provider(it, state) → values (nele,4), gradients (nele,4,12), Hessians
        (nele,4,12,12) of the gauge requestables of element type it, for the ElementCost
        types (their wrapped beams stay on the device: `skip` lists only the host types)."""
        self.model, self.dis = model, dis
        self.OX, self.OU, self.IA = OX, OU, IA
        self.nder = {"Λ": 1, "X": OX + 1, "U": OU + 1, "A": IA}
        self.ndof = {"Λ": model.ndof["X"], "X": model.ndof["X"], "U": model.ndof["U"], "A": model.ndof["A"]}
        self.skip = [it for it, E in enumerate(model.eletypes)
                     if getattr(E.cls, "host", False) or getattr(E.cls, "direct_host", False)]
        self.costed = [it for it, E in enumerate(model.eletypes) if getattr(E.cls, "host_cost", False)]
        self.types = sorted(self.skip + self.costed)
        self.gauge_provider = gauge_provider

    def _once_off(self, E):
        """Acost types: assembled once by assembleA! (Assemble.jl:507-519)."""
        return bool(getattr(E.cls, "once_off", False))

    def _chunks(self, it):
        E = self.model.eletypes[it]
        d = self.dis.dis[it]
        e0 = 0
        for nodes, p in zip(E.nodes, E.params):
            m = nodes.shape[0]
            yield p, d, slice(e0, e0 + m)
            e0 += m

    def step(self, state, Λgradient=True, pieces=True):
        """Per-step contributions: L1[α][der] dense vectors, and a list of COO pieces
        (α, i, β, j, rows, cols, vals) of L2 (0-based model dofs; none with pieces=False).
        Λgradient=False: host residual elements with no_second_order add only L1[Λ] = R to the
        gradient, as the reference's residual path does (DirectXUA.jl:104-119). With
        pieces=False the partials are first order (no Hessians are formed). self.value is left
        holding the step's summed Lagrangian value."""
        L1 = {c: [np.zeros(self.ndof[c]) for _ in range(self.nder[c])] for c in CLASSES}
        L2 = []
        value = 0.
        for it in self.types:
            E = self.model.eletypes[it]
            once = self._once_off(E)
            gd = None
            if it in self.costed:
                if self.gauge_provider is None:
                    raise MuscadeError(f"element type {it + 1} ({E.key[0]}) needs the device requestables")
                gd = self.gauge_provider(it, state)
            for p, d, sl in self._chunks(it):
                if gd is not None:
                    p = dict(p, _gd=tuple(None if a is None else a[sl] for a in gd))
                iX, iU, iA = d.X[sl] - 1, d.U[sl] - 1, d.A[sl] - 1
                m, Nx, Nu, Na = iX.shape[0], iX.shape[1], iU.shape[1], iA.shape[1]
                # partial layout and global dofs of each (class, derivative)
                layout, slot = [], 0
                for c, idx, n in (("Λ", iX, Nx), ("X", iX, Nx), ("U", iU, Nu), ("A", iA, Na)):
                    for der in range(self.nder[c] if c != "A" else self.IA):
                        layout.append((c, der, slot, idx))
                        slot += n
                npar = slot

                def seeded(c, der, idx, vals, scale):
                    for (cc, dd, s0, _) in layout:
                        if cc == c and dd == der:
                            return [Dual2.seed(vals[:, k], s0 + k, npar, scale[k], first_order=not pieces)
                                    for k in range(idx.shape[1])]
                    return [vals[:, k] for k in range(idx.shape[1])]

                Λ = seeded("Λ", 0, iX, state.Λ[0][iX], d.scaleΛ)
                X = [seeded("X", der, iX, state.X[der][iX], d.scaleX) for der in range(self.OX + 1)]
                U = [seeded("U", der, iU, state.U[der][iU], d.scaleU) for der in range(self.OU + 1)]
                A = seeded("A", 0, iA, state.A[iA], d.scaleA)
                if once:
                    # The per-step element loop reaches Acost elements too: its Acost method
                    # (Assemble.jl:487) is written for one element, but assemble! passes it the vector of
                    # a type's elements, so they take the generic Lagrangian path (DirectXUA.jl:152-171)
                    # at every step, with scaled partials, besides assembleA!'s once-off addition.
                    L = E.cls.acost(p, A)
                elif getattr(E.cls, "uses_SP", False):
                    L = E.cls.lagrangian(p, Λ, X, U, A, state.time, SP=state.SP)
                else:
                    L = E.cls.lagrangian(p, Λ, X, U, A, state.time)
                if not isinstance(L, Dual2):
                    value += float(np.sum(L))
                    continue
                if not np.all(np.isfinite(L.g)) or (L.h is not None and not np.all(np.isfinite(L.h))):
                    raise MuscadeError(f"lagrangian({E.key[0]},...) returned NaN",
                                       dict(ieletyp=it + 1, t=state.time))
                value += float(np.sum(L.v))
                rpath = (not Λgradient and getattr(E.cls, "residual_host", False) and E.cls.no_second_order)
                for (a, i, sa, ia) in layout:
                    na = ia.shape[1]
                    if not rpath or a == "Λ":
                        np.add.at(L1[a][i], ia, L.g[:, sa:sa + na])
                    if not pieces:
                        continue
                    for (b, j, sb, ib) in layout:
                        if a == "Λ" and b == "Λ":
                            continue
                        nb = ib.shape[1]
                        H = L.h[:, sa:sa + na, sb:sb + nb]
                        rows = np.broadcast_to(ia[:, :, None], (m, na, nb)).ravel()
                        cols = np.broadcast_to(ib[:, None, :], (m, na, nb)).ravel()
                        L2.append((a, i, b, j, rows, cols, H.ravel()))
        self.value = value
        return L1, L2

    def once(self, state):
        """Once-off A costs (assembleA!, src/Assemble.jl:507-519; Acost addin!, src/DirectXUA.jl:66-78):
        gradient and Hessian over the unscaled A-dofs."""
        gA = np.zeros(self.ndof["A"])
        pieces = []
        self.once_value = 0.
        for it in self.types:
            E = self.model.eletypes[it]
            if not self._once_off(E):
                continue
            for p, d, sl in self._chunks(it):
                iA = d.A[sl] - 1
                m, Na = iA.shape
                A = [Dual2.seed(state.A[iA[:, k]], k, Na) for k in range(Na)]
                C = E.cls.acost(p, A)
                if not isinstance(C, Dual2):
                    self.once_value += float(np.sum(C))
                    continue
                self.once_value += float(np.sum(C.v))
                np.add.at(gA, iA, C.g)
                rows = np.broadcast_to(iA[:, :, None], (m, Na, Na)).ravel()
                cols = np.broadcast_to(iA[:, None, :], (m, Na, Na)).ravel()
                pieces.append((rows, cols, C.h.ravel()))
        return gA, pieces


def _add_beam_dRdA(D, d, Λ, L1, L2, Λgradient):
    """The A column of the residual path for an A-scaled beam type (R reads A, as the
    reference's test Spring does, test/SomeElements.jl:162-186): L2[Λ,A] = ∂R/∂A·sΛ·sA and its
    transpose L2[A,Λ], and, with Λgradient, L1[A] += (L2[Λ,A])ᵀΛ as for X and U. D: ∂R/∂a of
    the type's elements at this step (nele × 12, unscaled; mh_beam_dRdA)."""
    rows = d.X - 1                                          # the Λ dofs of the element (its X dofs)
    col = np.repeat(d.A[:, 0] - 1, rows.shape[1])
    vals = (D * d.scaleΛ[None, :] * d.scaleA[0]).ravel()
    L2.append(("Λ", 0, "A", 0, rows.ravel(), col, vals))
    L2.append(("A", 0, "Λ", 0, col, rows.ravel(), vals))
    if Λgradient:
        np.add.at(L1["A"][0], col, vals * Λ[rows.ravel()])


class _Replica:
    """n disjoint copies of a model's device element types, numbered copy after copy: copy s
    owns nodes s·nnod.., X-dofs s·nX.. and U-dofs s·nU.. (just what upload_model reads)."""

    class _Type:
        def __init__(self, E, n, nnod):
            self.cls, self.key = E.cls, E.key
            self._nodes = np.concatenate([E.cat_nodes() + s * nnod for s in range(n)])
            self.params = [p for s in range(n) for p in E.params]
            self.nodes = [q for s in range(n) for q in E.nodes]
            self.nele = E.nele * n

        def cat_nodes(self):
            return self._nodes

    class _Dis:
        def __init__(self, d, n, nX, nU, shareA):
            self.X = np.concatenate([d.X + s * nX for s in range(n)])
            self.U = np.concatenate([d.U + s * nU for s in range(n)])
            # A-dofs are the model's, shared by every copy (one A for all steps)
            self.A = np.tile(d.A, (n, 1)) if shareA else np.zeros((self.X.shape[0], 0), np.int64)
            self.scaleX, self.scaleU, self.scaleA, self.scaleΛ = d.scaleX, d.scaleU, d.scaleA, d.scaleΛ

    def __init__(self, model, dis, types, n):
        nX, nU = model.ndof["X"], model.ndof["U"]
        shareA = any(model.eletypes[it].params and model.eletypes[it].params[0].get("amask") for it in types)
        self.nnod = model.nnod * n
        self._xyz = np.tile(model.coords(), (n, 1))
        self.ndof = {"X": nX * n, "U": nU * n, "A": model.ndof["A"] if shareA else 0}
        self.eletypes = [self._Type(model.eletypes[it], n, model.nnod) for it in types]
        self.dis = [self._Dis(dis.dis[it], n, nX, nU, shareA) for it in types]

    def coords(self):
        return self._xyz


class StepBatchAssembly:
    """The residual elements of all steps of one experiment in one device launch: AssemblyDirect
    on `nstep` disjoint copies of them (copy s carries step s's state), whose block-diagonal
    step blocks are cut back out. Each step block is bitwise what a launch of that step alone
    gives: the same element values, summed in the same element order."""

    def __init__(self, OX, OU, IA, model, dis, nstep, skip=(), device=0):
        types = [it for it in range(len(model.eletypes)) if it not in skip]
        for it in types:
            if model.eletypes[it].cls.kind not in (1, 2):   # beams only: bars read the time
                raise MuscadeError("step batching needs time-independent residual elements (beams)")
        self.rep = _Replica(model, dis, types, nstep)

        class _RDis:
            pass
        rdis = _RDis()
        rdis.dis = self.rep.dis
        self.nstep, self.nX, self.nU = nstep, model.ndof["X"], model.ndof["U"]
        self.OX, self.OU = OX, OU
        self.rindex = {it: k for k, it in enumerate(types)}
        self.out = AssemblyDirect(OX, OU, IA, self.rep, rdis, device=device, skip=())

    def assemble_steps(self, states, dbg=None):
        """L1, L2 of every step (lists over steps, as AssemblyDirect.assemble returns them)."""
        class _S:
            pass
        st = _S()
        st.time = states[0].time
        st.X = [np.concatenate([s.X[d] for s in states]) for d in range(self.OX + 1)]
        st.U = [np.concatenate([s.U[d] for s in states]) for d in range(self.OU + 1)]
        st.A = states[0].A
        lam = np.concatenate([s.Λ[0] for s in states])
        L1, L2 = self.out.assemble(st, lam, "matrices", 1.0, dbg=dbg)
        n = {"Λ": self.nX, "X": self.nX, "U": self.nU}
        res = []
        for k in range(self.nstep):
            l1 = {c: [v[k * n[c]:(k + 1) * n[c]] for v in vs] for c, vs in L1.items()}
            l2 = {}
            for (a, b), blocks in L2.items():
                ra, rb = slice(k * n[a], (k + 1) * n[a]), slice(k * n[b], (k + 1) * n[b])
                l2[(a, b)] = [[M[ra, rb] for M in row] for row in blocks]
            res.append((l1, l2))
        return res

    def beam_dRdA(self, it):
        """∂R/∂a of the A-scaled beam type it (a model index) at every step of the last
        `assemble_steps`: a list over steps of nele × 12 arrays (mh_beam_dRdA)."""
        D = self.out.beam_dRdA(self.rindex[it])
        return np.split(D, self.nstep)


class DeviceBigSystem:
    """DirectXUA's all-steps system assembled on the device (makepattern / preparebig /
    assemblebig!, src/DirectXUA.jl:245-383; SparseTools.jl:32-137) from the per-step blocks that
    a StepBatchAssembly leaves in its context, plus the host elements' pieces. The contribution
    list (row, column, source, finite-difference weight) is built once here and sorted on the
    device into the CSC pattern (mh_big_prepare); each iteration sums the contributions of every
    nz on the device (mh_big_assemble) and the host receives the finished values."""

    def __init__(self, batched, nv, block, dt, hcontrib, pattern=None):
        """pattern (a BigPattern): every one of its entries is added as a structural contribution
        of weight 0 after the real ones, so that the device pattern is Muscade's
        (makepattern/preparebig) and each nz's sum is unchanged."""
        out = batched.out
        self.ctx, self.nv = out.ctx, nv
        OX, nX, nU, n = batched.OX, batched.nX, batched.nU, batched.nstep
        cpXX, rvXX, _ = out.patterns[0]
        cpXU, rvXU, _ = out.patterns[1]
        cpUX, rvUX, _ = out.patterns[2]

        def coo(cp, rv):
            col = np.repeat(np.arange(cp.size - 1), np.diff(cp))
            return rv - 1, col
        R0, C0 = coo(cpXX, rvXX)
        nxx, nxu = R0.size, rvXU.size
        ndX = OX + 1
        rows, cols, src, w = [], [], [], []
        step = R0 // max(nX, 1)
        r0, c0 = R0 - step * nX, C0 - step * nX
        for s_ in range(n):
            k = np.nonzero(step == s_)[0]
            if k.size == 0:
                continue
            for d in range(ndX):
                sc = dt ** (-d)
                for Δ, wt in finitediff(d, n, s_ + 1):
                    # L2[Λ,X][1,d]: rows Λ of step s, columns X of step s+Δ
                    rows.append(block(s_ + 1, "Λ") + r0[k]); cols.append(block(s_ + 1 + Δ, "X") + c0[k])
                    src.append(d * nxx + k); w.append(np.full(k.size, wt * sc))
                    # L2[X,Λ][d,1]
                    rows.append(block(s_ + 1 + Δ, "X") + r0[k]); cols.append(block(s_ + 1, "Λ") + c0[k])
                    src.append((ndX + d) * nxx + k); w.append(np.full(k.size, wt * sc))
        if nU and nxu:
            R1, C1 = coo(cpXU, rvXU)
            st1 = R1 // nX
            rows.append(np.array([block(q + 1, "Λ") for q in st1], np.int64) + R1 - st1 * nX)
            cols.append(np.array([block(q + 1, "U") for q in st1], np.int64) + C1 - st1 * nU)
            src.append(2 * ndX * nxx + np.arange(nxu)); w.append(np.ones(nxu))
            R2, C2 = coo(cpUX, rvUX)
            st2 = R2 // nU
            rows.append(np.array([block(q + 1, "U") for q in st2], np.int64) + R2 - st2 * nU)
            cols.append(np.array([block(q + 1, "Λ") for q in st2], np.int64) + C2 - st2 * nX)
            src.append(2 * ndX * nxx + nxu + np.arange(R2.size)); w.append(np.ones(R2.size))
        ndev = 2 * ndX * nxx + 2 * nxu
        self.hsizes = [v.size for (_, _, _, v) in hcontrib]
        o = ndev
        for (r, c, wh, v) in hcontrib:
            rows.append(np.asarray(r, np.int64)); cols.append(np.asarray(c, np.int64))
            src.append(o + np.arange(v.size)); w.append(np.asarray(wh, float))
            o += v.size
        if pattern is not None:
            pr, pc = pattern.coords()
            rows.append(pr); cols.append(pc)
            src.append(np.full(pr.size, o, np.int64)); w.append(np.zeros(pr.size))
            o += 1                                   # one host value, always 0, for them all
        self.structural = pattern is not None
        cat = lambda a, t: np.ascontiguousarray(np.concatenate(a) if a else np.zeros(0), dtype=t)   # noqa: E731
        rows, cols, src, w = cat(rows, np.int64), cat(cols, np.int64), cat(src, np.int64), cat(w, np.float64)
        self.nhost = o - ndev
        nnz = ctypes.c_int64()
        self.ctx.check(lib.mh_big_prepare(self.ctx.h, nv, rows.size, i64ptr(rows), i64ptr(cols), i64ptr(src), dptr(w),
                                          self.nhost, ctypes.byref(nnz)), "big prepare")
        self.colptr = np.zeros(nv + 1, np.int64)
        self.rowval = np.zeros(nnz.value, np.int64)
        self.ctx.check(lib.mh_big_get_pattern(self.ctx.h, i64ptr(self.colptr), i64ptr(self.rowval)), "big pattern")
        self.colptr -= 1
        self.rowval -= 1
        self.ncontrib = rows.size

    def matches(self, hcontrib):
        return [v.size for (_, _, _, v) in hcontrib] == self.hsizes

    def _host_values(self, hcontrib):
        v = [v for (_, _, _, v) in hcontrib] + ([np.zeros(1)] if self.structural else [])
        return np.ascontiguousarray(np.concatenate(v) if v else np.zeros(0), dtype=np.float64)

    def prepare_solve(self, per_step, nstep, na):
        """Lay the system out for the device solve (mh_big_solve_prepare): dense diagonal blocks
        of the steps that finite differences couple, block tridiagonal, bordered by the A-dofs.
        Returns (block order, block count); MuscadeError when they do not fit in device memory."""
        m, nb = ctypes.c_int64(), ctypes.c_int64()
        self.ctx.check(lib.mh_big_solve_prepare(self.ctx.h, per_step, nstep, na, ctypes.byref(m), ctypes.byref(nb)),
                       "big solve prepare")
        return m.value, nb.value

    def solve(self, hcontrib, rhs, dbg=None):
        """Assemble the system on the device and solve Lvv·Δv = Lv there (mh_big_solve): only the
        right-hand side goes up and only Δv comes back."""
        hv = self._host_values(hcontrib)
        self.ctx.check(lib.mh_big_assemble(self.ctx.h, dptr(hv) if hv.size else None, None), "big assemble")
        x = np.zeros(self.nv)
        r = lib.mh_big_solve(self.ctx.h, dptr(as_f64(rhs)), dptr(x))
        if r == _lib.MH_ERR_SOLVE:
            raise MuscadeError(f"Lvv matrix factorization failed: {lib.mh_last_error(self.ctx.h).decode()}", dbg)
        self.ctx.check(r, "big solve")
        return x

    def assemble(self, hcontrib):
        hv = self._host_values(hcontrib)
        nz = np.zeros(self.rowval.size)
        self.ctx.check(lib.mh_big_assemble(self.ctx.h, dptr(hv) if hv.size else None, dptr(nz)), "big assemble")
        M = sp.csc_matrix((nz, self.rowval.copy(), self.colptr.copy()), shape=(self.nv, self.nv))
        return sparser(M)   # as the host path: sparser!(cLvv, Lvv, 1e-20) before the LU


def step_derivatives(V, order, Δt):
    """V⁽ᵈ⁾ of every step for d = 0..order (V: nstep × n), by the finite-difference kernels of
    src/FiniteDifferences.jl:8-34 (the derivatives DirectXUA's steps hold)."""
    n = V.shape[0]
    out = [V]
    for d in range(1, order + 1):
        W = np.zeros_like(V)
        s = Δt ** (-d)
        for k in range(1, n + 1):
            for Δs, w in finitediff(d, n, k):
                W[k - 1] += (w * s) * V[k - 1 + Δs]
        out.append(W)
    return out


def step_pullback(G, Δt):
    """The transpose of step_derivatives: Σ_d of the derivative-level gradients G[d] (nstep × n)
    mapped onto the steps, as assemblebig! adds L1 into Lv (src/DirectXUA.jl:336-357)."""
    n = G[0].shape[0]
    out = np.zeros_like(G[0])
    for d, Gd in enumerate(G):
        s = Δt ** (-d)
        for k in range(1, n + 1):
            for Δs, w in finitediff(d, n, k):
                out[k - 1 + Δs] += (w * s) * Gd[k - 1]
    return out


class _StepState:
    """The per-step state HostLagrangians reads (model, dis, time, Λ, X, U, A, SP)."""

    def __init__(self, model, dis, time, Λ, X, U, A, SP):
        self.model, self.dis, self.time = model, dis, time
        self.Λ, self.X, self.U, self.A, self.SP = Λ, X, U, A, SP


class AllStepsGradient:
    """The gradient of DirectXUA{OX,OU,IA}'s all-steps Lagrangian at given steps, by the device
    adjoint kernels, without the all-steps matrix:

        L = Σₖ [Λₖ·Rₖ(Xₖ, X′ₖ, X″ₖ, Uₖ, A) + host Lagrangians of step k] (+ once-off A costs),

    with X′, X″ (and U′, U″) the finite differences of the steps, as assemblebig! applies them
    (src/DirectXUA.jl:336-383, src/FiniteDifferences.jl:2-34). The residual elements (beams)
    of all steps are disjoint copies in one SweepX context whose A-dofs are shared: one
    assembly gives every Rₖ, one mh_vjp_lagrangian every Λₖ·∂Rₖ/∂X⁽ᵈ⁾ and Λₖ·∂Rₖ/∂Uₖ, and
    Σₖ Λₖ·∂Rₖ/∂A, recomputing each element's internal state instead of storing it. The host
    elements (constraints, loads, costs, gauge costs, A priors) add their gradients from
    HostLagrangians. The derivative-level gradients are mapped onto the steps' Xₖ through the
    finite-difference weights, the chain rule assemblebig! applies to Lv.

    Returns unscaled gradients: at unit dof scales, the all-steps Lv of solve(DirectXUA) with
    Λgradient=True, class by class. There is no counterpart in the reference (which forms Lv
    with forward-mode duals in its Newton loop)."""

    def __init__(self, OX, OU, IA, model, dis, nstep, Δt, device=0):
        from .assemble import AssemblySweepX
        self.OX, self.OU, self.IA = OX, OU, IA
        self.model, self.dis, self.nstep, self.Δt = model, dis, int(nstep), float(Δt)
        self.nX, self.nU, self.nA = model.ndof["X"], model.ndof["U"], model.ndof["A"]
        if OX > 0 and self.nstep < 6:
            raise MuscadeError("Number of steps must be ≥6")
        self.host = HostLagrangians(model, dis, OX, OU, IA)
        types = [it for it in range(len(model.eletypes)) if it not in self.host.skip]
        for it in types:
            if model.eletypes[it].cls.kind not in (_lib.MH_KIND_BEAM, _lib.MH_KIND_BEAM_U):
                raise MuscadeError("the all-steps adjoint takes beams as its residual elements")
        self.rep = _Replica(model, dis, types, self.nstep)
        self._rindex = {it: k for k, it in enumerate(types)}

        class _RDis:
            pass
        rdis = _RDis()
        rdis.dis = self.rep.dis
        self.out = AssemblySweepX(OX, self.rep, rdis, device=device)
        # the gauge requestables' derivatives of every step in one launch on the copies (X of all
        # steps is on the device after the assembly), handed to the host Lagrangians step by step
        self._gauge, self._step = {}, 0
        if self.host.costed:
            self.host.gauge_provider = lambda it, state: tuple(a[self._step] for a in self._gauge[it])
        self.timings = {}

    def __call__(self, Λ, X, U, A, time):
        """Λ, X (nstep × nX), U (nstep × nU): each step's multipliers, X and U; A (nA); time
        (nstep). Returns (L, dict of ∂L/∂Λ, ∂L/∂X, ∂L/∂U (nstep × n each) and ∂L/∂A (nA))."""
        clock = _time.perf_counter
        t0 = clock()
        n, nX, nU, dis = self.nstep, self.nX, self.nU, self.dis
        Λ, X = np.asarray(Λ, float).reshape(n, nX), np.asarray(X, float).reshape(n, nX)
        U = np.asarray(U, float).reshape(n, nU)
        A = np.asarray(A, float)
        Xd = step_derivatives(X, self.OX, self.Δt)
        Ud = step_derivatives(U, self.OU, self.Δt)
        st = _StepState(self.rep, None, float(time[0]), [Λ.ravel()], [x.ravel() for x in Xd], [Ud[0].ravel()], A, None)
        self.out.assemble(st, "iter", fetch=False)
        R = np.zeros(n * nX)
        self.out.ctx.check(lib.mh_get_rhs(self.out.ctx.h, dptr(R)), "rhs")
        gXd, gU0, gA = self.out.vjp(st, st.Λ[0], self.OX, upload=False)
        if self.host.costed:
            from .results import gauge_derivatives_on
            for it in self.host.costed:
                m = self.model.eletypes[it].nele
                # the Hessians are read only by costs built with exact=True
                hess = any(p.get("exact", False) for p, _, _ in self.host._chunks(it))
                v, g_, H = gauge_derivatives_on(self.out.ctx, self.out.groups[self._rindex[it]], None, n * m,
                                                upload=False, hessians=hess)
                self._gauge[it] = (v.reshape(n, m, 4), g_.reshape(n, m, 4, 12),
                                   H.reshape(n, m, 4, 12, 12) if hess else [None] * n)
        t1 = clock()
        R = R.reshape(n, nX)
        L = float(np.sum(Λ * R))
        gΛ = R.copy()
        gXd = [g.reshape(n, nX).copy() for g in gXd]
        gUd = [gU0.reshape(n, nU).copy()] + [np.zeros((n, nU)) for _ in range(self.OU)]
        gA = gA.copy()
        sc = {"Λ": dis.scaleΛ, "X": dis.scaleX, "U": dis.scaleU, "A": dis.scaleA}
        for k in range(n):
            sk = _StepState(self.model, dis, float(time[k]), [Λ[k]], [x[k] for x in Xd], [u[k] for u in Ud], A,
                            dict(γ=0., iter=1))
            self._step = k
            L1, _ = self.host.step(sk, Λgradient=True, pieces=False)
            L += self.host.value
            gΛ[k] += L1["Λ"][0] / sc["Λ"]
            for d in range(self.OX + 1):
                gXd[d][k] += L1["X"][d] / sc["X"]
            for d in range(self.OU + 1):
                if nU:
                    gUd[d][k] += L1["U"][d] / sc["U"]
            if self.IA and self.nA:
                gA += L1["A"][0] / sc["A"]
        if self.IA and self.nA:
            g1, _ = self.host.once(_StepState(self.model, dis, float(time[0]), [Λ[0]], [X[0]], [U[0]], A, None))
            gA += g1
            L += self.host.once_value
        grad = {"Λ": gΛ, "X": step_pullback(gXd, self.Δt), "U": step_pullback(gUd, self.Δt) if nU else np.zeros((n, 0)),
                "A": gA}
        self.timings = {"device_s": t1 - t0, "host_s": clock() - t1}
        return L, grad


class _NoDeviceKinds:
    """The per-step assembler of a model with no device element kind: nothing to add."""

    def assemble(self, state, Λ, mission="matrices", Δt=1.0, dbg=None):
        return {}, {}


class _DirectXUAMeta(type):
    def __getitem__(cls, key):
        OX, OU, IA = key
        if OX not in (0, 1, 2) or OU not in (0, 1, 2) or IA not in (0, 1):
            raise MuscadeError("DirectXUA{OX,OU,IA} needs OX, OU in 0..2 and IA in 0..1")
        return type(f"DirectXUA{{{OX},{OU},{IA}}}", (DirectXUA,), {"OX": OX, "OU": OU, "IA": IA})


class DirectXUA(metaclass=_DirectXUAMeta):
    """DirectXUA{OX,OU,IA}: DirectXUA[OX, OU, IA] (src/DirectXUA.jl:385-437)."""
    OX, OU, IA = 0, 0, 0

    @classmethod
    def run(cls, pstate, verbose, dbg, *, time, initialstate, maxiter=50, maxΔλ=1e-5, maxΔx=1e-5, maxΔu=1e-5,
            maxΔa=1e-5, saveiter=False, device=0, Assembly=AssemblyDirect, batch_steps=False, big="host", linsolve="host",
            stats=None, Λgradient=True):
        """`Assembly` is the per-step assembler of the device element kinds (prepare(AssemblyDirect{OX,OU,IA}),
        src/DirectXUA.jl:22-56). batch_steps: all steps of an experiment in one device launch per
        Newton iteration (StepBatchAssembly; beams only). big: "host" assembles the all-steps
        system on the host (assemblebig!, DirectXUA.jl:336-383), "device" on the device from the
        step-batched blocks (DeviceBigSystem; batch_steps and one experiment). linsolve: "host"
        factorizes the system with SciPy's sparse LU (lu(Lvv), DirectXUA.jl:489-490), "device"
        factorizes it where it was assembled (big="device"; DeviceBigSystem.solve, a dense block
        LU of the steps). stats (a dict) receives the iteration count and the seconds spent in
        device assembly, host assembly and the linear solve (lu_s).

        Λgradient: add Λᵀ∂R/∂(X,U) of the residual-path elements (the device kinds, and host
        residual elements with no_second_order) to the gradient, so that Λ converges to the
        multiplier (the default). Λgradient=False leaves it out, as DirectXUA.jl:104-119 does: the
        reference's own iterates."""
        clock = _time.perf_counter
        st_ = stats if stats is not None else {}
        st_.update(iterations=0, device_assembly_s=0., host_assembly_s=0., lu_s=0.)
        OX, OU, IA = cls.OX, cls.OU, cls.IA
        nexp = len(time)
        if len(initialstate) != nexp:
            raise MuscadeError("initialstate and time must be of the same length")
        nstep = [len(t) for t in time]
        Δt = [float(t[1] - t[0]) if len(t) > 1 else 1.0 for t in time]
        model, dis = initialstate[0].model, initialstate[0].dis
        nX, nU, nA = model.ndof["X"], model.ndof["U"], model.ndof["A"]
        nder = {"Λ": 1, "X": OX + 1, "U": OU + 1, "A": IA}
        classes = CLASSES if IA else CLASSES[:3]
        maxΔ2 = {"Λ": maxΔλ ** 2, "X": maxΔx ** 2, "U": maxΔu ** 2, "A": maxΔa ** 2}
        scale = {"Λ": dis.scaleΛ, "X": dis.scaleX, "U": dis.scaleU, "A": dis.scaleA}
        # state storage: state[iexp][istep], all sharing one A
        A = initialstate[0].A.copy()
        state = []
        for iexp, s0 in enumerate(initialstate):
            steps = []
            for t in time[iexp]:
                s = s0.copy(time=float(t)).with_ders(1, OX + 1, OU + 1)
                s.A = A
                steps.append(s)
            state.append(steps)
        stateiter = [] if saveiter else None
        pstate[0] = stateiter if saveiter else state

        if big not in ("host", "device"):
            raise MuscadeError(f"big must be 'host' or 'device', not {big!r}")
        if big == "device" and (not batch_steps or nexp != 1):
            raise MuscadeError("big='device' assembles the step-batched blocks of one experiment (batch_steps=True)")
        if linsolve not in ("host", "device"):
            raise MuscadeError(f"linsolve must be 'host' or 'device', not {linsolve!r}")
        if linsolve == "device" and big != "device":
            raise MuscadeError("linsolve='device' solves the system where it is assembled (big='device')")
        host = HostLagrangians(model, dis, OX, OU, IA)
        ascaled = [it for it, E in enumerate(model.eletypes) if it not in host.skip and E.params
                   and E.params[0].get("amask")]
        if batch_steps:
            batched = [StepBatchAssembly(OX, OU, IA, model, dis, n, skip=tuple(host.skip), device=device)
                       for n in nstep]
            provider = None
        else:
            if Assembly is AssemblyDirect and len(host.skip) == len(model.eletypes):
                out = _NoDeviceKinds()          # every element is evaluated on the host
            else:
                out = Assembly(OX, OU, IA, model, dis, device=device, skip=tuple(host.skip))
            provider = getattr(out, "gauge_derivatives", None)
        if host.costed:
            if provider is None:
                from .results import ElementResults
                provider = ElementResults(model, dis, device).gauge_derivatives
            host.gauge_provider = provider
        per_step = 2 * nX + nU
        exp0 = np.cumsum([0] + [n * per_step for n in nstep])
        nv = int(exp0[-1]) + (nA if IA else 0)   # A-dofs are unknowns only when IA = 1
        Aoff = int(exp0[-1])
        # Muscade's all-steps pattern (makepattern + preparebig, DirectXUA.jl:245-304)
        bigpat = BigPattern(IA, nstep, class_patterns(model, dis, OX, OU, IA))
        if bigpat.nv != nv:
            raise MuscadeError(f"all-steps pattern of order {bigpat.nv}, expected {nv}")
        st_["nnz"] = bigpat.nnz
        coff = {"Λ": 0, "X": nX, "U": 2 * nX}

        def block(iexp, istep, c):   # istep 1-based
            return int(exp0[iexp]) + (istep - 1) * per_step + coff[c]

        dbig = None
        for it in range(1, maxiter + 1):
            if verbose:
                print(f"\n    Iteration {it:3d}\n        Assembling", end="")
            rows, cols, vals = [], [], []
            hcontrib = []   # device big system: (rows, cols, w, values) of the host pieces, in order
            Lv = np.zeros(nv)
            if IA:
                gA, pieces = host.once(state[0][0])
                Lv[Aoff:Aoff + nA] += gA
                for r, c, v in pieces:
                    rows.append(r + Aoff); cols.append(c + Aoff); vals.append(v)
                    hcontrib.append((r + Aoff, c + Aoff, np.ones(v.size), v))

            def off(iexp, istep, c):
                return Aoff if c == "A" else block(iexp, istep, c)

            t_iter = clock()
            t_dev = 0.
            for iexp in range(nexp):
                dt = Δt[iexp]
                if batch_steps:
                    t0 = clock()
                    stepblocks = batched[iexp].assemble_steps(state[iexp], dbg=dict(dbg, iter=it))
                    dRdA = {ia: batched[iexp].beam_dRdA(ia) for ia in ascaled} if IA else {}
                    t_dev += clock() - t0
                for istep in range(1, nstep[iexp] + 1):
                    st = state[iexp][istep - 1]
                    st.SP = dict(γ=0., iter=it)
                    if batch_steps:
                        L1d, L2d = stepblocks[istep - 1]
                    else:
                        t0 = clock()
                        L1d, L2d = out.assemble(st, st.Λ[0], "matrices", 1.0, dbg=dict(dbg, iter=it, step=istep))
                        t_dev += clock() - t0
                    L1h, L2h = host.step(st, Λgradient)
                    # the device kinds take the residual path, which leaves Λᵀ∂R/∂(X,U) out of L1
                    # (src/DirectXUA.jl:104-119); add it from the same Jacobian blocks
                    for b, pair in (("X", ("Λ", "X")), ("U", ("Λ", "U"))):
                        if Λgradient and pair in L2d and L1h[b]:
                            for bder, M in enumerate(L2d[pair][0][:nder[b]]):
                                L1h[b][bder] = L1h[b][bder] + M.T @ st.Λ[0]
                    if IA:
                        for ia in ascaled:
                            D = dRdA[ia][istep - 1] if batch_steps else out.beam_dRdA(ia)
                            _add_beam_dRdA(D, dis.dis[ia], st.Λ[0], L1h, L2h, Λgradient)
                    for b in classes:
                        for bder in range(nder[b]):
                            v = L1h[b][bder]
                            if b in L1d and bder < len(L1d[b]):
                                v = v + L1d[b][bder]
                            s = dt ** (-bder)
                            for Δs, w in finitediff(bder, nstep[iexp], istep):
                                o = off(iexp, istep + Δs, b)
                                Lv[o:o + v.size] += w * s * v
                    if big == "device":
                        for (a, i, b, j, r, c, v) in L2h:
                            sc = dt ** (-i - j)
                            for Δa, wa in finitediff(i, nstep[iexp], istep):
                                for Δb, wb in finitediff(j, nstep[iexp], istep):
                                    hcontrib.append((r + off(iexp, istep + Δa, a), c + off(iexp, istep + Δb, b),
                                                     np.full(v.size, wa * wb * sc), v))
                        continue
                    for (a, b), blocks in L2d.items():
                        for i, row in enumerate(blocks):
                            for j, M in enumerate(row):
                                C = M.tocoo()
                                if C.nnz == 0:
                                    continue
                                s = dt ** (-i - j)
                                for Δa, wa in finitediff(i, nstep[iexp], istep):
                                    for Δb, wb in finitediff(j, nstep[iexp], istep):
                                        rows.append(C.row + off(iexp, istep + Δa, a))
                                        cols.append(C.col + off(iexp, istep + Δb, b))
                                        vals.append(C.data * (wa * wb * s))
                    for (a, i, b, j, r, c, v) in L2h:
                        s = dt ** (-i - j)
                        for Δa, wa in finitediff(i, nstep[iexp], istep):
                            for Δb, wb in finitediff(j, nstep[iexp], istep):
                                rows.append(r + off(iexp, istep + Δa, a))
                                cols.append(c + off(iexp, istep + Δb, b))
                                vals.append(v * (wa * wb * s))
            Δv = None
            t_lin = 0.
            if big == "device":
                if dbig is None or not dbig.matches(hcontrib):
                    dbig = DeviceBigSystem(batched[0], nv, lambda s_, c_: block(0, s_, c_), Δt[0], hcontrib,
                                           pattern=bigpat)
                    if linsolve == "device":
                        st_["block_order"], st_["blocks"] = dbig.prepare_solve(per_step, nstep[0], nv - Aoff)
                t0 = clock()
                if linsolve == "device":
                    Δv = dbig.solve(hcontrib, Lv, dict(dbg, iter=it))
                    t_lin = clock() - t0
                    st_["lu_s"] += t_lin
                else:
                    Lvv = dbig.assemble(hcontrib)
                    t_dev += clock() - t0
            else:
                Lvv = sparser(bigpat.matrix(np.concatenate(rows) if rows else np.zeros(0, np.int64),
                                            np.concatenate(cols) if cols else np.zeros(0, np.int64),
                                            np.concatenate(vals) if vals else np.zeros(0)))
            st_["device_assembly_s"] += t_dev
            st_["host_assembly_s"] += clock() - t_iter - t_dev - t_lin
            st_["iterations"] = it
            if verbose:
                print(", solving", end="")
            if Δv is None:
                t0 = clock()
                try:
                    Δv = spla.splu(Lvv).solve(Lv)
                except RuntimeError as e:
                    raise MuscadeError(f"Lvv matrix factorization failed: {e}", dict(dbg, iter=it))
                st_["lu_s"] += clock() - t0
            if verbose:
                print(", decrementing.")
            Δ2 = {c: 0. for c in classes}
            for iexp in range(nexp):
                dt = Δt[iexp]
                for istep in range(1, nstep[iexp] + 1):
                    st = state[iexp][istep - 1]
                    for b in ("Λ", "X", "U"):
                        n = nX if b != "U" else nU
                        if n == 0:
                            continue
                        target = st.Λ if b == "Λ" else (st.X if b == "X" else st.U)
                        for bder in range(nder[b]):
                            s = dt ** (-bder)
                            for Δs, w in finitediff(bder, nstep[iexp], istep):
                                o = block(iexp, istep + Δs, b)
                                Δb = Δv[o:o + n]
                                target[bder] -= Δb * (w * s) * scale[b]
                                if bder == 0:
                                    Δ2[b] = max(Δ2[b], float(Δb @ Δb))
            if IA and nA:
                Δa = Δv[Aoff:Aoff + nA]
                Δ2["A"] = float(Δa @ Δa)
                A -= Δa * scale["A"]
            if saveiter:
                stateiter.append(copy.deepcopy(state))
            if verbose:
                for c in classes:
                    print(f"        maxₜ(|Δ{c}|)={math.sqrt(Δ2[c]):7.1e} ≤ {math.sqrt(maxΔ2[c]):7.1e}")
            if all(Δ2[c] <= maxΔ2[c] for c in classes):
                if verbose:
                    print(f"\n    Converged in {it:3d} iterations.\n    nvar={nv}, nstep={sum(nstep)}")
                return
            if it == maxiter:
                raise MuscadeError(f"no convergence after {it:3d} iterations.", dbg)
