"""Linear solve of SweepX's Newton system on the device, for chain topologies.

The reference factorizes out.Lλx on the host and back-substitutes (src/SweepX.jl:205-209:
`lu(out.Lλx)`, `lu!`, `Δx = Lλx\\out.Lλ`), then decrements the state
(Newmarkβdecrement!{OX}, src/SweepX.jl:98-132). For a model whose multi-node elements join
their nodes into paths (mooring lines, risers, cantilevers; many of them batched in one
model), Lλx is block tridiagonal once the dofs are ordered node by node along the paths,
and the library solves it on the device by block cyclic reduction (include/muscade_hip.h,
mh_chain_prepare / mh_chain_solve). `chain_layout` finds that ordering once per model:

- λ pairs: the Lagrange-multiplier dof of each Hold (DofConstraint{:X},
  src/BasicElements.jl:426-438,484-488) is coupled to its held dof only; it leaves the block
  system (its row fixes the held dof, its value follows from the held dof's row).
- node blocks: every other X dof belongs to the block of its node (dofnode); blocks hold
  B = 3 or 6 positions (padded with identity rows), in the order of the nodes along each
  path, the paths one after another (in the order of their lowest-numbered end node).

A model whose elements join more than two nodes, or whose node graph has a branch or a
closed loop, is not a chain: `chain_layout` raises NotAChain and the host factorization
remains the way to solve it.
"""

import ctypes
import re

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components, depth_first_order

from . import _lib
from ._lib import as_f64, dptr, i64ptr, lib
from .model import MuscadeError


class NotAChain(MuscadeError):
    pass


class ChainLayout:
    """B, nblock; pos (nX, 1-based position or −p for the λ of pair p); pair_hpos (1-based);
    dummy_pos (1-based padding positions); node_order (0-based nodes in block order)."""

    def __init__(self, B, nblock, pos, pair_hpos, dummy_pos, node_order, pair_lambda):
        self.B, self.nblock = B, nblock
        self.pos, self.pair_hpos, self.dummy_pos = pos, pair_hpos, dummy_pos
        self.node_order, self.pair_lambda = node_order, pair_lambda


def hold_pairs(model, dis):
    """(λ dofs, held dofs), 0-based, of the model's Hold elements (element dofs [x, λ])."""
    lam, held = [], []
    for it, E in enumerate(model.eletypes):
        if getattr(E.cls, "kind", None) == _lib.MH_KIND_HOLD and E.nele:
            X = dis.dis[it].X
            held.append(X[:, 0] - 1)
            lam.append(X[:, 1] - 1)
    if not lam:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    return np.concatenate(lam).astype(np.int64), np.concatenate(held).astype(np.int64)


def chain_layout(model, dis, colptr, rowval):
    """The block layout of Lλx for a chain model (colptr, rowval: 1-based CSC pattern)."""
    nX, nnod = model.ndof["X"], model.nnod
    if nX == 0:
        raise NotAChain("the model has no X dofs")
    dn = model.dofnode("X").astype(np.int64)
    lam, held = hold_pairs(model, dis)
    islam = np.zeros(nX, bool)
    islam[lam] = True
    if np.unique(lam).size != lam.size:
        raise NotAChain("a λ dof belongs to several constraints")
    if np.unique(held).size != held.size:
        raise NotAChain("a dof is held by several constraints")
    if islam[held].any():
        raise NotAChain("a λ dof is itself held")
    if lam.size:
        c0 = colptr[lam] - 1
        cnt = colptr[lam + 1] - 1 - c0
        if cnt.max() > 2:
            raise NotAChain("a λ dof is coupled to more than its held dof")
        for j in range(2):
            m = cnt > j
            r = rowval[c0[m] + j] - 1
            if not ((r == lam[m]) | (r == held[m])).all():
                raise NotAChain("a λ dof is coupled to another dof than its held dof")
    # node graph of the elements' X dofs (λ dofs excluded)
    edges = []
    big = np.iinfo(np.int64).max
    for it, E in enumerate(model.eletypes):
        X = dis.dis[it].X
        if E.nele == 0 or X.shape[1] == 0:
            continue
        G = X.astype(np.int64) - 1
        valid = G >= 0
        valid[valid] = ~islam[G[valid]]
        nd = np.where(valid, dn[np.maximum(G, 0)], -1)
        mx = nd.max(axis=1)
        mn = np.where(valid, nd, big).min(axis=1)
        if not (~valid | (nd == mn[:, None]) | (nd == mx[:, None])).all():
            raise NotAChain(f"elements of type {it + 1} join more than two nodes")
        sel = valid.any(axis=1) & (mn != mx)
        if sel.any():
            edges.append(np.stack([mn[sel], mx[sel]], axis=1))
    if edges:
        key = np.unique(np.concatenate([e[:, 0] * nnod + e[:, 1] for e in edges]))
        Ed = np.stack([key // nnod, key % nnod], axis=1)
    else:
        Ed = np.zeros((0, 2), np.int64)
    deg = np.bincount(Ed.ravel(), minlength=nnod)
    if deg.size and deg.max() > 2:
        raise NotAChain("a node joins more than two elements' nodes (a branch)")
    hasX = np.zeros(nnod, bool)
    hasX[dn[~islam]] = True
    A = sp.csr_matrix((np.ones(2 * len(Ed)), (np.concatenate([Ed[:, 0], Ed[:, 1]]), np.concatenate([Ed[:, 1], Ed[:, 0]]))),
                      shape=(nnod, nnod))
    _, labels = connected_components(A, directed=False)
    ep = np.nonzero(hasX & (deg <= 1))[0]
    _, first = np.unique(labels[ep], return_index=True)
    ends = ep[first]                       # lowest-numbered end node of each path
    root = nnod
    Ar = sp.csr_matrix((np.ones(2 * len(Ed) + 2 * len(ends)),
                        (np.concatenate([Ed[:, 0], Ed[:, 1], np.full(len(ends), root), ends]),
                         np.concatenate([Ed[:, 1], Ed[:, 0], ends, np.full(len(ends), root)]))),
                       shape=(nnod + 1, nnod + 1))
    order = depth_first_order(Ar, root, directed=True, return_predecessors=False)[1:]
    order = order[hasX[order]].astype(np.int64)
    if order.size != int(hasX.sum()):
        raise NotAChain("the node graph has a closed loop")
    nb = order.size
    blk = np.full(nnod, -1, np.int64)
    blk[order] = np.arange(nb)
    cnt = np.bincount(dn[~islam], minlength=nnod)
    B = 3 if cnt.max() <= 3 else 6
    if cnt.max() > 6:
        raise NotAChain("a node has more than 6 X dofs besides λ dofs")
    dofs = np.nonzero(~islam)[0]
    key = blk[dn[dofs]]
    idx = np.lexsort((dofs, key))
    ks = key[idx]
    within = np.arange(ks.size) - np.searchsorted(ks, ks, side="left")
    pos = np.zeros(nX, np.int64)
    pos[dofs[idx]] = ks * B + within + 1
    pos[lam] = -(np.arange(lam.size) + 1)
    hpos = pos[held].copy()
    taken = np.zeros(nb * B, bool)
    taken[pos[dofs] - 1] = True
    dummy = np.nonzero(~taken)[0] + 1
    return ChainLayout(B, nb, pos, hpos, dummy.astype(np.int64), order, lam)


class ChainSolver:
    """Device solve of Lλx·Δx = Lλ for an AssemblySweepX whose model is a chain, and the
    device Newton-Newmark decrement (SweepX.jl:98-132,205-211)."""

    def __init__(self, out):
        self.out = out
        self.layout = L = chain_layout(out.model, out.dis, out.colptr, out.rowval)
        h = out.ctx.h
        out.ctx.check(lib.mh_chain_prepare(h, L.B, L.nblock, i64ptr(L.pos), L.pair_hpos.size,
                                           i64ptr(L.pair_hpos) if L.pair_hpos.size else None, L.dummy_pos.size,
                                           i64ptr(L.dummy_pos) if L.dummy_pos.size else None), "chain prepare")
        out.ctx.check(lib.mh_newton_prepare(h, dptr(as_f64(out.dis.scaleX))), "newton prepare")
        self.norms = np.zeros(2)

    def solve(self, dbg=None):
        """Δx from the last assembly, left on the device."""
        r = lib.mh_chain_solve(self.out.ctx.h)
        if r == _lib.MH_ERR_SOLVE:
            d = dict(dbg or {})
            d["reason"] = self.failure()
            raise MuscadeError("matrix factorization failed at step=%s, iiter=%s" % (d.get("step"), d.get("iiter")), d)
        self.out.ctx.check(r, "chain solve")

    def failure(self):
        """The library's reason for the last failed solve, with the model node of a singular
        diagonal block (1-based) where it names one."""
        msg = lib.mh_last_error(self.out.ctx.h).decode()
        m = re.search(r"diagonal block (\d+)", msg)
        if m and int(m.group(1)) < self.layout.node_order.size:
            msg += f", node {int(self.layout.node_order[int(m.group(1))]) + 1}"
        return msg

    def dx(self):
        x = np.zeros(self.out.nX)
        self.out.ctx.check(lib.mh_get_dx(self.out.ctx.h, dptr(x)), "Δx")
        return x

    def iteration_graph(self, OX, first, c6, dbg=None):
        """One whole device Newton iteration (assembly, solve, Newmark decrement) replayed from a
        captured HIP graph (mh_newton_iteration_graph); returns (Σ Δx², Σ Lλ²)."""
        c6 = as_f64(c6)
        err = _lib.mh_err()
        r = lib.mh_newton_iteration_graph(self.out.ctx.h, OX, int(first), dptr(c6), dptr(self.norms), ctypes.byref(err))
        d = dict(dbg or {})
        if r == _lib.MH_ERR_NAN:
            E = self.out.model.eletypes[err.igroup]
            raise MuscadeError(f"residual({E.key[0]},...) returned NaN in R, FB or derivatives",
                               dict(d, ieletyp=err.igroup + 1, iele=err.iele + 1))
        if r == _lib.MH_ERR_SOLVE:
            d["reason"] = self.failure()
            raise MuscadeError("matrix factorization failed at step=%s, iiter=%s" % (d.get("step"), d.get("iiter")), d)
        self.out.ctx.check(r, "Newton iteration")
        return float(self.norms[0]), float(self.norms[1])

    def newton_update(self, OX, first, c6):
        """Newmarkβdecrement!{OX} of the device state; returns (Σ Δx², Σ Lλ²)."""
        c6 = as_f64(c6)
        self.out.ctx.check(lib.mh_newton_update(self.out.ctx.h, OX, int(first), dptr(c6), dptr(self.norms)),
                           "newton update")
        return float(self.norms[0]), float(self.norms[1])

    def fetch_X(self, state, OX):
        """Copy the device X, X′, X″ (up to OX) into state.X."""
        for d in range(OX + 1):
            x = np.zeros(self.out.nX)
            self.out.ctx.check(lib.mh_get_X(self.out.ctx.h, d, dptr(x)), "X")
            state.X[d] = x

    def timings(self):
        ms = np.zeros(3)
        self.out.ctx.check(lib.mh_chain_timings(self.out.ctx.h, dptr(ms)), "chain timings")
        return ms
