"""Adjoint gradients of a limit state at a static equilibrium, on the device.

A design-point or reliability search needs dg/dA and dg/dU of a limit state g(X) at the
equilibrium X*(A, U), R(X*, A, U) = 0. With K = ∂R/∂X at X* (symmetric for the static
EulerBeam3D, Hold and DofLoad), K·μ = ∂g/∂X gives dg/dA = −μᵀ∂R/∂A and dg/dU = −μᵀ∂R/∂U.
The reference forms such gradients with forward-mode duals through the solver
(src/Diagnostic.jl:851-911 for an element); there is no adjoint solver in it. Here one
assembly, one chain solve (mh_adjoint_solve, Λ := −μ) and one adjoint launch
(mh_vjp_lagrangian) give the gradients of every model of a batch at once: the models of a
batch are disjoint chains of one system."""
import ctypes

import numpy as np

from . import _lib
from ._lib import as_f64, dptr, i64ptr, lib
from .assemble import AssemblySweepX
from .chainsolve import ChainSolver
from .model import MuscadeError


class LimitStateGradient:
    """dg/dA and dg/dU at static equilibria of a chain model (many cantilevers or risers batched
    in one model are one chain system)."""

    def __init__(self, model, dis, device=0):
        self.model = model
        self.out = AssemblySweepX(0, model, dis, device=device, graph=False)
        self.solver = ChainSolver(self.out)

    def __call__(self, state, dgdX, dbg=None, upload=True, check=True):
        """Returns (dg/dA (nA), dg/dU (nU), ∂L/∂X (nX) or None) for ∂g/∂X = dgdX at the
        equilibrium `state` (X, U, A). dgdX: a dense vector (nX) or its nonzeros as
        (1-based dofs, values). upload=False assembles at the state already on the device (the
        last call's). ∂L/∂X (check=True) is −(∂g/∂X)ᵀ up to the solve's accuracy (K
        symmetric), a check of the adjoint solve."""
        out = self.out
        out.assemble(state, "iter", dbg=dbg, fetch=False, upload=upload)
        if isinstance(dgdX, tuple):
            idx = np.ascontiguousarray(dgdX[0], dtype=np.int64)
            val = as_f64(np.asarray(dgdX[1], dtype=np.float64))
            r = lib.mh_adjoint_solve_sparse(out.ctx.h, idx.size, i64ptr(idx), dptr(val), -1.0)
        else:
            r = lib.mh_adjoint_solve(out.ctx.h, dptr(as_f64(dgdX)), -1.0)
        if r == _lib.MH_ERR_SOLVE:
            raise MuscadeError("adjoint solve failed: " + self.solver.failure(), dbg)
        out.ctx.check(r, "adjoint solve")
        err = _lib.mh_err()
        r = lib.mh_vjp_lagrangian(out.ctx.h, 0, ctypes.byref(err))
        if r == _lib.MH_ERR_NAN:
            raise MuscadeError("adjoint returned NaN", dict(dbg or {}, ieletyp=err.igroup + 1, iele=err.iele + 1))
        out.ctx.check(r, "adjoint")
        gX = np.zeros(self.model.ndof["X"]) if check else None
        gU = np.zeros(self.model.ndof["U"])
        gA = np.zeros(self.model.ndof["A"])
        out.ctx.check(lib.mh_get_vjp(out.ctx.h, dptr(gX) if check else None, dptr(gU) if gU.size else None,
                                     dptr(gA) if gA.size else None), "adjoint results")
        return gA, gU, gX
