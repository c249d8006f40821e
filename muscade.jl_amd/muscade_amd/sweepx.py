"""solve(SweepX{OX}; initialstate, time, β, γ, maxiter, maxΔx, maxLλ, saveiter) — src/SweepX.jl:134-226,
and the solve wrapper of src/Solve.jl:16-35.

The Newton/Newmark loop stays on the host like in the reference; assembly of
Lλ and Lλx runs on the GPU. The sparse factorization (UMFPACK `lu` in the reference)
runs on the host with SuperLU (linsolve="host"), or, for chain topologies, on the
device (linsolve="device", muscade_amd.chainsolve): then the Newmark decrement runs on
the device too, X stays there through the iterations of a step and each iteration
returns only |Δx|² and |Lλ|².
"""
import math
import sys
import time as _time

import numpy as np
import scipy.sparse.linalg as spla

from .assemble import AssemblySweepX, Newmarkβcoefficients
from .chainsolve import ChainSolver
from .model import MuscadeError


class _SweepXMeta(type):
    def __getitem__(cls, OX):
        if OX not in (0, 1, 2):
            raise MuscadeError("SweepX order must be 0, 1 or 2")
        return type(f"SweepX{{{OX}}}", (SweepX,), {"OX": OX})


class SweepX(metaclass=_SweepXMeta):
    """SweepX{OX}: SweepX[0] static Newton-Raphson, SweepX[1] first-order, SweepX[2] Newmark-β."""
    OX = 0

    @classmethod
    def run(cls, pstate, verbose, dbg, *, time, initialstate, β=0.25, γ=0.5, maxiter=50, maxΔx=1e-5,
            maxLλ=math.inf, saveiter=False, device=0, Assembly=AssemblySweepX, graph=True, linsolve="host", stats=None,
            savestep=1):
        """`Assembly` is the assembler type prepared for the model, as the type argument of
        prepare(AssemblySweepX{OX},model,dis) (SweepX.jl:24,188). graph: the device assembler
        replays captured HIP graphs for the Newton iterations of each time step; with
        linsolve="device" each whole iteration (assembly, solve, decrement) is one graph
        (graph="assembly": the assembly alone, the solve and decrement launched eagerly).
        linsolve: "host" (SuperLU) or "device" (chain topologies only; raises
        chainsolve.NotAChain otherwise). stats (a dict) receives the Newton iteration count, the
        seconds spent preparing the assembler (and the chain layout) and the graph captures.
        savestep (an extension; the reference keeps every step): keep the states of every
        savestep-th step and of the last one, the others left None. A long run of a large model
        then holds a few states on the host, and with linsolve="device" X leaves the device only
        at the kept steps."""
        OX = cls.OX
        model, dis = initialstate.model, initialstate.dis
        if int(savestep) != savestep or savestep < 1:
            raise MuscadeError("savestep must be a positive integer")
        nstep = len(time)

        def kept(step):
            return step % savestep == 0 or step == nstep
        if linsolve not in ("host", "device"):
            raise MuscadeError(f"linsolve must be 'host' or 'device', not {linsolve!r}")
        dev = linsolve == "device"
        if dev and Assembly is not AssemblySweepX:
            raise MuscadeError("linsolve='device' solves the device assembler's system (Assembly=AssemblySweepX)")
        if graph not in (True, False, "assembly"):
            raise MuscadeError("graph must be True, False or 'assembly'")
        kw = {"graph": bool(graph)} if Assembly is AssemblySweepX else {}
        t_prep = _time.perf_counter()
        out = Assembly(OX, model, dis, device=device, **kw)
        solver = ChainSolver(out) if dev else None
        # the whole Newton iteration as one replayed graph: device solve, no host element types
        whole = dev and graph is True and not getattr(out, "host_types", [])
        if stats is not None:
            stats["prepare_s"] = _time.perf_counter() - t_prep
        nX = out.nX
        scaleX = dis.scaleX
        cΔx2, cLλ2 = maxΔx ** 2, maxLλ ** 2
        state = initialstate.copy().with_ders(1, OX + 1, 1)
        states = [None] * (maxiter if saveiter else len(time))
        pstate[0] = states
        citer = 0
        a = b = None
        for step, t in enumerate(time, start=1):
            oldt = state.time
            state.time = float(t)
            Δt = state.time - oldt
            if Δt <= 0 and OX > 0:
                raise MuscadeError("Time step length not strictly positive at step=%3d" % step)
            out.c = Newmarkβcoefficients(OX, Δt, β, γ)
            c = out.c
            if dev:   # X, X′, X″ are on the device from the first step on
                out.set_state(state, state.time, X=(step == 1))
            for iiter in range(1, maxiter + 1):
                citer += 1
                if stats is not None:
                    stats["iterations"] = citer
                first = iiter == 1
                d = dict(dbg, solver="SweepX", step=step, iiter=iiter)
                if dev and whole:
                    Δx2, Lλ2 = solver.iteration_graph(OX, first, c.array(), d)
                elif dev:
                    out.assemble(state, "step" if first else "iter", dbg=d, fetch=False, upload=False)
                    solver.solve(d)
                    Δx2, Lλ2 = solver.newton_update(OX, first, c.array())
                if dev:
                    if not math.isfinite(Δx2):
                        raise MuscadeError("matrix factorization failed at step=%i, iiter=%i" % (step, iiter))
                    conv = Δx2 <= cΔx2 and Lλ2 <= cLλ2
                    if saveiter or (conv and kept(step)):
                        solver.fetch_X(state, OX)
                    if saveiter:
                        states[iiter - 1] = state.copy()
                    if conv:
                        if verbose:
                            print("    step %3d converged in %3d iterations. |Δx|=%7.1e |Lλ|=%7.1e"
                                  % (step, iiter, math.sqrt(Δx2), math.sqrt(Lλ2)))
                        if not saveiter and kept(step):
                            states[step - 1] = state.copy()
                        break
                    if iiter == maxiter:
                        raise MuscadeError("no convergence of step %3d after %3d iterations |Δx|=%g / %g, |Lλ|=%g / %g"
                                           % (step, iiter, math.sqrt(Δx2), maxΔx, math.sqrt(Lλ2), maxLλ))
                    continue
                Lλ, _ = out.assemble(state, "step" if first else "iter", dbg=d)
                try:
                    lu = spla.splu(out.Lλx())
                    Δx = lu.solve(Lλ)
                except Exception as e:   # SweepX.jl:207-209
                    raise MuscadeError("matrix factorization failed at step=%i, iiter=%i" % (step, iiter)) from e
                if not np.all(np.isfinite(Δx)):
                    raise MuscadeError("matrix factorization failed at step=%i, iiter=%i" % (step, iiter))
                Δx2, Lλ2 = float(np.sum(Δx ** 2)), float(np.sum(Lλ ** 2))
                # Newmarkβdecrement!{OX} (SweepX.jl:98-132) with the allXdofs group
                if OX == 2:
                    if first:
                        x1 = state.X[1] / scaleX
                        x2 = state.X[2] / scaleX
                        a = c.a2 * x1 + c.a3 * x2
                        b = c.b2 * x1 + c.b3 * x2
                        Δx1 = c.a1 * Δx + a
                        Δx2_ = c.b1 * Δx + b
                    else:
                        Δx1, Δx2_ = c.a1 * Δx, c.b1 * Δx
                    state.X[0] -= Δx * scaleX
                    state.X[1] -= Δx1 * scaleX
                    state.X[2] -= Δx2_ * scaleX
                elif OX == 1:
                    if first:
                        a = c.a2 * (state.X[1] / scaleX)
                        Δx1 = c.a1 * Δx + a
                    else:
                        Δx1 = c.a1 * Δx
                    state.X[0] -= Δx * scaleX
                    state.X[1] -= Δx1 * scaleX
                else:
                    state.X[0] -= Δx * scaleX
                if saveiter:
                    states[iiter - 1] = state.copy()
                if Δx2 <= cΔx2 and Lλ2 <= cLλ2:
                    if verbose:
                        print("    step %3d converged in %3d iterations. |Δx|=%7.1e |Lλ|=%7.1e"
                              % (step, iiter, math.sqrt(Δx2), math.sqrt(Lλ2)))
                    if not saveiter and kept(step):
                        states[step - 1] = state.copy()
                    break
                if iiter == maxiter:
                    raise MuscadeError("no convergence of step %3d after %3d iterations |Δx|=%g / %g, |Lλ|=%g / %g"
                                       % (step, iiter, math.sqrt(Δx2), maxΔx, math.sqrt(Lλ2), maxLλ))
        if stats is not None and hasattr(out, "graph_captures"):
            stats["graph_captures"] = out.graph_captures()
        if verbose:
            print("\n    nel=%d, ndof=%d, nstep=%d, niter=%d, niter/nstep=%5.2f"
                  % (model.getnele(), nX, len(time), citer, citer / max(len(time), 1)))


def solve(Solver, dbg=None, verbose=True, silenterror=False, catcherror=True, **kwargs):
    """solve(Solver; dbg, verbose, silenterror, catcherror, kwargs...) — src/Solve.jl:16-35.

    Returns the states computed so far, also when the solver raised."""
    pstate = [None]
    dbg = dict(dbg or {})
    if verbose:
        print(f"\n\n\nMuscade: {Solver.__name__} solver\n")
    t0 = _time.perf_counter()
    if catcherror:
        try:
            Solver.run(pstate, verbose, dict(dbg, solver=Solver.__name__), **kwargs)
        except MuscadeError as e:
            if not silenterror:
                print(f"\n\nMuscadeException: {e}\n\nAborting the analysis. Function 'solve' still returns any "
                      "results obtained before the exception.", file=sys.stderr)
    else:
        Solver.run(pstate, verbose, dict(dbg, solver=Solver.__name__), **kwargs)
    if verbose:
        print("    %s time: %.3f [s]\nMuscade done.\n\n" % (Solver.__name__, _time.perf_counter() - t0))
    return pstate[0]
