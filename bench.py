"""Benchmark: EulerBeam3D elements assembled per second (FP64 residual + tangent).

Workload: static assembly at scale — a 100×100×33-node space frame
(973,400 EulerBeam3D + 60,000 Hold elements, 2,040,000 X-dofs, seed 1001),
see muscade_amd/workloads.py. One step = one assemble!{:iter} of
AssemblySweepX{0}: element evaluation of every element plus the scatter into
Lλ and Lλx.nzval, from a device-resident state to device-resident outputs.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line. Kernel times come from HIP events recorded by the
library on the stream it launches on. Two secondary measurements ride along
in the same line (skip them with --no-dynamic / --no-adjoint): the dynamic
SweepX{2} assembly of 256 catenary chains ("dynamic") and batched vs plain
adjoint gradients of 512 uneven cantilevers ("adjoint").
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "muscade.jl_amd"))

HBM_PEAK = 8.0e12          # MI355X HBM3E, B/s
FP64_PEAK = 78.6e12        # MI355X FP64 vector, FLOP/s
# FP64 operations per beam of the static kernel (add/sub/mul/div/sqrt = 1, fma = 2),
# from the SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 counters of one assembly of this
# workload divided by the number of beams (profiles/space_frame_100x100x33_pmc.md).
F_ELEM = 6299
# The same for the dynamic beam of the catenary chains (OX = 2): value pass + direction pass
# per beam, and their HBM bytes per assembly from FETCH_SIZE/WRITE_SIZE
# (profiles/catenary_chains_256_kernels.md), by mission.
F_DYN = {"iter": 10890 + 75981, "step": 10890 + 84845}
DYN_TRAFFIC = {"iter": 9.17e9 + 13.77e9, "step": 9.17e9 + 14.39e9}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "space_frame_100x100x33_traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(nbeam, nhold, nnz, nX, N_c, N_cR, nKbuf):
    """Bytes each kernel must move at least (the roofline numerator)."""
    # eval, per beam: in X 12×8, dof map 12×4, nodes 2×4, coordinates 2×24, orient2 24, section 5×8;
    #                 out 45 tangent numbers ×8, R 12×8
    eval_beam = 96 + 48 + 8 + 48 + 24 + 40 + 45 * 8 + 96
    # Hold: in 2×8 + 2×4, out 4×8 + 2×8
    eval_b = nbeam * eval_beam + nhold * (16 + 8 + 32 + 16)
    # nz scatter: every element output read once (nKbuf doubles), one coded source
    # index per contribution, one segment offset and one write per nz
    scat_K = 8 * nKbuf + 4 * N_c + 4 * (nnz + 1) + 8 * nnz
    scat_R = 12 * N_cR + 8 * nX + 4 * (nX + 1)
    return eval_b, scat_K, scat_R


def calib_copy(n=1 << 27):
    """Time one streaming copy of n doubles (the counter-calibration kernel)."""
    import muscade_amd as mc
    from muscade_amd import _lib
    ctx = mc.DeviceContext(0)
    ms = np.zeros(1)
    ctx.check(_lib.lib.mh_calib_copy(ctx.h, n, _lib.dptr(ms)), "calib copy")
    print(json.dumps({"calib_copy_bytes_read": 8 * n, "calib_copy_bytes_written": 8 * n, "ms": ms[0],
                      "GB_s": 16 * n / (ms[0] * 1e-3) / 1e9}))


def cpu_baseline(model, state, nsample, nthreads=1):
    """The oracle (Muscade's nested-dual algorithm restated, add_value!/add_∂! in element
    order), timed on the first `nsample` beams of the same model: 1 thread (Muscade's
    assemble! is single-threaded), or the element evaluations on `nthreads` OpenMP threads
    followed by the same ordered scatter."""
    sys.path.insert(0, ROOT)
    from tests import _oracle as orc
    E = model.eletypes[0]
    d = state.dis.dis[0]
    n = min(nsample, d.X.shape[0])
    xyz = model.coords()
    nodes = E.cat_nodes()[:n]
    o2 = np.concatenate([p["orient2"] for p in E.params])[:n]
    mat = np.concatenate([p["mat"] for p in E.params])[:n]
    params = np.concatenate([xyz[nodes[:, 0], :3], xyz[nodes[:, 1], :3], o2, mat], axis=1)
    om = orc.OracleModel(model.ndof["X"])
    om.add_type(orc.K_BEAM, d.X[:n], None, None, d.scaleX, params)
    om.prepare()
    t0 = time.perf_counter()
    if nthreads == 1:
        r, rhs, nz, _ = om.assemble(0, 1, np.zeros(6), [state.X[0]])
    else:
        r, rhs, nz, _ = om.assemble0_parallel(nthreads, state.X[0])
    dt = time.perf_counter() - t0
    assert r == 0
    return n / dt, n, dt


def cpu_baseline_dynamic(model, state, nsample, c6):
    """The oracle's dynamic assembly (SweepX{2} :iter: three time levels of nested duals around the
    solver's partials, then add_value!/add_∂! in element order) on the first `nsample` beams of
    the same model, 1 thread."""
    sys.path.insert(0, ROOT)
    from tests import _oracle as orc
    E = model.eletypes[0]
    d = state.dis.dis[0]
    n = min(nsample, d.X.shape[0])
    xyz = model.coords()
    nodes = E.cat_nodes()[:n]
    o2 = np.concatenate([p["orient2"] for p in E.params])[:n]
    mat = np.concatenate([p["mat"] for p in E.params])[:n]
    params = np.concatenate([xyz[nodes[:, 0], :3], xyz[nodes[:, 1], :3], o2, mat], axis=1)
    om = orc.OracleModel(model.ndof["X"])
    om.add_type(orc.K_BEAM, d.X[:n], None, None, d.scaleX, params)
    om.prepare()
    t0 = time.perf_counter()
    r, rhs, nz, _ = om.assemble(2, 1, c6, state.X[:3])
    dt = time.perf_counter() - t0
    assert r == 0
    return n / dt, n, dt


def cpu_baseline_gradient(model, state, lam, nsample):
    """The reference's way to a gradient of L = Σ Λ_eᵀR_e: per element the forward-mode Jacobian
    (diffed_residual, nested duals) contracted with Λ_e, on the first `nsample` beams, 1 thread.
    Returns seconds per beam."""
    sys.path.insert(0, ROOT)
    from tests import _oracle as orc
    E = model.eletypes[0]
    d = state.dis.dis[0]
    n = min(nsample, d.X.shape[0])
    xyz = model.coords()
    nodes = E.cat_nodes()
    o2 = np.concatenate([p["orient2"] for p in E.params])
    mat = np.concatenate([p["mat"] for p in E.params])
    amask = E.params[0].get("amask", 0)
    if amask:
        mat = E.cls.scaled_section(mat, state.A[d.A[:, 0] - 1], amask)
    udof = d.U.shape[1] > 0
    t0 = time.perf_counter()
    for e in range(n):
        R, Ks, KU = orc.beam_diffed_residual(xyz[nodes[e, 0], :3], xyz[nodes[e, 1], :3], o2[e], mat[e],
                                             [state.X[0][d.X[e] - 1]], state.U[0][d.U[e] - 1] if udof else None)
        g = lam[d.X[e] - 1] @ Ks[0]
    dt = time.perf_counter() - t0
    return dt / n, n, dt


def host_threads():
    """Threads of the all-cores CPU baseline: OMP_NUM_THREADS if set (the host share of one
    GPU), else the visible cores, at most 16."""
    env = os.environ.get("OMP_NUM_THREADS")
    return max(1, int(env)) if env and env.isdigit() else max(1, min(16, os.cpu_count() or 1))


def dynamic_bench(args):
    """Secondary measurement: dynamic assembly (AssemblySweepX{2}, :iter) of the catenary
    chains (inertia, added mass, drag, roll inertia; K + a₁C + b₁M), 256 chains by default."""
    import muscade_amd as mc
    from muscade_amd import _lib
    from muscade_amd.workloads import catenary_chains
    t0 = time.perf_counter()
    model, state = catenary_chains(args.chains, args.chain_min, args.chain_max)
    nbeam = model.eletypes[0].nele
    log(f"dynamic model: {args.chains} chains, {nbeam} beams, nX={model.ndof['X']} ({time.perf_counter() - t0:.1f}s)")
    t0 = time.perf_counter()
    out = mc.AssemblySweepX(2, model, state.dis)
    t_prep = time.perf_counter() - t0
    log(f"dynamic prepare (nnz={out.nnz}): {t_prep:.2f}s")
    out.c = mc.Newmarkβcoefficients(2, 0.05, 0.25, 0.5)
    out.set_state(state, 0.0)
    h = out.ctx.h
    c6 = out.c.array()
    err = _lib.mh_err()
    _lib.lib.mh_assemble_sweepx(h, 0, 1, _lib.dptr(c6), ctypes.byref(err))   # the first assembly tunes the beam kernel
    bb, bms = ctypes.c_int(), (ctypes.c_float * 3)()
    if _lib.lib.mh_get_beam_tuning(h, ctypes.byref(bb), bms) == 0:
        tuning.update(beam_block=bb.value, beam_block_ms=[round(float(x), 4) for x in bms])
        log(f"beam kernel tuner: block {bb.value}, ms {list(bms)}")
    res = {}
    for mission, m in (("iter", 1), ("step", 0)):
        def step():
            r = _lib.lib.mh_assemble_sweepx(h, 2, m, _lib.dptr(c6), ctypes.byref(err))
            if r != 0:
                raise RuntimeError(_lib.lib.mh_last_error(h).decode())
        for _ in range(args.warmup):
            step()
        _lib.lib.mh_sync(h)
        ms = np.zeros(3)
        acc = np.zeros(3)
        nst = max(1, args.steps // 2)
        t0 = time.perf_counter()
        for _ in range(nst):
            step()
            _lib.lib.mh_last_timings(h, _lib.dptr(ms))
            acc += ms
        _lib.lib.mh_sync(h)
        el = time.perf_counter() - t0
        res[mission] = {"elements_per_s": nbeam * nst / el, "ms_per_step": el / nst * 1e3,
                        "kernels_ms": dict(zip(("eval", "scatter_nz", "scatter_rhs"), (acc / nst).tolist()))}
        t_eval = acc[0] / nst * 1e-3
        if t_eval > 0:
            ach = F_DYN[mission] * nbeam / t_eval / 1e12
            res[mission]["roofline_eval"] = {
                "bound": "fp64", "achieved": ach, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                "frac": ach / (FP64_PEAK / 1e12),
                "traffic": DYN_TRAFFIC[mission] if args.chains == 256 else None}
        log(f"dynamic {mission}: {el / nst * 1e3:.3f} ms/step, kernels {acc / nst}")
    if not args.no_chain_solve:
        res["newton_device"] = device_newton_iteration(args, out, state, c6)
    if not args.no_cpu_baseline:
        eps, ns, dts = cpu_baseline_dynamic(model, state, args.cpu_sample_dynamic, c6)
        res["cpu_baseline"] = {"value": eps, "unit": "elements/s", "cores": 1, "kind": "port",
                               "sample": f"first {ns} beams of the same chains, one sequential assemble!{{:iter}} of "
                                         f"AssemblySweepX{{2}} ({dts:.1f} s)"}
        log(f"dynamic cpu baseline: {eps:.0f} beams/s")
    return {"workload": f"catenary chains x{args.chains}, {nbeam} EulerBeam3D with inertia/added mass/drag/roll "
                        f"inertia, AssemblySweepX{{2}} (Newmark-beta, dt=0.05)",
            "nbeam": nbeam, "nX": out.nX, "nnz": out.nnz, "prepare_s": t_prep, **res}


def device_newton_iteration(args, out, state, c6):
    """One Newton iteration of SweepX{2} (the :step one) on the device for the catenary chains:
    assembly, block-cyclic-reduction solve of Lλx·Δx = Lλ and the Newmark decrement, X
    resident, two scalars back (SweepX.jl:202-211). Each timed iteration starts from the
    same uploaded state (upload untimed)."""
    from muscade_amd.chainsolve import ChainSolver
    t0 = time.perf_counter()
    cs = ChainSolver(out)
    t_lay = time.perf_counter() - t0
    L = cs.layout
    log(f"chain layout: {L.nblock} blocks of {L.B}, {L.pair_hpos.size} λ pairs ({t_lay:.1f}s)")

    def iteration():
        out.assemble(state, "step", fetch=False, upload=False)
        cs.solve()
        return cs.newton_update(2, True, c6)

    out.set_state(state, 0.0)
    iteration()
    n = max(2, args.steps // 4)
    wall, asm, sol = [], np.zeros(3), np.zeros(3)
    for _ in range(n):
        out.set_state(state, 0.0)
        out.ctx.check(_lib_sync(out), "sync")
        t0 = time.perf_counter()
        n2 = iteration()
        wall.append(time.perf_counter() - t0)
        asm += out.timings()
        sol += cs.timings()
    w = float(np.mean(wall)) * 1e3
    # full-size property: the device Δx solves the assembled system to a backward error
    # ‖Lλx·Δx − Lλ‖∞ / (‖Lλx‖∞‖Δx‖∞ + ‖Lλ‖∞) of a backward-stable solver (host check)
    out.set_state(state, 0.0)
    out.assemble(state, "step", upload=False)
    cs.solve()
    dx = cs.dx()
    K = out.Lλx()
    rn = np.bincount(out._indices0, weights=np.abs(out.nzval), minlength=out.nX).max()
    res_ = np.abs(K @ dx - out.Lλ).max()
    berr = float(res_ / (rn * np.abs(dx).max() + np.abs(out.Lλ).max()))
    rres = float(res_ / np.abs(out.Lλ).max())
    del K
    log(f"device Δx at full size: backward error {berr:.2e}, residual/|Lλ| {rres:.2e}")
    res = {"layout_s": t_lay, "nblock": L.nblock, "B": L.B, "npair": int(L.pair_hpos.size), "backward_error": berr,
           "relative_residual": rres,
           "iteration_ms": w, "assembly_kernels_ms": (asm / n).tolist(),
           "solve_ms": {"pack": sol[0] / n, "eliminate": sol[1] / n, "back_substitute": sol[2] / n},
           "norms": [float(np.sqrt(n2[0])), float(np.sqrt(n2[1]))]}
    log(f"device Newton iteration: {w:.2f} ms (solve: pack {sol[0] / n:.2f}, eliminate {sol[1] / n:.2f}, back "
        f"{sol[2] / n:.2f} ms), |Δx|={res['norms'][0]:.3e}")
    return res


def _lib_sync(out):
    from muscade_amd import _lib
    return _lib.lib.mh_sync(out.ctx.h)


def chain_solve_bench(args):
    """Device block-cyclic-reduction solve against the host sparse LU (SciPy SuperLU, as the
    reference's `lu` + `ldiv!`, SweepX.jl:205-209) on the same dynamic Lλx of a reduced
    catenary model (16 chains of 1000-4000 beams), both from the assembled matrix."""
    import scipy.sparse.linalg as spla
    import muscade_amd as mc
    from muscade_amd.chainsolve import ChainSolver
    from muscade_amd.workloads import catenary_chains
    model, state = catenary_chains(16, 1000, 4000, seed=1005)
    state.time = 1.0
    out = mc.AssemblySweepX(2, model, state.dis)
    out.c = mc.Newmarkβcoefficients(2, 0.05, 0.25, 0.5)
    out.assemble(state, "step")
    K, r = out.Lλx(), out.Lλ.copy()
    cs = ChainSolver(out)
    cs.solve()
    n = 10
    t0 = time.perf_counter()
    for _ in range(n):
        cs.solve()
    t_dev = (time.perf_counter() - t0) / n
    dx = cs.dx()
    t0 = time.perf_counter()
    lu = spla.splu(K.tocsc())
    t1 = time.perf_counter()
    ref = lu.solve(r)
    t2 = time.perf_counter()
    Ka = abs(K)
    denom = Ka.sum(axis=1).max() * np.abs(dx).max() + np.abs(r).max()
    res = {"workload": f"catenary chains x16, {model.eletypes[0].nele} EulerBeam3D, nX={out.nX}, dynamic Lλx (:step)",
           "device_solve_ms": t_dev * 1e3, "host_lu_ms": (t1 - t0) * 1e3, "host_solve_ms": (t2 - t1) * 1e3,
           "speedup_vs_host_lu_and_solve": (t2 - t0) / t_dev,
           "device_backward_error": float(np.abs(K @ dx - r).max() / denom),
           "host_backward_error": float(np.abs(K @ ref - r).max() /
                                        (Ka.sum(axis=1).max() * np.abs(ref).max() + np.abs(r).max()))}
    log(f"chain solve ({out.nX} dofs): device {t_dev * 1e3:.2f} ms, host LU {res['host_lu_ms']:.0f} ms + solve "
        f"{res['host_solve_ms']:.1f} ms; backward errors {res['device_backward_error']:.1e} / "
        f"{res['host_backward_error']:.1e}")
    return res


def adjoint_bench(args):
    """Secondary measurement: batched adjoint gradients vs the plain per-model path.
    The same `--models` uneven cantilevers (EulerBeam3D{true} with 8 U-loads each) either
    joined into one model (one adjoint launch per element kind for all of them) or one
    context per model called in a loop. A gradient set is (∂L/∂X, ∂L/∂U) of
    L = Σ_e Λ_eᵀ R_e for one model, computed on the device from resident X, U, Λ."""
    import muscade_amd as mc
    from muscade_amd import _lib
    from muscade_amd.workloads import cantilever_batch, cantilever_batch_specs
    specs = cantilever_batch_specs(args.models, args.model_min, args.model_max)
    t0 = time.perf_counter()
    [(mj, sj, lj)] = cantilever_batch(specs)
    sep = cantilever_batch(specs, joined=False)
    nbeam = mj.eletypes[0].nele
    log(f"adjoint batch: {len(specs)} models, {nbeam} beams ({time.perf_counter() - t0:.1f}s)")
    err = _lib.mh_err()
    outj = mc.AssemblySweepX(0, mj, sj.dis)
    outj.set_state(sj, 0.)
    outj.ctx.check(_lib.lib.mh_set_lambda(outj.ctx.h, _lib.dptr(lj)), "Λ")
    plain = []
    for m, s_, l_ in sep:
        o = mc.AssemblySweepX(0, m, s_.dis)
        o.set_state(s_, 0.)
        o.ctx.check(_lib.lib.mh_set_lambda(o.ctx.h, _lib.dptr(l_)), "Λ")
        plain.append(o)

    def batched():
        if _lib.lib.mh_vjp_lagrangian(outj.ctx.h, 0, ctypes.byref(err)) != 0:
            raise RuntimeError(_lib.lib.mh_last_error(outj.ctx.h).decode())

    def per_model():
        for o in plain:
            if _lib.lib.mh_vjp_lagrangian(o.ctx.h, 0, ctypes.byref(err)) != 0:
                raise RuntimeError(_lib.lib.mh_last_error(o.ctx.h).decode())

    res = {}
    for name, fn in (("batched", batched), ("plain", per_model)):
        for _ in range(args.warmup):
            fn()
        n = max(2, args.steps // 2)
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        el = (time.perf_counter() - t0) / n
        res[name] = {"grads_per_s": len(specs) / el, "ms_per_batch": el * 1e3}
        log(f"adjoint {name}: {el * 1e3:.3f} ms for {len(specs)} gradient sets")
    ms = np.zeros(3)
    batched()
    _lib.lib.mh_last_timings(outj.ctx.h, _lib.dptr(ms))
    # the batched gradients of a few models equal their plain ones
    gXj = np.zeros(mj.ndof["X"])
    _lib.lib.mh_get_vjp(outj.ctx.h, _lib.dptr(gXj), None, None)
    xo, worst = 0, 0.
    for k in range(min(4, len(specs))):
        m, s_, l_ = sep[k]
        g = np.zeros(m.ndof["X"])
        plain[k].ctx.check(_lib.lib.mh_get_vjp(plain[k].ctx.h, _lib.dptr(g), None, None), "vjp")
        nd = 6 * (specs[k]["ne"] + 1)
        worst = max(worst, float(np.abs(gXj[xo:xo + nd] - g[:nd]).max()))
        xo += nd
    cpu = None
    if not args.no_cpu_baseline:
        spb, ns, dts = cpu_baseline_gradient(mj, sj, lj, args.cpu_sample_gradient)
        cpu = {"value": len(specs) / (spb * nbeam), "unit": "gradient sets/s", "cores": 1, "kind": "port",
               "sample": f"first {ns} beams of the batch ({dts:.1f} s): per beam the nested-dual Jacobian "
                         f"(diffed_residual) contracted with Lambda, scaled to the {nbeam} beams"}
        log(f"adjoint cpu baseline: {cpu['value']:.3g} gradient sets/s")
    return {"workload": f"{len(specs)} cantilevers of U{{{args.model_min}..{args.model_max}}} EulerBeam3D{{true}} "
                        f"({nbeam} beams), 8 U-loads each; gradient set = dL/dX, dL/dU of L = sum Lambda.R",
            "cpu_baseline": cpu,
            "metric": "batched adjoint grads/s vs plain",
            "batched": res["batched"], "plain": res["plain"],
            "speedup": res["batched"]["grads_per_s"] / res["plain"]["grads_per_s"],
            "batched_kernels_ms": dict(zip(("elements", "scatter_X", "scatter_U"), ms.tolist())),
            "batched_vs_plain_max_abs_diff": worst}


def catenary_motion_bench(args):
    """The catenary chains in time (SweepX{2}, Newmark-β with β = 1/4, γ = 1/2, dt = 0.05 s): each
    chain's top moved horizontally, a·sin(2πt/T) with a ~ U(0.5, 3) m, T ~ U(6, 15) s, from rest
    in the as-meshed shape. Whole Newton iterations (assembly, device chain solve, device Newmark
    decrement) replayed from HIP graphs; X stays on the device through the steps and is fetched
    at the last one (savestep). `--motion-steps` steps are timed (the reference runs 1000)."""
    import muscade_amd as mc
    from muscade_amd.workloads import catenary_chains
    t0 = time.perf_counter()
    model, state = catenary_chains(args.chains, args.chain_min, args.chain_max, disp=0., rot=0., vel=0., acc=0.,
                                   motion=True)
    state.time = 0.
    nbeam = model.eletypes[0].nele
    log(f"catenary motion: {args.chains} chains, {nbeam} beams ({time.perf_counter() - t0:.1f}s)")
    dt = 0.05
    times = dt * np.arange(1, args.motion_steps + 1)
    st = {}
    t0 = time.perf_counter()
    out = mc.solve(mc.SweepX[2], initialstate=state, time=times, maxΔx=1e-5, maxiter=20, verbose=args.motion_progress,
                   catcherror=False, linsolve="device", graph=not args.eager, stats=st, savestep=args.motion_steps)
    el = time.perf_counter() - t0
    top_nodes = np.concatenate([E.cat_nodes()[:, 0] + 1 for E in model.eletypes if E.cls is mc.DofMotion])
    tops = model.getdofID("X", "t1", nodID=top_nodes)
    last = out[-1]
    return {"workload": f"catenary chains x{args.chains} ({nbeam} EulerBeam3D), top moved a*sin(2*pi*t/T), "
                        f"SweepX{{2}} dt=0.05 s from rest, device chain solve, whole Newton iterations as HIP graphs",
            "steps": int(args.motion_steps), "seconds_incl_prepare": el, "prepare_s": st.get("prepare_s"),
            "steps_per_s": args.motion_steps / (el - st.get("prepare_s", 0.)),
            "newton_iterations": st.get("iterations"), "graph_captures": st.get("graph_captures"),
            "iterations_per_step": st.get("iterations", 0) / args.motion_steps,
            "projected_1000_steps_s": st.get("prepare_s", 0.) + (el - st.get("prepare_s", 0.)) / args.motion_steps * 1000,
            "top_t1_at_last_step_m": {"min": float(last.X[0][tops - 1].min()), "max": float(last.X[0][tops - 1].max())}}


def design_point_bench(args):
    """Batched design-point search (the limit-state gradients of many models): `--models`
    tensioned risers, vertical cantilevers clamped at the bottom with a 1 MN dead top tension
    and the top held in twist (U{model_min..model_max} EulerBeam3DAU elements of 0.1 m, 16 stiffness parameters a as
    10^a on EA, EI₂, EI₃, GJ per sixteenth, 8 segments of U-loads each), at their static
    equilibrium (SweepX{0} with the device chain solve in 4 load steps, timed apart). A gradient set is dg/dA
    (16) and dg/dU (24) of one model's limit state g = lateral top displacement: assembly of
    the tangent at X*, the adjoint chain solve K·μ = ∂g/∂X and the adjoint launch with Λ = −μ.
    Batched: all models as one model (one of each launch). Plain: one context per model, in a
    loop, at the same equilibrium."""
    import muscade_amd as mc
    from muscade_amd.workloads import design_point_batch, design_point_equilibrium, design_point_specs
    specs = design_point_specs(args.models, args.model_min, args.model_max)
    t0 = time.perf_counter()
    [(mj, sj, tipj)] = design_point_batch(specs)
    sep = design_point_batch(specs, joined=False)
    nbeam = mj.eletypes[0].nele
    log(f"design point: {len(specs)} models, {nbeam} beams ({time.perf_counter() - t0:.1f}s)")
    st = {}
    t0 = time.perf_counter()
    eq = design_point_equilibrium(sj, stats=st)
    t_eq = time.perf_counter() - t0
    log(f"design point: equilibrium {t_eq:.2f}s ({st.get('iterations')} Newton iterations)")
    # the plain models at the same equilibrium: node dofs model after model, then the λ fields
    # (one λ of each Hold per model, field after field)
    nodal = np.cumsum([0] + [6 * (s["ne"] + 1) for s in specs])
    nm = len(specs)
    nlam = (mj.ndof["X"] - nodal[-1]) // nm
    plain = []
    for m, (mm, sm, tm) in enumerate(sep):
        s_ = sm.copy()
        s_.X[0][:6 * (specs[m]["ne"] + 1)] = eq.X[0][nodal[m]:nodal[m + 1]]
        s_.X[0][6 * (specs[m]["ne"] + 1):] = eq.X[0][nodal[-1] + m + nm * np.arange(nlam)]
        plain.append((mc.LimitStateGradient(mm, s_.dis), s_, (np.asarray(tm), np.ones(1))))
    dgj = np.zeros(mj.ndof["X"])
    dgj[tipj - 1] = 1.
    sparse_j = (tipj, np.ones(tipj.size))       # ∂g/∂X by its nonzeros: one per model
    lsg = mc.LimitStateGradient(mj, eq.dis)
    res = {}

    # timed with the equilibrium state resident on the device (uploaded by the warm-up call),
    # and once more with the state's upload (X, U, A from the host) in each call
    def batched(upload=False):
        return lsg(eq, sparse_j, upload=upload, check=False)

    def per_model(upload=False):
        return [f(s_, d_, upload=upload, check=False) for f, s_, d_ in plain]

    for name, fn in (("batched", batched), ("plain", per_model)):
        fn(upload=True)
        for _ in range(max(0, args.warmup - 1)):
            fn()
        n = max(2, args.steps // 4)
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        el = (time.perf_counter() - t0) / n
        t0 = time.perf_counter()
        fn(upload=True)
        elu = time.perf_counter() - t0
        res[name] = {"grads_per_s": nm / el, "ms_per_batch": el * 1e3, "ms_per_batch_with_state_upload": elu * 1e3}
        log(f"design point {name}: {el * 1e3:.2f} ms for {nm} gradient sets ({elu * 1e3:.2f} ms with the state upload)")
    gA, gU, gX = lsg(eq, dgj)
    worst, worst_x = 0., float(np.abs(gX + dgj).max())
    for m in range(min(8, nm)):
        a1, u1, _ = plain[m][0](plain[m][1], plain[m][2], upload=True, check=False)
        worst = max(worst, float(np.abs(gA[16 * m:16 * m + 16] - a1).max() / np.abs(a1).max()),
                    float(np.abs(gU[24 * m:24 * m + 24] - u1).max() / np.abs(u1).max()))
    cpu = None
    if not args.no_cpu_baseline:
        lam = np.zeros(mj.ndof["X"])
        lam[:] = np.random.default_rng(0).normal(size=mj.ndof["X"])
        spb, ns, dts = cpu_baseline_gradient(mj, eq, lam, args.cpu_sample_gradient)
        cpu = {"value": nm / (spb * nbeam), "unit": "gradient sets/s", "cores": 1, "kind": "port",
               "sample": f"first {ns} beams of the batch ({dts:.1f} s): the element part only (per beam the "
                         f"nested-dual Jacobian contracted with Lambda, scaled to the {nbeam} beams); the adjoint "
                         f"solve is not included, so this bounds the CPU rate from above"}
        log(f"design point cpu baseline: {cpu['value']:.3g} gradient sets/s (element part)")
    return {"workload": f"{nm} tensioned risers (vertical cantilevers, 1 MN top tension), "
                        f"U{{{args.model_min}..{args.model_max}}} EulerBeam3DAU of 0.1 m ({nbeam} beams), 16 stiffness "
                        f"parameters and 8 U-load segments each; gradient set = dg/dA, dg/dU of the lateral top "
                        f"displacement at equilibrium (|dx| <= 1e-6 m)",
            "metric": "batched adjoint grads/s vs plain",
            "batched": res["batched"], "plain": res["plain"],
            "speedup": res["batched"]["grads_per_s"] / res["plain"]["grads_per_s"],
            "equilibrium_s": t_eq, "equilibrium_iterations": st.get("iterations"),
            "batched_vs_plain_max_rel_diff": worst, "adjoint_check_max_abs": worst_x, "cpu_baseline": cpu,
            "tip_displacement_m": {"min": float(eq.X[0][tipj - 1].min()), "max": float(eq.X[0][tipj - 1].max())}}


def gauge_identification_bench(args):
    """The all-steps adjoint gradient of a load and stiffness identification from strain
    gauges, DirectXUA{2,0,1}: one catenary of `--gauge-nel` A-scaled beams (the chains' section
    and shape, EI scaled by 10^a per segment, 64 segments), a distributed load per length on
    every node (3 U-dofs per element), 50 bending-strain gauges (ElementCost, σ = 1e-6) over
    `--gauge-steps` steps of 0.05 s, priors on loads and parameters. The measured strains come
    from a forward SweepX{2} run with the true parameters and load (device chain solve), plus
    N(0, 1e-6) noise. Timed: L and ∂L/∂Λ, ∂L/∂X, ∂L/∂U, ∂L/∂A over all steps (AllStepsGradient:
    one assembly and one adjoint launch for every step's beams, the host Lagrangians per step),
    with central differences (h = 1e-6) at a few entries of each class."""
    import muscade_amd as mc
    from muscade_amd.results import ElementResults
    from muscade_amd.workloads import gauge_identification, gauge_identification_truth, gauge_strains
    nel, nstep, dt, nseg, ng = args.gauge_nel, args.gauge_steps, 0.05, 64, 50
    t0 = time.perf_counter()
    fwd, A, U, states = gauge_identification_truth(nel, nseg, nstep, dt, linsolve="device", graph=not args.eager)
    t_fwd = time.perf_counter() - t0
    log(f"gauge identification: forward SweepX{{2}} of {nel} beams over {nstep} steps {t_fwd:.1f}s")
    model0, gidx = gauge_identification(nel, nseg, ng, targets=np.zeros((nstep, ng)))
    st = mc.initialize(model0)
    it = next(i for i, E in enumerate(model0.eletypes) if E.cls is mc.ElementCost)
    G = np.concatenate([p["gauges"] for p in model0.eletypes[it].params])
    res_ = ElementResults(model0, st.dis)
    meas = []
    for s_ in states:
        st.X[0][:] = s_.X[0]
        v, _, _ = res_.gauge_derivatives(it, st)
        meas.append(gauge_strains(v, G)[:, 0])
    rng = np.random.default_rng(1003)
    meas = np.array(meas) + rng.normal(0., 1e-6, (nstep, ng))
    model, _ = gauge_identification(nel, nseg, ng, targets=meas)
    dis = mc.initialize(model).dis
    nX = model.ndof["X"]
    time_ = dt * np.arange(1, nstep + 1)
    X = np.array([s_.X[0] for s_ in states])
    Λ = rng.normal(0., 1e-3, (nstep, nX))
    Uv = np.tile(0.9 * U, (nstep, 1))
    Av = np.zeros(model.ndof["A"])
    t0 = time.perf_counter()
    grad = mc.AllStepsGradient(2, 0, 1, model, dis, nstep, dt)
    t_prep = time.perf_counter() - t0
    L, g = grad(Λ, X, Uv, Av, time_)
    n = max(2, args.steps // 8)
    t0 = time.perf_counter()
    dev = host = 0.
    for _ in range(n):
        L, g = grad(Λ, X, Uv, Av, time_)
        dev += grad.timings["device_s"]
        host += grad.timings["host_s"]
    el = (time.perf_counter() - t0) / n
    log(f"gauge identification: gradient {el:.3f}s (device part {dev / n:.3f}s, host Lagrangians {host / n:.3f}s)")
    # h = 1e-6 relative to the entry (at least 1e-6); the rounding of L bounds what a difference
    # quotient can resolve: about eps·|L|/h, reported beside the error
    fd_err = {}
    for cls, k, i in (("X", nstep // 2, 6 * (nel // 3) + 4), ("U", nstep // 3, 3 * (nel // 2) + 1), ("A", None, nseg // 2)):
        base = Av[i] if cls == "A" else {"X": X, "U": Uv}[cls][k, i]
        h = 1e-6 * max(1., abs(base))
        fd = 0.
        for sgn in (1., -1.):
            v = {"Λ": Λ, "X": X.copy(), "U": Uv.copy(), "A": Av.copy()}
            if cls == "A":
                v["A"][i] += sgn * h
            else:
                v[cls][k, i] += sgn * h
            fd += sgn * grad(v["Λ"], v["X"], v["U"], v["A"], time_)[0]
        fd /= 2 * h
        got = g["A"][i] if cls == "A" else g[cls][k, i]
        cmax = np.abs(g[cls]).max()
        fd_err[cls] = {"h": h, "adjoint": float(got), "central_difference": float(fd),
                       "err_vs_class_max": float(abs(got - fd) / cmax),
                       "rounding_floor_vs_class_max": float(np.finfo(float).eps * abs(L) / h / cmax)}
    log(f"gauge identification: central differences {fd_err}")
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, ROOT)
        from tests import _oracle as orc
        E, d = model.eletypes[0], dis.dis[0]
        xyz, nodes = model.coords(), E.cat_nodes()
        o2 = np.concatenate([p["orient2"] for p in E.params])
        mat = E.cls.scaled_section(np.concatenate([p["mat"] for p in E.params]), Av[d.A[:, 0] - 1],
                                   E.params[0]["amask"])
        Xd = mc.directxua.step_derivatives(X, 2, dt)
        k = nstep // 2
        ns = min(args.cpu_sample_dynamic, nel)
        t0 = time.perf_counter()
        for e in range(ns):
            idx = d.X[e] - 1
            R, Ks, KU = orc.beam_diffed_residual(xyz[nodes[e, 0], :3], xyz[nodes[e, 1], :3], o2[e], mat[e],
                                                 [x[k][idx] for x in Xd], Uv[k][d.U[e] - 1])
            gx = [Λ[k][idx] @ K_ for K_ in Ks]
        dts = time.perf_counter() - t0
        per = dts / ns
        cpu = {"value": 1. / (per * nel * nstep), "unit": "gradients/s", "cores": 1, "kind": "port",
               "sample": f"{ns} beams of step {k} ({dts:.1f} s): per beam the nested-dual Jacobians (X, X', X'', U) "
                         f"contracted with Lambda, scaled to {nel} beams x {nstep} steps; the host Lagrangians are "
                         f"not included, so this bounds the CPU rate from above"}
        log(f"gauge identification cpu baseline: {cpu['value']:.3g} gradients/s (element part)")
    return {"workload": f"load and stiffness identification from 50 strain gauges, DirectXUA{{2,0,1}}: "
                        f"{nel} EulerBeam3DAU catenary, {nseg} EI parameters, 3 load dofs per node, {nstep} steps",
            "unknowns": nstep * (2 * nX + model.ndof["U"]) + model.ndof["A"],
            "forward_sweepx_s": t_fwd, "prepare_s": t_prep, "gradients": n, "s_per_gradient": el,
            "device_s_per_gradient": dev / n, "host_s_per_gradient": host / n, "L": L,
            "central_differences": fd_err, "cpu_baseline": cpu}


def graph_bench(args):
    """Per-assembly time of small models (launch-bound), eager launches vs one captured HIP
    graph replayed (the Newton iterations of a time step): the 100-beam cantilever of the
    reference's CPU case and a 2000-beam one. Device-resident state, the NaN-flag readback
    included, outputs left on the device."""
    import muscade_amd as mc
    from muscade_amd import _lib
    from muscade_amd.workloads import cantilever
    res = {}
    for nel in (100, 2000):
        model = cantilever(nel=nel)
        state = mc.initialize(model)
        state.X[0][:] = np.random.default_rng(3).normal(size=model.ndof["X"]) * 1e-3
        state.time = 1.0
        row = {}
        for mode in ("eager", "graph"):
            out = mc.AssemblySweepX(0, model, state.dis, graph=(mode == "graph"))
            out.set_state(state, 1.0)
            fn = _lib.lib.mh_assemble_sweepx_graph if mode == "graph" else _lib.lib.mh_assemble_sweepx
            err = _lib.mh_err()
            c6 = out.c.array()
            for _ in range(20):
                fn(out.ctx.h, 0, 1, _lib.dptr(c6), ctypes.byref(err))
            n = 2000
            t0 = time.perf_counter()
            for _ in range(n):
                if fn(out.ctx.h, 0, 1, _lib.dptr(c6), ctypes.byref(err)) != 0:
                    raise RuntimeError(_lib.lib.mh_last_error(out.ctx.h).decode())
            row[mode + "_us"] = (time.perf_counter() - t0) / n * 1e6
            del out
        row["speedup"] = row["eager_us"] / row["graph_us"]
        res[f"cantilever_{nel}"] = row
        log(f"graph: {nel} beams: eager {row['eager_us']:.1f} us, graph {row['graph_us']:.1f} us per assembly")
    # whole device Newton iterations of SweepX{2} (assembly, chain solve, decrement) on one
    # 100-beam hanging chain with a moving top: eager, graphed assembly, one graph per iteration
    from muscade_amd.workloads import catenary_chains
    model, state = catenary_chains(1, 100, 100, seed=4, disp=0., rot=0., vel=0., acc=0., motion=True)
    state.time = 0.
    row = {}
    for name, g in (("eager", False), ("assembly_graph", "assembly"), ("iteration_graph", True)):
        kw = dict(initialstate=state, maxΔx=1e-6, maxiter=20, verbose=False, catcherror=False, linsolve="device",
                  graph=g)
        mc.solve(mc.SweepX[2], time=0.05 * np.arange(1, 4), **kw)       # warm-up
        st = {}
        t0 = time.perf_counter()
        mc.solve(mc.SweepX[2], time=0.05 * np.arange(1, 101), stats=st, **kw)
        el = time.perf_counter() - t0 - st.get("prepare_s", 0.)
        row[name + "_us_per_iteration"] = el / st["iterations"] * 1e6
        row["iterations"] = st["iterations"]
    row["speedup"] = row["eager_us_per_iteration"] / row["iteration_graph_us_per_iteration"]
    res["newton_chain_100"] = row
    log(f"graph: Newton iteration, 100-beam chain: {row}")
    return res


def newton_bench(args):
    """One Newton step of SweepX{0} (SweepX.jl:204-210) on a reduced lattice: device
    assembly, download of Lλ and Lλx, host sparse LU (SciPy SuperLU, as the reference's
    host `lu`) and back-substitution, each timed. Host factorization of the full
    100x100x33 frame is out of reach for SuperLU (its fill grows from 14M entries at
    20x20x6 to 115M at 30x30x10); the factorization stays on the host and is reported
    apart from the assembly."""
    import scipy.sparse.linalg as spla
    import muscade_amd as mc
    from muscade_amd.workloads import space_frame
    nx, ny, nz = 20, 20, 6
    model, state = space_frame(nx, ny, nz, seed=1001)
    out = mc.AssemblySweepX(0, model, state.dis)
    out.assemble(state, "iter")   # warm-up
    t0 = time.perf_counter()
    out.assemble(state, "iter", fetch=False)
    t1 = time.perf_counter()
    out.assemble(state, "iter")
    t2 = time.perf_counter()
    K = out.Lλx()
    t3 = time.perf_counter()
    lu = spla.splu(K)
    t4 = time.perf_counter()
    lu.solve(out.Lλ)
    t5 = time.perf_counter()
    res = {"workload": f"static space frame {nx}x{ny}x{nz} nodes, {model.eletypes[0].nele} EulerBeam3D, nX={out.nX}, "
                       f"nnz={out.nnz}", "assembly_ms": (t1 - t0) * 1e3, "assembly_with_download_ms": (t2 - t1) * 1e3,
           "host_lu_ms": (t4 - t3) * 1e3, "host_solve_ms": (t5 - t4) * 1e3, "lu_fill": int(lu.L.nnz + lu.U.nnz)}
    log(f"newton step ({nx}x{ny}x{nz}): assembly {res['assembly_ms']:.2f} ms, with download "
        f"{res['assembly_with_download_ms']:.2f} ms, host LU {res['host_lu_ms']:.0f} ms, solve {res['host_solve_ms']:.1f} ms")
    return res


def directxua_bench(args):
    """solve(DirectXUA{2,0,0}) of a load-identification problem: a 100-beam curved cantilever over
    40 steps, a U-load at the tip and a cost on the tip's deviation from a target motion.
    Residual elements per step on the GPU, one launch per step or all steps in one launch; the
    all-steps system assembled on the host or on the device; its LU on the host (SciPy) or on the
    device (dense block LU of the steps)."""
    import muscade_amd as mc
    from muscade_amd.workloads import chain_section
    nel, nstep = 100, 40
    model = mc.Model("directxua")
    x = np.linspace(0., 20., nel + 1)
    mc.addnode(model, np.stack([x, 0 * x, 0.02 * x * x], axis=1))
    mc.addelement(model, mc.EulerBeam3D, np.stack([np.arange(1, nel + 1), np.arange(2, nel + 2)], axis=1),
                  mat=chain_section(), orient2=np.array([0., 1., 0.]))
    for f in ("t1", "t2", "t3", "r1", "r2", "r3"):
        mc.addelement(model, mc.Hold, [1], field=f)
    # the load prior's Hessian (2e-6) must stay above DirectXUA's sparser! threshold, 1e-20 of the
    # largest entry of the all-steps matrix (about 3e10 here), or the U block loses its diagonal
    mc.addelement(model, mc.SingleUdof, [nel + 1], Xfield="t2", Ufield="ut2", cost=lambda u, t: 1e-6 * u * u)
    mc.addelement(model, mc.SingleDofCost, [nel + 1], dofclass="X", field="t2",
                  cost=lambda x, t: 1e3 * (x - 1e-2 * np.sin(t)) ** 2)
    st0 = mc.initialize(model, time=0.)
    time_ = np.arange(1, nstep + 1) * 0.05
    res = {"workload": f"load identification, {nel} EulerBeam3D x {nstep} steps, DirectXUA{{2,0,0}}",
           "unknowns": nstep * (2 * model.ndof["X"] + model.ndof["U"])}
    finals = {}
    for name, mode, big, lin in (("launch_per_step", False, "host", "host"),
                                 ("steps_in_one_launch", True, "host", "host"),
                                 ("steps_in_one_launch_big_system_on_device", True, "device", "host"),
                                 ("steps_in_one_launch_big_system_and_solve_on_device", True, "device", "device")):
        st = {}
        t0 = time.perf_counter()
        out = mc.solve(mc.DirectXUA[2, 0, 0], initialstate=[st0], time=[time_], maxiter=30, verbose=False,
                       catcherror=False, batch_steps=mode, big=big, linsolve=lin, stats=st)   # tolerances 1e-5
        st["total_s"] = time.perf_counter() - t0
        finals[name] = np.array([s_.U[0][0] for s_ in out[0]])
        res[name] = st
        log(f"directxua ({name}): {st}")
    ref = finals["launch_per_step"]
    res["max_rel_diff_of_identified_loads"] = {k: float(np.abs(v - ref).max() / np.abs(ref).max())
                                               for k, v in finals.items() if k != "launch_per_step"}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=100)
    ap.add_argument("--ny", type=int, default=100)
    ap.add_argument("--nz", type=int, default=33)
    ap.add_argument("--cpu-sample", type=int, default=40000)
    ap.add_argument("--cpu-sample-parallel", type=int, default=200000)
    ap.add_argument("--cpu-sample-dynamic", type=int, default=4000)
    ap.add_argument("--cpu-sample-gradient", type=int, default=20000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calib-copy", action="store_true", help="run only the counter-calibration copy kernel")
    ap.add_argument("--no-dynamic", action="store_true", help="skip the dynamic (SweepX{2}) assembly measurement")
    ap.add_argument("--only-dynamic", action="store_true", help="measure only the dynamic assembly")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the batched vs plain adjoint measurement")
    ap.add_argument("--no-graph", action="store_true", help="skip the eager vs HIP-graph small-model measurement")
    ap.add_argument("--no-newton", action="store_true", help="skip the reduced-lattice Newton step (host LU) timing")
    ap.add_argument("--no-directxua", action="store_true", help="skip the DirectXUA load-identification solve")
    ap.add_argument("--only-directxua", action="store_true", help="measure only the DirectXUA solve")
    ap.add_argument("--only-graph", action="store_true", help="measure only eager vs HIP-graph assembly")
    ap.add_argument("--no-chain-solve", action="store_true",
                    help="skip the device linear solve measurements (chain topologies)")
    ap.add_argument("--only-chain-solve", action="store_true", help="measure only the device vs host chain solve")
    ap.add_argument("--only-adjoint", action="store_true", help="measure only the batched adjoint gradients")
    ap.add_argument("--no-design-point", action="store_true", help="skip the batched limit-state gradients")
    ap.add_argument("--no-motion", action="store_true", help="skip the catenary chains run in time")
    ap.add_argument("--only-motion", action="store_true", help="measure only the catenary chains run in time")
    ap.add_argument("--motion-steps", type=int, default=10)
    ap.add_argument("--motion-progress", action="store_true",
                    help="print the solver's line per converged step (long motion runs)")
    ap.add_argument("--only-design-point", action="store_true", help="measure only the batched limit-state gradients")
    ap.add_argument("--eager", action="store_true",
                    help="SweepX solves with eager launches instead of replayed HIP graphs (for tracing tools)")
    ap.add_argument("--no-gauge", action="store_true", help="skip the gauge identification gradient")
    ap.add_argument("--only-gauge", action="store_true", help="measure only the gauge identification gradient")
    ap.add_argument("--gauge-nel", type=int, default=5000)
    ap.add_argument("--gauge-steps", type=int, default=200)
    ap.add_argument("--models", type=int, default=512)
    ap.add_argument("--model-min", type=int, default=500)
    ap.add_argument("--model-max", type=int, default=20000)
    ap.add_argument("--chains", type=int, default=256)
    ap.add_argument("--chain-min", type=int, default=1000)
    ap.add_argument("--chain-max", type=int, default=50000)
    args = ap.parse_args()
    if args.calib_copy:
        calib_copy()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "muscade-hip measures one GPU; runs on several GPUs are not implemented",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.only_dynamic:
        print(json.dumps({"dynamic": dynamic_bench(args)}))
        return
    if args.only_adjoint:
        print(json.dumps({"adjoint": adjoint_bench(args)}))
        return
    if args.only_motion:
        print(json.dumps({"catenary_motion": catenary_motion_bench(args)}))
        return
    if args.only_design_point:
        print(json.dumps({"design_point": design_point_bench(args)}))
        return
    if args.only_gauge:
        print(json.dumps({"gauge_identification": gauge_identification_bench(args)}))
        return
    if args.only_graph:
        print(json.dumps({"graph": graph_bench(args)}))
        return
    if args.only_chain_solve:
        print(json.dumps({"chain_solve": chain_solve_bench(args)}))
        return
    if args.only_directxua:
        print(json.dumps({"directxua": directxua_bench(args)}))
        return
    import muscade_amd as mc
    from muscade_amd import _lib
    from muscade_amd.workloads import space_frame

    t0 = time.perf_counter()
    model, state = space_frame(args.nx, args.ny, args.nz, seed=1001)
    nbeam = model.eletypes[0].nele
    nhold = sum(E.nele for E in model.eletypes[1:])
    log(f"model: {nbeam} beams, {nhold} holds, nX={model.ndof['X']} ({time.perf_counter() - t0:.1f}s)")
    t0 = time.perf_counter()
    out = mc.AssemblySweepX(0, model, state.dis)
    t_prep = time.perf_counter() - t0
    log(f"prepare (device pattern, nnz={out.nnz}): {t_prep:.2f}s")
    tv, tms = ctypes.c_int(), (ctypes.c_float * 16)()
    _lib.lib.mh_get_tuning(out.ctx.h, ctypes.byref(tv), tms)
    tuning = {"scatter_variant": tv.value, "scatter_variant_ms": [round(float(x), 4) for x in tms]}
    log(f"scatter tuner: {tuning}")
    out.set_state(state, 0.0)
    h = out.ctx.h
    c6 = out.c.array()
    err = _lib.mh_err()
    _lib.lib.mh_assemble_sweepx(h, 0, 1, _lib.dptr(c6), ctypes.byref(err))   # the first assembly tunes the beam kernel
    bb, bms = ctypes.c_int(), (ctypes.c_float * 3)()
    if _lib.lib.mh_get_beam_tuning(h, ctypes.byref(bb), bms) == 0:
        tuning.update(beam_block=bb.value, beam_block_ms=[round(float(x), 4) for x in bms])
        log(f"beam kernel tuner: block {bb.value}, ms {list(bms)}")

    def step():
        r = _lib.lib.mh_assemble_sweepx(h, 0, 1, _lib.dptr(c6), ctypes.byref(err))
        if r != 0:
            raise RuntimeError(_lib.lib.mh_last_error(h).decode())

    for _ in range(args.warmup):
        step()
    _lib.lib.mh_sync(h)
    ms = np.zeros(3)
    acc = np.zeros(3)
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        _lib.lib.mh_last_timings(h, _lib.dptr(ms))
        acc += ms
        if (i + 1) % max(1, args.steps // 4) == 0:
            log(f"  step {i + 1}/{args.steps}")
    _lib.lib.mh_sync(h)
    elapsed = time.perf_counter() - t0
    ms_step = elapsed / args.steps * 1e3
    value = nbeam * args.steps / elapsed
    t_eval, t_sk, t_sr = acc / args.steps * 1e-3   # s per launch group
    N_c = sum(E.nele * len(E.class_columns("X")) ** 2 for E in model.eletypes)
    N_cR = sum(E.nele * len(E.class_columns("X")) for E in model.eletypes)
    nKbuf = nbeam * 45 + sum(E.nele * len(E.class_columns("X")) ** 2 for E in model.eletypes[1:])
    b_eval, b_sk, b_sr = algorithmic_bytes(nbeam, nhold, out.nnz, out.nX, N_c, N_cR, nKbuf)
    rl_eval_hbm = {"bound": "hbm", "achieved": b_eval / t_eval / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s"}
    rl_scatter = {"bound": "hbm", "achieved": b_sk / t_sk / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s"}
    if F_ELEM:
        rl_eval = {"bound": "fp64", "achieved": F_ELEM * nbeam / t_eval / 1e12, "peak": FP64_PEAK / 1e12,
                   "unit": "TFLOP/s"}
    else:
        rl_eval = rl_eval_hbm
    traffic = {}
    if (args.nx, args.ny, args.nz) == (100, 100, 33) and os.path.exists(TRAFFIC_FILE):
        traffic = json.load(open(TRAFFIC_FILE))["bytes_per_launch"]
    for r_, key in ((rl_eval, "eval"), (rl_scatter, "scatter"), (rl_eval_hbm, "eval")):
        r_["frac"] = r_["achieved"] / r_["peak"]
        r_["traffic"] = traffic.get(key)
    dominant = ("eval", rl_eval) if t_eval >= t_sk else ("scatter", rl_scatter)
    log(f"per step: eval {t_eval * 1e3:.3f} ms, nz scatter {t_sk * 1e3:.3f} ms, rhs scatter {t_sr * 1e3:.3f} ms, "
        f"wall {ms_step:.3f} ms")
    # PCIe-inclusive rate (not `value`): state uploaded from host memory, rhs and nzval
    # downloaded to host memory, every step
    nz_host, rhs_host = np.zeros(out.nnz), np.zeros(out.nX)
    x_host = np.ascontiguousarray(state.X[0], dtype=np.float64)
    npc = max(2, min(5, args.steps))
    t0 = time.perf_counter()
    for _ in range(npc):
        _lib.lib.mh_set_X(h, 0, _lib.dptr(x_host))
        step()
        _lib.lib.mh_get_rhs(h, _lib.dptr(rhs_host))
        _lib.lib.mh_get_nzval(h, _lib.dptr(nz_host))
    t_pc = (time.perf_counter() - t0) / npc
    pcie = {"ms_per_step": t_pc * 1e3, "elements_per_s": nbeam / t_pc,
            "note": "H2D of X (16 MB) + assembly + D2H of rhs and nzval (673 MB) into pageable host memory"}
    log(f"PCIe-inclusive: {t_pc * 1e3:.2f} ms/step")
    # the same assemble!{:iter} replayed from a captured HIP graph (the Newton iterations of a
    # time step): bitwise the eager result; reported beside `value`, which is the eager path
    def gstep():
        r = _lib.lib.mh_assemble_sweepx_graph(h, 0, 1, _lib.dptr(c6), ctypes.byref(err))
        if r != 0:
            raise RuntimeError(_lib.lib.mh_last_error(h).decode())
    for _ in range(max(2, args.warmup)):
        gstep()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gstep()
    _lib.lib.mh_sync(h)
    t_g = (time.perf_counter() - t0) / args.steps
    graph_replay = {"ms_per_step": t_g * 1e3, "elements_per_s": nbeam / t_g}
    log(f"graph replay: {t_g * 1e3:.3f} ms/step")
    cpu = cpu1 = None
    if not args.no_cpu_baseline:
        nth = host_threads()
        log(f"cpu baseline (oracle, {nth} threads)...")
        eps, n, dt = cpu_baseline(model, state, args.cpu_sample_parallel, nth)
        cpu = {"value": eps, "unit": "elements/s", "cores": nth, "kind": "port",
               "sample": f"first {n} beams of the same 100x100x33 lattice, one assemble!{{:iter}} ({dt:.1f} s): "
                         f"nested forward-mode duals per element on {nth} OpenMP threads, then add_value!/add_∂! "
                         f"in element order"}
        log("cpu baseline (oracle, 1 thread)...")
        eps1, n1, dt1 = cpu_baseline(model, state, args.cpu_sample, 1)
        cpu1 = {"value": eps1, "unit": "elements/s", "cores": 1, "kind": "port",
                "sample": f"first {n1} beams of the same 100x100x33 lattice, one sequential assemble!{{:iter}} "
                          f"({dt1:.1f} s): nested forward-mode duals + sequential add_value!/add_∂! "
                          f"(Muscade's assemble! is single-threaded)"}
    line = {
        "metric": "EulerBeam3D elements assembled/s (FP64 R+K)",
        "value": value,
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"static space frame {args.nx}x{args.ny}x{args.nz} nodes, {nbeam} EulerBeam3D + "
                               f"{nhold} Hold, residual+tangent assembly (AssemblySweepX{{0}}, :iter)",
                   "nbeam": nbeam, "nX": out.nX, "nnz": out.nnz},
        "roofline": dict(dominant[1], kernel=dominant[0]),
        "kernels_ms": {"eval": t_eval * 1e3, "scatter_nz": t_sk * 1e3, "scatter_rhs": t_sr * 1e3},
        "roofline_eval": rl_eval,
        "roofline_eval_hbm": rl_eval_hbm,
        "roofline_scatter": rl_scatter,
        "prepare_s": t_prep,
        "tuning": tuning,
        "cpu_baseline": cpu,
        "cpu_baseline_1thread": cpu1,
        "pcie_inclusive": pcie,
        "graph_replay": graph_replay,
    }
    del out
    from muscade_amd.model import MuscadeError
    # secondary sections: a solver error in one is reported in its place and does not stop the line
    for flag, key, fn in (("no_dynamic", "dynamic", dynamic_bench), ("no_adjoint", "adjoint", adjoint_bench),
                          ("no_design_point", "design_point", design_point_bench),
                          ("no_gauge", "gauge_identification", gauge_identification_bench),
                          ("no_motion", "catenary_motion", catenary_motion_bench), ("no_graph", "graph", graph_bench),
                          ("no_newton", "newton_step", newton_bench), ("no_chain_solve", "chain_solve", chain_solve_bench),
                          ("no_directxua", "directxua", directxua_bench)):
        if getattr(args, flag):
            continue
        try:
            line[key] = fn(args)
        except (MuscadeError, RuntimeError) as e:
            log(f"{key}: {type(e).__name__}: {e}")
            line[key] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
