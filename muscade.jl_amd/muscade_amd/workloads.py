"""Synthetic models of the measured workloads (all inputs seeded, numpy PCG64).

space_frame: the static assembly-at-scale case — a 3D lattice of nx×ny×nz
nodes at spacing h, positions jittered by U(-1%,+1%)·h, EulerBeam3D members
along x, then y, then z (each in lexicographic order), section stiffnesses
× LogNormal(0, 0.2) per member, the base layer (k=0) clamped by six Holds
per node, random displacements N(0,1e-3 m) and rotations N(0,1e-2 rad).
100×100×33 nodes gives 973,400 beams and 2,040,000 X-dofs.

cantilever: the reference's static SweepX{0} case — 100 beams along x,
L=10 m, clamped at node 1, tip load ramped to 5e4 N in 10 steps.

catenary_chains: the dynamic SweepX{2} case — independent mooring-line-like
chains of EulerBeam3D (n_e ~ U{nmin..nmax} elements, horizontal span and depth
from a length ~ U(500,2000) m), EA=5e9 N, EI=2e7 Nm², w=1e3 N/m, with inertia,
added mass, linear and quadratic drag and roll inertia; bottom node pinned by
three Holds, a horizontal DofLoad at the top node; the state is the as-meshed
catenary plus small random displacements, velocities and accelerations.
256 chains with nmin=1000, nmax=50000 give about 6.5 million beams.

cantilever_batch: the batched-adjoint case — many uneven cantilevers
(n_e ~ U{nmin..nmax} EulerBeam3D{true} along x, h = 0.05 m, section stiffness
× LogNormal(0, 0.1) per model), each clamped at its first node and loaded by
8 distributed U-loads (one U-node per eighth of its length, 3 U-dofs each,
N(0, 1) kN/m), with a state of small random displacements and rotations and a
random adjoint vector Λ. The same models are built either one Model each (the
plain path) or all into one Model as disconnected parts (the batch).

gauge_identification: load and stiffness identification from strain gauges (DirectXUA{2,0,1}) —
one catenary of the chains' section and shape, n_e EulerBeam3DA{true} elements whose EI₂ and
EI₃ are scaled by 10^a per segment (nseg A-dofs), a distributed load per length on every node
(3 U-dofs, at each element's first node), anchor and top held, and ngauge bending-strain
gauges (offset 0.15 m, axial direction) on massless, stiffness-free sensors overlaid on beams
along the line, each an ElementCost ((ε − ε_measured)/σ)², σ = 1e-6, with priors 1e-6·|u|² on
the loads and a² on the parameters. The measured strains are those of a forward SweepX{2} run
with a ~ N(0, 0.1/ln 10) and a steady load (0.1, 0.3, 0) kN/m from rest in the as-meshed
shape, plus N(0, 1e-6) noise.

design_point_batch: the batched design-point search — uneven cantilevers of L = 10 m
(n_e ~ U{nmin..nmax} EulerBeam3DA{true} elements), whose section stiffnesses (EA, EI₂, EI₃,
GJ) are scaled per sixteenth of the length by 10^a, a ~ N(0, 0.1/ln 10) (a lognormal
stiffness factor of σ = 0.1), loaded by 8 distributed U-loads (one U-node per eighth, t3
~ N(1, 0.2)·q₀, t1 and t2 ~ N(0, 0.1)·q₀, q₀ = 1 kN/m); the limit state is the tip
displacement along z against a threshold.
"""
import numpy as np

from .elements import (STIFFNESS_FIELDS, BeamCrossSection, DofCost, DofLoad, DofMotion, ElementCost, EulerBeam3D,
                       EulerBeam3DAU, EulerBeam3DU, Hold, SingleAcost, StrainGaugeOnEulerBeam3D)
from .model import Model, MuscadeError, addelement, addnode, initialize, setdof
from .results import request

E_STEEL = 2.1e11
SECTION = dict(A=1e-2, I=1e-5, J=2e-5)
G_STEEL = E_STEEL / 2.6   # chosen: ν = 0.3


def base_section():
    EA = E_STEEL * SECTION["A"]
    EI = E_STEEL * SECTION["I"]
    GJ = G_STEEL * SECTION["J"]
    mu = 7850. * SECTION["A"]
    return BeamCrossSection(EA=EA, EI2=EI, EI3=EI, GJ=GJ, mu=mu, iota1=7850. * SECTION["J"])


def space_frame(nx, ny, nz, seed=1001, h=1.0, jitter=0.01, sigma=0.2, disp=1e-3, rot=1e-2):
    rng = np.random.default_rng(seed)
    model = Model("space_frame")
    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    # lexicographic node numbering with x fastest: node(i,j,k) = i + nx*(j + ny*k)
    ii, jj, kk = I.transpose(2, 1, 0).ravel(), J.transpose(2, 1, 0).ravel(), K.transpose(2, 1, 0).ravel()
    xyz = np.stack([ii, jj, kk], axis=1).astype(float) * h
    xyz += rng.uniform(-jitter * h, jitter * h, size=xyz.shape)
    addnode(model, xyz)
    nid = lambda i, j, k: 1 + i + nx * (j + ny * k)   # noqa: E731
    base = base_section().values
    members = []
    for axis, o2 in ((0, (0., 1., 0.)), (1, (0., 0., 1.)), (2, (0., 1., 0.))):
        a = [I, J, K][axis] < [nx, ny, nz][axis] - 1
        i0, j0, k0 = I[a], J[a], K[a]
        # lexicographic order of the member's first node (k slowest, then j, then i)
        order = np.lexsort((i0, j0, k0))
        i0, j0, k0 = i0[order], j0[order], k0[order]
        d = [0, 0, 0]
        d[axis] = 1
        n1 = nid(i0, j0, k0)
        n2 = nid(i0 + d[0], j0 + d[1], k0 + d[2])
        members.append((np.stack([n1, n2], axis=1), o2))
    for nodes, o2 in members:
        nele = nodes.shape[0]
        mat = np.broadcast_to(base, (nele, 16)).copy()
        mat[:, :4] *= rng.lognormal(0., sigma, size=(nele, 4))
        addelement(model, EulerBeam3D, nodes, mat=mat, orient2=np.array(o2))
    basenodes = nid(I[K == 0], J[K == 0], K[K == 0])
    basenodes = np.sort(basenodes)
    for f in ("t1", "t2", "t3", "r1", "r2", "r3"):
        addelement(model, Hold, basenodes.reshape(-1, 1), field=f)
    state = initialize(model)
    nnod = model.nnod
    allnodes = np.arange(1, nnod + 1)
    for f in ("t1", "t2", "t3"):
        setdof(state, rng.normal(0., disp, nnod), field=f, nodID=allnodes)
    for f in ("r1", "r2", "r3"):
        setdof(state, rng.normal(0., rot, nnod), field=f, nodID=allnodes)
    return model, state


def tipload(t, F0=5e3):
    return F0 * t


def cantilever(nel=100, L=10., F=5e4, nstep=10):
    """Straight cantilever along x, clamped at node 1, tip load along t3 ramped linearly in time."""
    model = Model("cantilever")
    x = np.linspace(0., L, nel + 1)
    addnode(model, np.stack([x, np.zeros_like(x), np.zeros_like(x)], axis=1))
    nodes = np.stack([np.arange(1, nel + 1), np.arange(2, nel + 2)], axis=1)
    addelement(model, EulerBeam3D, nodes, mat=base_section(), orient2=np.array([0., 1., 0.]))
    for f in ("t1", "t2", "t3", "r1", "r2", "r3"):
        addelement(model, Hold, [1], field=f)

    def load(t, F0=F / nstep):
        return F0 * t
    addelement(model, DofLoad, [nel + 1], field="t3", value=load)
    return model


def chain_section():
    """Section of the catenary chains. EA, EI and w are the workload's; μ, ι₁, GJ and the
    hydrodynamic coefficients are chosen (a 0.3 m steel-like line in water)."""
    return BeamCrossSection(EA=5e9, EI2=2e7, EI3=2e7, GJ=1.5e7, mu=1e3 / 9.81, iota1=0.5, w=1e3,
                            Ca1=5., Cl1=1., Cq1=10., Ca2=72., Cl2=5., Cq2=300., Ca3=72., Cl3=5., Cq3=300.)


def topload(t, F0=1e5, T=10.):
    return F0 * np.sin(2. * np.pi * t / T)


def top_motion(t, a=1., T=10.):
    """Prescribed horizontal motion of a chain's top: a·sin(2πt/T)."""
    return a * np.sin(2. * np.pi * t / T)


def catenary_chains(nchain=256, nmin=1000, nmax=50000, seed=1002, disp=1e-3, rot=1e-3, vel=0.1, acc=0.1,
                    motion=False):
    """motion=True: the top of each chain held vertically and transversally and moved
    horizontally, a·sin(2πt/T) with a ~ U(0.5, 3) m and T ~ U(6, 15) s per chain (DofMotion, the
    prescribed-motion DofConstraint of examples/DynamicBeamAnalysis.jl:114-117), instead of the
    horizontal top force."""
    rng = np.random.default_rng(seed)
    model = Model("catenary_chains")
    ne = rng.integers(nmin, nmax + 1, size=nchain)
    length = rng.uniform(500., 2000., size=nchain)
    coords, first = [], []
    base = 0
    for c in range(nchain):
        n = int(ne[c])
        s = np.linspace(0., 1., n + 1)
        span, depth = 0.8 * length[c], 0.5 * length[c]
        # a sagging curve from the anchor (0,y,0) to the top (span,y,depth), x increasing
        x = span * s
        z = depth * (np.cosh(1.5 * s) - 1.) / (np.cosh(1.5) - 1.)
        y = np.full_like(s, 50. * c)
        coords.append(np.stack([x, y, z], axis=1))
        first.append(base)
        base += n + 1
    addnode(model, np.concatenate(coords))
    sec = chain_section().values
    nodes = np.concatenate([np.stack([f + 1 + np.arange(int(n)), f + 2 + np.arange(int(n))], axis=1)
                            for f, n in zip(first, ne)])
    addelement(model, EulerBeam3D, nodes, mat=sec, orient2=np.array([0., 1., 0.]))
    anchors = np.array(first) + 1
    tops = np.array(first) + ne + 1
    for f in ("t1", "t2", "t3"):
        addelement(model, Hold, anchors.reshape(-1, 1), field=f)
    if motion:
        amp = rng.uniform(0.5, 3., size=nchain)
        per = rng.uniform(6., 15., size=nchain)
        for c in range(nchain):
            addelement(model, DofMotion, [int(tops[c])], field="t1", value=top_motion, a=float(amp[c]), T=float(per[c]))
        for f in ("t2", "t3"):
            addelement(model, Hold, tops.reshape(-1, 1), field=f)
    else:
        addelement(model, DofLoad, tops.reshape(-1, 1), field="t1", value=topload)
    state = initialize(model, nXder=3)
    nX = model.ndof["X"]
    state.X[0][:] = rng.normal(0., disp, nX)
    rdofs = np.concatenate([model.getdofID("X", f) for f in ("r1", "r2", "r3")]) - 1
    state.X[0][rdofs] = rng.normal(0., rot, rdofs.size)
    state.X[1][:] = rng.normal(0., vel, nX)
    state.X[2][:] = rng.normal(0., acc, nX)
    return model, state


def cantilever_batch_specs(nmodel=512, nmin=500, nmax=20000, seed=1004):
    """Per-model parameters of the batch: element count, stiffness factor, seeds."""
    rng = np.random.default_rng(seed)
    ne = rng.integers(nmin, nmax + 1, size=nmodel)
    stiff = rng.lognormal(0., 0.1, size=nmodel)
    return [dict(ne=int(ne[m]), stiff=float(stiff[m]), seed=int(seed * 1000 + m)) for m in range(nmodel)]


def _add_cantilever_u(model, spec, h=0.05, yoff=0.):
    """Add one cantilever of the batch to `model`; returns (first node, node count, nU-nodes)."""
    n = spec["ne"]
    x = np.arange(n + 1) * h
    base = model.nnod
    xyz = np.stack([x, np.full_like(x, yoff), np.zeros_like(x)], axis=1)
    useg = np.minimum((np.arange(n) * 8) // n, 7)
    uxyz = np.stack([(np.arange(8) + 0.5) * n * h / 8, np.full(8, yoff + 1.), np.zeros(8)], axis=1)
    addnode(model, np.concatenate([xyz, uxyz]))
    sec = base_section().values.copy()
    sec[:4] *= spec["stiff"]
    nodes = np.stack([base + 1 + np.arange(n), base + 2 + np.arange(n), base + n + 2 + useg], axis=1)
    addelement(model, EulerBeam3DU, nodes, mat=sec, orient2=np.array([0., 1., 0.]))
    return base + 1


def _spec_vectors(spec):
    """State and Λ of one cantilever in its own numbering: X = 6(n+1) node dofs (node-major,
    t1..r3) then the 6 λ of its Holds; U = 8 U-nodes × 3."""
    rng = np.random.default_rng(spec["seed"])
    nd = 6 * (spec["ne"] + 1)
    x = np.concatenate([rng.normal(0., 1e-3, nd), np.zeros(6)])
    lam = rng.normal(0., 1., nd + 6)
    u = rng.normal(0., 1e3, 24)
    return x, lam, u


def cantilever_batch(specs, joined=True):
    """The batch as one Model (joined) or as one Model per spec. Returns a list of
    (model, state, Λ); the Holds of a joined model come after all its beams, so its X-dofs
    are the models' node dofs one model after the other, then the λ field by field."""
    if joined:
        model = Model("cantilever_batch")
        firsts = [_add_cantilever_u(model, s, yoff=3. * m) for m, s in enumerate(specs)]
        for fld in ("t1", "t2", "t3", "r1", "r2", "r3"):
            addelement(model, Hold, np.array(firsts).reshape(-1, 1), field=fld)
        state = initialize(model)
        lam = np.zeros(model.ndof["X"])
        nbeam = sum(6 * (s["ne"] + 1) for s in specs)
        xo = 0
        for m, s in enumerate(specs):
            x, l, u = _spec_vectors(s)
            nd = 6 * (s["ne"] + 1)
            state.X[0][xo:xo + nd] = x[:nd]
            lam[xo:xo + nd] = l[:nd]
            lam[nbeam + m + len(specs) * np.arange(6)] = l[nd:]
            state.U[0][24 * m:24 * m + 24] = u
            xo += nd
        return [(model, state, lam)]
    out = []
    for s in specs:
        model = Model("cantilever")
        f = _add_cantilever_u(model, s)
        for fld in ("t1", "t2", "t3", "r1", "r2", "r3"):
            addelement(model, Hold, np.array([[f]]), field=fld)
        state = initialize(model)
        x, lam, u = _spec_vectors(s)
        state.X[0][:] = x
        state.U[0][:] = u
        out.append((model, state, lam))
    return out


def design_point_specs(nmodel=512, nmin=500, nmax=20000, seed=1005):
    """Per-model element counts and seeds of the design-point batch."""
    rng = np.random.default_rng(seed)
    ne = rng.integers(nmin, nmax + 1, size=nmodel)
    return [dict(ne=int(ne[m]), seed=int(seed * 1000 + m)) for m in range(nmodel)]


def riser_section():
    """Section of the design-point risers: the catenary chains' stiffnesses, no weight."""
    return BeamCrossSection(EA=5e9, EI2=2e7, EI3=2e7, GJ=1.5e7, mu=1e3 / 9.81, iota1=0.5)


def riser_tension(t, T=1e6):
    """Top tension of a design-point riser (a dead vertical load)."""
    return T


def _add_riser_a(model, spec, h=0.1, xoff=0.):
    """One riser of the design-point batch: a vertical EulerBeam3DAU cantilever of ne elements
    of h, clamped at the bottom (first node), the section's EA, EI₂, EI₃, GJ scaled per
    sixteenth by 10^a, 8 segments of U-loads (per length, global axes). Returns (bottom node,
    top node)."""
    n = spec["ne"]
    L = n * h
    z = np.arange(n + 1) * h
    base = model.nnod
    xyz = np.stack([np.full_like(z, xoff), np.zeros_like(z), z], axis=1)
    useg = np.minimum((np.arange(n) * 8) // n, 7)
    aseg = np.minimum((np.arange(n) * 16) // n, 15)
    uxyz = np.stack([np.full(8, xoff + 1.), np.zeros(8), (np.arange(8) + 0.5) * L / 8], axis=1)
    axyz = np.stack([np.full(16, xoff + 2.), np.zeros(16), (np.arange(16) + 0.5) * L / 16], axis=1)
    addnode(model, np.concatenate([xyz, uxyz, axyz]))
    nodes = np.stack([base + 1 + np.arange(n), base + 2 + np.arange(n), base + n + 2 + useg, base + n + 10 + aseg],
                     axis=1)
    addelement(model, EulerBeam3DAU, nodes, mat=riser_section(), orient2=np.array([0., 1., 0.]),
               scaled=STIFFNESS_FIELDS, afield="Ξk")
    return base + 1, base + n + 1


def _design_vectors(spec, T, h=0.1, drift=0.02):
    """U (8 U-nodes × t1, t2, t3) and A (16) of one riser: a lateral current along x,
    u ~ N(1, 0.2)·q with q = 2·drift·T/L (the taut-string tip drift q·L/(2T) = drift·L), small
    loads along y and z (N(0, 0.1)·q); a ~ N(0, 0.1/ln 10), so that 10^a is lognormal with σ = 0.1."""
    rng = np.random.default_rng(spec["seed"])
    q = 2. * drift * T / (spec["ne"] * h)
    u = np.zeros((8, 3))
    u[:, 0] = rng.normal(1., 0.2, 8) * q
    u[:, 1:] = rng.normal(0., 0.1, (8, 2)) * q
    a = rng.normal(0., 0.1 / np.log(10.), 16)
    return u.ravel(), a


def design_point_batch(specs, joined=True, T=1e6):
    """The batch as one Model (joined) or one Model per spec. Each model is a tensioned riser:
    a vertical cantilever clamped at the seabed with the dead top tension T (DofLoad along z)
    and its top held in twist, so that its lateral stiffness comes mostly from the tension,
    as for a production riser.
    (A horizontal 10 m cantilever cut into 20k elements is not solvable in FP64: the
    condition of its bending stiffness grows as ne⁴.) Returns a list of (model, state, tip)
    with the state's U and A set and tip the 1-based X-dofs of the lateral top displacements
    (t1, one per riser)."""
    def build(sp_list, name):
        model = Model(name)
        ends = [_add_riser_a(model, s_, xoff=10. * m) for m, s_ in enumerate(sp_list)]
        for fld in ("t1", "t2", "t3", "r1", "r2", "r3"):
            addelement(model, Hold, np.array([f for f, _ in ends]).reshape(-1, 1), field=fld)
        # the top is held in twist (a riser hangs from its vessel): without it the free twist
        # of a long riser is the softest mode and Newton wanders in it
        addelement(model, Hold, np.array([t for _, t in ends]).reshape(-1, 1), field="r3")
        addelement(model, DofLoad, np.array([t for _, t in ends]).reshape(-1, 1), field="t3", value=riser_tension,
                   T=T)
        state = initialize(model, time=0.)
        # start from the taut riser (axial strain T/EA), so that the first Newton iteration
        # already has the tension's lateral stiffness
        EA = riser_section().values[0]
        for m, s_ in enumerate(sp_list):
            u, a = _design_vectors(s_, T)
            state.U[0][24 * m:24 * m + 24] = u
            state.A[16 * m:16 * m + 16] = a
        dn = model.dofnode("X")
        t3 = model.getdofID("X", "t3")
        state.X[0][t3 - 1] = model.coords()[dn[t3 - 1], 2] * (T / EA)
        t1 = model.getdofID("X", "t1")
        at = np.full(model.nnod, -1, np.int64)
        at[dn[t1 - 1]] = t1
        tip = at[np.array([t for _, t in ends]) - 1]
        return model, state, tip
    if joined:
        return [build(specs, "design_point_batch")]
    return [build([s_], "riser") for s_ in specs]


def design_point_equilibrium(state, nload=4, maxΔx=1e-6, maxiter=30, linsolve="device", stats=None, **kw):
    """The static equilibrium of a design-point batch under its U-loads, by load stepping: the
    U-loads at 1/nload, 2/nload, … of their value, each SweepX{0} solve started from the last
    equilibrium (a taut line under its full lateral load in one Newton solve from the straight
    state overshoots: its axial force doubles in the first iterate). Returns the final state;
    stats["iterations"] totals the Newton iterations."""
    from .sweepx import SweepX, solve
    U = state.U[0].copy()
    st = state.copy()
    total = 0
    for k in range(1, nload + 1):
        st.U[0][:] = U * (k / nload)
        st.time = 0.
        sub = {}
        st = solve(SweepX[0], initialstate=st, time=[1.], maxΔx=maxΔx, maxiter=maxiter, verbose=False,
                   catcherror=False, linsolve=linsolve, stats=sub, **kw)[-1]
        total += sub.get("iterations", 0)
    if stats is not None:
        stats["iterations"] = total
    return st


GAUGE_SIGMA = 1e-6


def _gauge_cost(eleres, t, targets, dt, w):
    """((ε − ε_measured)/σ)² of each gauge sensor at the step of time t (steps at dt, 2dt, …)."""
    q = t / dt
    k = int(round(q)) - 1
    if abs(q - round(q)) > 1e-6 or not 0 <= k < len(targets):
        raise MuscadeError(f"no measured strains at t = {t} (steps of {dt} s, {len(targets)} of them)")
    e = eleres["ε"][0] - targets[k]
    return w * (e * e)


def _load_prior(X, U, A, t, w):
    return w * (U[0][0] * U[0][0] + U[0][1] * U[0][1] + U[0][2] * U[0][2])


def _a_prior(a, w):
    return w * a * a


def gauge_identification(nel=5000, nseg=64, ngauge=50, length=1000., targets=None, dt=0.05, wU=1e-6, wA=1.):
    """The identification model (targets: measured strains, nstep × ngauge; None: the forward
    model, without sensors and priors). Beams, then Holds, then sensors and priors, so that
    the X-, U- and A-dofs are numbered alike in both. Returns (model, gauge beam indices)."""
    model = Model("gauge_identification")
    s_ = np.linspace(0., 1., nel + 1)
    span, depth = 0.8 * length, 0.5 * length
    x = span * s_
    z = depth * (np.cosh(1.5 * s_) - 1.) / (np.cosh(1.5) - 1.)
    addnode(model, np.stack([x, np.zeros_like(s_), z], axis=1))
    seg = np.minimum((np.arange(nel) * nseg) // nel, nseg - 1)
    mid = (np.arange(nseg) + 0.5) / nseg
    addnode(model, np.stack([span * mid, np.full(nseg, 10.), np.zeros(nseg)], axis=1))
    e = np.arange(nel)
    nodes = np.stack([e + 1, e + 2, e + 1, nel + 2 + seg], axis=1)
    addelement(model, EulerBeam3DAU, nodes, mat=chain_section(), orient2=np.array([0., 1., 0.]),
               scaled=("EI₂", "EI₃"), afield="ΞEI")
    for fld in ("t1", "t2", "t3"):
        addelement(model, Hold, np.array([[1], [nel + 1]]), field=fld)
    gidx = np.round(np.linspace(0, nel - 1, ngauge + 2)[1:-1]).astype(np.int64)
    if targets is not None:
        targets = np.asarray(targets, float)
        kw = dict(P=np.array([[0.], [0.], [0.15]]), D=np.array([[1.], [0.], [0.]]),
                  elementkwargs=dict(mat=np.zeros(16), orient2=np.array([0., 1., 0.])))
        addelement(model, ElementCost, np.stack([gidx + 1, gidx + 2], axis=1), req=request("ε"), cost=_gauge_cost,
                   ElementType=StrainGaugeOnEulerBeam3D, elementkwargs=kw,
                   costargs=(targets, dt, 1. / GAUGE_SIGMA ** 2))
        addelement(model, DofCost, (e + 1).reshape(-1, 1), cost=_load_prior, uinod=(1, 1, 1),
                   ufield=("t1", "t2", "t3"), costargs=(wU,))
        addelement(model, SingleAcost, (nel + 2 + np.arange(nseg)).reshape(-1, 1), field="ΞEI", cost=_a_prior,
                   costargs=(wA,))
    return model, gidx


def gauge_identification_truth(nel=5000, nseg=64, nstep=200, dt=0.05, seed=1003, load=(100., 300., 0.), **solve_kw):
    """The truth of the identification: a ~ N(0, 0.1/ln 10) per segment, the steady load per
    length `load` on every element, and the forward SweepX{2} run from rest in the as-meshed
    shape (solve_kw: the assembler and linear solver). Returns (model, A, U, states)."""
    from .sweepx import SweepX, solve
    rng = np.random.default_rng(seed)
    model, _ = gauge_identification(nel, nseg, targets=None)
    st = initialize(model, time=0., nXder=3)
    A = rng.normal(0., 0.1 / np.log(10.), nseg)
    st.A[:] = A
    st.U[0][:] = np.tile(np.asarray(load, float), nel)
    times = dt * np.arange(1, nstep + 1)
    out = solve(SweepX[2], initialstate=st, time=times, maxΔx=1e-6, maxiter=20, verbose=False, catcherror=False,
                **solve_kw)
    return model, A, st.U[0].copy(), out


def gauge_strains(values, gauges):
    """Gauge strains ε = G·(εₐₓ, κ₁, κ₂, κ₃) of the sensors: values (m, 4) of their requestables
    at a state, gauges (m, ngauge, 4) their gauge factors (StrainGaugeOnEulerBeam3D's G)."""
    return np.einsum("mgk,mk->mg", gauges, values)
