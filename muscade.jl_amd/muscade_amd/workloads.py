"""Synthetic models of the measured workloads (all inputs seeded, numpy PCG64).

space_frame: the static assembly-at-scale case — a 3D lattice of nx×ny×nz
nodes at spacing h, positions jittered by U(-1%,+1%)·h, EulerBeam3D members
along x, then y, then z (each in lexicographic order), section stiffnesses
× LogNormal(0, 0.2) per member, the base layer (k=0) clamped by six Holds
per node, random displacements N(0,1e-3 m) and rotations N(0,1e-2 rad).
100×100×33 nodes gives 973,400 beams and 2,040,000 X-dofs.

cantilever: the reference's static SweepX{0} case — 100 beams along x,
L=10 m, clamped at node 1, tip load ramped to 5e4 N in 10 steps.
"""
import numpy as np

from .elements import BeamCrossSection, DofLoad, EulerBeam3D, Hold
from .model import Model, addelement, addnode, initialize, setdof

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
