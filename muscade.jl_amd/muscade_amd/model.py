"""Model description, dof numbering, Disassembler and State.

Host-side mirror of Muscade's src/ModelDescription.jl and the
`Disassembler`/`State`/`DofGroup` parts of src/Assemble.jl, written with numpy
so that models of millions of elements are numbered in seconds.

Numbering rules kept from the reference (these fix the asm maps and the CSC
pattern bit for bit):
  * element types are numbered in order of first appearance
    (getieletyp / addelement!, ModelDescription.jl:203-208);
  * within a class (:X, :U, :A) a dof is identified by (node, class, field)
    and receives the next number the first time an element refers to it,
    elements taken in the order they are added and each element's dofs in
    doflist order (ModelDescription.jl:226-249);
  * dof types (class, field) are registered in first-appearance order
    (ModelDescription.jl:228-233) and carry the scale of setscale!.
"""
import math

import numpy as np

CLASSES = ("X", "U", "A")
_SHIFT = 20   # key = node << _SHIFT | idoftyp


class MuscadeError(Exception):
    """Mirror of MuscadeException (src/Exceptions.jl:1-16): message + location info."""

    def __init__(self, msg, dbg=None):
        super().__init__(msg)
        self.msg = msg
        self.dbg = dict(dbg or {})

    def __str__(self):
        return self.msg + (f" {self.dbg}" if self.dbg else "")


class EleType:
    """One Muscade element type (an entry of model.eleobj): all elements of one concrete type."""

    def __init__(self, cls, key, doflist):
        self.cls = cls
        self.key = key
        self.inod, self.dofclass, self.field = doflist
        self.nodes = []     # chunks of (nele, nnod) 0-based node numbers
        self.dofid = []     # chunks of (nele, ndof) class-local 1-based dof numbers
        self.params = []    # chunks of per-element parameters (element class defines them)

    @property
    def nele(self):
        return sum(n.shape[0] for n in self.nodes)

    def cat_nodes(self):
        return np.concatenate(self.nodes) if self.nodes else np.zeros((0, max(self.inod, default=0)), np.int64)

    def cat_dofid(self):
        return np.concatenate(self.dofid) if self.dofid else np.zeros((0, len(self.inod)), np.int64)

    def class_columns(self, c):
        return [k for k, cc in enumerate(self.dofclass) if cc == c]


class Model:
    """Model(ID=:muscade_model) — ModelDescription.jl:150."""

    def __init__(self, ID="muscade_model"):
        self.ID = ID
        self._coords = []            # list of (n, dim) chunks
        self.nnod = 0
        self.eletypes = []           # first-appearance order
        self._typeindex = {}
        self.doftyp = []             # [class, field, scale]
        self._doftypindex = {}
        self.scaleΛ = 1.0
        self.locked = False
        # per class and dof type: node → dof number (0 = not numbered yet), grown on demand
        self._lut = {c: {} for c in CLASSES}
        self._dofnode = {c: [] for c in CLASSES}
        self._doftypof = {c: [] for c in CLASSES}
        self.ndof = {c: 0 for c in CLASSES}

    # ------------------------------------------------------------------ nodes
    def coords(self):
        if not self._coords:
            return np.zeros((0, 3))
        dim = max(c.shape[1] for c in self._coords)
        out = np.zeros((self.nnod, max(dim, 3)))
        i = 0
        for c in self._coords:
            out[i:i + c.shape[0], :c.shape[1]] = c
            i += c.shape[0]
        return out

    def _assert_unlocked(self):
        if self.locked:
            raise MuscadeError(f"model {self.ID} is initialized and can no longer be edited")

    # ---------------------------------------------------------------- dofs
    def _idoftyp(self, cls, field):
        k = (cls, field)
        if k not in self._doftypindex:
            self._doftypindex[k] = len(self.doftyp)
            self.doftyp.append([cls, field, 1.0])
        return self._doftypindex[k]

    def _lut_for(self, c, t, size):
        a = self._lut[c].get(t)
        if a is None or a.size < size:
            n = max(size, 2 * (a.size if a is not None else 0), 1024)
            b = np.zeros(n, np.int64)
            if a is not None:
                b[:a.size] = a
            self._lut[c][t] = a = b
        return a

    def _number(self, E, nodes):
        """First-encounter numbering of the dofs of a batch of elements (ModelDescription.jl:226-249)."""
        nele = nodes.shape[0]
        ndof = len(E.inod)
        idt = np.array([self._idoftyp(c, f) for c, f in zip(E.dofclass, E.field)], np.int64)
        dofid = np.zeros((nele, ndof), np.int64)
        nmax = int(nodes.max()) + 1 if nodes.size else 1
        for c in CLASSES:
            ks = E.class_columns(c)
            if not ks or nele == 0:
                continue
            cols = np.array([E.inod[k] - 1 for k in ks])
            nd = nodes[:, cols].astype(np.int64)             # nele × len(ks), element-major traversal
            ids = np.empty(nd.shape, np.int64)
            for q, k in enumerate(ks):
                ids[:, q] = self._lut_for(c, int(idt[k]), nmax)[nd[:, q]]
            flat_ids = ids.ravel()
            new = flat_ids == 0
            if new.any():
                keys = ((nd << _SHIFT) | idt[ks][None, :]).ravel()[new]
                uk, first = np.unique(keys, return_index=True)
                order = np.argsort(first, kind="stable")
                base = self.ndof[c]
                newid = np.empty(uk.size, np.int64)
                newid[order] = base + 1 + np.arange(uk.size)
                flat_ids[new] = newid[np.searchsorted(uk, keys)]
                self.ndof[c] += uk.size
                unode, utyp = uk >> _SHIFT, uk & ((1 << _SHIFT) - 1)
                self._dofnode[c].append(unode[order])
                self._doftypof[c].append(utyp[order])
                for t in np.unique(utyp):
                    m = utyp == t
                    self._lut_for(c, int(t), nmax)[unode[m]] = newid[m]
            dofid[:, ks] = flat_ids.reshape(nele, len(ks))
        return dofid

    def dofnode(self, c):
        return np.concatenate(self._dofnode[c]) if self._dofnode[c] else np.zeros(0, np.int64)

    def doftypof(self, c):
        return np.concatenate(self._doftypof[c]) if self._doftypof[c] else np.zeros(0, np.int64)

    def dofscale(self, c):
        sc = np.array([d[2] for d in self.doftyp]) if self.doftyp else np.zeros(0)
        return sc[self.doftypof(c)] if self.ndof[c] else np.zeros(0)

    def fieldof(self, c):
        names = [d[1] for d in self.doftyp]
        return [names[i] for i in self.doftypof(c)]

    def getndof(self, c=None):
        if c is None:
            return sum(self.ndof.values())
        return self.ndof[c]

    def getnele(self):
        return sum(E.nele for E in self.eletypes)

    def getdofID(self, cls, field, nodID=None):
        """Model dof numbers (1-based) of class/field, optionally restricted to nodes (ModelDescription.jl:133-138)."""
        if (cls, field) not in self._doftypindex:
            raise MuscadeError(f"The model has no dof of class {cls} and field {field}.")
        it = self._doftypindex[(cls, field)]
        ids = np.nonzero(self.doftypof(cls) == it)[0] + 1
        if nodID is not None:
            nodes = np.atleast_1d(np.asarray(nodID, np.int64)) - 1
            ids = ids[np.isin(self.dofnode(cls)[ids - 1], nodes)]
        return ids


def addnode(model, coord):
    """nodID = addnode!(model, coord) — ModelDescription.jl:166-173. Rows of a matrix are nodes.
    Returns 1-based node numbers (an int for a single vector)."""
    model._assert_unlocked()
    c = np.asarray(coord, dtype=np.float64)
    single = c.ndim == 1
    c = c.reshape(1, -1) if single else c
    ids = model.nnod + 1 + np.arange(c.shape[0])
    model._coords.append(c)
    model.nnod += c.shape[0]
    return int(ids[0]) if single else ids


def addelement(model, ElType, nodID, **kwargs):
    """eleID = addelement!(model, ElType, nodID; kwargs...) — ModelDescription.jl:195-263.

    nodID: one element's node numbers (vector) or a matrix with one element per row.
    kwargs are passed to the element constructor; beam materials and orient2 may
    also be given per element (arrays with one row per element)."""
    model._assert_unlocked()
    nod = np.asarray(nodID, dtype=np.int64)
    single = nod.ndim == 1
    nod = nod.reshape(1, -1) if single else nod
    nodes0 = nod - 1
    if nodes0.size and (nodes0.min() < 0 or nodes0.max() >= model.nnod):
        raise MuscadeError("addelement: node number out of range")
    key, doflist, params = ElType.construct(model, nodes0, **kwargs)
    nnod = max(doflist[0])
    if nod.shape[1] != nnod:
        raise MuscadeError(f"Connecting element of type {ElType.__name__}: Second dimension of inod ({nod.shape[1]}) "
                           f"must be equal to element's nnod ({nnod})")
    if key not in model._typeindex:
        model._typeindex[key] = len(model.eletypes)
        model.eletypes.append(EleType(ElType, key, doflist))
    ieletyp = model._typeindex[key]
    E = model.eletypes[ieletyp]
    iele0 = E.nele
    dofid = model._number(E, nodes0)
    E.nodes.append(nodes0)
    E.dofid.append(dofid)
    E.params.append(params)
    ids = [(ieletyp + 1, iele0 + 1 + i) for i in range(nod.shape[0])]
    return ids[0] if single else ids


def setscale(model, scale=None, Λscale=None):
    """setscale!(model; scale, Λscale) — ModelDescription.jl:287-299."""
    model._assert_unlocked()
    if scale is not None:
        for d in model.doftyp:
            if d[0] in scale and d[1] in scale[d[0]]:
                d[2] = float(scale[d[0]][d[1]])
    if Λscale is not None:
        model.scaleΛ = float(Λscale)


class EletypDisassembler:
    def __init__(self, X, U, A, sX, sU, sA, sΛ):
        self.X, self.U, self.A = X, U, A          # (nele, n) 1-based model dofs
        self.scaleX, self.scaleU, self.scaleA, self.scaleΛ = sX, sU, sA, sΛ


class Disassembler:
    """Disassembler(model) — src/Assemble.jl:32-99."""

    def __init__(self, model):
        self.dis = []
        for E in model.eletypes:
            d = E.cat_dofid()
            cols = {c: E.class_columns(c) for c in CLASSES}
            # scales come from the element type's first element (Assemble.jl:49-63)
            sc = {c: np.array([model.doftyp[model._doftypindex[(c, E.field[k])]][2] for k in cols[c]]) for c in CLASSES}
            self.dis.append(EletypDisassembler(d[:, cols["X"]], d[:, cols["U"]], d[:, cols["A"]],
                                               sc["X"], sc["U"], sc["A"], sc["X"] * model.scaleΛ))
        self.scaleX = model.dofscale("X")
        self.scaleΛ = self.scaleX * model.scaleΛ
        self.scaleU = model.dofscale("U")
        self.scaleA = model.dofscale("A")
        self.fieldX = model.fieldof("X")
        self.fieldU = model.fieldof("U")
        self.fieldA = model.fieldof("A")


class State:
    """State{nΛder,nXder,nUder} — src/Assemble.jl:106-159. Vectors are numpy arrays."""

    def __init__(self, time, Λ, X, U, A, SP, model, dis):
        self.time = time
        self.Λ = list(Λ)
        self.X = list(X)
        self.U = list(U)
        self.A = A
        self.SP = SP
        self.model = model
        self.dis = dis

    def copy(self, time=None):
        return State(self.time if time is None else time, [a.copy() for a in self.Λ], [a.copy() for a in self.X],
                     [a.copy() for a in self.U], self.A.copy(), self.SP, self.model, self.dis)

    def with_ders(self, nΛder, nXder, nUder):
        """State{nΛder,nXder,nUder}(s): shave off or zero-pad derivatives (Assemble.jl:131-144)."""
        def pad(v, n, size):
            return [v[i] if i < len(v) else np.zeros(size) for i in range(n)]
        return State(self.time, pad(self.Λ, nΛder, self.model.ndof["X"]), pad(self.X, nXder, self.model.ndof["X"]),
                     pad(self.U, nUder, self.model.ndof["U"]), self.A, self.SP, self.model, self.dis)


def initialize(model, nΛder=1, nXder=1, nUder=1, time=-math.inf):
    """initialstate = initialize!(model) — ModelDescription.jl:319-324."""
    model._assert_unlocked()
    model.locked = True
    dis = Disassembler(model)
    nX, nU, nA = model.ndof["X"], model.ndof["U"], model.ndof["A"]
    return State(time, [np.zeros(nX) for _ in range(nΛder)], [np.zeros(nX) for _ in range(nXder)],
                 [np.zeros(nU) for _ in range(nUder)], np.zeros(nA), None, model, dis)


def getdof(state, cls="X", field=None, nodID=None, order=0):
    """getdof(state; class, field, nodID, order) — src/Output.jl:21-54 (state or list of states)."""
    states = state if isinstance(state, (list, tuple)) else [state]
    c = "X" if cls == "Λ" else cls
    ids = states[0].model.getdofID(c, field, nodID)
    out = np.zeros((ids.size, len(states)))
    for i, s in enumerate(states):
        if cls == "A":
            out[:, i] = s.A[ids - 1]
            continue
        sc = {"Λ": s.Λ, "X": s.X, "U": s.U}[cls]
        if order < len(sc):
            out[:, i] = sc[order][ids - 1]
    return out[:, 0] if not isinstance(state, (list, tuple)) else out


def setdof(state, value, cls="X", field=None, nodID=None, order=0):
    """state = setdof!(state, value; class, field, nodID, order) — src/Output.jl:73-128."""
    c = "X" if cls == "Λ" else cls
    ids = state.model.getdofID(c, field, nodID)
    if cls == "A":
        state.A[ids - 1] = value
        return state
    sc = {"Λ": state.Λ, "X": state.X, "U": state.U}[cls]
    n = state.model.ndof[c]
    while len(sc) < order + 1:
        sc.append(np.zeros(n))
    sc[order][ids - 1] = value
    return state


class DofGroup:
    """DofGroup (src/Assemble.jl:167-188): a selection of model dofs stacked class after class in a
    solver vector, y[j] ↔ state[i]·scale. i*: model dofs (1-based) of each class in the group,
    j*: their positions in y (1-based), scale*, field* per group entry. nX, nU, nA are the
    model's dof counts."""

    def __init__(self, nX, nU, nA, iΛ, iX, iU, iA, jΛ, jX, jU, jA, sΛ, sX, sU, sA, fΛ, fX, fU, fA):
        self.nX, self.nU, self.nA = nX, nU, nA
        self.iΛ, self.iX, self.iU, self.iA = [np.asarray(v, np.int64) for v in (iΛ, iX, iU, iA)]
        self.jΛ, self.jX, self.jU, self.jA = jΛ, jX, jU, jA
        self.scaleΛ, self.scaleX, self.scaleU, self.scaleA = sΛ, sX, sU, sA
        self.fieldΛ, self.fieldX, self.fieldU, self.fieldA = fΛ, fX, fU, fA

    @classmethod
    def from_dis(cls, dis, iΛ, iX, iU, iA):
        """DofGroup(dis, iΛ, iX, iU, iA) (Assemble.jl:190-200): classes stacked, no permutation."""
        iΛ, iX, iU, iA = [np.asarray(list(v), np.int64) for v in (iΛ, iX, iU, iA)]
        nX, nU, nA = len(dis.scaleX), len(dis.scaleU), len(dis.scaleA)
        jΛ, jX, jU, jA = gradientpartition(len(iΛ), len(iX), len(iU), len(iA))
        pick = lambda a, i: a[i - 1] if len(i) else a[:0]   # noqa: E731
        pickf = lambda f, i: [f[k - 1] for k in i]          # noqa: E731
        return cls(nX, nU, nA, iΛ, iX, iU, iA, jΛ, jX, jU, jA, pick(dis.scaleΛ, iΛ), pick(dis.scaleX, iX),
                   pick(dis.scaleU, iU), pick(dis.scaleA, iA), pickf(dis.fieldX, iΛ), pickf(dis.fieldX, iX),
                   pickf(dis.fieldU, iU), pickf(dis.fieldA, iA))

    def getndof(self):
        return len(self.iΛ) + len(self.iX) + len(self.iU) + len(self.iA)


def gradientpartition(nΛ, nX, nU, nA, nXder=1, nUder=None):
    """Indices (1-based ranges) of the class partitions of a gradient (Assemble.jl:327-335)."""
    nUder = nXder if nUder is None else nUder
    return (range(1, nΛ + 1), range(nΛ + 1, nΛ + nX + 1), range(nΛ + nX * nXder + 1, nΛ + nX * nXder + nU + 1),
            range(nΛ + nX * nXder + nU * nUder + 1, nΛ + nX * nXder + nU * nUder + nA + 1))


def gradientstructure(dofgr, eledis):
    """Number of dofs of each class in the gradient an element returns (Assemble.jl:321-326)."""
    return (0 if len(dofgr.iΛ) == 0 else eledis.X.shape[1], 0 if len(dofgr.iX) == 0 else eledis.X.shape[1],
            0 if len(dofgr.iU) == 0 else eledis.U.shape[1], 0 if len(dofgr.iA) == 0 else eledis.A.shape[1])


def indexedstate(dofgr):
    """(Λ, X, U, A): for each model dof its position in the group (1-based), 0 if not in it
    (Assemble.jl:246-256)."""
    out = []
    for n, i, j in ((dofgr.nX, dofgr.iΛ, dofgr.jΛ), (dofgr.nX, dofgr.iX, dofgr.jX), (dofgr.nU, dofgr.iU, dofgr.jU),
                    (dofgr.nA, dofgr.iA, dofgr.jA)):
        v = np.zeros(n, np.int64)
        if len(i):
            v[i - 1] = np.asarray(list(j), np.int64)
        out.append(v)
    return tuple(out)


def _group(model, dis, Λ=False, X=False, U=False, A=False):
    r = lambda c, on: range(1, model.ndof[c] + 1) if on else range(0)   # noqa: E731
    return DofGroup.from_dis(dis, r("X", Λ), r("X", X), r("U", U), r("A", A))


def nodofs(model, dis):
    return _group(model, dis)


def allΛdofs(model, dis):
    return _group(model, dis, Λ=True)


def allXdofs(model, dis):
    """allXdofs(model, dis) (Assemble.jl:261): the group of SweepX."""
    return _group(model, dis, X=True)


def allUdofs(model, dis):
    return _group(model, dis, U=True)


def allAdofs(model, dis):
    return _group(model, dis, A=True)


def allXUdofs(model, dis):
    return _group(model, dis, X=True, U=True)


def allΛXUdofs(model, dis):
    return _group(model, dis, Λ=True, X=True, U=True)


def allΛXUAdofs(model, dis):
    return _group(model, dis, Λ=True, X=True, U=True, A=True)


def selecteddofs(model, dis, classes):
    """selecteddofs(model, dis, classes) (Assemble.jl:268-274)."""
    return _group(model, dis, Λ="Λ" in classes, X="X" in classes, U="U" in classes, A="A" in classes)


def decrement(state, ider, y, gr):
    """decrement!(state, ider, y, gr) (Assemble.jl:201-206), ider 1-based like the reference."""
    y = np.asarray(y, float)
    jl = lambda j: np.asarray(list(j), np.int64) - 1   # noqa: E731
    if ider <= len(state.Λ) and len(gr.iΛ):
        state.Λ[ider - 1][gr.iΛ - 1] -= y[jl(gr.jΛ)] * gr.scaleΛ
    if ider <= len(state.X) and len(gr.iX):
        state.X[ider - 1][gr.iX - 1] -= y[jl(gr.jX)] * gr.scaleX
    if ider <= len(state.U) and len(gr.iU):
        state.U[ider - 1][gr.iU - 1] -= y[jl(gr.jU)] * gr.scaleU
    if ider == 1 and len(gr.iA):
        state.A[gr.iA - 1] -= y[jl(gr.jA)] * gr.scaleA


def getdofs(state, ider, gr):
    """getdof!(state, ider, y, gr) (Assemble.jl:232-237): y[j] = state[i]/scale, ider 0-based as
    in the reference (the derivative ider+1)."""
    y = np.zeros(gr.getndof())
    jl = lambda j: np.asarray(list(j), np.int64) - 1   # noqa: E731
    if len(gr.iΛ):
        y[jl(gr.jΛ)] = state.Λ[ider][gr.iΛ - 1] / gr.scaleΛ
    if len(gr.iX):
        y[jl(gr.jX)] = state.X[ider][gr.iX - 1] / gr.scaleX
    if len(gr.iU):
        y[jl(gr.jU)] = state.U[ider][gr.iU - 1] / gr.scaleU
    if len(gr.iA):
        y[jl(gr.jA)] = state.A[gr.iA - 1] / gr.scaleA
    return y
