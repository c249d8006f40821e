"""ctypes binding of the CPU oracle (oracle/liboracle.so) — test infrastructure only.

The oracle restates Muscade's own algorithms (nested forward-mode duals,
sequential assembly) in C++; tests use it as the checker for the HIP path.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBPATH = os.path.join(ROOT, "oracle", "liboracle.so")

_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.POINTER(ctypes.c_int64)
_IP = ctypes.POINTER(ctypes.c_int)
_V = ctypes.c_void_p

K_BEAM, K_HOLD, K_DOFLOAD, K_ZERO, K_TURBINE, K_ANCHOR, K_BAR = 0, 1, 2, 3, 4, 5, 6

_SIG = {
    "orc_ipow_probe": (None, [_D, ctypes.c_int, _D]),
    "orc_sinc1_k": (ctypes.c_double, [ctypes.c_int, ctypes.c_double]),
    "orc_scac_d": (ctypes.c_double, [ctypes.c_int, ctypes.c_double]),
    "orc_rodrigues": (None, [_D, _D]),
    "orc_rodrigues_inv": (None, [_D, _D]),
    "orc_rodrigues_roundtrip": (None, [_D, _D, _D]),
    "orc_adjust": (None, [_D, _D, _D]),
    "orc_beam_construct": (ctypes.c_int, [_D, _D, _D, _D]),
    "orc_bar_construct": (None, [_D, _D, ctypes.c_double, _D]),
    "orc_bar_diffed_residual": (ctypes.c_int, [_D, _D, _D, ctypes.c_double, ctypes.c_int, ctypes.c_int, _D, _D,
                                               ctypes.c_double, _D, _D, _D]),
    "orc_model_set_time": (None, [_V, ctypes.c_double]),
    "orc_model_set_scales": (ctypes.c_int, [_V, ctypes.c_int, _D, _D]),
    "orc_model_prepare_direct": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I64]),
    "orc_model_direct_pattern": (None, [_V, ctypes.c_int, ctypes.c_int, _I64, _I64]),
    "orc_model_direct_sizes": (ctypes.c_int64, [_V, _I64]),
    "orc_model_assemble_direct": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_double, _D, _D, _D, _D, _D, _D, _D,
                                                 ctypes.c_double, _D, _D]),
    "orc_beam_diffed_residual": (ctypes.c_int, [_D, _D, _D, _D, ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, _D]),
    "orc_beam_residual_plain": (ctypes.c_int, [_D, _D, _D, _D, ctypes.c_int, _D, _D, _D, _D]),
    "orc_beam_gauge": (ctypes.c_int, [_D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D]),
    "orc_beam_results": (ctypes.c_int, [_D, _D, _D, _D, ctypes.c_int, _D, _D, _D]),
    "orc_beam_gauge2": (ctypes.c_int, [_D, _D, _D, _D, _D, _D]),
    "orc_anchor_eval": (ctypes.c_int, [_D, _D, _D, _D, _D, _D, _D, _D, _D]),
    "orc_number_dofs": (ctypes.c_int, [ctypes.c_int, _IP, _IP, _IP, ctypes.c_int, _IP, _IP, _IP, _IP, _IP, _IP, _IP,
                                       _I64, ctypes.c_int, _IP, _IP, _IP, _IP, _IP, _IP]),
    "orc_model_new": (_V, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]),
    "orc_model_free": (None, [_V]),
    "orc_model_add_type": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          _I64, _I64, _I64, _D, _D, ctypes.c_int]),
    "orc_model_set_dofload": (ctypes.c_int, [_V, ctypes.c_int, _D]),
    "orc_model_prepare_sweepx": (ctypes.c_int64, [_V]),
    "orc_model_get_pattern": (None, [_V, _I64, _I64]),
    "orc_model_get_asm": (None, [_V, ctypes.c_int, _I64, _I64]),
    "orc_model_assemble_sweepx": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, _D, _D, _D, _D, _IP,
                                                 _IP]),
    "orc_model_assemble_sweepx0_omp": (ctypes.c_int, [_V, ctypes.c_int, _D, _D, _D, _D, _D, _IP, _IP]),
}

lib = ctypes.CDLL(LIBPATH)
for _n, (_r, _a) in _SIG.items():
    getattr(lib, _n).restype = _r
    getattr(lib, _n).argtypes = _a


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def ip64(a):
    return None if a is None else a.ctypes.data_as(_I64)


def ip(a):
    return None if a is None else a.ctypes.data_as(_IP)


def ipow_probe(x8, b):
    out = np.zeros(8)
    lib.orc_ipow_probe(dp(f64(x8)), int(b), dp(out))
    return out


def beam_construct(c1, c2, orient2):
    out = np.zeros(53)
    r = lib.orc_beam_construct(dp(f64(c1)), dp(f64(c2)), dp(f64(orient2)), dp(out))
    if r:
        raise ValueError("orient2 nearly parallel")
    o = {}
    i = 0
    for name, n in [("cm", 3), ("rm", 9), ("zgp", 4), ("znod", 2), ("tgm", 3), ("tge", 3), ("ya", 4), ("yu", 4),
                    ("yv", 4), ("ka", 4), ("ku", 4), ("kv", 4), ("L", 1), ("dL", 4)]:
        o[name] = out[i:i + n].copy()
        i += n
    o["rm"] = o["rm"].reshape(3, 3, order="F")
    o["L"] = o["L"][0]
    return o


def beam_diffed_residual(c1, c2, orient2, mat16, X, U=None):
    """diffed_residual of one EulerBeam3D: X is a list of ND 12-vectors. Returns R, [∂R/∂X_der], ∂R/∂U."""
    ND = len(X)
    Xa = f64(np.concatenate([np.asarray(x, float) for x in X]))
    R = np.zeros(12)
    dX = np.zeros(ND * 144)
    dU = np.zeros(36)
    udof = 0 if U is None else 1
    r = lib.orc_beam_diffed_residual(dp(f64(c1)), dp(f64(c2)), dp(f64(orient2)), dp(f64(mat16)), udof, ND, dp(Xa),
                                     dp(f64(U) if U is not None else None), dp(R), dp(dX), dp(dU))
    if r:
        raise ValueError("orient2 nearly parallel")
    Ks = [dX[d * 144:(d + 1) * 144].reshape(12, 12, order="F") for d in range(ND)]
    return R, Ks, dU.reshape(12, 3, order="F")


EPS = float(np.finfo(float).eps)


def bar_construct(c1, c2, eps_s=EPS):
    out = np.zeros(26)
    lib.orc_bar_construct(dp(f64(c1)), dp(f64(c2)), eps_s, dp(out))
    names = [("cm", 3), ("tgm", 3), ("L0", 1), ("Ls", 1), ("wgp", 4), ("zgp", 4), ("znod", 2), ("psi1", 4), ("psi2", 4)]
    r, q = {}, 0
    for n, k in names:
        r[n] = out[q:q + k] if k > 1 else out[q]
        q += k
    return r


def bar_diffed_residual(c1, c2, mat9, X, U=None, t=0., eps_s=EPS):
    """diffed_residual of one Bar3D: X is a list of ND 6-vectors. Returns R, [∂R/∂X_der], ∂R/∂U."""
    ND = len(X)
    Xa = f64(np.concatenate([np.asarray(x, float) for x in X]))
    R = np.zeros(6)
    dX = np.zeros(ND * 36)
    dU = np.zeros(18)
    udof = 0 if U is None else 1
    lib.orc_bar_diffed_residual(dp(f64(c1)), dp(f64(c2)), dp(f64(mat9)), eps_s, udof, ND, dp(Xa),
                                dp(f64(U) if U is not None else None), t, dp(R), dp(dX), dp(dU))
    Ks = [dX[d * 36:(d + 1) * 36].reshape(6, 6, order="F") for d in range(ND)]
    return R, Ks, dU.reshape(6, 3, order="F")


def beam_residual_plain(c1, c2, orient2, mat16, X):
    ND = len(X)
    Xa = f64(np.concatenate([np.asarray(x, float) for x in X]))
    R = np.zeros(12)
    eps = np.zeros(1)
    kap = np.zeros(3)
    lib.orc_beam_residual_plain(dp(f64(c1)), dp(f64(c2)), dp(f64(orient2)), dp(f64(mat16)), ND, dp(Xa), dp(R), dp(eps),
                                dp(kap))
    return R, eps[0], kap


def beam_nres(ND):
    return 4 * ND + 4 * (6 * ND + 10)


def beam_results(c1, c2, orient2, mat16, X, U=None):
    """The requestables of one beam (layout of mh_beam_results) at X (list of ND vectors)."""
    ND = len(X)
    Xa = f64(np.concatenate([np.asarray(x, float) for x in X]))
    out = np.zeros(beam_nres(ND))
    r = lib.orc_beam_results(dp(f64(c1)), dp(f64(c2)), dp(f64(orient2)), dp(f64(mat16)), ND, dp(Xa),
                             dp(f64(U)) if U is not None else None, dp(out))
    assert r == 0
    return out


def upper_to_full(h78):
    """4×78 upper triangles (row-major) → 4×12×12 symmetric."""
    H = np.zeros(h78.shape[:-1] + (12, 12))
    iu = np.triu_indices(12)
    H[..., iu[0], iu[1]] = h78
    H[..., iu[1], iu[0]] = h78
    return H


def beam_gauge2(c1, c2, orient2, mat16, X):
    """(εₐₓ, κ₁, κ₂, κ₃) with gradients (4×12) and Hessians (4×12×12) w.r.t. X₀."""
    out = np.zeros(364)
    r = lib.orc_beam_gauge2(dp(f64(c1)), dp(f64(c2)), dp(f64(orient2)), dp(f64(mat16)), dp(f64(X)), dp(out))
    assert r == 0
    return out[:4], out[4:52].reshape(4, 12), upper_to_full(out[52:].reshape(4, 78))


def anchor_eval(p10, x, A):
    """AnchorLine: R (3), ∂R/∂(x, A) (3×5), Fh, ∂Fh (5), ∂²Fh (5×5), cr."""
    R, dR, Fh, dFh, HFh, cr = np.zeros(3), np.zeros(15), np.zeros(1), np.zeros(5), np.zeros(25), np.zeros(1)
    r = lib.orc_anchor_eval(dp(f64(p10)), dp(f64(x)), dp(f64(A)), dp(R), dp(dR), dp(Fh), dp(dFh), dp(HFh), dp(cr))
    assert r == 0
    return R, dR.reshape(3, 5), Fh[0], dFh, HFh.reshape(5, 5), cr[0]


def beam_gauge(c1, c2, orient2, mat16, X):
    R = np.zeros(12)
    K = np.zeros(144)
    eps = np.zeros(1)
    deps = np.zeros(12)
    kap = np.zeros(3)
    dkap = np.zeros(36)
    lib.orc_beam_gauge(dp(f64(c1)), dp(f64(c2)), dp(f64(orient2)), dp(f64(mat16)), dp(f64(X)), dp(R), dp(K), dp(eps),
                       dp(deps), dp(kap), dp(dkap))
    return R, K.reshape(12, 12, order="F"), eps[0], deps, kap, dkap.reshape(3, 12)


class OracleModel:
    """A model handed to the oracle's sequential SweepX assembler (Assemble.jl:470-505 restated)."""

    def __init__(self, nX, nU=0, nA=0):
        self.h = lib.orc_model_new(nX, nU, nA)
        self.nX = nX
        self.nU = nU
        self.ntypes = 0

    def add_type(self, kind, iX, iU=None, iA=None, scaleX=None, params=None, udof=0):
        iX = np.ascontiguousarray(iX, np.int64)
        nele, nX = iX.shape
        nU = 0 if iU is None else iU.shape[1]
        nA = 0 if iA is None else iA.shape[1]
        r = lib.orc_model_add_type(self.h, kind, nele, nX, nU, nA, ip64(iX),
                                   ip64(np.ascontiguousarray(iU, np.int64) if nU else None),
                                   ip64(np.ascontiguousarray(iA, np.int64) if nA else None),
                                   dp(f64(scaleX) if scaleX is not None else None),
                                   dp(f64(params) if params is not None else None), udof)
        if r:
            raise ValueError("orc_model_add_type failed")
        self.ntypes += 1
        return self.ntypes - 1

    def set_scales(self, itype, scaleL=None, scaleU=None):
        lib.orc_model_set_scales(self.h, itype, dp(f64(scaleL) if scaleL is not None else None),
                                 dp(f64(scaleU) if scaleU is not None else None))

    def prepare_direct(self, OX, OU, IA):
        """prepare(AssemblyDirect{OX,OU,IA}): patterns of the class pairs (Λ,X,U)²."""
        nnz = np.zeros(9, np.int64)
        r = lib.orc_model_prepare_direct(self.h, OX, OU, IA, ip64(nnz))
        if r:
            raise ValueError("orc_model_prepare_direct failed")
        self.dOX, self.dOU = OX, OU
        self.dnnz = nnz.reshape(3, 3)
        self.dpattern = {}
        nU = self.nU
        for a in range(3):
            for b in range(3):
                ncol = nU if b == 2 else self.nX
                cp = np.zeros(ncol + 1, np.int64)
                rv = np.zeros(self.dnnz[a, b], np.int64)
                lib.orc_model_direct_pattern(self.h, a, b, ip64(cp), ip64(rv))
                self.dpattern[(a, b)] = (cp, rv)
        return self.dnnz

    def assemble_direct(self, mission, dt, Lam, X, U=None, t=0.):
        nL1 = np.zeros(1, np.int64)
        nL2 = lib.orc_model_direct_sizes(self.h, ip64(nL1))
        L1 = np.zeros(nL1[0])
        L2 = np.zeros(nL2)
        Xs = [f64(x) for x in X] + [None] * (3 - len(X))
        Us = ([f64(u) for u in U] if U is not None else []) + [None] * (3 - (len(U) if U is not None else 0))
        r = lib.orc_model_assemble_direct(self.h, mission, dt, dp(f64(Lam)), dp(Xs[0]), dp(Xs[1]), dp(Xs[2]),
                                          dp(Us[0]), dp(Us[1]), dp(Us[2]), t, dp(L1), dp(L2))
        return r, L1, L2

    def set_time(self, t):
        lib.orc_model_set_time(self.h, float(t))

    def set_dofload(self, itype, F):
        lib.orc_model_set_dofload(self.h, itype, dp(f64(F)))

    def prepare(self):
        self.nnz = lib.orc_model_prepare_sweepx(self.h)
        self.colptr = np.zeros(self.nX + 1, np.int64)
        self.rowval = np.zeros(self.nnz, np.int64)
        lib.orc_model_get_pattern(self.h, ip64(self.colptr), ip64(self.rowval))
        return self.nnz

    def asm(self, itype, nele, nX):
        a1 = np.zeros(nele * nX, np.int64)
        a2 = np.zeros(nele * nX * nX, np.int64)
        lib.orc_model_get_asm(self.h, itype, ip64(a1), ip64(a2))
        return a1.reshape(nele, nX), a2.reshape(nele, nX * nX)

    def assemble(self, OX, mission, c6, X, U=None, A=None):
        rhs = np.zeros(self.nX)
        nz = np.zeros(self.nnz)
        et = np.zeros(1, np.intc)
        ee = np.zeros(1, np.intc)
        Xs = [f64(x) for x in X] + [None] * (3 - len(X))
        r = lib.orc_model_assemble_sweepx(self.h, OX, mission, dp(f64(c6)), dp(Xs[0]), dp(Xs[1]), dp(Xs[2]),
                                          dp(f64(U) if U is not None else None), dp(f64(A) if A is not None else None),
                                          dp(rhs), dp(nz), ip(et), ip(ee))
        return r, rhs, nz, (int(et[0]), int(ee[0]))

    def assemble0_parallel(self, nthreads, X0, U=None, A=None):
        """Static assemble!{:iter} with element evaluations on nthreads OpenMP threads
        (same results as assemble(0, 1, ...), bit for bit)."""
        rhs = np.zeros(self.nX)
        nz = np.zeros(self.nnz)
        et = np.zeros(1, np.intc)
        ee = np.zeros(1, np.intc)
        r = lib.orc_model_assemble_sweepx0_omp(self.h, int(nthreads), dp(f64(X0)), dp(f64(U) if U is not None else None),
                                               dp(f64(A) if A is not None else None), dp(rhs), dp(nz), ip(et), ip(ee))
        return r, rhs, nz, (int(et[0]), int(ee[0]))

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_model_free(self.h)
            self.h = None
