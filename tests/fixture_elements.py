"""The reference's test elements, restated as host residual elements (muscade_amd.ResidualElement),
so that the product's solver loops can be run against the reference's solver goldens:

- SdofOscillator, AdjustableSdofOscillator and Spring{D} (test/SomeElements.jl:160-226);
- El1 (test/TestDirectXUA.jl:9-21).

Test infrastructure: the residuals are the reference's formulas, written over Dual2 so that the
product's host-element path differentiates them."""
import numpy as np

from muscade_amd.elements import ResidualElement
from muscade_amd.hostcost import exp10


class SdofOscillator(ResidualElement):
    """R = −u + K₁x + K₂x² + C₁x′ + C₂x′² + M₁x″ + M₂x″² (SomeElements.jl:189-205)."""

    @classmethod
    def doflist(cls):
        return (1, 1), ("X", "U"), ("tx1", "tu1")

    @classmethod
    def make(cls, model, nodes0, K1=0., K2=0., C1=0., C2=0., M1=0., M2=0.):
        return dict(K1=K1, K2=K2, C1=C1, C2=C2, M1=M1, M2=M2)

    @staticmethod
    def residual(p, X, U, A, t):
        x, x1, x2, u = X[0][0], X[1][0], X[2][0], U[0][0]
        return [-u + p["K1"] * x + p["K2"] * x ** 2 + p["C1"] * x1 + p["C2"] * x1 ** 2 + p["M1"] * x2
                + p["M2"] * x2 ** 2]


class AdjustableSdofOscillator(ResidualElement):
    """R = −u + K·x + C·10^A₁·x′ + M·x″ (SomeElements.jl:207-226)."""

    @classmethod
    def doflist(cls):
        return (1, 1, 1), ("X", "U", "A"), ("tx1", "tu1", "ΞK")

    @classmethod
    def make(cls, model, nodes0, K=1., C=0., M=0.):
        return dict(K=K, C=C, M=M)

    @staticmethod
    def residual(p, X, U, A, t):
        x, x1, x2, u = X[0][0], X[1][0], X[2][0], U[0][0]
        C = p["C"] * exp10(A[0])
        return [-u + p["K"] * x + C * x1 + p["M"] * x2]


class Spring(ResidualElement):
    """Spring{D} (SomeElements.jl:166-186): L₀ = L·10^A₁, EA = EA·10^A₂ (the ☼-marked locals), T =
    EA(L − L₀)/L₀ with L the current length, R = (Δx/L·T, −Δx/L·T)."""
    no_second_order = False

    @classmethod
    def doflist(cls, D=None):
        raise NotImplementedError   # per dimension: Spring1, Spring2

    @classmethod
    def make(cls, model, nodes0, EA):
        c = model.coords()
        D = cls.D
        x1 = c[nodes0[:, 0], :D]
        x2 = c[nodes0[:, 1], :D]
        return dict(x1=x1, x2=x2, EA=float(EA), L=np.linalg.norm(x1 - x2, axis=1))

    @classmethod
    def residual(cls, p, X, U, A, t):
        D = cls.D
        L0 = p["L"] * exp10(A[0])
        EA = p["EA"] * exp10(A[1])
        dx = [(X[0][i] + p["x1"][:, i]) - (X[0][i + D] + p["x2"][:, i]) for i in range(D)]
        ss = dx[0] * dx[0]
        for i in range(1, D):
            ss = ss + dx[i] * dx[i]
        L = np.sqrt(ss)
        T = EA * (L - L0) / L0
        F1 = [d / L * T for d in dx]
        return F1 + [-f for f in F1]


def _spring(D):
    def doflist(cls):
        return ((1,) * D + (2,) * D + (3, 3), ("X",) * (2 * D) + ("A", "A"),
                tuple(f"tx{i + 1}" for i in range(D)) * 2 + ("ΞL₀", "ΞEI"))
    return type(f"Spring{D}", (Spring,), {"D": D, "doflist": classmethod(doflist)})


Spring1 = _spring(1)
Spring2 = _spring(2)


class El1(ResidualElement):
    """R = −u + K·x + C·10^ΞC·x′ + M·10^ΞM·x″ (TestDirectXUA.jl:9-21)."""

    @classmethod
    def doflist(cls):
        return (1, 1, 1, 1), ("X", "U", "A", "A"), ("tx1", "u", "ΞC", "ΞM")

    @classmethod
    def make(cls, model, nodes0, K, C, M):
        return dict(K=K, C=C, M=M)

    @staticmethod
    def residual(p, X, U, A, t):
        x, x1, x2, u = X[0][0], X[1][0], X[2][0], U[0][0]
        return [-u + p["K"] * x + p["C"] * exp10(A[0]) * x1 + p["M"] * exp10(A[1]) * x2]
