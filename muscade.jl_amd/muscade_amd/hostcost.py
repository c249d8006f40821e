"""User-cost elements evaluated on the host, and the second-order duals that differentiate them.

Muscade evaluates the Lagrangian of DofCost, Acost, SingleDofCost, SingleUdof and ElementCost
elements with nested forward-mode duals (src/BasicElements.jl:32-248; src/DirectXUA.jl:152-171:
`revariate{2}`, then `∂{2,Np}` gives the gradient and the Hessian over all partials). Their
cost functions are user callables, so they stay on the host, like any user `Functor`; only
their gradient and Hessian enter the assembled system.

`Dual2` carries, for a batch of m elements at once, a value (m,), a gradient (m, n) and a
Hessian (m, n, n) over n partials, and propagates them through +, −, ×, ÷, powers and the
numpy functions sin, cos, tan, exp, log, sqrt, tanh, arctan (chain rule to second order).
"""
import numpy as np


class Dual2:
    __array_priority__ = 100

    def __init__(self, v, g, h, hzero=False):
        self.v, self.g, self.h = v, g, h
        self.hzero = hzero          # h is known to be zero (a seeded partial)

    @staticmethod
    def seed(value, slot, n, scale=1.0, first_order=False):
        """The variable `value` (m,) as partial number `slot` of n, scaled: x = value + scale·δ.
        first_order: no second derivatives are carried (h = None through the arithmetic), for
        callers that want the gradient only."""
        v = np.asarray(value, dtype=float)
        g = np.zeros(v.shape + (n,))
        g[..., slot] = scale
        return Dual2(v, g, None if first_order else np.zeros(v.shape + (n, n)), hzero=True)

    # ---- helpers
    def _const(self, c):
        c = np.asarray(c, dtype=float)
        return Dual2(np.broadcast_to(c, self.v.shape).astype(float), np.zeros_like(self.g),
                     None if self.h is None else np.zeros_like(self.h), hzero=True)

    def _lift(self, o):
        return o if isinstance(o, Dual2) else self._const(o)

    def chain(self, f, f1, f2):
        """f(x) with f, f′, f″ evaluated at x.v."""
        g1 = f1[..., None] * self.g
        if self.h is None:
            return Dual2(f, g1, None)
        h = f1[..., None, None] * self.h + f2[..., None, None] * (self.g[..., :, None] * self.g[..., None, :])
        return Dual2(f, g1, h)

    # ---- arithmetic
    def __add__(self, o):
        if not isinstance(o, Dual2):
            return Dual2(self.v + o, self.g, self.h, self.hzero)
        h = None if self.h is None or o.h is None else self.h + o.h
        return Dual2(self.v + o.v, self.g + o.g, h, self.hzero and o.hzero)

    __radd__ = __add__

    def __neg__(self):
        return Dual2(-self.v, -self.g, None if self.h is None else -self.h, self.hzero)

    def __sub__(self, o):
        return self + (-o)

    def __rsub__(self, o):
        return (-self) + o

    def __mul__(self, o):
        if not isinstance(o, Dual2):
            o = np.asarray(o, dtype=float)
            return Dual2(self.v * o, self.g * o[..., None], None if self.h is None else self.h * o[..., None, None],
                         self.hzero)
        g = self.g * o.v[..., None] + o.g * self.v[..., None]
        if self.h is None or o.h is None:
            return Dual2(self.v * o.v, g, None)
        gg = self.g[..., :, None] * o.g[..., None, :]
        return Dual2(self.v * o.v, g,
                     self.h * o.v[..., None, None] + o.h * self.v[..., None, None] + gg + np.swapaxes(gg, -1, -2))

    __rmul__ = __mul__

    def reciprocal(self):
        x = self.v
        return self.chain(1. / x, -1. / x ** 2, 2. / x ** 3)

    def __truediv__(self, o):
        if not isinstance(o, Dual2):
            return self * (1. / np.asarray(o, dtype=float))
        return self * o.reciprocal()

    def __rtruediv__(self, o):
        return self.reciprocal() * o

    def __pow__(self, p):
        if isinstance(p, Dual2):
            return np.exp(p * np.log(self))
        x = self.v
        return self.chain(x ** p, p * x ** (p - 1), p * (p - 1) * x ** (p - 2) if p != 1 else np.zeros_like(x))

    def __rpow__(self, b):
        return np.exp(self * np.log(b))

    # ---- numpy functions
    _FUN = {
        np.sin: lambda x: (np.sin(x), np.cos(x), -np.sin(x)),
        np.cos: lambda x: (np.cos(x), -np.sin(x), -np.cos(x)),
        np.tan: lambda x: (np.tan(x), 1. / np.cos(x) ** 2, 2. * np.tan(x) / np.cos(x) ** 2),
        np.exp: lambda x: (np.exp(x), np.exp(x), np.exp(x)),
        np.log: lambda x: (np.log(x), 1. / x, -1. / x ** 2),
        np.sqrt: lambda x: (np.sqrt(x), 0.5 / np.sqrt(x), -0.25 / x ** 1.5),
        np.tanh: lambda x: (np.tanh(x), 1. - np.tanh(x) ** 2, -2. * np.tanh(x) * (1. - np.tanh(x) ** 2)),
        np.arctan: lambda x: (np.arctan(x), 1. / (1. + x * x), -2. * x / (1. + x * x) ** 2),
    }

    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        if method != "__call__":
            return NotImplemented
        if ufunc in self._FUN and len(inputs) == 1:
            return self.chain(*self._FUN[ufunc](self.v))
        if ufunc is np.add:
            return inputs[0] + inputs[1] if isinstance(inputs[0], Dual2) else inputs[1] + inputs[0]
        if ufunc is np.multiply:
            return inputs[0] * inputs[1] if isinstance(inputs[0], Dual2) else inputs[1] * inputs[0]
        if ufunc is np.subtract:
            return inputs[0] - inputs[1] if isinstance(inputs[0], Dual2) else -(inputs[1] - inputs[0])
        if ufunc is np.true_divide:
            return inputs[0] / inputs[1] if isinstance(inputs[0], Dual2) else inputs[1].__rtruediv__(inputs[0])
        if ufunc is np.power and isinstance(inputs[0], Dual2):
            return inputs[0] ** inputs[1]
        if ufunc is np.negative:
            return -inputs[0]
        return NotImplemented


def compose(v, g, H, x):
    """y = f(x) as a Dual2 from f's value v (m,), gradient g (m, k) and Hessian H (m, k, k)
    over k inputs x (a list of Dual2 or arrays): the chain rule to second order,
    ∇y = gᵀ∇x and ∇²y = ∇xᵀ H ∇x + Σ_k g_k ∇²x_k. H = None: f's Hessian left out (a
    Gauss-Newton composition)."""
    dx = [xi for xi in x if isinstance(xi, Dual2)]
    if not dx:
        return np.asarray(v, dtype=float)
    n = dx[0].g.shape[-1]
    m = np.shape(v)[0]
    G = np.stack([xi.g if isinstance(xi, Dual2) else np.zeros((m, n)) for xi in x], axis=1)          # m,k,n
    yg = np.einsum("mk,mkn->mn", g, G)
    if any(xi.h is None for xi in dx):   # first-order partials
        return Dual2(np.asarray(v, dtype=float), yg, None)
    # ∇xᵀ H ∇x as two contractions (H is zero for a Gauss-Newton composition), then the
    # inputs' own Hessians (zero for seeded partials)
    yh = np.einsum("mkn,mkp->mnp", G, np.einsum("mkl,mlp->mkp", H, G)) if H is not None else np.zeros((m, n, n))
    for k, xi in enumerate(x):
        if isinstance(xi, Dual2) and not xi.hzero:
            yh += g[:, k, None, None] * xi.h
    return Dual2(np.asarray(v, dtype=float), yg, yh)


def S(λ, g, γ):
    """Complementary slackness g·λ − γ (src/BasicElements.jl:289)."""
    return g * λ - γ


def KKT(λ, g, γ):
    """The pseudo-potential of src/BasicElements.jl:291-307: value 0, "gradient"
    λ·∇g + S(λ, g, γ)·∇λ, and as second derivatives the Jacobian of that gradient,
    H_ij = ∂_i(λ∇g + S∇λ)/∂p_j over the solver's nested partials (not symmetric)."""
    lg, gg = isinstance(λ, Dual2), isinstance(g, Dual2)
    if not lg and not gg:
        return 0.
    ref = λ if lg else g
    m, n = ref.g.shape
    first = (lg and λ.h is None) or (gg and g.h is None)
    lv = λ.v if lg else np.broadcast_to(np.asarray(λ, float), (m,))
    gv = g.v if gg else np.broadcast_to(np.asarray(g, float), (m,))
    lgr = λ.g if lg else np.zeros((m, n))
    ggr = g.g if gg else np.zeros((m, n))
    s = S(lv, gv, γ)
    grad = lv[:, None] * ggr + s[:, None] * lgr
    if first:
        return Dual2(np.zeros(m), grad, None)
    dS = gv[:, None] * lgr + lv[:, None] * ggr
    lh = λ.h if lg else np.zeros((m, n, n))
    gh = g.h if gg else np.zeros((m, n, n))
    H = (ggr[:, :, None] * lgr[:, None, :] + lv[:, None, None] * gh + lgr[:, :, None] * dS[:, None, :]
         + s[:, None, None] * lh)
    return Dual2(np.zeros(m), grad, H)


def exp10(x):
    """10^x for floats and Dual2 (Muscade's exp10 in test cost functions)."""
    return np.exp(x * np.log(10.)) if isinstance(x, Dual2) else 10. ** x


def call_cost(cost, args, costargs):
    """cost(args..., costargs...): a tuple splats positionally, a dict (NamedTuple) by name."""
    if isinstance(costargs, dict):
        return cost(*args, **costargs)
    return cost(*args, *tuple(costargs))
