"""Forward-mode dual numbers for the host-side tables.

Bolt differentiates with ForwardDiff: `ForwardDiff.derivative(f, x)` calls f
with a Dual number and reads its partial (test/runtests.jl:54-80 differentiate
bg.η(-5) and ih.csb²(0) with respect to Ω_b).  Here a CosmoParams field may hold
a Dual; Background then computes the tables together with their derivative
along the Dual parts (the C entry point bolt_host_tables_jvp, forward-mode AD
through Background and RECFAST), and its splines evaluate to Duals.

The device carries the derivative on: with a Dual background, hierarchy!
(rhs_batch), initial_conditions_batch, boltsolve_batch, source_grid and cltt
return Duals, computed by the tangent-linear solve and LOS (bolt_*_jvp).
plin, the polarization sources and the batched entry points take plain
cosmologies.
"""
import numpy as np


class Dual:
    """v + d·ε with ε² = 0 (one partial, like ForwardDiff.Dual{T,V,1}).
    v and d may be floats or numpy arrays of one shape."""

    __slots__ = ("v", "d")

    def __init__(self, v, d=0.0):
        self.v = v
        self.d = d

    @staticmethod
    def _p(o):
        return (o.v, o.d) if isinstance(o, Dual) else (o, 0.0)

    def __add__(self, o):
        v, d = Dual._p(o)
        return Dual(self.v + v, self.d + d)

    __radd__ = __add__

    def __sub__(self, o):
        v, d = Dual._p(o)
        return Dual(self.v - v, self.d - d)

    def __rsub__(self, o):
        v, d = Dual._p(o)
        return Dual(v - self.v, d - self.d)

    def __mul__(self, o):
        v, d = Dual._p(o)
        return Dual(self.v * v, self.d * v + self.v * d)

    __rmul__ = __mul__

    def __truediv__(self, o):
        v, d = Dual._p(o)
        q = self.v / v
        return Dual(q, (self.d - q * d) / v)

    def __rtruediv__(self, o):
        v, d = Dual._p(o)
        q = v / self.v
        return Dual(q, (d - q * self.d) / self.v)

    def __neg__(self):
        return Dual(-self.v, -self.d)

    def __pow__(self, p):
        if isinstance(p, Dual):
            raise TypeError("Dual ** Dual is not supported")
        return Dual(self.v ** p, p * self.v ** (p - 1) * self.d)

    def __lt__(self, o):
        return self.v < Dual._p(o)[0]

    def __le__(self, o):
        return self.v <= Dual._p(o)[0]

    def __gt__(self, o):
        return self.v > Dual._p(o)[0]

    def __ge__(self, o):
        return self.v >= Dual._p(o)[0]

    def __repr__(self):
        return f"Dual({self.v!r}, {self.d!r})"


def value(x):
    return x.v if isinstance(x, Dual) else x


def partial(x):
    return x.d if isinstance(x, Dual) else (np.zeros_like(x) if isinstance(x, np.ndarray) else 0.0)


def derivative(f, x):
    """ForwardDiff.derivative(f, x) for functions of one real that build their
    result from host-side quantities (Background, IonizationHistory)."""
    return partial(f(Dual(float(x), 1.0)))


PARAM_FIELDS = ("h", "Ω_r", "Ω_b", "Ω_c", "A", "n", "Y_p", "N_ν", "Σm_ν")


def tables_jacobian(par, x_grid=(-20.0, 0.0, 2001), nq=15):
    """The host tables and their derivatives with respect to every CosmoParams
    field in one pass (ForwardDiff's chunked Duals; bolt_host_tables_jacobian).
    Returns (Background(par), dcoef [9][NTAB][nx+2], dscal [9][NSCAL]) with the
    partials in PARAM_FIELDS order."""
    import ctypes
    from . import _abi
    from ._abi import check, lib, pd
    from .api import Background
    x0, x1, nx = float(x_grid[0]), float(x_grid[1]), int(x_grid[2])
    coef = np.zeros((_abi.NTAB, nx + 2))
    scal = np.zeros(_abi.NSCAL)
    dcoef = np.zeros((len(PARAM_FIELDS), _abi.NTAB, nx + 2))
    dscal = np.zeros((len(PARAM_FIELDS), _abi.NSCAL))
    xq, wq = np.zeros(nq), np.zeros(nq)
    check(lib().bolt_host_tables_jacobian(ctypes.byref(par.c_struct()), x0, x1, nx, int(nq), pd(coef), pd(scal),
                                          pd(dcoef), pd(dscal), pd(xq), pd(wq)))
    return Background(par, (x0, x1, nx), nq, _tables=(coef, scal, xq, wq)), dcoef, dscal
