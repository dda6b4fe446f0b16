"""Moments of spherical Bessel functions and the Filon rule on the device
(Bolt.jl src/bessel/moments.jl, interpolator.jl, integrator.jl).

I_m(x) = ∫_0^x t^m j_ν(t) dt.  `sph_bessel_interpolator` builds the moment
table on the GPU (bolt_moment_table_create); calling it evaluates the table
(spline in range, FP64 Maclaurin series below, Lommel asymptotic above), and
`integrate_sph_bessel_filon` integrates a quadratic against j_ν(kt) with it."""
import ctypes
import math

import numpy as np

from . import _abi
from ._abi import check, lib
from .api import default_context

_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)


def _arr(v):
    return np.ascontiguousarray(np.atleast_1d(np.asarray(v, dtype=np.float64)))


def _pd(a):
    return a.ctypes.data_as(_PD)


class MomentTable:
    """sph_bessel_interpolator's MomentTable (interpolator.jl:40-68), on the device"""

    def __init__(self, nu, order, kη_min, kη_max, N, ctx=None):
        self.ctx = ctx or default_context()
        self.nu, self.order, self.N = int(nu), int(order), int(N)
        self.kη_min, self.kη_max = float(kη_min), float(kη_max)
        h = ctypes.c_void_p()
        check(lib().bolt_moment_table_create(self.ctx.handle, self.nu, self.order, self.kη_min, self.kη_max,
                                             self.N, ctypes.byref(h)))
        self.handle = h

    def __call__(self, x):
        """the moments m = 0..order-1 at x: shape (order,) for a scalar x, (len(x), order) otherwise"""
        xs = _arr(x)
        out = np.zeros((len(xs), self.order))
        check(lib().bolt_moment_table_eval(self.handle, _pd(xs), len(xs), _pd(out)))
        return out[0] if np.ndim(x) == 0 else out

    def coefficients(self):
        c = np.zeros((self.order, self.N + 2))
        check(lib().bolt_moment_table_coefficients(self.handle, _pd(c)))
        return c

    def close(self):
        if getattr(self, "handle", None):
            lib().bolt_moment_table_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __repr__(self):
        return (f"MomentTable(ν={self.nu}) of order {self.order} with kη over "
                f"LinRange({self.kη_min}, {self.kη_max}, {self.N}) on the device")


def sph_bessel_interpolator(nu, order, kη_min, kη_max, N, ctx=None):
    """sph_bessel_interpolator(ν, order, kη_min, kη_max, N) (interpolator.jl:85-111)"""
    return MomentTable(nu, order, kη_min, kη_max, N, ctx)


def getnu(itp):
    return itp.nu


def getorder(itp):
    return itp.order


def integrate_sph_bessel_filon(f, fp, fpp, k, a, b, itp):
    """integrate_sph_bessel_filon(f, f′, f″, k, a, b, itp) (integrator.jl:5-18):
    ∫_a^b q(t) j_ν(kt) dt for the quadratic q with q(a) = f, q′(a) = f′, q″ = f″.
    Scalars give a float; arrays (broadcast together) an array."""
    args = np.broadcast_arrays(*(np.asarray(v, dtype=np.float64) for v in (f, fp, fpp, k, a, b)))
    flat = [np.ascontiguousarray(v.ravel()) for v in args]
    out = np.zeros(flat[0].size)
    check(lib().bolt_filon(itp.handle, *(_pd(v) for v in flat), flat[0].size, _pd(out)))
    return float(out[0]) if args[0].ndim == 0 else out.reshape(args[0].shape)


def filon_sum(itp, t, f, fp, fpp, ks):
    """Σ_i integrate_sph_bessel_filon over consecutive intervals (t_i, t_{i+1})
    with f, f′, f″ at t_i, for each k in ks (the loop of integrator.jl:20-38)"""
    t, f, fp, fpp, ks = (_arr(v) for v in (t, f, fp, fpp, ks))
    out = np.zeros(len(ks))
    check(lib().bolt_filon_sum(itp.handle, _pd(t), _pd(f), _pd(fp), _pd(fpp), len(t), _pd(ks), len(ks), _pd(out)))
    return out


def _moments(method, x, nu, m, ctx):
    x, m = np.broadcast_arrays(np.asarray(x, dtype=np.float64), np.asarray(m, dtype=np.float64))
    shape = x.shape
    xs, ms = np.ascontiguousarray(x.ravel()), np.ascontiguousarray(m.ravel())
    nus = np.ascontiguousarray(np.broadcast_to(np.asarray(nu, dtype=np.int32), shape).ravel())
    out = np.zeros(xs.size)
    check(lib().bolt_sph_moments((ctx or default_context()).handle, method, _pd(xs), nus.ctypes.data_as(_PI),
                                 _pd(ms), xs.size, _pd(out)))
    return float(out[0]) if len(shape) == 0 else out.reshape(shape)


def sph_j_moment_asymp(x, nu, m, ctx=None):
    """sph_j_moment_asymp(x, ν, m) (moments.jl:76-86), the Lommel asymptotic; x > ν"""
    return _moments(0, x, nu, m, ctx)


def sph_j_moment_maclaurin_1F2(x, nu, m, ctx=None):
    """sph_j_moment_maclaurin_₁F₂(x, ν, m) (moments.jl:97-102) in FP64"""
    return _moments(1, x, nu, m, ctx)


def sph_j_moment_series(x, nu, m, ctx=None):
    """what the tables use below x = 40: the series in double-double below 20,
    exact recurrences down to Si(x) from 20 to 40 (integer 0 ≤ m ≤ 3, ν ≤ 3)"""
    return _moments(2, x, nu, m, ctx)


def J_moment(x, nu, alpha, method="asymp", ctx=None):
    """∫_0^x t^α J_ν(t) dt for half-integer ν = n + ½ (J_moment_asymp,
    J_moment_weniger_₁F₂, moments.jl:24-49): √(2/π) times the spherical moment
    of order α + ½"""
    n = nu - 0.5
    if n != int(n) or n < 1:
        raise _abi.BoltError(_abi.BOLT_E_ARG, "J_moment: ν must be a half-integer ≥ 3/2")
    f = sph_j_moment_asymp if method == "asymp" else sph_j_moment_series
    return f(x, int(n), np.asarray(alpha, dtype=np.float64) + 0.5, ctx) / math.sqrt(math.pi / 2)
