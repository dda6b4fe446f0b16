"""ctypes binding of the CPU FP64 oracle (oracle/bolt_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libbolt_oracle.so")

_D = ctypes.c_double
_I = ctypes.c_int
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)


class OrcParams(ctypes.Structure):
    _fields_ = [(n, _D) for n in "h Omega_r Omega_b Omega_c A n Y_p N_nu Sum_m_nu".split()]


class OrcTables(ctypes.Structure):
    _fields_ = [("nx", _I), ("nq", _I), ("x0", _D), ("x1", _D), ("H0", _D), ("eta0", _D),
                ("rho_crit", _D), ("xq", _PD), ("wq", _PD), ("coef", _PD)]


class OrcCfg(ctypes.Structure):
    _fields_ = [("ell_g", _I), ("ell_nu", _I), ("ell_mnu", _I), ("nq", _I), ("reltol", _D),
                ("abstol", _D), ("max_steps", _I), ("lu_refine", _I)]


class OrcMomentTable(ctypes.Structure):
    _fields_ = [("nu", _I), ("order", _I), ("n", _I), ("xmin", _D), ("xmax", _D), ("coef", _PD)]


class OrcStats(ctypes.Structure):
    _fields_ = [("accepted", ctypes.c_int64), ("rejected", ctypes.c_int64),
                ("rhs_evals", ctypes.c_int64), ("factorizations", ctypes.c_int64), ("status", _I)]


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_state_length.argtypes = [ctypes.POINTER(OrcCfg)]
        L.orc_spline.argtypes = [ctypes.POINTER(OrcTables), _I, _D, _I]
        L.orc_spline.restype = _D
        P, T, C = ctypes.POINTER(OrcParams), ctypes.POINTER(OrcTables), ctypes.POINTER(OrcCfg)
        L.orc_rhs.argtypes = [P, T, C, _D, _D, _PD, _PD, _I]
        L.orc_jacobian.argtypes = [P, T, C, _D, _D, _PD]
        L.orc_initial_conditions.argtypes = [P, T, C, _D, _PD]
        L.orc_source.argtypes = [P, T, C, _D, _D, _PD, _PD]
        L.orc_source.restype = _D
        L.orc_boltsolve.argtypes = [P, T, C, _D, _PD, _I, _PD, ctypes.POINTER(OrcStats)]
        L.orc_boltsolve_steps.argtypes = [P, T, C, _D, _PD, _I, _PI, ctypes.POINTER(OrcStats)]
        L.orc_boltsolve_rsa.argtypes = [P, T, C, _D, _PD, ctypes.POINTER(OrcStats)]
        L.orc_source_column.argtypes = [P, T, C, _D, _PD, ctypes.POINTER(OrcStats)]
        L.orc_source_column_tp.argtypes = [P, T, C, _D, _PD, _PD, ctypes.POINTER(OrcStats)]
        L.orc_source_column_replay.argtypes = [P, T, C, _D, _PD, _I, _PD, ctypes.POINTER(OrcStats)]
        L.orc_source_P.argtypes = [P, T, C, _D, _D, _PD]
        L.orc_source_P.restype = _D
        L.orc_source_grid.argtypes = [P, T, C, _PD, _I, _PD, _PI, _I]
        L.orc_plin.argtypes = [P, T, C, _D, _D, ctypes.POINTER(OrcStats)]
        L.orc_plin.restype = _D
        L.orc_sph_bessel.argtypes = [_I, _D, _PD]
        L.orc_bessel_table.argtypes = [_I, _D, _PD]
        L.orc_bessel_table_eval.argtypes = [_PD, _D, _D]
        L.orc_bessel_table_eval.restype = _D
        for name in ("orc_sphj_moment_asymp", "orc_sphj_moment_maclaurin", "orc_sphj_moment_maclaurin_q"):
            getattr(L, name).argtypes = [_D, _I, _D]
            getattr(L, name).restype = _D
        L.orc_linrange.argtypes = [_D, _D, _I, _I]
        L.orc_linrange.restype = _D
        L.orc_J_moment_halfint.argtypes = [_D, _I, _D, _I]
        L.orc_J_moment_halfint.restype = _D
        L.orc_moment_table_build.argtypes = [ctypes.POINTER(OrcMomentTable), _I, _I, _D, _D, _I, _I]
        L.orc_moment_table_free.argtypes = [ctypes.POINTER(OrcMomentTable)]
        L.orc_moment_eval.argtypes = [ctypes.POINTER(OrcMomentTable), _D, _PD]
        L.orc_filon.argtypes = [_D, _D, _D, _D, _D, _D, ctypes.POINTER(OrcMomentTable)]
        L.orc_filon.restype = _D
        L.orc_filon_sum.argtypes = [ctypes.POINTER(OrcMomentTable), _PD, _PD, _PD, _PD, _I, _PD, _I, _PD, _I]
        L.orc_cltt.argtypes = [P, T, _PD, _PD, _I, _PD, _I, _PI, _I, _I, _PD, _I, _I, _I]
        L.orc_cl.argtypes = [P, T, _PD, _PD, _PD, _I, _PD, _I, _PI, _I, _I, _PD, _PD, _PD, _I, _I, _I]
        L.orc_tableau.argtypes = [_PD, _PD, _PD, _PD, _PD]
        L.orc_rho_sigma.argtypes = [P, T, _I, _D, _PD, _PD, _PD, _PD]
        L.orc_theta.argtypes = [P, T, _I, _D, _PD]
        L.orc_theta.restype = _D
        L.orc_omega_nu_planted.argtypes = [P, T, _I]
        L.orc_omega_nu_planted.restype = _D
        L.orc_plin_planted.argtypes = [P, T, C, _D, _D, ctypes.POINTER(OrcStats), _I]
        L.orc_plin_planted.restype = _D
        _lib = L
    return _lib


def _pd(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_PD)


class Oracle:
    """Oracle bound to one cosmology's parameters and tables.

    `par` is a sequence in CosmoParams order, `tables` a dict with keys
    nx, nq, x0, x1, H0, eta0, rho_crit, xq, wq, coef (numpy arrays)."""

    def __init__(self, par, tables):
        self.par = OrcParams(*[float(v) for v in par])
        self._keep = [np.ascontiguousarray(tables[k], dtype=np.float64) for k in ("xq", "wq", "coef")]
        self.tab = OrcTables(int(tables["nx"]), int(tables["nq"]), float(tables["x0"]), float(tables["x1"]),
                             float(tables["H0"]), float(tables["eta0"]), float(tables["rho_crit"]),
                             _pd(self._keep[0]), _pd(self._keep[1]), _pd(self._keep[2]))
        self.nx = int(tables["nx"])

    @staticmethod
    def cfg(ell_g=8, ell_nu=8, ell_mnu=10, nq=15, reltol=1e-6, abstol=1e-6, max_steps=0, lu_refine=1):
        return OrcCfg(ell_g, ell_nu, ell_mnu, nq, reltol, abstol, max_steps, lu_refine)

    def n(self, cfg):
        return lib().orc_state_length(ctypes.byref(cfg))

    def spline(self, tab, x, deriv=0):
        return lib().orc_spline(ctypes.byref(self.tab), tab, x, deriv)

    def rhs(self, cfg, k, x, u, mutate=False):
        u = np.array(u, dtype=np.float64)
        du = np.zeros_like(u)
        lib().orc_rhs(ctypes.byref(self.par), ctypes.byref(self.tab), ctypes.byref(cfg), k, x, _pd(u), _pd(du),
                      1 if mutate else 0)
        return du, u

    def jacobian(self, cfg, k, x):
        n = self.n(cfg)
        J = np.zeros((n, n))
        lib().orc_jacobian(ctypes.byref(self.par), ctypes.byref(self.tab), ctypes.byref(cfg), k, x, _pd(J))
        return J

    def initial_conditions(self, cfg, k):
        u = np.zeros(self.n(cfg))
        lib().orc_initial_conditions(ctypes.byref(self.par), ctypes.byref(self.tab), ctypes.byref(cfg), k, _pd(u))
        return u

    def source(self, cfg, k, x, u, du):
        u = np.ascontiguousarray(u, dtype=np.float64)
        du = np.ascontiguousarray(du, dtype=np.float64)
        return lib().orc_source(ctypes.byref(self.par), ctypes.byref(self.tab), ctypes.byref(cfg), k, x,
                                _pd(u), _pd(du))

    def boltsolve(self, cfg, k, x_save):
        xs = np.ascontiguousarray(x_save, dtype=np.float64)
        out = np.zeros((len(xs), self.n(cfg)))
        st = OrcStats()
        lib().orc_boltsolve(ctypes.byref(self.par), ctypes.byref(self.tab), ctypes.byref(cfg), k, _pd(xs), len(xs),
                            _pd(out), ctypes.byref(st))
        return out, st

    def boltsolve_rsa(self, cfg, k):
        out = np.zeros((self.nx, self.n(cfg)))
        st = OrcStats()
        lib().orc_boltsolve_rsa(ctypes.byref(self.par), ctypes.byref(self.tab), ctypes.byref(cfg), k, _pd(out),
                                ctypes.byref(st))
        return out, st

    def steps(self, cfg, k, max_xs=2000000):
        xs = np.zeros(max_xs)
        cnt = _I(0)
        st = OrcStats()
        lib().orc_boltsolve_steps(ctypes.byref(self.par), ctypes.byref(self.tab), ctypes.byref(cfg), k, _pd(xs),
                                  max_xs, ctypes.byref(cnt), ctypes.byref(st))
        return xs[:min(cnt.value, max_xs)], st

    def source_column(self, cfg, k):
        S = np.zeros(self.nx)
        st = OrcStats()
        lib().orc_source_column(ctypes.byref(self.par), ctypes.byref(self.tab), ctypes.byref(cfg), k, _pd(S),
                                ctypes.byref(st))
        return S, st

    def source_column_replay(self, cfg, k, step_xs):
        """source column along the accepted steps ending at step_xs (from steps())"""
        xs = np.ascontiguousarray(step_xs, dtype=np.float64)
        S = np.zeros(self.nx)
        st = OrcStats()
        lib().orc_source_column_replay(ctypes.byref(self.par), ctypes.byref(self.tab), ctypes.byref(cfg), k,
                                       _pd(xs), len(xs), _pd(S), ctypes.byref(st))
        return S, st

    def source_P(self, cfg, k, x, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        return lib().orc_source_P(ctypes.byref(self.par), ctypes.byref(self.tab), ctypes.byref(cfg), k, x, _pd(u))

    def source_column_tp(self, cfg, k):
        """temperature and polarization source columns from one solve"""
        S = np.zeros(self.nx)
        SP = np.zeros(self.nx)
        st = OrcStats()
        lib().orc_source_column_tp(ctypes.byref(self.par), ctypes.byref(self.tab), ctypes.byref(cfg), k, _pd(S),
                                   _pd(SP), ctypes.byref(st))
        return S, SP, st

    def source_grid(self, cfg, ks, nthreads=0):
        ks = np.ascontiguousarray(ks, dtype=np.float64)
        S = np.zeros((len(ks), self.nx))
        status = np.zeros(len(ks), dtype=np.int32)
        lib().orc_source_grid(ctypes.byref(self.par), ctypes.byref(self.tab), ctypes.byref(cfg), _pd(ks), len(ks),
                              _pd(S), status.ctypes.data_as(_PI), nthreads)
        return S, status

    def plin(self, cfg, k, x=0.0, plant=0):
        """P(k) at x (spectra.jl:163-198).  plant != 0 selects a deliberately
        wrong variant (orc_plin_planted), used only to show that a pin fails."""
        st = OrcStats()
        if plant:
            v = lib().orc_plin_planted(ctypes.byref(self.par), ctypes.byref(self.tab), ctypes.byref(cfg), k, x,
                                       ctypes.byref(st), plant)
        else:
            v = lib().orc_plin(ctypes.byref(self.par), ctypes.byref(self.tab), ctypes.byref(cfg), k, x,
                               ctypes.byref(st))
        return v, st

    def omega_nu(self, plant=0):
        """Ω_ν of plin's massive neutrino (spectra.jl:182-186)"""
        return lib().orc_omega_nu_planted(ctypes.byref(self.par), ctypes.byref(self.tab), plant)

    def theta(self, M1, a=1.0):
        M1 = np.ascontiguousarray(M1, dtype=np.float64)
        return lib().orc_theta(ctypes.byref(self.par), ctypes.byref(self.tab), len(M1), a, _pd(M1))

    def rho_sigma(self, M0, M2, a=1.0):
        M0 = np.ascontiguousarray(M0, dtype=np.float64)
        M2 = np.ascontiguousarray(M2, dtype=np.float64)
        rho, sig = _D(), _D()
        lib().orc_rho_sigma(ctypes.byref(self.par), ctypes.byref(self.tab), len(M0), a, _pd(M0), _pd(M2),
                            ctypes.byref(rho), ctypes.byref(sig))
        return rho.value, sig.value

    def cltt(self, S, k_src, kgrid, ells, bessel_mode=0, nthreads=0, kbeg=0, kend=None):
        S = np.ascontiguousarray(S, dtype=np.float64)
        k_src = np.ascontiguousarray(k_src, dtype=np.float64)
        kgrid = np.ascontiguousarray(kgrid, dtype=np.float64)
        ells = np.ascontiguousarray(ells, dtype=np.int32)
        cl = np.zeros(len(ells))
        if kend is None:
            kend = len(kgrid) - 1
        r = lib().orc_cltt(ctypes.byref(self.par), ctypes.byref(self.tab), _pd(S), _pd(k_src), len(k_src),
                           _pd(kgrid), len(kgrid), ells.ctypes.data_as(_PI), len(ells), bessel_mode, _pd(cl),
                           nthreads, kbeg, kend)
        assert r == 0
        return cl


    def cl(self, S, S_P, k_src, kgrid, ells, bessel_mode=0, nthreads=0, kbeg=0, kend=None):
        """(C_ℓ^TT, C_ℓ^TE, C_ℓ^EE); S_P may be None (then TE/EE are zero)"""
        S = np.ascontiguousarray(S, dtype=np.float64)
        SP = None if S_P is None else np.ascontiguousarray(S_P, dtype=np.float64)
        k_src = np.ascontiguousarray(k_src, dtype=np.float64)
        kgrid = np.ascontiguousarray(kgrid, dtype=np.float64)
        ells = np.ascontiguousarray(ells, dtype=np.int32)
        tt, te, ee = np.zeros(len(ells)), np.zeros(len(ells)), np.zeros(len(ells))
        if kend is None:
            kend = len(kgrid) - 1
        r = lib().orc_cl(ctypes.byref(self.par), ctypes.byref(self.tab), _pd(S), None if SP is None else _pd(SP),
                         _pd(k_src), len(k_src), _pd(kgrid), len(kgrid), ells.ctypes.data_as(_PI), len(ells),
                         bessel_mode, _pd(tt), _pd(te), _pd(ee), nthreads, kbeg, kend)
        assert r == 0
        return tt, te, ee


def tableau():
    A = np.zeros((6, 6)); b = np.zeros(6); bh = np.zeros(6); c = np.zeros(6); g = np.zeros(1)
    lib().orc_tableau(_pd(A), _pd(b), _pd(bh), _pd(c), _pd(g))
    return A, b, bh, c, g[0]


def sph_bessel(lmax, z):
    j = np.zeros(lmax + 1)
    lib().orc_sph_bessel(lmax, z, _pd(j))
    return j


def bessel_table(ell, zmax):
    c = np.zeros(5003)
    lib().orc_bessel_table(ell, zmax, _pd(c))
    return c


def bessel_table_eval(c, zmax, z):
    return lib().orc_bessel_table_eval(_pd(c), zmax, z)


# ---- moments of j_ν, their tables and the Filon rule (oracle/bessel_moments.c)
def sphj_moment(x, nu, m, method="asymp"):
    """∫_0^x t^m j_ν(t) dt: method "asymp" (Lommel), "maclaurin64" (FP64 ₁F₂
    series), "series" (the same series in __float128, the tables' nodes below 40)"""
    f = {"asymp": lib().orc_sphj_moment_asymp, "maclaurin64": lib().orc_sphj_moment_maclaurin,
         "series": lib().orc_sphj_moment_maclaurin_q}[method]
    return f(float(x), int(nu), float(m))


def J_moment_halfint(x, n, alpha, method="asymp"):
    """∫_0^x t^α J_{n+½}(t) dt"""
    return lib().orc_J_moment_halfint(float(x), int(n), float(alpha), 0 if method == "asymp" else 1)


class MomentTable:
    """sph_bessel_interpolator(ν, order, x_min, x_max, N) (interpolator.jl:85-111)"""

    def __init__(self, nu, order, xmin, xmax, n, nthreads=8):
        self.t = OrcMomentTable()
        r = lib().orc_moment_table_build(ctypes.byref(self.t), nu, order, float(xmin), float(xmax), int(n), nthreads)
        assert r == 0
        self.order = order

    def __call__(self, x):
        out = np.zeros(self.order)
        lib().orc_moment_eval(ctypes.byref(self.t), float(x), _pd(out))
        return out

    def coefficients(self):
        return np.ctypeslib.as_array(self.t.coef, shape=(self.order, self.t.n + 2)).copy()

    def filon(self, f, fp, fpp, k, a, b):
        return lib().orc_filon(float(f), float(fp), float(fpp), float(k), float(a), float(b), ctypes.byref(self.t))

    def filon_sum(self, t, f, fp, fpp, ks, nthreads=8):
        t, f, fp, fpp, ks = (np.ascontiguousarray(v, dtype=np.float64) for v in (t, f, fp, fpp, ks))
        out = np.zeros(len(ks))
        lib().orc_filon_sum(ctypes.byref(self.t), _pd(t), _pd(f), _pd(fp), _pd(fpp), len(t), _pd(ks), len(ks),
                            _pd(out), nthreads)
        return out

    def __del__(self):
        try:
            lib().orc_moment_table_free(ctypes.byref(self.t))
        except Exception:
            pass
