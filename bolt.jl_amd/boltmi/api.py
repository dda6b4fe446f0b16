"""Bolt.jl's hot-path API on top of libbolt_hip.so.

Names, argument meaning and defaults follow Bolt.jl (citations into src/):
CosmoParams (Bolt.jl:56-66), Background (background.jl:104-128),
IonizationHistory(RECFAST(bg), par, bg) (recfast.jl:674-726), Hierarchy
(perturbations.jl:7-21), boltsolve / boltsolve_rsa (perturbations.jl:25-111),
source_grid (spectra.jl:6-23), quadratic_k / log10_k (spectra.jl:60-68),
cltt (spectra.jl:84-96, 132-150) and plin (spectra.jl:163-198).

The background and ionization history are computed on the host (C++, FP64)
and uploaded once per cosmology; everything downstream runs on the GPU.

Python identifiers cannot hold subscript digits, primes or superscripts, so
Bolt's bg.ℋ′ / bg.η₀ / ih.csb² ... are spelled bg.H_p / bg.eta0 / ih.csb2 here
(bg.H, bg.H_p, bg.H_pp, bg.eta, bg.eta_p, bg.eta_pp, bg.rho0M, bg.H0, bg.eta0,
bg.rho_crit, bg.Omega_L; ih.Xe, ih.tau, ih.tau_p, ih.tau_pp, ih.g, ih.g_p,
ih.g_pp, ih.Tmat, ih.csb2, ih.tau0).
"""
import ctypes
import itertools
import math
import threading
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from ._abi import check, lib, pd, pi
from .dual import Dual, partial, value

# natural units (c = ħ = k_B = 1, Mpc): 1 eV in Mpc^-1 (Bolt.jl:51)
MASS_NATURAL = 1.602176634e-19 / (1.054571817e-34 * 299792458.0) * 3.0856775814913673e22


@dataclass
class CosmoParams:
    h: float = 0.7
    Ω_r: float = 5.0469e-5
    Ω_b: float = 0.046
    Ω_c: float = 0.224
    A: float = 2.097e-9
    n: float = 1.0
    Y_p: float = 0.24
    N_ν: float = 3.046
    Σm_ν: float = 0.06 * MASS_NATURAL

    def as_tuple(self):
        return (self.h, self.Ω_r, self.Ω_b, self.Ω_c, self.A, self.n, self.Y_p, self.N_ν, self.Σm_ν)

    def is_dual(self):
        return any(isinstance(v, Dual) for v in self.as_tuple())

    def c_struct(self):
        if self.is_dual():
            raise TypeError("this entry point takes a plain cosmology; Dual parameters reach the device through "
                            "Background (their tables' derivative) and cosmo_id (bolt_cosmo_upload_jvp)")
        return _abi.CosmoParamsC(*[float(v) for v in self.as_tuple()])

    def c_structs_dual(self):
        """(values, tangent direction) of a CosmoParams holding Dual fields."""
        return (_abi.CosmoParamsC(*[float(value(v)) for v in self.as_tuple()]),
                _abi.CosmoParamsC(*[float(partial(v)) for v in self.as_tuple()]))


class _Spline:
    """A table of the background / ionization history, callable at x."""

    def __init__(self, owner, tab, deriv=0):
        self._o, self._t, self._d = owner, tab, deriv

    def __call__(self, x):
        xs = np.atleast_1d(np.asarray(x, dtype=np.float64)).copy()
        out = np.empty_like(xs)
        check(lib().bolt_host_spline_eval(ctypes.byref(self._o.c_tables()), self._t, self._d, pd(xs), len(xs),
                                          pd(out)))
        if self._o.dcoef is None:
            return out if np.ndim(x) else float(out[0])
        # the spline is linear in its coefficients: its tangent is the spline of the tangent coefficients
        dout = np.empty_like(xs)
        check(lib().bolt_host_spline_eval(ctypes.byref(self._o.c_tables(tangent=True)), self._t, self._d, pd(xs),
                                          len(xs), pd(dout)))
        return Dual(out, dout) if np.ndim(x) else Dual(float(out[0]), float(dout[0]))


class Background:
    """Background(par; x_grid=-20.0:0.01:0.0, nq=15) with the RECFAST
    ionization history tables computed alongside (both host-side).

    If fields of par are Dual numbers, the tables' derivative along their
    partials is computed too (forward-mode AD, bolt_host_tables_jvp): the
    splines and scalars then evaluate to Duals, as under ForwardDiff."""

    def __init__(self, par: CosmoParams, x_grid=(-20.0, 0.0, 2001), nq=15, _tables=None):
        x0, x1, nx = x_grid
        self.par = par
        self.nx, self.nq, self.x0, self.x1 = int(nx), int(nq), float(x0), float(x1)
        self.coef = np.zeros((_abi.NTAB, self.nx + 2))
        self.scal = np.zeros(_abi.NSCAL)
        self.quad_pts = np.zeros(self.nq)
        self.quad_wts = np.zeros(self.nq)
        self.dcoef = self.dscal = None
        if _tables is not None:  # built by backgrounds()
            self.coef, self.scal, self.quad_pts, self.quad_wts = _tables
        elif par.is_dual():
            pv, pt = par.c_structs_dual()
            self.dcoef = np.zeros_like(self.coef)
            self.dscal = np.zeros_like(self.scal)
            check(lib().bolt_host_tables_jvp(ctypes.byref(pv), ctypes.byref(pt), self.x0, self.x1, self.nx,
                                             self.nq, pd(self.coef), pd(self.scal), pd(self.dcoef), pd(self.dscal),
                                             pd(self.quad_pts), pd(self.quad_wts)))
        else:
            check(lib().bolt_host_tables(ctypes.byref(par.c_struct()), self.x0, self.x1, self.nx, self.nq,
                                         pd(self.coef), pd(self.scal), pd(self.quad_pts), pd(self.quad_wts)))
        self.H0 = self._scalar(_abi.S_H0)
        self.eta0 = self._scalar(_abi.S_ETA0)
        self.rho_crit = self._scalar(_abi.S_RHO_CRIT)
        self.Omega_L = self._scalar(_abi.S_OMEGA_L)
        self.x_grid = np.array([(self.x0 * (self.nx - 1 - j) + self.x1 * j) / (self.nx - 1)
                                for j in range(self.nx)])
        self.H = _Spline(self, _abi.T_HC)
        self.H_p = _Spline(self, _abi.T_HC_P)
        self.H_pp = _Spline(self, _abi.T_HC_PP)
        self.eta = _Spline(self, _abi.T_ETA)
        self.eta_p = _Spline(self, _abi.T_ETA_P)
        self.eta_pp = _Spline(self, _abi.T_ETA_PP)
        self.rho0M = _Spline(self, _abi.T_RHO0M)
        self._ct = None
        self._ct_tangent = None
        # uploads of these tables: Context serial -> (weak reference to the
        # Context, cosmology id); released when this Background is collected
        self._uploads = {}
        weakref.finalize(self, _release_uploads, self._uploads)

    def _scalar(self, i):
        return float(self.scal[i]) if self.dscal is None else Dual(float(self.scal[i]), float(self.dscal[i]))

    def c_tables(self, tangent=False):
        if tangent:
            if self._ct_tangent is None:
                self._ct_tangent = _abi.TablesC(self.nx, self.nq, self.x0, self.x1,
                                                (ctypes.c_double * _abi.NSCAL)(*self.dscal), pd(self.quad_pts),
                                                pd(self.quad_wts), pd(self.dcoef))
            return self._ct_tangent
        if self._ct is None:
            self._ct = _abi.TablesC(self.nx, self.nq, self.x0, self.x1, (ctypes.c_double * _abi.NSCAL)(*self.scal),
                                    pd(self.quad_pts), pd(self.quad_wts), pd(self.coef))
        return self._ct

    def tables_dict(self):
        return dict(nx=self.nx, nq=self.nq, x0=self.x0, x1=self.x1, H0=self.H0, eta0=self.eta0,
                    rho_crit=self.rho_crit, xq=self.quad_pts, wq=self.quad_wts, coef=self.coef)

    @property
    def is_dual(self):
        return self.dcoef is not None

    def cosmo_id(self, ctx):
        """The id of these tables on ctx (uploaded once per context); a Dual
        background uploads its tangent too (bolt_cosmo_upload_jvp)."""
        if not ctx.handle:
            raise ValueError("the context is closed")
        ent = self._uploads.get(ctx.serial)
        if ent is None:
            cid = ctypes.c_int(-1)
            if self.is_dual:
                pv, pt = self.par.c_structs_dual()
                check(lib().bolt_cosmo_upload_jvp(ctx.handle, ctypes.byref(pv), ctypes.byref(pt),
                                                  ctypes.byref(self.c_tables()), pd(self.dcoef), pd(self.dscal),
                                                  ctypes.byref(cid)))
            else:
                check(lib().bolt_cosmo_upload(ctx.handle, ctypes.byref(self.par.c_struct()),
                                              ctypes.byref(self.c_tables()), ctypes.byref(cid)))
            ent = (weakref.ref(ctx), cid.value)
            self._uploads[ctx.serial] = ent
        return ent[1]


def _release_uploads(uploads):
    """free a collected Background's device tables on every live context"""
    for ref, cid in list(uploads.values()):
        ctx = ref()
        if ctx is not None and ctx.handle:
            try:
                lib().bolt_cosmo_release(ctx.handle, cid)
            except Exception:
                pass
    uploads.clear()


def backgrounds(pars, x_grid=(-20.0, 0.0, 2001), nq=15, threads=0, device=False, ctx=None):
    """[(par, Background(par), IonizationHistory(RECFAST(bg), par, bg)) for par
    in pars], the tables built on `threads` host threads (0: all hardware
    threads) by bolt_host_tables_batch; each bitwise what Background(par) builds.
    This is the table stage of a many-cosmology batch (source_grid_batch).
    device=True builds them on the GPU instead (bolt_device_tables_batch, one
    workgroup per cosmology), equal to the host's to RECFAST's ODE tolerance."""
    pars = list(pars)
    x0, x1, nx = float(x_grid[0]), float(x_grid[1]), int(x_grid[2])
    nc = len(pars)
    coef = np.zeros((nc, _abi.NTAB, nx + 2))
    scal = np.zeros((nc, _abi.NSCAL))
    xq, wq = np.zeros(nq), np.zeros(nq)
    status = np.zeros(max(nc, 1), dtype=np.int32)
    cp = (_abi.CosmoParamsC * max(nc, 1))(*[p.c_struct() for p in pars])
    if device:
        check(lib().bolt_device_tables_batch((ctx or default_context()).handle, cp, nc, x0, x1, nx, int(nq), pd(coef),
                                             pd(scal), pd(xq), pd(wq), pi(status)))
    else:
        check(lib().bolt_host_tables_batch(cp, nc, x0, x1, nx, int(nq), int(threads), pd(coef), pd(scal), pd(xq),
                                           pd(wq), pi(status)))
    out = []
    for c, par in enumerate(pars):
        bg = Background(par, (x0, x1, nx), nq, _tables=(coef[c], scal[c], xq.copy(), wq.copy()))
        out.append((par, bg, IonizationHistory(RECFAST(bg), par, bg)))
    return out


class RECFAST:
    """RECFAST(bg; kws...) marker (recfast.jl:22-124); the default constants are
    the only configuration the hot path consumes."""

    def __init__(self, bg: Background):
        self.bg = bg


class IonizationHistory:
    """IonizationHistory(RECFAST(bg), par, bg): the ionization tables."""

    def __init__(self, rec: RECFAST, par: CosmoParams, bg: Background):
        self.bg = bg
        self.tau0 = bg._scalar(_abi.S_TAU0)
        self.Xe = _Spline(bg, _abi.T_XE)
        self.tau = _Spline(bg, _abi.T_TAU)
        self.tau_p = _Spline(bg, _abi.T_TAU_P)
        self.tau_pp = _Spline(bg, _abi.T_TAU_PP)
        self.g = _Spline(bg, _abi.T_G)
        self.g_p = _Spline(bg, _abi.T_G_P)
        self.g_pp = _Spline(bg, _abi.T_G_PP)
        self.Tmat = _Spline(bg, _abi.T_TMAT)
        self.csb2 = _Spline(bg, _abi.T_CSB2)


def recfast_xe(par: CosmoParams, z, Tnow=-1.0):
    """Xe_RECFAST on redshifts z (recfast.jl:413-433)."""
    z = np.ascontiguousarray(np.atleast_1d(z), dtype=np.float64)
    out = np.zeros_like(z)
    check(lib().bolt_host_recfast_xe(ctypes.byref(par.c_struct()), Tnow, pd(z), len(z), pd(out)))
    return out


# ------------------------------------------------------------------- device
class Context:
    """One HIP context (device + stream) of libbolt_hip."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib().bolt_ctx_create(device, ctypes.byref(h)))
        self.handle = h
        self.lock = threading.Lock()
        # never reused, unlike id(): keys the uploads a Background keeps
        self.serial = next(_ctx_serial)

    def close(self):
        if self.handle:
            lib().bolt_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_ctx_serial = itertools.count(1)
_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class PerturbationIntegrator:
    pass


class BasicNewtonian(PerturbationIntegrator):
    """Bolt's Newtonian-gauge hierarchy; here it runs on the GPU."""


HIPNewtonian = BasicNewtonian


@dataclass
class Hierarchy:
    integrator: PerturbationIntegrator
    par: CosmoParams
    bg: Background
    ih: IonizationHistory
    k: float
    ℓᵧ: int = 8
    ℓ_ν: int = 8
    ℓ_mν: int = 10
    nq: int = 15

    def cfg(self, reltol=1e-6, abstol=1e-6, max_steps=0):
        return hier_cfg(self.ℓᵧ, self.ℓ_ν, self.ℓ_mν, self.nq, reltol, abstol, max_steps)

    @property
    def n(self):
        return 2 * (self.ℓᵧ + 1) + (self.ℓ_ν + 1) + (self.ℓ_mν + 1) * self.nq + 5


FP32_SOLVE = 1  # bolt_hier_cfg.flags: FP32 stage solves with FP64 iterative refinement


def precision_flags(precision="fp64", refine=0):
    """cfg.flags for a stage-solve precision: "fp64", or "fp32" (FP32 factors and
    substitutions, refined against the FP64 residual until every correction is
    below 1e-14 of its component, at most `refine` sweeps per stage; 0 selects 12)"""
    if precision == "fp64":
        return 0
    if precision == "fp32":
        if not 0 <= refine <= 15:
            raise ValueError("refine must be 0..15")
        return FP32_SOLVE | (refine << 4)
    raise ValueError(f"unknown precision {precision!r}")


def hier_cfg(ℓᵧ=8, ℓ_ν=8, ℓ_mν=10, nq=15, reltol=1e-6, abstol=1e-6, max_steps=0, flags=0):
    return _abi.HierCfgC(ℓᵧ, ℓ_ν, ℓ_mν, nq, reltol, abstol, max_steps, flags)


def state_length(cfg):
    return 2 * (cfg.ell_g + 1) + (cfg.ell_nu + 1) + (cfg.ell_mnu + 1) * cfg.nq + 5


class Solution:
    """Saved states of boltsolve; sol(x) for x in the saved set."""

    def __init__(self, xs, states, stats):
        self.t = np.asarray(xs)
        self.u = states
        self.stats = stats

    def __call__(self, x):
        i = np.nonzero(self.t == x)[0]
        if len(i) == 0:
            raise KeyError(f"x = {x} was not in saveat")
        return self.u[i[0]]


def _stats_dict(st):
    return dict(accepted=st.steps_accepted, rejected=st.steps_rejected, rhs_evals=st.rhs_evals,
                linear_solves=st.linear_solves, kernel_ms=st.kernel_ms)


def rhs_batch(par, bg, cfg, ks, x, U, mutate=False, ctx=None):
    """hierarchy! for modes ks at x; U: [n][B] (mode fastest). Returns (dU, U').
    With a Dual background and U a Dual (U, U̇): returns the Dual du (its
    partial is J U̇ + (∂f/∂p) U) and U."""
    ctx = ctx or default_context()
    if bg.is_dual:
        if mutate:
            raise ValueError("the tangent of hierarchy! is taken without the RSA write")
        Uv, Ud = (U.v, U.d) if isinstance(U, Dual) else (U, np.zeros_like(U))
        ks = np.ascontiguousarray(ks, dtype=np.float64)
        Uv = np.array(Uv, dtype=np.float64, order="C")
        Ud = np.array(Ud, dtype=np.float64, order="C")
        dU, dUd = np.zeros_like(Uv), np.zeros_like(Uv)
        with ctx.lock:
            cid = bg.cosmo_id(ctx)
            check(lib().bolt_rhs_batch(ctx.handle, cid, ctypes.byref(cfg), pd(ks), len(ks), x, pd(Uv), pd(dU), 0))
            check(lib().bolt_rhs_jvp_batch(ctx.handle, cid, ctypes.byref(cfg), pd(ks), len(ks), x, pd(Uv), pd(Ud),
                                           pd(dUd)))
        return Dual(dU, dUd), Uv
    ks = np.ascontiguousarray(ks, dtype=np.float64)
    U = np.array(U, dtype=np.float64, order="C")
    dU = np.zeros_like(U)
    with ctx.lock:
        check(lib().bolt_rhs_batch(ctx.handle, bg.cosmo_id(ctx), ctypes.byref(cfg), pd(ks), len(ks), x, pd(U), pd(dU),
                                   1 if mutate else 0))
    return dU, U


def initial_conditions_batch(par, bg, cfg, ks, ctx=None):
    """initial_conditions [n][B]; a Dual with a Dual background"""
    ctx = ctx or default_context()
    ks = np.ascontiguousarray(ks, dtype=np.float64)
    U = np.zeros((state_length(cfg), len(ks)))
    if bg.is_dual:
        Ud = np.zeros_like(U)
        with ctx.lock:
            check(lib().bolt_initial_conditions_jvp(ctx.handle, bg.cosmo_id(ctx), ctypes.byref(cfg), pd(ks), len(ks),
                                                    pd(U), pd(Ud)))
        return Dual(U, Ud)
    with ctx.lock:
        check(lib().bolt_initial_conditions(ctx.handle, bg.cosmo_id(ctx), ctypes.byref(cfg), pd(ks), len(ks), pd(U)))
    return U


def boltsolve_batch(par, bg, cfg, ks, x_save, ctx=None, raise_on_fail=True):
    """States [B][nsave][n] at x_save for modes ks, plus per-mode status and stats."""
    ctx = ctx or default_context()
    ks = np.ascontiguousarray(ks, dtype=np.float64)
    xs = np.ascontiguousarray(x_save, dtype=np.float64)
    n = state_length(cfg)
    out = np.zeros((len(ks), len(xs), n))
    status = np.zeros(len(ks), dtype=np.int32)
    st = _abi.StatsC()
    if bg.is_dual:  # the tangent-linear solve: states and their derivatives, as a Dual
        outd = np.zeros_like(out)
        with ctx.lock:
            rc = lib().bolt_boltsolve_jvp_batch(ctx.handle, bg.cosmo_id(ctx), ctypes.byref(cfg), pd(ks), len(ks),
                                                pd(xs), len(xs), pd(out), pd(outd), pi(status), ctypes.byref(st))
        if raise_on_fail or rc != _abi.BOLT_E_NUMERIC:
            check(rc)
        return Dual(out, outd), status, _stats_dict(st)
    with ctx.lock:
        rc = lib().bolt_boltsolve_batch(ctx.handle, bg.cosmo_id(ctx), ctypes.byref(cfg), pd(ks), len(ks), pd(xs),
                                        len(xs), pd(out), pi(status), ctypes.byref(st))
    if raise_on_fail or rc != _abi.BOLT_E_NUMERIC:
        check(rc)
    return out, status, _stats_dict(st)


def boltsolve(h: Hierarchy, ode_alg=None, reltol=1e-6, abstol=1e-6, saveat=None, ctx=None):
    """boltsolve(hierarchy; reltol, abstol) (perturbations.jl:25-33); the
    solution is kept at `saveat` (default: the background's x grid)."""
    xs = h.bg.x_grid if saveat is None else np.asarray(saveat, dtype=np.float64)
    out, status, st = boltsolve_batch(h.par, h.bg, h.cfg(reltol, abstol), [h.k], xs, ctx)
    return Solution(xs, out[0], st)


def boltsolve_rsa(h: Hierarchy, ode_alg=None, reltol=1e-6, abstol=1e-6, ctx=None):
    """results[pertlen, nx] with the RSA overwrite (perturbations.jl:86-111)."""
    ctx = ctx or default_context()
    cfg = h.cfg(reltol, abstol)
    ks = np.array([h.k], dtype=np.float64)
    out = np.zeros((1, h.bg.nx, h.n))
    status = np.zeros(1, dtype=np.int32)
    st = _abi.StatsC()
    with ctx.lock:
        check(lib().bolt_boltsolve_rsa_batch(ctx.handle, h.bg.cosmo_id(ctx), ctypes.byref(cfg), pd(ks), 1, pd(out),
                                             pi(status), ctypes.byref(st)))
    return out[0].T.copy()


class SourceGrid:
    """The interpolated source S(x, k): Gridded linear in x and k with Line
    extrapolation, like LinearInterpolation((x_grid, k_grid), grid,
    extrapolation_bc=Line()) (spectra.jl:21)."""

    def __init__(self, x_grid, k_grid, grid, stats=None):
        self.x_grid = np.asarray(x_grid)
        self.k_grid = np.asarray(k_grid)
        self.grid = grid          # [nk][nx]
        self.stats = stats

    @staticmethod
    def _locate(g, v):
        i = int(np.clip(np.searchsorted(g, v, side="right") - 1, 0, len(g) - 2))
        return i, (v - g[i]) / (g[i + 1] - g[i])

    def __call__(self, x, k):
        ix, tx = self._locate(self.x_grid, x)
        ik, tk = self._locate(self.k_grid, k)
        g = self.grid
        a = (1 - tx) * g[ik][ix] + tx * g[ik][ix + 1]
        b = (1 - tx) * g[ik + 1][ix] + tx * g[ik + 1][ix + 1]
        return (1 - tk) * a + tk * b


def source_grid(par, bg, ih, k_grid, integrator=None, ℓᵧ=8, reltol=1e-11, ℓ_ν=8, ℓ_mν=10, nq=15, abstol=1e-6,
                ctx=None, flags=0):
    """source_grid(par, bg, ih, k_grid, BasicNewtonian(); ℓᵧ=8, reltol=1e-11)
    (spectra.jl:6-23).  As in Bolt, ℓ_ν, ℓ_mν and nq keep the Hierarchy
    defaults unless given."""
    ctx = ctx or default_context()
    ks = np.ascontiguousarray(k_grid, dtype=np.float64)
    cfg = hier_cfg(ℓᵧ, ℓ_ν, ℓ_mν, nq, reltol, abstol, flags=flags)
    S = np.zeros((len(ks), bg.nx))
    status = np.zeros(len(ks), dtype=np.int32)
    st = _abi.StatsC()
    if bg.is_dual:  # ForwardDiff through source_grid: the grid is a Dual (values, derivatives)
        dS = np.zeros_like(S)
        with ctx.lock:
            check(lib().bolt_source_grid_jvp(ctx.handle, bg.cosmo_id(ctx), ctypes.byref(cfg), pd(ks), len(ks), pd(S),
                                             pd(dS), pi(status), ctypes.byref(st)))
        return SourceGrid(bg.x_grid, ks, Dual(S, dS), _stats_dict(st))
    with ctx.lock:
        check(lib().bolt_source_grid(ctx.handle, bg.cosmo_id(ctx), ctypes.byref(cfg), pd(ks), len(ks), pd(S),
                                     pi(status), ctypes.byref(st)))
    return SourceGrid(bg.x_grid, ks, S, _stats_dict(st))


def source_grid_P(par, bg, ih, k_grid, integrator=None, ℓᵧ=8, reltol=1e-11, ℓ_ν=8, ℓ_mν=10, nq=15, abstol=1e-6,
                  ctx=None):
    """source_grid_P(par, bg, ih, k_grid, BasicNewtonian(); ℓᵧ=8, reltol=1e-11)
    (spectra.jl:25-42): the polarization source (perturbations.jl:385-404) on
    every x node; ±Inf at the last node, where y = k(η(0) − η(x)) = 0."""
    return source_grids_TP(par, bg, ih, k_grid, integrator, ℓᵧ, reltol, ℓ_ν, ℓ_mν, nq, abstol, ctx,
                           want_t=False)[1]


def source_grids_TP(par, bg, ih, k_grid, integrator=None, ℓᵧ=8, reltol=1e-11, ℓ_ν=8, ℓ_mν=10, nq=15, abstol=1e-6,
                    ctx=None, want_t=True, want_p=True):
    """(source_grid(...), source_grid_P(...)) from one solve per k-mode."""
    ctx = ctx or default_context()
    ks = np.ascontiguousarray(k_grid, dtype=np.float64)
    cfg = hier_cfg(ℓᵧ, ℓ_ν, ℓ_mν, nq, reltol, abstol)
    S = np.zeros((len(ks), bg.nx)) if want_t else None
    SP = np.zeros((len(ks), bg.nx)) if want_p else None
    status = np.zeros(len(ks), dtype=np.int32)
    st = _abi.StatsC()
    with ctx.lock:
        check(lib().bolt_source_grid_tp(ctx.handle, bg.cosmo_id(ctx), ctypes.byref(cfg), pd(ks), len(ks),
                                        pd(S) if want_t else None, pd(SP) if want_p else None, pi(status),
                                        ctypes.byref(st)))
    stats = _stats_dict(st)
    return (SourceGrid(bg.x_grid, ks, S, stats) if want_t else None,
            SourceGrid(bg.x_grid, ks, SP, stats) if want_p else None)


def _per(v, n):
    return list(v) if isinstance(v, (list, tuple, np.ndarray)) else [v] * n


def source_grid_batch(cosmos, k_grids, ℓᵧ=8, reltol=1e-11, ℓ_ν=8, ℓ_mν=10, nq=15, abstol=1e-6, want_t=True,
                      want_p=False, ctx=None, flags=0):
    """source_grid (and/or source_grid_P) for many cosmologies in one launch.

    cosmos: [(par, bg, ih), ...]; k_grids: one k grid for all, or one per
    cosmology (equal lengths); hierarchy sizes and tolerances: a scalar or one
    per cosmology.  Returns ([SourceGrid] or None, [SourceGrid] or None), each
    grid identical to what source_grid gives for that cosmology alone."""
    ctx = ctx or default_context()
    nc = len(cosmos)
    kg = np.asarray(k_grids, dtype=np.float64)
    if kg.ndim == 1:
        kg = np.tile(kg, (nc, 1))
    kg = np.ascontiguousarray(kg)
    nk = kg.shape[1]
    cfgs = (_abi.HierCfgC * nc)()
    for c, (g, n_, m_, q_, rt, at) in enumerate(zip(_per(ℓᵧ, nc), _per(ℓ_ν, nc), _per(ℓ_mν, nc), _per(nq, nc),
                                                    _per(reltol, nc), _per(abstol, nc))):
        cfgs[c] = hier_cfg(g, n_, m_, q_, rt, at, flags=flags)
    nx = cosmos[0][1].nx
    S = np.zeros((nc, nk, nx)) if want_t else None
    SP = np.zeros((nc, nk, nx)) if want_p else None
    status = np.zeros((nc, nk), dtype=np.int32)
    st = _abi.StatsC()
    with ctx.lock:
        ids = np.array([bg.cosmo_id(ctx) for (_, bg, _) in cosmos], dtype=np.int32)
        check(lib().bolt_source_grid_batch(ctx.handle, nc, pi(ids), cfgs, pd(kg), nk, pd(S) if want_t else None,
                                           pd(SP) if want_p else None, pi(status), ctypes.byref(st)))
    stats = _stats_dict(st)
    out_t = [SourceGrid(cosmos[c][1].x_grid, kg[c], S[c], stats) for c in range(nc)] if want_t else None
    out_p = [SourceGrid(cosmos[c][1].x_grid, kg[c], SP[c], stats) for c in range(nc)] if want_p else None
    return out_t, out_p


def autotune(par, bg, cfg, k_grid, kgrid=None, ell_max=0, ctx=None):
    """Time every solver shape (rows per lane of the substitution scans, 5..8)
    on source_grid over k_grid with hierarchy cfg, and, given a dense kgrid
    and ell_max, every LOS shape (x slots per wavefront 1, 2, 4) on cltt for
    ℓ = 2..ell_max; later calls on this device use the fastest.  Returns the
    choices and the per-shape times (ms)."""
    ctx = ctx or default_context()
    ks = np.ascontiguousarray(k_grid, dtype=np.float64)
    kg = None if kgrid is None else np.ascontiguousarray(kgrid, dtype=np.float64)
    rep = _abi.TuneReportC()
    with ctx.lock:
        check(lib().bolt_autotune(ctx.handle, bg.cosmo_id(ctx), ctypes.byref(cfg), pd(ks), len(ks),
                                  None if kg is None else pd(kg), 0 if kg is None else len(kg), int(ell_max),
                                  ctypes.byref(rep)))
    return {"solver_sr": rep.solver_sr, "solver_ms": {sr: float(rep.solver_ms[sr - 5]) for sr in range(6, 9)},
            "los_spw": rep.los_spw, "los_ms": {spw: float(rep.los_ms[i]) for i, spw in enumerate((1, 2, 4))}}


def quadratic_k(kmin, kmax, nk):
    """quadratic_k (spectra.jl:60-63): k_i = kmin + (kmax-kmin)(i/nk)², i = 1..nk.
    Dual bounds are taken at their values (assume_nondual, spectra.jl:46-47, 61)."""
    kmin, kmax = value(kmin), value(kmax)
    return np.array([kmin + (kmax - kmin) * (i / nk) ** 2 for i in range(1, nk + 1)])


def log10_k(kmin, kmax, nk):
    """log10_k (spectra.jl:65-68), Dual bounds at their values."""
    kmin, kmax = value(kmin), value(kmax)
    return np.array([10 ** (math.log10(kmin) + math.log10(kmax / kmin) * (i - 1) / (nk - 1))
                     for i in range(1, nk + 1)])


def cltt_on_grid(ells, sf: SourceGrid, kgrid, par, bg, bessel_mode=0, ctx=None):
    """cltt(ℓ, s_itp, kgrid, par, bg) (spectra.jl:84-96) for a list of ℓ.  With a
    Dual background (and a Dual source grid from it) the result is a Dual: C_ℓ
    and its derivative along the parameters' partials (bolt_cltt_jvp)."""
    ctx = ctx or default_context()
    if bg.is_dual or isinstance(sf.grid, Dual):
        return _cltt_jvp(ells, sf, kgrid, bg, bessel_mode, ctx)
    ells_a = np.ascontiguousarray(np.atleast_1d(ells), dtype=np.int32)
    S = np.ascontiguousarray(sf.grid, dtype=np.float64)
    ksrc = np.ascontiguousarray(sf.k_grid, dtype=np.float64)
    kg = np.ascontiguousarray(kgrid, dtype=np.float64)
    cl = np.zeros(len(ells_a))
    with ctx.lock:
        check(lib().bolt_cltt(ctx.handle, bg.cosmo_id(ctx), pd(S), pd(ksrc), len(ksrc), pd(kg), len(kg), pi(ells_a),
                              len(ells_a), bessel_mode, pd(cl)))
    return cl if np.ndim(ells) else float(cl[0])


def _cltt_jvp(ells, sf, kgrid, bg, bessel_mode, ctx):
    if not bg.is_dual:
        raise TypeError("a Dual source grid needs the Dual background it was solved with (its η tables' derivative)")
    if bessel_mode != 0:
        raise ValueError("the derivative of C_ℓ is computed with the exact j_ℓ (bessel_mode 0)")
    ells_a = np.ascontiguousarray(np.atleast_1d(ells), dtype=np.int32)
    g = sf.grid
    S = np.ascontiguousarray(g.v if isinstance(g, Dual) else g, dtype=np.float64)
    dS = np.ascontiguousarray(g.d if isinstance(g, Dual) else np.zeros_like(S), dtype=np.float64)
    ksrc = np.ascontiguousarray(sf.k_grid, dtype=np.float64)
    kg = np.ascontiguousarray(kgrid, dtype=np.float64)
    cl = np.zeros(len(ells_a))
    dcl = np.zeros(len(ells_a))
    with ctx.lock:
        check(lib().bolt_cltt_jvp(ctx.handle, bg.cosmo_id(ctx), pd(S), pd(dS), pd(ksrc), len(ksrc), pd(kg), len(kg),
                                  pi(ells_a), len(ells_a), pd(cl), pd(dcl)))
    return Dual(cl, dcl) if np.ndim(ells) else Dual(float(cl[0]), float(dcl[0]))


def cltt(ells, par, bg, ih, sf, ctx=None):
    """cltt(ℓ, par, bg, ih, sf) and cltt(ℓ⃗, ...) (spectra.jl:132-150): the dense
    grid is quadratic_k(0.01H₀, 1000H₀, 5000)."""
    dense = quadratic_k(0.01 * bg.H0, 1000 * bg.H0, 5000)
    return cltt_on_grid(ells, sf, dense, par, bg, 0, ctx)


def cl_on_grid(ells, sf_t, sf_p, kgrid, par, bg, bessel_mode=0, ctx=None):
    """(C_ℓ^TT, C_ℓ^TE, C_ℓ^EE) from one LOS pass (spectra.jl:84-130); sf_t or
    sf_p may be None, and then the spectra that need it come back as None."""
    ctx = ctx or default_context()
    ells_a = np.ascontiguousarray(np.atleast_1d(ells), dtype=np.int32)
    src = sf_t if sf_t is not None else sf_p
    ksrc = np.ascontiguousarray(src.k_grid, dtype=np.float64)
    S = None if sf_t is None else np.ascontiguousarray(sf_t.grid, dtype=np.float64)
    SP = None if sf_p is None else np.ascontiguousarray(sf_p.grid, dtype=np.float64)
    if sf_t is not None and sf_p is not None and not np.array_equal(sf_t.k_grid, sf_p.k_grid):
        raise ValueError("the temperature and polarization grids must share k_grid")
    kg = np.ascontiguousarray(kgrid, dtype=np.float64)
    n = len(ells_a)
    tt = np.zeros(n) if S is not None else None
    te = np.zeros(n) if (S is not None and SP is not None) else None
    ee = np.zeros(n) if SP is not None else None
    with ctx.lock:
        check(lib().bolt_cl(ctx.handle, bg.cosmo_id(ctx), None if S is None else pd(S), None if SP is None else pd(SP),
                            pd(ksrc), len(ksrc), pd(kg), len(kg), pi(ells_a), n, bessel_mode,
                            None if tt is None else pd(tt), None if te is None else pd(te),
                            None if ee is None else pd(ee)))
    if np.ndim(ells):
        return tt, te, ee
    return tuple(None if v is None else float(v[0]) for v in (tt, te, ee))


def clte_on_grid(ells, sf_t, sf_e, kgrid, par, bg, ctx=None):
    """clte(ℓ, s_itp_t, s_itp_e, kgrid, par, bg) (spectra.jl:98-114)."""
    return cl_on_grid(ells, sf_t, sf_e, kgrid, par, bg, 0, ctx)[1]


def clee_on_grid(ells, sf_p, kgrid, par, bg, ctx=None):
    """clee(ℓ, s_itp_p, kgrid, par, bg) (spectra.jl:116-130)."""
    return cl_on_grid(ells, None, sf_p, kgrid, par, bg, 0, ctx)[2]


def clte(ells, par, bg, ih, sf, sf_P, ctx=None):
    """clte(ℓ, par, bg, ih, sf, sf_P) and clte(ℓ⃗, ...) (spectra.jl:137-140, 152-155)."""
    dense = quadratic_k(0.01 * bg.H0, 1000 * bg.H0, 5000)
    return clte_on_grid(ells, sf, sf_P, dense, par, bg, ctx)


def clee(ells, par, bg, ih, sf_P, ctx=None):
    """clee(ℓ, par, bg, ih, sf_P) and clee(ℓ⃗, ...) (spectra.jl:142-145, 157-160)."""
    dense = quadratic_k(0.01 * bg.H0, 1000 * bg.H0, 5000)
    return clee_on_grid(ells, sf_P, dense, par, bg, ctx)


def plin(k, par, bg, ih, n_q=15, ℓᵧ=50, ℓ_ν=50, ℓ_mν=20, x=0.0, reltol=1e-5, abstol=1e-6, ctx=None, flags=0):
    """plin (spectra.jl:163-198); k may be a scalar or an array (one batched launch)."""
    ctx = ctx or default_context()
    ks = np.ascontiguousarray(np.atleast_1d(k), dtype=np.float64)
    cfg = hier_cfg(ℓᵧ, ℓ_ν, ℓ_mν, n_q, reltol, abstol, flags=flags)
    pk = np.zeros(len(ks))
    status = np.zeros(len(ks), dtype=np.int32)
    st = _abi.StatsC()
    with ctx.lock:
        check(lib().bolt_plin(ctx.handle, bg.cosmo_id(ctx), ctypes.byref(cfg), pd(ks), len(ks), float(x), pd(pk),
                              pi(status), ctypes.byref(st)))
    return pk if np.ndim(k) else float(pk[0])


class Server:
    """Continuous batching of spectrum requests on one context.

    Each request carries its own cosmology; the library builds its tables on
    the host.  queue=False: a request runs on a free slot (one HIP stream with
    preallocated HBM), and requests on different slots run concurrently.
    queue=True: every request's k-modes join one persistent solver launch
    through a work queue the host fills while it runs, and each request's
    device memory is taken from an HBM arena by its own size
    (bolt_serve_start_queue).  Either way each result is the same as the
    computation run alone (plin / source_grid + cltt_on_grid on the same k
    grid)."""

    def __init__(self, slots=4, hbm_gib=8.0, ctx=None, queue=False, host_threads=4):
        self.ctx = ctx or default_context()
        self.queue = queue
        if queue:
            check(lib().bolt_serve_start_queue(self.ctx.handle, float(hbm_gib), int(host_threads)))
            self.slots = 0
        else:
            check(lib().bolt_serve_start(self.ctx.handle, int(slots), float(hbm_gib)))
            self.slots = lib().bolt_serve_slots(self.ctx.handle)
        self._out = {}

    def queue_info(self):
        """launches, requests served, of those admitted into a running launch, arena bytes and peak"""
        q = _abi.QueueInfoC()
        check(lib().bolt_serve_queue_info(self.ctx.handle, ctypes.byref(q)))
        return dict(launches=q.launches, served=q.served, joined_running=q.joined_running,
                    arena_bytes=q.arena_bytes, arena_peak_bytes=q.arena_peak_bytes,
                    solver_workgroups=q.solver_workgroups)

    def _submit(self, par, kind, nk, kmin, kmax, k, ell_max, nkd, cfg, nout):
        out = np.zeros(nout)
        kk = None if k is None else np.ascontiguousarray(k, dtype=np.float64)
        if kk is not None:
            nk = len(kk)
        req = _abi.RequestC(par.c_struct(), kind, int(nk), float(kmin), float(kmax),
                            None if kk is None else pd(kk), int(ell_max), int(nkd), cfg, pd(out))
        rid = ctypes.c_uint64()
        check(lib().bolt_submit(self.ctx.handle, ctypes.byref(req), ctypes.byref(rid)))
        self._out[rid.value] = (out, kk)
        return rid.value

    def submit_pk(self, par, nk=200, kmin=1e-4, kmax=0.5, n_q=15, ℓᵧ=50, ℓ_ν=50, ℓ_mν=20, reltol=1e-5,
                  abstol=1e-6, k=None):
        """plin at x = 0 (spectra.jl:163-198 defaults) on nk log-spaced k, or on the given k."""
        return self._submit(par, _abi.REQ_PK, nk, kmin, kmax, k, 0, 0, hier_cfg(ℓᵧ, ℓ_ν, ℓ_mν, n_q, reltol, abstol),
                            nk if k is None else len(k))

    def submit_cltt(self, par, nk=800, kmin=1e-4, kmax=0.5, ell_max=2500, nkd=5000, ℓᵧ=50, ℓ_ν=50, ℓ_mν=10, nq=15,
                    reltol=1e-11, abstol=1e-6, k=None):
        """source_grid on nk log-spaced k (or the given k), then cltt for ℓ = 2..ell_max."""
        return self._submit(par, _abi.REQ_CLTT, nk, kmin, kmax, k, ell_max, nkd,
                            hier_cfg(ℓᵧ, ℓ_ν, ℓ_mν, nq, reltol, abstol), ell_max - 1)

    @staticmethod
    def _info(r):
        return dict(state=r.state, status=r.status, queue_ms=r.queue_ms, host_ms=r.host_ms, device_ms=r.device_ms,
                    latency_ms=r.latency_ms, solve_ms=r.solve_ms, joined_running=bool(r.joined_running),
                    error=r.error.decode("utf-8", "replace"))

    def poll(self, rid):
        r = _abi.ResultC()
        check(lib().bolt_poll(self.ctx.handle, rid, ctypes.byref(r)))
        return self._info(r)

    def wait(self, rid):
        """(values, info) once the request is done; raises if it failed."""
        r = _abi.ResultC()
        check(lib().bolt_wait(self.ctx.handle, rid, ctypes.byref(r)))
        info = self._info(r)
        out = self._out.pop(rid)[0]
        if r.status != 0:
            raise _abi.BoltError(r.status, f"request {rid} failed: {info['error']}")
        return out, info

    def close(self):
        if self.ctx is not None and self.ctx.handle:
            lib().bolt_serve_stop(self.ctx.handle)
        self._out.clear()
