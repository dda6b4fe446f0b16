"""Host side: the C ABI library, the spline tables, Background and RECFAST.

These run without a GPU (host code only)."""
import ctypes
import os
import re

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

import boltmi
from boltmi import _abi
import cosmo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "bolt_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bolt_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    L = _abi.lib()
    names = header_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
        assert name in _abi.SIGNATURES, f"{name} has no ctypes signature"


def test_abi_version_and_state_length():
    L = _abi.lib()
    assert L.bolt_abi_version() == 1
    cfg = boltmi.hier_cfg(50, 50, 20, 15)
    assert L.bolt_state_length(ctypes.byref(cfg)) == 473  # perturbations.jl:90
    cfg = boltmi.hier_cfg(8, 8, 10, 15)
    assert L.bolt_state_length(ctypes.byref(cfg)) == 197


def test_no_gpu_reports_device_error_not_crash():
    # In a GPU-less container context creation must fail with an error code.
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and not os.path.exists("/dev/kfd"):
        with pytest.raises(_abi.BoltError) as e:
            boltmi.Context(0)
        assert e.value.code == _abi.BOLT_E_DEVICE


def test_invalid_arguments_are_rejected():
    par = boltmi.CosmoParams(Ω_r=-1.0)
    with pytest.raises(_abi.BoltError):
        boltmi.Background(par)


def test_gauss_legendre_matches_numpy():
    par, bg, ih = cosmo.background("default")
    x, w = np.polynomial.legendre.leggauss(15)
    np.testing.assert_allclose(bg.quad_pts, x, atol=2e-15)
    np.testing.assert_allclose(bg.quad_wts, w, atol=2e-15)


def test_tables_are_natural_cubic_splines():
    """Interpolations' BSpline(Cubic(Line(OnGrid()))) is the natural cubic spline
    through the node values; spline_∂ₓ re-splines its node derivatives."""
    par, bg, ih = cosmo.background("default")
    xg = bg.x_grid
    for tab in (bg.H, bg.eta, ih.tau, ih.g):
        y = tab(xg)
        ref = CubicSpline(xg, y, bc_type="natural")
        mid = 0.5 * (xg[1:] + xg[:-1])
        scale = np.max(np.abs(y))
        assert np.max(np.abs(tab(mid) - ref(mid))) < 1e-12 * scale
    dref = CubicSpline(xg, bg.H(xg), bc_type="natural")(xg, 1)
    assert np.max(np.abs(bg.H_p(xg) - dref)) < 1e-10 * np.max(np.abs(dref))
    d2ref = CubicSpline(xg, bg.H(xg), bc_type="natural")(xg, 2)
    assert np.max(np.abs(bg.H_pp(xg) - d2ref)) < 1e-8 * np.max(np.abs(d2ref))


def test_background_closed_forms():
    par, bg, ih = cosmo.background("nonu")
    # H₀ = h · 100 km/s/Mpc / c in Mpc⁻¹ (background.jl:6, Bolt.jl:49)
    assert bg.H0 == pytest.approx(0.7 * 100.0 / 299792.458, rel=1e-15)
    # ρ_crit = 3H₀²/8πG with G in Mpc² (CODATA 2018)
    G = 6.67430e-11 * 1.054571817e-34 / 299792458.0 ** 3 / 3.0856775814913673e22 ** 2
    assert bg.rho_crit == pytest.approx(3 * bg.H0 ** 2 / (8 * np.pi * G), rel=1e-14)
    # massless-limit neutrino density: 7/8 π²/15 T_ν⁴ (scripts/dev_fd_quadrature.jl:7),
    # up to the 15-point quadrature error
    Tnu = (par.N_ν / 3) ** 0.25 * (4 / 11) ** (1 / 3) * (15 / np.pi ** 2 * bg.rho_crit * par.Ω_r) ** 0.25
    a = 1e-7
    assert bg.rho0M(np.log(a)) * a ** 4 == pytest.approx(7 / 8 * np.pi ** 2 / 15 * Tnu ** 4, rel=1e-4)
    # radiation era: ℋ ≈ H₀ √Ω_rad / a
    x = -19.0
    om_rad = par.Ω_r * (1 + (7 * par.N_ν / 8) * (4 / 11) ** (4 / 3))
    assert bg.H(x) == pytest.approx(bg.H0 * np.sqrt(om_rad) / np.exp(x), rel=1e-3)
    # η(0) = ∫ dx/ℋ(x) (conformal time), direct quadrature vs the 15-point GL of
    # background.jl:72-77 (rel 1e-3: quadrature error of the reference's rule)
    from scipy.integrate import quad
    eta_quad = quad(lambda xx: 1.0 / bg.H(xx), -20.0, 0.0, limit=400)[0] + np.exp(-20.0) / (bg.H0 * np.sqrt(om_rad))
    assert bg.eta0 == pytest.approx(eta_quad, rel=1e-3)
    assert bg.eta0 == pytest.approx(bg.eta(0.0), rel=1e-12)


def test_recfast_matches_fortran_fixture(golden):
    """runtests.jl:38-48: X_e vs Fortran RECFAST at 1e-4 absolute (last 5 rows dropped)."""
    d = golden("recfast_xe.npz")
    z, xe = d["z"][:-5], d["xe"][:-5]
    par = boltmi.CosmoParams(Σm_ν=0.0, N_ν=3.0, Ω_r=5.042e-5)
    out = boltmi.recfast_xe(par, z, Tnow=2.725)
    assert np.max(np.abs(out - xe)) < 1e-4


def test_ionization_history_shape():
    par, bg, ih = cosmo.background("default")
    xg = bg.x_grid
    xe = ih.Xe(xg)
    fHe = 0.24 / (3.9715 * 0.76)
    assert xe[0] == pytest.approx(1 + 2 * fHe, rel=1e-12)        # z > 8000 (recfast.jl:415)
    assert xe[-1] == pytest.approx(1 + 2 * fHe, rel=1e-6)        # fully reionized H + He today
    assert np.all(ih.tau_p(xg) <= 0)                             # τ′ = -n_e σ_T a/ℋ
    assert ih.tau(0.0) == pytest.approx(0.0, abs=1e-12)
    # visibility integrates to ~1 over the grid (∫ g̃ dx = 1 - e^{-τ(x_0)}, up to
    # the trapezoid discretisation of τ)
    g = ih.g(xg)
    assert np.trapz(g, xg) == pytest.approx(1.0, rel=2e-3)


def _box(n, seed):
    """cosmologies drawn from the batch workload's parameter box"""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        h = rng.uniform(0.6, 0.8)
        out.append(boltmi.CosmoParams(h=h, Ω_b=rng.uniform(0.020, 0.024) / h ** 2, Ω_c=rng.uniform(0.10, 0.14) / h ** 2,
                                      n=rng.uniform(0.92, 1.0), A=np.exp(rng.uniform(2.9, 3.2)) / 1e10,
                                      Σm_ν=(rng.uniform(0.06, 0.3) if i % 2 else 0.0) * boltmi.MASS_NATURAL))
    return out


def test_host_tables_batch_is_bitwise_the_single_builds():
    pars = _box(6, 3)
    grid = (-20.0, 0.0, 501)
    built = boltmi.backgrounds(pars, x_grid=grid, nq=15, threads=3)
    for par, bg, ih in built:
        one = boltmi.Background(par, x_grid=grid, nq=15)
        assert np.array_equal(bg.coef, one.coef) and np.array_equal(bg.scal, one.scal)
        assert np.array_equal(bg.quad_pts, one.quad_pts) and np.array_equal(bg.quad_wts, one.quad_wts)
        assert ih.tau0 == one.scal[_abi.S_TAU0] and bg.eta(-3.0) == one.eta(-3.0)
    assert boltmi.backgrounds([], x_grid=grid) == []


def test_host_tables_batch_reports_each_failure_and_fills_the_rest():
    pars = _box(4, 5)
    pars[2] = boltmi.CosmoParams(h=-0.7)  # rejected by bolt_host_tables
    nx, nq = 201, 15
    cp = (_abi.CosmoParamsC * 4)(*[p.c_struct() for p in pars])
    coef = np.zeros((4, _abi.NTAB, nx + 2))
    scal = np.zeros((4, _abi.NSCAL))
    xq, wq = np.zeros(nq), np.zeros(nq)
    status = np.zeros(4, dtype=np.int32)
    L = _abi.lib()
    rc = L.bolt_host_tables_batch(cp, 4, -20.0, 0.0, nx, nq, 2, _abi.pd(coef), _abi.pd(scal), _abi.pd(xq),
                                  _abi.pd(wq), _abi.pi(status))
    assert rc == _abi.BOLT_E_NUMERIC and b"cosmology 2" in L.bolt_last_error()
    assert list(status) == [0, 0, _abi.BOLT_E_ARG, 0]
    for c in (0, 1, 3):
        one = boltmi.Background(pars[c], x_grid=(-20.0, 0.0, nx), nq=nq)
        assert np.array_equal(coef[c], one.coef)
    with pytest.raises(_abi.BoltError):
        boltmi.backgrounds(pars, x_grid=(-20.0, 0.0, nx))


def test_k_grids_quadratic_and_log10():
    """quadratic_k and log10_k (spectra.jl:60-68) by their defining properties:
    the quadratic grid's first point lies above kmin (i starts at 1), its last is
    kmax and its square-root spacing is uniform; the log grid runs from kmin to
    kmax with a constant ratio."""
    kmin, kmax, nk = 0.01 * 2.3349e-4, 1000 * 2.3349e-4, 5000
    q = boltmi.quadratic_k(kmin, kmax, nk)
    assert len(q) == nk and q[0] > kmin and q[-1] == kmax
    assert q[0] == pytest.approx(kmin + (kmax - kmin) / nk ** 2, rel=1e-15)
    s = np.sqrt((q - kmin) / (kmax - kmin))
    np.testing.assert_allclose(np.diff(s), 1.0 / nk, rtol=1e-9)
    g = boltmi.log10_k(1e-4, 0.5, 800)
    assert len(g) == 800 and g[0] == pytest.approx(1e-4, rel=1e-14) and g[-1] == pytest.approx(0.5, rel=1e-14)
    r = g[1:] / g[:-1]
    np.testing.assert_allclose(r, r[0], rtol=1e-12)
    np.testing.assert_allclose(g, np.logspace(-4, np.log10(0.5), 800), rtol=1e-13)
