"""Pins for the CPU oracle: the reference's own fixtures (CLASS perturbations,
CAMB C_ℓ), closed forms (KenCarp4 order conditions, super-horizon ICs,
Bessel functions, the single-shell LOS integral) and exact identities (the
Jacobian of a linear RHS)."""
from fractions import Fraction

import numpy as np
import pytest
from scipy.interpolate import CubicSpline
from scipy.special import spherical_jn

import boltmi
import cosmo
import pyoracle


# ------------------------------------------------------------------ KenCarp4
def _frac(v):
    return Fraction(v).limit_denominator(10 ** 12)


def test_kencarp4_tableau_order_conditions():
    """ESDIRK tableau: row sums = c, b of order 4 (8 conditions), b̂ of order 3
    but not 4, stiffly accurate, and R(∞) = 0 (L-stable)."""
    A, b, bh, c, g = pyoracle.tableau()
    Af = [[_frac(v) for v in row] for row in A]
    bf = [_frac(v) for v in b]
    bhf = [_frac(v) for v in bh]
    cf = [_frac(v) for v in c]
    for i in range(6):
        assert abs(float(sum(Af[i]) - cf[i])) < 1e-15
    def conds(w):
        c_ = cf
        Ac = [sum(Af[i][j] * c_[j] for j in range(6)) for i in range(6)]
        Ac2 = [sum(Af[i][j] * c_[j] ** 2 for j in range(6)) for i in range(6)]
        AAc = [sum(Af[i][j] * Ac[j] for j in range(6)) for i in range(6)]
        return [sum(w) - 1,
                sum(w[i] * c_[i] for i in range(6)) - Fraction(1, 2),
                sum(w[i] * c_[i] ** 2 for i in range(6)) - Fraction(1, 3),
                sum(w[i] * Ac[i] for i in range(6)) - Fraction(1, 6),
                sum(w[i] * c_[i] ** 3 for i in range(6)) - Fraction(1, 4),
                sum(w[i] * c_[i] * Ac[i] for i in range(6)) - Fraction(1, 8),
                sum(w[i] * Ac2[i] for i in range(6)) - Fraction(1, 12),
                sum(w[i] * AAc[i] for i in range(6)) - Fraction(1, 24)]
    assert max(abs(float(v)) for v in conds(bf)) < 1e-14
    ch = conds(bhf)
    assert max(abs(float(v)) for v in ch[:4]) < 1e-14
    assert max(abs(float(v)) for v in ch[4:]) > 1e-6
    assert all(A[5][j] == b[j] for j in range(6))       # stiffly accurate
    # R(z) = 1 + z bᵀ(I - zA)⁻¹ 1 -> 1 - bᵀA⁻¹1 as z -> ∞; with row 1 explicit,
    # use the implicit block
    Ai = np.array(A)[1:, 1:]
    bi = np.array(b)[1:]
    a1 = np.array(A)[1:, 0]
    # stage values for y' = λy with y0 = 1, z -> ∞: Y_2..6 solve A_i Y = -a1
    Y = np.linalg.solve(Ai, -a1)
    assert abs(Y[-1]) < 1e-13


# --------------------------------------------------------------- RHS & ICs
@pytest.mark.parametrize("kind", ["nonu", "massive"])
def test_jacobian_is_exact_for_the_linear_rhs(kind):
    o = cosmo.oracle(kind)
    cfg = cosmo.ocfg(50, 50, 20, 15)
    n = o.n(cfg)
    rng = np.random.default_rng(1)
    for x in [-19.5, -10.0, -7.2, -6.5, -2.0, -0.01]:
        for k in [1e-4, 0.01, 0.3]:
            J = o.jacobian(cfg, k, x)
            v = rng.standard_normal(n)
            f, _ = o.rhs(cfg, k, x, v)
            assert np.max(np.abs(J @ v - f)) <= 1e-12 * np.max(np.abs(f))
            # linearity: f(u + v) = f(u) + f(v)
            u = rng.standard_normal(n)
            fu, _ = o.rhs(cfg, k, x, u)
            fuv, _ = o.rhs(cfg, k, x, u + v)
            assert np.max(np.abs(fuv - fu - f)) <= 1e-12 * np.max(np.abs(fuv))


def test_superhorizon_initial_conditions():
    """Adiabatic ICs (perturbations.jl:274-338): Φ = (4f_ν+10)/(4f_ν+15),
    δ = δ_b = 3Θ₀, v = -3kΘ₁, and the metric is static outside the horizon."""
    par, bg, ih = cosmo.background("nonu")
    o = cosmo.oracle("nonu")
    cfg = cosmo.ocfg(50, 50, 20, 15)
    n = o.n(cfg)
    f_nu = 1 / (1 + 1 / (7 * par.N_ν / 8 * (4 / 11) ** (4 / 3)))
    for k in [1e-4, 1e-2, 0.2]:
        u = o.initial_conditions(cfg, k)
        iphi = n - 5
        assert u[iphi] == pytest.approx((4 * f_nu + 10) / (4 * f_nu + 15), rel=1e-15)
        assert u[iphi + 1] == pytest.approx(3 * u[0], rel=1e-15)
        assert u[iphi + 3] == u[iphi + 1]
        assert u[iphi + 2] == pytest.approx(-3 * k * u[1], rel=1e-15)
        du, _ = o.rhs(cfg, k, -20.0, u)
        # super-horizon (kη ≪ 1): Φ′/Φ and δ′/δ are O((kη)²) small
        assert abs(du[iphi] / u[iphi]) < 1e-3
        assert abs(du[iphi + 1] / u[iphi + 1]) < 1e-3


def test_rsa_switch_mutates_only_photon_and_nu_low_moments():
    o = cosmo.oracle("nonu")
    cfg = cosmo.ocfg(8, 8, 10, 15)
    n = o.n(cfg)
    u = np.random.default_rng(3).standard_normal(n)
    # RSA needs kη > 240 and -τ′ℋ/η > 100 (perturbations.jl:216): x = -11.5, k = 80
    du, um = o.rhs(cfg, 80.0, -11.5, u, mutate=True)
    changed = np.nonzero(um != u)[0]
    iN = 2 * 9
    assert set(changed) <= {0, 1, 2, iN, iN + 1, iN + 2}
    assert np.all(du[:2 * 9 + 9] == 0.0)       # Θ′, Θᵖ′, 𝒩′ zeroed (perturbations.jl:230-232)


# ------------------------------------------------------- reference fixtures
@pytest.mark.parametrize("tag", ["p03", "p01"])
def test_oracle_vs_class_fixture(golden, tag):
    """runtests.jl:83-147: boltsolve_rsa (ℓᵧ=ℓ_ν=50, ℓ_mν=20, tol 1e-9) vs
    CLASS Φ and δ_b at every CLASS x, 1e-3 relative."""
    par, bg, ih = cosmo.background("nonu")
    o = cosmo.oracle("nonu")
    d = golden(f"class_k{tag}.npz")
    k = par.h * float(d["k_hMpc"])
    res, st = o.boltsolve_rsa(cosmo.ocfg(50, 50, 20, 15, 1e-9, 1e-9), k)
    assert st.status == 0
    iphi = o.n(cosmo.ocfg(50, 50, 20, 15)) - 5
    cx = d["x"]
    phi = CubicSpline(bg.x_grid, res[:, iphi])(cx)
    db = CubicSpline(bg.x_grid, res[:, iphi + 3])(cx)
    assert np.max(np.abs(phi / CubicSpline(cx, d["phi"])(cx) - 1)) < 1e-3
    assert np.max(np.abs(-db / CubicSpline(cx, d["delta_b"])(cx) - 1)) < 1e-3


# ------------------------------------------------------------------- Bessel
def test_spherical_bessel_against_scipy_and_mpmath():
    import mpmath
    for z in [1e-4, 0.05, 0.7, 3.0, 10.0, 99.5, 257.3, 1000.0, 2499.7, 3300.0]:
        lmax = 2600
        j = pyoracle.sph_bessel(lmax, z)
        ref = spherical_jn(np.arange(lmax + 1), z)
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(j - ref)) < 1e-12 * scale, z
    for (l, z) in [(2, 0.3), (50, 40.0), (500, 480.0), (2500, 2400.0), (1000, 1500.0)]:
        j = pyoracle.sph_bessel(l, z)[l]
        ref = float(mpmath.sqrt(mpmath.pi / (2 * z)) * mpmath.besselj(l + 0.5, z))
        assert j == pytest.approx(ref, rel=1e-11, abs=1e-300)


def test_bessel_table_is_the_natural_spline_of_jl():
    """bessel_interpolator (spectra.jl:49-58): 5001-node natural cubic spline."""
    zmax = 3300.0
    for ell in [2, 100, 1200]:
        c = pyoracle.bessel_table(ell, zmax)
        nodes = np.linspace(0, zmax, 5001)
        ref = CubicSpline(nodes, spherical_jn(ell, nodes), bc_type="natural")
        for z in [0.1, 17.3, 1000.7, 2222.2, 3299.0]:
            assert pyoracle.bessel_table_eval(c, zmax, z) == pytest.approx(float(ref(z)), abs=1e-13)


def test_cltt_single_shell_known_answer():
    """A source on one x node at comoving distance r, n_s = 1: the LOS integral
    is 4πA(SΔx)² ∫ j_ℓ(kr)² dk/k = 4πA(SΔx)²/(2ℓ(ℓ+1))."""
    par, bg, ih = cosmo.background("default")
    p2 = boltmi.CosmoParams(**{**par.__dict__, "n": 1.0})
    o = pyoracle.Oracle(p2.as_tuple(), bg.tables_dict())
    j0 = 1300                                   # x = -7.0
    k_src = np.array([1e-7, 10.0])
    S = np.zeros((2, bg.nx))
    S[:, j0] = 1.0
    dense = boltmi.quadratic_k(0.01 * bg.H0, 1000 * bg.H0, 5000)
    ells = np.array([2, 10, 40])
    cl = o.cltt(S, k_src, dense, ells, 0, nthreads=8)
    dx = bg.x_grid[j0 + 1] - bg.x_grid[j0]
    expect = 4 * np.pi * p2.A * dx ** 2 / (2 * ells * (ells + 1))
    np.testing.assert_allclose(cl, expect, rtol=2e-3)


def test_cl_polarization_spin_factor_single_shell():
    """With the polarization source equal to the temperature one, TE = ℓð·TT
    and EE = ℓð²·TT, ℓð = √((ℓ+2)(ℓ+1)ℓ(ℓ−1)) (spectra.jl:98-130); TT itself
    is the single-shell closed form above."""
    par, bg, ih = cosmo.background("default")
    p2 = boltmi.CosmoParams(**{**par.__dict__, "n": 1.0})
    o = pyoracle.Oracle(p2.as_tuple(), bg.tables_dict())
    j0 = 1300
    k_src = np.array([1e-7, 10.0])
    S = np.zeros((2, bg.nx))
    S[:, j0] = 1.0
    dense = boltmi.quadratic_k(0.01 * bg.H0, 1000 * bg.H0, 5000)
    ells = np.array([2, 10, 40])
    tt, te, ee = o.cl(S, S, k_src, dense, ells, 0, nthreads=8)
    lf = np.sqrt((ells + 2.0) * (ells + 1) * ells * (ells - 1))
    np.testing.assert_allclose(te, lf * tt, rtol=1e-13)
    np.testing.assert_allclose(ee, lf ** 2 * tt, rtol=1e-13)
    np.testing.assert_array_equal(tt, o.cltt(S, k_src, dense, ells, 0, nthreads=8))
    dx = bg.x_grid[j0 + 1] - bg.x_grid[j0]
    np.testing.assert_allclose(tt, 4 * np.pi * p2.A * dx ** 2 / (2 * ells * (ells + 1)), rtol=2e-3)


def test_source_P_last_node_and_shape():
    """source_function_P = (3/(4y²)) g̃ Π with y = k(η(x_end) − η(x)): the
    value scales as 1/k² at fixed u and x, is linear in Π = Θ₂ + Θᵖ₂ + Θᵖ₀, and
    is infinite at the last node, where y = 0 (as Julia's 3/(4·0²))."""
    par, bg, ih = cosmo.background("default")
    o = cosmo.oracle("default")
    cfg = cosmo.ocfg(8, 8, 10, 15, 1e-6, 1e-6)
    n = o.n(cfg)
    u = np.zeros(n)
    ng = 8
    u[2] = 0.3             # Θ₂
    u[ng + 1] = 0.2        # Θᵖ₀
    u[ng + 3] = -0.1       # Θᵖ₂
    x = -6.9
    a = o.source_P(cfg, 0.01, x, u)
    b = o.source_P(cfg, 0.02, x, u)
    assert a != 0 and np.isfinite(a)
    np.testing.assert_allclose(b, a / 4, rtol=1e-14)
    u2 = u.copy(); u2[2] = 0.6
    np.testing.assert_allclose(o.source_P(cfg, 0.01, x, u2), a * 0.7 / 0.4, rtol=1e-13)
    assert np.isinf(o.source_P(cfg, 0.01, bg.x_grid[-1], u))


@pytest.mark.slow
def test_oracle_cl_vs_camb_fixture(golden):
    """runtests.jl:149-185: source_grid and source_grid_P (100 k, ℓᵧ=8,
    reltol 1e-11), then cltt and clee at ℓ = 10:10:2500 vs CAMB unlensed D_ℓ^TT
    and D_ℓ^EE at 1.1e-1 relative (TE is not asserted there: zero crossings)."""
    from concurrent.futures import ThreadPoolExecutor
    par, bg, ih = cosmo.background("default")
    o = cosmo.oracle("default")
    k_grid = boltmi.quadratic_k(0.1 * bg.H0, 1000 * bg.H0, 100)
    cfg = cosmo.ocfg(8, 8, 10, 15, 1e-11, 1e-6)
    with ThreadPoolExecutor(8) as ex:
        cols = list(ex.map(lambda k: o.source_column_tp(cfg, float(k)), k_grid))
    S = np.array([c[0] for c in cols])
    SP = np.array([c[1] for c in cols])
    S_only, status = o.source_grid(cfg, k_grid, nthreads=8)
    assert not status.any()
    np.testing.assert_array_equal(S, S_only)
    ells = np.arange(10, 2501, 10)
    dense = boltmi.quadratic_k(0.01 * bg.H0, 1000 * bg.H0, 5000)
    tt, te, ee = o.cl(S, SP, k_grid, dense, ells, 0, nthreads=8)
    d = golden("camb_cl.npz")
    fac = ells * (ells + 1) / (2 * np.pi)
    assert np.max(np.abs(fac * tt / np.interp(ells, d["ell"], d["tt"]) - 1)) < 1.1e-1
    assert np.max(np.abs(fac * ee / np.interp(ells, d["ell"], d["ee"]) - 1)) < 1.1e-1
