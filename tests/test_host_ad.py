"""Forward-mode AD through the host tables (Background, RECFAST, ionization
history): bolt_host_tables_jvp and boltmi.Dual / boltmi.derivative.

Bolt's own AD tests (test/runtests.jl:54-80) differentiate bg.η(-5) and
ih.csb²(0) with respect to Ω_b and compare with a central difference.  The
checks here pin every table and scalar against Richardson-extrapolated central
differences, the tangent map's linearity, the parameters the tables cannot
depend on, and bitwise equality of the primal with bolt_host_tables.

RECFAST's rate function switches formula at state thresholds (Heflag at
x_He = 0.98, the x_H branches, the Tmat tight-coupling test; recfast.jl:156-310).
Forward-mode AD, ForwardDiff's included, differentiates with those branch
decisions frozen, so it omits the jump a threshold crossing adds to the
derivative; a finite difference includes it.  Inside the recombination window
(700 < z < z_He_evo_start ~ 2800) the ionization tangents therefore differ from
finite differences at the percent level by construction, and are compared only
outside it.  Host code only: no GPU needed.
"""
import ctypes

import numpy as np
import pytest

import boltmi
from boltmi import _abi
from boltmi._abi import check, lib, pd

ZERO = dict(h=0.0, Ω_r=0.0, Ω_b=0.0, Ω_c=0.0, A=0.0, n=0.0, Y_p=0.0, N_ν=0.0, Σm_ν=0.0)
BG_TABS = [_abi.T_HC, _abi.T_HC_P, _abi.T_HC_PP, _abi.T_ETA, _abi.T_ETA_P, _abi.T_ETA_PP, _abi.T_RHO0M]
ION_TABS = [_abi.T_XE, _abi.T_TAU, _abi.T_TAU_P, _abi.T_G, _abi.T_G_P, _abi.T_TMAT, _abi.T_CSB2]
GRID = (-20.0, 0.0, 2001)
# spline-derivative tables are built by differencing node values (util.jl:12-13),
# which multiplies a tangent's rounding-level noise by ~1/dx per order (dx = 0.01)
ORDER = {_abi.T_HC_P: 1, _abi.T_HC_PP: 2, _abi.T_ETA_P: 1, _abi.T_ETA_PP: 2, _abi.T_TAU_P: 1, _abi.T_TAU_PP: 2,
         _abi.T_G_P: 1, _abi.T_G_PP: 2}


SOURCE = {_abi.T_HC_P: _abi.T_HC, _abi.T_HC_PP: _abi.T_HC, _abi.T_ETA_P: _abi.T_ETA, _abi.T_ETA_PP: _abi.T_ETA,
          _abi.T_TAU_P: _abi.T_TAU, _abi.T_TAU_PP: _abi.T_TAU, _abi.T_G_P: _abi.T_G, _abi.T_G_PP: _abi.T_G}


def rtol_for(t, base):
    return base * 100.0 ** ORDER.get(t, 0) / (10.0 if ORDER.get(t, 0) == 2 else 1.0)
MASSIVE = dict(h=0.67, Ω_b=0.049, Ω_c=0.26, Σm_ν=0.2 * boltmi.MASS_NATURAL, Y_p=0.25)


def jvp(par, direction, grid=GRID, nq=15):
    x0, x1, nx = grid
    c = np.zeros((_abi.NTAB, nx + 2))
    dc = np.zeros_like(c)
    s, ds = np.zeros(_abi.NSCAL), np.zeros(_abi.NSCAL)
    xq, wq = np.zeros(nq), np.zeros(nq)
    d = dict(ZERO)
    d.update(direction)
    check(lib().bolt_host_tables_jvp(ctypes.byref(par.c_struct()), ctypes.byref(boltmi.CosmoParams(**d).c_struct()),
                                     x0, x1, nx, nq, pd(c), pd(s), pd(dc), pd(ds), pd(xq), pd(wq)))
    return c, s, dc, ds


def tables(par, grid=GRID):
    bg = boltmi.Background(par, x_grid=grid)
    return bg.coef.copy(), bg.scal.copy()


def nodes(c):
    """spline values at the grid nodes from the padded coefficients"""
    return (c[..., :-2] + 4.0 * c[..., 1:-1] + c[..., 2:]) / 6.0


def fd(base, field, step):
    """central differences at step and step/2 and their Richardson extrapolation
    of the node values and scalars"""
    out = []
    for hh in (step, step / 2):
        p1, p2 = dict(base), dict(base)
        p1[field] = base[field] + hh
        p2[field] = base[field] - hh
        (c1, s1), (c2, s2) = tables(boltmi.CosmoParams(**p1)), tables(boltmi.CosmoParams(**p2))
        out.append(((nodes(c1) - nodes(c2)) / (2 * hh), (s1 - s2) / (2 * hh)))
    (f1, g1), (f2, g2) = out
    c0, _ = tables(boltmi.CosmoParams(**base))
    # rounding floor of a central difference at step/2 of values |v|; a
    # derivative table inherits its source table's rounding over dx^order
    v = np.abs(nodes(c0))
    dx = (GRID[1] - GRID[0]) / (GRID[2] - 1)
    for t, k in ORDER.items():
        v[t] = np.maximum(v[t], v[SOURCE[t]] / dx ** k)
    noise = 16.0 * np.finfo(float).eps * v / step
    return f1, f2, (4 * f2 - f1) / 3, (4 * g2 - g1) / 3, noise


def assert_tangent_close(got, f1, f2, rich, noise, rtol, what):
    """|got - rich| <= rtol * max|rich| + 8 |fd(h) - fd(h/2)| + noise at every
    node: the last two terms bound the finite difference's own error
    (truncation, and rounding where a table is huge and barely moves)."""
    scale = np.max(np.abs(rich))
    bound = rtol * scale + 8.0 * np.abs(f1 - f2) + noise
    bad = np.abs(got - rich) > bound
    assert not bad.any(), (f"{what}: {bad.sum()} nodes off, worst |d| {np.max(np.abs(got - rich)):.3e} "
                           f"scale {scale:.3e}")


def test_jvp_primal_is_bitwise_the_tables():
    for base in ({}, MASSIVE):
        par = boltmi.CosmoParams(**base)
        c, s, _, _ = jvp(par, {"Ω_b": 1.0})
        c0, s0 = tables(par)
        assert np.array_equal(c, c0) and np.array_equal(s, s0)


def test_jvp_is_linear_in_the_direction():
    """The tangent map is linear: exact scaling, and superposition to rounding.
    In RECFAST's ODE stages (700 < z < 2900 here: the helium and hydrogen
    stages start on Saha curves and cross rate branches) the explicit steps
    carry stiff transients of the tangent that amplify rounding, so
    superposition holds there to 1e-5 of the scale."""
    par = boltmi.CosmoParams(**MASSIVE)
    _, _, da, sa = jvp(par, {"Ω_b": 1.0})
    _, _, d2, s2 = jvp(par, {"Ω_b": 2.0})
    for t in range(_abi.NTAB):  # doubling is exact in floating point, except in the subnormal range
        assert np.max(np.abs(d2[t] - 2.0 * da[t])) <= 1e-15 * np.max(np.abs(da[t])) + 1e-300, t
    assert np.array_equal(s2, 2.0 * sa)
    _, _, db, sb = jvp(par, {"h": 1.0})
    _, _, dab, sab = jvp(par, {"Ω_b": 1.0, "h": -2.5})
    x = np.linspace(GRID[0], GRID[1], GRID[2])
    z = np.exp(-x) - 1.0
    near = (z > 700.0) & (z < 2900.0)
    for t in range(_abi.NTAB):
        got, ref = nodes(dab[t]), nodes(da[t]) - 2.5 * nodes(db[t])
        scale = max(np.max(np.abs(nodes(da[t]))), np.max(np.abs(nodes(db[t]))), 1e-300)
        assert np.max(np.abs(got - ref)[~near]) <= 1e-9 * scale, t
        assert np.max(np.abs(got - ref)[near]) <= 1e-5 * scale, t
    np.testing.assert_allclose(sab, sa - 2.5 * sb, rtol=1e-9, atol=0)


def test_jvp_directions_the_tables_do_not_depend_on():
    par = boltmi.CosmoParams(**MASSIVE)
    _, _, d, s = jvp(par, {"A": 1.0, "n": 1.0})  # primordial spectrum only
    assert not d.any() and not s.any()
    _, _, d, s = jvp(par, {"Y_p": 1.0})  # helium enters the ionization history only
    assert not d[BG_TABS].any()
    assert s[_abi.S_H0] == 0 and s[_abi.S_ETA0] == 0 and s[_abi.S_OMEGA_L] == 0
    assert d[_abi.T_XE].any() and d[_abi.T_CSB2].any()


@pytest.mark.parametrize("base,field,step", [
    ({}, "h", 1e-4), ({}, "Ω_b", 1e-4), ({}, "Ω_c", 1e-4), ({}, "Ω_r", 1e-8), ({}, "N_ν", 1e-3),
    (MASSIVE, "Σm_ν", 0.01 * boltmi.MASS_NATURAL), (MASSIVE, "h", 1e-4),
])
def test_background_tangents_match_finite_differences(base, field, step):
    b = dict(boltmi.CosmoParams(**base).__dict__)
    f1, f2, rich, srich, noise = fd(b, field, step)
    _, s, dc, ds = jvp(boltmi.CosmoParams(**b), {field: 1.0})
    got = nodes(dc)
    for t in BG_TABS:
        assert_tangent_close(got[t], f1[t], f2[t], rich[t], noise[t], rtol_for(t, 1e-6), f"{field} table {t}")
    for i in (_abi.S_H0, _abi.S_ETA0, _abi.S_RHO_CRIT, _abi.S_OMEGA_L):
        assert abs(ds[i] - srich[i]) <= 1e-7 * max(abs(srich[i]), 1e-300), (field, i, ds[i], srich[i])


@pytest.mark.parametrize("field,step", [("Ω_b", 1e-4), ("h", 1e-4), ("Y_p", 1e-4), ("Ω_c", 1e-4)])
def test_ionization_tangents_match_finite_differences_outside_recombination(field, step):
    b = dict(boltmi.CosmoParams().__dict__)
    f1, f2, rich, srich, noise = fd(b, field, step)
    _, s, dc, ds = jvp(boltmi.CosmoParams(**b), {field: 1.0})
    got = nodes(dc)
    x = np.linspace(GRID[0], GRID[1], GRID[2])
    z = np.exp(-x) - 1.0
    for t in ION_TABS:
        # the local tables (X_e, T_mat, c_s², and τ′, g̃ which are local in X_e)
        # away from the window; τ and g̃ integrate τ′ from x to 0 and are
        # compared over the late side only
        m = (z < 700.0) | ((z > 3000.0) & (t not in (_abi.T_TAU, _abi.T_G, _abi.T_G_P)))
        assert_tangent_close(got[t][m], f1[t][m], f2[t][m], rich[t][m], noise[t][m], rtol_for(t, 1e-5),
                             f"{field} table {t}")
    assert abs(ds[_abi.S_TAU0] - srich[_abi.S_TAU0]) <= 1e-7 * abs(srich[_abi.S_TAU0])


def test_reference_bg_fwddiff():
    """runtests.jl:54-64: d bg.η(-5) / dΩ_b on x_grid = -20:0.1:0"""
    def fbg(Ω_b):
        bg = boltmi.Background(boltmi.CosmoParams(Ω_b=Ω_b), x_grid=(-20.0, 0.0, 201), nq=15)
        return bg.eta(-5)

    D = 1e-3
    ad = boltmi.derivative(fbg, 0.046)
    assert ((fbg(0.046 + D) - fbg(0.046 - D)) / (2 * D) - ad) < 1e-5  # the reference's own criterion
    # the central difference at D = 1e-3 is off by ~1e-2 (truncation); Richardson pins the AD value
    rich = (4 * (fbg(0.046 + D / 2) - fbg(0.046 - D / 2)) / D - (fbg(0.046 + D) - fbg(0.046 - D)) / (2 * D)) / 3
    assert abs(ad - rich) <= 1e-9 * abs(rich)


def test_reference_ih_fwddiff():
    """runtests.jl:66-80: d ih.csb²(0) / dΩ_b"""
    def fih(Ω_b):
        par = boltmi.CosmoParams(Ω_b=Ω_b)
        bg = boltmi.Background(par, x_grid=(-20.0, 0.0, 201), nq=15)
        ih = boltmi.IonizationHistory(boltmi.RECFAST(bg), par, bg)
        return ih.csb2(0.0)

    D = 1e-3
    ad = boltmi.derivative(fih, 0.046)
    assert ((fih(0.046 + D) - fih(0.046 - D)) / (2 * D) - ad) < 1e-5  # the reference's own criterion
    D = 1e-4
    rich = (4 * (fih(0.046 + D / 2) - fih(0.046 - D / 2)) / D - (fih(0.046 + D) - fih(0.046 - D)) / (2 * D)) / 3
    assert abs(ad - rich) <= 1e-5 * abs(rich)


def test_dual_background_carries_its_tangent():
    """a Dual background is a Dual to the host code and is marked for the
    device's tangent upload; a plain-struct conversion of Dual parameters is refused"""
    bg = boltmi.Background(boltmi.CosmoParams(Ω_b=boltmi.Dual(0.046, 1.0)), x_grid=(-20.0, 0.0, 201))
    assert isinstance(bg.eta0, boltmi.Dual) and isinstance(bg.H(-1.0), boltmi.Dual)
    assert bg.is_dual and not boltmi.Background(boltmi.CosmoParams(), x_grid=(-20.0, 0.0, 201)).is_dual
    with pytest.raises(TypeError):
        boltmi.CosmoParams(h=boltmi.Dual(0.7, 1.0)).c_struct()


def test_jvp_rejects_bad_arguments():
    par = boltmi.CosmoParams().c_struct()
    bad = boltmi.CosmoParams(h=float("nan")).c_struct()
    c = np.zeros((_abi.NTAB, 203))
    s = np.zeros(_abi.NSCAL)
    q = np.zeros(15)
    rc = lib().bolt_host_tables_jvp(ctypes.byref(par), ctypes.byref(bad), -20.0, 0.0, 201, 15, pd(c), pd(s),
                                    pd(c), pd(s), pd(q), pd(q))
    assert rc == _abi.BOLT_E_ARG
    rc = lib().bolt_host_tables_jvp(ctypes.byref(par), None, -20.0, 0.0, 201, 15, pd(c), pd(s), pd(c), pd(s),
                                    pd(q), pd(q))
    assert rc == _abi.BOLT_E_ARG


def test_jacobian_partials_are_bitwise_the_single_direction_jvps():
    """bolt_host_tables_jacobian carries all nine partials in one pass; each is
    bitwise the one-direction JVP seeded with that field's unit vector."""
    import time
    par = boltmi.CosmoParams(**MASSIVE)
    t0 = time.perf_counter()
    bg, dcoef, dscal = boltmi.tables_jacobian(par, x_grid=GRID)
    t_jac = time.perf_counter() - t0
    c0, s0 = tables(par)
    assert np.array_equal(bg.coef, c0) and np.array_equal(bg.scal, s0)
    t0 = time.perf_counter()
    for p, field in enumerate(boltmi.PARAM_FIELDS):
        _, _, dc, ds = jvp(par, {field: 1.0})
        assert np.array_equal(dcoef[p], dc), field
        assert np.array_equal(dscal[p], ds), field
    t_jvp = time.perf_counter() - t0
    print(f"jacobian {t_jac:.2f} s, nine JVPs {t_jvp:.2f} s")
