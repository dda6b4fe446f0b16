"""Forward-mode derivatives on the device (ForwardDiff through Bolt's pipeline:
README.md, runtests.jl:54-80, examples/plot_deriv_cl.jl:100-125).

The device takes the cosmology's tables and scalars together with their tangent
along a parameter direction (bolt_host_tables_jvp, pinned in test_host_ad.py).
Each device derivative is compared with central differences of the CPU oracle
evaluated at exactly linearised inputs: parameters p ± ε·dp, tables
coef ± ε·dcoef and scalars scal ± ε·dscal.  For the solve, the oracle is
replayed along the accepted steps of its own adaptive solve at p (the step
sequence the device also takes), so the difference quotient converges to the
derivative of the fixed-step discrete solution, which is what the tangent-linear
solve computes; its error is O(ε²) plus rounding over ε."""
import numpy as np
import pytest

import boltmi
import cosmo
import pyoracle
from boltmi import _abi
from boltmi.dual import Dual, value

pytestmark = pytest.mark.gpu

MNU = 0.2 * boltmi.MASS_NATURAL
BASE = dict(h=0.67, Ω_r=5.0469e-5, Ω_b=0.049, Ω_c=0.26, A=2.1e-9, n=0.965, Y_p=0.24, N_ν=3.046, Σm_ν=MNU)
# a direction through every parameter the perturbations see (relative steps of order one)
DIR = dict(h=0.67, Ω_b=0.049, Ω_c=-0.26, Σm_ν=MNU, N_ν=0.3)
EPS = 1e-5


def dual_cosmology(direction, base=BASE):
    par = boltmi.CosmoParams(**{k: Dual(v, direction.get(k, 0.0)) for k, v in base.items()})
    bg = boltmi.Background(par)
    return par, bg


def perturbed_oracle(bg, base, direction, eps):
    """the oracle at p + eps·dp with tables coef + eps·dcoef, scal + eps·dscal"""
    names = ["h", "Ω_r", "Ω_b", "Ω_c", "A", "n", "Y_p", "N_ν", "Σm_ν"]
    pv = [base[k] + eps * direction.get(k, 0.0) for k in names]
    sc = bg.scal + eps * bg.dscal
    tabs = dict(nx=bg.nx, nq=bg.nq, x0=bg.x0, x1=bg.x1, H0=sc[_abi.S_H0], eta0=sc[_abi.S_ETA0],
                rho_crit=sc[_abi.S_RHO_CRIT], xq=bg.quad_pts, wq=bg.quad_wts, coef=bg.coef + eps * bg.dcoef)
    return pyoracle.Oracle(pv, tabs)


@pytest.fixture(scope="module")
def dual_bg():
    par, bg = dual_cosmology(DIR)
    return par, bg, perturbed_oracle(bg, BASE, DIR, 0.0), perturbed_oracle(bg, BASE, DIR, EPS), \
        perturbed_oracle(bg, BASE, DIR, -EPS)


def test_rhs_tangent_matches_oracle_central_differences(dual_bg):
    par, bg, o0, op, om = dual_bg
    cfg = boltmi.hier_cfg(8, 8, 10, 15)
    oc = cosmo.ocfg(8, 8, 10, 15)
    n = boltmi.state_length(cfg)
    ks = np.array([1e-3, 0.05, 0.3])
    rng = np.random.default_rng(7)
    U = rng.standard_normal((n, len(ks)))
    Ud = rng.standard_normal((n, len(ks)))
    for x in (-12.0, -6.5, -1.0):   # x = -1 at k = 0.3: the RSA branch zeroes the photon/ν rows
        dU, _ = boltmi.rhs_batch(par, bg, cfg, ks, x, Dual(U, Ud))
        for b, k in enumerate(ks):
            ref = o0.rhs(oc, float(k), x, U[:, b])[0]
            np.testing.assert_allclose(dU.v[:, b], ref, rtol=0, atol=1e-12 * np.max(np.abs(ref)))
            fd = (op.rhs(oc, float(k), x, U[:, b] + EPS * Ud[:, b])[0] -
                  om.rhs(oc, float(k), x, U[:, b] - EPS * Ud[:, b])[0]) / (2 * EPS)
            err = np.max(np.abs(dU.d[:, b] - fd)) / np.max(np.abs(fd))
            assert err < 1e-7, (x, k, err)


def test_initial_conditions_tangent_matches_oracle(dual_bg):
    par, bg, o0, op, om = dual_bg
    cfg = boltmi.hier_cfg(8, 8, 10, 15)
    oc = cosmo.ocfg(8, 8, 10, 15)
    ks = np.array([1e-4, 0.01, 0.3])
    U = boltmi.initial_conditions_batch(par, bg, cfg, ks)
    for b, k in enumerate(ks):
        ref = o0.initial_conditions(oc, float(k))
        np.testing.assert_allclose(U.v[:, b], ref, rtol=1e-13, atol=1e-300)
        fd = (op.initial_conditions(oc, float(k)) - om.initial_conditions(oc, float(k))) / (2 * EPS)
        err = np.max(np.abs(U.d[:, b] - fd)) / np.max(np.abs(fd))
        assert err < 1e-7, (k, err)


def test_source_grid_tangent_matches_replayed_oracle(dual_bg):
    """dS/dp of source_grid (spectra.jl:6-23) along the direction, every x node of
    three k-modes (the largest with the RSA switch), against Richardson-combined
    central differences of the oracle replaying its own accepted steps.  Needs the
    device's step sequence to be the oracle's: the accepted-step counts are
    compared too."""
    par, bg, o0, op, om = dual_bg
    o2p, o2m = perturbed_oracle(bg, BASE, DIR, 2 * EPS), perturbed_oracle(bg, BASE, DIR, -2 * EPS)
    ih = boltmi.IonizationHistory(boltmi.RECFAST(bg), par, bg)
    ks = np.array([2e-3, 0.03, 0.2])
    oc = cosmo.ocfg(8, 8, 10, 15, 1e-8, 1e-6)
    sf = boltmi.source_grid(par, bg, ih, ks, None, 8, 1e-8)
    assert isinstance(sf.grid, Dual)
    accepted = 0
    for i, k in enumerate(ks):
        xs, st = o0.steps(oc, float(k))
        accepted += st.accepted
        ref = o0.source_column(oc, float(k))[0]
        np.testing.assert_allclose(sf.grid.v[i], ref, rtol=0, atol=1e-6 * np.max(np.abs(ref)))
        fd = {}
        for e, (o_p, o_m) in ((EPS, (op, om)), (2 * EPS, (o2p, o2m))):
            sp, stp = o_p.source_column_replay(oc, float(k), xs)
            sm, stm = o_m.source_column_replay(oc, float(k), xs)
            assert stp.status == 0 and stm.status == 0
            fd[e] = (sp - sm) / (2 * e)
        # Richardson: the ε² terms cancel; the two quotients' difference bounds
        # what is left (largest at the visibility peak, where the tables'
        # tangents are steep, and where rounding over ε grows)
        rich = (4.0 * fd[EPS] - fd[2 * EPS]) / 3.0
        scale = np.max(np.abs(rich))
        bar = max(1e-6, 4.0 * np.max(np.abs(fd[EPS] - fd[2 * EPS])) / scale)
        err = np.max(np.abs(sf.grid.d[i] - rich)) / scale
        assert bar < 1e-4 and err < bar, (k, err, bar)
    assert sf.stats["accepted"] == accepted
    # the primal of the tangent launch is the plain solve's
    par0 = boltmi.CosmoParams(**BASE)
    bg0 = boltmi.Background(par0)
    S0 = boltmi.source_grid(par0, bg0, ih, ks, None, 8, 1e-8).grid
    np.testing.assert_allclose(sf.grid.v, S0, rtol=0, atol=1e-12 * np.max(np.abs(S0)))


def test_tangent_is_linear_in_the_direction():
    """the tangent map is linear: d(a·u + b·v) = a·d(u) + b·d(v) to rounding; and a
    direction the perturbations cannot see (A_s, n_s enter only the primordial
    spectrum, spectra.jl:92) gives exactly zero"""
    ks = np.array([5e-3, 0.1])
    d1 = dict(h=0.67, Ω_c=0.26)
    d2 = dict(Ω_b=0.049, Σm_ν=MNU)
    d12 = dict(h=0.67 * 2.0, Ω_c=0.26 * 2.0, Ω_b=-0.049 * 3.0, Σm_ν=-MNU * 3.0)
    grids = []
    for d in (d1, d2, d12):
        par, bg = dual_cosmology(d)
        ih = boltmi.IonizationHistory(boltmi.RECFAST(bg), par, bg)
        grids.append(boltmi.source_grid(par, bg, ih, ks, None, 8, 1e-8).grid)
    lin = 2.0 * grids[0].d - 3.0 * grids[1].d
    assert np.max(np.abs(grids[2].d - lin)) <= 1e-9 * np.max(np.abs(lin))
    par, bg = dual_cosmology(dict(A=2.1e-9, n=0.965))
    ih = boltmi.IonizationHistory(boltmi.RECFAST(bg), par, bg)
    assert np.all(boltmi.source_grid(par, bg, ih, ks, None, 8, 1e-8).grid.d == 0.0)


def test_tangent_solve_rejects_the_fp32_mode():
    pard, bgd = dual_cosmology(DIR)
    ihd = boltmi.IonizationHistory(boltmi.RECFAST(bgd), pard, bgd)
    with pytest.raises(_abi.BoltError):
        boltmi.source_grid(pard, bgd, ihd, np.array([0.01]), None, 8, 1e-8, flags=boltmi.precision_flags("fp32"))


def test_cltt_tangent_matches_oracle_central_differences(dual_bg):
    """dC_ℓ/dp (cltt, spectra.jl:84-96, under Duals) from the device's S and dS/dp:
    through S, the Bessel argument k(η₀ − η(x)) and P(k).  The oracle's cltt at
    (p ± ε·dp, tables ± ε·dtables, S ± ε·dS) is quadratic in S and smooth in η, so
    its central differences, Richardson-combined, are exact to O(ε⁴).  Per ℓ the
    error is measured against max(|dC_ℓ|, C_ℓ) (the direction moves parameters by
    O(1) relative amounts)."""
    par, bg, o0, op, om = dual_bg
    ih = boltmi.IonizationHistory(boltmi.RECFAST(bg), par, bg)
    H0 = value(bg.H0)
    ks = boltmi.quadratic_k(0.1 * H0, 1000 * H0, 40)
    sf = boltmi.source_grid(par, bg, ih, ks, None, 8, 1e-8)
    dense = boltmi.quadratic_k(0.01 * bg.H0, 1000 * bg.H0, 600)
    ells = np.array([2, 10, 37, 100, 300, 700, 1200])
    cl = boltmi.cltt_on_grid(ells, sf, dense, par, bg)
    S, dS = sf.grid.v, sf.grid.d
    ref = o0.cltt(S, ks, dense, ells, 0, nthreads=8)
    np.testing.assert_allclose(cl.v, ref, rtol=1e-6)
    # along this direction dη/dp is O(η) ~ 1e4 Mpc, so ε moves the Bessel argument
    # k(η₀ − η) by ~ε·1e3 rad: a plain central difference at ε = 1e-5 carries
    # (0.03)²/6 ~ 1e-4 of second-order error.  Richardson over ε = 1e-6, 2e-6.
    fd = {}
    for e in (1e-6, 2e-6):
        o_p, o_m = perturbed_oracle(bg, BASE, DIR, e), perturbed_oracle(bg, BASE, DIR, -e)
        fd[e] = (o_p.cltt(S + e * dS, ks, dense, ells, 0, nthreads=8) -
                 o_m.cltt(S - e * dS, ks, dense, ells, 0, nthreads=8)) / (2 * e)
    rich = (4.0 * fd[1e-6] - fd[2e-6]) / 3.0
    scale = np.maximum(np.abs(rich), ref)
    assert np.max(np.abs(fd[1e-6] - fd[2e-6]) / scale) < 1e-4   # the quotients themselves agree
    err = np.abs(cl.d - rich) / scale
    assert np.max(err) < 1e-5, err


def test_cltt_tangent_along_the_primordial_parameters():
    """A_s and n_s enter only P(k) = A (k/0.05)^(n−1) (spectra.jl:92): along A_s the
    derivative is C_ℓ/A_s·dA; along n_s it is Σ T² P ln(k/0.05) Δk/k, checked
    against the oracle's central difference in n_s"""
    ks = np.logspace(-4, np.log10(0.3), 60)
    ells = np.array([2, 20, 200, 800])
    A, ns = 2.1e-9, 0.965
    for d in (dict(A=A), dict(n=1.0)):
        base = dict(BASE, A=A, n=ns)
        par, bg = dual_cosmology(d, base)
        ih = boltmi.IonizationHistory(boltmi.RECFAST(bg), par, bg)
        sf = boltmi.source_grid(par, bg, ih, ks, None, 8, 1e-8)
        assert np.all(sf.grid.d == 0.0)
        dense = boltmi.quadratic_k(0.01 * bg.H0, 1000 * bg.H0, 400)
        cl = boltmi.cltt_on_grid(ells, sf, dense, par, bg)
        if "A" in d:
            np.testing.assert_allclose(cl.d, cl.v, rtol=1e-13)
        else:
            e = 1e-4
            fd = (perturbed_oracle(bg, base, d, e).cltt(sf.grid.v, ks, dense, ells, 0, nthreads=8) -
                  perturbed_oracle(bg, base, d, -e).cltt(sf.grid.v, ks, dense, ells, 0, nthreads=8)) / (2 * e)
            np.testing.assert_allclose(cl.d, fd, rtol=1e-6)
