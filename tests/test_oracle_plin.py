"""Pins for the oracle's plin and θ (spectra.jl:163-198, perturbations.jl:148-158).

Three independent checks, each chosen so that a plausible mistake in the
restatement fails it (the `plant` variants of the oracle put such mistakes in
on purpose and the tests assert that they are caught):

* the reference's own CAMB fixture test/data/camb_pk_x0.dat (linear P(k) at
  z = 0 for CosmoParams()'s cosmology, k in h/Mpc, P in (Mpc/h)³; written by
  test/camb_v_class_2022.ipynb, cells 3 and 5);
* the Poisson constraint of general relativity at x = 0:
  k²Φ = (3/2) H₀² Ω_m δ_m with δ_m the comoving-gauge matter density that
  plin squares, from the same solve's Φ (radiation contributes 1e-4 today);
* closed forms: Ω_ν h² = Σm_ν / 93.14 eV for plin's νfac, and the
  massless-limit moments θ = ρ = 4·(7/8)(π²/15)T_ν⁴ for ℳ_ℓ,ᵢ = −dlnf₀/dlnq(qᵢ).
"""
import math

import numpy as np
import pytest

import boltmi
import cosmo

# plant bits of orc_plin_planted
FLIP_GAUGE, NUFAC, NORM, THETA = 1, 2, 4, 8


def _camb():
    from conftest import GOLDEN
    import os
    return np.load(os.path.join(GOLDEN, "camb_pk.npz"), allow_pickle=False)


# every 100th node of the 1000-point fixture up to k = 0.18 h/Mpc
CAMB_IDX = [0, 100, 200, 300, 400, 500, 600, 650]


def _camb_errs(plant=0):
    par, bg, ih = cosmo.background("default")
    o = cosmo.oracle("default")
    d = _camb()
    kh, Ph = d["k_h"], d["P_h"]
    # ℓᵧ = ℓ_ν = ℓ_mν = 10 instead of plin's 50/50/20: at z = 0 the matter
    # density does not feel the photon and neutrino cut-offs at this level
    # (measured: 20/10 and 10/20 move P by < 1e-3), and n = 203 keeps the
    # dense-LU oracle at about a second per k
    cfg = cosmo.ocfg(10, 10, 10, 15, 1e-5, 1e-6)
    h = par.h
    errs = []
    for i in CAMB_IDX:
        P, st = o.plin(cfg, kh[i] * h, plant=plant)
        assert st.status == 0
        errs.append(P * h ** 3 / Ph[i] - 1.0)
    return np.array(errs)


def test_plin_matches_camb_fixture():
    """Bolt's P(k) vs CAMB (AccuracyBoost 3) for k = 1e-4..0.18 h/Mpc.  The
    oracle agrees to 3.4e-3 at the largest scale and ≤ 1.3e-3 elsewhere; the
    bound 5e-3 leaves room for CAMB's own accuracy, while the planted
    mistakes move P by factors (gauge sign at super-horizon k, the 2π²
    normalisation)."""
    errs = _camb_errs()
    assert np.max(np.abs(errs)) < 5e-3, errs
    for plant in (FLIP_GAUGE, NORM):
        assert np.max(np.abs(_camb_errs(plant))) > 5e-2, plant


def _poisson_ratio(k, plant=0):
    par, bg, ih = cosmo.background("default")
    o = cosmo.oracle("default")
    cfg = cosmo.ocfg(10, 10, 10, 15, 1e-8, 1e-8)
    n = o.n(cfg)
    P, st = o.plin(cfg, k, plant=plant)
    out, _ = o.boltsolve(cfg, k, np.array([0.0]))
    Phi = out[0][n - 5]   # u[iPhi] (perturbations.jl:123)
    dm = math.sqrt(P * k ** 3 / (2 * math.pi ** 2 * par.A * (k / 0.05) ** (par.n - 1)))
    Om = par.Ω_c + par.Ω_b + (par.Σm_ν / boltmi.MASS_NATURAL) / 93.14 / par.h ** 2
    return dm / (k * k * abs(Phi) / (1.5 * bg.H0 ** 2 * Om))


@pytest.mark.parametrize("k,tol", [(2e-5, 2e-3), (1e-4, 2e-3), (1e-3, 1e-4), (1e-2, 1e-4), (0.05, 1e-4)])
def test_plin_obeys_the_poisson_constraint(k, tol):
    """|δ_m| from plin equals k²|Φ|/((3/2)H₀²Ω_m).  Measured: 1.1e-3 off at
    k ≤ 1e-4 Mpc⁻¹ (k/ℋ₀ < 0.5: the initial conditions' O(kη) terms and the
    radiation share), ≤ 4e-5 from k = 1e-3 on.  Far outside the horizon the
    Newtonian δ is O(Φ) while the comoving one is O((k/ℋ)²Φ), so a wrong sign
    or factor in 3ℋv/k, or a wrong θ, breaks the constraint."""
    r = _poisson_ratio(k)
    assert abs(r - 1.0) < tol, r


def test_poisson_pin_catches_planted_mistakes():
    for plant in (FLIP_GAUGE, THETA):
        r = _poisson_ratio(1e-4, plant)
        assert abs(r - 1.0) > 1e-2, (plant, r)


def test_omega_nu_closed_form():
    """Ω_ν h² = Σm_ν / 93.14 eV (N_eff = 3.046, T_CMB = 2.7255 K); measured
    1.5e-3, the precision of the 93.14 eV figure for Bolt's Ω_r and ρ_crit"""
    par, bg, ih = cosmo.background("default")
    o = cosmo.oracle("default")
    m_eV = par.Σm_ν / boltmi.MASS_NATURAL
    assert o.omega_nu() * par.h ** 2 * 93.14 / m_eV == pytest.approx(1.0, abs=3e-3)
    assert abs(o.omega_nu(NUFAC) * par.h ** 2 * 93.14 / m_eV - 1.0) > 5e-2


def test_theta_and_rho_massless_closed_form():
    """For ℳ_ℓ,ᵢ = −dlnf₀/dlnq(qᵢ) in the massless limit the q-integrals
    4π∫q³f₀(−dlnf₀/dlnq)dq of θ (perturbations.jl:148-158) and ρ
    (perturbations.jl:127-145) both equal 4·(7/8)(π²/15)T_ν⁴ (integration by
    parts); the 15-point rule on [T_ν/30, 30T_ν] gets within 8.2e-5."""
    par, bg, ih = cosmo.background("default")
    o = cosmo.oracle("default")
    tb = bg.tables_dict()
    Tnu = (par.N_ν / 3) ** 0.25 * (4 / 11) ** (1 / 3) * (15 / np.pi ** 2 * bg.rho_crit * par.Ω_r) ** 0.25
    lo, hi = math.log10(Tnu / 30), math.log10(Tnu * 30)
    q = 10 ** (lo + (hi - lo) / 2 * (np.asarray(tb["xq"]) + 1))
    dl = -(q / Tnu) / (1 + np.exp(-q / Tnu))
    want = 4 * 7 / 8 * np.pi ** 2 / 15 * Tnu ** 4
    assert o.theta(-dl) == pytest.approx(want, rel=2e-4)
    rho, sig = o.rho_sigma(-dl, -dl, a=1e-12)
    assert rho == pytest.approx(want, rel=2e-4)
    assert sig == pytest.approx(want, rel=2e-4)
    # θ carries q³ and no ε: it does not change with the neutrino mass
    assert o.theta(-dl, a=1.0) == o.theta(-dl, a=1e-12)


def test_oracle_rejects_out_of_range_momentum_bins():
    """the oracle's per-bin scratch arrays hold 60 bins: larger nq reports
    status 4 / NaN instead of writing past them"""
    o = cosmo.oracle("default")
    cfg = cosmo.ocfg(8, 8, 4, 61)
    n = o.n(cfg)
    du, _ = o.rhs(cfg, 0.01, -5.0, np.ones(n))
    assert np.all(np.isnan(du))
    P, st = o.plin(cfg, 0.01)
    assert math.isnan(P) and st.status == 4
    assert math.isnan(o.theta(np.ones(61)))
