"""Background, RECFAST and ionization tables built on the GPU
(bolt_device_tables_batch, one workgroup per cosmology) against the host's
bolt_host_tables_batch, which the Fortran RECFAST, CLASS and CAMB fixtures pin
(tests/test_host_tables.py, tests/test_oracle*.py).

The background tables (ℋ, η, ρ₀ℳ and their derivatives) are the same
quadratures in the same order: they agree to a few ulps of the table.  The
ionization tables follow RECFAST's adaptive Dormand–Prince steps, which the
device's transcendental functions move by an ulp here and there; they agree to
the integration's tolerance (rtol 1e-11 on X_e, T_mat), carried through τ, g̃
and their spline derivatives."""
import math

import numpy as np
import pytest

import boltmi
from boltmi import _abi

pytestmark = pytest.mark.gpu

# bars relative to each table's largest entry.  A spline-derivative table is
# built from differences of its source's coefficients over 2Δx (or Δx²), with
# Δx = 0.01, so an ulp of the source becomes ~100 (or ~1e4) ulps of it.
BARS = {_abi.T_HC: 1e-13, _abi.T_ETA: 1e-13, _abi.T_RHO0M: 1e-13,
        _abi.T_HC_P: 1e-11, _abi.T_ETA_P: 1e-11, _abi.T_HC_PP: 1e-9, _abi.T_ETA_PP: 1e-9,
        _abi.T_XE: 2e-6, _abi.T_TAU: 1e-7, _abi.T_G: 1e-7, _abi.T_TMAT: 1e-7, _abi.T_CSB2: 1e-6,
        _abi.T_TAU_P: 1e-6, _abi.T_G_P: 1e-6, _abi.T_TAU_PP: 1e-5, _abi.T_G_PP: 1e-5}


def box(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for c in range(n):
        h = rng.uniform(0.6, 0.8)
        out.append(boltmi.CosmoParams(h=h, Ω_b=rng.uniform(0.020, 0.024) / h ** 2,
                                      Ω_c=rng.uniform(0.10, 0.14) / h ** 2, n=rng.uniform(0.92, 1.0),
                                      A=math.exp(rng.uniform(2.9, 3.2)) * 1e-10, N_ν=3.046,
                                      Σm_ν=(rng.uniform(0.06, 0.3) if c % 2 else 0.0) * boltmi.MASS_NATURAL))
    return out


def tables(cosmos):
    return np.array([bg.coef for _, bg, _ in cosmos]), np.array([bg.scal for _, bg, _ in cosmos])


def test_device_tables_match_the_host():
    pars = [boltmi.CosmoParams(), boltmi.CosmoParams(Σm_ν=0.0)] + box(7, 10)
    ch, sh = tables(boltmi.backgrounds(pars))
    cd, sd = tables(boltmi.backgrounds(pars, device=True))
    errs = {}
    for t, bar in BARS.items():
        scale = np.max(np.abs(ch[:, t]), axis=1, keepdims=True)
        errs[t] = float(np.max(np.abs(cd[:, t] - ch[:, t]) / scale))
    print("device vs host tables, max error / table max:", {t: f"{e:.1e}" for t, e in errs.items()})
    d = np.abs(cd[:, _abi.T_XE] - ch[:, _abi.T_XE])
    c, i = np.unravel_index(np.argmax(d), d.shape)
    x = -20.0 + 0.01 * (i - 1)
    print(f"largest X_e difference: cosmology {c}, coefficient {i} (x = {x:.2f}, z = {np.exp(-x) - 1:.1f}): "
          f"{ch[c, _abi.T_XE, i]:.9e} vs {cd[c, _abi.T_XE, i]:.9e}")
    for t, bar in BARS.items():
        assert errs[t] < bar, (t, errs[t], bar)
    assert np.array_equal(sd[:, _abi.S_H0], sh[:, _abi.S_H0])
    assert np.array_equal(sd[:, _abi.S_RHO_CRIT], sh[:, _abi.S_RHO_CRIT])
    assert np.array_equal(sd[:, _abi.S_OMEGA_L], sh[:, _abi.S_OMEGA_L])
    assert np.max(np.abs(sd[:, _abi.S_ETA0] / sh[:, _abi.S_ETA0] - 1)) < 1e-14
    assert np.array_equal(sd[:, _abi.S_TAU0], sh[:, _abi.S_TAU0])


def test_plin_through_device_tables():
    """P(k) (plin's defaults) from device-built tables against host-built ones:
    equal up to the two RECFAST solutions' difference, carried through a solve
    at plin's reltol 1e-5"""
    pars = [boltmi.CosmoParams(), box(3, 2)[1]]
    ks = np.array([1e-3, 0.01, 0.05, 0.2])
    for (p, bgh, ihh), (_, bgd, ihd) in zip(boltmi.backgrounds(pars), boltmi.backgrounds(pars, device=True)):
        a = boltmi.plin(ks, p, bgh, ihh)
        b = boltmi.plin(ks, p, bgd, ihd)
        assert np.max(np.abs(b / a - 1)) < 1e-5


def test_bad_arguments_are_refused():
    with pytest.raises(_abi.BoltError):
        boltmi.backgrounds([boltmi.CosmoParams(h=-0.7)], device=True)
    assert boltmi.backgrounds([], device=True) == []
