"""Bolt's own j_ℓ discretisation on the device (bessel_mode = 1): the natural
cubic spline of j_ℓ on 5001 nodes over [0, k_max η₀] (bessel_interpolator,
spectra.jl:49-58), evaluated at every (ℓ, k, x) term of the line-of-sight sum
(spectra.jl:70-96).  Checked against the oracle's spline mode on the same
source grids, against the exact-j_ℓ mode (they differ by the spline's
discretisation error), and through the reference's own test: runtests.jl:149-185
compare cltt/clee from source_grid(ℓᵧ = 8, 100 k) with CAMB at 1.1e-1."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import boltmi
import cosmo

pytestmark = pytest.mark.gpu


def test_spline_mode_matches_oracle_spline_mode():
    par, bg, ih = cosmo.background("default")
    o = cosmo.oracle("default")
    k_src = boltmi.quadratic_k(0.1 * bg.H0, 1000 * bg.H0, 40)
    cfg = cosmo.ocfg(8, 8, 10, 15, 1e-8, 1e-6)
    cols = [o.source_column_tp(cfg, float(k)) for k in k_src]
    S = np.array([c[0] for c in cols])
    SP = np.array([c[1] for c in cols])
    dense = boltmi.quadratic_k(0.01 * bg.H0, 1000 * bg.H0, 600)
    ells = np.array([2, 3, 10, 37, 100, 250, 500, 777, 1200, 37])
    sf_t = boltmi.SourceGrid(bg.x_grid, k_src, S)
    sf_p = boltmi.SourceGrid(bg.x_grid, k_src, SP)
    cl1 = boltmi.cltt_on_grid(ells, sf_t, dense, par, bg, bessel_mode=1)
    ref1 = o.cltt(S, k_src, dense, ells, 1, nthreads=8)
    np.testing.assert_allclose(cl1, ref1, rtol=1e-9)
    assert cl1[3] == cl1[-1]
    # the exact-j_ℓ sum differs by the spline's error (~1e-4 .. 1e-2 of j_ℓ's scale)
    cl0 = boltmi.cltt_on_grid(ells, sf_t, dense, par, bg, bessel_mode=0)
    assert np.max(np.abs(cl1 / cl0 - 1)) > 1e-5
    tt, te, ee = boltmi.cl_on_grid(ells, sf_t, sf_p, dense, par, bg, bessel_mode=1)
    rtt, rte, ree = o.cl(S, SP, k_src, dense, ells, 1, nthreads=8)
    np.testing.assert_allclose(tt, rtt, rtol=1e-9)
    np.testing.assert_allclose(ee, ree, rtol=1e-9)
    np.testing.assert_allclose(te, rte, rtol=1e-9, atol=1e-9 * np.max(np.abs(rte)))


def test_spline_mode_runtests_camb_comparison(golden):
    """runtests.jl:149-185 on the device: 100 k, ℓᵧ = 8, reltol 1e-11, cltt and
    clee at ℓ = 10:10:2500 with Bolt's spline j_ℓ, against CAMB unlensed D_ℓ at
    1.1e-1 relative."""
    par, bg, ih = cosmo.background("default")
    k_grid = boltmi.quadratic_k(0.1 * bg.H0, 1000 * bg.H0, 100)
    sf_t, sf_p = boltmi.source_grids_TP(par, bg, ih, k_grid, boltmi.BasicNewtonian(), 8, 1e-11)
    ells = np.arange(10, 2501, 10)
    dense = boltmi.quadratic_k(0.01 * bg.H0, 1000 * bg.H0, 5000)
    tt, te, ee = boltmi.cl_on_grid(ells, sf_t, sf_p, dense, par, bg, bessel_mode=1)
    d = golden("camb_cl.npz")
    fac = ells * (ells + 1) / (2 * np.pi)
    assert np.max(np.abs(fac * tt / np.interp(ells, d["ell"], d["tt"]) - 1)) < 1.1e-1
    assert np.max(np.abs(fac * ee / np.interp(ells, d["ell"], d["ee"]) - 1)) < 1.1e-1
