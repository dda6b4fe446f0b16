"""Every shape the autotuner chooses between gives the oracle's results: the
solver with 5, 6, 7 and 8 scan rows per lane (pinned per context through
BOLT_SOLVER_SR) at n = 323 and n = 473, and the LOS kernel with 1, 2 and 4
x slots per wavefront (BOLT_LOS_SPW); bolt_autotune itself returns a choice
among them and later calls use it."""
import math
import os

import numpy as np
import pytest

import boltmi
import cosmo

pytestmark = pytest.mark.gpu


def _ctx(var, val):
    old = os.environ.get(var)
    os.environ[var] = str(val)
    try:
        return boltmi.Context(0)
    finally:
        if old is None:
            del os.environ[var]
        else:
            os.environ[var] = old


@pytest.mark.parametrize("lm", [10, 20])
def test_solver_shapes_match_oracle(lm):
    par, bg, ih = cosmo.background("massive")
    o = cosmo.oracle("massive")
    # source_grid's own tolerance (spectra.jl:6): the shapes' sums differ in
    # roundoff, which can move a step decision, and at reltol 1e-11 that stays
    # far below the 1e-6 bar
    ks = np.array([0.003, 0.05, 0.15])
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(3) as ex:
        refs = list(ex.map(lambda k: o.source_column(cosmo.ocfg(50, 50, lm, 15, 1e-11, 1e-6), float(k)), ks))
    grids = []
    for sr in (6, 7, 8):  # five rows per lane is not launched (see SR_RUN_MIN)
        ctx = _ctx("BOLT_SOLVER_SR", sr)
        sf = boltmi.source_grid(par, bg, ih, ks, None, 50, 1e-11, ℓ_ν=50, ℓ_mν=lm, ctx=ctx)
        ctx.close()
        for i, (ref, ost) in enumerate(refs):
            err = np.max(np.abs(sf.grid[i] - ref)) / np.max(np.abs(ref))
            assert err < 1e-6, (sr, ks[i], err)
        grids.append(sf.grid)
    for g in grids[1:]:
        assert np.max(np.abs(g - grids[0])) <= 1e-7 * np.max(np.abs(grids[0]))


def test_los_shapes_match_oracle():
    par, bg, ih = cosmo.background("default")
    o = cosmo.oracle("default")
    k_src = boltmi.quadratic_k(0.1 * bg.H0, 1000 * bg.H0, 40)
    S, status = o.source_grid(cosmo.ocfg(8, 8, 10, 15, 1e-8, 1e-6), k_src, nthreads=8)
    sf = boltmi.SourceGrid(bg.x_grid, k_src, S)
    dense = boltmi.quadratic_k(0.01 * bg.H0, 1000 * bg.H0, 600)
    ells = np.array([2, 3, 10, 37, 100, 250, 500, 777, 1200])
    ref = o.cltt(S, k_src, dense, ells, 0, nthreads=8)
    for spw in (1, 2, 4):
        ctx = _ctx("BOLT_LOS_SPW", spw)
        cl = boltmi.cltt_on_grid(ells, sf, dense, par, bg, ctx=ctx)
        cl2 = boltmi.cltt_on_grid(ells, sf, dense, par, bg, ctx=ctx)
        ctx.close()
        np.testing.assert_allclose(cl, ref, rtol=1e-6)
        assert np.array_equal(cl, cl2)


def test_autotune_picks_a_shape_and_later_calls_use_it():
    par, bg, ih = cosmo.background("default")
    ks = np.logspace(-4, math.log10(0.2), 64)
    cfg = boltmi.hier_cfg(8, 8, 10, 15, 1e-8, 1e-6)
    dense = boltmi.quadratic_k(0.01 * bg.H0, 1000 * bg.H0, 800)
    rep = boltmi.autotune(par, bg, cfg, ks, dense, 600)
    assert rep["solver_sr"] in (6, 7, 8) and rep["los_spw"] in (1, 2, 4)
    assert all(t > 0 for t in rep["solver_ms"].values()) and all(t > 0 for t in rep["los_ms"].values())
    assert rep["solver_ms"][rep["solver_sr"]] == min(rep["solver_ms"].values())
    # the tuned shapes give the pinned shapes' results bit for bit
    sf = boltmi.source_grid(par, bg, ih, ks, None, 8, 1e-8)
    ctx = _ctx("BOLT_SOLVER_SR", rep["solver_sr"])
    pinned = boltmi.source_grid(par, bg, ih, ks, None, 8, 1e-8, ctx=ctx)
    ctx.close()
    assert np.array_equal(sf.grid, pinned.grid)
    ells = np.arange(2, 601)
    cl = boltmi.cltt_on_grid(ells, sf, dense, par, bg)
    ctx = _ctx("BOLT_LOS_SPW", rep["los_spw"])
    cl_p = boltmi.cltt_on_grid(ells, sf, dense, par, bg, ctx=ctx)
    ctx.close()
    assert np.array_equal(cl, cl_p)
