"""The full CMB workload against the CPU FP64 oracle, at its own settings.

Cosmology: flat ΛCDM, h = 0.7, Ω_b = 0.046, Ω_c = 0.224, n_s = 0.96,
A_s = 2.1e-9, N_eff = 3.046, massless ν.  Hierarchy ℓᵧ = ℓ_ν = 50, ℓ_mν = 10,
nq = 15 (n = 323); 800 log-spaced k in [1e-4, 0.5] Mpc⁻¹; source_grid on the
2001-node x grid at reltol 1e-11, abstol 1e-6 (spectra.jl:6-23); C_ℓ^TT for
ℓ = 2..2500 on quadratic_k(0.01H₀, 1000H₀, 5000) (spectra.jl:84-96, 132-150).

Bars: source columns within 1e-6 of their largest magnitude with the same
accepted/rejected step counts as the oracle; C_ℓ^TT within 1e-5 relative for
every ℓ, the oracle summing the same GPU source grid with exact j_ℓ.
"""
import math
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import boltmi
import cosmo

pytestmark = pytest.mark.gpu

NK = 800
# the smallest k, its neighbour, interior modes and the 5371-step k = 0.5 mode
COLS = [0, 1, 100, 250, 400, 550, 700, 799]


@pytest.fixture(scope="module")
def c2():
    par, bg, ih = cosmo.background("c1")
    ks = np.logspace(-4, math.log10(0.5), NK)
    print(f"# C2 source_grid: {NK} k, l_g = l_nu = 50, reltol 1e-11", flush=True)
    sf = boltmi.source_grid(par, bg, ih, ks, boltmi.BasicNewtonian(), 50, 1e-11, ℓ_ν=50, ℓ_mν=10, nq=15,
                            abstol=1e-6)
    return par, bg, ih, ks, sf


def test_c2_source_columns_match_oracle(c2):
    par, bg, ih, ks, sf = c2
    o = cosmo.oracle("c1")
    oc = cosmo.ocfg(50, 50, 10, 15, 1e-11, 1e-6)
    with ThreadPoolExecutor(8) as ex:
        refs = list(ex.map(lambda i: o.source_column(oc, float(ks[i])), COLS))
    for i, (ref, ost) in zip(COLS, refs):
        err = np.max(np.abs(sf.grid[i] - ref)) / np.max(np.abs(ref))
        assert err < 1e-6, (i, ks[i], err)
        # the same mode run alone: bitwise the column of the 800-mode launch
        # (whatever ran before it on its workgroup), and the oracle's step counts
        one = boltmi.source_grid(par, bg, ih, ks[i:i + 1], None, 50, 1e-11, ℓ_ν=50, ℓ_mν=10, nq=15, abstol=1e-6)
        assert np.array_equal(one.grid[0], sf.grid[i]), i
        assert (one.stats["accepted"], one.stats["rejected"]) == (ost.accepted, ost.rejected), i
    assert refs[-1][1].accepted > 5000   # k = 0.5: the longest mode of the launch


def test_c2_two_spectra_per_launch_bitwise(c2):
    """bench.py's default shape: two spectra of the cosmology in one persistent
    launch (1600 modes, largest k first) give bitwise the single-spectrum grid"""
    par, bg, ih, ks, sf = c2
    St, _ = boltmi.source_grid_batch([(par, bg, ih)] * 2, ks, ℓᵧ=50, reltol=1e-11, ℓ_ν=50, ℓ_mν=10, nq=15,
                                     abstol=1e-6)
    for g in St:
        assert np.array_equal(g.grid, sf.grid)


def test_c2_cltt_matches_oracle_every_ell(c2):
    par, bg, ih, ks, sf = c2
    o = cosmo.oracle("c1")
    ells = np.arange(2, 2501, dtype=np.int32)
    cl = boltmi.cltt(ells, par, bg, ih, sf)
    dense = boltmi.quadratic_k(0.01 * bg.H0, 1000 * bg.H0, 5000)
    print("# oracle cltt: 2499 l x 4999 k", flush=True)
    ref = o.cltt(sf.grid, ks, dense, ells, 0, nthreads=16)
    rel = np.abs(cl / ref - 1)
    print(f"# C2 C_l^TT max rel err {rel.max():.3e} at l = {ells[np.argmax(rel)]}", flush=True)
    assert np.all(rel < 1e-5), (rel.max(), ells[np.argmax(rel)])
