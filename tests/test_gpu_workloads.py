"""The batched-cosmology and serving workloads at their configured shapes.

Batch: 8 cosmologies drawn (seeded) from the parameter box h ∈ [0.6, 0.8],
Ω_b h² ∈ [0.020, 0.024], Ω_c h² ∈ [0.10, 0.14], n_s ∈ [0.92, 1.0],
ln 10¹⁰A_s ∈ [2.9, 3.2]; the odd ones with one massive ν (Σm_ν ∈ [0.06, 0.3]
eV, ℓ_mν = 20, n = 473), the even ones massless (ℓ_mν = 10, n = 323);
64 log k in [1e-4, 0.5] Mpc⁻¹ each, ℓᵧ = ℓ_ν = 50, reltol 1e-11: 512 modes of
two hierarchy sizes in one persistent launch.  Sampled modes are checked
against the CPU oracle (1e-6 of the column's largest value, equal step counts).

Serving: a seeded Poisson stream (λ = 50/s) of requests from the same box
(30 % massive), k grids of 50, 200 or 800 points, half P(k) and half C_ℓ^TT
to ℓ_max ∈ {500, 1500, 2500}; every result is bitwise the same request served
alone, and the C_ℓ ones also match the oracle summing the same source grid.
"""
import math
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import boltmi
import cosmo
import pyoracle

pytestmark = pytest.mark.gpu


def box_cosmo(rng, massive):
    h = rng.uniform(0.6, 0.8)
    return boltmi.CosmoParams(h=h, Ω_b=rng.uniform(0.020, 0.024) / h ** 2, Ω_c=rng.uniform(0.10, 0.14) / h ** 2,
                              n=rng.uniform(0.92, 1.0), A=math.exp(rng.uniform(2.9, 3.2)) * 1e-10, N_ν=3.046,
                              Σm_ν=(rng.uniform(0.06, 0.3) if massive else 0.0) * boltmi.MASS_NATURAL)


def test_batch_of_box_cosmologies_matches_oracle():
    rng = np.random.default_rng(2024)
    pars = [box_cosmo(rng, c % 2 == 1) for c in range(8)]
    lm = [20 if c % 2 == 1 else 10 for c in range(8)]
    cosmos = boltmi.backgrounds(pars, threads=8)
    ks = np.logspace(-4, math.log10(0.5), 64)
    print("# batch: 8 cosmologies x 64 k, one launch", flush=True)
    St, _ = boltmi.source_grid_batch(cosmos, ks, ℓᵧ=50, ℓ_ν=50, ℓ_mν=lm, reltol=1e-11)
    # two modes per cosmology (k ≈ 2.4e-4 and 0.034 Mpc⁻¹), against the oracle
    samples = [(c, i) for c in range(8) for i in (6, 50)]

    def oracle_col(ci):
        c, i = ci
        par, bg, _ = cosmos[c]
        o = pyoracle.Oracle(par.as_tuple(), bg.tables_dict())
        return o.source_column(pyoracle.Oracle.cfg(50, 50, lm[c], 15, 1e-11, 1e-6), float(ks[i]))

    with ThreadPoolExecutor(16) as ex:
        refs = list(ex.map(oracle_col, samples))
    for (c, i), (ref, ost) in zip(samples, refs):
        err = np.max(np.abs(St[c].grid[i] - ref)) / np.max(np.abs(ref))
        assert err < 1e-6, (c, i, err)
    # one cosmology of each hierarchy size alone: bitwise its rows of the batch,
    # and the oracle's step counts for the sampled modes
    for c in (0, 1):
        par, bg, ih = cosmos[c]
        one = boltmi.source_grid(par, bg, ih, ks, None, 50, 1e-11, ℓ_ν=50, ℓ_mν=lm[c])
        assert np.array_equal(one.grid, St[c].grid), c
        for i in (6, 50):
            single = boltmi.source_grid(par, bg, ih, ks[i:i + 1], None, 50, 1e-11, ℓ_ν=50, ℓ_mν=lm[c])
            ost = refs[samples.index((c, i))][1]
            assert (single.stats["accepted"], single.stats["rejected"]) == (ost.accepted, ost.rejected)


def test_poisson_request_stream_results_equal_alone():
    rng = np.random.default_rng(7)
    reqs = []
    for _ in range(10):
        massive = rng.uniform() < 0.3
        par = box_cosmo(rng, massive)
        nk = int(rng.choice([50, 200, 800]))
        lm = 20 if massive else 10
        if rng.uniform() < 0.5:
            reqs.append(("pk", par, dict(nk=nk), lm))
        else:
            reqs.append(("cltt", par, dict(nk=nk, ell_max=int(rng.choice([500, 1500, 2500])), ℓ_mν=lm), lm))
    gaps = rng.exponential(1.0 / 50.0, size=len(reqs))
    srv = boltmi.Server(slots=8, hbm_gib=8.0)

    def submit(kind, par, kw):
        return srv.submit_pk(par, **kw) if kind == "pk" else srv.submit_cltt(par, **kw)

    try:
        print("# serving: 10 requests at 50/s", flush=True)
        ids = []
        for (kind, par, kw, _), gap in zip(reqs, gaps):
            time.sleep(gap)
            ids.append(submit(kind, par, kw))
        mixed = [srv.wait(i) for i in ids]
        print("# serving: the same requests one at a time", flush=True)
        alone = [srv.wait(submit(kind, par, kw)) for kind, par, kw, _ in reqs]
    finally:
        srv.close()
    for (kind, par, kw, _), (vm, im), (va, ia) in zip(reqs, mixed, alone):
        assert im["status"] == 0 and ia["status"] == 0
        assert np.array_equal(vm, va), (kind, kw)
    # a C_ℓ request against the oracle on the same source grid
    j = next(i for i, r in enumerate(reqs) if r[0] == "cltt")
    kind, par, kw, lm = reqs[j]
    bg = boltmi.Background(par)
    ih = boltmi.IonizationHistory(boltmi.RECFAST(bg), par, bg)
    ks = np.logspace(-4, math.log10(0.5), kw["nk"])
    sf = boltmi.source_grid(par, bg, ih, ks, None, 50, 1e-11, ℓ_ν=50, ℓ_mν=lm)
    ells = np.arange(2, kw["ell_max"] + 1, dtype=np.int32)
    o = pyoracle.Oracle(par.as_tuple(), bg.tables_dict())
    ref = o.cltt(sf.grid, ks, boltmi.quadratic_k(0.01 * bg.H0, 1000 * bg.H0, 5000), ells, 0, nthreads=16)
    assert np.max(np.abs(mixed[j][0] / ref - 1)) < 1e-5
