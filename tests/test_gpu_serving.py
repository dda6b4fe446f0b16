"""The request server (bolt_serve_start / bolt_submit / bolt_poll / bolt_wait):
requests from several cosmologies, more than there are slots, in flight
together; every result equals the same computation run alone."""
import numpy as np
import pytest

import boltmi
import cosmo

pytestmark = pytest.mark.gpu


def test_served_results_equal_standalone():
    ks_cl = boltmi.quadratic_k(0.1 * 2.33e-4, 1000 * 2.33e-4, 30)
    ks_pk = np.logspace(-3, np.log10(0.3), 10)
    plan = [("cltt", "default"), ("pk", "massive"), ("cltt", "nonu"), ("pk", "c1"), ("cltt", "c1")]
    srv = boltmi.Server(slots=2, hbm_gib=2.0)
    try:
        assert srv.slots == 2
        ids = []
        for kind, name in plan:
            par, bg, ih = cosmo.background(name)
            if kind == "cltt":
                ids.append(srv.submit_cltt(par, k=ks_cl, ell_max=300, nkd=600, ℓᵧ=8, ℓ_ν=8, ℓ_mν=10,
                                           reltol=1e-8))
            else:
                ids.append(srv.submit_pk(par, k=ks_pk))
        got = [srv.wait(i) for i in ids]
    finally:
        srv.close()
    for (kind, name), (vals, info) in zip(plan, got):
        par, bg, ih = cosmo.background(name)
        assert info["state"] == 2 and info["status"] == 0
        assert info["latency_ms"] >= info["queue_ms"] >= 0
        if kind == "cltt":
            sf = boltmi.source_grid(par, bg, ih, ks_cl, boltmi.BasicNewtonian(), 8, 1e-8)
            ref = boltmi.cltt_on_grid(np.arange(2, 301), sf, boltmi.quadratic_k(0.01 * bg.H0, 1000 * bg.H0, 600),
                                      par, bg)
        else:
            ref = boltmi.plin(ks_pk, par, bg, ih)
        assert np.array_equal(vals, ref), (kind, name, np.max(np.abs(vals / ref - 1)))


def test_bad_request_is_rejected_before_queueing():
    srv = boltmi.Server(slots=1, hbm_gib=1.0)
    try:
        par, bg, ih = cosmo.background("default")
        with pytest.raises(boltmi._abi.BoltError):
            srv.submit_cltt(par, nk=1)
        with pytest.raises(boltmi._abi.BoltError):
            srv.submit_pk(par, k=[0.1, 0.05])
    finally:
        srv.close()
