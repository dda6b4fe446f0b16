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


def test_concurrent_contexts_match_one_context():
    """Calls on different contexts run concurrently (each context has its own
    stream and cached launch buffers); every result is bitwise what the same
    call gives alone, on the first call of a context and on later ones."""
    import threading

    par, bg, ih = cosmo.background("default")
    ks = np.logspace(-3, np.log10(0.2), 24)
    cfg = dict(ℓᵧ=8, reltol=1e-8)
    ref = boltmi.source_grid(par, bg, ih, ks, **cfg).grid
    ref_pk = boltmi.plin(ks[:6], par, bg, ih, ℓᵧ=12, ℓ_ν=12, ℓ_mν=8)
    ctxs = [boltmi.Context(0) for _ in range(3)]
    out = [[] for _ in ctxs]
    errs = []

    def work(j):
        try:
            for _ in range(2):
                out[j].append(boltmi.source_grid(par, bg, ih, ks, ctx=ctxs[j], **cfg).grid)
                out[j].append(boltmi.plin(ks[:6], par, bg, ih, ℓᵧ=12, ℓ_ν=12, ℓ_mν=8, ctx=ctxs[j]))
        except Exception as e:
            errs.append(e)

    th = [threading.Thread(target=work, args=(j,)) for j in range(len(ctxs))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for c in ctxs:
        c.close()
    assert not errs, errs
    for res in out:
        assert len(res) == 4
        for i in (0, 2):
            assert np.array_equal(res[i], ref)
        for i in (1, 3):
            assert np.array_equal(res[i], ref_pk)
