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


def _pinned_ctx():
    """a context whose solver runs six scan rows per lane, the shape of the
    queue server's launches, so that standalone results are comparable bitwise"""
    import os
    old = os.environ.get("BOLT_SOLVER_SR")
    os.environ["BOLT_SOLVER_SR"] = "6"
    try:
        return boltmi.Context(0)
    finally:
        if old is None:
            del os.environ["BOLT_SOLVER_SR"]
        else:
            os.environ["BOLT_SOLVER_SR"] = old


def test_queue_server_results_equal_standalone_and_join_a_running_launch():
    """bolt_serve_start_queue: every request's k-modes go through one persistent
    solver launch that the host feeds while it runs.  A large C_ℓ request is
    submitted first; small P(k) and C_ℓ requests submitted while it is being
    solved join that launch (their modes are published while it runs and it
    solves them).  Every result is bitwise the same computation run alone."""
    import time
    ks_cl = boltmi.quadratic_k(0.1 * 2.33e-4, 1000 * 2.33e-4, 30)
    ks_pk = np.logspace(-3, np.log10(0.3), 10)
    big_k = np.logspace(-4, np.log10(0.5), 800)
    plan = [("pk", "massive"), ("cltt", "nonu"), ("pk", "c1"), ("cltt", "c1")]
    srv = boltmi.Server(hbm_gib=2.0, queue=True)
    try:
        par, bg, ih = cosmo.background("default")
        big = srv.submit_cltt(par, k=big_k, ell_max=500, nkd=1000, ℓᵧ=50, ℓ_ν=50, ℓ_mν=10, reltol=1e-11)
        time.sleep(0.15)  # its tables are built and its modes published; the launch is running
        ids = []
        for kind, name in plan:
            p, b, i = cosmo.background(name)
            if kind == "cltt":
                ids.append(srv.submit_cltt(p, k=ks_cl, ell_max=300, nkd=600, ℓᵧ=8, ℓ_ν=8, ℓ_mν=10, reltol=1e-8))
            else:
                ids.append(srv.submit_pk(p, k=ks_pk))
        got = [srv.wait(i) for i in ids]
        got_big = srv.wait(big)
        qi = srv.queue_info()
    finally:
        srv.close()
    assert qi["served"] == len(plan) + 1
    assert 0 < qi["arena_peak_bytes"] <= qi["arena_bytes"]
    assert sum(info["joined_running"] for _, info in got) >= 1, [info for _, info in got]
    assert qi["launches"] < len(plan) + 1
    ctx = _pinned_ctx()
    try:
        for (kind, name), (vals, info) in zip(plan, got):
            par, bg, ih = cosmo.background(name)
            assert info["state"] == 2 and info["status"] == 0 and info["error"] == ""
            if kind == "cltt":
                sf = boltmi.source_grid(par, bg, ih, ks_cl, None, 8, 1e-8, ctx=ctx)
                ref = boltmi.cltt_on_grid(np.arange(2, 301), sf,
                                          boltmi.quadratic_k(0.01 * bg.H0, 1000 * bg.H0, 600), par, bg, ctx=ctx)
            else:
                ref = boltmi.plin(ks_pk, par, bg, ih, ctx=ctx)
            assert np.array_equal(vals, ref), (kind, name, np.max(np.abs(vals / ref - 1)))
        par, bg, ih = cosmo.background("default")
        sf = boltmi.source_grid(par, bg, ih, big_k, None, 50, 1e-11, ℓ_ν=50, ℓ_mν=10, ctx=ctx)
        ref = boltmi.cltt_on_grid(np.arange(2, 501), sf, boltmi.quadratic_k(0.01 * bg.H0, 1000 * bg.H0, 1000), par,
                                  bg, ctx=ctx)
        assert np.array_equal(got_big[0], ref)
    finally:
        ctx.close()


def test_queue_server_admits_by_footprint():
    """Requests whose memory together exceeds the arena wait for it and all
    complete; one that exceeds the arena alone is refused at submit."""
    probe = boltmi.Server(hbm_gib=1.0, queue=True)
    fixed = 2**30 - probe.queue_info()["arena_bytes"]   # the server's own work space
    probe.close()
    srv = boltmi.Server(hbm_gib=(fixed + 11.6e6) / 2**30, queue=True)   # an arena of ~11.6 MB
    try:
        par, bg, ih = cosmo.background("c1")
        ks = np.logspace(-3, np.log10(0.2), 300)   # 4.8 MB of source grid per request
        ids = [srv.submit_cltt(par, k=ks, ell_max=200, nkd=400, ℓᵧ=8, ℓ_ν=8, ℓ_mν=10, reltol=1e-6)
               for _ in range(6)]
        outs = [srv.wait(i)[0] for i in ids]
        qi = srv.queue_info()
        with pytest.raises(boltmi._abi.BoltError):
            srv.submit_cltt(par, k=np.logspace(-4, np.log10(0.5), 800), ell_max=2500, nkd=5000)
    finally:
        srv.close()
    assert qi["arena_peak_bytes"] <= qi["arena_bytes"] < 6 * 300 * 2001 * 8
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
