#!/usr/bin/env python3
"""Benchmark of the full C_ℓ^TT path on one MI355X.

One step = source_grid for 800 log-spaced k in [1e-4, 0.5] Mpc⁻¹ (ℓᵧ = ℓ_ν = 50,
ℓ_mν = 10, nq = 15, reltol 1e-11, abstol 1e-6, 2001 x nodes) followed by
cltt for ℓ = 2..2500 on quadratic_k(0.01H₀, 1000H₀, 5000), flat ΛCDM with
h = 0.7, Ω_b = 0.046, Ω_c = 0.224, n_s = 0.96, A_s = 2.1e-9, N_eff = 3.046,
massless ν.  Tables, k grids and the source grid stay resident in HBM.

Steps are independent spectra.  --per-launch B (default 2) solves B spectra in
one persistent solver launch (their 800·B modes share the work queue, largest k
first), followed by B LOS passes.  With --inflight N (default 2) such batches
run on N contexts, each with its own HIP stream, result buffers and host
thread, so a batch's solve can use the CUs that the previous one's longest
modes leave idle; value counts completed spectra per second of wall time, and
latency_ms gives the time of one batch as its caller sees it (every spectrum of
a batch completes with it).  --per-launch 1 --inflight 1 runs the spectra one
after another.

--batch measures the many-cosmology workload instead (k-mode solves/s of one
persistent launch over 64 cosmologies × 800 k), and --serve a Poisson request
stream; each prints its own JSON line.

Prints one JSON line (see the "roofline" and "cpu_baseline" objects).
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bolt.jl_amd"))

import numpy as np

METRIC = "k-mode solves/s + C_ℓ spectra/s vs plain FP64 GPU at 1e-5 rel err; p99 latency"
PEAK_FP64 = 78.6  # TFLOP/s, MI355X FP64 (vector and matrix)
PEAK_HBM = 8000.0  # GB/s

NK, LG, LN, LM, NQ = 800, 50, 50, 10, 15
RELTOL, ABSTOL = 1e-11, 1e-6
LMAX = 2500
NKD = 5000


def workload():
    import boltmi
    par = boltmi.CosmoParams(h=0.7, Ω_b=0.046, Ω_c=0.224, n=0.96, A=2.1e-9, N_ν=3.046, Σm_ν=0.0)
    bg = boltmi.Background(par)
    ih = boltmi.IonizationHistory(boltmi.RECFAST(bg), par, bg)
    ks = np.logspace(-4, math.log10(0.5), NK)
    cfg = boltmi.hier_cfg(LG, LN, LM, NQ, RELTOL, ABSTOL)
    ells = np.arange(2, LMAX + 1, dtype=np.int32)
    dense = boltmi.quadratic_k(0.01 * bg.H0, 1000 * bg.H0, NKD)
    return par, bg, ih, ks, cfg, ells, dense


def measured_traffic(kernel, per_launch=1):
    """HBM bytes per launch of `kernel` (solving per_launch spectra) from the
    committed rocprofv3 summary (FETCH_SIZE and WRITE_SIZE, scaled by a plain
    copy of known size), or None where that shape was not measured."""
    import re
    path = os.path.join(ROOT, "profiles", "k_boltsolve_cmb_full.md")
    if not os.path.exists(path):
        return None
    key = kernel if per_launch == 1 else f"{kernel}_x{per_launch}"
    m = re.search(r"traffic_bytes_per_launch\s+" + re.escape(key) + r"\s+([0-9.eE+-]+)", open(path).read())
    return float(m.group(1)) if m else None


def ode_flops(n, nq, steps, accepted, nmodes, nodes):
    """Algorithmic FP64 flops of the integrator as implemented (per step: five
    stage RHS + banded factor/solve + border, the smoothed error solve; per
    accepted step the FSAL RHS; per node the dense-output + RHS + source)."""
    rhs = 16 * n + 4 * nq + 40
    stage = rhs + 6 * n + 44 * n + 4 * (4 * nq + 40) + 6 * n
    per_step = 30 * n + 24 * n + 5 * stage + 9 * n + 6 * n + 4 * n
    return steps * per_step + accepted * rhs + nmodes * nodes * (8 * n + rhs + 120)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(par, bg, ks, ells, dense, threads, S_gpu, cl_gpu):
    """The CPU FP64 oracle beside the GPU, on the same inputs, and the GPU's
    error against it.

    Source grid: the oracle solves a bounded sample of the modes (32 of 800,
    spread over k), each on one thread of a `threads`-wide pool; per-mode times
    are summed and scaled to all modes.  The sampled columns are compared with
    the GPU's (largest |difference| / largest |value| per column).
    C_ℓ^TT: the oracle sums the GPU's whole source grid with exact j_ℓ for all
    2499 ℓ and 4999 dense k on `threads` threads, timed in full, and is compared
    with the GPU's C_ℓ ℓ by ℓ.  A 1-core figure comes from single-thread timings
    (the per-mode solves, and the LOS of 50 dense-k midpoints on one thread)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from concurrent.futures import ThreadPoolExecutor
    o = pyoracle.Oracle(par.as_tuple(), bg.tables_dict())
    cfg = pyoracle.Oracle.cfg(LG, LN, LM, NQ, RELTOL, ABSTOL, 0, 1)
    mode_idx = list(range(12, NK, 25))           # 32 of 800 modes, spread over k
    mode_idx[-1] = NK - 1                        # and the longest, k = 0.5
    cols = {}

    def one_mode(i):
        t0 = time.perf_counter()
        S, st = o.source_column(cfg, float(ks[i]))
        cols[i] = S
        return time.perf_counter() - t0

    with ThreadPoolExecutor(threads) as ex:
        t_modes = list(ex.map(one_mode, mode_idx))
    err_src = max(float(np.max(np.abs(S_gpu[i] - cols[i])) / np.max(np.abs(cols[i]))) for i in mode_idx)
    t0 = time.perf_counter()
    cl_ref = o.cltt(S_gpu, ks, dense, ells, 0, nthreads=threads)
    t_cl = time.perf_counter() - t0
    err_cl = float(np.max(np.abs(cl_gpu / cl_ref - 1)))
    kmid_idx = list(range(50, NKD - 1, 100))     # 50 of 4999 midpoints, one thread
    t0 = time.perf_counter()
    for ik in kmid_idx:
        o.cltt(S_gpu, ks, dense, ells, 0, nthreads=1, kbeg=ik, kend=ik + 1)
    t_cl1 = (time.perf_counter() - t0) * (NKD - 1) / len(kmid_idx)
    t_sg1 = sum(t_modes) * (NK / len(mode_idx))
    t_sg = t_sg1 / threads
    return {"value": 1.0 / (t_sg + t_cl), "unit": "spectra/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "one_core_value": 1.0 / (t_sg1 + t_cl1),
            "sample": f"oracle source_grid on {len(mode_idx)} of {NK} k-modes (one thread each, {threads} at a "
                      f"time; per-mode times summed and scaled to {NK}) + oracle cltt on the full GPU source grid "
                      f"(2499 l x 4999 k, {threads} threads, timed in full); one_core_value from single-thread "
                      f"timings (the sampled solves and 50 dense-k midpoints of the LOS)",
            "est_source_grid_s": t_sg, "cltt_s": t_cl, "one_core_source_grid_s": t_sg1, "one_core_cltt_s": t_cl1,
            "sample_cpu_s": sum(t_modes) + t_cl * threads,
            "gpu_max_err": {"source_columns_scaled": err_src, "cltt_rel": err_cl,
                            "columns_compared": len(mode_idx), "ells_compared": len(ells)}}


def plain_gpu_baseline(lane, L, cfg, d_k, value):
    """The plain one-mode-per-launch FP64 GPU form of the same spectrum: this
    build's own FP64 solver launched for one k-mode at a time (no batching
    across modes), then one LOS pass.  A sample of 40 of the 800 modes, spread
    over k, is launched one by one and timed by wall clock; the sum is scaled
    to 800 launches."""
    from boltmi import _abi
    import torch
    idx = list(range(10, NK, NK // 40))[:40]
    t_modes = 0.0
    for i in idx:
        ms = ctypes.c_float()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _abi.check(L.bolt_source_grid_dev(lane.ctx.handle, lane.cid, ctypes.byref(cfg), d_k[i:].data_ptr(), 1,
                                          lane.d_S[i:].data_ptr(), lane.d_status[i:].data_ptr(),
                                          lane.d_ctr[i:].data_ptr(), ctypes.byref(ms)))
        t_modes += time.perf_counter() - t0
    t_sg = t_modes * NK / len(idx)
    ms_los = ctypes.c_float()
    t0 = time.perf_counter()
    _abi.check(L.bolt_cltt_dev(lane.ctx.handle, lane.cid, lane.d_S.data_ptr(), d_k.data_ptr(), NK,
                               lane.d_dense.data_ptr(), NKD, _abi.pi(lane.ells), len(lane.ells),
                               lane.d_cl.data_ptr(), lane.d_steps.data_ptr(), ctypes.byref(ms_los)))
    t_los = time.perf_counter() - t0
    plain = 1.0 / (t_sg + t_los)
    if int(lane.d_status.max().item()) != 0:
        raise RuntimeError("a k-mode failed in the plain baseline")
    return {"value": plain, "unit": "spectra/s", "ratio": value / plain,
            "sample": f"{len(idx)} of {NK} k-modes launched one per launch (wall time, summed and scaled to {NK}), "
                      f"plus one LOS launch", "est_source_grid_s": t_sg, "los_s": t_los}


def c1(args):
    """The plain-method workload, measured apart from the default line.
    Cosmology: flat ΛCDM (h = 0.7, Ω_b = 0.046, Ω_c = 0.224, n_s = 0.96,
    A_s = 2.1e-9, N_eff = 3.046, massless ν).
    Hierarchy: ℓᵧ = ℓ_ν = ℓ_mν = 10, nq = 15 (n = 203).
    k: 100 log-spaced in [1e-4, 1] h/Mpc, solved at reltol = abstol = 1e-6.
    Output: states at 1000 x from -20 to 0 (δ_c and Φ among them), then P(k)
    for the same 100 k (plin, reltol 1e-5).
    One step is both; value is k-mode solves/s of wall time, states copied to
    the host included.  With --inflight N (default 3) the steps run on N contexts
    (each its own HIP stream and host thread), as the default line does.  Beside it: the
    oracle on a sample of the modes (its time and the largest scaled difference
    of δ_c and Φ)."""
    import torch
    import boltmi
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    par = boltmi.CosmoParams(h=0.7, Ω_b=0.046, Ω_c=0.224, n=0.96, A=2.1e-9, N_ν=3.046, Σm_ν=0.0)
    bg = boltmi.Background(par)
    ih = boltmi.IonizationHistory(boltmi.RECFAST(bg), par, bg)
    ks = 0.7 * np.logspace(-4, 0, 100)
    xs = np.linspace(-20.0, 0.0, 1000)
    cfg = boltmi.hier_cfg(10, 10, 10, NQ, 1e-6, 1e-6)
    n = boltmi.state_length(cfg)

    nin = args.inflight or 3
    ctxs = [boltmi.default_context()] + [boltmi.Context(0) for _ in range(max(1, nin) - 1)]

    def step(ctx):
        out, status, st = boltmi.boltsolve_batch(par, bg, cfg, ks, xs, ctx=ctx)
        pk = boltmi.plin(ks, par, bg, ih, n_q=NQ, ℓᵧ=10, ℓ_ν=10, ℓ_mν=10, ctx=ctx)
        return out, st, pk

    def run(nsteps):
        """nsteps steps spread over the contexts, each context on its own thread"""
        import threading
        counts = [nsteps // len(ctxs) + (1 if j < nsteps % len(ctxs) else 0) for j in range(len(ctxs))]
        res, err = [None] * len(ctxs), []

        def work(j):
            try:
                for _ in range(counts[j]):
                    res[j] = step(ctxs[j])
            except Exception as e:  # reported after the join
                err.append(e)

        th = [threading.Thread(target=work, args=(j,)) for j in range(len(ctxs))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        if err:
            raise err[0]
        return res[0]

    run(max(args.warmup, len(ctxs)))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out, st, pk = run(args.steps)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    line = {"metric": "k-mode solves/s, plain-method workload (solve with 1000 saved states + P(k))",
            "value": len(ks) / dt, "unit": "k-mode solves/s", "n_gpus": 1, "steps": args.steps,
            "ms_per_step": dt * 1e3, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "flat LCDM, l_g=l_nu=l_mnu=10, nq=15 (n=203), 100 log k in [1e-4,1] h/Mpc, "
                                   "reltol=abstol=1e-6, states at 1000 x, then P(k) at the same k",
                       "n_kmodes": len(ks), "state_length": n, "inflight": len(ctxs)},
            "inflight": len(ctxs), "solver_kernel_ms": st["kernel_ms"], "steps_accepted": st["accepted"]}
    if not args.no_cpu:
        import pyoracle
        from concurrent.futures import ThreadPoolExecutor
        o = pyoracle.Oracle(par.as_tuple(), bg.tables_dict())
        oc = pyoracle.Oracle.cfg(10, 10, 10, NQ, 1e-6, 1e-6, 0, 1)
        sample = list(range(0, 100, 10))
        thr = args.cpu_threads or min(16, os.cpu_count() or 1)

        def one(i):
            t = time.perf_counter()
            ref, _ = o.boltsolve(oc, float(ks[i]), xs)
            return time.perf_counter() - t, ref

        with ThreadPoolExecutor(thr) as ex:
            res = list(ex.map(one, sample))
        err = 0.0
        for (_, ref), i in zip(res, sample):
            for col in (n - 5, n - 4):  # Φ and δ_c
                err = max(err, float(np.max(np.abs(out[i][:, col] - ref[:, col])) / np.max(np.abs(ref[:, col]))))
        t_cpu = sum(t for t, _ in res) * len(ks) / len(sample) / thr
        line["cpu_baseline"] = {"value": len(ks) / t_cpu, "unit": "k-mode solves/s (solve only)", "cores": thr,
                                "kind": "port", "sample": f"oracle boltsolve on {len(sample)} of {len(ks)} modes, "
                                f"{thr} threads; per-mode times summed and scaled"}
        line["max_scaled_err_phi_delta_c"] = err
    print(json.dumps(line), flush=True)


def batch(args):
    """The many-cosmology workload, measured apart from the default line.
    --ncosmo cosmologies (64) are drawn from the parameter box:
    h in [0.6, 0.8], ω_b in [0.020, 0.024], ω_c in [0.10, 0.14], n_s in [0.92, 1.0],
    ln(10¹⁰A_s) in [2.9, 3.2]. Odd ones have one massive ν in [0.06, 0.3] eV
    with ℓ_mν = 20 (n = 473); even ones are massless with ℓ_mν = 10 (n = 323).
    Each gets 800 log k in [1e-4, 0.5] Mpc⁻¹ and source_grid at reltol 1e-11:
    51,200 (cosmology, k) pairs in one persistent launch, sources left in HBM.
    The host table stage (bolt_host_tables_batch on --cpu-threads threads) is
    timed separately."""
    import torch
    import boltmi
    from boltmi import _abi
    rng = np.random.default_rng(args.seed)
    pars, lm = [], []
    for i in range(args.ncosmo):
        h = rng.uniform(0.6, 0.8)
        massive = i % 2 == 1
        pars.append(boltmi.CosmoParams(h=h, Ω_b=rng.uniform(0.020, 0.024) / h ** 2, Ω_c=rng.uniform(0.10, 0.14) / h ** 2,
                                       n=rng.uniform(0.92, 1.0), A=math.exp(rng.uniform(2.9, 3.2)) / 1e10,
                                       Σm_ν=(rng.uniform(0.06, 0.3) if massive else 0.0) * boltmi.MASS_NATURAL))
        lm.append(20 if massive else 10)
    thr = args.cpu_threads or min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    cosmos = boltmi.backgrounds(pars, threads=thr)
    t_tab = time.perf_counter() - t0
    ctx = boltmi.default_context()
    L = _abi.lib()
    nc = len(cosmos)
    ids = np.array([bg.cosmo_id(ctx) for _, bg, _ in cosmos], dtype=np.int32)
    cfgs = (_abi.HierCfgC * nc)(*[boltmi.hier_cfg(LG, LN, m, NQ, RELTOL, ABSTOL) for m in lm])
    ks = np.logspace(-4, math.log10(0.5), NK)
    dev = torch.device("cuda:0")
    d_k = torch.tensor(np.tile(ks, (nc, 1)), dtype=torch.float64, device=dev)
    d_S = torch.empty((nc, NK, cosmos[0][1].nx), dtype=torch.float64, device=dev)
    d_status = torch.zeros((nc, NK), dtype=torch.int32, device=dev)
    d_ctr = torch.zeros((nc * NK, 4), dtype=torch.int64, device=dev)

    def run(m):
        ms = ctypes.c_float()
        _abi.check(L.bolt_source_grid_batch_dev(ctx.handle, m, _abi.pi(ids), cfgs, d_k.data_ptr(), NK,
                                                d_S.data_ptr(), None, d_status.data_ptr(), d_ctr.data_ptr(),
                                                ctypes.byref(ms)))
        return ms.value

    run(min(2, nc))  # warm-up on the first two cosmologies (one of each hierarchy)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms = run(nc)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(f"# batch: {nc} cosmologies x {NK} k in {dt:.2f} s (kernel {ms:.0f} ms), tables {t_tab:.2f} s on {thr} "
          f"threads", file=sys.stderr, flush=True)
    if int(d_status.max().item()) != 0:
        raise RuntimeError("a k-mode failed in the batch run")
    ctr = d_ctr.cpu().numpy()
    line = {
        "metric": "k-mode solves/s, many-cosmology batch (source_grid, one launch)", "value": nc * NK / dt,
        "unit": "k-mode solves/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{nc} cosmologies from the parameter box (half with one massive nu, l_mnu=20, "
                               f"n=473; half massless, l_mnu=10, n=323) x {NK} log k in [1e-4,0.5]/Mpc, "
                               f"l_g=l_nu=50, nq=15, reltol 1e-11, one persistent launch",
                   "n_cosmo": nc, "n_kmodes": nc * NK, "seed": args.seed},
        "batch_s": dt, "kernel_ms": ms, "source_grids_per_s": nc / dt,
        "ode_steps_total": int(ctr[:, 0].sum() + ctr[:, 1].sum()),
        "host_tables": {"seconds": t_tab, "threads": thr, "per_cosmology_ms": t_tab / nc * 1e3 * thr},
    }
    print(json.dumps(line), flush=True)


# hierarchy tuples (ℓᵧ, ℓ_ν, ℓ_mν, nq) of the precision sweep: n = 128, 224, 233, 256, 317
C4_TUPLES = [(20, 20, 5, 10), (12, 12, 8, 20), (40, 40, 6, 15), (16, 16, 4, 40), (30, 25, 6, 32)]


def c4(args):
    """Precision sweep: --ncosmo cosmologies (256) from the parameter box, half
    with one massive ν; hierarchy tuple c mod 5 of C4_TUPLES (n = 128..317);
    a k grid of 200, 400 or 800 log points in [1e-4, 0.5] Mpc⁻¹ drawn per
    cosmology.  Every cosmology's source grid (reltol 1e-11), C_ℓ^TT for
    ℓ = 2..2500 (5000 dense k) and P(k) at 200 log k (plin, reltol 1e-5) are
    computed with FP64 stage solves and with FP32 solves + FP64 refinement
    (the default: until every correction is below 1e-14 of its component, at
    most 12 sweeps, else the step is rejected).  Errors: FP32 against FP64 on
    every output of every cosmology; both against the CPU FP64 oracle on samples
    (P(k) at 6 k of 10 cosmologies, C_ℓ^TT at all ℓ for the two smallest
    hierarchies from the oracle's own source grids); and, for scale, P(k) at
    plin's reltol 1e-5 against FP64 P(k) at reltol 1e-9 (the integration's own
    error).  Solver counters (steps, rejections, RHS evaluations, linear solves)
    are summed over the source-grid modes.  FP16 matrix-core solves are not
    offered: the stage solve is banded plus rank 3, with no dense block."""
    import torch
    import boltmi
    from boltmi import _abi
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(args.seed)
    nc = args.ncosmo
    pars, tup, nks = [], [], []
    for c in range(nc):
        h = rng.uniform(0.6, 0.8)
        pars.append(boltmi.CosmoParams(h=h, Ω_b=rng.uniform(0.020, 0.024) / h ** 2, Ω_c=rng.uniform(0.10, 0.14) / h ** 2,
                                       n=rng.uniform(0.92, 1.0), A=math.exp(rng.uniform(2.9, 3.2)) / 1e10,
                                       Σm_ν=(rng.uniform(0.06, 0.3) if c % 2 else 0.0) * boltmi.MASS_NATURAL))
        tup.append(C4_TUPLES[c % len(C4_TUPLES)])
        nks.append(int(rng.choice([200, 400, 800])))
    thr = args.cpu_threads or min(16, os.cpu_count() or 1)
    cosmos = [None] * nc
    for t in C4_TUPLES:  # tables per momentum-bin count
        idx = [c for c in range(nc) if tup[c] == t]
        for c, cb in zip(idx, boltmi.backgrounds([pars[c] for c in idx], nq=t[3], threads=thr)):
            cosmos[c] = cb
    ctx = boltmi.default_context()
    L = _abi.lib()
    dev = torch.device("cuda:0")
    ells = np.arange(2, LMAX + 1, dtype=np.int32)
    kpk = np.logspace(-4, math.log10(0.5), 200)
    modes = {"fp64": 0, "fp32_refined": boltmi.precision_flags("fp32")}
    counters = {m: np.zeros(4, dtype=np.int64) for m in modes}
    out = {m: {"cl": [None] * nc, "pk": [None] * nc, "S": {}} for m in modes}
    timing = {}
    groups = sorted({(tup[c], nks[c]) for c in range(nc)})
    with ctx.lock:
        ids = [bg.cosmo_id(ctx) for _, bg, _ in cosmos]
    cl_pool = set()  # massless cosmologies of the two smallest hierarchies at nk = 200: the C_ℓ oracle samples
    for mname, flags in modes.items():
        t_sg = t_los = t_pk = 0.0
        nmodes = 0
        for t, nk in groups:
            idx = [c for c in range(nc) if tup[c] == t and nks[c] == nk]
            ks = np.logspace(-4, math.log10(0.5), nk)
            cfgs = (_abi.HierCfgC * len(idx))(*[boltmi.hier_cfg(t[0], t[1], t[2], t[3], 1e-11, 1e-6, flags=flags)
                                                for _ in idx])
            d_k = torch.tensor(np.tile(ks, (len(idx), 1)), dtype=torch.float64, device=dev)
            d_S = torch.empty((len(idx), nk, cosmos[idx[0]][1].nx), dtype=torch.float64, device=dev)
            d_st = torch.zeros((len(idx), nk), dtype=torch.int32, device=dev)
            d_ctr = torch.zeros((len(idx), nk, 4), dtype=torch.int64, device=dev)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ms = ctypes.c_float()
            with ctx.lock:
                _abi.check(L.bolt_source_grid_batch_dev(ctx.handle, len(idx), _abi.pi(np.array([ids[c] for c in idx],
                                                                                            dtype=np.int32)),
                                                        cfgs, d_k.data_ptr(), nk, d_S.data_ptr(), None, d_st.data_ptr(),
                                                        d_ctr.data_ptr(), ctypes.byref(ms)))
            torch.cuda.synchronize()
            t_sg += time.perf_counter() - t0
            counters[mname] += d_ctr.sum(dim=(0, 1)).cpu().numpy()
            nmodes += len(idx) * nk
            if int(d_st.max().item()) != 0:
                raise RuntimeError(f"a mode failed in the {mname} sweep")
            dense = None
            for j, c in enumerate(idx):
                bg = cosmos[c][1]
                dense = torch.tensor(boltmi.quadratic_k(0.01 * bg.H0, 1000 * bg.H0, NKD), device=dev)
                d_cl = torch.empty(len(ells), dtype=torch.float64, device=dev)
                t0 = time.perf_counter()
                with ctx.lock:
                    _abi.check(L.bolt_cltt_dev(ctx.handle, ids[c], d_S[j].data_ptr(), d_k[j].data_ptr(), nk,
                                               dense.data_ptr(), NKD, _abi.pi(ells), len(ells), d_cl.data_ptr(), None,
                                               None))
                t_los += time.perf_counter() - t0
                out[mname]["cl"][c] = d_cl.cpu().numpy()
                if nk == 200 and t in C4_TUPLES[:2] and c % 2 == 0:
                    cl_pool.add(c)
            del d_S
        for c in range(nc):
            par, bg, ih = cosmos[c]
            t = tup[c]
            t0 = time.perf_counter()
            out[mname]["pk"][c] = boltmi.plin(kpk, par, bg, ih, n_q=t[3], ℓᵧ=t[0], ℓ_ν=t[1], ℓ_mν=t[2], flags=flags)
            t_pk += time.perf_counter() - t0
        timing[mname] = {"source_grids_s": t_sg, "kmode_solves_per_s": nmodes / t_sg, "los_s": t_los, "plin_s": t_pk}
        print(f"# c4 {mname}: source grids {t_sg:.2f} s ({nmodes / t_sg:.0f} k-mode solves/s), LOS {t_los:.2f} s, "
              f"plin {t_pk:.2f} s", file=sys.stderr, flush=True)

    def rel(a, b):
        return float(np.max(np.abs(np.asarray(a) / np.asarray(b) - 1)))

    # the integration's own error at plin's reltol: FP64 P(k) at reltol 1e-9
    t0 = time.perf_counter()
    pk_tight = []
    for c in range(nc):
        par, bg, ih = cosmos[c]
        t = tup[c]
        pk_tight.append(boltmi.plin(kpk, par, bg, ih, n_q=t[3], ℓᵧ=t[0], ℓ_ν=t[1], ℓ_mν=t[2], reltol=1e-9))
    print(f"# c4 fp64 P(k) at reltol 1e-9: {time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)
    vs_tight = {m: max(rel(out[m]["pk"][c], pk_tight[c]) for c in range(nc)) for m in modes}

    fp32_vs_64 = {"pk": max(rel(out["fp32_refined"]["pk"][c], out["fp64"]["pk"][c]) for c in range(nc)),
                  "cltt": max(rel(out["fp32_refined"]["cl"][c], out["fp64"]["cl"][c]) for c in range(nc))}
    # oracle samples: P(k) at 6 k for 10 cosmologies spread over the tuples
    pk_samples = [(c, i) for c in range(0, nc, max(1, nc // 10))[:10] for i in (0, 40, 80, 120, 160, 199)]

    def pk_ref(ci):
        c, i = ci
        par, bg, _ = cosmos[c]
        t = tup[c]
        o = pyoracle.Oracle(par.as_tuple(), bg.tables_dict())
        return o.plin(pyoracle.Oracle.cfg(t[0], t[1], t[2], t[3], 1e-5, 1e-6), float(kpk[i]))[0]

    with ThreadPoolExecutor(thr) as ex:
        pk_o = list(ex.map(pk_ref, pk_samples))
    err_pk = {m: max(abs(out[m]["pk"][c][i] / r - 1) for (c, i), r in zip(pk_samples, pk_o)) for m in modes}
    # oracle samples: C_ℓ^TT from the oracle's own source grid, two smallest hierarchies at nk = 200
    cl_samples = sorted(cl_pool)[:2]
    err_cl = {m: 0.0 for m in modes}
    for c in cl_samples:
        par, bg, _ = cosmos[c]
        t = tup[c]
        o = pyoracle.Oracle(par.as_tuple(), bg.tables_dict())
        ks = np.logspace(-4, math.log10(0.5), 200)
        S, st = o.source_grid(pyoracle.Oracle.cfg(t[0], t[1], t[2], t[3], 1e-11, 1e-6), ks, nthreads=thr)
        ref = o.cltt(S, ks, boltmi.quadratic_k(0.01 * bg.H0, 1000 * bg.H0, NKD), ells, 0, nthreads=thr)
        for m in modes:
            err_cl[m] = max(err_cl[m], rel(out[m]["cl"][c], ref))
    line = {
        "metric": "precision sweep: k-mode solves/s and per-output max rel err vs the FP64 oracle",
        "value": timing["fp32_refined"]["kmode_solves_per_s"], "unit": "k-mode solves/s (FP32 solve + refinement)",
        "n_gpus": 1, "higher_is_better": True, "dtype": "f64 state, f32 stage solves", "data": "synthetic",
        "config": {"workload": f"{nc} cosmologies from the parameter box (half with one massive nu), hierarchy tuples "
                               f"{C4_TUPLES} (n = 128..317), nk in {{200, 400, 800}}, source_grid reltol 1e-11, "
                               f"C_l^TT l = 2..2500 on 5000 dense k, P(k) at 200 k (reltol 1e-5)",
                   "n_cosmo": nc, "seed": args.seed},
        "timing": timing,
        "solver_counters": {m: dict(zip(("steps_accepted", "steps_rejected", "rhs_evals", "linear_solves"),
                                        (int(v) for v in counters[m]))) for m in modes},
        "max_rel_err": {
            "fp32_refined_vs_fp64_all_cosmologies": fp32_vs_64,
            "pk_reltol_1e-5_vs_fp64_reltol_1e-9_all_cosmologies": vs_tight,
            "vs_oracle_pk": {m: float(v) for m, v in err_pk.items()},
            "vs_oracle_cltt": err_cl,
            "oracle_samples": {"pk": f"{len(pk_samples)} (cosmology, k) pairs", "cltt": f"cosmologies {cl_samples}, "
                               f"all 2499 l, oracle source grid at nk = 200"}},
        "fp16_matrix_core": "not offered: the stage solve is banded + rank-3, no dense matrix to put on MFMA",
    }
    print(json.dumps(line), flush=True)


def jvp(args):
    """Forward-mode derivative of the full-CMB spectrum on the device: source_grid
    (800 log k, ℓᵧ = ℓ_ν = 50, ℓ_mν = 10, reltol 1e-11) and C_ℓ^TT (ℓ = 2..2500 on
    5000 dense k) with one Dual direction (∂/∂Ω_b, as examples/plot_deriv_cl.jl),
    against the same computation without the derivative.  Reports the kernels'
    device times and the derivative's cost ratio."""
    import torch
    import boltmi
    from boltmi.dual import Dual, value
    kw = dict(h=0.7, Ω_b=0.046, Ω_c=0.224, n=0.96, A=2.1e-9, Σm_ν=0.06 * boltmi.MASS_NATURAL)
    par0 = boltmi.CosmoParams(**kw)
    bg0 = boltmi.Background(par0)
    ih0 = boltmi.IonizationHistory(boltmi.RECFAST(bg0), par0, bg0)
    pard = boltmi.CosmoParams(**dict(kw, Ω_b=Dual(0.046, 1.0)))
    bgd = boltmi.Background(pard)
    ihd = boltmi.IonizationHistory(boltmi.RECFAST(bgd), pard, bgd)
    ks = np.logspace(-4, math.log10(0.5), NK)
    dense = boltmi.quadratic_k(0.01 * bg0.H0, 1000 * bg0.H0, NKD)
    ells = np.arange(2, LMAX + 1)
    out = {}
    for name, (par, bg, ih) in (("primal", (par0, bg0, ih0)), ("with_derivative", (pard, bgd, ihd))):
        for rep in range(2):  # the first call loads the kernels
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sf = boltmi.source_grid(par, bg, ih, ks, None, 50, 1e-11, ℓ_ν=50, ℓ_mν=10)
            t1 = time.perf_counter()
            cl = boltmi.cltt_on_grid(ells, sf, dense, par, bg)
            t2 = time.perf_counter()
        out[name] = {"source_grid_s": t1 - t0, "solver_kernel_ms": sf.stats["kernel_ms"], "cltt_s": t2 - t1,
                     "steps_accepted": int(sf.stats["accepted"])}
        print(f"# jvp {name}: source_grid {t1 - t0:.3f} s (kernel {sf.stats['kernel_ms']:.1f} ms), "
              f"cltt {t2 - t1:.3f} s", file=sys.stderr, flush=True)
    line = {"metric": "derivative of the full-CMB spectrum along one parameter (source_grid + C_l^TT with "
                      "d/dOmega_b), cost relative to the spectrum alone",
            "value": out["with_derivative"]["solver_kernel_ms"] / out["primal"]["solver_kernel_ms"],
            "unit": "solver kernel time ratio (derivative / primal)", "n_gpus": 1, "higher_is_better": False,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": "800 log k in [1e-4, 0.5]/Mpc, l_g = l_nu = 50, l_mnu = 10, nq = 15, reltol 1e-11; "
                                   "C_l^TT l = 2..2500 on 5000 dense k; one Dual direction"},
            "timing": out}
    print(json.dumps(line), flush=True)


def tables(args):
    """Background + RECFAST tables for n cosmologies from the serving box (half
    with a massive ν): bolt_device_tables_batch (one wavefront per cosmology) and
    bolt_host_tables_batch on the host's threads, after one untimed device call."""
    import boltmi
    rng = np.random.default_rng(args.seed)

    def box(n):
        out = []
        for c in range(n):
            h = rng.uniform(0.6, 0.8)
            out.append(boltmi.CosmoParams(h=h, Ω_b=rng.uniform(0.020, 0.024) / h ** 2,
                                          Ω_c=rng.uniform(0.10, 0.14) / h ** 2, n=rng.uniform(0.92, 1.0),
                                          A=math.exp(rng.uniform(2.9, 3.2)) * 1e-10, N_ν=3.046,
                                          Σm_ν=(rng.uniform(0.06, 0.3) if c % 2 else 0.0) * boltmi.MASS_NATURAL))
        return out
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    boltmi.backgrounds(box(1), device=True)
    rows = {}
    for n in (1, 64, 256, 1024):
        p = box(n)
        t0 = time.perf_counter()
        boltmi.backgrounds(p, device=True)
        t1 = time.perf_counter()
        boltmi.backgrounds(p, threads=threads)
        t2 = time.perf_counter()
        rows[str(n)] = {"device_ms": (t1 - t0) * 1e3, "host_ms": (t2 - t1) * 1e3}
        print(f"# {n} cosmologies: device {rows[str(n)]['device_ms']:.0f} ms, host ({threads} threads) "
              f"{rows[str(n)]['host_ms']:.0f} ms", file=sys.stderr, flush=True)
    print(json.dumps({"metric": "Background + RECFAST tables, wall time per batch", "unit": "ms",
                      "host_threads": threads, "batches": rows}), flush=True)


def serve(args):
    print(json.dumps(serve_run(args.rate, args.requests, args.slots, args.hbm_gib, args.seed, args.queue)), flush=True)


def serve_run(rate, nreq, slots, hbm_gib, seed, queue=False):
    """Serving workload: seeded Poisson arrivals of spectrum requests, each with
    its own cosmology drawn from the box h ∈ [0.6, 0.8], Ω_b h² ∈ [0.020, 0.024],
    Ω_c h² ∈ [0.10, 0.14], n_s ∈ [0.92, 1.0], ln 10¹⁰A_s ∈ [2.9, 3.2], 30 % with
    one massive ν (Σm_ν ∈ [0.06, 0.3] eV, ℓ_mν = 20); k-grid size from
    {50, 200, 800}; half P(k) (plin defaults), half C_ℓ^TT to ℓ_max from
    {500, 1500, 2500} (ℓᵧ = ℓ_ν = 50, reltol 1e-11, 5000 dense k).  Latency is
    submit to result, host tables included.  queue=True serves through the
    device work queue (one persistent solver launch that requests join, HBM
    from an arena by request size, 8 host threads for the tables) instead of
    `slots` request slots."""
    import boltmi
    rng = np.random.default_rng(seed)
    srv = boltmi.Server(slots=slots, hbm_gib=hbm_gib, queue=queue, host_threads=8)
    reqs = []
    for _ in range(nreq):
        h = rng.uniform(0.6, 0.8)
        massive = rng.uniform() < 0.3
        par = boltmi.CosmoParams(h=h, Ω_b=rng.uniform(0.020, 0.024) / h**2, Ω_c=rng.uniform(0.10, 0.14) / h**2,
                                 n=rng.uniform(0.92, 1.0), A=math.exp(rng.uniform(2.9, 3.2)) * 1e-10,
                                 N_ν=3.046, Σm_ν=(rng.uniform(0.06, 0.3) if massive else 0.0) * boltmi.MASS_NATURAL)
        nk = int(rng.choice([50, 200, 800]))
        if rng.uniform() < 0.5:
            reqs.append(("pk", par, dict(nk=nk)))
        else:
            reqs.append(("cltt", par, dict(nk=nk, ell_max=int(rng.choice([500, 1500, 2500])),
                                           ℓ_mν=20 if massive else 10)))
    gaps = rng.exponential(1.0 / rate, size=len(reqs))
    try:
        # one small request of each kind first (not timed): loads the kernels
        warm = [srv.submit_pk(boltmi.CosmoParams(), nk=8), srv.submit_cltt(boltmi.CosmoParams(), nk=8, ell_max=50,
                                                                            nkd=200, ℓᵧ=8, ℓ_ν=8)]
        for w in warm:
            srv.wait(w)
        ids = []
        t0 = time.perf_counter()
        t_next = t0
        for (kind, par, kw), gap in zip(reqs, gaps):
            t_next += gap
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit_pk(par, **kw) if kind == "pk" else srv.submit_cltt(par, **kw))
        infos = [srv.wait(i)[1] for i in ids]
        t1 = time.perf_counter()
        qi = srv.queue_info() if queue else None
    finally:
        srv.close()
    lat = np.array([i["latency_ms"] for i in infos])
    line = {
        "metric": METRIC, "value": len(reqs) / (t1 - t0), "unit": "spectra/s", "n_gpus": 1, "steps": len(reqs),
        "warmup": 2, "ms_per_step": (t1 - t0) * 1e3 / len(reqs), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"serving: Poisson arrivals at {rate}/s, {len(reqs)} requests, "
                               + ("device work queue" if queue else f"{slots} slots")
                               + "; half P(k), half C_l^TT; 30% massive nu",
                   "seed": seed, "rate_per_s": rate},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(lat.max())},
        "queue_ms_p50": float(np.percentile([i["queue_ms"] for i in infos], 50)),
        "host_ms_p50": float(np.percentile([i["host_ms"] for i in infos], 50)),
        "device_ms_p50": float(np.percentile([i["device_ms"] for i in infos], 50)),
        "solve_ms_p50": float(np.percentile([i["solve_ms"] for i in infos], 50)),
        "roofline": None, "cpu_baseline": None,
    }
    if queue:
        line["queue"] = {"launches": qi["launches"], "joined_running": qi["joined_running"],
                         "arena_peak_mb": qi["arena_peak_bytes"] / 1e6, "solver_workgroups": qi["solver_workgroups"]}
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=12,
                    help="spectra timed (rounded up to a whole number of launches of --per-launch spectra)")
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--inflight", type=int, default=None,
                    help="independent batches in flight, each on its own context (HIP stream) and host thread "
                         "(default 2; 3 for --c1)")
    ap.add_argument("--per-launch", type=int, default=2, help="spectra solved by one persistent solver launch")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-mode-per-launch GPU baseline")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default solver and LOS shapes)")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the short serving leg (p50/p99 latency at two arrival rates) after the timed region")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this many MiB before the workload (reference "
                         "dispatch for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--c1", action="store_true", help="measure the plain-method workload (100 k, l=10, states + P(k))")
    ap.add_argument("--tables", action="store_true",
                    help="time the Background + RECFAST tables on the device and on the host for 1..1024 cosmologies")
    ap.add_argument("--jvp", action="store_true",
                    help="measure the forward-mode derivative of the full-CMB spectrum against the spectrum alone")
    ap.add_argument("--c4", action="store_true",
                    help="precision sweep: FP64 against FP32 + refinement stage solves over --ncosmo (256) cosmologies")
    ap.add_argument("--batch", action="store_true",
                    help="measure the many-cosmology batch (source_grid for --ncosmo cosmologies in one launch)")
    ap.add_argument("--ncosmo", type=int, default=0, help="--batch / --c4: number of cosmologies (64 / 256)")
    ap.add_argument("--serve", action="store_true",
                    help="measure the serving workload (Poisson request stream) instead of the full CMB path")
    ap.add_argument("--rate", type=float, default=50.0, help="--serve: mean request arrivals per second")
    ap.add_argument("--requests", type=int, default=48, help="--serve: number of requests")
    ap.add_argument("--slots", type=int, default=16, help="--serve: concurrent request slots")
    ap.add_argument("--hbm-gib", type=float, default=16.0, help="--serve: HBM bound for the slots")
    ap.add_argument("--seed", type=int, default=1, help="--serve: seed of the request stream")
    ap.add_argument("--queue", action="store_true", help="--serve: the device work-queue server instead of slots")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs the hot path on a single MI355X "
              "(multi-GPU sharding is described in DESIGN.md, not run).", file=sys.stderr)
        sys.exit(2)

    if args.serve:
        serve(args)
        return
    if args.tables:
        tables(args)
        return
    if args.batch:
        args.ncosmo = args.ncosmo or 64
        batch(args)
        return
    if args.c4:
        args.ncosmo = args.ncosmo or 256
        c4(args)
        return
    if args.c1:
        c1(args)
        return
    if args.jvp:
        jvp(args)
        return

    import torch
    import boltmi
    from boltmi import _abi

    par, bg, ih, ks, cfg, ells, dense = workload()
    nlanes = max(1, args.inflight or 2)
    B = max(1, args.per_launch)
    L = _abi.lib()
    dev = torch.device("cuda:0")
    d_k = torch.tensor(ks, dtype=torch.float64, device=dev)
    d_dense = torch.tensor(dense, dtype=torch.float64, device=dev)
    ells_c = np.ascontiguousarray(ells)

    class Lane:
        """one batch of B spectra in flight: a context (its own HIP stream) and its own result buffers"""

        def __init__(self, ctx):
            self.ctx = ctx
            self.cid = bg.cosmo_id(ctx)
            self.ids = (ctypes.c_int * B)(*([self.cid] * B))
            self.cfgs = (_abi.HierCfgC * B)(*([cfg] * B))
            self.d_kB = d_k.repeat(B).contiguous()
            self.d_SB = torch.empty((B, NK, bg.nx), dtype=torch.float64, device=dev)
            self.d_statusB = torch.zeros((B, NK), dtype=torch.int32, device=dev)
            self.d_ctrB = torch.zeros((B, NK, 4), dtype=torch.int64, device=dev)
            self.d_clB = torch.empty((B, len(ells)), dtype=torch.float64, device=dev)
            # the batch's first spectrum (the baselines and the error check use it)
            self.d_S, self.d_status, self.d_ctr, self.d_cl = self.d_SB[0], self.d_statusB[0], self.d_ctrB[0], self.d_clB[0]
            self.d_steps = torch.zeros(NKD - 1, dtype=torch.int64, device=dev)
            self.d_dense = d_dense
            self.ells = ells_c

    lanes = [Lane(boltmi.default_context())] + [Lane(boltmi.Context(0)) for _ in range(nlanes - 1)]
    d_ctr, d_steps = lanes[0].d_ctrB, lanes[0].d_steps
    if args.calib_copy_mib > 0:
        nb = args.calib_copy_mib << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    torch.cuda.synchronize()

    def step(ln):
        """B whole spectra on lane ln: one source_grid launch for all of them,
        then cltt from each grid; returns the solver launch's ms and the mean
        LOS ms per spectrum"""
        ms1 = ctypes.c_float()
        if B == 1:
            _abi.check(L.bolt_source_grid_dev(ln.ctx.handle, ln.cid, ctypes.byref(cfg), d_k.data_ptr(), NK,
                                              ln.d_S.data_ptr(), ln.d_status.data_ptr(), ln.d_ctr.data_ptr(),
                                              ctypes.byref(ms1)))
        else:
            _abi.check(L.bolt_source_grid_batch_dev(ln.ctx.handle, B, ln.ids, ln.cfgs, ln.d_kB.data_ptr(), NK,
                                                    ln.d_SB.data_ptr(), None, ln.d_statusB.data_ptr(),
                                                    ln.d_ctrB.data_ptr(), ctypes.byref(ms1)))
        los = 0.0
        for b in range(B):
            ms2 = ctypes.c_float()
            _abi.check(L.bolt_cltt_dev(ln.ctx.handle, ln.cid, ln.d_SB[b].data_ptr(), d_k.data_ptr(), NK,
                                       d_dense.data_ptr(), NKD, _abi.pi(ells_c), len(ells_c), ln.d_clB[b].data_ptr(),
                                       ln.d_steps.data_ptr(), ctypes.byref(ms2)))
            los += ms2.value
        return ms1.value, los / B

    def run(counts, label):
        """counts[j] steps on lane j, the lanes concurrently (the C calls release the GIL)"""
        import threading
        res = [[] for _ in lanes]
        err = []
        done = [0]
        lock = threading.Lock()

        def work(j):
            try:
                for _ in range(counts[j]):
                    t = time.perf_counter()
                    a, b = step(lanes[j])
                    res[j].append((a, b, (time.perf_counter() - t) * 1e3))
                    with lock:
                        done[0] += 1
                        print(f"# {label} {done[0]}/{sum(counts)} (lane {j}): solver {a:.1f} ms, LOS {b:.2f} ms",
                              file=sys.stderr, flush=True)
            except Exception as e:  # reported after the join
                err.append(e)

        th = [threading.Thread(target=work, args=(j,)) for j in range(len(lanes))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        if err:
            raise err[0]
        return [r for rr in res for r in rr]

    def split(k):
        """k spectra as launches of B per lane"""
        nl = -(-k // B)
        return [nl // len(lanes) + (1 if j < nl % len(lanes) else 0) for j in range(len(lanes))]

    # the project's autotuner picks the solver and LOS shapes for this workload
    # (untimed; the choice holds for every context on the device)
    tune = None
    if not args.no_tune:
        tune = boltmi.autotune(par, bg, cfg, ks, dense, LMAX, ctx=lanes[0].ctx)
        print(f"# autotune: solver SR {tune['solver_sr']} {tune['solver_ms']}, LOS slots/wave {tune['los_spw']} "
              f"{tune['los_ms']}", file=sys.stderr, flush=True)
    # every lane runs at least one untimed step (its context's first call sizes its cached buffers)
    run([max(c, 1) for c in split(args.warmup)], "warmup")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    counts = split(args.steps)
    nspec = sum(counts) * B
    timed = run(counts, "step")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kms = [(a, b) for a, b, _ in timed]
    lat = [w for _, _, w in timed]
    dt = (t1 - t0) / nspec
    for ln in lanes:
        if int(ln.d_statusB.max().item()) != 0:
            raise RuntimeError("a k-mode failed in the benchmark run")
    ctr = d_ctr.cpu().numpy().reshape(-1, 4)
    steps_total = int(ctr[:, 0].sum() + ctr[:, 1].sum())
    acc_total = int(ctr[:, 0].sum())
    n = boltmi.state_length(cfg)
    # per solver launch (B spectra)
    f_ode = ode_flops(n, NQ, steps_total, acc_total, NK * B, bg.nx)
    los_exec = int(d_steps.sum().item())
    nxi = int(np.sum(bg.x_grid > -8.0)) - 1
    f_los_nominal = 4.0 * len(ells) * (NKD - 1) * nxi
    ms_ode = float(np.mean([a for a, b in kms]))
    ms_los = float(np.mean([b for a, b in kms]))
    ode_tf = f_ode / (ms_ode * 1e-3) / 1e12
    los_tf = f_los_nominal / (ms_los * 1e-3) / 1e12
    los_tf_exec = 4.0 * los_exec / (ms_los * 1e-3) / 1e12
    if ms_ode >= ms_los:
        roof = {"kernel": "k_boltsolve", "bound": "fp64-vector", "achieved": ode_tf, "peak": PEAK_FP64,
                "unit": "TFLOP/s", "frac": ode_tf / PEAK_FP64, "traffic": measured_traffic("k_boltsolve", B),
                "chip_level_tflops": f_ode / B / dt / 1e12,
                "note": "FP64 vector peak (no MFMA instruction is used); algorithmic flops of the structured "
                        "ESDIRK solve per step/RHS/node, from the kernel's own step counters, for one launch "
                        "(per_launch spectra) over the kernel's average HIP-event duration (which with batches in "
                        "flight includes waiting for CUs); chip_level_tflops divides one spectrum's flops by the "
                        "wall time per spectrum"}
    else:
        roof = {"kernel": "k_los", "bound": "fp64-vector", "achieved": los_tf, "peak": PEAK_FP64, "unit": "TFLOP/s",
                "frac": los_tf / PEAK_FP64, "traffic": measured_traffic("k_los")}
    line = {
        "metric": METRIC, "value": 1.0 / dt, "unit": "spectra/s", "n_gpus": 1, "steps": nspec,
        "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "full CMB path: source_grid 800 log k in [1e-4,0.5]/Mpc (l_g=l_nu=50, l_mnu=10, "
                               "nq=15, reltol 1e-11) + C_l^TT l=2..2500 on quadratic_k(0.01H0,1000H0,5000)",
                   "n_kmodes": NK, "state_length": n, "n_ell": len(ells), "n_kdense": NKD,
                   "inflight": len(lanes), "per_launch": B},
        "kmode_solves_per_s": NK / dt,
        "inflight": len(lanes),
        "per_launch": B,
        "latency_ms": {"mean": float(np.mean(lat)), "max": float(np.max(lat)),
                       "note": "wall time of one batch of per_launch spectra on its lane"},
        "kernels_ms": {"k_boltsolve": ms_ode, "k_los": ms_los},
        "ode_steps_total": steps_total,
        "los": {"nominal_tflops": los_tf, "executed_tflops": los_tf_exec, "executed_steps": los_exec,
                "nominal_steps": f_los_nominal / 4.0},
        "roofline": roof,
        "autotune": tune,
    }
    if not args.no_plain:
        line["plain_gpu_baseline"] = plain_gpu_baseline(lanes[0], L, cfg, d_k, value=1.0 / dt)
    if not args.no_cpu:
        thr = args.cpu_threads or min(16, os.cpu_count() or 1)
        cb = cpu_baseline(par, bg, ks, ells, dense, thr, lanes[0].d_S.cpu().numpy(), lanes[0].d_cl.cpu().numpy())
        line["max_rel_err_vs_oracle"] = cb.pop("gpu_max_err")
        line["cpu_baseline"] = cb
    else:
        line["cpu_baseline"] = None
    if not args.no_serve:
        # after the timed region: p50/p99 of a seeded Poisson request stream at
        # the benchmark's arrival rate and at one below the device's capacity
        # (request slots, and the device work queue)
        legs = {}
        for queue in (False, True):
            for rate in (50.0, 4.0):
                r = serve_run(rate, 24, args.slots, args.hbm_gib, args.seed, queue)
                name = "queue" if queue else "slots"
                print(f"# serving ({name}) at {rate}/s: p50 {r['latency_ms']['p50']:.0f} ms, "
                      f"p99 {r['latency_ms']['p99']:.0f} ms, {r['value']:.2f} spectra/s", file=sys.stderr, flush=True)
                leg = {"spectra_per_s": r["value"], "latency_ms": r["latency_ms"], "queue_ms_p50": r["queue_ms_p50"],
                       "device_ms_p50": r["device_ms_p50"], "requests": r["steps"]}
                if queue:
                    leg.update(r["queue"])
                legs[f"{name}_rate_{rate:g}_per_s"] = leg
        line["serving"] = {"workload": "24 requests per rate: cosmologies from the parameter box (30% massive nu), "
                                       "nk in {50,200,800}, half P(k), half C_l^TT to l_max in {500,1500,2500}; "
                                       "latency from submit to result, host tables included; slots: "
                                       f"{args.slots} request slots; queue: one persistent solver launch fed "
                                       "through a device work queue",
                           **legs}
        line["p99_latency_ms"] = min(legs["slots_rate_50_per_s"]["latency_ms"]["p99"],
                                     legs["queue_rate_50_per_s"]["latency_ms"]["p99"])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
