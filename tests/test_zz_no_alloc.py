"""bolt_cltt_dev_noalloc (no allocation: the LOS work space is the library's
per-GPU device array) against bolt_cltt_dev on the same device inputs: bitwise
equal C_ℓ, on two calls in a row, the second one smaller; sizes beyond the
work space are rejected."""
import ctypes

import numpy as np
import pytest

import boltmi
import cosmo
from boltmi import _abi

pytestmark = pytest.mark.gpu


def test_noalloc_cltt_matches_cltt_dev_twice():
    import torch
    par, bg, ih = cosmo.background("default")
    ctx = boltmi.default_context()
    L = _abi.lib()
    ks = boltmi.quadratic_k(0.1 * bg.H0, 1000 * bg.H0, 60)
    sf = boltmi.source_grid(par, bg, ih, ks, None, 8, 1e-8)
    dev = torch.device("cuda:0")
    d_S = torch.tensor(sf.grid, device=dev)
    d_k = torch.tensor(ks, device=dev)
    with ctx.lock:
        cid = bg.cosmo_id(ctx)
    for nkd, lmax in ((3000, 2900), (700, 900)):
        d_dense = torch.tensor(boltmi.quadratic_k(0.01 * bg.H0, 1000 * bg.H0, nkd), device=dev)
        ells = np.concatenate([np.arange(2, lmax + 1), [10, 2, 10]]).astype(np.int32)
        outs = []
        for fn in (L.bolt_cltt_dev, L.bolt_cltt_dev_noalloc):
            d_cl = torch.full((len(ells),), float("nan"), dtype=torch.float64, device=dev)
            d_steps = torch.zeros(nkd - 1, dtype=torch.int64, device=dev)
            ms = ctypes.c_float()
            with ctx.lock:
                _abi.check(fn(ctx.handle, cid, d_S.data_ptr(), d_k.data_ptr(), len(ks), d_dense.data_ptr(), nkd,
                              _abi.pi(ells), len(ells), d_cl.data_ptr(), d_steps.data_ptr(), ctypes.byref(ms)))
            outs.append((d_cl.cpu().numpy(), d_steps.cpu().numpy()))
        (ref, ref_steps), (got, got_steps) = outs
        assert np.all(np.isfinite(got)) and np.all(got > 0)
        assert np.array_equal(got, ref), (nkd, lmax)
        assert np.array_equal(got_steps, ref_steps)
        assert got[-3] == got[8] and got[-1] == got[8] and got[-2] == got[0]   # repeated ℓ
    # beyond the work space: a clear error, no launch
    big = np.arange(2, 3002, dtype=np.int32)
    d_cl = torch.zeros(len(big), dtype=torch.float64, device=dev)
    with ctx.lock:
        rc = L.bolt_cltt_dev_noalloc(ctx.handle, cid, d_S.data_ptr(), d_k.data_ptr(), len(ks), d_dense.data_ptr(),
                                     700, _abi.pi(big), len(big), d_cl.data_ptr(), None, None)
    assert rc == _abi.BOLT_E_ARG
