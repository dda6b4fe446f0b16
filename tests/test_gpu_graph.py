"""A whole spectrum captured as one HIP graph (bolt_graph_spectrum_create):
every replay recomputes source_grid + cltt bit for bit as the eager device
calls do, also with other eager work on the device between the replays."""
import ctypes

import numpy as np
import pytest

import boltmi
import cosmo
from boltmi import _abi

pytestmark = pytest.mark.gpu


def test_graph_replays_equal_eager_calls():
    import torch
    par, bg, ih = cosmo.background("default")
    ctx = boltmi.default_context()
    L = _abi.lib()
    dev = torch.device("cuda:0")
    ks = np.logspace(-4, np.log10(0.3), 48)
    dense = boltmi.quadratic_k(0.01 * bg.H0, 1000 * bg.H0, 600)
    ells = np.arange(2, 701, dtype=np.int32)
    cfg = boltmi.hier_cfg(8, 8, 10, 15, 1e-8, 1e-6)
    d_k = torch.tensor(ks, device=dev)
    d_dense = torch.tensor(dense, device=dev)
    with ctx.lock:
        cid = bg.cosmo_id(ctx)

    def eager():
        d_S = torch.zeros((len(ks), bg.nx), dtype=torch.float64, device=dev)
        d_st = torch.zeros(len(ks), dtype=torch.int32, device=dev)
        d_cl = torch.zeros(len(ells), dtype=torch.float64, device=dev)
        with ctx.lock:
            _abi.check(L.bolt_source_grid_dev(ctx.handle, cid, ctypes.byref(cfg), d_k.data_ptr(), len(ks),
                                              d_S.data_ptr(), d_st.data_ptr(), None, None))
            _abi.check(L.bolt_cltt_dev(ctx.handle, cid, d_S.data_ptr(), d_k.data_ptr(), len(ks), d_dense.data_ptr(),
                                       len(dense), _abi.pi(ells), len(ells), d_cl.data_ptr(), None, None))
        return d_S.cpu().numpy(), d_cl.cpu().numpy()

    S_ref, cl_ref = eager()
    d_S = torch.zeros((len(ks), bg.nx), dtype=torch.float64, device=dev)
    d_st = torch.zeros(len(ks), dtype=torch.int32, device=dev)
    d_ctr = torch.zeros((len(ks), 4), dtype=torch.int64, device=dev)
    d_cl = torch.zeros(len(ells), dtype=torch.float64, device=dev)
    g = ctypes.c_void_p()
    with ctx.lock:
        _abi.check(L.bolt_graph_spectrum_create(ctx.handle, cid, ctypes.byref(cfg), d_k.data_ptr(), len(ks),
                                                d_dense.data_ptr(), len(dense), _abi.pi(ells), len(ells),
                                                d_S.data_ptr(), d_st.data_ptr(), d_ctr.data_ptr(), d_cl.data_ptr(),
                                                ctypes.byref(g)))
    try:
        for rep in range(3):
            d_S.fill_(float("nan"))
            d_cl.fill_(float("nan"))
            d_ctr.zero_()
            torch.cuda.synchronize()
            with ctx.lock:
                _abi.check(L.bolt_graph_launch(g))
            assert np.array_equal(d_S.cpu().numpy(), S_ref), rep
            assert np.array_equal(d_cl.cpu().numpy(), cl_ref), rep
            assert int(d_st.max().item()) == 0 and int(d_ctr[:, 0].min().item()) > 0
            # eager work of its own between the replays
            S2, cl2 = eager()
            assert np.array_equal(S2, S_ref) and np.array_equal(cl2, cl_ref)
    finally:
        with ctx.lock:
            L.bolt_graph_destroy(g)
