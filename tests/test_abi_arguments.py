"""The C-ABI entry points of the device tables, the Bessel moments and the
queue server reject NULL handles and bad sizes with an error code before they
touch a device, so these run on a machine without a GPU."""
import ctypes

import pytest

from boltmi import _abi

NULL = None


@pytest.fixture(scope="module")
def L():
    return _abi.lib()


def test_moment_tables_reject_null_handles(L):
    h = ctypes.c_void_p()
    assert L.bolt_moment_table_create(NULL, 3, 3, 0.0, 10.0, 100, ctypes.byref(h)) == _abi.BOLT_E_ARG
    assert L.bolt_moment_table_eval(NULL, NULL, 1, NULL) == _abi.BOLT_E_ARG
    assert L.bolt_moment_table_coefficients(NULL, NULL) == _abi.BOLT_E_ARG
    assert L.bolt_filon(NULL, NULL, NULL, NULL, NULL, NULL, NULL, 1, NULL) == _abi.BOLT_E_ARG
    assert L.bolt_filon_sum(NULL, NULL, NULL, NULL, NULL, 2, NULL, 1, NULL) == _abi.BOLT_E_ARG
    assert L.bolt_sph_moments(NULL, 0, NULL, NULL, NULL, 1, NULL) == _abi.BOLT_E_ARG
    assert L.bolt_last_error().startswith(b"bolt_sph_moments")


def test_device_tables_reject_null_context(L):
    out = (ctypes.c_double * 4)()
    st = (ctypes.c_int * 1)()
    pars = (_abi.CosmoParamsC * 1)()
    assert L.bolt_device_tables_batch(NULL, pars, 1, -20.0, 0.0, 2001, 15, out, out, out, out, st) == _abi.BOLT_E_ARG


def test_queue_server_and_batched_grid_reject_null_context(L):
    assert L.bolt_serve_start_queue(NULL, 1.0, 0) == _abi.BOLT_E_STATE
    info = _abi.QueueInfoC()
    assert L.bolt_serve_queue_info(NULL, ctypes.byref(info)) == _abi.BOLT_E_ARG
    ids = (ctypes.c_int * 1)()
    cfgs = (_abi.HierCfgC * 1)()
    ms = ctypes.c_float()
    assert L.bolt_source_grid_batch_dev(NULL, 1, ids, cfgs, NULL, 1, NULL, NULL, NULL, NULL,
                                        ctypes.byref(ms)) == _abi.BOLT_E_STATE
