"""The solver's per-k-mode LDS workspace (CPU, through bolt_ws_layout): its
regions never overlap, the slot that absorbs stores of rows past n belongs to
none of them, and every branch-free read past n (the substitution scans read
whole 384-row passes, the banded LU's last block up to 8 rows past a segment)
stays inside the bytes a launch reserves.  Covers every state length the
entry points accept with LDS to spare (n <= 515) and nq = 1..48, for the FP64
solver and for the FP32 solve mode (FP32 factors in half the bands' space, a
copy of the stage right-hand side after them)."""
import ctypes

import numpy as np

from boltmi import _abi


def layout(n, nq, fp32=0):
    cap = 128
    reg = (ctypes.c_longlong * (2 * cap))()
    nreg = ctypes.c_int()
    dump, tot, reach = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
    _abi.check(_abi.lib().bolt_ws_layout(n, nq, fp32, cap, reg, ctypes.byref(nreg), ctypes.byref(dump), ctypes.byref(tot),
                                         ctypes.byref(reach)))
    r = np.array(reg[:2 * nreg.value], dtype=np.int64).reshape(-1, 2)
    return r, dump.value, tot.value, reach.value


def test_workspace_regions_are_disjoint_and_reads_stay_inside():
    checked = 0
    for nq in range(1, 49):
        for n, fp32 in [(n, m) for n in list(range(2 * 4 + 3 + 3 + nq * 3 + 3, 516, 7)) + [323, 473, 515]
                        for m in (0, 1)]:
            r, dump, tot, reach = layout(n, nq, fp32)
            order = np.argsort(r[:, 0])
            rs = r[order]
            assert np.all(rs[:, 1] > rs[:, 0])
            assert np.all(rs[1:, 0] >= rs[:-1, 1]), (n, nq)          # pairwise disjoint
            assert rs[0, 0] >= 0 and rs[-1, 1] <= tot, (n, nq)
            inside = (r[:, 0] <= dump) & (dump < r[:, 1])
            assert not inside.any(), (n, nq)                          # nobody reads the dump slot
            assert dump + 8 <= tot
            assert reach <= tot, (n, nq, reach, tot)
            checked += 1
    assert checked > 2000


def test_workspace_of_the_benchmark_hierarchies_fits_lds():
    # n = 323 (l_g = l_nu = 50, l_mnu = 10) and 473 (l_mnu = 20): 104 KB and 147 KB
    for n, kb in ((323, 104), (473, 147)):
        _, _, tot, _ = layout(n, 15)
        assert abs(tot / 1024 - kb) < 1.5 and tot <= 160 * 1024 - 1024
