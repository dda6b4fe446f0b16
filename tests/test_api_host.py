"""Host-side logic of the Python API (no GPU): which uploads a Background
keeps, and when it releases them."""
import ctypes
import gc

import boltmi
from boltmi import api


class _FakeLib:
    """stands in for libbolt_hip's upload/release entry points"""

    def __init__(self):
        self.next = 0
        self.uploads = []
        self.released = []

    def bolt_cosmo_upload(self, handle, par, tables, cid):
        cid._obj.value = self.next
        self.uploads.append((handle.value, self.next))
        self.next += 1
        return 0

    def bolt_cosmo_release(self, handle, cid):
        self.released.append((handle.value, cid))
        return 0


class _FakeCtx:
    def __init__(self, h):
        self.handle = ctypes.c_void_p(h)
        self.serial = next(api._ctx_serial)


def test_uploads_are_keyed_by_context_serial_and_released(monkeypatch):
    bg = boltmi.Background(boltmi.CosmoParams(Σm_ν=0.0), x_grid=(-20.0, 0.0, 201))
    fake = _FakeLib()
    monkeypatch.setattr(api, "lib", lambda: fake)
    monkeypatch.setattr(api, "check", lambda rc: rc)
    a = _FakeCtx(0x1000)
    assert bg.cosmo_id(a) == 0 and bg.cosmo_id(a) == 0       # cached per context
    assert len(fake.uploads) == 1
    # a context that replaces a closed one at the same address is a new context
    a.handle = ctypes.c_void_p(None)
    del a
    gc.collect()
    b = _FakeCtx(0x1000)
    assert bg.cosmo_id(b) == 1 and len(fake.uploads) == 2
    c = _FakeCtx(0x2000)
    assert bg.cosmo_id(c) == 2
    # collecting the Background releases its uploads on the contexts still open
    del bg
    gc.collect()
    assert sorted(fake.released) == [(0x1000, 1), (0x2000, 2)]
