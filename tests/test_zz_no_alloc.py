"""GPU test of the allocation-free perft entry point
(zeta_perft_batch_device_noalloc): through the C-ABI, two calls in a row --
the second with a smaller input and a different depth -- give exactly the
counts of the allocating entry point (zeta_perft_batch_device) on the same
device buffers, and sizes beyond its limits are refused."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def noalloc(zeta, d_pos, n, depth, d_cnt, flags=0):
    L = zeta.lib()
    f = L.zeta_perft_batch_device_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    f.restype = ctypes.c_int
    return f(ctypes.c_void_p(d_pos), int(n), int(depth), ctypes.c_void_p(d_cnt), int(flags))


def test_noalloc_equals_allocating_entry_point(zeta):
    pos = zeta.playouts(2, 4096, 8, 40)
    dpos = zeta.DeviceBuffer(pos.nbytes)
    dpos.upload(pos)
    dref = zeta.DeviceBuffer(8 * len(pos))
    dna = zeta.DeviceBuffer(8 * len(pos))
    golden = np.load(os.path.join(HERE, "golden", "c2_perft4.npz"))["perft4"]
    junk = np.full(len(pos), 0xDEADBEEFDEADBEEF, dtype=np.uint64)
    # call 1: the whole 4096-root perft(4) workload (levels larger than the
    # scratch array's budget are walked in chunks); call 2: fewer roots, other
    # depth, over the previous call's scratch contents; call 3: no bulk counting
    for n, depth, flags in ((4096, 4, 0), (300, 3, 0), (64, 3, zeta.ZETA_PERFT_NOBULK)):
        print("no-alloc perft: roots", n, "depth", depth, "flags", flags, flush=True)
        dna.upload(junk)
        assert noalloc(zeta, dpos.ptr, n, depth, dna.ptr, flags) == 0, zeta.last_error()
        got = dna.download(np.zeros(n, dtype=np.uint64))
        zeta.perft_device(dpos.ptr, n, depth, dref.ptr, nobulk=bool(flags))
        want = dref.download(np.zeros(n, dtype=np.uint64))
        assert (got == want).all()
        if (n, depth, flags) == (4096, 4, 0):
            assert (got == golden).all()
    for a in (dpos, dref, dna):
        a.free()


def test_noalloc_refuses_sizes_beyond_its_limits(zeta):
    pos = zeta.playouts(2, 16, 8, 40)
    dpos = zeta.DeviceBuffer(pos.nbytes)
    dpos.upload(pos)
    dcnt = zeta.DeviceBuffer(8 * 16)
    assert noalloc(zeta, dpos.ptr, 65537, 3, dcnt.ptr) == -1
    assert "at most 65536" in zeta.last_error()
    assert noalloc(zeta, dpos.ptr, 16, 9, dcnt.ptr) == -1
    assert "at most 8" in zeta.last_error()
    assert noalloc(zeta, dpos.ptr, 16, 2, dcnt.ptr) == 0
    dpos.free()
    dcnt.free()
