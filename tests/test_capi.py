"""CPU tests of the C-ABI library: it loads, exports every function declared
in include/*.h, and fails loudly (error codes, messages) without a GPU."""
import ctypes
import os
import re

import pytest

import zeta_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"^\s*(?:int|bool|void|uint64_t|double)\s+\**(\w+)\s*\(", src, flags=re.M))
    return names


def test_exports_every_declared_symbol():
    L = zeta_amd.lib()
    names = _declared()
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_calls_before_init_fail_with_state_error():
    L = zeta_amd.lib()
    if zeta_amd._up:
        pytest.skip("library already initialised in this process")
    assert L.zeta_perft_batch(None, 0, 1, None) == -5  # ZETA_ESTATE
    buf = ctypes.create_string_buffer(256)
    L.zeta_last_error(buf, 256)
    assert b"not initialised" in buf.value


def test_init_without_gpu_fails_loudly():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    with pytest.raises(zeta_amd.ZetaError):
        zeta_amd.init()
