"""CPU tests of the C-ABI library: it loads, exports every function declared
in include/*.h, and fails loudly (error codes, messages) without a GPU."""
import ctypes
import os
import re
import subprocess

import pytest

import zeta_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return set(re.findall(r"^\s*(?:int|bool|void|uint64_t|double)\s+\**(\w+)\s*\(", src, flags=re.M))


def test_every_header_is_covered():
    assert sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")) == \
        ["zeta_clrun.h", "zeta_hip.h"]


def test_exports_every_declared_symbol():
    L = zeta_amd.lib()
    names = _declared("zeta_hip.h")
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_clrun_boundary_is_defined_in_the_front_end():
    # the six reference-named functions live in the shim linked into the
    # front end (they read the host program's globals), not in the library
    exe = os.path.join(ROOT, "zeta_amd", "bin", "zeta")
    if not os.path.exists(exe):
        pytest.fail("zeta_amd/bin/zeta not built")
    out = subprocess.run(["nm", exe], capture_output=True, text=True, check=True).stdout
    defined = {line.split()[-1] for line in out.splitlines() if len(line.split()) == 3 and line.split()[1] in "TDBR"}
    names = _declared("zeta_clrun.h")
    assert len(names) == 6
    assert names <= defined, names - defined
    for g in ("GLOBAL_BOARD", "COUNTERS", "PV", "GLOBAL_HASHHISTORY", "totalWorkUnits", "tt1_memory"):
        assert g in defined, g


def test_front_end_runs_without_a_gpu_and_fails_loudly():
    exe = os.path.join(ROOT, "zeta_amd", "bin", "zeta")
    p = subprocess.run([exe, "--version"], capture_output=True, text=True, timeout=30)
    assert p.returncode == 0 and "Zeta version" in p.stdout
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU runtime is present")
    p = subprocess.run([exe, "--config", os.path.join(ROOT, "zeta_amd", "host", "config.txt")], input="quit\n",
                       capture_output=True, text=True, timeout=60)
    assert p.returncode != 0
    assert "cl_init_device failed" in p.stderr


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
