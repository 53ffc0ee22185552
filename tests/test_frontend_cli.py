"""CPU tests of the engine's command line (zeta_amd/bin/zeta): the
reference's options as its getopt_long_only reads them (src/zeta.c:529-541,
one or two dashes, unique prefixes).  Every case here exits before the
engine touches a device."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENGINE = os.path.join(ROOT, "zeta_amd", "bin", "zeta")


def run(*args):
    if not os.path.exists(ENGINE):
        pytest.skip("zeta_amd/bin/zeta not built")
    return subprocess.run([ENGINE, *args], capture_output=True, text=True, timeout=60, stdin=subprocess.DEVNULL)


@pytest.mark.parametrize("flag", ["-v", "-version", "--version", "--vers"])
def test_version(flag):
    p = run(flag)
    assert p.returncode == 0 and p.stdout.startswith("Zeta version: ")


@pytest.mark.parametrize("flag", ["-h", "-help", "--help", "--he"])
def test_help(flag):
    p = run(flag)
    assert p.returncode == 0
    assert "--guessconfig" in p.stdout and "--config FILE" in p.stdout


@pytest.mark.parametrize("flag", ["-pl", "-dl", "--dl"])
def test_opencl_listings_name_the_hip_device(flag):
    # the OpenCL platform / device listings have no counterpart on HIP: the
    # engine says so and names the device -d picks
    p = run("-d", "3", flag)
    assert p.returncode == 0 and "device 3" in p.stdout


@pytest.mark.parametrize("args", [("-bogus",), ("-g",), ("-d",)])
def test_unknown_ambiguous_or_incomplete_options_print_the_help(args):
    # -g is a prefix of both guessconfig options; -d needs its argument
    p = run(*args)
    assert p.returncode == 1
    assert "Options:" in p.stdout
