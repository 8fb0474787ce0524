"""The super-k-mer cut of the binned Stage 2 count, built for the host from the same
header the device kernels include (spades-rs_amd/csrc/kc_superkmer.h): records tile the
windows, bins match a direct recomputation, and a k-mer and its reverse complement share a
bin (tests/superkmer_check.cpp)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT


def test_superkmer_cut_on_host(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not present")
    exe = tmp_path / "superkmer_check"
    subprocess.run([hipcc, "-O2", "-std=c++17", "-o", str(exe), os.path.join(ROOT, "tests", "superkmer_check.cpp")],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "failures 0" in r.stdout
