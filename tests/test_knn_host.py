"""The library's kNN search (avism_amd/csrc/knn_search.h, the code the HIP
kernels run) built for the host and checked against brute force, including
the tie rule (equal distances go to the lower input index), periodic images
and boxes smaller than the search."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


def test_knn_search_matches_brute_force(tmp_path):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = tmp_path / "knn_check"
    r = subprocess.run([cxx, "-O2", "-std=c++17", "-ffp-contract=off", str(HERE / "knn_host_check.cpp"),
                        "-o", str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
