"""CPU: the host code of the library (csrc/host.cpp: numbering, patterns,
two-space pattern, Project / lazyProject, Galerkin products, colouring,
sliced-ELL layouts) and the oracle, built with AddressSanitizer and
UndefinedBehaviorSanitizer (afepack_amd/sanitize/Makefile) and run over small
meshes of every element kind: no sanitizer report, no leak, and the drivers'
own structural and convergence checks hold."""
import os
import shutil
import subprocess

import pytest

from tests.conftest import ROOT


def test_host_code_and_oracle_clean_under_asan_ubsan():
    if shutil.which("g++") is None or shutil.which("gcc") is None or shutil.which("make") is None:
        pytest.skip("no host toolchain")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "afepack_amd", "sanitize"), "run"],
                       capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "host_check: ok" in out and "oracle_check: ok" in out
    assert "runtime error" not in out and "AddressSanitizer" not in out and "LeakSanitizer" not in out
