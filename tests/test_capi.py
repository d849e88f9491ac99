"""CPU: the C-ABI library loads, exports every entry point the public header
declares, and reports errors without a device."""
import ctypes
import os
import re

import pytest

import afepack_amd as afe

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(afe_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(afe.LIB_PATH)
    names = _declared()
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    from afepack_amd import _sigs
    assert _declared() <= set(_sigs.SIGS), sorted(_declared() - set(_sigs.SIGS))


def test_errors_without_device():
    lib = afe.load()
    with pytest.raises(afe.AfeError) as e:
        afe.mesh_square(0)
    assert e.value.code == -1
    assert b"bad mesh" in lib.afe_last_error()
    if not os.path.exists("/dev/kfd"):
        with pytest.raises(afe.AfeError) as e:
            afe.Context(0)
        assert e.value.code == -8


def test_tuning_validation():
    for bad in ((3, 1024, 4096), (0, 1001, 4096), (0, 32768, 4096), (0, 1024, -1)):
        with pytest.raises(afe.AfeError):
            afe.set_tuning(*bad)
    afe.set_tuning(2, 16384, 4096)
    afe.set_tuning(0, 2048, 4096)


def test_cpp_dropin_headers_compile(tmp_path):
    """The C++ drop-in headers compile this package's example programs (no
    device needed to build them; they run in the GPU tests)."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    inc = [f"-I{root}/afepack_amd/include", f"-I{root}/include"]
    for prog in ("moving_mesh_cg", "facade_check"):
        src = os.path.join(root, "afepack_amd", "examples", prog + ".cpp")
        r = subprocess.run(["g++", "-std=c++17", "-ffp-contract=off", "-Wall", "-fsyntax-only", *inc, src],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        assert "warning" not in r.stderr, r.stderr[-2000:]


REF_EXAMPLES = "/root/reference/examples"


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="AFEPack source tree not present")
@pytest.mark.parametrize("rel", ["poisson_equation/poisson_equation.cpp",
                                 "elliptic_equation/EllipticEquation.cpp",
                                 "elliptic_equation/main.cpp"])
def test_reference_example_sources_compile_against_dropin(rel):
    """Drop-in check: the reference's own example programs, read in place
    from the AFEPack tree (nothing is copied), compile unchanged against the
    drop-in headers."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = os.path.join(REF_EXAMPLES, rel)
    inc = [f"-I{root}/afepack_amd/include", f"-I{root}/include", f"-I{os.path.dirname(src)}"]
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", *inc, src], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
