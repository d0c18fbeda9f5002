"""CPU checks of the engine's host/device arithmetic helpers and of the C ABI
library (no GPU needed)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NATIVE = os.path.join(ROOT, "tests", "native")
HEADER = os.path.join(ROOT, "include", "chain_engine.h")
LIB = os.path.join(ROOT, "chaincompressor_amd", "libchain_engine.so")


def _build(src, tmp_path):
    exe = str(tmp_path / os.path.splitext(os.path.basename(src))[0])
    cmd = ["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", src, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_rcp31_division_exact(tmp_path):
    """floor(x / f) by multiplication, every quotient boundary, f <= 2^20."""
    exe = _build(os.path.join(NATIVE, "rcp31_check.cpp"), tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    assert "bad=0" in r.stdout


def test_rvn1_integer_rounding_matches_long_double(tmp_path):
    """The integer restatement of the reference's x87 long-double roundings
    (cpp/chain/lcps_cpp.cpp:596-641) against real long double arithmetic."""
    exe = _build(os.path.join(NATIVE, "rvn1_reduce_check.cpp"), tmp_path)
    r = subprocess.run([exe, "2000000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s+(chain_[a-z_0-9]+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip("libchain_engine.so not built")
    names = declared_functions()
    assert "chain_compress" in names and "chain_decompress" in names
    lib = ctypes.CDLL(LIB)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    from chaincompressor_amd import engine
    assert sorted(engine.EXPORTS) == names


def test_pybind_modules_expose_reference_names():
    import importlib
    try:
        comp = importlib.import_module("chaincompressor_amd.lcps_gpu_cpp")
        dec = importlib.import_module("chaincompressor_amd.lcps_gpu_decompress_ans")
    except ImportError:
        pytest.skip("pybind11 modules not built")
    assert hasattr(comp, "lcps_compress2_gpu") and hasattr(comp, "lcps_compress2")
    assert hasattr(dec, "lcps_decompress_gpu")
