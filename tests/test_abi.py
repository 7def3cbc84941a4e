"""The C ABI boundary: the library loads, exports every declared symbol, and
the GpuVTable provider behaves like NumPP's accessor (null without a device)."""
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "numpp_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(npph_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for need in ("npph_gemm", "npph_gesv_batched", "npph_gesv_refined_batched_dev", "npph_gpu_vtable"):
        assert need in syms


def test_library_exports_every_declared_symbol():
    from numpp_amd import _lib
    lib = _lib.load()
    for s in declared_symbols():
        assert hasattr(lib, s), s
    assert set(declared_symbols()) == set(_lib.SIGNATURES), "ctypes signatures out of sync with the header"


def test_exported_dynamic_symbols():
    from numpp_amd import _lib
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (npph_\w+)", out))
    assert set(declared_symbols()) <= exported
    assert "_ZN5numpp10gpu_vtableEv" in out  # numpp::gpu_vtable()


def test_no_device_contract():
    import numpp_amd
    from numpp_amd import _lib, vtable
    lib = _lib.load()
    if lib.npph_available():
        pytest.skip("a device is present: covered by the gpu tests")
    assert vtable.get() is None
    assert lib.npph_init() == _lib.NPPH_ERR_NO_DEVICE
    assert lib.npph_gemm(11, 1, 1, 1, None, None, None) == _lib.NPPH_ERR_NO_DEVICE
    assert "device" in _lib.last_error()
    with pytest.raises(numpp_amd.BackendUnavailable):
        numpp_amd.require_device()


def _build_cpp_check(tmp_path):
    from numpp_amd import _lib
    exe = str(tmp_path / "vtable_check")
    cmd = ["g++", "-std=c++20", "-O1", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "vtable_check.cpp"), _lib.LIB_PATH,
           "-Wl,-rpath," + os.path.dirname(_lib.LIB_PATH), "-o", exe]
    subprocess.run(cmd, check=True)
    return exe


def test_cpp_vtable_drop_in_no_device(tmp_path):
    from numpp_amd import _lib
    if _lib.load().npph_available():
        pytest.skip("device present")
    exe = _build_cpp_check(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "null" in r.stdout


@pytest.mark.gpu
def test_cpp_vtable_drop_in_device(tmp_path, hip):
    exe = _build_cpp_check(tmp_path)
    r = subprocess.run([exe, "--device"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "slots ok" in r.stdout
