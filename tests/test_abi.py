"""CPU checks of the C ABI: the built library exports every symbol include/uclchem_gpu.h declares."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "uclchem_gpu.h")
LIB = os.path.join(ROOT, "uclchem_amd", "libuclchem_gpu.so")


def declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(ucg_\w+)\s*\(", src, re.M)))


def test_header_declares_entry_points():
    names = declared()
    for n in ["ucg_init", "ucg_network_load", "ucg_solve_batch", "ucg_solve_batch_device", "ucg_rates",
              "ucg_rhs", "ucg_jacobian", "ucg_newton_solve", "ucg_last_error"]:
        assert n in names


def test_library_exports_all_declared_symbols():
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(ln.split()[-1] for ln in out.splitlines() if ln.strip())
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    L = ctypes.CDLL(LIB)           # loads without a GPU; no compute call is made here
    for n in declared():
        getattr(L, n)


def test_no_scalar_stores_in_device_code():
    """The device code has no scalar-cache stores or atomics (assembly of our own build)."""
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc missing")
    src = os.path.join(ROOT, "uclchem_amd", "csrc", "ucg_api.hip")
    asm = subprocess.run([hipcc, "-O3", "-std=c++17", "--offload-arch=gfx950", "--cuda-device-only", "-S",
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-o", "-", src],
                         capture_output=True, text=True)
    assert asm.returncode == 0, asm.stderr[-2000:]
    body = [ln.split(";")[0] for ln in asm.stdout.splitlines()]
    text = "\n".join(body)
    for bad in ("s_store_", "s_atomic_", "s_dcache_wb", "s_buffer_store", "s_scratch_store"):
        assert bad not in text, bad
    assert "v_pk_fma_f32" not in text and "v_pk_mul_f32" not in text and "v_pk_add_f32" not in text


def test_error_contract_without_compute():
    """Error paths of the C ABI need no GPU: negative UCG_E_* codes with a message in
    ucg_last_error() for a missing device, a missing network file, and bad network or queue
    handles (no kernel is launched by any of these calls)."""
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    codes = dict(re.findall(r"#define (UCG_E_\w+) (-?\d+)", open(HEADER).read()))
    L = ctypes.CDLL(LIB)
    L.ucg_last_error.restype = ctypes.c_char_p
    rc = L.ucg_init(0)
    assert rc in (0, int(codes["UCG_E_NODEVICE"]))
    if rc:
        assert b"device" in L.ucg_last_error()
    h = ctypes.c_int(-1)
    assert L.ucg_network_load_file(b"/nonexistent/net.ucgnet", ctypes.byref(h)) == int(codes["UCG_E_ARG"])
    assert b"nonexistent" in L.ucg_last_error()
    bad = int(codes["UCG_E_HANDLE"])
    assert L.ucg_solve_batch(77, 1, None, None, None, 0, None, None, None, None, 64) == bad
    assert L.ucg_cancel(77) == bad and L.ucg_network_free(77) == bad
    assert L.ucg_set_linear_solver(77, 0) == bad
    for f, args in ((L.ucg_queue_submit, (77, None, None)), (L.ucg_queue_test, (77, 0, None)),
                    (L.ucg_queue_close, (77,))):
        assert f(*args) == bad
        assert b"queue" in L.ucg_last_error()
