"""The flat network file (<name>.ucgnet) that the C ABI loads without Python.

CPU: an independent reader of the documented layout recovers every descriptor array
exactly, and the library's parser rejects damaged files with an error (parsing happens
before any device call).  GPU: a network loaded from the file behaves bit-identically
to the same tables uploaded from Python, and the Fortran example caller (ISO_C_BINDING)
returns bit-identical results."""
import ctypes
import os
import struct
import subprocess

import numpy as np
import pytest

from uclchem_amd import makerates_hip as mh

DT = {0: np.float64, 1: np.int32, 2: np.int8}


def read_ucgnet(path):
    raw = open(path, "rb").read()
    assert raw[:8] == b"UCGNET01"
    n = struct.unpack_from("<i", raw, 8)[0]
    off, out = 12, {}
    for _ in range(n):
        name = raw[off:off + 24].rstrip(b"\0").decode()
        dt, cnt = struct.unpack_from("<iq", raw, off + 24)
        off += 36
        a = np.frombuffer(raw, dtype=np.dtype(DT[dt]).newbyteorder("<"), count=cnt, offset=off)
        off += a.nbytes + (-a.nbytes) % 8
        out[name] = a
    assert off == len(raw)
    return out


@pytest.fixture(scope="module")
def netfile(tables, tmp_path_factory):
    p = str(tmp_path_factory.mktemp("net") / "default.ucgnet")
    mh.save_network_bin(tables, p)
    return p


def test_file_round_trip(tables, netfile):
    recs = read_ucgnet(netfile)
    sizes, arrays, ssp, srx = mh.device_descriptor(tables)
    assert list(recs["sizes"]) == [sizes[k] for k in mh.DESC_SIZES]
    assert np.array_equal(recs["special_species"], ssp)
    assert np.array_equal(recs["special_reactions"], srx)
    for k, v in arrays.items():
        assert recs[k].dtype == v.dtype and np.array_equal(recs[k], v.ravel()), k
    # partner expansion: surface species s has bulk partner bps[s], and back
    a = tables.arrays
    bps, spb = arrays["bulk_partner_of_surface"], arrays["surface_partner_of_bulk"]
    for s in a["surface_list"]:
        assert spb[bps[s]] == s


def _lib_or_skip():
    from uclchem_amd import gpu
    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("HIP library not built")
    return gpu.lib()


@pytest.mark.parametrize("damage", ["magic", "truncate", "missing"])
def test_parser_rejects_damaged_files(netfile, tmp_path, damage):
    L = _lib_or_skip()
    raw = bytearray(open(netfile, "rb").read())
    if damage == "magic":
        raw[:8] = b"NOTANET0"
    elif damage == "truncate":
        raw = raw[:len(raw) // 2]
    else:   # rename the 'alpha' record
        i = raw.index(b"alpha\0")
        raw[i:i + 5] = b"alphx"
    p = tmp_path / "bad.ucgnet"
    p.write_bytes(bytes(raw))
    h = ctypes.c_int(-1)
    rc = L.ucg_network_load_file(str(p).encode(), ctypes.byref(h))
    assert rc == -102, rc          # UCG_E_ARG, raised before any device call
    msg = L.ucg_last_error().decode()
    assert ("UCGNET01" in msg) or ("truncated" in msg) or ("lacks record alpha" in msg) or ("bad record" in msg), msg


@pytest.mark.gpu
def test_file_loaded_network_is_identical(tables, netfile, gpu_net):
    from uclchem_amd import gpu, workloads
    fnet = gpu.Network(tables, from_file=netfile)
    params = workloads.static_cloud_grid(4)
    a = gpu_net.solve(params)
    b = fnet.solve(params)
    assert np.array_equal(a["y"], b["y"]) and np.array_equal(a["stats"], b["stats"])


@pytest.mark.gpu
def test_fortran_caller(tables, netfile, gpu_net, tmp_path):
    from uclchem_amd import build
    from uclchem_amd.params import make_params
    if not os.path.exists(build.FORTRAN_EXAMPLE):
        pytest.skip("Fortran example not built (no Fortran compiler at build time)")
    out = tmp_path / "abund.bin"
    npts = 3
    r = subprocess.run([build.FORTRAN_EXAMPLE, netfile, str(npts), "64", str(out)], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    y_f = np.fromfile(out, dtype="<f8").reshape(npts, tables.neq)
    params = np.stack([make_params(initialdens=10 ** (2 + 5 * q / (npts - 1)), finaltime=1e6) for q in range(npts)])
    y_p = gpu_net.solve(params)["y"]
    assert np.array_equal(y_f, y_p)


def test_fortran_binding_covers_the_c_abi(tmp_path):
    """Every batch, control and request-queue entry of include/uclchem_gpu.h has a Fortran
    interface in uclchem_amd/fortran/uclchem_gpu.f90, and a caller using them compiles
    (compile only: nothing runs without a GPU)."""
    fc = "/opt/rocm/bin/amdflang"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    bdir = os.path.join(root, "uclchem_amd", "fortran", "_build")
    if not (os.path.exists(fc) and os.path.exists(os.path.join(bdir, "uclchem_gpu.mod"))):
        pytest.skip("Fortran binding not built")
    src = open(os.path.join(root, "uclchem_amd", "fortran", "uclchem_gpu.f90")).read()
    for name in ["ucg_init", "ucg_network_load_file", "ucg_network_info", "ucg_solve_batch", "ucg_cancel",
                 "ucg_queue_slot_bytes", "ucg_queue_open", "ucg_queue_submit", "ucg_queue_test",
                 "ucg_queue_result", "ucg_queue_close", "ucg_last_error"]:
        assert f'name="{name}"' in src, name
    prog = tmp_path / "queue_caller.f90"
    prog.write_text("""
PROGRAM queue_caller
    USE uclchem_gpu
    USE, INTRINSIC :: ISO_C_BINDING
    IMPLICIT NONE
    INTEGER(C_INT) :: rc, q, r, done, st, stats(8)
    INTEGER(C_SIZE_T) :: bytes
    REAL(C_DOUBLE) :: params(UCG_NPARAM), y(336)
    INTEGER :: handle
    rc = ucg_network_load_file_f('default.ucgnet', handle)
    CALL ucg_default_params(params)
    rc = ucg_queue_slot_bytes(handle, 64, bytes)
    rc = ucg_queue_open(handle, 16, 4, 64, 60.0d0, q)
    rc = ucg_queue_submit(q, params, r)
    rc = ucg_queue_test(q, r, done)
    rc = ucg_queue_result(q, r, y, st, stats)
    rc = ucg_queue_close(q)
    rc = ucg_cancel(handle)
END PROGRAM queue_caller
""")
    subprocess.check_call([fc, "-c", "-J", bdir, str(prog), "-o", str(tmp_path / "queue_caller.o")])
