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
