"""CPU checks that the C ABI validates every index table of a network before anything is
uploaded: a network file with one table corrupted (an index past its range, a pointer array
that is not monotone or does not end at its list's length) is refused with UCG_E_ARG and a
message naming the table; the intact file passes the checks (and then needs a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "uclchem_gpu.h")
LIB = os.path.join(ROOT, "uclchem_amd", "libuclchem_gpu.so")
CODES = {k: int(v) for k, v in re.findall(r"#define (UCG_E_\w+) (-?\d+)", open(HEADER).read())}


def _set(arrays, name, idx, value):
    a = arrays[name].copy()
    a.reshape(-1)[idx] = value
    arrays[name] = a


def corruptions(t):
    """(label, function editing (arrays, special_species, special_reactions)) pairs."""
    S, R, Q = t.nspec, t.nreac, t.neq
    a = t.arrays
    fr = int(a["type_ranges"][3, 0])          # first freeze-out reaction (T_FREEZE = 3)
    return [
        ("re past nspec", lambda A, s, r: _set(A, "re", 5, S + 3)),
        ("re below -1", lambda A, s, r: _set(A, "re", 7, -2)),
        ("pr past nspec", lambda A, s, r: _set(A, "pr", 11, S)),
        ("freeze-out without reactant", lambda A, s, r: _set(A, "re", 3 * fr, -1)),
        ("rtype", lambda A, s, r: _set(A, "rtype", 0, 99)),
        ("ice_kind", lambda A, s, r: _set(A, "ice_kind", 0, 5)),
        ("ops", lambda A, s, r: _set(A, "ops", 2, S + 1)),
        ("type_ranges reversed", lambda A, s, r: _set(A, "type_ranges", 1, int(A["type_ranges"].reshape(-1)[0]) - 1)),
        ("type_ranges past nreac", lambda A, s, r: _set(A, "type_ranges", 2 * 22 + 1, R)),
        ("surface_list", lambda A, s, r: _set(A, "surface_list", 0, -1)),
        ("bulk_list", lambda A, s, r: _set(A, "bulk_list", 0, S)),
        ("bulk_partner_of_surface", lambda A, s, r: _set(A, "bulk_partner_of_surface", int(a["surface_list"][0]), S + 2)),
        ("surface_partner_of_bulk", lambda A, s, r: _set(A, "surface_partner_of_bulk", int(a["bulk_list"][0]), -7)),
        ("freeze_partners", lambda A, s, r: _set(A, "freeze_partners", 0, R)),
        ("swap_list", lambda A, s, r: _set(A, "swap_list", 0, R + 4)),
        ("is_surface", lambda A, s, r: _set(A, "is_surface", 0, 3)),
        ("jnz_row", lambda A, s, r: _set(A, "jnz_row", 9, Q)),
        ("jnz_col", lambda A, s, r: _set(A, "jnz_col", 9, -1)),
        ("jnz_ptr not monotone", lambda A, s, r: _set(A, "jnz_ptr", 5, int(A["jnz_ptr"][7]) + 1)),
        ("aff_ptr end", lambda A, s, r: _set(A, "aff_ptr", R, int(A["aff_ptr"][R]) + 1)),
        ("aff_sp", lambda A, s, r: _set(A, "aff_sp", 4, Q)),
        ("sw_ptr start", lambda A, s, r: _set(A, "sw_ptr", 0, 1)),
        ("loss_idx", lambda A, s, r: _set(A, "loss_idx", 3, R)),
        ("prod_ptr end", lambda A, s, r: _set(A, "prod_ptr", S, int(A["prod_ptr"][S]) - 1)),
        ("special species", lambda A, s, r: s.__setitem__(2, Q)),
        ("special reaction", lambda A, s, r: r.__setitem__(6, R)),
        ("plan permutation", lambda A, s, r: _set(A, "nw_perm", 0, Q + 5)),
        ("plan assembly slot", lambda A, s, r: _set(A, "nw_jnz_slot", 0, -3)),
    ]


@pytest.fixture(scope="module")
def clib():
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    L = ctypes.CDLL(LIB)
    L.ucg_last_error.restype = ctypes.c_char_p
    return L


def test_corrupted_tables_are_refused(tables, clib, tmp_path, monkeypatch):
    from uclchem_amd import makerates_hip as mh
    real = mh.device_descriptor
    path = str(tmp_path / "net.ucgnet")
    h = ctypes.c_int(-1)
    mh.save_network_bin(tables, path)
    rc = clib.ucg_network_load_file(path.encode(), ctypes.byref(h))
    assert rc in (0, CODES["UCG_E_NODEVICE"]), clib.ucg_last_error()    # intact: checks pass
    if rc == 0:
        assert clib.ucg_network_free(h.value) == 0
    for label, edit in corruptions(tables):
        def bad_descriptor(t, edit=edit):
            sizes, arrays, ssp, srx = real(t)
            arrays = dict(arrays)
            ssp, srx = ssp.copy(), srx.copy()
            edit(arrays, ssp, srx)
            return sizes, arrays, ssp, srx
        monkeypatch.setattr(mh, "device_descriptor", bad_descriptor)
        mh.save_network_bin(tables, path)
        rc = clib.ucg_network_load_file(path.encode(), ctypes.byref(h))
        assert rc == CODES["UCG_E_ARG"], (label, rc, clib.ucg_last_error())
        assert clib.ucg_last_error(), label
    monkeypatch.setattr(mh, "device_descriptor", real)
