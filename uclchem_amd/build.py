"""Build libuclchem_gpu.so (HIP, gfx950) in-tree with hipcc, the network files the C ABI
loads directly, and the Fortran binding with its example caller."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libuclchem_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("UCG_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    # keep the compiler from emitting packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


FC = os.environ.get("UCG_FC", "/opt/rocm/bin/amdflang")
FDIR = os.path.join(HERE, "fortran")
FORTRAN_EXAMPLE = os.path.join(FDIR, "cloud_grid")
NET_DIR = os.path.join(HERE, "networks")


def build_fortran(verbose: bool = False) -> str | None:
    """The ISO_C_BINDING module and the example caller (needs a Fortran compiler)."""
    if not os.path.exists(FC):
        if verbose:
            print(f"{FC} not found: Fortran binding not built", flush=True)
        return None
    bdir = os.path.join(FDIR, "_build")
    os.makedirs(bdir, exist_ok=True)
    steps = [
        [FC, "-O2", "-J", bdir, "-c", os.path.join(FDIR, "uclchem_gpu.f90"), "-o", os.path.join(bdir, "uclchem_gpu.o")],
        [FC, "-O2", "-J", bdir, "-c", os.path.join(FDIR, "cloud_grid.f90"), "-o", os.path.join(bdir, "cloud_grid.o")],
        [FC, "-o", FORTRAN_EXAMPLE + ".tmp", os.path.join(bdir, "cloud_grid.o"), os.path.join(bdir, "uclchem_gpu.o"),
         "-L" + HERE, "-luclchem_gpu", "-Wl,-rpath,$ORIGIN/.."],
    ]
    for cmd in steps:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
    os.replace(FORTRAN_EXAMPLE + ".tmp", FORTRAN_EXAMPLE)
    return FORTRAN_EXAMPLE


def write_network_files(verbose: bool = False) -> None:
    """<name>.ucgnet beside every <name>.npz (the file the C ABI loads directly)."""
    from . import makerates_hip as mh
    for f in sorted(os.listdir(NET_DIR)):
        if f.endswith(".npz"):
            out = os.path.join(NET_DIR, f[:-4] + ".ucgnet")
            mh.save_network_bin(mh.load_network(os.path.join(NET_DIR, f)), out)
            if verbose:
                print(f"wrote {out}", flush=True)


def source_hash() -> str:
    """Short digest of the HIP sources (csrc/*.hip, *.h): names the kernel build a
    committed profile was measured on."""
    import hashlib
    h = hashlib.sha256()
    for f in sorted(os.listdir(CSRC)):
        if f.endswith((".hip", ".h")):
            h.update(f.encode())
            with open(os.path.join(CSRC, f), "rb") as fh:
                h.update(fh.read())
    return h.hexdigest()[:16]


def build(verbose: bool = False, extra=None) -> str:
    src = os.path.join(CSRC, "ucg_api.hip")
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + list(extra or []) + ["-o", tmp, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    write_network_files(verbose)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
