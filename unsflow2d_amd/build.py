"""In-tree build of the native parts.

* ``libufx.so`` -- the HIP kernels for gfx950 and the C-ABI (include/ufx.h),
  built with hipcc from ``csrc/``;
* ``fortran/`` -- the ISO_C_BINDING interface module and the Fortran host
  driver, built with amdflang against ``libufx.so`` (skipped when amdflang is
  absent).

Run ``python -m unsflow2d_amd.build``; the outputs are written next to their
sources so they travel with the tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libufx.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 arithmetic off: some of its forms have produced silently
    # wrong results on MI355X
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, cwd=cwd, check=True)


def _gpu_targets(path: str) -> set[str]:
    """Offload target ids embedded in a built object or library."""
    import re
    with open(path, "rb") as f:
        data = f.read()
    return {m.decode() for m in re.findall(rb"amdgcn-amd-amdhsa--(gfx[0-9a-z]+)", data)}


def build_lib() -> str:
    """Compile the kernels (hipcc, gfx950) and the host mesh code (g++, no
    floating-point contraction so it reproduces the reference bit for bit),
    then link libufx.so."""
    bdir = os.path.join(PKG, "_build")
    os.makedirs(bdir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include")]
    objs = []
    # the vertex-centred kernels and the device preprocessing evaluate every
    # expression as the reference does (no a*b+c contraction), so their FP64
    # results follow it bit for bit
    nofma = ["-ffp-contract=off"]
    extra = {"nc_solver.hip": nofma, "cc_mesh_dev.hip": nofma, "nc_mesh_dev.hip": nofma}
    for src in ("cc_solver.hip", "nc_solver.hip", "cc_mesh_dev.hip", "nc_mesh_dev.hip"):
        obj = os.path.join(bdir, src + ".o")
        _run([HIPCC, *HIP_FLAGS, *extra.get(src, []), *inc, "-c", os.path.join(CSRC, src), "-o",
              obj])
        tg = _gpu_targets(obj)
        if tg != {ARCH}:
            raise RuntimeError(f"{src}: expected code for {ARCH} only, found {sorted(tg)}")
        objs.append(obj)
    for src in ("cc_mesh.cpp", "nc_mesh.cpp", "cc_mesh_api.cpp"):
        obj = os.path.join(bdir, src + ".o")
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
              "-Wall", *inc, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_fortran() -> str | None:
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    if fc is None:
        print("amdflang not found: Fortran host not built", flush=True)
        return None
    fdir = os.path.join(PKG, "fortran")
    srcs = ["ufx_api.f90", "ufx_cc.f90", "flo2d_gpu.f90", "ufx_nc.f90", "nc2d_gpu.f90"]
    if not all(os.path.exists(os.path.join(fdir, s)) for s in srcs):
        return None
    for s in srcs:
        _run([fc, "-O2", "-ffree-form", "-module-dir", fdir, "-c", os.path.join(fdir, s),
              "-o", os.path.join(fdir, s.replace(".f90", ".o"))])
    obj = lambda s: os.path.join(fdir, s.replace(".f90", ".o"))
    link = ["-L", PKG, "-lufx", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,/opt/rocm/lib"]
    # the cell-centred host (the reference's CCFVM main) and the vertex-centred one (NCFVM main)
    _run([fc, "-o", os.path.join(fdir, "flo2d_gpu"), obj("ufx_api.f90"), obj("ufx_cc.f90"),
          obj("flo2d_gpu.f90"), *link])
    out = os.path.join(fdir, "nc2d_gpu")
    _run([fc, "-o", out, obj("ufx_api.f90"), obj("ufx_nc.f90"), obj("nc2d_gpu.f90"), *link])
    return out


def build_all() -> None:
    build_lib()
    build_fortran()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what:
        build_all()
    if "lib" in what:
        build_lib()
    if "fortran" in what:
        build_fortran()
