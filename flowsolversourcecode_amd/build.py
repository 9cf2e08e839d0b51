"""Compile the HIP/C++ library for gfx950: flowsolversourcecode_amd/libfsamd.so,
and, where amdflang is installed, the Fortran host of the device-resident
SIMPLE iteration (flowsolversourcecode_amd/fortran/_build/simple_case_device.x).

    python -m flowsolversourcecode_amd.build
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libfsamd.so")
FORTRAN = os.path.join(PKG, "fortran")
FORTRAN_HOST = os.path.join(FORTRAN, "_build", "simple_case_device.x")
SOURCES = ["fs_api.hip", "fs_solve.hip", "fs_amg.hip", "fs_mesh.hip", "fs_assemble.hip", "fs_simple.hip", "fs_probe.hip", "fs_mpnn.hip", "fs_polymesh.hip", "fs_periodic.hip", "fs_turb.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    # the device path must reproduce the reference's per-cell arithmetic:
    # no contraction of a*b+c into one rounding
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = []
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    for src in srcs:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        cmd = [HIPCC, *[f for f in FLAGS if f != "-shared"], "-c", src, "-o", obj,
               "-I", os.path.join(ROOT, "include"), "-I", CSRC]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    for src, p in procs:
        out = p.communicate()[0].decode()
        if p.returncode != 0:
            sys.stderr.write(out)
            raise RuntimeError(f"hipcc failed on {src}")
        if verbose and out.strip():
            sys.stderr.write(out)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    build_fortran_host()
    return LIB


def build_fortran_host():
    """fortran/SimpleCaseDevice.f90 with the interface modules, linked to
    libfsamd.so through a run path relative to the executable.  Returns the
    executable's path, or None where amdflang is absent."""
    fc = shutil.which("amdflang") or (os.path.exists("/opt/rocm/bin/amdflang") and "/opt/rocm/bin/amdflang")
    if not fc:
        return None
    out = os.path.join(FORTRAN, "_build")
    os.makedirs(out, exist_ok=True)
    objs = []
    for src in ("fsamd_iface.f90", "fsamd_case_iface.f90", "SimpleCaseDevice.f90"):
        obj = os.path.join(out, src + ".o")
        subprocess.run([fc, "-O2", "-module-dir", out, "-c", os.path.join(FORTRAN, src), "-o", obj], check=True)
        objs.append(obj)
    subprocess.run([fc, "-o", FORTRAN_HOST, *objs, "-L", PKG, "-lfsamd", "-Wl,-rpath,$ORIGIN/../.."], check=True)
    return FORTRAN_HOST


if __name__ == "__main__":
    print(build(verbose=True))
