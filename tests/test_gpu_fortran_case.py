"""The Fortran host of the device-resident SIMPLE iteration
(flowsolversourcecode_amd/fortran/SimpleCaseDevice.f90 over the
ISO_C_BINDING modules fsamd and fsamd_case): from a polyMesh directory on
disk, 8 outer iterations with the reference's ILU for momentum and p' give
the velocities, pressure and mass fluxes of the reference's own run bitwise,
and its restart file byte for byte (tests/golden/simple_hex_ilu.npz, written
by the reference's SolveMomentumX/Y/Z, SolveContinuity and writeSolution),
also when it resumes from the reference's restart file of 4 iterations."""
import importlib.util
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from conftest import golden
from flowsolversourcecode_amd import build as fb
from flowsolversourcecode_amd import meshgen

pytestmark = pytest.mark.skipif(not os.path.exists(fb.FORTRAN_HOST), reason="Fortran host not built (needs amdflang)")

_spec = importlib.util.spec_from_file_location(
    "make_simple_golden", os.path.join(os.path.dirname(__file__), "golden", "make_simple_golden.py"))
msg = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(msg)


@pytest.mark.gpu
def test_fortran_host_runs_the_reference_iterations_bitwise():
    g = golden("simple_hex_ilu.npz")
    spec, kw = msg.SIMPLE["hex_ilu"]
    pm, _, _, _, _ = msg.build_mesh(spec)
    with tempfile.TemporaryDirectory() as td:
        mesh_dir = os.path.join(td, "polyMesh")
        meshgen.write_polymesh(pm, mesh_dir)
        out = os.path.join(td, "run")
        print("fortran host: hex_ilu", kw["niter"], "iterations", flush=True)
        r = subprocess.run([fb.FORTRAN_HOST, mesh_dir, "ymax", str(kw["niter"]), out],
                           capture_output=True, text=True, timeout=120)
        sys.stdout.write(r.stdout + r.stderr)
        assert r.returncode == 0
        raw = np.fromfile(out + ".fields", dtype=np.float64)
        restart = open(out + ".restart", "rb").read()
    n, fi = len(g["u"]), len(g["mdot"])
    assert raw.size == 4 * n + fi
    for k, name in enumerate(("u", "v", "w", "p")):
        assert np.array_equal(raw[k * n:(k + 1) * n], g[name]), name
    assert np.array_equal(raw[4 * n:], g["mdot"])
    assert restart == g["restart"].tobytes()


@pytest.mark.gpu
def test_fortran_host_resumes_from_the_reference_restart_file():
    """Started from the reference's restart file after 4 iterations
    (simple_hex_ilu_half.npz), 4 more iterations of the Fortran host give the
    reference's 8-iteration fields and restart file bitwise."""
    half, g = golden("simple_hex_ilu_half.npz"), golden("simple_hex_ilu.npz")
    spec, _ = msg.SIMPLE["hex_ilu"]
    pm, _, _, _, _ = msg.build_mesh(spec)
    with tempfile.TemporaryDirectory() as td:
        mesh_dir = os.path.join(td, "polyMesh")
        meshgen.write_polymesh(pm, mesh_dir)
        rin = os.path.join(td, "half.restart")
        open(rin, "wb").write(half["restart"].tobytes())
        out = os.path.join(td, "run")
        print("fortran host: hex_ilu resumed after 4 iterations", flush=True)
        r = subprocess.run([fb.FORTRAN_HOST, mesh_dir, "ymax", "4", out, rin],
                           capture_output=True, text=True, timeout=120)
        sys.stdout.write(r.stdout + r.stderr)
        assert r.returncode == 0
        raw = np.fromfile(out + ".fields", dtype=np.float64)
        restart = open(out + ".restart", "rb").read()
    n = len(g["u"])
    for k, name in enumerate(("u", "v", "w", "p")):
        assert np.array_equal(raw[k * n:(k + 1) * n], g[name]), name
    assert np.array_equal(raw[4 * n:], g["mdot"])
    assert restart == g["restart"].tobytes()


def test_fortran_host_reports_a_missing_mesh():
    """Host-side failure (no device call is made): the library's message and
    exit status 1."""
    with tempfile.TemporaryDirectory() as td:
        r = subprocess.run([fb.FORTRAN_HOST, os.path.join(td, "none"), "ymax", "1", os.path.join(td, "run")],
                           capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "fs_polymesh_read" in r.stdout
