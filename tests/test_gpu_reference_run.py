"""Converged device fields against the reference's own SIMPLE run.

tests/golden/simple_*_converged.npz and simple_cavity2d_64.npz hold the
fields of the reference's SolveMomentumX/Y/Z and SolveContinuity, compiled
unmodified and iterated to convergence (tests/golden/make_simple_golden.py):
a perturbed hexahedral cavity, a perturbed tetrahedral cavity, the
hexahedral cavity with second-order upwind and limited gradients, the
tetrahedral cavity with SIMPLEC (urf_p = 1), and the
64 x 64 x 1 cavity with symmetry planes at Re = 100.  The device path
(multicolour Gauss-Seidel momentum, BiCGSTAB with the AMG V-cycle for p' to
1e-8) reaches the same fields to 1e-6 of their maximum."""
import importlib.util
import os

import numpy as np
import pytest

from conftest import golden

fs = pytest.importorskip("flowsolversourcecode_amd")
pytestmark = pytest.mark.gpu

_spec = importlib.util.spec_from_file_location(
    "make_simple_golden", os.path.join(os.path.dirname(__file__), "golden", "make_simple_golden.py"))
msg = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(msg)


@pytest.mark.parametrize("name,iters", [("hex_converged", 600), ("tet_converged", 600),
                                        ("hex_sou_converged", 600), ("tet_simplec_converged", 600),
                                        ("cavity2d_64", 600)])
@pytest.mark.parametrize("fp32", [0, 1])
def test_converged_fields_match_the_reference_run(name, iters, fp32):
    g = golden(f"simple_{name}.npz")
    spec, kw = msg.SIMPLE[name]
    pm, bu, bv, bw, types = msg.build_mesh(spec)
    dc = fs.Case(pm, bu, bv, bw, rho=1.0, mu=0.01)
    dc.configure(urf_u=kw.get("urf_u", 0.7), urf_p=kw.get("urf_p", 0.3), msolver="sor", mrrf=0.1, mitmax=10,
                 psolver="bicgs", prrf=1e-8, pitmax=300, plmg=fs.MG_PRECOND)
    dc.set_option("mg_fp32", fp32)
    if types:
        dc.set_patch_types(pm.patch_sizes, types)
    if kw.get("framework", "none") != "none":
        dc.set_convection(kw["framework"], kw["scheme"], kw.get("bleed", 0.0), kw.get("start_iter", 0))
    dc.set_gradient_limiters(kw.get("lim_mom", 0), kw.get("lim_cont", 0))
    dc.set_algorithm(kw.get("algorithm", "simple"))
    dc.iterate(iters)
    for k in ("u", "v", "w", "p"):
        a, b = dc.get(k), g[k]
        assert np.abs(a - b).max() <= 1e-6 * max(np.abs(b).max(), 1e-12), (k, np.abs(a - b).max())


def test_simplec_phases_bitwise():
    """SIMPLEC's D2 = V/(a_P + sum anb) on the device, and the p' system it
    shapes, bitwise the oracle's (which is bitwise the reference's)."""
    import pyoracle as po
    from flowsolversourcecode_amd import meshgen
    pm = meshgen.box_hex(6, 5, 4, perturb=0.15, seed=5)
    bu, bv, bw = meshgen.lid_driven_cavity_walls(pm)
    oc = po.OrCase(pm, bu, bv, bw, mu=0.01)
    dc = fs.Case(pm, bu, bv, bw, mu=0.01)
    oc.set_algorithm("simplec")
    dc.set_algorithm("simplec")
    rng = np.random.default_rng(5)
    for name, n, scale in (("u", dc.N, 1.0), ("v", dc.N, 1.0), ("w", dc.N, 1.0), ("p", dc.N, 3.0),
                           ("mdot", dc.Fi, 0.05), ("bp", dc.Fb, 3.0)):
        a = rng.standard_normal(n) * scale
        oc.set(name, a)
        dc.set(name, a)
    oc.momentum_gradients()
    dc.phase("momentum_gradients")
    for comp, name in enumerate("uvw"):
        oc.momentum_assemble(comp)
        dc.phase(f"assemble_{name}")
        for k in ("ac", "anb", "bc", "d" + name):
            assert np.array_equal(dc.get(k), oc.get(k)), (name, k)
        d2 = np.ctypeslib.as_array(getattr(oc.ptr.contents, f"d{name}2"), shape=(dc.N,))
        assert np.array_equal(dc.get(f"d{name}2"), d2), name
        assert not np.array_equal(d2, oc.get("d" + name))
    oc.pc_assemble()
    dc.phase("pc_assemble")
    for k in ("mdot", "ac", "anb", "bc"):
        assert np.array_equal(dc.get(k), oc.get(k)), k


@pytest.mark.parametrize("kind,gm,gc", [("hex", (1, 2), (2, 3)), ("tet", (2, 1), (1, 0)), ("hex", (2, 3), (2, 3))])
def test_gradient_methods_bitwise(kind, gm, gc):
    """Plain Green-Gauss (method 1) and other correction-sweep counts, split
    between u, v, w and p: the device gradients, momentum systems and p'
    system bitwise the oracle's."""
    import pyoracle as po
    from flowsolversourcecode_amd import meshgen
    pm = (meshgen.box_hex(6, 5, 4, perturb=0.15, seed=6) if kind == "hex"
          else meshgen.box_tet(3, 3, 2, perturb=0.15, seed=6))
    bu, bv, bw = meshgen.lid_driven_cavity_walls(pm)
    oc = po.OrCase(pm, bu, bv, bw, mu=0.01)
    dc = fs.Case(pm, bu, bv, bw, mu=0.01)
    for c in (oc, dc):
        c.set_gradient_method(gm, gc)
        c.set_gradient_limiters(1, 2)
    rng = np.random.default_rng(6)
    for name, n, scale in (("u", dc.N, 1.0), ("v", dc.N, 1.0), ("w", dc.N, 1.0), ("p", dc.N, 3.0),
                           ("mdot", dc.Fi, 0.05), ("bp", dc.Fb, 3.0)):
        a = rng.standard_normal(n) * scale
        oc.set(name, a)
        dc.set(name, a)
    oc.momentum_gradients()
    dc.phase("momentum_gradients")
    for k in ("gu", "gv", "gw", "gp", "bgp"):
        assert np.array_equal(dc.get(k), oc.get(k)), k
    for comp, name in enumerate("uvw"):
        oc.momentum_assemble(comp)
        dc.phase(f"assemble_{name}")
        for k in ("ac", "anb", "bc"):
            assert np.array_equal(dc.get(k), oc.get(k)), (name, k)
    oc.pc_assemble()
    dc.phase("pc_assemble")
    for k in ("mdot", "ac", "anb", "bc"):
        assert np.array_equal(dc.get(k), oc.get(k)), k


@pytest.mark.parametrize("kind,dec,simplec", [("hex", 1, False), ("tet", 2, True), ("hex", 2, False),
                                              ("tet", 1, True)])
def test_decompose_methods_bitwise(kind, dec, simplec):
    """MethodDecomposeS 1 and 2 on the device: the recomputed geometry
    (gDiff, T, gradient-correction offsets) and the momentum and p' systems
    bitwise the oracle's."""
    import pyoracle as po
    from flowsolversourcecode_amd import meshgen
    pm = (meshgen.box_hex(6, 5, 4, perturb=0.2, seed=7) if kind == "hex"
          else meshgen.box_tet(3, 3, 2, perturb=0.2, seed=7))
    bu, bv, bw = meshgen.lid_driven_cavity_walls(pm)
    oc = po.OrCase(pm, bu, bv, bw, mu=0.01)
    dc = fs.Case(pm, bu, bv, bw, mu=0.01)
    for c in (oc, dc):
        c.set_decompose(dec)
        c.set_algorithm("simplec" if simplec else "simple")
    for k in ("gdiff", "Tx", "Ty", "Tz"):
        assert np.array_equal(dc.get(k), oc.mesh.face[k]), k
    rng = np.random.default_rng(7)
    for name, n, scale in (("u", dc.N, 1.0), ("v", dc.N, 1.0), ("w", dc.N, 1.0), ("p", dc.N, 3.0),
                           ("mdot", dc.Fi, 0.05), ("bp", dc.Fb, 3.0)):
        a = rng.standard_normal(n) * scale
        oc.set(name, a)
        dc.set(name, a)
    oc.momentum_gradients()
    dc.phase("momentum_gradients")
    for k in ("gu", "gv", "gw", "gp"):
        assert np.array_equal(dc.get(k), oc.get(k)), k
    for comp, name in enumerate("uvw"):
        oc.momentum_assemble(comp)
        dc.phase(f"assemble_{name}")
        for k in ("ac", "anb", "bc"):
            assert np.array_equal(dc.get(k), oc.get(k)), (name, k)
    oc.pc_assemble()
    dc.phase("pc_assemble")
    for k in ("mdot", "ac", "anb", "bc"):
        assert np.array_equal(dc.get(k), oc.get(k)), k


@pytest.mark.parametrize("kind,gm,gc", [("hex", "upwind", "harmonic"), ("tet", "averagecorrected", "downwind"),
                                        ("hex", "harmonic", "upwind"), ("tet", "downwind", "average")])
def test_face_gradient_schemes_bitwise(kind, gm, gc):
    """GradientInterpolationSchemeMomentum / Continuity on the device: the
    momentum systems (face gradients of u, v, w) and the Rhie-Chow fluxes
    and p' system (face gradient of p) bitwise the oracle's."""
    import pyoracle as po
    from flowsolversourcecode_amd import meshgen
    pm = (meshgen.box_hex(6, 5, 4, perturb=0.2, seed=8) if kind == "hex"
          else meshgen.box_tet(3, 3, 2, perturb=0.2, seed=8))
    bu, bv, bw = meshgen.lid_driven_cavity_walls(pm)
    oc = po.OrCase(pm, bu, bv, bw, mu=0.01)
    dc = fs.Case(pm, bu, bv, bw, mu=0.01)
    for c in (oc, dc):
        c.set_gradient_interpolation(gm, gc)
        c.set_convection("tvd", "minmod")
    rng = np.random.default_rng(8)
    for name, n, scale in (("u", dc.N, 1.0), ("v", dc.N, 1.0), ("w", dc.N, 1.0), ("p", dc.N, 3.0),
                           ("mdot", dc.Fi, 0.05), ("bp", dc.Fb, 3.0)):
        a = rng.standard_normal(n) * scale
        a[rng.random(n) < 0.1] = 0.0
        oc.set(name, a)
        dc.set(name, a)
    oc.momentum_gradients()
    dc.phase("momentum_gradients")
    for comp, name in enumerate("uvw"):
        oc.momentum_assemble(comp)
        dc.phase(f"assemble_{name}")
        for k in ("ac", "anb", "bc"):
            assert np.array_equal(dc.get(k), oc.get(k)), (name, k)
    oc.pc_assemble()
    dc.phase("pc_assemble")
    for k in ("mdot", "ac", "anb", "bc"):
        assert np.array_equal(dc.get(k), oc.get(k)), k


@pytest.mark.parametrize("name", ["hex_ilu", "tet_ilu_sou_simplec", "cavity2d_8_ilu", "hex_ilu_options"])
def test_device_iterations_bitwise_the_reference_run(name):
    """With the reference's ILU for the momentum and p' solves (the device's
    'ilu' reproduces ILUMethod.for's iteration on the reference's numbering),
    eight whole device SIMPLE iterations equal the reference's own run bitwise:
    velocities, pressure and mass fluxes (simple_*ilu*.npz)."""
    g = golden(f"simple_{name}.npz")
    spec, kw = msg.SIMPLE[name]
    pm, bu, bv, bw, types = msg.build_mesh(spec)
    dc = fs.Case(pm, bu, bv, bw, rho=1.0, mu=0.01)
    dc.configure(urf_u=kw.get("urf_u", 0.7), urf_p=kw.get("urf_p", 0.3), msolver="ilu", mrrf=kw.get("mrrf", 0.1),
                 mitmax=kw.get("mitmax", 20), psolver="ilu", prrf=kw.get("prrf", 0.01),
                 pitmax=kw.get("pitmax", 200), plmg=0)
    if types:
        dc.set_patch_types(pm.patch_sizes, types)
    if kw.get("framework", "none") != "none":
        dc.set_convection(kw["framework"], kw["scheme"], kw.get("bleed", 0.0), kw.get("start_iter", 0))
    dc.set_gradient_limiters(kw.get("lim_mom", 0), kw.get("lim_cont", 0))
    dc.set_algorithm(kw.get("algorithm", "simple"))
    dc.set_gradient_method(kw.get("grad_mom", (2, 2)), kw.get("grad_cont", (2, 2)))
    dc.set_decompose(kw.get("decomp", 3))
    dc.set_gradient_interpolation(kw.get("gi_mom", "average"), kw.get("gi_cont", "averagecorrected"))
    dc.iterate(kw["niter"] - 1)
    dc.set_restart_capture(True)
    dc.iterate(1)
    for k in ("u", "v", "w", "p", "mdot"):
        assert np.array_equal(dc.get(k), g[k]), (name, k, np.abs(dc.get(k) - g[k]).max())
    # the restart file of the final state, byte for byte the reference's
    # writeSolution (WriteSolutionToContinue.for:3)
    import tempfile
    import pyoracle as po
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "restart.bin")
        dc.write_solution(path, pm.patch_sizes)
        mine = open(path, "rb").read()
    ref = g["restart"].tobytes()
    rm, rr = po.fortran_records(mine), po.fortran_records(ref)
    assert len(rm) == len(rr), (len(rm), len(rr))
    bad = [k for k in range(len(rr)) if rm[k] != rr[k]]
    assert not bad, (name, bad)
    assert mine == ref


def test_resume_from_the_reference_restart_file():
    """Read the reference's restart file after 4 iterations
    (simple_hex_ilu_half.npz), iterate 4 more on the device: the fields equal
    the reference's 8-iteration run bitwise (simple_hex_ilu.npz); a cut file
    and a file of another mesh are refused."""
    import tempfile
    half, full = golden("simple_hex_ilu_half.npz"), golden("simple_hex_ilu.npz")
    spec, kw = msg.SIMPLE["hex_ilu"]
    pm, bu, bv, bw, types = msg.build_mesh(spec)
    dc = fs.Case(pm, bu, bv, bw, rho=1.0, mu=0.01)
    dc.configure(urf_u=0.7, urf_p=0.3, msolver="ilu", mrrf=0.1, mitmax=20, psolver="ilu", prrf=0.01, pitmax=200,
                 plmg=0)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "restart.bin")
        open(path, "wb").write(half["restart"].tobytes())
        dc.read_solution(path, pm.patch_sizes)
        for k in ("u", "v", "w", "p", "mdot"):
            assert np.array_equal(dc.get(k), half[k]), k
        dc.iterate(4)
        for k in ("u", "v", "w", "p", "mdot"):
            assert np.array_equal(dc.get(k), full[k]), (k, np.abs(dc.get(k) - full[k]).max())
        # a cut record or stray bytes after the last record are refused
        for blob in (half["restart"].tobytes()[:-2], half["restart"].tobytes() + b"\0\0"):
            open(path, "wb").write(blob)
            with pytest.raises(fs.FsError):
                dc.read_solution(path, pm.patch_sizes)
        open(path, "wb").write(half["restart"].tobytes())
        from flowsolversourcecode_amd import meshgen
        pm2 = meshgen.box_hex(5, 5, 4, perturb=0.15, seed=3)
        other = fs.Case(pm2, *meshgen.lid_driven_cavity_walls(pm2), rho=1.0, mu=0.01)
        with pytest.raises(fs.FsError):
            other.read_solution(path, pm2.patch_sizes)
