"""Device SIMPLE path (include/fsamd_simple.h) against the oracle
(oracle/or_simple.c, or_geom.c) through the C ABI.

Every assembly phase is compared bitwise: the device forms each face flux
with the reference's operation order and gathers faces per cell in the
reference's face-loop order.  Linear solves on the device differ from the
reference's sequential sweeps, so whole iterations are compared at the
converged state, to 1e-6 relative (the field tolerance of the path)."""
import numpy as np
import pytest

import pyoracle as po
from flowsolversourcecode_amd import meshgen

fs = pytest.importorskip("flowsolversourcecode_amd")
pytestmark = pytest.mark.gpu


def _pair(shape=(7, 6, 5), perturb=0.15, seed=3, mu=0.01, **okw):
    pm = meshgen.box_hex(*shape, perturb=perturb, seed=seed)
    bu, bv, bw = meshgen.lid_driven_cavity_walls(pm)
    oc = po.OrCase(pm, bu, bv, bw, mu=mu, **okw)
    dc = fs.Case(pm, bu, bv, bw, mu=mu)
    return pm, oc, dc


def _random_state(oc, dc, seed=0):
    rng = np.random.default_rng(seed)
    N, Fi, Fb = dc.N, dc.Fi, dc.Fb
    for name, n, scale in (("u", N, 1.0), ("v", N, 1.0), ("w", N, 1.0), ("p", N, 3.0), ("mdot", Fi, 0.01),
                           ("bp", Fb, 3.0)):
        a = rng.standard_normal(n) * scale
        oc.set(name, a)
        dc.set(name, a)


def test_device_geometry_bitwise():
    pm, oc, dc = _pair()
    m = oc.mesh
    assert np.array_equal(dc.get("vol"), m.cell["vol"])
    for k, o in (("Sx", "Sx"), ("Sy", "Sy"), ("Sz", "Sz"), ("g", "g"), ("gdiff", "gdiff"), ("Tx", "Tx"),
                 ("dcf", "dcf"), ("A", "A")):
        assert np.array_equal(dc.get(k), m.face[o]), k


@pytest.mark.parametrize("seed", [0, 1])
def test_momentum_gradients_bitwise(seed):
    pm, oc, dc = _pair(seed=seed + 5)
    _random_state(oc, dc, seed)
    oc.momentum_gradients()
    dc.phase("momentum_gradients")
    for k in ("gu", "gv", "gw", "gp"):
        assert np.array_equal(dc.get(k), oc.get(k)), k
    # the face gradients are formed inside the face kernels (momentum fluxes,
    # Rhie-Chow) and checked through the assembled systems
    assert np.array_equal(dc.get("bgp"), oc.get("bgp"))


def test_momentum_and_pressure_correction_assembly_bitwise():
    pm, oc, dc = _pair()
    _random_state(oc, dc, 7)
    oc.momentum_gradients()
    dc.phase("momentum_gradients")
    for comp, name in enumerate("uvw"):
        oc.momentum_assemble(comp)
        dc.phase(f"assemble_{name}")
        for k in ("ac", "anb", "bc", "d" + name, name + "s"):
            assert np.array_equal(dc.get(k), oc.get(k)), (name, k)
    oc.pc_assemble()
    dc.phase("pc_assemble")
    for k in ("mdot", "ac", "anb", "bc"):
        assert np.array_equal(dc.get(k), oc.get(k)), k


def test_first_iterations_track_the_oracle_with_tight_solves():
    """With every linear system solved to round-off on both sides, whole
    SIMPLE iterations agree to ~1e-10 although the solvers differ."""
    pm, oc, dc = _pair(shape=(6, 5, 4), msolver=po.SOR, mrrf=1e-14, mitmax=4000, psolver=po.SOR, prrf=1e-14,
                       pitmax=20000)
    dc.configure(urf_u=0.7, urf_p=0.3, msolver="bicgs", mrrf=1e-14, mitmax=400, psolver="bicgs", prrf=1e-14,
                 pitmax=400, plmg=fs.MG_PRECOND)
    for it in range(3):
        oc.iterate(1)
        dc.iterate(1)
        for k in ("u", "v", "w", "p"):
            a, b = dc.get(k), oc.get(k)
            assert np.abs(a - b).max() <= 1e-9 * max(np.abs(b).max(), 1e-30), (it, k)


def test_converged_cavity_fields_agree_with_oracle():
    pm, oc, dc = _pair(shape=(8, 8, 8), perturb=0.15, seed=2, psolver=po.SOR, prrf=1e-4, pitmax=500)
    dc.configure(urf_u=0.7, urf_p=0.3, msolver="sor", mrrf=0.1, mitmax=20, psolver="bicgs", prrf=1e-8,
                 pitmax=300, plmg=fs.MG_PRECOND)
    oc.iterate(250)
    st = dc.iterate(250)
    assert oc.mass_imbalance() < 1e-12 and st["mass_l1"] < 1e-12, (oc.mass_imbalance(), st)
    for k in ("u", "v", "w", "p"):
        a, b = dc.get(k), oc.get(k)
        assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max(), k


def test_fp32_amg_pressure_correction_converges_to_same_fields():
    pm, oc, dc = _pair(shape=(8, 8, 8), perturb=0.15, seed=2)
    dc.configure(urf_u=0.7, urf_p=0.3, msolver="sor", mrrf=0.1, mitmax=20, psolver="bicgs", prrf=1e-8,
                 pitmax=300, plmg=fs.MG_PRECOND)
    ref = fs.Case(pm, *meshgen.lid_driven_cavity_walls(pm), mu=0.01)
    ref.configure(urf_u=0.7, urf_p=0.3, msolver="sor", mrrf=0.1, mitmax=20, psolver="bicgs", prrf=1e-8,
                  pitmax=300, plmg=fs.MG_PRECOND)
    dc.set_option("mg_fp32", 1)
    dc.iterate(250)
    ref.iterate(250)
    for k in ("u", "v", "w", "p"):
        a, b = dc.get(k), ref.get(k)
        assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max(), k
