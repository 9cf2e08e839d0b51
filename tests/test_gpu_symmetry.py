"""Symmetry boundaries on the device SIMPLE path against the oracle, through
the C ABI (fs_case_set_patch_types, include/fsamd_simple.h).

The assembly phases are compared bitwise on a perturbed, tilted mesh whose
front and back planes form one symmetry patch, so that every component of
the symmetry terms is non-zero and the z-normal the reference's symmetry
pressure-gradient projection reads from face i of patch i
(CorrectGradients.for:274) differs from the face's own on half the faces.
The two-dimensional cavity (64 x 64 x 1 cells, Re = 100, frontBack
symmetry) is compared at its converged state, to 1e-6, and against the
published centre-line velocities."""
import numpy as np
import pytest

import pyoracle as po
from flowsolversourcecode_amd import meshgen
from test_oracle_symmetry import ghia_errors

fs = pytest.importorskip("flowsolversourcecode_amd")
pytestmark = pytest.mark.gpu


def _tilted_pair(seed=3):
    pm = meshgen.box_hex(6, 5, 3, perturb=0.15, seed=seed)
    pm = meshgen.regroup(pm, [("lid", ["ymax"]), ("walls", ["xmin", "xmax", "ymin"]),
                              ("frontBack", ["zmin", "zmax"])])
    pm = meshgen.rotate_x(pm, 0.6)
    bu, bv, bw = meshgen.lid_driven_cavity_walls(pm, lid="lid", lid_velocity=(1.0, 0.2, -0.1))
    types = ["wall", "wall", "symmetry"]
    oc = po.OrCase(pm, bu, bv, bw, mu=0.01)
    dc = fs.Case(pm, bu, bv, bw, mu=0.01)
    oc.set_patch_types(pm.patch_sizes, types)
    dc.set_patch_types(pm.patch_sizes, types)
    rng = np.random.default_rng(seed)
    for name, n, scale in (("u", dc.N, 1.0), ("v", dc.N, 1.0), ("w", dc.N, 1.0), ("p", dc.N, 3.0),
                           ("mdot", dc.Fi, 0.01), ("bp", dc.Fb, 3.0)):
        a = rng.standard_normal(n) * scale
        oc.set(name, a)
        dc.set(name, a)
    return pm, oc, dc


@pytest.mark.parametrize("seed", [3, 4])
def test_symmetry_phases_bitwise(seed):
    pm, oc, dc = _tilted_pair(seed)
    oc.update_symmetry()
    dc.phase("update_symmetry")
    for k in ("bu", "bv", "bw"):
        assert np.array_equal(dc.get(k), oc.get(k)), k
    oc.momentum_gradients()
    dc.phase("momentum_gradients")
    for k in ("gu", "gv", "gw", "gp", "bgp"):
        assert np.array_equal(dc.get(k), oc.get(k)), k
    for comp, name in enumerate("uvw"):
        oc.momentum_assemble(comp)
        dc.phase(f"assemble_{name}")
        for k in ("ac", "anb", "bc", "d" + name):
            assert np.array_equal(dc.get(k), oc.get(k)), (name, k)
    oc.pc_assemble()
    dc.phase("pc_assemble")
    for k in ("mdot", "ac", "anb", "bc"):
        assert np.array_equal(dc.get(k), oc.get(k)), k


def test_patch_types_are_checked():
    m, (bu, bv, bw), types = meshgen.cavity_2d(4)
    dc = fs.Case(m, bu, bv, bw)
    with pytest.raises(fs.FsError):
        dc.set_patch_types(m.patch_sizes[:-1], types[:-1])
    with pytest.raises(fs.FsError):   # a symmetry patch 3 needs 3 faces
        dc.set_patch_types([m.patch_sizes[0], m.patch_sizes[1] + m.patch_sizes[2] - 2, 2],
                           ["wall", "wall", "symmetry"])


@pytest.mark.parametrize("fp32", [0, 1])
def test_cavity_2d_converged_fields_match_the_oracle(fp32):
    n, iters = 64, 600
    m, (bu, bv, bw), types = meshgen.cavity_2d(n)
    oc = po.OrCase(m, bu, bv, bw, rho=1.0, mu=0.01, urf_u=0.9, urf_p=0.1, msolver=po.SOR, mrrf=0.1, mitmax=10,
                   psolver=po.SOR, prrf=0.01, pitmax=100)
    oc.set_patch_types(m.patch_sizes, types)
    oc.iterate(iters)
    dc = fs.Case(m, bu, bv, bw, rho=1.0, mu=0.01)
    dc.set_patch_types(m.patch_sizes, types)
    dc.configure(urf_u=0.9, urf_p=0.1, msolver="sor", mrrf=0.1, mitmax=10, psolver="bicgs", prrf=1e-8,
                 pitmax=300, plmg=fs.MG_PRECOND)
    dc.set_option("mg_fp32", fp32)
    first = None
    for it in range(iters):
        st = dc.iterate(1)
        first = first or st["pc_initial_l1"]
        # every p' solve meets the stop rule until the flow has converged; by
        # then |bc|_1 is below 1e-6 of its first value and the relative target
        # 1e-8 lies under the round-off floor of the system (~1e-17)
        assert st["pc_unconverged"] == 0 or st["pc_initial_l1"] <= 1e-6 * first, (it, st)
    for k in ("u", "v", "p"):
        a, b = dc.get(k), oc.get(k)
        assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max(), (k, np.abs(a - b).max())
    assert np.all(dc.get("w") == 0.0)
    eu, ev = ghia_errors(dc.get("u"), dc.get("v"), n)
    assert eu < 0.015 and ev < 0.015
