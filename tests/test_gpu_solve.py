"""Device linear-solve path against the reference's golden vectors and the
oracle (oracle/or_solve.c, or_amg.c).  Runs through the C ABI of
libfsamd.so on an MI355X.

Tolerances: integer maps and the FP64 Galerkin operators are compared
exactly.  PBCG runs the reference's algorithm; its dot products are summed in
a different order, so residual histories agree to 1e-10 relative over the
first iterations (before round-off is amplified) and solutions to the stated
norms.  Solvers whose device form differs from the reference's sequential
sweeps (multicolour Gauss-Seidel, AMG cycles with parallel smoothers) are
compared at convergence against the dense direct solution.
"""
import numpy as np
import pytest

import pyoracle as po
import systems
from conftest import golden

fs = pytest.importorskip("flowsolversourcecode_amd")
pytestmark = pytest.mark.gpu

CASES = ["pneumann", "pshift", "momentum"]


def _mesh(g):
    return fs.Mesh(g["nbr"], g["nnbr"])


def _l1(g, dphi):
    A = systems.dense(g["nbr"], g["nnbr"], g["ac"], g["anb"])
    return np.abs(g["bc"] - A @ dphi).sum()


@pytest.mark.parametrize("case", CASES)
def test_initial_residual_norms(case):
    g = golden(f"solve_{case}.npz")
    m = _mesh(g)
    for nf in (1, 4):
        rc, fit, dphi, res4, st = m.solve_equation(nf, g["fit"], 1e-8, 0, "pbcg", 0, g["ac"], g["anb"], g["bc"])
        ref = po.residuals(g["nbr"], g["nnbr"], g["ac"], g["anb"], g["bc"], np.zeros(len(g["ac"])), g["fit"], nf)
        np.testing.assert_allclose(res4, ref, rtol=1e-14, atol=0)


@pytest.mark.parametrize("case", CASES)
def test_pbcg_follows_reference_history(case):
    g = golden(f"solve_{case}.npz")
    m = _mesh(g)
    m.set_option("check_every", 1)
    href = g["pbcg_hist"]
    k = min(12, len(href))
    rc, fit, dphi, res4, st = m.solve_equation("pressc", g["fit"], 1e-8, k, "pbcg", 0, g["ac"], g["anb"], g["bc"])
    assert st["iterations"] == k
    np.testing.assert_allclose(st["final_l1"], href[k - 1], rtol=1e-10)
    fo, dphi_o, _, hist_o = po.solve_equation(g["nbr"], g["nnbr"], g["ac"], g["anb"], g["bc"], g["fit"], 1e-8, k,
                                              po.PBCG)
    np.testing.assert_allclose(dphi, dphi_o, rtol=1e-9, atol=1e-12 * np.abs(dphi_o).max())
    np.testing.assert_allclose(fit, fo, rtol=1e-9, atol=1e-12 * np.abs(fo).max())


@pytest.mark.parametrize("case", CASES + ["pbig"])
def test_ilu_iterations_bitwise(case):
    """The reference's ILU iteration (exact D-ILU diagonal, level-scheduled
    substitutions in the sequential loops' operation order): dphi after a
    fixed number of iterations is bitwise the oracle's, which is bitwise the
    reference build's (golden ILU solution below)."""
    g = golden(f"solve_{case}.npz")
    m = _mesh(g)
    for k in (1, 3, 7):
        rc, fit, dphi, res4, st = m.solve_equation("pressc", g["fit"], 0.0, k, "ilu", 0, g["ac"], g["anb"], g["bc"])
        assert st["iterations"] == k
        fo, dphi_o, _, _ = po.solve_equation(g["nbr"], g["nnbr"], g["ac"], g["anb"], g["bc"], g["fit"], 0.0, k, po.ILU)
        assert np.array_equal(dphi, dphi_o), (k, np.abs(dphi - dphi_o).max())
    if "ilu_dphi" in g:
        rc, fit, dphi, res4, st = m.solve_equation("pressc", g["fit"], float(g["rrf"]), len(g["ilu_hist"]), "ilu", 0,
                                                   g["ac"], g["anb"], g["bc"])
        assert st["iterations"] == len(g["ilu_hist"])
        assert np.array_equal(dphi, g["ilu_dphi"])


@pytest.mark.parametrize("case", CASES + ["pbig"])
def test_amg_maps_and_galerkin_bitwise(case):
    g = golden(f"solve_{case}.npz")
    m = _mesh(g)
    m.set_option("mg_minparents", 2)
    rc, fit, dphi, res4, st = m.solve_equation("pressc", g["fit"], 1e-8, 200, "bicgs", fs.MG_PRECOND,
                                               g["ac"], g["anb"], g["bc"])
    L = m.amg_level(0)["levels"]
    assert L == len(g["mg_n"])
    parents, seeds, kids, kptr, nbs, nbptr, anbs, acs = [], [], [], [0], [], [0], [], []
    for l in range(L):
        a = m.amg_level(l)
        assert a["n"] == g["mg_n"][l]
        if l < L - 1:
            parents.append(a["parent"])
        if l > 0:
            seeds += list(a["seed"])
            for p in range(a["n"]):
                kids += list(a["child"][a["child_ptr"][p]:a["child_ptr"][p + 1]])
                kptr.append(len(kids))
        acs += list(a["ac"])
        for p in range(a["n"]):
            nbs += list(a["nbr"][:a["nnbr"][p], p])
            anbs += list(a["anb"][:a["nnbr"][p], p])
            nbptr.append(len(nbs))
    assert np.array_equal(np.concatenate(parents), g["mg_parent"])
    assert np.array_equal(seeds, g["mg_seed"])
    assert np.array_equal(kids, g["mg_child"]) and np.array_equal(kptr, g["mg_child_ptr"])
    assert np.array_equal(nbs, g["mg_nbr"]) and np.array_equal(nbptr, g["mg_nbr_ptr"])
    # coarse operators are produced on the device: bitwise equal to acMG/anbMG
    assert np.array_equal(acs, g["mg_ac"]) and np.array_equal(anbs, g["mg_anb"])


@pytest.mark.parametrize("case", ["pshift", "momentum"])
@pytest.mark.parametrize("tag", ["gmres", "bicgs"])
def test_device_krylov_against_the_reference_omp_solvers(case, tag):
    """The reference's 'gmres' and 'bicgs' (SolveAlgebraicSystemInParallel,
    tests/golden/solve_omp_krylov.npz) stop at their own fixed tolerances
    (GMRES_tol 1e-9; BICGSTAB_OMP 1e-5 and 100 iterations) whatever rrF says;
    the device's versions run to the caller's rrF (1e-8 here).  Expected:
    the same initial norms (1e-13), the device's true L1 residual within the
    stop rule (the reference's BiCGSTAB leaves 5e-4 of it on pshift), and the
    two solutions apart by at most the reference's distance from the exact
    solution plus 1e-6 of its largest entry."""
    g = golden(f"solve_{case}.npz")
    k = golden("solve_omp_krylov.npz")
    A = systems.dense(g["nbr"], g["nnbr"], g["ac"], g["anb"])
    x = np.linalg.solve(A, g["bc"])
    m = _mesh(g)
    rc, fit, dphi, res4, st = m.solve_equation("pressc", g["fit"], 1e-8, 400, tag, 0, g["ac"], g["anb"], g["bc"])
    assert rc == 0, st
    ref = k[f"{case}_{tag}_dphi"]
    l1, l1_ref = np.abs(g["bc"] - A @ dphi).sum(), np.abs(g["bc"] - A @ ref).sum()
    print(f"{case} {tag}: device L1 {l1 / np.abs(g['bc']).sum():.1e}, reference {l1_ref / np.abs(g['bc']).sum():.1e}",
          flush=True)
    np.testing.assert_allclose(res4, k[f"{case}_{tag}_res"], rtol=1e-13)
    assert l1 <= 1.0001e-8 * np.abs(g["bc"]).sum()
    assert np.abs(dphi - ref).max() <= np.abs(ref - x).max() + 1e-6 * np.abs(x).max()


def test_amg_maps_at_scale_bitwise():
    """The hierarchy of a 32^3 system (9 levels, minNumberofParents 1: the
    child caps of levels 2, 3, 4 and every later level in use,
    AgglomorateALevelAlgebraicly.for:79-105) against the reference build's
    (tests/golden/solve_amg_scale.npz, the system rebuilt from its seed):
    parents, seeds, children, coarse neighbour lists and coarse operators of
    every coarse level bitwise."""
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from make_solve_golden import SCALE
    g = golden("solve_amg_scale.npz")
    nbr, nnbr, ac, anb, bc = systems.quantized_pressure(*SCALE["dims"], SCALE["seed"])
    m = fs.Mesh(nbr, nnbr)
    m.set_option("mg_minparents", SCALE["minparents"])
    m.set_option("mg_maxlevels", SCALE["maxlevels"])
    print("amg maps at 32^3", flush=True)
    rc, fit, dphi, res4, st = m.solve_equation("pressc", np.zeros(len(ac)), 1e-8, 200, "bicgs", fs.MG_PRECOND,
                                               ac, anb, bc)
    L = m.amg_level(0)["levels"]
    assert L == len(g["mg_n"]) == 9
    parents, seeds, kids, kptr, nbs, nbptr, anbs, acs = [], [], [], [0], [], [0], [], []
    for l in range(L):
        a = m.amg_level(l)
        assert a["n"] == g["mg_n"][l]
        if l < L - 1:
            parents.append(a["parent"])
        if l == 0:
            continue
        seeds += list(a["seed"])
        for p in range(a["n"]):
            kids += list(a["child"][a["child_ptr"][p]:a["child_ptr"][p + 1]])
            kptr.append(len(kids))
        acs += list(a["ac"])
        for p in range(a["n"]):
            nbs += list(a["nbr"][:a["nnbr"][p], p])
            anbs += list(a["anb"][:a["nnbr"][p], p])
            nbptr.append(len(nbs))
    assert np.array_equal(np.concatenate(parents), g["mg_parent"])
    assert np.array_equal(seeds, g["mg_seed"])
    assert np.array_equal(kids, g["mg_child"]) and np.array_equal(kptr, g["mg_child_ptr"])
    assert np.array_equal(nbs, g["mg_nbr"]) and np.array_equal(nbptr, g["mg_nbr_ptr"])
    assert np.array_equal(acs, g["mg_ac"]) and np.array_equal(anbs, g["mg_anb"])
    assert rc == 0


@pytest.mark.parametrize("case", ["pshift", "momentum"])
@pytest.mark.parametrize("solver,lmg,fp32", [("sor", 0, 0), ("bicgs", 0, 0), ("bicgs", 2, 0), ("bicgs", 2, 1),
                                             ("sor", 1, 0), ("sor", 1, 1), ("gmres", 0, 0), ("gmres", 2, 1)])
def test_converged_solution_matches_direct(case, solver, lmg, fp32):
    g = golden(f"solve_{case}.npz")
    m = _mesh(g)
    m.set_option("mg_minparents", 2)
    m.set_option("mg_fp32", fp32)
    m.set_option("mg_maxcycles", 200)
    rc, fit, dphi, res4, st = m.solve_equation("pressc", g["fit"], 1e-8, 2000, solver, lmg, g["ac"], g["anb"], g["bc"])
    assert rc == 0, st
    A = systems.dense(g["nbr"], g["nnbr"], g["ac"], g["anb"])
    x = np.linalg.solve(A, g["bc"])
    assert st["final_l1"] <= 1e-8 * st["initial_l1"] * (1 + 1e-12)
    assert _l1(g, dphi) <= 1.0001e-8 * np.abs(g["bc"]).sum()
    np.testing.assert_allclose(dphi, x, rtol=0, atol=1e-6 * np.abs(x).max())
    np.testing.assert_allclose(fit, g["fit"] + dphi, rtol=1e-15)


def test_neumann_pressure_correction_converges_with_amg():
    """Pure-Neumann p' system (closed cavity): consistent right-hand side,
    singular operator; AMG-preconditioned BiCGSTAB reaches 1e-8 in L1."""
    nbr, nnbr, ac, anb, bc = systems.pressure_like(24, 20, 16, 7, shift=0.0)
    m = fs.Mesh(nbr, nnbr)
    for fp32 in (0, 1):
        m.set_option("mg_fp32", fp32)
        rc, fit, dphi, res4, st = m.solve_equation("pressc", np.zeros(len(ac)), 1e-8, 500, "bicgs", 2, ac, anb, bc)
        assert rc == 0, st
        A = systems.dense(nbr, nnbr, ac, anb)
        assert np.abs(bc - A @ dphi).sum() <= 1.0001e-8 * np.abs(bc).sum()


def test_invalid_solver_code_reports_error():
    g = golden("solve_pshift.npz")
    m = _mesh(g)
    with pytest.raises(fs.FsError):
        m.solve_equation("pressc", g["fit"], 1e-8, 10, 3, 0, g["ac"], g["anb"], g["bc"])


def _permuted(nbr, nnbr, ac, anb, bc, perm):
    """The same system with cell c renumbered perm[c]."""
    w, n = nbr.shape
    nbr2 = np.zeros_like(nbr)
    anb2 = np.zeros_like(anb)
    for s in range(w):
        k = nbr[s]
        nbr2[s, perm] = np.where(k > 0, perm[np.maximum(k, 1) - 1] + 1, 0)
        anb2[s, perm] = anb[s]
    out = [np.zeros_like(a) for a in (nnbr, ac, bc)]
    for o, a in zip(out, (nnbr, ac, bc)):
        o[perm] = a
    return nbr2, out[0], out[1], anb2, out[2]


@pytest.mark.parametrize("solver,lmg,shift", [("bicgs", 2, 0.05), ("cg", 2, 0.05), ("sor", 0, 1.0)])
def test_wide_numbering_uses_32bit_neighbours(solver, lmg, shift):
    """Neighbour offsets beyond +-32767 (a shuffled numbering of 38 400
    cells) take the 32-bit index path of the row loops; the 16-bit path
    serves the natural numbering.  Both reach the same solution."""
    nbr, nnbr, ac, anb, bc = systems.pressure_like(40, 40, 24, 21, shift=shift)
    perm = np.random.default_rng(3).permutation(len(ac))
    p = _permuted(nbr, nnbr, ac, anb, bc, perm)
    assert np.abs(np.where(p[0] > 0, p[0] - 1 - np.arange(len(ac))[None, :], 0)).max() > 32767
    sols = []
    for sys_ in ((nbr, nnbr, ac, anb, bc), p):
        m = fs.Mesh(sys_[0], sys_[1])
        rc, fit, dphi, res4, st = m.solve_equation("pressc", np.zeros(len(ac)), 1e-10, 3000, solver, lmg,
                                                   sys_[2], sys_[3], sys_[4])
        assert rc == 0, st
        sols.append(dphi)
    np.testing.assert_allclose(sols[1][perm], sols[0], rtol=0, atol=1e-7 * np.abs(sols[0]).max())


@pytest.mark.parametrize("solver", ["pbcg", "sor"])
def test_repeat_solve_runs_to_itmax(solver):
    """A solve repeated on one mesh starts with a batch the length of the
    previous iteration count, then batches of 2, 4, ... check_every.  The
    doubling stops at check_every: a repeat solve of more than 262
    iterations (31 batches) must run to itmax like the first.  PBCG
    stagnates on this system (the reference's 1e-10 added to its
    denominators), multicolour Gauss-Seidel is slow on it: neither reaches
    rrf = 1e-12 in 600 iterations, and the three calls are bitwise alike."""
    nbr, nnbr, ac, anb, bc = systems.pressure_like(14, 12, 10, 11, shift=1.0 if solver == "pbcg" else 0.05)
    fit0 = np.linspace(-1.0, 1.0, len(ac))
    m = fs.Mesh(nbr, nnbr)
    outs = [m.solve_equation("pressc", fit0, 1e-12, 600, solver, 0, ac, anb, bc) for _ in range(3)]
    for rc, fit, dphi, res4, st in outs:
        assert rc == 1 and st["iterations"] == 600, (rc, st)
        assert np.isfinite(dphi).all()
        assert np.array_equal(dphi, outs[0][2]) and st["final_l1"] == outs[0][4]["final_l1"]
