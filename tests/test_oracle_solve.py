"""The C restatement (oracle/) against the reference's own outputs.

Golden vectors in tests/golden/solve_*.npz come from the reference's Fortran
(SolveAlegbraicEquations.for, SORMethod.for, FactorizeUsingILU.for,
ILUMethod.for, PBCGMethod.for, AlgebraicMultiGrid/) compiled by
oracle/build_ref.sh.  Equality is exact: the restatement forms every sum in
the reference's order.  Independent mathematical pins follow at the end.
"""
import numpy as np
import pytest

import pyoracle as po
import systems
from conftest import golden

CASES = ["pneumann", "pshift", "momentum"]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("solver,tag", [(po.SOR, "sor"), (po.ILU, "ilu"), (po.PBCG, "pbcg")])
def test_solver_matches_reference_bitwise(case, solver, tag):
    g = golden(f"solve_{case}.npz")
    fit, dphi, res, hist = po.solve_equation(g["nbr"], g["nnbr"], g["ac"], g["anb"], g["bc"], g["fit"],
                                             float(g["rrf"]), 200, solver)
    assert np.array_equal(hist, g[f"{tag}_hist"])
    assert np.array_equal(dphi, g[f"{tag}_dphi"])
    assert np.array_equal(res, g[f"{tag}_res"])
    assert np.array_equal(fit, g[f"{tag}_fit"])


@pytest.mark.parametrize("case", CASES + ["pbig"])
def test_amg_maps_and_operators_match_reference(case):
    g = golden(f"solve_{case}.npz")
    h = po.Amg(g["nbr"], g["nnbr"], g["ac"], g["anb"], maxlevels=10, minparents=2)
    assert h.nlevels == len(g["mg_n"])
    parents, seeds, kids, kptr, nbs, nbptr, anbs, acs = [], [], [], [0], [], [0], [], []
    for l in range(h.nlevels):
        a = h.level(l)
        assert a["n"] == g["mg_n"][l]
        if l < h.nlevels - 1:
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
    assert np.array_equal(acs, g["mg_ac"]) and np.array_equal(anbs, g["mg_anb"])


def test_amg_maps_at_scale_match_reference():
    """The oracle's agglomeration of the 32^3 system of
    tests/golden/solve_amg_scale.npz (9 levels, every child cap in use):
    maps and coarse operators bitwise the reference build's."""
    import os, sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from make_solve_golden import SCALE
    g = golden("solve_amg_scale.npz")
    nbr, nnbr, ac, anb, bc = systems.quantized_pressure(*SCALE["dims"], SCALE["seed"])
    h = po.Amg(nbr, nnbr, ac, anb, maxlevels=SCALE["maxlevels"], minparents=SCALE["minparents"])
    assert h.nlevels == len(g["mg_n"]) == 9
    parents, seeds, kids, kptr, nbs, nbptr, anbs, acs = [], [], [], [0], [], [0], [], []
    for l in range(h.nlevels):
        a = h.level(l)
        assert a["n"] == g["mg_n"][l]
        if l < h.nlevels - 1:
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


@pytest.mark.parametrize("case", CASES + ["pbig"])
def test_amg_vcycle_solution_matches_reference(case):
    g = golden(f"solve_{case}.npz")
    h = po.Amg(g["nbr"], g["nnbr"], g["ac"], g["anb"], maxlevels=10, minparents=2)
    n = len(g["ac"])
    res0 = po.residuals(g["nbr"], g["nnbr"], g["ac"], g["anb"], g["bc"], np.zeros(n), g["fit"])
    dphi, _ = h.solve(g["bc"], res0[0], float(g["rrf"]), 30, 3, 5)
    assert np.array_equal(dphi, g["amg_dphi"])


@pytest.mark.parametrize("cycle", [1, 2])
@pytest.mark.parametrize("case", ["pshift", "pbig"])
def test_amg_f_and_w_cycles_match_reference(case, cycle):
    """F (1) and W (2) cycles (MultiGridCycle.for:55,133): the oracle's dphi
    after the cycle loop equals the reference build's SolveEquation with
    MultiGridCycleType 'fcycle' / 'wcycle' bitwise
    (tests/golden/solve_amg_cycles.npz, and ref_solve.x itself where it is
    built)."""
    g = golden(f"solve_{case}.npz")
    n = len(g["ac"])
    h = po.Amg(g["nbr"], g["nnbr"], g["ac"], g["anb"], maxlevels=10, minparents=2)
    res0 = po.residuals(g["nbr"], g["nnbr"], g["ac"], g["anb"], g["bc"], np.zeros(n), g["fit"])
    dphi, hist = h.solve(g["bc"], res0[0], float(g["rrf"]), 30, 2, 3, cycle=cycle)
    assert hist[-1] < 0.5 * res0[0]
    assert np.array_equal(dphi, golden("solve_amg_cycles.npz")[f"{case}_{'fw'[cycle - 1]}_dphi"])
    if po.ref_available():
        r = po.run_reference(g["nbr"], g["nnbr"], g["ac"], g["anb"], g["bc"], g["fit"], float(g["rrf"]), 30, 0,
                             mode=0, lmg=1, cycle=cycle, npre=2, npost=3, minparents=2)[0]
        assert np.array_equal(dphi, r["dphi"])


# ---------------------------------------------------------------- pins by maths

def test_galerkin_equals_PtAP():
    nbr, nnbr, ac, anb, bc = systems.momentum_like(9, 7, 5, 11)
    h = po.Amg(nbr, nnbr, ac, anb, maxlevels=4, minparents=2)
    A = systems.dense(nbr, nnbr, ac, anb)
    for l in range(1, h.nlevels):
        f, c = h.level(l - 1), h.level(l)
        P = np.zeros((f["n"], c["n"]))
        P[np.arange(f["n"]), f["parent"] - 1] = 1.0
        Ac = systems.dense(c["nbr"], c["nnbr"], c["ac"], c["anb"])
        np.testing.assert_allclose(Ac, P.T @ A @ P, rtol=1e-13, atol=1e-12)
        A = Ac


def test_aggregates_partition_the_level():
    nbr, nnbr, ac, anb, bc = systems.pressure_like(10, 9, 8, 5)
    h = po.Amg(nbr, nnbr, ac, anb, maxlevels=5, minparents=2)
    for l in range(1, h.nlevels):
        f, c = h.level(l - 1), h.level(l)
        members = []
        for p in range(c["n"]):
            if c["seed"][p]:
                members.append(c["seed"][p])
                assert f["parent"][c["seed"][p] - 1] == p + 1
            for ch in c["child"][c["child_ptr"][p]:c["child_ptr"][p + 1]]:
                members.append(ch)
                assert f["parent"][ch - 1] == p + 1
        assert sorted(members) == list(range(1, f["n"] + 1))


@pytest.mark.parametrize("solver", [po.SOR, po.ILU])
def test_stationary_solvers_reach_direct_solution(solver):
    nbr, nnbr, ac, anb, bc = systems.momentum_like(6, 5, 4, 21)
    A = systems.dense(nbr, nnbr, ac, anb)
    x = np.linalg.solve(A, bc)
    fit, dphi, res, hist = po.solve_equation(nbr, nnbr, ac, anb, bc, np.zeros(len(ac)), 1e-12, 500, solver)
    np.testing.assert_allclose(dphi, x, rtol=1e-9, atol=1e-11)
    assert hist[-1] <= 1e-12 * np.abs(bc).sum()


def test_ilu_diagonal_brute_force():
    """dc_j = 1/(a_jj - sum_{i<j} a_ji dc_i a_ij), the D-ILU recurrence, with
    the factor update skipped for row N-1 (1-based)."""
    nbr, nnbr, ac, anb, bc = systems.momentum_like(4, 3, 3, 8)
    A = systems.dense(nbr, nnbr, ac, anb)
    n = len(ac)
    d = ac.copy()
    for i in range(n):
        d[i] = 1.0 / d[i]
        if i == n - 2:
            continue
        for j in range(i + 1, n):
            if A[i, j] != 0.0:
                d[j] = d[j] - A[j, i] * d[i] * A[i, j]
    np.testing.assert_allclose(po.ilu_factor(nbr, nnbr, ac, anb), d, rtol=1e-15)


def test_residual_norms_closed_form():
    nbr, nnbr, ac, anb, bc = systems.momentum_like(5, 4, 3, 2)
    A = systems.dense(nbr, nnbr, ac, anb)
    rng = np.random.default_rng(0)
    dphi, fit = rng.standard_normal(len(ac)), rng.standard_normal(len(ac))
    r = np.abs(bc - A @ dphi)
    out = po.residuals(nbr, nnbr, ac, anb, bc, dphi, fit, nf=1)
    np.testing.assert_allclose(out[:3], [r.sum(), r.max(), np.sqrt((r * r).mean())], rtol=1e-13)
    np.testing.assert_allclose(out[3], (r / (np.abs(ac * fit).max() + 1e-15)).max(), rtol=1e-13)
