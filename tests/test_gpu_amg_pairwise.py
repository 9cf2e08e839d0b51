"""The pairwise aggregations of the device AMG (FS_OPT_MG_AGGREGATION 1 and
2; no reference counterpart, the reference's greedy agglomeration stays the
default and is pinned bitwise elsewhere).

Expected behaviour, written in each test: the level-1 map of one pairwise
pass equals the matching rule run here in plain Python on the exported fine
operator (each unmatched cell, in order, takes the unmatched neighbour of its
strongest negative coupling among those at least 0.25 of the row's
strongest); two passes give aggregates of at most 4 cells, each connected in
the fine graph; as the preconditioner of BiCGSTAB, CG and GMRES the cycle
drives the true L1 residual below 1e-8 of its initial value and reaches the
direct solution of a non-singular system to 1e-6 of its largest entry."""
import numpy as np
import pytest
import scipy.sparse.linalg as spl

import systems

fs = pytest.importorskip("flowsolversourcecode_amd")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def neumann():
    nbr, nnbr, ac, anb, bc = systems.pressure_like(32, 24, 20, 11, shift=0.0)
    return nbr, nnbr, ac, anb, bc, systems.sparse(nbr, nnbr, ac, anb)


@pytest.fixture(scope="module")
def shifted():
    nbr, nnbr, ac, anb, bc = systems.pressure_like(28, 20, 16, 12, shift=0.05)
    A = systems.sparse(nbr, nnbr, ac, anb)
    return nbr, nnbr, ac, anb, bc, A, spl.spsolve(A.tocsc(), bc)


def _mesh(nbr, nnbr, agg, **opts):
    m = fs.Mesh(nbr, nnbr)
    m.set_option("mg_aggregation", agg)
    for k, v in opts.items():
        m.set_option(k, v)
    return m


def _pairs(nbr, nnbr, anb, beta=0.25):
    """The matching rule in plain Python: 0-based pair number of each cell."""
    w, n = nbr.shape
    agg = [-1] * n
    na = 0
    for i in range(n):
        if agg[i] >= 0:
            continue
        row = [(int(nbr[l, i]) - 1, float(anb[l, i])) for l in range(int(nnbr[i])) if nbr[l, i] > 0]
        mx = max([0.0] + [-a for _, a in row])
        best, bv = -1, 0.0
        for j, a in row:
            s = -a
            if j == i or agg[j] >= 0 or not s > 0.0 or s < beta * mx:
                continue
            if s > bv:
                bv, best = s, j
        agg[i] = na
        if best >= 0:
            agg[best] = na
        na += 1
    ren, out = {}, []
    for i in range(n):
        ren.setdefault(agg[i], len(ren) + 1)
        out.append(ren[agg[i]])
    return np.array(out)


def test_one_pass_is_the_matching_rule(shifted):
    nbr, nnbr, ac, anb, bc, A, x = shifted
    m = _mesh(nbr, nnbr, 1)
    rc, fit, dphi, res4, st = m.solve_equation("pressc", np.zeros(len(ac)), 1e-8, 400, "bicgs", 2, ac, anb, bc)
    assert rc == 0, st
    lev0 = m.amg_level(0)
    assert np.array_equal(lev0["parent"], _pairs(nbr, nnbr, anb))
    lev1 = m.amg_level(1)
    assert lev1["n"] < 0.6 * len(ac) and (lev1["nchild"] <= 1).all()


def test_two_passes_give_connected_aggregates_of_at_most_four(shifted):
    nbr, nnbr, ac, anb, bc, A, x = shifted
    m = _mesh(nbr, nnbr, 2)
    rc, fit, dphi, res4, st = m.solve_equation("pressc", np.zeros(len(ac)), 1e-8, 400, "bicgs", 2, ac, anb, bc)
    assert rc == 0, st
    for level in range(2):
        f = m.amg_level(level)
        c = m.amg_level(level + 1)
        parent = f["parent"]
        assert (c["nchild"] <= 3).all() and c["n"] < 0.45 * f["n"], (level, c["n"], f["n"])
        members = {}
        for i, p in enumerate(parent):
            members.setdefault(int(p), []).append(i)
        for p, mem in members.items():
            assert c["seed"][p - 1] == mem[0] + 1  # the first fine cell seeds it
            s = set(mem)
            seen, todo = {mem[0]}, [mem[0]]
            while todo:
                i = todo.pop()
                for l in range(int(f["nnbr"][i])):
                    j = int(f["nbr"][l, i]) - 1
                    if j in s and j not in seen:
                        seen.add(j)
                        todo.append(j)
            assert seen == s, (level, p, mem)


@pytest.mark.parametrize("agg", [1, 2])
@pytest.mark.parametrize("solver", ["bicgs", "cg", "gmres"])
@pytest.mark.parametrize("fp32", [0, 1])
def test_pairwise_preconditioner_reaches_1e8(neumann, agg, solver, fp32):
    nbr, nnbr, ac, anb, bc, A = neumann
    m = _mesh(nbr, nnbr, agg, mg_fp32=fp32)
    rc, fit, dphi, res4, st = m.solve_equation("pressc", np.zeros(len(ac)), 1e-8, 400, solver, 2, ac, anb, bc)
    print(f"pairwise agg={agg} {solver} fp32={fp32}: {st['iterations']} iterations, {st['levels']} levels",
          flush=True)
    assert rc == 0, st
    assert np.abs(bc - A @ dphi).sum() <= 1.0001e-8 * np.abs(bc).sum()
    assert st["iterations"] < 80, st


@pytest.mark.parametrize("agg", [1, 2])
def test_pairwise_matches_direct(shifted, agg):
    nbr, nnbr, ac, anb, bc, A, x = shifted
    m = _mesh(nbr, nnbr, agg)
    rc, fit, dphi, res4, st = m.solve_equation("pressc", np.zeros(len(ac)), 1e-8, 400, "bicgs", 2, ac, anb, bc)
    assert rc == 0, st
    np.testing.assert_allclose(dphi, x, rtol=0, atol=1e-6 * np.abs(x).max())


def test_bad_aggregation_rejected():
    nbr, nnbr, ac, anb, bc = systems.pressure_like(6, 5, 4, 1, shift=0.1)
    m = fs.Mesh(nbr, nnbr)
    with pytest.raises(fs.FsError):
        m.set_option("mg_aggregation", 3)
