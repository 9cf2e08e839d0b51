"""Initial guesses of the p' solve (fs_solve_equation_device_from_guess,
the path of the learned warm start): a guess is first scaled to its multiple
with the smallest L2 residual, then dropped for the zero start when its L1
residual still exceeds that of the zero start.

Expected behaviour, written in each test: a garbage guess -- NaN or infinite
entries, as a diverged network would give, or huge noise -- is dropped, and
the solve is then bitwise the zero-start solve (same dphi, FiT, norms and
iteration count); a good guess (the solution itself, sign flipped and scaled)
is kept, rescaled, and converges in fewer iterations; every solve still
meets the stop rule of the zero start (1e-8 of the zero start's residual)."""
import numpy as np
import pytest

import systems

fs = pytest.importorskip("flowsolversourcecode_amd")
torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _t(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=DEV)


@pytest.fixture(scope="module")
def system():
    nbr, nnbr, ac, anb, bc = systems.pressure_like(24, 20, 16, 7, shift=0.02)
    return nbr, nnbr, ac, anb, bc, systems.sparse(nbr, nnbr, ac, anb)


def _solve(system, guess, solver, lmg=fs.MG_PRECOND):
    nbr, nnbr, ac, anb, bc, A = system
    m = fs.Mesh(nbr, nnbr)
    fit = _t(np.zeros(len(ac)))
    dphi = _t(np.zeros(len(ac)) if guess is None else guess)
    rc, res4, st = m.solve_equation_device("pressc", fit, 1e-8, 400, solver, lmg, _t(ac), _t(anb), _t(bc),
                                           dphi, from_guess=guess is not None)
    return rc, fit.cpu().numpy(), dphi.cpu().numpy(), res4, st


SOLVES = [("bicgs", fs.MG_PRECOND), ("cg", fs.MG_PRECOND), ("pbcg", fs.MG_PRECOND), ("gmres", fs.MG_PRECOND),
          ("sor", fs.MG_SOLVER), ("sor", fs.MG_OFF), ("ilu", fs.MG_OFF)]


@pytest.mark.parametrize("solver,lmg", SOLVES, ids=[f"{a}-{b}" for a, b in SOLVES])
@pytest.mark.parametrize("kind", ["nan", "inf", "noise"])
def test_garbage_guess_is_dropped(system, solver, lmg, kind):
    nbr, nnbr, ac, anb, bc, A = system
    n = len(ac)
    rng = np.random.default_rng(3)
    guess = {"nan": np.full(n, np.nan), "inf": np.where(np.arange(n) % 2, np.inf, -np.inf),
             "noise": 1e12 * rng.standard_normal(n)}[kind]
    if kind == "nan":
        guess[::3] = 1.0  # some finite entries among the NaN
    rc0, fit0, dphi0, res0, st0 = _solve(system, None, solver, lmg)
    rc, fit, dphi, res, st = _solve(system, guess, solver, lmg)
    print(f"{kind} guess, {solver} lmg {lmg}: rc {rc}, {st['iterations']} iterations (zero start rc {rc0}, "
          f"{st0['iterations']})", flush=True)
    assert rc == rc0 and rc in (0, 1), (rc, rc0)  # 1: itmax reached (plain SOR / ILU), from zero as well
    if rc0 == 0:
        assert np.abs(bc - A @ dphi).sum() <= 1.0001e-8 * np.abs(bc).sum()
    if kind != "noise":
        assert st["iterations"] == st0["iterations"]
        assert np.array_equal(dphi, dphi0) and np.array_equal(fit, fit0) and np.array_equal(res, res0)
    else:
        # noise scaled to its best multiple is almost zero: never worse than a
        # few extra iterations, and dropped outright when its residual exceeds
        # the zero start's
        assert st["iterations"] <= st0["iterations"] + 2 or rc0 == 1


@pytest.mark.parametrize("solver", ["bicgs", "cg"])
def test_good_guess_is_kept_and_saves_iterations(system, solver):
    nbr, nnbr, ac, anb, bc, A = system
    rc0, fit0, dphi0, res0, st0 = _solve(system, None, solver)
    near = -3.0 * dphi0 * (1.0 + 1e-4 * np.sin(np.arange(len(ac))))  # the solution, flipped, scaled, perturbed
    rc, fit, dphi, res, st = _solve(system, near, solver)
    print(f"good guess, {solver}: {st['iterations']} iterations (zero start {st0['iterations']})", flush=True)
    assert rc == 0
    assert np.abs(bc - A @ dphi).sum() <= 1.0001e-8 * np.abs(bc).sum()
    assert st["iterations"] < st0["iterations"]


def test_previous_p_guess_in_the_simple_loop():
    """FS_OPT_WARM_PREVIOUS: each p' solve of the device SIMPLE loop starts
    from the previous iteration's p' (1) or from the combination of the last
    4 p' fields with the smallest residual (4).  Every solve still meets the
    stop rule (no unconverged solve), with tight inner solves (momentum 1e-8,
    p' 1e-10) the fields stay within 1e-6 of the zero-start run, and both
    guesses save iterations (consecutive p' fields are nearly parallel)."""
    from flowsolversourcecode_amd import meshgen
    pm = meshgen.box_hex(24, 24, 24, perturb=0.15, seed=3)
    bu, bv, bw = meshgen.lid_driven_cavity_walls(pm)
    out = {}
    for prev in (0, 1, 4):
        c = fs.Case(pm, bu, bv, bw, rho=1.0, mu=1e-2)
        c.configure(urf_u=0.7, urf_p=0.3, msolver="sor", mrrf=1e-8, mitmax=400, psolver="cg", prrf=1e-10,
                    pitmax=400, plmg=fs.MG_PRECOND)
        c.set_option("warm_previous", prev)
        st = c.iterate(30)
        out[prev] = (st, c.get("u"), c.get("p"))
        print(f"warm_previous={prev}: {st['pc_iterations']} p' iterations over 30 SIMPLE iterations", flush=True)
        assert st["pc_unconverged"] == 0
    for prev in (1, 4):
        assert out[prev][0]["pc_iterations"] < out[0][0]["pc_iterations"]
        for k in (1, 2):
            a, b = out[prev][k], out[0][k]
            assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max()
