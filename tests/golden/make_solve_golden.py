"""Write golden vectors for the linear-solve path from the reference build.

Runs oracle/_ref/ref_solve.x (the reference's own SolveEquation, SOR, D-ILU,
PBCG and algebraic-multigrid sources, compiled by oracle/build_ref.sh) on small
seeded systems from tests/systems.py and stores inputs and outputs as .npz.
Run in the build container only (needs /root/reference):

    make -C oracle ref && python tests/golden/make_solve_golden.py [amg_scale | omp_krylov | amg_cycles]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, ".."))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import pyoracle as po  # noqa: E402
import systems  # noqa: E402

CASES = {
    "pneumann": lambda: systems.pressure_like(7, 6, 5, 1, shift=0.0),
    "pshift": lambda: systems.pressure_like(7, 6, 5, 2, shift=0.05),
    "momentum": lambda: systems.momentum_like(7, 6, 5, 3),
    "pbig": lambda: systems.pressure_like(16, 12, 10, 4),
}


def flatten_mg(levels):
    out = {"mg_n": np.array([lv["n"] for lv in levels], np.int32)}
    out["mg_parent"] = np.concatenate([lv["parent"] for lv in levels[:-1]]).astype(np.int32)
    seeds, kids, kptr, nb, nbptr, an, ac = [], [], [0], [], [0], [], []
    for lv in levels[1:]:
        seeds += lv["seed"]
        for c in lv["children"]:
            kids += c
            kptr.append(len(kids))
    for lv in levels:
        ac += list(lv["ac"])
        for p in range(lv["n"]):
            nb += lv["nbr"][p]
            an += lv["anb"][p]
            nbptr.append(len(nb))
    out.update(mg_seed=np.array(seeds, np.int32), mg_child=np.array(kids, np.int32),
               mg_child_ptr=np.array(kptr, np.int32), mg_nbr=np.array(nb, np.int32),
               mg_nbr_ptr=np.array(nbptr, np.int32), mg_anb=np.array(an), mg_ac=np.array(ac))
    return out


# the hierarchy at scale: 32^3 cells, 9 levels with minNumberofParents 1 (the
# child caps of levels 2, 3, 4 and beyond all in use); the system is rebuilt
# from its seed, the file holds the maps and the coarse operators only
SCALE = dict(dims=(32, 32, 32), seed=5, minparents=1, maxlevels=12)


def write_amg_scale():
    nbr, nnbr, ac, anb, bc = systems.quantized_pressure(*SCALE["dims"], SCALE["seed"])
    r = po.run_reference(nbr, nnbr, ac, anb, bc, np.zeros(len(ac)), 1e-8, 30, 0, mode=0, lmg=1,
                         minparents=SCALE["minparents"], maxlevels=SCALE["maxlevels"])[0]
    full = flatten_mg(r["mg"])
    n0 = len(ac)
    k0 = int(full["mg_nbr_ptr"][n0])  # the fine level's lists are the input itself
    rec = {k: full[k] for k in ("mg_n", "mg_parent", "mg_seed", "mg_child", "mg_child_ptr")}
    rec["mg_nbr"] = full["mg_nbr"][k0:]
    rec["mg_nbr_ptr"] = full["mg_nbr_ptr"][n0:] - k0
    rec["mg_anb"] = full["mg_anb"][k0:]
    rec["mg_ac"] = full["mg_ac"][n0:]
    path = os.path.join(HERE, "solve_amg_scale.npz")
    np.savez_compressed(path, **rec)
    print("wrote", path, os.path.getsize(path), "levels", list(full["mg_n"]))


def write_omp_krylov():
    """dphi and the residual norms of the reference's OpenMP 'gmres' and
    'bicgs' (SolveAlgebraicSystemInParallel, one thread) on the pshift and
    momentum systems: their fixed tolerances, not the caller's rrF, end
    these solves."""
    os.environ["OMP_NUM_THREADS"] = "1"
    out = {}
    for case in ("pshift", "momentum"):
        g = np.load(os.path.join(HERE, f"solve_{case}.npz"))
        for solver, tag in ((4, "gmres"), (5, "bicgs")):
            r = po.run_reference(g["nbr"], g["nnbr"], g["ac"], g["anb"], g["bc"], g["fit"], 1e-8, 200, solver,
                                 mode=0)[0]
            out[f"{case}_{tag}_dphi"], out[f"{case}_{tag}_res"] = r["dphi"], r["res"]
    path = os.path.join(HERE, "solve_omp_krylov.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path))


def write_amg_cycles():
    """dphi of the reference's multigrid as the solver with F and W cycles
    (2 pre- and 3 post-sweeps, minNumberofParents 2) on pshift and pbig."""
    out = {}
    for case in ("pshift", "pbig"):
        g = np.load(os.path.join(HERE, f"solve_{case}.npz"))
        for cycle, tag in ((1, "f"), (2, "w")):
            r = po.run_reference(g["nbr"], g["nnbr"], g["ac"], g["anb"], g["bc"], g["fit"], float(g["rrf"]), 30, 0,
                                 mode=0, lmg=1, cycle=cycle, npre=2, npost=3, minparents=2)[0]
            out[f"{case}_{tag}_dphi"] = r["dphi"]
    path = os.path.join(HERE, "solve_amg_cycles.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path))


def main():
    if not po.ref_available():
        sys.exit("reference build missing: run make -C oracle ref")
    if sys.argv[1:] == ["amg_scale"]:
        write_amg_scale()
        return
    if sys.argv[1:] == ["omp_krylov"]:
        write_omp_krylov()
        return
    if sys.argv[1:] == ["amg_cycles"]:
        write_amg_cycles()
        return
    for name, fn in CASES.items():
        nbr, nnbr, ac, anb, bc = fn()
        n = len(ac)
        fit = np.random.default_rng(9).standard_normal(n)
        rec = dict(nbr=nbr, nnbr=nnbr, ac=ac, anb=anb, bc=bc, fit=fit, rrf=1e-8)
        if name != "pbig":
            for solver, tag in ((0, "sor"), (1, "ilu"), (2, "pbcg")):
                r = po.run_reference(nbr, nnbr, ac, anb, bc, fit, 1e-8, 200, solver, mode=1)[0]
                rec[f"{tag}_dphi"], rec[f"{tag}_hist"] = r["dphi"], r["hist"]
                r0 = po.run_reference(nbr, nnbr, ac, anb, bc, fit, 1e-8, 200, solver, mode=0)[0]
                rec[f"{tag}_res"], rec[f"{tag}_fit"] = r0["res"], r0["fit"]
        r = po.run_reference(nbr, nnbr, ac, anb, bc, fit, 1e-8, 30, 0, mode=0, lmg=1, minparents=2)[0]
        rec["amg_dphi"], rec["amg_fit"] = r["dphi"], r["fit"]
        rec.update(flatten_mg(r["mg"]))
        path = os.path.join(HERE, f"solve_{name}.npz")
        np.savez_compressed(path, **rec)
        print("wrote", path, os.path.getsize(path))
    write_amg_scale()
    write_omp_krylov()
    write_amg_cycles()


if __name__ == "__main__":
    main()
