"""Write golden vectors for the linear-solve path from the reference build.

Runs oracle/_ref/ref_solve.x (the reference's own SolveEquation, SOR, D-ILU,
PBCG and algebraic-multigrid sources, compiled by oracle/build_ref.sh) on small
seeded systems from tests/systems.py and stores inputs and outputs as .npz.
Run in the build container only (needs /root/reference):

    make -C oracle ref && python tests/golden/make_solve_golden.py
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


def main():
    if not po.ref_available():
        sys.exit("reference build missing: run make -C oracle ref")
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


if __name__ == "__main__":
    main()
