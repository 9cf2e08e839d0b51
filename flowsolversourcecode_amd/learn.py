"""Learned warm start for the p' solve (fs_mpnn_*, include/fsamd_learn.h).

A message-passing network over the cell graph predicts the pressure
correction from bc, the p' operator and the cell volumes; its prediction
starts the p' Krylov solve, which still runs to the reference's stop rule
(SolveEquations/ContinuityEquation/SolveContinuityEquation.for:76), so the
network changes the iteration count and nothing else.  Training data are
(bc, converged p') snapshots the device SIMPLE loop writes into the
network's dataset while it runs.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _check, _dptr, _iptr, _setup_case_symbols, lib, meshgen, Case, MG_PRECOND

_dpp = C.POINTER(C.c_double)
_ipp = C.POINTER(C.c_int)


def _setup(L):
    if getattr(L, "_mpnn_ready", False):
        return
    _setup_case_symbols(L)
    L.fs_mpnn_create.argtypes = [C.c_int, C.c_int, _ipp, C.c_int, C.c_int, C.c_ulonglong, C.c_int,
                                 C.POINTER(C.c_void_p)]
    L.fs_mpnn_destroy.argtypes = [C.c_void_p]
    L.fs_mpnn_param_count.argtypes = [C.c_void_p]
    L.fs_mpnn_get_params.argtypes = [C.c_void_p, _dpp, C.c_int]
    L.fs_mpnn_set_params.argtypes = [C.c_void_p, _dpp, C.c_int]
    L.fs_mpnn_add_samples.argtypes = [C.c_void_p, C.c_int, _dpp, _dpp, _dpp, _dpp, _dpp, _dpp]
    L.fs_mpnn_sample_count.argtypes = [C.c_void_p]
    L.fs_mpnn_get_sample.argtypes = [C.c_void_p, C.c_int, _dpp, _dpp, _dpp, _dpp, _dpp, _dpp]
    L.fs_mpnn_loss_grad.argtypes = [C.c_void_p, _ipp, C.c_int, _dpp, _dpp]
    L.fs_mpnn_train_step.argtypes = [C.c_void_p, _ipp, C.c_int, C.c_double, _dpp]
    L.fs_mpnn_predict.argtypes = [C.c_void_p, _dpp, _dpp, _dpp, _dpp, _dpp, _dpp]
    L.fs_case_set_warm_start.argtypes = [C.c_void_p, C.c_void_p]
    L.fs_case_set_snapshot_sink.argtypes = [C.c_void_p, C.c_void_p]
    L._mpnn_ready = True


class Mpnn:
    """The warm-start network on the topology ``nbr`` ((W, N) reference
    layout: 1-based neighbours, 0 = none), with a device dataset of
    ``capacity`` samples.  use_prev adds the previous SIMPLE iteration's p'
    solution to the inputs (source and geometry only otherwise)."""

    def __init__(self, nbr, layers=3, seed=0, capacity=0, use_prev=False):
        L = lib()
        _setup(L)
        nbr = np.ascontiguousarray(nbr, dtype=np.int32)
        self.w, self.n = nbr.shape
        self.nbr = nbr
        h = C.c_void_p()
        _check(L.fs_mpnn_create(self.n, self.w, _iptr(nbr), layers, int(bool(use_prev)), seed, capacity, C.byref(h)),
               "fs_mpnn_create")
        self.h = h
        self.layers = layers
        self.use_prev = bool(use_prev)
        self.P = L.fs_mpnn_param_count(h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().fs_mpnn_destroy(self.h)
            self.h = None

    @classmethod
    def for_case(cls, case, **kw):
        """A network on the topology of a Case's mesh."""
        return cls(case.topology("enbr"), **kw)

    def params(self):
        out = np.zeros(self.P)
        _check(lib().fs_mpnn_get_params(self.h, _dptr(out), self.P), "fs_mpnn_get_params")
        return out

    def set_params(self, theta):
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        _check(lib().fs_mpnn_set_params(self.h, _dptr(theta), self.P), "fs_mpnn_set_params")

    def add_samples(self, b, x, ac, anb, vol, x_prev=None):
        """Arrays of shape (S, N) (anb: (S, W, N)); x_prev None = zeros."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (b, x, ac, anb, vol)]
        xp = None if x_prev is None else np.ascontiguousarray(x_prev, dtype=np.float64)
        s = arrs[0].shape[0]
        return _check(lib().fs_mpnn_add_samples(self.h, s, *[_dptr(a) for a in arrs], _dptr(xp) if xp is not None
                                                else None), "fs_mpnn_add_samples")

    @property
    def count(self):
        return lib().fs_mpnn_sample_count(self.h)

    def sample(self, k):
        """Sample k as stored on the device (FP32 values) as a dict of FP64 arrays."""
        b, x, ac, vol, xp = (np.zeros(self.n) for _ in range(5))
        anb = np.zeros((self.w, self.n))
        _check(lib().fs_mpnn_get_sample(self.h, k, _dptr(b), _dptr(x), _dptr(ac), _dptr(anb), _dptr(vol), _dptr(xp)),
               "fs_mpnn_get_sample")
        return {"b": b, "x": x, "ac": ac, "anb": anb, "vol": vol, "x_prev": xp}

    def loss_grad(self, ids, grad=True):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        loss = np.zeros(1)
        g = np.zeros(self.P) if grad else None
        _check(lib().fs_mpnn_loss_grad(self.h, _iptr(ids), len(ids), _dptr(loss), _dptr(g) if grad else None),
               "fs_mpnn_loss_grad")
        return float(loss[0]), g

    def train_step(self, ids, lr):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        loss = np.zeros(1)
        _check(lib().fs_mpnn_train_step(self.h, _iptr(ids), len(ids), float(lr), _dptr(loss)), "fs_mpnn_train_step")
        return float(loss[0])

    def train(self, epochs, batch=16, lr=2e-3, ids=None, seed=0, lr_final=None, log=None):
        """Adam over the dataset (or the sample ids given) for some epochs,
        the learning rate decaying geometrically to lr_final.  Returns the
        mean loss of each epoch."""
        rng = np.random.default_rng(seed)
        ids = np.arange(self.count) if ids is None else np.asarray(ids)
        steps = max(1, len(ids) // batch)
        total = epochs * steps
        lr_final = lr * 0.05 if lr_final is None else lr_final
        hist, k = [], 0
        for ep in range(epochs):
            perm = rng.permutation(ids)
            acc = 0.0
            for s in range(steps):
                cur = lr * (lr_final / lr) ** (k / max(1, total - 1))
                acc += self.train_step(perm[s * batch:(s + 1) * batch], cur)
                k += 1
            hist.append(acc / steps)
            if log:
                log(ep, hist[-1])
        return hist

    def predict(self, b, ac, anb, vol, x_prev=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (b, ac, anb, vol)]
        xp = None if x_prev is None else np.ascontiguousarray(x_prev, dtype=np.float64)
        x = np.zeros(self.n)
        _check(lib().fs_mpnn_predict(self.h, *[_dptr(a) for a in arrs], _dptr(xp) if xp is not None else None,
                                     _dptr(x)), "fs_mpnn_predict")
        return x


def attach(case, warm=None, sink=None):
    """Set (or clear, with None) the case's warm-start network and snapshot sink."""
    L = lib()
    _setup(L)
    _check(L.fs_case_set_warm_start(case.h, warm.h if warm is not None else None), "fs_case_set_warm_start")
    _check(L.fs_case_set_snapshot_sink(case.h, sink.h if sink is not None else None), "fs_case_set_snapshot_sink")


def cavity_case(n, re, perturb, seed, psolver="bicgs", fp32=1):
    """The lid-driven cavity of n^3 hexahedra at Reynolds number re (lid
    speed 1, side 1, rho 1), interior nodes perturbed by U(-perturb, perturb)
    of the spacing, SIMPLE settings of the benchmark."""
    pm = meshgen.box_hex(n, n, n, perturb=perturb, seed=seed)
    bu, bv, bw = meshgen.lid_driven_cavity_walls(pm)
    case = Case(pm, bu, bv, bw, rho=1.0, mu=1.0 / re)
    case.configure(urf_u=0.7, urf_p=0.3, msolver="sor", mrrf=0.1, mitmax=10, psolver=psolver, prrf=1e-8,
                   pitmax=200, plmg=MG_PRECOND)
    case.set_option("mg_fp32", fp32)
    return case


def collect(net, runs, n=16, iterations=60, log=None):
    """Runs cavity cases (list of (re, perturb, seed)) with the network as the
    snapshot sink: every p' solve appends (bc, p', ac, anb, vol)."""
    for re, perturb, seed in runs:
        case = cavity_case(n, re, perturb, seed)
        attach(case, sink=net)
        case.iterate(iterations)
        attach(case)
        if log:
            log(re, perturb, seed, net.count)
    return net.count


def iterations_with_and_without(net, run, n=16, iterations=60):
    """p' Krylov iterations of the same cavity run from a zero start and from
    the network's prediction; also the largest field difference."""
    out = {}
    fields = {}
    for mode in ("zero", "learned"):
        case = cavity_case(n, *run)
        attach(case, warm=net if mode == "learned" else None)
        its = []
        for _ in range(iterations):
            st = case.iterate(1)
            its.append(st["pc_iterations"])
        out[mode] = its
        fields[mode] = {k: case.get(k) for k in ("u", "v", "w", "p")}
    diff = max(np.abs(fields["zero"][k] - fields["learned"][k]).max() / max(np.abs(fields["zero"][k]).max(), 1e-30)
               for k in fields["zero"])
    return out, diff


def c1_case(re, perturb, seed, n=64, warm_previous=0, psolver="pbcg"):
    """A member of the two-dimensional cavity family of the C1 settings:
    n x n x 1 hexahedra, interior nodes perturbed by U(-perturb, perturb) of
    the spacing, lid speed 1 (Reynolds number re, rho 1), urf 0.9 / 0.1,
    momentum by Gauss-Seidel (rrF 0.1, <= 10 sweeps), p' by PBCG (C1's
    solver; `psolver` another Krylov method) with one AMG V-cycle as its
    preconditioner to 1e-8."""
    m, (bu, bv, bw), types = meshgen.cavity_2d(n, perturb=perturb, seed=seed)
    case = Case(m, bu, bv, bw, rho=1.0, mu=1.0 / re)
    case.configure(urf_u=0.9, urf_p=0.1, msolver="sor", mrrf=0.1, mitmax=10, psolver=psolver, prrf=1e-8,
                   pitmax=500, plmg=MG_PRECOND)
    case.set_patch_types(m.patch_sizes, types)
    case.set_option("warm_previous", warm_previous)
    return case


def c4_study(snapshots=20000, per_run=200, epochs=6, batch=16, heldout=4, iterations=100, seed=0, log=None,
             eval_psolver="bicgs"):
    """The learned form at its configured scale: `snapshots` (bc, p') pairs
    written by the SIMPLE loops of C1-family cavities (Re U(100, 2000),
    perturbation U(0, 0.2), fresh seeds; `per_run` outer iterations each,
    the previous p' among the inputs; their p' solved by BiCGSTAB + AMG to
    1e-8, which converges to the same p' as C1's PBCG in fewer launches), the
    network trained on the device, then held-out members of the family run
    `iterations` outer iterations three times: p' from zero, from the
    previous iteration's p' and from the network's guess.  The held-out
    solves use BiCGSTAB + AMG (`eval_psolver`): C1's PBCG + AMG stops at
    itmax 500 on every solve of this singular system without reaching 1e-8,
    so its iteration counts would not move.  Returns a dict of the p'
    iteration totals."""
    rng = np.random.default_rng(seed)
    probe = c1_case(1000.0, 0.1, 1)
    net = Mpnn.for_case(probe, layers=3, seed=7, capacity=snapshots, use_prev=True)
    del probe
    runs = 0
    while net.count < snapshots:
        re, pert = float(rng.uniform(100.0, 2000.0)), float(rng.uniform(0.0, 0.2))
        case = c1_case(re, pert, 1000 + runs, psolver="bicgs")
        attach(case, sink=net)
        case.iterate(min(per_run, snapshots - net.count))
        attach(case)
        runs += 1
        if log:
            log("collect", runs, net.count)
    hist = net.train(epochs=epochs, batch=batch, lr=2e-3, seed=seed, log=(lambda e, l: log("epoch", e, l))
                     if log else None)
    out = {"snapshots": net.count, "runs": runs, "epochs": epochs, "loss_first_epoch": hist[0],
           "loss_last_epoch": hist[-1], "heldout": []}
    for h in range(heldout):
        re, pert = float(rng.uniform(100.0, 2000.0)), float(rng.uniform(0.0, 0.2))
        rec = {"re": re, "perturb": pert}
        for mode in ("zero", "previous", "learned"):
            case = c1_case(re, pert, 900000 + h, warm_previous=int(mode == "previous"), psolver=eval_psolver)
            if mode == "learned":
                attach(case, warm=net)
            st = case.iterate(iterations)
            attach(case)
            rec[mode] = int(st["pc_iterations"])
            rec[mode + "_unconverged"] = int(st["pc_unconverged"])
        out["heldout"].append(rec)
        if log:
            log("heldout", rec)
    for mode in ("zero", "previous", "learned"):
        out[mode] = sum(r[mode] for r in out["heldout"])
    return out


if __name__ == "__main__":
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="the learned warm start on the C1 cavity family (C4)")
    ap.add_argument("--snapshots", type=int, default=20000)
    ap.add_argument("--per-run", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--heldout", type=int, default=4)
    ap.add_argument("--iterations", type=int, default=100)
    a = ap.parse_args()
    res = c4_study(a.snapshots, a.per_run, a.epochs, heldout=a.heldout, iterations=a.iterations,
                   log=lambda *x: print("#", *x, file=sys.stderr, flush=True))
    print(json.dumps(res))
