"""The kernel clock (fs_kernel_clock_*, include/fsamd.h): HIP events recorded
by the solve loop's own dispatches, for the roofline figure of the bench.
Counts and bytes are fixed by the loop's structure: one ELL product per
BiCGSTAB iteration, the algorithmic bytes of one product per launch (with
16-bit neighbour offsets); a
stride of k times every k-th launch."""
import numpy as np
import pytest

import systems

fs = pytest.importorskip("flowsolversourcecode_amd")
pytestmark = pytest.mark.gpu


def _run(stride):
    nbr, nnbr, ac, anb, bc = systems.pressure_like(32, 24, 20, 11, shift=0.0)
    m = fs.Mesh(nbr, nnbr)
    m.kernel_clock_start(["spmv", "fine_sweep"], capacity=4096, stride=stride)
    rc, fit, dphi, res4, st = m.solve_equation("pressc", np.zeros(len(ac)), 1e-8, 400, "bicgs", fs.MG_PRECOND,
                                               ac, anb, bc)
    assert rc == 0
    out = {k: m.kernel_clock_read(k) for k in ("spmv", "fine_sweep")}
    m.kernel_clock_start("off")
    return st, out, nbr, ac


def test_clock_counts_and_bytes():
    st, c, nbr, ac = _run(1)
    n, w = len(ac), nbr.shape[0]
    nnz = int((nbr > 0).sum())
    sp = c["spmv"]
    # one product per iteration that did work; later launches of the last
    # batch return at once and are counted apart
    assert sp["launches"] == st["iterations"]
    # neighbour offsets of this numbering fit 16 bits: 2 bytes per slot
    assert sp["bytes"] == pytest.approx(8.0 * n + 2.0 * n * w + 8.0 * nnz + 24.0 * n)
    assert sp["avg_ms"] > 0.0
    assert c["fine_sweep"]["launches"] > 0 and c["fine_sweep"]["avg_ms"] > 0.0


def test_clock_stride_samples_launches():
    st1, c1, _, _ = _run(1)
    st3, c3, _, _ = _run(3)
    assert st1["iterations"] == st3["iterations"]
    total = c1["fine_sweep"]["launches"] + c1["fine_sweep"]["untimed"]
    sampled = c3["fine_sweep"]["launches"] + c3["fine_sweep"]["untimed"]
    assert sampled == (total + 2) // 3
