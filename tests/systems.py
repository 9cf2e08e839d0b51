"""Synthetic ELL systems in the reference's layout for solver parity tests.

Cells of an nx*ny*nz block are numbered x fastest.  Each cell's slots list its
neighbours in the order -x,+x,-y,+y,-z,+z with boundary slots kept in place as
0 (the reference's ElementNeighbor convention: 1-based, 0 = boundary face).
"""
import numpy as np


def structured_pattern(nx, ny, nz):
    n = nx * ny * nz
    w = 6
    nbr = np.zeros((w, n), np.int32)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = i + nx * (j + ny * k)
                for s, (di, dj, dk) in enumerate(((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1))):
                    ii, jj, kk = i + di, j + dj, k + dk
                    if 0 <= ii < nx and 0 <= jj < ny and 0 <= kk < nz:
                        nbr[s, c] = ii + nx * (jj + ny * kk) + 1
    nnbr = np.full(n, w, np.int32)
    return nbr, nnbr


def mirror_slots(nbr):
    """slot of cell i in neighbour k's list, for every (slot, cell)."""
    w, n = nbr.shape
    mir = np.full((w, n), -1, np.int32)
    for s in range(w):
        for c in range(n):
            k = nbr[s, c]
            if k:
                for t in range(w):
                    if nbr[t, k - 1] == c + 1:
                        mir[s, c] = t
    return mir


def pressure_like(nx, ny, nz, seed, shift=0.0):
    """Symmetric M-matrix: anb = -g on each interior face (g random positive,
    identical from both sides), ac = sum(g) + shift (shift 0 = pure Neumann)."""
    rng = np.random.default_rng(seed)
    nbr, nnbr = structured_pattern(nx, ny, nz)
    w, n = nbr.shape
    anb = np.zeros((w, n))
    mir = mirror_slots(nbr)
    for s in range(w):
        for c in range(n):
            k = nbr[s, c]
            if k and c + 1 < k:
                g = rng.uniform(0.5, 2.0)
                anb[s, c] = -g
                anb[mir[s, c], k - 1] = -g
    ac = -anb.sum(axis=0) + shift
    bc = rng.standard_normal(n)
    if shift == 0.0:
        bc -= bc.mean()
    return nbr, nnbr, ac, anb, bc


def quantized_pressure(nx, ny, nz, seed):
    """pressure_like with every coupling rounded to a multiple of 1/256 (at
    least 1/256 in magnitude), ac = -sum(anb) and a zero-mean source of
    multiples of 1/8: every Galerkin sum of its coarse levels is exact in
    FP64, so a fixture of its hierarchy can hold only the maps and the coarse
    operators and rebuild the system from the seed."""
    nbr, nnbr, ac, anb, bc = pressure_like(nx, ny, nz, seed)
    anb = np.where(nbr > 0, np.minimum(-np.round(-anb * 256) / 256, -1 / 256), 0.0)
    ac = -anb.sum(axis=0)
    bc = np.random.default_rng(seed + 1000).integers(-8, 9, len(ac)) / 8.0
    bc = bc - bc.mean()
    return nbr, nnbr, ac, anb, bc


def momentum_like(nx, ny, nz, seed, urf=0.7):
    """Non-symmetric upwind convection-diffusion with implicit under-relaxation
    (ac/urf), the shape of a momentum system."""
    rng = np.random.default_rng(seed)
    nbr, nnbr = structured_pattern(nx, ny, nz)
    w, n = nbr.shape
    anb = np.zeros((w, n))
    ac = np.zeros(n)
    mir = mirror_slots(nbr)
    for s in range(w):
        for c in range(n):
            k = nbr[s, c]
            if k and c + 1 < k:
                d = rng.uniform(0.2, 1.0)
                m = rng.uniform(-1.0, 1.0)
                # owner c: FluxCf = d + max(m,0), FluxFf = -d - max(-m,0)
                anb[s, c] = -d - max(-m, 0.0)
                ac[c] += d + max(m, 0.0)
                anb[mir[s, c], k - 1] = -d - max(m, 0.0)
                ac[k - 1] += d + max(-m, 0.0)
    ac += rng.uniform(0.1, 0.5, n)  # wall / boundary contributions
    ac /= urf
    bc = rng.standard_normal(n)
    return nbr, nnbr, ac, anb, bc


def dense(nbr, nnbr, ac, anb):
    w, n = nbr.shape
    A = np.diag(np.asarray(ac, float).copy())
    for s in range(w):
        for c in range(n):
            if s < nnbr[c] and nbr[s, c]:
                A[c, nbr[s, c] - 1] += anb[s, c]
    return A


def sparse(nbr, nnbr, ac, anb):
    """The same operator as a scipy CSR matrix (for systems too large for dense)."""
    import scipy.sparse as sp
    w, n = nbr.shape
    rows = np.repeat(np.arange(n)[None, :], w, 0)
    m = (nbr > 0) & (np.arange(w)[:, None] < np.asarray(nnbr)[None, :])
    A = sp.csr_matrix((np.asarray(anb)[m], (rows[m], nbr[m] - 1)), shape=(n, n))
    return (A + sp.diags(np.asarray(ac, float))).tocsr()
