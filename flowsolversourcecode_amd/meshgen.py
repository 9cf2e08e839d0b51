"""Seeded synthetic meshes in polyMesh form (the reference reads OpenFOAM
polyMesh directories, SolveEquations/GridAndGeometry/ReadFoamMesh.f90:31-250).

``box_hex(nx, ny, nz, ...)`` builds an nx*ny*nz hexahedral block on
[0,lx]x[0,ly]x[0,lz] with the polyMesh conventions the reference relies on:

* interior faces first, ordered by owner and then by neighbour (owner <
  neighbour), with the face normal (right-hand node order) pointing from
  owner to neighbour;
* boundary faces after them, grouped by patch in the order xmin, xmax, ymin,
  ymax, zmin, zmax, each patch ordered by owner cell, normals outward.

Interior nodes can be moved by ``perturb`` times the local spacing per axis,
uniformly in [-perturb, perturb], from ``numpy.random.default_rng(seed)``
(boundary nodes stay on the box faces so the domain is unchanged).
"""
from __future__ import annotations

import numpy as np

PATCHES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


class PolyMesh:
    def __init__(self, points, face_ptr, face_nodes, owner, neighbour, patch_names, patch_sizes, shape):
        self.points = points            # (P, 3) float64
        self.face_ptr = face_ptr        # (F+1,) int32
        self.face_nodes = face_nodes    # (sum,) int32, 0-based
        self.owner = owner              # (F,) int32, 0-based
        self.neighbour = neighbour      # (Fi,) int32, 0-based
        self.patch_names = list(patch_names)
        self.patch_sizes = np.asarray(patch_sizes, np.int32)
        self.shape = shape

    @property
    def n_cells(self):
        return int(max(self.owner.max(), self.neighbour.max() if len(self.neighbour) else -1)) + 1

    @property
    def n_faces(self):
        return len(self.owner)

    @property
    def n_internal(self):
        return len(self.neighbour)

    def patch_face_range(self, name):
        i = self.patch_names.index(name)
        start = self.n_internal + int(self.patch_sizes[:i].sum())
        return start, start + int(self.patch_sizes[i])


def box_hex(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, perturb=0.0, seed=0):
    ix = np.arange(nx + 1)
    iy = np.arange(ny + 1)
    iz = np.arange(nz + 1)
    X, Y, Z = np.meshgrid(ix * (lx / nx), iy * (ly / ny), iz * (lz / nz), indexing="ij")
    # point index p(i,j,k) = i + (nx+1)*(j + (ny+1)*k): order points with i fastest
    pts = np.stack([X.transpose(2, 1, 0).ravel(), Y.transpose(2, 1, 0).ravel(), Z.transpose(2, 1, 0).ravel()], 1)
    if perturb > 0.0:
        rng = np.random.default_rng(seed)
        I, J, K = np.meshgrid(ix, iy, iz, indexing="ij")
        I, J, K = (a.transpose(2, 1, 0).ravel() for a in (I, J, K))
        interior = (I > 0) & (I < nx) & (J > 0) & (J < ny) & (K > 0) & (K < nz)
        d = rng.uniform(-perturb, perturb, size=(pts.shape[0], 3))
        d *= np.array([lx / nx, ly / ny, lz / nz])
        pts = pts + d * interior[:, None]

    def P(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    def C(i, j, k):
        return i + nx * (j + ny * k)

    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci, cj, ck = (a.transpose(2, 1, 0).ravel() for a in (ci, cj, ck))  # cell order, i fastest
    cell = C(ci, cj, ck)
    # interior faces of each cell towards +x, +y, +z (increasing neighbour index)
    fx_mask = ci < nx - 1
    fy_mask = cj < ny - 1
    fz_mask = ck < nz - 1
    quads, own, nbr, key = [], [], [], []
    # +x face at i+1: (i+1,j,k),(i+1,j+1,k),(i+1,j+1,k+1),(i+1,j,k+1): normal +x
    q = np.stack([P(ci + 1, cj, ck), P(ci + 1, cj + 1, ck), P(ci + 1, cj + 1, ck + 1), P(ci + 1, cj, ck + 1)], 1)
    quads.append(q[fx_mask]); own.append(cell[fx_mask]); nbr.append(cell[fx_mask] + 1)
    # +y face at j+1: (i,j+1,k),(i,j+1,k+1),(i+1,j+1,k+1),(i+1,j+1,k): normal +y
    q = np.stack([P(ci, cj + 1, ck), P(ci, cj + 1, ck + 1), P(ci + 1, cj + 1, ck + 1), P(ci + 1, cj + 1, ck)], 1)
    quads.append(q[fy_mask]); own.append(cell[fy_mask]); nbr.append(cell[fy_mask] + nx)
    # +z face at k+1: (i,j,k+1),(i+1,j,k+1),(i+1,j+1,k+1),(i,j+1,k+1): normal +z
    q = np.stack([P(ci, cj, ck + 1), P(ci + 1, cj, ck + 1), P(ci + 1, cj + 1, ck + 1), P(ci, cj + 1, ck + 1)], 1)
    quads.append(q[fz_mask]); own.append(cell[fz_mask]); nbr.append(cell[fz_mask] + nx * ny)
    q_int = np.concatenate(quads)
    o_int = np.concatenate(own)
    n_int = np.concatenate(nbr)
    order = np.lexsort((n_int, o_int))  # by owner, then neighbour
    q_int, o_int, n_int = q_int[order], o_int[order], n_int[order]

    bq, bo, sizes = [], [], []
    # outward normals
    m = ci == 0  # xmin: normal -x (reverse of the +x quad)
    q = np.stack([P(ci, cj, ck), P(ci, cj, ck + 1), P(ci, cj + 1, ck + 1), P(ci, cj + 1, ck)], 1)
    bq.append(q[m]); bo.append(cell[m]); sizes.append(int(m.sum()))
    m = ci == nx - 1
    q = np.stack([P(ci + 1, cj, ck), P(ci + 1, cj + 1, ck), P(ci + 1, cj + 1, ck + 1), P(ci + 1, cj, ck + 1)], 1)
    bq.append(q[m]); bo.append(cell[m]); sizes.append(int(m.sum()))
    m = cj == 0
    q = np.stack([P(ci, cj, ck), P(ci + 1, cj, ck), P(ci + 1, cj, ck + 1), P(ci, cj, ck + 1)], 1)
    bq.append(q[m]); bo.append(cell[m]); sizes.append(int(m.sum()))
    m = cj == ny - 1
    q = np.stack([P(ci, cj + 1, ck), P(ci, cj + 1, ck + 1), P(ci + 1, cj + 1, ck + 1), P(ci + 1, cj + 1, ck)], 1)
    bq.append(q[m]); bo.append(cell[m]); sizes.append(int(m.sum()))
    m = ck == 0
    q = np.stack([P(ci, cj, ck), P(ci, cj + 1, ck), P(ci + 1, cj + 1, ck), P(ci + 1, cj, ck)], 1)
    bq.append(q[m]); bo.append(cell[m]); sizes.append(int(m.sum()))
    m = ck == nz - 1
    q = np.stack([P(ci, cj, ck + 1), P(ci + 1, cj, ck + 1), P(ci + 1, cj + 1, ck + 1), P(ci, cj + 1, ck + 1)], 1)
    bq.append(q[m]); bo.append(cell[m]); sizes.append(int(m.sum()))

    quads = np.concatenate([q_int] + bq).astype(np.int32)
    owner = np.concatenate([o_int] + bo).astype(np.int32)
    face_ptr = (np.arange(len(quads) + 1) * 4).astype(np.int32)
    return PolyMesh(pts.astype(np.float64), face_ptr, quads.ravel().copy(), owner, n_int.astype(np.int32),
                    PATCHES, sizes, (nx, ny, nz))


# Kuhn split of a hexahedron into 6 tetrahedra along its diagonal v0-v6
# (local vertices v0 (i,j,k), v1 (i+1,j,k), v2 (i+1,j+1,k), v3 (i,j+1,k) and
# v4..v7 the same at k+1): the 6 monotone edge paths from v0 to v6.  Every
# hexahedron uses the same diagonal direction, so the split is conforming.
_KUHN = ((0, 1, 2, 6), (0, 1, 5, 6), (0, 4, 5, 6), (0, 4, 7, 6), (0, 3, 7, 6), (0, 3, 2, 6))


def box_tet(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, perturb=0.0, seed=0):
    """nx*ny*nz hexahedra of box_hex (same points, same perturbation) each
    split into 6 tetrahedra (Kuhn), as a polyMesh with box_hex's conventions:
    interior faces ordered by owner then neighbour with the normal from owner
    to neighbour, boundary faces grouped by patch (xmin, xmax, ymin, ymax,
    zmin, zmax), each patch ordered by owner, normals outward.  Cell 6 h + t
    is tetrahedron t of hexahedron h (h = i + nx (j + ny k))."""
    hexm = box_hex(nx, ny, nz, lx, ly, lz, perturb, seed)
    pts = hexm.points
    ix = np.arange(nx + 1)
    I, J, K = np.meshgrid(ix, np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    I, J, K = (a.transpose(2, 1, 0).ravel() for a in (I, J, K))

    def P(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci, cj, ck = (a.transpose(2, 1, 0).ravel() for a in (ci, cj, ck))  # hex order, i fastest
    corner = np.stack([P(ci, cj, ck), P(ci + 1, cj, ck), P(ci + 1, cj + 1, ck), P(ci, cj + 1, ck),
                       P(ci, cj, ck + 1), P(ci + 1, cj, ck + 1), P(ci + 1, cj + 1, ck + 1),
                       P(ci, cj + 1, ck + 1)], 1)
    nh = len(ci)
    tets = np.stack([corner[:, list(t)] for t in _KUHN], 1).reshape(6 * nh, 4)
    ncell = 6 * nh
    cent = pts[tets].mean(axis=1)
    # the 4 faces of every tetrahedron (the vertex opposite each one left out)
    tri = np.concatenate([tets[:, [1, 2, 3]], tets[:, [0, 2, 3]], tets[:, [0, 1, 3]], tets[:, [0, 1, 2]]])
    cell = np.tile(np.arange(ncell), 4)
    key = np.sort(tri, axis=1)
    order = np.lexsort((key[:, 2], key[:, 1], key[:, 0]))
    key, tri, cell = key[order], tri[order], cell[order]
    same = np.all(key[1:] == key[:-1], axis=1)
    first = np.nonzero(same)[0]                  # pairs (first, first + 1)
    paired = np.zeros(len(key), bool)
    paired[first] = paired[first + 1] = True
    c0, c1 = cell[first], cell[first + 1]
    own = np.minimum(c0, c1)
    nei = np.maximum(c0, c1)
    ftri = tri[first]

    def orient(t, frm, to):
        a, b, c = pts[t[:, 0]], pts[t[:, 1]], pts[t[:, 2]]
        nrm = np.cross(b - a, c - a)
        flip = (nrm * (to - frm)).sum(1) < 0
        t = t.copy()
        t[flip, 1], t[flip, 2] = t[flip, 2], t[flip, 1].copy()
        return t

    ftri = orient(ftri, cent[own], cent[nei])
    o = np.lexsort((nei, own))
    ftri, own, nei = ftri[o], own[o], nei[o]
    # boundary faces: outward normals, then patches
    btri, bcell = tri[~paired], cell[~paired]
    fc = pts[btri].mean(axis=1)
    btri = orient(btri, cent[bcell], fc)
    bI, bJ, bK = I[btri], J[btri], K[btri]
    sides = [np.all(bI == 0, 1), np.all(bI == nx, 1), np.all(bJ == 0, 1), np.all(bJ == ny, 1),
             np.all(bK == 0, 1), np.all(bK == nz, 1)]
    if not np.all(np.sum(sides, axis=0) == 1):
        raise RuntimeError("a boundary face lies on no single box side")
    bq, bo, sizes = [], [], []
    for m in sides:
        idx = np.nonzero(m)[0]
        idx = idx[np.argsort(bcell[idx], kind="stable")]
        bq.append(btri[idx])
        bo.append(bcell[idx])
        sizes.append(len(idx))
    faces = np.concatenate([ftri] + bq).astype(np.int32)
    owner = np.concatenate([own] + bo).astype(np.int32)
    face_ptr = (np.arange(len(faces) + 1) * 3).astype(np.int32)
    return PolyMesh(pts.copy(), face_ptr, faces.ravel().copy(), owner, nei.astype(np.int32), PATCHES, sizes,
                    (nx, ny, nz))


def graded_coordinates(n, length, ratio):
    """n + 1 node coordinates on [0, length], cells symmetric about the middle,
    each growing by `ratio` from either end toward the middle (n even)."""
    if n % 2:
        raise ValueError("graded_coordinates needs an even cell count")
    h = ratio ** np.arange(n // 2)
    h = np.concatenate([h, h[::-1]])
    y = np.concatenate([[0.0], np.cumsum(h)])
    return y * (length / y[-1])


def channel_tet(nx, ny, nz, lx=2.0 * np.pi, ly=2.0, lz=np.pi, grading=1.05, cells="tet"):
    """A plane channel for the periodic RANS case: nx*ny*nz hexahedra on
    [0,lx]x[0,ly]x[0,lz], graded in y by `grading` toward both walls (cells
    growing by that ratio from each wall to the centre plane), each split
    into 6 tetrahedra (Kuhn, box_tet's cell and face order).  x is the
    streamwise direction; the Kuhn split triangulates the x = 0 and x = lx
    planes alike, face for face.  Patches, in this order: outlet (x = lx) and
    inlet (x = 0), a translational periodic pair; walls (y = 0 and y = ly);
    sides (z = 0 and z = lz, symmetry planes).  Returns (mesh, patch types,
    pairs, translations): pairs[q] = the partner patch of patch q (-1 for
    none) and translations[q] = the vector that carries the partner patch's
    cells across patch q (the reference's x/y/zTranslation of the partner),
    the form SetBCPatches and RotationMatrixCoefficients read.  cells="hex"
    keeps the hexahedra unsplit."""
    m = (box_tet if cells == "tet" else box_hex)(nx, ny, nz, lx, ly, lz)
    pts = m.points.copy()
    j = (np.arange(len(pts)) // (nx + 1)) % (ny + 1)
    pts[:, 1] = graded_coordinates(ny, ly, grading)[j]
    m = PolyMesh(pts, m.face_ptr, m.face_nodes, m.owner, m.neighbour, m.patch_names, m.patch_sizes, m.shape)
    m = regroup(m, [("outlet", ["xmax"]), ("inlet", ["xmin"]), ("walls", ["ymin", "ymax"]),
                    ("sides", ["zmin", "zmax"])])
    types = ["periodic", "periodic", "wall", "symmetry"]
    pairs = [1, 0, -1, -1]
    translations = [(-lx, 0.0, 0.0), (lx, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)]
    return m, types, pairs, translations


def lid_driven_cavity_walls(mesh, lid="ymax", lid_velocity=(1.0, 0.0, 0.0)):
    """Wall velocity of every boundary face: the lid patch moves, the other
    walls are at rest.  Returns (bu, bv, bw), one entry per boundary face."""
    nb = mesh.n_faces - mesh.n_internal
    bu, bv, bw = np.zeros(nb), np.zeros(nb), np.zeros(nb)
    s, e = mesh.patch_face_range(lid)
    s -= mesh.n_internal
    e -= mesh.n_internal
    bu[s:e], bv[s:e], bw[s:e] = lid_velocity
    return bu, bv, bw


def regroup(mesh, groups):
    """The mesh with its boundary faces regrouped into new patches: groups is
    a list of (new name, [old patch names]); each new patch holds the faces of
    its old patches in the order given.  Interior faces are unchanged."""
    ni = mesh.n_internal
    used = [n for _, olds in groups for n in olds]
    if sorted(used) != sorted(mesh.patch_names):
        raise ValueError("groups must use every patch exactly once")
    order, sizes = [], []
    for _, olds in groups:
        k = 0
        for name in olds:
            s, e = mesh.patch_face_range(name)
            order.extend(range(s, e))
            k += e - s
        sizes.append(k)
    faces = np.concatenate([np.arange(ni), np.asarray(order, dtype=np.int64)])
    counts = np.diff(mesh.face_ptr)[faces]
    ptr = np.zeros(len(faces) + 1, np.int32)
    ptr[1:] = np.cumsum(counts)
    nodes = np.concatenate([mesh.face_nodes[mesh.face_ptr[f]:mesh.face_ptr[f + 1]] for f in faces]).astype(np.int32)
    return PolyMesh(mesh.points.copy(), ptr, nodes, mesh.owner[faces].astype(np.int32), mesh.neighbour.copy(),
                    [g for g, _ in groups], sizes, mesh.shape)


def rotate_x(mesh, angle):
    """The mesh rotated by angle (radians) about the x axis through the origin."""
    c, s = np.cos(angle), np.sin(angle)
    R = np.array([[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]])
    return PolyMesh(mesh.points @ R.T, mesh.face_ptr.copy(), mesh.face_nodes.copy(), mesh.owner.copy(),
                    mesh.neighbour.copy(), mesh.patch_names, mesh.patch_sizes.copy(), mesh.shape)


def quarter_turn_x(mesh):
    """The mesh turned by a right angle about the x axis, exactly:
    (x, y, z) -> (x, -z, y)."""
    P = mesh.points
    pts = np.stack([P[:, 0], -P[:, 2], P[:, 1]], 1)
    return PolyMesh(pts, mesh.face_ptr.copy(), mesh.face_nodes.copy(), mesh.owner.copy(), mesh.neighbour.copy(),
                    mesh.patch_names, mesh.patch_sizes.copy(), mesh.shape)


def cavity_2d(n, perturb=0.0, seed=0, split_front_back=False):
    """The two-dimensional lid-driven cavity: n x n x 1 hexahedra on
    [0,1]^2 x [0,1/n], patches lid (y = 1, moving at u = 1), walls (x = 0,
    x = 1, y = 0; at rest) and frontBack (z = 0 and z = 1/n, symmetry), or
    front and back as two symmetry patches.  Returns (mesh, (bu, bv, bw),
    patch types)."""
    m = box_hex(n, n, 1, lz=1.0 / n, perturb=perturb, seed=seed)
    fb = [("front", ["zmin"]), ("back", ["zmax"])] if split_front_back else [("frontBack", ["zmin", "zmax"])]
    m = regroup(m, [("lid", ["ymax"]), ("walls", ["xmin", "xmax", "ymin"])] + fb)
    types = ["wall", "wall"] + ["symmetry"] * len(fb)
    return m, lid_driven_cavity_walls(m, lid="lid"), types


# Ghia, Ghia & Shin, J. Comput. Phys. 48 (1982) 387-411, Tables I and II,
# Re = 100: (y, u) on the vertical centre line x = 0.5 and (x, v) on the
# horizontal centre line y = 0.5 of the two-dimensional lid-driven cavity.
GHIA_RE100_U = np.array([
    [0.9766, 0.84123], [0.9688, 0.78871], [0.9609, 0.73722], [0.9531, 0.68717], [0.8516, 0.23151],
    [0.7344, 0.00332], [0.6172, -0.13641], [0.5000, -0.20581], [0.4531, -0.21090], [0.2813, -0.15662],
    [0.1719, -0.10150], [0.1016, -0.06434], [0.0703, -0.04775], [0.0625, -0.04192], [0.0547, -0.03717]])
GHIA_RE100_V = np.array([
    [0.9688, -0.05906], [0.9609, -0.07391], [0.9531, -0.08864], [0.9453, -0.10313], [0.9063, -0.16914],
    [0.8594, -0.22445], [0.8047, -0.24533], [0.5000, 0.05454], [0.2344, 0.17527], [0.2266, 0.17507],
    [0.1563, 0.16077], [0.0938, 0.12317], [0.0781, 0.10890], [0.0703, 0.10091], [0.0625, 0.09233]])


def centreline_errors(u, v, n):
    """Largest deviation of the cavity_2d(n) centre-line velocities (cell
    (i, j) is i + n j; the centre lines fall between cell columns, n even)
    from the Re = 100 benchmark values."""
    u, v = np.asarray(u).reshape(n, n), np.asarray(v).reshape(n, n)
    yc = (np.arange(n) + 0.5) / n
    uc = 0.5 * (u[:, n // 2 - 1] + u[:, n // 2])
    vc = 0.5 * (v[n // 2 - 1, :] + v[n // 2, :])
    eu = np.abs(np.interp(GHIA_RE100_U[:, 0], yc, uc) - GHIA_RE100_U[:, 1]).max()
    ev = np.abs(np.interp(GHIA_RE100_V[:, 0], yc, vc) - GHIA_RE100_V[:, 1]).max()
    return float(eu), float(ev)


def write_polymesh(mesh, directory):
    """Writes the mesh as an OpenFOAM polyMesh directory in ASCII (the files
    the reference's Readit reads: points, faces, owner, neighbour, boundary)."""
    import os
    os.makedirs(directory, exist_ok=True)

    def header(f, cls, obj):
        f.write("FoamFile\n{\n    version     2.0;\n    format      ascii;\n"
                f"    class       {cls};\n    object      {obj};\n}}\n\n")

    with open(os.path.join(directory, "points"), "w") as f:
        header(f, "vectorField", "points")
        f.write(f"{len(mesh.points)}\n(\n")
        for p in mesh.points:
            f.write(f"({float(p[0])!r} {float(p[1])!r} {float(p[2])!r})\n")
        f.write(")\n")
    with open(os.path.join(directory, "faces"), "w") as f:
        header(f, "faceList", "faces")
        f.write(f"{mesh.n_faces}\n(\n")
        for k in range(mesh.n_faces):
            nodes = mesh.face_nodes[mesh.face_ptr[k]:mesh.face_ptr[k + 1]]
            f.write(f"{len(nodes)}(" + " ".join(str(int(v)) for v in nodes) + ")\n")
        f.write(")\n")
    for name, arr in (("owner", mesh.owner), ("neighbour", mesh.neighbour)):
        with open(os.path.join(directory, name), "w") as f:
            header(f, "labelList", name)
            f.write(f"{len(arr)}\n(\n")
            f.write("\n".join(str(int(v)) for v in arr))
            f.write("\n)\n")
    with open(os.path.join(directory, "boundary"), "w") as f:
        header(f, "polyBoundaryMesh", "boundary")
        f.write(f"{len(mesh.patch_names)}\n(\n")
        start = mesh.n_internal
        for name, size in zip(mesh.patch_names, mesh.patch_sizes):
            f.write(f"    {name}\n    {{\n        type            wall;\n        inGroups        List<word> 1(wall);\n"
                    f"        nFaces          {int(size)};\n        startFace       {start};\n    }}\n")
            start += int(size)
        f.write(")\n")
