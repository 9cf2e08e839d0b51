// Mesh topology and geometry in the reference's conventions (host side).
#pragma once

#include <vector>

namespace fs {

// polyMesh in memory, 0-based: faces 0..nif-1 interior (owner < neighbour by
// the polyMesh convention), then boundary faces patch by patch.
struct PolyInput {
  int npoints = 0, nfaces = 0, nif = 0;
  const double *x = nullptr, *y = nullptr, *z = nullptr;
  const int *fptr = nullptr, *fnodes = nullptr, *owner = nullptr, *neigh = nullptr;
};

// Face quantities over all faces (interior first): the reference's interior
// arrays (FaceAreax, FaceCentroidx, DistanceCF*, GFactorCF, FaceE*, gDiff,
// FaceT*) and boundary arrays (BFaceArea*, BDistanceCF*, BgDiff, ...) in one
// numbering.  For a boundary face d* is the owner-to-face vector and g = 1.
struct FaceGeom {
  std::vector<double> Sx, Sy, Sz, A, cx, cy, cz, nx, ny, nz, d, dx, dy, dz, ex, ey, ez, g, Ex,
      Ey, Ez, E, gdiff, Tx, Ty, Tz, T;
  std::vector<double> kx, ky, kz; // r_f - r_f' (interior only)
};

struct HostGeom {
  int N = 0, W = 0, F = 0, Fi = 0;
  std::vector<int> nef, gef, enbr, iown, inbr; // NumbOfElementFaces, NGlobalEFaces, ElementNeighbor, ...
  std::vector<int> cell_node_ptr, cell_nodes;
  FaceGeom face;
  std::vector<double> vol, xc, yc, zc;
};

void build_topology(const PolyInput &in, HostGeom &g);
void build_geometry(const PolyInput &in, HostGeom &g);
// the E/T split (MethodDecomposeS 1, 2 or 3) and the gradient-correction
// offsets of every face, recomputed in place; owner: 0-based owner per face
void decompose_faces(HostGeom &g, const int *owner, int method);

} // namespace fs
