// Host-side mesh preparation: the reference's polyMesh topology (Readit,
// SolveEquations/GridAndGeometry/ReadFoamMesh.f90:254-619) and geometry
// (ProcessGeometry, ProcessGeometryRoutine.for:3-499, MethodDecomposeS = 3),
// computed once per mesh in FP64 with the reference's operation order so the
// device works on bitwise the reference's geometry.
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "fs_geom.h"

namespace fs {

void build_topology(const PolyInput &in, HostGeom &g) {
  const int nf = in.nfaces, nif = in.nif;
  int nc = -1;
  for (int f = 0; f < nf; ++f) nc = std::max(nc, in.owner[f]);
  for (int f = 0; f < nif; ++f) nc = std::max(nc, in.neigh[f]);
  g.N = nc + 1;
  g.Fi = nif;
  g.F = nf;
  std::vector<int> count(g.N, 0);
  for (int f = 0; f < nif; ++f) {
    count[in.owner[f]]++;
    count[in.neigh[f]]++;
  }
  for (int f = nif; f < nf; ++f) count[in.owner[f]]++;
  g.W = 0;
  for (int c : count) g.W = std::max(g.W, c);
  const int N = g.N, W = g.W;
  // slots in face order: interior faces (owner then neighbour), then
  // boundary faces (ReadFoamMesh.f90:326-342)
  g.nef.assign(N, 0);
  g.gef.assign((size_t)N * W, 0);
  auto push = [&](int cell, int face) { g.gef[(size_t)(g.nef[cell]++) * N + cell] = face + 1; };
  for (int f = 0; f < nif; ++f) {
    push(in.owner[f], f);
    push(in.neigh[f], f);
  }
  for (int f = nif; f < nf; ++f) push(in.owner[f], f);
  g.enbr.assign((size_t)N * W, 0);
  g.iown.assign(nif, 0);
  g.inbr.assign(nif, 0);
  for (int f = 0; f < nif; ++f) { // :576-603
    const int o = in.owner[f], n = in.neigh[f];
    for (int j = 0; j < g.nef[o]; ++j)
      if (g.gef[(size_t)j * N + o] == f + 1) {
        g.iown[f] = j + 1;
        g.enbr[(size_t)j * N + o] = n + 1;
        break;
      }
    for (int j = 0; j < g.nef[n]; ++j)
      if (g.gef[(size_t)j * N + n] == f + 1) {
        g.inbr[f] = j + 1;
        g.enbr[(size_t)j * N + n] = o + 1;
        break;
      }
  }
  // per-cell unique node list in first-seen order (:446-476)
  g.cell_node_ptr.assign(N + 1, 0);
  g.cell_nodes.clear();
  std::vector<int> tmp;
  for (int c = 0; c < N; ++c) {
    tmp.clear();
    for (int j = 0; j < g.nef[c]; ++j) {
      const int f = g.gef[(size_t)j * N + c] - 1;
      for (int t = in.fptr[f]; t < in.fptr[f + 1]; ++t) {
        const int p = in.fnodes[t];
        bool seen = false;
        for (int q : tmp)
          if (q == p) seen = true;
        if (!seen) tmp.push_back(p);
      }
    }
    g.cell_nodes.insert(g.cell_nodes.end(), tmp.begin(), tmp.end());
    g.cell_node_ptr[c + 1] = (int)g.cell_nodes.size();
  }
}

void build_geometry(const PolyInput &in, HostGeom &g) {
  const int nf = g.F, nif = g.Fi, N = g.N;
  auto &F = g.face;
  for (auto *v : {&F.Sx, &F.Sy, &F.Sz, &F.A, &F.cx, &F.cy, &F.cz, &F.nx, &F.ny, &F.nz, &F.d, &F.dx,
                  &F.dy, &F.dz, &F.ex, &F.ey, &F.ez, &F.g, &F.Ex, &F.Ey, &F.Ez, &F.E, &F.gdiff, &F.Tx,
                  &F.Ty, &F.Tz, &F.T})
    v->assign(nf, 0.0);
  g.vol.assign(N, 0.0);
  g.xc.assign(N, 0.0);
  g.yc.assign(N, 0.0);
  g.zc.assign(N, 0.0);
  const double *X = in.x, *Y = in.y, *Z = in.z;
  // face area vector, area and centroid by centre-fan triangles
  for (int f = 0; f < nf; ++f) {
    const int b = in.fptr[f], e = in.fptr[f + 1];
    double cx = 0, cy = 0, cz = 0;
    for (int t = b; t < e; ++t) {
      cx = cx + X[in.fnodes[t]];
      cy = cy + Y[in.fnodes[t]];
      cz = cz + Z[in.fnodes[t]];
    }
    const int cnt = e - b;
    cx = cx / cnt;
    cy = cy / cnt;
    cz = cz / cnt;
    double sx = 0, sy = 0, sz = 0, ar = 0, mx = 0, my = 0, mz = 0;
    for (int t = b; t < e; ++t) {
      const int p2 = in.fnodes[t], p3 = in.fnodes[t + 1 < e ? t + 1 : b];
      const double tx = 0.5 * ((Y[p2] - cy) * (Z[p3] - cz) - (Y[p3] - cy) * (Z[p2] - cz));
      const double ty = -(0.5 * ((X[p2] - cx) * (Z[p3] - cz) - (X[p3] - cx) * (Z[p2] - cz)));
      const double tz = 0.5 * ((X[p2] - cx) * (Y[p3] - cy) - (X[p3] - cx) * (Y[p2] - cy));
      const double ta = std::sqrt(tx * tx + ty * ty + tz * tz);
      const double gx = (cx + X[p2] + X[p3]) / 3.0, gy = (cy + Y[p2] + Y[p3]) / 3.0,
                   gz = (cz + Z[p2] + Z[p3]) / 3.0;
      sx = sx + tx;
      sy = sy + ty;
      sz = sz + tz;
      ar = ar + ta;
      mx = mx + gx * ta;
      my = my + gy * ta;
      mz = mz + gz * ta;
    }
    F.Sx[f] = sx;
    F.Sy[f] = sy;
    F.Sz[f] = sz;
    F.A[f] = ar;
    F.cx[f] = mx / ar;
    F.cy[f] = my / ar;
    F.cz[f] = mz / ar;
  }
  // cells by pyramids on the node-average centre
  for (int c = 0; c < N; ++c) {
    const int nb = g.cell_node_ptr[c], ne = g.cell_node_ptr[c + 1];
    double x1 = 0, y1 = 0, z1 = 0;
    for (int t = nb; t < ne; ++t) {
      x1 = x1 + X[g.cell_nodes[t]];
      y1 = y1 + Y[g.cell_nodes[t]];
      z1 = z1 + Z[g.cell_nodes[t]];
    }
    x1 = x1 / (ne - nb);
    y1 = y1 / (ne - nb);
    z1 = z1 / (ne - nb);
    double v = 0, px = 0, py = 0, pz = 0;
    for (int j = 0; j < g.nef[c]; ++j) {
      const int f = g.gef[(size_t)j * N + c] - 1;
      const double x2 = F.cx[f], y2 = F.cy[f], z2 = F.cz[f];
      const double vt = std::fabs(F.Sx[f] * (x2 - x1) + F.Sy[f] * (y2 - y1) + F.Sz[f] * (z2 - z1)) / 3.0;
      v = v + vt;
      px = px + (0.75 * x2 + 0.25 * x1) * vt;
      py = py + (0.75 * y2 + 0.25 * y1) * vt;
      pz = pz + (0.75 * z2 + 0.25 * z1) * vt;
    }
    g.vol[c] = v;
    g.xc[c] = px / v;
    g.yc[c] = py / v;
    g.zc[c] = pz / v;
  }
  auto unit = [&](int f) {
    F.d[f] = std::sqrt(F.dx[f] * F.dx[f] + F.dy[f] * F.dy[f] + F.dz[f] * F.dz[f]);
    F.ex[f] = F.dx[f] / F.d[f];
    F.ey[f] = F.dy[f] / F.d[f];
    F.ez[f] = F.dz[f] / F.d[f];
  };
  auto orient = [&](int f) {
    if (F.dx[f] * F.Sx[f] + F.dy[f] * F.Sy[f] + F.dz[f] * F.Sz[f] < 0.) {
      F.Sx[f] = -F.Sx[f];
      F.Sy[f] = -F.Sy[f];
      F.Sz[f] = -F.Sz[f];
    }
    F.nx[f] = F.Sx[f] / F.A[f];
    F.ny[f] = F.Sy[f] / F.A[f];
    F.nz[f] = F.Sz[f] / F.A[f];
  };
  for (int f = 0; f < nif; ++f) {
    const int i = in.owner[f], j = in.neigh[f];
    F.dx[f] = g.xc[j] - g.xc[i];
    F.dy[f] = g.yc[j] - g.yc[i];
    F.dz[f] = g.zc[j] - g.zc[i];
    unit(f);
  }
  for (int f = 0; f < nif; ++f) orient(f);
  for (int f = 0; f < nif; ++f) {
    const int i = in.owner[f], j = in.neigh[f];
    const double ax = F.cx[f] - g.xc[i], ay = F.cy[f] - g.yc[i], az = F.cz[f] - g.zc[i];
    const double bx = F.cx[f] - g.xc[j], by = F.cy[f] - g.yc[j], bz = F.cz[f] - g.zc[j];
    const double d1 = std::sqrt(ax * ax + ay * ay + az * az), d2 = std::sqrt(bx * bx + by * by + bz * bz);
    F.g[f] = d2 / (d1 + d2);
  }
  for (int f = nif; f < nf; ++f) {
    const int o = in.owner[f];
    F.dx[f] = F.cx[f] - g.xc[o];
    F.dy[f] = F.cy[f] - g.yc[o];
    F.dz[f] = F.cz[f] - g.zc[o];
    unit(f);
    orient(f);
    F.g[f] = 1.0;
  }
  decompose_faces(g, in.owner, 3);
}

// The E/T split of every face's area vector (ProcessGeometryRoutine.for
// :442-500 interior, :519-575 boundary) by MethodDecomposeS: 1 minimum
// correction E = (e.S) e, 2 orthogonal correction E = |S| e, 3 over-relaxed
// E = |S|^2/(e.S) e; T = S - E, gDiff = |E|/d.  Then the point f' on the
// centroid line used by the face-centroid correction of the gradient
// (StandardGreenGaussCorrection1Gradient.for:111-119), kept as the offset
// r_f - r_f' of every interior face (it is formed with E).
void decompose_faces(HostGeom &g, const int *owner, int method) {
  const int nf = g.F, nif = g.Fi;
  auto &F = g.face;
  for (int f = 0; f < nf; ++f) {
    if (method == 1) {
      const double dot = F.ex[f] * F.Sx[f] + F.ey[f] * F.Sy[f] + F.ez[f] * F.Sz[f];
      F.Ex[f] = dot * F.ex[f];
      F.Ey[f] = dot * F.ey[f];
      F.Ez[f] = dot * F.ez[f];
      F.E[f] = std::fabs(dot);
    } else if (method == 2) {
      F.Ex[f] = F.A[f] * F.ex[f];
      F.Ey[f] = F.A[f] * F.ey[f];
      F.Ez[f] = F.A[f] * F.ez[f];
      F.E[f] = F.A[f];
    } else {
      const double dot = F.ex[f] * F.Sx[f] + F.ey[f] * F.Sy[f] + F.ez[f] * F.Sz[f];
      const double s = F.A[f] * F.A[f] / dot;
      F.Ex[f] = s * F.ex[f];
      F.Ey[f] = s * F.ey[f];
      F.Ez[f] = s * F.ez[f];
      F.E[f] = std::sqrt(F.Ex[f] * F.Ex[f] + F.Ey[f] * F.Ey[f] + F.Ez[f] * F.Ez[f]);
    }
    F.gdiff[f] = F.E[f] / F.d[f];
    F.Tx[f] = F.Sx[f] - F.Ex[f];
    F.Ty[f] = F.Sy[f] - F.Ey[f];
    F.Tz[f] = F.Sz[f] - F.Ez[f];
    F.T[f] = std::sqrt(F.Tx[f] * F.Tx[f] + F.Ty[f] * F.Ty[f] + F.Tz[f] * F.Tz[f]);
  }
  F.kx.assign(nif, 0.0);
  F.ky.assign(nif, 0.0);
  F.kz.assign(nif, 0.0);
  for (int f = 0; f < nif; ++f) {
    const int j = owner[f];
    const double edn = F.Ex[f] * F.nx[f] + F.Ey[f] * F.ny[f] + F.Ez[f] * F.nz[f];
    const double rdn = (F.cx[f] - g.xc[j]) * F.nx[f] + (F.cy[f] - g.yc[j]) * F.ny[f] +
                       (F.cz[f] - g.zc[j]) * F.nz[f];
    const double xp = g.xc[j] + (rdn / edn) * F.Ex[f];
    const double yp = g.yc[j] + (rdn / edn) * F.Ey[f];
    const double zp = g.zc[j] + (rdn / edn) * F.Ez[f];
    F.kx[f] = F.cx[f] - xp;
    F.ky[f] = F.cy[f] - yp;
    F.kz[f] = F.cz[f] - zp;
  }
}

} // namespace fs

// ---- host-only C ABI over the topology and geometry (no device calls)
#include <cstring>
#include <map>
#include <string>

#include "fsamd_simple.h"

struct fs_geom {
  fs::HostGeom g;
};

extern "C" int fs_geom_create(int npoints, const double *x, const double *y, const double *z,
                              int nfaces, const int *face_ptr, const int *face_nodes,
                              const int *owner, int nif, const int *neighbour, fs_geom **out) {
  if (!out || !x || !y || !z || !face_ptr || !face_nodes || !owner || !neighbour || nfaces <= 0 ||
      nif <= 0 || nif > nfaces || npoints <= 0)
    return -1;
  for (int f = 0; f < nfaces; ++f) {
    if (face_ptr[f + 1] - face_ptr[f] < 3 || owner[f] < 0 || (f < nif && neighbour[f] < 0)) return -1;
    for (int t = face_ptr[f]; t < face_ptr[f + 1]; ++t)
      if (face_nodes[t] < 0 || face_nodes[t] >= npoints) return -1;
  }
  auto *h = new fs_geom();
  fs::PolyInput in;
  in.npoints = npoints;
  in.nfaces = nfaces;
  in.nif = nif;
  in.x = x;
  in.y = y;
  in.z = z;
  in.fptr = face_ptr;
  in.fnodes = face_nodes;
  in.owner = owner;
  in.neigh = neighbour;
  fs::build_topology(in, h->g);
  fs::build_geometry(in, h->g);
  *out = h;
  return 0;
}

extern "C" int fs_geom_destroy(fs_geom *h) {
  delete h;
  return 0;
}

extern "C" int fs_geom_dims(fs_geom *h, int *dims) {
  if (!h || !dims) return -1;
  dims[0] = h->g.N;
  dims[1] = h->g.W;
  dims[2] = h->g.F;
  dims[3] = h->g.Fi;
  return 0;
}

extern "C" int fs_geom_get(fs_geom *h, const char *name, double *out, long long count) {
  if (!h || !name || !out) return -1;
  auto &F = h->g.face;
  const std::map<std::string, const std::vector<double> *> tab = {
      {"Sx", &F.Sx}, {"Sy", &F.Sy}, {"Sz", &F.Sz}, {"A", &F.A}, {"fcx", &F.cx}, {"fcy", &F.cy},
      {"fcz", &F.cz}, {"nx", &F.nx}, {"ny", &F.ny}, {"nz", &F.nz}, {"dcf", &F.d}, {"dx", &F.dx},
      {"dy", &F.dy}, {"dz", &F.dz}, {"ex", &F.ex}, {"ey", &F.ey}, {"ez", &F.ez}, {"g", &F.g},
      {"Ex", &F.Ex}, {"Ey", &F.Ey}, {"Ez", &F.Ez}, {"E", &F.E}, {"gdiff", &F.gdiff}, {"Tx", &F.Tx},
      {"Ty", &F.Ty}, {"Tz", &F.Tz}, {"T", &F.T}, {"vol", &h->g.vol}, {"xc", &h->g.xc},
      {"yc", &h->g.yc}, {"zc", &h->g.zc}};
  auto it = tab.find(name);
  if (it == tab.end() || (size_t)count != it->second->size()) return -1;
  std::memcpy(out, it->second->data(), sizeof(double) * (size_t)count);
  return 0;
}

extern "C" int fs_geom_topology(fs_geom *h, int which, int *out, long long count) {
  if (!h || !out) return -1;
  const auto &g = h->g;
  const std::vector<int> *v = which == 0 ? &g.nef : which == 1 ? &g.gef : which == 2 ? &g.enbr
                              : which == 3 ? &g.iown : which == 4 ? &g.inbr : nullptr;
  if (!v || (size_t)count != v->size()) return -1;
  std::copy(v->begin(), v->end(), out);
  return 0;
}
