/*
 * oracle/or_geom.c -- CPU restatement of the reference's polyMesh topology
 * (ReadFoamMesh.f90:254-619, Readit) and geometry (ProcessGeometryRoutine.for
 * :3-499, ProcessGeometry with MethodDecomposeS = 3).
 *
 * TEST INFRASTRUCTURE ONLY (see or_solve.c).
 *
 * Input: a polyMesh in memory -- points, face node lists (CSR), owner of every
 * face, neighbour of the interior faces (faces 0..nif-1 interior, then the
 * boundary faces patch by patch).  0-based in C; every cell number stored in
 * an output array the reference also stores is 1-based.
 * Face arrays cover all faces (interior then boundary) in one numbering,
 * which is the reference's NFacesTotal numbering; the reference keeps the
 * boundary part in (set, record) arrays.
 */
#include "or_simple.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static void *xc_alloc(size_t n, size_t s) { return calloc(n ? n : 1, s); }

void or_mesh_free(or_mesh *m) {
  free(m->x);
  free(m->y);
  free(m->z);
  free(m->fptr);
  free(m->fnodes);
  free(m->owner);
  free(m->neigh);
  free(m->nef);
  free(m->gef);
  free(m->enbr);
  free(m->iown);
  free(m->inbr);
  free(m->nen);
  free(m->len);
  free(m->Sx);
  free(m->Sy);
  free(m->Sz);
  free(m->A);
  free(m->fcx);
  free(m->fcy);
  free(m->fcz);
  free(m->nx);
  free(m->ny);
  free(m->nz);
  free(m->vol);
  free(m->xc);
  free(m->yc);
  free(m->zc);
  free(m->dcf);
  free(m->dx);
  free(m->dy);
  free(m->dz);
  free(m->ex);
  free(m->ey);
  free(m->ez);
  free(m->g);
  free(m->Ex);
  free(m->Ey);
  free(m->Ez);
  free(m->E);
  free(m->gdiff);
  free(m->Tx);
  free(m->Ty);
  free(m->Tz);
  free(m->T);
  free(m->patch_start);
  memset(m, 0, sizeof(*m));
}

/* Readit (ReadFoamMesh.f90:288-609): cell face slots in face order (each
 * interior face appended to its owner then to its neighbour, then boundary
 * faces to their owners), ElementNeighbor across each slot, iOwnerNeighbor /
 * iNeighborOwner, and the unique node list of each cell (:371-476). */
int or_mesh_topology(or_mesh *m) {
  const int nc = m->ncells, nf = m->nfaces, nif = m->nif;
  m->nef = (int *)xc_alloc(nc, sizeof(int));
  for (int f = 0; f < nif; ++f) {
    m->nef[m->owner[f]]++;
    m->nef[m->neigh[f]]++;
  }
  for (int f = nif; f < nf; ++f) m->nef[m->owner[f]]++;
  int w = -1;
  for (int c = 0; c < nc; ++c)
    if (m->nef[c] > w) w = m->nef[c];
  m->w = w;
  m->gef = (int *)xc_alloc((size_t)nc * w, sizeof(int));
  memset(m->nef, 0, sizeof(int) * nc);
  for (int f = 0; f < nif; ++f) {
    int o = m->owner[f], n = m->neigh[f];
    m->gef[(size_t)(m->nef[o]++) * nc + o] = f + 1;
    m->gef[(size_t)(m->nef[n]++) * nc + n] = f + 1;
  }
  for (int f = nif; f < nf; ++f) {
    int o = m->owner[f];
    m->gef[(size_t)(m->nef[o]++) * nc + o] = f + 1;
  }
  m->enbr = (int *)xc_alloc((size_t)nc * w, sizeof(int));
  m->iown = (int *)xc_alloc(nif, sizeof(int));
  m->inbr = (int *)xc_alloc(nif, sizeof(int));
  for (int f = 0; f < nif; ++f) { /* :576-603 */
    int o = m->owner[f], n = m->neigh[f];
    for (int j = 0; j < m->nef[o]; ++j)
      if (m->gef[(size_t)j * nc + o] == f + 1) {
        m->iown[f] = j + 1;
        m->enbr[(size_t)j * nc + o] = n + 1;
        break;
      }
    for (int j = 0; j < m->nef[n]; ++j)
      if (m->gef[(size_t)j * nc + n] == f + 1) {
        m->inbr[f] = j + 1;
        m->enbr[(size_t)j * nc + n] = o + 1;
        break;
      }
  }
  /* unique node list: first face's nodes, then unseen nodes of later faces
   * (:446-476; the node count is that of the unique set, :371-412) */
  int maxn = 0;
  m->nen = (int *)xc_alloc(nc, sizeof(int));
  int *tmp = (int *)xc_alloc(512, sizeof(int));
  m->len_w = 0;
  for (int pass = 0; pass < 2; ++pass) {
    for (int c = 0; c < nc; ++c) {
      int k1 = 0;
      for (int j = 0; j < m->nef[c]; ++j) {
        int f = m->gef[(size_t)j * nc + c] - 1;
        for (int t = m->fptr[f]; t < m->fptr[f + 1]; ++t) {
          int p = m->fnodes[t], seen = 0;
          for (int q = 0; q < k1; ++q)
            if (tmp[q] == p) seen = 1;
          if (!seen && k1 < 512) tmp[k1++] = p;
        }
      }
      if (pass == 0) {
        m->nen[c] = k1;
        if (k1 > maxn) maxn = k1;
      } else {
        for (int q = 0; q < k1; ++q) m->len[(size_t)q * nc + c] = tmp[q];
      }
    }
    if (pass == 0) {
      m->len_w = maxn;
      m->len = (int *)xc_alloc((size_t)nc * maxn, sizeof(int));
    }
  }
  free(tmp);
  return 0;
}

/* ProcessGeometry (ProcessGeometryRoutine.for:76-499) for MethodDecomposeS=3. */
int or_mesh_geometry(or_mesh *m) {
  const int nc = m->ncells, nf = m->nfaces, nif = m->nif;
#define A1(p, n) p = (double *)xc_alloc((n), sizeof(double))
  A1(m->Sx, nf); A1(m->Sy, nf); A1(m->Sz, nf); A1(m->A, nf);
  A1(m->fcx, nf); A1(m->fcy, nf); A1(m->fcz, nf);
  A1(m->nx, nf); A1(m->ny, nf); A1(m->nz, nf);
  A1(m->vol, nc); A1(m->xc, nc); A1(m->yc, nc); A1(m->zc, nc);
  A1(m->dcf, nf); A1(m->dx, nf); A1(m->dy, nf); A1(m->dz, nf);
  A1(m->ex, nf); A1(m->ey, nf); A1(m->ez, nf); A1(m->g, nf);
  A1(m->Ex, nf); A1(m->Ey, nf); A1(m->Ez, nf); A1(m->E, nf);
  A1(m->gdiff, nf); A1(m->Tx, nf); A1(m->Ty, nf); A1(m->Tz, nf); A1(m->T, nf);
#undef A1
  /* face area vectors and centroids by centre-fan triangles (:76-142 interior,
   * :162-228 boundary; same arithmetic) */
  for (int f = 0; f < nf; ++f) {
    const int b = m->fptr[f], e = m->fptr[f + 1], nn = e - b;
    double cx = 0, cy = 0, cz = 0;
    for (int t = b; t < e; ++t) {
      cx = cx + m->x[m->fnodes[t]];
      cy = cy + m->y[m->fnodes[t]];
      cz = cz + m->z[m->fnodes[t]];
    }
    cx = cx / nn;
    cy = cy / nn;
    cz = cz / nn;
    double ax = 0, ay = 0, az = 0, ar = 0, gx = 0, gy = 0, gz = 0;
    for (int t = b; t < e; ++t) {
      const int j2 = m->fnodes[t], j3 = (t + 1 < e) ? m->fnodes[t + 1] : m->fnodes[b];
      const double sx1 = 0.5 * ((m->y[j2] - cy) * (m->z[j3] - cz) - (m->y[j3] - cy) * (m->z[j2] - cz));
      const double sy1 = -(0.5 * ((m->x[j2] - cx) * (m->z[j3] - cz) - (m->x[j3] - cx) * (m->z[j2] - cz)));
      const double sz1 = 0.5 * ((m->x[j2] - cx) * (m->y[j3] - cy) - (m->x[j3] - cx) * (m->y[j2] - cy));
      const double area1 = sqrt(sx1 * sx1 + sy1 * sy1 + sz1 * sz1);
      const double x1 = (cx + m->x[j2] + m->x[j3]) / 3.0;
      const double y1 = (cy + m->y[j2] + m->y[j3]) / 3.0;
      const double z1 = (cz + m->z[j2] + m->z[j3]) / 3.0;
      ax = ax + sx1;
      ay = ay + sy1;
      az = az + sz1;
      ar = ar + area1;
      gx = gx + x1 * area1;
      gy = gy + y1 * area1;
      gz = gz + z1 * area1;
    }
    m->Sx[f] = ax;
    m->Sy[f] = ay;
    m->Sz[f] = az;
    m->A[f] = ar;
    m->fcx[f] = gx / ar;
    m->fcy[f] = gy / ar;
    m->fcz[f] = gz / ar;
  }
  /* cell volumes and centroids from face pyramids (:248-321) */
  for (int c = 0; c < nc; ++c) {
    double x1 = 0, y1 = 0, z1 = 0;
    for (int j = 0; j < m->nen[c]; ++j) {
      const int k = m->len[(size_t)j * nc + c];
      x1 = x1 + m->x[k];
      y1 = y1 + m->y[k];
      z1 = z1 + m->z[k];
    }
    x1 = x1 / m->nen[c];
    y1 = y1 / m->nen[c];
    z1 = z1 / m->nen[c];
    double v = 0, px = 0, py = 0, pz = 0;
    for (int j = 0; j < m->nef[c]; ++j) {
      const int f = m->gef[(size_t)j * nc + c] - 1;
      const double x2 = m->fcx[f], y2 = m->fcy[f], z2 = m->fcz[f];
      const double ddx = x2 - x1, ddy = y2 - y1, ddz = z2 - z1;
      const double xcp = 0.75 * x2 + 0.25 * x1, ycp = 0.75 * y2 + 0.25 * y1, zcp = 0.75 * z2 + 0.25 * z1;
      const double vt = fabs(m->Sx[f] * ddx + m->Sy[f] * ddy + m->Sz[f] * ddz) / 3.0;
      v = v + vt;
      px = px + xcp * vt;
      py = py + ycp * vt;
      pz = pz + zcp * vt;
    }
    m->vol[c] = v;
    m->xc[c] = px / v;
    m->yc[c] = py / v;
    m->zc[c] = pz / v;
  }
  /* interior: centroid distance, orientation owner->neighbour, unit normal,
   * interpolation factor (:323-387) */
  for (int f = 0; f < nif; ++f) {
    const int i = m->owner[f], j = m->neigh[f];
    m->dx[f] = m->xc[j] - m->xc[i];
    m->dy[f] = m->yc[j] - m->yc[i];
    m->dz[f] = m->zc[j] - m->zc[i];
    m->dcf[f] = sqrt(m->dx[f] * m->dx[f] + m->dy[f] * m->dy[f] + m->dz[f] * m->dz[f]);
    m->ex[f] = m->dx[f] / m->dcf[f];
    m->ey[f] = m->dy[f] / m->dcf[f];
    m->ez[f] = m->dz[f] / m->dcf[f];
  }
  for (int f = 0; f < nif; ++f) {
    const double dot = m->dx[f] * m->Sx[f] + m->dy[f] * m->Sy[f] + m->dz[f] * m->Sz[f];
    if (dot < 0.) {
      m->Sx[f] = -m->Sx[f];
      m->Sy[f] = -m->Sy[f];
      m->Sz[f] = -m->Sz[f];
    }
    m->nx[f] = m->Sx[f] / m->A[f];
    m->ny[f] = m->Sy[f] / m->A[f];
    m->nz[f] = m->Sz[f] / m->A[f];
  }
  for (int f = 0; f < nif; ++f) {
    const int i = m->owner[f], j = m->neigh[f];
    const double ax = m->fcx[f] - m->xc[i], ay = m->fcy[f] - m->yc[i], az = m->fcz[f] - m->zc[i];
    const double bx = m->fcx[f] - m->xc[j], by = m->fcy[f] - m->yc[j], bz = m->fcz[f] - m->zc[j];
    const double d1 = sqrt(ax * ax + ay * ay + az * az), d2 = sqrt(bx * bx + by * by + bz * bz);
    m->g[f] = d2 / (d1 + d2);
  }
  /* boundary: owner-to-face vector, outward orientation (:389-418) */
  for (int f = nif; f < nf; ++f) {
    const int o = m->owner[f];
    m->dx[f] = m->fcx[f] - m->xc[o];
    m->dy[f] = m->fcy[f] - m->yc[o];
    m->dz[f] = m->fcz[f] - m->zc[o];
    m->dcf[f] = sqrt(m->dx[f] * m->dx[f] + m->dy[f] * m->dy[f] + m->dz[f] * m->dz[f]);
    m->ex[f] = m->dx[f] / m->dcf[f];
    m->ey[f] = m->dy[f] / m->dcf[f];
    m->ez[f] = m->dz[f] / m->dcf[f];
    const double dot = m->dx[f] * m->Sx[f] + m->dy[f] * m->Sy[f] + m->dz[f] * m->Sz[f];
    if (dot < 0.) {
      m->Sx[f] = -m->Sx[f];
      m->Sy[f] = -m->Sy[f];
      m->Sz[f] = -m->Sz[f];
    }
    m->nx[f] = m->Sx[f] / m->A[f];
    m->ny[f] = m->Sy[f] / m->A[f];
    m->nz[f] = m->Sz[f] / m->A[f];
    m->g[f] = 1.0;
  }
  or_mesh_decompose(m, 3);
  return 0;
}

or_mesh *or_mesh_new(int npoints, const double *x, const double *y, const double *z, int nfaces,
                     const int *fptr, const int *fnodes, const int *owner, int nif,
                     const int *neigh, int npatches, const int *patch_size) {
  or_mesh *m = (or_mesh *)xc_alloc(1, sizeof(or_mesh));
  m->npoints = npoints;
  m->nfaces = nfaces;
  m->nif = nif;
  m->x = (double *)xc_alloc(npoints, sizeof(double));
  m->y = (double *)xc_alloc(npoints, sizeof(double));
  m->z = (double *)xc_alloc(npoints, sizeof(double));
  memcpy(m->x, x, sizeof(double) * npoints);
  memcpy(m->y, y, sizeof(double) * npoints);
  memcpy(m->z, z, sizeof(double) * npoints);
  m->fptr = (int *)xc_alloc(nfaces + 1, sizeof(int));
  memcpy(m->fptr, fptr, sizeof(int) * (nfaces + 1));
  m->fnodes = (int *)xc_alloc(fptr[nfaces], sizeof(int));
  memcpy(m->fnodes, fnodes, sizeof(int) * fptr[nfaces]);
  m->owner = (int *)xc_alloc(nfaces, sizeof(int));
  memcpy(m->owner, owner, sizeof(int) * nfaces);
  m->neigh = (int *)xc_alloc(nif, sizeof(int));
  memcpy(m->neigh, neigh, sizeof(int) * nif);
  int nc = -1;
  for (int f = 0; f < nfaces; ++f)
    if (owner[f] > nc) nc = owner[f];
  for (int f = 0; f < nif; ++f)
    if (neigh[f] > nc) nc = neigh[f];
  m->ncells = nc + 1;
  m->npatches = npatches;
  m->patch_start = (int *)xc_alloc(npatches + 1, sizeof(int));
  m->patch_start[0] = nif;
  for (int p = 0; p < npatches; ++p) m->patch_start[p + 1] = m->patch_start[p] + patch_size[p];
  or_mesh_topology(m);
  or_mesh_geometry(m);
  return m;
}

void or_mesh_delete(or_mesh *m) {
  or_mesh_free(m);
  free(m);
}

/* The E/T split by MethodDecomposeS (ProcessGeometryRoutine.for:442-500,
 * boundary :519-575): 1 E = (e.S) e, 2 E = |S| e, 3 E = |S|^2/(e.S) e. */
void or_mesh_decompose(or_mesh *m, int method) {
  for (int f = 0; f < m->nfaces; ++f) {
    if (method == 1) {
      const double dot = m->ex[f] * m->Sx[f] + m->ey[f] * m->Sy[f] + m->ez[f] * m->Sz[f];
      m->Ex[f] = dot * m->ex[f];
      m->Ey[f] = dot * m->ey[f];
      m->Ez[f] = dot * m->ez[f];
      m->E[f] = fabs(dot);
    } else if (method == 2) {
      m->Ex[f] = m->A[f] * m->ex[f];
      m->Ey[f] = m->A[f] * m->ey[f];
      m->Ez[f] = m->A[f] * m->ez[f];
      m->E[f] = m->A[f];
    } else {
      const double dot = m->ex[f] * m->Sx[f] + m->ey[f] * m->Sy[f] + m->ez[f] * m->Sz[f];
      const double s = m->A[f] * m->A[f] / dot;
      m->Ex[f] = s * m->ex[f];
      m->Ey[f] = s * m->ey[f];
      m->Ez[f] = s * m->ez[f];
      m->E[f] = sqrt(m->Ex[f] * m->Ex[f] + m->Ey[f] * m->Ey[f] + m->Ez[f] * m->Ez[f]);
    }
    m->gdiff[f] = m->E[f] / m->dcf[f];
    m->Tx[f] = m->Sx[f] - m->Ex[f];
    m->Ty[f] = m->Sy[f] - m->Ey[f];
    m->Tz[f] = m->Sz[f] - m->Ez[f];
    m->T[f] = sqrt(m->Tx[f] * m->Tx[f] + m->Ty[f] * m->Ty[f] + m->Tz[f] * m->Tz[f]);
  }
}
