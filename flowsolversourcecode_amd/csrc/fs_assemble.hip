// Face-flux assembly into the ELL system (AssembleGlobalMatrixFaceFluxes /
// AssembleGlobalMatrixElementFluxes, Assemble/AssembleGolbalElementAndFaceFluxes.for
// :43-125 and :4-39) as a per-cell gather instead of the reference's face
// scatter: no atomics, deterministic, and each cell adds its faces in the
// order the reference's face loop reaches them (interior faces by index, then
// boundary faces), so ac/anb/bc are bitwise the reference's.
#include <hip/hip_runtime.h>

#include <string>

#include "fs_common.h"
#include "fs_kernels.h"
#include "fs_solve.h"

namespace fs {

// role codes packed with the face index: code = face*4 + role
enum { ROLE_NONE = 0, ROLE_OWNER = 1, ROLE_NEIGH = 2, ROLE_BOUND = 3 };

__global__ void k_assemble_gather(int n, int w, const int *__restrict__ code,
                                  const double *__restrict__ cf, const double *__restrict__ ff,
                                  const double *__restrict__ tf, const double *__restrict__ ce,
                                  const double *__restrict__ te, double *__restrict__ ac,
                                  double *__restrict__ anb, double *__restrict__ bc) {
  const int i = xcd_block() * kBlock + threadIdx.x;
  if (i >= n) return;
  double a = ac[i], b = bc[i];
  for (int s = 0; s < w; ++s) {
    const int c = code[(size_t)s * n + i];
    const int role = c & 3, f = c >> 2;
    if (role == ROLE_OWNER) { // :60-66
      a = a + cf[f];
      anb[(size_t)s * n + i] = anb[(size_t)s * n + i] + ff[f];
      b = b - tf[f];
    } else if (role == ROLE_NEIGH) { // :67-72
      a = a - ff[f];
      anb[(size_t)s * n + i] = anb[(size_t)s * n + i] - cf[f];
      b = b + tf[f];
    } else if (role == ROLE_BOUND) { // :83-96
      a = a + cf[f];
      b = b - tf[f];
    }
  }
  if (ce) a = a + ce[i]; // AssembleGlobalMatrixElementFluxes :19-22
  if (te) b = b - te[i];
  ac[i] = a;
  bc[i] = b;
}

} // namespace fs

using fs::set_error;

extern "C" int fs_mesh_set_cell_faces(fs_mesh *m, int n_faces_total, const int *gef,
                                      const int *nef) {
  try {
    if (!m || !gef || !nef) {
      set_error("fs_mesh_set_cell_faces: NULL argument");
      return -1;
    }
    if (m->nif <= 0 || m->h_owner.empty()) {
      set_error("fs_mesh_set_cell_faces: the mesh was created without interior face owners");
      return -1;
    }
    const int n = m->n, w = m->w;
    std::vector<int> code((size_t)n * w, fs::ROLE_NONE);
    for (int i = 0; i < n; ++i) {
      if (nef[i] < 0 || nef[i] > w) {
        set_error("fs_mesh_set_cell_faces: NumbOfElementFaces out of range");
        return -1;
      }
      for (int s = 0; s < nef[i]; ++s) {
        const int f = gef[(size_t)s * n + i];
        if (f < 1 || f > n_faces_total) {
          set_error("fs_mesh_set_cell_faces: face number out of range");
          return -1;
        }
        int role;
        if (f <= m->nif) {
          if (m->h_owner[f - 1] == i + 1) role = fs::ROLE_OWNER;
          else if (m->h_neigh[f - 1] == i + 1) role = fs::ROLE_NEIGH;
          else {
            set_error("fs_mesh_set_cell_faces: interior face not adjacent to its cell");
            return -1;
          }
        } else {
          role = fs::ROLE_BOUND;
        }
        code[(size_t)s * n + i] = (f - 1) * 4 + role;
      }
    }
    m->nfaces_total = n_faces_total;
    m->d_face_code.upload(code, m->stream);
    FS_HIP(hipStreamSynchronize(m->stream));
    return 0;
  } catch (const std::exception &e) {
    set_error(e.what());
    return -2;
  }
}

extern "C" int fs_assemble_face_fluxes_device(fs_mesh *m, const double *cf, const double *ff,
                                              const double *tf, const double *ce,
                                              const double *te, double *ac, double *anb,
                                              double *bc) {
  if (!m || !cf || !ff || !tf || !ac || !anb || !bc) {
    set_error("fs_assemble_face_fluxes_device: NULL argument");
    return -1;
  }
  if (!m->d_face_code.p) {
    set_error("fs_assemble_face_fluxes_device: call fs_mesh_set_cell_faces first");
    return -1;
  }
  const int n = m->n;
  hipLaunchKernelGGL(fs::k_assemble_gather, dim3((n + fs::kBlock - 1) / fs::kBlock),
                     dim3(fs::kBlock), 0, m->stream, n, m->w, m->d_face_code.p, cf, ff, tf, ce,
                     te, ac, anb, bc);
  FS_HIP(hipGetLastError());
  FS_HIP(hipStreamSynchronize(m->stream));
  return 0;
}
