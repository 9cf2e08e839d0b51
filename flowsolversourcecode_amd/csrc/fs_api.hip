// C ABI entry points (include/fsamd.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <stdexcept>
#include <string>

#include "fs_common.h"
#include "fs_solve.h"

namespace fs {
static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
} // namespace fs

using fs::set_error;

#define FS_GUARD_BEGIN try {
#define FS_GUARD_END                                                           \
  }                                                                            \
  catch (const std::exception &e) {                                            \
    set_error(e.what());                                                       \
    return -2;                                                                 \
  }                                                                            \
  catch (...) {                                                                \
    set_error("unknown error");                                                \
    return -2;                                                                 \
  }

extern "C" {

const char *fs_last_error(void) { return fs::g_err.c_str(); }

int fs_kernel_clock_start(fs_mesh *m, int site_mask, int capacity) {
  if (!m) {
    set_error("fs_kernel_clock_start: mesh is NULL");
    return -1;
  }
  FS_GUARD_BEGIN
  return fs::kernel_clock_start(m, site_mask, capacity);
  FS_GUARD_END
}

int fs_kernel_clock_read(fs_mesh *m, int site, double *out) {
  if (!m || !out) {
    set_error("fs_kernel_clock_read: NULL argument");
    return -1;
  }
  FS_GUARD_BEGIN
  return fs::kernel_clock_read(m, site, out);
  FS_GUARD_END
}

int fs_kernel_clock_stride(fs_mesh *m, int stride) {
  if (!m) {
    set_error("fs_kernel_clock_stride: mesh is NULL");
    return -1;
  }
  FS_GUARD_BEGIN
  return fs::kernel_clock_stride(m, stride);
  FS_GUARD_END
}

int fs_kernel_clock_clear(fs_mesh *m) {
  if (!m) {
    set_error("fs_kernel_clock_clear: mesh is NULL");
    return -1;
  }
  FS_GUARD_BEGIN
  return fs::kernel_clock_clear(m);
  FS_GUARD_END
}
const char *fs_version(void) { return "fsamd 0.1 gfx950"; }

int fs_mesh_create(int n_cells, int width, const int *elem_nbr, const int *n_nbr, int n_ifaces,
                   const int *owner, const int *neighbour, const int *i_own_nbr,
                   const int *i_nbr_own, fs_mesh **out) {
  FS_GUARD_BEGIN
  if (!out) {
    set_error("fs_mesh_create: out is NULL");
    return -1;
  }
  *out = nullptr;
  if (n_cells <= 0 || width <= 0 || !elem_nbr || !n_nbr) {
    set_error("fs_mesh_create: need n_cells > 0, width > 0, ElementNeighbor and its counts");
    return -1;
  }
  // every neighbour index must name a cell: the device gathers through them
  for (int i = 0; i < n_cells; ++i) {
    if (n_nbr[i] < 0 || n_nbr[i] > width) {
      set_error("fs_mesh_create: NumberofElementNeighbors out of range at cell " + std::to_string(i + 1));
      return -1;
    }
  }
  for (size_t t = 0; t < (size_t)n_cells * width; ++t) {
    if (elem_nbr[t] < 0 || elem_nbr[t] > n_cells) {
      set_error("fs_mesh_create: ElementNeighbor entry out of range at linear index " + std::to_string(t));
      return -1;
    }
  }
  if (n_ifaces > 0 && owner && neighbour) {
    for (int f = 0; f < n_ifaces; ++f)
      if (owner[f] < 1 || owner[f] > n_cells || neighbour[f] < 1 || neighbour[f] > n_cells ||
          (i_own_nbr && (i_own_nbr[f] < 1 || i_own_nbr[f] > width)) ||
          (i_nbr_own && (i_nbr_own[f] < 1 || i_nbr_own[f] > width))) {
        set_error("fs_mesh_create: face arrays out of range at face " + std::to_string(f + 1));
        return -1;
      }
  }
  auto *m = new fs_mesh();
  m->n = n_cells;
  m->w = width;
  int rc = 0;
  try {
    FS_HIP_THROW(hipGetDevice(&m->device));
    FS_HIP_THROW(hipStreamCreate(&m->stream));
    m->h_nbr.assign(elem_nbr, elem_nbr + (size_t)n_cells * width);
    m->h_nnbr.assign(n_nbr, n_nbr + n_cells);
    m->h_nbr_dev.resize(m->h_nbr.size());
    for (int i = 0; i < n_cells; ++i)
      for (int j = 0; j < width; ++j) {
        const size_t t = (size_t)j * n_cells + i;
        const int k = m->h_nbr[t];
        m->h_nbr_dev[t] = (j < n_nbr[i] && k > 0) ? k - 1 : -1;
      }
    m->d_nbr.upload(m->h_nbr_dev, m->stream);
    {
      // 16-bit offsets when the mesh's numbering keeps every neighbour within
      // +-32767 cells (the row loops then read 2 bytes per slot instead of 4)
      const int wp = (width + 1) / 2;
      std::vector<int> d16((size_t)wp * n_cells, 0);
      bool fits = true;
      for (int i = 0; i < n_cells && fits; ++i)
        for (int j = 0; j < width; ++j) {
          const size_t t = (size_t)j * n_cells + i;
          const int k = m->h_nbr_dev[t];
          if (k < 0) continue;
          const int d = k - i;
          if (d == 0 || d < -32767 || d > 32767) {
            fits = false;
            break;
          }
          const unsigned h = (unsigned)(unsigned short)(short)d;
          d16[(size_t)(j / 2) * n_cells + i] |= (int)(j % 2 ? h << 16 : h);
        }
      if (fits) m->d_nbr16.upload(d16, m->stream);
    }
    if (n_ifaces > 0 && owner && neighbour) {
      m->nif = n_ifaces;
      m->h_owner.assign(owner, owner + n_ifaces);
      m->h_neigh.assign(neighbour, neighbour + n_ifaces);
      m->d_owner.upload(owner, n_ifaces, m->stream);
      m->d_neigh.upload(neighbour, n_ifaces, m->stream);
      if (i_own_nbr) m->d_iown.upload(i_own_nbr, n_ifaces, m->stream);
      if (i_nbr_own) m->d_inbr.upload(i_nbr_own, n_ifaces, m->stream);
    }
    rc = fs::mesh_prepare_solver(m);
  } catch (...) {
    fs::mesh_release_solver(m);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    throw;
  }
  *out = m;
  return rc;
  FS_GUARD_END
}

int fs_mesh_destroy(fs_mesh *m) {
  if (!m) return 0;
  (void)hipStreamSynchronize(m->stream);
  m->amg.reset();
  fs::mesh_release_solver(m);
  hipStream_t s = m->own_stream ? m->stream : nullptr;
  delete m;
  if (s) (void)hipStreamDestroy(s);
  return 0;
}

int fs_set_option(fs_mesh *m, int key, double value) {
  if (!m) {
    set_error("fs_set_option: mesh is NULL");
    return -1;
  }
  const int iv = (int)value;
  // keys that change the hierarchy's structure force a rebuild, but only when
  // their value actually changes
  const int old_levels = m->mg_maxlevels, old_parents = m->mg_minparents, old_tail = m->mg_tail_cells,
            old_sell = m->mg_sell_rows, old_agg = m->mg_agg;
  switch (key) {
  case FS_OPT_MG_PRESWEEP: m->mg_pre = iv < 0 ? 0 : iv; break;
  case FS_OPT_MG_POSTSWEEP: m->mg_post = iv < 0 ? 0 : iv; break;
  case FS_OPT_MG_MAXCYCLES: m->mg_maxcycles = iv; break;
  case FS_OPT_MG_MAXLEVELS: m->mg_maxlevels = iv; break;
  case FS_OPT_MG_MINPARENTS: m->mg_minparents = iv; break;
  case FS_OPT_MG_REAGGLOMERATE: m->mg_reagg = iv; break;
  case FS_OPT_MG_FP32: m->mg_fp32 = iv; break;
  case FS_OPT_MG_OMEGA: m->mg_omega = value; break;
  case FS_OPT_MG_SMOOTHER: m->mg_smoother = iv; break;
  case FS_OPT_GMRES_RESTART: m->gmres_m = iv; break;
  case FS_OPT_CHECK_EVERY: m->check_every = iv < 1 ? 1 : iv; break;
  case FS_OPT_MG_TAIL_CELLS: m->mg_tail_cells = iv; break;
  case FS_OPT_MG_COARSE_SWEEPS: m->mg_coarse_sweeps = iv < 1 ? 1 : iv; break;
  case FS_OPT_MG_ALPHA:
    if (!(value >= 0.0 && value <= 4.0)) {
      set_error("fs_set_option: MG_ALPHA must lie in [0, 4]");
      return -1;
    }
    m->mg_alpha = value;
    break;
  case FS_OPT_MG_COARSE_PRESWEEP: m->mg_cpre = iv < -1 ? -1 : iv; break;
  case FS_OPT_MG_COARSE_POSTSWEEP: m->mg_cpost = iv < -1 ? -1 : iv; break;
  case FS_OPT_MG_SELL_ROWS: m->mg_sell_rows = iv < 0 ? 0 : iv; break;
  case FS_OPT_MG_FUSE_PROLONG: m->mg_fuse_prolong = iv != 0; break;
  case FS_OPT_MG_CYCLE:
    if (iv < 0 || iv > 2) {
      set_error("fs_set_option: MG_CYCLE must be 0 (V), 1 (F) or 2 (W)");
      return -1;
    }
    m->mg_cycle = iv;
    break;
  case FS_OPT_WARM_PREVIOUS:
    if (iv < 0 || iv > fs::kProjMax) {
      set_error("fs_set_option: WARM_PREVIOUS must lie in [0, 4]");
      return -1;
    }
    m->warm_prev = iv;
    break;
  case FS_OPT_MG_AGGREGATION:
    if (iv < 0 || iv > 2) {
      set_error("fs_set_option: MG_AGGREGATION must be 0, 1 or 2");
      return -1;
    }
    m->mg_agg = iv;
    break;
  default:
    set_error("fs_set_option: unknown key " + std::to_string(key));
    return -1;
  }
  if (m->mg_maxlevels != old_levels || m->mg_minparents != old_parents || m->mg_tail_cells != old_tail ||
      m->mg_sell_rows != old_sell || m->mg_agg != old_agg)
    m->amg.reset();
  return 0;
}

int fs_solve_equation_device(fs_mesh *m, int nf, double *d_fit, double rrf, int itmax, int solver,
                             int lmultigrid, const double *d_ac, const double *d_anb,
                             const double *d_bc, double *d_dphi, double *res4,
                             fs_solve_stats *stats) {
  FS_GUARD_BEGIN
  if (!m || !d_fit || !d_ac || !d_anb || !d_bc || !d_dphi) {
    set_error("fs_solve_equation_device: NULL argument");
    return -1;
  }
  if (lmultigrid < FS_MG_OFF || lmultigrid > FS_MG_PRECOND) {
    set_error("fs_solve_equation_device: LMultigrid must be 0, 1 or 2");
    return -1;
  }
  return fs::solve_device(m, nf, d_fit, rrf, itmax, solver, lmultigrid, d_ac, d_anb, d_bc, d_dphi,
                          res4, stats);
  FS_GUARD_END
}

int fs_solve_equation_device_from_guess(fs_mesh *m, int nf, double *d_fit, double rrf, int itmax, int solver,
                                        int lmultigrid, const double *d_ac, const double *d_anb,
                                        const double *d_bc, double *d_dphi, double *res4,
                                        fs_solve_stats *stats) {
  FS_GUARD_BEGIN
  if (!m || !d_fit || !d_ac || !d_anb || !d_bc || !d_dphi) {
    set_error("fs_solve_equation_device_from_guess: NULL argument");
    return -1;
  }
  if (lmultigrid < FS_MG_OFF || lmultigrid > FS_MG_PRECOND) {
    set_error("fs_solve_equation_device_from_guess: LMultigrid must be 0, 1 or 2");
    return -1;
  }
  return fs::solve_device(m, nf, d_fit, rrf, itmax, solver, lmultigrid, d_ac, d_anb, d_bc, d_dphi,
                          res4, stats, true);
  FS_GUARD_END
}

int fs_solve_equation(fs_mesh *m, int nf, double *fit, double rrf, int itmax, int solver,
                      int lmultigrid, const double *ac, const double *anb, const double *bc,
                      double *dphi, double *res4, fs_solve_stats *stats) {
  FS_GUARD_BEGIN
  if (!m || !fit || !ac || !anb || !bc || !dphi) {
    set_error("fs_solve_equation: NULL argument");
    return -1;
  }
  const size_t n = m->n, nw = (size_t)m->n * m->w;
  fs::DevBuf<double> dfit, dac, danb, dbc, ddphi;
  dfit.upload(fit, n, m->stream);
  dac.upload(ac, n, m->stream);
  danb.upload(anb, nw, m->stream);
  dbc.upload(bc, n, m->stream);
  ddphi.alloc(n);
  const int rc = fs::solve_device(m, nf, dfit.p, rrf, itmax, solver, lmultigrid, dac.p, danb.p,
                                  dbc.p, ddphi.p, res4, stats);
  if (rc < 0) return rc; // host arrays untouched
  FS_HIP(hipMemcpyAsync(fit, dfit.p, sizeof(double) * n, hipMemcpyDeviceToHost, m->stream));
  FS_HIP(hipMemcpyAsync(dphi, ddphi.p, sizeof(double) * n, hipMemcpyDeviceToHost, m->stream));
  FS_HIP(hipStreamSynchronize(m->stream));
  return rc;
  FS_GUARD_END
}

// Inspection of the AMG hierarchy built by the last multigrid solve (tests).
int fs_amg_export(fs_mesh *m, int level, int *dims, int *parent, int *seed, int *nchild,
                  int *child_ptr, int *child, int *nbr, int *nnbr, double *ac, double *anb) {
  FS_GUARD_BEGIN
  if (!m) return -1;
  return fs::amg_export(m, level, dims, parent, seed, nchild, child_ptr, child, nbr, nnbr, ac, anb);
  FS_GUARD_END
}

} // extern "C"
