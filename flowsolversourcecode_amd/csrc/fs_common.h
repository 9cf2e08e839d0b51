// Internal declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "fsamd.h"

namespace fs {

void set_error(const std::string &msg);

#define FS_HIP(call)                                                          \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ::fs::set_error(std::string(#call) + ": " + hipGetErrorString(e_));     \
      return -2;                                                              \
    }                                                                         \
  } while (0)

#define FS_HIP_THROW(call)                                                    \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      throw std::runtime_error(std::string(#call) + ": " +                    \
                               hipGetErrorString(e_));                        \
  } while (0)

constexpr int kBlock = 256;     // threads per block for streaming kernels
constexpr int kPartials = 8192; // fixed number of reduction partials (one row per thread up to 2.1 M rows)

// Device buffer owning a hipMalloc allocation.
template <class T> struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count <= n && p) return;
    release();
    FS_HIP_THROW(hipMalloc(&p, sizeof(T) * (count ? count : 1)));
    n = count;
  }
  void upload(const T *h, size_t count, hipStream_t s) {
    alloc(count);
    FS_HIP_THROW(hipMemcpyAsync(p, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T> &h, hipStream_t s) { upload(h.data(), h.size(), s); }
};

struct Hierarchy; // fs_amg.hip

// Per-launch HIP-event timing of one kernel site of the solve loops: events
// are recorded by the dispatch itself (hipExtLaunchKernelGGL), tagged with
// the Krylov iteration that launched them, so launches after convergence
// (which return at once) can be told from the ones that did the work.
struct KernelClock {
  int sites = 0, cap = 0, used = 0;  // sites: bit (1 << site) per armed site
  int dropped[4] = {0, 0, 0, 0};
  int stride = 1;                    // time every stride-th launch of a site
  long long seen[4] = {0, 0, 0, 0};  // launches of each site since the start
  int cur_iter = -1;       // Krylov iteration being launched, -1 outside a loop
  int solve_first = 0;     // first record of the solve in progress
  std::vector<hipEvent_t> ev;  // 2 per record
  std::vector<int> iter, site;
  std::vector<double> bytes;
  std::vector<char> live;
};

} // namespace fs

struct fs_mesh {
  int n = 0, w = 0;
  int device = 0;           // the device current at fs_mesh_create
  hipStream_t stream = nullptr;
  bool own_stream = true;   // false: the stream belongs to another context
  std::vector<int> h_nbr;   // reference layout, 1-based, 0 = boundary
  std::vector<int> h_nnbr;
  std::vector<int> h_nbr_dev; // 0-based, -1 = none (the device layout)
  fs::DevBuf<int> d_nbr;    // 0-based, -1 = none (slots past nnbr as well)
  fs::DevBuf<int> d_nbr16;  // the same as int16 offsets k - i (0 = none), two
                            // slots per word, when every offset fits
  fs::DevBuf<int> d_mir;    // transpose map, linear anb index or -1
  // exact natural-order D-ILU schedule (FactorizeUsingILU.for:29-75)
  int ilu_nlevels = 0;
  std::vector<int> ilu_level_ptr;
  fs::DevBuf<int> d_ilu_cells;     // cells sorted by level
  fs::DevBuf<int> d_ilu_ptr;       // [n+1] per-cell lower terms
  fs::DevBuf<int> d_ilu_a, d_ilu_i, d_ilu_b; // anb(j,s) idx, i, anb(i,s') idx
  // face topology (optional)
  int nif = 0, nfaces_total = 0;
  std::vector<int> h_owner, h_neigh;
  fs::DevBuf<int> d_owner, d_neigh, d_iown, d_inbr;
  fs::DevBuf<int> d_face_code;  // per (cell, slot): face*4 + role
  // work vectors
  fs::DevBuf<double> v[12];
  fs::DevBuf<double> partial;  // kPartials * 5
  fs::DevBuf<double> scal;     // scalar slots
  fs::DevBuf<int> flags;       // [0] done, [1] iterations
  // The scratch the solve uses: the mesh's buffers above, or slices of the
  // library's static device array for the duration of a
  // fs_solve_equation_device_noalloc call.
  struct Scratch {
    double *v[12];
    double *dc;       // D-ILU diagonal
    double *partial;  // kPartials * 5
    double *scal;
    int *flags;
  } wk{};
  // options
  int mg_pre = 2, mg_post = 2, mg_maxcycles = 30, mg_maxlevels = 10;
  int mg_minparents = 10, mg_reagg = 300, mg_fp32 = 0, mg_smoother = 0;
  int mg_agg = 0;           // FS_OPT_MG_AGGREGATION
  int mg_fuse_prolong = 1;  // FS_OPT_MG_FUSE_PROLONG
  int mg_cycle = 0;         // FS_OPT_MG_CYCLE: 0 V, 1 F, 2 W
  int warm_prev = 0;        // FS_OPT_WARM_PREVIOUS (read by the SIMPLE case)
  double mg_omega = 0.8;
  int mg_tail_cells = 1024, mg_coarse_sweeps = 16;
  int mg_cpre = -1, mg_cpost = -1; // sweeps on the coarse levels (-1: as on the finest)
  int mg_sell_rows = 100000;       // coarse levels at least this large use sliced ELL
  double mg_alpha = 0.0;           // over-correction (0: 1 as solver, 1.7 as preconditioner)
  int gmres_m = 30, check_every = 8;
  int iter_pred[256] = {0}; // iterations of the last solve per (variable, solver)
  int ncalls_mg = 0;
  long long nnz = 0;                    // off-diagonal entries of the ELL pattern
  fs::KernelClock kclock;
  void *solve_work = nullptr;           // fs_solve.hip
  std::shared_ptr<fs::Hierarchy> amg;   // fs_amg.hip
};
