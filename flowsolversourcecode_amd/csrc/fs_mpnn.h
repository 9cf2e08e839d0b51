// Internal entry points of the warm-start network used by the SIMPLE case.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "fsamd_learn.h"

namespace fs {

// x = prediction for one system on the network's topology (device FP64
// arrays, the case's layouts), launched on `stream`
int mpnn_predict_device(fs_mpnn *n, hipStream_t stream, const double *b, const double *ac, const double *anb,
                        const double *vol, const double *xp, double *x);
// the network was built on this topology (0-based neighbours, -1 = none)
bool mpnn_matches(fs_mpnn *n, int n_cells, int width, const std::vector<int> &nbr0);
// appends one device system to the network's dataset; false when it is full
bool mpnn_capture(fs_mpnn *n, hipStream_t stream, const double *b, const double *x, const double *ac,
                  const double *anb, const double *vol, const double *xp);

} // namespace fs
