// High-resolution convection schemes and gradient limiting in the
// reference's arithmetic, shared by the device kernels and host validation.
//
// MIN/MAX follow the reference build's evaluation: left to right, a NaN
// operand after the first is passed over (so a 0/0 ratio of a flat field
// never reaches a limiter).
#pragma once

#include <hip/hip_runtime.h>

#include "fsamd_simple.h"

namespace fs {

__host__ __device__ inline double min2(double a, double b) { return (a < b || b != b) ? a : b; }
__host__ __device__ inline double max2(double a, double b) { return (a > b || b != b) ? a : b; }

// psi(r) of the TVD form (LimitersTVDType.for:44-115).
__host__ __device__ inline double tvd_psi(int scheme, double r) {
  switch (scheme) {
  case FS_CONV_DOWNWIND: return 2.;
  case FS_CONV_FROMM: return 0.5 * (1. + r);
  case FS_CONV_SOU: return r;
  case FS_CONV_CD: return 1.;
  case FS_CONV_QUICK: return (3. + r) / 4.;
  case FS_CONV_SUPERBEE: return max2(max2(0., min2(1., 2. * r)), min2(2., r));
  case FS_CONV_MINMOD: return max2(0., min2(1., r));
  case FS_CONV_OSHER: return max2(0., min2(2., r));
  case FS_CONV_VANLEER: return (r + fabs(r)) / (1. + fabs(r));
  case FS_CONV_MUSCL: return max2(0., min2(min2(2. * r, 0.5 * (1. + r)), 2.));
  case FS_CONV_SMART: return max2(0., min2(min2(2. * r, 0.75 * r + 0.25), 4.));
  case FS_CONV_HQUICK: return 2. * (r + fabs(r)) / (r + 3.);
  case FS_CONV_UMIST: return max2(0., min2(min2(min2(2. * r, 0.25 + 0.75 * r), 0.75 + 0.25 * r), 2.));
  case FS_CONV_CHARM: return r <= 0. ? 0. : r * (3. * r + 1.) / ((r + 1.) * (r + 1.));
  default: return 0.; // upwind
  }
}

// Normalised face value of the NVF form (LimitersNVFType.for:38-112);
// BettaConvection is the single-precision literal 0.1 (User0.F90:408).
__host__ __device__ inline double nvf_face(int scheme, double t) {
  const bool in01 = t >= 0. && t <= 1.;
  switch (scheme) {
  case FS_CONV_SMART: return in01 ? min2(min2(3. * t, 0.375 + 0.75 * t), (2. + t) / 3.) : t;
  case FS_CONV_MINMOD: return in01 ? min2(1.5 * t, 0.5 * t + 0.5) : t;
  case FS_CONV_BOUNDEDCD: return in01 ? 0.5 * t + 0.5 : t;
  case FS_CONV_CD: return 0.5 * t + 0.5;
  case FS_CONV_OSHER: return in01 ? min2(1.5 * t, 1.) : t;
  case FS_CONV_STOIC: return in01 ? min2(min2(0.5 * t + 0.5, 0.375 + 0.75 * t), (2. + t) / 3.) : t;
  case FS_CONV_MUSCL: return in01 ? min2(min2(2. * t, 0.25 + t), 1.) : t;
  case FS_CONV_SUPERBEE: return in01 ? min2(min2(min2(2. * t, 0.5 * t + 0.5), 1.5 * t), 1.) : t;
  case FS_CONV_DOWNWIND: return 1.;
  case FS_CONV_GAMMA: {
    const double beta = (double)0.1f;
    if (t > 0. && t < beta) {
      const double factor = 0.5 / beta;
      return (-factor * t + (1. + factor)) * t;
    }
    if (t >= beta && t < 1.) return 0.5 + 0.5 * t;
    return t;
  }
  default: return t; // upwind
  }
}

// The schemes each framework defines for a steady single-phase flow (the
// interface-capturing stacs, hric and cicsam need a volume fraction).
inline bool hr_scheme_defined(int framework, int scheme) {
  if (framework == FS_HR_TVD) return scheme >= FS_CONV_UPWIND && scheme <= FS_CONV_CHARM;
  if (framework == FS_HR_NVF) {
    switch (scheme) {
    case FS_CONV_UPWIND: case FS_CONV_SMART: case FS_CONV_MINMOD: case FS_CONV_BOUNDEDCD: case FS_CONV_CD:
    case FS_CONV_OSHER: case FS_CONV_STOIC: case FS_CONV_MUSCL: case FS_CONV_SUPERBEE: case FS_CONV_DOWNWIND:
    case FS_CONV_GAMMA: return true;
    default: return false;
    }
  }
  return framework == FS_HR_NONE;
}

// phi_HR - phi_C of one interior face (AssembleHRSchemeCorrection.for:39-105):
// phic/phid the upwind/downwind values, gu the upwind cell's gradient, d the
// distance vector from the upwind to the downwind centroid.
template <int FW>
__host__ __device__ inline double hr_increment(int scheme, double bleed, double phic, double phid,
                                               double gx, double gy, double gz, double dx, double dy,
                                               double dz) {
  const double small = (double)1.e-9f;
  const double phiu = phid - 2. * (gx * dx + gy * dy + gz * dz);
  double phihr;
  if constexpr (FW == FS_HR_NVF) {
    const double diff = phid - phiu;
    const double t = fabs(diff) < small ? 0. : (phic - phiu) / diff;
    phihr = nvf_face(scheme, t) * diff + phiu;
    if (scheme != FS_CONV_GAMMA) phihr = (1. - bleed) * phihr + bleed * phic;
  } else {
    const double diff = phid - phic;
    const double r = fabs(diff) < small ? 0. : (phic - phiu) / diff;
    phihr = phic + 0.5 * tvd_psi(scheme, r) * diff;
    phihr = (1. - bleed) * phihr + bleed * phic;
  }
  return phihr - phic;
}

} // namespace fs
