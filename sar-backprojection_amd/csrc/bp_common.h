// Device helpers shared by the FP64 and split-precision kernels.
#pragma once

#include "sarbp_internal.h"

namespace {

// One pixel-pulse term of process_point (sar_focuser_puntual.py:82-146).  Returns false when the
// range bin falls outside [0, nra-1).
__device__ __forceinline__ bool f64_term(const BpGeom &g, double x, double y, double z, double tx, double ty,
                                         double tz, const float2 *__restrict__ col, double &acc_re,
                                         double &acc_im) {
#pragma clang fp contract(off)
  const double dx = x - tx;                       // :82-84
  const double dy = y - ty;
  const double dz = z - tz;
  const double r2 = dx * dx + dy * dy + dz * dz;  // :88
  const double r = __builtin_sqrt(r2);            // :94
  const double q = __builtin_rint((r - g.near_range) / g.dra);  // :100-101, half to even
  if (!(q < (double)(g.nra - 1) && q >= 0.0)) return false;      // :106
  const double ph = g.kphase * r + g.beta * r2;   // :116
  double sn, cs;
  sincos(ph, &sn, &cs);                           // exp(-j ph) = (cs, -sn), :121
  const float2 d = col[(int)q];
  const double t1r = r2 * (double)d.x;            // :144, R2 * raw
  const double t1i = r2 * (double)d.y;
  acc_re += t1r * cs + t1i * sn;                  // t1r*cs - t1i*(-sn)
  acc_im += t1i * cs - t1r * sn;                  // t1r*(-sn) + t1i*cs
  return true;
}

__device__ __forceinline__ void store_mean(double2 *out, int32_t *ncnt, int64_t i, double acc_re, double acc_im,
                                           int32_t cnt) {
#pragma clang fp contract(off)
  double2 o;
  o.x = cnt > 0 ? acc_re / (double)cnt : 0.0;  // :150-152, componentwise mean
  o.y = cnt > 0 ? acc_im / (double)cnt : 0.0;
  out[i] = o;
  ncnt[i] = cnt;
}

// Range bin of one pixel-pulse pair exactly as process_point selects it
// (sar_focuser_puntual.py:82-106); -1 when outside [0, nra-1).
[[maybe_unused]] __device__ __forceinline__ int exact_bin(const BpGeom &g, double x, double y, double z, double tx, double ty,
                                      double tz) {
#pragma clang fp contract(off)
  const double dx = x - tx;
  const double dy = y - ty;
  const double dz = z - tz;
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double r = __builtin_sqrt(r2);
  const double q = __builtin_rint((r - g.near_range) / g.dra);
  return (q < (double)(g.nra - 1) && q >= 0.0) ? (int)q : -1;
}

}  // namespace
