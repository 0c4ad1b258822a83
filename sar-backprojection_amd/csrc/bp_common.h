// Device helpers shared by the FP64 and split-precision kernels.
#pragma once

#include "sarbp_internal.h"

namespace {

// ---- FP64 device math for the parity path --------------------------------------------------
//
// sqrt_rn: the correctly rounded square root (hardware 1/sqrt estimate, then the Newton-Raphson /
// FMA sequence whose last step rounds correctly, P. Markstein, "IA-64 and Elementary Functions",
// 2000, ch. 8) for arguments whose square root needs no rescaling; others take __builtin_sqrt.
__device__ __forceinline__ double sqrt_rn(double x) {
  if (!(x >= 0x1p-900 && x <= 0x1p900)) return __builtin_sqrt(x);
  const double y = __builtin_amdgcn_rsq(x);
  double gg = x * y, h = 0.5 * y;
  const double e = fma(-h, gg, 0.5);
  gg = fma(gg, e, gg);
  h = fma(h, e, h);
  double d = fma(-gg, gg, x);
  gg = fma(d, h, gg);
  d = fma(-gg, gg, x);
  return fma(d, h, gg);
}

// sin and cos of x for |x| < 2^30: Cody-Waite reduction by pi/2 in two parts (the first
// product and difference are exact with an FMA, so the reduced argument carries only its final
// rounding) and Taylor polynomials on |r| <= pi/4 (+ a rounding of n): sin to r^17, cos to r^16,
// truncation below 1e-18.  Larger |x| (never a slant-range phase) take the library sincos.
__device__ __forceinline__ void sincos_rd(double x, double &sn, double &cs) {
  if (!(__builtin_fabs(x) < 0x1p30)) {
    sincos(x, &sn, &cs);
    return;
  }
  constexpr double TWO_OVER_PI = 0.63661977236758134308;
  constexpr double PIO2_HI = 1.5707963267948966192;     // pi/2 rounded to double
  constexpr double PIO2_LO = 6.123233995736766036e-17;  // pi/2 - PIO2_HI
  const double n = __builtin_rint(x * TWO_OVER_PI);
  const double r = fma(-n, PIO2_LO, fma(-n, PIO2_HI, x));
  const double r2 = r * r;
  double ps = 1.0 / 355687428096000.0;  // 1/17!
  ps = fma(ps, r2, -1.0 / 1307674368000.0);
  ps = fma(ps, r2, 1.0 / 6227020800.0);
  ps = fma(ps, r2, -1.0 / 39916800.0);
  ps = fma(ps, r2, 1.0 / 362880.0);
  ps = fma(ps, r2, -1.0 / 5040.0);
  ps = fma(ps, r2, 1.0 / 120.0);
  ps = fma(ps, r2, -1.0 / 6.0);
  const double sr = fma(r * r2, ps, r);
  double pc = 1.0 / 20922789888000.0;  // 1/16!
  pc = fma(pc, r2, -1.0 / 87178291200.0);
  pc = fma(pc, r2, 1.0 / 479001600.0);
  pc = fma(pc, r2, -1.0 / 3628800.0);
  pc = fma(pc, r2, 1.0 / 40320.0);
  pc = fma(pc, r2, -1.0 / 720.0);
  pc = fma(pc, r2, 1.0 / 24.0);
  pc = fma(pc, r2, -0.5);
  const double cr = fma(r2, pc, 1.0);
  const int q = (int)n & 3;  // x = r + q pi/2 (mod 2 pi)
  const double s0 = (q & 1) ? cr : sr, c0 = (q & 1) ? sr : cr;
  sn = (q & 2) ? -s0 : s0;
  cs = ((q + 1) & 2) ? -c0 : c0;
}

// rint((r - near) / dra), the range bin of sar_focuser_puntual.py:100-101, from a product with
// 1/dra; the quotient is divided exactly only when the product lies within its error bound of a
// half-integer, where the two roundings could differ.
__device__ __forceinline__ double range_bin(const BpGeom &g, double r) {
#pragma clang fp contract(off)
  const double num = r - g.near_range;
  const double qe = num * g.inv_dra;
  const double q = __builtin_rint(qe);
  if (__builtin_fabs(qe - q) > fma(__builtin_fabs(qe), -0x1p-49, 0.5 - 0x1p-40)) return __builtin_rint(num / g.dra);
  return q;
}

// Range bin (as a double), R^2 and R of one pixel-pulse pair of process_point
// (sar_focuser_puntual.py:82-106).  Returns false when the bin falls outside [0, nra-1).
__device__ __forceinline__ bool f64_bin(const BpGeom &g, double x, double y, double z, double tx, double ty,
                                        double tz, double &q, double &r2, double &r) {
#pragma clang fp contract(off)
  const double dx = x - tx;                  // :82-84
  const double dy = y - ty;
  const double dz = z - tz;
  r2 = dx * dx + dy * dy + dz * dz;          // :88
  r = sqrt_rn(r2);                           // :94
  q = range_bin(g, r);                       // :100-101, half to even
  return q < (double)(g.nra - 1) && q >= 0.0;  // :106
}

// sin and cos of the phase of sar_focuser_puntual.py:116: exp(-j ph) = (cs, -sn) (:121).
__device__ __forceinline__ void f64_phase(const BpGeom &g, double r, double r2, double &sn, double &cs) {
#pragma clang fp contract(off)
  const double ph = g.kphase * r + g.beta * r2;
  sincos_rd(ph, sn, cs);
}

// term = R2 * raw * exp(-j ph), added to the running sums (:144-147).
__device__ __forceinline__ void f64_acc(double r2, double sn, double cs, float2 d, double &acc_re, double &acc_im) {
#pragma clang fp contract(off)
  const double t1r = r2 * (double)d.x;       // :144, R2 * raw
  const double t1i = r2 * (double)d.y;
  acc_re += t1r * cs + t1i * sn;             // t1r*cs - t1i*(-sn)
  acc_im += t1i * cs - t1r * sn;             // t1r*(-sn) + t1i*cs
}

// One pixel-pulse term against the pulse column col.  Returns false when the range bin falls
// outside [0, nra-1).
__device__ __forceinline__ bool f64_term(const BpGeom &g, double x, double y, double z, double tx, double ty,
                                         double tz, const float2 *__restrict__ col, double &acc_re,
                                         double &acc_im) {
  double q, r2, r, sn, cs;
  if (!f64_bin(g, x, y, z, tx, ty, tz, q, r2, r)) return false;
  const float2 d = col[(int)q];
  f64_phase(g, r, r2, sn, cs);
  f64_acc(r2, sn, cs, d, acc_re, acc_im);
  return true;
}

// Buffer resource of pulse a's RawData column (wave-uniform: scalar registers), bins [0, Nra-1):
// the last bin, which never contributes (sar_focuser_puntual.py:106), reads as zero too.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_column(const BpGeom &g, int a) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)(g.raw + (int64_t)a * g.nra), (short)0,
                                           (int)((g.nra - 1) * (int64_t)sizeof(float2)), 0x00020000);
}

__device__ __forceinline__ float2 raw_load(__amdgpu_buffer_rsrc_t col, unsigned byte_off) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(col, (int)byte_off, 0, 0));
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
