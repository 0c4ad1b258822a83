// Device helpers shared by the FP64 and split-precision kernels.
#pragma once

#include "sarbp_internal.h"

namespace {

// ---- FP64 device math for the parity path --------------------------------------------------
//
// sqrt_rn: the correctly rounded square root (hardware 1/sqrt estimate, then the Newton-Raphson /
// FMA sequence whose last step rounds correctly, P. Markstein, "IA-64 and Elementary Functions",
// 2000, ch. 8) for arguments whose square root needs no rescaling; others take __builtin_sqrt
// (WIDE = false: the caller guarantees x in [2^-900, 2^900] or a NaN, which comes out NaN either way).
template <bool WIDE = true>
__device__ __forceinline__ double sqrt_rn(double x) {
  if constexpr (WIDE)
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

// sin and cos of x for |x| < 2^25, table-driven (P. T. P. Tang, "Table-driven implementation of
// the logarithm / exponential function", ACM TOMS 15-16, 1989-90, applied to sin/cos): x = n h + r
// with h = pi/128 in two parts (the first product and difference are exact with an FMA),
// |r| <= pi/256; with (S, C) = (sin, cos)(j h), j = n mod 256, from tab:
//   sin x = S cos r + C sin r,  cos x = C cos r - S sin r,
// sin r to r^5 and cos r to r^6 (truncation below 1e-17 absolute).  tab may live in LDS or global
// memory (same values, same arithmetic).  Larger |x| (slant ranges beyond ~80 km at lambda =
// 3 cm) take the library sincos (WIDE = false: the caller guarantees |x| < 2^25, and the library
// path and its registers are left out).
// The two halves of the table method for |x| < 2^25: sincos_reduce finds r and the table entry
// (S, C) = (sin, cos)(j h), sincos_finish combines them; callers may issue the table read early and
// finish later.
__device__ __forceinline__ void sincos_reduce(double x, const double2 *tab, double2 &e, double &r) {
  constexpr double INV_H = 40.743665431525205957;            // 128/pi
  constexpr double H_HI = 3.141592653589793116 / 128.0;      // pi rounded to double, / 128
  constexpr double H_LO = 1.2246467991473531772e-16 / 128.0; // (pi - that) / 128
  // n = the integer nearest x / h, from one fused multiply-add: x * (128/pi) + 1.5 * 2^52 (|x| < 2^25,
  // so the sum lies in [2^52, 2^53), where its rounding keeps the integer part); the low word of
  // the sum holds n modulo 2^32, the table index
  constexpr double MAGIC = 6755399441055744.0;  // 1.5 * 2^52
  const double nm = fma(x, INV_H, MAGIC);
  const double n = nm - MAGIC;
  e = tab[(int)(unsigned)__builtin_bit_cast(unsigned long long, nm) & (SARBP_SCTAB - 1)];  // (sin, cos)(j h)
  r = fma(-n, H_LO, fma(-n, H_HI, x));
}
// fma(a, b, c) with all three operands in vector registers: the same IEEE fused multiply-add as
// fma(), written so that the compiler cannot turn it into the two-address form whose accumulator
// would be a per-use copy of the constant c (b and c are loop-invariant constants here, kept in
// registers across the loop)
__device__ __forceinline__ double fma_vvv(double a, double b, double c) {
  double d;
  __asm__("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}

__device__ __forceinline__ void sincos_finish(double2 e, double r, double &sn, double &cs) {
  const double r2 = r * r;
  const double sr = fma(r * r2, fma_vvv(r2, 1.0 / 120.0, -1.0 / 6.0), r);
  const double cr = fma(r2, fma(r2, fma_vvv(r2, -1.0 / 720.0, 1.0 / 24.0), -0.5), 1.0);
  sn = fma(e.x, cr, e.y * sr);
  cs = fma(e.y, cr, -(e.x * sr));
}

template <bool WIDE = true>
__device__ __forceinline__ void sincos_tab(double x, const double2 *tab, double &sn, double &cs) {
  if constexpr (WIDE)
    if (!(__builtin_fabs(x) < 0x1p25)) {
      sincos(x, &sn, &cs);
      return;
    }
  double2 e;
  double r;
  sincos_reduce(x, tab, e, r);
  sincos_finish(e, r, sn, cs);
}

// rint((r - near) / dra), the range bin of sar_focuser_puntual.py:100-101, from a product with
// 1/dra; the quotient is divided exactly only when the product lies within its error bound of a
// half-integer, where the two roundings could differ.  The product's error is below 2^-49 |qe| +
// 2^-40 (the rounding of 1/dra and of the product), under 2^-24 for |qe| < 2^24; beyond that both
// roundings are bins >= 2^24 - 1, outside every swath (nra <= 2^24), so one constant bound serves.
__device__ __forceinline__ double range_bin(const BpGeom &g, double r) {
#pragma clang fp contract(off)
  const double num = r - g.near_range;
  const double qe = num * g.inv_dra;
  const double q = __builtin_rint(qe);
  if (__builtin_fabs(qe - q) > 0.5 - 0x1p-24) return __builtin_rint(num / g.dra);
  return q;
}

// Range bin (as a double), R^2 and R of one pixel-pulse pair of process_point
// (sar_focuser_puntual.py:82-106), without the swath test.
__device__ __forceinline__ void f64_range(const BpGeom &g, double x, double y, double z, double tx, double ty,
                                          double tz, double &q, double &r2, double &r) {
#pragma clang fp contract(off)
  const double dx = x - tx;                  // :82-84
  const double dy = y - ty;
  const double dz = z - tz;
  r2 = dx * dx + dy * dy + dz * dz;          // :88
  r = sqrt_rn<false>(r2);                    // :94
  q = range_bin(g, r);                       // :100-101, half to even
}

// Range bin (as a double), R^2 and R of one pixel-pulse pair of process_point
// (sar_focuser_puntual.py:82-106).  Returns false when the bin falls outside [0, nra-1).
template <bool WIDE = true>
__device__ __forceinline__ bool f64_bin(const BpGeom &g, double x, double y, double z, double tx, double ty,
                                        double tz, double &q, double &r2, double &r) {
#pragma clang fp contract(off)
  const double dx = x - tx;                  // :82-84
  const double dy = y - ty;
  const double dz = z - tz;
  r2 = dx * dx + dy * dy + dz * dz;          // :88
  r = sqrt_rn<WIDE>(r2);                     // :94
  q = range_bin(g, r);                       // :100-101, half to even
  return q < (double)(g.nra - 1) && q >= (double)g.bin_lo;  // :106 (bin_lo: ind_line_ra_min, >= 0)
}

// sin and cos of the phase of sar_focuser_puntual.py:116: exp(-j ph) = (cs, -sn) (:121); tab is
// the dataset's sin/cos table (in LDS or global memory).
template <bool WIDE = true>
__device__ __forceinline__ void f64_phase(const BpGeom &g, const double2 *tab, double r, double r2, double &sn,
                                          double &cs) {
#pragma clang fp contract(off)
  const double ph = g.kphase * r + g.beta * r2;
  sincos_tab<WIDE>(ph, tab, sn, cs);
}

// Copies the dataset's sin/cos table into the block's LDS (all threads of the block call this).
__device__ __forceinline__ void load_sctab(const BpGeom &g, double2 *lds) {
  for (int i = threadIdx.x; i < SARBP_SCTAB; i += blockDim.x) lds[i] = g.sctab[i];
  __syncthreads();
}

// term = R2 * raw * exp(-j ph), added to the running sums (:144-147).
__device__ __forceinline__ void f64_acc(double r2, double sn, double cs, float2 d, double &acc_re, double &acc_im) {
#pragma clang fp contract(off)
  const double t1r = r2 * (double)d.x;       // :144, R2 * raw
  const double t1i = r2 * (double)d.y;
  // the two products of each component go into the running sum with fused multiply-adds (one
  // rounding each instead of three: as accurate or more than t1r*cs + t1i*sn then +=)
  acc_re = fma(t1r, cs, fma(t1i, sn, acc_re));    // + t1r*cs - t1i*(-sn)
  acc_im = fma(t1i, cs, fma(-t1r, sn, acc_im));   // + t1r*(-sn) + t1i*cs
}

// One pixel-pulse term against the pulse column col.  Returns false when the range bin falls
// outside [0, nra-1).
__device__ __forceinline__ bool f64_term(const BpGeom &g, const double2 *tab, double x, double y, double z,
                                         double tx, double ty, double tz, const float2 *__restrict__ col,
                                         double &acc_re, double &acc_im) {
  double q, r2, r, sn, cs;
  if (!f64_bin(g, x, y, z, tx, ty, tz, q, r2, r)) return false;
  const float2 d = col[(int)q];
  f64_phase(g, tab, r, r2, sn, cs);
  f64_acc(r2, sn, cs, d, acc_re, acc_im);
  return true;
}

// A gather byte offset beyond every column (Nra <= 2^24): the buffer's range check returns zero
// without touching memory.
constexpr unsigned RAW_NO_SAMPLE = 0x80000000u;

// Buffer resource of the RawData column starting at col (wave-uniform: scalar registers), bins
// [0, Nra-1): the last bin, which never contributes (sar_focuser_puntual.py:106), reads as zero too.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_column_at(const float2 *col, int64_t nra) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)col, (short)0, (int)((nra - 1) * (int64_t)sizeof(float2)),
                                           0x00020000);
}

// Buffer resource of pulse a's RawData column.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_column(const BpGeom &g, int a) {
  return raw_column_at(g.raw + (int64_t)a * g.nra, g.nra);
}

// Buffer resource of the split path's gathers: col points at range bin bin_lo of a pulse column, and
// nbin_bytes = (nra - 1 - bin_lo) * 8 covers the contributing bins [bin_lo, nra-1).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_split_column(const float2 *col, unsigned nbin_bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)col, (short)0, (int)nbin_bytes, 0x00020000);
}

__device__ __forceinline__ float2 raw_load(__amdgpu_buffer_rsrc_t col, unsigned byte_off) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(col, (int)byte_off, 0, 0));
}

// Buffer resource of four consecutive pulse columns of the split path: col points at range bin
// bin_lo of the first, and the range ends after bin nra-2 of the fourth (3 nra x 8 + nbin_bytes
// bytes), so every offset it accepts lies inside RawData.  Pulse k of the four is read with the
// scalar offset k x nra x 8, which the range check includes; the check does not stop an offset at
// the end of its own column, so callers pass offsets of bins inside [bin_lo, nra-1) only (or one
// beyond every column, which reads zero).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_split_group(const float2 *col, int64_t nra, unsigned nbin_bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)col, (short)0,
                                           (int)(3 * nra * (int64_t)sizeof(float2) + nbin_bytes), 0x00020000);
}

__device__ __forceinline__ float2 raw_load_at(__amdgpu_buffer_rsrc_t r, unsigned byte_off, unsigned soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)byte_off, (int)soff, 0));
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
// (sar_focuser_puntual.py:82-106); -1 when outside [bin_lo, nra-1).
[[maybe_unused]] __device__ __forceinline__ int exact_bin(const BpGeom &g, double x, double y, double z, double tx, double ty,
                                      double tz) {
#pragma clang fp contract(off)
  const double dx = x - tx;
  const double dy = y - ty;
  const double dz = z - tz;
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double r = __builtin_sqrt(r2);
  const double q = __builtin_rint((r - g.near_range) / g.dra);
  return (q < (double)(g.nra - 1) && q >= (double)g.bin_lo) ? (int)q : -1;
}

}  // namespace
