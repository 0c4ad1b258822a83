// Plain FP64 back-projection for gfx950 -- the parity path.
//
// Semantics: process_point, sar_focuser_puntual.py:21-155 (the Numba path), not the CUDA kernels
// of gpu_small_dataset.py / main_gpu.py, which round half away from zero and (main_gpu.py:129)
// accept the last range bin.  Every FP64 operation below is one IEEE rounding in the reference's
// order; this file is compiled with -ffp-contract=off so no multiply-add is fused, and sqrt and
// division are the correctly rounded f64 expansions.  The only difference left from the CPU path
// is the last-ulp behaviour of the device sin/cos.
#include "sarbp_internal.h"

#pragma clang fp contract(off)

namespace {

// One pixel per lane, each lane walking its own aperture in ascending pulse order.
__global__ __launch_bounds__(256) void k_bp_f64_lane(BpGeom g, PixelsView px, double2 *__restrict__ out,
                                                     int32_t *__restrict__ ncnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= px.n) return;
  double acc_re = 0.0, acc_im = 0.0;
  int32_t cnt = 0;
  const int32_t s = px.sint[i];
  if (px.mask[i] == 1 && s != -9999) {          // sar_focuser_puntual.py:46-53, 208-211
    int64_t lo, hi;
    sarbp_pulse_range(px.index[i], s, g.naz, lo, hi);
    const double x = px.x[i], y = px.y[i], z = px.z[i];
    const double lim = (double)(g.nra - 1);
    for (int64_t a = lo; a < hi; ++a) {
      const double4 t = g.traj[a];
      const double dx = x - t.x;                 // sar_focuser_puntual.py:82-84
      const double dy = y - t.y;
      const double dz = z - t.z;
      const double r2 = dx * dx + dy * dy + dz * dz;  // :88
      const double r = __builtin_sqrt(r2);            // :94
      const double q = __builtin_rint((r - g.near_range) / g.dra);  // :100-101, half to even
      if (!(q < lim && q >= 0.0)) continue;           // :106
      const double ph = g.kphase * r + g.beta * r2;   // :116
      double sn, cs;
      sincos(ph, &sn, &cs);                           // exp(-j ph) = (cs, -sn), :121
      const float2 d = g.raw[(int64_t)q + g.nra * a];
      const double t1r = r2 * (double)d.x;            // :144, R2 * raw
      const double t1i = r2 * (double)d.y;
      acc_re += t1r * cs + t1i * sn;                  // t1r*cs - t1i*(-sn)
      acc_im += t1i * cs - t1r * sn;                  // t1r*(-sn) + t1i*cs
      ++cnt;
    }
  }
  double2 o;
  o.x = cnt > 0 ? acc_re / (double)cnt : 0.0;        // :150-152, componentwise mean
  o.y = cnt > 0 ? acc_im / (double)cnt : 0.0;
  out[i] = o;
  ncnt[i] = cnt;
}

}  // namespace

int sarbp_launch_f64(const sarbp_dataset *ds, const sarbp_scene *sc, hipStream_t st) {
  if (sc->n == 0) return SARBP_OK;
  const int64_t blocks = (sc->n + 255) / 256;
  if (blocks > 0x7fffffff) return sarbp_fail(SARBP_EINVAL, "too many pixels for one launch");
  hipLaunchKernelGGL(k_bp_f64_lane, dim3((unsigned)blocks), dim3(256), 0, st, ds->geom(), sc->view(),
                     sc->out, sc->ncnt);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return sarbp_hip_fail(e, "k_bp_f64_lane launch");
  return SARBP_OK;
}
