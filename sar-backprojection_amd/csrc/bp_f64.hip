// Plain FP64 back-projection for gfx950 -- the parity path.
//
// Semantics: process_point, sar_focuser_puntual.py:21-155 (the Numba path), not the CUDA kernels
// of gpu_small_dataset.py / main_gpu.py, which round half away from zero and (main_gpu.py:129)
// accept the last range bin.  Every FP64 operation below is one IEEE rounding in the reference's
// order; this file is compiled with -ffp-contract=off so no multiply-add is fused, and sqrt and
// division are the correctly rounded f64 expansions.  The only difference left from the CPU path
// is the last-ulp behaviour of the device sin/cos.
//
// Two kernels:
//   k_bp_f64_lane  one pixel per lane, each lane walking its own aperture (no tiling);
//   k_bp_f64_tile  one 64-pixel tile per wavefront, walking the union of the tile's apertures
//                  pulse by pulse: the pulse index is wave-uniform, so the trajectory sample is a
//                  scalar load and the RawData gathers of neighbouring pixels fall in the same or
//                  adjacent range bins of one pulse column (few cache lines per load).
// Both accumulate each pixel's terms in ascending pulse order, as the reference does.
#include "bp_tile.h"

#pragma clang fp contract(off)

namespace {

__global__ __launch_bounds__(256) void k_bp_f64_lane(BpGeom g, PixelsView px, double2 *__restrict__ out,
                                                     int32_t *__restrict__ ncnt) {
  __shared__ double2 tab[SARBP_SCTAB];
  load_sctab(g, tab);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= px.n) return;
  double acc_re = 0.0, acc_im = 0.0;
  int32_t cnt = 0;
  const int32_t s = px.sint[i];
  if (px.mask[i] == 1 && s != -9999) {  // sar_focuser_puntual.py:46-53, 208-211
    int64_t lo, hi;
    sarbp_pulse_range(px.index[i], s, g.naz, lo, hi);
    const double x = px.x[i], y = px.y[i], z = px.z[i];
    for (int64_t a = lo; a < hi; ++a) {
      const double4 t = g.traj[a];
      if (f64_term(g, tab, x, y, z, t.x, t.y, t.z, g.raw + g.nra * a, acc_re, acc_im)) ++cnt;
    }
  }
  store_mean(out, ncnt, i, acc_re, acc_im, cnt);
}

// One wavefront per tile; block = 64 threads.  WIDE = false when the scene's distance bounds keep
// every R^2 and phase inside the direct ranges of sqrt_rn and sincos_tab (same bits, fewer
// registers: the library fallbacks are not compiled in).
template <bool WIDE>
__global__ __launch_bounds__(64) void k_bp_f64_tile(BpGeom g, PixelsView px, TileGrid tg, double2 *__restrict__ out,
                                                    int32_t *__restrict__ ncnt) {
  __shared__ double2 tab[SARBP_SCTAB];
  const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
  if (tile >= tg.ntiles()) return;  // whole block (one wavefront)
  load_sctab(g, tab);
  const TileLane t = tile_lane(g, px, tg, tile, threadIdx.x);
  double ar = 0.0, ai = 0.0;
  int32_t cnt = 0;
  f64_span<WIDE, 3>(g, tab, t, t.wlo, t.whi, ar, ai, cnt);
  if (t.valid) store_mean(out, ncnt, t.i, ar, ai, cnt);
}

// The per-pair arithmetic of the FP64 path (f64_bin, f64_phase) for n (pixel, pulse) pairs, for
// verification against the oracle: R2, R, bin (-1 outside [0, nra-1)), phase, sin, cos.
__global__ __launch_bounds__(256) void k_eval_pairs(BpGeom g, const double *__restrict__ x, const double *__restrict__ y,
                                                    const double *__restrict__ z, const int32_t *__restrict__ pulse,
                                                    int64_t n, double *__restrict__ r2o, double *__restrict__ ro,
                                                    int32_t *__restrict__ bino, double *__restrict__ pho,
                                                    double *__restrict__ sno, double *__restrict__ cso) {
  __shared__ double2 tab[SARBP_SCTAB];
  load_sctab(g, tab);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 t = g.traj[pulse[i]];  // pulse range checked on the host
  double q, r2, r, sn, cs;
  const bool in = f64_bin(g, x[i], y[i], z[i], t.x, t.y, t.z, q, r2, r);
  double ph;
  {
#pragma clang fp contract(off)
    ph = g.kphase * r + g.beta * r2;
  }
  f64_phase(g, tab, r, r2, sn, cs);
  r2o[i] = r2;
  ro[i] = r;
  bino[i] = in ? (int32_t)q : -1;
  pho[i] = ph;
  sno[i] = sn;
  cso[i] = cs;
}

}  // namespace

int sarbp_launch_eval_pairs(const sarbp_dataset *ds, const double *x, const double *y, const double *z,
                            const int32_t *pulse, int64_t n, double *r2, double *r, int32_t *bin, double *ph,
                            double *sn, double *cs, hipStream_t st) {
  if (n <= 0) return SARBP_OK;
  const int64_t blocks = (n + 255) / 256;
  if (blocks > 0x7fffffff) return sarbp_fail(SARBP_EINVAL, "too many pairs for one launch");
  hipLaunchKernelGGL(k_eval_pairs, dim3((unsigned)blocks), dim3(256), 0, st, ds->geom(), x, y, z, pulse, n, r2, r, bin,
                     ph, sn, cs);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return sarbp_hip_fail(e, "pair kernel launch");
  return SARBP_OK;
}

int sarbp_launch_f64(const sarbp_dataset *ds, const sarbp_scene *sc, const Tiling &t, hipStream_t st) {
  if (sc->n == 0) return SARBP_OK;
  if (t.kernel == 0) {
    const int64_t blocks = (sc->n + 255) / 256;
    if (blocks > 0x7fffffff) return sarbp_fail(SARBP_EINVAL, "too many pixels for one launch");
    hipLaunchKernelGGL(k_bp_f64_lane, dim3((unsigned)blocks), dim3(256), 0, st, sc->geom(), sc->view(), sc->out,
                       sc->ncnt);
  } else {
    const TileGrid tg = make_tile_grid(sc->dimx, sc->dimy, t.tile_x);
    const int64_t blocks = (tg.ntiles() + SARBP_NUM_XCD - 1) / SARBP_NUM_XCD * SARBP_NUM_XCD;
    if (blocks > 0x7fffffff) return sarbp_fail(SARBP_EINVAL, "too many tiles for one launch");
    hipLaunchKernelGGL(sc->f64_narrow ? k_bp_f64_tile<false> : k_bp_f64_tile<true>, dim3((unsigned)blocks), dim3(64),
                       0, st, sc->geom(), sc->view(), tg, sc->out, sc->ncnt);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return sarbp_hip_fail(e, "FP64 kernel launch");
  return SARBP_OK;
}
