// Split-precision back-projection for gfx950: one wavefront per 64-pixel tile, the split-form
// terms of bp_tile.h.  Same apertures, range bins and contributor counts as process_point
// (sar_focuser_puntual.py:21-155); FP64 only for the per-(tile, pulse) reference range and phase.
#include "bp_tile.h"

namespace {

// LDS records staged per refill: C pulses x TY rows = SPLIT_SCENE_RECS, rows C + 1 records apart
// (SPLIT_SCENE_RECS + TY records, about 6.2 KiB per wavefront).
constexpr int SPLIT_SCENE_RECS = 128;

// Five wavefronts per SIMD (at most 96 VGPRs): the loops fit without spills (measured on the
// 4000 x 250 x 40001 workload: 4.28 ms against 4.35 ms at six waves, where the loops spill a few
// values, and 4.77 ms at seven).
template <int ACCUM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_bp_split_tile(BpGeom g, PixelsView px, TileGrid tg,
                                                      double2 *__restrict__ out, int32_t *__restrict__ ncnt) {
  const int C = SPLIT_SCENE_RECS / tg.ty;  // LDS: SPLIT_SCENE_RECS records (dynamic)
  const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
  if (tile >= tg.ntiles()) return;  // whole wavefront
  const TileLane t = tile_lane(g, px, tg, tile, threadIdx.x);
  double ar = 0.0, ai = 0.0;
  int32_t cnt = 0;
  if (t.wlo < t.whi) {
    const SplitTile s = split_tile(g, tg, t, px.r0min);
    if (!s.use) {  // not a compact tile: FP64 terms
      f64_span(g, g.sctab, t, t.wlo, t.whi, ar, ai, cnt);  // table read from global memory (rare path)
    } else {
      float sr = 0.f, si = 0.f, cr = 0.f, ci = 0.f;
      split_span<ACCUM>(g, t, s, t.wlo, t.whi, 0, C, sr, si, cr, ci, ar, ai, cnt);
      split_total<ACCUM>(sr, si, cr, ci, ar, ai);
    }
  }
  if (t.valid) store_mean(out, ncnt, t.i, ar, ai, cnt);
}

}  // namespace

int sarbp_launch_split(const sarbp_dataset *ds, const sarbp_scene *sc, const Tiling &t, hipStream_t st) {
  if (sc->n == 0) return SARBP_OK;
  const TileGrid tg = make_tile_grid(sc->dimx, sc->dimy, t.tile_x);
  const int64_t blocks = (tg.ntiles() + SARBP_NUM_XCD - 1) / SARBP_NUM_XCD * SARBP_NUM_XCD;
  if (blocks > 0x7fffffff) return sarbp_fail(SARBP_EINVAL, "too many tiles for one launch");
  if (t.accum == 0)
    hipLaunchKernelGGL(k_bp_split_tile<0>, dim3((unsigned)blocks), dim3(64), (SPLIT_SCENE_RECS + tg.ty) * sizeof(PulseRec), st, sc->geom(), sc->view(), tg,
                       sc->out, sc->ncnt);
  else
    hipLaunchKernelGGL(k_bp_split_tile<1>, dim3((unsigned)blocks), dim3(64), (SPLIT_SCENE_RECS + tg.ty) * sizeof(PulseRec), st, sc->geom(), sc->view(), tg,
                       sc->out, sc->ncnt);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return sarbp_hip_fail(e, "split kernel launch");
  return SARBP_OK;
}
