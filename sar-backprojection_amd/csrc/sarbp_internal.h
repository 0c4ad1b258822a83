// Internal declarations shared by the HIP translation units of libsarbp.so.
#pragma once

#include <cmath>

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "sarbp.h"

// Constants of process_point (sar_focuser_puntual.py:35-40).  PI is float32(pi) held in a double,
// as the reference does (it copies IDL's single-precision !PI); C is the reference's rounded value.
#define SARBP_C_F64 2.9979246e8
#define SARBP_ALFA_F64 3.335987619777693e11
#define SARBP_PI32_F64 3.141592741012573242

// Geometry + RawData of one dataset as the kernels see it.
struct BpGeom {
  const float2 *raw;    // column-major (nra, naz): sample (r, a) at raw[r + nra*a]
  const double4 *traj;  // per pulse (tx, ty, tz, 0)
  int64_t nra;
  int64_t naz;
  double near_range;
  double dra;
  double inv_dra;       // 1/dra, used only where the result is re-checked exactly
  double kphase;        // (-4*PI32)/lambda        (sar_focuser_puntual.py:116)
  double beta;          // ((4*PI32)*ALFA)/(C*C)   (sar_focuser_puntual.py:40)
  const double2 *sctab; // SARBP_SCTAB entries (sin, cos)(j*pi/128), j = 0..255, for sincos_tab
  int32_t bin_lo;       // lowest range bin that contributes (process_point's ind_line_ra_min, clamped
                        // to [0, nra-1]; 0 unless a scene sets it)
};

// Entries of the sin/cos table of the FP64 path (one turn in steps of pi/128).
#define SARBP_SCTAB 256

// Pixel set in device memory (F-order flattened DEM grid).
struct PixelsView {
  const double *x;
  const double *y;
  const double *z;
  const uint8_t *mask;
  const int32_t *index;
  const int32_t *sint;
  int64_t n;
  float r0min;  // lower bound of every pixel-to-sensor distance (m), for the split form's series bound
};

struct sarbp_dataset {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::mutex timing_mu;  // the timed launches of the dataset's scenes and batches share ev0/ev1
  int64_t nra = 0, naz = 0;
  double near_range = 0, dra = 0, lambda_val = 0;
  float2 *raw = nullptr;
  double4 *traj = nullptr;
  double2 *sctab = nullptr;           // SARBP_SCTAB (sin, cos) pairs
  std::vector<double> htx, hty, htz;  // host copy of the trajectory (request planning)
  double tlo[3] = {0, 0, 0}, thi[3] = {0, 0, 0};  // bounding box of the trajectory
  int64_t bytes = 0;
  // requests submitted to a server against this dataset and not yet done or failed (preempted ones
  // included: they may be resumed); sarbp_dataset_destroy refuses while any is left
  std::atomic<int64_t> server_refs{0};
  // scenes and batches created on this dataset and not yet destroyed; sarbp_dataset_destroy refuses
  // while any is left (they use its stream and RawData)
  std::atomic<int64_t> child_refs{0};
  BpGeom geom() const;
};

// How a focusing pass maps pixels to wavefronts.  A tile is tile_x * (64/tile_x) pixels of the
// F-order grid (dimx fast); one 64-lane wavefront walks the union of its pixels' apertures pulse by
// pulse.  kernel 0 = one pixel per lane with per-lane apertures (no tiling).
struct Tiling {
  int kernel = 1;       // 0: per-lane apertures, 1: pulse-synchronous tiles
  int tile_x = 16;      // 64, 32, 16 or 8
  int chunk = 64;       // pulses of per-pulse reference data staged per LDS refill (split path)
  int accum = 0;        // split path: 0 = Kahan FP32, 1 = FP32 blocks folded into FP64
};

struct sarbp_scene {
  sarbp_dataset *ds = nullptr;
  int64_t n = 0;
  int64_t dimx = 0, dimy = 0;   // grid shape (F order); dimx = n, dimy = 1 for a plain list
  Tiling tiling_f64, tiling_split;
  double *x = nullptr, *y = nullptr, *z = nullptr;
  uint8_t *mask = nullptr;
  int32_t *index = nullptr, *sint = nullptr;
  double2 *out = nullptr;
  int32_t *ncnt = nullptr;
  int64_t bytes = 0;
  int32_t bin_lo = 0;  // lowest contributing range bin (sarbp_scene_set_bin_min)
  bool counted = false;  // counted in ds->child_refs
  float r0min = 0.f;   // lower bound of the pixel-to-sensor distances (sarbp_pixel_r0min)
  bool f64_narrow = false;  // every pair's R^2 and phase are within sarbp_narrow_ranges' bounds
  PixelsView view() const { return PixelsView{x, y, z, mask, index, sint, n, r0min}; }
  BpGeom geom() const;  // the dataset's geometry with this scene's range-bin floor
};

// Samples the device trajectory holds past pulse naz-1 (copies of the last one), so that kernels
// may read a few pulses ahead of a loop's end without clamping the index.
constexpr int SARBP_TRAJ_PAD = 8;

// Aperture of one pixel (sar_focuser_puntual.py:57-66): J = 2*ceil(s/2) + 1 pulses starting at
// index - ceil(|s|/2); integer forms of the reference's float ceil on an int32.
__host__ __device__ inline int64_t sarbp_ceil_half(int64_t s) { return s >= 0 ? (s + 1) / 2 : s / 2; }
__host__ __device__ inline void sarbp_aperture(int32_t index, int32_t sint, int64_t &a0, int64_t &J) {
  int64_t s = sint;
  int64_t h = sarbp_ceil_half(s < 0 ? -s : s);
  J = sarbp_ceil_half(s) * 2 + 1;
  a0 = (int64_t)index - h;
}
// Pulses actually used: a in [lo, hi) = [a0, a0+J) intersected with (0, naz-1).
__host__ __device__ inline void sarbp_pulse_range(int32_t index, int32_t sint, int64_t naz,
                                                  int64_t &lo, int64_t &hi) {
  int64_t a0, J;
  sarbp_aperture(index, sint, a0, J);
  lo = a0 > 1 ? a0 : 1;
  int64_t e = a0 + J;
  hi = e < naz - 1 ? e : naz - 1;
  if (J <= 0 || hi < lo) hi = lo;
}

Tiling default_tiling(int precision);

// Split form limits (bp_tile.h): the largest pixel-to-row-reference offset it focuses, and the offset
// beyond which its range series includes the fourth-order term.
constexpr float SPLIT_OMAX_M = 48.0f;
constexpr float SPLIT_FOURTH_M = 12.0f;
// Bound of the range series' truncation (m) for offsets up to omax from the reference point, at
// reference ranges of at least r0min: omax^4/(6.4 r0min^3) for the third-order form,
// omax^5/(8 r0min^4) with the fourth-order term (the suprema over offset directions are
// omax^4/(8 R0^3) and 0.083 omax^5/R0^4 while omax <= R0/6; checked in FP64 against exact distances
// on the CPU).  Host and device.
__host__ __device__ inline double split_truncation_m(double omax, double r0min) {
  if (!(r0min > 6.0 * omax)) return 1e30;  // the series needs |o| well inside R0
  const double q = omax / r0min, o4 = q * q * q * q * r0min;
  return omax > (double)SPLIT_FOURTH_M ? o4 * q / 8.0 : o4 / 6.4;  // margins 1.5x and 1.25x
}
// Phase (rad) that the truncation may reach before a tile takes the FP64 terms instead.
constexpr double SPLIT_TRUNC_MAX_RAD = 1e-4;

// Bounding boxes of a pixel set (host): of every pixel, and of the pixels that focus (mask == 1,
// sint != -9999), built over index ranges (in any order and split) and merged.
struct PixelBoxes {
  double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};     // every pixel
  double llo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, lhi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};  // pixels that focus
  bool finite = true;       // every coordinate is finite
  bool live_finite = true;  // every coordinate of the pixels that focus is finite
  bool any_live = false;
  // pixels [i0, i1); mask and sint may be null (then only the box of every pixel is kept)
  void add(const double *x, const double *y, const double *z, const uint8_t *mask, const int32_t *sint, int64_t i0,
           int64_t i1);
  void merge(const PixelBoxes &o);
};
// Lower bound of the distance from any pixel of the boxes to any trajectory sample of ds: the
// distance between the two bounding boxes (0 when they overlap, or when a coordinate is not finite),
// rounded down to float.
float sarbp_r0min_of(const sarbp_dataset *ds, const PixelBoxes &b);
// The same for n pixels.
float sarbp_pixel_r0min(const sarbp_dataset *ds, const double *x, const double *y, const double *z, int64_t n);
// sarbp_live_range_bounds from the boxes.
bool sarbp_range_bounds_of(const sarbp_dataset *ds, const PixelBoxes &b, double &rmin, double &rmax);
// Bounds [rmin, rmax] of the distances between the pixels that focus (mask == 1, sint != -9999)
// and every trajectory sample, from the two bounding boxes (rmin rounded down, rmax up); false
// when such a pixel has a non-finite coordinate.  No such pixel: true, rmin = rmax = 0.
bool sarbp_live_range_bounds(const sarbp_dataset *ds, const double *x, const double *y, const double *z,
                             const uint8_t *mask, const int32_t *sint, int64_t n, double &rmin, double &rmax);
// True when every pixel-pulse pair with distance in [r0min, rmax] has R^2 inside sqrt_rn's direct
// range [2^-900, 2^900] and a phase |k R + beta R^2| below 2^24 rad, so that the FP64 kernels may
// drop their wide-argument paths (library sqrt and sin/cos), which such pairs never take.
bool sarbp_narrow_ranges(const sarbp_dataset *ds, double r0min, double rmax);



// error plumbing (api.hip)
int sarbp_fail(int code, const std::string &msg);
int sarbp_hip_fail(hipError_t e, const char *what);

// kernel launchers (bp_f64.hip, bp_split.hip)
int sarbp_launch_f64(const sarbp_dataset *ds, const sarbp_scene *sc, const Tiling &t, hipStream_t st);
int sarbp_launch_split(const sarbp_dataset *ds, const sarbp_scene *sc, const Tiling &t, hipStream_t st);
int sarbp_launch_eval_pairs(const sarbp_dataset *ds, const double *x, const double *y, const double *z,
                            const int32_t *pulse, int64_t n, double *r2, double *r, int32_t *bin, double *ph,
                            double *sn, double *cs, hipStream_t st);

// Grid tile -> lanes.  Tiles are numbered x-fastest inside a band of tile rows; blocks are remapped
// so that each XCD (blocks are dealt round-robin to the 8 XCDs) works on a contiguous run of tiles
// and its L2 sees neighbouring apertures.
struct TileGrid {
  int64_t dimx, dimy;   // pixel grid
  int64_t ntx, nty;     // tiles along x and y
  int tx, ty;           // tile shape, tx * ty == 64
  __host__ __device__ int64_t ntiles() const { return ntx * nty; }
};
inline TileGrid make_tile_grid(int64_t dimx, int64_t dimy, int tile_x) {
  // widen the tile along x when the grid has fewer rows than the tile
  while (tile_x < 64 && 64 / tile_x > dimy) tile_x *= 2;
  TileGrid g;
  g.dimx = dimx;
  g.dimy = dimy;
  g.tx = tile_x;
  g.ty = 64 / tile_x;
  g.ntx = (dimx + g.tx - 1) / g.tx;
  g.nty = (dimy + g.ty - 1) / g.ty;
  return g;
}
constexpr int SARBP_NUM_XCD = 8;
__device__ inline int64_t xcd_tile(int64_t block, int64_t nblocks) {
  // block b runs on XCD b % 8; give XCD k the tiles [k*per, (k+1)*per)
  const int64_t per = (nblocks + SARBP_NUM_XCD - 1) / SARBP_NUM_XCD;
  const int64_t xcd = block % SARBP_NUM_XCD, slot = block / SARBP_NUM_XCD;
  return xcd * per + slot;
}

// wavefront reductions (64 lanes)
__device__ inline int wave_min_i32(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = min(v, __shfl_xor(v, m, 64));
  return v;
}
__device__ inline int wave_max_i32(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = max(v, __shfl_xor(v, m, 64));
  return v;
}
__device__ inline double wave_min_f64(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmin(v, __shfl_xor(v, m, 64));
  return v;
}
__device__ inline double wave_max_f64(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
  return v;
}

// ---- batched / served focusing: many pixel sets (requests) in one persistent launch ----------

// Partial focusing state of one pixel between launches (32 B): FP64 sums (the Kahan running sum
// of the split path is held exactly in ar/ai), Kahan compensations, contributor count.
struct BpState {
  double ar, ai;
  float cr, ci;
  int32_t cnt, pad;
};

// One request as the batch kernel sees it (device pointers).  Each request carries its dataset's
// geometry, so one launch can mix requests against different device-resident datasets.
struct BatchReqDev {
  BpGeom g;
  PixelsView px;
  double2 *out;
  int32_t *ncnt;
  BpState *state;    // partial sums between launches (serving), or the pulse segments' partial
                     // sums (slot k of pixel i at state[k * n + i]); may be null for one-shot batches
  TileGrid tg;
  int32_t precision;  // SARBP_PREC_*
  int32_t accum;      // split path accumulation (0 Kahan, 1 FP64 fold)
  int32_t nseg = 1;   // pulse segments per tile (state slots per pixel when > 1)
  int32_t f64_narrow = 0;  // FP64 terms may take the narrow-range form (sarbp_narrow_ranges)
};

// Work item: tiles [tile_lo, tile_lo + W) of one request (one per wavefront of the block), pulses
// [wlo + p_off, wlo + p_end) of each tile's union (p_end = INT32_MAX: to the end).  p_off and a
// finite p_end are multiples of the pulse chunk C.  slot < 0: the item continues the pixels' sums
// from the saved state (p_off > 0) or from zero, and writes the means when it reaches the end of
// the union.  slot >= 0: a pulse segment; its sums start from zero and go to partial slot `slot`
// (combined afterwards by sarbp_launch_combine).
struct BatchItem {
  int32_t req;
  int32_t tile_lo;
  int32_t p_off;
  int32_t p_end;
  int32_t slot;
};

constexpr int SARBP_BATCH_QUEUES = 8;  // one work queue per XCD, with stealing
// Queue heads sit 256 bytes apart (one per cache line and memory channel): every block of the
// persistent grid visits every queue, and heads sharing one line would serialise all those
// atomics in one place.
constexpr int SARBP_HEAD_STRIDE = 64;   // int32 elements between queue heads
constexpr size_t SARBP_HEADS_BYTES = (size_t)SARBP_BATCH_QUEUES * SARBP_HEAD_STRIDE * sizeof(int32_t);

// Which requests one batch launch holds: any mix, FP64 requests only, or split requests only (the
// last two run instantiations with fewer registers and more wavefronts per SIMD; same results).
constexpr int SARBP_BATCH_ANY = 0, SARBP_BATCH_F64 = 1, SARBP_BATCH_SPLIT = 2;
// The launch mode for a set of request precisions.
inline int sarbp_batch_mode(bool any_f64, bool any_split) {
  return any_f64 && any_split ? SARBP_BATCH_ANY : any_split ? SARBP_BATCH_SPLIT : SARBP_BATCH_F64;
}

// Launch the persistent batch kernel: `waves` wavefronts per block (1, 2 or 4), pulse chunk C,
// tiles of at most max_rows rows (LDS holds (C + 1) * max_rows pulse records per wavefront).
// d_heads: SARBP_HEADS_BYTES, zeroed by the launcher; d_seg: SARBP_BATCH_QUEUES + 1 item offsets;
// nitems: the number of items; mode: SARBP_BATCH_* covering the requests' precisions.  grid_blocks <= 0 sizes the grid from the occupancy, at most one
// block per item.
int sarbp_launch_batch(int device, const BatchReqDev *d_reqs, const BatchItem *d_items, const int32_t *d_seg,
                       int32_t *d_heads, int nitems, int mode, int waves, int chunk, int max_rows, int grid_blocks,
                       hipStream_t st);
size_t sarbp_batch_lds_bytes(int waves, int chunk, int max_rows);
constexpr size_t SARBP_BATCH_LDS_MAX = 64 * 1024;  // LDS of one batch block's pulse records
// Blocks of `waves` wavefronts of the batch kernel for `mode` that the device holds at once.
int sarbp_batch_resident_blocks(int device, int mode, int waves, int chunk, int max_rows, int *blocks);
// Pixel means from the partial sums of nseg pulse segments (part[k * n + i], added in the order
// k = 0 .. nseg-1): out[i], ncnt[i] for i < n.
int sarbp_launch_combine(const BpState *part, int nseg, int64_t n, double2 *out, int32_t *ncnt, hipStream_t st);
// The same for several segmented requests of one launch at once: reqs[list[k]] for k < count, each
// with nseg slots per pixel.  sarbp_launch_segments_clear zeroes their slots beforehand (a pixel
// contributes no slot beyond its own window, and a tile no item beyond its union).
int sarbp_launch_segments_clear(const BatchReqDev *d_reqs, const int32_t *d_list, int count, hipStream_t st);
int sarbp_launch_segments_combine(const BatchReqDev *d_reqs, const int32_t *d_list, int count, hipStream_t st);
