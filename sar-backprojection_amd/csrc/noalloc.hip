// Allocation-free batched focusing: sarbp_batch_focus_noalloc (include/sarbp.h).
//
// The same persistent launch as sarbp_batch_run (bp_batch.hip), for callers that cannot allocate
// device memory: the requests' pixel arrays and outputs are the caller's device buffers, and the
// launch's own device memory -- the request descriptors, the work list, the per-queue item offsets
// and the queue heads -- lives in `scratch`, a fixed array in the library's device image (one copy
// per GPU, held until the process ends).  The host stages the work list in a fixed host array, so
// no call allocates host or device memory.  The pixels' distance bound that the split form needs
// (PixelsView::r0min, computed on the host from host arrays elsewhere) is computed here on the device,
// with the same arithmetic, so results equal sarbp_batch_run's bit for bit.
#include <algorithm>
#include <climits>
#include <cstring>
#include <mutex>

#include "batch_internal.h"
#include "bp_common.h"

#pragma clang fp contract(off)

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)4 << 20;  // SARBP_NOALLOC_SCRATCH_BYTES in sarbp.h
static_assert(NOALLOC_SCRATCH_BYTES == (size_t)SARBP_NOALLOC_SCRATCH_BYTES, "scratch size mismatch");

// Work list of the launches of sarbp_batch_focus_noalloc on this GPU.  Host code never uses this name
// as an address: it uses the pointer hipGetSymbolAddress returned for the current device.
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;                      // one call at a time per process: it owns scratch and g_stage
char *g_scratch[MAX_DEVICES] = {};    // device address of scratch, per device (set once)
alignas(256) char g_stage[NOALLOC_SCRATCH_BYTES];  // host image of the uploaded part of scratch

// Layout of scratch: queue heads, queue offsets, request descriptors, work items.
constexpr size_t OFF_HEADS = 0;
constexpr size_t OFF_SEG = OFF_HEADS + SARBP_HEADS_BYTES;
constexpr size_t OFF_REQS = OFF_SEG + 256;
static_assert((SARBP_BATCH_QUEUES + 1) * sizeof(int32_t) <= 256, "queue offsets exceed their slot");

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// Lower bound of the pixel-to-sensor distances of each split request (one block per request): the
// distance between the bounding box of its pixels and the trajectory's, rounded down to float --
// sarbp_pixel_r0min's arithmetic (min/max are exact, the sum of squares runs in the same order,
// the square root is correctly rounded), written into the request's descriptor.
__global__ __launch_bounds__(1024) void k_pixel_r0min(BatchReqDev *reqs, const int32_t *which, double tlo0,
                                                      double tlo1, double tlo2, double thi0, double thi1,
                                                      double thi2) {
  __shared__ double part[16][6];
  BatchReqDev &r = reqs[which[blockIdx.x]];
  const PixelsView px = r.px;
  const double inf = __builtin_huge_val();
  double v[6] = {inf, inf, inf, -inf, -inf, -inf};  // lo x y z, hi x y z
  for (int64_t i = threadIdx.x; i < px.n; i += blockDim.x) {
    const double p[3] = {px.x[i], px.y[i], px.z[i]};
    for (int k = 0; k < 3; ++k) {
      v[k] = fmin(v[k], p[k]);
      v[k + 3] = fmax(v[k + 3], p[k]);
    }
  }
  for (int k = 0; k < 3; ++k) {
    v[k] = wave_min_f64(v[k]);
    v[k + 3] = wave_max_f64(v[k + 3]);
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0)
    for (int k = 0; k < 6; ++k) part[wave][k] = v[k];
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
    for (int k = 0; k < 3; ++k) {
      v[k] = fmin(v[k], part[w][k]);
      v[k + 3] = fmax(v[k + 3], part[w][k + 3]);
    }
  const double tlo[3] = {tlo0, tlo1, tlo2}, thi[3] = {thi0, thi1, thi2};
  double d2 = 0.0;
  for (int k = 0; k < 3; ++k) {
    const double gap = fmax(fmax(0.0, v[k] - thi[k]), tlo[k] - v[k + 3]);
    d2 += gap * gap;
  }
  const double d = sqrt_rn(d2);
  r.px.r0min = __builtin_isfinite(d) ? nextafterf((float)d, 0.f) : 0.f;
}

int scratch_for_device(int dev, char **out) {
  if (dev < 0 || dev >= MAX_DEVICES) return sarbp_fail(SARBP_EINVAL, "device index beyond the supported count");
  if (!g_scratch[dev]) {  // once per device, with that device current
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return sarbp_hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch address");
    g_scratch[dev] = (char *)p;
  }
  *out = g_scratch[dev];
  return SARBP_OK;
}

int focus_locked(sarbp_dataset *ds, const sarbp_request *reqs, int32_t nreq, double *const *out_c128,
                 int32_t *const *ncontrib, int waves, int ch, int max_rows, bool any_f64, bool any_split,
                 size_t nitems, size_t nsplit, size_t off_items, size_t off_which, size_t total);

}  // namespace

extern "C" int sarbp_batch_focus_noalloc(sarbp_dataset *ds, const sarbp_request *reqs, int32_t nreq,
                                         double *const *out_c128, int32_t *const *ncontrib, int32_t tile_px,
                                         int32_t pulse_chunk) {
  if (!ds) return sarbp_fail(SARBP_EINVAL, "dataset is null");
  if (nreq < 1 || !reqs || !out_c128 || !ncontrib) return sarbp_fail(SARBP_EINVAL, "bad request list");
  const int tp = tile_px ? tile_px : 64, ch = pulse_chunk ? pulse_chunk : 32;
  const int waves = tp / 64;
  int max_rows = 1;
  bool any_f64 = false, any_split = false;
  size_t nitems = 0, nsplit = 0;
  for (int32_t k = 0; k < nreq; ++k) {
    const sarbp_request &r = reqs[k];
    if (r.n < 0) return sarbp_fail(SARBP_EINVAL, "negative pixel count");
    if (r.n > 0 && (!r.x || !r.y || !r.z || !r.mask || !r.index || !r.sint || !out_c128[k] || !ncontrib[k]))
      return sarbp_fail(SARBP_EINVAL, "request pixel or output pointer is null");
    if (r.dimx < 0 || (r.n > 0 && r.dimx > 0 && r.n % r.dimx))
      return sarbp_fail(SARBP_EINVAL, "request pixel count must be a multiple of dimx");
    if (r.precision != SARBP_PREC_F64 && r.precision != SARBP_PREC_SPLIT)
      return sarbp_fail(SARBP_EINVAL, "unknown request precision");
    const int tx = r.tile_x ? r.tile_x : default_tiling(r.precision).tile_x;
    if (tx != 64 && tx != 32 && tx != 16 && tx != 8 && tx != 4) return sarbp_fail(SARBP_EINVAL, "bad tile_x");
    const int64_t dimx = r.dimx > 0 ? r.dimx : (r.n > 0 ? r.n : 1);
    const TileGrid tg = make_tile_grid(dimx, r.n > 0 ? r.n / dimx : 0, tx);
    (r.precision == SARBP_PREC_SPLIT ? any_split : any_f64) = true;
    if (r.precision == SARBP_PREC_SPLIT) {
      max_rows = std::max(max_rows, tg.ty);
      nsplit += r.n > 0;
    }
    nitems += (size_t)((tg.ntiles() + waves - 1) / waves);
  }
  if (!sarbp_batch_mapping_valid(tp, ch, max_rows))
    return sarbp_fail(SARBP_EINVAL, "invalid tile_px / pulse_chunk combination");
  const size_t off_items = OFF_REQS + align256((size_t)nreq * sizeof(BatchReqDev));
  const size_t off_which = off_items + align256(nitems * sizeof(BatchItem));
  const size_t total = off_which + nsplit * sizeof(int32_t);
  if (total > NOALLOC_SCRATCH_BYTES || nitems > (size_t)INT32_MAX)
    return sarbp_fail(SARBP_EINVAL, "the requests' work list exceeds the allocation-free scratch "
                                    "(SARBP_NOALLOC_SCRATCH_BYTES): use fewer or smaller requests");

  std::lock_guard<std::mutex> lk(g_mu);
  std::lock_guard<std::mutex> tl(ds->timing_mu);  // the dataset's stream, as its other timed launches
  int prev = 0;
  hipError_t e = hipGetDevice(&prev);
  if (e == hipSuccess && prev != ds->device) e = hipSetDevice(ds->device);
  if (e != hipSuccess) return sarbp_hip_fail(e, "set device");
  const int rc = focus_locked(ds, reqs, nreq, out_c128, ncontrib, waves, ch, max_rows, any_f64, any_split, nitems,
                              nsplit, off_items, off_which, total);
  if (prev != ds->device) (void)hipSetDevice(prev);
  return rc;
}

namespace {
int focus_locked(sarbp_dataset *ds, const sarbp_request *reqs, int32_t nreq, double *const *out_c128,
                 int32_t *const *ncontrib, int waves, int ch, int max_rows, bool any_f64, bool any_split,
                 size_t nitems, size_t nsplit, size_t off_items, size_t off_which, size_t total) {
  char *d = nullptr;
  int rc = scratch_for_device(ds->device, &d);
  if (rc) return rc;

  // host image: queue offsets, descriptors, items (dealt to the queues in runs of 8, in request and
  // tile order: sarbp_batch_run's dealing with equal costs), split requests' indices
  int32_t *seg = (int32_t *)(g_stage + OFF_SEG);
  BatchReqDev *hr = (BatchReqDev *)(g_stage + OFF_REQS);
  BatchItem *hi = (BatchItem *)(g_stage + off_items);
  int32_t *which = (int32_t *)(g_stage + off_which);
  constexpr size_t K = 8;
  size_t cnt[SARBP_BATCH_QUEUES] = {}, pos[SARBP_BATCH_QUEUES];
  for (size_t k = 0; k < nitems; ++k) cnt[(k / K) % SARBP_BATCH_QUEUES]++;
  size_t o = 0;
  for (int q = 0; q < SARBP_BATCH_QUEUES; ++q) {
    seg[q] = (int32_t)o;
    pos[q] = o;
    o += cnt[q];
  }
  seg[SARBP_BATCH_QUEUES] = (int32_t)o;
  const BpGeom g = ds->geom();
  size_t item = 0, ns = 0;
  for (int32_t k = 0; k < nreq; ++k) {
    const sarbp_request &r = reqs[k];
    const int tx = r.tile_x ? r.tile_x : default_tiling(r.precision).tile_x;
    const int64_t dimx = r.dimx > 0 ? r.dimx : (r.n > 0 ? r.n : 1);
    BatchReqDev q;
    std::memset(&q, 0, sizeof(q));
    q.g = g;
    q.px = PixelsView{r.x, r.y, r.z, r.mask, r.index, r.sint, r.n, 0.f};
    q.out = (double2 *)out_c128[k];
    q.ncnt = ncontrib[k];
    q.state = nullptr;
    q.tg = make_tile_grid(dimx, r.n > 0 ? r.n / dimx : 0, tx);
    q.precision = r.precision;
    q.accum = default_tiling(r.precision).accum;
    hr[k] = q;
    if (r.precision == SARBP_PREC_SPLIT && r.n > 0) which[ns++] = k;
    for (int64_t t = 0; t < q.tg.ntiles(); t += waves, ++item)
      hi[pos[(item / K) % SARBP_BATCH_QUEUES]++] = BatchItem{k, (int32_t)t, 0, INT32_MAX, -1};
  }
  (void)nsplit;
  hipStream_t st = ds->stream;
  hipError_t e = hipMemcpyAsync(d + OFF_SEG, g_stage + OFF_SEG, total - OFF_SEG, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return sarbp_hip_fail(e, "work-list upload");
  if (ns > 0) {
    hipLaunchKernelGGL(k_pixel_r0min, dim3((unsigned)ns), dim3(1024), 0, st, (BatchReqDev *)(d + OFF_REQS),
                       (const int32_t *)(d + off_which), ds->tlo[0], ds->tlo[1], ds->tlo[2], ds->thi[0], ds->thi[1],
                       ds->thi[2]);
    e = hipGetLastError();
    if (e != hipSuccess) return sarbp_hip_fail(e, "distance-bound kernel launch");
  }
  rc = sarbp_launch_batch(ds->device, (const BatchReqDev *)(d + OFF_REQS), (const BatchItem *)(d + off_items),
                          (const int32_t *)(d + OFF_SEG), (int32_t *)(d + OFF_HEADS), (int)nitems,
                          sarbp_batch_mode(any_f64, any_split), waves, ch, max_rows, 0, st);
  e = hipStreamSynchronize(st);  // the staging and scratch are reused by the next call
  if (rc) return rc;
  if (e != hipSuccess) return sarbp_hip_fail(e, "batch kernel");
  return SARBP_OK;
}
}  // namespace
