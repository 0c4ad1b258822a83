// Host side of batched focusing: device buffers of requests, work-item lists for the persistent
// kernel (bp_batch.hip), and the tile/pulse-chunk autotuner.
#include <algorithm>
#include <climits>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <vector>

#include "batch_internal.h"
#include "serve_policy.h"

namespace {
constexpr size_t ALIGN = 256;
size_t up(size_t v) { return (v + ALIGN - 1) / ALIGN * ALIGN; }

std::mutex g_tune_mu;
// (nra, naz, pixels, requests, rows of the widest split tile) -> (tile_px, chunk)
std::map<std::tuple<int64_t, int64_t, int64_t, int32_t, int>, std::pair<int, int>> g_tuned;
}  // namespace

char *DevicePool::get(size_t bytes, size_t *cap) {
  const size_t want = (bytes + ((size_t)1 << 20) - 1) >> 20 << 20;
  auto it = blocks.lower_bound(want);
  if (it != blocks.end() && it->first <= 2 * want) {
    char *p = it->second;
    *cap = it->first;
    cached -= it->first;
    blocks.erase(it);
    return p;
  }
  char *p = nullptr;
  hipError_t e = hipMalloc((void **)&p, want);
  if (e != hipSuccess && !blocks.empty()) {  // give the cached blocks back and retry once
    (void)hipGetLastError();
    release();
    e = hipMalloc((void **)&p, want);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  *cap = want;
  return p;
}

void DevicePool::put(char *p, size_t cap) {
  if (!p) return;
  if (cached + cap > limit) {
    (void)hipFree(p);
    return;
  }
  blocks.emplace(cap, p);
  cached += cap;
}

void DevicePool::release() {
  for (auto &b : blocks) (void)hipFree(b.second);
  blocks.clear();
  cached = 0;
}

void request_tiles_band(const sarbp_dataset *ds, const sarbp_request &r, const TileGrid &tg, int64_t by0, int64_t by1,
                        int *tile_len, int64_t &pp, SplitRows *rows) {
  const int64_t dimx = tg.dimx, dimy = tg.dimy;
  const double *htx = ds->htx.data(), *hty = ds->hty.data(), *htz = ds->htz.data();
  for (int64_t by = by0; by < by1; ++by) {
    const int64_t y0 = by * tg.ty, y1 = std::min(dimy, y0 + tg.ty);
    for (int64_t bx = 0; bx < tg.ntx; ++bx) {
      const int64_t x0 = bx * tg.tx, x1 = std::min(dimx, x0 + tg.tx);
      int tlo = INT_MAX, thi = INT_MIN;
      for (int64_t iy = y0; iy < y1; ++iy) {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        int64_t ends[2][3] = {{-1, 0, 0}, {-1, 0, 0}};  // first and last pixel of the row: i, a0, a1
        for (int64_t ix = x0; ix < x1; ++ix) {
          const int64_t i = ix + dimx * iy;
          if (r.mask[i] != 1 || r.sint[i] == -9999) continue;
          int64_t a0, a1;
          sarbp_pulse_range(r.index[i], r.sint[i], ds->naz, a0, a1);
          if (a1 <= a0) continue;
          pp += a1 - a0;
          tlo = std::min(tlo, (int)a0);
          thi = std::max(thi, (int)a1);
          if (!rows) continue;
          if (ends[0][0] < 0) ends[0][0] = i, ends[0][1] = a0, ends[0][2] = a1;
          ends[1][0] = i, ends[1][1] = a0, ends[1][2] = a1;
          const double p[3] = {r.x[i], r.y[i], r.z[i]};
          for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
          }
        }
        if (!rows || ends[0][0] < 0) continue;
        // the along-track direction cosine varies smoothly along a row: taken at the aperture ends
        // of the row's first and last pixels
        double c2max = 0.0;
        for (const auto &e : ends)
          for (int64_t a : {e[1], e[2] - 1}) {
            const double dx = r.x[e[0]] - htx[a], dy = r.y[e[0]] - hty[a], dz = r.z[e[0]] - htz[a];
            c2max = std::max(c2max, dx * dx / (dx * dx + dy * dy + dz * dz));
          }
        const double ox = 0.5 * (hi[0] - lo[0]), oy = 0.5 * (hi[1] - lo[1]), oz = 0.5 * (hi[2] - lo[2]);
        rows->any = true;
        rows->amax = std::max(rows->amax, ox * std::sqrt(c2max) + oy + oz);
        rows->omax = std::max(rows->omax, std::sqrt(ox * ox + oy * oy + oz * oz));
      }
      tile_len[(size_t)(by * tg.ntx + bx)] = thi > tlo ? thi - tlo : 0;
    }
  }
}

TileGrid request_tile_grid(const sarbp_request &r) {
  const int tx = r.tile_x ? r.tile_x : default_tiling(r.precision).tile_x;
  const int64_t dimx = r.dimx > 0 ? r.dimx : (r.n > 0 ? r.n : 1);
  const int64_t dimy = r.n > 0 ? r.n / dimx : 0;
  return make_tile_grid(dimx, dimy, tx);
}

void request_tiles(const sarbp_dataset *ds, const sarbp_request &r, TileGrid &tg, std::vector<int> &tile_len,
                   int64_t &pp, SplitRows *rows) {
  tg = request_tile_grid(r);
  tile_len.assign((size_t)tg.ntiles(), 0);
  pp = 0;
  if (rows) *rows = SplitRows();
  if (r.n <= 0) return;
  if (rows) rows->r0min = sarbp_pixel_r0min(ds, r.x, r.y, r.z, r.n);
  request_tiles_band(ds, r, tg, 0, tg.nty, tile_len.data(), pp, rows);
}

void reqbuf_layout(int64_t n64, int state_slots, size_t off[10]) {
  const size_t n = (size_t)n64;
  const size_t sz[9] = {n * 8, n * 8, n * 8, n, n * 4, n * 4, n * 16, n * 4, n * sizeof(BpState) * (size_t)state_slots};
  size_t o = 0;
  for (int k = 0; k < 9; ++k) {
    off[k] = o;
    o += up(sz[k]);
  }
  off[9] = o;
}

int reqbuf_create(sarbp_dataset *ds, const sarbp_request &r, int state_slots, int accum, ReqBuf &b,
                  hipStream_t st, DevicePool *pool, const std::vector<int> *tile_len, float r0min,
                  const void *packed, char *host_out, int narrow) {
  b = ReqBuf();
  if (r.n < 0) return sarbp_fail(SARBP_EINVAL, "negative pixel count");
  if (r.n > 0 && (!r.x || !r.y || !r.z || !r.mask || !r.index || !r.sint))
    return sarbp_fail(SARBP_EINVAL, "request pixel pointer is null");
  if (r.dimx < 0 || (r.n > 0 && r.dimx > 0 && r.n % r.dimx))
    return sarbp_fail(SARBP_EINVAL, "request pixel count must be a multiple of dimx");
  if (r.precision != SARBP_PREC_F64 && r.precision != SARBP_PREC_SPLIT)
    return sarbp_fail(SARBP_EINVAL, "unknown request precision");
  const int tx = r.tile_x ? r.tile_x : default_tiling(r.precision).tile_x;
  if (tx != 64 && tx != 32 && tx != 16 && tx != 8 && tx != 4) return sarbp_fail(SARBP_EINVAL, "bad tile_x");
  b.n = r.n;
  b.dimx = r.dimx > 0 ? r.dimx : (r.n > 0 ? r.n : 1);
  b.dimy = r.n > 0 ? r.n / b.dimx : 0;
  b.precision = r.precision;
  b.accum = accum;
  b.tg = make_tile_grid(b.dimx, b.dimy, tx);
  const size_t n = (size_t)r.n;
  size_t off[10];
  reqbuf_layout(r.n, state_slots, off);
  const size_t sz[6] = {n * 8, n * 8, n * 8, n, n * 4, n * 4};
  const size_t o = off[9];
  b.bytes = (int64_t)o;
  if (o > 0) {
    b.pool = pool;
    if (pool) {
      b.base = pool->get(o, &b.cap);
    } else {
      hipError_t e = hipMalloc((void **)&b.base, o);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        b.base = nullptr;
      }
      b.cap = o;
    }
    if (!b.base) return sarbp_fail(SARBP_ENOMEM, "request device allocation failed");
  }
  char *p = b.base;
  b.x = (double *)(p + off[0]);
  b.y = (double *)(p + off[1]);
  b.z = (double *)(p + off[2]);
  b.mask = (uint8_t *)(p + off[3]);
  b.index = (int32_t *)(p + off[4]);
  b.sint = (int32_t *)(p + off[5]);
  b.out = host_out ? (double2 *)host_out : (double2 *)(p + off[6]);
  b.ncnt = host_out ? (int32_t *)(host_out + (off[7] - off[6])) : (int32_t *)(p + off[7]);
  b.state = state_slots > 0 ? (BpState *)(p + off[8]) : nullptr;
  b.f64_narrow = narrow >= 0 ? narrow != 0 : request_narrow(ds, r);
  b.r0min = r0min >= 0.f ? r0min
                         : r.precision == SARBP_PREC_SPLIT ? sarbp_pixel_r0min(ds, r.x, r.y, r.z, r.n) : 0.f;
  const void *src[6] = {r.x, r.y, r.z, r.mask, r.index, r.sint};
  void *dst[6] = {b.x, b.y, b.z, b.mask, b.index, b.sint};
  for (int k = 0; k < 6 && n; ++k) {
    if (packed && k > 0) break;  // one copy of the packed image
    hipError_t e = packed ? hipMemcpyAsync(b.base, packed, off[6], hipMemcpyHostToDevice, st)
                          : hipMemcpyAsync(dst[k], src[k], sz[k], hipMemcpyHostToDevice, st);
    if (e != hipSuccess) {
      (void)hipStreamSynchronize(st);  // nothing in flight may still write the block
      reqbuf_free(b);
      return sarbp_hip_fail(e, "request upload");
    }
  }
  if (n && !host_out) {  // pixels of tiles that are never scheduled (no aperture) read as zero: out and ncnt at once
    hipError_t e = hipMemsetAsync(b.out, 0, off[8] - off[6], st);
    if (e != hipSuccess) {
      (void)hipStreamSynchronize(st);
      reqbuf_free(b);
      return sarbp_hip_fail(e, "request output reset");
    }
  }
  // per-tile union lengths (host, integers only): work estimate for scheduling
  if (tile_len) {
    b.tile_len = *tile_len;  // the caller keeps the request's pixel-pulse count (b.pp stays 0)
  } else {
    TileGrid tg;
    request_tiles(ds, r, tg, b.tile_len, b.pp);
  }
  return SARBP_OK;
}

void reqbuf_free(ReqBuf &b) {
  if (b.base) {
    if (b.pool)
      b.pool->put(b.base, b.cap);
    else
      (void)hipFree(b.base);
  }
  b.base = nullptr;
}

bool request_narrow(const sarbp_dataset *ds, const sarbp_request &r) {
  if (r.n <= 0) return true;
  double rmin = 0.0, rmax = 0.0;
  return sarbp_live_range_bounds(ds, r.x, r.y, r.z, r.mask, r.sint, r.n, rmin, rmax) &&
         sarbp_narrow_ranges(ds, rmin, rmax);
}

BatchReqDev reqbuf_dev(const sarbp_dataset *ds, const ReqBuf &b) {
  BatchReqDev d;
  d.g = ds->geom();
  d.px = PixelsView{b.x, b.y, b.z, b.mask, b.index, b.sint, b.n, b.r0min};
  d.out = b.out;
  d.ncnt = b.ncnt;
  d.state = b.state;
  d.tg = b.tg;
  d.precision = b.precision;
  d.accum = b.accum;
  d.f64_narrow = b.f64_narrow ? 1 : 0;
  return d;
}

void build_queues(const std::vector<WorkItemCost> &work, std::vector<BatchItem> &items, std::vector<int32_t> &seg) {
  deal_queues<SARBP_BATCH_QUEUES>(work, items, seg);
}

namespace {

static_assert(sizeof(BatchReqDev) % 4 == 0 && sizeof(BatchItem) % 4 == 0, "work lists are copied in 32-bit words");

// One piece of a work list: `words` 32-bit words from the staging block to device memory.
struct StageCopy {
  const uint32_t *src;
  uint32_t *dst;
  uint32_t words;
};

__global__ __launch_bounds__(256) void k_stage_in(StageCopy a, StageCopy b, StageCopy c, StageCopy d) {
  const StageCopy jobs[4] = {a, b, c, d};
  const uint32_t stride = gridDim.x * blockDim.x;
  for (int k = 0; k < 4; ++k)
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < jobs[k].words; i += stride) jobs[k].dst[i] = jobs[k].src[i];
}

}  // namespace

int ItemBuffers::upload(const std::vector<BatchReqDev> &reqs, const std::vector<BatchItem> &items,
                        const std::vector<int32_t> &seg, hipStream_t st, const std::vector<int32_t> *list) {
  auto grow = [](void **p, size_t &cap, size_t need) -> hipError_t {
    if (need <= cap) return hipSuccess;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    need = std::max(need, cap * 2);  // geometric growth: hipFree waits for the device
    cap = 0;
    hipError_t e = hipMalloc(p, need);
    if (e == hipSuccess) cap = need;
    return e;
  };
  auto up = [](size_t v) { return (v + 255) / 256 * 256; };
  const size_t rb = std::max<size_t>(1, reqs.size()) * sizeof(BatchReqDev);
  const size_t ib = std::max<size_t>(1, items.size()) * sizeof(BatchItem);
  const size_t sb = (SARBP_BATCH_QUEUES + 1) * sizeof(int32_t);
  const size_t nl = list ? list->size() : 0;
  const size_t lb = std::max<size_t>(1, nl) * sizeof(int32_t);
  hipError_t e = grow((void **)&d_reqs, cap_reqs, rb);
  if (e == hipSuccess) e = grow((void **)&d_items, cap_items, ib);
  if (e == hipSuccess && nl) e = grow((void **)&d_list, cap_list, lb);
  if (e == hipSuccess && !d_seg) e = hipMalloc((void **)&d_seg, sb);
  if (e == hipSuccess && !d_heads) e = hipMalloc((void **)&d_heads, SARBP_HEADS_BYTES);
  const size_t stage = up(rb) + up(ib) + up(sb) + up(lb);
  if (e == hipSuccess && stage > cap_stage) {  // page-locked, device-visible staging block
    const size_t want = std::max(stage, cap_stage * 2);
    if (h_stage) (void)hipHostFree(h_stage);
    h_stage = nullptr;
    cap_stage = 0;
    d_stage = nullptr;
    e = hipHostMalloc((void **)&h_stage, want, hipHostMallocMapped);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void **)&d_stage, h_stage, 0);
    if (e == hipSuccess) cap_stage = want;
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return sarbp_fail(SARBP_ENOMEM, "work-list allocation failed");
  }
  char *hr = h_stage, *hi = h_stage + up(rb), *hs = hi + up(ib), *hl = hs + up(sb);
  if (!reqs.empty()) std::memcpy(hr, reqs.data(), reqs.size() * sizeof(BatchReqDev));
  if (!items.empty()) std::memcpy(hi, items.data(), items.size() * sizeof(BatchItem));
  std::memcpy(hs, seg.data(), seg.size() * sizeof(int32_t));
  if (nl) std::memcpy(hl, list->data(), nl * sizeof(int32_t));
  // one kernel on the launch's stream reads the four pieces from the staging block (device-visible
  // host memory) into device memory: no DMA copy, whose engine, idle between rounds, has started
  // such a copy milliseconds late (a serving round of 0.6 ms took 8 ms, rocprofv3 trace)
  const char *ds = d_stage;
  const size_t off_i = hi - hr, off_s = hs - hr, off_l = hl - hr;
  StageCopy jobs[4] = {{(const uint32_t *)ds, (uint32_t *)d_reqs, (uint32_t)(reqs.size() * sizeof(BatchReqDev) / 4)},
                       {(const uint32_t *)(ds + off_i), (uint32_t *)d_items, (uint32_t)(items.size() * sizeof(BatchItem) / 4)},
                       {(const uint32_t *)(ds + off_s), (uint32_t *)d_seg, (uint32_t)seg.size()},
                       {(const uint32_t *)(ds + off_l), (uint32_t *)d_list, (uint32_t)nl}};
  const size_t words = std::max({(size_t)jobs[0].words, (size_t)jobs[1].words, (size_t)jobs[2].words, (size_t)jobs[3].words});
  const unsigned blocks = (unsigned)std::min<size_t>(256, std::max<size_t>(1, (words + 255) / 256));
  hipLaunchKernelGGL(k_stage_in, dim3(blocks), dim3(256), 0, st, jobs[0], jobs[1], jobs[2], jobs[3]);
  e = hipGetLastError();
  if (e != hipSuccess) return sarbp_hip_fail(e, "work-list upload");
  nitems = (int)items.size();
  return SARBP_OK;
}

void ItemBuffers::release() {
  if (d_reqs) (void)hipFree(d_reqs);
  if (d_items) (void)hipFree(d_items);
  if (d_seg) (void)hipFree(d_seg);
  if (d_heads) (void)hipFree(d_heads);
  if (d_list) (void)hipFree(d_list);
  if (h_stage) (void)hipHostFree(h_stage);
  *this = ItemBuffers();
}

struct sarbp_batch {
  sarbp_dataset *ds = nullptr;
  int max_rows = 1;
  std::vector<ReqBuf> reqs;
  ItemBuffers ib;
  int built_waves = 0;
  int tile_px = 64, chunk = 32;
  int64_t total_px = 0;
  bool any_f64 = false, any_split = false;
};

namespace {
int batch_build(sarbp_batch *b, int waves) {
  if (b->built_waves == waves) return SARBP_OK;
  std::vector<BatchReqDev> devs;
  std::vector<WorkItemCost> work;
  for (size_t r = 0; r < b->reqs.size(); ++r) {
    const ReqBuf &q = b->reqs[r];
    devs.push_back(reqbuf_dev(b->ds, q));
    const int64_t nt = q.tg.ntiles();
    for (int64_t t = 0; t < nt; t += waves) {
      int64_t c = 0;
      for (int64_t u = t; u < std::min(nt, t + waves); ++u) c = std::max<int64_t>(c, q.tile_len[(size_t)u]);
      work.push_back(WorkItemCost{BatchItem{(int32_t)r, (int32_t)t, 0, INT32_MAX, -1}, c});
    }
  }
  std::vector<BatchItem> items;
  std::vector<int32_t> seg;
  build_queues(work, items, seg);
  int rc = b->ib.upload(devs, items, seg, b->ds->stream);
  if (rc) return rc;
  b->built_waves = waves;
  return SARBP_OK;
}

int batch_time(sarbp_batch *b, int tile_px, int chunk, float *ms) {
  const int waves = tile_px / 64;
  int rc = batch_build(b, waves);
  if (rc) return rc;
  sarbp_dataset *ds = b->ds;
  std::lock_guard<std::mutex> lk(ds->timing_mu);
  hipError_t e = hipEventRecord(ds->ev0, ds->stream);
  if (e != hipSuccess) return sarbp_hip_fail(e, "event record");
  if (b->ib.nitems > 0) {
    rc = sarbp_launch_batch(ds->device, b->ib.d_reqs, b->ib.d_items, b->ib.d_seg, b->ib.d_heads, b->ib.nitems,
                            sarbp_batch_mode(b->any_f64, b->any_split), waves, chunk, b->max_rows, 0, ds->stream);
    if (rc) return rc;
  }
  e = hipEventRecord(ds->ev1, ds->stream);
  if (e == hipSuccess) e = hipEventSynchronize(ds->ev1);
  if (e != hipSuccess) return sarbp_hip_fail(e, "batch kernel");
  if (ms) {
    e = hipEventElapsedTime(ms, ds->ev0, ds->ev1);
    if (e != hipSuccess) return sarbp_hip_fail(e, "event timing");
  }
  return SARBP_OK;
}

bool valid_combo(int tile_px, int chunk, int max_rows) { return sarbp_batch_mapping_valid(tile_px, chunk, max_rows); }
}  // namespace

bool sarbp_batch_mapping_valid(int tile_px, int chunk, int max_rows) {
  if (tile_px != 64 && tile_px != 128 && tile_px != 256) return false;
  if (chunk < 8 || chunk > 1024 || chunk % 8) return false;
  return sarbp_batch_lds_bytes(tile_px / 64, chunk, max_rows) <= SARBP_BATCH_LDS_MAX;
}

extern "C" {

int sarbp_batch_create(sarbp_dataset *ds, const sarbp_request *reqs, int32_t nreq, sarbp_batch **out) {
  if (!out) return sarbp_fail(SARBP_EINVAL, "out is null");
  *out = nullptr;
  if (!ds) return sarbp_fail(SARBP_EINVAL, "dataset is null");
  if (nreq < 0 || (nreq > 0 && !reqs)) return sarbp_fail(SARBP_EINVAL, "bad request list");
  sarbp_batch *b = new (std::nothrow) sarbp_batch();
  if (!b) return sarbp_fail(SARBP_ENOMEM, "host allocation failed");
  b->ds = ds;
  ds->child_refs.fetch_add(1);  // released by sarbp_batch_destroy
  b->reqs.resize((size_t)nreq);
  for (int32_t r = 0; r < nreq; ++r) {
    int rc = reqbuf_create(ds, reqs[r], 0, default_tiling(reqs[r].precision).accum, b->reqs[(size_t)r],
                           ds->stream);
    if (rc) {
      sarbp_batch_destroy(b);
      return rc;
    }
    b->total_px += reqs[r].n;
    (reqs[r].precision == SARBP_PREC_SPLIT ? b->any_split : b->any_f64) = true;
    if (reqs[r].precision == SARBP_PREC_SPLIT)  // only split tiles stage LDS records
      b->max_rows = std::max(b->max_rows, b->reqs[(size_t)r].tg.ty);
  }
  hipError_t e = hipStreamSynchronize(ds->stream);
  if (e != hipSuccess) {
    sarbp_batch_destroy(b);
    return sarbp_hip_fail(e, "request upload");
  }
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(std::make_tuple(ds->nra, ds->naz, b->total_px, nreq, b->max_rows));
    if (it != g_tuned.end() && valid_combo(it->second.first, it->second.second, b->max_rows)) {
      b->tile_px = it->second.first;
      b->chunk = it->second.second;
    }
  }
  *out = b;
  return SARBP_OK;
}

int sarbp_batch_run(sarbp_batch *b, int32_t tile_px, int32_t pulse_chunk, float *gpu_ms) {
  if (!b) return sarbp_fail(SARBP_EINVAL, "null handle");
  const int tp = tile_px ? tile_px : b->tile_px;
  const int ch = pulse_chunk ? pulse_chunk : b->chunk;
  if (!valid_combo(tp, ch, b->max_rows)) return sarbp_fail(SARBP_EINVAL, "invalid tile_px / pulse_chunk combination");
  return batch_time(b, tp, ch, gpu_ms);
}

int sarbp_batch_autotune(sarbp_batch *b, int32_t reps, int32_t *best_tile_px, int32_t *best_chunk, float *best_ms) {
  if (!b) return sarbp_fail(SARBP_EINVAL, "null handle");
  if (reps < 1) reps = 1;
  float best = 0.f;
  int bt = 0, bc = 0;
  for (int tp : {64, 128, 256})
    for (int ch : {16, 32, 64, 128, 256, 512, 1024}) {
      if (!valid_combo(tp, ch, b->max_rows)) continue;
      float m = 0.f;
      for (int k = 0; k < reps; ++k) {
        float t = 0.f;
        int rc = batch_time(b, tp, ch, &t);
        if (rc) return rc;
        m = k == 0 ? t : std::min(m, t);
      }
      if (bt == 0 || m < best) {
        best = m;
        bt = tp;
        bc = ch;
      }
    }
  b->tile_px = bt;
  b->chunk = bc;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[std::make_tuple(b->ds->nra, b->ds->naz, b->total_px, (int32_t)b->reqs.size(), b->max_rows)] = {bt, bc};
  }
  if (best_tile_px) *best_tile_px = bt;
  if (best_chunk) *best_chunk = bc;
  if (best_ms) *best_ms = best;
  return SARBP_OK;
}

int sarbp_batch_download(sarbp_batch *b, int32_t req, double *out_c128, int32_t *ncontrib) {
  if (!b) return sarbp_fail(SARBP_EINVAL, "null handle");
  if (req < 0 || (size_t)req >= b->reqs.size()) return sarbp_fail(SARBP_EINVAL, "request index out of range");
  const ReqBuf &q = b->reqs[(size_t)req];
  hipStream_t st = b->ds->stream;
  hipError_t e = hipSuccess;
  if (out_c128 && q.n) e = hipMemcpyAsync(out_c128, q.out, (size_t)q.n * sizeof(double2), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && ncontrib && q.n)
    e = hipMemcpyAsync(ncontrib, q.ncnt, (size_t)q.n * sizeof(int32_t), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return sarbp_hip_fail(e, "batch download");
  return SARBP_OK;
}

int sarbp_batch_destroy(sarbp_batch *b) {
  if (!b) return SARBP_OK;
  if (b->ds && b->ds->stream) (void)hipStreamSynchronize(b->ds->stream);
  if (b->ds) b->ds->child_refs.fetch_sub(1);
  for (auto &q : b->reqs) reqbuf_free(q);
  b->ib.release();
  delete b;
  return SARBP_OK;
}

}  // extern "C"
