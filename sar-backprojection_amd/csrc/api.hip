// Host side of the C ABI declared in include/sarbp.h: device-resident datasets and pixel sets,
// focusing passes, error reporting.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "sarbp_internal.h"

namespace {
thread_local std::string g_last_error;

template <class T>
int dev_alloc(T **p, size_t count, int64_t &acct, const char *what) {
  *p = nullptr;
  if (count == 0) return SARBP_OK;
  hipError_t e = hipMalloc((void **)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return sarbp_fail(SARBP_ENOMEM, std::string("hipMalloc failed for ") + what + ": " + hipGetErrorString(e));
  }
  acct += (int64_t)(count * sizeof(T));
  return SARBP_OK;
}

void dev_free(void *p) {
  if (p) (void)hipFree(p);
}

int h2d(void *dst, const void *src, size_t bytes, hipStream_t st, const char *what) {
  if (bytes == 0) return SARBP_OK;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return sarbp_hip_fail(e, what);
  return SARBP_OK;
}

int check_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return sarbp_fail(SARBP_ENODEV, "no HIP device visible");
  }
  return SARBP_OK;
}

int upload_traj(sarbp_dataset *ds, const double *tx, const double *ty, const double *tz) {
  ds->htx.assign(tx, tx + ds->naz);
  ds->hty.assign(ty, ty + ds->naz);
  ds->htz.assign(tz, tz + ds->naz);
  const double *c[3] = {tx, ty, tz};
  for (int k = 0; k < 3; ++k) {
    ds->tlo[k] = *std::min_element(c[k], c[k] + ds->naz);
    ds->thi[k] = *std::max_element(c[k], c[k] + ds->naz);
  }
  std::vector<double4> t((size_t)(ds->naz + SARBP_TRAJ_PAD));
  for (int64_t a = 0; a < ds->naz + SARBP_TRAJ_PAD; ++a) {
    const int64_t k = std::min(a, ds->naz - 1);
    t[(size_t)a] = make_double4(tx[k], ty[k], tz[k], 0.0);
  }
  return h2d(ds->traj, t.data(), t.size() * sizeof(double4), ds->stream, "trajectory upload");
}

// (sin, cos)(j*pi/128) for j = 0..255 in extended precision, rounded once to double.
int upload_sctab(sarbp_dataset *ds) {
  std::vector<double2> t(SARBP_SCTAB);
  const long double h = acosl(-1.0L) / 128.0L;
  for (int j = 0; j < SARBP_SCTAB; ++j) t[(size_t)j] = make_double2((double)sinl(h * j), (double)cosl(h * j));
  return h2d(ds->sctab, t.data(), t.size() * sizeof(double2), ds->stream, "sin/cos table upload");
}

int dataset_init(int64_t nra, int64_t naz, const double *tx, const double *ty, const double *tz,
                 double near_range, double dra, double lambda_val, sarbp_dataset **out) {
  if (!out) return sarbp_fail(SARBP_EINVAL, "out is null");
  *out = nullptr;
  if (nra < 2 || naz < 3) return sarbp_fail(SARBP_EINVAL, "RawData must have nra >= 2 and naz >= 3");
  if (nra > ((int64_t)1 << 24) || naz > (int64_t)0x7fffffff)
    return sarbp_fail(SARBP_EINVAL, "nra must be at most 2^24 range bins and naz must fit in int32");
  if (!tx || !ty || !tz) return sarbp_fail(SARBP_EINVAL, "trajectory pointer is null");
  if (!(dra != 0.0) || !std::isfinite(dra) || !std::isfinite(near_range) || !(lambda_val != 0.0) ||
      !std::isfinite(lambda_val))
    return sarbp_fail(SARBP_EINVAL, "near_range, dra and lambda must be finite, dra and lambda non-zero");
  int rc = check_device();
  if (rc) return rc;
  sarbp_dataset *ds = new (std::nothrow) sarbp_dataset();
  if (!ds) return sarbp_fail(SARBP_ENOMEM, "host allocation failed");
  ds->nra = nra;
  ds->naz = naz;
  ds->near_range = near_range;
  ds->dra = dra;
  ds->lambda_val = lambda_val;
  hipError_t e = hipGetDevice(&ds->device);
  if (e == hipSuccess) e = hipStreamCreate(&ds->stream);
  if (e == hipSuccess) e = hipEventCreate(&ds->ev0);
  if (e == hipSuccess) e = hipEventCreate(&ds->ev1);
  if (e != hipSuccess) {
    rc = sarbp_hip_fail(e, "stream/event creation");
    sarbp_dataset_destroy(ds);
    return rc;
  }
  rc = dev_alloc(&ds->raw, (size_t)nra * (size_t)naz, ds->bytes, "RawData");
  if (!rc) rc = dev_alloc(&ds->traj, (size_t)(naz + SARBP_TRAJ_PAD), ds->bytes, "trajectory");
  if (!rc) rc = upload_traj(ds, tx, ty, tz);
  if (!rc) rc = dev_alloc(&ds->sctab, SARBP_SCTAB, ds->bytes, "sin/cos table");
  if (!rc) rc = upload_sctab(ds);
  if (rc) {
    sarbp_dataset_destroy(ds);
    return rc;
  }
  *out = ds;
  return SARBP_OK;
}
}  // namespace

int sarbp_fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

int sarbp_hip_fail(hipError_t e, const char *what) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  (void)hipGetLastError();
  return SARBP_EDEVICE;
}

Tiling default_tiling(int precision) {
  Tiling t;
  // fastest measured on the 4000 x 250 x 40001 workload: 8 x 8 tiles for FP64 (no per-row
  // records to stage), 16 x 4 for the split path (four row reference ranges per pulse)
  t.kernel = 1;
  t.tile_x = precision == SARBP_PREC_SPLIT ? 16 : 8;
  t.accum = 0;
  return t;
}

void PixelBoxes::add(const double *x, const double *y, const double *z, const uint8_t *mask, const int32_t *sint,
                     int64_t i0, int64_t i1) {
  for (int64_t i = i0; i < i1; ++i) {
    const double p[3] = {x[i], y[i], z[i]};
    const bool live = mask && mask[i] == 1 && sint[i] != -9999;
    any_live = any_live || live;
    for (int k = 0; k < 3; ++k) {
      const double v = p[k];
      if (!std::isfinite(v)) {
        finite = false;
        if (live) live_finite = false;
        continue;
      }
      lo[k] = std::min(lo[k], v);
      hi[k] = std::max(hi[k], v);
      if (live) {
        llo[k] = std::min(llo[k], v);
        lhi[k] = std::max(lhi[k], v);
      }
    }
  }
}

void PixelBoxes::merge(const PixelBoxes &o) {
  for (int k = 0; k < 3; ++k) {
    lo[k] = std::min(lo[k], o.lo[k]);
    hi[k] = std::max(hi[k], o.hi[k]);
    llo[k] = std::min(llo[k], o.llo[k]);
    lhi[k] = std::max(lhi[k], o.lhi[k]);
  }
  finite = finite && o.finite;
  live_finite = live_finite && o.live_finite;
  any_live = any_live || o.any_live;
}

float sarbp_r0min_of(const sarbp_dataset *ds, const PixelBoxes &b) {
  if (!b.finite || !(b.lo[0] <= b.hi[0])) return 0.f;  // no pixel, or a coordinate that is not finite
  double d2 = 0.0;
  for (int k = 0; k < 3; ++k) {
    const double gap = std::max({0.0, b.lo[k] - ds->thi[k], ds->tlo[k] - b.hi[k]});
    d2 += gap * gap;
  }
  const double d = std::sqrt(d2);
  if (!std::isfinite(d)) return 0.f;
  return std::nextafter((float)d, 0.f);  // rounded down
}

float sarbp_pixel_r0min(const sarbp_dataset *ds, const double *x, const double *y, const double *z, int64_t n) {
  if (n <= 0) return 0.f;
  PixelBoxes b;
  b.add(x, y, z, nullptr, nullptr, 0, n);
  return sarbp_r0min_of(ds, b);
}

bool sarbp_range_bounds_of(const sarbp_dataset *ds, const PixelBoxes &b, double &rmin, double &rmax) {
  rmin = rmax = 0.0;
  if (!b.live_finite) return false;
  if (!b.any_live) return true;
  double dmin = 0.0, dmax = 0.0;
  for (int k = 0; k < 3; ++k) {
    const double gap = std::max({0.0, b.llo[k] - ds->thi[k], ds->tlo[k] - b.lhi[k]});
    const double span = std::max(b.lhi[k] - ds->tlo[k], ds->thi[k] - b.llo[k]);
    dmin += gap * gap;
    dmax += span * span;
  }
  rmin = std::sqrt(dmin) * (1.0 - 1e-12);
  rmax = std::sqrt(dmax) * (1.0 + 1e-12);
  return std::isfinite(rmax);
}

bool sarbp_live_range_bounds(const sarbp_dataset *ds, const double *x, const double *y, const double *z,
                             const uint8_t *mask, const int32_t *sint, int64_t n, double &rmin, double &rmax) {
  PixelBoxes b;
  b.add(x, y, z, mask, sint, 0, n);
  return sarbp_range_bounds_of(ds, b, rmin, rmax);
}

bool sarbp_narrow_ranges(const sarbp_dataset *ds, double r0min, double rmax) {
  if (rmax == 0.0) return true;  // no pixel focuses
  if (!(r0min >= 0x1p-440 && rmax <= 0x1p440)) return false;
  const BpGeom g = ds->geom();
  return std::fabs(g.kphase) * rmax + std::fabs(g.beta) * rmax * rmax < 0x1p24;
}

BpGeom sarbp_scene::geom() const {
  BpGeom g = ds->geom();
  g.bin_lo = bin_lo;
  return g;
}

BpGeom sarbp_dataset::geom() const {
  BpGeom g;
  g.raw = raw;
  g.traj = traj;
  g.nra = nra;
  g.naz = naz;
  g.near_range = near_range;
  g.dra = dra;
  g.inv_dra = 1.0 / dra;
  // Same expressions and evaluation order as sar_focuser_puntual.py:40 and :116.
  const double pi = SARBP_PI32_F64;
  g.kphase = -4.0 * pi / lambda_val;
  g.beta = 4.0 * pi * SARBP_ALFA_F64 / (SARBP_C_F64 * SARBP_C_F64);
  g.sctab = sctab;
  g.bin_lo = 0;
  return g;
}

extern "C" {

int sarbp_abi_version(void) { return SARBP_ABI_VERSION; }

const char *sarbp_strerror(int status) {
  switch (status) {
    case SARBP_OK: return "ok";
    case SARBP_EINVAL: return "invalid argument";
    case SARBP_ENOMEM: return "out of memory";
    case SARBP_ELAUNCH: return "kernel launch failed";
    case SARBP_EDEVICE: return "HIP runtime error";
    case SARBP_ENODEV: return "no HIP device";
    case SARBP_EBUDGET: return "device-memory budget exceeded";
    case SARBP_ESTATE: return "invalid state";
    case SARBP_EAGAIN: return "not finished";
    default: return "unknown status";
  }
}

const char *sarbp_last_error(void) { return g_last_error.c_str(); }

int sarbp_device_count(int *count) {
  if (!count) return sarbp_fail(SARBP_EINVAL, "count is null");
  *count = 0;
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    (void)hipGetLastError();
  }
  return SARBP_OK;
}

int sarbp_dataset_create(const void *raw_c64, int64_t nra, int64_t naz, const double *tx, const double *ty,
                         const double *tz, double near_range, double dra, double lambda_val,
                         sarbp_dataset **out) {
  if (!raw_c64) return sarbp_fail(SARBP_EINVAL, "raw is null");
  int rc = dataset_init(nra, naz, tx, ty, tz, near_range, dra, lambda_val, out);
  if (rc) return rc;
  sarbp_dataset *ds = *out;
  rc = h2d(ds->raw, raw_c64, (size_t)nra * (size_t)naz * sizeof(float2), ds->stream, "RawData upload");
  if (rc) {
    sarbp_dataset_destroy(ds);
    *out = nullptr;
  }
  return rc;
}

int sarbp_dataset_create_streamed(sarbp_fill_fn fill, void *user, int64_t nra, int64_t naz, const double *tx,
                                  const double *ty, const double *tz, double near_range, double dra,
                                  double lambda_val, sarbp_dataset **out) {
  if (!fill) return sarbp_fail(SARBP_EINVAL, "fill callback is null");
  int rc = dataset_init(nra, naz, tx, ty, tz, near_range, dra, lambda_val, out);
  if (rc) return rc;
  sarbp_dataset *ds = *out;
  // Pinned staging in ~64 MiB pulse blocks, double-buffered so the callback overlaps the copy.
  const int64_t col = nra * (int64_t)sizeof(float2);
  int64_t blk = (int64_t)(64 << 20) / col;
  if (blk < 1) blk = 1;
  void *stage[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  hipError_t e = hipHostMalloc(&stage[0], (size_t)(blk * col), hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc(&stage[1], (size_t)(blk * col), hipHostMallocDefault);
  if (e == hipSuccess) e = hipEventCreate(&done[0]);
  if (e == hipSuccess) e = hipEventCreate(&done[1]);
  if (e != hipSuccess) rc = sarbp_hip_fail(e, "staging allocation");
  int k = 0;
  for (int64_t a0 = 0; !rc && a0 < naz; a0 += blk, k ^= 1) {
    const int64_t a1 = a0 + blk < naz ? a0 + blk : naz;
    e = hipEventSynchronize(done[k]);
    if (e != hipSuccess) { rc = sarbp_hip_fail(e, "staging wait"); break; }
    if (fill(user, a0, a1, stage[k]) != 0) { rc = sarbp_fail(SARBP_EINVAL, "fill callback failed"); break; }
    e = hipMemcpyAsync(ds->raw + nra * a0, stage[k], (size_t)((a1 - a0) * col), hipMemcpyHostToDevice,
                       ds->stream);
    if (e == hipSuccess) e = hipEventRecord(done[k], ds->stream);
    if (e != hipSuccess) rc = sarbp_hip_fail(e, "RawData block upload");
  }
  e = hipStreamSynchronize(ds->stream);
  if (!rc && e != hipSuccess) rc = sarbp_hip_fail(e, "RawData upload");
  for (int i = 0; i < 2; ++i) {
    if (stage[i]) (void)hipHostFree(stage[i]);
    if (done[i]) (void)hipEventDestroy(done[i]);
  }
  if (rc) {
    sarbp_dataset_destroy(ds);
    *out = nullptr;
  }
  return rc;
}

int sarbp_dataset_destroy(sarbp_dataset *ds) {
  if (!ds) return SARBP_OK;
  if (ds->server_refs.load() > 0)
    return sarbp_fail(SARBP_ESTATE, "the dataset has server requests that are not done or failed");
  if (ds->child_refs.load() > 0)
    return sarbp_fail(SARBP_ESTATE, "the dataset still has scenes or batches; destroy them first");
  if (ds->stream) (void)hipStreamSynchronize(ds->stream);
  dev_free(ds->raw);
  dev_free(ds->traj);
  dev_free(ds->sctab);
  if (ds->ev0) (void)hipEventDestroy(ds->ev0);
  if (ds->ev1) (void)hipEventDestroy(ds->ev1);
  if (ds->stream) (void)hipStreamDestroy(ds->stream);
  delete ds;
  return SARBP_OK;
}

int sarbp_eval_pairs(sarbp_dataset *ds, const double *x, const double *y, const double *z, const int32_t *pulse,
                     int64_t n, double *r2, double *r, int32_t *bin, double *ph, double *sn, double *cs) {
  if (!ds) return sarbp_fail(SARBP_EINVAL, "null dataset");
  if (n < 0 || (n > 0 && (!x || !y || !z || !pulse))) return sarbp_fail(SARBP_EINVAL, "bad pair arrays");
  for (int64_t i = 0; i < n; ++i)
    if (pulse[i] < 0 || pulse[i] >= ds->naz) return sarbp_fail(SARBP_EINVAL, "pulse index outside [0, naz)");
  if (n == 0) return SARBP_OK;
  const size_t un = (size_t)n;
  char *buf = nullptr;
  const size_t in_bytes = un * (3 * sizeof(double) + sizeof(int32_t));
  const size_t out_bytes = un * (5 * sizeof(double) + sizeof(int32_t));
  hipError_t e = hipMalloc((void **)&buf, in_bytes + out_bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return sarbp_fail(SARBP_ENOMEM, "pair buffers");
  }
  double *dx = (double *)buf, *dy = dx + un, *dz = dy + un;
  double *o_r2 = dz + un, *o_r = o_r2 + un, *o_ph = o_r + un, *o_sn = o_ph + un, *o_cs = o_sn + un;
  int32_t *dp = (int32_t *)(o_cs + un), *o_bin = dp + un;
  int rc = h2d(dx, x, un * 8, ds->stream, "pair upload");
  if (!rc) rc = h2d(dy, y, un * 8, ds->stream, "pair upload");
  if (!rc) rc = h2d(dz, z, un * 8, ds->stream, "pair upload");
  if (!rc) rc = h2d(dp, pulse, un * 4, ds->stream, "pair upload");
  if (!rc) rc = sarbp_launch_eval_pairs(ds, dx, dy, dz, dp, n, o_r2, o_r, o_bin, o_ph, o_sn, o_cs, ds->stream);
  struct Out {
    void *host;
    const void *dev;
    size_t bytes;
  } outs[6] = {{r2, o_r2, un * 8}, {r, o_r, un * 8}, {bin, o_bin, un * 4}, {ph, o_ph, un * 8}, {sn, o_sn, un * 8},
               {cs, o_cs, un * 8}};
  for (const Out &o : outs) {
    if (rc || !o.host) continue;
    e = hipMemcpyAsync(o.host, o.dev, o.bytes, hipMemcpyDeviceToHost, ds->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ds->stream);
    if (e != hipSuccess) rc = sarbp_hip_fail(e, "pair download");
  }
  if (!rc) {
    e = hipStreamSynchronize(ds->stream);
    if (e != hipSuccess) rc = sarbp_hip_fail(e, "pair kernel");
  }
  (void)hipStreamSynchronize(ds->stream);
  (void)hipFree(buf);
  return rc;
}

int sarbp_dataset_bytes(const sarbp_dataset *ds, int64_t *bytes) {
  if (!ds || !bytes) return sarbp_fail(SARBP_EINVAL, "null argument");
  *bytes = ds->bytes;
  return SARBP_OK;
}

int sarbp_scene_create(sarbp_dataset *ds, const double *x, const double *y, const double *z, const uint8_t *mask,
                       const int32_t *index, const int32_t *sint, int64_t n, int64_t dimx, sarbp_scene **out) {
  if (!out) return sarbp_fail(SARBP_EINVAL, "out is null");
  *out = nullptr;
  if (!ds) return sarbp_fail(SARBP_EINVAL, "dataset is null");
  if (n < 0) return sarbp_fail(SARBP_EINVAL, "negative pixel count");
  if (dimx < 0 || (n > 0 && dimx > 0 && n % dimx != 0))
    return sarbp_fail(SARBP_EINVAL, "pixel count must be a multiple of the grid's fast dimension dimx");
  if (n > 0 && (!x || !y || !z || !mask || !index || !sint))
    return sarbp_fail(SARBP_EINVAL, "pixel array pointer is null");
  sarbp_scene *sc = new (std::nothrow) sarbp_scene();
  if (!sc) return sarbp_fail(SARBP_ENOMEM, "host allocation failed");
  sc->ds = ds;
  sc->n = n;
  sc->dimx = dimx > 0 ? dimx : (n > 0 ? n : 1);
  sc->dimy = n > 0 ? n / sc->dimx : 0;
  sc->tiling_f64 = default_tiling(SARBP_PREC_F64);
  sc->tiling_split = default_tiling(SARBP_PREC_SPLIT);
  const size_t un = (size_t)n;
  int rc = dev_alloc(&sc->x, un, sc->bytes, "x");
  if (!rc) rc = dev_alloc(&sc->y, un, sc->bytes, "y");
  if (!rc) rc = dev_alloc(&sc->z, un, sc->bytes, "z");
  if (!rc) rc = dev_alloc(&sc->mask, un, sc->bytes, "mask");
  if (!rc) rc = dev_alloc(&sc->index, un, sc->bytes, "index");
  if (!rc) rc = dev_alloc(&sc->sint, un, sc->bytes, "sint");
  if (!rc) rc = dev_alloc(&sc->out, un, sc->bytes, "output");
  if (!rc) rc = dev_alloc(&sc->ncnt, un, sc->bytes, "counts");
  if (!rc) rc = h2d(sc->x, x, un * 8, ds->stream, "x upload");
  if (!rc) rc = h2d(sc->y, y, un * 8, ds->stream, "y upload");
  if (!rc) rc = h2d(sc->z, z, un * 8, ds->stream, "z upload");
  if (!rc) rc = h2d(sc->mask, mask, un, ds->stream, "mask upload");
  if (!rc) rc = h2d(sc->index, index, un * 4, ds->stream, "index upload");
  if (!rc) rc = h2d(sc->sint, sint, un * 4, ds->stream, "sint upload");
  if (rc) {
    sarbp_scene_destroy(sc);
    return rc;
  }
  sc->r0min = sarbp_pixel_r0min(ds, x, y, z, n);
  {
    double rmin = 0.0, rmax = 0.0;
    sc->f64_narrow = sarbp_live_range_bounds(ds, x, y, z, mask, sint, n, rmin, rmax) &&
                     sarbp_narrow_ranges(ds, rmin, rmax);
  }
  ds->child_refs.fetch_add(1);
  sc->counted = true;
  *out = sc;
  return SARBP_OK;
}

int sarbp_scene_destroy(sarbp_scene *sc) {
  if (!sc) return SARBP_OK;
  if (sc->ds && sc->ds->stream) (void)hipStreamSynchronize(sc->ds->stream);
  if (sc->counted) sc->ds->child_refs.fetch_sub(1);
  dev_free(sc->x);
  dev_free(sc->y);
  dev_free(sc->z);
  dev_free(sc->mask);
  dev_free(sc->index);
  dev_free(sc->sint);
  dev_free(sc->out);
  dev_free(sc->ncnt);
  delete sc;
  return SARBP_OK;
}

int sarbp_scene_focus(sarbp_dataset *ds, sarbp_scene *sc, int precision, float *gpu_ms) {
  if (!ds || !sc) return sarbp_fail(SARBP_EINVAL, "null handle");
  if (sc->ds != ds) return sarbp_fail(SARBP_EINVAL, "scene belongs to another dataset");
  if (precision != SARBP_PREC_F64 && precision != SARBP_PREC_SPLIT)
    return sarbp_fail(SARBP_EINVAL, "unknown precision");
  std::lock_guard<std::mutex> lk(ds->timing_mu);  // scenes of one dataset may be focused from several threads
  hipError_t e = hipEventRecord(ds->ev0, ds->stream);
  if (e != hipSuccess) return sarbp_hip_fail(e, "event record");
  int rc = precision == SARBP_PREC_SPLIT ? sarbp_launch_split(ds, sc, sc->tiling_split, ds->stream)
                                         : sarbp_launch_f64(ds, sc, sc->tiling_f64, ds->stream);
  if (rc) return rc;
  e = hipEventRecord(ds->ev1, ds->stream);
  if (e == hipSuccess) e = hipEventSynchronize(ds->ev1);
  if (e != hipSuccess) return sarbp_hip_fail(e, "focus kernel");
  if (gpu_ms) {
    float ms = 0.f;
    e = hipEventElapsedTime(&ms, ds->ev0, ds->ev1);
    if (e != hipSuccess) return sarbp_hip_fail(e, "event timing");
    *gpu_ms = ms;
  }
  return SARBP_OK;
}

int sarbp_scene_download(sarbp_scene *sc, double *out_c128, int32_t *ncontrib) {
  if (!sc) return sarbp_fail(SARBP_EINVAL, "null handle");
  hipStream_t st = sc->ds->stream;
  hipError_t e = hipSuccess;
  if (out_c128 && sc->n)
    e = hipMemcpyAsync(out_c128, sc->out, (size_t)sc->n * sizeof(double2), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && ncontrib && sc->n)
    e = hipMemcpyAsync(ncontrib, sc->ncnt, (size_t)sc->n * sizeof(int32_t), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return sarbp_hip_fail(e, "result download");
  return SARBP_OK;
}

int sarbp_scene_set_tiling(sarbp_scene *sc, int precision, int kernel, int tile_x, int accum) {
  if (!sc) return sarbp_fail(SARBP_EINVAL, "null handle");
  if (precision != SARBP_PREC_F64 && precision != SARBP_PREC_SPLIT)
    return sarbp_fail(SARBP_EINVAL, "unknown precision");
  if (tile_x != 64 && tile_x != 32 && tile_x != 16 && tile_x != 8 && tile_x != 4)
    return sarbp_fail(SARBP_EINVAL, "tile_x must be 64, 32, 16, 8 or 4");
  if (kernel < 0 || kernel > 1 || accum < 0 || accum > 1)
    return sarbp_fail(SARBP_EINVAL, "kernel and accum must be 0 or 1");
  if (precision == SARBP_PREC_SPLIT && kernel == 0)
    return sarbp_fail(SARBP_EINVAL, "the split path always uses tiles");
  Tiling &t = precision == SARBP_PREC_SPLIT ? sc->tiling_split : sc->tiling_f64;
  t.kernel = kernel;
  t.tile_x = tile_x;
  t.accum = accum;
  return SARBP_OK;
}

int sarbp_scene_set_bin_min(sarbp_scene *sc, int32_t bin_min) {
  if (!sc) return sarbp_fail(SARBP_EINVAL, "null handle");
  // bins below 0 never contribute (process_point's explicit bound check, sar_focuser_puntual.py:138),
  // and none at or above nra-1 (:106)
  const int64_t hi = sc->ds->nra - 1;
  sc->bin_lo = (int32_t)std::min<int64_t>(hi, std::max<int64_t>(0, bin_min));
  return SARBP_OK;
}

int sarbp_focus(sarbp_dataset *ds, const double *x, const double *y, const double *z, const uint8_t *mask,
                const int32_t *index, const int32_t *sint, int64_t n, int64_t dimx, int precision,
                double *out_c128, int32_t *ncontrib) {
  sarbp_scene *sc = nullptr;
  int rc = sarbp_scene_create(ds, x, y, z, mask, index, sint, n, dimx, &sc);
  if (rc) return rc;
  rc = sarbp_scene_focus(ds, sc, precision, nullptr);
  if (!rc) rc = sarbp_scene_download(sc, out_c128, ncontrib);
  sarbp_scene_destroy(sc);
  return rc;
}

int64_t sarbp_count_pp(const uint8_t *mask, const int32_t *index, const int32_t *sint, int64_t n, int64_t naz) {
  if (!mask || !index || !sint || n <= 0) return 0;
  int64_t tot = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (mask[i] != 1 || sint[i] == -9999) continue;
    int64_t lo, hi;
    sarbp_pulse_range(index[i], sint[i], naz, lo, hi);
    tot += hi - lo;
  }
  return tot;
}

}  // extern "C"
