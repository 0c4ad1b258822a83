// Persistent batch kernel: many requests (pixel sets of uneven size, aperture length and
// precision, each with its own device-resident dataset) focused in one launch.
//
// The grid is the number of blocks the device holds at once, at most one per work item; each block
// repeatedly takes the next item from one of SARBP_BATCH_QUEUES queues -- its home queue (block % 8,
// the XCD the block runs on) first, then the others -- until every queue is drained, so every block
// reaches the exit.  An item is W adjacent tiles of one request, one per wavefront; each wavefront
// focuses its tile over the item's pulse range with the FP64 or split terms of bp_tile.h, starting
// from zero or from the request's saved per-pixel state, and writes the pixel means (last range of
// the tile), saves the state for a later launch, or -- for a pulse segment -- stores the segment's
// partial sums for k_bp_combine.  The arithmetic per pixel is the same whatever the other requests
// in the launch, so a request's result does not depend on the mix.
#include <algorithm>
#include <climits>

#include "bp_tile.h"

#pragma clang fp contract(off)

namespace {

// MODE: which requests a launch may hold (SARBP_BATCH_ANY, _F64 or _SPLIT); it only removes code
// that the launch's requests cannot reach, so every item computes the same bits in every mode.
template <int MODE>
__device__ __forceinline__ void batch_tile(const BatchReqDev &r, const BatchItem &it, const double2 *tab,
                                           int64_t tile, int rec0, int C) {
  const BpGeom g = r.g;
  const TileLane t = tile_lane(g, r.px, r.tg, tile, threadIdx.x & 63);
  const bool segment = it.slot >= 0;  // partial sums of one pulse segment, into slot it.slot
  const bool first = segment || it.p_off == 0;
  if (t.wlo >= t.whi) {  // nothing to focus in this tile
    if (!segment && first && t.valid) store_mean(r.out, r.ncnt, t.i, 0.0, 0.0, 0);
    return;
  }
  const int p0 = t.wlo + it.p_off;
  if (p0 >= t.whi) return;  // this tile finished in an earlier range (a segment: its slot stays zero)
  const int p1 = (it.p_end == INT32_MAX || it.p_end >= t.whi - t.wlo) ? t.whi : t.wlo + it.p_end;
  const bool last = segment || p1 >= t.whi;  // final sums of this item's pulses
  double ar = 0.0, ai = 0.0;
  float cr = 0.f, ci = 0.f;
  int32_t cnt = 0;
  if (!first && t.valid) {
    const BpState st = r.state[t.i];
    ar = st.ar;
    ai = st.ai;
    cr = st.cr;
    ci = st.ci;
    cnt = st.cnt;
  }
  bool split = MODE != SARBP_BATCH_F64 && r.precision == SARBP_PREC_SPLIT;
  SplitTile s;
  if (split) {
    s = split_tile(g, r.tg, t, r.px.r0min);
    split = s.use;  // not a compact tile: FP64 terms (same decision every range)
  }
  if (!split) {
    if (r.f64_narrow)
      f64_span<false>(g, tab, t, p0, p1, ar, ai, cnt);
    else
      f64_span<true>(g, tab, t, p0, p1, ar, ai, cnt);
  } else if (r.accum == 0) {
    float sr = (float)ar, si = (float)ai;  // the Kahan running sum is held exactly in the state
    split_span<0>(g, t, s, p0, p1, rec0, C, sr, si, cr, ci, ar, ai, cnt);
    if (last) {
      split_total<0>(sr, si, cr, ci, ar, ai);
    } else {
      ar = sr;
      ai = si;
    }
  } else {
    float sr = 0.f, si = 0.f;  // chunk sums are folded at every chunk end, so none is pending here
    split_span<1>(g, t, s, p0, p1, rec0, C, sr, si, cr, ci, ar, ai, cnt);
  }
  if (!t.valid) return;
  if (segment) {
    BpState st;
    st.ar = ar;
    st.ai = ai;
    st.cr = 0.f;
    st.ci = 0.f;
    st.cnt = cnt;
    st.pad = 0;
    r.state[(int64_t)it.slot * r.px.n + t.i] = st;
  } else if (last) {
    store_mean(r.out, r.ncnt, t.i, ar, ai, cnt);
  } else {
    BpState st;
    st.ar = ar;
    st.ai = ai;
    st.cr = cr;
    st.ci = ci;
    st.cnt = cnt;
    st.pad = 0;
    r.state[t.i] = st;
  }
}

// C pulses of records per refill, up to max_rows rows per tile: (C + 1) * max_rows records per
// wavefront (rows C + 1 records apart).
template <int MODE>
__device__ __forceinline__ void batch_loop(const BatchReqDev *__restrict__ reqs, const BatchItem *__restrict__ items,
                                           const int32_t *__restrict__ seg, int32_t *heads, int C, int max_rows) {
  __shared__ int32_t next_item;
  __shared__ double2 tab[SARBP_SCTAB];  // every request's table holds the same values
  load_sctab(reqs[0].g, tab);
  const int wave = threadIdx.x >> 6;
  const int rec0 = wave * (C + 1) * max_rows;  // this wavefront's records in sarbp_lds_rec
  const int home = blockIdx.x % SARBP_BATCH_QUEUES;
  for (int qi = 0; qi < SARBP_BATCH_QUEUES;) {
    const int q = (home + qi) % SARBP_BATCH_QUEUES;
    if (threadIdx.x == 0) {  // heads only grow: a drained queue is seen without an atomic
      int32_t *head = heads + q * SARBP_HEAD_STRIDE;
      const int32_t seen = __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      next_item = seg[q] + seen >= seg[q + 1] ? seg[q + 1] : seg[q] + atomicAdd(head, 1);
    }
    __syncthreads();
    const int32_t k = next_item;
    __syncthreads();
    if (k >= seg[q + 1]) {  // queue q drained: move on, never come back
      ++qi;
      continue;
    }
    const BatchItem it = items[k];
    const BatchReqDev &r = reqs[it.req];
    const int64_t tile = (int64_t)it.tile_lo + wave;
    if (tile < r.tg.ntiles()) batch_tile<MODE>(r, it, tab, tile, rec0, C);
  }
}

// Any mix of requests: compiled for four wavefronts per SIMD (at most 128 VGPRs).  The FP64 and
// split paths together need 141 registers unbounded (three waves); five or six waves spill in the
// loops and are slower (64 mixed requests: 29.0 ms at four waves, 32.6 ms at three, 41 ms at five).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_bp_batch(
    const BatchReqDev *__restrict__ reqs, const BatchItem *__restrict__ items, const int32_t *__restrict__ seg,
    int32_t *heads, int C, int max_rows) {
  batch_loop<SARBP_BATCH_ANY>(reqs, items, seg, heads, C, max_rows);
}

// FP64 requests only: six wavefronts per SIMD (80 VGPRs; 32 FP64 requests 18.2 ms against 23.3 ms
// in the mixed kernel; seven waves spill and take twice as long).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_bp_batch_f64(
    const BatchReqDev *__restrict__ reqs, const BatchItem *__restrict__ items, const int32_t *__restrict__ seg,
    int32_t *heads, int C, int max_rows) {
  batch_loop<SARBP_BATCH_F64>(reqs, items, seg, heads, C, max_rows);
}

// Split requests only (with the FP64 terms of tiles that are not compact): five wavefronts per
// SIMD (96 VGPRs; 32 split requests 8.1 ms, against 8.5 ms in the mixed kernel and 8.9 ms at six
// waves, where the loops spill).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_bp_batch_split(
    const BatchReqDev *__restrict__ reqs, const BatchItem *__restrict__ items, const int32_t *__restrict__ seg,
    int32_t *heads, int C, int max_rows) {
  batch_loop<SARBP_BATCH_SPLIT>(reqs, items, seg, heads, C, max_rows);
}

// Clears the queue heads before a launch (a kernel rather than a memset node, so that the sequence
// also replays correctly when a caller captures it in a graph).
__global__ __launch_bounds__(256) void k_zero_heads(int32_t *heads) {
  for (int i = threadIdx.x; i < (int)(SARBP_HEADS_BYTES / sizeof(int32_t)); i += blockDim.x) heads[i] = 0;
}

__global__ __launch_bounds__(256) void k_bp_combine(const BpState *__restrict__ part, int nseg, int64_t n,
                                                    double2 *__restrict__ out, int32_t *__restrict__ ncnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double ar = 0.0, ai = 0.0;
  int32_t cnt = 0;
  for (int k = 0; k < nseg; ++k) {  // fixed order: the result does not depend on segment timing
    const BpState s = part[(int64_t)k * n + i];
    ar += s.ar;
    ai += s.ai;
    cnt += s.cnt;
  }
  store_mean(out, ncnt, i, ar, ai, cnt);
}

}  // namespace

namespace {

// One request per blockIdx.y; the blocks along x stride over its slots (clear) or pixels (combine).
// The combine's time is its result writes into the requests' host blocks, so it keeps few blocks per
// request and leaves the CUs to the batch kernel that runs beside it.
constexpr int SEG_BLOCKS_X = 64;
constexpr int COMBINE_BLOCKS_X = 8;

__global__ __launch_bounds__(256) void k_segments_clear(const BatchReqDev *__restrict__ reqs,
                                                        const int32_t *__restrict__ list) {
  const BatchReqDev &r = reqs[list[blockIdx.y]];
  const int64_t words = r.px.n * (int64_t)r.nseg * (int64_t)(sizeof(BpState) / sizeof(int4));
  int4 *p = reinterpret_cast<int4 *>(r.state);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = make_int4(0, 0, 0, 0);
}

__global__ __launch_bounds__(256) void k_segments_combine(const BatchReqDev *__restrict__ reqs,
                                                          const int32_t *__restrict__ list) {
  const BatchReqDev &r = reqs[list[blockIdx.y]];
  const int64_t n = r.px.n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double ar = 0.0, ai = 0.0;
    int32_t cnt = 0;
    for (int k = 0; k < r.nseg; ++k) {  // the order of k_bp_combine: the same bits
      const BpState s = r.state[(int64_t)k * n + i];
      ar += s.ar;
      ai += s.ai;
      cnt += s.cnt;
    }
    store_mean(r.out, r.ncnt, i, ar, ai, cnt);
  }
}

}  // namespace

int sarbp_launch_segments_clear(const BatchReqDev *d_reqs, const int32_t *d_list, int count, hipStream_t st) {
  if (count <= 0) return SARBP_OK;
  if (count > 65535) return sarbp_fail(SARBP_EINVAL, "too many segmented requests for one launch");
  hipLaunchKernelGGL(k_segments_clear, dim3(SEG_BLOCKS_X, (unsigned)count), dim3(256), 0, st, d_reqs, d_list);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return sarbp_hip_fail(e, "segment clear launch");
  return SARBP_OK;
}

int sarbp_launch_segments_combine(const BatchReqDev *d_reqs, const int32_t *d_list, int count, hipStream_t st) {
  if (count <= 0) return SARBP_OK;
  if (count > 65535) return sarbp_fail(SARBP_EINVAL, "too many segmented requests for one launch");
  hipLaunchKernelGGL(k_segments_combine, dim3(COMBINE_BLOCKS_X, (unsigned)count), dim3(256), 0, st, d_reqs, d_list);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return sarbp_hip_fail(e, "segment combine launch");
  return SARBP_OK;
}

int sarbp_launch_combine(const BpState *part, int nseg, int64_t n, double2 *out, int32_t *ncnt, hipStream_t st) {
  if (n <= 0) return SARBP_OK;
  if (nseg < 1) return sarbp_fail(SARBP_EINVAL, "segment count must be positive");
  if ((n + 255) / 256 > INT32_MAX) return sarbp_fail(SARBP_EINVAL, "too many pixels for one combine launch");
  hipLaunchKernelGGL(k_bp_combine, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, part, nseg, n, out, ncnt);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return sarbp_hip_fail(e, "combine kernel launch");
  return SARBP_OK;
}

size_t sarbp_batch_lds_bytes(int waves, int chunk, int max_rows) {
  return (size_t)waves * (size_t)(chunk + 1) * (size_t)max_rows * sizeof(PulseRec);
}

int sarbp_batch_resident_blocks(int device, int mode, int waves, int chunk, int max_rows, int *blocks) {
  auto kern = mode == SARBP_BATCH_F64 ? k_bp_batch_f64 : mode == SARBP_BATCH_SPLIT ? k_bp_batch_split : k_bp_batch;
  int per_cu = 0, ncu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, waves * 64,
                                                              sarbp_batch_lds_bytes(waves, chunk, max_rows));
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess) return sarbp_hip_fail(e, "occupancy query");
  *blocks = std::max(1, per_cu) * ncu;
  return SARBP_OK;
}

int sarbp_launch_batch(int device, const BatchReqDev *d_reqs, const BatchItem *d_items, const int32_t *d_seg,
                       int32_t *d_heads, int nitems, int mode, int waves, int chunk, int max_rows, int grid_blocks,
                       hipStream_t st) {
  if (mode != SARBP_BATCH_ANY && mode != SARBP_BATCH_F64 && mode != SARBP_BATCH_SPLIT)
    return sarbp_fail(SARBP_EINVAL, "unknown batch launch mode");
  auto kern = mode == SARBP_BATCH_F64 ? k_bp_batch_f64 : mode == SARBP_BATCH_SPLIT ? k_bp_batch_split : k_bp_batch;
  if (waves != 1 && waves != 2 && waves != 4) return sarbp_fail(SARBP_EINVAL, "waves per block must be 1, 2 or 4");
  if (chunk < 8 || chunk > 1024 || chunk % 8) return sarbp_fail(SARBP_EINVAL, "pulse chunk must be 8..1024, x8");
  if (max_rows < 1 || max_rows > 64) return sarbp_fail(SARBP_EINVAL, "bad tile row count");
  const size_t lds = sarbp_batch_lds_bytes(waves, chunk, max_rows);
  if (lds > SARBP_BATCH_LDS_MAX) return sarbp_fail(SARBP_EINVAL, "waves x pulse chunk x tile rows needs more than 64 KiB of LDS");
  if (nitems <= 0) return SARBP_OK;
  hipLaunchKernelGGL(k_zero_heads, dim3(1), dim3(256), 0, st, d_heads);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return sarbp_hip_fail(e, "queue reset");
  if (grid_blocks <= 0) {
    int per_cu = 0, ncu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, waves * 64, lds);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) return sarbp_hip_fail(e, "occupancy query");
    if (per_cu < 1) per_cu = 1;
    grid_blocks = per_cu * ncu;
  }
  grid_blocks = std::min(grid_blocks, nitems);  // a block beyond the item count would find no work
  hipLaunchKernelGGL(kern, dim3((unsigned)grid_blocks), dim3(waves * 64), lds, st, d_reqs, d_items, d_seg,
                     d_heads, chunk, max_rows);
  e = hipGetLastError();
  if (e != hipSuccess) return sarbp_hip_fail(e, "batch kernel launch");
  return SARBP_OK;
}
