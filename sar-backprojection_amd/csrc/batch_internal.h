// Host-side structures shared by batched focusing (batch.hip) and the serving loop (server.hip).
#pragma once

#include <map>
#include <vector>

#include "sarbp_internal.h"

// Device blocks kept for reuse by one owner thread (the serving loop allocates and releases one
// block per request; hipMalloc/hipFree on every request would sit on its latency path).  Blocks
// are sized in 1 MiB steps; a request takes the smallest cached block that is large enough and at
// most twice its size.  At most `limit` bytes stay cached.
struct DevicePool {
  std::multimap<size_t, char *> blocks;
  size_t cached = 0;
  size_t limit = (size_t)2 << 30;
  char *get(size_t bytes, size_t *cap);
  void put(char *p, size_t cap);
  void release();
};

// Device buffers of one request: one allocation holding pixels, outputs and (serving) state.
struct ReqBuf {
  char *base = nullptr;
  int64_t bytes = 0;
  size_t cap = 0;                 // size of the block at base
  DevicePool *pool = nullptr;     // the block returns here (or is freed when null)
  int64_t n = 0, dimx = 0, dimy = 0;
  int precision = 0, accum = 0;
  TileGrid tg{};
  double *x = nullptr, *y = nullptr, *z = nullptr;
  uint8_t *mask = nullptr;
  int32_t *index = nullptr, *sint = nullptr;
  double2 *out = nullptr;
  int32_t *ncnt = nullptr;
  BpState *state = nullptr;
  int64_t pp = 0;                 // pixel-pulse pairs of the request (0 when tile_len was given)
  float r0min = 0.f;              // lower bound of the pixel-to-sensor distances (sarbp_pixel_r0min)
  bool f64_narrow = false;        // the live pixels' distances satisfy sarbp_narrow_ranges
  std::vector<int> tile_len;      // union length of each tile (pulses)
};

// Geometry of a request's tile rows that bounds the split form's range-offset error: for each
// row of each tile (its pixels with an aperture), o = the half-extent of the pixels' coordinates
// and c = the largest along-track direction cosine |dx| / R at the aperture ends of the row's
// first and last pixels.
struct SplitRows {
  bool any = false;   // some row has a pixel with an aperture
  double amax = 0.0;  // largest o_x * c + o_y + o_z over the rows (metres)
  double omax = 0.0;  // largest |o| over the rows (metres)
  double r0min = 0.0; // lower bound of the request's pixel-to-sensor distances (metres)
};
// Tile grid, per-tile union lengths (pulses) and pixel-pulse pairs of a request, in one pass over
// the pixels tile by tile (host); with `rows`, also the tile rows' geometry above.
void request_tiles(const sarbp_dataset *ds, const sarbp_request &r, TileGrid &tg, std::vector<int> &tile_len,
                   int64_t &pp, SplitRows *rows = nullptr);
// The request's tile grid (its tile_x, or its precision's default).
TileGrid request_tile_grid(const sarbp_request &r);
// request_tiles over the tile rows [by0, by1) of tg: writes their entries of tile_len (ntiles
// entries), adds their pixel-pulse pairs to pp and, with `rows`, folds their rows' geometry into
// rows (any, amax, omax; not r0min).  Bands may run on different threads into separate pp and rows.
void request_tiles_band(const sarbp_dataset *ds, const sarbp_request &r, const TileGrid &tg, int64_t by0, int64_t by1,
                        int *tile_len, int64_t &pp, SplitRows *rows);
// Allocate and upload one request (asynchronously on st; the caller synchronises).  state_slots:
// BpState entries per pixel (0: none; serving: 1 for the state between rounds, or one per pulse
// segment).  pool: where the device block comes from (null: hipMalloc); tile_len: precomputed by
// request_tiles (null: computed here).
// Byte offsets of a request's device block (each array 256-byte aligned): x, y, z, mask, index, sint
// (the inputs, [0, off[6])), out, ncnt, state slots; off[9] is the block size.
void reqbuf_layout(int64_t n, int state_slots, size_t off[10]);
// r0min: the pixels' distance bound when the caller has it (negative: computed here for split
// requests; FP64 requests do not use it).  packed: a page-locked host image of the input arrays in
// the block's layout (uploaded with one copy; null: the six arrays of r are copied one by one).
// host_out: page-locked host memory, already zero, that the kernels write the means (n double2)
// and, 256-byte aligned after them, the counts (n int32) into, instead of the device block.
// narrow: whether the FP64 terms may take the narrow-range form (request_narrow; negative:
// decided here).
int reqbuf_create(sarbp_dataset *ds, const sarbp_request &r, int state_slots, int accum, ReqBuf &b, hipStream_t st,
                  DevicePool *pool = nullptr, const std::vector<int> *tile_len = nullptr, float r0min = -1.f,
                  const void *packed = nullptr, char *host_out = nullptr, int narrow = -1);
// sarbp_live_range_bounds and sarbp_narrow_ranges for a request's host pixel arrays.
bool request_narrow(const sarbp_dataset *ds, const sarbp_request &r);
void reqbuf_free(ReqBuf &b);
BatchReqDev reqbuf_dev(const sarbp_dataset *ds, const ReqBuf &b);

// tile_px in {64, 128, 256}, pulse_chunk 8..1024 (x8), and the records of max_rows tile rows fit the
// batch kernel's LDS.
bool sarbp_batch_mapping_valid(int tile_px, int chunk, int max_rows);

struct WorkItemCost {
  BatchItem item;
  int64_t cost;
};
// Order items longest first and deal them to the per-XCD queues.
void build_queues(const std::vector<WorkItemCost> &work, std::vector<BatchItem> &items, std::vector<int32_t> &seg);

// Device copies of a work list (grown on demand, reused across launches), uploaded from a
// page-locked staging block.  The staging block is rewritten by the next upload: the previous
// upload must have completed by then (every caller synchronises each launch).
struct ItemBuffers {
  BatchReqDev *d_reqs = nullptr;
  BatchItem *d_items = nullptr;
  int32_t *d_seg = nullptr, *d_heads = nullptr;
  int32_t *d_list = nullptr;  // indices into d_reqs (the serving loop's segmented requests)
  size_t cap_reqs = 0, cap_items = 0, cap_list = 0;
  char *h_stage = nullptr;        // page-locked staging block of the work list
  const char *d_stage = nullptr;  // the same block as the device addresses it
  size_t cap_stage = 0;
  int nitems = 0;
  int upload(const std::vector<BatchReqDev> &reqs, const std::vector<BatchItem> &items,
             const std::vector<int32_t> &seg, hipStream_t st, const std::vector<int32_t> *list = nullptr);
  void release();
};
