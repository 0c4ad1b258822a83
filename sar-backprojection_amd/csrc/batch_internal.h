// Host-side structures shared by batched focusing (batch.hip) and the serving loop (server.hip).
#pragma once

#include <vector>

#include "sarbp_internal.h"

// Device buffers of one request: one allocation holding pixels, outputs and (serving) state.
struct ReqBuf {
  char *base = nullptr;
  int64_t bytes = 0;
  int64_t n = 0, dimx = 0, dimy = 0;
  int precision = 0, accum = 0;
  TileGrid tg{};
  double *x = nullptr, *y = nullptr, *z = nullptr;
  uint8_t *mask = nullptr;
  int32_t *index = nullptr, *sint = nullptr;
  double2 *out = nullptr;
  int32_t *ncnt = nullptr;
  BpState *state = nullptr;
  int64_t pp = 0;                 // pixel-pulse pairs of the request
  std::vector<int> tile_len;      // union length of each tile (pulses)
};

// Allocate and upload one request (asynchronously on st; the caller synchronises).
int reqbuf_create(sarbp_dataset *ds, const sarbp_request &r, bool with_state, int accum, ReqBuf &b, hipStream_t st);
void reqbuf_free(ReqBuf &b);
BatchReqDev reqbuf_dev(const sarbp_dataset *ds, const ReqBuf &b);

struct WorkItemCost {
  BatchItem item;
  int64_t cost;
};
// Order items longest first and deal them to the per-XCD queues.
void build_queues(const std::vector<WorkItemCost> &work, std::vector<BatchItem> &items, std::vector<int32_t> &seg);

// Device copies of a work list (grown on demand, reused across launches).
struct ItemBuffers {
  BatchReqDev *d_reqs = nullptr;
  BatchItem *d_items = nullptr;
  int32_t *d_seg = nullptr, *d_heads = nullptr;
  size_t cap_reqs = 0, cap_items = 0;
  int nitems = 0;
  int upload(const std::vector<BatchReqDev> &reqs, const std::vector<BatchItem> &items,
             const std::vector<int32_t> &seg, hipStream_t st);
  void release();
};
