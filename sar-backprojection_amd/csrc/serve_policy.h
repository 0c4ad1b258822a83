// Scheduling rules of batched focusing and the serving loop that depend on integers only (no HIP):
// host code, also compiled on its own by the CPU tests.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <vector>

// Pulse segments per tile for a request of `live` tiles (union length > 0) on a device that holds
// `slots` wavefronts of the request's kernel at once.  The resident wavefronts take the items
// (tiles x segments) in turns, so the launch lasts about ceil(live K / slots) turns of max_len / K
// pulses; K is the smallest count within 5 % of the least such time, at most max_seg, with
// segments at least four pulse chunks long.  A request of more than 4 x slots tiles keeps one
// sequence per tile (and runs over several rounds).  Depends only on the request, the device and
// the configuration.
inline int segments_for(int64_t live, int max_len, int64_t slots, int chunk, int max_seg, int *seg_len) {
  *seg_len = 0;
  if (live <= 0 || max_seg <= 1 || slots <= 0 || live > 4 * slots) return 1;
  const int64_t kmax = std::min<int64_t>(max_seg, max_len / (4 * chunk));
  if (kmax <= 1) return 1;
  auto turns = [&](int64_t k) { return (double)((live * k + slots - 1) / slots) / (double)k; };
  double best = turns(1);
  for (int64_t k = 2; k <= kmax; ++k) best = std::min(best, turns(k));
  int64_t k = 1;
  while (turns(k) > 1.05 * best) ++k;
  if (k <= 1) return 1;
  const int64_t per = (max_len + k - 1) / k;
  const int len = (int)((per + chunk - 1) / chunk * chunk);
  *seg_len = len;
  return (max_len + len - 1) / len;
}

// Work list of a persistent launch: items (with .item and .cost, the pulses to walk) in NQ queues.
// Longest first (a counting sort over 32-pulse cost classes, list order within a class: the
// scheduling order only balances the load, results do not depend on it); then dealt to the
// queues in runs of 8 consecutive items, so each queue (XCD) works through neighbouring tiles and
// every queue sees the expensive items early.  items: the queues one after the other; seg: NQ + 1
// offsets into items.
template <int NQ, class Work, class Item>
void deal_queues(const std::vector<Work> &work, std::vector<Item> &items, std::vector<int32_t> &seg) {
  constexpr int NB = 1024;
  constexpr size_t K = 8;
  auto cls = [](int64_t c) { return NB - 1 - (int)std::min<int64_t>(NB - 1, std::max<int64_t>(0, c) >> 5); };
  const size_t n = work.size();
  std::vector<int32_t> start(NB + 1, 0);
  for (size_t k = 0; k < n; ++k) start[(size_t)cls(work[k].cost)]++;
  int32_t acc = 0;
  for (int c = 0; c <= NB; ++c) {
    const int32_t m = c < NB ? start[(size_t)c] : 0;
    start[(size_t)c] = acc;
    acc += m;
  }
  std::vector<int32_t> order(n);
  for (size_t k = 0; k < n; ++k) order[(size_t)start[(size_t)cls(work[k].cost)]++] = (int32_t)k;
  size_t cnt[NQ] = {}, pos[NQ];
  for (size_t k = 0; k < n; ++k) cnt[(k / K) % NQ]++;
  seg.assign(NQ + 1, 0);
  size_t o = 0;
  for (int q = 0; q < NQ; ++q) {
    seg[(size_t)q] = (int32_t)o;
    pos[q] = o;
    o += cnt[q];
  }
  seg[NQ] = (int32_t)o;
  items.resize(n);
  for (size_t k = 0; k < n; ++k) items[pos[(k / K) % NQ]++] = work[(size_t)order[k]].item;
}
