// Scheduling rules of the serving loop that depend on integers only (no HIP): host code, also
// compiled on its own by the CPU tests.
#pragma once

#include <algorithm>
#include <cstdint>

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
