"""Scheduling rules of batched focusing and serving (csrc/serve_policy.h), compiled on their own with
the host compiler.

Segment count.
For a request of `live` tiles whose longest tile union is `max_len` pulses, on a device holding
`slots` wavefronts: segments tile the union exactly (K segments of a whole number of chunks, at
least four chunks each, the last one non-empty), K never exceeds the limit, requests above four
device-fulls keep one sequence, and the launch time of the chosen segments -- ceil(live K / slots)
turns of one segment each -- is within 5 % of the best over every K (brute force), up to the
rounding of segments to whole chunks.
"""
import math
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

DRIVER = r'''
#include <cstdio>
#include <cstring>
#include "serve_policy.h"
struct It { int id; };
struct W { It item; long long cost; };
int main(int argc, char **argv) {
  if (argc > 1 && std::strcmp(argv[1], "deal") == 0) {  // n costs -> queue offsets, then item ids
    int n;
    while (std::scanf("%d", &n) == 1) {
      std::vector<W> w((size_t)n);
      for (int k = 0; k < n; ++k) { w[(size_t)k].item.id = k; std::scanf("%lld", &w[(size_t)k].cost); }
      std::vector<It> items;
      std::vector<int32_t> seg;
      deal_queues<8>(w, items, seg);
      for (int32_t v : seg) std::printf("%d ", v);
      std::printf("\n");
      for (const It &i : items) std::printf("%d ", i.id);
      std::printf("\n");
    }
    return 0;
  }
  long long live, slots; int max_len, chunk, max_seg;
  while (std::scanf("%lld %d %lld %d %d", &live, &max_len, &slots, &chunk, &max_seg) == 5) {
    int len = -1;
    const int k = segments_for(live, max_len, slots, chunk, max_seg, &len);
    std::printf("%d %d\n", k, len);
  }
  return 0;
}
'''


@pytest.fixture(scope="module")
def sched(tmp_path_factory):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("sched")
    src = d / "drv.cpp"
    src.write_text(DRIVER)
    exe = d / "drv"
    subprocess.run([cxx, "-O1", "-std=c++17", "-I", os.path.join(ROOT, "sar-backprojection_amd", "csrc"), "-o",
                    str(exe), str(src)], check=True)

    def run(cases):
        inp = "".join(f"{a} {b} {c} {e} {f}\n" for a, b, c, e, f in cases)
        out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
        return [tuple(map(int, ln.split())) for ln in out if ln.strip()]

    def deal(costs):
        inp = f"{len(costs)} " + " ".join(str(c) for c in costs) + "\n"
        out = subprocess.run([str(exe), "deal"], input=inp, capture_output=True, text=True, check=True).stdout
        lines = out.split("\n")
        return [int(v) for v in lines[0].split()], [int(v) for v in lines[1].split()]
    run.deal = deal
    return run


def turns(live, k, slots):
    return math.ceil(live * k / slots) / k


CASES = [(live, ml, slots, chunk, ms)
         for live in (1, 7, 64, 290, 1385, 2260, 4000, 6144, 15625, 24576, 24577, 100000)
         for ml in (64, 127, 128, 700, 1200, 3600, 4001)
         for slots in (5120, 6144)
         for chunk in (32, 64)
         for ms in (1, 2, 16)]


def test_segments_tile_the_union(sched):
    for (live, ml, slots, chunk, ms), (k, ln) in zip(CASES, sched(CASES)):
        assert 1 <= k <= max(1, ms), (live, ml, k)
        if k == 1:
            assert ln == 0
            continue
        assert ln % chunk == 0 and ln >= 4 * chunk, (live, ml, k, ln)
        assert (k - 1) * ln < ml <= k * ln, (live, ml, k, ln)  # exact cover, last segment non-empty


def test_segment_count_is_near_the_best(sched):
    for (live, ml, slots, chunk, ms), (k, ln) in zip(CASES, sched(CASES)):
        if live > 4 * slots or ms <= 1:
            assert k == 1
            continue
        kmax = min(ms, ml // (4 * chunk))
        if kmax <= 1:
            assert k == 1
            continue
        best = min(turns(live, j, slots) for j in range(1, kmax + 1))  # in units of max_len pulses
        assert k <= kmax
        if k == 1:
            assert turns(live, 1, slots) <= 1.05 * best
            continue
        # the launch (turns x segment length, in pulses) is within 5 % of the best count's, up to the
        # rounding of each segment to whole chunks
        n_turns = math.ceil(live * k / slots)
        assert n_turns * ln <= 1.05 * best * ml + n_turns * chunk, (live, ml, slots, chunk, ms, k, ln)


def test_known_cases(sched):
    # a few tiles: as many segments as the limit and the four-chunk minimum allow; the mean serving
    # request (1385 tiles, FP64 on 6144 wavefronts): 13; the largest (4000 tiles): 3; past four
    # device-fulls: one sequence
    got = sched([(1, 3600, 6144, 32, 16), (1385, 3600, 6144, 32, 16), (4000, 3700, 6144, 32, 16),
                 (30000, 3700, 6144, 32, 16), (64, 3600, 6144, 32, 1)])
    assert [g[0] for g in got] == [15, 13, 3, 1, 1]


def test_queues_partition_longest_first(sched):
    """deal_queues: every item exactly once; queue q holds items [seg[q], seg[q+1]); taken in the
    global longest-first order (32-pulse cost classes, list order within a class), runs of 8
    consecutive items go to queues 0, 1, ..., 7, 0, ... -- so each queue is itself longest first and
    every queue starts with some of the most expensive items."""
    import random
    rng = random.Random(3)
    for n in (0, 1, 7, 8, 9, 63, 64, 65, 1000, 7680):
        costs = [rng.choice((rng.randrange(0, 4000), 224, 100 + rng.randrange(7), 40000)) for _ in range(n)]
        seg, ids = sched.deal(costs)
        assert len(seg) == 9 and seg[0] == 0 and seg[-1] == n and all(a <= b for a, b in zip(seg, seg[1:]))
        assert sorted(ids) == list(range(n))
        cls = lambda c: min(1023, max(0, c) >> 5)  # noqa: E731
        order = sorted(range(n), key=lambda k: (-cls(costs[k]), k))  # longest first, stable
        want = [[] for _ in range(8)]
        for j, k in enumerate(order):
            want[(j // 8) % 8].append(k)
        got = [ids[seg[q]:seg[q + 1]] for q in range(8)]
        assert got == want, n
