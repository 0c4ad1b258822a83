// Wavefront-per-tile building blocks shared by the scene kernels (bp_f64.hip, bp_split.hip) and
// the persistent batch kernel (bp_batch.hip).
//
// A tile is up to 64 pixels of one request's F-order grid (tile_x along the fast axis); one
// wavefront owns it, lane = pixel.  The wavefront walks the union [wlo, whi) of its pixels'
// apertures pulse by pulse (the pulse index is wave-uniform), each lane adding only the pulses of
// its own aperture, in ascending order as process_point does (sar_focuser_puntual.py:134-147).
//
// FP64 terms: bp_common.h.  Split terms (FP64 per tile and pulse, FP32 per pixel and pulse):
//   per (tile row, pulse), FP64:  D = Pc - T(p), R0 = |D| for the row's reference point Pc (centre
//       of the bounding box of the row's pixels); the base phase phi(R0) = k*R0 + beta*R0^2 reduced to turns mod 1;
//       the bin offset (R0 - near)/dra split into integer B0 and fraction f0; and the FP32
//       coefficients u = D/R0, 1/(2R0), 1/(2R0^2), 1/(8R0^3), (k + 2 beta R0)/(2 pi), R0^2, 2R0.
//       Lanes stage the records of a chunk of pulses (x rows) into the wavefront's LDS.
//   per (pixel, pulse), FP32:  with o = P - Pc and a = u.o, t2 = |o|^2 - a^2,
//       dR = R - R0 = a + t2/(2R0) - a*t2/(2R0^2) [- t2(|o|^2 - 5a^2)/(8R0^3)]   (series of
//       R0*(sqrt(1+e)-1), e = (2 R0 a + |o|^2)/R0^2, to 3rd, or beyond 12 m offsets 4th, order);
//       bin = B0 + rint(f0 + dR/dra), re-decided in FP64 when within the error bound of a tie, so
//       bins and contributor counts equal the FP64 path's; phase turns = t0 + fract(c1*dR) +
//       beta/(2pi)*dR^2 into v_sin_f32 / v_cos_f32 (which take turns); weight R^2 = R0^2 + 2 R0 dR;
//       complex MAC in FP32 with Kahan compensation
//       (ACCUM 0) or FP32 sums over each 64-pulse chunk folded into FP64 (ACCUM 1).
//   A tile whose pixels lie more than SPLIT_OMAX_M from its centre (irregular pixel lists) takes the
//   FP64 terms: the series and FP32 offsets are only accurate for compact tiles.
#pragma once

#include <climits>

#include "bp_common.h"

namespace {

constexpr double INV_2PI = 0.15915494309189533577;

// Tile of one pixel grid, as seen by one lane.
struct TileLane {
  int64_t i = 0;       // flat pixel index
  bool valid = false;  // lane maps to a pixel of the grid
  double x = 0, y = 0, z = 0;
  int lo = 0x7fffffff, hi = -0x7fffffff;  // own pulses [lo, hi); empty when lo >= hi
  int len = 0;                            // hi - lo, or 0 when empty
  int wlo = 0, whi = 0;                   // union over the wavefront (uniform)
  int clo = 0, chi = 0;                   // pulses inside every non-empty aperture of the tile (uniform;
                                          // empty when clo >= chi)
};

__device__ __forceinline__ TileLane tile_lane(const BpGeom &g, const PixelsView &px, const TileGrid &tg,
                                              int64_t tile, int lane) {
  TileLane t;
  const int64_t ix = (tile % tg.ntx) * tg.tx + lane % tg.tx;
  const int64_t iy = (tile / tg.ntx) * tg.ty + lane / tg.tx;
  t.valid = ix < tg.dimx && iy < tg.dimy;
  t.i = ix + tg.dimx * iy;
  if (t.valid) {
    const int32_t s = px.sint[t.i];
    if (px.mask[t.i] == 1 && s != -9999) {  // sar_focuser_puntual.py:46-53
      int64_t l, h;
      sarbp_pulse_range(px.index[t.i], s, g.naz, l, h);
      if (h > l) {
        t.lo = (int)l;
        t.hi = (int)h;
        t.len = (int)(h - l);
      }
      t.x = px.x[t.i];
      t.y = px.y[t.i];
      t.z = px.z[t.i];
    }
  }
  t.wlo = __builtin_amdgcn_readfirstlane(wave_min_i32(t.lo));
  t.whi = __builtin_amdgcn_readfirstlane(wave_max_i32(t.hi));
  t.clo = __builtin_amdgcn_readfirstlane(wave_max_i32(t.len > 0 ? t.lo : INT_MIN));
  t.chi = __builtin_amdgcn_readfirstlane(wave_min_i32(t.len > 0 ? t.hi : INT_MAX));
  return t;
}

// FP64 terms of the narrow-range variant for pulses [p0, p1) of the union (tab: the sin/cos table),
// software-pipelined DEPTH (2 or 3) pulses deep: pulse a+DEPTH-1's range, bin and gather are issued
// before pulse a's phase, sin/cos and accumulation, so the gather's latency is covered by DEPTH-1
// pulses of arithmetic and each wave has DEPTH independent chains of FP64 work in flight.  The
// scene kernel runs three deep (2 % faster at seven wavefronts per SIMD than two deep at eight;
// measured); the batch kernels, whose register budget also holds the saved state and the other
// precision's path, two deep (three deep spills in their loops).  CORE: [p0, p1) lies inside every non-empty aperture of the tile, so a lane tests only
// that it has an aperture, not where the pulse lies in it.
template <bool CORE, int DEPTH>
__device__ __forceinline__ void f64_narrow_span(const BpGeom &g, const double2 *tab, const TileLane &t, int p0,
                                                int p1, double &ar, double &ai, int32_t &cnt) {
  if (p0 >= p1) return;
  // DEPTH register sets (A, B[, C]) rotate: while the other pulses' gathers are in flight one
  // pulse's phase and sums are computed.  Column and trajectory pointers advance by addition
  // (the device trajectory holds SARBP_TRAJ_PAD samples past the last pulse, so reading up to
  // seven pulses ahead needs no clamp; a column past the last pulse is only addressed with
  // RAW_NO_SAMPLE, which the range check turns into zero without an access).
  //
  // every lane issues its gather (a lane without a term reads RAW_NO_SAMPLE: zero, no memory
  // access), so the load sits on no branch and the wait before its use counts exactly the one
  // gather issued after it; the phase's argument reduction and its sin/cos table read are issued
  // there too, so that the LDS read is also in flight over the other pulse's work
  const bool act = t.len > 0;
  auto issue = [&](int a, const float2 *colp, double4 tr, bool &ok, double &rr, double &r2, float2 &d,
                   double2 &e) {
    const __amdgpu_buffer_rsrc_t col = raw_column_at(colp, g.nra);  // uniform
    double q = 0.0, r = 0.0;
    // the range and bin of every lane, on no branch (lanes outside their aperture are masked
    // off in the instructions either way)
    f64_range(g, t.x, t.y, t.z, tr.x, tr.y, tr.z, q, r2, r);
    // the aperture and swath tests of f64_bin as unsigned integer compares: q is finite here
    // (narrow ranges), and a q beyond the int32 range converts to INT_MIN or INT_MAX, which
    // lies outside [bin_lo, nra-1) either way (nra <= 2^24)
    const int iq = (int)q;
    const bool in_ap = CORE ? act : (unsigned)(a - t.lo) < (unsigned)t.len;
    ok = a < p1 && in_ap && (unsigned)(iq - g.bin_lo) < (unsigned)(g.nra - 1 - g.bin_lo);
    d = raw_load(col, ok ? (unsigned)iq << 3 : RAW_NO_SAMPLE);
    double ph;
    {
#pragma clang fp contract(off)
      ph = g.kphase * r + g.beta * r2;  // f64_phase's argument
    }
    sincos_reduce(ph, tab, e, rr);  // |ph| < 2^24 (narrow ranges); any table index is in bounds
  };
  auto consume = [&](bool ok, double rr, double r2, float2 d, double2 e) {
    if (ok) {
      double sn, cs;
      sincos_finish(e, rr, sn, cs);
      f64_acc(r2, sn, cs, d, ar, ai);
      ++cnt;
    }
  };
  const int64_t cstride = g.nra;
  const float2 *colA = g.raw + (int64_t)p0 * cstride;  // pulse a's column (set A)
  const double4 *tp = g.traj + p0;                     // tp[k]: pulse a + k
  bool okA, okB, okC;
  double rA = 0.0, r2A = 0.0, rB = 0.0, r2B = 0.0, rC = 0.0, r2C = 0.0;
  float2 dA, dB, dC;
  double2 eA, eB, eC;
  if constexpr (DEPTH == 2) {
    issue(p0, colA, tp[0], okA, rA, r2A, dA, eA);
    double4 t1 = tp[1], t2 = tp[2];
    for (int a = p0; a < p1; a += 2) {  // wave-uniform pulses a (set A) and a+1 (set B)
      issue(a + 1, colA + cstride, t1, okB, rB, r2B, dB, eB);
      const double4 t3 = tp[3];
      consume(okA, rA, r2A, dA, eA);
      colA += 2 * cstride;
      issue(a + 2, colA, t2, okA, rA, r2A, dA, eA);
      const double4 t4 = tp[4];
      consume(okB, rB, r2B, dB, eB);
      tp += 2;
      t1 = t3;
      t2 = t4;
    }
  } else {
    issue(p0, colA, tp[0], okA, rA, r2A, dA, eA);
    issue(p0 + 1, colA + cstride, tp[1], okB, rB, r2B, dB, eB);
    double4 t2 = tp[2], t3 = tp[3], t4 = tp[4];
    for (int a = p0; a < p1; a += 3) {  // wave-uniform pulses a, a+1, a+2 (sets A, B, C)
      issue(a + 2, colA + 2 * cstride, t2, okC, rC, r2C, dC, eC);
      const double4 t5 = tp[5];
      consume(okA, rA, r2A, dA, eA);
      issue(a + 3, colA + 3 * cstride, t3, okA, rA, r2A, dA, eA);
      const double4 t6 = tp[6];
      consume(okB, rB, r2B, dB, eB);
      issue(a + 4, colA + 4 * cstride, t4, okB, rB, r2B, dB, eB);
      const double4 t7 = tp[7];
      consume(okC, rC, r2C, dC, eC);
      colA += 3 * cstride;
      tp += 3;
      t2 = t5;
      t3 = t6;
      t4 = t7;
    }
  }
}

// FP64 terms for pulses [p0, p1) of the union (tab: the sin/cos table), in ascending pulse order.
// The trajectory sample is loaded one pulse ahead (scalar loads), and the RawData gather is issued
// as soon as the bin is known, before the phase and its sin/cos, which cover part of its latency.
// WIDE = false (the narrow-range variant, which has registers to spare) takes f64_narrow_span, the
// pulses inside every aperture of the tile without the per-lane aperture test; the terms and their
// order into the sums are the same.
template <bool WIDE = true, int DEPTH = 2>
__device__ __forceinline__ void f64_span(const BpGeom &g, const double2 *tab, const TileLane &t, int p0, int p1,
                                         double &ar, double &ai, int32_t &cnt) {
  if (p0 >= p1) return;
  if constexpr (!WIDE) {
    const int c0 = min(max(p0, t.clo), p1), c1 = max(min(p1, t.chi), c0);
    f64_narrow_span<false, DEPTH>(g, tab, t, p0, c0, ar, ai, cnt);
    f64_narrow_span<true, DEPTH>(g, tab, t, c0, c1, ar, ai, cnt);
    f64_narrow_span<false, DEPTH>(g, tab, t, c1, p1, ar, ai, cnt);
    return;
  }
  double4 tr = g.traj[p0];
  for (int a = p0; a < p1; ++a) {  // wave-uniform pulse
    const double4 trn = g.traj[a + 1 < p1 ? a + 1 : a];
    const __amdgpu_buffer_rsrc_t col = raw_column(g, a);
    double q, r2, r;
    if (a >= t.lo && a < t.hi && f64_bin<WIDE>(g, t.x, t.y, t.z, tr.x, tr.y, tr.z, q, r2, r)) {
      const float2 d = raw_load(col, (unsigned)(int)q << 3);
      double sn, cs;
      f64_phase<WIDE>(g, tab, r, r2, sn, cs);
      f64_acc(r2, sn, cs, d, ar, ai);
      ++cnt;
    }
    tr = trn;
  }
}

// 1.5 * 2^23: adding it rounds an FP32 value of magnitude below 2^22 to an integer.
constexpr float SPLIT_ROUND_MAGIC = 12582912.0f;

struct PulseRec {  // one pulse's tile-relative coefficients (3 x 16 B LDS reads)
  float4 a;        // ux, uy, uz, 1/(2R0)                                          (range offset)
  float4 b;        // 1/(2R0^2), f0 = q0 - B0, int bits: byte offset of bin B0 less 8 bits(magic),
                   // 1/(8R0^3)                                                      (bin)
  float4 c;        // t0 = frac(phi0/2pi), c1 = (k + 2 beta R0)/2pi, R0^2, 2R0       (phase and weight)
};

// Dynamic LDS of every kernel that stages split records (indexed, so addressing stays 32-bit).
extern __shared__ PulseRec sarbp_lds_rec[];

__device__ __forceinline__ int make_rec(const BpGeom &g, double xc, double yc, double zc, double4 t,
                                        PulseRec &r) {  // returns B0 (absolute bin)
  const double dx = xc - t.x, dy = yc - t.y, dz = zc - t.z;
  const double r0sq = dx * dx + dy * dy + dz * dz;
  // 1/R0 from the hardware reciprocal square root and one Newton step (relative error about 1.5 e^2
  // for an estimate within e; below 1e-13 for any e under 2^-22: R0 is only the expansion point, so
  // it needs no correct rounding -- the pixel ranges are R0 + dR exactly, and an error of 1e-13 R0
  // moves the reference bin offset by far less than the FP32 error the near-tie guard allows for)
  double inv = __builtin_amdgcn_rsq(r0sq);
  const double h = 0.5 * r0sq;
  inv = fma(inv, fma(-h * inv, inv, 0.5), inv);
  const double r0 = r0sq * inv;
  // bin offset, kept within +-2^25 bins so that 8 x bin stays far from int32 wrap-around (a
  // reference point that far outside the swath has no bin inside it either)
  const double q0 = fmin(fmax((r0 - g.near_range) * g.inv_dra, -33554432.0), 33554432.0);
  const double b0 = __builtin_rint(q0);
  const double ph0 = g.kphase * r0 + g.beta * r0sq;
  const double turns = ph0 * INV_2PI;
  const double t0 = turns - __builtin_floor(turns);
  const double c1 = (g.kphase + 2.0 * g.beta * r0) * INV_2PI;
  r.a = make_float4((float)(dx * inv), (float)(dy * inv), (float)(dz * inv), (float)(0.5 * inv));
  // byte offset of bin B0 from bin bin_lo, where the split path's column resources start
  r.b = make_float4((float)(0.5 * inv * inv), (float)(q0 - b0),
                    __uint_as_float((unsigned)((int)b0 - g.bin_lo) * 8u - 8u * (unsigned)__float_as_int(SPLIT_ROUND_MAGIC)),
                    (float)(0.125 * inv * inv * inv));
  r.c = make_float4((float)t0, (float)c1, (float)r0sq, (float)(2.0 * r0));
  return (int)b0;
}

// Row-relative quantities of the split form.  Each row of the tile (TX lanes at one grid y) has
// its own reference point, the centre of the bounding box of its focusing pixels, so the FP32
// offsets stay within half a tile width along the fast axis; records are staged per (pulse, row).
struct SplitTile {
  double sxc, syc, szc;  // reference point of the row this lane stages records for (row lane % TY)
  float ox, oy, oz, qo;  // offset of this lane's pixel from its row's reference point
  float omax;            // largest pixel-to-reference distance in the tile (uniform)
  bool use;              // the tile is compact enough for the split form (uniform)
  int core_lo, core_hi;  // pulses inside every non-empty aperture of the tile (uniform; may be empty)
  int margin;            // bins: a reference bin this far inside [0, Nra-1) keeps every pixel's bin inside
  float guard;           // |f0 + dR/dra - nearest integer| above this -> decide the bin in FP64
  int ty_shift;          // log2(rows per tile)
  int row;               // this lane's row
};

__device__ __forceinline__ double row_min_f64(double v, int tx) {
  for (int m = tx >> 1; m >= 1; m >>= 1) v = fmin(v, __shfl_xor(v, m, 64));
  return v;
}
__device__ __forceinline__ double row_max_f64(double v, int tx) {
  for (int m = tx >> 1; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
  return v;
}

// r0min: lower bound of the tile's pixel-to-sensor distances (PixelsView::r0min).
__device__ __forceinline__ SplitTile split_tile(const BpGeom &g, const TileGrid &tg, const TileLane &t, float r0min) {
  SplitTile s;
  const int lane = threadIdx.x & 63;
  const bool act = t.hi > t.lo;
  const double inf = __builtin_huge_val();
  const double lx = row_min_f64(act ? t.x : inf, tg.tx), hx = row_max_f64(act ? t.x : -inf, tg.tx);
  const double ly = row_min_f64(act ? t.y : inf, tg.tx), hy = row_max_f64(act ? t.y : -inf, tg.tx);
  const double lz = row_min_f64(act ? t.z : inf, tg.tx), hz = row_max_f64(act ? t.z : -inf, tg.tx);
  const bool row_empty = !(lx <= hx);
  const double xc = row_empty ? 0.0 : 0.5 * (lx + hx);  // rows without pixels: any finite point
  const double yc = row_empty ? 0.0 : 0.5 * (ly + hy);
  const double zc = row_empty ? 0.0 : 0.5 * (lz + hz);
  s.row = lane / tg.tx;
  s.ty_shift = __builtin_ctz((unsigned)tg.ty);
  const int src = (lane & (tg.ty - 1)) * tg.tx;  // first lane of the row this lane stages for
  s.sxc = __shfl(xc, src, 64);
  s.syc = __shfl(yc, src, 64);
  s.szc = __shfl(zc, src, 64);
  s.ox = act ? (float)(t.x - xc) : 0.f;
  s.oy = act ? (float)(t.y - yc) : 0.f;
  s.oz = act ? (float)(t.z - zc) : 0.f;
  s.qo = s.ox * s.ox + s.oy * s.oy + s.oz * s.oz;
  s.omax = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(sqrtf((float)wave_max_f64((double)s.qo)))));
  // error of f0 + dR/dra in bins: f0 and the FP32 sum ~1e-7, dR ~1.2e-7 * |o| m; margin >= 3x; plus
  // the series truncation
  const double trunc = split_truncation_m((double)s.omax, (double)r0min);
  s.guard = 0.5f - (2e-5f + 4e-7f * s.omax * (float)g.inv_dra) - (float)(fmin(trunc, 1.0) * g.inv_dra);
  // compact tiles only, whose series truncation stays far below the FP32 rounding; and |dR|/dra below
  // 2^21 bins, so that |f0 + dR/dra| < 2^22 for the rounding of split_bin
  s.use = s.omax <= SPLIT_OMAX_M && (double)s.omax * g.inv_dra < 2097152.0 &&
          trunc * __builtin_fabs(g.kphase) <= SPLIT_TRUNC_MAX_RAD;
  s.core_lo = t.clo;
  s.core_hi = t.chi;
  // |bin - B0| <= |f0| + |dR|/dra + 1 (an exact re-decision) with |dR| <= 1.01 |o|
  s.margin = s.use ? (int)((double)s.omax * g.inv_dra * 1.25) + 4 : INT_MAX;
  return s;
}

// Second half of one split term: phase, sin/cos and the weighted term R^2 d exp(-j phi), added to
// acc by four fused multiply-adds (adds nothing when the pulse does not contribute: its sample d is
// zero).  Computed on every lane: no branch in the hot loop.  The phase and weight coefficients are
// read from the record here, after the gather, so that only dR and the sample stay live while the
// gather is in flight.
__device__ __forceinline__ void split_mac(float dR, float2 d, const PulseRec *rk, float beta_t, float2 &acc) {
  const float4 rc = rk->c;
  const float t0 = rc.x, c1 = rc.y, r0sq = rc.z, two_r0 = rc.w;
  // the phase in turns, t0 + dR (c1 + beta' dR): the reference's (t0, in [0, 1)) plus the linear
  // and quadratic terms in two fused multiply-adds, reduced exactly.  The inner sum rounds to the
  // precision of c1 itself (FP32, as stored), the outer one to at most 2^-24 of |dR (c1 + beta' dR)|
  // (< ~100 turns: 4e-6 turns, 2.4e-5 rad); both vary from pulse to pulse
  const float tt = __builtin_amdgcn_fractf(fmaf(fmaf(beta_t, dR, c1), dR, t0));
  const float sn = __builtin_amdgcn_sinf(tt);  // sin(2 pi tt)
  const float cs = __builtin_amdgcn_cosf(tt);
  const float w = fmaf(two_r0, dR, r0sq);  // R^2 = R0^2 + 2 R0 dR (dR^2 / R0^2 < 1e-7: below FP32's resolution of R^2)
  const float wr = w * d.x, wi = w * d.y;
  acc.x = fmaf(wi, sn, fmaf(wr, cs, acc.x));   // + Re(w d) cos + Im(w d) sin
  acc.y = fmaf(-wr, sn, fmaf(wi, cs, acc.y));  // + Im(w d) cos - Re(w d) sin
}

// Compensated accumulation of one (pair) sum: Kahan (ACCUM 0) or plain FP32 within a chunk (ACCUM 1).
template <int ACCUM>
__device__ __forceinline__ void split_acc(float2 t, float &sr, float &si, float &cr, float &ci) {
  if (ACCUM == 0) {
    const float yr = t.x - cr, yi = t.y - ci;
    const float nr = sr + yr, ni = si + yi;
    cr = (nr - sr) - yr;
    ci = (ni - si) - yi;
    sr = nr;
    si = ni;
  } else {
    sr += t.x;
    si += t.y;
  }
}

// First half of one split term for pulse a (record rk): range offset, range bin, and the RawData
// gather.  The gather goes through a buffer resource of pulse a's column (base and size in scalar
// registers) with a per-lane byte offset; a lane whose pulse does not count (outside its aperture,
// or bin outside [0, Nra-1)) passes an offset beyond the column, which the buffer range check turns
// into a zero sample without a memory access.
struct SplitHalf {
  bool bad;  // inside the aperture but the bin is outside [0, Nra-1): not counted (rare)
  float dR;
  float2 d;
};

constexpr unsigned SPLIT_NO_SAMPLE = RAW_NO_SAMPLE;


// Range offset, bin and byte offset of one lane's pixel for the pulse of record rk (no
// aperture or swath test); fr = |qb - rint(qb)|: above s.guard the bin lies within the FP32 error
// bound of a rounding tie and must be decided exactly (exact_off).
template <bool FOURTH>
__device__ __forceinline__ void split_bin(const SplitTile &s, const PulseRec *rk, float inv_dra, SplitHalf &h,
                                          unsigned &off, float &fr) {
  const float4 ra = rk->a, rb = rk->b;
  // computed on every lane (lanes outside their aperture have zero offsets: finite, unused)
  const float av = fmaf(ra.z, s.oz, fmaf(ra.y, s.oy, ra.x * s.ox));
  const float t2 = fmaf(-av, av, s.qo);
  // dR = a + t2 (1/(2R0) - a/(2R0^2) [- (|o|^2 - 5a^2)/(8R0^3)])
  float c = fmaf(-av, rb.x, ra.w);
  if (FOURTH) c = fmaf(-fmaf(-5.f * av, av, s.qo), rb.w, c);
  const float dR = fmaf(t2, c, av);
  h.dR = dR;
  const float qb = fmaf(dR, inv_dra, rb.y);
  // rint(qb) for |qb| < 2^22: qb + 1.5 * 2^23 lies in [2^23, 2^24), where the sum's rounding (to
  // nearest, ties to even) keeps the integer part and the low mantissa bits hold rint(qb) +
  // 0x400000; both differences below are exact
  const float qm = qb + SPLIT_ROUND_MAGIC;
  const float rq = qm - SPLIT_ROUND_MAGIC;
  // byte offset of bin B0 + rq: rb.z holds 8 (B0 - bin_lo) - 8 bits(SPLIT_ROUND_MAGIC), modulo 2^32
  off = ((unsigned)__float_as_int(qm) << 3) + (unsigned)__float_as_int(rb.z);
  fr = fabsf(qb - rq);
}

// The byte offset of pulse a's bin decided in FP64 (the path of f64_bin), from bin_lo; -1 - bin_lo
// (far beyond the column) when the bin is outside [bin_lo, nra-1).
__device__ __forceinline__ unsigned exact_off(const BpGeom &g, const TileLane &t, int a) {
  const double4 tr = g.traj[a];
  const int eb = exact_bin(g, t.x, t.y, t.z, tr.x, tr.y, tr.z);  // -1: outside [bin_lo, nra-1)
  return (unsigned)(eb - g.bin_lo) << 3;
}

// One pulse of the general form: aperture and swath tests per lane, near-tie bins decided
// exactly, then the gather.
template <bool FOURTH>
__device__ __forceinline__ void split_first(const BpGeom &g, const TileLane &t, const SplitTile &s,
                                            const PulseRec *rk, int a, __amdgpu_buffer_rsrc_t col,
                                            float inv_dra, unsigned nbin_bytes, SplitHalf &h) {
  unsigned off;
  float fr;
  split_bin<FOURTH>(s, rk, inv_dra, h, off, fr);
  const bool act = a >= t.lo && a < t.hi;
  if (act && fr > s.guard) off = exact_off(g, t, a);  // near a rounding tie (rare)
  const bool ok = off < nbin_bytes;  // bin_lo <= bin < Nra-1
  h.bad = act && !ok;
  h.d = raw_load(col, act && ok ? off : SPLIT_NO_SAMPLE);
}

// Pulses are processed four at a time: the four gathers are in flight before any term is formed;
// the terms are added to acc in ascending pulse order.  Groups start at multiples of 4 from the
// chunk start, and chunks and caller ranges start at multiples of the chunk length (a multiple of
// 8) from the union start, so the grouping -- and the result -- does not depend on how a caller
// splits the pulse range.
//
// FAST: every lane with an aperture is inside it for the four pulses and every bin is inside the
// swath (the caller checks both), so no lane tests either; a lane without an aperture reads a sample
// at its row's reference point (its offsets are zero), in range or zero, and is never counted or
// stored (its count is zero, so its mean is written as 0).  Otherwise each lane tests its aperture
// and its bin against the swath, and a bin outside the swath inside the aperture is not counted.
// Either way the rare near-tie bins of the four pulses are decided on one branch.
template <bool FOURTH, bool FAST>
__device__ __forceinline__ void split_group(const BpGeom &g, const TileLane &t, const SplitTile &s,
                                            const PulseRec *rk, int a, const float2 *col, float inv_dra,
                                            unsigned nbin_bytes, float beta_t, float2 &acc, int32_t &cnt) {
  // col: bin bin_lo of the column of pulse a; one resource for the four columns, pulse k at the
  // scalar offset k x cs (every offset passed below is a bin inside the swath or beyond every column)
  const __amdgpu_buffer_rsrc_t cr = raw_split_group(col, g.nra, nbin_bytes);
  const unsigned cs = (unsigned)g.nra * 8u;
  SplitHalf X, Y, Z, W;
  unsigned oX, oY, oZ, oW;
  float fX, fY, fZ, fW;
  split_bin<FOURTH>(s, rk, inv_dra, X, oX, fX);
  split_bin<FOURTH>(s, rk + 1, inv_dra, Y, oY, fY);
  split_bin<FOURTH>(s, rk + 2, inv_dra, Z, oZ, fZ);
  split_bin<FOURTH>(s, rk + 3, inv_dra, W, oW, fW);
  // inside the lane's aperture (unsigned compare: an empty aperture has len 0)
  const bool aX = (unsigned)a - (unsigned)t.lo < (unsigned)t.len, aY = (unsigned)(a + 1) - (unsigned)t.lo < (unsigned)t.len,
             aZ = (unsigned)(a + 2) - (unsigned)t.lo < (unsigned)t.len, aW = (unsigned)(a + 3) - (unsigned)t.lo < (unsigned)t.len;
  if (fmaxf(fmaxf(fX, fY), fmaxf(fZ, fW)) > s.guard) {
    __asm__ volatile("");  // keeps the branch: near-tie bins are rare
    if (!FAST) {  // only lanes within their aperture re-decide
      if (aX && fX > s.guard) oX = exact_off(g, t, a);
      if (aY && fY > s.guard) oY = exact_off(g, t, a + 1);
      if (aZ && fZ > s.guard) oZ = exact_off(g, t, a + 2);
      if (aW && fW > s.guard) oW = exact_off(g, t, a + 3);
    } else {
      if (fX > s.guard) oX = exact_off(g, t, a);
      if (fY > s.guard) oY = exact_off(g, t, a + 1);
      if (fZ > s.guard) oZ = exact_off(g, t, a + 2);
      if (fW > s.guard) oW = exact_off(g, t, a + 3);
    }
  }
  X.bad = Y.bad = Z.bad = W.bad = false;
  if (FAST) {
    X.d = raw_load_at(cr, oX, 0u);
    Y.d = raw_load_at(cr, oY, cs);
    Z.d = raw_load_at(cr, oZ, 2u * cs);
    W.d = raw_load_at(cr, oW, 3u * cs);
  } else {
    const bool kX = oX < nbin_bytes, kY = oY < nbin_bytes, kZ = oZ < nbin_bytes, kW = oW < nbin_bytes;
    X.bad = aX && !kX;  // bin_lo <= bin < Nra-1 fails inside the aperture: not counted
    Y.bad = aY && !kY;
    Z.bad = aZ && !kZ;
    W.bad = aW && !kW;
    X.d = raw_load_at(cr, aX && kX ? oX : SPLIT_NO_SAMPLE, 0u);
    Y.d = raw_load_at(cr, aY && kY ? oY : SPLIT_NO_SAMPLE, cs);
    Z.d = raw_load_at(cr, aZ && kZ ? oZ : SPLIT_NO_SAMPLE, 2u * cs);
    W.d = raw_load_at(cr, aW && kW ? oW : SPLIT_NO_SAMPLE, 3u * cs);
  }
  split_mac(X.dR, X.d, rk, beta_t, acc);
  split_mac(Y.dR, Y.d, rk + 1, beta_t, acc);
  split_mac(Z.dR, Z.d, rk + 2, beta_t, acc);
  split_mac(W.dR, W.d, rk + 3, beta_t, acc);
  if (!FAST && (X.bad || Y.bad || Z.bad || W.bad)) {  // rare: a bin outside the swath
    __asm__ volatile("");  // keeps the branch (the count is not worth computing on every group)
    cnt -= (X.bad ? 1 : 0) + (Y.bad ? 1 : 0) + (Z.bad ? 1 : 0) + (W.bad ? 1 : 0);
  }
}

// Split terms for pulses [c0, p1), c0 on the chunk grid of the union (wlo + C*k), so chunk sums
// and staged records do not depend on where a caller splits the range.  The wavefront's
// TY * (C + 1) LDS records start at rec0: record rec0 + r*(C+1) + k for pulse c0+k and row r, so a
// lane reads its row's consecutive pulses at constant offsets, and the four rows' records of one
// pulse (C + 1 records = 12 (C + 1) dwords apart) fall in different LDS banks (lane l stages the
// records of pulses l / TY + 64/TY * m of row l % TY).  Sums: Kahan (sr, si, cr, ci) for ACCUM 0; for ACCUM 1
// (sr, si) hold the running chunk sum, folded into (ar, ai) at each chunk end.  Groups of four
// pulses inside every aperture of the tile, in a chunk whose reference bins all lie well inside
// the swath, take the fast form; both forms give the same terms.
template <int ACCUM, bool FOURTH>
__device__ __forceinline__ void split_span_t(const BpGeom &g, const TileLane &t, const SplitTile &s, int c0,
                                             int p1, int rec0, int C, float &sr, float &si, float &cr,
                                             float &ci, double &ar, double &ai, int32_t &cnt) {
  PulseRec *const rec = sarbp_lds_rec;
  const int lane = threadIdx.x & 63;
  const float inv_dra = (float)g.inv_dra;
  const float beta_t = (float)(g.beta * INV_2PI);
  const unsigned nbin_bytes = (unsigned)(g.nra - 1 - g.bin_lo) * 8u;
  const int nrec = C << s.ty_shift;
  const int rows_mask = (1 << s.ty_shift) - 1;
  const int rs = C + 1;  // row stride in records
  const int blo = g.bin_lo + s.margin, bhi = (int)(g.nra - 1) - 1 - s.margin;  // reference bins inside the swath
  // contributors: the pulses of this lane's aperture within [c0, p1), less the rare ones whose bin
  // falls outside the swath (counted as they occur)
  {
    const int lo = max(c0, t.lo), hi = min(p1, t.hi);  // compared, not subtracted: an empty aperture is
    cnt += hi > lo ? hi - lo : 0;                       // (INT_MAX, -INT_MAX)
  }
  // lane l stages the records m = l + 64 j: pulse c0 + (m >> ty_shift) of row m & rows_mask, so
  // from one j to the next the pulse advances by 64 >> ty_shift and the row stays the same
  const int pstep = 64 >> s.ty_shift;
  PulseRec *const srec = rec + rec0 + (lane & rows_mask) * rs + (lane >> s.ty_shift);
  for (; c0 < p1; c0 += C) {
    bool inside = true;
    {
      PulseRec *sr = srec;
      const double4 *tp = g.traj + c0 + (lane >> s.ty_shift);
      for (int m = lane, p = c0 + (lane >> s.ty_shift); m < nrec; m += 64, p += pstep, sr += pstep, tp += pstep) {
        if (p < p1) {
          const int b0 = make_rec(g, s.sxc, s.syc, s.szc, *tp, *sr);
          inside = inside && b0 >= blo && b0 <= bhi;
        }
      }
    }
    const bool swath = __builtin_amdgcn_ballot_w64(!inside) == 0;  // uniform
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int cn = min(C, p1 - c0);
    const PulseRec *rk = rec + rec0 + s.row * rs;
    const float2 *col = g.raw + (int64_t)c0 * g.nra + g.bin_lo;  // bin bin_lo of the column of pulse c0 + k
    // one group of four pulses in the fast or the general form, added to acc
    auto group = [&](int a, const PulseRec *rg, const float2 *cg, float2 &acc) {
      if (swath && a >= s.core_lo && a + 3 < s.core_hi)
        split_group<FOURTH, true>(g, t, s, rg, a, cg, inv_dra, nbin_bytes, beta_t, acc, cnt);
      else
        split_group<FOURTH, false>(g, t, s, rg, a, cg, inv_dra, nbin_bytes, beta_t, acc, cnt);
    };
    int k = 0;
    // eight pulses per compensated step: the terms of two groups (a multiple of 8 from the chunk
    // start) are summed by fused multiply-adds in pulse order, then the sum is accumulated
    for (; k + 7 < cn; k += 8, rk += 8, col += 8 * g.nra) {
      float2 acc = make_float2(0.f, 0.f);
      group(c0 + k, rk, col, acc);
      group(c0 + k + 4, rk + 4, col + 4 * g.nra, acc);
      split_acc<ACCUM>(acc, sr, si, cr, ci);
    }
    if (k + 3 < cn) {  // a last group of four (only at the end of the pulse range)
      float2 acc = make_float2(0.f, 0.f);
      group(c0 + k, rk, col, acc);
      split_acc<ACCUM>(acc, sr, si, cr, ci);
      k += 4;
      rk += 4;
      col += 4 * g.nra;
    }
    for (; k < cn; ++k, ++rk, col += g.nra) {  // tail (only at the end of the pulse range)
      SplitHalf X;
      split_first<FOURTH>(g, t, s, rk, c0 + k, raw_split_column(col, nbin_bytes), inv_dra, nbin_bytes, X);
      float2 acc = make_float2(0.f, 0.f);
      split_mac(X.dR, X.d, rk, beta_t, acc);
      split_acc<ACCUM>(acc, sr, si, cr, ci);
      cnt -= X.bad ? 1 : 0;
    }
    if (ACCUM == 1) {
      ar += (double)sr;
      ai += (double)si;
      sr = 0.f;
      si = 0.f;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();  // reads of this chunk finish before the next overwrites rec
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// rec0: first LDS record (of sarbp_lds_rec) owned by this wavefront; it owns C * TY records.
template <int ACCUM>
__device__ __forceinline__ void split_span(const BpGeom &g, const TileLane &t, const SplitTile &s, int c0, int p1,
                                           int rec0, int C, float &sr, float &si, float &cr, float &ci,
                                           double &ar, double &ai, int32_t &cnt) {
  if (s.omax > SPLIT_FOURTH_M)
    split_span_t<ACCUM, true>(g, t, s, c0, p1, rec0, C, sr, si, cr, ci, ar, ai, cnt);
  else
    split_span_t<ACCUM, false>(g, t, s, c0, p1, rec0, C, sr, si, cr, ci, ar, ai, cnt);
}

// Final sums of the split form as FP64.
template <int ACCUM>
__device__ __forceinline__ void split_total(float sr, float si, float cr, float ci, double &ar, double &ai) {
  if (ACCUM == 0) {
    ar = (double)sr - (double)cr;
    ai = (double)si - (double)ci;
  }
}

}  // namespace
