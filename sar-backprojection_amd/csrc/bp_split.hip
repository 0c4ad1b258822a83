// Split-precision back-projection for gfx950.
//
// Same per-pixel result as process_point (sar_focuser_puntual.py:21-155) -- same aperture, same
// range bins, same contributor counts -- with FP64 kept only where the absolute slant range and
// phase need it:
//
//   per (tile, pulse), FP64:  D = Pc - T(p), R0 = |D| for the tile's reference point Pc (the centre
//       of its pixels' bounding box); the base phase phi(R0) = k*R0 + beta*R0^2 reduced to turns
//       mod 1; the bin offset (R0 - near)/dra split into an integer and a fraction; and the FP32
//       coefficients u = D/R0, 1/(2R0), 1/(2R0^2), 1/(8R0^3), (k + 2 beta R0)/(2 pi), R0^2, 2R0.
//       Lane l of the wavefront computes pulse c0+l of a 64-pulse chunk into LDS.
//   per (pixel, pulse), FP32:  with o = P - Pc (per pixel, FP32) and a = u.o, t2 = |o|^2 - a^2,
//       dR = R - R0 = a + t2/(2R0) - a*t2/(2R0^2) [- t2(|o|^2 - 5a^2)/(8R0^3)]   (series of
//       R0*(sqrt(1+e)-1), e = (2 R0 a + |o|^2)/R0^2, to 3rd or 4th order in |o|/R0);
//       bin = B0 + rint(f0 + dR/dra), re-decided in FP64 when within the error bound of a tie;
//       phase turns = t0 + fract(c1*dR) + (exact product residual) + beta/(2pi)*dR^2, then the
//       hardware v_sin_f32 / v_cos_f32 (which take turns); weight R^2 = R0^2 + dR*(2R0 + dR);
//       complex MAC in FP32 with Kahan-compensated accumulation (accum = 0) or FP32 blocks of one
//       64-pulse chunk folded into FP64 (accum = 1).
// A tile whose pixels lie more than SPLIT_OMAX_M from its centre (irregular pixel lists) is
// focused with the FP64 term instead: the series and FP32 offsets are only accurate for compact
// tiles.
#include "bp_common.h"

namespace {

constexpr float SPLIT_OMAX_M = 48.0f;     // largest pixel-to-centre offset focused in split form
constexpr float SPLIT_FOURTH_M = 12.0f;   // include the 4th-order range term beyond this offset
constexpr double INV_2PI = 0.15915494309189533577;

struct PulseRec {  // one pulse's tile-relative coefficients, 12 floats (3 x 16 B LDS reads)
  float4 a;        // ux, uy, uz, 1/(2R0)
  float4 b;        // 1/(2R0^2), f0 = q0 - B0, B0 (int bits), t0 = frac(phi0/2pi)
  float4 c;        // c1 = (k + 2 beta R0)/2pi, R0^2, 2R0, 1/(8R0^3)
};

__device__ __forceinline__ void make_rec(const BpGeom &g, double xc, double yc, double zc, double4 t,
                                         PulseRec &r) {
  const double dx = xc - t.x, dy = yc - t.y, dz = zc - t.z;
  const double r0sq = dx * dx + dy * dy + dz * dz;
  const double r0 = __builtin_sqrt(r0sq);
  const double inv = 1.0 / r0;
  const double q0 = (r0 - g.near_range) * g.inv_dra;
  const double b0 = __builtin_rint(q0);
  const double ph0 = g.kphase * r0 + g.beta * r0sq;
  const double turns = ph0 * INV_2PI;
  const double t0 = turns - __builtin_floor(turns);
  const double c1 = (g.kphase + 2.0 * g.beta * r0) * INV_2PI;
  r.a = make_float4((float)(dx * inv), (float)(dy * inv), (float)(dz * inv), (float)(0.5 * inv));
  r.b = make_float4((float)(0.5 * inv * inv), (float)(q0 - b0), __int_as_float((int)b0), (float)t0);
  r.c = make_float4((float)c1, (float)r0sq, (float)(2.0 * r0), (float)(0.125 * inv * inv * inv));
}

template <int ACCUM>
__global__ __launch_bounds__(64) void k_bp_split_tile(BpGeom g, PixelsView px, TileGrid tg,
                                                      double2 *__restrict__ out, int32_t *__restrict__ ncnt) {
  __shared__ PulseRec rec[64];
  const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
  if (tile >= tg.ntiles()) return;  // whole wavefront
  const int lane = threadIdx.x;
  const int64_t ix = (tile % tg.ntx) * tg.tx + lane % tg.tx;
  const int64_t iy = (tile / tg.ntx) * tg.ty + lane / tg.tx;
  const bool valid = ix < tg.dimx && iy < tg.dimy;
  const int64_t i = ix + tg.dimx * iy;
  double x = 0, y = 0, z = 0;
  int lo = 0x7fffffff, hi = -0x7fffffff;
  if (valid) {
    const int32_t s = px.sint[i];
    if (px.mask[i] == 1 && s != -9999) {  // sar_focuser_puntual.py:46-53
      int64_t l, h;
      sarbp_pulse_range(px.index[i], s, g.naz, l, h);
      if (h > l) {
        lo = (int)l;
        hi = (int)h;
      }
      x = px.x[i];
      y = px.y[i];
      z = px.z[i];
    }
  }
  const bool act = hi > lo;
  const int wlo = __builtin_amdgcn_readfirstlane(wave_min_i32(lo));
  const int whi = __builtin_amdgcn_readfirstlane(wave_max_i32(hi));
  if (wlo >= whi) {  // nothing to focus in this tile
    if (valid) store_mean(out, ncnt, i, 0.0, 0.0, 0);
    return;
  }
  const double inf = __builtin_huge_val();
  const double xc = 0.5 * (wave_min_f64(act ? x : inf) + wave_max_f64(act ? x : -inf));
  const double yc = 0.5 * (wave_min_f64(act ? y : inf) + wave_max_f64(act ? y : -inf));
  const double zc = 0.5 * (wave_min_f64(act ? z : inf) + wave_max_f64(act ? z : -inf));
  const float ox = act ? (float)(x - xc) : 0.f;
  const float oy = act ? (float)(y - yc) : 0.f;
  const float oz = act ? (float)(z - zc) : 0.f;
  const float qo = ox * ox + oy * oy + oz * oz;
  const float omax =
      __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(sqrtf((float)wave_max_f64((double)qo)))));

  if (omax > SPLIT_OMAX_M) {  // not a compact tile: FP64 terms, ascending pulses
    double acc_re = 0.0, acc_im = 0.0;
    int32_t cnt = 0;
    for (int a = wlo; a < whi; ++a) {
      const double4 t = g.traj[a];
      if (a >= lo && a < hi && f64_term(g, x, y, z, t.x, t.y, t.z, g.raw + g.nra * (int64_t)a, acc_re, acc_im))
        ++cnt;
    }
    if (valid) store_mean(out, ncnt, i, acc_re, acc_im, cnt);
    return;
  }

  const bool fourth = omax > SPLIT_FOURTH_M;
  const float inv_dra = (float)g.inv_dra;
  // |error of f0 + dR/dra| stays far below this; closer to a half-integer -> decide in FP64
  const float guard = 0.5f - (2e-5f + 4e-7f * omax * inv_dra);
  const float beta_t = (float)(g.beta * INV_2PI);
  const unsigned nbins = (unsigned)(g.nra - 1);

  float sr = 0.f, si = 0.f, cr = 0.f, ci = 0.f;  // Kahan sums (ACCUM 0) or chunk block sums (1)
  double ar = 0.0, ai = 0.0;                     // FP64 fold (ACCUM 1)
  int32_t cnt = 0;

  for (int c0 = wlo; c0 < whi; c0 += 64) {
    const int p = c0 + lane;
    if (p < whi) make_rec(g, xc, yc, zc, g.traj[p], rec[lane]);
    __syncthreads();  // block == one wavefront: orders the LDS writes before the reads
    const int cn = min(64, whi - c0);
    for (int k = 0; k < cn; ++k) {
      const int a = c0 + k;
      const float4 ra = rec[k].a, rb = rec[k].b, rc = rec[k].c;
      if (a >= lo && a < hi) {
        const float av = fmaf(ra.z, oz, fmaf(ra.y, oy, ra.x * ox));
        const float t2 = fmaf(-av, av, qo);
        float dR = fmaf(t2, ra.w, av);
        dR = fmaf(-(av * t2), rb.x, dR);
        if (fourth) dR = fmaf(-t2 * fmaf(-5.f * av, av, qo), rc.w, dR);
        const float qb = fmaf(dR, inv_dra, rb.y);
        const float rq = __builtin_rintf(qb);
        int bin;
        if (fabsf(qb - rq) > guard) {  // near a rounding tie: decide exactly
          const double4 t = g.traj[a];
          bin = exact_bin(g, x, y, z, t.x, t.y, t.z);
        } else {
          bin = __float_as_int(rb.z) + (int)rq;
        }
        if ((unsigned)bin < nbins) {
          const float2 d = g.raw[(int64_t)a * g.nra + bin];
          const float ph = rc.x * dR;
          const float phl = fmaf(rc.x, dR, -ph);  // exact residual of the product
          const float tt = __builtin_amdgcn_fractf(ph) + (fmaf(beta_t, dR * dR, rb.w) + phl);
          const float sn = __builtin_amdgcn_sinf(tt);  // sin(2 pi tt)
          const float cs = __builtin_amdgcn_cosf(tt);
          const float w = fmaf(dR, rc.z + dR, rc.y);  // R^2
          const float wr = w * d.x, wi = w * d.y;
          const float tr = fmaf(wr, cs, wi * sn);     // R^2 d exp(-j phi)
          const float ti = fmaf(wi, cs, -(wr * sn));
          if (ACCUM == 0) {
            const float yr = tr - cr, yi = ti - ci;
            const float nr = sr + yr, ni = si + yi;
            cr = (nr - sr) - yr;
            ci = (ni - si) - yi;
            sr = nr;
            si = ni;
          } else {
            sr += tr;
            si += ti;
          }
          ++cnt;
        }
      }
    }
    if (ACCUM == 1) {
      ar += (double)sr;
      ai += (double)si;
      sr = 0.f;
      si = 0.f;
    }
    __syncthreads();  // reads of this chunk finish before the next chunk overwrites rec
  }
  if (ACCUM == 0) {
    ar = (double)sr - (double)cr;
    ai = (double)si - (double)ci;
  }
  if (valid) store_mean(out, ncnt, i, ar, ai, cnt);
}

}  // namespace

int sarbp_launch_split(const sarbp_dataset *ds, const sarbp_scene *sc, const Tiling &t, hipStream_t st) {
  if (sc->n == 0) return SARBP_OK;
  const TileGrid tg = make_tile_grid(sc->dimx, sc->dimy, t.tile_x);
  const int64_t blocks = (tg.ntiles() + SARBP_NUM_XCD - 1) / SARBP_NUM_XCD * SARBP_NUM_XCD;
  if (blocks > 0x7fffffff) return sarbp_fail(SARBP_EINVAL, "too many tiles for one launch");
  if (t.accum == 0)
    hipLaunchKernelGGL(k_bp_split_tile<0>, dim3((unsigned)blocks), dim3(64), 0, st, ds->geom(), sc->view(), tg,
                       sc->out, sc->ncnt);
  else
    hipLaunchKernelGGL(k_bp_split_tile<1>, dim3((unsigned)blocks), dim3(64), 0, st, ds->geom(), sc->view(), tg,
                       sc->out, sc->ncnt);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return sarbp_hip_fail(e, "split kernel launch");
  return SARBP_OK;
}
