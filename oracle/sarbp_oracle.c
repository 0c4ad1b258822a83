/*
 * sarbp_oracle.c -- CPU restatement of the reference back-projection, used ONLY as a checker.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only to check or to time the CPU baseline.  The product path
 * (sar-backprojection_amd/) never links, loads or calls it.
 *
 * What it restates (reference files, read as text):
 *   process_point        sar_focuser_puntual.py:21-155   (per-pixel FP64 back-projection)
 *   FocalizzatoreBpPx    sar_focuser_puntual.py:195-211  (Sint == -9999 -> mask = 0)
 *                        sar_focuser_puntual.py:259-268  (serial pixel loop, F-order flat index)
 *
 * Arithmetic contract (the reason this file is compiled with -O2 -ffp-contract=off, no fast-math):
 *   - every FP64 operation is evaluated in the reference's order, one IEEE rounding each;
 *   - R = sqrt(R2) and (R - near)/dra are correctly rounded (glibc sqrt, IEEE division);
 *   - np.round -> round-half-to-even (nearbyint in the default rounding mode);
 *   - exp(-j*phi) -> (cos(phi), -sin(phi)) from glibc;
 *   - the complex product (R2*d)*f is expanded as (ac - bd, ad + bc), d widened from complex64;
 *   - ascending-j sequential accumulation; the mean is acc/n componentwise.
 *
 * Pinning (tests/test_oracle.py): brute-force aperture enumeration, exact range-bin ties,
 * closed-form single-pulse values against mpmath, linearity in RawData, point-target focusing,
 * and bitwise agreement with the independent NumPy restatement in oracle/bp_numpy.py.
 * Numba's own complex-exp/complex-divide codegen is not importable here ("parity unpinned" at
 * the last-ulp level; well below the 1e-10 relative bound).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Constants of process_point, sar_focuser_puntual.py:35-40.  PI is float32(pi) held in a double,
 * as the reference does (IDL's single-precision !PI). */
static const double ORC_C = 2.9979246e8;
static const double ORC_ALFA = 3.335987619777693e11;
static const double ORC_PI = 3.141592741012573242;

/* Aperture of one pixel, sar_focuser_puntual.py:57-66: J = ceil(s/2)*2 + 1 pulses starting at
 * index - ceil(|s|/2).  Returns J (may be <= 0) and the start pulse. */
static void orc_aperture(int32_t index, int32_t sint, int64_t *start, int64_t *count)
{
    int64_t J = (int64_t)ceil((double)sint / 2.0) * 2 + 1;
    int64_t h = (int64_t)ceil(fabs((double)sint) / 2.0);
    *start = (int64_t)index - h;
    *count = J;
}

/* process_point, sar_focuser_puntual.py:21-155, with its ind_line_ra_min argument (sfp:27): the
 * range-bin filter keeps bins >= ind_line_ra_min (sfp:106) and the accumulation loop then skips bins
 * outside [0, Nra) (sfp:138).  raw is interleaved complex64, column-major (Nra, Naz): sample (r, a)
 * at raw[2*(r + Nra*a)].  Returns the contributor count n. */
int32_t sarbp_oracle_point_ex(double x, double y, double z, uint8_t mask, int32_t index, int32_t sint,
                              const double *tx, const double *ty, const double *tz,
                              const float *raw, int64_t Nra, int64_t Naz,
                              double near_range, double dra, double lambda_val, int32_t ind_line_ra_min,
                              double *out_re, double *out_im)
{
    const double beta = 4.0 * ORC_PI * ORC_ALFA / (ORC_C * ORC_C); /* sfp:40 */
    const double kphase = -4.0 * ORC_PI / lambda_val;               /* sfp:116, left operand */
    *out_re = 0.0;
    *out_im = 0.0;
    if (mask != 1) return 0;          /* sfp:46-47 */
    if (sint == -9999) return 0;      /* sfp:50-53 */
    int64_t a0, J;
    orc_aperture(index, sint, &a0, &J);
    double acc_re = 0.0, acc_im = 0.0;
    int32_t n = 0;
    for (int64_t j = 0; j < J; ++j) {
        int64_t a = a0 + j;
        if (!(a < Naz - 1 && a > 0)) continue;                 /* sfp:66 */
        double dx = x - tx[a];                                 /* sfp:82-84 */
        double dy = y - ty[a];
        double dz = z - tz[a];
        double r2 = dx * dx + dy * dy + dz * dz;               /* sfp:88, left to right */
        double r = sqrt(r2);                                   /* sfp:94 */
        double q = nearbyint((r - near_range) / dra);          /* sfp:100-101, half-even */
        if (!(q < (double)(Nra - 1) && q >= (double)ind_line_ra_min)) continue; /* sfp:106 */
        int64_t rb = (int64_t)q;
        if (!(rb >= 0 && rb < Nra)) continue;                  /* sfp:138 */
        double phase = kphase * r + beta * r2;                 /* sfp:116 */
        double fr = cos(phase);                                /* sfp:121: exp(-j*phase) */
        double fi = -sin(phase);
        const float *d = raw + 2 * (rb + Nra * a);             /* sfp:139 */
        double dre = (double)d[0], dim = (double)d[1];
        double t1r = r2 * dre, t1i = r2 * dim;                 /* sfp:144: R2 * raw */
        acc_re += t1r * fr - t1i * fi;                         /* (.) * conj filter */
        acc_im += t1r * fi + t1i * fr;
        n += 1;                                                /* sfp:145-146 */
    }
    if (n > 0) {                                               /* sfp:150-152 */
        *out_re = acc_re / (double)n;
        *out_im = acc_im / (double)n;
    }
    return n;
}

/* process_point with ind_line_ra_min = 0, as FocalizzatoreBpPx calls it (sar_focuser_puntual.py:214). */
int32_t sarbp_oracle_point(double x, double y, double z, uint8_t mask, int32_t index, int32_t sint,
                           const double *tx, const double *ty, const double *tz,
                           const float *raw, int64_t Nra, int64_t Naz,
                           double near_range, double dra, double lambda_val,
                           double *out_re, double *out_im)
{
    return sarbp_oracle_point_ex(x, y, z, mask, index, sint, tx, ty, tz, raw, Nra, Naz, near_range, dra,
                                 lambda_val, 0, out_re, out_im);
}

/* The per-(pixel, pulse) arithmetic of process_point (sar_focuser_puntual.py:82-121) for n pairs:
 * R2, R, the range bin (-1 when outside [0, Nra-1)), the phase and (sin, cos) of it.  Outputs
 * may be NULL.  Used to check the device arithmetic bit for bit. */
void sarbp_oracle_pairs(const double *x, const double *y, const double *z, const int32_t *pulse, int64_t n,
                        const double *tx, const double *ty, const double *tz, int64_t Nra,
                        double near_range, double dra, double lambda_val,
                        double *r2o, double *ro, int32_t *bino, double *pho, double *sno, double *cso,
                        int nthreads)
{
    const double beta = 4.0 * ORC_PI * ORC_ALFA / (ORC_C * ORC_C); /* sfp:40 */
    const double kphase = -4.0 * ORC_PI / lambda_val;               /* sfp:116 */
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (int64_t i = 0; i < n; ++i) {
        const int64_t a = pulse[i];
        double dx = x[i] - tx[a];
        double dy = y[i] - ty[a];
        double dz = z[i] - tz[a];
        double r2 = dx * dx + dy * dy + dz * dz;
        double r = sqrt(r2);
        double q = nearbyint((r - near_range) / dra);
        double ph = kphase * r + beta * r2;
        if (r2o) r2o[i] = r2;
        if (ro) ro[i] = r;
        if (bino) bino[i] = (q < (double)(Nra - 1) && q >= 0.0) ? (int32_t)q : -1;
        if (pho) pho[i] = ph;
        if (sno) sno[i] = sin(ph);
        if (cso) cso[i] = cos(ph);
    }
    (void)nthreads;
}

/* Pixel loop of FocalizzatoreBpPx (sar_focuser_puntual.py:208-211, 259-268) over F-order
 * flattened pixels.  out is interleaved complex128 [n]; ncontrib may be NULL.
 * nthreads <= 0 uses the OpenMP default; 1 reproduces the reference's serial loop. */
void sarbp_oracle_focus(const double *x, const double *y, const double *z, const uint8_t *mask,
                        const int32_t *index, const int32_t *sint, int64_t npix,
                        const double *tx, const double *ty, const double *tz,
                        const float *raw, int64_t Nra, int64_t Naz,
                        double near_range, double dra, double lambda_val,
                        double *out, int32_t *ncontrib, int nthreads)
{
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
#endif
    for (int64_t i = 0; i < npix; ++i) {
        uint8_t m = (sint[i] == -9999) ? 0 : mask[i];  /* sfp:208-211 */
        int32_t c = sarbp_oracle_point(x[i], y[i], z[i], m, index[i], sint[i], tx, ty, tz, raw,
                                       Nra, Naz, near_range, dra, lambda_val,
                                       &out[2 * i], &out[2 * i + 1]);
        if (ncontrib) ncontrib[i] = c;
    }
    (void)nthreads;
}

/* Contributing (pulse, range-bin, R) triples of one pixel in accumulation order, for the
 * bin-selection tests.  Returns how many were written (at most cap). */
int64_t sarbp_oracle_trace(double x, double y, double z, int32_t index, int32_t sint,
                           const double *tx, const double *ty, const double *tz,
                           int64_t Nra, int64_t Naz, double near_range, double dra,
                           int64_t *pulses, int64_t *bins, double *ranges, int64_t cap)
{
    if (sint == -9999) return 0;
    int64_t a0, J, k = 0;
    orc_aperture(index, sint, &a0, &J);
    for (int64_t j = 0; j < J && k < cap; ++j) {
        int64_t a = a0 + j;
        if (!(a < Naz - 1 && a > 0)) continue;
        double dx = x - tx[a], dy = y - ty[a], dz = z - tz[a];
        double r2 = dx * dx + dy * dy + dz * dz;
        double r = sqrt(r2);
        double q = nearbyint((r - near_range) / dra);
        if (!(q < (double)(Nra - 1) && q >= 0.0)) continue;
        pulses[k] = a;
        bins[k] = (int64_t)q;
        ranges[k] = r;
        ++k;
    }
    return k;
}

/* Number of geometry-evaluated pixel-pulse pairs: pixels with mask == 1 and sint != -9999,
 * pulses j < J with 0 < index + j - h < Naz - 1 (the unit of work of the throughput metric). */
int64_t sarbp_oracle_count_pp(const uint8_t *mask, const int32_t *index, const int32_t *sint,
                              int64_t npix, int64_t Naz)
{
    int64_t tot = 0;
    for (int64_t i = 0; i < npix; ++i) {
        if (mask[i] != 1 || sint[i] == -9999) continue;
        int64_t a0, J;
        orc_aperture(index[i], sint[i], &a0, &J);
        if (J <= 0) continue;
        int64_t lo = a0 > 1 ? a0 : 1;
        int64_t hi = a0 + J - 1 < Naz - 2 ? a0 + J - 1 : Naz - 2;
        if (hi >= lo) tot += hi - lo + 1;
    }
    return tot;
}
