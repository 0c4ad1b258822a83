/*
 * sarbp.h -- C ABI of the MI355X (gfx950) SAR back-projection library, libsarbp.so.
 *
 * Plain pointers and sizes only.  No exceptions cross this boundary: every entry point returns
 * an int status (SARBP_OK == 0, negative on failure); sarbp_last_error() gives a message for the
 * calling thread's last failure.  The caller owns every host buffer; the library owns the device
 * buffers behind the opaque handles.  Calls on one handle must not overlap; different handles may
 * be used from different threads.
 *
 * Layouts (identical to the reference's in-memory arrays):
 *   RawData   complex64, column-major (Nra, Naz): sample (r, a) at element r + Nra*a, each
 *             element two float32 (re, im) -- sar_focuser_puntual.py:480-491, main_gpu.py:375-381.
 *   pixels    F-order flattened DEM grid: x/y/z float64, mask uint8, index/sint int32
 *             -- sar_focuser_puntual.py:195-201.
 *   output    complex128 (re, im float64 pairs) per pixel, same flat order; reshape (dimx, dimy)
 *             in F order -- sar_focuser_puntual.py:306.
 */
#ifndef SARBP_H
#define SARBP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SARBP_ABI_VERSION 1

/* status codes */
#define SARBP_OK 0
#define SARBP_EINVAL -1    /* invalid argument (null pointer, bad size, mismatched shapes) */
#define SARBP_ENOMEM -2    /* device or pinned-host allocation failed */
#define SARBP_ELAUNCH -3   /* kernel launch failed */
#define SARBP_EDEVICE -4   /* HIP runtime error (device lost, copy failed, ...) */
#define SARBP_ENODEV -5    /* no HIP device */
#define SARBP_EBUDGET -6   /* request exceeds the device-memory admission budget */
#define SARBP_ESTATE -7    /* call not valid in the object's current state */
#define SARBP_EAGAIN -8    /* not finished yet (poll) */

/* arithmetic of a focusing pass */
#define SARBP_PREC_F64 0   /* plain FP64, reference operation order (parity path) */
#define SARBP_PREC_SPLIT 1 /* FP64 per-tile reference range + FP32 offsets, Kahan accumulation */

typedef struct sarbp_dataset sarbp_dataset; /* device-resident RawData + trajectory + scalars */
typedef struct sarbp_scene sarbp_scene;     /* device-resident pixel set + its output buffers */

int sarbp_abi_version(void);
const char *sarbp_strerror(int status);
/* Message of the calling thread's last failed call ("" if none). */
const char *sarbp_last_error(void);
/* Number of visible HIP devices. */
int sarbp_device_count(int *count);

/*
 * Upload RawData and the trajectory once; they stay in HBM until sarbp_dataset_destroy.
 * Replaces the per-call host transpose + H2D copies of FocalizzatoreBpCuPy
 * (gpu_small_dataset.py:149-159, 262-275) and the per-batch raw window staging of
 * SARSuperOptimizedFocuser.process_dataset (main_gpu.py:465-515).
 * raw: complex64 column-major (nra, naz); tx/ty/tz: float64[naz];
 * near_range, dra, lambda_val: NearRangeDbl / DeltaRangeDbl / LambdaDbl (sar_focuser_puntual.py:505-507).
 */
int sarbp_dataset_create(const void *raw_c64, int64_t nra, int64_t naz, const double *tx,
                         const double *ty, const double *tz, double near_range, double dra,
                         double lambda_val, sarbp_dataset **out);
/* Same, uploading RawData in pulse blocks from a caller callback (for memmapped files larger
 * than host RAM, main_gpu.py:372-385): fill(user, a0, a1, dst) writes pulses [a0, a1) as
 * complex64 column-major into dst and returns 0, or non-zero to abort. */
typedef int (*sarbp_fill_fn)(void *user, int64_t a0, int64_t a1, void *dst);
int sarbp_dataset_create_streamed(sarbp_fill_fn fill, void *user, int64_t nra, int64_t naz,
                                  const double *tx, const double *ty, const double *tz,
                                  double near_range, double dra, double lambda_val,
                                  sarbp_dataset **out);
/* Release the dataset's device memory.  SARBP_ESTATE (and the dataset stays valid) while a server
 * holds requests against it that are not done or failed: preempted requests count, since they may
 * be resumed; destroying the server fails and releases them. */
int sarbp_dataset_destroy(sarbp_dataset *ds);
/* Device bytes held by the dataset. */
int sarbp_dataset_bytes(const sarbp_dataset *ds, int64_t *bytes);

/*
 * Upload a pixel set (F-order flattened DEM grid) for repeated focusing.  Pixels whose sint is
 * -9999 are masked, as FocalizzatoreBpPx does (sar_focuser_puntual.py:208-211).  dimx is the
 * grid's fast (first) dimension, so pixel (ix, iy) is element ix + dimx*iy; 0 treats the pixels
 * as a plain list.  Tiles of neighbouring pixels are formed from this shape.
 */
int sarbp_scene_create(sarbp_dataset *ds, const double *x, const double *y, const double *z,
                       const uint8_t *mask, const int32_t *index, const int32_t *sint, int64_t n,
                       int64_t dimx, sarbp_scene **out);
/*
 * Work mapping of one precision (tuning; the defaults need no call).  kernel 1 = one 64-pixel
 * tile per wavefront walking the union of the tile's apertures pulse by pulse, tile_x x (64/tile_x)
 * pixels; kernel 0 = one pixel per lane (FP64 only).  accum (split only): 0 = Kahan FP32,
 * 1 = FP32 blocks of 64 pulses folded into FP64.
 */
int sarbp_scene_set_tiling(sarbp_scene *sc, int precision, int kernel, int tile_x, int accum);
/*
 * Lowest range bin that contributes, for the next focusing passes of the scene: process_point's
 * ind_line_ra_min (sar_focuser_puntual.py:27, 106).  A pulse counts only when its bin lies in
 * [max(bin_min, 0), Nra-1): the reference keeps bins >= ind_line_ra_min and then skips negative ones
 * in its accumulation loop (:138).  Default 0.
 */
int sarbp_scene_set_bin_min(sarbp_scene *sc, int32_t bin_min);
int sarbp_scene_destroy(sarbp_scene *sc);

/*
 * Focus every pixel of the scene: for each pixel the mean over its aperture of
 * R^2 * RawData[r, a] * exp(-j*phi), with the gates, range-bin rule and phase of
 * process_point (sar_focuser_puntual.py:21-155).  Results stay on the device until
 * sarbp_scene_download.  gpu_ms (nullable) receives the kernel time measured with HIP events
 * on the library's stream.
 * SARBP_PREC_F64 is the drop-in for process_point_kernel (gpu_small_dataset.py:9-108,
 * main_gpu.py:60-155) with the Numba path's semantics.
 */
int sarbp_scene_focus(sarbp_dataset *ds, sarbp_scene *sc, int precision, float *gpu_ms);
/* Copy results to the host: out_c128 (2*n float64, nullable), ncontrib (n int32, nullable). */
int sarbp_scene_download(sarbp_scene *sc, double *out_c128, int32_t *ncontrib);

/*
 * One-shot: scene_create + scene_focus + scene_download + scene_destroy.  The C-ABI behind
 * FocalizzatoreBpPx (sar_focuser_puntual.py:160) and FocalizzatoreBpCuPy(_Batched)
 * (gpu_small_dataset.py:113, 229).
 */
int sarbp_focus(sarbp_dataset *ds, const double *x, const double *y, const double *z,
                const uint8_t *mask, const int32_t *index, const int32_t *sint, int64_t n,
                int64_t dimx, int precision, double *out_c128, int32_t *ncontrib);

/* ---- many requests in one persistent launch ------------------------------------------------ */

/* One request: a pixel set with its own grid shape and precision.  Inputs are host arrays of n
 * elements (F-order flattened grid, dimx its fast dimension, 0 = plain list); they are copied at
 * creation/submission.  tile_x: 0 = default. */
typedef struct sarbp_request {
  const double *x, *y, *z;
  const uint8_t *mask;
  const int32_t *index, *sint;
  int64_t n;
  int64_t dimx;
  int32_t precision;
  int32_t tile_x;
} sarbp_request;

typedef struct sarbp_batch sarbp_batch;

/*
 * Upload a set of requests against one dataset.  Has no counterpart in the reference beyond the
 * pixel sub-batches of FocalizzatoreBpCuPy_Batched / process_dataset (gpu_small_dataset.py:285-336,
 * main_gpu.py:519-560), which are launched one after another.
 */
int sarbp_batch_create(sarbp_dataset *ds, const sarbp_request *reqs, int32_t nreq, sarbp_batch **out);
/* Focus every request in ONE persistent launch.  tile_px: pixels per work item (64, 128 or 256:
 * 1, 2 or 4 wavefronts of 64-pixel tiles); pulse_chunk: pulses of per-(pulse, tile row)
 * coefficients staged in LDS per refill (8..1024, multiple of 8; tile_px/64 * pulse_chunk *
 * rows per tile * 48 B <= 64 KiB).  0 for either uses the last autotuned (or default) value.
 * The results do not depend on tile_px or pulse_chunk.  gpu_ms: kernel time (nullable). */
int sarbp_batch_run(sarbp_batch *b, int32_t tile_px, int32_t pulse_chunk, float *gpu_ms);
/* Time every valid (tile_px, pulse_chunk) in {64,128,256} x {16,32,...,1024} on this batch
 * (reps launches each, best kept) and use the fastest for later runs of batches of this shape. */
int sarbp_batch_autotune(sarbp_batch *b, int32_t reps, int32_t *best_tile_px, int32_t *best_chunk,
                         float *best_ms);
int sarbp_batch_download(sarbp_batch *b, int32_t req, double *out_c128, int32_t *ncontrib);
int sarbp_batch_destroy(sarbp_batch *b);

/*
 * Allocation-free batch focusing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same persistent launch as sarbp_batch_run over nreq requests whose pixel arrays
 * (reqs[k].x .. .sint) are DEVICE pointers, writing request k's means to out_c128[k] (2*n float64)
 * and its contributor counts to ncontrib[k] (n int32), both device pointers; the results equal
 * sarbp_batch_create + sarbp_batch_run + sarbp_batch_download bit for bit.  tile_px / pulse_chunk
 * as sarbp_batch_run (0: 64 / 32).  The launch's work list (request descriptors, items, queue
 * offsets and heads) lives in a fixed device array of SARBP_NOALLOC_SCRATCH_BYTES per GPU, held
 * until the process ends: SARBP_EINVAL when the requests need more (about 20 B per 64-pixel tile
 * and 256 B per request: ~12 M pixels).  No call allocates host or device memory.  Calls share that
 * array: they are serialised within a process (another call waits), and the call returns when the
 * results are in place (it synchronises the dataset's stream).
 */
#define SARBP_NOALLOC_SCRATCH_BYTES (4 << 20)
int sarbp_batch_focus_noalloc(sarbp_dataset *ds, const sarbp_request *reqs, int32_t nreq,
                              double *const *out_c128, int32_t *const *ncontrib, int32_t tile_px,
                              int32_t pulse_chunk);

/*
 * Diagnostic: the FP64 path's arithmetic for one pixel-pulse pair (sar_focuser_puntual.py:82-121)
 * evaluated on the device for n pairs: r2 = (dx^2 + dy^2) + dz^2, r = sqrt(r2), bin =
 * rint((r - near)/dra) half-even (-1 when outside [0, nra-1)), ph = k r + beta r2 and (sin, cos)(ph),
 * with x, y, z float64[n] pixel coordinates and pulse int32[n] in [0, naz).  Outputs (each [n])
 * may be NULL.  No counterpart in the reference: it lets a caller verify the device arithmetic
 * against a CPU evaluation bit for bit (r2, r, bin, ph) and to the last ulp (sin, cos).
 */
int sarbp_eval_pairs(sarbp_dataset *ds, const double *x, const double *y, const double *z, const int32_t *pulse,
                     int64_t n, double *r2, double *r, int32_t *bin, double *ph, double *sn, double *cs);

/* ---- serving: continuous batching of requests against resident datasets -------------------- */

/*
 * A server owns one scheduling thread.  Submitted requests wait in a FIFO until admitted; admission
 * needs a free slot and device memory within hbm_budget_bytes (the request's pixel buffers,
 * outputs and partial state; the datasets themselves are resident already).  Each scheduling round
 * is ONE persistent launch over every admitted request, advancing each tile of each request by up to
 * the round's pulse span; requests that arrive while a round runs join the next round.  The partial
 * sums of every pixel stay on the device between rounds (the per-request checkpoint), so a request
 * can be preempted (its state copied to the host and its device memory released) and resumed later
 * with a bit-identical result; results never depend on which requests share a round.
 * A request too small to fill the device on its own (few tiles) has each tile's pulses split into
 * up to max_segments segments, focused in parallel by separate wavefronts in its first round and
 * added in segment order by a second kernel: its latency drops by up to that factor, and its
 * result still depends only on the request (not on the mix), though it can differ in the last bits
 * from the same request focused with its pulses in one sequence.
 * A split-precision request whose a-priori estimate of its RMS phase error (against the FP64
 * result, over pixels with |a| > 1e-5 max|a|) exceeds phase_err_limit_rad is focused in FP64
 * instead; the estimate comes from the pixels' tile geometry and bounds the measured error.  No counterpart in the reference, whose drivers process one dataset
 * pass at a time (main_gpu.py:415-575).
 */
typedef struct sarbp_server sarbp_server;

typedef struct sarbp_server_config {
  int64_t hbm_budget_bytes;   /* admission budget for request buffers (0: 200 GB) */
  double quantum_ms;          /* target GPU time of one scheduling round (0: 20 ms) */
  int32_t max_active;         /* admitted requests at once (0: 256) */
  int32_t pulse_chunk;        /* pulse granularity of round spans and LDS refills (0: 32) */
  double phase_err_limit_rad; /* split -> FP64 fallback threshold (0: 1e-3) */
  int32_t max_segments;       /* pulse segments per tile of a small request (0: 16; 1: never split) */
} sarbp_server_config;

#define SARBP_REQ_QUEUED 0
#define SARBP_REQ_RUNNING 1
#define SARBP_REQ_DONE 2
#define SARBP_REQ_PREEMPTED 3
#define SARBP_REQ_FAILED 4

typedef struct sarbp_request_info {
  int32_t state;             /* SARBP_REQ_* */
  int32_t precision;         /* precision actually used (after the FP64 fallback decision) */
  double phase_err_est_rad;  /* a-priori estimate of the split result's RMS phase error */
  double submit_s, admit_s, done_s; /* server clock (s since server creation); submit_s: entry to submit */
  int64_t pp;                /* pixel-pulse pairs of the request */
  int32_t rounds;            /* scheduling rounds the request took part in */
  int32_t preemptions;
  int32_t segments;          /* pulse segments per tile (1: the tiles' pulses run in order) */
  double round_s;            /* start of the request's last scheduling round (work list build) */
  double gpu_done_s;         /* its kernels of that round finished (then: result copy-out) */
} sarbp_request_info;

typedef struct sarbp_server_stats {
  int64_t submitted, completed, failed, rounds;
  int64_t pp_done;            /* pixel-pulse pairs of completed requests */
  double gpu_busy_s;          /* summed kernel time of all rounds */
  int64_t device_bytes_in_use;
  int32_t active, queued;
  /* the scheduling thread's own time, summed: admitting requests (allocation, uploads issued),
   * building rounds (work lists, per-XCD queues, their upload and the launches), and rounds from
   * their start to the end of their wait for the device */
  double admit_s, build_s, round_wall_s;
  int64_t records;  /* request records held (unreleased requests) */
} sarbp_server_stats;

int sarbp_server_create(const sarbp_server_config *cfg, sarbp_server **out);
/* Copy the request's inputs (into page-locked host memory the server reuses) and queue it.
 * out_c128 (2*n float64) and ncontrib (n int32, nullable) receive the result when the request
 * completes: they, and the dataset, must stay valid until the request is done or failed (or the
 * server is destroyed); sarbp_dataset_destroy refuses a dataset with such requests.  SARBP_EINVAL
 * for a split request whose tile rows x pulse_chunk records exceed the batch kernel's LDS.  SARBP_EBUDGET if the request alone exceeds the
 * budget.  The request's latency (sarbp_request_info: done_s - submit_s) starts at the entry to
 * this call.  Once a request is done or failed the server keeps only its record (state, timings,
 * counts; until sarbp_server_release): its copy of the inputs is released. */
int sarbp_server_submit(sarbp_server *s, sarbp_dataset *ds, const sarbp_request *req, double *out_c128,
                        int32_t *ncontrib, int64_t *id);
int sarbp_server_poll(sarbp_server *s, int64_t id, int32_t *state);
/* Block until done (SARBP_OK), failed (SARBP_EDEVICE) or timeout_s elapsed (SARBP_EAGAIN). */
int sarbp_server_wait(sarbp_server *s, int64_t id, double timeout_s);
/* Stop scheduling the request at the next round boundary, save its partial state on the host and
 * release its device memory; SARBP_ESTATE if it is already done.  A request already scheduled in
 * the round that completes it completes instead (the server queues the next round while one runs,
 * so that is the round on the device or the one queued behind it). */
int sarbp_server_preempt(sarbp_server *s, int64_t id);
/* Queue a preempted request again; it continues from its saved state. */
int sarbp_server_resume(sarbp_server *s, int64_t id);
int sarbp_server_info(sarbp_server *s, int64_t id, sarbp_request_info *info);
/* Drop the record of a done or failed request (its state, timings and counts): afterwards its id
 * is unknown.  SARBP_ESTATE while the request is queued, running, preempted, or waited on by
 * sarbp_server_wait in another thread.  A server that is never asked to release keeps one record
 * per request until it is destroyed. */
int sarbp_server_release(sarbp_server *s, int64_t id);
int sarbp_server_get_stats(sarbp_server *s, sarbp_server_stats *st);
/* Finish the current round, fail the requests still queued or running, join the thread. */
int sarbp_server_destroy(sarbp_server *s);

/*
 * Pixel-pulse pairs evaluated by a focusing pass (host-side, integers only): pixels with
 * mask == 1 and sint != -9999, pulses 0 < index + j - ceil(|sint|/2) < naz - 1 for
 * j < 2*ceil(sint/2) + 1.  The unit of the throughput metric.
 */
int64_t sarbp_count_pp(const uint8_t *mask, const int32_t *index, const int32_t *sint, int64_t n,
                       int64_t naz);

/*
 * A-priori estimate of the split path's RMS phase error (rad) for a request's pixels on a dataset,
 * against the FP64 result over the pixels with |a| > 1e-5 max|a|: the estimate behind the server's
 * FP64 fallback (measured errors stay below it).  Host-side, from the split tiling's rows (the
 * request's tile_x if its precision is SARBP_PREC_SPLIT, else the default); 1e9 when those rows are
 * too wide for the split form (the kernel takes FP64 terms there), 0 without contributing pixels.
 * No counterpart in the reference; for callers of sarbp_focus choosing a precision.
 */
int sarbp_split_error_estimate(sarbp_dataset *ds, const sarbp_request *req, double *rms_phase_rad);

#ifdef __cplusplus
}
#endif
#endif /* SARBP_H */
