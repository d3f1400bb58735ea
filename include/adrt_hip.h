/* C ABI of libadrt_hip.so: the MI355X engine for the ADRT oblique inverses of dsrim/adrt-pinv.
 *
 * Two layers.
 *  1. Reference-compatible row kernels on HOST buffers, same names and argument meaning as the C++ engine that
 *     fastcpp.py binds with ctypes (cpp/oblique.cpp:53-139; argtypes at fastcpp.py:25-31).  They copy in, run on the
 *     GPU and copy out.  `nthreads` is accepted and ignored.  Like the reference they return void; on a bad argument or
 *     a HIP failure they print one line to stderr and set adrt_last_error().
 *  2. A device-resident API (device pointers, explicit HIP streams, int status codes: 0 ok, -1 invalid argument,
 *     -2 HIP error, -3 out of memory, -4 bad state) for the row engine, the whole L_p / hybrid apply, hipGraph capture
 *     and the forward / transpose ADRT.
 *
 * Complex numbers are interleaved (re, im) doubles, the layout of numpy complex128 and std::complex<double>.
 * A stream argument of NULL means the default stream.  No call keeps a pointer it was given beyond the call, except
 * the plans, which copy their parameters to the device at creation.
 */
#ifndef ADRT_HIP_H
#define ADRT_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double re, im;
} adrt_c128;

/* ------------------------------------------------------------------ layer 1: reference-compatible (host buffers) */

/* X: K rows of length N = 2^n, in place.  transpose = 0: B_{m_hi} ... B_{m_lo}; 1: the adjoint.
 * Replaces bfly_rows of cpp/oblique.cpp:56 (bound at fastcpp.py:25). */
void bfly_rows(int n, int K, const adrt_c128* z, adrt_c128* X, int m_lo, int m_hi, int transpose, int nthreads);

/* Z[k] = p_k(M) Y[k] with deg[k] Chebyshev steps, constants a[k], c[k], M = B^H B over levels m_lo..m_hi;
 * *steps += sum_k max(deg[k]-1, 0) * 2 (m_hi - m_lo + 1) when steps != NULL.
 * Replaces cheb_rows of cpp/oblique.cpp:74-75 (bound at fastcpp.py:27). */
void cheb_rows(int n, int K, const adrt_c128* z, const int* deg, const double* a, const double* c, double tau,
               const adrt_c128* Y, adrt_c128* Z, int m_lo, int m_hi, int nthreads, long long* steps);

/* Certified upper bounds Lambda_k >= lambda_max(M_k) by the row-sum recursion over levels m_lo..n.
 * Replaces cert_rowsum of cpp/oblique.cpp:108 (bound at fastcpp.py:29). */
void cert_rowsum(int n, int K, const double* omega, int m_lo, double* out, int nthreads);

/* Parallel width of the engine: the compute-unit count of the current GPU.
 * Replaces omp_max_threads of cpp/oblique.cpp:132 (bound at fastcpp.py:31). */
int omp_max_threads(void);

/* Status-returning variants of the three kernels above (same arguments, no nthreads). */
int adrt_bfly_rows(int n, int K, const adrt_c128* z, adrt_c128* X, int m_lo, int m_hi, int transpose);
int adrt_cheb_rows(int n, int K, const adrt_c128* z, const int* deg, const double* a, const double* c, double tau,
                   const adrt_c128* Y, adrt_c128* Z, int m_lo, int m_hi, long long* steps);
int adrt_cert_rowsum(int n, int K, const double* omega, int m_lo, double* out);

/* ------------------------------------------------------------------ layer 2: device-resident API */

/* message of the last failure on this thread; the pointer stays valid until the next failing call on the thread */
const char* adrt_last_error(void);
int adrt_device_count(void);
int adrt_malloc(void** ptr, unsigned long long bytes);
int adrt_free(void* ptr);
/* kind: 0 host->host, 1 host->device, 2 device->host (synchronous), 3 device->device (stream-ordered),
 * 4 device->host stream-ordered (returns at once; for page-locked destinations from adrt_host_malloc) */
int adrt_memcpy(void* dst, const void* src, unsigned long long bytes, int kind, void* stream);
/* page-locked host memory: host<->device copies from / to it run asynchronously on the stream */
int adrt_host_malloc(void** ptr, unsigned long long bytes);
int adrt_host_free(void* ptr);
int adrt_memset(void* dst, int value, unsigned long long bytes, void* stream);
int adrt_stream_create(void** stream);
/* A stream on a hardware queue of its own.  HIP maps ordinary streams onto a small pool of hardware queues (4 by
 * default), so work queued on one stream can hold up another stream that shares its queue; a stream created with a CU
 * mask (here: every CU) gets a queue of its own (measured on MI355X).  For long-lived streams whose work must not wait
 * behind other streams', e.g. one per request size in a server.  Synchronises with the null stream like
 * hipStreamCreate's streams. */
int adrt_stream_create_own_queue(void** stream);
/* At most ADRT_MAX_OWN_QUEUES own-queue streams live at once (each holds a hardware queue beside HIP's pool); one
 * more fails with -4 until one is destroyed.  adrt_own_queue_count: how many are alive.  (No counterpart in the
 * reference, which has no device queues.) */
#define ADRT_MAX_OWN_QUEUES 4
int adrt_own_queue_count(void);
int adrt_stream_destroy(void* stream);
int adrt_stream_sync(void* stream);
int adrt_stream_wait(void* waiting, void* signal); /* `waiting` waits for the work enqueued so far on `signal` */
int adrt_stream_wait_event(void* stream, void* ev); /* later work on `stream` waits for the last record of `ev` */
int adrt_device_sync(void);
int adrt_device_mem_info(unsigned long long* free_bytes, unsigned long long* total_bytes);
int adrt_event_create(void** ev);
int adrt_event_destroy(void* ev);
int adrt_event_record(void* ev, void* stream);
int adrt_event_elapsed_ms(void* start, void* stop, float* ms);
int adrt_event_query(void* ev); /* 1 done, 0 pending, < 0 error */

/* Row engine with device-resident parameters: the plug-in surface of fastops.RowEngine (fastops.py:20-167).
 * Row r of a run uses parameter row r % L.  Rows whose parameter row is listed in dense_k use the exact filter
 * dense_F[i] (N x N, row-major, complex128) in P ops.  precision 0: complex128 rows; 1: complex64 rows with the
 * TwoSum-compensated Chebyshev iterate; 2: complex64 rows without the compensation.
 * ops: nops triples (kind, lo, hi), kind 0 = B, 1 = BT, 2 = P, applied left to right (fastops.py:156-167). */
typedef struct adrt_rows_plan adrt_rows_plan;
int adrt_rows_plan_create(adrt_rows_plan** out, int n, int L, const adrt_c128* z, const int* deg, const double* a,
                          const double* c, double tau, int n_dense, const int* dense_k, const adrt_c128* dense_F,
                          int precision);
int adrt_rows_plan_destroy(adrt_rows_plan* plan);
int adrt_rows_run(adrt_rows_plan* plan, const void* dY, void* dZ, long long rows, const int* ops, int nops,
                  void* stream);

/* Row-kernel variant per (n, class): 0 = built-in default, -1 = generic LDS kernel, 2..4 = register kernel with
 * 2^loge values per lane.  Class 0: FP64 passes, 1: FP32 passes, 2: the FP64 refinement passes of FP32 plans (unset:
 * as class 0).  Chosen by the autotuner; the environment variable ADRT_ROWS_LOGE sets every size. */
int adrt_rows_config_set(int n, int precision, int loge);
int adrt_rows_config_get(int n, int precision);

/* Dense-row GEMM placement per (n, precision) for passes whose dense rows take the MFMA path: 0 = built-in default
 * (one dense wave per SIMD inside the register row kernel's launch), -1 = a launch of its own after the rows,
 * 1..8 = at most that many dense waves per SIMD inside the row kernel's launch.  precision 2 / 3: the placement of
 * FP64 / FP32 passes over one sinogram (unset or 0: that of precision 0 / 1).  Chosen by the autotuner.  (No
 * counterpart in the reference, whose dense rows run inside its row loop.) */
int adrt_dense_config_set(int n, int precision, int mode);
int adrt_dense_config_get(int n, int precision);

/* Inverse plans: fastops.FastLp / FastHybrid (fastops.py:261-338) with the setup scalars given.
 * All image / sinogram buffers are float64 on the device:
 *   b: (S, 4, 2N-1, N) sinograms, x / y: (S, N, N) images. */
/* KIND_NORMAL: the normal operator R^T R of the ADRT row by row in frequency (the row Gram gram_freq, ops B 1..n then
 * BT 1..n on the packed quadrant pairs, patches/adrt-gram-freq.patch:220-251, 411-471) with backprojection R^T
 * (ops BT 1..n) and no filter or preconditioner: its apply is the fixed-degree Chebyshev semi-iteration on the normal
 * equations on [lo, hi] = [lam_min, lam_max] with outer_deg steps (lstsq_cheb, patch:474-545).  deg, a and c may be
 * null; n_dense must be 0. */
enum { ADRT_KIND_LP = 0, ADRT_KIND_HYBRID = 1, ADRT_KIND_NORMAL = 2 };
enum { ADRT_PRECOND_NONE = 0, ADRT_PRECOND_G1 = 1, ADRT_PRECOND_LORR = 2 };
enum { ADRT_PREC_F64 = 0, ADRT_PREC_F32 = 1, ADRT_PREC_F32_PLAIN = 2 };
typedef struct {
    int kind;           /* ADRT_KIND_* */
    int n;              /* N = 2^n, 1 <= n <= 11 (hybrid: n >= 2) */
    int precision;      /* ADRT_PREC_F64: every op in double.  ADRT_PREC_F32: offset FFTs, butterflies and Chebyshev
                           steps in float, the Chebyshev iterate accumulated with TwoSum compensation; dense table,
                           preconditioner and outer iteration in double.  ADRT_PREC_F32_PLAIN: ADRT_PREC_F32 without
                           the compensation */
    int precond;        /* hybrid only: ADRT_PRECOND_* */
    int g1_deg;         /* Chebyshev degree of G_1^{-1} (fastops default 14) */
    int outer_deg;      /* D, the outer Chebyshev degree */
    double lo, hi;      /* outer interval */
    double tau;         /* Tikhonov shift tau_fac * N */
    const double* lam_max; /* (L) symmetrised bounds; informational, may be NULL */
    const adrt_c128* z; /* (L) exp(i w_k), w = 2 pi fftfreq(2N) */
    const int* deg;     /* (L) inner degrees */
    const double* a;    /* (L) Chebyshev constants (fastspec.cheb_coeffs) */
    const double* c;    /* (L) */
    int n_dense;        /* frequencies with the exact filter */
    const int* dense_k; /* (n_dense) their indices in [0, L) */
    const adrt_c128* dense_F; /* (n_dense, N, N) row-major; may be NULL when dense_Wop and dense_Wback are given
                                 (plans with pass matrices never read it) */
    int max_batch;      /* sinograms per call; scratch is sized for it */
    /* optional: the whole op list of a dense row as one matrix (row-major (n_dense, N, N)): the operator pass
       (B1^H M' F B1 for G_h, M F for the frame) and the backprojection pass (B1^H F A'^H, F A^H).  When given,
       dense rows run as a batched GEMM instead of inside the row kernel.  For dense frequencies k and L - k whose z
       are exact conjugates, the GEMM applies the conjugate of k's matrices to the rows of L - k (the ops have real
       coefficients, so these are L - k's matrices) and does not read L - k's own. */
    const adrt_c128* dense_Wop;
    const adrt_c128* dense_Wback;
    int refine;         /* ADRT_PREC_F32 only: the backprojection and the last `refine` outer steps run in FP64 */
    int dense_on_device; /* 1: dense_F / dense_Wop / dense_Wback are DEVICE pointers to column-major matrices
                            (the row-major transposes), copied without a host round trip.  2: the same, but the plan
                            borrows them (no copy): the caller keeps them alive until adrt_plan_destroy, and plans of
                            one size share one set of tables */
    int g1_fp64;        /* ADRT_PREC_F32 / F32_PLAIN: 1 runs the G_1 preconditioner's inner iteration in double in the
                           FP32 outer steps too (default 0: in float there, its input rounded to float) */
    int l1_fp64;        /* ADRT_PREC_F32 hybrid plans: 1 runs the explicit level 1 (B 1..1, BT 1..1 of fastops.py:330), the
                           offset FFTs and the combine of the FP32 outer steps in double; the row kernel then reads and
                           writes complex128 rows and runs levels 2..n and the Chebyshev filter in float (default 0) */
    const int* dense_w; /* optional (n_dense), n >= 4, with dense_Wop / dense_Wback: the index of dense frequency i's
                           pass matrices in those tables, or -1 when it has none of its own; then L - k is dense, its
                           z the exact conjugate of k's, with an index >= 0, and the plan applies that partner's
                           matrices conjugated.  The tables hold max(dense_w) + 1 matrices.  NULL: frequency i's
                           matrices are entry i of n_dense */
} adrt_plan_desc;

typedef struct adrt_plan adrt_plan;
unsigned long long adrt_plan_bytes(const adrt_plan_desc* desc); /* device bytes a plan holds (admission control) */
int adrt_plan_create(adrt_plan** out, const adrt_plan_desc* desc);
int adrt_plan_destroy(adrt_plan* plan);

/* Contexts: a stream and a device-memory budget for one caller (a serving loop, a worker thread).  Plans created in
 * a context charge adrt_plan_bytes(desc) to its budget and return it when destroyed; adrt_ctx_malloc / adrt_ctx_free
 * do the same for buffers.  A request beyond the budget fails with -3 and allocates nothing.  budget_bytes 0 means no
 * limit.  adrt_ctx_destroy fails with -4 while plans or buffers of the context are alive.  One host thread at a time
 * per context; contexts are independent.  (The reference has no device state; this is the admission control of the
 * device-resident API.) */
typedef struct adrt_ctx adrt_ctx;
int adrt_ctx_create(adrt_ctx** out, unsigned long long budget_bytes);
int adrt_ctx_destroy(adrt_ctx* ctx);
void* adrt_ctx_stream(adrt_ctx* ctx); /* the context's stream (non-blocking); null for a null context */
int adrt_ctx_usage(adrt_ctx* ctx, unsigned long long* used_bytes, unsigned long long* budget_bytes);
int adrt_ctx_malloc(adrt_ctx* ctx, void** ptr, unsigned long long bytes);
int adrt_ctx_free(adrt_ctx* ctx, void* ptr);
int adrt_plan_create_in(adrt_ctx* ctx, adrt_plan** out, const adrt_plan_desc* desc);
/* replaces the outer interval and degree (set after the setup's eigenvalue estimate) */
int adrt_plan_set_outer(adrt_plan* plan, double lo, double hi, int outer_deg);
/* y = backproject(b) (fastops.py:283-285 / 333-335) */
int adrt_plan_backproject(adrt_plan* plan, const double* d_b, double* d_y, int S, void* stream);
/* y = G x: frame_window (L_p, fastops.py:277-281) or G_h (hybrid, fastops.py:325-331) */
int adrt_plan_operator(adrt_plan* plan, const double* d_x, double* d_y, int S, void* stream);
/* y = P x: g1_inv / lorr_inv (hybrid) or the identity (fasthybrid.py:101-130) */
int adrt_plan_precond(adrt_plan* plan, const double* d_x, double* d_y, int S, void* stream);
/* the row kernel alone on the rows of the plan's last packed pass (which: 0 backproject ops, 1 operator ops, in the
 * plan's precision; 2 / 3 the same in FP64, on rows an FP64 pass left, e.g. after an apply of an FP32 plan with
 * refine >= 1), written to a second buffer: the dominant kernel of an apply, for timing and tuning */
int adrt_plan_rows_pass(adrt_plan* plan, int which, int S, void* stream);
/* every later launch of the plan's row kernel stamps its own start and end into the two events (hipExtLaunchKernel;
 * NULL, NULL: stop), so the kernel alone can be timed inside a pass; not for a plan whose work is being captured into
 * a graph */
int adrt_plan_time_rows(adrt_plan* plan, void* ev_begin, void* ev_end);
/* the same with count event pairs taken in turn (launch i uses pair i mod count, counting from this call; count 0:
 * stop), so every row-kernel launch of a whole apply can be timed where it runs */
int adrt_plan_time_rows_ring(adrt_plan* plan, void* const* ev_begin, void* const* ev_end, int count);
/* row-kernel launches stamped since the last adrt_plan_time_rows / _ring call (a pass whose rows are all dense
 * launches no row kernel and stamps nothing), so a caller can check which launches its event pairs hold */
int adrt_plan_time_rows_count(adrt_plan* plan, long long* stamped);
/* the outer steps an apply of the plan runs in FP64 at its end: min(refine, outer_deg - 1) for ADRT_PREC_F32, else
 * 0 (FP32_PLAIN and FP64 plans take no refinement) */
int adrt_plan_refine_steps(adrt_plan* plan);
/* x = L b, the whole inverse (fastops.py:287-288 / 337-338); d_x must not alias d_b */
int adrt_plan_apply(adrt_plan* plan, const double* d_b, double* d_x, int S, void* stream);
/* Continuous batching at outer-step granularity.  No counterpart in the reference (its inverses take one sinogram
 * per call); for servers that let sinograms join a running batch between outer steps instead of waiting for the
 * whole batch.  One outer step of adrt_plan_apply for every lane i < S (S <= min(max_batch, 64)): lane i is at step
 * k_lane[i] (0 <= k < outer_deg), lanes ordered oldest first (k_lane non-increasing).  A lane at step 0 reads its
 * sinogram from d_b (lane i at d_b + i * 4 (2N-1) N) and keeps its right-hand side in the plan; a lane at step k >= 1
 * reads x_(k-1) from d_xcur and, for k >= 2, x_(k-2) from d_xnext.  Every lane writes its x_k to d_xnext (lane i at
 * offset i N^2), and the caller swaps d_xcur and d_xnext between ticks.  After outer_deg ticks a lane holds in
 * d_xnext the same bits adrt_plan_apply gives for its sinogram. */
int adrt_plan_tick(adrt_plan* plan, const int* k_lane, int S, const double* d_b, double* d_xcur, double* d_xnext,
                   void* stream);
/* moves lane `from` of a continuous batch to lane `to`: its right-hand side (in the plan) and its images in d_xa and
 * d_xb (the two iterate buffers of adrt_plan_tick); no counterpart in the reference */
int adrt_plan_lane_move(adrt_plan* plan, int from, int to, double* d_xa, double* d_xb, void* stream);

/* hipGraph capture of everything enqueued on `stream` (and streams joined through adrt_stream_wait). */
typedef struct adrt_graph adrt_graph;
int adrt_graph_begin(void* stream);
int adrt_graph_end(void* stream, adrt_graph** out);
int adrt_graph_launch(adrt_graph* graph, void* stream);
int adrt_graph_destroy(adrt_graph* graph);

/* Forward ADRT R: (S, N, N) -> (S, 4, 2N-1, N) and its transpose (adrtcore.py:122-137).
 * dtype 0: int32 in and out (forward), int32 in / int64 out (transpose); 1: float32; 2: float64.
 * d_work: adrt_work_bytes_fwd / adrt_work_bytes_T bytes of device scratch. */
unsigned long long adrt_work_bytes_fwd(int n, int dtype, int S);
unsigned long long adrt_work_bytes_T(int n, int dtype, int S);
int adrt_forward(int n, int dtype, const void* d_x, void* d_b, void* d_work, int S, void* stream);
int adrt_transpose(int n, int dtype, const void* d_b, void* d_x, void* d_work, int S, void* stream);

/* adrt_forward without a workspace argument, allocating nothing: for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference, whose
 * forward ADRT is NumPy on the host (adrtcore.py:122-128).  The levels use a static device array of
 * ADRT_NOALLOC_SCRATCH_MIB MiB that every GPU the library runs on holds until the process ends; the sinograms run in
 * chunks that fit it, so any S works and N up to 2048 in every dtype (a larger N, whose one sinogram needs more,
 * fails with -1 before any device work; adrt_forward has no such limit).  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other threads) race, so keep them on one stream or ordered.
 * No part of the array needs to start at zero (each level writes every entry it later reads).  The results are the
 * bits adrt_forward gives. */
#define ADRT_NOALLOC_SCRATCH_MIB 256
unsigned long long adrt_noalloc_scratch_bytes(void);
int adrt_forward_noalloc(int n, int dtype, const void* d_x, void* d_b, int S, void* stream);

/* Sinogram layout permutation, the device half of pkg_to_core / core_to_pkg (skimage_adrt_pkg.py:38-51): for each of S
 * float64 sinograms of M = 4 (2N-1) N entries, out[s][i] = in[s][src_of[i]] where src_of[i] >= 0, else 0.  d_src_of:
 * M int32 on the device, each in [-1, M) (the kernel does not check; the host builds the map from fingerprints of
 * both transforms); d_out must not alias d_in. */
int adrt_layout_permute(int n, const int* d_src_of, const double* d_in, double* d_out, int S, void* stream);

/* Least squares min ||R x - b||_2 by CGLS on the GPU: pinv.cgls (pinv.py:23-48), the cross-check of the inverses,
 * batched over S sinograms that each stop on their own test ||R^T r_j|| <= tol ||R^T r_0||.
 * d_b: (S, 4, 2N-1, N) and d_x: (S, N, N) float64 on the device; x0 = 1 starts from d_x, 0 from zero.  iters (host,
 * S ints) receives the iterations each sinogram ran; hist (host, S x hist_len doubles, NULL when hist_len = 0) the
 * history ||R^T r_j||, j = 0..iters (the reference's return value).  Where R^T b = 0 the result is x = 0 after no
 * iteration (the reference divides 0 by 0 there).  d_work: adrt_cgls_work_bytes(n, S, hist_len) bytes.  Synchronises
 * the stream once per iteration, so it cannot be captured into a graph. */
unsigned long long adrt_cgls_work_bytes(int n, int S, int hist_len);
int adrt_cgls(int n, const double* d_b, double* d_x, int S, double tol, int maxit, int x0, void* d_work, int* iters,
              double* hist, int hist_len, void* stream);
/* adrt_cgls with pinv.cgls's per-iteration callback (pinv.py:23, 43-44): after iteration `it` (from 0) has run and
 * its stream has synchronised, callback(it, d_x, gnorm, ran, S, user) is called on the calling thread.  d_x is the
 * device iterate (S, N, N) after that iteration, valid until the callback returns; ran[s] = 1 for the sinograms that
 * ran the iteration (those that had not stopped), gnorm[s] their ||R^T r|| after it (0 where ran[s] = 0).  The
 * callback may read d_x (e.g. copy it to the host) but must not write it or enqueue work on the stream. */
typedef void (*adrt_cgls_callback)(int it, const double* d_x, const double* gnorm, const int* ran, int S, void* user);
int adrt_cgls_cb(int n, const double* d_b, double* d_x, int S, double tol, int maxit, int x0, void* d_work,
                 int* iters, double* hist, int hist_len, void* stream, adrt_cgls_callback callback, void* user);

#ifdef __cplusplus
}
#endif

#endif /* ADRT_HIP_H */
