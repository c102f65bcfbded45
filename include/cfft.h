/*
 * cfft.h -- C ABI of the MI355X (gfx950) concrete-fft engine.
 *
 * This is the drop-in boundary for zama-ai/concrete-fft v0.5.1.  Every entry
 * point names the reference item it replaces (path:line in the reference
 * crate).  Conventions:
 *   - every function returns an int status (cfft_status); none aborts.  A
 *     Rust shim maps a non-zero status to `panic!` with the message of
 *     cfft_last_error_message(), which reproduces the reference's panic text.
 *   - plans are immutable after creation and may be shared between threads
 *     (the reference's Plan is Clone + Send + Sync, src/ordered.rs:186,
 *     src/unordered.rs:495).
 *   - complex buffers are the reference's `[c64]` layout: interleaved
 *     {re, im} IEEE-754 doubles (num_complex::Complex64, src/lib.rs:84).
 *     A batch is `batch` transforms of `fft_size` values, transform b
 *     starting at element b*stride (stride >= fft_size, in c64 units).
 *   - device entry points take device pointers and a stream (hipStream_t,
 *     NULL = default stream) and are stream ordered; *_host entry points take
 *     host pointers, copy through the device and synchronise.
 *   - no torch types, no C++ types: plain pointers and sizes only.
 */
#ifndef CFFT_H
#define CFFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CFFT_ABI_VERSION 1

typedef struct cfft_plan_s *cfft_plan;
typedef void *cfft_stream_t; /* hipStream_t */
typedef void *cfft_event_t;  /* hipEvent_t */

typedef enum {
  CFFT_SUCCESS = 0,
  CFFT_ERROR_NOT_POWER_OF_TWO = 1, /* assert!(n.is_power_of_two()) */
  CFFT_ERROR_SIZE = 2,             /* n out of the plan's supported range */
  CFFT_ERROR_INVALID_BASE = 3,     /* unordered base_n rules, src/unordered.rs:664-669 */
  CFFT_ERROR_INVALID_ALGO = 4,
  CFFT_ERROR_NULL_POINTER = 5,
  CFFT_ERROR_LENGTH_MISMATCH = 6,  /* assert_eq!(self.fft_size(), buf.len()) */
  CFFT_ERROR_WRONG_PLAN_KIND = 7,  /* e.g. fwd_monomial on an ordered plan */
  CFFT_ERROR_OUT_OF_MEMORY = 8,
  CFFT_ERROR_DEVICE = 9,           /* HIP runtime error */
  CFFT_ERROR_INVALID_ARGUMENT = 10,
  CFFT_ERROR_SIZE_OVERFLOW = 11,   /* dyn_stack::SizeOverflow */
  CFFT_ERROR_DESERIALIZE = 12,     /* serde invalid_length / malformed stream */
  CFFT_ERROR_NO_DEVICE = 13,
  CFFT_ERROR_BUSY = 14             /* server admission refused */
} cfft_status;

/* ordered::FftAlgo, src/ordered.rs:26-45 (same discriminant order). */
typedef enum {
  CFFT_DIF2 = 0,
  CFFT_DIT2 = 1,
  CFFT_DIF4 = 2,
  CFFT_DIT4 = 3,
  CFFT_DIF8 = 4,
  CFFT_DIT8 = 5,
  CFFT_DIF16 = 6,
  CFFT_DIT16 = 7
} cfft_algo;

/* ordered::Method (src/ordered.rs:48-59) / unordered::Method
 * (src/unordered.rs:526-537).  MEASURE is served by this library's GPU
 * autotuner: each candidate plan is timed for about measure_seconds on a
 * private stream (no device-wide synchronisation) on the shapes it serves --
 * the forward and inverse transform and, for unordered plans of
 * n <= 8192, the fused negacyclic product -- and the choice is kept in the
 * tuning cache (cfft_tune_cache_*), which later MEASURE plans of the same
 * kind and size on the same device and stack reuse without timing.  HIP
 * failures while timing are returned by the plan-creation call. */
typedef enum { CFFT_METHOD_USER_PROVIDED = 0, CFFT_METHOD_MEASURE = 1 } cfft_method;

/* ---- status / diagnostics ---- */
const char *cfft_status_string(int status);
/* Thread-local text of the last failure on this thread, worded like the
 * reference's panic message for the same contract violation. */
const char *cfft_last_error_message(void);
int cfft_abi_version(void);

/* ---- device plumbing (no reference counterpart: the reference has no
 *      device boundary; these let a caller without a GPU runtime binding
 *      drive the library) ---- */
int cfft_device_count(int *count);
int cfft_set_device(int device);
int cfft_device_alloc(size_t bytes, void **ptr);
int cfft_device_free(void *ptr);
int cfft_device_mem_info(size_t *free_bytes, size_t *total_bytes);
int cfft_memcpy_htod(void *dst, const void *src, size_t bytes, cfft_stream_t stream);
int cfft_memcpy_dtoh(void *dst, const void *src, size_t bytes, cfft_stream_t stream);
int cfft_memcpy_dtod(void *dst, const void *src, size_t bytes, cfft_stream_t stream);
int cfft_memset(void *dst, int value, size_t bytes, cfft_stream_t stream);
int cfft_stream_create(cfft_stream_t *stream);
int cfft_stream_destroy(cfft_stream_t stream);
int cfft_stream_synchronize(cfft_stream_t stream);
int cfft_device_synchronize(void);
int cfft_event_create(cfft_event_t *ev);
int cfft_event_destroy(cfft_event_t ev);
int cfft_event_record(cfft_event_t ev, cfft_stream_t stream);
int cfft_event_synchronize(cfft_event_t ev);
int cfft_event_elapsed_ms(cfft_event_t start, cfft_event_t stop, float *ms);

/* ---- plans ---- */
/* ordered::Plan::new(n, method), src/ordered.rs:242-278: n a power of two,
 * n <= 2^10.  `algo` is used for USER_PROVIDED; MEASURE times all 8 algos on
 * the device for about `measure_seconds` each (Method::Measure(Duration)),
 * or reuses the tuning cache's choice. */
int cfft_ordered_plan_create(size_t n, int method, int algo, double measure_seconds,
                             cfft_plan *out);
/* The same ordered plan for 2^11 <= n <= 2^16 (and any smaller n).  The
 * reference documents the ordered plan up to 2^16 (src/ordered.rs:185) but
 * Plan::new asserts n < 2^11 (src/ordered.rs:244), so no reference output
 * exists at these sizes: the plan runs the reference's ordered recursion
 * (Stockham cores of the chosen radix, same twiddle table) extended to n,
 * checked against the closed-form DFT.  Up to 2^13 a transform runs in one
 * workgroup's LDS; above, each pass runs over global memory through a device
 * scratch of batch x n points from the library's stream-ordered cache.  n
 * above 2^16 returns CFFT_ERROR_SIZE.  cfft_ordered_plan_create keeps the
 * reference's assertion. */
int cfft_ordered_plan_create_large(size_t n, int method, int algo, double measure_seconds,
                                   cfft_plan *out);
/* unordered::Plan::new(n, method), src/unordered.rs:659-747: n a power of
 * two; base_n a power of two with base_n <= n, base_n <= 2^10 and
 * base_n >= 32 unless base_n == n.  MEASURE picks (base_algo, base_n) on the
 * device. */
int cfft_unordered_plan_create(size_t n, int method, int base_algo, size_t base_n,
                               double measure_seconds, cfft_plan *out);
/* fft128::Plan::new(n), src/fft128/mod.rs:1864-1885: n a power of two,
 * n >= 32 (and n <= 2^20 on the device).  Up to 2048 points (and the
 * 4096-point inverse) a transform runs in one workgroup's LDS; above, the
 * stages above the 2048-point blocks run over global memory (before the
 * blocks forward, after them inverse), with the same butterflies and so
 * the same bits. */
int cfft128_plan_create(size_t n, cfft_plan *out);
int cfft_plan_destroy(cfft_plan plan);

/* ---- tuning cache (no reference counterpart: the reference measures
 *      again in every process) ----
 * Entries map (device key, kind, n) to (algo, base_n); kind 0 = ordered,
 * 1 = unordered.  The device key names the device model, ISA,
 * compute-unit count, HIP runtime version and ABI version.  The default
 * file is $CFFT_TUNE_CACHE, else $XDG_CACHE_HOME/cfft_amd/tune-cache.tsv,
 * else $HOME/.cache/cfft_amd/tune-cache.tsv ("" or "off" in CFFT_TUNE_CACHE:
 * no file).  It is read once, at the first MEASURE lookup, and rewritten
 * (merged, through a rename) after each new measurement; a file that cannot
 * be written leaves the choice in the process.  The file is a text table:
 * one line per entry, device key, kind, n, algo and base_n separated by
 * tabs; '#' lines are comments; lines that do not parse or name a choice
 * no plan accepts are skipped. */
/* The default file's path ("" when disabled); `len` gets its length. */
int cfft_tune_cache_path(char *buf, size_t cap, size_t *len);
/* The current device's key (needs a device). */
int cfft_tune_device_key(char *buf, size_t cap, size_t *len);
/* Adds the entries of `path` to the process table (they replace entries
 * of the same key); `entries` gets the number of entries read. */
int cfft_tune_cache_import(const char *path, size_t *entries);
/* Writes the process table to `path`. */
int cfft_tune_cache_export(const char *path);
/* Empties the process table; the default file is not read again. */
int cfft_tune_cache_clear(void);
int cfft_tune_cache_size(size_t *entries);
/* found = 1 and (algo, base_n) when the process table holds the key. */
int cfft_tune_cache_lookup(const char *device_key, int kind, size_t n, int *found, int *algo,
                           size_t *base_n);

/* Plan::fft_size, src/ordered.rs:290 / src/unordered.rs:765 /
 * src/fft128/mod.rs:1896 */
int cfft_plan_fft_size(cfft_plan plan, size_t *n);
/* ordered Plan::algo (src/ordered.rs:304) -> (algo, n);
 * unordered Plan::algo (src/unordered.rs:783) -> (base_algo, base_n) */
int cfft_plan_algo(cfft_plan plan, int *algo, size_t *base_n);
/* Plan::fft_scratch (src/ordered.rs:320, src/unordered.rs:798): bytes and
 * alignment of the scratch the reference asks its caller for (n resp. base_n
 * c64, cache-line aligned).  The device path needs none (it works in LDS);
 * the figure is reported for drop-in fidelity. */
int cfft_plan_fft_scratch(cfft_plan plan, size_t *bytes, size_t *align);
/* Position of frequency i in the plan's permuted Fourier order:
 * idx[i] = bit_rev_twice(log2 n, log2 base_n, i) (src/unordered.rs:1046-1051).
 * Identity for ordered plans. */
int cfft_plan_permutation(cfft_plan plan, size_t *idx);

/* ---- f64 transforms ---- */
/* ordered Plan::fwd / Plan::inv (src/ordered.rs:342-373) and unordered
 * Plan::fwd / Plan::inv (src/unordered.rs:826-839, :927-940): in place,
 * unnormalised; unordered fwd leaves the result in the plan's permuted order
 * and inv expects it.  Results are bit-identical to the reference's f64
 * path for the same plan (algo, base_n). */
int cfft_fwd(cfft_plan plan, void *buf, size_t batch, size_t stride, cfft_stream_t stream);
int cfft_inv(cfft_plan plan, void *buf, size_t batch, size_t stride, cfft_stream_t stream);
int cfft_fwd_host(cfft_plan plan, double *buf, size_t batch);
int cfft_inv_host(cfft_plan plan, double *buf, size_t batch);
/* unordered Plan::fwd_monomial(degree, buf), src/unordered.rs:844-900:
 * forward transform of X^degree in the plan's permuted order. */
int cfft_fwd_monomial(cfft_plan plan, size_t degree, void *buf, cfft_stream_t stream);
int cfft_fwd_monomial_host(cfft_plan plan, size_t degree, double *buf);

/* ---- Fourier-buffer serde (feature "serde", src/unordered.rs:942-1036) ---- */
/* Gather a permuted-order buffer into standard order (what
 * serialize_fourier_buffer emits element by element) / scatter standard order
 * into the plan's permuted order (deserialize_fourier_buffer).  Device. */
int cfft_fourier_to_standard(cfft_plan plan, const void *permuted, void *standard, size_t batch,
                             cfft_stream_t stream);
int cfft_fourier_from_standard(cfft_plan plan, const void *standard, void *permuted,
                               size_t batch, cfft_stream_t stream);
/* Byte stream of serialize_fourier_buffer through bincode 1.3 `options()`
 * (varint sequence length, then n x (re, im) little-endian f64), as in
 * test_serde (src/unordered.rs:9411-9466).  Host buffers.  With out == NULL
 * only *out_len is set. */
int cfft_serialize_fourier_buffer(cfft_plan plan, const double *buf, uint8_t *out, size_t out_cap,
                                  size_t *out_len);
int cfft_deserialize_fourier_buffer(cfft_plan plan, const uint8_t *bytes, size_t len,
                                    double *buf);

/* ---- fft128 (feature "fft128", src/fft128/mod.rs) ---- */
/* fft128::Plan::fwd / inv (src/fft128/mod.rs:1905-1960): split double-double
 * arrays re0, re1, im0, im1 (hi/lo words), each batch x n, in place; fwd
 * leaves bit-reversed order, inv expects it; unnormalised.  Device. */
int cfft128_fwd(cfft_plan plan, double *re0, double *re1, double *im0, double *im1, size_t batch,
                cfft_stream_t stream);
int cfft128_inv(cfft_plan plan, double *re0, double *re1, double *im0, double *im1, size_t batch,
                cfft_stream_t stream);
int cfft128_fwd_host(cfft_plan plan, double *re0, double *re1, double *im0, double *im1,
                     size_t batch);
int cfft128_inv_host(cfft_plan plan, double *re0, double *re1, double *im0, double *im1,
                     size_t batch);
/* #[doc(hidden)] free functions negacyclic_{fwd,inv}_fft{,_scalar,_avxfma}
 * (src/fft128/mod.rs:351-1102), with caller-provided twiddles (device
 * pointers, n each; 8 <= n <= 2^20).  mul_variant 0 reproduces the scalar
 * path's double-double product, 1 the AVX2/AVX-512 path's fused form. */
int cfft128_negacyclic_fwd_fft(size_t n, double *re0, double *re1, double *im0, double *im1,
                               const double *tw_re0, const double *tw_re1, const double *tw_im0,
                               const double *tw_im1, size_t batch, int mul_variant,
                               cfft_stream_t stream);
int cfft128_negacyclic_inv_fft(size_t n, double *re0, double *re1, double *im0, double *im1,
                               const double *tw_re0, const double *tw_re1, const double *tw_im0,
                               const double *tw_im1, size_t batch, int mul_variant,
                               cfft_stream_t stream);
/* init_negacyclic_twiddles (src/fft128/mod.rs:1804-1828), host arrays of n. */
int cfft128_init_negacyclic_twiddles(size_t n, double *re0, double *re1, double *im0,
                                     double *im1);

/* ---- fused negacyclic products (caller-side in the reference: TFHE-rs
 *      builds them on unordered::Plan; the fold layout follows
 *      src/fft128/mod.rs:1996-2016) ---- */
/* C = A * B mod (X^N + 1), N = 2 * fft_size(plan), coefficients on the
 * 2^-64 torus: A u64 torus coefficients, B signed integer coefficients
 * (gadget digits), C u64 torus.  One kernel: fold, twist, two forward
 * transforms, pointwise product, inverse, untwist, 1/n and torus rounding.
 * Bit-identical to the reference-based CPU product for the same plan. */
int cfft_negacyclic_polymul_u64(cfft_plan plan, const uint64_t *a, const int64_t *b, uint64_t *c,
                                size_t batch, cfft_stream_t stream);
/* Exact C = A * B mod (X^N + 1, 2^64) for u64 A, B: balanced 16-bit limbs,
 * one f64 negacyclic transform per limb, per-weight Fourier products,
 * integer rounding and carries.  `margin` (optional, device, batch doubles)
 * receives max |x - nearest integer| over the rounded products of each
 * polynomial: the result is exact while it stays < 0.5.  The call does not
 * inspect the margins (that would synchronise the stream): a caller that
 * needs the guarantee passes `margin` and checks it; for 4 <= N <= 4096 it
 * stays far below 0.5 (measured 0.0098 at N = 4096 with every limb at
 * -2^15, the largest |D_3|). */
int cfft_exact_polymul_u64(cfft_plan plan, const uint64_t *a, const uint64_t *b, uint64_t *c,
                           double *margin, size_t batch, cfft_stream_t stream);
/* The same exact product over the double-double transform of an fft128 plan
 * of fft_size N/2 (the product the reference's fft128 test computes,
 * src/fft128/mod.rs:1996-2058): balanced 32-bit limbs, 4 forward transforms,
 * a0 b0 and a0 b1 + a1 b0 in the Fourier domain, 2 inverse transforms,
 * rounding and carries mod 2^64.  Exact for N <= 8192; `margin` as above. */
int cfft128_exact_polymul_u64(cfft_plan plan, const uint64_t *a, const uint64_t *b, uint64_t *c,
                              double *margin, size_t batch, cfft_stream_t stream);

/* ---- external product (TFHE GGSW x GLWE; caller-side in the reference) ---- */
/* Transform a GGSW key of k1*levels rows x k1 columns of u64 torus
 * polynomials (key: [row][col][N] u64, row = p * levels + i for GLWE poly p
 * and gadget level i) into the Fourier domain of `plan` (key_fourier:
 * [row][col][n] c64 in the plan's permuted order, the layout
 * fourier_to_standard / serialize_fourier_buffer take).  Device buffers. */
int cfft_ggsw_to_fourier(cfft_plan plan, const uint64_t *key, int k1, int levels, void *key_fourier,
                         cfft_stream_t stream);
/* out (batch x k1 polys of N u64) = GGSW(key_fourier) (x) GLWE(glwe, batch x
 * k1 polys): gadget decomposition (round to the nearest multiple of
 * 2^(64 - levels*base_log), ties up; balanced digits in [-2^(B-1), 2^(B-1)),
 * level 0 most significant), forward transforms, Fourier multiply-accumulate
 * against the resident key (acc += key * digit, rows in order from acc = 0),
 * inverse, untwist, torus rounding.  1 <= k1 <= 64, 1 <= base_log <= 32,
 * levels * base_log <= 64.  The digit spectra go through a device scratch
 * taken from the library's stream-ordered cache (allocated on first use).
 * Device buffers. */
int cfft_external_product(cfft_plan plan, const void *key_fourier, int k1, int levels,
                          int base_log, const uint64_t *glwe, uint64_t *out, size_t batch,
                          cfft_stream_t stream);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  cfft_external_product with the same arguments and the
 * same bits, allocating nothing: the spectra go through a fixed per-device
 * scratch of cfft_noalloc_scratch_bytes() (32 MiB) held until the process
 * ends, in passes of as many GLWEs as it holds (k1 * levels * N/2 * 16 bytes
 * each); CFFT_ERROR_SIZE if one GLWE's spectra exceed it.  Every call on one
 * device shares that device's scratch: calls that overlap on a device race
 * (serialize them, e.g. on one stream). */
int cfft_external_product_noalloc(cfft_plan plan, const void *key_fourier, int k1, int levels,
                                  int base_log, const uint64_t *glwe, uint64_t *out,
                                  size_t batch, cfft_stream_t stream);
size_t cfft_noalloc_scratch_bytes(void);

/* ---- uneven batch (no counterpart in the reference, which runs one plan
 *      per call): f64 negacyclic round trips of mixed N in one launch ---- */
/* A job: N = 2^log2_N real coefficients at in + in_offset (doubles) are
 * folded, twisted and forward transformed with its size class's unordered
 * plan into fourier + fourier_offset (N/2 c64, the plan's order); the inverse
 * transform and untwist (with 1/n) write N doubles at out + out_offset.
 * 5 <= log2_N <= 14. */
typedef struct {
  uint64_t in_offset;
  uint64_t fourier_offset;
  uint64_t out_offset;
  uint32_t log2_N;
  uint32_t reserved;
} cfft_uneven_job;
typedef struct cfft_uneven_s *cfft_uneven;
/* Builds a schedule for a job list (host array) over buffers of the stated
 * lengths (in and out in doubles, fourier in c64; every job is checked to lie
 * inside them).  method USER_PROVIDED: every size class uses base_algo with
 * base_n = min(n, 1024), and the Dif16 tiles run from the device queue;
 * MEASURE: each class uses the fastest of a candidate set of plans, timed on
 * the device for about measure_seconds each (the choice replaces
 * Method::Measure, src/unordered.rs:553-640, for this workload), then the
 * device queue and the static deal of the Dif16 tiles are timed on the whole
 * schedule for about measure_seconds each (on temporary buffers of the
 * stated lengths, when they fit in a quarter of free device memory) and the
 * faster is kept.  The timing runs on the schedule's private stream.  Blocks
 * until built. */
int cfft_uneven_create(const cfft_uneven_job *jobs, size_t njobs, size_t in_len,
                       size_t fourier_len, size_t out_len, int method, int base_algo,
                       double measure_seconds, cfft_uneven *out);
int cfft_uneven_destroy(cfft_uneven u);
/* Size classes whose plan is Dif16 with base_n = min(n, 1024) (N = 512 ..
 * 16384) run in one persistent launch of the register-resident Dif16 tile
 * body; the other classes in a second persistent launch of the generic body
 * over a device tile queue.  Dif16 tiles are taken from a device queue or
 * dealt round-robin to the workgroups (cfft_uneven_tile_schedule). */
#define CFFT_UNEVEN_DEVICE_QUEUE 0
#define CFFT_UNEVEN_STATIC_DEAL 1
/* The persistent launches replayed from a HIP graph captured for these
 * buffer addresses (re-captured when they change).  Runs of one schedule
 * must not overlap (they share its queue counters). */
int cfft_uneven_run(cfft_uneven u, const double *in, double *fourier, double *out,
                    cfft_stream_t stream);
/* The same launches issued directly on the stream (no graph; a caller may
 * capture them in its own graph: no allocation, copy or memset). */
int cfft_uneven_run_direct(cfft_uneven u, const double *in, double *fourier, double *out,
                           cfft_stream_t stream);
/* CFFT_UNEVEN_DEVICE_QUEUE or CFFT_UNEVEN_STATIC_DEAL: query, or override
 * the choice made at creation (results are the same bits either way). */
int cfft_uneven_tile_schedule(cfft_uneven u, int *kind);
int cfft_uneven_set_tile_schedule(cfft_uneven u, int kind);
/* The plain form: one launch per size class, one workgroup per tile. */
int cfft_uneven_run_plain(cfft_uneven u, const double *in, double *fourier, double *out,
                          cfft_stream_t stream);
/* The plan a size class runs with (which fixes its Fourier order). */
int cfft_uneven_class_algo(cfft_uneven u, uint32_t log2_N, int *algo, size_t *base_n);
int cfft_uneven_info(cfft_uneven u, size_t *njobs, size_t *ntiles, size_t *nclasses,
                     uint32_t *grid);

/* ---- request server (no counterpart in the reference): continuous batching
 *      of negacyclic products and external products on one device ---- */
/* A dispatcher thread groups waiting requests by (kind, N, key) in arrival
 * order and launches a batch whenever one of `streams` launch slots is free
 * and the device pool (hbm_budget_bytes, carved once from free device
 * memory) and the pinned staging pool have room for it; requests that do not
 * fit wait for running batches to release memory.  A batch's device block
 * holds its inputs, its outputs and, for a small external-product batch, the
 * Fourier scratch of the split form, so serving allocates nothing per call.  Batches on different
 * streams overlap on the device.  Each result is bit-identical to the same
 * request computed alone (the batched kernels treat products independently). */
typedef struct {
  size_t hbm_budget_bytes; /* 0: min(1 GiB, free device memory / 4) */
  size_t pinned_bytes;     /* 0: same as hbm_budget_bytes */
  uint32_t max_batch;      /* requests per launch, 0: 1024 */
  uint32_t streams;        /* launch slots, 0: 4 (at most 16) */
} cfft_server_config;
typedef enum { CFFT_REQUEST_POLYMUL = 0, CFFT_REQUEST_EXTERNAL_PRODUCT = 1 } cfft_request_kind;
/* OR-ed into cfft_request.kind: in0, in1 and out are device buffers of the
 * server's device (8-byte aligned; checked at submission).  The batch's
 * inputs are gathered from them and its results scattered back by kernels
 * on the device, so nothing but a table of addresses crosses the bus: a
 * chain of requests on one device accumulator (a CMux chain of blind
 * rotation) stays on the device.  The inputs must be complete when the
 * request is submitted (the server does not wait on the caller's streams);
 * the output is complete when the ticket is done.  Device requests batch
 * with device requests of the same kind, N and key, up to 32767 a batch. */
#define CFFT_REQUEST_DEVICE_BUFFERS 0x100
/* Host buffers (or device buffers, with CFFT_REQUEST_DEVICE_BUFFERS), owned
 * by the caller and valid until the request completes.
 * POLYMUL: in0 = A (N u64 torus), in1 = B (N i64), out = N u64 (as
 * cfft_negacyclic_polymul_u64 with a Dif16 plan, base_n = min(N/2, 1024)).
 * EXTERNAL_PRODUCT: key_id from cfft_server_add_key, in0 = GLWE (k1 x N
 * u64), out = k1 x N u64 (as cfft_external_product). */
typedef struct {
  int kind;
  int key_id;
  size_t N;
  const void *in0;
  const void *in1;
  void *out;
} cfft_request;
typedef struct {
  uint64_t batches;
  uint64_t completed;
  uint64_t admission_waits; /* batch formations deferred for lack of free memory */
  uint64_t max_inflight;
  size_t hbm_budget_bytes;
  uint32_t streams;
} cfft_server_statistics;
typedef struct cfft_server_s *cfft_server;
int cfft_server_create(const cfft_server_config *config, cfft_server *out);
/* Drains submitted requests, then stops the dispatcher and frees everything. */
int cfft_server_destroy(cfft_server s);
/* Registers a GGSW key (host, [row][col][N] u64 torus) for external products
 * of size N; it is transformed once and stays resident in the Fourier domain. */
int cfft_server_add_key(cfft_server s, size_t N, int k1, int levels, int base_log,
                        const uint64_t *key, int *key_id);
/* Non-blocking; CFFT_ERROR_BUSY if the request alone exceeds the budget. */
int cfft_server_submit(cfft_server s, const cfft_request *request, uint64_t *ticket);
int cfft_server_poll(cfft_server s, uint64_t ticket, int *done);
/* Blocks until the request is done and its output written; forgets the ticket.
 * One thread at a time may wait on a ticket: a second waiter gets
 * CFFT_ERROR_INVALID_ARGUMENT. */
int cfft_server_wait(cfft_server s, uint64_t ticket);
/* Gives up a ticket without waiting (for callers that poll): forgotten now if
 * its request is done, else as soon as it completes; its output is still
 * written.  A ticket somebody waits on cannot be released. */
int cfft_server_release(cfft_server s, uint64_t ticket);
/* Open-loop replay: submits request i at arrival_s[i] seconds after the start
 * (from this thread), waits for all, and stores each latency from its
 * scheduled arrival to its completion in latency_s[i]. */
int cfft_server_replay(cfft_server s, const cfft_request *requests, size_t n,
                       const double *arrival_s, double *latency_s);
int cfft_server_stats(cfft_server s, cfft_server_statistics *stats);

#ifdef __cplusplus
}
#endif
#endif
