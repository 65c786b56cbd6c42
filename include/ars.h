/*
 * ars.h -- C ABI of the MI355X (gfx950) sharded additive-FFT Reed-Solomon
 * codec.  Drop-in replacement for the sharded encode / erasure-recover path of
 * the `rs` crate of vkomenda/additive-fft-reed-solomon.
 *
 * Code and layout are the reference's:
 *   - GF(2^8) modulo 0x11d, evaluation points omega_0..omega_{n-1} spanned by
 *     the Cantor basis, LCH novel-polynomial-basis FFT (rs/src/kernel/mod.rs:7-24);
 *   - a codeword is n shards of shard_len bytes, shard i at [i*L, (i+1)*L);
 *     parity shards 0..T-1 first, message shards T..n-1 after (T = n - k)
 *     (rs/src/poly_11d_lut.rs:100-121, rs/benches/recover_erasure_shards.rs:100-101).
 *
 * (n, k) are run-time arguments (the reference takes N, T as const generics,
 * rs/src/codec.rs:11).  n must be a power of two in [2, 256] and 1 <= k < n.
 * The reference's contract also makes T = n - k a power of two
 * (rs/src/lib.rs:16-19); codes whose T is not a power of two but whose k is
 * (e.g. n = 256, k = 16) are accepted as an extension: they compute the same
 * code (the unique systematic codeword with deg f < k) by a different route.
 *
 * All pointers named d_* are device pointers (HBM).  shard_len may be any
 * value in [1, 4 GiB); shard_len % 4 == 0 with 4-byte aligned pointers takes
 * the dword path.  Codewords of 4 GiB and more (n * shard_len >= 2^32) run on
 * the dword path only (ars_encode, ars_recover, ars_recover_with_plan,
 * ars_recover_noalloc, ars_recompute_data: ARS_E_UNSUPPORTED otherwise) and on
 * any buffers through the *_host variants; ars_decode, batches and the server
 * keep n * shard_len below 4 GiB (ARS_E_INVAL).  No entry point allocates on the hot path except
 * ars_recover (a 6.4 KB erasure plan buffer, allocated on a context's first
 * call on each stream and kept until ars_destroy; per call, stream-ordered,
 * on hipStreamPerThread, under graph capture and past 64 streams; launches of
 * at most one workgroup per compute unit need none: their workgroups compute
 * the plan themselves, in one launch) and the *_host variants (staging
 * buffers).
 *
 * Return codes: >= 0 success (ars_recover: 1 = the reference's `true`,
 * 0 = `false` because more than T erasures), < 0 = ARS_E_* below, for inputs
 * the reference would panic on or silently mis-compute.
 */
#ifndef ARS_H
#define ARS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ARS_OK 0
#define ARS_TRUE 1
#define ARS_FALSE 0
#define ARS_E_INVAL (-1)       /* bad (n, k), shard_len == 0, or bad flags */
#define ARS_E_POSITION (-2)    /* erasure position >= n */
#define ARS_E_DUPLICATE (-3)   /* repeated erasure position */
#define ARS_E_NULL (-4)        /* null pointer where data is required */
#define ARS_E_HIP (-5)         /* HIP runtime error (message: ars_last_error) */
#define ARS_E_NOMEM (-6)       /* device or host allocation failed */
#define ARS_E_UNSUPPORTED (-7) /* neither T nor k is a power of two, or a codeword of
                                  4 GiB or more off the dword path */
#define ARS_E_BUDGET (-8)      /* serving: request exceeds the HBM budget */
#define ARS_E_STATE (-9)       /* serving / batch object used out of order */

/* Encode flags. */
#define ARS_ENC_ACCUMULATE 1u  /* XOR the result into parity the way the
                                  reference does (rs/src/codec.rs:259-264):
                                  parity_out = FFT_T(parity_in ^ sum_i IFFT_T(chunk_i)).
                                  Default: overwrite, equal to the reference
                                  called with zeroed parity. */

typedef struct ars_ctx ars_ctx;

/* Create a context on HIP device `device`: uploads the twiddle tables and
 * keeps them resident.  One context may be used from several host threads on
 * different streams. */
int ars_create(int device, ars_ctx **out);
int ars_destroy(ars_ctx *ctx);
const char *ars_last_error(void);
const char *ars_version(void);

/* Codec::encode_systematic_sharded (rs/src/codec.rs:241-265).
 * d_message: k*L bytes (message shards T..n-1), d_parity: T*L bytes.
 * The reference's `workspace` argument is not needed (tiles live in LDS).
 * Asynchronous on `stream` (a hipStream_t; NULL = default stream). */
int ars_encode(ars_ctx *ctx, uint32_t n, uint32_t k, const uint8_t *d_message,
               uint8_t *d_parity, size_t shard_len, uint32_t flags, void *stream);

/* Codec::recover_erasure_shards (rs/src/codec.rs:479-569).
 * d_received: n*L bytes, in place.  positions: HOST array of e shard indices
 * (any order).  Writes every erased shard p; as in the reference, whatever p
 * held before is XORed into the result (zero-filled erasures come back as
 * the original shard).  Returns 1, 0 if e > n - k (received untouched), or
 * ARS_E_*.  Asynchronous on `stream`.  Calls on one stream, from any number
 * of host threads, share the context's plan buffer for that stream; their
 * kernels run in issue order. */
int ars_recover(ars_ctx *ctx, uint32_t n, uint32_t k, uint8_t *d_received, size_t shard_len,
                const uint8_t *positions, uint32_t e, void *stream);

/* Two-step recover for callers that reuse one erasure pattern or want the
 * per-pattern work off the critical path (serving, batching, timing the
 * main kernel alone).  ars_recover_plan validates the pattern exactly like
 * ars_recover and writes, on `stream`, a device-resident plan of
 * ars_recover_plan_size() bytes to d_plan: per shard, the multiplier
 * Pi(omega_i) = prod_{p in E} (omega_i + omega_p) or, for erased shards,
 * 1 / Pi'(omega_i) -- the reference's erasure locator / denominators
 * (rs/src/codec.rs:436-465, 527-536) in evaluated form.
 * ars_recover_with_plan then runs the fused recover for that pattern; the
 * plan must come from ars_recover_plan with the same (n, k).  Return codes
 * as for ars_recover (the plan call returns 0 for e > T and 1 otherwise;
 * with_plan returns 1). */
size_t ars_recover_plan_size(void);
int ars_recover_plan(ars_ctx *ctx, uint32_t n, uint32_t k, const uint8_t *positions, uint32_t e,
                     void *d_plan, void *stream);
int ars_recover_with_plan(ars_ctx *ctx, uint32_t n, uint32_t k, uint8_t *d_received,
                          size_t shard_len, const void *d_plan, void *stream);

/* ars_recover without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  Same arguments,
 * semantics and return codes as ars_recover; it calls no hipMalloc, hipFree or
 * allocating library routine.  Launches of more than one workgroup per
 * compute unit put the erasure plan in a scratch array of the library (one
 * plan, 6.5 KiB, for n <= 256; larger n is rejected with ARS_E_INVAL by the
 * (n, k) check) that each GPU holds until the process ends.  Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (other streams,
 * other host threads, other contexts on that device) race, so serialize them
 * on one stream.  Each call's plan kernel rewrites every entry its recover
 * kernel reads, so nothing carries over from the previous call. */
int ars_recover_noalloc(ars_ctx *ctx, uint32_t n, uint32_t k, uint8_t *d_received, size_t shard_len,
                        const uint8_t *positions, uint32_t e, void *stream);

/* Codec::recompute_data_from_parity_sharded (rs/src/codec.rs:321-340):
 * n = 2T only; rewrites the message shards T..n-1 from the parity shards. */
int ars_recompute_data(ars_ctx *ctx, uint32_t n, uint32_t k, uint8_t *d_received,
                       size_t shard_len, void *stream);

/* ------------------------------------------------------------------------
 * Sharded error decoding.  The reference decodes errors (not only erasures)
 * one received word at a time, Codec::decode_systematic_scalar
 * (rs/src/codec.rs:341-420), and has no sharded form (README.md:58).  Here
 * every byte column of the n x L codeword is one received word, decoded in
 * place exactly as decode_systematic_scalar decodes it: left as it is when
 * its syndrome is zero or when the reference returns false, corrected
 * otherwise.  The syndromes, the key equation of each column with a non-zero
 * syndrome, the root search and the error values all run on the GPU; the call
 * returns after one synchronisation of `stream` (for the counts).
 *
 * T = n - k must be a power of two >= 2 (the reference's decoder; its HGCD
 * needs t_log >= 1).  flags: 0 reproduces the reference byte for byte,
 * including its error values, which use poly::deriv_lnh
 * (rs/src/poly_arith.rs:91-116) where the formal derivative belongs and are
 * wrong for many patterns of up to T/2 errors; ARS_DEC_FORMAL_DERIVATIVE uses
 * the formal derivative and corrects every pattern of up to T/2 errors.
 * ARS_DEC_HOST_KEY_EQUATION solves the key equations on the host instead of
 * the GPU (same results; for cross-checking).
 * d_status (optional, device, L bytes): per column 0 syndrome zero, 1 the
 * reference's true after the key equation, 2 the reference's false.
 * counts (optional, host, 3 values): columns with status 0, 1, 2.
 * Returns ARS_TRUE when no column returned false, ARS_FALSE otherwise. */
#define ARS_DEC_FORMAL_DERIVATIVE 1u
#define ARS_DEC_HOST_KEY_EQUATION 2u
int ars_decode(ars_ctx *ctx, uint32_t n, uint32_t k, uint8_t *d_received, size_t shard_len,
               uint32_t flags, uint8_t *d_status, uint64_t *counts, void *stream);

/* ------------------------------------------------------------------------
 * Batches: many codewords of uneven (n, k), shard lengths and erasure
 * patterns in one persistent kernel launch, replayed from a hipGraph.
 * No counterpart in the reference (it processes one codeword per call); each
 * job has exactly the semantics of ars_encode / ars_recover on its own.
 * ------------------------------------------------------------------------ */
#define ARS_OP_ENCODE 1u
#define ARS_OP_RECOVER 2u

typedef struct ars_job {
    uint32_t op;               /* ARS_OP_ENCODE or ARS_OP_RECOVER */
    uint32_t n, k;
    uint32_t flags;            /* encode: ARS_ENC_ACCUMULATE */
    uint64_t shard_len;
    const uint8_t *d_message;  /* encode: k * shard_len bytes (device) */
    uint8_t *d_parity;         /* encode: (n - k) * shard_len bytes (device) */
    uint8_t *d_codeword;       /* recover: n * shard_len bytes (device), in place */
    const uint8_t *positions;  /* recover: e erased shard indices (host) */
    uint32_t e;
    int32_t status;            /* out: what ars_encode / ars_recover would return */
} ars_job;

typedef struct ars_batch ars_batch;

typedef struct ars_batch_info {
    uint32_t jobs, items, segments, plans, fallback_jobs, grid;
    uint64_t algorithmic_bytes;  /* sum over jobs of read + written shard bytes */
} ars_batch_info;

#define ARS_BATCH_NO_GRAPH 1u  /* launch the kernels directly instead of replaying the graph */

/* A batch object owns a device arena for descriptors and erasure plans and a
 * captured graph; it serves one run at a time: a run's descriptor upload
 * waits, on the device, for the previous run on the same object, whatever
 * stream that used (use several objects for overlap).  Calls on one object
 * from several host threads at once are not supported. */
int ars_batch_create(ars_ctx *ctx, uint32_t max_jobs, ars_batch **out);
int ars_batch_destroy(ars_batch *b);
/* Validates every job (setting its status), then launches all valid jobs as
 * one batch on `stream`, asynchronously.  Jobs the batched kernel cannot take
 * (shard_len % 4 != 0 or unaligned pointers) run as plain calls on the same
 * stream.  Returns ARS_OK or a negative code for a failure of the batch
 * itself (the per-job failures are in the statuses). */
int ars_batch_run(ars_batch *b, ars_job *jobs, uint32_t count, uint32_t flags, void *stream);
int ars_batch_info_get(const ars_batch *b, ars_batch_info *info);
/* Persistent workgroups per compute unit (0 = occupancy maximum); the
 * autotuner sets this. */
int ars_batch_set_blocks_per_cu(ars_batch *b, uint32_t blocks_per_cu);
/* The same jobs as one ars_encode / ars_recover call each, in order, on
 * `stream` (the one-codeword-per-launch form the batch is measured against). */
int ars_run_jobs_plain(ars_ctx *ctx, ars_job *jobs, uint32_t count, void *stream);

/* ------------------------------------------------------------------------
 * Serving: a request loop with continuous batching.  No counterpart in the
 * reference.  Requests (encode / recover of any supported shape) arrive at
 * any time from any thread; a dispatcher thread admits them against an HBM
 * budget (a device pool that holds each admitted request's codeword), groups
 * whatever is admitted into a batch and launches it on one of
 * `pipeline_depth` streams as soon as one is free, so requests that arrive
 * while others run join the next batch and overlap with it on the device.
 * Each request gets exactly the single-call result.
 * ------------------------------------------------------------------------ */
typedef struct ars_server ars_server;

typedef struct ars_server_config {
    uint64_t hbm_budget;       /* bytes of device pool for request codewords (0 = 64 GiB) */
    uint32_t max_batch_jobs;   /* jobs per launched batch (0 = 1024) */
    uint32_t pipeline_depth;   /* batches in flight at once, one stream each (0 = 2) */
    uint64_t direct_min_bytes; /* requests whose codeword (n * shard_len) has at least this
                                  many bytes run as plain ars_encode / ars_recover calls on
                                  the batch's stream instead of in the batched kernel, where
                                  launch overhead is negligible against their work
                                  (0 = 16 MiB; UINT64_MAX = never) */
    uint32_t min_batch_jobs;   /* a batch with fewer jobs left for the kernel runs them as
                                  plain calls too, its fixed cost exceeding theirs (0 = 4
                                  while another batch is in flight, 16 on an otherwise idle
                                  GPU; 1 = always use the kernel); never more than
                                  max_batch_jobs */
    uint32_t reserved;         /* 0 */
} ars_server_config;

typedef struct ars_request {
    uint32_t op;               /* ARS_OP_ENCODE or ARS_OP_RECOVER */
    uint32_t n, k, flags;
    uint64_t shard_len;
    const uint8_t *positions;  /* recover: e shard indices (host; copied at submit) */
    uint32_t e;
    uint32_t pad;
    const uint8_t *host_in;    /* encode: message k*L; recover: codeword n*L; copied in on
                                  admission.  NULL: the pool slot keeps whatever it holds
                                  (synthetic load). */
    uint8_t *host_out;         /* encode: parity T*L; recover: codeword n*L; copied out on
                                  completion.  NULL: no copy back. */
} ars_request;

typedef struct ars_server_stats {
    uint64_t submitted, completed, rejected;
    uint64_t batches;                /* slot launches that ran the batched kernel */
    uint64_t algorithmic_bytes;      /* over completed requests */
    double p50_latency_s, p99_latency_s, max_latency_s;
    double mean_batch_jobs;          /* jobs per batched-kernel launch */
    double first_arrival_s, last_completion_s;  /* on the server's clock */
    uint64_t batched_jobs;           /* jobs run in the batched kernel */
    uint64_t direct_jobs;            /* jobs run as plain calls */
} ars_server_stats;

int ars_server_create(ars_ctx *ctx, const ars_server_config *cfg, ars_server **out);
/* Stops the dispatcher after draining every submitted request. */
int ars_server_destroy(ars_server *s);
/* Thread-safe.  Host buffers must stay valid until the request completes.
 * Returns ARS_OK and a ticket, or ARS_E_BUDGET if the request alone exceeds
 * the budget (never admissible). */
int ars_server_submit(ars_server *s, const ars_request *req, uint64_t *ticket);
/* Blocks until the ticket completes; status is what ars_encode / ars_recover
 * would return; latency = completion - submission, in seconds. */
int ars_server_wait(ars_server *s, uint64_t ticket, int32_t *status, double *latency_s);
/* Blocks until every submitted request has completed. */
int ars_server_drain(ars_server *s);
/* Latency percentiles over requests completed since the last reset. */
int ars_server_stats_get(ars_server *s, ars_server_stats *out);
int ars_server_stats_reset(ars_server *s);
/* Load generator: submits reqs[i] at arrival_s[i] seconds after the call
 * (from a thread of its own, so arrivals do not wait on completions), waits
 * for all, and returns per-request statuses and latencies. */
int ars_server_replay(ars_server *s, const ars_request *reqs, const double *arrival_s, uint32_t count,
                      int32_t *statuses, double *latencies_s);

/* The plain-call baseline for ars_server_replay, with no counterpart in the
 * reference: the same trace served by one call per request, in arrival order
 * on `stream` (input copy if host_in, ars_encode / ars_recover, output copy if
 * host_out), without batching.  Request i is issued at arrival_s[i] seconds
 * after the call; a second thread timestamps completions.  d_buffer
 * (buffer_bytes of device memory, at least n * shard_len of every request)
 * holds one request's codeword at a time: the stream runs the requests in
 * order.  statuses[i] = what ars_encode / ars_recover returned (or the copy's
 * error); latencies_s[i] = completion - arrival, queueing included. */
int ars_plain_replay(ars_ctx *ctx, const ars_request *reqs, const double *arrival_s, uint32_t count,
                     uint8_t *d_buffer, size_t buffer_bytes, int32_t *statuses, double *latencies_s,
                     void *stream);

/* ------------------------------------------------------------------------
 * Autotuner.  For transforms of 32+ points the fused kernels come in tile
 * variants: column quads per workgroup (16, 32 or 64) and LDS exchange buffers
 * (1: half the LDS, extra barriers; 2: one barrier per exchange, and for
 * recover the formal derivative's low-bit terms taken without an LDS copy).
 * ars_autotune times every variant for one (n, k, shard_len) on scratch
 * buffers (three interleaved rounds, the fastest round of each) and keeps the
 * fastest in the context for that operation and transform size, or the
 * default variant when none beats it by more than 2%.
 * ars_batch_autotune picks the persistent grid (workgroups
 * per compute unit) for a representative batch.
 * ------------------------------------------------------------------------ */
typedef struct ars_tune_result {
    uint32_t op;          /* ARS_OP_ENCODE or ARS_OP_RECOVER */
    uint32_t log2_size;   /* transform size: T for encode, n for recover */
    uint32_t quads;       /* chosen column quads per workgroup (0: not tunable) */
    uint32_t buffers;     /* chosen LDS exchange buffers */
    double best_us;       /* kernel time of the chosen variant (mean of 5 launches, fastest round) */
    double default_us;    /* the same for the default variant */
} ars_tune_result;

int ars_autotune(ars_ctx *ctx, uint32_t n, uint32_t k, size_t shard_len, ars_tune_result *encode_out,
                 ars_tune_result *recover_out, void *stream);
int ars_tune_set(ars_ctx *ctx, uint32_t op, uint32_t log2_size, uint32_t quads, uint32_t buffers);
/* Runs the batch 16 times on the jobs' own buffers (encode rewrites the
 * parity, recover the erased shards, with the same result every time) and
 * keeps the fastest blocks-per-CU setting in the batch object. */
int ars_batch_autotune(ars_batch *b, ars_job *jobs, uint32_t count, uint32_t *best_blocks_per_cu,
                       void *stream);

/* Host-slice variants: same semantics on host buffers, synchronous.  They
 * stream the slices through the GPU in column chunks over pinned staging
 * (PCIe-bound, not the HBM roofline path). */
int ars_encode_host(ars_ctx *ctx, uint32_t n, uint32_t k, const uint8_t *message,
                    uint8_t *parity, size_t shard_len, uint32_t flags);
int ars_recover_host(ars_ctx *ctx, uint32_t n, uint32_t k, uint8_t *received, size_t shard_len,
                     const uint8_t *positions, uint32_t e);

#ifdef __cplusplus
}
#endif

#endif /* ARS_H */
