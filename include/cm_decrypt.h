/*
 * cm_decrypt.h -- C ABI of the MI355X Classic McEliece batched decryption engine.
 *
 * Plain C types only (pointers, sizes, integers).  All buffers are caller-owned.
 * Every function returns CMD_OK (0) or a negative error code; the library never
 * calls exit().  A context may be used from several host threads (calls are
 * serialised by an internal lock); a server object is thread-safe.
 *
 * Reference interface replaced by each entry point (paths relative to the
 * reference repository):
 *   - the file contract of the drivers: ct_<p>.bin (KATNUM x CIPHERTEXTBYTES, raw),
 *     sk_<p>.bin (SECRETKEYBYTES), errorstream0_<p>.bin (text)
 *     (CPU/param/param6960119/src/util.c:30-58, decrypt.c:55-66, 138);
 *   - the C functions initialisation / compute_inverses / synd / bm / root /
 *     support_gen (CPU/include/common/root.h:10-25, bm.h:13, GPU_Baseline/include/common/root.h:13-28);
 *   - the GPU drivers' batch loop decrypt()/decrypt_all()
 *     (GPU_Baseline/param/param8192128/src/cuda/Decrypt.cu:177-324,
 *      GPU_Baseline/param/param6960119/src/cuda/decrypt.cu:259-434,
 *      GPU_Optimised/param/param8192128/src/cuda/Decrypt.cu:220-314).
 */
#ifndef CM_DECRYPT_H
#define CM_DECRYPT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------------------- */
#define CMD_OK             0
#define CMD_ERR_FILE      (-1)   /* KAT_FILE_OPEN_ERROR  (CPU/param/param6960119/include/common.h:11) */
#define CMD_ERR_CRYPTO    (-4)   /* KAT_CRYPTO_FAILURE   (common.h:12): bad KATNUM / bad key bytes */
#define CMD_ERR_PARAM     (-10)  /* unknown parameter set id */
#define CMD_ERR_ARG       (-11)  /* invalid argument (null pointer, stride too small, ...) */
#define CMD_ERR_HIP       (-12)  /* HIP runtime error (message in cmd_last_error) */
#define CMD_ERR_NODEV     (-13)  /* no usable GPU */
#define CMD_ERR_BUDGET    (-14)  /* HBM budget of the context would be exceeded */
#define CMD_ERR_KEY       (-15)  /* unknown or unloaded key id, or key of another parameter set */
#define CMD_ERR_FULL      (-16)  /* server queue full */
#define CMD_ERR_NOMEM     (-17)  /* host allocation failed */
#define CMD_ERR_SINGULAR  (-18)  /* generator: key has no systematic form */
#define CMD_ERR_PENDING   (-19)  /* server: request not finished yet */
#define CMD_ERR_TIMEOUT   (-20)  /* server: wait timed out */

/* Text for an error code; cmd_last_error() gives the calling thread's last detail. */
const char *cmd_strerror(int code);
const char *cmd_last_error(void);

/* ---- parameter sets ---------------------------------------------------------------- */
/* ids: 348864, 460896, 6688128, 6960119, 8192128 (CPU/param/param<p>/include/common.h) */
typedef struct {
    int32_t id, m, n, t;
    int32_t ct_bytes;    /* crypto_kem_CIPHERTEXTBYTES == SYND_BYTES */
    int32_t sb;          /* syndrome input bits, 8 * SYND_BYTES */
    int32_t sk_bytes;    /* crypto_kem_SECRETKEYBYTES */
    int32_t cond_bytes;  /* COND_BYTES */
    int32_t e_bytes;     /* n / 8 */
} cmd_param_info;

int cmd_param_get(int32_t param_id, cmd_param_info *out);

/* ---- host-only key handling (no GPU needed) ------------------------------------------ */

/* support_gen(gf *s, const unsigned char *c): CPU/src/common/root_13.c:181-211,
 * CPU/param/param348864/src/root.c:336-385.  L receives n elements. */
int cmd_support_gen(int32_t param_id, const uint8_t *cond, uint16_t *L);

/* initialisation() key part + compute_inverses(): CPU/param/param6960119/src/util.c:60-100.
 * g: t+1 coefficients (monic), L: n, table: sb x 2t (row-major); any may be NULL. */
int cmd_key_expand_host(int32_t param_id, const uint8_t *sk, size_t sk_len,
                        uint16_t *g, uint16_t *L, uint16_t *table);

/* write_error_positions (CPU/param/param6960119/src/decrypt.c:55-66) for `count` error
 * vectors of n/8 bytes, e_stride bytes apart: " %d" per set bit ascending, '\n' per vector.
 * Writes at most cap bytes to buf; *written receives the full length (call with cap 0 to size). */
int cmd_errorstream_format(int32_t param_id, const uint8_t *e, size_t e_stride, size_t count,
                           char *buf, size_t cap, size_t *written);
/* The same text from CMD_MODE_POSITIONS rows (t uint16 each, pos_stride bytes apart; 0xFFFF
 * ends a row early). */
int cmd_errorstream_format_positions(int32_t param_id, const uint16_t *pos, size_t pos_stride, size_t count,
                                     char *buf, size_t cap, size_t *written);

/* Writes (append != 0: appends) the errorstream file; the reference drivers reset it first
 * (decrypt.c:138) or append per ciphertext (CPU/param/param8192128/src/Decrypt.c:161). */
int cmd_errorstream_write(int32_t param_id, const char *path, const uint8_t *e, size_t e_stride,
                          size_t count, int append);
int cmd_errorstream_write_positions(int32_t param_id, const char *path, const uint16_t *pos,
                                    size_t pos_stride, size_t count, int append);

/* ---- seeded test-vector generator (replaces kem/: keypair, gen_e, syndrome) --------------
 * kem/common/core/operations.c:36-107 and encrypt.c:40-120 define the formats; the
 * randomness here is this library's own seeded generator, not the NIST DRBG. */

/* Writes sk_bytes to sk: 32 B seed | ff ff ff ff 00 00 00 00 | g (t x 2 B LE) | Benes
 * condition bits of a seeded permutation | n/8 B s.  g is a seeded monic irreducible
 * polynomial; keys whose first m*t parity-check columns are singular are redrawn. */
int cmd_gen_keypair(int32_t param_id, uint64_t seed, uint8_t *sk);

/* `count` ciphertexts for key sk, each with a uniformly random error of exactly `weight`
 * distinct positions in [0, n) (weight <= t; weight < 0 means t).  ct: count x ct_bytes,
 * e: count x e_bytes (the error vectors, the expected decoder output for weight == t). */
int cmd_gen_encrypt(int32_t param_id, const uint8_t *sk, size_t sk_len, uint64_t seed,
                    size_t count, int32_t weight, uint8_t *ct, uint8_t *e);

/* Benes condition bits for a permutation pi of {0..2^m-1} (controlbitsfrompermutation,
 * kem/common/core/controlbits.c:180-220): applying the network to the identity yields pi,
 * and support_gen over these bits gives L[i] = bitrev(pi[i]).  cond: (2m-1)*2^(m-4) bytes. */
int cmd_controlbits(int32_t m, const int16_t *pi, uint8_t *cond);

/* Many keys at once with `threads` host threads: sk_out holds n_keys x sk_bytes; key k uses
 * seed base_seed + k (skipping to later seeds like cmd_gen_keypair when needed). */
int cmd_gen_keys(int32_t param_id, uint64_t base_seed, size_t n_keys, uint8_t *sk_out, int threads);

/* ---- GPU context and resident keys -------------------------------------------------- */
typedef struct cmd_ctx cmd_ctx;

#define CMD_MODE_PLAIN 0   /* three launches per key group: synd, bm, chien (uint16 GF) */
#define CMD_MODE_FUSED 1   /* one persistent synd->bm->chien launch for any mix of sets/keys */
#define CMD_MODE_KEM   2   /* fused decryption + decapsulation check + SHAKE256: the output row
                              of an item is its 32-byte session key (crypto_kem_dec) */
#define CMD_KEM_KEY_BYTES 32
#define CMD_MODE_POSITIONS 3  /* fused decryption; the output row of an item is its error
                                 positions: t uint16 ascending, padded with 0xFFFF (2t bytes,
                                 the errorstream's numbers without the n/8-byte bitmap) */

/* device: HIP ordinal; hbm_budget_bytes: 0 = 90% of the free device memory. */
int cmd_ctx_create(int32_t device, uint64_t hbm_budget_bytes, cmd_ctx **out);
int cmd_ctx_destroy(cmd_ctx *ctx);
/* Device memory charged against the budget (keys + work buffers) and the budget. */
int cmd_ctx_memory(cmd_ctx *ctx, uint64_t *used, uint64_t *budget);
/* Changes the HBM budget (admission control) of a live context; nothing already allocated
 * is freed, later allocations that would exceed it fail with CMD_ERR_BUDGET.  No counterpart
 * in the reference, which allocates fixed global arrays. */
int cmd_ctx_set_budget(cmd_ctx *ctx, uint64_t hbm_budget_bytes);

/* Uploads one key's Goppa coefficients and control bits and expands them on the GPU
 * (support, syndrome table, bit planes, Benes masks); the key stays resident until unloaded.
 * Replaces initialisation() + compute_inverses() + InitializeC()
 * (GPU_Baseline/param/param8192128/include/decrypt.h:147-173). */
int cmd_key_load(cmd_ctx *ctx, int32_t param_id, const uint8_t *sk, size_t sk_len, int32_t *key_id);
int cmd_key_unload(cmd_ctx *ctx, int32_t key_id);
/* Copies a resident key's device arrays to host memory (any pointer may be NULL):
 * table sb x 2t uint16 (L_i^j / g(L_i)^2), L n uint16, lbits m x *lbits_stride uint32 (bit k of
 * L[32w + b] at bit b of word [k][w]), benes (2m-1) x 2^(m-5) uint32 swap masks.  No
 * counterpart in the reference: a diagnostic that lets callers and tests compare the
 * on-device key expansion of cmd_key_load with cmd_key_expand_host. */
int cmd_key_export(cmd_ctx *ctx, int32_t key_id, uint16_t *table, uint16_t *L, uint32_t *lbits, uint32_t *benes,
                   int32_t *lbits_stride);
/* Device bytes one resident key of this set occupies. */
int cmd_key_device_bytes(int32_t param_id, uint64_t *bytes);

/* The same key pairs as cmd_gen_keys (byte for byte), with the systematic-form test of every
 * attempt on the GPU: host threads draw each attempt's Goppa polynomial and support
 * permutation, the GPU builds the first mt parity-check columns of all attempts of a round and
 * decides their invertibility by Gauss-Jordan (the work of the reference's pk_gen.c:35-149 in
 * crypto_kem_keypair, kem/common/core/operations.c:36-107), and each key takes its first
 * successful attempt in seed order. */
int cmd_gen_keys_device(cmd_ctx *ctx, int32_t param_id, uint64_t base_seed, size_t n_keys, uint8_t *sk_out,
                        int threads);

/* The same ciphertexts as cmd_gen_encrypt, for resident keys of one set, computed on the GPU:
 * ct rows k*per_key .. (k+1)*per_key-1 (SYND_BYTES each) are cmd_gen_encrypt(param, sk of
 * key_ids[k], seeds[k], per_key, weight) and e (optional, n/8 bytes per row) its errors.
 * Per key the parity-check columns and the inverse of their first mt columns (Gauss-Jordan,
 * the work the reference's pk_gen.c:35-149 does on the CPU) are built on the device.
 * CMD_ERR_SINGULAR when a key has no systematic form. */
int cmd_gen_encrypt_device(cmd_ctx *ctx, size_t n_keys, const int32_t *key_ids, const uint64_t *seeds,
                           size_t per_key, int32_t weight, uint8_t *ct, uint8_t *e);

/* Batch decryption, host buffers (copied in and out, blocking).  Item i uses key key_ids[i]
 * and its ciphertext at ct + i*ct_stride; its error vector (n/8 bytes, the reference's
 * build_error_vector, decrypt.c:43-53) goes to e_out + i*e_stride, and the rest of each
 * e_stride row is zero.  Sets and keys may mix.  Replaces the reference's chunked batch
 * driver (GPU_Baseline/param/param8192128/src/cuda/Decrypt.cu:177-324): fused, KEM and
 * positions modes run chunks of up to 65536 items through a three-stage pipeline (copy in,
 * kernels, copy out of different chunks overlap); buffers from cmd_host_alloc are copied by
 * DMA directly, others through pinned staging.  In CMD_MODE_KEM the row of an item is its
 * 32-byte session key, in CMD_MODE_POSITIONS its t uint16 error positions. */
int cmd_decrypt(cmd_ctx *ctx, int32_t mode, size_t count, const int32_t *key_ids,
                const uint8_t *ct, size_t ct_stride, uint8_t *e_out, size_t e_stride);

/* Pinned host memory for cmd_decrypt / cmd_kem_dec buffers (ciphertexts read straight into
 * it, results formatted straight from it), freed with cmd_host_free or with the context.
 * No counterpart in the reference, which stages through pageable memory. */
int cmd_host_alloc(cmd_ctx *ctx, size_t bytes, void **ptr);
int cmd_host_free(cmd_ctx *ctx, void *ptr);

/* Same on device buffers already resident in HBM, enqueued on `stream` (a hipStream_t,
 * NULL = the context's stream); key_ids stay on the host.  Only the n/8 bytes of each row
 * are written.  Returns once the work is complete. */
int cmd_decrypt_device(cmd_ctx *ctx, int32_t mode, size_t count, const int32_t *key_ids,
                       const void *d_ct, size_t ct_stride, void *d_e, size_t e_stride, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * fused decryption as cmd_decrypt_device with CMD_MODE_FUSED, but the launch's work list
 * (tiles, item order, per-item keys, work-queue counter) goes to a fixed device array of the
 * library (about 6 MiB per GPU, held until the process ends) instead of memory allocated per
 * call; nothing is allocated on the device.  At most CMD_NOALLOC_MAX_ITEMS items per call
 * (CMD_ERR_ARG beyond).  All calls on one GPU share that array: calls that overlap on one
 * GPU (several contexts or processes' threads at once) race; one context serialises its own
 * calls, and each call returns once its kernels are done. */
#define CMD_NOALLOC_MAX_ITEMS 262144
int cmd_decrypt_device_noalloc(cmd_ctx *ctx, size_t count, const int32_t *key_ids, const void *d_ct,
                               size_t ct_stride, void *d_e, size_t e_stride, void *stream);

/* KEM decapsulation, host buffers (crypto_kem_dec of Classic McEliece, batched).  Item i:
 * ciphertext ct + i*ct_stride of key key_ids[i]; its 32-byte session key goes to
 * key_out + i*key_stride.  The key is SHAKE256(1 || e || c) when the decoded e has weight t
 * and the same syndrome over the whole support as c, else SHAKE256(0 || s || c) with the
 * secret key's s (implicit rejection; both cases run the same instructions).  The success key
 * is the one the reference's encapsulation derives (kem/common/core/operations.c:15-34);
 * decapsulation itself has no counterpart in the reference, which stops at e.  Same as
 * cmd_decrypt with CMD_MODE_KEM; cmd_decrypt_device and batches take that mode too. */
int cmd_kem_dec(cmd_ctx *ctx, size_t count, const int32_t *key_ids, const uint8_t *ct, size_t ct_stride,
                uint8_t *key_out, size_t key_stride);

/* ---- prepared batches (fixed key assignment, replayed as one hipGraph) --------------- */
typedef struct cmd_batch cmd_batch;

/* Plans a batch for `mode` (item grouping, schedules, device work list); with
 * use_graph != 0 the launch sequence is captured into a hipGraph on first run. */
int cmd_batch_create(cmd_ctx *ctx, int32_t mode, size_t count, const int32_t *key_ids,
                     size_t ct_stride, size_t e_stride, int32_t use_graph, cmd_batch **out);
int cmd_batch_run(cmd_batch *b, const void *d_ct, void *d_e, void *stream);
int cmd_batch_destroy(cmd_batch *b);

/* ---- per-stage entry points (reference synd / bm / root on the GPU) ------------------- */
/* synd (CPU/param/param6960119/src/util.c:102-114) for count ciphertexts of one key;
 * s_out: count x 2t uint16. */
int cmd_stage_synd(cmd_ctx *ctx, int32_t key_id, size_t count, const uint8_t *ct,
                   size_t ct_stride, uint16_t *s_out);
/* bm (CPU/src/common/bm.c:12-68): s: count x 2t, loc_out: count x (t+1), reversed like bm(). */
int cmd_stage_bm(cmd_ctx *ctx, int32_t param_id, size_t count, const uint16_t *s, uint16_t *loc_out);
/* root + build_error_vector (root_13.c:233-240, decrypt.c:43-53): loc: count x (t+1). */
int cmd_stage_root(cmd_ctx *ctx, int32_t key_id, size_t count, const uint16_t *loc, uint8_t *e_out,
                   size_t e_stride);
/* Syndrome of n-bit words over the whole support, s_j(r) = sum_{i<n} r_i L_i^j / g(L_i)^2 for
 * j < 2t (the decapsulation check's s_cmp; the reference's synd, util.c:102-114, covers the
 * first sb positions, where the two agree).  r: count rows of n/8 bytes at r_stride; s_out:
 * count x 2t uint16.  Computed by the transposed additive FFT of CMD_MODE_KEM. */
int cmd_stage_synd_full(cmd_ctx *ctx, int32_t key_id, size_t count, const uint8_t *r, size_t r_stride,
                        uint16_t *s_out);

/* ---- schedules (autotuner) ----------------------------------------------------------- */
/* Fused-kernel schedule per set: ciphertexts per tile, wavefronts per workgroup, wavefronts
 * per tile, persistent groups per CU, and the root-search algorithm (0: Horner at each support
 * element as root_13.c:216-240 does; 1: additive FFT over all 2^m field elements followed by
 * the key's Benes network).  All give the same error vectors; cmd_autotune measures candidates
 * on the GPU.  A launch over several sets takes the wavefront layout of its heaviest set and
 * each set's own tile width. */
typedef struct {
    int32_t tile_cts;        /* ciphertexts per tile: 1, 2, 4, 8, 16 or 32 (the kernel is
                                bit-sliced 32 wide; narrower tiles spread small batches wider) */
    int32_t waves_per_group; /* wavefronts per workgroup: 1, 2 or 4 (2 when waves_per_ct is 2) */
    int32_t waves_per_ct;    /* wavefronts per tile: 1, or 2 (the tile's syndrome columns and
                                Berlekamp-Massey lane blocks split between a pair of wavefronts,
                                four wavefronts per SIMD; additive-FFT root search only) */
    int32_t groups_per_cu;   /* persistent workgroups per compute unit */
    int32_t root_alg;        /* 0 Horner, 1 additive FFT + Benes */
} cmd_schedule;

int cmd_schedule_get(cmd_ctx *ctx, int32_t param_id, cmd_schedule *out);
int cmd_schedule_set(cmd_ctx *ctx, int32_t param_id, const cmd_schedule *s);
/* on != 0 (the default): a fused/KEM/positions batch too small to give about four tiles per
 * CU at the schedules' tile widths halves them (not below 4) until it does.  Off: the
 * schedules' widths apply as set. */
int cmd_set_tile_fit(cmd_ctx *ctx, int32_t on);
/* Times each candidate schedule on `count` synthetic items of the set (resident keys
 * required: key_ids of that set) and keeps the fastest; best_us receives its time. */
int cmd_autotune(cmd_ctx *ctx, int32_t param_id, size_t count, const int32_t *key_ids,
                 int32_t reps, cmd_schedule *best, double *best_us);
/* JSON persistence of the tuned schedules. */
int cmd_schedule_save(cmd_ctx *ctx, const char *path);
int cmd_schedule_load(cmd_ctx *ctx, const char *path);

/* ---- serving loop ------------------------------------------------------------------- */
typedef struct cmd_server cmd_server;

typedef struct {
    int32_t  max_batch;        /* items per launch (upper bound) */
    int32_t  max_inflight;     /* launches on the device at once (streams) */
    int32_t  queue_capacity;   /* pending requests before submit returns CMD_ERR_FULL */
    int32_t  min_batch;        /* launch as soon as this many are pending ... */
    double   max_wait_us;      /* ... or the oldest pending request waited this long */
    uint64_t io_budget_bytes;  /* HBM for in-flight ct/e buffers (admission control) */
} cmd_server_config;

typedef struct {
    uint64_t submitted, completed, rejected, launches;
    uint64_t failed;                   /* requests finished with an error (their launch failed) */
    double   mean_batch;
    double   p50_us, p99_us, p999_us, max_us;  /* submit -> result available */
    double   queue_p99_us, queue_max_us;       /* submit -> launch enqueued */
    double   device_p99_us, device_max_us;     /* launch enqueued -> completion seen (per launch) */
    double   gpu_p99_us, gpu_max_us;           /* input copy start -> results in host memory, on
                                                  the launch stream (HIP events, per launch) */
    double   busy_s;                   /* time with at least one launch in flight */
    /* Longest single step of the worker thread, by kind (wall time, microseconds): planning
     * and enqueueing a launch, retiring one (result copies, bookkeeping), one pass of
     * completion queries, and the overshoot of its timed wait for work beyond the time asked
     * for.  A long wait overshoot with short steps means the thread was not scheduled. */
    double   step_launch_max_us, step_retire_max_us, step_poll_max_us, wait_overshoot_max_us;
} cmd_server_stats;

int cmd_server_create(cmd_ctx *ctx, const cmd_server_config *cfg, cmd_server **out);
/* Copies the ciphertext (ct_bytes of the key's set); *req receives the request id.  The
 * request decrypts (CMD_MODE_FUSED); its result is the n/8-byte error vector. */
int cmd_server_submit(cmd_server *s, int32_t key_id, const uint8_t *ct, uint64_t *req);
/* Submits count decryption requests at once (item i: key_ids[i], ciphertext at
 * ct + i*ct_stride); reqs[i] receives the ids.  Keys and strides are checked, and the keys
 * held against cmd_key_unload, before anything is queued; if the queue fills, the first
 * *accepted requests stay submitted and CMD_ERR_FULL is returned. */
int cmd_server_submit_batch(cmd_server *s, size_t count, const int32_t *key_ids, const uint8_t *ct,
                            size_t ct_stride, uint64_t *reqs, size_t *accepted);
/* The same for requests of `mode`: CMD_MODE_FUSED (result: the n/8-byte error vector) or
 * CMD_MODE_KEM (result: the 32-byte session key of crypto_kem_dec, as cmd_kem_dec computes
 * it).  Both kinds share launches: a launch's decapsulations run the re-encryption check and
 * SHAKE256 after its fused decryption kernel. */
int cmd_server_submit_batch_mode(cmd_server *s, int32_t mode, size_t count, const int32_t *key_ids,
                                 const uint8_t *ct, size_t ct_stride, uint64_t *reqs, size_t *accepted);
/* Bytes of a request's result (n/8 for a decryption, 32 for a decapsulation). */
int cmd_server_result_bytes(cmd_server *s, uint64_t req, size_t *bytes);
/* CMD_OK and the result (out_cap >= its size, else CMD_ERR_ARG) once done, CMD_ERR_PENDING
 * before; frees the request.  A request whose launch failed returns that launch's error code
 * (and is freed); the server keeps serving other launches. */
int cmd_server_poll(cmd_server *s, uint64_t req, uint8_t *out, size_t out_cap);
/* Blocks until done or timeout_us elapses (< 0: no limit). */
int cmd_server_wait(cmd_server *s, uint64_t req, uint8_t *out, size_t out_cap, double timeout_us);
/* Waits for count requests and writes their results to out + i*out_stride (the requests are
 * freed); CMD_ERR_TIMEOUT once timeout_us (< 0: no limit) has passed.  The first failed
 * request's code is returned after every request has been waited for. */
int cmd_server_wait_batch(cmd_server *s, size_t count, const uint64_t *reqs, uint8_t *out, size_t out_stride,
                          double timeout_us);
int cmd_server_stats_get(cmd_server *s, cmd_server_stats *out);
/* Per-request latency samples of the requests retired since the last reset, in retirement
 * order: total (submit -> result) and queue (submit -> launch) microseconds.  Copies at most
 * cap of each; *n receives the number available. */
int cmd_server_latencies(cmd_server *s, double *total_us, double *queue_us, size_t cap, size_t *n);
/* One record per launch retired since the last reset (those whose device work completed),
 * in retirement order: where a launch's time went, on the host clock except gpu_us. */
typedef struct {
    double   t_us;        /* requests taken from the queue, microseconds since server creation */
    double   enqueue_us;  /* planning and enqueueing the copy and kernels (worker thread) */
    double   gpu_us;      /* on the launch stream: input copy start -> results in host memory */
    double   seen_us;     /* enqueue done -> completion seen by the worker (contains gpu_us) */
    double   retire_us;   /* completion seen -> results delivered to the requests */
    uint32_t items, pad;
} cmd_server_launch;
int cmd_server_launch_log(cmd_server *s, cmd_server_launch *out, size_t cap, size_t *n);
int cmd_server_reset_stats(cmd_server *s);
/* Stops admission, drains in-flight work, frees everything. */
int cmd_server_destroy(cmd_server *s);

/* ---- measurement helpers -------------------------------------------------------------- */
/* Average duration in microseconds of the kernels of the last cmd_batch_run /
 * cmd_decrypt_device launches, and of each pipeline chunk of cmd_decrypt, made with timing
 * on (HIP events on the launch stream), and resets the accumulator. */
int cmd_set_kernel_timing(cmd_ctx *ctx, int32_t on);
int cmd_kernel_timing(cmd_ctx *ctx, double *avg_us, uint64_t *launches);

/* Measured VALU issue rate of the context's device in 32-bit lane operations per second
 * (chains of single v_xor_b32 instructions over all CUs): the peak the decoder's integer
 * work is compared with.  No counterpart in the reference; a measurement aid. */
int cmd_valu_peak(cmd_ctx *ctx, double *lane_ops_per_s);

/* VALU issue rate in wave instructions per SIMD per clock (at the device's peak clock) with
 * waves_per_simd (1..8) wavefronts per SIMD, each lane running `chains` (1, 2, 4 or 8)
 * independent chains of v_xor_b32; and the dependent-chain latencies of one wavefront in
 * shader-clock ticks per step (v_xor_b32, DPP row shift + v_xor_b32, v_readlane + scalar +
 * v_xor_b32).  No counterpart in the reference; measurement aids for what bounds a
 * wavefront's issue. */
int cmd_valu_issue_rate(cmd_ctx *ctx, int32_t waves_per_simd, int32_t chains, double *wave_insts_per_clk);
int cmd_chain_latency(cmd_ctx *ctx, double ticks[3]);

/* Library build identification (kernel variants compiled in). */
const char *cmd_build_info(void);

#ifdef __cplusplus
}
#endif

#endif /* CM_DECRYPT_H */
