/*
 * gpurecsplit_amd.h -- C ABI of the MI355X (gfx950) RecSplit construction engine.
 *
 * A RecSplit minimal perfect hash function (MPHF) maps n distinct 128-bit key
 * hashes {first, second} onto [0, n).  This library builds it on one AMD GPU and
 * returns a handle whose serialized bytes follow the reference's stream layout
 * (leaf size, bucket size, key count, Golomb-Rice descriptor vector, double
 * Elias-Fano structure), and which answers queries on the host or, batched, on
 * the GPU.
 *
 * Reference interfaces these entry points replace (paths relative to the
 * ByteHamster/GpuRecSplit repository root):
 *   rs_build / rs_build_device
 *       GPURecSplit(vector<hash128_t>& keys, size_t bucket_size, int numThreads)
 *       include/function/GPURecSplit.cuh:484-489 (construction: hash_gen, :637-815)
 *   rs_mphf_serialize / rs_mphf_serialized_size
 *       operator<<  include/function/GPURecSplit.cuh:817-825
 *   rs_mphf_deserialize
 *       operator>>  include/function/GPURecSplit.cuh:827-841
 *   rs_mphf_query
 *       size_t operator()(const hash128_t&)
 *       include/function/AbstractParallelRecSplit.hpp:409-466
 *   rs_mphf_bits
 *       getBits()   include/function/AbstractParallelRecSplit.hpp:476-478
 *       (without the 8*sizeof(object) term, which the C++ wrapper adds)
 *   rs_query_device
 *       batched operator() on the GPU (no reference counterpart; the reference
 *       queries one key at a time on the host)
 *
 * Errors are returned as status codes (the reference prints and calls exit() on
 * CUDA errors, GPURecSplit.cuh:73-80, loops forever on duplicate keys, :436,
 * and throws std::out_of_range for a bucket of 3000 or more keys, :682).
 * rs_last_error() gives a message for the calling thread's last failure.
 *
 * Keys are passed as n pairs of uint64 (first, second), little-endian, i.e. the
 * memory layout of the reference's hash128_t array.  The caller's keys are
 * never modified (the reference sorts the caller's vector in place,
 * AbstractParallelRecSplit.hpp:344).
 *
 * Thread safety: a context serialises builds issued through it; use one context
 * per host thread for concurrent builds on one device.
 */
#ifndef GPURECSPLIT_AMD_H
#define GPURECSPLIT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rs_ctx rs_ctx;
typedef struct rs_mphf rs_mphf;

enum rs_status {
    RS_OK = 0,
    RS_ERR_ARG = 1,               /* invalid argument (n == 0, leaf size outside [2, 24], ...) */
    RS_ERR_HIP = 2,               /* HIP runtime error */
    RS_ERR_BUCKET_TOO_LARGE = 3,  /* a bucket holds 3000 or more keys */
    RS_ERR_DUPLICATE_KEYS = 4,    /* two keys in one bucket share `second` */
    RS_ERR_SEED_OVERFLOW = 5,     /* a node search passed the 2^44 seed limit */
    RS_ERR_EF_OVERFLOW = 6,       /* Elias-Fano jump offset >= 2^16 (reference aborts) */
    RS_ERR_NOMEM = 7,             /* host or device allocation failed */
    RS_ERR_FORMAT = 8,            /* malformed serialized stream or leaf-size mismatch */
    RS_ERR_NO_DEVICE = 9,         /* reserved (no longer returned: device copies are made on demand) */
    RS_ERR_TIMEOUT = 10,          /* the build passed the context's time limit */
    RS_ERR_INTERNAL = 11,         /* an internal consistency check failed */
    RS_ERR_PENDING = 12           /* rs_wait: the job did not finish within the timeout */
};

/* Per-build measurements (HIP events on the build stream). */
typedef struct rs_build_stats {
    double total_ms;          /* host wall clock of the whole build call */
    double partition_ms;      /* histogram + scans + scatter + per-bucket sort */
    double search_ms[8];      /* search kernels: [0] fanout-2 levels, [1] upper
                                 aggregation, [2] lower aggregation, [3] rotation
                                 leaves, [4] brute-force leaves */
    double split_ms;          /* partition of keys between tree levels */
    double encode_ms;         /* Golomb-Rice + Elias-Fano */
    uint64_t evals[8];        /* algorithmic key-seed evaluations per kind (same index) */
    uint64_t nodes[8];        /* nodes per kind */
    uint64_t executed_evals[8]; /* key-seed evaluations the kernels actually ran, per kind */
    uint64_t num_buckets;
    uint64_t num_nodes;
    uint64_t rice_bits;
    uint32_t max_bucket_size;
    uint32_t search_waves;    /* persistent waves of the full search grid */
    /* Per kind (same index), filled when the context was created with the
     * environment variable RS_WAVE_TIMES=1 (else 0): every search wave records its
     * start and exit on the device clock.  search_busy = share of the resident
     * waves' time, from the first start to the last exit of the launch, spent
     * inside the search loop (1 - idle share of the launch's tail);
     * search_tail_ms = time from the median wave's exit to the last exit. */
    double search_busy[8];
    double search_tail_ms[8];
    /* candidate seeds tested per kind (algorithmic, from the output seeds): x + 1
     * per split or brute-force node with seed x, (floor(x/L) + 1) * L per rotation
     * leaf (each pass over its L keys tests L rotations, i.e. L seeds) */
    uint64_t seeds[8];
    /* per kind: the most persistent waves a launch of that kind ran with (levels
     * of little expected work run on fewer than search_waves) */
    uint32_t kind_waves[8];
} rs_build_stats;

/* Context bound to one GPU (device ordinal as seen by HIP). */
int rs_ctx_create(int device, rs_ctx **out);
void rs_ctx_destroy(rs_ctx *ctx);

/* Wall-clock limit for each later build on ctx, in seconds (0: none, the
 * default; the environment variable RS_TIME_LIMIT sets it at creation).  Past
 * the limit the seed-search kernels stop at their next work claim and the build
 * returns RS_ERR_TIMEOUT; the context stays usable.  No reference counterpart
 * (the reference search runs until every seed is found, GPURecSplit.cuh:508-577);
 * it is for servers that must bound a job's time. */
int rs_ctx_set_time_limit(rs_ctx *ctx, double seconds);

/* Construction from host keys (n pairs first,second).  leaf_size in [2, 24],
 * bucket_size >= 1, rotate != 0 selects rotation fitting for full leaves
 * (USE_BIJECTIONS_ROTATE, GPURecSplit.cuh:409). */
int rs_build(rs_ctx *ctx, const uint64_t *keys, size_t n, uint32_t leaf_size, size_t bucket_size,
             int rotate, rs_mphf **out);

/* Same, with the keys already resident in device memory of ctx's GPU. */
int rs_build_device(rs_ctx *ctx, const uint64_t *d_keys, size_t n, uint32_t leaf_size,
                    size_t bucket_size, int rotate, rs_mphf **out);

/* Allocation-free construction: the same build as rs_build_device, for callers that
 * cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc).  No reference counterpart.  The call synchronises the library's own
 * stream between its phases (it reads sizes back to the host), so it cannot be
 * recorded by a caller's stream capture.  Every device
 * buffer of the build comes from a fixed 64 MiB scratch array that the library
 * holds per GPU until the process ends; the call makes no hipMalloc, hipFree or
 * pinned/torch allocation.  Inputs bounded by 75*n + 89*ceil(n/bucket_size) +
 * 2^20 <= 64 MiB (about 400,000 keys at any bucket size, 870,000 at b = 100) are
 * accepted; larger ones fail with RS_ERR_ARG.  The structure is written
 * serialized (rs_mphf_serialize's layout) into out[0, *out_len); if cap is too
 * small the call fails with RS_ERR_ARG and *out_len holds the size needed.
 * All calls on one GPU share its scratch array: the library serialises them with
 * a lock (overlapping uses of the array would race), and each call clears what it
 * needs before use. */
int rs_build_device_noalloc(rs_ctx *ctx, const uint64_t *d_keys, size_t n, uint32_t leaf_size, size_t bucket_size,
                            int rotate, uint8_t *out, size_t cap, size_t *out_len);

/* Batched construction: num_sets independent MPHFs in one set of launches.  Set i
 * owns keys [set_offsets[i], set_offsets[i+1]) of d_keys (device memory;
 * set_offsets is a host array of num_sets+1 increasing offsets, every set
 * non-empty).  All sets share leaf_size, bucket_size and rotate; out receives
 * num_sets handles (host-queryable, no device copy), stats the batch's
 * measurements (may be NULL).  Each handle equals what rs_build_device returns
 * for its set alone.  No reference counterpart (the reference builds one MPHF
 * per constructor call, GPURecSplit.cuh:484-489). */
int rs_build_batch(rs_ctx *ctx, const uint64_t *d_keys, const uint64_t *set_offsets, size_t num_sets,
                   uint32_t leaf_size, size_t bucket_size, int rotate, rs_mphf **out, rs_build_stats *stats);

/* Launch tuning of the seed search (no reference counterpart; the reference fixes
 * its launch constants, GPURecSplit.cuh:420-425).  rs_autotune times builds of a
 * synthetic sample (sample_keys SplitMix64 keys; 0 sizes it to ~40 ms per build)
 * for the leaf size, bucket size and flags, picks per node kind the candidates per
 * lane and per claimed chunk, then the workgroups per CU, and makes the result the
 * context's tuning for that leaf size; report (may be NULL) receives a JSON
 * summary.  RS_AUTOTUNE=1 in the environment tunes each leaf size at its first
 * build on a context.  A tuning is a string "u2=4,g2=2,bpc=8,..." (the RS_TUNE
 * syntax): rs_ctx_set_tune applies one to a leaf size (0: the default for all
 * leaf sizes), rs_ctx_get_tune returns the one in effect. */
int rs_autotune(rs_ctx *ctx, uint32_t leaf_size, size_t bucket_size, int rotate, uint64_t sample_keys, char *report,
                size_t cap);
int rs_ctx_set_tune(rs_ctx *ctx, uint32_t leaf_size, const char *tune);
int rs_ctx_get_tune(rs_ctx *ctx, uint32_t leaf_size, char *buf, size_t cap);

/* Device memory the context holds for reuse by later builds (grow-only buffers),
 * and their release (which also frees the pinned host staging of batch results).
 * No reference counterpart. */
uint64_t rs_ctx_device_bytes(rs_ctx *ctx);
int rs_ctx_release(rs_ctx *ctx);

/* ---------------------------------------------------------------------------
 * Build plans: the construction of one MPHF of a fixed key count, queued on the
 * caller's HIP stream with no device allocation and no host synchronisation, so
 * that it can be recorded by stream capture into a hipGraph and replayed for new
 * keys.  No reference counterpart (the reference constructor is synchronous,
 * GPURecSplit.cuh:484-489).
 *
 * rs_plan_create sizes every device buffer for n keys (one allocation, about
 * 90 bytes per key; rs_plan_device_bytes) and uploads the leaf size's tables.
 * rs_plan_enqueue queues a build of the n keys (first, second pairs) at d_keys, in
 * device memory of the plan's GPU, on `stream` (a hipStream_t; NULL: the null
 * stream); it only launches kernels, and the keys must stay valid until the
 * stream's work is done.  After the caller has synchronised the stream,
 * rs_plan_result copies the serialized structure (rs_mphf_serialize's layout,
 * byte-identical to rs_build_device's) into out[0, *len) (out == NULL: only *len)
 * and reports the build's errors (duplicate keys, a bucket of 3000 keys, ...; a
 * structure beyond the plan's generous size bounds fails with RS_ERR_ARG).
 * Enqueues of one plan share its buffers: a second enqueue must not start before
 * the first one's result has been taken; use one plan per concurrent stream.
 * Destroy plans before their context.
 * ------------------------------------------------------------------------- */
typedef struct rs_plan rs_plan;

int rs_plan_create(rs_ctx *ctx, size_t n, uint32_t leaf_size, size_t bucket_size, int rotate, rs_plan **out);
int rs_plan_enqueue(rs_plan *plan, const uint64_t *d_keys, void *stream);
int rs_plan_result(rs_plan *plan, uint8_t *out, size_t cap, size_t *len);
size_t rs_plan_device_bytes(const rs_plan *plan);
void rs_plan_destroy(rs_plan *plan);

/* ---------------------------------------------------------------------------
 * Construction server: asynchronous jobs with continuous batching on one GPU.
 * No reference counterpart (the reference builds one MPHF per synchronous
 * constructor call, GPURecSplit.cuh:484-489).
 *
 * A job's keys are copied into the server's device key pool when it is submitted
 * (or, while the HBM budget is exhausted, kept in host memory and uploaded as
 * memory frees up).  `slots` build contexts, each with its own HIP stream, take
 * the oldest ready job together with every ready job of the same (leaf size,
 * bucket size, rotation), up to max_batch_keys, and build them as one batch;
 * jobs that arrive while batches run are taken by the next free slot, and the
 * batches of different slots overlap on the device.  A batch starts only when
 * the key pool, every slot's buffers and the batch's worst-case working set fit
 * hbm_budget_bytes.  Every result equals the standalone build of its keys; a
 * failing batch is retried job by job, so only the failing job reports an error.
 * ------------------------------------------------------------------------- */
typedef struct rs_server rs_server;

typedef struct rs_server_config {
    int device;                /* HIP device ordinal */
    uint32_t slots;            /* concurrent batch builds; 0: 4 */
    uint64_t hbm_budget_bytes; /* device bytes for keys and build buffers; 0: 64 GiB */
    uint64_t max_batch_keys;   /* keys per batch; 0: 50,000,000 */
    double time_limit_s;       /* per batch build (rs_ctx_set_time_limit); 0: none */
} rs_server_config;

enum rs_job_state { RS_JOB_QUEUED = 0, RS_JOB_RUNNING = 1, RS_JOB_DONE = 2, RS_JOB_FAILED = 3 };

typedef struct rs_job_info {
    int state;          /* rs_job_state */
    int status;         /* rs_status of a finished job */
    double t_submit;    /* seconds on the server clock (from rs_server_create) */
    double t_start;     /* -1 until the job's batch starts */
    double t_done;      /* -1 until the job finishes */
    uint64_t n;
    uint32_t leaf;
    uint64_t bucket;
    uint64_t batch_id;  /* batch the job was built in (0: not started) */
    uint32_t batch_jobs;
} rs_job_info;

typedef struct rs_serve_stats {
    uint64_t jobs_submitted, jobs_done, jobs_failed, keys_done;
    uint64_t batches;           /* batch builds started */
    uint64_t batches_split;     /* failed batches retried job by job */
    uint64_t committed_bytes;   /* key pool + slot buffers + growth bound of running batches */
    uint64_t peak_committed_bytes;
    uint64_t key_pool_bytes, key_pool_used_bytes;
    uint64_t jobs_waiting_for_memory, jobs_ready;
    double uptime_s;
} rs_serve_stats;

int rs_server_create(const rs_server_config *cfg, rs_server **out);
/* Finishes every submitted job, then frees the server (results not taken are freed). */
void rs_server_destroy(rs_server *s);
/* Host keys (n pairs); the call copies them, so the buffer may be reused at once. */
int rs_submit_job(rs_server *s, const uint64_t *keys, size_t n, uint32_t leaf_size, size_t bucket_size, int rotate,
                  uint64_t *job_id);
/* Keys already in device memory of the server's GPU; they are used in place and
 * must stay valid until rs_wait returns for the job. */
int rs_submit_job_device(rs_server *s, const uint64_t *d_keys, size_t n, uint32_t leaf_size, size_t bucket_size,
                         int rotate, uint64_t *job_id);
/* num_sets host key sets (as rs_build_batch) submitted as separate jobs. */
int rs_submit_batch(rs_server *s, const uint64_t *keys, const uint64_t *set_offsets, size_t num_sets,
                    uint32_t leaf_size, size_t bucket_size, int rotate, uint64_t *job_ids);
/* State of a job (not blocking). */
int rs_poll(rs_server *s, uint64_t job_id, rs_job_info *info);
/* Waits up to timeout_s seconds (< 0: no limit) for the job.  On success the
 * job's MPHF is handed to the caller (*out, free with rs_mphf_free) and the job
 * is forgotten; a failed job returns its status and is forgotten; a job still
 * running returns RS_ERR_PENDING and stays. */
int rs_wait(rs_server *s, uint64_t job_id, double timeout_s, rs_mphf **out, rs_job_info *info);
int rs_server_stats(rs_server *s, rs_serve_stats *out);

/* ---------------------------------------------------------------------------
 * String keys.  The reference replaces every string key by the 128-bit hash
 * first_hash(key, len) = SpookyHash V2 Hash128 with seeds 0, taken as
 * {first = h1, second = h0} (AbstractParallelRecSplit.hpp:250,
 * include/util/Hash128.h:32-36); these entry points do the same.
 * ------------------------------------------------------------------------- */

/* SpookyHash::Hash128: *hash1, *hash2 are the seeds on entry and the hash on exit. */
void rs_spooky128(const void *data, size_t len, uint64_t *hash1, uint64_t *hash2);
/* first_hash of one key: out2[0] = first, out2[1] = second. */
void rs_first_hash(const void *data, size_t len, uint64_t *out2);
/* first_hash of n keys on the host: key i is bytes [offsets[i], offsets[i+1]). */
void rs_hash_strings(const uint8_t *bytes, const uint64_t *offsets, size_t n, uint64_t *out);
/* The same on the GPU, everything in device memory of ctx's GPU. */
int rs_hash_strings_device(rs_ctx *ctx, const uint8_t *d_bytes, const uint64_t *d_offsets, size_t n,
                           uint64_t *d_keys);
/* Construction from string keys (host bytes + n+1 offsets): the keys are hashed on
 * the GPU, then built as rs_build does.  Replaces GPURecSplit(const vector<string>&,
 * size_t bucket_size, int numThreads), include/function/GPURecSplit.cuh:443-478. */
int rs_build_strings(rs_ctx *ctx, const uint8_t *bytes, const uint64_t *offsets, size_t n, uint32_t leaf_size,
                     size_t bucket_size, int rotate, rs_mphf **out);
/* operator()(const string&), AbstractParallelRecSplit.hpp:473. */
uint64_t rs_mphf_query_string(const rs_mphf *h, const void *data, size_t len);

/* Serialization in the reference's stream layout. */
size_t rs_mphf_serialized_size(const rs_mphf *h);
int rs_mphf_serialize(const rs_mphf *h, uint8_t *buf, size_t cap);
int rs_mphf_deserialize(const uint8_t *buf, size_t len, uint32_t leaf_size, int rotate, rs_mphf **out);

/* Host query of one key: its index in [0, n). */
uint64_t rs_mphf_query(const rs_mphf *h, uint64_t first, uint64_t second);
/* Host query of n keys. */
void rs_mphf_query_batch(const rs_mphf *h, const uint64_t *keys, size_t n, uint64_t *out);
/* Batched query on the GPU: d_keys (n pairs) and d_out (n uint64) in device memory
 * of ctx's GPU.  Any handle works: one without a copy on ctx's GPU (deserialized,
 * batch results, built on another device) is uploaded on its first device query
 * and keeps that copy until freed. */
int rs_query_device(rs_ctx *ctx, const rs_mphf *h, const uint64_t *d_keys, size_t n, uint64_t *d_out);

/* Space of the structure in bits (Elias-Fano + descriptors). */
uint64_t rs_mphf_bits(const rs_mphf *h);
uint64_t rs_mphf_num_keys(const rs_mphf *h);
int rs_mphf_stats(const rs_mphf *h, rs_build_stats *out);
void rs_mphf_free(rs_mphf *h);

/* Golomb-Rice memo for a leaf size: golomb parameter << 27 | subtree nodes << 16 |
 * subtree fixed bits, for bucket sizes 0..2999 (AbstractParallelRecSplit.hpp:191-227,
 * golombRiceMemo.hpp). */
int rs_golomb_memo(uint32_t leaf_size, uint32_t *out3000);

const char *rs_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GPURECSPLIT_AMD_H */
