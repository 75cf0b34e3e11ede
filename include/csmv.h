/*
 * csmv.h -- C-ABI of the MI355X (gfx950) CSMV engine.
 *
 * The reference has no FFI; its boundary is a host init/teardown pair, the
 * message-ring constructor and the device transaction API.  Each entry point
 * below names the reference interface it replaces.  All functions return a
 * csmv_status; none exits the process or throws.  The engine owns device
 * memory; the caller owns every host buffer it passes.
 */
#ifndef CSMV_H
#define CSMV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CSMV_OK = 0,
    CSMV_E_HIP = 1,        /* a HIP runtime call failed (see csmv_last_error) */
    CSMV_E_ARG = 2,        /* invalid argument */
    CSMV_E_CAPACITY = 3,   /* size beyond a configured or device limit */
    CSMV_E_TIMEOUT = 4,    /* a device wait exceeded the spin timeout */
    CSMV_E_NODEVICE = 5,   /* no gfx950 device visible */
    CSMV_E_STATE = 6,      /* call out of order (e.g. run before generate) */
    CSMV_E_LIVELOCK = 7,   /* a transaction exceeded the attempt limit */
    CSMV_E_IO = 8          /* a log file could not be written or read, or is malformed */
} csmv_status;

/* Commit-path variants.  They are feature flags of one code path; the named
 * presets reproduce the reference's optimisation directories:
 *   vanilla           csmv_optimizations/vanilla/csmv_sh           SCAN | LANE  | SERVER_WB
 *   batchedAtrInsert  csmv_optimizations/batchedAtrInsert/csmv_sh  SCAN | BATCH | CLIENT_WB
 *   clientWriteback   csmv_optimizations/clientWriteback/csmv_sh   SCAN | BATCH | SERVER_WB
 *   collabValidation  csmv_optimizations/collabValidation/csmv_sh  COLLAB | BATCH | CLIENT_WB
 * (the directory named clientWriteback is the one whose server writes back).
 * STAMP validation is this engine's own mode: an item-indexed commit-stamp
 * table replaces the linear record scan. */
enum {
    CSMV_VAL_SCAN = 0x0,      /* per-lane serial record-ring scan (TXValidate, bank/csmv_sh/API.cu:88-144) */
    CSMV_VAL_COLLAB = 0x1,    /* 64 lanes scan one client's window (collabValidation/csmv_sh/API.cu:193-266) */
    CSMV_VAL_STAMP = 0x2,     /* per-item last-commit stamp */
    CSMV_VAL_MASK = 0x3,
    CSMV_INSERT_LANE = 0x0,   /* per-lane record insertion (bank/csmv_sh/API.cu:193-272) */
    CSMV_INSERT_BATCH = 0x4,  /* one batched insertion per wave (batchedAtrInsert/csmv_sh/API.cu:266-389) */
    CSMV_WB_CLIENT = 0x0,     /* clients install versions (batchedAtrInsert/csmv_sh/API.cu:231-263) */
    CSMV_WB_SERVER = 0x8,     /* the server installs versions (vanilla/csmv_sh/API.cu:269-272) */
    CSMV_SERIAL_PIPE = 0x10   /* pipelined serialized lane (bank driver, with CSMV_VARIANT_STAMP
                                 only; no counterpart in the reference): the serialized-lane
                                 clients (four transactions a lane, requests of up to 256)
                                 stage their requests' producer chains and split them
                                 by commit partition; committer waves apply each account's
                                 request-wide delta with one ordered 64-bit atomic, install the
                                 versions and reply with the values every read sees, and the
                                 clients execute and log their transactions on them.  Every
                                 transfer must be routed to the serialized lane (the normal lane
                                 then carries audits only; routing every transaction leaves
                                 none to abort); at most 4096 client waves.  A serving run's
                                 bank partition with it runs every transaction on the lane
                                 (read sets of 2..64).  No escalation, cache partitions,
                                 caller kernels or allocation-free runs. */
};

#define CSMV_VARIANT_VANILLA          (CSMV_VAL_SCAN | CSMV_INSERT_LANE | CSMV_WB_SERVER)
#define CSMV_VARIANT_BATCHED_ATR      (CSMV_VAL_SCAN | CSMV_INSERT_BATCH | CSMV_WB_CLIENT)
#define CSMV_VARIANT_CLIENT_WRITEBACK (CSMV_VAL_SCAN | CSMV_INSERT_BATCH | CSMV_WB_SERVER)
#define CSMV_VARIANT_COLLAB           (CSMV_VAL_COLLAB | CSMV_INSERT_BATCH | CSMV_WB_CLIENT)
#define CSMV_VARIANT_STAMP            (CSMV_VAL_STAMP | CSMV_INSERT_BATCH | CSMV_WB_CLIENT)
#define CSMV_VARIANT_STAMP_PIPE       (CSMV_VARIANT_STAMP | CSMV_SERIAL_PIPE)

typedef struct {
    uint32_t n_items;          /* data items (accounts); VertionedDataItem count */
    uint32_t depth;            /* versions per item (MaxVersions, API.cuh:18), 1..64 */
    uint32_t atr_entries;      /* record ring capacity (TXRecordSize, API.cuh:17); power of two */
    uint32_t client_waves;     /* persistent client wavefronts (64 lanes each) */
    uint32_t wg_waves;         /* waves per workgroup, 3..8 (3..16 with CSMV_SERIAL_PIPE); the
                                  server workgroup uses all */
    uint32_t variant;          /* CSMV_VAL_* | CSMV_INSERT_* | CSMV_WB_* (| CSMV_SERIAL_PIPE) */
    uint32_t timeout_ms;       /* bound of every device-side wait (0 = 4000) */
    uint32_t max_attempts;     /* per-transaction attempt bound (0 = 1<<20) */
    uint32_t stage_ahead;      /* requests staged ahead of the committer (0 = 3) */
    uint32_t lds_kib;          /* server LDS budget in KiB (0 = device maximum) */
    uint32_t n_serial_waves;   /* client waves of the serialized lane (routed transfers) */
    uint32_t begin_wait;       /* 1: a wave's snapshot covers its own last commit
                                  (TXBegin's prevTs wait, bank/csmv_sh/API.cu:283) */
    uint32_t client_wg_waves;  /* client waves per client workgroup, 1..wg_waves (0 = wg_waves):
                                  fewer spread the clients over more CUs, each with its own
                                  memory pipeline; the other waves of the workgroup leave */
    uint32_t escalate_after;   /* forward progress of the bank drivers' normal lane: a
                                  transaction that has failed this many attempts is sent on
                                  the serialized lane in its wave's next request (executed
                                  by the commit server in commit order, so it cannot abort).
                                  0 = never: pure arrival order with unchanged operands,
                                  bounded by max_attempts (the reference retries in arrival
                                  order, bank/csmv_sh/exe.cu:142-269).  Needs the STAMP
                                  variant with client write-back; the cache driver has no
                                  serialized lane and ignores it. */
    uint32_t commit_parts;     /* pipelined serialized lane (CSMV_SERIAL_PIPE): commit server
                                  workgroups, 1..8 (0 = 1).  Items are split among them by a
                                  hash; each applies, installs and publishes its own items'
                                  writes for every request, in request order, and the clock is
                                  the least of their published progress. */
} csmv_config;

typedef struct {
    uint64_t seed;
    uint64_t n_txns;
    uint32_t n_reads;      /* reads per transfer (2..4 for csmv_bank_run, 2..64 for the
                              serving loop); the first two are the transfer pair */
    uint32_t n_reads_ro;   /* reads per audit (1..64) */
    uint32_t pct_ro;       /* percent audits 0..100 */
    uint32_t dist;         /* 0 uniform, 1 Zipf */
    uint32_t amount_max;   /* amount uniform in 1..amount_max (1..255) */
    uint32_t n_reads_min;  /* 0: transfers read n_reads; else uniform in n_reads_min..n_reads */
    double theta;          /* Zipf exponent */
} csmv_bank_spec;

typedef struct {
    uint64_t commits_update;
    uint64_t commits_ro;
    uint64_t aborts_dataage;     /* no version <= snapshot left in the ring */
    uint64_t aborts_readwrite;   /* read item written by an earlier batch after the snapshot */
    uint64_t aborts_prevalid;    /* read item written by a lower lane of the same batch */
    uint64_t aborts_recordage;   /* snapshot older than the record ring covers */
    uint64_t batches;            /* commit requests served */
    uint64_t final_ts;           /* last commit timestamp */
    double kernel_ms;            /* persistent kernel time (HIP events) */
    uint32_t error;              /* device error word (0 = none) */
    uint32_t client_lanes;
    /* timing counters in nanoseconds (the reference's time_rate, API.cuh:121-135):
     * [0] committer idle, [1] committer busy, [2] receivers waiting for the
     * committer, [3] receivers waiting for requests, [4] receivers staging,
     * [5] client waves waiting for replies, [6] client wave rounds, [7] rounds (count),
     * and, with phase timers compiled in (-DCSMV_PHASE_TIMERS=1), the committer's
     * stages: [8] mailbox read and progress publication, [9] committed values /
     * validation, [10] serialized-lane chain sums, [11] timestamps, tables, replies and
     * log, [12] re-reading a mailbox the read-ahead found unstaged, [13] waiting for a
     * completion slot, [14] requests the read-ahead missed (count); [15] unused */
    double prof_ns[16];
    uint32_t stamp_buckets;      /* LDS commit-stamp buckets (STAMP mode) */
    uint32_t stamp_exact;        /* 1 if every item has its own bucket */
    uint64_t aborts_unknown;     /* serialized lane: an operand's committed value was not at hand */
    uint32_t lds_bytes;          /* server workgroup LDS */
    uint32_t reserved;
    uint64_t escalated;          /* normal-lane transactions escalated to the serialized lane */
} csmv_stats;

const char *csmv_status_str(csmv_status s);

typedef struct csmv_engine csmv_engine;

/* Replaces TXInit's allocations (bank/csmv_sh/API.cu:293-333) and create_gbc
 * (bank/sync_lib/one_phase_server.h:13-42). */
csmv_status csmv_create(const csmv_config *cfg, csmv_engine **out);
/* Replaces TXEnd's frees (API.cu:354-358) and free_gbc (msg_aux.h:94-102). */
csmv_status csmv_destroy(csmv_engine *e);
/* Text of the last failure on this engine ("" if none). */
const char *csmv_last_error(const csmv_engine *e);

/* TXInit's data upload (API.cu:310-322): item i gets one version (ts 0, values[i]). */
csmv_status csmv_load_i32(csmv_engine *e, const int32_t *values, size_t n);
/* Same with every item set to one value. */
csmv_status csmv_fill_i32(csmv_engine *e, int32_t value);
/* TXEnd's read-back (API.cu:337-348): the head value of every item. */
csmv_status csmv_read_state_i32(csmv_engine *e, int32_t *out, size_t n);

/* Host: Zipf CDF scaled to 2^31 (cdf[n-1] = 2^31), the table the device
 * generator searches.  Needs no device. */
csmv_status csmv_zipf_cdf(uint32_t n, double theta, uint32_t *cdf);

/* Generates the bank stream on the device (the driver's operand draws,
 * bank/csmv_sh/exe.cu:134-201, as a resident stream). */
csmv_status csmv_bank_generate(csmv_engine *e, const csmv_bank_spec *spec);
/* Copies the stream back: op[n_txns], addr[rmax][n_txns]. */
csmv_status csmv_bank_read_stream(csmv_engine *e, uint32_t *op, uint32_t *addr);
/* Routes the resident stream: transfers with hot[k] != 0 go to the serialized lane
 * (executed by the commit server in commit order), everything else to the normal
 * lane.  hot = NULL removes the routing.  The reference has no counterpart: its
 * drivers run every transaction through the same client path
 * (bank/csmv_sh/exe.cu:142-269). */
csmv_status csmv_bank_route(csmv_engine *e, const uint8_t *hot, size_t n);
/* Runs the whole resident stream through the persistent kernel (replaces
 * parent_kernel/server_kernel/client_kernel, bank/csmv_sh/exe.cu:89-338). */
csmv_status csmv_bank_run(csmv_engine *e, csmv_stats *stats);
/* Commit log: ts[n_txns] (commit ts; audit snapshot), v[n_reads][n_txns]
 * (transfer reads; v[k] = audit sum), aborts[n_txns] (failed attempts). */
csmv_status csmv_bank_read_log(csmv_engine *e, uint32_t *ts, int32_t *v,
                               uint32_t *aborts);

/* Memcached-style cache (memcachedGPU/memcdSH.cu:486-768): n_sets x n_ways slots of 12
 * int32 items (field-major: state, key low, key high, LRU timestamp, 8 value words), so
 * the engine must be created with n_items = 12 * n_sets * n_ways; the final cache
 * contents are csmv_read_state_i32. */
typedef struct {
    uint64_t seed;
    uint64_t n_reqs;
    uint64_t n_keys;       /* keyspace (< 2^31) */
    uint32_t n_sets;       /* power of two */
    uint32_t n_ways;       /* 1..8 */
    uint32_t pct_get;      /* percent GET, rest SET */
    uint32_t batch;        /* requests per LRU clock tick */
    uint32_t dist;         /* 0 uniform, 1 Zipf over key ranks */
    uint32_t reserved;
    double theta;
} csmv_memcd_spec;

/* Request stream on the device (the host key generation loop, memcdSH.cu:586-595). */
csmv_status csmv_memcd_generate(csmv_engine *e, const csmv_memcd_spec *spec);
/* A request stream from host arrays (key[n_reqs], op[n_reqs]: 0 GET, 1 SET), e.g. one
 * partition of a stream split by the caller: the engine uses spec's n_reqs, n_sets, n_ways
 * and batch (request k runs with LRU clock k / batch + 1 and writes value words of (key, k));
 * its seed, keyspace and distribution are only recorded. */
csmv_status csmv_memcd_load(csmv_engine *e, const csmv_memcd_spec *spec, const uint64_t *key,
                            const uint32_t *op);
csmv_status csmv_memcd_read_stream(csmv_engine *e, uint64_t *key, uint32_t *op);
/* Serves the whole stream with one persistent kernel (replaces the per-batch memcdTx
 * launches with a persistent server_kernel, memcdSH.cu:603-651). */
csmv_status csmv_memcd_run(csmv_engine *e, csmv_stats *stats);
/* Per request: commit ts (writers) or snapshot (read-only), result word (bit 0 hit,
 * bit 1 wrote, bits 4..7 way), GET value checksum, failed attempts, latency in
 * wall-clock ticks from pick-up to commit. */
csmv_status csmv_memcd_read_log(csmv_engine *e, uint32_t *ts, uint32_t *res, uint32_t *vsum,
                                uint32_t *aborts, uint32_t *lat_ticks);
/* Rate of the device wall clock used for latencies (kHz). */
csmv_status csmv_wallclock_khz(csmv_engine *e, uint32_t *khz);

/* ---- caller-supplied kernels -------------------------------------------------------------
 * The reference's drivers are the caller's own kernels: a client kernel calls TXBegin /
 * TXRead / TXReadOnly / TXWrite / TXCommit (collabValidation/csmv_sh/API.cuh:135-151) beside
 * a server kernel built with the application's hook (bank/csmv_sh/exe.cu:56-79, 89-338).
 * Here the caller writes one persistent kernel
 *     __global__ void k(csmv::Ctx c, csmv::ServerLayout L, ...);
 * whose workgroup 0 runs csmv::server_workgroup<R, W, false>(c, L, lds) (csmv_server.hpp)
 * while its other waves take csmv::client_wave(c), run transactions with csmv_device.hpp's
 * tx_begin / tx_read / tx_read_only / tx_write / tx_commit on Tx<R, W>, and end with
 * csmv::client_finish.  csmv_kernel_prepare checks the wire and the grid (every workgroup
 * resident), clears the engine's run state, fills the context and layout the kernel takes
 * (sizes must be sizeof(csmv::Ctx) and sizeof(csmv::ServerLayout)), sets the kernel's LDS
 * attribute and records the start event on the engine's stream; the caller launches
 *     k<<<launch.grid, launch.block, launch.lds_bytes, (hipStream_t)launch.stream>>>(c, L, ...)
 * and calls csmv_kernel_finish, which waits for the kernel and returns its statistics.  The
 * pool must be freshly loaded; the caller owns its stream, its log buffers and the kernel's
 * other arguments.  Such kernels run the normal lane (no serialized lane, no escalation). */
typedef struct {
    const void *kernel;       /* the caller's __global__ function */
    uint32_t reads;           /* R of the kernel's Tx<R, W> (1..64) */
    uint32_t writes;          /* W (1..32, with 1 + R + 2 W <= 69) */
} csmv_kernel_spec;

typedef struct {
    uint32_t grid;            /* 1 server workgroup + the client workgroups */
    uint32_t block;           /* wg_waves * 64 threads */
    uint32_t lds_bytes;       /* dynamic LDS of every workgroup */
    uint32_t reserved;
    void *stream;             /* the engine's hipStream_t: the launch goes on it */
} csmv_launch;

csmv_status csmv_kernel_prepare(csmv_engine *e, const csmv_kernel_spec *spec, void *ctx,
                                size_t ctx_bytes, void *layout, size_t layout_bytes,
                                csmv_launch *launch);
csmv_status csmv_kernel_finish(csmv_engine *e, csmv_stats *stats);

/* ---- bank run without allocation -----------------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * csmv_bank_run of a bank stream without an engine, on device buffers the caller owns.
 * Everything else the run needs (version pool, request and completion rings, replies,
 * mailboxes, control words, committed-value tables) lives in one fixed device array of
 * CSMV_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; the call makes no
 * hipMalloc / hipFree.  A configuration whose layout does not fit the array is
 * CSMV_E_CAPACITY (for example 16,384 accounts at depth 8 with 64 client waves fit).
 * Every call on a GPU uses that GPU's one array, which keeps the last call's contents (each
 * call clears what must start at zero): calls that overlap on one GPU, from several threads
 * or streams, race and must not be made.  The run goes on the caller's stream (NULL: the
 * null stream) and the call returns when it is done; kernel_ms is the host-timed run of the
 * persistent kernel.  Normal lane only (n_serial_waves must be 0; escalate_after allowed);
 * transfers read 2..4 accounts and audits up to 64, as for csmv_bank_run. */
#define CSMV_NOALLOC_SCRATCH_BYTES (4u << 20)
typedef struct {
    const int32_t *init;       /* [n_items] initial balances (device) */
    int32_t *final_state;      /* [n_items] head values after the run (device) */
    const uint32_t *op;        /* [n_txns] kind | amount << 8 | reads << 16 (device) */
    const uint32_t *addr;      /* [rmax][n_txns] operands (device) */
    uint64_t n_txns;           /* 1 .. 2^31 - 1 */
    uint32_t *log_ts;          /* [n_txns] commit ts; audit: snapshot (device) */
    int32_t *log_v;            /* [reads][n_txns] transfer reads; log_v[k] = audit sum (device) */
    uint32_t *log_ab;          /* [n_txns] failed attempts (device) */
} csmv_bank_buffers;

csmv_status csmv_bank_run_noalloc(const csmv_config *cfg, const csmv_bank_buffers *buf,
                                  void *stream, csmv_stats *stats);

/* Same-address device-scope atomic rate (atomics per second), measured by a
 * microbenchmark kernel: the ceiling unit for the per-lane commit protocol. */
csmv_status csmv_atomic_rate(csmv_engine *e, uint32_t waves, uint32_t iters,
                             double *atomics_per_s);

/* ---- serving loop ----------------------------------------------------------------------
 * One persistent kernel serves a cache partition (an engine holding a cache request
 * stream, csmv_memcd_generate) and a bank partition (an engine holding a transfer-only
 * bank stream, read sets up to 64) as requests arrive.  The reference serves cache
 * requests batch by batch from a host loop against a persistent server
 * (memcachedGPU/memcdSH.cu:603-651); here arrivals are a Poisson process per class
 * (rate = offered_rps * pct_cache / 100 for cache requests, the rest for bank
 * transactions), requests wait in device-resident queues, a lane that commits takes the
 * next arrived request at once (continuous batching), and admission is capped by the
 * in-flight version-pool bytes: 12 bytes per item a request may install (cache GET 1,
 * SET 12, bank transfer 2).  The partitions hold disjoint data, so each has its own
 * commit server and clock.  Either engine may be NULL (that class served alone, at the
 * same arrival times).  Both need fresh pools, the STAMP variant with client
 * write-back, and the same wg_waves. */
typedef struct {
    uint64_t seed;          /* arrival process */
    double offered_rps;     /* offered load over both classes, requests per second */
    uint32_t pct_cache;     /* percent of the offered load that is cache requests */
    uint32_t reserved;
    uint64_t pool_budget;   /* cap on admitted in-flight version bytes (0 = none, else >= 1024) */
} csmv_serve_spec;

typedef struct {
    csmv_stats cache;       /* cache partition (its commit server's counters) */
    csmv_stats bank;        /* bank partition */
    double kernel_ms;       /* the serving kernel, HIP events */
    uint64_t admit_stalls;  /* wave rounds whose arrived requests the budget refused */
    uint64_t peak_inflight; /* largest admitted in-flight version bytes */
    uint64_t final_inflight;/* admitted bytes left at the end (0 when every request committed) */
    uint32_t wallclock_khz; /* tick rate of arrivals and latencies */
    uint32_t reserved;
} csmv_serve_stats;

csmv_status csmv_serve_run(csmv_engine *cache, csmv_engine *bank, const csmv_serve_spec *spec,
                           csmv_serve_stats *stats);
/* The same with the cache class split over n_cache engines (0..7 together with the bank
 * partition), each its own commit server and clock: partition i holds a disjoint range of the
 * cache's sets and the requests that map to it (csmv_memcd_load).  The class's arrivals are
 * thinned by partition: each partition's Poisson rate is the class rate times its share of
 * the class's requests.  stats->cache sums the partitions' counters. */
csmv_status csmv_serve_run_multi(csmv_engine *const *caches, uint32_t n_cache, csmv_engine *bank,
                                 const csmv_serve_spec *spec, csmv_serve_stats *stats);
/* Per request of the stream an engine served: arrival tick, and latency from arrival to
 * commit in wall-clock ticks (queueing, admission and every attempt included). */
csmv_status csmv_serve_read_arrivals(csmv_engine *e, uint32_t *arr_ticks, size_t n);
csmv_status csmv_serve_read_latency(csmv_engine *e, uint32_t *lat_ticks, size_t n);

/* ---- learned contention predictor ----------------------------------------------------
 * No counterpart in the reference, whose only abort record is the per-class totals of
 * Statistics (bank/csmv_sh/API.cuh:112-119) and whose drivers dispatch in arrival order
 * (bank/csmv_sh/exe.cu:142-269).  The predictor only chooses the lane of a transfer.
 *
 * Abort counters: CSMV_ABORT_BUCKETS uint32 buckets; an item x maps to bucket
 * h0(x) mod 4096.  Features of a transfer: CSMV_PRED_FEATURES bf16 values (32 hashed
 * read/write-set buckets, 32 hashed floor(log2(1 + counter)) / 8 terms).  Network:
 * 64 -> 64 (ReLU) -> 1 (sigmoid), bf16 weights, fp32 accumulation; parameter vector of
 * CSMV_PRED_PARAMS floats: W1[f][h] (f-major), b1[64], w2[64], b2. */
#define CSMV_ABORT_BUCKETS 4096
#define CSMV_PRED_FEATURES 64
#define CSMV_PRED_PARAMS 4225

/* Glorot-uniform weights and zero biases from a seed; resets the optimizer. */
csmv_status csmv_pred_init(csmv_engine *e, uint64_t seed);
csmv_status csmv_pred_set_params(csmv_engine *e, const float *params, size_t n);
csmv_status csmv_pred_get_params(csmv_engine *e, float *params, size_t n);
/* counters = (counters >> decay_shift) + the abort log of the last completed bank run
 * (failed attempts of every transfer, added to the bucket of each item it touches);
 * decay_shift >= 32 clears first. */
csmv_status csmv_abort_fold(csmv_engine *e, uint32_t decay_shift);
csmv_status csmv_abort_read_counters(csmv_engine *e, uint32_t *out, size_t n);
csmv_status csmv_abort_load_counters(csmv_engine *e, const uint32_t *in, size_t n);
/* Appends the features (under the current counters) and labels (aborted at least once
 * in the last completed run) of up to max_samples transfers of the resident stream, in
 * stream order, to the device training set; *n_total = rows held. */
csmv_status csmv_pred_collect(csmv_engine *e, uint64_t max_samples, uint64_t *n_total);
csmv_status csmv_pred_read_dataset(csmv_engine *e, uint16_t *X, float *y, size_t n);
csmv_status csmv_pred_clear_dataset(csmv_engine *e);
/* steps full-batch Adam steps (beta 0.9 / 0.999, eps 1e-8) on the mean binary
 * cross-entropy of the training set; loss[steps] (nullable) = loss before each step. */
csmv_status csmv_pred_train(csmv_engine *e, uint32_t steps, float lr, float *loss);
/* Host-buffer forms of the forward and backward kernels (X: n x 64 bf16 bits):
 * p[n]; grad[CSMV_PRED_PARAMS] of the mean loss and the loss itself. */
csmv_status csmv_pred_forward(csmv_engine *e, const uint16_t *X, size_t n, float *p);
csmv_status csmv_pred_grad(csmv_engine *e, const uint16_t *X, const float *y, size_t n,
                           float *grad, float *loss);
/* Predicted abort probability of every transaction of the resident stream (0 for
 * audits) under the current counters; p (nullable) gets n_txns floats. */
csmv_status csmv_pred_infer(csmv_engine *e, float tau, float *p, size_t n);
/* Learned routing of the resident stream: transfers with p > tau go to the serialized
 * lane, the rest stay on the normal lane, both in stream order. */
csmv_status csmv_bank_route_learned(csmv_engine *e, float tau, uint64_t *n_hot);
/* The current routing lists (either pointer may be NULL). */
csmv_status csmv_bank_read_route(csmv_engine *e, uint32_t *norm, uint32_t *hot,
                                 uint64_t *n_norm, uint64_t *n_hot);

/* ---- log files ------------------------------------------------------------------------------
 * The reference's only output is the statistics printf (bank/csmv_sh/exe.cu:470-500); here a
 * run's logs can be kept on disk.  Every file is a 64-byte header followed by little-endian
 * arrays, in the order listed per kind:
 *   magic "CSMVLOG" + NUL (8 bytes), u32 version (CSMV_LOG_VERSION), u32 kind, u64 n (entries),
 *   u32 rows, u32 buckets, u64 seed (of the stream the log belongs to), u64 checksum
 *   (FNV-1a 64 over the arrays), 16 bytes zero.
 *   CSMV_LOG_BANK   commit log of a bank stream: ts[n] u32 (commit ts; audit: snapshot),
 *                   v[rows][n] i32 (transfer reads; v[0][k] = audit sum), aborts[n] u32
 *   CSMV_LOG_CACHE  commit log of a cache stream: ts[n], res[n], vsum[n], aborts[n] (u32);
 *                   rows = 0
 *   CSMV_LOG_ABORTS abort log: classes[8] u64 (commits update / read-only, aborts DataAge /
 *                   ReadWrite / pre-validation / RecordAge / unknown, escalated),
 *                   attempts[n] u32 (failed attempts per transaction), counters[buckets] u32
 *                   (per-bucket abort counters, CSMV_ABORT_BUCKETS or 0); rows = 0
 * Writers take host arrays; readers called with NULL arrays return the header only (so the
 * caller can size its buffers), and check magic, version, kind, sizes and checksum
 * (CSMV_E_IO otherwise).  No device is needed. */
#define CSMV_LOG_VERSION 1u
enum { CSMV_LOG_BANK = 1, CSMV_LOG_CACHE = 2, CSMV_LOG_ABORTS = 3 };

typedef struct {
    char magic[8];
    uint32_t version;
    uint32_t kind;
    uint64_t n;
    uint32_t rows;
    uint32_t buckets;
    uint64_t seed;
    uint64_t checksum;
    uint8_t reserved[16];
} csmv_log_header;

csmv_status csmv_log_write_bank(const char *path, uint64_t seed, uint64_t n, uint32_t rows,
                                const uint32_t *ts, const int32_t *v, const uint32_t *aborts);
csmv_status csmv_log_read_bank(const char *path, csmv_log_header *hdr, uint32_t *ts, int32_t *v,
                               uint32_t *aborts);
csmv_status csmv_log_write_cache(const char *path, uint64_t seed, uint64_t n, const uint32_t *ts,
                                 const uint32_t *res, const uint32_t *vsum, const uint32_t *aborts);
csmv_status csmv_log_read_cache(const char *path, csmv_log_header *hdr, uint32_t *ts, uint32_t *res,
                                uint32_t *vsum, uint32_t *aborts);
csmv_status csmv_log_write_aborts(const char *path, uint64_t seed, const csmv_stats *stats,
                                  uint64_t n, const uint32_t *attempts, const uint32_t *counters,
                                  uint32_t buckets);
csmv_status csmv_log_read_aborts(const char *path, csmv_log_header *hdr, uint64_t *classes,
                                 uint32_t *attempts, uint32_t *counters);

#ifdef __cplusplus
}
#endif
#endif
