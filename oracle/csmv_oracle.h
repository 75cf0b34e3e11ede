/*
 * csmv_oracle.h -- CPU oracle for the CSMV hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP engine, its
 * C-ABI library, bench.py's GPU leg) may include, link or call this code.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, and only as the checker (or as the timed CPU baseline).
 *
 * Contents
 *   - random streams: splitmix64, the reference LCG RAND_R_FNC
 *     (bank/csmv_sh/API.cuh:37-53), a Zipf CDF and its inverse-CDF draw;
 *   - the bank transaction stream (configs "bank plain" and "bank at
 *     scale"), restating the transfer/audit bodies of
 *     bank/csmv_sh/exe.cu:147-207;
 *   - a serial-replay checker: replays committed update transactions in
 *     commit-timestamp order and checks every logged read, every audit
 *     snapshot, the final state and conservation of the total
 *     (bank/csmv_sh/API.cu:335-350 is the reference's only check);
 *   - the memcached-style cache stream and its serial replay
 *     (memcachedGPU/memcdSH.cu:299-448);
 *   - a multi-threaded CPU restatement of the CSMV protocol
 *     (bank/csmv_sh/API.cu:5-291, bank/csmv_sh/exe.cu:56-79,127-277)
 *     used as the reported CPU baseline.
 */
#ifndef CSMV_ORACLE_H
#define CSMV_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random streams ---------------------------------------------------- */
uint64_t orc_splitmix64(uint64_t x);
/* One step of the reference LCG; returns a 31-bit value, advances *state. */
uint32_t orc_rand_r_fnc(uint64_t *state);
/* Per-transaction generator state: splitmix64(seed + k). */
uint64_t orc_txn_state(uint64_t seed, uint64_t k);

/* Zipf(theta) over n ranks: cdf[i] = floor(2^31 * H(i+1)/H(n)), cdf[n-1] = 2^31. */
int orc_zipf_cdf(uint32_t n, double theta, uint32_t *cdf);
/* First i with r < cdf[i] (r in [0, 2^31)). */
uint32_t orc_zipf_rank(const uint32_t *cdf, uint32_t n, uint32_t r);
/* Rank -> item index (odd-multiplier bijection when n is a power of two). */
uint32_t orc_scramble(uint32_t rank, uint32_t n);

/* ---- bank stream -------------------------------------------------------- */
enum { ORC_KIND_TRANSFER = 0, ORC_KIND_AUDIT = 1 };
enum { ORC_DIST_UNIFORM = 0, ORC_DIST_ZIPF = 1 };

typedef struct {
    uint64_t seed;
    uint64_t n_txns;
    uint32_t n_accounts;
    uint32_t n_reads;      /* reads per transfer (>= 2; first two are the pair) */
    uint32_t n_reads_ro;   /* reads per audit */
    uint32_t pct_ro;       /* percent of audits, 0..100 */
    uint32_t dist;         /* ORC_DIST_* */
    uint32_t amount_max;   /* transfer amount uniform in 1..amount_max */
    uint32_t n_reads_min;  /* 0: transfers read n_reads; else uniform n_reads_min..n_reads */
    double   theta;        /* Zipf exponent */
} orc_bank_spec;

/* op[k] = kind | amount << 8 | n_reads << 16.  addr is [rmax][n_txns] with
 * rmax = max(n_reads, n_reads_ro); unused entries are 0xFFFFFFFF.
 * cdf may be NULL for the uniform distribution. */
int orc_bank_gen(const orc_bank_spec *s, const uint32_t *cdf,
                 uint32_t *op, uint32_t *addr);

typedef struct {
    uint64_t n_updates;
    uint64_t n_audits;
    uint64_t bad_ts;          /* commit ts outside 1..U or duplicated */
    uint64_t bad_snapshot;    /* audit snapshot outside 0..U */
    uint64_t bad_reads;       /* logged update reads != replay state */
    uint64_t bad_audits;      /* logged audit sums != replay state at snapshot */
    uint64_t bad_final;       /* device final values != replay final values */
    int64_t  total;           /* sum of replayed final balances */
    int64_t  expected_total;  /* n_accounts * init */
    int64_t  first_bad_txn;   /* -1 if none */
} orc_replay_report;

/* Replays the log (log_ts[k]: commit ts for transfers, snapshot for audits;
 * log_v[j*n_txns + k]: j-th read of a transfer, log_v[k] = audit sum) and
 * compares with final[] (may be NULL).  Returns the number of failures. */
int64_t orc_bank_replay(const orc_bank_spec *s, const uint32_t *op,
                        const uint32_t *addr, int32_t init,
                        const uint32_t *log_ts, const int32_t *log_v,
                        const int32_t *final, int32_t *replayed,
                        orc_replay_report *rep);

/* Serial execution of the stream in the given order (NULL = stream order);
 * writes the final balances and, optionally, a log in the engine's format
 * (commit ts 1..U in execution order). */
int orc_bank_serial(const orc_bank_spec *s, const uint32_t *op,
                    const uint32_t *addr, int32_t init, const uint64_t *order,
                    int32_t *state, uint32_t *log_ts, int32_t *log_v);

/* ---- multi-threaded CPU CSMV (the CPU baseline) ------------------------- */
typedef struct {
    uint32_t depth;          /* versions per item */
    uint32_t atr_entries;    /* ATR ring capacity */
    uint32_t n_clients;      /* client threads (one server thread is added) */
    uint32_t pin;            /* pin threads to cores 0..n */
} orc_cpu_cfg;

typedef struct {
    uint64_t commits_update, commits_ro;
    uint64_t aborts_dataage, aborts_readwrite, aborts_recordage;
    double   seconds;
    uint32_t final_ts;
    uint32_t threads;
} orc_cpu_stats;

/* Runs transactions [0, n_run) of the stream (n_run <= s->n_txns) with the
 * CSMV protocol on host threads.  Fills state[n_accounts] with head values and
 * the log in the engine's format.  Returns 0 on success. */
int orc_bank_cpu_csmv(const orc_bank_spec *s, const uint32_t *op,
                      const uint32_t *addr, int32_t init, uint64_t n_run,
                      const orc_cpu_cfg *cfg, int32_t *state,
                      uint32_t *log_ts, int32_t *log_v, orc_cpu_stats *st);

/* ---- memcached-style cache (config "memcached KV") -------------------------
 * Restates memcachedGPU/memcdSH.cu:299-448 with 64-bit keys and 32-byte values.
 * The cache is n_sets x n_ways slots; slot s has 12 int32 items, stored field-major
 * (item = field * n_slots + s): state, key low, key high, LRU timestamp, 8 value
 * words.  Request k carries the LRU clock k / batch + 1.
 *   GET: probe the ways for a valid matching key; on a hit read the value and, if
 *        the slot's timestamp is older than the clock, write the clock (the only
 *        write of a GET, so most GETs are read-only).
 *   SET: read state/key/timestamp of every way, pick the slot holding the key,
 *        else the first invalid way, else the least recently used (lowest way on a
 *        tie), and write key, value, clock and VALID.
 * Result word: bit 0 hit (GET) or key present (SET), bit 1 the request wrote,
 * bits 4..7 the way. */
enum { ORC_MC_GET = 0, ORC_MC_SET = 1 };
enum { ORC_MC_STATE = 0, ORC_MC_KLO = 1, ORC_MC_KHI = 2, ORC_MC_TS = 3, ORC_MC_V0 = 4,
       ORC_MC_FIELDS = 12 };

typedef struct {
    uint64_t seed;
    uint64_t n_reqs;
    uint64_t n_keys;       /* keyspace size (< 2^31) */
    uint32_t n_sets;       /* power of two */
    uint32_t n_ways;       /* 1..8 */
    uint32_t pct_get;      /* percent GET, rest SET */
    uint32_t batch;        /* requests per batch */
    uint32_t dist;         /* ORC_DIST_* over key ranks */
    uint32_t reserved;
    double   theta;
} orc_memcd_spec;

uint64_t orc_memcd_key(uint32_t rank);
uint32_t orc_memcd_set_of(uint64_t key, uint32_t n_sets);
uint32_t orc_memcd_value_word(uint64_t key, uint64_t k, uint32_t w);
/* req_key[k], req_op[k]; cdf (over n_keys ranks) may be NULL for uniform keys. */
int orc_memcd_gen(const orc_memcd_spec *s, const uint32_t *cdf, uint64_t *req_key,
                  uint32_t *req_op);

typedef struct {
    uint64_t n_updates, n_readonly, hits, misses, sets;
    uint64_t bad_ts, bad_snapshot, bad_result, bad_kind, bad_final;
    int64_t first_bad_req;
} orc_memcd_report;

/* Serial execution in the given order (NULL = stream order): final items
 * [12 * n_slots] and a log in the engine's format. */
int orc_memcd_serial(const orc_memcd_spec *s, const uint64_t *req_key,
                     const uint32_t *req_op, const uint64_t *order, int32_t *items,
                     uint32_t *log_ts, uint32_t *log_res, uint32_t *log_vsum);
/* Replay of an engine log (ts: commit ts of writers, snapshot of read-only
 * requests); compares results and the final items (may be NULL). */
int64_t orc_memcd_replay(const orc_memcd_spec *s, const uint64_t *req_key,
                         const uint32_t *req_op, const uint32_t *log_ts,
                         const uint32_t *log_res, const uint32_t *log_vsum,
                         const int32_t *final, orc_memcd_report *rep);
/* The same two from a given initial cache (init: 12 * n_slots items; NULL = empty). */
int orc_memcd_serial_from(const orc_memcd_spec *s, const int32_t *init,
                          const uint64_t *req_key, const uint32_t *req_op,
                          const uint64_t *order, int32_t *items, uint32_t *log_ts,
                          uint32_t *log_res, uint32_t *log_vsum);
int64_t orc_memcd_replay_from(const orc_memcd_spec *s, const int32_t *init,
                              const uint64_t *req_key, const uint32_t *req_op,
                              const uint32_t *log_ts, const uint32_t *log_res,
                              const uint32_t *log_vsum, const int32_t *final,
                              orc_memcd_report *rep);

#ifdef __cplusplus
}
#endif
#endif
