// csmv_device.hpp -- device-side transaction API of the gfx950 CSMV engine.
//
// Shape kept from the reference's device API (collabValidation/csmv_sh/API.cuh:135-151,
// memcachedGPU/memcdSH.cu:154-231):
//   tx_begin      TXBegin      snapshot timestamp from the global clock
//   tx_read       TXRead       newest version <= snapshot, read set logged, read-own-write
//   tx_read_only  TXReadOnly   same, nothing logged
//   tx_write      TXWrite      write-set upsert (buffered in registers)
//   tx_commit     TXCommit     wave-synchronous commit; returns the commit timestamp
//                              (>= 1) or 0 on abort, like the server's reply
// One transaction per lane.  tx_begin and tx_commit must be called by all 64 lanes of
// the wavefront (uniform control flow); a lane with nothing to commit passes req=false.
//
// Memory layout (structure of arrays in HBM):
//   ht[i]                  head | tail << 16 of item i's version ring
//   slot[i * depth + d]    version << 32 | value (one 64-bit word, read and written
//                          atomically, so a reader never sees a torn pair); an item's
//                          ring is contiguous, so a ring walk stays in one or two lines
// Cross-workgroup traffic uses agent-scope atomics: gfx950's eight XCDs have private
// L2 caches, and agent scope is what keeps them coherent.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace csmv {

constexpr uint32_t kNone = 0xFFFFFFFFu;
// request-ring flag: the request's transactions are executed by the commit server
constexpr uint32_t kDelegateBit = 0x80000000u;
constexpr uint32_t kWave = 64;
// bank stream: op[k] = kind | amount << 8 | reads << 16
constexpr uint32_t kKindTransfer = 0, kKindAudit = 1;
// Serialized lane: a read-only transaction is sent with this amount; the server reads its
// items in commit order without assigning it a timestamp, and replies with its snapshot
// (the timestamp of the last commit it sees) tagged by kSnapshotBit.  Commit timestamps
// stay below 2^31.
constexpr uint32_t kReadOnlyAmt = 0xFFFFFFFFu;
constexpr uint32_t kSnapshotBit = 0x80000000u;

// ---- wave64 primitives --------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id()
{
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// An empty asm statement that takes every value of a round of loads: nothing computed from
// them can be scheduled before the last one has arrived, so the compiler issues the round
// as one batch and waits once, instead of interleaving each load with its use (which it
// does under register pressure, one memory latency per load).
#define CSMV_PIN(i) "+v"(a[i])
template <typename V, int N>
__device__ __forceinline__ void pin_round(V (&a)[N])
{
    if constexpr (N == 4) asm volatile("" : CSMV_PIN(0), CSMV_PIN(1), CSMV_PIN(2), CSMV_PIN(3));
    else if constexpr (N == 8)
        asm volatile("" : CSMV_PIN(0), CSMV_PIN(1), CSMV_PIN(2), CSMV_PIN(3), CSMV_PIN(4), CSMV_PIN(5),
                     CSMV_PIN(6), CSMV_PIN(7));
    else if constexpr (N == 9)
        asm volatile("" : CSMV_PIN(0), CSMV_PIN(1), CSMV_PIN(2), CSMV_PIN(3), CSMV_PIN(4), CSMV_PIN(5),
                     CSMV_PIN(6), CSMV_PIN(7), CSMV_PIN(8));
    else if constexpr (N == 16)
        asm volatile("" : CSMV_PIN(0), CSMV_PIN(1), CSMV_PIN(2), CSMV_PIN(3), CSMV_PIN(4), CSMV_PIN(5),
                     CSMV_PIN(6), CSMV_PIN(7), CSMV_PIN(8), CSMV_PIN(9), CSMV_PIN(10), CSMV_PIN(11),
                     CSMV_PIN(12), CSMV_PIN(13), CSMV_PIN(14), CSMV_PIN(15));
    else {
        // other sizes in groups of eight (each group one batch)
#pragma unroll
        for (int i = 0; i + 7 < N; i += 8)
            asm volatile("" : CSMV_PIN(i), CSMV_PIN(i + 1), CSMV_PIN(i + 2), CSMV_PIN(i + 3),
                         CSMV_PIN(i + 4), CSMV_PIN(i + 5), CSMV_PIN(i + 6), CSMV_PIN(i + 7));
#pragma unroll
        for (int i = N & ~7; i < N; ++i) asm volatile("" : CSMV_PIN(i));
    }
}
#undef CSMV_PIN
// Number of set bits of m below this lane (the reference's popc(ballot & lanemask_lt)).
__device__ __forceinline__ uint32_t prefix_count(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint32_t bcast(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t bcast64(uint64_t v)
{
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t readlane(uint32_t v, uint32_t l)
{
    return __builtin_amdgcn_readlane(v, l);
}
__device__ __forceinline__ uint64_t readlane64(uint64_t v, uint32_t l)
{
    uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
    uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

// ---- lane exchange over DPP ------------------------------------------------------------
// Wave scans and reductions move values between lanes with DPP row shifts and row
// broadcasts, which the VALU executes as an operand modifier (hipcc folds the move into
// the combining instruction, e.g. v_min_u32_dpp); a shuffle (ds_bpermute) goes through the
// LDS pipeline instead.  GFX9 DPP controls: row_shr:k = 0x110 + k, row_mirror = 0x140,
// row_half_mirror = 0x141, row_bcast:15 = 0x142, row_bcast:31 = 0x143, quad_perm = 0..0xFF.
// Every lane of the wave must be active.
namespace dpp {
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowMirror = 0x140, kRowHalfMirror = 0x141;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143;
constexpr int kQuadSwap1 = 0xB1;   // quad_perm [1,0,3,2]: lane ^ 1
constexpr int kQuadSwap2 = 0x4E;   // quad_perm [2,3,0,1]: lane ^ 2
}  // namespace dpp

// v of the lane the DPP control names; a lane whose source lies outside its row, or whose
// row is masked off, gets `id`.
template <int Ctrl, int RowMask = 0xF>
__device__ __forceinline__ uint32_t dpp_mov(uint32_t id, uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, Ctrl, RowMask, 0xF, false);
}

// Inclusive scan of op over the wave in lane order (id: op's identity).  Within each row
// of 16 lanes, Hillis-Steele steps at distances 1, 2, 4, 8; then the total of row 0 goes
// into row 1 and that of row 2 into row 3 (row_bcast:15 on rows 1 and 3), and the total of
// lanes 0..31 into rows 2 and 3 (row_bcast:31).
template <typename Op>
__device__ __forceinline__ uint32_t wave_scan(uint32_t v, uint32_t id, Op op)
{
    v = op(v, dpp_mov<dpp::kRowShr1>(id, v));
    v = op(v, dpp_mov<dpp::kRowShr2>(id, v));
    v = op(v, dpp_mov<dpp::kRowShr4>(id, v));
    v = op(v, dpp_mov<dpp::kRowShr8>(id, v));
    v = op(v, dpp_mov<dpp::kRowBcast15, 0xA>(id, v));
    v = op(v, dpp_mov<dpp::kRowBcast31, 0xC>(id, v));
    return v;
}

struct OpMin {
    __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a < b ? a : b; }
};
struct OpMax {
    __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a > b ? a : b; }
};
struct OpOr {
    __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a | b; }
};
struct OpAdd {
    __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a + b; }
};

// v of lane src, for any src (a gather such as a producer lookup): ds_bpermute, which
// crosses lanes through the LDS unit without touching LDS memory.  The source lane must be
// active.
__device__ __forceinline__ uint32_t lane_gather(uint32_t v, uint32_t src)
{
    return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)v);
}

// Reductions: the scan's last lane, read into a scalar register (uniform across the wave).
__device__ __forceinline__ uint32_t wave_min(uint32_t v)
{
    return readlane(wave_scan(v, 0xFFFFFFFFu, OpMin{}), 63);
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v)
{
    return readlane(wave_scan(v, 0u, OpMax{}), 63);
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
    return readlane(wave_scan(v, 0u, OpAdd{}), 63);
}
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) { return wave_scan(v, 0u, OpAdd{}); }
__device__ __forceinline__ uint64_t wave_or64(uint64_t v)
{
    const uint32_t lo = readlane(wave_scan((uint32_t)v, 0u, OpOr{}), 63);
    const uint32_t hi = readlane(wave_scan((uint32_t)(v >> 32), 0u, OpOr{}), 63);
    return ((uint64_t)hi << 32) | lo;
}

// ---- agent-scope atomics --------------------------------------------------------------
template <typename T> __device__ __forceinline__ T ld_rlx(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T> __device__ __forceinline__ T ld_acq(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T> __device__ __forceinline__ void st_rlx(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T> __device__ __forceinline__ void st_rel(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T> __device__ __forceinline__ T fetch_add(T *p, T v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void fence_acq() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
// Waits until every memory operation the wave has issued is complete: its agent-scope
// stores are then at the device's point of coherence (the wait an agent-scope release
// makes, without the release's write-back of the L2's dirty lines).
__device__ __forceinline__ void wait_stores() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
__device__ __forceinline__ void fence_rel() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }

// ---- commit-path variant bits (csmv.h CSMV_VAL_* / CSMV_INSERT_* / CSMV_WB_*) -----------
constexpr uint32_t kValScan = 0x0, kValCollab = 0x1, kValStamp = 0x2;
constexpr uint32_t kInsertBatch = 0x4, kServerWriteback = 0x8;
// Pipelined serialized lane (bank driver, STAMP preset): clients stage their requests'
// producer chains, the committer applies each item's request-wide delta with one ordered
// atomic and installs the versions itself.
constexpr uint32_t kSerialPipe = 0x10;

// ---- device error codes (ctx.err) ------------------------------------------------------
enum : uint32_t {
    kErrReplyTimeout = 1,
    kErrRingTimeout = 2,
    kErrMailboxTimeout = 3,
    kErrAdvanceTimeout = 4,
    kErrSlotTimeout = 5,
    kErrDoneRingTimeout = 6,
    kErrLivelock = 7,
    kErrReadSetOverflow = 8,
    kErrPipeNormal = 9,      // a normal-lane update reached the pipelined serialized lane's server
};

// Engine view passed by value to every kernel.
struct Ctx {
    // version pool
    uint32_t *ht;
    uint64_t *slot;
    uint32_t n_items, depth;
    // control words (each on its own 128-byte line)
    uint32_t *gc;            // global clock: every commit <= gc is installed
    uint32_t *err;           // first device error code
    uint32_t *clients_done;  // client waves that have finished
    uint64_t *next_txn;      // next stream index to hand out
    uint64_t *req_tail;      // commit requests issued
    uint64_t *stats;         // kStat* counters
    // messaging
    uint64_t *ring;          // [Q] (n+1) << 32 | client wave
    uint64_t *reply;         // [client_waves * 64] (n+1) << 32 | commit ts
    uint64_t *done;          // [Q] (n+1) << 32 | last ts of batch n (write-back finished)
    uint32_t *mbox;          // [client_waves][rows][64] request mailboxes
    uint64_t *sv;            // [n_items] newest commit of each item: stamp << 32 | value
                             // (server-private; the stamp is 0 before the run's first commit)
    uint64_t *reply_v;       // [client_waves][2W][64] (n+1) << 32 | new value, then predecessor
                             // stamp of each write (delegated); word-major per wave, so one
                             // store or load of 64 lanes covers 512 contiguous bytes
    uint64_t *prof;          // kProf* timing counters (wall_clock64 ticks)
    uint32_t qmask;
    uint32_t n_client_waves, wg_waves;
    uint32_t client_wg_waves; // client waves per client workgroup
    uint32_t variant, atr_entries;
    uint32_t stage_ahead;    // receivers stage request n once request n - stage_ahead is taken
    uint32_t n_serial_waves; // client waves of the serialized lane (the last ones)
    uint32_t begin_wait;     // a wave's next snapshot covers its own last commit
    uint64_t timeout_ticks;
    uint32_t max_attempts;
    uint32_t escalate;       // failed attempts after which a bank transaction is delegated (0 = never)
    // resident bank stream and its commit log
    const uint32_t *op;
    const uint32_t *addr;
    uint64_t n_txns;
    // routing: the normal lane takes norm_list[i] (i itself when norm_list is null),
    // the serialized lane takes hot_list[i]
    const uint32_t *norm_list;
    const uint32_t *hot_list;
    uint64_t n_norm, n_hot;
    uint64_t *next_hot;
    // memcached-style cache: request stream and results
    const uint64_t *req_key;
    const uint32_t *req_op;
    uint32_t n_sets, n_ways, n_slots, mc_batch;
    uint32_t *log_res;     // bit 0 hit / key present, bit 1 wrote, bits 4..7 way
    uint32_t *log_vsum;    // GET hit: value checksum
    uint32_t *log_lat;     // pick-up to commit, wall-clock ticks
    uint32_t *log_ts;
    int32_t *log_v;
    uint32_t *log_ab;
    // serving loop (arr == nullptr: the whole stream is present at launch)
    const uint32_t *arr;   // arrival tick of each request, counted from the serving start
    uint64_t *t_start;     // serving start tick (published by the first client wave)
    uint64_t *inflight;    // admitted in-flight version bytes (shared by the partitions)
    uint64_t budget;       // admission cap on *inflight (0 = none)
    uint64_t *serve_ctr;   // [0] admission rounds refused, [1] peak in-flight bytes
    // pipelined serialized lane: spare words that idle lanes' atomics and stores target,
    // so every step issues the same memory instructions ([kSinkWords] u64)
    uint64_t *sink;
    uint32_t *pipe_prog;   // [commit partitions] timestamp each has published through
    uint32_t pipe_parts;   // commit partitions (pipelined lane; 0 elsewhere)
    uint32_t pipe_heads;   // pipelined lane: 1 = snapshot readers run beside it, so the head
                           // words move during the run (0: set from the install counts after)
};
constexpr uint32_t kSinkWords = 32768;

enum : uint32_t {
    kStatCommitUpd = 0,
    kStatCommitRo,
    kStatAbortAge,
    kStatAbortRw,
    kStatAbortPrevalid,
    kStatAbortRecord,
    kStatBatches,
    kStatFinalTs,
    kStatAbortUnknown,   // serialized lane: an operand's committed value was not at hand
    kStatEscalated,      // normal-lane transactions sent on the serialized lane
    kStatCount
};

// Timing counters (wall_clock64 ticks), the counterpart of the reference's per-thread
// time_rate fields (bank/csmv_sh/API.cuh:121-135).
enum : uint32_t {
    kProfCommitIdle = 0,   // committer waiting for a staged request
    kProfCommitBusy,       // committer deciding
    kProfRecvWindow,       // receivers waiting for the committer to come close
    kProfRecvRing,         // receivers waiting for requests
    kProfRecvBusy,         // receivers staging
    kProfClientReply,      // client waves waiting for replies (sum over waves)
    kProfClientRound,      // client wave rounds (sum over waves)
    kProfClientRounds,     // number of client rounds
    kProfCommitRead,       // committer: mailbox to registers + publication of progress
    kProfCommitResolve,    // committer: committed values / validation
    kProfCommitChain,      // committer: serialized-lane chain sums
    kProfCommitFinish,     // committer: timestamps, tables, replies, log
    kProfCommitSpin,       // committer: re-reading a mailbox the read-ahead found unstaged
    kProfCommitAdv,        // committer: waiting for the advancer to free a completion slot
    kProfCommitMiss,       // committer: requests the read-ahead did not find staged (count)
    kProfCount
};

// Bounded wait: every spin loop in the engine ends, either by its condition or by the
// device timeout, which records an error code that makes all other roles leave too.  The
// clock is first read on the first tick, so a wait whose condition already holds costs no
// clock read.
struct Spin {
    uint64_t t0;
    uint32_t it;
    __device__ Spin() : t0(0), it(0) {}
    __device__ bool tick(const Ctx &c, uint32_t code)
    {
        if (it == 0) t0 = wall_clock64();
        __builtin_amdgcn_s_sleep(1);
        if ((++it & 31u) == 0) {
            if (ld_rlx(c.err)) return false;
            if (wall_clock64() - t0 > c.timeout_ticks) {
                if (lane_id() == 0) {
                    uint32_t expected = 0;
                    __hip_atomic_compare_exchange_strong(c.err, &expected, code,
                                                         __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
                }
                return false;
            }
        }
        return true;
    }
};

__device__ __forceinline__ void raise_error(const Ctx &c, uint32_t code)
{
    uint32_t expected = 0;
    __hip_atomic_compare_exchange_strong(c.err, &expected, code, __ATOMIC_RELAXED,
                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- client waves of a persistent launch ---------------------------------------------------
// Workgroup 0 of a launch is the commit server; client wave cw is wave cw mod
// client_wg_waves of workgroup 1 + cw / client_wg_waves.  Returns the calling wave's client
// index, or kNone for a wave without a role (it leaves at once and never calls
// client_finish).
__device__ __forceinline__ uint32_t client_wave(const Ctx &c)
{
    if (blockIdx.x == 0) return kNone;
    const uint32_t wave = threadIdx.x / kWave;
    if (wave >= c.client_wg_waves) return kNone;
    const uint32_t cw = (blockIdx.x - 1) * c.client_wg_waves + wave;
    return cw < c.n_client_waves ? cw : kNone;
}

// A client wave's last act (all 64 lanes): each lane's commits (updates, read-only) and the
// attempts it abandoned before sending (no version old enough, DataAge) go into the engine
// statistics; then the wave's exit is published, which the commit server's receivers wait
// for before they leave.  Every client wave of a launch must call it once.
__device__ __forceinline__ void client_finish(const Ctx &c, uint64_t commits_upd, uint64_t commits_ro,
                                              uint64_t aborts_dataage)
{
    if (commits_upd) fetch_add(&c.stats[kStatCommitUpd], commits_upd);
    if (commits_ro) fetch_add(&c.stats[kStatCommitRo], commits_ro);
    if (aborts_dataage) fetch_add(&c.stats[kStatAbortAge], aborts_dataage);
    fence_rel();
    if (lane_id() == 0) __hip_atomic_fetch_add(c.clients_done, 1u, __ATOMIC_RELEASE,
                                               __HIP_MEMORY_SCOPE_AGENT);
}

// ---- per-lane transaction state ----------------------------------------------------------
template <int R, int W> struct Tx {
    uint32_t snap;
    uint32_t nr, nw;
    bool aborted;
    bool overflow;
    uint32_t rs[R];
    uint32_t wa[W];
    int32_t wv[W];
};

// Newest version <= snap of item; false when the ring holds none ("DataAge",
// bank/csmv_sh/API.cu:20-24 and 26-40).  The head word is read once; a slot rewritten
// during the walk carries a version newer than the snapshot and is skipped.
__device__ __forceinline__ bool pool_read(const Ctx &c, uint32_t item, uint32_t snap,
                                          int32_t &out)
{
    const uint32_t ht = ld_rlx(&c.ht[item]);
    fence_acq();   // the slots the head names are read after it (see pool_read_chunk)
    const uint32_t h = ht & 0xFFFFu, t = ht >> 16;
    uint32_t j = h;
    for (uint32_t d = 0; d < c.depth; ++d) {
        const uint64_t x = ld_rlx(&c.slot[(size_t)item * c.depth + j]);
        if ((uint32_t)(x >> 32) <= snap) {
            out = (int32_t)(uint32_t)x;
            return true;
        }
        if (j == t) break;
        j = (j == 0) ? c.depth - 1 : j - 1;
    }
    return false;
}

// The ring walk for an item whose head was seen newer than snap: the head now is that
// version or a newer one, so the walk starts one slot below it (an item with a single
// version has none at or below snap).
__device__ __forceinline__ bool pool_read_below(const Ctx &c, uint32_t item, uint32_t snap,
                                                int32_t &out)
{
    const uint32_t ht = ld_rlx(&c.ht[item]);
    fence_acq();
    const uint32_t h = ht & 0xFFFFu, t = ht >> 16;
    if (h == t) return false;
    uint32_t j = (h == 0) ? c.depth - 1 : h - 1;
    for (uint32_t d = 1; d < c.depth; ++d) {
        const uint64_t x = ld_rlx(&c.slot[(size_t)item * c.depth + j]);
        if ((uint32_t)(x >> 32) <= snap) {
            out = (int32_t)(uint32_t)x;
            return true;
        }
        if (j == t) break;
        j = (j == 0) ? c.depth - 1 : j - 1;
    }
    return false;
}

// TXBegin (collabValidation/csmv_sh/API.cu:581-586): wave-uniform snapshot.
// With wait_ts > 0 the snapshot is taken once the clock covers the wave's own last
// commit (TXBegin's prevTs wait, bank/csmv_sh/API.cu:283); otherwise a wave re-reading
// items it just wrote would validate against its own commits.
template <int R, int W>
__device__ __forceinline__ bool tx_begin(const Ctx &c, Tx<R, W> &tx, uint32_t wait_ts = 0)
{
    uint32_t g = bcast(ld_acq(c.gc));
    if (g < wait_ts) {
        Spin sp;
        while ((g = bcast(ld_acq(c.gc))) < wait_ts)
            if (!sp.tick(c, kErrReplyTimeout)) return false;
    }
    tx.snap = g;
    tx.nr = 0;
    tx.nw = 0;
    tx.aborted = false;
    tx.overflow = false;
#pragma unroll
    for (int r = 0; r < R; ++r) tx.rs[r] = kNone;
#pragma unroll
    for (int w = 0; w < W; ++w) { tx.wa[w] = kNone; tx.wv[w] = 0; }
    return true;
}

// TXWrite (bank/csmv_sh/API.cu:69-85): upsert into the register write set.
template <int R, int W>
__device__ __forceinline__ void tx_write(Tx<R, W> &tx, uint32_t item, int32_t v)
{
    // Register arrays are only updated with constant indices and selects, so they stay
    // in VGPRs (a predicated store a[n] = x would be folded into scratch memory).
    bool hit = false;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const bool m = (uint32_t)w < tx.nw && tx.wa[w] == item;
        tx.wv[w] = m ? v : tx.wv[w];
        hit |= m;
    }
    if (!hit) {
        if (tx.nw >= (uint32_t)W) { tx.overflow = true; tx.aborted = true; return; }
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const bool m = (uint32_t)w == tx.nw;
            tx.wa[w] = m ? item : tx.wa[w];
            tx.wv[w] = m ? v : tx.wv[w];
        }
        tx.nw++;
    }
}

template <int R, int W>
__device__ __forceinline__ void log_read(Tx<R, W> &tx, uint32_t item)
{
    if (tx.nr >= (uint32_t)R) { tx.overflow = true; tx.aborted = true; return; }
#pragma unroll
    for (int r = 0; r < R; ++r) tx.rs[r] = (uint32_t)r == tx.nr ? item : tx.rs[r];
    tx.nr++;
}

// TXRead (bank/csmv_sh/API.cu:5-41).
template <int R, int W>
__device__ __forceinline__ int32_t tx_read(const Ctx &c, Tx<R, W> &tx, uint32_t item)
{
    if (tx.aborted) return 0;
    bool own = false;
    int32_t v = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const bool m = (uint32_t)w < tx.nw && tx.wa[w] == item;
        v = m ? tx.wv[w] : v;
        own |= m;
    }
    if (!own && !pool_read(c, item, tx.snap, v)) { tx.aborted = true; return 0; }
    log_read(tx, item);
    return v;
}

// TXReadOnly (bank/csmv_sh/API.cu:43-67).
template <int R, int W>
__device__ __forceinline__ int32_t tx_read_only(const Ctx &c, Tx<R, W> &tx, uint32_t item)
{
    if (tx.aborted) return 0;
    int32_t v = 0;
    if (!pool_read(c, item, tx.snap, v)) { tx.aborted = true; return 0; }
    return v;
}

// Client write-back (TXWriteBack, bank/csmv_sh/API.cu:147-190): evict the oldest version
// when the ring is full, write (ts, value) at head+1, then publish the new head.  The
// global clock is advanced by the server workgroup once every lower batch is installed.
template <int R, int W>
__device__ __forceinline__ void writeback(const Ctx &c, const uint32_t *wa, const int32_t *wv,
                                          uint32_t nw, uint32_t ts)
{
    // the head words of every write in one batch of loads (pin_round), then the stores
    uint32_t nh[W], hw[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const bool on = ts && (uint32_t)w < nw && wa[w] != kNone;
        hw[w] = ld_rlx(&c.ht[on ? wa[w] : 0u]);
    }
    pin_round(hw);
#pragma unroll
    for (int w = 0; w < W; ++w) {
        nh[w] = 0;
        if (ts && (uint32_t)w < nw && wa[w] != kNone) {
            const uint32_t item = wa[w];
            const uint32_t ht = hw[w];
            uint32_t h = ht & 0xFFFFu, t = ht >> 16;
            h = (h + 1 == c.depth) ? 0 : h + 1;
            if (h == t) t = (t + 1 == c.depth) ? 0 : t + 1;
            st_rlx(&c.slot[(size_t)item * c.depth + h],
                   ((uint64_t)ts << 32) | (uint32_t)wv[w]);
            nh[w] = h | (t << 16);
        }
    }
    fence_rel();
#pragma unroll
    for (int w = 0; w < W; ++w)
        if (ts && (uint32_t)w < nw && wa[w] != kNone) st_rlx(&c.ht[wa[w]], nh[w]);
}

template <int R, int W> constexpr uint32_t mbox_rows() { return 1 + R + 2 * W; }

// TXCommit (memcachedGPU/memcdSH.cu:154-231) for one wavefront.  Lanes with req=true
// send (snapshot, read set, write set) through the wave's mailbox; one agent-scope
// fetch_add reserves the request slot for all 64 lanes, one release store publishes it.
// Returns the commit timestamp (0 = aborted).  On a device error *fail is set and the
// caller must leave its loop.
//
// With delegate=true the request carries transactions for the serialized lane: the
// server executes them (rs = operands, wv = transaction id and amount) and returns the
// written values, which replace tx.wv before the client installs them.
//
// A driver whose conflicts are covered by a subset of its writes may send only that
// subset and install the full write set: install_n >= 0 selects install_wa / install_wv
// (install_n entries) instead of the transaction's own write set.
//
// A driver whose read set is a column of its stream sets kStreamRows and passes rows_src
// (read r at rows_src[r * rows_stride], tx.nr entries) instead of keeping tx.rs.
template <int R, int W, int WI = W, bool kStreamRows = false>
__device__ __forceinline__ uint32_t tx_commit(const Ctx &c, uint32_t cw, Tx<R, W> &tx, bool req, bool *fail,
                                               uint64_t *wait_ticks, bool delegate = false,
                                               const uint32_t *install_wa = nullptr,
                                               const int32_t *install_wv = nullptr,
                                               int install_n = -1,
                                               const uint32_t *rows_src = nullptr,
                                               uint64_t rows_stride = 0)
{
    const uint32_t lane = lane_id();
    constexpr uint32_t rows = mbox_rows<R, W>();
    uint32_t *mb = c.mbox + (size_t)cw * rows * kWave;
    st_rlx(&mb[lane], req ? tx.snap : kNone);
#pragma unroll 16
    for (int r = 0; r < R; ++r) {
        const bool on = req && (uint32_t)r < tx.nr;
        uint32_t x;
        if constexpr (kStreamRows) x = on ? rows_src[(uint64_t)r * rows_stride] : kNone;
        else x = on ? tx.rs[r] : kNone;
        st_rlx(&mb[(1 + r) * kWave + lane], x);
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
        st_rlx(&mb[(1 + R + w) * kWave + lane], (req && (uint32_t)w < tx.nw) ? tx.wa[w] : kNone);
        st_rlx(&mb[(1 + R + W + w) * kWave + lane], (uint32_t)tx.wv[w]);
    }
    fence_rel();
    uint64_t n = 0;
    if (lane == 0) {
        n = fetch_add(c.req_tail, (uint64_t)1);
        st_rel(&c.ring[n & c.qmask], ((n + 1) << 32) | cw | (delegate ? kDelegateBit : 0u));
    }
    n = bcast64(n);

    uint64_t v;
    const uint64_t tw0 = wall_clock64();
    Spin sp;
    // While the request is in flight only lane 0 polls, and only its own reply word: the
    // committer writes the 64 reply words of a request with one store instruction, after
    // the request's other reply words, so once lane 0's word carries the request's tag
    // the wave reads everything once (one memory request per poll instead of 64 lanes'
    // worth of tagged words).
    {
        const uint64_t *r0 = c.reply + (size_t)cw * kWave;
        uint64_t w0 = 0;
        for (;;) {
            if (lane == 0) w0 = ld_rlx(r0);
            if ((bcast64(w0) >> 32) == n + 1) break;
            if (!sp.tick(c, kErrReplyTimeout)) { *fail = true; return 0; }
        }
    }
    // serialized lane: the reply word and, for each write, the new value and the stamp
    // its install must follow (kNone: the write was superseded inside the request and is
    // not installed), all 2W + 1 words in one batch of loads per poll; every word carries
    // the request's tag, so a word read ahead of its update is seen as stale
    uint64_t xy[2 * W + 1];
    for (;;) {
        if (delegate) {
            const uint64_t *rv = c.reply_v + (size_t)cw * 2 * W * kWave + lane;
#pragma unroll
            for (int q = 0; q < 2 * W; ++q) xy[q] = ld_rlx(&rv[(size_t)q * kWave]);
        }
        xy[2 * W] = ld_rlx(&c.reply[(size_t)cw * kWave + lane]);
        pin_round(xy);
        v = xy[2 * W];
        bool stale = (v >> 32) != n + 1;
        if (delegate) {
            const bool live = req && (uint32_t)v != 0u;
#pragma unroll
            for (int w = 0; w < W; ++w)
                stale |= live && (uint32_t)w < tx.nw &&
                         ((xy[w] >> 32) != n + 1 || (xy[W + w] >> 32) != n + 1);
        }
        if (ballot(stale) == 0) break;
        if (!sp.tick(c, kErrReplyTimeout)) { *fail = true; return 0; }
    }
    const uint32_t ts = req ? (uint32_t)v : 0u;
    if (delegate && ballot(ts != 0)) {
        uint32_t prev[W];
        bool pend[W];
#pragma unroll
        for (int w = 0; w < W; ++w) {
            tx.wv[w] = (int32_t)(uint32_t)xy[w];
            prev[w] = (uint32_t)xy[W + w];
            pend[w] = ts != 0 && (uint32_t)w < tx.nw && prev[w] != kNone && tx.wa[w] != kNone;
            if (!pend[w]) tx.wa[w] = kNone;
        }
        // installs of one item happen in commit order: each write is installed as soon as
        // the item's head carries its predecessor's stamp (the write of the same item by
        // the previous request that wrote it), independently of the wave's other writes
        for (;;) {
            bool more = false, any = false;
            uint32_t nh[W];
            bool go[W];
            // head words, then head slots, of every pending write: two batches of loads
            uint32_t hw[W];
            uint64_t hs[W];
#pragma unroll
            for (int w = 0; w < W; ++w) hw[w] = ld_rlx(&c.ht[pend[w] ? tx.wa[w] : 0u]);
            pin_round(hw);
#pragma unroll
            for (int w = 0; w < W; ++w)
                hs[w] = ld_rlx(&c.slot[(size_t)(pend[w] ? tx.wa[w] : 0u) * c.depth + (hw[w] & 0xFFFFu)]);
            pin_round(hs);
#pragma unroll
            for (int w = 0; w < W; ++w) {
                go[w] = false;
                nh[w] = 0;
                if (pend[w]) {
                    const uint32_t item = tx.wa[w];
                    const uint32_t ht = hw[w];
                    const uint64_t head = hs[w];
                    if ((uint32_t)(head >> 32) == prev[w]) {
                        uint32_t h = ht & 0xFFFFu, t = ht >> 16;
                        h = (h + 1 == c.depth) ? 0 : h + 1;
                        if (h == t) t = (t + 1 == c.depth) ? 0 : t + 1;
                        st_rlx(&c.slot[(size_t)item * c.depth + h],
                               ((uint64_t)ts << 32) | (uint32_t)tx.wv[w]);
                        nh[w] = h | (t << 16);
                        go[w] = true;
                        any = true;
                    } else {
                        more = true;
                    }
                }
            }
            if (ballot(any)) {
                fence_rel();
#pragma unroll
                for (int w = 0; w < W; ++w)
                    if (go[w]) {
                        st_rlx(&c.ht[tx.wa[w]], nh[w]);
                        pend[w] = false;
                        tx.wa[w] = kNone;
                    }
            }
            if (ballot(more) == 0) break;
            if (!sp.tick(c, kErrReplyTimeout)) { *fail = true; return 0; }
        }
    }
    *wait_ticks += wall_clock64() - tw0;
    if ((c.variant & kServerWriteback) == 0) {
        // The completion word of a request is written by exactly one side: the committer
        // when no lane committed an update (read-only lanes of the serialized lane reply
        // with their snapshot, tagged by kSnapshotBit, and install nothing), the client
        // otherwise, after its installs.
        const uint64_t cm = ballot(ts != 0 && (ts & kSnapshotBit) == 0);
        if (cm) {
            // (the serialized lane has installed every write above)
            if (delegate) {
            } else if (install_n >= 0) {
                writeback<R, WI>(c, install_wa, install_wv, (uint32_t)install_n, ts);
            } else {
                writeback<R, W>(c, tx.wa, tx.wv, tx.nw, ts);
            }
            // the highest committed update holds the request's last timestamp
            const uint32_t last_lane = 63u - (uint32_t)__builtin_clzll(cm);
            const uint32_t end_ts = readlane(ts, last_lane);
            fence_rel();
            if (lane == 0) st_rlx(&c.done[n & c.qmask], ((n + 1) << 32) | end_ts);
        }
    }
    return ts;
}

}  // namespace csmv
