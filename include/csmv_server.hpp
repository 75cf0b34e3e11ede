// csmv_server.hpp -- the commit-server workgroup of the gfx950 CSMV engine.
//
// A caller's own persistent kernel runs it in workgroup 0 beside its client waves (see
// csmv_kernel_prepare in csmv.h): server_workgroup<R, W, false>(ctx, layout, lds), with
// the wire R / W of the clients' Tx<R, W>, the context and layout the engine prepared and
// the workgroup's dynamic LDS.
//
// The reference runs one persistent server block: a leader warp scans a ready bitmask
// and hands 32-message chunks to worker warps through shared memory
// (bank/sync_lib/msg_passing.h:85-370, one_phase_server.h:44-108); each worker lane
// validates and inserts its client's record (TXAddToRecord, bank/csmv_sh/API.cu:193-272)
// and answers through a per-warp scoreboard (bank/csmv_sh/exe.cu:56-79).
//
// Here the server is ONE workgroup whose waves have fixed roles:
//   wave 0       committer: takes requests strictly in arrival order from LDS
//                mailboxes and does only what must be serial: the check against
//                commits decided since the request was staged, the lane-order
//                resolution of conflicts inside the request, timestamps, stamping,
//                the reply.  A single wave issues about one VALU op per 4 cycles, so
//                everything that can be computed per request in parallel is moved
//                off this wave.
//   wave 1       clock advancer (client write-back): scans 64 completion words per
//                step and moves the global clock over every installed batch.
//   waves 2..    receivers: request n goes to receiver n mod M, which polls the
//                request ring, copies the client's mailbox into LDS mailbox n mod NMB
//                and precomputes (a) each lane's conflicts with lower lanes of the
//                same request (an LDS hash of the request's write set; the
//                reference's TXPreValidation, batchedAtrInsert/csmv_sh/API.cu:94-191)
//                and (b) in hashed-stamp mode, each lane's conflicts with commits
//                already stamped when the request was staged.
//
// Commit stamps (validation mode STAMP): stamp[x] is the timestamp of the last commit
// that wrote item x.  A lane with snapshot s and read/write items X may commit iff no
// x in X has stamp[x] > s -- the reference's record-ring test (TXValidate,
// bank/csmv_sh/API.cu:104-122) with an unbounded ring, so there are no RecordAge
// aborts.  When all items fit, the exact stamps live in LDS.  Otherwise the exact
// table is in HBM and LDS holds B buckets, bucket[h(x)] = max stamp of the items
// hashing there; the receiver gathers exact stamps as of commit T_g, and the committer
// flags x when bucket[h(x)] > T_g (written after staging).  That is sound, and wrong
// only towards aborting when another item shares the bucket.
#pragma once

#include "csmv_device.hpp"

namespace csmv {

enum : uint32_t {
    kLdsCnext = 0,    // requests whose mailbox the committer has consumed
    kLdsCdone = 1,    // committer finished
    kLdsCfinal = 2,   // requests served (valid once kLdsCdone)
    kLdsAdv = 3,      // batches the advancer has moved the clock over
    kLdsStampTs = 4,  // every stamp <= this value is visible in the HBM stamp table
    // pipelined serialized lane
    kLdsPub = 7,      // requests whose head words and clock the publisher has published
    kLdsEnd = 8,      // 1 + the number of requests, once the stream has ended (0 before)
    kLdsTok = 10,     // 64-bit issue token: the request whose atomics go next | the commit
                      // timestamps assigned before it << 32
    kLdsHdr = 16
};
constexpr uint32_t kExitBit = 0x80000000u;

// Server LDS layout (word offsets), computed on the host.
struct ServerLayout {
    uint32_t nmb;          // LDS mailboxes
    uint32_t mb_rows;      // rows of 64 words per LDS mailbox
    uint32_t mb_off;       // mailboxes
    uint32_t state_off;    // mailbox state words [nmb]
    uint32_t wave_off;     // client wave (| kDelegateBit) per mailbox [nmb]
    uint32_t tg_off;       // T_g per mailbox [nmb]
    uint32_t bucket_off;   // commit-stamp buckets [nbuckets]
    uint32_t nbuckets;     // power of two (STAMP mode)
    uint32_t exact;        // bucket i is item i (n_items <= nbuckets)
    uint32_t val_off;      // exact mode: value of every item [n_items]
    uint32_t sup_off;      // serialized lane: per-receiver chain scratch [M][2 * 64 * RO + 64 * W]
    uint32_t jmp_off;      // serialized lane: committer's base values [64 * RO], new values [64 * W]
    uint32_t atr_off;      // record ring [W][E] (SCAN / COLLAB)
    uint32_t hash_off;     // per-receiver hash tables [M][3][hsz]
    uint32_t hsz;          // power of two >= 2 * 64 * W (the writes of one request); with
                           // collaborative validation >= 2 * 64 * (R + W) (all its items)
    uint32_t harr;         // arrays per receiver hash: key, writer lanes (lo, hi), with
                           // collaborative validation the lanes touching the item (lo, hi),
                           // and for a wide wire a bitmap of the written items (last array)
    uint32_t words;        // total
    uint32_t pipe_off;     // pipelined serialized lane: install records [kPipeRing][2W][64], the
                           // timestamps after each of those requests [kPipeRing] and their
                           // ready words [kPipeRing]
    uint32_t parts;        // pipelined serialized lane: commit server workgroups (item partitions)
};

// The commit partition of an item (pipelined serialized lane with `parts` server workgroups).
__device__ __forceinline__ uint32_t mix32(uint32_t x);
__device__ __forceinline__ uint32_t pipe_part(uint32_t x, uint32_t parts)
{
    return parts == 1 ? 0u : (mix32(x) >> 8) % parts;
}

// Pipelined serialized lane wire.  A request's chain nodes (4 reads x 64 lanes) are split
// by the partition of their item into one list per commit partition: per client wave and
// partition a box of kPipeBox words in HBM, word 0 the header (entries | update lanes <<
// 16), then the entries' items, weights and meta words, kPipeMaxE of each.  A meta word
// holds the entry's root entry (the entry whose atomic returns the value before the
// request: itself, or the canonical node of a written item), the root flag (the entry
// issues its item's atomic), the install flag (the canonical node of a written item: it
// installs the item's last write) and that last writer's rank among the update lanes.
// A partition replies with a value per entry (the value the node reads) and, once those
// stores are complete, a header word per client wave (tag | the request's first timestamp).
// The ring word of a request carries the client wave (bits 0-11), per partition its
// entries in 32-entry chunks (4 bits each from bit 12), so a receiver loads what the list
// holds in one round, and its tag ((n + 1) mod 2^20, from bit 44; the ring holds fewer
// words).
constexpr uint32_t kPipeParts = 8;                       // commit partitions at most
constexpr uint32_t kPipeMaxE = 4 * kWave;                // entries of one request
constexpr uint32_t kPipeItem = kWave, kPipeWgt = kPipeItem + kPipeMaxE, kPipeMeta = kPipeWgt + kPipeMaxE;
constexpr uint32_t kPipeBox = kPipeMeta + kPipeMaxE;    // words per (client wave, partition)
constexpr uint32_t kPipeRootBit = 1u << 8, kPipeInstBit = 1u << 9, kPipeRankShift = 10;
constexpr uint32_t kPipeCwMask = 0xFFFu, kPipeChunkShift = 12, kPipeChunk = 32, kPipeTagShift = 44;
__host__ __device__ constexpr uint64_t pipe_ring_tag(uint64_t n) { return (n + 1) & 0xFFFFFull; }
// entries of partition q in a ring word
__host__ __device__ constexpr uint32_t pipe_ring_entries(uint64_t w, uint32_t q)
{
    return (uint32_t)((w >> (kPipeChunkShift + 4 * q)) & 0xFu) * kPipeChunk;
}
// the ring word's chunk field for partition q's cnt entries
__host__ __device__ constexpr uint64_t pipe_ring_chunks(uint32_t cnt, uint32_t q)
{
    return (uint64_t)((cnt + kPipeChunk - 1) / kPipeChunk) << (kPipeChunkShift + 4 * q);
}
// Install records of this many requests wait for the publisher: per request the item and
// head word of each entry, the entry count, the timestamp after the request, the ready word
// and the word that says its installs are issued.
constexpr uint32_t kPipeRing = 16;
__host__ __device__ constexpr uint32_t pipe_words() { return kPipeRing * (2 * kPipeMaxE + 4); }
// Receiver and committer waves of a pipelined server workgroup of wg waves (one more
// publishes): a receiver takes up to two requests per memory round trip, so a quarter of
// the waves receive (16 waves: 4 receivers, 11 committers, the publisher).
__host__ __device__ constexpr uint32_t pipe_receivers(uint32_t wg) { return wg >= 8 ? wg / 4 : 1; }
__host__ __device__ constexpr uint32_t pipe_committers(uint32_t wg) { return wg - 1 - pipe_receivers(wg); }
// Row indices inside an LDS mailbox.  For a delegated request the wv rows carry the
// transaction id and amount, and prod packs each read's producer (the lower lane whose
// write it must see: lane | slot << 6, 0xFF = none).
// A wide wire (read sets beyond 16) keeps its read set in the staged rows instead of
// registers, validates with stamps only, and on the serialized lane runs the producer
// chains over the transfer operands (the first W reads) alone: the other reads take the
// value their producer wrote, or the committed value.
template <int R> constexpr bool kWideWire = R > 16;
// serialized-lane chain nodes per lane
template <int R, int W> constexpr int kChainReads = kWideWire<R> ? W : R;

// Rows of 64 words in one LDS mailbox for a wire of R reads and W writes (host and
// device): without the serialized lane the rows up to pref, with it the chain rows too.
__host__ __device__ constexpr uint32_t chain_reads(uint32_t R, uint32_t W) { return R > 16 ? W : R; }
__host__ __device__ constexpr uint32_t mailbox_rows(uint32_t R, uint32_t W, bool delegation)
{
    return delegation ? 5 + R + 3 * W + (R + 3) / 4 + 2 * chain_reads(R, W) : 4 + R + 2 * W;
}

template <int R, int W> struct Rows {
    static constexpr int RO = kChainReads<R, W>;
    static constexpr uint32_t snap = 0;
    static constexpr uint32_t items = 1;              // R reads then W writes
    static constexpr uint32_t wv = 1 + R + W;
    static constexpr uint32_t conf_lo = 1 + R + 2 * W;
    static constexpr uint32_t conf_hi = 2 + R + 2 * W;
    static constexpr uint32_t pref = 3 + R + 2 * W;
    static constexpr uint32_t prod = 4 + R + 2 * W;   // (R + 3) / 4 rows of producer codes
    // producer chains resolved at staging: each chain node's root node and the sum of the
    // deltas between them, and the bit mask of this lane's superseded write slots
    static constexpr uint32_t root = prod + (R + 3) / 4;
    static constexpr uint32_t wgt = root + RO;
    static constexpr uint32_t supf = wgt + RO;
    // stamp-bucket index (exact mode: the item) of each write
    static constexpr uint32_t bkt = supf + 1;
    static constexpr uint32_t count = bkt + W;
    static_assert(count == mailbox_rows(R, W, true) && pref + 1 == mailbox_rows(R, W, false),
                  "host layout rows follow the staged rows");
};

__device__ __forceinline__ uint32_t lds_ld_acq(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_st_rel(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void wg_fence() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup"); }
// Orders LDS accesses between lanes of one wave (DS instructions of a wave execute in
// order, so this only constrains the compiler).
__device__ __forceinline__ void wave_fence() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

// The committer's phase timers (read / resolve / chain / finish) cost a clock read each on
// the commit path, so they are compiled in only on request (-DCSMV_PHASE_TIMERS=1).  The
// roles' busy and waiting sums (one clock read per role and step) are kept unless
// -DCSMV_ROLE_TIMERS=0: without them the serialized-lane commit path runs about 2 % faster
// (code generation, not the reads themselves: reading the clocks on one request in eight
// gained nothing).
#ifndef CSMV_PHASE_TIMERS
#define CSMV_PHASE_TIMERS 0
#endif
constexpr bool kPhaseTimers = CSMV_PHASE_TIMERS != 0;
__device__ __forceinline__ uint64_t phase_clock() { return kPhaseTimers ? wall_clock64() : 0ull; }
#ifndef CSMV_ROLE_TIMERS
#define CSMV_ROLE_TIMERS 1
#endif
constexpr bool kRoleTimers = CSMV_ROLE_TIMERS != 0 || kPhaseTimers;
__device__ __forceinline__ uint64_t role_clock() { return kRoleTimers ? wall_clock64() : 0ull; }

__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ uint32_t bucket_of(const ServerLayout &L, uint32_t x)
{
    return L.exact ? x : (mix32(x) & (L.nbuckets - 1));
}

template <int R, int W>
__device__ __forceinline__ bool hits(uint32_t x, const uint32_t *rs, const uint32_t *wa)
{
    bool h = false;
#pragma unroll
    for (int r = 0; r < R; ++r) h |= rs[r] == x;
#pragma unroll
    for (int w = 0; w < W; ++w) h |= wa[w] == x;
    return h;
}

// A running sum held in lane 0 of a vector register (the other lanes add zero).
struct LaneSum {
    uint64_t v = 0;
    __device__ void add(uint64_t x) { v += lane_id() == 0 ? x : 0ull; }
};

// ---- committer --------------------------------------------------------------------------
// Committed-value tables.  Exact mode (every item fits LDS): stamp and value of every
// item in LDS.  Hashed mode: one 64-bit word per item in HBM, c.sv[x] = stamp << 32 |
// value of x's newest commit (written only by the committer, so it holds every commit for
// this wave's own later loads; receivers read the stamp half), and LDS holds the stamp
// buckets that filter the normal lane's validation.  One word per item makes an operand's
// value and stamp one load and an install one store: the committer's gathers and scatters
// are line-granular, and its CU's memory pipeline sets the pace of the commit path.
struct ValueTables {
    uint32_t *bucket, *lval;
};

__device__ __forceinline__ uint64_t sv_load(const uint64_t *sv, uint32_t x)
{
    return __hip_atomic_load(&sv[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t sv_stamp(uint64_t w) { return (uint32_t)(w >> 32); }
// The stamp half alone: a 32-bit load of the word's upper half (little-endian), through a
// type that may alias the word.  The committer stores whole words, so the half read is one
// of the stamps it stored.
typedef uint32_t __attribute__((may_alias)) sv_half_t;
__device__ __forceinline__ uint32_t sv_stamp_load(const uint64_t *sv, uint32_t x)
{
    return __hip_atomic_load((const sv_half_t *)&sv[x] + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// The value half alone (the word's lower half).
__device__ __forceinline__ int32_t sv_value_load(const uint64_t *sv, uint32_t x)
{
    return (int32_t)__hip_atomic_load((const sv_half_t *)&sv[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ int32_t sv_value(uint64_t w) { return (int32_t)(uint32_t)w; }

// Records a commit of value v at timestamp ts for item x: the exact LDS value (the exact
// stamp is the bucket), or the HBM word.  At most one lane of a request records an item
// (the last writer of a serialized-lane chain; a normal-lane lane commits only if no lower
// committed lane writes what it reads or writes), so the store needs no atomic max.
// (The HBM store is an atomic one so that the two branches cannot be merged into one
// flat store, whose completion later LDS waits would also wait for.)
__device__ __forceinline__ void record_value(uint64_t *sv, const ServerLayout &L, const ValueTables &V,
                                             uint32_t x, uint32_t ts, int32_t v)
{
    if (L.exact) V.lval[x] = (uint32_t)v;
    else __hip_atomic_store(&sv[x], ((uint64_t)ts << 32) | (uint32_t)v, __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The bank transfer body executed by the serialized lane on the values the items held
// before this transaction (in[j]); read-own-write as bank/csmv_sh/API.cu:9-18.  A
// read-only transaction (upd false) reads in[j] and writes nothing.
template <int R, int W>
__device__ __forceinline__ void transfer_exec(const uint32_t *it, uint32_t amt, const int32_t *in,
                                              int32_t *rd, const uint32_t *wa, int32_t *nv, bool upd)
{
    // written with selects only (no per-lane branches): a read-only lane never records
    // an own write, so its reads are in[j] unchanged and it writes nothing (wa = kNone)
    uint32_t wi[2] = {kNone, kNone};
    int32_t wvv[2] = {0, 0};
#pragma unroll
    for (int j = 0; j < R; ++j) {
        int32_t v = in[j];
        v = (wi[0] == it[j]) ? wvv[0] : v;
        v = (wi[1] == it[j]) ? wvv[1] : v;
        rd[j] = it[j] == kNone ? 0 : v;
        const bool w_on = upd && it[j] != kNone;
        if (j == 0) {
            wi[0] = w_on ? it[0] : kNone;
            wvv[0] = (int32_t)((uint32_t)v - amt);
        }
        if (j == 1) {
            const int32_t nvb = (int32_t)((uint32_t)v + amt);
            const bool same = w_on && wi[0] == it[1];
            wvv[0] = same ? nvb : wvv[0];
            wi[1] = (w_on && !same) ? it[1] : kNone;
            wvv[1] = nvb;
        }
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
        nv[w] = 0;
        if (wa[w] == wi[0]) nv[w] = wvv[0];
        if (w < 2 && wa[w] == wi[1] && wa[w] != kNone) nv[w] = wvv[1];
    }
}

// kStampCw: the committer is compiled for the STAMP preset (stamp validation, batched
// insertion, client write-back) alone, so the other variants' code and the values it
// keeps live drop out of the loop (fewer scalar registers spilled on the hot path).
template <int R, int W, bool D, bool kStampCw = false>
__device__ __forceinline__ void committer(const Ctx &c, const ServerLayout &L, uint32_t *lds)
{
    using RW = Rows<R, W>;
    // a wide wire is validated with stamps only, and its read set stays in the staged
    // rows (LDS) instead of registers
    constexpr bool kWide = kWideWire<R>;
    const uint32_t lane = lane_id();
    const uint32_t E = c.atr_entries, emask = E - 1;
    const uint32_t vmode = (kWide || kStampCw) ? kValStamp : (c.variant & 0x3u);
    const bool batched = kStampCw || (c.variant & kInsertBatch) != 0;
    const bool server_wb = !kStampCw && (c.variant & kServerWriteback) != 0;
    uint32_t *mbst = lds + L.state_off;
    uint32_t *mbwave = lds + L.wave_off;
    uint32_t *mbtg = lds + L.tg_off;
    uint32_t *atr = lds + L.atr_off;
    ValueTables V;
    V.bucket = lds + L.bucket_off;
    V.lval = lds + L.val_off;
    // serialized lane: base values per chain node (RO per lane), new values per write slot
    constexpr int RO = kChainReads<R, W>;
    int32_t *jbase = (int32_t *)(lds + L.jmp_off);
    int32_t *jnv = jbase + kWave * RO;
    uint64_t *const sv = c.sv;
    uint32_t *const log_ts = c.log_ts;
    int32_t *const log_v = c.log_v;
    uint64_t *const reply = c.reply;
    uint64_t *const reply_v = c.reply_v;
    const uint64_t n_txns = c.n_txns;
    const uint32_t Q = c.qmask + 1;
    const uint64_t lt = (1ull << lane) - 1;

    uint32_t last = 0;  // newest assigned commit timestamp
    // statistics and timing sums, kept in lane 0 of vector registers: as wave-uniform
    // values they would take 20 scalar registers for the whole loop and push the loop's
    // other uniform values out to spill slots
    LaneSum n_rw, n_pv, n_rec, n_unk, t_idle, t_busy, t_read, t_res, t_chain, t_fin, t_spin, t_adv,
        n_miss;
    uint32_t n = 0, mbi = 0, adv_seen = 0;
    uint64_t t_prev = role_clock();   // end of the previous request
    // staged rows (a wide wire keeps its read rows in LDS and reads them on use)
    constexpr int kDr = D && !kWide ? (R + 3) / 4 + W + 2 * RO + 1 : 0;
    constexpr int kNr = kWide ? 3 + 2 * W + 3 : 3 + R + 2 * W + 3 + kDr;
    uint32_t row[kNr];
    // The state word of mailbox m, then every staged row, as one batch of LDS reads: the
    // DS instructions of a wave execute in order and the receiver writes the rows before
    // it releases the state, so rows read after a state that shows the request are that
    // request's (a batch read too early is read again).
    auto read_into = [&](uint32_t m, uint32_t (&row)[kNr]) -> uint32_t {
        const uint32_t *bx = lds + L.mb_off + m * L.mb_rows * kWave;
        const uint32_t st = __hip_atomic_load(&mbst[m], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        wave_fence();
        int q = 0;
        row[q++] = mbwave[m];
        row[q++] = mbtg[m];
        row[q++] = bx[RW::snap * kWave + lane];
        if constexpr (!kWide) {
#pragma unroll
            for (int r = 0; r < R; ++r) row[q++] = bx[(RW::items + r) * kWave + lane];
        }
#pragma unroll
        for (int w = 0; w < W; ++w) row[q++] = bx[(RW::items + R + w) * kWave + lane];
#pragma unroll
        for (int w = 0; w < W; ++w) row[q++] = bx[(RW::wv + w) * kWave + lane];
        row[q++] = bx[RW::conf_hi * kWave + lane];
        row[q++] = bx[RW::conf_lo * kWave + lane];
        row[q++] = bx[RW::pref * kWave + lane];
        if constexpr (kDr > 0) {
#pragma unroll
            for (int j = 0; j < (R + 3) / 4; ++j) row[q++] = bx[(RW::prod + j) * kWave + lane];
#pragma unroll
            for (int w = 0; w < W; ++w) row[q++] = bx[(RW::bkt + w) * kWave + lane];
#pragma unroll
            for (int r = 0; r < RO; ++r) row[q++] = bx[(RW::root + r) * kWave + lane];
#pragma unroll
            for (int r = 0; r < RO; ++r) row[q++] = bx[(RW::wgt + r) * kWave + lane];
            row[q++] = bx[RW::supf * kWave + lane];
        }
        pin_round(row);
        return st;
    };
    // A narrow wire's next request is read ahead while the current one's HBM loads are in
    // flight (its mailbox cannot be restaged before the committer publishes that it took
    // it); a wide wire's mailbox is read on use, so it is read when its turn comes.
    constexpr bool kAhead = !kWide;
    uint32_t ahead[kNr];
    uint32_t ahead_st = 0;
    for (;; ++n, mbi = mbi + 1 == L.nmb ? 0u : mbi + 1) {
        uint32_t s = 0;
        bool bad = false;
        const uint32_t *box = lds + L.mb_off + mbi * L.mb_rows * kWave;
        if (kAhead && (ahead_st & ~kExitBit) == n + 1) {
            s = ahead_st;
#pragma unroll
            for (int i = 0; i < kNr; ++i) row[i] = ahead[i];
        } else {
            const uint64_t ts0 = phase_clock();
            Spin sp;
            for (;;) {
                s = read_into(mbi, row);
                if ((s & ~kExitBit) == n + 1) break;
                if (!sp.tick(c, kErrMailboxTimeout)) { bad = true; break; }
            }
            if constexpr (kPhaseTimers) {
                t_spin.add(phase_clock() - ts0);
                n_miss.add(1);
            }
        }
        if (bad || (s & kExitBit)) break;
        // the completion slot n mod Q must be consumed (the advancer's position is read
        // again only when the last value seen does not cover it)
        if (!server_wb && n >= Q && adv_seen < n - Q + 1) {
            const uint64_t ta0 = phase_clock();
            Spin sp;
            while ((adv_seen = lds_ld_acq(&lds[kLdsAdv])) < n - Q + 1)
                if (!sp.tick(c, kErrDoneRingTimeout)) { bad = true; break; }
            if (bad) break;
            if constexpr (kPhaseTimers) t_adv.add(phase_clock() - ta0);
        }
        const uint64_t t1 = role_clock();
        int q = 0;
        const uint32_t cwr = row[q++];
        const uint32_t cw = cwr & ~kDelegateBit;
        const bool delegate = D && (cwr & kDelegateBit) != 0;
        const uint32_t tg = row[q++];
        const uint32_t snap = row[q++];
        uint32_t rs[kWide ? 1 : R], wa[W], prod[kWide ? 1 : (R + 3) / 4], bkw[W];
        uint32_t root[RO], supf = 0;
        int32_t wv[W], wgt[RO];
        if constexpr (!kWide) {
#pragma unroll
            for (int r = 0; r < R; ++r) rs[r] = row[q++];
        }
        auto RS = [&](int r) -> uint32_t {
            if constexpr (kWide) return box[(RW::items + r) * kWave + lane];
            else return rs[r];
        };
#pragma unroll
        for (int w = 0; w < W; ++w) wa[w] = row[q++];
#pragma unroll
        for (int w = 0; w < W; ++w) wv[w] = (int32_t)row[q++];
        const uint32_t chi = row[q++];
        const uint64_t conf = ((uint64_t)chi << 32) | row[q++];
        const bool pref = row[q++] != 0;
        if constexpr (kDr > 0) {
#pragma unroll
            for (int j = 0; j < (R + 3) / 4; ++j) prod[j] = row[q++];
#pragma unroll
            for (int w = 0; w < W; ++w) bkw[w] = row[q++];
#pragma unroll
            for (int r = 0; r < RO; ++r) root[r] = row[q++];
#pragma unroll
            for (int r = 0; r < RO; ++r) wgt[r] = (int32_t)row[q++];
            supf = row[q++];
        } else if (delegate) {
            // wide wire: the delegate rows of the chain nodes and writes
#pragma unroll
            for (int r = 0; r < RO; ++r) {
                root[r] = box[(RW::root + r) * kWave + lane];
                wgt[r] = (int32_t)box[(RW::wgt + r) * kWave + lane];
            }
            supf = box[RW::supf * kWave + lane];
#pragma unroll
            for (int w = 0; w < W; ++w) bkw[w] = box[(RW::bkt + w) * kWave + lane];
        }
        // producer code of read r: lane | write slot << 6 (0xFF: no lower writer)
        auto PROD = [&](int r) -> uint32_t {
            const uint32_t word = kWide ? box[(RW::prod + r / 4) * kWave + lane] : prod[kWide ? 0 : r / 4];
            return (word >> (8 * (r % 4))) & 0xFFu;
        };
        // hashed mode: the committed values of the chain nodes and the stamps of the
        // writes from the HBM tables (which hold this wave's own earlier stores), issued
        // as soon as the request's rows are in registers
        uint64_t hsv[RO];
        if (delegate && !L.exact) {
#pragma unroll
            for (int r = 0; r < RO; ++r) {
                const uint32_t x = RS(r);
                hsv[r] = sv_load(sv, x != kNone ? x : 0u);
            }
        }
        // mailbox consumed; every stamp <= last is in the HBM tables for the waves of this
        // workgroup that acquire either word (one release fence orders both stores, so the
        // wave waits for its pending LDS operations once)
        if (lane == 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __hip_atomic_store(&lds[kLdsStampTs], last, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_store(&lds[kLdsCnext], n + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        if constexpr (kAhead) ahead_st = read_into(mbi + 1 == L.nmb ? 0u : mbi + 1, ahead);
        const uint64_t tr = phase_clock();
        uint64_t tres = tr, tchain = tr;
        const bool valid = snap != kNone;
        uint32_t myts = 0;
        uint32_t nc = 0;

        if (delegate) {
            // ---- serialized lane: the server executes the transactions ----------------
            const uint32_t k = (uint32_t)wv[0], amt = (uint32_t)wv[1];
            // Each chain node's root and the sum of the deltas from it were resolved at
            // staging (stage_chains); the root's value before the request is the one thing
            // that depends on the commit order: the committed value (the exact LDS table,
            // or the HBM table this wave writes).  Chain nodes are the first RO reads
            // (every read of a narrow wire, the transfer operands of a wide one).
            uint32_t ops[RO];
#pragma unroll
            for (int r = 0; r < RO; ++r) ops[r] = RS(r);
            bool is_root[RO];
#pragma unroll
            for (int r = 0; r < RO; ++r) is_root[r] = valid && ops[r] != kNone && root[r] == lane * RO + r;
            int32_t base[RO];
            uint32_t prev[W];
            // One round of loads, issued for every node and write whether needed or not
            // (a load of item 0 stands in for an empty one), then selects: the branch-free
            // form waits once instead of once per item.
            if (L.exact) {
                int32_t lv[RO];
                uint32_t ls[W];
#pragma unroll
                for (int r = 0; r < RO; ++r) lv[r] = (int32_t)V.lval[ops[r] != kNone ? ops[r] : 0u];
#pragma unroll
                for (int w = 0; w < W; ++w) ls[w] = V.bucket[wa[w] != kNone ? wa[w] : 0u];
#pragma unroll
                for (int r = 0; r < RO; ++r) base[r] = is_root[r] ? lv[r] : 0;
#pragma unroll
                for (int w = 0; w < W; ++w) prev[w] = (valid && wa[w] != kNone) ? ls[w] : kNone;
            } else {
                // the committed value and stamp: the HBM words hold those of every commit
                // (this wave's own stores included), loaded when the request was read.  A
                // write's stamp comes with the read of the same item (a transfer reads what
                // it writes); a write no read covers loads its word here.
#pragma unroll
                for (int r = 0; r < RO; ++r) base[r] = is_root[r] ? sv_value(hsv[r]) : 0;
                bool miss[W], any_miss = false;
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    uint32_t st = 0;
                    bool found = false;
#pragma unroll
                    for (int r = 0; r < RO; ++r) {
                        const bool m = !found && ops[r] == wa[w];
                        st = m ? sv_stamp(hsv[r]) : st;
                        found |= m;
                    }
                    miss[w] = valid && wa[w] != kNone && !found;
                    any_miss |= miss[w];
                    prev[w] = (valid && wa[w] != kNone) ? st : kNone;
                }
                if (ballot(any_miss)) {
                    uint32_t ex[W];
#pragma unroll
                    for (int w = 0; w < W; ++w) ex[w] = sv_stamp_load(sv, miss[w] ? wa[w] : 0u);
#pragma unroll
                    for (int w = 0; w < W; ++w) prev[w] = miss[w] ? ex[w] : prev[w];
                }
            }
            tres = phase_clock();
#pragma unroll
            for (int r = 0; r < RO; ++r) jbase[lane * RO + r] = base[r];
            wave_fence();
            int32_t in[RO], rd[RO], nv[W];
#pragma unroll
            for (int r = 0; r < RO; ++r) in[r] = jbase[root[r]] + wgt[r];
            const bool okx = valid;
            tchain = phase_clock();
            // read-only lanes take no timestamp: each reads the state after the commits
            // of the lanes below it and replies with that snapshot
            const bool ro = amt == kReadOnlyAmt;
            transfer_exec<RO, W>(ops, amt, in, rd, wa, nv, !ro);
            const uint64_t committed = ballot(okx && !ro);
            nc = (uint32_t)__builtin_popcountll(committed);
            const uint32_t below = last + (uint32_t)__builtin_popcountll(committed & lt);
            if (okx) myts = ro ? (kSnapshotBit | below) : below + 1;
            last += nc;
            // only the last write of an item in the request is installed (staged flags);
            // no snapshot can fall inside the request's timestamp range, since the clock
            // moves over whole requests
            bool inst[W];
#pragma unroll
            for (int w = 0; w < W; ++w)
                inst[w] = okx && wa[w] != kNone && ((supf >> w) & 1u) == 0;
            if constexpr (kWide) {
#pragma unroll
                for (int w = 0; w < W; ++w) jnv[lane * W + w] = nv[w];
                wave_fence();
                // the other reads see the write of their producer (a lower lane), this
                // lane's own write, or the committed value before the request; their
                // committed values are read before this request's values are recorded.
                // Chunks of 16 reads, pipelined: a chunk's gathers are issued before the
                // previous chunk's values are used (the loop is unrolled, so the compiler's
                // wait counts cover exactly the loads they need).
                constexpr int kNc = (R - RO + 15) / 16;
                uint32_t cx[2][16], cneed[2];
                int32_t cv[2][16];
                auto prep = [&](int ci, uint32_t (&x)[16], uint32_t &need, int32_t (&v)[16]) {
                    const int c0 = RO + 16 * ci;
#pragma unroll
                    for (int j = 0; j < 16; ++j) x[j] = (c0 + j < R) ? RS(c0 + j) : kNone;
                    // the committed values the reads need (a read that does not need one
                    // loads item 0)
                    need = 0;
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        const bool own = x[j] != kNone && (x[j] == wa[0] || x[j] == wa[1]);
                        const uint32_t pj = (c0 + j < R) ? PROD(c0 + j) : 0xFFu;
                        if (okx && x[j] != kNone && !own && pj == 0xFFu) need |= 1u << j;
                    }
                    if (L.exact) {
#pragma unroll
                        for (int j = 0; j < 16; ++j)
                            v[j] = (int32_t)V.lval[((need >> j) & 1u) ? x[j] : 0u];
                    } else {
#pragma unroll
                        for (int j = 0; j < 16; ++j) v[j] = sv_value_load(sv, ((need >> j) & 1u) ? x[j] : 0u);
                    }
                };
                prep(0, cx[0], cneed[0], cv[0]);
#pragma unroll
                for (int ci = 0; ci < kNc; ++ci) {
                    const int cur = ci & 1;
                    if (ci + 1 < kNc) prep(ci + 1, cx[cur ^ 1], cneed[cur ^ 1], cv[cur ^ 1]);
                    const int c0 = RO + 16 * ci;
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        if (!okx || ro || cx[cur][j] == kNone) continue;
                        const uint32_t pj = PROD(c0 + j);
                        int32_t val = ((cneed[cur] >> j) & 1u) ? cv[cur][j] : 0;
                        if (pj != 0xFFu) val = jnv[(pj & 63u) * W + (pj >> 6)];
#pragma unroll
                        for (int w = 0; w < W; ++w) val = (wa[w] == cx[cur][j]) ? nv[w] : val;
                        log_v[(uint64_t)(c0 + j) * n_txns + k] = val;
                    }
                }
            }
            // the last writer of each item records it and raises the item's bucket: its
            // timestamp is the highest of the item's writers in the request (lower writers
            // of a hot item would only queue on the same LDS word)
#pragma unroll
            for (int w = 0; w < W; ++w)
                if (inst[w]) {
                    record_value(sv, L, V, wa[w], myts, nv[w]);
                    __hip_atomic_fetch_max(&V.bucket[bkw[w]], myts, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            if (okx) {
                // a transfer logs its timestamp and every read; an audit its snapshot and
                // the sum of its reads (exe.cu:213-223), in row 0
                uint32_t sum = 0;
#pragma unroll
                for (int r = 0; r < RO; ++r) sum += (uint32_t)rd[r];
                log_ts[k] = ro ? below : myts;
                log_v[k] = ro ? (int32_t)sum : rd[0];
#pragma unroll
                for (int r = 1; r < RO; ++r)
                    if (!ro && ops[r] != kNone) log_v[(uint64_t)r * n_txns + k] = rd[r];
            }
            uint64_t *rv = reply_v + (size_t)cw * 2 * W * kWave + lane;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                st_rlx(&rv[(size_t)w * kWave], ((uint64_t)(n + 1) << 32) | (uint32_t)nv[w]);
                st_rlx(&rv[(size_t)(W + w) * kWave], ((uint64_t)(n + 1) << 32) | (inst[w] ? prev[w] : kNone));
            }
        } else {
            bool rec = false, conflict = false;
            if (vmode == kValStamp) {
                if (valid && L.exact) {
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (RS(r) != kNone) conflict |= V.bucket[RS(r)] > snap;
#pragma unroll
                    for (int w = 0; w < W; ++w)
                        if (wa[w] != kNone) conflict |= V.bucket[wa[w]] > snap;
                } else if (valid) {
                    // stamps <= T_g were checked at staging (pref); an item whose bucket
                    // was written since is checked against its exact stamp in HBM
                    conflict = pref;
                    uint32_t ex[R + W];
                    if constexpr (kWide) {
#pragma unroll
                        for (int r = 0; r < R + W; ++r) {
                            const uint32_t x = r < R ? RS(r) : wa[r - R];
                            ex[r] = 0;
                            if (x != kNone && V.bucket[bucket_of(L, x)] > tg) ex[r] = sv_stamp_load(sv, x);
                        }
                    } else {
                        // the flagged items' stamps in one batch of loads (an item not
                        // flagged reads item 0 and is ignored), skipped when no lane has one
                        uint32_t bx[R + W];
                        bool fl[R + W], anyf = false;
#pragma unroll
                        for (int r = 0; r < R + W; ++r) {
                            const uint32_t x = r < R ? RS(r) : wa[r - R];
                            bx[r] = V.bucket[x != kNone ? bucket_of(L, x) : 0u];
                        }
#pragma unroll
                        for (int r = 0; r < R + W; ++r) {
                            const uint32_t x = r < R ? RS(r) : wa[r - R];
                            fl[r] = x != kNone && bx[r] > tg;
                            anyf |= fl[r];
                        }
#pragma unroll
                        for (int r = 0; r < R + W; ++r) ex[r] = 0;
                        if (ballot(anyf)) {
#pragma unroll
                            for (int r = 0; r < R + W; ++r) {
                                const uint32_t x = r < R ? RS(r) : wa[r - R];
                                ex[r] = sv_stamp_load(sv, fl[r] ? x : 0u);
                            }
                            pin_round(ex);
#pragma unroll
                            for (int r = 0; r < R + W; ++r) ex[r] = fl[r] ? ex[r] : 0u;
                        }
                    }
#pragma unroll
                    for (int r = 0; r < R + W; ++r) conflict |= ex[r] > snap;
                }
            } else {
                // record ring indexed by timestamp: entry t lives at t mod E and the ring
                // covers (last - E, last]; an older snapshot cannot be validated
                // ("RecordAge", bank/csmv_sh/API.cu:93-100).
                rec = valid && (last - snap) > E;
                const bool live = valid && !rec;
                const uint64_t lm = ballot(live);
                if (lm) {
                    // TXValidate: every lane walks back from the newest record to its own
                    // snapshot (bank/csmv_sh/API.cu:104-122); lanes read the same record at
                    // the same step, so the LDS read is a broadcast.  With collaborative
                    // validation the receiver has already scanned the records up to T_g
                    // (pref), so the walk stops at T_g.
                    uint32_t lo;
                    if (vmode == kValCollab) {
                        conflict = live && pref;
                        lo = tg;
                    } else {
                        lo = wave_min(live ? snap : kNone);
                    }
                    for (uint32_t t = last; t > lo; --t) {
                        const bool chk = live && !conflict && t > snap;
                        if (ballot(chk) == 0) break;
                        const uint32_t e = t & emask;
                        if (chk) {
#pragma unroll
                            for (int w = 0; w < W; ++w) {
                                const uint32_t x = atr[w * E + e];
                                if (x != kNone) {
                                    bool h = false;
#pragma unroll
                                    for (int r = 0; r < R; ++r) h |= RS(r) == x;
#pragma unroll
                                    for (int q = 0; q < W; ++q) h |= wa[q] == x;
                                    conflict |= h;
                                }
                            }
                        }
                    }
                }
            }
            const bool ok = valid && !rec && !conflict;
            n_rec.add(__builtin_popcountll(ballot(rec)));
            n_rw.add(__builtin_popcountll(ballot(valid && !rec && conflict)));

            // Lane order inside the request: lane L commits only if no lower lane it
            // conflicts with commits.  Lanes without such conflicts are final at once;
            // the others are decided in ascending order.
            const uint64_t m = ballot(ok);
            uint64_t committed = m;
            uint64_t it = ballot(ok && (conf & m) != 0);
            committed &= ~it;
            while (it) {
                const uint32_t i = (uint32_t)__builtin_ctzll(it);
                it &= it - 1;
                if ((readlane64(conf, i) & committed) == 0) committed |= 1ull << i;
            }
            n_pv.add(__builtin_popcountll(m & ~committed));
            const bool mine = (committed >> lane) & 1ull;
            nc = (uint32_t)__builtin_popcountll(committed);
            auto record = [&](uint32_t ts) {
                if (vmode == kValStamp) {
#pragma unroll
                    for (int w = 0; w < W; ++w)
                        if (wa[w] != kNone) {
                            __hip_atomic_fetch_max(&V.bucket[bucket_of(L, wa[w])], ts, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                            record_value(sv, L, V, wa[w], ts, wv[w]);
                        }
                } else {
#pragma unroll
                    for (int w = 0; w < W; ++w) atr[w * E + (ts & emask)] = wa[w];
                }
            };
            if (batched) {
                // one batched insertion (batchedAtrInsert/csmv_sh/API.cu:296-377):
                // consecutive timestamps in lane order, offset = popc(committed & lt)
                if (mine) {
                    myts = last + 1 + (uint32_t)__builtin_popcountll(committed & lt);
                    record(myts);
                }
                last += nc;
            } else {
                // per-lane insertion (bank/csmv_sh/API.cu:219-265): one record and one
                // head publication per committing lane, in lane order.
                uint64_t jt = committed;
                while (jt) {
                    const uint32_t i = (uint32_t)__builtin_ctzll(jt);
                    jt &= jt - 1;
                    ++last;
                    if (lane == i) {
                        myts = last;
                        record(myts);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                }
            }
            if (server_wb && nc) {
                // the server installs the versions and moves the clock itself
                // (vanilla/csmv_sh/API.cu:269-272, clientWriteback/csmv_sh/API.cu:386-393)
                writeback<R, W>(c, wa, wv, (uint32_t)W, myts);
                fence_rel();
                if (lane == 0) st_rlx(c.gc, last);
            }
        }
        st_rlx(&reply[(size_t)cw * kWave + lane], ((uint64_t)(n + 1) << 32) | myts);
        if (!server_wb && nc == 0 && lane == 0)
            st_rlx(&c.done[n & c.qmask], ((uint64_t)(n + 1) << 32) | last);
        const uint64_t t2 = role_clock();
        t_idle.add(t1 - t_prev);
        t_busy.add(t2 - t1);
        t_prev = t2;
        if constexpr (kPhaseTimers) {
            t_read.add(tr - t1);
            if (delegate) {
                t_res.add(tres - tr);
                t_chain.add(tchain - tres);
                t_fin.add(t2 - tchain);
            } else {
                t_res.add(t2 - tr);
            }
        }
    }
    if (lane == 0) {
        lds[kLdsCfinal] = n;
        lds_st_rel(&lds[kLdsCdone], 1u);
        fetch_add(&c.stats[kStatAbortRw], n_rw.v);
        fetch_add(&c.stats[kStatAbortPrevalid], n_pv.v);
        fetch_add(&c.stats[kStatAbortRecord], n_rec.v);
        fetch_add(&c.stats[kStatAbortUnknown], n_unk.v);
        fetch_add(&c.stats[kStatBatches], (uint64_t)n);
        st_rlx(&c.stats[kStatFinalTs], (uint64_t)last);
        fetch_add(&c.prof[kProfCommitIdle], t_idle.v);
        fetch_add(&c.prof[kProfCommitBusy], t_busy.v);
        fetch_add(&c.prof[kProfCommitRead], t_read.v);
        fetch_add(&c.prof[kProfCommitResolve], t_res.v);
        fetch_add(&c.prof[kProfCommitChain], t_chain.v);
        fetch_add(&c.prof[kProfCommitFinish], t_fin.v);
        if constexpr (kPhaseTimers) {
            fetch_add(&c.prof[kProfCommitSpin], t_spin.v);
            fetch_add(&c.prof[kProfCommitAdv], t_adv.v);
            fetch_add(&c.prof[kProfCommitMiss], n_miss.v);
        }
    }
}

// ---- clock advancer -----------------------------------------------------------------------
// In-order clock advance (bank/csmv_sh/API.cu:177-185 spins each committer on
// GC == ts-1; batchedAtrInsert/csmv_sh/API.cu:257-261 does it once per warp).  Here one
// wave reads 64 consecutive completion words per step and moves the clock over the
// whole completed prefix with a single release store.
__device__ __forceinline__ void advancer(const Ctx &c, uint32_t *lds)
{
    const uint32_t lane = lane_id();
    uint32_t next = 0;
    Spin sp;
    for (;;) {
        const uint32_t idx = next + lane;
        const uint64_t w = ld_rlx(&c.done[idx & c.qmask]);
        const uint64_t okm = ballot((w >> 32) == (uint64_t)idx + 1);
        const uint32_t p = okm == ~0ull ? 64u : (uint32_t)__builtin_ctzll(~okm);
        if (p) {
            fence_acq();
            const uint32_t g = readlane((uint32_t)w, p - 1);
            if (lane == 0) st_rel(c.gc, g);
            next += p;
            if (lane == 0) lds_st_rel(&lds[kLdsAdv], next);
            sp = Spin();
            continue;
        }
        if (lds_ld_acq(&lds[kLdsCdone]) && next >= lds[kLdsCfinal]) break;
        if (!sp.tick(c, kErrAdvanceTimeout)) break;
    }
}

// ---- receivers ------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t hash_claim(uint32_t *hkey, uint32_t hmask, uint32_t x)
{
    uint32_t h = mix32(x) & hmask;
    for (uint32_t p = 0; p <= hmask; ++p) {
        const uint32_t old = atomicCAS(&hkey[h], kNone, x);
        if (old == kNone || old == x) return h;
        h = (h + 1) & hmask;
    }
    return kNone;
}

__device__ __forceinline__ uint32_t hash_find(const uint32_t *hkey, uint32_t hmask, uint32_t x)
{
    uint32_t h = mix32(x) & hmask;
    for (uint32_t p = 0; p <= hmask; ++p) {
        const uint32_t k = __hip_atomic_load(&hkey[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (k == x) return h;
        if (k == kNone) return kNone;
        h = (h + 1) & hmask;
    }
    return kNone;
}

// Collaborative validation at staging: windows of at most this many records are walked by
// every lane in step; longer ones are split across the lanes through the item hash.
constexpr uint32_t kCollabWalk = 64;

// Serialized lane, at staging: resolve the producer chains of a delegated request from its
// staged rows.  A chain node is one of the first RO reads of a lane; its producer (staged
// code: lane | write slot << 6) is the highest lower lane writing the item, i.e. that lane's
// node for the slot.  A transfer changes its first item by -amt and its second by +amt (0
// when both are one item), so the value a node sees is its chain root's value before the
// request plus the deltas of the writers on the way.  Pointer jumping over this wave's LDS
// scratch gives every node its root and that sum in ceil(log2(chain)) rounds; a write slot
// is superseded when a higher lane reads the item from it and writes it too.  Only the
// root's value depends on the commit order, so the committer is left with one gather.
template <int R, int W>
__device__ __forceinline__ void stage_chains(uint32_t *dst, uint32_t *scr, uint32_t lane)
{
    using RW = Rows<R, W>;
    constexpr int RO = RW::RO;
    uint32_t *snext = scr;
    int32_t *sw = (int32_t *)(scr + kWave * RO);
    uint32_t *ssup = scr + 2 * kWave * RO;
    wave_fence();
    const bool valid = dst[RW::snap * kWave + lane] != kNone;
    const uint32_t amt = dst[(RW::wv + 1) * kWave + lane];
    uint32_t ops[RO], wa[W], pj[RO];
#pragma unroll
    for (int r = 0; r < RO; ++r) {
        ops[r] = dst[(RW::items + r) * kWave + lane];
        pj[r] = (dst[(RW::prod + r / 4) * kWave + lane] >> (8 * (r % 4))) & 0xFFu;
    }
#pragma unroll
    for (int w = 0; w < W; ++w) wa[w] = dst[(RW::items + R + w) * kWave + lane];
    // the producer's delta for the slot: shuffled from the producer lane on every lane
    const int32_t d0 = (wa[1] == kNone) ? 0 : -(int32_t)amt, d1 = (int32_t)amt;
    uint32_t nxt[RO];
    int32_t wgt[RO];
#pragma unroll
    for (int r = 0; r < RO; ++r) {
        const uint32_t node = lane * RO + r;
        const bool has_p = pj[r] != 0xFFu;
        const int srcl = (int)(has_p ? (pj[r] & 63u) : lane);
        const int32_t q0 = (int32_t)lane_gather((uint32_t)d0, (uint32_t)srcl),
                      q1 = (int32_t)lane_gather((uint32_t)d1, (uint32_t)srcl);
        nxt[r] = has_p ? (pj[r] & 63u) * RO + (pj[r] >> 6) : node;
        wgt[r] = has_p ? ((pj[r] >> 6) == 0 ? q0 : q1) : 0;
        snext[node] = nxt[r];
        sw[node] = wgt[r];
    }
#pragma unroll
    for (int w = 0; w < W; ++w) ssup[lane * W + w] = 0;
    wave_fence();
    for (uint32_t round = 0; round < 6; ++round) {
        bool more = false;
        uint32_t nn[RO];
        int32_t ww[RO];
#pragma unroll
        for (int r = 0; r < RO; ++r) {
            const uint32_t node = lane * RO + r;
            const uint32_t t = nxt[r];
            const uint32_t tt = snext[t];
            ww[r] = wgt[r] + (t != node ? sw[t] : 0);
            nn[r] = tt;
            more |= tt != t;
        }
        wave_fence();
#pragma unroll
        for (int r = 0; r < RO; ++r) {
            nxt[r] = nn[r];
            wgt[r] = ww[r];
            snext[lane * RO + r] = nn[r];
            sw[lane * RO + r] = ww[r];
        }
        wave_fence();
        if (ballot(more) == 0) break;
    }
#pragma unroll
    for (int r = 0; r < RO; ++r) {
        bool mine_w = false;
#pragma unroll
        for (int w = 0; w < W; ++w) mine_w |= wa[w] == ops[r];
        if (valid && pj[r] != 0xFFu && mine_w) ssup[(pj[r] & 63u) * W + (pj[r] >> 6)] = 1u;
    }
    wave_fence();
    uint32_t supf = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) supf |= ssup[lane * W + w] ? (1u << w) : 0u;
#pragma unroll
    for (int r = 0; r < RO; ++r) {
        dst[(RW::root + r) * kWave + lane] = nxt[r];
        dst[(RW::wgt + r) * kWave + lane] = (uint32_t)wgt[r];
    }
    dst[RW::supf * kWave + lane] = supf;
    wave_fence();
}

template <int R, int W, bool D>
__device__ __forceinline__ void receiver(const Ctx &c, const ServerLayout &L, uint32_t *lds,
                                         uint32_t rid, uint32_t M)
{
    using RW = Rows<R, W>;
    constexpr bool kWide = kWideWire<R>;
    const bool stamp_mode = (c.variant & 0x3u) == kValStamp;
    const uint32_t lane = lane_id();
    const uint64_t lt = (1ull << lane) - 1;
    const uint64_t *const sv = c.sv;
    uint32_t *mbst = lds + L.state_off;
    uint32_t *mbwave = lds + L.wave_off;
    uint32_t *mbtg = lds + L.tg_off;
    // per-receiver hash: key, write-lane mask (lo, hi)
    uint32_t *hkey = lds + L.hash_off + rid * L.harr * L.hsz;
    uint32_t *hwlo = hkey + L.hsz, *hwhi = hkey + 2 * L.hsz;
    // collaborative validation: lanes reading or writing each item
    const bool collab = (c.variant & 0x3u) == kValCollab;
    uint32_t *hrlo = hkey + 3 * L.hsz, *hrhi = hkey + 4 * L.hsz;
    // wide wires: a bitmap of the request's written items (hsz words; bit index from the
    // upper bits of the item's hash), so that a read no write can match skips the probe
    uint32_t *hbits = hkey + (L.harr - 1) * L.hsz;
    const bool hb = kWide && L.harr > (collab ? 5u : 3u);
    auto hbit = [&](uint32_t m) -> uint32_t { return 1u << ((m >> 16) & 31u); };
    auto hword = [&](uint32_t m) -> uint32_t { return (m >> 21) & (L.hsz - 1); };
    const uint32_t E = c.atr_entries, emask = E - 1;
    const uint32_t *atr = lds + L.atr_off;
    const uint32_t hmask = L.hsz - 1;
    const uint32_t mybit = 1u << (lane & 31);
    uint32_t *chain_scr = lds + L.sup_off +
                          rid * (2 * kWave * Rows<R, W>::RO + kWave * W);
    auto wmask = [&](uint32_t h) -> uint64_t {
        return ((uint64_t)__hip_atomic_load(&hwhi[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) << 32) |
               __hip_atomic_load(&hwlo[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    constexpr uint32_t in_rows = mbox_rows<R, W>();
    uint64_t t_window = 0, t_ring = 0, t_busy = 0;
    for (uint32_t n = rid;; n += M) {
        const uint32_t mbi = n % L.nmb;
        const uint64_t t0 = role_clock();
        // stage request n once its mailbox is free and the committer is at most
        // stage_ahead requests behind it (a short window keeps T_g fresh).  The committer
        // publishes its progress as soon as a narrow wire's rows are in registers, but it
        // keeps reading a wide wire's mailbox (read rows, producer codes) until the
        // request is decided: a wide mailbox is free once the committer has taken the
        // request after the one that used it.
        constexpr uint32_t kLag = kWide ? 2u : 1u;
        const uint32_t need = n + kLag > L.nmb ? n + kLag - L.nmb : 0;
        const uint32_t need2 = n > c.stage_ahead ? n - c.stage_ahead : 0;
        const uint32_t want = need > need2 ? need : need2;
        if (want) {
            // a receiver past the last request waits here until the committer ends
            Spin sp;
            bool bad = false;
            while (lds_ld_acq(&lds[kLdsCnext]) < want)
                if (lds_ld_acq(&lds[kLdsCdone]) || !sp.tick(c, kErrSlotTimeout)) {
                    bad = true;
                    break;
                }
            if (bad) break;
        }
        const uint64_t t1 = role_clock();
        uint64_t e;
        bool leave = false;
        {
            Spin sp;
            for (uint32_t polls = 0;; ++polls) {
                e = ld_rlx(&c.ring[n & c.qmask]);
                if ((e >> 32) == (uint64_t)n + 1) break;
                if ((polls & 7u) == 7u && ld_acq(c.clients_done) == c.n_client_waves &&
                    ld_acq(c.req_tail) <= n) {
                    // every client has left and request n will never come
                    if (ld_rlx(c.req_tail) == n && lane == 0)
                        lds_st_rel(&mbst[mbi], kExitBit | (n + 1));
                    leave = true;
                    break;
                }
                if (!sp.tick(c, kErrRingTimeout)) { leave = true; break; }
            }
        }
        if (leave) break;
        const uint64_t t2 = role_clock();
        fence_acq();
        const uint32_t cwr = (uint32_t)e;
        const uint32_t cw = cwr & ~kDelegateBit;
        const bool delegate = D && (cwr & kDelegateBit) != 0;
        const uint32_t *src = c.mbox + (size_t)cw * in_rows * kWave;
        uint32_t *dst = lds + L.mb_off + mbi * L.mb_rows * kWave;
        // commits already visible in the HBM tables when the request is staged
        const uint32_t tg = lds_ld_acq(&lds[kLdsStampTs]);
        if constexpr (kWide) {
            // wide wire (normal lane, STAMP validation): the writes go into the hash
            // first, then the read rows stream through in chunks: load, prefetch the
            // exact stamps, look up the lower-lane writers, copy to the staging rows
            const uint32_t snap = ld_rlx(&src[lane]);
            const bool valid = snap != kNone;
            uint32_t wa[W], wv[W], slot_w[W];
#pragma unroll
            for (int w = 0; w < W; ++w) {
                wa[w] = ld_rlx(&src[(1 + R + w) * kWave + lane]);
                wv[w] = ld_rlx(&src[(1 + R + W + w) * kWave + lane]);
                slot_w[w] = kNone;
                if (valid && wa[w] != kNone) {
                    slot_w[w] = hash_claim(hkey, hmask, wa[w]);
                    if (slot_w[w] != kNone) {
                        atomicOr(lane < 32 ? &hwlo[slot_w[w]] : &hwhi[slot_w[w]], mybit);
                        if (hb) {
                            const uint32_t m = mix32(wa[w]);
                            atomicOr(&hbits[hword(m)], hbit(m));
                        }
                    }
                }
            }
            wg_fence();
            uint64_t conf = 0;
            uint32_t p = 0;
            constexpr int kChunk = 8;
            static_assert(R % kChunk == 0, "wide wires stage whole chunks of reads");
            // the read rows, a chunk at a time; the next chunk's rows are loaded while this
            // chunk is looked up (for the serialized lane nothing else in the loop goes to
            // memory, so waiting for them waits for nothing more)
            uint32_t xn[kChunk];
#pragma unroll
            for (int j = 0; j < kChunk; ++j) xn[j] = ld_rlx(&src[(1 + j) * kWave + lane]);
#pragma unroll 1
            for (int c0 = 0; c0 < R; c0 += kChunk) {
                uint32_t x[kChunk], st[kChunk], plane[kChunk];
#pragma unroll
                for (int j = 0; j < kChunk; ++j) x[j] = xn[j];
                if (c0 + kChunk < R) {
#pragma unroll
                    for (int j = 0; j < kChunk; ++j) xn[j] = ld_rlx(&src[(1 + c0 + kChunk + j) * kWave + lane]);
                }
                // staging-time stamps of the normal lane (one pinned round of loads; an item
                // not wanted reads item 0); the serialized lane needs none
                if (!L.exact && !delegate) {
                    pin_round(x);
#pragma unroll
                    for (int j = 0; j < kChunk; ++j) {
                        const bool want = valid && x[j] != kNone;
                        st[j] = sv_stamp_load(sv, want ? x[j] : 0u);
                    }
                    pin_round(st);
#pragma unroll
                    for (int j = 0; j < kChunk; ++j) {
                        const bool want = valid && x[j] != kNone;
                        st[j] = want ? st[j] : 0u;
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < kChunk; ++j) st[j] = 0u;
                }
                // the chunk's slots in the write hash, probed together: one LDS round per
                // probe step for all its reads (a read per probe loop would wait the
                // longest probe of the wave once per read)
                uint32_t fs[kChunk], hs[kChunk], mj[kChunk], bw[kChunk];
                bool fp[kChunk];
#pragma unroll
                for (int j = 0; j < kChunk; ++j) mj[j] = mix32(x[j]);
                if (hb) {
#pragma unroll
                    for (int j = 0; j < kChunk; ++j) bw[j] = hbits[hword(mj[j])];
                }
#pragma unroll
                for (int j = 0; j < kChunk; ++j) {
                    fs[j] = kNone;
                    hs[j] = mj[j] & hmask;
                    fp[j] = valid && x[j] != kNone && (!hb || (bw[j] & hbit(mj[j])) != 0);
                }
                for (uint32_t step = 0; step <= hmask; ++step) {
                    uint32_t k[kChunk];
#pragma unroll
                    for (int j = 0; j < kChunk; ++j)
                        k[j] = __hip_atomic_load(&hkey[fp[j] ? hs[j] : 0u], __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP);
                    bool more = false;
#pragma unroll
                    for (int j = 0; j < kChunk; ++j) {
                        const bool hit = fp[j] && k[j] == x[j];
                        const bool next = fp[j] && k[j] != x[j] && k[j] != kNone;
                        fs[j] = hit ? hs[j] : fs[j];
                        hs[j] = next ? ((hs[j] + 1) & hmask) : hs[j];
                        fp[j] = next;
                        more |= next;
                    }
                    if (ballot(more) == 0) break;
                }
                uint64_t wm[kChunk];
#pragma unroll
                for (int j = 0; j < kChunk; ++j) wm[j] = wmask(fs[j] != kNone ? fs[j] : 0u);
#pragma unroll
                for (int j = 0; j < kChunk; ++j) {
                    wm[j] = fs[j] != kNone ? (wm[j] & lt) : 0ull;
                    conf |= wm[j];
                    plane[j] = wm[j] ? 63u - (uint32_t)__builtin_clzll(wm[j]) : 0xFFu;
                    p = st[j] > p ? st[j] : p;
                    dst[(RW::items + c0 + j) * kWave + lane] = x[j];
                }
                if (delegate) {
                    // producer code: the lower writer's lane and the write slot holding
                    // the item (the shuffles run on every lane)
                    uint32_t word[kChunk / 4] = {};
#pragma unroll
                    for (int j = 0; j < kChunk; ++j) {
                        const uint32_t pl = plane[j];
                        const int srcl = (int)(pl == 0xFFu ? lane : pl);
                        uint32_t code = 0xFFu;
#pragma unroll
                        for (int w = 0; w < W; ++w) {
                            const uint32_t pw = lane_gather(wa[w], (uint32_t)srcl);
                            code = (pl != 0xFFu && pw == x[j]) ? (pl | ((uint32_t)w << 6)) : code;
                        }
                        word[j / 4] |= code << (8 * (j % 4));
                    }
#pragma unroll
                    for (int q = 0; q < kChunk / 4; ++q)
                        dst[(RW::prod + c0 / 4 + q) * kWave + lane] = word[q];
                }
            }
            // the write rows: lower lanes writing the same items, staging-time stamps
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const uint32_t x = wa[w];
                const uint32_t st = (!L.exact && !delegate && valid && x != kNone) ? sv_stamp_load(sv, x) : 0u;
                p = st > p ? st : p;
                if (valid && x != kNone) {
                    const uint32_t h = hash_find(hkey, hmask, x);
                    if (h != kNone) conf |= wmask(h) & lt;
                }
                dst[(RW::items + R + w) * kWave + lane] = x;
                if (delegate) dst[(RW::bkt + w) * kWave + lane] = x != kNone ? bucket_of(L, x) : 0u;
            }
            wg_fence();
#pragma unroll
            for (int w = 0; w < W; ++w)
                if (slot_w[w] != kNone) {
                    hkey[slot_w[w]] = kNone;
                    hwlo[slot_w[w]] = 0; hwhi[slot_w[w]] = 0;
                    // every bit of the bitmap is one of this request's writes
                    if (hb) hbits[hword(mix32(wa[w]))] = 0;
                }
            dst[RW::snap * kWave + lane] = snap;
#pragma unroll
            for (int w = 0; w < W; ++w) dst[(RW::wv + w) * kWave + lane] = wv[w];
            dst[RW::conf_lo * kWave + lane] = (uint32_t)conf;
            dst[RW::conf_hi * kWave + lane] = (uint32_t)(conf >> 32);
            dst[RW::pref * kWave + lane] = (!L.exact && valid && p > snap) ? 1u : 0u;
        } else {
            const uint32_t snap = ld_rlx(&src[lane]);
            uint32_t it[R + W], wv[W];
    #pragma unroll
            for (int r = 0; r < R + W; ++r) it[r] = ld_rlx(&src[(1 + r) * kWave + lane]);
    #pragma unroll
            for (int w = 0; w < W; ++w) wv[w] = ld_rlx(&src[(1 + R + W + w) * kWave + lane]);
            pin_round(it);
            const bool valid = snap != kNone;

            bool pref = false;
            // normal lane: the exact stamps as of T_g (every load is issued, an empty entry
            // reads item 0, and the round is pinned, so it costs one memory latency)
            if (!L.exact && !delegate) {
                if (stamp_mode) {
                    uint32_t v[R + W];
    #pragma unroll
                    for (int r = 0; r < R + W; ++r)
                        v[r] = sv_stamp_load(sv, it[r] != kNone ? it[r] : 0u);
                    pin_round(v);
                    uint32_t p = 0;
    #pragma unroll
                    for (int r = 0; r < R + W; ++r) p = (it[r] != kNone && v[r] > p) ? v[r] : p;
                    pref = valid && p > snap;
                }
            }

            // Conflicts with lower lanes of the request.  Insert the write set into an LDS
            // hash of item -> lane masks, then look up: conf = lower lanes that write an
            // item this lane reads or writes; for a delegated request, the highest such
            // lane per read is its producer.  The hash is this wave's own, and one wave's
            // LDS operations take effect in program order, so only the compiler needs
            // holding back between the steps.  Every step issues its loads or swaps for
            // all items at once (an idle item swaps kNone for kNone at slot 0, which
            // changes nothing) and waits once.
            uint32_t mx[R + W];
    #pragma unroll
            for (int r = 0; r < R + W; ++r) mx[r] = mix32(it[r]);
            uint32_t slot_w[W], hp[W];
            bool pend[W];
    #pragma unroll
            for (int w = 0; w < W; ++w) {
                slot_w[w] = kNone;
                hp[w] = mx[R + w] & hmask;
                pend[w] = valid && it[R + w] != kNone;
            }
            for (uint32_t step = 0; step <= hmask; ++step) {
                uint32_t old[W];
    #pragma unroll
                for (int w = 0; w < W; ++w) {
                    old[w] = kNone;
                    __hip_atomic_compare_exchange_strong(&hkey[pend[w] ? hp[w] : 0u], &old[w],
                                                         pend[w] ? it[R + w] : kNone, __ATOMIC_RELAXED,
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                // (selects only: per-lane branches cost more than the step itself)
                bool more = false;
    #pragma unroll
                for (int w = 0; w < W; ++w) {
                    const bool got = pend[w] && (old[w] == kNone || old[w] == it[R + w]);
                    slot_w[w] = got ? hp[w] : slot_w[w];
                    hp[w] = (pend[w] && !got) ? ((hp[w] + 1) & hmask) : hp[w];
                    pend[w] = pend[w] && !got;
                    more |= pend[w];
                }
                if (ballot(more) == 0) break;
            }
    #pragma unroll
            for (int w = 0; w < W; ++w)
                if (slot_w[w] != kNone) atomicOr(lane < 32 ? &hwlo[slot_w[w]] : &hwhi[slot_w[w]], mybit);
            asm volatile("" ::: "memory");
            // the reads' slots (a write's slot is the one it claimed)
            uint32_t fs[R], hs[R];
            bool fp[R];
    #pragma unroll
            for (int r = 0; r < R; ++r) {
                fs[r] = kNone;
                hs[r] = mx[r] & hmask;
                fp[r] = valid && it[r] != kNone;
            }
            for (uint32_t step = 0; step <= hmask; ++step) {
                uint32_t k[R];
    #pragma unroll
                for (int r = 0; r < R; ++r)
                    k[r] = __hip_atomic_load(&hkey[fp[r] ? hs[r] : 0u], __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP);
                bool more = false;
    #pragma unroll
                for (int r = 0; r < R; ++r) {
                    const bool hit = fp[r] && k[r] == it[r];
                    const bool next = fp[r] && k[r] != it[r] && k[r] != kNone;
                    fs[r] = hit ? hs[r] : fs[r];
                    hs[r] = next ? ((hs[r] + 1) & hmask) : hs[r];
                    fp[r] = next;
                    more |= next;
                }
                if (ballot(more) == 0) break;
            }
            uint64_t conf = 0;
            uint32_t prod[(R + 3) / 4];
    #pragma unroll
            for (int q = 0; q < (R + 3) / 4; ++q) prod[q] = 0xFFFFFFFFu;
            // producer of read r: the highest lower lane writing the item (0xFF: none)
            uint32_t plane[R];
            {
                uint64_t wm[R + W];
    #pragma unroll
                for (int r = 0; r < R + W; ++r) {
                    const uint32_t h = r < R ? fs[r < R ? r : 0] : slot_w[r < R ? 0 : r - R];
                    const uint32_t at = h != kNone ? h : 0u;
                    const uint64_t m = ((uint64_t)__hip_atomic_load(&hwhi[at], __ATOMIC_RELAXED,
                                                                    __HIP_MEMORY_SCOPE_WORKGROUP) << 32) |
                                       __hip_atomic_load(&hwlo[at], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    wm[r] = h != kNone ? (m & lt) : 0ull;
                }
    #pragma unroll
                for (int r = 0; r < R + W; ++r) conf |= wm[r];
    #pragma unroll
                for (int r = 0; r < R; ++r)
                    plane[r] = (delegate && wm[r]) ? 63u - (uint32_t)__builtin_clzll(wm[r]) : 0xFFu;
            }
            if (delegate) {
                // which write slot of the producer holds the item; the shuffles run on
                // every lane (a ds_bpermute from an inactive lane is undefined)
    #pragma unroll
                for (int r = 0; r < R; ++r) {
                    const uint32_t p = plane[r];
                    const int srcl = (int)(p == 0xFFu ? lane : p);
                    uint32_t code = p;
    #pragma unroll
                    for (int w = 0; w < W; ++w) {
                        const uint32_t pw = lane_gather(it[R + w], (uint32_t)srcl);
                        code = (p != 0xFFu && pw == it[r]) ? (p | ((uint32_t)w << 6)) : code;
                    }
                    const uint32_t sh = 8u * (r % 4);
                    prod[r / 4] = (prod[r / 4] & ~(0xFFu << sh)) | (code << sh);
                }
            }
            asm volatile("" ::: "memory");
            // Collaborative validation at staging (collabValidation/csmv_sh/API.cu:212-264
            // splits one client's window over a warp, one client after another).  Here the
            // records committed before the request was staged, ts in (snap, T_g], are
            // scanned once for the whole request, the 64 lanes splitting the window: every
            // item a record wrote is looked up in an LDS hash of the request's items (item ->
            // lanes reading or writing it), and a lane conflicts when a record newer than
            // its snapshot wrote one of its items.  The committer checks only (T_g, last].
            uint32_t slot_r[R];
    #pragma unroll
            for (int r = 0; r < R; ++r) slot_r[r] = kNone;
            // a record older than T_g - E belongs to lanes the committer fails with
            // RecordAge (its ring no longer covers their window)
            uint32_t clo = collab ? wave_min(valid ? snap : kNone) : kNone;
            if (collab && clo != kNone && tg > E && clo < tg - E) clo = tg - E;
            const uint32_t cwin = (clo != kNone && tg > clo) ? tg - clo : 0u;
            if (collab && cwin <= kCollabWalk) {
                // a short window: every lane walks it, reading the same record at each step
                bool cf = false;
                for (uint32_t t = tg; t > clo; --t) {
                    const bool chk = valid && !cf && t > snap;
                    if (ballot(chk) == 0) break;
                    const uint32_t e = t & emask;
                    if (chk) {
    #pragma unroll
                        for (int w = 0; w < W; ++w) {
                            const uint32_t x = atr[w * E + e];
                            bool h = false;
    #pragma unroll
                            for (int r = 0; r < R + W; ++r) h |= it[r] == x;
                            cf |= x != kNone && h;
                        }
                    }
                }
                pref = valid && cf;
            } else if (collab) {
                // a slot for every read as well (the writes hold theirs)
                uint32_t hq[R];
                bool pq[R];
    #pragma unroll
                for (int r = 0; r < R; ++r) {
                    hq[r] = mx[r] & hmask;
                    pq[r] = valid && it[r] != kNone;
                }
                for (uint32_t step = 0; step <= hmask; ++step) {
                    uint32_t old[R];
    #pragma unroll
                    for (int r = 0; r < R; ++r) {
                        old[r] = kNone;
                        __hip_atomic_compare_exchange_strong(&hkey[pq[r] ? hq[r] : 0u], &old[r],
                                                             pq[r] ? it[r] : kNone, __ATOMIC_RELAXED,
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                    bool more = false;
    #pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const bool got = pq[r] && (old[r] == kNone || old[r] == it[r]);
                        slot_r[r] = got ? hq[r] : slot_r[r];
                        hq[r] = (pq[r] && !got) ? ((hq[r] + 1) & hmask) : hq[r];
                        pq[r] = pq[r] && !got;
                        more |= pq[r];
                    }
                    if (ballot(more) == 0) break;
                }
    #pragma unroll
                for (int r = 0; r < R; ++r)
                    if (slot_r[r] != kNone) atomicOr(lane < 32 ? &hrlo[slot_r[r]] : &hrhi[slot_r[r]], mybit);
    #pragma unroll
                for (int w = 0; w < W; ++w)
                    if (slot_w[w] != kNone) atomicOr(lane < 32 ? &hrlo[slot_w[w]] : &hrhi[slot_w[w]], mybit);
                // the lanes' snapshots, read by the lanes that scan the window
                dst[RW::snap * kWave + lane] = snap;
                asm volatile("" ::: "memory");
                const uint32_t lo = clo;
                uint64_t cm = 0;
                for (uint32_t top = tg; top > lo && lo != kNone; top = (top - lo > kWave) ? top - kWave : lo) {
                    const uint32_t t = top - lane;
                    if (lane < top - lo) {
                        const uint32_t e = t & emask;
    #pragma unroll
                        for (int w = 0; w < W; ++w) {
                            const uint32_t x = atr[w * E + e];
                            const uint32_t h = x != kNone ? hash_find(hkey, hmask, x) : kNone;
                            if (h != kNone) {
                                uint64_t m = ((uint64_t)__hip_atomic_load(&hrhi[h], __ATOMIC_RELAXED,
                                                                          __HIP_MEMORY_SCOPE_WORKGROUP) << 32) |
                                             __hip_atomic_load(&hrlo[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                                m &= ~cm;
                                while (m) {
                                    const uint32_t b = (uint32_t)__builtin_ctzll(m);
                                    m &= m - 1;
                                    if (dst[RW::snap * kWave + b] < t) cm |= 1ull << b;
                                }
                            }
                        }
                    }
                }
                cm = wave_or64(cm);
                pref = valid && ((cm >> lane) & 1ull) != 0;
                asm volatile("" ::: "memory");
            }
    #pragma unroll
            for (int w = 0; w < W; ++w)
                if (slot_w[w] != kNone) {
                    hkey[slot_w[w]] = kNone;
                    hwlo[slot_w[w]] = 0; hwhi[slot_w[w]] = 0;
                    if (collab) { hrlo[slot_w[w]] = 0; hrhi[slot_w[w]] = 0; }
                }
    #pragma unroll
            for (int r = 0; r < R; ++r)
                if (slot_r[r] != kNone) {
                    hkey[slot_r[r]] = kNone;
                    hrlo[slot_r[r]] = 0; hrhi[slot_r[r]] = 0;
                }
    
            dst[RW::snap * kWave + lane] = snap;
    #pragma unroll
            for (int r = 0; r < R + W; ++r) dst[(RW::items + r) * kWave + lane] = it[r];
    #pragma unroll
            for (int w = 0; w < W; ++w) dst[(RW::wv + w) * kWave + lane] = wv[w];
            dst[RW::conf_lo * kWave + lane] = (uint32_t)conf;
            dst[RW::conf_hi * kWave + lane] = (uint32_t)(conf >> 32);
            dst[RW::pref * kWave + lane] = pref ? 1u : 0u;
            if (delegate) {
    #pragma unroll
                for (int q = 0; q < (R + 3) / 4; ++q) dst[(RW::prod + q) * kWave + lane] = prod[q];
    #pragma unroll
                for (int w = 0; w < W; ++w)
                    dst[(RW::bkt + w) * kWave + lane] =
                        it[R + w] == kNone ? 0u : L.exact ? it[R + w] : (mx[R + w] & (L.nbuckets - 1));
            }
        }
        if (delegate) stage_chains<R, W>(dst, chain_scr, lane);
        if (lane == 0) {
            mbwave[mbi] = cwr;
            mbtg[mbi] = tg;
        }
        wg_fence();
        if (lane == 0) lds_st_rel(&mbst[mbi], n + 1);
        const uint64_t t3 = role_clock();
        t_window += t1 - t0;
        t_ring += t2 - t1;
        t_busy += t3 - t2;
    }
    if (lane == 0) {
        fetch_add(&c.prof[kProfRecvWindow], t_window);
        fetch_add(&c.prof[kProfRecvRing], t_ring);
        fetch_add(&c.prof[kProfRecvBusy], t_busy);
    }
}

// ---- pipelined serialized lane --------------------------------------------------------------
// The serialized lane above resolves each request's values with a round of loads of the
// committed table, so the committer waits for that round once per request; and its clients
// install the versions, each install of an item waiting for the previous one on another CU
// (about 1.2 us per hand-off, measured; a Zipf-hot account is written by every request).
// The pipelined form (variant bit kSerialPipe, bank transfers) takes both waits off the
// commit path:
//  - a bank transfer changes an item by a delta known before it runs (-amt, +amt), so an
//    item's value after a request is its value before plus the sum of the request's deltas
//    on it.  The client stages its request's producer chains; each written item has one
//    canonical node (the lowest read of it with no lower writer), which carries the
//    request-wide delta.  The committer adds that delta, and one install, to the item's
//    word (value << 32 | installs) with one 64-bit atomic and gets the word before the
//    request back.  A wave's atomics on one word are performed in the order it issues
//    them, so request n+1's atomics go out before request n's have returned: the commit
//    order of the values is kept by the L2, not by the committer waiting.
//  - the committer installs the last write of each item itself (the install count gives
//    the ring slot), publishes the head words once those slot stores are complete, and the
//    clock once the head words are, so a reader's snapshot only ever covers whole requests
//    whose versions are in place.  Nothing waits for completion: a store issued before a
//    request's atomics is complete once those atomics have returned (one wave's memory
//    operations complete in issue order), which the committer waits for anyway.
// The reference has no counterpart: its server validates and records every client's
// commit (bank/csmv_sh/API.cu:88-272) and its clients write back their own versions.

// Client side (TXCommit for a wave of routed transfers, bank/csmv_sh/exe.cu:184-201 as the
// body), S transactions a lane: a request carries up to 64 S transactions in the order of
// virtual lanes v = s * 64 + lane (every lane's first transaction, then every lane's
// second, ...), so each token step, atomic round, reply and hand-off of the commit servers
// serves 64 S.  The wave stages the request's producer chains in its LDS scratch (writer
// masks of 64 S bits, producer codes virtual lane | write slot << 8, nodes v * R + r), writes
// each partition's entry list to the wave's boxes, sends, and waits for the replies: the
// values the nodes read, from which it runs its transactions and logs them (commit
// timestamp, or kSnapshotBit | snapshot for a read-only transaction, in ts_out; 0 for one
// the request did not carry).  A request carries the virtual lanes, in order, whose nodes
// keep every partition's list within kPipeMaxE; the others wait for the next request.
// The server installs the versions.  Scratch (pipe_client_words, cleared by
// pipe_client_init): the write hash (key, 2 S writer-mask words, canonical node,
// request-wide delta, last writer's rank) of pipe_client_hsz slots, two buffers of (next
// node, weight) per node, its entry and hash slot, per write slot its superseded flag, and
// the staged lists (a block of kPipeMaxE per partition).
// Transactions a lane of the pipelined client, and the write-hash size its request needs
// (the writes fill at most half of it).
constexpr int kPipeSlots = 4;
__host__ __device__ constexpr uint32_t pow2_ceil_u32(uint32_t v)
{
    uint32_t q = 1;
    while (q < v) q <<= 1;
    return q;
}
__host__ __device__ constexpr uint32_t pipe_client_hsz(uint32_t W, uint32_t S)
{
    return pow2_ceil_u32(2 * S * kWave * W);
}
__host__ __device__ constexpr uint32_t pipe_client_words(uint32_t R, uint32_t W, uint32_t S)
{
    return (4 + 2 * S) * pipe_client_hsz(W, S) + 8 * S * kWave * R + 3 * S * kWave * W + S * kWave +
           3 * kPipeParts * kPipeMaxE + kPipeParts;
}

template <int R, int W, int S>
__device__ __forceinline__ void pipe_client_init(uint32_t *scr)
{
    constexpr uint32_t hsz = pipe_client_hsz(W, S);
    const uint32_t lane = lane_id();
    for (uint32_t i = lane; i < hsz; i += kWave) {
        scr[i] = kNone;
#pragma unroll
        for (uint32_t q = 1; q <= 2 * S; ++q) scr[q * hsz + i] = 0;
        scr[(1 + 2 * S) * hsz + i] = kNone;
        scr[(2 + 2 * S) * hsz + i] = 0;
    }
    wave_fence();
}

template <int R, int W, int S>
__device__ __forceinline__ void tx_commit_pipe(const Ctx &c, const ServerLayout &L, uint32_t *scr,
                                                uint32_t cw, const Tx<R, W> (&tx)[S],
                                                const bool (&req)[S], uint32_t (&ts_out)[S],
                                                bool *fail, uint64_t *wait_ticks)
{
    static_assert(!kWideWire<R> && W == 2 && R * kWave <= kPipeMaxE && S >= 1 && S <= 4,
                  "the pipelined lane carries bank transfers");
    constexpr int RO = R;
    constexpr uint32_t kNoP = 0xFFFFu;            // no producer
    constexpr uint32_t NV = S * kWave, NN = NV * RO;   // virtual lanes, nodes
    constexpr uint32_t hsz = pipe_client_hsz(W, S), hmask = hsz - 1;
    const uint32_t lane = lane_id();
    const uint64_t lt = (1ull << lane) - 1;
    const uint32_t P = L.parts;
    // the write hash
    uint32_t *hkey = scr, *hw = scr + hsz;        // hw[i * hsz + slot]: virtual lanes 32 i ..
    uint32_t *canon = scr + (1 + 2 * S) * hsz;
    int32_t *htot = (int32_t *)(scr + (2 + 2 * S) * hsz);
    uint32_t *hrank = scr + (3 + 2 * S) * hsz;
    // per node (v * R + r): chain buffers A and B (next node, weight), the producer code
    // (then the node's entry), hash slot, item and partition
    uint32_t *snext = scr + (4 + 2 * S) * hsz;
    int32_t *sw = (int32_t *)(snext + NN);
    uint32_t *e2 = snext + 4 * NN;
    uint32_t *sfs = e2 + NN;
    uint32_t *sops = sfs + NN;
    uint32_t *spq = sops + NN;
    // per write slot (v * W + w): item, hash slot, superseded flag; per virtual lane: amount
    uint32_t *swa = spq + NN;
    uint32_t *sslot = swa + NV * W;
    uint32_t *ssup = sslot + NV * W;
    uint32_t *samt = ssup + NV * W;
    uint32_t *sitem = samt + NV;                  // [kPipeParts][kPipeMaxE]
    uint32_t *swgt = sitem + kPipeParts * kPipeMaxE;
    uint32_t *smeta = swgt + kPipeParts * kPipeMaxE;
    uint32_t *scnt = smeta + kPipeParts * kPipeMaxE;   // [kPipeParts] entries placed so far
    static_assert(kPipeParts == 8, "entry placement splits on three partition bits");
    auto nd = [&](int s, int r) -> uint32_t { return ((uint32_t)s * kWave + lane) * RO + (uint32_t)r; };
    auto vl = [&](int s) -> uint32_t { return (uint32_t)s * kWave + lane; };
    // a virtual lane's deltas on its two write slots
    auto delta = [&](uint32_t v, uint32_t w) -> int32_t {
        const int32_t a = (int32_t)samt[v];
        return w == 0 ? (swa[v * W + 1] == kNone ? 0 : -a) : a;
    };

    // 1. the operands into LDS; the virtual lanes this request carries: an order prefix
    //    whose nodes keep every partition's list within kPipeMaxE (every lane takes part in
    //    the scans).  Each step below keeps one slot's values in registers at a time.
    bool on[S], ro[S];
    uint32_t kx[S];
    {
        uint32_t pq[S][RO];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            on[s] = req[s];
            ro[s] = (uint32_t)tx[s].wv[1] == kReadOnlyAmt;
            kx[s] = (uint32_t)tx[s].wv[0];
            samt[vl(s)] = (uint32_t)tx[s].wv[1];
#pragma unroll
            for (int r = 0; r < RO; ++r) {
                const uint32_t x = (req[s] && (uint32_t)r < tx[s].nr) ? tx[s].rs[r] : kNone;
                pq[s][r] = x != kNone ? pipe_part(x, P) : kNone;
                sops[nd(s, r)] = x;
                spq[nd(s, r)] = pq[s][r];
            }
#pragma unroll
            for (int w = 0; w < W; ++w)
                swa[vl(s) * W + w] = (req[s] && (uint32_t)w < tx[s].nw) ? tx[s].wa[w] : kNone;
        }
#pragma unroll
        for (uint32_t t = 0; t < kPipeParts; ++t) {
            if (t >= P) break;
            uint32_t below = 0;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                uint32_t n = 0;
#pragma unroll
                for (int r = 0; r < RO; ++r) n += pq[s][r] == t ? 1u : 0u;
                const uint32_t ps = below + wave_incl_scan(n);
                on[s] = on[s] && ps <= kPipeMaxE;
                below = readlane(ps, 63);
            }
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s)
        if (!on[s]) {
#pragma unroll
            for (int r = 0; r < RO; ++r) {
                sops[nd(s, r)] = kNone;
                spq[nd(s, r)] = kNone;
            }
#pragma unroll
            for (int w = 0; w < W; ++w) swa[vl(s) * W + w] = kNone;
        }
    uint32_t rank[S], nc = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const uint64_t u = ballot(on[s] && !ro[s]);
        rank[s] = nc + (uint32_t)__builtin_popcountll(u & lt);
        nc += (uint32_t)__builtin_popcountll(u);
    }
    wave_fence();

    // 2. the write sets into the hash (item -> writer virtual lanes)
    {
        uint32_t a[S][W], slot_w[S][W], hp[S][W];
        bool pend[S][W];
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int w = 0; w < W; ++w) {
                a[s][w] = swa[vl(s) * W + w];
                slot_w[s][w] = kNone;
                hp[s][w] = mix32(a[s][w]) & hmask;
                pend[s][w] = a[s][w] != kNone;
            }
        for (uint32_t step = 0; step <= hmask; ++step) {
            bool more = false;
#pragma unroll
            for (int s = 0; s < S; ++s)
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    uint32_t old = kNone;
                    __hip_atomic_compare_exchange_strong(&hkey[pend[s][w] ? hp[s][w] : 0u], &old,
                                                         pend[s][w] ? a[s][w] : kNone, __ATOMIC_RELAXED,
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    const bool got = pend[s][w] && (old == kNone || old == a[s][w]);
                    slot_w[s][w] = got ? hp[s][w] : slot_w[s][w];
                    hp[s][w] = (pend[s][w] && !got) ? ((hp[s][w] + 1) & hmask) : hp[s][w];
                    pend[s][w] = pend[s][w] && !got;
                    more |= pend[s][w];
                }
            if (ballot(more) == 0) break;
        }
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int w = 0; w < W; ++w) {
                sslot[vl(s) * W + w] = slot_w[s][w];
                if (slot_w[s][w] != kNone)
                    atomicOr(&hw[(vl(s) >> 5) * hsz + slot_w[s][w]], 1u << (vl(s) & 31));
            }
    }
    wave_fence();

    // 3. per slot: each read's hash slot and producer (the highest lower virtual lane
    //    writing the item, and its write slot: v | w << 8), and each written item's
    //    canonical node (the lowest node of the item without a producer)
#pragma unroll
    for (int s = 0; s < S; ++s) {
        uint32_t x[RO], fs[RO], hs[RO];
        bool fp[RO];
#pragma unroll
        for (int r = 0; r < RO; ++r) {
            x[r] = sops[nd(s, r)];
            fs[r] = kNone;
            hs[r] = mix32(x[r]) & hmask;
            fp[r] = x[r] != kNone;
        }
        for (uint32_t step = 0; step <= hmask; ++step) {
            bool more = false;
#pragma unroll
            for (int r = 0; r < RO; ++r) {
                const uint32_t kk = __hip_atomic_load(&hkey[fp[r] ? hs[r] : 0u], __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_WORKGROUP);
                const bool hit = fp[r] && kk == x[r];
                const bool next = fp[r] && kk != x[r] && kk != kNone;
                fs[r] = hit ? hs[r] : fs[r];
                hs[r] = next ? ((hs[r] + 1) & hmask) : hs[r];
                fp[r] = next;
                more |= next;
            }
            if (ballot(more) == 0) break;
        }
#pragma unroll
        for (int r = 0; r < RO; ++r) {
            const uint32_t at = fs[r] != kNone ? fs[r] : 0u;
            uint32_t pv = kNoP;
            // the virtual lanes below v: mask words 0 .. 2 s + 1, the last two as the lanes
            // below the lane (the earlier words whole)
#pragma unroll
            for (int i = 2 * s + 1; i >= 0; --i) {
                const uint32_t lo = (uint32_t)i * 32u;
                const uint32_t below = i < 2 * s ? 0xFFFFFFFFu : (i & 1) ? (uint32_t)(lt >> 32) : (uint32_t)lt;
                const uint32_t raw = __hip_atomic_load(&hw[(uint32_t)i * hsz + at], __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                const uint32_t m = fs[r] != kNone ? (raw & below) : 0u;
                pv = (pv == kNoP && m) ? lo + 31u - (uint32_t)__builtin_clz(m) : pv;
            }
            uint32_t code = kNoP;
            if (pv != kNoP) {
#pragma unroll
                for (int w = 0; w < W; ++w) code = swa[pv * W + w] == x[r] ? (pv | ((uint32_t)w << 8)) : code;
            }
            e2[nd(s, r)] = code;
            sfs[nd(s, r)] = fs[r];
            if (fs[r] != kNone && code == kNoP) atomicMin(&canon[fs[r]], nd(s, r));
        }
        // one slot's loads at a time (fewer registers in flight)
        asm volatile("" ::: "memory");
    }
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int w = 0; w < W; ++w) ssup[vl(s) * W + w] = 0;
    wave_fence();

    // 4. superseded write slots (a write slot is superseded when a higher virtual lane reads
    //    the item from it and writes it too), then the producer chains by pointer jumping
    //    over two buffers of (next node, weight): each node's chain root and the sum of the
    //    writers' deltas between them
    uint32_t *nA = snext, *nB = snext + 2 * NN;
    int32_t *wA = sw, *wB = sw + 2 * NN;
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int r = 0; r < RO; ++r) {
            const uint32_t node = nd(s, r);
            const uint32_t code = e2[node];
            const bool has_p = code != kNoP;
            const uint32_t pv = has_p ? (code & 0xFFu) : 0u, pw = code >> 8;
            nA[node] = has_p ? pv * RO + pw : node;
            wA[node] = has_p ? delta(pv, pw & 1u) : 0;
            const uint32_t x = sops[node];
            bool mine_w = false;
#pragma unroll
            for (int w = 0; w < W; ++w) mine_w |= swa[vl(s) * W + w] == x && x != kNone;
            if (has_p && mine_w) ssup[pv * W + pw] = 1u;
        }
    wave_fence();
    for (uint32_t round = 0; round < 8; ++round) {
        bool more = false;
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int r = 0; r < RO; ++r) {
                const uint32_t node = nd(s, r);
                const uint32_t t = nA[node];
                const uint32_t tt = nA[t];
                wB[node] = wA[node] + (t != node ? wA[t] : 0);
                nB[node] = tt;
                more |= tt != t;
            }
        wave_fence();
        uint32_t *const tn = nA;
        int32_t *const tw = wA;
        nA = nB;
        wA = wB;
        nB = tn;
        wB = tw;
        if (ballot(more) == 0) break;
    }

    // 5. each written item's request-wide delta and last writer's rank, from its last
    //    (unsuperseded) write: node w reads write slot w's item, so that write's value is the
    //    value before the request + the node's chain sum + the slot's delta
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint32_t sl = sslot[vl(s) * W + w];
            if (sl != kNone && ssup[vl(s) * W + w] == 0) {
                htot[sl] = wA[nd(s, w)] + delta(vl(s), (uint32_t)w);
                hrank[sl] = rank[s];
            }
        }
    wave_fence();

    // 6. the entries: node (v, r) of an item of partition q becomes entry e of q's list,
    //    staged at q * kPipeMaxE + e; the node's place (q << 16 | e, kept in chain buffer B,
    //    which the chains no longer need) for reading its value back.  A node of an item no
    //    one writes is its own root.
    //    Root entries come first in each list, so the committer's atomics fill the fewest
    //    instructions (the token is held while they issue).
    //    A node's entry: the lanes before it with the same partition (three ballots on the
    //    partition's bits) after the entries placed so far (scnt, which the group's last
    //    lane moves on; one wave's LDS operations are performed in order).
    uint32_t *const splace = nB;
    if (lane < kPipeParts) scnt[lane] = 0;
    wave_fence();
#pragma unroll
    for (int pass = 0; pass < 2; ++pass)
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int r = 0; r < RO; ++r) {
                const uint32_t node = nd(s, r);
                const uint32_t q0 = spq[node];
                const uint32_t f = sfs[node];
                const bool root = (f != kNone ? canon[f] : node) == node;
                const bool v = q0 != kNone && root == (pass == 0);
                uint64_t m = ballot(v);
#pragma unroll
                for (uint32_t b = 0; b < 3; ++b) {
                    const bool bit = ((q0 >> b) & 1u) != 0;
                    const uint64_t mb = ballot(v && bit);
                    m &= bit ? mb : ~mb;
                }
                if (v) {
                    const uint32_t e = __hip_atomic_load(&scnt[q0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) +
                                       (uint32_t)__builtin_popcountll(m & lt);
                    if (((m >> lane) >> 1) == 0)
                        __hip_atomic_store(&scnt[q0], e + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    e2[node] = e;
                    splace[node] = (q0 << 16) | e;
                } else if (pass == 0 && q0 == kNone) {
                    e2[node] = 0;
                    splace[node] = kNone;
                }
            }
    wave_fence();
    uint32_t cnt[kPipeParts];
#pragma unroll
    for (uint32_t t = 0; t < kPipeParts; ++t) cnt[t] = scnt[t];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int r = 0; r < RO; ++r) {
            const uint32_t node = nd(s, r);
            const uint32_t pl = splace[node];
            const uint32_t f = sfs[node];
            const bool written = f != kNone;
            const uint32_t at = written ? f : 0u;
            const uint32_t cn = canon[at];
            const uint32_t rt = written ? cn : node;
            const bool is_canon = written && cn == node;
            const bool is_root = rt == node;
            const uint32_t meta = e2[rt < NN ? rt : 0u] | (is_root ? kPipeRootBit : 0u) |
                                  (is_canon ? kPipeInstBit | (hrank[at] << kPipeRankShift) : 0u);
            const int32_t wout = is_canon ? htot[at] : wA[node];
            if (pl != kNone) {
                const uint32_t o = (pl >> 16) * kPipeMaxE + (pl & 0xFFFFu);
                sitem[o] = sops[node];
                swgt[o] = (uint32_t)wout;
                smeta[o] = meta;
            }
        }
    wave_fence();

    // 7. the lists into the boxes a row at a time, the headers, the hash slots back to
    //    empty; send once the box stores are complete (no release fence: it would write
    //    back the XCD's dirty L2 lines)
    // (an opaque scalar base: the stores take it and immediate offsets, instead of per-lane
    // addresses computed once outside the client's loop and kept in scratch)
    uint64_t bxa = bcast64((uint64_t)(uintptr_t)(c.mbox + (size_t)cw * kPipeParts * kPipeBox));
    asm volatile("" : "+s"(bxa));
    uint32_t *bx = (uint32_t *)(uintptr_t)bxa;
    uint64_t chunks = 0;
#pragma unroll
    for (uint32_t q = 0; q < kPipeParts; ++q) {
        chunks |= pipe_ring_chunks(cnt[q], q);
        if (q >= P) continue;
        uint32_t *b = bx + q * kPipeBox;
#pragma unroll
        for (uint32_t g = 0; g < kPipeMaxE / kWave; ++g) {
            const uint32_t x = g * kWave + lane;
            if (x < cnt[q]) {
                st_rlx(&b[kPipeItem + x], sitem[q * kPipeMaxE + x]);
                st_rlx(&b[kPipeWgt + x], swgt[q * kPipeMaxE + x]);
                st_rlx(&b[kPipeMeta + x], smeta[q * kPipeMaxE + x]);
            }
        }
        if (lane == 0) st_rlx(&b[0], cnt[q] | (nc << 16));
    }
    wave_fence();
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint32_t sl = sslot[vl(s) * W + w];
            if (sl != kNone) {
                hkey[sl] = kNone;
#pragma unroll
                for (uint32_t i = 0; i < 2 * S; ++i) hw[i * hsz + sl] = 0;
                canon[sl] = kNone;
                htot[sl] = 0;
            }
        }
    wave_fence();
    bool any = false;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        ts_out[s] = 0;
        any = any || on[s];
    }
    if (ballot(any) == 0) return;
    wait_stores();
    uint64_t n = 0;
    if (lane == 0) {
        n = fetch_add(c.req_tail, (uint64_t)1);
        st_rlx(&c.ring[n & c.qmask], (pipe_ring_tag(n) << kPipeTagShift) | chunks | cw);
    }
    n = bcast64(n);

    // 8. every partition's header word, then the value of each of the lane's nodes
    const uint64_t tw0 = wall_clock64();
    Spin sp;
    const uint64_t *hdr = c.reply + (size_t)cw * kWave;
    uint64_t h0 = 0;
    for (uint32_t pt = 0; pt < P; ++pt) {
        uint64_t w0 = 0;
        for (;;) {
            if (lane == 0) w0 = ld_rlx(&hdr[pt]);
            w0 = bcast64(w0);
            if ((w0 >> 32) == n + 1) break;
            if (!sp.tick(c, kErrReplyTimeout)) { *fail = true; return; }
        }
        h0 = pt == 0 ? w0 : h0;
    }
    const uint32_t *rv = (const uint32_t *)c.reply_v + (size_t)cw * kPipeParts * kPipeMaxE;
    int32_t in[S][RO];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int r = 0; r < RO; ++r) {
            in[s][r] = 0;
            const uint32_t pl = splace[nd(s, r)];
            if (pl != kNone) in[s][r] = (int32_t)ld_rlx(&rv[(pl >> 16) * kPipeMaxE + (pl & 0xFFFFu)]);
        }
    *wait_ticks += wall_clock64() - tw0;

    // 9. the transactions on the values before them, in commit order, and their log
#pragma unroll
    for (int s = 0; s < S; ++s) {
        uint32_t ops[RO], wa[W];
#pragma unroll
        for (int r = 0; r < RO; ++r) ops[r] = sops[nd(s, r)];
#pragma unroll
        for (int w = 0; w < W; ++w) wa[w] = swa[vl(s) * W + w];
        int32_t rd[RO], nv[W];
        transfer_exec<RO, W>(ops, samt[vl(s)], in[s], rd, wa, nv, !ro[s]);
        const uint32_t below = (uint32_t)h0 + rank[s];
        const uint32_t myts = ro[s] ? (kSnapshotBit | below) : below + 1;
        if (on[s]) {
            uint32_t sum = 0;
#pragma unroll
            for (int r = 0; r < RO; ++r) sum += (uint32_t)rd[r];
            c.log_ts[kx[s]] = ro[s] ? below : myts;
            c.log_v[kx[s]] = ro[s] ? (int32_t)sum : rd[0];
#pragma unroll
            for (int r = 1; r < RO; ++r)
                if (!ro[s] && ops[r] != kNone) c.log_v[(uint64_t)r * c.n_txns + kx[s]] = rd[r];
            ts_out[s] = myts;
        }
    }
}

// Global-memory accesses through an address chosen by a select (an active lane's target or
// its sink).  The address-space cast keeps them global instructions: a generic (flat) one
// would also count against the LDS wait counter, so every later LDS wait would wait for it.
typedef __attribute__((address_space(1))) uint64_t g_u64;
typedef __attribute__((address_space(1))) uint32_t g_u32;
__device__ __forceinline__ uint64_t g_fetch_add(uintptr_t a, uint64_t v)
{
    return __hip_atomic_fetch_add((g_u64 *)a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void g_st_rlx64(uintptr_t a, uint64_t v)
{
    __hip_atomic_store((g_u64 *)a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void g_st_rlx32(uintptr_t a, uint32_t v)
{
    __hip_atomic_store((g_u32 *)a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void g_st32(uintptr_t a, uint32_t v) { *(g_u32 *)a = v; }

// Waits until an atomic's word has returned (an empty asm statement that reads it: the
// compiler waits for the load before it, and one wave's memory operations complete in
// issue order, so everything issued before the atomic is complete too).
__device__ __forceinline__ void wait_word(uint64_t v) { asm volatile("" ::"v"(v)); }

// TXBegin on the pipelined lane: the snapshot is the least timestamp every commit partition
// has published its versions through (each partition's progress word, stored once its
// versions and head words are in place).
__device__ __forceinline__ uint32_t pipe_snapshot(const Ctx &c)
{
    uint32_t g = kNone;
    for (uint32_t o = 0; o < c.pipe_parts; ++o) {
        const uint32_t v = ld_acq(&c.pipe_prog[o]);
        g = v < g ? v : g;
    }
    return bcast(g);
}

// The pipelined committers: K waves of a server workgroup, wave k taking requests
// n = k, k + K, ...  With several server workgroups (commit partitions) each one runs every
// request on the entries of its items (pipe_part): a transfer changes each item by its own
// delta, so an item's chain, value, installs and head words never need another partition.
// The committers' only serial step is the issue of a request's atomics, passed on by an LDS
// token: request n's atomics are issued once request n-1's are (the token's workgroup
// release and acquire order the two waves' read-modify-writes on every item they share),
// and the token carries the request's first timestamp.  The rest runs in parallel across
// the waves: wait for the words, give each entry its value (its root's word before the
// request plus the entry's weight), install the last write of each written item, reply
// with the values, and hand the head words to the publisher once the stores are complete.
// Lane i takes entries i, i + 64, ...; idle lanes issue nothing (their instructions are
// masked).
template <int R, int W>
__device__ __forceinline__ void committer_pipe(const Ctx &c, const ServerLayout &L, uint32_t *lds,
                                               uint32_t part, uint32_t k, uint32_t K)
{
    constexpr int G = kPipeMaxE / kWave;
    const uint32_t lane = lane_id();
    const uint32_t D = c.depth;
    const bool dpow2 = (D & (D - 1)) == 0;
    uint32_t *mbst = lds + L.state_off;
    uint32_t *mbwave = lds + L.wave_off;
    uint32_t *mbfree = lds + L.tg_off;                      // [nmb] 1 + request finished in it
    int32_t *jv = (int32_t *)(lds + L.jmp_off + k * kPipeMaxE);   // [kPipeMaxE] words before
    uint32_t *irec = lds + L.pipe_off;                      // [kPipeRing][2][kPipeMaxE]
    uint32_t *icnt = irec + kPipeRing * 2 * kPipeMaxE;      // [kPipeRing] entries
    uint32_t *ilast = icnt + kPipeRing;                     // [kPipeRing]
    uint32_t *iready = ilast + kPipeRing;                   // [kPipeRing]
    uint32_t *iissued = iready + kPipeRing;                 // [kPipeRing] 1 + request
    uint64_t *const tab = c.sv;
    // a word per committer for its completion probe
    uint64_t *const probe_at = c.sink + (size_t)(part * 16 + k) * 16;
    auto mod_d = [&](uint32_t j) -> uint32_t { return dpow2 ? (j & (D - 1)) : j % D; };
    LaneSum t_idle, t_busy, t_tok, t_atom, t_fin, t_room, t_probe;
    bool held = false;                          // a finished request not yet handed over
    uint32_t held_n = 0, held_qi = 0, held_last = 0, held_cw = 0, held_base = 0;
    uint64_t tail_probe = 0;                    // issued after its stores
    // a finished request's stores (versions, reply values) are complete once the probe
    // load issued after them has returned (one wave's memory operations complete in issue
    // order): then its reply header goes out and its installs go to the publisher
    auto hand_over = [&]() {
        wait_word(bcast64(tail_probe));
        if (lane == 0) {
            st_rlx(&c.reply[(size_t)held_cw * kWave + part], ((uint64_t)(held_n + 1) << 32) | held_base);
            ilast[held_qi] = held_last;
            lds_st_rel(&iready[held_qi], held_n + 1);
        }
        held = false;
    };
    uint32_t n = k;
    for (;; n += K) {
        const uint32_t m = n % L.nmb;
        const uint64_t t0 = role_clock();
        // request n staged, or the end of the stream (the finished request is handed over
        // while nothing is staged: its client may be waiting for it)
        bool end = false, bad = false;
        {
            Spin sp;
            for (;;) {
                const uint32_t st = lds_ld_acq(&mbst[m]);
                if ((st & ~kExitBit) == n + 1) {
                    if (st & kExitBit) {
                        end = true;
                        if (lane == 0) lds_st_rel(&lds[kLdsEnd], n + 1);
                    }
                    break;
                }
                if (held) hand_over();
                const uint32_t e = lds_ld_acq(&lds[kLdsEnd]);
                if (e && n + 1 >= e) { end = true; break; }
                if (!sp.tick(c, kErrMailboxTimeout)) { bad = true; break; }
            }
        }
        if (end || bad) break;
        const uint32_t cw = mbwave[m];
        const uint64_t t1 = role_clock();
        const uint32_t *box = lds + L.mb_off + m * L.mb_rows * kWave;
        const uint32_t hdr = box[0];
        const uint32_t cnt = hdr & 0x1FFu, nc = hdr >> 16;
        uint32_t item[G], meta[G];
        int32_t wgt[G];
        bool act[G], at[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const uint32_t e = g * kWave + lane;
            act[g] = e < cnt;
            item[g] = box[kPipeItem + e];
            wgt[g] = (int32_t)box[kPipeWgt + e];
            meta[g] = box[kPipeMeta + e];
        }
        // the mailbox is free again once its rows are in registers
        if (lane == 0) lds_st_rel(&mbfree[m], n + 1);
        // request n's atomics after request n-1's: each root entry adds its item's
        // request-wide delta (and one install when the item is written) and gets the word
        // before the request.  The addresses and addends are ready before the token
        // arrives, and the token carries the request's first timestamp, so the hand-off is
        // one LDS load and one LDS store.
        uint64_t addr[G], add[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            at[g] = act[g] && (meta[g] & kPipeRootBit) != 0;
            addr[g] = (uint64_t)(uintptr_t)(tab + (at[g] ? item[g] : 0u));
            add[g] = ((uint64_t)(uint32_t)wgt[g] << 32) | ((meta[g] & kPipeInstBit) ? 1u : 0u);
        }
        pin_round(addr);
        pin_round(add);
        // otherwise the hand-over waits here, its probe's latency spent on the steps above
        if (held) hand_over();
        const uint64_t tp0 = phase_clock();
        uint32_t base = 0;
        {
            Spin sp;
            for (;;) {
                const uint64_t tok = __hip_atomic_load((uint64_t *)&lds[kLdsTok], __ATOMIC_ACQUIRE,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                if ((uint32_t)tok == n) {
                    base = (uint32_t)(tok >> 32);
                    break;
                }
                if (!sp.tick(c, kErrMailboxTimeout)) { bad = true; break; }
            }
        }
        if (bad) break;
        uint64_t ret[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            ret[g] = 0;
            if (at[g]) ret[g] = g_fetch_add((uintptr_t)addr[g], add[g]);
        }
        if (lane == 0)
            __hip_atomic_store((uint64_t *)&lds[kLdsTok], ((uint64_t)(base + nc) << 32) | (n + 1),
                               __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        const uint64_t tp1 = phase_clock();
        // the words before the request, to every entry of their items
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (at[g]) jv[g * kWave + lane] = (int32_t)(uint32_t)(ret[g] >> 32);
        wave_fence();
        const uint64_t tp2 = phase_clock();
        int32_t val[G];
#pragma unroll
        for (int g = 0; g < G; ++g)
            val[g] = jv[act[g] ? (meta[g] & 0xFFu) : 0u] + (at[g] ? 0 : wgt[g]);
        wave_fence();
        // the install records wait for a free ring entry (the publisher is done with the
        // request kPipeRing before)
        const uint64_t tp3 = phase_clock();
        const uint32_t qi = n % kPipeRing;
        if (n >= kPipeRing) {
            Spin sp;
            while (lds_ld_acq(&lds[kLdsPub]) < n + 1 - kPipeRing)
                if (!sp.tick(c, kErrSlotTimeout)) { bad = true; break; }
        }
        if (bad) break;
        const uint64_t tp3b = phase_clock();
        // an installing entry writes its item's last value (the word before + the
        // request-wide delta) with its last writer's timestamp to the ring slot after the
        // item's installs so far (j-th install: slot j mod depth; the tail moves once the
        // ring is full).  Installs j and j + depth of an item share a slot; they come from
        // requests at least depth apart, made by different committer waves of this
        // workgroup.  With depth >= kPipeRing install j is complete before install j + depth
        // is stored (the wait above); with a shallower ring the installs wait until those of
        // request n - depth are issued: the memory operations this CU's waves issue in turn
        // reach the L2 in that order (as the token relies on for the atomics), so install j
        // lands first.  (Request n - depth + kPipeRing marks the same word only once the
        // publisher is done with request n - depth.)
        if (D < kPipeRing && n >= D) {
            Spin sp;
            while (__hip_atomic_load(&iissued[(n - D) % kPipeRing], __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_WORKGROUP) < n - D + 1)
                if (!sp.tick(c, kErrSlotTimeout)) { bad = true; break; }
        }
        if (bad) break;
        asm volatile("" ::: "memory");
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const uint32_t e = g * kWave + lane;
            const bool inst = at[g] && (meta[g] & kPipeInstBit) != 0;
            const uint32_t j = (uint32_t)ret[g] + 1;
            const uint32_t pos = mod_d(j);
            const uint32_t tail = j < D ? 0u : mod_d(j + 1);
            const uint32_t ts = base + ((meta[g] >> kPipeRankShift) & 255u) + 1;
            if (inst)
                st_rlx(c.slot + (size_t)item[g] * D + pos,
                       ((uint64_t)ts << 32) | (uint32_t)(val[g] + wgt[g]));
            if (act[g]) {
                irec[qi * 2 * kPipeMaxE + e] = inst ? item[g] : kNone;
                irec[(qi * 2 + 1) * kPipeMaxE + e] = pos | (tail << 16);
            }
        }
        // (the mark after the install stores' issue)
        asm volatile("" ::: "memory");
        if (lane == 0) {
            icnt[qi] = cnt;
            __hip_atomic_store(&iissued[qi], n + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        // the reply: each entry's value now; this partition's header word (tag and the
        // request's first timestamp) once the values are complete (hand_over)
        uint32_t *rv = (uint32_t *)c.reply_v + ((size_t)cw * kPipeParts + part) * kPipeMaxE;
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (act[g]) st_rlx(&rv[g * kWave + lane], (uint32_t)val[g]);
        const uint64_t tp4 = phase_clock();
        // the probe after the stores (a load of one word by every lane: one memory request,
        // and no partial update of the register the hand-over waits on)
        asm volatile("" ::: "memory");
        tail_probe = __hip_atomic_load(probe_at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        held = true;
        held_n = n;
        held_qi = qi;
        held_last = base + nc;
        held_cw = cw;
        held_base = base;
        const uint64_t t2 = role_clock();
        t_idle.add(t1 - t0);
        t_busy.add(t2 - t1);
        if constexpr (kPhaseTimers) {
            t_tok.add(tp1 - tp0);
            t_atom.add(tp2 - tp1);
            t_fin.add(tp3 - tp2);
            t_room.add(tp3b - tp3);
            t_probe.add(tp4 - tp3b);
        }
    }
    if (held) hand_over();
    if (lane == 0) {
        fetch_add(&c.prof[kProfCommitIdle], t_idle.v);
        fetch_add(&c.prof[kProfCommitBusy], t_busy.v);
        if constexpr (kPhaseTimers) {
            // phases of the pipelined committers: token wait and atomic issue, the atomics'
            // return, the values, ring room, stores
            fetch_add(&c.prof[kProfCommitRead], t_tok.v);
            fetch_add(&c.prof[kProfCommitResolve], t_atom.v);
            fetch_add(&c.prof[kProfCommitChain], t_fin.v);
            fetch_add(&c.prof[kProfCommitSpin], t_room.v);
            fetch_add(&c.prof[kProfCommitFinish], t_probe.v);
        }
    }
}

// The publisher (one wave of each pipelined server workgroup): takes the requests whose
// installs are complete in request order, as many as are ready, publishes their head words
// (in request order, so an item's head only moves forward; only when normal-lane
// transactions read snapshots during the run), waits for those stores, and publishes the
// partition's progress; a snapshot is the least progress over the
// partitions, so a reader's snapshot only ever covers whole requests whose versions are in
// place everywhere.
template <int R, int W>
__device__ __forceinline__ void publisher_pipe(const Ctx &c, const ServerLayout &L, uint32_t *lds,
                                               uint32_t part)
{
    constexpr int G = kPipeMaxE / kWave;
    const uint32_t lane = lane_id();
    uint32_t *irec = lds + L.pipe_off;
    uint32_t *icnt = irec + kPipeRing * 2 * kPipeMaxE;
    uint32_t *ilast = icnt + kPipeRing;
    uint32_t *iready = ilast + kPipeRing;
    uint64_t *const probe_at = c.sink + (size_t)(part * 16 + 15) * 16;
    const bool heads = c.pipe_heads != 0;
    uint32_t q = 0, gc = 0;
    bool bad = false;
    LaneSum p_batches, p_wait, p_busy;
    for (;;) {
        const uint64_t tq0 = phase_clock();
        // the ready requests from q on (at most half the ring at a time)
        uint32_t qe = q;
        {
            Spin sp;
            for (;;) {
                while (qe < q + kPipeRing / 2 && lds_ld_acq(&iready[qe % kPipeRing]) == qe + 1) ++qe;
                if (qe > q) break;
                const uint32_t e = lds_ld_acq(&lds[kLdsEnd]);
                if (e && q + 1 >= e) break;
                if (!sp.tick(c, kErrSlotTimeout)) { bad = true; break; }
            }
        }
        if (bad || qe == q) break;
        const uint64_t tq1 = phase_clock();
        // (a run without normal-lane transactions has no snapshot reader: its head words are
        // set from the install counts after the run)
        for (uint32_t x = q; heads && x < qe; ++x) {
            const uint32_t qi = x % kPipeRing;
            const uint32_t cn = icnt[qi];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const uint32_t e = g * kWave + lane;
                if (e < cn) {
                    const uint32_t item = irec[qi * 2 * kPipeMaxE + e];
                    const uint32_t nh = irec[(qi * 2 + 1) * kPipeMaxE + e];
                    if (item != kNone) st_rlx(c.ht + item, nh);
                }
            }
        }
        gc = ilast[(qe - 1) % kPipeRing];
        if (heads) {
            // the head words are complete once a probe issued after them has returned
            asm volatile("" ::: "memory");
            uint64_t probe = 0;
            if (lane == 0) probe = g_fetch_add((uintptr_t)probe_at, 0ull);
            wait_word(bcast64(probe));
        }
        if (lane == 0) {
            // this partition's versions through timestamp gc are in place (every version and
            // head word is an agent-scope store, complete by the committers' probes and this
            // wave's); a snapshot is the least of the partitions' progress words
            // (pipe_snapshot)
            st_rlx(&c.pipe_prog[part], gc);
            lds_st_rel(&lds[kLdsPub], qe);
        }
        q = qe;
        if constexpr (kPhaseTimers) {
            p_batches.add(1);
            p_wait.add(tq1 - tq0);
            p_busy.add(phase_clock() - tq1);
        }
    }
    if constexpr (kPhaseTimers)
        if (lane == 0) {
            // the publishers' batches, waiting and busy time (the receivers' counters: the
            // pipelined receivers keep none)
            fetch_add(&c.prof[kProfRecvWindow], p_batches.v);
            fetch_add(&c.prof[kProfRecvRing], p_wait.v);
            fetch_add(&c.prof[kProfRecvBusy], p_busy.v);
        }
    if (lane == 0) {
        lds[kLdsCfinal] = q;
        lds_st_rel(&lds[kLdsCdone], 1u);
        if (part == 0) {
            const uint32_t fin = (uint32_t)(((uint64_t *)&lds[kLdsTok])[0] >> 32);
            fetch_add(&c.stats[kStatBatches], (uint64_t)q);
            st_rlx(&c.stats[kStatFinalTs], (uint64_t)fin);
            st_rlx(c.gc, fin);
        }
    }
}

// A receiver of the pipelined lane: takes requests n = rid, rid + M, ... and copies this
// partition's entry list from the client's box into LDS mailbox n mod nmb: the header and
// the list's 32-entry chunks (their number from the ring word) in one round of loads.  Each
// step polls the ring words of its next two requests, issues the box loads of those that
// have arrived, and stores the lists loaded in the previous step into their mailboxes (once
// a committer has finished each mailbox's previous request) while those loads are in
// flight: up to two requests per memory round trip.
template <int R, int W>
__device__ __forceinline__ void receiver_pipe(const Ctx &c, const ServerLayout &L, uint32_t *lds,
                                              uint32_t part, uint32_t rid, uint32_t M)
{
    constexpr int G = kPipeMaxE / kWave;
    const uint32_t lane = lane_id();
    uint32_t *mbst = lds + L.state_off;
    uint32_t *mbwave = lds + L.wave_off;
    uint32_t *mbfree = lds + L.tg_off;
    // the lists in flight (at most two), as loaded
    uint32_t va[3 * G], vb[3 * G];
    uint32_t ha[1] = {0}, hb[1] = {0};
    uint32_t np = 0;            // requests in flight: pn, pn + M
    uint32_t pn = 0;
    uint64_t pca = 0, pcb = 0;
    LaneSum r_free, r_ring;     // phase timers: waiting for a mailbox, for the ring
    // stores a loaded list into the mailbox of request q
    auto stage = [&](uint32_t q, uint64_t cwr, const uint32_t *h, const uint32_t *v) -> bool {
        const uint32_t mbi = q % L.nmb;
        if (q >= L.nmb) {
            const uint64_t tf = phase_clock();
            Spin sf;
            while (lds_ld_acq(&mbfree[mbi]) < q + 1 - L.nmb)
                if (lds_ld_acq(&lds[kLdsCdone]) || !sf.tick(c, kErrSlotTimeout)) return false;
            if constexpr (kPhaseTimers) r_free.add(phase_clock() - tf);
        }
        uint32_t *dst = lds + L.mb_off + mbi * L.mb_rows * kWave;
        const uint32_t cnt = h[0] & 0x1FFu;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const uint32_t x = g * kWave + lane;
            if (x < cnt)
#pragma unroll
                for (int f = 0; f < 3; ++f) dst[kPipeItem + f * kPipeMaxE + x] = v[f * G + g];
        }
        if (lane == 0) {
            dst[0] = h[0];
            mbwave[mbi] = (uint32_t)cwr & kPipeCwMask;
        }
        wg_fence();
        if (lane == 0) lds_st_rel(&mbst[mbi], q + 1);
        return true;
    };
    // issues the box loads of a request whose ring word has arrived (their address from
    // that word, so issued after it; agent-scope loads, which read the device's point of
    // coherence: no acquire, whose invalidation would take this XCD's L2 lines, the
    // committers' value words among them)
    auto fetch = [&](uint64_t cwr, uint32_t *h, uint32_t *v) {
        const uint32_t C = pipe_ring_entries(cwr, part);
        const uint32_t *ps = c.mbox + ((size_t)((uint32_t)cwr & kPipeCwMask) * kPipeParts + part) * kPipeBox;
        h[0] = ld_rlx(&ps[0]);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const uint32_t x = g * kWave + lane;
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                v[f * G + g] = 0;
                if (x < C) v[f * G + g] = ld_rlx(&ps[kPipeItem + f * kPipeMaxE + x]);
            }
        }
    };
    uint32_t n = rid;           // the next request to fetch
    Spin sp;
    for (uint32_t polls = 0;; ++polls) {
        const uint64_t tr = phase_clock();
        const uint64_t e0 = ld_rlx(&c.ring[n & c.qmask]);
        const uint64_t e1 = ld_rlx(&c.ring[(n + M) & c.qmask]);
        if (np >= 1 && !stage(pn, pca, ha, va)) break;
        if (np == 2 && !stage(pn + M, pcb, hb, vb)) break;
        np = 0;
        const uint64_t a = bcast64(e0), b = bcast64(e1);   // one word each: wave-wide branches
        if ((a >> kPipeTagShift) == pipe_ring_tag(n)) {
            pn = n;
            pca = a;
            fetch(pca, ha, va);
            np = 1;
            if ((b >> kPipeTagShift) == pipe_ring_tag(n + M)) {
                pcb = b;
                fetch(pcb, hb, vb);
                np = 2;
            }
            n += np * M;
            polls = 0;
            sp = Spin();
            continue;
        }
        // (the acquire only once every client has left: an acquire invalidates this XCD's
        // L2 lines, the committers' value words among them)
        if ((polls & 7u) == 7u && ld_rlx(c.clients_done) == c.n_client_waves) {
            fence_acq();
            if (ld_acq(c.req_tail) <= n) {
                // every client has left and request n will never come
                if (ld_rlx(c.req_tail) == n && lane == 0) {
                    const uint32_t mbi = n % L.nmb;
                    bool ok = true;
                    if (n >= L.nmb) {
                        Spin sf;
                        while (lds_ld_acq(&mbfree[mbi]) < n + 1 - L.nmb)
                            if (lds_ld_acq(&lds[kLdsCdone]) || !sf.tick(c, kErrSlotTimeout)) {
                                ok = false;
                                break;
                            }
                    }
                    if (ok) lds_st_rel(&mbst[mbi], kExitBit | (n + 1));
                }
                break;
            }
        }
        if constexpr (kPhaseTimers) r_ring.add(phase_clock() - tr);
        if (!sp.tick(c, kErrRingTimeout)) break;
    }
    if constexpr (kPhaseTimers)
        if (lane == 0) {
            // the pipelined receivers' waits for a free mailbox and for the ring (slots the
            // pipelined committers leave unused)
            fetch_add(&c.prof[kProfCommitMiss], r_free.v);
            fetch_add(&c.prof[kProfCommitAdv], r_ring.v);
        }
}

// The server workgroup of commit partition `part` of the pipelined lane: K committers
// (pipe_committers), one publisher (it moves the clock: no advancer), the other waves
// receive.
template <int R, int W>
__device__ __forceinline__ void pipe_server_workgroup(const Ctx &c, const ServerLayout &L, uint32_t *lds,
                                                      uint32_t part)
{
    const uint32_t wave = threadIdx.x / kWave;
    for (uint32_t i = threadIdx.x; i < L.mb_off; i += blockDim.x) lds[i] = 0;
    // the install ring too: its ready words must not match a request of this run
    for (uint32_t i = threadIdx.x; i < pipe_words(); i += blockDim.x) lds[L.pipe_off + i] = 0;
    __syncthreads();
    const uint32_t K = pipe_committers(c.wg_waves);
    if (wave < K) committer_pipe<R, W>(c, L, lds, part, wave, K);
    else if (wave == K) publisher_pipe<R, W>(c, L, lds, part);
    else receiver_pipe<R, W>(c, L, lds, part, wave - K - 1, c.wg_waves - K - 1);
}

// D: the serialized lane (delegated requests) is compiled in.
template <int R, int W, bool D>
__device__ __forceinline__ void server_workgroup(const Ctx &c, const ServerLayout &L, uint32_t *lds)
{
    const uint32_t wave = threadIdx.x / kWave;
    for (uint32_t i = threadIdx.x; i < L.mb_off; i += blockDim.x) lds[i] = 0;
    if ((c.variant & 0x3u) == kValStamp) {
        for (uint32_t i = threadIdx.x; i < L.nbuckets; i += blockDim.x) lds[L.bucket_off + i] = 0;
        if (L.exact)
            for (uint32_t i = threadIdx.x; i < c.n_items; i += blockDim.x)
                lds[L.val_off + i] = (uint32_t)c.sv[i];
    }
    const uint32_t M = c.wg_waves - 2;
    for (uint32_t r = 0; r < M; ++r)
        for (uint32_t i = threadIdx.x; i < L.hsz; i += blockDim.x) {
            uint32_t *h = lds + L.hash_off + r * L.harr * L.hsz;
            h[i] = kNone;
            for (uint32_t q = 1; q < L.harr; ++q) h[q * L.hsz + i] = 0;
        }
    __syncthreads();
    // Waves are spread over the CU's four SIMDs in turn, so wave 4 shares the committer's
    // SIMD: it runs the advancer (mostly asleep in its polls) rather than a receiver,
    // whose hash and chain work would take issue slots from the committer.
    const uint32_t adv = c.wg_waves >= 5 ? 4u : 1u;
    if (wave == 0) {
        if (c.variant == (kValStamp | kInsertBatch)) committer<R, W, D, true>(c, L, lds);
        else committer<R, W, D>(c, L, lds);
    }
    else if (wave == adv) {
        if ((c.variant & kServerWriteback) == 0) advancer(c, lds);
    } else receiver<R, W, D>(c, L, lds, wave - 1 - (wave > adv ? 1u : 0u), M);
    // exact mode: the committed values go back to the HBM words (stamp 0, as a run's
    // stamps restart), so a later run without a new fill starts from this run's state, as
    // in hashed mode, where the words are the table
    if ((c.variant & 0x3u) == kValStamp && L.exact) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < c.n_items; i += blockDim.x) c.sv[i] = lds[L.val_off + i];
    }
}

}  // namespace csmv
