// serve.hpp -- the persistent serving loop's client side: arrival-gated pick-up,
// admission by in-flight version-pool bytes, latency from arrival to commit, and the
// wide bank transaction (2..64 reads) with one batched snapshot read.
//
// The reference serves memcached requests batch by batch from a host loop, one kernel
// launch per batch against a persistent server (memcachedGPU/memcdSH.cu:603-651).  Here
// requests wait in device-resident queues (the resident streams in arrival order, each
// with its arrival tick) and are admitted lane by lane: a lane that commits takes the
// next arrived request in the same round (continuous batching), so requests of
// different ages and sizes share a wavefront and overlap on the device.
#pragma once

#include "csmv_device.hpp"
#include "csmv_server.hpp"

namespace csmv {

// Version-pool bytes a request may install: (version, value) + head word per item.
constexpr uint32_t kVersionBytes = 12;

// The serving start: the first client wave to run publishes its clock; arrival ticks
// count from there.
__device__ __forceinline__ uint64_t serve_start(const Ctx &c)
{
    uint64_t t = 0;
    if (lane_id() == 0) {
        const uint64_t now = wall_clock64();
        uint64_t expected = 0;
        t = __hip_atomic_compare_exchange_strong(c.t_start, &expected, now, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                ? now
                : expected;
    }
    return bcast64(t);
}

// Admission.  Lanes holding an arrived request ask for `need` bytes of in-flight
// versions.  One fetch_add claims the bytes of every such lane; the longest lane-order
// prefix of them that fits under the budget above the counter's previous value is
// admitted and the rest is returned at once.  The admitted bytes never exceed the budget
// (the counter can pass it only by bytes about to be returned, which makes concurrent
// waves refuse, never admit, too much).  A round that admits nothing counts as a stall;
// the lanes ask again next round.
__device__ __forceinline__ void serve_gate(const Ctx &c, uint64_t t0, bool has, bool &adm,
                                           uint64_t k, uint32_t need, uint32_t &held,
                                           uint64_t &stalls)
{
    const uint64_t now = wall_clock64() - t0;
    const bool cand = has && !adm && (uint64_t)c.arr[k] <= now;
    if (ballot(cand) == 0) return;
    if (c.budget == 0) {
        if (cand) { adm = true; held = 0; }
        return;
    }
    const uint32_t incl = wave_incl_scan(cand ? need : 0u);
    const uint32_t want = readlane(incl, 63);
    uint64_t old = 0;
    if (lane_id() == 0) old = fetch_add(c.inflight, (uint64_t)want);
    old = bcast64(old);
    const uint64_t room = c.budget > old ? c.budget - old : 0;
    const uint64_t fit = ballot(cand && incl <= room);
    const uint32_t take = fit ? readlane(incl, 63u - (uint32_t)__builtin_clzll(fit)) : 0u;
    if (lane_id() == 0) {
        if (want > take) fetch_add(c.inflight, (uint64_t)0 - (uint64_t)(want - take));
        if (take) atomicMax((unsigned long long *)&c.serve_ctr[1], (unsigned long long)(old + take));
    }
    if (cand && incl <= take) {
        adm = true;
        held = need;
    }
    if (take == 0) ++stalls;
}

// Returns the bytes of lanes that committed this round to the budget.
__device__ __forceinline__ void serve_release(const Ctx &c, bool done, uint32_t &held)
{
    if (c.budget == 0) return;
    const uint32_t s = wave_sum(done ? held : 0u);
    if (done) held = 0;
    if (s && lane_id() == 0) __hip_atomic_fetch_add(c.inflight, (uint64_t)0 - s, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
}

// Snapshot reads of up to N items (kNone entries are skipped): all head words, then all
// head slots, as two rounds of independent loads; a head newer than the snapshot falls
// back to the ring walk below the head (pool_read_below).
template <int N>
__device__ __forceinline__ bool pool_read_chunk(const Ctx &c, const uint32_t (&it)[N], uint32_t snap,
                                                int32_t (&out)[N])
{
    // every load of a round is issued (an empty entry reads item 0) and the round is pinned
    // (pin_round): otherwise the compiler waits for each load before issuing the next.
    // An installer writes the slot, then (after a release fence) the head word; the acquire
    // fence between the two rounds makes a head seen here order the slot it names, so a
    // new head never pairs with the slot's evicted contents.  (On gfx950 it is one
    // buffer_inv sc1 per round; the bank headline measured the same with and without it.)
    uint32_t ht[N];
    uint64_t x[N];
#pragma unroll
    for (int j = 0; j < N; ++j) ht[j] = ld_rlx(&c.ht[it[j] != kNone ? it[j] : 0u]);
    pin_round(ht);
    fence_acq();
#pragma unroll
    for (int j = 0; j < N; ++j)
        x[j] = ld_rlx(&c.slot[(size_t)(it[j] != kNone ? it[j] : 0u) * c.depth + (ht[j] & 0xFFFFu)]);
    pin_round(x);
    bool ok = true;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        int32_t v = (int32_t)(uint32_t)x[j];
        if (it[j] != kNone && (uint32_t)(x[j] >> 32) > snap) ok = pool_read_below(c, it[j], snap, v) && ok;
        out[j] = v;
    }
    return ok;
}

// Snapshot reads of up to N items: all head words, then all head slots, as two rounds of
// independent loads.  An item in walk (bit j) whose head is newer than the snapshot is
// walked back below the head (ok turns false when no version is old enough); any other
// such item is returned (bit j), its value in out being the head's, not the snapshot's.
template <int N>
__device__ __forceinline__ uint32_t pool_read_some(const Ctx &c, const uint32_t (&it)[N], uint32_t snap,
                                                   int32_t (&out)[N], uint32_t walk, bool &ok)
{
    uint32_t ht[N];
    uint64_t x[N];
#pragma unroll
    for (int j = 0; j < N; ++j) ht[j] = ld_rlx(&c.ht[it[j] != kNone ? it[j] : 0u]);
    pin_round(ht);
    fence_acq();   // as in pool_read_chunk
#pragma unroll
    for (int j = 0; j < N; ++j)
        x[j] = ld_rlx(&c.slot[(size_t)(it[j] != kNone ? it[j] : 0u) * c.depth + (ht[j] & 0xFFFFu)]);
    pin_round(x);
    uint32_t newer = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        int32_t v = (int32_t)(uint32_t)x[j];
        if (it[j] != kNone && (uint32_t)(x[j] >> 32) > snap) {
            if ((walk >> j) & 1u) ok = pool_read_below(c, it[j], snap, v) && ok;
            else newer |= 1u << j;
        }
        out[j] = v;
    }
    return newer;
}

// A serving-loop bank transaction as the serialized lane carries it: the transfer pair as
// the write set, the stream index and amount in the write-value rows; the read set stays
// in the stream (tx_commit copies its rows into the mailbox).
template <int R, int W>
__device__ __forceinline__ void serve_delegated_tx(const Ctx &c, Tx<R, W> &tx, bool on, uint64_t k,
                                                   uint32_t op)
{
    tx.snap = on ? 0u : kNone;
    tx.nw = 0;
    tx.aborted = false;
    tx.overflow = false;
#pragma unroll
    for (int w = 0; w < W; ++w) { tx.wa[w] = kNone; tx.wv[w] = 0; }
    tx.nr = on ? min((op >> 16) & 0xFFu, (uint32_t)R) : 0u;
    if (on) {
        tx_write(tx, c.addr[k], 0);
        tx_write(tx, c.addr[c.n_txns + k], 0);
        tx.wv[0] = (int32_t)(uint32_t)k;
        tx.wv[1] = (int32_t)((op >> 8) & 0xFFu);
    }
}

// Serving-loop bank transaction: a transfer over the first two of nr accounts
// (2 <= nr <= R) that also reads the other nr - 2 (bank/csmv_sh/exe.cu:184-201 widened
// to the serving mix's read sets).  Same commit path and log as the bank driver, plus
// the latency from arrival to commit.
template <int R, int W> __device__ __forceinline__ void bank_serve_client(const Ctx &c, uint32_t cw)
{
    const uint32_t lane = lane_id();
    const uint64_t T = c.n_txns;
    const uint64_t NT = c.n_norm;
    const uint64_t t0 = serve_start(c);
    bool has = false, adm = false, exhausted = false;
    uint64_t k = 0;
    uint32_t op = 0, ab = 0, my_last = 0, held = 0;
    uint64_t n_upd = 0, n_age = 0, n_esc = 0, stalls = 0, t_reply = 0, t_round = 0, rounds = 0;
    for (;;) {
        if (ld_rlx(c.err)) break;
        const uint64_t tr0 = wall_clock64();
        if (!exhausted) {
            const uint64_t need = ballot(!has);
            if (need) {
                const uint32_t cnt = (uint32_t)__builtin_popcountll(need);
                uint64_t base = 0;
                if (lane == 0) base = fetch_add(c.next_txn, (uint64_t)cnt);
                base = bcast64(base);
                if (base + cnt >= NT) exhausted = true;
                if (!has) {
                    const uint64_t mine = base + prefix_count(need);
                    if (mine < NT) {
                        has = true;
                        adm = false;
                        k = c.norm_list ? c.norm_list[mine] : mine;
                        op = c.op[k];
                        ab = 0;
                    }
                }
            }
        }
        if (ballot(has) == 0) break;
        serve_gate(c, t0, has, adm, k, W * kVersionBytes, held, stalls);
        const bool act = has && adm;
        if (ballot(act) == 0) {
            __builtin_amdgcn_s_sleep(2);
            continue;
        }
        if (c.escalate) {
            // forward progress: transactions that have failed c.escalate attempts go to the
            // serialized lane in this round's request; the other lanes sit it out
            const bool esc = act && ab >= c.escalate;
            if (ballot(esc)) {
                Tx<R, W> tx;
                serve_delegated_tx<R, W>(c, tx, esc, k, op);
                bool fail = false;
                const uint32_t ts = tx_commit<R, W, W, true>(c, cw, tx, esc, &fail, &t_reply, true,
                                                             nullptr, nullptr, -1, c.addr + k, T);
                if (fail) break;
                const uint64_t cmu = ballot(ts != 0);
                if (cmu) my_last = readlane(ts, 63u - (uint32_t)__builtin_clzll(cmu));
                bool done = false;
                if (esc) {
                    if (ts) {
                        c.log_ab[k] = ab;
                        c.log_lat[k] = (uint32_t)(wall_clock64() - t0 - c.arr[k]);
                        done = true;
                        has = false;
                        ++n_upd;
                        ++n_esc;
                    } else {
                        ++ab;
                    }
                }
                serve_release(c, done, held);
                t_round += wall_clock64() - tr0;
                ++rounds;
                continue;
            }
        }

        Tx<R, W> tx;
        if (!tx_begin(c, tx, c.begin_wait ? my_last : 0u)) break;
        const uint32_t amt = (op >> 8) & 0xFFu;
        const uint32_t nr = act ? min((op >> 16) & 0xFFu, (uint32_t)R) : 0u;
        // program order over snapshot values, 16 reads at a time: read a, write a - amt,
        // read b (own write if b == a), write b + amt, then the other reads (own writes
        // first).  The read set stays in the stream (the mailbox rows are copied from it
        // at commit) and each read value goes to the log at once: only the committed
        // attempt's rows are kept.
        bool ok = true;
        constexpr int kChunk = 16;
#pragma unroll 1
        for (int c0 = 0; c0 < R; c0 += kChunk) {
            if ((uint32_t)c0 < nr) {
                uint32_t it[kChunk];
                int32_t v[kChunk];
#pragma unroll
                for (int j = 0; j < kChunk; ++j)
                    it[j] = (uint32_t)(c0 + j) < nr ? c.addr[(uint64_t)(c0 + j) * T + k] : kNone;
                ok = pool_read_chunk<kChunk>(c, it, tx.snap, v) && ok;
#pragma unroll
                for (int j = 0; j < kChunk; ++j) {
                    if (it[j] == kNone) continue;
                    int32_t x = v[j];
#pragma unroll
                    for (int w = 0; w < W; ++w)
                        x = ((uint32_t)w < tx.nw && tx.wa[w] == it[j]) ? tx.wv[w] : x;
                    if (c0 + j == 0) tx_write(tx, it[j], (int32_t)((uint32_t)x - amt));
                    else if (c0 + j == 1) tx_write(tx, it[j], (int32_t)((uint32_t)x + amt));
                    c.log_v[(uint64_t)(c0 + j) * T + k] = x;
                }
            }
        }
        tx.nr = nr;
        tx.aborted = tx.aborted || !ok;
        if (tx.overflow) raise_error(c, kErrReadSetOverflow);
        if (act && tx.aborted) { ++ab; ++n_age; }
        const bool req = act && !tx.aborted;
        bool done = false;
        if (ballot(req)) {
            bool fail = false;
            const uint32_t ts = tx_commit<R, W, W, true>(c, cw, tx, req, &fail, &t_reply, false,
                                                         nullptr, nullptr, -1, c.addr + k, T);
            if (fail) break;
            const uint64_t cm = ballot(ts != 0);
            if (cm) my_last = readlane(ts, 63u - (uint32_t)__builtin_clzll(cm));
            if (req) {
                if (ts) {
                    c.log_ts[k] = ts;
                    c.log_ab[k] = ab;
                    c.log_lat[k] = (uint32_t)(wall_clock64() - t0 - c.arr[k]);
                    done = true;
                    has = false;
                    ++n_upd;
                } else {
                    ++ab;
                }
            }
        }
        serve_release(c, done, held);
        if (has && ab > c.max_attempts) raise_error(c, kErrLivelock);
        t_round += wall_clock64() - tr0;
        ++rounds;
    }
    if (lane == 0) {
        fetch_add(&c.prof[kProfClientReply], t_reply);
        fetch_add(&c.prof[kProfClientRound], t_round);
        fetch_add(&c.prof[kProfClientRounds], rounds);
        if (stalls) fetch_add(&c.serve_ctr[0], stalls);
    }
    if (n_upd) fetch_add(&c.stats[kStatCommitUpd], n_upd);
    if (n_age) fetch_add(&c.stats[kStatAbortAge], n_age);
    if (n_esc) fetch_add(&c.stats[kStatEscalated], n_esc);
    fence_rel();
    if (lane == 0) __hip_atomic_fetch_add(c.clients_done, 1u, __ATOMIC_RELEASE,
                                          __HIP_MEMORY_SCOPE_AGENT);
}

// Serving-loop serialized lane for routed bank transactions: the wave hands them to the
// commit server, which executes them in commit order on the newest committed values (the
// transfer pair through its producer chains, the other reads from their producer or the
// committed value) and logs them; the wave installs the returned versions in commit order.
template <int R, int W> __device__ __forceinline__ void bank_serve_serial_client(const Ctx &c, uint32_t cw)
{
    const uint32_t lane = lane_id();
    const uint64_t T = c.n_txns;
    const uint64_t NH = c.n_hot;
    const uint64_t t0 = serve_start(c);
    bool has = false, adm = false, exhausted = false;
    uint64_t k = 0;
    uint32_t op = 0, ab = 0, held = 0;
    uint64_t n_upd = 0, stalls = 0, t_reply = 0, t_round = 0, rounds = 0;
    for (;;) {
        if (ld_rlx(c.err)) break;
        const uint64_t tr0 = wall_clock64();
        if (!exhausted) {
            const uint64_t need = ballot(!has);
            if (need) {
                const uint32_t cnt = (uint32_t)__builtin_popcountll(need);
                uint64_t base = 0;
                if (lane == 0) base = fetch_add(c.next_hot, (uint64_t)cnt);
                base = bcast64(base);
                if (base + cnt >= NH) exhausted = true;
                if (!has) {
                    const uint64_t mine = base + prefix_count(need);
                    if (mine < NH) {
                        has = true;
                        adm = false;
                        k = c.hot_list[mine];
                        op = c.op[k];
                        ab = 0;
                    }
                }
            }
        }
        if (ballot(has) == 0) break;
        serve_gate(c, t0, has, adm, k, W * kVersionBytes, held, stalls);
        const bool act = has && adm;
        if (ballot(act) == 0) {
            __builtin_amdgcn_s_sleep(2);
            continue;
        }
        Tx<R, W> tx;
        serve_delegated_tx<R, W>(c, tx, act, k, op);
        bool fail = false;
        const uint32_t ts = tx_commit<R, W, W, true>(c, cw, tx, act, &fail, &t_reply, true, nullptr,
                                                     nullptr, -1, c.addr + k, T);
        if (fail) break;
        bool done = false;
        if (act) {
            if (ts) {
                c.log_ab[k] = ab;
                c.log_lat[k] = (uint32_t)(wall_clock64() - t0 - c.arr[k]);
                done = true;
                has = false;
                ++n_upd;
            } else {
                ++ab;
            }
        }
        serve_release(c, done, held);
        if (has && ab > c.max_attempts) raise_error(c, kErrLivelock);
        t_round += wall_clock64() - tr0;
        ++rounds;
    }
    if (lane == 0) {
        fetch_add(&c.prof[kProfClientReply], t_reply);
        fetch_add(&c.prof[kProfClientRound], t_round);
        fetch_add(&c.prof[kProfClientRounds], rounds);
        if (stalls) fetch_add(&c.serve_ctr[0], stalls);
    }
    if (n_upd) fetch_add(&c.stats[kStatCommitUpd], n_upd);
    fence_rel();
    if (lane == 0) __hip_atomic_fetch_add(c.clients_done, 1u, __ATOMIC_RELEASE,
                                          __HIP_MEMORY_SCOPE_AGENT);
}

// ---- the serving bank class on the pipelined serialized lane --------------------------
// A serving bank transaction reads 2..R accounts and transfers over the first two.  On the
// pipelined lane (the bank engine's variant carries kSerialPipe) its nodes are the chain
// nodes of the transfer pair (as in tx_commit_pipe) plus one leaf per other read: a leaf
// takes the value its producer (the highest lower lane writing the account) wrote, or the
// value before the request, which its root entry's atomic returns; a leaf reading the
// transaction's own write takes that write on the client and sends nothing.  A request
// carries the lanes, in lane order, whose nodes keep every partition's list within
// kPipeMaxE; the other lanes wait for the next request.
//
// A client wave's LDS scratch (pipe_wide_words): the write hash (key, writer lanes lo / hi,
// canonical node, request-wide delta, last writer's rank), per chain node (2 per lane) its
// root, weight, own write delta, superseded flag and entry, the staged lists (a block of
// kPipeMaxE per partition: items, weights, meta words), the read sets (R rows of 64) and
// each read's (partition << 16 | entry).
__host__ __device__ constexpr uint32_t pipe_wide_words(uint32_t hsz, uint32_t R)
{
    return 6 * hsz + 5 * 2 * kWave + 3 * kPipeParts * kPipeMaxE + 2 * R * kWave;
}

template <int R, int W>
__device__ __forceinline__ uint32_t tx_commit_pipe_wide(const Ctx &c, const ServerLayout &L,
                                                        uint32_t *scr, uint32_t cw, bool req,
                                                        uint64_t k, uint32_t op, bool *fail,
                                                        uint64_t *wait_ticks)
{
    static_assert(W == 2, "the pipelined lane carries transfers");
    constexpr int CN = 2;   // chain nodes per lane: the transfer pair
    const uint32_t lane = lane_id();
    const uint64_t lt = (1ull << lane) - 1;
    const uint64_t T = c.n_txns;
    const uint32_t hsz = L.hsz, hmask = hsz - 1, P = L.parts;
    uint32_t *hkey = scr, *hwlo = scr + hsz, *hwhi = scr + 2 * hsz, *canon = scr + 3 * hsz;
    int32_t *htot = (int32_t *)(scr + 4 * hsz);
    uint32_t *hrank = scr + 5 * hsz;
    uint32_t *snext = scr + 6 * hsz;                    // [64 * CN] root node
    int32_t *sw = (int32_t *)(snext + kWave * CN);      // [64 * CN] weight from the root
    int32_t *sdel = sw + kWave * CN;                    // [64 * CN] the node's own write delta
    uint32_t *ssup = (uint32_t *)(sdel + kWave * CN);   // [64 * CN] superseded flags
    uint32_t *e2 = ssup + kWave * CN;                   // [64 * CN] entry of each chain node
    uint32_t *sitem = e2 + kWave * CN;                  // [kPipeParts][kPipeMaxE]
    uint32_t *swgt = sitem + kPipeParts * kPipeMaxE;
    uint32_t *smeta = swgt + kPipeParts * kPipeMaxE;
    uint32_t *rsx = smeta + kPipeParts * kPipeMaxE;     // [R][64] read sets
    uint32_t *le = rsx + R * kWave;                     // [R][64] (partition << 16 | entry)
    const uint32_t mybit = 1u << (lane & 31);

    // 1. the read sets into LDS (16 rows a round; a row past a lane's reads loads its last
    //    read again, and a lane without a transaction row 0 of transaction 0, so no load
    //    leaves the stream; both keep kNone)
    const uint32_t nr = req ? min((op >> 16) & 0xFFu, (uint32_t)R) : 0u;
    const uint32_t amt = (op >> 8) & 0xFFu;
    const uint64_t kk = req ? k : 0;
#pragma unroll 1
    for (uint32_t j0 = 0; j0 < (uint32_t)R; j0 += 16) {
        if (ballot(j0 < nr) == 0) break;
        uint32_t v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = c.addr[(uint64_t)min(j0 + j, nr ? nr - 1 : 0u) * T + kk];
        pin_round(v);
#pragma unroll
        for (int j = 0; j < 16; ++j) rsx[(j0 + j) * kWave + lane] = j0 + j < nr ? v[j] : kNone;
    }
    wave_fence();
    const uint32_t a0 = nr ? rsx[lane] : kNone, a1 = nr > 1 ? rsx[kWave + lane] : kNone;

    // 2. the lanes this request carries: a lane-order prefix whose nodes keep every
    //    partition's list within kPipeMaxE (a leaf on an own write is no node)
    uint32_t cq[kPipeParts];
#pragma unroll
    for (uint32_t t = 0; t < kPipeParts; ++t) cq[t] = 0;
    for (uint32_t j = 0; j < nr; ++j) {
        const uint32_t x = rsx[j * kWave + lane];
        if (j >= CN && (x == a0 || x == a1)) continue;
        const uint32_t q = pipe_part(x, P);
#pragma unroll
        for (uint32_t t = 0; t < kPipeParts; ++t) cq[t] += q == t ? 1u : 0u;
    }
    // (every lane takes part in the scans: a lane-prefix scan reads the lanes before it)
    bool on = req;
#pragma unroll
    for (uint32_t t = 0; t < kPipeParts; ++t) {
        if (t >= P) break;
        const uint32_t pre = wave_incl_scan(cq[t]);
        on = on && pre <= kPipeMaxE;
    }

    // 3. the write set into the hash (item -> writer lanes)
    uint32_t wa[W];
    wa[0] = on ? a0 : kNone;
    wa[1] = (on && a1 != a0) ? a1 : kNone;
    const int32_t d0 = (wa[1] == kNone) ? 0 : -(int32_t)amt, d1 = (int32_t)amt;
    uint32_t slot_w[W], hp[W];
    bool pend[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
        slot_w[w] = kNone;
        hp[w] = mix32(wa[w]) & hmask;
        pend[w] = wa[w] != kNone;
    }
    for (uint32_t step = 0; step <= hmask; ++step) {
        uint32_t old[W];
#pragma unroll
        for (int w = 0; w < W; ++w) {
            old[w] = kNone;
            __hip_atomic_compare_exchange_strong(&hkey[pend[w] ? hp[w] : 0u], &old[w],
                                                 pend[w] ? wa[w] : kNone, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        bool more = false;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const bool got = pend[w] && (old[w] == kNone || old[w] == wa[w]);
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
    wave_fence();

    // looks an account up in the write hash: its slot, or kNone when no lane writes it
    auto find = [&](uint32_t x) -> uint32_t {
        uint32_t hs = mix32(x) & hmask, fs = kNone;
        bool fp = x != kNone;
        for (uint32_t step = 0; step <= hmask; ++step) {
            const uint32_t kx = __hip_atomic_load(&hkey[fp ? hs : 0u], __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_WORKGROUP);
            const bool hit = fp && kx == x;
            const bool next = fp && kx != x && kx != kNone;
            fs = hit ? hs : fs;
            hs = next ? ((hs + 1) & hmask) : hs;
            fp = next;
            if (ballot(next) == 0) break;
        }
        return fs;
    };
    // the producer of a read of x (hash slot fs): the highest lower lane writing x and the
    // write slot (lane | slot << 6), 0xFF = none
    auto producer = [&](uint32_t x, uint32_t fs) -> uint32_t {
        const uint32_t at = fs != kNone ? fs : 0u;
        const uint64_t m = ((uint64_t)__hip_atomic_load(&hwhi[at], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) << 32) |
                           __hip_atomic_load(&hwlo[at], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        const uint64_t lower = fs != kNone ? (m & lt) : 0ull;
        const uint32_t pl = lower ? 63u - (uint32_t)__builtin_clzll(lower) : 0xFFu;
        const uint32_t srcl = pl == 0xFFu ? lane : pl;
        uint32_t code = 0xFFu;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint32_t pw = lane_gather(wa[w], srcl);
            code = (pl != 0xFFu && pw == x) ? (pl | ((uint32_t)w << 6)) : code;
        }
        return code;
    };

    // 4. the chain nodes (the transfer pair): producers, canonical nodes, chain sums by
    //    pointer jumping, superseded writes, each written item's request-wide delta and
    //    last writer's rank (as tx_commit_pipe, over two nodes a lane)
    uint32_t ops[CN], fs[CN], pj[CN];
    ops[0] = on ? a0 : kNone;
    ops[1] = on ? a1 : kNone;
#pragma unroll
    for (int r = 0; r < CN; ++r) {
        fs[r] = find(ops[r]);
        pj[r] = producer(ops[r], fs[r]);
    }
#pragma unroll
    for (int r = 0; r < CN; ++r)
        if (fs[r] != kNone && pj[r] == 0xFFu) atomicMin(&canon[fs[r]], lane * CN + (uint32_t)r);
    uint32_t nxt[CN];
    int32_t wgt[CN];
#pragma unroll
    for (int r = 0; r < CN; ++r) {
        const uint32_t node = lane * CN + r;
        const bool has_p = pj[r] != 0xFFu;
        const uint32_t srcl = has_p ? (pj[r] & 63u) : lane;
        const int32_t q0 = (int32_t)lane_gather((uint32_t)d0, srcl),
                      q1 = (int32_t)lane_gather((uint32_t)d1, srcl);
        nxt[r] = has_p ? (pj[r] & 63u) * CN + (pj[r] >> 6) : node;
        wgt[r] = has_p ? ((pj[r] >> 6) == 0 ? q0 : q1) : 0;
        snext[node] = nxt[r];
        sw[node] = wgt[r];
        ssup[node] = 0;
    }
    wave_fence();
    for (uint32_t round = 0; round < 8; ++round) {
        bool more = false;
        uint32_t nn[CN];
        int32_t ww[CN];
#pragma unroll
        for (int r = 0; r < CN; ++r) {
            const uint32_t node = lane * CN + r;
            const uint32_t t = nxt[r];
            const uint32_t tt = snext[t];
            ww[r] = wgt[r] + (t != node ? sw[t] : 0);
            nn[r] = tt;
            more |= tt != t;
        }
        wave_fence();
#pragma unroll
        for (int r = 0; r < CN; ++r) {
            nxt[r] = nn[r];
            wgt[r] = ww[r];
            snext[lane * CN + r] = nn[r];
            sw[lane * CN + r] = ww[r];
        }
        wave_fence();
        if (ballot(more) == 0) break;
    }
#pragma unroll
    for (int r = 0; r < CN; ++r) {
        bool mine_w = false;
#pragma unroll
        for (int w = 0; w < W; ++w) mine_w |= wa[w] == ops[r] && ops[r] != kNone;
        if (pj[r] != 0xFFu && mine_w) ssup[(pj[r] & 63u) * CN + (pj[r] >> 6)] = 1u;
    }
    wave_fence();
    uint32_t supf = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) supf |= ssup[lane * CN + w] ? (1u << w) : 0u;
    const uint64_t umask = ballot(on);
    const uint32_t rank = (uint32_t)__builtin_popcountll(umask & lt);
#pragma unroll
    for (int w = 0; w < W; ++w)
        if (wa[w] != kNone && ((supf >> w) & 1u) == 0) {
            htot[slot_w[w]] = wgt[w] + (w == 0 ? d0 : d1);
            hrank[slot_w[w]] = rank;
        }
    wave_fence();

    // 5. the entries, in rounds of one node a lane (the chain nodes first, then the
    //    leaves): entry e of partition q's list is staged at q * kPipeMaxE + e
    uint32_t cnt[kPipeParts];
#pragma unroll
    for (uint32_t t = 0; t < kPipeParts; ++t) cnt[t] = 0;
    auto place = [&](bool node_on, uint32_t x) -> uint32_t {
        const uint32_t q = node_on ? pipe_part(x, P) : kNone;
        uint32_t e = 0;
#pragma unroll
        for (uint32_t t = 0; t < kPipeParts; ++t) {
            if (t >= P) break;
            const uint64_t mq = ballot(q == t);
            if (q == t) e = cnt[t] + (uint32_t)__builtin_popcountll(mq & lt);
            cnt[t] += (uint32_t)__builtin_popcountll(mq);
        }
        return node_on ? (q << 16) | e : kNone;
    };
    uint32_t ce[CN];
#pragma unroll
    for (int r = 0; r < CN; ++r) {
        ce[r] = place(ops[r] != kNone, ops[r]);
        e2[lane * CN + r] = ce[r] & 0xFFFFu;
        le[r * kWave + lane] = ce[r];
    }
    wave_fence();
#pragma unroll
    for (int r = 0; r < CN; ++r) {
        const uint32_t node = lane * CN + r;
        const bool written = fs[r] != kNone;
        const uint32_t at = written ? fs[r] : 0u;
        const uint32_t cn = canon[at];
        const uint32_t rt = written ? cn : nxt[r];
        const bool is_canon = written && cn == node;
        const bool is_root = rt == node;
        snext[node] = rt;               // for the leaves: root and weight from it
        sw[node] = wgt[r];
        sdel[node] = r == 0 ? d0 : d1;
        const uint32_t meta = e2[rt & (kWave * CN - 1)] | (is_root ? kPipeRootBit : 0u) |
                              (is_canon ? kPipeInstBit | (hrank[at] << kPipeRankShift) : 0u);
        const int32_t wout = is_canon ? htot[at] : wgt[r];
        if (ce[r] != kNone) {
            const uint32_t o = (ce[r] >> 16) * kPipeMaxE + (ce[r] & 0xFFFFu);
            sitem[o] = ops[r];
            swgt[o] = (uint32_t)wout;
            smeta[o] = meta;
        }
    }
    wave_fence();
#pragma unroll 1
    for (uint32_t j = CN; j < (uint32_t)R; ++j) {
        const uint32_t x = j < nr ? rsx[j * kWave + lane] : kNone;
        const bool leaf = on && x != kNone && x != a0 && x != a1;
        if (ballot(leaf) == 0) {
            if (ballot(j < nr) == 0) break;
            le[j * kWave + lane] = kNone;
            continue;
        }
        const uint32_t f = find(leaf ? x : kNone);
        const uint32_t pc = producer(leaf ? x : kNone, f);
        const uint32_t pe = place(leaf, x);
        uint32_t root_e = pe & 0xFFFFu, meta_root = kPipeRootBit;
        int32_t lw = 0;
        if (pc != 0xFFu) {
            // the producer's write: its node's value plus its delta
            const uint32_t pn = (pc & 63u) * CN + (pc >> 6);
            root_e = e2[snext[pn] & (kWave * CN - 1)];
            lw = sw[pn] + sdel[pn];
            meta_root = 0;
        } else if (f != kNone) {
            // written only by higher lanes: the value before the request
            root_e = e2[canon[f] & (kWave * CN - 1)];
            meta_root = 0;
        }
        le[j * kWave + lane] = leaf ? pe : kNone;
        if (leaf) {
            const uint32_t o = (pe >> 16) * kPipeMaxE + (pe & 0xFFFFu);
            sitem[o] = x;
            swgt[o] = (uint32_t)lw;
            smeta[o] = root_e | meta_root;
        }
    }
    wave_fence();

    // 6. the lists into the boxes a row at a time, the headers, then the hash slots back to
    //    empty; send once the box stores are complete (no release fence: it would write
    //    back the XCD's dirty L2 lines)
    uint32_t *bx = c.mbox + (size_t)cw * kPipeParts * kPipeBox;
    const uint32_t nc = (uint32_t)__builtin_popcountll(umask);
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
    for (int w = 0; w < W; ++w)
        if (slot_w[w] != kNone) {
            hkey[slot_w[w]] = kNone;
            hwlo[slot_w[w]] = 0;
            hwhi[slot_w[w]] = 0;
            canon[slot_w[w]] = kNone;
            htot[slot_w[w]] = 0;
        }
    wave_fence();
    if (nc == 0) return 0u;
    wait_stores();
    uint64_t n = 0;
    if (lane == 0) {
        n = fetch_add(c.req_tail, (uint64_t)1);
        st_rlx(&c.ring[n & c.qmask], (pipe_ring_tag(n) << kPipeTagShift) | chunks | cw);
    }
    n = bcast64(n);

    // 7. every partition's header word (tag and the request's first timestamp), then the
    //    value of each of the lane's nodes
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
            if (!sp.tick(c, kErrReplyTimeout)) { *fail = true; return 0; }
        }
        h0 = pt == 0 ? w0 : h0;
    }
    *wait_ticks += wall_clock64() - tw0;
    const uint32_t *rv = (const uint32_t *)c.reply_v + (size_t)cw * kPipeParts * kPipeMaxE;

    // 8. the transaction in program order: read a, write a - amt, read b (own write if
    //    b == a), write b + amt, then the other reads (own writes first); the log
    const uint32_t ts = (uint32_t)h0 + rank + 1;
    if (on) {
        const int32_t x0 = (int32_t)ld_rlx(&rv[(ce[0] >> 16) * kPipeMaxE + (ce[0] & 0xFFFFu)]);
        const int32_t w0v = (int32_t)((uint32_t)x0 - amt);
        const int32_t x1 = a1 == a0 ? w0v : (int32_t)ld_rlx(&rv[(ce[1] >> 16) * kPipeMaxE + (ce[1] & 0xFFFFu)]);
        const int32_t w1v = (int32_t)((uint32_t)x1 + amt);
        const int32_t va0 = a1 == a0 ? w1v : w0v;   // a's value after the transfer
        c.log_v[k] = x0;
        c.log_v[T + k] = x1;
        for (uint32_t j = CN; j < nr; ++j) {
            const uint32_t x = rsx[j * kWave + lane];
            const uint32_t p = le[j * kWave + lane];
            int32_t v;
            if (x == a1) v = w1v;
            else if (x == a0) v = va0;
            else v = (int32_t)ld_rlx(&rv[(p >> 16) * kPipeMaxE + (p & 0xFFFFu)]);
            c.log_v[(uint64_t)j * T + k] = v;
        }
        c.log_ts[k] = ts;
    }
    return on ? ts : 0u;
}

// Serving-loop bank transactions on the pipelined lane: every bank wave takes
// transactions in arrival order, admits them by version-pool bytes and sends the lanes a
// request can carry; the commit servers execute them in commit order (no transaction
// aborts) and the wave logs them with their latency from arrival to commit.
template <int R, int W>
__device__ __forceinline__ void bank_serve_pipe_client(const Ctx &c, const ServerLayout &L, uint32_t *scr,
                                                       uint32_t cw)
{
    const uint32_t lane = lane_id();
    const uint64_t NT = c.n_norm;
    const uint64_t t0 = serve_start(c);
    for (uint32_t i = lane; i < L.hsz; i += kWave) {
        scr[i] = kNone;
        scr[L.hsz + i] = 0;
        scr[2 * L.hsz + i] = 0;
        scr[3 * L.hsz + i] = kNone;
        scr[4 * L.hsz + i] = 0;
    }
    wave_fence();
    bool has = false, adm = false, exhausted = false;
    uint64_t k = 0;
    uint32_t op = 0, held = 0;
    uint64_t n_upd = 0, stalls = 0, t_reply = 0, t_round = 0, rounds = 0;
    for (;;) {
        if (ld_rlx(c.err)) break;
        const uint64_t tr0 = wall_clock64();
        if (!exhausted) {
            const uint64_t need = ballot(!has);
            if (need) {
                const uint32_t cnt = (uint32_t)__builtin_popcountll(need);
                uint64_t base = 0;
                if (lane == 0) base = fetch_add(c.next_txn, (uint64_t)cnt);
                base = bcast64(base);
                if (base + cnt >= NT) exhausted = true;
                if (!has) {
                    const uint64_t mine = base + prefix_count(need);
                    if (mine < NT) {
                        has = true;
                        adm = false;
                        k = c.norm_list ? c.norm_list[mine] : mine;
                        op = c.op[k];
                    }
                }
            }
        }
        if (ballot(has) == 0) break;
        serve_gate(c, t0, has, adm, k, W * kVersionBytes, held, stalls);
        const bool act = has && adm;
        if (ballot(act) == 0) {
            __builtin_amdgcn_s_sleep(2);
            continue;
        }
        bool fail = false;
        const uint32_t ts = tx_commit_pipe_wide<R, W>(c, L, scr, cw, act, k, op, &fail, &t_reply);
        if (fail) break;
        bool done = false;
        if (act && ts) {
            c.log_ab[k] = 0;
            c.log_lat[k] = (uint32_t)(wall_clock64() - t0 - c.arr[k]);
            done = true;
            has = false;
            ++n_upd;
        }
        serve_release(c, done, held);
        t_round += wall_clock64() - tr0;
        ++rounds;
    }
    if (lane == 0) {
        fetch_add(&c.prof[kProfClientReply], t_reply);
        fetch_add(&c.prof[kProfClientRound], t_round);
        fetch_add(&c.prof[kProfClientRounds], rounds);
        if (stalls) fetch_add(&c.serve_ctr[0], stalls);
    }
    if (n_upd) fetch_add(&c.stats[kStatCommitUpd], n_upd);
    fence_rel();
    if (lane == 0) __hip_atomic_fetch_add(c.clients_done, 1u, __ATOMIC_RELEASE,
                                          __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace csmv
