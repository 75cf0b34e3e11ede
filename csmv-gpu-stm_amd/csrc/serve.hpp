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

}  // namespace csmv
