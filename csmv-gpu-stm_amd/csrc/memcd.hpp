// memcd.hpp -- memcached-style set-associative cache driver over the CSMV device API.
//
// The reference's cache (memcachedGPU/memcdSH.cu:299-448) keeps four item arrays per
// slot; here a slot of the n_sets x n_ways cache has 12 int32 items stored field-major
// (item = field * n_slots + slot): state, key low, key high, LRU timestamp, 8 value words
// (64-bit keys, 32-byte values).  Request k runs with LRU clock k / batch + 1.
//
//   GET: probe the ways for a valid matching key (memcdSH.cu:420-429); on a hit read the
//        value and, when the slot's timestamp is older than the clock, write the clock.
//   SET: read state/key/timestamp of every way and pick the slot holding the key, else
//        the first invalid way, else the least recently used one (lowest way on a tie);
//        write key, value, clock, VALID (memcdSH.cu:365-387, without its INVALID = 0
//        test that is always true, memcdSH.cu:376).
//
// Validation uses covering items: every write of a key or value also writes the slot's
// state and timestamp, so a GET logs the state words it probed and the hit slot's
// timestamp, a SET the state and timestamp of every way; keys and values are read at the
// snapshot without logging.  The request carries those reads and the covering writes
// (timestamp, state); the client installs all written items.
#pragma once

#include "csmv_device.hpp"
#include "serve.hpp"

namespace csmv {

constexpr uint32_t kMcState = 0, kMcKlo = 1, kMcKhi = 2, kMcTs = 3, kMcV0 = 4, kMcFields = 12;
constexpr uint32_t kMcGet = 0, kMcSet = 1;
constexpr int kMcR = 16;   // logged reads on the wire
constexpr int kMcW = 2;    // covering writes on the wire
constexpr int kMcInstall = 12;

__host__ __device__ __forceinline__ uint32_t mc_set_of(uint64_t key, uint32_t n_sets)
{
    uint64_t z = key + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (uint32_t)(z & (uint64_t)(n_sets - 1));
}

__host__ __device__ __forceinline__ uint32_t mc_value_word(uint64_t key, uint64_t k, uint32_t w)
{
    uint64_t z = (key ^ (k * 8 + w)) + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (uint32_t)(z ^ (z >> 31));
}

template <int R, int W> __device__ __forceinline__ void memcd_client(const Ctx &c, uint32_t cw)
{
    const uint32_t lane = lane_id();
    const uint64_t NT = c.n_txns;
    const uint32_t ns = c.n_slots, ways = c.n_ways;
    // serving loop: requests become eligible at their arrival tick and are admitted by
    // in-flight version bytes; latency counts from arrival
    const bool serving = c.arr != nullptr;
    const uint64_t t0 = serving ? serve_start(c) : 0;
    bool has = false, adm = false, exhausted = false;
    uint64_t k = 0, t_pick = 0;
    uint32_t ab = 0, my_last = 0, held = 0;
    uint64_t n_upd = 0, n_ro = 0, n_age = 0, stalls = 0, t_reply = 0, t_round = 0, rounds = 0;
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
                        k = mine;
                        ab = 0;
                        t_pick = wall_clock64();
                    }
                }
            }
        }
        if (ballot(has) == 0) break;
        if (serving) {
            const uint32_t need =
                has ? (c.req_op[k] == kMcSet ? kMcInstall : 1u) * kVersionBytes : 0u;
            serve_gate(c, t0, has, adm, k, need, held, stalls);
        }
        const bool act = has && (!serving || adm);
        if (serving && ballot(act) == 0) {
            __builtin_amdgcn_s_sleep(2);
            continue;
        }

        Tx<R, W> tx;
        if (!tx_begin(c, tx, c.begin_wait ? my_last : 0u)) break;
        uint32_t res = 0, vsum = 0;
        uint32_t iwa[kMcInstall];
        int32_t iwv[kMcInstall];
#pragma unroll
        for (int i = 0; i < kMcInstall; ++i) { iwa[i] = kNone; iwv[i] = 0; }
        int inw = 0;
        if (act) {
            const uint64_t key = c.req_key[k];
            const uint32_t op = c.req_op[k];
            const uint32_t lo = (uint32_t)key, hi = (uint32_t)(key >> 32);
            const uint32_t set = mc_set_of(key, c.n_sets);
            const uint32_t clock = (uint32_t)(k / c.mc_batch + 1);
            // states, keys and timestamps of every way at the snapshot, in two batches of
            // loads shared by GET and SET lanes (one pass for a wave that holds both)
            uint32_t ia[16], ib[16];
            int32_t va[16], vb[16];
#pragma unroll
            for (int w = 0; w < 8; ++w) {
                const bool on = (uint32_t)w < ways;
                const uint32_t slot = set * ways + (uint32_t)w;
                ia[w] = on ? kMcState * ns + slot : kNone;
                ia[8 + w] = on ? kMcKlo * ns + slot : kNone;
                ib[w] = on ? kMcKhi * ns + slot : kNone;
                ib[8 + w] = on ? kMcTs * ns + slot : kNone;
            }
            // A GET walks back the states and keys newer than its snapshot (a read-only GET
            // commits at its snapshot) and, once its hit is known, that slot's timestamp; it
            // never uses the other ways' timestamps.  A SET logs the state and timestamp of
            // every way, and every write of a key writes the slot's state too, so a SET that
            // read anything written since its snapshot is doomed at validation: it aborts at
            // once instead of walking.
            const bool get = op == kMcGet;
            bool ok = true;
            const uint32_t na = pool_read_some<16>(c, ia, tx.snap, va, get ? 0xFFFFu : 0u, ok);
            const uint32_t nb = pool_read_some<16>(c, ib, tx.snap, vb, get ? 0x00FFu : 0u, ok);
            if (!ok) tx.aborted = true;
            if (op == kMcGet) {
                // the hit way, then its value; the states are logged up to the hit way
                // (the ones a sequential probe reads) and the hit slot's timestamp
                int hit = -1;
#pragma unroll
                for (int w = 7; w >= 0; --w)
                    if ((uint32_t)w < ways && va[w] == 1 && (uint32_t)va[8 + w] == lo &&
                        (uint32_t)vb[w] == hi)
                        hit = w;
#pragma unroll
                for (int w = 0; w < 8; ++w)
                    if ((uint32_t)w < ways && (hit < 0 || w <= hit)) log_read(tx, ia[w]);
                if (hit >= 0 && ((nb >> (8 + hit)) & 1u) != 0) {
                    int32_t tw = 0;
                    if (!pool_read_below(c, kMcTs * ns + set * ways + (uint32_t)hit, tx.snap, tw))
                        tx.aborted = true;
#pragma unroll
                    for (int w = 0; w < 8; ++w) vb[8 + w] = (w == hit) ? tw : vb[8 + w];
                }
                if (hit >= 0 && !tx.aborted) {
                    const uint32_t slot = set * ways + (uint32_t)hit;
                    uint32_t iv[8];
                    int32_t vv[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) iv[j] = (kMcV0 + j) * ns + slot;
                    if (!pool_read_chunk<8>(c, iv, tx.snap, vv)) tx.aborted = true;
                    uint32_t acc = 0;
#pragma unroll
                    for (int j = 0; j < 8; ++j) acc += (uint32_t)vv[j] * (2u * j + 1u);
                    uint32_t t = 0;
#pragma unroll
                    for (int w = 0; w < 8; ++w) t = (w == hit) ? (uint32_t)vb[8 + w] : t;
                    log_read(tx, kMcTs * ns + slot);
                    res = 1u | ((uint32_t)hit << 4);
                    vsum = acc;
                    if (t < clock && !tx.aborted) {
                        res |= 2u;
                        tx_write(tx, kMcTs * ns + slot, (int32_t)clock);
                        iwa[0] = kMcTs * ns + slot;
                        iwv[0] = (int32_t)clock;
                        inw = 1;
                    }
                }
            } else {
                // SET: the way holding the key, else the first invalid way, else the least
                // recently used; states and timestamps are logged
                if (na != 0 || nb != 0) tx.aborted = true;
                int found = -1, invalid = -1, lru = 0;
                uint32_t lru_ts = kNone;
#pragma unroll
                for (int w = 0; w < 8; ++w) {
                    if ((uint32_t)w >= ways) continue;
                    const int32_t st = va[w];
                    const uint32_t ts = (uint32_t)vb[8 + w];
                    log_read(tx, ia[w]);
                    log_read(tx, ib[8 + w]);
                    if (st == 1 && (uint32_t)va[8 + w] == lo && (uint32_t)vb[w] == hi && found < 0)
                        found = w;
                    if (st != 1 && invalid < 0) invalid = w;
                    if (ts < lru_ts) { lru_ts = ts; lru = w; }
                }
                const uint32_t w = found >= 0 ? (uint32_t)found
                                 : invalid >= 0 ? (uint32_t)invalid : (uint32_t)lru;
                const uint32_t slot = set * ways + w;
                res = (found >= 0 ? 1u : 0u) | 2u | (w << 4);
                tx_write(tx, kMcTs * ns + slot, (int32_t)clock);
                tx_write(tx, kMcState * ns + slot, 1);
                iwa[0] = kMcKlo * ns + slot; iwv[0] = (int32_t)lo;
                iwa[1] = kMcKhi * ns + slot; iwv[1] = (int32_t)hi;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    iwa[2 + j] = (kMcV0 + j) * ns + slot;
                    iwv[2 + j] = (int32_t)mc_value_word(key, k, (uint32_t)j);
                }
                iwa[10] = kMcTs * ns + slot; iwv[10] = (int32_t)clock;
                iwa[11] = kMcState * ns + slot; iwv[11] = 1;
                inw = kMcInstall;
            }
        }
        if (tx.overflow) raise_error(c, kErrReadSetOverflow);
        const bool ok_exec = act && !tx.aborted;
        if (act && tx.aborted) { ++ab; ++n_age; }
        bool done = false;
        if (ok_exec && !(res & 2u)) {
            // read-only: commits at its snapshot without a round trip (memcdSH.cu:176-183)
            c.log_ts[k] = tx.snap;
            c.log_res[k] = res;
            c.log_vsum[k] = vsum;
            c.log_ab[k] = ab;
            c.log_lat[k] = (uint32_t)(wall_clock64() - (serving ? t0 + c.arr[k] : t_pick));
            done = true;
            has = false;
            ++n_ro;
        }
        const bool req = ok_exec && (res & 2u);
        if (ballot(req)) {
            bool fail = false;
            const uint32_t ts = tx_commit<R, W, kMcInstall>(c, cw, tx, req, &fail, &t_reply, false,
                                                            iwa, iwv, req ? inw : 0);
            if (fail) break;
            const uint64_t cm = ballot(ts != 0);
            if (cm) my_last = readlane(ts, 63u - (uint32_t)__builtin_clzll(cm));
            if (req) {
                if (ts) {
                    c.log_ts[k] = ts;
                    c.log_res[k] = res;
                    c.log_vsum[k] = vsum;
                    c.log_ab[k] = ab;
                    c.log_lat[k] = (uint32_t)(wall_clock64() - (serving ? t0 + c.arr[k] : t_pick));
                    done = true;
                    has = false;
                    ++n_upd;
                } else {
                    ++ab;
                }
            }
        }
        if (serving) serve_release(c, done, held);
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
    if (n_ro) fetch_add(&c.stats[kStatCommitRo], n_ro);
    if (n_age) fetch_add(&c.stats[kStatAbortAge], n_age);
    fence_rel();
    if (lane == 0) __hip_atomic_fetch_add(c.clients_done, 1u, __ATOMIC_RELEASE,
                                          __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace csmv
