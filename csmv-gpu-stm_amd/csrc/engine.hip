// engine.hip -- gfx950 CSMV engine: bank driver, persistent kernel, host engine, C-ABI.
//
// One persistent kernel replaces the reference's device-side launch of server and client
// grids (parent_kernel, bank/csmv_sh/exe.cu:319-338; HIP has no dynamic parallelism):
// workgroup 0 is the commit server (csmv_server.hpp), every other workgroup holds client
// wavefronts.  The grid never exceeds what the device keeps resident, so spinning
// clients cannot starve the server, and every wait is bounded (Spin).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "csmv.h"
#include "csmv_device.hpp"
#include "csmv_server.hpp"
#include "memcd.hpp"
#include "mlp.hpp"
#include "serve.hpp"

namespace csmv {

// Reads of one update transaction on the wire, writes of one transaction.
constexpr int kBankR = 4;
constexpr int kBankW = 2;
// serving loop: bank transactions read 2..64 accounts
constexpr int kServeBankR = 64;
// reply values per client wave: the serialized lane's rows (a new value and a previous
// head per write and lane) or the pipelined lane's entry values of every partition
constexpr uint32_t kReplyVWords = std::max<uint32_t>(2 * kBankW * kWave, kPipeParts * kPipeMaxE);
// the pipelined lane's boxes of a client wave (one per partition), in mailbox rows
constexpr uint32_t kPipeBoxRows = (kPipeParts * kPipeBox + kWave - 1) / kWave;

// ---- stream generation ----------------------------------------------------------------------
__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// RAND_R_FNC (bank/csmv_sh/API.cuh:37-53): three LCG steps (a = 1103515245, c = 12345,
// mod 2^64) combining 11 + 10 + 10 bits taken from bit 16 of each state.
__host__ __device__ __forceinline__ uint32_t rand_r_fnc(uint64_t &state)
{
    uint64_t next = state * 1103515245ull + 12345ull;
    uint64_t result = (next >> 16) & 2047u;
    next = next * 1103515245ull + 12345ull;
    result = (result << 10) ^ ((next >> 16) & 1023u);
    next = next * 1103515245ull + 12345ull;
    result = (result << 10) ^ ((next >> 16) & 1023u);
    state = next;
    return (uint32_t)result;
}

__device__ __forceinline__ uint32_t zipf_rank(const uint32_t *cdf, uint32_t n, uint32_t r)
{
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (r < cdf[mid]) hi = mid; else lo = mid + 1;
    }
    return lo;
}

__device__ __forceinline__ uint32_t scramble(uint32_t rank, uint32_t n)
{
    return (n & (n - 1)) == 0 ? (rank * 2654435761u) & (n - 1) : rank;
}

struct GenArgs {
    uint64_t seed, n_txns;
    uint32_t n_items, n_reads, n_reads_ro, ro_thr, dist, amount_max, rmax, n_reads_min;
    const uint32_t *cdf;
    uint32_t *op, *addr;
};

// One lane per transaction k; state = splitmix64(seed + k), then the driver's draws:
// read-only test rnd < prRead*0xFFFF (bank/csmv_sh/exe.cu:134-139), amount, the
// read-set size when it varies (serving mix), operands.
__global__ void __launch_bounds__(256) bank_generate_kernel(GenArgs g)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g.n_txns) return;
    uint64_t st = splitmix64(g.seed + k);
    const uint32_t rnd = rand_r_fnc(st) & 0xFFFFu;
    const uint32_t kind = rnd < g.ro_thr ? kKindAudit : kKindTransfer;
    uint32_t amt = 0, nr = g.n_reads_ro;
    if (kind == kKindTransfer) {
        amt = 1 + rand_r_fnc(st) % g.amount_max;
        nr = g.n_reads;
        if (g.n_reads_min) nr = g.n_reads_min + rand_r_fnc(st) % (g.n_reads - g.n_reads_min + 1);
    }
    g.op[k] = kind | (amt << 8) | (nr << 16);
    for (uint32_t j = 0; j < g.rmax; ++j) {
        uint32_t it = kNone;
        if (j < nr) {
            const uint32_t r = rand_r_fnc(st);
            it = g.dist == 1 ? scramble(zipf_rank(g.cdf, g.n_items, r), g.n_items)
                             : r % g.n_items;
        }
        g.addr[(uint64_t)j * g.n_txns + k] = it;
    }
}

__global__ void __launch_bounds__(256) pool_fill_kernel(uint32_t *ht, uint64_t *slot, uint64_t *sv,
                                                        uint32_t n, uint32_t depth,
                                                        const int32_t *values, int32_t fill)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ht[i] = 0;
    const int32_t v = values ? values[i] : fill;
    slot[(size_t)i * depth] = (uint32_t)v;
    sv[i] = (uint32_t)v;   // stamp 0
    for (uint32_t d = 1; d < depth; ++d) slot[(size_t)i * depth + d] = 0;
}

// A run's commit stamps start at zero; the committed values carry over from the pool.
__global__ void __launch_bounds__(256) stamp_clear_kernel(uint64_t *sv, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) sv[i] &= 0xFFFFFFFFull;
}

// Pipelined serialized lane: an item's committer word is its newest value << 32 | the
// number of installs its version ring has taken (head h and tail t give it: t stays 0 until
// the ring is full, after which t = h + 1).
__global__ void __launch_bounds__(256) pipe_init_kernel(const uint32_t *ht, const uint64_t *slot,
                                                        uint64_t *sv, uint32_t n, uint32_t depth)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t h = ht[i] & 0xFFFFu, t = ht[i] >> 16;
    const uint32_t v = (uint32_t)slot[(size_t)i * depth + h];
    sv[i] = ((uint64_t)v << 32) | (t == 0 ? h : h + depth);
}

// Pipelined serialized lane run with no normal-lane transaction (no snapshot reader during
// the run): the commit servers install the versions but leave the head words, which the
// install counts give at the end (j installs: head j mod depth, tail 0 until the ring is
// full, then head + 1).
__global__ void __launch_bounds__(256) pipe_heads_kernel(const uint64_t *sv, uint32_t *ht, uint32_t n,
                                                         uint32_t depth)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = (uint32_t)sv[i];
    ht[i] = (j % depth) | ((j < depth ? 0u : (j + 1) % depth) << 16);
}

__global__ void __launch_bounds__(256) pool_head_kernel(const uint32_t *ht, const uint64_t *slot,
                                                        uint32_t n, uint32_t depth, int32_t *out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = (int32_t)(uint32_t)slot[(size_t)i * depth + (ht[i] & 0xFFFFu)];
}

// ---- bank driver (client wavefronts) -----------------------------------------------------------
// A bank transaction as the serialized lane carries it (the commit server executes it in
// commit order): the operands as the read set, the transfer pair as the write set, and
// the stream index and amount (kReadOnlyAmt for an audit) in the write-value rows.
template <int R, int W>
__device__ __forceinline__ void delegated_tx(const Ctx &c, Tx<R, W> &tx, bool on, uint64_t k,
                                             uint32_t op)
{
    tx.snap = on ? 0u : kNone;
    tx.nr = 0;
    tx.nw = 0;
    tx.aborted = false;
    tx.overflow = false;
#pragma unroll
    for (int r = 0; r < R; ++r) tx.rs[r] = kNone;
#pragma unroll
    for (int w = 0; w < W; ++w) { tx.wa[w] = kNone; tx.wv[w] = 0; }
    const uint32_t amt = (op >> 8) & 0xFFu, nr = (op >> 16) & 0xFFu;
    const bool audit = (op & 0xFFu) == kKindAudit;
    if (on) {
        // the operand rows in one batch of loads (a row past nr reloads row 0)
        uint32_t ops[R];
#pragma unroll
        for (int j = 0; j < R; ++j) ops[j] = c.addr[(uint64_t)((uint32_t)j < nr ? j : 0) * c.n_txns + k];
        pin_round(ops);
#pragma unroll
        for (int j = 0; j < R; ++j)
            if ((uint32_t)j < nr) {
                tx.rs[j] = ops[j];
                if (j < 2 && !audit) tx_write(tx, ops[j], 0);
            }
        tx.nr = nr;
        tx.wv[0] = (int32_t)(uint32_t)k;
        tx.wv[1] = (int32_t)(audit ? kReadOnlyAmt : amt);
    }
}

// delegated_tx for S transactions a lane, their operand rows in one batch of loads (an idle
// slot loads stream index 0's rows, in bounds)
template <int R, int W, int S>
__device__ __forceinline__ void delegated_txs(const Ctx &c, Tx<R, W> (&tx)[S], const bool (&on)[S],
                                              const uint64_t (&k)[S], const uint32_t (&op)[S])
{
    uint32_t ops[S][R];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const uint32_t nr = on[s] ? (op[s] >> 16) & 0xFFu : 0u;
        const uint64_t ks = on[s] ? k[s] : 0;
#pragma unroll
        for (int j = 0; j < R; ++j) ops[s][j] = c.addr[(uint64_t)((uint32_t)j < nr ? j : 0) * c.n_txns + ks];
    }
#pragma unroll
    for (int s = 0; s < S; ++s) pin_round(ops[s]);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        Tx<R, W> &t = tx[s];
        t.snap = on[s] ? 0u : kNone;
        t.nr = 0;
        t.nw = 0;
        t.aborted = false;
        t.overflow = false;
#pragma unroll
        for (int r = 0; r < R; ++r) t.rs[r] = kNone;
#pragma unroll
        for (int w = 0; w < W; ++w) { t.wa[w] = kNone; t.wv[w] = 0; }
        const uint32_t amt = (op[s] >> 8) & 0xFFu, nr = (op[s] >> 16) & 0xFFu;
        const bool audit = (op[s] & 0xFFu) == kKindAudit;
        if (on[s]) {
#pragma unroll
            for (int j = 0; j < R; ++j)
                if ((uint32_t)j < nr) {
                    t.rs[j] = ops[s][j];
                    if (j < 2 && !audit) tx_write(t, ops[s][j], 0);
                }
            t.nr = nr;
            t.wv[0] = (int32_t)(uint32_t)k[s];
            t.wv[1] = (int32_t)(audit ? kReadOnlyAmt : amt);
        }
    }
}

// client_kernel (bank/csmv_sh/exe.cu:89-284) on a resident stream: a lane holds one
// transaction until it commits; lanes that finish are refilled with one fetch_add per
// wave (ballot + prefix count).  Audits commit locally at their snapshot
// (exe.cu:213-223); transfers go through tx_commit.  With c.escalate set, a transaction
// that has failed that many attempts is sent on the serialized lane instead (forward
// progress: the reference retries in arrival order until it commits, exe.cu:142-269).
template <int R, int W, bool kPipe = false>
__device__ __forceinline__ void bank_client(const Ctx &c, uint32_t cw)
{
    const uint32_t lane = lane_id();
    const uint64_t T = c.n_txns;
    const uint64_t NT = c.n_norm;
    bool has = false, exhausted = false;
    uint64_t k = 0;
    uint32_t op = 0, ab = 0, my_last = 0;
    uint64_t n_upd = 0, n_ro = 0, n_age = 0, n_esc = 0, t_reply = 0, t_round = 0, rounds = 0;
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
                        k = c.norm_list ? c.norm_list[mine] : mine;
                        op = c.op[k];
                        ab = 0;
                    }
                }
            }
        }
        if (ballot(has) == 0) break;

        if (c.escalate) {
            // this round's request carries the wave's escalated lanes (transfers, and audits
            // whose reads fit the wire) on the serialized lane; the other lanes sit it out
            const bool esc = has && ab >= c.escalate &&
                             ((op & 0xFFu) == kKindTransfer || ((op >> 16) & 0xFFu) <= (uint32_t)R);
            if (ballot(esc)) {
                Tx<R, W> tx;
                delegated_tx<R, W>(c, tx, esc, k, op);
                bool fail = false;
                const uint32_t ts = tx_commit<R, W>(c, cw, tx, esc, &fail, &t_reply, true);
                if (fail) break;
                const uint64_t cmu = ballot(ts != 0 && (ts & kSnapshotBit) == 0);
                if (cmu) my_last = readlane(ts, 63u - (uint32_t)__builtin_clzll(cmu));
                if (esc) {
                    if (ts) {
                        c.log_ab[k] = ab;
                        has = false;
                        ++n_esc;
                        if (ts & kSnapshotBit) ++n_ro;
                        else ++n_upd;
                    } else {
                        ++ab;
                    }
                }
                t_round += wall_clock64() - tr0;
                ++rounds;
                continue;
            }
        }

        Tx<R, W> tx;
        // the pipelined lane's commit partitions publish progress words instead of one clock
        // (and this lane commits no updates to wait for)
        if (!tx_begin(c, tx, (c.begin_wait && !kPipe) ? my_last : 0u)) break;
        if constexpr (kPipe) tx.snap = pipe_snapshot(c);
        const uint32_t kind = op & 0xFFu, amt = (op >> 8) & 0xFFu, nr = (op >> 16) & 0xFFu;
        const bool ro = has && kind == kKindAudit;
        const bool upd = has && kind == kKindTransfer;
        if constexpr (kPipe) {
            // the pipelined lane's servers take only staged entry lists: a transfer left on
            // this lane is a routing error
            if (ballot(upd)) {
                raise_error(c, kErrPipeNormal);
                break;
            }
        }
        int32_t reads[R];
#pragma unroll
        for (int j = 0; j < R; ++j) reads[j] = 0;
        uint32_t sum = 0;
        if (ro) {
            // audit: snapshot reads, 16 at a time (two rounds of independent loads each)
            bool ok = true;
#pragma unroll 1
            for (uint32_t c0 = 0; c0 < nr; c0 += 16) {
                uint32_t it[16];
                int32_t v[16];
                // (every row load issued, in bounds: a read past nr loads row 0)
#pragma unroll
                for (int j = 0; j < 16; ++j) it[j] = c.addr[(uint64_t)(c0 + j < nr ? c0 + j : 0u) * T + k];
                pin_round(it);
#pragma unroll
                for (int j = 0; j < 16; ++j) it[j] = c0 + j < nr ? it[j] : kNone;
                ok = pool_read_chunk<16>(c, it, tx.snap, v) && ok;
#pragma unroll
                for (int j = 0; j < 16; ++j) sum += it[j] != kNone ? (uint32_t)v[j] : 0u;
            }
            if (!ok) tx.aborted = true;
        } else if (upd) {
            // transfer body (exe.cu:184-201) over one batch of snapshot reads: read a,
            // write a-amt, read b (own write if b == a), write b+amt, then the extra reads
            // of the wider transaction (own writes first); every read is logged
            uint32_t it[R];
            int32_t v[R];
#pragma unroll
            for (int j = 0; j < R; ++j) it[j] = c.addr[(uint64_t)((uint32_t)j < nr ? j : 0) * T + k];
            pin_round(it);
#pragma unroll
            for (int j = 0; j < R; ++j) it[j] = (uint32_t)j < nr ? it[j] : kNone;
            if (!pool_read_chunk<R>(c, it, tx.snap, v)) tx.aborted = true;
#pragma unroll
            for (int j = 0; j < R; ++j) {
                if ((uint32_t)j < nr) {
                    int32_t x = v[j];
#pragma unroll
                    for (int w = 0; w < W; ++w)
                        x = ((uint32_t)w < tx.nw && tx.wa[w] == it[j]) ? tx.wv[w] : x;
                    reads[j] = x;
                    log_read(tx, it[j]);
                    if (j == 0) tx_write(tx, it[j], (int32_t)((uint32_t)x - amt));
                    else if (j == 1) tx_write(tx, it[j], (int32_t)((uint32_t)x + amt));
                }
            }
        }
        if (tx.overflow) raise_error(c, kErrReadSetOverflow);
        if (ro) {
            if (!tx.aborted) {
                c.log_ts[k] = tx.snap;
                c.log_v[k] = (int32_t)sum;
                c.log_ab[k] = ab;
                has = false;
                ++n_ro;
            } else {
                ++ab;
                ++n_age;
            }
        }
        if (upd && tx.aborted) { ++ab; ++n_age; }
        const bool req = upd && !tx.aborted;
        if (ballot(req)) {
            bool fail = false;
            const uint32_t ts = tx_commit<R, W>(c, cw, tx, req, &fail, &t_reply);
            if (fail) break;
            const uint64_t cm = ballot(ts != 0);
            if (cm) my_last = readlane(ts, 63u - (uint32_t)__builtin_clzll(cm));
            if (req) {
                if (ts) {
                    c.log_ts[k] = ts;
#pragma unroll
                    for (int j = 0; j < R; ++j)
                        if ((uint32_t)j < nr) c.log_v[(uint64_t)j * T + k] = reads[j];
                    c.log_ab[k] = ab;
                    has = false;
                    ++n_upd;
                } else {
                    ++ab;
                }
            }
        }
        if (has && ab > c.max_attempts) raise_error(c, kErrLivelock);
        t_round += wall_clock64() - tr0;
        ++rounds;
    }
    if (lane == 0) {
        fetch_add(&c.prof[kProfClientReply], t_reply);
        fetch_add(&c.prof[kProfClientRound], t_round);
        fetch_add(&c.prof[kProfClientRounds], rounds);
    }
    if (n_upd) fetch_add(&c.stats[kStatCommitUpd], n_upd);
    if (n_ro) fetch_add(&c.stats[kStatCommitRo], n_ro);
    if (n_age) fetch_add(&c.stats[kStatAbortAge], n_age);
    if (n_esc) fetch_add(&c.stats[kStatEscalated], n_esc);
    fence_rel();
    if (lane == 0) __hip_atomic_fetch_add(c.clients_done, 1u, __ATOMIC_RELEASE,
                                          __HIP_MEMORY_SCOPE_AGENT);
}

// Serialized lane: the wave hands routed transfers to the commit server, which executes
// them in commit order on the newest committed values (no snapshot, so no conflict
// window); the wave installs the versions the server returns.
template <int R, int W>
__device__ __forceinline__ void bank_serial_client(const Ctx &c, uint32_t cw)
{
    const uint32_t lane = lane_id();
    const uint64_t NH = c.n_hot;
    bool has = false, exhausted = false;
    uint64_t k = 0;
    uint32_t op = 0, ab = 0;
    uint64_t n_upd = 0, n_ro = 0, t_reply = 0, t_round = 0, rounds = 0;
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
                        k = c.hot_list[mine];
                        op = c.op[k];
                        ab = 0;
                    }
                }
            }
        }
        if (ballot(has) == 0) break;
        Tx<R, W> tx;
        delegated_tx<R, W>(c, tx, has, k, op);
        bool fail = false;
        const uint32_t ts = tx_commit<R, W>(c, cw, tx, has, &fail, &t_reply, true);
        if (fail) break;
        if (has) {
            if (ts) {
                c.log_ab[k] = ab;
                has = false;
                if (ts & kSnapshotBit) ++n_ro;
                else ++n_upd;
            } else {
                ++ab;
            }
        }
        if (has && ab > c.max_attempts) raise_error(c, kErrLivelock);
        t_round += wall_clock64() - tr0;
        ++rounds;
    }
    if (lane == 0) {
        fetch_add(&c.prof[kProfClientReply], t_reply);
        fetch_add(&c.prof[kProfClientRound], t_round);
        fetch_add(&c.prof[kProfClientRounds], rounds);
    }
    if (n_upd) fetch_add(&c.stats[kStatCommitUpd], n_upd);
    if (n_ro) fetch_add(&c.stats[kStatCommitRo], n_ro);
    fence_rel();
    if (lane == 0) __hip_atomic_fetch_add(c.clients_done, 1u, __ATOMIC_RELEASE,
                                          __HIP_MEMORY_SCOPE_AGENT);
}

// The pipelined lane's client wave: kPipeSlots transactions a lane (the lower slots the
// older), a request of up to 64 kPipeSlots (tx_commit_pipe).  A transaction the request
// could not carry (its partition's list was full) stays in its slot for the next request;
// nothing aborts.
template <int R, int W>
__device__ __forceinline__ void bank_pipe_client(const Ctx &c, const ServerLayout &L, uint32_t *scr,
                                                 uint32_t cw)
{
    constexpr int S = kPipeSlots;
    const uint32_t lane = lane_id();
    const uint64_t NH = c.n_hot;
    pipe_client_init<R, W, S>(scr);
    bool has[S], exhausted = false;
    uint64_t k[S];
    uint32_t op[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        has[s] = false;
        k[s] = 0;
        op[s] = 0;
    }
    // the held transactions to the lowest slots, in order
    auto compact = [&]() {
#pragma unroll
        for (int pass = 0; pass + 1 < S; ++pass)
#pragma unroll
            for (int s = 0; s + 1 < S; ++s)
                if (!has[s] && has[s + 1]) {
                    has[s] = true;
                    has[s + 1] = false;
                    k[s] = k[s + 1];
                    op[s] = op[s + 1];
                }
    };
    uint64_t n_upd = 0, n_ro = 0, t_reply = 0, t_round = 0, rounds = 0;
    for (;;) {
        if (ld_rlx(c.err)) break;
        const uint64_t tr0 = wall_clock64();
        compact();
        if (!exhausted) {
            uint32_t want = 0;
#pragma unroll
            for (int s = 0; s < S; ++s) want += has[s] ? 0u : 1u;
            const uint32_t incl = wave_incl_scan(want);
            const uint32_t total = readlane(incl, 63);
            if (total) {
                uint64_t base = 0;
                if (lane == 0) base = fetch_add(c.next_hot, (uint64_t)total);
                base = bcast64(base);
                if (base + total >= NH) exhausted = true;
                // every slot's loads in one batch per level (stream index, then kind, then
                // operands): three round trips a round whatever the slot count
                uint64_t mine = base + (incl - want);
                bool take[S];
                uint64_t at[S];
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    take[s] = !has[s] && mine < NH;
                    at[s] = take[s] ? mine : 0;
                    mine += has[s] ? 0 : 1;
                }
                uint64_t kn[S];
#pragma unroll
                for (int s = 0; s < S; ++s) kn[s] = c.hot_list[at[s]];
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    k[s] = take[s] ? kn[s] : k[s];
                    has[s] = has[s] || take[s];
                }
                uint32_t on_[S];
#pragma unroll
                for (int s = 0; s < S; ++s) on_[s] = c.op[k[s]];
#pragma unroll
                for (int s = 0; s < S; ++s) op[s] = take[s] ? on_[s] : op[s];
                compact();
            }
        }
        bool any = false;
#pragma unroll
        for (int s = 0; s < S; ++s) any = any || has[s];
        if (ballot(any) == 0) break;
        Tx<R, W> tx[S];
        delegated_txs<R, W, S>(c, tx, has, k, op);
        bool fail = false;
        uint32_t ts[S];
        tx_commit_pipe<R, W, S>(c, L, scr, cw, tx, has, ts, &fail, &t_reply);
        if (fail) break;
#pragma unroll
        for (int s = 0; s < S; ++s)
            if (has[s] && ts[s]) {
                c.log_ab[k[s]] = 0;
                has[s] = false;
                if (ts[s] & kSnapshotBit) ++n_ro;
                else ++n_upd;
            }
        t_round += wall_clock64() - tr0;
        ++rounds;
    }
    if (lane == 0) {
        fetch_add(&c.prof[kProfClientReply], t_reply);
        fetch_add(&c.prof[kProfClientRound], t_round);
        fetch_add(&c.prof[kProfClientRounds], rounds);
    }
    if (n_upd) fetch_add(&c.stats[kStatCommitUpd], n_upd);
    if (n_ro) fetch_add(&c.stats[kStatCommitRo], n_ro);
    fence_rel();
    if (lane == 0) __hip_atomic_fetch_add(c.clients_done, 1u, __ATOMIC_RELEASE,
                                          __HIP_MEMORY_SCOPE_AGENT);
}

template <int R, int W>
__global__ void __launch_bounds__(512) bank_kernel(Ctx c, ServerLayout L)
{
    extern __shared__ uint32_t lds[];
    if (blockIdx.x == 0) {
        server_workgroup<R, W, true>(c, L, lds);
        return;
    }
    const uint32_t wave = threadIdx.x / kWave;
    if (wave >= c.client_wg_waves) return;
    const uint32_t cw = (blockIdx.x - 1) * c.client_wg_waves + wave;
    if (cw >= c.n_client_waves) return;
    if (cw >= c.n_client_waves - c.n_serial_waves) bank_serial_client<R, W>(c, cw);
    else bank_client<R, W>(c, cw);
}

// The pipelined serialized lane's kernel (variant bit kSerialPipe): its own server
// workgroup of up to 12 waves (several committers), clients that stage their chains in
// their workgroup's LDS, and the normal lane's audits.
template <int R, int W>
__global__ void __launch_bounds__(1024) bank_pipe_kernel(Ctx c, ServerLayout L)
{
    extern __shared__ uint32_t lds[];
    if (blockIdx.x < L.parts) {
        pipe_server_workgroup<R, W>(c, L, lds, blockIdx.x);
        return;
    }
    const uint32_t wave = threadIdx.x / kWave;
    if (wave >= c.client_wg_waves) return;
    const uint32_t cw = (blockIdx.x - L.parts) * c.client_wg_waves + wave;
    if (cw >= c.n_client_waves) return;
    if (cw >= c.n_client_waves - c.n_serial_waves)
        bank_pipe_client<R, W>(c, L, lds + wave * pipe_client_words(R, W, kPipeSlots), cw);
    else bank_client<R, W, true>(c, cw);
}

struct McGenArgs {
    uint64_t seed, n_reqs, n_keys;
    uint32_t get_thr, dist;
    const uint32_t *cdf;
    uint64_t *key;
    uint32_t *op;
};

// Request k: GET/SET as the driver's rnd > prRead test (memcachedGPU/memcdSH.cu:335-343),
// then a key rank from the Zipf table over the keyspace and a 64-bit key of that rank.
__global__ void __launch_bounds__(256) memcd_generate_kernel(McGenArgs g)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g.n_reqs) return;
    uint64_t st = splitmix64(g.seed + k);
    const uint32_t rnd = rand_r_fnc(st) & 0xFFFFu;
    g.op[k] = rnd < g.get_thr ? kMcGet : kMcSet;
    const uint32_t r = rand_r_fnc(st);
    const uint32_t rank = g.dist == 1 ? zipf_rank(g.cdf, (uint32_t)g.n_keys, r)
                                      : (uint32_t)(r % g.n_keys);
    g.key[k] = splitmix64(0x6D656D6364ull ^ ((uint64_t)rank << 1));
}

template <int R, int W>
__global__ void __launch_bounds__(512) memcd_kernel(Ctx c, ServerLayout L)
{
    extern __shared__ uint32_t lds[];
    if (blockIdx.x == 0) {
        server_workgroup<R, W, false>(c, L, lds);
        return;
    }
    const uint32_t wave = threadIdx.x / kWave;
    if (wave >= c.client_wg_waves) return;
    const uint32_t cw = (blockIdx.x - 1) * c.client_wg_waves + wave;
    if (cw < c.n_client_waves) memcd_client<R, W>(c, cw);
}

// Serving loop: one persistent kernel with a commit server per partition and the client
// workgroups of each partition.  Partitions hold disjoint data (cache partitions disjoint
// ranges of sets, the bank partition the accounts), so each keeps its own clock and commit
// order.  Partitions 0 .. n_cache - 1 are cache partitions; the bank partition, if any, is
// the last.  Every partition's context travels in the kernel arguments.
constexpr uint32_t kMaxServeParts = 8;
struct ServeArgs {
    Ctx c[kMaxServeParts];
    ServerLayout L[kMaxServeParts];
    uint32_t n_cache, has_bank;
    uint32_t bank_pipe;                // the bank partition runs the pipelined lane
    uint32_t wg_end[kMaxServeParts];   // client workgroups of partitions 0..i (prefix sums)
};
static_assert(sizeof(ServeArgs) <= 4096, "the serving kernel's arguments fit the kernarg limit");

// kPipeBank: the bank partition runs the pipelined lane (a kernel of its own, so that the
// other bank roles' registers do not add to it).
template <int RM, int WM, int RB, int WB, bool kPipeBank>
__global__ void __launch_bounds__(512) serve_kernel(ServeArgs a)
{
    extern __shared__ uint32_t lds[];
    const uint32_t n_part = a.n_cache + a.has_bank;
    // server workgroups: one per cache partition, then the bank partition's (one per
    // commit partition of a pipelined bank lane)
    const uint32_t n_srv = a.n_cache + (a.has_bank ? (kPipeBank ? a.L[a.n_cache].parts : 1u) : 0u);
    if (blockIdx.x < n_srv) {
        const uint32_t p = blockIdx.x;
        if (p < a.n_cache) server_workgroup<RM, WM, false>(a.c[p], a.L[p], lds);
        else if constexpr (kPipeBank)
            pipe_server_workgroup<RB, WB>(a.c[a.n_cache], a.L[a.n_cache], lds, p - a.n_cache);
        else server_workgroup<RB, WB, true>(a.c[p], a.L[p], lds);
        return;
    }
    const uint32_t b = blockIdx.x - n_srv;
    uint32_t p = 0;
    while (p + 1 < n_part && b >= a.wg_end[p]) ++p;
    const uint32_t b0 = p ? a.wg_end[p - 1] : 0u;
    const Ctx &c = a.c[p];
    const uint32_t wave = threadIdx.x / kWave;
    if (wave >= c.client_wg_waves) return;
    const uint32_t cw = (b - b0) * c.client_wg_waves + wave;
    if (cw >= c.n_client_waves) return;
    if (p < a.n_cache) memcd_client<RM, WM>(c, cw);
    else if constexpr (kPipeBank)
        bank_serve_pipe_client<RB, WB>(c, a.L[p], lds + wave * pipe_wide_words(a.L[p].hsz, RB), cw);
    else if (cw >= c.n_client_waves - c.n_serial_waves) bank_serve_serial_client<RB, WB>(c, cw);
    else bank_serve_client<RB, WB>(c, cw);
}

// Same-address agent-scope atomic rate: lane 0 of every wave adds to one word.
__global__ void __launch_bounds__(256) atomic_rate_kernel(uint64_t *p, uint32_t iters)
{
    if (lane_id() != 0) return;
    for (uint32_t i = 0; i < iters; ++i)
        __hip_atomic_fetch_add(p, (uint64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace csmv

// =============================================================================================
// host engine
// =============================================================================================
using namespace csmv;

struct csmv_engine {
    csmv_config cfg{};
    int device = 0;
    uint32_t n_cu = 0;
    uint32_t lds_max = 0;
    uint32_t wallclock_khz = 0;
    uint32_t Q = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // pool
    uint32_t *ht = nullptr;
    uint64_t *slot = nullptr;
    bool pool_loaded = false;
    // control + messaging
    uint8_t *ctl = nullptr;
    uint64_t *ring = nullptr, *reply = nullptr, *done = nullptr;
    uint32_t *mbox = nullptr;
    uint64_t *sv = nullptr;
    uint64_t *reply_v = nullptr;
    uint64_t *sink = nullptr;   // pipelined serialized lane: idle lanes' targets
    // routing of the resident stream (empty = everything on the normal lane)
    uint32_t *norm_list = nullptr, *hot_list = nullptr;
    uint64_t n_norm = 0, n_hot = 0;
    bool routed = false;
    // memcached-style cache stream and log
    csmv_memcd_spec mspec{};
    bool has_mc = false;
    uint64_t mc_cap = 0;
    uint64_t *mc_key = nullptr;
    uint32_t *mc_op = nullptr, *mc_cdf = nullptr;
    uint32_t *mc_ts = nullptr, *mc_res = nullptr, *mc_vsum = nullptr, *mc_ab = nullptr,
             *mc_lat = nullptr;
    // bank stream + log
    csmv_bank_spec spec{};
    bool has_stream = false;
    uint32_t rmax = 0, log_rows = 0;
    uint64_t stream_cap = 0;
    uint32_t *op = nullptr, *addr = nullptr, *cdf = nullptr;
    uint32_t *log_ts = nullptr, *log_ab = nullptr, *log_lat = nullptr;
    int32_t *log_v = nullptr;
    // serving loop: arrival ticks of the stream this engine serves, and its class
    uint32_t *arr = nullptr;
    uint64_t arr_cap = 0;
    uint32_t served = 0;   // 0 none, 1 cache requests, 2 bank transactions
    bool bank_logged = false;   // log_* hold a completed run of the resident stream
    int32_t *tmp = nullptr;
    // learned contention predictor (allocated on first use)
    float *mlp_params = nullptr, *mlp_m = nullptr, *mlp_v = nullptr, *mlp_grad = nullptr,
          *mlp_part = nullptr, *mlp_loss = nullptr;
    uint16_t *mlp_w1t = nullptr;
    uint32_t mlp_step = 0, mlp_loss_cap = 0;
    uint32_t *abort_ctr = nullptr;
    uint16_t *ds_x = nullptr;
    float *ds_y = nullptr;
    uint64_t ds_n = 0, ds_cap = 0;
    uint16_t *io_x = nullptr;
    float *io_y = nullptr, *io_p = nullptr;
    uint64_t io_cap = 0;
    // per-transaction scheduler buffers and the routing partition
    float *pred_p = nullptr;
    uint8_t *pred_flag = nullptr;
    uint32_t *part_cnt = nullptr, *scr_a = nullptr, *scr_b = nullptr;
    uint64_t *part_total = nullptr;
    uint64_t pred_cap = 0, route_cap = 0;
    // request mailbox rows per client wave (the widest wire the engine serves)
    uint32_t mbox_rows = 0;
    // a caller-supplied kernel prepared by csmv_kernel_prepare and not yet finished
    ServerLayout ext_layout{};
    bool ext_prepared = false;
    std::string err;
};

namespace {

constexpr size_t kCtlBytes = 4096;
constexpr size_t kOffGc = 0, kOffErr = 128, kOffDone = 256, kOffNext = 384, kOffTail = 512,
                 kOffStats = 640, kOffProf = 1024, kOffNextHot = 1536,
                 // serving loop: start tick, in-flight version bytes, [stalls, peak]
                 kOffServeStart = 2048, kOffInflight = 2176, kOffServeCtr = 2304,
                 // pipelined serialized lane: each commit partition's published timestamp
                 kOffPipeProg = 2432;

csmv_status fail(csmv_engine *e, csmv_status s, const std::string &msg)
{
    if (e) e->err = msg;
    return s;
}

#define HIPCHK(e, call)                                                                     \
    do {                                                                                    \
        hipError_t _r = (call);                                                             \
        if (_r != hipSuccess)                                                               \
            return fail((e), CSMV_E_HIP,                                                    \
                        std::string(#call) + ": " + hipGetErrorString(_r));                 \
    } while (0)

template <typename T> csmv_status dalloc(csmv_engine *e, T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t r = hipMalloc((void **)p, sizeof(T) * count);
    if (r != hipSuccess)
        return fail(e, r == hipErrorOutOfMemory ? CSMV_E_CAPACITY : CSMV_E_HIP,
                    std::string("hipMalloc: ") + hipGetErrorString(r));
    return CSMV_OK;
}

template <typename T> void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

// Applies the configuration defaults and checks the ranges (CSMV_E_ARG when out of range).
csmv_status normalize_config(csmv_config &cfg)
{
    if (cfg.timeout_ms == 0) cfg.timeout_ms = 4000;
    if (cfg.max_attempts == 0) cfg.max_attempts = 1u << 20;
    if (cfg.wg_waves == 0) cfg.wg_waves = 8;
    if (cfg.stage_ahead == 0) cfg.stage_ahead = 3;
    if (cfg.client_wg_waves == 0) cfg.client_wg_waves = cfg.wg_waves;
    if (cfg.commit_parts == 0) cfg.commit_parts = 1;
    if (cfg.commit_parts > kPipeParts || (cfg.commit_parts > 1 && !(cfg.variant & kSerialPipe))) return CSMV_E_ARG;
    if (cfg.n_serial_waves >= cfg.client_waves && cfg.n_serial_waves) return CSMV_E_ARG;
    if (cfg.n_items == 0 || cfg.depth == 0 || cfg.depth > 64 || cfg.client_waves == 0 ||
        cfg.wg_waves < 3 || cfg.wg_waves > ((cfg.variant & kSerialPipe) ? 16u : 8u) ||
        ((cfg.variant & kSerialPipe) && cfg.client_waves > kPipeCwMask + 1) ||
        cfg.client_wg_waves > cfg.wg_waves ||
        cfg.stage_ahead > 64 || cfg.atr_entries < 64 ||
        (cfg.atr_entries & (cfg.atr_entries - 1)) || cfg.atr_entries > (1u << 16) ||
        (cfg.variant & ~0x1Fu) || (cfg.variant & 3u) == 3u ||
        ((cfg.variant & kSerialPipe) && (cfg.variant & 0xFu) != (kValStamp | kInsertBatch)) ||
        (cfg.escalate_after && cfg.variant != (kValStamp | kInsertBatch)))
        return CSMV_E_ARG;
    return CSMV_OK;
}

uint32_t pow2_at_least(uint64_t v)
{
    uint32_t q = 1;
    while (q < v) q <<= 1;
    return q;
}

// Both routing lists hold up to T transactions.
csmv_status route_reserve(csmv_engine *e, uint64_t T)
{
    if (T <= e->route_cap) return CSMV_OK;
    dfree(e->norm_list);
    dfree(e->hot_list);
    e->route_cap = 0;
    csmv_status st;
    if ((st = dalloc(e, &e->norm_list, T)) || (st = dalloc(e, &e->hot_list, T))) return st;
    e->route_cap = T;
    return CSMV_OK;
}

uint32_t pow2_floor(uint64_t v)
{
    uint32_t q = 1;
    while ((uint64_t)q * 2 <= v) q <<= 1;
    return q;
}

// Server LDS layout for a request wire of R reads and W writes.
ServerLayout server_layout_at(const csmv_config &cfg, uint32_t lds_max, uint32_t R, uint32_t W,
                              bool delegation, uint32_t nmb, uint32_t hshift)
{
    ServerLayout L{};
    L.parts = 1;
    const uint32_t vmode = cfg.variant & 3u;
    if (delegation && (cfg.variant & kSerialPipe)) {
        // pipelined serialized lane: the receivers copy the entry lists the clients staged
        // (no receiver hash or chain scratch, no stamp buckets); each committer's words per
        // entry, and the install records the publisher takes.  The clients' staging scratch
        // lives in their own workgroups' LDS, the same size (a wide wire's clients also
        // keep their read sets there).
        L.nmb = nmb ? nmb : 16;
        L.mb_rows = kPipeBox / kWave;
        L.state_off = kLdsHdr;
        L.wave_off = L.state_off + L.nmb;
        L.tg_off = L.wave_off + L.nmb;
        L.mb_off = (L.tg_off + L.nmb + 63) & ~63u;
        uint32_t w = L.mb_off + L.nmb * L.mb_rows * kWave;
        L.hsz = pow2_at_least(kWave * 2 * W) << 1;
        L.harr = 0;
        L.hash_off = L.sup_off = w;
        L.jmp_off = w;
        w += pipe_committers(cfg.wg_waves) * kPipeMaxE;
        L.pipe_off = w;
        w += pipe_words();
        L.atr_off = L.bucket_off = L.val_off = w;
        L.words = std::max<uint32_t>(w, cfg.client_wg_waves * (R > 16 ? pipe_wide_words(L.hsz, R)
                                                                      : pipe_client_words(R, W, kPipeSlots)));
        L.parts = cfg.commit_parts;
        return L;
    }
    const uint32_t M = cfg.wg_waves - 2;
    // enough mailboxes for the receivers and for the staging window
    L.nmb = nmb ? nmb : std::min<uint32_t>(std::max<uint32_t>({M + 2, cfg.stage_ahead + 2, 4u}), 16);
    L.mb_rows = mailbox_rows(R, W, delegation);
    L.state_off = kLdsHdr;
    L.wave_off = L.state_off + L.nmb;
    L.tg_off = L.wave_off + L.nmb;
    L.mb_off = (L.tg_off + L.nmb + 63) & ~63u;
    uint32_t w = L.mb_off + L.nmb * L.mb_rows * kWave;
    // the writes of a request fill at most 1/2 >> hshift of a receiver's hash
    // collaborative validation hashes every item of a request (lanes touching it: load at
    // most 3/4, or 3/8 when the next shift fits), the other modes its writes (lanes
    // writing it: load at most 1/2)
    const bool collab = vmode == kValCollab;
    L.hsz = pow2_at_least(kWave * (collab ? R + W : 2 * W)) << hshift;
    // a wide wire's receivers also keep a bitmap of the request's written items
    L.harr = (collab ? 5 : 3) + (R > 16 ? 1 : 0);
    L.hash_off = w;
    w += M * L.harr * L.hsz;
    // serialized lane: per-receiver chain scratch (next, weight per chain node; superseded
    // flags per write slot) and the committer's exchange of base and new values
    const uint32_t RO = chain_reads(R, W);
    L.sup_off = w;
    if (delegation) w += M * (2 * kWave * RO + kWave * W);
    L.jmp_off = w;
    if (delegation) w += kWave * RO + kWave * W;
    L.atr_off = w;
    if (vmode != kValStamp) w += W * cfg.atr_entries;
    L.bucket_off = w;
    L.val_off = w;
    L.nbuckets = 0;
    L.exact = 0;
    if (vmode == kValStamp) {
        const uint64_t budget = lds_max / 4;
        const uint64_t room = budget > w ? budget - w : 0;
        if (2ull * pow2_at_least(cfg.n_items) <= room) {
            // exact: stamp and value of every item
            L.exact = 1;
            L.nbuckets = pow2_at_least(cfg.n_items);
            L.val_off = L.bucket_off + L.nbuckets;
            w += L.nbuckets + cfg.n_items;
        } else {
            // hashed stamps; exact stamps and values stay in HBM
            L.nbuckets = pow2_floor(room);
            w += L.nbuckets;
        }
    }
    L.words = w;
    return L;
}

// A wave probes its receiver hash until its longest probe sequence ends, so the hash is
// kept sparse (load at most 1/8, else 1/4, else 1/2) as long as the stamp buckets keep
// 8192 entries and the layout fits.
ServerLayout server_layout_cfg(const csmv_config &cfg, uint32_t lds_max, uint32_t R, uint32_t W,
                               bool delegation, uint32_t nmb = 0)
{
    const bool stamp = (cfg.variant & 3u) == kValStamp;
    if (delegation && (cfg.variant & kSerialPipe))
        return server_layout_at(cfg, lds_max, R, W, delegation, nmb, 0);
    for (uint32_t hshift = 2; hshift > 0; --hshift) {
        const ServerLayout L = server_layout_at(cfg, lds_max, R, W, delegation, nmb, hshift);
        if ((uint64_t)L.words * 4 <= lds_max && (!stamp || L.exact || L.nbuckets >= 8192)) return L;
    }
    return server_layout_at(cfg, lds_max, R, W, delegation, nmb, 0);
}

ServerLayout server_layout_rt(const csmv_engine *e, uint32_t R, uint32_t W, bool delegation,
                              uint32_t nmb = 0)
{
    return server_layout_cfg(e->cfg, e->lds_max, R, W, delegation, nmb);
}

template <int R, int W>
ServerLayout server_layout(const csmv_engine *e, bool delegation, uint32_t nmb = 0)
{
    return server_layout_rt(e, (uint32_t)R, (uint32_t)W, delegation, nmb);
}

}  // namespace

extern "C" {

const char *csmv_status_str(csmv_status s)
{
    switch (s) {
    case CSMV_OK: return "ok";
    case CSMV_E_HIP: return "HIP runtime error";
    case CSMV_E_ARG: return "invalid argument";
    case CSMV_E_CAPACITY: return "capacity exceeded";
    case CSMV_E_TIMEOUT: return "device wait timed out";
    case CSMV_E_NODEVICE: return "no gfx950 device";
    case CSMV_E_STATE: return "call out of order";
    case CSMV_E_LIVELOCK: return "transaction attempt limit exceeded";
    case CSMV_E_IO: return "log file I/O failed or file malformed";
    }
    return "unknown status";
}

const char *csmv_last_error(const csmv_engine *e) { return e ? e->err.c_str() : ""; }

csmv_status csmv_zipf_cdf(uint32_t n, double theta, uint32_t *cdf)
{
    if (n == 0 || !cdf || !(theta >= 0.0)) return CSMV_E_ARG;
    double *h = new (std::nothrow) double[n];
    if (!h) return CSMV_E_CAPACITY;
    double acc = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        acc += std::pow((double)(i + 1), -theta);
        h[i] = acc;
    }
    const double total = h[n - 1];
    for (uint32_t i = 0; i + 1 < n; ++i) {
        double v = std::floor(2147483648.0 * (h[i] / total));
        if (v > 2147483648.0) v = 2147483648.0;
        cdf[i] = (uint32_t)v;
    }
    cdf[n - 1] = 0x80000000u;
    delete[] h;
    return CSMV_OK;
}

csmv_status csmv_create(const csmv_config *cfg_in, csmv_engine **out)
{
    if (!cfg_in || !out) return CSMV_E_ARG;
    *out = nullptr;
    csmv_config cfg = *cfg_in;
    if (normalize_config(cfg)) return CSMV_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CSMV_E_NODEVICE;
    csmv_engine *e = new (std::nothrow) csmv_engine();
    if (!e) return CSMV_E_CAPACITY;
    e->cfg = cfg;
    csmv_status s;
    hipDeviceProp_t prop;
    if (hipGetDevice(&e->device) != hipSuccess ||
        hipGetDeviceProperties(&prop, e->device) != hipSuccess) {
        delete e;
        return CSMV_E_HIP;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        delete e;
        return CSMV_E_NODEVICE;
    }
    e->n_cu = (uint32_t)prop.multiProcessorCount;
    e->lds_max = (uint32_t)prop.sharedMemPerBlock;
    if (cfg.lds_kib) e->lds_max = std::min<uint32_t>(e->lds_max, cfg.lds_kib * 1024u);
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, e->device) != hipSuccess ||
        khz <= 0)
        khz = 100000;
    e->wallclock_khz = (uint32_t)khz;
    e->Q = pow2_at_least(std::max<uint64_t>(4096, 2ull * cfg.client_waves));
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess) {
        csmv_destroy(e);
        return CSMV_E_HIP;
    }
    // request mailboxes sized for the widest wire of the drivers
    const uint32_t rows = std::max({mbox_rows<kBankR, kBankW>(), mbox_rows<kMcR, kMcW>(),
                                    mbox_rows<kServeBankR, kBankW>(), kPipeBoxRows});
    static_assert(kMcW <= kBankW, "reply values are sized for the bank wire");
    e->mbox_rows = rows;
    if ((s = dalloc(e, &e->ht, cfg.n_items)) ||
        (s = dalloc(e, &e->slot, (size_t)cfg.n_items * cfg.depth)) ||
        (s = dalloc(e, &e->ctl, kCtlBytes)) || (s = dalloc(e, &e->ring, e->Q)) ||
        (s = dalloc(e, &e->done, e->Q)) ||
        (s = dalloc(e, &e->reply, (size_t)cfg.client_waves * kWave)) ||
        (s = dalloc(e, &e->mbox, (size_t)cfg.client_waves * rows * kWave)) ||
        (s = dalloc(e, &e->sv, cfg.n_items)) || (s = dalloc(e, &e->tmp, cfg.n_items)) ||
        (s = dalloc(e, &e->sink, kSinkWords)) ||
        (s = dalloc(e, &e->reply_v, (size_t)cfg.client_waves * kReplyVWords))) {
        csmv_destroy(e);
        return s;
    }
    *out = e;
    return CSMV_OK;
}

csmv_status csmv_destroy(csmv_engine *e)
{
    if (!e) return CSMV_E_ARG;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    dfree(e->ht); dfree(e->slot); dfree(e->ctl); dfree(e->ring); dfree(e->done);
    dfree(e->reply); dfree(e->mbox); dfree(e->sv); dfree(e->tmp); dfree(e->sink);
    dfree(e->reply_v); dfree(e->norm_list); dfree(e->hot_list);
    dfree(e->mc_key); dfree(e->mc_op); dfree(e->mc_cdf); dfree(e->mc_ts); dfree(e->mc_res);
    dfree(e->mc_vsum); dfree(e->mc_ab); dfree(e->mc_lat);
    dfree(e->op); dfree(e->addr); dfree(e->cdf); dfree(e->log_ts); dfree(e->log_v);
    dfree(e->log_ab); dfree(e->log_lat); dfree(e->arr);
    dfree(e->mlp_params); dfree(e->mlp_m); dfree(e->mlp_v); dfree(e->mlp_grad);
    dfree(e->mlp_part); dfree(e->mlp_loss); dfree(e->mlp_w1t); dfree(e->abort_ctr);
    dfree(e->ds_x); dfree(e->ds_y); dfree(e->io_x); dfree(e->io_y); dfree(e->io_p);
    dfree(e->pred_p); dfree(e->pred_flag); dfree(e->part_cnt); dfree(e->scr_a);
    dfree(e->scr_b); dfree(e->part_total);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return CSMV_OK;
}

static csmv_status fill_pool(csmv_engine *e, const int32_t *dvals, int32_t v)
{
    const uint32_t n = e->cfg.n_items;
    hipLaunchKernelGGL(pool_fill_kernel, dim3((n + 255) / 256), dim3(256), 0, e->stream, e->ht,
                       e->slot, e->sv, n, e->cfg.depth, dvals, v);
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipStreamSynchronize(e->stream));
    e->pool_loaded = true;
    return CSMV_OK;
}

csmv_status csmv_load_i32(csmv_engine *e, const int32_t *values, size_t n)
{
    if (!e || !values || n != e->cfg.n_items) return fail(e, CSMV_E_ARG, "load: size mismatch");
    HIPCHK(e, hipMemcpyAsync(e->tmp, values, n * 4, hipMemcpyHostToDevice, e->stream));
    return fill_pool(e, e->tmp, 0);
}

csmv_status csmv_fill_i32(csmv_engine *e, int32_t value)
{
    if (!e) return CSMV_E_ARG;
    return fill_pool(e, nullptr, value);
}

csmv_status csmv_read_state_i32(csmv_engine *e, int32_t *out, size_t n)
{
    if (!e || !out || n != e->cfg.n_items) return fail(e, CSMV_E_ARG, "read_state: size mismatch");
    hipLaunchKernelGGL(pool_head_kernel, dim3((n + 255) / 256), dim3(256), 0, e->stream, e->ht,
                       e->slot, (uint32_t)n, e->cfg.depth, e->tmp);
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipMemcpyAsync(out, e->tmp, n * 4, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return CSMV_OK;
}

csmv_status csmv_bank_generate(csmv_engine *e, const csmv_bank_spec *s)
{
    if (!e || !s) return CSMV_E_ARG;
    if (s->n_txns == 0 || s->n_txns >= (1ull << 31) || s->n_reads < 2 ||
        s->n_reads > (uint32_t)kServeBankR || s->n_reads_ro < 1 || s->n_reads_ro > 64 ||
        s->pct_ro > 100 || s->dist > 1 || s->amount_max == 0 || s->amount_max > 255 ||
        (s->n_reads_min && (s->n_reads_min < 2 || s->n_reads_min > s->n_reads)) ||
        (s->dist == 1 && !(s->theta >= 0.0)))
        return fail(e, CSMV_E_ARG, "bank spec out of range");
    const uint32_t rmax = std::max(s->n_reads, s->n_reads_ro);
    if (s->n_txns > e->stream_cap || rmax > e->rmax || s->n_reads > e->log_rows) {
        dfree(e->op); dfree(e->addr); dfree(e->log_ts); dfree(e->log_v); dfree(e->log_ab);
        dfree(e->log_lat);
        e->stream_cap = 0;
        csmv_status st;
        if ((st = dalloc(e, &e->op, s->n_txns)) || (st = dalloc(e, &e->addr, rmax * s->n_txns)) ||
            (st = dalloc(e, &e->log_ts, s->n_txns)) ||
            (st = dalloc(e, &e->log_v, (size_t)s->n_reads * s->n_txns)) ||
            (st = dalloc(e, &e->log_ab, s->n_txns)) || (st = dalloc(e, &e->log_lat, s->n_txns)))
            return st;
        e->stream_cap = s->n_txns;
        e->rmax = rmax;
        e->log_rows = s->n_reads;
    }
    dfree(e->cdf);
    if (s->dist == 1) {
        uint32_t *h = new (std::nothrow) uint32_t[e->cfg.n_items];
        if (!h) return fail(e, CSMV_E_CAPACITY, "cdf host buffer");
        csmv_zipf_cdf(e->cfg.n_items, s->theta, h);
        csmv_status st = dalloc(e, &e->cdf, e->cfg.n_items);
        if (st) { delete[] h; return st; }
        hipError_t r = hipMemcpy(e->cdf, h, e->cfg.n_items * 4, hipMemcpyHostToDevice);
        delete[] h;
        HIPCHK(e, r);
    }
    GenArgs g{};
    g.seed = s->seed;
    g.n_txns = s->n_txns;
    g.n_items = e->cfg.n_items;
    g.n_reads = s->n_reads;
    g.n_reads_ro = s->n_reads_ro;
    g.ro_thr = (uint32_t)((uint64_t)s->pct_ro * 0xFFFFu / 100u);
    g.dist = s->dist;
    g.amount_max = s->amount_max;
    g.rmax = rmax;
    g.n_reads_min = s->n_reads_min;
    g.cdf = e->cdf;
    g.op = e->op;
    g.addr = e->addr;
    hipLaunchKernelGGL(bank_generate_kernel, dim3((unsigned)((s->n_txns + 255) / 256)), dim3(256),
                       0, e->stream, g);
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipStreamSynchronize(e->stream));
    e->spec = *s;
    e->has_stream = true;
    e->bank_logged = false;
    e->routed = false;
    return CSMV_OK;
}

csmv_status csmv_bank_route(csmv_engine *e, const uint8_t *hot, size_t n)
{
    if (!e) return CSMV_E_ARG;
    if (!e->has_stream) return fail(e, CSMV_E_STATE, "route before generate");
    if (!hot) {
        e->routed = false;
        return CSMV_OK;
    }
    const uint64_t T = e->spec.n_txns;
    if (n != T) return fail(e, CSMV_E_ARG, "route: one flag per transaction");
    uint32_t *op = new (std::nothrow) uint32_t[T];
    uint32_t *lists = new (std::nothrow) uint32_t[T];
    if (!op || !lists) { delete[] op; delete[] lists; return fail(e, CSMV_E_CAPACITY, "route buffers"); }
    hipError_t r = hipMemcpy(op, e->op, T * 4, hipMemcpyDeviceToHost);
    if (r != hipSuccess) { delete[] op; delete[] lists; HIPCHK(e, r); }
    // normal lane first (stream order), then the serialized lane (stream order); an audit
    // goes to the serialized lane when its read set fits the lane's wire
    auto serial = [&](uint64_t k) {
        return hot[k] && ((op[k] & 0xFFu) == kKindTransfer || ((op[k] >> 16) & 0xFFu) <= (uint32_t)kBankR);
    };
    uint64_t nn = 0;
    for (uint64_t k = 0; k < T; ++k)
        if (!serial(k)) lists[nn++] = (uint32_t)k;
    uint64_t nh = 0;
    for (uint64_t k = 0; k < T; ++k)
        if (serial(k)) lists[nn + nh++] = (uint32_t)k;
    delete[] op;
    e->routed = false;
    const csmv_status st = route_reserve(e, T);
    if (st) {
        delete[] lists;
        return st;
    }
    if (nn) r = hipMemcpy(e->norm_list, lists, nn * 4, hipMemcpyHostToDevice);
    if (r == hipSuccess && nh) r = hipMemcpy(e->hot_list, lists + nn, nh * 4, hipMemcpyHostToDevice);
    delete[] lists;
    HIPCHK(e, r);
    e->n_norm = nn;
    e->n_hot = nh;
    e->routed = true;
    return CSMV_OK;
}

csmv_status csmv_bank_read_stream(csmv_engine *e, uint32_t *op, uint32_t *addr)
{
    if (!e || !op || !addr) return CSMV_E_ARG;
    if (!e->has_stream) return fail(e, CSMV_E_STATE, "no stream generated");
    const uint64_t T = e->spec.n_txns;
    const uint32_t rmax = std::max(e->spec.n_reads, e->spec.n_reads_ro);
    HIPCHK(e, hipMemcpy(op, e->op, T * 4, hipMemcpyDeviceToHost));
    HIPCHK(e, hipMemcpy(addr, e->addr, (size_t)rmax * T * 4, hipMemcpyDeviceToHost));
    return CSMV_OK;
}

}  // extern "C"

// The configuration fields of a kernel context.
static void ctx_config(Ctx &c, const csmv_config &cfg, uint32_t khz, uint32_t Q)
{
    c.n_items = cfg.n_items;
    c.depth = cfg.depth;
    c.qmask = Q - 1;
    c.n_client_waves = cfg.client_waves;
    c.wg_waves = cfg.wg_waves;
    c.client_wg_waves = cfg.client_wg_waves;
    c.variant = cfg.variant;
    c.atr_entries = cfg.atr_entries;
    c.stage_ahead = cfg.stage_ahead;
    c.n_serial_waves = cfg.n_serial_waves;
    c.begin_wait = cfg.begin_wait;
    c.timeout_ticks = (uint64_t)cfg.timeout_ms * khz;
    c.max_attempts = cfg.max_attempts;
    c.escalate = cfg.escalate_after;
}

// The control-word pointers of a kernel context over a control block.
static void ctx_control(Ctx &c, uint8_t *ctl)
{
    c.gc = (uint32_t *)(ctl + kOffGc);
    c.err = (uint32_t *)(ctl + kOffErr);
    c.clients_done = (uint32_t *)(ctl + kOffDone);
    c.next_txn = (uint64_t *)(ctl + kOffNext);
    c.req_tail = (uint64_t *)(ctl + kOffTail);
    c.stats = (uint64_t *)(ctl + kOffStats);
    c.prof = (uint64_t *)(ctl + kOffProf);
    c.next_hot = (uint64_t *)(ctl + kOffNextHot);
    c.pipe_prog = (uint32_t *)(ctl + kOffPipeProg);
}

// Statistics of a launch from a host copy of its control block.
static void decode_stats(const uint8_t *host_ctl, const ServerLayout &L, float ms,
                         uint32_t client_waves, uint32_t khz, csmv_stats *out)
{
    const uint64_t *st = (const uint64_t *)(host_ctl + kOffStats);
    std::memset(out, 0, sizeof(*out));
    out->commits_update = st[kStatCommitUpd];
    out->commits_ro = st[kStatCommitRo];
    out->aborts_dataage = st[kStatAbortAge];
    out->aborts_readwrite = st[kStatAbortRw];
    out->aborts_prevalid = st[kStatAbortPrevalid];
    out->aborts_recordage = st[kStatAbortRecord];
    out->batches = st[kStatBatches];
    out->final_ts = st[kStatFinalTs];
    out->kernel_ms = ms;
    std::memcpy(&out->error, host_ctl + kOffErr, 4);
    out->client_lanes = client_waves * kWave;
    const uint64_t *pf = (const uint64_t *)(host_ctl + kOffProf);
    const double ns_per_tick = 1e6 / (double)khz;
    for (int i = 0; i < kProfCount; ++i) out->prof_ns[i] = (double)pf[i] * ns_per_tick;
    out->prof_ns[kProfClientRounds] = (double)pf[kProfClientRounds];
    out->prof_ns[kProfCommitMiss] = (double)pf[kProfCommitMiss];
    out->stamp_buckets = L.nbuckets;
    out->stamp_exact = L.exact;
    out->aborts_unknown = st[kStatAbortUnknown];
    out->lds_bytes = (uint32_t)L.words * 4;
    out->escalated = st[kStatEscalated];
}

static void base_ctx(csmv_engine *e, Ctx &c)
{
    ctx_config(c, e->cfg, e->wallclock_khz, e->Q);
    ctx_control(c, e->ctl);
    c.ht = e->ht;
    c.slot = e->slot;
    c.ring = e->ring;
    c.reply = e->reply;
    c.done = e->done;
    c.mbox = e->mbox;
    c.sv = e->sv;
    c.reply_v = e->reply_v;
    c.sink = e->sink;
}

// Clears the control words and message rings of one engine before a launch.
static csmv_status reset_run_state(csmv_engine *e, const ServerLayout &L)
{
    HIPCHK(e, hipMemsetAsync(e->ctl, 0, kCtlBytes, e->stream));
    HIPCHK(e, hipMemsetAsync(e->ring, 0, e->Q * 8, e->stream));
    HIPCHK(e, hipMemsetAsync(e->done, 0, e->Q * 8, e->stream));
    HIPCHK(e, hipMemsetAsync(e->reply, 0, (size_t)e->cfg.client_waves * kWave * 8, e->stream));
    HIPCHK(e, hipMemsetAsync(e->reply_v, 0, (size_t)e->cfg.client_waves * kReplyVWords * 8,
                             e->stream));
    const uint32_t n = e->cfg.n_items;
    if (e->cfg.variant & kSerialPipe) {
        hipLaunchKernelGGL(pipe_init_kernel, dim3((n + 255) / 256), dim3(256), 0, e->stream, e->ht,
                           e->slot, e->sv, n, e->cfg.depth);
        HIPCHK(e, hipGetLastError());
    } else if ((e->cfg.variant & 3u) == kValStamp && !L.exact) {
        hipLaunchKernelGGL(stamp_clear_kernel, dim3((n + 255) / 256), dim3(256), 0, e->stream, e->sv, n);
        HIPCHK(e, hipGetLastError());
    }
    return CSMV_OK;
}

// Reads one engine's counters after a launch; returns the device error word.
static csmv_status collect_stats(csmv_engine *e, const ServerLayout &L, float ms, csmv_stats *out,
                                 uint32_t *derr)
{
    uint8_t host_ctl[kCtlBytes];
    HIPCHK(e, hipMemcpy(host_ctl, e->ctl, kCtlBytes, hipMemcpyDeviceToHost));
    std::memcpy(derr, host_ctl + kOffErr, 4);
    if (out) decode_stats(host_ctl, L, ms, e->cfg.client_waves, e->wallclock_khz, out);
    return CSMV_OK;
}

static csmv_status device_error(csmv_engine *e, uint32_t derr)
{
    if (derr == kErrLivelock) return fail(e, CSMV_E_LIVELOCK, "transaction attempt limit");
    if (derr == kErrReadSetOverflow) return fail(e, CSMV_E_CAPACITY, "read/write set overflow");
    if (derr == kErrPipeNormal)
        return fail(e, CSMV_E_ARG, "a normal-lane update reached the pipelined serialized lane "
                                   "(route every transfer to it)");
    if (derr) return fail(e, CSMV_E_TIMEOUT, "device wait timed out, code " + std::to_string(derr));
    return CSMV_OK;
}

// Launches one persistent kernel (server workgroup 0 + client workgroups) over the
// prepared context and collects its statistics.
template <typename K>
static csmv_status launch_persistent(csmv_engine *e, K kern, const Ctx &c, const ServerLayout &L,
                                     csmv_stats *out, const std::function<void()> &after = {})
{
    const size_t lds_bytes = (size_t)L.words * 4;
    if (lds_bytes > e->lds_max)
        return fail(e, CSMV_E_CAPACITY, "server LDS layout of " + std::to_string(lds_bytes) +
                                            " bytes exceeds the budget (reduce atr_entries)");
    const uint32_t block = e->cfg.wg_waves * kWave;
    const uint32_t cpw = e->cfg.client_wg_waves;
    const uint32_t client_wgs = (e->cfg.client_waves + cpw - 1) / cpw;
    const uint32_t grid = L.parts + client_wgs;
    HIPCHK(e, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds_bytes));
    int per_cu = 0;
    HIPCHK(e, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, (int)block, lds_bytes));
    if ((uint64_t)per_cu * e->n_cu < grid)
        return fail(e, CSMV_E_CAPACITY,
                    "grid of " + std::to_string(grid) + " workgroups exceeds the " +
                        std::to_string((uint64_t)per_cu * e->n_cu) + " the device keeps resident");
    csmv_status s = reset_run_state(e, L);
    if (s) return s;
    HIPCHK(e, hipStreamSynchronize(e->stream));

    HIPCHK(e, hipEventRecord(e->ev0, e->stream));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds_bytes, e->stream, c, L);
    HIPCHK(e, hipGetLastError());
    // work that completes the run (inside the timed region)
    if (after) {
        after();
        HIPCHK(e, hipGetLastError());
    }
    HIPCHK(e, hipEventRecord(e->ev1, e->stream));
    HIPCHK(e, hipEventSynchronize(e->ev1));
    float ms = 0.f;
    HIPCHK(e, hipEventElapsedTime(&ms, e->ev0, e->ev1));
    e->pool_loaded = false;
    uint32_t derr = 0;
    if ((s = collect_stats(e, L, ms, out, &derr))) return s;
    return device_error(e, derr);
}

extern "C" {

csmv_status csmv_bank_run(csmv_engine *e, csmv_stats *out)
{
    if (!e) return CSMV_E_ARG;
    if (!e->has_stream) return fail(e, CSMV_E_STATE, "run before generate");
    if (!e->pool_loaded) return fail(e, CSMV_E_STATE, "run needs a freshly loaded pool");
    if (e->spec.n_reads > (uint32_t)kBankR)
        return fail(e, CSMV_E_ARG, "transfers of more than 4 reads run in the serving loop");
    const ServerLayout L = server_layout<kBankR, kBankW>(e, true);
    Ctx c{};
    base_ctx(e, c);
    c.op = e->op;
    c.addr = e->addr;
    c.n_txns = e->spec.n_txns;
    c.log_ts = e->log_ts;
    c.log_v = e->log_v;
    c.log_ab = e->log_ab;
    c.norm_list = e->routed ? e->norm_list : nullptr;
    c.hot_list = e->routed ? e->hot_list : nullptr;
    c.n_norm = e->routed ? e->n_norm : e->spec.n_txns;
    c.n_hot = e->routed ? e->n_hot : 0;
    c.pipe_parts = (e->cfg.variant & kSerialPipe) ? e->cfg.commit_parts : 0;
    c.pipe_heads = c.n_norm != 0;
    if (c.n_hot && (c.n_serial_waves == 0 || (e->cfg.variant & 3u) != kValStamp ||
                    (e->cfg.variant & kServerWriteback)))
        return fail(e, CSMV_E_ARG, "the serialized lane needs n_serial_waves > 0 and the "
                                   "STAMP variant with client write-back");
    if ((e->cfg.variant & kSerialPipe) && (c.n_hot == 0 || c.n_serial_waves == 0))
        return fail(e, CSMV_E_ARG, "the pipelined serialized lane needs routed transfers and "
                                   "n_serial_waves > 0");
    if ((e->cfg.variant & kSerialPipe) &&
        (uint64_t)e->cfg.client_wg_waves * pipe_client_words(kBankR, kBankW, kPipeSlots) * 4 > e->lds_max)
        return fail(e, CSMV_E_CAPACITY,
                    "the pipelined lane's client waves stage their requests in " +
                        std::to_string(pipe_client_words(kBankR, kBankW, kPipeSlots) * 4 / 1024) +
                        " KiB of LDS each: client_wg_waves of " +
                        std::to_string(e->cfg.client_wg_waves) + " do not fit (use 1)");
    csmv_status st;
    if ((e->cfg.variant & kSerialPipe) && c.n_norm == 0) {
        // no snapshot reader during the run: the head words follow from the install counts
        const uint32_t n = e->cfg.n_items;
        st = launch_persistent(e, bank_pipe_kernel<kBankR, kBankW>, c, L, out, [&]() {
            hipLaunchKernelGGL(pipe_heads_kernel, dim3((n + 255) / 256), dim3(256), 0, e->stream,
                               e->sv, e->ht, n, e->cfg.depth);
        });
    } else if (e->cfg.variant & kSerialPipe) {
        st = launch_persistent(e, bank_pipe_kernel<kBankR, kBankW>, c, L, out);
    } else {
        st = launch_persistent(e, bank_kernel<kBankR, kBankW>, c, L, out);
    }
    e->bank_logged = st == CSMV_OK;
    return st;
}

// ---- memcached-style cache ---------------------------------------------------------------
}  // extern "C"

// Device arrays of a cache request stream of T requests and its log.
static csmv_status mc_reserve(csmv_engine *e, uint64_t T)
{
    if (T <= e->mc_cap) return CSMV_OK;
    dfree(e->mc_key); dfree(e->mc_op); dfree(e->mc_ts); dfree(e->mc_res); dfree(e->mc_vsum);
    dfree(e->mc_ab); dfree(e->mc_lat);
    e->mc_cap = 0;
    csmv_status st;
    if ((st = dalloc(e, &e->mc_key, T)) || (st = dalloc(e, &e->mc_op, T)) ||
        (st = dalloc(e, &e->mc_ts, T)) || (st = dalloc(e, &e->mc_res, T)) ||
        (st = dalloc(e, &e->mc_vsum, T)) || (st = dalloc(e, &e->mc_ab, T)) ||
        (st = dalloc(e, &e->mc_lat, T)))
        return st;
    e->mc_cap = T;
    return CSMV_OK;
}

extern "C" {

csmv_status csmv_memcd_load(csmv_engine *e, const csmv_memcd_spec *s, const uint64_t *key,
                            const uint32_t *op)
{
    if (!e || !s || !key || !op) return CSMV_E_ARG;
    if (s->n_reqs == 0 || s->n_reqs >= (1ull << 31) || s->n_sets == 0 ||
        (s->n_sets & (s->n_sets - 1)) || s->n_ways == 0 || s->n_ways > 8 || s->batch == 0)
        return fail(e, CSMV_E_ARG, "memcd spec out of range");
    if ((uint64_t)kMcFields * s->n_sets * s->n_ways != e->cfg.n_items)
        return fail(e, CSMV_E_ARG, "memcd: the engine needs n_items = 12 * n_sets * n_ways");
    for (uint64_t k = 0; k < s->n_reqs; ++k)
        if (op[k] != kMcGet && op[k] != kMcSet) return fail(e, CSMV_E_ARG, "memcd: request op not GET or SET");
    csmv_status st = mc_reserve(e, s->n_reqs);
    if (st) return st;
    HIPCHK(e, hipMemcpy(e->mc_key, key, s->n_reqs * 8, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(e->mc_op, op, s->n_reqs * 4, hipMemcpyHostToDevice));
    e->mspec = *s;
    e->has_mc = true;
    return CSMV_OK;
}

csmv_status csmv_memcd_generate(csmv_engine *e, const csmv_memcd_spec *s)
{
    if (!e || !s) return CSMV_E_ARG;
    if (s->n_reqs == 0 || s->n_reqs >= (1ull << 31) || s->n_keys == 0 ||
        s->n_keys >= (1ull << 31) || s->n_sets == 0 || (s->n_sets & (s->n_sets - 1)) ||
        s->n_ways == 0 || s->n_ways > 8 || s->pct_get > 100 || s->batch == 0 || s->dist > 1 ||
        (s->dist == 1 && !(s->theta >= 0.0)))
        return fail(e, CSMV_E_ARG, "memcd spec out of range");
    if ((uint64_t)kMcFields * s->n_sets * s->n_ways != e->cfg.n_items)
        return fail(e, CSMV_E_ARG, "memcd: the engine needs n_items = 12 * n_sets * n_ways");
    const uint64_t T = s->n_reqs;
    csmv_status rs = mc_reserve(e, T);
    if (rs) return rs;
    dfree(e->mc_cdf);
    if (s->dist == 1) {
        uint32_t *h = new (std::nothrow) uint32_t[s->n_keys];
        if (!h) return fail(e, CSMV_E_CAPACITY, "cdf host buffer");
        csmv_zipf_cdf((uint32_t)s->n_keys, s->theta, h);
        csmv_status st = dalloc(e, &e->mc_cdf, s->n_keys);
        if (st) { delete[] h; return st; }
        hipError_t r = hipMemcpy(e->mc_cdf, h, s->n_keys * 4, hipMemcpyHostToDevice);
        delete[] h;
        HIPCHK(e, r);
    }
    McGenArgs g{};
    g.seed = s->seed;
    g.n_reqs = T;
    g.n_keys = s->n_keys;
    g.get_thr = (uint32_t)((uint64_t)s->pct_get * 0xFFFFu / 100u);
    g.dist = s->dist;
    g.cdf = e->mc_cdf;
    g.key = e->mc_key;
    g.op = e->mc_op;
    hipLaunchKernelGGL(memcd_generate_kernel, dim3((unsigned)((T + 255) / 256)), dim3(256), 0,
                       e->stream, g);
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipStreamSynchronize(e->stream));
    e->mspec = *s;
    e->has_mc = true;
    return CSMV_OK;
}

csmv_status csmv_memcd_read_stream(csmv_engine *e, uint64_t *key, uint32_t *op)
{
    if (!e || !key || !op) return CSMV_E_ARG;
    if (!e->has_mc) return fail(e, CSMV_E_STATE, "no memcd stream");
    HIPCHK(e, hipMemcpy(key, e->mc_key, e->mspec.n_reqs * 8, hipMemcpyDeviceToHost));
    HIPCHK(e, hipMemcpy(op, e->mc_op, e->mspec.n_reqs * 4, hipMemcpyDeviceToHost));
    return CSMV_OK;
}

csmv_status csmv_memcd_run(csmv_engine *e, csmv_stats *out)
{
    if (!e) return CSMV_E_ARG;
    if (!e->has_mc) return fail(e, CSMV_E_STATE, "run before generate");
    if (!e->pool_loaded) return fail(e, CSMV_E_STATE, "run needs a freshly loaded pool");
    if ((e->cfg.variant & 3u) != kValStamp || (e->cfg.variant & (kServerWriteback | kSerialPipe)))
        return fail(e, CSMV_E_ARG, "the cache driver runs the STAMP variant with client write-back");
    const ServerLayout L = server_layout<kMcR, kMcW>(e, false);
    Ctx c{};
    base_ctx(e, c);
    c.n_serial_waves = 0;
    c.n_txns = e->mspec.n_reqs;
    c.req_key = e->mc_key;
    c.req_op = e->mc_op;
    c.n_sets = e->mspec.n_sets;
    c.n_ways = e->mspec.n_ways;
    c.n_slots = e->mspec.n_sets * e->mspec.n_ways;
    c.mc_batch = e->mspec.batch;
    c.log_ts = e->mc_ts;
    c.log_res = e->mc_res;
    c.log_vsum = e->mc_vsum;
    c.log_ab = e->mc_ab;
    c.log_lat = e->mc_lat;
    c.n_norm = c.n_txns;
    return launch_persistent(e, memcd_kernel<kMcR, kMcW>, c, L, out);
}

csmv_status csmv_memcd_read_log(csmv_engine *e, uint32_t *ts, uint32_t *res, uint32_t *vsum,
                                uint32_t *aborts, uint32_t *lat_ticks)
{
    if (!e) return CSMV_E_ARG;
    if (!e->has_mc) return fail(e, CSMV_E_STATE, "no memcd stream");
    const uint64_t T = e->mspec.n_reqs;
    if (ts) HIPCHK(e, hipMemcpy(ts, e->mc_ts, T * 4, hipMemcpyDeviceToHost));
    if (res) HIPCHK(e, hipMemcpy(res, e->mc_res, T * 4, hipMemcpyDeviceToHost));
    if (vsum) HIPCHK(e, hipMemcpy(vsum, e->mc_vsum, T * 4, hipMemcpyDeviceToHost));
    if (aborts) HIPCHK(e, hipMemcpy(aborts, e->mc_ab, T * 4, hipMemcpyDeviceToHost));
    if (lat_ticks) HIPCHK(e, hipMemcpy(lat_ticks, e->mc_lat, T * 4, hipMemcpyDeviceToHost));
    return CSMV_OK;
}

csmv_status csmv_wallclock_khz(csmv_engine *e, uint32_t *khz)
{
    if (!e || !khz) return CSMV_E_ARG;
    *khz = e->wallclock_khz;
    return CSMV_OK;
}

csmv_status csmv_bank_read_log(csmv_engine *e, uint32_t *ts, int32_t *v, uint32_t *aborts)
{
    if (!e) return CSMV_E_ARG;
    if (!e->has_stream) return fail(e, CSMV_E_STATE, "no stream");
    const uint64_t T = e->spec.n_txns;
    if (ts) HIPCHK(e, hipMemcpy(ts, e->log_ts, T * 4, hipMemcpyDeviceToHost));
    if (v) HIPCHK(e, hipMemcpy(v, e->log_v, (size_t)e->spec.n_reads * T * 4, hipMemcpyDeviceToHost));
    if (aborts) HIPCHK(e, hipMemcpy(aborts, e->log_ab, T * 4, hipMemcpyDeviceToHost));
    return CSMV_OK;
}

csmv_status csmv_kernel_prepare(csmv_engine *e, const csmv_kernel_spec *sp, void *ctx,
                                size_t ctx_bytes, void *layout, size_t layout_bytes,
                                csmv_launch *out)
{
    if (!e) return CSMV_E_ARG;
    if (!sp || !sp->kernel || !ctx || !layout || !out)
        return fail(e, CSMV_E_ARG, "kernel_prepare: null argument");
    if (ctx_bytes != sizeof(Ctx) || layout_bytes != sizeof(ServerLayout))
        return fail(e, CSMV_E_ARG, "kernel_prepare: context / layout buffers must be sizeof(csmv::Ctx) = " +
                                       std::to_string(sizeof(Ctx)) + " and sizeof(csmv::ServerLayout) = " +
                                       std::to_string(sizeof(ServerLayout)) + " bytes");
    const uint32_t R = sp->reads, W = sp->writes;
    if (R == 0 || R > 64 || W == 0 || W > 32 || 1 + R + 2 * W > e->mbox_rows)
        return fail(e, CSMV_E_ARG, "kernel_prepare: wire beyond the request mailboxes");
    if (e->ext_prepared) return fail(e, CSMV_E_STATE, "kernel_prepare: a prepared kernel was not finished");
    if (e->cfg.variant & kSerialPipe)
        return fail(e, CSMV_E_ARG, "kernel_prepare: caller kernels run the normal lane");
    if (!e->pool_loaded) return fail(e, CSMV_E_STATE, "kernel_prepare needs a freshly loaded pool");
    const ServerLayout L = server_layout_rt(e, R, W, false);
    const size_t lds_bytes = (size_t)L.words * 4;
    if (lds_bytes > e->lds_max)
        return fail(e, CSMV_E_CAPACITY, "server LDS layout of " + std::to_string(lds_bytes) +
                                            " bytes exceeds the budget");
    const uint32_t block = e->cfg.wg_waves * kWave;
    const uint32_t cpw = e->cfg.client_wg_waves;
    const uint32_t grid = 1 + (e->cfg.client_waves + cpw - 1) / cpw;
    HIPCHK(e, hipFuncSetAttribute(sp->kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds_bytes));
    int per_cu = 0;
    HIPCHK(e, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sp->kernel, (int)block, lds_bytes));
    if ((uint64_t)per_cu * e->n_cu < grid)
        return fail(e, CSMV_E_CAPACITY, "grid of " + std::to_string(grid) +
                                            " workgroups exceeds what the device keeps resident");
    Ctx c{};
    base_ctx(e, c);
    c.n_serial_waves = 0;
    c.escalate = 0;
    csmv_status s = reset_run_state(e, L);
    if (s) return s;
    HIPCHK(e, hipStreamSynchronize(e->stream));
    std::memcpy(ctx, &c, sizeof(Ctx));
    std::memcpy(layout, &L, sizeof(ServerLayout));
    HIPCHK(e, hipEventRecord(e->ev0, e->stream));
    e->ext_layout = L;
    e->ext_prepared = true;
    out->grid = grid;
    out->block = block;
    out->lds_bytes = (uint32_t)lds_bytes;
    out->reserved = 0;
    out->stream = (void *)e->stream;
    return CSMV_OK;
}

csmv_status csmv_kernel_finish(csmv_engine *e, csmv_stats *out)
{
    if (!e) return CSMV_E_ARG;
    if (!e->ext_prepared) return fail(e, CSMV_E_STATE, "kernel_finish without kernel_prepare");
    e->ext_prepared = false;
    e->pool_loaded = false;
    HIPCHK(e, hipEventRecord(e->ev1, e->stream));
    HIPCHK(e, hipEventSynchronize(e->ev1));
    float ms = 0.f;
    HIPCHK(e, hipEventElapsedTime(&ms, e->ev0, e->ev1));
    uint32_t derr = 0;
    csmv_status s = collect_stats(e, e->ext_layout, ms, out, &derr);
    if (s) return s;
    return device_error(e, derr);
}

csmv_status csmv_atomic_rate(csmv_engine *e, uint32_t waves, uint32_t iters, double *rate)
{
    if (!e || !rate || waves == 0 || iters == 0 || waves > 65536) return CSMV_E_ARG;
    uint64_t *p = (uint64_t *)(e->ctl + kOffStats);
    HIPCHK(e, hipMemsetAsync(e->ctl, 0, kCtlBytes, e->stream));
    const uint32_t blocks = (waves + 3) / 4;
    hipLaunchKernelGGL(atomic_rate_kernel, dim3(blocks), dim3(256), 0, e->stream, p, iters);
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipEventRecord(e->ev0, e->stream));
    hipLaunchKernelGGL(atomic_rate_kernel, dim3(blocks), dim3(256), 0, e->stream, p, iters);
    HIPCHK(e, hipEventRecord(e->ev1, e->stream));
    HIPCHK(e, hipEventSynchronize(e->ev1));
    float ms = 0.f;
    HIPCHK(e, hipEventElapsedTime(&ms, e->ev0, e->ev1));
    *rate = (double)blocks * 4.0 * iters / (ms * 1e-3);
    return CSMV_OK;
}

}  // extern "C"

// =============================================================================================
// learned contention predictor (mlp.hpp): abort-log counters, training set, training,
// inference and the device routing partition
// =============================================================================================
namespace {

constexpr float kAdamBeta1 = 0.9f, kAdamBeta2 = 0.999f, kAdamEps = 1e-8f;

csmv_status pred_ensure(csmv_engine *e)
{
    if (e->mlp_params) return CSMV_OK;
    csmv_status st;
    if ((st = dalloc(e, &e->mlp_params, kParamCount)) || (st = dalloc(e, &e->mlp_m, kParamCount)) ||
        (st = dalloc(e, &e->mlp_v, kParamCount)) || (st = dalloc(e, &e->mlp_grad, kPartStride)) ||
        (st = dalloc(e, &e->mlp_part, (size_t)kMlpGradBlocks * kMlpWaves * kPartStride)) ||
        (st = dalloc(e, &e->mlp_w1t, (size_t)kFeat * kHidden)) ||
        (st = dalloc(e, &e->abort_ctr, kAbortBuckets)) || (st = dalloc(e, &e->part_total, 1))) {
        dfree(e->mlp_params);
        return st;
    }
    HIPCHK(e, hipMemsetAsync(e->mlp_params, 0, kParamCount * 4, e->stream));
    HIPCHK(e, hipMemsetAsync(e->mlp_m, 0, kParamCount * 4, e->stream));
    HIPCHK(e, hipMemsetAsync(e->mlp_v, 0, kParamCount * 4, e->stream));
    HIPCHK(e, hipMemsetAsync(e->mlp_w1t, 0, kFeat * kHidden * 2, e->stream));
    HIPCHK(e, hipMemsetAsync(e->abort_ctr, 0, kAbortBuckets * 4, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return CSMV_OK;
}

csmv_status pred_upload(csmv_engine *e, const float *params)
{
    HIPCHK(e, hipMemcpyAsync(e->mlp_params, params, kParamCount * 4, hipMemcpyHostToDevice,
                             e->stream));
    HIPCHK(e, hipMemsetAsync(e->mlp_m, 0, kParamCount * 4, e->stream));
    HIPCHK(e, hipMemsetAsync(e->mlp_v, 0, kParamCount * 4, e->stream));
    hipLaunchKernelGGL(mlp_sync_kernel, dim3((kFeat * kHidden + 255) / 256), dim3(256), 0,
                       e->stream, e->mlp_params, e->mlp_w1t);
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipStreamSynchronize(e->stream));
    e->mlp_step = 0;
    return CSMV_OK;
}

// per-transaction buffers of the resident stream
csmv_status pred_reserve(csmv_engine *e, uint64_t T)
{
    if (T <= e->pred_cap) return CSMV_OK;
    dfree(e->pred_p); dfree(e->pred_flag); dfree(e->part_cnt); dfree(e->scr_a); dfree(e->scr_b);
    e->pred_cap = 0;
    csmv_status st;
    if ((st = dalloc(e, &e->pred_p, T)) || (st = dalloc(e, &e->pred_flag, T)) ||
        (st = dalloc(e, &e->part_cnt, (T + kPartItems - 1) / kPartItems)) ||
        (st = dalloc(e, &e->scr_a, T)) || (st = dalloc(e, &e->scr_b, T)))
        return st;
    e->pred_cap = T;
    return CSMV_OK;
}

csmv_status io_reserve(csmv_engine *e, uint64_t n)
{
    if (n <= e->io_cap) return CSMV_OK;
    dfree(e->io_x); dfree(e->io_y); dfree(e->io_p);
    e->io_cap = 0;
    csmv_status st;
    if ((st = dalloc(e, &e->io_x, n * kFeat)) || (st = dalloc(e, &e->io_y, n)) ||
        (st = dalloc(e, &e->io_p, n)))
        return st;
    e->io_cap = n;
    return CSMV_OK;
}

// Stable partition of [0, T) by pred_flag into (hot, rest); returns the hot count.
csmv_status partition(csmv_engine *e, uint64_t T, uint32_t *hot, uint32_t *rest, uint64_t *n_hot)
{
    const uint32_t nb = (uint32_t)((T + kPartItems - 1) / kPartItems);
    hipLaunchKernelGGL(part_count_kernel, dim3(nb), dim3(256), 0, e->stream, e->pred_flag, T,
                       e->part_cnt);
    hipLaunchKernelGGL(part_scan_kernel, dim3(1), dim3(1024), 0, e->stream, e->part_cnt, nb,
                       e->part_total);
    hipLaunchKernelGGL(part_scatter_kernel, dim3(nb), dim3(256), 0, e->stream, e->pred_flag, T,
                       e->part_cnt, hot, rest);
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipMemcpyAsync(n_hot, e->part_total, 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return CSMV_OK;
}

__global__ void __launch_bounds__(256) transfer_flag_kernel(const uint32_t *op, uint64_t T,
                                                            uint8_t *flag)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < T) flag[k] = (op[k] & 0xFFu) == kKindTransfer ? 1 : 0;
}

// one gradient pass over n rows of (X, y): partials -> e->mlp_grad (+ loss at the end)
csmv_status grad_pass(csmv_engine *e, const uint16_t *X, const float *y, uint64_t n)
{
    const uint64_t ntiles = (n + 63) / 64;
    const uint32_t blocks =
        (uint32_t)std::min<uint64_t>(kMlpGradBlocks, (ntiles + kMlpWaves - 1) / kMlpWaves);
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(64 * kMlpWaves), 0, e->stream, X, y,
                       n, e->mlp_w1t, e->mlp_params, e->mlp_part);
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((kPartStride + 255) / 256), dim3(256), 0,
                       e->stream, e->mlp_part, blocks * kMlpWaves, e->mlp_grad);
    HIPCHK(e, hipGetLastError());
    return CSMV_OK;
}

}  // namespace

extern "C" {

csmv_status csmv_pred_init(csmv_engine *e, uint64_t seed)
{
    if (!e) return CSMV_E_ARG;
    csmv_status st = pred_ensure(e);
    if (st) return st;
    // Glorot-uniform weights from a counter-based stream, zero biases
    float h[kParamCount];
    const double a1 = std::sqrt(6.0 / (kFeat + kHidden)), a2 = std::sqrt(6.0 / (kHidden + 1));
    for (uint32_t i = 0; i < kParamCount; ++i) {
        const double u = (double)(splitmix64(seed * 0x10001ull + i) >> 11) * 0x1.0p-53;
        if (i < kParamB1) h[i] = (float)((2.0 * u - 1.0) * a1);
        else if (i >= kParamW2 && i < kParamB2) h[i] = (float)((2.0 * u - 1.0) * a2);
        else h[i] = 0.f;
    }
    return pred_upload(e, h);
}

csmv_status csmv_pred_set_params(csmv_engine *e, const float *params, size_t n)
{
    if (!e || !params || n != kParamCount) return fail(e, CSMV_E_ARG, "params: size mismatch");
    csmv_status st = pred_ensure(e);
    if (st) return st;
    return pred_upload(e, params);
}

csmv_status csmv_pred_get_params(csmv_engine *e, float *params, size_t n)
{
    if (!e || !params || n != kParamCount) return fail(e, CSMV_E_ARG, "params: size mismatch");
    csmv_status st = pred_ensure(e);
    if (st) return st;
    HIPCHK(e, hipMemcpy(params, e->mlp_params, kParamCount * 4, hipMemcpyDeviceToHost));
    return CSMV_OK;
}

csmv_status csmv_abort_fold(csmv_engine *e, uint32_t decay_shift)
{
    if (!e) return CSMV_E_ARG;
    csmv_status st = pred_ensure(e);
    if (st) return st;
    if (!e->bank_logged) return fail(e, CSMV_E_STATE, "fold needs a completed bank run");
    const uint64_t T = e->spec.n_txns;
    hipLaunchKernelGGL(abort_decay_kernel, dim3(kAbortBuckets / 256), dim3(256), 0, e->stream,
                       e->abort_ctr, decay_shift);
    hipLaunchKernelGGL(abort_fold_kernel, dim3((unsigned)((T + 255) / 256)), dim3(256), 0,
                       e->stream, e->op, e->addr, e->log_ab, T, e->abort_ctr);
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return CSMV_OK;
}

csmv_status csmv_abort_read_counters(csmv_engine *e, uint32_t *out, size_t n)
{
    if (!e || !out || n != kAbortBuckets) return fail(e, CSMV_E_ARG, "counters: size mismatch");
    csmv_status st = pred_ensure(e);
    if (st) return st;
    HIPCHK(e, hipMemcpy(out, e->abort_ctr, kAbortBuckets * 4, hipMemcpyDeviceToHost));
    return CSMV_OK;
}

csmv_status csmv_abort_load_counters(csmv_engine *e, const uint32_t *in, size_t n)
{
    if (!e || !in || n != kAbortBuckets) return fail(e, CSMV_E_ARG, "counters: size mismatch");
    csmv_status st = pred_ensure(e);
    if (st) return st;
    HIPCHK(e, hipMemcpy(e->abort_ctr, in, kAbortBuckets * 4, hipMemcpyHostToDevice));
    return CSMV_OK;
}

csmv_status csmv_pred_collect(csmv_engine *e, uint64_t max_samples, uint64_t *n_total)
{
    if (!e) return CSMV_E_ARG;
    csmv_status st = pred_ensure(e);
    if (st) return st;
    if (!e->bank_logged) return fail(e, CSMV_E_STATE, "collect needs a completed bank run");
    const uint64_t T = e->spec.n_txns;
    if ((st = pred_reserve(e, T))) return st;
    hipLaunchKernelGGL(transfer_flag_kernel, dim3((unsigned)((T + 255) / 256)), dim3(256), 0,
                       e->stream, e->op, T, e->pred_flag);
    HIPCHK(e, hipGetLastError());
    uint64_t n_tr = 0;
    if ((st = partition(e, T, e->scr_a, e->scr_b, &n_tr))) return st;
    const uint64_t n = std::min(n_tr, max_samples);
    if (e->ds_n + n > e->ds_cap) {
        const uint64_t cap = std::max<uint64_t>(e->ds_n + n, 2 * e->ds_cap);
        uint16_t *nx = nullptr;
        float *ny = nullptr;
        if ((st = dalloc(e, &nx, cap * kFeat)) || (st = dalloc(e, &ny, cap))) {
            dfree(nx);
            return st;
        }
        if (e->ds_n) {
            HIPCHK(e, hipMemcpyAsync(nx, e->ds_x, e->ds_n * kFeat * 2, hipMemcpyDeviceToDevice,
                                     e->stream));
            HIPCHK(e, hipMemcpyAsync(ny, e->ds_y, e->ds_n * 4, hipMemcpyDeviceToDevice, e->stream));
            HIPCHK(e, hipStreamSynchronize(e->stream));
        }
        dfree(e->ds_x);
        dfree(e->ds_y);
        e->ds_x = nx;
        e->ds_y = ny;
        e->ds_cap = cap;
    }
    if (n) {
        const uint64_t tiles = (n + 63) / 64;
        hipLaunchKernelGGL(features_kernel, dim3((unsigned)((tiles + kMlpWaves - 1) / kMlpWaves)),
                           dim3(64 * kMlpWaves), 0, e->stream, e->op, e->addr, T, e->log_ab,
                           e->scr_a, n, e->abort_ctr, e->ds_x + e->ds_n * kFeat,
                           e->ds_y + e->ds_n);
        HIPCHK(e, hipGetLastError());
        HIPCHK(e, hipStreamSynchronize(e->stream));
    }
    e->ds_n += n;
    if (n_total) *n_total = e->ds_n;
    return CSMV_OK;
}

csmv_status csmv_pred_read_dataset(csmv_engine *e, uint16_t *X, float *y, size_t n)
{
    if (!e || n > e->ds_n) return fail(e, CSMV_E_ARG, "dataset: n exceeds the collected rows");
    if (X && n) HIPCHK(e, hipMemcpy(X, e->ds_x, n * kFeat * 2, hipMemcpyDeviceToHost));
    if (y && n) HIPCHK(e, hipMemcpy(y, e->ds_y, n * 4, hipMemcpyDeviceToHost));
    return CSMV_OK;
}

csmv_status csmv_pred_clear_dataset(csmv_engine *e)
{
    if (!e) return CSMV_E_ARG;
    e->ds_n = 0;
    return CSMV_OK;
}

csmv_status csmv_pred_train(csmv_engine *e, uint32_t steps, float lr, float *loss)
{
    if (!e || steps == 0 || !(lr > 0.f)) return CSMV_E_ARG;
    csmv_status st = pred_ensure(e);
    if (st) return st;
    if (e->ds_n == 0) return fail(e, CSMV_E_STATE, "train needs a collected dataset");
    if (steps > e->mlp_loss_cap) {
        dfree(e->mlp_loss);
        e->mlp_loss_cap = 0;
        if ((st = dalloc(e, &e->mlp_loss, steps))) return st;
        e->mlp_loss_cap = steps;
    }
    for (uint32_t s = 0; s < steps; ++s) {
        if ((st = grad_pass(e, e->ds_x, e->ds_y, e->ds_n))) return st;
        const uint32_t t = ++e->mlp_step;
        const float c1 = (float)(1.0 - std::pow((double)kAdamBeta1, (double)t));
        const float c2 = (float)(1.0 - std::pow((double)kAdamBeta2, (double)t));
        hipLaunchKernelGGL(mlp_adam_kernel, dim3((kParamCount + 255) / 256), dim3(256), 0,
                           e->stream, e->mlp_params, e->mlp_m, e->mlp_v, e->mlp_grad, e->mlp_w1t,
                           lr, kAdamBeta1, kAdamBeta2, kAdamEps, c1, c2);
        // the loss of this pass (before the update)
        HIPCHK(e, hipMemcpyAsync(e->mlp_loss + s, e->mlp_grad + kParamCount, 4,
                                 hipMemcpyDeviceToDevice, e->stream));
        HIPCHK(e, hipGetLastError());
    }
    if (loss) HIPCHK(e, hipMemcpyAsync(loss, e->mlp_loss, steps * 4, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return CSMV_OK;
}

csmv_status csmv_pred_forward(csmv_engine *e, const uint16_t *X, size_t n, float *p)
{
    if (!e || !X || !p || n == 0) return CSMV_E_ARG;
    csmv_status st = pred_ensure(e);
    if (st || (st = io_reserve(e, n))) return st;
    HIPCHK(e, hipMemcpyAsync(e->io_x, X, n * kFeat * 2, hipMemcpyHostToDevice, e->stream));
    const uint64_t tiles = (n + 63) / 64;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)((tiles + kMlpWaves - 1) / kMlpWaves)),
                       dim3(64 * kMlpWaves), 0, e->stream, e->io_x, (uint64_t)n, e->mlp_w1t,
                       e->mlp_params, e->io_p);
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipMemcpyAsync(p, e->io_p, n * 4, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return CSMV_OK;
}

csmv_status csmv_pred_grad(csmv_engine *e, const uint16_t *X, const float *y, size_t n,
                           float *grad, float *loss)
{
    if (!e || !X || !y || !grad || n == 0) return CSMV_E_ARG;
    csmv_status st = pred_ensure(e);
    if (st || (st = io_reserve(e, n))) return st;
    HIPCHK(e, hipMemcpyAsync(e->io_x, X, n * kFeat * 2, hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipMemcpyAsync(e->io_y, y, n * 4, hipMemcpyHostToDevice, e->stream));
    if ((st = grad_pass(e, e->io_x, e->io_y, n))) return st;
    float h[kPartStride];
    HIPCHK(e, hipMemcpyAsync(h, e->mlp_grad, kPartStride * 4, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    std::memcpy(grad, h, kParamCount * 4);
    if (loss) *loss = h[kParamCount];
    return CSMV_OK;
}

csmv_status csmv_pred_infer(csmv_engine *e, float tau, float *p, size_t n)
{
    if (!e) return CSMV_E_ARG;
    if (!e->has_stream) return fail(e, CSMV_E_STATE, "infer before generate");
    const uint64_t T = e->spec.n_txns;
    if (p && n != T) return fail(e, CSMV_E_ARG, "infer: one probability per transaction");
    csmv_status st = pred_ensure(e);
    if (st || (st = pred_reserve(e, T))) return st;
    const uint64_t tiles = (T + 63) / 64;
    hipLaunchKernelGGL(mlp_infer_kernel, dim3((unsigned)((tiles + kMlpWaves - 1) / kMlpWaves)),
                       dim3(64 * kMlpWaves), 0, e->stream, e->op, e->addr, T, e->abort_ctr,
                       e->mlp_w1t, e->mlp_params, tau, e->pred_p, e->pred_flag);
    HIPCHK(e, hipGetLastError());
    if (p) HIPCHK(e, hipMemcpyAsync(p, e->pred_p, T * 4, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return CSMV_OK;
}

csmv_status csmv_bank_route_learned(csmv_engine *e, float tau, uint64_t *n_hot)
{
    csmv_status st = csmv_pred_infer(e, tau, nullptr, 0);
    if (st) return st;
    const uint64_t T = e->spec.n_txns;
    e->routed = false;
    if ((st = route_reserve(e, T))) return st;
    uint64_t nh = 0;
    if ((st = partition(e, T, e->hot_list, e->norm_list, &nh))) return st;
    e->n_hot = nh;
    e->n_norm = T - nh;
    e->routed = true;
    if (n_hot) *n_hot = nh;
    return CSMV_OK;
}

csmv_status csmv_bank_read_route(csmv_engine *e, uint32_t *norm, uint32_t *hot, uint64_t *n_norm,
                                 uint64_t *n_hot)
{
    if (!e) return CSMV_E_ARG;
    if (!e->routed) return fail(e, CSMV_E_STATE, "stream is not routed");
    if (n_norm) *n_norm = e->n_norm;
    if (n_hot) *n_hot = e->n_hot;
    if (norm && e->n_norm) HIPCHK(e, hipMemcpy(norm, e->norm_list, e->n_norm * 4, hipMemcpyDeviceToHost));
    if (hot && e->n_hot) HIPCHK(e, hipMemcpy(hot, e->hot_list, e->n_hot * 4, hipMemcpyDeviceToHost));
    return CSMV_OK;
}

}  // extern "C"

// =============================================================================================
// serving loop (serve.hpp): Poisson arrivals, admission by in-flight version bytes,
// continuous batching over a cache partition and a bank partition
// =============================================================================================
namespace {

// Poisson arrivals at `rate` per second as wall-clock ticks from the serving start:
// exponential gaps -ln(u) / rate with u from splitmix64(seed + k) in (0, 1].
csmv_status poisson_arrivals(csmv_engine *e, uint64_t seed, double rate, uint64_t n)
{
    if (n > e->arr_cap) {
        dfree(e->arr);
        e->arr_cap = 0;
        csmv_status st = dalloc(e, &e->arr, n);
        if (st) return st;
        e->arr_cap = n;
    }
    uint32_t *h = new (std::nothrow) uint32_t[n];
    if (!h) return fail(e, CSMV_E_CAPACITY, "arrival host buffer");
    const double ticks_per_s = (double)e->wallclock_khz * 1e3;
    double t = 0.0;
    for (uint64_t k = 0; k < n; ++k) {
        const double u = (double)((splitmix64(seed + k) >> 11) + 1) * 0x1.0p-53;
        t += -std::log(u) / rate;
        const double tick = t * ticks_per_s;
        if (tick >= 4294967295.0) {
            delete[] h;
            return fail(e, CSMV_E_CAPACITY, "arrival span beyond 2^32 clock ticks");
        }
        h[k] = (uint32_t)tick;
    }
    hipError_t r = hipMemcpy(e->arr, h, n * 4, hipMemcpyHostToDevice);
    delete[] h;
    HIPCHK(e, r);
    return CSMV_OK;
}

// Server layout of a partition: fewer mailboxes (down to 4) when the wire is wide, so
// that the stamp buckets keep at least 4096 entries.
template <int R, int W> csmv_status fit_layout(csmv_engine *e, ServerLayout &L, bool delegation)
{
    uint32_t nmb = 0;
    for (;;) {
        L = server_layout<R, W>(e, delegation, nmb);
        const bool pipe = delegation && (e->cfg.variant & kSerialPipe);
        const bool fits = (uint64_t)L.words * 4 <= e->lds_max && (pipe || L.exact || L.nbuckets >= 4096);
        if (fits) return CSMV_OK;
        if (L.nmb <= 4)
            return fail(e, CSMV_E_CAPACITY, "serving server layout does not fit the LDS budget");
        nmb = L.nmb - 1;
    }
}

}  // namespace

extern "C" {

csmv_status csmv_serve_run_multi(csmv_engine *const *caches, uint32_t n_cache, csmv_engine *bank,
                                 const csmv_serve_spec *sp, csmv_serve_stats *out)
{
    if (n_cache && !caches) return CSMV_E_ARG;
    csmv_engine *p = n_cache ? caches[0] : bank;   // holds the shared serving words
    if (!p) return CSMV_E_ARG;
    if (n_cache + (bank ? 1u : 0u) > kMaxServeParts)
        return fail(p, CSMV_E_CAPACITY, "at most 7 cache partitions beside the bank partition");
    // each class arrives at its share of the offered load whether or not the other class
    // is served, so a partition served alone sees the arrivals it sees in the mix
    const double rate_c = sp ? sp->offered_rps * sp->pct_cache / 100.0 : 0.0;
    const double rate_b = sp ? sp->offered_rps * (100 - sp->pct_cache) / 100.0 : 0.0;
    if (!sp || !(sp->offered_rps > 0.0) || sp->pct_cache > 100 || (n_cache && !(rate_c > 0.0)) ||
        (bank && !(rate_b > 0.0)) || (sp->pool_budget && sp->pool_budget < 1024))
        return fail(p, CSMV_E_ARG, "serve spec out of range");
    std::vector<csmv_engine *> parts;
    for (uint32_t i = 0; i < n_cache; ++i) {
        if (!caches[i]) return fail(p, CSMV_E_ARG, "null cache partition");
        parts.push_back(caches[i]);
    }
    if (bank) parts.push_back(bank);
    for (size_t i = 0; i < parts.size(); ++i)
        for (size_t j = 0; j < i; ++j)
            if (parts[i] == parts[j]) return fail(p, CSMV_E_ARG, "an engine serves one partition");
    for (csmv_engine *e : parts) {
        if ((e->cfg.variant & 3u) != kValStamp || (e->cfg.variant & kServerWriteback) ||
            ((e->cfg.variant & kSerialPipe) && e != bank))
            return fail(e, CSMV_E_ARG, "serving runs the STAMP variant with client write-back "
                                       "(the pipelined lane for the bank partition only)");
        if (!e->pool_loaded) return fail(e, CSMV_E_STATE, "serve needs a freshly loaded pool");
        if (e->device != p->device || e->cfg.wg_waves != p->cfg.wg_waves)
            return fail(e, CSMV_E_ARG, "partitions must share the device and wg_waves");
    }
    for (uint32_t i = 0; i < n_cache; ++i)
        if (!caches[i]->has_mc) return fail(caches[i], CSMV_E_STATE, "cache partition has no stream");
    if (bank && (!bank->has_stream || bank->spec.pct_ro != 0))
        return fail(bank, CSMV_E_STATE, "bank partition needs a transfer-only stream");
    // Arrivals: one Poisson process per class; a cache request's partition depends on its
    // key alone, so splitting the cache class by partition thins it into independent Poisson
    // processes, each at the class rate times the partition's share of the requests.
    uint64_t n_cache_reqs = 0;
    for (uint32_t i = 0; i < n_cache; ++i) n_cache_reqs += caches[i]->mspec.n_reqs;
    csmv_status st;
    ServeArgs args{};
    args.n_cache = n_cache;
    args.has_bank = bank ? 1u : 0u;
    uint32_t words = 0, wgs = 0;
    for (uint32_t i = 0; i < n_cache; ++i) {
        csmv_engine *e = caches[i];
        const double share = (double)e->mspec.n_reqs / (double)n_cache_reqs;
        const uint64_t seed = i == 0 ? splitmix64(sp->seed * 2 + 1) : splitmix64(sp->seed * 2 + 1 + 1000003ull * i);
        if ((st = poisson_arrivals(e, seed, rate_c * share, e->mspec.n_reqs))) return st;
        ServerLayout &L = args.L[i];
        if ((st = fit_layout<kMcR, kMcW>(e, L, false))) return st;
        Ctx &c = args.c[i];
        base_ctx(e, c);
        c.n_serial_waves = 0;
        c.escalate = 0;
        c.n_txns = e->mspec.n_reqs;
        c.req_key = e->mc_key;
        c.req_op = e->mc_op;
        c.n_sets = e->mspec.n_sets;
        c.n_ways = e->mspec.n_ways;
        c.n_slots = e->mspec.n_sets * e->mspec.n_ways;
        c.mc_batch = e->mspec.batch;
        c.log_ts = e->mc_ts;
        c.log_res = e->mc_res;
        c.log_vsum = e->mc_vsum;
        c.log_ab = e->mc_ab;
        c.log_lat = e->mc_lat;
        c.n_norm = c.n_txns;
        c.arr = e->arr;
        words = std::max(words, L.words);
        wgs += (e->cfg.client_waves + e->cfg.client_wg_waves - 1) / e->cfg.client_wg_waves;
        args.wg_end[i] = wgs;
    }
    if (bank) {
        const uint32_t bi = n_cache;
        if ((st = poisson_arrivals(bank, splitmix64(sp->seed * 2 + 2), rate_b, bank->spec.n_txns)))
            return st;
        // the pipelined lane takes every bank transaction (no routing, no normal lane)
        const bool pipe = (bank->cfg.variant & kSerialPipe) != 0;
        const bool serial = !pipe && bank->routed && bank->n_hot > 0;
        if (serial && bank->cfg.n_serial_waves == 0)
            return fail(bank, CSMV_E_ARG, "a routed bank partition needs n_serial_waves > 0");
        // escalated transactions travel on the serialized lane too
        const bool delegation = pipe || serial || bank->cfg.escalate_after != 0;
        ServerLayout &L = args.L[bi];
        if ((st = fit_layout<kServeBankR, kBankW>(bank, L, delegation))) return st;
        Ctx &c = args.c[bi];
        base_ctx(bank, c);
        args.bank_pipe = pipe ? 1u : 0u;
        c.pipe_parts = pipe ? L.parts : 0u;
        c.pipe_heads = 0;
        c.n_serial_waves = serial ? bank->cfg.n_serial_waves : 0;
        c.hot_list = serial ? bank->hot_list : nullptr;
        c.n_hot = serial ? bank->n_hot : 0;
        c.op = bank->op;
        c.addr = bank->addr;
        c.n_txns = bank->spec.n_txns;
        c.log_ts = bank->log_ts;
        c.log_v = bank->log_v;
        c.log_ab = bank->log_ab;
        c.log_lat = bank->log_lat;
        c.norm_list = (bank->routed && !pipe) ? bank->norm_list : nullptr;
        c.n_norm = (bank->routed && !pipe) ? bank->n_norm : bank->spec.n_txns;
        c.arr = bank->arr;
        words = std::max(words, L.words);
        wgs += (bank->cfg.client_waves + bank->cfg.client_wg_waves - 1) / bank->cfg.client_wg_waves;
        args.wg_end[bi] = wgs;
    }
    for (size_t i = 0; i < parts.size(); ++i) {
        args.c[i].t_start = (uint64_t *)(p->ctl + kOffServeStart);
        args.c[i].inflight = (uint64_t *)(p->ctl + kOffInflight);
        args.c[i].serve_ctr = (uint64_t *)(p->ctl + kOffServeCtr);
        args.c[i].budget = sp->pool_budget;
    }
    const size_t lds_bytes = (size_t)words * 4;
    if (lds_bytes > p->lds_max) return fail(p, CSMV_E_CAPACITY, "serving LDS layout too large");
    const uint32_t wg = p->cfg.wg_waves;
    // a pipelined bank partition has a server workgroup per commit partition
    const uint32_t bank_srv = args.bank_pipe ? args.L[n_cache].parts : (bank ? 1u : 0u);
    const uint32_t grid = n_cache + bank_srv + wgs;
    auto kern = args.bank_pipe ? serve_kernel<kMcR, kMcW, kServeBankR, kBankW, true>
                               : serve_kernel<kMcR, kMcW, kServeBankR, kBankW, false>;
    HIPCHK(p, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds_bytes));
    int per_cu = 0;
    HIPCHK(p, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, (int)(wg * kWave),
                                                           lds_bytes));
    if ((uint64_t)per_cu * p->n_cu < grid)
        return fail(p, CSMV_E_CAPACITY, "serving grid exceeds what the device keeps resident");
    for (size_t i = 0; i < parts.size(); ++i)
        if ((st = reset_run_state(parts[i], args.L[i]))) return st;
    for (csmv_engine *e : parts) HIPCHK(e, hipStreamSynchronize(e->stream));

    HIPCHK(p, hipEventRecord(p->ev0, p->stream));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(wg * kWave), lds_bytes, p->stream, args);
    HIPCHK(p, hipGetLastError());
    if (args.bank_pipe) {
        // no snapshot reader ran beside the pipelined bank lane: its head words follow from
        // the install counts
        const uint32_t n = bank->cfg.n_items;
        hipLaunchKernelGGL(pipe_heads_kernel, dim3((n + 255) / 256), dim3(256), 0, p->stream, bank->sv,
                           bank->ht, n, bank->cfg.depth);
        HIPCHK(p, hipGetLastError());
    }
    HIPCHK(p, hipEventRecord(p->ev1, p->stream));
    HIPCHK(p, hipEventSynchronize(p->ev1));
    float ms = 0.f;
    HIPCHK(p, hipEventElapsedTime(&ms, p->ev0, p->ev1));
    csmv_serve_stats o{};
    std::vector<uint32_t> derr(parts.size(), 0);
    for (size_t i = 0; i < parts.size(); ++i) {
        csmv_engine *e = parts[i];
        e->pool_loaded = false;
        const bool is_bank = bank && i + 1 == parts.size();
        e->served = is_bank ? 2 : 1;
        if (is_bank) e->bank_logged = false;
        csmv_stats one{};
        if ((st = collect_stats(e, args.L[i], ms, &one, &derr[i]))) return st;
        if (is_bank) {
            o.bank = one;
        } else if (i == 0) {
            o.cache = one;
        } else {
            // the cache class's counters summed over its partitions
            o.cache.commits_update += one.commits_update;
            o.cache.commits_ro += one.commits_ro;
            o.cache.aborts_dataage += one.aborts_dataage;
            o.cache.aborts_readwrite += one.aborts_readwrite;
            o.cache.aborts_prevalid += one.aborts_prevalid;
            o.cache.aborts_recordage += one.aborts_recordage;
            o.cache.aborts_unknown += one.aborts_unknown;
            o.cache.batches += one.batches;
            o.cache.client_lanes += one.client_lanes;
            o.cache.error = o.cache.error ? o.cache.error : one.error;
            for (int k = 0; k < 16; ++k) o.cache.prof_ns[k] += one.prof_ns[k];
        }
    }
    uint64_t words3[3] = {0, 0, 0};
    uint64_t inflight = 0;
    HIPCHK(p, hipMemcpy(words3, p->ctl + kOffServeCtr, 16, hipMemcpyDeviceToHost));
    HIPCHK(p, hipMemcpy(&inflight, p->ctl + kOffInflight, 8, hipMemcpyDeviceToHost));
    o.kernel_ms = ms;
    o.admit_stalls = words3[0];
    o.peak_inflight = words3[1];
    o.final_inflight = inflight;
    o.wallclock_khz = p->wallclock_khz;
    if (out) *out = o;
    for (size_t i = 0; i < parts.size(); ++i)
        if ((st = device_error(parts[i], derr[i]))) return st;
    // the bank log of a completed run feeds the abort counters like csmv_bank_run's
    if (bank) bank->bank_logged = true;
    return CSMV_OK;
}

csmv_status csmv_serve_run(csmv_engine *cache, csmv_engine *bank, const csmv_serve_spec *sp,
                           csmv_serve_stats *out)
{
    if ((!cache && !bank) || cache == bank) return fail(cache ? cache : bank, CSMV_E_ARG, "serve: two distinct partitions");
    return csmv_serve_run_multi(cache ? &cache : nullptr, cache ? 1u : 0u, bank, sp, out);
}

csmv_status csmv_serve_read_arrivals(csmv_engine *e, uint32_t *arr_ticks, size_t n)
{
    if (!e || !arr_ticks) return CSMV_E_ARG;
    const uint64_t T = e->served == 1 ? e->mspec.n_reqs : e->served == 2 ? e->spec.n_txns : 0;
    if (!T || n != T) return fail(e, CSMV_E_ARG, "arrivals: one tick per request of the served stream");
    HIPCHK(e, hipMemcpy(arr_ticks, e->arr, T * 4, hipMemcpyDeviceToHost));
    return CSMV_OK;
}

csmv_status csmv_serve_read_latency(csmv_engine *e, uint32_t *lat_ticks, size_t n)
{
    if (!e || !lat_ticks) return CSMV_E_ARG;
    const uint64_t T = e->served == 1 ? e->mspec.n_reqs : e->served == 2 ? e->spec.n_txns : 0;
    if (!T || n != T) return fail(e, CSMV_E_ARG, "latency: one entry per request of the served stream");
    HIPCHK(e, hipMemcpy(lat_ticks, e->served == 1 ? e->mc_lat : e->log_lat, T * 4,
                        hipMemcpyDeviceToHost));
    return CSMV_OK;
}

}  // extern "C"

// =============================================================================================
// bank run without allocation: every buffer beyond the caller's inputs and outputs lives in
// one file-scope device array (its address per device from hipGetSymbolAddress)
// =============================================================================================
namespace {

static_assert(CSMV_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch is a whole number of 256-byte blocks");
// In host code `scratch` names a host-side stand-in; only the address hipGetSymbolAddress
// returns for the current device may be handed to kernels or copies.
static __device__ __align__(256) uint8_t scratch[CSMV_NOALLOC_SCRATCH_BYTES];

struct ScratchDevice {
    bool ready;
    uint8_t *base;        // this device's copy of scratch
    uint32_t n_cu, lds_max, khz;
};
constexpr int kMaxDevices = 64;
ScratchDevice g_scratch_dev[kMaxDevices];
std::mutex g_scratch_mu;

// Once per device, with that device current: the array's address and the properties the
// launch needs.
csmv_status scratch_device(int dev, ScratchDevice **out)
{
    if (dev < 0 || dev >= kMaxDevices) return CSMV_E_CAPACITY;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    ScratchDevice &d = g_scratch_dev[dev];
    if (!d.ready) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return CSMV_E_HIP;
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return CSMV_E_NODEVICE;
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr)
            return CSMV_E_HIP;
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
            khz = 100000;
        d.base = (uint8_t *)p;
        d.n_cu = (uint32_t)prop.multiProcessorCount;
        d.lds_max = (uint32_t)prop.sharedMemPerBlock;
        d.khz = (uint32_t)khz;
        d.ready = true;
    }
    *out = &d;
    return CSMV_OK;
}

// Byte offsets of one run's buffers inside the array.  The words that must start at zero
// (control block, rings, replies) come first, so one clear covers [0, zero_end); the pool
// fill sets the rings and the committed words (stamp 0).
struct ScratchMap {
    size_t ctl, ring, done, reply, reply_v, zero_end, mbox, ht, slot, sv, total;
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

ScratchMap scratch_map(const csmv_config &cfg, uint32_t Q)
{
    ScratchMap m{};
    const size_t waves = cfg.client_waves, n = cfg.n_items;
    size_t o = 0;
    m.ctl = o; o = align256(o + kCtlBytes);
    m.ring = o; o = align256(o + (size_t)Q * 8);
    m.done = o; o = align256(o + (size_t)Q * 8);
    m.reply = o; o = align256(o + waves * kWave * 8);
    m.reply_v = o; o = align256(o + waves * kWave * 2 * kBankW * 8);
    m.zero_end = o;
    m.mbox = o; o = align256(o + waves * mbox_rows<kBankR, kBankW>() * kWave * 4);
    m.ht = o; o = align256(o + n * 4);
    m.slot = o; o = align256(o + n * cfg.depth * 8);
    m.sv = o; o = align256(o + n * 8);
    m.total = o;
    return m;
}

__global__ void __launch_bounds__(256) scratch_clear_kernel(uint64_t *p, size_t n_words)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_words;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

}  // namespace

extern "C" csmv_status csmv_bank_run_noalloc(const csmv_config *cfg_in, const csmv_bank_buffers *b,
                                             void *stream, csmv_stats *out)
{
    if (!cfg_in || !b) return CSMV_E_ARG;
    csmv_config cfg = *cfg_in;
    if (normalize_config(cfg) || cfg.n_serial_waves || (cfg.variant & kSerialPipe)) return CSMV_E_ARG;
    if (!b->init || !b->final_state || !b->op || !b->addr || !b->log_ts || !b->log_v || !b->log_ab ||
        b->n_txns == 0 || b->n_txns >= (1ull << 31))
        return CSMV_E_ARG;
    const uint32_t Q = pow2_at_least(std::max<uint64_t>(4096, 2ull * cfg.client_waves));
    const ScratchMap m = scratch_map(cfg, Q);
    if (m.total > CSMV_NOALLOC_SCRATCH_BYTES) return CSMV_E_CAPACITY;
    int ndev = 0, dev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0 || hipGetDevice(&dev) != hipSuccess)
        return CSMV_E_NODEVICE;
    ScratchDevice *d = nullptr;
    csmv_status s = scratch_device(dev, &d);
    if (s) return s;
    const uint32_t lds_max = cfg.lds_kib ? std::min<uint32_t>(d->lds_max, cfg.lds_kib * 1024u) : d->lds_max;
    const ServerLayout L = server_layout_cfg(cfg, lds_max, kBankR, kBankW, true);
    const size_t lds_bytes = (size_t)L.words * 4;
    if (lds_bytes > lds_max) return CSMV_E_CAPACITY;
    const uint32_t block = cfg.wg_waves * kWave;
    const uint32_t grid = 1 + (cfg.client_waves + cfg.client_wg_waves - 1) / cfg.client_wg_waves;
    auto kern = bank_kernel<kBankR, kBankW>;
    if (hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds_bytes) != hipSuccess)
        return CSMV_E_HIP;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, (int)block, lds_bytes) != hipSuccess)
        return CSMV_E_HIP;
    if ((uint64_t)per_cu * d->n_cu < grid) return CSMV_E_CAPACITY;

    uint8_t *base = d->base;
    Ctx c{};
    ctx_config(c, cfg, d->khz, Q);
    ctx_control(c, base + m.ctl);
    c.ring = (uint64_t *)(base + m.ring);
    c.done = (uint64_t *)(base + m.done);
    c.reply = (uint64_t *)(base + m.reply);
    c.reply_v = (uint64_t *)(base + m.reply_v);
    c.mbox = (uint32_t *)(base + m.mbox);
    c.ht = (uint32_t *)(base + m.ht);
    c.slot = (uint64_t *)(base + m.slot);
    c.sv = (uint64_t *)(base + m.sv);
    c.op = b->op;
    c.addr = b->addr;
    c.n_txns = b->n_txns;
    c.n_norm = b->n_txns;
    c.log_ts = b->log_ts;
    c.log_v = b->log_v;
    c.log_ab = b->log_ab;

    hipStream_t st = (hipStream_t)stream;
    const uint32_t n = cfg.n_items;
    // the zero-initialised words by a fill kernel (no memset node if the caller captures),
    // then the pool from the initial balances
    hipLaunchKernelGGL(scratch_clear_kernel, dim3(256), dim3(256), 0, st, (uint64_t *)base,
                       m.zero_end / 8);
    hipLaunchKernelGGL(pool_fill_kernel, dim3((n + 255) / 256), dim3(256), 0, st, c.ht, c.slot,
                       c.sv, n, cfg.depth, b->init, 0);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) return CSMV_E_HIP;
    const auto t0 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds_bytes, st, c, L);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) return CSMV_E_HIP;
    const auto t1 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL(pool_head_kernel, dim3((n + 255) / 256), dim3(256), 0, st, c.ht, c.slot, n,
                       cfg.depth, b->final_state);
    uint8_t host_ctl[kCtlBytes];
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(host_ctl, base + m.ctl, kCtlBytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return CSMV_E_HIP;
    csmv_stats local;
    decode_stats(host_ctl, L, (float)std::chrono::duration<double, std::milli>(t1 - t0).count(),
                 cfg.client_waves, d->khz, &local);
    if (out) *out = local;
    if (local.error == kErrLivelock) return CSMV_E_LIVELOCK;
    if (local.error == kErrReadSetOverflow) return CSMV_E_CAPACITY;
    if (local.error) return CSMV_E_TIMEOUT;
    return CSMV_OK;
}
