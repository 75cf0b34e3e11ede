// server.hpp -- the commit-server workgroup of the gfx950 CSMV engine.
//
// The reference runs one persistent server block: a leader warp scans a ready bitmask
// and hands 32-message chunks to worker warps through shared memory
// (bank/sync_lib/msg_passing.h:85-370, one_phase_server.h:44-108); each worker lane
// validates and inserts its client's record (TXAddToRecord, bank/csmv_sh/API.cu:193-272)
// and answers through a per-warp scoreboard (bank/csmv_sh/exe.cu:56-79).
//
// Here the server is ONE workgroup whose waves have fixed roles:
//   wave 0       committer: takes requests strictly in arrival order from LDS
//                mailboxes, validates the 64 lanes of a request at once, resolves
//                conflicts between lanes of the same request in lane order, assigns
//                consecutive timestamps, records the write sets and replies;
//   wave 1       clock advancer (client write-back): scans 64 completion words per
//                step and moves the global clock over every installed batch;
//   waves 2..    receivers: request n goes to receiver n mod M, which polls the
//                request ring and copies the client's mailbox into LDS mailbox
//                n mod NMB, so the committer never waits on HBM for request data.
#pragma once

#include "csmv_device.hpp"

namespace csmv {

enum : uint32_t { kLdsCnext = 0, kLdsCdone = 1, kLdsCfinal = 2, kLdsAdv = 3, kLdsHdr = 8 };
constexpr uint32_t kExitBit = 0x80000000u;

__device__ __forceinline__ uint32_t lds_ld_acq(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_st_rel(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ uint32_t wave_min(uint32_t v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        uint32_t o = (uint32_t)__shfl_xor((int)v, m, 64);
        v = o < v ? o : v;
    }
    return v;
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

// ---- committer --------------------------------------------------------------------------
template <int R, int W> __device__ __forceinline__ void committer(const Ctx &c, uint32_t *lds)
{
    constexpr uint32_t rows = mbox_rows<R, W>();
    const uint32_t lane = lane_id();
    const uint32_t E = c.atr_entries, emask = E - 1;
    const uint32_t vmode = c.variant & 0x3u;
    const bool batched = (c.variant & kInsertBatch) != 0;
    const bool server_wb = (c.variant & kServerWriteback) != 0;
    uint32_t *mbst = lds + kLdsHdr;
    uint32_t *mbwave = lds + kLdsHdr + c.nmb;
    uint32_t *atr = lds + c.lds_atr_off;
    uint32_t *lstamp = lds + c.lds_stamp_off;
    const uint32_t Q = c.qmask + 1;

    uint32_t last = 0;  // newest assigned commit timestamp
    uint64_t n_rw = 0, n_pv = 0, n_rec = 0;
    uint32_t n = 0;
    for (;; ++n) {
        const uint32_t mbi = n % c.nmb;
        uint32_t s = 0;
        bool bad = false;
        {
            Spin sp;
            for (;;) {
                s = lds_ld_acq(&mbst[mbi]);
                if ((s & ~kExitBit) == n + 1) break;
                if (!sp.tick(c, kErrMailboxTimeout)) { bad = true; break; }
            }
        }
        if (bad || (s & kExitBit)) break;
        if (!server_wb && n >= Q) {   // the completion slot n mod Q must be consumed
            Spin sp;
            while (lds_ld_acq(&lds[kLdsAdv]) < n - Q + 1)
                if (!sp.tick(c, kErrDoneRingTimeout)) { bad = true; break; }
            if (bad) break;
        }
        const uint32_t cw = mbwave[mbi];
        const uint32_t *box = lds + c.lds_mb_off + mbi * rows * kWave;
        const uint32_t snap = box[lane];
        uint32_t rs[R], wa[W];
        int32_t wv[W];
#pragma unroll
        for (int r = 0; r < R; ++r) rs[r] = box[(1 + r) * kWave + lane];
#pragma unroll
        for (int w = 0; w < W; ++w) {
            wa[w] = box[(1 + R + w) * kWave + lane];
            wv[w] = server_wb ? (int32_t)box[(1 + R + W + w) * kWave + lane] : 0;
        }
        // the mailbox is in registers now: hand it back to the receivers
        if (lane == 0) lds_st_rel(&lds[kLdsCnext], n + 1);

        const bool valid = snap != kNone;
        bool rec = false, conflict = false;
        if (vmode == kValStamp) {
            if (valid) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (rs[r] != kNone)
                        conflict |= (c.stamp_in_lds ? lstamp[rs[r]] : c.stamps[rs[r]]) > snap;
#pragma unroll
                for (int w = 0; w < W; ++w)
                    if (wa[w] != kNone)
                        conflict |= (c.stamp_in_lds ? lstamp[wa[w]] : c.stamps[wa[w]]) > snap;
            }
        } else {
            // record ring indexed by timestamp: entry t lives at t mod E and the ring
            // covers (last - E, last]; an older snapshot cannot be validated
            // ("RecordAge", bank/csmv_sh/API.cu:93-100).
            rec = valid && (last - snap) > E;
            const bool live = valid && !rec;
            const uint64_t lm = ballot(live);
            if (lm && vmode == kValScan) {
                // TXValidate: every lane walks back from the newest record to its
                // own snapshot (bank/csmv_sh/API.cu:104-122); lanes read the same
                // record at the same step, so the LDS read is a broadcast.
                const uint32_t lo = wave_min(live ? snap : kNone);
                for (uint32_t t = last; t > lo; --t) {
                    const bool chk = live && !conflict && t > snap;
                    if (ballot(chk) == 0) break;
                    const uint32_t e = t & emask;
                    if (chk) {
#pragma unroll
                        for (int w = 0; w < W; ++w) {
                            const uint32_t x = atr[w * E + e];
                            if (x != kNone) conflict |= hits<R, W>(x, rs, wa);
                        }
                    }
                }
            } else if (lm) {
                // collaborative validation (collabValidation/csmv_sh/API.cu:212-264):
                // the 64 lanes split one client lane's window, lane l taking records
                // last - l, last - l - 64, ...; client lanes are served in lane order.
                uint64_t it = lm;
                while (it) {
                    const uint32_t cl = (uint32_t)__builtin_ctzll(it);
                    it &= it - 1;
                    const uint32_t sc = readlane(snap, cl);
                    uint32_t rsc[R], wac[W];
#pragma unroll
                    for (int r = 0; r < R; ++r) rsc[r] = readlane(rs[r], cl);
#pragma unroll
                    for (int w = 0; w < W; ++w) wac[w] = readlane(wa[w], cl);
                    const uint32_t K = last - sc;
                    bool cf = false;
                    for (uint32_t base = 0; base < K; base += kWave) {
                        const uint32_t d = base + lane;
                        if (d < K) {
                            const uint32_t e = (last - d) & emask;
#pragma unroll
                            for (int w = 0; w < W; ++w) {
                                const uint32_t x = atr[w * E + e];
                                if (x != kNone) cf |= hits<R, W>(x, rsc, wac);
                            }
                        }
                        if (ballot(cf)) break;
                    }
                    const bool any = ballot(cf) != 0;
                    if (lane == cl) conflict = any;
                }
            }
        }
        const bool ok = valid && !rec && !conflict;
        n_rec += __builtin_popcountll(ballot(rec));
        n_rw += __builtin_popcountll(ballot(valid && !rec && conflict));

        // Lanes of one request, in lane order (TXPreValidation,
        // batchedAtrInsert/csmv_sh/API.cu:94-191): lane L conflicts with lower lane i
        // when i writes an item L reads or writes; L commits only if no conflicting
        // lower lane commits.
        const uint64_t m = ballot(ok);
        uint64_t conf = 0;
        {
            uint64_t it = m;
            while (it) {
                const uint32_t i = (uint32_t)__builtin_ctzll(it);
                it &= it - 1;
                bool h = false;
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    const uint32_t x = readlane(wa[w], i);
                    if (x != kNone) h |= hits<R, W>(x, rs, wa);
                }
                if (h && lane > i) conf |= 1ull << i;
            }
        }
        uint64_t committed = m;
        if (ballot(ok && conf != 0)) {
            committed = 0;
            uint64_t it = m;
            while (it) {
                const uint32_t i = (uint32_t)__builtin_ctzll(it);
                it &= it - 1;
                if ((readlane64(conf, i) & committed) == 0) committed |= 1ull << i;
            }
        }
        n_pv += __builtin_popcountll(m & ~committed);
        const bool mine = (committed >> lane) & 1ull;
        const uint32_t nc = (uint32_t)__builtin_popcountll(committed);
        uint32_t myts = 0;
        if (batched) {
            // one batched insertion (batchedAtrInsert/csmv_sh/API.cu:296-377):
            // consecutive timestamps in lane order, offset = popc(committed & lt)
            if (mine) {
                myts = last + 1 + prefix_count(committed);
                if (vmode == kValStamp) {
#pragma unroll
                    for (int w = 0; w < W; ++w)
                        if (wa[w] != kNone) {
                            if (c.stamp_in_lds) lstamp[wa[w]] = myts;
                            else c.stamps[wa[w]] = myts;
                        }
                } else {
#pragma unroll
                    for (int w = 0; w < W; ++w) atr[w * E + (myts & emask)] = wa[w];
                }
            }
            last += nc;
        } else {
            // per-lane insertion (bank/csmv_sh/API.cu:219-265): one record and one
            // head publication per committing lane, in lane order.
            uint64_t it = committed;
            while (it) {
                const uint32_t i = (uint32_t)__builtin_ctzll(it);
                it &= it - 1;
                ++last;
                if (lane == i) {
                    myts = last;
                    if (vmode == kValStamp) {
#pragma unroll
                        for (int w = 0; w < W; ++w)
                            if (wa[w] != kNone) {
                                if (c.stamp_in_lds) lstamp[wa[w]] = myts;
                                else c.stamps[wa[w]] = myts;
                            }
                    } else {
#pragma unroll
                        for (int w = 0; w < W; ++w) atr[w * E + (myts & emask)] = wa[w];
                    }
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
        st_rlx(&c.reply[(size_t)cw * kWave + lane], ((uint64_t)(n + 1) << 32) | myts);
        if (!server_wb && nc == 0 && lane == 0)
            st_rlx(&c.done[n & c.qmask], ((uint64_t)(n + 1) << 32) | last);
    }
    if (lane == 0) {
        lds[kLdsCfinal] = n;
        lds_st_rel(&lds[kLdsCdone], 1u);
        fetch_add(&c.stats[kStatAbortRw], n_rw);
        fetch_add(&c.stats[kStatAbortPrevalid], n_pv);
        fetch_add(&c.stats[kStatAbortRecord], n_rec);
        fetch_add(&c.stats[kStatBatches], (uint64_t)n);
        st_rlx(&c.stats[kStatFinalTs], (uint64_t)last);
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
template <int R, int W>
__device__ __forceinline__ void receiver(const Ctx &c, uint32_t *lds, uint32_t rid, uint32_t M)
{
    constexpr uint32_t rows = mbox_rows<R, W>();
    const uint32_t rows_copy = (c.variant & kServerWriteback) ? rows : 1 + R + W;
    const uint32_t lane = lane_id();
    uint32_t *mbst = lds + kLdsHdr;
    uint32_t *mbwave = lds + kLdsHdr + c.nmb;
    for (uint32_t n = rid;; n += M) {
        const uint32_t mbi = n % c.nmb;
        if (n >= c.nmb) {
            Spin sp;
            bool bad = false;
            while (lds_ld_acq(&lds[kLdsCnext]) < n - c.nmb + 1)
                if (!sp.tick(c, kErrSlotTimeout)) { bad = true; break; }
            if (bad) return;
        }
        uint64_t e;
        {
            Spin sp;
            for (;;) {
                e = ld_rlx(&c.ring[n & c.qmask]);
                if ((e >> 32) == (uint64_t)n + 1) break;
                if (ld_acq(c.clients_done) == c.n_client_waves && ld_acq(c.req_tail) <= n) {
                    // every client has left and request n will never come
                    if (ld_rlx(c.req_tail) == n && lane == 0)
                        lds_st_rel(&mbst[mbi], kExitBit | (n + 1));
                    return;
                }
                if (!sp.tick(c, kErrRingTimeout)) return;
            }
        }
        fence_acq();
        const uint32_t cw = (uint32_t)e;
        const uint32_t *src = c.mbox + (size_t)cw * rows * kWave;
        uint32_t *dst = lds + c.lds_mb_off + mbi * rows * kWave;
        for (uint32_t row = 0; row < rows_copy; ++row)
            dst[row * kWave + lane] = ld_rlx(&src[row * kWave + lane]);
        if (lane == 0) mbwave[mbi] = cw;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (lane == 0) lds_st_rel(&mbst[mbi], n + 1);
    }
}

template <int R, int W> __device__ __forceinline__ void server_workgroup(const Ctx &c, uint32_t *lds)
{
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t hdr = c.lds_mb_off;
    for (uint32_t i = threadIdx.x; i < hdr; i += blockDim.x) lds[i] = 0;
    if (c.stamp_in_lds)
        for (uint32_t i = threadIdx.x; i < c.n_items; i += blockDim.x) lds[c.lds_stamp_off + i] = 0;
    __syncthreads();
    if (wave == 0) committer<R, W>(c, lds);
    else if (wave == 1) {
        if ((c.variant & kServerWriteback) == 0) advancer(c, lds);
    } else receiver<R, W>(c, lds, wave - 2, c.wg_waves - 2);
}

}  // namespace csmv
