// c1_driver.hip -- a bank driver written only against the engine's public headers
// (include/csmv.h, csmv_device.hpp, csmv_server.hpp), the way the reference's bank driver
// is written against its device API: client_kernel calls TXBegin / TXRead / TXReadOnly /
// TXWrite and commits beside the server kernel (bank/csmv_sh/exe.cu:89-338).  Here both
// are one persistent kernel: workgroup 0 runs the engine's commit server, the other waves
// run the transactions of a resident bank stream with one transaction per lane.
//
// Test code: it shows that a caller's own kernel can run the engine's transactions
// through csmv_kernel_prepare / csmv_kernel_finish, and it exercises the per-item
// tx_read / tx_read_only of the device API.
#include <hip/hip_runtime.h>

#include "csmv.h"
#include "csmv_device.hpp"
#include "csmv_server.hpp"

using namespace csmv;

namespace {

// the bank-plain shape: transfers read and write two accounts
constexpr int kR = 2, kW = 2;

struct Stream {
    const uint32_t *op;    // kind | amount << 8 | reads << 16 (the engine's bank stream format)
    const uint32_t *addr;  // [rmax][n]: read j of transaction k at addr[j * n + k]
    uint64_t n;
    uint64_t *next;        // next stream index to hand out
    uint32_t *log_ts;      // commit ts (audit: snapshot)
    int32_t *log_v;        // [2][n] transfer reads; log_v[k] = audit sum
    uint32_t *log_ab;      // failed attempts
};

__global__ void __launch_bounds__(512) c1_kernel(Ctx c, ServerLayout L, Stream s)
{
    extern __shared__ uint32_t lds[];
    if (blockIdx.x == 0) {
        server_workgroup<kR, kW, false>(c, L, lds);
        return;
    }
    const uint32_t cw = client_wave(c);
    if (cw == kNone) return;
    const uint32_t lane = lane_id();
    bool has = false;
    uint64_t k = 0;
    uint32_t op = 0, ab = 0, last = 0;
    uint64_t n_upd = 0, n_ro = 0, n_age = 0, wait = 0;
    for (;;) {
        if (ld_rlx(c.err)) break;
        // lanes without a transaction take the next ones (one fetch_add per wave)
        const uint64_t need = ballot(!has);
        if (need) {
            uint64_t base = 0;
            if (lane == 0) base = fetch_add(s.next, (uint64_t)__builtin_popcountll(need));
            base = bcast64(base);
            if (!has) {
                const uint64_t mine = base + prefix_count(need);
                if (mine < s.n) {
                    has = true;
                    k = mine;
                    op = s.op[k];
                    ab = 0;
                }
            }
        }
        if (ballot(has) == 0) break;
        // TXBegin with the prevTs wait: the snapshot covers the wave's own last commit
        Tx<kR, kW> tx;
        if (!tx_begin(c, tx, last)) break;
        const uint32_t kind = op & 0xFFu, amt = (op >> 8) & 0xFFu, nr = (op >> 16) & 0xFFu;
        bool req = false;
        int32_t va = 0, vb = 0;
        if (has && kind == kKindAudit) {
            // audit: read-only, committed at its snapshot without a round trip
            uint32_t sum = 0;
            for (uint32_t j = 0; j < nr && !tx.aborted; ++j)
                sum += (uint32_t)tx_read_only(c, tx, s.addr[(uint64_t)j * s.n + k]);
            if (!tx.aborted) {
                s.log_ts[k] = tx.snap;
                s.log_v[k] = (int32_t)sum;
                s.log_ab[k] = ab;
                has = false;
                ++n_ro;
            } else {
                ++ab;
                ++n_age;
            }
        } else if (has) {
            // transfer: read a, write a - amt, read b (the own write when b == a), write b + amt
            const uint32_t a = s.addr[k], b = s.addr[s.n + k];
            va = tx_read(c, tx, a);
            tx_write(tx, a, (int32_t)((uint32_t)va - amt));
            vb = tx_read(c, tx, b);
            tx_write(tx, b, (int32_t)((uint32_t)vb + amt));
            if (tx.aborted) {
                ++ab;
                ++n_age;
            } else {
                req = true;
            }
        }
        if (ballot(req)) {
            bool fail = false;
            const uint32_t ts = tx_commit<kR, kW>(c, cw, tx, req, &fail, &wait);
            if (fail) break;
            const uint64_t cm = ballot(ts != 0);
            if (cm) last = readlane(ts, 63u - (uint32_t)__builtin_clzll(cm));
            if (req) {
                if (ts) {
                    s.log_ts[k] = ts;
                    s.log_v[k] = va;
                    s.log_v[s.n + k] = vb;
                    s.log_ab[k] = ab;
                    has = false;
                    ++n_upd;
                } else {
                    ++ab;
                }
            }
        }
    }
    client_finish(c, n_upd, n_ro, n_age);
}

template <typename T> hipError_t upload(T **d, const T *h, size_t n)
{
    hipError_t r = hipMalloc((void **)d, n * sizeof(T));
    if (r == hipSuccess && h) r = hipMemcpy(*d, h, n * sizeof(T), hipMemcpyHostToDevice);
    return r;
}

}  // namespace

// Runs transactions [0, n) of a bank stream (host arrays op[n], addr[rmax][n]) with the
// kernel above on the engine e (pool freshly loaded); fills the log (log_ts[n],
// log_v[2][n], log_ab[n]) and the statistics.  Returns a csmv_status.
extern "C" int c1drv_run(csmv_engine *e, const uint32_t *op, const uint32_t *addr, uint64_t n,
                         uint32_t rmax, uint32_t *log_ts, int32_t *log_v, uint32_t *log_ab,
                         csmv_stats *stats)
{
    if (!e || !op || !addr || n == 0 || rmax < 2 || !log_ts || !log_v || !log_ab) return CSMV_E_ARG;
    Stream s{};
    uint32_t *d_op = nullptr, *d_addr = nullptr, *d_ts = nullptr, *d_ab = nullptr;
    int32_t *d_v = nullptr;
    uint64_t *d_next = nullptr;
    int st = CSMV_E_HIP;
    if (upload(&d_op, op, n) == hipSuccess && upload(&d_addr, addr, (size_t)rmax * n) == hipSuccess &&
        upload<uint64_t>(&d_next, nullptr, 1) == hipSuccess &&
        hipMemset(d_next, 0, 8) == hipSuccess && upload<uint32_t>(&d_ts, nullptr, n) == hipSuccess &&
        upload<int32_t>(&d_v, nullptr, 2 * n) == hipSuccess &&
        upload<uint32_t>(&d_ab, nullptr, n) == hipSuccess) {
        s.op = d_op;
        s.addr = d_addr;
        s.n = n;
        s.next = d_next;
        s.log_ts = d_ts;
        s.log_v = d_v;
        s.log_ab = d_ab;
        const csmv_kernel_spec spec{(const void *)c1_kernel, (uint32_t)kR, (uint32_t)kW};
        Ctx ctx;
        ServerLayout L;
        csmv_launch ln;
        st = csmv_kernel_prepare(e, &spec, &ctx, sizeof ctx, &L, sizeof L, &ln);
        if (st == CSMV_OK) {
            hipLaunchKernelGGL(c1_kernel, dim3(ln.grid), dim3(ln.block), ln.lds_bytes,
                               (hipStream_t)ln.stream, ctx, L, s);
            const hipError_t lr = hipGetLastError();
            st = csmv_kernel_finish(e, stats);
            if (lr != hipSuccess) st = CSMV_E_HIP;
        }
        if (st == CSMV_OK &&
            (hipMemcpy(log_ts, d_ts, n * 4, hipMemcpyDeviceToHost) != hipSuccess ||
             hipMemcpy(log_v, d_v, 2 * n * 4, hipMemcpyDeviceToHost) != hipSuccess ||
             hipMemcpy(log_ab, d_ab, n * 4, hipMemcpyDeviceToHost) != hipSuccess))
            st = CSMV_E_HIP;
    }
    (void)hipFree(d_op);
    (void)hipFree(d_addr);
    (void)hipFree(d_next);
    (void)hipFree(d_ts);
    (void)hipFree(d_v);
    (void)hipFree(d_ab);
    return st;
}
