/*
 * csmv_oracle.c -- CPU oracle for the CSMV hot path (test infrastructure).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this code, and only as the checker or the timed CPU baseline.
 * See csmv_oracle.h for the list of functions and what each restates.
 */
#define _GNU_SOURCE
#include "csmv_oracle.h"

#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define NONE32 0xFFFFFFFFu

/* ---- random streams ---------------------------------------------------- */

/* SplitMix64 output function (Steele, Lea, Flood, OOPSLA 2014): the value
 * returned by the generator seeded with x on its first call. */
uint64_t orc_splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* RAND_R_FNC, bank/csmv_sh/API.cuh:37-53: three steps of the LCG
 * next = next*1103515245 + 12345 (mod 2^64), combining bits 16..26,
 * 16..25 and 16..25 of the three states into an 11+10+10-bit value. */
uint32_t orc_rand_r_fnc(uint64_t *state)
{
    uint64_t next = *state;
    uint64_t result;
    next = next * 1103515245ull + 12345ull;
    result = (next / 65536) % 2048;
    next = next * 1103515245ull + 12345ull;
    result <<= 10;
    result ^= (next / 65536) % 1024;
    next = next * 1103515245ull + 12345ull;
    result <<= 10;
    result ^= (next / 65536) % 1024;
    *state = next;
    return (uint32_t)result;
}

/* The reference seeds each client thread with seed + tid
 * (bank/csmv_sh/exe.cu:105); adjacent LCG seeds give correlated draws, so a
 * transaction's state is the splitmix64 image of seed + k instead. */
uint64_t orc_txn_state(uint64_t seed, uint64_t k)
{
    return orc_splitmix64(seed + k);
}

int orc_zipf_cdf(uint32_t n, double theta, uint32_t *cdf)
{
    if (n == 0 || !cdf) return -1;
    double *h = (double *)malloc(sizeof(double) * n);
    if (!h) return -1;
    double acc = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        acc += pow((double)(i + 1), -theta);
        h[i] = acc;
    }
    const double total = h[n - 1];
    for (uint32_t i = 0; i + 1 < n; ++i) {
        double v = floor(2147483648.0 * (h[i] / total));
        if (v > 2147483648.0) v = 2147483648.0;
        cdf[i] = (uint32_t)v;
    }
    cdf[n - 1] = 0x80000000u;
    free(h);
    return 0;
}

uint32_t orc_zipf_rank(const uint32_t *cdf, uint32_t n, uint32_t r)
{
    uint32_t lo = 0, hi = n - 1;   /* answer in [lo, hi] */
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (r < cdf[mid]) hi = mid; else lo = mid + 1;
    }
    return lo;
}

uint32_t orc_scramble(uint32_t rank, uint32_t n)
{
    if (n && (n & (n - 1)) == 0) return (rank * 2654435761u) & (n - 1);
    return rank;
}

static inline uint32_t draw_item(uint64_t *st, uint32_t n, uint32_t dist,
                                 const uint32_t *cdf)
{
    uint32_t r = orc_rand_r_fnc(st);
    if (dist == ORC_DIST_ZIPF) return orc_scramble(orc_zipf_rank(cdf, n, r), n);
    return r % n;
}

/* ---- bank stream -------------------------------------------------------- */

static inline uint32_t rmax_of(const orc_bank_spec *s)
{
    return s->n_reads > s->n_reads_ro ? s->n_reads : s->n_reads_ro;
}

/* Transfer/audit choice and operand draws follow the reference driver
 * (bank/csmv_sh/exe.cu:134-139 for the read-only test rnd < prRead*0xFFFF,
 * exe.cu:184-201 for the two-account transfer); the amount is drawn as
 * 1 + RAND % amount_max, then (n_reads_min > 0) the read-set size as
 * n_reads_min + RAND % (n_reads - n_reads_min + 1). */
int orc_bank_gen(const orc_bank_spec *s, const uint32_t *cdf, uint32_t *op,
                 uint32_t *addr)
{
    if (!s || !op || !addr || s->n_accounts == 0 || s->n_reads < 2 ||
        s->n_reads > 255 || s->n_reads_ro > 255 || s->pct_ro > 100 ||
        s->amount_max == 0 || s->amount_max > 255 ||
        (s->n_reads_min && (s->n_reads_min < 2 || s->n_reads_min > s->n_reads)))
        return -1;
    if (s->dist == ORC_DIST_ZIPF && !cdf) return -1;
    const uint32_t rmax = rmax_of(s);
    const uint64_t T = s->n_txns;
    const uint32_t ro_thr = (uint32_t)((uint64_t)s->pct_ro * 0xFFFFu / 100u);
    for (uint64_t k = 0; k < T; ++k) {
        uint64_t st = orc_txn_state(s->seed, k);
        uint32_t rnd = orc_rand_r_fnc(&st) & 0xFFFFu;
        uint32_t kind = rnd < ro_thr ? ORC_KIND_AUDIT : ORC_KIND_TRANSFER;
        uint32_t amt = 0, nr;
        if (kind == ORC_KIND_TRANSFER) {
            amt = 1 + orc_rand_r_fnc(&st) % s->amount_max;
            nr = s->n_reads;
            /* read-set size uniform in n_reads_min..n_reads (the serving mix) */
            if (s->n_reads_min)
                nr = s->n_reads_min + orc_rand_r_fnc(&st) % (s->n_reads - s->n_reads_min + 1);
        } else {
            nr = s->n_reads_ro;
        }
        op[k] = kind | (amt << 8) | (nr << 16);
        for (uint32_t j = 0; j < rmax; ++j)
            addr[(uint64_t)j * T + k] =
                j < nr ? draw_item(&st, s->n_accounts, s->dist, cdf) : NONE32;
    }
    return 0;
}

/* A transaction body executed against a plain array (serial semantics).
 * Transfer (bank/csmv_sh/exe.cu:184-201 with amount amt):
 *   v0 = read(a); write(a, v0 - amt); v1 = read(b); write(b, v1 + amt);
 *   v_j = read(addr_j) for j >= 2.  Reads consult the write buffer first
 *   (read-own-write, bank/csmv_sh/API.cu:9-18).
 * Audit: sum of read(addr_j), 32-bit wrap-around. */
typedef struct {
    uint32_t n;
    uint32_t a[2];
    int32_t v[2];
} wbuf2;

static inline int32_t wb_lookup(const wbuf2 *w, uint32_t item, int *hit)
{
    for (uint32_t i = 0; i < w->n; ++i)
        if (w->a[i] == item) { *hit = 1; return w->v[i]; }
    *hit = 0;
    return 0;
}

static inline void wb_put(wbuf2 *w, uint32_t item, int32_t v)
{
    for (uint32_t i = 0; i < w->n; ++i)
        if (w->a[i] == item) { w->v[i] = v; return; }
    w->a[w->n] = item;
    w->v[w->n] = v;
    w->n++;
}

static inline int32_t add32(int32_t a, int32_t b)
{
    return (int32_t)((uint32_t)a + (uint32_t)b);
}

static void transfer_body(const int32_t *state, uint32_t amt, uint32_t nr,
                          const uint32_t *addr, uint64_t stride, uint64_t k,
                          int32_t *reads, wbuf2 *w)
{
    w->n = 0;
    for (uint32_t j = 0; j < nr; ++j) {
        uint32_t it = addr[(uint64_t)j * stride + k];
        int hit;
        int32_t v = wb_lookup(w, it, &hit);
        if (!hit) v = state[it];
        reads[j] = v;
        if (j == 0) wb_put(w, it, add32(v, -(int32_t)amt));
        else if (j == 1) wb_put(w, it, add32(v, (int32_t)amt));
    }
}

static int32_t audit_body(const int32_t *state, uint32_t nr,
                          const uint32_t *addr, uint64_t stride, uint64_t k)
{
    uint32_t sum = 0;
    for (uint32_t j = 0; j < nr; ++j)
        sum += (uint32_t)state[addr[(uint64_t)j * stride + k]];
    return (int32_t)sum;
}

int orc_bank_serial(const orc_bank_spec *s, const uint32_t *op,
                    const uint32_t *addr, int32_t init, const uint64_t *order,
                    int32_t *state, uint32_t *log_ts, int32_t *log_v)
{
    const uint64_t T = s->n_txns;
    for (uint32_t i = 0; i < s->n_accounts; ++i) state[i] = init;
    uint32_t ts = 0;
    int32_t reads[256];
    for (uint64_t q = 0; q < T; ++q) {
        uint64_t k = order ? order[q] : q;
        uint32_t kind = op[k] & 0xFF, amt = (op[k] >> 8) & 0xFF,
                 nr = (op[k] >> 16) & 0xFF;
        if (kind == ORC_KIND_AUDIT) {
            int32_t sum = audit_body(state, nr, addr, T, k);
            if (log_ts) log_ts[k] = ts;
            if (log_v) log_v[k] = sum;
        } else {
            wbuf2 w;
            transfer_body(state, amt, nr, addr, T, k, reads, &w);
            for (uint32_t i = 0; i < w.n; ++i) state[w.a[i]] = w.v[i];
            ++ts;
            if (log_ts) log_ts[k] = ts;
            if (log_v)
                for (uint32_t j = 0; j < nr; ++j) log_v[(uint64_t)j * T + k] = reads[j];
        }
    }
    return 0;
}

/* Serial replay in commit-timestamp order.  Closed-form parity: committed
 * update transactions carry the dense timestamps 1..U (each exactly once);
 * transaction t's logged reads must equal the state after t-1 commits, an
 * audit with snapshot s must see the state after s commits, and the final
 * state must equal the device's head values (TXEnd, bank/csmv_sh/API.cu:335-350,
 * checks only the total). */
int64_t orc_bank_replay(const orc_bank_spec *s, const uint32_t *op,
                        const uint32_t *addr, int32_t init,
                        const uint32_t *log_ts, const int32_t *log_v,
                        const int32_t *final, int32_t *replayed,
                        orc_replay_report *rep)
{
    const uint64_t T = s->n_txns;
    const uint32_t N = s->n_accounts;
    memset(rep, 0, sizeof(*rep));
    rep->first_bad_txn = -1;
    uint64_t U = 0, A = 0;
    for (uint64_t k = 0; k < T; ++k) {
        if ((op[k] & 0xFF) == ORC_KIND_AUDIT) ++A; else ++U;
    }
    rep->n_updates = U;
    rep->n_audits = A;
#define BAD(field, k) do { rep->field++; if (rep->first_bad_txn < 0) rep->first_bad_txn = (int64_t)(k); } while (0)
    uint64_t *order = (uint64_t *)malloc(sizeof(uint64_t) * (U + 1));
    uint64_t *acount = (uint64_t *)calloc(U + 2, sizeof(uint64_t));
    uint64_t *alist = (uint64_t *)malloc(sizeof(uint64_t) * (A + 1));
    int32_t *state = (int32_t *)malloc(sizeof(int32_t) * N);
    if (!order || !acount || !alist || !state) {
        free(order); free(acount); free(alist); free(state);
        return -1;
    }
    for (uint64_t t = 0; t <= U; ++t) order[t] = UINT64_MAX;
    for (uint64_t k = 0; k < T; ++k) {
        uint32_t ts = log_ts[k];
        if ((op[k] & 0xFF) == ORC_KIND_AUDIT) {
            if (ts > U) BAD(bad_snapshot, k); else acount[ts + 1]++;
        } else {
            if (ts == 0 || ts > U || order[ts] != UINT64_MAX) BAD(bad_ts, k);
            else order[ts] = k;
        }
    }
    /* bucket audits by snapshot (counting sort) */
    for (uint64_t t = 1; t <= U + 1; ++t) acount[t] += acount[t - 1];
    {
        uint64_t *fill = (uint64_t *)malloc(sizeof(uint64_t) * (U + 2));
        memcpy(fill, acount, sizeof(uint64_t) * (U + 2));
        for (uint64_t k = 0; k < T; ++k)
            if ((op[k] & 0xFF) == ORC_KIND_AUDIT && log_ts[k] <= U)
                alist[fill[log_ts[k]]++] = k;
        free(fill);
    }
    for (uint32_t i = 0; i < N; ++i) state[i] = init;
    int32_t reads[256];
    for (uint64_t t = 0; t <= U; ++t) {
        for (uint64_t q = acount[t]; q < acount[t + 1]; ++q) {
            uint64_t k = alist[q];
            int32_t sum = audit_body(state, (op[k] >> 16) & 0xFF, addr, T, k);
            if (sum != log_v[k]) BAD(bad_audits, k);
        }
        if (t == U) break;
        uint64_t k = order[t + 1];
        if (k == UINT64_MAX) continue;   /* hole already reported */
        uint32_t amt = (op[k] >> 8) & 0xFF, nr = (op[k] >> 16) & 0xFF;
        wbuf2 w;
        transfer_body(state, amt, nr, addr, T, k, reads, &w);
        int ok = 1;
        for (uint32_t j = 0; j < nr; ++j)
            if (reads[j] != log_v[(uint64_t)j * T + k]) ok = 0;
        if (!ok) BAD(bad_reads, k);
        for (uint32_t i = 0; i < w.n; ++i) state[w.a[i]] = w.v[i];
    }
    int64_t total = 0;
    for (uint32_t i = 0; i < N; ++i) {
        total += state[i];
        if (final && final[i] != state[i]) BAD(bad_final, i);
    }
    rep->total = total;
    rep->expected_total = (int64_t)N * (int64_t)init;
    if (replayed) memcpy(replayed, state, sizeof(int32_t) * N);
#undef BAD
    free(order); free(acount); free(alist); free(state);
    int64_t fails = (int64_t)(rep->bad_ts + rep->bad_snapshot + rep->bad_reads +
                              rep->bad_audits + rep->bad_final);
    if (rep->total != rep->expected_total) fails++;
    return fails;
}

/* ---- multi-threaded CPU CSMV ------------------------------------------- */
/*
 * Restates the reference protocol on host threads:
 *  - version rings per item (VertionedDataItem, bank/csmv_sh/API.cuh:61-67),
 *    kept as a packed (head|tail<<16) word and (version<<32|value) slots;
 *  - TXBegin takes the snapshot from the global clock (API.cu:275-291);
 *  - TXRead walks from head to tail for the newest version <= snapshot,
 *    aborting ("DataAge") when none is left (API.cu:5-41);
 *  - update transactions send (client id, snapshot) through a request ring
 *    to ONE server thread (sync_lib one-phase server, exe.cu:56-79);
 *  - the server validates the read set against records newer than the
 *    snapshot (TXValidate, API.cu:88-144; "RecordAge" when the record ring
 *    no longer covers the window), assigns the next timestamp and appends
 *    the write set (TXAddToRecord, API.cu:193-272);
 *  - the client installs its versions and advances the clock in commit order
 *    (TXWriteBack, API.cu:147-190).
 */
typedef struct {
    uint32_t snap;
    uint32_t nr, nw;
    uint32_t rs[256];
    uint32_t wa[2];
    int32_t wv[2];
} cpu_mailbox;

typedef struct {
    const orc_bank_spec *s;
    const uint32_t *op, *addr;
    uint64_t n_run;
    uint32_t N, D, E, Q, nc;
    _Atomic uint32_t *ht;
    _Atomic uint64_t *slot;
    _Atomic uint32_t gc;
    _Atomic uint64_t next_txn;
    _Atomic uint64_t req_tail;
    _Atomic uint64_t *req;
    _Atomic uint64_t *reply;
    cpu_mailbox *mb;
    _Atomic uint32_t clients_done;
    _Atomic int failed;
    uint32_t *log_ts;
    int32_t *log_v;
    _Atomic uint64_t c_upd, c_ro, a_age, a_rw, a_rec;
    uint32_t final_ts;
} cpu_ctx;

typedef struct {
    cpu_ctx *c;
    uint32_t id;
    int core;
} cpu_arg;

static inline void cpu_relax(unsigned *spins)
{
    if (++*spins > 256) { sched_yield(); *spins = 0; }
    else __builtin_ia32_pause();
}

static void pin_to(int core)
{
    if (core < 0) return;
    cpu_set_t set;
    CPU_ZERO(&set);
    CPU_SET(core, &set);
    pthread_setaffinity_np(pthread_self(), sizeof(set), &set);
}

static int cpu_read(cpu_ctx *c, uint32_t item, uint32_t snap, int32_t *out)
{
    uint32_t ht = atomic_load_explicit(&c->ht[item], memory_order_acquire);
    uint32_t h = ht & 0xFFFF, t = ht >> 16;
    for (uint32_t d = 0; d < c->D; ++d) {
        uint32_t j = (h + c->D - d) % c->D;
        uint64_t x = atomic_load_explicit(&c->slot[(uint64_t)j * c->N + item],
                                          memory_order_acquire);
        if ((uint32_t)(x >> 32) <= snap) { *out = (int32_t)(uint32_t)x; return 1; }
        if (j == t) break;
    }
    return 0;
}

static void cpu_writeback(cpu_ctx *c, const uint32_t *wa, const int32_t *wv,
                          uint32_t nw, uint32_t ts)
{
    uint32_t newht[2];
    for (uint32_t i = 0; i < nw; ++i) {
        uint32_t ht = atomic_load_explicit(&c->ht[wa[i]], memory_order_relaxed);
        uint32_t h = ht & 0xFFFF, t = ht >> 16;
        uint32_t n = (h + 1) % c->D;
        if (n == t) t = (t + 1) % c->D;
        atomic_store_explicit(&c->slot[(uint64_t)n * c->N + wa[i]],
                              ((uint64_t)ts << 32) | (uint32_t)wv[i],
                              memory_order_release);
        newht[i] = n | (t << 16);
    }
    for (uint32_t i = 0; i < nw; ++i)
        atomic_store_explicit(&c->ht[wa[i]], newht[i], memory_order_release);
}

static void *cpu_client(void *p)
{
    cpu_arg *a = (cpu_arg *)p;
    cpu_ctx *c = a->c;
    pin_to(a->core);
    const uint64_t T = c->s->n_txns;
    cpu_mailbox *mb = &c->mb[a->id];
    int32_t reads[256];
    for (;;) {
        uint64_t k = atomic_fetch_add_explicit(&c->next_txn, 1, memory_order_relaxed);
        if (k >= c->n_run || atomic_load_explicit(&c->failed, memory_order_relaxed)) break;
        uint32_t kind = c->op[k] & 0xFF, amt = (c->op[k] >> 8) & 0xFF,
                 nr = (c->op[k] >> 16) & 0xFF;
        for (;;) {
            uint32_t snap = atomic_load_explicit(&c->gc, memory_order_acquire);
            int ok = 1;
            if (kind == ORC_KIND_AUDIT) {
                uint32_t sum = 0;
                for (uint32_t j = 0; j < nr && ok; ++j) {
                    int32_t v;
                    ok = cpu_read(c, c->addr[(uint64_t)j * T + k], snap, &v);
                    sum += (uint32_t)v;
                }
                if (!ok) { atomic_fetch_add(&c->a_age, 1); continue; }
                c->log_ts[k] = snap;
                c->log_v[k] = (int32_t)sum;
                atomic_fetch_add_explicit(&c->c_ro, 1, memory_order_relaxed);
                break;
            }
            wbuf2 w;
            w.n = 0;
            for (uint32_t j = 0; j < nr && ok; ++j) {
                uint32_t it = c->addr[(uint64_t)j * T + k];
                int hit;
                int32_t v = wb_lookup(&w, it, &hit);
                if (!hit) ok = cpu_read(c, it, snap, &v);
                reads[j] = v;
                if (j == 0) wb_put(&w, it, add32(v, -(int32_t)amt));
                else if (j == 1) wb_put(&w, it, add32(v, (int32_t)amt));
            }
            if (!ok) { atomic_fetch_add(&c->a_age, 1); continue; }
            mb->snap = snap;
            mb->nr = nr;
            for (uint32_t j = 0; j < nr; ++j) mb->rs[j] = c->addr[(uint64_t)j * T + k];
            mb->nw = w.n;
            for (uint32_t i = 0; i < w.n; ++i) { mb->wa[i] = w.a[i]; mb->wv[i] = w.v[i]; }
            uint64_t n = atomic_fetch_add_explicit(&c->req_tail, 1, memory_order_relaxed);
            atomic_store_explicit(&c->req[n % c->Q], ((n + 1) << 32) | a->id,
                                  memory_order_release);
            uint64_t r;
            unsigned spins = 0;
            while (((r = atomic_load_explicit(&c->reply[a->id], memory_order_acquire)) >> 32) != n + 1) {
                if (atomic_load_explicit(&c->failed, memory_order_relaxed)) return NULL;
                cpu_relax(&spins);
            }
            uint32_t ts = (uint32_t)r;
            if (!ts) continue;   /* aborted: the server counted the class */
            cpu_writeback(c, w.a, w.v, w.n, ts);
            spins = 0;
            while (atomic_load_explicit(&c->gc, memory_order_acquire) != ts - 1) {
                if (atomic_load_explicit(&c->failed, memory_order_relaxed)) return NULL;
                cpu_relax(&spins);
            }
            atomic_store_explicit(&c->gc, ts, memory_order_release);
            c->log_ts[k] = ts;
            for (uint32_t j = 0; j < nr; ++j) c->log_v[(uint64_t)j * T + k] = reads[j];
            atomic_fetch_add_explicit(&c->c_upd, 1, memory_order_relaxed);
            break;
        }
    }
    atomic_fetch_add_explicit(&c->clients_done, 1, memory_order_release);
    return NULL;
}

static void *cpu_server(void *p)
{
    cpu_arg *a = (cpu_arg *)p;
    cpu_ctx *c = a->c;
    pin_to(a->core);
    const uint32_t E = c->E;
    uint32_t *atr = (uint32_t *)malloc(sizeof(uint32_t) * 2 * E);
    uint32_t last = 0;
    for (uint64_t n = 0;; ++n) {
        uint64_t e;
        unsigned spins = 0;
        for (;;) {
            e = atomic_load_explicit(&c->req[n % c->Q], memory_order_acquire);
            if ((e >> 32) == n + 1) break;
            if (atomic_load_explicit(&c->clients_done, memory_order_acquire) == c->nc &&
                atomic_load_explicit(&c->req_tail, memory_order_acquire) <= n) {
                c->final_ts = last;
                free(atr);
                return NULL;
            }
            cpu_relax(&spins);
        }
        uint32_t id = (uint32_t)e;
        cpu_mailbox *mb = &c->mb[id];
        uint32_t snap = mb->snap, result = 0;
        if (last - snap > E) {
            atomic_fetch_add_explicit(&c->a_rec, 1, memory_order_relaxed);
        } else {
            int conflict = 0;
            for (uint32_t t = last; t > snap && !conflict; --t) {
                const uint32_t *w = &atr[2 * (t % E)];
                for (uint32_t x = 0; x < 2 && !conflict; ++x) {
                    if (w[x] == NONE32) continue;
                    for (uint32_t j = 0; j < mb->nr; ++j)
                        if (mb->rs[j] == w[x]) { conflict = 1; break; }
                }
            }
            if (conflict) {
                atomic_fetch_add_explicit(&c->a_rw, 1, memory_order_relaxed);
            } else {
                result = ++last;
                uint32_t *w = &atr[2 * (result % E)];
                w[0] = mb->nw > 0 ? mb->wa[0] : NONE32;
                w[1] = mb->nw > 1 ? mb->wa[1] : NONE32;
            }
        }
        atomic_store_explicit(&c->reply[id], ((n + 1) << 32) | result, memory_order_release);
    }
}

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

int orc_bank_cpu_csmv(const orc_bank_spec *s, const uint32_t *op,
                      const uint32_t *addr, int32_t init, uint64_t n_run,
                      const orc_cpu_cfg *cfg, int32_t *state,
                      uint32_t *log_ts, int32_t *log_v, orc_cpu_stats *st)
{
    if (!s || !cfg || cfg->n_clients == 0 || cfg->depth < 1 || cfg->depth > 0xFFFF ||
        cfg->atr_entries == 0 || n_run > s->n_txns || s->n_reads > 256 ||
        s->n_reads_ro > 256)
        return -1;
    cpu_ctx c;
    memset(&c, 0, sizeof(c));
    c.s = s; c.op = op; c.addr = addr; c.n_run = n_run;
    c.N = s->n_accounts; c.D = cfg->depth; c.E = cfg->atr_entries;
    c.nc = cfg->n_clients;
    c.Q = 1;
    while (c.Q < 2 * c.nc) c.Q <<= 1;
    c.ht = (_Atomic uint32_t *)calloc(c.N, sizeof(uint32_t));
    c.slot = (_Atomic uint64_t *)calloc((size_t)c.N * c.D, sizeof(uint64_t));
    c.req = (_Atomic uint64_t *)calloc(c.Q, sizeof(uint64_t));
    c.reply = (_Atomic uint64_t *)calloc(c.nc, sizeof(uint64_t));
    c.mb = (cpu_mailbox *)calloc(c.nc, sizeof(cpu_mailbox));
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (c.nc + 1));
    cpu_arg *args = (cpu_arg *)malloc(sizeof(cpu_arg) * (c.nc + 1));
    if (!c.ht || !c.slot || !c.req || !c.reply || !c.mb || !th || !args) return -2;
    for (uint32_t i = 0; i < c.N; ++i) {
        c.ht[i] = 0;
        c.slot[i] = (uint32_t)init;   /* slot 0 = (version 0, init) */
    }
    c.log_ts = log_ts;
    c.log_v = log_v;
    /* pinning: the server on the first CPU the process may run on, client i on the
     * (i + 1)-th (wrapping when there are more threads than allowed CPUs) */
    int cpus[CPU_SETSIZE], ncpu = 0;
    {
        cpu_set_t allowed;
        CPU_ZERO(&allowed);
        if (sched_getaffinity(0, sizeof(allowed), &allowed) == 0)
            for (int k = 0; k < CPU_SETSIZE; ++k)
                if (CPU_ISSET(k, &allowed)) cpus[ncpu++] = k;
        if (ncpu == 0) cpus[ncpu++] = 0;
    }
    double t0 = now_s();
    args[0].c = &c; args[0].id = 0; args[0].core = cfg->pin ? cpus[0] : -1;
    pthread_create(&th[0], NULL, cpu_server, &args[0]);
    for (uint32_t i = 0; i < c.nc; ++i) {
        args[i + 1].c = &c; args[i + 1].id = i;
        args[i + 1].core = cfg->pin ? cpus[(i + 1) % (uint32_t)ncpu] : -1;
        pthread_create(&th[i + 1], NULL, cpu_client, &args[i + 1]);
    }
    for (uint32_t i = 0; i <= c.nc; ++i) pthread_join(th[i], NULL);
    double t1 = now_s();
    for (uint32_t i = 0; i < c.N; ++i) {
        uint32_t h = c.ht[i] & 0xFFFF;
        state[i] = (int32_t)(uint32_t)c.slot[(uint64_t)h * c.N + i];
    }
    if (st) {
        st->commits_update = c.c_upd; st->commits_ro = c.c_ro;
        st->aborts_dataage = c.a_age; st->aborts_readwrite = c.a_rw;
        st->aborts_recordage = c.a_rec; st->seconds = t1 - t0;
        st->final_ts = c.final_ts; st->threads = c.nc + 1;
    }
    free((void *)c.ht); free((void *)c.slot); free((void *)c.req);
    free((void *)c.reply); free(c.mb); free(th); free(args);
    return 0;
}

/* ---- memcached-style cache ------------------------------------------------ */

uint64_t orc_memcd_key(uint32_t rank)
{
    return orc_splitmix64(0x6D656D6364ull ^ ((uint64_t)rank << 1));
}

uint32_t orc_memcd_set_of(uint64_t key, uint32_t n_sets)
{
    return (uint32_t)(orc_splitmix64(key) & (uint64_t)(n_sets - 1));
}

uint32_t orc_memcd_value_word(uint64_t key, uint64_t k, uint32_t w)
{
    return (uint32_t)orc_splitmix64(key ^ (k * 8 + w));
}

/* GET/SET choice as the driver's rnd > prRead test (memcachedGPU/memcdSH.cu:335-343);
 * key rank from the Zipf table over the keyspace. */
int orc_memcd_gen(const orc_memcd_spec *s, const uint32_t *cdf, uint64_t *req_key,
                  uint32_t *req_op)
{
    if (!s || !req_key || !req_op || s->n_keys == 0 || s->n_keys >= (1ull << 31) ||
        s->pct_get > 100 || (s->dist == ORC_DIST_ZIPF && !cdf))
        return -1;
    const uint32_t get_thr = (uint32_t)((uint64_t)s->pct_get * 0xFFFFu / 100u);
    for (uint64_t k = 0; k < s->n_reqs; ++k) {
        uint64_t st = orc_txn_state(s->seed, k);
        uint32_t rnd = orc_rand_r_fnc(&st) & 0xFFFFu;
        req_op[k] = rnd < get_thr ? ORC_MC_GET : ORC_MC_SET;
        uint32_t r = orc_rand_r_fnc(&st);
        uint32_t rank = s->dist == ORC_DIST_ZIPF ? orc_zipf_rank(cdf, (uint32_t)s->n_keys, r)
                                                 : (uint32_t)(r % s->n_keys);
        req_key[k] = orc_memcd_key(rank);
    }
    return 0;
}

typedef struct {
    uint32_t res;      /* bit0 hit, bit1 wrote, bits 4..7 way */
    uint32_t vsum;
    uint32_t nw;
    uint32_t wi[12];   /* written items */
    int32_t wv[12];
} mc_out;

static inline uint32_t mc_vsum_acc(uint32_t acc, uint32_t j, uint32_t v)
{
    return acc + v * (2u * j + 1u);
}

/* One request against a plain item array (field-major, see the header). */
static void memcd_body(const orc_memcd_spec *s, const int32_t *it, uint64_t key, uint32_t op,
                       uint64_t k, mc_out *o)
{
    const uint32_t ns = s->n_sets * s->n_ways;
    const uint32_t set = orc_memcd_set_of(key, s->n_sets);
    const uint32_t lo = (uint32_t)key, hi = (uint32_t)(key >> 32);
    const uint32_t clock = (uint32_t)(k / s->batch + 1);
#define F(f, slot) ((uint32_t)(f) * ns + (slot))
    o->res = 0; o->vsum = 0; o->nw = 0;
    if (op == ORC_MC_GET) {
        for (uint32_t w = 0; w < s->n_ways; ++w) {
            uint32_t slot = set * s->n_ways + w;
            if (it[F(ORC_MC_STATE, slot)] != 1) continue;
            if ((uint32_t)it[F(ORC_MC_KLO, slot)] != lo || (uint32_t)it[F(ORC_MC_KHI, slot)] != hi)
                continue;
            o->res = 1u | (w << 4);
            uint32_t acc = 0;
            for (uint32_t j = 0; j < 8; ++j)
                acc = mc_vsum_acc(acc, j, (uint32_t)it[F(ORC_MC_V0 + j, slot)]);
            o->vsum = acc;
            if ((uint32_t)it[F(ORC_MC_TS, slot)] < clock) {
                o->res |= 2u;
                o->wi[0] = F(ORC_MC_TS, slot);
                o->wv[0] = (int32_t)clock;
                o->nw = 1;
            }
            return;
        }
        return;
    }
    /* SET: key present -> that way; else first invalid way; else least recent */
    int32_t found = -1, invalid = -1, lru = 0;
    uint32_t lru_ts = 0xFFFFFFFFu;
    for (uint32_t w = 0; w < s->n_ways; ++w) {
        uint32_t slot = set * s->n_ways + w;
        int32_t stv = it[F(ORC_MC_STATE, slot)];
        uint32_t klo = (uint32_t)it[F(ORC_MC_KLO, slot)], khi = (uint32_t)it[F(ORC_MC_KHI, slot)];
        uint32_t ts = (uint32_t)it[F(ORC_MC_TS, slot)];
        if (stv == 1 && klo == lo && khi == hi && found < 0) found = (int32_t)w;
        if (stv != 1 && invalid < 0) invalid = (int32_t)w;
        if (ts < lru_ts) { lru_ts = ts; lru = (int32_t)w; }
    }
    uint32_t w = found >= 0 ? (uint32_t)found : invalid >= 0 ? (uint32_t)invalid : (uint32_t)lru;
    uint32_t slot = set * s->n_ways + w;
    o->res = (found >= 0 ? 1u : 0u) | 2u | (w << 4);
    o->wi[0] = F(ORC_MC_KLO, slot); o->wv[0] = (int32_t)lo;
    o->wi[1] = F(ORC_MC_KHI, slot); o->wv[1] = (int32_t)hi;
    for (uint32_t j = 0; j < 8; ++j) {
        o->wi[2 + j] = F(ORC_MC_V0 + j, slot);
        o->wv[2 + j] = (int32_t)orc_memcd_value_word(key, k, j);
    }
    o->wi[10] = F(ORC_MC_TS, slot); o->wv[10] = (int32_t)clock;
    o->wi[11] = F(ORC_MC_STATE, slot); o->wv[11] = 1;
    o->nw = 12;
#undef F
}

int orc_memcd_serial(const orc_memcd_spec *s, const uint64_t *req_key,
                     const uint32_t *req_op, const uint64_t *order, int32_t *items,
                     uint32_t *log_ts, uint32_t *log_res, uint32_t *log_vsum)
{
    return orc_memcd_serial_from(s, NULL, req_key, req_op, order, items, log_ts, log_res,
                                 log_vsum);
}

int orc_memcd_serial_from(const orc_memcd_spec *s, const int32_t *init,
                          const uint64_t *req_key, const uint32_t *req_op,
                          const uint64_t *order, int32_t *items, uint32_t *log_ts,
                          uint32_t *log_res, uint32_t *log_vsum)
{
    const uint64_t n_items = (uint64_t)ORC_MC_FIELDS * s->n_sets * s->n_ways;
    if (init) memcpy(items, init, sizeof(int32_t) * n_items);
    else memset(items, 0, sizeof(int32_t) * n_items);
    uint32_t ts = 0;
    for (uint64_t q = 0; q < s->n_reqs; ++q) {
        uint64_t k = order ? order[q] : q;
        mc_out o;
        memcd_body(s, items, req_key[k], req_op[k], k, &o);
        for (uint32_t i = 0; i < o.nw; ++i) items[o.wi[i]] = o.wv[i];
        if (o.nw) ++ts;
        if (log_ts) log_ts[k] = ts;
        if (log_res) log_res[k] = o.res;
        if (log_vsum) log_vsum[k] = o.vsum;
    }
    return 0;
}

int64_t orc_memcd_replay(const orc_memcd_spec *s, const uint64_t *req_key,
                         const uint32_t *req_op, const uint32_t *log_ts,
                         const uint32_t *log_res, const uint32_t *log_vsum,
                         const int32_t *final, orc_memcd_report *rep)
{
    return orc_memcd_replay_from(s, NULL, req_key, req_op, log_ts, log_res, log_vsum, final,
                                 rep);
}

int64_t orc_memcd_replay_from(const orc_memcd_spec *s, const int32_t *init,
                              const uint64_t *req_key, const uint32_t *req_op,
                              const uint32_t *log_ts, const uint32_t *log_res,
                              const uint32_t *log_vsum, const int32_t *final,
                              orc_memcd_report *rep)
{
    const uint64_t R = s->n_reqs;
    const uint64_t n_items = (uint64_t)ORC_MC_FIELDS * s->n_sets * s->n_ways;
    memset(rep, 0, sizeof(*rep));
    rep->first_bad_req = -1;
#define BAD(field, k) do { rep->field++; if (rep->first_bad_req < 0) rep->first_bad_req = (int64_t)(k); } while (0)
    uint64_t U = 0;
    for (uint64_t k = 0; k < R; ++k) if (log_res[k] & 2u) ++U;
    rep->n_updates = U;
    rep->n_readonly = R - U;
    uint64_t *order = (uint64_t *)malloc(sizeof(uint64_t) * (U + 1));
    uint64_t *cnt = (uint64_t *)calloc(U + 2, sizeof(uint64_t));
    uint64_t *lst = (uint64_t *)malloc(sizeof(uint64_t) * (R - U + 1));
    int32_t *it = (int32_t *)calloc(n_items, sizeof(int32_t));
    if (!order || !cnt || !lst || !it) {
        free(order); free(cnt); free(lst); free(it);
        return -1;
    }
    if (init) memcpy(it, init, sizeof(int32_t) * n_items);
    for (uint64_t t = 0; t <= U; ++t) order[t] = UINT64_MAX;
    for (uint64_t k = 0; k < R; ++k) {
        uint32_t ts = log_ts[k];
        if (log_res[k] & 2u) {
            if (ts == 0 || ts > U || order[ts] != UINT64_MAX) BAD(bad_ts, k);
            else order[ts] = k;
        } else {
            if (ts > U) BAD(bad_snapshot, k); else cnt[ts + 1]++;
        }
    }
    for (uint64_t t = 1; t <= U + 1; ++t) cnt[t] += cnt[t - 1];
    {
        uint64_t *fill = (uint64_t *)malloc(sizeof(uint64_t) * (U + 2));
        memcpy(fill, cnt, sizeof(uint64_t) * (U + 2));
        for (uint64_t k = 0; k < R; ++k)
            if (!(log_res[k] & 2u) && log_ts[k] <= U) lst[fill[log_ts[k]]++] = k;
        free(fill);
    }
    for (uint64_t t = 0; t <= U; ++t) {
        for (uint64_t q = cnt[t]; q < cnt[t + 1]; ++q) {
            uint64_t k = lst[q];
            mc_out o;
            memcd_body(s, it, req_key[k], req_op[k], k, &o);
            if (o.nw) BAD(bad_kind, k);
            else if (o.res != log_res[k] || o.vsum != log_vsum[k]) BAD(bad_result, k);
            if (req_op[k] == ORC_MC_GET) { if (o.res & 1u) rep->hits++; else rep->misses++; }
        }
        if (t == U) break;
        uint64_t k = order[t + 1];
        if (k == UINT64_MAX) continue;
        mc_out o;
        memcd_body(s, it, req_key[k], req_op[k], k, &o);
        if (!o.nw) BAD(bad_kind, k);
        else if (o.res != log_res[k] || o.vsum != log_vsum[k]) BAD(bad_result, k);
        if (req_op[k] == ORC_MC_GET) { if (o.res & 1u) rep->hits++; else rep->misses++; }
        else rep->sets++;
        for (uint32_t i = 0; i < o.nw; ++i) it[o.wi[i]] = o.wv[i];
    }
    if (final)
        for (uint64_t i = 0; i < n_items; ++i)
            if (final[i] != it[i]) BAD(bad_final, i);
#undef BAD
    free(order); free(cnt); free(lst); free(it);
    return (int64_t)(rep->bad_ts + rep->bad_snapshot + rep->bad_result + rep->bad_kind +
                     rep->bad_final);
}
