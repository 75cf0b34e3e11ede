// logfile.cpp -- on-disk commit logs and abort logs (host code, no device).
//
// The reference keeps no log: a run ends with the statistics printf
// (bank/csmv_sh/exe.cu:470-500).  The engine's logs (commit timestamps, read values,
// failed attempts, per-class abort totals, per-bucket abort counters) are device arrays
// read back through the C-ABI; this file gives them a file format: a 64-byte header, then
// little-endian arrays (layout in csmv.h).
#include <cstdio>
#include <cstring>

#include "csmv.h"

namespace {

constexpr char kMagic[8] = {'C', 'S', 'M', 'V', 'L', 'O', 'G', '\0'};
static_assert(sizeof(csmv_log_header) == 64, "the header is 64 bytes");
static_assert(sizeof(csmv_stats) > 0, "");

// The format is little-endian; gfx950 hosts (x86-64) are too, so arrays are written as
// they lie in memory.
constexpr bool host_little_endian()
{
    return __BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__;
}
static_assert(host_little_endian(), "log files are written on little-endian hosts");

struct Fnv {
    uint64_t h = 0xcbf29ce484222325ull;
    void add(const void *p, size_t n)
    {
        const uint8_t *b = (const uint8_t *)p;
        for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 0x100000001b3ull;
    }
};

struct Part {
    const void *src;   // writer: the array
    void *dst;         // reader: where it goes (nullptr: header only)
    size_t bytes;
};

csmv_status write_file(const char *path, csmv_log_header h, const Part *parts, int n_parts)
{
    Fnv f;
    for (int i = 0; i < n_parts; ++i) f.add(parts[i].src, parts[i].bytes);
    h.checksum = f.h;
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return CSMV_E_IO;
    bool ok = std::fwrite(&h, sizeof h, 1, fp) == 1;
    for (int i = 0; ok && i < n_parts; ++i)
        ok = parts[i].bytes == 0 || std::fwrite(parts[i].src, 1, parts[i].bytes, fp) == parts[i].bytes;
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? CSMV_OK : CSMV_E_IO;
}

csmv_status read_header(FILE *fp, uint32_t kind, csmv_log_header *h)
{
    if (std::fread(h, sizeof *h, 1, fp) != 1) return CSMV_E_IO;
    if (std::memcmp(h->magic, kMagic, 8) != 0 || h->version != CSMV_LOG_VERSION || h->kind != kind)
        return CSMV_E_IO;
    return CSMV_OK;
}

// Reads the arrays into the parts (every dst non-null) and checks the size and checksum.
csmv_status read_parts(FILE *fp, const csmv_log_header &h, const Part *parts, int n_parts)
{
    Fnv f;
    for (int i = 0; i < n_parts; ++i) {
        if (parts[i].bytes && std::fread(parts[i].dst, 1, parts[i].bytes, fp) != parts[i].bytes)
            return CSMV_E_IO;
        f.add(parts[i].dst, parts[i].bytes);
    }
    uint8_t extra;
    if (std::fread(&extra, 1, 1, fp) != 0) return CSMV_E_IO;   // trailing bytes
    return f.h == h.checksum ? CSMV_OK : CSMV_E_IO;
}

csmv_log_header header(uint32_t kind, uint64_t seed, uint64_t n, uint32_t rows, uint32_t buckets)
{
    csmv_log_header h;
    std::memset(&h, 0, sizeof h);
    std::memcpy(h.magic, kMagic, 8);
    h.version = CSMV_LOG_VERSION;
    h.kind = kind;
    h.n = n;
    h.rows = rows;
    h.buckets = buckets;
    h.seed = seed;
    return h;
}

// Opens a log of the given kind, returns its header, and reads its arrays when the caller
// passed every destination (all of them null: header only; some null: argument error).
template <int N>
csmv_status read_file(const char *path, uint32_t kind, csmv_log_header *hdr,
                      Part (&parts)[N], size_t (*sizes)(const csmv_log_header &, int))
{
    if (!path) return CSMV_E_ARG;
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return CSMV_E_IO;
    csmv_log_header h;
    csmv_status s = read_header(fp, kind, &h);
    int given = 0;
    for (int i = 0; i < N; ++i) given += parts[i].dst != nullptr;
    if (s == CSMV_OK && given != 0 && given != N) s = CSMV_E_ARG;
    if (s == CSMV_OK && given == N) {
        for (int i = 0; i < N; ++i) parts[i].bytes = sizes(h, i);
        s = read_parts(fp, h, parts, N);
    }
    std::fclose(fp);
    if (s == CSMV_OK && hdr) *hdr = h;
    return s;
}

constexpr uint64_t kMaxEntries = 1ull << 40;

}  // namespace

extern "C" {

csmv_status csmv_log_write_bank(const char *path, uint64_t seed, uint64_t n, uint32_t rows,
                                const uint32_t *ts, const int32_t *v, const uint32_t *aborts)
{
    if (!path || !ts || !v || !aborts || n == 0 || n > kMaxEntries || rows == 0 || rows > 64)
        return CSMV_E_ARG;
    const Part parts[3] = {{ts, nullptr, n * 4}, {v, nullptr, (size_t)rows * n * 4}, {aborts, nullptr, n * 4}};
    return write_file(path, header(CSMV_LOG_BANK, seed, n, rows, 0), parts, 3);
}

csmv_status csmv_log_read_bank(const char *path, csmv_log_header *hdr, uint32_t *ts, int32_t *v,
                               uint32_t *aborts)
{
    Part parts[3] = {{nullptr, ts, 0}, {nullptr, v, 0}, {nullptr, aborts, 0}};
    return read_file(path, CSMV_LOG_BANK, hdr, parts, [](const csmv_log_header &h, int i) -> size_t {
        return i == 1 ? (size_t)h.rows * h.n * 4 : h.n * 4;
    });
}

csmv_status csmv_log_write_cache(const char *path, uint64_t seed, uint64_t n, const uint32_t *ts,
                                 const uint32_t *res, const uint32_t *vsum, const uint32_t *aborts)
{
    if (!path || !ts || !res || !vsum || !aborts || n == 0 || n > kMaxEntries) return CSMV_E_ARG;
    const Part parts[4] = {{ts, nullptr, n * 4}, {res, nullptr, n * 4}, {vsum, nullptr, n * 4},
                           {aborts, nullptr, n * 4}};
    return write_file(path, header(CSMV_LOG_CACHE, seed, n, 0, 0), parts, 4);
}

csmv_status csmv_log_read_cache(const char *path, csmv_log_header *hdr, uint32_t *ts, uint32_t *res,
                                uint32_t *vsum, uint32_t *aborts)
{
    Part parts[4] = {{nullptr, ts, 0}, {nullptr, res, 0}, {nullptr, vsum, 0}, {nullptr, aborts, 0}};
    return read_file(path, CSMV_LOG_CACHE, hdr, parts,
                     [](const csmv_log_header &h, int) -> size_t { return h.n * 4; });
}

csmv_status csmv_log_write_aborts(const char *path, uint64_t seed, const csmv_stats *st, uint64_t n,
                                  const uint32_t *attempts, const uint32_t *counters, uint32_t buckets)
{
    if (!path || !st || !attempts || n == 0 || n > kMaxEntries || (buckets && !counters) ||
        (buckets != 0 && buckets != CSMV_ABORT_BUCKETS))
        return CSMV_E_ARG;
    const uint64_t classes[8] = {st->commits_update,   st->commits_ro,       st->aborts_dataage,
                                 st->aborts_readwrite, st->aborts_prevalid,  st->aborts_recordage,
                                 st->aborts_unknown,   st->escalated};
    const Part parts[3] = {{classes, nullptr, sizeof classes}, {attempts, nullptr, n * 4},
                           {counters, nullptr, (size_t)buckets * 4}};
    return write_file(path, header(CSMV_LOG_ABORTS, seed, n, 0, buckets), parts, 3);
}

csmv_status csmv_log_read_aborts(const char *path, csmv_log_header *hdr, uint64_t *classes,
                                 uint32_t *attempts, uint32_t *counters)
{
    // counters may stay null for a log without them
    if (!path) return CSMV_E_ARG;
    csmv_log_header h;
    Part head[1] = {{nullptr, nullptr, 0}};
    csmv_status s = read_file(path, CSMV_LOG_ABORTS, &h, head,
                              [](const csmv_log_header &, int) -> size_t { return 0; });
    if (s) return s;
    if (hdr) *hdr = h;
    if (!classes && !attempts && !counters) return CSMV_OK;
    if (!classes || !attempts || (h.buckets && !counters)) return CSMV_E_ARG;
    uint32_t dummy = 0;
    Part parts[3] = {{nullptr, classes, 0}, {nullptr, attempts, 0}, {nullptr, h.buckets ? counters : &dummy, 0}};
    return read_file(path, CSMV_LOG_ABORTS, nullptr, parts, [](const csmv_log_header &hh, int i) -> size_t {
        return i == 0 ? 64 : i == 1 ? hh.n * 4 : (size_t)hh.buckets * 4;
    });
}

}  // extern "C"
