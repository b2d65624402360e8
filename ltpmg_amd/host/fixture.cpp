// fixture.cpp -- versioned binary batch fixtures (include/ltpmg_fixture.h).
#include <cstdio>
#include <cstring>

#include "ltpmg_fixture.h"
#include "rng.h"

namespace {

constexpr char kMagic[8] = {'L', 'T', 'P', 'M', 'G', 'F', 'X', '\0'};

struct File {
    std::FILE *f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

// mix64 chain over a byte range as little-endian 64-bit words; `word`
// holds a partial word between calls so the chain sees one byte stream
struct Checksum {
    uint64_t h = ltpmg_host::kGolden, word = 0;
    uint32_t fill = 0;
    void bytes(const void *p, size_t n) {
        const uint8_t *b = static_cast<const uint8_t *>(p);
        for (size_t i = 0; i < n; ++i) {
            word |= (uint64_t)b[i] << (8 * fill);
            if (++fill == 8) {
                h = ltpmg_host::mix64(h ^ word);
                word = 0;
                fill = 0;
            }
        }
    }
    uint64_t done() {
        if (fill) h = ltpmg_host::mix64(h ^ word);
        fill = 0;
        word = 0;
        return h;
    }
};

// zeros after meta until its bytes are a multiple of 8
size_t meta_pad(uint32_t n_ops) { return (8 - ((size_t)n_ops * 2) % 8) % 8; }

uint64_t arrays_checksum(const ltpmg_batch *b) {
    static const uint8_t zeros[8] = {};
    Checksum c;
    c.bytes(b->txn_off, ((size_t)b->n_txns + 1) * 4);
    if (b->n_ops) {
        c.bytes(b->rows, (size_t)b->n_ops * 4);
        c.bytes(b->meta, (size_t)b->n_ops * 2);
    }
    c.bytes(zeros, meta_pad(b->n_ops));
    if (b->n_ops) c.bytes(b->args, (size_t)b->n_ops * 8);
    return c.done();
}

bool put(std::FILE *f, const void *p, size_t n) { return n == 0 || std::fwrite(p, 1, n, f) == n; }
bool get(std::FILE *f, void *p, size_t n) { return n == 0 || std::fread(p, 1, n, f) == n; }

// Reads everything before the arrays; returns false on any mismatch.
bool read_header(std::FILE *f, ltpmg_fixture_info *info, uint32_t *n_txns, uint32_t *n_ops, uint64_t *sum) {
    char magic[8];
    uint32_t version = 0, header_bytes = 0;
    if (!get(f, magic, 8) || std::memcmp(magic, kMagic, 8) != 0) return false;
    if (!get(f, &version, 4) || version == 0 || version > LTPMG_FIXTURE_VERSION) return false;
    if (!get(f, &header_bytes, 4)) return false;
    ltpmg_fixture_info in;
    std::memset(&in, 0, sizeof in);
    if (!get(f, &in.seed, 8) || !get(f, &in.mode, 4) || !get(f, &in.n_tables, 4) ||
        in.n_tables > LTPMG_MAX_TABLES)
        return false;
    if (version >= 2 && (!get(f, &in.append_mask, 4) || (in.append_mask >> in.n_tables) != 0)) return false;
    for (uint32_t t = 0; t < in.n_tables; ++t) {
        ltpmg_table_desc &d = in.tables[t];
        if (!get(f, &d.n_rows, 4) || !get(f, &d.n_cols, 4) || d.n_cols > LTPMG_MAX_COLS ||
            !get(f, d.col_bytes, (size_t)d.n_cols * 4))
            return false;
        if (((in.append_mask >> t) & 1u) && (!get(f, &in.n_present[t], 4) || in.n_present[t] > d.n_rows))
            return false;
    }
    uint32_t nt = 0, no = 0;
    uint64_t s = 0;
    if (!get(f, &nt, 4) || !get(f, &no, 4) || !get(f, &s, 8)) return false;
    if ((uint64_t)std::ftell(f) != header_bytes) return false;
    if (info) *info = in;
    *n_txns = nt;
    *n_ops = no;
    *sum = s;
    return true;
}

}  // namespace

extern "C" int ltpmg_fixture_write(const char *path, const ltpmg_fixture_info *info, const ltpmg_batch *b) {
    if (!path || !info || !b || !b->txn_off || (b->n_ops && (!b->rows || !b->meta || !b->args)) ||
        info->n_tables > LTPMG_MAX_TABLES || (info->append_mask >> info->n_tables) != 0)
        return LTPMG_EINVAL;
    for (uint32_t t = 0; t < info->n_tables; ++t)
        if (info->tables[t].n_cols > LTPMG_MAX_COLS ||
            (((info->append_mask >> t) & 1u) && info->n_present[t] > info->tables[t].n_rows))
            return LTPMG_EINVAL;
    uint32_t header = 8 + 4 + 4 + 8 + 4 + 4 + 4 + 4 + 4 + 8;
    for (uint32_t t = 0; t < info->n_tables; ++t)
        header += 8 + 4 * info->tables[t].n_cols + (((info->append_mask >> t) & 1u) ? 4 : 0);
    File fl;
    fl.f = std::fopen(path, "wb");
    if (!fl.f) return LTPMG_EINVAL;
    const uint32_t version = LTPMG_FIXTURE_VERSION;
    const uint64_t sum = arrays_checksum(b);
    static const uint8_t zeros[8] = {};
    bool ok = put(fl.f, kMagic, 8) && put(fl.f, &version, 4) && put(fl.f, &header, 4) &&
              put(fl.f, &info->seed, 8) && put(fl.f, &info->mode, 4) && put(fl.f, &info->n_tables, 4) &&
              put(fl.f, &info->append_mask, 4);
    for (uint32_t t = 0; ok && t < info->n_tables; ++t) {
        const ltpmg_table_desc &d = info->tables[t];
        ok = put(fl.f, &d.n_rows, 4) && put(fl.f, &d.n_cols, 4) && put(fl.f, d.col_bytes, (size_t)d.n_cols * 4);
        if (ok && ((info->append_mask >> t) & 1u)) ok = put(fl.f, &info->n_present[t], 4);
    }
    ok = ok && put(fl.f, &b->n_txns, 4) && put(fl.f, &b->n_ops, 4) && put(fl.f, &sum, 8) &&
         put(fl.f, b->txn_off, ((size_t)b->n_txns + 1) * 4) && put(fl.f, b->rows, (size_t)b->n_ops * 4) &&
         put(fl.f, b->meta, (size_t)b->n_ops * 2) && put(fl.f, zeros, meta_pad(b->n_ops)) &&
         put(fl.f, b->args, (size_t)b->n_ops * 8);
    ok = ok && std::fflush(fl.f) == 0;
    return ok ? LTPMG_OK : LTPMG_EINVAL;
}

extern "C" int ltpmg_fixture_probe(const char *path, ltpmg_fixture_info *info, uint32_t *n_txns, uint32_t *n_ops) {
    if (!path || !n_txns || !n_ops) return LTPMG_EINVAL;
    File fl;
    fl.f = std::fopen(path, "rb");
    uint64_t sum = 0;
    if (!fl.f || !read_header(fl.f, info, n_txns, n_ops, &sum)) return LTPMG_EINVAL;
    return LTPMG_OK;
}

extern "C" int ltpmg_fixture_read(const char *path, ltpmg_batch *b) {
    if (!path || !b || !b->txn_off || (b->n_ops && (!b->rows || !b->meta || !b->args))) return LTPMG_EINVAL;
    File fl;
    fl.f = std::fopen(path, "rb");
    uint32_t nt = 0, no = 0;
    uint64_t sum = 0;
    if (!fl.f || !read_header(fl.f, nullptr, &nt, &no, &sum) || nt != b->n_txns || no != b->n_ops)
        return LTPMG_EINVAL;
    // the batch struct's arrays are const for the engine; here the caller's
    // buffers are the destination
    uint8_t pad[8];
    bool ok = get(fl.f, const_cast<uint32_t *>(b->txn_off), ((size_t)nt + 1) * 4) &&
              get(fl.f, const_cast<uint32_t *>(b->rows), (size_t)no * 4) &&
              get(fl.f, const_cast<uint16_t *>(b->meta), (size_t)no * 2) && get(fl.f, pad, meta_pad(no)) &&
              get(fl.f, const_cast<uint64_t *>(b->args), (size_t)no * 8);
    if (!ok || std::fgetc(fl.f) != EOF) return LTPMG_EINVAL;
    return arrays_checksum(b) == sum ? LTPMG_OK : LTPMG_EINVAL;
}
