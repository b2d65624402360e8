/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY: serial-replay oracle for the batched
 * transaction path.  Definitions and reference citations are in oracle.h.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

#define GOLDEN 0x9E3779B97F4A7C15ull

uint64_t orc_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t orc_stream(uint64_t s, uint32_t j) {
    return orc_mix64(s + (uint64_t)(j + 1) * GOLDEN);
}

uint64_t orc_cellseed(uint64_t seed, uint32_t table, uint32_t row, uint32_t col) {
    uint64_t id = ((uint64_t)row << 10) | ((uint64_t)table << 6) | (uint64_t)col;
    return orc_mix64(seed + GOLDEN * (id + 1));
}

void orc_bytes_from_stream(uint64_t s, uint32_t width, uint8_t *out) {
    /* little-endian bytes of stream(s,0), stream(s,1), ... */
    for (uint32_t w = 0; 8 * w < width; ++w) {
        uint64_t v = orc_stream(s, w);
        for (uint32_t k = 0; k < 8 && 8 * w + k < width; ++k) out[8 * w + k] = (uint8_t)(v >> (8 * k));
    }
}

void orc_fill_column(uint64_t seed, uint32_t table, uint32_t col, uint32_t width,
                     uint32_t row0, uint32_t nrows, uint8_t *out) {
    for (uint32_t i = 0; i < nrows; ++i)
        orc_bytes_from_stream(orc_cellseed(seed, table, row0 + i, col), width,
                              out + (uint64_t)i * width);
}

void orc_fill_rows(uint64_t seed, uint32_t table, uint32_t col, uint32_t width,
                   const uint32_t *rows, uint32_t n, uint8_t *col_base) {
    for (uint32_t i = 0; i < n; ++i)
        orc_bytes_from_stream(orc_cellseed(seed, table, rows[i], col), width,
                              col_base + (uint64_t)rows[i] * width);
}

static int op_fields(const orc_table *tables, uint32_t n_tables, uint32_t row, uint16_t m,
                     uint32_t *kind, uint32_t *tab, uint32_t *col) {
    *kind = m & 3u;
    *col = (m >> 2) & 63u;
    *tab = (m >> 8) & 15u;
    if ((m >> 12) != 0) return -1;
    if (*tab >= n_tables || tables[*tab].n_rows == 0) return -1;
    if (*kind == ORC_KIND_ROW) {
        if (!tables[*tab].present) return -1;
        if (*col == ORC_ROW_DELETE) return row < tables[*tab].n_rows ? 0 : -1;
        return (*col == ORC_ROW_INSERT && row == 0) ? 0 : -1;
    }
    if (*col >= tables[*tab].n_cols || row >= tables[*tab].n_rows) return -1;
    return 0;
}

static uint64_t rowseed(uint64_t arg, uint32_t col) {
    return orc_mix64(arg + (uint64_t)(col + 1) * 0xD1B54A32D192ED03ull);
}

/* Apply one op to the current image, in place (the serial step). */
static void apply_op(orc_table *t, uint32_t kind, uint32_t col, uint32_t row, uint64_t arg,
                     uint8_t *res_slot, uint32_t res_stride) {
    uint32_t w = t->width[col];
    uint8_t *cell = t->col[col] + (uint64_t)row * w;
    if (kind == ORC_KIND_READ) {
        memset(res_slot, 0, res_stride);
        memcpy(res_slot, cell, w);
    } else if (kind == ORC_KIND_UPDATE) {
        orc_bytes_from_stream(arg, w, cell);
    } else { /* RMW_ADD: every little-endian u32 word += (u32)arg */
        for (uint32_t b = 0; b < w; b += 4) {
            uint32_t v = (uint32_t)cell[b] | ((uint32_t)cell[b + 1] << 8) |
                         ((uint32_t)cell[b + 2] << 16) | ((uint32_t)cell[b + 3] << 24);
            v += (uint32_t)arg;
            cell[b] = (uint8_t)v;
            cell[b + 1] = (uint8_t)(v >> 8);
            cell[b + 2] = (uint8_t)(v >> 16);
            cell[b + 3] = (uint8_t)(v >> 24);
        }
    }
}

int orc_replay(orc_table *tables, uint32_t n_tables, uint32_t n_txns,
               const uint32_t *txn_off, const uint32_t *rows, const uint16_t *meta,
               const uint64_t *args, uint32_t mode, uint8_t *commit,
               uint8_t *results, uint32_t res_stride, uint8_t *found) {
    if (mode != ORC_MODE_A && mode != ORC_MODE_O) return -1;
    if (txn_off[0] != 0) return -1;
    for (uint32_t t = 0; t < n_txns; ++t)
        if (txn_off[t + 1] < txn_off[t]) return -1;
    uint32_t n_ops = txn_off[n_txns];
    for (uint32_t i = 0; i < n_ops; ++i) {
        uint32_t k, tb, c;
        if (op_fields(tables, n_tables, rows[i], meta[i], &k, &tb, &c)) return -2;
        if (k == ORC_KIND_READ && tables[tb].width[c] > res_stride) return -3;
        if (k == ORC_KIND_ROW && res_stride < 4) return -3;
    }
    if (found) memset(found, 0, n_ops);

    if (mode == ORC_MODE_O) {
        for (uint32_t t = 0; t < n_txns; ++t) commit[t] = 1;
    } else {
        /* W[k] = min TID writing row k, per table (the commented atomicMin rule). */
        uint32_t *W[16];
        for (uint32_t tb = 0; tb < n_tables; ++tb) {
            W[tb] = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(tables[tb].n_rows ? tables[tb].n_rows : 1));
            if (!W[tb]) return -4;
            memset(W[tb], 0xff, sizeof(uint32_t) * (size_t)tables[tb].n_rows);
        }
        for (uint32_t t = 0; t < n_txns; ++t)
            for (uint32_t i = txn_off[t]; i < txn_off[t + 1]; ++i) {
                uint32_t k, tb, c;
                op_fields(tables, n_tables, rows[i], meta[i], &k, &tb, &c);
                if (k == ORC_KIND_READ || (k == ORC_KIND_ROW && c == ORC_ROW_INSERT)) continue;
                if (W[tb][rows[i]] > t) W[tb][rows[i]] = t;
            }
        for (uint32_t t = 0; t < n_txns; ++t) {
            uint8_t ok = 1;
            for (uint32_t i = txn_off[t]; i < txn_off[t + 1]; ++i) {
                uint32_t k, tb, c;
                op_fields(tables, n_tables, rows[i], meta[i], &k, &tb, &c);
                if (k == ORC_KIND_ROW && c == ORC_ROW_INSERT) continue;  /* no existing row */
                if (W[tb][rows[i]] < t) ok = 0;
            }
            commit[t] = ok;
        }
        for (uint32_t tb = 0; tb < n_tables; ++tb) free(W[tb]);
    }

    /* Serial replay of the committed txns in ascending TID, ops in txn order.
     * INSERTs are collected and appended after the last txn, in op order. */
    uint32_t n_ins = 0;
    for (uint32_t t = 0; t < n_txns; ++t) {
        if (!commit[t]) continue;
        for (uint32_t i = txn_off[t]; i < txn_off[t + 1]; ++i) {
            uint32_t k, tb, c;
            op_fields(tables, n_tables, rows[i], meta[i], &k, &tb, &c);
            orc_table *tt = &tables[tb];
            uint8_t *slot = results + (uint64_t)i * res_stride;
            if (k == ORC_KIND_ROW) {
                if (c == ORC_ROW_INSERT) {
                    ++n_ins;
                } else if (tt->present[rows[i]]) {
                    tt->present[rows[i]] = 0;
                    if (found) found[i] = 1;
                }
                continue;
            }
            if (tt->present && !tt->present[rows[i]]) {
                if (k == ORC_KIND_READ) memset(slot, 0, res_stride);
                continue;
            }
            apply_op(tt, k, c, rows[i], args[i], slot, res_stride);
            if (found) found[i] = 1;
        }
    }
    int rc = 0;
    for (uint32_t t = 0; t < n_txns && n_ins; ++t) {
        if (!commit[t]) continue;
        for (uint32_t i = txn_off[t]; i < txn_off[t + 1]; ++i) {
            uint32_t k, tb, c;
            op_fields(tables, n_tables, rows[i], meta[i], &k, &tb, &c);
            if (k != ORC_KIND_ROW || c != ORC_ROW_INSERT) continue;
            orc_table *tt = &tables[tb];
            if (tt->cursor >= tt->n_rows) {
                rc = -5;
                continue;
            }
            const uint32_t row = tt->cursor++;
            for (uint32_t col = 0; col < tt->n_cols; ++col)
                orc_bytes_from_stream(rowseed(args[i], col), tt->width[col], tt->col[col] + (uint64_t)row * tt->width[col]);
            tt->present[row] = 1;
            uint8_t *slot = results + (uint64_t)i * res_stride;
            memset(slot, 0, res_stride);
            slot[0] = (uint8_t)row;
            slot[1] = (uint8_t)(row >> 8);
            slot[2] = (uint8_t)(row >> 16);
            slot[3] = (uint8_t)(row >> 24);
            if (found) found[i] = 1;
        }
    }
    return rc;
}

int orc_commit_pairwise(uint32_t n_txns, const uint32_t *txn_off, const uint32_t *rows,
                        const uint16_t *meta, uint8_t *commit) {
    for (uint32_t t = 0; t < n_txns; ++t) {
        uint8_t ok = 1;
        for (uint32_t s = 0; s < t && ok; ++s)
            for (uint32_t i = txn_off[s]; i < txn_off[s + 1] && ok; ++i) {
                /* READs and INSERTs write no existing row */
                if ((meta[i] & 3u) == ORC_KIND_READ || (meta[i] & 0xFFu) == ORC_KIND_ROW) continue;
                uint32_t tab_s = (meta[i] >> 8) & 15u;
                for (uint32_t j = txn_off[t]; j < txn_off[t + 1]; ++j)
                    if ((meta[j] & 0xFFu) != ORC_KIND_ROW && rows[j] == rows[i] &&
                        ((meta[j] >> 8) & 15u) == tab_s) { ok = 0; break; }
            }
        commit[t] = ok;
    }
    return 0;
}
