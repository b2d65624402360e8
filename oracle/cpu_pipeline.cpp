// cpu_pipeline.cpp -- TEST INFRASTRUCTURE / CPU BASELINE ONLY.
//
// A multi-threaded C++ restatement of the reference's per-batch pipeline, run
// on host cores to give the reported CPU baseline.  The reference's own CPU
// executor (cuda/src/gpulauncher.cu:708-753, cuda/src/threadpool.cu) is
// disabled (launch commented out at gpulauncher.cu:771) and only handles the
// TEST benchmark's 0xffffffff ops, so it cannot be timed; this file restates
// the GPU pipeline's stages instead:
//
//   Mode O (every txn commits, per-row TID order):
//     histogram  cnt[row]++                      gpulauncher.cu:413-417 (register_cc)
//     scan       segment offset per row          gpulauncher.cu:1558-1589 (txn_prefix_offset)
//     scatter    op index into row segment       gpulauncher.cu:2106-2437 (*_analyze)
//     sort       each segment by TID             gpulauncher.cu:2439-2523 (filter_commit)
//     apply      each row's ops in order         gpulauncher.cu:2525-2700 (commit)
//   Mode A (lowest TID wins):
//     reserve    atomic min of writer TID/row    gpulauncher.cu:414 (commented atomicMin)
//     validate   commit iff no lower writer      gpulauncher.cu:2130-2142 (commented rule)
//     reads of committed txns from the snapshot, then their writes.
//
// Outputs equal orc_replay (checked in tests/test_oracle.py).
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

namespace {

template <class F>
void parallel_for(uint64_t n, uint32_t n_threads, F &&f) {
    if (n_threads <= 1 || n < 4096) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    uint64_t chunk = (n + n_threads - 1) / n_threads;
    for (uint32_t i = 0; i < n_threads; ++i) {
        uint64_t a = i * chunk, b = std::min<uint64_t>(n, a + chunk);
        if (a >= b) break;
        th.emplace_back([&f, a, b] { f(a, b); });
    }
    for (auto &t : th) t.join();
}

struct OpRef {
    uint32_t kind, tab, col;
};

inline OpRef decode(uint16_t m) { return {m & 3u, (m >> 8) & 15u, (m >> 2) & 63u}; }

void exec_op(orc_table *tables, const OpRef &o, uint32_t row, uint64_t arg, uint8_t *res,
             uint32_t res_stride) {
    orc_table &t = tables[o.tab];
    uint32_t w = t.width[o.col];
    uint8_t *cell = t.col[o.col] + (uint64_t)row * w;
    if (o.kind == ORC_KIND_READ) {
        std::memset(res, 0, res_stride);
        std::memcpy(res, cell, w);
    } else if (o.kind == ORC_KIND_UPDATE) {
        orc_bytes_from_stream(arg, w, cell);
    } else {
        for (uint32_t b = 0; b < w; b += 4) {
            uint32_t v;
            std::memcpy(&v, cell + b, 4);
            v += (uint32_t)arg;
            std::memcpy(cell + b, &v, 4);
        }
    }
}

}  // namespace

extern "C" double orc_cpu_pipeline(orc_table *tables, uint32_t n_tables, uint32_t n_txns,
                                   const uint32_t *txn_off, const uint32_t *rows,
                                   const uint16_t *meta, const uint64_t *args, uint32_t mode,
                                   uint8_t *commit, uint8_t *results, uint32_t res_stride,
                                   uint32_t n_threads) {
    if (n_threads == 0) n_threads = std::max(1u, std::thread::hardware_concurrency());
    const uint32_t n_ops = txn_off[n_txns];
    // global row id = table base + row, so one histogram covers all tables
    std::vector<uint64_t> base(n_tables + 1, 0);
    for (uint32_t t = 0; t < n_tables; ++t) base[t + 1] = base[t] + tables[t].n_rows;
    const uint64_t total_rows = base[n_tables];
    for (uint32_t t = 0; t < n_tables; ++t)
        if (tables[t].present) return -1.0;  // append tables: orc_replay only
    for (uint32_t i = 0; i < n_ops; ++i) {
        OpRef o = decode(meta[i]);
        if (o.kind > ORC_KIND_RMW_ADD || o.tab >= n_tables || rows[i] >= tables[o.tab].n_rows ||
            o.col >= tables[o.tab].n_cols)
            return -1.0;
    }
    // op -> txn, built outside the timed span (the GPU path derives it per tile)
    std::vector<uint32_t> op_txn(n_ops);
    for (uint32_t t = 0; t < n_txns; ++t)
        for (uint32_t i = txn_off[t]; i < txn_off[t + 1]; ++i) op_txn[i] = t;

    auto t0 = std::chrono::steady_clock::now();
    if (mode == ORC_MODE_O) {
        std::vector<std::atomic<uint32_t>> cnt(total_rows);
        parallel_for(total_rows, n_threads, [&](uint64_t a, uint64_t b) {
            for (uint64_t r = a; r < b; ++r) cnt[r].store(0, std::memory_order_relaxed);
        });
        parallel_for(n_ops, n_threads, [&](uint64_t a, uint64_t b) {
            for (uint64_t i = a; i < b; ++i) {
                OpRef o = decode(meta[i]);
                cnt[base[o.tab] + rows[i]].fetch_add(1, std::memory_order_relaxed);
            }
        });
        // exclusive scan (sequential: the reference allocates segments with one
        // global atomic, gpulauncher.cu:1580-1584)
        std::vector<uint32_t> off(total_rows + 1);
        uint32_t run = 0;
        for (uint64_t r = 0; r < total_rows; ++r) {
            off[r] = run;
            run += cnt[r].load(std::memory_order_relaxed);
        }
        off[total_rows] = run;
        std::vector<std::atomic<uint32_t>> cur(total_rows);
        parallel_for(total_rows, n_threads, [&](uint64_t a, uint64_t b) {
            for (uint64_t r = a; r < b; ++r) cur[r].store(off[r], std::memory_order_relaxed);
        });
        std::vector<uint32_t> seg(n_ops);
        parallel_for(n_ops, n_threads, [&](uint64_t a, uint64_t b) {
            for (uint64_t i = a; i < b; ++i) {
                OpRef o = decode(meta[i]);
                seg[cur[base[o.tab] + rows[i]].fetch_add(1, std::memory_order_relaxed)] = (uint32_t)i;
            }
        });
        // per-row sort by (TID, op index) == op index, then ordered apply
        parallel_for(total_rows, n_threads, [&](uint64_t a, uint64_t b) {
            uint32_t tb = 0;
            for (uint64_t r = a; r < b; ++r) {
                uint32_t s = off[r], e = off[r + 1];
                if (s == e) continue;
                std::sort(seg.begin() + s, seg.begin() + e);
                while (r >= base[tb + 1]) ++tb;
                uint32_t row = (uint32_t)(r - base[tb]);
                for (uint32_t k = s; k < e; ++k) {
                    uint32_t i = seg[k];
                    exec_op(tables, decode(meta[i]), row, args[i],
                            results + (uint64_t)i * res_stride, res_stride);
                }
            }
        });
        for (uint32_t t = 0; t < n_txns; ++t) commit[t] = 1;
    } else if (mode == ORC_MODE_A) {
        std::vector<std::atomic<uint32_t>> W(total_rows);
        parallel_for(total_rows, n_threads, [&](uint64_t a, uint64_t b) {
            for (uint64_t r = a; r < b; ++r) W[r].store(0xffffffffu, std::memory_order_relaxed);
        });
        parallel_for(n_ops, n_threads, [&](uint64_t a, uint64_t b) {
            for (uint64_t i = a; i < b; ++i) {
                OpRef o = decode(meta[i]);
                if (o.kind == ORC_KIND_READ) continue;
                std::atomic<uint32_t> &w = W[base[o.tab] + rows[i]];
                uint32_t t = op_txn[i], cur_w = w.load(std::memory_order_relaxed);
                while (t < cur_w && !w.compare_exchange_weak(cur_w, t, std::memory_order_relaxed)) {
                }
            }
        });
        parallel_for(n_txns, n_threads, [&](uint64_t a, uint64_t b) {
            for (uint64_t t = a; t < b; ++t) {
                uint8_t ok = 1;
                for (uint32_t i = txn_off[t]; i < txn_off[t + 1]; ++i) {
                    OpRef o = decode(meta[i]);
                    if (W[base[o.tab] + rows[i]].load(std::memory_order_relaxed) < t) ok = 0;
                }
                commit[t] = ok;
            }
        });
        // committed reads see the snapshot: all reads before any write
        parallel_for(n_ops, n_threads, [&](uint64_t a, uint64_t b) {
            for (uint64_t i = a; i < b; ++i) {
                OpRef o = decode(meta[i]);
                if (o.kind == ORC_KIND_READ && commit[op_txn[i]])
                    exec_op(tables, o, rows[i], args[i], results + i * res_stride, res_stride);
            }
        });
        parallel_for(n_ops, n_threads, [&](uint64_t a, uint64_t b) {
            for (uint64_t i = a; i < b; ++i) {
                OpRef o = decode(meta[i]);
                if (o.kind != ORC_KIND_READ && commit[op_txn[i]])
                    exec_op(tables, o, rows[i], args[i], results + i * res_stride, res_stride);
            }
        });
    } else {
        return -1.0;
    }
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}
