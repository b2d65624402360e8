// server.cpp -- open-loop serving over the engine (include/ltpmg_serve.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <thread>
#include <vector>

#include "ltpmg.h"
#include "ltpmg_serve.h"

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

using Clock = std::chrono::steady_clock;

// Consecutive arrival ids [first, first + n): arrivals join the queue as
// spans, so the host's work per transaction is a few array accesses.
struct Span {
    uint32_t first;
    uint32_t n;
};

struct Flight {
    uint32_t slot;
    uint64_t first;      // entry index of its first transaction (cut order)
    uint32_t n;
    uint32_t n_ops;
    uint64_t bytes;      // device bytes the batch holds (admission)
    size_t span0, n_spans;  // its spans in TID order, in `cut_spans`
};

// A pinned copy of the pool's op arrays, made before the clock starts: cut
// batches go up from it by DMA (ltpmg_submit_runs) without a host copy.
struct PinnedPool {
    ltpmg_batch b{};
    void *rows = nullptr, *meta = nullptr, *args = nullptr;
    ~PinnedPool() {
        ltpmg_pinned_free(rows);
        ltpmg_pinned_free(meta);
        ltpmg_pinned_free(args);
    }
    int init(const ltpmg_batch &pool) {
        const size_t n = pool.n_ops;
        if (ltpmg_pinned_alloc(n * 4, &rows) != LTPMG_OK || ltpmg_pinned_alloc(n * 2, &meta) != LTPMG_OK ||
            ltpmg_pinned_alloc(n * 8, &args) != LTPMG_OK)
            return LTPMG_ENOMEM;
        std::memcpy(rows, pool.rows, n * 4);
        std::memcpy(meta, pool.meta, n * 2);
        std::memcpy(args, pool.args, n * 8);
        b = ltpmg_batch{pool.n_txns, pool.n_ops, pool.txn_off, static_cast<const uint32_t *>(rows),
                        static_cast<const uint16_t *>(meta), static_cast<const uint64_t *>(args)};
        return LTPMG_OK;
    }
};

double pct(std::vector<double> &v, double q) {
    if (v.empty()) return 0.0;
    size_t k = (size_t)std::min<double>((double)v.size() - 1, std::floor(q * (double)(v.size() - 1) + 0.5));
    std::nth_element(v.begin(), v.begin() + (long)k, v.end());
    return v[k];
}

}  // namespace

extern "C" int ltpmg_serve_run(ltpmg_engine *e, const ltpmg_serve_cfg *cfg, const ltpmg_batch *pool,
                               ltpmg_serve_stats *out, ltpmg_serve_record *rec) {
    if (!e || !cfg || !pool || !out || pool->n_txns == 0) return LTPMG_EINVAL;
    if (cfg->max_batch == 0 || cfg->n_inflight == 0 || cfg->n_inflight > LTPMG_MAX_SLOTS ||
        (cfg->mode != LTPMG_MODE_A && cfg->mode != LTPMG_MODE_O) || cfg->arrival_rate < 0)
        return LTPMG_EINVAL;
    if (rec && rec->results && cfg->n_arrivals > pool->n_txns) return LTPMG_EINVAL;
    int rc = ltpmg_check_batch(e, pool);
    if (rc != LTPMG_OK) return rc;
    std::memset(out, 0, sizeof *out);
    const uint32_t stride = ltpmg_result_stride(e);
    const uint32_t n = cfg->n_arrivals;

    // scheduled arrival times (s): exponential gaps from a SplitMix64 stream
    std::vector<double> arrive(n, 0.0);
    if (cfg->arrival_rate > 0) {
        double t = 0.0;
        uint64_t s = mix64(cfg->seed + kGolden);
        for (uint32_t i = 0; i < n; ++i) {
            s += kGolden;
            double u = (double)(mix64(s) >> 11) * (1.0 / 9007199254740992.0);
            t += -std::log1p(-u) / cfg->arrival_rate;
            arrive[i] = t;
        }
    }

    // size every slot for the largest batch before the clock starts, so no
    // batch pays for device or pinned allocations while requests wait
    {
        uint32_t max_len = 1;
        for (uint32_t t = 0; t < pool->n_txns; ++t)
            max_len = std::max(max_len, pool->txn_off[t + 1] - pool->txn_off[t]);
        const uint64_t cap_ops = std::min<uint64_t>((uint64_t)cfg->max_batch * max_len, (1u << 29) - 1);
        for (uint32_t s = 0; s < cfg->n_inflight; ++s) {
            ltpmg_staging st;
            rc = ltpmg_stage_batch(e, s, cfg->max_batch, (uint32_t)cap_ops, &st);
            if (rc != LTPMG_OK) return rc;
        }
    }

    PinnedPool pinned;
    rc = pinned.init(*pool);
    if (rc != LTPMG_OK) return rc;
    // device bytes of a batch: its ops and result slots, the two key / value
    // buffers of the ordered sort, its offsets and commit flags
    const uint64_t op_bytes = 14 + (uint64_t)stride + 16, txn_bytes = 5;
    auto batch_bytes = [&](uint64_t txns, uint64_t ops) { return ops * op_bytes + (txns + 1) * txn_bytes; };
    uint64_t inflight_bytes = 0, peak_bytes = 0, holds = 0;

    std::deque<Span> queue;  // waiting transactions, in TID order for the next cut
    uint64_t queued = 0;
    std::vector<Span> cut_spans;  // spans of the batches in flight, in cut order
    std::vector<Span> retry;
    std::vector<ltpmg_op_run> runs;
    std::vector<uint8_t> commit_buf, res_buf;
    std::deque<Flight> inflight;
    std::vector<uint32_t> free_slots;
    for (uint32_t s = cfg->n_inflight; s-- > 0;) free_slots.push_back(s);
    std::vector<double> latency(n);
    uint32_t next = 0, done = 0;
    uint64_t cut_entries = 0;
    double pack_s = 0.0, device_ms = 0.0;
    uint64_t batches = 0, retries = 0, batch_txn_sum = 0;
    if (rec) {
        rec->n_batches = rec->n_entries = rec->overflow = 0;
        if (rec->batch_off && rec->cap_batches) rec->batch_off[0] = 0;
    }
    const uint32_t P = pool->n_txns;
    // ops of pool txns [pt, pt + c), which do not wrap
    auto pool_ops = [&](uint32_t pt, uint32_t c) { return pool->txn_off[pt + c] - pool->txn_off[pt]; };

    const auto t0 = Clock::now();
    auto now_s = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
    double last_done = 0.0;

    while (done < n) {
        bool progressed = false;
        const double now = now_s();
        if (next < n && arrive[next] <= now) {
            const uint32_t upto = (uint32_t)(std::upper_bound(arrive.begin() + next, arrive.end(), now) - arrive.begin());
            if (!queue.empty() && queue.back().first + queue.back().n == next) queue.back().n += upto - next;
            else queue.push_back(Span{next, upto - next});
            queued += upto - next;
            next = upto;
        }

        // completions, in cut order
        while (!inflight.empty()) {
            const Flight f = inflight.front();
            int q = ltpmg_query(e, f.slot);
            if (q == 1) break;
            if (q != LTPMG_OK) return q;
            const double t_done = now_s();
            last_done = t_done;
            commit_buf.resize(f.n);
            const bool want_res = rec && rec->results;
            if (want_res) res_buf.resize((size_t)f.n_ops * stride);
            const bool all_commit = cfg->mode == LTPMG_MODE_O;
            if (all_commit && !want_res) {
                // the ordered mode commits every transaction: no flags to copy back
                std::fill(commit_buf.begin(), commit_buf.end(), (uint8_t)1);
            } else {
                rc = ltpmg_fetch_results(e, f.slot, commit_buf.data(), want_res ? res_buf.data() : nullptr,
                                         want_res ? res_buf.size() : 0);
                if (rc != LTPMG_OK) return rc;
            }
            ltpmg_batch_stats st;
            if (ltpmg_stats(e, f.slot, &st) == LTPMG_OK) device_ms += st.exec_ms;
            retry.clear();
            if (all_commit && !want_res && !(rec && rec->batch_commit)) {
                for (size_t si = f.span0; si < f.span0 + f.n_spans; ++si) {
                    const Span sp = cut_spans[si];
                    for (uint32_t i = 0; i < sp.n; ++i) latency[done + i] = t_done - arrive[sp.first + i];
                    done += sp.n;
                }
            } else {
                uint32_t k = 0;
                size_t op = 0;
                for (size_t si = f.span0; si < f.span0 + f.n_spans; ++si) {
                    const Span sp = cut_spans[si];
                    uint32_t pt = sp.first % P;
                    for (uint32_t i = 0; i < sp.n; ++i, ++k) {
                        const uint32_t id = sp.first + i, len = pool_ops(pt, 1);
                        if (rec && rec->batch_commit && f.first + k < rec->cap_entries)
                            rec->batch_commit[f.first + k] = commit_buf[k];
                        if (commit_buf[k]) {
                            latency[done++] = t_done - arrive[id];
                            if (want_res)
                                std::memcpy(rec->results + (size_t)pool->txn_off[pt] * stride,
                                            res_buf.data() + op * stride, (size_t)len * stride);
                        } else if (!retry.empty() && retry.back().first + retry.back().n == id) {
                            ++retry.back().n;
                        } else {
                            retry.push_back(Span{id, 1});
                        }
                        op += len;
                        if (++pt == P) pt = 0;
                    }
                }
            }
            for (const Span &r : retry) retries += r.n;
            for (size_t q2 = retry.size(); q2-- > 0;) {
                queue.push_front(retry[q2]);
                queued += retry[q2].n;
            }
            inflight_bytes -= f.bytes;
            free_slots.push_back(f.slot);
            inflight.pop_front();
            // the spans of batches no longer in flight are dropped now and
            // then (the flights' span indices move down with them)
            const size_t live0 = inflight.empty() ? cut_spans.size() : inflight.front().span0;
            if (live0 == cut_spans.size()) {
                cut_spans.clear();
            } else if (live0 > 4096 && live0 * 2 > cut_spans.size()) {
                cut_spans.erase(cut_spans.begin(), cut_spans.begin() + (long)live0);
                for (Flight &g : inflight) g.span0 -= live0;
            }
            progressed = true;
        }

        // cut a batch into a free slot
        if (!free_slots.empty() && queued) {
            bool cut;
            if (cfg->min_batch) {
                cut = queued >= cfg->min_batch || (next == n && inflight.empty());
            } else {
                cut = queued >= cfg->max_batch ||
                      (now - arrive[queue.front().first]) * 1e6 >= cfg->max_wait_us || next == n;
            }
            if (cut) {
                const auto p0 = Clock::now();
                // admission: the batch's device bytes must fit the in-flight
                // budget beside the batches already on the device; a batch
                // that does not fit is cut smaller, or waits for one to finish.
                // The batch is taken from the queue in pieces of consecutive
                // pool txns (a span ends, or the pool wraps).
                const uint32_t want = (uint32_t)std::min<uint64_t>(queued, cfg->max_batch);
                uint32_t k = 0;
                uint64_t n_ops = 0;
                const size_t span0 = cut_spans.size();
                bool held = false;
                for (size_t qi = 0; qi < queue.size() && k < want && !held;) {
                    Span &sp = queue[qi];
                    const uint32_t pt = sp.first % P;
                    uint32_t c = std::min({sp.n, want - k, P - pt});
                    uint64_t ops = pool_ops(pt, c);
                    if (cfg->hbm_inflight_bytes &&
                        inflight_bytes + batch_bytes(k + c, n_ops + ops) > cfg->hbm_inflight_bytes) {
                        // the most txns of this piece that fit (ops grow with c)
                        uint32_t lo = 0, hi = c;
                        while (lo < hi) {
                            const uint32_t mid = (lo + hi + 1) / 2;
                            if (inflight_bytes + batch_bytes(k + mid, n_ops + pool_ops(pt, mid)) >
                                cfg->hbm_inflight_bytes)
                                hi = mid - 1;
                            else
                                lo = mid;
                        }
                        c = lo;
                        ops = pool_ops(pt, c);
                        held = true;
                    }
                    if (c == 0) break;
                    if (cut_spans.size() > span0 && cut_spans.back().first + cut_spans.back().n == sp.first)
                        cut_spans.back().n += c;
                    else
                        cut_spans.push_back(Span{sp.first, c});
                    k += c;
                    n_ops += ops;
                    sp.first += c;
                    sp.n -= c;
                    if (sp.n == 0) ++qi;
                }
                while (!queue.empty() && queue.front().n == 0) queue.pop_front();
                if (held) ++holds;
                if (k == 0) {
                    if (inflight.empty()) return LTPMG_EBUDGET;  // one transaction exceeds the budget
                    std::this_thread::yield();
                    continue;
                }
                queued -= k;
                const uint32_t slot = free_slots.back();
                ltpmg_staging st;
                rc = ltpmg_stage_batch(e, slot, k, (uint32_t)n_ops, &st);
                if (rc != LTPMG_OK) return rc;
                // offsets, and the ops as runs of the pinned pool: each piece
                // is a run (consecutive pieces that continue in the pool merge)
                runs.clear();
                st.txn_off[0] = 0;
                uint32_t q = 0;
                for (size_t si = span0; si < cut_spans.size(); ++si) {
                    const Span sp = cut_spans[si];
                    for (uint32_t i = 0; i < sp.n;) {
                        const uint32_t pt = (sp.first + i) % P, c = std::min(sp.n - i, P - pt);
                        const uint32_t base = st.txn_off[q], src0 = pool->txn_off[pt];
                        for (uint32_t j = 1; j <= c; ++j) st.txn_off[q + j] = base + (pool->txn_off[pt + j] - src0);
                        const uint32_t len = st.txn_off[q + c] - base;
                        if (!runs.empty() && runs.back().src_op + runs.back().n_ops == src0)
                            runs.back().n_ops += len;
                        else
                            runs.push_back(ltpmg_op_run{src0, base, len});
                        q += c;
                        i += c;
                    }
                }
                rc = ltpmg_submit_runs(e, slot, &pinned.b, runs.data(), (uint32_t)runs.size());  // the pool was checked
                if (rc == LTPMG_OK) rc = ltpmg_execute(e, slot, cfg->mode);
                if (rc != LTPMG_OK) return rc;
                free_slots.pop_back();
                const uint64_t bytes = batch_bytes(k, n_ops);
                inflight_bytes += bytes;
                peak_bytes = std::max(peak_bytes, inflight_bytes);
                inflight.push_back(Flight{slot, cut_entries, k, (uint32_t)n_ops, bytes, span0, cut_spans.size() - span0});
                ++batches;
                batch_txn_sum += k;
                if (rec && rec->batch_off && rec->batch_txns && !rec->overflow) {
                    if (rec->n_batches < rec->cap_batches && cut_entries + k <= rec->cap_entries) {
                        uint64_t at = cut_entries;
                        for (size_t si = span0; si < cut_spans.size(); ++si)
                            for (uint32_t i = 0; i < cut_spans[si].n; ++i) rec->batch_txns[at++] = cut_spans[si].first + i;
                        rec->batch_off[++rec->n_batches] = (uint32_t)(cut_entries + k);
                        rec->n_entries = (uint32_t)(cut_entries + k);
                    } else {
                        rec->overflow = 1;
                    }
                }
                cut_entries += k;
                pack_s += std::chrono::duration<double>(Clock::now() - p0).count();
                progressed = true;
            }
        }
        if (!progressed) std::this_thread::yield();
    }
    latency.resize(done);

    out->offered = n;
    out->committed = done;
    out->aborted_attempts = retries;
    out->batches = batches;
    out->wall_s = last_done - (n ? arrive[0] : 0.0);
    out->committed_per_s = out->wall_s > 0 ? done / out->wall_s : 0.0;
    out->p50_ms = 1e3 * pct(latency, 0.50);
    out->p90_ms = 1e3 * pct(latency, 0.90);
    out->p99_ms = 1e3 * pct(latency, 0.99);
    out->p999_ms = 1e3 * pct(latency, 0.999);
    out->max_ms = latency.empty() ? 0.0 : 1e3 * *std::max_element(latency.begin(), latency.end());
    out->mean_batch_txns = batches ? (double)batch_txn_sum / batches : 0.0;
    out->host_pack_s = pack_s;
    out->device_ms = device_ms;
    out->peak_inflight_bytes = peak_bytes;
    out->budget_holds = holds;
    return LTPMG_OK;
}
