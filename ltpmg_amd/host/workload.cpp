// workload.cpp -- seeded, thread-count-independent KV batch generator.
// Semantics are documented in include/ltpmg_workload.h.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

#include "ltpmg.h"
#include "ltpmg_workload.h"

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed, uint64_t stream) : s(mix64(seed + kGolden * (stream + 1))) {}
    uint64_t next() {
        s += kGolden;
        return mix64(s);
    }
    // uniform in [0, n) by multiply-shift
    uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// zeta(n, theta), summed in fixed 2^20-term chunks so the result does not
// depend on how many threads add the chunks up.
double zeta_exact(uint64_t n, double theta) {
    static std::mutex mu;
    static std::map<std::pair<uint64_t, double>, double> cache;
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = cache.find({n, theta});
        if (it != cache.end()) return it->second;
    }
    const uint64_t chunk = 1u << 20, n_chunks = (n + chunk - 1) / chunk;
    std::vector<double> part(n_chunks, 0.0);
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::thread> th;
    for (unsigned w = 0; w < std::min<uint64_t>(hw, n_chunks); ++w)
        th.emplace_back([&, w] {
            for (uint64_t c = w; c < n_chunks; c += hw) {
                double s = 0.0;
                uint64_t a = c * chunk + 1, b = std::min(n, (c + 1) * chunk);
                for (uint64_t i = a; i <= b; ++i) s += std::pow((double)i, -theta);
                part[c] = s;
            }
        });
    for (auto &t : th) t.join();
    double z = 0.0;
    for (double p : part) z += p;
    std::lock_guard<std::mutex> g(mu);
    cache[{n, theta}] = z;
    return z;
}

struct Zipf {
    uint64_t n;
    double theta, zetan, alpha, eta, half_pow;
    Zipf(uint64_t n_, double th) : n(n_), theta(th) {
        zetan = zeta_exact(n, theta);
        double zeta2 = 1.0 + std::pow(0.5, theta);
        alpha = 1.0 / (1.0 - theta);
        eta = (1.0 - std::pow(2.0 / (double)n, 1.0 - theta)) / (1.0 - zeta2 / zetan);
        half_pow = std::pow(0.5, theta);
    }
    uint64_t draw(Rng &r) const {
        double u = r.unit(), uz = u * zetan;
        if (uz < 1.0) return 0;
        if (uz < 1.0 + half_pow) return 1;
        uint64_t k = (uint64_t)((double)n * std::pow(eta * u - eta + 1.0, alpha));
        return k < n ? k : n - 1;
    }
};

struct Chunk {
    std::vector<uint32_t> len, rows;
    std::vector<uint16_t> meta;
    std::vector<uint64_t> args;
};

}  // namespace

extern "C" double ltpmg_zeta(uint64_t n, double theta) { return zeta_exact(n, theta); }

extern "C" uint64_t ltpmg_gen_kv_max_ops(const ltpmg_kv_spec *sp) {
    if (!sp) return 0;
    uint64_t per = sp->read_all_cols ? std::max<uint32_t>(1, sp->n_cols) : 1;
    return (uint64_t)sp->n_txns * sp->ops_max * per;
}

extern "C" int ltpmg_gen_kv(const ltpmg_kv_spec *sp, uint32_t *txn_off, uint32_t *rows,
                            uint16_t *meta, uint64_t *args, uint64_t max_ops,
                            uint32_t *n_ops_out, uint32_t n_threads) {
    if (!sp || !txn_off || !n_ops_out) return LTPMG_EINVAL;
    if (sp->ops_min < 1 || sp->ops_max < sp->ops_min || sp->n_rows < sp->ops_max ||
        sp->n_cols < 1 || sp->n_cols > LTPMG_MAX_COLS || sp->read_pct > 100 ||
        sp->table_id >= LTPMG_MAX_TABLES || sp->zipf_theta < 0.0 || sp->zipf_theta >= 1.0 ||
        (sp->write_kind != LTPMG_OP_UPDATE && sp->write_kind != LTPMG_OP_RMW_ADD))
        return LTPMG_EINVAL;
    const uint64_t per_read = sp->read_all_cols ? sp->n_cols : 1;
    if ((uint64_t)sp->ops_max * per_read > LTPMG_MAX_TXN_OPS) return LTPMG_EINVAL;
    if (n_threads == 0) n_threads = std::max(1u, std::thread::hardware_concurrency());
    const bool zipf = sp->zipf_theta > 0.0;
    Zipf zp(zipf ? sp->n_rows : 2, zipf ? sp->zipf_theta : 0.5);

    const uint32_t n_chunks = std::min<uint32_t>(n_threads, std::max<uint32_t>(1, sp->n_txns));
    std::vector<Chunk> chunks(n_chunks);
    auto gen = [&](uint32_t c) {
        uint64_t t0 = (uint64_t)sp->n_txns * c / n_chunks, t1 = (uint64_t)sp->n_txns * (c + 1) / n_chunks;
        Chunk &ch = chunks[c];
        ch.len.reserve(t1 - t0);
        std::vector<uint32_t> keys;
        for (uint64_t t = t0; t < t1; ++t) {
            Rng r(sp->seed, t);
            uint32_t n_logical = sp->ops_min + (uint32_t)r.below(sp->ops_max - sp->ops_min + 1);
            keys.clear();
            uint32_t len = 0;
            for (uint32_t j = 0; j < n_logical; ++j) {
                uint32_t k;
                do {
                    k = (uint32_t)(zipf ? zp.draw(r) : r.below(sp->n_rows));
                } while (std::find(keys.begin(), keys.end(), k) != keys.end());
                keys.push_back(k);
                bool rd = r.below(100) < sp->read_pct;
                uint32_t col = (uint32_t)r.below(sp->n_cols);
                uint64_t a = r.next();
                if (rd && sp->read_all_cols) {
                    for (uint32_t cc = 0; cc < sp->n_cols; ++cc) {
                        ch.rows.push_back(k);
                        ch.meta.push_back(LTPMG_META(LTPMG_OP_READ, sp->table_id, cc));
                        ch.args.push_back(0);
                    }
                    len += sp->n_cols;
                } else {
                    uint32_t kind = rd ? LTPMG_OP_READ : sp->write_kind;
                    ch.rows.push_back(k);
                    ch.meta.push_back(LTPMG_META(kind, sp->table_id, col));
                    ch.args.push_back(kind == LTPMG_OP_RMW_ADD ? (uint64_t)(uint32_t)sp->rmw_delta
                                      : kind == LTPMG_OP_UPDATE ? a : 0);
                    len += 1;
                }
            }
            ch.len.push_back(len);
        }
    };
    std::vector<std::thread> th;
    for (uint32_t c = 0; c < n_chunks; ++c) th.emplace_back(gen, c);
    for (auto &t : th) t.join();

    uint64_t total = 0;
    for (auto &ch : chunks) total += ch.rows.size();
    if (total > max_ops || total >= (1ull << 30)) return LTPMG_EINVAL;
    uint64_t op = 0, t = 0;
    txn_off[0] = 0;
    for (auto &ch : chunks) {
        if (!ch.rows.empty()) {
            if (!rows || !meta || !args) return LTPMG_EINVAL;
            std::memcpy(rows + op, ch.rows.data(), ch.rows.size() * 4);
            std::memcpy(meta + op, ch.meta.data(), ch.meta.size() * 2);
            std::memcpy(args + op, ch.args.data(), ch.args.size() * 8);
        }
        for (uint32_t l : ch.len) {
            op += l;
            txn_off[++t] = (uint32_t)op;
        }
    }
    *n_ops_out = (uint32_t)total;
    return LTPMG_OK;
}
