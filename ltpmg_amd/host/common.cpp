// common.cpp -- flags, logging and timers of the host API
// (the reference's src/common.cpp; flag names and defaults kept, except
// deviceIDs, which defaults to the one GPU of this build).
#include "common.h"
#include "run_options.h"

#include <ctime>
#include <map>

DEFINE_string(benchmark, "YCSB_E", "YCSB_A/B/C/E, KV, KV_C1, TPCC_PART, OE (TPCC_ALL/TEST: sizes only).");
DEFINE_uint32(table_size, 1, "YCSB: table_size * 1,000,000 rows; TPC-C: warehouses.");
DEFINE_uint32(batch_size, 8192, "transactions per batch.");
DEFINE_uint32(epoch_tp, 10, "batches executed (the same batch each epoch).");
DEFINE_uint32(epoch_sync, 1, "kept for compatibility (unused, as in the reference).");
DEFINE_string(deviceIDs, "0", "GPU id; this build runs one GPU.");
DEFINE_uint32(neworder_percent, 50, "percentage of neworder transactions (TPC-C sizing).");
DEFINE_string(log_path, "~/ltpmg/log", "path of log (kept for compatibility).");
DEFINE_double(zipf_config, 1.0, "zipf theta; values >= 1 are clamped to 0.99.");
DEFINE_string(data_distribution, "unif", "unif or zipf.");
DEFINE_string(cc_mode, "O", "O: ordered, every txn commits (per-row TID order); A: lowest TID wins.");
DEFINE_uint64(seed, 0x4C54504D47000001ull, "seed of the table fill and the workload.");
DEFINE_uint32(kv_rows, 100000, "KV / KV_C1: rows.");
DEFINE_uint32(kv_fields, 10, "KV / KV_C1: fields per row.");
DEFINE_uint32(kv_field_bytes, 100, "KV: bytes per field (KV_C1 uses 4).");
DEFINE_uint32(kv_ops, 16, "KV / KV_C1: logical ops per txn.");
DEFINE_uint32(read_percent, 50, "KV / KV_C1: percent of ops that READ.");
DEFINE_string(dump_path, "", "write epoch 0's batch, results and image digest under this prefix.");
DEFINE_uint32(append_cap, 1u << 20, "order entry: rows the append-only tables can take, in transactions.");
DEFINE_uint32(oe_preload, 3000, "order entry: orders per district at table creation (at most 3000).");
DEFINE_string(schema_dir, "", "directory of <benchmark>.json table schemas when Param's path is absent "
                             "(default: the schemas/ directory beside the library).");
DEFINE_uint32(minloglevel, 0, "drop log lines below this severity (0 INFO .. 3 FATAL).");

namespace ltpmg_log {
int min_level = 0;
}

namespace ltpmg_flags {

namespace {
struct Entry {
    Kind kind;
    void *ptr;
    const char *help;
};
std::map<std::string, Entry> &registry() {
    static std::map<std::string, Entry> r;
    return r;
}

bool assign(const Entry &e, const std::string &v) {
    try {
        size_t used = 0;
        switch (e.kind) {
            case Kind::Str: *static_cast<std::string *>(e.ptr) = v; return true;
            case Kind::U32: {
                unsigned long long x = std::stoull(v, &used, 0);
                if (used != v.size() || x > 0xFFFFFFFFull) return false;
                *static_cast<uint32_t *>(e.ptr) = (uint32_t)x;
                return true;
            }
            case Kind::U64: {
                unsigned long long x = std::stoull(v, &used, 0);
                if (used != v.size()) return false;
                *static_cast<uint64_t *>(e.ptr) = x;
                return true;
            }
            case Kind::Dbl: {
                double x = std::stod(v, &used);
                if (used != v.size()) return false;
                *static_cast<double *>(e.ptr) = x;
                return true;
            }
            case Kind::Bool:
                if (v == "1" || v == "true") *static_cast<bool *>(e.ptr) = true;
                else if (v == "0" || v == "false") *static_cast<bool *>(e.ptr) = false;
                else return false;
                return true;
        }
    } catch (...) {
    }
    return false;
}
}  // namespace

Reg::Reg(const char *name, Kind kind, void *ptr, const char *help) {
    registry()[name] = Entry{kind, ptr, help};
}

void print_help(std::ostream &os) {
    for (auto &kv : registry()) os << "  --" << kv.first << "  " << kv.second.help << "\n";
}

bool parse(int *argc, char ***argv) {
    std::vector<char *> keep{(*argv)[0]};
    for (int i = 1; i < *argc; ++i) {
        std::string a = (*argv)[i];
        if (a.rfind("--", 0) != 0 && a.rfind("-", 0) != 0) {
            keep.push_back((*argv)[i]);
            continue;
        }
        std::string body = a.substr(a.rfind("--", 0) == 0 ? 2 : 1), name = body, value;
        bool has_value = false;
        auto eq = body.find('=');
        if (eq != std::string::npos) {
            name = body.substr(0, eq);
            value = body.substr(eq + 1);
            has_value = true;
        }
        if (name == "help") {
            print_help(std::cout);
            std::exit(0);
        }
        auto it = registry().find(name);
        if (it == registry().end()) {
            std::cerr << "unknown flag --" << name << "\n";
            return false;
        }
        if (!has_value) {
            if (it->second.kind == Kind::Bool) {
                value = "true";
            } else if (i + 1 < *argc) {
                value = (*argv)[++i];
            } else {
                std::cerr << "flag --" << name << " needs a value\n";
                return false;
            }
        }
        if (!assign(it->second, value)) {
            std::cerr << "bad value '" << value << "' for --" << name << "\n";
            return false;
        }
    }
    for (size_t i = 0; i < keep.size(); ++i) (*argv)[i] = keep[i];
    *argc = (int)keep.size();
    return true;
}

}  // namespace ltpmg_flags

void initialize_dependency(int argc, char **argv) {
    if (!ltpmg_flags::parse(&argc, &argv)) {
        ltpmg_flags::print_help(std::cerr);
        std::exit(2);
    }
    ltpmg_log::min_level = (int)FLAGS_minloglevel;
    // the GPU library (libltpmg_gpu.so) reads this build's settings from the
    // environment: it sees only the reference's Param
    ltpmg_host::export_gpu_options();
}

void free_dependency() {}

long long current_time() {
    timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (long long)t.tv_sec * 1000000 + t.tv_nsec / 1000;
}

float duration(long long start_t, long long end_t) { return (float)(end_t - start_t) / 1e6f; }

void sleep(float time_t) {
    long long start_t = current_time();
    while (duration(start_t, current_time()) < time_t) {
    }
}
