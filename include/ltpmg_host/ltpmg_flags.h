// ltpmg_flags.h -- the small part of gflags and glog that the host API uses.
//
// The reference host code uses gflags (DEFINE_*/DECLARE_*, FLAGS_*,
// ParseCommandLineFlags) and glog (LOG(severity) streams), src/common.cpp:3-43.
// Neither library is part of this image, so these macros keep the same
// spelling over a tiny registry: --name=value and --name value are parsed,
// unknown flags are an error, LOG(FATAL) aborts after printing.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <sstream>
#include <string>

namespace ltpmg_flags {

enum class Kind { Str, U32, U64, Dbl, Bool };

struct Reg {
    Reg(const char *name, Kind kind, void *ptr, const char *help);
};

// Parses and removes recognised flags from argv.  Returns false (after
// printing why) on an unknown flag or a malformed value.
bool parse(int *argc, char ***argv);
void print_help(std::ostream &os);

}  // namespace ltpmg_flags

#define DECLARE_string(n) extern std::string FLAGS_##n
#define DECLARE_uint32(n) extern uint32_t FLAGS_##n
#define DECLARE_uint64(n) extern uint64_t FLAGS_##n
#define DECLARE_double(n) extern double FLAGS_##n
#define DECLARE_bool(n) extern bool FLAGS_##n

#define LTPMG_DEFINE_FLAG_(type, kind, n, v, help) \
    type FLAGS_##n = v;                           \
    static ::ltpmg_flags::Reg ltpmg_flag_reg_##n(#n, ::ltpmg_flags::Kind::kind, &FLAGS_##n, help)
#define DEFINE_string(n, v, help) LTPMG_DEFINE_FLAG_(std::string, Str, n, v, help)
#define DEFINE_uint32(n, v, help) LTPMG_DEFINE_FLAG_(uint32_t, U32, n, v, help)
#define DEFINE_uint64(n, v, help) LTPMG_DEFINE_FLAG_(uint64_t, U64, n, v, help)
#define DEFINE_double(n, v, help) LTPMG_DEFINE_FLAG_(double, Dbl, n, v, help)
#define DEFINE_bool(n, v, help) LTPMG_DEFINE_FLAG_(bool, Bool, n, v, help)

namespace ltpmg_log {

enum Severity { INFO = 0, WARNING = 1, ERROR = 2, FATAL = 3 };

// Messages below this severity are dropped (set from --minloglevel).
extern int min_level;

class Line {
public:
    Line(Severity s, const char *file, int line) : sev_(s) {
        static const char tag[] = "IWEF";
        if (sev_ >= min_level) os_ << tag[sev_] << ' ' << basename(file) << ':' << line << "] ";
    }
    ~Line() {
        if (sev_ >= min_level || sev_ == FATAL) {
            os_ << '\n';
            std::cerr << os_.str() << std::flush;
        }
        if (sev_ == FATAL) std::abort();
    }
    std::ostream &stream() { return os_; }

private:
    static const char *basename(const char *p) {
        const char *b = p;
        for (; *p; ++p)
            if (*p == '/') b = p + 1;
        return b;
    }
    Severity sev_;
    std::ostringstream os_;
};

}  // namespace ltpmg_log

#define LOG(sev) ::ltpmg_log::Line(::ltpmg_log::sev, __FILE__, __LINE__).stream()
