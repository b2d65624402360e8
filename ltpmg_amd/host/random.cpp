// random.cpp -- seeded random source of the host API (see random.h).
#include "random.h"

#include "ltpmg_workload.h"

namespace {
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
}  // namespace

Random::Random() { init_seed(0); }

Random::~Random() {}

void Random::init_seed(uint64_t seed) { seed_ = mix64(seed + kGolden); }

void Random::set_seed(uint64_t seed) { seed_ = seed; }

uint64_t Random::get_seed() { return seed_; }

uint64_t Random::next() {
    seed_ += kGolden;
    return mix64(seed_);
}

uint64_t Random::next(uint32_t bits) { return bits == 0 ? 0 : next() >> (64 - bits); }

double Random::next_double() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }

uint64_t Random::uniform_dist(uint64_t a, uint64_t b) {
    if (b <= a) return a;
    uint64_t span = b - a + 1;
    if (span == 0) return next();  // the whole 64-bit range
    return a + (uint64_t)(((unsigned __int128)next() * span) >> 64);
}

std::string Random::rand_str(std::size_t length, const std::string &str) {
    std::string r;
    r.reserve(length);
    for (std::size_t i = 0; i < length; ++i) r += str[uniform_dist(0, str.size() - 1)];
    return r;
}

std::string Random::a_string(std::size_t min_len, std::size_t max_len) {
    return rand_str(uniform_dist(min_len, max_len), alpha());
}

// TPC-C NURand(A, x, y)
uint64_t Random::non_uniform_distribution(uint64_t A, uint64_t x, uint64_t y) {
    return (uniform_dist(0, A) | uniform_dist(x, y)) % (y - x + 1) + x;
}

std::string Random::n_string(std::size_t min_len, std::size_t max_len) {
    return rand_str(uniform_dist(min_len, max_len), numeric());
}

std::string Random::rand_zip() { return rand_str(4, numeric()) + "11111"; }

std::string Random::rand_last_name(int n) {
    const auto &s = customer_last_names();
    return s[(n / 100) % 10] + s[(n / 10) % 10] + s[n % 10];
}

std::string Random::rand_C_LAST() { return rand_last_name((int)non_uniform_distribution(255, 0, 999)); }

uint32_t Random::rand_test(uint64_t a, uint64_t b) { return (uint32_t)uniform_dist(a, b); }

void Random::init_rand_zipf(uint64_t n, double s) {
    if (n < 2) n = 2;
    if (s >= 1.0) s = 0.99;  // Gray et al.'s method needs theta < 1
    if (s <= 0.0) s = 0.01;
    zipf_n_ = n;
    zipf_theta_ = s;
    zipf_zetan_ = ltpmg_zeta(n, s);
    zipf_alpha_ = 1.0 / (1.0 - s);
    double zeta2 = 1.0 + std::pow(0.5, s);
    zipf_eta_ = (1.0 - std::pow(2.0 / (double)n, 1.0 - s)) / (1.0 - zeta2 / zipf_zetan_);
}

uint32_t Random::rand_zipf(uint64_t a, uint64_t b) {
    uint64_t n = b - a + 1;
    if (n != zipf_n_) init_rand_zipf(n, zipf_theta_ > 0 ? zipf_theta_ : 0.99);
    double u = next_double(), uz = u * zipf_zetan_;
    uint64_t k;
    if (uz < 1.0) k = 0;
    else if (uz < 1.0 + std::pow(0.5, zipf_theta_)) k = 1;
    else k = (uint64_t)((double)n * std::pow(zipf_eta_ * u - zipf_eta_ + 1.0, zipf_alpha_));
    if (k >= n) k = n - 1;
    return (uint32_t)(a + k);
}
