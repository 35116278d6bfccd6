// Shared host-side helpers: error reporting across the C ABI, deterministic
// random streams, and a small parallel-for over chunks.
#pragma once

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/medusa_c.h"

namespace medusa {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

[[noreturn]] void fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void set_last_error(const char* msg);
void clear_last_error();

#define MDS_CHECK(cond, code, ...)                  \
    do {                                            \
        if (!(cond)) ::medusa::fail(code, __VA_ARGS__); \
    } while (0)

// Runs f() and converts exceptions into a status code + last-error text.
template <class F>
int guard(F&& f) {
    try {
        clear_last_error();
        f();
        return MEDUSA_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return MEDUSA_ERR_NOMEM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return MEDUSA_ERR_INTERNAL;
    }
}

// ---- deterministic random streams --------------------------------------
// splitmix64 seeds a xoshiro256** state; stream s of seed k is independent of
// how many threads generate it, so every generator is reproducible.
inline uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    Rng(uint64_t seed, uint64_t stream) {
        uint64_t x = seed * 0x9E3779B97F4A7C15ull + stream * 0xD1B54A32D192ED03ull + 0x632BE59BD9B4E019ull;
        for (auto& w : s) w = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    inline uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform double in [0,1) with 53 random bits
    inline double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    // uniform integer in [0, n) by multiply-high (n < 2^32)
    inline uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
    // uniform integer in [0, n) for 64-bit n
    inline uint64_t below64(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
};

// Parallel loop over [0, count) items in the given number of threads; each
// worker takes a contiguous block.  f(begin, end) must be thread-safe.
template <class F>
void parallel_blocks(uint64_t count, F&& f, unsigned max_threads = 0) {
    unsigned hw = std::thread::hardware_concurrency();
    if (hw == 0) hw = 1;
    if (max_threads && hw > max_threads) hw = max_threads;
    if (hw > 32) hw = 32;
    if (count < 2 * (uint64_t)hw || hw == 1) {
        f((uint64_t)0, count);
        return;
    }
    std::vector<std::thread> ts;
    const uint64_t per = (count + hw - 1) / hw;
    for (unsigned t = 0; t < hw; ++t) {
        const uint64_t b = per * t, e = std::min<uint64_t>(count, b + per);
        if (b >= e) break;
        ts.emplace_back([&f, b, e] { f(b, e); });
    }
    for (auto& t : ts) t.join();
}

}  // namespace medusa
