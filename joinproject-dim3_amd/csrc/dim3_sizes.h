// Size rules of the device key tables (host-only, plain C++: the CPU tests
// compile this header on its own).
//
// A relation holds at most kMaxTuples tuples, so that every slot index and
// slot count of a key table fits 32 bits: an open-addressing table for n keys
// has key_table_cap(n) = the power of two >= 2n slots (at most 2^31) plus the
// slot standing for the key -1; a dense value range is mapped onto a power of
// two of at most 2^31 slots, else the attribute is hashed.
#pragma once

#include <stdint.h>

namespace dim3s {

constexpr uint64_t kMaxTuples = (1ull << 30) - 1;
constexpr uint64_t kMaxSlots = 1ull << 31;

// smallest power of two >= v (at least 64)
inline uint64_t pow2_at_least(uint64_t v) {
    uint64_t c = 64;
    while (c < v) c <<= 1;
    return c;
}

// slots of the open-addressing table of n keys (the key -1 slot not counted)
inline uint64_t key_table_cap(uint64_t n) { return pow2_at_least(2 * n); }

// a value range of `span` values over n tuples is mapped directly (no table)
// when it is at most 2n + 64 values wide and its power of two fits 2^31
inline bool range_fits(uint64_t span, uint64_t n) {
    return span <= 2 * n + 64 && pow2_at_least(span) <= kMaxSlots;
}

// the sampled keys of EstimateJoinCardinality with `bits` sample bits: at most
// the S size and at most the 2^(32 - bits) int32 values whose low bits are zero
inline uint64_t estimate_table_cap(uint64_t nS, uint32_t bits) {
    const uint64_t vals = bits >= 32 ? 1ull : (1ull << (32 - bits));
    return key_table_cap(nS < vals ? nS : vals);
}

}  // namespace dim3s
