// Offsets of the exclusive scans over per-particle counts (cell occupancy, candidates, kept pairs). Counts are summed
// in 64 bits and stored as int32 offsets that saturate at INT32_MAX instead of wrapping, so offsets never turn
// negative and stay non-decreasing. Every capacity they are checked against is below 2^31, so a source whose offset
// saturated writes nothing and is flagged as a capacity overflow.
//
// The functions are constexpr, which makes them callable from device code and from plain host C++ alike (the CPU unit
// test of this logic compiles this header with g++).
#pragma once

#include <stdint.h>

namespace csg {

constexpr int32_t kOffsetMax = 0x7fffffff;

// a running 64-bit offset as stored in an int32 offset array
constexpr int32_t saturateOffset(long long v) { return v > kOffsetMax ? kOffsetMax : (int32_t)v; }

// whether candidate `pos` of a source whose list starts at `base` may be written into a buffer of `cap` entries
constexpr bool candidateSlotValid(long long base, long long pos, long long cap) {
    return base >= 0 && base < kOffsetMax && base + pos < cap;
}

// whether a source with `count` candidates starting at `base` loses some of them to a buffer of `cap` entries
constexpr bool candidatesOverflow(long long base, long long count, long long cap) {
    return count > 0 && (base < 0 || base >= kOffsetMax || base + count > cap);
}

}  // namespace csg
