// diso_scan_word.h - aggregate words of the single-pass scan's control block.
//
// The control block holds a header (an epoch counter) and two buffers of 64-bit words, one word
// per count block. Call number `epoch` uses buffer (epoch & 1): every block publishes its
// aggregate there, tagged with tag(epoch) = 1 + epoch % 3, and reads its predecessors' words from
// it; it clears its share of the other buffer for the next call; the call's last block advances
// the epoch. A word counts as published for a call only when its tag is that call's tag, so a
// word written late by a block of an earlier call (one whose last block gave up waiting for it)
// is never taken for a word of a later call that reuses the buffer: the buffer is reused two
// calls later, whose tag differs from the writer's. The same word then reads as "not published"
// and is either overwritten by the current call's block or counted around (the reader recounts
// that block's range).
//
// Word layout: [63:62] tag (0 = empty), [61:59] sample flags, [58:29] face / patch count,
// [28:0] vertex / quad count.
//
// Plain C++ (host and device), so the protocol is testable on the CPU.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define DISO_HD __host__ __device__
#else
#define DISO_HD
#endif

namespace diso_scan {

constexpr int kCountBits = 29;  // vertex / quad count
constexpr int kFaceBits = 30;   // face / patch count

DISO_HD constexpr uint32_t buffer_of(uint32_t epoch) { return epoch & 1u; }
DISO_HD constexpr uint32_t tag_of(uint32_t epoch) { return 1u + epoch % 3u; }

DISO_HD constexpr uint64_t pack(int64_t v, int64_t f, uint32_t flags, uint32_t tag) {
    return ((uint64_t)tag << 62) | ((uint64_t)(flags & 7u) << 59) | ((uint64_t)f << kCountBits) | (uint64_t)v;
}
DISO_HD constexpr bool published(uint64_t w, uint32_t tag) { return (uint32_t)(w >> 62) == tag; }
DISO_HD constexpr int64_t count_v(uint64_t w) { return (int64_t)(w & ((1ull << kCountBits) - 1)); }
DISO_HD constexpr int64_t count_f(uint64_t w) { return (int64_t)((w >> kCountBits) & ((1ull << kFaceBits) - 1)); }
DISO_HD constexpr uint32_t flags_of(uint64_t w) { return (uint32_t)((w >> 59) & 7u); }

}  // namespace diso_scan
