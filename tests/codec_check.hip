// Host-side check of the handle codec in svdag_device.h (built and run by
// test_codec_cpu.py; no GPU needed).  For every item size and material width
// it checks that the packed slot-code widths equal bit_width(line_slots - 1),
// that node and leaf handles round-trip pointer and size at the first, a
// middle and the last line a table may hold, that no handle equals the
// sentinel, and that the largest pointer fits 32 bits.  Prints "ok" or the
// first failure.
#include "svdag_device.h"

#include <cstdio>

using namespace svdag;

static uint32_t bw(uint32_t v)
{
    uint32_t b = 0;
    while (v) {
        ++b;
        v >>= 1;
    }
    return b;
}

int main()
{
    for (uint32_t s = 2; s <= kMaxItemWords; ++s) {
        const uint32_t want = bw(line_slots(s) - 1u);
        if (slot_code_bits(s) != want) {
            std::printf("slot_code_bits(%u) = %u, want %u\n", s, slot_code_bits(s), want);
            return 1;
        }
    }
    const uint32_t Ms[3] = { 0, 4, 8 };
    for (uint32_t M : Ms) {
        DagDev d {};
        d.M = M;
        d.minLeafWords = 2 + (M ? 1 : 0);
        d.maxLeafWords = 2 + (64 * M + 31) / 32;
        d.leafSizeBits = bw(d.maxLeafWords - d.minLeafWords);
        const uint32_t maxWords = d.maxLeafWords > 9 ? d.maxLeafWords : 9;
        for (uint32_t s = 2; s <= maxWords; ++s) {
            const uint64_t limit = table_line_limit(s, d.minLeafWords, d.maxLeafWords, d.leafSizeBits);
            if (limit == 0 || ((limit - 1) << kSlotBits | (kMaxSlots - 1)) > 0xFFFFFFFFull) {
                std::printf("M=%u s=%u: line limit %llu out of the pointer range\n", M, s, (unsigned long long)limit);
                return 1;
            }
            const bool node = s <= 9, leaf = s >= d.minLeafWords && s <= d.maxLeafWords;
            const uint64_t lines[4] = { 0, 1, limit / 2, limit - 1 };
            for (uint64_t line : lines) {
                for (uint32_t slot = 0; slot < line_slots(s); ++slot) {
                    const uint32_t ptr = (uint32_t)(line << kSlotBits) | slot;
                    if (index_to_ptr(ptr_to_index(ptr, s), s) != ptr) {
                        std::printf("M=%u s=%u: index round trip of %#x\n", M, s, ptr);
                        return 1;
                    }
                    if (node) {
                        const uint32_t h = node_handle(ptr, s);
                        if (h == kSentinel || node_handle_size(h) != s || node_handle_ptr(h) != ptr) {
                            std::printf("M=%u s=%u: node handle %#x of %#x\n", M, s, h, ptr);
                            return 1;
                        }
                    }
                    if (leaf) {
                        const uint32_t h = leaf_handle(d, ptr, s);
                        if (h == kSentinel || leaf_handle_size(d, h) != s || leaf_handle_ptr(d, h) != ptr) {
                            std::printf("M=%u s=%u: leaf handle %#x of %#x\n", M, s, h, ptr);
                            return 1;
                        }
                    }
                }
            }
            std::printf("M=%u s=%u slots=%u code bits=%u line limit=%llu\n", M, s, line_slots(s), slot_code_bits(s),
                        (unsigned long long)limit);
        }
    }
    std::printf("ok\n");
    return 0;
}
