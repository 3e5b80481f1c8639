// Host-side internals shared by the translation units of libsvdag.
#pragma once
#include "svdag_device.h"
#include "svdag.h"
#include <hip/hip_runtime.h>
#include <cstdio>
#include <string>
#include <vector>

namespace svdag {

// Evaluate a HIP call; on failure record the message and return SVDAG_E_DEVICE.
#define SVDAG_HIP(call)                                                          \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) {                                                  \
            ::svdag::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
            return SVDAG_E_DEVICE;                                               \
        }                                                                        \
    } while (0)

#define SVDAG_TRY(expr)          \
    do {                         \
        int r_ = (expr);         \
        if (r_ != SVDAG_OK)      \
            return r_;           \
    } while (0)

void set_error(const std::string& msg);
// SVDAG_E_NO_DEVICE (with the message set) when no HIP device is visible.
int require_device();

// One slab hash table: device storage plus the host copy of its descriptor.
struct HostTable {
    TableDev dev {};
    uint32_t itemWords = 0;
    uint32_t numBuckets = 0;
    uint32_t slabCap = 0;
};

// withRefs: the table holds DAG items, with per-slot reference counts and
// roles (dev.refs, dev.roles) kept beside the lines and grown with them.
int table_alloc(HostTable& t, uint32_t itemWords, uint32_t numBuckets, uint32_t slabCap, hipStream_t stream,
                bool withRefs = false);
void table_free(HostTable& t);
// Grow the slab arena of a table to newCap slabs, preserving slab ids.
int table_grow(HostTable& t, uint32_t newCap, hipStream_t stream);
int table_read_alloc(const HostTable& t, AllocState& out, hipStream_t stream);

struct TableSet {
    uint32_t minWords = 0, maxWords = 0;
    std::vector<HostTable> tables;   // index = words - minWords
    TableDev* devTables = nullptr;   // device copy of the descriptors
    int upload(hipStream_t stream);
    void release();
};

// Scratch of the XCD-binned find (hashset.hip): bin records, per-(segment,
// XCD) counters and descriptors, and an overflow flag.
struct BinScratch {
    uint32_t* buf = nullptr;
    size_t bufWords = 0;
    uint32_t* count = nullptr;    // [64 segments][8], 128 bytes apart
    void* desc = nullptr;         // BinDesc[64][8]
    uint32_t* overflow = nullptr;
    void release();
};

// Launch helpers implemented in hashset.hip.
// ev0 / ev1 (optional) are recorded right before the first and after the
// last kernel of the operation, after every descriptor upload.
int launch_bulk(const TableSet& ts, const uint32_t* words, const svdag_segment* segs, uint32_t nseg,
                uint32_t* out, int variant, bool insert, hipStream_t stream, uint32_t* segBuf, BinScratch* bins,
                hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);

inline uint32_t grid_for(uint64_t work, uint32_t perBlock, uint32_t cap = 8192)
{
    uint64_t g = (work + perBlock - 1) / perBlock;
    if (g < 1)
        g = 1;
    if (g > cap)
        g = cap;
    return (uint32_t)g;
}

} // namespace svdag
