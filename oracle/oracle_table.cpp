// ORACLE - test infrastructure only.  Nothing in the product links or calls
// this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load liboracle.so, and only as the checker / the timed CPU baseline.
//
// CPU restatement of the reference's default GPU hash table,
// CompactAccelerationHashTable on GpuHashTableBase with
// HASH_TABLE_STORE_SLABS_IN_TABLE=1 and HASH_TABLE_HASH_METHOD=SlabHashXor
// (configuration/gpu_hash_dag_definitions.h:1-10), host code path:
//   slab  = 8 tag words (32 x 8-bit "acceleration" tags, 0 = empty) | next |
//           active mask | 32 items        (compact_acceleration_hash_table.h:519-541)
//   slab size = 10 + 32 * itemWords      (compact_acceleration_hash_table.cu:5-13)
//   bucket = xorHash32xN(item) % numBuckets          (gpu_hash_table_base.h:258-273)
//   xorHash32xN = XOR_i ((2717583853*w_i + i*1117415409) mod 4294967291) (utils.h:386-403)
//   tag  = (XOR_i (32-i)*w_i) & 0xFF, 0 -> 1   (compact_acceleration_hash_table.h:493-503)
//   add  = first empty tag in chain order, else push a new slab after the
//          bucket's head slab                 (compact_acceleration_hash_table.h:27-61;
//                                              gpu_hash_table_base.h:128-170)
//   find = chain walk, tag match then word compare (compact_acceleration_hash_table.h:205-231)
//   pointer = slab << 6 | inBase << 5 | idx   (gpu_hash_table_base.h:114-125)
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr uint32_t kItemsPerSlab = 32;
constexpr uint32_t kTagWords = 8;
constexpr uint32_t kNextOff = 8;
constexpr uint32_t kActiveOff = 9;
constexpr uint32_t kItemsOff = 10;
constexpr uint32_t kNextSentinel = 0xFFFFFFFFu;
constexpr uint32_t kBaseNode = 0xFFFFFFFEu;
constexpr uint32_t kNotFound = 0xFFFFFFFFu;

uint32_t xor_hash(const uint32_t* p, uint32_t n)
{
    uint32_t seed = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t x = 2717583853u * p[i] + i * 1117415409u;
        seed ^= x % 4294967291u;
    }
    return seed;
}

uint32_t accel_hash(const uint32_t* p, uint32_t n)
{
    uint32_t out = 0;
    for (uint32_t i = 0; i < n; ++i)
        out ^= (32u - i) * p[i];
    out &= 0xFFu;
    return out == 0 ? 1u : out;
}

struct Table {
    uint32_t itemWords, numBuckets, slabWords;
    std::vector<uint32_t> table;   // numBuckets inline head slabs
    std::vector<uint32_t> pool;    // overflow slabs, slabWords each

    Table(uint32_t s, uint32_t nb)
        : itemWords(s)
        , numBuckets(nb)
        , slabWords(10 + kItemsPerSlab * s)
        , table((size_t)nb * slabWords, 0u)
    {
        for (uint32_t b = 0; b < nb; ++b)
            table[(size_t)b * slabWords + kNextOff] = kNextSentinel;
    }

    uint32_t* slab_ptr(uint32_t slabIdx, uint32_t bucket)
    {
        return slabIdx == kBaseNode ? &table[(size_t)bucket * slabWords] : &pool[(size_t)slabIdx * slabWords];
    }
    const uint32_t* slab_ptr(uint32_t slabIdx, uint32_t bucket) const
    {
        return slabIdx == kBaseNode ? &table[(size_t)bucket * slabWords] : &pool[(size_t)slabIdx * slabWords];
    }
    static uint32_t encode(uint32_t bucket, uint32_t slabIdx, uint32_t idx)
    {
        if (slabIdx == kBaseNode)
            return (bucket << 6) | (1u << 5) | idx;
        return (slabIdx << 6) | idx;
    }

    uint32_t add(const uint32_t* item)
    {
        const uint32_t tag = accel_hash(item, itemWords);
        const uint32_t bucket = xor_hash(item, itemWords) % numBuckets;
        uint32_t slabIdx = kBaseNode;
        while (true) {
            if (slabIdx == kNextSentinel) {
                // New slab: zeroed tags, linked right after the head slab.
                const uint32_t id = (uint32_t)(pool.size() / slabWords);
                pool.resize(pool.size() + slabWords, 0u);
                uint32_t* head = slab_ptr(kBaseNode, bucket);
                pool[(size_t)id * slabWords + kNextOff] = head[kNextOff];
                head[kNextOff] = id;
                slabIdx = id;
            }
            uint32_t* slab = slab_ptr(slabIdx, bucket);
            for (uint32_t i = 0; i < kTagWords; ++i) {
                for (uint32_t j = 0; j < 4; ++j) {
                    if (((slab[i] >> (8 * j)) & 0xFFu) == 0) {
                        slab[i] |= tag << (8 * j);
                        const uint32_t idx = i * 4 + j;
                        std::memcpy(&slab[kItemsOff + idx * itemWords], item, itemWords * 4);
                        return encode(bucket, slabIdx, idx);
                    }
                }
            }
            slabIdx = slab[kNextOff];
        }
    }

    uint32_t find(const uint32_t* item) const
    {
        const uint32_t tag = accel_hash(item, itemWords);
        const uint32_t bucket = xor_hash(item, itemWords) % numBuckets;
        uint32_t slabIdx = kBaseNode;
        while (slabIdx != kNextSentinel) {
            const uint32_t* slab = slab_ptr(slabIdx, bucket);
            for (uint32_t i = 0; i < kTagWords; ++i) {
                for (uint32_t j = 0; j < 4; ++j) {
                    if (((slab[i] >> (8 * j)) & 0xFFu) == tag) {
                        const uint32_t idx = i * 4 + j;
                        if (std::memcmp(&slab[kItemsOff + idx * itemWords], item, itemWords * 4) == 0)
                            return encode(bucket, slabIdx, idx);
                    }
                }
            }
            slabIdx = slab[kNextOff];
        }
        return kNotFound;
    }

    const uint32_t* decode(uint32_t ptr) const
    {
        const uint32_t idx = ptr & 31u;
        const uint32_t inBase = (ptr >> 5) & 1u;
        const uint32_t slab = ptr >> 6;
        const uint32_t* p = inBase ? &table[(size_t)slab * slabWords] : &pool[(size_t)slab * slabWords];
        return p + kItemsOff + idx * itemWords;
    }
};

} // namespace

extern "C" {

uint32_t ora_xor_hash(const uint32_t* item, uint32_t n) { return xor_hash(item, n); }
uint32_t ora_accel_hash(const uint32_t* item, uint32_t n) { return accel_hash(item, n); }

void* ora_table_create(uint32_t itemWords, uint32_t numBuckets)
{
    if (itemWords == 0 || numBuckets == 0)
        return nullptr;
    return new Table(itemWords, numBuckets);
}
void ora_table_destroy(void* t) { delete static_cast<Table*>(t); }

void ora_table_add(void* tp, const uint32_t* items, uint32_t n, uint32_t* out)
{
    Table* t = static_cast<Table*>(tp);
    for (uint32_t i = 0; i < n; ++i)
        out[i] = t->add(items + (size_t)i * t->itemWords);
}

// Read-only finds over `threads` host threads (the CPU baseline's timed loop).
void ora_table_find(void* tp, const uint32_t* items, uint32_t n, uint32_t* out, int threads)
{
    const Table* t = static_cast<const Table*>(tp);
    if (threads <= 1) {
        for (uint32_t i = 0; i < n; ++i)
            out[i] = t->find(items + (size_t)i * t->itemWords);
        return;
    }
    std::vector<std::thread> pool;
    const uint32_t chunk = (n + threads - 1) / threads;
    for (int k = 0; k < threads; ++k) {
        const uint32_t b = k * chunk, e = std::min<uint64_t>((uint64_t)b + chunk, n);
        pool.emplace_back([=] {
            for (uint32_t i = b; i < e; ++i)
                out[i] = t->find(items + (size_t)i * t->itemWords);
        });
    }
    for (auto& th : pool)
        th.join();
}

void ora_table_decode(void* tp, const uint32_t* ptrs, uint32_t n, uint32_t* outItems)
{
    const Table* t = static_cast<const Table*>(tp);
    for (uint32_t i = 0; i < n; ++i)
        std::memcpy(outItems + (size_t)i * t->itemWords, t->decode(ptrs[i]), t->itemWords * 4);
}

uint32_t ora_table_overflow_slabs(void* tp)
{
    const Table* t = static_cast<const Table*>(tp);
    return (uint32_t)(t->pool.size() / t->slabWords);
}

// Raw memory of the table: the head slabs (numBuckets * slabWords words) and
// the overflow slabs in allocation order (overflow slab k is slab id k).
void ora_table_raw(void* tp, uint32_t* heads, uint32_t* pool)
{
    const Table* t = static_cast<const Table*>(tp);
    if (heads)
        std::memcpy(heads, t->table.data(), t->table.size() * 4);
    if (pool)
        std::memcpy(pool, t->pool.data(), t->pool.size() * 4);
}

} // extern "C"
