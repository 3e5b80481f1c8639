// Export of the reachable DAG and its canonical form (host side).
//
// Canonical form: every reachable (level, item) is re-expressed with its
// children replaced by their canonical ids at the level below; the unique
// contents of each level are sorted (size, then words) and numbered by rank.
// The dump is, for level = leaf level down to 0: [level, count, then per item
// size, words...], followed by the root's canonical id.  Allocation order and
// handle values therefore do not affect it.  hash = FNV-1a 64 of the dump words.
#include "dag_internal.h"
#include <algorithm>
#include <cstring>
#include <map>
#include <unordered_map>
#include <unordered_set>

namespace svdag {

struct HostDag {
    uint32_t levels = 0, leafLevel = 0, M = 0, minLeaf = 0, leafSizeBits = 0;
    std::vector<std::vector<uint32_t>> items; // per table: slots [fresh][kSlots][words]
    std::vector<std::vector<uint32_t>> tags;  // per table: tag lines [fresh][32]
    std::vector<uint32_t> fresh;
    std::vector<uint32_t> fullyFilledLevel0;  // level-0 fully-filled node per material

    // Words of a stored item; nullptr if the pointer is out of range or its
    // slot is free (tag 0), i.e. the handle dangles.
    const uint32_t* item(uint32_t s, uint32_t ptr) const
    {
        const uint32_t slab = ptr >> kSlotBits, slot = ptr & ((1u << kSlotBits) - 1u);
        if (s < 2 || s - 2 >= items.size() || slab >= fresh[s - 2] || slot >= kSlots)
            return nullptr;
        const uint32_t tag = (tags[s - 2][(size_t)slab * kTagWords + slot / 4] >> (8 * (slot % 4))) & 0xFFu;
        if (!tag)
            return nullptr;
        return items[s - 2].data() + ((size_t)slab * kSlots + slot) * s;
    }
};

static int copy_dag_to_host(svdag_ctx* ctx, HostDag& h)
{
    const DagDev& d = ctx->dag;
    h.levels = d.levels;
    h.leafLevel = d.leafLevel;
    h.M = d.M;
    h.minLeaf = d.minLeafWords;
    h.leafSizeBits = d.leafSizeBits;
    h.items.resize(ctx->ts.tables.size());
    h.tags.resize(ctx->ts.tables.size());
    h.fresh.resize(ctx->ts.tables.size());
    h.fullyFilledLevel0.resize(d.numMaterials);
    SVDAG_HIP(hipMemcpyAsync(h.fullyFilledLevel0.data(), ctx->dFullyFilled, d.numMaterials * 4, hipMemcpyDeviceToHost,
                             ctx->stream));
    for (size_t i = 0; i < ctx->ts.tables.size(); ++i) {
        const HostTable& t = ctx->ts.tables[i];
        AllocState a {};
        SVDAG_TRY(table_read_alloc(t, a, ctx->stream));
        const uint32_t fresh = std::min(a.fresh, t.slabCap);
        h.fresh[i] = fresh;
        h.items[i].resize((size_t)fresh * kSlots * t.itemWords);
        h.tags[i].resize((size_t)fresh * kTagWords);
        SVDAG_HIP(hipMemcpyAsync(h.items[i].data(), t.dev.items, h.items[i].size() * 4, hipMemcpyDeviceToHost,
                                 ctx->stream));
        SVDAG_HIP(hipMemcpyAsync(h.tags[i].data(), t.dev.tags, h.tags[i].size() * 4, hipMemcpyDeviceToHost,
                                 ctx->stream));
    }
    SVDAG_HIP(hipStreamSynchronize(ctx->stream));
    return SVDAG_OK;
}

struct Rec {
    uint32_t level, handle;
    std::vector<uint32_t> words;
};

static uint32_t h_leaf_size(const HostDag& h, uint32_t handle)
{
    return h.leafSizeBits ? h.minLeaf + (handle >> (32u - h.leafSizeBits)) : h.minLeaf;
}
static uint32_t h_leaf_ptr(const HostDag& h, uint32_t handle)
{
    return h.leafSizeBits ? handle & ((1u << (32u - h.leafSizeBits)) - 1u) : handle;
}

// Breadth-first walk of every (level, handle) reachable from the roots.
static int reachable(const HostDag& h, const std::vector<uint32_t>& roots, std::vector<std::vector<Rec>>& levels)
{
    levels.assign(h.leafLevel + 1, {});
    std::vector<uint32_t> cur;
    for (uint32_t r : roots)
        if (r != kSentinel)
            cur.push_back(r);
    for (uint32_t L = 0; L <= h.leafLevel; ++L) {
        std::unordered_set<uint32_t> seen;
        std::vector<uint32_t> next;
        for (uint32_t hd : cur) {
            if (!seen.insert(hd).second)
                continue;
            const bool leaf = L == h.leafLevel;
            const uint32_t s = leaf ? h_leaf_size(h, hd) : node_handle_size(hd);
            const uint32_t* p = h.item(s, leaf ? h_leaf_ptr(h, hd) : node_handle_ptr(hd));
            if (!p) {
                set_error("export: reachable handle points at a free or out-of-range slot");
                return SVDAG_E_INVALID;
            }
            Rec r { L, hd, std::vector<uint32_t>(p, p + s) };
            if (!leaf)
                for (uint32_t j = 1; j < s; ++j)
                    next.push_back(p[j]);
            levels[L].push_back(std::move(r));
        }
        cur.swap(next);
    }
    return SVDAG_OK;
}

static uint64_t fnv1a(const std::vector<uint32_t>& words)
{
    uint64_t x = 1469598103934665603ull;
    for (uint32_t w : words)
        for (int b = 0; b < 4; ++b) {
            x ^= (w >> (8 * b)) & 0xFFu;
            x *= 1099511628211ull;
        }
    return x;
}

static bool content_less(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b)
{
    if (a.size() != b.size())
        return a.size() < b.size();
    return a < b;
}

static void canonicalize(const HostDag& h, uint32_t root, const std::vector<std::vector<Rec>>& levels,
                         std::vector<uint32_t>& dump, std::vector<uint32_t>& counts)
{
    std::unordered_map<uint32_t, uint32_t> idBelow;
    counts.assign(h.leafLevel + 1, 0);
    for (int L = (int)h.leafLevel; L >= 0; --L) {
        std::vector<std::vector<uint32_t>> contents;
        contents.reserve(levels[L].size());
        for (const Rec& r : levels[L]) {
            std::vector<uint32_t> c = r.words;
            if ((uint32_t)L != h.leafLevel)
                for (size_t j = 1; j < c.size(); ++j)
                    c[j] = idBelow.at(c[j]);
            contents.push_back(std::move(c));
        }
        std::vector<std::vector<uint32_t>> uniq = contents;
        std::sort(uniq.begin(), uniq.end(), content_less);
        uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
        std::unordered_map<uint32_t, uint32_t> ids;
        for (size_t i = 0; i < levels[L].size(); ++i) {
            const auto it = std::lower_bound(uniq.begin(), uniq.end(), contents[i], content_less);
            ids[levels[L][i].handle] = (uint32_t)(it - uniq.begin());
        }
        dump.push_back((uint32_t)L);
        dump.push_back((uint32_t)uniq.size());
        for (const auto& u : uniq) {
            dump.push_back((uint32_t)u.size());
            dump.insert(dump.end(), u.begin(), u.end());
        }
        counts[L] = (uint32_t)uniq.size();
        idBelow.swap(ids);
    }
    dump.push_back(root == kSentinel ? kSentinel : idBelow.at(root));
}

} // namespace svdag

using namespace svdag;

extern "C" {

int svdag_export_tables(svdag_ctx* ctx, uint32_t* buf, size_t* len)
{
    if (!ctx || !len)
        return SVDAG_E_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HostDag h;
    SVDAG_TRY(copy_dag_to_host(ctx, h));
    std::vector<std::vector<Rec>> levels;
    std::vector<uint32_t> roots { ctx->root };
    roots.insert(roots.end(), h.fullyFilledLevel0.begin(), h.fullyFilledLevel0.end());
    SVDAG_TRY(reachable(h, roots, levels));
    std::vector<uint32_t> out { h.levels, h.M, ctx->root, 0u };
    uint32_t n = 0;
    for (const auto& lv : levels)
        for (const Rec& r : lv) {
            out.push_back(r.level);
            out.push_back(r.handle);
            out.push_back((uint32_t)r.words.size());
            out.insert(out.end(), r.words.begin(), r.words.end());
            ++n;
        }
    out[3] = n;
    const size_t need = out.size();
    if (buf && *len >= need)
        std::memcpy(buf, out.data(), need * 4);
    *len = need;
    return SVDAG_OK;
}

int svdag_export_canonical(svdag_ctx* ctx, uint64_t* hash, uint32_t* buf, size_t* len, uint32_t* counts)
{
    if (!ctx || !hash)
        return SVDAG_E_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HostDag h;
    SVDAG_TRY(copy_dag_to_host(ctx, h));
    std::vector<std::vector<Rec>> levels;
    SVDAG_TRY(reachable(h, { ctx->root }, levels));
    std::vector<uint32_t> dump, cnt;
    canonicalize(h, ctx->root, levels, dump, cnt);
    *hash = fnv1a(dump);
    if (counts)
        for (uint32_t L = 0; L < 32; ++L)
            counts[L] = L < cnt.size() ? cnt[L] : 0u;
    if (len) {
        if (buf && *len >= dump.size())
            std::memcpy(buf, dump.data(), dump.size() * 4);
        *len = dump.size();
    }
    return SVDAG_OK;
}

} // extern "C"
