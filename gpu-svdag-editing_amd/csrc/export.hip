// Export of the reachable DAG and its canonical form (host side).
//
// Canonical form: every reachable (level, item) is re-expressed with its
// children replaced by their canonical ids at the level below; the unique
// contents of each level are sorted (size, then words) and numbered by rank.
// The dump is, for level = leaf level down to 0: [level, count, then per item
// size, words...], followed by the root's canonical id.  Allocation order and
// handle values therefore do not affect it.  hash = FNV-1a 64 of the dump words.
#include "dag_internal.h"
#include "host_threads.h"
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <map>
#include <unordered_map>
#include <unordered_set>
#include <atomic>
#include <functional>
#include <thread>

namespace svdag {

struct HostDag {
    uint32_t levels = 0, leafLevel = 0, M = 0, minLeaf = 0, leafSizeBits = 0;
    std::vector<std::vector<uint32_t>> lines; // per table: [fresh][32] (tags, next, items)
    std::vector<uint32_t> fresh, slots, tagWords;
    std::vector<uint32_t> fullyFilledLevel0;  // level-0 fully-filled node per material

    // Words of a stored item; nullptr if the pointer is out of range or its
    // slot is free (tag 0), i.e. the handle dangles.
    const uint32_t* item(uint32_t s, uint32_t ptr) const
    {
        const uint32_t slab = ptr >> kSlotBits, slot = ptr & ((1u << kSlotBits) - 1u);
        if (s < 2 || s - 2 >= lines.size() || slab >= fresh[s - 2] || slot >= slots[s - 2])
            return nullptr;
        const uint32_t* line = lines[s - 2].data() + (size_t)slab * kLineWords;
        if (!((line[slot / 4] >> (8 * (slot % 4))) & 0xFFu))
            return nullptr;
        return line + tagWords[s - 2] + 1u + slot * s;
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
    h.lines.resize(ctx->ts.tables.size());
    h.fresh.resize(ctx->ts.tables.size());
    h.slots.resize(ctx->ts.tables.size());
    h.tagWords.resize(ctx->ts.tables.size());
    h.fullyFilledLevel0.resize(d.numMaterials);
    SVDAG_HIP(hipMemcpyAsync(h.fullyFilledLevel0.data(), ctx->dFullyFilled, d.numMaterials * 4, hipMemcpyDeviceToHost,
                             ctx->stream));
    for (size_t i = 0; i < ctx->ts.tables.size(); ++i) {
        const HostTable& t = ctx->ts.tables[i];
        AllocState a {};
        SVDAG_TRY(table_read_alloc(t, a, ctx->stream));
        const uint32_t fresh = std::min(a.fresh, t.slabCap);
        h.fresh[i] = fresh;
        h.slots[i] = t.dev.slots;
        h.tagWords[i] = t.dev.tagWords;
        h.lines[i].resize((size_t)fresh * kLineWords);
        SVDAG_HIP(hipMemcpyAsync(h.lines[i].data(), t.dev.lines, h.lines[i].size() * 4, hipMemcpyDeviceToHost,
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
    const uint32_t idx = h.leafSizeBits ? handle & ((1u << (32u - h.leafSizeBits)) - 1u) : handle;
    return index_to_ptr(idx, h_leaf_size(h, handle));
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

uint64_t fnv1a(const std::vector<uint32_t>& words)
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

// The canonical dump of a large DAG without per-item allocations: the
// reachable handles of every level as a sorted array (children gathered,
// sorted, deduplicated), contents in one flat array per level with the
// children replaced by their canonical ids (a hash map of the level below),
// and the level's items ordered by a 64-bit key of (size, word 0, top of
// word 1) with full comparisons only on ties.  Gathers and sorts run on
// host threads.  Same dump as canonicalize(): contents sorted by (size,
// words), numbered by rank.
int canonical_dump(svdag_ctx* ctx, std::vector<uint32_t>& dump, std::vector<uint32_t>& counts)
{
    HostDag h;
    SVDAG_TRY(copy_dag_to_host(ctx, h));
    const uint32_t leafLevel = h.leafLevel;
    const uint32_t T = export_threads();
    std::vector<std::vector<uint32_t>> lvl(leafLevel + 1);
    auto itemOf = [&](uint32_t L, uint32_t hd, uint32_t& s) -> const uint32_t* {
        const bool leaf = L == leafLevel;
        s = leaf ? h_leaf_size(h, hd) : node_handle_size(hd);
        return h.item(s, leaf ? h_leaf_ptr(h, hd) : node_handle_ptr(hd));
    };
    std::atomic<bool> bad { false };
    if (ctx->root != kSentinel)
        lvl[0].push_back(ctx->root);
    for (uint32_t L = 0; L <= leafLevel; ++L) {
        std::vector<uint32_t>& cur = lvl[L];
        parallel_sort(cur, T, std::less<uint32_t>());
        cur.erase(std::unique(cur.begin(), cur.end()), cur.end());
        if (L == leafLevel)
            break;
        std::vector<std::vector<uint32_t>> part(T);
        parallel_chunks(cur.size(), T, [&](size_t b, size_t e, uint32_t t) {
            for (size_t i = b; i < e; ++i) {
                uint32_t s = 0;
                const uint32_t* p = itemOf(L, cur[i], s);
                if (!p) {
                    bad = true;
                    return;
                }
                part[t].insert(part[t].end(), p + 1, p + s);
            }
        });
        if (bad) {
            set_error("export: reachable handle points at a free or out-of-range slot");
            return SVDAG_E_INVALID;
        }
        std::vector<uint32_t>& next = lvl[L + 1];
        for (auto& q : part)
            next.insert(next.end(), q.begin(), q.end());
    }
    counts.assign(leafLevel + 1, 0);
    std::vector<uint32_t> idBelow; // canonical id of lvl[L + 1][i]
    // Open-addressing map from the level below's handles to their canonical
    // ids (handles are never kSentinel, which marks a free slot).
    std::vector<uint64_t> idMap;
    uint64_t idMask = 0;
    auto idOf = [&](uint32_t w) -> uint32_t {
        for (uint64_t q = mix32(w) & idMask;; q = (q + 1) & idMask) {
            const uint32_t key = (uint32_t)(idMap[q] >> 32);
            if (key == w)
                return (uint32_t)idMap[q];
            if (key == kSentinel)
                return kSentinel;
        }
    };
    for (int L = (int)leafLevel; L >= 0; --L) {
        const std::vector<uint32_t>& hs = lvl[L];
        const size_t n = hs.size();
        const bool leaf = (uint32_t)L == leafLevel;
        // Item sizes, then offsets, then the contents with children mapped.
        std::vector<uint64_t> off(n + 1, 0);
        parallel_chunks(n, T, [&](size_t b, size_t e, uint32_t) {
            for (size_t i = b; i < e; ++i) {
                uint32_t s = 0;
                if (!itemOf((uint32_t)L, hs[i], s))
                    bad = true;
                off[i + 1] = s;
            }
        });
        if (bad) {
            set_error("export: reachable handle points at a free or out-of-range slot");
            return SVDAG_E_INVALID;
        }
        for (size_t i = 0; i < n; ++i)
            off[i + 1] += off[i];
        std::vector<uint32_t> words(off[n]);
        parallel_chunks(n, T, [&](size_t b, size_t e, uint32_t) {
            for (size_t i = b; i < e; ++i) {
                uint32_t s = 0;
                const uint32_t* p = itemOf((uint32_t)L, hs[i], s);
                uint32_t* w = words.data() + off[i];
                w[0] = p[0];
                for (uint32_t j = 1; j < s; ++j) {
                    w[j] = leaf ? p[j] : idOf(p[j]);
                    if (w[j] == kSentinel && !leaf)
                        bad = true;
                }
            }
        });
        if (bad) {
            set_error("export: a child handle is missing from the level below");
            return SVDAG_E_INVALID;
        }
        struct Key {
            uint64_t k;
            uint32_t i;
        };
        std::vector<Key> order(n);
        parallel_chunks(n, T, [&](size_t b, size_t e, uint32_t) {
            for (size_t i = b; i < e; ++i) {
                const uint32_t s = (uint32_t)(off[i + 1] - off[i]);
                const uint32_t* w = words.data() + off[i];
                order[i] = { ((uint64_t)s << 59) | ((uint64_t)w[0] << 27) | (uint64_t)(s > 1 ? w[1] >> 5 : 0u),
                             (uint32_t)i };
            }
        });
        auto less = [&](const Key& a, const Key& b) {
            if (a.k != b.k)
                return a.k < b.k;
            const uint32_t* wa = words.data() + off[a.i];
            const uint32_t* wb = words.data() + off[b.i];
            const uint32_t s = (uint32_t)(off[a.i + 1] - off[a.i]); // equal keys: equal sizes
            return std::lexicographical_compare(wa + 1, wa + s, wb + 1, wb + s);
        };
        parallel_sort(order, T, less);
        std::vector<uint32_t> ids(n);
        dump.push_back((uint32_t)L);
        const size_t countAt = dump.size();
        dump.push_back(0u);
        uint32_t uniq = 0;
        for (size_t r = 0; r < n; ++r) {
            const Key& cur = order[r];
            const bool same = r > 0 && !less(order[r - 1], cur);
            if (!same) {
                const uint32_t s = (uint32_t)(off[cur.i + 1] - off[cur.i]);
                dump.push_back(s);
                dump.insert(dump.end(), words.begin() + (ptrdiff_t)off[cur.i], words.begin() + (ptrdiff_t)off[cur.i + 1]);
                ++uniq;
            }
            ids[cur.i] = uniq - 1u;
        }
        dump[countAt] = uniq;
        counts[L] = uniq;
        idBelow.swap(ids);
        uint64_t cap = 64;
        while (cap < 2 * (uint64_t)n)
            cap <<= 1;
        idMap.assign(cap, (uint64_t)kSentinel << 32);
        idMask = cap - 1;
        for (size_t i = 0; i < n; ++i) {
            uint64_t q = mix32(hs[i]) & idMask;
            while ((uint32_t)(idMap[q] >> 32) != kSentinel)
                q = (q + 1) & idMask;
            idMap[q] = ((uint64_t)hs[i] << 32) | idBelow[i];
        }
    }
    dump.push_back(ctx->root == kSentinel ? kSentinel : idBelow.at(0));
    return SVDAG_OK;
}

// Find-or-insert one level's items (unique among themselves) into the DAG's
// tables with the bulk kernels; fills the handles in input order.
static int insert_level(svdag_ctx* ctx, const std::vector<std::vector<uint32_t>>& items, bool leaf,
                        std::vector<uint32_t>& handles)
{
    const size_t n = items.size();
    handles.assign(n, kSentinel);
    if (n == 0)
        return SVDAG_OK;
    uint32_t* segBuf = nullptr;
    SVDAG_HIP(hipMalloc(&segBuf, 64 * 5 * 4));
    struct Free {
        std::vector<void*> p;
        ~Free()
        {
            for (void* q : p)
                if (q)
                    (void)hipFree(q);
        }
    } fr { { segBuf } };
    // Items grouped by size into segments; `order` maps segment rows back.
    auto run = [&](const std::vector<size_t>& which, bool insert, std::vector<uint32_t>& ptrs) -> int {
        std::vector<std::vector<size_t>> bySize(kMaxItemWords + 1);
        for (size_t i : which)
            bySize[items[i].size()].push_back(i);
        std::vector<uint32_t> words;
        std::vector<svdag_segment> segs;
        std::vector<size_t> order;
        for (uint32_t sz = 2; sz <= kMaxItemWords; ++sz) {
            if (bySize[sz].empty())
                continue;
            segs.push_back({ sz, (uint32_t)bySize[sz].size(), (uint32_t)words.size(), (uint32_t)order.size() });
            for (size_t i : bySize[sz]) {
                words.insert(words.end(), items[i].begin(), items[i].end());
                order.push_back(i);
            }
        }
        ptrs.assign(items.size(), kNotFound);
        if (order.empty())
            return SVDAG_OK;
        uint32_t *dw = nullptr, *dout = nullptr;
        SVDAG_HIP(hipMalloc(&dw, words.size() * 4));
        fr.p.push_back(dw);
        SVDAG_HIP(hipMalloc(&dout, order.size() * 4));
        fr.p.push_back(dout);
        SVDAG_HIP(hipMemcpy(dw, words.data(), words.size() * 4, hipMemcpyHostToDevice));
        SVDAG_TRY(launch_bulk(ctx->ts, dw, segs.data(), (uint32_t)segs.size(), dout, 16, insert, ctx->stream, segBuf,
                              nullptr));
        std::vector<uint32_t> out(order.size());
        SVDAG_HIP(hipMemcpyAsync(out.data(), dout, out.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
        SVDAG_HIP(hipStreamSynchronize(ctx->stream));
        for (size_t k = 0; k < order.size(); ++k)
            ptrs[order[k]] = out[k];
        return SVDAG_OK;
    };
    std::vector<size_t> all(n);
    for (size_t i = 0; i < n; ++i)
        all[i] = i;
    std::vector<uint32_t> found;
    SVDAG_TRY(run(all, false, found));
    std::vector<size_t> missing;
    for (size_t i = 0; i < n; ++i)
        if (found[i] == kNotFound)
            missing.push_back(i);
    for (int attempt = 0; attempt < 8 && !missing.empty(); ++attempt) {
        std::vector<uint32_t> ins;
        SVDAG_TRY(run(missing, true, ins));
        bool oom = false;
        for (auto& t : ctx->ts.tables) {
            AllocState a {};
            SVDAG_TRY(table_read_alloc(t, a, ctx->stream));
            oom |= a.oom != 0;
        }
        if (!oom) {
            for (size_t i : missing)
                found[i] = ins[i];
            missing.clear();
            break;
        }
        // Some inserts failed: grow the arenas and find-or-insert the rest
        // again (the ones that landed are found this time).
        {
            FrameQuiesce q(ctx);
            for (auto& t : ctx->ts.tables)
                SVDAG_TRY(table_grow(
                    t, (uint32_t)std::min<uint64_t>(2ull * t.slabCap, ctx->slabLimit[t.itemWords - 2]), ctx->stream));
            SVDAG_TRY(ctx_refresh_tables(ctx));
        }
        std::vector<uint32_t> again;
        SVDAG_TRY(run(missing, false, again));
        std::vector<size_t> still;
        for (size_t i : missing) {
            if (again[i] != kNotFound)
                found[i] = again[i];
            else
                still.push_back(i);
        }
        missing.swap(still);
    }
    if (!missing.empty()) {
        set_error("import: slab arenas could not hold the DAG");
        return SVDAG_E_OOM;
    }
    for (size_t i = 0; i < n; ++i) {
        const uint32_t sz = (uint32_t)items[i].size();
        handles[i] = leaf ? leaf_handle(ctx->dag, found[i], sz) : node_handle(found[i], sz);
    }
    return SVDAG_OK;
}

int import_dump(svdag_ctx* ctx, const uint32_t* d, size_t len)
{
    const DagDev& dag = ctx->dag;
    size_t p = 0;
    auto need = [&](size_t k) { return p + k <= len; };
    std::vector<uint32_t> below; // handles of the level below, by canonical id
    for (int L = (int)dag.leafLevel; L >= 0; --L) {
        if (!need(2) || d[p] != (uint32_t)L) {
            set_error("import: level " + std::to_string(L) + " missing or out of order");
            return SVDAG_E_INVALID;
        }
        const uint32_t count = d[p + 1];
        p += 2;
        const bool leaf = (uint32_t)L == dag.leafLevel;
        std::vector<std::vector<uint32_t>> items;
        items.reserve(count);
        for (uint32_t i = 0; i < count; ++i) {
            if (!need(1)) {
                set_error("import: truncated");
                return SVDAG_E_INVALID;
            }
            const uint32_t sz = d[p++];
            if (sz < 2 || sz > kMaxItemWords || !need(sz)) {
                set_error("import: bad item size");
                return SVDAG_E_INVALID;
            }
            std::vector<uint32_t> w(d + p, d + p + sz);
            p += sz;
            if (leaf) {
                if (sz != leaf_size_from_words(w[0], w[1], dag.M) || (w[0] == 0 && w[1] == 0)) {
                    set_error("import: leaf size does not match its mask");
                    return SVDAG_E_INVALID;
                }
            } else {
                if ((w[0] & 0xFFu) == 0 || sz != node_size_from_header(w[0]) || header_payload(w[0], dag.M) != 0) {
                    set_error("import: node header does not match its size");
                    return SVDAG_E_INVALID;
                }
                for (uint32_t j = 1; j < sz; ++j) {
                    if (w[j] >= below.size()) {
                        set_error("import: child id out of range");
                        return SVDAG_E_INVALID;
                    }
                    w[j] = below[w[j]];
                }
            }
            items.push_back(std::move(w));
        }
        std::vector<uint32_t> handles;
        SVDAG_TRY(insert_level(ctx, items, leaf, handles));
        below.swap(handles);
    }
    if (!need(1)) {
        set_error("import: root id missing");
        return SVDAG_E_INVALID;
    }
    const uint32_t rid = d[p++];
    if (p != len || (rid != kSentinel && rid >= below.size())) {
        set_error("import: bad root id or trailing words");
        return SVDAG_E_INVALID;
    }
    // Bulk inserts keep no reference counts: the next collection marks and
    // recounts.
    ctx->refsDirty = true;
    root_log(ctx);
    ctx->root = rid == kSentinel ? kSentinel : below[rid];
    SVDAG_HIP(hipMemcpyAsync(ctx->dRoot, &ctx->root, 4, hipMemcpyHostToDevice, ctx->stream));
    SVDAG_HIP(hipStreamSynchronize(ctx->stream));
    ctx->undo.clear();
    ctx->redo.clear();
    return SVDAG_OK;
}

constexpr uint32_t kFileMagic = 0x47445653u; // "SVDG"
constexpr uint32_t kFileVersion = 1u;

} // namespace svdag

using namespace svdag;

extern "C" {

int svdag_import_canonical(svdag_ctx* ctx, const uint32_t* buf, size_t len_words)
{
    if (!ctx || (!buf && len_words))
        return SVDAG_E_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    return import_dump(ctx, buf, len_words);
}

int svdag_save(svdag_ctx* ctx, const char* path)
{
    if (!ctx || !path)
        return SVDAG_E_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    std::vector<uint32_t> dump, cnt;
    SVDAG_TRY(canonical_dump(ctx, dump, cnt));
    const uint64_t hash = fnv1a(dump);
    const uint32_t head[6] = { kFileMagic, kFileVersion, ctx->dag.levels, ctx->dag.M, (uint32_t)(dump.size() & 0xFFFFFFFFu),
                               (uint32_t)(dump.size() >> 32) };
    FILE* f = std::fopen(path, "wb");
    if (!f) {
        set_error(std::string("svdag_save: cannot open ") + path);
        return SVDAG_E_INVALID;
    }
    bool ok = std::fwrite(head, 4, 6, f) == 6 && std::fwrite(dump.data(), 4, dump.size(), f) == dump.size()
        && std::fwrite(&hash, 8, 1, f) == 1;
    ok &= std::fclose(f) == 0;
    if (!ok) {
        set_error(std::string("svdag_save: write failed: ") + path);
        return SVDAG_E_INVALID;
    }
    return SVDAG_OK;
}

int svdag_load(svdag_ctx* ctx, const char* path)
{
    if (!ctx || !path)
        return SVDAG_E_INVALID;
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        set_error(std::string("svdag_load: cannot open ") + path);
        return SVDAG_E_INVALID;
    }
    uint32_t head[6] = {};
    std::vector<uint32_t> dump;
    uint64_t hash = 0;
    bool ok = std::fread(head, 4, 6, f) == 6 && head[0] == kFileMagic && head[1] == kFileVersion;
    const uint64_t n = ok ? ((uint64_t)head[5] << 32 | head[4]) : 0;
    if (ok && n < (1ull << 34)) {
        dump.resize(n);
        ok = std::fread(dump.data(), 4, n, f) == n && std::fread(&hash, 8, 1, f) == 1;
    } else {
        ok = false;
    }
    std::fclose(f);
    if (!ok) {
        set_error(std::string("svdag_load: not a DAG file of this format: ") + path);
        return SVDAG_E_INVALID;
    }
    if (head[2] != ctx->dag.levels || head[3] != ctx->dag.M) {
        set_error("svdag_load: the file's depth / material bits differ from the context's");
        return SVDAG_E_INVALID;
    }
    if (fnv1a(dump) != hash) {
        set_error("svdag_load: checksum mismatch");
        return SVDAG_E_INVALID;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    return import_dump(ctx, dump.data(), dump.size());
}


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
    std::vector<uint32_t> dump, cnt;
    SVDAG_TRY(canonical_dump(ctx, dump, cnt));
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
