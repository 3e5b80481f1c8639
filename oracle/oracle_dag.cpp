// ORACLE - test infrastructure only.  Nothing in the product links or calls
// this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load liboracle.so, and only as the checker.
//
// CPU restatement of the reference's GPU editable-DAG path, level by level as
// the reference runs it, over a plain hash-consed item store (content -> id;
// ids play the role of handles, which canonicalisation makes irrelevant):
//   encoding            my_gpu_base_dag.h:86-351 (header bits 344-351,
//                       LeafBuilder 102-150, get_material 293-316)
//   fully-filled chain  my_gpu_hash_dag.cu:38-63
//   initial DAG rule    my_gpu_hash_dag_factory.cpp:317-424 (full flag iff all
//                       8 children full with one material, 357-360, 406-407)
//   editors             my_gpu_dag_editors.h:45-119, 127-253, 401-536
//   edit SVO            create_edit_svo.cu:167-303 (inner), 320-426 (leaf)
//   merge               my_gpu_hash_dag_edits.cu:44-150 (updateChildPointers,
//                       discover_fully_filled gated at 337/589), 282-310, 626,
//                       and the no-leaf skip at 651
//   get_value           dag_utils.h:146-182
//   tracer              dag_tracer.cu:59-101, tracer.cu:45-273, 353-411
//   shadows / occlusion tracer.cu:274-351 (any-hit walk), 600-691,
//                       tracer.h:106-109, PCG_RNG utils.h:45-82
//   GC mark             my_gpu_hash_dag.cu:364-458 (roots: root + level-0
//                       fully-filled node of every material)
// Float rules: every float/double operation is written out in the
// reference's order; this file is compiled with -ffp-contract=off.  std::fma
// stands exactly where nvcc's default --fmad=true has a single possible
// fusion (an add whose one operand is a product): the sphere overlap's
// dist -= d*d, the per-pixel ray setup, the hit position and the shadow /
// occlusion ray origins.  The same sites are fmaf / fma in the HIP kernels.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <numbers>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace {

constexpr uint32_t SENT = 0xFFFFFFFFu;

struct VecHash {
    size_t operator()(const std::vector<uint32_t>& v) const
    {
        uint64_t h = 1469598103934665603ull ^ v.size();
        for (uint32_t x : v) {
            h ^= x;
            h *= 1099511628211ull;
            h ^= h >> 29;
        }
        return (size_t)h;
    }
};

int popc(uint32_t x) { return __builtin_popcount(x); }
int popcll(uint64_t x) { return __builtin_popcountll(x); }

// ------------------------------------------------------------------ encoding
struct Enc {
    uint32_t M;
    uint32_t header(uint32_t mask) const { return mask & 0xFFu; }
    uint32_t header_payload(uint32_t mask, uint32_t payload) const { return (mask & 0xFFu) | (payload << (9 + M)); }
    uint32_t header_full(uint32_t mask, uint32_t mat) const { return (mask & 0xFFu) | (1u << 8) | (M ? mat << 9 : 0u); }
    bool is_full(uint32_t header, uint32_t& mat) const
    {
        mat = M ? (header >> 9) & ((1u << M) - 1u) : 0u;
        return (header >> 8) & 1u;
    }
    uint32_t payload(uint32_t header) const { return header >> (9 + M); }
    uint32_t leaf_size(const uint32_t* leaf) const
    {
        const uint32_t n = popc(leaf[0]) + popc(leaf[1]);
        return 2 + ((n * M + 31) >> 5);
    }
    bool material(const uint32_t* leaf, uint32_t voxel, uint32_t& mat) const
    {
        const uint64_t bits = (uint64_t)leaf[0] | ((uint64_t)leaf[1] << 32);
        if (!(bits & (1ull << voxel)))
            return false;
        if (M == 0) {
            mat = 0;
            return true;
        }
        const uint32_t off = popcll(bits & ((1ull << voxel) - 1ull)) * M;
        const uint32_t mask = (1u << M) - 1u;
        mat = (leaf[2 + (off >> 5)] >> (off & 31)) & mask;
        if ((off & 31) + M > 32)
            mat |= (leaf[3 + (off >> 5)] << (32 - (off & 31))) & mask;
        return true;
    }
    uint32_t max_leaf() const { return 2 + (64 * M + 31) / 32; }
};

// LeafBuilder (my_gpu_base_dag.h:102-150): voxels in bit order, materials
// packed in set-voxel order, straddling words when 32 % M != 0.
struct LeafBuilder {
    std::vector<uint32_t> w;
    uint32_t M, voxel = 0, offset = 0;
    explicit LeafBuilder(uint32_t M_)
        : w(2 + (64 * M_ + 31) / 32, 0u)
        , M(M_)
    {
    }
    void set(uint32_t value)
    {
        w[voxel >> 5] |= 1u << (voxel & 31);
        if (M) {
            const uint32_t bits = offset * M;
            const uint32_t within = bits & 31;
            w[2 + (bits >> 5)] |= value << within;
            if (within + M > 32)
                w[3 + (bits >> 5)] = value >> (32 - within);
            ++offset;
        }
    }
    void next() { ++voxel; }
    std::vector<uint32_t> finish() const
    {
        std::vector<uint32_t> out(w.begin(), w.begin() + 2 + (offset * M + 31) / 32);
        return out;
    }
};

// ------------------------------------------------------------------ store
struct Dag {
    uint32_t levels, leafLevel, M, numMat;
    Enc enc;
    std::vector<std::vector<uint32_t>> items;
    std::unordered_map<std::vector<uint32_t>, uint32_t, VecHash> index;
    std::vector<uint32_t> ff; // [level][material]
    uint32_t root = SENT;

    Dag(uint32_t L, uint32_t M_)
        : levels(L)
        , leafLevel(L - 2)
        , M(M_)
        , numMat(1u << M_)
        , enc { M_ }
        , ff((size_t)L * (1u << M_), SENT)
    {
        for (uint32_t m = 0; m < numMat; ++m) {
            LeafBuilder b(M);
            for (int i = 0; i < 64; ++i) {
                b.set(m);
                b.next();
            }
            ff[leafLevel * numMat + m] = add(b.finish());
        }
        for (int level = (int)leafLevel - 1; level >= 0; --level)
            for (uint32_t m = 0; m < numMat; ++m) {
                std::vector<uint32_t> n { enc.header_full(0xFF, m) };
                for (int c = 0; c < 8; ++c)
                    n.push_back(ff[(level + 1) * numMat + m]);
                ff[level * numMat + m] = add(n);
            }
    }
    uint32_t add(const std::vector<uint32_t>& w)
    {
        auto it = index.find(w);
        if (it != index.end())
            return it->second;
        const uint32_t id = (uint32_t)items.size();
        items.push_back(w);
        index.emplace(w, id);
        return id;
    }
    const uint32_t* get(uint32_t id) const { return items[id].data(); }
    bool leaf_full(const uint32_t* leaf, uint32_t& mat) const
    {
        // is_leaf_fully_filled: all 64 set and every material equal to the first.
        if (leaf[0] != 0xFFFFFFFFu || leaf[1] != 0xFFFFFFFFu)
            return false;
        uint32_t first = 0;
        enc.material(leaf, 0, first);
        for (uint32_t v = 1; v < 64; ++v) {
            uint32_t m = 0;
            enc.material(leaf, v, m);
            if (m != first)
                return false;
        }
        mat = first;
        return true;
    }
};

// ------------------------------------------------------------------ terrain
double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
double lerp(double t, double a, double b) { return a + t * (b - a); }
double grad(uint32_t h, double x, double z)
{
    static const double gx[8] = { 1, -1, 1, -1, 1, -1, 0, 0 };
    static const double gz[8] = { 1, 1, -1, -1, 0, 0, 1, -1 };
    const uint32_t k = h & 7u;
    // Written so that the zero components contribute exactly nothing.
    if (k < 4)
        return gx[k] * x + gz[k] * z;
    if (k < 6)
        return gx[k] * x;
    return gz[k] * z;
}

struct Perlin {
    uint32_t P[512];
    explicit Perlin(uint64_t seed)
    {
        uint32_t p[256];
        for (uint32_t i = 0; i < 256; ++i)
            p[i] = i;
        uint64_t s = seed;
        for (uint32_t i = 255; i >= 1; --i) {
            s += 0x9E3779B97F4A7C15ull; // splitmix64
            uint64_t z = s;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            const uint32_t j = (uint32_t)(z % (i + 1));
            std::swap(p[i], p[j]);
        }
        for (uint32_t i = 0; i < 512; ++i)
            P[i] = p[i & 255];
    }
    double noise(double x, double z) const
    {
        const double X = std::floor(x), Z = std::floor(z);
        const uint32_t xi = (uint32_t)(int64_t)X & 255u, zi = (uint32_t)(int64_t)Z & 255u;
        const double xf = x - X, zf = z - Z;
        const double u = fade(xf), v = fade(zf);
        const double a = lerp(u, grad(P[P[xi] + zi], xf, zf), grad(P[P[xi + 1] + zi], xf - 1.0, zf));
        const double b = lerp(u, grad(P[P[xi] + zi + 1], xf, zf - 1.0), grad(P[P[xi + 1] + zi + 1], xf - 1.0, zf - 1.0));
        return lerp(v, a, b);
    }
};

struct TerrainSpec {
    uint64_t seed;
    double base, amplitude, frequency;
    uint32_t octaves, material;
};

int32_t height_at(const Perlin& pn, const TerrainSpec& t, uint32_t x, uint32_t z)
{
    double sum = 0.0, amp = 1.0, freq = 1.0, norm = 0.0;
    for (uint32_t o = 0; o < t.octaves; ++o) {
        sum = sum + amp * pn.noise(((double)x * t.frequency) * freq, ((double)z * t.frequency) * freq);
        norm = norm + amp;
        amp = amp * 0.5;
        freq = freq * 2.0;
    }
    return (int32_t)std::floor(t.base + t.amplitude * (sum / norm));
}

struct Terrain {
    uint32_t side;
    std::vector<int32_t> h;
    int32_t at(uint32_t x, uint32_t z) const { return h[(size_t)z * side + x]; }
};

// Factory rule over the heightfield: recursive octree, hash-consed, a node
// is flagged full iff its 8 children are full with one material.  Regions
// entirely solid (uniform material) or empty are decided from the height
// range, which yields exactly what the full recursion would.
struct Built {
    uint32_t handle;
    bool full;
    uint32_t mat;
};

// Height-band material rule of svdag_build_terrain_banded (a synthetic scene's
// rule with no reference counterpart): band y >> b plus a seeded 2-bit offset
// per 2^b column tile, masked to the material range.
struct Bands {
    bool on;
    uint32_t b, seed, material, mask;
    uint32_t at(uint32_t x, uint32_t y, uint32_t z) const
    {
        if (!on)
            return material;
        uint32_t h = seed ^ ((x >> b) * 0x9E3779B1u) ^ ((z >> b) * 0x85EBCA77u);
        h = (h ^ (h >> 15)) * 0x2C1B3C6Du;
        h ^= h >> 12;
        return (material + (y >> b) + (h & 3u)) & mask;
    }
};

Built build_terrain(Dag& d, const Terrain& t, const Bands& bands, uint32_t level, uint32_t px, uint32_t py, uint32_t pz)
{
    const uint32_t depth = d.levels - level;
    const uint32_t S = 1u << depth;
    const uint32_t x0 = px * S, z0 = pz * S;
    const int64_t y0 = (int64_t)py * S, y1 = y0 + S - 1;
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    for (uint32_t z = z0; z < z0 + S; ++z)
        for (uint32_t x = x0; x < x0 + S; ++x) {
            lo = std::min(lo, t.at(x, z));
            hi = std::max(hi, t.at(x, z));
        }
    if (y0 > hi)
        return { SENT, false, 0 };
    // A solid region within one band cube is that material's full node; a
    // wider one is built from its children (the factory rule decides).
    if (y1 <= lo && !(bands.on && depth > bands.b)) {
        const uint32_t mat = bands.at(x0, (uint32_t)y0, z0);
        return { d.ff[level * d.numMat + mat], true, mat };
    }
    if (level == d.leafLevel) {
        LeafBuilder b(d.M);
        for (uint32_t v = 0; v < 64; ++v) {
            const uint32_t x = (px << 2) | (((v >> 5) & 1u) << 1) | ((v >> 2) & 1u);
            const uint32_t y = (py << 2) | (((v >> 4) & 1u) << 1) | ((v >> 1) & 1u);
            const uint32_t z = (pz << 2) | (((v >> 3) & 1u) << 1) | (v & 1u);
            if ((int64_t)y <= t.at(x, z))
                b.set(bands.at(x, y, z));
            b.next();
        }
        const auto w = b.finish();
        if (w[0] == 0 && w[1] == 0)
            return { SENT, false, 0 };
        uint32_t fm = 0;
        const bool full = d.leaf_full(w.data(), fm);
        return { d.add(w), full, fm };
    }
    std::vector<uint32_t> node { 0u };
    uint32_t mask = 0;
    bool allFull = true;
    uint32_t fm = SENT;
    for (uint32_t c = 0; c < 8; ++c) {
        const Built b = build_terrain(d, t, bands, level + 1, (px << 1) | ((c >> 2) & 1u), (py << 1) | ((c >> 1) & 1u),
                                      (pz << 1) | (c & 1u));
        if (b.handle == SENT) {
            allFull = false;
            continue;
        }
        mask |= 1u << c;
        node.push_back(b.handle);
        if (!b.full || (fm != SENT && b.mat != fm))
            allFull = false;
        fm = b.full ? b.mat : fm;
    }
    if (!mask)
        return { SENT, false, 0 };
    const bool full = allFull && mask == 0xFF;
    node[0] = full ? d.enc.header_full(mask, fm) : d.enc.header(mask);
    return { d.add(node), full, full ? fm : 0u };
}

// ------------------------------------------------------------------ editors
enum Kind { SPHERE_ADD = 0, SPHERE_REMOVE = 1, BOX_ADD = 2, BOX_REMOVE = 3, PAINT = 4, COPY = 5 };
enum Op { SUBDIVIDE, FILL, EMPTY, KEEP };
enum State { PARTIAL, FILLED, EMPTY_ST };

struct Editor {
    uint32_t kind, material;
    float c[3], radius, radius2;       // sphere
    float bmin[3], bmax[3];            // box / copy destination
    float d2s[3];
    int32_t id2s[3];
    // Rotated copy (no reference counterpart): quarter turns about +Y, the
    // destination pivot pv, the source pivot qv = pv + id2s, and the
    // destination voxels [ivmin, ivmax] whose unit boxes overlap [bmin, bmax).
    uint32_t rot;
    int32_t pv[3], qv[3], ivmin[3], ivmax[3];
};

Editor make_editor(uint32_t kind, uint32_t material, const float* center, float radius, const float* src, const float* dst,
                   uint32_t rot = 0)
{
    Editor e {};
    e.kind = kind;
    e.material = material;
    // MyGpuSphereEditorBase: centre + 0.5, radius >= 0.1 (my_gpu_dag_editors.h:406-411).
    for (int a = 0; a < 3; ++a)
        e.c[a] = center[a] + 0.5f;
    e.radius = std::max(radius, 0.1f);
    e.radius2 = e.radius * e.radius;
    // MyGpuBoxEditorBase (50-55); MyGpuCopyEditor centres it at dest + (center - source) (157-164).
    float bc[3] = { center[0], center[1], center[2] };
    if (kind == COPY)
        for (int a = 0; a < 3; ++a) {
            bc[a] = dst[a] + (center[a] - src[a]);
            e.d2s[a] = -dst[a] + src[a];
            e.id2s[a] = (int32_t)e.d2s[a];
        }
    for (int a = 0; a < 3; ++a) {
        e.bmin[a] = bc[a] - radius;
        e.bmax[a] = radius == 0 ? bc[a] + 1.0f : bc[a] + radius;
    }
    e.rot = kind == COPY ? (rot & 3u) : 0u;
    if (e.rot)
        for (int a = 0; a < 3; ++a) {
            // voxel v overlaps [bmin, bmax) iff bmin - 1 < v < bmax
            e.ivmin[a] = (int32_t)std::floor(e.bmin[a]);
            e.ivmax[a] = (int32_t)std::ceil(e.bmax[a]) - 1;
            e.pv[a] = (int32_t)std::floor(bc[a] + 0.5f);
            e.qv[a] = e.pv[a] + e.id2s[a];
        }
    return e;
}

// Rotated copy map of an offset u from the destination pivot: k quarter turns
// of R(x, y, z) = (-z - 1, y, x), the turn about +Y that carries voxel cells
// onto voxel cells.  No reference counterpart.
void rot_cell(uint32_t k, const int64_t* u, int64_t* out)
{
    int64_t x = u[0], z = u[2];
    for (uint32_t i = 0; i < k; ++i) {
        const int64_t nx = -z - 1, nz = x;
        x = nx;
        z = nz;
    }
    out[0] = x;
    out[1] = u[1];
    out[2] = z;
}
// Source voxel of destination voxel v under the rotated copy, clamped at 0
// per axis like the translation (my_gpu_dag_editors.h:231).
void rot_source(const Editor& e, const int64_t* v, int64_t* s)
{
    const int64_t u[3] = { v[0] - e.pv[0], v[1] - e.pv[1], v[2] - e.pv[2] };
    int64_t r[3];
    rot_cell(e.rot, u, r);
    for (int a = 0; a < 3; ++a) {
        s[a] = e.qv[a] + r[a];
        if (s[a] < 0)
            s[a] = 0;
    }
}

float sq(float f) { return f * f; }

bool sphere_overlap(const Editor& e, const float* mn, const float* mx)
{
    float dist = e.radius2;
    for (int a = 0; a < 3; ++a) {
        // dist -= squared(d), contracted as nvcc's --fmad=true does (one fusion possible).
        if (e.c[a] < mn[a]) {
            const float d = e.c[a] - mn[a];
            dist = std::fma(-d, d, dist);
        } else if (e.c[a] > mx[a]) {
            const float d = e.c[a] - mx[a];
            dist = std::fma(-d, d, dist);
        }
    }
    return dist > 0;
}
bool sphere_contained(const Editor& e, const float* mn, const float* mx)
{
    return std::max(sq(mn[0] - e.c[0]), sq(mx[0] - e.c[0])) + std::max(sq(mn[1] - e.c[1]), sq(mx[1] - e.c[1]))
        + std::max(sq(mn[2] - e.c[2]), sq(mx[2] - e.c[2]))
        < e.radius2;
}
bool box_overlap(const float* bmn, const float* bmx, const float* mn, const float* mx)
{
    for (int a = 0; a < 3; ++a)
        if (bmn[a] >= mx[a] || bmx[a] <= mn[a] || mn[a] >= bmx[a] || mx[a] <= bmn[a])
            return false;
    return true;
}
bool box_contained(const float* bmn, const float* bmx, const float* mn, const float* mx)
{
    for (int a = 0; a < 3; ++a)
        if (!(bmn[a] <= mn[a] && bmx[a] >= mx[a]))
            return false;
    return true;
}

void box_of(const uint32_t* p, uint32_t depth, float* mn, float* mx)
{
    for (int a = 0; a < 3; ++a) {
        mn[a] = float(p[a] << depth);
        mx[a] = mn[a] + float(1u << depth);
    }
}

Op should_edit(const Editor& e, const uint32_t* p, uint32_t depth, State st)
{
    float mn[3], mx[3];
    box_of(p, depth, mn, mx);
    switch (e.kind) {
    case SPHERE_ADD:
    case SPHERE_REMOVE: {
        const bool overlaps = sphere_overlap(e, mn, mx);
        const float smn[3] = { mn[0] + 1.0f, mn[1] + 1.0f, mn[2] + 1.0f };
        const float smx[3] = { mx[0] - 1.0f, mx[1] - 1.0f, mx[2] - 1.0f };
        const bool contained = sphere_contained(e, smn, smx);
        if (!overlaps)
            return KEEP;
        if (contained)
            return e.kind == SPHERE_ADD ? FILL : EMPTY;
        return SUBDIVIDE;
    }
    case BOX_ADD:
    case BOX_REMOVE:
        if (box_contained(e.bmin, e.bmax, mn, mx))
            return e.kind == BOX_ADD ? FILL : EMPTY;
        if (box_overlap(e.bmin, e.bmax, mn, mx))
            return SUBDIVIDE;
        return KEEP;
    case PAINT: {
        const bool overlaps = sphere_overlap(e, mn, mx);
        const bool contained = sphere_contained(e, mn, mx);
        if (!overlaps)
            return KEEP;
        if (contained && st == FILLED)
            return FILL;
        if (st == EMPTY_ST)
            return EMPTY;
        return SUBDIVIDE;
    }
    case COPY:
        return box_overlap(e.bmin, e.bmax, mn, mx) ? SUBDIVIDE : KEEP;
    }
    return KEEP;
}

struct Parent {
    uint32_t p[3];
    uint32_t h;
};
struct CopyState {
    Parent par[8];
    int n = 0;
};

bool aabb_overlap(const float* lmn, const float* lmx, const float* rmn, const float* rmx)
{
    for (int a = 0; a < 3; ++a)
        if (lmn[a] >= rmx[a] || lmx[a] <= rmn[a] || rmn[a] >= lmx[a] || rmx[a] <= lmn[a])
            return false;
    return true;
}

// Rotated copy, node level: the destination voxels of the node inside the
// edit box form an integer range; its image under the map (corners mapped,
// then ordered, then clamped at 0) is the source range, and the source
// children overlapping it become the parents of the next level.  Exact
// integer arithmetic, so every destination voxel's source cell lies in one
// of the collected parents.
Op should_edit_copy_rot(const Editor& e, const Dag& d, const uint32_t* p, uint32_t depth, const CopyState& in, CopyState& out)
{
    out.n = 0;
    int64_t lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        const int64_t n0 = (int64_t)p[a] << depth, n1 = n0 + ((int64_t)1 << depth) - 1;
        lo[a] = std::max<int64_t>(n0, e.ivmin[a]);
        hi[a] = std::min<int64_t>(n1, e.ivmax[a]);
        if (lo[a] > hi[a])
            return KEEP;
    }
    int64_t a0[3], a1[3], slo[3], shi[3];
    rot_source(e, lo, a0);
    rot_source(e, hi, a1);
    for (int a = 0; a < 3; ++a) {
        slo[a] = std::min(a0[a], a1[a]);
        shi[a] = std::max(a0[a], a1[a]);
    }
    for (int k = 0; k < in.n; ++k) {
        const uint32_t* node = d.get(in.par[k].h);
        const uint32_t mask = node[0] & 0xFFu;
        for (uint32_t c = 0; c < 8; ++c) {
            if (!(mask & (1u << c)))
                continue;
            const uint32_t cp[3] = { (in.par[k].p[0] << 1) | ((c >> 2) & 1u), (in.par[k].p[1] << 1) | ((c >> 1) & 1u),
                                     (in.par[k].p[2] << 1) | (c & 1u) };
            bool ov = true;
            for (int a = 0; a < 3; ++a) {
                const int64_t c0 = (int64_t)cp[a] << depth, c1 = c0 + ((int64_t)1 << depth) - 1;
                ov = ov && c0 <= shi[a] && c1 >= slo[a];
            }
            if (ov && out.n < 8) {
                Parent& q = out.par[out.n++];
                std::memcpy(q.p, cp, sizeof(cp));
                q.h = node[1 + popc(mask & ((1u << c) - 1u))];
            }
        }
    }
    return out.n == 0 ? KEEP : SUBDIVIDE;
}

Op should_edit_copy(const Editor& e, const Dag& d, const uint32_t* p, uint32_t depth, const CopyState& in, CopyState& out)
{
    if (e.rot)
        return should_edit_copy_rot(e, d, p, depth, in, out);
    float mn[3], mx[3];
    box_of(p, depth, mn, mx);
    out.n = 0;
    if (!box_overlap(e.bmin, e.bmax, mn, mx))
        return KEEP;
    const float size = float(1u << depth);
    float smn[3], smx[3];
    for (int a = 0; a < 3; ++a) {
        smn[a] = (mn[a] > e.bmin[a] ? mn[a] : e.bmin[a]) + e.d2s[a];
        const float lim = e.bmax[a] + e.d2s[a];
        const float cand = smn[a] + size;
        smx[a] = cand < lim ? cand : lim;
    }
    for (int k = 0; k < in.n; ++k) {
        const uint32_t* node = d.get(in.par[k].h);
        const uint32_t mask = node[0] & 0xFFu;
        for (uint32_t c = 0; c < 8; ++c) {
            if (!(mask & (1u << c)))
                continue;
            const uint32_t cp[3] = { (in.par[k].p[0] << 1) | ((c >> 2) & 1u), (in.par[k].p[1] << 1) | ((c >> 1) & 1u),
                                     (in.par[k].p[2] << 1) | (c & 1u) };
            float cmn[3], cmx[3];
            box_of(cp, depth, cmn, cmx);
            if (aabb_overlap(cmn, cmx, smn, smx) && out.n < 8) {
                Parent& q = out.par[out.n++];
                std::memcpy(q.p, cp, sizeof(cp));
                q.h = node[1 + popc(mask & ((1u << c) - 1u))];
            }
        }
    }
    return out.n == 0 ? KEEP : SUBDIVIDE;
}

bool copy_value(const Editor& e, const Dag& d, const uint32_t* v, const CopyState& st, bool hasOld, uint32_t oldMat,
                uint32_t& mat)
{
    int64_t s[3];
    if (e.rot) {
        const int64_t vv[3] = { v[0], v[1], v[2] };
        rot_source(e, vv, s);
    } else {
        for (int a = 0; a < 3; ++a) {
            s[a] = (int32_t)v[a] + e.id2s[a];
            if (s[a] < 0)
                s[a] = 0;
        }
    }
    const Parent* found = nullptr;
    for (int k = 0; k < st.n; ++k)
        if (st.par[k].p[0] == (uint32_t)s[0] >> 2 && st.par[k].p[1] == (uint32_t)s[1] >> 2 && st.par[k].p[2] == (uint32_t)s[2] >> 2)
            found = &st.par[k];
    mat = oldMat;
    if (!found)
        return hasOld;
    const uint32_t bit = ((s[0] & 1) ? 4u : 0u) | ((s[1] & 1) ? 2u : 0u) | ((s[2] & 1) ? 1u : 0u) | ((s[0] & 2) ? 32u : 0u)
        | ((s[1] & 2) ? 16u : 0u) | ((s[2] & 2) ? 8u : 0u);
    uint32_t m;
    if (d.enc.material(d.get(found->h), bit, m)) {
        mat = m;
        return true;
    }
    return hasOld;
}

// ------------------------------------------------------------------ edit
struct Work {
    uint32_t p[3];
    uint32_t h;
    CopyState st;
};

// Returns 1 when the edit was merged, 0 when it made no leaf (skipped).
int apply_edit(Dag& d, const Editor& e, uint32_t* svoCounts)
{
    const bool copy = e.kind == COPY;
    std::vector<std::vector<Work>> work(d.leafLevel + 1);
    std::vector<std::vector<std::vector<uint32_t>>> nodes(d.leafLevel); // SVO nodes per level
    std::vector<std::vector<uint32_t>> imask(d.leafLevel);
    Work w0 {};
    w0.h = d.root;
    if (copy) {
        w0.st.n = 1;
        w0.st.par[0] = Parent { { 0, 0, 0 }, d.root };
    }
    work[0].push_back(w0);
    // Top-down (create_edit_svo.cu): children classified in child order.
    for (uint32_t L = 0; L < d.leafLevel; ++L) {
        const uint32_t childLevel = L + 1, childDepth = d.levels - childLevel;
        for (const Work& w : work[L]) {
            uint32_t childH[8], mask = 0;
            const uint32_t* old = w.h != SENT ? d.get(w.h) : nullptr;
            const uint32_t om = old ? old[0] & 0xFFu : 0u;
            uint32_t subMask = 0;
            for (uint32_t c = 0; c < 8; ++c) {
                uint32_t ch = (om >> c) & 1u ? old[1 + popc(om & ((1u << c) - 1u))] : SENT;
                const uint32_t cp[3] = { (w.p[0] << 1) | ((c >> 2) & 1u), (w.p[1] << 1) | ((c >> 1) & 1u), (w.p[2] << 1) | (c & 1u) };
                State st = PARTIAL;
                if (ch == SENT)
                    st = EMPTY_ST;
                else if (childLevel != d.leafLevel) {
                    uint32_t fm;
                    st = d.enc.is_full(d.get(ch)[0], fm) ? FILLED : PARTIAL;
                }
                Op op;
                CopyState cs;
                if (copy)
                    op = should_edit_copy(e, d, cp, childDepth, w.st, cs);
                else
                    op = should_edit(e, cp, childDepth, st);
                if (op == SUBDIVIDE) {
                    Work nw {};
                    std::memcpy(nw.p, cp, sizeof(cp));
                    nw.h = ch;
                    nw.st = cs;
                    ch = (uint32_t)work[childLevel].size();
                    work[childLevel].push_back(nw);
                    subMask |= 1u << c;
                } else if (op == EMPTY) {
                    ch = SENT;
                } else if (op == FILL) {
                    ch = d.ff[childLevel * d.numMat + e.material];
                }
                childH[c] = ch;
                if (ch != SENT)
                    mask |= 1u << c;
            }
            std::vector<uint32_t> n;
            if (mask) {
                n.push_back(d.enc.header_payload(mask, subMask));
                for (uint32_t c = 0; c < 8; ++c)
                    if (mask & (1u << c))
                        n.push_back(childH[c]);
            }
            nodes[L].push_back(n);
            imask[L].push_back(subMask);
        }
    }
    // Leaves: one voxel per lane in bit order (create_edit_svo.cu:320-426).
    std::vector<std::vector<uint32_t>> leaves;
    for (const Work& w : work[d.leafLevel]) {
        const uint32_t* old = w.h != SENT ? d.get(w.h) : nullptr;
        LeafBuilder b(d.M);
        for (uint32_t v = 0; v < 64; ++v) {
            uint32_t oldMat = 0;
            const bool hasOld = old && d.enc.material(old, v, oldMat);
            const uint32_t vp[3] = { (w.p[0] << 2) | (((v >> 5) & 1u) << 1) | ((v >> 2) & 1u),
                                     (w.p[1] << 2) | (((v >> 4) & 1u) << 1) | ((v >> 1) & 1u),
                                     (w.p[2] << 2) | (((v >> 3) & 1u) << 1) | (v & 1u) };
            const Op op = should_edit(e, vp, 0, hasOld ? FILLED : EMPTY_ST);
            bool has = false;
            uint32_t mat = 0;
            if (op == FILL || op == SUBDIVIDE) {
                if (copy)
                    has = copy_value(e, d, vp, w.st, hasOld, oldMat, mat);
                else if (e.kind == SPHERE_ADD || e.kind == BOX_ADD || e.kind == PAINT) {
                    has = true;
                    mat = e.material;
                }
            } else if (op == KEEP) {
                has = hasOld;
                mat = oldMat;
            }
            if (has)
                b.set(mat);
            b.next();
        }
        leaves.push_back(b.finish());
    }
    for (uint32_t L = 0; L <= d.leafLevel && svoCounts; ++L)
        svoCounts[L] = (uint32_t)work[L].size();
    if (leaves.empty())
        return 0; // my_gpu_hash_dag_edits.cu:651
    // Bottom-up merge with the fully-filled discovery gate.
    std::vector<uint32_t> below(leaves.size());
    uint32_t fullBelow = 0;
    for (size_t i = 0; i < leaves.size(); ++i) {
        const auto& l = leaves[i];
        uint32_t fm;
        if (d.leaf_full(l.data(), fm))
            ++fullBelow;
        below[i] = (l[0] == 0 && l[1] == 0) ? SENT : d.add(l);
    }
    bool gate = fullBelow > 0;
    for (int L = (int)d.leafLevel - 1; L >= 0; --L) {
        std::vector<uint32_t> cur(nodes[L].size());
        uint32_t fullHere = 0;
        for (size_t i = 0; i < nodes[L].size(); ++i) {
            const auto& n = nodes[L][i];
            if (n.empty()) {
                cur[i] = SENT;
                continue;
            }
            const uint32_t mask = n[0] & 0xFFu, im = imask[L][i];
            std::vector<uint32_t> out { 0u };
            uint32_t newMask = 0;
            size_t k = 1;
            for (uint32_t c = 0; c < 8; ++c) {
                if (!(mask & (1u << c)))
                    continue;
                uint32_t h = n[k++];
                if (im & (1u << c))
                    h = below[h];
                if (h != SENT) {
                    out.push_back(h);
                    newMask |= 1u << c;
                }
            }
            bool full = false;
            uint32_t fm = 0;
            if (gate && newMask == 0xFF) {
                full = true;
                for (uint32_t j = 0; j < 8 && full; ++j) {
                    uint32_t m = 0;
                    const bool f = ((uint32_t)L + 1 == d.leafLevel) ? d.leaf_full(d.get(out[1 + j]), m)
                                                                   : d.enc.is_full(d.get(out[1 + j])[0], m);
                    if (j == 0)
                        fm = m;
                    full = f && m == fm;
                }
            }
            out[0] = full ? d.enc.header_full(0xFF, fm) : d.enc.header(newMask);
            if (full)
                ++fullHere;
            cur[i] = newMask ? d.add(out) : SENT;
        }
        if (gate)
            gate = fullHere > 0;
        below.swap(cur);
    }
    d.root = below[0];
    return 1;
}

// ------------------------------------------------------------------ queries / canonical
uint32_t get_value(const Dag& d, uint32_t x, uint32_t y, uint32_t z)
{
    uint32_t h = d.root;
    if (h == SENT)
        return SENT;
    for (uint32_t level = 0; level < d.leafLevel; ++level) {
        const uint32_t* n = d.get(h);
        const uint32_t mask = n[0] & 0xFFu;
        const uint32_t sh = d.levels - (level + 1);
        const uint32_t c = (((x >> sh) & 1u) ? 4u : 0u) | (((y >> sh) & 1u) ? 2u : 0u) | (((z >> sh) & 1u) ? 1u : 0u);
        if (!(mask & (1u << c)))
            return SENT;
        h = n[1 + popc(mask & ((1u << c) - 1u))];
    }
    const uint32_t v = ((x & 1) ? 4u : 0u) | ((y & 1) ? 2u : 0u) | ((z & 1) ? 1u : 0u) | ((x & 2) ? 32u : 0u)
        | ((y & 2) ? 16u : 0u) | ((z & 2) ? 8u : 0u);
    uint32_t m;
    return d.enc.material(d.get(h), v, m) ? m : SENT;
}

// Item reader over either the oracle store or an exported record list.
using Reader = std::function<const std::vector<uint32_t>*(uint32_t level, uint32_t handle)>;

uint64_t fnv(const std::vector<uint32_t>& words)
{
    uint64_t x = 1469598103934665603ull;
    for (uint32_t w : words)
        for (int b = 0; b < 4; ++b) {
            x ^= (w >> (8 * b)) & 0xFFu;
            x *= 1099511628211ull;
        }
    return x;
}

// Canonical dump: per level (leaf level first) [level, count, (size, words)...]
// with unique contents sorted by (size, words) and children replaced by the
// child's rank; then the root's rank.
std::vector<uint32_t> canonical(uint32_t leafLevel, uint32_t root, const Reader& read, std::vector<uint32_t>& counts)
{
    std::vector<std::vector<uint32_t>> handles(leafLevel + 1);
    if (root != SENT) {
        std::vector<uint32_t> cur { root };
        for (uint32_t L = 0; L <= leafLevel; ++L) {
            std::sort(cur.begin(), cur.end());
            cur.erase(std::unique(cur.begin(), cur.end()), cur.end());
            handles[L] = cur;
            std::vector<uint32_t> next;
            if (L < leafLevel)
                for (uint32_t h : cur) {
                    const auto* w = read(L, h);
                    next.insert(next.end(), w->begin() + 1, w->end());
                }
            cur.swap(next);
        }
    }
    std::vector<uint32_t> dump;
    std::unordered_map<uint32_t, uint32_t> rankBelow;
    counts.assign(leafLevel + 1, 0);
    for (int L = (int)leafLevel; L >= 0; --L) {
        std::vector<std::vector<uint32_t>> content;
        for (uint32_t h : handles[L]) {
            std::vector<uint32_t> c = *read(L, h);
            if ((uint32_t)L < leafLevel)
                for (size_t j = 1; j < c.size(); ++j)
                    c[j] = rankBelow.at(c[j]);
            content.push_back(c);
        }
        auto less = [](const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) {
            return a.size() != b.size() ? a.size() < b.size() : a < b;
        };
        std::vector<std::vector<uint32_t>> sorted = content;
        std::sort(sorted.begin(), sorted.end(), less);
        sorted.erase(std::unique(sorted.begin(), sorted.end()), sorted.end());
        std::unordered_map<uint32_t, uint32_t> rank;
        for (size_t i = 0; i < handles[L].size(); ++i)
            rank[handles[L][i]] = (uint32_t)(std::lower_bound(sorted.begin(), sorted.end(), content[i], less) - sorted.begin());
        dump.push_back((uint32_t)L);
        dump.push_back((uint32_t)sorted.size());
        for (const auto& s : sorted) {
            dump.push_back((uint32_t)s.size());
            dump.insert(dump.end(), s.begin(), s.end());
        }
        counts[L] = (uint32_t)sorted.size();
        rankBelow.swap(rank);
    }
    dump.push_back(root == SENT ? SENT : rankBelow.at(root));
    return dump;
}

// ------------------------------------------------------------------ tracer
struct Ray {
    float o[3], d[3], inv[3], tmin, tmax;
};

uint32_t next_child(uint32_t order, uint32_t mask)
{
    for (uint32_t c = 0; c < 8; ++c)
        if (mask & (1u << (c ^ order)))
            return c ^ order;
    return 0;
}
float maxf(float a, float b) { return a > b ? a : b; }
float minf(float a, float b) { return a < b ? a : b; }

uint32_t isect_mask(bool isRoot, uint32_t levels, uint32_t level, const uint32_t* path, const Ray& r)
{
    const uint32_t shift = levels - level;
    const float radius = float(1u << (shift - 1));
    float cr[3], tmid[3], pmin[3], pmax[3];
    for (int a = 0; a < 3; ++a) {
        const float center = radius + float(path[a] << shift);
        cr[a] = center - r.o[a];
        tmid[a] = cr[a] * r.inv[a];
        const float sr = radius * std::fabs(r.inv[a]);
        pmin[a] = tmid[a] - sr;
        pmax[a] = tmid[a] + sr;
    }
    float tmin = maxf(maxf(pmin[0], maxf(pmin[1], pmin[2])), 0.0f);
    float tmax = minf(pmax[0], minf(pmax[1], pmax[2]));
    tmin = maxf(tmin, r.tmin);
    tmax = minf(tmax, r.tmax);
    if (isRoot && tmin >= tmax)
        return 0;
    uint32_t mask = 0;
    const float tm = 0.5f * (tmin + tmax);
    mask |= 1u << ((tm * r.d[0] >= cr[0] ? 4 : 0) + (tm * r.d[1] >= cr[1] ? 2 : 0) + (tm * r.d[2] >= cr[2] ? 1 : 0));
    const float eps = 1e-4f;
    static const uint32_t hiMask[3] = { 0xF0, 0xCC, 0xAA }, loMask[3] = { 0x0F, 0x33, 0x55 };
    for (int a = 0; a < 3; ++a) {
        if (!(tmin <= tmid[a] && tmid[a] <= tmax))
            continue;
        const int b0 = a == 0 ? 1 : 0, b1 = a == 2 ? 1 : 2;
        uint32_t A = 0, B = 0;
        const float q0 = tmid[a] * r.d[b0], q1 = tmid[a] * r.d[b1];
        if (q0 >= cr[b0] - eps) A |= hiMask[b0];
        if (q0 <= cr[b0] + eps) A |= loMask[b0];
        if (q1 >= cr[b1] - eps) B |= hiMask[b1];
        if (q1 <= cr[b1] + eps) B |= loMask[b1];
        mask |= A & B;
    }
    return mask & 0xFFu;
}

struct Hit {
    uint32_t x, y, z, material;
    float t, px, py, pz;
};

// intersect_ray + createSurfaceInteraction (tracer.cu:186-273, 353-389) for
// any ray: the first set voxel along it, its material, the entry-face t and
// position, and (optionally) the surface code (entry axis and normal sign).
Hit trace_ray(const Dag& d, const Ray& r, uint32_t* surfCode = nullptr)
{
    uint32_t path[3] = { 0, 0, 0 };
    uint32_t material = SENT;
    if (d.root != SENT) {
        const uint32_t order = (r.d[0] < 0.f ? 4 : 0) + (r.d[1] < 0.f ? 2 : 0) + (r.d[2] < 0.f ? 1 : 0);
        struct Entry {
            uint32_t index, childMask, visitMask;
        };
        std::vector<Entry> stack(d.levels);
        Entry cache { d.root, d.get(d.root)[0] & 0xFFu, 0 };
        cache.visitMask = cache.childMask & isect_mask(true, d.levels, 0, path, r);
        uint32_t level = 0, addr = 0;
        uint64_t leafBits = 0;
        for (;;) {
            uint32_t nl = level;
            while (nl > 0 && !cache.visitMask)
                cache = stack[--nl];
            if (nl == 0 && !cache.visitMask) {
                path[0] = path[1] = path[2] = 0;
                break;
            }
            for (int k = 0; k < 3; ++k)
                path[k] >>= (level - nl);
            level = nl;
            const uint32_t c = next_child(order, cache.visitMask);
            cache.visitMask &= ~(1u << c);
            path[0] = (path[0] << 1) | ((c >> 2) & 1u);
            path[1] = (path[1] << 1) | ((c >> 1) & 1u);
            path[2] = (path[2] << 1) | (c & 1u);
            stack[level] = cache;
            ++level;
            if (level == d.levels) {
                // morton & 63 == leaf voxel bit (tracer.cu:237-241)
                const uint32_t v = ((path[0] & 1) << 2) | ((path[1] & 1) << 1) | (path[2] & 1) | ((path[0] & 2) << 4)
                    | ((path[1] & 2) << 3) | ((path[2] & 2) << 2);
                uint32_t m;
                material = d.enc.material(d.get(addr), v, m) ? m : SENT;
                break;
            }
            if (level < d.leafLevel) {
                const uint32_t* n = d.get(cache.index);
                cache.index = n[1 + popc(cache.childMask & ((1u << c) - 1u))];
                cache.childMask = d.get(cache.index)[0] & 0xFFu;
            } else if (level == d.leafLevel) {
                const uint32_t* n = d.get(cache.index);
                addr = n[1 + popc(cache.childMask & ((1u << c) - 1u))];
                const uint32_t* leaf = d.get(addr);
                leafBits = (uint64_t)leaf[0] | ((uint64_t)leaf[1] << 32);
                // Leaf::get_first_child_mask (base_dag.h:153-189): non-empty bytes.
                uint32_t m = 0;
                for (int b = 0; b < 8; ++b)
                    if ((leafBits >> (8 * b)) & 0xFF)
                        m |= 1u << b;
                cache.childMask = m;
            } else {
                // get_second_child_mask (base_dag.h:190-195).
                const uint32_t word = (c & 4) ? (uint32_t)(leafBits >> 32) : (uint32_t)leafBits;
                cache.childMask = (word >> ((c & 3) * 8)) & 0xFFu;
            }
            cache.visitMask = cache.childMask & isect_mask(false, d.levels, level, path, r);
        }
    }
    float dist[3];
    for (int k = 0; k < 3; ++k) {
        const float bmn = (float)path[k], bmx = bmn + 1.0f;
        const float t1 = (bmn - r.o[k]) * r.inv[k], t2 = (bmx - r.o[k]) * r.inv[k];
        dist[k] = minf(t1, t2);
    }
    const int axis = dist[0] > dist[1] ? (dist[0] > dist[2] ? 0 : 2) : (dist[1] > dist[2] ? 1 : 2);
    const float t = dist[axis];
    // createSurfaceInteraction: the normal points against the ray on the
    // entry axis; the null path (0,0,0) is no surface (path.h:61-64).
    if (surfCode)
        *surfCode = (path[0] | path[1] | path[2]) == 0 ? 0u : 1u + 2u * (uint32_t)axis + (r.d[axis] < 0.0f ? 1u : 0u);
    // origin + t * direction (tracer.cu:388), fused.
    return { path[0], path[1], path[2], material, t, std::fma(t, r.d[0], r.o[0]), std::fma(t, r.d[1], r.o[1]),
             std::fma(t, r.d[2], r.o[2]) };
}

Hit trace_pixel(const Dag& d, const double* cam, const double* rayMin, const double* ddx, const double* ddy, uint32_t px, uint32_t py,
                uint32_t* surfCode = nullptr)
{
    double a[3];
    // rayMin + px*ddx + py*ddy - cam (tracer.cu:404); each add has one product
    // operand, which nvcc's --fmad=true fuses.
    for (int k = 0; k < 3; ++k)
        a[k] = std::fma((double)py, ddy[k], std::fma((double)px, ddx[k], rayMin[k])) - cam[k];
    const double len = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    const double inv = 1 / len;
    Ray r;
    for (int k = 0; k < 3; ++k) {
        r.o[k] = (float)cam[k];
        r.d[k] = (float)(inv * a[k]);
        r.inv[k] = 1.0f / r.d[k];
    }
    r.tmin = 0.0f;
    r.tmax = FLT_MAX;
    return trace_ray(d, r, surfCode);
}

// intersect_ray_node_out_of_order (tracer.cu:274-351): any set voxel along
// the ray; children are visited highest visit bit first.
bool occluded_ray(const Dag& d, const Ray& r)
{
    if (d.root == SENT)
        return false;
    struct Entry {
        uint32_t index, childMask, visitMask;
    };
    std::vector<Entry> stack(d.levels);
    uint32_t path[3] = { 0, 0, 0 };
    Entry cache { d.root, d.get(d.root)[0] & 0xFFu, 0 };
    cache.visitMask = cache.childMask & isect_mask(true, d.levels, 0, path, r);
    uint32_t level = 0;
    uint64_t leafBits = 0;
    for (;;) {
        uint32_t nl = level;
        while (nl > 0 && !cache.visitMask)
            cache = stack[--nl];
        if (nl == 0 && !cache.visitMask)
            return false;
        for (int k = 0; k < 3; ++k)
            path[k] >>= (level - nl);
        level = nl;
        uint32_t c = 7;
        while (!(cache.visitMask & (1u << c)))
            --c; // 31 - clz(visitMask)
        cache.visitMask &= ~(1u << c);
        path[0] = (path[0] << 1) | ((c >> 2) & 1u);
        path[1] = (path[1] << 1) | ((c >> 1) & 1u);
        path[2] = (path[2] << 1) | (c & 1u);
        stack[level] = cache;
        ++level;
        if (level == d.levels)
            return true;
        if (level < d.leafLevel) {
            const uint32_t* n = d.get(cache.index);
            cache.index = n[1 + popc(cache.childMask & ((1u << c) - 1u))];
            cache.childMask = d.get(cache.index)[0] & 0xFFu;
        } else if (level == d.leafLevel) {
            const uint32_t* n = d.get(cache.index);
            const uint32_t* leaf = d.get(n[1 + popc(cache.childMask & ((1u << c) - 1u))]);
            leafBits = (uint64_t)leaf[0] | ((uint64_t)leaf[1] << 32);
            uint32_t m = 0;
            for (int b = 0; b < 8; ++b)
                if ((leafBits >> (8 * b)) & 0xFF)
                    m |= 1u << b;
            cache.childMask = m;
        } else {
            const uint32_t word = (c & 4) ? (uint32_t)(leafBits >> 32) : (uint32_t)leafBits;
            cache.childMask = (word >> ((c & 3) * 8)) & 0xFFu;
        }
        cache.visitMask = cache.childMask & isect_mask(false, d.levels, level, path, r);
    }
}

Ray make_ray(const float* o, const float* dir, float tmax)
{
    Ray r;
    for (int k = 0; k < 3; ++k) {
        r.o[k] = o[k];
        r.d[k] = dir[k];
        r.inv[k] = 1.0f / dir[k];
    }
    r.tmin = 0.0f;
    r.tmax = tmax;
    return r;
}

// PCG32 XSH-RR (O'Neill 2014) with the reference's seeding (utils.h:45-82).
struct Pcg32 {
    uint64_t state, inc;
    uint32_t next()
    {
        const uint64_t old = state;
        state = old * 6364136223846793005ull + (inc | 1ull);
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
};
Pcg32 pcg_seeded(uint64_t stream)
{
    Pcg32 g { 0x853c49e6748fea9bull, stream };
    g.next();
    g.next();
    return g;
}
float pcg_float(Pcg32& g) { return (float)g.next() / 4294967296.0f; }

// sin/cos of 2*pi*u: nearest quarter turn plus Taylor series (degree 9 / 10)
// of the remainder; the same float operations as the GPU sampler.
void sincos_turns(float u, float& sn, float& cs)
{
    const float q = 4.0f * u;
    const float k = std::floor(q + 0.5f);
    const float x = (q - k) * 1.57079632679489662f;
    const float x2 = x * x;
    const float sp = x * (1.0f + x2 * (-0.166666667f + x2 * (8.33333333e-3f + x2 * (-1.98412698e-4f + x2 * 2.75573192e-6f))));
    const float cp = 1.0f + x2 * (-0.5f + x2 * (4.16666667e-2f + x2 * (-1.38888889e-3f + x2 * (2.48015873e-5f + x2 * -2.75573192e-7f))));
    switch ((int)k & 3) {
    case 0: sn = sp; cs = cp; break;
    case 1: sn = cp; cs = -sp; break;
    case 2: sn = -sp; cs = -cp; break;
    default: sn = -cp; cs = sp; break;
    }
}

// Normal / dpdu / dpdv of a surface code (createSurfaceInteraction).
void surface_frame(uint32_t code, float* n, float* du, float* dv)
{
    const uint32_t axis = (code - 1) >> 1;
    for (int k = 0; k < 3; ++k)
        n[k] = du[k] = dv[k] = 0.0f;
    n[axis] = ((code - 1) & 1) ? 1.0f : -1.0f;
    if (axis == 0) { du[2] = 1.0f; dv[1] = 1.0f; }
    else if (axis == 1) { du[0] = 1.0f; dv[2] = 1.0f; }
    else { du[0] = 1.0f; dv[1] = 1.0f; }
}

// trace_ambient_occlusion's per-pixel body (tracer.cu:661-691).
float ao_at(const Dag& d, const float* p, uint32_t code, uint32_t samples, float bias, float len, uint64_t stream)
{
    float n[3], du[3], dv[3];
    surface_frame(code, n, du, dv);
    float o[3];
    for (int k = 0; k < 3; ++k)
        o[k] = std::fma(n[k], bias, p[k]); // position + normal * shadowBias, fused (tracer.cu:675)
    Pcg32 g = pcg_seeded(stream);
    uint32_t open = 0;
    for (uint32_t s = 0; s < samples; ++s) {
        const float ux = pcg_float(g), uy = pcg_float(g);
        const float rr = std::sqrt(ux);
        float sn, cs;
        sincos_turns(uy, sn, cs);
        const float zz = 1.0f - ux;
        const float local[3] = { rr * cs, rr * sn, std::sqrt(0.0f < zz ? zz : 0.0f) };
        float dir[3];
        for (int k = 0; k < 3; ++k)
            dir[k] = (local[0] * du[k] + local[1] * dv[k]) + local[2] * n[k];
        if (!occluded_ray(d, make_ray(o, dir, len)))
            ++open;
    }
    return (float)open / (float)samples;
}

// sun_direction() (tracer.cu:600-603) = normalize(float3(0.3f, 1, 0.5f)):
// (1 / sqrt(dot(v, v))) * v in float, rounded per operation (cuda_math.h:115-118;
// constant operands are folded without fusion).
void sun_dir(float* L)
{
    const float v[3] = { 0.3f, 1.0f, 0.5f };
    const float dot = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    const float inv = 1.0f / std::sqrt(dot);
    for (int k = 0; k < 3; ++k)
        L[k] = inv * v[k];
}

} // namespace

extern "C" {

void* ora_dag_create(uint32_t levels, uint32_t M)
{
    if (levels < 4 || levels > 20 || (M != 0 && M != 4 && M != 8))
        return nullptr;
    return new Dag(levels, M);
}
void ora_dag_destroy(void* d) { delete static_cast<Dag*>(d); }
uint32_t ora_dag_root(void* d) { return static_cast<Dag*>(d)->root; }
void ora_dag_set_root(void* d, uint32_t r) { static_cast<Dag*>(d)->root = r; }
uint64_t ora_dag_item_count(void* d) { return static_cast<Dag*>(d)->items.size(); }

void ora_terrain_heights(uint64_t seed, double base, double amp, double freq, uint32_t octaves, uint32_t x0, uint32_t z0,
                         uint32_t w, uint32_t dd, int32_t* out)
{
    const Perlin pn(seed);
    const TerrainSpec t { seed, base, amp, freq, octaves, 0 };
    for (uint32_t z = 0; z < dd; ++z)
        for (uint32_t x = 0; x < w; ++x)
            out[(size_t)z * w + x] = height_at(pn, t, x0 + x, z0 + z);
}

// band_log2 == 0: one material; else the band rule of svdag_build_terrain_banded.
int ora_dag_build_terrain(void* dp, uint64_t seed, double base, double amp, double freq, uint32_t octaves, uint32_t material,
                          uint32_t band_log2, uint32_t band_seed)
{
    Dag& d = *static_cast<Dag*>(dp);
    const Perlin pn(seed);
    const TerrainSpec spec { seed, base, amp, freq, octaves, material };
    Terrain t;
    t.side = 1u << d.levels;
    t.h.resize((size_t)t.side * t.side);
    for (uint32_t z = 0; z < t.side; ++z)
        for (uint32_t x = 0; x < t.side; ++x)
            t.h[(size_t)z * t.side + x] = height_at(pn, spec, x, z);
    const Bands bands { band_log2 != 0, band_log2, band_seed, material, d.numMat - 1u };
    d.root = build_terrain(d, t, bands, 0, 0, 0, 0).handle;
    return 0;
}

// Returns 1 if merged, 0 if skipped (no leaf).
int ora_dag_edit(void* dp, uint32_t kind, uint32_t material, const float* center, float radius, const float* src,
                 const float* dst, uint32_t rotation, uint32_t* svoCounts)
{
    Dag& d = *static_cast<Dag*>(dp);
    const float zero[3] = { 0, 0, 0 };
    const Editor e = make_editor(kind, material, center, radius, src ? src : zero, dst ? dst : zero, rotation);
    return apply_edit(d, e, svoCounts);
}

void ora_dag_query(void* dp, const uint32_t* xyz, uint32_t n, uint32_t* out)
{
    const Dag& d = *static_cast<Dag*>(dp);
    for (uint32_t i = 0; i < n; ++i)
        out[i] = get_value(d, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

static size_t emit(const std::vector<uint32_t>& dump, uint32_t* buf, size_t* len)
{
    if (buf && len && *len >= dump.size())
        std::memcpy(buf, dump.data(), dump.size() * 4);
    if (len)
        *len = dump.size();
    return dump.size();
}

uint64_t ora_dag_canonical(void* dp, uint32_t* buf, size_t* len, uint32_t* counts)
{
    const Dag& d = *static_cast<Dag*>(dp);
    std::vector<uint32_t> cnt;
    const auto dump = canonical(d.leafLevel, d.root, [&](uint32_t, uint32_t h) { return &d.items[h]; }, cnt);
    for (size_t i = 0; i < cnt.size() && counts; ++i)
        counts[i] = cnt[i];
    emit(dump, buf, len);
    return fnv(dump);
}

// Canonical form of an exported record list [levels, M, root, n, (level, handle, size, words)...].
uint64_t ora_canonical_from_records(const uint32_t* rec, size_t nwords, uint32_t* buf, size_t* len, uint32_t* counts)
{
    const uint32_t levels = rec[0], root = rec[2], n = rec[3];
    std::vector<std::unordered_map<uint32_t, std::vector<uint32_t>>> byLevel(levels);
    size_t p = 4;
    for (uint32_t i = 0; i < n && p + 3 <= nwords; ++i) {
        const uint32_t L = rec[p], h = rec[p + 1], s = rec[p + 2];
        byLevel[L][h] = std::vector<uint32_t>(rec + p + 3, rec + p + 3 + s);
        p += 3 + s;
    }
    std::vector<uint32_t> cnt;
    const auto dump = canonical(levels - 2, root, [&](uint32_t L, uint32_t h) { return &byLevel[L].at(h); }, cnt);
    for (size_t i = 0; i < cnt.size() && counts; ++i)
        counts[i] = cnt[i];
    emit(dump, buf, len);
    return fnv(dump);
}

// Reachable item count from root plus the level-0 fully-filled nodes (GC roots).
uint64_t ora_dag_live_items(void* dp)
{
    const Dag& d = *static_cast<Dag*>(dp);
    std::unordered_set<uint32_t> seen;
    std::vector<std::pair<uint32_t, uint32_t>> stack;
    if (d.root != SENT)
        stack.push_back({ 0, d.root });
    for (uint32_t m = 0; m < d.numMat; ++m)
        stack.push_back({ 0, d.ff[m] });
    std::unordered_set<uint64_t> visited;
    while (!stack.empty()) {
        auto [L, h] = stack.back();
        stack.pop_back();
        if (!visited.insert(((uint64_t)L << 32) | h).second)
            continue;
        seen.insert(h);
        if (L < d.leafLevel) {
            const auto& w = d.items[h];
            for (size_t j = 1; j < w.size(); ++j)
                stack.push_back({ L + 1, w[j] });
        }
    }
    return seen.size();
}

// get_trace_params (dag_tracer.cu:59-101) + trace_paths for every pixel.
void ora_dag_trace(void* dp, const double* position, const double* rotation, double fovDeg, const double* bmin,
                   const double* bmax, uint32_t W, uint32_t H, uint32_t* out)
{
    const Dag& d = *static_cast<Dag*>(dp);
    const double right[3] = { -rotation[0], -rotation[1], -rotation[2] };
    const double up[3] = { rotation[3], rotation[4], rotation[5] };
    const double fwd[3] = { rotation[6], rotation[7], rotation[8] };
    const double fov = fovDeg / 2.0 * (double(M_PI) / 180.);
    const double aspect = double(W) / double(H);
    double cam[3], rayMin[3], ddx[3], ddy[3];
    for (int k = 0; k < 3; ++k) {
        const double X = right[k] * std::sin(fov) * aspect, Y = up[k] * std::sin(fov), Z = fwd[k] * std::cos(fov);
        const double bl = position[k] + Z - Y - X, br = position[k] + Z - Y + X, tl = position[k] + Z + Y - X;
        const double tr = -bmin[k], sc = double(1 << d.levels) / (bmax[k] - bmin[k]);
        cam[k] = (position[k] + tr) * sc;
        rayMin[k] = (bl + tr) * sc;
        ddx[k] = ((br + tr) * sc - rayMin[k]) * (1.0 / W);
        ddy[k] = ((tl + tr) * sc - rayMin[k]) * (1.0 / H);
    }
    for (uint32_t py = 0; py < H; ++py)
        for (uint32_t px = 0; px < W; ++px) {
            const Hit h = trace_pixel(d, cam, rayMin, ddx, ddy, px, py);
            std::memcpy(out + ((size_t)(H - 1 - py) * W + px) * 8, &h, sizeof(h));
        }
}

// Primary trace + trace_shadows + trace_ambient_occlusion for every pixel;
// outputs in the stored (row H-1-y) layout, 0 where there is no surface.
void ora_dag_shade(void* dp, const double* position, const double* rotation, double fovDeg, const double* bmin,
                   const double* bmax, uint32_t W, uint32_t H, float bias, uint32_t aoSamples, float aoLength,
                   uint64_t seed, float* outSun, float* outAo)
{
    const Dag& d = *static_cast<Dag*>(dp);
    const double right[3] = { -rotation[0], -rotation[1], -rotation[2] };
    const double up[3] = { rotation[3], rotation[4], rotation[5] };
    const double fwd[3] = { rotation[6], rotation[7], rotation[8] };
    const double fov = fovDeg / 2.0 * (double(M_PI) / 180.);
    const double aspect = double(W) / double(H);
    double cam[3], rayMin[3], ddx[3], ddy[3];
    for (int k = 0; k < 3; ++k) {
        const double X = right[k] * std::sin(fov) * aspect, Y = up[k] * std::sin(fov), Z = fwd[k] * std::cos(fov);
        const double bl = position[k] + Z - Y - X, br = position[k] + Z - Y + X, tl = position[k] + Z + Y - X;
        const double tr = -bmin[k], sc = double(1 << d.levels) / (bmax[k] - bmin[k]);
        cam[k] = (position[k] + tr) * sc;
        rayMin[k] = (bl + tr) * sc;
        ddx[k] = ((br + tr) * sc - rayMin[k]) * (1.0 / W);
        ddy[k] = ((tl + tr) * sc - rayMin[k]) * (1.0 / H);
    }
    float L[3];
    sun_dir(L);
    for (uint32_t py = 0; py < H; ++py)
        for (uint32_t px = 0; px < W; ++px) {
            uint32_t code = 0;
            const Hit h = trace_pixel(d, cam, rayMin, ddx, ddy, px, py, &code);
            const size_t i = (size_t)(H - 1 - py) * W + px;
            if (!code) {
                outSun[i] = 0.0f;
                outAo[i] = 0.0f;
                continue;
            }
            const float p[3] = { h.px, h.py, h.pz };
            float n[3], du[3], dv[3];
            surface_frame(code, n, du, dv);
            float o[3];
            for (int k = 0; k < 3; ++k)
                o[k] = std::fma(bias, L[k], p[k]); // position + shadowBias * L, fused (tracer.cu:621)
            float ndl = (n[0] * L[0] + n[1] * L[1]) + n[2] * L[2];
            ndl = ndl < 0.0f ? 0.0f : ndl;
            outSun[i] = occluded_ray(d, make_ray(o, L, FLT_MAX)) ? 0.0f : ndl;
            outAo[i] = aoSamples ? ao_at(d, p, code, aoSamples, bias, aoLength, (uint64_t)i * (seed > 1 ? seed : 1))
                                 : 0.0f;
        }
}

// Tracer::trace_path_tracing (tracer.cu:758-842) for every stored pixel i =
// y * W + x, with the same camera as the primary trace.  Colours are one per
// material (palette, 3 floats each; a material without one gets the checker
// (1, b, 1 - b)), made linear with accurateSRGBToLinear.  At depth 0 without
// pixel integration the surface is the stored primary surface of pixel i and
// its colour the colour pass's RGBA8 (colors); otherwise the camera ray of
// sample (x, y) (+ a jitter from the pixel's PCG32 stream when integrating)
// through row H - 1 - y.  Each bounce: sky -> Lo += c * sky * env, stop; else
// a shadow ray from position + 0.1 * normal toward the sun adds
// c * brdf(L) * sunColor * env when it reaches nothing, and the path goes on
// along a uniform hemisphere sample (z = u.x, phi = 2 pi u.y, pdf 1 / (2 pi)):
// c *= brdf(dir) / pdf, brdf(d) = max(0, N.d) / pi * Kd.  accum (3 floats per
// pixel) holds the sums of earlier calls and receives the new ones; the
// image is min(accum / (samples + accumulated), 1) made sRGB, as RGBA8.
// Products whose sum nvcc fuses (one product per add) are fused here.
extern "C" float ora_srgb_to_linear(float c);
extern "C" float ora_linear_to_srgb(float c);
extern "C" uint32_t ora_rgb888(const float* c);

void ora_dag_path_trace(void* dp, const double* position, const double* rotation, double fovDeg, const double* bmin,
                        const double* bmax, uint32_t W, uint32_t H, const uint32_t* colors, const float* palette,
                        uint32_t numColors, float env, uint32_t numSamples, uint32_t maxDepth, uint32_t integrate,
                        uint64_t seed, float* accum, uint32_t numAccumulated, uint32_t* out)
{
    const Dag& d = *static_cast<Dag*>(dp);
    const double right[3] = { -rotation[0], -rotation[1], -rotation[2] };
    const double up[3] = { rotation[3], rotation[4], rotation[5] };
    const double fwd[3] = { rotation[6], rotation[7], rotation[8] };
    const double fov = fovDeg / 2.0 * (double(M_PI) / 180.);
    const double aspect = double(W) / double(H);
    double cam[3], rayMin[3], ddx[3], ddy[3];
    for (int k = 0; k < 3; ++k) {
        const double X = right[k] * std::sin(fov) * aspect, Y = up[k] * std::sin(fov), Z = fwd[k] * std::cos(fov);
        const double bl = position[k] + Z - Y - X, br = position[k] + Z - Y + X, tl = position[k] + Z + Y - X;
        const double tr = -bmin[k], sc = double(1 << d.levels) / (bmax[k] - bmin[k]);
        cam[k] = (position[k] + tr) * sc;
        rayMin[k] = (bl + tr) * sc;
        ddx[k] = ((br + tr) * sc - rayMin[k]) * (1.0 / W);
        ddy[k] = ((tl + tr) * sc - rayMin[k]) * (1.0 / H);
    }
    float L[3];
    sun_dir(L);
    const float sky[3] = { 187.0f / 255.0f, 242.0f / 255.0f, 250.0f / 255.0f };
    const float sunColor[3] = { 0.95f, 0.94f, 0.47f };
    const float invPi = std::numbers::inv_pi_v<float>;
    const float pdf = 0.5f * invPi;
    const float eps = 0.1f;
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            Pcg32 g = pcg_seeded((uint64_t)(uint32_t)i * (seed > 1 ? seed : 1));
            float Lo[3] = { accum[3 * i], accum[3 * i + 1], accum[3 * i + 2] };
            for (uint32_t s = 0; s < numSamples; ++s) {
                float sx = (float)x, sy = (float)y;
                if (integrate) {
                    sx += pcg_float(g);
                    sy += pcg_float(g);
                }
                const double row = (double)((float)(H - 1) - sy);
                double a[3];
                for (int k = 0; k < 3; ++k)
                    a[k] = std::fma(row, ddy[k], std::fma((double)sx, ddx[k], rayMin[k])) - cam[k];
                const double len = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
                const double inv = 1 / len;
                float o[3], dir[3];
                for (int k = 0; k < 3; ++k) {
                    o[k] = (float)cam[k];
                    dir[k] = (float)(inv * a[k]);
                }
                float c[3] = { 1.0f, 1.0f, 1.0f };
                for (uint32_t depth = 0; depth < maxDepth; ++depth) {
                    Hit h;
                    uint32_t code = 0;
                    float kd[3];
                    if (depth == 0 && !integrate) {
                        // The stored primary surface of pixel i (camera row H - 1 - y).
                        h = trace_pixel(d, cam, rayMin, ddx, ddy, x, H - 1 - y, &code);
                        for (int k = 0; k < 3; ++k)
                            kd[k] = ora_srgb_to_linear((float)((colors[i] >> (8 * k)) & 0xFFu) / 255.0f);
                    } else {
                        h = trace_ray(d, make_ray(o, dir, FLT_MAX), &code);
                        if (code) {
                            float col[3];
                            if (h.material >= numColors) {
                                const float b = (float)((h.x ^ h.y ^ h.z) & 1u);
                                col[0] = 1.0f;
                                col[1] = b;
                                col[2] = 1.0f - b;
                            } else {
                                for (int k = 0; k < 3; ++k)
                                    col[k] = palette[3 * h.material + k];
                            }
                            for (int k = 0; k < 3; ++k)
                                kd[k] = ora_srgb_to_linear(col[k]);
                        }
                    }
                    if (!code) {
                        for (int k = 0; k < 3; ++k)
                            Lo[k] = std::fma(c[k] * sky[k], env, Lo[k]);
                        break;
                    }
                    float n[3], du[3], dv[3];
                    surface_frame(code, n, du, dv);
                    const float p[3] = { h.px, h.py, h.pz };
                    float so[3];
                    for (int k = 0; k < 3; ++k)
                        so[k] = std::fma(eps, n[k], p[k]);
                    if (!occluded_ray(d, make_ray(so, L, FLT_MAX))) {
                        float ndl = (n[0] * L[0] + n[1] * L[1]) + n[2] * L[2];
                        ndl = 0.0f > ndl ? 0.0f : ndl;
                        for (int k = 0; k < 3; ++k)
                            Lo[k] = std::fma((c[k] * ((ndl * invPi) * kd[k])) * sunColor[k], env, Lo[k]);
                    }
                    const float ux = pcg_float(g), uy = pcg_float(g);
                    const float rr = std::sqrt(std::max(0.0f, std::fma(-ux, ux, 1.0f)));
                    float sn, cs;
                    sincos_turns(uy, sn, cs);
                    const float local[3] = { rr * cs, rr * sn, ux };
                    for (int k = 0; k < 3; ++k) {
                        dir[k] = (local[0] * du[k] + local[1] * dv[k]) + local[2] * n[k];
                        o[k] = so[k];
                    }
                    float ndd = (n[0] * dir[0] + n[1] * dir[1]) + n[2] * dir[2];
                    ndd = 0.0f > ndd ? 0.0f : ndd;
                    for (int k = 0; k < 3; ++k)
                        c[k] = c[k] * (((ndd * invPi) * kd[k]) / pdf);
                }
            }
            float v[3];
            for (int k = 0; k < 3; ++k) {
                accum[3 * i + k] = Lo[k];
                const float m = Lo[k] / (float)(numSamples + numAccumulated);
                v[k] = ora_linear_to_srgb(m < 1.0f ? m : 1.0f);
            }
            out[i] = ora_rgb888(v);
        }
}

// Any-hit query of one ray (for the brute-force checks).
int ora_dag_occluded(void* dp, const float* origin, const float* dir, float tmax)
{
    return occluded_ray(*static_cast<Dag*>(dp), make_ray(origin, dir, tmax)) ? 1 : 0;
}

// Occlusion at an explicit surface point with a surface code.
float ora_dag_ao_at(void* dp, const float* p, uint32_t code, uint32_t samples, float bias, float len, uint64_t stream)
{
    return ao_at(*static_cast<Dag*>(dp), p, code, samples, bias, len, stream);
}

// PCG32 outputs from an explicit (state, inc), and from the reference seeding.
void ora_pcg32_raw(uint64_t state, uint64_t inc, uint32_t n, uint32_t* out)
{
    Pcg32 g { state, inc };
    for (uint32_t i = 0; i < n; ++i)
        out[i] = g.next();
}
void ora_pcg32_seeded(uint64_t stream, uint32_t n, uint32_t* out)
{
    Pcg32 g = pcg_seeded(stream);
    for (uint32_t i = 0; i < n; ++i)
        out[i] = g.next();
}
void ora_sincos_turns(float u, float* sn, float* cs) { sincos_turns(u, *sn, *cs); }
void ora_sun_dir(float* L) { sun_dir(L); }

// Encoding helpers exposed for the reference's own encoding tests.
uint32_t ora_leaf_build(uint32_t M, const uint32_t* present, const uint32_t* materials, uint32_t* out)
{
    LeafBuilder b(M);
    for (uint32_t v = 0; v < 64; ++v) {
        if (present[v])
            b.set(materials[v]);
        b.next();
    }
    const auto w = b.finish();
    std::memcpy(out, w.data(), w.size() * 4);
    return (uint32_t)w.size();
}
int ora_leaf_material(uint32_t M, const uint32_t* leaf, uint32_t voxel, uint32_t* mat)
{
    const Enc e { M };
    return e.material(leaf, voxel, *mat) ? 1 : 0;
}
uint32_t ora_leaf_size(uint32_t M, const uint32_t* leaf) { return Enc { M }.leaf_size(leaf); }

} // extern "C"

extern "C" {

// Node header encode/decode (my_gpu_base_dag.h:152-189).
uint32_t ora_header_make(uint32_t M, uint32_t mask, uint32_t payload) { return Enc { M }.header_payload(mask, payload); }
void ora_header_decode(uint32_t M, uint32_t header, uint32_t* mask, uint32_t* payload, uint32_t* full, uint32_t* mat)
{
    const Enc e { M };
    *mask = header & 0xFFu;
    *payload = e.payload(header);
    *full = e.is_full(header, *mat) ? 1u : 0u;
}

// MyGpuSphereEditorBase::test_node_overlap for a box (my_gpu_dag_editors.h:420-440).
int ora_sphere_overlap(const float* center, float radius, const float* mn, const float* mx)
{
    const float zero[3] = { 0, 0, 0 };
    const Editor e = make_editor(SPHERE_ADD, 0, center, radius, zero, zero);
    return sphere_overlap(e, mn, mx) ? 1 : 0;
}

// Factory rule (my_gpu_hash_dag_factory.cpp:317-424) over a dense voxel grid:
// vox[(x * side + y) * side + z] = material, or 0xFF for empty.
uint32_t ora_dag_build_dense(void* dp, const uint8_t* vox)
{
    Dag& d = *static_cast<Dag*>(dp);
    const uint32_t side = 1u << d.levels;
    std::function<Built(uint32_t, uint32_t, uint32_t, uint32_t)> rec = [&](uint32_t level, uint32_t px, uint32_t py,
                                                                          uint32_t pz) -> Built {
        if (level == d.leafLevel) {
            LeafBuilder b(d.M);
            for (uint32_t v = 0; v < 64; ++v) {
                const uint32_t x = (px << 2) | (((v >> 5) & 1u) << 1) | ((v >> 2) & 1u);
                const uint32_t y = (py << 2) | (((v >> 4) & 1u) << 1) | ((v >> 1) & 1u);
                const uint32_t z = (pz << 2) | (((v >> 3) & 1u) << 1) | (v & 1u);
                const uint8_t m = vox[((size_t)x * side + y) * side + z];
                if (m != 0xFF)
                    b.set(m);
                b.next();
            }
            const auto w = b.finish();
            if (w[0] == 0 && w[1] == 0)
                return { SENT, false, 0 };
            uint32_t fm = 0;
            const bool full = d.leaf_full(w.data(), fm);
            return { d.add(w), full, fm };
        }
        std::vector<uint32_t> node { 0u };
        uint32_t mask = 0, fm = SENT;
        bool allFull = true;
        for (uint32_t c = 0; c < 8; ++c) {
            const Built b = rec(level + 1, (px << 1) | ((c >> 2) & 1u), (py << 1) | ((c >> 1) & 1u), (pz << 1) | (c & 1u));
            if (b.handle == SENT) {
                allFull = false;
                continue;
            }
            mask |= 1u << c;
            node.push_back(b.handle);
            if (!b.full || (fm != SENT && b.mat != fm))
                allFull = false;
            fm = b.full ? b.mat : fm;
        }
        if (!mask)
            return { SENT, false, 0 };
        const bool full = allFull && mask == 0xFF;
        node[0] = full ? d.enc.header_full(mask, fm) : d.enc.header(mask);
        return { d.add(node), full, full ? fm : 0u };
    };
    d.root = rec(0, 0, 0, 0).handle;
    return d.root;
}

} // extern "C"
