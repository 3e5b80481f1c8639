// Edit tools evaluated on the device: node classification (should_edit) and
// the per-voxel new value (get_new_value).
//
// The float arithmetic is the reference's, operation by operation.  The build
// contracts nothing by itself; the one site where nvcc's default
// (--fmad=true) has a single possible fusion, `dist -= squared(d)` in the
// sphere overlap test (my_gpu_dag_editors.h:425-437), is written as
// fmaf(-d, d, dist).  Everything else is rounded operation by operation:
//   sphere  my_gpu_dag_editors.h:401-496   (centre + 0.5, radius >= 0.1,
//           overlap = r^2 - sum of squared axis gaps > 0, Fill/Empty when the
//           box shrunk by one voxel is inside)
//   box     my_gpu_dag_editors.h:45-119    ([c-r, c+r), r == 0 -> [c, c+1))
//   paint   my_gpu_dag_editors.h:498-536   (Fill only over flagged full nodes)
//   copy    my_gpu_dag_editors.h:127-253   (translation; up to 8 source parents;
//           quarter turns about +Y have no reference counterpart; done in
//           integer voxel arithmetic: decide_copy_rot / rot_source)
// The terrain tool (kind 6) has no reference counterpart: it classifies nodes
// against a min/max height pyramid so that a heightfield builds through the
// same pipeline as an edit.
#pragma once
#include "svdag_device.h"

namespace svdag {

enum EditKind : uint32_t {
    EK_SPHERE_ADD = 0,
    EK_SPHERE_REMOVE = 1,
    EK_BOX_ADD = 2,
    EK_BOX_REMOVE = 3,
    EK_PAINT = 4,
    EK_COPY = 5,
    EK_TERRAIN = 6,
    EK_DENSE = 7
};
// Dense-grid pyramid codes besides the materials 0..255.
constexpr uint16_t kDenseEmpty = 0x100;
constexpr uint16_t kDenseMixed = 0x101;
enum Op : uint32_t { OP_SUBDIVIDE = 0, OP_FILL = 1, OP_EMPTY = 2, OP_KEEP = 3 };
enum NState : uint32_t { NS_PARTIAL = 0, NS_FILLED = 1, NS_EMPTY = 2 };

constexpr uint32_t kMaxLevels = 24;

struct EditorDev {
    uint32_t kind;
    uint32_t material;
    // sphere / paint
    float cx, cy, cz;
    float radius, r2;
    // box and copy destination bounds
    float bmin[3], bmax[3];
    // copy
    float d2s[3];
    int32_t id2s[3];
    // rotated copy (rot != 0, no reference counterpart): quarter turns about
    // +Y, destination pivot pv, source pivot qv = pv + id2s, and the
    // destination voxels [ivmin, ivmax] whose unit boxes overlap the box
    uint32_t rot;
    int32_t pv[3], qv[3], ivmin[3], ivmax[3];
    // terrain: height pyramid, level k holds (2^D >> k)^2 cells, row-major in z.
    // hmax holds levels 1..D only (level 0's maxima are hmin's heights) and is
    // indexed by pyrOffset[k] - pyrOffset[1].
    uint32_t levels;
    const int32_t* hmin;
    const int32_t* hmax;
    uint64_t pyrOffset[kMaxLevels];
    // terrain height bands: cubes of 2^bandLog2 voxels hold one material
    // (band_material); banded == 0 gives every voxel `material`
    uint32_t banded, bandLog2, bandSeed, materialMask;
    // terrain tile [tileMin, tileMax) in x and z: nodes outside keep their
    // content, nodes straddling the edge subdivide (tiled == 0: no tile)
    uint32_t tiled;
    uint32_t tileMin[2], tileMax[2];
    // dense grid of side 2^denseLog2 at the origin: pyramid level k holds
    // (side >> k)^3 codes indexed (x * s + y) * s + z (pyrOffset as above)
    const uint16_t* dense;
    uint32_t denseLog2;
};

// One copy-editor source parent: path and handle (handle == kSentinel ends the list).
struct CopyParent {
    uint32_t x, y, z, h;
};
struct CopyState {
    CopyParent p[8];
};

__host__ __device__ inline float fmax_ref(float a, float b) { return a > b ? a : b; }
__host__ __device__ inline float fmin_ref(float a, float b) { return a < b ? a : b; }

__host__ __device__ inline bool sphere_overlap(const EditorDev& e, const float mn[3], const float mx[3])
{
    float dist = e.r2;
    const float c[3] = { e.cx, e.cy, e.cz };
    for (int a = 0; a < 3; ++a) {
        if (c[a] < mn[a]) {
            const float d = c[a] - mn[a];
            dist = fmaf(-d, d, dist);
        } else if (c[a] > mx[a]) {
            const float d = c[a] - mx[a];
            dist = fmaf(-d, d, dist);
        }
    }
    return dist > 0.0f;
}

__host__ __device__ inline bool sphere_contained(const EditorDev& e, const float mn[3], const float mx[3])
{
    const float c[3] = { e.cx, e.cy, e.cz };
    float m[3];
    for (int a = 0; a < 3; ++a) {
        const float lo = mn[a] - c[a];
        const float hi = mx[a] - c[a];
        const float l2 = lo * lo, h2 = hi * hi;
        m[a] = l2 > h2 ? l2 : h2; // std::max(l2, h2)
    }
    return (m[0] + m[1]) + m[2] < e.r2;
}

__host__ __device__ inline bool box_overlap(const float bmin[3], const float bmax[3], const float mn[3], const float mx[3])
{
    return !(bmin[0] >= mx[0] || bmin[1] >= mx[1] || bmin[2] >= mx[2] || bmax[0] <= mn[0] || bmax[1] <= mn[1]
             || bmax[2] <= mn[2] || mn[0] >= bmax[0] || mn[1] >= bmax[1] || mn[2] >= bmax[2] || mx[0] <= bmin[0]
             || mx[1] <= bmin[1] || mx[2] <= bmin[2]);
}

__host__ __device__ inline bool box_contained(const float bmin[3], const float bmax[3], const float mn[3], const float mx[3])
{
    return bmin[0] <= mn[0] && bmin[1] <= mn[1] && bmin[2] <= mn[2] && bmax[0] >= mx[0] && bmax[1] >= mx[1]
        && bmax[2] >= mx[2];
}

// Node box [p << depth, (p << depth) + 2^depth) in float, as Path::as_position.
__host__ __device__ inline void node_box(uint32_t x, uint32_t y, uint32_t z, uint32_t depth, float mn[3], float mx[3])
{
    mn[0] = (float)(x << depth);
    mn[1] = (float)(y << depth);
    mn[2] = (float)(z << depth);
    const float size = (float)(1u << depth);
    mx[0] = mn[0] + size;
    mx[1] = mn[1] + size;
    mx[2] = mn[2] + size;
}

// Terrain material of voxel (x, y, z): with bands, the band index y >> b plus
// a seeded 2-bit offset per 2^b column tile, over the material range.
// No reference counterpart (a synthetic scene's material rule).
__host__ __device__ inline uint32_t band_material(const EditorDev& e, uint32_t x, uint32_t y, uint32_t z)
{
    if (!e.banded)
        return e.material;
    const uint32_t b = e.bandLog2;
    uint32_t h = e.bandSeed ^ ((x >> b) * 0x9E3779B1u) ^ ((z >> b) * 0x85EBCA77u);
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    return (e.material + (y >> b) + (h & 3u)) & e.materialMask;
}

// should_edit for the stateless tools (and the copy tool's voxel level).
__device__ inline Op decide(const EditorDev& e, uint32_t x, uint32_t y, uint32_t z, uint32_t depth, NState st,
                            uint32_t& fillMat)
{
    float mn[3], mx[3];
    node_box(x, y, z, depth, mn, mx);
    fillMat = e.material;
    switch (e.kind) {
    case EK_SPHERE_ADD:
    case EK_SPHERE_REMOVE: {
        if (!sphere_overlap(e, mn, mx))
            return OP_KEEP;
        const float smn[3] = { mn[0] + 1.0f, mn[1] + 1.0f, mn[2] + 1.0f };
        const float smx[3] = { mx[0] - 1.0f, mx[1] - 1.0f, mx[2] - 1.0f };
        if (sphere_contained(e, smn, smx))
            return e.kind == EK_SPHERE_ADD ? OP_FILL : OP_EMPTY;
        return OP_SUBDIVIDE;
    }
    case EK_BOX_ADD:
    case EK_BOX_REMOVE:
        if (box_contained(e.bmin, e.bmax, mn, mx))
            return e.kind == EK_BOX_ADD ? OP_FILL : OP_EMPTY;
        if (box_overlap(e.bmin, e.bmax, mn, mx))
            return OP_SUBDIVIDE;
        return OP_KEEP;
    case EK_PAINT: {
        const bool overlaps = sphere_overlap(e, mn, mx);
        const bool contained = sphere_contained(e, mn, mx);
        if (!overlaps)
            return OP_KEEP;
        if (contained && st == NS_FILLED)
            return OP_FILL;
        if (st == NS_EMPTY)
            return OP_EMPTY;
        return OP_SUBDIVIDE;
    }
    case EK_COPY:
        return box_overlap(e.bmin, e.bmax, mn, mx) ? OP_SUBDIVIDE : OP_KEEP;
    case EK_DENSE: {
        // A cell of one material is filled with it, an empty cell is empty,
        // a mixed cell is subdivided.  Nodes larger than the grid hold it at
        // their low corner (the origin node) or nothing.
        if (depth > e.denseLog2) {
            if ((x | y | z) != 0u)
                return OP_EMPTY;
            return e.dense[e.pyrOffset[e.denseLog2]] == kDenseEmpty ? OP_EMPTY : OP_SUBDIVIDE;
        }
        const uint64_t s = 1ull << (e.denseLog2 - depth);
        if (x >= s || y >= s || z >= s)
            return OP_EMPTY;
        const uint32_t c = e.dense[e.pyrOffset[depth] + ((uint64_t)x * s + y) * s + z];
        if (c == kDenseEmpty)
            return OP_EMPTY;
        if (c == kDenseMixed)
            return OP_SUBDIVIDE;
        fillMat = c;
        return OP_FILL;
    }
    case EK_TERRAIN: {
        if (e.tiled) {
            const uint64_t x0 = (uint64_t)x << depth, z0 = (uint64_t)z << depth, w = (uint64_t)1 << depth;
            if (x0 + w <= e.tileMin[0] || x0 >= e.tileMax[0] || z0 + w <= e.tileMin[1] || z0 >= e.tileMax[1])
                return OP_KEEP;
            if (x0 < e.tileMin[0] || x0 + w > e.tileMax[0] || z0 < e.tileMin[1] || z0 + w > e.tileMax[1])
                return OP_SUBDIVIDE;
        }
        const uint64_t width = (uint64_t)1 << (e.levels - depth);
        const uint64_t idx = e.pyrOffset[depth] + (uint64_t)z * width + x;
        const int64_t lo = e.hmin[idx];
        const int64_t hi = depth == 0 ? lo : e.hmax[idx - e.pyrOffset[1]];
        const int64_t y0 = (int64_t)y << depth;
        const int64_t y1 = y0 + ((int64_t)1 << depth) - 1;
        if (y1 <= lo) {
            // A solid node wider than a band cube spans several materials.
            if (e.banded && depth > e.bandLog2)
                return OP_SUBDIVIDE;
            fillMat = band_material(e, x << depth, y << depth, z << depth);
            return OP_FILL;
        }
        if (y0 > hi)
            return OP_EMPTY;
        return OP_SUBDIVIDE;
    }
    default:
        return OP_KEEP;
    }
}

// get_new_value for the stateless tools: true and a material when the voxel
// ends up set.
__device__ inline bool new_value(const EditorDev& e, bool hasOld, uint32_t oldMat, uint32_t& mat)
{
    switch (e.kind) {
    case EK_SPHERE_ADD:
    case EK_BOX_ADD:
    case EK_PAINT:
    case EK_TERRAIN:
        mat = e.material;
        return true;
    default:
        (void)hasOld;
        (void)oldMat;
        return false;
    }
}

__host__ __device__ inline bool aabb_overlap(const float lmn[3], const float lmx[3], const float rmn[3], const float rmx[3])
{
    return !(lmn[0] >= rmx[0] || lmn[1] >= rmx[1] || lmn[2] >= rmx[2] || lmx[0] <= rmn[0] || lmx[1] <= rmn[1]
             || lmx[2] <= rmn[2] || rmn[0] >= lmx[0] || rmn[1] >= lmx[1] || rmn[2] >= lmx[2] || rmx[0] <= lmn[0]
             || rmx[1] <= lmn[1] || rmx[2] <= lmn[2]);
}

// Rotated copy: source voxel of destination voxel v, k quarter turns of
// R(x, y, z) = (-z - 1, y, x) (the turn about +Y that carries voxel cells
// onto voxel cells) about the pivots, clamped at 0 per axis like the
// translating copy (my_gpu_dag_editors.h:231).  No reference counterpart.
__host__ __device__ inline void rot_source(const EditorDev& e, const int64_t v[3], int64_t s[3])
{
    int64_t x = v[0] - e.pv[0], z = v[2] - e.pv[2];
    for (uint32_t i = 0; i < e.rot; ++i) {
        const int64_t nx = -z - 1;
        z = x;
        x = nx;
    }
    s[0] = e.qv[0] + x;
    s[1] = e.qv[1] + (v[1] - e.pv[1]);
    s[2] = e.qv[2] + z;
    for (int a = 0; a < 3; ++a)
        s[a] = s[a] > 0 ? s[a] : 0;
}

// Rotated copy, node level.  The node's destination voxels inside the edit
// box are an integer range; its image under rot_source (both corners mapped,
// then ordered) is the source range, and the children of the parents'
// source nodes that overlap it are the next parents.  Integer arithmetic
// throughout, so every destination voxel's source lies in a collected parent.
__device__ inline Op decide_copy_rot(const EditorDev& e, const DagDev& dag, uint32_t x, uint32_t y, uint32_t z,
                                     uint32_t depth, const CopyState& in, CopyState& out)
{
    for (int k = 0; k < 8; ++k)
        out.p[k].h = kSentinel;
    const uint32_t p[3] = { x, y, z };
    int64_t lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        const int64_t n0 = (int64_t)p[a] << depth;
        const int64_t n1 = n0 + ((int64_t)1 << depth) - 1;
        lo[a] = n0 > e.ivmin[a] ? n0 : e.ivmin[a];
        hi[a] = n1 < e.ivmax[a] ? n1 : e.ivmax[a];
        if (lo[a] > hi[a])
            return OP_KEEP;
    }
    int64_t a0[3], a1[3], slo[3], shi[3];
    rot_source(e, lo, a0);
    rot_source(e, hi, a1);
    for (int a = 0; a < 3; ++a) {
        slo[a] = a0[a] < a1[a] ? a0[a] : a1[a];
        shi[a] = a0[a] < a1[a] ? a1[a] : a0[a];
    }
    int nOut = 0;
    for (int k = 0; k < 8; ++k) {
        const CopyParent& par = in.p[k];
        if (par.h == kSentinel)
            break;
        const uint32_t* node = dag_node_ptr(dag, par.h);
        const uint32_t mask = node[0] & 0xFFu;
        for (uint32_t c = 0; c < 8; ++c) {
            if (!(mask & (1u << c)))
                continue;
            const uint32_t cp[3] = { (par.x << 1) | ((c >> 2) & 1u), (par.y << 1) | ((c >> 1) & 1u),
                                     (par.z << 1) | (c & 1u) };
            bool ov = true;
            for (int a = 0; a < 3; ++a) {
                const int64_t c0 = (int64_t)cp[a] << depth;
                const int64_t c1 = c0 + ((int64_t)1 << depth) - 1;
                ov = ov && c0 <= shi[a] && c1 >= slo[a];
            }
            if (ov && nOut < 8) {
                out.p[nOut].x = cp[0];
                out.p[nOut].y = cp[1];
                out.p[nOut].z = cp[2];
                out.p[nOut].h = node[1 + __builtin_popcount(mask & ((1u << c) - 1u))];
                ++nOut;
            }
        }
    }
    return nOut == 0 ? OP_KEEP : OP_SUBDIVIDE;
}

// Copy tool, stateful should_edit for an inner-level child at childDepth:
// collects the source children of the parent's source nodes that overlap the
// shifted destination region (my_gpu_dag_editors.h:173-217).
__device__ inline Op decide_copy(const EditorDev& e, const DagDev& dag, uint32_t x, uint32_t y, uint32_t z,
                                 uint32_t depth, const CopyState& in, CopyState& out)
{
    if (e.rot)
        return decide_copy_rot(e, dag, x, y, z, depth, in, out);
    float mn[3], mx[3];
    node_box(x, y, z, depth, mn, mx);
    for (int k = 0; k < 8; ++k)
        out.p[k].h = kSentinel;
    if (!box_overlap(e.bmin, e.bmax, mn, mx))
        return OP_KEEP;
    const float size = (float)(1u << depth);
    float smn[3], smx[3];
    for (int a = 0; a < 3; ++a) {
        smn[a] = fmax_ref(mn[a], e.bmin[a]) + e.d2s[a];
        smx[a] = fmin_ref(smn[a] + size, e.bmax[a] + e.d2s[a]);
    }
    int nOut = 0;
    for (int k = 0; k < 8; ++k) {
        const CopyParent& par = in.p[k];
        if (par.h == kSentinel)
            break;
        const uint32_t* node = dag_node_ptr(dag, par.h);
        const uint32_t mask = node[0] & 0xFFu;
        for (uint32_t c = 0; c < 8; ++c) {
            if (!(mask & (1u << c)))
                continue;
            const uint32_t cx = (par.x << 1) | ((c >> 2) & 1u);
            const uint32_t cy = (par.y << 1) | ((c >> 1) & 1u);
            const uint32_t cz = (par.z << 1) | (c & 1u);
            float cmn[3], cmx[3];
            node_box(cx, cy, cz, depth, cmn, cmx);
            if (aabb_overlap(cmn, cmx, smn, smx) && nOut < 8) {
                out.p[nOut].x = cx;
                out.p[nOut].y = cy;
                out.p[nOut].z = cz;
                out.p[nOut].h = node[1 + __builtin_popcount(mask & ((1u << c) - 1u))];
                ++nOut;
            }
        }
    }
    return nOut == 0 ? OP_KEEP : OP_SUBDIVIDE;
}

// Copy tool voxel value (my_gpu_dag_editors.h:225-252): the source voxel's
// material if set, else the old value; the source position clamps at 0.
__device__ inline bool copy_value(const EditorDev& e, const DagDev& dag, uint32_t vx, uint32_t vy, uint32_t vz,
                                  const CopyState& st, bool hasOld, uint32_t oldMat, uint32_t& mat)
{
    int32_t s[3];
    if (e.rot) {
        const int64_t v[3] = { vx, vy, vz };
        int64_t r[3];
        rot_source(e, v, r);
        // beyond the volume no parent matches (coordinates stay below 2^31)
        for (int a = 0; a < 3; ++a)
            s[a] = (int32_t)(r[a] < 0x7FFFFFFF ? r[a] : 0x7FFFFFFF);
    } else {
        s[0] = (int32_t)vx + e.id2s[0];
        s[1] = (int32_t)vy + e.id2s[1];
        s[2] = (int32_t)vz + e.id2s[2];
        for (int a = 0; a < 3; ++a)
            s[a] = s[a] > 0 ? s[a] : 0;
    }
    const uint32_t px = (uint32_t)s[0] >> 2, py = (uint32_t)s[1] >> 2, pz = (uint32_t)s[2] >> 2;
    const CopyParent* found = nullptr;
    for (int k = 0; k < 8; ++k) {
        if (st.p[k].h == kSentinel)
            break;
        if (st.p[k].x == px && st.p[k].y == py && st.p[k].z == pz)
            found = &st.p[k];
    }
    mat = oldMat;
    if (!found)
        return hasOld;
    const uint32_t bit = (((uint32_t)s[0] & 1u) << 2) | (((uint32_t)s[1] & 1u) << 1) | ((uint32_t)s[2] & 1u)
        | (((uint32_t)s[0] & 2u) << 4) | (((uint32_t)s[1] & 2u) << 3) | (((uint32_t)s[2] & 2u) << 2);
    uint32_t m;
    if (leaf_material(dag_leaf_ptr(dag, found->h), bit, dag.M, m)) {
        mat = m;
        return true;
    }
    return hasOld;
}

} // namespace svdag
