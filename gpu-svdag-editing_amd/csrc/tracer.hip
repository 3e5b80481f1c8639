// Primary-ray tracer over the editable DAG.
//
// reference: DAGTracer::resolve_paths / get_trace_params (dag_tracer.cu:59-134),
// Tracer::trace_paths, intersect_ray, compute_intersection_mask,
// createSurfaceInteraction (tracer.cu:45-273, 353-411).
//
// One lane per pixel, 16x16 pixel tiles per 256-lane block.  The per-level
// traversal stack (node handle, child mask, visit mask) is kept in LDS laid
// out [level][lane] so that a wave's pushes and pops hit 64 consecutive
// 8-byte slots.  Float arithmetic follows the reference operation by
// operation.  Nothing is contracted by the compiler; explicit fma/fmaf stand
// where nvcc's default --fmad=true has exactly one possible fusion (the
// per-pixel ray setup, the hit position, the shadow and occlusion ray
// origins).  Sites with two products in one sum (the slab bounds
// tmid +- radius*|invDir|, dot products) stay rounded per operation: which
// product nvcc would fuse there is not fixed by its documentation.  The
// camera setup is double precision.
#include "dag_internal.h"
#include <numbers>
#include <numeric>
#include <cfloat>
#include <cmath>
#include <cstring>

namespace svdag {

struct TraceParams {
    double cam[3], rayMin[3], ddx[3], ddy[3];
    uint32_t width, height;
};

struct Ray {
    float o[3], d[3], inv[3];
    float tmin, tmax;
};

__device__ inline uint32_t next_child(uint32_t order, uint32_t mask)
{
    for (uint32_t c = 0; c < 8; ++c) {
        const uint32_t ci = c ^ order;
        if (mask & (1u << ci))
            return ci;
    }
    return 0;
}

template <bool kRoot>
__device__ inline uint32_t intersection_mask(uint32_t levels, uint32_t level, uint32_t px, uint32_t py, uint32_t pz,
                                             const Ray& ray)
{
    const uint32_t shift = levels - level;
    const float radius = (float)(1u << (shift - 1));
    const float pos[3] = { (float)(px << shift), (float)(py << shift), (float)(pz << shift) };
    float cr[3], tmid[3], pmin[3], pmax[3];
    for (int a = 0; a < 3; ++a) {
        const float center = radius + pos[a];
        cr[a] = center - ray.o[a];
        tmid[a] = cr[a] * ray.inv[a];
        const float slab = radius * fabsf(ray.inv[a]);
        pmin[a] = tmid[a] - slab;
        pmax[a] = tmid[a] + slab;
    }
    float tmin = fmax_ref(pmin[0], fmax_ref(pmin[1], pmin[2]));
    tmin = fmax_ref(tmin, 0.0f);
    float tmax = fmin_ref(pmax[0], fmin_ref(pmax[1], pmax[2]));
    tmin = fmax_ref(tmin, ray.tmin);
    tmax = fmin_ref(tmax, ray.tmax);
    if (kRoot && tmin >= tmax)
        return 0;
    uint32_t mask = 0;
    {
        const float tm = 0.5f * (tmin + tmax);
        const float p0 = tm * ray.d[0], p1 = tm * ray.d[1], p2 = tm * ray.d[2];
        const uint32_t first = (p0 >= cr[0] ? 4u : 0u) + (p1 >= cr[1] ? 2u : 0u) + (p2 >= cr[2] ? 1u : 0u);
        mask |= 1u << first;
    }
    const float eps = 1e-4f;
    if (tmin <= tmid[0] && tmid[0] <= tmax) {
        const float p1 = tmid[0] * ray.d[1], p2 = tmid[0] * ray.d[2];
        uint32_t A = 0, B = 0;
        if (p1 >= cr[1] - eps) A |= 0xCCu;
        if (p1 <= cr[1] + eps) A |= 0x33u;
        if (p2 >= cr[2] - eps) B |= 0xAAu;
        if (p2 <= cr[2] + eps) B |= 0x55u;
        mask |= A & B;
    }
    if (tmin <= tmid[1] && tmid[1] <= tmax) {
        const float p0 = tmid[1] * ray.d[0], p2 = tmid[1] * ray.d[2];
        uint32_t C = 0, D = 0;
        if (p0 >= cr[0] - eps) C |= 0xF0u;
        if (p0 <= cr[0] + eps) C |= 0x0Fu;
        if (p2 >= cr[2] - eps) D |= 0xAAu;
        if (p2 <= cr[2] + eps) D |= 0x55u;
        mask |= C & D;
    }
    if (tmin <= tmid[2] && tmid[2] <= tmax) {
        const float p0 = tmid[2] * ray.d[0], p1 = tmid[2] * ray.d[1];
        uint32_t E = 0, F = 0;
        if (p0 >= cr[0] - eps) E |= 0xF0u;
        if (p0 <= cr[0] + eps) E |= 0x0Fu;
        if (p1 >= cr[1] - eps) F |= 0xCCu;
        if (p1 <= cr[1] + eps) F |= 0x33u;
        mask |= E & F;
    }
    return mask & 0xFFu;
}

constexpr uint32_t kTraceBlock = 256;

// intersect_ray (tracer.cu:186-273): the first set voxel along the ray, in
// visit order, and its material; (0, 0, 0) and kSentinel when there is none.
// stk is this lane's column of the block's [level][lane] LDS stack.
__device__ inline void closest_voxel(const DagDev& dag, uint32_t root, const Ray& ray, uint2 (*stk)[kTraceBlock],
                                     uint32_t tid, uint32_t& x, uint32_t& y, uint32_t& z, uint32_t& material)
{
    const uint32_t levels = dag.levels;
    const uint32_t leafLevel = dag.leafLevel;
    const uint32_t order = (ray.d[0] < 0.f ? 4u : 0u) + (ray.d[1] < 0.f ? 2u : 0u) + (ray.d[2] < 0.f ? 1u : 0u);
    uint32_t level = 0;
    x = y = z = 0;
    material = kSentinel;

    if (root != kSentinel) {
        uint32_t cIndex = root;
        uint32_t cChild = dag_node_ptr(dag, root)[0] & 0xFFu;
        uint32_t cVisit = cChild & intersection_mask<true>(levels, 0, 0, 0, 0, ray);
        uint32_t addr = 0;
        uint64_t leafBits = 0;
        for (;;) {
            // Ascend while the current node has no child left to visit.
            uint32_t newLevel = level;
            while (newLevel > 0 && !cVisit) {
                --newLevel;
                const uint2 e = stk[newLevel][tid];
                cIndex = e.x;
                cChild = e.y & 0xFFu;
                cVisit = (e.y >> 8) & 0xFFu;
            }
            if (newLevel == 0 && !cVisit) {
                x = y = z = 0;
                break;
            }
            const uint32_t up = level - newLevel;
            x >>= up;
            y >>= up;
            z >>= up;
            level = newLevel;

            const uint32_t nc = next_child(order, cVisit);
            cVisit &= ~(1u << nc);
            x = (x << 1) | ((nc >> 2) & 1u);
            y = (y << 1) | ((nc >> 1) & 1u);
            z = (z << 1) | (nc & 1u);
            stk[level][tid] = make_uint2(cIndex, cChild | (cVisit << 8));
            ++level;

            if (level == levels) {
                const uint32_t v = ((x & 1u) << 2) | ((y & 1u) << 1) | (z & 1u) | ((x & 2u) << 4) | ((y & 2u) << 3)
                    | ((z & 2u) << 2);
                uint32_t m;
                material = leaf_material(dag_leaf_ptr(dag, addr), v, dag.M, m) ? m : kSentinel;
                break;
            }
            if (level < leafLevel) {
                const uint32_t* node = dag_node_ptr(dag, cIndex);
                cIndex = node[1 + __builtin_popcount(cChild & ((1u << nc) - 1u))];
                cChild = dag_node_ptr(dag, cIndex)[0] & 0xFFu;
            } else if (level == leafLevel) {
                const uint32_t* node = dag_node_ptr(dag, cIndex);
                addr = node[1 + __builtin_popcount(cChild & ((1u << nc) - 1u))];
                const uint32_t* leaf = dag_leaf_ptr(dag, addr);
                leafBits = (uint64_t)leaf[0] | ((uint64_t)leaf[1] << 32);
                uint32_t fm = 0;
                for (uint32_t b = 0; b < 8; ++b)
                    fm |= ((leafBits >> (8 * b)) & 0xFFull) ? (1u << b) : 0u;
                cChild = fm;
            } else {
                cChild = (uint32_t)(leafBits >> (8 * nc)) & 0xFFu;
            }
            cVisit = cChild & intersection_mask<false>(levels, level, x, y, z, ray);
        }
    } else {
        x = y = z = 0;
    }

}

// createSurfaceInteraction (tracer.cu:353-389): the entry face of voxel
// (x, y, z) along the ray: t, the position origin + t * direction (one
// fusion possible), and the surface code 1 + 2 * axis + (the normal, which
// points against the ray, is +axis); 0 for the null path (0, 0, 0), which
// is no surface (Path::is_null, path.h:61-64).
__device__ inline uint32_t entry_face(const Ray& ray, uint32_t x, uint32_t y, uint32_t z, float& t, float pos[3])
{
    const float bmin[3] = { (float)x, (float)y, (float)z };
    float dists[3];
    for (int k = 0; k < 3; ++k) {
        const float bmax = bmin[k] + 1.0f;
        const float t1 = (bmin[k] - ray.o[k]) * ray.inv[k];
        const float t2 = (bmax - ray.o[k]) * ray.inv[k];
        dists[k] = fmin_ref(t1, t2);
    }
    const int axis = dists[0] > dists[1] ? (dists[0] > dists[2] ? 0 : 2) : (dists[1] > dists[2] ? 1 : 2);
    t = dists[axis];
    for (int k = 0; k < 3; ++k)
        pos[k] = fmaf(t, ray.d[k], ray.o[k]);
    return (x | y | z) == 0u ? 0u : 1u + 2u * (uint32_t)axis + (ray.d[axis] < 0.0f ? 1u : 0u);
}

__global__ void __launch_bounds__(kTraceBlock) trace_kernel(DagDev dag, uint32_t root, TraceParams p, svdag_hit* out,
                                                             uint8_t* surf)
{
    __shared__ uint2 stack[20][kTraceBlock]; // levels <= 20
    const uint32_t tid = threadIdx.x;
    const uint32_t tilesX = (p.width + 15) / 16;
    const uint32_t px = (blockIdx.x % tilesX) * 16 + (tid & 15u);
    const uint32_t py = (blockIdx.x / tilesX) * 16 + (tid >> 4);
    if (px >= p.width || py >= p.height)
        return;

    // Per-pixel ray in double, normalised, then rounded to float (tracer.cu:402-405).
    // rayMin + px*ddx + py*ddy has one fusion per add under --fmad=true:
    // fma(py, ddy, fma(px, ddx, rayMin)).
    double a[3];
    for (int k = 0; k < 3; ++k)
        a[k] = fma((double)py, p.ddy[k], fma((double)px, p.ddx[k], p.rayMin[k])) - p.cam[k];
    const double len = sqrt((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]);
    const double inv = 1.0 / len;
    Ray ray;
    for (int k = 0; k < 3; ++k) {
        ray.o[k] = (float)p.cam[k];
        ray.d[k] = (float)(inv * a[k]);
        ray.inv[k] = 1.0f / ray.d[k];
    }
    ray.tmin = 0.0f;
    ray.tmax = FLT_MAX;

    uint32_t x, y, z, material;
    closest_voxel(dag, root, ray, stack, tid, x, y, z, material);

    svdag_hit h;
    h.x = x;
    h.y = y;
    h.z = z;
    h.material = material;
    float pos[3];
    const uint32_t code = entry_face(ray, x, y, z, h.t, pos);
    h.px = pos[0];
    h.py = pos[1];
    h.pz = pos[2];
    const size_t at = (size_t)(p.height - 1 - py) * p.width + px;
    out[at] = h;
    if (surf)
        surf[at] = (uint8_t)code;
}

// ------------------------------------------------------------------ sampling
// PCG32: XSH-RR output over a 64-bit LCG (O'Neill 2014), seeded as the
// reference seeds it (utils.h:45-82): the stream value is the increment, the
// state starts at PCG's default initializer and two outputs are dropped.
struct Pcg32 {
    uint64_t state, inc;
    __device__ explicit Pcg32(uint64_t stream) : state(0x853c49e6748fea9bull), inc(stream)
    {
        next();
        next();
    }
    __device__ uint32_t next()
    {
        const uint64_t old = state;
        state = old * 6364136223846793005ull + (inc | 1ull);
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    // float(u32) / float(UINT32_MAX); the divisor rounds to 2^32.
    __device__ float next_float() { return (float)next() / 4294967296.0f; }
};

// sin and cos of 2*pi*u for u in [0, 1].  u is split into the nearest
// quarter turn k and a remainder f of at most an eighth of a turn; sin and
// cos of f*pi/2 come from their Taylor series to degree 9 and 10 (remainder
// below 2e-9 on [-pi/4, pi/4]), then rotate by k quarter turns.  This is our
// own formulation; the oracle evaluates the same float operations, so the
// sampled directions agree bit for bit.
__device__ inline void sincos_turns(float u, float& sn, float& cs)
{
    const float q = 4.0f * u;
    const float k = floorf(q + 0.5f);
    const float x = (q - k) * 1.57079632679489662f;
    const float x2 = x * x;
    const float sp = x * (1.0f + x2 * (-0.166666667f + x2 * (8.33333333e-3f + x2 * (-1.98412698e-4f + x2 * 2.75573192e-6f))));
    const float cp = 1.0f + x2 * (-0.5f + x2 * (4.16666667e-2f + x2 * (-1.38888889e-3f + x2 * (2.48015873e-5f + x2 * -2.75573192e-7f))));
    switch ((int)k & 3) {
    case 0:
        sn = sp;
        cs = cp;
        break;
    case 1:
        sn = cp;
        cs = -sp;
        break;
    case 2:
        sn = -sp;
        cs = -cp;
        break;
    default:
        sn = -cp;
        cs = sp;
        break;
    }
}

// cosineSampleHemisphere (tracer.cu:645-659): r = sqrt(u.x), theta = 2*pi*u.y.
__device__ inline void cosine_hemisphere(float ux, float uy, float out[3])
{
    const float r = sqrtf(ux);
    float sn, cs;
    sincos_turns(uy, sn, cs);
    const float zz = 1.0f - ux;
    out[0] = r * cs;
    out[1] = r * sn;
    out[2] = sqrtf(0.0f < zz ? zz : 0.0f);
}

// ------------------------------------------------------------------ shadows and ambient occlusion
// reference: Tracer::trace_shadows / trace_ambient_occlusion (tracer.cu:613-691),
// intersect_ray_node_out_of_order (274-351), cosineSampleHemisphere (645-659),
// SurfaceInteraction::transformDirectionToWorld (tracer.h:106-109) and
// Utils::PCG_RNG (utils.h:45-82).  Both passes read the surface of the last
// primary trace: the hit position and the entry-face normal.

// Any set voxel along the ray: children are taken highest visit bit first
// and the first voxel reached ends the walk.  stk is this lane's column of
// the block's [level][lane] LDS stack.
__device__ inline bool occluded(const DagDev& dag, uint32_t root, const Ray& ray, uint2 (*stk)[kTraceBlock],
                                uint32_t tid)
{
    if (root == kSentinel)
        return false;
    const uint32_t levels = dag.levels, leafLevel = dag.leafLevel;
    uint32_t level = 0, x = 0, y = 0, z = 0;
    uint32_t cIndex = root;
    uint32_t cChild = dag_node_ptr(dag, root)[0] & 0xFFu;
    uint32_t cVisit = cChild & intersection_mask<true>(levels, 0, 0, 0, 0, ray);
    uint64_t leafBits = 0;
    for (;;) {
        uint32_t newLevel = level;
        while (newLevel > 0 && !cVisit) {
            --newLevel;
            const uint2 e = stk[newLevel][tid];
            cIndex = e.x;
            cChild = e.y & 0xFFu;
            cVisit = (e.y >> 8) & 0xFFu;
        }
        if (newLevel == 0 && !cVisit)
            return false;
        const uint32_t up = level - newLevel;
        x >>= up;
        y >>= up;
        z >>= up;
        level = newLevel;
        const uint32_t nc = 31u - (uint32_t)__clz((int)cVisit);
        cVisit &= ~(1u << nc);
        x = (x << 1) | ((nc >> 2) & 1u);
        y = (y << 1) | ((nc >> 1) & 1u);
        z = (z << 1) | (nc & 1u);
        stk[level][tid] = make_uint2(cIndex, cChild | (cVisit << 8));
        ++level;
        if (level == levels)
            return true;
        if (level < leafLevel) {
            const uint32_t* node = dag_node_ptr(dag, cIndex);
            cIndex = node[1 + __builtin_popcount(cChild & ((1u << nc) - 1u))];
            cChild = dag_node_ptr(dag, cIndex)[0] & 0xFFu;
        } else if (level == leafLevel) {
            const uint32_t* node = dag_node_ptr(dag, cIndex);
            const uint32_t* leaf = dag_leaf_ptr(dag, node[1 + __builtin_popcount(cChild & ((1u << nc) - 1u))]);
            leafBits = (uint64_t)leaf[0] | ((uint64_t)leaf[1] << 32);
            uint32_t fm = 0;
            for (uint32_t b = 0; b < 8; ++b)
                fm |= ((leafBits >> (8 * b)) & 0xFFull) ? (1u << b) : 0u;
            cChild = fm;
        } else {
            cChild = (uint32_t)(leafBits >> (8 * nc)) & 0xFFu;
        }
        cVisit = cChild & intersection_mask<false>(levels, level, x, y, z, ray);
    }
}

__device__ inline void make_ray(Ray& r, const float o[3], const float d[3], float tmax)
{
    for (int k = 0; k < 3; ++k) {
        r.o[k] = o[k];
        r.d[k] = d[k];
        r.inv[k] = 1.0f / d[k];
    }
    r.tmin = 0.0f;
    r.tmax = tmax;
}

// Normal, dpdu and dpdv of a surface code (createSurfaceInteraction,
// tracer.cu:353-389).
__device__ inline void surface_frame(uint32_t code, float n[3], float du[3], float dv[3])
{
    const uint32_t axis = (code - 1u) >> 1;
    for (int k = 0; k < 3; ++k)
        n[k] = du[k] = dv[k] = 0.0f;
    n[axis] = ((code - 1u) & 1u) ? 1.0f : -1.0f;
    if (axis == 0) {
        du[2] = 1.0f;
        dv[1] = 1.0f;
    } else if (axis == 1) {
        du[0] = 1.0f;
        dv[2] = 1.0f;
    } else {
        du[0] = 1.0f;
        dv[1] = 1.0f;
    }
}

__global__ void __launch_bounds__(kTraceBlock) shadow_kernel(DagDev dag, uint32_t root, const svdag_hit* hits,
                                                             const uint8_t* surf, uint32_t n, float bias, float3 sun,
                                                             float* out)
{
    __shared__ uint2 stk[20][kTraceBlock];
    const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t code = surf[i];
    if (code == 0u) {
        out[i] = 0.0f;
        return;
    }
    const svdag_hit h = hits[i];
    const float L[3] = { sun.x, sun.y, sun.z };
    const float p[3] = { h.px, h.py, h.pz };
    // si.position + shadowBias * sun_direction() (tracer.cu:621), one fusion possible.
    float o[3];
    for (int k = 0; k < 3; ++k)
        o[k] = fmaf(bias, L[k], p[k]);
    Ray r;
    make_ray(r, o, L, FLT_MAX);
    float nrm[3], du[3], dv[3];
    surface_frame(code, nrm, du, dv);
    float ndl = (nrm[0] * L[0] + nrm[1] * L[1]) + nrm[2] * L[2];
    ndl = ndl < 0.0f ? 0.0f : ndl; // std::max(NdotL, 0.0f)
    out[i] = occluded(dag, root, r, stk, threadIdx.x) ? 0.0f : ndl;
}

__global__ void __launch_bounds__(kTraceBlock) ao_kernel(DagDev dag, uint32_t root, const svdag_hit* hits,
                                                         const uint8_t* surf, uint32_t n, uint32_t numSamples,
                                                         float bias, float rayLength, uint64_t seed, float* out)
{
    __shared__ uint2 stk[20][kTraceBlock];
    const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t code = surf[i];
    if (code == 0u) {
        out[i] = 0.0f;
        return;
    }
    const svdag_hit h = hits[i];
    const float p[3] = { h.px, h.py, h.pz };
    float nrm[3], du[3], dv[3];
    surface_frame(code, nrm, du, dv);
    // si.position + si.normal * shadowBias (tracer.cu:675), one fusion possible.
    float o[3];
    for (int k = 0; k < 3; ++k)
        o[k] = fmaf(nrm[k], bias, p[k]);
    // The sample stream is keyed by the pixel's index in the stored image.
    Pcg32 rng((uint64_t)i * (seed > 1ull ? seed : 1ull));
    uint32_t hitsUnoccluded = 0;
    for (uint32_t s = 0; s < numSamples; ++s) {
        const float ux = rng.next_float(), uy = rng.next_float();
        float local[3];
        cosine_hemisphere(ux, uy, local);
        float d[3];
        for (int k = 0; k < 3; ++k)
            d[k] = (local[0] * du[k] + local[1] * dv[k]) + local[2] * nrm[k];
        Ray r;
        make_ray(r, o, d, rayLength);
        if (!occluded(dag, root, r, stk, threadIdx.x))
            ++hitsUnoccluded;
    }
    out[i] = (float)hitsUnoccluded / (float)numSamples;
}

// ---------------------------------------------------------------- colour and lighting
// ColorUtils (color_utils.h:31-45, 233-252), clamp (cuda_math.h:44-45) and
// the light colours of tracer.cu:598-611, in float as there.
__device__ inline float clamp01(float f)
{
    return f < 0.0f ? 0.0f : (f > 1.0f ? 1.0f : f);
}
__device__ inline uint32_t to_rgb888(const float c[3])
{
    return (uint32_t)(clamp01(c[0]) * 255.0f) | ((uint32_t)(clamp01(c[1]) * 255.0f) << 8)
           | ((uint32_t)(clamp01(c[2]) * 255.0f) << 16) | 0xff000000u;
}
__device__ inline float srgb_to_linear(float c)
{
    const float lo = c / 12.92f;
    const float hi = powf((c + 0.055f) / 1.055f, 2.4f);
    return c <= 0.04045f ? lo : hi;
}
// pow(|c|, 1/2.4) * 1.055 - 0.055: the one product-then-add, fused as nvcc
// fuses it by default.
__device__ inline float linear_to_srgb(float c)
{
    const float lo = c * 12.92f;
    const float hi = fmaf(powf(fabsf(c), 1.0f / 2.4f), 1.055f, -0.055f);
    return c <= 0.0031308f ? lo : hi;
}
__device__ inline float sky_channel(int k)
{
    return (k == 0 ? 187.0f : (k == 1 ? 242.0f : 250.0f)) / 255.0f;
}
__device__ inline float sun_channel(int k)
{
    return k == 0 ? 0.95f : (k == 1 ? 0.94f : 0.47f);
}

// Tracer::trace_colors -> computeColorAtSurface (tracer.cu:413-596) for a DAG
// with materials, each material's textures being one colour (texture images
// are not part of this library): no surface -> 0; a material without a
// colour -> the invalid-colour checker (1, b, 1 - b), b = (x ^ y ^ z) & 1 of
// the hit voxel; else the material's colour; stored as RGBA8.
__global__ void __launch_bounds__(256) colors_kernel(const svdag_hit* hits, uint32_t n, const float* palette,
                                                     uint32_t numColors, uint32_t* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const svdag_hit h = hits[i];
    float c[3] = { 0.0f, 0.0f, 0.0f };
    if (h.x != 0u || h.y != 0u || h.z != 0u) {
        if (h.material >= numColors) {
            const float b = (float)((h.x ^ h.y ^ h.z) & 1u);
            c[0] = 1.0f;
            c[1] = b;
            c[2] = 1.0f - b;
        } else {
            for (int k = 0; k < 3; ++k)
                c[k] = palette[3 * h.material + k];
        }
    }
    out[i] = to_rgb888(c);
}

struct LightingDev {
    double cam[3], rayMin[3], ddx[3], ddy[3];
    float sunDir[3];
    float sunBrightness, fogDensity;
    uint32_t width, applyAo, applySun;
};

// Tracer::trace_lighting (tracer.cu:709-750) with applyFog (tracer.cu:693-707);
// the tool overlay (TOOL_OVERLAY) is not built.  Pixel (x, y) of the stored
// image reads the stored surface, occlusion and sun at its own index, and
// reconstructs its fog ray from (x, y) as the reference does.
__global__ void __launch_bounds__(256) lighting_kernel(const svdag_hit* hits, uint32_t n, const uint32_t* colors,
                                                       const float* ao, const float* sun, LightingDev p,
                                                       uint32_t* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const svdag_hit h = hits[i];
    float c[3];
    if (h.x == 0u && h.y == 0u && h.z == 0u) {
        for (int k = 0; k < 3; ++k)
            c[k] = sky_channel(k);
    } else {
        const uint32_t col = colors[i];
        const float occl = p.applyAo ? ao[i] : 1.0f;
        const float vis = p.applySun ? sun[i] : 1.0f;
        for (int k = 0; k < 3; ++k) {
            const float kd = srgb_to_linear((float)((col >> (8 * k)) & 0xFFu) / 255.0f);
            const float ambient = (occl * 0.5f) * sky_channel(k);
            const float direct = (vis * sun_channel(k)) * p.sunBrightness;
            c[k] = (ambient + direct) * kd;
        }
        if (p.fogDensity > 0.0f) {
            const uint32_t x = i % p.width, y = i / p.width;
            double d[3];
            for (int k = 0; k < 3; ++k)
                d[k] = fma((double)y, p.ddy[k], fma((double)x, p.ddx[k], p.rayMin[k])) - p.cam[k];
            const double len = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
            float dir[3], o[3];
            for (int k = 0; k < 3; ++k) {
                dir[k] = (float)(d[k] * (1.0 / len));
                o[k] = (float)p.cam[k];
            }
            const float pos[3] = { h.px, h.py, h.pz };
            float dx[3];
            for (int k = 0; k < 3; ++k)
                dx[k] = o[k] - pos[k];
            const float dist = sqrtf((dx[0] * dx[0] + dx[1] * dx[1]) + dx[2] * dx[2]);
            const float fogAmount = 1.0f - expf(-dist * (p.fogDensity * 0.00001f));
            const float sd = (dir[0] * p.sunDir[0] + dir[1] * p.sunDir[1]) + dir[2] * p.sunDir[2];
            const float sunAmount = 1.01f * fmaxf(sd, 0.0f);
            const float w = powf(sunAmount, 30.0f);
            const float f = clamp01(fogAmount);
            for (int k = 0; k < 3; ++k) {
                const float fogColor = sky_channel(k) * (1.0f - w) + 1.0f * w;
                c[k] = c[k] * (1.0f - f) + fogColor * f;
            }
        }
    }
    float s[3];
    for (int k = 0; k < 3; ++k)
        s[k] = linear_to_srgb(c[k]);
    out[i] = to_rgb888(s);
}

struct PathDev {
    TraceParams cam;
    float sunDir[3];
    float env;
    uint32_t numSamples, maxDepth, integrate, numAccumulated, numColors;
    uint64_t seed;
};

// Tracer::trace_path_tracing (tracer.cu:758-842), one lane per stored pixel
// i = y * W + x.  The PCG32 stream is i * max(seed, 1).  Sample camera rays
// go through (x, H - 1 - y) (plus a jitter when integrating); at depth 0
// without integration the path starts on the stored primary surface of
// pixel i with the colour pass's colour.  Per bounce: nothing hit -> sky
// times env, stop; else next-event estimation toward the sun from
// position + 0.1 * normal, then a uniform hemisphere sample (z = u.x,
// phi = 2 pi u.y, pdf 1 / (2 pi)) continues the path with
// c *= brdf / pdf, brdf(d) = max(0, N.d) / pi * Kd, Kd the material's colour
// made linear.  Sums with one product per add are fused as nvcc fuses them;
// the products with the axis-aligned frame vectors are exact either way.
__global__ void __launch_bounds__(kTraceBlock) path_kernel(DagDev dag, uint32_t root, PathDev p, const svdag_hit* hits,
                                                           const uint8_t* surf, const uint32_t* colors,
                                                           const float* palette, float* accum, uint32_t* out)
{
    __shared__ uint2 stack[20][kTraceBlock];
    const uint32_t tid = threadIdx.x;
    const uint32_t W = p.cam.width, H = p.cam.height;
    const uint32_t i = blockIdx.x * kTraceBlock + tid;
    if (i >= W * H)
        return;
    const uint32_t x = i % W, y = i / W;
    Pcg32 rng((uint64_t)i * (p.seed > 1ull ? p.seed : 1ull));
    float Lo[3] = { accum[3 * i], accum[3 * i + 1], accum[3 * i + 2] };
    const float invPi = std::numbers::inv_pi_v<float>;
    const float pdf = 0.5f * invPi;
    const float eps = 0.1f;
    for (uint32_t s = 0; s < p.numSamples; ++s) {
        float sx = (float)x, sy = (float)y;
        if (p.integrate) {
            sx += rng.next_float();
            sy += rng.next_float();
        }
        const double row = (double)((float)(H - 1) - sy);
        double a[3];
        for (int k = 0; k < 3; ++k)
            a[k] = fma(row, p.cam.ddy[k], fma((double)sx, p.cam.ddx[k], p.cam.rayMin[k])) - p.cam.cam[k];
        const double len = sqrt((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]);
        const double inv = 1.0 / len;
        float o[3], dir[3];
        for (int k = 0; k < 3; ++k) {
            o[k] = (float)p.cam.cam[k];
            dir[k] = (float)(inv * a[k]);
        }
        float c[3] = { 1.0f, 1.0f, 1.0f };
        for (uint32_t depth = 0; depth < p.maxDepth; ++depth) {
            uint32_t code = 0;
            float pos[3] = { 0.0f, 0.0f, 0.0f }, kd[3] = { 0.0f, 0.0f, 0.0f };
            if (depth == 0 && !p.integrate) {
                const svdag_hit h = hits[i];
                code = surf[i];
                pos[0] = h.px;
                pos[1] = h.py;
                pos[2] = h.pz;
                for (int k = 0; k < 3; ++k)
                    kd[k] = srgb_to_linear((float)((colors[i] >> (8 * k)) & 0xFFu) / 255.0f);
            } else {
                Ray r;
                make_ray(r, o, dir, FLT_MAX);
                uint32_t hx, hy, hz, m;
                closest_voxel(dag, root, r, stack, tid, hx, hy, hz, m);
                float t;
                code = entry_face(r, hx, hy, hz, t, pos);
                if (code) {
                    float col[3];
                    if (m >= p.numColors) {
                        const float b = (float)((hx ^ hy ^ hz) & 1u);
                        col[0] = 1.0f;
                        col[1] = b;
                        col[2] = 1.0f - b;
                    } else {
                        for (int k = 0; k < 3; ++k)
                            col[k] = palette[3 * m + k];
                    }
                    for (int k = 0; k < 3; ++k)
                        kd[k] = srgb_to_linear(col[k]);
                }
            }
            if (!code) {
                for (int k = 0; k < 3; ++k)
                    Lo[k] = fmaf(c[k] * sky_channel(k), p.env, Lo[k]);
                break;
            }
            float n[3], du[3], dv[3];
            surface_frame(code, n, du, dv);
            float so[3];
            for (int k = 0; k < 3; ++k)
                so[k] = fmaf(eps, n[k], pos[k]);
            Ray sr;
            make_ray(sr, so, p.sunDir, FLT_MAX);
            if (!occluded(dag, root, sr, stack, tid)) {
                float ndl = (n[0] * p.sunDir[0] + n[1] * p.sunDir[1]) + n[2] * p.sunDir[2];
                ndl = 0.0f > ndl ? 0.0f : ndl;
                for (int k = 0; k < 3; ++k)
                    Lo[k] = fmaf((c[k] * ((ndl * invPi) * kd[k])) * sun_channel(k), p.env, Lo[k]);
            }
            const float ux = rng.next_float(), uy = rng.next_float();
            const float q = fmaf(-ux, ux, 1.0f);
            const float rr = sqrtf(0.0f < q ? q : 0.0f);
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
        const float m = Lo[k] / (float)(p.numSamples + p.numAccumulated);
        v[k] = linear_to_srgb(m < 1.0f ? m : 1.0f);
    }
    out[i] = to_rgb888(v);
}

// BlurKernel (blur.cu:8-34): 2 * half + 1 weights of gaussian(half + 0.5,
// sigma, x), sigma = 0.3 (half - 1) + 0.8 passed as the variance, each
// divided by their std::reduce sum; host float arithmetic, as there.
struct BlurWeights {
    float w[9];
};
static BlurWeights blur_weights()
{
    constexpr uint32_t half = 4; // DAGTracer's ambientOcclusionBlurKernel(4) (dag_tracer.cu:19)
    const float sigma = 0.3f * (float)(half - 1) + 0.8f;
    const float mean = (float)half + 0.5f;
    std::vector<float> k(2 * half + 1);
    for (size_t x = 0; x < k.size(); ++x) {
        const float norm = 1.0f / std::sqrt(2 * std::numbers::pi_v<float> * sigma);
        const float d = (float)x - mean;
        k[x] = norm * std::exp(-0.5f * d * d / sigma);
    }
    const float sum = std::reduce(k.begin(), k.end(), 0.0f, std::plus<float>());
    BlurWeights b {};
    for (size_t x = 0; x < k.size(); ++x)
        b.w[x] = k[x] / sum;
    return b;
}

// horizontalKernel / verticalKernel (blur.cu:36-66): taps outside the image
// are skipped (no renormalisation); sum += weight * value, one fusion.
template <bool kVertical>
__global__ void __launch_bounds__(256) blur_kernel(const float* in, float* out, uint32_t W, uint32_t H, BlurWeights b)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * H)
        return;
    const int x = (int)(i % W), y = (int)(i / W);
    float sum = 0.0f;
    for (int d = -4; d <= 4; ++d) {
        const int nx = kVertical ? x : x + d, ny = kVertical ? y + d : y;
        if (kVertical ? (ny >= 0 && ny < (int)H) : (nx >= 0 && nx < (int)W))
            sum = fmaf(b.w[d + 4], in[(size_t)ny * W + nx], sum);
    }
    out[i] = sum;
}

// get_trace_params (dag_tracer.cu:59-101), double precision.
static TraceParams trace_params(const svdag_camera& c, uint32_t levels, uint32_t W, uint32_t H)
{
    const double* R = c.rotation;
    const double right[3] = { -R[0], -R[1], -R[2] };
    const double up[3] = { R[3], R[4], R[5] };
    const double fwd[3] = { R[6], R[7], R[8] };
    const double fov = c.fov_degrees / 2.0 * (double(M_PI) / 180.);
    const double aspect = double(W) / double(H);
    TraceParams p {};
    const double sf = sin(fov), cf = cos(fov);
    double bl[3], br[3], tl[3];
    for (int k = 0; k < 3; ++k) {
        const double X = right[k] * sf * aspect;
        const double Y = up[k] * sf;
        const double Z = fwd[k] * cf;
        bl[k] = c.position[k] + Z - Y - X;
        br[k] = c.position[k] + Z - Y + X;
        tl[k] = c.position[k] + Z + Y - X;
    }
    for (int k = 0; k < 3; ++k) {
        const double tr = -c.bounds_min[k];
        const double sc = double(1u << levels) / (c.bounds_max[k] - c.bounds_min[k]);
        const double fpos = (c.position[k] + tr) * sc;
        const double fbl = (bl[k] + tr) * sc;
        const double ftl = (tl[k] + tr) * sc;
        const double fbr = (br[k] + tr) * sc;
        p.cam[k] = fpos;
        p.rayMin[k] = fbl;
        p.ddx[k] = (fbr - fbl) * (1.0 / W);
        p.ddy[k] = (ftl - fbl) * (1.0 / H);
    }
    p.width = W;
    p.height = H;
    return p;
}

// Enqueue a trace of the DAG rooted at `root` (a value, so concurrent edits
// that publish new roots cannot change what this frame sees).
int launch_trace(svdag_ctx* ctx, const svdag_camera& cam, uint32_t W, uint32_t H, uint32_t root, hipStream_t stream,
                 svdag_hit* outDev, uint8_t* surfDev)
{
    const TraceParams p = trace_params(cam, ctx->dag.levels, W, H);
    const uint32_t tiles = ((W + 15) / 16) * ((H + 15) / 16);
    trace_kernel<<<tiles, kTraceBlock, 0, stream>>>(ctx->dag, root, p, outDev, surfDev);
    SVDAG_HIP(hipGetLastError());
    return SVDAG_OK;
}

} // namespace svdag

using namespace svdag;

extern "C" int svdag_trace_primary(svdag_ctx* ctx, const svdag_camera* cam, uint32_t W, uint32_t H, svdag_hit* out,
                                   float* ms)
{
    if (!ctx || !cam || !out || W == 0 || H == 0 || (uint64_t)W * H > (1ull << 28)) {
        set_error("svdag_trace_primary: bad arguments");
        return SVDAG_E_INVALID;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    const size_t n = (size_t)W * H;
    if (ctx->hitCap < n) {
        for (void** b : { (void**)&ctx->dHits, (void**)&ctx->dSurf, (void**)&ctx->dShade, (void**)&ctx->dAo,
                          (void**)&ctx->dColor }) {
            if (*b)
                (void)hipFree(*b);
            *b = nullptr;
        }
        ctx->hitCap = 0;
        SVDAG_HIP(hipMalloc(&ctx->dHits, n * sizeof(svdag_hit)));
        SVDAG_HIP(hipMalloc(&ctx->dSurf, n));
        SVDAG_HIP(hipMalloc(&ctx->dShade, n * sizeof(float)));
        SVDAG_HIP(hipMalloc(&ctx->dAo, n * sizeof(float)));
        SVDAG_HIP(hipMalloc(&ctx->dColor, n * sizeof(uint32_t)));
        ctx->hitCap = n;
    }
    ctx->traceW = ctx->traceH = 0;
    ctx->sunValid = ctx->aoValid = ctx->colorValid = false;
    {
        const TraceParams tp = trace_params(*cam, ctx->dag.levels, W, H);
        for (int k = 0; k < 3; ++k) {
            ctx->traceCam[k] = tp.cam[k];
            ctx->traceCam[3 + k] = tp.rayMin[k];
            ctx->traceCam[6 + k] = tp.ddx[k];
            ctx->traceCam[9 + k] = tp.ddy[k];
        }
    }
    SVDAG_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    SVDAG_TRY(launch_trace(ctx, *cam, W, H, ctx->root, ctx->stream, (svdag_hit*)ctx->dHits, ctx->dSurf));
    SVDAG_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    SVDAG_HIP(hipMemcpyAsync(out, ctx->dHits, n * sizeof(svdag_hit), hipMemcpyDeviceToHost, ctx->stream));
    SVDAG_HIP(hipStreamSynchronize(ctx->stream));
    ctx->traceW = W;
    ctx->traceH = H;
    ctx->traceRoot = ctx->root;
    if (ms)
        SVDAG_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return SVDAG_OK;
}

// Sun direction of the reference (tracer.cu:600-603): normalize(float3(0.3f, 1, 0.5f))
// with normalize(a) = (1 / length(a)) * a and length(a) = sqrt(dot(a, a)), all
// in float (cuda_math.h:115-118).  The operands are constants, which the
// compiler folds operation by operation, so nothing is fused.
static float3 sun_direction()
{
    const float v[3] = { 0.3f, 1.0f, 0.5f };
    const float dot = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    const float inv = 1.0f / std::sqrt(dot);
    return make_float3(inv * v[0], inv * v[1], inv * v[2]);
}

static int shade_pass(svdag_ctx* ctx, int kind, uint32_t numSamples, float bias, float rayLength, uint64_t seed,
                      float* out, float* ms)
{
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (ctx->traceW == 0 || !ctx->dShade) {
        set_error("shadows / ambient occlusion need a primary trace first (svdag_trace_primary)");
        return SVDAG_E_INVALID;
    }
    const uint32_t n = ctx->traceW * ctx->traceH;
    const uint32_t grid = (n + kTraceBlock - 1) / kTraceBlock;
    float* dst = kind == 0 ? ctx->dShade : ctx->dAo;
    SVDAG_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if (kind == 0)
        shadow_kernel<<<grid, kTraceBlock, 0, ctx->stream>>>(ctx->dag, ctx->traceRoot, (const svdag_hit*)ctx->dHits,
                                                             ctx->dSurf, n, bias, sun_direction(), dst);
    else
        ao_kernel<<<grid, kTraceBlock, 0, ctx->stream>>>(ctx->dag, ctx->traceRoot, (const svdag_hit*)ctx->dHits,
                                                         ctx->dSurf, n, numSamples, bias, rayLength, seed, dst);
    SVDAG_HIP(hipGetLastError());
    SVDAG_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    SVDAG_HIP(hipMemcpyAsync(out, dst, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    SVDAG_HIP(hipStreamSynchronize(ctx->stream));
    (kind == 0 ? ctx->sunValid : ctx->aoValid) = true;
    if (ms)
        SVDAG_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return SVDAG_OK;
}

extern "C" int svdag_trace_shadows(svdag_ctx* ctx, float shadow_bias, float* out, float* ms)
{
    if (!ctx || !out || !std::isfinite(shadow_bias)) {
        set_error("svdag_trace_shadows: bad arguments");
        return SVDAG_E_INVALID;
    }
    return shade_pass(ctx, 0, 0, shadow_bias, 0.0f, 0, out, ms);
}

extern "C" int svdag_trace_ambient_occlusion(svdag_ctx* ctx, uint32_t num_samples, float shadow_bias,
                                             float ao_ray_length, uint64_t seed, float* out, float* ms)
{
    if (!ctx || !out || num_samples == 0 || num_samples > 4096 || !std::isfinite(shadow_bias)
        || !(ao_ray_length > 0.0f)) {
        set_error("svdag_trace_ambient_occlusion: bad arguments");
        return SVDAG_E_INVALID;
    }
    return shade_pass(ctx, 1, num_samples, shadow_bias, ao_ray_length, seed, out, ms);
}

extern "C" int svdag_trace_colors(svdag_ctx* ctx, const float* material_rgb, uint32_t num_materials, uint32_t* out,
                                  float* ms)
{
    if (!ctx || (num_materials && !material_rgb) || num_materials > 65536) {
        set_error("svdag_trace_colors: bad arguments");
        return SVDAG_E_INVALID;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (ctx->traceW == 0 || !ctx->dColor) {
        set_error("svdag_trace_colors needs a primary trace first (svdag_trace_primary)");
        return SVDAG_E_INVALID;
    }
    const uint32_t n = ctx->traceW * ctx->traceH;
    float* dPal = nullptr;
    SVDAG_HIP(hipMalloc(&dPal, std::max<size_t>(num_materials, 1) * 3 * sizeof(float)));
    struct Free {
        float* p;
        ~Free() { (void)hipFree(p); }
    } fr { dPal };
    if (num_materials)
        SVDAG_HIP(hipMemcpyAsync(dPal, material_rgb, (size_t)num_materials * 3 * sizeof(float), hipMemcpyHostToDevice,
                                 ctx->stream));
    SVDAG_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    colors_kernel<<<(n + 255) / 256, 256, 0, ctx->stream>>>((const svdag_hit*)ctx->dHits, n, dPal, num_materials,
                                                            ctx->dColor);
    SVDAG_HIP(hipGetLastError());
    SVDAG_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    if (out)
        SVDAG_HIP(hipMemcpyAsync(out, ctx->dColor, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    SVDAG_HIP(hipStreamSynchronize(ctx->stream));
    ctx->colorValid = true;
    if (ms)
        SVDAG_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return SVDAG_OK;
}

extern "C" int svdag_trace_lighting(svdag_ctx* ctx, const svdag_lighting* params, uint32_t* out, float* ms)
{
    if (!ctx || !params || !out || !std::isfinite(params->fog_density) || !std::isfinite(params->sun_brightness)) {
        set_error("svdag_trace_lighting: bad arguments");
        return SVDAG_E_INVALID;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (ctx->traceW == 0 || !ctx->colorValid) {
        set_error("svdag_trace_lighting needs a primary trace and its colour pass first (svdag_trace_colors)");
        return SVDAG_E_INVALID;
    }
    if ((params->apply_ambient_occlusion && !ctx->aoValid) || (params->apply_shadows && !ctx->sunValid)) {
        set_error("svdag_trace_lighting: the occlusion / shadow pass it applies has not run on this trace");
        return SVDAG_E_INVALID;
    }
    const uint32_t n = ctx->traceW * ctx->traceH;
    LightingDev p {};
    for (int k = 0; k < 3; ++k) {
        p.cam[k] = ctx->traceCam[k];
        p.rayMin[k] = ctx->traceCam[3 + k];
        p.ddx[k] = ctx->traceCam[6 + k];
        p.ddy[k] = ctx->traceCam[9 + k];
    }
    const float3 sd = sun_direction();
    p.sunDir[0] = sd.x;
    p.sunDir[1] = sd.y;
    p.sunDir[2] = sd.z;
    p.sunBrightness = params->sun_brightness;
    p.fogDensity = params->fog_density;
    p.width = ctx->traceW;
    p.applyAo = params->apply_ambient_occlusion ? 1u : 0u;
    p.applySun = params->apply_shadows ? 1u : 0u;
    uint32_t* dOut = nullptr;
    SVDAG_HIP(hipMalloc(&dOut, (size_t)n * 4));
    struct Free {
        uint32_t* p;
        ~Free() { (void)hipFree(p); }
    } fr { dOut };
    SVDAG_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    lighting_kernel<<<(n + 255) / 256, 256, 0, ctx->stream>>>((const svdag_hit*)ctx->dHits, n, ctx->dColor, ctx->dAo,
                                                              ctx->dShade, p, dOut);
    SVDAG_HIP(hipGetLastError());
    SVDAG_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    SVDAG_HIP(hipMemcpyAsync(out, dOut, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    SVDAG_HIP(hipStreamSynchronize(ctx->stream));
    if (ms)
        SVDAG_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return SVDAG_OK;
}

extern "C" int svdag_blur_ambient_occlusion(svdag_ctx* ctx, float* out, float* ms)
{
    if (!ctx)
        return SVDAG_E_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (ctx->traceW == 0 || !ctx->aoValid) {
        set_error("svdag_blur_ambient_occlusion needs an ambient-occlusion pass on the current trace first");
        return SVDAG_E_INVALID;
    }
    const uint32_t W = ctx->traceW, H = ctx->traceH, n = W * H;
    if (W % 8 || H % 8) {
        set_error("svdag_blur_ambient_occlusion: width and height must be multiples of 8 (as BlurKernel::apply)");
        return SVDAG_E_INVALID;
    }
    float* dScratch = nullptr;
    SVDAG_HIP(hipMalloc(&dScratch, (size_t)n * sizeof(float)));
    struct Free {
        float* p;
        ~Free() { (void)hipFree(p); }
    } fr { dScratch };
    const BlurWeights b = blur_weights();
    SVDAG_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    blur_kernel<false><<<(n + 255) / 256, 256, 0, ctx->stream>>>(ctx->dAo, dScratch, W, H, b);
    blur_kernel<true><<<(n + 255) / 256, 256, 0, ctx->stream>>>(dScratch, ctx->dAo, W, H, b);
    SVDAG_HIP(hipGetLastError());
    SVDAG_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    if (out)
        SVDAG_HIP(hipMemcpyAsync(out, ctx->dAo, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    SVDAG_HIP(hipStreamSynchronize(ctx->stream));
    if (ms)
        SVDAG_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return SVDAG_OK;
}

extern "C" int svdag_trace_path_tracing(svdag_ctx* ctx, const svdag_path_tracing* params, const float* material_rgb,
                                        uint32_t num_materials, uint32_t* out, uint32_t* accumulated, float* ms)
{
    if (!ctx || !params || !out || (num_materials && !material_rgb) || num_materials > 65536
        || !std::isfinite(params->environment_brightness) || params->num_samples > 4096
        || params->max_path_depth > 64) {
        set_error("svdag_trace_path_tracing: bad arguments");
        return SVDAG_E_INVALID;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (ctx->traceW == 0 || (!params->integrate_pixel && !ctx->colorValid)) {
        set_error("svdag_trace_path_tracing needs a primary trace (and, without pixel integration, its colour pass)");
        return SVDAG_E_INVALID;
    }
    const uint32_t n = ctx->traceW * ctx->traceH;
    if (ctx->accumCap < n) {
        if (ctx->dAccum)
            (void)hipFree(ctx->dAccum);
        ctx->dAccum = nullptr;
        ctx->accumCap = 0;
        SVDAG_HIP(hipMalloc(&ctx->dAccum, (size_t)n * 3 * sizeof(float)));
        ctx->accumCap = n;
        ctx->accumValid = false;
    }
    // A new accumulation when asked for, or when the image, the traced root
    // or the parameters changed (the reference clears on any change of its
    // parameters or root, dag_tracer.cu:286-296).
    const bool same = params->accumulate && ctx->accumValid && ctx->accumW == ctx->traceW
        && ctx->accumH == ctx->traceH && ctx->accumRoot == ctx->traceRoot
        && std::memcmp(ctx->accumCam, ctx->traceCam, sizeof(ctx->traceCam)) == 0
        && ctx->accumEnv == params->environment_brightness && ctx->accumDepth == params->max_path_depth
        && ctx->accumIntegrate == (params->integrate_pixel ? 1u : 0u);
    if (!same) {
        SVDAG_HIP(hipMemsetAsync(ctx->dAccum, 0, (size_t)n * 3 * sizeof(float), ctx->stream));
        ctx->numAccumulated = 0;
    }
    float* dPal = nullptr;
    SVDAG_HIP(hipMalloc(&dPal, std::max<size_t>(num_materials, 1) * 3 * sizeof(float)));
    uint32_t* dOut = nullptr;
    if (hipMalloc(&dOut, (size_t)n * 4) != hipSuccess) {
        (void)hipFree(dPal);
        set_error("svdag_trace_path_tracing: out of device memory");
        return SVDAG_E_DEVICE;
    }
    struct Free {
        void* a;
        void* b;
        ~Free()
        {
            (void)hipFree(a);
            (void)hipFree(b);
        }
    } fr { dPal, dOut };
    if (num_materials)
        SVDAG_HIP(hipMemcpyAsync(dPal, material_rgb, (size_t)num_materials * 3 * sizeof(float), hipMemcpyHostToDevice,
                                 ctx->stream));
    PathDev p {};
    for (int k = 0; k < 3; ++k) {
        p.cam.cam[k] = ctx->traceCam[k];
        p.cam.rayMin[k] = ctx->traceCam[3 + k];
        p.cam.ddx[k] = ctx->traceCam[6 + k];
        p.cam.ddy[k] = ctx->traceCam[9 + k];
    }
    p.cam.width = ctx->traceW;
    p.cam.height = ctx->traceH;
    const float3 sd = sun_direction();
    p.sunDir[0] = sd.x;
    p.sunDir[1] = sd.y;
    p.sunDir[2] = sd.z;
    p.env = params->environment_brightness;
    p.numSamples = params->num_samples;
    p.maxDepth = params->max_path_depth;
    p.integrate = params->integrate_pixel ? 1u : 0u;
    p.numAccumulated = ctx->numAccumulated;
    // Past 256 accumulated paths the reference traces no more samples (the
    // image only re-averages, dag_tracer.cu:297-302).
    if (ctx->numAccumulated > 256)
        p.numSamples = 0;
    p.numColors = num_materials;
    p.seed = params->seed;
    SVDAG_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    path_kernel<<<(n + kTraceBlock - 1) / kTraceBlock, kTraceBlock, 0, ctx->stream>>>(
        ctx->dag, ctx->traceRoot, p, (const svdag_hit*)ctx->dHits, ctx->dSurf, ctx->dColor, dPal, ctx->dAccum, dOut);
    SVDAG_HIP(hipGetLastError());
    SVDAG_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    SVDAG_HIP(hipMemcpyAsync(out, dOut, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    SVDAG_HIP(hipStreamSynchronize(ctx->stream));
    ctx->numAccumulated += p.numSamples;
    ctx->accumValid = true;
    ctx->accumW = ctx->traceW;
    ctx->accumH = ctx->traceH;
    ctx->accumRoot = ctx->traceRoot;
    std::memcpy(ctx->accumCam, ctx->traceCam, sizeof(ctx->traceCam));
    ctx->accumEnv = params->environment_brightness;
    ctx->accumDepth = params->max_path_depth;
    ctx->accumIntegrate = p.integrate;
    if (accumulated)
        *accumulated = ctx->numAccumulated;
    if (ms)
        SVDAG_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return SVDAG_OK;
}
