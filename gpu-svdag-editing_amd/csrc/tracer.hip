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
// operation with contraction off; the camera setup is double precision.
#include "dag_internal.h"
#include <cfloat>
#include <cmath>

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

__global__ void __launch_bounds__(kTraceBlock) trace_kernel(DagDev dag, uint32_t root, TraceParams p, svdag_hit* out)
{
    __shared__ uint2 stack[20][kTraceBlock]; // levels <= 20
    const uint32_t tid = threadIdx.x;
    const uint32_t tilesX = (p.width + 15) / 16;
    const uint32_t px = (blockIdx.x % tilesX) * 16 + (tid & 15u);
    const uint32_t py = (blockIdx.x / tilesX) * 16 + (tid >> 4);
    if (px >= p.width || py >= p.height)
        return;

    // Per-pixel ray in double, normalised, then rounded to float (tracer.cu:402-405).
    double a[3];
    for (int k = 0; k < 3; ++k)
        a[k] = ((p.rayMin[k] + (double)px * p.ddx[k]) + (double)py * p.ddy[k]) - p.cam[k];
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

    const uint32_t levels = dag.levels;
    const uint32_t leafLevel = dag.leafLevel;
    const uint32_t order = (ray.d[0] < 0.f ? 4u : 0u) + (ray.d[1] < 0.f ? 2u : 0u) + (ray.d[2] < 0.f ? 1u : 0u);
    uint32_t level = 0;
    uint32_t x = 0, y = 0, z = 0;
    uint32_t material = kSentinel;

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
                const uint2 e = stack[newLevel][tid];
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
            stack[level][tid] = make_uint2(cIndex, cChild | (cVisit << 8));
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

    // createSurfaceInteraction (tracer.cu:353-389): entry face of the hit voxel.
    const float bmin[3] = { (float)x, (float)y, (float)z };
    float dists[3];
    for (int k = 0; k < 3; ++k) {
        const float bmax = bmin[k] + 1.0f;
        const float t1 = (bmin[k] - ray.o[k]) * ray.inv[k];
        const float t2 = (bmax - ray.o[k]) * ray.inv[k];
        dists[k] = fmin_ref(t1, t2);
    }
    const int axis = dists[0] > dists[1] ? (dists[0] > dists[2] ? 0 : 2) : (dists[1] > dists[2] ? 1 : 2);
    const float t = dists[axis];
    svdag_hit h;
    h.x = x;
    h.y = y;
    h.z = z;
    h.material = material;
    h.t = t;
    h.px = ray.o[0] + t * ray.d[0];
    h.py = ray.o[1] + t * ray.d[1];
    h.pz = ray.o[2] + t * ray.d[2];
    out[(size_t)(p.height - 1 - py) * p.width + px] = h;
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
                 svdag_hit* outDev)
{
    const TraceParams p = trace_params(cam, ctx->dag.levels, W, H);
    const uint32_t tiles = ((W + 15) / 16) * ((H + 15) / 16);
    trace_kernel<<<tiles, kTraceBlock, 0, stream>>>(ctx->dag, root, p, outDev);
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
        if (ctx->dHits)
            (void)hipFree(ctx->dHits);
        ctx->dHits = nullptr;
        SVDAG_HIP(hipMalloc(&ctx->dHits, n * sizeof(svdag_hit)));
        ctx->hitCap = n;
    }
    SVDAG_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    SVDAG_TRY(launch_trace(ctx, *cam, W, H, ctx->root, ctx->stream, (svdag_hit*)ctx->dHits));
    SVDAG_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    SVDAG_HIP(hipMemcpyAsync(out, ctx->dHits, n * sizeof(svdag_hit), hipMemcpyDeviceToHost, ctx->stream));
    SVDAG_HIP(hipStreamSynchronize(ctx->stream));
    if (ms)
        SVDAG_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return SVDAG_OK;
}
