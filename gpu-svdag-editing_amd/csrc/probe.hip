// Memory-system probes: the practical ceilings the slab-hash find is measured
// against.  A stream read gives the sequential HBM rate; a random-line read
// touches random 128-byte lines the way a find probe does (the tags and the
// link, in the line's first 32 bytes), so it gives the rate at which the
// memory system serves random lines.  No reference counterpart.
#include "svdag_internal.h"
#include "svdag_device.h"
#include "svdag.h"
#include <algorithm>

namespace svdag {

__global__ void __launch_bounds__(256) probe_stream(const uint4* p, uint64_t n16, uint32_t* out)
{
    uint32_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint4 v = p[i];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    for (int m = 32; m >= 1; m >>= 1)
        acc ^= (uint32_t)__shfl_xor((int)acc, m, 64);
    if ((threadIdx.x & 63u) == 0)
        out[(blockIdx.x * blockDim.x + threadIdx.x) >> 6] = acc;
}

__global__ void __launch_bounds__(256) probe_lines(const uint32_t* lines, uint32_t nLines, uint64_t reads, uint32_t* out)
{
    uint32_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < reads; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t line = mix32((uint32_t)i * 0x9E3779B1u + 0x7F4A7C15u) % nLines;
        const uint32_t* l = lines + (size_t)line * kLineWords;
        const uint4 v = *reinterpret_cast<const uint4*>(l);
        acc ^= v.x ^ v.y ^ v.z ^ v.w ^ l[4]; // a probe's line head: tags and link, one 32-byte sector
    }
    for (int m = 32; m >= 1; m >>= 1)
        acc ^= (uint32_t)__shfl_xor((int)acc, m, 64);
    if ((threadIdx.x & 63u) == 0)
        out[(blockIdx.x * blockDim.x + threadIdx.x) >> 6] = acc;
}

__global__ void probe_fill(uint4* p, uint64_t n16)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = make_uint4((uint32_t)i, (uint32_t)(i >> 32), 0x5A5A5A5Au, (uint32_t)i * 2654435761u);
}

} // namespace svdag

using namespace svdag;

extern "C" int svdag_hbm_probe(uint32_t kind, uint64_t bytes, uint64_t reads, float* ms)
{
    if (!ms || kind > SVDAG_PROBE_RANDOM_LINES || bytes < 4096 || (bytes >> 7) > 0xFFFFFFFFull
        || (kind == SVDAG_PROBE_RANDOM_LINES && reads == 0)) {
        set_error("svdag_hbm_probe: bad arguments");
        return SVDAG_E_INVALID;
    }
    SVDAG_TRY(require_device());
    bytes &= ~(uint64_t)127;
    const uint64_t n16 = bytes / 16;
    const uint32_t grid = 8192, block = 256;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    void* buf = nullptr;
    uint32_t* out = nullptr;
    int rc = SVDAG_OK;
    auto fail = [&](hipError_t e, const char* what) {
        set_error(std::string("svdag_hbm_probe: ") + what + ": " + hipGetErrorString(e));
        rc = SVDAG_E_DEVICE;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess)
        fail(e, "stream");
    else if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess)
        fail(e, "events");
    else if ((e = hipMalloc(&buf, bytes)) != hipSuccess)
        fail(e, "buffer");
    else if ((e = hipMalloc(&out, (size_t)grid * block / 64 * 4)) != hipSuccess)
        fail(e, "output");
    if (rc == SVDAG_OK) {
        probe_fill<<<grid, block, 0, st>>>(reinterpret_cast<uint4*>(buf), n16);
        float best = 1e30f;
        for (int it = 0; it < 6 && rc == SVDAG_OK; ++it) {
            (void)hipEventRecord(e0, st);
            if (kind == SVDAG_PROBE_STREAM)
                probe_stream<<<grid, block, 0, st>>>(reinterpret_cast<const uint4*>(buf), n16, out);
            else
                probe_lines<<<grid, block, 0, st>>>(reinterpret_cast<const uint32_t*>(buf), (uint32_t)(bytes >> 7), reads,
                                                    out);
            (void)hipEventRecord(e1, st);
            if ((e = hipEventSynchronize(e1)) != hipSuccess) {
                fail(e, "kernel");
                break;
            }
            float t = 0.0f;
            (void)hipEventElapsedTime(&t, e0, e1);
            if (it > 0) // the first launch warms up
                best = std::min(best, t);
        }
        *ms = best;
    }
    if (out)
        (void)hipFree(out);
    if (buf)
        (void)hipFree(buf);
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    if (st)
        (void)hipStreamDestroy(st);
    return rc;
}
