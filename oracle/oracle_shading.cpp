// ORACLE - test infrastructure only.  Nothing in the product links or calls
// this code; only tests/ load liboracle.so, as the checker.
//
// CPU restatement of the passes that follow the primary trace and the
// shadow / occlusion passes, for a DAG with materials:
//   colour   Tracer::trace_colors -> computeColorAtSurface (tracer.cu:413-596)
//            with each material's textures one colour; invalid material ->
//            the checker (1, b, 1 - b), b = (x ^ y ^ z) & 1; no surface -> 0;
//   blur     BlurKernel(4) (blur.cu:8-85, dag_tracer.cu:19, 189-196): 9
//            weights gaussian(4.5, sigma, x), sigma = 0.3 * 3 + 0.8 used as
//            the variance, divided by their std::reduce sum; horizontal then
//            vertical, taps outside the image skipped, sum += w * v;
//   lighting Tracer::trace_lighting + applyFog (tracer.cu:693-750): sky where
//            nothing was hit, else (ao * 0.5 * sky + sun * sunColor *
//            brightness) * Kd, Kd = the colour made linear; optional fog;
//            made sRGB;
//   ColorUtils::float3_to_rgb888, rgb888_to_float3, accurateSRGBToLinear,
//   accurateLinearToSRGB (color_utils.h:31-45, 233-252); clamp (cuda_math.h:44).
// Where nvcc's default contraction has exactly one product-then-add to fuse
// (the sRGB encode, the blur's running sum, the fog ray's two steps) it is
// written as std::fma; everything else is rounded operation by operation.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numbers>
#include <numeric>
#include <vector>

extern "C" void ora_sun_dir(float* L);

namespace {

struct Rec {
    uint32_t x, y, z, material;
    float t, px, py, pz;
};

float clamp01(float f) { return f < 0.0f ? 0.0f : (f > 1.0f ? 1.0f : f); }

uint32_t rgb888(const float c[3])
{
    return (uint32_t)(clamp01(c[0]) * 255.0f) | ((uint32_t)(clamp01(c[1]) * 255.0f) << 8)
           | ((uint32_t)(clamp01(c[2]) * 255.0f) << 16) | 0xff000000u;
}

float srgb_to_linear(float c)
{
    const float lo = c / 12.92f;
    const float hi = std::pow((c + 0.055f) / 1.055f, 2.4f);
    return c <= 0.04045f ? lo : hi;
}

float linear_to_srgb(float c)
{
    const float lo = c * 12.92f;
    const float hi = std::fma(std::pow(std::fabs(c), 1.0f / 2.4f), 1.055f, -0.055f);
    return c <= 0.0031308f ? lo : hi;
}

const float kSky[3] = { 187.0f / 255.0f, 242.0f / 255.0f, 250.0f / 255.0f };
const float kSunColor[3] = { 0.95f, 0.94f, 0.47f };

Rec rec_at(const uint32_t* hits, size_t i)
{
    Rec r;
    std::memcpy(&r, hits + i * 8, sizeof(r));
    return r;
}

bool no_surface(const Rec& h) { return h.x == 0 && h.y == 0 && h.z == 0; }

void blur_weights(float* w)
{
    const uint32_t half = 4;
    const float sigma = 0.3f * (float)(half - 1) + 0.8f;
    const float mean = (float)half + 0.5f;
    std::vector<float> k(2 * half + 1);
    for (size_t x = 0; x < k.size(); ++x) {
        const float norm = 1.0f / std::sqrt(2 * std::numbers::pi_v<float> * sigma);
        const float d = (float)x - mean;
        k[x] = norm * std::exp(-0.5f * d * d / sigma);
    }
    const float sum = std::reduce(k.begin(), k.end(), 0.0f, std::plus<float>());
    for (size_t x = 0; x < k.size(); ++x)
        w[x] = k[x] / sum;
}

} // namespace

extern "C" {

float ora_srgb_to_linear(float c) { return srgb_to_linear(c); }
float ora_linear_to_srgb(float c) { return linear_to_srgb(c); }
uint32_t ora_rgb888(const float* c) { return rgb888(c); }
void ora_blur_weights(float* w9) { blur_weights(w9); }

void ora_colors(const uint32_t* hits, uint32_t n, const float* palette, uint32_t numColors, uint32_t* out)
{
    for (uint32_t i = 0; i < n; ++i) {
        const Rec h = rec_at(hits, i);
        float c[3] = { 0.0f, 0.0f, 0.0f };
        if (!no_surface(h)) {
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
        out[i] = rgb888(c);
    }
}

void ora_blur(const float* in, uint32_t W, uint32_t H, float* out)
{
    float w[9];
    blur_weights(w);
    std::vector<float> mid((size_t)W * H);
    for (int pass = 0; pass < 2; ++pass) {
        const float* src = pass == 0 ? in : mid.data();
        float* dst = pass == 0 ? mid.data() : out;
        for (uint32_t y = 0; y < H; ++y)
            for (uint32_t x = 0; x < W; ++x) {
                float sum = 0.0f;
                for (int d = -4; d <= 4; ++d) {
                    const int nx = pass == 0 ? (int)x + d : (int)x;
                    const int ny = pass == 0 ? (int)y : (int)y + d;
                    if (nx < 0 || ny < 0 || nx >= (int)W || ny >= (int)H)
                        continue;
                    sum = std::fma(w[d + 4], src[(size_t)ny * W + nx], sum);
                }
                dst[(size_t)y * W + x] = sum;
            }
    }
}

// ao / sun may be null (not applied).  The camera is the primary trace's:
// get_trace_params (dag_tracer.cu:59-101) in double.
void ora_lighting(const uint32_t* hits, uint32_t W, uint32_t H, const uint32_t* colors, const float* ao,
                  const float* sun, const double* position, const double* rotation, double fovDeg,
                  const double* bmin, const double* bmax, uint32_t levels, float sunBrightness, float fogDensity,
                  uint32_t* out)
{
    const double right[3] = { -rotation[0], -rotation[1], -rotation[2] };
    const double up[3] = { rotation[3], rotation[4], rotation[5] };
    const double fwd[3] = { rotation[6], rotation[7], rotation[8] };
    const double fov = fovDeg / 2.0 * (double(M_PI) / 180.);
    const double aspect = double(W) / double(H);
    double cam[3], rayMin[3], ddx[3], ddy[3];
    for (int k = 0; k < 3; ++k) {
        const double X = right[k] * std::sin(fov) * aspect, Y = up[k] * std::sin(fov), Z = fwd[k] * std::cos(fov);
        const double bl = position[k] + Z - Y - X, br = position[k] + Z - Y + X, tl = position[k] + Z + Y - X;
        const double tr = -bmin[k], sc = double(1u << levels) / (bmax[k] - bmin[k]);
        cam[k] = (position[k] + tr) * sc;
        rayMin[k] = (bl + tr) * sc;
        ddx[k] = ((br + tr) * sc - rayMin[k]) * (1.0 / W);
        ddy[k] = ((tl + tr) * sc - rayMin[k]) * (1.0 / H);
    }
    float L[3];
    ora_sun_dir(L);
    const uint32_t n = W * H;
    for (uint32_t i = 0; i < n; ++i) {
        const Rec h = rec_at(hits, i);
        float c[3];
        if (no_surface(h)) {
            for (int k = 0; k < 3; ++k)
                c[k] = kSky[k];
        } else {
            const float occl = ao ? ao[i] : 1.0f;
            const float vis = sun ? sun[i] : 1.0f;
            for (int k = 0; k < 3; ++k) {
                const float kd = srgb_to_linear((float)((colors[i] >> (8 * k)) & 0xFFu) / 255.0f);
                const float ambient = (occl * 0.5f) * kSky[k];
                const float direct = (vis * kSunColor[k]) * sunBrightness;
                c[k] = (ambient + direct) * kd;
            }
            if (fogDensity > 0.0f) {
                // The camera ray of stored pixel (x, y), as trace_lighting
                // rebuilds it (no row flip there).
                const uint32_t x = i % W, y = i / W;
                double d[3];
                for (int k = 0; k < 3; ++k)
                    d[k] = std::fma((double)y, ddy[k], std::fma((double)x, ddx[k], rayMin[k])) - cam[k];
                const double len = std::sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
                float dir[3], o[3];
                for (int k = 0; k < 3; ++k) {
                    dir[k] = (float)(d[k] * (1.0 / len));
                    o[k] = (float)cam[k];
                }
                const float dx = o[0] - h.px, dy = o[1] - h.py, dz = o[2] - h.pz;
                const float dist = std::sqrt((dx * dx + dy * dy) + dz * dz);
                const float fogAmount = 1.0f - std::exp(-dist * (fogDensity * 0.00001f));
                const float sd = (dir[0] * L[0] + dir[1] * L[1]) + dir[2] * L[2];
                const float sunAmount = 1.01f * (sd > 0.0f ? sd : 0.0f);
                const float w = std::pow(sunAmount, 30.0f);
                const float f = clamp01(fogAmount);
                for (int k = 0; k < 3; ++k) {
                    const float fogColor = kSky[k] * (1.0f - w) + 1.0f * w;
                    c[k] = c[k] * (1.0f - f) + fogColor * f;
                }
            }
        }
        float s[3];
        for (int k = 0; k < 3; ++k)
            s[k] = linear_to_srgb(c[k]);
        out[i] = rgb888(s);
    }
}

} // extern "C"
