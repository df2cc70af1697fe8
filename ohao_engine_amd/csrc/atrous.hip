// Variance-guided a-trous wavelet filter (SVGF's spatial pass), the reference's
// DenoiseMode::Atrous comparator (AtrousDenoiser, ohao/render/rt/denoise/
// atrous_denoise.hpp:44-67; shaders/rt/rt_svgf_atrous.comp:42-114), run on the
// accumulated frame. Its inputs are what the reference's temporal pass
// (rt_svgf_temporal.comp:87-120) hands over for a still camera, one sample per
// frame:
//   * colour = the mean of the per-sample display values (ACES(x/2)^(1/2.2)),
//   * variance = E[l^2] - E[l]^2 of the per-sample display luminance l,
//     raised to the 3x3 spatial luminance variance while fewer than 4 samples
//     are in (the temporal pass's bootstrap, rt_svgf_temporal.comp:105-120);
//   * normal / depth = the first-hit AOVs (mean normal, mean hit distance
//     over the hits; background 1e30 when fewer than half the samples hit);
//   * 5 iterations, steps 1, 2, 4, 8, 16, B3-spline taps, the reference's
//     sigmaL 0.4, sigmaNormal 0.3, sigmaDepth 2.0 (atrous_denoise.cpp:38-40).
// The colour and moments are exact means over all samples where the
// reference's exponential average caps the history at 20 frames (alpha 0.05).
#include "ohpt_internal.h"

namespace ohpt {
namespace {

constexpr float kLw[3] = {0.2126f, 0.7152f, 0.0722f};

// Per-pixel records: cv = (display colour rgb, luminance variance), geo =
// (normal xyz, depth): one 16-byte load each per filter tap.
__global__ __launch_bounds__(256) void k_svgf_prep(const float4* __restrict__ acc, const float4* __restrict__ disp,
                                                   const float4* __restrict__ aa, const float4* __restrict__ an,
                                                   uint32_t n, float4* __restrict__ cv, float4* __restrict__ geo) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float cnt = acc[i].w;
    const float4 d = disp[i], al = aa[i], nm = an[i];
    const float inv = cnt > 0.0f ? 1.0f / cnt : 0.0f;
    const float cx = d.x * inv, cy = d.y * inv, cz = d.z * inv;
    const float m1 = kLw[0] * cx + kLw[1] * cy + kLw[2] * cz;
    cv[i] = make_float4(cx, cy, cz, fmaxf(0.0f, d.w * inv - m1 * m1));
    float nx = nm.x * inv, ny = nm.y * inv, nz = nm.z * inv;
    const float len = sqrtf(nx * nx + ny * ny + nz * nz);
    if (len > 0.0f) { nx /= len; ny /= len; nz /= len; }
    const float hits = nm.w;   // samples whose primary ray hit (the AOV hit flag summed)
    const float depth = (hits * 2.0f > cnt && hits > 0.0f) ? al.w / hits : 1e30f;
    geo[i] = make_float4(nx, ny, nz, depth);
}

__device__ __forceinline__ float luma(float4 c) { return kLw[0] * c.x + kLw[1] * c.y + kLw[2] * c.z; }

// Blocks of 64 x 4 pixels: a wave reads one row segment of 64 records per tap.
constexpr int kBX = 64, kBY = 4;

// Short accumulations (< 4 samples): the moment variance is unreliable, so the
// variance is at least the 3x3 luminance variance of the display image (taps
// clamped to the edge, 9 of them), rt_svgf_temporal.comp:105-120.
__global__ __launch_bounds__(256) void k_svgf_bootstrap(const float4* __restrict__ acc, const float4* __restrict__ cvin,
                                                        int W, int H, float4* __restrict__ cvout) {
    const int x = blockIdx.x * kBX + (threadIdx.x % kBX), y = blockIdx.y * kBY + (threadIdx.x / kBX);
    if (x >= W || y >= H) return;
    const int p = y * W + x;
    float4 c = cvin[p];
    if (acc[p].w < 4.0f) {
        float s = 0.0f, s2 = 0.0f;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int qx = min(max(x + dx, 0), W - 1), qy = min(max(y + dy, 0), H - 1);
                const float l = luma(cvin[qy * W + qx]);
                s += l;
                s2 += l * l;
            }
        const float mm = s / 9.0f;
        c.w = fmaxf(c.w, fmaxf(0.0f, s2 / 9.0f - mm * mm));
    }
    cvout[p] = c;
}

// One iteration, rt_svgf_atrous.comp:49-115. The three edge stops are one
// exponential of the summed exponents. The last iteration writes alpha 1.
__global__ __launch_bounds__(256) void k_svgf_atrous(const float4* __restrict__ cvin, const float4* __restrict__ geo,
                                                     int W, int H, int step, int last, float sigma_l, float sigma_n,
                                                     float sigma_d, float4* __restrict__ cvout) {
    const int x = blockIdx.x * kBX + (threadIdx.x % kBX), y = blockIdx.y * kBY + (threadIdx.x / kBX);
    if (x >= W || y >= H) return;
    const int p = y * W + x;
    const float4 cc = cvin[p], cg = geo[p];
    const float cl = luma(cc);
    // 3x3 Gaussian prefilter of the variance for the luminance edge stop (weights sum to 1)
    const float gk[3] = {0.25f, 0.5f, 0.25f};
    float cvar = 0.0f;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            const int qx = min(max(x + dx, 0), W - 1), qy = min(max(y + dy, 0), H - 1);
            cvar += gk[dx + 1] * gk[dy + 1] * cvin[qy * W + qx].w;
        }
    const float inv_l = 1.0f / (sqrtf(fmaxf(cvar, 1e-8f)) * sigma_l + 1e-6f);
    const float inv_n = 1.0f / (sigma_n * sigma_n + 1e-4f);
    const float inv_d = 1.0f / (sigma_d * sigma_d + 1e-4f);
    const float kb[5] = {1.0f / 16.0f, 4.0f / 16.0f, 6.0f / 16.0f, 4.0f / 16.0f, 1.0f / 16.0f};
    float sr = 0.0f, sg = 0.0f, sb = 0.0f, ws = 0.0f, vs = 0.0f;
#pragma unroll
    for (int dy = -2; dy <= 2; ++dy) {
        const int qy = min(max(y + dy * step, 0), H - 1);
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            const int q = qy * W + min(max(x + dx * step, 0), W - 1);
            const float4 sc = cvin[q], sgeo = geo[q];
            const float nd = fmaxf(1.0f - (cg.x * sgeo.x + cg.y * sgeo.y + cg.z * sgeo.z), 0.0f);
            const float e = fabsf(cl - luma(sc)) * inv_l + nd * inv_n + fabsf(cg.w - sgeo.w) * inv_d;
            const float w = kb[dx + 2] * kb[dy + 2] * expf(-e);
            sr += sc.x * w;
            sg += sc.y * w;
            sb += sc.z * w;
            ws += w;
            vs += w * w * sc.w;
        }
    }
    float4 o;
    if (ws > 1e-6f) {
        const float iw = 1.0f / ws;
        o = make_float4(sr * iw, sg * iw, sb * iw, vs * iw * iw);
    } else {
        o = make_float4(cc.x, cc.y, cc.z, cvar);
    }
    if (last) o.w = 1.0f;
    cvout[p] = o;
}

}  // namespace

hipError_t svgf_atrous(const float4* accum, const float4* disp, const float4* aov_albedo,
                       const float4* aov_normal, uint32_t W, uint32_t H, float4* out, void* scratch,
                       hipStream_t s) {
    const uint32_t n = W * H;
    if (n == 0) return hipSuccess;
    float4* c0 = reinterpret_cast<float4*>(scratch);
    float4* c1 = c0 + n;
    float4* geo = c1 + n;
    k_svgf_prep<<<(n + 255) / 256, 256, 0, s>>>(accum, disp, aov_albedo, aov_normal, n, c0, geo);
    const dim3 grid((W + kBX - 1) / kBX, (H + kBY - 1) / kBY);
    k_svgf_bootstrap<<<grid, 256, 0, s>>>(accum, c0, (int)W, (int)H, c1);
    // ping-pong c1 -> out -> c0 -> out -> c0 -> out: the odd count of 5 ends in `out`
    static_assert(kSvgfIterations % 2 == 1, "the ping-pong below ends in `out` for an odd count");
    const float4* ci = c1;
    for (int it = 0; it < kSvgfIterations; ++it) {
        float4* co = (it % 2 == 0) ? out : c0;
        k_svgf_atrous<<<grid, 256, 0, s>>>(ci, geo, (int)W, (int)H, 1 << it, it == kSvgfIterations - 1 ? 1 : 0,
                                           kSvgfSigmaL, kSvgfSigmaN, kSvgfSigmaD, co);
        ci = co;
    }
    return hipGetLastError();
}

size_t svgf_scratch_bytes(uint32_t W, uint32_t H) { return (size_t)W * H * 48; }

}  // namespace ohpt
