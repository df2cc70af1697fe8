// Environment importance-sampling tables on the GPU (EnvCDF::build,
// ohao/render/rt/env_cdf.cpp:13-61): per-row prefix sums of
// luminance * sin(theta_row_centre), normalised, plus the marginal CDF over
// rows and the integral. Each row is summed sequentially in column order, as
// the reference does, so the tables differ from it only through sinf.
#pragma clang fp contract(off)

#include "ohpt_internal.h"

namespace ohpt {
namespace {

__global__ __launch_bounds__(64) void k_env_rows(const float4* __restrict__ env, uint32_t W, uint32_t H,
                                                 float* __restrict__ cond, float* __restrict__ row_sum) {
    const uint32_t y = blockIdx.x * 64 + threadIdx.x;
    if (y >= H) return;
    const float kPiF = 3.14159265358979323846f;
    const float theta = kPiF * ((float)y + 0.5f) / (float)H;
    const float s = sinf(theta);
    const float4* row = env + (size_t)y * W;
    float* c = cond + (size_t)y * W;
    float acc = 0.0f;
    for (uint32_t x = 0; x < W; ++x) {
        const float4 p = row[x];
        const float l = 0.2126f * p.x + 0.7152f * p.y + 0.0722f * p.z;
        acc += l * s;
        c[x] = acc;
    }
    if (acc > 0.0f) {
        for (uint32_t x = 0; x < W; ++x) c[x] /= acc;
    } else {
        for (uint32_t x = 0; x < W; ++x) c[x] = (float)(x + 1) / (float)W;
    }
    row_sum[y] = acc;
}

__global__ void k_env_marginal(const float* __restrict__ row_sum, uint32_t H, float* __restrict__ marg,
                               float* __restrict__ integral) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float total = 0.0f;
    for (uint32_t y = 0; y < H; ++y) { total += row_sum[y]; marg[y] = total; }
    *integral = total;
    if (total > 0.0f) {
        for (uint32_t y = 0; y < H; ++y) marg[y] /= total;
    } else {
        for (uint32_t y = 0; y < H; ++y) marg[y] = (float)(y + 1) / (float)H;
    }
}

// Guide tables for CDF inversion (the cutpoint method): for a CDF of n
// entries and M = 2^log2m cells, guide[k] = first index i with cdf[i] >= k/M
// (k/M is exact), guide[M] likewise for 1. A lookup of u in cell
// k = floor(u*M) (exact: M is a power of two) then searches only
// [guide[k], guide[k+1]], which holds the answer of the full search, so the
// sampled texel is the same as with the full binary search.
__global__ __launch_bounds__(256) void k_env_guide(const float* __restrict__ cdf, uint32_t n, uint32_t rows,
                                                   uint32_t log2m, uint32_t* __restrict__ guide) {
    const uint32_t m = 1u << log2m;
    const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (uint64_t)rows * (m + 1)) return;
    const uint32_t row = (uint32_t)(t / (m + 1)), k = (uint32_t)(t % (m + 1));
    guide[t] = cdf_lower_bound(cdf + (size_t)row * n, n, ldexpf((float)k, -(int)log2m));
}

__global__ __launch_bounds__(256) void k_env_sample(DevScene S, const float* __restrict__ u, float4* __restrict__ out,
                                                    uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    f3 d = mk(0.0f, 0.0f, 0.0f);
    float pdf = 0.0f;
    if (!env_sample(S, u[2 * i], u[2 * i + 1], d, pdf)) pdf = 0.0f;
    out[i] = make_float4(d.x, d.y, d.z, pdf);
}

}  // namespace

void launch_env_sample(const DevScene& S, const float* u, float4* out, uint32_t n, hipStream_t s) {
    if (n == 0 || S.env_w == 0) return;
    k_env_sample<<<(n + 255) / 256, 256, 0, s>>>(S, u, out, n);
}

uint32_t env_guide_log2(uint32_t n) {
    uint32_t l = 0;
    while ((2u << l) <= n && l < 16) ++l;   // largest power of two <= n, at most 2^16 cells
    return l;
}

hipError_t env_cdf_build(const float4* env, uint32_t W, uint32_t H, float* marg, float* cond, float* row_sum,
                         float* integral_dev, uint32_t* marg_guide, uint32_t* cond_guide, hipStream_t s) {
    if (W == 0 || H == 0) return hipSuccess;
    k_env_rows<<<(H + 63) / 64, 64, 0, s>>>(env, W, H, cond, row_sum);
    k_env_marginal<<<1, 64, 0, s>>>(row_sum, H, marg, integral_dev);
    const uint32_t lm = env_guide_log2(H), lc = env_guide_log2(W);
    const uint64_t nm = (1ull << lm) + 1, nc = ((1ull << lc) + 1) * H;
    k_env_guide<<<(uint32_t)((nm + 255) / 256), 256, 0, s>>>(marg, H, 1u, lm, marg_guide);
    k_env_guide<<<(uint32_t)((nc + 255) / 256), 256, 0, s>>>(cond, W, H, lc, cond_guide);
    return hipGetLastError();
}

}  // namespace ohpt
