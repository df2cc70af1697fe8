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

}  // namespace

hipError_t env_cdf_build(const float4* env, uint32_t W, uint32_t H, float* marg, float* cond, float* row_sum,
                         float* integral_dev, hipStream_t s) {
    if (W == 0 || H == 0) return hipSuccess;
    k_env_rows<<<(H + 63) / 64, 64, 0, s>>>(env, W, H, cond, row_sum);
    k_env_marginal<<<1, 64, 0, s>>>(row_sum, H, marg, integral_dev);
    return hipGetLastError();
}

}  // namespace ohpt
