// ORACLE — TEST INFRASTRUCTURE ONLY.
// C entry points over the reference's own CPU sampler and env-CDF sources
// (ohao/render/rt/{env_cdf,sobol_generator,owen_scramble}.cpp), compiled
// from the reference tree by oracle/Makefile into oracle/_ref/. Used only to
// pin the oracle's restatement of those functions.
#include "render/rt/env_cdf.hpp"
#include "render/rt/owen_scramble.hpp"
#include "render/rt/sobol_generator.hpp"

#include <cstring>

extern "C" {

float ref_sobol_sample1d(uint32_t index, uint32_t dim) { return ohao::SobolGenerator::sample1D(index, dim); }

void ref_sobol_dirs(uint32_t* out) {
    std::memcpy(out, ohao::SobolGenerator::directionNumbers(), 4 * 32 * sizeof(uint32_t));
}

uint32_t ref_owen(uint32_t v, uint32_t seed) { return ohao::owenScramble(v, seed); }

void ref_env_cdf(const float* rgba, int W, int H, float* marg, float* cond, float* integral) {
    ohao::EnvCDF cdf;
    cdf.build(rgba, W, H);
    std::memcpy(marg, cdf.marginalCDF().data(), cdf.marginalCDF().size() * sizeof(float));
    std::memcpy(cond, cdf.conditionalCDF().data(), cdf.conditionalCDF().size() * sizeof(float));
    *integral = cdf.integral();
}

}  // extern "C"
