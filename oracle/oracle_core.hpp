// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the path-tracing core, used by tests/, the smoke check and
// bench.py's cpu_baseline leg as the checker and the CPU baseline. Nothing in
// the product links, loads or calls this code.
//
// Parity pins per function (what checks each part; see tests/test_oracle_*.py):
//   pcg32                      O'Neill 2014 reference output vectors (pcg32-demo, seed 42/54)
//   sobol_int / owen / hashPixel  reference KATs tests/renderer/sobol_test.cpp:10-112 and
//                              bit-exact vs the reference C++ built into oracle/_ref
//   env_cdf_build              bit-exact vs ohao/render/rt/env_cdf.cpp (oracle/_ref) and the
//                              properties of tests/renderer/env_cdf_test.cpp:9-61
//   env_sample / env_pdf       quadrature: integral of pdf over the sphere = 1; histogram test
//   lbvh_build                 Karras invariants, contiguous leaf ranges, box containment
//   trace (FP32 LBVH)          equals FP64 brute force except listed t-ties
//   bsdf eval/sample/pdf       quadrature of the pdf = 1; VNDF histogram; reciprocity
//   integrator                 closed forms: diffuse/GGX furnace, emissive enclosure
//                              L = Le * sum_{k<=D} rho^k
//
// Written against the reference's semantics (citations are paths in the
// reference tree); every floating-point operation in the traversal is spelled
// out in the same order as the HIP kernels so the FP32 traversal can be
// compared bit for bit. Build with -ffp-contract=off.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

#include "../include/ohpt.h"

namespace oracle {

// ---------------------------------------------------------------- samplers
// PCG32, M.E. O'Neill 2014 (pcg-random.org, pcg32_srandom_r / pcg32_random_r).
struct Pcg32 {
    uint64_t state = 0, inc = 1;
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((0u - rot) & 31u));
    }
    void seed(uint64_t initstate, uint64_t initseq) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
};

// splitmix64 finaliser (Steele, Lea, Flood 2014) for per-(sample, vertex) state.
inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

inline float u32_to_unit(uint32_t v) { return (float)(v >> 8) * (1.0f / 16777216.0f); }

// Sobol direction numbers (Joe-Kuo new-joe-kuo-6.21201), dims 0..3,
// as tabulated in ohao/render/rt/sobol_generator.cpp:14-54.
extern const uint32_t kSobolDirs[4 * 32];

// sobolIntForIndex, sobol_generator.cpp:59-68: XOR of direction numbers of the set bits.
inline uint32_t sobol_int(uint32_t index, uint32_t dim) {
    uint32_t r = 0;
    const uint32_t* d = kSobolDirs + dim * 32;
    for (uint32_t bit = 0; index != 0; ++bit, index >>= 1)
        if (index & 1u) r ^= d[bit];
    return r;
}

inline uint32_t reverse_bits(uint32_t v) {
    v = (v << 16) | (v >> 16);
    v = ((v & 0x00FF00FFu) << 8) | ((v & 0xFF00FF00u) >> 8);
    v = ((v & 0x0F0F0F0Fu) << 4) | ((v & 0xF0F0F0F0u) >> 4);
    v = ((v & 0x33333333u) << 2) | ((v & 0xCCCCCCCCu) >> 2);
    v = ((v & 0x55555555u) << 1) | ((v & 0xAAAAAAAAu) >> 1);
    return v;
}

// Burley 2020 hash-based Owen scramble: owen_scramble.cpp:12-31,
// sampler_sobol.glsl:13-36 (laine-karras style hash on reversed bits).
inline uint32_t owen_scramble(uint32_t v, uint32_t seed) {
    v = reverse_bits(v);
    v ^= v * 0x3d20adeau;
    v += seed;
    v *= (seed >> 16) | 1u;
    v ^= v * 0x05526c56u;
    v ^= v * 0x53a22864u;
    return reverse_bits(v);
}

// Pixel hash: sampler_sobol.glsl:51-58 (murmur3 finaliser).
inline uint32_t hash_pixel(uint32_t x, uint32_t y) {
    uint32_t h = (x * 0x1b873593u) ^ (y * 0xcc9e2d51u);
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

// getSample1D_sobol, sampler_sobol.glsl:69-77: padded 4D Sobol, per-pad Owen seed.
inline float sobol_owen_sample(uint32_t px, uint32_t py, uint32_t index, uint32_t dim) {
    uint32_t pad = dim >> 2, local = dim & 3u;
    uint32_t seed = hash_pixel(px, py) ^ (pad * 0x9e3779b9u);
    return u32_to_unit(owen_scramble(sobol_int(index, local), seed));
}

// Dimension layout of one path sample (shared with the HIP kernels by spec):
//   dims 0,1          camera jitter
//   dims 2+9b+k       vertex b: k=0 light pick, 1-2 light point, 3-4 env,
//                     5 russian roulette, 6 lobe, 7-8 direction
constexpr uint32_t kDimsPerVertex = 9;

struct VertexSamples { float u[kDimsPerVertex]; };

inline void draw_camera(uint32_t sampler, uint32_t px, uint32_t py, uint32_t width,
                        uint32_t sample_index, float out[2]) {
    if (sampler == OHPT_SAMPLER_SOBOL_OWEN) {
        out[0] = sobol_owen_sample(px, py, sample_index, 0);
        out[1] = sobol_owen_sample(px, py, sample_index, 1);
    } else {
        Pcg32 r;
        r.seed(splitmix64(((uint64_t)sample_index << 16) | 0u), (uint64_t)py * width + px);
        out[0] = u32_to_unit(r.next());
        out[1] = u32_to_unit(r.next());
    }
}

inline void draw_vertex(uint32_t sampler, uint32_t px, uint32_t py, uint32_t width,
                        uint32_t sample_index, uint32_t vertex, VertexSamples& vs) {
    if (sampler == OHPT_SAMPLER_SOBOL_OWEN) {
        for (uint32_t k = 0; k < kDimsPerVertex; ++k)
            vs.u[k] = sobol_owen_sample(px, py, sample_index, 2 + kDimsPerVertex * vertex + k);
    } else {
        Pcg32 r;
        r.seed(splitmix64(((uint64_t)sample_index << 16) | (uint64_t)(vertex + 1u)),
               (uint64_t)py * width + px);
        for (uint32_t k = 0; k < kDimsPerVertex; ++k) vs.u[k] = u32_to_unit(r.next());
    }
}

// ---------------------------------------------------------------- env CDF
// EnvCDF::build, ohao/render/rt/env_cdf.cpp:13-61, same loop order and types.
void env_cdf_build(const float* rgba, int W, int H, std::vector<float>& marg,
                   std::vector<float>& cond, float& integral);

// ---------------------------------------------------------------- LBVH
struct BvhNode {            // 64 B, identical to the device layout
    float b[12];            // child0 lo xyz, child0 hi xyz, child1 lo xyz, child1 hi xyz
    uint32_t child[2];      // bit31 set = leaf (sorted triangle slot)
    uint32_t parent;
    uint32_t pad;
};
static_assert(sizeof(BvhNode) == 64, "node layout");
constexpr uint32_t kLeafBit = 0x80000000u;

// 4-wide node (128 B), identical to the device layout: the binary LBVH
// collapsed two levels at a time.
struct Bvh4Node {
    float lox[4], loy[4], loz[4], hix[4], hiy[4], hiz[4];   // child boxes
    uint32_t child[4];      // bit31 set = leaf (sorted triangle slot); kEmpty = unused slot
    uint32_t bin_node;      // binary internal node this node was collapsed from
    uint32_t count;         // used slots
    uint32_t pad[2];
};
static_assert(sizeof(Bvh4Node) == 128, "node4 layout");
constexpr uint32_t kEmpty = 0xFFFFFFFFu;

struct Bvh {
    uint32_t n = 0;
    std::vector<BvhNode> nodes;      // n-1 internal nodes (1 when n == 1)
    std::vector<Bvh4Node> nodes4;    // collapsed tree traversed by trace/occlusion queries
    std::vector<float> tri;          // sorted triangles, 12 floats each (v0 xyz, prim, v1 xyz, 0, v2 xyz, 0)
    std::vector<uint32_t> prim;      // sorted slot -> original triangle id
    std::vector<uint64_t> keys;      // morton << 32 | prim, sorted
};

uint32_t morton30(float cx, float cy, float cz, const float cmin[3], const float cmax[3]);
void lbvh_build(const float* tri_pos /* n*9 */, uint32_t n, Bvh& out);   // binary LBVH + 4-wide collapse
// Refit: same topology and leaf order, triangles re-read from tri_pos (caller's
// order), boxes recomputed bottom-up, 4-wide nodes re-collapsed.
void lbvh_refit(const float* tri_pos /* n*9 */, uint32_t n, Bvh& bvh);
// Collapse (called by lbvh_build): every internal node at even depth becomes a
// 4-wide node, numbered in binary-node order; its slots are its binary
// children in order, an internal child replaced by that child's two children.
void bvh4_collapse(Bvh& bvh);

struct HitF32 { float t; uint32_t prim; float u, v; };
// FP32 closest hit / occlusion through the 4-wide tree, op-for-op as the HIP traversal kernels.
// One ray against one triangle (v: 3 vertices x xyz): the same FP32 watertight test.
bool tri_hit_f32(const float o[3], const float d[3], const float v[9], float tmin, float tmax, float tuv[3]);
HitF32 trace_f32(const Bvh& bvh, const float o[3], const float d[3], float tmin, float tmax,
                 uint32_t* node_visits, uint32_t* tri_tests);
bool occluded_f32(const Bvh& bvh, const float o[3], const float d[3], float tmin, float tmax);

// FP64 watertight test over all triangles (brute force). Reports the nearest t
// and every triangle whose t lies within rel_tie of it.
struct HitF64 { double t; uint32_t prim; double u, v; uint32_t n_ties; };
HitF64 brute_closest_f64(const float* tri_pos, uint32_t n, const double o[3], const double d[3],
                         double tmin, double tmax, double rel_tie, uint32_t* ties, uint32_t max_ties);

}  // namespace oracle
