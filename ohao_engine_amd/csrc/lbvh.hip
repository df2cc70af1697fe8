// LBVH builder for gfx950: 30-bit Morton codes of triangle centroids, a stable
// LSD radix sort with LDS-local ranking (8-bit digits, wave64 ballot match),
// Karras 2012 hierarchy over (morton << 32 | prim) keys, and a bottom-up AABB
// refit with acquire/release counters. Replaces the driver-built BLAS/TLAS of
// ohao/render/rt/rt_acceleration_structure.cpp:205-309, :419-535.
//
// Every floating-point step is IEEE add/mul/div/min/max without contraction,
// so the tree (topology and boxes) is bit-identical to a sequential build.
#pragma clang fp contract(off)

#include "ohpt_internal.h"

namespace ohpt {
namespace {

constexpr uint32_t kRadixBlock = 256;
constexpr uint32_t kRadixItems = 16;
constexpr uint32_t kRadixTile = kRadixBlock * kRadixItems;
constexpr uint32_t kRadixDigits = 256;

__device__ __forceinline__ uint32_t expand_bits10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__device__ __forceinline__ uint32_t quantize10(float c, float lo, float hi) {
    const float ext = hi - lo;
    float x = 0.5f;
    if (ext > 0.0f) x = (c - lo) / ext;
    const float f = x * 1024.0f;
    if (!(f > 0.0f)) return 0u;
    if (f >= 1023.0f) return 1023u;
    return (uint32_t)f;
}

__device__ __forceinline__ void tri_bounds(const float* t, float lo[3], float hi[3]) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = fminf(fminf(t[a], t[3 + a]), t[6 + a]);
        hi[a] = fmaxf(fmaxf(t[a], t[3 + a]), t[6 + a]);
    }
}

// --- centroid bounds: per-block partial min/max, then one block folds them
__global__ __launch_bounds__(256) void k_centroid_bounds(const float* __restrict__ tp, uint32_t n,
                                                         float* __restrict__ partial) {
    __shared__ float smin[3][256], smax[3][256];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        float lo[3], hi[3];
        tri_bounds(tp + (size_t)i * 9, lo, hi);
        for (int a = 0; a < 3; ++a) {
            const float c = (lo[a] + hi[a]) * 0.5f;
            mn[a] = fminf(mn[a], c);
            mx[a] = fmaxf(mx[a], c);
        }
    }
    for (int a = 0; a < 3; ++a) { smin[a][threadIdx.x] = mn[a]; smax[a][threadIdx.x] = mx[a]; }
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int a = 0; a < 3; ++a) {
                smin[a][threadIdx.x] = fminf(smin[a][threadIdx.x], smin[a][threadIdx.x + s]);
                smax[a][threadIdx.x] = fmaxf(smax[a][threadIdx.x], smax[a][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int a = 0; a < 3; ++a) {
            partial[blockIdx.x * 6 + a] = smin[a][0];
            partial[blockIdx.x * 6 + 3 + a] = smax[a][0];
        }
}

__global__ __launch_bounds__(256) void k_fold_bounds(float* __restrict__ partial, uint32_t nparts) {
    __shared__ float s[6][256];
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = threadIdx.x; i < nparts; i += 256)
        for (int a = 0; a < 3; ++a) {
            v[a] = fminf(v[a], partial[i * 6 + a]);
            v[3 + a] = fmaxf(v[3 + a], partial[i * 6 + 3 + a]);
        }
    for (int a = 0; a < 6; ++a) s[a][threadIdx.x] = v[a];
    __syncthreads();
    for (uint32_t st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int a = 0; a < 3; ++a) {
                s[a][threadIdx.x] = fminf(s[a][threadIdx.x], s[a][threadIdx.x + st]);
                s[3 + a][threadIdx.x] = fmaxf(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + st]);
            }
        __syncthreads();
    }
    // result stored after the partials region at index nparts
    if (threadIdx.x < 6) partial[(size_t)nparts * 6 + threadIdx.x] = s[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_morton(const float* __restrict__ tp, uint32_t n,
                                                const float* __restrict__ bounds, uint32_t* __restrict__ keys,
                                                uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float lo[3], hi[3];
    tri_bounds(tp + (size_t)i * 9, lo, hi);
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) q[a] = quantize10((lo[a] + hi[a]) * 0.5f, bounds[a], bounds[3 + a]);
    keys[i] = (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
    vals[i] = i;
}

// --- stable LSD radix sort -------------------------------------------------
__global__ __launch_bounds__(kRadixBlock) void k_radix_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                            uint32_t shift, uint32_t nblocks,
                                                            uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kRadixDigits];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kRadixTile;
    for (uint32_t r = 0; r < kRadixItems; ++r) {
        const uint32_t idx = base + r * kRadixBlock + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of m entries by one block of 1024 threads
__global__ __launch_bounds__(1024) void k_scan_exclusive(uint32_t* __restrict__ data, uint32_t m) {
    __shared__ uint32_t part[1024];
    const uint32_t per = (m + 1023) / 1024;
    const uint32_t b = threadIdx.x * per, e = min(b + per, m);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += data[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t v = data[i];
        data[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(kRadixBlock) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                               const uint32_t* __restrict__ vin,
                                                               uint32_t* __restrict__ kout,
                                                               uint32_t* __restrict__ vout,
                                                               const uint32_t* __restrict__ offsets, uint32_t n,
                                                               uint32_t shift, uint32_t nblocks) {
    __shared__ uint32_t s_base[kRadixDigits];
    __shared__ uint32_t s_wave[kRadixBlock / 64][kRadixDigits];
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63u;
    s_base[t] = offsets[t * nblocks + blockIdx.x];
    const uint32_t base = blockIdx.x * kRadixTile;
    for (uint32_t r = 0; r < kRadixItems; ++r) {
        for (uint32_t w = 0; w < kRadixBlock / 64; ++w) s_wave[w][t] = 0u;
        __syncthreads();
        const uint32_t idx = base + r * kRadixBlock + t;
        const bool valid = idx < n;
        const uint32_t key = valid ? kin[idx] : 0u;
        const uint32_t digit = (key >> shift) & 0xFFu;
        uint64_t peers = __ballot(valid);
        for (uint32_t bit = 0; bit < 8; ++bit) {
            const bool set = (digit >> bit) & 1u;
            const uint64_t bm = __ballot(valid && set);
            peers &= set ? bm : ~bm;
        }
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(peers >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)peers, 0u));
        const int leader = __ffsll((unsigned long long)peers) - 1;
        if (valid && (int)lane == leader) s_wave[wave][digit] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t run = s_base[t];
            for (uint32_t w = 0; w < kRadixBlock / 64; ++w) {
                const uint32_t c = s_wave[w][t];
                s_wave[w][t] = run;
                run += c;
            }
            s_base[t] = run;
        }
        __syncthreads();
        if (valid) {
            const uint32_t pos = s_wave[wave][digit] + rank;
            kout[pos] = key;
            vout[pos] = vin[idx];
        }
        __syncthreads();
    }
}

// --- sorted geometry ----------------------------------------------------------
__global__ __launch_bounds__(256) void k_gather(const float* __restrict__ tp, const float* __restrict__ tn,
                                                const float* __restrict__ tuv, const uint32_t* __restrict__ tm,
                                                const uint32_t* __restrict__ sorted_prim, uint32_t n,
                                                float4* __restrict__ tris, float4* __restrict__ attrs,
                                                uint32_t* __restrict__ prim_out, uint32_t* __restrict__ slot_of) {
    const uint32_t s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const uint32_t p = sorted_prim[s];
    const float* t = tp + (size_t)p * 9;
    tris[(size_t)s * 3 + 0] = make_float4(t[0], t[1], t[2], __uint_as_float(p));
    tris[(size_t)s * 3 + 1] = make_float4(t[3], t[4], t[5], 0.0f);
    tris[(size_t)s * 3 + 2] = make_float4(t[6], t[7], t[8], 0.0f);
    float nn[9];
    if (tn) {
        for (int k = 0; k < 9; ++k) nn[k] = tn[(size_t)p * 9 + k];
    } else {
        const float e1x = t[3] - t[0], e1y = t[4] - t[1], e1z = t[5] - t[2];
        const float e2x = t[6] - t[0], e2y = t[7] - t[1], e2z = t[8] - t[2];
        const float cx = e1y * e2z - e1z * e2y, cy = e1z * e2x - e1x * e2z, cz = e1x * e2y - e1y * e2x;
        for (int k = 0; k < 3; ++k) { nn[k * 3] = cx; nn[k * 3 + 1] = cy; nn[k * 3 + 2] = cz; }
    }
    float uv[6] = {0, 0, 0, 0, 0, 0};
    if (tuv)
        for (int k = 0; k < 6; ++k) uv[k] = tuv[(size_t)p * 6 + k];
    const uint32_t m = tm ? tm[p] : 0u;
    attrs[(size_t)s * 4 + 0] = make_float4(nn[0], nn[1], nn[2], nn[3]);
    attrs[(size_t)s * 4 + 1] = make_float4(nn[4], nn[5], nn[6], nn[7]);
    attrs[(size_t)s * 4 + 2] = make_float4(nn[8], uv[0], uv[1], uv[2]);
    attrs[(size_t)s * 4 + 3] = make_float4(uv[3], uv[4], uv[5], __uint_as_float(m));
    prim_out[s] = p;
    slot_of[p] = s;
}

// --- Karras 2012 hierarchy ------------------------------------------------------
__device__ __forceinline__ int kdelta(const uint32_t* __restrict__ mk, const uint32_t* __restrict__ pk,
                                      int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = ((uint64_t)mk[i] << 32) | pk[i];
    const uint64_t b = ((uint64_t)mk[j] << 32) | pk[j];
    return __clzll((long long)(a ^ b));
}

__global__ __launch_bounds__(256) void k_karras(const uint32_t* __restrict__ mk, const uint32_t* __restrict__ pk,
                                                uint32_t n, uint32_t* __restrict__ nodes_u,
                                                uint32_t* __restrict__ node_parent,
                                                uint32_t* __restrict__ leaf_parent) {
    const int64_t i = blockIdx.x * 256 + threadIdx.x;
    const int64_t N = n;
    if (i >= N - 1) return;
    const int d = (kdelta(mk, pk, N, i, i + 1) - kdelta(mk, pk, N, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = kdelta(mk, pk, N, i, i - d);
    int64_t lmax = 2;
    while (kdelta(mk, pk, N, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (kdelta(mk, pk, N, i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = kdelta(mk, pk, N, i, j);
    int64_t s = 0, div = 2, t;
    do {
        t = (l + div - 1) / div;
        if (kdelta(mk, pk, N, i, i + (s + t) * d) > dnode) s += t;
        div *= 2;
    } while (t > 1);
    const int64_t gamma = i + s * d + min(d, 0);
    const int64_t lo = min(i, j), hi = max(i, j);
    const uint32_t c0 = (lo == gamma) ? (kLeafBit | (uint32_t)gamma) : (uint32_t)gamma;
    const uint32_t c1 = (hi == gamma + 1) ? (kLeafBit | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
    nodes_u[(size_t)i * 16 + 12] = c0;
    nodes_u[(size_t)i * 16 + 13] = c1;
    nodes_u[(size_t)i * 16 + 15] = 0u;
    if (c0 & kLeafBit) leaf_parent[c0 & ~kLeafBit] = (uint32_t)i; else node_parent[c0] = (uint32_t)i;
    if (c1 & kLeafBit) leaf_parent[c1 & ~kLeafBit] = (uint32_t)i; else node_parent[c1] = (uint32_t)i;
    if (i == 0) node_parent[0] = kMiss;
}

__global__ void k_single_leaf(uint32_t* nodes_u, uint32_t* node_parent, uint32_t* leaf_parent) {
    nodes_u[12] = kLeafBit;
    nodes_u[13] = kLeafBit;
    nodes_u[15] = 0u;
    node_parent[0] = kMiss;
    leaf_parent[0] = 0u;
}

__device__ __forceinline__ float ld_coherent(const float* p) {
    return __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t*>(p), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ void st_coherent(float* p, float v) {
    __hip_atomic_store(reinterpret_cast<uint32_t*>(p), __float_as_uint(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void child_box(uint32_t c, const float4* __restrict__ tris, const float* nbox,
                                          float lo[3], float hi[3]) {
    if (c & kLeafBit) {
        const size_t s = c & ~kLeafBit;
        const float4 a = tris[s * 3], b = tris[s * 3 + 1], e = tris[s * 3 + 2];
        const float v[9] = {a.x, a.y, a.z, b.x, b.y, b.z, e.x, e.y, e.z};
        tri_bounds(v, lo, hi);
    } else {
        for (int k = 0; k < 3; ++k) {
            lo[k] = ld_coherent(&nbox[(size_t)c * 6 + k]);
            hi[k] = ld_coherent(&nbox[(size_t)c * 6 + 3 + k]);
        }
    }
}

// One thread per leaf climbs to the root; the second arrival at a node builds it.
__global__ __launch_bounds__(256) void k_refit(const float4* __restrict__ tris, uint32_t n,
                                               float* __restrict__ nodes_f, const uint32_t* __restrict__ node_parent,
                                               const uint32_t* __restrict__ leaf_parent, uint32_t* __restrict__ flags,
                                               float* __restrict__ nbox) {
    const uint32_t s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    uint32_t node = leaf_parent[s];
    const uint32_t* nodes_u = reinterpret_cast<const uint32_t*>(nodes_f);
    while (node != kMiss) {
        const uint32_t prev = __hip_atomic_fetch_add(&flags[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t need = (n == 1) ? 0u : 1u;  // a single-leaf root has one arrival
        if (prev < need) return;
        const uint32_t c0 = nodes_u[(size_t)node * 16 + 12], c1 = nodes_u[(size_t)node * 16 + 13];
        float lo0[3], hi0[3], lo1[3], hi1[3];
        child_box(c0, tris, nbox, lo0, hi0);
        child_box(c1, tris, nbox, lo1, hi1);
        float* nd = nodes_f + (size_t)node * 16;
        for (int k = 0; k < 3; ++k) {
            nd[k] = lo0[k]; nd[3 + k] = hi0[k]; nd[6 + k] = lo1[k]; nd[9 + k] = hi1[k];
            st_coherent(&nbox[(size_t)node * 6 + k], fminf(lo0[k], lo1[k]));
            st_coherent(&nbox[(size_t)node * 6 + 3 + k], fmaxf(hi0[k], hi1[k]));
        }
        reinterpret_cast<uint32_t*>(nd)[14] = node_parent[node];
        __threadfence();
        node = node_parent[node];
    }
}

__global__ __launch_bounds__(256) void k_depth(const uint32_t* __restrict__ node_parent,
                                               const uint32_t* __restrict__ leaf_parent, uint32_t n,
                                               uint32_t* __restrict__ out) {
    const uint32_t s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    uint32_t d = 1, node = leaf_parent[s];
    while (node != kMiss && d < 4096) { node = node_parent[node]; ++d; }
    atomicMax(out, d);
}

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

struct ScratchLayout {
    size_t keys0, vals0, keys1, vals1, hist, partial, node_parent, leaf_parent, flags, nbox, depth, total;
};
ScratchLayout layout(uint32_t n) {
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const uint32_t nb = div_up(n, kRadixTile);
    ScratchLayout L{};
    size_t o = 0;
    L.keys0 = o; o += al((size_t)n * 4);
    L.vals0 = o; o += al((size_t)n * 4);
    L.keys1 = o; o += al((size_t)n * 4);
    L.vals1 = o; o += al((size_t)n * 4);
    L.hist = o; o += al((size_t)kRadixDigits * nb * 4 + 4);
    L.partial = o; o += al((size_t)(1024 + 1) * 6 * 4);
    L.node_parent = o; o += al((size_t)n * 4 + 4);
    L.leaf_parent = o; o += al((size_t)n * 4 + 4);
    L.flags = o; o += al((size_t)n * 4 + 4);
    L.nbox = o; o += al((size_t)n * 24 + 24);
    L.depth = o; o += al(4);
    L.total = o;
    return L;
}

__global__ void k_fill(uint32_t* p, uint32_t v, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

size_t radix_sort_scratch_bytes(uint32_t n) {
    const uint32_t nb = div_up(n, kRadixTile);
    return ((size_t)n * 8 + 256) + ((size_t)kRadixDigits * nb * 4 + 256);
}

hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t n, uint32_t key_bits, void* scratch,
                            hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint32_t nb = div_up(n, kRadixTile);
    uint32_t* k1 = reinterpret_cast<uint32_t*>(scratch);
    uint32_t* v1 = k1 + n;
    uint32_t* hist = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(scratch) + (((size_t)n * 8 + 255) & ~size_t(255)));
    uint32_t *ka = keys, *va = vals, *kb = k1, *vb = v1;
    const uint32_t passes = (key_bits + 7) / 8;
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t shift = p * 8;
        k_radix_hist<<<nb, kRadixBlock, 0, s>>>(ka, n, shift, nb, hist);
        k_scan_exclusive<<<1, 1024, 0, s>>>(hist, kRadixDigits * nb);
        k_radix_scatter<<<nb, kRadixBlock, 0, s>>>(ka, va, kb, vb, hist, n, shift, nb);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        hipMemcpyAsync(keys, ka, (size_t)n * 4, hipMemcpyDeviceToDevice, s);
        hipMemcpyAsync(vals, va, (size_t)n * 4, hipMemcpyDeviceToDevice, s);
    }
    return hipGetLastError();
}

size_t lbvh_scratch_bytes(uint32_t n) { return layout(n).total; }

hipError_t lbvh_build(const LbvhBuffers& b, hipStream_t s, uint32_t* depth_out) {
    const uint32_t n = b.n;
    if (n == 0) return hipSuccess;
    const ScratchLayout L = layout(n);
    char* sc = reinterpret_cast<char*>(b.scratch);
    uint32_t* keys0 = reinterpret_cast<uint32_t*>(sc + L.keys0);
    uint32_t* vals0 = reinterpret_cast<uint32_t*>(sc + L.vals0);
    uint32_t* keys1 = reinterpret_cast<uint32_t*>(sc + L.keys1);
    uint32_t* vals1 = reinterpret_cast<uint32_t*>(sc + L.vals1);
    uint32_t* hist = reinterpret_cast<uint32_t*>(sc + L.hist);
    float* partial = reinterpret_cast<float*>(sc + L.partial);
    uint32_t* node_parent = reinterpret_cast<uint32_t*>(sc + L.node_parent);
    uint32_t* leaf_parent = reinterpret_cast<uint32_t*>(sc + L.leaf_parent);
    uint32_t* flags = reinterpret_cast<uint32_t*>(sc + L.flags);
    float* nbox = reinterpret_cast<float*>(sc + L.nbox);
    uint32_t* depth = reinterpret_cast<uint32_t*>(sc + L.depth);

    const uint32_t nparts = std::min<uint32_t>(1024u, div_up(n, 256));
    k_centroid_bounds<<<nparts, 256, 0, s>>>(b.tri_pos, n, partial);
    k_fold_bounds<<<1, 256, 0, s>>>(partial, nparts);
    const float* bounds = partial + (size_t)nparts * 6;
    k_morton<<<div_up(n, 256), 256, 0, s>>>(b.tri_pos, n, bounds, keys0, vals0);

    // 30-bit keys: four 8-bit passes (the last covers bits 24..29)
    const uint32_t nb = div_up(n, kRadixTile);
    uint32_t *ka = keys0, *va = vals0, *kb = keys1, *vb = vals1;
    for (uint32_t p = 0; p < 4; ++p) {
        const uint32_t shift = p * 8;
        k_radix_hist<<<nb, kRadixBlock, 0, s>>>(ka, n, shift, nb, hist);
        k_scan_exclusive<<<1, 1024, 0, s>>>(hist, kRadixDigits * nb);
        k_radix_scatter<<<nb, kRadixBlock, 0, s>>>(ka, va, kb, vb, hist, n, shift, nb);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    k_gather<<<div_up(n, 256), 256, 0, s>>>(b.tri_pos, b.tri_nrm, b.tri_uv, b.tri_mat, va, n, b.tris, b.attrs,
                                             b.prim, b.slot_of);
    uint32_t* nodes_u = reinterpret_cast<uint32_t*>(b.nodes);
    const uint32_t ni = n > 1 ? n - 1 : 1;
    k_fill<<<div_up(ni, 256), 256, 0, s>>>(flags, 0u, ni);
    k_fill<<<1, 256, 0, s>>>(depth, 0u, 1);
    if (n == 1) {
        k_single_leaf<<<1, 1, 0, s>>>(nodes_u, node_parent, leaf_parent);
    } else {
        k_karras<<<div_up(n - 1, 256), 256, 0, s>>>(ka, va, n, nodes_u, node_parent, leaf_parent);
    }
    k_refit<<<div_up(n, 256), 256, 0, s>>>(b.tris, n, reinterpret_cast<float*>(b.nodes), node_parent, leaf_parent,
                                           flags, nbox);
    k_depth<<<div_up(n, 256), 256, 0, s>>>(node_parent, leaf_parent, n, depth);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (depth_out) {
        e = hipMemcpyAsync(depth_out, depth, 4, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(s);
    }
    return e;
}

}  // namespace ohpt
