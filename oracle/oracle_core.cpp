// ORACLE — TEST INFRASTRUCTURE ONLY (see oracle_core.hpp). Build with -ffp-contract=off.
#include "oracle_core.hpp"

#include <limits>
#include <stdexcept>
#include <numeric>

namespace oracle {

// Joe-Kuo new-joe-kuo-6.21201 direction numbers for Sobol dims 0..3. The
// sampler must reproduce the reference sequence exactly, so this is the same
// table as ohao/render/rt/sobol_generator.cpp:14-54 (checked bit for bit
// against that file built into oracle/_ref).
const uint32_t kSobolDirs[4 * 32] = {
    0x80000000u, 0x40000000u, 0x20000000u, 0x10000000u, 0x08000000u, 0x04000000u, 0x02000000u, 0x01000000u,
    0x00800000u, 0x00400000u, 0x00200000u, 0x00100000u, 0x00080000u, 0x00040000u, 0x00020000u, 0x00010000u,
    0x00008000u, 0x00004000u, 0x00002000u, 0x00001000u, 0x00000800u, 0x00000400u, 0x00000200u, 0x00000100u,
    0x00000080u, 0x00000040u, 0x00000020u, 0x00000010u, 0x00000008u, 0x00000004u, 0x00000002u, 0x00000001u,

    0x80000000u, 0xC0000000u, 0xA0000000u, 0xF0000000u, 0x88000000u, 0xCC000000u, 0xAA000000u, 0xFF000000u,
    0x80800000u, 0xC0C00000u, 0xA0A00000u, 0xF0F00000u, 0x88880000u, 0xCCCC0000u, 0xAAAA0000u, 0xFFFF0000u,
    0x80008000u, 0xC000C000u, 0xA000A000u, 0xF000F000u, 0x88008800u, 0xCC00CC00u, 0xAA00AA00u, 0xFF00FF00u,
    0x80808080u, 0xC0C0C0C0u, 0xA0A0A0A0u, 0xF0F0F0F0u, 0x88888888u, 0xCCCCCCCCu, 0xAAAAAAAAu, 0xFFFFFFFFu,

    0x80000000u, 0x40000000u, 0xE0000000u, 0x50000000u, 0x98000000u, 0xBC000000u, 0xAE000000u, 0x25000000u,
    0x5B800000u, 0x91400000u, 0x65600000u, 0x2E500000u, 0x91180000u, 0x88D40000u, 0xDD360000u, 0x02950000u,
    0x4C888000u, 0x09EB4000u, 0x86CA6000u, 0x13B15000u, 0x42ADD800u, 0x3DF8D400u, 0x2FE57E00u, 0x56EB8100u,
    0x8C41A680u, 0xE68455C0u, 0x3C9CE8E0u, 0xB44DBFF0u, 0x7C27F6C8u, 0x5239A6ACu, 0x853B0EDEu, 0x52EB2EF9u,

    0x80000000u, 0xC0000000u, 0x20000000u, 0x50000000u, 0xF8000000u, 0x74000000u, 0xA2000000u, 0x93000000u,
    0xD8800000u, 0x25400000u, 0x59E00000u, 0xE6D00000u, 0x78080000u, 0xB40C0000u, 0x82020000u, 0xC3050000u,
    0x208F8000u, 0x51CBC000u, 0xFBEA2000u, 0x75AD5000u, 0xA00AF800u, 0x90077400u, 0xD800A200u, 0x25009300u,
    0x59E08480u, 0xE6D0CEC0u, 0x78088520u, 0xB40CEBB0u, 0x82028348u, 0xC305C274u, 0x208FB552u, 0x51CB9AE9u,
};

// ------------------------------------------------------------------ env CDF
void env_cdf_build(const float* px, int W, int H, std::vector<float>& marg,
                   std::vector<float>& cond, float& integral) {
    cond.assign((size_t)W * H, 0.0f);
    marg.assign((size_t)H, 0.0f);
    std::vector<float> rowSum((size_t)H, 0.0f);
    const float kPi = 3.14159265358979323846f;
    for (int y = 0; y < H; ++y) {
        const float theta = kPi * (float(y) + 0.5f) / float(H);
        const float s = std::sin(theta);
        float acc = 0.0f;
        for (int x = 0; x < W; ++x) {
            const float* p = px + ((size_t)y * W + x) * 4;
            const float lum = 0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2];
            acc += lum * s;
            cond[(size_t)y * W + x] = acc;
        }
        if (acc > 0.0f) {
            for (int x = 0; x < W; ++x) cond[(size_t)y * W + x] /= acc;
        } else {
            for (int x = 0; x < W; ++x) cond[(size_t)y * W + x] = float(x + 1) / float(W);
        }
        rowSum[y] = acc;
    }
    float total = 0.0f;
    for (int y = 0; y < H; ++y) { total += rowSum[y]; marg[y] = total; }
    integral = total;
    if (total > 0.0f) {
        for (int y = 0; y < H; ++y) marg[y] /= total;
    } else {
        for (int y = 0; y < H; ++y) marg[y] = float(y + 1) / float(H);
    }
}

// ------------------------------------------------------------------ LBVH
static inline uint32_t expand_bits10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

static inline uint32_t quantize10(float c, float lo, float hi) {
    const float ext = hi - lo;
    float x = 0.5f;
    if (ext > 0.0f) x = (c - lo) / ext;
    const float f = x * 1024.0f;
    if (!(f > 0.0f)) return 0u;
    if (f >= 1023.0f) return 1023u;
    return (uint32_t)f;
}

uint32_t morton30(float cx, float cy, float cz, const float cmin[3], const float cmax[3]) {
    const uint32_t qx = quantize10(cx, cmin[0], cmax[0]);
    const uint32_t qy = quantize10(cy, cmin[1], cmax[1]);
    const uint32_t qz = quantize10(cz, cmin[2], cmax[2]);
    return (expand_bits10(qx) << 2) | (expand_bits10(qy) << 1) | expand_bits10(qz);
}

static inline int delta(const std::vector<uint64_t>& k, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    return __builtin_clzll(k[(size_t)i] ^ k[(size_t)j]);
}

static void tri_box(const float* t, float lo[3], float hi[3]) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = std::fmin(std::fmin(t[a], t[3 + a]), t[6 + a]);
        hi[a] = std::fmax(std::fmax(t[a], t[3 + a]), t[6 + a]);
    }
}

static void refit_boxes(Bvh& out);

void lbvh_build(const float* tp, uint32_t n, Bvh& out) {
    out = Bvh{};
    out.n = n;
    if (n == 0) return;
    // centroid bounds
    float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    std::vector<float> cen((size_t)n * 3);
    for (uint32_t i = 0; i < n; ++i) {
        float lo[3], hi[3];
        tri_box(tp + (size_t)i * 9, lo, hi);
        for (int a = 0; a < 3; ++a) {
            const float c = (lo[a] + hi[a]) * 0.5f;
            cen[(size_t)i * 3 + a] = c;
            cmin[a] = std::fmin(cmin[a], c);
            cmax[a] = std::fmax(cmax[a], c);
        }
    }
    std::vector<uint64_t> keys(n);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t m = morton30(cen[i * 3], cen[i * 3 + 1], cen[i * 3 + 2], cmin, cmax);
        keys[i] = ((uint64_t)m << 32) | i;
    }
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return (keys[a] >> 32) < (keys[b] >> 32); });
    out.prim = order;
    out.keys.resize(n);
    out.tri.assign((size_t)n * 12, 0.0f);
    for (uint32_t s = 0; s < n; ++s) {
        const uint32_t p = order[s];
        out.keys[s] = keys[p];
        const float* t = tp + (size_t)p * 9;
        float* d = &out.tri[(size_t)s * 12];
        d[0] = t[0]; d[1] = t[1]; d[2] = t[2]; std::memcpy(&d[3], &p, 4);
        d[4] = t[3]; d[5] = t[4]; d[6] = t[5]; d[7] = 0.0f;
        d[8] = t[6]; d[9] = t[7]; d[10] = t[8]; d[11] = 0.0f;
    }
    const uint32_t ni = n > 1 ? n - 1 : 1;
    out.nodes.assign(ni, BvhNode{});
    for (auto& nd : out.nodes) { nd.parent = 0xFFFFFFFFu; nd.pad = 0; }
    if (n == 1) {
        out.nodes[0].child[0] = kLeafBit | 0u;
        out.nodes[0].child[1] = kLeafBit | 0u;
    } else {
        // Karras 2012, "Maximizing parallelism in the construction of BVHs, octrees and k-d trees".
        const int64_t N = n;
        for (int64_t i = 0; i < N - 1; ++i) {
            const int d = (delta(out.keys, N, i, i + 1) - delta(out.keys, N, i, i - 1)) >= 0 ? 1 : -1;
            const int dmin = delta(out.keys, N, i, i - d);
            int64_t lmax = 2;
            while (delta(out.keys, N, i, i + lmax * d) > dmin) lmax *= 2;
            int64_t l = 0;
            for (int64_t t = lmax / 2; t >= 1; t /= 2)
                if (delta(out.keys, N, i, i + (l + t) * d) > dmin) l += t;
            const int64_t j = i + l * d;
            const int dnode = delta(out.keys, N, i, j);
            int64_t s = 0;
            int64_t div = 2;
            int64_t t;
            do {
                t = (l + div - 1) / div;
                if (delta(out.keys, N, i, i + (s + t) * d) > dnode) s += t;
                div *= 2;
            } while (t > 1);
            const int64_t gamma = i + s * d + std::min(d, 0);
            const int64_t lo = std::min(i, j), hi = std::max(i, j);
            BvhNode& nd = out.nodes[(size_t)i];
            nd.child[0] = (lo == gamma) ? (kLeafBit | (uint32_t)gamma) : (uint32_t)gamma;
            nd.child[1] = (hi == gamma + 1) ? (kLeafBit | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
            for (int c = 0; c < 2; ++c)
                if (!(nd.child[c] & kLeafBit)) out.nodes[nd.child[c]].parent = (uint32_t)i;
        }
    }
    refit_boxes(out);
    bvh4_collapse(out);
}

// Bottom-up box refit over the committed topology, post-order (children
// before parents), root = node 0.
static void refit_boxes(Bvh& out) {
    const uint32_t ni = (uint32_t)out.nodes.size();
    std::vector<uint32_t> post;
    post.reserve(ni);
    {
        std::vector<std::pair<uint32_t, int>> st;
        st.push_back({0u, 0});
        while (!st.empty()) {
            auto& top = st.back();
            if (top.second < 2) {
                const uint32_t c = out.nodes[top.first].child[top.second++];
                if (!(c & kLeafBit)) st.push_back({c, 0});
            } else {
                post.push_back(top.first);
                st.pop_back();
            }
        }
    }
    for (uint32_t idx : post) {
        BvhNode& nd = out.nodes[idx];
        for (int c = 0; c < 2; ++c) {
            float lo[3], hi[3];
            const uint32_t ch = nd.child[c];
            if (ch & kLeafBit) {
                const float* t = &out.tri[(size_t)(ch & ~kLeafBit) * 12];
                const float v[9] = {t[0], t[1], t[2], t[4], t[5], t[6], t[8], t[9], t[10]};
                tri_box(v, lo, hi);
            } else {
                const BvhNode& cn = out.nodes[ch];
                for (int a = 0; a < 3; ++a) {
                    lo[a] = std::fmin(cn.b[a], cn.b[6 + a]);
                    hi[a] = std::fmax(cn.b[3 + a], cn.b[9 + a]);
                }
            }
            for (int a = 0; a < 3; ++a) { nd.b[c * 6 + a] = lo[a]; nd.b[c * 6 + 3 + a] = hi[a]; }
        }
    }
}

void lbvh_refit(const float* tp, uint32_t n, Bvh& bvh) {
    if (n != bvh.n) throw std::runtime_error("lbvh_refit: triangle count changed");
    if (n == 0) return;
    for (uint32_t s = 0; s < n; ++s) {
        const uint32_t p = bvh.prim[s];
        const float* t = tp + (size_t)p * 9;
        float* d = &bvh.tri[(size_t)s * 12];
        d[0] = t[0]; d[1] = t[1]; d[2] = t[2];
        d[4] = t[3]; d[5] = t[4]; d[6] = t[5];
        d[8] = t[6]; d[9] = t[7]; d[10] = t[8];
    }
    refit_boxes(bvh);
    bvh4_collapse(bvh);
}

// ------------------------------------------------------------------ FP32 traversal
// Robust slab test (tmax scaled by 1 + 2*gamma(3), Ize 2013 / pbrt-v3) and the
// watertight ray/triangle test of Woop, Benthin, Wald (JCGT 2013); edge
// functions that round to zero are re-evaluated with the exact-sign
// difference of products below (the paper re-evaluates in double precision).
namespace {

struct RayPre {
    float o[3], inv[3];
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

constexpr float kGamma3Scale = 1.0f + 2.0f * (3.0f * 5.9604645e-08f) / (1.0f - 3.0f * 5.9604645e-08f);

inline void ray_setup(const float o[3], const float d_in[3], RayPre& r) {
    float d[3];
    for (int a = 0; a < 3; ++a) {
        float v = d_in[a];
        if (std::fabs(v) < 1e-20f) v = std::copysign(1e-20f, v);
        d[a] = v;
        r.o[a] = o[a];
        r.inv[a] = 1.0f / v;
    }
    const float ax = std::fabs(d[0]), ay = std::fabs(d[1]), az = std::fabs(d[2]);
    int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (d[kz] < 0.0f) std::swap(kx, ky);
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sx = d[kx] / d[kz];
    r.Sy = d[ky] / d[kz];
    r.Sz = 1.0f / d[kz];
}

// returns entry distance or +inf when missed
inline float box_hit(const RayPre& r, const float* lo, const float* hi, float tmin, float tmax) {
    float t0 = tmin, t1 = tmax;
    for (int a = 0; a < 3; ++a) {
        const float ta = (lo[a] - r.o[a]) * r.inv[a];
        const float tb = (hi[a] - r.o[a]) * r.inv[a];
        const float tn = std::fmin(ta, tb);
        const float tf = std::fmax(ta, tb) * kGamma3Scale;
        t0 = std::fmax(t0, tn);
        t1 = std::fmin(t1, tf);
    }
    return (t0 <= t1) ? t0 : INFINITY;
}

inline float comp(const float* v, int k) { return v[k]; }

// a*b - c*d by Kahan's algorithm with fused multiply-adds (error-free
// product residual; relative error below 2 ulp, Jeannerod, Louvet & Muller
// 2013): exactly zero only when the exact value is, otherwise of the exact
// sign -- what the watertight test needs when a float edge function rounds
// to zero.
inline float diff_of_products(float a, float b, float c, float d) {
    const float w = c * d;
    const float e = std::fma(-c, d, w);
    const float f = std::fma(a, b, -w);
    return f + e;
}

// returns true and updates t,u,v on a hit with tmin < t < tmax
inline bool tri_hit(const RayPre& r, const float* v0, const float* v1, const float* v2,
                    float tmin, float tmax, float& t_out, float& u_out, float& v_out) {
    float A[3], B[3], C[3];
    for (int a = 0; a < 3; ++a) { A[a] = v0[a] - r.o[a]; B[a] = v1[a] - r.o[a]; C[a] = v2[a] - r.o[a]; }
    const float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
    const float Ax = comp(A, r.kx) - r.Sx * Akz;
    const float Ay = comp(A, r.ky) - r.Sy * Akz;
    const float Bx = comp(B, r.kx) - r.Sx * Bkz;
    const float By = comp(B, r.ky) - r.Sy * Bkz;
    const float Cx = comp(C, r.kx) - r.Sx * Ckz;
    const float Cy = comp(C, r.ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = diff_of_products(Cx, By, Cy, Bx);
        V = diff_of_products(Ax, Cy, Ay, Cx);
        W = diff_of_products(Bx, Ay, By, Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    if (det < 0.0f) { T = -T; det = -det; U = -U; V = -V; W = -W; }
    if (!(T > tmin * det) || !(T < tmax * det)) return false;
    const float rcp = 1.0f / det;
    t_out = T * rcp;
    u_out = V * rcp;
    v_out = W * rcp;
    return true;
}

}  // namespace

// ---------------------------------------------------------------- 4-wide collapse
void bvh4_collapse(Bvh& bvh) {
    bvh.nodes4.clear();
    if (bvh.n == 0) return;
    const size_t ni = bvh.nodes.size();
    std::vector<uint32_t> depth(ni, 0), map(ni, kEmpty);
    {
        std::vector<uint32_t> st{0u};
        while (!st.empty()) {
            const uint32_t i = st.back();
            st.pop_back();
            for (int c = 0; c < 2; ++c) {
                const uint32_t ch = bvh.nodes[i].child[c];
                if (!(ch & kLeafBit)) { depth[ch] = depth[i] + 1; st.push_back(ch); }
            }
        }
    }
    uint32_t cnt = 0;
    for (size_t i = 0; i < ni; ++i)
        if ((depth[i] & 1u) == 0) map[i] = cnt++;
    bvh.nodes4.assign(cnt, Bvh4Node{});
    for (size_t i = 0; i < ni; ++i) {
        if (map[i] == kEmpty) continue;
        Bvh4Node& q = bvh.nodes4[map[i]];
        for (int k = 0; k < 4; ++k) {
            q.lox[k] = q.loy[k] = q.loz[k] = 0.0f;
            q.hix[k] = q.hiy[k] = q.hiz[k] = 0.0f;
            q.child[k] = kEmpty;
        }
        q.bin_node = (uint32_t)i;
        q.pad[0] = q.pad[1] = 0;
        uint32_t m = 0;
        auto put = [&](const float* b, uint32_t ref) {
            q.lox[m] = b[0]; q.loy[m] = b[1]; q.loz[m] = b[2];
            q.hix[m] = b[3]; q.hiy[m] = b[4]; q.hiz[m] = b[5];
            q.child[m] = ref;
            ++m;
        };
        const BvhNode& nd = bvh.nodes[i];
        for (int c = 0; c < 2; ++c) {
            const uint32_t ch = nd.child[c];
            if (ch & kLeafBit) {
                put(&nd.b[6 * c], ch);
            } else {
                const BvhNode& a = bvh.nodes[ch];
                for (int cc = 0; cc < 2; ++cc) {
                    const uint32_t g = a.child[cc];
                    put(&a.b[6 * cc], (g & kLeafBit) ? g : map[g]);
                }
            }
        }
        q.count = m;
    }
}

// Child order key: the entry distance's bits mapped to an unsigned total
// order with the two low bits replaced by the slot, so the keys of one node
// are distinct and entries within four ulps order by slot; +inf (missed or
// empty) sorts last.
static inline uint32_t order_key(float e, uint32_t slot) {
    uint32_t b;
    std::memcpy(&b, &e, 4);
    const uint32_t m = (b & 0x80000000u) ? 0xFFFFFFFFu : 0x80000000u;
    return ((b ^ m) & ~3u) | slot;
}

// Traversal order (identical in the HIP traversal kernels): at each 4-wide
// node the used child boxes are tested against [tmin, t_closest] and the hit
// children ordered by order_key; hit leaves are intersected in that order;
// internal children whose entry distance is still within the closest hit
// continue, the nearest next and the others pushed farthest first.
template <bool kAnyHit>
static HitF32 traverse4(const Bvh& bvh, const float o[3], const float d[3], float tmin, float tmax,
                        uint32_t* node_visits, uint32_t* tri_tests, bool* occ) {
    HitF32 h{tmax, 0xFFFFFFFFu, 0.0f, 0.0f};
    uint32_t nv = 0, nt = 0;
    if (occ) *occ = false;
    if (bvh.n == 0) return h;
    RayPre r;
    ray_setup(o, d, r);
    uint32_t stack[256];
    int sp = 0;
    uint32_t cur = 0;
    float tclose = tmax;
    for (;;) {
        ++nv;
        const Bvh4Node& nd = bvh.nodes4[cur];
        float e[4];
        for (int k = 0; k < 4; ++k) {
            if (nd.child[k] == kEmpty) { e[k] = INFINITY; continue; }
            const float lo[3] = {nd.lox[k], nd.loy[k], nd.loz[k]}, hi[3] = {nd.hix[k], nd.hiy[k], nd.hiz[k]};
            e[k] = box_hit(r, lo, hi, tmin, tclose);
        }
        uint32_t key[4];
        for (int k = 0; k < 4; ++k) key[k] = order_key(e[k], (uint32_t)k);
        std::sort(key, key + 4);
        int ord[4];
        for (int j = 0; j < 4; ++j) ord[j] = (int)(key[j] & 3u);
        for (int j = 0; j < 4; ++j) {
            const int k = ord[j];
            if (e[k] == INFINITY) break;
            if (!(nd.child[k] & kLeafBit)) continue;
            const float* t = &bvh.tri[(size_t)(nd.child[k] & ~kLeafBit) * 12];
            float tt, uu, vv;
            ++nt;
            if (tri_hit(r, t, t + 4, t + 8, tmin, tclose, tt, uu, vv)) {
                uint32_t p; std::memcpy(&p, &t[3], 4);
                h = HitF32{tt, p, uu, vv};
                tclose = tt;
                if (kAnyHit) {
                    if (occ) *occ = true;
                    if (node_visits) *node_visits = nv;
                    if (tri_tests) *tri_tests = nt;
                    return h;
                }
            }
        }
        uint32_t go[4];
        int ng = 0;
        for (int j = 0; j < 4; ++j) {
            const int k = ord[j];
            if (e[k] == INFINITY) break;
            if (!(nd.child[k] & kLeafBit) && e[k] <= tclose) go[ng++] = nd.child[k];
        }
        if (ng == 0) {
            if (sp == 0) break;
            cur = stack[--sp];
        } else {
            cur = go[0];
            for (int j = ng - 1; j >= 1; --j) stack[sp++] = go[j];
        }
    }
    if (node_visits) *node_visits = nv;
    if (tri_tests) *tri_tests = nt;
    return h;
}

bool tri_hit_f32(const float o[3], const float d[3], const float v[9], float tmin, float tmax, float tuv[3]) {
    RayPre r;
    ray_setup(o, d, r);
    return tri_hit(r, v, v + 3, v + 6, tmin, tmax, tuv[0], tuv[1], tuv[2]);
}

HitF32 trace_f32(const Bvh& bvh, const float o[3], const float d[3], float tmin, float tmax,
                 uint32_t* node_visits, uint32_t* tri_tests) {
    return traverse4<false>(bvh, o, d, tmin, tmax, node_visits, tri_tests, nullptr);
}

bool occluded_f32(const Bvh& bvh, const float o[3], const float d[3], float tmin, float tmax) {
    bool occ = false;
    traverse4<true>(bvh, o, d, tmin, tmax, nullptr, nullptr, &occ);
    return occ;
}

// ------------------------------------------------------------------ FP64 brute force
static bool tri_hit_f64(const double o[3], const double d[3], const float* t, double tmin,
                        double tmax, double& tout, double& uout, double& vout) {
    // Same watertight formulation in double precision.
    const double ax = std::fabs(d[0]), ay = std::fabs(d[1]), az = std::fabs(d[2]);
    int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (d[kz] < 0.0) std::swap(kx, ky);
    if (d[kz] == 0.0) return false;
    const double Sx = d[kx] / d[kz], Sy = d[ky] / d[kz], Sz = 1.0 / d[kz];
    double A[3], B[3], C[3];
    for (int a = 0; a < 3; ++a) {
        A[a] = (double)t[a] - o[a];
        B[a] = (double)t[3 + a] - o[a];
        C[a] = (double)t[6 + a] - o[a];
    }
    const double Ax = A[kx] - Sx * A[kz], Ay = A[ky] - Sy * A[kz];
    const double Bx = B[kx] - Sx * B[kz], By = B[ky] - Sy * B[kz];
    const double Cx = C[kx] - Sx * C[kz], Cy = C[ky] - Sy * C[kz];
    double U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if ((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0)) return false;
    double det = U + V + W;
    if (det == 0.0) return false;
    double T = U * Sz * A[kz] + V * Sz * B[kz] + W * Sz * C[kz];
    if (det < 0) { T = -T; det = -det; U = -U; V = -V; W = -W; }
    const double tt = T / det;
    if (!(tt > tmin) || !(tt < tmax)) return false;
    tout = tt; uout = V / det; vout = W / det;
    return true;
}

HitF64 brute_closest_f64(const float* tp, uint32_t n, const double o[3], const double d[3],
                         double tmin, double tmax, double rel_tie, uint32_t* ties,
                         uint32_t max_ties) {
    HitF64 best{tmax, 0xFFFFFFFFu, 0, 0, 0};
    std::vector<std::pair<double, uint32_t>> hits;
    for (uint32_t i = 0; i < n; ++i) {
        double t, u, v;
        if (tri_hit_f64(o, d, tp + (size_t)i * 9, tmin, tmax, t, u, v)) {
            hits.push_back({t, i});
            if (t < best.t) best = HitF64{t, i, u, v, 0};
        }
    }
    uint32_t nt = 0;
    if (best.prim != 0xFFFFFFFFu) {
        for (auto& h : hits) {
            if (h.second == best.prim) continue;
            if (std::fabs(h.first - best.t) <= rel_tie * std::fabs(best.t)) {
                if (ties && nt < max_ties) ties[nt] = h.second;
                ++nt;
            }
        }
    }
    best.n_ties = nt;
    return best;
}

}  // namespace oracle
