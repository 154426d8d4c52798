// rt_core.hpp — arithmetic shared by every kernel of the tracer.
//
// All of it is compiled with -ffp-contract=off and IEEE division/sqrt, so each
// float operation rounds exactly once, in the order written.  That order is the
// reference's: dot products sum from 0, left to right
// (software_renderer/math/vector.h:235-241), cross products as in
// vector.h:288-297, Moller-Trumbore as in graphic/triangle.cpp:21-45.  The CPU
// restatement in oracle/ uses the same order, which is what makes hit ids,
// t, u, v and whole paths bit-identical between the two.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rt.h"

#define RT_HD __host__ __device__ __forceinline__

namespace rt {

struct f3 { float x, y, z; };

RT_HD f3 mk(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
RT_HD f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
RT_HD f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
RT_HD f3 mulv(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
RT_HD f3 scale(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
// vector.h:235-241: T sum = 0; sum += a[i]*b[i]
RT_HD float dot(f3 a, f3 b) {
    float s = 0.0f;
    s = s + a.x * b.x;
    s = s + a.y * b.y;
    s = s + a.z * b.z;
    return s;
}
// vector.h:288-297
RT_HD f3 cross(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// vector.h:202-225: |v|^2 summed in float, sqrt and 1/|v| in double, each
// component multiplied in double and rounded back to float.
RT_HD f3 normalize_ref(f3 a) {
    float m2 = dot(a, a);
    double m = sqrt((double)m2);
    if (m != 0.0) {
        double f = 1.0 / m;
        a.x = (float)((double)a.x * f);
        a.y = (float)((double)a.y * f);
        a.z = (float)((double)a.z * f);
    }
    return a;
}

// Moller-Trumbore exactly as triangle.cpp:21-45 (two-sided, EPSILON 1e-7,
// inclusive edges).  Returns true and t/u/v on a hit with t > 1e-7.
RT_HD bool intersect_mt(f3 o, f3 d, f3 v0, f3 e0, f3 e1, float& t, float& u, float& v) {
    const float EPS = 1e-7f;
    f3 h = cross(d, e1);
    float a = dot(e0, h);
    if (a > -EPS && a < EPS) return false;
    float f = 1.0f / a;
    f3 s = sub(o, v0);
    u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f) return false;
    f3 q = cross(s, e0);
    v = f * dot(d, q);
    if (v < 0.0f || u + v > 1.0f) return false;
    t = f * dot(e1, q);
    return t > EPS;
}

// The same test with division-free early rejection.  v_rcp_f32 (1 ulp)
// gives approximate u, v, t; a candidate is rejected early only when the
// approximation is outside the acceptance region by more than 2^-20 of its
// magnitude, which bounds the difference to the correctly rounded values, or
// when its t lies beyond `tcull` (no triangle there can beat the current
// best, the same argument as the box culling).  Everything that survives runs
// the exact test above, so the accepted set and the returned t, u, v are
// those of intersect_mt.
__device__ __forceinline__ bool intersect_mt_fast(f3 o, f3 d, f3 v0, f3 e0, f3 e1, float tmin, float tmax,
                                                  float tcull, float& t, float& u, float& v) {
    const float EPS = 1e-7f;
    f3 h = cross(d, e1);
    float a = dot(e0, h);
    if (a > -EPS && a < EPS) return false;
    f3 s = sub(o, v0);
    const float su = dot(s, h);
    f3 q = cross(s, e0);
    const float sv = dot(d, q);
    const float st = dot(e1, q);
    if (fabsf(a) < 1e30f) {
        const float r = __builtin_amdgcn_rcpf(a);
        const float ua = r * su, va = r * sv, ta = r * st;
        const float tu = fabsf(ua) * 0x1p-20f + 0x1p-100f;
        const float tv = fabsf(va) * 0x1p-20f + 0x1p-100f;
        const float tt = fabsf(ta) * 0x1p-20f + 0x1p-100f;
        if (ua < -tu || ua > 1.0f + tu) return false;
        if (va < -tv || ua + va > 1.0f + tu + tv + 0x1p-20f) return false;
        if (ta + tt < EPS || ta + tt < tmin || ta - tt > tcull || ta - tt > tmax) return false;
    }
    const float f = 1.0f / a;
    u = f * su;
    if (u < 0.0f || u > 1.0f) return false;
    v = f * sv;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = f * st;
    return t > EPS;
}

// Hit point o + t*d (triangle.cpp:39) and its squared distance to o
// (graphics_object.cpp:77 -> vector.h:227-229, (o - p).magnitudeSquared()).
RT_HD f3 hit_point(f3 o, f3 d, float t) { return add(o, scale(t, d)); }
RT_HD float dist2(f3 o, f3 p) { f3 w = sub(o, p); return dot(w, w); }

// Closest-hit order: smaller squared distance wins; exact ties go to the
// lower triangle index.
RT_HD bool closer(float d2a, uint32_t ida, float d2b, uint32_t idb) {
    return d2a < d2b || (d2a == d2b && ida < idb);
}

// ---------------------------------------------------------------------------
// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11).  Counter-based: the stream of path p, sample s,
// bounce b is Philox(key = seed, ctr = (p, s, b, lane)), independent of how
// paths are scheduled.
struct u4 { uint32_t x, y, z, w; };

RT_HD void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

RT_HD u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(M0, c.x, hi0, lo0);
        mulhilo(M1, c.z, hi1, lo1);
        u4 n;
        n.x = hi1 ^ c.y ^ k0;
        n.y = lo1;
        n.z = hi0 ^ c.w ^ k1;
        n.w = lo0;
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// uniform float in [0, 1) with 24 random bits, like
// std::uniform_real_distribution<float>(0, 1) (math/random.cpp:4-11).
RT_HD float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// sin and cos of 2*pi*x for x = k * 2^-24 in [0, 1).  Written from the Taylor
// series on one octant (|a| <= pi/4, truncation < 2e-9) with exact octant
// reduction (8x is exact for such x); the same float operations on host and
// device, so both sides get the same bits.
RT_HD void sincos_2pi(float x, float& s, float& c) {
    float e = x * 8.0f;               // exact
    int o = (int)e;                   // octant 0..7
    float r = e - (float)o;           // exact, in [0, 1)
    if (o & 1) r = 1.0f - r;          // exact
    float a = r * 0x1.921fb6p-1f;     // r * pi/4
    float a2 = a * a;
    // coefficients: float(+-1/k!) for the odd (sin) and even (cos) terms
    float ps = a * (1.0f + a2 * (-0x1.555556p-3f + a2 * (0x1.111112p-7f +
               a2 * (-0x1.a01a02p-13f + a2 * 0x1.71de3ap-19f))));
    float pc = 1.0f + a2 * (-0.5f + a2 * (0x1.555556p-5f + a2 * (-0x1.6c16c2p-10f +
               a2 * (0x1.a01a02p-16f + a2 * -0x1.27e4fcp-22f))));
    // within the quadrant q = o >> 1 the angle is (o & 1) ? pi/2 - a : a
    float sq = (o & 1) ? pc : ps;
    float cq = (o & 1) ? ps : pc;
    switch (o >> 1) {
        case 0: s = sq;  c = cq;  break;
        case 1: s = cq;  c = -sq; break;
        case 2: s = -sq; c = -cq; break;
        default: s = -cq; c = sq; break;
    }
}

// Uniform direction on the unit sphere, flipped into the hemisphere of n
// (math/random.cpp:16-35): theta = 2*pi*xi1, z = 2*xi2 - 1,
// (sqrt(1-z^2) cos theta, sqrt(1-z^2) sin theta, z); negate when d.n < 0.
RT_HD f3 hemisphere_dir(float xi1, float xi2, f3 n) {
    float st, ct;
    sincos_2pi(xi1, st, ct);
    float z = 2.0f * xi2 - 1.0f;
    float r = sqrtf(1.0f - z * z);
    f3 d = mk(r * ct, r * st, z);
    if (dot(d, n) < 0.0f) d = mk(-d.x, -d.y, -d.z);
    return d;
}

// reflect (vector.h:319-322): the scalar 2.0*N.I is double, each product
// with N is rounded to float.
RT_HD f3 reflect_ref(f3 I, f3 N) {
    double k = 2.0 * (double)dot(N, I);
    f3 kn = mk((float)(k * (double)N.x), (float)(k * (double)N.y), (float)(k * (double)N.z));
    return sub(I, kn);
}

// refract / customRefract (vector.h:324-345) with the same float/double
// promotions as the reference.
RT_HD f3 refract_ref(f3 I, f3 N, float eta) {
    float ndi = dot(N, I);
    float k = (float)(1.0 - (double)(eta * eta) * (1.0 - (double)(ndi * ndi)));
    if (k < 0.0f) return mk(0.0f, 0.0f, 0.0f);
    f3 ei = scale(eta, I);
    double m = (double)(eta * ndi) + sqrt((double)k);
    f3 mn = mk((float)(m * (double)N.x), (float)(m * (double)N.y), (float)(m * (double)N.z));
    return sub(ei, mn);
}

// sin(acos(x)) for the branch test of customRefract, as sqrt((1 - x)(1 + x))
// in double: for a float x both factors are exact, so the result carries two
// roundings and is the same on host and device (libm's sin and acos differ
// between the two).  NaN for |x| > 1, as acos gives.  It stays within 2^-50
// of glibc's sin(acos(x)) over [-1, 1] (tested), so the float angle it yields
// can differ from the reference's only by one rounding near a float boundary.
RT_HD double sin_acos_ref(double x) { return sqrt((1.0 - x) * (1.0 + x)); }

RT_HD f3 custom_refract_ref(f3 I, f3 N, float eta) {
    if ((double)dot(N, I) > 0.0) {
        N = mk(-N.x, -N.y, -N.z);
    } else {
        eta = (float)(1.0 / (double)eta);
    }
    float angle = (float)(sin_acos_ref((double)dot(N, I)) * (double)eta);
    if (-1.0f < angle && angle < 1.0f) return refract_ref(I, N, eta);
    return reflect_ref(I, N);
}

// Row-vector product v * M (matrix.h:245-253): ret[j] += m[i][j] * v[i],
// i = 0..3, starting from 0.
RT_HD void vec4_mat(const float* m, const float v[4], float out[4]) {
    for (int j = 0; j < 4; ++j) out[j] = 0.0f;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out[j] = out[j] + m[4 * i + j] * v[i];
}

// Primary ray direction of pixel (px, py), path_tracer.cpp:19-23.
RT_HD f3 primary_dir(const float* view_inv, const float* proj_inv, uint32_t px, uint32_t py,
                     uint32_t w, uint32_t h) {
    float cx = (float)px + 0.5f, cy = (float)py + 0.5f;
    float ux = cx / (float)w, uy = cy / (float)h;
    float dx = 2.0f * ux - 1.0f, dy = 2.0f * uy - 1.0f;
    float v4[4] = {dx, dy, 1.0f, 1.0f}, t4[4];
    vec4_mat(proj_inv, v4, t4);
    f3 target = normalize_ref(mk(t4[0], t4[1], t4[2]));
    float w4[4] = {target.x, target.y, target.z, 0.0f}, d4[4];
    vec4_mat(view_inv, w4, d4);
    return normalize_ref(mk(d4[0], d4[1], d4[2]));
}

// Triangle area 0.5 |(p1 - p0) x (p2 - p0)| (light sampling).
RT_HD float tri_area(f3 p0, f3 p1, f3 p2) {
    const f3 c = cross(sub(p1, p0), sub(p2, p0));
    return 0.5f * sqrtf(dot(c, c));
}

// Reciprocal used only for box culling: zero components are replaced by a
// tiny signed value so the slab test never forms 0 * inf.
RT_HD float safe_rcp(float d) {
    const float tiny = 8.27180613e-25f;  // 2^-80
    float a = fabsf(d) < tiny ? copysignf(tiny, d) : d;
    return 1.0f / a;
}

// Dense positions for the flagged lanes of a workgroup, for K counters at
// once, with one atomic per counter and workgroup: same-address atomics from
// every wave of a launch over 2^20 slots queue up at the memory-side atomic
// unit, and the K atomics here are independent, so they overlap.  pos[k] is
// meaningful for lanes with flag[k] set; lanes keep their order inside the
// workgroup.  Every thread of the workgroup (BLOCK threads) must call it; lds
// holds K * (BLOCK / 64 + 1) words.
template <uint32_t BLOCK, int K>
__device__ __forceinline__ void block_alloc(const bool (&flag)[K], uint32_t* const (&counter)[K], uint32_t* lds,
                                            uint32_t lane, uint64_t lt, uint32_t (&pos)[K]) {
    constexpr uint32_t NW = BLOCK / 64;
    const uint32_t w = threadIdx.x >> 6;
    uint64_t m[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        m[k] = __ballot(flag[k]);
        if (lane == 0) lds[k * (NW + 1) + w] = (uint32_t)__popcll(m[k]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            tot[k] = 0;
            for (uint32_t i = 0; i < NW; ++i) {
                const uint32_t c = lds[k * (NW + 1) + i];
                lds[k * (NW + 1) + i] = tot[k];
                tot[k] += c;
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) lds[k * (NW + 1) + NW] = tot[k] ? atomicAdd(counter[k], tot[k]) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) pos[k] = lds[k * (NW + 1) + NW] + lds[k * (NW + 1) + w] + (uint32_t)__popcll(m[k] & lt);
    __syncthreads();  // lds is reused by the next call
}

// As block_alloc, for per-lane counts: pos[k] is the first of the cnt[k]
// consecutive positions the lane gets from *counter[k].
template <uint32_t BLOCK, int K>
__device__ __forceinline__ void block_alloc_counts(const uint32_t (&cnt)[K], uint32_t* const (&counter)[K],
                                                   uint32_t* lds, uint32_t (&pos)[K]) {
    constexpr uint32_t NW = BLOCK / 64;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t x[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        x[k] = cnt[k];  // inclusive prefix over the wave
#pragma unroll
        for (uint32_t o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x[k], o);
            if (lane >= o) x[k] += y;
        }
        if (lane == 63u) lds[k * (NW + 1) + w] = x[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            tot[k] = 0;
            for (uint32_t i = 0; i < NW; ++i) {
                const uint32_t c = lds[k * (NW + 1) + i];
                lds[k * (NW + 1) + i] = tot[k];
                tot[k] += c;
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) lds[k * (NW + 1) + NW] = tot[k] ? atomicAdd(counter[k], tot[k]) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) pos[k] = lds[k * (NW + 1) + NW] + lds[k * (NW + 1) + w] + x[k] - cnt[k];
    __syncthreads();  // lds is reused by the next call
}

}  // namespace rt
