// lbvh.hip — 4-wide BVH built on the GPU.
//
// Replaces the reference's agglomerative build (software_renderer/math/
// bounding_volume_hierachy.cpp:9-27,47-75: repeatedly merge the closest pair of
// centroids, O(n^3)) with a linear BVH:
//   1. 30-bit Morton code of each triangle's box centre (k_morton)
//   2. LSD radix sort, 4 passes of 8 bits, stable (k_rs_hist, k_scan_*,
//      k_rs_scatter)
//   3. binary hierarchy from the sorted codes (Karras, "Maximizing
//      parallelism in the construction of BVHs, octrees, and k-d trees",
//      HPG 2012) (k_karras); with RT_BUILD_CLUSTER, parallel locally-ordered
//      clustering of the sorted triangles instead (PLOC, k_ploc_*), its
//      leaves then renumbered in depth-first order so that every subtree
//      covers a contiguous run
//   4. node boxes: unions of sorted leaf ranges (block unions, a sparse table
//      over blocks, one range query per node) (k_node_boxes); the clustering
//      makes its unions as it merges
//   5. top-down collapse to 4-wide nodes: a node opens its largest-area inner
//      child until it has four children; a subtree of <= leaf_size triangles
//      becomes a leaf over a contiguous range of the sorted order when its
//      surface-area cost is no more than splitting it once (k_collapse)
//   6. leaf-ordered triangle records v0 / edge0 / edge1 / id (k_tri_records)
//   7. the traversal's 64-byte nodes: each node's child boxes quantised to
//      8 bits against the node's own box, rounded outward (k_quantize)
// The topology does not affect results: the traversal's answer is defined by
// the closest-hit rule alone (see trace.hip).  Boxes are padded outward so the
// box test never rejects a triangle the intersection test would accept.
#include <algorithm>

#include "rt_internal.hpp"
#include "rt_core.hpp"

namespace rt {
namespace {

constexpr int RS_BLOCK = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;

// counters[] slots
constexpr int CTR_NODES = 0;
constexpr int CTR_Q0 = 1;      // entries of the collapse queues 0 and 1 (one level reads one, appends to the other)
constexpr int CTR_LEAVES = 2;
constexpr int CTR_ERROR = 3;   // set when a structural bound is exceeded (never for a valid tree)
constexpr int CTR_Q1 = 4;
constexpr int CTR_DEPTH = 5;   // levels with a non-empty input queue

__device__ __forceinline__ uint32_t expand10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__device__ __forceinline__ float quant(float x, float lo, float inv_ext) {
    float q = (x - lo) * inv_ext * 1024.0f;
    return fminf(fmaxf(q, 0.0f), 1023.0f);
}

__global__ void k_morton(uint32_t n, const float* __restrict__ pos, const uint32_t* __restrict__ idx,
                         float lx, float ly, float lz, float ix, float iy, float iz,
                         uint32_t* __restrict__ keys, uint32_t* __restrict__ vals, float* __restrict__ leafbox) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t a = idx[3 * i], b = idx[3 * i + 1], c = idx[3 * i + 2];
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        float pa = pos[3 * a + k], pb = pos[3 * b + k], pc = pos[3 * c + k];
        lo[k] = fminf(pa, fminf(pb, pc));
        hi[k] = fmaxf(pa, fmaxf(pb, pc));
    }
    for (int k = 0; k < 3; ++k) {
        leafbox[6 * i + k] = lo[k];
        leafbox[6 * i + 3 + k] = hi[k];
    }
    uint32_t qx = (uint32_t)quant(0.5f * (lo[0] + hi[0]), lx, ix);
    uint32_t qy = (uint32_t)quant(0.5f * (lo[1] + hi[1]), ly, iy);
    uint32_t qz = (uint32_t)quant(0.5f * (lo[2] + hi[2]), lz, iz);
    keys[i] = (expand10(qx) << 2) | (expand10(qy) << 1) | expand10(qz);
    vals[i] = i;
}

__global__ void __launch_bounds__(RS_BLOCK) k_rs_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift,
                                                      uint32_t nblocks, uint32_t* __restrict__ hist) {
    __shared__ uint32_t cnt[256];
    const int t = threadIdx.x;
    cnt[t] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * RS_TILE;
    for (int j = 0; j < RS_ITEMS; ++j) {
        uint32_t i = base + j * RS_BLOCK + t;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(uint32_t)t * nblocks + blockIdx.x] = cnt[t];
}

// Exclusive scan of `total` counters in place, in three coalesced passes:
// every 1024-counter tile is scanned by one block (k_scan_tiles), the tile
// totals by one block (k_scan_totals), and the totals' prefixes added back
// (k_scan_add).
constexpr uint32_t SCAN_TILE = 1024;

// exclusive scan of 256 values held one per thread; returns the total
__device__ __forceinline__ uint32_t block_scan_256(uint32_t& v, uint32_t* sh) {
    const uint32_t t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (uint32_t off = 1; off < 256; off <<= 1) {
        const uint32_t x = t >= off ? sh[t - off] : 0u;
        __syncthreads();
        sh[t] += x;
        __syncthreads();
    }
    const uint32_t incl = sh[t], total = sh[255];
    __syncthreads();
    v = incl - v;
    return total;
}

__global__ void __launch_bounds__(256) k_scan_tiles(uint32_t* __restrict__ h, uint32_t total,
                                                    uint32_t* __restrict__ tsum) {
    __shared__ uint32_t sh[256];
    const uint32_t base = blockIdx.x * SCAN_TILE + 4u * threadIdx.x;
    uint32_t x[4], s = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        x[j] = base + j < total ? h[base + j] : 0u;
        s += x[j];
    }
    uint32_t pre = s;
    const uint32_t tot = block_scan_256(pre, sh);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (base + j < total) h[base + j] = pre;
        pre += x[j];
    }
    if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(256) k_scan_totals(uint32_t* __restrict__ tsum, uint32_t ntiles) {
    __shared__ uint32_t sh[256];
    uint32_t carry = 0;
    for (uint32_t b = 0; b < ntiles; b += 256) {
        const uint32_t i = b + threadIdx.x;
        uint32_t v = i < ntiles ? tsum[i] : 0u;
        const uint32_t tot = block_scan_256(v, sh);
        if (i < ntiles) tsum[i] = v + carry;
        carry += tot;
    }
}

__global__ void __launch_bounds__(256) k_scan_add(uint32_t* __restrict__ h, uint32_t total,
                                                  const uint32_t* __restrict__ tsum) {
    const uint32_t add = tsum[blockIdx.x];
    const uint32_t base = blockIdx.x * SCAN_TILE + 4u * threadIdx.x;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (base + j < total) h[base + j] += add;
}

// Stable scatter: each of the RS_ITEMS rounds ranks 256 consecutive keys.
// Within a wave, keys with equal digits are found with 8 ballots; across the
// 4 waves of the block the per-wave counts are prefix-summed in LDS.
__global__ void __launch_bounds__(RS_BLOCK) k_rs_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                         uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                         uint32_t n, int shift, uint32_t nblocks,
                                                         const uint32_t* __restrict__ hist) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[RS_BLOCK / 64][256];
    const int t = threadIdx.x;
    const int w = t >> 6, lane = t & 63;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    run[t] = hist[(uint32_t)t * nblocks + blockIdx.x];
    const uint32_t base = blockIdx.x * RS_TILE;
    for (int j = 0; j < RS_ITEMS; ++j) {
        for (int k = 0; k < RS_BLOCK / 64; ++k) wcnt[k][t] = 0;
        __syncthreads();
        uint32_t i = base + j * RS_BLOCK + t;
        bool valid = i < n;
        uint32_t key = valid ? kin[i] : 0u;
        uint32_t val = valid ? vin[i] : 0u;
        uint32_t dig = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
        for (int bit = 0; bit < 8; ++bit) {
            int on = (dig >> bit) & 1;
            uint64_t bal = __ballot(on);
            peers &= on ? bal : ~bal;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wcnt[w][dig] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t off = run[dig] + rank;
            for (int k = 0; k < w; ++k) off += wcnt[k][dig];
            kout[off] = key;
            vout[off] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int k = 0; k < RS_BLOCK / 64; ++k) add += wcnt[k][t];
        run[t] += add;
        __syncthreads();
    }
}

// inner layout (uint32 arrays of length n): left, right, first, last, iparent, lparent
struct Inner {
    uint32_t *left, *right, *first, *last, *iparent, *lparent;
};

__device__ __forceinline__ int kdelta(const uint32_t* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if (a != b) return __clz(a ^ b);
    return 32 + __clz((uint32_t)i ^ (uint32_t)j);
}

__global__ void k_karras(const uint32_t* __restrict__ keys, int n, Inner in) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (kdelta(keys, n, i, i + 1) - kdelta(keys, n, i, i - 1)) > 0 ? 1 : -1;
    int dmin = kdelta(keys, n, i, i - d);
    int lmax = 2;
    while (kdelta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (kdelta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = kdelta(keys, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (kdelta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + min(d, 0);
    int lo = min(i, j), hi = max(i, j);
    uint32_t left = (lo == gamma) ? (RT_LEAF_BIT | (uint32_t)gamma) : (uint32_t)gamma;
    uint32_t right = (hi == gamma + 1) ? (RT_LEAF_BIT | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
    in.left[i] = left;
    in.right[i] = right;
    in.first[i] = (uint32_t)lo;
    in.last[i] = (uint32_t)hi;
    if (left & RT_LEAF_BIT) in.lparent[left & ~RT_LEAF_BIT] = (uint32_t)i; else in.iparent[left] = (uint32_t)i;
    if (right & RT_LEAF_BIT) in.lparent[right & ~RT_LEAF_BIT] = (uint32_t)i; else in.iparent[right] = (uint32_t)i;
    if (i == 0) in.iparent[0] = 0xFFFFFFFFu;
}

__device__ __forceinline__ void load_box(uint32_t ref, const uint32_t* __restrict__ vals, const float* leafbox,
                                         const float* innerbox, float b[6]) {
    const float* src = (ref & RT_LEAF_BIT) ? leafbox + 6u * vals[ref & ~RT_LEAF_BIT] : innerbox + 6u * ref;
    for (int k = 0; k < 6; ++k) b[k] = __builtin_nontemporal_load(src + k);
}

// Node boxes as range unions.  A Karras node covers the contiguous sorted
// leaves [first, last], so its box is the union of those leaves' boxes; min
// and max are exact, so this equals the union of its children's boxes
// without a bottom-up walk (whose per-node device-scope synchronisation
// dominated the build).  Leaves are grouped in blocks of 64; a sparse table
// over block boxes answers any run of whole blocks with two lookups, and a
// node adds at most 126 single leaves at the ends.
constexpr uint32_t BOX_BLOCK = 64;

__device__ __forceinline__ void box_union(float a[6], const float* b) {
    for (int k = 0; k < 3; ++k) {
        a[k] = fminf(a[k], b[k]);
        a[3 + k] = fmaxf(a[3 + k], b[3 + k]);
    }
}

// sorted leaf boxes and the unions of each block of 64 (table level 0)
__global__ void k_sorted_boxes(uint32_t n, const uint32_t* __restrict__ vals, const float* __restrict__ leafbox,
                               float* __restrict__ sbox) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* src = leafbox + 6u * vals[i];
    for (int k = 0; k < 6; ++k) sbox[6u * i + k] = src[k];
}

__global__ void k_block_boxes(uint32_t n, uint32_t nb, const float* __restrict__ sbox, float* __restrict__ t0) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nb) return;
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    const uint32_t e = min(n, (j + 1) * BOX_BLOCK);
    for (uint32_t i = j * BOX_BLOCK; i < e; ++i) box_union(b, sbox + 6u * i);
    for (int k = 0; k < 6; ++k) t0[6u * j + k] = b[k];
}

// table level k: union of level k-1 at j and j + half (where that exists)
__global__ void k_box_table_level(uint32_t nb, uint32_t half, const float* __restrict__ prev, float* __restrict__ cur) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nb) return;
    float b[6];
    for (int k = 0; k < 6; ++k) b[k] = prev[6u * j + k];
    if (j + half < nb) box_union(b, prev + 6u * (j + half));
    for (int k = 0; k < 6; ++k) cur[6u * j + k] = b[k];
}

__global__ void k_node_boxes(uint32_t ninner, Inner in, const float* __restrict__ sbox, const float* __restrict__ table,
                             uint32_t nb, float* __restrict__ innerbox) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ninner) return;
    const uint32_t a = in.first[p], z = in.last[p];
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    const uint32_t ja = (a + BOX_BLOCK - 1) / BOX_BLOCK;   // first whole block
    const uint32_t jz1 = (z + 1) / BOX_BLOCK;               // one past the last whole block
    if (ja >= jz1) {
        for (uint32_t i = a; i <= z; ++i) box_union(b, sbox + 6u * i);
    } else {
        for (uint32_t i = a; i < ja * BOX_BLOCK; ++i) box_union(b, sbox + 6u * i);
        for (uint32_t i = jz1 * BOX_BLOCK; i <= z; ++i) box_union(b, sbox + 6u * i);
        const uint32_t len = jz1 - ja;
        const uint32_t k = 31u - (uint32_t)__builtin_clz(len);
        const float* lvl = table + (size_t)k * nb * 6u;
        box_union(b, lvl + 6u * ja);
        box_union(b, lvl + 6u * (jz1 - (1u << k)));
    }
    for (int k = 0; k < 6; ++k) innerbox[6u * p + k] = b[k];
}

// ---- PLOC: binary hierarchy by parallel locally-ordered clustering --------
//
// Meister and Bittner, "Parallel locally-ordered clustering for bounding
// volume hierarchy construction", IEEE TVCG 24(12), 2018.  The clusters start
// as the triangles in Morton order.  Every iteration each cluster finds, among
// the PLOC_R clusters on either side in that order, the one whose union box
// has the smallest surface area (ties: the lower index); mutual nearest
// neighbours merge into a new inner node, which takes the lower one's place,
// and the list is compacted in order (two exclusive scans give the new nodes'
// numbers and the survivors' positions).  The best pair of the whole list
// under (area, lower index, higher index) is always mutual, so every
// iteration merges at least once and n - 1 merges end at one cluster.
// Inner nodes are numbered from n - 2 down in the order they are made, so
// the root is node 0 and every node is made in a later iteration than its
// children.  During the build in.last[x] holds x's triangle count.  Then a
// top-down pass, one launch per iteration from the last, gives every node
// the range [first, last] of positions it covers in depth-first leaf order
// and writes the triangle ids in that order: every subtree is a contiguous
// run again, as the collapse and the triangle records expect.  Nothing here
// depends on scheduling: the tree is a function of the triangles.
constexpr int PLOC_R = 16;
constexpr int PLOC_BLOCK = 256;
constexpr int CTR_PM = 8;    // live clusters at the start of an iteration: slot 8 + parity
constexpr int CTR_PC = 10;   // inner nodes made before an iteration: slot 10 + parity
constexpr int CTR_CUM = 16;  // counters[16 + t]: inner nodes made before iteration t
constexpr uint32_t PLOC_MAX_ITERS = 240;

__device__ __forceinline__ float union_half_area(const float* a, const float* b) {
    const float x = fmaxf(a[3], b[3]) - fminf(a[0], b[0]);
    const float y = fmaxf(a[4], b[4]) - fminf(a[1], b[1]);
    const float z = fmaxf(a[5], b[5]) - fminf(a[2], b[2]);
    return x * y + y * z + z * x;
}

__device__ __forceinline__ uint32_t ref_size(uint32_t ref, const uint32_t* last) {
    return (ref & RT_LEAF_BIT) ? 1u : last[ref];
}

// initial clusters: leaf i of the Morton order
__global__ void k_ploc_init(uint32_t n, uint32_t* __restrict__ cr) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cr[i] = RT_LEAF_BIT | i;
}

// nn[i]: the neighbour within PLOC_R of cluster i with the smallest union
// area (the first of equal ones); a lone cluster is its own neighbour
__global__ void __launch_bounds__(PLOC_BLOCK) k_ploc_nn(const float* __restrict__ cb, const uint32_t* __restrict__ ctr,
                                                        int par, uint32_t* __restrict__ nn) {
    __shared__ float sb[(PLOC_BLOCK + 2 * PLOC_R) * 6];
    const uint32_t m = ctr[CTR_PM + par];
    for (uint32_t b0 = blockIdx.x * PLOC_BLOCK; b0 < m; b0 += gridDim.x * PLOC_BLOCK) {
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < (PLOC_BLOCK + 2 * PLOC_R) * 6; k += PLOC_BLOCK) {
            const int64_t g = (int64_t)b0 - PLOC_R + (int64_t)(k / 6);
            sb[k] = (g >= 0 && g < (int64_t)m) ? cb[(size_t)g * 6 + k % 6] : 0.0f;
        }
        __syncthreads();
        const uint32_t i = b0 + threadIdx.x;
        if (i < m) {
            const float* bi = sb + (threadIdx.x + PLOC_R) * 6;
            float best = INFINITY;
            uint32_t bj = i;
            for (int d = -PLOC_R; d <= PLOC_R; ++d) {
                const int64_t j = (int64_t)i + d;
                if (d == 0 || j < 0 || j >= (int64_t)m) continue;
                const float a = union_half_area(bi, sb + (threadIdx.x + PLOC_R + d) * 6);
                if (a < best) {
                    best = a;
                    bj = (uint32_t)j;
                }
            }
            nn[i] = bj;
        }
    }
}

// lead[i] = 1 where cluster i merges with a higher mutual neighbour, keep[i]
// = 0 where it is absorbed by a lower one; entries m .. m_ub - 1 are zero
__global__ void k_ploc_flags(const uint32_t* __restrict__ nn, const uint32_t* __restrict__ ctr, int par, uint32_t m_ub,
                             uint32_t* __restrict__ lead, uint32_t* __restrict__ keep) {
    const uint32_t m = ctr[CTR_PM + par];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m_ub; i += gridDim.x * blockDim.x) {
        uint32_t l = 0u, k = 0u;
        if (i < m) {
            const uint32_t j = nn[i];
            const bool mutual = j != i && nn[j] == i;
            l = (mutual && i < j) ? 1u : 0u;
            k = (mutual && i > j) ? 0u : 1u;
        }
        lead[i] = l;
        keep[i] = k;
    }
}

// merges and compaction; lead_x / keep_x are the exclusive scans of the flags
__global__ void k_ploc_merge(const uint32_t* __restrict__ cr, const float* __restrict__ cb,
                             const uint32_t* __restrict__ nn, const uint32_t* __restrict__ lead_x,
                             const uint32_t* __restrict__ keep_x, uint32_t* ctr, int par, uint32_t n, uint32_t t,
                             Inner in, float* __restrict__ innerbox, uint32_t* __restrict__ cr_out,
                             float* __restrict__ cb_out) {
    const uint32_t m = ctr[CTR_PM + par], made = ctr[CTR_PC + par];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        const uint32_t j = nn[i];
        const bool mutual = j != i && nn[j] == i;
        const uint32_t p = keep_x[i];
        if (mutual && i < j) {
            const uint32_t slot = (n - 2u) - (made + lead_x[i]);
            const uint32_t a = cr[i], b = cr[j];
            in.left[slot] = a;
            in.right[slot] = b;
            in.last[slot] = ref_size(a, in.last) + ref_size(b, in.last);
            cr_out[p] = slot;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float lo = fminf(cb[6u * i + k], cb[6u * j + k]), hi = fmaxf(cb[6u * i + 3 + k], cb[6u * j + 3 + k]);
                innerbox[6u * slot + k] = lo;
                innerbox[6u * slot + 3 + k] = hi;
                cb_out[6u * p + k] = lo;
                cb_out[6u * p + 3 + k] = hi;
            }
        } else if (!mutual) {
            cr_out[p] = cr[i];
#pragma unroll
            for (int k = 0; k < 6; ++k) cb_out[6u * p + k] = cb[6u * i + k];
        }
        if (i == m - 1u) {
            const uint32_t made_next = made + lead_x[i] + ((mutual && i < j) ? 1u : 0u);
            ctr[CTR_PM + (par ^ 1)] = p + ((mutual && i > j) ? 0u : 1u);
            ctr[CTR_PC + (par ^ 1)] = made_next;
            ctr[CTR_CUM + t + 1u] = made_next;
        }
    }
}

// the nodes made in iteration t (their parents are done): children's first
// positions, the triangle ids of leaf children in depth-first order, and the
// node's last position
__global__ void k_ploc_topdown(const uint32_t* __restrict__ ctr, uint32_t t, uint32_t n, Inner in,
                               const uint32_t* __restrict__ vals_in, uint32_t* __restrict__ vals_out) {
    const uint32_t lo = (n - 1u) - ctr[CTR_CUM + t + 1u], hi = (n - 1u) - ctr[CTR_CUM + t];
    for (uint32_t s = lo + blockIdx.x * blockDim.x + threadIdx.x; s < hi; s += gridDim.x * blockDim.x) {
        const uint32_t f = in.first[s], a = in.left[s], b = in.right[s], cnt = in.last[s];
        const uint32_t ca = ref_size(a, in.last);
        if (a & RT_LEAF_BIT) {
            vals_out[f] = vals_in[a & ~RT_LEAF_BIT];
            in.left[s] = RT_LEAF_BIT | f;
        } else {
            in.first[a] = f;
        }
        if (b & RT_LEAF_BIT) {
            vals_out[f + ca] = vals_in[b & ~RT_LEAF_BIT];
            in.right[s] = RT_LEAF_BIT | (f + ca);
        } else {
            in.first[b] = f + ca;
        }
        in.last[s] = f + cnt - 1u;
    }
}

struct Entry {
    uint32_t ref, first, count;
    bool leaf;
    float b[6];
};

__device__ __forceinline__ float half_area(const float b[6]) {
    float x = b[3] - b[0], y = b[4] - b[1], z = b[5] - b[2];
    return x * y + y * z + z * x;
}

__device__ __forceinline__ uint32_t ref_count(uint32_t ref, const Inner& in) {
    return (ref & RT_LEAF_BIT) ? 1u : in.last[ref] - in.first[ref] + 1u;
}

// A subtree becomes a leaf when it has at most leaf_size triangles and the
// surface-area cost of testing them all is no more than that of one more
// node step plus its two halves as leaves:
//   count * A  <=  alpha * A + count_l * A_l + count_r * A_r
// (alpha = cost of a node step over a triangle step; a large alpha makes
// every small subtree a leaf).  Triangles that overlap little and sit apart
// (a soup) split; the two triangles of a heightfield quad stay together.
__device__ __forceinline__ void make_entry(uint32_t ref, const Inner& in, const uint32_t* vals, const float* leafbox,
                                           const float* innerbox, uint32_t leaf_size, float alpha, Entry& e) {
    e.ref = ref;
    if (ref & RT_LEAF_BIT) {
        e.first = ref & ~RT_LEAF_BIT;
        e.count = 1;
    } else {
        e.first = in.first[ref];
        e.count = in.last[ref] - e.first + 1;
    }
    load_box(ref, vals, leafbox, innerbox, e.b);
    if (ref & RT_LEAF_BIT) {
        e.leaf = true;
    } else if (e.count > leaf_size) {
        e.leaf = false;
    } else {
        float bl[6], br[6];
        const uint32_t l = in.left[ref], r = in.right[ref];
        load_box(l, vals, leafbox, innerbox, bl);
        load_box(r, vals, leafbox, innerbox, br);
        const float a = half_area(e.b);
        const float split = alpha * a + (float)ref_count(l, in) * half_area(bl) + (float)ref_count(r, in) * half_area(br);
        e.leaf = (float)e.count * a <= split;
    }
}

// One item per thread: (binary inner node, 4-wide slot, parent slot).  New
// node slots, next-level queue entries and the leaf count are taken with one
// atomic per counter and workgroup.
constexpr uint32_t COLLAPSE_BLOCK = 128;

__global__ void __launch_bounds__(COLLAPSE_BLOCK) k_collapse(const uint4* __restrict__ qin,
                                                             const uint32_t* __restrict__ nq_dev,
                                                             uint4* __restrict__ qout, Inner in,
                                                             const uint32_t* __restrict__ vals, const float* leafbox,
                                                             const float* innerbox, uint32_t leaf_size, float alpha,
                                                             float pad, rt_bvh4_node* __restrict__ nodes,
                                                             uint32_t max_nodes, uint32_t max_queue,
                                                             uint32_t* counters, int out_slot) {
    __shared__ uint32_t lds_alloc[3 * (COLLAPSE_BLOCK / 64 + 1)];
    // the level's queue length, read on the device; the grid strides over it
    const uint32_t nq = *nq_dev;
    for (uint32_t k0 = blockIdx.x * COLLAPSE_BLOCK; k0 < nq; k0 += gridDim.x * COLLAPSE_BLOCK) {
        const uint32_t k = k0 + threadIdx.x;
        const bool active = k < nq;
        uint4 it = make_uint4(0u, 0u, 0u, 0u);
        Entry e[4];
        int cnt = 0;
        if (active) {
            it = qin[k];
            cnt = 2;
            make_entry(in.left[it.x], in, vals, leafbox, innerbox, leaf_size, alpha, e[0]);
            make_entry(in.right[it.x], in, vals, leafbox, innerbox, leaf_size, alpha, e[1]);
            while (cnt < 4) {
                int best = -1;
                float best_a = -1.0f;
                for (int c = 0; c < cnt; ++c) {
                    if (!e[c].leaf) {
                        float a = half_area(e[c].b);
                        if (a > best_a) { best_a = a; best = c; }
                    }
                }
                if (best < 0) break;
                uint32_t r = e[best].ref;
                make_entry(in.left[r], in, vals, leafbox, innerbox, leaf_size, alpha, e[best]);
                make_entry(in.right[r], in, vals, leafbox, innerbox, leaf_size, alpha, e[cnt]);
                ++cnt;
            }
        }
        uint32_t inner = 0, leaves = 0;
        for (int c = 0; c < cnt; ++c) {
            if (e[c].leaf) ++leaves;
            else ++inner;
        }
        const uint32_t counts[3] = {inner, inner, leaves};
        uint32_t* const ctr[3] = {&counters[CTR_NODES], &counters[out_slot], &counters[CTR_LEAVES]};
        uint32_t base[3];
        block_alloc_counts<COLLAPSE_BLOCK, 3>(counts, ctr, lds_alloc, base);
        if (!active) continue;
        rt_bvh4_node nd;
        uint32_t j = 0;
        for (int c = 0; c < 4; ++c) {
            if (c < cnt) {
                nd.lo_x[c] = e[c].b[0] - pad; nd.hi_x[c] = e[c].b[3] + pad;
                nd.lo_y[c] = e[c].b[1] - pad; nd.hi_y[c] = e[c].b[4] + pad;
                nd.lo_z[c] = e[c].b[2] - pad; nd.hi_z[c] = e[c].b[5] + pad;
                if (e[c].leaf) {
                    nd.child[c] = RT_LEAF_BIT | ((e[c].count - 1u) << RT_LEAF_COUNT_SHIFT) | e[c].first;
                } else {
                    const uint32_t slot = base[0] + j, q = base[1] + j;
                    ++j;
                    if (slot < max_nodes && q < max_queue) {
                        qout[q] = make_uint4(e[c].ref, slot, it.y, 0u);
                        nd.child[c] = slot;
                    } else {
                        counters[CTR_ERROR] = 1u;
                        nd.child[c] = RT_CHILD_EMPTY;
                    }
                }
            } else {
                nd.lo_x[c] = nd.lo_y[c] = nd.lo_z[c] = INFINITY;
                nd.hi_x[c] = nd.hi_y[c] = nd.hi_z[c] = -INFINITY;
                nd.child[c] = RT_CHILD_EMPTY;
            }
        }
        nd.meta[0] = it.z;
        nd.meta[1] = nd.meta[2] = nd.meta[3] = 0u;
        nodes[it.y] = nd;
    }
}

// Before a collapse level: the queue it appends to starts empty, and a
// non-empty input queue counts one level of depth.
__global__ void k_level_prep(uint32_t* counters, int in_slot, int out_slot) {
    if (threadIdx.x == 0) {
        if (counters[in_slot] > 0) counters[CTR_DEPTH] += 1u;
        counters[out_slot] = 0u;
    }
}

// Root of a scene with <= leaf_size triangles: one leaf child.
__global__ void k_root_leaf(uint32_t n, const float* __restrict__ leafbox, const uint32_t* __restrict__ vals, float pad,
                            rt_bvh4_node* nodes) {
    rt_bvh4_node nd;
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint32_t j = 0; j < n; ++j) {
        const float* s = leafbox + 6u * vals[j];
        for (int k = 0; k < 3; ++k) { b[k] = fminf(b[k], s[k]); b[3 + k] = fmaxf(b[3 + k], s[3 + k]); }
    }
    for (int c = 0; c < 4; ++c) {
        nd.lo_x[c] = nd.lo_y[c] = nd.lo_z[c] = INFINITY;
        nd.hi_x[c] = nd.hi_y[c] = nd.hi_z[c] = -INFINITY;
        nd.child[c] = RT_CHILD_EMPTY;
    }
    nd.lo_x[0] = b[0] - pad; nd.lo_y[0] = b[1] - pad; nd.lo_z[0] = b[2] - pad;
    nd.hi_x[0] = b[3] + pad; nd.hi_y[0] = b[4] + pad; nd.hi_z[0] = b[5] + pad;
    nd.child[0] = RT_LEAF_BIT | ((n - 1u) << RT_LEAF_COUNT_SHIFT);
    nd.meta[0] = 0xFFFFFFFFu;
    nd.meta[1] = nd.meta[2] = nd.meta[3] = 0u;
    nodes[0] = nd;
}

// Leaf-ordered records: edge0 = v1 - v0, edge1 = v2 - v0 as the Triangle
// constructor computes them (triangle.cpp:4-8).
__global__ void k_tri_records(uint32_t n, const uint32_t* __restrict__ vals, const float* __restrict__ pos,
                              const uint32_t* __restrict__ idx, rt_tri_rec* __restrict__ tris) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t t = vals[j];
    uint32_t a = idx[3 * t], b = idx[3 * t + 1], c = idx[3 * t + 2];
    rt_tri_rec r;
    for (int k = 0; k < 3; ++k) {
        float p0 = pos[3 * a + k];
        r.v0[k] = p0;
        r.e0[k] = pos[3 * b + k] - p0;
        r.e1[k] = pos[3 * c + k] - p0;
    }
    r.id = t;
    r.pad0 = r.pad1 = 0u;
    tris[j] = r;
}

inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

}  // namespace

// The traversal's 64-byte node (rt_bvh4q_node) from a build node: per axis
// the origin is the low corner of the union of the valid children's boxes
// and the scale the least power of two with 255 steps covering its extent;
// low bounds round down and high bounds up (in double, where the offsets
// from the origin are exact), so the quantised boxes contain the float ones.
__global__ void k_quantize(const rt_bvh4_node* __restrict__ nodes, uint32_t n, rt_bvh4q_node* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const rt_bvh4_node nd = nodes[i];
    const float* lo3[3] = {nd.lo_x, nd.lo_y, nd.lo_z};
    const float* hi3[3] = {nd.hi_x, nd.hi_y, nd.hi_z};
    rt_bvh4q_node q;
    uint8_t* qlo[3] = {q.lo_x, q.lo_y, q.lo_z};
    uint8_t* qhi[3] = {q.hi_x, q.hi_y, q.hi_z};
    for (int k = 0; k < 3; ++k) {
        float lo = INFINITY, hi = -INFINITY;
        for (int c = 0; c < 4; ++c)
            if (nd.child[c] != RT_CHILD_EMPTY) {
                lo = fminf(lo, lo3[k][c]);
                hi = fmaxf(hi, hi3[k][c]);
            }
        if (!(lo <= hi)) lo = hi = 0.0f;
        const double ext = (double)hi - (double)lo;
        int e = -100;
        if (ext > 0.0) {
            const double step = ext / 255.0;
            e = ilogb(step);
            if (ldexp(1.0, e) < step) ++e;
            e = e < -100 ? -100 : (e > 100 ? 100 : e);
        }
        q.origin[k] = lo;
        q.exp[k] = (int8_t)e;
        for (int c = 0; c < 4; ++c) {
            double a = 255.0, b = 0.0;
            if (nd.child[c] != RT_CHILD_EMPTY) {
                a = fmin(fmax(floor(ldexp((double)lo3[k][c] - (double)lo, -e)), 0.0), 255.0);
                b = fmin(fmax(ceil(ldexp((double)hi3[k][c] - (double)lo, -e)), 0.0), 255.0);
            }
            qlo[k][c] = (uint8_t)a;
            qhi[k][c] = (uint8_t)b;
        }
    }
    q.pad = 0;
    q.parent = nd.meta[0];
    q.reserved = 0;
    for (int c = 0; c < 4; ++c) q.child[c] = nd.child[c];
    out[i] = q;
}

int radix_sort_pairs(rt_ctx* c, uint32_t* keys[2], uint32_t* vals[2], uint32_t n, uint32_t passes, int* cur_out) {
    int cur = 0;
    *cur_out = 0;
    if (n == 0 || passes == 0) return RT_OK;
    hipStream_t st = c->stream;
    const uint32_t nblocks = blocks_for(n, RS_TILE);
    const uint32_t scan_tiles = (256u * nblocks + SCAN_TILE - 1) / SCAN_TILE;
    int rc;
    if ((rc = ensure(c, c->hist, 4ull * 256 * nblocks + 4ull * scan_tiles))) return rc;
    uint32_t* hist = (uint32_t*)c->hist.ptr;
    uint32_t* tsum = hist + 256u * nblocks;
    for (uint32_t pass = 0; pass < passes; ++pass) {
        const int shift = (int)pass * 8;
        hipLaunchKernelGGL(k_rs_hist, dim3(nblocks), dim3(RS_BLOCK), 0, st, keys[cur], n, shift, nblocks, hist);
        hipLaunchKernelGGL(k_scan_tiles, dim3(scan_tiles), dim3(256), 0, st, hist, 256u * nblocks, tsum);
        hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(256), 0, st, tsum, scan_tiles);
        hipLaunchKernelGGL(k_scan_add, dim3(scan_tiles), dim3(256), 0, st, hist, 256u * nblocks, (const uint32_t*)tsum);
        hipLaunchKernelGGL(k_rs_scatter, dim3(nblocks), dim3(RS_BLOCK), 0, st, keys[cur], vals[cur], keys[cur ^ 1],
                           vals[cur ^ 1], n, shift, nblocks, (const uint32_t*)hist);
        cur ^= 1;
    }
    RT_HIP(c, hipGetLastError());
    *cur_out = cur;
    return RT_OK;
}

// PLOC hierarchy over the n triangles in Morton order (vals_morton); writes
// in (left, right, first, last), the inner boxes and vals_tree, the triangle
// ids in depth-first leaf order.  cr: two n-entry cluster buffers.  The host
// reads each iteration's cluster count from a pinned copy one iteration
// behind (it never drains the stream); the iteration launched after the last
// merge finds one cluster and changes nothing.
int build_ploc(rt_ctx* c, uint32_t n, const uint32_t* vals_morton, uint32_t* vals_tree, uint32_t* cr[2],
               const float* leafbox, const Inner& in, float* innerbox) {
    const size_t N = n;
    hipStream_t st = c->stream;
    const uint32_t stiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    int rc;
    if ((rc = ensure(c, c->boxtab, 48ull * N)) || (rc = ensure(c, c->hist, 8ull * stiles + 64)) ||
        (rc = ensure_snap(c)))
        return rc;
    float* cb[2] = {(float*)c->boxtab.ptr, (float*)c->boxtab.ptr + 6ull * N};
    uint32_t* nn = (uint32_t*)c->queue[0].ptr;
    uint32_t* lead = (uint32_t*)c->queue[1].ptr;
    uint32_t* keep = lead + N;
    uint32_t* tsum0 = (uint32_t*)c->hist.ptr;
    uint32_t* tsum1 = tsum0 + stiles;
    uint32_t* ctr = (uint32_t*)c->counters.ptr;
    hipLaunchKernelGGL(k_sorted_boxes, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, vals_morton, leafbox, cb[0]);
    hipLaunchKernelGGL(k_ploc_init, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, cr[0]);
    uint32_t init[CTR_CUM + 1] = {};
    init[CTR_PM] = n;
    RT_HIP(c, hipMemcpyAsync(ctr, init, sizeof(init), hipMemcpyHostToDevice, st));
    const uint32_t gcap = 16u * (uint32_t)c->num_cus;
    uint32_t m_ub = n, iters = 0;
    for (uint32_t t = 0;; ++t) {
        if (t >= PLOC_MAX_ITERS) return fail(c, RT_E_DEVICE, "rt_build: clustering did not converge");
        const int par = (int)(t & 1u);
        const uint32_t g = std::min(blocks_for(m_ub, PLOC_BLOCK), gcap);
        hipLaunchKernelGGL(k_ploc_nn, dim3(g), dim3(PLOC_BLOCK), 0, st, (const float*)cb[par], (const uint32_t*)ctr, par,
                           nn);
        hipLaunchKernelGGL(k_ploc_flags, dim3(g), dim3(256), 0, st, (const uint32_t*)nn, (const uint32_t*)ctr, par, m_ub,
                           lead, keep);
        const uint32_t tiles = (m_ub + SCAN_TILE - 1) / SCAN_TILE;
        uint32_t* arr[2] = {lead, keep};
        uint32_t* ts[2] = {tsum0, tsum1};
        for (int k = 0; k < 2; ++k) {
            hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(256), 0, st, arr[k], m_ub, ts[k]);
            hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(256), 0, st, ts[k], tiles);
            hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(256), 0, st, arr[k], m_ub, (const uint32_t*)ts[k]);
        }
        hipLaunchKernelGGL(k_ploc_merge, dim3(g), dim3(256), 0, st, (const uint32_t*)cr[par], (const float*)cb[par],
                           (const uint32_t*)nn, (const uint32_t*)lead, (const uint32_t*)keep, ctr, par, n, t, in,
                           innerbox, cr[par ^ 1], cb[par ^ 1]);
        RT_HIP(c, hipGetLastError());
        RT_HIP(c, hipMemcpyAsync(c->snap + 16 * (t & 1u), ctr + CTR_PM, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                 st));
        RT_HIP(c, hipEventRecord(c->snap_ev[t & 1u], st));
        iters = t + 1;
        if (t == 0) continue;
        RT_HIP(c, hipEventSynchronize(c->snap_ev[(t - 1) & 1u]));
        const uint32_t* h = c->snap + 16 * ((t - 1) & 1u);
        const uint32_t m_after = h[((t - 1) & 1u) ^ 1u];  // clusters after iteration t - 1
        if (m_after <= 1u) break;
        m_ub = m_after;  // bounds the clusters of iteration t + 1
    }
    RT_HIP(c, hipMemsetAsync(in.first, 0, sizeof(uint32_t), st));  // the root covers positions from 0
    const uint32_t gt = std::min(blocks_for(n / 2 + 1, 256), gcap);
    for (uint32_t t = iters; t-- > 0;)
        hipLaunchKernelGGL(k_ploc_topdown, dim3(gt), dim3(256), 0, st, (const uint32_t*)ctr, t, n, in, vals_morton,
                           vals_tree);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}

int build_lbvh(rt_ctx* c, uint32_t leaf_size, uint32_t flags) {
    const uint32_t n = c->T;
    const size_t N = n;
    int rc;
    if ((rc = ensure(c, c->keys[0], 4 * N)) || (rc = ensure(c, c->keys[1], 4 * N)) ||
        (rc = ensure(c, c->vals[0], 4 * N)) || (rc = ensure(c, c->vals[1], 4 * N)) ||
        (rc = ensure(c, c->leafbox, 24 * N)) || (rc = ensure(c, c->inner, 6 * 4 * N)) ||
        (rc = ensure(c, c->innerbox, 24 * N)) ||
        (rc = ensure(c, c->queue[0], 16 * (N / 2 + 2))) || (rc = ensure(c, c->queue[1], 16 * (N / 2 + 2))) ||
        (rc = ensure(c, c->counters, 4 * (CTR_CUM + PLOC_MAX_ITERS + 2))) || (rc = ensure(c, c->nodes, sizeof(rt_bvh4_node) * (N + 1))) ||
        (rc = ensure(c, c->tris, sizeof(rt_tri_rec) * N)))
        return rc;
    hipStream_t st = c->stream;
    RT_HIP(c, hipEventRecord(c->ev[0], st));

    float ext[3], inv[3];
    for (int k = 0; k < 3; ++k) {
        ext[k] = c->scene_hi[k] - c->scene_lo[k];
        inv[k] = ext[k] > 0.0f ? 1.0f / ext[k] : 0.0f;
    }
    float smax = 0.0f;
    for (int k = 0; k < 3; ++k) smax = fmaxf(smax, fmaxf(fabsf(c->scene_lo[k]), fabsf(c->scene_hi[k])));
    // outward padding of every box: 2^-16 of the largest coordinate magnitude
    const float pad = smax * 1.52587890625e-05f + 1e-30f;

    uint32_t* keys[2] = {(uint32_t*)c->keys[0].ptr, (uint32_t*)c->keys[1].ptr};
    uint32_t* vals[2] = {(uint32_t*)c->vals[0].ptr, (uint32_t*)c->vals[1].ptr};
    float* leafbox = (float*)c->leafbox.ptr;
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, (const float*)c->pos.ptr,
                       (const uint32_t*)c->idx.ptr, c->scene_lo[0], c->scene_lo[1], c->scene_lo[2], inv[0], inv[1],
                       inv[2], keys[0], vals[0], leafbox);

    int cur = 0;
    if ((rc = radix_sort_pairs(c, keys, vals, n, 4, &cur))) return rc;

    uint32_t* ib = (uint32_t*)c->inner.ptr;
    Inner in{ib, ib + N, ib + 2 * N, ib + 3 * N, ib + 4 * N, ib + 5 * N};
    float* innerbox = (float*)c->innerbox.ptr;
    rt_bvh4_node* nodes = (rt_bvh4_node*)c->nodes.ptr;
    uint32_t* counters = (uint32_t*)c->counters.ptr;

    const uint32_t* tv = vals[cur];  // triangle ids in leaf order
    uint32_t depth = 1, num_nodes = 1, num_leaves = 1;
    std::vector<uint32_t> level_start = {0u, 1u};  // first node of each level (the collapse allocates level by level)
    if (n <= leaf_size) {
        hipLaunchKernelGGL(k_root_leaf, dim3(1), dim3(1), 0, st, n, (const float*)leafbox, (const uint32_t*)vals[cur],
                           pad, nodes);
    } else if (flags & RT_BUILD_CLUSTER) {
        if ((rc = build_ploc(c, n, vals[cur], vals[cur ^ 1], keys, leafbox, in, innerbox))) return rc;
        tv = vals[cur ^ 1];
    } else {
        hipLaunchKernelGGL(k_karras, dim3(blocks_for(n - 1, 256)), dim3(256), 0, st, (const uint32_t*)keys[cur], (int)n,
                           in);
        const uint32_t nb = (n + BOX_BLOCK - 1) / BOX_BLOCK;
        uint32_t levels = 1;
        while ((1u << levels) <= nb) ++levels;  // table levels 0 .. floor(log2(nb))
        if ((rc = ensure(c, c->boxtab, 24ull * N + 24ull * nb * levels))) return rc;
        float* sbox = (float*)c->boxtab.ptr;
        float* table = sbox + 6ull * N;
        hipLaunchKernelGGL(k_sorted_boxes, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, (const uint32_t*)vals[cur],
                           (const float*)leafbox, sbox);
        hipLaunchKernelGGL(k_block_boxes, dim3(blocks_for(nb, 256)), dim3(256), 0, st, n, nb, (const float*)sbox, table);
        for (uint32_t k = 1; k < levels; ++k)
            hipLaunchKernelGGL(k_box_table_level, dim3(blocks_for(nb, 256)), dim3(256), 0, st, nb, 1u << (k - 1),
                               (const float*)(table + 6ull * nb * (k - 1)), table + 6ull * nb * k);
        hipLaunchKernelGGL(k_node_boxes, dim3(blocks_for(n - 1, 256)), dim3(256), 0, st, n - 1, in, (const float*)sbox,
                           (const float*)table, nb, innerbox);
    }
    hipLaunchKernelGGL(k_tri_records, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, tv, (const float*)c->pos.ptr,
                       (const uint32_t*)c->idx.ptr, (rt_tri_rec*)c->tris.ptr);
    if (n > leaf_size) {
        uint32_t init[6] = {1u, 1u, 0u, 0u, 0u, 0u};  // root node; queue 0 holds the root's entry
        RT_HIP(c, hipMemcpyAsync(counters, init, sizeof(init), hipMemcpyHostToDevice, st));
        uint4 root = make_uint4(0u, 0u, 0xFFFFFFFFu, 0u);
        RT_HIP(c, hipMemcpyAsync(c->queue[0].ptr, &root, sizeof(root), hipMemcpyHostToDevice, st));
        // One collapse level per launch, each reading its queue's length on
        // the device (a persistent grid strides over it).  The host copies the
        // counters to pinned memory after each level and reads the previous
        // level's copy, so it never drains the stream; the level launched
        // after the last non-empty one finds an empty queue and does nothing.
        if ((rc = ensure_snap(c))) return rc;
        const uint32_t max_nodes = (uint32_t)(N + 1), max_queue = (uint32_t)(N / 2 + 2);
        const uint32_t cgrid = 8u * (uint32_t)c->num_cus;
        int q = 0;
        for (uint32_t level = 0;; ++level) {
            if (level >= 130) return fail(c, RT_E_DEVICE, "rt_build: BVH deeper than 128 levels");
            const int in_slot = q ? CTR_Q1 : CTR_Q0, out_slot = q ? CTR_Q0 : CTR_Q1;
            hipLaunchKernelGGL(k_level_prep, dim3(1), dim3(64), 0, st, counters, in_slot, out_slot);
            hipLaunchKernelGGL(k_collapse, dim3(cgrid), dim3(COLLAPSE_BLOCK), 0, st, (const uint4*)c->queue[q].ptr,
                               (const uint32_t*)&counters[in_slot], (uint4*)c->queue[q ^ 1].ptr, in,
                               tv, (const float*)leafbox, (const float*)innerbox, leaf_size,
                               c->bvh_leaf_alpha, pad, nodes, max_nodes, max_queue, counters, out_slot);
            RT_HIP(c, hipGetLastError());
            RT_HIP(c, hipMemcpyAsync(c->snap + 16 * (level & 1), counters, 8 * sizeof(uint32_t),
                                     hipMemcpyDeviceToHost, st));
            RT_HIP(c, hipEventRecord(c->snap_ev[level & 1], st));
            q ^= 1;
            if (level == 0) continue;
            RT_HIP(c, hipEventSynchronize(c->snap_ev[(level - 1) & 1]));
            const uint32_t* h = c->snap + 16 * ((level - 1) & 1);
            if (h[CTR_ERROR]) return fail(c, RT_E_DEVICE, "rt_build: BVH collapse exceeded its structural bounds");
            // after level l - 1, the nodes allocated so far are levels 0 .. l
            level_start.push_back(h[CTR_NODES]);
            // the queue level - 1 appended to is the one this level read
            if (h[(level - 1) & 1 ? CTR_Q0 : CTR_Q1] == 0) break;
        }
        uint32_t h[8];
        RT_HIP(c, hipMemcpyAsync(h, counters, sizeof(h), hipMemcpyDeviceToHost, st));
        RT_HIP(c, hipStreamSynchronize(st));
        num_nodes = h[CTR_NODES];
        num_leaves = h[CTR_LEAVES];
        depth = h[CTR_DEPTH];
        if (h[CTR_ERROR] || num_nodes > max_nodes)
            return fail(c, RT_E_DEVICE, "rt_build: BVH collapse exceeded its structural bounds");
    }
    RT_HIP(c, hipGetLastError());
    if ((rc = ensure(c, c->qnodes, sizeof(rt_bvh4q_node) * ((size_t)num_nodes + 1)))) return rc;
    hipLaunchKernelGGL(k_quantize, dim3(blocks_for(num_nodes, 256)), dim3(256), 0, st,
                       (const rt_bvh4_node*)c->nodes.ptr, num_nodes, (rt_bvh4q_node*)c->qnodes.ptr);
    RT_HIP(c, hipGetLastError());
    RT_HIP(c, hipEventRecord(c->ev[1], st));
    RT_HIP(c, hipEventSynchronize(c->ev[1]));
    float ms = 0.0f;
    RT_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));

    c->num_nodes = num_nodes;
    c->bvh_depth = depth;
    if (level_start.back() != num_nodes) level_start.push_back(num_nodes);
    set_coop_splits(c, level_start);
    c->built = true;
    c->info.num_nodes = num_nodes;
    c->info.depth = depth;
    c->info.num_triangles = n;
    c->info.build_ms = ms;
    c->info.leaf_size = leaf_size;
    c->info.num_leaves = num_leaves;
    return RT_OK;
}

}  // namespace rt
