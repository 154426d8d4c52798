// nrc.hip — learned radiance cache: a small MLP on BF16 MFMA with FP32
// accumulation, trained online (Mueller, Rousselle, Novak, Keller, "Real-time
// neural radiance caching for path tracing", SIGGRAPH 2021).  It stands in
// for the reference renderer's DDGI indirect term (res/shader/
// majercik2019_final_raygen.rgen:43-91,124), which has no software-renderer
// counterpart.
//
// Network: input = frequency encoding of the position normalised to the
// scene box (12 frequencies x sin/cos x 3 axes = 72) + normal (3) + incoming
// direction (3) = 78, padded to 96; layers 96->64, 4 x (64->64), 64->16 (3
// used); ReLU on the hidden layers, linear output.  Loss: relative L2,
// (y - t)^2 / (y_sg^2 + 0.01), averaged over samples (y_sg = y without
// gradient).  Optimiser: Adam on FP32 master weights; the BF16 copy the
// kernels read is refreshed after every step.
//
// Kernels:
//   k_mlp_infer  4 waves, weights -> LDS once, then persistent tiles of 128
//                samples (32 per wave) carried through encoding and 6 layers
//                in registers
//   k_mlp_train  8 waves (4 for small batches), tiles of 16 samples per
//                wave, at most one workgroup per CU striding over the tiles:
//                forward and the activation gradients in registers, the
//                weight gradients of each layer from a [feature][sample]
//                staging buffer in LDS; each workgroup sums its own partial
//                gradient over its tiles in registers and writes it once,
//                through LDS as whole 256-byte rows of 64-parameter blocks
//                (no atomics, no read-modify-write of global memory)
//   k_adam       sums the partials in a fixed order (one run per wave,
//                then their sum: deterministic), Adam update, BF16 refresh
// Every GEMM is a sum of 16x16 output tiles of v_mfma_f32_16x16x32_bf16.
#include <math.h>
#include <stdint.h>

#include <vector>

#include "rt_internal.hpp"
#include "rt_core.hpp"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

namespace rt {
namespace nrc {

constexpr int L = 6;                                   // layers
constexpr int KD[L] = {96, 64, 64, 64, 64, 64};        // input width (padded)
constexpr int ND[L] = {64, 64, 64, 64, 64, 16};        // output width (padded)
constexpr int K0_VALID = 78, N5_VALID = 3;
constexpr int FREQS = 12;

// fp32 parameter layout: [W0 b0 W1 b1 ... W5 b5], W_l row-major [n][k]
constexpr int offw(int l) { int o = 0; for (int i = 0; i < l; ++i) o += ND[i] * KD[i] + ND[i]; return o; }
constexpr int offb(int l) { return offw(l) + ND[l] * KD[l]; }
constexpr int NPARAM = offw(L);
// bf16 weight layout: [W0 W1 ... W5]
constexpr int offwb(int l) { int o = 0; for (int i = 0; i < l; ++i) o += ND[i] * KD[i]; return o; }
constexpr int NWB = offwb(L);
static_assert(NPARAM == 23888, "parameter count");
// Per-workgroup partial gradients are stored by blocks of 64 parameters:
// block b of every workgroup's partial, then block b + 1, so that the sum
// over workgroups of one block reads one contiguous run.
constexpr int PART_BLOCKS = (NPARAM + 63) / 64;
constexpr int PART_STRIDE = 64 * PART_BLOCKS;  // floats per workgroup, the last block padded
__device__ __forceinline__ size_t part_index(int block, uint32_t wg, uint32_t nparts) {
    return ((size_t)block * nparts + wg) * 64;
}
static_assert(NWB == 23552, "bf16 weight count");


struct Geo {
    float lo[3], inv[3];
};

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ __bf16 tobf(float x) { return (__bf16)x; }
// a and b rounded to BF16 (to nearest even) in one dword, a in the low half
// (one v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t pack2(float a, float b) {
    const f32x2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}
// max(x, 0) of both BF16 halves as one packed 16-bit integer max: a BF16
// with the sign bit set is a negative integer (-0 gives +0)
__device__ __forceinline__ uint32_t relu2(uint32_t x) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2, x), s16x2{0, 0}));
}
// per BF16 half: g where the ReLU output h (never negative) is non-zero,
// else +0; g times min(h, 1) as packed 16-bit integers
__device__ __forceinline__ uint32_t mask2(uint32_t g, uint32_t h) {
    const u16x2 m = __builtin_elementwise_min(__builtin_bit_cast(u16x2, h), u16x2{1, 1});
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, g) * m);
}
// ReLU of the 4 consecutive outputs an accumulator holds, as BF16 pairs:
// slots d, d + 1 of a column-operand fragment
__device__ __forceinline__ void relu_tile(const f32x4& a, u32x4& x, int d) {
    x[d] = relu2(pack2(a[0], a[1]));
    x[d + 1] = relu2(pack2(a[2], a[3]));
}

// Features of k block ks that lane group kg supplies for one sample (the
// input layer's weight columns are stored in this order).  kg < 3: sin and
// cos of pi x 2^(4 ks + j), j = 0..3, interleaved, for the sample's
// coordinate x on axis kg normalised to the scene box: an anchor from the
// hardware sine and cosine (v_sin_f32 / v_cos_f32, argument in revolutions)
// of the fractional part of x 2^(4 ks - 1) (the scaling and the subtraction
// of the floor are exact), then three angle doublings; the error stays far
// under the BF16 step the features are rounded to (the inference and
// gradient tests compare against the oracle's exact encoding).  kg = 3: the
// normal and the direction (nd) in block 0, zero padding
// after.
__device__ __forceinline__ bf16x8 feature_block(int kg, int ks, float x, const float (&nd)[6]) {
    u32x4 v;
    if (kg < 3) {
        const float y = ldexpf(x, 4 * ks - 1);
        const float f = y - floorf(y);  // revolutions
        float sn = __builtin_amdgcn_sinf(f), cs = __builtin_amdgcn_cosf(f);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            v[j] = pack2(sn, cs);
            const float s2 = 2.0f * sn * cs, c2 = (cs - sn) * (cs + sn);
            sn = s2;
            cs = c2;
        }
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = ks == 0 && j < 3 ? pack2(nd[2 * j], nd[2 * j + 1]) : 0u;
    }
    return __builtin_bit_cast(bf16x8, v);
}

// The inputs lane group kg needs of sample sm (zeros past the end; n >= 1):
// the coordinate on axis kg normalised to the scene box (kg < 3), the normal
// and the direction (used by kg = 3).  Every lane issues the same loads.
__device__ __forceinline__ void load_features(const float* __restrict__ feats, uint32_t sm, uint32_t n, int kg,
                                              const Geo& G, float& xa, float (&nd)[6]) {
    const bool v = sm < n;
    const float* f = feats + (size_t)(v ? sm : 0u) * 9;
    const int a = kg < 3 ? kg : 0;
    const float lo = a == 0 ? G.lo[0] : (a == 1 ? G.lo[1] : G.lo[2]);
    const float inv = a == 0 ? G.inv[0] : (a == 1 ? G.inv[1] : G.inv[2]);
    const float x = f[a];
    xa = (v && kg < 3) ? (x - lo) * inv : 0.0f;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const float e = f[3 + j];
        nd[j] = v ? e : 0.0f;
    }
}

// ---- inference ------------------------------------------------------------
//
// 256 threads (4 waves) share one copy of the BF16 weights in LDS; each wave
// owns 32 samples of a 128-sample tile and keeps their activations in
// registers from the encoding to the output.  Every product is computed
// transposed, C^T = W X^T with the weights as the row operand: on
// v_mfma_f32_16x16x32_bf16 lane (mi = l % 16, kg = l / 16) then holds outputs
// 16 nt + 4 kg + i (i < 4) of sample 16 mt + mi, and the next layer's column
// operand wants, for its 32-wide k block ks, k slots 8 kg + j of the same
// sample.  A product sums over k in any order, so each layer's weight
// columns are stored permuted (slot 8 kg + j of block ks holds input
// 32 ks + 4 kg + j for j < 4 and 32 ks + 16 + 4 kg + j - 4 otherwise): the
// slots a lane must supply are exactly the outputs it holds, and a layer's
// ReLU output becomes the next layer's operand without leaving the lane.  The
// input layer's columns are permuted the same way to the encoding: lanes of
// group kg < 3 compute axis kg's 24 sin/cos features (block ks = anchor ks),
// group 3 the normal, direction and zero padding.
constexpr int INF_NT = 256;                       // threads per workgroup
constexpr int INF_MT = 2;                         // 16-sample groups a wave carries through the layers
constexpr int INF_TILE = INF_NT / 64 * 16 * INF_MT;  // samples per workgroup tile

// Every layer's BF16 weights are kept in the order the kernels read them, so
// that a workgroup fills LDS with one straight copy:
//  - forward, [W0 .. W5]: per layer, operand fragments of 16 rows x 32 slots
//    (blk), slots permuted so that a lane's 8 slots of one row are the inputs
//    it holds (enc_slot for the input layer, wslot for the others);
//  - transposed, [W1^T .. W5^T] (training only): per layer, fragments of 16
//    rows (inputs) x 32 slots (outputs), slot wslot(o) of row k holding W[o][k],
//    the row operand of the activation gradients W^T dZ; the output layer's
//    16 outputs fill half of one block, the other slots are zero.
constexpr int offwt(int l) { return NWB + (l - 1) * 64 * 64; }  // l = 1..5
constexpr int NWALL = NWB + 4 * 64 * 64 + 64 * 32;

// index of slot p (0..K-1) of row n in a matrix of N rows stored as 16 x 32
// operand fragments (k block major, then row tile), each fragment in lane
// order: lane (mi, kg) of a v_mfma_f32_16x16x32_bf16 operand, which supplies
// row mi and slots 8 kg .. 8 kg + 7, finds them as 16 bytes at lane index
// 16 kg + mi, so a wave reads a fragment as one contiguous kilobyte (no LDS
// bank conflicts)
__device__ __forceinline__ int blk(int n, int p, int N) {
    return (((p >> 5) * (N >> 4) + (n >> 4)) * 64 + ((p >> 3) & 3) * 16 + (n & 15)) * 8 + (p & 7);
}
// slot of encoding feature k (0..95) in the input layer's rows: slot
// (ks, 8 kg + j) holds feature 24 kg + 8 ks + j
__device__ __forceinline__ int enc_slot(int k) {
    const int kg = k / 24, r = k % 24;
    return 32 * (r >> 3) + 8 * kg + (r & 7);
}
// slot of input k in the rows of a layer l >= 1: slot 8 kg + j of block ks
// holds input 32 ks + 4 kg + j for j < 4 and 32 ks + 16 + 4 kg + j - 4
// otherwise
__device__ __forceinline__ int wslot(int k) {
    const int q = k & 31;
    return (k & ~31) + 8 * ((q & 15) >> 2) + 4 * (q >> 4) + (q & 3);
}

// global -> LDS copy of the first COUNT (multiple of 8) BF16 weights: every
// thread's loads are issued before its first LDS write
template <int NTH, int COUNT>
__device__ __forceinline__ void copy_weights(const __bf16* __restrict__ Wb, __bf16* sW) {
    constexpr int V = COUNT / 8, IT = (V + NTH - 1) / NTH;
    u32x4 r[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int c = min((int)threadIdx.x + i * NTH, V - 1);
        r[i] = ((const u32x4*)Wb)[c];
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int c = threadIdx.x + i * NTH;
        if (c < V) ((u32x4*)sW)[c] = r[i];
    }
}

// acc[mt][nt] = (FIRST ? bias[nt] : acc[mt][nt]) + W[16 nt.., block ks] . X[mt]^T
// for the wave's samples (the bias enters as the first product's addend)
template <int N, int NT_, bool FIRST, int MT = INF_MT>
__device__ __forceinline__ void layer_block(const __bf16* sWl, int ks, const bf16x8 (&x)[MT], int lane,
                                            const f32x4 (&bias)[NT_], f32x4 (&acc)[MT][NT_]) {
    const int mi = lane & 15, kg = lane >> 4;
    bf16x8 w[NT_];
#pragma unroll
    for (int nt = 0; nt < NT_; ++nt) w[nt] = *(const bf16x8*)(sWl + blk(16 * nt + mi, 32 * ks + 8 * kg, N));
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT_; ++nt)
            acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[nt], x[mt], FIRST ? bias[nt] : acc[mt][nt], 0,
                                                                  0, 0);
}

// The biases in LDS, layer l's at sB[64 l ..] (the output layer's 16 last):
// a lane reads them with one base address and constant offsets.
constexpr int NBIAS = (L - 1) * 64 + 16;
template <int NTH>
__device__ __forceinline__ void copy_bias(const float* __restrict__ Wf, float* sB) {
#pragma unroll
    for (int e = threadIdx.x; e < NBIAS; e += NTH) {
        int src = 0;
#pragma unroll
        for (int l = 0; l < L; ++l)
            if ((e >> 6) == l) src = offb(l) + (e & 63);
        sB[e] = Wf[src];
    }
}

// bias of outputs 16 nt + 4 kg + i (i < 4) of layer l
template <int NT_>
__device__ __forceinline__ void load_bias(const float* sB, int l, int kg, f32x4 (&b)[NT_]) {
#pragma unroll
    for (int nt = 0; nt < NT_; ++nt) b[nt] = *(const f32x4*)(sB + 64 * l + 16 * nt + 4 * kg);
}

// SCATTER: out[3 * idx[s] + c] += qw[3 * s + c] * max(y_c, 0) (atomically) in
// place of out[3 * s + c] = y_c.  n_dev (may be NULL): the sample count is
// read from device memory (at most n_max).
template <bool SCATTER>
__global__ void __launch_bounds__(INF_NT) __attribute__((amdgpu_waves_per_eu(3))) k_mlp_infer(const float* __restrict__ feats, uint32_t n_max,
                                                      const uint32_t* __restrict__ n_dev,
                                                      const __bf16* __restrict__ Wb, const float* __restrict__ Wf,
                                                      Geo G, float* __restrict__ out, const float* __restrict__ qw,
                                                      const uint32_t* __restrict__ idx) {
    const uint32_t n = n_dev ? min(*n_dev, n_max) : n_max;
    if (n == 0) return;
    __shared__ __attribute__((aligned(16))) __bf16 sW[NWB];
    __shared__ __attribute__((aligned(16))) float sB[NBIAS];
    copy_weights<INF_NT, NWB>(Wb, sW);
    copy_bias<INF_NT>(Wf, sB);
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int mi = lane & 15, kg = lane >> 4;
    const uint32_t tiles = (n + INF_TILE - 1) / INF_TILE;
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        // keeps the weight fragments' LDS loads inside the loop (hoisted, all
        // six layers' fragments would not fit in registers)
        __asm__ __volatile__("" ::: "memory");
        const uint32_t s0 = t * INF_TILE + 16 * INF_MT * w;  // this wave's first sample
        f32x4 acc[INF_MT][4];
        f32x4 bias[4];
        load_bias<4>(sB, 0, kg, bias);
        // input layer, one k block per encoding anchor
        float xa[INF_MT], nd[INF_MT][6];
#pragma unroll
        for (int mt = 0; mt < INF_MT; ++mt) load_features(feats, s0 + 16 * mt + mi, n, kg, G, xa[mt], nd[mt]);
#pragma unroll
        for (int ks = 0; ks < 3; ++ks) {
            bf16x8 x[INF_MT];
#pragma unroll
            for (int mt = 0; mt < INF_MT; ++mt) x[mt] = feature_block(kg, ks, xa[mt], nd[mt]);
            if (ks == 0) layer_block<64, 4, true>(sW + offwb(0), ks, x, lane, bias, acc);
            else layer_block<64, 4, false>(sW + offwb(0), ks, x, lane, bias, acc);
        }
        // hidden layers 1..4 and the output layer, activations in registers
        f32x4 o5[INF_MT][1];
#pragma unroll
        for (int l = 1; l < L; ++l) {
            u32x4 xd[2][INF_MT];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                for (int mt = 0; mt < INF_MT; ++mt) relu_tile(acc[mt][nt], xd[nt >> 1][mt], 2 * (nt & 1));
            bf16x8 x[2][INF_MT];
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
#pragma unroll
                for (int mt = 0; mt < INF_MT; ++mt) x[ks][mt] = __builtin_bit_cast(bf16x8, xd[ks][mt]);
            if (l < L - 1) {
                load_bias<4>(sB, l, kg, bias);
                layer_block<64, 4, true>(sW + offwb(l), 0, x[0], lane, bias, acc);
                layer_block<64, 4, false>(sW + offwb(l), 1, x[1], lane, bias, acc);
            } else {
                f32x4 b5[1];
                load_bias<1>(sB, l, kg, b5);
                layer_block<16, 1, true>(sW + offwb(l), 0, x[0], lane, b5, o5);
                layer_block<16, 1, false>(sW + offwb(l), 1, x[1], lane, b5, o5);
            }
        }
        // lanes with kg == 0 hold outputs 0..3 of sample 16 mt + mi
        if (kg == 0) {
#pragma unroll
            for (int mt = 0; mt < INF_MT; ++mt) {
                const uint32_t sm = s0 + 16 * mt + mi;
                if (sm < n) {
#pragma unroll
                    for (int c = 0; c < N5_VALID; ++c) {
                        const float y = o5[mt][0][c];
                        if (SCATTER) atomicAdd(&out[3 * (size_t)idx[sm] + c], qw[3 * (size_t)sm + c] * fmaxf(y, 0.0f));
                        else out[(size_t)sm * 3 + c] = y;
                    }
                }
            }
        }
    }
}

// ---- training ---------------------------------------------------------------
//
// A workgroup of NW waves (8, or 4 for small batches) takes tiles of 16 NW
// samples, 16 per wave, at most one workgroup per CU striding over the tiles.
// Per tile:
//  - forward as in inference (one 16-sample group per wave); the five ReLU
//    outputs stay in registers as column operands, the encoding is also
//    written to its own staging buffer, [feature][sample];
//  - loss, and the output gradient dZ5 in the same column-operand layout;
//  - backward, layer l = 5 .. 0: the wave writes dZ_l and the layer's input
//    H_{l-1} (l > 0) into the staging buffers transposed, [feature][sample]; computes dZ_{l-1} = (W_l^T dZ_l) masked by
//    H_{l-1} > 0 in registers (the transposed weights are the row operand and
//    dZ_l the column operand, so the product lands in the column-operand
//    layout again, slots permuted as in the forward); then, after a barrier,
//    its share of dW_l = sum over the tile's samples of dZ_l H_{l-1}^T from
//    the staging buffers (one 16-row block of outputs, a group of 16-column
//    input tiles).  A bias gradient is the same product against a column
//    operand that is one in column 0 and zero elsewhere; for the input layer
//    it is the encoding's padding feature ENC_ONE, staged as 1 (the forward
//    sees 0 there and its weights stay 0).
// Each thread sums its share of the gradient over the workgroup's tiles, in
// tile order, in registers and writes it once (no atomics, no
// read-modify-write of global memory): the result does not depend on
// scheduling.
constexpr int ENC_ONE = 78;

template <int NW>
struct TrainShape {
    static constexpr int NTH = 64 * NW;           // threads per workgroup
    static constexpr int TILE = 16 * NW;          // samples per tile
    static constexpr int SR = TILE + 8;           // staging row pitch (BF16): consecutive rows 16 bytes apart in bank order
    static constexpr int G = NW / 4;              // waves sharing a 16-row block of outputs
    static constexpr int BH = 4 / G;              // hidden layers: input tiles per wave
    static constexpr int B0 = 6 / G;              // input layer: input tiles per wave
    static constexpr int B5 = (5 + NW - 1) / NW;  // output layer: tiles per wave (4 weight tiles, 1 bias tile)
};

// writes the values of column-operand fragment v (k block ks) of the lane's
// sample column into S[feature][col], for features below nrows
__device__ __forceinline__ void stage_act(__bf16* S, int sr, const bf16x8& v, int ks, int kg, int col, int nrows) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int f = 32 * ks + (j < 4 ? 4 * kg + j : 16 + 4 * kg + j - 4);
        if (f < nrows) S[f * sr + col] = v[j];
    }
}

// 16 x 32 operand fragment of rows r0 .. r0 + 15 and samples 32 kk .. 32 kk + 31
// of a staging buffer
__device__ __forceinline__ bf16x8 stage_frag(const __bf16* S, int sr, int r0, int kk, int mi, int kg) {
    return *(const bf16x8*)(S + (r0 + mi) * sr + 32 * kk + 8 * kg);
}

template <int NW>
__global__ void __launch_bounds__(64 * NW) k_mlp_train(const float* __restrict__ feats, const float* __restrict__ targets,
                                                       uint32_t n, const __bf16* __restrict__ Wb,
                                                       const float* __restrict__ Wf, Geo G, float* __restrict__ partial,
                                                       float* __restrict__ loss_part) {
    using S = TrainShape<NW>;
    constexpr int SR = S::SR;
    constexpr int KS = S::TILE / 32;  // k steps of a weight-gradient product
    // one LDS block: the weights and the three staging buffers while the
    // tiles run, the workgroup's whole partial gradient (fp32) at the end
    constexpr int LDS_MAIN = 2 * (NWALL + 64 * SR + 64 * SR + 96 * SR);
    static_assert(LDS_MAIN >= 4 * PART_STRIDE, "the partial gradient must fit the LDS block");
    __shared__ __attribute__((aligned(16))) unsigned char smem[LDS_MAIN];
    __bf16* const sW = (__bf16*)smem;                   // forward and transposed weights
    __bf16* const sZ = sW + NWALL;                      // dZ_l, [output][sample]
    __bf16* const sH = sZ + 64 * SR;                    // H_{l-1}, [input][sample]
    __bf16* const sX = sH + 64 * SR;                    // the encoding, [feature][sample]
    __shared__ __attribute__((aligned(16))) float sB[NBIAS];
    __shared__ float sLoss[NW];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int mi = lane & 15, kg = lane >> 4;
    const int rb = w % 4, cg = w / 4;  // this wave's 16-row block of outputs and its group of input tiles
    const int col = 16 * w + mi;       // this lane's sample column in a tile
    const uint32_t tiles = (n + S::TILE - 1) / S::TILE;
    // a tile's inputs (coordinate, normal and direction, targets), loaded one
    // tile ahead: the first tile's before the weight copy, each later one's
    // at the start of the tile before it, so that the loads overlap work
    float nx_xa, nx_nd[6], nx_tg[N5_VALID];
    auto load_in = [&](uint32_t t) {
        const uint32_t sm = t * S::TILE + col;
        load_features(feats, sm, n, kg, G, nx_xa, nx_nd);
        const size_t r = sm < n ? sm : 0u;
#pragma unroll
        for (int c = 0; c < N5_VALID; ++c) {
            const float e = targets[r * 3 + c];
            nx_tg[c] = sm < n ? e : 0.0f;
        }
    };
    load_in(blockIdx.x);
    copy_weights<S::NTH, NWALL>(Wb, sW);
    copy_bias<S::NTH>(Wf, sB);
    bf16x8 ones;  // column operand of the bias products
#pragma unroll
    for (int j = 0; j < 8; ++j) ones[j] = tobf(mi == 0 ? 1.0f : 0.0f);
    // this thread's share of the workgroup's gradient: input layer tiles,
    // hidden layer tiles (the last one the bias tile, used when cg == 0),
    // output layer tiles
    f32x4 g0[S::B0], gh[L - 2][S::BH + 1], g5[S::B5];
#pragma unroll
    for (int j = 0; j < S::B0; ++j) g0[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int l = 0; l < L - 2; ++l)
#pragma unroll
        for (int j = 0; j <= S::BH; ++j) gh[l][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int j = 0; j < S::B5; ++j) g5[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float lacc = 0.0f;
    const float two_over_n = 2.0f / (float)n;
    __syncthreads();  // weights in LDS
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        // keeps the weight fragments' LDS loads inside the loop
        __asm__ __volatile__("" ::: "memory");
        const bool valid = tile * S::TILE + col < n;
        const float xa = nx_xa;
        float nd[6], tg[N5_VALID];
#pragma unroll
        for (int j = 0; j < 6; ++j) nd[j] = nx_nd[j];
#pragma unroll
        for (int c = 0; c < N5_VALID; ++c) tg[c] = nx_tg[c];
        if (tile + gridDim.x < tiles) load_in(tile + gridDim.x);
        // ---- forward
        f32x4 acc[1][4];
        f32x4 bias[4];
        load_bias<4>(sB, 0, kg, bias);
#pragma unroll
        for (int ks = 0; ks < 3; ++ks) {
            const bf16x8 xk[1] = {feature_block(kg, ks, xa, nd)};
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int f = 24 * kg + 8 * ks + j;
                sX[f * SR + col] = f == ENC_ONE ? tobf(1.0f) : xk[0][j];
            }
            if (ks == 0) layer_block<64, 4, true, 1>(sW + offwb(0), ks, xk, lane, bias, acc);
            else layer_block<64, 4, false, 1>(sW + offwb(0), ks, xk, lane, bias, acc);
        }
        u32x4 h[L - 1][2];  // ReLU outputs of layers 0 .. 4, k blocks 0 and 1 (BF16 pairs)
        f32x4 o5[1][1];
#pragma unroll
        for (int l = 1; l < L; ++l) {
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) relu_tile(acc[0][nt], h[l - 1][nt >> 1], 2 * (nt & 1));
            const bf16x8 xk0[1] = {__builtin_bit_cast(bf16x8, h[l - 1][0])};
            const bf16x8 xk1[1] = {__builtin_bit_cast(bf16x8, h[l - 1][1])};
            if (l < L - 1) {
                load_bias<4>(sB, l, kg, bias);
                layer_block<64, 4, true, 1>(sW + offwb(l), 0, xk0, lane, bias, acc);
                layer_block<64, 4, false, 1>(sW + offwb(l), 1, xk1, lane, bias, acc);
            } else {
                f32x4 b5[1];
                load_bias<1>(sB, l, kg, b5);
                layer_block<16, 1, true, 1>(sW + offwb(l), 0, xk0, lane, b5, o5);
                layer_block<16, 1, false, 1>(sW + offwb(l), 1, xk1, lane, b5, o5);
            }
        }
        // ---- loss: lane (mi, kg) holds outputs 4 kg + i of its sample;
        // with q = (y - t) / (y^2 + 0.01), the loss term is (y - t) q and
        // dL/dy = q 2 / n on the 3 valid outputs, which are slots 0 .. 2 of
        // the output layer's gradient fragment
        float lsum = 0.0f, dy[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int c = 4 * kg + i;
            if (c < N5_VALID && valid) {
                const float y = o5[0][0][i];
                const float t = tg[i];
                const float r = y - t;
                const float q = r / (y * y + 0.01f);
                lsum += r * q;
                dy[i] = q * two_over_n;
            }
        }
        u32x4 dz[2] = {u32x4{pack2(dy[0], dy[1]), pack2(dy[2], dy[3]), 0u, 0u}, u32x4{0u, 0u, 0u, 0u}};
        for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
        if (lane == 0) sLoss[w] = lsum;

        // ---- backward
#pragma unroll
        for (int l = L - 1; l >= 0; --l) {
            stage_act(sZ, SR, __builtin_bit_cast(bf16x8, dz[0]), 0, kg, col, ND[l]);
            if (l < L - 1) stage_act(sZ, SR, __builtin_bit_cast(bf16x8, dz[1]), 1, kg, col, 64);
            if (l > 0) {
                stage_act(sH, SR, __builtin_bit_cast(bf16x8, h[l - 1][0]), 0, kg, col, 64);
                stage_act(sH, SR, __builtin_bit_cast(bf16x8, h[l - 1][1]), 1, kg, col, 64);
            }
            // dZ_{l-1} = (W_l^T dZ_l) masked by H_{l-1} > 0, in registers
            u32x4 dzn[2];
            if (l > 0) {
                f32x4 g[4];
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) g[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int ks = 0; ks < (l == L - 1 ? 1 : 2); ++ks)
#pragma unroll
                    for (int mt = 0; mt < 4; ++mt) {
                        const bf16x8 a = *(const bf16x8*)(sW + offwt(l) + blk(16 * mt + mi, 32 * ks + 8 * kg, 64));
                        g[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(bf16x8, dz[ks]), g[mt],
                                                                        0, 0, 0);
                    }
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) {
                    const int ks = mt >> 1, d = 2 * (mt & 1);
                    dzn[ks][d] = mask2(pack2(g[mt][0], g[mt][1]), h[l - 1][ks][d]);
                    dzn[ks][d + 1] = mask2(pack2(g[mt][2], g[mt][3]), h[l - 1][ks][d + 1]);
                }
            }
            __syncthreads();  // staging written
            if (l == L - 1 && threadIdx.x == 0)
                for (int q = 0; q < NW; ++q) lacc += sLoss[q];
            // this wave's share of dW_l (and the bias) over the tile's samples:
            // every operand fragment first, then the products
            if (l == L - 1) {
                bf16x8 af[KS], bfr[KS][S::B5];
#pragma unroll
                for (int kk = 0; kk < KS; ++kk) {
                    af[kk] = stage_frag(sZ, SR, 0, kk, mi, kg);
#pragma unroll
                    for (int j = 0; j < S::B5; ++j) {
                        const int b = w + NW * j;
                        bfr[kk][j] = b < 4 ? stage_frag(sH, SR, 16 * b, kk, mi, kg) : ones;
                    }
                }
#pragma unroll
                for (int kk = 0; kk < KS; ++kk)
#pragma unroll
                    for (int j = 0; j < S::B5; ++j)
                        if (w + NW * j <= 4) g5[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[kk], bfr[kk][j], g5[j], 0, 0, 0);
            } else if (l == 0) {
                bf16x8 af[KS], bfr[KS][S::B0];
#pragma unroll
                for (int kk = 0; kk < KS; ++kk) {
                    af[kk] = stage_frag(sZ, SR, 16 * rb, kk, mi, kg);
#pragma unroll
                    for (int j = 0; j < S::B0; ++j) bfr[kk][j] = stage_frag(sX, SR, 16 * (cg * S::B0 + j), kk, mi, kg);
                }
#pragma unroll
                for (int kk = 0; kk < KS; ++kk)
#pragma unroll
                    for (int j = 0; j < S::B0; ++j) g0[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[kk], bfr[kk][j], g0[j], 0, 0, 0);
            } else {
                bf16x8 af[KS], bfr[KS][S::BH];
#pragma unroll
                for (int kk = 0; kk < KS; ++kk) {
                    af[kk] = stage_frag(sZ, SR, 16 * rb, kk, mi, kg);
#pragma unroll
                    for (int j = 0; j < S::BH; ++j) bfr[kk][j] = stage_frag(sH, SR, 16 * (cg * S::BH + j), kk, mi, kg);
                }
#pragma unroll
                for (int kk = 0; kk < KS; ++kk) {
#pragma unroll
                    for (int j = 0; j < S::BH; ++j)
                        gh[l - 1][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[kk], bfr[kk][j], gh[l - 1][j], 0, 0, 0);
                    if (cg == 0)
                        gh[l - 1][S::BH] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[kk], ones, gh[l - 1][S::BH], 0, 0, 0);
                }
            }
            __syncthreads();  // staging read
            if (l > 0) {
                dz[0] = dzn[0];
                dz[1] = dzn[1];
            }
        }
    }  // tiles
    // the workgroup's partial gradient: first into LDS in parameter order
    // (lane (mi, kg) of tile (row block r, column tile b) holds rows
    // 16 r + 4 kg + i, column 16 b + mi), then out as whole 256-byte rows of
    // the blocked layout.  (The lane index passes through an empty asm so
    // that the compiler does not compute these addresses before the loop
    // and hold them across it.)
    __syncthreads();  // every wave is done with the weights and the staging buffers
    float* const sP = (float*)smem;
    auto put = [&](int q, float v) { sP[q] = v; };
    int lane_o = lane;
    __asm__ __volatile__("" : "+v"(lane_o));
    const int mo = lane_o & 15, ko = lane_o >> 4;
#pragma unroll
    for (int j = 0; j < S::B0; ++j) {
        const int c = 16 * (cg * S::B0 + j) + mo;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = 16 * rb + 4 * ko + i;
            if (c == ENC_ONE) {
                put(offb(0) + r, g0[j][i]);
                put(offw(0) + r * KD[0] + c, 0.0f);
            } else {
                put(offw(0) + r * KD[0] + c, g0[j][i]);
            }
        }
    }
#pragma unroll
    for (int l = 1; l < L - 1; ++l) {
#pragma unroll
        for (int j = 0; j < S::BH; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                put(offw(l) + (16 * rb + 4 * ko + i) * 64 + 16 * (cg * S::BH + j) + mo, gh[l - 1][j][i]);
        if (cg == 0 && mo == 0)
#pragma unroll
            for (int i = 0; i < 4; ++i) put(offb(l) + 16 * rb + 4 * ko + i, gh[l - 1][S::BH][i]);
    }
#pragma unroll
    for (int j = 0; j < S::B5; ++j) {
        const int b = w + NW * j;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (b < 4) put(offw(L - 1) + (4 * ko + i) * 64 + 16 * b + mo, g5[j][i]);
            else if (b == 4 && mo == 0) put(offb(L - 1) + 4 * ko + i, g5[j][i]);
        }
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = lacc;
    if (threadIdx.x < PART_STRIDE - NPARAM) sP[NPARAM + threadIdx.x] = 0.0f;  // the last block's padding
    __syncthreads();
    for (int f = threadIdx.x; f < PART_STRIDE / 4; f += S::NTH) {
        const int p = 4 * f;
        // streaming stores: k_adam reads each row once, right after
        __builtin_nontemporal_store(*(const f32x4*)(sP + p),
                                    (f32x4*)(partial + part_index(p / 64, blockIdx.x, gridDim.x) + p % 64));
    }
}

// Adam with bias correction on the summed partial gradients (workgroup order,
// so the sum is deterministic); refreshes the bf16 weight copy.  With lr = 0
// it only writes the summed gradient (gradient queries).
// parameter p's BF16 copies (weights only; the biases stay FP32): the
// forward layout and, for layers 1 .. 5, the transposed one
__device__ __forceinline__ void store_bf16(__bf16* Wb, int p, float w) {
#pragma unroll
    for (int l = 0; l < L; ++l) {
        const int o = offw(l), nw = ND[l] * KD[l];
        if (p >= o && p < o + nw) {
            const int r = (p - o) / KD[l], k = (p - o) % KD[l];
            const __bf16 b = tobf(w);
            Wb[offwb(l) + blk(r, l == 0 ? enc_slot(k) : wslot(k), ND[l])] = b;
            if (l > 0) Wb[offwt(l) + blk(k, wslot(r), 64)] = b;
        }
    }
}

// A block of ADAM_NT threads serves 64 parameters: wave q sums partials
// q, q + QN, q + 2 QN, ... in order (each load one coalesced 256-byte row,
// all of them in flight together), then wave 0 adds the QN sums in wave
// order.
constexpr int ADAM_NT = 512;
constexpr int QN = ADAM_NT / 64;
__global__ void __launch_bounds__(ADAM_NT) k_adam(const float* __restrict__ partial, uint32_t nparts,
                                                  float* __restrict__ grad, float* __restrict__ Wf,
                                                  float* __restrict__ m, float* __restrict__ v,
                                                  __bf16* __restrict__ Wb, float lr, float b1, float b2, float eps,
                                                  float bc1, float bc2, int update) {
    __shared__ float s[QN][64];
    const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int p = blockIdx.x * 64 + lane;
    float g = 0.0f;
    if (p < NPARAM) {
        // up to 256 partials: every load of the wave in flight at once
        constexpr int PW = 256 / QN;
        float x[PW];
#pragma unroll
        for (int j = 0; j < PW; ++j) {
            const uint32_t w = q + QN * j;
            x[j] = w < nparts ? __builtin_nontemporal_load(partial + part_index(blockIdx.x, w, nparts) + lane) : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < PW; ++j) g += x[j];
        for (uint32_t w = q + QN * PW; w < nparts; w += QN) g += partial[part_index(blockIdx.x, w, nparts) + lane];
    }
    s[q][lane] = g;
    __syncthreads();
    if (q != 0 || p >= NPARAM) return;
    g = s[0][lane];
#pragma unroll
    for (int k = 1; k < QN; ++k) g += s[k][lane];
    grad[p] = g;
    if (!update) return;
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float wnew = Wf[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    Wf[p] = wnew;
    store_bf16(Wb, p, wnew);
}

__global__ void k_refresh_bf16(const float* __restrict__ Wf, __bf16* __restrict__ Wb) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    store_bf16(Wb, p, Wf[p]);
}

// batch sample j <- record b + stride * j
__global__ void k_gather(const float* __restrict__ feats, const float* __restrict__ tgt, uint32_t b, uint32_t stride,
                         uint32_t n, float* __restrict__ bf, float* __restrict__ bt) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const size_t r = (size_t)b + (size_t)stride * j;
#pragma unroll
    for (int k = 0; k < 9; ++k) bf[9 * (size_t)j + k] = feats[9 * r + k];
#pragma unroll
    for (int k = 0; k < 3; ++k) bt[3 * (size_t)j + k] = tgt[3 * r + k];
}

}  // namespace nrc
}  // namespace rt

// ---------------------------------------------------------------- C ABI
using namespace rt;
using namespace rt::nrc;

struct rt_nrc {
    rt_ctx* ctx = nullptr;
    rt_nrc_config cfg{};
    Geo geo{};
    DevBuf wf, wb, m, v, grad, partial, loss, feats, targets, out;
    uint32_t steps = 0;
    int infer_grid = 0;
};

namespace {

int nrc_fail(rt_nrc* s, int code, const std::string& msg) { return fail(s->ctx, code, msg); }

uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int stage(rt_nrc* s, DevBuf& b, const void* src, size_t bytes, int ptr_kind, const void** out) {
    if (ptr_kind == RT_PTR_DEVICE) {
        *out = src;
        return RT_OK;
    }
    int rc = ensure(s->ctx, b, bytes);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, s->ctx->stream);
    if (e != hipSuccess) return hip_fail(s->ctx, e, "hipMemcpyAsync");
    *out = b.ptr;
    return RT_OK;
}

// one gradient evaluation over n samples (and an Adam step when update)
int run_train(rt_nrc* s, const float* feats, const float* targets, uint32_t n, int ptr_kind, float* loss, float* grad_out,
              bool update) {
    rt_ctx* c = s->ctx;
    if (n == 0) return nrc_fail(s, RT_E_INVALID, "rt_nrc: no samples");
    if (!feats || !targets) return nrc_fail(s, RT_E_INVALID, "rt_nrc: NULL samples");
    const void *df, *dt;
    int rc;
    if ((rc = stage(s, s->feats, feats, (size_t)n * 36, ptr_kind, &df)) ||
        (rc = stage(s, s->targets, targets, (size_t)n * 12, ptr_kind, &dt)))
        return rc;
    // one workgroup per CU at most (142 KB of LDS with 8 waves, 111 KB with
    // 4); each strides over tiles, so k_adam sums at most num_cus partial
    // gradients.  8-wave tiles when they fill every CU, else 4-wave ones.
    const bool wide = (n + TrainShape<8>::TILE - 1) / TrainShape<8>::TILE >= (uint32_t)c->num_cus;
    const uint32_t tile = wide ? TrainShape<8>::TILE : TrainShape<4>::TILE;
    const uint32_t tiles = (n + tile - 1) / tile;
    const uint32_t wgs = tiles < (uint32_t)c->num_cus ? tiles : (uint32_t)c->num_cus;
    if ((rc = ensure(c, s->partial, (size_t)wgs * PART_STRIDE * 4)) || (rc = ensure(c, s->loss, (size_t)wgs * 4))) return rc;
    if (wide)
        hipLaunchKernelGGL(k_mlp_train<8>, dim3(wgs), dim3(TrainShape<8>::NTH), 0, c->stream, (const float*)df,
                           (const float*)dt, n, (const __bf16*)s->wb.ptr, (const float*)s->wf.ptr, s->geo,
                           (float*)s->partial.ptr, (float*)s->loss.ptr);
    else
        hipLaunchKernelGGL(k_mlp_train<4>, dim3(wgs), dim3(TrainShape<4>::NTH), 0, c->stream, (const float*)df,
                           (const float*)dt, n, (const __bf16*)s->wb.ptr, (const float*)s->wf.ptr, s->geo,
                           (float*)s->partial.ptr, (float*)s->loss.ptr);
    RT_HIP(c, hipGetLastError());
    float bc1 = 1.0f, bc2 = 1.0f;
    if (update) {
        s->steps += 1;
        bc1 = 1.0f - powf(s->cfg.beta1, (float)s->steps);
        bc2 = 1.0f - powf(s->cfg.beta2, (float)s->steps);
    }
    hipLaunchKernelGGL(k_adam, dim3((NPARAM + 63) / 64), dim3(ADAM_NT), 0, c->stream, (const float*)s->partial.ptr, wgs,
                       (float*)s->grad.ptr, (float*)s->wf.ptr, (float*)s->m.ptr, (float*)s->v.ptr,
                       (__bf16*)s->wb.ptr, s->cfg.lr, s->cfg.beta1, s->cfg.beta2, s->cfg.eps, bc1, bc2,
                       update ? 1 : 0);
    RT_HIP(c, hipGetLastError());
    // untimed: two event markers would cost a tenth of a short step
    c->last_op = 0;
    c->timing = rt_timing{};
    if (grad_out) RT_HIP(c, hipMemcpyAsync(grad_out, s->grad.ptr, NPARAM * 4, hipMemcpyDeviceToHost, c->stream));
    if (loss) {
        std::vector<float> lp(wgs);
        RT_HIP(c, hipMemcpyAsync(lp.data(), s->loss.ptr, (size_t)wgs * 4, hipMemcpyDeviceToHost, c->stream));
        RT_HIP(c, hipStreamSynchronize(c->stream));
        double acc = 0.0;
        for (float x : lp) acc += x;
        *loss = (float)(acc / ((double)n * N5_VALID));
    } else if (grad_out || ptr_kind == RT_PTR_HOST) {
        RT_HIP(c, hipStreamSynchronize(c->stream));
    }
    return RT_OK;
}

}  // namespace

namespace rt {
// used by the path tracer's cache mode (pathtrace.hip)
int nrc_infer_device(rt_nrc* s, const float* d_feats, uint32_t n, float* d_out) {
    rt_ctx* c = s->ctx;
    if (n == 0) return RT_OK;
    const uint32_t tiles = (n + INF_TILE - 1) / INF_TILE;
    const uint32_t grid = tiles < (uint32_t)s->infer_grid ? tiles : (uint32_t)s->infer_grid;
    hipLaunchKernelGGL(k_mlp_infer<false>, dim3(grid), dim3(INF_NT), 0, c->stream, d_feats, n, nullptr, (const __bf16*)s->wb.ptr,
                       (const float*)s->wf.ptr, s->geo, d_out, nullptr, nullptr);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}
int nrc_infer_scatter(rt_nrc* s, const float* d_feats, const float* d_w, const uint32_t* d_idx, uint32_t n,
                      float* d_acc, const uint32_t* d_n) {
    rt_ctx* c = s->ctx;
    if (n == 0) return RT_OK;
    const uint32_t tiles = (n + INF_TILE - 1) / INF_TILE;
    const uint32_t grid = tiles < (uint32_t)s->infer_grid ? tiles : (uint32_t)s->infer_grid;
    hipLaunchKernelGGL(k_mlp_infer<true>, dim3(grid), dim3(INF_NT), 0, c->stream, d_feats, n, d_n,
                       (const __bf16*)s->wb.ptr, (const float*)s->wf.ptr, s->geo, d_acc, d_w, d_idx);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}
int nrc_train_records(rt_nrc* s, const float* d_feats, const float* d_targets, uint32_t count, uint32_t batch,
                      uint32_t max_steps, uint32_t* steps, float* last_loss) {
    rt_ctx* c = s->ctx;
    *steps = 0;
    if (count == 0 || max_steps == 0 || batch == 0) return RT_OK;
    const uint32_t n = count < batch ? count : batch;
    const uint32_t stride = count / n;
    const uint32_t ns = stride < max_steps ? stride : max_steps;
    int rc;
    if ((rc = ensure(c, s->feats, (size_t)n * 36)) || (rc = ensure(c, s->targets, (size_t)n * 12))) return rc;
    for (uint32_t b = 0; b < ns; ++b) {
        hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, c->stream, d_feats, d_targets, b, stride, n,
                           (float*)s->feats.ptr, (float*)s->targets.ptr);
        RT_HIP(c, hipGetLastError());
        if ((rc = run_train(s, (const float*)s->feats.ptr, (const float*)s->targets.ptr, n, RT_PTR_DEVICE,
                            b + 1 == ns ? last_loss : nullptr, nullptr, true)))
            return rc;
    }
    *steps = ns;
    return RT_OK;
}
rt_ctx* nrc_ctx(const rt_nrc* s) { return s->ctx; }
void nrc_detach(rt_nrc* s) {
    rt_ctx* c = s->ctx;
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    DevBuf* all[] = {&s->wf, &s->wb, &s->m, &s->v, &s->grad, &s->partial, &s->loss, &s->feats, &s->targets, &s->out};
    for (DevBuf* b : all) {
        if (b->ptr) (void)hipFree(b->ptr);
        b->ptr = nullptr;
        b->bytes = 0;
    }
    for (size_t i = 0; i < c->caches.size(); ++i)
        if (c->caches[i] == s) {
            c->caches.erase(c->caches.begin() + (long)i);
            break;
        }
    s->ctx = nullptr;
}
int nrc_train_device(rt_nrc* s, const float* d_feats, const float* d_targets, uint32_t n) {
    return run_train(s, d_feats, d_targets, n, RT_PTR_DEVICE, nullptr, nullptr, true);
}
}  // namespace rt

extern "C" {

int rt_nrc_create(rt_ctx* c, const rt_nrc_config* cfg, rt_nrc** out) {
    if (!c) return RT_E_INVALID;
    if (!out) return fail(c, RT_E_INVALID, "rt_nrc_create: out is NULL");
    *out = nullptr;
    rt_nrc* s = new (std::nothrow) rt_nrc();
    if (!s) return fail(c, RT_E_NOMEM, "rt_nrc_create: host allocation failed");
    s->ctx = c;
    if (cfg) s->cfg = *cfg;
    if (s->cfg.lr == 0.0f) s->cfg.lr = 1e-2f;
    if (s->cfg.beta1 == 0.0f) s->cfg.beta1 = 0.9f;
    if (s->cfg.beta2 == 0.0f) s->cfg.beta2 = 0.99f;
    if (s->cfg.eps == 0.0f) s->cfg.eps = 1e-8f;
    float lo[3], hi[3];
    bool given = false;
    for (int k = 0; k < 3; ++k) given |= s->cfg.bounds_lo[k] != s->cfg.bounds_hi[k];
    for (int k = 0; k < 3; ++k) {
        lo[k] = given ? s->cfg.bounds_lo[k] : c->scene_lo[k];
        hi[k] = given ? s->cfg.bounds_hi[k] : c->scene_hi[k];
        const float ext = hi[k] - lo[k];
        s->geo.lo[k] = lo[k];
        s->geo.inv[k] = ext > 0.0f ? 1.0f / ext : 0.0f;
    }
    (void)hipSetDevice(c->device);
    int rc;
    if ((rc = ensure(c, s->wf, NPARAM * 4)) || (rc = ensure(c, s->wb, NWALL * 2)) || (rc = ensure(c, s->m, NPARAM * 4)) ||
        (rc = ensure(c, s->v, NPARAM * 4)) || (rc = ensure(c, s->grad, NPARAM * 4))) {
        rt_nrc_destroy(s);
        return rc;
    }
    // He-uniform weights on the valid inputs / outputs, zero biases and padding
    std::vector<float> w(NPARAM, 0.0f);
    uint64_t st = s->cfg.seed ^ 0x5EEDull;
    for (int l = 0; l < L; ++l) {
        const int kv = (l == 0) ? K0_VALID : KD[l];
        const int nv = (l == L - 1) ? N5_VALID : ND[l];
        const float a = sqrtf(6.0f / (float)kv);
        for (int nn = 0; nn < nv; ++nn)
            for (int k = 0; k < kv; ++k) {
                const double u = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
                w[offw(l) + nn * KD[l] + k] = (float)((2.0 * u - 1.0) * a);
            }
    }
    hipError_t e = hipMemcpy(s->wf.ptr, w.data(), NPARAM * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(s->m.ptr, 0, NPARAM * 4);
    if (e == hipSuccess) e = hipMemset(s->v.ptr, 0, NPARAM * 4);
    // the transposed output layer's unused slots stay zero
    if (e == hipSuccess) e = hipMemset(s->wb.ptr, 0, NWALL * 2);
    if (e != hipSuccess) {
        rt_nrc_destroy(s);
        return hip_fail(c, e, "rt_nrc_create");
    }
    hipLaunchKernelGGL(k_refresh_bf16, dim3((NPARAM + 255) / 256), dim3(256), 0, c->stream, (const float*)s->wf.ptr,
                       (__bf16*)s->wb.ptr);
    int per_cu = 0;
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mlp_infer<false>, INF_NT, 0);
    s->infer_grid = c->num_cus * (per_cu > 0 ? per_cu : 1);
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        rt_nrc_destroy(s);
        return hip_fail(c, e, "rt_nrc_create");
    }
    c->caches.push_back(s);
    *out = s;
    return RT_OK;
}

void rt_nrc_destroy(rt_nrc* s) {
    if (!s) return;
    nrc_detach(s);
    delete s;
}

int rt_nrc_params(rt_nrc* s, float* params, uint32_t* count, int set) {
    if (!s || !count) return RT_E_INVALID;
    if (!s->ctx) return RT_E_STATE;
    const uint32_t cap = *count;
    *count = NPARAM;
    if (!params) return RT_OK;
    if (cap < NPARAM) return nrc_fail(s, RT_E_INVALID, "rt_nrc_params: capacity below the parameter count");
    rt_ctx* c = s->ctx;
    (void)hipSetDevice(c->device);
    if (set) {
        RT_HIP(c, hipMemcpyAsync(s->wf.ptr, params, NPARAM * 4, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_refresh_bf16, dim3((NPARAM + 255) / 256), dim3(256), 0, c->stream, (const float*)s->wf.ptr,
                           (__bf16*)s->wb.ptr);
        RT_HIP(c, hipGetLastError());
        RT_HIP(c, hipMemsetAsync(s->m.ptr, 0, NPARAM * 4, c->stream));
        RT_HIP(c, hipMemsetAsync(s->v.ptr, 0, NPARAM * 4, c->stream));
        s->steps = 0;
    } else {
        RT_HIP(c, hipMemcpyAsync(params, s->wf.ptr, NPARAM * 4, hipMemcpyDeviceToHost, c->stream));
    }
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_nrc_infer(rt_nrc* s, const float* feats, uint32_t n, float* out, int ptr_kind) {
    if (!s) return RT_E_INVALID;
    if (!s->ctx) return RT_E_STATE;
    if (n > 0 && (!feats || !out)) return nrc_fail(s, RT_E_INVALID, "rt_nrc_infer: NULL argument");
    if (ptr_kind != RT_PTR_HOST && ptr_kind != RT_PTR_DEVICE) return nrc_fail(s, RT_E_INVALID, "rt_nrc_infer: bad ptr_kind");
    if (n == 0) return RT_OK;
    rt_ctx* c = s->ctx;
    (void)hipSetDevice(c->device);
    const void* df;
    int rc;
    if ((rc = stage(s, s->feats, feats, (size_t)n * 36, ptr_kind, &df))) return rc;
    float* dout = out;
    if (ptr_kind == RT_PTR_HOST) {
        if ((rc = ensure(c, s->out, (size_t)n * 12))) return rc;
        dout = (float*)s->out.ptr;
    }
    if ((rc = nrc_infer_device(s, (const float*)df, n, dout))) return rc;
    c->last_op = 0;  // untimed, as training
    c->timing = rt_timing{};
    if (ptr_kind == RT_PTR_HOST) {
        RT_HIP(c, hipMemcpyAsync(out, dout, (size_t)n * 12, hipMemcpyDeviceToHost, c->stream));
        RT_HIP(c, hipStreamSynchronize(c->stream));
    }
    return RT_OK;
}

int rt_nrc_gradient(rt_nrc* s, const float* feats, const float* targets, uint32_t n, float* grad, float* loss,
                    int ptr_kind) {
    if (!s) return RT_E_INVALID;
    if (!s->ctx) return RT_E_STATE;
    if (!grad) return nrc_fail(s, RT_E_INVALID, "rt_nrc_gradient: NULL gradient output");
    (void)hipSetDevice(s->ctx->device);
    return run_train(s, feats, targets, n, ptr_kind, loss, grad, false);
}

int rt_nrc_train(rt_nrc* s, const float* feats, const float* targets, uint32_t n, float* loss, int ptr_kind) {
    if (!s) return RT_E_INVALID;
    if (!s->ctx) return RT_E_STATE;
    (void)hipSetDevice(s->ctx->device);
    return run_train(s, feats, targets, n, ptr_kind, loss, nullptr, true);
}

}  // extern "C"
