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
//   k_mlp_train  4 waves, tiles of 64 samples, at most one workgroup per CU
//                striding over the tiles: encode and forward in registers as
//                in inference, activations also stored transposed in LDS;
//                loss; backward with the gradients in both orientations;
//                bias gradients come out of the weight-gradient products
//                through a row of ones under the activations; each
//                workgroup sums its own partial gradient over its tiles in
//                registers and writes it once (no atomics, no
//                read-modify-write of global memory)
//   k_adam       sums the partials in a fixed order (eight interleaved runs,
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
constexpr int TILE = 64;                               // samples per workgroup tile
constexpr int NT = 256;                                // threads per workgroup

// fp32 parameter layout: [W0 b0 W1 b1 ... W5 b5], W_l row-major [n][k]
constexpr int offw(int l) { int o = 0; for (int i = 0; i < l; ++i) o += ND[i] * KD[i] + ND[i]; return o; }
constexpr int offb(int l) { return offw(l) + ND[l] * KD[l]; }
constexpr int NPARAM = offw(L);
// bf16 weight layout: [W0 W1 ... W5]
constexpr int offwb(int l) { int o = 0; for (int i = 0; i < l; ++i) o += ND[i] * KD[i]; return o; }
constexpr int NWB = offwb(L);
// training: 16 x 16 weight-gradient tiles a wave accumulates per layer and
// their offsets: K / 16 + 1 for the hidden layers (the last tile's first
// column is the bias gradient: the activations carry a row of ones there),
// 6 for the input layer (the encoding's padding column 78 is the ones row),
// and for the 16-row output layer one weight tile plus its bias tile
constexpr int GW_OFF[L] = {0, 6, 11, 16, 21, 26};
constexpr int GW_TILES = 28;
constexpr int BIAS_COL0 = 78;  // input layer: padding column that carries the ones row
static_assert(NPARAM == 23888, "parameter count");
static_assert(NWB == 23552, "bf16 weight count");


struct Geo {
    float lo[3], inv[3];
};

__device__ __forceinline__ __bf16 tobf(float x) { return (__bf16)x; }
// max(x, 0) as one integer max on the bits: a float with the sign bit set is
// a negative integer (-0 and negative NaNs give +0)
__device__ __forceinline__ float relu(float x) { return __int_as_float(max(__float_as_int(x), 0)); }

// Features of k block ks that lane group kg supplies for one sample (the
// input layer's weight columns are stored in this order).  kg < 3: sin and
// cos of pi x 2^(4 ks + j), j = 0..3, interleaved, for the sample's
// coordinate x on axis kg normalised to the scene box: an anchor from the
// hardware sine and cosine (v_sin_f32 / v_cos_f32, argument in revolutions)
// of the fractional part of x 2^(4 ks - 1) (the scaling and the subtraction
// of the floor are exact), then three angle doublings; the error stays far
// under the BF16 step the features are rounded to (the inference and
// gradient tests compare against the oracle's exact encoding).  kg = 3: the
// normal and the direction (nd, read only when ok) in block 0, zero padding
// after.
__device__ __forceinline__ bf16x8 feature_block(int kg, int ks, float x, const float* nd, bool ok) {
    bf16x8 v;
    if (kg < 3) {
        const float y = ldexpf(x, 4 * ks - 1);
        const float f = y - floorf(y);  // revolutions
        float sn = __builtin_amdgcn_sinf(f), cs = __builtin_amdgcn_cosf(f);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            v[2 * j] = tobf(sn);
            v[2 * j + 1] = tobf(cs);
            const float s2 = 2.0f * sn * cs, c2 = (cs - sn) * (cs + sn);
            sn = s2;
            cs = c2;
        }
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = tobf(ks == 0 && j < 6 && ok ? nd[j] : 0.0f);
    }
    return v;
}

// Training buffers in LDS: activations and gradients transposed,
// [feature][sample] with rows of TS, and gradients also [sample][feature]
// with rows of HS; rows padded by 8 BF16 (16 bytes) so that the 16 rows of an
// operand fragment start in different LDS banks.
constexpr int HS = 64 + 8, TS = 64 + 8;

// 16x16 tile of C = A * B over K (multiple of 32), both operands with k
// contiguous: A(m, k) = A[m * sAm + k], B(k, n) = B[n * sBn + k].  Lane l
// supplies A row m0 + l%16 and B column n0 + l%16, k = 8*(l/16) + j (one
// 16-byte load each); on return acc[i] = C(m0 + 4*(l/16) + i, n0 + l%16).
__device__ __forceinline__ f32x4 tile_mm(const __bf16* A, int sAm, const __bf16* B, int sBn, int m0, int n0, int K,
                                         int lane) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const int mi = lane & 15, kg = lane >> 4;
    for (int k0 = 0; k0 < K; k0 += 32) {
        const bf16x8 a = *(const bf16x8*)(A + (m0 + mi) * sAm + k0 + 8 * kg);
        const bf16x8 b = *(const bf16x8*)(B + (n0 + mi) * sBn + k0 + 8 * kg);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    return acc;
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

// blocked index of slot p (0..K-1) of row n in a matrix of N rows: K/32
// blocks of [N][32], so a lane's 8 slots of one row are one 16-byte load and
// the 16 rows of a fragment are 1 KB contiguous
__device__ __forceinline__ int blk(int n, int p, int N) { return (p >> 5) * N * 32 + n * 32 + (p & 31); }

// weights, row-major [n][k] per layer in global memory -> blocked and
// permuted in LDS (see above)
__device__ __forceinline__ void load_weights_blocked(const __bf16* __restrict__ Wb, __bf16* sW) {
    // input layer: slots (ks, 8 kg .. 8 kg + 7) <- features 24 kg + 8 ks ..
    for (int c = threadIdx.x; c < ND[0] * 12; c += INF_NT) {
        const int n = c / 12, ks = (c % 12) >> 2, kg = c & 3;
        *(uint4*)(sW + offwb(0) + blk(n, 32 * ks + 8 * kg, ND[0])) =
            *(const uint4*)(Wb + offwb(0) + n * KD[0] + 24 * kg + 8 * ks);
    }
    // hidden and output layers: slots 8 kg + j <- inputs 4 kg + j, 16 + 4 kg + j - 4
#pragma unroll
    for (int l = 1; l < L; ++l) {
        const int K = KD[l], N = ND[l];
        const int chunks = N * (K / 8);
        for (int c = threadIdx.x; c < chunks; c += INF_NT) {
            const int n = c / (K / 8), ks = (c % (K / 8)) >> 2, kg = c & 3;
            const __bf16* src = Wb + offwb(l) + n * K + 32 * ks + 4 * kg;
            uint2* dst = (uint2*)(sW + offwb(l) + blk(n, 32 * ks + 8 * kg, N));
            dst[0] = *(const uint2*)src;
            dst[1] = *(const uint2*)(src + 16);
        }
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

// bias of outputs 16 nt + 4 kg + i (i < 4) of layer l
template <int NT_>
__device__ __forceinline__ void load_bias(const float* __restrict__ Wf, int l, int kg, f32x4 (&b)[NT_]) {
#pragma unroll
    for (int nt = 0; nt < NT_; ++nt) {
        const float4 v = *(const float4*)(Wf + offb(l) + 16 * nt + 4 * kg);
        b[nt] = f32x4{v.x, v.y, v.z, v.w};
    }
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
    load_weights_blocked(Wb, sW);
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
        load_bias<4>(Wf, 0, kg, bias);
        // input layer, one k block per encoding anchor
        float xa[INF_MT];  // group kg < 3: the normalised axis coordinate of each sample
#pragma unroll
        for (int mt = 0; mt < INF_MT; ++mt) {
            const uint32_t sm = s0 + 16 * mt + mi;
            xa[mt] = (kg < 3 && sm < n) ? (feats[(size_t)sm * 9 + kg] - G.lo[kg]) * G.inv[kg] : 0.0f;
        }
#pragma unroll
        for (int ks = 0; ks < 3; ++ks) {
            bf16x8 x[INF_MT];
#pragma unroll
            for (int mt = 0; mt < INF_MT; ++mt) {
                const uint32_t sm = s0 + 16 * mt + mi;
                x[mt] = feature_block(kg, ks, xa[mt], feats + (size_t)sm * 9 + 3, sm < n);
            }
            if (ks == 0) layer_block<64, 4, true>(sW + offwb(0), ks, x, lane, bias, acc);
            else layer_block<64, 4, false>(sW + offwb(0), ks, x, lane, bias, acc);
        }
        // hidden layers 1..4 and the output layer, activations in registers
        f32x4 o5[INF_MT][1];
#pragma unroll
        for (int l = 1; l < L; ++l) {
            bf16x8 x[2][INF_MT];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                for (int mt = 0; mt < INF_MT; ++mt) {
                    const int ks = nt >> 1, h = 4 * (nt & 1);
#pragma unroll
                    for (int i = 0; i < 4; ++i) x[ks][mt][h + i] = tobf(relu(acc[mt][nt][i]));
                }
            if (l < L - 1) {
                load_bias<4>(Wf, l, kg, bias);
                layer_block<64, 4, true>(sW + offwb(l), 0, x[0], lane, bias, acc);
                layer_block<64, 4, false>(sW + offwb(l), 1, x[1], lane, bias, acc);
            } else {
                f32x4 b5[1];
                load_bias<1>(Wf, l, kg, b5);
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

// Slot of input k in layer l's blocked, permuted LDS rows (the inference
// layout: see load_weights_blocked), for layers l >= 1.
__device__ __forceinline__ int wslot(int k) {
    const int q = k & 31;
    return (k & ~31) + 8 * ((q & 15) >> 2) + 4 * (q >> 4) + (q & 3);
}

// One tile of 64 samples, 4 waves (wave w: samples 16 w .. 16 w + 15).
// Forward as in inference, activations in registers, biases as the first
// addend; every activation (and the encoding) is also written to LDS
// transposed, [feature][sample], for the weight gradients, whose products
// sum over samples.  The output-layer and hidden-layer gradients are kept in
// both orientations: [sample][feature] for the activation gradients (their
// row operand), [feature][sample] for the weight gradients.
__global__ void __launch_bounds__(NT) k_mlp_train(const float* __restrict__ feats, const float* __restrict__ targets,
                                                  uint32_t n, const __bf16* __restrict__ Wb,
                                                  const float* __restrict__ Wf, Geo G, float* __restrict__ partial,
                                                  float* __restrict__ loss_part) {
    __shared__ __attribute__((aligned(16))) __bf16 sW[NWB];             // inference layout
    __shared__ __attribute__((aligned(16))) __bf16 sXT[96 * TS];        // encoding, [feature][sample]
    __shared__ __attribute__((aligned(16))) __bf16 sHT[L - 1][80 * TS];  // ReLU activations, [feature][sample], then a row of ones
    __shared__ __attribute__((aligned(16))) __bf16 sD[2][TILE * HS];     // gradients, [sample][feature]
    __shared__ __attribute__((aligned(16))) __bf16 sDT[2][64 * TS];      // gradients, [feature][sample]
    __shared__ float sLoss[NT / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int mi = lane & 15, kg = lane >> 4;
    float* part = partial + (size_t)blockIdx.x * NPARAM;
    load_weights_blocked(Wb, sW);
    float lacc = 0.0f;
    // this thread's share of the workgroup's gradient, summed over its tiles
    // in tile order in registers and written once at the end: the weight
    // tiles of every layer (wave w owns rows 16 w .. 16 w + 15 of each dW,
    // GW_OFF[l] + kt indexes layer l's column tile kt) and, for threads < 64,
    // one bias element per layer
    f32x4 gw[GW_TILES];
#pragma unroll
    for (int i = 0; i < GW_TILES; ++i) gw[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // activation rows 64..79: ones, then zeros, so that each layer's last
    // weight-gradient tile sums the output gradients over samples (the bias
    // gradient) in its first column
    for (int e = threadIdx.x; e < (L - 1) * 16 * TS; e += NT) {
        const int l = e / (16 * TS), r = (e / TS) % 16;
        sHT[l][(64 + r) * TS + e % TS] = tobf(r == 0 ? 1.0f : 0.0f);
    }
    const uint32_t tiles = (n + TILE - 1) / TILE;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        // keeps the weight fragments' LDS loads inside the loop
        __asm__ __volatile__("" ::: "memory");
        const uint32_t s0 = tile * TILE;
        const int sw = 16 * w + mi;             // this lane's sample in the tile
        const bool valid = s0 + sw < n;
        __syncthreads();  // the previous tile's backward pass is done with the buffers
        // ---- forward
        f32x4 acc[1][4];
        f32x4 bias[4];
        load_bias<4>(Wf, 0, kg, bias);
        const float xa = (kg < 3 && valid) ? (feats[(size_t)(s0 + sw) * 9 + kg] - G.lo[kg]) * G.inv[kg] : 0.0f;
#pragma unroll
        for (int ks = 0; ks < 3; ++ks) {
            bf16x8 x[1];
            x[0] = feature_block(kg, ks, xa, feats + (size_t)(s0 + sw) * 9 + 3, valid);
            // slot (ks, 8 kg + j) holds feature 24 kg + 8 ks + j; the padding
            // feature BIAS_COL0 is stored as 1 for the weight gradients
            // (the forward product sees it as 0)
#pragma unroll
            for (int j = 0; j < 8; ++j)
                sXT[(24 * kg + 8 * ks + j) * TS + sw] = (24 * kg + 8 * ks + j == BIAS_COL0) ? tobf(1.0f) : x[0][j];
            if (ks == 0) layer_block<64, 4, true, 1>(sW + offwb(0), ks, x, lane, bias, acc);
            else layer_block<64, 4, false, 1>(sW + offwb(0), ks, x, lane, bias, acc);
        }
        f32x4 o5[1][1];
#pragma unroll
        for (int l = 1; l < L; ++l) {
            bf16x8 x[2][1];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const int ks = nt >> 1, h = 4 * (nt & 1);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const __bf16 v = tobf(relu(acc[0][nt][i]));
                    x[ks][0][h + i] = v;
                    sHT[l - 1][(16 * nt + 4 * kg + i) * TS + sw] = v;
                }
            }
            if (l < L - 1) {
                load_bias<4>(Wf, l, kg, bias);
                layer_block<64, 4, true, 1>(sW + offwb(l), 0, x[0], lane, bias, acc);
                layer_block<64, 4, false, 1>(sW + offwb(l), 1, x[1], lane, bias, acc);
            } else {
                f32x4 b5[1];
                load_bias<1>(Wf, l, kg, b5);
                layer_block<16, 1, true, 1>(sW + offwb(l), 0, x[0], lane, b5, o5);
                layer_block<16, 1, false, 1>(sW + offwb(l), 1, x[1], lane, b5, o5);
            }
        }
        // ---- loss: lane (mi, kg) holds outputs 4 kg + i of its sample;
        // dL/dy = 2 (y - t) / (y^2 + 0.01) / n on the 3 valid outputs
        float lsum = 0.0f;
        {
            bf16x4 dyt;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int c = 4 * kg + i;
                float dy = 0.0f;
                if (c < N5_VALID && valid) {
                    const float y = o5[0][0][i];
                    const float t = targets[(size_t)(s0 + sw) * 3 + c];
                    const float den = y * y + 0.01f;
                    const float r = y - t;
                    lsum += r * r / den;
                    dy = 2.0f * r / den / (float)n;
                }
                dyt[i] = tobf(dy);
                sDT[0][c * TS + sw] = dyt[i];
            }
            *(bf16x4*)(sD[0] + sw * HS + 4 * kg) = dyt;
        }
        for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
        if (lane == 0) sLoss[w] = lsum;
        __syncthreads();
        if (threadIdx.x == 0) lacc += sLoss[0] + sLoss[1] + sLoss[2] + sLoss[3];

        // ---- backward
        // layer 5: dW5[n][k] = sum_s dY[s][n] H4[s][k] (16 x 64), one 16-column tile per wave
        {
            const f32x4 g = tile_mm(sDT[0], TS, sHT[L - 2], TS, 0, 16 * w, 64, lane);
#pragma unroll
            for (int i = 0; i < 4; ++i) gw[GW_OFF[L - 1]][i] += g[i];
            if (w == 0) {  // columns 64..79: the ones row, so column 64 is the bias gradient
                const f32x4 gbias = tile_mm(sDT[0], TS, sHT[L - 2], TS, 0, 64, 64, lane);
#pragma unroll
                for (int i = 0; i < 4; ++i) gw[GW_OFF[L - 1] + 1][i] += gbias[i];
            }
        }
        // dH_{l-1}[s][k] = sum_n dZ[s][n] W_l[n][k] (n < nvalid), masked by H_{l-1} > 0, into both layouts
        auto act_grad = [&](int l, int nvalid, const __bf16* dZ, __bf16* dst, __bf16* dstT) {
            const int N = ND[l];
            const __bf16* Wl = sW + offwb(l);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const int col = 16 * nt + mi;
                const int slot = wslot(col);
                const __bf16* wcol = Wl + (slot >> 5) * N * 32 + (slot & 31);  // W_l[n][col] at wcol[n * 32]
                f32x4 g = {0.0f, 0.0f, 0.0f, 0.0f};
                for (int k0 = 0; k0 < ((nvalid + 31) & ~31); k0 += 32) {
                    const int kb = k0 + 8 * kg;
                    bf16x8 av, bv;
                    if (kb < nvalid) av = *(const bf16x8*)(dZ + (16 * w + mi) * HS + kb);
                    else av = bf16x8{};
#pragma unroll
                    for (int j = 0; j < 8; ++j) bv[j] = kb + j < nvalid ? wcol[(kb + j) * 32] : tobf(0.0f);
                    g = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, g, 0, 0, 0);
                }
                // g[i] = dH(sample 16 w + 4 kg + i, feature col)
                const bf16x4 h = *(const bf16x4*)(sHT[l - 1] + col * TS + 16 * w + 4 * kg);
                bf16x4 gt;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    gt[i] = tobf((float)h[i] > 0.0f ? g[i] : 0.0f);
                    dst[(16 * w + 4 * kg + i) * HS + col] = gt[i];
                }
                *(bf16x4*)(dstT + col * TS + 16 * w + 4 * kg) = gt;
            }
        };
        act_grad(L - 1, N5_VALID, sD[0], sD[1], sDT[1]);
        __syncthreads();
        int cur = 1;
#pragma unroll
        for (int l = L - 2; l >= 0; --l) {
            const __bf16* inT = (l == 0) ? sXT : sHT[l - 1];
            const int K = KD[l];
            // dW_l[n][k] = sum_s dZ[s][n] in[s][k]: rows n in 16-row tiles (one per wave), K / 16
            // column tiles, plus the ones-row tile (bias) for the hidden layers
            const int KT = (l == 0) ? K / 16 : K / 16 + 1;
#pragma unroll
            for (int kt = 0; kt < KT; ++kt) {
                const f32x4 g = tile_mm(sDT[cur], TS, inT, TS, 16 * w, 16 * kt, 64, lane);
#pragma unroll
                for (int i = 0; i < 4; ++i) gw[GW_OFF[l] + kt][i] += g[i];
            }
            if (l > 0) {
                act_grad(l, 64, sD[cur], sD[cur ^ 1], sDT[cur ^ 1]);
                __syncthreads();
                cur ^= 1;
            }
        }
    }  // tiles
    // the workgroup's partial gradient, written once; the ones-row columns
    // go to the biases
#pragma unroll
    for (int l = 0; l < L; ++l) {
        const int K = KD[l];
        const int nkt = (l == L - 1) ? 1 : K / 16;
#pragma unroll
        for (int kt = 0; kt < nkt; ++kt) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                // layer 5 (16 x 64): wave w holds columns 16 w .. 16 w + 15 of all 16 rows
                const int row = (l == L - 1) ? 4 * kg + i : 16 * w + 4 * kg + i;
                const int col = (l == L - 1) ? 16 * w + mi : 16 * kt + mi;
                const float g = gw[GW_OFF[l] + kt][i];
                if (l == 0 && col == BIAS_COL0) {
                    part[offb(0) + row] = g;
                    part[offw(0) + row * K + col] = 0.0f;
                } else {
                    part[offw(l) + row * K + col] = g;
                }
            }
        }
        if (l > 0 && mi == 0 && (l < L - 1 || w == 0)) {
            const int bt = (l == L - 1) ? GW_OFF[l] + 1 : GW_OFF[l] + K / 16;
#pragma unroll
            for (int i = 0; i < 4; ++i) part[offb(l) + ((l == L - 1) ? 4 * kg + i : 16 * w + 4 * kg + i)] = gw[bt][i];
        }
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = lacc;
}

// Adam with bias correction on the summed partial gradients (workgroup order,
// so the sum is deterministic); refreshes the bf16 weight copy.  With lr = 0
// it only writes the summed gradient (gradient queries).
// parameter p's BF16 copy (weights only; the biases stay FP32)
__device__ __forceinline__ void store_bf16(__bf16* Wb, int p, float w) {
    for (int l = 0; l < L; ++l) {
        const int o = offw(l), nw = ND[l] * KD[l];
        if (p >= o && p < o + nw) Wb[offwb(l) + (p - o)] = tobf(w);
    }
}

// A block of ADAM_NT threads serves 64 parameters: wave q sums partials
// q, q + QN, q + 2 QN, ... in order (each load one coalesced 256-byte row,
// eight in flight), then wave 0 adds the QN sums in wave order.
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
        uint32_t w = q;
        for (; w + 7 * QN < nparts; w += 8 * QN) {
            float x[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) x[j] = partial[(size_t)(w + QN * j) * NPARAM + p];
#pragma unroll
            for (int j = 0; j < 8; ++j) g += x[j];
        }
        for (; w < nparts; w += QN) g += partial[(size_t)w * NPARAM + p];
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
    const uint32_t tiles = (n + TILE - 1) / TILE;
    // one workgroup per CU at most (the kernel holds 116 KB of LDS); each
    // strides over tiles, so k_adam sums at most num_cus partial gradients
    const uint32_t wgs = tiles < (uint32_t)c->num_cus ? tiles : (uint32_t)c->num_cus;
    if ((rc = ensure(c, s->partial, (size_t)wgs * NPARAM * 4)) || (rc = ensure(c, s->loss, (size_t)wgs * 4))) return rc;
    RT_HIP(c, hipEventRecord(c->ev[2], c->stream));
    hipLaunchKernelGGL(k_mlp_train, dim3(wgs), dim3(NT), 0, c->stream, (const float*)df, (const float*)dt, n,
                       (const __bf16*)s->wb.ptr, (const float*)s->wf.ptr, s->geo, (float*)s->partial.ptr,
                       (float*)s->loss.ptr);
    RT_HIP(c, hipGetLastError());
    RT_HIP(c, hipEventRecord(c->ev[3], c->stream));
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
    c->last_op = 3;
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
    if ((rc = ensure(c, s->wf, NPARAM * 4)) || (rc = ensure(c, s->wb, NWB * 2)) || (rc = ensure(c, s->m, NPARAM * 4)) ||
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
    RT_HIP(c, hipEventRecord(c->ev[2], c->stream));
    if ((rc = nrc_infer_device(s, (const float*)df, n, dout))) return rc;
    RT_HIP(c, hipEventRecord(c->ev[3], c->stream));
    c->last_op = 3;
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
