// learned_guess.hip -- a per-site network that predicts the induced dipoles
// before the CG solve, trained on the GPU on dipoles the solve converged.
//
// The reference has no counterpart: its ConjugateGradient starts from Pi = 1
// (src/forces/Induction.cpp:2404-2410).  Here CG starts from
//     Pi_i = alpha_i E0_i + F_i^T (sd * y_i + mean),
// y_i the network's output for site i's descriptor and F_i the site's frame,
// and then refines to the same tolerance, so the solve still decides the
// result; a poor guess costs iterations, not accuracy.
//
// Site frame: z along the H-O-H bisector, x along H1->H2 made perpendicular
// to z, y = z x x; H1 and O use it, H2 its mirror image (-x, y, z).
// Descriptor, 32 values in the site's frame:
//   0-2   alpha E0                      3-5  alpha (T alpha E0)
//   6-29  alpha sum q_j r / r^3 over intermolecular partners in the shells
//         r < 0.3, 0.4, 0.5, 0.6 nm, split by q_j > 0 and q_j <= 0
//   30    alpha                         31   1 for a positive site, else 0
// Network: 32 -> 128 -> 128 -> 128 -> 3, SiLU after the hidden layers, on
// standardised inputs and outputs.
//
// Matrix products run on the matrix cores as
// v_mfma_f32_16x16x16_bf16: BF16 operands, FP32 accumulation.  Fragment
// layout of lane l (r = l % 16, g = l / 16):
//   A[r][4g .. 4g+3],  B[4g .. 4g+3][r],  D[4g + t][r] for t = 0..3.
// One wave carries 16 rows through every layer; activations pass between
// layers through LDS, because D's layout is A's transposed.
//  * The training step is one kernel per batch: forward, loss, backward to
//    the first layer.  It keeps the pre-activations in registers and writes
//    the BF16 layer inputs and back-propagated gradients feature-major.
//  * The weight gradient then runs one wave per 16 x 16 tile of dW and batch
//    slice, each tile an MFMA product over the batch.
//  * Adam sums the slices in a fixed order, so a training run is
//    deterministic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "smd_internal.h"

namespace smd {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int MF = 32, MH = 128, MO = 3, MOP = 16;   // features, hidden width, outputs (padded)
// FP32 master parameters: the four weight matrices [out][in] (the last one
// padded to 16 rows that stay zero), then the four bias vectors.
constexpr int OFF_W1 = 0, OFF_W2 = 4096, OFF_W3 = 20480, OFF_W4 = 36864, W_INT = 38912;
constexpr int OFF_B1 = 38912, OFF_B2 = 39040, OFF_B3 = 39168, OFF_B4 = 39296, P_INT = 39312;
// the transposed BF16 copies [in][out] sit at the same offsets
constexpr int NPUB = 37635;        // W1 b1 W2 b2 W3 b3 W4[3] b4[3]
constexpr int NNORM = 2 * MF + 2 * MO;
constexpr int MLP_WAVES = 4;       // waves (of 16 rows) per block
constexpr int LDS_STRIDE = 136;    // BF16 per LDS row: 128 + 8, rows stay 16-byte aligned
constexpr int WG_TILES = 16 + 64 + 64 + 8;
constexpr int WG_SPLITS = 16;      // batch slices of the weight gradient
constexpr int DESC_LANES = 16;
constexpr double SQRT_COULOMB_LG = 11.78708976;   // SQRTCOULOMBCONSTANT, src/NEMD_defs.h:103
constexpr double POLMIN_LG = 0.00000001;          // Induction.cpp:195

__device__ __forceinline__ uint16_t f2bf(float f) {   // nearest, ties to even
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(short h) { return __uint_as_float((uint32_t)(uint16_t)h << 16); }

__device__ __forceinline__ v4f mfma16(v4s a, v4s b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

struct MlpArgs {
    int rows, rows16;          // valid rows, rows padded to 16
    const int32_t* idx;        // row -> sample of x and tgt (null: identity)
    const uint16_t* x;         // [*][32] standardised inputs, BF16
    const uint16_t* wb;        // BF16 weights [out][in]
    const uint16_t* wt;        // BF16 weights [in][out]
    const float* w;            // FP32 masters (biases)
    float* y;                  // [rows][3]
    // training
    const float* tgt;          // [*][3] standardised targets
    uint16_t* xt;              // [32][rows16] layer-1 inputs, feature-major
    uint16_t* ht;              // [3][128][rows16] inputs of layers 2-4
    uint16_t* dzt;             // [3][128][rows16] dL/dz of layers 1-3, then [16][rows16] of layer 4
    double* loss_part;         // [blocks] sums of squared errors
    float scale;               // 2 / (3 rows)
};

__device__ __forceinline__ float silu_f(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu_f(float z) {
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// acc[c] = bias of the 16 columns of tile c, for every row
__device__ __forceinline__ void bias_init(v4f (&acc)[8], const float* b, int r) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const float v = b[c * 16 + r];
        acc[c] = v4f{v, v, v, v};
    }
}

// acc += A (16 x K, from LDS rows) . W^T tiles, W [128][K] BF16 row-major
template <int K>
__device__ __forceinline__ void dense_lds(v4f (&acc)[8], const uint16_t* lds, const uint16_t* W, int r, int g) {
#pragma unroll
    for (int ks = 0; ks < K / 16; ++ks) {
        const v4s av = *reinterpret_cast<const v4s*>(lds + r * LDS_STRIDE + ks * 16 + 4 * g);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const v4s bv = *reinterpret_cast<const v4s*>(W + (size_t)(c * 16 + r) * K + ks * 16 + 4 * g);
            acc[c] = mfma16(av, bv, acc[c]);
        }
    }
}

// SiLU of a hidden layer's pre-activations into LDS (row 4g+t, column
// 16c+r); when training, the BF16 activations also go feature-major to ht.
template <bool TRAIN>
__device__ __forceinline__ void hidden_out(const MlpArgs& a, int l, const v4f (&acc)[8], uint16_t* lds, int row0,
                                           bool live, int r, int g) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int col = c * 16 + r;
        v4s hb;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint16_t h = f2bf(silu_f(acc[c][t]));
            hb[t] = (short)h;
            lds[(4 * g + t) * LDS_STRIDE + col] = h;
        }
        if (TRAIN && live)
            *reinterpret_cast<v4s*>(a.ht + (size_t)l * MH * a.rows16 + (size_t)col * a.rows16 + row0 + 4 * g) = hb;
    }
}

// Backward through a hidden layer: dz = (dh) * silu'(z), dh[16 x 128] from
// the A fragments in LDS (K columns) times W^T tiles Wt [128][K] of the layer
// above; dz to LDS (when out_lds) and feature-major to dzt.
template <int K>
__device__ __forceinline__ void back_layer(const MlpArgs& a, int l, const uint16_t* lds_in, uint16_t* lds_out,
                                           const uint16_t* Wt, const v4f (&z)[8], int row0, bool live, int r,
                                           int g) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        v4f acc{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < K / 16; ++ks) {
            const v4s av = *reinterpret_cast<const v4s*>(lds_in + r * LDS_STRIDE + ks * 16 + 4 * g);
            const v4s bv = *reinterpret_cast<const v4s*>(Wt + (size_t)(c * 16 + r) * K + ks * 16 + 4 * g);
            acc = mfma16(av, bv, acc);
        }
        const int col = c * 16 + r;
        v4s db;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint16_t d = f2bf(acc[t] * dsilu_f(z[c][t]));
            db[t] = (short)d;
            if (lds_out) lds_out[(4 * g + t) * LDS_STRIDE + col] = d;
        }
        if (live)
            *reinterpret_cast<v4s*>(a.dzt + (size_t)l * MH * a.rows16 + (size_t)col * a.rows16 + row0 + 4 * g) = db;
    }
}

// Forward of 16 rows per wave; with TRAIN also the loss and the backward pass.
// Every wave runs every step (rows past rows16 compute on zeros and store
// nothing) so that the block barriers stay uniform.
template <bool TRAIN>
__global__ __launch_bounds__(256) void k_mlp_rows(MlpArgs a) {
    __shared__ __align__(16) uint16_t act[MLP_WAVES][2][16 * LDS_STRIDE];
    __shared__ double lsum[MLP_WAVES];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int row0 = (blockIdx.x * MLP_WAVES + wv) * 16;
    const bool live = row0 < a.rows16;
    uint16_t* L0 = act[wv][0];
    uint16_t* L1 = act[wv][1];

    v4s ax[2];
    {
        const int row = row0 + r;
        if (live && row < a.rows) {
            const int s = a.idx ? a.idx[row] : row;
            const uint16_t* xr = a.x + (size_t)s * MF;
            ax[0] = *reinterpret_cast<const v4s*>(xr + 4 * g);
            ax[1] = *reinterpret_cast<const v4s*>(xr + 16 + 4 * g);
        } else {
            ax[0] = v4s{0, 0, 0, 0};
            ax[1] = v4s{0, 0, 0, 0};
        }
        if (TRAIN && live)
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
#pragma unroll
                for (int t = 0; t < 4; ++t) a.xt[(size_t)(ks * 16 + 4 * g + t) * a.rows16 + row] = (uint16_t)ax[ks][t];
    }
    v4f z1[8], z2[8], z3[8];
    bias_init(z1, a.w + OFF_B1, r);
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const v4s bv = *reinterpret_cast<const v4s*>(a.wb + OFF_W1 + (size_t)(c * 16 + r) * MF + ks * 16 + 4 * g);
            z1[c] = mfma16(ax[ks], bv, z1[c]);
        }
    hidden_out<TRAIN>(a, 0, z1, L0, row0, live, r, g);
    __syncthreads();
    bias_init(z2, a.w + OFF_B2, r);
    dense_lds<MH>(z2, L0, a.wb + OFF_W2, r, g);
    hidden_out<TRAIN>(a, 1, z2, L1, row0, live, r, g);
    __syncthreads();
    bias_init(z3, a.w + OFF_B3, r);
    dense_lds<MH>(z3, L1, a.wb + OFF_W3, r, g);
    __syncthreads();                       // every lane has read L0 (layer 2's input)
    hidden_out<TRAIN>(a, 2, z3, L0, row0, live, r, g);
    __syncthreads();
    v4f o;
    {
        const float b = a.w[OFF_B4 + r];
        o = v4f{b, b, b, b};
#pragma unroll
        for (int ks = 0; ks < MH / 16; ++ks) {
            const v4s av = *reinterpret_cast<const v4s*>(L0 + r * LDS_STRIDE + ks * 16 + 4 * g);
            const v4s bv = *reinterpret_cast<const v4s*>(a.wb + OFF_W4 + (size_t)r * MH + ks * 16 + 4 * g);
            o = mfma16(av, bv, o);
        }
    }
    // o[t] = output r of row row0 + 4g + t
    if (live && r < MO)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int row = row0 + 4 * g + t;
            if (row < a.rows) a.y[(size_t)row * MO + r] = o[t];
        }
    if (!TRAIN) return;

    // loss and dL/dy; dL/dy of padded rows and outputs is zero
    double se = 0.0;
    v4s dyb;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int row = row0 + 4 * g + t;
        float d = 0.f;
        if (live && r < MO && row < a.rows) {
            const int s = a.idx ? a.idx[row] : row;
            d = o[t] - a.tgt[(size_t)s * MO + r];
            se += (double)d * (double)d;
        }
        const uint16_t h = f2bf(a.scale * d);
        dyb[t] = (short)h;
    }
    __syncthreads();                       // every lane has read L0 (layer 4's input)
#pragma unroll
    for (int t = 0; t < 4; ++t) L1[(4 * g + t) * LDS_STRIDE + r] = (uint16_t)dyb[t];
    if (live)
        *reinterpret_cast<v4s*>(a.dzt + (size_t)3 * MH * a.rows16 + (size_t)r * a.rows16 + row0 + 4 * g) = dyb;
    se = wave_sum(se);
    if (lane == 0) lsum[wv] = se;
    __syncthreads();
    if (threadIdx.x == 0) a.loss_part[blockIdx.x] = (lsum[0] + lsum[1]) + (lsum[2] + lsum[3]);
    // layer 3: dh3 = dy . W4 (K = 16), W4^T [128][16]
    back_layer<MOP>(a, 2, L1, L0, a.wt + OFF_W4, z3, row0, live, r, g);
    __syncthreads();
    back_layer<MH>(a, 1, L0, L1, a.wt + OFF_W3, z2, row0, live, r, g);
    __syncthreads();
    back_layer<MH>(a, 0, L1, nullptr, a.wt + OFF_W2, z1, row0, live, r, g);
}

// ---- inference: weights staged in LDS, 16x16x32 matrix products ----------

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

__device__ __forceinline__ v4f mfma32(v8bf a, v8bf b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// v_mfma_f32_16x16x32_bf16 fragments of lane l (r = l % 16, g = l / 16):
//   A[r][8g .. 8g+7],  B[8g .. 8g+7][r],  D[4g + t][r] for t = 0..3.
// Each block stages the four BF16 weight matrices in LDS once (rows padded
// by 8 so that the 16 rows a fragment read touches fall in different banks)
// and carries INF_WAVES x 16 rows through the network, 16 per wave.
constexpr int INF_WAVES = 8;
constexpr int W1_STRIDE = MF + 8;                       // BF16 per LDS row of W1
constexpr int INF_W1 = 0, INF_W2 = MH * W1_STRIDE, INF_W3 = INF_W2 + MH * LDS_STRIDE,
              INF_W4 = INF_W3 + MH * LDS_STRIDE, INF_WSIZE = INF_W4 + MOP * LDS_STRIDE;

__device__ __forceinline__ v8bf ld8(const uint16_t* p) { return *reinterpret_cast<const v8bf*>(p); }

// rows x cols BF16 (row-major, cols a multiple of 8) from global to LDS with
// the given LDS row stride, 16 bytes per thread and step
__device__ __forceinline__ void stage(uint16_t* dst, int stride, const uint16_t* src, int rows, int cols) {
    const int per_row = cols / 8;
    for (int t = threadIdx.x; t < rows * per_row; t += blockDim.x) {
        const int row = t / per_row, c8 = t % per_row;
        *reinterpret_cast<uint4*>(dst + row * stride + 8 * c8) =
            *reinterpret_cast<const uint4*>(src + (size_t)row * cols + 8 * c8);
    }
}

// acc += A (16 x 128, LDS rows) . W^T over 8 column tiles, W [128][128] in LDS
__device__ __forceinline__ void dense128(v4f (&acc)[8], const uint16_t* lds, const uint16_t* W, int r, int g) {
#pragma unroll
    for (int ks = 0; ks < MH / 32; ++ks) {
        const v8bf av = ld8(lds + r * LDS_STRIDE + ks * 32 + 8 * g);
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[c] = mfma32(av, ld8(W + (c * 16 + r) * LDS_STRIDE + ks * 32 + 8 * g), acc[c]);
    }
}

__device__ __forceinline__ void silu_lds(const v4f (&acc)[8], uint16_t* lds, int r, int g) {
#pragma unroll
    for (int c = 0; c < 8; ++c)
#pragma unroll
        for (int t = 0; t < 4; ++t) lds[(4 * g + t) * LDS_STRIDE + c * 16 + r] = f2bf(silu_f(acc[c][t]));
}

__global__ __launch_bounds__(INF_WAVES * WAVE) void k_mlp_infer(MlpArgs a) {
    __shared__ __align__(16) uint16_t wl[INF_WSIZE];
    __shared__ __align__(16) uint16_t act[INF_WAVES][2][16 * LDS_STRIDE];
    stage(wl + INF_W1, W1_STRIDE, a.wb + OFF_W1, MH, MF);
    stage(wl + INF_W2, LDS_STRIDE, a.wb + OFF_W2, MH, MH);
    stage(wl + INF_W3, LDS_STRIDE, a.wb + OFF_W3, MH, MH);
    stage(wl + INF_W4, LDS_STRIDE, a.wb + OFF_W4, MOP, MH);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int row0 = (blockIdx.x * INF_WAVES + wv) * 16;
    uint16_t* L0 = act[wv][0];
    uint16_t* L1 = act[wv][1];
    uint4 xin = make_uint4(0u, 0u, 0u, 0u);   // rows past the end compute on zeros
    {
        const int row = row0 + r;
        if (row < a.rows) {
            const int sidx = a.idx ? a.idx[row] : row;
            xin = *reinterpret_cast<const uint4*>(a.x + (size_t)sidx * MF + 8 * g);
        }
    }
    const v8bf ax = __builtin_bit_cast(v8bf, xin);
    __syncthreads();   // the weights are in LDS
    // every wave runs every layer, so the barriers between them stay uniform
    v4f z[8];
    bias_init(z, a.w + OFF_B1, r);
#pragma unroll
    for (int c = 0; c < 8; ++c) z[c] = mfma32(ax, ld8(wl + INF_W1 + (c * 16 + r) * W1_STRIDE + 8 * g), z[c]);
    silu_lds(z, L0, r, g);
    __syncthreads();
    bias_init(z, a.w + OFF_B2, r);
    dense128(z, L0, wl + INF_W2, r, g);
    silu_lds(z, L1, r, g);
    __syncthreads();
    bias_init(z, a.w + OFF_B3, r);
    dense128(z, L1, wl + INF_W3, r, g);
    __syncthreads();   // every lane has read L0 (layer 2's input)
    silu_lds(z, L0, r, g);
    __syncthreads();
    v4f o;
    {
        const float b = a.w[OFF_B4 + r];
        o = v4f{b, b, b, b};
#pragma unroll
        for (int ks = 0; ks < MH / 32; ++ks)
            o = mfma32(ld8(L0 + r * LDS_STRIDE + ks * 32 + 8 * g), ld8(wl + INF_W4 + r * LDS_STRIDE + ks * 32 + 8 * g), o);
    }
    if (r < MO)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int row = row0 + 4 * g + t;
            if (row < a.rows) a.y[(size_t)row * MO + r] = o[t];
        }
}

// dW[m][k] = sum_n dz[m][n] in[k][n] over one batch slice, one wave per tile;
// tiles with k0 = 0 also sum dz for the bias.  Writes every entry of its
// slice's partial vector.
__global__ __launch_bounds__(64) void k_mlp_wgrad(MlpArgs a, float* part, int chunk) {
    int t = blockIdx.x, l;
    if (t < 16) l = 0;
    else if (t < 80) { l = 1; t -= 16; }
    else if (t < 144) { l = 2; t -= 80; }
    else { l = 3; t -= 144; }
    const int K = l == 0 ? MF : MH;
    const int kt_n = K / 16, mt = t / kt_n, kt = t % kt_n;
    const int woff = l == 0 ? OFF_W1 : l == 1 ? OFF_W2 : l == 2 ? OFF_W3 : OFF_W4;
    const int boff = l == 0 ? OFF_B1 : l == 1 ? OFF_B2 : l == 2 ? OFF_B3 : OFF_B4;
    const uint16_t* dz = a.dzt + (size_t)l * MH * a.rows16;
    const uint16_t* in = l == 0 ? a.xt : a.ht + (size_t)(l - 1) * MH * a.rows16;
    const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
    const uint16_t* arow = dz + (size_t)(mt * 16 + r) * a.rows16 + 4 * g;
    const uint16_t* brow = in + (size_t)(kt * 16 + r) * a.rows16 + 4 * g;
    const int n0 = blockIdx.y * chunk, n1 = min(a.rows16, n0 + chunk);
    v4f acc{0.f, 0.f, 0.f, 0.f};
    float bs = 0.f;
    for (int n = n0; n < n1; n += 16) {
        const v4s av = *reinterpret_cast<const v4s*>(arow + n);
        const v4s bv = *reinterpret_cast<const v4s*>(brow + n);
        acc = mfma16(av, bv, acc);
        bs += (bf2f(av[0]) + bf2f(av[1])) + (bf2f(av[2]) + bf2f(av[3]));
    }
    float* p = part + (size_t)blockIdx.y * P_INT;
#pragma unroll
    for (int q = 0; q < 4; ++q) p[woff + (size_t)(mt * 16 + 4 * g + q) * K + kt * 16 + r] = acc[q];
    if (kt == 0) {
        bs += __shfl_xor(bs, 16);
        bs += __shfl_xor(bs, 32);
        if (g == 0) p[boff + mt * 16 + r] = bs;
    }
}

// Index of master parameter p < W_INT in the transposed copy.
__device__ __forceinline__ int transposed_index(int p) {
    if (p < OFF_W2) return OFF_W1 + (p % MF) * MH + p / MF;
    if (p < OFF_W3) { const int q = p - OFF_W2; return OFF_W2 + (q % MH) * MH + q / MH; }
    if (p < OFF_W4) { const int q = p - OFF_W3; return OFF_W3 + (q % MH) * MH + q / MH; }
    const int q = p - OFF_W4;
    return OFF_W4 + (q % MH) * MOP + q / MH;
}

struct AdamArgs {
    float* w; float* m; float* v;
    uint16_t* wb; uint16_t* wt;
    const float* part; int splits;
    float* grad_out;           // when set: store the summed gradient, update nothing
    float lr, b1, b2, eps, c1, c2;   // c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)
    int update;                // 0: only refresh the BF16 copies from w
};

__global__ __launch_bounds__(256) void k_mlp_adam(AdamArgs a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P_INT) return;
    float w = a.w[p];
    if (a.part) {
        float gsum = 0.f;
        for (int s = 0; s < a.splits; ++s) gsum += a.part[(size_t)s * P_INT + p];
        if (a.grad_out) { a.grad_out[p] = gsum; return; }
        if (a.update) {
            const float m = a.b1 * a.m[p] + (1.f - a.b1) * gsum;
            const float v = a.b2 * a.v[p] + (1.f - a.b2) * gsum * gsum;
            a.m[p] = m;
            a.v[p] = v;
            w -= a.lr * (m * a.c1) / (sqrtf(v * a.c2) + a.eps);
            a.w[p] = w;
        }
    }
    if (p < W_INT) {
        const uint16_t h = f2bf(w);
        a.wb[p] = h;
        a.wt[transposed_index(p)] = h;
    }
}

// ---- descriptors -----------------------------------------------------------------

struct DescArgs {
    int n, nmol, stride;
    Box box;
    const int32_t* mol_start;
    const int32_t* mol;
    const double* pos;
    const double* q;
    const double* polar;
    const double* e0;
    int tstride;               // the solve's coupling table: partners inside cutoffInd in list order
    const int32_t* tcnt;
    const int32_t* tj;
    // mixed solves: the solve's hi/lo FP32 site records (x_hi, y_hi, z_hi, 1/sqrt(alpha)),
    // (x_lo, y_lo, z_lo, q / sqrt(Coulomb constant)); null for FP64 descriptors
    const float4* rec32;
    float Lh[3], Ll[3], invL[3];
    const float* shells;       // the charge shells from the solve's mixed field pass [n][24], else null
    double* frame;             // [n][9] rows x, y, z of each site's frame
    const double* x;           // alpha E0 [3n]
    const double* z;           // A (alpha E0) [3n]
    float* raw;                // [n][32] when set
    uint16_t* xn;              // [n][32] standardised BF16 when set
    const float* norm;         // mean[32], sd[32], out mean[3], out sd[3]
};

__global__ __launch_bounds__(256) void k_site_frames(DescArgs a) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.nmol) return;
    const int a0 = a.mol_start[m];
    const double* p = a.pos + 3 * a0;
    double u1[3], u2[3], zz[3], d[3];
    double l1 = 0.0, l2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        u1[k] = p[k] - p[3 + k];
        u2[k] = p[6 + k] - p[3 + k];
        l1 += u1[k] * u1[k];
        l2 += u2[k] * u2[k];
    }
    l1 = sqrt(l1);
    l2 = sqrt(l2);
    double lz = 0.0;
    for (int k = 0; k < 3; ++k) {
        zz[k] = u1[k] / l1 + u2[k] / l2;
        lz += zz[k] * zz[k];
    }
    lz = sqrt(lz);
    double dz = 0.0;
    for (int k = 0; k < 3; ++k) {
        zz[k] /= lz;
        d[k] = p[6 + k] - p[k];
        dz += d[k] * zz[k];
    }
    double lx = 0.0;
    for (int k = 0; k < 3; ++k) {
        d[k] -= dz * zz[k];
        lx += d[k] * d[k];
    }
    lx = sqrt(lx);
    for (int k = 0; k < 3; ++k) d[k] /= lx;
    const double yy[3] = {zz[1] * d[2] - zz[2] * d[1], zz[2] * d[0] - zz[0] * d[2], zz[0] * d[1] - zz[1] * d[0]};
    for (int s = 0; s < 3; ++s) {
        double* f = a.frame + 9 * (size_t)(a0 + s);
        const double sx = s == 2 ? -1.0 : 1.0;
        for (int k = 0; k < 3; ++k) {
            f[k] = sx * d[k];
            f[3 + k] = yy[k];
            f[6 + k] = zz[k];
        }
    }
}

__device__ __forceinline__ void rot3(const double* f, const double v[3], double o[3]) {
    for (int k = 0; k < 3; ++k) o[k] = f[3 * k] * v[0] + f[3 * k + 1] * v[1] + f[3 * k + 2] * v[2];
}

// Writes site i's 32 descriptors from its shell sums acc (FP64, site i's
// sums over its partners in the lab frame).
__device__ __forceinline__ void descriptor_row(const DescArgs& a, int i, const double (&acc)[8][3]);

// The shell sums of the mixed solves' descriptors: FP32 pair math from the
// hi/lo records (no position, charge or molecule gathers beyond the two
// records; partners of the same 3-site molecule are j / 3 == i / 3), four
// entries per lane in flight, FP32 sums per lane, FP64 across the lanes.
template <int G>
__global__ __launch_bounds__(256) void k_descriptors32(DescArgs a) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = tid / G, sub = tid % G;
    float acc[8][3];
#pragma unroll
    for (int b = 0; b < 8; ++b) acc[b][0] = acc[b][1] = acc[b][2] = 0.0f;
    if (i < a.n) {
        const int cnt = a.tcnt[i];
        const float4 hi = a.rec32[2 * i], lo = a.rec32[2 * i + 1];
        const int32_t* row = a.tj + (int64_t)i * a.tstride;
        const int mi = i / 3;
        auto term = [&](int j, float4 h, float4 l) {
            const float hx = hi.x - h.x, hy = hi.y - h.y, hz = hi.z - h.z;
            const float nx = rintf(hx * a.invL[0]), ny = rintf(hy * a.invL[1]), nz = rintf(hz * a.invL[2]);
            const float dx = (hx - nx * a.Lh[0]) + ((lo.x - l.x) - nx * a.Ll[0]);
            const float dy = (hy - ny * a.Lh[1]) + ((lo.y - l.y) - ny * a.Ll[1]);
            const float dz = (hz - nz * a.Lh[2]) + ((lo.z - l.z) - nz * a.Ll[2]);
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (j / 3 == mi || !(r2 < 0.36f)) return;
            const int shell = r2 < 0.09f ? 0 : r2 < 0.16f ? 1 : r2 < 0.25f ? 2 : 3;
            const float qj = l.w;
            const int b = 2 * shell + (qj > 0.0f ? 0 : 1);
            const float r_1 = __builtin_amdgcn_rsqf(r2);
            const float sc = qj * r_1 * r_1 * r_1;
#pragma unroll
            for (int bb = 0; bb < 8; ++bb)
                if (bb == b) {
                    acc[bb][0] += sc * dx;
                    acc[bb][1] += sc * dy;
                    acc[bb][2] += sc * dz;
                }
        };
        int k = sub;
        for (; k + 3 * G < cnt; k += 4 * G) {
            int jj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) jj[u] = row[k + u * G];
            float4 h[4], l[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                h[u] = a.rec32[2 * jj[u]];
                l[u] = a.rec32[2 * jj[u] + 1];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) term(jj[u], h[u], l[u]);
        }
        for (; k < cnt; k += G) {
            const int j = row[k];
            term(j, a.rec32[2 * j], a.rec32[2 * j + 1]);
        }
    }
    double acc64[8][3];
#pragma unroll
    for (int b = 0; b < 8; ++b)
#pragma unroll
        for (int x = 0; x < 3; ++x) acc64[b][x] = (double)acc[b][x];
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
        for (int b = 0; b < 8; ++b)
#pragma unroll
            for (int x = 0; x < 3; ++x) acc64[b][x] += __shfl_xor(acc64[b][x], o, G);
    if (i >= a.n || sub != 0) return;
    descriptor_row(a, i, acc64);
}

// Descriptors from the shells the mixed field pass already summed: thread per site.
__global__ __launch_bounds__(256) void k_descriptors_shells(DescArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double acc[8][3];
#pragma unroll
    for (int b = 0; b < 8; ++b)
#pragma unroll
        for (int x = 0; x < 3; ++x) acc[b][x] = (double)a.shells[24 * (size_t)i + 3 * b + x];
    descriptor_row(a, i, acc);
}

template <int G>
__global__ __launch_bounds__(256) void k_descriptors(DescArgs a) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = tid / G, sub = tid % G;
    double acc[8][3];
#pragma unroll
    for (int b = 0; b < 8; ++b) acc[b][0] = acc[b][1] = acc[b][2] = 0.0;
    if (i < a.n) {
        // the coupling table's row (cutoffInd >= 0.6 nm holds every partner
        // the shells need; its excluded partners are skipped as same-molecule)
        const int cnt = a.tcnt[i];
        const double xi = a.pos[3 * i], yi = a.pos[3 * i + 1], zi = a.pos[3 * i + 2];
        const int32_t* row = a.tj + (int64_t)i * a.tstride;
        const int mi = a.mol[i];
        for (int k = sub; k < cnt; k += G) {
            const int j = row[k];
            if (a.mol[j] == mi) continue;
            double dx = xi - a.pos[3 * j], dy = yi - a.pos[3 * j + 1], dz = zi - a.pos[3 * j + 2];
            pbc_near(dx, dy, dz, a.box);
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 < 0.36)) continue;
            const int shell = r2 < 0.09 ? 0 : r2 < 0.16 ? 1 : r2 < 0.25 ? 2 : 3;
            const double qj = a.q[j] / SQRT_COULOMB_LG;
            const int b = 2 * shell + (qj > 0.0 ? 0 : 1);
            const double s = qj / (r2 * sqrt(r2));
#pragma unroll
            for (int bb = 0; bb < 8; ++bb)
                if (bb == b) {
                    acc[bb][0] += s * dx;
                    acc[bb][1] += s * dy;
                    acc[bb][2] += s * dz;
                }
        }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
        for (int b = 0; b < 8; ++b)
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[b][k] += __shfl_xor(acc[b][k], o, G);
    if (i >= a.n || sub != 0) return;
    descriptor_row(a, i, acc);
}

__device__ __forceinline__ void descriptor_row(const DescArgs& a, int i, const double (&acc)[8][3]) {
    const double* F = a.frame + 9 * (size_t)i;
    const double al = a.polar[i];
    const double alc = al > POLMIN_LG ? al : POLMIN_LG;
    float f[MF];
    double v[3], o3[3];
    for (int k = 0; k < 3; ++k) v[k] = a.x[3 * i + k];
    rot3(F, v, o3);
    for (int k = 0; k < 3; ++k) f[k] = (float)o3[k];
    // A x = x / alpha - T x, so alpha T x = alpha (x / alpha - A x)
    for (int k = 0; k < 3; ++k) v[k] = al * (a.x[3 * i + k] / alc - a.z[3 * i + k]);
    rot3(F, v, o3);
    for (int k = 0; k < 3; ++k) f[3 + k] = (float)o3[k];
    for (int b = 0; b < 8; ++b) {
        for (int k = 0; k < 3; ++k) v[k] = al * acc[b][k];
        rot3(F, v, o3);
        for (int k = 0; k < 3; ++k) f[6 + 3 * b + k] = (float)o3[k];
    }
    f[30] = (float)al;
    f[31] = a.q[i] > 0.0 ? 1.0f : 0.0f;
    if (a.raw)
        for (int k = 0; k < MF; ++k) a.raw[(size_t)i * MF + k] = f[k];
    if (a.xn)
        for (int k = 0; k < MF; ++k) a.xn[(size_t)i * MF + k] = f2bf((f[k] - a.norm[k]) / a.norm[MF + k]);
}

// Training targets F_i (Pi_i - alpha_i E0_i).
__global__ __launch_bounds__(256) void k_targets(int n, const double* frame, const double* polar, const double* e0,
                                                 const double* mu, float* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v[3], o[3];
    for (int k = 0; k < 3; ++k) v[k] = mu[3 * i + k] - polar[i] * e0[3 * i + k];
    rot3(frame + 9 * (size_t)i, v, o);
    for (int k = 0; k < 3; ++k) out[(size_t)i * MO + k] = (float)o[k];
}

// Pi_i = alpha_i E0_i + F_i^T (sd * y_i + mean)
__global__ __launch_bounds__(256) void k_guess(int n, const double* frame, const double* polar, const double* e0,
                                               const float* y, const float* norm, double* x, float* x32) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* F = frame + 9 * (size_t)i;
    double loc[3];
    for (int k = 0; k < 3; ++k) loc[k] = (double)norm[2 * MF + MO + k] * y[(size_t)i * MO + k] + norm[2 * MF + k];
    for (int k = 0; k < 3; ++k) {
        const double v = polar[i] * e0[3 * i + k] + F[k] * loc[0] + F[3 + k] * loc[1] + F[6 + k] * loc[2];
        x[3 * i + k] = v;
        if (x32) x32[4 * i + k] = (float)v;   // the mixed solve's FP32 copy (float4 per site)
    }
}

// Per-column mean and standard deviation of rows x cols FP32 data, one block
// per column, FP64 sums in a fixed order.
__global__ __launch_bounds__(256) void k_col_stats(const float* data, int64_t rows, int cols, int col0,
                                                   float* norm, int mean_at, int sd_at) {
    __shared__ double red[256];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int64_t r = threadIdx.x; r < rows; r += 256) s += data[r * cols + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double mean = red[0] / (double)rows;
    __syncthreads();
    double v = 0.0;
    for (int64_t r = threadIdx.x; r < rows; r += 256) {
        const double d = data[r * cols + c] - mean;
        v += d * d;
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double sd = sqrt(red[0] / (double)rows);
        if (!(sd > 1e-6 * fabs(mean)) || !(sd > 1e-30)) sd = 1.0;   // a constant column passes through centred
        norm[mean_at + col0 + c] = (float)mean;
        norm[sd_at + col0 + c] = (float)sd;
    }
}

__global__ __launch_bounds__(256) void k_standardise(int64_t rows, const float* raw_x, const float* raw_t,
                                                     const float* norm, uint16_t* xn, float* tn) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < rows * MF) {
        const int k = (int)(e % MF);
        xn[e] = f2bf((raw_x[e] - norm[k]) / norm[MF + k]);
    }
    if (e < rows * MO) {
        const int k = (int)(e % MO);
        tn[e] = (raw_t[e] - norm[2 * MF + k]) / norm[2 * MF + MO + k];
    }
}

// ---- host state --------------------------------------------------------------------

struct Mlp {
    float* w = nullptr;        // FP32 masters [P_INT]
    float* am = nullptr;       // Adam moments [P_INT]
    float* av = nullptr;
    uint16_t* wb = nullptr;    // BF16 [W_INT]
    uint16_t* wt = nullptr;    // BF16 transposed [W_INT]
    float* norm = nullptr;     // [NNORM]
    bool norm_set = false;
    float* part = nullptr;     // [WG_SPLITS][P_INT]
    float* grad = nullptr;     // [P_INT]
    // row buffers for up to cap rows (a multiple of 16)
    int64_t cap = 0;
    uint16_t* x = nullptr;     // [cap][32]
    float* y = nullptr;        // [cap][3]
    float* tgt = nullptr;      // [cap][3]
    uint16_t* xt = nullptr;    // [32][cap]
    uint16_t* ht = nullptr;    // [3][128][cap]
    uint16_t* dzt = nullptr;   // [3 * 128 + 16][cap]
    double* loss_part = nullptr;   // [cap / 64]
    // per-site buffers of the context's n sites
    double* frame = nullptr;   // [n][9]
    double* xa = nullptr;      // [3n] alpha E0
    double* za = nullptr;      // [3n] A alpha E0
    // training set
    int64_t ds_rows = 0, ds_cap = 0;
    float* ds_x = nullptr;     // [ds_cap][32] raw descriptors
    float* ds_t = nullptr;     // [ds_cap][3] raw targets
    uint16_t* ds_xn = nullptr; // standardised copies
    float* ds_tn = nullptr;
    int32_t* perm = nullptr;   // [ds_cap]
    int64_t adam_t = 0;
};

static void free_rows(Mlp* m) {
    void* ps[] = {m->x, m->y, m->tgt, m->xt, m->ht, m->dzt, m->loss_part};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    m->x = nullptr; m->y = nullptr; m->tgt = nullptr; m->xt = nullptr; m->ht = nullptr; m->dzt = nullptr;
    m->loss_part = nullptr;
    m->cap = 0;
}

static void free_dataset(Mlp* m) {
    void* ps[] = {m->ds_x, m->ds_t, m->ds_xn, m->ds_tn, m->perm};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    m->ds_x = nullptr; m->ds_t = nullptr; m->ds_xn = nullptr; m->ds_tn = nullptr; m->perm = nullptr;
    m->ds_rows = m->ds_cap = 0;
}

void mlp_release(Ctx* c) {
    Mlp* m = c->mlp;
    if (!m) return;
    free_rows(m);
    free_dataset(m);
    void* ps[] = {m->w, m->am, m->av, m->wb, m->wt, m->norm, m->part, m->grad, m->frame, m->xa, m->za};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    delete m;
    c->mlp = nullptr;
}

static int64_t round16(int64_t v) { return (v + 15) / 16 * 16; }

static int ensure_rows(Ctx* c, int64_t rows16) {
    Mlp* m = c->mlp;
    if (rows16 <= m->cap) return SMD_OK;
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    free_rows(m);
    const int64_t cap = std::max<int64_t>(rows16, 1024);
    SMD_HIP(c, hipMalloc(&m->x, sizeof(uint16_t) * MF * cap));
    SMD_HIP(c, hipMalloc(&m->y, sizeof(float) * MO * cap));
    SMD_HIP(c, hipMalloc(&m->tgt, sizeof(float) * MO * cap));
    SMD_HIP(c, hipMalloc(&m->xt, sizeof(uint16_t) * MF * cap));
    SMD_HIP(c, hipMalloc(&m->ht, sizeof(uint16_t) * 3 * MH * cap));
    SMD_HIP(c, hipMalloc(&m->dzt, sizeof(uint16_t) * (3 * MH + MOP) * cap));
    SMD_HIP(c, hipMalloc(&m->loss_part, sizeof(double) * (cap / 64 + 1)));
    // rows 3..15 of layer 4's dL/dz are never written and must read as zero
    SMD_HIP(c, hipMemsetAsync(m->dzt, 0, sizeof(uint16_t) * (3 * MH + MOP) * cap, c->stream));
    m->cap = cap;
    return SMD_OK;
}

static MlpArgs rows_args(Mlp* m, int64_t rows, const int32_t* idx, const uint16_t* x, const float* tgt) {
    MlpArgs a{};
    a.rows = (int)rows;
    a.rows16 = (int)round16(rows);
    a.idx = idx;
    a.x = x;
    a.wb = m->wb;
    a.wt = m->wt;
    a.w = m->w;
    a.y = m->y;
    a.tgt = tgt;
    a.xt = m->xt;
    a.ht = m->ht;
    a.dzt = m->dzt;
    a.loss_part = m->loss_part;
    a.scale = rows > 0 ? (float)(2.0 / (3.0 * (double)rows)) : 0.f;
    return a;
}

static unsigned row_blocks(int64_t rows16) { return (unsigned)((rows16 + 16 * MLP_WAVES - 1) / (16 * MLP_WAVES)); }

// One forward + backward over rows, then the weight gradient's slices into part.
static int train_pass(Ctx* c, const MlpArgs& a) {
    Mlp* m = c->mlp;
    hipLaunchKernelGGL(k_mlp_rows<true>, dim3(row_blocks(a.rows16)), dim3(256), 0, c->stream, a);
    const int slices = std::max(1, std::min(WG_SPLITS, a.rows16 / 16));
    const int chunk = (int)round16((a.rows16 + slices - 1) / slices);
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(WG_TILES, WG_SPLITS), dim3(64), 0, c->stream, a, m->part, chunk);
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

static AdamArgs adam_args(Mlp* m) {
    AdamArgs a{};
    a.w = m->w; a.m = m->am; a.v = m->av; a.wb = m->wb; a.wt = m->wt;
    a.part = m->part; a.splits = WG_SPLITS;
    a.b1 = 0.9f; a.b2 = 0.999f; a.eps = 1e-8f;
    return a;
}

static int refresh_copies(Ctx* c) {
    AdamArgs a = adam_args(c->mlp);
    a.part = nullptr;
    hipLaunchKernelGGL(k_mlp_adam, dim3((P_INT + 255) / 256), dim3(256), 0, c->stream, a);
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

// public order W1 b1 W2 b2 W3 b3 W4[3][128] b4[3] <-> the padded master layout
static void pub_map(std::vector<int>& to) {
    to.clear();
    for (int p = 0; p < MH * MF; ++p) to.push_back(OFF_W1 + p);
    for (int p = 0; p < MH; ++p) to.push_back(OFF_B1 + p);
    for (int p = 0; p < MH * MH; ++p) to.push_back(OFF_W2 + p);
    for (int p = 0; p < MH; ++p) to.push_back(OFF_B2 + p);
    for (int p = 0; p < MH * MH; ++p) to.push_back(OFF_W3 + p);
    for (int p = 0; p < MH; ++p) to.push_back(OFF_B3 + p);
    for (int p = 0; p < MO * MH; ++p) to.push_back(OFF_W4 + p);
    for (int p = 0; p < MO; ++p) to.push_back(OFF_B4 + p);
}

static DescArgs desc_args(Ctx* c) {
    DescArgs d{};
    d.n = c->n;
    d.nmol = c->nmol;
    d.stride = c->stride;
    d.box = c->box;
    d.mol_start = c->d_mol_start;
    d.mol = c->d_mol;
    d.pos = c->d_pos;
    d.q = c->d_q;
    d.polar = c->d_polar;
    d.e0 = c->d_e0;
    d.tstride = c->it_stride;
    d.tcnt = c->d_it_cnt;
    d.tj = c->d_it_j;
    d.frame = c->mlp->frame;
    d.norm = c->mlp->norm;
    d.rec32 = c->ind_last.prec == SMD_PREC_MIXED ? c->d_ind_rec32 : nullptr;
    d.shells = c->ind_shells_valid ? c->d_shell32 : nullptr;
    for (int k = 0; k < 3; ++k) {
        const double L = c->box.L[k];
        d.Lh[k] = (float)L;
        d.Ll[k] = (float)(L - (double)d.Lh[k]);
        d.invL[k] = (float)(1.0 / L);
    }
    return d;
}

static int check_sites(Ctx* c) {
    for (int mm = 0; mm < c->nmol; ++mm)
        if (c->h_mol_start[mm + 1] - c->h_mol_start[mm] != 3)
            return fail(c, SMD_ENOTSUP, "the learned guess is implemented for 3-site water");
    if (!c->list_valid) return fail(c, SMD_ESTATE, "no pair list");
    if (c->list_cutoff < 0.6)
        return fail(c, SMD_ENOTSUP, "the learned guess's descriptors need a pair-list cutoff of at least 0.6 nm");
    if (!c->d_it_cnt || !c->ind_last.do_min || c->ind_last.cutoff_ind < 0.6)
        return fail(c, SMD_ENOTSUP, "the learned guess needs a CG solve with cutoffInd of at least 0.6 nm");
    return SMD_OK;
}

// Frames and descriptors of the current state from x = alpha E0 and z = A x.
static int descriptors_run(Ctx* c, const double* x, const double* z, float* raw, uint16_t* xn) {
    DescArgs d = desc_args(c);
    d.x = x;
    d.z = z;
    d.raw = raw;
    d.xn = xn;
    hipLaunchKernelGGL(k_site_frames, dim3((c->nmol + 255) / 256), dim3(256), 0, c->stream, d);
    const unsigned grid = (unsigned)(((int64_t)c->n * DESC_LANES + 255) / 256);
    if (d.shells) hipLaunchKernelGGL(k_descriptors_shells, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, d);
    else if (d.rec32) hipLaunchKernelGGL(k_descriptors32<DESC_LANES>, dim3(grid), dim3(256), 0, c->stream, d);
    else hipLaunchKernelGGL(k_descriptors<DESC_LANES>, dim3(grid), dim3(256), 0, c->stream, d);
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

// The guess of smd_induce: x holds alpha E0 and z = A x on entry; x gets the guess.
int mlp_guess(Ctx* c, double* x, const double* z, float* x32) {
    Mlp* m = c->mlp;
    if (!m || !m->norm_set) return fail(c, SMD_ESTATE, "learned guess: no trained network (smd_mlp_train or smd_mlp_set_norm)");
    int rc;
    if ((rc = check_sites(c))) return rc;
    if ((rc = ensure_rows(c, round16(c->n)))) return rc;
    if ((rc = descriptors_run(c, x, z, nullptr, m->x))) return rc;
    MlpArgs a = rows_args(m, c->n, nullptr, m->x, nullptr);
    hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((a.rows16 + 16 * INF_WAVES - 1) / (16 * INF_WAVES))),
                       dim3(INF_WAVES * WAVE), 0, c->stream, a);
    hipLaunchKernelGGL(k_guess, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->n, m->frame, c->d_polar,
                       c->d_e0, m->y, m->norm, x, x32);
    SMD_HIP(c, hipGetLastError());
    return SMD_OK;
}

static int require(Ctx* c) {
    if (!c->mlp) return fail(c, SMD_ESTATE, "learned guess not initialised (smd_mlp_init)");
    return SMD_OK;
}

// x = alpha E0 and z = A x of the last converged smd_induce, then descriptors.
static int state_descriptors(Ctx* c, float* raw) {
    if (!c->ind_last_converged) return fail(c, SMD_ESTATE, "no converged smd_induce on this state");
    int rc;
    if ((rc = check_sites(c))) return rc;
    Mlp* m = c->mlp;
    if ((rc = induce_direct_apply(c, &c->ind_last, m->xa, m->za))) return rc;
    return descriptors_run(c, m->xa, m->za, raw, nullptr);
}

}  // namespace smd

using namespace smd;

extern "C" {

int smd_mlp_init(smd_ctx* c, uint64_t seed) {
    if (!c) return SMD_EINVAL;
    if (c->n <= 0) return fail(c, SMD_ESTATE, "no system");
    if (!c->mlp) {
        Mlp* m = new Mlp();
        c->mlp = m;
        SMD_HIP(c, hipMalloc(&m->w, sizeof(float) * P_INT));
        SMD_HIP(c, hipMalloc(&m->am, sizeof(float) * P_INT));
        SMD_HIP(c, hipMalloc(&m->av, sizeof(float) * P_INT));
        SMD_HIP(c, hipMalloc(&m->wb, sizeof(uint16_t) * W_INT));
        SMD_HIP(c, hipMalloc(&m->wt, sizeof(uint16_t) * W_INT));
        SMD_HIP(c, hipMalloc(&m->norm, sizeof(float) * NNORM));
        SMD_HIP(c, hipMalloc(&m->part, sizeof(float) * WG_SPLITS * P_INT));
        SMD_HIP(c, hipMalloc(&m->grad, sizeof(float) * P_INT));
        SMD_HIP(c, hipMalloc(&m->frame, sizeof(double) * 9 * (size_t)c->n));
        SMD_HIP(c, hipMalloc(&m->xa, sizeof(double) * 3 * (size_t)c->n));
        SMD_HIP(c, hipMalloc(&m->za, sizeof(double) * 3 * (size_t)c->n));
    }
    Mlp* m = c->mlp;
    // uniform +-sqrt(6 / fan_in) weights, the last layer's scaled by 0.1; zero biases
    std::vector<float> w(P_INT, 0.f);
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    const int offs[4] = {OFF_W1, OFF_W2, OFF_W3, OFF_W4}, ins[4] = {MF, MH, MH, MH}, outs[4] = {MH, MH, MH, MO};
    for (int l = 0; l < 4; ++l) {
        const double lim = std::sqrt(6.0 / ins[l]) * (l == 3 ? 0.1 : 1.0);
        for (int o = 0; o < outs[l]; ++o)
            for (int i = 0; i < ins[l]; ++i) w[offs[l] + o * ins[l] + i] = (float)(lim * u(rng));
    }
    std::vector<float> norm(NNORM, 0.f);
    for (int k = 0; k < MF; ++k) norm[MF + k] = 1.f;
    for (int k = 0; k < MO; ++k) norm[2 * MF + MO + k] = 1.f;
    SMD_HIP(c, hipMemcpyAsync(m->w, w.data(), sizeof(float) * P_INT, hipMemcpyHostToDevice, c->stream));
    SMD_HIP(c, hipMemcpyAsync(m->norm, norm.data(), sizeof(float) * NNORM, hipMemcpyHostToDevice, c->stream));
    SMD_HIP(c, hipMemsetAsync(m->am, 0, sizeof(float) * P_INT, c->stream));
    SMD_HIP(c, hipMemsetAsync(m->av, 0, sizeof(float) * P_INT, c->stream));
    m->norm_set = false;
    m->adam_t = 0;
    int rc = refresh_copies(c);
    if (rc) return rc;
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    return SMD_OK;
}

int smd_mlp_set_params(smd_ctx* c, const float* params, int64_t count) {
    if (!c || !params) return SMD_EINVAL;
    int rc;
    if ((rc = require(c))) return rc;
    if (count != NPUB) return fail(c, SMD_EINVAL, "smd_mlp_set_params: %lld parameters, expected %d", (long long)count, NPUB);
    std::vector<float> w(P_INT, 0.f);
    std::vector<int> to;
    pub_map(to);
    for (int k = 0; k < NPUB; ++k) w[to[k]] = params[k];
    SMD_HIP(c, hipMemcpyAsync(c->mlp->w, w.data(), sizeof(float) * P_INT, hipMemcpyHostToDevice, c->stream));
    if ((rc = refresh_copies(c))) return rc;
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    return SMD_OK;
}

int smd_mlp_get_params(smd_ctx* c, float* params, int64_t count) {
    if (!c || !params) return SMD_EINVAL;
    int rc;
    if ((rc = require(c))) return rc;
    if (count != NPUB) return fail(c, SMD_EINVAL, "smd_mlp_get_params: %lld parameters, expected %d", (long long)count, NPUB);
    std::vector<float> w(P_INT);
    SMD_HIP(c, hipMemcpyAsync(w.data(), c->mlp->w, sizeof(float) * P_INT, hipMemcpyDeviceToHost, c->stream));
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<int> to;
    pub_map(to);
    for (int k = 0; k < NPUB; ++k) params[k] = w[to[k]];
    return SMD_OK;
}

int smd_mlp_set_norm(smd_ctx* c, const float* norm, int64_t count) {
    if (!c || !norm) return SMD_EINVAL;
    int rc;
    if ((rc = require(c))) return rc;
    if (count != NNORM) return fail(c, SMD_EINVAL, "smd_mlp_set_norm: %lld values, expected %d", (long long)count, NNORM);
    for (int k = 0; k < MF; ++k)
        if (!(norm[MF + k] > 0.f)) return fail(c, SMD_EINVAL, "smd_mlp_set_norm: feature sd %d is not positive", k);
    for (int k = 0; k < MO; ++k)
        if (!(norm[2 * MF + MO + k] > 0.f)) return fail(c, SMD_EINVAL, "smd_mlp_set_norm: output sd %d is not positive", k);
    SMD_HIP(c, hipMemcpyAsync(c->mlp->norm, norm, sizeof(float) * NNORM, hipMemcpyHostToDevice, c->stream));
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    c->mlp->norm_set = true;
    return SMD_OK;
}

int smd_mlp_get_norm(smd_ctx* c, float* norm, int64_t count) {
    if (!c || !norm) return SMD_EINVAL;
    int rc;
    if ((rc = require(c))) return rc;
    if (count != NNORM) return fail(c, SMD_EINVAL, "smd_mlp_get_norm: %lld values, expected %d", (long long)count, NNORM);
    SMD_HIP(c, hipMemcpyAsync(norm, c->mlp->norm, sizeof(float) * NNORM, hipMemcpyDeviceToHost, c->stream));
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    return SMD_OK;
}

static int upload_rows(Ctx* c, const float* x, int64_t rows) {
    std::vector<uint16_t> xb((size_t)rows * MF);
    for (size_t k = 0; k < xb.size(); ++k) {   // the kernels' rounding, on the host
        uint32_t u;
        std::memcpy(&u, &x[k], 4);
        u += 0x7FFFu + ((u >> 16) & 1u);
        xb[k] = (uint16_t)(u >> 16);
    }
    SMD_HIP(c, hipMemcpyAsync(c->mlp->x, xb.data(), sizeof(uint16_t) * xb.size(), hipMemcpyHostToDevice, c->stream));
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    return SMD_OK;
}

int smd_mlp_forward(smd_ctx* c, const float* x, int64_t rows, float* y) {
    if (!c || !x || !y) return SMD_EINVAL;
    int rc;
    if ((rc = require(c))) return rc;
    if (rows <= 0 || rows > (1 << 26)) return fail(c, SMD_EINVAL, "smd_mlp_forward: rows %lld", (long long)rows);
    if ((rc = ensure_rows(c, round16(rows)))) return rc;
    if ((rc = upload_rows(c, x, rows))) return rc;
    Mlp* m = c->mlp;
    MlpArgs a = rows_args(m, rows, nullptr, m->x, nullptr);
    hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((a.rows16 + 16 * INF_WAVES - 1) / (16 * INF_WAVES))),
                       dim3(INF_WAVES * WAVE), 0, c->stream, a);
    SMD_HIP(c, hipGetLastError());
    SMD_HIP(c, hipMemcpyAsync(y, m->y, sizeof(float) * MO * rows, hipMemcpyDeviceToHost, c->stream));
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    return SMD_OK;
}

int smd_mlp_loss_grad(smd_ctx* c, const float* x, const float* target, int64_t rows, double* loss, float* grad) {
    if (!c || !x || !target) return SMD_EINVAL;
    int rc;
    if ((rc = require(c))) return rc;
    if (rows <= 0 || rows > (1 << 26)) return fail(c, SMD_EINVAL, "smd_mlp_loss_grad: rows %lld", (long long)rows);
    Mlp* m = c->mlp;
    if ((rc = ensure_rows(c, round16(rows)))) return rc;
    if ((rc = upload_rows(c, x, rows))) return rc;
    SMD_HIP(c, hipMemcpyAsync(m->tgt, target, sizeof(float) * MO * rows, hipMemcpyHostToDevice, c->stream));
    MlpArgs a = rows_args(m, rows, nullptr, m->x, m->tgt);
    if ((rc = train_pass(c, a))) return rc;
    AdamArgs ad = adam_args(m);
    ad.grad_out = m->grad;
    hipLaunchKernelGGL(k_mlp_adam, dim3((P_INT + 255) / 256), dim3(256), 0, c->stream, ad);
    SMD_HIP(c, hipGetLastError());
    const unsigned nb = row_blocks(a.rows16);
    std::vector<double> lp(nb);
    std::vector<float> g(P_INT);
    SMD_HIP(c, hipMemcpyAsync(lp.data(), m->loss_part, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
    SMD_HIP(c, hipMemcpyAsync(g.data(), m->grad, sizeof(float) * P_INT, hipMemcpyDeviceToHost, c->stream));
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    if (loss) *loss = std::accumulate(lp.begin(), lp.end(), 0.0) / (3.0 * (double)rows);
    if (grad) {
        std::vector<int> to;
        pub_map(to);
        for (int k = 0; k < NPUB; ++k) grad[k] = g[to[k]];
    }
    return SMD_OK;
}

int smd_mlp_descriptors(smd_ctx* c, float* desc, int64_t n_atoms) {
    if (!c || !desc) return SMD_EINVAL;
    int rc;
    if ((rc = require(c))) return rc;
    if (n_atoms != c->n) return fail(c, SMD_EINVAL, "smd_mlp_descriptors: %lld atoms, context has %d", (long long)n_atoms, c->n);
    if ((rc = ensure_rows(c, round16(c->n)))) return rc;
    float* raw = nullptr;
    SMD_HIP(c, hipMalloc(&raw, sizeof(float) * MF * (size_t)c->n));
    rc = state_descriptors(c, raw);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(desc, raw, sizeof(float) * MF * (size_t)c->n, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(c, e, "smd_mlp_descriptors copy");
    }
    (void)hipFree(raw);
    return rc;
}

int smd_mlp_add_frame(smd_ctx* c, int64_t* total_rows) {
    if (!c) return SMD_EINVAL;
    int rc;
    if ((rc = require(c))) return rc;
    Mlp* m = c->mlp;
    const int64_t need = m->ds_rows + c->n;
    if (need > m->ds_cap) {
        const int64_t cap = std::max<int64_t>(need, 2 * m->ds_cap);
        float *nx = nullptr, *nt = nullptr;
        SMD_HIP(c, hipMalloc(&nx, sizeof(float) * MF * cap));
        SMD_HIP(c, hipMalloc(&nt, sizeof(float) * MO * cap));
        if (m->ds_rows) {
            SMD_HIP(c, hipMemcpyAsync(nx, m->ds_x, sizeof(float) * MF * m->ds_rows, hipMemcpyDeviceToDevice, c->stream));
            SMD_HIP(c, hipMemcpyAsync(nt, m->ds_t, sizeof(float) * MO * m->ds_rows, hipMemcpyDeviceToDevice, c->stream));
        }
        SMD_HIP(c, hipStreamSynchronize(c->stream));
        const int64_t keep = m->ds_rows;
        free_dataset(m);
        m->ds_x = nx;
        m->ds_t = nt;
        m->ds_rows = keep;
        m->ds_cap = cap;
    }
    if ((rc = state_descriptors(c, m->ds_x + (size_t)MF * m->ds_rows))) return rc;
    hipLaunchKernelGGL(k_targets, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->n, m->frame, c->d_polar,
                       c->d_e0, c->d_mu, m->ds_t + (size_t)MO * m->ds_rows);
    SMD_HIP(c, hipGetLastError());
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    m->ds_rows = need;
    if (total_rows) *total_rows = need;
    return SMD_OK;
}

int smd_mlp_clear_frames(smd_ctx* c) {
    if (!c) return SMD_EINVAL;
    int rc;
    if ((rc = require(c))) return rc;
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    free_dataset(c->mlp);
    return SMD_OK;
}

int smd_mlp_train(smd_ctx* c, const smd_mlp_train_params* p, smd_mlp_train_report* rep) {
    if (!c || !p) return SMD_EINVAL;
    int rc;
    if ((rc = require(c))) return rc;
    Mlp* m = c->mlp;
    if (m->ds_rows < 16) return fail(c, SMD_ESTATE, "smd_mlp_train: no training frames (smd_mlp_add_frame)");
    if (p->epochs <= 0 || p->batch < 16 || p->batch > (1 << 20) || !(p->lr > 0.f) || !(p->lr_final >= 0.f))
        return fail(c, SMD_EINVAL, "smd_mlp_train: epochs %d, batch %d, lr %g, lr_final %g", p->epochs, p->batch,
                    (double)p->lr, (double)p->lr_final);
    const int64_t N = m->ds_rows;
    const int64_t B = std::min<int64_t>(p->batch, N);
    if ((rc = ensure_rows(c, round16(B)))) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SMD_HIP(c, hipEventCreate(&e0));
    SMD_HIP(c, hipEventCreate(&e1));
    SMD_HIP(c, hipEventRecord(e0, c->stream));
    if (p->renormalize || !m->norm_set) {
        hipLaunchKernelGGL(k_col_stats, dim3(MF), dim3(256), 0, c->stream, m->ds_x, N, MF, 0, m->norm, 0, MF);
        hipLaunchKernelGGL(k_col_stats, dim3(MO), dim3(256), 0, c->stream, m->ds_t, N, MO, 0, m->norm, 2 * MF,
                           2 * MF + MO);
        SMD_HIP(c, hipGetLastError());
        m->norm_set = true;
    }
    if (!m->ds_xn) {
        SMD_HIP(c, hipMalloc(&m->ds_xn, sizeof(uint16_t) * MF * m->ds_cap));
        SMD_HIP(c, hipMalloc(&m->ds_tn, sizeof(float) * MO * m->ds_cap));
        SMD_HIP(c, hipMalloc(&m->perm, sizeof(int32_t) * m->ds_cap));
    }
    hipLaunchKernelGGL(k_standardise, dim3((unsigned)((N * MF + 255) / 256)), dim3(256), 0, c->stream, N, m->ds_x,
                       m->ds_t, m->norm, m->ds_xn, m->ds_tn);
    SMD_HIP(c, hipMemsetAsync(m->am, 0, sizeof(float) * P_INT, c->stream));
    SMD_HIP(c, hipMemsetAsync(m->av, 0, sizeof(float) * P_INT, c->stream));
    m->adam_t = 0;
    const int64_t per_epoch = (N + B - 1) / B;
    const int64_t total = per_epoch * p->epochs;
    std::vector<int32_t> perm(N);
    std::mt19937_64 rng(p->seed);
    AdamArgs ad = adam_args(m);
    ad.update = 1;
    double loss_first = 0.0, loss_last = 0.0;
    for (int ep = 0; ep < p->epochs; ++ep) {
        std::iota(perm.begin(), perm.end(), 0);
        for (int64_t k = N - 1; k > 0; --k) {   // Fisher-Yates on the seeded generator
            const int64_t j = (int64_t)(rng() % (uint64_t)(k + 1));
            std::swap(perm[k], perm[j]);
        }
        SMD_HIP(c, hipMemcpyAsync(m->perm, perm.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, c->stream));
        SMD_HIP(c, hipStreamSynchronize(c->stream));   // perm is reused next epoch
        double ep_loss = 0.0;
        const bool record = ep == 0 || ep == p->epochs - 1;
        for (int64_t b = 0; b < per_epoch; ++b) {
            const int64_t r0 = b * B, rows = std::min(B, N - r0);
            MlpArgs a = rows_args(m, rows, m->perm + r0, m->ds_xn, m->ds_tn);
            if ((rc = train_pass(c, a))) break;
            ++m->adam_t;
            const double t = (double)m->adam_t;
            const double frac = (double)(m->adam_t - 1) / (double)std::max<int64_t>(1, total - 1);
            ad.lr = (float)(p->lr_final + 0.5 * (p->lr - p->lr_final) * (1.0 + std::cos(3.14159265358979323846 * frac)));
            ad.c1 = (float)(1.0 / (1.0 - std::pow(0.9, t)));
            ad.c2 = (float)(1.0 / (1.0 - std::pow(0.999, t)));
            hipLaunchKernelGGL(k_mlp_adam, dim3((P_INT + 255) / 256), dim3(256), 0, c->stream, ad);
            if (record) {
                const unsigned nb = row_blocks(a.rows16);
                std::vector<double> lp(nb);
                SMD_HIP(c, hipMemcpyAsync(lp.data(), m->loss_part, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
                SMD_HIP(c, hipStreamSynchronize(c->stream));
                ep_loss += std::accumulate(lp.begin(), lp.end(), 0.0);
            }
        }
        if (rc) break;
        SMD_HIP(c, hipGetLastError());
        if (ep == 0) loss_first = ep_loss / (3.0 * (double)N);
        if (ep == p->epochs - 1) loss_last = ep_loss / (3.0 * (double)N);
    }
    SMD_HIP(c, hipEventRecord(e1, c->stream));
    SMD_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (rep) {
        rep->samples = N;
        rep->steps = (int32_t)total;
        rep->loss_first = loss_first;
        rep->loss_last = loss_last;
        rep->ms = ms;
    }
    return SMD_OK;
}

}  // extern "C"
