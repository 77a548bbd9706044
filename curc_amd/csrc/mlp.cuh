// Candidate scorer network on gfx950: 32 -> 128 -> 128 -> 1 with ReLU,
// bfloat16 operands and float32 accumulation on the matrix cores
// (v_mfma_f32_32x32x16_bf16), float32 master weights and Adam.
//
// MFMA 32x32x16 operand layout (one wave): lane l holds
//   A[l % 32][8 * (l / 32) + 0..7]        (A is 32 x 16, M x K)
//   B[8 * (l / 32) + 0..7][l % 32]        (B is 16 x 32, K x N)
// and accumulator register r of lane l is
//   D[8 * (r / 4) + 4 * (l / 32) + (r % 4)][l % 32].
// Operands are staged so that each lane's 8 K-values are 16 contiguous
// bytes: weights transposed to [out][in] for the forward products, kept
// [in][out] for the input-gradient product, activations and gradients
// transposed to [unit][sample] for the weight-gradient products.
#pragma once
#include "kernels.cuh"

namespace curc {

constexpr uint32_t MLP_IN = 32, MLP_H = 128;
// flat parameter layout (float32): W1[32][128], b1[128], W2[128][128], b2[128], w3[128], b3
constexpr uint32_t MLP_OFF_W1 = 0;
constexpr uint32_t MLP_OFF_B1 = MLP_OFF_W1 + MLP_IN * MLP_H;
constexpr uint32_t MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr uint32_t MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr uint32_t MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr uint32_t MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;
constexpr uint32_t MLP_NPARAMS = MLP_OFF_B3 + 1;
constexpr uint32_t MLP_TILE = 32;  // samples per wave and MFMA tile

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// float -> bfloat16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ floatx16 mfma(bf16x8 a, bf16x8 b, floatx16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ uint32_t acc_row(uint32_t r, uint32_t lane) {
    return 8u * (r >> 2) + 4u * (lane >> 5) + (r & 3u);
}

inline uint32_t mlp_grid(uint64_t n) {
    const uint64_t tiles = (n + MLP_TILE - 1) / MLP_TILE;
    return (uint32_t)(tiles < 4 ? 1 : ((tiles + 3) / 4 > 1024 ? 1024 : (tiles + 3) / 4));
}

// Device copies of one network: float32 master parameters, Adam moments,
// gradient accumulator, and bfloat16 weight copies in both orders.
struct MlpDev {
    float* p;        // [MLP_NPARAMS]
    float* m;        // Adam first moment
    float* v;        // Adam second moment
    float* g;        // gradient accumulator
    uint16_t* w1t;   // [128][32]  W1 transposed
    uint16_t* w2t;   // [128][128] W2 transposed ([out][in])
    uint16_t* w2;    // [128][128] W2 ([in][out])
};

// bfloat16 copies of the weights after an update.
__global__ void k_mlp_pack(MlpDev net) {
    GRID_STRIDE(i, (uint64_t)MLP_IN * MLP_H) {
        const uint32_t in = (uint32_t)i / MLP_H, out = (uint32_t)i % MLP_H;
        net.w1t[out * MLP_IN + in] = f2bf(net.p[MLP_OFF_W1 + i]);
    }
    GRID_STRIDE(i, (uint64_t)MLP_H * MLP_H) {
        const uint32_t in = (uint32_t)i / MLP_H, out = (uint32_t)i % MLP_H;
        const uint16_t b = f2bf(net.p[MLP_OFF_W2 + i]);
        net.w2t[out * MLP_H + in] = b;
        net.w2[i] = b;
    }
}

// Forward of one 32-sample tile held by one wave.  x: the tile's features
// as bf16 [32][32] (sample-major) in LDS; w1t / w2t in LDS; h1 scratch
// [32][128] bf16 in LDS.  Returns the lane's accumulators of layer 2
// (pre-activation z2, bias included) in acc2[4] (column blocks of 32) and
// writes h1 = relu(z1) to LDS.
__device__ __forceinline__ void mlp_tile_forward(const uint16_t* x, const uint16_t* w1t, const uint16_t* w2t,
                                                 const float* b1, const float* b2, uint16_t* h1,
                                                 floatx16 (&acc2)[4]) {
    const uint32_t lane = __lane_id();
    const uint32_t lr = lane & 31u, lk = 8u * (lane >> 5);
    // layer 1: [32 x 32] . [32 x 128]
#pragma unroll
    for (uint32_t nb = 0; nb < 4; ++nb) {
        floatx16 acc = {};
#pragma unroll
        for (uint32_t k0 = 0; k0 < MLP_IN; k0 += 16)
            acc = mfma(load8(x + lr * MLP_IN + k0 + lk), load8(w1t + (nb * 32 + lr) * MLP_IN + k0 + lk), acc);
        const uint32_t col = nb * 32 + lr;
        const float bias = b1[col];
#pragma unroll
        for (uint32_t r = 0; r < 16; ++r) {
            const float z = acc[r] + bias;
            h1[acc_row(r, lane) * MLP_H + col] = f2bf(z > 0.f ? z : 0.f);
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    // layer 2: [32 x 128] . [128 x 128]
#pragma unroll
    for (uint32_t nb = 0; nb < 4; ++nb) {
        floatx16 acc = {};
#pragma unroll
        for (uint32_t k0 = 0; k0 < MLP_H; k0 += 16)
            acc = mfma(load8(h1 + lr * MLP_H + k0 + lk), load8(w2t + (nb * 32 + lr) * MLP_H + k0 + lk), acc);
        const float bias = b2[nb * 32 + lr];
#pragma unroll
        for (uint32_t r = 0; r < 16; ++r) acc[r] += bias;
        acc2[nb] = acc;
    }
}

// Logits of a tile from layer-2 accumulators: s[row] = relu(z2[row]) . w3 + b3.
// part: LDS scratch [32][33] floats.  Lanes 0..31 return the logit of
// sample `lane`.
__device__ __forceinline__ float mlp_tile_logit(const floatx16 (&acc2)[4], const float* w3, float b3, float* part) {
    const uint32_t lane = __lane_id();
    const uint32_t lr = lane & 31u;
    float ps[16];
#pragma unroll
    for (uint32_t r = 0; r < 16; ++r) ps[r] = 0.f;
#pragma unroll
    for (uint32_t nb = 0; nb < 4; ++nb) {
        const float w = w3[nb * 32 + lr];
#pragma unroll
        for (uint32_t r = 0; r < 16; ++r) {
            const float h = acc2[nb][r] > 0.f ? acc2[nb][r] : 0.f;
            ps[r] += h * w;
        }
    }
#pragma unroll
    for (uint32_t r = 0; r < 16; ++r) part[acc_row(r, lane) * 33 + lr] = ps[r];
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    float s = b3;
    if (lane < 32)
        for (uint32_t c = 0; c < 32; ++c) s += part[lane * 33 + c];
    return s;
}

// LDS layout shared by the forward kernels (per workgroup of 4 waves).
struct alignas(16) MlpLds {
    alignas(16) uint16_t w1t[MLP_H * MLP_IN];
    alignas(16) uint16_t w2t[MLP_H * MLP_H];
    alignas(16) float b1[MLP_H];
    alignas(16) float b2[MLP_H];
    alignas(16) float w3[MLP_H];
    alignas(16) uint16_t x[4][MLP_TILE * MLP_IN];
    alignas(16) uint16_t h1[4][MLP_TILE * MLP_H];
    alignas(16) float part[4][MLP_TILE * 33];
};

__device__ __forceinline__ void mlp_stage_weights(const MlpDev& net, MlpLds& L) {
    for (uint32_t i = threadIdx.x; i < MLP_H * MLP_IN; i += blockDim.x) L.w1t[i] = net.w1t[i];
    for (uint32_t i = threadIdx.x; i < MLP_H * MLP_H; i += blockDim.x) L.w2t[i] = net.w2t[i];
    for (uint32_t i = threadIdx.x; i < MLP_H; i += blockDim.x) {
        L.b1[i] = net.p[MLP_OFF_B1 + i];
        L.b2[i] = net.p[MLP_OFF_B2 + i];
        L.w3[i] = net.p[MLP_OFF_W3 + i];
    }
}

// Inference over n samples (x: float32 [n][32]); writes logits.
// Workgroups of MLP_INF_WAVES waves share one LDS copy of the bf16 weights;
// each wave keeps only its tile's hidden layer in LDS (8 KB, reused for
// the logit partials), and takes its tile's inputs straight from global
// memory into the MFMA operand registers.  136 KB of LDS per workgroup,
// one workgroup per CU, MLP_INF_WAVES waves per CU.
constexpr uint32_t MLP_INF_WAVES = 12;

struct alignas(16) MlpInferLds {
    alignas(16) uint16_t w1t[MLP_H * MLP_IN];
    alignas(16) uint16_t w2t[MLP_H * MLP_H];
    alignas(16) uint16_t h1[MLP_INF_WAVES][MLP_TILE * MLP_H];
};
static_assert(sizeof(float) * MLP_TILE * 33 <= sizeof(uint16_t) * MLP_TILE * MLP_H, "logit partials fit h1");

inline uint32_t mlp_infer_grid(uint64_t n) {
    const uint64_t tiles = (n + MLP_TILE - 1) / MLP_TILE;
    const uint64_t wg = (tiles + MLP_INF_WAVES - 1) / MLP_INF_WAVES;
    return (uint32_t)(wg < 1 ? 1 : (wg > 512 ? 512 : wg));
}

// A operand of layer 1 for this lane: features 8*(lane/32) + k0 + 0..7 of
// sample s0 + lane % 32 (zeros past n).
__device__ __forceinline__ bf16x8 mlp_load_x(const float* __restrict__ x, uint64_t n, uint64_t s0, uint32_t k0) {
    const uint32_t lane = __lane_id();
    const uint64_t smp = s0 + (lane & 31u);
    bf16x8 a;
    if (smp < n) {
        const float4* q = reinterpret_cast<const float4*>(x + smp * MLP_IN + k0 + 8u * (lane >> 5));
        const float4 u = q[0], v = q[1];
        const uint16_t b[8] = {f2bf(u.x), f2bf(u.y), f2bf(u.z), f2bf(u.w), f2bf(v.x), f2bf(v.y), f2bf(v.z), f2bf(v.w)};
        uint4 p;
        p.x = (uint32_t)b[0] | ((uint32_t)b[1] << 16);
        p.y = (uint32_t)b[2] | ((uint32_t)b[3] << 16);
        p.z = (uint32_t)b[4] | ((uint32_t)b[5] << 16);
        p.w = (uint32_t)b[6] | ((uint32_t)b[7] << 16);
        a = __builtin_bit_cast(bf16x8, p);
    } else {
        a = __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0));
    }
    return a;
}

__global__ void __launch_bounds__(64 * MLP_INF_WAVES) k_mlp_infer(MlpDev net, const float* __restrict__ x,
                                                                   uint64_t n_host, const uint32_t* __restrict__ d_n,
                                                                   float* __restrict__ logit) {
    __shared__ MlpInferLds L;
    const uint64_t n = d_n ? (uint64_t)*d_n : n_host;
    for (uint32_t i = threadIdx.x; i < MLP_H * MLP_IN; i += blockDim.x) L.w1t[i] = net.w1t[i];
    for (uint32_t i = threadIdx.x; i < MLP_H * MLP_H; i += blockDim.x) L.w2t[i] = net.w2t[i];
    __syncthreads();
    const uint32_t w = threadIdx.x >> 6, lane = __lane_id();
    const uint32_t lr = lane & 31u, lk = 8u * (lane >> 5);
    const float* b1 = net.p + MLP_OFF_B1;
    const float* b2 = net.p + MLP_OFF_B2;
    const float* w3 = net.p + MLP_OFF_W3;
    const float b3 = net.p[MLP_OFF_B3];
    uint16_t* h1 = L.h1[w];
    const uint64_t ntile = (n + MLP_TILE - 1) / MLP_TILE;
    for (uint64_t t = (uint64_t)blockIdx.x * MLP_INF_WAVES + w; t < ntile; t += (uint64_t)gridDim.x * MLP_INF_WAVES) {
        const uint64_t s0 = t * MLP_TILE;
        const bf16x8 a0 = mlp_load_x(x, n, s0, 0), a1 = mlp_load_x(x, n, s0, 16);
        // layer 1: [32 x 32] . [32 x 128] -> relu -> h1 (bf16, LDS)
#pragma unroll
        for (uint32_t nb = 0; nb < 4; ++nb) {
            floatx16 acc = {};
            acc = mfma(a0, load8(L.w1t + (nb * 32 + lr) * MLP_IN + 0 + lk), acc);
            acc = mfma(a1, load8(L.w1t + (nb * 32 + lr) * MLP_IN + 16 + lk), acc);
            const uint32_t col = nb * 32 + lr;
            const float bias = b1[col];
#pragma unroll
            for (uint32_t r = 0; r < 16; ++r) {
                const float z = acc[r] + bias;
                h1[acc_row(r, lane) * MLP_H + col] = f2bf(z > 0.f ? z : 0.f);
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        // layer 2: [32 x 128] . [128 x 128]
        floatx16 acc2[4];
#pragma unroll
        for (uint32_t nb = 0; nb < 4; ++nb) {
            floatx16 acc = {};
#pragma unroll
            for (uint32_t k0 = 0; k0 < MLP_H; k0 += 16)
                acc = mfma(load8(h1 + lr * MLP_H + k0 + lk), load8(L.w2t + (nb * 32 + lr) * MLP_H + k0 + lk), acc);
            const float bias = b2[nb * 32 + lr];
#pragma unroll
            for (uint32_t r = 0; r < 16; ++r) acc[r] += bias;
            acc2[nb] = acc;
        }
        // every lane's reads of h1 are done before the partials overwrite it
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        const float sc = mlp_tile_logit(acc2, w3, b3, reinterpret_cast<float*>(h1));
        if (lane < 32 && s0 + lane < n) logit[s0 + lane] = sc;
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
    }
}

// Training forward: per sample the bf16 inputs transposed (xt [32][n]),
// h1 transposed (h1t [128][n]), relu(z2) as float32 (h2 [n][128]) and
// dlogit = (sigmoid(s) - y) / n_total; logits too.
__global__ void __launch_bounds__(256) k_mlp_train_fwd(MlpDev net, const float* __restrict__ x,
                                                       const float* __restrict__ y, uint64_t n, float inv_total,
                                                       uint16_t* __restrict__ xt, uint16_t* __restrict__ h1t,
                                                       float* __restrict__ h2, float* __restrict__ ds,
                                                       float* __restrict__ logit) {
    __shared__ MlpLds L;
    mlp_stage_weights(net, L);
    __syncthreads();
    const uint32_t w = threadIdx.x >> 6, lane = __lane_id();
    const uint32_t lr = lane & 31u;
    const float b3 = net.p[MLP_OFF_B3];
    const uint64_t ntile = (n + MLP_TILE - 1) / MLP_TILE;
    for (uint64_t t = (uint64_t)blockIdx.x * 4 + w; t < ntile; t += (uint64_t)gridDim.x * 4) {
        const uint64_t s0 = t * MLP_TILE;
        for (uint32_t q = lane; q < MLP_TILE * MLP_IN; q += 64) {
            const uint64_t smp = s0 + q / MLP_IN;
            const uint16_t b = smp < n ? f2bf(x[smp * MLP_IN + q % MLP_IN]) : (uint16_t)0;
            L.x[w][q] = b;
            if (smp < n) xt[(uint64_t)(q % MLP_IN) * n + smp] = b;
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        floatx16 acc2[4];
        mlp_tile_forward(L.x[w], L.w1t, L.w2t, L.b1, L.b2, L.h1[w], acc2);
        for (uint32_t q = lane; q < MLP_TILE * MLP_H; q += 64) {
            const uint64_t smp = s0 + q / MLP_H;
            if (smp < n) h1t[(uint64_t)(q % MLP_H) * n + smp] = L.h1[w][q];
        }
#pragma unroll
        for (uint32_t nb = 0; nb < 4; ++nb)
#pragma unroll
            for (uint32_t r = 0; r < 16; ++r) {
                const uint64_t smp = s0 + acc_row(r, lane);
                const float h = acc2[nb][r] > 0.f ? acc2[nb][r] : 0.f;
                if (smp < n) h2[smp * MLP_H + nb * 32 + lr] = h;
            }
        const float s = mlp_tile_logit(acc2, L.w3, b3, L.part[w]);
        if (lane < 32 && s0 + lane < n) {
            const uint64_t smp = s0 + lane;
            logit[smp] = s;
            ds[smp] = (1.f / (1.f + __expf(-s)) - y[smp]) * inv_total;
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
    }
}

// dz2 = ds * w3 (where h2 > 0), stored transposed (dz2t [128][n], bf16);
// dz1 = (dz2 . W2^T) masked by h1 > 0, stored transposed (dz1t [128][n]);
// gradients of w3, b3 and b2 accumulated into net.g.
__global__ void __launch_bounds__(256) k_mlp_train_bwd(MlpDev net, uint64_t n, const float* __restrict__ h2,
                                                       const float* __restrict__ ds,
                                                       const uint16_t* __restrict__ h1t,
                                                       uint16_t* __restrict__ dz2t, uint16_t* __restrict__ dz1t) {
    __shared__ alignas(16) uint16_t w2[MLP_H * MLP_H];
    __shared__ float w3[MLP_H];
    __shared__ alignas(16) uint16_t dz2[4][MLP_TILE * MLP_H];
    __shared__ float gw3[MLP_H], gb2[MLP_H], gb3;
    for (uint32_t i = threadIdx.x; i < MLP_H * MLP_H; i += blockDim.x) w2[i] = net.w2[i];
    for (uint32_t i = threadIdx.x; i < MLP_H; i += blockDim.x) {
        w3[i] = net.p[MLP_OFF_W3 + i];
        gw3[i] = 0.f;
        gb2[i] = 0.f;
    }
    if (threadIdx.x == 0) gb3 = 0.f;
    __syncthreads();
    const uint32_t w = threadIdx.x >> 6, lane = __lane_id();
    const uint32_t lr = lane & 31u, lk = 8u * (lane >> 5);
    const uint64_t ntile = (n + MLP_TILE - 1) / MLP_TILE;
    for (uint64_t t = (uint64_t)blockIdx.x * 4 + w; t < ntile; t += (uint64_t)gridDim.x * 4) {
        const uint64_t s0 = t * MLP_TILE;
        // dz2 for the tile: lane covers columns lane and lane + 64
        for (uint32_t q = lane; q < MLP_TILE * MLP_H; q += 64) {
            const uint32_t row = q / MLP_H, col = q % MLP_H;
            const uint64_t smp = s0 + row;
            float d = 0.f;
            if (smp < n) {
                const float h = h2[smp * MLP_H + col];
                const float dsv = ds[smp];
                d = h > 0.f ? dsv * w3[col] : 0.f;
                atomicAdd(&gw3[col], h * dsv);
                atomicAdd(&gb2[col], d);
                if (col == 0) atomicAdd(&gb3, dsv);
            }
            const uint16_t b = f2bf(d);
            dz2[w][q] = b;
            if (smp < n) dz2t[(uint64_t)col * n + smp] = b;
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        // dz1 = dz2 . W2^T: B[k = out][n = in] = W2[in][out]
#pragma unroll
        for (uint32_t nb = 0; nb < 4; ++nb) {
            floatx16 acc = {};
#pragma unroll
            for (uint32_t k0 = 0; k0 < MLP_H; k0 += 16)
                acc = mfma(load8(dz2[w] + lr * MLP_H + k0 + lk), load8(w2 + (nb * 32 + lr) * MLP_H + k0 + lk), acc);
            const uint32_t col = nb * 32 + lr;
#pragma unroll
            for (uint32_t r = 0; r < 16; ++r) {
                const uint64_t smp = s0 + acc_row(r, lane);
                if (smp < n) {
                    const bool on = bf2f(h1t[(uint64_t)col * n + smp]) > 0.f;
                    dz1t[(uint64_t)col * n + smp] = f2bf(on ? acc[r] : 0.f);
                }
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < MLP_H; i += blockDim.x) {
        atomicAdd(&net.g[MLP_OFF_W3 + i], gw3[i]);
        atomicAdd(&net.g[MLP_OFF_B2 + i], gb2[i]);
    }
    if (threadIdx.x == 0) atomicAdd(&net.g[MLP_OFF_B3], gb3);
}

// Weight gradients as contractions over samples, split over the grid:
// block (tile, part) computes one 32 x 32 output tile of
//   dW2[in][out] = sum_s h1[s][in] * dz2[s][out]   (16 tiles), or
//   dW1[in][out] = sum_s x[s][in]  * dz1[s][out]   (4 tiles)
// over a slice of the samples, and adds it into net.g.  One wave per block.
__global__ void __launch_bounds__(64) k_mlp_train_wgrad(MlpDev net, uint64_t n, const uint16_t* __restrict__ xt,
                                                        const uint16_t* __restrict__ h1t,
                                                        const uint16_t* __restrict__ dz2t,
                                                        const uint16_t* __restrict__ dz1t, uint32_t parts) {
    const uint32_t tile = blockIdx.x, part = blockIdx.y;
    const uint32_t lane = __lane_id();
    const uint32_t lr = lane & 31u, lk = 8u * (lane >> 5);
    const bool is_w2 = tile < 16;
    const uint32_t mb = is_w2 ? tile / 4 : 0, nb = is_w2 ? tile % 4 : tile - 16;
    const uint16_t* A = is_w2 ? h1t : xt;     // [in][n]
    const uint16_t* B = is_w2 ? dz2t : dz1t;  // [out][n]
    const uint64_t per = ((n + parts - 1) / parts + 15) / 16 * 16;
    const uint64_t k_lo = (uint64_t)part * per, k_hi = min(n, k_lo + per);
    floatx16 acc = {};
    const uint64_t arow = (uint64_t)(mb * 32 + lr) * n, brow = (uint64_t)(nb * 32 + lr) * n;
    for (uint64_t k0 = k_lo; k0 < k_hi; k0 += 16) {
        uint16_t a8[8], b8[8];
#pragma unroll
        for (uint32_t q = 0; q < 8; ++q) {
            const uint64_t k = k0 + lk + q;
            a8[q] = k < k_hi ? A[arow + k] : (uint16_t)0;
            b8[q] = k < k_hi ? B[brow + k] : (uint16_t)0;
        }
        acc = mfma(__builtin_bit_cast(bf16x8, a8), __builtin_bit_cast(bf16x8, b8), acc);
    }
    const uint32_t base = is_w2 ? MLP_OFF_W2 : MLP_OFF_W1;
#pragma unroll
    for (uint32_t r = 0; r < 16; ++r) {
        const uint32_t in = mb * 32 + acc_row(r, lane), out = nb * 32 + lr;
        atomicAdd(&net.g[base + in * MLP_H + out], acc[r]);
    }
}

// b1's gradient: sum over samples of dz1, one thread per unit.
__global__ void k_mlp_train_b1(MlpDev net, uint64_t n, const uint16_t* __restrict__ dz1t) {
    __shared__ float red[256];
    const uint32_t unit = blockIdx.x;
    float s = 0.f;
    for (uint64_t k = threadIdx.x; k < n; k += blockDim.x) s += bf2f(dz1t[(uint64_t)unit * n + k]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(&net.g[MLP_OFF_B1 + unit], red[0]);
}

// Adam (float32), then the gradient buffer is cleared for the next step.
__global__ void k_mlp_adam(MlpDev net, float lr, float b1, float b2, float eps, float c1, float c2) {
    GRID_STRIDE(i, (uint64_t)MLP_NPARAMS) {
        const float g = net.g[i];
        const float m = b1 * net.m[i] + (1.f - b1) * g;
        const float v = b2 * net.v[i] + (1.f - b2) * g * g;
        net.m[i] = m;
        net.v[i] = v;
        net.p[i] -= lr * (m * c1) / (sqrtf(v * c2) + eps);
        net.g[i] = 0.f;
    }
}

}  // namespace curc

namespace curc {
// Mean binary cross-entropy of sigmoid(logit) (for reporting): sums into *out.
__global__ void k_bce_sum(const float* __restrict__ logit, const float* __restrict__ y, uint64_t n,
                          float* __restrict__ out) {
    float s = 0.f;
    GRID_STRIDE(i, n) {
        const float z = logit[i];
        s += fmaxf(z, 0.f) + log1pf(__expf(-fabsf(z))) - y[i] * z;
    }
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
    if (__lane_id() == 0) atomicAdd(out, s);
}
}  // namespace curc
