// Per-particle MLP 8 -> 64 -> 64 -> 64 -> 1 (tanh-form GELU) on the gfx950
// matrix cores: BF16 operands, FP32 accumulation (v_mfma_f32_16x16x32_bf16).
//
// Formulation: activations are kept "transposed", units x particles.  A layer
// z = W h + b is the MFMA D = A*B with A = W (out x in, from weight fragments)
// and B = h (in x 16 particles).  The D layout of a 16x16 tile is
//   lane l = 16*g + c holds D[4g + j][c], j = 0..3,
// so two consecutive 16-row tiles of one layer's output, rounded to BF16, are
// exactly the B operand of the next layer when K slot (g, j) of the 32-deep
// product carries logical index kmap(g, j) = j < 4 ? 4g + j : 16 + 4g + j - 4.
// A and B place the same logical index in the same slot, so the product does
// not depend on the hardware's own K order.
//
// Kernels: fragment packing (FP32 master weights -> BF16 A fragments),
// forward (inference; in training also stores pre/post activations),
// backward through the activations (dh = W^T dz, same chaining with the
// transposed fragments), weight/bias gradients (MFMA with K = particles),
// and Adam.
#include "sph_common.h"
#include "mlp.h"
#include "sphgpu.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

namespace sphgpu {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int kmap(int g, int j) { return j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4); }

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }   // round to nearest even

// GELU, tanh form: 0.5 x (1 + tanh(u)), u = c (x + 0.044715 x^3),
// c = sqrt(2/pi), evaluated as x sigma(2u) = x / (1 + exp(-2u)) (the same
// function: one exponential and one hardware reciprocal per activation)
__device__ __forceinline__ float gelu_sigma(float x) {
    // exp(-2u) = exp2(t), t = x (k1 + k2 x^2), k1 = -2 c log2(e), k2 = 0.044715 k1
    const float k1 = -2.0f * 0.7978845608028654f * 1.4426950408889634f;
    const float k2 = 0.044715f * k1;
    const float t = x * (k1 + k2 * (x * x));
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(t));
}
__device__ __forceinline__ float gelu(float x) { return x * gelu_sigma(x); }
// d/dx x sigma(2u) = sigma + 2 x sigma (1 - sigma) c (1 + 3 * 0.044715 x^2)
__device__ __forceinline__ float gelu_grad(float x) {
    const float c = 0.7978845608028654f;
    const float sg = gelu_sigma(x);
    return sg + 2.0f * x * sg * (1.0f - sg) * c * (1.0f + 3.0f * 0.044715f * x * x);
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// --------------------------------------------------------------- layout ----
// parameter block (FP32 master): W1[64][8] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[1][64] b4[1]
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + 64 * 8;
constexpr int OFF_W2 = OFF_B1 + 64, OFF_B2 = OFF_W2 + 64 * 64;
constexpr int OFF_W3 = OFF_B2 + 64, OFF_B3 = OFF_W3 + 64 * 64;
constexpr int OFF_W4 = OFF_B3 + 64, OFF_B4 = OFF_W4 + 64;
static_assert(OFF_B4 + 1 == MLP_PARAMS, "parameter count");

struct LayerDim { int off_w, in, out; };
__constant__ LayerDim kLayers[4] = {{OFF_W1, 8, 64}, {OFF_W2, 64, 64}, {OFF_W3, 64, 64}, {OFF_W4, 64, 1}};

// fragment counts: forward A = W (out x in): [ceil(out/16) M tiles] x [ceil(in/32) K steps]
//   L1 4x1 (ids 0-3), L2 4x2 (4-11), L3 4x2 (12-19), L4 1x2 (20-21);
// backward A = W^T (in x out): W4^T 4x1 (0-3), W3^T 4x2 (4-11), W2^T 4x2 (12-19)
// (the first layer needs no input gradient).

// one thread per (fragment, lane): 64 lanes x 8 BF16 values
__global__ void k_mlp_pack(const float* __restrict__ P, bf16x8* __restrict__ fwd, bf16x8* __restrict__ bwd) {
    int fid = blockIdx.x, lane = threadIdx.x;
    int g = lane >> 4, c = lane & 15;
    bf16x8 v;
    if (fid < MLP_FWD_FRAGS) {
        int layer = fid < 4 ? 0 : fid < 12 ? 1 : fid < 20 ? 2 : 3;
        const int base[4] = {0, 4, 12, 20};
        const int ks[4] = {1, 2, 2, 2};
        int rel = fid - base[layer], t = rel / ks[layer], s = rel % ks[layer];
        LayerDim L = kLayers[layer];
        int out = 16 * t + c;
        for (int j = 0; j < 8; ++j) {
            int in = 32 * s + kmap(g, j);
            float w = (out < L.out && in < L.in) ? P[L.off_w + out * L.in + in] : 0.0f;
            v[j] = to_bf16(w);
        }
        fwd[fid * 64 + lane] = v;
    } else {
        int b = fid - MLP_FWD_FRAGS;
        // backward fragments: layer 3 (W4^T: 64 x 1, 4 tiles x 1 kstep) -> 0..3,
        // layer 2 (W3^T: 4 x 2) -> 4..11, layer 1 (W2^T: 4 x 2) -> 12..19
        int layer = b < 4 ? 3 : b < 12 ? 2 : 1;
        int base = b < 4 ? 0 : b < 12 ? 4 : 12;
        int ks = b < 4 ? 1 : 2;
        int rel = b - base, t = rel / ks, s = rel % ks;
        LayerDim L = kLayers[layer];   // W is out x in; A = W^T is in x out
        int in = 16 * t + c;
        for (int j = 0; j < 8; ++j) {
            int out = 32 * s + kmap(g, j);
            float w = (out < L.out && in < L.in) ? P[L.off_w + out * L.in + in] : 0.0f;
            v[j] = to_bf16(w);
        }
        bwd[b * 64 + lane] = v;
    }
}

// B operand for K step s from two D tiles (t = 2s, 2s+1) of activations
__device__ __forceinline__ bf16x8 chain(const float (&h)[4][4], int s) {
    bf16x8 b;
    for (int j = 0; j < 4; ++j) {
        b[j] = to_bf16(h[2 * s][j]);
        b[j + 4] = to_bf16(h[2 * s + 1][j]);
    }
    return b;
}

// forward of 16 particles [p0, p0+16); optionally stores z/h rows for training
struct FwdOut {
    float z[3][4][4];   // pre-activations (layer, tile, j)
    float y;
};

__device__ __forceinline__ void forward_tile(const bf16x8* __restrict__ F, const float* __restrict__ P,
                                             const float* __restrict__ feat, int p0, int n, FwdOut& o,
                                             float* __restrict__ store) {
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    const int p = p0 + c;
    bf16x8 x;
    for (int j = 0; j < 8; ++j) x[j] = to_bf16(0.0f);
    if (g < 2 && p < n) {
        const float4 f = reinterpret_cast<const float4*>(feat)[p * 2 + g];
        x[0] = to_bf16(f.x); x[1] = to_bf16(f.y); x[2] = to_bf16(f.z); x[3] = to_bf16(f.w);
    }
    float h[4][4];
    // layer 1
    for (int t = 0; t < 4; ++t) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(F[(0 + t) * 64 + lane], x, acc);
        for (int j = 0; j < 4; ++j) {
            int u = 16 * t + 4 * g + j;
            float z = acc[j] + P[OFF_B1 + u];
            o.z[0][t][j] = z;
            h[t][j] = gelu(z);
        }
    }
    // layers 2, 3
    const int fbase[2] = {4, 12}, boff[2] = {OFF_B2, OFF_B3};
    for (int L = 0; L < 2; ++L) {
        bf16x8 b0 = chain(h, 0), b1 = chain(h, 1);
        for (int t = 0; t < 4; ++t) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(F[(fbase[L] + 2 * t) * 64 + lane], b0, acc);
            acc = mfma(F[(fbase[L] + 2 * t + 1) * 64 + lane], b1, acc);
            for (int j = 0; j < 4; ++j) {
                int u = 16 * t + 4 * g + j;
                float z = acc[j] + P[boff[L] + u];
                o.z[L + 1][t][j] = z;
                h[t][j] = gelu(z);
            }
        }
    }
    // layer 4 (1 output row)
    {
        bf16x8 b0 = chain(h, 0), b1 = chain(h, 1);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(F[20 * 64 + lane], b0, acc);
        acc = mfma(F[21 * 64 + lane], b1, acc);
        o.y = acc[0] + P[OFF_B4];       // row 0: lanes g == 0, j == 0
    }
    if (store && p < n) {
        // rows [p][layer*64 + unit]: z1 z2 z3, then y at 192
        for (int L = 0; L < 3; ++L)
            for (int t = 0; t < 4; ++t)
                for (int j = 0; j < 4; ++j) store[(size_t)p * MLP_STORE + L * 64 + 16 * t + 4 * g + j] = o.z[L][t][j];
        if (g == 0) store[(size_t)p * MLP_STORE + 192] = o.y;
    }
}

__global__ __launch_bounds__(256) void k_mlp_forward(const bf16x8* __restrict__ F, const float* __restrict__ P,
                                                   const float* __restrict__ feat, float* __restrict__ y, int n,
                                                   float* __restrict__ store) {
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int ntile = (n + 15) / 16;
    for (int tile = wave; tile < ntile; tile += nw) {
        FwdOut o;
        forward_tile(F, P, feat, tile * 16, n, o, store);
        int p = tile * 16 + (lane & 15);
        if ((lane >> 4) == 0 && p < n) y[p] = o.y;
    }
}

// Inference only: the 22 weight fragments stay in registers and the biases
// in LDS for all the tiles of 16 particles a wave processes (the training
// forward above re-reads them per tile).  gate1: a tile whose 16 particles
// all have feature 1 below zero is skipped and its outputs are not written
// (the DFSPH warm start reads the output only where Drho/Dt > 0, and feature
// 1 is asinh of a positive multiple of Drho/Dt, negative only where it is).
__global__ __launch_bounds__(256, 3) void k_mlp_infer(const bf16x8* __restrict__ F, const float* __restrict__ P,
                                                 const float* __restrict__ feat, float* __restrict__ y, int n,
                                                 int gate1) {
    __shared__ float sb[3 * 64];
    for (int k = threadIdx.x; k < 3 * 64; k += 256) {
        const int off[3] = {OFF_B1, OFF_B2, OFF_B3};
        sb[k] = P[off[k >> 6] + (k & 63)];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    bf16x8 w[MLP_FWD_FRAGS];
#pragma unroll
    for (int f = 0; f < MLP_FWD_FRAGS; ++f) w[f] = F[f * 64 + lane];
    const float b4 = P[OFF_B4];
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    const int ntile = (n + 15) / 16;
    for (int tile = wave; tile < ntile; tile += nw) {
        const int p = tile * 16 + c;
        bf16x8 x;
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = to_bf16(0.0f);
        bool need = false;
        if (g < 2 && p < n) {
            const float4 f = reinterpret_cast<const float4*>(feat)[p * 2 + g];
            x[0] = to_bf16(f.x); x[1] = to_bf16(f.y); x[2] = to_bf16(f.z); x[3] = to_bf16(f.w);
            need = g == 0 && !(f.y < 0.0f);
        }
        if (gate1 && __ballot(need) == 0) continue;   // uniform per wave
        float h[4][4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(w[t], x, acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) h[t][j] = gelu(acc[j] + sb[16 * t + 4 * g + j]);
        }
#pragma unroll
        for (int L = 0; L < 2; ++L) {
            const bf16x8 b0 = chain(h, 0), b1 = chain(h, 1);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = mfma(w[4 + 8 * L + 2 * t], b0, acc);
                acc = mfma(w[4 + 8 * L + 2 * t + 1], b1, acc);
#pragma unroll
                for (int j = 0; j < 4; ++j) h[t][j] = gelu(acc[j] + sb[64 * (L + 1) + 16 * t + 4 * g + j]);
            }
        }
        const bf16x8 b0 = chain(h, 0), b1 = chain(h, 1);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(w[20], b0, acc);
        acc = mfma(w[21], b1, acc);
        if (g == 0 && p < n) y[p] = acc[0] + b4;
    }
}

// backward through the activations for 16 particles: dL/dy -> dz3, dz2, dz1,
// stored as rows [p][layer*64 + unit] in dz (FP32).  Pre-activations come
// from the forward store.
__global__ __launch_bounds__(256) void k_mlp_backward(const bf16x8* __restrict__ B, const float* __restrict__ store,
                                                    const float* __restrict__ dy, float* __restrict__ dz, int n) {
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    const int ntile = (n + 15) / 16;
    for (int tile = wave; tile < ntile; tile += nw) {
        const int p = tile * 16 + c;
        const bool ok = p < n;
        // B operand: dz4 (1 x P) in K slot (0, 0)
        bf16x8 b;
        for (int j = 0; j < 8; ++j) b[j] = to_bf16(0.0f);
        if (g == 0 && ok) b[0] = to_bf16(dy[p]);
        float d[4][4];
        // dh3 = W4^T dz4 ; dz3 = dh3 * gelu'(z3)
        for (int t = 0; t < 4; ++t) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(B[(0 + t) * 64 + lane], b, acc);
            for (int j = 0; j < 4; ++j) {
                int u = 16 * t + 4 * g + j;
                float z = ok ? store[(size_t)p * MLP_STORE + 2 * 64 + u] : 0.0f;
                d[t][j] = acc[j] * gelu_grad(z);
                if (ok) dz[(size_t)p * MLP_STORE + 2 * 64 + u] = d[t][j];
            }
        }
        // dh2 = W3^T dz3 -> dz2 ; dh1 = W2^T dz2 -> dz1
        const int fb[2] = {4, 12}, zl[2] = {1, 0};
        for (int L = 0; L < 2; ++L) {
            bf16x8 b0 = chain(d, 0), b1 = chain(d, 1);
            float e[4][4];
            for (int t = 0; t < 4; ++t) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = mfma(B[(fb[L] + 2 * t) * 64 + lane], b0, acc);
                acc = mfma(B[(fb[L] + 2 * t + 1) * 64 + lane], b1, acc);
                for (int j = 0; j < 4; ++j) {
                    int u = 16 * t + 4 * g + j;
                    float z = ok ? store[(size_t)p * MLP_STORE + zl[L] * 64 + u] : 0.0f;
                    e[t][j] = acc[j] * gelu_grad(z);
                    if (ok) dz[(size_t)p * MLP_STORE + zl[L] * 64 + u] = e[t][j];
                }
            }
            for (int t = 0; t < 4; ++t)
                for (int j = 0; j < 4; ++j) d[t][j] = e[t][j];
        }
        if (g == 0 && ok) dz[(size_t)p * MLP_STORE + 192] = dy[p];
    }
}

// Tile t of the weight/bias gradient: 16x16 blocks of dW (out x in) per
// layer, then bias blocks.  L1 4x1, L2 4x4, L3 4x4, L4 1x4 (W), then bias
// tiles 4+4+4+1.
struct GradTile { int L, to, ti, bias; };
__device__ __forceinline__ GradTile grad_tile(int t) {
    GradTile r{0, 0, 0, 0};
    int k = t;
    const int wt[4] = {4, 16, 16, 4};
    while (r.L < 4 && k >= wt[r.L]) { k -= wt[r.L]; ++r.L; }
    if (r.L < 4) {
        int nin = (r.L == 0) ? 1 : 4;
        r.to = k / nin;
        r.ti = k % nin;
    } else {
        r.bias = 1;
        const int bt[4] = {4, 4, 4, 1};
        r.L = 0;
        while (r.L < 3 && k >= bt[r.L]) { k -= bt[r.L]; ++r.L; }
        r.to = k;
    }
    return r;
}

// Weight and bias gradients, K = samples split into slices of
// MLP_WGRAD_SLICE: one block per slice.  For each chunk of 32 samples the
// block stages in LDS, once, the chunk's pre-activation gradients dz (the A
// operand: output unit x sample) and its layer inputs h = features,
// GELU(z1), GELU(z2), GELU(z3) (the B operand: sample x input unit), both as
// BF16 and unit-major so that a fragment's eight samples are two 8-byte LDS
// reads: thread t loads column t of the chunk's rows (coalesced across the
// wave), applies GELU once and writes its unit's 32 samples with 8-byte
// stores.  The next chunk's loads are issued before this chunk's tiles run;
// the four waves accumulate the 53 tiles (each wave every fourth) on the
// matrix cores.  k_mlp_wgrad_reduce sums the slices in a fixed order
// (deterministic).
constexpr int WG_CHUNK = 32;
constexpr int WG_ROW = WG_CHUNK + 4;     // LDS row pitch in BF16 (8-byte aligned, staggered banks)
constexpr int WG_DZ = 193;               // dz rows: z1 z2 z3 (192) and y
constexpr int WG_H = 200;                // input rows: features (8), h1, h2, h3
constexpr int WG_TILES_PER_WAVE = (MLP_WGRAD_TILES + 3) / 4;

__device__ __forceinline__ bf16x8 lds_frag(const __bf16* row, int g) {
    // samples kmap(g, 0..3) = 4g..4g+3 and kmap(g, 4..7) = 16+4g..16+4g+3
    const uint2 lo = *reinterpret_cast<const uint2*>(row + 4 * g);
    const uint2 hi = *reinterpret_cast<const uint2*>(row + 16 + 4 * g);
    bf16x8 v;
    __builtin_memcpy(&v, &lo, 8);
    __builtin_memcpy(reinterpret_cast<char*>(&v) + 8, &hi, 8);
    return v;
}

// One chunk's operands in registers: thread t holds column t of the chunk's
// 32 store rows (z1 z2 z3) and 32 dz rows, and feature (t & 7) of sample
// t >> 3.  Each load instruction reads 64 consecutive floats of one row.
struct WgCols {
    float z[WG_CHUNK];
    float d[WG_CHUNK];
    float f;
};

__device__ __forceinline__ void wg_load(const float* __restrict__ feat, const float* __restrict__ store,
                                        const float* __restrict__ dz, int q, int p1, int tid, WgCols& C) {
#pragma unroll
    for (int s = 0; s < WG_CHUNK; ++s) {
        const size_t row = (size_t)(q + s) * MLP_STORE + tid;
        const bool ok = q + s < p1;
        C.z[s] = ok && tid < 192 ? store[row] : 0.0f;
        C.d[s] = ok && tid < WG_DZ ? dz[row] : 0.0f;
    }
    const int p = q + (tid >> 3);
    C.f = p < p1 ? feat[(size_t)p * 8 + (tid & 7)] : 0.0f;
}

// four BF16 of a unit-major LDS row (one 8-byte store)
__device__ __forceinline__ void st_bf16x4(__bf16* dst, float a, float b, float c, float d) {
    const __bf16 v[4] = {to_bf16(a), to_bf16(b), to_bf16(c), to_bf16(d)};
    uint2 u;
    __builtin_memcpy(&u, v, 8);
    *reinterpret_cast<uint2*>(dst) = u;
}

__device__ __forceinline__ void wg_store(const WgCols& C, int q, int p1, int tid, __bf16* sA, __bf16* sB) {
    if (tid < 192) {
        __bf16* row = sB + (8 + tid) * WG_ROW;
#pragma unroll
        for (int s = 0; s < WG_CHUNK; s += 4)
            st_bf16x4(row + s, gelu(C.z[s]), gelu(C.z[s + 1]), gelu(C.z[s + 2]), gelu(C.z[s + 3]));
    }
    if (tid < WG_DZ) {
        __bf16* row = sA + tid * WG_ROW;
#pragma unroll
        for (int s = 0; s < WG_CHUNK; s += 4) st_bf16x4(row + s, C.d[s], C.d[s + 1], C.d[s + 2], C.d[s + 3]);
    }
    sB[(tid & 7) * WG_ROW + (tid >> 3)] = to_bf16(C.f);
    if (tid < WG_CHUNK) sB[WG_H * WG_ROW + tid] = to_bf16(q + tid < p1 ? 1.0f : 0.0f);
}

__global__ __launch_bounds__(256) void k_mlp_wgrad(const float* __restrict__ feat, const float* __restrict__ store,
                                                 const float* __restrict__ dz, float* __restrict__ part, int n) {
    __shared__ __attribute__((aligned(16))) __bf16 sA[WG_DZ * WG_ROW];
    __shared__ __attribute__((aligned(16))) __bf16 sB[(WG_H + 1) * WG_ROW];   // last row: bias input 1 / 0
    const int slice = blockIdx.x, tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63, g = lane >> 4, c = lane & 15;
    f32x4 acc[WG_TILES_PER_WAVE];
#pragma unroll
    for (int k = 0; k < WG_TILES_PER_WAVE; ++k) acc[k] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int p0 = slice * MLP_WGRAD_SLICE, p1 = min(n, p0 + MLP_WGRAD_SLICE);
    WgCols C;
    wg_load(feat, store, dz, p0, p1, tid, C);
    for (int q = p0; q < p1; q += WG_CHUNK) {
        __syncthreads();   // the previous chunk's fragments are read
        wg_store(C, q, p1, tid, sA, sB);
        __syncthreads();
        // the next chunk's loads are in flight while this one's tiles run
        if (q + WG_CHUNK < p1) wg_load(feat, store, dz, q + WG_CHUNK, p1, tid, C);
#pragma unroll
        for (int k = 0; k < WG_TILES_PER_WAVE; ++k) {
            const int t = wave + 4 * k;
            if (t >= MLP_WGRAD_TILES) break;
            const GradTile T = grad_tile(t);
            const int nin = (T.L == 0) ? 8 : 64, nout = (T.L == 3) ? 1 : 64;
            const int zoff = (T.L < 3) ? T.L * 64 : 192;            // dz rows of layer L's output
            const int hoff = (T.L == 0) ? 0 : 8 + (T.L - 1) * 64;   // rows of layer L's input
            const int orow = 16 * T.to + c, icol = 16 * T.ti + c;
            const bf16x8 zero = bf16x8{};
            const bf16x8 a = orow < nout ? lds_frag(sA + (zoff + orow) * WG_ROW, g) : zero;
            const bf16x8 b = T.bias ? lds_frag(sB + WG_H * WG_ROW, g)
                                    : (icol < nin ? lds_frag(sB + (hoff + icol) * WG_ROW, g) : zero);
            acc[k] = mfma(a, b, acc[k]);
        }
    }
    // D[4g+j][c] of each tile
#pragma unroll
    for (int k = 0; k < WG_TILES_PER_WAVE; ++k) {
        const int t = wave + 4 * k;
        if (t >= MLP_WGRAD_TILES) break;
        float* out = part + ((size_t)slice * MLP_WGRAD_TILES + t) * 256;
        for (int j = 0; j < 4; ++j) out[(4 * g + j) * 16 + c] = acc[k][j];
    }
}

// Sum of the slices' partial tiles: eight threads per gradient element,
// thread group r summing slices r, r + 8, r + 16, ... in order, then the
// eight sums added in group order (a fixed order: deterministic).  Block b
// covers 32 elements of tile b / 8.
__global__ __launch_bounds__(256) void k_mlp_wgrad_reduce(const float* __restrict__ part, float* __restrict__ grad,
                                                        int slices, float scale) {
    __shared__ float red[8][32];
    const int tile = blockIdx.x >> 3, r = threadIdx.x >> 5, lane = threadIdx.x & 31;
    const int e = ((blockIdx.x & 7) << 5) | lane;
    float s = 0.0f;
#pragma unroll 4
    for (int k = r; k < slices; k += 8) s += part[((size_t)k * MLP_WGRAD_TILES + tile) * 256 + e];
    red[r][lane] = s;
    __syncthreads();
    if (r != 0) return;
#pragma unroll
    for (int k = 1; k < 8; ++k) s += red[k][lane];
    const GradTile T = grad_tile(tile);
    const int nin = (T.L == 0) ? 8 : 64, nout = (T.L == 3) ? 1 : 64;
    const int row = e >> 4, col = e & 15;
    const int out = 16 * T.to + row, in = 16 * T.ti + col;
    if (out >= nout) return;
    if (T.bias ? col != 0 : in >= nin) return;
    const int woff[4] = {OFF_W1, OFF_W2, OFF_W3, OFF_W4}, boff[4] = {OFF_B1, OFF_B2, OFF_B3, OFF_B4};
    if (T.bias) grad[boff[T.L] + out] = s * scale;
    else grad[woff[T.L] + out * nin + in] = s * scale;
}

// MSE loss gradient dL/dy = 2 (y - t) / n (times `scale` for normalised targets)
// loss = mean(w e^2), e = y - target, w = over where e > 0 else 1
__global__ void k_mlp_loss(const float* __restrict__ store, const float* __restrict__ target,
                           float* __restrict__ dy, double* __restrict__ lossPart, int n, float over) {
    __shared__ double lds[4];
    int i = blockIdx.x * 256 + threadIdx.x;
    double l = 0.0;
    if (i < n) {
        float y = store[(size_t)i * MLP_STORE + 192];
        float e = y - target[i];
        float w = e > 0.0f ? over : 1.0f;
        dy[i] = 2.0f * w * e / (float)n;
        l = (double)w * (double)e * (double)e;
    }
    double s = block_sum_256(l, lds);
    if (threadIdx.x == 0) lossPart[blockIdx.x] = s;
}

__global__ void k_mlp_adam(float* __restrict__ P, float* __restrict__ M, float* __restrict__ V,
                           const float* __restrict__ grad, float lr, float b1, float b2, float eps, float c1,
                           float c2) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= MLP_PARAMS) return;
    float gr = grad[i];
    float m = b1 * M[i] + (1.0f - b1) * gr;
    float v = b2 * V[i] + (1.0f - b2) * gr * gr;
    M[i] = m;
    V[i] = v;
    P[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

}  // namespace

// ------------------------------------------------------------------ host ----
void mlp_pack(MlpDev& m, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(MLP_FWD_FRAGS + MLP_BWD_FRAGS), dim3(64), 0, st, (const float*)m.params,
                       (bf16x8*)m.fwd, (bf16x8*)m.bwd);
}

void mlp_forward(const MlpDev& m, const float* feat, float* y, int n, float* store, hipStream_t st, bool gate1) {
    if (n <= 0) return;
    int waves = (n + 15) / 16;
    if (!store) {
        // about three waves per SIMD, each looping over tiles
        const int blocks = std::min((waves + 3) / 4, 768);
        hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, st, (const bf16x8*)m.fwd, (const float*)m.params,
                           feat, y, n, gate1 ? 1 : 0);
        return;
    }
    int blocks = std::min((waves + 3) / 4, 8192);
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(256), 0, st, (const bf16x8*)m.fwd,
                       (const float*)m.params, feat, y, n, store);
}

void mlp_grad(MlpDev& m, const float* feat, const float* target, int n, hipStream_t st) {
    mlp_forward(m, feat, m.scratchY, n, m.store, st);
    int lb = (n + 255) / 256;
    hipLaunchKernelGGL(k_mlp_loss, dim3(lb), dim3(256), 0, st, (const float*)m.store, target, m.dy, m.lossPart, n,
                       m.overWeight);
    int waves = (n + 15) / 16;
    int blocks = std::min((waves + 3) / 4, 8192);
    hipLaunchKernelGGL(k_mlp_backward, dim3(blocks), dim3(256), 0, st, (const bf16x8*)m.bwd,
                       (const float*)m.store, (const float*)m.dy, m.dz, n);
    const int slices = (n + MLP_WGRAD_SLICE - 1) / MLP_WGRAD_SLICE;
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(slices), dim3(256), 0, st, feat, (const float*)m.store,
                       (const float*)m.dz, m.gradPart, n);
    hipLaunchKernelGGL(k_mlp_wgrad_reduce, dim3(8 * MLP_WGRAD_TILES), dim3(256), 0, st, (const float*)m.gradPart,
                       m.grad, slices, 1.0f);
}

void mlp_adam(MlpDev& m, float lr, hipStream_t st) {
    m.adamStep += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float c1 = 1.0f - std::pow(b1, (float)m.adamStep), c2 = 1.0f - std::pow(b2, (float)m.adamStep);
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_PARAMS + 255) / 256), dim3(256), 0, st, m.params, m.adamM, m.adamV,
                       (const float*)m.grad, lr, b1, b2, eps, c1, c2);
    mlp_pack(m, st);
}

// Glorot-uniform weights, zero biases, from a 64-bit LCG stream
void mlp_init_params(std::vector<float>& p, unsigned long long seed) {
    p.assign(MLP_PARAMS, 0.0f);
    unsigned long long s = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    auto uni = [&]() {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        return double(s >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    const int off[4] = {OFF_W1, OFF_W2, OFF_W3, OFF_W4}, in[4] = {8, 64, 64, 64}, out[4] = {64, 64, 64, 1};
    for (int L = 0; L < 4; ++L) {
        double a = std::sqrt(6.0 / double(in[L] + out[L]));
        for (int k = 0; k < in[L] * out[L]; ++k) p[off[L] + k] = float(a * uni());
    }
}

int mlp_alloc(MlpDev& m, int cap) {
    m.cap = std::max(cap, 16);
    auto al = [](void** p, size_t b) { return hipMalloc(p, std::max<size_t>(b, 256)) == hipSuccess; };
    bool ok = al((void**)&m.params, sizeof(float) * MLP_PARAMS) && al((void**)&m.adamM, sizeof(float) * MLP_PARAMS) &&
              al((void**)&m.adamV, sizeof(float) * MLP_PARAMS) && al((void**)&m.grad, sizeof(float) * MLP_PARAMS) &&
              al(&m.fwd, 16 * 64 * MLP_FWD_FRAGS) && al(&m.bwd, 16 * 64 * MLP_BWD_FRAGS) &&
              al((void**)&m.store, sizeof(float) * MLP_STORE * (size_t)m.cap) &&
              al((void**)&m.dz, sizeof(float) * MLP_STORE * (size_t)m.cap) &&
              al((void**)&m.dy, sizeof(float) * m.cap) && al((void**)&m.scratchY, sizeof(float) * m.cap) &&
              al((void**)&m.lossPart, sizeof(double) * (m.cap / 256 + 2)) &&
              al((void**)&m.gradPart, sizeof(float) * 256 * MLP_WGRAD_TILES *
                                          (size_t)((m.cap + MLP_WGRAD_SLICE - 1) / MLP_WGRAD_SLICE));
    if (!ok) return -1;
    hipMemset(m.adamM, 0, sizeof(float) * MLP_PARAMS);
    hipMemset(m.adamV, 0, sizeof(float) * MLP_PARAMS);
    hipMemset(m.dz, 0, sizeof(float) * MLP_STORE * (size_t)m.cap);
    m.adamStep = 0;
    return 0;
}

void mlp_free(MlpDev& m) {
    void* ps[] = {m.params, m.adamM, m.adamV, m.grad, m.fwd, m.bwd, m.store, m.dz, m.dy, m.scratchY, m.lossPart,
                  m.gradPart};
    for (void* p : ps)
        if (p) hipFree(p);
    m = MlpDev{};
}

double mlp_loss_host(const MlpDev& m, int n, hipStream_t st) {
    int lb = (n + 255) / 256;
    std::vector<double> part(lb);
    hipMemcpyAsync(part.data(), m.lossPart, sizeof(double) * lb, hipMemcpyDeviceToHost, st);
    hipStreamSynchronize(st);
    double s = 0.0;
    for (double v : part) s += v;
    return n > 0 ? s / n : 0.0;
}

// ---- gradient on caller-provided device buffers, scratch in a static array
// Every buffer mlp_grad needs beyond its inputs and outputs (weight
// fragments, activations, their gradients, loss and gradient partials) lives
// in one file-scope device array, so the entry point allocates nothing.
// Each GPU holds its own copy until the process ends; host code reaches it
// only through the address hipGetSymbolAddress returns on that device.
constexpr int MLP_DEV_MAX = SPHGPU_MLP_DEVICE_MAX;
constexpr size_t al256(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t SCR_FWD = 0;
constexpr size_t SCR_BWD = SCR_FWD + al256(16 * 64 * MLP_FWD_FRAGS);
constexpr size_t SCR_STORE = SCR_BWD + al256(16 * 64 * MLP_BWD_FRAGS);
constexpr size_t SCR_DZ = SCR_STORE + al256(sizeof(float) * MLP_STORE * (size_t)MLP_DEV_MAX);
constexpr size_t SCR_DY = SCR_DZ + al256(sizeof(float) * MLP_STORE * (size_t)MLP_DEV_MAX);
constexpr size_t SCR_Y = SCR_DY + al256(sizeof(float) * MLP_DEV_MAX);
constexpr size_t SCR_LOSS = SCR_Y + al256(sizeof(float) * MLP_DEV_MAX);
constexpr size_t SCR_GPART = SCR_LOSS + al256(sizeof(double) * (MLP_DEV_MAX / 256 + 2));
constexpr size_t SCR_BYTES =
    SCR_GPART + al256(sizeof(float) * 256 * MLP_WGRAD_TILES * ((MLP_DEV_MAX + MLP_WGRAD_SLICE - 1) / MLP_WGRAD_SLICE));

static __device__ __align__(256) unsigned char scratch[SCR_BYTES];

namespace {
// mean loss: the block partials summed in order by one thread (the order
// mlp_loss_host sums them in), divided by n
__global__ void k_mlp_loss_final(const double* __restrict__ part, int parts, int n, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int k = 0; k < parts; ++k) s += part[k];
    *out = s / n;
}
}  // namespace

// device address of `scratch` on the current device, looked up once per device
static int scratch_on_device(unsigned char** out) {
    static std::mutex mu;
    static unsigned char* ptr[64] = {};
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return SPHGPU_E_DEVICE;
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SPHGPU_E_DEVICE;
        ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = ptr[dev];
    return SPHGPU_OK;
}

}  // namespace sphgpu

// ------------------------------------------------------------- C-ABI ----
struct sphgpu_mlp {
    sphgpu::MlpDev m;
    hipStream_t st = nullptr;
    float* feat = nullptr;
    float* target = nullptr;
    float* y = nullptr;
};

extern "C" int sphgpu_mlp_create(int capacity, unsigned long long seed, sphgpu_mlp** out) {
    if (!out || capacity <= 0) return SPHGPU_E_ARG;
    *out = nullptr;
    if (!sphgpu_device_available()) return SPHGPU_E_DEVICE;
    sphgpu_mlp* h = new sphgpu_mlp();
    if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess || sphgpu::mlp_alloc(h->m, capacity) ||
        hipMalloc((void**)&h->feat, sizeof(float) * 8 * (size_t)h->m.cap) != hipSuccess ||
        hipMalloc((void**)&h->target, sizeof(float) * (size_t)h->m.cap) != hipSuccess ||
        hipMalloc((void**)&h->y, sizeof(float) * (size_t)h->m.cap) != hipSuccess) {
        sphgpu_mlp_destroy(h);
        return SPHGPU_E_DEVICE;
    }
    std::vector<float> p;
    sphgpu::mlp_init_params(p, seed);
    return sphgpu_mlp_set_params(h, p.data()) == 0 ? (*out = h, SPHGPU_OK) : (sphgpu_mlp_destroy(h), SPHGPU_E_DEVICE);
}

extern "C" void sphgpu_mlp_destroy(sphgpu_mlp* h) {
    if (!h) return;
    if (h->st) hipStreamSynchronize(h->st);
    sphgpu::mlp_free(h->m);
    if (h->feat) hipFree(h->feat);
    if (h->target) hipFree(h->target);
    if (h->y) hipFree(h->y);
    if (h->st) hipStreamDestroy(h->st);
    delete h;
}

extern "C" int sphgpu_mlp_set_params(sphgpu_mlp* h, const float* params) {
    if (!h || !params) return SPHGPU_E_ARG;
    if (hipMemcpyAsync(h->m.params, params, sizeof(float) * sphgpu::MLP_PARAMS, hipMemcpyHostToDevice, h->st) != hipSuccess)
        return SPHGPU_E_DEVICE;
    sphgpu::mlp_pack(h->m, h->st);
    return hipStreamSynchronize(h->st) == hipSuccess ? SPHGPU_OK : SPHGPU_E_DEVICE;
}

extern "C" int sphgpu_mlp_get_params(sphgpu_mlp* h, float* params) {
    if (!h || !params) return SPHGPU_E_ARG;
    hipMemcpyAsync(params, h->m.params, sizeof(float) * sphgpu::MLP_PARAMS, hipMemcpyDeviceToHost, h->st);
    return hipStreamSynchronize(h->st) == hipSuccess ? SPHGPU_OK : SPHGPU_E_DEVICE;
}

extern "C" int sphgpu_mlp_set_loss(sphgpu_mlp* h, float over_weight) {
    if (!h || !(over_weight > 0.0f) || !std::isfinite(over_weight)) return SPHGPU_E_ARG;
    h->m.overWeight = over_weight;
    return SPHGPU_OK;
}

extern "C" int sphgpu_mlp_forward(sphgpu_mlp* h, long long n, const float* feat, float* out) {
    if (!h || n < 0 || (n > 0 && (!feat || !out))) return SPHGPU_E_ARG;
    for (long long o = 0; o < n; o += h->m.cap) {
        int k = int(std::min<long long>(h->m.cap, n - o));
        hipMemcpyAsync(h->feat, feat + 8 * o, sizeof(float) * 8 * k, hipMemcpyHostToDevice, h->st);
        sphgpu::mlp_forward(h->m, h->feat, h->y, k, nullptr, h->st);
        hipMemcpyAsync(out + o, h->y, sizeof(float) * k, hipMemcpyDeviceToHost, h->st);
    }
    return hipStreamSynchronize(h->st) == hipSuccess && hipGetLastError() == hipSuccess ? SPHGPU_OK : SPHGPU_E_DEVICE;
}

extern "C" int sphgpu_mlp_gradient(sphgpu_mlp* h, long long n, const float* feat, const float* target, float* grad,
                                   double* loss) {
    if (!h || n <= 0 || n > h->m.cap || !feat || !target || !grad) return SPHGPU_E_ARG;
    int k = int(n);
    hipMemcpyAsync(h->feat, feat, sizeof(float) * 8 * k, hipMemcpyHostToDevice, h->st);
    hipMemcpyAsync(h->target, target, sizeof(float) * k, hipMemcpyHostToDevice, h->st);
    sphgpu::mlp_grad(h->m, h->feat, h->target, k, h->st);
    hipMemcpyAsync(grad, h->m.grad, sizeof(float) * sphgpu::MLP_PARAMS, hipMemcpyDeviceToHost, h->st);
    double l = sphgpu::mlp_loss_host(h->m, k, h->st);
    if (loss) *loss = l;
    return hipGetLastError() == hipSuccess ? SPHGPU_OK : SPHGPU_E_DEVICE;
}

extern "C" int sphgpu_mlp_train(sphgpu_mlp* h, long long n, const float* feat, const float* target, int iters, float lr,
                                double* loss) {
    if (!h || n <= 0 || n > h->m.cap || !feat || !target || iters < 0) return SPHGPU_E_ARG;
    int k = int(n);
    hipMemcpyAsync(h->feat, feat, sizeof(float) * 8 * k, hipMemcpyHostToDevice, h->st);
    hipMemcpyAsync(h->target, target, sizeof(float) * k, hipMemcpyHostToDevice, h->st);
    for (int it = 0; it < iters; ++it) {
        sphgpu::mlp_grad(h->m, h->feat, h->target, k, h->st);
        sphgpu::mlp_adam(h->m, lr, h->st);
    }
    sphgpu::mlp_grad(h->m, h->feat, h->target, k, h->st);
    double l = sphgpu::mlp_loss_host(h->m, k, h->st);
    if (loss) *loss = l;
    return hipGetLastError() == hipSuccess ? SPHGPU_OK : SPHGPU_E_DEVICE;
}

extern "C" int sphgpu_mlp_gradient_device(const float* params, long long n, const float* feat, const float* target,
                                          float over_weight, float* grad, double* loss, void* stream) {
    using namespace sphgpu;
    if (!params || !feat || !target || !grad || !loss || n <= 0 || !(over_weight > 0.0f) ||
        !std::isfinite(over_weight))
        return SPHGPU_E_ARG;
    if (n > MLP_DEV_MAX) return SPHGPU_E_CAPACITY;
    unsigned char* base = nullptr;
    if (int rc = scratch_on_device(&base)) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    MlpDev m;
    m.params = const_cast<float*>(params);   // read only: no Adam step here
    m.grad = grad;
    m.fwd = base + SCR_FWD;
    m.bwd = base + SCR_BWD;
    m.store = reinterpret_cast<float*>(base + SCR_STORE);
    m.dz = reinterpret_cast<float*>(base + SCR_DZ);
    m.dy = reinterpret_cast<float*>(base + SCR_DY);
    m.scratchY = reinterpret_cast<float*>(base + SCR_Y);
    m.lossPart = reinterpret_cast<double*>(base + SCR_LOSS);
    m.gradPart = reinterpret_cast<float*>(base + SCR_GPART);
    m.overWeight = over_weight;
    m.cap = MLP_DEV_MAX;
    const int k = int(n);
    // every scratch word the kernels read is written earlier in this call:
    // fragments by the pack, activations and gradients for samples < n, loss
    // partials per block, gradient partials per slice
    mlp_pack(m, st);
    mlp_grad(m, feat, target, k, st);
    hipLaunchKernelGGL(k_mlp_loss_final, dim3(1), dim3(64), 0, st, (const double*)m.lossPart, (k + 255) / 256, k,
                       loss);
    return hipGetLastError() == hipSuccess ? SPHGPU_OK : SPHGPU_E_DEVICE;
}
