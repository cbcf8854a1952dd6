// Device state of the per-particle MLP 8 -> 64 -> 64 -> 64 -> 1 (mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace sphgpu {

constexpr int MLP_IN = 8;
constexpr int MLP_PARAMS = 64 * 8 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1;   // 8961
constexpr int MLP_FWD_FRAGS = 22;    // BF16 A fragments of the forward pass
constexpr int MLP_BWD_FRAGS = 20;    // BF16 A fragments of the backward pass
constexpr int MLP_STORE = 196;       // per-sample row: z1 z2 z3 (3 x 64), y at 192
constexpr int MLP_WGRAD_TILES = 53;  // 40 weight tiles + 13 bias tiles
constexpr int MLP_WGRAD_SLICE = 128; // samples per weight-gradient block

struct MlpDev {
    float* params = nullptr;    // FP32 master parameters (W1 b1 W2 b2 W3 b3 W4 b4)
    float* adamM = nullptr;
    float* adamV = nullptr;
    float* grad = nullptr;
    void* fwd = nullptr;        // MLP_FWD_FRAGS x 64 lanes x 8 BF16
    void* bwd = nullptr;        // MLP_BWD_FRAGS x 64 lanes x 8 BF16
    float* store = nullptr;     // cap x MLP_STORE: pre-activations / output
    float* dz = nullptr;        // cap x MLP_STORE: gradients w.r.t. pre-activations
    float* dy = nullptr;        // cap
    float* scratchY = nullptr;  // cap
    double* lossPart = nullptr; // cap / 256 + 1
    float* gradPart = nullptr;  // slices x MLP_WGRAD_TILES x 256 partial gradient tiles
    float overWeight = 1.0f;    // loss weight where the prediction exceeds the target
    int cap = 0;                // samples per gradient evaluation
    int adamStep = 0;
};

void mlp_pack(MlpDev& m, hipStream_t st);
// y[i] = MLP(feat[8 i .. 8 i + 7]); store (cap x MLP_STORE) receives activations when non-null
// gate1 (inference only): tiles of 16 particles whose feature 1 is negative
// for all 16 are skipped, their outputs left as they were
void mlp_forward(const MlpDev& m, const float* feat, float* y, int n, float* store, hipStream_t st,
                 bool gate1 = false);
// gradient of mean((MLP(feat) - target)^2) into m.grad (n <= m.cap); loss parts in m.lossPart
void mlp_grad(MlpDev& m, const float* feat, const float* target, int n, hipStream_t st);
void mlp_adam(MlpDev& m, float lr, hipStream_t st);

}  // namespace sphgpu

#include <vector>
namespace sphgpu {
void mlp_init_params(std::vector<float>& p, unsigned long long seed);
int mlp_alloc(MlpDev& m, int cap);
void mlp_free(MlpDev& m);
double mlp_loss_host(const MlpDev& m, int n, hipStream_t st);
}  // namespace sphgpu
