// Interface of the learned-correction MLP (mgsb_mlp.hip) used by the solver.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define MLP_HIDDEN 128

namespace mgsb {

struct Mlp;

int mlp_nparams(int in, int out);
int mlp_create(Mlp** out, int in, int outd, uint32_t seed, hipStream_t st);
void mlp_destroy(Mlp* p);
int mlp_params_count(const Mlp* p);
int mlp_in(const Mlp* p);
int mlp_out(const Mlp* p);
int mlp_set_params(Mlp* p, const float* w, int n, hipStream_t st);
int mlp_get_params(Mlp* p, float* w, int n, hipStream_t st);
// Fused BF16-MFMA inference: X (n x in, device) -> Y (row stride ldy, device).
int mlp_infer(Mlp* p, const float* X, int n, float* Y, int ldy, hipStream_t st, const int* skip);
// Value at a coarse vertex of a fine-level field u through its
// coarse-in-fine binding (four fine vertices id, weights w), summed in
// corner order.
__device__ __forceinline__ float3 inject4(const float4* __restrict__ u, int4 id, float4 w) {
    const int ids[4] = {id.x, id.y, id.z, id.w};
    const float ws[4] = {w.x, w.y, w.z, w.w};
    float dx = 0.f, dy = 0.f, dz = 0.f;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const float4 a = u[ids[q]];
        dx = dx + ws[q] * a.x;
        dy = dy + ws[q] * a.y;
        dz = dz + ws[q] * a.z;
    }
    return make_float3(dx, dy, dz);
}
// The learned coarse correction in one launch (features, network, prior and
// mask): epred[j] = vcl[j].z * (e1_j + y_j * out_sd + out_mu), y = mlp of the
// standardised features (injected fine velocity vf, rest position X1, last
// coarse deltas e1 = inject(D1), e2 = inject(D2)) of each of the n level-1
// vertices; every injection goes through the binding (aids, aw).
constexpr int kCoarseFeat = 12;
struct CoarseIO {
    const float4* vf;
    const int4* aids;
    const float4* aw;
    const float4* X1;
    const float4* D1;  // fine level
    const float4* D2;
    const float* norm;  // in_mu[12] in_sd[12] out_mu[3] out_sd[3]
    const float4* vcl;
    float4* epred;
};
int mlp_infer_coarse(Mlp* p, int n, const CoarseIO& io, hipStream_t st);
// Test paths: gradients of the mean squared error, and the GEMM-path forward.
int mlp_grad(Mlp* p, const float* X_host, const float* T_host, int n, float* grad_host, double* loss, hipStream_t st);
int mlp_forward_train_path(Mlp* p, const float* X_host, int n, float* Y_host, hipStream_t st);
// Adam training on device-resident samples X (n x in), T (n x out).
int mlp_train(Mlp* p, const float* X, const float* T, long n, int epochs, int batch, float lr, uint32_t seed,
              float* loss_hist, hipStream_t st);

}  // namespace mgsb
