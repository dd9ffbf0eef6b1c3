// Small MLP (in -> H -> H -> H -> out, ReLU) for the learned coarse
// correction: BF16 MFMA (v_mfma_f32_16x16x16_bf16) with FP32 accumulation,
// FP32 master weights, Adam.  Training and inference kernels are written
// here; nothing comes from a library.
//
// MFMA 16x16x16 operand layout (one wave, 64 lanes, lane l):
//   A: row l%16, k = 4*(l/16)+q   B: col l%16, k = 4*(l/16)+q   (q = 0..3)
//   D: col l%16, row 4*(l/16)+q
// Operands are rounded FP32 -> BF16 (nearest-even) as they are loaded.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mgsb_device.h"
#include "mgsb_internal.h"
#include "mgsb_mlp.h"

namespace mgsb {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (short)(u >> 16);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(short b) { return __uint_as_float(((uint32_t)(unsigned short)b) << 16); }

#define MLPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) {                                                               \
            set_error(std::string("HIP error ") + hipGetErrorString(_e) + " at " #expr);      \
            return MGSB_ERR_HIP;                                                              \
        }                                                                                     \
    } while (0)

// C(m,n) = sum_k A(m,k) B(k,n) (+ bias[n]); A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn].  One wave per 16x16 tile of C, four tiles per
// workgroup.  act: 0 none, 1 ReLU into H (Z keeps the pre-activation).
__global__ void __launch_bounds__(256) k_gemm_bf16(int M, int N, int K, const float* __restrict__ A, long sam,
                                                   long sak, const float* __restrict__ B, long sbk, long sbn,
                                                   const float* __restrict__ bias, float* __restrict__ Z,
                                                   float* __restrict__ H, int ldc, int act) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tiles_n = (N + 15) >> 4;
    const int tile = blockIdx.x * 4 + wave;
    const int tm = tile / tiles_n, tn = tile % tiles_n;
    if (tm * 16 >= M) return;
    const int row = tm * 16 + (lane & 15), col = tn * 16 + (lane & 15), kq = (lane >> 4) * 4;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int k = k0 + kq + q;
            a[q] = f2bf((row < M && k < K) ? A[row * sam + k * sak] : 0.f);
            b[q] = f2bf((col < N && k < K) ? B[k * sbk + col * sbn] : 0.f);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const int oc = tn * 16 + (lane & 15);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int orow = tm * 16 + (lane >> 4) * 4 + q;
        if (orow < M && oc < N) {
            float v = acc[q] + (bias ? bias[oc] : 0.f);
            Z[(long)orow * ldc + oc] = v;
            if (act == 1) H[(long)orow * ldc + oc] = v > 0.f ? v : 0.f;
        }
    }
}

// Split-K form of k_gemm_bf16 for products with few output tiles and a long
// K (the weight gradients, K = batch rows): block (x, c) runs four tiles over
// the K chunk [c*kc, min(K, (c+1)*kc)) and stores each lane's four MFMA
// accumulators as part[(c*tiles + tile)*256 + lane*4 + q].
__global__ void __launch_bounds__(256) k_gemm_bf16_part(int M, int N, int K, int kc, const float* __restrict__ A,
                                                        long sam, long sak, const float* __restrict__ B, long sbk,
                                                        long sbn, float* __restrict__ part) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tiles_n = (N + 15) >> 4, tiles = ((M + 15) >> 4) * tiles_n;
    const int tile = blockIdx.x * 4 + wave;
    if (tile >= tiles) return;
    const int tm = tile / tiles_n, tn = tile % tiles_n;
    const int row = tm * 16 + (lane & 15), col = tn * 16 + (lane & 15), kq = (lane >> 4) * 4;
    const int kb = blockIdx.y * kc, ke = min(K, kb + kc);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = kb; k0 < ke; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int k = k0 + kq + q;
            a[q] = f2bf((row < M && k < ke) ? A[row * sam + k * sak] : 0.f);
            b[q] = f2bf((col < N && k < ke) ? B[k * sbk + col * sbn] : 0.f);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    *reinterpret_cast<f32x4*>(part + ((size_t)blockIdx.y * tiles + tile) * 256 + lane * 4) = acc;
}

// Second stage of the split-K product: each lane adds its accumulators over
// the S chunks in chunk order (the same sum on every run), then the output
// mapping, bias and activation of k_gemm_bf16.
__global__ void __launch_bounds__(256) k_gemm_splitk_final(int M, int N, int S, const float* __restrict__ part,
                                                           const float* __restrict__ bias, float* __restrict__ Z,
                                                           float* __restrict__ H, int ldc, int act) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tiles_n = (N + 15) >> 4, tiles = ((M + 15) >> 4) * tiles_n;
    const int tile = blockIdx.x * 4 + wave;
    if (tile >= tiles) return;
    const int tm = tile / tiles_n, tn = tile % tiles_n;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int c = 0; c < S; c++) acc += *reinterpret_cast<const f32x4*>(part + ((size_t)c * tiles + tile) * 256 + lane * 4);
    const int oc = tn * 16 + (lane & 15);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int orow = tm * 16 + (lane >> 4) * 4 + q;
        if (orow < M && oc < N) {
            float v = acc[q] + (bias ? bias[oc] : 0.f);
            Z[(long)orow * ldc + oc] = v;
            if (act == 1) H[(long)orow * ldc + oc] = v > 0.f ? v : 0.f;
        }
    }
}

// dY = 2 (Y - T) / (n*out); per-block partial sums of (Y - T)^2.
__global__ void __launch_bounds__(256) k_mse(int n, const float* __restrict__ Y, const float* __restrict__ T,
                                             float* __restrict__ dY, float scale, double* __restrict__ part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const float d = Y[i] - T[i];
        dY[i] = 2.f * d * scale;
        v = (double)d * (double)d;
    }
    __shared__ double sh[256];
    sh[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (int t = 0; t < (int)blockDim.x; t++) a += sh[t];
        part[blockIdx.x] = a;
    }
}

__global__ void __launch_bounds__(256) k_relu_bwd(int n, const float* __restrict__ Z, float* __restrict__ dA) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(Z[i] > 0.f)) dA[i] = 0.f;
}

// db[j] = sum_rows dZ[r][j] in two deterministic stages.  Stage 1: block
// (cx, ry) sums rows [ry*kColRows, +kColRows) of columns [cx*64, +64): each of
// the 4 row lanes of a column sums every 4th row in order, the 4 lane sums
// are added in lane order, and the chunk sum goes to part[ry][j].  Stage 2
// adds the chunk sums of each column in chunk order.
constexpr int kColRows = 128;
__global__ void __launch_bounds__(256) k_colsum_part(int rows, int cols, const float* __restrict__ dZ,
                                                     float* __restrict__ part) {
    __shared__ float sh[4][64];
    const int c = threadIdx.x & 63, lane = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + c;
    const int r0 = blockIdx.y * kColRows, r1 = min(rows, r0 + kColRows);
    float a = 0.f;
    if (j < cols)
        for (int r = r0 + lane; r < r1; r += 4) a += dZ[(long)r * cols + j];
    sh[lane][c] = a;
    __syncthreads();
    if (lane == 0 && j < cols) part[(long)blockIdx.y * cols + j] = ((sh[0][c] + sh[1][c]) + sh[2][c]) + sh[3][c];
}

__global__ void __launch_bounds__(256) k_colsum_final(int nchunks, int cols, const float* __restrict__ part,
                                                      float* __restrict__ db) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    float a = 0.f;
    for (int q = 0; q < nchunks; q++) a += part[(long)q * cols + j];
    db[j] = a;
}

__global__ void __launch_bounds__(256) k_adam(int n, float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                              float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Rows idx[b] of X (n x in) and T (n x out) into a batch.
__global__ void __launch_bounds__(256) k_gather_rows(int nb, const int* __restrict__ idx, const float* __restrict__ X,
                                                     int in, const float* __restrict__ T, int out,
                                                     float* __restrict__ Xb, float* __restrict__ Tb) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const long r = idx[b];
    for (int c = 0; c < in; c++) Xb[(long)b * in + c] = X[r * in + c];
    for (int c = 0; c < out; c++) Tb[(long)b * out + c] = T[r * out + c];
}

// BF16 copies of the weights, zero-padded to multiples of 16, for inference.
// BF16 copy of a K x N weight (row-major FP32), zero-padded to Kp x Np and
// stored in the order the 16x16x16 MFMA reads its B operand: fragment
// (tile t of 16 columns, K block kb) is 64 lanes x 4 consecutive values,
// lane l holding W[kb*16 + (l/16)*4 + q][t*16 + l%16], q = 0..3, so a
// wave loads one fragment as one coalesced 512-byte read.
__global__ void __launch_bounds__(256) k_pack_bf16(int K, int N, int Kp, int Np, const float* __restrict__ W,
                                                   short* __restrict__ Wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Kp * Np) return;
    const int nkb = Kp / 16;
    const int frag = i >> 8, lane = (i >> 2) & 63, q = i & 3;
    const int t = frag / nkb, kb = frag % nkb;
    const int k = kb * 16 + (lane >> 4) * 4 + q, n = t * 16 + (lane & 15);
    Wb[i] = f2bf((k < K && n < N) ? W[k * N + n] : 0.f);
}
__device__ __forceinline__ bf16x4 frag_b(const short* __restrict__ Wb, int nkb, int t, int kb, int lane) {
    return *reinterpret_cast<const bf16x4*>(Wb + ((size_t)(t * nkb + kb) * 64 + lane) * 4);
}

// Fused inference: one workgroup of kMlpWaves waves per 16 samples through
// all four layers.  Each wave computes kH / 16 / kMlpWaves of a hidden
// layer's eight 16-column output tiles; the layer's activations stay in LDS
// as BF16 and a barrier separates the layers.  Wave 0 computes the output layer (one tile).  Every
// tile accumulates its K blocks in order, as a one-wave-per-tile product
// would.  All of a wave's weight fragments (k_pack_bf16's order) are loaded
// at the start, in one round trip, instead of one dependent load per
// product.  Y = mlp(X).
constexpr int kH = MLP_HIDDEN;
// waves per inference workgroup: each computes kH / 16 / kMlpWaves of a
// hidden layer's 16-column tiles.  Eight waves, one tile each: 7.0-7.3 us per
// launch of the learned stage on the 1,025-vertex level against 7.4-7.5 with
// four and 8.6 with two (with the activations' rows unpadded, whose reads
// conflicted 16-way, eight had measured slower: 10.6 against 9.0)
constexpr int kMlpWaves = 8;

// Row-major FP32 inputs X (n x in) and outputs Y (row stride ldy).
struct DenseIn {
    const float* X;
    int n, in;
    __device__ void load4(int row, int k0, float (&v)[4], float* /*prior*/) const {
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = (row < n && k0 + q < in) ? X[(long)row * in + k0 + q] : 0.f;
    }
};
struct DenseOut {
    float* Y;
    int n, out, ldy;
    __device__ float2 pre(int /*col*/) const { return make_float2(0.f, 1.f); }
    __device__ void store(int row, int col, float v, const float* /*prior*/, float2 /*pre*/) const {
        if (row < n && col < out) Y[(long)row * ldy + col] = v;
    }
};

template <class In, class Out>
__global__ void __launch_bounds__(64 * kMlpWaves) k_mlp_infer(int n, int nout, In in, Out out, const short* __restrict__ W1,
                                                   const short* __restrict__ W2, const short* __restrict__ W3,
                                                   const short* __restrict__ W4, const float* __restrict__ b1,
                                                   const float* __restrict__ b2, const float* __restrict__ b3,
                                                   const float* __restrict__ b4, const int* __restrict__ skip) {
    // the flag loads with the weights and biases; a skipped launch returns
    // before the inputs' loads
    const int sk = load_flag(skip);
    // rows padded by 8 bytes: the 16 rows a wave's A-fragment read spans
    // (and the four rows of an activation store) fall in different banks
    __shared__ __align__(16) short hs[2][16][kH + 4];
    // per-row values the input functor hands to the output functor (the
    // coarse correction's prior), written by wave 0's loads
    __shared__ float pri[16][4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int base = blockIdx.x * 16;
    const int r = lane & 15, kq = (lane >> 4) * 4;
    constexpr int kTilesPerWave = kH / 16 / kMlpWaves;
    constexpr int kKb = kH / 16;
    // every weight fragment this wave multiplies by, issued up front
    bf16x4 w1[kTilesPerWave], w23[2][kTilesPerWave][kKb], w4[kKb];
#pragma unroll
    for (int tt = 0; tt < kTilesPerWave; tt++) {
        const int t = wave * kTilesPerWave + tt;
        w1[tt] = frag_b(W1, 1, t, 0, lane);
#pragma unroll
        for (int kb = 0; kb < kKb; kb++) {
            w23[0][tt][kb] = frag_b(W2, kKb, t, kb, lane);
            w23[1][tt][kb] = frag_b(W3, kKb, t, kb, lane);
        }
    }
    if (wave == 0)
#pragma unroll
        for (int kb = 0; kb < kKb; kb++) w4[kb] = frag_b(W4, kKb, 0, kb, lane);
    // the biases and the output's constants in the same round trip
    float bb[3][kTilesPerWave];
#pragma unroll
    for (int tt = 0; tt < kTilesPerWave; tt++) {
        const int t = wave * kTilesPerWave + tt;
        bb[0][tt] = b1[t * 16 + r];
        bb[1][tt] = b2[t * 16 + r];
        bb[2][tt] = b3[t * 16 + r];
    }
    const float bias4 = (wave == 0 && r < nout) ? b4[r] : 0.f;
    const float2 opre = (wave == 0 && r < nout) ? out.pre(r) : make_float2(0.f, 1.f);
    if (sk) return;  // the same flag for the whole workgroup
    // layer 1: K = 16 (inputs zero-padded)
    {
        bf16x4 a;
        float x4[4];
        in.load4(base + r, kq, x4, wave == 0 ? &pri[r][0] : nullptr);
#pragma unroll
        for (int q = 0; q < 4; q++) a[q] = f2bf(x4[q]);
#pragma unroll
        for (int tt = 0; tt < kTilesPerWave; tt++) {
            const int t = wave * kTilesPerWave + tt;
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, w1[tt], acc, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float v = acc[q] + bb[0][tt];
                hs[0][kq + q][t * 16 + r] = f2bf(v > 0.f ? v : 0.f);
            }
        }
    }
    __syncthreads();
    // layers 2 and 3: K = H
#pragma unroll
    for (int L = 0; L < 2; L++) {
        const int src = L & 1, dst = src ^ 1;
#pragma unroll
        for (int tt = 0; tt < kTilesPerWave; tt++) {
            const int t = wave * kTilesPerWave + tt;
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < kKb; kb++) {
                const bf16x4 a = *reinterpret_cast<const bf16x4*>(&hs[src][r][kb * 16 + kq]);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, w23[L][tt][kb], acc, 0, 0, 0);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float v = acc[q] + bb[1 + L][tt];
                hs[dst][kq + q][t * 16 + r] = f2bf(v > 0.f ? v : 0.f);
            }
        }
        __syncthreads();
    }
    // layer 4: N = 16 (outputs zero-padded); source buffer is 0 after 2 layers
    if (wave != 0) return;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < kKb; kb++) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(&hs[0][r][kb * 16 + kq]);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, w4[kb], acc, 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < 4; q++)
        if (r < nout) out.store(base + kq + q, r, acc[q] + bias4, &pri[kq + q][0], opre);
}

// ---------------------------------------------------------------------------
struct Mlp {
    int in = 0, out = 0;
    int off[8];            // offsets of W1 b1 W2 b2 W3 b3 W4 b4 in the flat vector
    int nparams = 0;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // device, nparams each
    short *W1b = nullptr, *W2b = nullptr, *W3b = nullptr, *W4b = nullptr;
    long adam_t = 0;
    // training scratch
    int cap = 0;
    float *Xb = nullptr, *Tb = nullptr, *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr, *Z3 = nullptr,
          *H3 = nullptr, *Y = nullptr, *dY = nullptr, *dA = nullptr, *dB = nullptr;
    int* idx = nullptr;
    double* part = nullptr;
    float* colpart = nullptr;  // column-sum chunk partials
    float* kpart = nullptr;    // split-K partial tiles
    size_t kpart_cap = 0;      // floats
    std::vector<void*> owned;
};

static int mlp_alloc(Mlp* p, void** ptr, size_t bytes) {
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 16);
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return MGSB_ERR_NOMEM;
    }
    p->owned.push_back(*ptr);
    return MGSB_OK;
}

int mlp_nparams(int in, int out) {
    return in * kH + kH + 2 * (kH * kH + kH) + kH * out + out;
}

static void set_offsets(Mlp* p) {
    int o = 0;
    const int sizes[8] = {p->in * kH, kH, kH * kH, kH, kH * kH, kH, kH * p->out, p->out};
    for (int i = 0; i < 8; i++) {
        p->off[i] = o;
        o += sizes[i];
    }
    p->nparams = o;
}

int mlp_pack(Mlp* p, hipStream_t st) {
    const int bs = 256;
    k_pack_bf16<<<(16 * kH + bs - 1) / bs, bs, 0, st>>>(p->in, kH, 16, kH, p->w + p->off[0], p->W1b);
    k_pack_bf16<<<(kH * kH + bs - 1) / bs, bs, 0, st>>>(kH, kH, kH, kH, p->w + p->off[2], p->W2b);
    k_pack_bf16<<<(kH * kH + bs - 1) / bs, bs, 0, st>>>(kH, kH, kH, kH, p->w + p->off[4], p->W3b);
    k_pack_bf16<<<(kH * 16 + bs - 1) / bs, bs, 0, st>>>(kH, p->out, kH, 16, p->w + p->off[6], p->W4b);
    MLPCHK(hipGetLastError());
    return MGSB_OK;
}

int mlp_create(Mlp** out, int in, int outd, uint32_t seed, hipStream_t st) {
    if (in < 1 || in > 16 || outd < 1 || outd > 16) {
        set_error("mlp_create: input and output widths must be in [1, 16]");
        return MGSB_ERR_ARG;
    }
    Mlp* p = new Mlp();
    p->in = in;
    p->out = outd;
    set_offsets(p);
    int rc;
    const size_t nb = sizeof(float) * p->nparams;
    if ((rc = mlp_alloc(p, (void**)&p->w, nb)) || (rc = mlp_alloc(p, (void**)&p->g, nb)) ||
        (rc = mlp_alloc(p, (void**)&p->m, nb)) || (rc = mlp_alloc(p, (void**)&p->v, nb)) ||
        (rc = mlp_alloc(p, (void**)&p->W1b, sizeof(short) * 16 * kH)) ||
        (rc = mlp_alloc(p, (void**)&p->W2b, sizeof(short) * kH * kH)) ||
        (rc = mlp_alloc(p, (void**)&p->W3b, sizeof(short) * kH * kH)) ||
        (rc = mlp_alloc(p, (void**)&p->W4b, sizeof(short) * kH * 16))) {
        mlp_destroy(p);
        return rc;
    }
    // He-uniform init from a seeded xorshift stream; biases 0
    std::vector<float> w(p->nparams, 0.f);
    uint64_t x = 0x9E3779B97F4A7C15ull ^ seed;
    auto next = [&]() {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return (double)(x >> 11) * (1.0 / 9007199254740992.0);
    };
    const int fan_in[4] = {in, kH, kH, kH};
    const int cnt[4] = {in * kH, kH * kH, kH * kH, kH * outd};
    for (int L = 0; L < 4; L++) {
        const double lim = std::sqrt(6.0 / fan_in[L]) * (L == 3 ? 0.1 : 1.0);
        for (int i = 0; i < cnt[L]; i++) w[p->off[2 * L] + i] = (float)((2 * next() - 1) * lim);
    }
    rc = mlp_set_params(p, w.data(), p->nparams, st);
    if (rc) {
        mlp_destroy(p);
        return rc;
    }
    *out = p;
    return MGSB_OK;
}

void mlp_destroy(Mlp* p) {
    if (!p) return;
    for (void* q : p->owned) (void)hipFree(q);
    if (p->kpart) (void)hipFree(p->kpart);
    delete p;
}

int mlp_params_count(const Mlp* p) { return p->nparams; }
int mlp_in(const Mlp* p) { return p->in; }
int mlp_out(const Mlp* p) { return p->out; }

int mlp_set_params(Mlp* p, const float* w, int n, hipStream_t st) {
    if (n != p->nparams) {
        set_error("mlp_set_params: expected " + std::to_string(p->nparams) + " parameters");
        return MGSB_ERR_ARG;
    }
    MLPCHK(hipMemcpyAsync(p->w, w, sizeof(float) * n, hipMemcpyHostToDevice, st));
    MLPCHK(hipMemsetAsync(p->m, 0, sizeof(float) * n, st));
    MLPCHK(hipMemsetAsync(p->v, 0, sizeof(float) * n, st));
    p->adam_t = 0;
    int rc = mlp_pack(p, st);
    if (rc) return rc;
    MLPCHK(hipStreamSynchronize(st));
    return MGSB_OK;
}

int mlp_get_params(Mlp* p, float* w, int n, hipStream_t st) {
    if (n != p->nparams) {
        set_error("mlp_get_params: expected " + std::to_string(p->nparams) + " parameters");
        return MGSB_ERR_ARG;
    }
    MLPCHK(hipMemcpyAsync(w, p->w, sizeof(float) * n, hipMemcpyDeviceToHost, st));
    MLPCHK(hipStreamSynchronize(st));
    return MGSB_OK;
}

int mlp_infer(Mlp* p, const float* X, int n, float* Y, int ldy, hipStream_t st, const int* skip) {
    const int blocks = (n + 15) / 16;  // one workgroup per 16 samples
    if (blocks == 0) return MGSB_OK;
    k_mlp_infer<<<blocks, 64 * kMlpWaves, 0, st>>>(n, p->out, DenseIn{X, n, p->in}, DenseOut{Y, n, p->out, ldy}, p->W1b, p->W2b, p->W3b,
                                        p->W4b, p->w + p->off[1], p->w + p->off[3], p->w + p->off[5],
                                        p->w + p->off[7], skip);
    MLPCHK(hipGetLastError());
    return MGSB_OK;
}

// The learned coarse correction as one launch: the features of each level-1
// vertex (fine velocity injected through the coarse-in-fine binding, rest
// position, the last two coarse deltas), standardised, through the network,
// and its output de-standardised onto the last delta and masked by the
// level's Dirichlet flags (the features are k_features' values).
struct CoarseIn {
    CoarseIO io;
    int n;
    // prior (when not null): the row's injected last delta, kept for the output
    __device__ void load4(int row, int k0, float (&v)[4], float* prior) const {
        float raw[4] = {0.f, 0.f, 0.f, 0.f};
        if (row < n) {
            if (k0 == 0) {
                const int4 id = io.aids[row];
                const float4 w = io.aw[row];
                const float4 a = io.vf[id.x], b = io.vf[id.y], c = io.vf[id.z], d = io.vf[id.w];
                raw[0] = w.x * a.x + w.y * b.x + w.z * c.x + w.w * d.x;
                raw[1] = w.x * a.y + w.y * b.y + w.z * c.y + w.w * d.y;
                raw[2] = w.x * a.z + w.y * b.z + w.z * c.z + w.w * d.z;
                raw[3] = io.X1[row].x;
                if (prior) prior[3] = io.vcl[row].z;  // the output's Dirichlet mask
            } else if (k0 == 4) {
                const float4 X = io.X1[row];
                const float3 e = inject4(io.D1, io.aids[row], io.aw[row]);
                raw[0] = X.y;
                raw[1] = X.z;
                raw[2] = e.x;
                raw[3] = e.y;
                if (prior) {
                    prior[0] = e.x;
                    prior[1] = e.y;
                }
            } else if (k0 == 8) {
                const int4 id = io.aids[row];
                const float4 w = io.aw[row];
                const float3 e = inject4(io.D1, id, w), f = inject4(io.D2, id, w);
                raw[0] = e.z;
                raw[1] = f.x;
                raw[2] = f.y;
                raw[3] = f.z;
                if (prior) prior[2] = e.z;
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int k = k0 + q;
            v[q] = (row < n && k < kCoarseFeat) ? (raw[q] - io.norm[k]) / io.norm[kCoarseFeat + k] : 0.f;
        }
    }
};
struct CoarseOut {
    CoarseIO io;
    int n;
    // the output column's (mean, sd), loaded with the weights
    __device__ float2 pre(int col) const {
        return make_float2(io.norm[2 * kCoarseFeat + col], io.norm[2 * kCoarseFeat + 3 + col]);
    }
    // prior: the row's injected last delta and its mask, as load4 read them
    __device__ void store(int row, int col, float v, const float* prior, float2 ms) const {
        if (row >= n || col >= 3) return;
        reinterpret_cast<float*>(io.epred + row)[col] = prior[3] * (prior[col] + (v * ms.y + ms.x));
    }
};

int mlp_infer_coarse(Mlp* p, int n, const CoarseIO& io, hipStream_t st) {
    if (p->in != kCoarseFeat || p->out != 3) {
        set_error("mlp_infer_coarse: the network must map 12 features to 3 outputs");
        return MGSB_ERR_STATE;
    }
    const int blocks = (n + 15) / 16;
    if (blocks == 0) return MGSB_OK;
    k_mlp_infer<<<blocks, 64 * kMlpWaves, 0, st>>>(n, 3, CoarseIn{io, n}, CoarseOut{io, n}, p->W1b, p->W2b, p->W3b, p->W4b,
                                        p->w + p->off[1], p->w + p->off[3], p->w + p->off[5], p->w + p->off[7],
                                        nullptr);
    MLPCHK(hipGetLastError());
    return MGSB_OK;
}

static void colsum(Mlp* p, int rows, int cols, const float* dZ, float* db, hipStream_t st) {
    const int nch = (rows + kColRows - 1) / kColRows;
    k_colsum_part<<<dim3((cols + 63) / 64, nch), 256, 0, st>>>(rows, cols, dZ, p->colpart);
    k_colsum_final<<<(cols + 255) / 256, 256, 0, st>>>(nch, cols, p->colpart, db);
}

static int ensure_cap(Mlp* p, int cap) {
    if (p->cap >= cap) return MGSB_OK;
    // scratch is only ever grown; old buffers stay owned until destroy
    int rc;
    const size_t fh = sizeof(float) * (size_t)cap * kH;
    if ((rc = mlp_alloc(p, (void**)&p->Xb, sizeof(float) * (size_t)cap * p->in)) ||
        (rc = mlp_alloc(p, (void**)&p->Tb, sizeof(float) * (size_t)cap * p->out)) ||
        (rc = mlp_alloc(p, (void**)&p->Z1, fh)) || (rc = mlp_alloc(p, (void**)&p->H1, fh)) ||
        (rc = mlp_alloc(p, (void**)&p->Z2, fh)) || (rc = mlp_alloc(p, (void**)&p->H2, fh)) ||
        (rc = mlp_alloc(p, (void**)&p->Z3, fh)) || (rc = mlp_alloc(p, (void**)&p->H3, fh)) ||
        (rc = mlp_alloc(p, (void**)&p->Y, sizeof(float) * (size_t)cap * p->out)) ||
        (rc = mlp_alloc(p, (void**)&p->dY, sizeof(float) * (size_t)cap * p->out)) ||
        (rc = mlp_alloc(p, (void**)&p->dA, fh)) || (rc = mlp_alloc(p, (void**)&p->dB, fh)) ||
        (rc = mlp_alloc(p, (void**)&p->idx, sizeof(int) * (size_t)cap)) ||
        (rc = mlp_alloc(p, (void**)&p->part, sizeof(double) * ((size_t)cap * p->out / 256 + 2))) ||
        (rc = mlp_alloc(p, (void**)&p->colpart, sizeof(float) * ((size_t)cap / kColRows + 1) * kH)))
        return rc;
    p->cap = cap;
    return MGSB_OK;
}

// Products with few output tiles and a long K (fewer than kSplitTiles tiles
// of 16x16, K of at least 2 kSplitChunk) are split over K into chunks of
// kSplitChunk, at most kSplitMax of them, and reduced in a second launch.
constexpr int kSplitTiles = 256, kSplitChunk = 256, kSplitMax = 64;

static int gemm(Mlp* p, int M, int N, int K, const float* A, long sam, long sak, const float* B, long sbk, long sbn,
                const float* bias, float* Z, float* H, int ldc, int act, hipStream_t st) {
    const int tiles = ((M + 15) / 16) * ((N + 15) / 16);
    if (tiles < kSplitTiles && K >= 2 * kSplitChunk) {
        int S = (K + kSplitChunk - 1) / kSplitChunk;
        S = S < kSplitMax ? S : kSplitMax;
        const int kc = ((K + S - 1) / S + 15) / 16 * 16;  // chunks of whole 16-steps
        S = (K + kc - 1) / kc;
        const size_t need = (size_t)S * tiles * 256;
        if (need > p->kpart_cap) {
            if (p->kpart) MLPCHK(hipFree(p->kpart));
            p->kpart = nullptr;
            p->kpart_cap = 0;
            MLPCHK(hipMalloc(&p->kpart, need * sizeof(float)));
            p->kpart_cap = need;
        }
        k_gemm_bf16_part<<<dim3((tiles + 3) / 4, S), 256, 0, st>>>(M, N, K, kc, A, sam, sak, B, sbk, sbn, p->kpart);
        MLPCHK(hipGetLastError());
        k_gemm_splitk_final<<<(tiles + 3) / 4, 256, 0, st>>>(M, N, S, p->kpart, bias, Z, H, ldc, act);
        MLPCHK(hipGetLastError());
        return MGSB_OK;
    }
    k_gemm_bf16<<<(tiles + 3) / 4, 256, 0, st>>>(M, N, K, A, sam, sak, B, sbk, sbn, bias, Z, H, ldc, act);
    MLPCHK(hipGetLastError());
    return MGSB_OK;
}

// Forward + backward on nb rows already in Xb / Tb; gradients into p->g;
// returns the batch MSE (mean over nb*out) in *loss when loss != nullptr.
static int fwd_bwd(Mlp* p, int nb, hipStream_t st, double* loss) {
    const int in = p->in, out = p->out;
    float* W1 = p->w + p->off[0];
    float* W2 = p->w + p->off[2];
    float* W3 = p->w + p->off[4];
    float* W4 = p->w + p->off[6];
    int rc;
    if ((rc = gemm(p, nb, kH, in, p->Xb, in, 1, W1, kH, 1, p->w + p->off[1], p->Z1, p->H1, kH, 1, st))) return rc;
    if ((rc = gemm(p, nb, kH, kH, p->H1, kH, 1, W2, kH, 1, p->w + p->off[3], p->Z2, p->H2, kH, 1, st))) return rc;
    if ((rc = gemm(p, nb, kH, kH, p->H2, kH, 1, W3, kH, 1, p->w + p->off[5], p->Z3, p->H3, kH, 1, st))) return rc;
    if ((rc = gemm(p, nb, out, kH, p->H3, kH, 1, W4, out, 1, p->w + p->off[7], p->Y, nullptr, out, 0, st))) return rc;
    const int ny = nb * out;
    const int nbl = (ny + 255) / 256;
    k_mse<<<nbl, 256, 0, st>>>(ny, p->Y, p->Tb, p->dY, 1.f / (float)ny, p->part);
    MLPCHK(hipGetLastError());
    float* g = p->g;
    // layer 4: dW4 = H3^T dY, db4 = colsum dY, dA3 = dY W4^T
    if ((rc = gemm(p, kH, out, nb, p->H3, 1, kH, p->dY, out, 1, nullptr, g + p->off[6], nullptr, out, 0, st))) return rc;
    colsum(p, nb, out, p->dY, g + p->off[7], st);
    if ((rc = gemm(p, nb, kH, out, p->dY, out, 1, W4, 1, out, nullptr, p->dA, nullptr, kH, 0, st))) return rc;
    const int nh = nb * kH;
    k_relu_bwd<<<(nh + 255) / 256, 256, 0, st>>>(nh, p->Z3, p->dA);
    // layer 3
    if ((rc = gemm(p, kH, kH, nb, p->H2, 1, kH, p->dA, kH, 1, nullptr, g + p->off[4], nullptr, kH, 0, st))) return rc;
    colsum(p, nb, kH, p->dA, g + p->off[5], st);
    if ((rc = gemm(p, nb, kH, kH, p->dA, kH, 1, W3, 1, kH, nullptr, p->dB, nullptr, kH, 0, st))) return rc;
    k_relu_bwd<<<(nh + 255) / 256, 256, 0, st>>>(nh, p->Z2, p->dB);
    // layer 2
    if ((rc = gemm(p, kH, kH, nb, p->H1, 1, kH, p->dB, kH, 1, nullptr, g + p->off[2], nullptr, kH, 0, st))) return rc;
    colsum(p, nb, kH, p->dB, g + p->off[3], st);
    if ((rc = gemm(p, nb, kH, kH, p->dB, kH, 1, W2, 1, kH, nullptr, p->dA, nullptr, kH, 0, st))) return rc;
    k_relu_bwd<<<(nh + 255) / 256, 256, 0, st>>>(nh, p->Z1, p->dA);
    // layer 1
    if ((rc = gemm(p, in, kH, nb, p->Xb, 1, in, p->dA, kH, 1, nullptr, g + p->off[0], nullptr, kH, 0, st))) return rc;
    colsum(p, nb, kH, p->dA, g + p->off[1], st);
    MLPCHK(hipGetLastError());
    if (loss) {
        std::vector<double> part(nbl);
        MLPCHK(hipMemcpyAsync(part.data(), p->part, sizeof(double) * nbl, hipMemcpyDeviceToHost, st));
        MLPCHK(hipStreamSynchronize(st));
        double a = 0.0;
        for (double v : part) a += v;
        *loss = a / ny;
    }
    return MGSB_OK;
}

int mlp_grad(Mlp* p, const float* X_host, const float* T_host, int n, float* grad_host, double* loss,
             hipStream_t st) {
    int rc = ensure_cap(p, n);
    if (rc) return rc;
    MLPCHK(hipMemcpyAsync(p->Xb, X_host, sizeof(float) * (size_t)n * p->in, hipMemcpyHostToDevice, st));
    MLPCHK(hipMemcpyAsync(p->Tb, T_host, sizeof(float) * (size_t)n * p->out, hipMemcpyHostToDevice, st));
    if ((rc = fwd_bwd(p, n, st, loss))) return rc;
    MLPCHK(hipMemcpyAsync(grad_host, p->g, sizeof(float) * p->nparams, hipMemcpyDeviceToHost, st));
    MLPCHK(hipStreamSynchronize(st));
    return MGSB_OK;
}

int mlp_forward_train_path(Mlp* p, const float* X_host, int n, float* Y_host, hipStream_t st) {
    int rc = ensure_cap(p, n);
    if (rc) return rc;
    MLPCHK(hipMemcpyAsync(p->Xb, X_host, sizeof(float) * (size_t)n * p->in, hipMemcpyHostToDevice, st));
    MLPCHK(hipMemsetAsync(p->Tb, 0, sizeof(float) * (size_t)n * p->out, st));
    if ((rc = fwd_bwd(p, n, st, nullptr))) return rc;
    MLPCHK(hipMemcpyAsync(Y_host, p->Y, sizeof(float) * (size_t)n * p->out, hipMemcpyDeviceToHost, st));
    MLPCHK(hipStreamSynchronize(st));
    return MGSB_OK;
}

int mlp_train(Mlp* p, const float* X, const float* T, long n, int epochs, int batch, float lr, uint32_t seed,
              float* loss_hist, hipStream_t st) {
    if (n < 1 || batch < 16 || epochs < 1) {
        set_error("mlp_train: need n >= 1, batch >= 16, epochs >= 1");
        return MGSB_ERR_ARG;
    }
    batch = (int)std::min<long>(batch, n);
    int rc = ensure_cap(p, batch);
    if (rc) return rc;
    std::vector<int> perm(n);
    for (long i = 0; i < n; i++) perm[i] = (int)i;
    uint64_t x = 0xD1B54A32D192ED03ull ^ seed;
    auto rnd = [&]() {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return x;
    };
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    std::vector<int> hidx(batch);
    for (int ep = 0; ep < epochs; ep++) {
        for (long i = n - 1; i > 0; i--) std::swap(perm[i], perm[rnd() % (uint64_t)(i + 1)]);
        double acc = 0.0;
        long nbat = 0;
        for (long s0 = 0; s0 + batch <= n; s0 += batch) {
            for (int b = 0; b < batch; b++) hidx[b] = perm[s0 + b];
            MLPCHK(hipMemcpyAsync(p->idx, hidx.data(), sizeof(int) * batch, hipMemcpyHostToDevice, st));
            k_gather_rows<<<(batch + 255) / 256, 256, 0, st>>>(batch, p->idx, X, p->in, T, p->out, p->Xb, p->Tb);
            MLPCHK(hipGetLastError());
            double l = 0.0;
            if ((rc = fwd_bwd(p, batch, st, &l))) return rc;
            acc += l;
            nbat++;
            p->adam_t++;
            const float c1 = 1.f - std::pow(b1, (float)p->adam_t), c2 = 1.f - std::pow(b2, (float)p->adam_t);
            k_adam<<<(p->nparams + 255) / 256, 256, 0, st>>>(p->nparams, p->w, p->g, p->m, p->v, lr, b1, b2, eps, c1,
                                                             c2);
            MLPCHK(hipGetLastError());
        }
        if (loss_hist) loss_hist[ep] = (float)(acc / std::max<long>(1, nbat));
    }
    rc = mlp_pack(p, st);
    if (rc) return rc;
    MLPCHK(hipStreamSynchronize(st));
    return MGSB_OK;
}

}  // namespace mgsb
