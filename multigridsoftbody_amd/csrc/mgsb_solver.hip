// libmgsb device solver: hand-written gfx950 kernels for the projective-
// dynamics (PD) ARAP tetrahedral solver and the host orchestration of its
// step schedules.  See include/mgsb.h for the C ABI.
//
// Data layout in HBM (per level):
//   per vertex : float4 x, x_prev, s, x_last, v   (xyz + pad; 16 B aligned
//                gathers), float4 vc = (m/h^2, diag, fixflag, m)
//   per tet    : int4 tet, int4 slot, 9 SoA streams of Dm^-1 (B[k*T+t]),
//                float vol
//   per (tet, corner) in vertex-sorted order : float4 cf  (corner forces)
//   vertex -> (tet,corner) CSR row pointer rp[V+1]
// The per-tet kernel (one tet per lane) writes its 4 corner forces to their
// CSR slots; the per-vertex kernel sums its contiguous CSR range in tet order
// (deterministic, no atomics) and fuses collision response and the Jacobi /
// under-relaxation / Chebyshev update.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "mgsb_device.h"
#include "mgsb_internal.h"
#include "mgsb_mlp.h"

using namespace mgsb;

#define HIPCHK(expr)                                                                                      \
    do {                                                                                                  \
        hipError_t _e = (expr);                                                                           \
        if (_e != hipSuccess) {                                                                           \
            set_error(std::string("HIP error ") + hipGetErrorString(_e) + " at " #expr);                  \
            return MGSB_ERR_HIP;                                                                          \
        }                                                                                                 \
    } while (0)

namespace {

constexpr int kTetModePolar = 0;  // calculateIF
constexpr int kTetModeAc = 1;     // calTetIFV1: drop |f| < 1e-5
constexpr int kTetModeLinear = 2; // -F(e) AC w, no clamp: linear operator on corrections
constexpr int kTetModeStable = 3; // calculateIF with the deviatoric-invariant polar

constexpr int kVertPoint = 1;     // sphere point DCD + plane
constexpr int kVertTri = 2;       // gather triangle DCD contributions
constexpr int kVertWriteCol = 4;  // write collision flags / depth / energy
constexpr int kVertContactJacobi = 8;  // V-cycle: exact 3x3 block step, no Chebyshev, where a collider acts
constexpr int kVertDcdOnly = 16;       // collision detection only (with kVertWriteCol): positions untouched

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Sum of the xyz of cf[e0..e1) in index order.  Eight loads are issued
// before the adds that consume them, so a vertex waits on memory about once
// per eight entries instead of once per entry; the summation order (and the
// result, bit for bit) is that of the plain loop.  (Measured on the medium
// box: issuing a whole 24-entry row at once, or staging a block's rows in
// LDS with coalesced loads, made the vertex kernels slower.)
__device__ __forceinline__ void gather_sum(const float3* __restrict__ cf, int e0, int e1, float& fx, float& fy,
                                           float& fz) {
    int e = e0;
    for (; e + 8 <= e1; e += 8) {
        float3 q[8];
#pragma unroll
        for (int k = 0; k < 8; k++) q[k] = cf[e + k];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            fx = fx + q[k].x;
            fy = fy + q[k].y;
            fz = fz + q[k].z;
        }
    }
    const int rem = e1 - e;
    float3 q[8];
#pragma unroll
    for (int k = 0; k < 8; k++) q[k] = k < rem ? cf[e + k] : make_float3(0.f, 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < 8; k++)
        if (k < rem) {
            fx = fx + q[k].x;
            fy = fy + q[k].y;
            fz = fz + q[k].z;
        }
}

// Sum of the xyz of cf[e0..e1) by a group of 2^lsh adjacent lanes (sub = the
// lane's index in the group): lane sub adds entries e0 + sub, e0 + sub + L, ...
// (up to eight loads in flight), then the group combines its partials by an
// xor butterfly, which leaves the identical total in every lane of the group
// (a + b == b + a).  lsh = 0 is gather_sum.  A row is usually read in one
// round trip instead of three, and a vertex kernel launches L times the
// waves, which the medium box's 20k vertices need to cover 256 CUs.
__device__ __forceinline__ float3 gather_lanes(const float3* __restrict__ cf, int e0, int e1, int sub, int lsh) {
    float fx = 0.f, fy = 0.f, fz = 0.f;
    if (lsh == 0) {
        gather_sum(cf, e0, e1, fx, fy, fz);
        return make_float3(fx, fy, fz);
    }
    const int L = 1 << lsh;
    int e = e0 + sub;
    for (; e + 7 * L < e1; e += 8 * L) {
        float3 q[8];
#pragma unroll
        for (int k = 0; k < 8; k++) q[k] = cf[e + k * L];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            fx = fx + q[k].x;
            fy = fy + q[k].y;
            fz = fz + q[k].z;
        }
    }
    const int rem = e < e1 ? (e1 - e + L - 1) >> lsh : 0;
    float3 q[8];
#pragma unroll
    for (int k = 0; k < 8; k++) q[k] = k < rem ? cf[e + k * L] : make_float3(0.f, 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < 8; k++)
        if (k < rem) {
            fx = fx + q[k].x;
            fy = fy + q[k].y;
            fz = fz + q[k].z;
        }
    for (int m = 1; m < L; m <<= 1) {
        fx = fx + __shfl_xor(fx, m, 64);
        fy = fy + __shfl_xor(fy, m, 64);
        fz = fz + __shfl_xor(fz, m, 64);
    }
    return make_float3(fx, fy, fz);
}

// (P e)_i: the coarse vector e interpolated at fine vertex i through its
// barycentric binding in a coarse tet (ids, weights).
__device__ __forceinline__ float3 prolong_at(int i, const float4* __restrict__ e, const int4* __restrict__ ids,
                                             const float4* __restrict__ w) {
    const int4 id = ids[i];
    const float4 ww = w[i];
    const float4 a = e[id.x], b = e[id.y], c = e[id.z], d = e[id.w];
    return make_float3(ww.x * a.x + ww.y * b.x + ww.z * c.x + ww.w * d.x, ww.x * a.y + ww.y * b.y + ww.z * c.y + ww.w * d.y,
                       ww.x * a.z + ww.y * b.z + ww.z * c.z + ww.w * d.z);
}

// prolong_at with the binding (coarse tet corners, weights) already loaded:
// the same sums in the same order.
__device__ __forceinline__ float3 prolong_w(int4 id, float4 ww, const float4* __restrict__ e) {
    const float4 a = e[id.x], b = e[id.y], c = e[id.z], d = e[id.w];
    return make_float3(ww.x * a.x + ww.y * b.x + ww.z * c.x + ww.w * d.x, ww.x * a.y + ww.y * b.y + ww.z * c.y + ww.w * d.y,
                       ww.x * a.z + ww.y * b.z + ww.z * c.z + ww.w * d.z);
}

// calculateST (PDSolver.cu:72-116).  Block 0 / lane 0 first advances sphere
// centres by vel*dt (UpdateCollider before Step, Simulator.cpp:146-149).
// split: x and x_prev take the in-step correction u (0, or with U1/U2 the
// temporal extrapolation of the last two steps' corrections).
__global__ void __launch_bounds__(256) k_st(int V, float4* __restrict__ x, float4* __restrict__ v,
                                            float4* __restrict__ xlast, float4* __restrict__ s,
                                            float4* __restrict__ xprev, const float4* __restrict__ vc, float dt,
                                            float damp, float gx, float gy, float gz, DevColliders* col, int advance,
                                            int split, const float4* __restrict__ U1, const float4* __restrict__ U2,
                                            const float4* __restrict__ epred, const int4* __restrict__ iids,
                                            const float4* __restrict__ iw) {
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (advance && blockIdx.x == 0 && threadIdx.x == 0) {
        for (int q = 0; q < col->nsph; q++) {
            float4 c = col->sph[q];
            const float4 w = col->vel[q];
            c.x = c.x + w.x * dt;
            c.y = c.y + w.y * dt;
            c.z = c.z + w.z * dt;
            col->sph[q] = c;
        }
    }
    if (i >= V) return;
    float4 p = x[i], u = v[i];
    const float fix = vc[i].z;
    xlast[i] = p;
    const float df = damp * fix;
    u.x = u.x * df;
    u.y = u.y * df;
    u.z = u.z * df;
    u.x = u.x + gx * dt * fix;
    u.y = u.y + gy * dt * fix;
    u.z = u.z + gz * dt * fix;
    p.x = p.x + u.x * dt * fix;
    p.y = p.y + u.y * dt * fix;
    p.z = p.z + u.z * dt * fix;
    v[i] = u;
    s[i] = p;
    if (split) {  // x and x_prev hold the in-step correction u = x - s
        // U1 != nullptr: temporal warm start u_t = fix (2 u_{n-1} - u_{n-2});
        // epred != nullptr: plus the prolongated learned coarse correction
        float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
        if (U1) {
            const float4 a = U1[i], b = U2[i];
            q = make_float4(fix * (2.f * a.x - b.x), fix * (2.f * a.y - b.y), fix * (2.f * a.z - b.z), 0.f);
        }
        if (epred) {
            const float3 e = prolong_at(i, epred, iids, iw);
            q = make_float4(q.x + fix * e.x, q.y + fix * e.y, q.z + fix * e.z, 0.f);
        }
        x[i] = q;
        xprev[i] = q;
    } else {
        x[i] = p;
        xprev[i] = p;
    }
}

struct VState;
__device__ void vstate_step_done(VState* st);

// End of a split-position step: x = s + u (and x_prev = s + u_prev), then
// calculateV; block 0 folds the step's outcome into the run aggregates.
// D1 != nullptr (learned schedule): D2 = D1, D1 = what the step's correction
// added to its temporal extrapolation, u - vfix (2 u1 - u2) (u when
// temporal == 0), per fine vertex; the learned stage injects both at the
// coarse vertices as its last two coarse deltas.
__global__ void __launch_bounds__(256) k_finish_split(int V, float4* __restrict__ x, float4* __restrict__ xprev,
                                                      const float4* __restrict__ s, const float4* __restrict__ xlast,
                                                      float4* __restrict__ v, float dt, VState* vs,
                                                      float4* __restrict__ U1, float4* __restrict__ U2,
                                                      const float4* __restrict__ vc, float4* __restrict__ D1,
                                                      float4* __restrict__ D2, int temporal,
                                                      const int* __restrict__ skip) {
    // the skip flag first (a skipped launch loads nothing), then the vertex's
    // streams in one round trip (lanes past V read vertex V - 1)
    if (load_flag(skip)) return;
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int ic = min(i, V - 1);
    const float4 u = x[ic], up = xprev[ic], st = s[ic], q = xlast[ic];
    float4 u1 = make_float4(0.f, 0.f, 0.f, 0.f), u2 = u1, d1 = u1;
    float f = 0.f;
    if (U1) {
        u1 = U1[ic];
        u2 = U2[ic];
        f = vc[ic].z;
    }
    if (D1) d1 = D1[ic];
    if (vs && blockIdx.x == 0 && threadIdx.x == 0) vstate_step_done(vs);
    if (i >= V) return;
    if (D1) {
        float4 a = u;
        if (temporal) {
            a.x = a.x - f * (2.f * u1.x - u2.x);
            a.y = a.y - f * (2.f * u1.y - u2.y);
            a.z = a.z - f * (2.f * u1.z - u2.z);
        }
        D2[i] = d1;
        D1[i] = make_float4(a.x, a.y, a.z, 0.f);
    }
    if (U1) {  // corrections of the last two steps (temporal warm start)
        U2[i] = u1;
        U1[i] = make_float4(u.x, u.y, u.z, 0.f);
    }
    const float4 p = make_float4(st.x + u.x, st.y + u.y, st.z + u.z, 0.f);
    x[i] = p;
    xprev[i] = make_float4(st.x + up.x, st.y + up.y, st.z + up.z, 0.f);
    v[i] = make_float4((p.x - q.x) / dt, (p.y - q.y) / dt, (p.z - q.z) / dt, 0.f);
}

// calculateV (PDSolver.cu:245-252).
__global__ void __launch_bounds__(256) k_vel(int V, const float4* __restrict__ x, const float4* __restrict__ xlast,
                                             float4* __restrict__ v, float dt) {
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const float4 p = x[i], q = xlast[i];
    v[i] = make_float4((p.x - q.x) / dt, (p.y - q.y) / dt, (p.z - q.z) / dt, 0.f);
}

// DCDByTriangle_sphere_V2 (PDSolver.cu:612-680) for one triangle and one
// sphere; returns the response centre - hit in mv, true on contact.
__device__ __forceinline__ bool tri_sphere_hit(float3 A, float3 B, float3 C, float3 n, float3 center, float radius,
                                               float3& mv) {
    const float distance = fabsf(dot3(n, sub3(center, A)));
    if (distance > radius) return false;
    const float3 projected = sub3(center, mul3(n, dot3(n, sub3(center, A))));
    float3 hit;
    if (point_in_triangle(projected, A, B, C)) {
        const float sign = dot3(n, sub3(center, A)) < 0 ? -1.f : 1.f;
        hit = add3(center, mul3(mul3(n, radius - distance), sign));
        mv = sub3(center, hit);
        return true;
    }
    const float3 tv[3] = {A, B, C};
    for (int i = 0; i < 3; i++) {
        const float3 Bg = tv[i], En = tv[(i + 1) % 3];
        float3 ed = sub3(En, Bg);
        const float l = len3(ed);
        ed = div3(ed, len3(ed));
        const float3 bc = sub3(center, Bg);
        const float t = dot3(bc, ed);
        if (t < 0 || t > l) continue;
        const float3 closest = add3(Bg, mul3(ed, t));
        float3 radial = sub3(center, closest);
        const float dist = len3(radial);
        if (dist >= radius) continue;
        const float3 Oth = tv[(i + 2) % 3];
        if (dot3(cross3(ed, sub3(Oth, Bg)), cross3(ed, sub3(projected, Bg))) >= 0) continue;
        radial = div3(radial, len3(radial));
        hit = add3(closest, mul3(radial, radius));
        mv = sub3(center, hit);
        return true;
    }
    for (int i = 0; i < 3; i++) {
        float3 VO = sub3(center, tv[i]);
        if (len3(VO) > radius) continue;
        VO = div3(VO, len3(VO));
        hit = add3(tv[i], mul3(VO, radius));
        mv = sub3(center, hit);
        return true;
    }
    return false;
}

// The stable mode's products (it has no counterpart in the reference, so it
// keeps no reference operation order): F = D Dm^-1 and the corner forces
// (R - F) AC with their multiply-adds fused, and the weight vol k applied as
// one factor after the NaN -> 0 and [-10, 10] clamp.
__device__ __forceinline__ void stable_products_F(const float (&D)[9], const float (&b)[9], float (&F)[9]) {
#pragma clang fp contract(fast)
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) F[r * 3 + c] = D[r * 3 + 0] * b[c] + D[r * 3 + 1] * b[3 + c] + D[r * 3 + 2] * b[6 + c];
}
__device__ __forceinline__ void stable_forces(const float (&RF)[9], const float (&AC)[12], float vk, float (&fo)[12]) {
#pragma clang fp contract(fast)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            float acc = RF[i * 3 + 0] * AC[0 * 4 + j] + RF[i * 3 + 1] * AC[1 * 4 + j] + RF[i * 3 + 2] * AC[2 * 4 + j];
            acc = isnan(acc) ? 0.f : fminf(fmaxf(acc, -10.f), 10.f);
            fo[j * 3 + i] = acc * vk;
        }
}

// Per-tet local step of tet g: F = Ds Dm^-1, R - F by the Franca polar
// decomposition, P = (R - F) AC with AC rebuilt from Dm^-1, NaN -> 0, clamp to
// [-10, 10], times vol * k (calculateIF, PDSolver.cu:128-183; calTetIFV1's
// |f| < 1e-5 drop in Ac mode, PDSolverAc.cu:630-662).  Corner j's force goes
// to fo[3j..3j+2].  sref != nullptr: split positions (x holds u, see below).
// The local step from the corners' positions p0..p3 (and, split, their s0..s3),
// Dm^-1 in b, rest volume vt and stiffness kk.
template <int MODE>
__device__ __forceinline__ void tet_forces_at(const float (&b)[9], float vt, float kk, float4 p0, float4 p1,
                                              float4 p2, float4 p3, bool split, float4 s0, float4 s1, float4 s2,
                                              float4 s3, float (&fo)[12]) {
    float D[9];
    D[0] = p1.x - p0.x; D[1] = p2.x - p0.x; D[2] = p3.x - p0.x;
    D[3] = p1.y - p0.y; D[4] = p2.y - p0.y; D[5] = p3.y - p0.y;
    D[6] = p1.z - p0.z; D[7] = p2.z - p0.z; D[8] = p3.z - p0.z;
    // Split positions x = s + u (V-cycle steps, sref = s): x holds u and the
    // edge matrix is Ds(s) + Ds(u), each from its own corner differences, so
    // the small in-step correction keeps its low bits; at u = 0 it is Ds(s)
    // exactly, the same F(s) in every pass of the step.  One product with
    // Dm^-1 instead of F(s) + Ds(u) Dm^-1 (45 fewer instructions per tet).
    if (split) {
        D[0] = (s1.x - s0.x) + D[0]; D[1] = (s2.x - s0.x) + D[1]; D[2] = (s3.x - s0.x) + D[2];
        D[3] = (s1.y - s0.y) + D[3]; D[4] = (s2.y - s0.y) + D[4]; D[5] = (s3.y - s0.y) + D[5];
        D[6] = (s1.z - s0.z) + D[6]; D[7] = (s2.z - s0.z) + D[7]; D[8] = (s3.z - s0.z) + D[8];
    }
    float F[9];
    if (MODE == kTetModeStable) {
        stable_products_F(D, b, F);
    } else {
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 3; c++)
                F[r * 3 + c] = D[r * 3 + 0] * b[c] + D[r * 3 + 1] * b[3 + c] + D[r * 3 + 2] * b[6 + c];
    }
    float RF[9];
    if (MODE == kTetModeLinear) {
#pragma unroll
        for (int q = 0; q < 9; q++) RF[q] = -F[q];
    } else if (MODE == kTetModeStable) {
        polar_rf_stable(F, RF);
    } else {
        polar_rf<MODE == kTetModePolar>(F, RF);
    }
    // AC[r][0] = -B[0][r] - B[1][r] - B[2][r], AC[r][j] = B[j-1][r]
    float AC[12];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        AC[r * 4 + 0] = -b[r] - b[3 + r] - b[6 + r];
        AC[r * 4 + 1] = b[r];
        AC[r * 4 + 2] = b[3 + r];
        AC[r * 4 + 3] = b[6 + r];
    }
    if (MODE == kTetModeStable) {
        stable_forces(RF, AC, vt * kk, fo);
        return;
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            float acc = RF[i * 3 + 0] * AC[0 * 4 + j] + RF[i * 3 + 1] * AC[1 * 4 + j] + RF[i * 3 + 2] * AC[2 * 4 + j];
            if (MODE != kTetModeLinear) {
                // NaN -> 0, then clamp to [-10, 10] (one v_med3_f32)
                acc = isnan(acc) ? 0.f : fminf(fmaxf(acc, -10.f), 10.f);
            }
            float r = acc * vt * kk;
            // |r| < 1e-5 against the double literal is exactly |r| <= 1e-5f
            if (MODE == kTetModeAc && fabsf(r) <= 1e-5f) r = 0.f;
            fo[j * 3 + i] = r;
        }
    }
}

// The per-tet streams of tet g: Dm^-1 (B), rest volume, ARAP weight.
struct TetIn {
    float b[9];
    float vt, kk;
};
__device__ __forceinline__ TetIn tet_streams(int g, int T, const float* __restrict__ B, const float* __restrict__ vol,
                                             float k, const float* __restrict__ tk) {
    TetIn in;
#pragma unroll
    for (int q = 0; q < 9; q++) in.b[q] = B[(size_t)q * T + g];
    in.vt = vol[g];
    in.kk = tk ? tk[g] : k;  // per-body stiffness when several bodies have their own E
    return in;
}
// The corner gathers of x (and s) and the local step's forces.
template <int MODE>
__device__ __forceinline__ void tet_forces_in(const TetIn& in, int4 id, const float4* __restrict__ x,
                                              const float4* __restrict__ sref, float (&fo)[12]) {
    const float4 p0 = x[id.x], p1 = x[id.y], p2 = x[id.z], p3 = x[id.w];
    float4 s0 = make_float4(0.f, 0.f, 0.f, 0.f), s1 = s0, s2 = s0, s3 = s0;
    if (sref) {
        s0 = sref[id.x];
        s1 = sref[id.y];
        s2 = sref[id.z];
        s3 = sref[id.w];
    }
    tet_forces_at<MODE>(in.b, in.vt, in.kk, p0, p1, p2, p3, sref != nullptr, s0, s1, s2, s3, fo);
}
template <int MODE>
__device__ __forceinline__ void tet_local_forces(int g, int T, int4 id, const float4* __restrict__ x,
                                                 const float* __restrict__ B, const float* __restrict__ vol,
                                                 float k, const float4* __restrict__ sref,
                                                 const float* __restrict__ tk, float (&fo)[12]) {
    // every load is issued up front: the per-tet streams and the index (id,
    // loaded by the caller) in the first round trip, the corner gathers of x
    // (and s) in the second
    const TetIn in = tet_streams(g, T, B, vol, k, tk);
    tet_forces_in<MODE>(in, id, x, sref, fo);
}

template <int MODE>
__device__ __forceinline__ void tet_local_step(int g, int T, const float4* __restrict__ x,
                                               const int4* __restrict__ tet, const float* __restrict__ B,
                                               const float* __restrict__ vol, const int4* __restrict__ slot,
                                               float3* __restrict__ cf, float k, const float4* __restrict__ sref,
                                               const float* __restrict__ tk) {
    const int4 id = tet[g];
    const int4 sl = slot[g];
    float fo[12];
    tet_local_forces<MODE>(g, T, id, x, B, vol, k, sref, tk, fo);
    cf[sl.x] = make_float3(fo[0], fo[1], fo[2]);
    cf[sl.y] = make_float3(fo[3], fo[4], fo[5]);
    cf[sl.z] = make_float3(fo[6], fo[7], fo[8]);
    cf[sl.w] = make_float3(fo[9], fo[10], fo[11]);
}

// One tile of kTileTets consecutive tets per workgroup: the tile's corner
// forces go to LDS at their positions in the tile's vertex-sorted corner
// array (tls), then each of the tile's vertices sums its row there in (tet,
// corner) order and writes one partial to its slot in the vertex-sorted
// (tile, vertex) order (tvs).  The vertex kernels gather those partials, a
// row of ~2 per vertex instead of ~20 corners, and the tet kernel's scattered
// stores shrink by the same factor.  Deterministic: fixed sums, no atomics.
// The tiled kernels read each tet as one int4 whose entries carry the corner's
// vertex id in the low 22 bits and its position in the tile's corner array
// in the high 10 (levels of at most 2^22 vertices run tiled).
struct TileLds {
    float x[4 * kTileTets], y[4 * kTileTets], z[4 * kTileTets];
};
constexpr int kTileIdBits = 22;
constexpr unsigned kTileIdMask = (1u << kTileIdBits) - 1u;
__device__ __forceinline__ int4 tile_ids(int4 q) {
    return make_int4((int)((unsigned)q.x & kTileIdMask), (int)((unsigned)q.y & kTileIdMask),
                     (int)((unsigned)q.z & kTileIdMask), (int)((unsigned)q.w & kTileIdMask));
}

__device__ __forceinline__ void tile_put(TileLds& L, int4 q, const float (&fo)[12]) {
    const int p[4] = {(int)((unsigned)q.x >> kTileIdBits), (int)((unsigned)q.y >> kTileIdBits),
                      (int)((unsigned)q.z >> kTileIdBits), (int)((unsigned)q.w >> kTileIdBits)};
#pragma unroll
    for (int j = 0; j < 4; j++) {
        L.x[p[j]] = fo[j * 3 + 0];
        L.y[p[j]] = fo[j * 3 + 1];
        L.z[p[j]] = fo[j * 3 + 2];
    }
}

// For lane threadIdx.x, its tile vertex's row of the tile's corner array,
// its partial's slot and whether the lane holds a vertex, from the padded
// per-tile rows (tpad[tile * stride + lane]: (slot, row start | row end << 16),
// zeros past the tile's vertex count; a vertex of the tile has at least one
// corner there, so row end > row start marks a live lane).  8 bytes a lane:
// the padding lanes cost no more than the compact rows they replace.  The address depends on the tile
// alone, so the load goes out with the tet streams and the kernel's flags in
// its first round trip, instead of after a load of the tile's vertex range.
// Loaded before the local step, so tile_sum's first pass reads no index
// after the barrier.
struct TileRow {
    int r0, r1, slot, live;
};
// A tiled level's per-tile metadata as the tiled kernels take it.
struct TileMeta {
    const int2* tpad;  // padded rows (tile_row), stride entries per tile
    const int *tvo, *tvs, *tlr;
    int stride;  // entries per tile in tpad: the largest tile's vertex count rounded up to 16, at most kTileTets
    int over;    // 1 if a tile holds more than kTileTets vertices
};
__device__ __forceinline__ TileRow tile_row(int tile, int stride, const int2* __restrict__ tpad) {
    // lanes past the stride read the tile's last entry and drop it: no branch
    // around the load, which would wait for it at the branch's end
    const int l = threadIdx.x;
    const int2 e = tpad[(size_t)tile * stride + min(l, stride - 1)];
    const int r0 = e.y & 0xffff, r1 = (int)((unsigned)e.y >> 16);
    return TileRow{r0, r1, e.x, (l < stride && r1 > r0) ? 1 : 0};
}
__device__ __forceinline__ float3 tile_row_sum(const TileLds& L, int r0, int r1) {
    // in row order; four entries' LDS reads in flight before their sums
    float fx = 0.f, fy = 0.f, fz = 0.f;
    int e = r0;
    for (; e + 4 <= r1; e += 4) {
        float x[4], y[4], z[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            x[k] = L.x[e + k];
            y[k] = L.y[e + k];
            z[k] = L.z[e + k];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            fx = fx + x[k];
            fy = fy + y[k];
            fz = fz + z[k];
        }
    }
    for (; e < r1; e++) {
        fx = fx + L.x[e];
        fy = fy + L.y[e];
        fz = fz + L.z[e];
    }
    return make_float3(fx, fy, fz);
}
// over: the level has a tile of more than kTileTets vertices (none on the
// generated meshes); its lanes then also take vertices kTileTets, ... of the
// tile's list (tvo, tvs, tlr).
__device__ __forceinline__ void tile_sum(const TileLds& L, const TileRow& t, int tile, int over,
                                         const int* __restrict__ tvo, const int* __restrict__ tvs,
                                         const int* __restrict__ tlr, float3* __restrict__ pcf) {
    if (t.live) pcf[t.slot] = tile_row_sum(L, t.r0, t.r1);
    if (!over) return;
    const int v0 = tvo[tile], nv = tvo[tile + 1] - v0;
    const int* lr = tlr + v0 + tile;
    for (int l = threadIdx.x + kTileTets; l < nv; l += kTileTets)
        pcf[tvs[v0 + l]] = tile_row_sum(L, lr[l], lr[l + 1]);
}

// Triangle-sphere DCD of surface triangle t (PDSolver.cu:612-680): writes
// (k*(centre-hit), k) summed over the spheres to each of its corners' slots.
__device__ __forceinline__ void tri_lane(int t, int ntri, const float4* __restrict__ x, const int4* __restrict__ tri,
                                         const int4* __restrict__ tslot, float4* __restrict__ tcf,
                                         const DevColliders* __restrict__ col) {
    if (t >= ntri) return;
    const int4 id = tri[t];
    const float3 A = f3(x[id.x]), Bq = f3(x[id.y]), Cq = f3(x[id.z]);
    float3 n = cross3(sub3(Bq, A), sub3(Cq, A));
    n = div3(n, len3(n));
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    const float kc = col->kcol;
    for (int q = 0; q < col->nsph; q++) {
        if (col->vel[q].w == 0.f) continue;
        const float4 sp = col->sph[q];
        float3 mv;
        if (tri_sphere_hit(A, Bq, Cq, n, make_float3(sp.x, sp.y, sp.z), sp.w, mv)) {
            acc.x = acc.x + kc * mv.x;
            acc.y = acc.y + kc * mv.y;
            acc.z = acc.z + kc * mv.z;
            acc.w = acc.w + kc;
        }
    }
    const int4 sl = tslot[t];
    tcf[sl.x] = acc;
    tcf[sl.y] = acc;
    tcf[sl.z] = acc;
}

// Per-tet local step kernel, one tet per lane (tet_local_step).  TRI: lanes
// T..T+S-1 (when ntri > 0) run the triangle-sphere DCD (PDSolver.cu:612-680)
// and write (k*(centre-hit), k) per corner.  The instances without it hold
// 55 VGPRs instead of 78: 8 waves per SIMD instead of 6.
template <int MODE, bool TRI>
__global__ void __launch_bounds__(256) k_tet(int T, const float4* __restrict__ x, const int4* __restrict__ tet,
                                             const float* __restrict__ B, const float* __restrict__ vol,
                                             const int4* __restrict__ slot, float3* __restrict__ cf, float k,
                                             int ntri, const int4* __restrict__ tri, const int4* __restrict__ tslot,
                                             float4* __restrict__ tcf, const DevColliders* __restrict__ col,
                                             const int* __restrict__ skip, const float4* __restrict__ sref,
                                             const float* __restrict__ tk) {
    if (skip && *skip) return;
    const int g = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (g >= T) {
        if (TRI) tri_lane(g - T, ntri, x, tri, tslot, tcf, col);
        return;
    }
    tet_local_step<MODE>(g, T, x, tet, B, vol, slot, cf, k, sref, tk);
}

// The tiled form of k_tet (tile_put / tile_sum): workgroups 0..ntiles-1 run
// one tile each (kTileTets lanes), the ones after them the triangle lanes.
// The skip flag is tested before any other load: a skipped launch (the
// recompute after a rejected warm start, skipped on most steps) would
// otherwise stream 56 B per tet for nothing, and on the levels this kernel
// serves (the fused sweeps take the small ones) that costs more than the
// flag's round trip.  The tile's rows and the tet's streams then load in one
// round trip.
template <int MODE, bool TRI>
__global__ void __launch_bounds__(kTileTets) k_tet_tiled(
    int T, int ntiles, const float4* __restrict__ x, const int4* __restrict__ tetp, const float* __restrict__ B,
    const float* __restrict__ vol, TileMeta tm, float3* __restrict__ pcf, float k, int ntri,
    const int4* __restrict__ tri, const int4* __restrict__ tslot, float4* __restrict__ tcf,
    const DevColliders* __restrict__ col, const int* __restrict__ skip, const float4* __restrict__ sref,
    const float* __restrict__ tk) {
    if (load_flag(skip)) return;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    if (blk >= ntiles) {
        if (TRI) tri_lane((blk - ntiles) * kTileTets + threadIdx.x, ntri, x, tri, tslot, tcf, col);
        return;
    }
    __shared__ TileLds L;
    const int g = blk * kTileTets + threadIdx.x;
    const TileRow tr = tile_row(blk, tm.stride, tm.tpad);
    // branch-free loads: the last tile's lanes past T read tet T - 1 (tile_row)
    const int gc = min(g, T - 1);
    const int4 q = tetp[gc];
    const TetIn in = tet_streams(gc, T, B, vol, k, tk);
    if (g < T) {
        float fo[12];
        tet_forces_in<MODE>(in, tile_ids(q), x, sref, fo);
        tile_put(L, q, fo);
    }
    __syncthreads();
    tile_sum(L, tr, blk, tm.over, tm.tvo, tm.tvs, tm.tlr, pcf);
}

// Start of a temporally warm-started V-cycle step: lanes 0..T-1 evaluate the
// local step at the extrapolation (split, u = x, forces to cf), lanes
// T..2T-1 at the inertial prediction s itself (forces to cf2), so that one
// launch serves both residual norms the start needs.
template <int MODE>
__global__ void __launch_bounds__(256) k_tet_dual(int T, const float4* __restrict__ x, const float4* __restrict__ s,
                                                  const int4* __restrict__ tet, const float* __restrict__ B,
                                                  const float* __restrict__ vol, const int4* __restrict__ slot,
                                                  float3* __restrict__ cf, float3* __restrict__ cf2, float k,
                                                  const float* __restrict__ tk) {
    const int g = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (g < T)
        tet_local_step<MODE>(g, T, x, tet, B, vol, slot, cf, k, s, tk);
    else if (g < 2 * T)
        tet_local_step<MODE>(g - T, T, s, tet, B, vol, slot, cf2, k, nullptr, tk);
}

// The tiled form of k_tet_dual: workgroups 0..ntiles-1 take the tiles at the
// extrapolation (partials to pcf), ntiles..2 ntiles-1 the same tiles at s
// (pcf2).
template <int MODE>
__global__ void __launch_bounds__(kTileTets) k_tet_dual_tiled(
    int T, int ntiles, const float4* __restrict__ x, const float4* __restrict__ s, const int4* __restrict__ tetp,
    const float* __restrict__ B, const float* __restrict__ vol, TileMeta tm, float3* __restrict__ pcf,
    float3* __restrict__ pcf2, float k, const float* __restrict__ tk) {
    const int b0 = xcd_block(blockIdx.x, gridDim.x);
    const bool at_s = b0 >= ntiles;
    const int blk = at_s ? b0 - ntiles : b0;
    __shared__ TileLds L;
    const int g = blk * kTileTets + threadIdx.x;
    const TileRow tr = tile_row(blk, tm.stride, tm.tpad);
    if (g < T) {
        const int4 q = tetp[g];
        float fo[12];
        tet_local_forces<MODE>(g, T, tile_ids(q), at_s ? s : x, B, vol, k, at_s ? nullptr : s, tk, fo);
        tile_put(L, q, fo);
    }
    __syncthreads();
    tile_sum(L, tr, blk, tm.over, tm.tvo, tm.tvs, tm.tlr, at_s ? pcf2 : pcf);
}

// calculatePOS's under-relaxation (next - pos) * 0.6 + pos (PDSolver.cu:
// 219-221): the float difference times the double literal 0.6 plus the float
// position, evaluated in double and rounded to float once.  Here the
// constant is the double word relax_hi + relax_lo (two floats, 48 bits), the
// product is a TwoProduct and the sum a TwoSum, so the result is the
// correctly rounded float of the double expression but for a tie within
// about 2^-48 |next - pos| of a rounding boundary.
__device__ __forceinline__ float relax_step(float d, float pos, float relax_hi, float relax_lo) {
    float ph, pe;
    two_prod(d, relax_hi, ph, pe);
    pe = __builtin_fmaf(d, relax_lo, pe);
    float sh, se;
    two_sum(ph, pos, sh, se);
    return sh + (se + pe);
}

// Vertex update of vertex i from its summed corner forces fg: collision
// response, then calculatePOS (PDSolver.cu:192-235): Jacobi step,
// under-relaxation, Chebyshev extrapolation with omega.  Sphere point DCD is
// DCDByPoint_sphere (PDSolver.cu:277-331) including its 1.5x radius and t/2.
// p = x (or the correction u when split), pp = x_prev, st = s, c = (m/h^2,
// diag, fix, m).  Returns the new position (p itself with kVertDcdOnly) and
// the vertex's collision flag, depth and energy.
__device__ __forceinline__ float4 vertex_step(float4 p, float4 pp, float4 st, float4 c, float3 fg, int i,
                                              const int* __restrict__ trp, const float4* __restrict__ tcf,
                                              const DevColliders* __restrict__ col, float omega, float relax_hi,
                                              float relax_lo, int flags, int split, int& flag, float& depth,
                                              float& energy) {
    const float fx = fg.x, fy = fg.y, fz = fg.z;
    // position seen by the colliders, and s - x of the inertial term
    const float4 pa = split ? make_float4(st.x + p.x, st.y + p.y, st.z + p.z, 0.f) : p;
    const float sx = split ? -p.x : st.x - p.x, sy = split ? -p.y : st.y - p.y, sz = split ? -p.z : st.z - p.z;
    float cdx = 0.f, cdy = 0.f, cdz = 0.f, cfx = 0.f, cfy = 0.f, cfz = 0.f;
    float cxy = 0.f, cxz = 0.f, cyz = 0.f;  // off-diagonals of the contacts' k n n^T
    flag = 0;
    depth = 0.f;
    energy = 0.f;
    if (flags & kVertTri) {
        const int t0 = trp[i], t1 = trp[i + 1];
        for (int e = t0; e < t1; e++) {
            const float4 q = tcf[e];
            if (q.w != 0.f) {
                flag += 1;
                cdx = cdx + q.w;
                cdy = cdy + q.w;
                cdz = cdz + q.w;
                cfx = cfx + q.x;
                cfy = cfy + q.y;
                cfz = cfz + q.z;
            }
        }
    }
    if (flags & kVertPoint) {
        const float kc = col->kcol;
        const float3 P = f3(pa);
        for (int q = 0; q < col->nsph; q++) {
            if (col->vel[q].w == 0.f) continue;
            const float4 sp = col->sph[q];
            const float3 center = make_float3(sp.x, sp.y, sp.z);
            // split: offsets from the centre as (s - c) + u, so the in-step
            // correction keeps its low bits against the 1e6-scale penalty
            const float3 cp = split ? add3(sub3(f3(st), center), f3(p)) : sub3(P, center);
            const float3 dir = div3(cp, len3(cp));
            const float distance = len3(cp);
            const float oldR = sp.w;
            const float r = sp.w * 1.5f;
            if (distance < r) {
                const float bq = 2 * dot3(cp, dir);
                const float cq = dot3(cp, cp) - r * r;
                float t = (-bq + sqrtf(bq * bq - 4.f * cq)) / 2.f;
                if (t > 0.5f * oldR) {
                    flag = 1;
                    depth = t - 0.5f * oldR;
                }
                t = t / 2.0f;
                const float3 colP = add3(P, mul3(dir, t));
                const float3 ccp = split ? add3(cp, mul3(dir, t)) : sub3(colP, center);
                const float3 n = div3(ccp, len3(ccp));
                const float nd = dot3(n, dir);
                cdx = cdx + kc * n.x * n.x;
                cdy = cdy + kc * n.y * n.y;
                cdz = cdz + kc * n.z * n.z;
                cxy = cxy + kc * n.x * n.y;
                cxz = cxz + kc * n.x * n.z;
                cyz = cyz + kc * n.y * n.z;
                cfx = cfx + kc * t * nd * n.x;
                cfy = cfy + kc * t * nd * n.y;
                cfz = cfz + kc * t * nd * n.z;
                const float ee = split ? -t * nd : dot3(sub3(P, colP), n);
                energy = energy + kc * (ee * ee);
            }
        }
        if (col->plane_on) {
            const float4 pl = col->plane;
            const float3 n = make_float3(pl.x, pl.y, pl.z);
            const float sd = split ? (dot3(n, f3(st)) - pl.w) + dot3(n, f3(p)) : dot3(n, P) - pl.w;
            if (sd < 0.f) {
                const float dp = -sd;
                const float kp = col->kplane;
                flag = 1;
                if (dp > depth) depth = dp;
                cdx = cdx + kp * n.x * n.x;
                cdy = cdy + kp * n.y * n.y;
                cdz = cdz + kp * n.z * n.z;
                cxy = cxy + kp * n.x * n.y;
                cxz = cxz + kp * n.x * n.z;
                cyz = cyz + kp * n.y * n.z;
                cfx = cfx + kp * dp * n.x;
                cfy = cfy + kp * dp * n.y;
                cfz = cfz + kp * dp * n.z;
                energy = energy + kp * (dp * dp);
            }
        }
    }
    if (flags & kVertDcdOnly) return p;
    const float dc = c.x, dg = c.y, fix = c.z;
    if ((flags & kVertContactJacobi) && (cdx + cdy + cdz) > 0.f) {
        // V-cycle sweeps, a vertex a collider acts on: the Jacobi step solves
        // its 3x3 block (a I + K) d = g exactly, K = sum of the contacts'
        // k n n^T (the per-axis diagonal k n_c^2 also pins the directions
        // tangent to an oblique contact), with no Chebyshev extrapolation
        // (that of the stiff penalty direction kept the contact rim from
        // converging).  Adjugate inverse of the SPD block.
        const float a = dg + dc;
        const float gx = dc * sx + (fx + cfx), gy = dc * sy + (fy + cfy), gz = dc * sz + (fz + cfz);
        const float m00 = a + cdx, m11 = a + cdy, m22 = a + cdz;
        const float c00 = m11 * m22 - cyz * cyz, c01 = cxz * cyz - cxy * m22, c02 = cxy * cyz - cxz * m11;
        const float c11 = m00 * m22 - cxz * cxz, c12 = cxy * cxz - m00 * cyz, c22 = m00 * m11 - cxy * cxy;
        const float det = m00 * c00 + cxy * c01 + cxz * c02;
        const float dx = (c00 * gx + c01 * gy + c02 * gz) / det;
        const float dy = (c01 * gx + c11 * gy + c12 * gz) / det;
        const float dz = (c02 * gx + c12 * gy + c22 * gz) / det;
        const float bx = relax_step(dx * fix + p.x - p.x, p.x, relax_hi, relax_lo);
        const float by = relax_step(dy * fix + p.y - p.y, p.y, relax_hi, relax_lo);
        const float bz = relax_step(dz * fix + p.z - p.z, p.z, relax_hi, relax_lo);
        return make_float4(bx, by, bz, 0.f);
    }
    float nx = (dc * sx + (fx + cfx)) / (dg + cdx + dc) * fix + p.x;
    float ny = (dc * sy + (fy + cfy)) / (dg + cdy + dc) * fix + p.y;
    float nz = (dc * sz + (fz + cfz)) / (dg + cdz + dc) * fix + p.z;
    nx = relax_step(nx - p.x, p.x, relax_hi, relax_lo);
    ny = relax_step(ny - p.y, p.y, relax_hi, relax_lo);
    nz = relax_step(nz - p.z, p.z, relax_hi, relax_lo);
    nx = omega * (nx - pp.x) + pp.x;
    ny = omega * (ny - pp.y) + pp.y;
    nz = omega * (nz - pp.z) + pp.z;
    return make_float4(nx, ny, nz, 0.f);
}


// Per-vertex update kernel: gather the CSR range of corner forces (tet order,
// or the tiles' partials), then vertex_step.
__global__ void __launch_bounds__(256) k_vert(int V, float4* __restrict__ x, float4* __restrict__ xprev,
                                              const float4* __restrict__ s, const float4* __restrict__ vc,
                                              const int* __restrict__ rp, const float3* __restrict__ cf,
                                              const int* __restrict__ trp, const float4* __restrict__ tcf,
                                              const DevColliders* __restrict__ col, float omega, float relax_hi,
                                              float relax_lo, int flags, int* __restrict__ cflag, float* __restrict__ cdepth,
                                              float* __restrict__ cenergy, const int* __restrict__ skip, int split,
                                              int lsh, const int* __restrict__ calm) {
    // the skip flag first (a skipped launch loads nothing), then the stall
    // flag, the per-vertex streams and the row bounds in one round trip
    // (lanes past V read vertex V - 1: no branch around the loads)
    if (load_flag(skip)) return;
    const int cm = load_flag(calm);
    // 2^lsh lanes per vertex gather its row (gather_lanes); lane 0 finishes
    const int g = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int i = g >> lsh, sub = g & ((1 << lsh) - 1);
    const int ic = min(i, V - 1);
    const float4 p = x[ic];  // x, or the correction u when split
    const float4 st = s[ic];
    const float4 c = vc[ic];  // (m/h^2, diag, fix, m)
    const float4 pp = xprev[ic];
    const int e0 = rp[ic], e1 = rp[ic + 1];
    if (i >= V) return;
    if (cm) omega = 1.0f;  // stall guard: plain (damped) Jacobi for the rest of the step
    const float3 fg = gather_lanes(cf, e0, e1, sub, lsh);
    if (sub) return;
    int flag;
    float depth, energy;
    const float4 q = vertex_step(p, pp, st, c, fg, i, trp, tcf, col, omega, relax_hi, relax_lo, flags, split, flag,
                                 depth, energy);
    if (flags & kVertWriteCol) {
        cflag[i] = flag;
        cdepth[i] = depth;
        cenergy[i] = energy;
    }
    if (flags & kVertDcdOnly) return;
    xprev[i] = p;
    x[i] = q;
}

// One sweep of the V-cycle's fine smoother in one launch (tiled layout,
// split positions).  Phase 1, over the tile's vertices: the vertex update of
// k_vert from the previous local step's partials pin (JACOBI), or the coarse
// correction of k_prolong_add (PROLONG), into LDS with the vertex's s; the
// vertex's first tile (its partial leads its row of prp) stores it to xout.
// Phase 2, over the tile's tets: the local step at the new positions read
// from LDS, its partials to pout as k_tet_tiled writes them.  A vertex shared
// by several tiles is updated by each of them from the same inputs in the
// same order, so every copy equals the owner's.  Positions rotate over three
// buffers (xcur and xprv read, xout written) and partials over two, so no
// launch reads what it writes.  tetq: per corner, its tile-local vertex index
// (low 16 bits) and its position in the tile's corner array (high 16); tvqp:
// per (tile, vertex), padded as the tile rows (tile_row), the vertex, its row
// of partials and whether the tile owns it.
constexpr int kFuseJacobi = 0, kFuseProlong = 1;
// tvqp's second word: the vertex's first partial (low kTvqRowBits) and its
// count of partials (the rest)
constexpr int kTvqRowBits = 26;
constexpr int kTvqRowMask = (1 << kTvqRowBits) - 1;
constexpr int kFuseMaxNv = 256;  // vertices of a tile the fused sweep holds in LDS
static_assert(kFuseMaxNv <= kTileTets, "one lane per tile vertex: the sums' rows load with the vertex");
template <int MODE, int PHASE>
__global__ void __launch_bounds__(kTileTets) k_sweep_fused(
    int T, const float4* __restrict__ xcur, const float4* __restrict__ xprv, float4* __restrict__ xout,
    const float4* __restrict__ s, const float4* __restrict__ vc, const float3* __restrict__ pin, TileMeta tm,
    const int2* __restrict__ tvqp, const int4* __restrict__ tetq,
    const float* __restrict__ B, const float* __restrict__ vol, float k, const float* __restrict__ tk,
    float3* __restrict__ pout, const DevColliders* __restrict__ col, int flags, float omega, float relax_hi,
    float relax_lo, int* __restrict__ cflag, float* __restrict__ cdepth, float* __restrict__ cenergy,
    const float4* __restrict__ ec, const int4* __restrict__ tiid, const float4* __restrict__ tiw,
    const float* __restrict__ amask, const int* __restrict__ skip, const int* __restrict__ calm,
    const int* __restrict__ jrev, const float4* __restrict__ rs) {
    // jrev, rs (PROLONG after k_restrict_judged): when the judge rejected the
    // warm start, the prolongation starts from u = 0 with the active set at s
    // (rs.w) in place of x and the mask at the warm start.
    // First round trip: the flags, the tile's rows (padded per tile, so their
    // addresses need no earlier load), with PROLONG the vertices' bindings
    // into the coarse level (tiid, tiw, padded the same way), and the tet's
    // streams
    const int sk = load_flag(skip);
    const int cm = load_flag(calm);
    const int rv = load_flag(jrev);
    __shared__ float4 lu[kFuseMaxNv], ls[kFuseMaxNv];  // the tile's new positions and s
    __shared__ TileLds L;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    // lane l takes tile vertex l in both phases (a tile has at most
    // kFuseMaxNv vertices, checked at set-up)
    const int l = threadIdx.x;
    const TileRow tr = tile_row(blk, tm.stride, tm.tpad);
    // (vertex | owner << 31, row of partials start | count << 26), padded as tpad
    const int2 tq = tvqp[(size_t)blk * tm.stride + min(l, tm.stride - 1)];
    int4 bid = make_int4(0, 0, 0, 0);
    float4 bw = make_float4(0.f, 0.f, 0.f, 0.f);
    if (PHASE == kFuseProlong) {
        bid = tiid[(size_t)blk * tm.stride + min(l, tm.stride - 1)];
        bw = tiw[(size_t)blk * tm.stride + min(l, tm.stride - 1)];
    }
    // the tet's streams; the last tile's lanes past T read tet T - 1 and
    // drop its forces (branch-free loads, as tile_row)
    const int g = blk * kTileTets + l;
    const int gc = min(g, T - 1);
    const int4 q = tetq[gc];
    float b[9];
#pragma unroll
    for (int c = 0; c < 9; c++) b[c] = B[(size_t)c * T + gc];
    const float vt = vol[gc];
    const float kk = tk ? tk[gc] : k;
    if (sk) return;  // the same flag for the whole workgroup
    if (cm) omega = 1.0f;  // stall guard (k_vert)
    if (tr.live) {
        // second round trip: the vertex's streams and its row of partials
        const int v = tq.x & 0x7fffffff;
        const int r0 = tq.y & kTvqRowMask, r1 = r0 + (int)((unsigned)tq.y >> kTvqRowBits);
        const bool owner = tq.x < 0;
        const float4 st = s[v];
        float4 pn;
        if (PHASE == kFuseProlong) {
            const float3 e = prolong_w(bid, bw, ec);
            const float m = amask ? vc[v].z * (rv ? rs[v].w : amask[v]) : vc[v].z;
            pn = rv ? make_float4(0.f, 0.f, 0.f, 0.f) : xcur[v];
            pn.x = pn.x + m * e.x;
            pn.y = pn.y + m * e.y;
            pn.z = pn.z + m * e.z;
        } else {
            const float4 p = xcur[v], pp = xprv[v], c = vc[v];
            float fx = 0.f, fy = 0.f, fz = 0.f;
            gather_sum(pin, r0, r1, fx, fy, fz);
            int flag;
            float depth, energy;
            pn = vertex_step(p, pp, st, c, make_float3(fx, fy, fz), v, nullptr, nullptr, col, omega, relax_hi,
                             relax_lo, flags & ~kVertTri, 1, flag, depth, energy);
            if (owner && (flags & kVertWriteCol)) {
                cflag[v] = flag;
                cdepth[v] = depth;
                cenergy[v] = energy;
            }
        }
        lu[l] = pn;
        ls[l] = st;
        if (owner) xout[v] = pn;
    }
    __syncthreads();
    if (g < T) {
        const int lv[4] = {q.x & 0xffff, q.y & 0xffff, q.z & 0xffff, q.w & 0xffff};
        float fo[12];
        tet_forces_at<MODE>(b, vt, kk, lu[lv[0]], lu[lv[1]], lu[lv[2]], lu[lv[3]], true, ls[lv[0]], ls[lv[1]],
                            ls[lv[2]], ls[lv[3]], fo);
        const int pos[4] = {(int)((unsigned)q.x >> 16), (int)((unsigned)q.y >> 16), (int)((unsigned)q.z >> 16),
                            (int)((unsigned)q.w >> 16)};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            L.x[pos[j]] = fo[j * 3 + 0];
            L.y[pos[j]] = fo[j * 3 + 1];
            L.z[pos[j]] = fo[j * 3 + 2];
        }
    }
    __syncthreads();
    if (tr.live) pout[tr.slot] = tile_row_sum(L, tr.r0, tr.r1);
}

// The start of a warm-started V-cycle step in one launch (tiled layout, the
// levels that run fused sweeps): k_st and k_tet_dual_tiled.  Workgroups
// 0..ntiles-1 take the tiles at the warm start, ntiles..2 ntiles-1 the same
// tiles at s.  Phase 1, over the tile's vertices: calculateST
// (PDSolver.cu:72-116) in k_st's operation order, s and the warm start's
// correction q (temporal extrapolation fix (2 U1 - U2), plus the prolongated
// learned correction epred) into LDS; the vertex's owner tile of the first
// half stores x_last = x, s and q (to xw: x is read by the other tiles of
// this launch, so q goes to a spare buffer that the dual residual copies
// into x and x_prev).  The damped velocity k_st stores is not stored: the
// V-cycle's finish overwrites v from positions before anything reads it.
// Phase 2: the local step at s + q (first half) or at s (second half), from
// LDS, its partials as k_tet_dual_tiled writes them (pcf, pcf2).  Block 0
// advances the sphere centres by vel dt, as k_st does.
template <int MODE>
__global__ void __launch_bounds__(kTileTets) k_start_fused(
    int T, int ntiles, const float4* __restrict__ x, const float4* __restrict__ v, float4* __restrict__ xlast,
    float4* __restrict__ s, float4* __restrict__ xw, const float4* __restrict__ vc, float dt, float damp, float gx,
    float gy, float gz, DevColliders* col, const float4* __restrict__ U1, const float4* __restrict__ U2,
    const float4* __restrict__ epred, const int4* __restrict__ tiid, const float4* __restrict__ tiw, TileMeta tm,
    const int2* __restrict__ tvqp, const int4* __restrict__ tetq, const float* __restrict__ B,
    const float* __restrict__ vol, float k, const float* __restrict__ tk, float3* __restrict__ pcf,
    float3* __restrict__ pcf2) {
    __shared__ float4 lu[kFuseMaxNv], ls[kFuseMaxNv];  // the tile's q and s
    __shared__ TileLds L;
    const int b0 = xcd_block(blockIdx.x, gridDim.x);
    const bool at_s = b0 >= ntiles;
    const int blk = at_s ? b0 - ntiles : b0;
    const int l = threadIdx.x;
    if (blockIdx.x == 0 && l == 0) {
        for (int q = 0; q < col->nsph; q++) {
            float4 c = col->sph[q];
            const float4 w = col->vel[q];
            c.x = c.x + w.x * dt;
            c.y = c.y + w.y * dt;
            c.z = c.z + w.z * dt;
            col->sph[q] = c;
        }
    }
    // first round trip: the tile's rows and the tet's streams (branch-free,
    // as k_sweep_fused)
    const TileRow tr = tile_row(blk, tm.stride, tm.tpad);
    const int2 tq = tvqp[(size_t)blk * tm.stride + min(l, tm.stride - 1)];
    // the learned correction's binding, padded as the rows (first half only:
    // the second needs s alone)
    int4 bid = make_int4(0, 0, 0, 0);
    float4 bw = make_float4(0.f, 0.f, 0.f, 0.f);
    if (epred && !at_s) {
        bid = tiid[(size_t)blk * tm.stride + min(l, tm.stride - 1)];
        bw = tiw[(size_t)blk * tm.stride + min(l, tm.stride - 1)];
    }
    const int g = blk * kTileTets + l;
    const int gc = min(g, T - 1);
    const int4 qt = tetq[gc];
    float b[9];
#pragma unroll
    for (int c = 0; c < 9; c++) b[c] = B[(size_t)c * T + gc];
    const float vt = vol[gc];
    const float kk = tk ? tk[gc] : k;
    if (tr.live) {
        const int vv = tq.x & 0x7fffffff;
        const float4 p0 = x[vv];
        float4 p = p0, u = v[vv];
        const float fix = vc[vv].z;
        const float df = damp * fix;
        u.x = u.x * df;
        u.y = u.y * df;
        u.z = u.z * df;
        u.x = u.x + gx * dt * fix;
        u.y = u.y + gy * dt * fix;
        u.z = u.z + gz * dt * fix;
        p.x = p.x + u.x * dt * fix;
        p.y = p.y + u.y * dt * fix;
        p.z = p.z + u.z * dt * fix;
        float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
        if (U1 && !at_s) {
            const float4 a = U1[vv], bq = U2[vv];
            q = make_float4(fix * (2.f * a.x - bq.x), fix * (2.f * a.y - bq.y), fix * (2.f * a.z - bq.z), 0.f);
        }
        if (epred && !at_s) {
            const float3 e = prolong_w(bid, bw, epred);
            q = make_float4(q.x + fix * e.x, q.y + fix * e.y, q.z + fix * e.z, 0.f);
        }
        lu[l] = q;
        ls[l] = p;
        if (tq.x < 0 && !at_s) {
            xlast[vv] = p0;
            s[vv] = p;
            xw[vv] = q;
        }
    }
    __syncthreads();
    if (g < T) {
        const int lv[4] = {qt.x & 0xffff, qt.y & 0xffff, qt.z & 0xffff, qt.w & 0xffff};
        float fo[12];
        if (at_s) {
            const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
            tet_forces_at<MODE>(b, vt, kk, ls[lv[0]], ls[lv[1]], ls[lv[2]], ls[lv[3]], false, z, z, z, z, fo);
        } else {
            tet_forces_at<MODE>(b, vt, kk, lu[lv[0]], lu[lv[1]], lu[lv[2]], lu[lv[3]], true, ls[lv[0]], ls[lv[1]],
                                ls[lv[2]], ls[lv[3]], fo);
        }
        const int pos[4] = {(int)((unsigned)qt.x >> 16), (int)((unsigned)qt.y >> 16), (int)((unsigned)qt.z >> 16),
                            (int)((unsigned)qt.w >> 16)};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            L.x[pos[j]] = fo[j * 3 + 0];
            L.y[pos[j]] = fo[j * 3 + 1];
            L.z[pos[j]] = fo[j * 3 + 2];
        }
    }
    __syncthreads();
    if (tr.live) (at_s ? pcf2 : pcf)[tr.slot] = tile_row_sum(L, tr.r0, tr.r1);
}

// updatePointInTet (PDSolver.cu:724-744): dst.x = dst.x_prev = sum w src.x.
__global__ void __launch_bounds__(256) k_interp(int N, float4* __restrict__ dx, float4* __restrict__ dxprev,
                                                const float4* __restrict__ src, const int4* __restrict__ ids,
                                                const float4* __restrict__ w) {
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int4 id = ids[i];
    const float4 ww = w[i];
    const float4 a = src[id.x], b = src[id.y], c = src[id.z], d = src[id.w];
    const float4 r = make_float4(ww.x * a.x + ww.y * b.x + ww.z * c.x + ww.w * d.x,
                                 ww.x * a.y + ww.y * b.y + ww.z * c.y + ww.w * d.y,
                                 ww.x * a.z + ww.y * b.z + ww.z * c.z + ww.w * d.z, 0.f);
    dx[i] = r;
    dxprev[i] = r;
}

// updateMapping (PDSolver.cu:764-785).
__global__ void __launch_bounds__(256) k_update_mapping(int N, const float4* __restrict__ xf,
                                                        const float4* __restrict__ xc, const int4* __restrict__ ids,
                                                        float4* __restrict__ w) {
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int4 id = ids[i];
    float ww[4];
    barycentric4(f3(xf[i]), f3(xc[id.x]), f3(xc[id.y]), f3(xc[id.z]), f3(xc[id.w]), ww);
    w[i] = make_float4(ww[0], ww[1], ww[2], ww[3]);
}

// UpdateTriNormal (PDSolver.cu:394-425): unit normal per surface triangle.
__global__ void __launch_bounds__(256) k_tri_normal(int S, const float4* __restrict__ x, const int4* __restrict__ tri,
                                                    float4* __restrict__ tn) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= S) return;
    const int4 id = tri[t];
    const float3 A = f3(x[id.x]), B = f3(x[id.y]), C = f3(x[id.z]);
    float3 n = cross3(sub3(B, A), sub3(C, A));
    n = div3(n, len3(n));
    tn[t] = make_float4(n.x, n.y, n.z, 0.f);
}

// AvgOutsideTetVertNormal (PDSolver.cu:436-447) over a deterministic
// vertex -> triangle gather instead of atomics.
__global__ void __launch_bounds__(256) k_vert_normal(int V, const int* __restrict__ trp,
                                                     const int* __restrict__ trid, const float4* __restrict__ tn,
                                                     float4* __restrict__ vn) {
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= V) return;
    float nx = 0.f, ny = 0.f, nz = 0.f;
    for (int e = trp[i]; e < trp[i + 1]; e++) {
        const float4 q = tn[trid[e]];
        nx = nx + q.x;
        ny = ny + q.y;
        nz = nz + q.z;
    }
    const float l = sqrtf(nx * nx + ny * ny + nz * nz);
    if (l > 1.1920928955078125e-7f) {
        nx = nx / l;
        ny = ny / l;
        nz = nz / l;
    }
    vn[i] = make_float4(nx, ny, nz, 0.f);
}

// Block reduction of NV doubles (wave64 shuffles, then LDS across waves).
template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], double* __restrict__ out) {
    __shared__ double sh[NV][8];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        double a = v[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
        if (lane == 0) sh[q][wid] = a;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const int nw = (blockDim.x + 63) >> 6;
        double a = 0.0;
        for (int w = 0; w < nw; w++) a += sh[threadIdx.x][w];
        out[(size_t)blockIdx.x * NV + threadIdx.x] = a;
    }
}

// Final pass of a two-stage reduction: one block sums nb partials of NV values
// in block order (deterministic).
template <int NV>
__global__ void __launch_bounds__(256) k_reduce_final(int nb, const double* __restrict__ part,
                                                      double* __restrict__ out) {
    double v[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
        for (int q = 0; q < NV; q++) v[q] += part[(size_t)b * NV + q];
    __shared__ double sh[NV][256];
#pragma unroll
    for (int q = 0; q < NV; q++) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    if (threadIdx.x < NV) {
        double a = 0.0;
        for (int t = 0; t < (int)blockDim.x; t++) a += sh[threadIdx.x][t];
        out[threadIdx.x] = a;
    }
}

// runCalEnergy (PDSolver.cu:816-913), per-vertex part: dX^2, m |x-s|^2,
// collision energy, collided count (FP64 partial sums).
__global__ void __launch_bounds__(256) k_energy_vert(int V, const float4* __restrict__ x,
                                                     const float4* __restrict__ xprev, const float4* __restrict__ s,
                                                     const float4* __restrict__ vc, const int* __restrict__ cflag,
                                                     const float* __restrict__ cenergy, double* __restrict__ part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < V) {
        const float4 p = x[i], q = xprev[i], st = s[i];
        const double ax = (double)p.x - q.x, ay = (double)p.y - q.y, az = (double)p.z - q.z;
        const double bx = (double)p.x - st.x, by = (double)p.y - st.y, bz = (double)p.z - st.z;
        v[0] = ax * ax + ay * ay + az * az;
        v[1] = (double)vc[i].w * (bx * bx + by * by + bz * bz);
        v[2] = (double)cenergy[i];
        v[3] = cflag[i] > 0 ? 1.0 : 0.0;
    }
    block_reduce_store<4>(v, part);
}

// Per-tet part: k vol / 2 |F - R|^2.
__global__ void __launch_bounds__(256) k_energy_tet(int T, const float4* __restrict__ x,
                                                    const int4* __restrict__ tet, const float* __restrict__ B,
                                                    const float* __restrict__ vol, float k, const float* __restrict__ tk,
                                                    double* __restrict__ part) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (g < T) {
        const int4 id = tet[g];
        const float4 p0 = x[id.x], p1 = x[id.y], p2 = x[id.z], p3 = x[id.w];
        float b[9];
        for (int q = 0; q < 9; q++) b[q] = B[(size_t)q * T + g];
        float D[9] = {p1.x - p0.x, p2.x - p0.x, p3.x - p0.x, p1.y - p0.y, p2.y - p0.y,
                      p3.y - p0.y, p1.z - p0.z, p2.z - p0.z, p3.z - p0.z};
        float F[9], RF[9];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) F[r * 3 + c] = D[r * 3 + 0] * b[c] + D[r * 3 + 1] * b[3 + c] + D[r * 3 + 2] * b[6 + c];
        polar_rf<true>(F, RF);  // runCalEnergy calls GetRotation_D (PDSolver.cu:883)
        double e = 0.0;
        for (int q = 0; q < 9; q++) e += (double)RF[q] * (double)RF[q];
        v[0] = e * (double)vol[g] * (double)(tk ? tk[g] : k) / 2.0;
    }
    block_reduce_store<1>(v, part);
}

// runCalEnergy's per-vertex Ep density (calEveryVertEp, PDSolver.cu:893-924),
// per-tet part: e = k vol / 2 |F - R|^2, negated where det F < 1, and the
// current signed volume (GetVolumn, utils.cu:254-275), written to the tet's
// four corner slots of the CSR corner array so that each vertex sums its
// incident tets in ascending tet order, the reference's loop order.
__global__ void __launch_bounds__(256) k_ep_tet(int T, const float4* __restrict__ x, const int4* __restrict__ tet,
                                                const float* __restrict__ B, const float* __restrict__ vol, float k,
                                                const float* __restrict__ tk, const int4* __restrict__ slot,
                                                float3* __restrict__ cf) {
    const int g = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (g >= T) return;
    const int4 id = tet[g];
    const float4 p0 = x[id.x], p1 = x[id.y], p2 = x[id.z], p3 = x[id.w];
    float b[9];
    for (int q = 0; q < 9; q++) b[q] = B[(size_t)q * T + g];
    float D[9] = {p1.x - p0.x, p2.x - p0.x, p3.x - p0.x, p1.y - p0.y, p2.y - p0.y,
                  p3.y - p0.y, p1.z - p0.z, p2.z - p0.z, p3.z - p0.z};
    float F[9], RF[9];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) F[r * 3 + c] = D[r * 3 + 0] * b[c] + D[r * 3 + 1] * b[3 + c] + D[r * 3 + 2] * b[6 + c];
    polar_rf<true>(F, RF);
    float e = 0.f;
    for (int q = 0; q < 9; q++) e = e + RF[q] * RF[q];
    e = e * vol[g] * (tk ? tk[g] : k) / 2;
    const float det = F[0] * F[4] * F[8] + F[1] * F[5] * F[6] + F[3] * F[7] * F[2] - F[2] * F[4] * F[6] -
                      F[1] * F[3] * F[8] - F[0] * F[5] * F[7];
    if (det < 1.0f) e = -e;
    // GetVolumn: (AB, AC, AD) as columns, det by the first column's cofactors
    const float r0 = D[4] * D[8] - D[7] * D[5], r1 = D[7] * D[2] - D[1] * D[8], r2 = D[1] * D[5] - D[4] * D[2];
    const float v = (D[0] * r0 + D[3] * r1 + D[6] * r2) / 6.0f;
    const int4 sl = slot[g];
    const float3 q = make_float3(e, v, 0.f);
    cf[sl.x] = q;
    cf[sl.y] = q;
    cf[sl.z] = q;
    cf[sl.w] = q;
}

// Per-vertex part: density = sum of the incident tets' signed e / sum of
// their current volumes, both summed in ascending tet order.
__global__ void __launch_bounds__(256) k_ep_vert(int V, const int* __restrict__ rp, const float3* __restrict__ cf,
                                                 float* __restrict__ dens) {
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= V) return;
    float es = 0.f, vs = 0.f;
    for (int e = rp[i]; e < rp[i + 1]; e++) {
        const float3 q = cf[e];
        es = es + q.x;
        vs = vs + q.y;
    }
    dens[i] = es / vs;
}

// Scratch of the allocation-free energy diagnostics (mgsb_energy_noalloc):
// the FP64 block partials of the two reductions and their results, in a
// fixed device array instead of the handle's partials (which grow by
// hipMalloc on demand).  Every GPU the library runs on holds its own copy
// until the process ends; in host code `scratch` is only a symbol, and its
// address on each device comes from hipGetSymbolAddress (scratch_setup).
constexpr size_t kScratchDoubles = 131072;  // 1 MiB
constexpr size_t kScratchResults = 16;      // the last 16 doubles take the reduction results
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kHistMax = 64;  // checks of a step kept for mgsb_step_history
struct VState {
    double res0;    // ||g(s)||
    double res;     // ||g(x)|| at the last check
    int converged;  // res <= tol * res0
    int cycles;     // full cycles completed this step
    int check;      // 0: no checks, 1: after each cycle, 2: also after pre-smoothing
    float tol;
    // run aggregates since the last reset (updated at the end of each step)
    double max_rel;
    long long total_cycles;
    int nsteps;
    int nfail;      // steps that ended with res > tol * res0
    int fallback;   // learned schedule: the warm start was rejected this step
    int nfallback;
    double res_ws;  // residual right after the learned correction
    int halves;     // half-cycles run this step
    int total_halves;
    double res_t;   // residual at the temporal extrapolation
    int accepted;   // warm-start stage x holds: 0 s, 1 temporal, 2 temporal + learned
    int revert;     // pending revert of x: 0 none, 1 to s, 2 to the temporal extrapolation
    int cf_ok;      // corner forces are those of the current x (or the step has converged)
    double rel;     // convergence measure at the last check: res / res0, or with several
                    // bodies the largest per-body res_b / res0_b
    int maxh;       // half-cycle cap of a step (2 max_cycles, + 1 with coarse_first)
    int notdone;    // the step has neither converged nor reached the cap (no checks: 0);
                    // the finish piece, enqueued before the host reads the verdict, skips
    double slack;   // the learned stage is kept while its residual <= slack * res
    int cfirst;     // the step's first half-cycle is a coarse correction (coarse_first)
    int guard;      // stall_guard: a check whose residual exceeds kStallRatio times the previous one's ...
    int calm;       // ... turns the fine smoother's Chebyshev extrapolation off for the rest of the step
    int nhist;      // checks of the current step so far, and their relative residuals
    double hist[kHistMax];
};

// Check modes of the fused residual reduction.
constexpr int kChkNone = -1;
constexpr int kChkStart = 0;  // g(s): res0
constexpr int kChkCycle = 1;  // end of a cycle
constexpr int kChkWarm = 2;   // after the learned warm start (judge, may reject it)
constexpr int kChkMid = 3;    // after pre-smoothing
constexpr int kChkTemporal = 4;  // after the temporal extrapolation

// stall_guard: a check whose residual exceeds kStallRatio times the previous
// check's turns the fine smoother's Chebyshev extrapolation off for the rest
// of the step.  Steps that stall next to inverted tets (the blob) jump by 4-18x
// between checks; contact steps (the large box) creep up by at most ~2.8x
// while they converge, which the guard leaves alone.
constexpr double kStallRatio = 3.0;

// The warm-start judge: a stage is kept only if its residual r is at most the
// best one so far (times slack for the learned stage); a non-finite residual
// is never kept.
__device__ __forceinline__ bool judge_keeps(double r, double best, int mode, double slack) {
    const double bound = mode == kChkWarm ? slack * best : best;
    return r <= bound;
}

// Convergence bookkeeping from the squared residual norm t.  hm (host-mapped,
// optional) mirrors (converged, cycles, halves, fallback, accepted) for the
// host's cycle loop.  rel_in >= 0: the convergence measure of several bodies
// (largest per-body relative residual), tested as rel_in <= tol; otherwise
// the test is res <= tol * res0 on the whole-mesh norms.
__device__ void vstate_apply(VState* st, double t, int mode, int* hm, double rel_in = -1.0) {
    const double r = sqrt(t);
    auto ok = [&](double res) {
        return rel_in >= 0.0 ? rel_in <= (double)st->tol : res <= (double)st->tol * st->res0;
    };
    auto rel_of = [&](double res) { return rel_in >= 0.0 ? rel_in : (st->res0 > 0.0 ? res / st->res0 : 0.0); };
    if (mode == kChkStart) {
        st->res0 = r;
        st->res = r;
        st->rel = r > 0.0 ? 1.0 : 0.0;
        st->cycles = 0;
        st->halves = 0;
        st->fallback = 0;
        st->res_t = 0.0;
        st->res_ws = 0.0;
        st->accepted = 0;
        st->revert = 0;
        st->cf_ok = 1;
        st->calm = 0;
        st->converged = (st->check && r == 0.0) ? 1 : 0;
    } else if (mode == kChkTemporal || mode == kChkWarm) {
        // a warm-start stage is kept only if it lowers the best residual so
        // far; a non-finite residual (a diverged network) is never kept
        if (mode == kChkTemporal)
            st->res_t = r;
        else
            st->res_ws = r;
        if (!judge_keeps(r, st->res, mode, st->slack)) {
            st->revert = st->accepted == 1 ? 2 : 1;
            st->fallback |= mode == kChkTemporal ? 1 : 2;
            st->cf_ok = 0;
        } else {
            st->res = r;
            st->rel = rel_of(r);
            st->accepted = mode == kChkTemporal ? 1 : 2;
            st->revert = 0;
            st->cf_ok = 1;
            if (st->check && ok(r)) st->converged = 1;
        }
    } else {
        if (st->guard && !(r <= kStallRatio * st->res)) st->calm = 1;  // the iteration jumps: no extrapolation
        st->res = r;
        st->rel = rel_of(r);
        if (mode == kChkCycle) {
            // with coarse_first the first check follows the coarse half alone
            st->halves += (st->check == 2 || (st->cfirst && st->cycles == 0)) ? 1 : 2;
            st->cycles += 1;
        } else {
            st->halves += 1;
        }
        if (st->check && ok(r)) st->converged = 1;
    }
    st->notdone = (st->check == 0 || st->converged || (st->maxh > 0 && st->halves >= st->maxh)) ? 0 : 1;
    if (mode == kChkStart) st->nhist = 0;
    if (st->nhist < kHistMax) st->hist[st->nhist++] = st->res0 > 0.0 ? r / st->res0 : 0.0;
    (void)hm;  // the check kernel publishes the verdict with its count (bump_seq)
}

// End of a step: fold the step's outcome into the run aggregates.
__device__ void vstate_step_done(VState* st) {
    // a non-finite residual counts as an infinitely bad step, not as none
    const double rel = st->rel == st->rel ? st->rel : __builtin_huge_val();
    if (rel > st->max_rel) st->max_rel = rel;
    st->total_cycles += st->cycles;
    st->total_halves += st->halves;
    st->nsteps += 1;
    if (rel > (double)st->tol) st->nfail += 1;
    st->nfallback += st->fallback != 0;
}

// Workgroup sum of v in a fixed order: a shuffle tree in each wave, then the
// waves' sums in wave order; thread 0 holds the totals.
template <int NV>
__device__ __forceinline__ void wg_sum_fixed(double (&v)[NV]) {
    __shared__ double sh[NV][16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_down(v[q], off, 64);
        if (lane == 0) sh[q][wid] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NV; q++) {
            double t = 0.0;
            for (int w = 0; w < nw; w++) t += sh[q][w];
            v[q] = t;
        }
}

// Every check kernel, skipped or not, counts itself in seqc and publishes the
// count with the step's verdict as ONE 64-bit store to host-mapped memory
// (low word: the count; high word: converged, fallback bits, accepted stage,
// half-cycles, cycles).  The host waits for the count of the checks it has
// enqueued, instead of for the whole stream, and reads the verdict from the
// same word, so no fence has to order separate mirror writes.
__device__ __forceinline__ unsigned pack_verdict(const VState* st) {
    const unsigned h = (unsigned)min(st->halves, 255), c = (unsigned)min(st->cycles, 255);
    return (st->converged ? 1u : 0u) | ((unsigned)st->fallback & 3u) << 1 | ((unsigned)st->accepted & 3u) << 3 |
           h << 8 | c << 16;
}
__device__ __forceinline__ void bump_seq(unsigned* seqc, int* hm, const VState* st) {
    const unsigned n = atomicAdd(seqc, 1u) + 1u;
    if (hm) {
        const unsigned long long w = (unsigned long long)pack_verdict(st) << 32 | n;
        __hip_atomic_store(reinterpret_cast<unsigned long long*>(hm), w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Second stage of the residual-norm reduction: one workgroup sums the nb
// block partials in a fixed order (lane-strided, then wg_sum_fixed: the same
// order on every run) and applies the check.  A separate launch: completing the norm inside the producing
// kernel (the last block to arrive, found with a device-scope counter) costs
// far more on this GPU, since the blocks' atomics on one counter serialise.
constexpr int kCheckBatch = 4;
__global__ void __launch_bounds__(256) k_check(int nb, const double* __restrict__ part, VState* st, int mode,
                                               int* hm, const int* __restrict__ skip, unsigned* seqc) {
    // the flag and up to kCheckBatch partials per lane in one round trip
    // (addresses past nb read the last partial, which is dropped)
    const int sk = load_flag(skip);
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < nb; b += kCheckBatch * blockDim.x) {
        double q[kCheckBatch];
#pragma unroll
        for (int k = 0; k < kCheckBatch; k++) q[k] = part[min(b + k * (int)blockDim.x, nb - 1)];
        if (sk) break;
#pragma unroll
        for (int k = 0; k < kCheckBatch; k++)
            if (b + k * (int)blockDim.x < nb) v[0] += q[k];
    }
    if (sk) {
        if (threadIdx.x == 0) bump_seq(seqc, hm, st);
        return;
    }
    wg_sum_fixed<1>(v);
    if (threadIdx.x == 0) {
        vstate_apply(st, v[0], mode, hm);
        bump_seq(seqc, hm, st);
    }
}

// Nonlinear residual g = m/h^2 (s - x) + f_el + f_col of a vertex on free
// DOFs from its summed corner forces f: p is x (or the correction u when
// split), st = s, c = (m/h^2, diag, fix, m).  contact: a collider acts on the
// vertex (the active set).
__device__ __forceinline__ float3 vertex_residual(float4 p, float4 st, float4 c, float fx, float fy, float fz,
                                                  const DevColliders* __restrict__ col, int flags, int split,
                                                  bool& contact) {
    const float4 pa = split ? make_float4(st.x + p.x, st.y + p.y, st.z + p.z, 0.f) : p;
    const float sx = split ? -p.x : st.x - p.x, sy = split ? -p.y : st.y - p.y, sz = split ? -p.z : st.z - p.z;
    float cfx = 0.f, cfy = 0.f, cfz = 0.f;
    contact = false;
    if (flags & kVertPoint) {
        const float kc = col->kcol;
        const float3 P = f3(pa);
        for (int q = 0; q < col->nsph; q++) {
            if (col->vel[q].w == 0.f) continue;
            const float4 sp = col->sph[q];
            const float3 center = make_float3(sp.x, sp.y, sp.z);
            const float3 cp = split ? add3(sub3(f3(st), center), f3(p)) : sub3(P, center);
            const float3 dir = div3(cp, len3(cp));
            const float distance = len3(cp);
            const float rr = sp.w * 1.5f;
            if (distance < rr) {
                contact = true;
                const float bq = 2 * dot3(cp, dir);
                const float cq = dot3(cp, cp) - rr * rr;
                float t = (-bq + sqrtf(bq * bq - 4.f * cq)) / 2.f;
                t = t / 2.0f;
                const float3 colP = add3(P, mul3(dir, t));
                const float3 ccp = split ? add3(cp, mul3(dir, t)) : sub3(colP, center);
                const float3 n = div3(ccp, len3(ccp));
                const float nd = dot3(n, dir);
                cfx = cfx + kc * t * nd * n.x;
                cfy = cfy + kc * t * nd * n.y;
                cfz = cfz + kc * t * nd * n.z;
            }
        }
        if (col->plane_on) {
            const float4 pl = col->plane;
            const float sd = split ? ((pl.x * st.x + pl.y * st.y + pl.z * st.z) - pl.w) +
                                         (pl.x * p.x + pl.y * p.y + pl.z * p.z)
                                   : pl.x * pa.x + pl.y * pa.y + pl.z * pa.z - pl.w;
            if (sd < 0.f) {
                contact = true;
                cfx = cfx + col->kplane * (-sd) * pl.x;
                cfy = cfy + col->kplane * (-sd) * pl.y;
                cfz = cfz + col->kplane * (-sd) * pl.z;
            }
        }
    }
    const float fix = c.z;
    const float gx = (c.x * sx + (fx + cfx)) * fix;
    const float gy = (c.x * sy + (fy + cfy)) * fix;
    const float gz = (c.x * sz + (fz + cfz)) * fix;
    return make_float3(gx, gy, gz);
}

// Nonlinear residual over free DOFs (vertex_residual); writes g (float4) when
// r != nullptr and the FP64 block partials of |g|^2 when part != nullptr
// (k_check completes the norm).
// After fused sweeps the positions sit in a rotated buffer: xcanon != nullptr
// takes x (the level's x buffer), and xprev takes xpsrc[i] instead of x when
// xpsrc != nullptr (the level's x_prev buffer keeps the smoother's history).
// These buffers may alias each other and x: no __restrict__ on them, and every
// vertex reads before it writes.
__global__ void __launch_bounds__(256) k_resid(int V, const float4* x, const float4* __restrict__ s,
                                               const float4* __restrict__ vc, const int* __restrict__ rp,
                                               const float3* __restrict__ cf, const DevColliders* __restrict__ col,
                                               int flags, float4* __restrict__ r, double* __restrict__ part,
                                               float4* xprev, const int* __restrict__ skip, int split,
                                               double* __restrict__ g2, float* __restrict__ cmask, int lsh,
                                               float4* xcanon = nullptr, const float4* xpsrc = nullptr) {
    // the skip flag first (a skipped launch loads nothing), then the vertex's
    // streams and its row bounds in one round trip (lanes past V read vertex
    // V - 1 and drop it: no branch around the loads)
    if (load_flag(skip)) return;
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = gi >> lsh, sub = gi & ((1 << lsh) - 1);  // 2^lsh lanes per vertex
    const int ic = min(i, V - 1);
    const float4 p = x[ic], st = s[ic], c = vc[ic];
    // x_prev's new value: loaded from a valid buffer either way (a select
    // between xpsrc[i] and the local p would put p on the stack)
    const float4 pl = (xpsrc ? xpsrc : x)[ic];
    const int e0 = rp[ic], e1 = rp[ic + 1];
    double v[1] = {0.0};
    if (i < V) {
        const float4 pq = xpsrc ? pl : p;
        const float3 f = gather_lanes(cf, e0, e1, sub, lsh);
        if (sub == 0) {
            bool contact;
            const float3 g = vertex_residual(p, st, c, f.x, f.y, f.z, col, flags, split, contact);
            // the active set is left out of the restriction: the coarse operators
            // carry no penalty stiffness (cmask also masks the coarse corrections)
            const float am = contact ? 0.f : 1.f;
            if (cmask) cmask[i] = am;
            if (r) r[i] = cmask ? make_float4(g.x * am, g.y * am, g.z * am, 0.f) : make_float4(g.x, g.y, g.z, 0.f);
            if (xcanon) xcanon[i] = p;
            if (xprev) xprev[i] = pq;
            v[0] = (double)g.x * g.x + (double)g.y * g.y + (double)g.z * g.z;
            if (g2) g2[i] = v[0];
        }
    }
    if (part) block_reduce_store<1>(v, part);
}

// Both residuals of a temporally warm-started start on split positions: at
// the extrapolation u = x with the forces cf (g written to r when
// r != nullptr, active set to cmask) and at s itself (u = 0) with the forces
// cf2 (g(s) and its active-set mask written to rs when rs != nullptr).  part[2b] holds block b's partial of |g(s)|^2, part[2b + 1] that of
// |g(s + u)|^2.
__global__ void __launch_bounds__(256) k_resid_dual(int V, const float4* __restrict__ x,
                                                    const float4* __restrict__ s, const float4* __restrict__ vc,
                                                    const int* __restrict__ rp, const float3* __restrict__ cf,
                                                    const float3* __restrict__ cf2,
                                                    const DevColliders* __restrict__ col, int flags,
                                                    float4* __restrict__ r, double* __restrict__ part,
                                                    float* __restrict__ cmask, int lsh, float4* __restrict__ rs,
                                                    float4* __restrict__ xo = nullptr,
                                                    float4* __restrict__ xpo = nullptr) {
    // xo, xpo (k_start_fused): x holds the warm start's correction in a spare
    // buffer; it is copied into x and x_prev here
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = gi >> lsh, sub = gi & ((1 << lsh) - 1);  // 2^lsh lanes per vertex
    double v[2] = {0.0, 0.0};
    if (i < V) {
        const float4 p = x[i], st = s[i], c = vc[i];
        const int r0 = rp[i], r1 = rp[i + 1];
        const float3 a = gather_lanes(cf2, r0, r1, sub, lsh);  // forces at s
        const float3 b = gather_lanes(cf, r0, r1, sub, lsh);   // forces at the extrapolation
        if (sub == 0) {
            bool contact;
            const float3 g0 =
                vertex_residual(make_float4(0.f, 0.f, 0.f, 0.f), st, c, a.x, a.y, a.z, col, flags, 1, contact);
            v[0] = (double)g0.x * g0.x + (double)g0.y * g0.y + (double)g0.z * g0.z;
            if (rs) {  // g(s) masked by the active set at s, the mask in w (for k_revert)
                const float am0 = contact ? 0.f : 1.f;
                rs[i] = make_float4(g0.x * am0, g0.y * am0, g0.z * am0, am0);
            }
            const float3 g = vertex_residual(p, st, c, b.x, b.y, b.z, col, flags, 1, contact);
            const float am = contact ? 0.f : 1.f;
            if (cmask) cmask[i] = am;
            if (r) r[i] = cmask ? make_float4(g.x * am, g.y * am, g.z * am, 0.f) : make_float4(g.x, g.y, g.z, 0.f);
            v[1] = (double)g.x * g.x + (double)g.y * g.y + (double)g.z * g.z;
            if (xo) {
                xo[i] = p;
                xpo[i] = p;
            }
        }
    }
    block_reduce_store<2>(v, part);
}

// Per-body residual norms of several bodies (each a contiguous vertex range
// vb[b] .. vb[b+1] of level 0): block b sums g2 over its body in a fixed order;
// the block that arrives last (device-scope counter) adds the bodies in body
// order, forms the largest per-body relative residual (res0b holds the
// per-body norms of g(s), written at the step's start check) and applies the
// check through vstate_apply.
constexpr int kBodyCheckThreads = 1024;
__global__ void __launch_bounds__(kBodyCheckThreads) k_body_check(const int* __restrict__ vb, const double* __restrict__ g2,
                                                    double* __restrict__ bodysq, double* __restrict__ res0b,
                                                    unsigned* count, VState* st, int mode, int* hm,
                                                    const int* __restrict__ skip, unsigned* seqc) {
    if (skip && *skip) {
        if (blockIdx.x == 0 && threadIdx.x == 0) bump_seq(seqc, hm, st);
        return;
    }
    __shared__ double sh[kBodyCheckThreads];
    __shared__ int s_last;
    const int b = blockIdx.x;
    // a body has up to ~50k vertices: eight loads in flight per lane
    const int e1 = vb[b + 1], stride = blockDim.x;
    double a = 0.0;
    int i = vb[b] + threadIdx.x;
    for (; i + 7 * stride < e1; i += 8 * stride) {
        double q[8];
#pragma unroll
        for (int k = 0; k < 8; k++) q[k] = g2[i + k * stride];
#pragma unroll
        for (int k = 0; k < 8; k++) a += q[k];
    }
    for (; i < e1; i += stride) a += g2[i];
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int w = blockDim.x >> 1; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bodysq[b] = sh[0];
        __threadfence();
        s_last = atomicAdd(count, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last) return;  // the last workgroup to arrive combines the bodies
    __threadfence();
    // the bodies' norms in parallel (one lane per body, strided), the total in
    // a fixed order and the largest relative residual
    double v[1] = {0.0};
    double worst = 0.0;
    for (int q = threadIdx.x; q < (int)gridDim.x; q += blockDim.x) {
        const double sq = __hip_atomic_load(bodysq + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        v[0] += sq;
        const double r = sqrt(sq);
        if (mode == kChkStart) {
            res0b[q] = r;
        } else {
            const double r0 = res0b[q];
            const double rel = r0 > 0.0 ? r / r0 : (r > 0.0 ? 1e300 : 0.0);
            worst = rel > worst ? rel : worst;
        }
    }
    wg_sum_fixed<1>(v);
    __shared__ double s_worst[kBodyCheckThreads / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(worst, off, 64);
        worst = o > worst ? o : worst;
    }
    if ((threadIdx.x & 63) == 0) s_worst[threadIdx.x >> 6] = worst;
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) worst = s_worst[w] > worst ? s_worst[w] : worst;
    *count = 0u;
    vstate_apply(st, v[0], mode, hm, worst);
    bump_seq(seqc, hm, st);
}

// ---------------------------------------------------------------------------
// V-cycle kernels.  Every kernel of a cycle returns at once when the step's
// convergence flag is set, so a captured step with max_cycles cycles costs
// only empty launches after convergence.
// ---------------------------------------------------------------------------
// Warm start: x = u_t = fix (2 u_{n-1} - u_{n-2}), the temporal extrapolation
// of the last two steps' corrections (split positions: x holds u = x - s).
__global__ void __launch_bounds__(256) k_warm_temporal(int V, float4* __restrict__ x, float4* __restrict__ xprev,
                                                       const float4* __restrict__ U1, const float4* __restrict__ U2,
                                                       const float4* __restrict__ vc, const int* __restrict__ skip,
                                                       const float4* __restrict__ epred, const int4* __restrict__ iids,
                                                       const float4* __restrict__ iw) {
    if (skip && *skip) return;
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const float f = vc[i].z;
    float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
    if (U1) {
        const float4 a = U1[i], b = U2[i];
        q = make_float4(f * (2.f * a.x - b.x), f * (2.f * a.y - b.y), f * (2.f * a.z - b.z), 0.f);
    }
    if (epred) {
        const float3 e = prolong_at(i, epred, iids, iw);
        q = make_float4(q.x + f * e.x, q.y + f * e.y, q.z + f * e.z, 0.f);
    }
    x[i] = q;
    xprev[i] = q;
}

// Judge of a rejected warm-start stage: x returns to s (u = 0) or to the
// temporal extrapolation.  cf2 != nullptr holds the corner forces at s (the
// dual start): on a return to s each vertex copies its CSR range of them
// into cf, so the forces match x again without another local step, and with
// rs != nullptr the residual and active set at s into r and cmask.
__global__ void __launch_bounds__(256) k_revert(int V, float4* __restrict__ x, float4* __restrict__ xprev,
                                                const float4* __restrict__ U1, const float4* __restrict__ U2,
                                                const float4* __restrict__ vc, const VState* st,
                                                const int* __restrict__ rp, const float3* __restrict__ cf2,
                                                float3* __restrict__ cf, float4* __restrict__ r,
                                                const float4* __restrict__ rs, float* __restrict__ cmask) {
    const int rv = st->revert;
    if (!rv) return;
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= V) return;
    if (cf2 && rv == 1)
        for (int e = rp[i]; e < rp[i + 1]; e++) cf[e] = cf2[e];
    if (rs && rv == 1) {  // the residual at s (and its active set) for a coarse-first half
        const float4 g = rs[i];
        r[i] = make_float4(g.x, g.y, g.z, 0.f);
        cmask[i] = g.w;
    }
    float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
    if (rv == 2) {
        const float4 a = U1[i], b = U2[i];
        const float f = vc[i].z;
        q = make_float4(f * (2.f * a.x - b.x), f * (2.f * a.y - b.y), f * (2.f * a.z - b.z), 0.f);
    }
    x[i] = q;
    xprev[i] = q;
}

// Both checks of a temporally warm-started start and the judge's revert in
// one launch (the dual start): res0 = |g(s)| from k_resid_dual's partials
// part[2b], then the judge of the warm start from part[2b + 1]; converged
// already at s, x (which holds the warm start) returns to s without a
// verdict.  Every workgroup sums the partials in the same fixed order
// (lane-strided over blockDim 256, then wg_sum_fixed, as k_check) and takes
// the same verdict (judge_keeps, as vstate_apply decides it: after the start
// check the best residual is res0 and no stage is accepted, so a rejection
// returns to s); workgroup 0 applies it to the step state and publishes it.  vstate_apply at the start check resets every field the
// judge reads but check, tol and slack, so the copies decide alike.
__global__ void __launch_bounds__(256) k_judge_revert(int nb, const double* __restrict__ part, VState* st, int* hm,
                                                      unsigned* seqc, int judge, int V, float4* __restrict__ x,
                                                      float4* __restrict__ xprev, const int* __restrict__ rp,
                                                      const float3* __restrict__ cf2, float3* __restrict__ cf,
                                                      float4* __restrict__ r, const float4* __restrict__ rs,
                                                      float* __restrict__ cmask) {
    // the partials (kCheckBatch per lane), the check state the verdict reads
    // and the vertex's row bounds in one round trip
    const int chk = st->check;
    const double slack = st->slack;
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int ic = min(i, V - 1);
    const int e0 = rp[ic], e1 = rp[ic + 1];
    double v[2] = {0.0, 0.0};
    for (int q = threadIdx.x; q < nb; q += kCheckBatch * blockDim.x) {
        double a[kCheckBatch], b[kCheckBatch];
#pragma unroll
        for (int k = 0; k < kCheckBatch; k++) {
            const int qc = min(q + k * (int)blockDim.x, nb - 1);
            a[k] = part[2 * qc];
            b[k] = part[2 * qc + 1];
        }
#pragma unroll
        for (int k = 0; k < kCheckBatch; k++)
            if (q + k * (int)blockDim.x < nb) {
                v[0] += a[k];
                v[1] += b[k];
            }
    }
    wg_sum_fixed<2>(v);
    __shared__ int s_rv;
    if (threadIdx.x == 0) {
        if (blockIdx.x == 0) {
            vstate_apply(st, v[0], kChkStart, hm);
            if (!st->converged)
                vstate_apply(st, v[1], judge, hm);
            else
                st->revert = 1;
            s_rv = st->revert;
            bump_seq(seqc, hm, st);
        } else {  // the same decision from the same sums (vstate_apply at the start, then the judge)
            const double r0 = sqrt(v[0]);
            const bool conv = chk && r0 == 0.0;
            s_rv = (conv || !judge_keeps(sqrt(v[1]), r0, judge, slack)) ? 1 : 0;
        }
    }
    __syncthreads();
    const int rv = s_rv;  // the dual start rejects to s only
    if (!rv) return;
    if (i >= V) return;
    for (int e = e0; e < e1; e++) cf[e] = cf2[e];
    if (rs) {
        const float4 g = rs[i];
        r[i] = make_float4(g.x, g.y, g.z, 0.f);
        cmask[i] = g.w;
    }
    x[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    xprev[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Features of the learned coarse correction at each vertex of level 1, all
// known before the step starts: the fine velocity injected through the
// coarse-in-fine binding (3), the rest position (3), and the coarse deltas
// of the last two steps (3 + 3): the fine corrections beyond the temporal
// extrapolation (D1, D2, kept by k_finish_split) injected the same way;
// normalised with (mu, sd).  Raw values optionally recorded.
constexpr int kFeat = 12;
// k_judge_revert folded into the coarse-first half that follows a fused
// warm start (same captured sequence): every workgroup sums the dual
// residual's partials in k_judge_revert's order (256 lanes, kCheckBatch
// partials a lane, wg_sum_fixed) and takes the same verdict; workgroup 0
// applies res0 and the judge to the check state and publishes the count and
// verdict.  A converged start ends the launch (later kernels skip on the
// converged flag); a rejected warm start restricts g(s) (rs) instead of g at
// the warm start (rf), and the prolongation sweep after it starts from u = 0
// with the active set at s (k_sweep_fused's jrev).  Those are the only
// readers of what k_judge_revert's return to s writes (x, x_prev, the
// partials, r, the active-set mask) before the half's own sweeps and
// residual write them again.  Then the restriction as k_restrict.
__global__ void __launch_bounds__(256) k_restrict_judged(int Vc, const int* __restrict__ rtp,
                                                         const int* __restrict__ rti, const float* __restrict__ rtw,
                                                         const float4* __restrict__ rf, const float4* __restrict__ rs,
                                                         const float4* __restrict__ vcl, float4* __restrict__ rc, int nb,
                                                         const double* __restrict__ part, VState* st, int* hm,
                                                         unsigned* seqc, int judge) {
    const int chk = st->check;
    const double slack = st->slack;
    const float tol = st->tol;
    double v[2] = {0.0, 0.0};
    for (int q = threadIdx.x; q < nb; q += kCheckBatch * blockDim.x) {
        double a[kCheckBatch], b[kCheckBatch];
#pragma unroll
        for (int k = 0; k < kCheckBatch; k++) {
            const int qc = min(q + k * (int)blockDim.x, nb - 1);
            a[k] = part[2 * qc];
            b[k] = part[2 * qc + 1];
        }
#pragma unroll
        for (int k = 0; k < kCheckBatch; k++)
            if (q + k * (int)blockDim.x < nb) {
                v[0] += a[k];
                v[1] += b[k];
            }
    }
    wg_sum_fixed<2>(v);
    __shared__ int s_mode;  // 0 the warm start kept, 1 rejected (back to s), 2 the step has converged
    if (threadIdx.x == 0) {
        int mode;
        if (blockIdx.x == 0) {
            vstate_apply(st, v[0], kChkStart, hm);
            if (!st->converged)
                vstate_apply(st, v[1], judge, hm);
            else
                st->revert = 1;
            mode = st->converged ? 2 : (st->revert ? 1 : 0);
            bump_seq(seqc, hm, st);
        } else {  // the same decision from the same sums (vstate_apply at the start, then the judge)
            const double r0 = sqrt(v[0]), r1 = sqrt(v[1]);
            const bool keep = judge_keeps(r1, r0, judge, slack);
            if (chk && r0 == 0.0)
                mode = 2;
            else if (!keep)
                mode = 1;
            else
                mode = (chk && r1 <= (double)tol * r0) ? 2 : 0;
        }
        s_mode = mode;
    }
    __syncthreads();
    const int mode = s_mode;
    if (mode == 2) return;
    const float4* __restrict__ src = mode == 1 ? rs : rf;
    const int j = (xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (j >= Vc) return;  // whole wave
    const int e0 = rtp[j], e1 = rtp[j + 1];
    const float fx = vcl[j].z;
    float ax = 0.f, ay = 0.f, az = 0.f;
    for (int e = e0 + lane; e < e1; e += 4 * 64) {
        float w[4];
        int id[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const bool on = e + k * 64 < e1;
            w[k] = on ? rtw[e + k * 64] : 0.f;
            id[k] = on ? rti[e + k * 64] : 0;
        }
        float4 q[4];
#pragma unroll
        for (int k = 0; k < 4; k++) q[k] = e + k * 64 < e1 ? src[id[k]] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (e + k * 64 < e1) {
                ax = ax + w[k] * q[k].x;
                ay = ay + w[k] * q[k].y;
                az = az + w[k] * q[k].z;
            }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        ax += __shfl_down(ax, off, 64);
        ay += __shfl_down(ay, off, 64);
        az += __shfl_down(az, off, 64);
    }
    if (lane == 0) rc[j] = make_float4(ax * fx, ay * fx, az * fx, 0.f);
}

__global__ void __launch_bounds__(256) k_features(int Vc, const float4* __restrict__ vf, const int4* __restrict__ aids,
                                                  const float4* __restrict__ aw, const float4* __restrict__ X0,
                                                  const float4* __restrict__ D1, const float4* __restrict__ D2,
                                                  const float* __restrict__ norm, float* __restrict__ feat,
                                                  float* __restrict__ rec) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= Vc) return;
    const int4 id = aids[j];
    const float4 w = aw[j];
    const float4 a = vf[id.x], b = vf[id.y], c = vf[id.z], d = vf[id.w];
    const float4 X = X0[j];
    const float3 ep = inject4(D1, id, w), ep2 = inject4(D2, id, w);
    const float raw[kFeat] = {w.x * a.x + w.y * b.x + w.z * c.x + w.w * d.x,
                              w.x * a.y + w.y * b.y + w.z * c.y + w.w * d.y,
                              w.x * a.z + w.y * b.z + w.z * c.z + w.w * d.z,
                              X.x, X.y, X.z, ep.x, ep.y, ep.z, ep2.x, ep2.y, ep2.z};
#pragma unroll
    for (int q = 0; q < kFeat; q++) {
        if (rec) rec[(long)j * kFeat + q] = raw[q];
        feat[(long)j * kFeat + q] = (raw[q] - norm[q]) / norm[kFeat + q];
    }
}

// Training target of a recorded step, after k_finish_split has shifted the
// history: the change of the coarse delta, inject(D1) - inject(D2) (this
// step's against the last step's, which is the network's prior).
__global__ void __launch_bounds__(256) k_target(int Vc, const float4* __restrict__ D1, const float4* __restrict__ D2,
                                                const int4* __restrict__ aids, const float4* __restrict__ aw,
                                                float* __restrict__ rec, const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= Vc) return;
    const int4 id = aids[j];
    const float4 w = aw[j];
    const float3 e = inject4(D1, id, w), e1 = inject4(D2, id, w);
    rec[(long)j * 3 + 0] = e.x - e1.x;
    rec[(long)j * 3 + 1] = e.y - e1.y;
    rec[(long)j * 3 + 2] = e.z - e1.z;
}

// r_c[j] = vfix_j * sum_i w_ij r_f[i]  (P^T, no atomics).  One wave per
// coarse vertex: lane l sums entries l, l+64, ... of the row, then a fixed
// shuffle tree combines the lanes (deterministic order).
__global__ void __launch_bounds__(256) k_restrict(int Vc, const int* __restrict__ rtp, const int* __restrict__ rti,
                                                  const float* __restrict__ rtw, const float4* __restrict__ rf,
                                                  const float4* __restrict__ vcl, float4* __restrict__ rc,
                                                  const int* __restrict__ skip) {
    const int j = (xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (j >= Vc) return;  // whole wave
    // the flag, the row bounds and the row's mask in one round trip
    const int sk = load_flag(skip);
    const int e0 = rtp[j], e1 = rtp[j + 1];
    const float fx = vcl[j].z;
    if (sk) return;
    float ax = 0.f, ay = 0.f, az = 0.f;
    // a lane's entries e = start + lane, + 64, ... in that order, four of
    // them (rows up to 256 entries) in one round trip of loads
    for (int e = e0 + lane; e < e1; e += 4 * 64) {
        float w[4];
        int id[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const bool on = e + k * 64 < e1;
            w[k] = on ? rtw[e + k * 64] : 0.f;
            id[k] = on ? rti[e + k * 64] : 0;
        }
        float4 q[4];
#pragma unroll
        for (int k = 0; k < 4; k++) q[k] = e + k * 64 < e1 ? rf[id[k]] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (e + k * 64 < e1) {
                ax = ax + w[k] * q[k].x;
                ay = ay + w[k] * q[k].y;
                az = az + w[k] * q[k].z;
            }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        ax += __shfl_down(ax, off, 64);
        ay += __shfl_down(ay, off, 64);
        az += __shfl_down(az, off, 64);
    }
    if (lane == 0) rc[j] = make_float4(ax * fx, ay * fx, az * fx, 0.f);
}

// dst[i] += mask_i * (P src)[i]; dst2 (optional) receives the new value.
__global__ void __launch_bounds__(256) k_prolong_add(int N, float4* __restrict__ dst, float4* __restrict__ dst2,
                                                     const float4* __restrict__ src, const int4* __restrict__ ids,
                                                     const float4* __restrict__ w, const float4* __restrict__ maskv,
                                                     const int* __restrict__ skip, const float* __restrict__ amask) {
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (load_flag(skip)) return;
    const int4 id = ids[i];
    const float4 ww = w[i];
    const float4 a = src[id.x], b = src[id.y], c = src[id.z], d = src[id.w];
    const float vx = ww.x * a.x + ww.y * b.x + ww.z * c.x + ww.w * d.x;
    const float vy = ww.x * a.y + ww.y * b.y + ww.z * c.y + ww.w * d.y;
    const float vz = ww.x * a.z + ww.y * b.z + ww.z * c.z + ww.w * d.z;
    const float m = amask ? maskv[i].z * amask[i] : maskv[i].z;
    float4 o = dst[i];
    o.x = o.x + m * vx;
    o.y = o.y + m * vy;
    o.z = o.z + m * vz;
    dst[i] = o;
    if (dst2) dst2[i] = o;
}

// Row i of the scalar level matrix times a float4 field (columns ascending;
// loads issued four at a time, sums in row order).
__device__ __forceinline__ float3 spmv_row(int i, const int* __restrict__ ap, const int* __restrict__ aj,
                                           const float* __restrict__ av, const float4* __restrict__ x) {
    float sx = 0.f, sy = 0.f, sz = 0.f;
    const int q0 = ap[i], q1 = ap[i + 1];
    for (int q = q0; q < q1; q += 16) {
        const int n = q1 - q;
        int j[16];
        float a[16];
#pragma unroll
        for (int k = 0; k < 16; k++) {
            j[k] = k < n ? aj[q + k] : 0;
            a[k] = k < n ? av[q + k] : 0.f;
        }
        float4 xv[16];
#pragma unroll
        for (int k = 0; k < 16; k++) xv[k] = k < n ? x[j[k]] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 16; k++)
            if (k < n) {
                sx = sx + a[k] * xv[k].x;
                sy = sy + a[k] * xv[k].y;
                sz = sz + a[k] * xv[k].z;
            }
    }
    return make_float3(sx, sy, sz);
}

// One damped-Jacobi + Chebyshev sweep on A e = r for vertex i.  e is read at
// neighbours, so the new value goes into ep[i] (which only vertex i reads);
// the caller then swaps the roles of e and ep.  vcl = (A_ii, 0, vfix, m).
__device__ __forceinline__ void jacobi_cheb_row(int i, const int* __restrict__ ap, const int* __restrict__ aj,
                                                const float* __restrict__ av, float4* __restrict__ e,
                                                float4* __restrict__ ep, const float4* __restrict__ r,
                                                const float4* __restrict__ vcl, float omega, float relax,
                                                int first_zero) {
    const float4 c = vcl[i];
    const float d = c.x;
    const float fx = c.z;
    const float4 rr = r[i];
    if (first_zero) {
        float yx = fx * (relax * (rr.x / d));
        float yy = fx * (relax * (rr.y / d));
        float yz = fx * (relax * (rr.z / d));
        ep[i] = make_float4(omega * yx, omega * yy, omega * yz, 0.f);
        e[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const float3 Ae = spmv_row(i, ap, aj, av, e);
    const float4 ec = e[i], pc = ep[i];
    float yx = ec.x + fx * (relax * ((rr.x - Ae.x) / d));
    float yy = ec.y + fx * (relax * ((rr.y - Ae.y) / d));
    float yz = ec.z + fx * (relax * ((rr.z - Ae.z) / d));
    yx = omega * (yx - pc.x) + pc.x;
    yy = omega * (yy - pc.y) + pc.y;
    yz = omega * (yz - pc.z) + pc.z;
    ep[i] = make_float4(yx, yy, yz, 0.f);
}

// Row i of A e for a level stored as column-major ELL (w entries per row in
// the CSR row's order, rows shorter than w padded with (i, 0)): the row's
// loads are coalesced across lanes and e is read from LDS.  Same summation
// order as spmv_row (a padded entry adds a zero product).
__device__ __forceinline__ float3 spmv_row_ell(int i, int V, int w, const int* __restrict__ ajE,
                                               const float* __restrict__ avE, const float4* e) {
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (int k0 = 0; k0 < w; k0 += 8) {
        int j[8];
        float a[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const bool on = k0 + k < w;
            j[k] = on ? ajE[(size_t)(k0 + k) * V + i] : i;
            a[k] = on ? avE[(size_t)(k0 + k) * V + i] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float4 x = e[j[k]];
            sx = sx + a[k] * x.x;
            sy = sy + a[k] * x.y;
            sz = sz + a[k] * x.z;
        }
    }
    return make_float3(sx, sy, sz);
}

// jacobi_cheb_row on an ELL level whose e / ep live in LDS.
__device__ __forceinline__ void jacobi_cheb_row_ell(int i, int V, int w, const int* __restrict__ ajE,
                                                    const float* __restrict__ avE, float4* e, float4* ep,
                                                    const float4* __restrict__ r, const float4* __restrict__ vcl,
                                                    float omega, float relax, int first_zero) {
    const float4 c = vcl[i];
    const float d = c.x;
    const float fx = c.z;
    const float4 rr = r[i];
    if (first_zero) {
        float yx = fx * (relax * (rr.x / d));
        float yy = fx * (relax * (rr.y / d));
        float yz = fx * (relax * (rr.z / d));
        ep[i] = make_float4(omega * yx, omega * yy, omega * yz, 0.f);
        e[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const float3 Ae = spmv_row_ell(i, V, w, ajE, avE, e);
    const float4 ec = e[i], pc = ep[i];
    float yx = ec.x + fx * (relax * ((rr.x - Ae.x) / d));
    float yy = ec.y + fx * (relax * ((rr.y - Ae.y) / d));
    float yz = ec.z + fx * (relax * ((rr.z - Ae.z) / d));
    yx = omega * (yx - pc.x) + pc.x;
    yy = omega * (yy - pc.y) + pc.y;
    yz = omega * (yz - pc.z) + pc.z;
    ep[i] = make_float4(yx, yy, yz, 0.f);
}

__global__ void __launch_bounds__(256) k_spmv_smooth(int V, const int* __restrict__ ap, const int* __restrict__ aj,
                                                     const float* __restrict__ av, float4* __restrict__ e,
                                                     float4* __restrict__ ep, const float4* __restrict__ r,
                                                     const float4* __restrict__ vcl, float omega, float relax,
                                                     int first_zero, const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= V) return;
    jacobi_cheb_row(i, ap, aj, av, e, ep, r, vcl, omega, relax, first_zero);
}

// rr = vfix (r - A e).
__global__ void __launch_bounds__(256) k_spmv_resid(int V, const int* __restrict__ ap, const int* __restrict__ aj,
                                                    const float* __restrict__ av, const float4* __restrict__ e,
                                                    const float4* __restrict__ r, const float4* __restrict__ vcl,
                                                    float4* __restrict__ rr, const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const float3 Ae = spmv_row(i, ap, aj, av, e);
    const float4 rv = r[i];
    const float fx = vcl[i].z;
    rr[i] = make_float4(fx * (rv.x - Ae.x), fx * (rv.y - Ae.y), fx * (rv.z - Ae.z), 0.f);
}

__global__ void __launch_bounds__(256) k_zero4(int n, float4* __restrict__ a, const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

constexpr int kMaxDense = 1024;          // dense rows kept in LDS
constexpr size_t kMaxDenseEntries = 1u << 25;  // 256 MiB of block factors at most

__device__ __forceinline__ void wave_sum3(double& a0, double& a1, double& a2) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a0 += __shfl_down(a0, off, 64);
        a1 += __shfl_down(a1, off, 64);
        a2 += __shfl_down(a2, off, 64);
    }
}

// Coarsest solve e = W^T (W r), block by block (one block per connected
// component; W_b = C_b^-1 with A_b = C_b C_b^T, FP64 Cholesky).  r is in
// rb[c * ld + row] on entry and e on exit, yb is scratch (LDS, or device
// memory when the rows do not fit); rowinfo[row] = (first row of the block,
// block size, offset of W_b).  One wave per row with the lanes across the
// row (W_b and Wt_b = W_b^T both row-major), so all loads of a row are in
// flight at once; each row is summed by a fixed shuffle tree.  When no block
// has more than 64 rows (many small bodies) one LANE takes each row instead
// and sums it in index order: the rows of a 64-body batch then take three
// passes of the workgroup instead of 160 dependent row loads per wave.
constexpr int kLaneRowMax = 64;
__device__ __forceinline__ void dense_solve_block(int m, const int4* __restrict__ rowinfo,
                                                  const double* __restrict__ W, const double* __restrict__ Wt,
                                                  double* rb, double* yb, int ld, int maxblk) {
    if (maxblk <= kLaneRowMax) {
        for (int row = threadIdx.x; row < m; row += blockDim.x) {
            const int4 ri = rowinfo[row];
            const int i = row - ri.x;
            const double* wr = W + ri.z + (size_t)i * ri.y;
            const double* r0 = rb + ri.x;
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll 8
            for (int k = 0; k <= i; k++) {
                const double w = wr[k];
                a0 += w * r0[k];
                a1 += w * r0[ld + k];
                a2 += w * r0[2 * ld + k];
            }
            yb[row] = a0;
            yb[ld + row] = a1;
            yb[2 * ld + row] = a2;
        }
        __syncthreads();
        for (int row = threadIdx.x; row < m; row += blockDim.x) {
            const int4 ri = rowinfo[row];
            const int i = row - ri.x;
            const double* wr = Wt + ri.z + (size_t)i * ri.y;
            const double* y0 = yb + ri.x;
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll 8
            for (int k = i; k < ri.y; k++) {
                const double w = wr[k];
                a0 += w * y0[k];
                a1 += w * y0[ld + k];
                a2 += w * y0[2 * ld + k];
            }
            rb[row] = a0;
            rb[ld + row] = a1;
            rb[2 * ld + row] = a2;
        }
        __syncthreads();
        return;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int row = wave; row < m; row += nw) {
        const int4 ri = rowinfo[row];
        const int i = row - ri.x;
        const double* wr = W + ri.z + (size_t)i * ri.y;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int k = lane; k <= i; k += 64) {
            const double w = wr[k];
            a0 += w * rb[ri.x + k];
            a1 += w * rb[ld + ri.x + k];
            a2 += w * rb[2 * ld + ri.x + k];
        }
        wave_sum3(a0, a1, a2);
        if (lane == 0) {
            yb[row] = a0;
            yb[ld + row] = a1;
            yb[2 * ld + row] = a2;
        }
    }
    __syncthreads();
    for (int row = wave; row < m; row += nw) {
        const int4 ri = rowinfo[row];
        const int i = row - ri.x;
        const double* wr = Wt + ri.z + (size_t)i * ri.y;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int k = i + lane; k < ri.y; k += 64) {
            const double w = wr[k];
            a0 += w * yb[ri.x + k];
            a1 += w * yb[ld + ri.x + k];
            a2 += w * yb[2 * ld + ri.x + k];
        }
        wave_sum3(a0, a1, a2);
        if (lane == 0) {
            rb[row] = a0;
            rb[ld + row] = a1;
            rb[2 * ld + row] = a2;
        }
    }
    __syncthreads();
}

// The coarsest level's direct solve in one workgroup: gather r into dense
// rows, dense_solve_block, scatter e (pinned vertices get 0).  big: device
// scratch of 6 n doubles used when n > kMaxDense.
__device__ __forceinline__ void coarsest_direct(int V, int n, const int* __restrict__ dense,
                                                const int4* __restrict__ rowinfo, const double* __restrict__ W,
                                                const double* __restrict__ Wt, const float4* __restrict__ r,
                                                float4* __restrict__ e, double* big, int maxblk) {
    __shared__ double lds[6 * kMaxDense];
    double* rb = n <= kMaxDense ? lds : big;
    const int ld = n <= kMaxDense ? kMaxDense : n;
    double* yb = rb + 3 * ld;
    for (int i = threadIdx.x; i < V; i += blockDim.x) {
        const int d = dense[i];
        if (d >= 0) {
            const float4 q = r[i];
            rb[d] = q.x;
            rb[ld + d] = q.y;
            rb[2 * ld + d] = q.z;
        }
    }
    __syncthreads();
    dense_solve_block(n, rowinfo, W, Wt, rb, yb, ld, maxblk);
    for (int i = threadIdx.x; i < V; i += blockDim.x) {
        const int d = dense[i];
        e[i] = d >= 0 ? make_float4((float)rb[d], (float)rb[ld + d], (float)rb[2 * ld + d], 0.f)
                      : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();
}

// Coarsest level on its own (when it is not part of the one-workgroup tail).
__global__ void __launch_bounds__(256) k_coarse_solve(int Vc, int n, const int* __restrict__ dense,
                                                      const int4* __restrict__ rowinfo, const double* __restrict__ W,
                                                      const double* __restrict__ Wt, const float4* __restrict__ r,
                                                      float4* __restrict__ e, double* big, int maxblk,
                                                      const int* __restrict__ skip) {
    if (skip && *skip) return;
    coarsest_direct(Vc, n, dense, rowinfo, W, Wt, r, e, big, maxblk);
}

// The coarsest solve of many small bodies on its own: one wave per block
// (blk = first row, rows <= 64, W offset), one lane per row, the same sums in
// the same order as dense_solve_block's lane path, r and e addressed through
// the row's vertex.  Pinned vertices' e is zeroed beforehand.  Replaces the
// one-workgroup solve when the coarsest level is the whole tail: for the
// 64-body batch that workgroup ran 2,560 rows through device-memory vectors.
constexpr int kBlockRowsMax = 64;
__global__ void __launch_bounds__(kBlockRowsMax) k_coarse_solve_blocks(const int4* __restrict__ blk,
                                                                       const int* __restrict__ rowvert,
                                                                       const double* __restrict__ W,
                                                                       const double* __restrict__ Wt,
                                                                       const float4* __restrict__ r,
                                                                       float4* __restrict__ e,
                                                                       const int* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double rb[3][kBlockRowsMax], yb[3][kBlockRowsMax];
    const int4 b = blk[blockIdx.x];
    const int i = threadIdx.x;
    const bool on = i < b.y;
    const int v = on ? rowvert[b.x + i] : 0;
    if (on) {
        const float4 q = r[v];
        rb[0][i] = q.x;
        rb[1][i] = q.y;
        rb[2][i] = q.z;
    }
    __syncthreads();
    if (on) {
        const double* wr = W + b.z + (size_t)i * b.y;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int k = 0; k <= i; k++) {
            const double w = wr[k];
            a0 += w * rb[0][k];
            a1 += w * rb[1][k];
            a2 += w * rb[2][k];
        }
        yb[0][i] = a0;
        yb[1][i] = a1;
        yb[2][i] = a2;
    }
    __syncthreads();
    if (on) {
        const double* wr = Wt + b.z + (size_t)i * b.y;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int k = i; k < b.y; k++) {
            const double w = wr[k];
            a0 += w * yb[0][k];
            a1 += w * yb[1][k];
            a2 += w * yb[2][k];
        }
        e[v] = make_float4((float)a0, (float)a1, (float)a2, 0.f);
    }
}

// Small coarse levels of the linear V-cycle in ONE workgroup: level 0 of the
// descriptor array already holds its right-hand side r; the workgroup runs
// pre-smoothing, residual, restriction, ... down to the coarsest level, the
// coarsest solve, then prolongation and post-smoothing back up, with
// workgroup barriers between phases instead of kernel launches.  On return
// the first level's correction is in e0.
struct TailLevel {
    int V, nfree;
    int maxblk;  // largest block of the coarsest factor
    const int *ap, *aj;
    const float* av;
    const float4* vcl;
    float4 *e0, *e1, *r, *rr;
    const int *rtp, *rti;   // restriction into this level from the previous one
    const float* rtw;
    const int4* iids;       // prolongation from the next level into this one
    const float4* iw;
    const int* dense;
    const int4* rowinfo;
    const double *W, *Wt;
    double* big;  // dense-vector scratch when nfree > kMaxDense
    float rho;
    int ellw;     // first tail level only: ELL width of ajE / avE (0: CSR, vectors in device memory)
    const int* ajE;
    const float* avE;
};

// The first tail level keeps its correction (and the Chebyshev partner) in
// LDS up to this many vertices.
constexpr int kTailLdsV = 2048;

__global__ void __launch_bounds__(1024) k_coarse_tail(const TailLevel* __restrict__ tl, int n, int sweeps,
                                                      float relax, int direct, const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int tid = threadIdx.x, nt = blockDim.x;
    unsigned flips = 0u;  // bit q: level q's correction is in e1 (else e0)
    auto flip = [&](int q) { return (flips >> q) & 1u; };
    // level 0 with an ELL matrix: its e0 / e1 are the two LDS buffers
    __shared__ float4 e_lds[2][kTailLdsV];
    const bool lds0 = tl[0].ellw > 0;
    auto ecur = [&](int q) -> float4* {
        const TailLevel& L = tl[q];
        if (q == 0 && lds0) return e_lds[flip(0)];
        return flip(q) ? L.e1 : L.e0;
    };
    auto eprv = [&](int q) -> float4* {
        const TailLevel& L = tl[q];
        if (q == 0 && lds0) return e_lds[flip(0) ^ 1u];
        return flip(q) ? L.e0 : L.e1;
    };
    auto smooth = [&](int q, int ns, bool zero) {
        const TailLevel& L = tl[q];
        const bool ell = q == 0 && lds0;
        float omega = 1.0f;
        for (int k = 0; k < ns; k++) {
            omega = 4 / (4 - L.rho * L.rho * omega);
            float4* e = ecur(q);
            float4* ep = eprv(q);
            for (int i = tid; i < L.V; i += nt) {
                if (ell)
                    jacobi_cheb_row_ell(i, L.V, L.ellw, L.ajE, L.avE, e, ep, L.r, L.vcl, omega, relax, zero && k == 0);
                else
                    jacobi_cheb_row(i, L.ap, L.aj, L.av, e, ep, L.r, L.vcl, omega, relax, zero && k == 0);
            }
            __syncthreads();
            flips ^= 1u << q;
        }
    };
    // down
    for (int q = 0; q + 1 < n; q++) {
        const TailLevel& L = tl[q];
        smooth(q, sweeps, true);
        const float4* e = ecur(q);
        const bool ell = q == 0 && lds0;
        for (int i = tid; i < L.V; i += nt) {
            const float3 Ae = ell ? spmv_row_ell(i, L.V, L.ellw, L.ajE, L.avE, e) : spmv_row(i, L.ap, L.aj, L.av, e);
            const float4 rv = L.r[i];
            const float fx = L.vcl[i].z;
            L.rr[i] = make_float4(fx * (rv.x - Ae.x), fx * (rv.y - Ae.y), fx * (rv.z - Ae.z), 0.f);
        }
        __syncthreads();
        const TailLevel& C = tl[q + 1];
        {  // one wave per coarse vertex, lanes stride the row, shuffle tree
            const int wave = tid >> 6, lane = tid & 63, nw = nt >> 6;
            for (int j = wave; j < C.V; j += nw) {
                float ax = 0.f, ay = 0.f, az = 0.f;
                for (int t = C.rtp[j] + lane; t < C.rtp[j + 1]; t += 64) {
                    const float w = C.rtw[t];
                    const float4 v = L.rr[C.rti[t]];
                    ax = ax + w * v.x;
                    ay = ay + w * v.y;
                    az = az + w * v.z;
                }
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    ax += __shfl_down(ax, off, 64);
                    ay += __shfl_down(ay, off, 64);
                    az += __shfl_down(az, off, 64);
                }
                if (lane == 0) {
                    const float fx = C.vcl[j].z;
                    C.r[j] = make_float4(ax * fx, ay * fx, az * fx, 0.f);
                }
            }
        }
        __syncthreads();
    }
    // coarsest
    {
        const TailLevel& L = tl[n - 1];
        if (direct && L.nfree >= 0) {
            coarsest_direct(L.V, L.nfree, L.dense, L.rowinfo, L.W, L.Wt, L.r, L.e0, L.big, L.maxblk);
            flips &= ~(1u << (n - 1));
            __syncthreads();
        } else {
            smooth(n - 1, sweeps > 0 ? sweeps : 1, true);
        }
    }
    // up
    for (int q = n - 2; q >= 0; q--) {
        const TailLevel& L = tl[q];
        const float4* ec = ecur(q + 1);
        float4* e = ecur(q);
        float4* ep = eprv(q);
        for (int i = tid; i < L.V; i += nt) {
            const int4 id = L.iids[i];
            const float4 ww = L.iw[i];
            const float4 a = ec[id.x], b = ec[id.y], c = ec[id.z], d = ec[id.w];
            const float vx = ww.x * a.x + ww.y * b.x + ww.z * c.x + ww.w * d.x;
            const float vy = ww.x * a.y + ww.y * b.y + ww.z * c.y + ww.w * d.y;
            const float vz = ww.x * a.z + ww.y * b.z + ww.z * c.z + ww.w * d.z;
            const float m = L.vcl[i].z;
            float4 o = e[i];
            o.x = o.x + m * vx;
            o.y = o.y + m * vy;
            o.z = o.z + m * vz;
            e[i] = o;
            ep[i] = o;
        }
        __syncthreads();
        smooth(q, sweeps, false);
    }
    // leave the first level's correction in e0
    if (lds0 || flip(0)) {
        const TailLevel& L = tl[0];
        const float4* e = ecur(0);
        for (int i = tid; i < L.V; i += nt) L.e0[i] = e[i];
    }
}

// The coarse tail is linear in the residual of its first level (fixed
// Chebyshev weights, fixed transfers, the direct coarsest solve), and the
// three components go through the same scalar operator: e = Bt r with one
// V x V matrix Bt.  For a first tail level of at most kTailDenseMaxV
// vertices Bt is formed at set-up, column by column, by running the tail on
// unit right-hand sides (k_unit_rhs, k_store_cols), and each coarse
// correction is then one matrix-vector product: one wave per row, the row
// read coalesced, a fixed shuffle tree combining the lanes.  The tail's
// dozen dependent phases in one workgroup become one launch over the whole
// GPU; the product rounds differently from the sequence of sparse steps.
constexpr int kTailDenseMaxV = 2048;

__global__ void __launch_bounds__(256) k_unit_rhs(int V, float4* __restrict__ r, int c0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    r[i] = make_float4(i == c0 ? 1.f : 0.f, i == c0 + 1 ? 1.f : 0.f, i == c0 + 2 ? 1.f : 0.f, 0.f);
}

__global__ void __launch_bounds__(256) k_store_cols(int V, const float4* __restrict__ e, float* __restrict__ Bt,
                                                    int c0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const float4 v = e[i];
    float* row = Bt + (size_t)i * V;
    row[c0] = v.x;
    if (c0 + 1 < V) row[c0 + 1] = v.y;
    if (c0 + 2 < V) row[c0 + 2] = v.z;
}

constexpr int kTailBatch = 17;  // a 1,025-vertex first tail level in one batch
__global__ void __launch_bounds__(256) k_tail_gemv(int V, const float* __restrict__ Bt, const float4* __restrict__ r,
                                                   float4* __restrict__ e, const int* __restrict__ skip) {
    const int row = (xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (row >= V) return;
    const float* b = Bt + (size_t)row * V;
    float ax = 0.f, ay = 0.f, az = 0.f;
    // a lane's entries j = lane, lane + 64, ... in that order; kTailBatch of
    // them, with the flag, in one round trip of loads before their products
    // (rows up to kTailBatch * 64 entries; addresses past the row read its
    // last entry, whose product is dropped)
    const int sk = load_flag(skip);
    for (int j = lane; j < V; j += kTailBatch * 64) {
        float w[kTailBatch];
        float4 q[kTailBatch];
#pragma unroll
        for (int k = 0; k < kTailBatch; k++) {
            const int jc = min(j + k * 64, V - 1);
            w[k] = b[jc];
            q[k] = r[jc];
        }
        if (sk) return;
#pragma unroll
        for (int k = 0; k < kTailBatch; k++)
            if (j + k * 64 < V) {
                ax = ax + w[k] * q[k].x;
                ay = ay + w[k] * q[k].y;
                az = az + w[k] * q[k].z;
            }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        ax += __shfl_down(ax, off, 64);
        ay += __shfl_down(ay, off, 64);
        az += __shfl_down(az, off, 64);
    }
    if (lane == 0) e[row] = make_float4(ax, ay, az, 0.f);
}

inline int nblk(int n, int bs) { return (n + bs - 1) / bs; }

// ---------------------------------------------------------------------------
// Host structures
// ---------------------------------------------------------------------------
template <class T>
int dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
        return MGSB_ERR_NOMEM;
    }
    return MGSB_OK;
}

// Device address of `scratch` per device ordinal, looked up once per device
// with that device current (at mgsb_create).
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxScratchDevices] = {};

int scratch_setup(int device) {
    if (device < 0 || device >= kMaxScratchDevices) {
        set_error("device ordinal beyond the scratch table (64 devices)");
        return MGSB_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return MGSB_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        return MGSB_ERR_HIP;
    }
    g_scratch[device] = static_cast<double*>(ptr);
    return MGSB_OK;
}

double* scratch_of(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch[device];
}

struct DevLevel {
    HostLevel h;
    int V = 0, T = 0, S = 0;
    float4 *x = nullptr, *xprev = nullptr, *s = nullptr, *xlast = nullptr, *v = nullptr, *vc = nullptr;
    int4 *tet = nullptr, *slot = nullptr;
    float *B = nullptr, *vol = nullptr;
    int* rp = nullptr;
    float3* cf = nullptr;  // corner forces in vertex-sorted (CSR) order
    // tiled local step (HostLevel::prp ...): per-(tile, vertex) partials
    int ntiles = 0, npart = 0;
    int *prp = nullptr, *tvo = nullptr, *tvs = nullptr, *tlr = nullptr;
    int4* tetp = nullptr;  // tets with each corner's position in its tile's corner array (nullptr: not tiled)
    int2* tpad = nullptr;  // padded tile rows (tile_row), tstride entries per tile
    int tstride = 0, tover = 0;
    float3 *pcf = nullptr, *pcf2 = nullptr;
    // fused sweeps (k_sweep_fused): per corner its tile-local vertex index and
    // corner position; the tiles' vertex lists; a third position buffer
    // (nullptr: the level does not run fused sweeps)
    int4* tetq = nullptr;
    int2* tvqp = nullptr;  // per (tile, vertex), padded as tpad: vertex id | owner << 31, prp row start | count << 26
    // level 0 of a V-cycle with fused sweeps: each tile vertex's binding into
    // level 1 (coarse tet corners, weights), padded as tpad
    int4* tiid = nullptr;
    float4* tiw = nullptr;
    float4* xz = nullptr;
    int4 *tri = nullptr, *tslot = nullptr;
    int *trp = nullptr, *trid = nullptr;
    float4 *tcf = nullptr, *tn = nullptr, *vn = nullptr;
    int* cflag = nullptr;
    float *cdepth = nullptr, *cenergy = nullptr;
    // V-cycle state: residual / correction of this level, Dirichlet mask of
    // the correction in vcl = (m/h^2, diag, vfix, m), restriction into this
    // level from the next finer one (P^T as CSR), coarsest dense factor.
    float4 *r = nullptr, *e = nullptr, *ep = nullptr, *rr = nullptr, *vcl = nullptr;
    int *rtp = nullptr, *rti = nullptr;
    float* rtw = nullptr;
    int *ap = nullptr, *aj = nullptr;  // scalar level matrix A = m/h^2 + L (coarse levels)
    float* av = nullptr;
    float* tailB = nullptr;            // first tail level: the tail as a dense V x V matrix (or nullptr)
    int* ajE = nullptr;                // ... as column-major ELL (first level of the coarse tail)
    float* avE = nullptr;
    int ellw = 0;
    std::vector<float> adiag;          // its diagonal
    float4 *U1 = nullptr, *U2 = nullptr;  // corrections u = x - s of the last two steps
    float* cmask = nullptr;               // 0 where a collider acts (active set) at the last residual
    float3* cf2 = nullptr;                // level 0: corner forces at s (temporally warm-started start)
    int cur = 0;                       // which of e / ep holds the correction
    int* dense = nullptr;
    double *W = nullptr, *Wt = nullptr;  // coarsest block factors W_b = C_b^-1 and transposes
    int4* rowinfo = nullptr;             // per dense row: (block start, block size, W offset, 0)
    double* big = nullptr;               // dense-vector scratch (6 n) when n > kMaxDense
    float* tk = nullptr;                 // per-tet stiffness (nullptr: p.stiffness)
    int nfree = -1;
    int maxblk = 0;                      // largest block of the coarsest factor (rows)
    int4* blk = nullptr;                 // per block: (first row, rows, W offset, 0)
    int* rowvert = nullptr;              // vertex of each dense row
    int nblocks = 0;
    std::vector<float> vfix;
    double rho = 0.0;
    std::vector<void*> owned;
};

struct LevelMap {  // between level l (fine) and l+1 (coarse)
    int4* iids = nullptr;   // fine vertex -> coarse tet corners
    float4* iw = nullptr;
    int4* aids = nullptr;   // coarse vertex -> fine tet corners
    float4* aw = nullptr;
    std::vector<int32_t> h_iids, h_aids;
    std::vector<float> h_iw, h_aw;
};

}  // namespace

// Pieces of a V-cycle step: the start (with or without the learned stage),
// the two half-cycles, and the finish (with or without the coarse delta that
// only the learned stage reads as a feature).
enum { kPieceStart = 0, kPiecePre = 1, kPiecePost = 2, kPieceFinish = 3, kPieceStartGated = 4, kPieceFinishLite = 5 };

struct mgsb_solver {
    mgsb_params p{};
    float relax_key = NAN, relax_hi = 0.f, relax_lo = 0.f;  // p.relax as a double word (launch_vert)
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<DevLevel> L;
    std::vector<LevelMap> M;
    DevColliders hcol{};
    DevColliders* dcol = nullptr;
    double* dpart = nullptr;  // reduction partials
    size_t npart = 0;
    double* dred = nullptr;   // reduction results (16)
    bool graph_valid = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
    size_t ev_used = 0;
    double time_ms = 0.0;
    int64_t time_launches = 0;
    mgsb_step_stats last{};
    VState* vstate = nullptr;
    double* vpart = nullptr;  // partials of the level-0 residual norm
    int* hmirror = nullptr;      // host-mapped 64-bit word: checks run (low), last verdict (high)
    unsigned* seqc = nullptr;    // device count of check kernels run
    unsigned chk_launched = 0;   // check kernels enqueued (graph replays included)
    // V-cycle steps run as pieces (start, pre-smoothing half, coarse + post
    // half, finish).  The host launches the start, the halves the previous
    // step needed and the finish as one graph, then reads the convergence
    // mirror and adds halves (and a finish) only when the step needs them.
    // Each distinct piece sequence is captured once.  Fewer, longer graphs:
    // on MI355X a dependent kernel inside a graph costs about 2 us, while
    // back-to-back launches of a one-kernel graph run at about 16 us each.
    struct PieceGraph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        unsigned checks = 0;  // check kernels in the sequence
    };
    std::map<std::string, PieceGraph> pseq;
    bool precaptured = false;  // the finite set of first rounds and follow-ups is in pseq
    // the piece being enqueued is followed by a coarse-first half in the same
    // round (enqueue_seq); a fused warm start then leaves its judge to that
    // half's restriction (k_restrict_judged), whose arguments are kept here
    bool next_is_post = false;
    int judge_folded = 0;  // 0 none pending, else the judge's check mode
    int judge_nb = 0;      // the dual residual's block count
    int halves_guess = 4;
    // learned schedule: after kGateAfter consecutive steps whose learned
    // warm start the judge undid, steps start without the MLP stage for
    // gate_skip steps (kGateFirst, doubling up to kGateMax after each further
    // such run; a kept stage resets both)
    int gate_rejects = 0, gate_skip = 0, gate_left = 0;
    // the stage runs while its steps end after the first half-cycle or need
    // fewer half-cycles than the last step without it (off_halves, -1: none
    // yet); otherwise, and every kGateProbe steps with it, the next step runs
    // without it, and the stage resumes after a step without it that needed
    // two or more
    bool stage_on = true;
    int off_halves = -1, since_probe = 0;
    // bodies: connected components of level 0 (body_of[l][v] on every level,
    // through the coarse-in-fine bindings); when they are contiguous vertex
    // ranges the V-cycle checks each body's own relative residual
    int nbodies = 1;
    bool body_checks = false;
    std::vector<std::vector<int32_t>> body_of;
    std::vector<int32_t> h_vb;
    int* vb = nullptr;
    double *g2 = nullptr, *bodysq = nullptr, *res0b = nullptr;
    unsigned* bcount = nullptr;
    TailLevel* tail = nullptr; // descriptors of the one-workgroup coarse tail
    int tail_from = -1;        // first level handled by the tail kernel (-1: none)
    // learned coarse correction (level 1)
    Mlp* mlp = nullptr;
    float *feat = nullptr, *pred = nullptr, *norm = nullptr;
    float4 *epred = nullptr, *X1 = nullptr;
    float4 *D1 = nullptr, *D2 = nullptr;  // fine level: the last two steps' corrections beyond the extrapolation
    float *recX = nullptr, *recT = nullptr;  // recorded samples (raw)
    int rec_cap = 0, rec_steps = 0;           // in steps of V1 samples
    bool recording = false;
    std::vector<void*> owned;

    ~mgsb_solver() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto& kv : pseq) {
            if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
            if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
        }
        if (hmirror) (void)hipHostFree(hmirror);
        for (auto& e : ev_pool) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        for (auto& lv : L) {
            for (void* q : lv.owned) (void)hipFree(q);
            if (lv.W) (void)hipFree(lv.W);
            if (lv.Wt) (void)hipFree(lv.Wt);
            if (lv.rowinfo) (void)hipFree(lv.rowinfo);
            if (lv.big) (void)hipFree(lv.big);
            if (lv.blk) (void)hipFree(lv.blk);
            if (lv.rowvert) (void)hipFree(lv.rowvert);
        }
        for (void* q : owned) (void)hipFree(q);
        if (mlp) mlp_destroy(mlp);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <class T>
int up(mgsb_solver* s, std::vector<void*>& own, T** d, const T* h, size_t n) {
    int rc = dalloc(d, n);
    if (rc) return rc;
    own.push_back((void*)*d);
    if (h && n) HIPCHK(hipMemcpyAsync(*d, h, n * sizeof(T), hipMemcpyHostToDevice, s->stream));
    return MGSB_OK;
}

int upload_vc(mgsb_solver* s, DevLevel& d) {
    std::vector<float4> vc(d.V);
    const float dt = s->p.dt;
    for (int i = 0; i < d.V; i++) vc[i] = make_float4(d.h.mass[i] / (dt * dt), d.h.diag[i], d.h.fix[i], d.h.mass[i]);
    HIPCHK(hipMemcpyAsync(d.vc, vc.data(), sizeof(float4) * d.V, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return MGSB_OK;
}

int upload_level(mgsb_solver* s, DevLevel& d) {
    const HostLevel& h = d.h;
    d.V = h.V;
    d.T = h.T;
    d.S = h.S;
    int rc;
    std::vector<float4> x4(h.V);
    for (int i = 0; i < h.V; i++) x4[i] = make_float4(h.X[i * 3], h.X[i * 3 + 1], h.X[i * 3 + 2], 0.f);
    std::vector<float4> zero(h.V, make_float4(0.f, 0.f, 0.f, 0.f));
    if ((rc = up(s, d.owned, &d.x, x4.data(), h.V))) return rc;
    if ((rc = up(s, d.owned, &d.xprev, x4.data(), h.V))) return rc;
    if ((rc = up(s, d.owned, &d.s, x4.data(), h.V))) return rc;
    if ((rc = up(s, d.owned, &d.xlast, x4.data(), h.V))) return rc;
    if ((rc = up(s, d.owned, &d.v, zero.data(), h.V))) return rc;
    if ((rc = up<float4>(s, d.owned, &d.vc, nullptr, h.V))) return rc;
    std::vector<int4> t4(h.T), s4(h.T);
    for (int t = 0; t < h.T; t++) {
        t4[t] = make_int4(h.tets[t * 4], h.tets[t * 4 + 1], h.tets[t * 4 + 2], h.tets[t * 4 + 3]);
        s4[t] = make_int4(h.slot[t * 4], h.slot[t * 4 + 1], h.slot[t * 4 + 2], h.slot[t * 4 + 3]);
    }
    if ((rc = up(s, d.owned, &d.tet, t4.data(), h.T))) return rc;
    if ((rc = up(s, d.owned, &d.slot, s4.data(), h.T))) return rc;
    std::vector<float> bs((size_t)h.T * 9);
    for (int t = 0; t < h.T; t++)
        for (int q = 0; q < 9; q++) bs[(size_t)q * h.T + t] = h.B[(size_t)t * 9 + q];
    if ((rc = up(s, d.owned, &d.B, bs.data(), bs.size()))) return rc;
    if ((rc = up(s, d.owned, &d.vol, h.vol.data(), h.T))) return rc;
    if ((rc = up(s, d.owned, &d.rp, h.rp.data(), h.rp.size()))) return rc;
    if ((rc = up<float3>(s, d.owned, &d.cf, nullptr, (size_t)h.T * 4))) return rc;
    d.ntiles = (int)h.tvo.size() - 1;
    d.npart = h.tvo.back();
    if ((rc = up(s, d.owned, &d.prp, h.prp.data(), h.prp.size()))) return rc;
    if ((rc = up(s, d.owned, &d.tvo, h.tvo.data(), h.tvo.size()))) return rc;
    if ((rc = up(s, d.owned, &d.tvs, h.tvs.data(), h.tvs.size()))) return rc;
    if ((rc = up(s, d.owned, &d.tlr, h.tlr.data(), h.tlr.size()))) return rc;
    if ((rc = up<float3>(s, d.owned, &d.pcf, nullptr, (size_t)d.npart))) return rc;
    {
        // padded tile rows: tile k's lane l < its vertex count holds (row
        // start, row end, slot, 1) of tile vertex l, the rest zeros
        int maxnv = 0;
        for (int k = 0; k < d.ntiles; k++) maxnv = std::max(maxnv, h.tvo[(size_t)k + 1] - h.tvo[(size_t)k]);
        d.tstride = std::min(kTileTets, std::max(16, (maxnv + 15) / 16 * 16));
        d.tover = maxnv > kTileTets ? 1 : 0;
        std::vector<int2> tp((size_t)std::max(1, d.ntiles) * d.tstride, make_int2(0, 0));
        for (int k = 0; k < d.ntiles; k++) {
            const int v0 = h.tvo[(size_t)k], nv = std::min(kTileTets, h.tvo[(size_t)k + 1] - v0);
            for (int l = 0; l < nv; l++) {
                // rows of the tile's corner array: 4 kTileTets entries, so 16 bits each
                const int r0 = h.tlr[(size_t)v0 + k + l], r1 = h.tlr[(size_t)v0 + k + l + 1];
                tp[(size_t)k * d.tstride + l] = make_int2(h.tvs[(size_t)v0 + l], r0 | r1 << 16);
            }
        }
        if ((rc = up(s, d.owned, &d.tpad, tp.data(), tp.size()))) return rc;
    }
    if (h.V <= (1 << 22)) {
        std::vector<int4> tp(h.T);
        for (int t = 0; t < h.T; t++) {
            int c[4];
            for (int j = 0; j < 4; j++) {
                const uint32_t pos = (h.tls[(size_t)t * 2 + (j >> 1)] >> (16 * (j & 1))) & 0xffffu;
                c[j] = (int)((uint32_t)h.tets[(size_t)t * 4 + j] | (pos << 22));
            }
            tp[t] = make_int4(c[0], c[1], c[2], c[3]);
        }
        if ((rc = up(s, d.owned, &d.tetp, tp.data(), tp.size()))) return rc;
        // fused sweeps hold a tile's vertices in LDS and need every vertex in a tile
        int maxnv = 0;
        for (int k = 0; k < d.ntiles; k++) maxnv = std::max(maxnv, h.tvo[(size_t)k + 1] - h.tvo[(size_t)k]);
        bool covered = true;
        for (int v = 0; v < h.V && covered; v++) covered = h.prp[(size_t)v + 1] > h.prp[(size_t)v];
        // tvqp's packing: partial rows below 2^kTvqRowBits, at most 31 tiles a vertex
        bool packs = (int64_t)h.prp.back() < ((int64_t)1 << kTvqRowBits);
        for (int v = 0; v < h.V && packs; v++) packs = h.prp[(size_t)v + 1] - h.prp[(size_t)v] < 32;
        if (maxnv <= kFuseMaxNv && covered && packs) {
            std::vector<int4> tq(h.T);
            for (int t = 0; t < h.T; t++) {
                int c[4];
                for (int j = 0; j < 4; j++) {
                    const uint32_t sh = 16 * (j & 1);
                    const uint32_t pos = (h.tls[(size_t)t * 2 + (j >> 1)] >> sh) & 0xffffu;
                    const uint32_t lv = (h.tlv[(size_t)t * 2 + (j >> 1)] >> sh) & 0xffffu;
                    c[j] = (int)(lv | (pos << 16));
                }
                tq[t] = make_int4(c[0], c[1], c[2], c[3]);
            }
            if ((rc = up(s, d.owned, &d.tetq, tq.data(), tq.size()))) return rc;
            std::vector<int2> tv((size_t)std::max(1, d.ntiles) * d.tstride, make_int2(0, 0));
            for (int k = 0; k < d.ntiles; k++) {
                const int v0 = h.tvo[(size_t)k], nv = h.tvo[(size_t)k + 1] - v0;
                for (int l = 0; l < nv; l++) {
                    const size_t q = (size_t)v0 + l;
                    const int v = h.tvid[q];
                    const int r0 = h.prp[(size_t)v], r1 = h.prp[(size_t)v + 1];
                    // the first tile of a row owns the vertex
                    const unsigned own = h.tvs[q] == r0 ? 0x80000000u : 0u;
                    tv[(size_t)k * d.tstride + l] = make_int2((int)((unsigned)v | own), r0 | (r1 - r0) << kTvqRowBits);
                }
            }
            if ((rc = up(s, d.owned, &d.tvqp, tv.data(), tv.size()))) return rc;
        }
    }
    std::vector<int4> tr4(std::max(1, h.S)), ts4(std::max(1, h.S));
    for (int t = 0; t < h.S; t++) {
        tr4[t] = make_int4((int)h.tris[t * 3], (int)h.tris[t * 3 + 1], (int)h.tris[t * 3 + 2], 0);
        ts4[t] = make_int4(h.tslot[t * 3], h.tslot[t * 3 + 1], h.tslot[t * 3 + 2], 0);
    }
    // vertex -> triangle ids in CSR order (for vertex normals)
    std::vector<int32_t> trid(std::max<size_t>(1, (size_t)h.S * 3));
    for (int t = 0; t < h.S; t++)
        for (int c = 0; c < 3; c++) trid[h.tslot[t * 3 + c]] = t;
    if ((rc = up(s, d.owned, &d.tri, tr4.data(), tr4.size()))) return rc;
    if ((rc = up(s, d.owned, &d.tslot, ts4.data(), ts4.size()))) return rc;
    if ((rc = up(s, d.owned, &d.trp, h.trp.data(), h.trp.size()))) return rc;
    if ((rc = up(s, d.owned, &d.trid, trid.data(), trid.size()))) return rc;
    std::vector<float4> tz(std::max<size_t>(1, (size_t)h.S * 3), make_float4(0.f, 0.f, 0.f, 0.f));
    if ((rc = up(s, d.owned, &d.tcf, tz.data(), tz.size()))) return rc;
    if ((rc = up<float4>(s, d.owned, &d.tn, nullptr, std::max(1, h.S)))) return rc;
    if ((rc = up<float4>(s, d.owned, &d.vn, nullptr, h.V))) return rc;
    std::vector<int> iz(h.V, 0);
    std::vector<float> fz(h.V, 0.f);
    if ((rc = up(s, d.owned, &d.cflag, iz.data(), h.V))) return rc;
    if ((rc = up(s, d.owned, &d.cdepth, fz.data(), h.V))) return rc;
    if ((rc = up(s, d.owned, &d.cenergy, fz.data(), h.V))) return rc;
    return upload_vc(s, d);
}

int push_colliders(mgsb_solver* s) {
    HIPCHK(hipMemcpyAsync(s->dcol, &s->hcol, sizeof(DevColliders), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return MGSB_OK;
}

int pull_colliders(mgsb_solver* s) {
    HIPCHK(hipMemcpyAsync(&s->hcol, s->dcol, sizeof(DevColliders), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return MGSB_OK;
}

// coarse_first applies to steps with checks: the first half-cycle is a
// coarse-correction half, and the half-cycle cap is the first coarse half
// plus max_cycles cycles
bool coarse_first(const mgsb_solver* s) { return s->p.coarse_first != 0 && s->p.check_every > 0; }
int max_halves(const mgsb_solver* s) { return 2 * s->p.max_cycles + (coarse_first(s) ? 1 : 0); }

int push_vstate(mgsb_solver* s) {
    if (!s->vstate) return MGSB_OK;
    VState v{};
    v.check = s->p.check_every;
    v.tol = s->p.tol;
    v.maxh = max_halves(s);
    v.cfirst = coarse_first(s) ? 1 : 0;
    v.guard = s->p.stall_guard ? 1 : 0;
    v.slack = s->p.learn_slack > 0.f ? (double)s->p.learn_slack : 1.0;
    HIPCHK(hipMemcpyAsync(s->vstate, &v, sizeof(VState), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return MGSB_OK;
}

int block_tet(const mgsb_solver* s) { return s->p.block_tet > 0 ? s->p.block_tet : 256; }
int block_vert(const mgsb_solver* s) { return s->p.block_vert > 0 ? s->p.block_vert : 256; }
// lanes per vertex of the gathering vertex kernels (k_vert, k_resid,
// k_resid_dual) as a shift: vert_lanes 1, 2, 4 or 8 (0: 4)
int vert_lsh(const mgsb_solver* s) {
    // default: one lane for the ~2.5 partials a vertex has in the tiled
    // layout, four for its ~20 corners in the per-corner one
    const int l = s->p.vert_lanes > 0 ? s->p.vert_lanes : (s->p.tile_gather ? 1 : 4);
    return l >= 8 ? 3 : l >= 4 ? 2 : l >= 2 ? 1 : 0;
}
constexpr int kMaxVertLanes = 8;

// --- launch helpers (all on s->stream) --------------------------------------
int timed_begin(mgsb_solver* s, int kernel) {
    if (!s->timing || kernel != 0) return MGSB_OK;
    if (s->ev_used == s->ev_pool.size()) {
        hipEvent_t a, b;
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
        s->ev_pool.push_back({a, b});
    }
    HIPCHK(hipEventRecord(s->ev_pool[s->ev_used].first, s->stream));
    return MGSB_OK;
}
int timed_end(mgsb_solver* s, int kernel) {
    if (!s->timing || kernel != 0) return MGSB_OK;
    HIPCHK(hipEventRecord(s->ev_pool[s->ev_used].second, s->stream));
    s->ev_used++;
    return MGSB_OK;
}

TileMeta tile_meta(const DevLevel& d) { return TileMeta{d.tpad, d.tvo, d.tvs, d.tlr, d.tstride, d.tover}; }

// split: x holds the in-step correction u and the kernel adds F(s) (V-cycle steps).
template <int MODE>
void launch_tet_mode(mgsb_solver* s, DevLevel& d, int nb, int bs, const float4* xx, int ntri, const int* skip,
                     const float4* sref, bool tiled) {
    if (tiled) {
        auto k = ntri ? k_tet_tiled<MODE, true> : k_tet_tiled<MODE, false>;
        k<<<d.ntiles + nblk(ntri, kTileTets), kTileTets, 0, s->stream>>>(
            d.T, d.ntiles, xx, d.tetp, d.B, d.vol, tile_meta(d), d.pcf, s->p.stiffness, ntri, d.tri,
            d.tslot, d.tcf, s->dcol, skip, sref, d.tk);
        return;
    }
    auto k = ntri ? k_tet<MODE, true> : k_tet<MODE, false>;
    k<<<nb, bs, 0, s->stream>>>(d.T, xx, d.tet, d.B, d.vol, d.slot, d.cf, s->p.stiffness, ntri, d.tri, d.tslot,
                                d.tcf, s->dcol, skip, sref, d.tk);
}

// The per-vertex gathers read the (tile, vertex) partials of the tiled local
// step, or the per-corner forces.
bool tiled(const mgsb_solver* s, const DevLevel& d) { return s->p.tile_gather != 0 && d.tetp != nullptr; }
const int* gath_rp(const mgsb_solver* s, const DevLevel& d) { return tiled(s, d) ? d.prp : d.rp; }
float3* gath_cf(const mgsb_solver* s, const DevLevel& d) { return tiled(s, d) ? d.pcf : d.cf; }
float3* gath_cf2(const mgsb_solver* s, const DevLevel& d) { return tiled(s, d) ? d.pcf2 : d.cf2; }

int launch_tet(mgsb_solver* s, DevLevel& d, int mode, bool with_tri, const float4* xsrc = nullptr,
               const int* skip = nullptr, bool split = false, bool corners = false) {
    const float4* sref = split ? d.s : nullptr;
    const int bs = block_tet(s);
    const int ntri = with_tri ? d.S : 0;
    const int n = d.T + ntri;
    const float4* xx = xsrc ? xsrc : d.x;
    // timing covers the finest level's local step only (kernel 0)
    const int kid = (&d == &s->L[0] && mode != kTetModeLinear) ? 0 : 1;
    int rc = timed_begin(s, kid);
    if (rc) return rc;
    const int nb = nblk(n, bs);
    const bool til = tiled(s, d) && !corners;  // corners: the per-corner layout (test readback)
    switch (mode) {
        case kTetModePolar: launch_tet_mode<kTetModePolar>(s, d, nb, bs, xx, ntri, skip, sref, til); break;
        case kTetModeAc: launch_tet_mode<kTetModeAc>(s, d, nb, bs, xx, ntri, skip, sref, til); break;
        case kTetModeStable: launch_tet_mode<kTetModeStable>(s, d, nb, bs, xx, ntri, skip, sref, til); break;
        default: launch_tet_mode<kTetModeLinear>(s, d, nb, bs, xx, ntri, skip, sref, til);
    }
    HIPCHK(hipGetLastError());
    return timed_end(s, kid);
}

// MGSB_IF_* -> per-tet kernel mode
int tet_mode_of(int if_mode) {
    return if_mode == MGSB_IF_AC ? kTetModeAc : (if_mode == MGSB_IF_STABLE ? kTetModeStable : kTetModePolar);
}

// Local step at the temporal extrapolation (forces to cf) and at s (cf2) in
// one launch of 2T lanes.
int launch_tet_dual(mgsb_solver* s, DevLevel& d) {
    if (tiled(s, d)) {
        const int nb = 2 * d.ntiles;
        auto k = tet_mode_of(s->p.if_mode) == kTetModePolar
                     ? k_tet_dual_tiled<kTetModePolar>
                     : (tet_mode_of(s->p.if_mode) == kTetModeAc ? k_tet_dual_tiled<kTetModeAc>
                                                                  : k_tet_dual_tiled<kTetModeStable>);
        k<<<nb, kTileTets, 0, s->stream>>>(d.T, d.ntiles, d.x, d.s, d.tetp, d.B, d.vol, tile_meta(d), d.pcf,
                                           d.pcf2, s->p.stiffness, d.tk);
        HIPCHK(hipGetLastError());
        return MGSB_OK;
    }
    const int bs = block_tet(s);
    const int nb = nblk(2 * d.T, bs);
    switch (tet_mode_of(s->p.if_mode)) {
        case kTetModePolar:
            k_tet_dual<kTetModePolar><<<nb, bs, 0, s->stream>>>(d.T, d.x, d.s, d.tet, d.B, d.vol, d.slot, d.cf, d.cf2,
                                                                s->p.stiffness, d.tk);
            break;
        case kTetModeAc:
            k_tet_dual<kTetModeAc><<<nb, bs, 0, s->stream>>>(d.T, d.x, d.s, d.tet, d.B, d.vol, d.slot, d.cf, d.cf2,
                                                             s->p.stiffness, d.tk);
            break;
        default:
            k_tet_dual<kTetModeStable><<<nb, bs, 0, s->stream>>>(d.T, d.x, d.s, d.tet, d.B, d.vol, d.slot, d.cf,
                                                                 d.cf2, s->p.stiffness, d.tk);
    }
    HIPCHK(hipGetLastError());
    return MGSB_OK;
}

// The double a float parameter stands for: the double nearest to the
// shortest decimal that rounds to it (0.6f -> 0.6), as the constants the
// reference writes in its kernels are double literals (PDSolver.cu:219).
double decimal_double(float f) {
    char buf[32];
    for (int prec = 1; prec <= 9; prec++) {
        std::snprintf(buf, sizeof buf, "%.*g", prec, (double)f);
        const double d = std::strtod(buf, nullptr);
        if ((float)d == f) return d;
    }
    return (double)f;
}

void relax_consts(mgsb_solver* s) {
    if (!(s->p.relax == s->relax_key)) {  // the relaxation constant as a double word
        s->relax_key = s->p.relax;
        const double r = decimal_double(s->p.relax);
        s->relax_hi = (float)r;
        s->relax_lo = (float)(r - (double)s->relax_hi);
    }
}

// the stall guard's flag for the V-cycle's fine sweeps (split positions)
const int* calm_flag(const mgsb_solver* s, int split) {
    return (split && s->vstate && s->p.stall_guard) ? &s->vstate->calm : nullptr;
}

int launch_vert(mgsb_solver* s, DevLevel& d, float omega, int flags, const int* skip = nullptr, int split = 0) {
    const int bs = block_vert(s), lsh = vert_lsh(s);
    relax_consts(s);
    k_vert<<<nblk(d.V << lsh, bs), bs, 0, s->stream>>>(d.V, d.x, d.xprev, d.s, d.vc, gath_rp(s, d), gath_cf(s, d),
                                                        d.trp, d.tcf, s->dcol,
                                                        omega, s->relax_hi, s->relax_lo, flags, d.cflag, d.cdepth,
                                                        d.cenergy, skip, split, lsh, calm_flag(s, split));
    HIPCHK(hipGetLastError());
    return MGSB_OK;
}

// warm: split positions start from the temporal extrapolation (U1, U2);
// epred: plus the prolongated learned coarse correction (level 0 only)
int launch_st(mgsb_solver* s, DevLevel& d, int advance, int split = 0, bool warm = false,
              const float4* epred = nullptr) {
    const int bs = block_vert(s);
    k_st<<<nblk(d.V, bs), bs, 0, s->stream>>>(d.V, d.x, d.v, d.xlast, d.s, d.xprev, d.vc, s->p.dt, s->p.damping,
                                               s->p.gravity[0], s->p.gravity[1], s->p.gravity[2], s->dcol, advance,
                                               split, warm ? d.U1 : nullptr, warm ? d.U2 : nullptr, epred,
                                               epred ? s->M[0].iids : nullptr, epred ? s->M[0].iw : nullptr);
    HIPCHK(hipGetLastError());
    return MGSB_OK;
}

int launch_vel(mgsb_solver* s, DevLevel& d) {
    const int bs = block_vert(s);
    k_vel<<<nblk(d.V, bs), bs, 0, s->stream>>>(d.V, d.x, d.xlast, d.v, s->p.dt);
    HIPCHK(hipGetLastError());
    return MGSB_OK;
}

// One reference PD iteration on a level (runClearTemp, DCD, IF, omega, POS).
int pd_iteration(mgsb_solver* s, DevLevel& d, int if_mode, float omega, bool tri_dcd, bool last) {
    int rc = launch_tet(s, d, tet_mode_of(if_mode), tri_dcd);
    if (rc) return rc;
    int flags = tri_dcd ? kVertTri : kVertPoint;
    if (last) flags |= kVertWriteCol;
    return launch_vert(s, d, omega, flags);
}

// PDSolver::Step (PDSolver.cpp:243-293).
int enqueue_step_single(mgsb_solver* s) {
    DevLevel& d = s->L[0];
    int rc = launch_st(s, d, 1);
    if (rc) return rc;
    float omega = 1.0f;
    const float rho = s->p.rho;
    for (int i = 0; i < s->p.iters; i++) {
        omega = 4 / (4 - rho * rho * omega);
        if ((rc = pd_iteration(s, d, s->p.if_mode, omega, false, i == s->p.iters - 1))) return rc;
    }
    return launch_vel(s, d);
}

// PDSolver_MG::Step (PDSolver_MG.cpp:183-253) between levels 1 (coarse) and 0.
int enqueue_step_twolevel(mgsb_solver* s) {
    DevLevel& f = s->L[0];
    DevLevel& c = s->L[1];
    LevelMap& m = s->M[0];
    int rc;
    if ((rc = launch_st(s, c, 1))) return rc;
    if ((rc = launch_st(s, f, 0))) return rc;
    const float rho = s->p.rho;
    float omega = 1.0f;
    for (int i = 0; i < s->p.iters_coarse; i++) {
        omega = 4 / (4 - rho * rho * omega);
        if ((rc = pd_iteration(s, c, s->p.if_mode, omega, s->p.tri_dcd_coarse != 0, i == s->p.iters_coarse - 1)))
            return rc;
    }
    const int bs = block_vert(s);
    k_interp<<<nblk(f.V, bs), bs, 0, s->stream>>>(f.V, f.x, f.xprev, c.x, m.iids, m.iw);
    HIPCHK(hipGetLastError());
    omega = 1.0f;
    for (int i = 0; i < s->p.iters; i++) {
        omega = 4 / (4 - rho * rho * omega);
        if ((rc = pd_iteration(s, f, s->p.if_mode, omega, false, i == s->p.iters - 1))) return rc;
    }
    k_interp<<<nblk(c.V, bs), bs, 0, s->stream>>>(c.V, c.x, c.xprev, f.x, m.aids, m.aw);
    HIPCHK(hipGetLastError());
    k_update_mapping<<<nblk(f.V, bs), bs, 0, s->stream>>>(f.V, f.x, c.x, m.iids, m.iw);
    HIPCHK(hipGetLastError());
    if ((rc = launch_vel(s, c))) return rc;
    return launch_vel(s, f);
}

// --- V-cycle -------------------------------------------------------------------
float cheb_next(float omega, float rho) { return 4 / (4 - rho * rho * omega); }

constexpr int kTailMaxV = 4096;  // levels this small run inside the one-workgroup tail

float4* cur_e(DevLevel& d) { return d.cur ? d.ep : d.e; }
float4* cur_ep(DevLevel& d) { return d.cur ? d.e : d.ep; }

int lin_smooth(mgsb_solver* s, int l, int sweeps, bool zero, const int* skip) {
    DevLevel& d = s->L[l];
    const int bs = block_vert(s);
    const float rho = (float)d.rho;
    float omega = 1.0f;
    for (int k = 0; k < sweeps; k++) {
        omega = cheb_next(omega, rho);
        k_spmv_smooth<<<nblk(d.V, bs), bs, 0, s->stream>>>(d.V, d.ap, d.aj, d.av, cur_e(d), cur_ep(d), d.r, d.vcl,
                                                           omega, s->p.relax_coarse, (zero && k == 0) ? 1 : 0, skip);
        HIPCHK(hipGetLastError());
        d.cur ^= 1;
    }
    return MGSB_OK;
}

// Linear V-cycle on A_l e_l = r_l (l >= 1); the result is in cur_e(L[l]).
int vcycle_lin(mgsb_solver* s, int l, const int* skip) {
    DevLevel& d = s->L[l];
    const int nl = (int)s->L.size();
    const int bs = block_vert(s);
    int rc;
    // the coarsest level alone, many small blocks: one wave per block
    if (l == nl - 1 && s->p.coarse_direct && d.nfree >= 0 && d.maxblk <= kBlockRowsMax && d.nblocks > 1) {
        d.cur = 0;
        k_zero4<<<nblk(d.V, bs), bs, 0, s->stream>>>(d.V, d.e, skip);
        k_coarse_solve_blocks<<<d.nblocks, kBlockRowsMax, 0, s->stream>>>(d.blk, d.rowvert, d.W, d.Wt, d.r, d.e,
                                                                          skip);
        HIPCHK(hipGetLastError());
        return MGSB_OK;
    }
    if (l == s->tail_from) {
        if (d.tailB && s->p.coarse_dense)
            k_tail_gemv<<<nblk(d.V * 64, 256), 256, 0, s->stream>>>(d.V, d.tailB, d.r, d.e, skip);
        else
            k_coarse_tail<<<1, 1024, 0, s->stream>>>(s->tail, nl - l, s->p.coarse_sweeps, s->p.relax_coarse,
                                                     s->p.coarse_direct, skip);
        HIPCHK(hipGetLastError());
        d.cur = 0;
        return MGSB_OK;
    }
    d.cur = 0;
    if (l == nl - 1) {
        if (s->p.coarse_direct && d.nfree >= 0) {
            k_coarse_solve<<<1, 256, 0, s->stream>>>(d.V, d.nfree, d.dense, d.rowinfo, d.W, d.Wt, d.r, d.e, d.big,
                                                     d.maxblk, skip);
            HIPCHK(hipGetLastError());
            return MGSB_OK;
        }
        return lin_smooth(s, l, std::max(1, s->p.coarse_sweeps), true, skip);
    }
    if ((rc = lin_smooth(s, l, s->p.coarse_sweeps, true, skip))) return rc;
    k_spmv_resid<<<nblk(d.V, bs), bs, 0, s->stream>>>(d.V, d.ap, d.aj, d.av, cur_e(d), d.r, d.vcl, d.rr, skip);
    HIPCHK(hipGetLastError());
    DevLevel& c = s->L[l + 1];
    k_restrict<<<nblk(c.V * 64, 256), 256, 0, s->stream>>>(c.V, c.rtp, c.rti, c.rtw, d.rr, c.vcl, c.r, skip);
    HIPCHK(hipGetLastError());
    if ((rc = vcycle_lin(s, l + 1, skip))) return rc;
    LevelMap& m = s->M[l];
    k_prolong_add<<<nblk(d.V, bs), bs, 0, s->stream>>>(d.V, cur_e(d), cur_ep(d), cur_e(c), m.iids, m.iw, d.vcl, skip,
                                                       nullptr);
    HIPCHK(hipGetLastError());
    return lin_smooth(s, l, s->p.coarse_sweeps, false, skip);
}

// Fused fine-level sweeps (k_sweep_fused) run on a tiled level set up for
// them, with one lane per vertex.  Within a half-cycle the positions rotate
// over x, x_prev and xz and the partials over pcf and pcf2; a half-cycle
// starts with u in x (x_prev its history) and the partials at x in pcf, and
// its residual kernel puts x and x_prev back.
constexpr int kFuseAutoMaxV = 65536;
bool fused(const mgsb_solver* s, const DevLevel& d) {
    const bool want = s->p.fuse_sweeps == 2 || (s->p.fuse_sweeps == 1 && d.V <= kFuseAutoMaxV);
    return want && tiled(s, d) && d.tetq && d.xz && vert_lsh(s) == 0;
}
struct FuseBufs {
    float4* cur;   // positions (u) now
    float4* prev;  // the smoother's previous iterate
    int pin;       // the partials at cur: 0 pcf, 1 pcf2
};
float4* fuse_spare(DevLevel& d, const float4* a, const float4* b) {
    for (float4* q : {d.x, d.xprev, d.xz})
        if (q != a && q != b) return q;
    return nullptr;
}
template <int MODE, int PHASE>
void launch_fused_mode(mgsb_solver* s, DevLevel& d, const FuseBufs& fb, float4* out, float omega, int flags,
                       const int* skip, const float4* ec, const float* amask, const int* jrev) {
    LevelMap& m = s->M[0];
    k_sweep_fused<MODE, PHASE><<<d.ntiles, kTileTets, 0, s->stream>>>(
        d.T, fb.cur, fb.prev, out, d.s, d.vc, fb.pin ? d.pcf2 : d.pcf, tile_meta(d), d.tvqp, d.tetq,
        d.B, d.vol, s->p.stiffness, d.tk, fb.pin ? d.pcf : d.pcf2, s->dcol, flags, omega, s->relax_hi, s->relax_lo,
        d.cflag, d.cdepth, d.cenergy, ec, d.tiid, d.tiw, amask, skip, calm_flag(s, 1), jrev, jrev ? d.rr : nullptr);
}
// One fused sweep: JACOBI (vertex update with omega, then the local step) or
// PROLONG (x += mask P ec, then the local step; the smoother restarts).
template <int PHASE>
int launch_fused(mgsb_solver* s, DevLevel& d, FuseBufs& fb, float omega, int flags, const int* skip,
                 const float4* ec = nullptr, const float* amask = nullptr, const int* jrev = nullptr) {
    relax_consts(s);
    float4* out = fuse_spare(d, fb.cur, fb.prev);
    switch (tet_mode_of(s->p.if_mode)) {
        case kTetModePolar: launch_fused_mode<kTetModePolar, PHASE>(s, d, fb, out, omega, flags, skip, ec, amask, jrev); break;
        case kTetModeAc: launch_fused_mode<kTetModeAc, PHASE>(s, d, fb, out, omega, flags, skip, ec, amask, jrev); break;
        default: launch_fused_mode<kTetModeStable, PHASE>(s, d, fb, out, omega, flags, skip, ec, amask, jrev);
    }
    HIPCHK(hipGetLastError());
    fb.prev = PHASE == kFuseProlong ? out : fb.cur;
    fb.cur = out;
    fb.pin ^= 1;
    return MGSB_OK;
}

// Fine-level residual on split positions (x holds u = x - s).  first: the
// step's first evaluation, at s itself.  chk: the check applied to the norm
// (kChkNone: no norm).  fb != nullptr: after fused sweeps, the residual at
// fb->cur from the partials fb->pin; x and x_prev return to their buffers
// (x_prev: x with restart, else the smoother's previous iterate).
int fine_residual(mgsb_solver* s, bool need_tet, bool write_r, int chk, bool restart, const int* skip,
                  bool first = false, const FuseBufs* fb = nullptr) {
    DevLevel& d = s->L[0];
    int rc;
    if (need_tet && (rc = launch_tet(s, d, tet_mode_of(s->p.if_mode), false, first ? d.s : d.x, skip, !first)))
        return rc;
    const int lsh = vert_lsh(s), nb = nblk(d.V << lsh, 256);
    const float4* xin = fb ? fb->cur : d.x;
    const float3* cfin = fb ? (fb->pin ? d.pcf2 : d.pcf) : gath_cf(s, d);
    float4* xcanon = (fb && fb->cur != d.x) ? d.x : nullptr;
    float4* xpout = restart ? d.xprev : nullptr;
    const float4* xpsrc = nullptr;
    if (fb) {
        if (restart) {
            xpout = fb->cur != d.xprev ? d.xprev : nullptr;
        } else if (fb->prev != d.xprev) {
            xpout = d.xprev;
            xpsrc = fb->prev;
        }
    }
    if (s->body_checks && chk != kChkNone) {  // per-body norms, then the check
        k_resid<<<nb, 256, 0, s->stream>>>(d.V, xin, d.s, d.vc, gath_rp(s, d), cfin, s->dcol, kVertPoint,
                                           write_r ? d.r : nullptr, nullptr, xpout, skip, 1, s->g2, d.cmask, lsh,
                                           xcanon, xpsrc);
        HIPCHK(hipGetLastError());
        k_body_check<<<s->nbodies, kBodyCheckThreads, 0, s->stream>>>(s->vb, s->g2, s->bodysq, s->res0b, s->bcount,
                                                                      s->vstate, chk, s->hmirror, skip, s->seqc);
        s->chk_launched++;
        HIPCHK(hipGetLastError());
        return MGSB_OK;
    }
    k_resid<<<nb, 256, 0, s->stream>>>(d.V, xin, d.s, d.vc, gath_rp(s, d), cfin, s->dcol, kVertPoint,
                                       write_r ? d.r : nullptr, chk != kChkNone ? s->vpart : nullptr, xpout, skip, 1,
                                       nullptr, d.cmask, lsh, xcanon, xpsrc);
    HIPCHK(hipGetLastError());
    if (chk != kChkNone) {
        k_check<<<1, 256, 0, s->stream>>>(nb, s->vpart, s->vstate, chk, s->hmirror, skip, s->seqc);
        s->chk_launched++;
        HIPCHK(hipGetLastError());
    }
    return MGSB_OK;
}

// A residual-correction V-cycle step (include/mgsb.h, MGSB_SCHED_VCYCLE) is
// enqueued as four pieces.  Every kernel after the start returns at once when
// the step's convergence flag is set.
constexpr int kGateAfter = 8, kGateFirst = 8, kGateMax = 256, kGateProbe = 64;

bool temporal_warm(const mgsb_solver* s) {
    return s->p.warm_start == 1 || (s->p.schedule == MGSB_SCHED_LEARNED && s->p.warm_start != 2);
}


// Inertial prediction, g(s) and res0, then the warm start: the temporal
// extrapolation of the last corrections (warm_start = 1 or learned), and with
// learned = true the MLP's coarse correction on top.  The network reads only
// what is known before the step (velocity, rest position, the last two
// steps' coarse deltas), so it runs before the inertial prediction and its
// prolongated correction enters the warm start with the extrapolation:
// the stage costs the feature and inference launches, and the one judge of
// the warm start (vstate_apply: kept if its residual is at most learn_slack
// times g(s)'s, else undone by k_revert) covers both.
// k_start_fused on level 0 (fused(s, d) holds): the inertial prediction with
// sphere advance and the warm start (temporal extrapolation, plus epred when
// the learned stage ran), then both local steps.
int launch_start_fused(mgsb_solver* s, DevLevel& d, bool temporal, const float4* epred) {
    auto k = tet_mode_of(s->p.if_mode) == kTetModePolar
                 ? k_start_fused<kTetModePolar>
                 : (tet_mode_of(s->p.if_mode) == kTetModeAc ? k_start_fused<kTetModeAc> : k_start_fused<kTetModeStable>);
    k<<<2 * d.ntiles, kTileTets, 0, s->stream>>>(
        d.T, d.ntiles, d.x, d.v, d.xlast, d.s, d.xz, d.vc, s->p.dt, s->p.damping, s->p.gravity[0], s->p.gravity[1],
        s->p.gravity[2], s->dcol, temporal ? d.U1 : nullptr, temporal ? d.U2 : nullptr, epred,
        epred ? d.tiid : nullptr, epred ? d.tiw : nullptr, tile_meta(d), d.tvqp, d.tetq, d.B, d.vol,
        s->p.stiffness, d.tk, d.pcf, d.pcf2);
    HIPCHK(hipGetLastError());
    return MGSB_OK;
}

int enqueue_vc_start(mgsb_solver* s, bool learned) {
    DevLevel& f = s->L[0];
    DevLevel& c1 = s->L[1];
    const int* skip = &s->vstate->converged;
    const int nbv = nblk(f.V, 256);
    const int bs = block_vert(s);
    const bool features = s->recording;  // (the learned stage computes its own)
    const bool temporal = temporal_warm(s);
    const bool warm = temporal || learned;
    // g(s) and g(warm start) from one local-step launch, one residual launch
    // and one check (per-body checks keep the sequential form, which reduces
    // each residual per body)
    const bool dual = warm && !s->body_checks;
    const int judge = learned ? kChkWarm : kChkTemporal;
    const float4* ep = learned ? s->epred : nullptr;
    float* recX = s->recording ? s->recX + (size_t)s->rec_steps * c1.V * kFeat : nullptr;
    // a coarse-first step restricts the start's residual: keep it in r
    const bool cf = coarse_first(s);
    int rc;
    if (features) {  // features at the step's start (recorded even if the step converges at once)
        k_features<<<nblk(c1.V, bs), bs, 0, s->stream>>>(c1.V, f.v, s->M[0].aids, s->M[0].aw, s->X1, s->D1, s->D2,
                                                         s->norm, s->feat, recX);
        HIPCHK(hipGetLastError());
    }
    if (learned) {  // features, network, prior and mask in one launch
        const CoarseIO io{f.v, s->M[0].aids, s->M[0].aw, s->X1, s->D1, s->D2, s->norm, c1.vcl, s->epred};
        if ((rc = mlp_infer_coarse(s->mlp, c1.V, io, s->stream))) return rc;
    }
    if (dual) {
        // on a level that runs fused sweeps, the inertial prediction and both
        // local steps in one launch; the warm start's correction then sits in
        // xz until the dual residual copies it into x and x_prev
        const bool fst = fused(s, f);
        if (fst) {
            if ((rc = launch_start_fused(s, f, temporal, ep))) return rc;
        } else {
            if ((rc = launch_st(s, f, 1, 1, temporal, ep))) return rc;
            if ((rc = launch_tet_dual(s, f))) return rc;
        }
        const int lsh = vert_lsh(s), nbr = nblk(f.V << lsh, 256);
        k_resid_dual<<<nbr, 256, 0, s->stream>>>(f.V, fst ? f.xz : f.x, f.s, f.vc, gath_rp(s, f), gath_cf(s, f),
                                                 gath_cf2(s, f), s->dcol, kVertPoint, cf ? f.r : nullptr, s->vpart,
                                                 f.cmask, lsh, cf ? f.rr : nullptr, fst ? f.x : nullptr,
                                                 fst ? f.xprev : nullptr);
        HIPCHK(hipGetLastError());
        if (fst && cf && s->next_is_post) {  // the judge runs in the next half's restriction
            s->judge_folded = judge;
            s->judge_nb = nbr;
            return MGSB_OK;
        }
        // res0, the judge and its revert to s in one launch
        k_judge_revert<<<nbv, 256, 0, s->stream>>>(nbr, s->vpart, s->vstate, s->hmirror, s->seqc, judge, f.V, f.x,
                                                   f.xprev, gath_rp(s, f), gath_cf2(s, f), gath_cf(s, f), f.r,
                                                   cf ? f.rr : nullptr, f.cmask);
        s->chk_launched++;
        HIPCHK(hipGetLastError());
        return MGSB_OK;
    }
    if ((rc = launch_st(s, f, 1, 1, false))) return rc;
    if ((rc = fine_residual(s, true, cf, kChkStart, true, nullptr, true))) return rc;
    if (warm) {
        k_warm_temporal<<<nbv, 256, 0, s->stream>>>(f.V, f.x, f.xprev, temporal ? f.U1 : nullptr,
                                                    temporal ? f.U2 : nullptr, f.vc, skip, ep,
                                                    ep ? s->M[0].iids : nullptr, ep ? s->M[0].iw : nullptr);
        HIPCHK(hipGetLastError());
        if ((rc = fine_residual(s, true, cf, judge, true, skip))) return rc;
        k_revert<<<nbv, 256, 0, s->stream>>>(f.V, f.x, f.xprev, f.U1, f.U2, f.vc, s->vstate, gath_rp(s, f),
                                             nullptr, gath_cf(s, f), nullptr, nullptr, nullptr);
        HIPCHK(hipGetLastError());
        // after a revert the corner forces belong to the rejected x: recompute
        if ((rc = launch_tet(s, f, tet_mode_of(s->p.if_mode), false, nullptr, &s->vstate->cf_ok, true))) return rc;
        if (cf) {  // ... and so does the residual a coarse-first half restricts
            const int lsh = vert_lsh(s);
            k_resid<<<nblk(f.V << lsh, 256), 256, 0, s->stream>>>(
                f.V, f.x, f.s, f.vc, gath_rp(s, f), gath_cf(s, f), s->dcol, kVertPoint, f.r, nullptr, nullptr,
                &s->vstate->cf_ok, 1, nullptr, f.cmask, lsh);
            HIPCHK(hipGetLastError());
        }
    }
    return MGSB_OK;
}

// First half of a cycle: Chebyshev-accelerated nonlinear Jacobi sweeps, then
// the residual that the coarse correction restricts (checked when
// check_every == 2).
int enqueue_vc_pre(mgsb_solver* s) {
    DevLevel& f = s->L[0];
    const int* skip = &s->vstate->converged;
    const float rs = s->p.rho_smooth;
    int rc;
    float omega = 1.0f;
    if (fused(s, f)) {  // each sweep's vertex update inside the next local step's launch
        FuseBufs fb{f.x, f.xprev, 0};
        for (int k = 0; k < s->p.pre_sweeps; k++) {
            omega = cheb_next(omega, rs);
            if ((rc = launch_fused<kFuseJacobi>(s, f, fb, omega, kVertPoint | kVertContactJacobi, skip))) return rc;
        }
        return fine_residual(s, false, true, s->p.check_every == 2 ? kChkMid : kChkNone, false, skip, false, &fb);
    }
    for (int k = 0; k < s->p.pre_sweeps; k++) {
        omega = cheb_next(omega, rs);
        if (k > 0 && (rc = launch_tet(s, f, tet_mode_of(s->p.if_mode), false, nullptr, skip, true))) return rc;
        if ((rc = launch_vert(s, f, omega, kVertPoint | kVertContactJacobi, skip, 1))) return rc;
    }
    return fine_residual(s, s->p.pre_sweeps > 0, true, s->p.check_every == 2 ? kChkMid : kChkNone, false, skip);
}

// Second half: restriction, the linear V-cycle on the coarse levels,
// prolongation, post-smoothing and the end-of-cycle residual check.
int enqueue_vc_post(mgsb_solver* s) {
    DevLevel& f = s->L[0];
    DevLevel& c1 = s->L[1];
    const int* skip = &s->vstate->converged;
    const int bs = block_vert(s);
    const float rs = s->p.rho_smooth;
    int rc;
    const int judge = s->judge_folded;
    s->judge_folded = 0;
    if (judge) {  // the fused warm start's judge (k_restrict_judged)
        k_restrict_judged<<<nblk(c1.V * 64, 256), 256, 0, s->stream>>>(
            c1.V, c1.rtp, c1.rti, c1.rtw, f.r, f.rr, c1.vcl, c1.r, s->judge_nb, s->vpart, s->vstate, s->hmirror,
            s->seqc, judge);
        s->chk_launched++;
    } else {
        k_restrict<<<nblk(c1.V * 64, 256), 256, 0, s->stream>>>(c1.V, c1.rtp, c1.rti, c1.rtw, f.r, c1.vcl, c1.r, skip);
    }
    HIPCHK(hipGetLastError());
    if ((rc = vcycle_lin(s, 1, skip))) return rc;
    float omega = 1.0f;
    if (fused(s, f)) {  // the prolongation, and each sweep's vertex update, inside a local step's launch
        // the partials alternate so that the last sweep's land in pcf
        FuseBufs fb{f.x, f.xprev, (1 + s->p.post_sweeps) & 1};
        if ((rc = launch_fused<kFuseProlong>(s, f, fb, 1.0f, 0, skip, cur_e(c1), f.cmask,
                                             judge ? &s->vstate->revert : nullptr)))
            return rc;
        for (int k = 0; k < s->p.post_sweeps; k++) {
            omega = cheb_next(omega, rs);
            const int fl = kVertPoint | kVertContactJacobi | (k == s->p.post_sweeps - 1 ? kVertWriteCol : 0);
            if ((rc = launch_fused<kFuseJacobi>(s, f, fb, omega, fl, skip))) return rc;
        }
        return fine_residual(s, false, false, kChkCycle, true, skip, false, &fb);
    }
    k_prolong_add<<<nblk(f.V, bs), bs, 0, s->stream>>>(f.V, f.x, f.xprev, cur_e(c1), s->M[0].iids, s->M[0].iw, f.vc,
                                                       skip, f.cmask);
    HIPCHK(hipGetLastError());
    for (int k = 0; k < s->p.post_sweeps; k++) {
        omega = cheb_next(omega, rs);
        if ((rc = launch_tet(s, f, tet_mode_of(s->p.if_mode), false, nullptr, skip, true))) return rc;
        if ((rc = launch_vert(s, f, omega,
                              kVertPoint | kVertContactJacobi | (k == s->p.post_sweeps - 1 ? kVertWriteCol : 0), skip,
                              1)))
            return rc;
    }
    return fine_residual(s, true, false, kChkCycle, true, skip);
}

// Coarse correction achieved (features of the next step / training target;
// x = s + u, velocities, the delta history of the learned features, run
// aggregates; then the training target when the step is recorded (left out
// when lite).
int enqueue_vc_finish(mgsb_solver* s, bool lite) {
    DevLevel& f = s->L[0];
    DevLevel& c1 = s->L[1];
    const int bs = block_vert(s);
    k_finish_split<<<nblk(f.V, bs), bs, 0, s->stream>>>(f.V, f.x, f.xprev, f.s, f.xlast, f.v, s->p.dt, s->vstate,
                                                        f.U1, f.U2, f.vc, s->D1, s->D2, temporal_warm(s) ? 1 : 0,
                                                        &s->vstate->notdone);
    HIPCHK(hipGetLastError());
    if (!lite && s->recording) {
        k_target<<<nblk(c1.V, bs), bs, 0, s->stream>>>(c1.V, s->D1, s->D2, s->M[0].aids, s->M[0].aw,
                                                       s->recT + (size_t)s->rec_steps * c1.V * 3,
                                                       &s->vstate->notdone);
        HIPCHK(hipGetLastError());
    }
    return MGSB_OK;
}

int enqueue_piece(mgsb_solver* s, int piece) {
    const bool learned = s->p.schedule == MGSB_SCHED_LEARNED;
    switch (piece) {
        case kPieceStart: return enqueue_vc_start(s, learned);
        case kPieceStartGated: return enqueue_vc_start(s, false);
        case kPieceFinishLite: return enqueue_vc_finish(s, true);
        case kPiecePre: return enqueue_vc_pre(s);
        case kPiecePost: return enqueue_vc_post(s);
        default: return enqueue_vc_finish(s, false);
    }
}

// Enqueues the pieces seq in order, telling each whether a coarse-first
// half follows it in the same round.
int enqueue_seq(mgsb_solver* s, const std::string& seq) {
    int rc = MGSB_OK;
    for (size_t i = 0; i < seq.size() && rc == MGSB_OK; i++) {
        s->next_is_post = i + 1 < seq.size() && seq[i + 1] == (char)kPiecePost;
        rc = enqueue_piece(s, seq[i]);
    }
    s->next_is_post = false;
    s->judge_folded = 0;
    return rc;
}

// Captures the pieces seq (kPiece* codes, one char each) as one graph and
// keeps it (no launch).
int capture_pieces(mgsb_solver* s, const std::string& seq, std::map<std::string, mgsb_solver::PieceGraph>::iterator& it) {
    it = s->pseq.find(seq);
    if (it == s->pseq.end()) {
        HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        const unsigned before = s->chk_launched;
        const int rc = enqueue_seq(s, seq);
        mgsb_solver::PieceGraph pg;
        pg.checks = s->chk_launched - before;  // captured, not run yet
        s->chk_launched = before;
        const hipError_t ce = hipStreamEndCapture(s->stream, &pg.graph);
        if (rc || ce != hipSuccess) {
            if (pg.graph) (void)hipGraphDestroy(pg.graph);
            if (rc) return rc;
            HIPCHK(ce);
        }
        const hipError_t ie = hipGraphInstantiate(&pg.exec, pg.graph, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            (void)hipGraphDestroy(pg.graph);
            HIPCHK(ie);
        }
        it = s->pseq.emplace(seq, pg).first;
    }
    return MGSB_OK;
}

// Enqueues the pieces seq in order: eagerly, or as one graph (captured at
// the sequence's first use unless pre-captured).
int run_pieces(mgsb_solver* s, const std::string& seq, bool graphable) {
    if (!graphable) return enqueue_seq(s, seq);
    std::map<std::string, mgsb_solver::PieceGraph>::iterator it;
    if (int rc = capture_pieces(s, seq, it)) return rc;
    HIPCHK(hipGraphLaunch(it->second.exec, s->stream));
    s->chk_launched += it->second.checks;
    return MGSB_OK;
}

// The piece sequences step_vcycle can launch form a finite set: the first
// round of a step (start or gated start, the halves guessed from the last
// step, finish or lite finish) and the rounds after a verdict (one half, or
// a cycle, plus a finish).  They are captured before the first step runs, so
// that no step pays for a capture and instantiation inside a timed loop
// (one capture costs far more than a step).  First rounds are captured up to
// kPrecaptureHalves halves; longer guesses are captured at first use.
constexpr int kPrecaptureHalves = 8;

// The piece of a step's half-cycle h (0-based): pre-smoothing and coarse-
// correction halves alternate, the first a coarse one with coarse_first.
char half_piece(const mgsb_solver* s, int h) {
    return (char)(((h & 1) != 0) != coarse_first(s) ? kPiecePost : kPiecePre);
}
// With checks after whole cycles only, a sequence of h halves must not end
// with a pre-smoothing half (it carries no check).
bool ends_unchecked(const mgsb_solver* s, int h) {
    return s->p.check_every == 1 && h > 0 && half_piece(s, h - 1) == kPiecePre;
}

int precapture_pieces(mgsb_solver* s) {
    const bool learned = s->p.schedule == MGSB_SCHED_LEARNED;
    const int check = s->p.check_every;
    const int maxh = max_halves(s);
    const std::string starts = learned ? std::string{(char)kPieceStart, (char)kPieceStartGated}
                                       : std::string{(char)kPieceStart};
    // only a recording (which runs eagerly) takes the full finish (fin_for)
    const std::string fins(1, (char)kPieceFinishLite);
    auto halves = [s](char st, int h) {
        std::string q(1, st);
        for (int k = 0; k < h; k++) q.push_back(half_piece(s, k));
        return q;
    };
    std::vector<std::string> seqs;
    for (const char st : starts) {
        if (!check) {  // no checks: start, max_cycles cycles, finish
            for (const char f : fins) seqs.push_back(halves(st, maxh) + f);
            continue;
        }
        for (int h = 0; h <= std::min(maxh, kPrecaptureHalves); h++) {
            if (ends_unchecked(s, h)) continue;
            for (const char f : fins) seqs.push_back(halves(st, h) + f);
        }
    }
    if (check)
        for (const char f : fins) {
            if (check == 2) {
                seqs.push_back(std::string{(char)kPiecePre, f});
                seqs.push_back(std::string{(char)kPiecePost, f});
            } else {
                seqs.push_back(std::string{(char)kPiecePre, (char)kPiecePost, f});
            }
        }
    HIPCHK(hipStreamSynchronize(s->stream));
    for (const std::string& q : seqs) {
        std::map<std::string, mgsb_solver::PieceGraph>::iterator it;
        if (int rc = capture_pieces(s, q, it)) return rc;
    }
    s->precaptured = true;
    return MGSB_OK;
}

// Waits until every check kernel enqueued so far has run (the host mirror's
// count reaches chk_launched), spinning on the coherent mirror; the stream
// itself may still be busy.  Falls back to the stream's completion.
unsigned long long mirror_word(const mgsb_solver* s) {
    return __atomic_load_n(reinterpret_cast<const unsigned long long*>(s->hmirror), __ATOMIC_ACQUIRE);
}
// The verdict published with the last check that ran (bump_seq).
struct Verdict {
    bool converged;
    int fallback, accepted, halves, cycles;
};
Verdict mirror_verdict(const mgsb_solver* s) {
    const unsigned v = (unsigned)(mirror_word(s) >> 32);
    return Verdict{(v & 1u) != 0, (int)((v >> 1) & 3u), (int)((v >> 3) & 3u), (int)((v >> 8) & 255u),
                   (int)((v >> 16) & 255u)};
}

int wait_checks(mgsb_solver* s) {
    const unsigned want = s->chk_launched;
    for (long it = 1;; it++) {
        if ((unsigned)mirror_word(s) == want) return MGSB_OK;
        if ((it & 4095) == 0) {
            const hipError_t q = hipStreamQuery(s->stream);
            if (q == hipSuccess) {  // all work done: the mirror is final
                if ((unsigned)mirror_word(s) == want) return MGSB_OK;
                set_error("step: the stream drained with a published check count other than the one enqueued");
                return MGSB_ERR_STATE;
            }
            if (q != hipErrorNotReady) HIPCHK(q);
        }
    }
}

// One V-cycle step.  Without checks all max_cycles cycles are enqueued.  With
// checks the host enqueues the number of half-cycles the previous step
// needed and the finish piece behind them, then waits for the checks' verdict
// on the host mirror and adds halves (and another finish) one check point at
// a time until the step has converged or max_cycles cycles have run.  The
// finish kernels run only once the device has marked the step done, so the
// finish overlaps the host's wait, and no launches are spent on a converged
// step beyond a wrong guess.
int step_vcycle(mgsb_solver* s, bool graphable) {
    int rc;
    if (graphable && !s->precaptured && (rc = precapture_pieces(s))) return rc;
    const bool learned = s->p.schedule == MGSB_SCHED_LEARNED;
    const bool gate_on = learned && s->p.learn_gate && !s->recording;
    // no MLP stage this step: the gate is shut after rejections, or the
    // stage is idle (below)
    const bool gated = gate_on && (s->gate_left > 0 || !s->stage_on);
    if (gate_on && s->gate_left > 0) s->gate_left--;
    std::string seq(1, (char)(gated ? kPieceStartGated : kPieceStart));
    const int check = s->p.check_every;
    const int maxh = max_halves(s);
    // every finish keeps the delta history of the learned features; only a
    // recorded step also writes its training target
    auto fin_for = [&](int) { return s->recording ? kPieceFinish : kPieceFinishLite; };
    int h = 0;
    int target = check ? std::min(maxh, std::max(0, s->halves_guess)) : maxh;
    bool finished = false;
    for (;;) {
        for (; h < target; h++) seq.push_back(half_piece(s, h));
        if (!check) break;
        if (ends_unchecked(s, h)) {  // a pre half carries no check: finish the cycle
            target = h + 1;
            continue;
        }
        seq.push_back((char)fin_for(h));  // the finish runs only if the step is done
        if ((rc = run_pieces(s, seq, graphable))) return rc;
        seq.clear();
        if ((rc = wait_checks(s))) return rc;
        const Verdict vd = mirror_verdict(s);
        if (vd.converged || h >= maxh) {
            s->halves_guess = vd.converged ? vd.halves : maxh;
            finished = true;
            break;
        }
        target = h + (check == 2 ? 1 : 2);
    }
    if (gate_on && check && finished && gated && s->gate_left == 0) {  // a step without the stage
        // the warm start without it needed more than the first half-cycle:
        // what is left may be the smooth error a coarse correction removes
        s->off_halves = mirror_verdict(s).halves;
        s->stage_on = s->off_halves >= 2;
        s->since_probe = 0;
    }
    if (gate_on && check && finished && !gated) {  // the mirror holds this step's verdict
        const Verdict vd = mirror_verdict(s);
        // the stage keeps running while its steps end after the first
        // half-cycle or need fewer than the last step without it; otherwise
        // (and every kGateProbe steps) the next step runs without it
        s->stage_on = (vd.halves <= 1 || (s->off_halves >= 0 && vd.halves < s->off_halves)) &&
                      ++s->since_probe < kGateProbe;
        if (vd.fallback & 2) {  // the judge undid the stage's correction
            if (++s->gate_rejects >= kGateAfter) {
                s->gate_skip = s->gate_skip ? std::min(2 * s->gate_skip, kGateMax) : kGateFirst;
                s->gate_left = s->gate_skip;
                s->gate_rejects = 0;
            }
        } else if (vd.accepted == 2) {
            s->gate_rejects = 0;
            s->gate_skip = 0;
        }
    }
    if (!finished) seq.push_back((char)fin_for(h));
    if (!seq.empty() && (rc = run_pieces(s, seq, graphable))) return rc;
    if (s->recording) s->rec_steps++;  // the finish enqueued above recorded this step's target
    return MGSB_OK;
}

int enqueue_step(mgsb_solver* s) {
    switch (s->p.schedule) {
        case MGSB_SCHED_SINGLE: return enqueue_step_single(s);
        case MGSB_SCHED_TWOLEVEL: return enqueue_step_twolevel(s);
        default: set_error("schedule not available"); return MGSB_ERR_ARG;
    }
}

// Coarse Dirichlet masks, vcl, rho and the coarsest factor from the current
// level-0 fixflags (V-cycle set-up; rerun when fixflags or dt change).
int invalidate_graph(mgsb_solver* s);

int setup_vcycle(mgsb_solver* s) {
    const int nl = (int)s->L.size();
    if (nl < 2) return MGSB_OK;
    // captured pieces hold the device pointers rebuilt below
    if (int rc = invalidate_graph(s)) return rc;
    s->L[0].vfix = s->L[0].h.fix;
    const float dt = s->p.dt;
    // one-workgroup tail: the longest suffix of levels (>= 1) with V <= kTailMaxV
    s->tail_from = -1;
    for (int l = nl - 1; l >= 1 && s->L[l].V <= kTailMaxV; l--) s->tail_from = l;
    for (int l = 1; l < nl; l++) {
        DevLevel& c = s->L[l];
        derive_vfix(s->L[l - 1].vfix, s->M[l - 1].h_aids, s->M[l - 1].h_aw, c.V, c.vfix);
        std::vector<int32_t> ap, aj;
        std::vector<float> av;
        build_level_matrix(c.h, s->p.stiffness, dt, ap, aj, av);
        c.adiag.assign(c.V, 0.f);
        for (int i = 0; i < c.V; i++)
            for (int q = ap[i]; q < ap[i + 1]; q++)
                if (aj[q] == i) c.adiag[i] = av[q];
        if (!c.ap) {
            int rc = up(s, c.owned, &c.ap, ap.data(), ap.size());
            if (rc) return rc;
            if ((rc = up(s, c.owned, &c.aj, aj.data(), aj.size()))) return rc;
            if ((rc = up(s, c.owned, &c.av, av.data(), av.size()))) return rc;
        } else {
            HIPCHK(hipMemcpyAsync(c.av, av.data(), sizeof(float) * av.size(), hipMemcpyHostToDevice, s->stream));
        }
        if (l == s->tail_from && l < nl - 1 && c.V <= kTailLdsV) {  // the tail's first level: ELL copy
            int w = 0;
            for (int i = 0; i < c.V; i++) w = std::max(w, ap[i + 1] - ap[i]);
            std::vector<int32_t> je((size_t)w * c.V);
            std::vector<float> ve((size_t)w * c.V);
            for (int i = 0; i < c.V; i++)
                for (int k = 0; k < w; k++) {
                    const int q = ap[i] + k;
                    const bool on = q < ap[i + 1];
                    je[(size_t)k * c.V + i] = on ? aj[q] : i;
                    ve[(size_t)k * c.V + i] = on ? av[q] : 0.f;
                }
            int rc = MGSB_OK;
            if (!c.ajE) {
                if ((rc = up(s, c.owned, &c.ajE, je.data(), je.size()))) return rc;
                if ((rc = up(s, c.owned, &c.avE, ve.data(), ve.size()))) return rc;
            } else {
                HIPCHK(hipMemcpyAsync(c.avE, ve.data(), sizeof(float) * ve.size(), hipMemcpyHostToDevice, s->stream));
            }
            c.ellw = w;
            HIPCHK(hipStreamSynchronize(s->stream));
        }
        std::vector<float4> vcl(c.V);
        for (int i = 0; i < c.V; i++) vcl[i] = make_float4(c.adiag[i], 0.f, c.vfix[i], c.h.mass[i]);
        HIPCHK(hipMemcpyAsync(c.vcl, vcl.data(), sizeof(float4) * c.V, hipMemcpyHostToDevice, s->stream));
        c.rho = s->p.rho_coarse > 0 ? s->p.rho_coarse
                                     : power_rho(c.h, c.vfix, dt, s->p.relax_coarse, s->p.stiffness);
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    DevLevel& cc = s->L[nl - 1];
    std::vector<int32_t> dense, rowinfo;
    std::vector<double> W;
    int n = 0;
    cc.nfree = -1;
    for (double** q : {&cc.W, &cc.Wt, &cc.big}) {
        if (*q) HIPCHK(hipFree(*q));
        *q = nullptr;
    }
    if (cc.rowinfo) HIPCHK(hipFree(cc.rowinfo));
    cc.rowinfo = nullptr;
    if (cc.blk) HIPCHK(hipFree(cc.blk));
    if (cc.rowvert) HIPCHK(hipFree(cc.rowvert));
    cc.blk = nullptr;
    cc.rowvert = nullptr;
    cc.nblocks = 0;
    if (coarse_cholesky(cc.h, cc.vfix, dt, s->p.stiffness, dense, W, rowinfo, n) == 0 && n > 0 &&
        W.size() <= kMaxDenseEntries) {
        int rc = dalloc(&cc.W, W.size());
        if (rc) return rc;
        if ((rc = dalloc(&cc.Wt, W.size()))) return rc;
        if ((rc = dalloc(&cc.rowinfo, (size_t)n))) return rc;
        if (n > kMaxDense && (rc = dalloc(&cc.big, (size_t)6 * n))) return rc;
        std::vector<double> Wt(W.size());
        for (int r = 0; r < n; r++) {  // transpose block by block
            const int nb = rowinfo[(size_t)r * 4 + 1], off = rowinfo[(size_t)r * 4 + 2];
            const int i = r - rowinfo[(size_t)r * 4 + 0];
            for (int j = 0; j < nb; j++) Wt[(size_t)off + (size_t)i * nb + j] = W[(size_t)off + (size_t)j * nb + i];
        }
        HIPCHK(hipMemcpyAsync(cc.W, W.data(), sizeof(double) * W.size(), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(cc.Wt, Wt.data(), sizeof(double) * Wt.size(), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(cc.rowinfo, rowinfo.data(), sizeof(int32_t) * rowinfo.size(), hipMemcpyHostToDevice,
                              s->stream));
        HIPCHK(hipMemcpyAsync(cc.dense, dense.data(), sizeof(int) * dense.size(), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        cc.nfree = n;
        cc.maxblk = 0;
        for (int r = 0; r < n; r++) cc.maxblk = std::max(cc.maxblk, (int)rowinfo[(size_t)r * 4 + 1]);
        // per-block table and row -> vertex map of the one-wave-per-block solve
        std::vector<int32_t> blk, rowvert(n, 0);
        for (int r = 0; r < n; r++)
            if (rowinfo[(size_t)r * 4 + 0] == r)
                blk.insert(blk.end(), {r, rowinfo[(size_t)r * 4 + 1], rowinfo[(size_t)r * 4 + 2], 0});
        for (int i = 0; i < cc.V; i++)
            if (dense[i] >= 0) rowvert[dense[i]] = i;
        cc.nblocks = (int)(blk.size() / 4);
        if ((rc = dalloc(&cc.blk, (size_t)cc.nblocks))) return rc;
        if ((rc = dalloc(&cc.rowvert, (size_t)n))) return rc;
        HIPCHK(hipMemcpy(cc.blk, blk.data(), sizeof(int32_t) * blk.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(cc.rowvert, rowvert.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    }
    if (s->tail_from >= 1) {
        std::vector<TailLevel> tl;
        for (int l = s->tail_from; l < nl; l++) {
            DevLevel& d = s->L[l];
            TailLevel t{};
            t.V = d.V;
            t.nfree = d.nfree;
            t.maxblk = d.maxblk;
            t.ap = d.ap;
            t.aj = d.aj;
            t.av = d.av;
            t.vcl = d.vcl;
            t.e0 = d.e;
            t.e1 = d.ep;
            t.r = d.r;
            t.rr = d.rr;
            t.rtp = d.rtp;
            t.rti = d.rti;
            t.rtw = d.rtw;
            t.iids = l + 1 < nl ? s->M[l].iids : nullptr;
            t.iw = l + 1 < nl ? s->M[l].iw : nullptr;
            t.dense = d.dense;
            t.W = d.W;
            t.Wt = d.Wt;
            t.rowinfo = d.rowinfo;
            t.big = d.big;
            t.rho = (float)d.rho;
            t.ellw = l == s->tail_from ? d.ellw : 0;
            t.ajE = d.ajE;
            t.avE = d.avE;
            tl.push_back(t);
        }
        if (!s->tail) {
            int rc = dalloc(&s->tail, 8);
            if (rc) return rc;
            s->owned.push_back(s->tail);
        }
        HIPCHK(hipMemcpyAsync(s->tail, tl.data(), sizeof(TailLevel) * tl.size(), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        // the tail as a dense matrix, three columns per run on unit vectors
        DevLevel& t0 = s->L[s->tail_from];
        if (s->p.coarse_dense && s->tail_from < nl - 1 && t0.V <= kTailDenseMaxV) {
            if (!t0.tailB) {
                int rc = dalloc(&t0.tailB, (size_t)t0.V * t0.V);
                if (rc) return rc;
                t0.owned.push_back(t0.tailB);
            }
            const int nb = nblk(t0.V, 256);
            for (int c0 = 0; c0 < t0.V; c0 += 3) {
                k_unit_rhs<<<nb, 256, 0, s->stream>>>(t0.V, t0.r, c0);
                k_coarse_tail<<<1, 1024, 0, s->stream>>>(s->tail, nl - s->tail_from, s->p.coarse_sweeps,
                                                         s->p.relax_coarse, s->p.coarse_direct, nullptr);
                k_store_cols<<<nb, 256, 0, s->stream>>>(t0.V, t0.e, t0.tailB, c0);
            }
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(s->stream));
        }
    }
    return MGSB_OK;
}

int alloc_vcycle(mgsb_solver* s) {
    const int nl = (int)s->L.size();
    int rc;
    for (int l = 0; l < nl; l++) {
        DevLevel& d = s->L[l];
        if ((rc = up<float4>(s, d.owned, &d.r, nullptr, d.V))) return rc;
        if (l == 0) {
            std::vector<float4> z(d.V, make_float4(0.f, 0.f, 0.f, 0.f));
            if ((rc = up(s, d.owned, &d.U1, z.data(), d.V))) return rc;
            if ((rc = up(s, d.owned, &d.U2, z.data(), d.V))) return rc;
            std::vector<float> one(d.V, 1.f);
            if ((rc = up(s, d.owned, &d.cmask, one.data(), d.V))) return rc;
            // the residual at s of a warm-started start, for a coarse-first revert
            if ((rc = up<float4>(s, d.owned, &d.rr, nullptr, d.V))) return rc;
            if (d.tetq && !d.xz && (rc = up<float4>(s, d.owned, &d.xz, nullptr, d.V))) return rc;
            if ((rc = up<float3>(s, d.owned, &d.cf2, nullptr, (size_t)d.T * 4))) return rc;
            if ((rc = up<float3>(s, d.owned, &d.pcf2, nullptr, (size_t)d.npart))) return rc;
            continue;
        }
        if ((rc = up<float4>(s, d.owned, &d.e, nullptr, d.V))) return rc;
        if ((rc = up<float4>(s, d.owned, &d.ep, nullptr, d.V))) return rc;
        if ((rc = up<float4>(s, d.owned, &d.rr, nullptr, d.V))) return rc;
        if ((rc = up<float4>(s, d.owned, &d.vcl, nullptr, d.V))) return rc;
        if ((rc = up<int>(s, d.owned, &d.dense, nullptr, d.V))) return rc;
        std::vector<int32_t> rtp, rti;
        std::vector<float> rtw;
        build_restriction(s->L[l - 1].V, d.V, s->M[l - 1].h_iids, s->M[l - 1].h_iw, rtp, rti, rtw);
        if ((rc = up(s, d.owned, &d.rtp, rtp.data(), rtp.size()))) return rc;
        if ((rc = up(s, d.owned, &d.rti, rti.data(), std::max<size_t>(1, rti.size())))) return rc;
        if ((rc = up(s, d.owned, &d.rtw, rtw.data(), std::max<size_t>(1, rtw.size())))) return rc;
    }
    {
        DevLevel& c1 = s->L[1];
        if ((rc = up<float>(s, s->owned, &s->feat, nullptr, (size_t)c1.V * kFeat))) return rc;
        if ((rc = up<float>(s, s->owned, &s->pred, nullptr, (size_t)c1.V * 3))) return rc;
        std::vector<float> norm(2 * kFeat + 6, 0.f);
        for (int q = 0; q < kFeat; q++) norm[kFeat + q] = 1.f;
        norm[2 * kFeat + 3] = norm[2 * kFeat + 4] = norm[2 * kFeat + 5] = 1.f;
        if ((rc = up(s, s->owned, &s->norm, norm.data(), norm.size()))) return rc;
        std::vector<float4> z(c1.V, make_float4(0.f, 0.f, 0.f, 0.f)), x1(c1.V);
        for (int i = 0; i < c1.V; i++) x1[i] = make_float4(c1.h.X[i * 3], c1.h.X[i * 3 + 1], c1.h.X[i * 3 + 2], 0.f);
        if ((rc = up(s, s->owned, &s->epred, z.data(), c1.V))) return rc;
        if ((rc = up(s, s->owned, &s->X1, x1.data(), c1.V))) return rc;
        std::vector<float4> zf(s->L[0].V, make_float4(0.f, 0.f, 0.f, 0.f));
        if ((rc = up(s, s->owned, &s->D1, zf.data(), zf.size()))) return rc;
        if ((rc = up(s, s->owned, &s->D2, zf.data(), zf.size()))) return rc;
        if ((rc = mlp_create(&s->mlp, kFeat, 3, 1234u, s->stream))) return rc;
    }
    if ((rc = dalloc(&s->vstate, 1))) return rc;
    s->owned.push_back(s->vstate);
    // two norms (dual start) per block of the residual kernels at any lane count
    if ((rc = dalloc(&s->vpart, 2 * (size_t)nblk(s->L[0].V * kMaxVertLanes, 256)))) return rc;
    s->owned.push_back(s->vpart);
    if ((rc = dalloc(&s->seqc, 1))) return rc;
    s->owned.push_back(s->seqc);
    HIPCHK(hipMemsetAsync(s->seqc, 0, sizeof(unsigned), s->stream));
    s->chk_launched = 0;
    HIPCHK(hipHostMalloc((void**)&s->hmirror, 8 * sizeof(int), hipHostMallocCoherent));
    std::memset(s->hmirror, 0, 8 * sizeof(int));
    return push_vstate(s);
}

int ensure_partials(mgsb_solver* s, size_t n) {
    if (s->npart >= n) return MGSB_OK;
    if (s->dpart) HIPCHK(hipFree(s->dpart));
    s->dpart = nullptr;
    int rc = dalloc(&s->dpart, n);
    if (rc) return rc;
    s->npart = n;
    return MGSB_OK;
}

int reduce_finish(mgsb_solver* s, int nb, int nv, double* host_out) {
    if (nv == 1)
        k_reduce_final<1><<<1, 256, 0, s->stream>>>(nb, s->dpart, s->dred);
    else
        k_reduce_final<4><<<1, 256, 0, s->stream>>>(nb, s->dpart, s->dred);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host_out, s->dred, sizeof(double) * nv, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return MGSB_OK;
}

// ||g(x)|| of level 0 (FP64).
int residual_norm(mgsb_solver* s, double* out) {
    DevLevel& d = s->L[0];
    int rc = launch_tet(s, d, tet_mode_of(s->p.if_mode), false);
    if (rc) return rc;
    const int lsh = vert_lsh(s), nb = nblk(d.V << lsh, 256);
    if ((rc = ensure_partials(s, (size_t)nb * 4))) return rc;
    k_resid<<<nb, 256, 0, s->stream>>>(d.V, d.x, d.s, d.vc, gath_rp(s, d), gath_cf(s, d), s->dcol, kVertPoint, nullptr, s->dpart,
                                       nullptr, nullptr, 0, nullptr, nullptr, lsh);
    HIPCHK(hipGetLastError());
    double sq = 0;
    if ((rc = reduce_finish(s, nb, 1, &sq))) return rc;
    *out = std::sqrt(sq);
    return MGSB_OK;
}

int check_level(mgsb_solver* s, int level) {
    if (!s) {
        set_error("null solver");
        return MGSB_ERR_ARG;
    }
    if (level < 0 || level >= (int)s->L.size()) {
        set_error("level out of range");
        return MGSB_ERR_ARG;
    }
    return MGSB_OK;
}

int read_f4(mgsb_solver* s, const float4* d, int n, float* out) {
    std::vector<float4> h(n);
    HIPCHK(hipMemcpyAsync(h.data(), d, sizeof(float4) * n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (int i = 0; i < n; i++) {
        out[i * 3] = h[i].x;
        out[i * 3 + 1] = h[i].y;
        out[i * 3 + 2] = h[i].z;
    }
    return MGSB_OK;
}

int write_f4(mgsb_solver* s, float4* d, int n, const float* in) {
    std::vector<float4> h(n);
    for (int i = 0; i < n; i++) h[i] = make_float4(in[i * 3], in[i * 3 + 1], in[i * 3 + 2], 0.f);
    HIPCHK(hipMemcpyAsync(d, h.data(), sizeof(float4) * n, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return MGSB_OK;
}

int validate_params(const mgsb_params* p, int nlevels) {
    if (!(p->dt > 0) || p->iters < 0 || p->iters_coarse < 0) {
        set_error("invalid params: dt must be > 0, iteration counts >= 0");
        return MGSB_ERR_ARG;
    }
    if (p->schedule == MGSB_SCHED_TWOLEVEL && nlevels != 2) {
        set_error("two-level schedule needs exactly 2 levels");
        return MGSB_ERR_ARG;
    }
    if ((p->schedule == MGSB_SCHED_VCYCLE || p->schedule == MGSB_SCHED_LEARNED) && nlevels < 2) {
        set_error("V-cycle schedules need at least 2 levels");
        return MGSB_ERR_ARG;
    }
    if (p->pre_sweeps < 0 || p->post_sweeps < 0 || p->coarse_sweeps < 1 || p->max_cycles < 0) {
        set_error("sweep counts must be >= 0 (coarse_sweeps >= 1), max_cycles >= 0");
        return MGSB_ERR_ARG;
    }
    if (p->warm_start < 0 || p->warm_start > 2 || p->learn_gate < 0 || p->learn_gate > 1) {
        set_error("warm_start must be 0, 1 or 2 and learn_gate 0 or 1");
        return MGSB_ERR_ARG;
    }
    if (p->check_every < 0 || p->check_every > 2) {
        set_error("check_every must be 0 (no checks), 1 (after each cycle) or 2 (also after pre-smoothing)");
        return MGSB_ERR_ARG;
    }
    if (p->coarse_first < 0 || p->coarse_first > 1 || p->fuse_sweeps < 0 || p->fuse_sweeps > 2 ||
        p->stall_guard < 0 || p->stall_guard > 1) {
        set_error("coarse_first and stall_guard must be 0 or 1, fuse_sweeps 0, 1 or 2");
        return MGSB_ERR_ARG;
    }
    if (nlevels > 8) {
        set_error("at most 8 levels");
        return MGSB_ERR_ARG;
    }
    if (p->schedule < 0 || p->schedule > MGSB_SCHED_LEARNED) {
        set_error("unknown schedule");
        return MGSB_ERR_ARG;
    }
    if (p->if_mode != MGSB_IF_PLAIN && p->if_mode != MGSB_IF_AC && p->if_mode != MGSB_IF_STABLE) {
        set_error("unknown if_mode");
        return MGSB_ERR_ARG;
    }
    const int bt = p->block_tet, bv = p->block_vert;
    auto okb = [](int b) { return b == 0 || (b >= 64 && b <= 256 && b % 64 == 0); };
    if (!okb(bt) || !okb(bv)) {
        set_error("block sizes must be 0 or a multiple of 64 in [64, 256]");
        return MGSB_ERR_ARG;
    }
    if (!(p->learn_slack >= 0.f) || (p->learn_slack > 0.f && p->learn_slack < 1.f)) {
        set_error("learn_slack must be 0 (reads as 1) or >= 1");
        return MGSB_ERR_ARG;
    }
    const int vl = p->vert_lanes;
    if (vl != 0 && vl != 1 && vl != 2 && vl != 4 && vl != 8) {
        set_error("vert_lanes must be 0, 1, 2, 4 or 8");
        return MGSB_ERR_ARG;
    }
    return MGSB_OK;
}

int invalidate_graph(mgsb_solver* s) {
    // a step returns with its last graph possibly still running: no
    // executable graph is destroyed before the stream has drained
    if (s->stream && (s->gexec || !s->pseq.empty())) HIPCHK(hipStreamSynchronize(s->stream));
    if (s->gexec) HIPCHK(hipGraphExecDestroy(s->gexec));
    if (s->graph) HIPCHK(hipGraphDestroy(s->graph));
    s->gexec = nullptr;
    s->graph = nullptr;
    s->graph_valid = false;
    s->gate_rejects = s->gate_skip = s->gate_left = 0;
    s->stage_on = true;
    s->off_halves = -1;
    s->since_probe = 0;
    for (auto& kv : s->pseq) {
        if (kv.second.exec) HIPCHK(hipGraphExecDestroy(kv.second.exec));
        if (kv.second.graph) HIPCHK(hipGraphDestroy(kv.second.graph));
    }
    s->pseq.clear();
    s->precaptured = false;
    return MGSB_OK;
}

}  // namespace

extern "C" {

int mgsb_create(const mgsb_mesh* levels, int32_t nlevels, const mgsb_params* p, mgsb_solver** out) {
    try {
        if (!levels || nlevels < 1 || !p || !out) {
            set_error("mgsb_create: bad arguments");
            return MGSB_ERR_ARG;
        }
        *out = nullptr;
        int rc = validate_params(p, nlevels);
        if (rc) return rc;
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        if (p->device < 0 || p->device >= ndev) {
            set_error("mgsb_create: no HIP device " + std::to_string(p->device));
            return MGSB_ERR_HIP;
        }
        HIPCHK(hipSetDevice(p->device));
        if ((rc = scratch_setup(p->device))) return rc;
        auto s = std::make_unique<mgsb_solver>();
        s->p = *p;
        s->device = p->device;
        HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        s->L.resize(nlevels);
        for (int l = 0; l < nlevels; l++) {
            if ((rc = build_level(levels[l], p->stiffness, p->density, s->L[l].h))) return rc;
        }
        s->M.resize(nlevels - 1);
        for (int l = 0; l + 1 < nlevels; l++) {
            HostLevel& fine = s->L[l].h;
            HostLevel& coarse = s->L[l + 1].h;
            LevelMap& m = s->M[l];
            int uf = bind_points_interp(coarse, fine.X.data(), fine.V, m.h_iids, m.h_iw);
            if (uf) {
                set_error("mgsb_create: " + std::to_string(uf) + " vertices of level " + std::to_string(l) +
                          " lie in no tet of level " + std::to_string(l + 1));
                return MGSB_ERR_MESH;
            }
            int uc = bind_points_average(fine, coarse.X.data(), coarse.V, m.h_aids, m.h_aw);
            if (uc) {
                // a covering coarse level (e.g. nested voxel levels) has vertices outside
                // the finer mesh: they follow their nearest finer vertex (weight 1);
                // the reference leaves such points at weight 0 (PDSolver_MG.cpp:133-168)
                for (int j = 0; j < coarse.V; j++) {
                    const float* w = &m.h_aw[(size_t)j * 4];
                    if (w[0] != 0.f || w[1] != 0.f || w[2] != 0.f || w[3] != 0.f) continue;
                    const float* P = &coarse.X[(size_t)j * 3];
                    int best = 0;
                    double bd = 1e300;
                    for (int i = 0; i < fine.V; i++) {
                        const double dx = fine.X[(size_t)i * 3] - P[0], dy = fine.X[(size_t)i * 3 + 1] - P[1],
                                     dz = fine.X[(size_t)i * 3 + 2] - P[2];
                        const double d = dx * dx + dy * dy + dz * dz;
                        if (d < bd) {
                            bd = d;
                            best = i;
                        }
                    }
                    for (int q = 0; q < 4; q++) m.h_aids[(size_t)j * 4 + q] = best;
                    m.h_aw[(size_t)j * 4] = 1.f;
                }
            }
        }
        {  // bodies
            std::vector<int32_t> comp;
            const HostLevel& h0 = s->L[0].h;
            s->nbodies = mesh_components(h0.V, h0.T, h0.tets.data(), comp);
            s->body_of.assign(nlevels, {});
            s->body_of[0] = comp;
            for (int l = 1; l < nlevels; l++) {
                const auto& aids = s->M[l - 1].h_aids;
                auto& bo = s->body_of[l];
                bo.resize(s->L[l].h.V);
                for (int j = 0; j < s->L[l].h.V; j++) bo[j] = s->body_of[l - 1][aids[(size_t)j * 4]];
            }
            bool contiguous = true;
            for (int i = 1; i < h0.V; i++) contiguous = contiguous && comp[i] >= comp[i - 1];
            s->body_checks = s->nbodies > 1 && contiguous && nlevels >= 2;
            s->h_vb.assign(s->nbodies + 1, h0.V);
            for (int i = h0.V - 1; i >= 0; i--) s->h_vb[comp[i]] = i;
        }
        for (auto& d : s->L)
            if ((rc = upload_level(s.get(), d))) return rc;
        for (auto& m : s->M) {
            std::vector<int4> i4(m.h_iids.size() / 4), a4(m.h_aids.size() / 4);
            std::vector<float4> w4(i4.size()), b4(a4.size());
            for (size_t i = 0; i < i4.size(); i++) {
                i4[i] = make_int4(m.h_iids[i * 4], m.h_iids[i * 4 + 1], m.h_iids[i * 4 + 2], m.h_iids[i * 4 + 3]);
                w4[i] = make_float4(m.h_iw[i * 4], m.h_iw[i * 4 + 1], m.h_iw[i * 4 + 2], m.h_iw[i * 4 + 3]);
            }
            for (size_t i = 0; i < a4.size(); i++) {
                a4[i] = make_int4(m.h_aids[i * 4], m.h_aids[i * 4 + 1], m.h_aids[i * 4 + 2], m.h_aids[i * 4 + 3]);
                b4[i] = make_float4(m.h_aw[i * 4], m.h_aw[i * 4 + 1], m.h_aw[i * 4 + 2], m.h_aw[i * 4 + 3]);
            }
            if ((rc = up(s.get(), s->owned, &m.iids, i4.data(), i4.size()))) return rc;
            if ((rc = up(s.get(), s->owned, &m.iw, w4.data(), w4.size()))) return rc;
            if ((rc = up(s.get(), s->owned, &m.aids, a4.data(), a4.size()))) return rc;
            if ((rc = up(s.get(), s->owned, &m.aw, b4.data(), b4.size()))) return rc;
        }
        if (nlevels >= 2 && s->L[0].tvqp) {
            // level 0's binding into level 1 per (tile, vertex), padded as the
            // tile rows: the fused prolongation and the fused start load it
            // with the rows instead of after the vertex id
            DevLevel& d = s->L[0];
            const HostLevel& h = d.h;
            const LevelMap& m = s->M[0];
            std::vector<int4> ti((size_t)std::max(1, d.ntiles) * d.tstride, make_int4(0, 0, 0, 0));
            std::vector<float4> tw(ti.size(), make_float4(0.f, 0.f, 0.f, 0.f));
            for (int k = 0; k < d.ntiles; k++) {
                const int v0 = h.tvo[(size_t)k], nv = h.tvo[(size_t)k + 1] - v0;
                for (int l = 0; l < nv; l++) {
                    const size_t v = (size_t)h.tvid[(size_t)v0 + l], q = (size_t)k * d.tstride + l;
                    ti[q] = make_int4(m.h_iids[v * 4], m.h_iids[v * 4 + 1], m.h_iids[v * 4 + 2], m.h_iids[v * 4 + 3]);
                    tw[q] = make_float4(m.h_iw[v * 4], m.h_iw[v * 4 + 1], m.h_iw[v * 4 + 2], m.h_iw[v * 4 + 3]);
                }
            }
            if ((rc = up(s.get(), s->owned, &d.tiid, ti.data(), ti.size()))) return rc;
            if ((rc = up(s.get(), s->owned, &d.tiw, tw.data(), tw.size()))) return rc;
        }
        if ((rc = dalloc(&s->dcol, 1))) return rc;
        s->owned.push_back(s->dcol);
        std::memset(&s->hcol, 0, sizeof(DevColliders));
        s->hcol.kcol = p->kcol;
        if ((rc = push_colliders(s.get()))) return rc;
        if ((rc = dalloc(&s->dred, 16))) return rc;
        s->owned.push_back(s->dred);
        if (nlevels >= 2) {
            if ((rc = alloc_vcycle(s.get()))) return rc;
            if ((rc = setup_vcycle(s.get()))) return rc;
        }
        if (s->body_checks) {
            const int nb = s->nbodies;
            if ((rc = up(s.get(), s->owned, &s->vb, s->h_vb.data(), s->h_vb.size()))) return rc;
            if ((rc = up<double>(s.get(), s->owned, &s->g2, nullptr, (size_t)s->L[0].V))) return rc;
            if ((rc = up<double>(s.get(), s->owned, &s->bodysq, nullptr, (size_t)nb))) return rc;
            if ((rc = up<double>(s.get(), s->owned, &s->res0b, nullptr, (size_t)nb))) return rc;
            std::vector<unsigned> z(1, 0u);
            if ((rc = up(s.get(), s->owned, &s->bcount, z.data(), 1))) return rc;
        }
        HIPCHK(hipStreamSynchronize(s->stream));
        *out = s.release();
        return MGSB_OK;
    } catch (const std::exception& e) {
        set_error(std::string("mgsb_create: ") + e.what());
        return MGSB_ERR_NOMEM;
    }
}

void mgsb_destroy(mgsb_solver* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->dpart) (void)hipFree(s->dpart);
    delete s;
}

int mgsb_set_params(mgsb_solver* s, const mgsb_params* p) {
    if (!s || !p) {
        set_error("mgsb_set_params: null argument");
        return MGSB_ERR_ARG;
    }
    int rc = validate_params(p, (int)s->L.size());
    if (rc) return rc;
    if (p->stiffness != s->p.stiffness || p->density != s->p.density || p->device != s->p.device) {
        set_error("mgsb_set_params: stiffness, density and device are fixed at create");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    const bool dt_changed = p->dt != s->p.dt;
    s->p = *p;
    s->hcol.kcol = p->kcol;
    if ((rc = pull_colliders(s))) return rc;
    s->hcol.kcol = p->kcol;
    if ((rc = push_colliders(s))) return rc;
    if (dt_changed)
        for (auto& d : s->L)
            if ((rc = upload_vc(s, d))) return rc;
    if (s->L.size() >= 2) {
        if ((rc = setup_vcycle(s))) return rc;
        if ((rc = push_vstate(s))) return rc;
    }
    return invalidate_graph(s);
}

int mgsb_set_fixed(mgsb_solver* s, int32_t level, const float* fixflag) {
    int rc = check_level(s, level);
    if (rc) return rc;
    if (!fixflag) {
        set_error("mgsb_set_fixed: null array");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    DevLevel& d = s->L[level];
    for (int i = 0; i < d.V; i++) d.h.fix[i] = fixflag[i];
    int rc2 = upload_vc(s, d);
    if (rc2) return rc2;
    return s->L.size() >= 2 ? setup_vcycle(s) : MGSB_OK;
}

int mgsb_fix_spheres(mgsb_solver* s, const float* spheres, int32_t n) {
    if (!s || (n > 0 && !spheres) || n < 0) {
        set_error("mgsb_fix_spheres: bad arguments");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    for (auto& d : s->L) {
        for (int i = 0; i < d.V; i++) {
            const float* X = &d.h.X[(size_t)i * 3];
            for (int q = 0; q < n; q++) {
                const float* c = spheres + q * 4;
                const float dx = X[0] - c[0], dy = X[1] - c[1], dz = X[2] - c[2];
                if (std::sqrt(dx * dx + dy * dy + dz * dz) < c[3]) {
                    d.h.fix[i] = 0.f;
                    break;
                }
            }
        }
        int rc = upload_vc(s, d);
        if (rc) return rc;
    }
    return s->L.size() >= 2 ? setup_vcycle(s) : MGSB_OK;
}

int mgsb_fix_plane(mgsb_solver* s, int32_t axis, float value, float tol) {
    if (!s || axis < 0 || axis > 2) {
        set_error("mgsb_fix_plane: bad arguments");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    for (auto& d : s->L) {
        for (int i = 0; i < d.V; i++)
            if (std::fabs(d.h.X[(size_t)i * 3 + axis] - value) < tol) d.h.fix[i] = 0.f;
        int rc = upload_vc(s, d);
        if (rc) return rc;
    }
    return s->L.size() >= 2 ? setup_vcycle(s) : MGSB_OK;
}

int mgsb_set_sphere(mgsb_solver* s, int32_t id, const float center[3], float radius, const float vel[3],
                    int32_t active) {
    if (!s || id < 0 || id >= MGSB_MAX_SPHERES || !center) {
        set_error("mgsb_set_sphere: bad arguments");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    int rc = pull_colliders(s);
    if (rc) return rc;
    s->hcol.sph[id] = make_float4(center[0], center[1], center[2], radius);
    s->hcol.vel[id] = make_float4(vel ? vel[0] : 0.f, vel ? vel[1] : 0.f, vel ? vel[2] : 0.f, active ? 1.f : 0.f);
    if (id + 1 > s->hcol.nsph) s->hcol.nsph = id + 1;
    return push_colliders(s);
}

int mgsb_set_plane(mgsb_solver* s, const float n[3], float d, float k, int32_t active) {
    if (!s || !n) {
        set_error("mgsb_set_plane: bad arguments");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    int rc = pull_colliders(s);
    if (rc) return rc;
    s->hcol.plane = make_float4(n[0], n[1], n[2], d);
    s->hcol.kplane = k;
    s->hcol.plane_on = active ? 1 : 0;
    return push_colliders(s);
}

int mgsb_set_velocity(mgsb_solver* s, const float* v) {
    if (!s || !v) {
        set_error("mgsb_set_velocity: bad arguments");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    return write_f4(s, s->L[0].v, s->L[0].V, v);
}

int mgsb_set_positions(mgsb_solver* s, int32_t level, const float* x) {
    int rc = check_level(s, level);
    if (rc) return rc;
    if (!x) {
        set_error("mgsb_set_positions: null array");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    if (level == 0 && s->L[0].U1) {  // the warm starts' history no longer applies
        HIPCHK(hipMemsetAsync(s->L[0].U1, 0, sizeof(float4) * s->L[0].V, s->stream));
        HIPCHK(hipMemsetAsync(s->L[0].U2, 0, sizeof(float4) * s->L[0].V, s->stream));
    }
    if (level == 0 && s->D1) {
        HIPCHK(hipMemsetAsync(s->D1, 0, sizeof(float4) * s->L[0].V, s->stream));
        HIPCHK(hipMemsetAsync(s->D2, 0, sizeof(float4) * s->L[0].V, s->stream));
    }
    return write_f4(s, s->L[level].x, s->L[level].V, x);
}

int mgsb_step(mgsb_solver* s, int32_t nsteps, mgsb_step_stats* stats) {
    if (!s || nsteps < 0) {
        set_error("mgsb_step: bad arguments");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    int rc;
    const bool vcycle = s->p.schedule == MGSB_SCHED_VCYCLE || s->p.schedule == MGSB_SCHED_LEARNED;
    for (int st = 0; st < nsteps; st++) {
        if (s->recording && s->rec_steps >= s->rec_cap) s->recording = false;
        const bool graphable = s->p.use_graph && !s->timing && !s->recording;
        if (vcycle) {
            if ((rc = step_vcycle(s, graphable))) return rc;
        } else if (graphable) {
            if (!s->graph_valid) {
                HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
                rc = enqueue_step(s);
                hipGraph_t g = nullptr;
                hipError_t ce = hipStreamEndCapture(s->stream, &g);
                if (rc) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                HIPCHK(ce);
                s->graph = g;
                HIPCHK(hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0));
                s->graph_valid = true;
            }
            HIPCHK(hipGraphLaunch(s->gexec, s->stream));
        } else {
            if ((rc = enqueue_step(s))) return rc;
        }
    }
    s->last.steps = nsteps;
    if (s->vstate && (s->p.schedule == MGSB_SCHED_VCYCLE || s->p.schedule == MGSB_SCHED_LEARNED) && stats) {
        VState v;
        HIPCHK(hipMemcpyAsync(&v, s->vstate, sizeof(VState), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        s->last.res0 = v.res0;
        s->last.res = v.res;
        s->last.rel = v.rel;
        s->last.cycles = v.cycles;
        s->last.halves = v.halves;
        s->last.fallback = v.fallback;
    }
    if (stats) *stats = s->last;
    HIPCHK(hipGetLastError());
    return MGSB_OK;
}

int mgsb_sync(mgsb_solver* s) {
    if (!s) {
        set_error("mgsb_sync: null solver");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    return MGSB_OK;
}

int mgsb_read_positions(mgsb_solver* s, int32_t level, float* xyz) {
    int rc = check_level(s, level);
    if (rc) return rc;
    if (!xyz) {
        set_error("mgsb_read_positions: null array");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    return read_f4(s, s->L[level].x, s->L[level].V, xyz);
}

int mgsb_read_velocity(mgsb_solver* s, int32_t level, float* v) {
    int rc = check_level(s, level);
    if (rc) return rc;
    if (!v) {
        set_error("mgsb_read_velocity: null array");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    return read_f4(s, s->L[level].v, s->L[level].V, v);
}

int mgsb_read_fixed(mgsb_solver* s, int32_t level, float* fixflag) {
    int rc = check_level(s, level);
    if (rc) return rc;
    if (!fixflag) {
        set_error("mgsb_read_fixed: null array");
        return MGSB_ERR_ARG;
    }
    const std::vector<float>& f = s->L[level].h.fix;
    std::copy(f.begin(), f.end(), fixflag);
    return MGSB_OK;
}

static int compute_normals(mgsb_solver* s) {
    DevLevel& d = s->L[0];
    const int bs = 256;
    if (d.S > 0) {
        k_tri_normal<<<nblk(d.S, bs), bs, 0, s->stream>>>(d.S, d.x, d.tri, d.tn);
        HIPCHK(hipGetLastError());
    }
    k_vert_normal<<<nblk(d.V, bs), bs, 0, s->stream>>>(d.V, d.trp, d.trid, d.tn, d.vn);
    HIPCHK(hipGetLastError());
    return MGSB_OK;
}

int mgsb_read_normals(mgsb_solver* s, float* n) {
    if (!s || !n) {
        set_error("mgsb_read_normals: bad arguments");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    int rc = compute_normals(s);
    if (rc) return rc;
    return read_f4(s, s->L[0].vn, s->L[0].V, n);
}

int mgsb_read_render9(mgsb_solver* s, float* v9) {
    if (!s || !v9) {
        set_error("mgsb_read_render9: bad arguments");
        return MGSB_ERR_ARG;
    }
    const int V = s->L[0].V;
    std::vector<float> x((size_t)V * 3), n((size_t)V * 3);
    int rc = mgsb_read_positions(s, 0, x.data());
    if (rc) return rc;
    if ((rc = mgsb_read_normals(s, n.data()))) return rc;
    for (int i = 0; i < V; i++) {
        for (int c = 0; c < 3; c++) {
            v9[(size_t)i * 9 + c] = x[(size_t)i * 3 + c];
            v9[(size_t)i * 9 + 3 + c] = n[(size_t)i * 3 + c];
            v9[(size_t)i * 9 + 6 + c] = 0.f;
        }
    }
    return MGSB_OK;
}

int mgsb_read_collision(mgsb_solver* s, int32_t* flags, float* depth) {
    if (!s) {
        set_error("mgsb_read_collision: null solver");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    DevLevel& d = s->L[0];
    if (flags) HIPCHK(hipMemcpyAsync(flags, d.cflag, sizeof(int) * d.V, hipMemcpyDeviceToHost, s->stream));
    if (depth) HIPCHK(hipMemcpyAsync(depth, d.cdepth, sizeof(float) * d.V, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return MGSB_OK;
}

int mgsb_energy(mgsb_solver* s, mgsb_energy_t* e) {
    if (!s || !e) {
        set_error("mgsb_energy: bad arguments");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    DevLevel& d = s->L[0];
    const int nbv = nblk(d.V, 256), nbt = nblk(d.T, 256);
    int rc = ensure_partials(s, (size_t)std::max(nbv, nbt) * 4);
    if (rc) return rc;
    k_energy_vert<<<nbv, 256, 0, s->stream>>>(d.V, d.x, d.xprev, d.s, d.vc, d.cflag, d.cenergy, s->dpart);
    HIPCHK(hipGetLastError());
    double v4[4];
    if ((rc = reduce_finish(s, nbv, 4, v4))) return rc;
    k_energy_tet<<<nbt, 256, 0, s->stream>>>(d.T, d.x, d.tet, d.B, d.vol, s->p.stiffness, d.tk, s->dpart);
    HIPCHK(hipGetLastError());
    double ep = 0;
    if ((rc = reduce_finish(s, nbt, 1, &ep))) return rc;
    const double dt = s->p.dt;
    e->dX = std::sqrt(v4[0]);
    e->Ek = v4[1] / (2 * dt * dt);
    e->Ep = ep;
    e->Ec = v4[2];
    e->Nc = v4[3];
    return MGSB_OK;
}

int mgsb_energy_noalloc(mgsb_solver* s, mgsb_energy_t* e) {
    if (!s || !e) {
        set_error("mgsb_energy_noalloc: bad arguments");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    double* sp = scratch_of(s->device);
    if (!sp) {
        set_error("mgsb_energy_noalloc: no scratch address for this device");
        return MGSB_ERR_STATE;
    }
    DevLevel& d = s->L[0];
    const int nbv = nblk(d.V, 256), nbt = nblk(d.T, 256);
    const size_t cap = (kScratchDoubles - kScratchResults) / 4;  // blocks of 256 items
    if ((size_t)std::max(nbv, nbt) > cap) {
        set_error("mgsb_energy_noalloc: the finest level has more than " + std::to_string(cap * 256) +
                  " vertices or tets (fixed scratch of 1 MiB); use mgsb_energy");
        return MGSB_ERR_ARG;
    }
    // every block writes its own partials and each final pass writes its
    // results, so nothing in the scratch has to start at zero
    double* part = sp;
    double* res = sp + (kScratchDoubles - kScratchResults);
    k_energy_vert<<<nbv, 256, 0, s->stream>>>(d.V, d.x, d.xprev, d.s, d.vc, d.cflag, d.cenergy, part);
    HIPCHK(hipGetLastError());
    k_reduce_final<4><<<1, 256, 0, s->stream>>>(nbv, part, res);
    HIPCHK(hipGetLastError());
    k_energy_tet<<<nbt, 256, 0, s->stream>>>(d.T, d.x, d.tet, d.B, d.vol, s->p.stiffness, d.tk, part);
    HIPCHK(hipGetLastError());
    k_reduce_final<1><<<1, 256, 0, s->stream>>>(nbt, part, res + 4);
    HIPCHK(hipGetLastError());
    double out[5];
    HIPCHK(hipMemcpyAsync(out, res, sizeof(out), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    const double dt = s->p.dt;
    e->dX = std::sqrt(out[0]);
    e->Ek = out[1] / (2 * dt * dt);
    e->Ep = out[4];
    e->Ec = out[2];
    e->Nc = out[3];
    return MGSB_OK;
}

}  // extern "C"

namespace {

// The five energy terms of level 0 into out[0..5) on the device
// (dX^2, m |x - s|^2, Ec, Nc, Ep), enqueued without a host wait.
int energy_enqueue(mgsb_solver* s, double* out) {
    DevLevel& d = s->L[0];
    const int nbv = nblk(d.V, 256), nbt = nblk(d.T, 256);
    int rc = ensure_partials(s, (size_t)std::max(nbv, nbt) * 4);
    if (rc) return rc;
    k_energy_vert<<<nbv, 256, 0, s->stream>>>(d.V, d.x, d.xprev, d.s, d.vc, d.cflag, d.cenergy, s->dpart);
    HIPCHK(hipGetLastError());
    k_reduce_final<4><<<1, 256, 0, s->stream>>>(nbv, s->dpart, out);
    HIPCHK(hipGetLastError());
    k_energy_tet<<<nbt, 256, 0, s->stream>>>(d.T, d.x, d.tet, d.B, d.vol, s->p.stiffness, d.tk, s->dpart);
    HIPCHK(hipGetLastError());
    k_reduce_final<1><<<1, 256, 0, s->stream>>>(nbt, s->dpart, out + 4);
    HIPCHK(hipGetLastError());
    return MGSB_OK;
}

void energy_terms(const double* r, double dt, mgsb_energy_t* e) {
    e->dX = std::sqrt(r[0]);
    e->Ek = r[1] / (2 * dt * dt);
    e->Ep = r[4];
    e->Ec = r[2];
    e->Nc = r[3];
}

// Point DCD (spheres and plane) at the current positions of level 0, writing
// the collision flags, depths and energies (DCDByPoint, PDSolver.cpp:228-233).
int dcd_refresh(mgsb_solver* s) {
    DevLevel& d = s->L[0];
    return launch_vert(s, d, 1.0f, kVertPoint | kVertWriteCol | kVertDcdOnly);
}

// PDSolver::StepForConvergence (PDSolver.cpp:202-221) on level 0: save the
// velocity, calculateST with the colliders where the next step will see
// them, `iters` reference iterations with calculateIF's local step and the
// energy after each, then the positions and velocity restored
// (runResetPosVel, PDSolver.cu:270-273).  E receives iters + 1 totals
// Ek + Ep + Ec (E[0] at s, after a DCD at s), last the terms of the final
// iterate.  Nothing else of the solver's state is changed that the next
// step does not rebuild.
int convergence_reference(mgsb_solver* s, int iters, double* E, mgsb_energy_t* last) {
    DevLevel& d = s->L[0];
    int rc = pull_colliders(s);
    if (rc) return rc;
    const DevColliders saved = s->hcol;
    for (int q = 0; q < s->hcol.nsph; q++) {  // the advance the next step's calculateST applies
        float4& c = s->hcol.sph[q];
        const float4 w = s->hcol.vel[q];
        c.x = c.x + w.x * s->p.dt;
        c.y = c.y + w.y * s->p.dt;
        c.z = c.z + w.z * s->p.dt;
    }
    if ((rc = push_colliders(s))) return rc;
    float4* vbak = nullptr;
    double* er = nullptr;
    if ((rc = dalloc(&vbak, (size_t)d.V))) return rc;
    if ((rc = dalloc(&er, (size_t)(iters + 1) * 5))) {
        (void)hipFree(vbak);
        return rc;
    }
    std::vector<double> h((size_t)(iters + 1) * 5);
    auto body = [&]() -> int {
        int r;
        HIPCHK(hipMemcpyAsync(vbak, d.v, sizeof(float4) * d.V, hipMemcpyDeviceToDevice, s->stream));
        if ((r = launch_st(s, d, 0))) return r;
        if ((r = dcd_refresh(s))) return r;
        if ((r = energy_enqueue(s, er))) return r;
        float omega = 1.0f;
        const float rho = s->p.rho;
        for (int i = 0; i < iters; i++) {
            omega = 4 / (4 - rho * rho * omega);
            if ((r = pd_iteration(s, d, MGSB_IF_PLAIN, omega, false, true))) return r;
            if ((r = energy_enqueue(s, er + (size_t)(i + 1) * 5))) return r;
        }
        HIPCHK(hipMemcpyAsync(d.x, d.xlast, sizeof(float4) * d.V, hipMemcpyDeviceToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(d.v, vbak, sizeof(float4) * d.V, hipMemcpyDeviceToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(h.data(), er, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        return MGSB_OK;
    };
    const bool was = s->timing;
    s->timing = false;
    rc = body();
    s->timing = was;
    (void)hipFree(vbak);
    (void)hipFree(er);
    if (rc) return rc;
    s->hcol = saved;
    if ((rc = push_colliders(s))) return rc;
    for (int i = 0; i <= iters; i++) {
        mgsb_energy_t t;
        energy_terms(&h[(size_t)i * 5], s->p.dt, &t);
        if (E) E[i] = t.Ek + t.Ep + t.Ec;
        if (last && i == iters) *last = t;
    }
    return MGSB_OK;
}

}  // namespace

extern "C" {

int mgsb_step_convergence(mgsb_solver* s, int32_t iters, double* E, mgsb_energy_t* last) {
    if (!s || iters < 1 || !E) {
        set_error("mgsb_step_convergence: iters >= 1 and E (iters + 1 doubles) required");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    return convergence_reference(s, iters, E, last);
}

int mgsb_step_error(mgsb_solver* s, int32_t iters, double* err, double* E) {
    if (!s || iters < 1 || !err) {
        set_error("mgsb_step_error: iters >= 1 and err required");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    std::vector<double> Eh((size_t)iters + 1);
    int rc = convergence_reference(s, iters, Eh.data(), nullptr);
    if (rc) return rc;
    if ((rc = mgsb_step(s, 1, nullptr))) return rc;
    if ((rc = dcd_refresh(s))) return rc;
    double* er = nullptr;
    if ((rc = dalloc(&er, 5))) return rc;
    double r[5];
    rc = energy_enqueue(s, er);
    hipError_t e = hipSuccess;
    if (!rc) e = hipMemcpyAsync(r, er, sizeof(r), hipMemcpyDeviceToHost, s->stream);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(er);
    if (rc) return rc;
    HIPCHK(e);
    mgsb_energy_t t;
    energy_terms(r, s->p.dt, &t);
    const double E0 = Eh[0], Ec = Eh[(size_t)iters], Ee = t.Ek + t.Ep + t.Ec;
    *err = (Ee - Ec) / (E0 - Ec);
    if (E) {
        E[0] = E0;
        E[1] = Ec;
        E[2] = Ee;
    }
    return MGSB_OK;
}

int mgsb_read_ep_density(mgsb_solver* s, float* density) {
    if (!s || !density) {
        set_error("mgsb_read_ep_density: bad arguments");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    DevLevel& d = s->L[0];
    float* dd = nullptr;
    int rc = dalloc(&dd, (size_t)d.V);
    if (rc) return rc;
    const int bs = 256;
    k_ep_tet<<<nblk(d.T, bs), bs, 0, s->stream>>>(d.T, d.x, d.tet, d.B, d.vol, s->p.stiffness, d.tk, d.slot, d.cf);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        k_ep_vert<<<nblk(d.V, bs), bs, 0, s->stream>>>(d.V, d.rp, d.cf, dd);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(density, dd, sizeof(float) * d.V, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(dd);
    HIPCHK(e);
    return MGSB_OK;
}

int mgsb_residual(mgsb_solver* s, double* res) {
    if (!s || !res) {
        set_error("mgsb_residual: bad arguments");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    return residual_norm(s, res);
}

int mgsb_level_info(mgsb_solver* s, int32_t level, int32_t* nverts, int32_t* ntets, int32_t* ntris) {
    int rc = check_level(s, level);
    if (rc) return rc;
    if (nverts) *nverts = s->L[level].V;
    if (ntets) *ntets = s->L[level].T;
    if (ntris) *ntris = s->L[level].S;
    return MGSB_OK;
}

int mgsb_step_history(mgsb_solver* s, double* rel, int32_t n, int32_t* count) {
    if (!s || !count || n < 0 || (n > 0 && !rel)) {
        set_error("mgsb_step_history: bad arguments");
        return MGSB_ERR_ARG;
    }
    *count = 0;
    if (!s->vstate) return MGSB_OK;
    HIPCHK(hipSetDevice(s->device));
    VState v;
    HIPCHK(hipMemcpyAsync(&v, s->vstate, sizeof(VState), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    *count = v.nhist;
    for (int i = 0; i < std::min<int>(n, v.nhist); i++) rel[i] = v.hist[i];
    return MGSB_OK;
}

int mgsb_run_stats(mgsb_solver* s, mgsb_run_stats_t* out, int32_t reset) {
    if (!s || !out) {
        set_error("mgsb_run_stats: bad arguments");
        return MGSB_ERR_ARG;
    }
    std::memset(out, 0, sizeof(*out));
    if (!s->vstate) return MGSB_OK;
    HIPCHK(hipSetDevice(s->device));
    VState v;
    HIPCHK(hipMemcpyAsync(&v, s->vstate, sizeof(VState), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    out->max_rel = v.max_rel;
    out->total_cycles = v.total_cycles;
    out->steps = v.nsteps;
    out->failed_steps = v.nfail;
    out->fallbacks = v.nfallback;
    out->total_halves = v.total_halves;
    if (reset) {
        v.max_rel = 0.0;
        v.total_cycles = 0;
        v.total_halves = 0;
        v.nsteps = 0;
        v.nfail = 0;
        v.nfallback = 0;
        HIPCHK(hipMemcpyAsync(s->vstate, &v, sizeof(VState), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    return MGSB_OK;
}

static int need_mlp(mgsb_solver* s, const char* what) {
    if (!s || !s->mlp) {
        set_error(std::string(what) + ": the solver has no learned correction (needs >= 2 levels)");
        return MGSB_ERR_STATE;
    }
    HIPCHK(hipSetDevice(s->device));
    return MGSB_OK;
}

int mgsb_learn_record(mgsb_solver* s, int32_t max_steps) {
    int rc = need_mlp(s, "mgsb_learn_record");
    if (rc) return rc;
    if (max_steps < 0) {
        set_error("mgsb_learn_record: max_steps < 0");
        return MGSB_ERR_ARG;
    }
    if (max_steps == 0) {
        s->recording = false;
        return MGSB_OK;
    }
    const int V1 = s->L[1].V;
    if (max_steps > s->rec_cap) {
        float *nx = nullptr, *nt = nullptr;
        if ((rc = dalloc(&nx, (size_t)max_steps * V1 * kFeat))) return rc;
        if ((rc = dalloc(&nt, (size_t)max_steps * V1 * 3))) {
            (void)hipFree(nx);
            return rc;
        }
        if (s->rec_steps > 0) {
            HIPCHK(hipMemcpyAsync(nx, s->recX, sizeof(float) * (size_t)s->rec_steps * V1 * kFeat,
                                  hipMemcpyDeviceToDevice, s->stream));
            HIPCHK(hipMemcpyAsync(nt, s->recT, sizeof(float) * (size_t)s->rec_steps * V1 * 3, hipMemcpyDeviceToDevice,
                                  s->stream));
            HIPCHK(hipStreamSynchronize(s->stream));
        }
        if (s->recX) HIPCHK(hipFree(s->recX));
        if (s->recT) HIPCHK(hipFree(s->recT));
        s->recX = nx;
        s->recT = nt;
        s->rec_cap = max_steps;
    }
    s->recording = true;
    return invalidate_graph(s);
}

int mgsb_learn_samples(mgsb_solver* s, int64_t* n) {
    int rc = need_mlp(s, "mgsb_learn_samples");
    if (rc) return rc;
    if (!n) {
        set_error("mgsb_learn_samples: null output");
        return MGSB_ERR_ARG;
    }
    *n = (int64_t)s->rec_steps * s->L[1].V;
    return MGSB_OK;
}

int mgsb_learn_clear(mgsb_solver* s) {
    int rc = need_mlp(s, "mgsb_learn_clear");
    if (rc) return rc;
    s->rec_steps = 0;
    return MGSB_OK;
}

int mgsb_learn_train(mgsb_solver* s, int32_t epochs, int32_t batch, float lr, uint32_t seed, float* loss_hist) {
    int rc = need_mlp(s, "mgsb_learn_train");
    if (rc) return rc;
    const long n = (long)s->rec_steps * s->L[1].V;
    if (n < batch || epochs < 1 || !(lr > 0)) {
        set_error("mgsb_learn_train: need recorded samples >= batch, epochs >= 1, lr > 0");
        return MGSB_ERR_ARG;
    }
    std::vector<float> X((size_t)n * kFeat), T((size_t)n * 3);
    HIPCHK(hipMemcpyAsync(X.data(), s->recX, sizeof(float) * X.size(), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(T.data(), s->recT, sizeof(float) * T.size(), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    // per-feature / per-output mean and standard deviation (FP64), then
    // standardise; constant features get sd 1
    std::vector<float> norm(2 * kFeat + 6);
    auto stats = [&](const std::vector<float>& A, int w, int c, float& mu, float& sd) {
        double m = 0, q = 0;
        for (long i = 0; i < n; i++) m += A[(size_t)i * w + c];
        m /= n;
        for (long i = 0; i < n; i++) {
            double d = A[(size_t)i * w + c] - m;
            q += d * d;
        }
        q = std::sqrt(q / n);
        mu = (float)m;
        sd = q > 1e-12 ? (float)q : 1.f;
    };
    for (int c = 0; c < kFeat; c++) stats(X, kFeat, c, norm[c], norm[kFeat + c]);
    for (int c = 0; c < 3; c++) stats(T, 3, c, norm[2 * kFeat + c], norm[2 * kFeat + 3 + c]);
    for (long i = 0; i < n; i++) {
        for (int c = 0; c < kFeat; c++) X[(size_t)i * kFeat + c] = (X[(size_t)i * kFeat + c] - norm[c]) / norm[kFeat + c];
        for (int c = 0; c < 3; c++)
            T[(size_t)i * 3 + c] = (T[(size_t)i * 3 + c] - norm[2 * kFeat + c]) / norm[2 * kFeat + 3 + c];
    }
    float *dX = nullptr, *dT = nullptr;
    if ((rc = dalloc(&dX, X.size()))) return rc;
    if ((rc = dalloc(&dT, T.size()))) {
        (void)hipFree(dX);
        return rc;
    }
    HIPCHK(hipMemcpyAsync(dX, X.data(), sizeof(float) * X.size(), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(dT, T.data(), sizeof(float) * T.size(), hipMemcpyHostToDevice, s->stream));
    rc = mlp_train(s->mlp, dX, dT, n, epochs, batch, lr, seed, loss_hist, s->stream);
    (void)hipStreamSynchronize(s->stream);
    (void)hipFree(dX);
    (void)hipFree(dT);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(s->norm, norm.data(), sizeof(float) * norm.size(), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return invalidate_graph(s);
}

int mgsb_learn_nparams(mgsb_solver* s, int32_t* n) {
    int rc = need_mlp(s, "mgsb_learn_nparams");
    if (rc) return rc;
    if (!n) {
        set_error("mgsb_learn_nparams: null output");
        return MGSB_ERR_ARG;
    }
    *n = mlp_params_count(s->mlp);
    return MGSB_OK;
}

int mgsb_learn_weights(mgsb_solver* s, float* w, int32_t n, int32_t set) {
    int rc = need_mlp(s, "mgsb_learn_weights");
    if (rc) return rc;
    if (!w) {
        set_error("mgsb_learn_weights: null array");
        return MGSB_ERR_ARG;
    }
    rc = set ? mlp_set_params(s->mlp, w, n, s->stream) : mlp_get_params(s->mlp, w, n, s->stream);
    if (rc) return rc;
    return set ? invalidate_graph(s) : MGSB_OK;
}

int mgsb_learn_norm(mgsb_solver* s, float* norm, int32_t set) {
    int rc = need_mlp(s, "mgsb_learn_norm");
    if (rc) return rc;
    if (!norm) {
        set_error("mgsb_learn_norm: null array");
        return MGSB_ERR_ARG;
    }
    const size_t nb = sizeof(float) * (2 * kFeat + 6);
    if (set)
        HIPCHK(hipMemcpyAsync(s->norm, norm, nb, hipMemcpyHostToDevice, s->stream));
    else
        HIPCHK(hipMemcpyAsync(norm, s->norm, nb, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return MGSB_OK;
}

int mgsb_learn_last(mgsb_solver* s, float* e1) {
    int rc = need_mlp(s, "mgsb_learn_last");
    if (rc) return rc;
    if (!e1) {
        set_error("mgsb_learn_last: null array");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    return read_f4(s, s->epred, s->L[1].V, e1);
}

int mgsb_mlp_forward(mgsb_solver* s, const float* x, int32_t n, float* y, int32_t path) {
    int rc = need_mlp(s, "mgsb_mlp_forward");
    if (rc) return rc;
    if (!x || !y || n < 1) {
        set_error("mgsb_mlp_forward: bad arguments");
        return MGSB_ERR_ARG;
    }
    if (path == 1) return mlp_forward_train_path(s->mlp, x, n, y, s->stream);
    float *dx = nullptr, *dy = nullptr;
    if ((rc = dalloc(&dx, (size_t)n * kFeat))) return rc;
    if ((rc = dalloc(&dy, (size_t)n * 3))) {
        (void)hipFree(dx);
        return rc;
    }
    HIPCHK(hipMemcpyAsync(dx, x, sizeof(float) * (size_t)n * kFeat, hipMemcpyHostToDevice, s->stream));
    rc = mlp_infer(s->mlp, dx, n, dy, 3, s->stream, nullptr);
    if (!rc) {
        HIPCHK(hipMemcpyAsync(y, dy, sizeof(float) * (size_t)n * 3, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    (void)hipFree(dx);
    (void)hipFree(dy);
    return rc;
}

int mgsb_mlp_grad(mgsb_solver* s, const float* x, const float* t, int32_t n, float* grad, double* loss) {
    int rc = need_mlp(s, "mgsb_mlp_grad");
    if (rc) return rc;
    if (!x || !t || !grad || n < 1) {
        set_error("mgsb_mlp_grad: bad arguments");
        return MGSB_ERR_ARG;
    }
    return mlp_grad(s->mlp, x, t, n, grad, loss, s->stream);
}

int mgsb_level_rho(mgsb_solver* s, int32_t level, double* rho) {
    int rc = check_level(s, level);
    if (rc) return rc;
    if (!rho) {
        set_error("mgsb_level_rho: null output");
        return MGSB_ERR_ARG;
    }
    *rho = s->L[level].rho;
    return MGSB_OK;
}

int mgsb_coarse_nfree(mgsb_solver* s, int32_t* nfree) {
    if (!s || !nfree) {
        set_error("mgsb_coarse_nfree: bad arguments");
        return MGSB_ERR_ARG;
    }
    *nfree = s->L.size() >= 2 ? s->L.back().nfree : -1;
    return MGSB_OK;
}

int mgsb_set_timing(mgsb_solver* s, int32_t on) {
    if (!s) {
        set_error("mgsb_set_timing: null solver");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->timing = on != 0;
    s->ev_used = 0;
    s->time_ms = 0.0;
    s->time_launches = 0;
    return MGSB_OK;
}

int mgsb_kernel_time(mgsb_solver* s, int32_t kernel, double* ms, int64_t* launches) {
    if (!s || kernel != 0) {
        set_error("mgsb_kernel_time: only kernel 0 (per-tet local step) is timed");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (size_t i = 0; i < s->ev_used; i++) {
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, s->ev_pool[i].first, s->ev_pool[i].second));
        s->time_ms += t;
        s->time_launches++;
    }
    s->ev_used = 0;
    if (ms) *ms = s->time_ms;
    if (launches) *launches = s->time_launches;
    return MGSB_OK;
}

}  // extern "C"

namespace {

// Device state a step reads and writes (positions, velocities, the temporal
// history, the learned features' previous delta, colliders, convergence
// state), copied out so that trial steps can be undone.
struct StateSnap {
    std::vector<std::pair<float4*, size_t>> where;
    std::vector<std::vector<float4>> bufs;
    VState v{};
    DevColliders col{};
    int halves_guess = 0, gate_rejects = 0, gate_skip = 0, gate_left = 0, off_halves = -1, since_probe = 0;
    bool stage_on = true;
};

int snap_take(mgsb_solver* s, StateSnap& S) {
    HIPCHK(hipStreamSynchronize(s->stream));
    S.where.clear();
    for (auto& d : s->L)
        for (float4* q : {d.x, d.xprev, d.s, d.xlast, d.v}) S.where.push_back({q, (size_t)d.V});
    DevLevel& f = s->L[0];
    if (f.U1) {
        S.where.push_back({f.U1, (size_t)f.V});
        S.where.push_back({f.U2, (size_t)f.V});
    }
    if (s->D1) {
        S.where.push_back({s->D1, (size_t)f.V});
        S.where.push_back({s->D2, (size_t)f.V});
    }
    S.bufs.resize(S.where.size());
    for (size_t i = 0; i < S.where.size(); i++) {
        S.bufs[i].resize(S.where[i].second);
        HIPCHK(hipMemcpy(S.bufs[i].data(), S.where[i].first, sizeof(float4) * S.where[i].second,
                         hipMemcpyDeviceToHost));
    }
    if (s->vstate) HIPCHK(hipMemcpy(&S.v, s->vstate, sizeof(VState), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&S.col, s->dcol, sizeof(DevColliders), hipMemcpyDeviceToHost));
    S.halves_guess = s->halves_guess;
    S.gate_rejects = s->gate_rejects;
    S.gate_skip = s->gate_skip;
    S.gate_left = s->gate_left;
    S.stage_on = s->stage_on;
    S.off_halves = s->off_halves;
    S.since_probe = s->since_probe;
    return MGSB_OK;
}

int snap_restore(mgsb_solver* s, const StateSnap& S) {
    HIPCHK(hipStreamSynchronize(s->stream));
    for (size_t i = 0; i < S.where.size(); i++)
        HIPCHK(hipMemcpy(S.where[i].first, S.bufs[i].data(), sizeof(float4) * S.where[i].second,
                         hipMemcpyHostToDevice));
    if (s->vstate) HIPCHK(hipMemcpy(s->vstate, &S.v, sizeof(VState), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->dcol, &S.col, sizeof(DevColliders), hipMemcpyHostToDevice));
    s->hcol = S.col;
    s->halves_guess = S.halves_guess;
    s->gate_rejects = S.gate_rejects;
    s->gate_skip = S.gate_skip;
    s->gate_left = S.gate_left;
    s->stage_on = S.stage_on;
    s->off_halves = S.off_halves;
    s->since_probe = S.since_probe;
    return MGSB_OK;
}

}  // namespace

extern "C" {

int mgsb_autotune(mgsb_solver* s, int32_t trial_steps, mgsb_tune_result* out) {
    if (!s || trial_steps < 1 || !out) {
        set_error("mgsb_autotune: bad arguments (trial_steps >= 1)");
        return MGSB_ERR_ARG;
    }
    if (s->recording || s->timing) {
        set_error("mgsb_autotune: stop recording / kernel timing first");
        return MGSB_ERR_STATE;
    }
    HIPCHK(hipSetDevice(s->device));
    StateSnap S0;
    int rc = snap_take(s, S0);
    if (rc) return rc;
    const mgsb_params p0 = s->p;
    const bool vcycle = s->p.schedule == MGSB_SCHED_VCYCLE || s->p.schedule == MGSB_SCHED_LEARNED;
    int trials = 0;
    // 1. Launch shapes by kernel time: one fine-level sweep (local step and
    // vertex update) as the schedule runs it, 20 back-to-back between HIP
    // events.  A V-cycle sweep starts from u = 0 at the saved s, as the
    // first sweep of a step does.
    DevLevel& f = s->L[0];
    const int tmode = tet_mode_of(s->p.if_mode);
    const int vflags = vcycle ? (kVertPoint | kVertContactJacobi) : kVertPoint;
    const int split = vcycle ? 1 : 0;
    hipEvent_t ea, eb;
    HIPCHK(hipEventCreate(&ea));
    HIPCHK(hipEventCreate(&eb));
    auto sweep_us = [&](double& us) -> int {
        constexpr int reps = 20;
        int r = MGSB_OK;
        for (int pass = 0; pass < 2 && !r; pass++) {  // pass 0 warms up
            if (vcycle) {
                HIPCHK(hipMemsetAsync(f.x, 0, sizeof(float4) * f.V, s->stream));
                HIPCHK(hipMemsetAsync(f.xprev, 0, sizeof(float4) * f.V, s->stream));
            }
            HIPCHK(hipEventRecord(ea, s->stream));
            if (vcycle && s->L.size() >= 2 && fused(s, f)) {  // the vertex update inside the local step
                FuseBufs fb{f.x, f.xprev, 0};
                for (int k = 0; k < (pass ? reps : 1) && !r; k++)
                    r = launch_fused<kFuseJacobi>(s, f, fb, 1.0f, vflags, nullptr);
            } else {
                for (int k = 0; k < (pass ? reps : 1) && !r; k++) {
                    r = launch_tet(s, f, tmode, false, nullptr, nullptr, split != 0);
                    if (!r) r = launch_vert(s, f, 1.0f, vflags, nullptr, split);
                }
            }
            HIPCHK(hipEventRecord(eb, s->stream));
        }
        if (r) return r;
        HIPCHK(hipEventSynchronize(eb));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ea, eb));
        us = 1e3 * (double)ms / reps;
        trials++;
        return MGSB_OK;
    };
    struct Shape {
        int tg, bt, bv, vl, fz;
    };
    std::vector<Shape> shapes{{p0.tile_gather ? 1 : 0, p0.block_tet, p0.block_vert, p0.vert_lanes, p0.fuse_sweeps}};
    for (int bv : {64, 128, 256})
        for (int vl : {1, 2}) shapes.push_back({1, p0.block_tet, bv, vl, 0});
    if (vcycle) shapes.push_back({1, p0.block_tet, p0.block_vert, 1, 2});  // fused sweeps
    for (int bt : {64, 128, 256})
        for (int bv : {64, 128, 256})
            for (int vl : {2, 4, 8}) shapes.push_back({0, bt, bv, vl, 0});
    double base_us = 0.0, best_us = 1e300;
    Shape best = shapes[0];
    for (size_t q = 0; q < shapes.size(); q++) {
        s->p.tile_gather = shapes[q].tg;
        s->p.block_tet = shapes[q].bt;
        s->p.block_vert = shapes[q].bv;
        s->p.vert_lanes = shapes[q].vl;
        s->p.fuse_sweeps = shapes[q].fz;
        double us = 0.0;
        if ((rc = sweep_us(us))) break;
        if (q == 0) base_us = us;
        if (us < best_us) {
            best_us = us;
            best = shapes[q];
        }
    }
    (void)hipEventDestroy(ea);
    (void)hipEventDestroy(eb);
    if (rc) {
        s->p = p0;
        (void)snap_restore(s, S0);
        return rc;
    }
    if (!(best_us < 0.98 * base_us)) {  // a win within the timing noise keeps the current shape
        best = shapes[0];
        best_us = base_us;
    }
    s->p.tile_gather = best.tg;
    s->p.block_tet = best.bt;
    s->p.block_vert = best.bv;
    s->p.vert_lanes = best.vl;
    s->p.fuse_sweeps = best.fz;
    // 2. Steps per second over a window of trial_steps steps that starts
    // trial_steps steps after the saved state (past the first steps, whose
    // residual histories differ most from the run's), each candidate from
    // the same state, best of two runs; on V-cycle schedules the sweep count
    // (3, 4, 5, 6 pre- and post-sweeps) is chosen there, kept only if its
    // steps all met the target and it beat the current count by 3 %.
    if ((rc = snap_restore(s, S0)) || (rc = invalidate_graph(s))) return rc;
    if ((rc = mgsb_step(s, trial_steps, nullptr))) return rc;
    StateSnap S1;
    if ((rc = snap_take(s, S1))) return rc;
    auto measure = [&](double& ms, bool& ok) -> int {
        ms = 1e300;
        ok = true;
        for (int rep = 0; rep < 2; rep++) {
            int r;
            if ((r = snap_restore(s, S1)) || (r = invalidate_graph(s))) return r;
            if ((r = mgsb_step(s, 1, nullptr))) return r;  // graph capture, warm caches
            HIPCHK(hipStreamSynchronize(s->stream));
            VState a{}, b{};
            if (s->vstate) HIPCHK(hipMemcpy(&a, s->vstate, sizeof(VState), hipMemcpyDeviceToHost));
            const auto t0 = std::chrono::steady_clock::now();
            if ((r = mgsb_step(s, trial_steps, nullptr))) return r;
            HIPCHK(hipStreamSynchronize(s->stream));
            const auto t1 = std::chrono::steady_clock::now();
            if (s->vstate) HIPCHK(hipMemcpy(&b, s->vstate, sizeof(VState), hipMemcpyDeviceToHost));
            ms = std::min(ms, std::chrono::duration<double, std::milli>(t1 - t0).count() / trial_steps);
            ok = ok && b.nfail == a.nfail;
        }
        trials++;
        return MGSB_OK;
    };
    const mgsb_params shaped = s->p;
    double ms_shaped = 0.0, ms = 0.0, ms_default = 0.0;
    bool ok = false;
    s->p = p0;
    if ((rc = measure(ms_default, ok))) return rc;
    s->p = shaped;
    if ((rc = measure(ms_shaped, ok))) return rc;
    double best_ms = ms_shaped;
    mgsb_params bestp = shaped;
    if (vcycle)
        for (int sw : {3, 4, 5, 6}) {
            if (sw == shaped.pre_sweeps && sw == shaped.post_sweeps) continue;
            s->p = shaped;
            s->p.pre_sweeps = s->p.post_sweeps = sw;
            if ((rc = measure(ms, ok))) return rc;
            if (ok && ms < 0.97 * best_ms) {
                best_ms = ms;
                bestp = s->p;
            }
        }
    if (best_ms > ms_default) {  // the window disagrees with the kernel timing: keep what was there
        best_ms = ms_default;
        bestp = p0;
    }
    s->p = bestp;
    // the resolved values, so that the parameters read back are the ones in use
    s->p.block_tet = block_tet(s);
    s->p.block_vert = block_vert(s);
    s->p.vert_lanes = 1 << vert_lsh(s);
    if ((rc = snap_restore(s, S0)) || (rc = invalidate_graph(s))) return rc;
    out->block_tet = block_tet(s);
    out->block_vert = block_vert(s);
    out->pre_sweeps = s->p.pre_sweeps;
    out->post_sweeps = s->p.post_sweeps;
    out->ms_per_step = best_ms;
    out->ms_per_step_default = ms_default;
    out->trials = trials;
    out->vert_lanes = 1 << vert_lsh(s);
    out->tile_gather = s->p.tile_gather;
    out->sweep_us = best_us;
    out->sweep_us_default = base_us;
    out->fuse_sweeps = fused(s, s->L[0]) ? 1 : 0;
    return MGSB_OK;
}

int mgsb_body_count(mgsb_solver* s, int32_t* nbodies, int32_t* per_body_checks) {
    if (!s || !nbodies) {
        set_error("mgsb_body_count: bad arguments");
        return MGSB_ERR_ARG;
    }
    *nbodies = s->nbodies;
    if (per_body_checks) *per_body_checks = s->body_checks ? 1 : 0;
    return MGSB_OK;
}

int mgsb_set_body_stiffness(mgsb_solver* s, const float* E, int32_t n) {
    if (!s || !E || n != s->nbodies) {
        set_error("mgsb_set_body_stiffness: need one stiffness per body (mgsb_body_count)");
        return MGSB_ERR_ARG;
    }
    for (int b = 0; b < n; b++)
        if (!(E[b] > 0.f)) {
            set_error("mgsb_set_body_stiffness: stiffness must be > 0");
            return MGSB_ERR_ARG;
        }
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    int rc;
    for (size_t l = 0; l < s->L.size(); l++) {
        DevLevel& d = s->L[l];
        HostLevel& h = d.h;
        h.tk.resize(h.T);
        for (int t = 0; t < h.T; t++) h.tk[t] = E[s->body_of[l][h.tets[(size_t)t * 4]]];
        recompute_diag(h, s->p.stiffness);
        if (!d.tk && (rc = up<float>(s, d.owned, &d.tk, nullptr, (size_t)h.T))) return rc;
        HIPCHK(hipMemcpyAsync(d.tk, h.tk.data(), sizeof(float) * h.T, hipMemcpyHostToDevice, s->stream));
        if ((rc = upload_vc(s, d))) return rc;
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    if (s->L.size() >= 2 && (rc = setup_vcycle(s))) return rc;
    return invalidate_graph(s);
}

int mgsb_body_residuals(mgsb_solver* s, double* res0, double* res, int32_t n) {
    if (!s || n != s->nbodies || !res0 || !res) {
        set_error("mgsb_body_residuals: need arrays of mgsb_body_count entries");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    if (!s->body_checks) {
        if (n != 1 || !s->vstate) {
            set_error("mgsb_body_residuals: per-body residuals need contiguous bodies and a V-cycle schedule");
            return MGSB_ERR_STATE;
        }
        VState v;
        HIPCHK(hipMemcpyAsync(&v, s->vstate, sizeof(VState), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        res0[0] = v.res0;
        res[0] = v.res;
        return MGSB_OK;
    }
    std::vector<double> sq(n);
    HIPCHK(hipMemcpyAsync(res0, s->res0b, sizeof(double) * n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(sq.data(), s->bodysq, sizeof(double) * n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (int b = 0; b < n; b++) res[b] = std::sqrt(sq[b]);
    return MGSB_OK;
}

int mgsb_time_kernel(mgsb_solver* s, int32_t kernel, int32_t reps, double* avg_us) {
    if (!s || kernel < 0 || kernel > 1 || reps < 1 || !avg_us) {
        set_error("mgsb_time_kernel: kernel must be 0 or 1, reps >= 1, avg_us non-null");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    const bool split = s->p.schedule == MGSB_SCHED_VCYCLE || s->p.schedule == MGSB_SCHED_LEARNED;
    DevLevel& d = s->L[0];
    if (kernel == 1 && !(split && s->L.size() >= 2 && fused(s, d))) {
        set_error("mgsb_time_kernel: kernel 1 (fused sweep) needs a V-cycle schedule on a level that fuses its sweeps");
        return MGSB_ERR_STATE;
    }
    const bool was = s->timing;
    s->timing = false;
    // kernel 1 reads x, x_prev and pcf and writes only xz and pcf2 (free between steps)
    auto one = [&]() -> int {
        if (kernel == 0) return launch_tet(s, d, tet_mode_of(s->p.if_mode), false, nullptr, nullptr, split);
        FuseBufs fb{d.x, d.xprev, 0};
        return launch_fused<kFuseJacobi>(s, d, fb, 1.0f, kVertPoint | kVertContactJacobi, nullptr);
    };
    HIPCHK(hipStreamSynchronize(s->stream));
    int rc = one();  // warm-up
    if (rc) return rc;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, s->stream));
    for (int r = 0; r < reps && !rc; r++) rc = one();
    hipError_t e = hipEventRecord(b, s->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    s->timing = was;
    if (rc) return rc;
    HIPCHK(e);
    *avg_us = 1e3 * (double)ms / reps;
    return MGSB_OK;
}

int mgsb_tet_forces(mgsb_solver* s, int32_t level, int32_t if_mode, const float* x, float* cf) {
    if (!s || level < 0 || level >= (int)s->L.size() || if_mode < 0 || if_mode > MGSB_IF_LINEAR || !x || !cf) {
        set_error("mgsb_tet_forces: bad level, if_mode (0..3) or null array");
        return MGSB_ERR_ARG;
    }
    HIPCHK(hipSetDevice(s->device));
    DevLevel& d = s->L[level];
    float4* xd = nullptr;
    int rc = dalloc(&xd, (size_t)d.V);
    if (rc) return rc;
    const int mode = if_mode == MGSB_IF_LINEAR ? kTetModeLinear : tet_mode_of(if_mode);
    const bool was = s->timing;
    s->timing = false;
    rc = write_f4(s, xd, d.V, x);
    if (!rc) rc = launch_tet(s, d, mode, false, xd, nullptr, false, true);
    s->timing = was;
    std::vector<float3> h((size_t)d.T * 4);
    hipError_t e = hipSuccess;
    if (!rc) e = hipMemcpyAsync(h.data(), d.cf, sizeof(float3) * h.size(), hipMemcpyDeviceToHost, s->stream);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(xd);
    if (rc) return rc;
    HIPCHK(e);
    for (int t = 0; t < d.T; t++)
        for (int j = 0; j < 4; j++) {
            const float3 q = h[(size_t)d.h.slot[(size_t)t * 4 + j]];
            float* o = cf + (size_t)t * 12 + j * 3;
            o[0] = q.x;
            o[1] = q.y;
            o[2] = q.z;
        }
    return MGSB_OK;
}

}  // extern "C"
