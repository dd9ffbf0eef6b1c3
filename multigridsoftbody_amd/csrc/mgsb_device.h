// Device-side math of libmgsb (gfx950).  Every function follows the
// reference's FP32 operation order; the library is compiled with
// -ffp-contract=off so no multiply-add is fused behind the source's back.
#pragma once

#include <hip/hip_runtime.h>

namespace mgsb {

#define MGSB_MAX_SPH 8

// Colliders live in device memory so a captured step graph sees moves made
// between replays (sphere centres advance on device each step).
struct DevColliders {
    int nsph;
    int plane_on;
    float kcol;
    float kplane;
    float4 sph[MGSB_MAX_SPH];  // cx cy cz r
    float4 vel[MGSB_MAX_SPH];  // vx vy vz active
    float4 plane;              // nx ny nz d
};

// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx % 8).  This
// remaps them so each XCD walks one contiguous chunk of the index space and
// neighbouring tets/vertices share that XCD's L2.
__device__ __forceinline__ int xcd_block(int bid, int nb) {
    const int x = bid & 7, local = bid >> 3;
    const int per = nb >> 3, rem = nb & 7;
    return x * per + (x < rem ? x : rem) + local;
}

// A kernel's device-side flags (skip, stall guard): a null flag reads the
// zero word below.  The kernels read their flags with their first loads and
// test them after issuing those loads.  The address is a select of two
// uniform global addresses, so the read is one scalar load in the kernel's
// first batch, and its test waits for scalar loads only (a test of the
// pointer ahead of the load costs two more scalar round trips).
__device__ int kFlagZero = 0;  // never written
__device__ __forceinline__ int load_flag(const int* __restrict__ f) { return *(f ? f : &kFlagZero); }

__device__ __forceinline__ float3 f3(float4 a) { return make_float3(a.x, a.y, a.z); }
__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 add3(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 mul3(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 div3(float3 a, float s) { return make_float3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float len3(float3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
// The reference compares floats against the double literal 1e-6; for a float
// x that is exactly x <= 1e-6f (no float lies in (1e-6f, 1e-6]).
__device__ __forceinline__ bool lt_1e6(float x) { return x <= 1e-6f; }

// Error-free transformations in FP32: a + b = s + e exactly (Knuth's
// TwoSum, TAOCP vol. 2, 4.2.2) and a * b = p + e exactly (Dekker's
// TwoProduct, "A floating-point technique for extending the available
// precision", 1971, with the product's error taken by one fused
// multiply-add).
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void two_prod(float a, float b, float& p, float& e) {
    p = a * b;
    e = __builtin_fmaf(a, b, -p);
}

// l = I_c (I_c^2 - 4.5 II_c) + 13.5 III_c as GetRotation_D evaluates it
// (utils.cu:161): I_c^2 is a float product, the double literals 4.5 and 13.5
// carry everything else into double, and the double result is rounded to
// float.  l is the cancelling numerator of cos(3 phi) (three O(27) terms,
// result O(k^1.5)), so an FP32 evaluation loses it at small strains (at
// strain 1e-3 its error is hundreds of times k^1.5).  Here every term is
// kept exactly as a sum of floats: 4.5 II_c and 13.5 III_c as TwoProducts,
// I_c^2 - 4.5 II_c as dh + dl (exact: the difference needs about 30 bits),
// I_c (dh + dl) as four TwoProduct parts.  After the leading parts cancel
// (TwoSum, exact), the small parts are summed with their rounding errors
// carried (cascaded TwoSum, Ogita-Rump-Oishi "Accurate sum and dot
// product", 2005), and the result is rounded once.  Over 1.4M random F
// (strains 1e-4 to 20, half of them inverted) this equals the
// double-evaluated l rounded to float in all but one case (one ulp).
__device__ __forceinline__ float franca_l_double(float I_c, float I_c2, float II_c, float III_c) {
    float ph, pl;
    two_prod(4.5f, II_c, ph, pl);
    float dh, e1;
    two_sum(I_c2, -ph, dh, e1);
    const float dl = e1 - pl;
    float mh, m1, m2h, m2l;
    two_prod(I_c, dh, mh, m1);
    two_prod(I_c, dl, m2h, m2l);
    float qh, ql;
    two_prod(13.5f, III_c, qh, ql);
    float sh, e2;
    two_sum(mh, qh, sh, e2);
    float t1, c1, t2, c2, t3, c3;
    two_sum(e2, m1, t1, c1);
    two_sum(t1, m2h, t2, c2);
    two_sum(t2, ql, t3, c3);
    float th, tl;
    two_sum(t3, ((c1 + c2) + c3) + m2l, th, tl);
    float r, cr;
    two_sum(sh, th, r, cr);
    return r + (cr + tl);
}

// Franca closed-form polar decomposition R = F U^-1 from the invariants of
// C = F^T F.  REF_D selects the variant:
//   true : GetRotation_D (utils.cu:123-226), the local step of calculateIF
//          and of runCalEnergy: l to double accuracy (franca_l_double) and
//          III_u = +-sqrt(III_c) with the sign of det F (utils.cu:171-172);
//   false: getTetFRV2 (PDSolverAc.cu:494-595): l in FP32 with float
//          constants and III_u = det F.
// (In FP32, sqrtf(det * det) == |det| barring underflow, so the two III_u
// agree wherever the guards below let the formula run.)  Fallbacks return
// R = I where the reference does.  Writes R - F into RF.
template <bool REF_D>
__device__ __forceinline__ void polar_rf(const float F[9], float RF[9]) {
    float C[9], C2[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) C[i * 3 + j] = F[0 * 3 + i] * F[0 * 3 + j] + F[1 * 3 + i] * F[1 * 3 + j] + F[2 * 3 + i] * F[2 * 3 + j];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            C2[i * 3 + j] = C[i * 3 + 0] * C[j * 3 + 0] + C[i * 3 + 1] * C[j * 3 + 1] + C[i * 3 + 2] * C[j * 3 + 2];
    const float det = F[0] * F[4] * F[8] + F[1] * F[5] * F[6] + F[3] * F[7] * F[2] - F[2] * F[4] * F[6] -
                      F[1] * F[3] * F[8] - F[0] * F[5] * F[7];
    const float I_c = C[0] + C[4] + C[8];
    const float I_c2 = I_c * I_c;
    const float II_c = 0.5f * (I_c2 - C2[0] - C2[4] - C2[8]);
    const float III_c = det * det;
    const float k = I_c2 - 3 * II_c;
    float iu[9];
    bool ok = true;
    if (k < 1e-6f) {
        if (lt_1e6(I_c)) {
            ok = false;
        } else {
            const float inv_lambda = 1 / sqrtf(I_c / 3);
#pragma unroll
            for (int i = 0; i < 9; i++) iu[i] = 0.f;
            iu[0] = iu[4] = iu[8] = inv_lambda;
        }
    } else {
        const float l = REF_D ? franca_l_double(I_c, I_c2, II_c, III_c) : I_c * (I_c * I_c - 4.5f * II_c) + 13.5f * III_c;
        const float k_root = sqrtf(k);
        float value = l / (k * k_root);
        value = value < -1.0f ? -1.0f : value;
        value = value > 1.0f ? 1.0f : value;
        const float phi = acosf(value);
        const float lambda2 = (I_c + 2 * k_root * cosf(phi / 3)) / 3;
        const float lambda = sqrtf(lambda2);
        const float III_u = REF_D ? copysignf(sqrtf(III_c), det < 0.f ? -1.f : 1.f) : det;
        float I_u = 0.f, II_u = 0.f;
        if (lt_1e6(fabsf(lambda))) ok = false;
        if (ok && lt_1e6(-lambda2 + I_c + 2 * III_u / lambda)) ok = false;
        if (ok) {
            I_u = lambda + sqrtf(-lambda2 + I_c + 2 * III_u / lambda);
            II_u = (I_u * I_u - I_c) * 0.5f;
            if (lt_1e6(fabsf(I_u * II_u - III_u))) ok = false;
        }
        if (ok && lt_1e6(fabsf(III_u))) ok = false;
        if (ok) {
            float inv_rate = 1 / (I_u * II_u - III_u);
            const float fdiag = I_u * III_u * inv_rate;
            const float fc = (I_u * I_u - II_u) * inv_rate;
            float U[9];
#pragma unroll
            for (int i = 0; i < 9; i++) U[i] = fc * C[i] - inv_rate * C2[i];
            U[0] += fdiag;
            U[4] += fdiag;
            U[8] += fdiag;
            inv_rate = 1 / III_u;
            const float gdiag = II_u * inv_rate;
            const float gu = -I_u * inv_rate;
#pragma unroll
            for (int i = 0; i < 9; i++) iu[i] = gu * U[i] + inv_rate * C[i];
            iu[0] += gdiag;
            iu[4] += gdiag;
            iu[8] += gdiag;
        }
    }
    if (ok) {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                RF[i * 3 + j] = (F[i * 3 + 0] * iu[0 * 3 + j] + F[i * 3 + 1] * iu[1 * 3 + j] + F[i * 3 + 2] * iu[2 * 3 + j]) -
                                F[i * 3 + j];
    } else {
#pragma unroll
        for (int i = 0; i < 9; i++) RF[i] = ((i % 4) == 0 ? 1.f : 0.f) - F[i];
    }
}

// The same Franca formulas evaluated from deviatoric invariants: with
// C' = C - I_c/3 I, k = I_c^2 - 3 II_c = 1.5 tr(C'^2) and l = 13.5 det C'.
// Algebraically identical to polar_rf, but k no longer comes from the
// difference of two O(10) numbers, so small anisotropic strains survive in
// FP32 (the reference form drops them below k < 1e-6).  Isotropic branch only
// for k < 1e-30.  Writes R - F into RF.  This form has no counterpart in the
// reference (it serves the V-cycle schedules, whose residual iteration the
// reference form stalls), so it is free of the reference's operation order:
// its multiply-adds fuse, most divisions and square roots are the
// hardware's reciprocal, reciprocal square root and square root (about 1
// ulp each), and cos(phi / 3), phi / 3 in [0, pi / 3], is the hardware
// cosine: the local step is bound by instruction issue on large meshes.
// lambda and I_u stay IEEE (below).  C and C^2 are formed from their six
// distinct entries, U^-1 = gu U + q C + gdiag I with U = fc C - r C^2 +
// fdiag I is applied as one combination a C + b C^2 + g I, and R - F is
// F (U^-1 - I): the same formulas in fewer instructions.
__device__ __forceinline__ float rcp_hw(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrt_hw(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ void polar_rf_stable(const float F[9], float RF[9]) {
#pragma clang fp contract(fast)
    // C = F^T F is symmetric: its six distinct entries
    const float c00 = F[0] * F[0] + F[3] * F[3] + F[6] * F[6];
    const float c01 = F[0] * F[1] + F[3] * F[4] + F[6] * F[7];
    const float c02 = F[0] * F[2] + F[3] * F[5] + F[6] * F[8];
    const float c11 = F[1] * F[1] + F[4] * F[4] + F[7] * F[7];
    const float c12 = F[1] * F[2] + F[4] * F[5] + F[7] * F[8];
    const float c22 = F[2] * F[2] + F[5] * F[5] + F[8] * F[8];
    // det F by cofactors of the first row
    const float det = F[0] * (F[4] * F[8] - F[5] * F[7]) + F[1] * (F[5] * F[6] - F[3] * F[8]) +
                      F[2] * (F[3] * F[7] - F[4] * F[6]);
    const float I_c = c00 + c11 + c22;
    const float m = I_c * (1.0f / 3);
    const float D0 = c00 - m, D4 = c11 - m, D8 = c22 - m;
    const float k = 1.5f * (D0 * D0 + D4 * D4 + D8 * D8 + 2 * (c01 * c01 + c02 * c02 + c12 * c12));
    // U^-1 = a C + b C^2 + g I (symmetric); RF = F (U^-1 - I) = R - F
    float a = 0.f, b = 0.f, g = 0.f;
    bool ok = true;
    if (k < 1e-30f) {
        if (lt_1e6(I_c))
            ok = false;
        else
            g = __builtin_amdgcn_rsqf(m);
    } else {
        const float l = 13.5f * (D0 * (D4 * D8 - c12 * c12) - c01 * (c01 * D8 - c12 * c02) +
                                 c02 * (c01 * c12 - D4 * c02));
        const float k_root = sqrt_hw(k);
        float value = l * rcp_hw(k * k_root);
        value = value < -1.0f ? -1.0f : value;
        value = value > 1.0f ? 1.0f : value;
        const float phi = acosf(value);
        const float lambda2 = m + 2.0f / 3 * k_root * __cosf(phi * (1.0f / 3));  // phi / 3 in [0, pi/3]
        // IEEE from here to I_u: on an inverted tet near a reflection the
        // difference under the root cancels and 1 / (I_u II_u - III_u) below
        // amplifies every error in lambda and in that root
        const float lambda = sqrtf(lambda2);
        const float III_u = det;
        const float t2 = 2 * III_u / lambda;
        float I_u = 0.f, II_u = 0.f;
        if (lt_1e6(fabsf(lambda))) ok = false;
        if (ok && lt_1e6(-lambda2 + I_c + t2)) ok = false;
        if (ok) {
            I_u = lambda + sqrtf(-lambda2 + I_c + t2);
            II_u = (I_u * I_u - I_c) * 0.5f;
            if (lt_1e6(fabsf(I_u * II_u - III_u))) ok = false;
        }
        if (ok && lt_1e6(fabsf(III_u))) ok = false;
        if (ok) {
            // U = fc C - r C^2 + fdiag I and U^-1 = gu U + q C + gdiag I,
            // folded into one combination of C, C^2 and I
            const float r = rcp_hw(I_u * II_u - III_u);
            const float fdiag = I_u * III_u * r;
            const float fc = (I_u * I_u - II_u) * r;
            const float q = rcp_hw(III_u);
            const float gdiag = II_u * q;
            const float gu = -I_u * q;
            a = gu * fc + q;
            b = -(gu * r);
            g = gu * fdiag + gdiag;
        }
    }
    if (!ok) {
#pragma unroll
        for (int i = 0; i < 9; i++) RF[i] = ((i % 4) == 0 ? 1.f : 0.f) - F[i];
        return;
    }
    // C^2 (symmetric) only where it enters
    const float q00 = c00 * c00 + c01 * c01 + c02 * c02;
    const float q01 = c00 * c01 + c01 * c11 + c02 * c12;
    const float q02 = c00 * c02 + c01 * c12 + c02 * c22;
    const float q11 = c01 * c01 + c11 * c11 + c12 * c12;
    const float q12 = c01 * c02 + c11 * c12 + c12 * c22;
    const float q22 = c02 * c02 + c12 * c12 + c22 * c22;
    const float gm = g - 1.0f;
    const float M[9] = {a * c00 + b * q00 + gm, a * c01 + b * q01,      a * c02 + b * q02,
                        a * c01 + b * q01,      a * c11 + b * q11 + gm, a * c12 + b * q12,
                        a * c02 + b * q02,      a * c12 + b * q12,      a * c22 + b * q22 + gm};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            RF[i * 3 + j] = F[i * 3 + 0] * M[0 * 3 + j] + F[i * 3 + 1] * M[1 * 3 + j] + F[i * 3 + 2] * M[2 * 3 + j];
}

// barycentricCoordinate on four points (utils.cu:335-364), GetVolumn inline.
__device__ __forceinline__ void barycentric4(float3 P, float3 a, float3 b, float3 c, float3 d, float w[4]) {
    const float3 AB = sub3(b, a), AC = sub3(c, a), AD = sub3(d, a);
    const float M0 = AB.x, M1 = AC.x, M2 = AD.x, M3 = AB.y, M4 = AC.y, M5 = AD.y, M6 = AB.z, M7 = AC.z, M8 = AD.z;
    const float R0 = M4 * M8 - M7 * M5;
    const float R1 = M7 * M2 - M1 * M8;
    const float R2 = M1 * M5 - M4 * M2;
    const float Vol = (M0 * R0 + M3 * R1 + M6 * R2) / 6.0f;
    const float3 center = mul3(add3(add3(add3(a, b), c), d), 0.25f);
    const float3 tv[4] = {a, b, c, d};
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float3 q = tv[i], f0 = tv[(i + 1) & 3], f1 = tv[(i + 2) & 3], f2 = tv[(i + 3) & 3];
        float3 n = cross3(sub3(f1, f0), sub3(f2, f0));
        if (dot3(n, sub3(q, f0)) > 0) n = make_float3(-n.x, -n.y, -n.z);
        const float l = len3(n);
        const float area = l * 0.5f;
        n = div3(n, l);
        w[i] = 0.25f - dot3(sub3(P, center), n) * area / (3 * Vol);
    }
}

// pointInTriangle (utils.cu:366-383).
__device__ __forceinline__ bool point_in_triangle(float3 P, float3 A, float3 B, float3 C) {
    const float3 v0 = sub3(C, A), v1 = sub3(B, A), v2 = sub3(P, A);
    const float d00 = dot3(v0, v0), d01 = dot3(v0, v1), d02 = dot3(v0, v2), d11 = dot3(v1, v1), d12 = dot3(v1, v2);
    const float inv = 1.0f / (d00 * d11 - d01 * d01);
    const float u = (d11 * d02 - d01 * d12) * inv;
    const float w = (d00 * d12 - d01 * d02) * inv;
    return (u >= 0) && (w >= 0) && (u + w <= 1.0f);
}

}  // namespace mgsb
