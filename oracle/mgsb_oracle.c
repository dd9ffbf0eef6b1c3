/*
 * mgsb_oracle.c -- CPU restatement of the reference projective-dynamics (PD)
 * tetrahedral solver of buuugyellow/MultiGridSoftBody.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: tests, the smoke check
 * and the CPU-baseline leg of the benchmark call it as the checker.  The
 * product (libmgsb, HIP) never links, loads or calls it.
 *
 * The same source is compiled twice:
 *   - default      : `real` = float, prefix orc32_.  Follows the reference's
 *                    FP32 operation order, including the places where the
 *                    reference promotes to double through a double literal
 *                    (GetRotation_D's `l`, calculatePOS's 0.6 relaxation).
 *   - ORC_DOUBLE   : `real` = double, prefix orc64_.  The FP64 yardstick.
 *
 * Summation order of scattered forces: the reference scatters with atomicAdd
 * (order not defined).  Here every per-vertex sum runs over incident tets in
 * increasing tet index, i.e. the order of a sequential scatter loop over tets.
 * The OpenMP path computes per-tet contributions in parallel, then gathers
 * per vertex in that same order, so 1 thread and N threads agree bitwise.
 *
 * Citations are to files of the reference repository.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef ORC_DOUBLE
typedef float real;
#define PFX(n) orc32_##n
#define R_SQRT(x) sqrtf(x)
#define R_ACOS(x) acosf(x)
#define R_COS(x) cosf(x)
#define R_FABS(x) fabsf(x)
#else
typedef double real;
#define PFX(n) orc64_##n
#define R_SQRT(x) sqrt(x)
#define R_ACOS(x) acos(x)
#define R_COS(x) cos(x)
#define R_FABS(x) fabs(x)
#endif

#define ORC_MAX_SPHERES 8

/* ------------------------------------------------------------------------- */
/* Parameters (always float at the interface, converted to `real` inside).    */
/* ------------------------------------------------------------------------- */
typedef struct orc_params {
    float dt, damping, stiffness, density, rho;
    double relax;         /* 0.6 in the reference, a double literal          */
    float gravity[3];
    float kcol;           /* collision stiffness (PDSolver.cpp:171)            */
    int iters;            /* PD iterations per step (fine / single level)      */
    int iters_coarse;     /* coarse iterations in the two-level schedule       */
    int if_mode;          /* 0: calculateIF (PDSolver.cu:128); 1: Ac path      */
                          /*    getTetFRV2 + calTetIFV1 (PDSolverAc.cu:447,630)*/
                          /* 3: calculateIF with the deviatoric-invariant polar */
    int tri_dcd_coarse;   /* two-level schedule: triangle DCD on coarse level  */
    int nspheres;
    float sphere[ORC_MAX_SPHERES][8]; /* cx cy cz r active vx vy vz           */
    float plane[5];       /* nx ny nz d k (k>0 = active); not in the reference */
} orc_params;

/* ------------------------------------------------------------------------- */
/* Small vector helpers, same association as utils.cu:277-296.                */
/* ------------------------------------------------------------------------- */
typedef struct { real x, y, z; } v3;
static inline v3 v3m(real x, real y, real z) { v3 r = {x, y, z}; return r; }
static inline v3 vsub(v3 a, v3 b) { return v3m(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vadd(v3 a, v3 b) { return v3m(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vmul(v3 a, real s) { return v3m(a.x * s, a.y * s, a.z * s); }
static inline v3 vdiv(v3 a, real s) { return v3m(a.x / s, a.y / s, a.z / s); }
static inline real vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) {
    return v3m(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline real vlen(v3 a) { return R_SQRT(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline v3 vnormalize(v3 a) { return vdiv(a, vlen(a)); }
static inline v3 vload(const real* p) { return v3m(p[0], p[1], p[2]); }

/* ------------------------------------------------------------------------- */
/* Matrix_Inverse_3 (utils.cu:233-252): adjugate / det, returns det.          */
/* The reference exits on |det| < 1e-6; here the caller gets the det back.    */
/* ------------------------------------------------------------------------- */
static real mat_inverse3(const real* A, real* R) {
    R[0] = A[4] * A[8] - A[7] * A[5];
    R[1] = A[7] * A[2] - A[1] * A[8];
    R[2] = A[1] * A[5] - A[4] * A[2];
    R[3] = A[5] * A[6] - A[3] * A[8];
    R[4] = A[0] * A[8] - A[2] * A[6];
    R[5] = A[2] * A[3] - A[0] * A[5];
    R[6] = A[3] * A[7] - A[4] * A[6];
    R[7] = A[1] * A[6] - A[0] * A[7];
    R[8] = A[0] * A[4] - A[1] * A[3];
    real det = A[0] * R[0] + A[3] * R[1] + A[6] * R[2];
    if (R_FABS(det) < 1e-6) return det;
    real inv_det = 1 / det;
    for (int i = 0; i < 9; i++) R[i] *= inv_det;
    return det;
}

/* ------------------------------------------------------------------------- */
/* Franca closed-form polar decomposition R = F U^{-1}.                       */
/* variant 0: GetRotation_D (utils.cu:123-226), `l` evaluated in double as    */
/*            the reference's double literals make it.                        */
/* variant 1: getTetFRV2 (PDSolverAc.cu:494-595), all-float constants.        */
/* Fallbacks return R = I exactly where the reference does.  Returns det F.   */
/* ------------------------------------------------------------------------- */
static real polar(const real F[9], real R[9], int variant) {
    for (int i = 0; i < 9; i++) R[i] = 0;
    R[0] = R[4] = R[8] = 1;
    real C[9], C2[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            real acc = 0;
            for (int k = 0; k < 3; k++) acc += F[k * 3 + i] * F[k * 3 + j];
            C[i * 3 + j] = acc;
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            real acc = 0;
            for (int k = 0; k < 3; k++) acc += C[i * 3 + k] * C[j * 3 + k];
            C2[i * 3 + j] = acc;
        }
    real det = F[0] * F[4] * F[8] + F[1] * F[5] * F[6] + F[3] * F[7] * F[2] - F[2] * F[4] * F[6] -
               F[1] * F[3] * F[8] - F[0] * F[5] * F[7];
    real I_c = C[0] + C[4] + C[8];
    real I_c2 = I_c * I_c;
    real II_c = (real)0.5 * (I_c2 - C2[0] - C2[4] - C2[8]);
    real III_c = det * det;
    real k = I_c2 - 3 * II_c;
    real inv_U[9];
    if (variant == 2) {
        /* Same Franca formulas from deviatoric invariants: with C' = C - I_c/3 I,
         * k = I_c^2 - 3 II_c = 1.5 tr(C'^2) and l = 13.5 det C', which keeps
         * the anisotropic part of small strains that the O(1) difference
         * loses in FP32; the isotropic branch is kept for k == 0 only. */
        real m = I_c / 3;
        real D0 = C[0] - m, D4 = C[4] - m, D8 = C[8] - m;
        k = (real)1.5 * (D0 * D0 + D4 * D4 + D8 * D8 + 2 * (C[1] * C[1] + C[2] * C[2] + C[5] * C[5]));
        if (k < (real)1e-30f) {
            if (I_c < 1e-6) return det;
            real inv_lambda = 1 / R_SQRT(m);
            for (int i = 0; i < 9; i++) inv_U[i] = 0;
            inv_U[0] = inv_U[4] = inv_U[8] = inv_lambda;
        } else {
            real l = (real)13.5 * (D0 * (D4 * D8 - C[5] * C[5]) - C[1] * (C[1] * D8 - C[5] * C[2]) +
                                   C[2] * (C[1] * C[5] - D4 * C[2]));
            real k_root = R_SQRT(k);
            real value = l / (k * k_root);
            if (value < -1.0) value = -1.0;
            if (value > 1.0) value = 1.0;
            real phi = R_ACOS(value);
            real lambda2 = m + (real)2 / 3 * k_root * R_COS(phi / 3);
            real lambda = R_SQRT(lambda2);
            real III_u = det;
            if (R_FABS(lambda) < 1e-6) return det;
            if (-lambda2 + I_c + 2 * III_u / lambda < 1e-6) return det;
            real I_u = lambda + R_SQRT(-lambda2 + I_c + 2 * III_u / lambda);
            real II_u = (I_u * I_u - I_c) * (real)0.5;
            if (R_FABS(I_u * II_u - III_u) < 1e-6) return det;
            if (R_FABS(III_u) < 1e-6) return det;
            real inv_rate = 1 / (I_u * II_u - III_u);
            real fdiag = I_u * III_u * inv_rate;
            real fc = (I_u * I_u - II_u) * inv_rate;
            real U[9];
            for (int i = 0; i < 9; i++) U[i] = fc * C[i] - inv_rate * C2[i];
            U[0] += fdiag; U[4] += fdiag; U[8] += fdiag;
            inv_rate = 1 / III_u;
            real gdiag = II_u * inv_rate, gu = -I_u * inv_rate;
            for (int i = 0; i < 9; i++) inv_U[i] = gu * U[i] + inv_rate * C[i];
            inv_U[0] += gdiag; inv_U[4] += gdiag; inv_U[8] += gdiag;
        }
    } else if (k < (real)1e-6f) {
        if (I_c < 1e-6) return det;
        real inv_lambda = 1 / R_SQRT(I_c / 3);
        for (int i = 0; i < 9; i++) inv_U[i] = 0;
        inv_U[0] = inv_U[4] = inv_U[8] = inv_lambda;
    } else {
        real l;
        if (variant == 0)
            l = (real)((double)I_c * ((double)(I_c * I_c) - 4.5 * (double)II_c) + 13.5 * (double)III_c);
        else
            l = I_c * (I_c * I_c - (real)4.5f * II_c) + (real)13.5f * III_c;
        real k_root = R_SQRT(k);
        real value = l / (k * k_root);
        if (value < -1.0) value = -1.0;
        if (value > 1.0) value = 1.0;
        real phi = R_ACOS(value);
        real lambda2 = (I_c + 2 * k_root * R_COS(phi / 3)) / 3;
        real lambda = R_SQRT(lambda2);
        real III_u;
        if (variant == 0) {
            III_u = R_SQRT(III_c);
            if (det < 0) III_u = -III_u;
        } else {
            III_u = det;
        }
        if (R_FABS(lambda) < 1e-6) return det;
        if (-lambda2 + I_c + 2 * III_u / lambda < 1e-6) return det;
        real I_u = lambda + R_SQRT(-lambda2 + I_c + 2 * III_u / lambda);
        real II_u = (I_u * I_u - I_c) * (real)0.5;
        if (R_FABS(I_u * II_u - III_u) < 1e-6) return det;
        real inv_rate = 1 / (I_u * II_u - III_u);
        real factor = I_u * III_u * inv_rate;
        real U[9];
        for (int i = 0; i < 9; i++) U[i] = 0;
        U[0] = U[4] = U[8] = factor;
        factor = (I_u * I_u - II_u) * inv_rate;
        for (int i = 0; i < 9; i++) U[i] += factor * C[i] - inv_rate * C2[i];
        if (R_FABS(III_u) < 1e-6) return det;
        inv_rate = 1 / III_u;
        factor = II_u * inv_rate;
        for (int i = 0; i < 9; i++) inv_U[i] = 0;
        inv_U[0] = inv_U[4] = inv_U[8] = factor;
        factor = -I_u * inv_rate;
        for (int i = 0; i < 9; i++) inv_U[i] += factor * U[i] + inv_rate * C[i];
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            real acc = 0;
            for (int k2 = 0; k2 < 3; k2++) acc += F[i * 3 + k2] * inv_U[k2 * 3 + j];
            R[i * 3 + j] = acc;
        }
    return det;
}

/* ------------------------------------------------------------------------- */
/* Level: one tet mesh with the reference's per-level state (PDSolver.h:8-68). */
/* ------------------------------------------------------------------------- */
typedef struct orc_level {
    int V, T, S;
    int* tets;
    unsigned* tris;
    real *x0, *x, *v, *fext, *xlast, *s, *xprev;
    real *B, *AC, *vol, *mass, *diag, *fix, *center, *faceN, *faceA;
    real *force, *cdiag, *cforce, *cenergy, *cdepth;
    real* coff;      /* off-diagonals (xy, xz, yz) of the contacts' k n n^T  */
    int* cflag;
    real* cf;        /* per tet 12: corner j component i at [t*12 + j*3 + i] */
    int *rp, *ri;    /* vertex -> incident (t*4+c), t ascending               */
    int *trp, *tri;  /* vertex -> incident (tri*3+c), tri ascending           */
    real *tc;        /* per tri corner: 3 force + 1 diag                      */
    real *trin;      /* tri normals                                           */
    real *vnormal;   /* vertex normals                                        */
    float stiffness, density;
    real* tk;        /* per-tet stiffness (bodies of their own E) or NULL     */
    int bad_tet;     /* index of first degenerate tet or -1                   */
} orc_level;

/* ARAP weight per unit volume of tet t. */
static inline real tet_k(const orc_level* L, int t) { return L->tk ? L->tk[t] : (real)L->stiffness; }

static void* xcalloc(size_t n, size_t s) { return calloc(n ? n : 1, s); }

static void build_csr(int V, int n_items, int corners, const int* items, int** rp_out, int** ri_out) {
    int* rp = (int*)xcalloc((size_t)V + 1, sizeof(int));
    for (int e = 0; e < n_items * corners; e++) rp[items[e] + 1]++;
    for (int i = 0; i < V; i++) rp[i + 1] += rp[i];
    int* fill = (int*)xcalloc((size_t)V, sizeof(int));
    int* ri = (int*)xcalloc((size_t)n_items * corners, sizeof(int));
    for (int e = 0; e < n_items * corners; e++) {  /* ascending item order */
        int v = items[e];
        ri[rp[v] + fill[v]++] = e;
    }
    free(fill);
    *rp_out = rp;
    *ri_out = ri;
}

/* InitVolumeConstraint (PDSolver.cpp:12-132). */
static void precompute(orc_level* L) {
    int T = L->T;
    real k = (real)L->stiffness;
    L->bad_tet = -1;
    for (int i = 0; i < T; i++) {
        int v0 = L->tets[i * 4 + 0], v1 = L->tets[i * 4 + 1], v2 = L->tets[i * 4 + 2], v3i = L->tets[i * 4 + 3];
        const real* X = L->x0;
        real D[9];
        D[0] = X[v1 * 3 + 0] - X[v0 * 3 + 0];
        D[1] = X[v2 * 3 + 0] - X[v0 * 3 + 0];
        D[2] = X[v3i * 3 + 0] - X[v0 * 3 + 0];
        D[3] = X[v1 * 3 + 1] - X[v0 * 3 + 1];
        D[4] = X[v2 * 3 + 1] - X[v0 * 3 + 1];
        D[5] = X[v3i * 3 + 1] - X[v0 * 3 + 1];
        D[6] = X[v1 * 3 + 2] - X[v0 * 3 + 2];
        D[7] = X[v2 * 3 + 2] - X[v0 * 3 + 2];
        D[8] = X[v3i * 3 + 2] - X[v0 * 3 + 2];
        real det = mat_inverse3(D, &L->B[i * 9]);
        if (R_FABS(det) < 1e-6 && L->bad_tet < 0) L->bad_tet = i;
        L->vol[i] = R_FABS(det) / 6;
        real vq = L->vol[i] / 4;
        L->mass[v0] += vq;
        L->mass[v1] += vq;
        L->mass[v2] += vq;
        L->mass[v3i] += vq;
        const real* b = &L->B[i * 9];
        real* ac = &L->AC[i * 12];
        ac[0] = -b[0] - b[3] - b[6]; ac[1] = b[0]; ac[2] = b[3]; ac[3] = b[6];
        ac[4] = -b[1] - b[4] - b[7]; ac[5] = b[1]; ac[6] = b[4]; ac[7] = b[7];
        ac[8] = -b[2] - b[5] - b[8]; ac[9] = b[2]; ac[10] = b[5]; ac[11] = b[8];
        int vs[4] = {v0, v1, v2, v3i};
        for (int j = 0; j < 4; j++)
            for (int r = 0; r < 3; r++) L->diag[vs[j]] += ac[r * 4 + j] * ac[r * 4 + j] * L->vol[i] * k;
        for (int c = 0; c < 3; c++) {
            real acc = 0;
            for (int j = 0; j < 4; j++) acc += X[vs[j] * 3 + c];
            L->center[i * 3 + c] = acc / 4;
        }
        int faces[4][3] = {{v1, v2, v3i}, {v0, v2, v3i}, {v0, v1, v3i}, {v0, v1, v2}};
        for (int f = 0; f < 4; f++) {
            v3 A = vload(X + faces[f][0] * 3), Bq = vload(X + faces[f][1] * 3), Cq = vload(X + faces[f][2] * 3);
            v3 Dq = vload(X + vs[f] * 3);
            v3 n = vcross(vsub(Bq, A), vsub(Cq, A));
            if (vdot(n, vsub(Dq, A)) > 0) n = v3m(-n.x, -n.y, -n.z);
            real len = vlen(n);
            L->faceA[i * 4 + f] = (real)0.5 * len;
            n = vdiv(n, len);
            L->faceN[i * 12 + f * 3 + 0] = n.x;
            L->faceN[i * 12 + f * 3 + 1] = n.y;
            L->faceN[i * 12 + f * 3 + 2] = n.z;
        }
    }
    real rho = (real)L->density;
    if (rho != 1)
        for (int i = 0; i < L->V; i++) L->mass[i] *= rho;
}

orc_level* PFX(level_create)(int V, int T, int S, const float* pos, const int* tets, const unsigned* tris,
                            float stiffness, float density) {
    orc_level* L = (orc_level*)calloc(1, sizeof(orc_level));
    L->V = V; L->T = T; L->S = S;
    L->stiffness = stiffness; L->density = density;
    L->tets = (int*)xcalloc((size_t)T * 4, sizeof(int));
    memcpy(L->tets, tets, sizeof(int) * 4 * (size_t)T);
    L->tris = (unsigned*)xcalloc((size_t)S * 3, sizeof(unsigned));
    if (S) memcpy(L->tris, tris, sizeof(unsigned) * 3 * (size_t)S);
    size_t v3n = (size_t)V * 3;
    real** vec3s[] = {&L->x0, &L->x, &L->v, &L->fext, &L->xlast, &L->s, &L->xprev, &L->force, &L->cdiag,
                      &L->cforce, &L->vnormal, &L->coff};
    for (size_t q = 0; q < sizeof(vec3s) / sizeof(vec3s[0]); q++) *vec3s[q] = (real*)xcalloc(v3n, sizeof(real));
    L->mass = (real*)xcalloc(V, sizeof(real));
    L->diag = (real*)xcalloc(V, sizeof(real));
    L->fix = (real*)xcalloc(V, sizeof(real));
    L->cenergy = (real*)xcalloc(V, sizeof(real));
    L->cdepth = (real*)xcalloc(V, sizeof(real));
    L->cflag = (int*)xcalloc(V, sizeof(int));
    L->B = (real*)xcalloc((size_t)T * 9, sizeof(real));
    L->AC = (real*)xcalloc((size_t)T * 12, sizeof(real));
    L->vol = (real*)xcalloc(T, sizeof(real));
    L->center = (real*)xcalloc((size_t)T * 3, sizeof(real));
    L->faceN = (real*)xcalloc((size_t)T * 12, sizeof(real));
    L->faceA = (real*)xcalloc((size_t)T * 4, sizeof(real));
    L->cf = (real*)xcalloc((size_t)T * 12, sizeof(real));
    L->tc = (real*)xcalloc((size_t)S * 12, sizeof(real));
    L->trin = (real*)xcalloc((size_t)S * 3, sizeof(real));
    for (size_t i = 0; i < v3n; i++) L->x0[i] = L->x[i] = (real)pos[i];
    for (int i = 0; i < V; i++) L->fix[i] = 1;
    precompute(L);
    build_csr(V, T, 4, L->tets, &L->rp, &L->ri);
    build_csr(V, S, 3, (const int*)L->tris, &L->trp, &L->tri);
    return L;
}

void PFX(level_free)(orc_level* L) {
    if (!L) return;
    void* ps[] = {L->tets, L->tris, L->x0, L->x, L->v, L->fext, L->xlast, L->s, L->xprev, L->B, L->AC, L->vol,
                  L->mass, L->diag, L->fix, L->center, L->faceN, L->faceA, L->force, L->cdiag, L->cforce,
                  L->cenergy, L->cdepth, L->cflag, L->cf, L->rp, L->ri, L->trp, L->tri, L->tc, L->trin,
                  L->vnormal, L->tk, L->coff};
    for (size_t i = 0; i < sizeof(ps) / sizeof(ps[0]); i++) free(ps[i]);
    free(L);
}

/* Per-tet stiffness (several bodies, each with its own E); recomputes the
 * Jacobi diagonal in the precompute's tet order. */
void PFX(level_set_tet_k)(orc_level* L, const float* k) {
    if (!L->tk) L->tk = (real*)xcalloc((size_t)L->T, sizeof(real));
    for (int t = 0; t < L->T; t++) L->tk[t] = (real)k[t];
    for (int i = 0; i < L->V; i++) L->diag[i] = 0;
    for (int t = 0; t < L->T; t++) {
        const real* ac = &L->AC[t * 12];
        for (int j = 0; j < 4; j++)
            for (int r = 0; r < 3; r++)
                L->diag[L->tets[t * 4 + j]] += ac[r * 4 + j] * ac[r * 4 + j] * L->vol[t] * L->tk[t];
    }
}

int PFX(level_bad_tet)(const orc_level* L) { return L->bad_tet; }

void PFX(set_threads)(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* fixflag: 1 = free, 0 = pinned (PDSolver.cpp:134-162). */
void PFX(level_set_fix)(orc_level* L, const float* fix) {
    for (int i = 0; i < L->V; i++) L->fix[i] = (real)fix[i];
}
void PFX(level_set_velocity)(orc_level* L, const float* v) {
    for (int i = 0; i < L->V * 3; i++) L->v[i] = (real)v[i];
}
void PFX(level_set_positions)(orc_level* L, const float* x) {
    for (int i = 0; i < L->V * 3; i++) L->x[i] = (real)x[i];
}
void PFX(level_set_s)(orc_level* L, const float* s) {
    for (int i = 0; i < L->V * 3; i++) L->s[i] = (real)s[i];
}

/* which: 0 x, 1 v, 2 s, 3 xprev, 4 xlast, 5 mass, 6 diag, 7 vol, 8 B(9T), 9 AC(12T),
 *        10 cdiag, 11 cforce, 12 cflag(as real), 13 cdepth, 14 cenergy, 15 vnormal,
 *        16 trin, 17 fix, 18 force, 19 center(3T), 20 faceN(12T), 21 faceA(4T) */
static const real* level_array(const orc_level* L, int which, size_t* n) {
    switch (which) {
        case 0: *n = (size_t)L->V * 3; return L->x;
        case 1: *n = (size_t)L->V * 3; return L->v;
        case 2: *n = (size_t)L->V * 3; return L->s;
        case 3: *n = (size_t)L->V * 3; return L->xprev;
        case 4: *n = (size_t)L->V * 3; return L->xlast;
        case 5: *n = (size_t)L->V; return L->mass;
        case 6: *n = (size_t)L->V; return L->diag;
        case 7: *n = (size_t)L->T; return L->vol;
        case 8: *n = (size_t)L->T * 9; return L->B;
        case 9: *n = (size_t)L->T * 12; return L->AC;
        case 10: *n = (size_t)L->V * 3; return L->cdiag;
        case 11: *n = (size_t)L->V * 3; return L->cforce;
        case 13: *n = (size_t)L->V; return L->cdepth;
        case 14: *n = (size_t)L->V; return L->cenergy;
        case 15: *n = (size_t)L->V * 3; return L->vnormal;
        case 16: *n = (size_t)L->S * 3; return L->trin;
        case 17: *n = (size_t)L->V; return L->fix;
        case 18: *n = (size_t)L->V * 3; return L->force;
        case 19: *n = (size_t)L->T * 3; return L->center;
        case 20: *n = (size_t)L->T * 12; return L->faceN;
        case 21: *n = (size_t)L->T * 4; return L->faceA;
        default: *n = 0; return NULL;
    }
}

int PFX(level_get)(const orc_level* L, int which, double* out) {
    if (which == 12) {
        for (int i = 0; i < L->V; i++) out[i] = L->cflag[i];
        return L->V;
    }
    size_t n;
    const real* a = level_array(L, which, &n);
    if (!a) return -1;
    for (size_t i = 0; i < n; i++) out[i] = (double)a[i];
    return (int)n;
}

/* ------------------------------------------------------------------------- */
/* calculateST (PDSolver.cu:72-116).                                           */
/* ------------------------------------------------------------------------- */
static void calc_st(orc_level* L, const orc_params* p) {
    real dt = p->dt, damp = p->damping;
    real g[3] = {p->gravity[0], p->gravity[1], p->gravity[2]};
#pragma omp parallel for schedule(static)
    for (int i = 0; i < L->V; i++) {
        real fix = L->fix[i];
        for (int c = 0; c < 3; c++) {
            int q = i * 3 + c;
            L->xlast[q] = L->x[q];
            L->v[q] *= damp * fix;
            L->v[q] += g[c] * dt * fix;
            L->v[q] += L->fext[q] * dt * fix;
            L->x[q] += L->v[q] * dt * fix;
            L->s[q] = L->x[q];
            L->xprev[q] = L->x[q];
            L->fext[q] = 0;
        }
    }
}

/* calculateV (PDSolver.cu:245-252). */
static void calc_v(orc_level* L, const orc_params* p) {
    real dt = p->dt;
#pragma omp parallel for schedule(static)
    for (int q = 0; q < L->V * 3; q++) L->v[q] = (L->x[q] - L->xlast[q]) / dt;
}

/* Per-tet elastic contribution: calculateIF (PDSolver.cu:128-183, mode 0) or
 * getTetFRV2 + calTetIFV1 (PDSolverAc.cu:447-662, mode 1: float constants and
 * contributions below 1e-5 in magnitude dropped). */
static void tet_forces_on(orc_level* L, const real* Xsrc, int mode) {
#pragma omp parallel for schedule(static)
    for (int t = 0; t < L->T; t++) {
        const real k = tet_k(L, t);
        const int* id = &L->tets[t * 4];
        const real* X = Xsrc;
        real D[9];
        D[0] = X[id[1] * 3 + 0] - X[id[0] * 3 + 0];
        D[1] = X[id[2] * 3 + 0] - X[id[0] * 3 + 0];
        D[2] = X[id[3] * 3 + 0] - X[id[0] * 3 + 0];
        D[3] = X[id[1] * 3 + 1] - X[id[0] * 3 + 1];
        D[4] = X[id[2] * 3 + 1] - X[id[0] * 3 + 1];
        D[5] = X[id[3] * 3 + 1] - X[id[0] * 3 + 1];
        D[6] = X[id[1] * 3 + 2] - X[id[0] * 3 + 2];
        D[7] = X[id[2] * 3 + 2] - X[id[0] * 3 + 2];
        D[8] = X[id[3] * 3 + 2] - X[id[0] * 3 + 2];
        const real* B = &L->B[t * 9];
        real F[9];
        F[0] = D[0] * B[0] + D[1] * B[3] + D[2] * B[6];
        F[1] = D[0] * B[1] + D[1] * B[4] + D[2] * B[7];
        F[2] = D[0] * B[2] + D[1] * B[5] + D[2] * B[8];
        F[3] = D[3] * B[0] + D[4] * B[3] + D[5] * B[6];
        F[4] = D[3] * B[1] + D[4] * B[4] + D[5] * B[7];
        F[5] = D[3] * B[2] + D[4] * B[5] + D[5] * B[8];
        F[6] = D[6] * B[0] + D[7] * B[3] + D[8] * B[6];
        F[7] = D[6] * B[1] + D[7] * B[4] + D[8] * B[7];
        F[8] = D[6] * B[2] + D[7] * B[5] + D[8] * B[8];
        real R[9];
        if (mode == 2) {  /* linear operator on a correction: R := 0, no clamp */
            for (int i = 0; i < 9; i++) R[i] = -F[i];
        } else {
            polar(F, R, mode == 1 ? 1 : (mode == 3 ? 2 : 0));
            for (int i = 0; i < 9; i++) R[i] = R[i] - F[i];
        }
        const real* AC = &L->AC[t * 12];
        real vol = L->vol[t];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 4; j++) {
                real acc = 0;
                for (int m = 0; m < 3; m++) acc += R[i * 3 + m] * AC[m * 4 + j];
                if (mode != 2) {
                    if (isnan(acc)) acc = 0;
                    acc = acc > 10 ? 10 : acc;
                    acc = acc < -10 ? -10 : acc;
                }
                real r = acc * vol * k;
                if (mode == 1 && R_FABS(r) < 1e-5) r = 0;
                L->cf[t * 12 + j * 3 + i] = r;
            }
    }
}

static void tet_forces(orc_level* L, int mode) { tet_forces_on(L, L->x, mode); }

/* Gather in increasing tet order == sequential scatter order. */
static void gather_forces(orc_level* L) {
#pragma omp parallel for schedule(static)
    for (int v = 0; v < L->V; v++) {
        real f[3] = {0, 0, 0};
        for (int e = L->rp[v]; e < L->rp[v + 1]; e++) {
            int te = L->ri[e];
            int t = te >> 2, c = te & 3;
            for (int i = 0; i < 3; i++) f[i] += L->cf[t * 12 + c * 3 + i];
        }
        for (int i = 0; i < 3; i++) L->force[v * 3 + i] = f[i];
    }
}

/* runClearCollision (PDSolver.cu:691-698). */
static void clear_collision(orc_level* L) {
    memset(L->cdiag, 0, sizeof(real) * 3 * (size_t)L->V);
    memset(L->coff, 0, sizeof(real) * 3 * (size_t)L->V);
    memset(L->cforce, 0, sizeof(real) * 3 * (size_t)L->V);
    memset(L->cenergy, 0, sizeof(real) * (size_t)L->V);
    memset(L->cdepth, 0, sizeof(real) * (size_t)L->V);
    memset(L->cflag, 0, sizeof(int) * (size_t)L->V);
}

/* DCDByPoint_sphere (PDSolver.cu:277-331), one active sphere, all vertices. */
static void dcd_point_sphere(orc_level* L, v3 center, real radius, real kc) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < L->V; i++) {
        v3 p = vload(L->x + i * 3);
        v3 cp = vsub(p, center);
        v3 dir = vdiv(cp, vlen(cp));
        real distance = vlen(cp);
        real oldR = radius;
        real r = radius * (real)1.5f;
        if (distance < r) {
            real a = 1;
            real b = 2 * vdot(cp, dir);
            real c = vdot(cp, cp) - r * r;
            real t = (-b + R_SQRT(b * b - 4 * a * c)) / (2 * a);
            if (t > (real)0.5f * oldR) {
                L->cflag[i] = 1;
                L->cdepth[i] = t - (real)0.5f * oldR;
            }
            t = t / (real)2.0f;
            v3 colP = vadd(p, vmul(dir, t));
            v3 ccp = vsub(colP, center);
            v3 n = vdiv(ccp, vlen(ccp));
            real nd = vdot(n, dir);
            L->cdiag[i * 3 + 0] += kc * n.x * n.x;
            L->cdiag[i * 3 + 1] += kc * n.y * n.y;
            L->cdiag[i * 3 + 2] += kc * n.z * n.z;
            L->coff[i * 3 + 0] += kc * n.x * n.y;
            L->coff[i * 3 + 1] += kc * n.x * n.z;
            L->coff[i * 3 + 2] += kc * n.y * n.z;
            L->cforce[i * 3 + 0] += kc * t * nd * n.x;
            L->cforce[i * 3 + 1] += kc * t * nd * n.y;
            L->cforce[i * 3 + 2] += kc * t * nd * n.z;
            real e = vdot(vsub(p, colP), n);
            L->cenergy[i] += kc * (e * e);
        }
    }
}

/* Ground plane n.x = d as a penalty constraint k((x - q).n)^2 with q the
 * projection on the plane.  Not in the reference; same response form as the
 * sphere's (PDSolver.cu:314-321). */
static void dcd_point_plane(orc_level* L, v3 n, real d, real kp) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < L->V; i++) {
        v3 p = vload(L->x + i * 3);
        real sd = vdot(n, p) - d;
        if (sd < 0) {
            real depth = -sd;
            L->cflag[i] = 1;
            if (depth > L->cdepth[i]) L->cdepth[i] = depth;
            L->cdiag[i * 3 + 0] += kp * n.x * n.x;
            L->cdiag[i * 3 + 1] += kp * n.y * n.y;
            L->cdiag[i * 3 + 2] += kp * n.z * n.z;
            L->coff[i * 3 + 0] += kp * n.x * n.y;
            L->coff[i * 3 + 1] += kp * n.x * n.z;
            L->coff[i * 3 + 2] += kp * n.y * n.z;
            L->cforce[i * 3 + 0] += kp * depth * n.x;
            L->cforce[i * 3 + 1] += kp * depth * n.y;
            L->cforce[i * 3 + 2] += kp * depth * n.z;
            L->cenergy[i] += kp * (depth * depth);
        }
    }
}

/* pointInTriangle (utils.cu:366-383). */
static int point_in_triangle(v3 P, v3 A, v3 B, v3 C) {
    v3 v0 = vsub(C, A), v1 = vsub(B, A), v2 = vsub(P, A);
    real d00 = vdot(v0, v0), d01 = vdot(v0, v1), d02 = vdot(v0, v2), d11 = vdot(v1, v1), d12 = vdot(v1, v2);
    real inv = (real)1.0f / (d00 * d11 - d01 * d01);
    real u = (d11 * d02 - d01 * d12) * inv;
    real w = (d00 * d12 - d01 * d02) * inv;
    return (u >= 0) && (w >= 0) && (u + w <= (real)1.0f);
}

/* UpdateTriNormal (PDSolver.cu:394-425) without the vertex accumulation. */
static void tri_normals(orc_level* L) {
#pragma omp parallel for schedule(static)
    for (int t = 0; t < L->S; t++) {
        v3 A = vload(L->x + L->tris[t * 3 + 0] * 3), B = vload(L->x + L->tris[t * 3 + 1] * 3),
           C = vload(L->x + L->tris[t * 3 + 2] * 3);
        v3 n = vnormalize(vcross(vsub(B, A), vsub(C, A)));
        L->trin[t * 3 + 0] = n.x;
        L->trin[t * 3 + 1] = n.y;
        L->trin[t * 3 + 2] = n.z;
    }
}

/* DCDByTriangle_sphere_V2 (PDSolver.cu:612-680) with the response of
 * DCDByTriangle_sphere_response (:462-501): each colliding triangle adds k to
 * the collision diagonal and k*(center - hit) to the collision force of its
 * three vertices.  Contributions are stored per (tri, corner) and gathered. */
static void dcd_tri_sphere(orc_level* L, v3 center, real radius, real kc) {
#pragma omp parallel for schedule(static)
    for (int t = 0; t < L->S; t++) {
        v3 A = vload(L->x + L->tris[t * 3 + 0] * 3), B = vload(L->x + L->tris[t * 3 + 1] * 3),
           C = vload(L->x + L->tris[t * 3 + 2] * 3);
        v3 n = vload(L->trin + t * 3);
        real distance = R_FABS(vdot(n, vsub(center, A)));
        if (distance > radius) continue;
        v3 projected = vsub(center, vmul(n, vdot(n, vsub(center, A))));
        int hitf = 0;
        v3 hit = v3m(0, 0, 0);
        if (point_in_triangle(projected, A, B, C)) {
            int sign = 1;
            if (vdot(n, vsub(center, A)) < 0) sign = -1;
            hit = vadd(center, vmul(vmul(n, radius - distance), (real)sign));
            hitf = 1;
        }
        v3 tv[3] = {A, B, C};
        for (int i = 0; i < 3 && !hitf; i++) {
            v3 Bg = tv[i], En = tv[(i + 1) % 3];
            v3 ed = vsub(En, Bg);
            real len = vlen(ed);
            ed = vnormalize(ed);
            v3 bc = vsub(center, Bg);
            real tt = vdot(bc, ed);
            if (tt < 0 || tt > len) continue;
            v3 closest = vadd(Bg, vmul(ed, tt));
            v3 radial = vsub(center, closest);
            real dist = vlen(radial);
            if (dist >= radius) continue;
            v3 Oth = tv[(i + 2) % 3];
            v3 BO = vsub(Oth, Bg), BP = vsub(projected, Bg);
            if (vdot(vcross(ed, BO), vcross(ed, BP)) >= 0) continue;
            radial = vnormalize(radial);
            hit = vadd(closest, vmul(radial, radius));
            hitf = 1;
        }
        for (int i = 0; i < 3 && !hitf; i++) {
            v3 Vq = tv[i];
            v3 VO = vsub(center, Vq);
            if (vlen(VO) > radius) continue;
            VO = vnormalize(VO);
            hit = vadd(Vq, vmul(VO, radius));
            hitf = 1;
        }
        if (!hitf) continue;
        v3 mv = vsub(center, hit);
        for (int c = 0; c < 3; c++) {
            L->tc[t * 12 + c * 4 + 0] += kc * mv.x;
            L->tc[t * 12 + c * 4 + 1] += kc * mv.y;
            L->tc[t * 12 + c * 4 + 2] += kc * mv.z;
            L->tc[t * 12 + c * 4 + 3] += kc;
        }
    }
}

static void gather_tri_collision(orc_level* L) {
#pragma omp parallel for schedule(static)
    for (int v = 0; v < L->V; v++) {
        for (int e = L->trp[v]; e < L->trp[v + 1]; e++) {
            int te = L->tri[e];
            int t = te / 3, c = te % 3;
            const real* q = &L->tc[t * 12 + c * 4];
            if (q[3] != 0) {
                L->cflag[v] += 1;
                for (int i = 0; i < 3; i++) {
                    L->cdiag[v * 3 + i] += q[3];
                    L->cforce[v * 3 + i] += q[i];
                }
            }
        }
    }
}

/* DCDByPoint (PDSolver.cpp:228-233) + plane. */
static void dcd_point(orc_level* L, const orc_params* p) {
    clear_collision(L);
    for (int s = 0; s < p->nspheres; s++) {
        const float* sp = p->sphere[s];
        if (sp[4] != 0) dcd_point_sphere(L, v3m(sp[0], sp[1], sp[2]), sp[3], p->kcol);
    }
    if (p->plane[4] > 0) dcd_point_plane(L, v3m(p->plane[0], p->plane[1], p->plane[2]), p->plane[3], p->plane[4]);
}

/* DCDByTriangle (PDSolver.cpp:235-241). */
static void dcd_triangle(orc_level* L, const orc_params* p) {
    clear_collision(L);
    tri_normals(L);
    memset(L->tc, 0, sizeof(real) * 12 * (size_t)L->S);
    int any = 0;
    for (int s = 0; s < p->nspheres; s++) {
        const float* sp = p->sphere[s];
        if (sp[4] != 0) {
            dcd_tri_sphere(L, v3m(sp[0], sp[1], sp[2]), sp[3], p->kcol);
            any = 1;
        }
    }
    if (any) gather_tri_collision(L);
}

/* calculatePOS (PDSolver.cu:192-235).  The 0.6 under-relaxation is a double
 * literal in the reference, so the FP32 oracle evaluates that line in double. */
/* contact_jacobi (V-cycle only, not the reference): vertices a collider acts
 * on take an exact 3x3 block Jacobi step without Chebyshev extrapolation;
 * the extrapolation of their stiff penalty direction kept the contact rim
 * from converging, and the per-axis diagonal pinned their tangential motion. */
static void calc_pos(orc_level* L, const orc_params* p, real omega, int contact_jacobi) {
    real dt = p->dt;
    double relax = p->relax;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < L->V; i++) {
        real fix = L->fix[i];
        real dc = L->mass[i] / (dt * dt);
        const real* cd = L->cdiag + i * 3;
        if (contact_jacobi && cd[0] + cd[1] + cd[2] > 0) {
            /* V-cycle sweeps, a vertex a collider acts on: the Jacobi step
             * solves its 3x3 block (a I + K) d = g exactly, K = sum of the
             * contacts' k n n^T (the per-axis diagonal k n_c^2 of calculatePOS
             * also pins the directions tangent to an oblique contact), with
             * no Chebyshev extrapolation.  Adjugate inverse of the SPD block. */
            const real* co = L->coff + i * 3;
            real a = L->diag[i] + dc;
            real g[3];
            for (int c = 0; c < 3; c++)
                g[c] = dc * (L->s[i * 3 + c] - L->x[i * 3 + c]) + (L->force[i * 3 + c] + L->cforce[i * 3 + c]);
            real m00 = a + cd[0], m11 = a + cd[1], m22 = a + cd[2], m01 = co[0], m02 = co[1], m12 = co[2];
            real c00 = m11 * m22 - m12 * m12, c01 = m02 * m12 - m01 * m22, c02 = m01 * m12 - m02 * m11;
            real c11 = m00 * m22 - m02 * m02, c12 = m01 * m02 - m00 * m12, c22 = m00 * m11 - m01 * m01;
            real det = m00 * c00 + m01 * c01 + m02 * c02;
            real d[3] = {(c00 * g[0] + c01 * g[1] + c02 * g[2]) / det, (c01 * g[0] + c11 * g[1] + c12 * g[2]) / det,
                         (c02 * g[0] + c12 * g[1] + c22 * g[2]) / det};
            for (int c = 0; c < 3; c++) {
                int q = i * 3 + c;
                real pos = L->x[q];
                real next = d[c] * fix + pos;
                next = (real)((double)(next - pos) * relax + (double)pos);
                L->xprev[q] = pos;
                L->x[q] = next;
            }
            continue;
        }
        for (int c = 0; c < 3; c++) {
            int q = i * 3 + c;
            real pos = L->x[q];
            real elem = L->force[q] + L->cforce[q];
            real next = (dc * (L->s[q] - pos) + elem) / (L->diag[i] + L->cdiag[q] + dc) * fix + pos;
            next = (real)((double)(next - pos) * relax + (double)pos);
            next = omega * (next - L->xprev[q]) + L->xprev[q];
            L->xprev[q] = pos;
            L->x[q] = next;
        }
    }
}

static void move_spheres(orc_params* p) {
    for (int s = 0; s < p->nspheres; s++) {
        float* sp = p->sphere[s];
        sp[0] += sp[5] * p->dt;
        sp[1] += sp[6] * p->dt;
        sp[2] += sp[7] * p->dt;
    }
}

/* One PD iteration on a level: runClearTemp; DCD; IF; omega; POS. */
static void pd_iteration(orc_level* L, const orc_params* p, real omega, int tri_dcd, int contact_jacobi) {
    if (tri_dcd)
        dcd_triangle(L, p);
    else
        dcd_point(L, p);
    tet_forces(L, p->if_mode);
    gather_forces(L);
    calc_pos(L, p, omega, contact_jacobi);
}

/* PDSolver::Step (PDSolver.cpp:243-293); sphere colliders move first, as
 * Simulator::UpdateCollider does before Step (Simulator.cpp:146-149).
 * `p` is updated in place (sphere centres advance). */
void PFX(step_single)(orc_level* L, orc_params* p, int nsteps) {
    for (int st = 0; st < nsteps; st++) {
        move_spheres(p);
        calc_st(L, p);
        real omega = 1;
        real rho = p->rho;
        for (int i = 0; i < p->iters; i++) {
            omega = 4 / (4 - rho * rho * omega);
            pd_iteration(L, p, omega, 0, 0);
        }
        calc_v(L, p);
    }
}

/* ------------------------------------------------------------------------- */
/* Two-level scheme (PDSolver_MG.cpp).                                        */
/* ------------------------------------------------------------------------- */
/* GetVolumn (utils.cu:254-275). */
static real get_volume(v3 A, v3 B, v3 C, v3 D) {
    v3 AB = vsub(B, A), AC = vsub(C, A), AD = vsub(D, A);
    real M[9] = {AB.x, AC.x, AD.x, AB.y, AC.y, AD.y, AB.z, AC.z, AD.z};
    real R0 = M[4] * M[8] - M[7] * M[5];
    real R1 = M[7] * M[2] - M[1] * M[8];
    real R2 = M[1] * M[5] - M[4] * M[2];
    real det = M[0] * R0 + M[3] * R1 + M[6] * R2;
    return det / (real)6.0f;
}

/* barycentricCoordinate, overload on four points (utils.cu:335-364). */
static void barycentric4(v3 P, v3 A, v3 B, v3 C, v3 D, real* w) {
    real V = get_volume(A, B, C, D);
    v3 center = vmul(vadd(vadd(vadd(A, B), C), D), (real)0.25f);
    v3 tv[4] = {A, B, C, D};
    static const int pf[4][4] = {{0, 1, 2, 3}, {1, 2, 3, 0}, {2, 3, 0, 1}, {3, 0, 1, 2}};
    for (int i = 0; i < 4; i++) {
        v3 p = tv[pf[i][0]], f0 = tv[pf[i][1]], f1 = tv[pf[i][2]], f2 = tv[pf[i][3]];
        v3 n = vcross(vsub(f1, f0), vsub(f2, f0));
        if (vdot(n, vsub(p, f0)) > 0) n = v3m(-n.x, -n.y, -n.z);
        real len = vlen(n);
        real area = len * (real)0.5f;
        n = vdiv(n, len);
        w[i] = (real)0.25f - vdot(vsub(P, center), n) * area / (3 * V);
    }
}

void PFX(barycentric4)(const float* P, const float* tet12, float* w) {
    real ww[4];
    barycentric4(v3m(P[0], P[1], P[2]), v3m(tet12[0], tet12[1], tet12[2]), v3m(tet12[3], tet12[4], tet12[5]),
                 v3m(tet12[6], tet12[7], tet12[8]), v3m(tet12[9], tet12[10], tet12[11]), ww);
    for (int i = 0; i < 4; i++) w[i] = (float)ww[i];
}

typedef struct orc_mg {
    orc_level *coarse, *fine;
    int* iids;   /* fine vertex -> 4 coarse vertex ids   */
    real* iw;    /* weights                              */
    int* aids;   /* coarse vertex -> 4 fine vertex ids   */
    real* aw;
    int n_unfound_fine, n_unfound_coarse;
} orc_mg;

/* PDSolver_MG::interpolate (PDSolver_MG.cpp:5-58): each fine vertex binds to
 * the first coarse tet (in index order) whose four weights all lie in
 * [-1e-5, 1+1e-5].  Brute force, O(V_fine * T_coarse). */
static void mg_interpolate_map(orc_mg* M) {
    orc_level *c = M->coarse, *f = M->fine;
    M->n_unfound_fine = 0;
    for (int vi = 0; vi < f->V; vi++) {
        v3 P = vload(f->x + vi * 3);
        int found = 0;
        for (int t = 0; t < c->T && !found; t++) {
            const int* id = &c->tets[t * 4];
            real w[4];
            barycentric4(P, vload(c->x + id[0] * 3), vload(c->x + id[1] * 3), vload(c->x + id[2] * 3),
                         vload(c->x + id[3] * 3), w);
            int in = 1;
            for (int i = 0; i < 4; i++)
                if (w[i] < -1e-5 || w[i] > 1 + 1e-5) in = 0;
            if (in) {
                for (int i = 0; i < 4; i++) {
                    M->iids[vi * 4 + i] = id[i];
                    M->iw[vi * 4 + i] = w[i];
                }
                found = 1;
            }
        }
        if (!found) M->n_unfound_fine++;
    }
}

/* pointInTet (utils.cu:298-311). */
static int point_in_tet(const real* tv12, const real* faceN12, v3 P) {
    for (int k = 0; k < 4; k++) {
        int ia = (k + 1) % 4;
        v3 A = vload(tv12 + ia * 3);
        v3 n = vload(faceN12 + k * 3);
        if (vdot(vsub(P, A), n) > 1e-6) return 0;
    }
    return 1;
}

/* Coarse-in-fine binding (PDSolver_MG.cpp:133-168): first fine tet (index
 * order) passing pointInTet, weights from the precomputed-face barycentric
 * overload (utils.cu:313-333). */
static void mg_average_map(orc_mg* M) {
    orc_level *c = M->coarse, *f = M->fine;
    M->n_unfound_coarse = 0;
    for (int vc = 0; vc < c->V; vc++) {
        v3 P = vload(c->x + vc * 3);
        int found = 0;
        for (int t = 0; t < f->T && !found; t++) {
            const int* id = &f->tets[t * 4];
            real tv[12];
            for (int i = 0; i < 4; i++)
                for (int q = 0; q < 3; q++) tv[i * 3 + q] = f->x[id[i] * 3 + q];
            if (!point_in_tet(tv, f->faceN + t * 12, P)) continue;
            v3 rB = vload(f->center + t * 3);
            real V = f->vol[t];
            for (int i = 0; i < 4; i++) {
                v3 n = vload(f->faceN + t * 12 + i * 3);
                M->aids[vc * 4 + i] = id[i];
                M->aw[vc * 4 + i] = (real)0.25f - vdot(vsub(P, rB), n) * f->faceA[t * 4 + i] / (3 * V);
            }
            found = 1;
        }
        if (!found) M->n_unfound_coarse++;
    }
}

orc_mg* PFX(mg_create)(orc_level* coarse, orc_level* fine) {
    orc_mg* M = (orc_mg*)calloc(1, sizeof(orc_mg));
    M->coarse = coarse;
    M->fine = fine;
    M->iids = (int*)xcalloc((size_t)fine->V * 4, sizeof(int));
    M->iw = (real*)xcalloc((size_t)fine->V * 4, sizeof(real));
    M->aids = (int*)xcalloc((size_t)coarse->V * 4, sizeof(int));
    M->aw = (real*)xcalloc((size_t)coarse->V * 4, sizeof(real));
    mg_interpolate_map(M);
    mg_average_map(M);
    return M;
}

void PFX(mg_free)(orc_mg* M) {
    if (!M) return;
    free(M->iids); free(M->iw); free(M->aids); free(M->aw);
    free(M);
}

int PFX(mg_unfound)(const orc_mg* M, int which) { return which ? M->n_unfound_coarse : M->n_unfound_fine; }

void PFX(mg_get_maps)(const orc_mg* M, int* iids, float* iw, int* aids, float* aw) {
    for (int i = 0; i < M->fine->V * 4; i++) { iids[i] = M->iids[i]; iw[i] = (float)M->iw[i]; }
    for (int i = 0; i < M->coarse->V * 4; i++) { aids[i] = M->aids[i]; aw[i] = (float)M->aw[i]; }
}

/* updatePointInTet (PDSolver.cu:724-744): dst.x = dst.xprev = sum w src.x. */
static void update_point_in_tet(orc_level* dst, const orc_level* src, const int* ids, const real* w) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < dst->V; i++) {
        const int* id = ids + i * 4;
        const real* ww = w + i * 4;
        for (int c = 0; c < 3; c++) {
            real r = ww[0] * src->x[id[0] * 3 + c] + ww[1] * src->x[id[1] * 3 + c] + ww[2] * src->x[id[2] * 3 + c] +
                     ww[3] * src->x[id[3] * 3 + c];
            dst->x[i * 3 + c] = r;
            dst->xprev[i * 3 + c] = r;
        }
    }
}

/* updateMapping (PDSolver.cu:764-785). */
static void update_mapping(orc_mg* M) {
    orc_level *c = M->coarse, *f = M->fine;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < f->V; i++) {
        const int* id = M->iids + i * 4;
        real w[4];
        barycentric4(vload(f->x + i * 3), vload(c->x + id[0] * 3), vload(c->x + id[1] * 3), vload(c->x + id[2] * 3),
                     vload(c->x + id[3] * 3), w);
        for (int q = 0; q < 4; q++) M->iw[i * 4 + q] = w[q];
    }
}

/* PDSolver_MG::Step (PDSolver_MG.cpp:183-253). The same parameter block serves
 * both levels, as both PDSolver::Init calls set the same constants. */
void PFX(step_mg)(orc_mg* M, orc_params* p, int nsteps) {
    orc_level *c = M->coarse, *f = M->fine;
    real rho = p->rho;
    for (int st = 0; st < nsteps; st++) {
        move_spheres(p);
        calc_st(c, p);
        calc_st(f, p);
        real omega = 1;
        int mode = p->if_mode;
        orc_params pc = *p;
        pc.if_mode = mode;
        for (int i = 0; i < p->iters_coarse; i++) {
            omega = 4 / (4 - rho * rho * omega);
            pd_iteration(c, &pc, omega, p->tri_dcd_coarse, 0);
        }
        update_point_in_tet(f, c, M->iids, M->iw);
        omega = 1;
        for (int i = 0; i < p->iters; i++) {
            omega = 4 / (4 - rho * rho * omega);
            pd_iteration(f, p, omega, 0, 0);
        }
        update_point_in_tet(c, f, M->aids, M->aw);
        update_mapping(M);
        calc_v(c, p);
        calc_v(f, p);
    }
}

/* ------------------------------------------------------------------------- */
/* Diagnostics.                                                               */
/* ------------------------------------------------------------------------- */
/* runCalEnergy (PDSolver.cu:816-913): out = {Ek, Ep, Ec, Nc, dX}, in double. */
void PFX(energy)(const orc_level* L, const orc_params* p, double* out) {
    double Ek = 0, Ep = 0, Ec = 0, Nc = 0, dX = 0;
    double dt = p->dt;
    for (int i = 0; i < L->V; i++) {
        double dx2 = 0, ds2 = 0;
        for (int c = 0; c < 3; c++) {
            double a = (double)L->x[i * 3 + c] - (double)L->xprev[i * 3 + c];
            double b = (double)L->x[i * 3 + c] - (double)L->s[i * 3 + c];
            dx2 += a * a;
            ds2 += b * b;
        }
        dX += dx2;
        Ek += (double)L->mass[i] * ds2;
    }
    dX = sqrt(dX);
    Ek = Ek / (2 * dt * dt);
    for (int t = 0; t < L->T; t++) {
        const int* id = &L->tets[t * 4];
        const real* X = L->x;
        real D[9];
        for (int r = 0; r < 3; r++)
            for (int q = 0; q < 3; q++) D[r * 3 + q] = X[id[q + 1] * 3 + r] - X[id[0] * 3 + r];
        const real* B = &L->B[t * 9];
        real F[9], R[9];
        for (int r = 0; r < 3; r++)
            for (int q = 0; q < 3; q++) F[r * 3 + q] = D[r * 3 + 0] * B[q] + D[r * 3 + 1] * B[3 + q] + D[r * 3 + 2] * B[6 + q];
        polar(F, R, 0);
        double e = 0;
        for (int i = 0; i < 9; i++) e += ((double)F[i] - R[i]) * ((double)F[i] - R[i]);
        Ep += e * (double)L->vol[t] * (double)tet_k(L, t) / 2;
    }
    for (int i = 0; i < L->V; i++) {
        Ec += L->cenergy[i];
        Nc += L->cflag[i] > 0 ? 1 : 0;
    }
    out[0] = Ek; out[1] = Ep; out[2] = Ec; out[3] = Nc; out[4] = dX;
}

/* PDSolver::StepForConvergence (PDSolver.cpp:202-221): save the velocity,
 * calculateST (colliders advanced as the next step will see them), `iters`
 * iterations of DCDByPoint + calculateIF (mode 0) + omega + calculatePOS
 * with runCalEnergy after each, then x = x_last and v restored
 * (runResetPosVel, PDSolver.cu:270-273).  E[0] = Ek + Ep + Ec at s after a
 * DCD at s (PDSolver::Step's log, PDSolver.cpp:252-255), E[i] after
 * iteration i.  `p` is not changed. */
void PFX(step_convergence)(orc_level* L, const orc_params* p0, int iters, double* E) {
    orc_params p = *p0;
    move_spheres(&p);
    size_t n3 = (size_t)L->V * 3;
    real* vbak = (real*)xcalloc(n3, sizeof(real));
    memcpy(vbak, L->v, sizeof(real) * n3);
    calc_st(L, &p);
    dcd_point(L, &p);
    double e[5];
    PFX(energy)(L, &p, e);
    E[0] = e[0] + e[1] + e[2];
    orc_params pp = p;
    pp.if_mode = 0;
    real omega = 1, rho = p.rho;
    for (int i = 0; i < iters; i++) {
        omega = 4 / (4 - rho * rho * omega);
        pd_iteration(L, &pp, omega, 0, 0);
        PFX(energy)(L, &pp, e);
        E[i + 1] = e[0] + e[1] + e[2];
    }
    memcpy(L->x, L->xlast, sizeof(real) * n3);
    memcpy(L->v, vbak, sizeof(real) * n3);
    free(vbak);
}

/* runCalEnergy's per-vertex Ep density (calEveryVertEp, PDSolver.cu:893-924):
 * per tet e = sum (F - R)^2 (float differences squared in double) * vol * k / 2,
 * negated where det F < 1; EpSum and VSum (current signed volume, GetVolumn)
 * are float accumulators summed in tet order; density = EpSum / VSum. */
void PFX(ep_density)(const orc_level* L, float* out) {
    float* es = (float*)xcalloc((size_t)L->V, sizeof(float));
    float* vs = (float*)xcalloc((size_t)L->V, sizeof(float));
    for (int t = 0; t < L->T; t++) {
        const int* id = &L->tets[t * 4];
        const real* X = L->x;
        real D[9];
        for (int r = 0; r < 3; r++)
            for (int q = 0; q < 3; q++) D[r * 3 + q] = X[id[q + 1] * 3 + r] - X[id[0] * 3 + r];
        const real* B = &L->B[t * 9];
        real F[9], R[9];
        for (int r = 0; r < 3; r++)
            for (int q = 0; q < 3; q++) F[r * 3 + q] = D[r * 3 + 0] * B[q] + D[r * 3 + 1] * B[3 + q] + D[r * 3 + 2] * B[6 + q];
        real det = polar(F, R, 0);
        double e = 0;
        for (int i = 0; i < 9; i++) {
            real d = F[i] - R[i];
            e += (double)d * (double)d;
        }
        e = e * (double)L->vol[t] * (double)tet_k(L, t) / 2;
        if (det < 1.0f) e = -e;
        v3 a = vload(X + id[0] * 3), b = vload(X + id[1] * 3), c = vload(X + id[2] * 3), d4 = vload(X + id[3] * 3);
        float vol = (float)get_volume(a, b, c, d4);
        for (int j = 0; j < 4; j++) {
            es[id[j]] = (float)((double)es[id[j]] + e);
            vs[id[j]] += vol;
        }
    }
    for (int i = 0; i < L->V; i++) out[i] = es[i] / vs[i];
    free(es);
    free(vs);
}

/* Vertex normals, runUpdateOutsideTetVertNormal (PDSolver.cu:394-460):
 * unit triangle normals summed per vertex then normalised if |n| > FLT_EPSILON. */
void PFX(vertex_normals)(orc_level* L) {
    tri_normals(L);
    for (int v = 0; v < L->V; v++) {
        real n[3] = {0, 0, 0};
        for (int e = L->trp[v]; e < L->trp[v + 1]; e++) {
            int t = L->tri[e] / 3;
            for (int i = 0; i < 3; i++) n[i] += L->trin[t * 3 + i];
        }
        real len = R_SQRT(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (len > (real)1.1920928955078125e-7) {
            for (int i = 0; i < 3; i++) n[i] /= len;
        }
        for (int i = 0; i < 3; i++) L->vnormal[v * 3 + i] = n[i];
    }
}

/* Nonlinear PD residual g(x) = M/h^2 (s - x) + f_el(x) + f_col(x) over free
 * DOFs, FP64 norm.  The residual the V-cycle drives down; the reference has no
 * residual (it logs energies), so this is the build's own definition. */
double PFX(residual_norm)(orc_level* L, const orc_params* p) {
    dcd_point(L, p);
    tet_forces(L, 0);
    gather_forces(L);
    double acc = 0;
    real dt = p->dt;
    for (int i = 0; i < L->V; i++) {
        if (L->fix[i] == 0) continue;
        real dc = L->mass[i] / (dt * dt);
        for (int c = 0; c < 3; c++) {
            int q = i * 3 + c;
            double g = (double)dc * ((double)L->s[q] - (double)L->x[q]) + (double)L->force[q] + (double)L->cforce[q];
            acc += g * g;
        }
    }
    return sqrt(acc);
}

/* Unit-level entry points for tests. */
float PFX(polar_f)(const float* F, float* R, int variant) {
    real FF[9], RR[9];
    for (int i = 0; i < 9; i++) FF[i] = F[i];
    real det = polar(FF, RR, variant);
    for (int i = 0; i < 9; i++) R[i] = (float)RR[i];
    return (float)det;
}

float PFX(inverse3_f)(const float* A, float* R) {
    real AA[9], RR[9];
    for (int i = 0; i < 9; i++) AA[i] = A[i];
    real det = mat_inverse3(AA, RR);
    for (int i = 0; i < 9; i++) R[i] = (float)RR[i];
    return (float)det;
}

/* Per-tet corner forces at the current positions (12 per tet, corner-major). */
void PFX(tet_corner_forces)(orc_level* L, int mode, double* out) {
    tet_forces(L, mode);
    for (size_t i = 0; i < (size_t)L->T * 12; i++) out[i] = (double)L->cf[i];
}

/* One PD sweep on the current state: point DCD, local step, gather and the
 * Jacobi step (contact_jacobi: the V-cycle's block step on contact
 * vertices). */
void PFX(sweep)(orc_level* L, const orc_params* p, double omega, int contact_jacobi) {
    pd_iteration(L, p, (real)omega, 0, contact_jacobi);
}

/* One collision pass on the current positions (point DCD or triangle DCD). */
void PFX(collide)(orc_level* L, const orc_params* p, int tri) {
    if (tri)
        dcd_triangle(L, p);
    else
        dcd_point(L, p);
}

/* ========================================================================= */
/* Residual-correction V-cycle (this build's multigrid schedule; the         */
/* reference's own two-level scheme is step_mg above).                      */
/*                                                                           */
/* Per step: inertial prediction; then up to max_cycles cycles of            */
/*   pre  : `pre` nonlinear Jacobi/Chebyshev sweeps on level 0 (the          */
/*          reference sweep: local step + calculatePOS),                     */
/*   g    : residual g = m/h^2 (s - x) + f_el + f_col on free DOFs,          */
/*   coarse correction: r_1 = P^T g; linear V-cycle on A_l e = r_l for       */
/*          l >= 1 (A = m/h^2 + L, L the ARAP stiffness at R = 0), damped   */
/*          Jacobi with Chebyshev acceleration (rho per level by power       */
/*          iteration), FP64 Cholesky on the coarsest level;                 */
/*   x += P e_1 (free DOFs), post-smoothing, residual norm (FP64).           */
/* The cycle loop stops once ||g|| <= tol * ||g(s)|| when `check` is set.    */
/* ========================================================================= */
#define ORC_MAXL 8

typedef struct orc_vparams {
    int pre, post, coarse_sweeps, max_cycles, coarse_direct, check;
    float tol, rho_smooth, relax_coarse, rho_coarse;
    int coarse_first; /* with checks: the first half-cycle is the coarse correction */
    int stall_guard;  /* a check above the previous one turns the fine Chebyshev off for the step */
} orc_vparams;

typedef struct orc_ml {
    int nl;
    orc_level* L[ORC_MAXL];
    int* pid[ORC_MAXL];  /* level l vertex -> 4 vertex ids of level l+1 */
    real* pw[ORC_MAXL];
    real* vfix[ORC_MAXL];
    real *e[ORC_MAXL], *ep[ORC_MAXL], *r[ORC_MAXL], *q[ORC_MAXL];
    double rho[ORC_MAXL];
    int nfree;
    int* dense;          /* coarsest vertex -> dense index or -1 */
    double* chol;        /* lower Cholesky factor, nfree^2       */
    double res0, res;
    int cycles, halves;
    float relax_coarse, dt;
    real* warm;          /* pending warm start: level-1 correction (V1 x 3) */
    real* cmask;         /* level 0: 0 where a collider acts (active set), else 1 */
    real* fmask;         /* level 0: fix * cmask, the coarse corrections' mask      */
    real* warm0;         /* ... and a fine-level correction (V0 x 3)        */
    int has_warm0;
    int has_warm, fallback;
    double res_ws, res_t;
} orc_ml;

/* q = A e = m/h^2 e - f_lin(e) on level l (f_lin from the linear tet pass). */
static void ml_apply(orc_ml* M, int l, const real* e, real* q) {
    orc_level* L = M->L[l];
    tet_forces_on(L, e, 2);
    gather_forces(L);
    real dt = M->dt;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < L->V; i++) {
        real dc = L->mass[i] / (dt * dt);
        for (int c = 0; c < 3; c++) q[i * 3 + c] = dc * e[i * 3 + c] - L->force[i * 3 + c];
    }
}

/* Spectral radius of J = I - relax D^-1 A on free DOFs of level l, power
 * iteration (100 steps, fixed start vector), FP64. */
static double ml_power_rho(orc_ml* M, int l) {
    orc_level* L = M->L[l];
    int V = L->V;
    double dt = M->dt, relax = M->relax_coarse;
    real* w = (real*)xcalloc((size_t)V * 3, sizeof(real));
    real* q = (real*)xcalloc((size_t)V * 3, sizeof(real));
    double* wd = (double*)xcalloc((size_t)V, sizeof(double));
    for (int i = 0; i < V; i++) wd[i] = M->vfix[l][i] * (1.0 + 0.5 * sin(1.7 * i));
    double lam = 0;
    for (int it = 0; it < 100; it++) {
        double nw = 0;
        for (int i = 0; i < V; i++) nw += wd[i] * wd[i];
        nw = sqrt(nw);
        if (nw == 0) break;
        for (int i = 0; i < V; i++) {
            wd[i] /= nw;
            for (int c = 0; c < 3; c++) w[i * 3 + c] = (real)wd[i];
        }
        ml_apply(M, l, w, q);
        double ny = 0;
        for (int i = 0; i < V; i++) {
            double d = (double)L->mass[i] / (dt * dt) + (double)L->diag[i];
            double y = M->vfix[l][i] * (wd[i] - relax * (double)q[i * 3] / d);
            wd[i] = y;
            ny += y * y;
        }
        lam = sqrt(ny);
    }
    free(w); free(q); free(wd);
    return lam;
}

/* Dense A on the free vertices of the coarsest level, FP64 Cholesky. */
static int ml_factor_coarsest(orc_ml* M) {
    int l = M->nl - 1;
    orc_level* L = M->L[l];
    int V = L->V, n = 0;
    M->dense = (int*)xcalloc((size_t)V, sizeof(int));
    for (int i = 0; i < V; i++) M->dense[i] = M->vfix[l][i] != 0 ? n++ : -1;
    M->nfree = n;
    double* A = (double*)xcalloc((size_t)n * n, sizeof(double));
    double dt = M->dt;
    for (int i = 0; i < V; i++)
        if (M->dense[i] >= 0) A[(size_t)M->dense[i] * n + M->dense[i]] += (double)L->mass[i] / (dt * dt);
    for (int t = 0; t < L->T; t++) {
        const real* AC = &L->AC[t * 12];
        double w = (double)L->vol[t] * (double)tet_k(L, t);
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) {
                int ia = M->dense[L->tets[t * 4 + a]], ib = M->dense[L->tets[t * 4 + b]];
                if (ia < 0 || ib < 0) continue;
                double s = 0;
                for (int r = 0; r < 3; r++) s += (double)AC[r * 4 + a] * (double)AC[r * 4 + b];
                A[(size_t)ia * n + ib] += w * s;
            }
    }
    for (int j = 0; j < n; j++) {
        double d = A[(size_t)j * n + j];
        for (int p = 0; p < j; p++) d -= A[(size_t)j * n + p] * A[(size_t)j * n + p];
        if (d <= 0) { free(A); return -1; }
        d = sqrt(d);
        A[(size_t)j * n + j] = d;
        for (int i = j + 1; i < n; i++) {
            double s = A[(size_t)i * n + j];
            for (int p = 0; p < j; p++) s -= A[(size_t)i * n + p] * A[(size_t)j * n + p];
            A[(size_t)i * n + j] = s / d;
        }
    }
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) A[(size_t)i * n + j] = 0;
    M->chol = A;
    return 0;
}

static void ml_solve_coarsest(orc_ml* M, const real* r, real* e) {
    int l = M->nl - 1, n = M->nfree, V = M->L[l]->V;
    double* y = (double*)xcalloc((size_t)n * 3, sizeof(double));
    for (int i = 0; i < V; i++)
        if (M->dense[i] >= 0)
            for (int c = 0; c < 3; c++) y[(size_t)M->dense[i] * 3 + c] = r[i * 3 + c];
    const double* Lc = M->chol;
    for (int c = 0; c < 3; c++) {
        for (int i = 0; i < n; i++) {
            double s = y[(size_t)i * 3 + c];
            for (int p = 0; p < i; p++) s -= Lc[(size_t)i * n + p] * y[(size_t)p * 3 + c];
            y[(size_t)i * 3 + c] = s / Lc[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; i--) {
            double s = y[(size_t)i * 3 + c];
            for (int p = i + 1; p < n; p++) s -= Lc[(size_t)p * n + i] * y[(size_t)p * 3 + c];
            y[(size_t)i * 3 + c] = s / Lc[(size_t)i * n + i];
        }
    }
    for (int i = 0; i < V; i++)
        for (int c = 0; c < 3; c++) e[i * 3 + c] = M->dense[i] >= 0 ? (real)y[(size_t)M->dense[i] * 3 + c] : 0;
    free(y);
}

/* r_{l+1} = P_l^T r_l, accumulated over fine vertices in ascending order. */
static void ml_restrict(orc_ml* M, int l, const real* rf, real* rc) {
    orc_level* C = M->L[l + 1];
    memset(rc, 0, sizeof(real) * 3 * (size_t)C->V);
    for (int i = 0; i < M->L[l]->V; i++)
        for (int q = 0; q < 4; q++) {
            int j = M->pid[l][i * 4 + q];
            real w = M->pw[l][i * 4 + q];
            for (int c = 0; c < 3; c++) rc[j * 3 + c] += w * rf[i * 3 + c];
        }
    for (int j = 0; j < C->V; j++)
        for (int c = 0; c < 3; c++) rc[j * 3 + c] *= M->vfix[l + 1][j];
}

/* dst (level l) += mask * P_l src (level l+1). */
static void ml_prolong_add(orc_ml* M, int l, const real* src, real* dst, const real* mask) {
    orc_level* F = M->L[l];
#pragma omp parallel for schedule(static)
    for (int i = 0; i < F->V; i++) {
        const int* id = M->pid[l] + i * 4;
        const real* w = M->pw[l] + i * 4;
        for (int c = 0; c < 3; c++) {
            real v = w[0] * src[id[0] * 3 + c] + w[1] * src[id[1] * 3 + c] + w[2] * src[id[2] * 3 + c] +
                     w[3] * src[id[3] * 3 + c];
            dst[i * 3 + c] = dst[i * 3 + c] + mask[i] * v;
        }
    }
}

/* Damped Jacobi + Chebyshev sweeps on A_l e = r_l; `zero` = e starts at 0. */
static void ml_smooth(orc_ml* M, int l, int sweeps, int zero) {
    orc_level* L = M->L[l];
    real *e = M->e[l], *ep = M->ep[l], *r = M->r[l], *q = M->q[l];
    real relax = M->relax_coarse, dt = M->dt;
    real rho = (real)M->rho[l];
    real omega = 1;
    for (int k = 0; k < sweeps; k++) {
        omega = 4 / (4 - rho * rho * omega);
        int first_zero = zero && k == 0;
        if (!first_zero) ml_apply(M, l, e, q);
        for (int i = 0; i < L->V; i++) {
            real d = L->mass[i] / (dt * dt) + L->diag[i];
            real fx = M->vfix[l][i];
            for (int c = 0; c < 3; c++) {
                int j = i * 3 + c;
                real ec = first_zero ? 0 : e[j];
                real qc = first_zero ? 0 : q[j];
                real epc = first_zero ? 0 : ep[j];
                real y = ec + fx * (relax * ((r[j] - qc) / d));
                y = omega * (y - epc) + epc;
                ep[j] = ec;
                e[j] = y;
            }
        }
    }
}

static void ml_vcycle_lin(orc_ml* M, int l, const orc_vparams* vp) {
    orc_level* L = M->L[l];
    size_t n3 = (size_t)L->V * 3;
    if (l == M->nl - 1) {
        if (vp->coarse_direct && M->chol) {
            ml_solve_coarsest(M, M->r[l], M->e[l]);
        } else {
            ml_smooth(M, l, vp->coarse_sweeps, 1);
        }
        return;
    }
    ml_smooth(M, l, vp->coarse_sweeps, 1);
    ml_apply(M, l, M->e[l], M->q[l]);
    real* rr = (real*)xcalloc(n3, sizeof(real));
    for (int i = 0; i < L->V; i++)
        for (int c = 0; c < 3; c++) rr[i * 3 + c] = M->vfix[l][i] * (M->r[l][i * 3 + c] - M->q[l][i * 3 + c]);
    ml_restrict(M, l, rr, M->r[l + 1]);
    free(rr);
    ml_vcycle_lin(M, l + 1, vp);
    ml_prolong_add(M, l, M->e[l + 1], M->e[l], M->vfix[l]);
    memcpy(M->ep[l], M->e[l], sizeof(real) * n3);
    /* post-smoothing from the corrected e (not zero) */
    {
        orc_level* Lq = L;
        real *e = M->e[l], *ep = M->ep[l], *r = M->r[l], *q = M->q[l];
        real relax = M->relax_coarse, dt = M->dt, rho = (real)M->rho[l], omega = 1;
        for (int k = 0; k < vp->coarse_sweeps; k++) {
            omega = 4 / (4 - rho * rho * omega);
            ml_apply(M, l, e, q);
            for (int i = 0; i < Lq->V; i++) {
                real d = Lq->mass[i] / (dt * dt) + Lq->diag[i];
                real fx = M->vfix[l][i];
                for (int c = 0; c < 3; c++) {
                    int j = i * 3 + c;
                    real y = e[j] + fx * (relax * ((r[j] - q[j]) / d));
                    y = omega * (y - ep[j]) + ep[j];
                    ep[j] = e[j];
                    e[j] = y;
                }
            }
        }
    }
}

/* g(x) on level 0 into r[0] (free DOFs), returns ||g|| (FP64); also restarts
 * the fine Chebyshev history (x_prev = x). */
/* Nonlinear residual on level 0 (stored in r[0]) and its norm; also restarts
 * x_prev.  Vertices a collider acts on form the active set: their residual
 * is left out of the restriction and the coarse corrections skip them
 * (cmask = 0), since the coarse operators carry no penalty stiffness. */
static double ml_fine_residual(orc_ml* M, const orc_params* p) {
    orc_level* L = M->L[0];
    dcd_point(L, p);
    tet_forces(L, p->if_mode);
    gather_forces(L);
    if (!M->cmask) M->cmask = (real*)xcalloc((size_t)L->V, sizeof(real));
    double acc = 0;
    real dt = p->dt;
    for (int i = 0; i < L->V; i++) {
        real dc = L->mass[i] / (dt * dt);
        real fx = L->fix[i];
        M->cmask[i] = (L->cdiag[i * 3] + L->cdiag[i * 3 + 1] + L->cdiag[i * 3 + 2] > 0) ? 0 : 1;
        for (int c = 0; c < 3; c++) {
            int q = i * 3 + c;
            real g = (dc * (L->s[q] - L->x[q]) + (L->force[q] + L->cforce[q])) * fx;
            M->r[0][q] = g * M->cmask[i];
            acc += (double)g * (double)g;
            L->xprev[q] = L->x[q];
        }
    }
    return sqrt(acc);
}

/* Mask of the fine-level coarse corrections: Dirichlet flags times the
 * active-set mask of the last residual evaluation. */
static const real* ml_fine_mask(orc_ml* M) {
    orc_level* L = M->L[0];
    if (!M->fmask) M->fmask = (real*)xcalloc((size_t)L->V, sizeof(real));
    for (int i = 0; i < L->V; i++) M->fmask[i] = L->fix[i] * (M->cmask ? M->cmask[i] : 1);
    return M->fmask;
}

orc_ml* PFX(ml_create)(orc_level** levels, int nl, float relax_coarse, float dt, float rho_coarse) {
    if (nl < 2 || nl > ORC_MAXL) return NULL;
    orc_ml* M = (orc_ml*)calloc(1, sizeof(orc_ml));
    M->nl = nl;
    M->relax_coarse = relax_coarse;
    M->dt = dt;
    for (int l = 0; l < nl; l++) {
        M->L[l] = levels[l];
        size_t n3 = (size_t)levels[l]->V * 3;
        M->e[l] = (real*)xcalloc(n3, sizeof(real));
        M->ep[l] = (real*)xcalloc(n3, sizeof(real));
        M->r[l] = (real*)xcalloc(n3, sizeof(real));
        M->q[l] = (real*)xcalloc(n3, sizeof(real));
        M->vfix[l] = (real*)xcalloc(levels[l]->V, sizeof(real));
    }
    for (int i = 0; i < levels[0]->V; i++) M->vfix[0][i] = levels[0]->fix[i];
    for (int l = 0; l + 1 < nl; l++) {
        orc_mg tmp;
        memset(&tmp, 0, sizeof(tmp));
        tmp.coarse = levels[l + 1];
        tmp.fine = levels[l];
        tmp.iids = (int*)xcalloc((size_t)levels[l]->V * 4, sizeof(int));
        tmp.iw = (real*)xcalloc((size_t)levels[l]->V * 4, sizeof(real));
        tmp.aids = (int*)xcalloc((size_t)levels[l + 1]->V * 4, sizeof(int));
        tmp.aw = (real*)xcalloc((size_t)levels[l + 1]->V * 4, sizeof(real));
        mg_interpolate_map(&tmp);
        mg_average_map(&tmp);
        M->pid[l] = tmp.iids;
        M->pw[l] = tmp.iw;
        /* coarse vertex pinned iff its largest-weight fine corner is pinned */
        for (int j = 0; j < levels[l + 1]->V; j++) {
            int best = 0;
            for (int q = 1; q < 4; q++)
                if (tmp.aw[j * 4 + q] > tmp.aw[j * 4 + best]) best = q;
            M->vfix[l + 1][j] = M->vfix[l][tmp.aids[j * 4 + best]] != 0 ? 1 : 0;
        }
        free(tmp.aids);
        free(tmp.aw);
    }
    for (int l = 1; l < nl; l++) M->rho[l] = rho_coarse > 0 ? rho_coarse : ml_power_rho(M, l);
    if (ml_factor_coarsest(M)) {
        free(M->chol);
        M->chol = NULL;
    }
    return M;
}

void PFX(ml_free)(orc_ml* M) {
    if (!M) return;
    for (int l = 0; l < M->nl; l++) {
        free(M->e[l]); free(M->ep[l]); free(M->r[l]); free(M->q[l]); free(M->vfix[l]);
        free(M->pid[l]); free(M->pw[l]);
    }
    free(M->dense);
    free(M->chol);
    free(M->warm);
    free(M->warm0);
    free(M->cmask);
    free(M->fmask);
    free(M);
}

double PFX(ml_rho)(const orc_ml* M, int l) { return M->rho[l]; }
int PFX(ml_nfree)(const orc_ml* M) { return M->chol ? M->nfree : -1; }
void PFX(ml_get_vfix)(const orc_ml* M, int l, float* out) {
    for (int i = 0; i < M->L[l]->V; i++) out[i] = (float)M->vfix[l][i];
}
/* Active-set mask of the last fine residual evaluation (1 where no collider
 * acts, 0 where one does); all ones before the first evaluation. */
void PFX(ml_get_mask)(const orc_ml* M, float* out) {
    for (int i = 0; i < M->L[0]->V; i++) out[i] = M->cmask ? (float)M->cmask[i] : 1.0f;
}
void PFX(ml_stats)(const orc_ml* M, double* out) {
    out[0] = M->res0;
    out[1] = M->res;
    out[2] = M->cycles;
    out[3] = M->halves;
    out[4] = M->fallback;
    out[5] = M->res_ws;
    out[6] = M->res_t;
}

/* Warm start of the next step: x = s + fix * u0 + P e1, with e1 the
 * learned schedule's coarse correction (already masked by the level-1
 * Dirichlet flags) and u0 a fine-level correction; either may be NULL. */
void PFX(ml_set_warm)(orc_ml* M, const float* e1, const float* u0) {
    int V1 = M->L[1]->V, V0 = M->L[0]->V;
    M->has_warm = 0;
    M->has_warm0 = 0;
    if (e1) {
        if (!M->warm) M->warm = (real*)xcalloc((size_t)V1 * 3, sizeof(real));
        for (int i = 0; i < V1 * 3; i++) M->warm[i] = e1[i];
        M->has_warm = 1;
    }
    if (u0) {
        if (!M->warm0) M->warm0 = (real*)xcalloc((size_t)V0 * 3, sizeof(real));
        for (int i = 0; i < V0 * 3; i++) M->warm0[i] = u0[i];
        M->has_warm0 = 1;
    }
}

/* Coarse-correction half of a cycle: the masked residual M->r[0] of the last
 * fine evaluation restricted, the linear V-cycle on the coarse levels,
 * prolongation, post-smoothing; returns the residual after it.  calm: the
 * stall guard has turned the Chebyshev extrapolation off (omega 1). */
static double ml_coarse_half(orc_ml* M, orc_params* p, const orc_vparams* vp, int calm) {
    orc_level* L0 = M->L[0];
    real rs = vp->rho_smooth;
    ml_restrict(M, 0, M->r[0], M->r[1]);
    ml_vcycle_lin(M, 1, vp);
    ml_prolong_add(M, 0, M->e[1], L0->x, ml_fine_mask(M));
    memcpy(L0->xprev, L0->x, sizeof(real) * 3 * (size_t)L0->V);
    real omega = 1;
    for (int k = 0; k < vp->post; k++) {
        omega = 4 / (4 - rs * rs * omega);
        pd_iteration(L0, p, calm ? 1 : omega, 0, 1);
    }
    return ml_fine_residual(M, p);
}

/* One residual-correction V-cycle step (the device schedule MGSB_SCHED_VCYCLE).
 * check: 0 no checks (max_cycles cycles), 1 after every cycle, 2 also after
 * the pre-smoothing half (the residual the restriction uses).  cycles counts
 * completed cycles, halves the half-cycles run (pre-smoothing half, coarse
 * correction + post-smoothing half).  coarse_first (with checks): the first
 * half is a coarse-correction half on the start's residual (counted as one
 * half and one cycle), then max_cycles cycles at most.  stall_guard: once a
 * check's residual exceeds 3 times the previous one (the start's, or the
 * accepted warm start's), the fine sweeps drop the Chebyshev extrapolation
 * (omega 1) for the rest of the step. */
void PFX(step_vcycle)(orc_ml* M, orc_params* p, const orc_vparams* vp, int nsteps) {
    orc_level* L0 = M->L[0];
    real rs = vp->rho_smooth;
    const int cfirst = vp->coarse_first && vp->check;
    for (int st = 0; st < nsteps; st++) {
        move_spheres(p);
        calc_st(L0, p);
        M->res0 = ml_fine_residual(M, p);
        M->res = M->res0;
        M->cycles = 0;
        M->halves = 0;
        M->fallback = 0;
        M->res_ws = 0.0;
        M->res_t = 0.0;
        int conv = vp->check && M->res0 == 0.0;
        /* warm start in two stages, each kept only if it lowers the best
         * residual so far: x = s + fix u0 (temporal), then x += P e1 (learned);
         * a non-finite residual (a diverged network) is never kept */
        int accepted = 0;
        if (M->has_warm0 && !conv) {
            for (int i = 0; i < L0->V; i++)
                for (int c = 0; c < 3; c++) L0->x[i * 3 + c] = L0->x[i * 3 + c] + L0->fix[i] * M->warm0[i * 3 + c];
            double rt = ml_fine_residual(M, p);
            M->res_t = rt;
            if (!(rt <= M->res)) {
                M->fallback |= 1;
                memcpy(L0->x, L0->s, sizeof(real) * 3 * (size_t)L0->V);
                ml_fine_residual(M, p);
            } else {
                M->res = rt;
                accepted = 1;
                conv = vp->check && rt <= (double)vp->tol * M->res0;
            }
        }
        if (M->has_warm && !conv) {
            ml_prolong_add(M, 0, M->warm, L0->x, ml_fine_mask(M));
            double rw = ml_fine_residual(M, p);
            M->res_ws = rw;
            if (!(rw <= M->res)) {
                M->fallback |= 2;
                memcpy(L0->x, L0->s, sizeof(real) * 3 * (size_t)L0->V);
                if (accepted == 1)
                    for (int i = 0; i < L0->V; i++)
                        for (int c = 0; c < 3; c++)
                            L0->x[i * 3 + c] = L0->x[i * 3 + c] + L0->fix[i] * M->warm0[i * 3 + c];
                ml_fine_residual(M, p);
            } else {
                M->res = rw;
                conv = vp->check && rw <= (double)vp->tol * M->res0;
            }
        }
        M->has_warm = 0;
        M->has_warm0 = 0;
        int calm = 0;
#define ORC_CHECK(r) \
    do { \
        if (vp->stall_guard && !((r) <= 3.0 * M->res)) calm = 1; \
        M->res = (r); \
    } while (0)
        if (cfirst && !conv) {
            double rc = ml_coarse_half(M, p, vp, calm);
            ORC_CHECK(rc);
            M->cycles++;
            M->halves++;
            conv = M->res <= (double)vp->tol * M->res0;
        }
        for (int c = 0; c < vp->max_cycles && !conv; c++) {
            real omega = 1;
            for (int k = 0; k < vp->pre; k++) {
                omega = 4 / (4 - rs * rs * omega);
                pd_iteration(L0, p, calm ? 1 : omega, 0, 1);
            }
            double rm = ml_fine_residual(M, p);
            if (vp->check == 2) {
                ORC_CHECK(rm);
                M->halves++;
                if (rm <= (double)vp->tol * M->res0) break;
            }
            double rc = ml_coarse_half(M, p, vp, calm);
            if (vp->check)
                ORC_CHECK(rc);
            else
                M->res = rc;
            M->cycles++;
            M->halves += vp->check == 2 ? 1 : 2;
            conv = vp->check && M->res <= (double)vp->tol * M->res0;
        }
#undef ORC_CHECK
        calc_v(L0, p);
    }
}

/* `sweeps` damped-Jacobi + Chebyshev sweeps on A_l e = r from e = 0 on level
 * l (ml_smooth; test entry point). */
void PFX(ml_lin_smooth)(orc_ml* M, int l, int sweeps, const float* r, float* e) {
    int V = M->L[l]->V;
    for (int i = 0; i < V * 3; i++) M->r[l][i] = r[i];
    ml_smooth(M, l, sweeps, 1);
    for (int i = 0; i < V * 3; i++) e[i] = (float)M->e[l][i];
}

/* Coarsest-level solve on a given right-hand side (test entry point). */
void PFX(ml_coarse_solve)(orc_ml* M, const float* r, float* e) {
    int l = M->nl - 1, V = M->L[l]->V;
    real* rr = (real*)xcalloc((size_t)V * 3, sizeof(real));
    real* ee = (real*)xcalloc((size_t)V * 3, sizeof(real));
    for (int i = 0; i < V * 3; i++) rr[i] = r[i];
    ml_solve_coarsest(M, rr, ee);
    for (int i = 0; i < V * 3; i++) e[i] = (float)ee[i];
    free(rr);
    free(ee);
}

/* ========================================================================= */
/* MLP of the learned coarse correction: in -> H -> H -> H -> out, ReLU.     */
/* FP64 accumulation; with bf16 != 0 every GEMM operand is rounded to BF16    */
/* (nearest even, via float) where the device kernels round it.  Parameter    */
/* layout: W1[in][H] b1[H] W2[H][H] b2 W3[H][H] b3 W4[H][out] b4[out].         */
/* ========================================================================= */
#define ORC_MLP_H 128

static double orc_bf(double v, int on) {
    if (!on) return v;
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return f;
}

typedef struct {
    int in, out, off[8];
} orc_mlp_shape;

static orc_mlp_shape orc_mlp_layout(int in, int out) {
    orc_mlp_shape s;
    s.in = in;
    s.out = out;
    int sizes[8] = {in * ORC_MLP_H, ORC_MLP_H, ORC_MLP_H * ORC_MLP_H, ORC_MLP_H, ORC_MLP_H * ORC_MLP_H, ORC_MLP_H,
                    ORC_MLP_H * out, out};
    int o = 0;
    for (int i = 0; i < 8; i++) {
        s.off[i] = o;
        o += sizes[i];
    }
    return s;
}

/* Z[n][N] = bf(A)[n][K] bf(W)[K][N] + b */
static void orc_dense(const double* A, int n, int K, const float* W, const float* b, int N, double* Z, int bf) {
    for (int r = 0; r < n; r++)
        for (int j = 0; j < N; j++) {
            double acc = 0;
            for (int k = 0; k < K; k++) acc += orc_bf(A[(size_t)r * K + k], bf) * orc_bf(W[(size_t)k * N + j], bf);
            Z[(size_t)r * N + j] = acc + b[j];
        }
}

void PFX(mlp_forward)(const float* w, int in, int out, const float* x, int n, double* y, int bf) {
    orc_mlp_shape s = orc_mlp_layout(in, out);
    int H = ORC_MLP_H;
    double* X = (double*)xcalloc((size_t)n * in, sizeof(double));
    double* A = (double*)xcalloc((size_t)n * H, sizeof(double));
    double* B = (double*)xcalloc((size_t)n * H, sizeof(double));
    for (size_t i = 0; i < (size_t)n * in; i++) X[i] = x[i];
    orc_dense(X, n, in, w + s.off[0], w + s.off[1], H, A, bf);
    for (size_t i = 0; i < (size_t)n * H; i++) A[i] = A[i] > 0 ? A[i] : 0;
    orc_dense(A, n, H, w + s.off[2], w + s.off[3], H, B, bf);
    for (size_t i = 0; i < (size_t)n * H; i++) B[i] = B[i] > 0 ? B[i] : 0;
    orc_dense(B, n, H, w + s.off[4], w + s.off[5], H, A, bf);
    for (size_t i = 0; i < (size_t)n * H; i++) A[i] = A[i] > 0 ? A[i] : 0;
    orc_dense(A, n, H, w + s.off[6], w + s.off[7], out, y, bf);
    free(X); free(A); free(B);
}

/* Mean squared error over n*out and its gradient w.r.t. the parameters. */
double PFX(mlp_grad)(const float* w, int in, int out, const float* x, const float* t, int n, double* grad, int bf) {
    orc_mlp_shape s = orc_mlp_layout(in, out);
    int H = ORC_MLP_H;
    size_t nh = (size_t)n * H;
    double* X = (double*)xcalloc((size_t)n * in, sizeof(double));
    double *Z1 = (double*)xcalloc(nh, 8), *H1 = (double*)xcalloc(nh, 8), *Z2 = (double*)xcalloc(nh, 8),
           *H2 = (double*)xcalloc(nh, 8), *Z3 = (double*)xcalloc(nh, 8), *H3 = (double*)xcalloc(nh, 8);
    double* Y = (double*)xcalloc((size_t)n * out, 8);
    double* dY = (double*)xcalloc((size_t)n * out, 8);
    double *dA = (double*)xcalloc(nh, 8), *dB = (double*)xcalloc(nh, 8);
    for (size_t i = 0; i < (size_t)n * in; i++) X[i] = x[i];
    orc_dense(X, n, in, w + s.off[0], w + s.off[1], H, Z1, bf);
    for (size_t i = 0; i < nh; i++) H1[i] = Z1[i] > 0 ? Z1[i] : 0;
    orc_dense(H1, n, H, w + s.off[2], w + s.off[3], H, Z2, bf);
    for (size_t i = 0; i < nh; i++) H2[i] = Z2[i] > 0 ? Z2[i] : 0;
    orc_dense(H2, n, H, w + s.off[4], w + s.off[5], H, Z3, bf);
    for (size_t i = 0; i < nh; i++) H3[i] = Z3[i] > 0 ? Z3[i] : 0;
    orc_dense(H3, n, H, w + s.off[6], w + s.off[7], out, Y, bf);
    double loss = 0;
    size_t ny = (size_t)n * out;
    for (size_t i = 0; i < ny; i++) {
        double d = Y[i] - t[i];
        loss += d * d;
        dY[i] = 2 * d / (double)ny;
    }
    loss /= (double)ny;
    /* one layer back: dW = bf(Hin)^T bf(dZ), db = sum dZ, dIn = bf(dZ) bf(W)^T */
#define ORC_BACK(Hin, K, dZ, N, Wo, bo, dIn)                                                            \
    do {                                                                                               \
        for (int k = 0; k < (K); k++)                                                                  \
            for (int j = 0; j < (N); j++) {                                                            \
                double a = 0;                                                                          \
                for (int r = 0; r < n; r++)                                                            \
                    a += orc_bf((Hin)[(size_t)r * (K) + k], bf) * orc_bf((dZ)[(size_t)r * (N) + j], bf); \
                grad[(Wo) + (size_t)k * (N) + j] = a;                                                  \
            }                                                                                          \
        for (int j = 0; j < (N); j++) {                                                                \
            double a = 0;                                                                              \
            for (int r = 0; r < n; r++) a += (dZ)[(size_t)r * (N) + j];                                \
            grad[(bo) + j] = a;                                                                        \
        }                                                                                              \
        if (dIn)                                                                                       \
            for (int r = 0; r < n; r++)                                                                \
                for (int k = 0; k < (K); k++) {                                                        \
                    double a = 0;                                                                      \
                    for (int j = 0; j < (N); j++)                                                      \
                        a += orc_bf((dZ)[(size_t)r * (N) + j], bf) * orc_bf(w[(Wo) + (size_t)k * (N) + j], bf); \
                    ((double*)(dIn))[(size_t)r * (K) + k] = a;                                         \
                }                                                                                      \
    } while (0)
    ORC_BACK(H3, H, dY, out, s.off[6], s.off[7], dA);
    for (size_t i = 0; i < nh; i++)
        if (!(Z3[i] > 0)) dA[i] = 0;
    ORC_BACK(H2, H, dA, H, s.off[4], s.off[5], dB);
    for (size_t i = 0; i < nh; i++)
        if (!(Z2[i] > 0)) dB[i] = 0;
    ORC_BACK(H1, H, dB, H, s.off[2], s.off[3], dA);
    for (size_t i = 0; i < nh; i++)
        if (!(Z1[i] > 0)) dA[i] = 0;
    ORC_BACK(X, in, dA, H, s.off[0], s.off[1], (double*)0);
#undef ORC_BACK
    free(X); free(Z1); free(H1); free(Z2); free(H2); free(Z3); free(H3); free(Y); free(dY); free(dA); free(dB);
    return loss;
}
