// Host side of libmgsb: gmsh .msh loading, surface extraction, structured
// box / voxel blob generation, rest-state precompute and the deterministic
// gather structures.  All of it runs once, at solver construction.
//
// Floating-point order in the precompute follows the reference's source text
// (PDSolver.cpp:12-132, utils.cu:233-364) so that device results can be
// compared with a CPU restatement bit for bit where the maths allows.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "mgsb_internal.h"

namespace mgsb {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

namespace {

template <class T>
T* dup_array(const std::vector<T>& v) {
    T* p = static_cast<T*>(std::malloc(sizeof(T) * (v.empty() ? 1 : v.size())));
    if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

struct V3 {
    float x, y, z;
};
inline V3 v3(const float* p) { return {p[0], p[1], p[2]}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float len(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// Node counts of gmsh 2.2 element types (gmsh file-format documentation).
int gmsh_nodes_of_type(int type) {
    switch (type) {
        case 1: return 2;   // line
        case 2: return 3;   // triangle
        case 3: return 4;   // quad
        case 4: return 4;   // tet
        case 5: return 8;   // hex
        case 6: return 6;   // prism
        case 7: return 5;   // pyramid
        case 8: return 3;   // 2nd-order line
        case 9: return 6;   // 2nd-order triangle
        case 10: return 9;  // 2nd-order quad
        case 11: return 10; // 2nd-order tet
        case 15: return 1;  // point
        default: return -1;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Surface extraction, SoftObject::TetFaceExtraction (SoftObject.cpp:154-251).
// A face is on the surface iff exactly one tet has it.  Windings per tet are
// (v0,v2,v1|v3), (v0,v1,v3|v2), (v0,v3,v2|v1), (v1,v2,v3|v0); the output
// order is the reference std::set's order, i.e. sorted keys descending.
// ---------------------------------------------------------------------------
static void extract_surface(int T, const int32_t* tets, std::vector<uint32_t>& tris, std::vector<uint32_t>& opp) {
    struct Face {
        int k[3];
        int f[3];
        int opp;
    };
    std::vector<Face> faces;
    faces.reserve((size_t)T * 4);
    for (int i = 0; i < T; i++) {
        int v0 = tets[i * 4 + 0], v1 = tets[i * 4 + 1], v2 = tets[i * 4 + 2], v3i = tets[i * 4 + 3];
        int fv[4][4] = {{v0, v2, v1, v3i}, {v0, v1, v3i, v2}, {v0, v3i, v2, v1}, {v1, v2, v3i, v0}};
        for (auto& q : fv) {
            Face F;
            F.f[0] = q[0]; F.f[1] = q[1]; F.f[2] = q[2];
            F.opp = q[3];
            int a = q[0], b = q[1], c = q[2];
            if (a > b) std::swap(a, b);
            if (b > c) std::swap(b, c);
            if (a > b) std::swap(a, b);
            F.k[0] = a; F.k[1] = b; F.k[2] = c;
            faces.push_back(F);
        }
    }
    std::vector<size_t> order(faces.size());
    std::iota(order.begin(), order.end(), 0);
    auto keyless = [&](size_t a, size_t b) {
        const Face &A = faces[a], &B = faces[b];
        if (A.k[0] != B.k[0]) return A.k[0] > B.k[0];
        if (A.k[1] != B.k[1]) return A.k[1] > B.k[1];
        if (A.k[2] != B.k[2]) return A.k[2] > B.k[2];
        return a < b;
    };
    std::sort(order.begin(), order.end(), keyless);
    tris.clear();
    opp.clear();
    size_t i = 0;
    while (i < order.size()) {
        size_t j = i + 1;
        const Face& A = faces[order[i]];
        while (j < order.size()) {
            const Face& B = faces[order[j]];
            if (B.k[0] != A.k[0] || B.k[1] != A.k[1] || B.k[2] != A.k[2]) break;
            j++;
        }
        if (j - i == 1) {
            tris.push_back((uint32_t)A.f[0]);
            tris.push_back((uint32_t)A.f[1]);
            tris.push_back((uint32_t)A.f[2]);
            opp.push_back((uint32_t)A.opp);
        }
        i = j;
    }
}

// ---------------------------------------------------------------------------
// Precompute, InitVolumeConstraint (PDSolver.cpp:12-132).
// ---------------------------------------------------------------------------
static float matrix_inverse3(const float* A, float* R) {  // utils.cu:233-252
    R[0] = A[4] * A[8] - A[7] * A[5];
    R[1] = A[7] * A[2] - A[1] * A[8];
    R[2] = A[1] * A[5] - A[4] * A[2];
    R[3] = A[5] * A[6] - A[3] * A[8];
    R[4] = A[0] * A[8] - A[2] * A[6];
    R[5] = A[2] * A[3] - A[0] * A[5];
    R[6] = A[3] * A[7] - A[4] * A[6];
    R[7] = A[1] * A[6] - A[0] * A[7];
    R[8] = A[0] * A[4] - A[1] * A[3];
    float det = A[0] * R[0] + A[3] * R[1] + A[6] * R[2];
    if (std::fabs(det) < 1e-6) return det;
    float inv_det = 1 / det;
    for (int i = 0; i < 9; i++) R[i] *= inv_det;
    return det;
}

static void build_incidence(int V, int n_items, int corners, const int32_t* items, std::vector<int32_t>& rp,
                            std::vector<int32_t>& slot) {
    rp.assign((size_t)V + 1, 0);
    for (int e = 0; e < n_items * corners; e++) rp[items[e] + 1]++;
    for (int i = 0; i < V; i++) rp[i + 1] += rp[i];
    std::vector<int32_t> fill((size_t)V, 0);
    slot.assign((size_t)n_items * corners, 0);
    for (int e = 0; e < n_items * corners; e++) {
        int v = items[e];
        slot[e] = rp[v] + fill[v]++;
    }
}

void build_tiles(HostLevel& L) {
    const int T = L.T, V = L.V, nt = (T + kTileTets - 1) / kTileTets;
    L.tvo.assign((size_t)nt + 1, 0);
    L.tlr.clear();
    L.tls.assign((size_t)T * 2, 0u);
    L.tlv.assign((size_t)T * 2, 0u);
    std::vector<int32_t> local((size_t)V, -1), cnt((size_t)V, 0), all, verts, fill;
    for (int k = 0; k < nt; k++) {
        const int t0 = k * kTileTets, t1 = std::min(T, t0 + kTileTets);
        verts.clear();
        for (int e = t0 * 4; e < t1 * 4; e++) {
            const int v = L.tets[(size_t)e];
            if (local[v] < 0) {
                local[v] = 0;
                verts.push_back(v);
            }
        }
        std::sort(verts.begin(), verts.end());
        const int nv = (int)verts.size();
        for (int i = 0; i < nv; i++) local[verts[i]] = i;
        fill.assign((size_t)nv + 1, 0);
        for (int e = t0 * 4; e < t1 * 4; e++) fill[(size_t)local[L.tets[(size_t)e]] + 1]++;
        for (int i = 0; i < nv; i++) fill[(size_t)i + 1] += fill[(size_t)i];
        L.tlr.insert(L.tlr.end(), fill.begin(), fill.end());
        for (int t = t0; t < t1; t++)
            for (int j = 0; j < 4; j++) {
                const int lv = local[L.tets[(size_t)t * 4 + j]];
                const uint32_t pos = (uint32_t)fill[(size_t)lv]++;
                L.tls[(size_t)t * 2 + (j >> 1)] |= pos << (16 * (j & 1));
                L.tlv[(size_t)t * 2 + (j >> 1)] |= (uint32_t)lv << (16 * (j & 1));
            }
        for (int v : verts) {
            cnt[v]++;
            all.push_back(v);
            local[v] = -1;
        }
        L.tvo[(size_t)k + 1] = L.tvo[(size_t)k] + nv;
    }
    L.prp.assign((size_t)V + 1, 0);
    for (int v = 0; v < V; v++) L.prp[(size_t)v + 1] = L.prp[(size_t)v] + cnt[v];
    std::vector<int32_t> at(L.prp.begin(), L.prp.end() - 1);
    L.tvs.resize(all.size());
    for (size_t q = 0; q < all.size(); q++) L.tvs[q] = at[all[q]]++;  // tiles ascending in each row
    L.tvid = all;
}

int build_level(const mgsb_mesh& m, float stiffness, float density, HostLevel& L) {
    if (m.nverts <= 0 || m.ntets <= 0 || !m.pos || !m.tets) {
        set_error("build_level: empty mesh");
        return MGSB_ERR_ARG;
    }
    L.V = m.nverts;
    L.T = m.ntets;
    L.X.assign(m.pos, m.pos + (size_t)L.V * 3);
    L.tets.assign(m.tets, m.tets + (size_t)L.T * 4);
    for (int e = 0; e < L.T * 4; e++) {
        if (L.tets[e] < 0 || L.tets[e] >= L.V) {
            set_error("build_level: tet " + std::to_string(e / 4) + " has vertex id out of range");
            return MGSB_ERR_MESH;
        }
    }
    if (m.ntris > 0 && m.tris) {
        L.S = m.ntris;
        L.tris.assign(m.tris, m.tris + (size_t)L.S * 3);
        for (int e = 0; e < L.S * 3; e++)
            if ((int64_t)L.tris[e] >= L.V) {
                set_error("build_level: surface triangle vertex id out of range");
                return MGSB_ERR_MESH;
            }
    } else {
        std::vector<uint32_t> opp;
        extract_surface(L.T, L.tets.data(), L.tris, opp);
        L.S = (int)L.tris.size() / 3;
    }
    int T = L.T, V = L.V;
    L.B.assign((size_t)T * 9, 0.f);
    L.AC.assign((size_t)T * 12, 0.f);
    L.vol.assign(T, 0.f);
    L.mass.assign(V, 0.f);
    L.diag.assign(V, 0.f);
    L.center.assign((size_t)T * 3, 0.f);
    L.faceN.assign((size_t)T * 12, 0.f);
    L.faceA.assign((size_t)T * 4, 0.f);
    L.fix.assign(V, 1.f);
    L.bad_tet = -1;
    const float k = stiffness;
    const float* X = L.X.data();
    for (int i = 0; i < T; i++) {
        const int v0 = L.tets[i * 4 + 0], v1 = L.tets[i * 4 + 1], v2 = L.tets[i * 4 + 2], v3i = L.tets[i * 4 + 3];
        float D[9];
        D[0] = X[v1 * 3 + 0] - X[v0 * 3 + 0];
        D[1] = X[v2 * 3 + 0] - X[v0 * 3 + 0];
        D[2] = X[v3i * 3 + 0] - X[v0 * 3 + 0];
        D[3] = X[v1 * 3 + 1] - X[v0 * 3 + 1];
        D[4] = X[v2 * 3 + 1] - X[v0 * 3 + 1];
        D[5] = X[v3i * 3 + 1] - X[v0 * 3 + 1];
        D[6] = X[v1 * 3 + 2] - X[v0 * 3 + 2];
        D[7] = X[v2 * 3 + 2] - X[v0 * 3 + 2];
        D[8] = X[v3i * 3 + 2] - X[v0 * 3 + 2];
        float det = matrix_inverse3(D, &L.B[(size_t)i * 9]);
        if (std::fabs(det) < 1e-6 && L.bad_tet < 0) L.bad_tet = i;
        L.vol[i] = std::fabs(det) / 6;
        float vq = L.vol[i] / 4;
        L.mass[v0] += vq;
        L.mass[v1] += vq;
        L.mass[v2] += vq;
        L.mass[v3i] += vq;
        const float* b = &L.B[(size_t)i * 9];
        float* ac = &L.AC[(size_t)i * 12];
        ac[0] = -b[0] - b[3] - b[6]; ac[1] = b[0]; ac[2] = b[3]; ac[3] = b[6];
        ac[4] = -b[1] - b[4] - b[7]; ac[5] = b[1]; ac[6] = b[4]; ac[7] = b[7];
        ac[8] = -b[2] - b[5] - b[8]; ac[9] = b[2]; ac[10] = b[5]; ac[11] = b[8];
        const int vs[4] = {v0, v1, v2, v3i};
        for (int j = 0; j < 4; j++)
            for (int r = 0; r < 3; r++) L.diag[vs[j]] += ac[r * 4 + j] * ac[r * 4 + j] * L.vol[i] * k;
        for (int c = 0; c < 3; c++) {
            float acc = 0;
            for (int j = 0; j < 4; j++) acc += X[vs[j] * 3 + c];
            L.center[(size_t)i * 3 + c] = acc / 4;
        }
        const int faces[4][3] = {{v1, v2, v3i}, {v0, v2, v3i}, {v0, v1, v3i}, {v0, v1, v2}};
        for (int f = 0; f < 4; f++) {
            V3 A = v3(X + faces[f][0] * 3), Bq = v3(X + faces[f][1] * 3), Cq = v3(X + faces[f][2] * 3);
            V3 Dq = v3(X + vs[f] * 3);
            V3 n = cross(Bq - A, Cq - A);
            if (dot(n, Dq - A) > 0) n = {-n.x, -n.y, -n.z};
            float l = len(n);
            L.faceA[(size_t)i * 4 + f] = 0.5f * l;
            n = n / l;
            L.faceN[(size_t)i * 12 + f * 3 + 0] = n.x;
            L.faceN[(size_t)i * 12 + f * 3 + 1] = n.y;
            L.faceN[(size_t)i * 12 + f * 3 + 2] = n.z;
        }
    }
    if (density != 1.f)
        for (int i = 0; i < V; i++) L.mass[i] *= density;
    if (L.bad_tet >= 0) {
        set_error("build_level: degenerate rest tet " + std::to_string(L.bad_tet) +
                  " (|det Dm| < 1e-6; the reference exits here, utils.cu:244-248)");
        return MGSB_ERR_MESH;
    }
    build_incidence(V, T, 4, L.tets.data(), L.rp, L.slot);
    build_incidence(V, L.S, 3, reinterpret_cast<const int32_t*>(L.tris.data()), L.trp, L.tslot);
    build_tiles(L);
    return MGSB_OK;
}

void recompute_diag(HostLevel& L, float k) {
    L.diag.assign(L.V, 0.f);
    for (int i = 0; i < L.T; i++) {
        const float kt = L.tk.empty() ? k : L.tk[(size_t)i];
        const float* ac = &L.AC[(size_t)i * 12];
        for (int j = 0; j < 4; j++)
            for (int r = 0; r < 3; r++)
                L.diag[L.tets[(size_t)i * 4 + j]] += ac[r * 4 + j] * ac[r * 4 + j] * L.vol[i] * kt;
    }
}

int mesh_components(int V, int T, const int32_t* tets, std::vector<int32_t>& comp) {
    std::vector<int32_t> par(V);
    for (int i = 0; i < V; i++) par[i] = i;
    auto find = [&](int a) {
        while (par[a] != a) a = par[a] = par[par[a]];
        return a;
    };
    for (int t = 0; t < T; t++)
        for (int c = 1; c < 4; c++) {
            int a = find(tets[(size_t)t * 4]), b = find(tets[(size_t)t * 4 + c]);
            if (a != b) par[std::max(a, b)] = std::min(a, b);  // root = smallest vertex
        }
    comp.assign(V, -1);
    std::vector<int32_t> id(V, -1);
    int n = 0;
    for (int i = 0; i < V; i++) {
        int r = find(i);
        if (id[r] < 0) id[r] = n++;
        comp[i] = id[r];
    }
    return n;
}

// ---------------------------------------------------------------------------
// Barycentric bindings between levels (PDSolver_MG.cpp:5-58, 133-168).
// ---------------------------------------------------------------------------
void barycentric4(const float* P, const float* A, const float* B, const float* C, const float* D, float* w) {
    // GetVolumn (utils.cu:254-275)
    V3 a = v3(A), b = v3(B), c = v3(C), d = v3(D);
    V3 AB = b - a, ACv = c - a, AD = d - a;
    float M[9] = {AB.x, ACv.x, AD.x, AB.y, ACv.y, AD.y, AB.z, ACv.z, AD.z};
    float R0 = M[4] * M[8] - M[7] * M[5];
    float R1 = M[7] * M[2] - M[1] * M[8];
    float R2 = M[1] * M[5] - M[4] * M[2];
    float Vol = (M[0] * R0 + M[3] * R1 + M[6] * R2) / 6.0f;
    V3 center = (((a + b) + c) + d) * 0.25f;
    V3 tv[4] = {a, b, c, d};
    static const int pf[4][4] = {{0, 1, 2, 3}, {1, 2, 3, 0}, {2, 3, 0, 1}, {3, 0, 1, 2}};
    V3 p = v3(P);
    for (int i = 0; i < 4; i++) {
        V3 q = tv[pf[i][0]], f0 = tv[pf[i][1]], f1 = tv[pf[i][2]], f2 = tv[pf[i][3]];
        V3 n = cross(f1 - f0, f2 - f0);
        if (dot(n, q - f0) > 0) n = {-n.x, -n.y, -n.z};
        float l = len(n);
        float area = l * 0.5f;
        n = n / l;
        w[i] = 0.25f - dot(p - center, n) * area / (3 * Vol);
    }
}

namespace {

// Uniform grid over tet bounding boxes; candidates come back in ascending tet
// index so "first accepted in index order" matches a brute-force scan.
struct TetGrid {
    float lo[3], cell[3];
    int n[3];
    std::vector<int32_t> start, items;

    void build(const HostLevel& L, float inflate_rel) {
        float hi[3];
        for (int c = 0; c < 3; c++) {
            lo[c] = 1e30f;
            hi[c] = -1e30f;
        }
        for (int v = 0; v < L.V; v++)
            for (int c = 0; c < 3; c++) {
                lo[c] = std::min(lo[c], L.X[v * 3 + c]);
                hi[c] = std::max(hi[c], L.X[v * 3 + c]);
            }
        double ext[3], diagl = 0;
        for (int c = 0; c < 3; c++) {
            ext[c] = std::max(1e-6, (double)hi[c] - lo[c]);
            diagl += ext[c] * ext[c];
        }
        diagl = std::sqrt(diagl);
        float pad = (float)(diagl * inflate_rel);
        for (int c = 0; c < 3; c++) {
            lo[c] -= 2 * pad;
            hi[c] += 2 * pad;
        }
        // about 2 tets per cell
        double volb = 1;
        for (int c = 0; c < 3; c++) volb *= ((double)hi[c] - lo[c]);
        double target = std::max(1.0, (double)L.T / 2.0);
        double h = std::cbrt(volb / target);
        for (int c = 0; c < 3; c++) {
            n[c] = std::max(1, std::min(1024, (int)std::ceil(((double)hi[c] - lo[c]) / h)));
            cell[c] = (hi[c] - lo[c]) / n[c];
        }
        size_t ncell = (size_t)n[0] * n[1] * n[2];
        std::vector<int32_t> cnt(ncell + 1, 0);
        auto range = [&](int t, int* a, int* b) {
            float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
            for (int j = 0; j < 4; j++) {
                int v = L.tets[t * 4 + j];
                for (int c = 0; c < 3; c++) {
                    mn[c] = std::min(mn[c], L.X[v * 3 + c]);
                    mx[c] = std::max(mx[c], L.X[v * 3 + c]);
                }
            }
            for (int c = 0; c < 3; c++) {
                float ex = (mx[c] - mn[c]) * inflate_rel * 10 + pad;
                a[c] = std::max(0, std::min(n[c] - 1, (int)std::floor((mn[c] - ex - lo[c]) / cell[c])));
                b[c] = std::max(0, std::min(n[c] - 1, (int)std::floor((mx[c] + ex - lo[c]) / cell[c])));
            }
        };
        for (int t = 0; t < L.T; t++) {
            int a[3], b[3];
            range(t, a, b);
            for (int z = a[2]; z <= b[2]; z++)
                for (int y = a[1]; y <= b[1]; y++)
                    for (int x = a[0]; x <= b[0]; x++) cnt[((size_t)z * n[1] + y) * n[0] + x + 1]++;
        }
        for (size_t i = 0; i < ncell; i++) cnt[i + 1] += cnt[i];
        start = cnt;
        items.assign(cnt[ncell], 0);
        std::vector<int32_t> fill(ncell, 0);
        for (int t = 0; t < L.T; t++) {  // ascending t within each cell
            int a[3], b[3];
            range(t, a, b);
            for (int z = a[2]; z <= b[2]; z++)
                for (int y = a[1]; y <= b[1]; y++)
                    for (int x = a[0]; x <= b[0]; x++) {
                        size_t ci = ((size_t)z * n[1] + y) * n[0] + x;
                        items[start[ci] + fill[ci]++] = t;
                    }
        }
    }
    // returns cell index or -1 if outside the grid
    long cell_of(const float* p) const {
        int q[3];
        for (int c = 0; c < 3; c++) {
            float f = (p[c] - lo[c]) / cell[c];
            if (!(f >= 0) || f >= n[c] + 1e-3f) return -1;
            q[c] = std::min(n[c] - 1, (int)std::floor(f));
        }
        return ((long)q[2] * n[1] + q[1]) * n[0] + q[0];
    }
};

}  // namespace

int bind_points_interp(const HostLevel& mesh, const float* pts, int n, std::vector<int32_t>& ids,
                       std::vector<float>& w) {
    TetGrid g;
    g.build(mesh, 1e-4f);
    ids.assign((size_t)n * 4, 0);
    w.assign((size_t)n * 4, 0.f);
    int unfound = 0;
    for (int i = 0; i < n; i++) {
        const float* P = pts + (size_t)i * 3;
        long ci = g.cell_of(P);
        bool found = false;
        if (ci >= 0) {
            for (int e = g.start[ci]; e < g.start[ci + 1] && !found; e++) {
                int t = g.items[e];
                const int32_t* id = &mesh.tets[(size_t)t * 4];
                float ww[4];
                barycentric4(P, &mesh.X[id[0] * 3], &mesh.X[id[1] * 3], &mesh.X[id[2] * 3], &mesh.X[id[3] * 3], ww);
                bool in = true;
                for (int q = 0; q < 4; q++)
                    if (ww[q] < -1e-5 || ww[q] > 1 + 1e-5) in = false;
                if (in) {
                    for (int q = 0; q < 4; q++) {
                        ids[(size_t)i * 4 + q] = id[q];
                        w[(size_t)i * 4 + q] = ww[q];
                    }
                    found = true;
                }
            }
        }
        if (!found) unfound++;
    }
    return unfound;
}

int bind_points_average(const HostLevel& mesh, const float* pts, int n, std::vector<int32_t>& ids,
                        std::vector<float>& w) {
    TetGrid g;
    g.build(mesh, 1e-4f);
    ids.assign((size_t)n * 4, 0);
    w.assign((size_t)n * 4, 0.f);
    int unfound = 0;
    for (int i = 0; i < n; i++) {
        const float* P = pts + (size_t)i * 3;
        V3 p = v3(P);
        long ci = g.cell_of(P);
        bool found = false;
        if (ci >= 0) {
            for (int e = g.start[ci]; e < g.start[ci + 1] && !found; e++) {
                int t = g.items[e];
                const int32_t* id = &mesh.tets[(size_t)t * 4];
                bool in = true;
                for (int k = 0; k < 4 && in; k++) {  // pointInTet, utils.cu:298-311
                    int ia = (k + 1) % 4;
                    V3 A = v3(&mesh.X[id[ia] * 3]);
                    V3 nrm = v3(&mesh.faceN[(size_t)t * 12 + k * 3]);
                    if (dot(p - A, nrm) > 1e-6) in = false;
                }
                if (!in) continue;
                V3 rB = v3(&mesh.center[(size_t)t * 3]);
                float Vt = mesh.vol[t];
                for (int q = 0; q < 4; q++) {  // utils.cu:313-333
                    V3 nrm = v3(&mesh.faceN[(size_t)t * 12 + q * 3]);
                    ids[(size_t)i * 4 + q] = id[q];
                    w[(size_t)i * 4 + q] = 0.25f - dot(p - rB, nrm) * mesh.faceA[(size_t)t * 4 + q] / (3 * Vt);
                }
                found = true;
            }
        }
        if (!found) unfound++;
    }
    return unfound;
}

}  // namespace mgsb

using namespace mgsb;

extern "C" {

const char* mgsb_last_error(void) { return g_err.c_str(); }

void mgsb_mesh_free(mgsb_mesh* m) {
    if (!m) return;
    std::free(m->pos);
    std::free(m->tets);
    std::free(m->tris);
    std::free(m->tri_opp);
    std::memset(m, 0, sizeof(*m));
}

int mgsb_extract_surface(mgsb_mesh* m) {
    if (!m || m->ntets <= 0 || !m->tets) {
        set_error("mgsb_extract_surface: mesh has no tets");
        return MGSB_ERR_ARG;
    }
    std::vector<uint32_t> tris, opp;
    extract_surface(m->ntets, m->tets, tris, opp);
    std::free(m->tris);
    std::free(m->tri_opp);
    m->tris = dup_array(tris);
    m->tri_opp = dup_array(opp);
    m->ntris = (int32_t)opp.size();
    return MGSB_OK;
}

int mgsb_load_msh(const char* path, mgsb_mesh* out) {
    if (!path || !out) {
        set_error("mgsb_load_msh: null argument");
        return MGSB_ERR_ARG;
    }
    std::memset(out, 0, sizeof(*out));
    std::ifstream f(path);
    if (!f) {
        set_error(std::string("mgsb_load_msh: cannot open ") + path);
        return MGSB_ERR_IO;
    }
    std::vector<float> pos;
    std::vector<int32_t> tets;
    std::unordered_map<long, int32_t> id_of;
    bool sequential = true;
    std::string tok;
    bool have_nodes = false, have_elems = false;
    while (f >> tok) {
        if (tok == "$Nodes") {
            long n;
            if (!(f >> n) || n < 0) {
                set_error("mgsb_load_msh: bad $Nodes count");
                return MGSB_ERR_MESH;
            }
            pos.reserve((size_t)n * 3);
            for (long i = 0; i < n; i++) {
                long idx;
                double x, y, z;
                if (!(f >> idx >> x >> y >> z)) {
                    set_error("mgsb_load_msh: truncated $Nodes");
                    return MGSB_ERR_MESH;
                }
                if (idx != i + 1) sequential = false;
                id_of[idx] = (int32_t)i;
                pos.push_back((float)x);
                pos.push_back((float)y);
                pos.push_back((float)z);
            }
            have_nodes = true;
        } else if (tok == "$Elements") {
            long n;
            if (!(f >> n) || n < 0) {
                set_error("mgsb_load_msh: bad $Elements count");
                return MGSB_ERR_MESH;
            }
            for (long i = 0; i < n; i++) {
                long idx;
                int type, ntags;
                if (!(f >> idx >> type >> ntags) || ntags < 0) {
                    set_error("mgsb_load_msh: truncated $Elements");
                    return MGSB_ERR_MESH;
                }
                for (int q = 0; q < ntags; q++) {
                    long tg;
                    f >> tg;
                }
                int nn = gmsh_nodes_of_type(type);
                if (nn < 0) {
                    set_error("mgsb_load_msh: unsupported element type " + std::to_string(type));
                    return MGSB_ERR_MESH;
                }
                long nodes[10];
                for (int q = 0; q < nn; q++) f >> nodes[q];
                if (!f) {
                    set_error("mgsb_load_msh: truncated element list");
                    return MGSB_ERR_MESH;
                }
                if (type == 4) {
                    for (int q = 0; q < 4; q++) tets.push_back((int32_t)nodes[q]);
                }
            }
            have_elems = true;
        }
    }
    if (!have_nodes || !have_elems) {
        set_error("mgsb_load_msh: missing $Nodes or $Elements section");
        return MGSB_ERR_MESH;
    }
    for (auto& t : tets) {
        if (sequential) {
            t = t - 1;  // SoftObject.cpp:120-123
        } else {
            auto it = id_of.find(t);
            if (it == id_of.end()) {
                set_error("mgsb_load_msh: element references unknown node");
                return MGSB_ERR_MESH;
            }
            t = it->second;
        }
        if (t < 0 || t >= (int32_t)(pos.size() / 3)) {
            set_error("mgsb_load_msh: node id out of range");
            return MGSB_ERR_MESH;
        }
    }
    out->nverts = (int32_t)(pos.size() / 3);
    out->ntets = (int32_t)(tets.size() / 4);
    out->pos = dup_array(pos);
    out->tets = dup_array(tets);
    if (out->ntets == 0) {
        set_error("mgsb_load_msh: no tetrahedra (type 4) in file");
        mgsb_mesh_free(out);
        return MGSB_ERR_MESH;
    }
    return mgsb_extract_surface(out);
}

// Kuhn (Freudenthal) split of each hex cell into 6 tets along the diagonal
// (0,0,0)-(1,1,1): for each permutation (a,b,c) of the axes the path
// 000 -> +e_a -> +e_a+e_b -> 111.  Conforming across cells; odd permutations
// swap corners 1 and 2 so every tet has positive orientation.
int mgsb_make_box(int32_t nx, int32_t ny, int32_t nz, const float lo[3], const float hi[3], mgsb_mesh* out) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || !lo || !hi || !out) {
        set_error("mgsb_make_box: bad arguments");
        return MGSB_ERR_ARG;
    }
    std::memset(out, 0, sizeof(*out));
    const long NX = nx + 1, NY = ny + 1, NZ = nz + 1;
    const long V = NX * NY * NZ;
    const long T = 6L * nx * ny * nz;
    if (V > (1L << 30) || T > (1L << 29)) {
        set_error("mgsb_make_box: mesh too large");
        return MGSB_ERR_ARG;
    }
    std::vector<float> pos((size_t)V * 3);
    auto vid = [&](long i, long j, long k) { return (int32_t)((j * NZ + k) * NX + i); };  // y slowest
    for (long j = 0; j < NY; j++)
        for (long k = 0; k < NZ; k++)
            for (long i = 0; i < NX; i++) {
                int32_t v = vid(i, j, k);
                pos[(size_t)v * 3 + 0] = lo[0] + (hi[0] - lo[0]) * (float)i / (float)nx;
                pos[(size_t)v * 3 + 1] = lo[1] + (hi[1] - lo[1]) * (float)j / (float)ny;
                pos[(size_t)v * 3 + 2] = lo[2] + (hi[2] - lo[2]) * (float)k / (float)nz;
            }
    static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    static const int odd[6] = {0, 1, 1, 0, 0, 1};
    std::vector<int32_t> tets;
    tets.reserve((size_t)T * 4);
    for (long j = 0; j < ny; j++)
        for (long k = 0; k < nz; k++)
            for (long i = 0; i < nx; i++) {
                for (int p = 0; p < 6; p++) {
                    long c[3] = {0, 0, 0};
                    int32_t q[4];
                    q[0] = vid(i, j, k);
                    for (int s = 0; s < 3; s++) {
                        c[perms[p][s]] = 1;
                        q[s + 1] = vid(i + c[0], j + c[1], k + c[2]);
                    }
                    if (odd[p]) std::swap(q[1], q[2]);
                    for (int s = 0; s < 4; s++) tets.push_back(q[s]);
                }
            }
    out->nverts = (int32_t)V;
    out->ntets = (int32_t)T;
    out->pos = dup_array(pos);
    out->tets = dup_array(tets);
    return mgsb_extract_surface(out);
}

// Voxelised superquadric blob.  A voxel (integer cell (i,j,k), size h, cell
// centre at ((i+0.5)h, ...)) is kept if its centre satisfies
//   |x/a|^p + |y/b|^p + |z/c|^p <= r(dir)^p,
//   r(dir) = 1 + amp * sum_{m<8} sin(f_m . dir + phase_m) / 8,
// with f_m, phase_m from a seeded PCG32 stream.  Cell coordinates are integer
// multiples of h from the origin, so blobs at h, 3h, 9h are nested lattices.
// Vertices are deduplicated lattice points; 6 Kuhn tets per voxel.
int mgsb_make_blob(float a, float b, float c, float p, float amp, uint32_t seed, float h, mgsb_mesh* out) {
    return mgsb_make_blob_level(a, b, c, p, amp, seed, h, 1, out);
}

// Level of a nested hierarchy: cells of size coarsen * h on the same lattice
// origin; a cell is kept iff any of its coarsen^3 cells of size h is inside
// the shape, so each level covers every finer one (its vertices lie in the
// coarser level's tets, as the V-cycle's prolongation needs).
int mgsb_make_blob_level(float a, float b, float c, float p, float amp, uint32_t seed, float h, int32_t coarsen,
                         mgsb_mesh* out) {
    if (!(a > 0 && b > 0 && c > 0 && p > 0 && h > 0) || !out || coarsen < 1 || coarsen > 64) {
        set_error("mgsb_make_blob: bad arguments");
        return MGSB_ERR_ARG;
    }
    std::memset(out, 0, sizeof(*out));
    // PCG32 (O'Neill 2014), XSH-RR output.
    uint64_t state = 0x853c49e6748fea9bULL ^ ((uint64_t)seed << 1), inc = ((uint64_t)seed << 1) | 1u;
    auto pcg = [&]() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    };
    auto unif = [&]() { return (pcg() >> 8) * (1.0 / 16777216.0); };
    double fr[8][3], ph[8];
    for (int m = 0; m < 8; m++) {
        for (int q = 0; q < 3; q++) fr[m][q] = (unif() * 2 - 1) * 3.0;
        ph[m] = unif() * 6.283185307179586;
    }
    const double rmax = 1 + std::fabs(amp);
    const int ni = (int)std::ceil(a * rmax / h) + 1, nj = (int)std::ceil(b * rmax / h) + 1,
              nk = (int)std::ceil(c * rmax / h) + 1;
    auto inside_fine = [&](int i, int j, int k) {
        double x = (i + 0.5) * h, y = (j + 0.5) * h, z = (k + 0.5) * h;
        double rr = std::sqrt(x * x + y * y + z * z);
        double d[3] = {0, 0, 0};
        if (rr > 0) {
            d[0] = x / rr; d[1] = y / rr; d[2] = z / rr;
        }
        double s = 0;
        for (int m = 0; m < 8; m++) s += std::sin(fr[m][0] * d[0] + fr[m][1] * d[1] + fr[m][2] * d[2] + ph[m]);
        double r = 1 + amp * s / 8;
        double lhs = std::pow(std::fabs(x / a), p) + std::pow(std::fabs(y / b), p) + std::pow(std::fabs(z / c), p);
        return lhs <= std::pow(r, (double)p);
    };
    // coarse lattice: cell I covers fine cells coarsen*I .. coarsen*I + coarsen - 1
    const int C = coarsen;
    const int ci = (ni + C - 1) / C + 1, cj = (nj + C - 1) / C + 1, ck = (nk + C - 1) / C + 1;
    auto inside = [&](int I, int J, int K) {
        for (int dj = 0; dj < C; dj++)
            for (int dk = 0; dk < C; dk++)
                for (int di = 0; di < C; di++)
                    if (inside_fine(I * C + di, J * C + dj, K * C + dk)) return true;
        return false;
    };
    const double hc = (double)h * C;
    const long W = 2L * ci + 1, H = 2L * cj + 1;  // lattice points from -ci..ci
    std::unordered_map<long, int32_t> vmap;
    std::vector<float> pos;
    std::vector<int32_t> tets;
    auto vert = [&](int i, int j, int k) {
        long key = ((long)(k + ck) * H + (j + cj)) * W + (i + ci);
        auto it = vmap.find(key);
        if (it != vmap.end()) return it->second;
        int32_t id = (int32_t)(pos.size() / 3);
        vmap.emplace(key, id);
        pos.push_back((float)(i * hc));
        pos.push_back((float)(j * hc));
        pos.push_back((float)(k * hc));
        return id;
    };
    static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    static const int odd[6] = {0, 1, 1, 0, 0, 1};
    for (int j = -cj; j < cj; j++)
        for (int k = -ck; k < ck; k++)
            for (int i = -ci; i < ci; i++) {
                if (!inside(i, j, k)) continue;
                for (int pq = 0; pq < 6; pq++) {
                    int cc[3] = {0, 0, 0};
                    int32_t q[4];
                    q[0] = vert(i, j, k);
                    for (int s = 0; s < 3; s++) {
                        cc[perms[pq][s]] = 1;
                        q[s + 1] = vert(i + cc[0], j + cc[1], k + cc[2]);
                    }
                    if (odd[pq]) std::swap(q[1], q[2]);
                    for (int s = 0; s < 4; s++) tets.push_back(q[s]);
                }
            }
    if (tets.empty()) {
        set_error("mgsb_make_blob: no voxel inside the shape");
        return MGSB_ERR_ARG;
    }
    out->nverts = (int32_t)(pos.size() / 3);
    out->ntets = (int32_t)(tets.size() / 4);
    out->pos = dup_array(pos);
    out->tets = dup_array(tets);
    return mgsb_extract_surface(out);
}

int mgsb_precompute(const mgsb_mesh* m, float stiffness, float density, float* B, float* AC, float* vol,
                    float* mass, float* diag) {
    if (!m) {
        set_error("mgsb_precompute: null mesh");
        return MGSB_ERR_ARG;
    }
    try {
        HostLevel L;
        int rc = build_level(*m, stiffness, density, L);
        if (rc) return rc;
        if (B) std::copy(L.B.begin(), L.B.end(), B);
        if (AC) std::copy(L.AC.begin(), L.AC.end(), AC);
        if (vol) std::copy(L.vol.begin(), L.vol.end(), vol);
        if (mass) std::copy(L.mass.begin(), L.mass.end(), mass);
        if (diag) std::copy(L.diag.begin(), L.diag.end(), diag);
        return MGSB_OK;
    } catch (const std::exception& e) {
        set_error(std::string("mgsb_precompute: ") + e.what());
        return MGSB_ERR_NOMEM;
    }
}

int mgsb_mesh_tiles(const mgsb_mesh* m, int32_t sizes[2], int32_t* prp, int32_t* tvo, int32_t* tvs, int32_t* tlr,
                    uint32_t* tls) {
    if (!m || !sizes) {
        set_error("mgsb_mesh_tiles: null mesh or sizes");
        return MGSB_ERR_ARG;
    }
    try {
        HostLevel L;
        int rc = build_level(*m, 1.f, 1.f, L);
        if (rc) return rc;
        sizes[0] = (int32_t)L.tvo.size() - 1;
        sizes[1] = L.tvo.back();
        if (prp) std::copy(L.prp.begin(), L.prp.end(), prp);
        if (tvo) std::copy(L.tvo.begin(), L.tvo.end(), tvo);
        if (tvs) std::copy(L.tvs.begin(), L.tvs.end(), tvs);
        if (tlr) std::copy(L.tlr.begin(), L.tlr.end(), tlr);
        if (tls) std::copy(L.tls.begin(), L.tls.end(), tls);
        return MGSB_OK;
    } catch (const std::exception& e) {
        set_error(std::string("mgsb_mesh_tiles: ") + e.what());
        return MGSB_ERR_NOMEM;
    }
}

int mgsb_bind_points(const mgsb_mesh* m, const float* pts, int32_t n, int32_t mode, int32_t* ids, float* w) {
    if (!m || (n > 0 && (!pts || !ids || !w)) || n < 0 || (mode != 0 && mode != 1)) {
        set_error("mgsb_bind_points: bad arguments");
        return MGSB_ERR_ARG;
    }
    try {
        HostLevel L;
        int rc = build_level(*m, 1.f, 1.f, L);
        if (rc) return rc;
        std::vector<int32_t> vi;
        std::vector<float> vw;
        int unf = mode == 0 ? bind_points_interp(L, pts, n, vi, vw) : bind_points_average(L, pts, n, vi, vw);
        std::copy(vi.begin(), vi.end(), ids);
        std::copy(vw.begin(), vw.end(), w);
        return unf;
    } catch (const std::exception& e) {
        set_error(std::string("mgsb_bind_points: ") + e.what());
        return MGSB_ERR_NOMEM;
    }
}

void mgsb_default_params(mgsb_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->dt = 1.0f / 60.0f;
    p->damping = 0.95f;
    p->stiffness = 8000.0f;
    p->density = 1.0f;
    p->rho = 0.9992f;
    p->relax = 0.6f;
    p->gravity[0] = 0.0f;
    p->gravity[1] = -9.8f;
    p->gravity[2] = 0.0f;
    p->kcol = 1000.0f;
    p->iters = 20;
    p->iters_coarse = 20;
    p->schedule = MGSB_SCHED_SINGLE;
    p->if_mode = MGSB_IF_AC;
    p->tri_dcd_coarse = 1;
    p->pre_sweeps = 5;
    p->post_sweeps = 5;
    p->coarse_sweeps = 2;
    p->max_cycles = 8;
    p->tol = 1e-3f;
    p->check_every = 0;
    p->use_graph = 0;
    p->block_tet = 0;
    p->block_vert = 0;
    p->device = 0;
    p->rho_smooth = 0.9f;
    p->relax_coarse = 0.6f;
    p->rho_coarse = 0.0f;
    p->coarse_direct = 1;
    p->warm_start = 0;
    p->learn_gate = 1;
    p->vert_lanes = 0;
    p->learn_slack = 1.0f;
    p->coarse_dense = 1;
    p->tile_gather = 1;
    p->coarse_first = 0;
    p->fuse_sweeps = 1;
    p->stall_guard = 1;
}

const char* mgsb_version(void) { return "mgsb 0.1 (gfx950)"; }

}  // extern "C"
