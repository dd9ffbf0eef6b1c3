// Internal host-side structures of libmgsb shared by the mesh/precompute code
// (mgsb_host.cpp) and the device solver (mgsb_solver.hip).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "mgsb.h"

namespace mgsb {

void set_error(const std::string& msg);

// One tet mesh level after InitVolumeConstraint (PDSolver.cpp:12-132), plus
// the deterministic gather structures.
struct HostLevel {
    int V = 0, T = 0, S = 0;
    std::vector<float> X;         // rest positions, 3V
    std::vector<int32_t> tets;    // 4T
    std::vector<uint32_t> tris;   // 3S
    std::vector<float> B;         // Dm^-1, 9T row-major per tet
    std::vector<float> AC;        // 3x4 gradient matrix, 12T
    std::vector<float> vol;       // T
    std::vector<float> mass;      // V (density applied)
    std::vector<float> diag;      // V, Jacobi diagonal sum k vol AC^2
    std::vector<float> center;    // 3T
    std::vector<float> faceN;     // 12T, outward unit normal of the face opposite corner f
    std::vector<float> faceA;     // 4T
    std::vector<float> fix;       // V, 1 free / 0 pinned
    // vertex -> incident (tet, corner), tets ascending.  slot[t*4+c] is the
    // position of (t,c) in that order.
    std::vector<int32_t> rp;      // V+1
    std::vector<int32_t> slot;    // 4T
    // vertex -> incident (surface tri, corner), tris ascending.
    std::vector<int32_t> trp;     // V+1
    std::vector<int32_t> tslot;   // 3S
    // Tiles of kTileTets consecutive tets, one workgroup each: the tile sums
    // its corner forces per vertex and writes one partial per (tile, vertex).
    // prp: vertex -> its partials, tiles ascending.  Tile k's vertices
    // (ascending ids) are tvo[k] .. tvo[k+1]-1 of tvs (the position of each
    // one's partial in prp order); tlr holds each tile's nv+1 row starts into
    // its corner array (at tvo[k] + k), and tls the position of corner j of
    // tet t in that array (16 bits each, corners 0,1 in tls[2t], 2,3 in
    // tls[2t+1]); a row lists its corners in (tet, corner) order.
    std::vector<int32_t> prp;     // V+1
    std::vector<int32_t> tvo;     // ntiles+1
    std::vector<int32_t> tvs;     // tvo[ntiles]
    std::vector<int32_t> tlr;     // tvo[ntiles] + ntiles
    std::vector<uint32_t> tls;    // 2T
    // the tiles' vertex lists in tile order (tvo ranges), and each corner's
    // index in its tile's list (16 bits each, packed as tls)
    std::vector<int32_t> tvid;    // tvo[ntiles]
    std::vector<uint32_t> tlv;    // 2T
    std::vector<float> tk;        // per-tet ARAP weight per unit volume (empty: the scalar stiffness)
    int bad_tet = -1;
};

inline double tet_k(const HostLevel& L, int t, double k) { return L.tk.empty() ? k : (double)L.tk[(size_t)t]; }

// Jacobi diagonal sum_t k_t vol_t AC^2 recomputed in the precompute's tet
// order (after per-tet stiffness changes).
void recompute_diag(HostLevel& L, float k);

// Connected components of a tet mesh (union of tet corners), numbered in
// order of their smallest vertex index; returns the component count.
int mesh_components(int V, int T, const int32_t* tets, std::vector<int32_t>& comp);

constexpr int kTileTets = 256;
// The tile structures of a level with tets and V set (see HostLevel).
void build_tiles(HostLevel& L);

// Builds a level: copies the mesh, precomputes, builds gather structures.
// stiffness = ARAP weight per unit volume, density scales the lumped mass.
int build_level(const mgsb_mesh& m, float stiffness, float density, HostLevel& L);

// Barycentric binding of `n` points to tets of `mesh` following
// PDSolver_MG::interpolate (PDSolver_MG.cpp:5-58): first tet in index order
// whose weights (barycentricCoordinate, utils.cu:335-364) all lie in
// [-1e-5, 1+1e-5].  Spatial-hash candidates replace the brute-force scan;
// the lowest accepted index is the same.  Returns the number of points with no
// tet (those get ids 0 and weights 0, as the reference leaves them).
int bind_points_interp(const HostLevel& mesh, const float* pts, int n, std::vector<int32_t>& ids,
                       std::vector<float>& w);

// Coarse-in-fine binding of PDSolver_MG::Init (PDSolver_MG.cpp:133-168):
// first tet in index order passing pointInTet (utils.cu:298-311), weights from
// the precomputed-face barycentric form (utils.cu:313-333).
int bind_points_average(const HostLevel& mesh, const float* pts, int n, std::vector<int32_t>& ids,
                        std::vector<float>& w);

// barycentricCoordinate on four points (utils.cu:335-364).
void barycentric4(const float* P, const float* A, const float* B, const float* C, const float* D, float* w);

// V-cycle set-up (mgsb_mg.cpp).
void build_restriction(int Vf, int Vc, const std::vector<int32_t>& iids, const std::vector<float>& iw,
                       std::vector<int32_t>& rtp, std::vector<int32_t>& rti, std::vector<float>& rtw);
void derive_vfix(const std::vector<float>& vfix_fine, const std::vector<int32_t>& aids, const std::vector<float>& aw,
                 int Vc, std::vector<float>& vfix_coarse);
double power_rho(const HostLevel& L, const std::vector<float>& vfix, double dt, double relax, double k);
// Block-diagonal FP64 Cholesky of the coarsest level, one dense block per
// connected component: dense[i] = row of vertex i (-1 pinned), rows grouped
// by component; W holds W_b = C_b^-1 (row-major n_b x n_b) back to back and
// rowinfo[4 r .. 4 r + 3] = (first row of r's block, n_b, offset of W_b, 0).
int coarse_cholesky(const HostLevel& L, const std::vector<float>& vfix, double dt, double k,
                    std::vector<int32_t>& dense, std::vector<double>& W, std::vector<int32_t>& rowinfo, int& n);
void build_level_matrix(const HostLevel& L, double k, double dt, std::vector<int32_t>& ap,
                        std::vector<int32_t>& aj, std::vector<float>& av);

}  // namespace mgsb
