// Host-side set-up of the residual-correction V-cycle: restriction as the
// transposed prolongation (CSR, fine vertices ascending, so device sums are
// deterministic), Dirichlet masks of the coarse corrections, the Chebyshev
// spectral radius per coarse level, and the FP64 Cholesky factor of the
// coarsest level.  All of it runs once per create (and on dt changes).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "mgsb_internal.h"

namespace mgsb {

void build_restriction(int Vf, int Vc, const std::vector<int32_t>& iids, const std::vector<float>& iw,
                       std::vector<int32_t>& rtp, std::vector<int32_t>& rti, std::vector<float>& rtw) {
    rtp.assign((size_t)Vc + 1, 0);
    for (int i = 0; i < Vf; i++)
        for (int q = 0; q < 4; q++) rtp[iids[(size_t)i * 4 + q] + 1]++;
    for (int j = 0; j < Vc; j++) rtp[j + 1] += rtp[j];
    rti.assign((size_t)rtp[Vc], 0);
    rtw.assign((size_t)rtp[Vc], 0.f);
    std::vector<int32_t> fill((size_t)Vc, 0);
    for (int i = 0; i < Vf; i++)  // ascending fine vertex
        for (int q = 0; q < 4; q++) {
            int j = iids[(size_t)i * 4 + q];
            int e = rtp[j] + fill[j]++;
            rti[e] = i;
            rtw[e] = iw[(size_t)i * 4 + q];
        }
}

// A coarse vertex is pinned iff the largest-weight corner of its binding into
// the finer level (coarse-in-fine map) is pinned; on nested grids that is the
// coincident fine vertex.
void derive_vfix(const std::vector<float>& vfix_fine, const std::vector<int32_t>& aids, const std::vector<float>& aw,
                 int Vc, std::vector<float>& vfix_coarse) {
    vfix_coarse.assign(Vc, 1.f);
    for (int j = 0; j < Vc; j++) {
        int best = 0;
        for (int q = 1; q < 4; q++)
            if (aw[(size_t)j * 4 + q] > aw[(size_t)j * 4 + best]) best = q;
        vfix_coarse[j] = vfix_fine[aids[(size_t)j * 4 + best]] != 0.f ? 1.f : 0.f;
    }
}

// (L w)_a = sum_t k vol sum_k (Ds(w) B)_k AC[k][a] for a scalar field w.
static void apply_stiffness(const HostLevel& L, double k, const std::vector<double>& w, std::vector<double>& out) {
    std::fill(out.begin(), out.end(), 0.0);
    for (int t = 0; t < L.T; t++) {
        const int32_t* id = &L.tets[(size_t)t * 4];
        const float* B = &L.B[(size_t)t * 9];
        const float* AC = &L.AC[(size_t)t * 12];
        double d[3] = {w[id[1]] - w[id[0]], w[id[2]] - w[id[0]], w[id[3]] - w[id[0]]};
        double f[3];
        for (int c = 0; c < 3; c++) f[c] = d[0] * B[c] + d[1] * B[3 + c] + d[2] * B[6 + c];
        double kv = tet_k(L, t, k) * L.vol[t];
        for (int a = 0; a < 4; a++) out[id[a]] += kv * (f[0] * AC[a] + f[1] * AC[4 + a] + f[2] * AC[8 + a]);
    }
}

// Spectral radius of J = I - relax D^-1 A on the free vertices (A = m/h^2 +
// L, D its diagonal), 100 power-iteration steps from a fixed start vector.
double power_rho(const HostLevel& L, const std::vector<float>& vfix, double dt, double relax, double k) {
    const int V = L.V;
    std::vector<double> w(V), q(V);
    for (int i = 0; i < V; i++) w[i] = vfix[i] * (1.0 + 0.5 * std::sin(1.7 * i));
    double lam = 0;
    for (int it = 0; it < 100; it++) {
        double nw = 0;
        for (int i = 0; i < V; i++) nw += w[i] * w[i];
        nw = std::sqrt(nw);
        if (nw == 0) break;
        for (int i = 0; i < V; i++) w[i] /= nw;
        apply_stiffness(L, k, w, q);
        double ny = 0;
        for (int i = 0; i < V; i++) {
            double dc = (double)L.mass[i] / (dt * dt);
            double Aw = dc * w[i] + q[i];
            double y = vfix[i] * (w[i] - relax * Aw / (dc + (double)L.diag[i]));
            w[i] = y;
            ny += y * y;
        }
        lam = std::sqrt(ny);
    }
    return lam;
}

// Dense A = m/h^2 + L on the free vertices of each connected component,
// FP64 Cholesky A_b = C_b C_b^T, and W_b = C_b^-1 (lower triangular), so that
// A_b^-1 r = W_b^T (W_b r).  Components do not couple, so the blocks are the
// whole factor.
int coarse_cholesky(const HostLevel& L, const std::vector<float>& vfix, double dt, double k,
                    std::vector<int32_t>& dense, std::vector<double>& W, std::vector<int32_t>& rowinfo, int& n) {
    const int V = L.V;
    std::vector<int32_t> comp;
    const int nc = mesh_components(V, L.T, L.tets.data(), comp);
    // rows grouped by component, vertices ascending within a component
    std::vector<int32_t> cstart(nc + 1, 0);
    for (int i = 0; i < V; i++)
        if (vfix[i] != 0.f) cstart[comp[i] + 1]++;
    for (int c = 0; c < nc; c++) cstart[c + 1] += cstart[c];
    n = cstart[nc];
    dense.assign(V, -1);
    {
        std::vector<int32_t> fill(cstart.begin(), cstart.end() - 1);
        for (int i = 0; i < V; i++)
            if (vfix[i] != 0.f) dense[i] = fill[comp[i]]++;
    }
    W.clear();
    rowinfo.assign((size_t)n * 4, 0);
    if (n == 0) return 0;
    std::vector<size_t> woff(nc + 1, 0);
    for (int c = 0; c < nc; c++) {
        const size_t nb = (size_t)(cstart[c + 1] - cstart[c]);
        woff[c + 1] = woff[c] + nb * nb;
    }
    if (woff[nc] > (size_t)INT32_MAX) return -1;
    for (int c = 0; c < nc; c++)  // O(n_b^3) host factorisation: bounded block size
        if (cstart[c + 1] - cstart[c] > 2048) return -2;
    W.assign(woff[nc], 0.0);
    std::vector<std::vector<int32_t>> ctets(nc);
    for (int t = 0; t < L.T; t++) ctets[comp[L.tets[(size_t)t * 4]]].push_back(t);
    for (int c = 0; c < nc; c++) {
        const int r0 = cstart[c], nb = cstart[c + 1] - cstart[c];
        for (int r = r0; r < r0 + nb; r++) {
            rowinfo[(size_t)r * 4 + 0] = r0;
            rowinfo[(size_t)r * 4 + 1] = nb;
            rowinfo[(size_t)r * 4 + 2] = (int32_t)woff[c];
        }
        if (nb == 0) continue;
        std::vector<double> A((size_t)nb * nb, 0.0);
        for (int i = 0; i < V; i++)
            if (comp[i] == c && dense[i] >= 0)
                A[(size_t)(dense[i] - r0) * nb + (dense[i] - r0)] += (double)L.mass[i] / (dt * dt);
        for (int t : ctets[c]) {
            const float* AC = &L.AC[(size_t)t * 12];
            double w = (double)L.vol[t] * tet_k(L, t, k);
            for (int a = 0; a < 4; a++)
                for (int b = 0; b < 4; b++) {
                    int ia = dense[L.tets[(size_t)t * 4 + a]], ib = dense[L.tets[(size_t)t * 4 + b]];
                    if (ia < 0 || ib < 0) continue;
                    double s = 0;
                    for (int r = 0; r < 3; r++) s += (double)AC[r * 4 + a] * (double)AC[r * 4 + b];
                    A[(size_t)(ia - r0) * nb + (ib - r0)] += w * s;
                }
        }
        for (int j = 0; j < nb; j++) {
            double d = A[(size_t)j * nb + j];
            for (int p = 0; p < j; p++) d -= A[(size_t)j * nb + p] * A[(size_t)j * nb + p];
            if (!(d > 0)) return -1;
            d = std::sqrt(d);
            A[(size_t)j * nb + j] = d;
            for (int i = j + 1; i < nb; i++) {
                double s = A[(size_t)i * nb + j];
                for (int p = 0; p < j; p++) s -= A[(size_t)i * nb + p] * A[(size_t)j * nb + p];
                A[(size_t)i * nb + j] = s / d;
            }
        }
        // W_b = C_b^-1 by forward substitution on the identity, column by column
        double* Wb = W.data() + woff[c];
        for (int col = 0; col < nb; col++)
            for (int i = col; i < nb; i++) {
                double s = (i == col) ? 1.0 : 0.0;
                for (int p = col; p < i; p++) s -= A[(size_t)i * nb + p] * Wb[(size_t)p * nb + col];
                Wb[(size_t)i * nb + col] = s / A[(size_t)i * nb + i];
            }
    }
    return 0;
}

}  // namespace mgsb

namespace mgsb {

// Scalar CSR matrix of A = m/h^2 I + L on all vertices of a level (the same
// operator acts on x, y and z), L_ab = sum_t k vol_t sum_r AC_t[r][a] AC_t[r][b]
// accumulated in FP64, columns ascending per row.
void build_level_matrix(const HostLevel& L, double k, double dt, std::vector<int32_t>& ap,
                        std::vector<int32_t>& aj, std::vector<float>& av) {
    const int V = L.V;
    std::vector<std::vector<std::pair<int, double>>> rows(V);
    for (int t = 0; t < L.T; t++) {
        const float* AC = &L.AC[(size_t)t * 12];
        const double w = (double)L.vol[t] * tet_k(L, t, k);
        for (int a = 0; a < 4; a++) {
            const int va = L.tets[(size_t)t * 4 + a];
            for (int b = 0; b < 4; b++) {
                const int vb = L.tets[(size_t)t * 4 + b];
                double s = 0;
                for (int r = 0; r < 3; r++) s += (double)AC[r * 4 + a] * (double)AC[r * 4 + b];
                rows[va].push_back({vb, w * s});
            }
        }
    }
    ap.assign((size_t)V + 1, 0);
    aj.clear();
    av.clear();
    for (int i = 0; i < V; i++) {
        auto& r = rows[i];
        r.push_back({i, (double)L.mass[i] / (dt * dt)});
        std::sort(r.begin(), r.end(), [](const std::pair<int, double>& x, const std::pair<int, double>& y) {
            return x.first < y.first;
        });
        size_t q = 0;
        while (q < r.size()) {
            int j = r[q].first;
            double s = 0;
            while (q < r.size() && r[q].first == j) s += r[q++].second;
            aj.push_back(j);
            av.push_back((float)s);
        }
        ap[i + 1] = (int32_t)aj.size();
    }
}

}  // namespace mgsb
