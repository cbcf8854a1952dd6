/* Test oracle (not part of the product): CPU restatement of the reference's
 * offline surfacer, in FP64, serial, particle order.
 *
 *   surface_setup  Source/surface_field.cpp:8-43
 *   surface_aniso  surface_field.cpp:83-176 (grid_insert 178-192,
 *                  grid_neighbors 194-226, ker_spline 233-250, weight_wij 253-260)
 *   surface_isotro surface_field.cpp:45-75
 *   marchingCubes  marching_cubes.cpp:45-120, with a case table built from the
 *                  cube's faces (a restatement of the product's rule, pinned by
 *                  closure, orientation and area tests), and every cube of the
 *                  grid visited once: the reference's y sweep reads the first
 *                  row twice and never reaches the last (its datacache rows are
 *                  refilled before gridi advances), which is not kept.
 *
 * The SVD of the symmetric positive covariance is its eigen decomposition,
 * computed here by cyclic Jacobi in double.  The reference caps the
 * covariance list at 1024 neighbours (surface_field.cpp:114-117); no cap here
 * (equal whenever a particle has at most 1024 neighbours within 2 r_h). */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace {

struct Setup {
    double fmin[3], cell, vol, rh, hf, ks, kr, kn, lambda;
    int fnum[3], Ne, hcells;
    double gmin[3], gdiv;
    int gnum[3];
};

Setup setup(const float mn[3], const float mx[3], float rp_, float rh_, float cell_) {
    Setup S{};
    const double rp = rp_, rh = rh_, cell = cell_;
    S.vol = rp * rp * rp;
    S.rh = rh;
    S.cell = cell;
    S.hf = 2 * rh;
    for (int a = 0; a < 3; ++a) {
        S.fmin[a] = double(mn[a]) - rh * 0.75;
        S.fnum[a] = int((double(mx[a]) - double(mn[a]) + rh * 1.5) / cell) + 1;
        S.gmin[a] = double(mn[a]) - S.hf * 1.25;
        S.gnum[a] = int((double(mx[a]) - double(mn[a]) + S.hf * 2.5) / S.hf) + 1;
    }
    S.gdiv = 1.0 / S.hf;
    S.kr = 4.0;
    S.ks = 1.0 / (0.16 * S.hf * S.hf) * 1.6;
    S.kn = 0.95;
    S.Ne = 25;
    S.lambda = 1.0;
    S.hcells = int(rh / cell);
    return S;
}

double ker_spline(const Setup& S, double r) {
    const double h = S.rh;
    const double p1 = 8.0 / (M_PI * std::pow(h, 3.0)) * S.vol, p2 = -48.0 / (M_PI * std::pow(h, 5.0)) * S.vol;
    const double p3 = 48.0 / (M_PI * std::pow(h, 6.0)) * S.vol, p4 = 16.0 / (M_PI * std::pow(h, 6.0)) * S.vol;
    return r < 0.5 * h ? p1 + p2 * r * r + p3 * r * r * r : p4 * (h - r) * (h - r) * (h - r);
}

void jacobi(double a[3][3], double V[3][3], double ev[3]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) V[i][j] = i == j;
    for (int sweep = 0; sweep < 50; ++sweep) {
        const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        if (off < 1e-300) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (a[p][q] == 0) continue;
                const double th = (a[q][q] - a[p][p]) / (2 * a[p][q]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
                const double c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < 3; ++k) {
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < 3; ++i) ev[i] = a[i][i];
}

/* ---- marching cubes table: the faces' low runs chained into loops ---- */
const int kC[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}, {1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}};
const int kE[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6}, {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

struct McTab {
    uint16_t edges[256];
    std::vector<int> tris[256];
};

McTab make_table() {
    McTab T;
    auto edge = [](int a, int b) {
        for (int e = 0; e < 12; ++e)
            if ((kE[e][0] == a && kE[e][1] == b) || (kE[e][0] == b && kE[e][1] == a)) return e;
        return -1;
    };
    // faces, corners counter-clockwise about the outward normal
    std::vector<std::vector<int>> F;
    for (int ax = 0; ax < 3; ++ax)
        for (int side = 0; side < 2; ++side) {
            std::vector<int> cs;
            for (int c = 0; c < 8; ++c)
                if (kC[c][ax] == side) cs.push_back(c);
            const int u = (ax + 1) % 3, v = (ax + 2) % 3;
            const double sg = side ? 1 : -1;
            std::sort(cs.begin(), cs.end(), [&](int a, int b) {
                return std::atan2(sg * (kC[a][v] - 0.5), kC[a][u] - 0.5) <
                       std::atan2(sg * (kC[b][v] - 0.5), kC[b][u] - 0.5);
            });
            F.push_back(cs);
        }
    for (int ci = 0; ci < 256; ++ci) {
        auto lo = [&](int c) { return (ci >> c) & 1; };
        uint16_t em = 0;
        for (int e = 0; e < 12; ++e)
            if (lo(kE[e][0]) != lo(kE[e][1])) em |= uint16_t(1 << e);
        T.edges[ci] = em;
        int nxt[12];
        std::fill(nxt, nxt + 12, -1);
        for (const auto& f : F)
            for (int k = 0; k < 4; ++k) {
                if (!(lo(f[k]) && !lo(f[(k + 1) % 4]))) continue;
                int j = k;
                while (lo(f[(j + 3) % 4])) j = (j + 3) % 4;
                nxt[edge(f[k], f[(k + 1) % 4])] = edge(f[(j + 3) % 4], f[j]);
            }
        bool used[12] = {};
        std::vector<int> tri;
        for (int e = 0; e < 12; ++e) {
            if (!((em >> e) & 1) || used[e]) continue;
            std::vector<int> L;
            for (int c = e; !used[c]; c = nxt[c]) {
                used[c] = true;
                L.push_back(c);
            }
            /* fan apex: the first vertex whose fan has no triangle inside a
             * cube face (all six end corners share a fixed coordinate) */
            auto flat = [&](int a, int b, int c) {
                for (int ax = 0; ax < 3; ++ax)
                    for (int sd = 0; sd < 2; ++sd) {
                        bool all = true;
                        for (int e : {a, b, c})
                            all = all && kC[kE[e][0]][ax] == sd && kC[kE[e][1]][ax] == sd;
                        if (all) return true;
                    }
                return false;
            };
            const size_t m = L.size();
            size_t s0 = 0;
            for (; s0 < m; ++s0) {
                bool any = false;
                for (size_t k = 1; k + 1 < m; ++k) any = any || flat(L[s0], L[(s0 + k) % m], L[(s0 + k + 1) % m]);
                if (!any) break;
            }
            if (s0 == m) throw std::logic_error("no fan without a face triangle");
            for (size_t k = 1; k + 1 < m; ++k) {
                const int t[3] = {L[s0], L[(s0 + k) % m], L[(s0 + k + 1) % m]};
                tri.insert(tri.end(), t, t + 3);
            }
        }
        T.tris[ci] = tri;
    }
    /* one winding for every case: case 1's triangle faces corner 0 (the
     * origin; the triangle lies towards +x, +y, +z of it) */
    const std::vector<int>& q = T.tris[1];
    double P[3][3];
    for (int m = 0; m < 3; ++m)
        for (int d = 0; d < 3; ++d) P[m][d] = 0.5 * (kC[kE[q[m]][0]][d] + kC[kE[q[m]][1]][d]);
    const double u[3] = {P[1][0] - P[0][0], P[1][1] - P[0][1], P[1][2] - P[0][2]};
    const double w[3] = {P[2][0] - P[0][0], P[2][1] - P[0][1], P[2][2] - P[0][2]};
    const double n[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
    if (n[0] + n[1] + n[2] > 0)
        for (auto& tv : T.tris)
            for (size_t k = 0; k < tv.size(); k += 3) std::swap(tv[k + 1], tv[k + 2]);
    return T;
}

}  // namespace

extern "C" {

int orc_surface_grid(const float mn[3], const float mx[3], float rp, float rh, float cell, double gmin[3],
                     int gnum[3]) {
    const Setup S = setup(mn, mx, rp, rh, cell);
    for (int a = 0; a < 3; ++a) {
        gmin[a] = S.fmin[a];
        gnum[a] = S.fnum[a];
    }
    return 0;
}

/* aniso != 0: surface_aniso, else surface_isotro.  field: fnum product
 * doubles.  xbar (3n), G (9n), det (n), nbr (n) optional (aniso only). */
int orc_surface_field(const float mn[3], const float mx[3], float rp, float rh, float cell, int aniso, long long n,
                      const float* xyz, double* field, double* xbar, double* Gout, double* det, int* nbr) {
    const Setup S = setup(mn, mx, rp, rh, cell);
    const size_t nf = size_t(S.fnum[0]) * S.fnum[1] * S.fnum[2];
    std::fill(field, field + nf, 0.0);
    std::vector<double> P(size_t(3 * n));
    for (long long i = 0; i < 3 * n; ++i) P[i] = xyz[i];
    // particle grid (grid_insert), cells of h_field, clamped to the grid
    auto cellc = [&](const double* p, int a) {
        int c = int((p[a] - S.gmin[a]) * S.gdiv);
        return std::min(std::max(c, 0), S.gnum[a] - 1);
    };
    const size_t ncell = size_t(S.gnum[0]) * S.gnum[1] * S.gnum[2];
    std::vector<long long> first(ncell, -1), next(size_t(n), -1);
    for (long long i = 0; i < n; ++i) {
        const size_t c = (size_t(cellc(&P[3 * i], 2)) * S.gnum[1] + cellc(&P[3 * i], 1)) * S.gnum[0] + cellc(&P[3 * i], 0);
        next[i] = first[c];
        first[c] = i;
    }
    const double hf2 = S.hf * S.hf, h2 = S.rh * S.rh;
    for (long long i = 0; i < n; ++i) {
        const double* pi = &P[3 * i];
        double xb[3] = {pi[0], pi[1], pi[2]};
        double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, s0;
        int cnt = 0;
        if (aniso) {
            double sw = 0, sx[3] = {0, 0, 0};
            std::vector<std::pair<long long, double>> nb;
            const int cx = cellc(pi, 0), cy = cellc(pi, 1), cz = cellc(pi, 2);
            for (int z = std::max(cz - 1, 0); z <= std::min(cz + 1, S.gnum[2] - 1); ++z)
                for (int y = std::max(cy - 1, 0); y <= std::min(cy + 1, S.gnum[1] - 1); ++y)
                    for (int x = std::max(cx - 1, 0); x <= std::min(cx + 1, S.gnum[0] - 1); ++x)
                        for (long long j = first[(size_t(z) * S.gnum[1] + y) * S.gnum[0] + x]; j >= 0; j = next[j]) {
                            if (j == i) continue;
                            const double* pj = &P[3 * j];
                            const double d2 = (pi[0] - pj[0]) * (pi[0] - pj[0]) + (pi[1] - pj[1]) * (pi[1] - pj[1]) +
                                              (pi[2] - pj[2]) * (pi[2] - pj[2]);
                            if (d2 > hf2) continue;
                            const double r = std::sqrt(d2), w = 1 - r * r * r / (S.hf * S.hf * S.hf);   /* eq. (11) */
                            for (int a = 0; a < 3; ++a) sx[a] += pj[a] * w;
                            sw += w;
                            nb.push_back({j, w});
                        }
            cnt = int(nb.size());
            if (cnt >= S.Ne) {
                double C[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                for (int a = 0; a < 3; ++a) xb[a] = pi[a] * (1 - S.lambda) + sx[a] / sw * S.lambda;  /* eqs. (6), (10) */
                for (const auto& q : nb) {
                    const double* pj = &P[3 * q.first];
                    const double d[3] = {pj[0] - xb[0], pj[1] - xb[1], pj[2] - xb[2]};
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) C[a][b] += d[a] * d[b] * q.second;
                }
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) C[a][b] /= sw;                                 /* eq. (9) */
                double V[3][3], ev[3];
                jacobi(C, V, ev);
                int o[3] = {0, 1, 2};
                std::sort(o, o + 3, [&](int a, int b) { return ev[a] > ev[b]; });
                double s[3] = {std::max(ev[o[0]], 0.0), std::max(ev[o[1]], 0.0), std::max(ev[o[2]], 0.0)};
                s[1] = std::max(s[1], s[0] / S.kr);
                s[2] = std::max(s[2], s[0] / S.kr);                                            /* eq. (15) */
                for (double& v : s) v *= S.ks;
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        double g = 0;
                        for (int k = 0; k < 3; ++k) g += V[a][o[k]] / s[k] * V[b][o[k]];
                        G[a][b] = g;
                    }
                s0 = s[0];
            } else {
                const double g = S.kn * (3 - 2 * cnt / double(S.Ne));
                for (int a = 0; a < 3; ++a) G[a][a] = g;
                s0 = 1 / S.kn;
            }
        } else {
            for (int a = 0; a < 3; ++a) G[a][a] = 1;
            s0 = 1;
        }
        const double dG = G[0][0] * (G[1][1] * G[2][2] - G[1][2] * G[2][1]) -
                          G[0][1] * (G[1][0] * G[2][2] - G[1][2] * G[2][0]) +
                          G[0][2] * (G[1][0] * G[2][1] - G[1][1] * G[2][0]);
        if (aniso) {
            if (xbar) for (int a = 0; a < 3; ++a) xbar[3 * i + a] = xb[a];
            if (Gout) for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) Gout[9 * i + 3 * a + b] = G[a][b];
            if (det) det[i] = dG;
            if (nbr) nbr[i] = cnt;
        }
        const int ext = aniso ? int(S.hcells * s0) : S.hcells;
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            const int p = int((xb[a] - S.fmin[a]) / S.cell);
            lo[a] = std::max(p - ext, 0);
            hi[a] = std::min(p + ext + 1, S.fnum[a] - 1);
        }
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    const double d[3] = {S.fmin[0] + x * S.cell - xb[0], S.fmin[1] + y * S.cell - xb[1],
                                         S.fmin[2] + z * S.cell - xb[2]};
                    double u[3];
                    for (int a = 0; a < 3; ++a) u[a] = G[a][0] * d[0] + G[a][1] * d[1] + G[a][2] * d[2];
                    const double r2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
                    if (r2 <= h2)
                        field[(size_t(z) * S.fnum[1] + y) * S.fnum[0] + x] +=
                            (aniso ? dG : 1.0) * ker_spline(S, std::sqrt(r2));
                }
    }
    return 0;
}

/* marchingCubes over every cube, triangles (9 doubles each) in cube order */
long long orc_marching_cubes(const float* f, const int num[3], const double mn[3], double cell, float iso,
                             double* tris, long long cap) {
    static const McTab T = make_table();
    const int nx = num[0], ny = num[1], nz = num[2];
    long long nt = 0;
    for (int z = 0; z + 1 < nz; ++z)
        for (int y = 0; y + 1 < ny; ++y)
            for (int x = 0; x + 1 < nx; ++x) {
                double v[8];
                int ci = 0;
                for (int k = 0; k < 8; ++k) {
                    v[k] = f[(size_t(z + kC[k][2]) * ny + (y + kC[k][1])) * nx + (x + kC[k][0])];
                    if (v[k] < iso) ci |= 1 << k;
                }
                const std::vector<int>& t = T.tris[ci];
                for (size_t m = 0; m < t.size(); ++m) {
                    int a = kE[t[m]][0], b = kE[t[m]][1];
                    /* from the grid edge's lower node (same vertex from both cubes) */
                    if (kC[a][0] + kC[a][1] + kC[a][2] > kC[b][0] + kC[b][1] + kC[b][2]) std::swap(a, b);
                    const double mu = (double(iso) - v[a]) / (v[b] - v[a]);
                    if (tris && nt + (long long)(m / 3) < cap) {
                        double* o = tris + 9 * (nt + (long long)(m / 3)) + 3 * (m % 3);
                        for (int d = 0; d < 3; ++d) {
                            const double pa = mn[d] + ((d == 0 ? x : d == 1 ? y : z) + kC[a][d]) * cell;
                            const double pb = mn[d] + ((d == 0 ? x : d == 1 ? y : z) + kC[b][d]) * cell;
                            o[d] = pa + mu * (pb - pa);
                        }
                    }
                }
                nt += (long long)(t.size() / 3);
            }
    return nt;
}

/* the table, for comparison with the product's */
int orc_marching_cubes_table(unsigned short* edges, unsigned char* ntris, signed char* tris) {
    static const McTab T = make_table();
    for (int c = 0; c < 256; ++c) {
        edges[c] = T.edges[c];
        ntris[c] = (unsigned char)(T.tris[c].size() / 3);
        for (int k = 0; k < 16; ++k) tris[16 * c + k] = k < int(T.tris[c].size()) ? (signed char)T.tris[c][k] : -1;
    }
    return 0;
}

}  // extern "C"
