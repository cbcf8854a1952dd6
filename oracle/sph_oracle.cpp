/*
 * sph_oracle.cpp -- CPU restatement of the reference SPH particle loops.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP path, and the timed CPU
 * baseline).  It keeps the reference's loop structure on purpose:
 *   - AoS particles (FluidParticle / BoundaryParticle, SphBase.h:66-199);
 *   - a uniform grid of cell size smoothRadius_h with per-cell linked lists
 *     filled by a SERIAL head insertion, solids before fluids
 *     (SphBase.cpp:991-1089);
 *   - per-particle neighbour lists capped at 48 (SphBase.h:628-646);
 *   - OpenMP `schedule(dynamic, max(100, N/160))` (SphBase.h:675-676);
 *   - serial convergence sums (DFSPH.cpp:626-631, 888-896).
 * Deviations from the reference (each one is a reference defect; see the
 * comments at the use site):
 *   - sums that decide convergence accumulate in double (an FP32 serial sum
 *     of 1e6 values near 1000 carries ~1e-4 relative error, the size of the
 *     tolerance it is compared with);
 *   - the DFSPH density loop has a cap (none at DFSPH.cpp:228);
 *   - IISPH: per-iteration reset of the error sum and iteration counter,
 *     boundary rho_adv uses v_adv, the warm start p_l = 0.5 p(t-dt) is kept,
 *     a_adv is integrated, a_adv reads this step's densities
 *     (IISph_M.cpp:81, 269, 277, 283-284, 350; SphBase.cpp:387-388);
 *   - IISPH: the pressure force from boundary neighbours uses p_a/rho_a^2,
 *     the term the pressure solve predicts, not 2 p_a/rho_a^2
 *     (IISph_M.cpp:558-559 against 127, 271, 338), and so does the reaction
 *     force on dynamic solids (2 p_b/rho_b^2 at IISph_M.cpp:596);
 *   - IISPH phases are solved jointly (the reference iterates phase by phase
 *     with stale cross-phase pressures, IISph_M.cpp:85);
 *   - IISPH with several phases, only with multiphase_consistent = 1 (off
 *     by default): the solve's d_ij p_j and d_ji terms between phases take
 *     the masses of the number-density pressure force the reference
 *     applies (IISph_M.cpp:526-546: m_j^2 / m_i and m_i^2 / m_j) instead of
 *     the particle's own mass (IISph_M.cpp:262, 302, 329), so the solved
 *     pressure is the one the force applies; within a phase nothing changes;
 *   - IISPH convergence error: the mean over all particles of
 *     (rho_adv + a_ii p + sum_j ...)/rho0 - 1 where p is not clamped, and 0
 *     where it is.  The reference sums the predicted density of the
 *     unclamped particles only, divides by all of them and subtracts rho0
 *     (IISph_M.cpp:345-354), so every clamped particle counts as an error of
 *     -rho0 and a free surface ends the loop at once; without clamped
 *     particles the two are the same number;
 *   - a zero distance pair is counted once per neighbour search and
 *     contributes W(0) to densities and nothing to gradient terms (the
 *     reference counts it per force evaluation, WcSph.cpp:160, and divides
 *     0/0 in the alpha and solver passes, DFSPH.cpp:157).
 * Surface tension / adhesion (m_TH.r, m_TH.bt) are zero in every workload
 * and are not restated.  The one-argument kernels use h = smoothRadius_h/2
 * (SphBase.h:808-829), the value ker_update_numbers() would set.
 *
 * Build with -ffp-contract=off so that dis_square() keeps the reference's
 * rounding ((dx*dx + dy*dy) + dz*dz, Vec23.h:170-171).
 */
#include "sph_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>
#include <omp.h>

namespace {

template <class T>
struct V3 {
    T x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3 operator*(T s) const { return V3(x * s, y * s, z * s); }
    V3& operator+=(const V3& o) { x += o.x; y += o.y; z += o.z; return *this; }
    V3& operator-=(const V3& o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
    T dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }      /* Vec23.h:139-140 */
    T len_square() const { return x * x + y * y + z * z; }
    T dis_square(const V3& o) const {                                      /* Vec23.h:170-171 */
        return (x - o.x) * (x - o.x) + (y - o.y) * (y - o.y) + (z - o.z) * (z - o.z);
    }
    V3 cross(const V3& a) const { return V3(y * a.z - z * a.y, z * a.x - x * a.z, x * a.y - y * a.x); }
    bool inside(const V3& lo, const V3& hi) const {                        /* Vec23.h:193-194 */
        return lo.x <= x && lo.y <= y && lo.z <= z && x <= hi.x && y <= hi.y && z <= hi.z;
    }
};

template <class T>
struct FluidPart {           /* the fields of FluidParticle (SphBase.h:66-165) the loops touch */
    V3<T> x, v, a, a_adv, v_adv, dii, sum_dijpj, a_p, xp;
    T rho = 0, p = 0, p_l = 0, aii = 0, rho_adv = 0;
    T alpha = 0, kappa = 0, kappaV = 0, divdev = 0;
    T kappa_sum = 0, kappaV_sum = 0;   /* accumulated kappa*dt^2 (warm start state) */
    T d = 0;
    T beta = 1;                        /* IISph per-particle rest-density factor (SphBase.h:98) */
    int id = 0;
};

template <class T>
struct BoundPart {           /* BoundaryParticle (SphBase.h:168-199) */
    V3<T> x, v, force;
    T volume = 0, kerSumFromSelf = 0, d = 0;
    T rho = 0, alpha = 0, rho_adv = 0, kappa = 0, kappaV = 0;   /* IISph DF_* solvers */
};

template <class T>
struct Fluid {
    T rho0, visc_alpha;
    std::vector<FluidPart<T>> parts;
};

template <class T>
struct Solid {
    bool dynamic, rigid;
    T visc_alpha;
    std::vector<BoundPart<T>> parts;
    std::vector<V3<T>> inner;   /* innerPositions: body-frame samples of a dynamic solid */
    V3<T> com;                  /* centre of mass of the last pose */
    bool posePending = false;   /* pose set by the host, applied at the next update point */
    float pT[16] = {}, pCom[3] = {}, pV[3] = {}, pW[3] = {};
};

struct PartIdx {             /* SphBase.h:596-617 */
    int t, i;
    bool valid() const { return t != 0; }
    bool isFluid() const { return t < 0; }
    int toFluidI() const { return -t - 1; }
    int toSolidI() const { return t - 1; }
    bool operator==(const PartIdx& o) const { return t == o.t && i == o.i; }
};

template <class T>
struct Neigb { PartIdx pidx; T dis; };

template <class T>
struct NeigbStr {            /* SphBase.h:628-646, cap 48 */
    static const int maxNeigbNum = 48;
    std::array<Neigb<T>, maxNeigbNum> neigs;
    int num = 0;
    void push(const PartIdx& j, T dis, int& overflow) {
        if (num < maxNeigbNum) { neigs[num].pidx = j; neigs[num].dis = dis; ++num; }
        else ++overflow;
    }
};

static int dyn_chunk(int n) { return std::max(100, n / 160); }  /* SphBase.h:675-676 */

template <class T>
class Sim {
public:
    orc_params P;
    T r, H, kh, sigma, sigma_h, W0;
    V3<T> spMin, spMax, grav;
    std::vector<Fluid<T>> fluids;
    std::vector<Solid<T>> solids;
    std::vector<std::vector<NeigbStr<T>>> nbF, nbS;
    int frame = 0;
    bool needSearch = false;   /* fluid appended since the last neighbour search */
    T dt_cur, dt_prev = 0;
    int zeroDis = 0, tooMany = 0;

    /* SpaceGridStruct (SphBase.h:682-705) */
    struct Grid {
        V3<T> min; T divisor = 0; int num[3] = {0, 0, 0}; long long factor[3] = {1, 1, 1};
        std::vector<long long> offset;
        std::vector<PartIdx> first;
        std::vector<std::vector<PartIdx>> nextF, nextS;
        bool init = false;
        long long cellOf(const V3<T>& p) const {                          /* SphBase.h:700-701 */
            int cx = int((p.x - min.x) * divisor), cy = int((p.y - min.y) * divisor),
                cz = int((p.z - min.z) * divisor);
            return cx * factor[0] + cy * factor[1] + cz * factor[2];
        }
    } G;

    explicit Sim(const orc_params& p) : P(p) {
        r = T(P.spacing_r);
        H = T(P.smooth_radius_h);
        kh = H / 2;                                                        /* SphBase.h:812 */
        sigma = T(1) / T(M_PI) / (kh * kh * kh);                           /* SphBase.h:814-816 */
        sigma_h = T(1) / T(M_PI) / (kh * kh * kh) / kh;                    /* SphBase.h:817-819 */
        W0 = sigma;
        spMin = V3<T>(P.space_min[0], P.space_min[1], P.space_min[2]);
        spMax = V3<T>(P.space_max[0], P.space_max[1], P.space_max[2]);
        grav = V3<T>(P.gravity[0], P.gravity[1], P.gravity[2]);
        dt_cur = T(P.dt);
    }

    /* cubic spline, SphBase.h:840-853 with h = smoothRadius_h/2 */
    T W(T rr) const {
        T q = rr / kh, qm2 = 2 - q;
        if (!(q >= 0 && q <= 2)) return 0;
        return q <= 1 ? sigma * (1 - T(1.5) * q * q + T(0.75) * q * q * q)
                      : sigma * (T(0.25) * qm2 * qm2 * qm2);
    }
    /* |dW/dr|, SphBase.h:876-889 */
    T Wg(T rr) const {
        T q = rr / kh, qm2 = 2 - q;
        if (!(q >= 0 && q <= 2)) return 0;
        return q <= 1 ? sigma_h * (3 * q - T(2.25) * q * q) : sigma_h * T(0.75) * qm2 * qm2;
    }

    T mass(int k) const { return fluids[k].rho0 * (r * r * r); }         /* fm0 = rho0*v0, DFSPH.cpp:88,119 */

    const FluidPart<T>& fp(const PartIdx& j) const { return fluids[j.toFluidI()].parts[j.i]; }
    BoundPart<T>& bp(const PartIdx& j) { return solids[j.toSolidI()].parts[j.i]; }
    const BoundPart<T>& bp(const PartIdx& j) const { return solids[j.toSolidI()].parts[j.i]; }
    const V3<T>& posOf(const PartIdx& j) const { return j.isFluid() ? fp(j).x : bp(j).x; }
    T dOf(const PartIdx& j) const { return j.isFluid() ? fp(j).d : bp(j).d; }
    const PartIdx& nextOf(const PartIdx& j) const {
        return j.isFluid() ? G.nextF[j.toFluidI()][j.i] : G.nextS[j.toSolidI()][j.i];
    }

    /* SphBase::SpaceGridStruct::initialize, SphBase.cpp:963-988 */
    void gridInit() {
        const int gridt = 1;
        T rh = H;
        G.min = V3<T>(spMin.x - rh * 2, spMin.y - rh * 2, spMin.z - rh * 2);
        G.divisor = gridt / rh;
        V3<T> ext((spMax.x - spMin.x + rh * 2) * G.divisor, (spMax.y - spMin.y + rh * 2) * G.divisor,
                  (spMax.z - spMin.z + rh * 2) * G.divisor);
        G.num[0] = int(ext.x) + 2; G.num[1] = int(ext.y) + 2; G.num[2] = int(ext.z) + 2;
        G.factor[0] = 1; G.factor[1] = G.num[0]; G.factor[2] = (long long)G.num[0] * G.num[1];
        G.first.assign((size_t)G.num[0] * G.num[1] * G.num[2], PartIdx{0, -1});
        G.offset.clear();
        for (int dz = -gridt; dz <= gridt; ++dz)
            for (int dy = -gridt; dy <= gridt; ++dy)
                for (int dx = -gridt; dx <= gridt; ++dx)
                    G.offset.push_back(dx * G.factor[0] + dy * G.factor[1] + dz * G.factor[2]);
        G.init = true;
    }

    /* SphBase::gridSetup, SphBase.cpp:991-1089 (serial head insertion) */
    void gridSetup() {
        if (!G.init) gridInit();
        int nf = int(fluids.size()), ns = int(solids.size());
        G.nextF.resize(nf); nbF.resize(nf);
        for (int k = 0; k < nf; ++k) { G.nextF[k].resize(fluids[k].parts.size()); nbF[k].resize(fluids[k].parts.size()); }
        G.nextS.resize(ns); nbS.resize(ns);
        for (int k = 0; k < ns; ++k) { G.nextS[k].resize(solids[k].parts.size()); nbS[k].resize(solids[k].parts.size()); }
        std::fill(G.first.begin(), G.first.end(), PartIdx{0, -1});
        std::vector<long long> cell;
        for (int k = 0; k < ns; ++k) {
            auto& ps = solids[k].parts; int n = int(ps.size()); cell.resize(n);
#pragma omp parallel for schedule(dynamic, 100)
            for (int i = 0; i < n; ++i) cell[i] = G.cellOf(ps[i].x);
            for (int i = 0; i < n; ++i) {
                G.nextS[k][i] = G.first[cell[i]];
                G.first[cell[i]] = PartIdx{k + 1, i};
            }
        }
        for (int k = 0; k < nf; ++k) {
            auto& ps = fluids[k].parts; int n = int(ps.size()); cell.resize(n);
#pragma omp parallel for schedule(dynamic, 100)
            for (int i = 0; i < n; ++i) cell[i] = G.cellOf(ps[i].x);
            for (int i = 0; i < n; ++i) {
                G.nextF[k][i] = G.first[cell[i]];
                G.first[cell[i]] = PartIdx{-(k + 1), i};
            }
        }
    }

    /* SphBase::neighbourSearch, SphBase.cpp:849-960 */
    void neighbourSearch(bool fluidPart = true, bool solidPart = true, bool RR_same = false) {
        gridSetup();
        if (fluidPart) needSearch = false;
        int n_off = int(G.offset.size());
        int overflow = 0, zero = 0;
        if (fluidPart)
            for (int k = 0; k < int(fluids.size()); ++k) {
                auto& ps = fluids[k].parts; int n = int(ps.size());
#pragma omp parallel for schedule(dynamic, dyn_chunk(n)) reduction(+ : overflow, zero)
                for (int ii = 0; ii < n; ++ii) {
                    PartIdx i{-(k + 1), ii};
                    NeigbStr<T>& L = nbF[k][ii]; L.num = 0;
                    const V3<T>& pa = ps[ii].x;
                    long long jo = G.cellOf(pa);
                    for (int o = 0; o < n_off; ++o)
                        for (PartIdx j = G.first[jo + G.offset[o]]; j.valid(); j = nextOf(j)) {
                            if (i == j) continue;
                            T h = dOf(j) + ps[ii].d;
                            T h2 = h * h;                  /* (real_t)pow(h,2): exact square rounded once */
                            T len2 = pa.dis_square(posOf(j));
                            if (len2 < h2) {
                                if (len2 == 0) ++zero;
                                L.push(j, std::sqrt(len2), overflow);
                            }
                        }
                }
            }
        if (solidPart)
            for (int k = 0; k < int(solids.size()); ++k) {
                auto& ps = solids[k].parts; int n = int(ps.size());
                bool rigid = solids[k].rigid;
#pragma omp parallel for schedule(dynamic, dyn_chunk(n)) reduction(+ : overflow)
                for (int ii = 0; ii < n; ++ii) {
                    PartIdx i{k + 1, ii};
                    NeigbStr<T>& L = nbS[k][ii]; L.num = 0;
                    const V3<T>& pa = ps[ii].x;
                    long long jo = G.cellOf(pa);
                    for (int o = 0; o < n_off; ++o)
                        for (PartIdx j = G.first[jo + G.offset[o]]; j.valid(); j = nextOf(j)) {
                            if (i == j) continue;
                            if (!RR_same && rigid && i.t == j.t) continue;
                            T h = dOf(j) + ps[ii].d;
                            T h2 = h * h;
                            T len2 = pa.dis_square(posOf(j));
                            if (len2 < h2) L.push(j, std::sqrt(len2), overflow);
                        }
                }
            }
        tooMany += overflow;
        zeroDis += zero;   /* m_EC.zeroDis: coincident fluid-particle neighbour pairs of this search */
    }

    /* SphBase::prepareSolidPartWeight, SphBase.cpp:766-813 */
    void prepareSolidPartWeight() {
        neighbourSearch(false, true, true);
        for (int k = 0; k < int(solids.size()); ++k) {
            auto& ps = solids[k].parts;
            bool rigid = solids[k].rigid;
            for (size_t i = 0; i < ps.size(); ++i) {
                const NeigbStr<T>& L = nbS[k][i];
                T sum = W0, sum2 = W0;
                for (int j = 0; j < L.num; ++j) {
                    const PartIdx& pj = L.neigs[j].pidx;
                    if (rigid && pj.t == k + 1) sum += W(L.neigs[j].dis);
                    if (!pj.isFluid()) sum2 += W(L.neigs[j].dis);
                }
                ps[i].kerSumFromSelf = rigid ? sum : W0;
                ps[i].volume = 1 / sum2;
            }
        }
    }

    /* SphBase::updateSolidPartWeight, SphBase.cpp:817-841 */
    void updateSolidPartWeight() {
        for (int k = 0; k < int(solids.size()); ++k) {
            auto& ps = solids[k].parts; int n = int(ps.size());
#pragma omp parallel for
            for (int i = 0; i < n; ++i) {
                const NeigbStr<T>& L = nbS[k][i];
                T vol = ps[i].kerSumFromSelf;
                for (int j = 0; j < L.num; ++j)
                    if (!L.neigs[j].pidx.isFluid()) vol += W(L.neigs[j].dis);
                ps[i].volume = 1 / vol;
            }
        }
    }

    int numFluid() const { int n = 0; for (auto& f : fluids) n += int(f.parts.size()); return n; }

    /* SphFluid::removeOutsideParts, SphBase.cpp:26-48 (serial swap compaction) */
    int removeOutside(std::vector<FluidPart<T>>& fp_) {
        if (fp_.empty()) return 0;
        int num = int(fp_.size()), f = 0, b = num - 1;
        while (f < b) {
            while (f < b && fp_[f].x.inside(spMin, spMax)) ++f;
            while (b > f && !fp_[b].x.inside(spMin, spMax)) --b;
            if (f < b) { fp_[f] = fp_[b]; ++f; --b; }
        }
        if (!fp_[f].x.inside(spMin, spMax)) --f;
        fp_.resize(f + 1);
        return num - (f + 1);
    }

    /* ------------------------------ rigids ----------------------------- */
    /* SphSolid::updatePaticles, SphBase.cpp:95-128: world position from the
     * OpenGL column-major transform (with the projective divide), velocity
     * v + w x (x - com) */
    void updateParticles(int k, const float* Tm, const float* c, const float* lv, const float* av) {
        Solid<T>& so = solids[k];
        so.com = V3<T>(c[0], c[1], c[2]);
        V3<T> vl(lv[0], lv[1], lv[2]), w(av[0], av[1], av[2]);
        for (size_t i = 0; i < so.inner.size(); ++i) {
            const V3<T>& p0 = so.inner[i];
            V3<T>& bp = so.parts[i].x;
            bp.x = T(Tm[0]) * p0.x + T(Tm[4]) * p0.y + T(Tm[8]) * p0.z + T(Tm[12]);
            bp.y = T(Tm[1]) * p0.x + T(Tm[5]) * p0.y + T(Tm[9]) * p0.z + T(Tm[13]);
            bp.z = T(Tm[2]) * p0.x + T(Tm[6]) * p0.y + T(Tm[10]) * p0.z + T(Tm[14]);
            T alpha = T(Tm[3]) * p0.x + T(Tm[7]) * p0.y + T(Tm[11]) * p0.z + T(Tm[15]);
            bp.x /= alpha; bp.y /= alpha; bp.z /= alpha;
            V3<T> rr = bp - so.com;
            so.parts[i].v = vl + V3<T>(w.y * rr.z - w.z * rr.y, w.z * rr.x - w.x * rr.z, w.x * rr.y - w.y * rr.x);
        }
    }

    /* Reaction force on the particles of dynamic solids from their fluid
     * neighbours: WcSph::wc_boundPartForce_f (WcSph.cpp:314-362, loop
     * 438-460) and IISph_M::ii_boundPartForcePressure_f (IISph_M.cpp:586-607,
     * loop 660-681).  The pressure factor is 2 for WCSPH (as its fluid-side
     * boundary force, WcSph.cpp:285) and 1 for IISPH, matching the fluid-side
     * IISPH boundary force this restatement uses (see the header); adhesion is
     * off (m_TH.bt = 0). */
    void rigidReaction(bool iisph) {
        T v0 = r * r * r, Hs = H, cs = T(P.sound_speed);
        const T pf = iisph ? T(1) : T(2);
        for (int k = 0; k < int(solids.size()); ++k) {
            Solid<T>& so = solids[k];
            if (!so.dynamic) continue;
            T ralpha = so.visc_alpha;
            int n = int(so.parts.size());
#pragma omp parallel for
            for (int i = 0; i < n; ++i) {
                BoundPart<T>& a = so.parts[i];
                a.force = V3<T>();
                const NeigbStr<T>& L = nbS[k][i];
                for (int j = 0; j < L.num; ++j) {
                    const auto& nb = L.neigs[j];
                    if (!nb.pidx.isFluid()) continue;
                    T dis = nb.dis;
                    if (dis == 0) continue;
                    const FluidPart<T>& b = fp(nb.pidx);
                    T frho0 = fluids[nb.pidx.toFluidI()].rho0, fm0 = frho0 * v0;
                    V3<T> xab = a.x - b.x;
                    T grad = -Wg(dis) / dis, force = 0;
                    if (b.p > 0) force += grad * (-fm0 * frho0 * a.volume * (b.p / (b.rho * b.rho) * pf));
                    T pro = (a.v - b.v).dot(xab);
                    if (pro < 0) {
                        T nu = 2 * ralpha * Hs * cs / (b.rho * 2);
                        T pi = -nu * pro / (dis * dis + T(0.01) * Hs * Hs);
                        force += grad * (-fm0 * frho0 * a.volume * pi);
                    }
                    a.force += xab * force;
                }
            }
        }
    }

    /* SphSolid::rigidForceJoin, SphBase.cpp:131-143, summed in double */
    void rigidForceJoin(int k, double* f, double* t) const {
        const Solid<T>& so = solids[k];
        double tf[3] = {0, 0, 0}, tt[3] = {0, 0, 0};
        for (const auto& bp : so.parts) {
            double fx = double(bp.force.x), fy = double(bp.force.y), fz = double(bp.force.z);
            double rx = double(bp.x.x) - double(so.com.x), ry = double(bp.x.y) - double(so.com.y),
                   rz = double(bp.x.z) - double(so.com.z);
            tf[0] += fx; tf[1] += fy; tf[2] += fz;
            tt[0] += ry * fz - rz * fy; tt[1] += rz * fx - rx * fz; tt[2] += rx * fy - ry * fx;
        }
        for (int q = 0; q < 3; ++q) { f[q] = tf[q]; t[q] = tt[q]; }
    }

    /* ------------------------------ WCSPH ------------------------------ */
    /* WcSph::wc_computePressure, WcSph.cpp:99-153 (normals omitted: surface tension is off) */
    void wc_computePressure() {
        T v0 = r * r * r;
        for (int k = 0; k < int(fluids.size()); ++k) {
            auto& ps = fluids[k].parts; int n = int(ps.size());
            T rho0 = fluids[k].rho0, fm0 = rho0 * v0, wm0 = W0 * fm0;
            T B = rho0 * T(P.sound_speed) * T(P.sound_speed) / T(P.gamma);
#pragma omp parallel for schedule(dynamic, dyn_chunk(n))
            for (int i = 0; i < n; ++i) {
                FluidPart<T>& a = ps[i];
                T density = wm0;
                const NeigbStr<T>& L = nbF[k][i];
                for (int j = 0; j < L.num; ++j) {
                    const auto& nb = L.neigs[j];
                    if (nb.pidx.isFluid()) density += fm0 * W(nb.dis);
                    else density += rho0 * bp(nb.pidx).volume * W(nb.dis);
                }
                a.rho = density;
                a.p = B * (std::pow(density / rho0, T(P.gamma)) - 1);
                if (a.p < 0) a.p = 0;                                      /* SphBase.h:32 */
            }
        }
    }

    /* WcSph::wc_computeForce + wc_fluidPartForce_fsame/_fdiff/_bound, WcSph.cpp:156-312, 364-436 */
    void wc_computeForce() {
        T v0 = r * r * r;
        T Hs = H, cs = T(P.sound_speed);
        int zd = 0;
        for (int k = 0; k < int(fluids.size()); ++k) {
            auto& ps = fluids[k].parts; int n = int(ps.size());
            T rho0 = fluids[k].rho0, fm0 = rho0 * v0, alpha = fluids[k].visc_alpha;
#pragma omp parallel for schedule(dynamic, dyn_chunk(n)) reduction(+ : zd)
            for (int i = 0; i < n; ++i) {
                FluidPart<T>& a = ps[i];
                a.a = grav;
                const NeigbStr<T>& L = nbF[k][i];
                for (int j = 0; j < L.num; ++j) {
                    const auto& nb = L.neigs[j];
                    T dis = nb.dis;
                    if (dis == 0) continue;
                    if (nb.pidx.isFluid()) {
                        const FluidPart<T>& b = fp(nb.pidx);
                        int kb = nb.pidx.toFluidI();
                        V3<T> xab = a.x - b.x;
                        T grad = -Wg(dis) / dis;
                        T acce;
                        T pro = (a.v - b.v).dot(xab);
                        if (kb == k) {
                            acce = grad * (-fm0 * (a.p / (a.rho * a.rho) + b.p / (b.rho * b.rho)));
                            if (pro < 0) {
                                T nu = 2 * alpha * Hs * cs / (a.rho + b.rho);
                                T pi = -nu * pro / (dis * dis + T(0.01) * Hs * Hs);
                                acce += grad * (-fm0 * pi);
                            }
                        } else {
                            T fmb = v0 * fluids[kb].rho0;
                            T al = (alpha + fluids[kb].visc_alpha) / 2;
                            T da = a.rho / fm0, db = b.rho / fmb;
                            acce = grad * (-1 / fm0 * (a.p / (da * da) + b.p / (db * db)));
                            if (pro < 0) {
                                T nu = 2 * al * Hs * cs / (a.rho + b.rho);
                                T pi = -nu * pro / (dis * dis + T(0.01) * Hs * Hs);
                                acce += grad * (-(fm0 + fmb) / 2 * pi);
                            }
                        }
                        a.a += xab * acce;
                    } else {
                        const BoundPart<T>& b = bp(nb.pidx);
                        T ralpha = solids[nb.pidx.toSolidI()].visc_alpha;
                        V3<T> xab = a.x - b.x;
                        T grad = -Wg(dis) / dis, acce = 0;
                        if (a.p > 0) acce += grad * (-rho0 * b.volume * (a.p / (a.rho * a.rho) * 2));
                        T pro = (a.v - b.v).dot(xab);
                        if (pro < 0) {
                            T nu = 2 * ralpha * Hs * cs / (a.rho * 2);
                            T pi = -nu * pro / (dis * dis + T(0.01) * Hs * Hs);
                            acce += grad * (-rho0 * b.volume * pi);
                        }
                        a.a += xab * acce;
                    }
                }
            }
        }
        zeroDis += zd;
    }

    /* SphBase::updateFluids, SphBase.cpp:377-402 (WCSPH branch) */
    int updateFluidsWC() {
        int removed = 0;
        T dt = dt_cur;
        for (auto& f : fluids) {
            int n = int(f.parts.size());
#pragma omp parallel for
            for (int i = 0; i < n; ++i) {
                f.parts[i].v += f.parts[i].a * dt;
                f.parts[i].x += f.parts[i].v * dt;
            }
            removed += removeOutside(f.parts);
        }
        return removed;
    }

    /* ------------------------------ DFSPH ------------------------------ */
    /* DFSPH::DF_constantDensitySolverFirst, DFSPH.cpp:84-193 */
    void df_first() {
        T v0 = r * r * r;
        int zd = 0;
        for (int k = 0; k < int(fluids.size()); ++k) {
            auto& ps = fluids[k].parts; int n = int(ps.size());
            T rho0 = fluids[k].rho0, fm0 = rho0 * v0, wm0 = W0 * fm0;
#pragma omp parallel for schedule(dynamic, dyn_chunk(n)) reduction(+ : zd)
            for (int i = 0; i < n; ++i) {
                FluidPart<T>& a = ps[i];
                T density = wm0, a2 = 0;
                V3<T> a1;
                const NeigbStr<T>& L = nbF[k][i];
                for (int j = 0; j < L.num; ++j) {
                    const auto& nb = L.neigs[j];
                    if (nb.dis == 0) {                 /* coincident: W(0) only, no gradient */
                        density += nb.pidx.isFluid() ? fm0 * W0 : rho0 * bp(nb.pidx).volume * W0;
                        continue;
                    }
                    T gs = -Wg(nb.dis);
                    if (nb.pidx.isFluid()) {
                        density += fm0 * W(nb.dis);
                        V3<T> g = (a.x - fp(nb.pidx).x) * (gs / nb.dis);
                        a1 += g * fm0;
                        a2 += (gs * fm0) * (gs * fm0);
                    } else {
                        const BoundPart<T>& b = bp(nb.pidx);
                        density += rho0 * b.volume * W(nb.dis);
                        V3<T> g = (a.x - b.x) * (gs / nb.dis);
                        a1 += g * (rho0 * b.volume);
                    }
                }
                a.rho = density;
                T den = a1.len_square() + a2;
                if (den < T(1.0e-6)) den = T(1.0e-6);
                a.alpha = a.rho / den;
            }
        }
        zeroDis += zd;
    }

    /* DFSPH::DF_forceExceptPressure + helpers, DFSPH.cpp:1246-1359, 1419-1475 */
    void df_forceExceptPressure() {
        T v0 = r * r * r, Hs = H, nu = T(P.nu), cs = T(P.sound_speed);
        int zd = 0;
        for (int k = 0; k < int(fluids.size()); ++k) {
            auto& ps = fluids[k].parts; int n = int(ps.size());
            T rho0 = fluids[k].rho0, fm0 = rho0 * v0, alpha = fluids[k].visc_alpha;
#pragma omp parallel for schedule(dynamic, dyn_chunk(n)) reduction(+ : zd)
            for (int i = 0; i < n; ++i) {
                FluidPart<T>& a = ps[i];
                a.p = 0;
                a.a_adv = grav;
                const NeigbStr<T>& L = nbF[k][i];
                for (int j = 0; j < L.num; ++j) {
                    const auto& nb = L.neigs[j];
                    T dis = nb.dis;
                    if (dis == 0) continue;
                    if (nb.pidx.isFluid()) {
                        const FluidPart<T>& b = fp(nb.pidx);
                        int kb = nb.pidx.toFluidI();
                        V3<T> xab = a.x - b.x;
                        T pro = (a.v - b.v).dot(xab);
                        if (kb == k) {                                     /* DFSPH.cpp:1246-1275 */
                            if (pro < 0) {
                                V3<T> g2 = xab * (-Wg(dis) / dis);
                                a.a_adv += g2 * (pro / (dis * dis + T(0.01) * Hs * Hs) * fm0 / b.rho * 10 * nu);
                            }
                        } else {                                           /* DFSPH.cpp:1278-1303 */
                            T fmb = v0 * fluids[kb].rho0, al = (alpha + fluids[kb].visc_alpha) / 2;
                            if (pro < 0) {
                                T nuv = 2 * al * Hs * cs / (a.rho + b.rho);
                                T pi = -nuv * pro / (dis * dis + T(0.01) * Hs * Hs);
                                T acce = (-Wg(dis) / dis) * (-(fm0 + fmb) / 2 * pi);
                                a.a_adv += xab * acce;
                            }
                        }
                    } else {                                               /* DFSPH.cpp:1334-1359 */
                        BoundPart<T>& b = bp(nb.pidx);
                        V3<T> xab = a.x - b.x;
                        T pro = (a.v - b.v).dot(xab);
                        if (pro < 0) {
                            V3<T> g2 = xab * (-Wg(dis) / dis);
                            V3<T> tmp = g2 * (pro / (dis * dis + T(0.01) * Hs * Hs) * (rho0 * b.volume) / a.rho * 10 * nu);
                            a.a_adv += tmp;
                        }
                    }
                }
            }
        }
        zeroDis += zd;
    }

    /* pass B of either DFSPH loop: v* -= dt * sum m (k_a/rho_a + k_b/rho_b) gradW (+ boundary),
     * DFSPH.cpp:647-693 and 840-883.  `which` selects the kappa field. */
    void df_velocityCorrection(int which, T dt, const std::vector<std::vector<T>>* kapOverride = nullptr) {
        T v0 = r * r * r;
        for (int k = 0; k < int(fluids.size()); ++k) {
            auto& ps = fluids[k].parts; int n = int(ps.size());
            T rho0 = fluids[k].rho0, fm0 = rho0 * v0;
            auto kap = [&](const FluidPart<T>& q, int kk, int idx) -> T {
                if (kapOverride) return (*kapOverride)[kk][idx];
                return which == 0 ? q.kappa : q.kappaV;
            };
#pragma omp parallel for schedule(dynamic, dyn_chunk(n))
            for (int i = 0; i < n; ++i) {
                FluidPart<T>& a = ps[i];
                T ka = kap(a, k, i);
                const NeigbStr<T>& L = nbF[k][i];
                for (int j = 0; j < L.num; ++j) {
                    const auto& nb = L.neigs[j];
                    if (nb.dis == 0) continue;
                    T gs = -Wg(nb.dis);
                    if (nb.pidx.isFluid()) {
                        const FluidPart<T>& b = fp(nb.pidx);
                        T kb = kap(b, nb.pidx.toFluidI(), nb.pidx.i);
                        V3<T> g = (a.x - b.x) * (gs / nb.dis);
                        a.v_adv = a.v_adv - g * (fm0 * ((ka / a.rho) + (kb / b.rho)) * dt);
                    } else {
                        const BoundPart<T>& b = bp(nb.pidx);
                        V3<T> g = (a.x - b.x) * (gs / nb.dis);
                        V3<T> acc_p = g * (rho0 * b.volume * (ka / a.rho));
                        a.v_adv = a.v_adv - acc_p * dt;
                    }
                }
            }
        }
    }

    /* Dynamic rigid step 4 (DFSPH.cpp:674-690): the boundary term of this
     * iteration's velocity correction, acc_p = gradW rho0 V_b kappa_a/rho_a,
     * added to the dynamic boundary particle, times the fluid particle's mass
     * (step 5, DFSPH.cpp:723-737, there m_Fluids[0]'s for every phase; here
     * each phase's own, which keeps Newton's third law for two phases).
     * Serial: the reference's += from a parallel loop races. */
    void df_rigidReaction() {
        bool any = false;
        for (auto& so : solids) any |= so.dynamic;
        if (!any) return;
        T v0 = r * r * r;
        for (int k = 0; k < int(fluids.size()); ++k) {
            const auto& ps = fluids[k].parts;
            T rho0 = fluids[k].rho0, fm0 = rho0 * v0;
            for (size_t i = 0; i < ps.size(); ++i) {
                const FluidPart<T>& a = ps[i];
                const NeigbStr<T>& L = nbF[k][i];
                for (int j = 0; j < L.num; ++j) {
                    const auto& nb = L.neigs[j];
                    if (nb.dis == 0 || nb.pidx.isFluid()) continue;
                    Solid<T>& so = solids[nb.pidx.toSolidI()];
                    if (!so.dynamic) continue;
                    BoundPart<T>& b = so.parts[nb.pidx.i];
                    V3<T> g = (a.x - b.x) * (-Wg(nb.dis) / nb.dis);
                    b.force += g * (rho0 * b.volume * (a.kappa / a.rho) * fm0);
                }
            }
        }
    }

    /* the host's pose (set_solid_pose) reaches the samples here: at init, and
     * at the point of SphWithBullet::updateSolids in the step */
    void applyPendingPoses() {
        for (int k = 0; k < int(solids.size()); ++k)
            if (solids[k].posePending) {
                Solid<T>& so = solids[k];
                updateParticles(k, so.pT, so.pCom, so.pV, so.pW);
                so.posePending = false;
            }
    }

    /* sum_j m (v*_a - v*_b).gradW + sum_b rho0 V_b (v*_a - v_b).gradW for particle i of
     * fluid k: the predicted density change rate of DFSPH.cpp:589-617 / 795-825. */
    T densityChange(int k, int i) const {
        const auto& ps = fluids[k].parts;
        const FluidPart<T>& a = ps[i];
        T rho0 = fluids[k].rho0, fm0 = rho0 * r * r * r, d = 0;
        const NeigbStr<T>& L = nbF[k][i];
        for (int j = 0; j < L.num; ++j) {
            const auto& nb = L.neigs[j];
            if (nb.dis == 0) continue;
            T gs = -Wg(nb.dis);
            if (nb.pidx.isFluid()) {
                const FluidPart<T>& b = fp(nb.pidx);
                d += ((a.x - b.x) * (gs / nb.dis)).dot(a.v_adv - b.v_adv) * fm0;
            } else {
                const BoundPart<T>& b = bp(nb.pidx);
                d += ((a.x - b.x) * (gs / nb.dis)).dot(a.v_adv - b.v) * (rho0 * b.volume);
            }
        }
        return d;
    }

    /* Warm start of a DFSPH loop (Bender & Koschier; not in the reference):
     * a scaled copy of the previous step's accumulated stiffness (one half
     * for the density loop, warm_scale for the divergence loop), only where
     * the solve itself would apply stiffness (rho* > rho0 for the density
     * loop; Drho/Dt > 0 and rho + dt Drho/Dt >= rho0 for the divergence
     * loop), applied as one velocity correction.  The accumulator restarts
     * from the applied guess. */
    void warmStart(int which, T dt) {
        std::vector<std::vector<T>> k0(fluids.size());
        for (size_t k = 0; k < fluids.size(); ++k) {
            auto& ps = fluids[k].parts;
            int n = int(ps.size());
            k0[k].assign(n, T(0));
            T rho0 = fluids[k].rho0;
#pragma omp parallel for schedule(dynamic, dyn_chunk(n))
            for (int i = 0; i < n; ++i) {
                T d = densityChange(int(k), i);
                FluidPart<T>& a = ps[i];
                if (which == 0) {
                    bool on = a.rho + dt * d > rho0;
                    a.kappa_sum = on ? T(0.5) * a.kappa_sum : T(0);
                    k0[k][i] = a.kappa_sum / (dt * dt);
                } else {
                    /* where the solve itself would apply stiffness (DFSPH.cpp:829-832) */
                    bool on = d > 0 && !(a.rho + d * dt < rho0);
                    a.kappaV_sum = on ? T(P.warm_scale) * a.kappaV_sum : T(0);
                    k0[k][i] = a.kappaV_sum / dt;
                }
            }
        }
        df_velocityCorrection(which, dt, &k0);
    }

    /* pass A of the density loop, DFSPH.cpp:572-643: rho*, clamp, kappa.  Returns sum rho* (double). */
    double df_densityAdv(T dt, int& nOut) {
        T v0 = r * r * r;
        double sum = 0; nOut = 0;
        for (int k = 0; k < int(fluids.size()); ++k) {
            auto& ps = fluids[k].parts; int n = int(ps.size());
            T rho0 = fluids[k].rho0, fm0 = rho0 * v0;
#pragma omp parallel for schedule(dynamic, dyn_chunk(n))
            for (int i = 0; i < n; ++i) {
                FluidPart<T>& a = ps[i];
                T dadv = a.rho;
                const NeigbStr<T>& L = nbF[k][i];
                for (int j = 0; j < L.num; ++j) {
                    const auto& nb = L.neigs[j];
                    if (nb.dis == 0) continue;
                    T gs = -Wg(nb.dis);
                    if (nb.pidx.isFluid()) {
                        const FluidPart<T>& b = fp(nb.pidx);
                        V3<T> g = (a.x - b.x) * (gs / nb.dis);
                        dadv += g.dot(a.v_adv - b.v_adv) * fm0 * dt;
                    } else {
                        const BoundPart<T>& b = bp(nb.pidx);
                        V3<T> g = (a.x - b.x) * (gs / nb.dis);
                        dadv += g.dot(a.v_adv - b.v) * rho0 * b.volume * dt;
                    }
                }
                if (dadv < rho0) dadv = rho0;
                a.rho_adv = dadv;
            }
            /* serial sum (DFSPH.cpp:626-631), accumulated in double */
            for (int i = 0; i < n; ++i) sum += double(ps[i].rho_adv) / double(rho0);
            /* kappa (DFSPH.cpp:636-643) */
            for (int i = 0; i < n; ++i) {
                ps[i].kappa = (ps[i].rho_adv - rho0) / (dt * dt) * ps[i].alpha;
                ps[i].kappa_sum += ps[i].kappa * dt * dt;
            }
            nOut += n;
        }
        return sum;
    }

    /* DFSPH::DF_constantDensitySolverSecond, DFSPH.cpp:195-738 (vorticity options omitted) */
    void df_second(orc_stats* st) {
        T dt = dt_cur;
        df_forceExceptPressure();
        for (auto& f : fluids) {
            int n = int(f.parts.size());
#pragma omp parallel for
            for (int i = 0; i < n; ++i) f.parts[i].v_adv = f.parts[i].v + f.parts[i].a_adv * dt;
        }
        /* The density loop is not warm-started: its kappa only pushes
         * (rho* is clamped at rho0, DFSPH.cpp:620-623), so an over-estimated
         * initial guess is never taken back and adds energy. */
        for (auto& f : fluids) for (auto& p : f.parts) p.kappa_sum = 0;
        /* Dynamic rigid step 1 (DFSPH.cpp:90-104), here at the start of the
         * density solve rather than in DF_constantDensitySolverFirst: the
         * same forces reach updateSolids, and they stay readable after the
         * step */
        for (auto& so : solids) for (auto& b : so.parts) b.force = V3<T>();
        double err = 1; int it = 0;
        int maxIt = P.max_iter_density > 0 ? P.max_iter_density : std::numeric_limits<int>::max();
        while (err > double(P.density_tol) || it < P.min_iter_density) {
            if (it >= maxIt) { st->max_iter_hit = 1; break; }
            int n = 0;
            double s = df_densityAdv(dt, n);
            df_velocityCorrection(0, dt);
            df_rigidReaction();
            err = n > 0 ? s / n - 1.0 : 0.0;                              /* DFSPH.cpp:696 */
            ++it;
        }
        st->iter_density = it;
        st->density_err = err;
        for (auto& f : fluids) {
            int n = int(f.parts.size());
#pragma omp parallel for
            for (int i = 0; i < n; ++i) {
                f.parts[i].v = f.parts[i].v_adv;
                f.parts[i].x += f.parts[i].v * dt;
            }
            st->removed += removeOutside(f.parts);
        }
    }

    /* DFSPH::DF_divergenceFreeSolver, DFSPH.cpp:740-911 (stream correction omitted) */
    void df_divergence(orc_stats* st) {
        T v0 = r * r * r, dt = dt_cur;
        for (auto& f : fluids) for (auto& p : f.parts) p.v_adv = p.v;
        if (P.warm_start == 1 && frame > 0) warmStart(1, dt);
        else for (auto& f : fluids) for (auto& p : f.parts) p.kappaV_sum = 0;
        double avg = 0; int it = 0;
        while (avg > double(P.divergence_tol) || it < P.min_iter_div) {
            if (it > P.max_iter_div) { st->max_iter_hit = 1; break; }   /* DFSPH.cpp:772-773 */
            /* The first iteration of a warm-started solve clamps the
             * accumulated stiffness at zero rather than the increment, so
             * an over-large guess is taken back where the fluid now
             * expands; every later iteration is the reference's (increment
             * >= 0).  Projecting every iteration is unstable (signed Jacobi
             * oscillates), so only the guess is corrected this way. */
            const bool projected = P.warm_start != 0 && it == 0;
            double sum = 0; int ntot = 0;
            for (int k = 0; k < int(fluids.size()); ++k) {
                auto& ps = fluids[k].parts; int n = int(ps.size());
                T rho0 = fluids[k].rho0, fm0 = rho0 * v0;
#pragma omp parallel for schedule(dynamic, dyn_chunk(n))
                for (int i = 0; i < n; ++i) {
                    FluidPart<T>& a = ps[i];
                    T dd = 0;
                    const NeigbStr<T>& L = nbF[k][i];
                    for (int j = 0; j < L.num; ++j) {
                        const auto& nb = L.neigs[j];
                        if (nb.dis == 0) continue;
                        T gs = -Wg(nb.dis);
                        if (nb.pidx.isFluid()) {
                            const FluidPart<T>& b = fp(nb.pidx);
                            V3<T> g = (a.x - b.x) * (gs / nb.dis);
                            dd += g.dot(a.v_adv - b.v_adv) * fm0;
                        } else {
                            const BoundPart<T>& b = bp(nb.pidx);
                            V3<T> g = (a.x - b.x) * (gs / nb.dis);
                            dd += g.dot(a.v_adv - b.v) * (rho0 * b.volume);
                        }
                    }
                    if (!projected) {
                        if (dd < 0 || (a.rho + dd * dt) < rho0) dd = 0;   /* DFSPH.cpp:829-832 */
                        a.divdev = dd;
                        a.kappaV = dd * a.alpha / dt;                      /* DFSPH.cpp:835 */
                        a.kappaV_sum += a.kappaV * dt;
                    } else {
                        /* clamp the accumulated stiffness at zero instead of
                         * each increment: an over-large initial guess is
                         * taken back where the fluid now expands */
                        /* below rest density the solve applies no stiffness:
                         * the whole guess is taken back */
                        T inc = dd * a.alpha;
                        if ((a.rho + dd * dt) < rho0 || inc < -a.kappaV_sum) inc = -a.kappaV_sum;
                        a.kappaV = inc / dt;
                        a.kappaV_sum += inc;
                        a.divdev = std::abs(inc) / a.alpha;
                    }
                }
                for (int i = 0; i < n; ++i) sum += double(ps[i].divdev);
                ntot += n;
            }
            df_velocityCorrection(1, dt);
            avg = ntot > 0 ? sum / ntot : 0.0;
            ++it;
        }
        st->iter_divergence = it;
        st->divergence_err = avg;
        for (auto& f : fluids) for (auto& p : f.parts) p.v = p.v_adv;
    }

    /* ------------------------------ IISPH ------------------------------ */
    /* IISph_M::ii_forceExceptPressure + helpers, IISph_M.cpp:364-476 */
    void ii_forceExceptPressure() {
        T v0 = r * r * r, Hs = H, nu = T(P.nu), cs = T(P.sound_speed);
        int zd = 0;
        for (int k = 0; k < int(fluids.size()); ++k) {
            auto& ps = fluids[k].parts; int n = int(ps.size());
            T rho0 = fluids[k].rho0, fm0 = rho0 * v0, alpha = fluids[k].visc_alpha;
#pragma omp parallel for schedule(dynamic, dyn_chunk(n)) reduction(+ : zd)
            for (int i = 0; i < n; ++i) {
                FluidPart<T>& a = ps[i];
                a.a_adv = grav;
                const NeigbStr<T>& L = nbF[k][i];
                for (int j = 0; j < L.num; ++j) {
                    const auto& nb = L.neigs[j];
                    T dis = nb.dis;
                    if (dis == 0) continue;
                    if (nb.pidx.isFluid()) {
                        const FluidPart<T>& b = fp(nb.pidx);
                        int kb = nb.pidx.toFluidI();
                        V3<T> xab = a.x - b.x;
                        T pro = (a.v - b.v).dot(xab);
                        if (kb == k) {                                     /* IISph_M.cpp:364-385: no pro<0 gate */
                            V3<T> g2 = xab * (-Wg(dis) / dis);
                            a.a_adv += g2 * (pro / (dis * dis + T(0.01) * Hs * Hs) * fm0 / b.rho * 10 * nu);
                        } else {                                           /* IISph_M.cpp:388-405 */
                            T fmb = v0 * fluids[kb].rho0, al = (alpha + fluids[kb].visc_alpha) / 2;
                            if (pro < 0) {
                                T nuv = 2 * al * Hs * cs / (a.rho + b.rho);
                                T pi = -nuv * pro / (dis * dis + T(0.01) * Hs * Hs);
                                a.a_adv += xab * ((-Wg(dis) / dis) * (-(fm0 + fmb) / 2 * pi));
                            }
                        }
                    } else {                                               /* IISph_M.cpp:408-428 */
                        const BoundPart<T>& b = bp(nb.pidx);
                        V3<T> xab = a.x - b.x;
                        T pro = (a.v - b.v).dot(xab);
                        V3<T> g2 = xab * (-Wg(dis) / dis);
                        a.a_adv += g2 * (pro / (dis * dis + T(0.01) * Hs * Hs) * (rho0 * b.volume) / a.rho * 10 * nu);
                    }
                }
            }
        }
        zeroDis += zd;
    }

    /* IISph_M::ii_computePressure, IISph_M.cpp:78-360, all phases solved jointly */
    void ii_computePressure(orc_stats* st) {
        T v0 = r * r * r, dt = dt_cur;
        /* density first (IISph_M.cpp:101-112), so that a_adv reads this step's rho */
        for (int k = 0; k < int(fluids.size()); ++k) {
            auto& ps = fluids[k].parts; int n = int(ps.size());
            T rho0 = fluids[k].rho0, fm0 = rho0 * v0, wm0 = W0 * fm0;
#pragma omp parallel for schedule(dynamic, dyn_chunk(n))
            for (int i = 0; i < n; ++i) {
                FluidPart<T>& a = ps[i];
                T density = wm0;
                const NeigbStr<T>& L = nbF[k][i];
                for (int j = 0; j < L.num; ++j) {
                    const auto& nb = L.neigs[j];
                    if (nb.pidx.isFluid()) density += fm0 * W(nb.dis);
                    else density += rho0 * bp(nb.pidx).volume * W(nb.dis);
                }
                a.rho = density;
            }
        }
        ii_forceExceptPressure();
        /* d_ii and v_adv, IISph_M.cpp:114-135 */
        for (int k = 0; k < int(fluids.size()); ++k) {
            auto& ps = fluids[k].parts; int n = int(ps.size());
            T rho0 = fluids[k].rho0, fm0 = rho0 * v0;
#pragma omp parallel for schedule(dynamic, dyn_chunk(n))
            for (int i = 0; i < n; ++i) {
                FluidPart<T>& a = ps[i];
                V3<T> t1;
                const NeigbStr<T>& L = nbF[k][i];
                for (int j = 0; j < L.num; ++j) {
                    const auto& nb = L.neigs[j];
                    if (nb.dis == 0) continue;
                    if (nb.pidx.isFluid()) {
                        V3<T> g = (a.x - fp(nb.pidx).x) * (-Wg(nb.dis) / nb.dis);
                        t1 += g * (-fm0 / (a.rho * a.rho));
                    } else {
                        const BoundPart<T>& b = bp(nb.pidx);
                        V3<T> g = (a.x - b.x) * (-Wg(nb.dis) / nb.dis);
                        t1 += g * (-(rho0 * b.volume) / (a.rho * a.rho));
                    }
                }
                a.dii = t1 * (dt * dt);
                a.v_adv = a.v + a.a_adv * dt;
            }
        }
        /* a_ii, rho_adv, p_l, IISph_M.cpp:243-278 */
        for (int k = 0; k < int(fluids.size()); ++k) {
            auto& ps = fluids[k].parts; int n = int(ps.size());
            T rho0 = fluids[k].rho0, fm0 = rho0 * v0;
#pragma omp parallel for schedule(dynamic, dyn_chunk(n))
            for (int i = 0; i < n; ++i) {
                FluidPart<T>& a = ps[i];
                T t1 = 0, t2 = 0;
                const NeigbStr<T>& L = nbF[k][i];
                for (int j = 0; j < L.num; ++j) {
                    const auto& nb = L.neigs[j];
                    if (nb.dis == 0) continue;
                    if (nb.pidx.isFluid()) {
                        const FluidPart<T>& b = fp(nb.pidx);
                        V3<T> g = (a.x - b.x) * (-Wg(nb.dis) / nb.dis);
                        t1 += g.dot(a.v_adv - b.v_adv) * fm0 * dt;
                        /* d_ji (IISph_M.cpp:262: m_i); consistent: m_i^2 / m_j */
                        T fmb = v0 * fluids[nb.pidx.toFluidI()].rho0;
                        T cm = P.multiphase_consistent ? fm0 * fm0 / fmb : fm0;
                        V3<T> dji = g * (cm * dt * dt / (a.rho * a.rho));
                        t2 += (a.dii - dji).dot(g) * fm0;
                    } else {
                        const BoundPart<T>& b = bp(nb.pidx);
                        V3<T> g = (a.x - b.x) * (-Wg(nb.dis) / nb.dis);
                        t1 += g.dot(a.v_adv - b.v) * (rho0 * b.volume) * dt;  /* v_adv: IISph_M.cpp:269 uses v */
                        t2 += a.dii.dot(g) * (rho0 * b.volume);
                    }
                }
                if (t2 == 0) t2 = 1;
                a.aii = t2;
                a.rho_adv = a.rho + t1;
                a.p_l = T(0.5) * a.p;          /* warm start from p(t-dt), IISph_M.cpp:277 */
            }
        }
        /* relaxed Jacobi, IISph_M.cpp:280-358 */
        T omega = T(P.omega);
        int l = 0; double avgErr = 1e10;
        int maxIt = P.max_iter_density > 0 ? P.max_iter_density : std::numeric_limits<int>::max();
        int ntot = numFluid();
        while (avgErr > double(P.density_tol) || l < P.min_iter_density) {
            if (l >= maxIt) { st->max_iter_hit = 1; break; }
            for (int k = 0; k < int(fluids.size()); ++k) {
                auto& ps = fluids[k].parts; int n = int(ps.size());
#pragma omp parallel for schedule(dynamic, dyn_chunk(n))
                for (int i = 0; i < n; ++i) {
                    FluidPart<T>& a = ps[i];
                    V3<T> t1;
                    const NeigbStr<T>& L = nbF[k][i];
                    for (int j = 0; j < L.num; ++j) {
                        const auto& nb = L.neigs[j];
                        if (nb.dis == 0 || !nb.pidx.isFluid()) continue;
                        const FluidPart<T>& b = fp(nb.pidx);
                        /* d_ij p_j: IISph_M.cpp:302 takes the particle's own mass m_i; with
                         * multiphase_consistent the number-density force's m_j^2 / m_i */
                        T fmb = v0 * fluids[nb.pidx.toFluidI()].rho0, fma = v0 * fluids[k].rho0;
                        T cm = P.multiphase_consistent ? fmb * fmb / fma : fma;
                        V3<T> g = (a.x - b.x) * (-Wg(nb.dis) / nb.dis);
                        t1 += g * (-(cm * b.p_l) / (b.rho * b.rho));
                    }
                    a.sum_dijpj = t1 * (dt * dt);
                }
            }
            double sumErr = 0;
            std::vector<std::vector<T>> newps(fluids.size()), errss(fluids.size());
            for (int k = 0; k < int(fluids.size()); ++k) {
                auto& ps = fluids[k].parts; int n = int(ps.size());
                T rho0 = fluids[k].rho0, fm0 = rho0 * v0;
                std::vector<T>& newp = newps[k];
                std::vector<T>& errs = errss[k];
                newp.resize(n);
                errs.resize(n);
#pragma omp parallel for schedule(dynamic, dyn_chunk(n))
                for (int i = 0; i < n; ++i) {
                    FluidPart<T>& a = ps[i];
                    T fin = 0;
                    const NeigbStr<T>& L = nbF[k][i];
                    for (int j = 0; j < L.num; ++j) {
                        const auto& nb = L.neigs[j];
                        if (nb.dis == 0) continue;
                        if (nb.pidx.isFluid()) {
                            const FluidPart<T>& b = fp(nb.pidx);
                            V3<T> g = (a.x - b.x) * (-Wg(nb.dis) / nb.dis);
                            T fmb = v0 * fluids[nb.pidx.toFluidI()].rho0;
                            T cm = P.multiphase_consistent ? fm0 * fm0 / fmb : fm0;   /* IISph_M.cpp:329 */
                            V3<T> dji = g * (dt * dt * cm / (a.rho * a.rho));
                            V3<T> t1 = (a.sum_dijpj - b.dii * b.p_l) - (b.sum_dijpj - dji * a.p_l);
                            fin += fm0 * t1.dot(g);
                        } else {
                            const BoundPart<T>& b = bp(nb.pidx);
                            V3<T> g = (a.x - b.x) * (-Wg(nb.dis) / nb.dis);
                            fin += rho0 * b.volume * a.sum_dijpj.dot(g);
                        }
                    }
                    T t = (1 - omega) * a.p_l + (omega / a.aii) * (rho0 - a.rho_adv - fin);
                    if (t < 0) { t = 0; errs[i] = 0; }
                    else errs[i] = (a.rho_adv + a.aii * a.p_l + fin - rho0) / rho0;
                    newp[i] = t;
                }
            }
            /* Jacobi over all phases: publish after the whole sweep */
            for (int k = 0; k < int(fluids.size()); ++k) {
                auto& ps = fluids[k].parts;
                for (size_t i = 0; i < ps.size(); ++i) {
                    ps[i].p_l = newps[k][i]; ps[i].p = newps[k][i];
                    sumErr += double(errss[k][i]);
                }
            }
            avgErr = ntot > 0 ? sumErr / ntot : 0.0;
            ++l;
        }
        st->iter_density = l;
        st->density_err = avgErr;
    }

    /* IISph_M::ii_computeForcePressure, IISph_M.cpp:478-684 */
    void ii_computeForcePressure() {
        T v0 = r * r * r, Hs = H, cs = T(P.sound_speed);
        for (int k = 0; k < int(fluids.size()); ++k) {
            auto& ps = fluids[k].parts; int n = int(ps.size());
            T rho0 = fluids[k].rho0, fm0 = rho0 * v0, alpha = fluids[k].visc_alpha;
#pragma omp parallel for schedule(dynamic, dyn_chunk(n))
            for (int i = 0; i < n; ++i) {
                FluidPart<T>& a = ps[i];
                a.a_p = V3<T>();
                const NeigbStr<T>& L = nbF[k][i];
                for (int j = 0; j < L.num; ++j) {
                    const auto& nb = L.neigs[j];
                    T dis = nb.dis;
                    if (dis == 0) continue;
                    if (nb.pidx.isFluid()) {
                        const FluidPart<T>& b = fp(nb.pidx);
                        int kb = nb.pidx.toFluidI();
                        V3<T> xab = a.x - b.x;
                        T grad = -Wg(dis) / dis, acce;
                        if (kb == k) {
                            acce = grad * (-fm0 * (a.p / (a.rho * a.rho) + b.p / (b.rho * b.rho)));
                        } else {
                            T fmb = v0 * fluids[kb].rho0, al = (alpha + fluids[kb].visc_alpha) / 2;
                            T da = a.rho / fm0, db = b.rho / fmb;
                            acce = grad * (-1 / fm0 * (a.p / (da * da) + b.p / (db * db)));
                            T pro = (a.v - b.v).dot(xab);
                            if (pro < 0) {
                                T nuv = 2 * al * Hs * cs / (a.rho + b.rho);
                                T pi = -nuv * pro / (dis * dis + T(0.01) * Hs * Hs);
                                acce += grad * (-(fm0 + fmb) / 2 * pi);
                            }
                        }
                        a.a_p += xab * acce;
                    } else {
                        const BoundPart<T>& b = bp(nb.pidx);
                        V3<T> xab = a.x - b.x;
                        T grad = -Wg(dis) / dis, acce = 0;
                        /* p_a/rho_a^2, the boundary term the pressure solve predicts (IISph_M.cpp:127, 271, 338);
                         * IISph_M.cpp:558-559 applies 2 p_a/rho_a^2, twice the push the solve accounted for */
                        if (a.p > 0) acce += grad * (-rho0 * b.volume * (a.p / (a.rho * a.rho)));
                        a.a_p += xab * acce;
                    }
                }
            }
        }
    }

    int updateFluidsII() {
        int removed = 0;
        T dt = dt_cur;
        for (auto& f : fluids) {
            int n = int(f.parts.size());
#pragma omp parallel for
            for (int i = 0; i < n; ++i) {
                /* v(t+dt) = v_adv + dt a_p  (SphBase.cpp:387-388 drops a_adv) */
                f.parts[i].v = f.parts[i].v_adv + f.parts[i].a_p * dt;
                f.parts[i].x += f.parts[i].v * dt;
            }
            removed += removeOutside(f.parts);
        }
        return removed;
    }

    /* ------------------------------ step ------------------------------ */
    void cflDt() {
        if (P.cfl <= 0) { dt_cur = T(P.dt); return; }
        T vmax = 0;
        for (auto& f : fluids) for (auto& p : f.parts) vmax = std::max(vmax, std::sqrt(p.v.len_square()));
        T dt = T(P.dt);
        if (vmax > 0) dt = std::min(dt, T(P.cfl) * r / vmax);
        dt = std::max(dt, T(P.dt_min));
        dt_cur = dt;
    }

    /* ------------------------------ PCISPH ----------------------------- */
    /* PciSph::pci_updateDelta (PciSph.cpp:133-162): the SP09 stiffness from
     * a prototype particle with a full neighbourhood on the uniform lattice
     * of spacing r, delta = -1 / (beta (-|sum gradW|^2 - sum |gradW|^2)),
     * beta = 2 v0^2; the solve uses delta / dt^2.  gradW = x (dW/dr) / |x|. */
    T pciDelta() const {
        const T Hs = H;
        const int n = int(std::ceil(double(Hs) / double(r)));
        V3<T> gsum;
        T g2 = 0;
        for (int z = -n; z <= n; ++z)
            for (int y = -n; y <= n; ++y)
                for (int x = -n; x <= n; ++x) {
                    if (x == 0 && y == 0 && z == 0) continue;
                    V3<T> vf(T(x) * r, T(y) * r, T(z) * r);
                    T dis = std::sqrt(vf.len_square());
                    if (dis < Hs) {
                        T grad = -Wg(dis);
                        gsum += vf * (grad / dis);
                        g2 += grad * grad;
                    }
                }
        T v0 = r * r * r, beta = 2 * v0 * v0;
        return -1 / (beta * (-gsum.dot(gsum) - g2));
    }

    /* PciSph::pci_forceExceptPressure (PciSph.cpp:235-281) with the three
     * viscosity helpers (172-232): p = 0, a_p = 0, a = g + viscosity, the
     * densities of the previous step's last prediction */
    void pci_forceExceptPressure() {
        T v0 = r * r * r, Hs = H, cs = T(P.sound_speed);
        int zd = 0;
        for (int k = 0; k < int(fluids.size()); ++k) {
            auto& ps = fluids[k].parts; int n = int(ps.size());
            T rho0 = fluids[k].rho0, fm0 = rho0 * v0, alpha = fluids[k].visc_alpha;
#pragma omp parallel for schedule(dynamic, dyn_chunk(n)) reduction(+ : zd)
            for (int i = 0; i < n; ++i) {
                FluidPart<T>& a = ps[i];
                a.p = 0; a.a_p = V3<T>(); a.a = grav;
                const NeigbStr<T>& L = nbF[k][i];
                for (int j = 0; j < L.num; ++j) {
                    const auto& nb = L.neigs[j];
                    T dis = nb.dis;
                    if (dis == 0) continue;
                    T grad = -Wg(dis) / dis;
                    if (nb.pidx.isFluid()) {
                        const FluidPart<T>& b = fp(nb.pidx);
                        int kb = nb.pidx.toFluidI();
                        V3<T> xab = a.x - b.x;
                        T pro = (a.v - b.v).dot(xab);
                        if (pro >= 0) continue;
                        T fmb = v0 * fluids[kb].rho0;
                        T al = kb == k ? alpha : (alpha + fluids[kb].visc_alpha) / 2;
                        T nu = 2 * al * Hs * cs / (a.rho + b.rho);
                        T pi = -nu * pro / (dis * dis + T(0.01) * Hs * Hs);
                        T acce = grad * (kb == k ? -fm0 * pi : -(fm0 + fmb) / 2 * pi);
                        a.a += xab * acce;
                    } else {
                        const BoundPart<T>& b = bp(nb.pidx);
                        T ralpha = solids[nb.pidx.toSolidI()].visc_alpha;
                        V3<T> xab = a.x - b.x;
                        T pro = (a.v - b.v).dot(xab);
                        if (pro >= 0) continue;
                        T nu = 2 * ralpha * Hs * cs / (a.rho * 2);
                        T pi = -nu * pro / (dis * dis + T(0.01) * Hs * Hs);
                        a.a += xab * (grad * (-rho0 * b.volume * pi));
                    }
                }
            }
        }
        zeroDis += zd;
    }

    /* one prediction-correction iteration (PciSph.cpp:111-115):
     * pci_predictPosition (289-306), pci_predictDensity (310-352, negative
     * errors clamped: NEGLECT_NEGATIVE_PRESSURE, SphBase.h:32),
     * pci_forcePressure (399-436 with 356-397).  Returns the largest relative
     * density error; *avg the mean. */
    T pci_iteration(T deltaT, double* avg) {
        T v0 = r * r * r, dt = dt_cur, Hs = H;
        for (auto& f : fluids) {
            int n = int(f.parts.size());
#pragma omp parallel for
            for (int i = 0; i < n; ++i) {
                FluidPart<T>& q = f.parts[i];
                q.xp = q.x + (q.v + (q.a + q.a_p) * dt) * dt;
            }
        }
        T emax = 0; double esum = 0; long long cnt = 0;
        for (int k = 0; k < int(fluids.size()); ++k) {
            auto& ps = fluids[k].parts; int n = int(ps.size());
            T rho0 = fluids[k].rho0, fm0 = rho0 * v0, wm0 = W0 * fm0;
            T kmax = 0;
#pragma omp parallel for schedule(dynamic, dyn_chunk(n)) reduction(+ : esum, cnt) reduction(max : kmax)
            for (int i = 0; i < n; ++i) {
                FluidPart<T>& a = ps[i];
                T density = wm0;
                const NeigbStr<T>& L = nbF[k][i];
                for (int j = 0; j < L.num; ++j) {
                    const auto& nb = L.neigs[j];
                    if (nb.pidx.isFluid()) {
                        T dis = std::sqrt((a.xp - fp(nb.pidx).xp).len_square());
                        if (dis < Hs) density += fm0 * W(dis);     /* a's mass: SP08 number density */
                    } else {
                        const BoundPart<T>& b = bp(nb.pidx);
                        T dis = std::sqrt((a.xp - b.x).len_square());
                        if (dis < Hs) density += rho0 * b.volume * W(dis);
                    }
                }
                a.rho = density;
                T err = density - rho0;
                if (err < 0) err = 0;
                a.p += deltaT * err;
                kmax = std::max(kmax, err / rho0);
                esum += double(err / rho0); ++cnt;
            }
            emax = std::max(emax, kmax);
        }
        for (int k = 0; k < int(fluids.size()); ++k) {
            auto& ps = fluids[k].parts; int n = int(ps.size());
            T fm0 = fluids[k].rho0 * v0, rho0 = fluids[k].rho0;
#pragma omp parallel for schedule(dynamic, dyn_chunk(n))
            for (int i = 0; i < n; ++i) {
                FluidPart<T>& a = ps[i];
                a.a_p = V3<T>();
                const NeigbStr<T>& L = nbF[k][i];
                for (int j = 0; j < L.num; ++j) {
                    const auto& nb = L.neigs[j];
                    T dis = nb.dis;
                    if (dis == 0) continue;
                    T grad = -Wg(dis) / dis;
                    if (nb.pidx.isFluid()) {
                        const FluidPart<T>& b = fp(nb.pidx);
                        int kb = nb.pidx.toFluidI();
                        V3<T> xab = a.x - b.x;
                        T acce;
                        if (kb == k) {
                            acce = grad * (-fm0 * (a.p / (a.rho * a.rho) + b.p / (b.rho * b.rho)));
                        } else {                                   /* SP08, PciSph.cpp:369-381 */
                            T fmb = v0 * fluids[kb].rho0, da = a.rho / fm0, db = b.rho / fmb;
                            acce = grad * (-1 / fm0 * (a.p / (da * da) + b.p / (db * db)));
                        }
                        a.a_p += xab * acce;
                    } else if (a.p > 0) {                         /* PciSph.cpp:384-397 */
                        const BoundPart<T>& b = bp(nb.pidx);
                        a.a_p += (a.x - b.x) * (grad * (-rho0 * b.volume * (2 * a.p / (a.rho * a.rho))));
                    }
                }
            }
        }
        *avg = cnt ? esum / double(cnt) : 0.0;
        return emax;
    }

    /* PciSph::pci_computeForce (PciSph.cpp:97-126) and pci_forceTotal
     * (463-500): iterations while (max error >= eta || it < min) && it < max */
    void pci_computeForce(orc_stats* st) {
        const T deltaT = pciDelta() / (dt_cur * dt_cur);
        pci_forceExceptPressure();
        int it = 0; T emax = 0; double avg = 0;
        const int maxIt = P.max_iter_density > 0 ? P.max_iter_density : std::numeric_limits<int>::max();
        do {
            emax = pci_iteration(deltaT, &avg);
            ++it;
        } while ((emax >= T(P.density_tol) || it < P.min_iter_density) && it < maxIt);
        st->iter_density = it;
        st->density_err = double(emax);
        st->max_iter_hit = (emax >= T(P.density_tol) && it >= maxIt) ? 1 : 0;
        for (auto& f : fluids) {
            int n = int(f.parts.size());
#pragma omp parallel for
            for (int i = 0; i < n; ++i) f.parts[i].a += f.parts[i].a_p;
        }
    }

    /* ---- IISph class, header-consistent DFSPH (IISph.cpp:100-288, 816-1217) ----
     * Per-particle mass m = rho0 beta d^3 and rest density rho0 beta
     * (IISph.cpp:110-115); boundary particles carry mass boundary_density V_b
     * and rest density boundary_density (IISph.cpp:130; BoundPart::restDensity
     * is the scene's value, Demo-mySPH.h:64, and m_TH.boundaryDensity its
     * default, SphBase.h:576).  adjust_density (m_TH.adjustDensity,
     * SphBase.h:575) makes the fluid-fluid density and rho_adv terms use the
     * particle's own mass (a number density).  Boundary particles get their
     * own kappa and kappaV each iteration; the reference computes them for
     * m_Solids[k] with k the fluid index (IISph.cpp:847, 1069), here for every
     * solid.  Every particle has d = spacing_r, so each pair's kernel h =
     * (d_a + d_b)/2 is the one-argument kernel's. */
    T iidfMass(int k, const FluidPart<T>& q) const { return fluids[k].rho0 * q.beta * r * r * r; }
    T iidfRest(int k, const FluidPart<T>& q) const { return fluids[k].rho0 * q.beta; }
    T bDensity() const { return T(P.boundary_density); }
    T bMass(const BoundPart<T>& b) const { return bDensity() * b.volume; }

    /* IISph::DF_prepareAttribute, IISph.cpp:100-288: densities and alpha of
     * fluid and boundary particles */
    void iidf_prepare() {
        const bool adjust = P.adjust_density != 0;
        int zd = 0;
        for (int k = 0; k < int(fluids.size()); ++k) {
            auto& ps = fluids[k].parts; int n = int(ps.size());
#pragma omp parallel for schedule(dynamic, dyn_chunk(n)) reduction(+ : zd)
            for (int i = 0; i < n; ++i) {
                FluidPart<T>& a = ps[i];
                const T ma = iidfMass(k, a), resta = iidfRest(k, a);
                T density = W0 * ma, a2 = 0;
                V3<T> a1;
                const NeigbStr<T>& L = nbF[k][i];
                for (int j = 0; j < L.num; ++j) {
                    const auto& nb = L.neigs[j];
                    const T w = nb.dis == 0 ? W0 : W(nb.dis);
                    if (nb.pidx.isFluid()) {
                        const FluidPart<T>& b = fp(nb.pidx);
                        const T mb = iidfMass(nb.pidx.toFluidI(), b);
                        density += (adjust ? ma : mb) * w;                     /* IISph.cpp:163-164 */
                        if (nb.dis == 0) continue;
                        T gs = -Wg(nb.dis);
                        a1 += (a.x - b.x) * (gs / nb.dis) * mb;
                        a2 += (gs * mb) * (gs * mb);
                    } else {
                        const Solid<T>& so = solids[nb.pidx.toSolidI()];
                        const BoundPart<T>& b = so.parts[nb.pidx.i];
                        density += resta * b.volume * w;                        /* IISph.cpp:187 */
                        if (nb.dis == 0) continue;
                        T gs = -Wg(nb.dis);
                        a1 += (a.x - b.x) * (gs / nb.dis) * (bDensity() * b.volume);   /* IISph.cpp:192 */
                        if (so.dynamic) a2 += (gs * bMass(b)) * (gs * bMass(b));       /* IISph.cpp:193-195 */
                    }
                }
                a.rho = density;
                T den = a1.len_square() + a2;
                if (den < T(1.0e-6)) den = T(1.0e-6);
                a.alpha = density / den;
            }
        }
        zeroDis += zd;
        for (int k = 0; k < int(solids.size()); ++k) {                    /* IISph.cpp:234-287 */
            auto& ps = solids[k].parts; int n = int(ps.size());
#pragma omp parallel for schedule(dynamic, dyn_chunk(n))
            for (int i = 0; i < n; ++i) {
                BoundPart<T>& a = ps[i];
                const T ma = bMass(a);
                T density = W0 * ma, a2 = 0;
                const NeigbStr<T>& L = nbS[k][i];
                for (int j = 0; j < L.num; ++j) {
                    const auto& nb = L.neigs[j];
                    const T w = nb.dis == 0 ? W0 : W(nb.dis);
                    if (nb.pidx.isFluid()) {
                        const FluidPart<T>& b = fp(nb.pidx);
                        const T mb = iidfMass(nb.pidx.toFluidI(), b);
                        density += (P.adjust_density ? ma : mb) * w;
                        if (nb.dis == 0) continue;
                        T gs = -Wg(nb.dis);
                        a2 += (gs * mb) * (gs * mb);
                    } else {
                        density += bDensity() * bp(nb.pidx).volume * w;
                    }
                }
                a.rho = density;
                T den = a2 < T(1.0e-6) ? T(1.0e-6) : a2;
                a.alpha = density / den;
            }
        }
    }

    /* sum over a boundary particle's neighbours of gradW.(v_b - u_j) m_j,
     * with u = v_adv for fluid neighbours (the boundary passes of both loops,
     * IISph.cpp:859-899 with fluidMass, 1079-1110) */
    T iidfBoundaryRate(int k, int i, bool adjust) const {
        const BoundPart<T>& a = solids[k].parts[i];
        T d = 0;
        const NeigbStr<T>& L = nbS[k][i];
        for (int j = 0; j < L.num; ++j) {
            const auto& nb = L.neigs[j];
            if (nb.dis == 0) continue;
            T gs = -Wg(nb.dis);
            if (nb.pidx.isFluid()) {
                const FluidPart<T>& b = fp(nb.pidx);
                V3<T> g = (a.x - b.x) * (gs / nb.dis);
                d += g.dot(a.v - b.v_adv) * (adjust ? bMass(a) : iidfMass(nb.pidx.toFluidI(), b));
            } else {
                const BoundPart<T>& b = bp(nb.pidx);
                V3<T> g = (a.x - b.x) * (gs / nb.dis);
                d += g.dot(a.v - b.v) * (bDensity() * b.volume);
            }
        }
        return d;
    }

    /* the velocity pass of both loops (IISph.cpp:966-1001, 1166-1206):
     * v_adv -= dt sum gradW m_b (k_a/rho_a + k_b/rho_b), boundary neighbours
     * with their own kappa and mass boundary_density V_b */
    void iidf_velocity(int which, T dt) {
        for (int k = 0; k < int(fluids.size()); ++k) {
            auto& ps = fluids[k].parts; int n = int(ps.size());
#pragma omp parallel for schedule(dynamic, dyn_chunk(n))
            for (int i = 0; i < n; ++i) {
                FluidPart<T>& a = ps[i];
                const T ka = (which == 0 ? a.kappa : a.kappaV) / a.rho;
                const NeigbStr<T>& L = nbF[k][i];
                for (int j = 0; j < L.num; ++j) {
                    const auto& nb = L.neigs[j];
                    if (nb.dis == 0) continue;
                    T gs = -Wg(nb.dis);
                    if (nb.pidx.isFluid()) {
                        const FluidPart<T>& b = fp(nb.pidx);
                        const T kb = (which == 0 ? b.kappa : b.kappaV) / b.rho;
                        V3<T> g = (a.x - b.x) * (gs / nb.dis);
                        a.v_adv = a.v_adv - g * (iidfMass(nb.pidx.toFluidI(), b) * (ka + kb) * dt);
                    } else {
                        const BoundPart<T>& b = bp(nb.pidx);
                        const T kb = (which == 0 ? b.kappa : b.kappaV) / b.rho;
                        V3<T> g = (a.x - b.x) * (gs / nb.dis);
                        a.v_adv = a.v_adv - g * (bDensity() * b.volume * (ka + kb) * dt);
                    }
                }
            }
        }
    }

    /* IISph::DF_divergenceFreeSolver, IISph.cpp:1039-1217 */
    void iidf_divergence(orc_stats* st) {
        const T dt = dt_cur;
        for (auto& f : fluids) for (auto& p : f.parts) p.v_adv = p.v;
        double avg = 10; int it = 0;
        while (avg > double(P.divergence_tol) || it < P.min_iter_div) {
            if (it > P.max_iter_div) { st->max_iter_hit = 1; break; }     /* IISph.cpp:1062 */
            ++it;
            for (int k = 0; k < int(solids.size()); ++k) {                /* every solid (IISph.cpp:1069) */
                auto& ps = solids[k].parts; int n = int(ps.size());
#pragma omp parallel for
                for (int i = 0; i < n; ++i) {
                    BoundPart<T>& a = ps[i];
                    T dd = iidfBoundaryRate(k, i, false);
                    if (dd < 0 || (a.rho + dd * dt) < bDensity()) dd = 0;
                    a.kappaV = dd * a.alpha / dt;
                }
            }
            double sum = 0; int ntot = 0;
            for (int k = 0; k < int(fluids.size()); ++k) {
                auto& ps = fluids[k].parts; int n = int(ps.size());
#pragma omp parallel for schedule(dynamic, dyn_chunk(n))
                for (int i = 0; i < n; ++i) {
                    FluidPart<T>& a = ps[i];
                    const T resta = iidfRest(k, a);
                    T dd = 0;
                    const NeigbStr<T>& L = nbF[k][i];
                    for (int j = 0; j < L.num; ++j) {
                        const auto& nb = L.neigs[j];
                        if (nb.dis == 0) continue;
                        T gs = -Wg(nb.dis);
                        if (nb.pidx.isFluid()) {
                            const FluidPart<T>& b = fp(nb.pidx);
                            V3<T> g = (a.x - b.x) * (gs / nb.dis);
                            dd += g.dot(a.v_adv - b.v_adv) * iidfMass(nb.pidx.toFluidI(), b);
                        } else {
                            const BoundPart<T>& b = bp(nb.pidx);
                            V3<T> g = (a.x - b.x) * (gs / nb.dis);
                            dd += g.dot(a.v_adv - b.v) * (resta * b.volume);    /* IISph.cpp:1156 */
                        }
                    }
                    if (dd < 0 || (a.rho + dd * dt) < resta) dd = 0;          /* IISph.cpp:1160-1162 */
                    a.divdev = dd;
                    a.kappaV = dd * a.alpha / dt;
                }
                for (int i = 0; i < n; ++i) sum += double(ps[i].divdev);     /* IISph.cpp:1204-1205 */
                ntot += n;
            }
            iidf_velocity(1, dt);
            avg = ntot > 0 ? sum / ntot * double(dt) : 0.0;                  /* IISph.cpp:1207 */
        }
        st->iter_divergence = it;
        st->divergence_err = avg;
        for (auto& f : fluids) for (auto& p : f.parts) p.v = p.v_adv;
    }

    /* IISph::DF_constantDensitySolver, IISph.cpp:816-1037 */
    void iidf_density(orc_stats* st) {
        const T dt = dt_cur;
        const bool adjust = P.adjust_density != 0;
        for (auto& f : fluids) {
            int n = int(f.parts.size());
#pragma omp parallel for
            for (int i = 0; i < n; ++i) f.parts[i].v_adv = f.parts[i].v + f.parts[i].a_adv * dt;
        }
        double err = 1; int it = 0;
        while (err > double(P.density_tol) || it < P.min_iter_density) {
            if (it > P.max_iter_density) { st->max_iter_hit = 1; break; }   /* IISph.cpp:840 */
            ++it;
            for (int k = 0; k < int(solids.size()); ++k) {                /* every solid (IISph.cpp:847) */
                auto& ps = solids[k].parts; int n = int(ps.size());
#pragma omp parallel for
                for (int i = 0; i < n; ++i) {
                    BoundPart<T>& a = ps[i];
                    T radv = a.rho + iidfBoundaryRate(k, i, adjust) * dt;
                    if (radv < bDensity()) radv = bDensity();
                    a.rho_adv = radv;
                    a.kappa = (radv - bDensity()) / (dt * dt) * a.alpha;
                }
            }
            double sum = 0; int ntot = 0;
            for (int k = 0; k < int(fluids.size()); ++k) {
                auto& ps = fluids[k].parts; int n = int(ps.size());
#pragma omp parallel for schedule(dynamic, dyn_chunk(n))
                for (int i = 0; i < n; ++i) {
                    FluidPart<T>& a = ps[i];
                    const T ma = iidfMass(k, a), resta = iidfRest(k, a);
                    T radv = a.rho;
                    const NeigbStr<T>& L = nbF[k][i];
                    for (int j = 0; j < L.num; ++j) {
                        const auto& nb = L.neigs[j];
                        if (nb.dis == 0) continue;
                        T gs = -Wg(nb.dis);
                        if (nb.pidx.isFluid()) {
                            const FluidPart<T>& b = fp(nb.pidx);
                            V3<T> g = (a.x - b.x) * (gs / nb.dis);
                            radv += g.dot(a.v_adv - b.v_adv) * (adjust ? ma : iidfMass(nb.pidx.toFluidI(), b)) * dt;
                        } else {
                            const BoundPart<T>& b = bp(nb.pidx);
                            V3<T> g = (a.x - b.x) * (gs / nb.dis);
                            radv += g.dot(a.v_adv - b.v) * bDensity() * b.volume * dt;   /* IISph.cpp:948 */
                        }
                    }
                    if (radv < resta) radv = resta;
                    a.rho_adv = radv;
                    a.kappa = (radv - resta) / (dt * dt) * a.alpha;              /* IISph.cpp:962-965 */
                }
                for (int i = 0; i < n; ++i)                                   /* IISph.cpp:956-957 */
                    sum += (double(ps[i].rho_adv) - double(iidfRest(k, ps[i]))) / double(iidfRest(k, ps[i]));
                ntot += n;
            }
            iidf_velocity(0, dt);
            err = ntot > 0 ? sum / ntot : 0.0;                                  /* IISph.cpp:1004 */
        }
        st->iter_density = it;
        st->density_err = err;
        /* acce_presure = (v_adv - v)/dt (IISph.cpp:1008-1011), and updateFluids
         * under the IISph define: v += acce_presure dt, x += v dt
         * (SphBase.cpp:387-391) */
        for (auto& f : fluids) {
            int n = int(f.parts.size());
#pragma omp parallel for
            for (int i = 0; i < n; ++i) {
                FluidPart<T>& a = f.parts[i];
                a.a_p = (a.v_adv - a.v) * (T(1) / dt);
                a.v = a.v_adv;
                a.x += a.v * dt;
            }
            st->removed += removeOutside(f.parts);
        }
    }

    void fillStats(orc_stats* st) {
        double se = 0, me = -1e300, ke = 0; long long n = 0; int nan = 0;
        T v0 = r * r * r;
        for (auto& f : fluids)
            for (auto& p : f.parts) {
                double e = double(p.rho) / double(f.rho0) - 1.0;
                se += e; me = std::max(me, e); ++n;
                ke += 0.5 * double(f.rho0 * v0) * double(p.v.len_square());
                if (!std::isfinite(double(p.x.x + p.x.y + p.x.z + p.v.x + p.v.y + p.v.z))) ++nan;
            }
        st->rho_mean_err = n ? se / n : 0;
        st->rho_max_err = n ? me : 0;
        st->kinetic_energy = ke;
        st->nan_count = nan;
        st->dt = float(dt_cur);
    }

    /* update point (SphWithBullet::updateSolids): the host callback, then
     * the poses it (or the host before the step) set */
    int updateSolids(int (*cb)(void*), void* user) {
        int rc = cb ? cb(user) : 0;
        applyPendingPoses();
        return rc;
    }

    int step(orc_stats* st, int (*cb)(void*) = nullptr, void* user = nullptr) {
        std::memset(st, 0, sizeof(*st));
        int zd0 = zeroDis, tm0 = tooMany, rc = 0;
        cflDt();
        if (P.solver == ORC_WCSPH) {                                       /* WcSph.cpp:23-96 */
            applyPendingPoses();      /* poses set after the previous update point */
            neighbourSearch();
            updateSolidPartWeight();
            wc_computePressure();
            wc_computeForce();
            rigidReaction(false);
            st->removed = updateFluidsWC();
            rc = updateSolids(cb, user);                                   /* WcSph.cpp:85 */
            fillStats(st);
        } else if (P.solver == ORC_DFSPH) {                                /* DFSPH.cpp:9-75 */
            /* after an append the density solve needs lists and densities
             * that cover the new particles (in the reference the step after
             * a stepEvent emitter reads lists sized before it) */
            if (frame == 0 || needSearch) { neighbourSearch(); updateSolidPartWeight(); df_first(); }
            df_second(st);
            rc = updateSolids(cb, user);                                   /* DFSPH.cpp:54 */
            neighbourSearch();
            updateSolidPartWeight();
            df_first();
            df_divergence(st);
            fillStats(st);
        } else if (P.solver == ORC_IISPH_DF) {
            /* IISph::sphStep with the DF_* solvers in place of the VF_* ones
             * (the commented-out calls at IISph.cpp:59-60): search, V_b,
             * DF_prepareAttribute, the non-pressure acceleration (DFSPH's
             * DF_forceExceptPressure: the IISph class computes it in its VF_*
             * advection step, IISph.cpp:1511-1567), the divergence-free solve,
             * the constant-density solve, updateFluids, updateSolids */
            applyPendingPoses();
            neighbourSearch();
            updateSolidPartWeight();
            iidf_prepare();
            df_forceExceptPressure();
            iidf_divergence(st);
            iidf_density(st);
            rc = updateSolids(cb, user);
            fillStats(st);
        } else if (P.solver == ORC_PCISPH) {                               /* PciSph.cpp:31-95 */
            applyPendingPoses();
            if (frame == 0)      /* the host adds fluid at rest density (SphWithIO.cpp:1175) */
                for (auto& f : fluids) for (auto& q : f.parts) q.rho = f.rho0;
            neighbourSearch();
            updateSolidPartWeight();
            pci_computeForce(st);
            rigidReaction(false);           /* pci_boundPartForce_f = the WCSPH form, PciSph.cpp:440-460 */
            st->removed = updateFluidsWC();
            rc = updateSolids(cb, user);
            fillStats(st);
        } else {                                                           /* IISph_M.cpp:10-75 */
            applyPendingPoses();
            neighbourSearch();
            updateSolidPartWeight();
            ii_computePressure(st);
            ii_computeForcePressure();
            rigidReaction(true);
            st->removed = updateFluidsII();
            rc = updateSolids(cb, user);                                   /* IISph_M.cpp:64 */
            fillStats(st);
        }
        st->zero_dis = zeroDis - zd0;
        st->too_many_neigb = tooMany - tm0;
        st->dt = float(dt_cur);
        dt_prev = dt_cur;
        ++frame;
        return rc;
    }
};

struct SimBase {
    int bits;
    void* p;
};

}  // namespace

template <class T>
static void kernel_eval_t(double h, int n, const double* r, double* W, double* dW) {
    orc_params p{};
    p.spacing_r = float(h);
    p.smooth_radius_h = float(2 * h);
    p.dt = 1e-3f;
    Sim<T> S(p);
    S.kh = T(h);
    S.sigma = T(1) / T(M_PI) / (S.kh * S.kh * S.kh);
    S.sigma_h = S.sigma / S.kh;
    for (int i = 0; i < n; ++i) {
        W[i] = double(S.W(T(r[i])));
        dW[i] = double(S.Wg(T(r[i])));
    }
}

struct orc_sim {
    int bits;
    std::unique_ptr<Sim<float>> f;
    std::unique_ptr<Sim<double>> d;
};

template <class F>
static int dispatch(orc_sim* s, F&& fn) {
    if (!s) return -1;
    if (s->bits == 32) return fn(*s->f);
    return fn(*s->d);
}
template <class F>
static int dispatchc(const orc_sim* s, F&& fn) {
    if (!s) return -1;
    if (s->bits == 32) return fn(*s->f);
    return fn(*s->d);
}

extern "C" {

orc_sim* orc_create(int bits, const orc_params* p) {
    if (!p || (bits != 32 && bits != 64)) return nullptr;
    orc_sim* s = new orc_sim();
    s->bits = bits;
    if (bits == 32) s->f.reset(new Sim<float>(*p));
    else s->d.reset(new Sim<double>(*p));
    return s;
}

void orc_destroy(orc_sim* s) { delete s; }

int orc_add_fluid(orc_sim* s, int n, const float* xyz, const float* vel, const int* ids,
                  float rho0, float visc) {
    return dispatch(s, [&](auto& S) {
        using T = std::decay_t<decltype(S.r)>;
        Fluid<T> f; f.rho0 = T(rho0); f.visc_alpha = T(visc);
        f.parts.resize(n);
        for (int i = 0; i < n; ++i) {
            auto& q = f.parts[i];
            q.x = V3<T>(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
            if (vel) q.v = V3<T>(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
            q.d = T(S.P.spacing_r);
            q.id = ids ? ids[i] : i;
        }
        S.fluids.push_back(std::move(f));
        return int(S.fluids.size()) - 1;
    });
}

int orc_append_fluid(orc_sim* s, int k, int n, const float* xyz, const float* vel, const int* ids) {
    return dispatch(s, [&](auto& S) {
        using T = std::decay_t<decltype(S.r)>;
        if (k < 0 || k >= int(S.fluids.size()) || n < 0 || !xyz || !ids) return -1;
        auto& ps = S.fluids[k].parts;
        for (int i = 0; i < n; ++i) {
            FluidPart<T> q;
            q.x = V3<T>(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
            if (!q.x.inside(S.spMin, S.spMax)) return -1;
            if (vel) q.v = V3<T>(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
            q.d = T(S.P.spacing_r);
            q.id = ids[i];
            q.rho = S.fluids[k].rho0;
            ps.push_back(q);
        }
        S.needSearch = true;
        return 0;
    });
}

int orc_set_fluid_beta(orc_sim* s, int k, int n, const float* beta) {
    return dispatch(s, [&](auto& S) {
        using T = std::decay_t<decltype(S.r)>;
        if (k < 0 || k >= int(S.fluids.size()) || !beta || n != int(S.fluids[k].parts.size())) return -1;
        for (int i = 0; i < n; ++i) S.fluids[k].parts[i].beta = T(beta[i]);
        return 0;
    });
}

int orc_add_solid(orc_sim* s, int n, const float* xyz, int dynamic, float visc) {
    return dispatch(s, [&](auto& S) {
        using T = std::decay_t<decltype(S.r)>;
        Solid<T> so; so.dynamic = dynamic != 0; so.rigid = true; so.visc_alpha = T(visc);
        so.parts.resize(n);
        for (int i = 0; i < n; ++i) {
            so.parts[i].x = V3<T>(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
            so.parts[i].d = T(S.P.spacing_r);
        }
        if (so.dynamic) {   /* samples in body coordinates; identity pose until one is set */
            so.inner.resize(n);
            for (int i = 0; i < n; ++i) so.inner[i] = so.parts[i].x;
        }
        S.solids.push_back(std::move(so));
        return int(S.solids.size()) - 1;
    });
}

int orc_set_solid_pose(orc_sim* s, int k, const float* T16, const float* com, const float* v, const float* w) {
    return dispatch(s, [&](auto& S) {
        if (k < 0 || k >= int(S.solids.size()) || !S.solids[k].dynamic) return -1;
        auto& so = S.solids[k];
        std::memcpy(so.pT, T16, sizeof(so.pT));
        std::memcpy(so.pCom, com, sizeof(so.pCom));
        std::memcpy(so.pV, v, sizeof(so.pV));
        std::memcpy(so.pW, w, sizeof(so.pW));
        so.posePending = true;
        return 0;
    });
}

int orc_solid_force_torque(orc_sim* s, int k, double* f, double* t) {
    return dispatch(s, [&](auto& S) {
        if (k < 0 || k >= int(S.solids.size())) return -1;
        S.rigidForceJoin(k, f, t);
        return 0;
    });
}

int orc_init(orc_sim* s) {
    return dispatch(s, [&](auto& S) {
        // SphWithIO::addFluidCuboid / SphSolid::adjustSpace grow the domain to
        // contain every particle (SphWithIO.cpp:161-162, SphBase.cpp:149-161);
        // here the domain is given, so particles outside it are rejected.
        S.applyPendingPoses();
        for (auto& f : S.fluids)
            for (auto& q : f.parts)
                if (!q.x.inside(S.spMin, S.spMax)) return -1;
        for (auto& so : S.solids)
            for (auto& q : so.parts)
                if (!q.x.inside(S.spMin, S.spMax)) return -1;
        S.prepareSolidPartWeight();
        return 0;
    });
}

int orc_step(orc_sim* s, orc_stats* st) {
    return dispatch(s, [&](auto& S) { return S.step(st); });
}

int orc_pci_delta(const orc_sim* s, double* delta) {
    return dispatchc(s, [&](const auto& S) { *delta = double(S.pciDelta()); return 0; });
}

int orc_step_dt(const orc_sim* s, double* dt) {
    return dispatchc(s, [&](const auto& S) { *dt = double(S.dt_cur); return 0; });
}

int orc_step_cb(orc_sim* s, orc_stats* st, int (*cb)(void*), void* user) {
    return dispatch(s, [&](auto& S) { return S.step(st, cb, user); });
}

int orc_num_fluid(const orc_sim* s, int k) {
    return dispatchc(s, [&](const auto& S) {
        return (k < 0 || k >= int(S.fluids.size())) ? -1 : int(S.fluids[k].parts.size());
    });
}

int orc_get_fluid(const orc_sim* s, int k, int field, double* out) {
    return dispatchc(s, [&](const auto& S) {
        if (k < 0 || k >= int(S.fluids.size())) return -1;
        const auto& ps = S.fluids[k].parts;
        size_t n = ps.size();
        for (size_t i = 0; i < n; ++i) {
            const auto& q = ps[i];
            switch (field) {
                case ORC_F_POS: out[3 * i] = q.x.x; out[3 * i + 1] = q.x.y; out[3 * i + 2] = q.x.z; break;
                case ORC_F_VEL: out[3 * i] = q.v.x; out[3 * i + 1] = q.v.y; out[3 * i + 2] = q.v.z; break;
                case ORC_F_RHO: out[i] = q.rho; break;
                case ORC_F_PRESSURE: out[i] = q.p; break;
                case ORC_F_ACC:
                    if (S.P.solver == ORC_IISPH) { out[3 * i] = q.a_p.x; out[3 * i + 1] = q.a_p.y; out[3 * i + 2] = q.a_p.z; }
                    else { out[3 * i] = q.a.x; out[3 * i + 1] = q.a.y; out[3 * i + 2] = q.a.z; }
                    break;
                case ORC_F_ALPHA: out[i] = q.alpha; break;
                case ORC_F_KAPPA: out[i] = q.kappa; break;
                case ORC_F_ID: out[i] = q.id; break;
                case ORC_F_AADV: out[3 * i] = q.a_adv.x; out[3 * i + 1] = q.a_adv.y; out[3 * i + 2] = q.a_adv.z; break;
                case ORC_F_KAPPA_SUM: out[i] = q.kappa_sum; break;
                case ORC_F_AII: out[i] = q.aii; break;
                case ORC_F_RHO_ADV: out[i] = q.rho_adv; break;
                case ORC_F_KAPPAV: out[i] = q.kappaV; break;
                case ORC_F_DIVDEV: out[i] = q.divdev; break;
                case ORC_F_VADV: out[3 * i] = q.v_adv.x; out[3 * i + 1] = q.v_adv.y; out[3 * i + 2] = q.v_adv.z; break;
                default: return -2;
            }
        }
        return int(n);
    });
}

int orc_set_fluid_state(orc_sim* s, int k, const double* xyz, const double* vel) {
    return dispatch(s, [&](auto& S) {
        using T = std::decay_t<decltype(S.r)>;
        if (k < 0 || k >= int(S.fluids.size()) || !xyz) return -1;
        auto& ps = S.fluids[k].parts;
        for (size_t i = 0; i < ps.size(); ++i) {
            ps[i].x = V3<T>(T(xyz[3 * i]), T(xyz[3 * i + 1]), T(xyz[3 * i + 2]));
            if (vel) ps[i].v = V3<T>(T(vel[3 * i]), T(vel[3 * i + 1]), T(vel[3 * i + 2]));
        }
        return int(ps.size());
    });
}

int orc_get_boundary_volume(const orc_sim* s, int k, double* out) {
    return dispatchc(s, [&](const auto& S) {
        if (k < 0 || k >= int(S.solids.size())) return -1;
        const auto& ps = S.solids[k].parts;
        for (size_t i = 0; i < ps.size(); ++i) out[i] = ps[i].volume;
        return int(ps.size());
    });
}

int orc_neighbours(orc_sim* s, int k, int cap, int* counts, int* kinds, int* ids, float* dis) {
    return dispatch(s, [&](auto& S) {
        int before = S.tooMany;
        S.neighbourSearch();
        if (k < 0 || k >= int(S.fluids.size())) return -1;
        std::vector<int> solidBase(S.solids.size() + 1, 0);
        for (size_t q = 0; q < S.solids.size(); ++q) solidBase[q + 1] = solidBase[q] + int(S.solids[q].parts.size());
        const auto& L = S.nbF[k];
        for (size_t i = 0; i < L.size(); ++i) {
            int c = std::min(L[i].num, cap);
            counts[i] = L[i].num;
            for (int j = 0; j < c; ++j) {
                const auto& nb = L[i].neigs[j];
                size_t o = i * size_t(cap) + j;
                if (nb.pidx.isFluid()) { kinds[o] = 0; ids[o] = S.fp(nb.pidx).id; }
                else { kinds[o] = 1; ids[o] = solidBase[nb.pidx.toSolidI()] + nb.pidx.i; }
                dis[o] = float(nb.dis);
            }
        }
        return S.tooMany - before;
    });
}

int orc_compute_density(orc_sim* s) {
    return dispatch(s, [&](auto& S) {
        S.neighbourSearch();
        S.updateSolidPartWeight();
        if (S.P.solver == ORC_IISPH_DF) S.iidf_prepare();
        else S.df_first();
        return 0;
    });
}

int orc_get_boundary_field(const orc_sim* s, int k, int field, double* out) {
    return dispatchc(s, [&](const auto& S) {
        if (k < 0 || k >= int(S.solids.size()) || field < 0 || field > 4) return -1;
        const auto& ps = S.solids[k].parts;
        for (size_t i = 0; i < ps.size(); ++i) {
            const auto& q = ps[i];
            out[i] = field == 0 ? q.rho : field == 1 ? q.alpha : field == 2 ? q.kappa : field == 3 ? q.rho_adv : q.kappaV;
        }
        return int(ps.size());
    });
}

void orc_kernel_eval(int bits, double h, int n, const double* r, double* W, double* dW) {
    if (bits == 32) kernel_eval_t<float>(h, n, r, W, dW);
    else kernel_eval_t<double>(h, n, r, W, dW);
}

void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
int orc_get_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
