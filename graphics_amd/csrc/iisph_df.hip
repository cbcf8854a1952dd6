// The IISph class's header-consistent DFSPH (IISph.cpp:100-288
// DF_prepareAttribute, 816-1037 DF_constantDensitySolver, 1039-1217
// DF_divergenceFreeSolver), one thread per particle.
//
// Against the DFSPH class (solvers.hip) it differs in:
//   - a per-particle mass m = rho0 beta d^3 and rest density rho0 beta
//     (IISph.cpp:110-115); beta rides in the ksum array (reordered with the
//     particles, unused by this solver otherwise), the mass of each slot is
//     recomputed every step into pm;
//   - m_TH.adjustDensity (Dev::adjustDensity): the fluid-fluid density and
//     predicted-density terms use the particle's own mass (a number density);
//   - boundary terms use m_TH.boundaryDensity (Dev::boundaryDensity), which is
//     also every boundary particle's rest density, and boundary particles
//     carry their own density, alpha and kappa (bk = kappa/rho per boundary
//     particle), which the fluid's velocity pass reads;
//   - the divergence-free solve runs first, at the same positions, then the
//     constant-density solve, then updateFluids (the order of the commented
//     calls at IISph.cpp:59-60).
// Boundary particles find their neighbours by scanning the 27 cells around
// them (fluid through cellF, boundary through cellB, the reference's
// predicate, no pair with a particle of the same solid: every solid is
// rigid); fluid particles use their neighbour lists.
#include "sph_common.h"

namespace sphgpu {

namespace {

__shared__ int g_idftab[NB16_TAB];

struct SlotI {
    int i, tile, scene;
    bool live, packed;
};

__device__ __forceinline__ SlotI slot_idf(const Dev& D, bool gated) {
    SlotI s;
    s.tile = tile_of_block(D.tileMap);
    s.i = s.tile * TILE + threadIdx.x;
    s.scene = D.tileScene[s.tile];
    s.live = (int)threadIdx.x < D.tileValid[s.tile];
    if (gated) s.live = s.live && D.active[s.scene] != 0;
    if (threadIdx.x < NB16_TAB) g_idftab[threadIdx.x] = D.nbTab[(size_t)s.tile * NB16_TAB + threadIdx.x];
    __syncthreads();
    s.packed = __builtin_amdgcn_readfirstlane(g_idftab[NB16_FLAG]) != 0;
    return s;
}

__device__ __forceinline__ int nf_cnt(int c) { return c & 0xff; }
__device__ __forceinline__ int nb_cnt(int c) { return (c >> 16) & 0xff; }

// every listed fluid neighbour j of slot s (padding entries, the particle
// itself, included: they contribute nothing to gradient sums and are skipped
// by index in density sums)
template <class F>
__device__ __forceinline__ void fluid_list(const Dev& D, const SlotI& s, int n, F&& f) {
    if (s.packed) {
        const int4* col = D.nbF16 + s.i;
        const int n4 = (n + 3) & ~3;
        for (int k = 0; k < n4; ++k) {
            const int4 g = col[(size_t)(k >> 3) * D.ncap];
            const int w = (k & 7) >> 1;
            const unsigned word = (unsigned)(w == 0 ? g.x : w == 1 ? g.y : w == 2 ? g.z : g.w);
            const unsigned e = (word >> (16 * (k & 1))) & 0xffffu;
            f(g_idftab[e >> NB16_OFF_BITS] + (int)e);
        }
    } else {
        const int* col = D.nbF;
        const int n4 = (n + 3) & ~3;
        for (int k = 0; k < n4; ++k) f(col[nb_slot(k, s.i, D.ncap)]);
    }
}

template <class F>
__device__ __forceinline__ void boundary_list(const Dev& D, const SlotI& s, int n, F&& f) {
    if (s.packed) {
        const int4* col = D.nbB16 + s.i;
        for (int k = 0; k < n; ++k) {
            const int4 g = col[(size_t)(k >> 3) * D.ncap];
            const int w = (k & 7) >> 1;
            const unsigned word = (unsigned)(w == 0 ? g.x : w == 1 ? g.y : w == 2 ? g.z : g.w);
            const unsigned e = (word >> (16 * (k & 1))) & 0xffffu;
            f(g_idftab[32 + (e >> NB16_OFF_BITS)] + (int)e);
        }
    } else {
        for (int k = 0; k < n; ++k) f(D.nbB[nb_slot(k, s.i, D.ncap)]);
    }
}

__device__ __forceinline__ float rest_of(const Dev& D, int i, int ph) { return D.rho0[ph] * D.ksum[i]; }

// pair vector and the kernel gradient factor -|dW/dr|/r
struct Pair {
    float dx, dy, dz, r, g;
};
__device__ __forceinline__ Pair pair_of(const Dev& D, float4 a, float4 b) {
    Pair p;
    p.dx = a.x - b.x;
    p.dy = a.y - b.y;
    p.dz = a.z - b.z;
    p.r = fast_sqrt(dot3f(p.dx, p.dy, p.dz, p.dx, p.dy, p.dz));
    p.g = kernel_gradfac(p.r, D);
    return p;
}

// ---------------------------------------------------------------- masses ----
// m = rho0 beta v0 of every live slot (IISph.cpp:110-115)
__global__ __launch_bounds__(256) void k_idf_mass(Dev D) {
    const int tile = blockIdx.x, i = tile * TILE + threadIdx.x;
    if ((int)threadIdx.x >= D.tileValid[tile]) return;
    const int ph = (int)D.x[i].w;
    D.pm[i] = D.rho0[ph] * D.ksum[i] * D.v0;
}

// ------------------------------------------------------ DF_prepareAttribute --
// fluid density and alpha (IISph.cpp:143-200); publishes rho in v.w for the
// non-pressure acceleration pass
__global__ __launch_bounds__(256) void k_idf_prepare_fluid(Dev D) {
    const SlotI s = slot_idf(D, false);
    if (!s.live) return;
    const int i = s.i;
    const float4 pa = D.x[i];
    const int ph = (int)pa.w;
    const float ma = D.pm[i], resta = rest_of(D, i, ph);
    const int c = D.cnt[i];
    Kahan rho;
    rho.init(ma * D.W0);
    float a1x = 0.f, a1y = 0.f, a1z = 0.f, a2 = 0.f;
    fluid_list(D, s, nf_cnt(c), [&](int j) {
        if (j == i) return;
        const Pair p = pair_of(D, pa, D.x[j]);
        const float mb = D.pm[j];
        rho.add((D.adjustDensity ? ma : mb) * kernel_W(p.r, D));     // IISph.cpp:163-164
        const float gm = p.g * mb;
        a1x += p.dx * gm; a1y += p.dy * gm; a1z += p.dz * gm;
        const float gs = gm * p.r;                                   // |dW/dr| m_b
        a2 += gs * gs;
    });
    boundary_list(D, s, nb_cnt(c), [&](int j) {
        const float4 pb = D.xb[j];
        const Pair p = pair_of(D, pa, pb);
        rho.add(resta * pb.w * kernel_W(p.r, D));                     // IISph.cpp:187
        const float gm = p.g * D.boundaryDensity * pb.w;             // IISph.cpp:192
        a1x += p.dx * gm; a1y += p.dy * gm; a1z += p.dz * gm;
        if (D.anyDynamic && j < D.nb && D.solids[D.bsolid[j]].dynamic) {   // IISph.cpp:193-195
            const float gs = gm * p.r;
            a2 += gs * gs;
        }
    });
    const float density = rho.s;
    float den = dot3f(a1x, a1y, a1z, a1x, a1y, a1z) + a2;
    if (den < 1.0e-6f) den = 1.0e-6f;
    D.rho[i] = density;
    D.alpha[i] = density / den;
    D.v[i].w = density;
}

// every fluid and boundary neighbour of boundary particle b, from the 27
// cells around it: fluid(j, pair), boundary(j, pair, V_j)
template <class FF, class FB>
__device__ __forceinline__ void boundary_stencil(const Dev& D, int b, float4 pa, FF&& fluid, FB&& bound) {
    const SceneGrid G = D.grids[D.bscene[b]];
    const int sa = D.bsolid[b];
    int cx, cy, cz;
    cell_of(G, pa, cx, cy, cz);
    for (int dz = -1; dz <= 1; ++dz) {
        const int z = cz + dz;
        if (z < 0 || z >= G.nz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = cy + dy;
            if (y < 0 || y >= G.ny) continue;
            const uint32_t k0 = key_of(G, D.rowRank, cx, y, z);
            const int2 rf = run3(D.cellF, k0, cx, G.nx);
            for (int j = rf.x; j < rf.y; ++j) {
                const float4 pb = D.x[j];
                if (dis_square_exact(pa, pb) < D.h2) fluid(j, pair_of(D, pa, pb));
            }
            const int2 rb = run3(D.cellB, k0, cx, G.nx);
            for (int j = rb.x; j < rb.y; ++j) {
                if (j == b || D.bsolid[j] == sa) continue;
                const float4 pb = D.xb[j];
                if (dis_square_exact(pa, pb) < D.h2) bound(j, pair_of(D, pa, pb), pb.w);
            }
        }
    }
}

// boundary density and alpha (IISph.cpp:234-287); the mass of a boundary
// particle is boundaryDensity V_b (IISph.cpp:130)
__global__ __launch_bounds__(256) void k_idf_prepare_boundary(Dev D) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= D.nb) return;
    const float4 pa = D.xb[b];
    const float ma = D.boundaryDensity * pa.w;
    Kahan rho;
    rho.init(ma * D.W0);
    float a2 = 0.0f;
    boundary_stencil(D, b, pa,
        [&](int j, const Pair& p) {
            const float mf = D.pm[j];
            rho.add((D.adjustDensity ? ma : mf) * kernel_W(p.r, D));
            const float gs = p.g * p.r * mf;
            a2 += gs * gs;
        },
        [&](int, const Pair& p, float vol) { rho.add(D.boundaryDensity * vol * kernel_W(p.r, D)); });
    const float density = rho.s;
    D.brho[b] = density;
    D.balpha[b] = density / (a2 < 1.0e-6f ? 1.0e-6f : a2);
}

// -------------------------------------------------------- solver passes ----
// mode 0: v_adv = v (the divergence solve starts from v, IISph.cpp:1056);
// mode 1: v = the divergence-free v_adv, v_adv = v + dt a_adv
// (IISph.cpp:833-836).  v_adv.w carries the particle's mass for the
// neighbours' gathers.
__global__ __launch_bounds__(256) void k_idf_vadv(Dev D, int mode) {
    const int tile = blockIdx.x, i = tile * TILE + threadIdx.x;
    if ((int)threadIdx.x >= D.tileValid[tile]) return;
    const float dt = D.dt[D.tileScene[tile]];
    if (mode == 0) {
        const float4 v = D.v[i];
        D.vs[i] = make_float4(v.x, v.y, v.z, D.pm[i]);
    } else {
        const float4 u = D.vs[i], a = D.f4a[i];
        D.v[i] = make_float4(u.x, u.y, u.z, D.rho[i]);
        D.vs[i] = make_float4(u.x + a.x * dt, u.y + a.y * dt, u.z + a.z * dt, u.w);
    }
}

// boundary particles' own stiffness, kappa/rho into bk:
// mode 0 density (IISph.cpp:849-904), mode 1 divergence (1069-1115)
__global__ __launch_bounds__(256) void k_idf_kappa_boundary(Dev D, int mode) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= D.nb) return;
    const int sc = D.bscene[b];
    if (!D.active[sc]) return;
    const float4 pa = D.xb[b], va = D.vb[b];
    const float ma = D.boundaryDensity * pa.w, rest = D.boundaryDensity;
    const bool adjust = mode == 0 && D.adjustDensity;
    float sf = 0.0f, sb = 0.0f;
    boundary_stencil(D, b, pa,
        [&](int j, const Pair& p) {
            const float4 u = D.vs[j];
            sf += (adjust ? ma : u.w) * p.g * dot3f(va.x - u.x, va.y - u.y, va.z - u.z, p.dx, p.dy, p.dz);
        },
        [&](int j, const Pair& p, float vol) {
            const float4 u = D.vb[j];
            sb += vol * p.g * dot3f(va.x - u.x, va.y - u.y, va.z - u.z, p.dx, p.dy, p.dz);
        });
    const float dt = D.dt[sc], rho = D.brho[b];
    const float d = sf + D.boundaryDensity * sb;
    float kr;
    if (mode == 0) {
        float radv = rho + d * dt;
        if (radv < rest) radv = rest;
        kr = (radv - rest) / (dt * dt) * D.balpha[b] / rho;
    } else {
        const float dd = (d < 0.0f || (rho + d * dt) < rest) ? 0.0f : d;
        kr = dd * D.balpha[b] / dt / rho;
    }
    D.bk[b] = kr;
}

// fluid stiffness (mode 0 density, IISph.cpp:910-965; mode 1 divergence,
// 1120-1165) -> xk = (x, kappa/rho); partial sums of the loop's error:
// (rho_adv - rest)/rest (IISph.cpp:956-957), or Drho/Dt dt (1204-1207)
__global__ __launch_bounds__(256) void k_idf_kappa_fluid(Dev D, int mode) {
    __shared__ double lds[4];
    const SlotI s = slot_idf(D, true);
    double contrib = 0.0;
    if (s.live) {
        const int i = s.i;
        const float4 pa = D.x[i], va = D.vs[i];
        const int ph = (int)pa.w;
        const float ma = va.w, rest = rest_of(D, i, ph);
        const int c = D.cnt[i];
        const bool adjust = mode == 0 && D.adjustDensity;
        float sf = 0.0f, sb = 0.0f;
        fluid_list(D, s, nf_cnt(c), [&](int j) {
            const float4 u = D.vs[j];
            const Pair p = pair_of(D, pa, D.x[j]);
            sf += (adjust ? ma : u.w) * p.g * dot3f(va.x - u.x, va.y - u.y, va.z - u.z, p.dx, p.dy, p.dz);
        });
        boundary_list(D, s, nb_cnt(c), [&](int j) {
            const float4 pb = D.xb[j], u = D.vb[j];
            const Pair p = pair_of(D, pa, pb);
            sb += pb.w * p.g * dot3f(va.x - u.x, va.y - u.y, va.z - u.z, p.dx, p.dy, p.dz);
        });
        const float dt = D.dt[s.scene], rho = D.rho[i];
        float kr;
        if (mode == 0) {
            float radv = rho + dt * (sf + D.boundaryDensity * sb);      // IISph.cpp:948
            if (radv < rest) radv = rest;
            kr = (radv - rest) / (dt * dt) * D.alpha[i] / rho;
            contrib = ((double)radv - (double)rest) / (double)rest;
        } else {
            float dd = sf + rest * sb;                                  // IISph.cpp:1156
            if (dd < 0.0f || (rho + dd * dt) < rest) dd = 0.0f;
            kr = dd * D.alpha[i] / dt / rho;
            contrib = (double)dd * (double)dt;
        }
        D.xk[i] = make_float4(pa.x, pa.y, pa.z, kr);
    }
    const double t = block_sum_256(contrib, lds);
    if (threadIdx.x == 0) D.partial[2 * s.tile] = t;
}

// velocity pass of both loops (IISph.cpp:966-1001, 1166-1206):
// v_adv -= dt (sum_f m_b (k_a/rho_a + k_b/rho_b) gradW
//              + sum_b boundaryDensity V_b (k_a/rho_a + k_b/rho_b) gradW)
__global__ __launch_bounds__(256) void k_idf_velocity(Dev D) {
    const SlotI s = slot_idf(D, true);
    if (!s.live) return;
    const int i = s.i;
    const float4 pa = D.xk[i];
    const float ka = pa.w;
    const int c = D.cnt[i];
    float fx = 0.f, fy = 0.f, fz = 0.f, bx = 0.f, by = 0.f, bz = 0.f;
    fluid_list(D, s, nf_cnt(c), [&](int j) {
        const float4 pb = D.xk[j];
        const Pair p = pair_of(D, pa, pb);
        const float w = p.g * D.pm[j] * (ka + pb.w);
        fx += p.dx * w; fy += p.dy * w; fz += p.dz * w;
    });
    boundary_list(D, s, nb_cnt(c), [&](int j) {
        const float4 pb = D.xb[j];
        const Pair p = pair_of(D, pa, pb);
        const float w = p.g * pb.w * (ka + D.bk[j]);
        bx += p.dx * w; by += p.dy * w; bz += p.dz * w;
    });
    const float dt = D.dt[s.scene], cb = dt * D.boundaryDensity;
    float4 v = D.vs[i];
    v.x -= dt * fx + cb * bx;
    v.y -= dt * fy + cb * by;
    v.z -= dt * fz + cb * bz;
    D.vs[i] = v;
}

inline int tiles_of(const Dev& D) { return D.npad / TILE; }
inline int bblocks(const Dev& D) { return (D.nb + 255) / 256; }

}  // namespace

void launch_idf_prepare(const Dev& D, hipStream_t st) {
    const int t = tiles_of(D);
    if (t > 0) {
        hipLaunchKernelGGL(k_idf_mass, dim3(t), dim3(256), 0, st, D);
        hipLaunchKernelGGL(k_idf_prepare_fluid, dim3(t), dim3(256), 0, st, D);
    }
    if (D.nb > 0) hipLaunchKernelGGL(k_idf_prepare_boundary, dim3(bblocks(D)), dim3(256), 0, st, D);
}

void launch_idf_vadv(const Dev& D, int mode, hipStream_t st) {
    if (tiles_of(D) > 0) hipLaunchKernelGGL(k_idf_vadv, dim3(tiles_of(D)), dim3(256), 0, st, D, mode);
}

void launch_idf_kappa(const Dev& D, int mode, hipStream_t st) {
    if (D.nb > 0) hipLaunchKernelGGL(k_idf_kappa_boundary, dim3(bblocks(D)), dim3(256), 0, st, D, mode);
    if (tiles_of(D) > 0) hipLaunchKernelGGL(k_idf_kappa_fluid, dim3(tiles_of(D)), dim3(256), 0, st, D, mode);
}

void launch_idf_velocity(const Dev& D, hipStream_t st) {
    if (tiles_of(D) > 0) hipLaunchKernelGGL(k_idf_velocity, dim3(tiles_of(D)), dim3(256), 0, st, D);
}

}  // namespace sphgpu
