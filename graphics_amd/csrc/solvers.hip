// Per-particle neighbour passes of the three solvers.  Each kernel is one
// loop over particles of the reference (cited per kernel), one thread per
// particle, neighbours read through the padded ELL lists built by grid.hip.
// What a pass reads from its neighbours is packed into one or two 16-byte
// records (xk: position + the pass's scalar; v: velocity + density) so that
// each neighbour costs one or two gathers.  Sums that decide convergence are
// reduced in FP64: a fixed-order block tree per tile (deterministic), then a
// fixed-order per-scene reduction.
#include "sph_common.h"

namespace sphgpu {

namespace {

struct Slot {
    int i, tile, scene;
    bool live;
    bool packed;   // the tile's neighbour lists are in the packed 16-bit form
};

__device__ __forceinline__ Slot slot_of(const Dev& D, bool gated) {
    Slot s;
    s.tile = tile_of_block(D.tileMap);
    s.i = s.tile * TILE + threadIdx.x;
    s.scene = D.tileScene[s.tile];
    s.live = (int)threadIdx.x < D.tileValid[s.tile];
    if (gated) s.live = s.live && D.active[s.scene] != 0;
    s.packed = false;
    return s;
}

// The tile's list table (fluid slots 0..31, boundary 32..63) and the window
// offset of each fluid slot ([31]: the window size), staged in LDS by the
// kernels that read neighbour lists.
__shared__ int g_nbtab[NB16_TAB];
__shared__ int g_winoff[32];

// slot_of for kernels that read neighbour lists: every thread of the block
// calls it before any returns (it stages the tile's tables in LDS, both with
// one round trip).  A gated block whose scene has converged loads nothing
// (its particles are not live): solver iterations launched past convergence
// return at once.
__device__ __forceinline__ Slot slot_nb(const Dev& D, bool gated) {
    Slot s = slot_of(D, gated);
    if (gated && !D.active[s.scene]) return s;   // uniform per block
    if (threadIdx.x < NB16_TAB) g_nbtab[threadIdx.x] = D.nbTab[(size_t)s.tile * NB16_TAB + threadIdx.x];
    else if (threadIdx.x < NB16_TAB + 32)
        g_winoff[threadIdx.x - NB16_TAB] = D.nbWin[(size_t)s.tile * 32 + threadIdx.x - NB16_TAB];
    __syncthreads();
    s.packed = __builtin_amdgcn_readfirstlane(g_nbtab[NB16_FLAG]) != 0;
    return s;
}

// A list group is read once per pass.  Where the lists and records together
// fit the last-level cache the plain load lets the next pass find the list
// there; on larger scenes a streaming load keeps the list from pushing the
// gathered records out (Dev::listNT, chosen by the autotuner).
typedef int v4i_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int4 ld_group(const int4* p, bool nt) {
    if (nt) {
        const v4i_t v = __builtin_nontemporal_load(reinterpret_cast<const v4i_t*>(p));
        return make_int4(v.x, v.y, v.z, v.w);
    }
    return *p;
}

// Neighbour loops in split form: load(j) issues the gather(s) of neighbour j
// and returns the record(s), use(j, r) consumes them.  A batch of neighbours is gathered first, then the next list group is
// requested, then the batch is used: a wave waits for its memory returns in
// issue order, so a list load issued ahead of the gathers would hold up every
// batch.  W = 8: one batch of eight per packed group (single-record passes);
// W = 4: two batches of four (two-record passes: eight gathers in flight
// either way).  Entries past the padded count still gather (from the padding
// index) but are not used, so every pass sums exactly the listed entries.
template <int W, int TOFF, class LD, class US>
__device__ __forceinline__ void each_in_list16_lu(const int4* __restrict__ list, int ncap, int i, int n, int padj,
                                                  bool nt, LD&& load, US&& use) {
    static_assert(W == 4 || W == 8, "batch of four or eight");
    const int4* __restrict__ col = list + i;
    const int ng = (n + 7) >> 3;
    if (ng <= 0) return;
    const int n4 = (n + 3) & ~3;
    int4 cur = ld_group(&col[0], nt);
    for (int g = 0; g < ng; ++g) {
        const unsigned w[4] = {(unsigned)cur.x, (unsigned)cur.y, (unsigned)cur.z, (unsigned)cur.w};
        const bool second = 8 * g + 4 < n4;   // entries 8g+4..8g+7 are listed
        int j[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const unsigned e = (w[k >> 1] >> (16 * (k & 1))) & 0xffffu;
            const int jj = g_nbtab[TOFF + (e >> NB16_OFF_BITS)] + (int)e;
            j[k] = (k < 4 || second) ? jj : padj;
        }
        int4 nxt = cur;
        if (W == 8) {
            decltype(load(0)) r[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) r[k] = load(j[k]);
            __builtin_amdgcn_sched_barrier(0);
            if (g + 1 < ng) nxt = ld_group(&col[(size_t)(g + 1) * ncap], nt);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < 4; ++k) use(j[k], r[k]);
            if (second) {
#pragma unroll
                for (int k = 4; k < 8; ++k) use(j[k], r[k]);
            }
        } else {
            decltype(load(0)) r[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) r[k] = load(j[k]);
#pragma unroll
            for (int k = 0; k < 4; ++k) use(j[k], r[k]);
#pragma unroll
            for (int k = 0; k < 4; ++k) r[k] = load(j[4 + k]);
            __builtin_amdgcn_sched_barrier(0);
            if (g + 1 < ng) nxt = ld_group(&col[(size_t)(g + 1) * ncap], nt);
            __builtin_amdgcn_sched_barrier(0);
            if (second) {
#pragma unroll
                for (int k = 0; k < 4; ++k) use(j[4 + k], r[k]);
            }
        }
        cur = nxt;
    }
}

// the same for 32-bit lists (groups of four, all listed or padding)
template <class LD, class US>
__device__ __forceinline__ void each_in_list_lu(const int* __restrict__ list, int ncap, int i, int n, bool nt,
                                                LD&& load, US&& use) {
    const int4* __restrict__ col = reinterpret_cast<const int4*>(list) + i;
    const int ng = (n + 3) >> 2;
    if (ng <= 0) return;
    int4 cur = ld_group(&col[0], nt);
    for (int g = 0; g < ng; ++g) {
        const int j[4] = {cur.x, cur.y, cur.z, cur.w};
        decltype(load(0)) r[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) r[k] = load(j[k]);
        int4 nxt = cur;
        __builtin_amdgcn_sched_barrier(0);
        if (g + 1 < ng) nxt = ld_group(&col[(size_t)(g + 1) * ncap], nt);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < 4; ++k) use(j[k], r[k]);
        cur = nxt;
    }
}

template <int W, class LD, class US>
__device__ __forceinline__ void each_fluid_lu(const Dev& D, const Slot& s, int nf, LD&& load, US&& use) {
    const bool nt = __builtin_amdgcn_readfirstlane(D.listNT) != 0;
    if (s.packed) each_in_list16_lu<W, 0>(D.nbF16, D.ncap, s.i, nf, s.i, nt, load, use);
    else each_in_list_lu(D.nbF, D.ncap, s.i, nf, nt, load, use);
}

template <int W, class LD, class US>
__device__ __forceinline__ void each_boundary_lu(const Dev& D, const Slot& s, int nb, LD&& load, US&& use) {
    const bool nt = __builtin_amdgcn_readfirstlane(D.listNT) != 0;
    if (s.packed) each_in_list16_lu<W, 32>(D.nbB16, D.ncap, s.i, nb, D.nb, nt, load, use);
    else each_in_list_lu(D.nbB, D.ncap, s.i, nb, nt, load, use);
}

// two gather records of one neighbour
struct Rec2 {
    float4 a, b;
};

// ---- LDS windows --------------------------------------------------------
// The fluid neighbours of a packed tile all lie in its slot runs (one
// contiguous run of the sorted particles per (own row, stencil row) pair), so
// the concatenation of those runs -- the tile's window, about ten times the
// tile's own 256 particles -- holds every record the tile's neighbour loops
// read.  A pass stages the window in LDS with coalesced loads (a particle
// has about 28 fluid neighbours, so a record is read by about three lanes
// of the tile) and the loops then read it from
// LDS instead of gathering per lane through the L1, taking the window index
// of each neighbour from the window-index copy of the fluid lists (nbW16).
// Tiles with 32-bit lists, or whose window exceeds WIN_MAX records, keep the
// per-lane gathers (the same sums in the same order, so both forms give
// identical results).

// Window size of the block's tile (after slot_nb) for the pass `bit`, or 0
// when the tile's loops gather from memory.
__device__ __forceinline__ int win_layout(const Dev& D, const Slot& s, int bit) {
    if (!s.packed || !(D.winMask & bit)) return 0;   // uniform per block
    const int n = __builtin_amdgcn_readfirstlane(g_winoff[31]);
    return n <= WIN_MAX ? n : 0;
}

// sorted particle index of window record k (0 <= k < window size): the slot
// holding k is the largest s with g_winoff[s] <= k (empty slots share their
// offset with the next one)
__device__ __forceinline__ int win_particle(int k) {
    int s = 0;
#pragma unroll
    for (int b = 16; b > 0; b >>= 1)
        if (g_winoff[s + b] <= k) s += b;
    return g_nbtab[s] + (s << NB16_OFF_BITS) + (k - g_winoff[s]);
}

// Stages the window: load(j) reads what a record needs from particle j,
// store(k, r) writes it to window record k.  A window holds at most
// WIN_MAX = 12 x 256 records, so each thread requests all of its records
// before it stores the first.
template <class LD, class ST>
__device__ __forceinline__ void win_stage(int n, LD&& load, ST&& store) {
    constexpr int DEPTH = WIN_MAX / 256;
    decltype(load(0)) r[DEPTH];
#pragma unroll
    for (int u = 0; u < DEPTH; ++u) {
        const int k = threadIdx.x + 256 * u;
        if (k < n) r[u] = load(win_particle(k));
    }
#pragma unroll
    for (int u = 0; u < DEPTH; ++u) {
        const int k = threadIdx.x + 256 * u;
        if (k < n) store(k, r[u]);
    }
    __syncthreads();
}

// The packed fluid list of one particle in registers: every group is
// requested at once, before the pass stages its window, so the list loads
// overlap the staging instead of waiting one group at a time.
constexpr int NB16_GROUPS = NBF_CAP / 8;
struct List16 {
    int4 g[NB16_GROUPS];
};
__device__ __forceinline__ void list16_load(const int4* __restrict__ list, int ncap, int i, int n, bool nt,
                                            List16& L) {
    const int ng = (n + 7) >> 3;
#pragma unroll
    for (int g = 0; g < NB16_GROUPS; ++g)
        L.g[g] = g < ng ? ld_group(&list[(size_t)g * ncap + i], nt) : make_int4(0, 0, 0, 0);
}

// Neighbour loop over a window-index list (in registers) whose records are
// in the LDS window: load(k) reads window record k, use(k, r) consumes it.
// Entries past the padded count are read from record 0 and not used.
template <class LD, class US>
__device__ __forceinline__ void each_in_win16(const List16& L, int n, LD&& load, US&& use) {
    const int ng = (n + 7) >> 3;
    const int n4 = (n + 3) & ~3;
#pragma unroll
    for (int g = 0; g < NB16_GROUPS; ++g) {
        if (g >= ng) break;
        const int4 cur = L.g[g];
        const unsigned w[4] = {(unsigned)cur.x, (unsigned)cur.y, (unsigned)cur.z, (unsigned)cur.w};
        const bool second = 8 * g + 4 < n4;
        int k[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int kq = (int)((w[q >> 1] >> (16 * (q & 1))) & 0xffffu);
            k[q] = (q < 4 || second) ? kq : 0;
        }
        decltype(load(0)) r[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) r[q] = load(k[q]);
#pragma unroll
        for (int q = 0; q < 4; ++q) use(k[q], r[q]);
        if (second) {
#pragma unroll
            for (int q = 4; q < 8; ++q) use(k[q], r[q]);
        }
    }
}

__device__ __forceinline__ int phase_of(float4 x) { return (int)x.w; }

__device__ __forceinline__ void write_partial(const Dev& D, int tile, int q, double v, double* lds) {
    double t = block_sum_256(v, lds);
    if (threadIdx.x == 0) D.partial[2 * tile + q] = t;
}

// packed counts: stored fluid entries | fluid neighbours found (the stored
// ones plus any beyond the list capacity, up to 255) << 8 | boundary << 16 |
// the particle's phase << 24
__device__ __forceinline__ int nf_of(int c) { return c & 0xff; }
__device__ __forceinline__ int nf_real(int c) { return (c >> 8) & 0xff; }
__device__ __forceinline__ int nb_of(int c) { return (c >> 16) & 0xff; }
__device__ __forceinline__ int phase_of_cnt(int c) { return c >> 24; }

// The window-index list of a live particle of a tile that may stage a
// window for the pass `bit` (packed, the pass's window on); the window itself
// is decided by win_layout (the list builder writes this list for exactly the
// tiles whose window it accepts).
__device__ __forceinline__ void list16_prefetch(const Dev& D, const Slot& s, int c, List16& L, int bit) {
    if (s.live && s.packed && (D.winMask & bit)) {
        const bool nt = __builtin_amdgcn_readfirstlane(D.listNT) != 0;
        list16_load(D.nbW16, D.ncap, s.i, nf_of(c), nt, L);
    }
}

// ---- two tiles per block ------------------------------------------------
// The LDS-window builds take two consecutive tiles per block: the second
// tile's window records are requested into registers before the first
// tile's neighbour loops run, so their latency is spent on the first tile's
// work instead of in a second staging phase.  Each particle's own values are
// loaded before them (a wait for a later load waits for every earlier one).
struct Rec3x2 {
    float3 a, b;
};
__device__ __forceinline__ Slot slot_tile(const Dev& D, int tile) {
    Slot s;
    s.tile = tile;
    s.i = tile * TILE + threadIdx.x;
    s.scene = D.tileScene[tile];
    s.live = (int)threadIdx.x < D.tileValid[tile] && D.active[s.scene] != 0;
    s.packed = false;
    return s;
}
// win_particle with the tables in `tab` (list table, then window offsets)
__device__ __forceinline__ int win_particle_tab(const int* tab, int k) {
    const int* off = tab + NB16_TAB;
    int s = 0;
#pragma unroll
    for (int b = 16; b > 0; b >>= 1)
        if (off[s + b] <= k) s += b;
    return tab[s] + (s << NB16_OFF_BITS) + (k - off[s]);
}

// The LDS-window build of a solver-loop pass: two consecutive tiles per
// block.  own(s, c) loads what a particle reads of itself, load(j) a window
// record of particle j, store(k, r) writes window record k, body(s, c, nw,
// L, own) runs the tile's particles (nw > 0: their fluid neighbours are in
// the window) and returns the particle's part of the tile's partial sum,
// which PARTIAL blocks write.  The second tile's records are requested
// before the first tile's body runs and stored after it.
template <int BIT, bool PARTIAL, class OwnF, class LoadF, class StoreF, class BodyF>
__device__ __forceinline__ void window_pair(const Dev& D, OwnF&& own, LoadF&& load, StoreF&& store, BodyF&& body) {
    __shared__ double lds[4];
    __shared__ int tabB[NB16_TAB + 32];   // the second tile's list table and window offsets
    const int tA = 2 * tile_of_block(D.tileMap), tB = tA + 1;
    const bool hasB = tB < D.npad / TILE;
    Slot sA = slot_tile(D, tA);
    Slot sB = slot_tile(D, hasB ? tB : tA);
    const bool actA = D.active[sA.scene] != 0;
    const bool actB = hasB && D.active[sB.scene] != 0;
    if (!hasB) sB.live = false;
    const int t = threadIdx.x;
    // both tiles' tables with one round trip
    if (actA) {
        if (t < NB16_TAB) g_nbtab[t] = D.nbTab[(size_t)tA * NB16_TAB + t];
        else if (t < NB16_TAB + 32) g_winoff[t - NB16_TAB] = D.nbWin[(size_t)tA * 32 + t - NB16_TAB];
    }
    if (actB) {
        if (t >= 128 && t < 128 + NB16_TAB) tabB[t - 128] = D.nbTab[(size_t)tB * NB16_TAB + t - 128];
        else if (t >= 128 + NB16_TAB && t < 128 + NB16_TAB + 32)
            tabB[t - 128] = D.nbWin[(size_t)tB * 32 + t - 128 - NB16_TAB];
    }
    __syncthreads();
    sA.packed = actA && __builtin_amdgcn_readfirstlane(g_nbtab[NB16_FLAG]) != 0;
    sB.packed = actB && __builtin_amdgcn_readfirstlane(tabB[NB16_FLAG]) != 0;
    // the first tile: list and own values, then its window
    const int cA = sA.live ? D.cnt[sA.i] : 0;
    List16 LA;
    list16_prefetch(D, sA, cA, LA, BIT);
    const auto oA = own(sA, cA);
    const int nwA = win_layout(D, sA, BIT);
    if (nwA) win_stage(nwA, load, store);
    // the second tile's window records, in flight during the first tile's body
    int nwB = 0;
    if (sB.packed && (D.winMask & BIT)) {
        nwB = __builtin_amdgcn_readfirstlane(tabB[NB16_TAB + 31]);
        if (nwB > WIN_MAX) nwB = 0;
    }
    constexpr int DEPTH = WIN_MAX / 256;
    decltype(load(0)) rB[DEPTH];
#pragma unroll
    for (int u = 0; u < DEPTH; ++u) {
        const int k = t + 256 * u;
        if (k < nwB) rB[u] = load(win_particle_tab(tabB, k));
    }
    const double sumA = body(sA, cA, nwA, LA, oA);
    if (PARTIAL) {
        if (actA) write_partial(D, tA, 0, sumA, lds);
        else if (t == 0) D.partial[2 * tA] = 0.0;
    }
    if (!hasB) return;   // uniform
    __syncthreads();     // every lane is done with the first tile's window and tables
    if (actB) {
        if (t < NB16_TAB) g_nbtab[t] = tabB[t];
        else if (t < NB16_TAB + 32) g_winoff[t - NB16_TAB] = tabB[t];
    }
#pragma unroll
    for (int u = 0; u < DEPTH; ++u) {
        const int k = t + 256 * u;
        if (k < nwB) store(k, rB[u]);
    }
    __syncthreads();
    const int cB = sB.live ? D.cnt[sB.i] : 0;
    List16 LB;
    list16_prefetch(D, sB, cB, LB, BIT);
    const auto oB = own(sB, cB);
    const double sumB = body(sB, cB, nwB, LB, oB);
    if (PARTIAL) {
        if (actB) write_partial(D, tB, 0, sumB, lds);
        else if (t == 0) D.partial[2 * tB] = 0.0;
    }
}

__device__ __forceinline__ bool inside_domain(const SceneGrid& G, float4 x) {
    return G.smin[0] <= x.x && G.smin[1] <= x.y && G.smin[2] <= x.z && x.x <= G.smax[0] &&
           x.y <= G.smax[1] && x.z <= G.smax[2];
}

// ---------------------------------------------------------------- CFL ----
__global__ __launch_bounds__(256) void k_cfl_partial(Dev D) {
    __shared__ float lds[4];
    Slot s = slot_of(D, false);
    float v2 = 0.0f;
    if (s.live) {
        float4 v = D.v[s.i];
        v2 = v.x * v.x + v.y * v.y + v.z * v.z;
    }
    float m = block_max_256(v2, lds);
    if (threadIdx.x == 0) D.vmaxPartial[s.tile] = m;
}

__global__ __launch_bounds__(256) void k_cfl_reduce(Dev D, float cfl, float spacing, float dtMax,
                                                  float dtMin) {
    __shared__ float lds[4];
    int sc = blockIdx.x;
    int t0 = D.scenePadStart[sc] / TILE, t1 = D.scenePadStart[sc + 1] / TILE;
    float m = 0.0f;
    for (int t = t0 + threadIdx.x; t < t1; t += 256) m = fmaxf(m, D.vmaxPartial[t]);
    m = block_max_256(m, lds);
    if (threadIdx.x == 0) {
        float dt = dtMax;
        if (cfl > 0.0f) {
            float vmax = sqrtf(m);
            if (vmax > 0.0f) dt = fminf(dtMax, cfl * spacing / vmax);
            dt = fmaxf(dt, dtMin);
        }
        D.dt[sc] = dt;
    }
}

// ------------------------------------------------------- scene reduce ----
// mode 0: DFSPH density  err = sum(rho*/rho0)/n - 1          (DFSPH.cpp:696)
// mode 1: DFSPH divergence / IISPH error   err = sum/n      (DFSPH.cpp:896, IISph_M.cpp:353)
// mode 2: PCISPH  err = max error/rho0 (second partial)     (PciSph.cpp:344-350)
// Loop continuation (while err > tol || it < minIt; PCISPH err >= tol,
// PciSph.cpp:116), stopping at `cap`.
// Writes the scene's continuation flag to out (the active array itself, or
// the next iteration's buffer of the DFSPH loops; a scene already done stays
// done).  Every thread of the block calls it.
__device__ __forceinline__ void scene_reduce_block(const Dev& D, int sc, int mode, float tol, int minIt, int cap,
                                                   int* out) {
    __shared__ double lds[4];
    if (!D.active[sc]) {   // uniform per block
        if (threadIdx.x == 0 && out != D.active) out[sc] = 0;
        return;
    }
    int t0 = D.scenePadStart[sc] / TILE, t1 = D.scenePadStart[sc + 1] / TILE;
    __shared__ float ldsf[4];
    double s = 0.0;
    float mx = 0.0f;
    // each thread sums tiles t0 + tid, + 256, + 512, ... in that order; the
    // loads of eight of them are issued together
    for (int t = t0 + threadIdx.x; t < t1; t += 256 * 8) {
        double v[8];
        float w[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int tt = t + 256 * k;
            v[k] = tt < t1 ? D.partial[2 * tt] : 0.0;
            w[k] = (mode == 2 && tt < t1) ? (float)D.partial[2 * tt + 1] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (t + 256 * k < t1) s += v[k];
            mx = fmaxf(mx, w[k]);
        }
    }
    s = block_sum_256(s, lds);
    mx = block_max_256(mx, ldsf);
    if (threadIdx.x == 0) {
        int n = D.sceneCount[sc];
        double err = n > 0 ? s / n : 0.0;
        if (mode == 0) err = n > 0 ? err - 1.0 : 0.0;
        if (mode == 2) err = (double)mx;
        int it = D.iters[sc] + 1;
        D.iters[sc] = it;
        D.sceneErr[sc] = err;
        bool cont = (mode == 2 ? err >= (double)tol : err > (double)tol) || (it < minIt);
        if (cont && it >= cap) {
            cont = false;
            D.counters[C_COUNT - 1] = 1;  // max-iteration flag
        }
        out[sc] = cont ? 1 : 0;
    }
}

__global__ __launch_bounds__(256) void k_scene_reduce(Dev D, int mode, float tol, int minIt, int cap) {
    scene_reduce_block(D, blockIdx.x, mode, tol, minIt, cap, D.activeNext ? D.activeNext : D.active);
}

// ------------------------------------------------------------- stats -----
__global__ __launch_bounds__(256) void k_stats_partial(Dev D, double* part) {
    __shared__ double lds[4];
    __shared__ float ldf[4];
    Slot s = slot_of(D, false);
    double e = 0.0, ke = 0.0, n = 0.0;
    float emax = -1e30f;
    int nan = 0;
    if (s.live && D.id[s.i] >= 0) {
        float4 x = D.x[s.i], v = D.v[s.i];
        int ph = phase_of(x);
        float r = D.rho[s.i];
        float ee = r / D.rho0[ph] - 1.0f;
        e = (double)ee;
        emax = ee;
        n = 1.0;
        ke = 0.5 * (double)D.mass[ph] * (double)(v.x * v.x + v.y * v.y + v.z * v.z);
        if (!isfinite(x.x + x.y + x.z + v.x + v.y + v.z)) nan = 1;
    }
    double se = block_sum_256(e, lds);
    double sk = block_sum_256(ke, lds);
    double sn = block_sum_256(n, lds);
    float mx = block_max_256(emax, ldf);
    if (nan) atomicAdd(&D.counters[C_NAN], 1);
    if (threadIdx.x == 0) {
        part[4 * s.tile] = se;
        part[4 * s.tile + 1] = sk;
        part[4 * s.tile + 2] = (double)mx;
        part[4 * s.tile + 3] = sn;
    }
}

// per scene: [mean(rho/rho0-1), max(rho/rho0-1), kinetic energy, count]
__global__ __launch_bounds__(256) void k_stats_reduce(Dev D, const double* part, double* out) {
    __shared__ double lds[4];
    __shared__ float ldf[4];
    int sc = blockIdx.x;
    int t0 = D.scenePadStart[sc] / TILE, t1 = D.scenePadStart[sc + 1] / TILE;
    double se = 0.0, sk = 0.0, sn = 0.0;
    float mx = -1e30f;
    for (int t = t0 + threadIdx.x; t < t1; t += 256) {
        se += part[4 * t];
        sk += part[4 * t + 1];
        mx = fmaxf(mx, (float)part[4 * t + 2]);
        sn += part[4 * t + 3];
    }
    se = block_sum_256(se, lds);
    sk = block_sum_256(sk, lds);
    sn = block_sum_256(sn, lds);
    mx = block_max_256(mx, ldf);
    if (threadIdx.x == 0) {
        out[4 * sc] = sn > 0 ? se / sn : 0.0;
        out[4 * sc + 1] = sn > 0 ? (double)mx : 0.0;
        out[4 * sc + 2] = sk;
        out[4 * sc + 3] = sn;
    }
}

// ------------------------------------------------------ shared passes ----
// density m W(0) + sum_f m W + sum_b rho0 V_b W, Kahan-compensated
// (WcSph.cpp:121-138, DFSPH.cpp:132-169, IISph_M.cpp:101-112); coincident
// pairs contribute W(0), the padding (self) entries nothing.
__device__ __forceinline__ float density_sum(const Dev& D, const Slot& s, float4 pa, float m, float rho0, int c) {
    const int i = s.i;
    Kahan rho;
    rho.init(m * D.W0);
    each_fluid_lu<8>(D, s, nf_of(c), [&](int j) { return D.x[j]; }, [&](int j, float4 pb) {
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        float w = kernel_W(fast_sqrt(dx * dx + dy * dy + dz * dz), D);
        rho.add(j != i ? m * w : 0.0f);
    });
    each_boundary_lu<8>(D, s, nb_of(c), [&](int j) { return D.xb[j]; }, [&](int j, float4 pb) {
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        rho.add(rho0 * pb.w * kernel_W(fast_sqrt(dx * dx + dy * dy + dz * dz), D));
    });
    return rho.s;
}

// one pair's term g(r) (v*_a - v*_b).x_ab of the density change rate
__device__ __forceinline__ float dchange_term(const Dev& D, float4 pa, float4 va, float bx, float by, float bz,
                                              float ux, float uy, float uz) {
    const float dx = pa.x - bx, dy = pa.y - by, dz = pa.z - bz;
    const float g = kernel_gradfac_sq(dot3f(dx, dy, dz, dx, dy, dz), D);
    return g * dot3f(va.x - ux, va.y - uy, va.z - uz, dx, dy, dz);
}

// sum_b V_b gradW.(v*_a - v_b) over the boundary neighbours
__device__ __forceinline__ float dchange_boundary(const Dev& D, const Slot& s, float4 pa, float4 va, int c) {
    float sb = 0.0f;
    each_boundary_lu<4>(D, s, nb_of(c), [&](int j) { return Rec2{D.xb[j], D.vb[j]}; }, [&](int j, Rec2 rec) {
        sb += rec.a.w * dchange_term(D, pa, va, rec.a.x, rec.a.y, rec.a.z, rec.b.x, rec.b.y, rec.b.z);
    });
    return sb;
}

// sum_f m gradW.(v*_a - v*_b) + sum_b rho0 V_b gradW.(v*_a - v_b): the
// predicted density change rate (DFSPH.cpp:589-617, 795-825)
__device__ __forceinline__ float density_change(const Dev& D, const Slot& s, float4 pa, float4 va, float m,
                                                float rho0, int c) {
    float sf = 0.0f, sb = 0.0f;
    each_fluid_lu<4>(D, s, nf_of(c), [&](int j) { return Rec2{D.x[j], D.vs[j]}; }, [&](int j, Rec2 rec) {
        sf += dchange_term(D, pa, va, rec.a.x, rec.a.y, rec.a.z, rec.b.x, rec.b.y, rec.b.z);
    });
    sb = dchange_boundary(D, s, pa, va, c);
    return m * sf + rho0 * sb;
}

// ============================================================ WCSPH ======
// WcSph::wc_computePressure (WcSph.cpp:99-153): rho (Kahan), Tait p >= 0.
// Publishes the force pass's gather records: xk = (x, p/rho^2), v.w = rho.
__global__ __launch_bounds__(256) void k_wc_pressure(Dev D) {
    Slot s = slot_nb(D, false);
    if (!s.live) return;
    const int i = s.i;
    float4 pa = D.x[i];
    int ph = phase_of(pa);
    float m = D.mass[ph], rho0 = D.rho0[ph];
    float density = density_sum(D, s, pa, m, rho0, D.cnt[i]);
    float p = D.tait_B[ph] * (powf(density / rho0, D.gamma) - 1.0f);
    p = p < 0.0f ? 0.0f : p;
    D.rho[i] = density;
    D.f1[i] = p;
    D.xk[i] = make_float4(pa.x, pa.y, pa.z, p / (density * density));
    D.v[i].w = density;
}

// WcSph::wc_computeForce (WcSph.cpp:156-312, 364-436) fused with the
// Euler-Cromer update and domain removal of SphBase::updateFluids
// (SphBase.cpp:377-402): writes x2/v2, the caller swaps buffers.
__global__ __launch_bounds__(256) void k_wc_force(Dev D) {
    Slot s = slot_nb(D, false);
    if (!s.live) return;
    const int i = s.i;
    float4 pa = D.x[i], va = D.v[i];
    int ph = phase_of(pa);
    float m = D.mass[ph], rho0 = D.rho0[ph], alpha = D.visc[ph];
    float rhoa = va.w, pra = D.f1[i];
    float pa_r2 = pra / (rhoa * rhoa);
    const float Hcs2 = 2.0f * D.Hs * D.cs;
    const bool multi = D.nphases > 1;
    float ax = D.gx, ay = D.gy, az = D.gz;
    int c = D.cnt[i];
    each_fluid_lu<4>(D, s, nf_of(c), [&](int j) { return Rec2{D.xk[j], D.v[j]}; }, [&](int j, Rec2 rec) {
        const float4 pb = rec.a, vb = rec.b;   // (x, p/rho^2), (v, rho)
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        float r2 = dx * dx + dy * dy + dz * dz;
        float g = kernel_gradfac(fast_sqrt(r2), D);
        float pro = (va.x - vb.x) * dx + (va.y - vb.y) * dy + (va.z - vb.z) * dz;
        int phb = multi ? phase_of(D.x[j]) : ph;
        float acce;
        if (phb == ph) {
            float nu = alpha * Hcs2 / (rhoa + vb.w);
            float pi = pro < 0.0f ? -nu * pro * fast_rcp(r2 + D.visc_eps) : 0.0f;
            acce = g * (-m * (pa_r2 + pb.w + pi));
        } else {                                      /* SP08 multiphase, WcSph.cpp:220-260 */
            float mb = D.mass[phb], al = 0.5f * (alpha + D.visc[phb]);
            // p/delta^2 with delta = rho/m: p m^2 / rho^2
            acce = g * (-1.0f / m * (pa_r2 * m * m + pb.w * mb * mb));
            float nu = al * Hcs2 / (rhoa + vb.w);
            float pi = pro < 0.0f ? -nu * pro * fast_rcp(r2 + D.visc_eps) : 0.0f;
            acce += g * (-0.5f * (m + mb) * pi);
        }
        ax += dx * acce; ay += dy * acce; az += dz * acce;
    });
    each_boundary_lu<4>(D, s, nb_of(c), [&](int j) { return Rec2{D.xb[j], D.vb[j]}; }, [&](int j, Rec2 rec) {
        const float4 pb = rec.a, vb = rec.b;
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        float r2 = dx * dx + dy * dy + dz * dz;
        float g = kernel_gradfac(fast_sqrt(r2), D);
        float acce = pra > 0.0f ? g * (-rho0 * pb.w * (pa_r2 * 2.0f)) : 0.0f;
        float pro = (va.x - vb.x) * dx + (va.y - vb.y) * dy + (va.z - vb.z) * dz;
        float nu = D.solid_visc * Hcs2 / (rhoa * 2.0f);
        float pi = pro < 0.0f ? -nu * pro * fast_rcp(r2 + D.visc_eps) : 0.0f;
        acce += g * (-rho0 * pb.w * pi);
        ax += dx * acce; ay += dy * acce; az += dz * acce;
    });
    float dt = D.dt[s.scene];
    D.f4a[i] = make_float4(ax, ay, az, 0.0f);
    float4 vn = make_float4(va.x + ax * dt, va.y + ay * dt, va.z + az * dt, 0.0f);
    float4 xn = make_float4(pa.x + vn.x * dt, pa.y + vn.y * dt, pa.z + vn.z * dt, pa.w);
    D.v2[i] = vn;
    D.x2[i] = xn;
    if (!inside_domain(D.grids[s.scene], xn)) {
        D.id[i] = -1;
        atomicAdd(&D.counters[C_REMOVED], 1);
    }
}

// ============================================================ DFSPH ======
// DFSPH::DF_constantDensitySolverFirst (DFSPH.cpp:84-193): rho (Kahan) and alpha.
template <bool WARM>
__global__ __launch_bounds__(256) void k_df_density_alpha(Dev D, float* __restrict__ feat) {
    Slot s = slot_nb(D, false);
    if (!s.live) return;
    const int i = s.i;
    float4 pa = D.x[i];
    int ph = phase_of(pa);
    float m = D.mass[ph], rho0 = D.rho0[ph];
    Kahan rho;
    rho.init(m * D.W0);
    float a1x = 0.f, a1y = 0.f, a1z = 0.f, a2 = 0.f;
    int c = D.cnt[i];
    // WARM: also Drho/Dt with v* = v (density_change) and the nearest
    // boundary distance, for the warm start of the divergence loop
    const float4 va = WARM ? D.v[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float sp = D.spacing;
    float sf = 0.0f, sb = 0.0f, bmin2 = 4.0f * sp * sp;
    auto fluid = [&](int j, float4 pb, float4 vb) {
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        float r = fast_sqrt(dx * dx + dy * dy + dz * dz);
        rho.add(j != i ? m * kernel_W(r, D) : 0.0f);
        float g = kernel_gradfac(r, D);
        float gm = g * m;
        a1x += dx * gm; a1y += dy * gm; a1z += dz * gm;
        float gs = gm * r;                     // |dW/dr| m (zero at r = 0)
        a2 += gs * gs;
        if (WARM) sf += g * ((va.x - vb.x) * dx + (va.y - vb.y) * dy + (va.z - vb.z) * dz);
    };
    auto boundary = [&](float4 pb, float4 vb) {
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        float r2 = dx * dx + dy * dy + dz * dz;
        float r = fast_sqrt(r2);
        float psi = rho0 * pb.w;
        rho.add(psi * kernel_W(r, D));
        float g = kernel_gradfac(r, D);
        float gm = g * psi;
        a1x += dx * gm; a1y += dy * gm; a1z += dz * gm;
        if (WARM) {
            sb += pb.w * g * ((va.x - vb.x) * dx + (va.y - vb.y) * dy + (va.z - vb.z) * dz);
            bmin2 = fminf(bmin2, r2);
        }
    };
    if (WARM) {
        each_fluid_lu<4>(D, s, nf_of(c), [&](int j) { return Rec2{D.x[j], D.v[j]}; },
                         [&](int j, Rec2 rec) { fluid(j, rec.a, rec.b); });
        each_boundary_lu<4>(D, s, nb_of(c), [&](int j) { return Rec2{D.xb[j], D.vb[j]}; },
                            [&](int j, Rec2 rec) { boundary(rec.a, rec.b); });
    } else {
        each_fluid_lu<8>(D, s, nf_of(c), [&](int j) { return D.x[j]; },
                         [&](int j, float4 pb) { fluid(j, pb, pb); });
        each_boundary_lu<8>(D, s, nb_of(c), [&](int j) { return D.xb[j]; },
                            [&](int j, float4 pb) { boundary(pb, pb); });
    }
    float density = rho.s;
    float den = a1x * a1x + a1y * a1y + a1z * a1z + a2;
    if (den < 1.0e-6f) den = 1.0e-6f;
    const float alpha = density / den;
    D.rho[i] = density;
    D.alpha[i] = alpha;
    D.v[i].w = density;
    if (!WARM) return;
    const float d = m * sf + rho0 * sb;
    D.f2[i] = d;
    if (!feat) return;
    const float dt = D.dt[s.scene];
    float4* f = reinterpret_cast<float4*>(feat) + 2 * (size_t)i;
    f[0] = make_float4((density / rho0 - 1.0f) * 100.0f, asinhf(d * dt / rho0 * 1000.0f),
                       (float)(nf_real(c) + nb_of(c)) * (1.0f / 30.0f),
                       log1pf(fmaxf(D.kvsum[i] * dt / (alpha * rho0) * 1000.0f, 0.0f)));
    f[1] = make_float4(alpha / D.alpha_lat, sqrtf(va.x * va.x + va.y * va.y + va.z * va.z) * dt / sp,
                       sqrtf(bmin2) / sp, dt / D.dt_max);
}

// DFSPH::DF_forceExceptPressure (DFSPH.cpp:1419-1475; helpers 1246-1359) and
// v* = v + dt a_adv (DFSPH.cpp:557-563).  The density loop starts from zero
// kappa: it only pushes (rho* >= rho0, DFSPH.cpp:620-623), so a warm guess
// that overshoots would never be taken back.
__global__ __launch_bounds__(256) void k_df_aadv(Dev D) {
    Slot s = slot_nb(D, false);
    if (!s.live) return;
    const int i = s.i;
    float4 pa = D.x[i], va = D.v[i];
    int ph = phase_of(pa);
    float m = D.mass[ph], rho0 = D.rho0[ph], alpha = D.visc[ph];
    float rhoa = va.w;
    const float nu10 = 10.0f * D.nu;
    const float Hcs2 = 2.0f * D.Hs * D.cs;
    const bool multi = D.nphases > 1;
    float ax = D.gx, ay = D.gy, az = D.gz;
    int c = D.cnt[i];
    each_fluid_lu<4>(D, s, nf_of(c), [&](int j) { return Rec2{D.x[j], D.v[j]}; }, [&](int j, Rec2 rec) {
        const float4 pb = rec.a, vb = rec.b;   // vb = (v, rho)
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        float r2 = dx * dx + dy * dy + dz * dz;
        float pro = (va.x - vb.x) * dx + (va.y - vb.y) * dy + (va.z - vb.z) * dz;
        float g = kernel_gradfac(fast_sqrt(r2), D);
        int phb = multi ? phase_of(pb) : ph;
        float f;
        if (phb == ph) {
            f = g * (pro * fast_rcp(r2 + D.visc_eps) * m * fast_rcp(vb.w) * nu10);
        } else {                                         /* DFSPH.cpp:1278-1303 */
            float mb = D.mass[phb], al = 0.5f * (alpha + D.visc[phb]);
            float nuv = al * Hcs2 / (rhoa + vb.w);
            f = g * (-0.5f * (m + mb) * (-nuv * pro * fast_rcp(r2 + D.visc_eps)));
        }
        f = pro < 0.0f ? f : 0.0f;
        ax += dx * f; ay += dy * f; az += dz * f;
    });
    each_boundary_lu<4>(D, s, nb_of(c), [&](int j) { return Rec2{D.xb[j], D.vb[j]}; }, [&](int j, Rec2 rec) {
        const float4 pb = rec.a, vb = rec.b;
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        float r2 = dx * dx + dy * dy + dz * dz;
        float pro = (va.x - vb.x) * dx + (va.y - vb.y) * dy + (va.z - vb.z) * dz;
        float g = kernel_gradfac(fast_sqrt(r2), D);
        float f = g * (pro * fast_rcp(r2 + D.visc_eps) * (rho0 * pb.w) / rhoa * nu10);
        f = pro < 0.0f ? f : 0.0f;
        ax += dx * f; ay += dy * f; az += dz * f;
    });
    float dt = D.dt[s.scene];
    D.f4a[i] = make_float4(ax, ay, az, 0.0f);
    D.vs[i] = make_float4(va.x + ax * dt, va.y + ay * dt, va.z + az * dt, rhoa);   // w: rho, as in v
    if (D.solver != SOLVER_IISPH_DF) D.ksum[i] = 0.0f;   // (there ksum carries beta)
}

// Warm start of the divergence loop (not in the reference).  The stiffness
// the loop accumulates, kappaV*dt summed over iterations, is carried per
// particle in kvsum; its dimensionless form
//     s = kvsum dt / (alpha rho0) * 1000
// (the relative density change per step the solve removed, in 1e-3) is
// heavy-tailed, so the model works with u = log(1 + s), and it is residual:
// it predicts the change y = u - u_prev from the previous step's u_prev
// (feature 3), so an untrained network starts from the previous stiffness.
//
// Features of the learned model, per particle, at the start of the solve:
//   0 (rho/rho0 - 1) * 100            density deviation
//   1 asinh(Drho/Dt dt / rho0 * 1000) velocity divergence (with the current v)
//   2 (fluid + boundary neighbours)/30
//   3 log(1 + previous step's s)      previous stiffness
//   4 alpha / alpha_lattice           alpha factor
//   5 |v| dt / spacing
//   6 distance to the nearest boundary particle / spacing (2 if none)
//   7 dt / dt_max
// k_df_density_alpha<true> computes them with rho and alpha (one neighbour
// pass) and always stores Drho/Dt in f2 for the gate of k_df_warm_apply.

// mode 0: no warm start (accumulator restarts at zero); mode 1: warm_scale
// times the previous step's accumulated stiffness K_prev (Bender & Koschier
// use one half); mode 2: learn_scale times the model's prediction
// (s = exp(u_prev + y) - 1), bounded by max(K_prev, K_first), K_first = Drho/Dt alpha
// being the first increment the solve would make: the model may lower the
// guess anywhere but raise it only as far as one of those two.  Modes 1
// and 2 apply only where the solve itself would apply stiffness; the guess
// becomes xk.w = kappaV0/rho for one velocity correction, and the
// accumulator restarts from it.
__global__ __launch_bounds__(256) void k_df_warm_apply(Dev D, const float* __restrict__ y,
                                                     const float* __restrict__ feat, int mode) {
    Slot s = slot_of(D, false);
    if (!s.live) return;
    const int i = s.i;
    if (mode == 0) {
        D.kvsum[i] = 0.0f;
        return;
    }
    float4 pa = D.x[i];
    int ph = phase_of(pa);
    float dt = D.dt[s.scene];
    float kv = 0.0f, d = D.f2[i];
    // only where the solve itself would apply stiffness (DFSPH.cpp:829-832)
    if (d > 0.0f && !(D.rho[i] + d * dt < D.rho0[ph])) {
        const float kprev = D.kvsum[i];
        if (mode == 1) {
            kv = D.warm_scale * kprev;
        } else {
            const float alpha = D.alpha[i];
            float u = feat[8 * (size_t)i + 3] + y[i];
            float kp = expm1f(fminf(fmaxf(u, 0.0f), 12.0f)) * 1.0e-3f * alpha * D.rho0[ph] / dt;
            kv = D.learn_scale * fminf(kp, fmaxf(kprev, d * alpha));
        }
    }
    D.kvsum[i] = kv;
    D.xk[i] = make_float4(pa.x, pa.y, pa.z, kv / dt / D.rho[i]);
}

// converged stiffness of the divergence solve in the model's units, as the
// change from the previous step's (feature 3)
__global__ __launch_bounds__(256) void k_df_targets(Dev D, const float* __restrict__ feat, float* __restrict__ tgt) {
    Slot s = slot_of(D, false);
    if (!s.live) return;
    const int i = s.i;
    int ph = phase_of(D.x[i]);
    tgt[i] = log1pf(fmaxf(D.kvsum[i] * D.dt[s.scene] / (D.alpha[i] * D.rho0[ph]) * 1000.0f, 0.0f)) -
             feat[8 * (size_t)i + 3];
}

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// minibatch of `n` live particles drawn uniformly (hashed slots, retried
// until a live slot is hit); copies features and targets.
__global__ __launch_bounds__(256) void k_mlp_gather(Dev D, const float* __restrict__ feat,
                                                  const float* __restrict__ tgt, float* __restrict__ bfeat,
                                                  float* __restrict__ btgt, int n, uint32_t seed) {
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    int slot = -1;
    for (int tries = 0; tries < 16 && slot < 0; ++tries) {
        uint32_t r = hash32(seed * 0x9e3779b9u + (uint32_t)k * 16u + (uint32_t)tries);
        int q = (int)(r % (uint32_t)D.npad);
        int t = q / TILE;
        if ((q - t * TILE) < D.tileValid[t] && D.id[q] >= 0) slot = q;
    }
    if (slot < 0) slot = 0;
    const float4* f = reinterpret_cast<const float4*>(feat) + 2 * (size_t)slot;
    float4* o = reinterpret_cast<float4*>(bfeat) + 2 * (size_t)k;
    o[0] = f[0];
    o[1] = f[1];
    btgt[k] = tgt[slot];
}

// Window records of the two-record passes: position and v* of each
// neighbour, 24 bytes in two LDS arrays (x, y) and (z, v*x, v*y, v*z), each
// read with one aligned LDS load.
struct WinXV {
    float2* xy;
    float4* zv;
};
struct RecW {
    float2 a;
    float4 b;
};

// a two-record window record: x, y, z of one array and the first three
// components of another (v* or u), 24 bytes
__device__ __forceinline__ Rec3x2 load_xyz2(const float4* __restrict__ a, const float4* __restrict__ b, int j) {
    return Rec3x2{*reinterpret_cast<const float3*>(&a[j]), *reinterpret_cast<const float3*>(&b[j])};
}
__device__ __forceinline__ void store_xyz2(WinXV w, int k, const Rec3x2& r) {
    w.xy[k] = make_float2(r.a.x, r.a.y);
    w.zv[k] = make_float4(r.a.z, r.b.x, r.b.y, r.b.z);
}

__device__ __forceinline__ void stage_xv(const Dev& D, int n, WinXV w) {
    win_stage(n, [&](int j) { return Rec2{D.x[j], D.vs[j]}; }, [&](int k, Rec2 r) {
        w.xy[k] = make_float2(r.a.x, r.a.y);
        w.zv[k] = make_float4(r.a.z, r.b.x, r.b.y, r.b.z);
    });
}

// density_change with the fluid records read from the LDS window
__device__ __forceinline__ float density_change_win(const Dev& D, const Slot& s, float4 pa, float4 va, float m,
                                                    float rho0, int c, WinXV w, const List16& L) {
    float sf = 0.0f;
    each_in_win16(L, nf_of(c), [&](int k) { return RecW{w.xy[k], w.zv[k]}; }, [&](int, RecW rec) {
        sf += dchange_term(D, pa, va, rec.a.x, rec.a.y, rec.b.x, rec.b.y, rec.b.z, rec.b.w);
    });
    const float sb = dchange_boundary(D, s, pa, va, c);
    return m * sf + rho0 * sb;
}

// LDS of the two-record window (WIN_MAX records of 8 + 16 bytes)
#define WIN_XV_DECL(N)                           \
    __shared__ float2 s_wxy[N];                  \
    __shared__ float4 s_wzv[N];                  \
    const WinXV win{s_wxy, s_wzv};

// ---- pass A of the DFSPH loops -------------------------------------------
// Density loop (MODE 0, DFSPH.cpp:572-643): rho* = rho + dt Drho/Dt, clamp
// rho* >= rho0, kappa = (rho*-rho0)/dt^2 alpha -> xk.w = kappa/rho; partial
// sum of rho*/rho0.  Divergence loop (MODE 1, DFSPH.cpp:780-836): Drho/Dt,
// clamp, kappaV = Drho/Dt alpha / dt -> xk.w = kappaV/rho; partial sum of
// Drho/Dt.  projected (first divergence iteration of a warm-started solve):
// the accumulated stiffness kvsum is clamped at zero instead of the
// increment, so a guess that was too large is taken back where the fluid
// now expands; the error is the mean applied |increment| / alpha.  Later
// iterations are the reference's (projecting every iteration is unstable:
// signed Jacobi oscillates).

// what a particle of pass A reads of itself
struct OwnAB {
    float4 pa, va;
    float m, rho0, dt, alpha, acc;
};
template <int MODE>
__device__ __forceinline__ OwnAB own_ab(const Dev& D, const Slot& s) {
    OwnAB o{};
    if (!s.live) return o;
    o.pa = D.x[s.i];
    o.va = D.vs[s.i];   // v* carries rho in w (as v does), one array fewer to stream
    o.dt = D.dt[s.scene];
    o.alpha = D.alpha[s.i];
    o.acc = MODE == 0 ? D.ksum[s.i] : D.kvsum[s.i];
    const int ph = phase_of(o.pa);
    o.m = D.mass[ph];
    o.rho0 = D.rho0[ph];
    return o;
}

// one particle of pass A; nw > 0: its fluid neighbours are in the LDS window
template <int MODE>
__device__ __forceinline__ double pass_ab_particle(const Dev& D, const Slot& s, const OwnAB& o, int c, int nw,
                                                   WinXV win, const List16& L, int projected) {
    if (!s.live) return 0.0;
    const int i = s.i;
    const float4 pa = o.pa, va = o.va;
    const float m = o.m, rho0 = o.rho0, dt = o.dt, rhoa = va.w;
    float dd = nw ? density_change_win(D, s, pa, va, m, rho0, c, win, L) : density_change(D, s, pa, va, m, rho0, c);
    if (MODE == 0) {
        float rs = rhoa + dt * dd;
        if (rs < rho0) rs = rho0;
        const float dev = (rs - rho0) * o.alpha;   // kappa * dt^2
        D.xk[i] = make_float4(pa.x, pa.y, pa.z, dev / (dt * dt) / rhoa);
        D.ksum[i] = o.acc + dev;
        return (double)rs / (double)rho0;
    }
    if (!projected) {
        if (dd < 0.0f || (rhoa + dd * dt) < rho0) dd = 0.0f;
        const float kv = dd * o.alpha / dt;
        D.xk[i] = make_float4(pa.x, pa.y, pa.z, kv / rhoa);
        D.kvsum[i] = o.acc + kv * dt;
        return (double)dd;
    }
    // below rest density the solve applies no stiffness: the whole guess is
    // taken back
    const float alpha = o.alpha, K = o.acc;
    float inc = dd * alpha;
    if ((rhoa + dd * dt) < rho0 || inc < -K) inc = -K;
    const float kv = inc / dt;
    D.xk[i] = make_float4(pa.x, pa.y, pa.z, kv / rhoa);
    D.kvsum[i] = K + inc;
    return (double)(fabsf(inc) / alpha);
}

// Gathers-only build: one tile per block.
template <int MODE>
__device__ __forceinline__ void pass_ab_gather(const Dev& D, int projected) {
    __shared__ double lds[4];
    Slot s = slot_nb(D, true);
    if (!D.active[s.scene]) {   // uniform: a converged scene's tiles contribute nothing
        if (threadIdx.x == 0) D.partial[2 * s.tile] = 0.0;
        return;
    }
    const int c = s.live ? D.cnt[s.i] : 0;
    const OwnAB o = own_ab<MODE>(D, s);
    const List16 L{};
    write_partial(D, s.tile, 0, pass_ab_particle<MODE>(D, s, o, c, 0, WinXV{nullptr, nullptr}, L, projected), lds);
}

// LDS-window build: two tiles per block (window_pair)
template <int MODE>
__device__ __forceinline__ void pass_ab_window_pair(const Dev& D, int projected) {
    WIN_XV_DECL(WIN_MAX)
    window_pair<MODE == 0 ? WIN_DENSITY : WIN_DIVERGENCE, true>(
        D, [&](const Slot& s, int) { return own_ab<MODE>(D, s); }, [&](int j) { return load_xyz2(D.x, D.vs, j); },
        [&](int k, const Rec3x2& r) { store_xyz2(win, k, r); },
        [&](const Slot& s, int c, int nw, const List16& L, const OwnAB& o) {
            return pass_ab_particle<MODE>(D, s, o, c, nw, win, L, projected);
        });
}

template <bool WIN>
__global__ __launch_bounds__(256, WIN ? 2 : 1) void k_df_pass_density(Dev D) {
    if constexpr (WIN) pass_ab_window_pair<0>(D, 0);
    else pass_ab_gather<0>(D, 0);
}

template <bool WIN>
__global__ __launch_bounds__(256, WIN ? 2 : 1) void k_df_pass_divergence(Dev D, int projected) {
    if constexpr (WIN) pass_ab_window_pair<1>(D, projected);
    else pass_ab_gather<1>(D, projected);
}

// pass B of both loops (DFSPH.cpp:647-693, 840-883):
// v* -= dt (m sum (kr_a + kr_b) gradW + rho0 sum V_b kr_a gradW), kr = xk.w
struct OwnV {
    float4 pa, v;
    float m, rho0, dt;
};
__device__ __forceinline__ OwnV own_v(const Dev& D, const Slot& s, int c) {
    OwnV o{};
    if (!s.live) return o;
    o.pa = D.xk[s.i];
    o.v = D.vs[s.i];
    o.dt = D.dt[s.scene];
    const int ph = phase_of_cnt(c);   // (the phase from the count word: no read of x)
    o.m = D.mass[ph];
    o.rho0 = D.rho0[ph];
    return o;
}
// one particle of pass B; nw > 0: its fluid neighbours are in the window wxk
__device__ __forceinline__ void velocity_particle(const Dev& D, const Slot& s, const OwnV& o, int c, int nw,
                                                  const float4* wxk, const List16& L) {
    if (!s.live) return;
    const float4 pa = o.pa;
    const float ka = pa.w;
    float fx = 0.f, fy = 0.f, fz = 0.f, bx = 0.f, by = 0.f, bz = 0.f;
    auto fluid = [&](int, float4 pb) {
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        float g = kernel_gradfac_sq(dot3f(dx, dy, dz, dx, dy, dz), D) * (ka + pb.w);
        fx += dx * g; fy += dy * g; fz += dz * g;
    };
    if (nw) {
        each_in_win16(L, nf_of(c), [&](int k) { return wxk[k]; }, fluid);
    } else {
        each_fluid_lu<8>(D, s, nf_of(c), [&](int j) { return D.xk[j]; }, fluid);
    }
    each_boundary_lu<8>(D, s, nb_of(c), [&](int j) { return D.xb[j]; }, [&](int j, float4 pb) {
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        float g = kernel_gradfac_sq(dx * dx + dy * dy + dz * dz, D) * pb.w;
        bx += dx * g; by += dy * g; bz += dz * g;
    });
    const float cf = o.dt * o.m, cb = o.dt * o.rho0 * ka;
    float4 v = o.v;
    v.x -= cf * fx + cb * bx;
    v.y -= cf * fy + cb * by;
    v.z -= cf * fz + cb * bz;
    D.vs[s.i] = v;
}

template <bool WIN>
__global__ __launch_bounds__(256, WIN ? 2 : 1) void k_df_pass_velocity(Dev D) {
    // the loop's convergence test on pass A's partial sums (scene b on
    // blocks b, b + gridDim, ...)
    if (D.redMode >= 0)
        for (int sc = blockIdx.x; sc < D.nscenes; sc += gridDim.x)
            scene_reduce_block(D, sc, D.redMode, D.redTol, D.redMinIt, D.redCap, D.activeNext);
    if constexpr (WIN) {
        __shared__ float4 s_wxk[WIN_MAX];   // window records: xk of each neighbour
        window_pair<WIN_VELOCITY, false>(
            D, [&](const Slot& s, int c) { return own_v(D, s, c); }, [&](int j) { return D.xk[j]; },
            [&](int k, const float4& r) { s_wxk[k] = r; },
            [&](const Slot& s, int c, int nw, const List16& L, const OwnV& o) {
                velocity_particle(D, s, o, c, nw, s_wxk, L);
                return 0.0;
            });
    } else {
        Slot s = slot_nb(D, true);
        if (!D.active[s.scene]) return;   // uniform per block
        const int c = s.live ? D.cnt[s.i] : 0;
        const OwnV o = own_v(D, s, c);
        const List16 L{};
        velocity_particle(D, s, o, c, 0, nullptr, L);
    }
}

// v = v*, x += dt v, removal outside [spaceMin, spaceMax] (DFSPH.cpp:705-718)
__global__ __launch_bounds__(256) void k_df_integrate(Dev D) {
    Slot s = slot_of(D, false);
    if (!s.live) return;
    const int i = s.i;
    float dt = D.dt[s.scene];
    float4 v = D.vs[i];
    float4 x = D.x[i];
    x.x += v.x * dt; x.y += v.y * dt; x.z += v.z * dt;
    D.v[i] = v;
    D.x[i] = x;
    if (!inside_domain(D.grids[s.scene], x)) {
        D.id[i] = -1;
        atomicAdd(&D.counters[C_REMOVED], 1);
    }
}

// ============================================================ IISPH ======
// density (IISph_M.cpp:101-112), computed before a_adv so that a_adv reads
// this step's densities; published as v.w for the a_adv gathers.
__global__ __launch_bounds__(256) void k_ii_density(Dev D) {
    Slot s = slot_nb(D, false);
    if (!s.live) return;
    const int i = s.i;
    float4 pa = D.x[i];
    int ph = phase_of(pa);
    float density = density_sum(D, s, pa, D.mass[ph], D.rho0[ph], D.cnt[i]);
    D.rho[i] = density;
    D.v[i].w = density;
}

// a_adv (IISph_M.cpp:364-476), d_ii (114-130) and v_adv (135)
__global__ __launch_bounds__(256) void k_ii_aadv_dii(Dev D) {
    Slot s = slot_nb(D, false);
    if (!s.live) return;
    const int i = s.i;
    float4 pa = D.x[i], va = D.v[i];
    int ph = phase_of(pa);
    float m = D.mass[ph], rho0 = D.rho0[ph], alpha = D.visc[ph];
    float rhoa = va.w;
    const float nu10 = 10.0f * D.nu, Hcs2 = 2.0f * D.Hs * D.cs;
    const bool multi = D.nphases > 1;
    float ax = D.gx, ay = D.gy, az = D.gz;
    float dfx = 0.f, dfy = 0.f, dfz = 0.f, dbx = 0.f, dby = 0.f, dbz = 0.f;
    int c = D.cnt[i];
    each_fluid_lu<4>(D, s, nf_of(c), [&](int j) { return Rec2{D.x[j], D.v[j]}; }, [&](int j, Rec2 rec) {
        const float4 pb = rec.a, vb = rec.b;   // vb = (v, rho)
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        float r2 = dx * dx + dy * dy + dz * dz;
        float g = kernel_gradfac(fast_sqrt(r2), D);
        float pro = (va.x - vb.x) * dx + (va.y - vb.y) * dy + (va.z - vb.z) * dz;
        int phb = multi ? phase_of(pb) : ph;
        float f;
        if (phb == ph) {
            f = g * (pro * fast_rcp(r2 + D.visc_eps) * m * fast_rcp(vb.w) * nu10);   // no pro<0 gate (IISph_M.cpp:377)
        } else {
            float mb = D.mass[phb], al = 0.5f * (alpha + D.visc[phb]);
            float nuv = al * Hcs2 / (rhoa + vb.w);
            f = pro < 0.0f ? g * (-0.5f * (m + mb) * (-nuv * pro * fast_rcp(r2 + D.visc_eps))) : 0.0f;
        }
        ax += dx * f; ay += dy * f; az += dz * f;
        dfx += dx * g; dfy += dy * g; dfz += dz * g;
    });
    each_boundary_lu<4>(D, s, nb_of(c), [&](int j) { return Rec2{D.xb[j], D.vb[j]}; }, [&](int j, Rec2 rec) {
        const float4 pb = rec.a, vb = rec.b;
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        float r2 = dx * dx + dy * dy + dz * dz;
        float g = kernel_gradfac(fast_sqrt(r2), D);
        float pro = (va.x - vb.x) * dx + (va.y - vb.y) * dy + (va.z - vb.z) * dz;
        float f = g * (pro * fast_rcp(r2 + D.visc_eps) * (rho0 * pb.w) / rhoa * nu10);
        ax += dx * f; ay += dy * f; az += dz * f;
        float gw = g * pb.w;
        dbx += dx * gw; dby += dy * gw; dbz += dz * gw;
    });
    float dt = D.dt[s.scene];
    float cf = -m / (rhoa * rhoa) * dt * dt, cb = -rho0 / (rhoa * rhoa) * dt * dt;
    D.f4a[i] = make_float4(cf * dfx + cb * dbx, cf * dfy + cb * dby, cf * dfz + cb * dbz, 0.0f);
    D.vs[i] = make_float4(va.x + ax * dt, va.y + ay * dt, va.z + az * dt, 0.0f);
}

// a_ii, rho_adv, p_l = 0.5 p(t-dt) (IISph_M.cpp:243-278); publishes
// xk = (x, p_l/rho^2) for the sum_j d_ij p_j gathers.  With feat (the
// learned warm start) also the model's eight features per particle:
//   0 (rho/rho0 - 1) * 100              density deviation
//   1 (rho_adv/rho0 - 1) * 100          predicted density deviation
//   2 (fluid + boundary neighbours)/30
//   3 log(1 + s_prev), s = -a_ii p / rho0 * 1000   previous pressure, as the
//     relative density change it removes (in 1e-3)
//   4 a_ii / (dt^2 a_lat)               a_ii against a lattice particle's
//   5 |v_adv| dt / spacing
//   6 distance to the nearest boundary particle / spacing (2 if none)
//   7 dt / dt_max
template <bool WARM>
__global__ __launch_bounds__(256) void k_ii_aii(Dev D, float* __restrict__ feat) {
    Slot s = slot_nb(D, false);
    if (!s.live) return;
    const int i = s.i;
    float4 pa = D.x[i], va = D.vs[i], dii = D.f4a[i];
    int ph = phase_of(pa);
    float m = D.mass[ph], rho0 = D.rho0[ph];
    float rhoa = D.rho[i];
    float dt = D.dt[s.scene];
    float cji = m * dt * dt / (rhoa * rhoa);
    float t1f = 0.f, t1b = 0.f, t2f = 0.f, t2b = 0.f, cself = 0.f;
    const float sp = D.spacing;
    float bmin2 = 4.0f * sp * sp;
    int c = D.cnt[i];
    const bool cons = D.mpCons && D.nphases > 1;
    each_fluid_lu<4>(D, s, nf_of(c), [&](int j) { return Rec2{D.x[j], D.vs[j]}; }, [&](int j, Rec2 rec) {
        const float4 pb = rec.a, vb = rec.b;
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        float g = kernel_gradfac(fast_sqrt(dx * dx + dy * dy + dz * dz), D);
        float gx = dx * g, gy = dy * g, gz = dz * g;
        t1f += gx * (va.x - vb.x) + gy * (va.y - vb.y) + gz * (va.z - vb.z);
        if (!cons) {
            t2f += (dii.x - cji * gx) * gx + (dii.y - cji * gy) * gy + (dii.z - cji * gz) * gz;
        } else {
            // d_ji of the number-density force: dt^2 m_a^2 / (m_b rho_a^2) gradW
            const float cb = cji * (m / D.mass[phase_of(pb)]);
            const float g2 = gx * gx + gy * gy + gz * gz;
            t2f += dii.x * gx + dii.y * gy + dii.z * gz - cb * g2;
            cself += cb * g2;
        }
    });
    each_boundary_lu<4>(D, s, nb_of(c), [&](int j) { return Rec2{D.xb[j], D.vb[j]}; }, [&](int j, Rec2 rec) {
        const float4 pb = rec.a, vb = rec.b;
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        float g = kernel_gradfac(fast_sqrt(r2), D) * pb.w;
        float gx = dx * g, gy = dy * g, gz = dz * g;
        t1b += gx * (va.x - vb.x) + gy * (va.y - vb.y) + gz * (va.z - vb.z);   // v_adv (IISph_M.cpp:269 uses v)
        t2b += dii.x * gx + dii.y * gy + dii.z * gz;
        if (WARM) bmin2 = fminf(bmin2, r2);
    });
    float aii = m * t2f + rho0 * t2b;
    if (aii == 0.0f) aii = 1.0f;
    D.alpha[i] = aii;
    const float radv = rhoa + dt * (m * t1f + rho0 * t1b);
    D.kr0[i] = radv;
    const float pprev = D.ksum[i];
    float pl = 0.5f * pprev;
    D.kr[i] = pl;
    D.xk[i] = make_float4(pa.x, pa.y, pa.z, cons ? pl * (m * m / (rhoa * rhoa)) : pl / (rhoa * rhoa));
    if (cons) D.f2[i] = m * cself;   // sum_b m_a d_ba.gradW_ab: the p_a term of the Jacobi update
    if (!WARM) return;
    float4* f = reinterpret_cast<float4*>(feat) + 2 * (size_t)i;
    f[0] = make_float4((rhoa / rho0 - 1.0f) * 100.0f, (radv / rho0 - 1.0f) * 100.0f,
                       (float)(nf_real(c) + nb_of(c)) * (1.0f / 30.0f),
                       log1pf(fmaxf(-aii * pprev / rho0 * 1000.0f, 0.0f)));
    f[1] = make_float4(aii / (dt * dt * D.aii_lat), sqrtf(va.x * va.x + va.y * va.y + va.z * va.z) * dt / sp,
                       sqrtf(bmin2) / sp, dt / D.dt_max);
}

// Learned start of the IISPH pressure solve (not in the reference, whose
// start is p_l = 0.5 p(t-dt), IISph_M.cpp:277): p_l = the model's converged
// pressure, from its prediction y of the change u - u_prev of
// u = log(1 + s), s = -a_ii p / rho0 * 1000; the guess may fall anywhere
// below but rise only as far as the larger of the previous pressure and
// the particle's own Jacobi estimate (rho0 - rho_adv)/a_ii.  It is taken
// whole: the Jacobi iteration corrects a start in both directions (unlike
// the DFSPH loops, which only add stiffness), and on the 8M two-phase tank
// 0.5, 0.7, 0.85 and 1.0 of it need 11.4, 9.5, 8.2 and 7.1 iterations.
__global__ __launch_bounds__(256) void k_ii_warm_apply(Dev D, const float* __restrict__ y,
                                                     const float* __restrict__ feat) {
    Slot s = slot_of(D, false);
    if (!s.live) return;
    const int i = s.i;
    const float4 pa = D.x[i];
    const int ph = phase_of(pa);
    const float rho0 = D.rho0[ph], aii = D.alpha[i], rhoa = D.rho[i];
    const float u = feat[8 * (size_t)i + 3] + y[i];
    float pl = 0.0f;
    if (aii < 0.0f) {
        const float pm = expm1f(fminf(fmaxf(u, 0.0f), 12.0f)) * 1.0e-3f * rho0 / (-aii);
        const float pj = fmaxf((rho0 - D.kr0[i]) / aii, 0.0f);
        pl = fminf(pm, fmaxf(D.ksum[i], pj));
    }
    D.kr[i] = pl;
    const float m = D.mass[ph];
    D.xk[i].w = (D.mpCons && D.nphases > 1) ? pl * (m * m / (rhoa * rhoa)) : pl / (rhoa * rhoa);
}

// converged pressure in the model's units, as the change from feature 3
__global__ __launch_bounds__(256) void k_ii_targets(Dev D, const float* __restrict__ feat, float* __restrict__ tgt) {
    Slot s = slot_of(D, false);
    if (!s.live) return;
    const int i = s.i;
    const float rho0 = D.rho0[phase_of(D.x[i])];
    tgt[i] = log1pf(fmaxf(-D.alpha[i] * D.kr[i] / rho0 * 1000.0f, 0.0f)) - feat[8 * (size_t)i + 3];
}

// sum_j d_ij p_j = dt^2 sum_f -m p_j/rho_j^2 gradW (IISph_M.cpp:289-309);
// stores u = d_ii p_l + sum_j d_ij p_j, the per-particle part of the
// Jacobi update that neighbours read.
struct OwnS {
    float4 pa, dii;
    float m, dt, pl;
};
__device__ __forceinline__ OwnS own_s(const Dev& D, const Slot& s, int c) {
    OwnS o{};
    if (!s.live) return o;
    // own position from the xk records the pass gathers (xyz = x during the
    // loop), the phase from the count word: no read of x
    o.pa = D.xk[s.i];
    o.dii = D.f4a[s.i];
    o.pl = D.kr[s.i];
    o.dt = D.dt[s.scene];
    o.m = D.mass[phase_of_cnt(c)];
    return o;
}
__device__ __forceinline__ void sumdij_particle(const Dev& D, const Slot& s, const OwnS& o, int c, int nw,
                                                const float4* wxk, const List16& L) {
    if (!s.live) return;
    const float4 pa = o.pa;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    auto fluid = [&](int, float4 pb) {
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        float g = kernel_gradfac_sq(dot3f(dx, dy, dz, dx, dy, dz), D) * pb.w;
        sx += dx * g; sy += dy * g; sz += dz * g;
    };
    if (nw) {
        each_in_win16(L, nf_of(c), [&](int k) { return wxk[k]; }, fluid);
    } else {
        each_fluid_lu<8>(D, s, nf_of(c), [&](int j) { return D.xk[j]; }, fluid);
    }
    const float dt = o.dt, m = o.m;
    // the records carry p/rho^2, or m^2 p/rho^2 with the number-density
    // force between phases (whose d_ij is dt^2 m_j^2 / (m_i rho_j^2) gradW)
    const float cf = (D.mpCons && D.nphases > 1) ? -dt * dt / m : -m * dt * dt;
    const float4 dii = o.dii;
    const float pl = o.pl;
    D.f4b[s.i] = make_float4(dii.x * pl + cf * sx, dii.y * pl + cf * sy, dii.z * pl + cf * sz, 0.0f);
}

template <bool WIN>
__global__ __launch_bounds__(256, WIN ? 2 : 1) void k_ii_sumdij(Dev D) {
    if constexpr (WIN) {
        __shared__ float4 s_wxk[WIN_MAX];   // window records: xk of each neighbour
        window_pair<WIN_SUMDIJ, false>(
            D, [&](const Slot& s, int c) { return own_s(D, s, c); }, [&](int j) { return D.xk[j]; },
            [&](int k, const float4& r) { s_wxk[k] = r; },
            [&](const Slot& s, int c, int nw, const List16& L, const OwnS& o) {
                sumdij_particle(D, s, o, c, nw, s_wxk, L);
                return 0.0;
            });
    } else {
        Slot s = slot_nb(D, true);
        if (!D.active[s.scene]) return;   // uniform per block
        const int c = s.live ? D.cnt[s.i] : 0;
        const OwnS o = own_s(D, s, c);
        const List16 L{};
        sumdij_particle(D, s, o, c, 0, nullptr, L);
    }
}

struct OwnP {
    float4 pa, ua, dia;
    float m, rho0, rhoa, pla, dt, aii, radv, f2;
};
__device__ __forceinline__ OwnP own_p(const Dev& D, const Slot& s) {
    OwnP o{};
    if (!s.live) return o;
    const int i = s.i;
    o.pa = D.x[i];
    o.ua = D.f4b[i];
    o.dia = D.f4a[i];
    o.rhoa = D.rho[i];
    o.pla = D.kr[i];
    o.dt = D.dt[s.scene];
    o.aii = D.alpha[i];
    o.radv = D.kr0[i];
    o.f2 = (D.mpCons && D.nphases > 1) ? D.f2[i] : 0.0f;
    const int ph = phase_of(o.pa);
    o.m = D.mass[ph];
    o.rho0 = D.rho0[ph];
    return o;
}
// one particle's Jacobi update; nw > 0: its fluid neighbours (x and u =
// d_ii p_l + sum_j d_ij p_j) are in the window
__device__ __forceinline__ double pressure_particle(const Dev& D, const Slot& s, const OwnP& o, int c, int nw,
                                                    WinXV win, const List16& L) {
    if (!s.live) return 0.0;
    const int i = s.i;
    const float4 pa = o.pa, ua = o.ua, dia = o.dia;
    const float m = o.m, rho0 = o.rho0, rhoa = o.rhoa, pla = o.pla, dt = o.dt;
    float cji = dt * dt * m / (rhoa * rhoa);
    // own sum_j d_ij p_j
    float sax = ua.x - dia.x * pla, say = ua.y - dia.y * pla, saz = ua.z - dia.z * pla;
    float ff = 0.f, fb = 0.f;
    // with the number-density force between phases d_ji depends on the
    // neighbour's mass; its part sum_b m_a d_ba.gradW_ab p_a is then taken
    // whole from k_ii_aii (f2) instead of per pair
    const bool cons = D.mpCons && D.nphases > 1;
    const float cj = cons ? 0.0f : cji;
    auto fluid = [&](float bx, float by, float bz, float ux, float uy, float uz) {
        float dx = pa.x - bx, dy = pa.y - by, dz = pa.z - bz;
        float g = kernel_gradfac_sq(dot3f(dx, dy, dz, dx, dy, dz), D);
        float gx = dx * g, gy = dy * g, gz = dz * g;
        // (sum d_ij p_j|a - d_jj p_j) - (sum d_jk p_k|b - d_ji p_a)
        float tx = sax - ux + cj * gx * pla;
        float ty = say - uy + cj * gy * pla;
        float tz = saz - uz + cj * gz * pla;
        ff += dot3f(tx, ty, tz, gx, gy, gz);
    };
    if (nw) {
        each_in_win16(L, nf_of(c), [&](int k) { return RecW{win.xy[k], win.zv[k]}; },
                      [&](int, RecW r) { fluid(r.a.x, r.a.y, r.b.x, r.b.y, r.b.z, r.b.w); });
    } else {
        each_fluid_lu<4>(D, s, nf_of(c), [&](int j) { return Rec2{D.x[j], D.f4b[j]}; }, [&](int j, Rec2 rec) {
            fluid(rec.a.x, rec.a.y, rec.a.z, rec.b.x, rec.b.y, rec.b.z);
        });
    }
    each_boundary_lu<8>(D, s, nb_of(c), [&](int j) { return D.xb[j]; }, [&](int j, float4 pb) {
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        float g = kernel_gradfac_sq(dx * dx + dy * dy + dz * dz, D) * pb.w;
        fb += (sax * dx + say * dy + saz * dz) * g;
    });
    float fin = m * ff + rho0 * fb;
    if (cons) fin += o.f2 * pla;
    const float aii = o.aii, radv = o.radv;
    float t = (1.0f - D.omega) * pla + (D.omega / aii) * (rho0 - radv - fin);
    double contrib = 0.0;
    if (t < 0.0f) {
        t = 0.0f;
    } else {
        contrib = ((double)radv + (double)aii * (double)pla + (double)fin - (double)rho0) / (double)rho0;
    }
    // the new p_l, published in place: this pass gathers x and
    // sum_j d_ij p_j of its neighbours, never their p_l
    D.f1[i] = t;
    D.kr[i] = t;
    D.xk[i].w = cons ? t * (m * m / (rhoa * rhoa)) : t / (rhoa * rhoa);
    return contrib;
}

// relaxed Jacobi update (IISph_M.cpp:312-351) into f1; partial sum of the
// relative density error of the particles that are not clamped.
template <bool WIN>
__global__ __launch_bounds__(256, WIN ? 2 : 1) void k_ii_pressure(Dev D) {
    if constexpr (WIN) {
        WIN_XV_DECL(WIN_MAX)
        window_pair<WIN_PRESSURE, true>(
            D, [&](const Slot& s, int) { return own_p(D, s); }, [&](int j) { return load_xyz2(D.x, D.f4b, j); },
            [&](int k, const Rec3x2& r) { store_xyz2(win, k, r); },
            [&](const Slot& s, int c, int nw, const List16& L, const OwnP& o) {
                return pressure_particle(D, s, o, c, nw, win, L);
            });
    } else {
        __shared__ double lds[4];
        Slot s = slot_nb(D, true);
        if (!D.active[s.scene]) {   // uniform: a converged scene's tiles contribute nothing
            if (threadIdx.x == 0) D.partial[2 * s.tile] = 0.0;
            return;
        }
        const int c = s.live ? D.cnt[s.i] : 0;
        const OwnP o = own_p(D, s);
        const List16 L{};
        write_partial(D, s.tile, 0, pressure_particle(D, s, o, c, 0, WinXV{nullptr, nullptr}, L), lds);
    }
}

// IISph_M::ii_computeForcePressure (IISph_M.cpp:478-684) fused with the
// update v = v_adv + dt a_p, x += dt v and domain removal.
// gate 0: every scene; gate 1 (end of a captured step): the scenes whose
// solver loop has ended, marked as integrated; gate 2 (after the loop was
// continued outside the capture): the scenes not yet integrated.
__global__ __launch_bounds__(256) void k_ii_force(Dev D, int gate) {
    Slot s = slot_nb(D, false);
    if (gate == 1) {                          // uniform per block
        if (D.active[s.scene]) return;
        if (threadIdx.x == 0) D.integ[s.scene] = 1;
    } else if (gate == 2 && D.integ[s.scene]) {
        return;
    }
    if (!s.live) return;
    const int i = s.i;
    float4 pa = D.x[i];
    int ph = phase_of(pa);
    float m = D.mass[ph], rho0 = D.rho0[ph], alpha = D.visc[ph];
    float rhoa = D.rho[i], pra = D.kr[i];
    float pa_r2 = pra / (rhoa * rhoa);
    const float Hcs2 = 2.0f * D.Hs * D.cs;
    const bool multi = D.nphases > 1;
    const bool cons = D.mpCons && multi;   // neighbour records carry m^2 p/rho^2
    float ax = 0.f, ay = 0.f, az = 0.f;
    int c = D.cnt[i];
    each_fluid_lu<8>(D, s, nf_of(c), [&](int j) { return D.xk[j]; }, [&](int j, float4 pb) {           // (x, p/rho^2)
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        float r2 = dx * dx + dy * dy + dz * dz;
        float g = kernel_gradfac(fast_sqrt(r2), D);
        int phb = multi ? phase_of(D.x[j]) : ph;
        float acce;
        if (phb == ph) {
            acce = g * (-m * (pa_r2 + (cons ? pb.w / (m * m) : pb.w)));
        } else {
            float4 va = D.v[i], vb = D.v[j];
            float mb = D.mass[phb], al = 0.5f * (alpha + D.visc[phb]);
            acce = g * (-1.0f / m * (pa_r2 * m * m + (cons ? pb.w : pb.w * mb * mb)));
            float pro = (va.x - vb.x) * dx + (va.y - vb.y) * dy + (va.z - vb.z) * dz;
            float nuv = al * Hcs2 / (rhoa + vb.w);
            float pi = pro < 0.0f ? -nuv * pro * fast_rcp(r2 + D.visc_eps) : 0.0f;
            acce += g * (-0.5f * (m + mb) * pi);
        }
        ax += dx * acce; ay += dy * acce; az += dz * acce;
    });
    if (pra > 0.0f) {
        each_boundary_lu<8>(D, s, nb_of(c), [&](int j) { return D.xb[j]; }, [&](int j, float4 pb) {
            float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
            float g = kernel_gradfac(fast_sqrt(dx * dx + dy * dy + dz * dz), D);
            // p_a/rho_a^2: the boundary term the pressure solve predicts
            // (IISph_M.cpp:127, 271, 338), not the 2 p_a/rho_a^2 of IISph_M.cpp:558-559
            float acce = g * (-rho0 * pb.w * pa_r2);
            ax += dx * acce; ay += dy * acce; az += dz * acce;
        });
    }
    float dt = D.dt[s.scene];
    float4 vadv = D.vs[i];
    D.f4b[i] = make_float4(ax, ay, az, 0.0f);
    float4 vn = make_float4(vadv.x + ax * dt, vadv.y + ay * dt, vadv.z + az * dt, 0.0f);
    float4 xn = make_float4(pa.x + vn.x * dt, pa.y + vn.y * dt, pa.z + vn.z * dt, pa.w);
    D.v2[i] = vn;
    D.x2[i] = xn;
    D.ksum[i] = pra;
    if (!inside_domain(D.grids[s.scene], xn)) {
        D.id[i] = -1;
        atomicAdd(&D.counters[C_REMOVED], 1);
    }
}

// =========================================================== PCISPH ======
// Solenthaler & Pajarola's predictive-corrective loop as PciSph restates it
// (PciSph.cpp:97-126).  Gather records: xk = (x*, phase) for the density
// pass, vs = (x, p/rho*^2) for the pressure pass; v.w carries the last
// predicted density for the next step's viscosity.  The position prediction
// (pci_predictPosition, PciSph.cpp:289-306) of each iteration is fused into
// the pass before it (k_pci_aext for the first, k_pci_pressure after).

// frame 0: no prediction exists yet; fluid is added at rest density
// (SphWithIO.cpp:1175 sets FluidPart::density = 1000 before adding)
__global__ __launch_bounds__(256) void k_pci_rest_density(Dev D) {
    Slot s = slot_of(D, false);
    if (!s.live) return;
    D.v[s.i].w = D.rho0[phase_of(D.x[s.i])];
}

// pci_forceExceptPressure (PciSph.cpp:235-281, viscosity helpers 172-232):
// a = g + viscosity, p = 0; first prediction x* = x + (v + a dt) dt
__global__ __launch_bounds__(256) void k_pci_aext(Dev D) {
    Slot s = slot_nb(D, false);
    if (!s.live) return;
    const int i = s.i;
    float4 pa = D.x[i], va = D.v[i];
    int ph = phase_of(pa);
    float m = D.mass[ph], alpha = D.visc[ph], rho0 = D.rho0[ph];
    float rhoa = va.w;
    const float Hcs2 = 2.0f * D.Hs * D.cs;
    const bool multi = D.nphases > 1;
    float ax = D.gx, ay = D.gy, az = D.gz;
    int c = D.cnt[i];
    each_fluid_lu<4>(D, s, nf_of(c), [&](int j) { return Rec2{D.x[j], D.v[j]}; }, [&](int j, Rec2 rec) {
        const float4 pb = rec.a, vb = rec.b;   // vb = (v, rho)
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        float r2 = dx * dx + dy * dy + dz * dz;
        float pro = (va.x - vb.x) * dx + (va.y - vb.y) * dy + (va.z - vb.z) * dz;
        float g = kernel_gradfac(fast_sqrt(r2), D);
        int phb = multi ? phase_of(pb) : ph;
        float mm = m, al = alpha;
        if (phb != ph) {                                 /* PciSph.cpp:186-201 */
            mm = 0.5f * (m + D.mass[phb]);
            al = 0.5f * (alpha + D.visc[phb]);
        }
        float nu = al * Hcs2 / (rhoa + vb.w);
        float f = g * (-mm * (-nu * pro * fast_rcp(r2 + D.visc_eps)));
        f = pro < 0.0f ? f : 0.0f;
        ax += dx * f; ay += dy * f; az += dz * f;
    });
    each_boundary_lu<4>(D, s, nb_of(c), [&](int j) { return Rec2{D.xb[j], D.vb[j]}; }, [&](int j, Rec2 rec) {
        const float4 pb = rec.a, vb = rec.b;
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        float r2 = dx * dx + dy * dy + dz * dz;
        float pro = (va.x - vb.x) * dx + (va.y - vb.y) * dy + (va.z - vb.z) * dz;
        float g = kernel_gradfac(fast_sqrt(r2), D);
        float nu = D.solid_visc * Hcs2 / (rhoa * 2.0f);
        float f = g * (-rho0 * pb.w * (-nu * pro * fast_rcp(r2 + D.visc_eps)));
        f = pro < 0.0f ? f : 0.0f;
        ax += dx * f; ay += dy * f; az += dz * f;
    });
    const float dt = D.dt[s.scene];
    D.f4a[i] = make_float4(ax, ay, az, 0.0f);
    D.f1[i] = 0.0f;
    D.xk[i] = make_float4(pa.x + (va.x + ax * dt) * dt, pa.y + (va.y + ay * dt) * dt,
                          pa.z + (va.z + az * dt) * dt, pa.w);
}

// pci_predictDensity (PciSph.cpp:310-352): rho* at the predicted positions
// over the neighbours of x(t), every fluid neighbour with a's mass (SP08
// number density), Kahan; error clamped at 0 (NEGLECT_NEGATIVE_PRESSURE,
// SphBase.h:32); p += delta / dt^2 * error.  Partials: sum and max of
// error / rho0.
__global__ __launch_bounds__(256) void k_pci_density(Dev D) {
    __shared__ double lds[4];
    __shared__ float ldsf[4];
    Slot s = slot_nb(D, true);
    double contrib = 0.0;
    float emax = 0.0f;
    if (s.live) {
        const int i = s.i;
        const float4 xa = D.xk[i];
        int ph = phase_of(xa);
        float m = D.mass[ph], rho0 = D.rho0[ph];
        Kahan rho;
        rho.init(m * D.W0);
        int c = D.cnt[i];
        each_fluid_lu<8>(D, s, nf_of(c), [&](int j) { return D.xk[j]; }, [&](int j, float4 pb) {
            float dx = xa.x - pb.x, dy = xa.y - pb.y, dz = xa.z - pb.z;
            float w = kernel_W(fast_sqrt(dx * dx + dy * dy + dz * dz), D);
            rho.add(j != i ? m * w : 0.0f);
        });
        each_boundary_lu<8>(D, s, nb_of(c), [&](int j) { return D.xb[j]; }, [&](int j, float4 pb) {
            float dx = xa.x - pb.x, dy = xa.y - pb.y, dz = xa.z - pb.z;
            rho.add(rho0 * pb.w * kernel_W(fast_sqrt(dx * dx + dy * dy + dz * dz), D));
        });
        const float density = rho.s;
        const float err = fmaxf(density - rho0, 0.0f);
        const float dt = D.dt[s.scene];
        const float p = D.f1[i] + D.pci_delta / (dt * dt) * err;
        const float4 pa = D.x[i];
        D.f1[i] = p;
        D.rho[i] = density;
        D.v[i].w = density;
        D.vs[i] = make_float4(pa.x, pa.y, pa.z, p / (density * density));
        emax = err / rho0;
        contrib = (double)emax;
    }
    write_partial(D, s.tile, 0, contrib, lds);
    const float mx = block_max_256(emax, ldsf);
    if (threadIdx.x == 0) D.partial[2 * s.tile + 1] = (double)mx;
}

// pci_forcePressure (PciSph.cpp:399-436 with 356-397): a_p from the (x,
// p/rho*^2) records; across phases p/delta^2 with delta = rho/m; boundary
// 2 p_a/rho_a^2 when p_a > 0.  Then the next prediction x* = x + (v + (a +
// a_p) dt) dt.
__global__ __launch_bounds__(256) void k_pci_pressure(Dev D) {
    Slot s = slot_nb(D, true);
    if (!s.live) return;
    const int i = s.i;
    const float4 pa = D.vs[i];
    const float4 xa = D.x[i];
    int ph = phase_of(xa);
    float m = D.mass[ph], rho0 = D.rho0[ph];
    const float qa = pa.w;
    const bool multi = D.nphases > 1;
    float ax = 0.f, ay = 0.f, az = 0.f;
    int c = D.cnt[i];
    each_fluid_lu<8>(D, s, nf_of(c), [&](int j) { return D.vs[j]; }, [&](int j, float4 pb) {
        float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
        float g = kernel_gradfac(fast_sqrt(dx * dx + dy * dy + dz * dz), D);
        float acce;
        int phb = multi ? phase_of(D.x[j]) : ph;
        if (phb == ph) {
            acce = g * (-m * (qa + pb.w));
        } else {                                         /* SP08, PciSph.cpp:369-381 */
            float mb = D.mass[phb];
            acce = g * (-1.0f / m * (qa * m * m + pb.w * mb * mb));
        }
        ax += dx * acce; ay += dy * acce; az += dz * acce;
    });
    if (qa > 0.0f) {
        each_boundary_lu<8>(D, s, nb_of(c), [&](int j) { return D.xb[j]; }, [&](int j, float4 pb) {
            float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
            float g = kernel_gradfac(fast_sqrt(dx * dx + dy * dy + dz * dz), D);
            float acce = g * (-rho0 * pb.w * (2.0f * qa));
            ax += dx * acce; ay += dy * acce; az += dz * acce;
        });
    }
    const float dt = D.dt[s.scene];
    const float4 va = D.v[i], a = D.f4a[i];
    D.f4b[i] = make_float4(ax, ay, az, 0.0f);
    D.xk[i] = make_float4(xa.x + (va.x + (a.x + ax) * dt) * dt, xa.y + (va.y + (a.y + ay) * dt) * dt,
                          xa.z + (va.z + (a.z + az) * dt) * dt, xa.w);
}

// pci_forceTotal (PciSph.cpp:463-474) and SphBase::updateFluids
// (SphBase.cpp:377-402): a += a_p, v += a dt, x += v dt, removal; in place
__global__ __launch_bounds__(256) void k_pci_integrate(Dev D) {
    Slot s = slot_of(D, false);
    if (!s.live) return;
    const int i = s.i;
    const float dt = D.dt[s.scene];
    const float4 a0 = D.f4a[i], ap = D.f4b[i];
    const float4 a = make_float4(a0.x + ap.x, a0.y + ap.y, a0.z + ap.z, 0.0f);
    float4 v = D.v[i], x = D.x[i];
    v.x += a.x * dt; v.y += a.y * dt; v.z += a.z * dt;
    x.x += v.x * dt; x.y += v.y * dt; x.z += v.z * dt;
    D.f4a[i] = a;
    D.v[i] = v;
    D.x[i] = x;
    if (!inside_domain(D.grids[s.scene], x)) {
        D.id[i] = -1;
        atomicAdd(&D.counters[C_REMOVED], 1);
    }
}

// plain streaming copy (calibration of the byte counters, and what a copy
// reaches): four 16-byte loads per thread in flight before their stores
__global__ __launch_bounds__(256) void k_copy_probe(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
    const size_t stride = (size_t)gridDim.x * 256;
    size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    for (; k + 3 * stride < n; k += 4 * stride) {
        const float4 r0 = a[k], r1 = a[k + stride], r2 = a[k + 2 * stride], r3 = a[k + 3 * stride];
        b[k] = r0;
        b[k + stride] = r1;
        b[k + 2 * stride] = r2;
        b[k + 3 * stride] = r3;
    }
    for (; k < n; k += stride) b[k] = a[k];
}

// start of a solver loop: every scene active, iteration counts zero
__global__ __launch_bounds__(64) void k_loop_start(Dev D) {
    const int sc = blockIdx.x * 64 + threadIdx.x;
    if (sc >= D.nscenes) return;
    D.active[sc] = 1;
    D.iters[sc] = 0;
    D.integ[sc] = 0;
}

__global__ void k_fill_i32(int* p, int v, int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

inline int tiles_of(const Dev& D) { return D.npad / TILE; }

}  // namespace

void launch_copy_probe(const float4* a, float4* b, size_t n, hipStream_t st) {
    int blocks = int(std::min<size_t>((n + 1023) / 1024, 65536));
    if (n > 0) hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, st, a, b, n);
}

void launch_copy_f4(float4* dst, const float4* src, size_t n, hipStream_t st) { launch_copy_probe(src, dst, n, st); }

void launch_cfl(const Dev& D, float cfl, float spacing, float dtMax, float dtMin, hipStream_t st) {
    if (tiles_of(D) > 0) hipLaunchKernelGGL(k_cfl_partial, dim3(tiles_of(D)), dim3(256), 0, st, D);
    hipLaunchKernelGGL(k_cfl_reduce, dim3(D.nscenes), dim3(256), 0, st, D, cfl, spacing, dtMax, dtMin);
}
void launch_loop_start(const Dev& D, hipStream_t st) {
    hipLaunchKernelGGL(k_loop_start, dim3((D.nscenes + 63) / 64), dim3(64), 0, st, D);
}
void launch_scene_reduce(const Dev& D, int mode, float tol, int minIt, int cap, hipStream_t st) {
    hipLaunchKernelGGL(k_scene_reduce, dim3(D.nscenes), dim3(256), 0, st, D, mode, tol, minIt, cap);
}
void launch_stats(const Dev& D, double* part, double* out, hipStream_t st) {
    if (tiles_of(D) > 0) hipLaunchKernelGGL(k_stats_partial, dim3(tiles_of(D)), dim3(256), 0, st, D, part);
    hipLaunchKernelGGL(k_stats_reduce, dim3(D.nscenes), dim3(256), 0, st, D, (const double*)part, out);
}
void launch_fill_i32(int* p, int v, int n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_fill_i32, dim3((n + 255) / 256), dim3(256), 0, st, p, v, n);
}

thread_local LaunchTiming g_launch_timing;

// One block of 256 threads per tile; `lds` bytes of dynamic LDS that the
// kernel does not use cap how many blocks share a CU (the gather builds of
// the solver-loop passes run fastest below their register-limited occupancy)
#define SPH_LAUNCH_GL(kern, blocks, lds, ...)                                                          \
    do {                                                                                               \
        if ((blocks) > 0) {                                                                            \
            const LaunchTiming t_ = g_launch_timing;                                                   \
            if (t_.start) {                                                                            \
                g_launch_timing = LaunchTiming{};                                                      \
                hipExtLaunchKernelGGL(kern, dim3(blocks), dim3(256), (lds), st, t_.start, t_.stop, 0, D, \
                                      ##__VA_ARGS__);                                                  \
            } else {                                                                                   \
                hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), (lds), st, D, ##__VA_ARGS__);        \
            }                                                                                          \
        }                                                                                              \
    } while (0)
#define SPH_LAUNCH_L(kern, lds, ...) SPH_LAUNCH_GL(kern, tiles_of(D), lds, ##__VA_ARGS__)
#define SPH_LAUNCH(kern, ...) SPH_LAUNCH_L(kern, 0, ##__VA_ARGS__)

void launch_wc_pressure(const Dev& D, hipStream_t st) { SPH_LAUNCH(k_wc_pressure); }
void launch_wc_force_integrate(const Dev& D, hipStream_t st) { SPH_LAUNCH(k_wc_force); }
void launch_df_density_alpha(const Dev& D, hipStream_t st) { SPH_LAUNCH(k_df_density_alpha<false>, nullptr); }
void launch_df_density_alpha_warm(const Dev& D, float* feat, hipStream_t st) {
    SPH_LAUNCH(k_df_density_alpha<true>, feat);
}
void launch_df_aadv(const Dev& D, hipStream_t st) { SPH_LAUNCH(k_df_aadv); }
void launch_df_warm_apply(const Dev& D, const float* y, const float* feat, int mode, hipStream_t st) {
    SPH_LAUNCH(k_df_warm_apply, y, feat, mode);
}
void launch_df_targets(const Dev& D, const float* feat, float* tgt, hipStream_t st) {
    SPH_LAUNCH(k_df_targets, feat, tgt);
}
void launch_mlp_gather(const Dev& D, const float* feat, const float* tgt, float* bfeat, float* btgt, int n,
                       uint32_t seed, hipStream_t st) {
    if (n > 0)
        hipLaunchKernelGGL(k_mlp_gather, dim3((n + 255) / 256), dim3(256), 0, st, D, feat, tgt, bfeat, btgt, n, seed);
}
// Each window pass has two builds: with the LDS window (48-72 KB of LDS per
// block, two tiles per block) and without (per-lane
// gathers only: no window LDS and fewer registers, so more waves per SIMD);
// Dev::winMask picks per pass, and Dev::gatherLds caps the gather build's
// blocks per CU.
void launch_df_pass_density(const Dev& D, hipStream_t st) {
    if (D.winMask & WIN_DENSITY) SPH_LAUNCH_GL(k_df_pass_density<true>, (tiles_of(D) + 1) / 2, 0);
    else SPH_LAUNCH_L(k_df_pass_density<false>, D.gatherLds[0]);
}
void launch_df_pass_divergence(const Dev& D, int projected, hipStream_t st) {
    if (D.winMask & WIN_DIVERGENCE) SPH_LAUNCH_GL(k_df_pass_divergence<true>, (tiles_of(D) + 1) / 2, 0, projected);
    else SPH_LAUNCH_L(k_df_pass_divergence<false>, D.gatherLds[1], projected);
}
void launch_df_pass_velocity(const Dev& D, hipStream_t st) {
    if (D.winMask & WIN_VELOCITY) SPH_LAUNCH_GL(k_df_pass_velocity<true>, (tiles_of(D) + 1) / 2, 0);
    else SPH_LAUNCH_L(k_df_pass_velocity<false>, D.gatherLds[2]);
}
void launch_df_integrate(const Dev& D, hipStream_t st) { SPH_LAUNCH(k_df_integrate); }
void launch_ii_density(const Dev& D, hipStream_t st) { SPH_LAUNCH(k_ii_density); }
void launch_ii_aadv_dii(const Dev& D, hipStream_t st) { SPH_LAUNCH(k_ii_aadv_dii); }
void launch_ii_aii(const Dev& D, float* feat, hipStream_t st) {
    if (feat) SPH_LAUNCH(k_ii_aii<true>, feat);
    else SPH_LAUNCH(k_ii_aii<false>, feat);
}
void launch_ii_warm_apply(const Dev& D, const float* y, const float* feat, hipStream_t st) {
    SPH_LAUNCH(k_ii_warm_apply, y, feat);
}
void launch_ii_targets(const Dev& D, const float* feat, float* tgt, hipStream_t st) { SPH_LAUNCH(k_ii_targets, feat, tgt); }
void launch_ii_sumdij(const Dev& D, hipStream_t st) {
    if (D.winMask & WIN_SUMDIJ) SPH_LAUNCH_GL(k_ii_sumdij<true>, (tiles_of(D) + 1) / 2, 0);
    else SPH_LAUNCH_L(k_ii_sumdij<false>, D.gatherLds[3]);
}
void launch_ii_pressure(const Dev& D, hipStream_t st) {
    if (D.winMask & WIN_PRESSURE) SPH_LAUNCH_GL(k_ii_pressure<true>, (tiles_of(D) + 1) / 2, 0);
    else SPH_LAUNCH_L(k_ii_pressure<false>, D.gatherLds[4]);
}
void launch_ii_force_integrate(const Dev& D, int gate, hipStream_t st) { SPH_LAUNCH(k_ii_force, gate); }
void launch_pci_rest_density(const Dev& D, hipStream_t st) { SPH_LAUNCH(k_pci_rest_density); }
void launch_pci_aext(const Dev& D, hipStream_t st) { SPH_LAUNCH(k_pci_aext); }
void launch_pci_density(const Dev& D, hipStream_t st) { SPH_LAUNCH(k_pci_density); }
void launch_pci_pressure(const Dev& D, hipStream_t st) { SPH_LAUNCH(k_pci_pressure); }
void launch_pci_integrate(const Dev& D, hipStream_t st) { SPH_LAUNCH(k_pci_integrate); }

}  // namespace sphgpu
