// Internal definitions shared by the HIP translation units of libsphgpu.so.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cstdint>

namespace sphgpu {

constexpr int TILE = 256;            // particles per tile; every particle kernel uses 256 threads
constexpr int NBF_CAP = 64;          // fluid neighbours kept per particle (ELL list width)
// solver-loop passes that can stage their tile's neighbour window in LDS
// (bits of Dev::winMask, chosen per pass by the autotuner)
constexpr int WIN_DENSITY = 1, WIN_DIVERGENCE = 2, WIN_VELOCITY = 4, WIN_SUMDIJ = 8, WIN_PRESSURE = 16;
constexpr int WIN_ALL = 31;
constexpr int NBB_CAP = 48;          // boundary neighbours kept per particle
constexpr int REF_NEIGHBOUR_CAP = 48;  // NeigbStr::maxNeigbNum (SphBase.h:630)
constexpr int MAX_PHASES = 8;
constexpr uint32_t KEY_INVALID = 0xFFFFFFFFu;
constexpr int SOLVER_IISPH_DF = 4;   // SPHGPU_IISPH_DF: the IISph class's DF_* solvers

// Uniform grid of one scene (SphBase::SpaceGridStruct, SphBase.h:682-705).
// Cells are numbered row-major along x inside "rows" (cy, cz); rows are
// ranked in 2-D Z-order (Morton on cy, cz) so that a run of three x-adjacent
// cells is contiguous in the sorted particle order while rows that are
// neighbours in y/z stay close in memory.
// Pose of a solid (SphSolid::updatePaticles, SphBase.cpp:95-128): OpenGL
// column-major 4x4 transform of the body-frame samples, centre of mass,
// linear and angular velocity; static solids keep their samples as given.
struct SolidPose {
    float T[16];
    float com[3], v[3], w[3];
    float visc;      // SphSolid::viscosity_alpha
    int dynamic;
    int firstId;     // original boundary indices [firstId, endId) belong to this solid
    int endId;
    int pad[3];
};

struct SceneGrid {
    float minx, miny, minz;  // spaceMin - 2*smoothRadius_h (SphBase.cpp:968)
    float divisor;           // 1/smoothRadius_h (SphBase.cpp:969)
    float smin[3], smax[3];  // spaceMin / spaceMax: particles outside are removed
    int nx, ny, nz;          // cell counts (SphBase.cpp:970)
    uint32_t keyBase;        // first key of this scene
    int rowBase;             // offset of this scene's row-rank table
    int pad;
};

// Indices of device counters (int32, reset per step).
enum Counter {
    C_ZERO_DIS = 0,
    C_TOO_MANY = 1,
    C_LIST_OVF = 2,
    C_REMOVED = 3,
    C_NAN = 4,
    C_COUNT = 8
};

// Everything a kernel may need, passed by value as the kernel argument.
struct Dev {
    // ---- sizes
    int ncap;          // fluid slot capacity (padded), ELL stride
    int npad;          // fluid slots in use (upper bound, multiple of TILE)
    int nb;            // boundary particles
    int nscenes;
    int nphases;
    int solver;

    // ---- kernel constants (SphBase.h:808-829, h = smoothRadius_h / 2)
    float kh, inv_kh, sigma, sigma_h, W0;
    float gA, gB, gU, gV;   // kernel_gradfac_sq's spline pieces in r
    float h2;          // neighbour cutoff fl((d+d)^2), SphBase.cpp:883-884
    float Hs;          // smoothRadius_h (artificial viscosity length, WcSph.cpp:182)
    float visc_eps;    // 0.01*Hs*Hs (WcSph.cpp:184)
    float v0;          // spacing^3 (DFSPH.cpp:88)
    float spacing;     // m_TH.spacing_r
    float alpha_lat;   // DFSPH alpha of an interior lattice particle (feature scale)
    float aii_lat;     // IISPH a_ii / dt^2 of an interior lattice particle (feature scale)
    float dt_max;
    float warm_scale;  // PREVIOUS warm start: scale of the previous stiffness
    float learn_scale; // LEARNED warm start: scale of the bounded prediction
    int mpCons;        // IISPH with several phases: pressure solve consistent with the
                       // number-density pressure force between phases (xk.w = m^2 p/rho^2)
    int tileMap;       // 1: contiguous tile runs per XCD, 0: round robin (autotuned)
    int nbStage;       // neighbour-list groups staged in LDS: 6 or 8 (autotuned)
    int listNT;        // 1: passes read list groups with streaming loads (autotuned)
    int winMask;       // solver-loop passes (WIN_* bits) that stage packed tiles' neighbour
                       // windows in LDS; the others gather per lane (autotuned)
    int gatherLds[5];  // per window pass (bit order): unused dynamic LDS bytes of its gather
                       // build, capping its blocks per CU (0: register-limited; autotuned)
    float gx, gy, gz;
    float cs, gamma, nu, omega;
    float rho0[MAX_PHASES], mass[MAX_PHASES], visc[MAX_PHASES], tait_B[MAX_PHASES];
    float pci_delta;   // PCISPH stiffness before the division by dt^2
    float boundaryDensity;   // m_TH.boundaryDensity (IISph DF_* solvers)
    int adjustDensity;       // m_TH.adjustDensity (IISph DF_* solvers)
    float solid_visc;  // viscosity_alpha of the boundary (one value for all static solids)

    // ---- fluid state (sorted, padded slots)
    float4* x;         // xyz, w = phase
    float4* v;         // xyz, w = density of the last density pass (gathered with v)
    int* id;           // original id, -1 for padding / removed
    float* ksum;       // DFSPH accumulated kappa*dt^2; IISPH pressure (warm start)
    float* kvsum;      // DFSPH accumulated kappaV*dt
    uint32_t* key;     // cell key of each slot (KEY_INVALID for padding)
    // alternate buffers (reorder / integration targets)
    float4* x2;
    float4* v2;
    int* id2;
    float* ksum2;
    float* kvsum2;
    uint32_t* key2;

    // ---- per-step fields
    float4* xk;        // packed gather record: xyz + per-pass scalar of the neighbour
                       // (DFSPH kappa/rho, WCSPH p/rho^2, IISPH p_l/rho^2)
    float4* vs;        // DFSPH v*, IISPH v_adv
    float4* f4a;       // WCSPH acceleration, IISPH d_ii, DFSPH a_adv
    float4* f4b;       // IISPH u = d_ii p_l + sum_j d_ij p_j; pressure acceleration
    float* rho;
    float* alpha;      // DFSPH alpha; IISPH a_ii
    float* kr;         // kappa/rho (pass B input); IISPH p_l
    float* kr0;        // warm-start kappa/rho; IISPH rho_adv
    float* f1;         // WCSPH pressure; IISPH p_l (next)
    float* f2;         // per-particle scratch (divergence deviation, errors)
    float* pm;         // IISph DF_*: mass of each slot, rho0 beta v0
    float* brho;       // IISph DF_*: boundary particles' density, alpha, kappa/rho
    float* balpha;
    float* bk;

    // ---- neighbour lists.  Packed form: 16-bit entries, eight per int4,
    // group g of particle i at g*ncap + i, decoded through the tile's base
    // table nbTab[64*tile ..] (fluid slots 0..31, boundary slots 32..63).
    // Tiles whose runs do not fit the packed form keep 32-bit entries in
    // nbF/nbB (groups of four, same layout); nbTab[64*tile + 31] says which.
    int4* nbF16;
    int4* nbB16;
    // The fluid lists again, as window indices: a packed tile whose window
    // (its fluid slot runs concatenated in slot order, WIN_MAX records at
    // most) fits gets each entry as the neighbour's index in that window, so
    // a pass that stages the window in LDS reads its records with no table.
    int4* nbW16;
    int* nbTab;
    int* nbWin;        // per tile, 32 ints: offset of each fluid slot's run in the tile's
                       // window (packed tiles; [31], the flag slot, whose run is
                       // empty, holds the window size; 0 for other tiles)
    int* nbF;
    int* nbB;
    int* cnt;          // stored fluid entries | fluid neighbours found << 8 | boundary << 16

    // ---- boundary (sorted by cell once)
    float4* xb;        // xyz, w = Akinci volume V_b
    float4* vb;        // boundary velocity
    int* bid;          // original boundary index
    int* bsolid;       // solid index
    int* bscene;
    float* bkself;     // kerSumFromSelf
    uint32_t* bkey;
    // dynamic solids (rigid coupling)
    int nsolids;
    int anyDynamic;
    const SolidPose* solids;   // [nsolids]
    float4* bloc;              // body-frame sample of a dynamic solid's particle
    float4* bforce;            // reaction force on the particle (last step)
    int* brank;                // slot of original boundary index k
    float4 *xb2, *vb2, *bloc2; // re-sort targets
    int *bid2, *bsolid2, *bscene2;
    float* bkself2;
    uint32_t* bkey2;

    // ---- grid
    const SceneGrid* grids;
    const int* rowRank;
    int2* cellF;       // [start, end) of fluid slots per key
    int2* cellB;       // [start, end) of boundary particles per key
    uint32_t totalKeys;

    // ---- scene layout
    int* sceneStart;     // first sorted rank of each scene (unpadded)
    int* sceneCount;     // live particles per scene
    int* scenePadStart;  // first padded slot of each scene (multiple of TILE), [nscenes+1]
    int* tileScene;
    int* tileValid;
    const uint32_t* sceneKeyBase;  // [nscenes+1]

    // ---- per-scene solver scalars
    float* dt;           // per scene
    int* active;         // per scene: 1 while its solver loop runs
    int* active1;        // the DFSPH loops' second flag buffer (odd iterations)
    int* iters;          // per scene: iterations of the current loop
    int* integ;          // per scene: 1 once a captured IISPH step integrated it
    double* sceneErr;    // per scene
    double* partial;     // per tile partial sums (2 per tile)
    // The DFSPH loops run their convergence test inside the velocity pass
    // (pass B): its first nscenes blocks reduce the partial sums pass A left
    // and write the flag of the next iteration to activeNext, while every
    // block gates on `active`, the flag pass A ran under (the launcher swaps
    // the two buffers by iteration parity).  redMode < 0: no test.
    int* activeNext;
    int redMode, redMinIt, redCap;
    float redTol;
    int* counters;       // Counter
    float* vmaxPartial;  // per tile
};

// ---------------------------------------------------------------------------
// device helpers

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// cell coordinate exactly as SphBase.h:700-701: int((p - min) * divisor), FP32, truncation.
__device__ __forceinline__ int cell_coord(float p, float mn, float div) {
#pragma clang fp contract(off)
    float t = (p - mn) * div;
    return int(t);
}

// cell coordinates of a point inside scene grid G, clamped into the grid
__device__ __forceinline__ void cell_of(const SceneGrid& G, float4 p, int& cx, int& cy, int& cz) {
    cx = cell_coord(p.x, G.minx, G.divisor);
    cy = cell_coord(p.y, G.miny, G.divisor);
    cz = cell_coord(p.z, G.minz, G.divisor);
    cx = min(max(cx, 0), G.nx - 1);
    cy = min(max(cy, 0), G.ny - 1);
    cz = min(max(cz, 0), G.nz - 1);
}

__device__ __forceinline__ uint32_t key_of(const SceneGrid& G, const int* rowRank, int cx, int cy, int cz) {
    int row = rowRank[G.rowBase + cz * G.ny + cy];
    return G.keyBase + (uint32_t)row * (uint32_t)G.nx + (uint32_t)cx;
}

// union of the [start,end) ranges of three consecutive keys (empty = start >= end)
__device__ __forceinline__ int2 run3(const int2* __restrict__ table, uint32_t k0, int cx, int nx) {
    int s = 0x7fffffff, e = 0;
    for (int dx = -1; dx <= 1; ++dx) {
        int c = cx + dx;
        if (c < 0 || c >= nx) continue;
        int2 r = table[k0 + dx];
        if (r.y > r.x) {
            s = min(s, r.x);
            e = max(e, r.y);
        }
    }
    return make_int2(s, e);
}

// squared distance exactly as Vec23.h:170-171 ((dx*dx + dy*dy) + dz*dz, no FMA)
__device__ __forceinline__ float dis_square_exact(float4 a, float4 b) {
#pragma clang fp contract(off)
    float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return dx * dx + dy * dy + dz * dz;
}

// Workgroups are dispatched to the 8 XCDs round-robin by workgroup id, and
// each XCD has its own L2.  Tiles next to each other in key order gather the
// same neighbour particles, so XCD x is given one contiguous run of tiles:
// workgroup b runs on XCD b % 8 as the (b / 8)-th tile of that XCD's run.
// A bijection on [0, n) for any n.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcd_block(int b, int n) {
    const int x = b % NUM_XCD, q = n / NUM_XCD, r = n % NUM_XCD;
    return x * q + min(x, r) + b / NUM_XCD;
}

// tile of a workgroup under the mapping the autotuner chose (Dev::tileMap)
__device__ __forceinline__ int tile_of_block(int map) {
    return map ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
}

__device__ __forceinline__ uint32_t morton2(uint32_t a, uint32_t b) {
    auto spread = [](uint32_t v) {
        v &= 0xFFFFu;
        v = (v | (v << 8)) & 0x00FF00FFu;
        v = (v | (v << 4)) & 0x0F0F0F0Fu;
        v = (v | (v << 2)) & 0x33333333u;
        v = (v | (v << 1)) & 0x55555555u;
        return v;
    };
    return spread(a) | (spread(b) << 1);
}

// cubic spline W(r), SphBase.h:855-871, with q = r/h.
__device__ __forceinline__ float kernel_W(float r, const Dev& D) {
    float q = r * D.inv_kh;
    float qm2 = 2.0f - q;
    float w1 = D.sigma * (1.0f - 1.5f * q * q + 0.75f * q * q * q);
    float w2 = D.sigma * (0.25f * qm2 * qm2 * qm2);
    float w = q <= 1.0f ? w1 : w2;
    return q <= 2.0f ? w : 0.0f;
}

// Hardware square root and reciprocal (v_sqrt_f32 / v_rcp_f32, 1 ulp) for
// the solver passes, whose bar is 1e-5 relative; the neighbour predicate and
// the test dump keep correctly rounded arithmetic.
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// -|dW/dr| / r  (SphBase.h:890-906 divided by r), the scalar that turns x_ab
// into grad W_ab (DFSPH.cpp:156-157).  For q <= 1 the 1/r cancels:
// -(sigma_h/h)(3 - 2.25 q); for 1 < q <= 2: -0.75 sigma_h (2-q)^2 / r.
// At r = 0 the q <= 1 branch is selected (finite).
__device__ __forceinline__ float kernel_gradfac(float r, const Dev& D) {
    float q = r * D.inv_kh;
    float qm2 = 2.0f - q;
    float g1 = -(D.sigma_h * D.inv_kh) * (3.0f - 2.25f * q);
    float g2 = -0.75f * D.sigma_h * qm2 * qm2 * fast_rcp(r);
    float g = q <= 1.0f ? g1 : g2;
    return q <= 2.0f ? g : 0.0f;
}

// kernel_gradfac from r^2 for the solver-loop passes: one reciprocal square
// root gives r and 1/r of the pair, and both spline pieces are written in r
// with the constants folded (Dev::gA..gV): for r <= h,
// -(sigma_h/h)(3 - 2.25 r/h) = gA + gB r; for h < r, -0.75 sigma_h (2 - r/h)^2 / r
// = -u^2 / r with u = sqrt(0.75 sigma_h)(2 - r/h) = gU r + gV, clamped at 0
// past the support (ten VALU instructions a pair).  r^2 is floored at 1e-30
// so that coincident particles, whose pair vector x_ab is zero, get the
// finite r <= h value and add nothing.
constexpr float R2_FLOOR = 1.0e-30f;
__device__ __forceinline__ float kernel_gradfac_sq(float r2, const Dev& D) {
    const float ri = __builtin_amdgcn_rsqf(fmaxf(r2, R2_FLOOR));
    const float r = r2 * ri;
    const float g1 = __builtin_fmaf(r, D.gB, D.gA);
    const float u = fmaxf(__builtin_fmaf(r, D.gU, D.gV), 0.0f);
    const float g2 = -(u * u) * ri;
    return r <= D.kh ? g1 : g2;
}

// a.x b.x + a.y b.y + a.z b.z in one fixed rounding order,
// fma(az, bz, fma(ay, by, ax bx)).  Left to the compiler, the contraction of
// a three-term sum may fuse a different product first in different loop
// bodies; pair terms evaluated by two code paths (LDS window or gathers)
// must round identically.
__device__ __forceinline__ float dot3f(float ax, float ay, float az, float bx, float by, float bz) {
    return __builtin_fmaf(az, bz, __builtin_fmaf(ay, by, __fmul_rn(ax, bx)));
}

// Kahan-compensated FP32 accumulator.
struct Kahan {
    float s, c;
    __device__ __forceinline__ void init(float v) { s = v; c = 0.0f; }
    __device__ __forceinline__ void add(float t) {
#pragma clang fp contract(off)
        float y = t - c;
        float u = s + y;
        c = (u - s) - y;
        s = u;
    }
};

// Deterministic FP64 block sum over 256 threads (4 waves): fixed shuffle tree
// per wave, then waves in order.  Every thread of the block must call it.
__device__ __forceinline__ double block_sum_256(double v, double* lds) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane_id() == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = ((lds[0] + lds[1]) + lds[2]) + lds[3];
    __syncthreads();
    return r;
}

__device__ __forceinline__ float block_max_256(float v, float* lds) {
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    if (lane_id() == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = 0.0f;
    if (threadIdx.x == 0) r = fmaxf(fmaxf(lds[0], lds[1]), fmaxf(lds[2], lds[3]));
    __syncthreads();
    return r;
}

// scene of a key: largest s with keyBase[s] <= key (binary search over nscenes+1 bases)
__device__ __forceinline__ int scene_of_key(uint32_t key, const uint32_t* base, int ns) {
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= key) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Neighbour lists are padded to a multiple of 4 entries: fluid lists with the
// particle itself (distance 0: no gradient contribution; densities skip it by
// index), boundary lists with the dummy boundary particle D.nb placed far
// away (W = gradW = 0).
// 32-bit ELL neighbour lists stored in groups of four entries: entries 4g..4g+3 of
// particle i are the int4 at group g, (g * ncap + i).  A wave reads one group
// of its 64 particles with one 16-byte load per lane (1 KiB contiguous).
__device__ __forceinline__ size_t nb_slot(int k, int i, int ncap) {
    return ((size_t)(k >> 2) * ncap + i) * 4 + (k & 3);
}

// Packed lists.  The neighbours a particle can have lie in nine row runs
// (cells cx-1..cx+1 of the rows y+dy, z+dz).  A tile spans at most
// NB16_ROWS distinct own rows; slot = own-row ordinal * 9 + row, and an entry
// is (slot << NB16_OFF_BITS) + (j - base[slot]) with base[slot] the smallest
// run start over the tile's particles of that slot.  The table holds
// base[slot] - (slot << NB16_OFF_BITS), so j = table[e >> NB16_OFF_BITS] + e.
// Boundary padding uses slot NB16_DUMMY, whose base is the dummy particle.
constexpr int NB16_OFF_BITS = 11;
constexpr int NB16_ROWS = 3;
constexpr int NB16_DUMMY = 31;
constexpr int NB16_FLAG = 31;          // fluid table slot 31: 1 = tile packed
constexpr int NB16_TAB = 64;           // ints of table per tile

__device__ __forceinline__ size_t nb16_slot(int k, int i, int ncap) {
    return ((size_t)(k >> 3) * ncap + i) * 8 + (k & 7);
}

constexpr float DUMMY_FAR = 1.0e15f;   // coordinate of the dummy boundary particle

// Records of the LDS window of a packed tile at most (12 x 256): a larger
// window is not staged and the tile's loops gather from memory.
constexpr int WIN_MAX = 3072;
// records past the end of a cell run that the list builder may read (its
// candidate loads come four at a time); the position arrays are allocated
// that much longer
constexpr int X_TAIL = 3;

// ---------------------------------------------------------------------------
// host-side launchers implemented in the .hip files

// Events of the launch being timed: set by the host's timing scope; the
// first kernel launched through SPH_LAUNCH while they are set records them
// at its own dispatch boundaries (kernel time only) and clears them.
struct LaunchTiming {
    hipEvent_t start = nullptr, stop = nullptr;
};
extern thread_local LaunchTiming g_launch_timing;

// sort.hip
struct SortScratch {
    uint32_t* keys2;
    uint32_t* vals2;
    uint32_t* hist;       // 256 * maxTiles
    uint32_t* scanTmp;    // scan block sums
    int maxItems;
};
size_t sort_scratch_bytes(int maxItems);
void sort_scratch_bind(SortScratch& s, void* mem, int maxItems);
// Stable LSD radix sort of (keys, vals) on the lowest `bits` key bits.
// Result ends in keys/vals (ping-pong through the scratch buffers).
void radix_sort_pairs(uint32_t* keys, uint32_t* vals, int n, int bits, SortScratch& s,
                      hipStream_t st);
// exclusive scan of n uint32 in place; tmp needs scan_tmp_words(n) words.
int scan_tmp_words(int n);
void exclusive_scan_u32(uint32_t* data, int n, uint32_t* tmp, hipStream_t st);

// grid.hip
void launch_fluid_keys_into(const Dev& D, uint32_t* keys, uint32_t* vals, hipStream_t st);
void launch_scene_bounds(const Dev& D, const uint32_t* sortedKeys, int n, hipStream_t st);
void launch_scene_layout(const Dev& D, int* hostOut, hipStream_t st);
void launch_reorder(const Dev& D, const uint32_t* sortedKeys, const uint32_t* sortedVals, int n,
                    int oldPad, hipStream_t st);
void launch_cell_reset(const Dev& D, const uint32_t* keys, int n, int2* table, hipStream_t st);
void launch_cell_build(const Dev& D, const uint32_t* keys, int n, int2* table, hipStream_t st);
void launch_neighbours(const Dev& D, hipStream_t st);
// one of the two LDS staging depths of the list builder, for the autotuner
void launch_neighbours_variant(const Dev& D, int groups, hipStream_t st);
void launch_boundary_keys_from(const Dev& D, const float4* xs, const int* scene, uint32_t* keys,
                               uint32_t* vals, hipStream_t st);
void launch_boundary_reorder(const Dev& D, const float4* srcX, const int* srcSolid,
                             const int* srcScene, const uint32_t* sortedKeys,
                             const uint32_t* sortedVals, hipStream_t st);
void launch_boundary_volume(const Dev& D, hipStream_t st);
// rigid coupling: world positions/velocities from the poses; re-sort of the
// boundary after they moved (into the *2 arrays); V_b update; reaction forces
// on dynamic solids' particles; per-solid force and torque (double[6] each)
void launch_rigid_pose(const Dev& D, hipStream_t st);
void launch_boundary_init_local(const Dev& D, const float4* local, hipStream_t st);
void launch_boundary_regather(const Dev& D, const uint32_t* sk, const uint32_t* sv, hipStream_t st);
void launch_boundary_volume_update(const Dev& D, hipStream_t st);
void launch_rigid_reaction(const Dev& D, int mode, hipStream_t st);   // 0 WCSPH, 1 IISPH, 2 DFSPH
void launch_solid_reduce(const Dev& D, double* out, hipStream_t st);
void launch_neighbour_dump_ranked(const Dev& D, int cap, int* counts, int* kinds, long long* ids,
                                  float* dis, const int* rankOf, hipStream_t st);

// solvers.hip
void launch_cfl(const Dev& D, float cfl, float spacing, float dtMax, float dtMin, hipStream_t st);
void launch_wc_pressure(const Dev& D, hipStream_t st);
void launch_wc_force_integrate(const Dev& D, hipStream_t st);
void launch_df_density_alpha(const Dev& D, hipStream_t st);
void launch_df_aadv(const Dev& D, hipStream_t st);
void launch_df_pass_density(const Dev& D, hipStream_t st);
// density and alpha plus the warm start's inputs: Drho/Dt (v* = v) into f2
// and, with feat, the learned model's eight features per particle
void launch_df_density_alpha_warm(const Dev& D, float* feat, hipStream_t st);
void launch_df_warm_apply(const Dev& D, const float* y, const float* feat, int mode, hipStream_t st);
void launch_df_targets(const Dev& D, const float* feat, float* tgt, hipStream_t st);
void launch_mlp_gather(const Dev& D, const float* feat, const float* tgt, float* bfeat, float* btgt, int n,
                       uint32_t seed, hipStream_t st);
void launch_df_pass_divergence(const Dev& D, int projected, hipStream_t st);
void launch_df_pass_velocity(const Dev& D, hipStream_t st);
void launch_df_integrate(const Dev& D, hipStream_t st);
void launch_ii_density(const Dev& D, hipStream_t st);
void launch_ii_aadv_dii(const Dev& D, hipStream_t st);
// with feat: also the learned warm start's eight features per particle
void launch_ii_aii(const Dev& D, float* feat, hipStream_t st);
void launch_ii_warm_apply(const Dev& D, const float* y, const float* feat, hipStream_t st);
void launch_ii_targets(const Dev& D, const float* feat, float* tgt, hipStream_t st);
void launch_ii_sumdij(const Dev& D, hipStream_t st);
void launch_ii_pressure(const Dev& D, hipStream_t st);
void launch_ii_force_integrate(const Dev& D, int gate, hipStream_t st);
void launch_pci_rest_density(const Dev& D, hipStream_t st);
void launch_pci_aext(const Dev& D, hipStream_t st);
void launch_pci_density(const Dev& D, hipStream_t st);
void launch_pci_pressure(const Dev& D, hipStream_t st);
void launch_pci_integrate(const Dev& D, hipStream_t st);
void launch_scene_reduce(const Dev& D, int mode, float tol, int minIt, int cap, hipStream_t st);
// start of a solver loop: every scene active, iteration counts zero
void launch_loop_start(const Dev& D, hipStream_t st);
void launch_stats(const Dev& D, double* part, double* out, hipStream_t st);
void launch_fill_i32(int* p, int v, int n, hipStream_t st);
void launch_copy_probe(const float4* a, float4* b, size_t n, hipStream_t st);
// dst[i] = src[i] with a kernel (graph-capturable, no copy engine)
void launch_copy_f4(float4* dst, const float4* src, size_t n, hipStream_t st);

// iisph_df.hip: the IISph class's DF_* solvers
void launch_idf_prepare(const Dev& D, hipStream_t st);
void launch_idf_vadv(const Dev& D, int mode, hipStream_t st);
void launch_idf_kappa(const Dev& D, int mode, hipStream_t st);
void launch_idf_velocity(const Dev& D, hipStream_t st);

}  // namespace sphgpu
