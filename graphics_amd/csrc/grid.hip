// Uniform-grid neighbour search on the GPU (the reference's SphBase::gridSetup
// + neighbourSearch, SphBase.cpp:849-1089, re-designed): cell keys -> radix
// sort -> scene layout padded to whole tiles -> SoA reorder -> per-cell
// [start,end) tables -> ELL neighbour lists built from 9 contiguous x-runs of
// three cells each.  The pair test keeps the reference's FP32 rounding so
// that the neighbour sets are bit-identical.
#include "sph_common.h"

#include <type_traits>

namespace sphgpu {

namespace {

__device__ __forceinline__ bool slot_valid(const Dev& D, int i, int* scene) {
    int tile = i / TILE;
    *scene = D.tileScene[tile];
    return (i - tile * TILE) < D.tileValid[tile];
}


__global__ __launch_bounds__(256) void k_fluid_keys(Dev D, uint32_t* keys, uint32_t* vals) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= D.npad) return;
    uint32_t k = KEY_INVALID;
    if (D.id[i] >= 0) {
        float4 p = D.x[i];
        int s = D.tileScene[i / TILE];
        const SceneGrid G = D.grids[s];
        int cx, cy, cz;
        cell_of(G, p, cx, cy, cz);
        k = key_of(G, D.rowRank, cx, cy, cz);
    }
    keys[i] = k;
    vals[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_scene_bounds(Dev D, const uint32_t* __restrict__ sk, int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint32_t k = sk[i];
    if (k == KEY_INVALID) return;
    int s = scene_of_key(k, D.sceneKeyBase, D.nscenes);
    bool first = (i == 0) || sk[i - 1] == KEY_INVALID ||
                 scene_of_key(sk[i - 1], D.sceneKeyBase, D.nscenes) != s;
    bool last = (i == n - 1) || sk[i + 1] == KEY_INVALID ||
                scene_of_key(sk[i + 1], D.sceneKeyBase, D.nscenes) != s;
    if (first) D.sceneStart[s] = i;
    if (last) D.sceneCount[s] = i + 1;   // end for now; the layout kernel turns it into a count
}

// single thread: counts, padded starts, tile tables.  hostOut (mapped to the
// host by the caller via a device buffer) receives [npad_new, n_live].
__global__ void k_scene_layout(Dev D, int* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int pad = 0, live = 0;
    for (int s = 0; s < D.nscenes; ++s) {
        int st = D.sceneStart[s], en = D.sceneCount[s];
        int c = (en > st) ? en - st : 0;
        if (st < 0) c = 0;
        D.sceneCount[s] = c;
        D.scenePadStart[s] = pad;
        pad += (c + TILE - 1) / TILE * TILE;
        live += c;
    }
    D.scenePadStart[D.nscenes] = pad;
    out[0] = pad;
    out[1] = live;
}

__global__ __launch_bounds__(256) void k_tiles(Dev D, int numTiles) {
    int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= numTiles) return;
    int slot = t * TILE;
    int s = 0;
    // largest s with padStart[s] <= slot
    int lo = 0, hi = D.nscenes - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (D.scenePadStart[mid] <= slot) lo = mid; else hi = mid - 1;
    }
    s = lo;
    int valid = 0;
    if (slot < D.scenePadStart[D.nscenes]) {
        int local = slot - D.scenePadStart[s];
        valid = min(max(D.sceneCount[s] - local, 0), TILE);
    }
    D.tileScene[t] = s;
    D.tileValid[t] = valid;
}

// scatter sorted rank r to its padded slot; also marks padding slots
__global__ __launch_bounds__(256) void k_reorder(Dev D, const uint32_t* __restrict__ sk,
                                               const uint32_t* __restrict__ sv, int n) {
    int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    uint32_t k = sk[r];
    if (k == KEY_INVALID) return;
    int s = scene_of_key(k, D.sceneKeyBase, D.nscenes);
    int p = r - D.sceneStart[s] + D.scenePadStart[s];
    int src = (int)sv[r];
    D.x2[p] = D.x[src];
    D.v2[p] = D.v[src];
    D.id2[p] = D.id[src];
    D.ksum2[p] = D.ksum[src];
    D.kvsum2[p] = D.kvsum[src];
    D.key2[p] = k;
}

__global__ __launch_bounds__(256) void k_mark_padding(Dev D, int limit) {
    int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= limit) return;
    int tile = t / TILE;
    if ((t - tile * TILE) >= D.tileValid[tile]) {
        D.id2[t] = -1;
        D.key2[t] = KEY_INVALID;
    }
}

__global__ __launch_bounds__(256) void k_cell_reset(const uint32_t* __restrict__ keys, int n,
                                                  int2* __restrict__ table) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint32_t k = keys[i];
    if (k == KEY_INVALID) return;
    if (i == 0 || keys[i - 1] != k) table[k] = make_int2(0, 0);
}

__global__ __launch_bounds__(256) void k_cell_build(const uint32_t* __restrict__ keys, int n,
                                                  int2* __restrict__ table) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint32_t k = keys[i];
    if (k == KEY_INVALID) return;
    if (i == 0 || keys[i - 1] != k) table[k].x = i;
    if (i == n - 1 || keys[i + 1] != k) table[k].y = i + 1;
}


// Lists are built in LDS, one column per thread, and written out group by
// group (16 bytes per particle), so each store instruction of a wave writes
// whole cache lines.  A tile whose particles span at most NB16_ROWS own rows,
// and whose row runs each fit NB16_OFF_BITS of offset, gets packed 16-bit
// lists (8 entries per group); any other tile gets 32-bit lists (4 per
// group).  NB_STAGE_GROUPS groups are staged in LDS and the rest written to
// global memory directly (6 or 8 groups: LDS per block against blocks per
// CU, chosen by the autotuner).  Both forms hold the same entries in the
// same order.
template <int NB_STAGE_GROUPS>
__global__ __launch_bounds__(256) void k_neighbours(Dev D) {
    __shared__ int4 stage[NB_STAGE_GROUPS][256];
    __shared__ int sBase[2 * 32], sEnd[2 * 32];
    __shared__ int sRow[256];
    __shared__ int sWave[4];
    __shared__ int sOverflow;
    __shared__ int sWin[33];   // window offset of each fluid slot - (slot << NB16_OFF_BITS); [32] = size
    const int tid = threadIdx.x;
    const int tile = tile_of_block(D.tileMap);
    const int i = tile * 256 + tid;
    int scene = 0;
    const bool act = i < D.npad && slot_valid(D, i, &scene);
    int* st = reinterpret_cast<int*>(&stage[0][0]);
    uint16_t* st16 = reinterpret_cast<uint16_t*>(&stage[0][0]);
    const int stride = D.ncap;
    int cx = 0, cy = 0, cz = 0;
    float4 pa = make_float4(0.f, 0.f, 0.f, 0.f);
    SceneGrid G{};
    if (act) {
        G = D.grids[scene];
        pa = D.x[i];
        cell_of(G, pa, cx, cy, cz);
    }
    // ordinal of the particle's own row among the tile's rows: the valid
    // slots are a prefix of the tile in key order, so one row's particles
    // are contiguous
    const int row = act ? cz * G.ny + cy : -1;
    sRow[tid] = row;
    if (tid < 64) {
        sBase[tid] = INT_MAX;
        sEnd[tid] = 0;
    }
    if (tid == 0) sOverflow = 0;
    __syncthreads();
    const bool first = act && (tid == 0 || sRow[tid - 1] != row);
    const unsigned long long bal = __ballot(first);
    const int lane = tid & 63, wave = tid >> 6;
    if (lane == 0) sWave[wave] = __popcll(bal);
    __syncthreads();
    int ord = __popcll(bal & ((2ull << lane) - 1ull)) - 1;
    for (int w = 0; w < wave; ++w) ord += sWave[w];
    bool packed = sWave[0] + sWave[1] + sWave[2] + sWave[3] <= NB16_ROWS;
    // the nine fluid and boundary row runs (rows y+dy, z+dz, r = 3 (dz+1) +
    // (dy+1)), looked up once and kept in registers; empty: (0, 0)
    int2 runF[9], runB[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        runF[r] = make_int2(0, 0);
        runB[r] = make_int2(0, 0);
        const int z = cz + r / 3 - 1, y = cy + r % 3 - 1;
        if (act && z >= 0 && z < G.nz && y >= 0 && y < G.ny) {
            const uint32_t k0 = key_of(G, D.rowRank, cx, y, z);
            const int2 rf = run3(D.cellF, k0, cx, G.nx);
            const int2 rb = run3(D.cellB, k0, cx, G.nx);
            if (rf.y > rf.x) runF[r] = rf;
            if (rb.y > rb.x) runB[r] = rb;
        }
    }
    // per slot: the smallest run start and the largest run end of the tile
    if (packed && act) {
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            const int slot = ord * 9 + r;
            if (runF[r].y > runF[r].x) {
                atomicMin(&sBase[slot], runF[r].x);
                atomicMax(&sEnd[slot], runF[r].y);
            }
            if (runB[r].y > runB[r].x) {
                atomicMin(&sBase[32 + slot], runB[r].x);
                atomicMax(&sEnd[32 + slot], runB[r].y);
            }
        }
    }
    __syncthreads();
    if (tid < 64 && sBase[tid] != INT_MAX && sEnd[tid] - sBase[tid] > (1 << NB16_OFF_BITS)) sOverflow = 1;
    __syncthreads();
    packed = packed && sOverflow == 0;
    // the tile's window: its fluid slot runs concatenated in slot order (the
    // order in which the solver passes stage it); entry e of slot s is
    // record e + sWin[s] of it
    if (tid < 64) {
        const int len = (packed && tid < 32 && tid != NB16_FLAG && sBase[tid] != INT_MAX) ? sEnd[tid] - sBase[tid] : 0;
        int incl = len;
        for (int off = 1; off < 32; off <<= 1) {
            const int t = __shfl_up(incl, off, 64);
            if (tid >= off) incl += t;
        }
        if (tid < 32) sWin[tid] = (incl - len) - (tid << NB16_OFF_BITS);
        if (tid == 31) sWin[32] = incl;
    }
    __syncthreads();
    const bool win = packed && sWin[32] <= WIN_MAX;
    auto to_win = [&](unsigned e) { return (e + (unsigned)sWin[e >> NB16_OFF_BITS]) & 0xffffu; };

    // staged entry k of this thread: element k * 256 + tid of the stage
    // (entries of one thread 256 elements apart, so one store instruction of
    // a wave writes consecutive elements and its address is one operation)
    auto put = [&](int* list, int k, int j) {
        if (k < NB_STAGE_GROUPS * 4) st[(k << 8) + tid] = j;
        else list[nb_slot(k, i, stride)] = j;
    };
    // bias(slot) = (slot << NB16_OFF_BITS) - base[slot]: an entry is j + bias
    auto bias16 = [&](int slot) { return packed ? (slot << NB16_OFF_BITS) - sBase[slot] : 0; };
    auto put16 = [&](int4* list, int k, int bias, int j) {
        const uint16_t e = uint16_t(j + bias);
        if (k < NB_STAGE_GROUPS * 8) {
            st16[(k << 8) + tid] = e;
        } else {
            reinterpret_cast<uint16_t*>(list)[nb16_slot(k, i, stride)] = e;
            if (win && list == D.nbF16) reinterpret_cast<uint16_t*>(D.nbW16)[nb16_slot(k, i, stride)] = uint16_t(to_win(e));
        }
    };
    auto flush = [&](int* list, int count) {
        __syncthreads();
        const int ng = min((count + 3) >> 2, NB_STAGE_GROUPS);
        int4* out = reinterpret_cast<int4*>(list);
        for (int g = 0; g < ng; ++g) {
            const int b = (4 * g) << 8;
            out[(size_t)g * stride + i] = make_int4(st[b + tid], st[b + 256 + tid], st[b + 512 + tid], st[b + 768 + tid]);
        }
        __syncthreads();
    };
    auto flush16 = [&](int4* out, int count) {
        __syncthreads();
        const int ng = min((count + 7) >> 3, NB_STAGE_GROUPS);
        const bool towin = win && out == D.nbF16;
        for (int g = 0; g < ng; ++g) {
            unsigned w[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int b = (8 * g + 2 * q) << 8;
                w[q] = (unsigned)st16[b + tid] | (unsigned)st16[b + 256 + tid] << 16;
            }
            out[(size_t)g * stride + i] = make_int4((int)w[0], (int)w[1], (int)w[2], (int)w[3]);
            if (towin) {
                // the same group as window indices (entries past the padded
                // count are converted too, and never read)
                int r[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) r[q] = (int)(to_win(w[q] & 0xffffu) | to_win(w[q] >> 16) << 16);
                D.nbW16[(size_t)g * stride + i] = make_int4(r[0], r[1], r[2], r[3]);
            }
        }
        __syncthreads();
    };
    int nbn = 0, over = 0, zero = 0;
    // Candidate loop of one kind (fluid or boundary), candidates four at a
    // time (four loads in flight, tested in order).  A candidate's entry is
    // stored in LDS at the next free entry before the test decides (a miss
    // is overwritten by the next hit, or lies past the padded count that
    // the passes read), so the common path has no branch; entries past the
    // LDS stage go to global memory and entries past the list width are
    // counted (both rare, in a branch of their own).  n: entries found.
    auto candidates = [&](auto packed_c, const int2* runs, const float4* __restrict__ xs, int selfj, int cap,
                          int slot0, int4* list16, int* list32, int& n) {
        constexpr bool PK = decltype(packed_c)::value;
        constexpr int STAGE = PK ? NB_STAGE_GROUPS * 8 : NB_STAGE_GROUPS * 4;
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            const int2 rr = runs[r];
            const int bias = PK && rr.y > rr.x ? bias16(slot0 + ord * 9 + r) : 0;
            for (int j0 = rr.x; j0 < rr.y; j0 += 4) {
                // (up to X_TAIL records past the run are read, and not used:
                // the position arrays are that much longer)
                const float4* __restrict__ xg = xs + j0;
                float4 q[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) q[k] = xg[k];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int j = j0 + k;
                    const float l2 = dis_square_exact(pa, q[k]);
                    const bool hit = j < rr.y && j != selfj && l2 < D.h2;
                    if (n < STAGE) {
                        if (PK) st16[(n << 8) + tid] = uint16_t(j + bias);
                        else st[(n << 8) + tid] = j;
                    }
                    if (hit && n >= STAGE) {
                        if (n < cap) {
                            if (PK) put16(list16, n, bias, j);
                            else put(list32, n, j);
                        } else {
                            ++over;
                        }
                    }
                    zero += (hit && l2 == 0.0f);
                    n += hit;
                }
            }
        }
    };
    using PackedT = std::integral_constant<bool, true>;
    using PlainT = std::integral_constant<bool, false>;
    // fluid neighbours (SphBase.cpp:867-905); the particle itself is skipped
    int nf = 0;
    if (packed) candidates(PackedT{}, runF, D.x, i, NBF_CAP, 0, D.nbF16, nullptr, nf);
    else candidates(PlainT{}, runF, D.x, i, NBF_CAP, 0, nullptr, D.nbF, nf);
    const int stored = min(nf, NBF_CAP);
    const int cf = act ? stored : 0;
    // pad: the particle itself
    for (int k = cf; k < ((cf + 3) & ~3); ++k) {
        if (packed) put16(D.nbF16, k, bias16(ord * 9 + 4), i);
        else put(D.nbF, k, i);
    }
    if (packed) flush16(D.nbF16, cf);
    else flush(D.nbF, cf);
    // boundary neighbours (SphBase.cpp:908-930); no index is skipped
    if (act) {
        if (packed) candidates(PackedT{}, runB, D.xb, -1, NBB_CAP, 32, D.nbB16, nullptr, nbn);
        else candidates(PlainT{}, runB, D.xb, -1, NBB_CAP, 32, nullptr, D.nbB, nbn);
    }
    const int cb = act ? min(nbn, NBB_CAP) : 0;
    // pad: the far dummy particle
    for (int k = cb; k < ((cb + 3) & ~3); ++k) {
        if (packed) {
            const uint16_t e = uint16_t(NB16_DUMMY << NB16_OFF_BITS);
            if (k < NB_STAGE_GROUPS * 8) st16[(k << 8) + tid] = e;
            else reinterpret_cast<uint16_t*>(D.nbB16)[nb16_slot(k, i, stride)] = e;
        }
        else put(D.nbB, k, D.nb);
    }
    if (packed) flush16(D.nbB16, cb);
    else flush(D.nbB, cb);
    if (tid < NB16_TAB) {
        const int slot = tid & 31;
        int v;
        if (tid == NB16_FLAG) v = packed ? 1 : 0;
        else if (tid == 32 + NB16_DUMMY) v = D.nb - (NB16_DUMMY << NB16_OFF_BITS);
        else v = sBase[tid] != INT_MAX ? sBase[tid] - (slot << NB16_OFF_BITS) : 0;
        D.nbTab[(size_t)tile * NB16_TAB + tid] = v;
        if (tid < 32) D.nbWin[(size_t)tile * 32 + tid] = packed ? sWin[tid] + (tid << NB16_OFF_BITS) : 0;
    }
    // stored fluid entries | fluid neighbours found << 8 | boundary << 16 | phase << 24
    if (i < D.npad) D.cnt[i] = cf | (act ? min(nf, 255) : 0) << 8 | (cb << 16) | ((act ? int(pa.w) : 0) << 24);
    if (!act) return;
    int tm = nf + nbn - REF_NEIGHBOUR_CAP;
    if (tm > 0) atomicAdd(&D.counters[C_TOO_MANY], tm);
    if (over > 0) atomicAdd(&D.counters[C_LIST_OVF], over);
    if (zero > 0) atomicAdd(&D.counters[C_ZERO_DIS], zero);
}

// ---- boundary --------------------------------------------------------------

__global__ __launch_bounds__(256) void k_boundary_keys(Dev D, const float4* __restrict__ xs,
                                                     const int* __restrict__ scene, uint32_t* keys,
                                                     uint32_t* vals) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= D.nb) return;
    const SceneGrid G = D.grids[scene[i]];
    int cx, cy, cz;
    cell_of(G, xs[i], cx, cy, cz);
    keys[i] = key_of(G, D.rowRank, cx, cy, cz);
    vals[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_boundary_reorder(Dev D, const float4* __restrict__ xs,
                                                        const int* __restrict__ solid,
                                                        const int* __restrict__ scene,
                                                        const uint32_t* __restrict__ sk,
                                                        const uint32_t* __restrict__ sv) {
    int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= D.nb) return;
    int src = (int)sv[r];
    D.xb[r] = xs[src];
    D.vb[r] = make_float4(0.f, 0.f, 0.f, 0.f);
    D.bid[r] = src;
    D.bsolid[r] = solid[src];
    D.bscene[r] = scene[src];
    D.bkey[r] = sk[r];
}

// SphBase::prepareSolidPartWeight (SphBase.cpp:766-813): V_b = 1/(W(0) + sum over
// boundary neighbours W(r)); kerSumFromSelf = W(0) + same-solid part.
__global__ __launch_bounds__(256) void k_boundary_volume(Dev D) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= D.nb) return;
    const SceneGrid G = D.grids[D.bscene[i]];
    float4 pa = D.xb[i];
    int sa = D.bsolid[i];
    int cx, cy, cz;
    cell_of(G, pa, cx, cy, cz);
    float sumAll = D.W0, sumSelf = D.W0;
    for (int dz = -1; dz <= 1; ++dz) {
        int z = cz + dz;
        if (z < 0 || z >= G.nz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            int y = cy + dy;
            if (y < 0 || y >= G.ny) continue;
            uint32_t k0 = key_of(G, D.rowRank, cx, y, z);
            int2 rb = run3(D.cellB, k0, cx, G.nx);
            for (int j = rb.x; j < rb.y; ++j) {
                if (j == i) continue;
                float4 pb = D.xb[j];
                float l2 = dis_square_exact(pa, pb);
                if (l2 < D.h2) {
                    float w = kernel_W(sqrtf(l2), D);
                    sumAll += w;
                    if (D.bsolid[j] == sa) sumSelf += w;
                }
            }
        }
    }
    pa.w = 1.0f / sumAll;
    D.xb[i] = pa;
    D.bkself[i] = sumSelf;
}

// ---- dynamic solids -----------------------------------------------------------
// SphSolid::updatePaticles (SphBase.cpp:95-128): x = T p0 (projective divide),
// v = v_lin + w x (x - com), for the particles of dynamic solids
__global__ __launch_bounds__(256) void k_rigid_pose(Dev D) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= D.nb) return;
    const SolidPose& S = D.solids[D.bsolid[i]];
    if (!S.dynamic) return;
    const float4 p = D.bloc[i];
    const float* T = S.T;
    float x = T[0] * p.x + T[4] * p.y + T[8] * p.z + T[12];
    float y = T[1] * p.x + T[5] * p.y + T[9] * p.z + T[13];
    float z = T[2] * p.x + T[6] * p.y + T[10] * p.z + T[14];
    const float a = T[3] * p.x + T[7] * p.y + T[11] * p.z + T[15];
    x /= a; y /= a; z /= a;
    const float rx = x - S.com[0], ry = y - S.com[1], rz = z - S.com[2];
    D.xb[i] = make_float4(x, y, z, D.xb[i].w);
    D.vb[i] = make_float4(S.v[0] + (S.w[1] * rz - S.w[2] * ry), S.v[1] + (S.w[2] * rx - S.w[0] * rz),
                          S.v[2] + (S.w[0] * ry - S.w[1] * rx), 0.0f);
}

// body-frame samples (original order) into sorted order; slot of each original index
__global__ __launch_bounds__(256) void k_boundary_init_local(Dev D, const float4* __restrict__ local) {
    int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= D.nb) return;
    const int k = D.bid[r];
    D.bloc[r] = local[k];
    D.brank[k] = r;
}

// boundary arrays gathered into cell-key order after the solids moved
__global__ __launch_bounds__(256) void k_boundary_regather(Dev D, const uint32_t* __restrict__ sk,
                                                         const uint32_t* __restrict__ sv) {
    int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= D.nb) return;
    const int src = (int)sv[r];
    D.xb2[r] = D.xb[src];
    D.vb2[r] = D.vb[src];
    D.bloc2[r] = D.bloc[src];
    D.bid2[r] = D.bid[src];
    D.bsolid2[r] = D.bsolid[src];
    D.bscene2[r] = D.bscene[src];
    D.bkself2[r] = D.bkself[src];
    D.bkey2[r] = sk[r];
    D.brank[D.bid[src]] = r;
}

// SphBase::updateSolidPartWeight (SphBase.cpp:817-841): V_b = 1 /
// (kerSumFromSelf + sum of W over boundary neighbours of other solids; a
// rigid body's own particles keep the part fixed at the start, kerSumFromSelf)
__global__ __launch_bounds__(256) void k_boundary_volume_update(Dev D) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= D.nb) return;
    const SceneGrid G = D.grids[D.bscene[i]];
    float4 pa = D.xb[i];
    const int sa = D.bsolid[i];
    int cx, cy, cz;
    cell_of(G, pa, cx, cy, cz);
    float vol = D.bkself[i];
    for (int dz = -1; dz <= 1; ++dz) {
        int z = cz + dz;
        if (z < 0 || z >= G.nz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            int y = cy + dy;
            if (y < 0 || y >= G.ny) continue;
            int2 rb = run3(D.cellB, key_of(G, D.rowRank, cx, y, z), cx, G.nx);
            for (int j = rb.x; j < rb.y; ++j) {
                if (j == i || D.bsolid[j] == sa) continue;
                float l2 = dis_square_exact(pa, D.xb[j]);
                if (l2 < D.h2) vol += kernel_W(sqrtf(l2), D);
            }
        }
    }
    pa.w = 1.0f / vol;
    D.xb[i] = pa;
}

// Reaction force on the particles of dynamic solids from their fluid
// neighbours, summed from the boundary particle's side (a fixed order).
// mode 0 WCSPH, WcSph::wc_boundPartForce_f (WcSph.cpp:314-362): pressure x2
// (its fluid-side boundary force, WcSph.cpp:285) and artificial viscosity;
// fluid pressure f1, density v.w.
// mode 1 IISPH, IISph_M::ii_boundPartForcePressure_f (IISph_M.cpp:586-607):
// pressure x1 (the fluid-side boundary force of this implementation) and
// viscosity; pressure kr, density rho.
// mode 2 DFSPH, the constant-density solve's boundary term accumulated over
// its iterations (DFSPH.cpp:674-690, scaled by the fluid mass at 723-737):
// F_b = sum_j m_j rho0_j V_b (sum_it kappa_j / rho_j) grad_j W_jb, with
// sum_it kappa_j dt^2 = ksum_j; no viscosity.
__global__ __launch_bounds__(256) void k_rigid_reaction(Dev D, int mode) {
    int a = blockIdx.x * 256 + threadIdx.x;
    if (a >= D.nb) return;
    const SolidPose& S = D.solids[D.bsolid[a]];
    if (!S.dynamic) return;
    const SceneGrid G = D.grids[D.bscene[a]];
    const float4 pa = D.xb[a], va = D.vb[a];
    const bool iisph = mode == 1, dfsph = mode == 2;
    const float pf = iisph ? 1.0f : 2.0f;
    const float nuc = 2.0f * S.visc * D.Hs * D.cs;
    const float dt = D.dt[D.bscene[a]];
    const float rdt2 = 1.0f / (dt * dt);
    int cx, cy, cz;
    cell_of(G, pa, cx, cy, cz);
    float fx = 0.f, fy = 0.f, fz = 0.f;
    for (int dz = -1; dz <= 1; ++dz) {
        int z = cz + dz;
        if (z < 0 || z >= G.nz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            int y = cy + dy;
            if (y < 0 || y >= G.ny) continue;
            int2 rf = run3(D.cellF, key_of(G, D.rowRank, cx, y, z), cx, G.nx);
            for (int j = rf.x; j < rf.y; ++j) {
                const float4 pb = D.x[j];
                const float l2 = dis_square_exact(pa, pb);
                if (!(l2 < D.h2) || l2 == 0.0f) continue;
                const int ph = (int)pb.w;
                const float frho0 = D.rho0[ph], fm0 = D.mass[ph];
                const float dx = pa.x - pb.x, dy2 = pa.y - pb.y, dz2 = pa.z - pb.z;
                const float g = kernel_gradfac(sqrtf(l2), D);
                if (dfsph) {
                    const float k = D.ksum[j] * rdt2 / D.rho[j];
                    const float force = -g * (fm0 * frho0 * pa.w * k);
                    fx += dx * force; fy += dy2 * force; fz += dz2 * force;
                    continue;
                }
                const float4 vb = D.v[j];
                const float rhob = iisph ? D.rho[j] : vb.w;
                const float pb_p = iisph ? D.kr[j] : D.f1[j];
                const float cvol = -fm0 * frho0 * pa.w;
                float force = pb_p > 0.0f ? g * (cvol * (pb_p / (rhob * rhob) * pf)) : 0.0f;
                const float pro = (va.x - vb.x) * dx + (va.y - vb.y) * dy2 + (va.z - vb.z) * dz2;
                if (pro < 0.0f) {
                    const float nu = nuc / (rhob * 2.0f);
                    const float pi = -nu * pro / (l2 + D.visc_eps);
                    force += g * (cvol * pi);
                }
                fx += dx * force; fy += dy2 * force; fz += dz2 * force;
            }
        }
    }
    D.bforce[a] = make_float4(fx, fy, fz, 0.0f);
}

// SphSolid::rigidForceJoin (SphBase.cpp:131-143) per dynamic solid: one block
// per solid over its particles in original order, in double, fixed tree
__global__ __launch_bounds__(256) void k_solid_reduce(Dev D, double* __restrict__ out) {
    __shared__ double lds[6][4];
    const int s = blockIdx.x;
    const SolidPose& S = D.solids[s];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (S.dynamic) {
        for (int k = S.firstId + (int)threadIdx.x; k < S.endId; k += 256) {
            const int r = D.brank[k];
            const float4 f = D.bforce[r], x = D.xb[r];
            const double rx = (double)x.x - S.com[0], ry = (double)x.y - S.com[1], rz = (double)x.z - S.com[2];
            acc[0] += f.x; acc[1] += f.y; acc[2] += f.z;
            acc[3] += ry * f.z - rz * f.y;
            acc[4] += rz * f.x - rx * f.z;
            acc[5] += rx * f.y - ry * f.x;
        }
    }
    for (int q = 0; q < 6; ++q) {
        double v = acc[q];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if ((threadIdx.x & 63) == 0) lds[q][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int q = threadIdx.x;
        out[6 * s + q] = ((lds[q][0] + lds[q][1]) + lds[q][2]) + lds[q][3];
    }
}

// test dump: the stored neighbour lists (the ones every pass iterates), with
// original ids and correctly rounded distances
__global__ __launch_bounds__(256) void k_neighbour_dump(Dev D, int cap, int* counts, int* kinds,
                                                      long long* ids, float* dis, const int* rankOf) {
    int i = blockIdx.x * 256 + threadIdx.x;
    int scene;
    if (i >= D.npad || !slot_valid(D, i, &scene)) return;
    int o = rankOf[i];
    if (o < 0) return;
    const float4 pa = D.x[i];
    const int cnt = D.cnt[i], nf = cnt & 0xFF, nbn = (cnt >> 16) & 0xFF;
    const int* tab = D.nbTab + (size_t)(i / TILE) * NB16_TAB;
    const bool packed = tab[NB16_FLAG] != 0;
    const uint16_t* f16 = reinterpret_cast<const uint16_t*>(D.nbF16);
    const uint16_t* b16 = reinterpret_cast<const uint16_t*>(D.nbB16);
    int c = 0;
    for (int k = 0; k < nf; ++k) {
        int j;
        if (packed) {
            const int e = f16[nb16_slot(k, i, D.ncap)];
            j = tab[e >> NB16_OFF_BITS] + e;
        } else {
            j = D.nbF[nb_slot(k, i, D.ncap)];
        }
        if (c < cap) {
            size_t q = (size_t)o * cap + c;
            kinds[q] = 0; ids[q] = D.id[j]; dis[q] = sqrtf(dis_square_exact(pa, D.x[j]));
        }
        ++c;
    }
    for (int k = 0; k < nbn; ++k) {
        int j;
        if (packed) {
            const int e = b16[nb16_slot(k, i, D.ncap)];
            j = tab[32 + (e >> NB16_OFF_BITS)] + e;
        } else {
            j = D.nbB[nb_slot(k, i, D.ncap)];
        }
        if (c < cap) {
            size_t q = (size_t)o * cap + c;
            kinds[q] = 1; ids[q] = D.bid[j]; dis[q] = sqrtf(dis_square_exact(pa, D.xb[j]));
        }
        ++c;
    }
    counts[o] = c;
}

inline int blocks_for(int n) { return (n + 255) / 256; }

}  // namespace

void launch_fluid_keys_into(const Dev& D, uint32_t* keys, uint32_t* vals, hipStream_t st) {
    hipLaunchKernelGGL(k_fluid_keys, dim3(blocks_for(D.npad)), dim3(256), 0, st, D, keys, vals);
}

void launch_scene_bounds(const Dev& D, const uint32_t* sk, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_scene_bounds, dim3(blocks_for(n)), dim3(256), 0, st, D, sk, n);
}

void launch_scene_layout(const Dev& D, int* out, hipStream_t st) {
    hipLaunchKernelGGL(k_scene_layout, dim3(1), dim3(64), 0, st, D, out);
    int numTiles = D.ncap / TILE;
    hipLaunchKernelGGL(k_tiles, dim3(blocks_for(numTiles)), dim3(256), 0, st, D, numTiles);
}

void launch_reorder(const Dev& D, const uint32_t* sk, const uint32_t* sv, int n, int oldPad,
                    hipStream_t st) {
    int limit = oldPad > 0 ? oldPad : n;
    hipLaunchKernelGGL(k_mark_padding, dim3(blocks_for(limit)), dim3(256), 0, st, D, limit);
    hipLaunchKernelGGL(k_reorder, dim3(blocks_for(n)), dim3(256), 0, st, D, sk, sv, n);
}

void launch_cell_reset(const Dev& D, const uint32_t* keys, int n, int2* table, hipStream_t st) {
    (void)D;
    if (n > 0) hipLaunchKernelGGL(k_cell_reset, dim3(blocks_for(n)), dim3(256), 0, st, keys, n, table);
}

void launch_cell_build(const Dev& D, const uint32_t* keys, int n, int2* table, hipStream_t st) {
    (void)D;
    if (n > 0) hipLaunchKernelGGL(k_cell_build, dim3(blocks_for(n)), dim3(256), 0, st, keys, n, table);
}

void launch_neighbours_variant(const Dev& D, int groups, hipStream_t st) {
    const dim3 g(blocks_for(D.npad)), b(256);
    if (groups >= 8) hipLaunchKernelGGL(k_neighbours<8>, g, b, 0, st, D);
    else hipLaunchKernelGGL(k_neighbours<6>, g, b, 0, st, D);
}

void launch_neighbours(const Dev& D, hipStream_t st) { launch_neighbours_variant(D, D.nbStage, st); }

void launch_boundary_keys_from(const Dev& D, const float4* xs, const int* scene, uint32_t* keys,
                               uint32_t* vals, hipStream_t st) {
    hipLaunchKernelGGL(k_boundary_keys, dim3(blocks_for(D.nb)), dim3(256), 0, st, D, xs, scene, keys, vals);
}

void launch_boundary_reorder(const Dev& D, const float4* srcX, const int* srcSolid, const int* srcScene,
                             const uint32_t* sk, const uint32_t* sv, hipStream_t st) {
    hipLaunchKernelGGL(k_boundary_reorder, dim3(blocks_for(D.nb)), dim3(256), 0, st, D, srcX, srcSolid,
                       srcScene, sk, sv);
}

void launch_boundary_volume(const Dev& D, hipStream_t st) {
    hipLaunchKernelGGL(k_boundary_volume, dim3(blocks_for(D.nb)), dim3(256), 0, st, D);
}

void launch_rigid_reaction(const Dev& D, int mode, hipStream_t st) {
    if (D.nb > 0) hipLaunchKernelGGL(k_rigid_reaction, dim3(blocks_for(D.nb)), dim3(256), 0, st, D, mode);
}

void launch_solid_reduce(const Dev& D, double* out, hipStream_t st) {
    if (D.nsolids > 0) hipLaunchKernelGGL(k_solid_reduce, dim3(D.nsolids), dim3(256), 0, st, D, out);
}

void launch_boundary_init_local(const Dev& D, const float4* local, hipStream_t st) {
    if (D.nb > 0) hipLaunchKernelGGL(k_boundary_init_local, dim3(blocks_for(D.nb)), dim3(256), 0, st, D, local);
}

void launch_rigid_pose(const Dev& D, hipStream_t st) {
    if (D.nb > 0) hipLaunchKernelGGL(k_rigid_pose, dim3(blocks_for(D.nb)), dim3(256), 0, st, D);
}

void launch_boundary_regather(const Dev& D, const uint32_t* sk, const uint32_t* sv, hipStream_t st) {
    if (D.nb > 0) hipLaunchKernelGGL(k_boundary_regather, dim3(blocks_for(D.nb)), dim3(256), 0, st, D, sk, sv);
}

void launch_boundary_volume_update(const Dev& D, hipStream_t st) {
    if (D.nb > 0) hipLaunchKernelGGL(k_boundary_volume_update, dim3(blocks_for(D.nb)), dim3(256), 0, st, D);
}

void launch_neighbour_dump_ranked(const Dev& D, int cap, int* counts, int* kinds, long long* ids,
                                  float* dis, const int* rankOf, hipStream_t st) {
    hipLaunchKernelGGL(k_neighbour_dump, dim3(blocks_for(D.npad)), dim3(256), 0, st, D, cap, counts,
                       kinds, ids, dis, rankOf);
}

}  // namespace sphgpu
