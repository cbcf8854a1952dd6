// Stable LSD radix sort of (uint32 key, uint32 value) pairs and a device-wide
// exclusive scan, written for wave64.
//
// One sort pass handles up to 8 key bits.  A tile is 2048 items: 4 waves x
// 8 sub-chunks x 64 lanes, item (w, k, lane) at tile offset w*512 + k*64 + lane.
// Ranking inside a sub-chunk uses ballots on the digit bits (lanes with the
// same digit form a match mask; rank = popcount of lower matching lanes), and
// a per-wave LDS counter per digit carries the count across sub-chunks, so
// the order inside a tile is the input order (stable).  Tiles are combined
// through a digit-major histogram that is exclusive-scanned.
#include "sph_common.h"

namespace sphgpu {

namespace {

constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 8;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;   // 2048
constexpr int SCAN_TILE = 2048;

// rank one 64-item sub-chunk: returns the match mask of this lane's digit
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid, int nbits) {
    uint64_t m = __ballot(valid);
    for (int b = 0; b < nbits; ++b) {
        bool bit = (d >> b) & 1u;
        uint64_t bal = __ballot(valid && bit);
        m &= bit ? bal : ~bal;
    }
    return m;
}

__global__ __launch_bounds__(SORT_THREADS) void k_rs_hist(const uint32_t* __restrict__ keys, int n,
                                                        int shift, int nbits, uint32_t* __restrict__ hist,
                                                        int numTiles) {
    __shared__ uint32_t cnt[4][256];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    for (int i = tid; i < 4 * 256; i += SORT_THREADS) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * SORT_TILE + w * 512;
    const uint32_t mask = (1u << nbits) - 1u;
    for (int k = 0; k < SORT_ITEMS; ++k) {
        long long idx = base + k * 64 + lane;
        bool valid = idx < n;
        uint32_t d = valid ? (keys[idx] >> shift) & mask : 0u;
        uint64_t m = match_digit(d, valid, nbits);
        bool leader = valid && (__ffsll((long long)m) - 1) == lane;
        if (leader) cnt[w][d] += (uint32_t)__popcll(m);
    }
    __syncthreads();
    for (int d = tid; d < 256; d += SORT_THREADS) {
        uint32_t t = cnt[0][d] + cnt[1][d] + cnt[2][d] + cnt[3][d];
        hist[(size_t)d * numTiles + blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(SORT_THREADS) void k_rs_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, int n, int shift, int nbits, const uint32_t* __restrict__ offs,
    int numTiles) {
    __shared__ uint32_t cnt[4][256];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    for (int i = tid; i < 4 * 256; i += SORT_THREADS) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * SORT_TILE + w * 512;
    const uint32_t mask = (1u << nbits) - 1u;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t key[SORT_ITEMS], val[SORT_ITEMS], rank[SORT_ITEMS];
    for (int k = 0; k < SORT_ITEMS; ++k) {
        long long idx = base + k * 64 + lane;
        bool valid = idx < n;
        key[k] = valid ? kin[idx] : 0u;
        val[k] = valid ? vin[idx] : 0u;
        uint32_t d = (key[k] >> shift) & mask;
        uint64_t m = match_digit(d, valid, nbits);
        uint32_t before = valid ? cnt[w][d] : 0u;
        rank[k] = before + (uint32_t)__popcll(m & lt);
        bool leader = valid && (__ffsll((long long)m) - 1) == lane;
        if (leader) cnt[w][d] = before + (uint32_t)__popcll(m);
    }
    __syncthreads();
    for (int d = tid; d < 256; d += SORT_THREADS) {
        uint32_t run = (d <= (int)mask) ? offs[(size_t)d * numTiles + blockIdx.x] : 0u;
        for (int ww = 0; ww < 4; ++ww) {
            uint32_t c = cnt[ww][d];
            cnt[ww][d] = run;
            run += c;
        }
    }
    __syncthreads();
    for (int k = 0; k < SORT_ITEMS; ++k) {
        long long idx = base + k * 64 + lane;
        if (idx < n) {
            uint32_t d = (key[k] >> shift) & mask;
            uint32_t pos = cnt[w][d] + rank[k];
            kout[pos] = key[k];
            vout[pos] = val[k];
        }
    }
}

// ---- exclusive scan -------------------------------------------------------

__device__ __forceinline__ uint32_t block_exclusive_scan_256(uint32_t v, uint32_t* lds, uint32_t* total) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) lds[w] = x;
    __syncthreads();
    uint32_t wpre = 0;
    for (int i = 0; i < w; ++i) wpre += lds[i];
    uint32_t tot = lds[0] + lds[1] + lds[2] + lds[3];
    __syncthreads();
    if (total) *total = tot;
    return wpre + x - v;
}

__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* __restrict__ d, int n,
                                                   uint32_t* __restrict__ sums) {
    __shared__ uint32_t lds[4];
    long long base = (long long)blockIdx.x * SCAN_TILE + threadIdx.x * 8;
    uint32_t s = 0;
    for (int k = 0; k < 8; ++k)
        if (base + k < n) s += d[base + k];
    uint32_t tot;
    block_exclusive_scan_256(s, lds, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_scan_down(uint32_t* __restrict__ d, int n,
                                                 const uint32_t* __restrict__ sums) {
    __shared__ uint32_t lds[4];
    long long base = (long long)blockIdx.x * SCAN_TILE + threadIdx.x * 8;
    uint32_t v[8], s = 0;
    for (int k = 0; k < 8; ++k) {
        v[k] = (base + k < n) ? d[base + k] : 0u;
        s += v[k];
    }
    uint32_t pre = block_exclusive_scan_256(s, lds, nullptr) + (sums ? sums[blockIdx.x] : 0u);
    for (int k = 0; k < 8; ++k) {
        if (base + k < n) d[base + k] = pre;
        pre += v[k];
    }
}

}  // namespace

int scan_tmp_words(int n) {
    int total = 0;
    while (n > SCAN_TILE) {
        n = (n + SCAN_TILE - 1) / SCAN_TILE;
        total += n;
    }
    return total + 1;
}

void exclusive_scan_u32(uint32_t* data, int n, uint32_t* tmp, hipStream_t st) {
    if (n <= 0) return;
    int blocks = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (blocks == 1) {
        hipLaunchKernelGGL(k_scan_down, dim3(1), dim3(256), 0, st, data, n, (const uint32_t*)nullptr);
        return;
    }
    hipLaunchKernelGGL(k_scan_reduce, dim3(blocks), dim3(256), 0, st, (const uint32_t*)data, n, tmp);
    exclusive_scan_u32(tmp, blocks, tmp + blocks, st);
    hipLaunchKernelGGL(k_scan_down, dim3(blocks), dim3(256), 0, st, data, n, (const uint32_t*)tmp);
}

__global__ __launch_bounds__(256) void k_copy_pairs(const uint32_t* __restrict__ ka, const uint32_t* __restrict__ va,
                                                  uint32_t* __restrict__ kb, uint32_t* __restrict__ vb, int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        kb[i] = ka[i];
        vb[i] = va[i];
    }
}

size_t sort_scratch_bytes(int maxItems) {
    int tiles = (maxItems + SORT_TILE - 1) / SORT_TILE;
    size_t words = 2 * (size_t)maxItems + 256 * (size_t)tiles + scan_tmp_words(256 * tiles) + 64;
    return words * sizeof(uint32_t);
}

void sort_scratch_bind(SortScratch& s, void* mem, int maxItems) {
    int tiles = (maxItems + SORT_TILE - 1) / SORT_TILE;
    uint32_t* p = static_cast<uint32_t*>(mem);
    s.keys2 = p; p += maxItems;
    s.vals2 = p; p += maxItems;
    s.hist = p; p += 256 * (size_t)tiles;
    s.scanTmp = p;
    s.maxItems = maxItems;
}

void radix_sort_pairs(uint32_t* keys, uint32_t* vals, int n, int bits, SortScratch& s, hipStream_t st) {
    if (n <= 1 || bits <= 0) return;
    int tiles = (n + SORT_TILE - 1) / SORT_TILE;
    uint32_t *ka = keys, *va = vals, *kb = s.keys2, *vb = s.vals2;
    int passes = (bits + 7) / 8;
    for (int p = 0; p < passes; ++p) {
        int shift = 8 * p;
        int nbits = (p == passes - 1) ? bits - shift : 8;
        hipLaunchKernelGGL(k_rs_hist, dim3(tiles), dim3(SORT_THREADS), 0, st, (const uint32_t*)ka, n,
                           shift, nbits, s.hist, tiles);
        exclusive_scan_u32(s.hist, 256 * tiles, s.scanTmp, st);
        hipLaunchKernelGGL(k_rs_scatter, dim3(tiles), dim3(SORT_THREADS), 0, st, (const uint32_t*)ka,
                           (const uint32_t*)va, kb, vb, n, shift, nbits, (const uint32_t*)s.hist, tiles);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        // a kernel, not a memcpy node, so that a captured step holds kernels only
        hipLaunchKernelGGL(k_copy_pairs, dim3((n + 255) / 256), dim3(256), 0, st, (const uint32_t*)ka,
                           (const uint32_t*)va, keys, vals, n);
    }
}

}  // namespace sphgpu
