// Surface reconstruction on the device: the reference's offline surfacer
// (Source/re.c: surface_field::surface_setup / surface_aniso /
// surface_isotro, surface_field.cpp:8-226; marchingCubes,
// marching_cubes.cpp:45-120) as HIP kernels:
//
//   1. particles binned by a radix sort into cells of h_field = 2 r_h;
//   2. k_sf_aniso, one thread per particle: Yu & Turk's weighted mean x_bar
//      (eq. 10, lambda = 1) and covariance (eq. 9) over the neighbours within
//      h_field (weights 1 - (r/h_field)^3, eq. 11), symmetric eigen
//      decomposition (cyclic Jacobi; equal to the SVD of the symmetric
//      positive covariance), singular values clamped to sigma0 / k_r and
//      scaled by k_s (eq. 15), G = U diag(1/sigma) U^T; fewer than N_e = 25
//      neighbours: G = k_n (3 - 2 n / N_e) I at the particle;
//   3. k_sf_splat, one block per particle: det(G) W(|G (p - x_bar)|) added to
//      every field node p of the ellipsoid's box (FP32 atomics, as the
//      reference accumulates RE_TYPE = float);
//   4. marching cubes: a count pass per cube, an exclusive scan, an emit pass
//      writing triangles in cube order (x fastest, then y, then z).
#include "mc_table.h"
#include "sph_common.h"
#include "sphgpu.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace sphgpu;

namespace {

struct SurfParams {
    float fmin[3];        // field grid origin
    int fnum[3];          // field grid nodes per axis
    float cell;           // field grid spacing
    float gmin[3];        // particle grid origin
    int gnum[3];          // particle grid cells per axis
    float ginv;           // 1 / h_field
    float hfield, hfield2, hf3inv;
    float rh, rh2;        // kernel radius of the field (radius_h)
    int hcells;           // int(radius_h / cell)
    float ks, kr, kn, lambda;
    int Ne;
    float k1, k2, k3, k4; // spline constants times the particle volume
};

struct SurfRec {          // per particle: ellipsoid centre, symmetric G, det(G), box half-width
    float x, y, z, det;
    float g00, g01, g02, g11;
    float g12, g22;
    int ext, pad;
};

__device__ __forceinline__ int cell_axis(float v, float mn, float inv, int n) {
    int c = int((v - mn) * inv);
    return min(max(c, 0), n - 1);
}

__global__ __launch_bounds__(256) void k_sf_keys(SurfParams S, const float4* __restrict__ x, int n,
                                                 uint32_t* keys, uint32_t* vals) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float4 p = x[i];
    int cx = cell_axis(p.x, S.gmin[0], S.ginv, S.gnum[0]);
    int cy = cell_axis(p.y, S.gmin[1], S.ginv, S.gnum[1]);
    int cz = cell_axis(p.z, S.gmin[2], S.ginv, S.gnum[2]);
    keys[i] = (uint32_t)((cz * S.gnum[1] + cy) * S.gnum[0] + cx);
    vals[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_sf_gather(const float4* __restrict__ x, const uint32_t* __restrict__ vals,
                                                   int n, float4* xs) {
    int r = blockIdx.x * 256 + threadIdx.x;
    if (r < n) xs[r] = x[vals[r]];
}

__global__ __launch_bounds__(256) void k_sf_fill_i2(int2* p, int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = make_int2(0, 0);
}

__global__ __launch_bounds__(256) void k_sf_ranges(const uint32_t* __restrict__ keys, int n, int2* ranges) {
    int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    uint32_t k = keys[r];
    if (r == 0 || keys[r - 1] != k) ranges[k].x = r;
    if (r == n - 1 || keys[r + 1] != k) ranges[k].y = r + 1;
}

// symmetric 3x3 eigen decomposition by cyclic Jacobi rotations (a fixed
// number of sweeps: deterministic); eigenvectors in the columns of V
__device__ void jacobi3(float a[3][3], float V[3][3], float ev[3]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) V[i][j] = i == j ? 1.0f : 0.0f;
    for (int sweep = 0; sweep < 8; ++sweep) {
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                const float apq = a[p][q];
                if (fabsf(apq) < 1e-30f) continue;
                const float theta = (a[q][q] - a[p][p]) / (2.0f * apq);
                const float t = (theta >= 0.0f ? 1.0f : -1.0f) / (fabsf(theta) + sqrtf(theta * theta + 1.0f));
                const float c = 1.0f / sqrtf(t * t + 1.0f), s = t * c;
                for (int k = 0; k < 3; ++k) {   // A <- A J (columns p, q)
                    const float akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {   // A <- J^T A (rows p, q)
                    const float apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    const float vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < 3; ++i) ev[i] = a[i][i];
}

__global__ __launch_bounds__(256) void k_sf_aniso(SurfParams S, const float4* __restrict__ xs,
                                                  const int2* __restrict__ ranges, int n, SurfRec* rec) {
    int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    const float4 pi = xs[r];
    const int cx = cell_axis(pi.x, S.gmin[0], S.ginv, S.gnum[0]);
    const int cy = cell_axis(pi.y, S.gmin[1], S.ginv, S.gnum[1]);
    const int cz = cell_axis(pi.z, S.gmin[2], S.ginv, S.gnum[2]);
    // pass 1: weighted mean (relative to the particle), neighbour count
    float sw = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f;
    int cnt = 0;
    for (int z = max(cz - 1, 0); z <= min(cz + 1, S.gnum[2] - 1); ++z)
        for (int y = max(cy - 1, 0); y <= min(cy + 1, S.gnum[1] - 1); ++y)
            for (int x = max(cx - 1, 0); x <= min(cx + 1, S.gnum[0] - 1); ++x) {
                const int2 rg = ranges[(z * S.gnum[1] + y) * S.gnum[0] + x];
                for (int j = rg.x; j < rg.y; ++j) {
                    if (j == r) continue;
                    const float4 pj = xs[j];
                    const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
                    const float d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 > S.hfield2) continue;
                    const float d = sqrtf(d2);
                    const float w = 1.0f - d * d * d * S.hf3inv;
                    sw += w; sx += w * dx; sy += w * dy; sz += w * dz;
                    ++cnt;
                }
            }
    SurfRec o;
    float sig0;
    if (cnt >= S.Ne) {
        const float mx = S.lambda * sx / sw, my = S.lambda * sy / sw, mz = S.lambda * sz / sw;
        // pass 2: covariance about x_bar (eq. 9 with x_bar of eq. 6)
        float c00 = 0.f, c01 = 0.f, c02 = 0.f, c11 = 0.f, c12 = 0.f, c22 = 0.f;
        for (int z = max(cz - 1, 0); z <= min(cz + 1, S.gnum[2] - 1); ++z)
            for (int y = max(cy - 1, 0); y <= min(cy + 1, S.gnum[1] - 1); ++y)
                for (int x = max(cx - 1, 0); x <= min(cx + 1, S.gnum[0] - 1); ++x) {
                    const int2 rg = ranges[(z * S.gnum[1] + y) * S.gnum[0] + x];
                    for (int j = rg.x; j < rg.y; ++j) {
                        if (j == r) continue;
                        const float4 pj = xs[j];
                        const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
                        const float d2 = dx * dx + dy * dy + dz * dz;
                        if (d2 > S.hfield2) continue;
                        const float d = sqrtf(d2);
                        const float w = 1.0f - d * d * d * S.hf3inv;
                        const float ex = dx - mx, ey = dy - my, ez = dz - mz;
                        c00 += w * ex * ex; c01 += w * ex * ey; c02 += w * ex * ez;
                        c11 += w * ey * ey; c12 += w * ey * ez; c22 += w * ez * ez;
                    }
                }
        const float iw = 1.0f / sw;
        float A[3][3] = {{c00 * iw, c01 * iw, c02 * iw}, {c01 * iw, c11 * iw, c12 * iw}, {c02 * iw, c12 * iw, c22 * iw}};
        float V[3][3], ev[3];
        jacobi3(A, V, ev);
        // singular values in descending order (JAMA's SVD order)
        int o0 = 0, o1 = 1, o2 = 2;
        if (ev[o0] < ev[o1]) { int t = o0; o0 = o1; o1 = t; }
        if (ev[o1] < ev[o2]) { int t = o1; o1 = o2; o2 = t; }
        if (ev[o0] < ev[o1]) { int t = o0; o0 = o1; o1 = t; }
        float s0 = fmaxf(ev[o0], 0.0f), s1 = fmaxf(ev[o1], 0.0f), s2 = fmaxf(ev[o2], 0.0f);
        s1 = fmaxf(s1, s0 / S.kr);
        s2 = fmaxf(s2, s0 / S.kr);
        s0 *= S.ks; s1 *= S.ks; s2 *= S.ks;
        const float i0 = 1.0f / s0, i1 = 1.0f / s1, i2 = 1.0f / s2;
        const int oo[3] = {o0, o1, o2};
        const float is[3] = {i0, i1, i2};
        float G[3][3];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                float g = 0.0f;
                for (int k = 0; k < 3; ++k) g += V[a][oo[k]] * is[k] * V[b][oo[k]];
                G[a][b] = g;
            }
        o.x = pi.x + mx; o.y = pi.y + my; o.z = pi.z + mz;
        o.det = i0 * i1 * i2;
        o.g00 = G[0][0]; o.g01 = 0.5f * (G[0][1] + G[1][0]); o.g02 = 0.5f * (G[0][2] + G[2][0]);
        o.g11 = G[1][1]; o.g12 = 0.5f * (G[1][2] + G[2][1]); o.g22 = G[2][2];
        sig0 = s0;
    } else {
        const float g = S.kn * (3.0f - 2.0f * float(cnt) / float(S.Ne));
        o.x = pi.x; o.y = pi.y; o.z = pi.z;
        o.det = g * g * g;
        o.g00 = g; o.g01 = 0.0f; o.g02 = 0.0f; o.g11 = g; o.g12 = 0.0f; o.g22 = g;
        sig0 = 1.0f / S.kn;
    }
    o.ext = int(float(S.hcells) * sig0);
    o.pad = cnt;
    rec[r] = o;
}

// cubic spline of the field (surface_field.cpp:233-250), times the particle volume
__device__ __forceinline__ float ker_field(const SurfParams& S, float r) {
    return r < 0.5f * S.rh ? S.k1 + S.k2 * r * r + S.k3 * r * r * r
                           : S.k4 * (S.rh - r) * (S.rh - r) * (S.rh - r);
}

// one block per particle over the nodes of its box [p - ext, p + ext + 1]
__global__ __launch_bounds__(256) void k_sf_splat(SurfParams S, const SurfRec* __restrict__ rec, int n,
                                                  float* __restrict__ field) {
    const int r = blockIdx.x;
    if (r >= n) return;
    const SurfRec o = rec[r];
    // The nodes with |G d| <= rh lie in the ellipsoid's bounding box, whose
    // half-width along axis a is rh sqrt(sum_b (G^-1)_ab^2) (G symmetric);
    // it is searched instead of the cube of half-width ext (the largest
    // stretch along every axis), one node of margin added for rounding.
    const float c00 = o.g11 * o.g22 - o.g12 * o.g12, c01 = o.g02 * o.g12 - o.g01 * o.g22;
    const float c02 = o.g01 * o.g12 - o.g02 * o.g11, c11 = o.g00 * o.g22 - o.g02 * o.g02;
    const float c12 = o.g01 * o.g02 - o.g00 * o.g12, c22 = o.g00 * o.g11 - o.g01 * o.g01;
    const float dg = o.g00 * c00 + o.g01 * c01 + o.g02 * c02;   // det G
    const float half[3] = {sqrtf(c00 * c00 + c01 * c01 + c02 * c02), sqrtf(c01 * c01 + c11 * c11 + c12 * c12),
                           sqrtf(c02 * c02 + c12 * c12 + c22 * c22)};
    int lo[3], hi[3];
    const float xb[3] = {o.x, o.y, o.z};
    for (int a = 0; a < 3; ++a) {
        const int p = int((xb[a] - S.fmin[a]) / S.cell);
        int e = o.ext;
        if (dg > 0.0f) {
            const float w = S.rh * half[a] / dg / S.cell;    // half-width in nodes
            if (w < float(o.ext)) e = min(o.ext, int(w) + 2);
        }
        lo[a] = max(p - e, 0);
        hi[a] = min(p + e + 1, S.fnum[a] - 1);
    }
    const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
    if (nx <= 0 || ny <= 0 || nz <= 0) return;
    const int total = nx * ny * nz;
    for (int t = threadIdx.x; t < total; t += 256) {
        const int ix = lo[0] + t % nx, iy = lo[1] + (t / nx) % ny, iz = lo[2] + t / (nx * ny);
        const float dx = S.fmin[0] + ix * S.cell - o.x;
        const float dy = S.fmin[1] + iy * S.cell - o.y;
        const float dz = S.fmin[2] + iz * S.cell - o.z;
        const float ux = o.g00 * dx + o.g01 * dy + o.g02 * dz;
        const float uy = o.g01 * dx + o.g11 * dy + o.g12 * dz;
        const float uz = o.g02 * dx + o.g12 * dy + o.g22 * dz;
        const float q2 = ux * ux + uy * uy + uz * uz;
        if (q2 <= S.rh2) atomicAdd(&field[((size_t)iz * S.fnum[1] + iy) * S.fnum[0] + ix], o.det * ker_field(S, sqrtf(q2)));
    }
}

// surface_isotro (surface_field.cpp:45-75): W(|p - x|) over the box of
// half-width int(r_h / cell) around each particle
__global__ __launch_bounds__(256) void k_sf_isotro(SurfParams S, const float4* __restrict__ x, int n,
                                                   float* __restrict__ field) {
    const int r = blockIdx.x;
    if (r >= n) return;
    const float4 p = x[r];
    const float xb[3] = {p.x, p.y, p.z};
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        const int c = int((xb[a] - S.fmin[a]) / S.cell);
        lo[a] = max(c - S.hcells, 0);
        hi[a] = min(c + S.hcells + 1, S.fnum[a] - 1);
    }
    const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
    if (nx <= 0 || ny <= 0 || nz <= 0) return;
    const int total = nx * ny * nz;
    for (int t = threadIdx.x; t < total; t += 256) {
        const int ix = lo[0] + t % nx, iy = lo[1] + (t / nx) % ny, iz = lo[2] + t / (nx * ny);
        const float dx = p.x - (S.fmin[0] + ix * S.cell);
        const float dy = p.y - (S.fmin[1] + iy * S.cell);
        const float dz = p.z - (S.fmin[2] + iz * S.cell);
        const float d2 = dx * dx + dy * dy + dz * dz;
        if (d2 <= S.rh2) atomicAdd(&field[((size_t)iz * S.fnum[1] + iy) * S.fnum[0] + ix], ker_field(S, sqrtf(d2)));
    }
}

struct McTable {
    unsigned short edges[256];
    signed char tris[256][16];
    unsigned char ntris[256];
};
__constant__ McTable c_mc;

__device__ __forceinline__ int mc_case(const float* __restrict__ f, const int N[3], int ix, int iy, int iz, float iso,
                                       float v[8]) {
    const size_t sy = N[0], sz = (size_t)N[0] * N[1];
    const size_t base = (size_t)iz * sz + (size_t)iy * sy + ix;
    const int off[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}, {1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}};
    int ci = 0;
    for (int k = 0; k < 8; ++k) {
        v[k] = f[base + off[k][0] + off[k][1] * sy + off[k][2] * sz];
        if (v[k] < iso) ci |= 1 << k;
    }
    return ci;
}

__global__ __launch_bounds__(256) void k_mc_count(const float* __restrict__ f, int nx, int ny, int nz, float iso,
                                                  uint32_t* __restrict__ count) {
    const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t cubes = (size_t)(nx - 1) * (ny - 1) * (nz - 1);
    if (c >= cubes) return;
    const int ix = int(c % (nx - 1)), iy = int((c / (nx - 1)) % (ny - 1)), iz = int(c / ((size_t)(nx - 1) * (ny - 1)));
    const int N[3] = {nx, ny, nz};
    float v[8];
    count[c] = c_mc.ntris[mc_case(f, N, ix, iy, iz, iso, v)];
}

__global__ __launch_bounds__(256) void k_mc_emit(const float* __restrict__ f, int nx, int ny, int nz, float iso,
                                                 float mnx, float mny, float mnz, float cell,
                                                 const uint32_t* __restrict__ offset, float* __restrict__ out,
                                                 long long cap) {
    const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t cubes = (size_t)(nx - 1) * (ny - 1) * (nz - 1);
    if (c >= cubes) return;
    const int ix = int(c % (nx - 1)), iy = int((c / (nx - 1)) % (ny - 1)), iz = int(c / ((size_t)(nx - 1) * (ny - 1)));
    const int N[3] = {nx, ny, nz};
    float v[8];
    const int ci = mc_case(f, N, ix, iy, iz, iso, v);
    const int nt = c_mc.ntris[ci];
    if (nt == 0) return;
    const int off[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}, {1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}};
    const int E[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6}, {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
    const long long first = offset[c];
    for (int t = 0; t < nt; ++t) {
        if (first + t >= cap) return;
        for (int m = 0; m < 3; ++m) {
            const int e = c_mc.tris[ci][3 * t + m];
            int a = E[e][0], b = E[e][1];
            // VertexInterp (marching_cubes.cpp:40-41) from the grid edge's
            // lower node, with node positions from integer indices: the
            // cubes on either side of an edge compute the same vertex
            if (off[a][0] + off[a][1] + off[a][2] > off[b][0] + off[b][1] + off[b][2]) {
                const int tmp = a; a = b; b = tmp;
            }
            const float mu = (iso - v[a]) / (v[b] - v[a]);
            const float ax = mnx + float(ix + off[a][0]) * cell, ay = mny + float(iy + off[a][1]) * cell,
                        az = mnz + float(iz + off[a][2]) * cell;
            const float bx = mnx + float(ix + off[b][0]) * cell, by = mny + float(iy + off[b][1]) * cell,
                        bz = mnz + float(iz + off[b][2]) * cell;
            float* o = out + 9 * (first + t) + 3 * m;
            o[0] = ax + mu * (bx - ax);
            o[1] = ay + mu * (by - ay);
            o[2] = az + mu * (bz - az);
        }
    }
}

}  // namespace

struct sphgpu_surfacer {
    SurfParams S{};
    double fmin[3], cell;
    int fnum[3];
    size_t nfield = 0, ncells = 0;
    hipStream_t st = nullptr;
    float* field = nullptr;
    int cap = 0;
    float4 *x = nullptr, *xs = nullptr;
    uint32_t *keys = nullptr, *vals = nullptr;
    int2* ranges = nullptr;
    SurfRec* rec = nullptr;
    void* sortMem = nullptr;
    SortScratch sort{};
    int keyBits = 1;
    uint32_t* mcCount = nullptr;
    uint32_t* mcTmp = nullptr;
    size_t mcCap = 0;
    int nrec = 0;
    std::string err;
};

namespace {

int sfail(sphgpu_surfacer* s, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (s) s->err = buf;
    return code;
}

#define SCHECK(s, expr)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) return sfail(s, SPHGPU_E_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

int ensure_particles(sphgpu_surfacer* s, int n) {
    if (n <= s->cap) return 0;
    (void)hipFree(s->x); (void)hipFree(s->xs); (void)hipFree(s->keys); (void)hipFree(s->vals);
    (void)hipFree(s->rec); (void)hipFree(s->sortMem);
    s->x = s->xs = nullptr; s->keys = s->vals = nullptr; s->rec = nullptr; s->sortMem = nullptr;
    const int cap = std::max(n, 1024);
    SCHECK(s, hipMalloc(&s->x, sizeof(float4) * cap));
    SCHECK(s, hipMalloc(&s->xs, sizeof(float4) * cap));
    SCHECK(s, hipMalloc(&s->keys, sizeof(uint32_t) * cap));
    SCHECK(s, hipMalloc(&s->vals, sizeof(uint32_t) * cap));
    SCHECK(s, hipMalloc(&s->rec, sizeof(SurfRec) * cap));
    SCHECK(s, hipMalloc(&s->sortMem, sort_scratch_bytes(cap)));
    sort_scratch_bind(s->sort, s->sortMem, cap);
    s->cap = cap;
    return 0;
}

// particles to the device, binned; per-particle ellipsoids when aniso
int prepare(sphgpu_surfacer* s, long long n, const float* xyz, bool aniso) {
    if (n < 0 || (n > 0 && !xyz)) return sfail(s, SPHGPU_E_ARG, "bad particle array");
    if (n > (1 << 30)) return sfail(s, SPHGPU_E_CAPACITY, "too many particles");
    if (int rc = ensure_particles(s, int(n))) return rc;
    std::vector<float4> h(static_cast<size_t>(n));
    for (long long i = 0; i < n; ++i) {
        h[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0.0f);
        if (!std::isfinite(h[i].x) || !std::isfinite(h[i].y) || !std::isfinite(h[i].z))
            return sfail(s, SPHGPU_E_ARG, "non-finite position %lld", i);
    }
    SCHECK(s, hipMemcpyAsync(s->x, h.data(), sizeof(float4) * n, hipMemcpyHostToDevice, s->st));
    SCHECK(s, hipMemsetAsync(s->field, 0, sizeof(float) * s->nfield, s->st));
    s->nrec = int(n);
    if (!aniso || n == 0) return 0;
    const int blocks = int((n + 255) / 256);
    hipLaunchKernelGGL(k_sf_keys, dim3(blocks), dim3(256), 0, s->st, s->S, (const float4*)s->x, int(n), s->keys, s->vals);
    radix_sort_pairs(s->keys, s->vals, int(n), s->keyBits, s->sort, s->st);
    hipLaunchKernelGGL(k_sf_gather, dim3(blocks), dim3(256), 0, s->st, (const float4*)s->x, (const uint32_t*)s->vals,
                       int(n), s->xs);
    hipLaunchKernelGGL(k_sf_fill_i2, dim3(int((s->ncells + 255) / 256)), dim3(256), 0, s->st, s->ranges, int(s->ncells));
    hipLaunchKernelGGL(k_sf_ranges, dim3(blocks), dim3(256), 0, s->st, (const uint32_t*)s->keys, int(n), s->ranges);
    hipLaunchKernelGGL(k_sf_aniso, dim3(blocks), dim3(256), 0, s->st, s->S, (const float4*)s->xs,
                       (const int2*)s->ranges, int(n), s->rec);
    return 0;
}

int upload_table(sphgpu_surfacer* s) {
    const mctable::Table& t = mctable::table();
    McTable m{};
    std::memcpy(m.edges, t.edges, sizeof(m.edges));
    for (int c = 0; c < 256; ++c) {
        m.ntris[c] = t.ntris[c];
        for (int k = 0; k < 16; ++k) m.tris[c][k] = t.tris[c][k];
    }
    SCHECK(s, hipMemcpyToSymbolAsync(HIP_SYMBOL(c_mc), &m, sizeof(m), 0, hipMemcpyHostToDevice, s->st));
    SCHECK(s, hipStreamSynchronize(s->st));
    return 0;
}

}  // namespace

extern "C" {

int sphgpu_marching_cubes_table(unsigned short* edges, unsigned char* ntris, signed char* tris) {
    const mctable::Table& t = mctable::table();
    if (edges) std::memcpy(edges, t.edges, sizeof(t.edges));
    if (ntris) std::memcpy(ntris, t.ntris, sizeof(t.ntris));
    if (tris)
        for (int c = 0; c < 256; ++c) std::memcpy(tris + 16 * c, t.tris[c], 16);
    return SPHGPU_OK;
}

int sphgpu_surfacer_create(const float mn[3], const float mx[3], float radius_p, float radius_h, float cell_size,
                           sphgpu_surfacer** out) {
    if (!mn || !mx || !out) return SPHGPU_E_ARG;
    *out = nullptr;
    if (!(radius_p > 0) || !(radius_h > 0) || !(cell_size > 0)) return SPHGPU_E_ARG;
    for (int a = 0; a < 3; ++a)
        if (!(mx[a] > mn[a])) return SPHGPU_E_ARG;
    auto* s = new sphgpu_surfacer();
    SurfParams& S = s->S;
    // surface_setup (surface_field.cpp:8-43), in double on the host
    const double rp = radius_p, rh = radius_h, cell = cell_size, vol = rp * rp * rp, hf = 2.0 * rh;
    size_t nf = 1, nc = 1;
    for (int a = 0; a < 3; ++a) {
        s->fmin[a] = double(mn[a]) - rh * 0.75;
        s->fnum[a] = int((double(mx[a]) - double(mn[a]) + rh * 1.5) / cell) + 1;
        S.fmin[a] = float(s->fmin[a]);
        S.fnum[a] = s->fnum[a];
        nf *= size_t(s->fnum[a]);
        S.gmin[a] = float(double(mn[a]) - hf * 1.25);
        S.gnum[a] = int((double(mx[a]) - double(mn[a]) + hf * 2.5) / hf) + 1;
        nc *= size_t(S.gnum[a]);
    }
    if (nf > (size_t(1) << 31) || nc > (size_t(1) << 30)) {
        delete s;
        return SPHGPU_E_CAPACITY;
    }
    s->cell = cell;
    s->nfield = nf;
    s->ncells = nc;
    S.cell = float(cell);
    S.ginv = float(1.0 / hf);
    S.hfield = float(hf);
    S.hfield2 = float(hf * hf);
    S.hf3inv = float(1.0 / (hf * hf * hf));
    S.rh = float(rh);
    S.rh2 = float(rh * rh);
    S.hcells = int(rh / cell);
    S.kr = 4.0f;                                   // eq. (15), surface_field.cpp:33-37
    S.ks = float(1.0 / (0.16 * hf * hf) * 1.6);
    S.kn = 0.95f;
    S.Ne = 25;
    S.lambda = 1.0f;                               // eq. (6), surface_field.cpp:39
    S.k1 = float(8.0 / (M_PI * std::pow(rh, 3.0)) * vol);
    S.k2 = float(-48.0 / (M_PI * std::pow(rh, 5.0)) * vol);
    S.k3 = float(48.0 / (M_PI * std::pow(rh, 6.0)) * vol);
    S.k4 = float(16.0 / (M_PI * std::pow(rh, 6.0)) * vol);
    while ((1ull << s->keyBits) <= nc) ++s->keyBits;
    if (!sphgpu_device_available()) {
        *out = s;   // parameters are queryable; compute calls report E_DEVICE
        s->st = nullptr;
        return SPHGPU_OK;
    }
    if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&s->field, sizeof(float) * nf) != hipSuccess ||
        hipMalloc(&s->ranges, sizeof(int2) * nc) != hipSuccess || upload_table(s) != 0) {
        sphgpu_surfacer_destroy(s);
        return SPHGPU_E_DEVICE;
    }
    *out = s;
    return SPHGPU_OK;
}

void sphgpu_surfacer_destroy(sphgpu_surfacer* s) {
    if (!s) return;
    if (s->st) (void)hipStreamSynchronize(s->st);
    (void)hipFree(s->field); (void)hipFree(s->ranges); (void)hipFree(s->x); (void)hipFree(s->xs);
    (void)hipFree(s->keys); (void)hipFree(s->vals); (void)hipFree(s->rec); (void)hipFree(s->sortMem);
    (void)hipFree(s->mcCount); (void)hipFree(s->mcTmp);
    if (s->st) (void)hipStreamDestroy(s->st);
    delete s;
}

const char* sphgpu_surfacer_last_error(const sphgpu_surfacer* s) { return s ? s->err.c_str() : "null surfacer"; }

int sphgpu_surfacer_grid(const sphgpu_surfacer* s, double grid_min[3], int grid_num[3], double* cell_size) {
    if (!s) return SPHGPU_E_ARG;
    for (int a = 0; a < 3; ++a) {
        if (grid_min) grid_min[a] = s->fmin[a];
        if (grid_num) grid_num[a] = s->fnum[a];
    }
    if (cell_size) *cell_size = s->cell;
    return SPHGPU_OK;
}

int sphgpu_surfacer_aniso(sphgpu_surfacer* s, long long n, const float* xyz) {
    if (!s) return SPHGPU_E_ARG;
    if (!s->st) return sfail(s, SPHGPU_E_DEVICE, "no gfx9 HIP device available");
    if (int rc = prepare(s, n, xyz, true)) return rc;
    if (n > 0)
        hipLaunchKernelGGL(k_sf_splat, dim3(unsigned(n)), dim3(256), 0, s->st, s->S, (const SurfRec*)s->rec, int(n),
                           s->field);
    SCHECK(s, hipGetLastError());
    SCHECK(s, hipStreamSynchronize(s->st));
    return SPHGPU_OK;
}

int sphgpu_surfacer_isotro(sphgpu_surfacer* s, long long n, const float* xyz) {
    if (!s) return SPHGPU_E_ARG;
    if (!s->st) return sfail(s, SPHGPU_E_DEVICE, "no gfx9 HIP device available");
    if (int rc = prepare(s, n, xyz, false)) return rc;
    if (n > 0)
        hipLaunchKernelGGL(k_sf_isotro, dim3(unsigned(n)), dim3(256), 0, s->st, s->S, (const float4*)s->x, int(n),
                           s->field);
    SCHECK(s, hipGetLastError());
    SCHECK(s, hipStreamSynchronize(s->st));
    return SPHGPU_OK;
}

int sphgpu_surfacer_particles(sphgpu_surfacer* s, float* xbar, float* G, float* det, int* neighbours) {
    if (!s) return SPHGPU_E_ARG;
    if (!s->st) return sfail(s, SPHGPU_E_DEVICE, "no gfx9 HIP device available");
    const int n = s->nrec;
    std::vector<SurfRec> r(size_t(std::max(n, 0)));
    std::vector<uint32_t> order(size_t(std::max(n, 0)));
    if (n > 0) {
        SCHECK(s, hipMemcpyAsync(r.data(), s->rec, sizeof(SurfRec) * n, hipMemcpyDeviceToHost, s->st));
        SCHECK(s, hipMemcpyAsync(order.data(), s->vals, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s->st));
    }
    SCHECK(s, hipStreamSynchronize(s->st));
    for (int k = 0; k < n; ++k) {
        const int i = int(order[k]);
        const SurfRec& o = r[k];
        if (xbar) { xbar[3 * i] = o.x; xbar[3 * i + 1] = o.y; xbar[3 * i + 2] = o.z; }
        if (G) {
            const float g[9] = {o.g00, o.g01, o.g02, o.g01, o.g11, o.g12, o.g02, o.g12, o.g22};
            std::memcpy(G + 9 * (size_t)i, g, sizeof(g));
        }
        if (det) det[i] = o.det;
        if (neighbours) neighbours[i] = o.pad;
    }
    return SPHGPU_OK;
}

int sphgpu_surfacer_get_field(sphgpu_surfacer* s, float* out) {
    if (!s || !out) return SPHGPU_E_ARG;
    if (!s->st) return sfail(s, SPHGPU_E_DEVICE, "no gfx9 HIP device available");
    SCHECK(s, hipMemcpyAsync(out, s->field, sizeof(float) * s->nfield, hipMemcpyDeviceToHost, s->st));
    SCHECK(s, hipStreamSynchronize(s->st));
    return SPHGPU_OK;
}

int sphgpu_surfacer_set_field(sphgpu_surfacer* s, const float* in) {
    if (!s || !in) return SPHGPU_E_ARG;
    if (!s->st) return sfail(s, SPHGPU_E_DEVICE, "no gfx9 HIP device available");
    SCHECK(s, hipMemcpyAsync(s->field, in, sizeof(float) * s->nfield, hipMemcpyHostToDevice, s->st));
    SCHECK(s, hipStreamSynchronize(s->st));
    return SPHGPU_OK;
}

int sphgpu_surfacer_marching_cubes(sphgpu_surfacer* s, float iso, float* tris, long long cap, long long* ntris) {
    if (!s || !ntris) return SPHGPU_E_ARG;
    if (!s->st) return sfail(s, SPHGPU_E_DEVICE, "no gfx9 HIP device available");
    if (!std::isfinite(iso)) return sfail(s, SPHGPU_E_ARG, "non-finite iso level");
    const int nx = s->fnum[0], ny = s->fnum[1], nz = s->fnum[2];
    *ntris = 0;
    if (nx < 2 || ny < 2 || nz < 2) return SPHGPU_OK;
    const size_t cubes = size_t(nx - 1) * (ny - 1) * (nz - 1);
    if (cubes >= (size_t(1) << 31)) return sfail(s, SPHGPU_E_CAPACITY, "too many cubes");
    if (s->mcCap < cubes) {
        (void)hipFree(s->mcCount); (void)hipFree(s->mcTmp);
        s->mcCount = s->mcTmp = nullptr;
        SCHECK(s, hipMalloc(&s->mcCount, sizeof(uint32_t) * (cubes + 1)));
        SCHECK(s, hipMalloc(&s->mcTmp, sizeof(uint32_t) * (size_t)scan_tmp_words(int(cubes + 1))));
        s->mcCap = cubes;
    }
    const unsigned blocks = unsigned((cubes + 255) / 256);
    hipLaunchKernelGGL(k_mc_count, dim3(blocks), dim3(256), 0, s->st, (const float*)s->field, nx, ny, nz, iso,
                       s->mcCount);
    SCHECK(s, hipMemsetAsync(s->mcCount + cubes, 0, sizeof(uint32_t), s->st));
    exclusive_scan_u32(s->mcCount, int(cubes + 1), s->mcTmp, s->st);
    uint32_t total = 0;
    SCHECK(s, hipMemcpyAsync(&total, s->mcCount + cubes, sizeof(uint32_t), hipMemcpyDeviceToHost, s->st));
    SCHECK(s, hipStreamSynchronize(s->st));
    *ntris = total;
    if (!tris || total == 0) return SPHGPU_OK;
    const long long keep = std::min<long long>(cap, total);
    float* dout = nullptr;
    SCHECK(s, hipMalloc(&dout, sizeof(float) * 9 * std::max<long long>(keep, 1)));
    hipLaunchKernelGGL(k_mc_emit, dim3(blocks), dim3(256), 0, s->st, (const float*)s->field, nx, ny, nz, iso,
                       float(s->fmin[0]), float(s->fmin[1]), float(s->fmin[2]), float(s->cell),
                       (const uint32_t*)s->mcCount, dout, keep);
    hipError_t e1 = hipMemcpyAsync(tris, dout, sizeof(float) * 9 * keep, hipMemcpyDeviceToHost, s->st);
    hipError_t e2 = hipStreamSynchronize(s->st);
    (void)hipFree(dout);
    if (e1 != hipSuccess || e2 != hipSuccess) return sfail(s, SPHGPU_E_DEVICE, "marching cubes readback failed");
    if (keep < (long long)total) return sfail(s, SPHGPU_E_CAPACITY, "%u triangles, room for %lld", total, cap);
    return SPHGPU_OK;
}

}  // extern "C"
