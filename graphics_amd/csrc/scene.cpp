// Host-side scene samplers: the lattice fluid block and the box-surface
// boundary sampler of the reference's scene builder, plus seeded jitter.
#include "sphgpu.h"

#include <cmath>
#include <cstdint>

namespace {

// PCG32 (O'Neill, XSH-RR 64/32), seeded as pcg32_srandom_r(initstate, initseq).
struct Pcg32 {
    uint64_t state = 0, inc = 0;
    Pcg32(uint64_t seed, uint64_t seq) {
        inc = (seq << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xorshifted = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = uint32_t(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((~rot + 1u) & 31u));
    }
    // uniform in [-1, 1)
    double sym() { return 2.0 * (double(next()) * (1.0 / 4294967296.0)) - 1.0; }
};

}  // namespace

extern "C" long long sphgpu_scene_lattice(const int n[3], const float origin[3], float spacing,
                                          float jitter, unsigned long long seed,
                                          unsigned long long stream, float* xyz) {
    if (!n || !origin || n[0] < 0 || n[1] < 0 || n[2] < 0) return -1;
    long long total = (long long)n[0] * n[1] * n[2];
    if (!xyz) return total;
    Pcg32 rng(seed, stream);
    const double s = spacing, amp = double(jitter) * s;
    long long o = 0;
    // x fastest, then y, then z: the insertion order of SphWithIO.cpp:175-192
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                double px = origin[0] + i * s, py = origin[1] + j * s, pz = origin[2] + k * s;
                if (amp != 0.0) {
                    px += amp * rng.sym();
                    py += amp * rng.sym();
                    pz += amp * rng.sym();
                }
                xyz[3 * o] = float(px);
                xyz[3 * o + 1] = float(py);
                xyz[3 * o + 2] = float(pz);
                ++o;
            }
    return total;
}

// Box-surface lattice points (SphWithIO::makeCubeSampleParts, SphWithIO.cpp:489-522):
// per axis num = ceil(2*half/r) intervals of pro = 2*half/num; every lattice
// point with a coordinate index at 0 or num is kept, except the open top face
// (index num along y with the x/z indices interior) when have_top == 0.
extern "C" long long sphgpu_scene_box(const float center[3], const float half[3], float spacing,
                                      int have_top, float* xyz) {
    if (!center || !half || !(spacing > 0)) return -1;
    int num[3];
    float pro[3], lo[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = -half[a];
        // The reference takes ceil() of the FP32 quotient; for extents that are
        // whole multiples of r (0.6/0.02) that quotient can round just above
        // the integer and add one interval.  A relative slack of 1e-5 keeps
        // the sample spacing equal to r in that case.
        double q = double(half[a] - lo[a]) / double(spacing);
        num[a] = int(std::ceil(q * (1.0 - 1e-5)));
        pro[a] = num[a] > 0 ? (half[a] - lo[a]) / num[a] : 1.0f;
    }
    long long o = 0;
    for (int k = 0; k <= num[2]; ++k)
        for (int j = 0; j <= num[1]; ++j)
            for (int i = 0; i <= num[0]; ++i) {
                int pi[3] = {i, j, k};
                bool interior = true;
                for (int a = 0; a < 3; ++a) interior = interior && (pi[a] > 0 && pi[a] < num[a]);
                if (interior) continue;
                if (!have_top && j == num[1]) {
                    // top face: skip when the point below it would be interior
                    bool below_interior = (i > 0 && i < num[0]) && (j - 1 > 0 && j - 1 < num[1]) &&
                                          (k > 0 && k < num[2]);
                    if (below_interior) continue;
                }
                if (xyz) {
                    xyz[3 * o] = center[0] + (lo[0] + float(i) * pro[0]);
                    xyz[3 * o + 1] = center[1] + (lo[1] + float(j) * pro[1]);
                    xyz[3 * o + 2] = center[2] + (lo[2] + float(k) * pro[2]);
                }
                ++o;
            }
    return o;
}
