// Marching-cubes case table built from the cube's geometry (see mc_table.h).
#include "mc_table.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

namespace mctable {

const int kCorner[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}, {1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}};
const int kEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6}, {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

namespace {

int edge_of(int a, int b) {
    for (int e = 0; e < 12; ++e)
        if ((kEdge[e][0] == a && kEdge[e][1] == b) || (kEdge[e][0] == b && kEdge[e][1] == a)) return e;
    throw std::logic_error("corners not adjacent");
}

// true when the three edges lie in one face of the cube (a face is a
// coordinate fixed at 0 or 1 for all six end corners)
bool in_one_face(int e0, int e1, int e2) {
    const int es[3] = {e0, e1, e2};
    for (int axis = 0; axis < 3; ++axis)
        for (int side = 0; side < 2; ++side) {
            bool all = true;
            for (int e : es)
                for (int k = 0; k < 2; ++k) all = all && kCorner[kEdge[e][k]][axis] == side;
            if (all) return true;
        }
    return false;
}

// the six faces, corners counter-clockwise seen from outside the cube
std::vector<std::vector<int>> faces() {
    std::vector<std::vector<int>> out;
    for (int axis = 0; axis < 3; ++axis)
        for (int side = 0; side < 2; ++side) {
            std::vector<int> cs;
            for (int c = 0; c < 8; ++c)
                if (kCorner[c][axis] == side) cs.push_back(c);
            // in-plane axes u, v with (u, v, n) right-handed for the outward normal
            const int u = (axis + 1) % 3, v = (axis + 2) % 3;
            const double sgn = side ? 1.0 : -1.0;
            std::sort(cs.begin(), cs.end(), [&](int a, int b) {
                const double aa = std::atan2(sgn * (kCorner[a][v] - 0.5), kCorner[a][u] - 0.5);
                const double bb = std::atan2(sgn * (kCorner[b][v] - 0.5), kCorner[b][u] - 0.5);
                return aa < bb;
            });
            out.push_back(cs);
        }
    return out;
}

Table build() {
    Table t{};
    const auto F = faces();
    for (int ci = 0; ci < 256; ++ci) {
        auto low = [&](int c) { return (ci >> c) & 1; };
        uint16_t em = 0;
        for (int e = 0; e < 12; ++e)
            if (low(kEdge[e][0]) != low(kEdge[e][1])) em |= uint16_t(1u << e);
        t.edges[ci] = em;
        // per face: segments from the crossing leaving a run of low corners
        // (counter-clockwise) back to the crossing entering it
        int next[12];
        std::fill(next, next + 12, -1);
        for (const auto& f : F) {
            for (int k = 0; k < 4; ++k) {
                const int a = f[k], b = f[(k + 1) % 4];
                if (!(low(a) && !low(b))) continue;          // exit of a low run at edge a-b
                // walk backwards to the entry of this run
                int j = k;
                while (low(f[(j + 3) % 4]) && f[(j + 3) % 4] != b) j = (j + 3) % 4;
                const int p = f[(j + 3) % 4], q = f[j];          // entry edge p (high) -> q (low)
                next[edge_of(a, b)] = edge_of(p, q);
            }
        }
        // chain into loops
        std::vector<std::vector<int>> loops;
        bool used[12] = {false};
        for (int e = 0; e < 12; ++e) {
            if (!((em >> e) & 1) || used[e]) continue;
            std::vector<int> loop;
            int cur = e;
            while (!used[cur]) {
                used[cur] = true;
                loop.push_back(cur);
                cur = next[cur];
                if (cur < 0) throw std::logic_error("open loop");
            }
            if (cur != e) throw std::logic_error("loop does not close");
            loops.push_back(loop);
        }
        // fan triangles in loop order: the face rule orients every loop the
        // same way relative to its low corners, so one winding fits all cases.
        // The fan starts at the first loop vertex for which no triangle lies
        // in a cube face (a loop crossing an ambiguous face twice would
        // otherwise put a triangle in the face, doubled by the neighbour).
        std::vector<int> tri;
        for (const auto& L : loops) {
            const size_t m = L.size();
            size_t start = 0;
            for (; start < m; ++start) {
                bool flat = false;
                for (size_t k = 1; k + 1 < m && !flat; ++k)
                    flat = in_one_face(L[start], L[(start + k) % m], L[(start + k + 1) % m]);
                if (!flat) break;
            }
            if (start == m) throw std::logic_error("no fan without a face triangle");
            for (size_t k = 1; k + 1 < m; ++k) {
                tri.push_back(L[start]);
                tri.push_back(L[(start + k) % m]);
                tri.push_back(L[(start + k + 1) % m]);
            }
        }
        if (tri.size() > size_t(3 * kMaxTris)) throw std::logic_error("too many triangles");
        t.ntris[ci] = uint8_t(tri.size() / 3);
        std::fill(t.tris[ci], t.tris[ci] + 3 * kMaxTris + 1, int8_t(-1));
        for (size_t k = 0; k < tri.size(); ++k) t.tris[ci][k] = int8_t(tri[k]);
    }
    // the winding that turns case 1's triangle (corner 0 below iso) towards
    // corner 0, applied to every case
    {
        const int8_t* q = t.tris[1];
        double P[3][3];
        for (int m = 0; m < 3; ++m)
            for (int d = 0; d < 3; ++d) P[m][d] = 0.5 * (kCorner[kEdge[q[m]][0]][d] + kCorner[kEdge[q[m]][1]][d]);
        const double u[3] = {P[1][0] - P[0][0], P[1][1] - P[0][1], P[1][2] - P[0][2]};
        const double w[3] = {P[2][0] - P[0][0], P[2][1] - P[0][1], P[2][2] - P[0][2]};
        const double n[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
        // corner 0 is the origin; the triangle lies towards +x, +y, +z of it
        if (n[0] + n[1] + n[2] > 0)
            for (int ci = 0; ci < 256; ++ci)
                for (int k = 0; k < 3 * t.ntris[ci]; k += 3) std::swap(t.tris[ci][k + 1], t.tris[ci][k + 2]);
    }
    return t;
}

}  // namespace

const Table& table() {
    static const Table t = build();
    return t;
}

}  // namespace mctable
