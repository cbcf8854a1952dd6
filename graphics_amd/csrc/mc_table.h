// Marching-cubes case table, generated at load time (mc_table.cpp).
//
// Cube corners and edges follow the numbering of Source/marching_cubes.cpp
// (corner k at offset (dx, dy, dz)): 0 (0,0,0), 1 (0,0,1), 2 (0,1,1),
// 3 (0,1,0), 4 (1,0,0), 5 (1,0,1), 6 (1,1,1), 7 (1,1,0); edges 0 (0-1),
// 1 (1-2), 2 (2-3), 3 (3-0), 4 (4-5), 5 (5-6), 6 (6-7), 7 (7-4), 8 (0-4),
// 9 (1-5), 10 (2-6), 11 (3-7).  Bit k of a case is set when corner k's
// value is below the iso level.
#pragma once

#include <cstdint>

namespace mctable {

constexpr int kMaxTris = 5;

struct Table {
    uint16_t edges[256];                 // crossed edges of each case
    int8_t tris[256][3 * kMaxTris + 1];  // edge triples, -1 terminated
    uint8_t ntris[256];
};

// The table, built once: per face the crossings pair up around each run of
// below-iso corners (an ambiguous face separates its two below corners, a
// choice that depends on the face alone, so neighbouring cubes agree and the
// surface is closed); segments chain across the cube's faces into loops,
// each fanned into triangles whose normals point from the above-iso corners
// to the below-iso ones (out of the fluid for a density-like field).
const Table& table();

// corner offsets and edge end points
extern const int kCorner[8][3];
extern const int kEdge[12][2];

}  // namespace mctable
