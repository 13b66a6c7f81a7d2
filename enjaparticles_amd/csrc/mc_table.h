// Marching-cubes case table, generated at run time (no stored table).
//
// Corner k of a cube sits at offset (k & 1, (k >> 1) & 1, (k >> 2) & 1).
// Edge e joins corners kMcEdge[e][0] < kMcEdge[e][1], which differ in one
// bit: edges 0-3 run along x, 4-7 along y, 8-11 along z.  Case m has bit k
// set when corner k lies above the iso level.
//
// For every face, each maximal run of consecutive above-iso corners (in the
// face's corner cycle) is cut off by one segment joining the two crossing
// edges that bound the run; on a face whose above-iso corners are diagonal
// this separates them.  The rule depends only on the face's four values, so
// the two cubes sharing a face cut it the same way and the surface is closed.
// Each segment is directed with its run on the left as seen from outside the
// cube, so the two cubes sharing a face traverse its segments in opposite
// directions and the surface is consistently oriented.  The segments chain
// into loops of crossing edges, starting at their lowest edge; each loop is
// triangulated without a diagonal between two vertices on one face (which
// would lie in the face, on the neighbour's surface), wound so normals
// (right-hand rule) point from the above-iso corners to the others.
#pragma once

#include <cstdint>

namespace rtps {

constexpr int kMcMaxTris = 5;   // triangles per case (the builder fails if a case needs more)

struct McTable {
    int8_t tri[256][16];   // edge ids, three per triangle, -1 padded
    uint8_t ntri[256];
};

constexpr int kMcEdge[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7},    // x
                                {0, 2}, {1, 3}, {4, 6}, {5, 7},    // y
                                {0, 4}, {1, 5}, {2, 6}, {3, 7}};   // z

// Builds the table; false if a case needs more than kMcMaxTris triangles.
bool build_mc_table(McTable& t);
// The process-wide table (built on first use).
const McTable& mc_table();

}  // namespace rtps
