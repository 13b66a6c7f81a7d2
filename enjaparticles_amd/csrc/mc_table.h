// Marching-cubes case table of SPH::postProcess: the standard table of
// Lorensen & Cline's marching cubes in the layout traversehp.cl:74 holds it
// (Bourke's tabulation, 16 entries per case), with the reference's corner and
// edge numbering, so the same field gives the reference's triangles.
//
// Corner k of a cube (bit k of the case, set when the node value exceeds the
// iso level, classifycubes.cl:44-55) sits at lattice offset kMcCorner[k].
// Edge e runs from kMcEdgeEnds[e][0..2] to kMcEdgeEnds[e][3..5]
// (traversehp.cl:35-72); a vertex is interpolated from the first end toward
// the second.  Case c lists its triangles' edges, three per triangle, as the
// 4-bit fields of kMcCaseWords[c] from the lowest up, ended by 0xF: one 8-byte
// word per case (2 KB in all), a single scalar load per cube on the GPU.
#pragma once

#include <cstdint>

namespace rtps {

constexpr int kMcMaxTris = 5;   // triangles per case

struct McTable {
    int8_t tri[256][16];   // edge ids, three per triangle, -1 padded
    uint8_t ntri[256];
};

constexpr int8_t kMcCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1},
                                    {0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}};

constexpr int8_t kMcEdgeEnds[12][6] = {
    {0, 0, 0, 1, 0, 0}, {1, 0, 0, 1, 0, 1}, {1, 0, 1, 0, 0, 1}, {0, 0, 1, 0, 0, 0},
    {0, 1, 0, 1, 1, 0}, {1, 1, 0, 1, 1, 1}, {1, 1, 1, 0, 1, 1}, {0, 1, 1, 0, 1, 0},
    {0, 0, 0, 0, 1, 0}, {1, 0, 0, 1, 1, 0}, {1, 0, 1, 1, 1, 1}, {0, 0, 1, 0, 1, 1}};

extern const uint64_t kMcCaseWords[256];

// The unpacked table (tri / ntri), for the device upload and the C ABI.
const McTable& mc_table();

}  // namespace rtps
