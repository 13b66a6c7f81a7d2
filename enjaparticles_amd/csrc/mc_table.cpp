// Marching-cubes case table builder (see mc_table.h for the rule).

#include "mc_table.h"

#include <cstring>
#include <mutex>
#include <utility>
#include <vector>

namespace rtps {

namespace {

// Corner cycles of the six faces (adjacent corners share an edge).
constexpr int kFace[6][4] = {{0, 2, 6, 4}, {1, 3, 7, 5},    // x = 0, x = 1
                             {0, 1, 5, 4}, {2, 3, 7, 6},    // y = 0, y = 1
                             {0, 1, 3, 2}, {4, 5, 7, 6}};   // z = 0, z = 1

int edge_of(int a, int b) {
    if (a > b) std::swap(a, b);
    for (int e = 0; e < 12; ++e)
        if (kMcEdge[e][0] == a && kMcEdge[e][1] == b) return e;
    return -1;
}

void corner_pos(int k, double p[3]) {
    p[0] = k & 1;
    p[1] = (k >> 1) & 1;
    p[2] = (k >> 2) & 1;
}

// Two cube edges on a common face (each edge lies on the two faces fixing
// one of the coordinates its corners share).
bool share_face(int e1, int e2) {
    const int a1 = kMcEdge[e1][0], b1 = kMcEdge[e1][1], a2 = kMcEdge[e2][0], b2 = kMcEdge[e2][1];
    for (int axis = 0; axis < 3; ++axis) {
        const int bit = 1 << axis;
        if ((a1 & bit) != (b1 & bit) || (a2 & bit) != (b2 & bit)) continue;   // axis not fixed on both
        if ((a1 & bit) == (a2 & bit)) return true;
    }
    return false;
}

// Triangulates the loop polygon poly (vertices in loop order) without a
// diagonal between two vertices on one cube face: such a diagonal would lie in
// the face, where the neighbouring cube has its own surface.  Side (p0, p_last)
// takes the first apex j (in loop order) whose diagonals are allowed and whose
// two sub-polygons triangulate; triangles (a, b, c) in loop order are stored
// as (a, c, b).  False if no such triangulation exists.
bool triangulate(const std::vector<int>& poly, std::vector<int>& out) {
    const size_t k = poly.size();
    if (k < 3) return true;
    if (k == 3) {
        out.insert(out.end(), {poly[0], poly[2], poly[1]});
        return true;
    }
    for (size_t j = 1; j + 1 < k; ++j) {
        if (j > 1 && share_face(poly[0], poly[j])) continue;
        if (j + 2 < k && share_face(poly[j], poly[k - 1])) continue;
        std::vector<int> left(poly.begin(), poly.begin() + j + 1), right(poly.begin() + j, poly.end());
        std::vector<int> tris;
        if (!triangulate(left, tris) || !triangulate(right, tris)) continue;
        tris.insert(tris.end(), {poly[0], poly[k - 1], poly[j]});
        out.insert(out.end(), tris.begin(), tris.end());
        return true;
    }
    return false;
}

}  // namespace

bool build_mc_table(McTable& t) {
    std::memset(t.tri, -1, sizeof(t.tri));
    std::memset(t.ntri, 0, sizeof(t.ntri));
    // face corner cycles, counter-clockwise seen from outside the cube
    int face[6][4];
    for (int f = 0; f < 6; ++f) {
        double p[4][3], c[3] = {0, 0, 0};
        for (int i = 0; i < 4; ++i) {
            face[f][i] = kFace[f][i];
            corner_pos(kFace[f][i], p[i]);
            for (int k = 0; k < 3; ++k) c[k] += 0.25 * p[i][k];
        }
        const double u[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
        const double v[3] = {p[2][0] - p[1][0], p[2][1] - p[1][1], p[2][2] - p[1][2]};
        const double n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        const double out = n[0] * (c[0] - 0.5) + n[1] * (c[1] - 0.5) + n[2] * (c[2] - 0.5);
        if (out < 0) std::swap(face[f][1], face[f][3]);
    }
    for (int m = 0; m < 256; ++m) {
        auto above = [&](int k) { return (m >> k) & 1; };
        // directed segments: on each face, from the edge leaving a run of
        // above-iso corners to the edge entering it (the run on the left seen
        // from outside); a crossing edge lies on two faces, leaving the run
        // on one and entering one on the other, so next[] is a permutation
        int next[12];
        for (int e = 0; e < 12; ++e) next[e] = -1;
        for (int f = 0; f < 6; ++f) {
            const int* c = face[f];
            for (int i = 0; i < 4; ++i) {
                if (!above(c[i]) || above(c[(i + 3) % 4])) continue;   // i starts a run
                int j = i;
                int len = 0;
                while (above(c[(j + 1) % 4]) && len < 4) { j = (j + 1) % 4; ++len; }
                if (len >= 3) break;   // whole face above
                const int e_in = edge_of(c[(i + 3) % 4], c[i]);
                const int e_out = edge_of(c[j], c[(j + 1) % 4]);
                if (next[e_out] != -1) return false;
                next[e_out] = e_in;
            }
        }
        bool seen[12] = {false};
        int nt = 0;
        for (int e0 = 0; e0 < 12; ++e0) {
            if (next[e0] < 0 || seen[e0]) continue;
            std::vector<int> loop;
            int cur = e0;
            do {
                loop.push_back(cur);
                seen[cur] = true;
                cur = next[cur];
            } while (cur >= 0 && cur != e0 && loop.size() <= 12);
            if (cur != e0) return false;
            // wound so the right-hand normal points from the above-iso
            // corners to the others
            std::vector<int> tris;
            if (!triangulate(loop, tris)) return false;
            for (size_t q = 0; q < tris.size(); q += 3) {
                if (nt >= kMcMaxTris) return false;
                for (int v = 0; v < 3; ++v) t.tri[m][3 * nt + v] = (int8_t)tris[q + v];
                ++nt;
            }
        }
        t.ntri[m] = (uint8_t)nt;
    }
    return true;
}

const McTable& mc_table() {
    static McTable table;
    static std::once_flag once;
    static bool ok = false;
    std::call_once(once, [] { ok = build_mc_table(table); });
    (void)ok;   // the builder succeeds (every case needs at most 5 triangles; checked by the table test)
    return table;
}

}  // namespace rtps
