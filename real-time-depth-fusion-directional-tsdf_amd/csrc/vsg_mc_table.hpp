// Marching-cubes triangle table, generated from the cube's geometry.
//
// Cube corners: corner c is at (c & 1, (c >> 1) & 1, (c >> 2) & 1).  Edges join two
// corners differing in one coordinate and run from the lower to the higher one
// (kMcEdge), so a cube and its neighbour compute the vertex of a shared edge from the
// same endpoints in the same order.  Case k: bit c set when corner c is inside
// (sdf < iso).  For each case the iso-surface's boundary on each face is a set of
// segments between the face's sign-changing edges; a face with all four edges
// changing (inside corners on one diagonal) cuts off each inside corner, a rule that
// depends on the face alone, so the two cubes sharing the face agree.  Every
// sign-changing edge lies on two faces, so the segments link into closed loops; each
// loop is fanned into triangles oriented so that their normals (v1 - v0) x (v2 - v0)
// point towards the outside corners.  Row k lists the triangles' edges, three per
// triangle, then -1 (at most 10 triangles).
#pragma once

#include <cstdint>
#include <cstring>

namespace vsg {

constexpr int kMcRow = 32;

// Edge e joins corners kMcEdge[e][0] < kMcEdge[e][1] (differing in one bit).
constexpr int kMcEdge[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7},   // along x
                                {0, 2}, {1, 3}, {4, 6}, {5, 7},   // along y
                                {0, 4}, {1, 5}, {2, 6}, {3, 7}};  // along z

// The corners of edge e follow from its number: axis e >> 2, and the corner bits of the
// two other axes e & 3 with a zero inserted at the axis (the GPU kernels decode it so).
constexpr bool mc_edge_numbering_ok() {
  for (int e = 0; e < 12; ++e) {
    const int axis = e >> 2, j = e & 3;
    const int a = ((j >> axis) << (axis + 1)) | (j & ((1 << axis) - 1));
    if (a != kMcEdge[e][0] || (a | (1 << axis)) != kMcEdge[e][1]) return false;
  }
  return true;
}
static_assert(mc_edge_numbering_ok(), "edge numbering");

inline int mc_edge_of(int a, int b) {
  if (a > b) {
    const int t = a;
    a = b;
    b = t;
  }
  for (int e = 0; e < 12; ++e)
    if (kMcEdge[e][0] == a && kMcEdge[e][1] == b) return e;
  return -1;
}

// True when edges e1 and e2 lie on a common face of the cube (a fixed coordinate in
// common: an edge along axis a holds the other two coordinates fixed).
inline bool mc_share_face(int e1, int e2) {
  const int a1 = kMcEdge[e1][0], b1 = kMcEdge[e1][1];
  const int a2 = kMcEdge[e2][0], b2 = kMcEdge[e2][1];
  for (int d = 0; d < 3; ++d) {
    const int m = 1 << d;
    if ((a1 & m) == (b1 & m) && (a2 & m) == (b2 & m) && (a1 & m) == (a2 & m)) return true;
  }
  return false;
}

// Triangulates the polygon loop[0..n) (loop order kept) so that no diagonal joins two
// edges on a common cube face: such a chord would lie on the face, where the
// neighbouring cube's surface meets it.  Writes n - 2 triangles to out (3 ints each);
// false when no such triangulation exists.
inline bool mc_triangulate(const int* loop, int n, int* out) {
  if (n < 3) return n == 0;
  if (n == 3) {
    out[0] = loop[0], out[1] = loop[1], out[2] = loop[2];
    return true;
  }
  // the triangle on the side (loop[0], loop[1]) has its apex at some loop[k]
  for (int k = 2; k < n; ++k) {
    if (k > 2 && mc_share_face(loop[1], loop[k])) continue;
    if (k < n - 1 && mc_share_face(loop[k], loop[0])) continue;
    int right[12], left[12], nr = 0, nl = 0;
    for (int i = 1; i <= k; ++i) right[nr++] = loop[i];   // loop[1..k]
    for (int i = k; i < n; ++i) left[nl++] = loop[i];     // loop[k..n) then loop[0]
    left[nl++] = loop[0];
    int* o = out;
    o[0] = loop[0], o[1] = loop[1], o[2] = loop[k];
    o += 3;
    if (nr >= 3) {
      if (!mc_triangulate(right, nr, o)) continue;
      o += 3 * (nr - 2);
    }
    if (nl >= 3 && !mc_triangulate(left, nl, o)) continue;
    return true;
  }
  return false;
}

// Fills tri[256][kMcRow].  Returns the largest number of triangles of a case, or -1
// when some loop has no triangulation free of chords on the cube's faces.
inline int build_mc_table(int8_t tri[256][kMcRow]) {
  // the six faces as corner cycles (each consecutive pair, and last-first, an edge)
  static const int face[6][4] = {{0, 2, 6, 4}, {1, 3, 7, 5}, {0, 1, 5, 4}, {2, 3, 7, 6}, {0, 1, 3, 2}, {4, 5, 7, 6}};
  int most = 0;
  for (int k = 0; k < 256; ++k) {
    std::memset(tri[k], -1, kMcRow);
    auto inside = [&](int c) { return ((k >> c) & 1) != 0; };
    // links: each sign-changing edge gets its two neighbours along the boundary
    int nb[12][2], nn[12];
    for (int e = 0; e < 12; ++e) nn[e] = 0;
    for (int f = 0; f < 6; ++f) {
      int fe[4];
      bool ch[4];
      int nch = 0;
      for (int i = 0; i < 4; ++i) {
        const int a = face[f][i], b = face[f][(i + 1) & 3];
        fe[i] = mc_edge_of(a, b);
        ch[i] = inside(a) != inside(b);
        nch += ch[i] ? 1 : 0;
      }
      auto link = [&](int e1, int e2) {
        nb[e1][nn[e1]++] = e2;
        nb[e2][nn[e2]++] = e1;
      };
      if (nch == 2) {
        int s[2], m = 0;
        for (int i = 0; i < 4; ++i)
          if (ch[i]) s[m++] = fe[i];
        link(s[0], s[1]);
      } else if (nch == 4) {
        // cut off each inside corner: corner i of the cycle lies between edges i - 1 and i
        for (int i = 0; i < 4; ++i)
          if (inside(face[f][i])) link(fe[(i + 3) & 3], fe[i]);
      }
    }
    // loops
    bool used[12] = {};
    int n = 0;
    for (int e0 = 0; e0 < 12; ++e0) {
      if (nn[e0] != 2 || used[e0]) continue;
      int loop[12], len = 0, prev = -1, cur = e0;
      do {
        loop[len++] = cur;
        used[cur] = true;
        const int nx = nb[cur][0] != prev ? nb[cur][0] : nb[cur][1];
        prev = cur;
        cur = nx;
      } while (cur != e0 && len < 12);
      // orientation: the loop's area normal (Newell) against the inside-to-outside
      // directions of its edges
      double P[12][3], out[3] = {0, 0, 0}, area[3] = {0, 0, 0};
      for (int i = 0; i < len; ++i) {
        const int a = kMcEdge[loop[i]][0], b = kMcEdge[loop[i]][1];
        for (int d = 0; d < 3; ++d) {
          const double pa = (a >> d) & 1, pb = (b >> d) & 1;
          P[i][d] = 0.5 * (pa + pb);
          out[d] += inside(a) ? (pb - pa) : (pa - pb);
        }
      }
      for (int i = 0; i < len; ++i) {
        const double* p = P[i];
        const double* q = P[(i + 1) % len];
        area[0] += (p[1] - q[1]) * (p[2] + q[2]);
        area[1] += (p[2] - q[2]) * (p[0] + q[0]);
        area[2] += (p[0] - q[0]) * (p[1] + q[1]);
      }
      const bool flip = area[0] * out[0] + area[1] * out[1] + area[2] * out[2] < 0.0;
      int tris[3 * 10];
      if (!mc_triangulate(loop, len, tris)) return -1;
      for (int i = 0; i + 2 < len; ++i) {
        tri[k][3 * n] = (int8_t)tris[3 * i];
        tri[k][3 * n + 1] = (int8_t)tris[3 * i + (flip ? 2 : 1)];
        tri[k][3 * n + 2] = (int8_t)tris[3 * i + (flip ? 1 : 2)];
        ++n;
      }
    }
    if (n > most) most = n;
    if (n > 10) return -1;
  }
  return most;
}

}  // namespace vsg
