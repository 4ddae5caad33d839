// Host-side geometry structures of libbem_hip.
#pragma once

#include <algorithm>
#include <cstdint>
#include <utility>
#include <vector>

#include "bem_common.h"

namespace bemgeo {

struct BodyMesh {
  int nv = 0, nt = 0, ne = 0;
  std::vector<double> xyz;      // nv x 3
  std::vector<int32_t> tri;     // nt x 3
  std::vector<int32_t> edges;   // ne x 4 {p1, p2, t1, t2}
  std::vector<std::pair<int, int>> close;  // (a, b) triangle pairs sharing a vertex
};

struct Tree {
  int T = 0, E = 0;
  std::vector<TriGeom> tg;
  std::vector<RwgRec> rw;
  std::vector<std::pair<int, int>> close;
  std::vector<int> tri_off, rwg_off;
};

// Reference edge table; returns the edge count or -1 on a bad vertex id.
int build_edges(int nv, int nt, const int32_t *tri, std::vector<int32_t> &edges);
void build_close_pairs(BodyMesh &B);
// Union of bodies (body k translated by shifts[3k..3k+2]) in one index space.
void build_tree(const std::vector<const BodyMesh *> &bodies, const double *shifts, Tree &out);
// Flux surface of one body (translated by shift, its RWGs numbered from
// rwg_off): per triangle its vertices in the triangle's own order and the
// RWGs of its edges e12, e23, e31 with the signed edge length.  Returns -1
// when an edge of a triangle has no RWG (an open surface).
int flux_table(const BodyMesh &B, const double *shift, int rwg_off, std::vector<FluxTri> &out);
// The reference's RWGTree order of one body (rwgtree.cpp): order[0..n) are
// edge indices (build_edges order) leaf by leaf, leaf_sizes the RWG counts of
// the leaves.  order needs room for every edge.  Returns 0, -1 on a bad
// vertex id, -2 when the surface is not closed, -3 when the plane search
// does not converge.
int rwg_tree_order(int nv, const double *xyz, int nt, const int32_t *tri, int min_size, double min_radius,
                   int32_t *order, int &n, std::vector<int32_t> &leaf_sizes);

}  // namespace bemgeo
