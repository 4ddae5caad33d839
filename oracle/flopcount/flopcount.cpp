// Operation counts of the reference's assembly per triangle-pair class,
// from the oracle's own source compiled onto counting scalars
// (counting.hpp; build recipe: oracle/Makefile target `flopcount`).  Test
// infrastructure: the per-pair figures the benchmark's assembly rates use.
//
// Input (stdin): nv nt, nv lines "x y z", nt lines "a b c" (one closed body).
// Output (stdout, JSON), for the pairs of that body's self block:
//   per class (close, far6, far10, far15, far21, far36): pairs, and the
//   counted operations of lk_triangle (classification and integrals,
//   src/Impedance.cpp:276-364) summed over them, as a total and split into
//   add / mul / div / sqrt / exp / log / atan2 / sincos;
//   close pairs also by kind (self, edge neighbour, vertex neighbour);
//   scatter: the operations of the scatter into Z per pair
//   (src/Impedance.cpp:365-384), from one test row of oracle_zts.
// With a file name as argument, the Zts of that scatter pass (test row 0,
// eps = 3 + 0.4i, mu = 1, omega = 0.5; (2 rwgs) x (2 rwgs) complex128,
// column-major) is written there, so a test can check that the counting
// build computes the oracle's numbers bit for bit.
// The far tiers other than the 6-point rule are never chosen (the
// thresholds decrease); one pair at each is counted anyway, with the rule
// forced, so that every class has a figure.
#include "counting.hpp"

#include <cstdio>
#include <string>
#include <vector>

#include "oracle_counted.inc"

static uint64_t total(const OpCount &c) {
  return c.add + c.mul + c.div + c.sqrt + c.exp + c.log + c.atan2 + c.sincos;
}
static OpCount diff(const OpCount &a, const OpCount &b) {
  return {a.add - b.add, a.mul - b.mul, a.div - b.div, a.sqrt - b.sqrt,
          a.exp - b.exp, a.log - b.log, a.atan2 - b.atan2, a.sincos - b.sincos};
}
static void acc(OpCount &s, const OpCount &d) {
  s.add += d.add; s.mul += d.mul; s.div += d.div; s.sqrt += d.sqrt;
  s.exp += d.exp; s.log += d.log; s.atan2 += d.atan2; s.sincos += d.sincos;
}
static std::string js(const char *name, uint64_t n, const OpCount &c) {
  char buf[512];
  std::snprintf(buf, sizeof buf,
                "\"%s\": {\"pairs\": %llu, \"ops\": %llu, \"add\": %llu, \"mul\": %llu, \"div\": %llu, "
                "\"sqrt\": %llu, \"exp\": %llu, \"log\": %llu, \"atan2\": %llu, \"sincos\": %llu}",
                name, (unsigned long long)n, (unsigned long long)total(c), (unsigned long long)c.add,
                (unsigned long long)c.mul, (unsigned long long)c.div, (unsigned long long)c.sqrt,
                (unsigned long long)c.exp, (unsigned long long)c.log, (unsigned long long)c.atan2,
                (unsigned long long)c.sincos);
  return buf;
}

int main(int argc, char **argv) {
  int nv, nt;
  if (std::scanf("%d %d", &nv, &nt) != 2 || nv <= 0 || nt <= 0) return 2;
  std::vector<F> xyz(3 * (size_t)nv);
  for (auto &x : xyz) {
    double d;
    if (std::scanf("%lf", &d) != 1) return 2;
    x = F(d);
  }
  std::vector<int> tri(3 * (size_t)nt);
  for (auto &t : tri)
    if (std::scanf("%d", &t) != 1) return 2;
  std::vector<int> edges(4 * 3 * (size_t)nt);
  const int ne = oracle_mesh_edges(nv, nt, tri.data(), 3 * nt, edges.data());
  if (ne * 2 != 3 * nt) return 3;

  const double omega = 0.5;   // any lossy medium: the counts do not depend on it
  const OpCount t0 = g_ops;
  tree T;
  tree_build(&T, nv, xyz.data(), nt, tri.data(), ne, edges.data());
  green g;
  green_init(&g, F(3.0) + I * F(0.4), F(1.0) + I * F(0.0), F(omega));
  const OpCount setup = diff(g_ops, t0);

  // every pair of the self block, by class; close pairs also by kind
  OpCount cls_ops[6] = {}, kind_ops[3] = {};
  uint64_t cls_n[6] = {}, kind_n[3] = {};
  std::vector<CF> L(9), K(9);
  for (int ta = 0; ta < nt; ta++)
    for (int tb = 0; tb < nt; tb++) {
      const int c = oracle_pair_class(T.xyz, T.tri + 3 * ta, T.tri + 3 * tb, &T.temax[ta], &T.temax[tb], 1);
      if (c < 0 || c > 5) return 4;
      const OpCount b = g_ops;
      lk_triangle(&g, &T, ta, tb, 1, L.data(), K.data());
      const OpCount d = diff(g_ops, b);
      acc(cls_ops[c], d);
      cls_n[c]++;
      if (c == 0) {
        const int kind = ta == tb ? 0 : (shared_points(T.tri + 3 * ta, T.tri + 3 * tb) == 2 ? 1 : 2);
        acc(kind_ops[kind], d);
        kind_n[kind]++;
      }
    }
  // the tiers never chosen: one far pair with each rule forced
  {
    int ta = -1, tb = -1;
    for (int a = 0; a < nt && ta < 0; a++)
      for (int b = 0; b < nt; b++)
        if (oracle_pair_class(T.xyz, T.tri + 3 * a, T.tri + 3 * b, &T.temax[a], &T.temax[b], 1) == 1) {
          ta = a;
          tb = b;
          break;
        }
    static const int offs[6] = {0, BEM_QUAD_OFF_6, BEM_QUAD_OFF_10, BEM_QUAD_OFF_15, BEM_QUAD_OFF_21,
                                BEM_QUAD_OFF_36};
    static const int npts[6] = {0, 6, 10, 15, 21, 36};
    for (int c = 2; c <= 5; c++) {
      if (cls_n[c] || ta < 0) continue;
      const OpCount b = g_ops;
      (void)oracle_pair_class(T.xyz, T.tri + 3 * ta, T.tri + 3 * tb, &T.temax[ta], &T.temax[tb], 1);
      for (int i = 0; i < 9; i++) L[i] = K[i] = 0;
      lk_far(&g, &T.tlist[ta], &T.tlist[tb], offs[c], npts[c], L.data(), K.data());
      cls_ops[c] = diff(g_ops, b);
      cls_n[c] = 1;
    }
  }
  // scatter per pair: one test row of oracle_zts less its pairs, its set-up
  OpCount row_pairs = {};
  for (int tb = 0; tb < nt; tb++) {
    const OpCount b = g_ops;
    lk_triangle(&g, &T, 0, tb, 1, L.data(), K.data());
    acc(row_pairs, diff(g_ops, b));
  }
  std::vector<CF> Z((size_t)2 * ne * 2 * ne);
  std::vector<int> body(nt, 0);
  const OpCount b = g_ops;
  oracle_zts(nv, (const F *)xyz.data(), nt, tri.data(), body.data(), ne, edges.data(), F(3.0), F(0.4), F(1.0),
             F(0.0), F(omega), 0, 1, (F *)Z.data(), 2 * ne);
  OpCount row = diff(g_ops, b);
  row = diff(row, row_pairs);
  row = diff(row, setup);
  OpCount scatter = row;   // over nt pairs of the row
  if (argc > 1) {
    FILE *f = std::fopen(argv[1], "wb");
    if (!f) return 5;
    std::fwrite(Z.data(), sizeof(CF), Z.size(), f);
    std::fclose(f);
  }

  static const char *names[6] = {"close", "far6", "far10", "far15", "far21", "far36"};
  std::printf("{\n \"triangles\": %d, \"rwgs\": %d,\n", nt, ne);
  for (int c = 0; c < 6; c++) std::printf(" %s,\n", js(names[c], cls_n[c], cls_ops[c]).c_str());
  std::printf(" %s,\n", js("close_self", kind_n[0], kind_ops[0]).c_str());
  std::printf(" %s,\n", js("close_edge", kind_n[1], kind_ops[1]).c_str());
  std::printf(" %s,\n", js("close_vertex", kind_n[2], kind_ops[2]).c_str());
  std::printf(" %s\n}\n", js("scatter", (uint64_t)nt, scatter).c_str());
  tree_free(&T);
  return 0;
}
