// transfer_sweep: a transfer-style driver written against the reference's
// API shape (RWGTree, Options, FrequencySweep, CalcHeatTransfer; compare
// src/transfer.cpp:3-29), linked against libbem_hip through bem_compat.
//
// usage: transfer_sweep <body1 mesh> <body2 mesh> <freqs file> <scale m> <T1> <T2> <gx> <gy> <gz>
// mesh file: "nv nt" then nv lines "x y z" then nt lines "a b c" (0-based).
// Both bodies use the SiC-like Lorentz permittivity of the benchmark
// configurations.  Prints one line per frequency: "a gx gy gz eV P".
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bem_compat.h"

class LorentzSiC : public Dielectric {
 public:
  Complex eps(Real w) {
    Complex i(0, 1);
    return 6.7 * (w * w - 0.120 * 0.120 + i * 0.00059 * w) / (w * w - 0.0983 * 0.0983 + i * 0.00059 * w);
  }
  Complex mu(Real) { return 1.0; }
};

static bool read_mesh(const char *fn, std::vector<double> &xyz, std::vector<int> &tri) {
  FILE *f = std::fopen(fn, "r");
  if (!f) return false;
  int nv, nt;
  if (std::fscanf(f, "%d %d", &nv, &nt) != 2) return false;
  xyz.resize(3 * (size_t)nv);
  tri.resize(3 * (size_t)nt);
  for (auto &x : xyz)
    if (std::fscanf(f, "%lf", &x) != 1) return false;
  for (auto &t : tri)
    if (std::fscanf(f, "%d", &t) != 1) return false;
  std::fclose(f);
  return true;
}

int main(int argc, char **argv) {
  if (argc != 10) {
    std::fprintf(stderr, "usage: transfer_sweep <mesh1> <mesh2> <freqs> <scale> <T1> <T2> <gx> <gy> <gz>\n");
    return 2;
  }
  std::vector<double> x1, x2;
  std::vector<int> t1, t2;
  if (!read_mesh(argv[1], x1, t1) || !read_mesh(argv[2], x2, t2)) {
    std::fprintf(stderr, "cannot read meshes\n");
    return 2;
  }
  FrequencySweep sweep(argv[3]);
  Real a = std::atof(argv[4]), T1 = std::atof(argv[5]), T2 = std::atof(argv[6]);
  Real gap[3] = {std::atof(argv[7]), std::atof(argv[8]), std::atof(argv[9])};
  LorentzSiC diel1, diel2;
  Options options;
  RWGTree s(x1, t1, &diel1, gap);
  RWGTree t(x2, t2, &diel2);
  RWGTree top(&s, &t);
  for (Real eV : sweep.f) {
    Real P = CalcHeatTransfer(a, eV, T1, T2, &top, &s, &t, options);
    std::printf("%.15g %.15g %.15g %.15g %.15g %.14g\n", a, gap[0], gap[1], gap[2], eV, P);
  }
  top.sub[0] = top.sub[1] = nullptr;
  return 0;
}
