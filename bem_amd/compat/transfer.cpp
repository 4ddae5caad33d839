// transfer: the reference's single-frequency driver (src/transfer.cpp:3-29)
// with its command line, over parseShapes and CalcHeatTransfer of the compat
// layer.  The reference computes P and discards it; this driver prints it as
// "eV P" (%.15g %.14g).
//
// usage: transfer <x3d body 1> <x3d body 2> <eV> <a (m)> <T1 (K)> <T2 (K)> <gap x y z>
#include <cstdio>
#include <cstdlib>

#include "bem_compat.h"

int main(int argc, char **argv) {
  if (argc != 10) {
    std::fprintf(stderr,
                 "Usage: transfer <x3d file for body 1> <x3d file for body 2> <eV> <scale factor applied to x3d "
                 "files (in meters)> <T_1 (in K)> <T_2 (in K)> <gap x y z>\n");
    return 2;
  }
  const Real eV = std::atof(argv[3]), a = std::atof(argv[4]), T1 = std::atof(argv[5]), T2 = std::atof(argv[6]);
  const Vector3 gap(std::atof(argv[7]), std::atof(argv[8]), std::atof(argv[9]));
  try {
    Options options;
    RWGTree *top = nullptr, *s = nullptr, *t = nullptr;
    parseShapes(top, s, t, argv[1], argv[2], gap, options);
    const Real P = CalcHeatTransfer(a, eV, T1, T2, top, s, t, options);
    std::printf("%.15g %.14g\n", eV, P);
    delete top;
    delete s;
    delete t;
    return 0;
  } catch (const bem::Error &e) {
    std::fprintf(stderr, "transfer: %s\n", e.what());
    return 1;
  }
}
