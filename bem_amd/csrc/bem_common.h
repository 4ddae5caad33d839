// Shared definitions for the HIP kernels and the host runtime of libbem_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define BEM_ONETHIRD 0.33333333333333333333333333333333
#define BEM_ONEOVER4PI 0.079577471545947667884441881686257

// Pair classes of the reference's LKTriangle (src/Impedance.cpp:285-361).
// The far-tier thresholds decrease (12, 8, 4, 2 in src/Options.cpp:23-26), so
// a pair failing d*12 > emax^2 fails the 8/4/2 tests as well: only the 6-point
// and the 36-point far rules are reachable.
enum { BEM_CLASS_CLOSE = 0, BEM_CLASS_G6 = 1, BEM_CLASS_G36 = 5 };

// Number of tile planes per triangle pair: 9 complex L + 9 complex K entries,
// stored as separate real planes (structure of arrays) so that neighbouring
// pairs write neighbouring addresses.
#define BEM_TILE_PLANES 36

// Triangle record of one "tree" (a set of bodies sharing one RWG index
// space).  All coordinates are global (the body translation applied), as the
// reference applies the gap to the body-1 vertices before meshing
// (src/HeatTransfer.cpp:67-69).
struct TriGeom {
  // frame half-RWG (the first half pushed for this triangle, src/Impedance.cpp:397-401)
  double p0[3], p1[3], p2[3];
  double v01[3], v12[3], v20[3], v02[3];
  double n[3], m01[3], m12[3], m20[3];
  double normv01i, normv12i, normv20i, Ai, elen;
  // the up-to-3 half-RWGs on this triangle, in push order
  double hp0[3][3];
  // centroid and longest side (tier selection, src/Impedance.cpp:329-335)
  double centre[3];
  double emax;
  int32_t nh;
  int32_t body;
  int32_t vid[3];   // global vertex ids (shared-vertex test)
  int32_t pad;
};

// Device-side medium constants (src/GreensFunction.cpp:17-30).
struct Medium {
  double ksq_re, ksq_im;    // k^2 = eps mu omega^2
  double ik_re, ik_im;      // i k
  double k2_re, k2_im;      // -k^2/2
  double Zeps_re, Zeps_im;  // 1/(i omega eps)
  double Zmu_re, Zmu_im;    // 1/(i omega mu)
};

// One RWG of a tree: its two half-RWGs as (triangle, slot in that triangle's
// half list) and the edge length.
struct RwgRec {
  int32_t tp, sp, tn, sn;
  double len;
};

// One point dipole: moment j (real) at p in the medium md of its region; it
// drives the triangles [t_lo, t_hi) of the tree (every triangle for the
// exterior, one body's for an interior dipole).
struct DipoleSrc {
  double p[3], j[3];
  Medium md;
  int t_lo, t_hi, external, pad;
};

// One triangle of a flux surface: its vertices P1, P2, P3 in the triangle's
// own order, and for the edges e12, e23, e31 the RWG index and the signed
// edge length (+len when the triangle is the RWG's positive half).
struct FluxTri {
  double P[3][3];
  double c[3];
  int32_t e[3];
  int32_t pad;
};

#define HIP_CHECK_RET(expr)                                   \
  do {                                                        \
    hipError_t _e = (expr);                                   \
    if (_e != hipSuccess) return (int)_e;                     \
  } while (0)
