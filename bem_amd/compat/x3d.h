// Shape files in X3D (XML), read as the reference's ShapeParser does
// (src/ShapeParser.cpp:79-147), with a small XML reader of this project's own
// in place of xerces.
//
// What is read:
//  - every <Coordinate point="x y z, x y z, ..."> appends its points, each
//    mapped through the Transform stack in effect (below);
//  - every <IndexedFaceSet coordIndex="a b c -1, ..."> appends one triangle
//    (a, b, c) per comma-separated group of at least four integers;
//  - <Dielectric material="..." key="value" ...> names the body's medium;
//    its other attributes are the medium's parameters.
// Indices of a second IndexedFaceSet are not offset by the points before it
// (the reference appends both lists unchanged).
//
// Transform stack: every element pushes a transform when it opens and pops
// one when it closes; <Transform> carries center, rotation, scale,
// scaleOrientation and translation (others push the identity).  A point is
// mapped by each transform of the stack from the outermost to the innermost,
// each as P' = T + C + R SR S (-SR)(P - C), the rotation (x y z phi) applied
// with the matrix of src/util/Rotation.cpp:71-78 (the transpose of the
// right-handed rotation by phi about the unnormalised axis).
//
// Numbers are read as the reference reads them: space-separated, each group
// parsed from its first character onward (src/util/Vector3.h:30-46,
// src/FaceIndex.cpp:4-13); a group with too few numbers is skipped.  XML
// attribute values have tabs and line breaks turned into spaces and the five
// predefined entities and character references decoded, as an XML parser
// delivers them.
#pragma once

#include <string>
#include <utility>
#include <vector>

struct X3DShape {
  std::vector<double> coord;  // 3 per point, transformed
  std::vector<int> face;      // 3 per triangle, indices into coord
  bool dielectric = false;    // a <Dielectric> element was seen
  std::string material;       // its material attribute
  std::vector<std::pair<std::string, std::string>> params;  // its other attributes
};

// Appends the shapes of an X3D document to shape.  Throws bem::Error
// (BEM_E_INVALID) on an unreadable file or malformed XML.
void read_x3d(const char *path, X3DShape &shape);
void read_x3d_text(const std::string &text, X3DShape &shape);

// C entry points of libbem_compat.so for callers outside C++ (bindings,
// tests); neither needs a device.  Both return 0 or BEM_E_INVALID.
extern "C" {
// Shapes of one file: the counts go to *nv / *nt; points (nv x 3) and
// triangles (nt x 3) are written when xyz / tri are non-null, in which case
// *nv / *nt hold their capacity on entry; the <Dielectric> material name
// (empty without one) to material[0..matcap).
int bemc_x3d_read(const char *path, double *xyz, int *nv, int *tri, int *nt, char *material, int matcap);
// eps and mu (re, im, re, im) of a medium of the material library at eV and
// temperature T; params as "key=value" pairs separated by ';' (or null).
int bemc_dielectric(const char *name, const char *params, double eV, double T, double *epsmu4);
}
