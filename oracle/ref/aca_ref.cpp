// aca_ref.cpp -- driver around the reference's own adaptive cross
// approximation, aca(DenseMatrix&, vector<int>&, vector<int>&, R eps2)
// (src/util/DenseMatrix.h:213-279), compiled together with the reference
// sources (see oracle/ref/Makefile).  Test infrastructure only: it pins the
// oracle's restatement (oracle/oracle.py: aca) and never ships.
//
// Input (binary, stdin): int32 m, int32 n, float64 eps2, then m x n
// complex128 column-major.  Output (binary, file argv[1]): int32 count, then
// count int32 row pivots, then count int32 column pivots.
#include <cstdio>
#include <vector>
#include "DenseMatrix.h"
#include "ReferenceCountingPointer.h"

int main(int argc, char **argv) {
  if (argc != 2) return 2;
  int mn[2];
  double eps2;
  if (fread(mn, sizeof(int), 2, stdin) != 2 || fread(&eps2, sizeof(double), 1, stdin) != 1) return 3;
  const int m = mn[0], n = mn[1];
  initReferenceCountingPointerLock();
  ComplexDenseMatrix A(m, n);
  std::vector<double> col(2 * (size_t)m);
  for (int j = 0; j < n; j++) {
    if (fread(col.data(), sizeof(double), 2 * (size_t)m, stdin) != 2 * (size_t)m) return 3;
    for (int i = 0; i < m; i++) A(i, j) = Complex(col[2 * i], col[2 * i + 1]);
  }
  std::vector<int> r, c;
  aca(A, r, c, eps2);
  FILE *out = fopen(argv[1], "wb");
  if (!out) return 4;
  int cnt = (int)r.size();
  fwrite(&cnt, sizeof(int), 1, out);
  fwrite(r.data(), sizeof(int), r.size(), out);
  fwrite(c.data(), sizeof(int), c.size(), out);
  fclose(out);
  return 0;
}
