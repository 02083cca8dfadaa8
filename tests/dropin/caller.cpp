// A caller compiled against the generated kernels' C signatures (as
// examples/tensor/helium_stabilization/benchmark/benchmark.cpp:11-14 is), linked against
// libpyprop_amd.so instead: it reads one multiply's arguments from a file, calls the
// drop-in symbol with the generated wrapper's argument list (generator.py:1428-1451) and
// writes dest.
//
// input (little endian): int32 kind (0 simp_simp_bandnh, 1 diag_bandnh, 2 simp_bandnh),
// int32 rank, int32 potential/source/dest extents [rank] each, float64 scaling, complex128
// potential, source, dest, then per Simp rank int32 count and int32 pairs [count][2].
// output: complex128 dest.
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pyprop_amd_dropin.h"

typedef std::complex<double> cplx;

template <class T>
static void get(FILE* f, T* p, size_t n) {
  if (fread(p, sizeof(T), n, f) != n) {
    std::fprintf(stderr, "caller: short input\n");
    std::exit(2);
  }
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: caller <input> <output>\n");
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int kind, rank;
  get(f, &kind, 1);
  get(f, &rank, 1);
  if (rank < 2 || rank > 3 || kind < 0 || kind > 2) return 2;
  int pe[3], se[3], de[3];
  get(f, pe, rank);
  get(f, se, rank);
  get(f, de, rank);
  double scaling;
  get(f, &scaling, 1);
  size_t np = 1, ns = 1, nd = 1;
  for (int r = 0; r < rank; ++r) {
    np *= pe[r];
    ns *= se[r];
    nd *= de[r];
  }
  std::vector<cplx> potential(np), source(ns), dest(nd);
  get(f, potential.data(), np);
  get(f, source.data(), ns);
  get(f, dest.data(), nd);
  const int nsimp = kind == 0 ? 2 : (kind == 2 ? 1 : 0);
  std::vector<int> pairs[2];
  int count[2] = {0, 0}, two = 2;
  for (int q = 0; q < nsimp; ++q) {
    get(f, &count[q], 1);
    pairs[q].resize(2 * (size_t)count[q]);
    get(f, pairs[q].data(), pairs[q].size());
  }
  std::fclose(f);
  if (kind == 0)
    tensorpotentialmultiply_simp_simp_bandnh_(potential.data(), &pe[0], &pe[1], &pe[2], &scaling, source.data(),
                                              &se[0], &se[1], &se[2], dest.data(), &de[0], &de[1], &de[2],
                                              pairs[0].data(), &count[0], &two, pairs[1].data(), &count[1], &two);
  else if (kind == 1)
    tensorpotentialmultiply_diag_bandnh_(potential.data(), &pe[0], &pe[1], &scaling, source.data(), &se[0], &se[1],
                                         dest.data(), &de[0], &de[1]);
  else
    tensorpotentialmultiply_simp_bandnh_(potential.data(), &pe[0], &pe[1], &scaling, source.data(), &se[0], &se[1],
                                         dest.data(), &de[0], &de[1], pairs[0].data(), &count[0], &two);
  FILE* o = std::fopen(argv[2], "wb");
  if (!o) return 2;
  std::fwrite(dest.data(), sizeof(cplx), nd, o);
  std::fclose(o);
  return 0;
}
