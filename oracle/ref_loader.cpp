// Loads a CSR file pair with the reference's own loader (graph.h / graph.hpp,
// instantiated as main.cu:50-52 does: int64 files narrowed to int32) and
// writes what it holds in memory as raw little-endian int32 arrays:
//   ref_loader <beg_pos.bin> <csr.bin> <out.bin>
// out.bin = int32 n, int32 m, int32 beg_pos[n + 1], int32 csr[m].
// Built from the reference sources in place by oracle/Makefile (target ref);
// the output goes to oracle/_ref/ and is used only by tests.
#include "graph.h"

#include <cstdio>

int main(int argc, char **argv) {
  if (argc != 4) return 2;
  graph<long, long, double, int, int, double> g(argv[1], argv[2], nullptr);
  FILE *f = std::fopen(argv[3], "wb");
  if (!f) return 3;
  int hdr[2] = {g.vert_count, g.edge_count};
  std::fwrite(hdr, sizeof(int), 2, f);
  std::fwrite(g.beg_pos, sizeof(int), size_t(g.vert_count) + 1, f);
  std::fwrite(g.csr, sizeof(int), size_t(g.edge_count), f);
  std::fclose(f);
  return 0;
}
