// ent_main: command-line driver with the reference's interface (main.cu:37-88):
//   ent_main <beg_pos.bin> <csr.bin>
// loads the reference's int64 CSR files, picks 64 distinct non-isolated roots
// with glibc rand() from its default seed (main.cu:55-78), and runs the
// compatibility entry point, which prints per-root and average TEPS.
// Extra form (no files needed): ent_main --rmat <scale> [edgefactor] [seed]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ent_bfs.h"

int main(int argc, char **argv) {
  std::printf("Input format: ./exe beg_pos(binary, signed long) csr_file(binary, signed long)\n");
  ent_csr g{};
  int rc;
  if (argc >= 3 && std::strcmp(argv[1], "--rmat") == 0) {
    const int scale = std::atoi(argv[2]);
    const int ef = argc > 3 ? std::atoi(argv[3]) : 16;
    const unsigned long long seed = argc > 4 ? std::strtoull(argv[4], nullptr, 10) : 22;
    rc = ent_gen_rmat(scale, ef, 0.57, 0.19, 0.19, seed, 1, &g);
  } else if (argc == 3) {
    rc = ent_csr_read_bin(argv[1], argv[2], &g);
  } else {
    std::printf("Wrong input\n");
    return -1;
  }
  if (rc) {
    std::fprintf(stderr, "graph load failed: %s\n", ent_last_error());
    return 1;
  }
  std::printf("Graph load (success): %lld verts, %lld edges\n", (long long)g.n, (long long)g.m);
  std::vector<int32_t> roots(64);
  rc = ent_pick_roots_glibc(&g, 64, roots.data());
  if (!rc) rc = ent_bfs_gpu_coalescing_mem(roots.data(), 64, g.beg_pos, g.adj, g.n, g.m, 0);
  if (rc) std::fprintf(stderr, "error: %s\n", ent_last_error());
  ent_csr_free(&g);
  return rc ? 1 : 0;
}
