// ent_convert: drop-in for the reference's converter binary
// (tuple_text_to_binary_csr/tuple_text_to_bin.cpp:56-315). Same positional
// arguments (tuple file, reverse 0/1, lines to skip) and the same output
// files next to the input (<file>_beg_pos.bin, _csr.bin, _deg.bin,
// _head.bin, _weight.bin, all int64); the CSR is built on the GPU.
// Options the reference lacks: --no-loops, --sort, --dedup, --device N,
// --out PREFIX.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ent_convert.h"
#include "ent_graph.h"

int main(int argc, char **argv) {
  std::vector<const char *> pos;
  int flags = 0, device = 0;
  const char *prefix = nullptr;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--no-loops")) flags |= ENT_CSR_NO_LOOPS;
    else if (!std::strcmp(argv[i], "--sort")) flags |= ENT_CSR_SORT;
    else if (!std::strcmp(argv[i], "--dedup")) flags |= ENT_CSR_DEDUP;
    else if (!std::strcmp(argv[i], "--device") && i + 1 < argc) device = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--out") && i + 1 < argc) prefix = argv[++i];
    else pos.push_back(argv[i]);
  }
  if (pos.size() != 3) {
    std::fprintf(stderr,
                 "usage: %s tuple_file reverse(1 reverse, 0 not) lines_to_skip "
                 "[--no-loops] [--sort] [--dedup] [--device N] [--out PREFIX]\n",
                 argv[0]);
    return 2;
  }
  if (std::atol(pos[1]) == 1) flags |= ENT_CSR_REVERSE;
  const long long skip = std::atoll(pos[2]);
  ent_tuples t;
  int rc = ent_tuples_read(pos[0], skip, &t);
  if (rc) {
    std::fprintf(stderr, "error: %s\n", ent_last_error());
    return 1;
  }
  std::printf("tuples: %lld\nmin vertex id: %lld\nmax vertex id: %lld\nvert count: %lld\n", (long long)t.ne,
              (long long)t.vmin, (long long)(t.vmin + t.n - 1), (long long)t.n);
  // One weight per tuple, rand() % 63 + 1 in file order from glibc's default
  // seed, as the reference draws them (tuple_text_to_bin.cpp:230-234).
  std::vector<int32_t> weight(size_t(t.ne > 0 ? t.ne : 1));
  std::srand(1);
  for (int64_t i = 0; i < t.ne; ++i) weight[size_t(i)] = std::rand() % 63 + 1;
  ent_csr g;
  int32_t *w = nullptr;
  ent_build_stats st;
  rc = ent_csr_build(device, t.n, t.ne, t.src, t.dst, weight.data(), flags, &g, &w, &st);
  ent_tuples_free(&t);
  if (rc) {
    std::fprintf(stderr, "error: %s\n", ent_last_error());
    return 1;
  }
  std::printf("edge count: %lld\nbuild: %.3f ms on device %d\n", (long long)g.m, st.device_ms, device);
  const std::string out = prefix ? prefix : pos[0];
  rc = ent_csr_write_ref_files(&g, w, out.c_str());
  std::free(w);
  ent_csr_free(&g);
  if (rc) {
    std::fprintf(stderr, "error: %s\n", ent_last_error());
    return 1;
  }
  std::printf("wrote %s_{beg_pos,csr,deg,head,weight}.bin\n", out.c_str());
  return 0;
}
