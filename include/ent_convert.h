/*
 * ent_convert.h - edge lists to CSR on the GPU, and the reference's text
 * tuple format.
 *
 * Replaces the reference's offline converter
 * (tuple_text_to_binary_csr/tuple_text_to_bin.cpp:56-315), which reads a
 * text file of "u v" tuples, shifts ids by the smallest one, and writes a CSR
 * whose rows hold their tuples in file order, duplicates and self-loops kept
 * (both directions with reverse = 1). Here the CSR is built on the device by
 * radix sort, with options the reference lacks: self-loop removal, sorted
 * rows and duplicate removal. The ent_convert CLI is the drop-in for the
 * converter binary.
 *
 * All functions return ENT_OK (0) or a negative ENT_E* code (ent_graph.h);
 * ent_last_error() describes the failure.
 */
#ifndef ENT_CONVERT_H
#define ENT_CONVERT_H

#include <stdint.h>

#include "ent_bfs.h"
#include "ent_graph.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Construction flags. Without ENT_CSR_SORT / ENT_CSR_DEDUP each row keeps
 * its entries in tuple order: tuple i contributes (u, v) to row u and, with
 * ENT_CSR_REVERSE, (v, u) to row v right after it (the reference's order,
 * tuple_text_to_bin.cpp:222-246). */
#define ENT_CSR_REVERSE 1  /* also store every tuple reversed (reverse = 1) */
#define ENT_CSR_NO_LOOPS 2 /* drop tuples (u, u) */
#define ENT_CSR_SORT 4     /* rows in ascending neighbour order */
#define ENT_CSR_DEDUP 8    /* keep the first of repeated neighbours; implies SORT */
/* The symmetric simple graph the BFS benchmarks use. */
#define ENT_CSR_SIMPLE (ENT_CSR_REVERSE | ENT_CSR_NO_LOOPS | ENT_CSR_DEDUP)

/* Tuples read from a text file; arrays owned by the library. */
typedef struct ent_tuples {
  int64_t n;    /* vmax - vmin + 1 */
  int64_t ne;   /* tuples */
  int64_t vmin; /* smallest id in the file; stored ids are id - vmin */
  int32_t *src; /* ne */
  int32_t *dst; /* ne */
} ent_tuples;

/* Reads a text tuple file as the reference converter does
 * (tuple_text_to_bin.cpp:66-147): skips skip_lines lines (a newline at byte 0
 * does not end a line there either), then reads tokens separated by ' ',
 * '\t' or '\n' with atol semantics, two per tuple. The end of the file also
 * ends a token (the reference needs a trailing separator). Multi-threaded.
 * ENT_EIO if the file cannot be read, ENT_EINVAL for fewer lines than
 * skip_lines, an odd token count, or ids spanning 2^31 or more. */
int ent_tuples_read(const char *path, int64_t skip_lines, ent_tuples *out);
void ent_tuples_free(ent_tuples *t);

/* The ef * 2^scale raw R-MAT tuples behind ent_gen_rmat (same arguments and
 * counter-based RNG), self-loops included and nothing merged:
 * ent_csr_build(..., ENT_CSR_SIMPLE) of them equals ent_gen_rmat's CSR. */
int ent_gen_rmat_tuples(int32_t scale, int32_t edgefactor, double a, double b,
                        double c, uint64_t seed, int32_t relabel,
                        int32_t *src, int32_t *dst);

/* Timing of one device construction. */
typedef struct ent_build_stats {
  int64_t records;   /* sorted records: ne, or 2 ne with ENT_CSR_REVERSE */
  int64_t m;         /* CSR entries */
  int32_t key_bits;  /* radix sort key width */
  int32_t pad;
  double device_ms;  /* sort and CSR assembly, HIP events on the build stream */
} ent_build_stats;

/* CSR of ne tuples over n vertices, built on GPU `device` from host arrays.
 * weight (NULL or ne entries) follows its tuple; *weight_out (if weight_out
 * is not NULL) receives m entries, malloc'd, to release with free(). The
 * CSR is returned in *out (release with ent_csr_free). st may be NULL. */
int ent_csr_build(int32_t device, int64_t n, int64_t ne, const int32_t *src,
                  const int32_t *dst, const int32_t *weight, int32_t flags,
                  ent_csr *out, int32_t **weight_out, ent_build_stats *st);

/* The same on device-resident arrays of `device`: src_dev, dst_dev (ne
 * each), beg_pos_dev (n + 1), adj_dev (adj_capacity >= records). *m_out
 * receives m. Scratch memory is allocated and released inside. The ids are
 * not checked here (they are for host arrays): ids outside [0, n) give a
 * wrong CSR, though every write stays inside the given arrays. */
int ent_csr_build_device(int32_t device, int64_t n, int64_t ne,
                         const int32_t *src_dev, const int32_t *dst_dev,
                         int32_t flags, int64_t *beg_pos_dev, int32_t *adj_dev,
                         int64_t adj_capacity, int64_t *m_out,
                         ent_build_stats *st);

/* A BFS handle whose CSR is built on the device from host tuples (no host
 * CSR); upload_flags as ent_graph_upload_ex. */
int ent_graph_upload_edges(const int32_t *src, const int32_t *dst, int64_t n,
                           int64_t ne, int32_t flags, int32_t device,
                           int32_t upload_flags, ent_graph **out);

/* The reference converter's output files for prefix P
 * (tuple_text_to_bin.cpp:153-184), all int64: P_beg_pos.bin (n + 1),
 * P_csr.bin (m), P_deg.bin (n), P_head.bin (m, the row of each entry) and,
 * when weight is not NULL, P_weight.bin (m). The reference's head file holds
 * the row only for entries of the forward direction and 0 for reversed
 * ones. */
int ent_csr_write_ref_files(const ent_csr *g, const int32_t *weight,
                            const char *prefix);

#ifdef __cplusplus
}
#endif

#endif /* ENT_CONVERT_H */
