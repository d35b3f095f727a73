/*
 * tbfs_gen.h -- deterministic synthetic graphs (host side of libturbobfs).
 *
 * No reference counterpart: the reference reads Matrix Market files only.
 * These generators produce the benchmark workloads.  Every generator uses
 * integer arithmetic on a seeded splitmix64 stream, so a (parameters, seed)
 * pair names one graph on every machine.  Outputs are symmetric CSCs
 * (undirected graphs), rows ascending in each column, without self-loops or
 * duplicate entries.  Functions return 0 on success, 1 for bad arguments,
 * 3 when memory runs out.
 */
#ifndef TBFS_GEN_H
#define TBFS_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tbfs_csc_host tbfs_csc_host;

/* Erdos-Renyi G(n, p): every pair i < j joined with probability p. */
int tbfs_gen_er(int32_t n, double p, uint64_t seed, tbfs_csc_host **out);

/* R-MAT: edge_factor * 2^scale samples, each choosing `scale` quadrants with
 * probabilities a, b, c, 1-a-b-c; symmetrized, deduplicated, self-loops
 * dropped; vertex ids randomly permuted when relabel != 0. */
int tbfs_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b,
                  double c, uint64_t seed, int32_t relabel,
                  tbfs_csc_host **out);

/* w x h 4-neighbour grid with round(remove_frac * E) of its E edges removed
 * uniformly at random; vertex (x, y) is id y*w + x unless relabel != 0. */
int tbfs_gen_grid(int32_t w, int32_t h, double remove_frac, uint64_t seed,
                  int32_t relabel, tbfs_csc_host **out);

/* Watts-Strogatz: ring of n vertices with k (even) nearest neighbours,
 * rewiring probability p. */
int tbfs_gen_ws(int32_t n, int32_t k, double p, uint64_t seed,
                tbfs_csc_host **out);

int tbfs_csc_host_info(const tbfs_csc_host *g, int32_t *n, int32_t *nz);
/* Copy CP (n+1 entries) and IC (nz entries) out; either may be NULL. */
int tbfs_csc_host_copy(const tbfs_csc_host *g, int32_t *CP, int32_t *IC);
void tbfs_csc_host_free(tbfs_csc_host *g);

#ifdef __cplusplus
}
#endif
#endif /* TBFS_GEN_H */
