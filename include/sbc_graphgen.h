/*
 * sbc_graphgen.h -- host-side graph inputs for the streaming-BC engine (C ABI).
 *
 * Everything here is plain C on the host: deterministic graph generators
 * (R-MAT, Chung-Lu), the DIMACS10 reader/writer with the reference's phantom
 * vertex 0 convention, largest-component extraction, root sampling and
 * insertion-stream generation.  The CPU oracle, the reference baseline and the
 * device library all consume the same CSR these functions produce, so one
 * seeded call gives every consumer identical inputs.
 *
 * Conventions (shared with src/sbcMain_utils.c:382-422 of the reference):
 *   - vertex ids are 1..n, vertex 0 is an isolated phantom, NV = n + 1;
 *   - the graph is undirected and simple: every edge appears in both
 *     adjacency lists, no self loops, no duplicates;
 *   - neighbour order is the CSR order (DIMACS file order for files read,
 *     ascending neighbour id for generated graphs); inserted edges are appended
 *     after it in stream order, as stinger_insert_edge does
 *     (src/stinger/stinger.c:778-869).
 *
 * Random numbers come from splitmix64 seeding a xoshiro256** stream; glibc
 * rand() is never used, so streams are identical on every host.
 */
#ifndef SBC_GRAPHGEN_H
#define SBC_GRAPHGEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Host CSR graph.  rowptr has nv+1 entries, col has ne entries (directed
 * slots, i.e. twice the undirected edge count). */
typedef struct sbcHostGraph {
    int64_t nv;
    int64_t ne;
    int64_t *rowptr;
    int64_t *col;
} sbcHostGraph;

/* Undirected R-MAT with 2^scale vertices and `draws` edge draws using the
 * quadrant probabilities (a, b, c, 1-a-b-c).  Self loops and duplicates are
 * dropped.  keep_lcc != 0 keeps only the largest connected component and
 * relabels it 1..n' in increasing original-id order; otherwise all 2^scale
 * vertices are kept as ids 1..2^scale. */
sbcHostGraph *sbcGenRmat(int scale, int64_t draws, double a, double b, double c,
                         uint64_t seed, int keep_lcc);

/* Chung-Lu power-law graph: n vertices with expected degrees proportional to
 * (i + i0)^(-1/(gamma-1)), scaled to mean degree `mean_deg`; n*mean_deg/2 edge
 * endpoint pairs are drawn in proportion to the weights, self loops and
 * duplicates dropped.  keep_lcc as for sbcGenRmat. */
sbcHostGraph *sbcGenChungLu(int64_t n, double gamma, double mean_deg, uint64_t seed,
                            int keep_lcc);

/* Builds a graph from an undirected edge list of 1-based ids in [1, n]
 * (self loops and duplicates dropped, adjacency sorted ascending). */
sbcHostGraph *sbcGraphFromEdges(int64_t n, int64_t m, const int64_t *eu, const int64_t *ev);

/* DIMACS10: header "n m [fmt]", then line u (1..n) lists u's neighbours.
 * Lines starting with '%' are comments.  Returns NULL on a malformed file. */
sbcHostGraph *sbcReadDimacs(const char *path);
int sbcWriteDimacs(const sbcHostGraph *g, const char *path);

void sbcHostGraphFree(sbcHostGraph *g);

/* Size of the largest connected component (ties: the one holding the
 * smallest vertex id) and, if members != NULL, a 0/1 flag per vertex. */
int64_t sbcLargestComponent(const sbcHostGraph *g, uint8_t *members);

/* K distinct roots, drawn uniformly (partial Fisher-Yates) from the vertices
 * of the largest component, or from all NV ids when from_lcc == 0.  When
 * K == NV the roots are 0..NV-1 in order (exact BC), matching PickRoots in
 * src/sbcMain_utils.c:425-456.  Returns 0 on success. */
int sbcPickRoots(const sbcHostGraph *g, int64_t K, uint64_t seed, int from_lcc, int64_t *roots);

/* `count` insertions (u[i], v[i]): uniformly random pairs of distinct
 * vertices of the largest component that are neither adjacent in g nor
 * already chosen earlier in the stream.  Returns the number generated. */
int64_t sbcGenInsertions(const sbcHostGraph *g, int64_t count, uint64_t seed, int64_t *u, int64_t *v);

/* Counts of an adjacency (for tests): 1 if v is a neighbour of u. */
int sbcHostHasEdge(const sbcHostGraph *g, int64_t u, int64_t v);

#ifdef __cplusplus
}
#endif

#endif /* SBC_GRAPHGEN_H */
