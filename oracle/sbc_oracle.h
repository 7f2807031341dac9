/*
 * sbc_oracle.h -- CPU oracle for the streaming-BC edge-insertion path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a single-threaded C restatement of the
 * reference algorithm used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  Nothing in streamingbc_amd/
 * links, loads or calls it.
 *
 * What it restates (reference file:line):
 *   orc_static(..., ORC_PUSH)   BrandesSingleTree + BrandesApproxCaseParallel
 *                               fold, NT=1 (src/streamingbc_init.c:60-84, 87-183)
 *   orc_static(..., ORC_PULL)   same BFS; delta summed per vertex over its
 *                               children in adjacency order (the summation order
 *                               of the FINE ascent, src/streamingbc_internal_fine.c:191-207)
 *   orc_insert(..., ORC_COARSE) insertEdgeStreamingBC with granularity COARSE,
 *                               NO_BALANCE, NT=1 (src/streamingbc.c:88-226) calling
 *                               addEdgeWithoutMovementBrandes (src/streamingbc_internal.c:11-248)
 *                               and moveUpTreeBrandes (src/streamingbc_internal.c:251-603)
 *   orc_insert(..., ORC_FINE)   the same with granularity FINE, run sequentially:
 *                               addEdgeWithoutMovementBrandesFG / moveUpTreeBrandesFG
 *                               (src/streamingbc_internal_fine.c:11-340, 343-1016), i.e.
 *                               level-synchronous descent and pull-form ascent
 *   orc_delete                  deleteEdgeStreamingBC, NT=1, NO_BALANCE (src/streamingbc.c:258-416):
 *                               the reference's per-root dispatch; the adjacent case
 *                               (src/streamingbc_internal.c:607-868) as the FINE update with
 *                               negated added paths, the shallower endpoint dropping the
 *                               deeper one's old term; the movement case (880-1401) by the
 *                               static pull pass over the post-deletion graph, or, in
 *                               orc_delete_mode(ORC_MOVE_IN_PLACE), by the device's in-place
 *                               rule (movers, their new levels, path counts downwards, deltas
 *                               upwards where an input changed), which is not the reference's
 *   orc_graph_remove            stinger_remove_edge x2 as called by DeleteEdge
 *                               (src/sbcMain_utils.c:473-481; src/stinger/stinger.c:1055-1090)
 *   orc_graph_insert            stinger_insert_edge x2 as called by InsertEdge
 *                               (src/sbcMain_utils.c:458-471; src/stinger/stinger.c:778-869):
 *                               neighbour appended at the end of each list
 *
 * How each part is pinned (the tests that check it):
 *   - static push: bit-exact against the reference compiled from its sources
 *     (oracle/_ref), 2x networkx unnormalised BC fixtures, closed forms on path,
 *     star, cycle and complete graphs, brute-force path counting on tiny graphs;
 *   - static pull: dist/sigma equal to push, delta within 1e-12, closed forms;
 *   - insert coarse / fine: bit-exact (level, sigma, delta, BC) against the
 *     reference's COARSE / FINE update at NT=1, NO_BALANCE (golden fixtures in
 *     tests/golden written by the compiled reference), and equal to a static
 *     recompute of the post-insertion graph (dist/sigma exact, delta/BC 1e-9);
 *   - delete: case counts equal and level/sigma bit-exact, delta/BC within
 *     1e-9, against the reference's deleteEdgeStreamingBC (golden fixtures
 *     written by the compiled reference), and equal to a static recompute of
 *     the post-deletion graph; in place: bit-exact equal to the rebuild on
 *     every movement-case root whose tree before the deletion was the static
 *     pass's, level/sigma bit-exact and delta within 1e-12 after any mix of
 *     updates, and the reference's fixtures as above.
 *
 * Conventions: level and sigma are int64 with LONG_MAX for unreachable
 * vertices, delta and BC are double, BC excludes each tree's own root.
 */
#ifndef SBC_ORACLE_H
#define SBC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_PUSH = 0, ORC_PULL = 1 };
enum { ORC_COARSE = 0, ORC_FINE = 1 };

typedef struct orc_graph orc_graph;
typedef struct orc_forest orc_forest;

/* Per (root, insertion) record written by orc_insert. */
typedef struct {
    int32_t kase;      /* 0: gap 0, 1: |gap| == 1, 2: |gap| >= 2 */
    int32_t gap;       /* level[u] - level[v] (clamped to int32) */
    int64_t processed; /* vertex visits, descent + ascent (dynamicTraverseVerticeCounter) */
    int64_t scanned;   /* adjacency entries scanned (dynamicTraverseEdgeCounter) */
    int64_t affected;  /* |D u U u S| */
    int64_t moved;     /* vertices whose level changed */
    int64_t descent;   /* |D| */
} orc_item;

orc_graph *orc_graph_new(int64_t nv, const int64_t *rowptr, const int64_t *col);
void orc_graph_free(orc_graph *g);
/* Appends v to u's list and u to v's list.  Returns 1 if inserted, 0 if the
 * edge already exists, -1 for a self loop or out-of-range id. */
int orc_graph_insert(orc_graph *g, int64_t u, int64_t v);
/* Removes v from u's list and u from v's list, keeping the order of the other
 * neighbours (STINGER leaves a hole that traversals skip).  Returns 1 if
 * removed, -1 if absent, a self loop or an out-of-range id. */
int orc_graph_remove(orc_graph *g, int64_t u, int64_t v);
int64_t orc_graph_degree(const orc_graph *g, int64_t v);

orc_forest *orc_forest_new(int64_t nv, int64_t K, const int64_t *roots);
void orc_forest_free(orc_forest *f);

void orc_static(orc_forest *f, const orc_graph *g, int order);
/* The edge (u,v) must already be in g.  items: K records or NULL. */
void orc_insert(orc_forest *f, const orc_graph *g, int64_t u, int64_t v, int mode, orc_item *items);
/* The edge (u,v) must already be removed from g.  counts: roots of the
 * same-level, adjacent and movement cases.  Movement-case roots are rebuilt
 * by the static pull pass (orc_delete, ORC_MOVE_REBUILD) or updated in place
 * (ORC_MOVE_IN_PLACE): the same levels and path counts, and a delta
 * recomputed by the pull sum exactly where one of its inputs changed. */
enum { ORC_MOVE_REBUILD = 0, ORC_MOVE_IN_PLACE = 1 };
void orc_delete(orc_forest *f, const orc_graph *g, int64_t u, int64_t v, int64_t counts[3]);
void orc_delete_mode(orc_forest *f, const orc_graph *g, int64_t u, int64_t v, int64_t counts[3], int move);

void orc_get_tree(const orc_forest *f, int64_t ri, int64_t *level, int64_t *sigma, double *delta);
void orc_set_tree(orc_forest *f, int64_t ri, const int64_t *level, const int64_t *sigma, const double *delta);
void orc_get_bc(const orc_forest *f, double *bc);
void orc_set_bc(orc_forest *f, const double *bc);
/* 1 if any sigma addition overflowed int64 since the forest was created. */
int orc_sigma_overflowed(const orc_forest *f);

/* BALANCE estimate of each of K roots for a window of n insertions, from the
 * roots' levels (K x nv, LONG_MAX unreachable) over a CSR graph: the sum over
 * the window of 2 * edgesBelow[s] + edgesAbove[s] of the deeper endpoint s
 * (src/streamingbc.c:98-122; src/streamingbc_init.c:112-169).  Used only to
 * order roots in scheduling comparisons. */
void orc_balance_work(int64_t nv, const int64_t *rowptr, const int64_t *col, int64_t K, const int64_t *levels,
                      int64_t n, const int64_t *u, const int64_t *v, int64_t *work);
/* edgesAbove / edgesBelow of one tree given its levels L[nv], as the static
 * pass computes them (src/streamingbc_init.c:112-169): pinned against the
 * reference's own arrays after its static pass. */
void orc_edge_counts(int64_t nv, const int64_t *rowptr, const int64_t *col, const int64_t *L, int64_t *above,
                     int64_t *below);

#ifdef __cplusplus
}
#endif

#endif
