/*
 * ref_driver.c -- a small C driver around the reference StreamingBC library,
 * compiled together with the reference's own sources into oracle/_ref/ by
 * oracle/build_ref.sh.  TEST INFRASTRUCTURE ONLY: it writes golden fixtures
 * in this container and times the reference's OpenMP update as the CPU
 * baseline; nothing in streamingbc_amd/ loads it.
 *
 * The shipped driver never runs the update (src/sbcMain.c:108-178 is commented
 * out), so this file plays the part of InsertEdge + updateEdgeNEW
 * (src/sbcMain_utils.c:458-471, 556-574): insert both directions into STINGER,
 * then time insertEdgeStreamingBC alone with CLOCK_MONOTONIC.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <omp.h>

#include "stinger.h"
#include "streamingbc.h"
#include "streamingbc_aux.h"

typedef struct {
    struct stinger *G;
    bcForest *F;
    extraArraysPerThread **eAPT;
    int64_t NV, NK, NT;
    int64_t *roots;
    int exact;
    StreamingExtraInfo last;
} ref_ctx;

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* CSR with int64 rowptr[nv+1] / col[ne]; NK == nv selects the exact forest. */
void *ref_new(int64_t nv, int64_t ne, const int64_t *rowptr, const int64_t *col,
              int64_t NK, const int64_t *roots, int64_t NT)
{
    ref_ctx *c = (ref_ctx *)calloc(1, sizeof(ref_ctx));
    csrGraph csr;
    csr.NV = nv;
    csr.NE = ne;
    csr.vertexPointerArray = (int64_t *)rowptr;
    csr.edgeArray = (int64_t *)col;
    CreateStingerFromCSR(&csr, &c->G);
    c->NV = nv;
    c->NK = NK;
    c->NT = NT;
    c->roots = (int64_t *)malloc((size_t)NK * sizeof(int64_t));
    memcpy(c->roots, roots, (size_t)NK * sizeof(int64_t));
    c->exact = (NK == nv);
    c->F = c->exact ? streamingBCCreateForestExact(nv)
                    : streamingBCCreateForestApproximate(nv, NK, c->roots);
    c->eAPT = streamingBCCreateAuxilary(NT, nv);
    return c;
}

/* returns wall seconds of the static pass */
double ref_static(void *h)
{
    ref_ctx *c = (ref_ctx *)h;
    /* BrandesExactParallel does not set the team size itself but indexes
     * eAPT by thread number, so the team must not exceed NT */
    omp_set_num_threads((int)c->NT);
    double t0 = now_s();
    if (c->exact) streamingBCInitStaticExact(c->F, c->G, c->NT, c->eAPT);
    else streamingBCInitStaticApproximate(c->F, c->G, c->NT, c->eAPT, c->NK, c->roots);
    return now_s() - t0;
}

/* Inserts (u,v) into STINGER (both directions) and runs the update.
 * Returns the seconds spent in insertEdgeStreamingBC, or -1 if the edge was
 * already present (the reference update is then not called). */
double ref_insert(void *h, int64_t u, int64_t v, int64_t lb, int64_t gran)
{
    ref_ctx *c = (ref_ctx *)h;
    int r1 = stinger_insert_edge(c->G, 0, u, v, 0, 0);
    int r2 = stinger_insert_edge(c->G, 0, v, u, 0, 0);
    if (r1 != 1 || r2 != 1) return -1.0;
    omp_set_num_threads((int)c->NT);
    double t0 = now_s();
    c->last = insertEdgeStreamingBC(c->F, c->G, u, v, c->roots, c->NK, c->NV, c->NT, c->eAPT,
                                    (uint32_t)lb, (uint32_t)gran);
    return now_s() - t0;
}

/* Removes (u,v) from STINGER (both directions) and runs deleteEdgeStreamingBC,
 * as DeleteEdge does (src/sbcMain_utils.c:473-591).  counts3 gets the
 * returned sameLevel, adjacent and movement counts.  Returns the seconds spent
 * in deleteEdgeStreamingBC, or -1 if the edge was not present. */
double ref_delete(void *h, int64_t u, int64_t v, int64_t lb, int64_t gran, int64_t *counts3)
{
    ref_ctx *c = (ref_ctx *)h;
    int r1 = stinger_remove_edge(c->G, 0, u, v);
    int r2 = stinger_remove_edge(c->G, 0, v, u);
    if (r1 != 1 || r2 != 1) return -1.0;
    omp_set_num_threads((int)c->NT);
    double t0 = now_s();
    c->last = deleteEdgeStreamingBC(c->F, c->G, u, v, c->roots, c->NK, c->NV, c->NT, c->eAPT,
                                    (uint32_t)lb, (uint32_t)gran);
    double t = now_s() - t0;
    counts3[0] = c->last.sameLevel;
    counts3[1] = c->last.adjacent;
    counts3[2] = c->last.movement;
    return t;
}

void ref_get_tree(void *h, int64_t ri, int64_t *level, int64_t *sigma, double *delta)
{
    ref_ctx *c = (ref_ctx *)h;
    bcTree *t = c->F->forest[c->roots[ri]];
    for (int64_t v = 0; v < c->NV; v++) {
        if (level) level[v] = t->vArr[v].level;
        if (sigma) sigma[v] = t->vArr[v].sigma;
        if (delta) delta[v] = t->vArr[v].delta;
    }
}

/* edgesAbove / edgesBelow of root index ri's tree (src/streamingbc_aux.h:21-33) */
void ref_get_edge_counts(void *h, int64_t ri, int64_t *above, int64_t *below)
{
    ref_ctx *c = (ref_ctx *)h;
    bcTree *t = c->F->forest[c->roots[ri]];
    for (int64_t v = 0; v < c->NV; v++) {
        above[v] = t->vArr[v].edgesAbove;
        below[v] = t->vArr[v].edgesBelow;
    }
}

void ref_get_bc(void *h, double *bc)
{
    ref_ctx *c = (ref_ctx *)h;
    memcpy(bc, c->F->totalBC, (size_t)c->NV * sizeof(double));
}

/* traversal counters accumulated over all threads since creation */
void ref_counters(void *h, int64_t *out4)
{
    ref_ctx *c = (ref_ctx *)h;
    memset(out4, 0, 4 * sizeof(int64_t));
    for (int64_t t = 0; t < c->NT; t++) {
        out4[0] += c->eAPT[t]->dynamicTraverseVerticeCounter;
        out4[1] += c->eAPT[t]->dynamicTraverseEdgeCounter;
        out4[2] += c->eAPT[t]->adjacentCounter;
        out4[3] += c->eAPT[t]->movementCounter;
    }
}

void ref_free(void *h)
{
    ref_ctx *c = (ref_ctx *)h;
    if (!c) return;
    streamingBCDeleteAuxilary(c->eAPT, c->NT, c->NV);
    if (c->exact) streamingBCDeleteForestExact(&c->F);
    else streamingBCDeleteForestApproximate(&c->F, c->NK, c->roots);
    stinger_free_all(c->G);
    free(c->roots);
    free(c);
}
