/*
 * sbc_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY, never linked into the
 * product).  See sbc_oracle.h for what each function restates and how it is
 * pinned.  Compiled with -ffp-contract=off so every FP64 expression rounds
 * exactly as written, like the reference built with plain gcc -O3.
 */
#include "sbc_oracle.h"

#include <limits.h>
#include <stdlib.h>
#include <string.h>

#define INF_LEVEL LONG_MAX /* INFINITY_MY, src/streamingbc_aux.h:11 */

struct orc_graph {
    int64_t nv;
    int64_t *deg;
    int64_t *cap;
    int64_t **adj;
};

struct orc_forest {
    int64_t nv, K;
    int64_t *roots;
    int64_t *level; /* K x nv */
    int64_t *sigma; /* K x nv */
    double *delta;  /* K x nv */
    double *bc;     /* nv */
    int ovf;
    /* per-call scratch, reset on the touched entries only */
    int64_t *touched, *nsig, *diff, *mdelta, *imoved, *nlev;
    double *ndel, *bcpart;
    int64_t *qd, *qu, *qs;
    int64_t *lvl_d, *lvl_s; /* per-frontier offsets (fine mode) */
};

/* ------------------------------------------------------------------ graph */

orc_graph *orc_graph_new(int64_t nv, const int64_t *rowptr, const int64_t *col)
{
    orc_graph *g = (orc_graph *)calloc(1, sizeof(orc_graph));
    g->nv = nv;
    g->deg = (int64_t *)calloc((size_t)nv, sizeof(int64_t));
    g->cap = (int64_t *)calloc((size_t)nv, sizeof(int64_t));
    g->adj = (int64_t **)calloc((size_t)nv, sizeof(int64_t *));
    for (int64_t v = 0; v < nv; v++) {
        int64_t d = rowptr[v + 1] - rowptr[v];
        g->deg[v] = d;
        g->cap[v] = d + 4;
        g->adj[v] = (int64_t *)malloc((size_t)g->cap[v] * sizeof(int64_t));
        memcpy(g->adj[v], col + rowptr[v], (size_t)d * sizeof(int64_t));
    }
    return g;
}

void orc_graph_free(orc_graph *g)
{
    if (!g) return;
    for (int64_t v = 0; v < g->nv; v++) free(g->adj[v]);
    free(g->adj);
    free(g->deg);
    free(g->cap);
    free(g);
}

int64_t orc_graph_degree(const orc_graph *g, int64_t v) { return g->deg[v]; }

static void push_adj(orc_graph *g, int64_t u, int64_t v)
{
    if (g->deg[u] == g->cap[u]) {
        g->cap[u] = 2 * g->cap[u] + 4;
        g->adj[u] = (int64_t *)realloc(g->adj[u], (size_t)g->cap[u] * sizeof(int64_t));
    }
    g->adj[u][g->deg[u]++] = v;
}

int orc_graph_insert(orc_graph *g, int64_t u, int64_t v)
{
    if (u == v || u < 0 || v < 0 || u >= g->nv || v >= g->nv) return -1;
    for (int64_t i = 0; i < g->deg[u]; i++)
        if (g->adj[u][i] == v) return 0;
    push_adj(g, u, v);
    push_adj(g, v, u);
    return 1;
}

/* Drops v from u's list keeping the order of the rest: STINGER marks the
 * slot deleted and every traversal skips it (src/stinger/stinger.c:1055-1090,
 * src/stinger/stinger-traversal.h:61-76). */
static int drop_adj(orc_graph *g, int64_t u, int64_t v)
{
    for (int64_t i = 0; i < g->deg[u]; i++)
        if (g->adj[u][i] == v) {
            memmove(&g->adj[u][i], &g->adj[u][i + 1], (size_t)(g->deg[u] - i - 1) * sizeof(int64_t));
            g->deg[u]--;
            return 1;
        }
    return 0;
}

int orc_graph_remove(orc_graph *g, int64_t u, int64_t v)
{
    if (u == v || u < 0 || v < 0 || u >= g->nv || v >= g->nv) return -1;
    if (!drop_adj(g, u, v)) return -1;
    drop_adj(g, v, u);
    return 1;
}

/* ------------------------------------------------------------------ forest */

orc_forest *orc_forest_new(int64_t nv, int64_t K, const int64_t *roots)
{
    orc_forest *f = (orc_forest *)calloc(1, sizeof(orc_forest));
    size_t n = (size_t)nv, kn = (size_t)K * n;
    f->nv = nv;
    f->K = K;
    f->roots = (int64_t *)malloc((size_t)K * sizeof(int64_t));
    memcpy(f->roots, roots, (size_t)K * sizeof(int64_t));
    f->level = (int64_t *)malloc(kn * sizeof(int64_t));
    f->sigma = (int64_t *)malloc(kn * sizeof(int64_t));
    f->delta = (double *)calloc(kn, sizeof(double));
    f->bc = (double *)calloc(n, sizeof(double));
    for (size_t i = 0; i < kn; i++) { f->level[i] = INF_LEVEL; f->sigma[i] = INF_LEVEL; }
    f->touched = (int64_t *)calloc(n, sizeof(int64_t));
    f->nsig = (int64_t *)calloc(n, sizeof(int64_t));
    f->diff = (int64_t *)calloc(n, sizeof(int64_t));
    f->mdelta = (int64_t *)calloc(n, sizeof(int64_t));
    f->imoved = (int64_t *)malloc(n * sizeof(int64_t));
    f->nlev = (int64_t *)calloc(n, sizeof(int64_t));
    for (size_t i = 0; i < n; i++) f->imoved[i] = -1;
    f->ndel = (double *)calloc(n, sizeof(double));
    f->bcpart = (double *)calloc(n, sizeof(double));
    f->qd = (int64_t *)malloc(n * sizeof(int64_t));
    f->qu = (int64_t *)malloc(n * sizeof(int64_t));
    f->qs = (int64_t *)malloc(n * sizeof(int64_t));
    f->lvl_d = (int64_t *)malloc((n + 2) * sizeof(int64_t));
    f->lvl_s = (int64_t *)malloc((n + 2) * sizeof(int64_t));
    return f;
}

void orc_forest_free(orc_forest *f)
{
    if (!f) return;
    free(f->roots); free(f->level); free(f->sigma); free(f->delta); free(f->bc);
    free(f->touched); free(f->nsig); free(f->diff); free(f->mdelta); free(f->imoved);
    free(f->nlev); free(f->ndel); free(f->bcpart); free(f->qd); free(f->qu); free(f->qs);
    free(f->lvl_d); free(f->lvl_s);
    free(f);
}

void orc_get_tree(const orc_forest *f, int64_t ri, int64_t *level, int64_t *sigma, double *delta)
{
    size_t o = (size_t)ri * (size_t)f->nv, n = (size_t)f->nv;
    if (level) memcpy(level, f->level + o, n * sizeof(int64_t));
    if (sigma) memcpy(sigma, f->sigma + o, n * sizeof(int64_t));
    if (delta) memcpy(delta, f->delta + o, n * sizeof(double));
}

void orc_set_tree(orc_forest *f, int64_t ri, const int64_t *level, const int64_t *sigma, const double *delta)
{
    size_t o = (size_t)ri * (size_t)f->nv, n = (size_t)f->nv;
    memcpy(f->level + o, level, n * sizeof(int64_t));
    memcpy(f->sigma + o, sigma, n * sizeof(int64_t));
    memcpy(f->delta + o, delta, n * sizeof(double));
}

void orc_get_bc(const orc_forest *f, double *bc) { memcpy(bc, f->bc, (size_t)f->nv * sizeof(double)); }
void orc_set_bc(orc_forest *f, const double *bc) { memcpy(f->bc, bc, (size_t)f->nv * sizeof(double)); }
int orc_sigma_overflowed(const orc_forest *f) { return f->ovf; }

/* sigma additions wrap like two's complement hardware; overflow is recorded */
static inline int64_t sig_add(orc_forest *f, int64_t a, int64_t b)
{
    int64_t r;
    if (__builtin_add_overflow(a, b, &r)) f->ovf = 1;
    return (int64_t)((uint64_t)a + (uint64_t)b);
}

/* (sa / sb) * (d + 1), the dependency term of src/streamingbc_init.c:159-161 */
static inline double dep(int64_t sa, int64_t sb, double d)
{
    return ((double)sa / (double)sb) * (double)(d + 1);
}

/* fold the per-insertion partials into BC: src/streamingbc.c:211-218 (NT = 1) */
static void fold_bc(orc_forest *f)
{
    for (int64_t v = 0; v < f->nv; v++) {
        f->bc[v] += f->bcpart[v];
        f->bcpart[v] = 0.0;
    }
}

/* ------------------------------------------------------------------ static */

/* BFS of src/streamingbc_init.c:92-148.  Writes the visit order to q and
 * returns its length. */
static int64_t static_bfs(orc_forest *f, const orc_graph *g, int64_t ri, int64_t *q)
{
    int64_t nv = f->nv, r = f->roots[ri];
    int64_t *L = f->level + (size_t)ri * nv, *S = f->sigma + (size_t)ri * nv;
    double *D = f->delta + (size_t)ri * nv;
    for (int64_t j = 0; j < nv; j++) { L[j] = INF_LEVEL; S[j] = INF_LEVEL; D[j] = 0; }
    L[r] = 0;
    S[r] = 1;
    int64_t qh = 0, qt = 0;
    q[qt++] = r;
    while (qh < qt) {
        int64_t c = q[qh++];
        for (int64_t e = 0; e < g->deg[c]; e++) {
            int64_t k = g->adj[c][e];
            if (L[k] > L[c]) {
                if (L[k] == INF_LEVEL) {
                    L[k] = L[c] + 1;
                    q[qt++] = k;
                }
                if (S[k] == INF_LEVEL) S[k] = S[c];
                else S[k] = sig_add(f, S[k], S[c]);
            }
        }
    }
    return qt;
}

/* One root's tree from scratch; the visit order is left in f->qd and its
 * length returned. */
static int64_t static_root(orc_forest *f, const orc_graph *g, int64_t ri, int order)
{
    int64_t nv = f->nv;
    int64_t *L = f->level + (size_t)ri * nv, *S = f->sigma + (size_t)ri * nv;
    double *D = f->delta + (size_t)ri * nv;
    int64_t n = static_bfs(f, g, ri, f->qd);
    if (order == ORC_PUSH) {
        /* stack ascent, src/streamingbc_init.c:155-180 */
        for (int64_t i = n - 1; i >= 0; i--) {
            int64_t c = f->qd[i];
            for (int64_t e = 0; e < g->deg[c]; e++) {
                int64_t k = g->adj[c][e];
                if (L[k] == L[c] - 1) D[k] += dep(S[k], S[c], D[c]);
            }
        }
    } else {
        /* pull: each vertex sums its children in adjacency order */
        for (int64_t i = n - 1; i >= 0; i--) {
            int64_t k = f->qd[i];
            double acc = 0.0;
            for (int64_t e = 0; e < g->deg[k]; e++) {
                int64_t c = g->adj[k][e];
                if (L[c] == L[k] + 1) acc += dep(S[k], S[c], D[c]);
            }
            D[k] = acc;
        }
    }
    return n;
}

void orc_static(orc_forest *f, const orc_graph *g, int order)
{
    int64_t nv = f->nv;
    for (int64_t v = 0; v < nv; v++) f->bc[v] = 0.0;
    for (int64_t ri = 0; ri < f->K; ri++) {
        int64_t r = f->roots[ri];
        double *D = f->delta + (size_t)ri * nv;
        int64_t n = static_root(f, g, ri, order);
        /* BC partial in stack order, src/streamingbc_init.c:175-177 */
        for (int64_t i = n - 1; i >= 0; i--) {
            int64_t c = f->qd[i];
            if (c != r) f->bcpart[c] += D[c];
        }
    }
    fold_bc(f);
}

static void fine_update(orc_forest *f, const orc_graph *g, int64_t ri, int64_t s, int64_t p,
                        int moved, int del, orc_item *it);

/* deleteEdgeStreamingBC, src/streamingbc.c:258-416 (NT = 1, NO_BALANCE), for
 * (u,v) already removed from g.  Per root the reference's dispatch (310-345):
 * endpoints on the same level, nothing changes; otherwise s (the deeper
 * endpoint) either keeps a neighbour one level up or not.
 *  - Adjacent case (removeEdgeWithoutMovementBrandes, 607-868 of
 *    src/streamingbc_internal.c; the FG form, src/streamingbc_internal_fine.c:
 *    1019-1400): no level changes; s and its descendants lose sigma[p] paths
 *    (the level-synchronous descent with negative added paths), and
 *    dependencies are pulled again level by level.  p, which no scan reaches
 *    through the removed edge, is marked once s's level is done, and its pull
 *    starts from its old dependency less s's old term.
 *  - Movement case (moveDownTreeBrandes, 880-1401): s moves down or out of
 *    reach; the root's tree is recomputed by the static pull pass over g, the
 *    tree that case ends in, and the change of each dependency added to BC. */
/* ------------------------------------------------------------------ movement case in place */

/* The movement case of deleting (s, p) from root ri's tree (p one level above
 * s and s left without a parent), updated in place to the static pull tree
 * of the new graph: the rule the device's in-place update follows (stated
 * here, not taken from the reference, whose moveDownTreeBrandes,
 * src/streamingbc_internal.c:880-1401, moves the subtree down in its own
 * summation order).  Levels and path counts come out as the static pass's;
 * a delta is recomputed by the pull sum over its children in adjacency order
 * exactly when its level, its path count, its children or their counts or
 * deltas changed, so it equals the static pass's delta whenever the tree
 * before the deletion was the static pass's.  flg (zero) and lst are nv
 * scratch entries; flg is zero again on return. */
enum { MV_LIST = 1, MV_SEEN = 2, MV_MOVER = 4, MV_SIGSEED = 8, MV_SIGPROP = 16, MV_SIGCH = 32, MV_DEL = 64,
       MV_DONE = 128 };

static void mv_flag(int64_t *flg, int64_t *lst, int64_t *n, int64_t y, int64_t fl)
{
    if (!(flg[y] & MV_LIST)) lst[(*n)++] = y;
    flg[y] |= fl | MV_LIST;
}

static void move_in_place(orc_forest *f, const orc_graph *g, int64_t ri, int64_t s, int64_t p, int64_t *flg,
                          int64_t *lst)
{
    int64_t nv = f->nv, r = f->roots[ri];
    int64_t *L = f->level + (size_t)ri * nv, *S = f->sigma + (size_t)ri * nv;
    double *D = f->delta + (size_t)ri * nv;
    int64_t n = 0;
    mv_flag(flg, lst, &n, s, MV_SEEN | MV_MOVER);
    /* movers, old level by old level */
    for (int64_t lv = L[s];; lv++) {
        int64_t t = n;
        for (int64_t i = 0; i < t; i++) {
            int64_t w = lst[i];
            if (!(flg[w] & MV_MOVER) || L[w] != lv) continue;
            for (int64_t e = 0; e < g->deg[w]; e++) {
                int64_t y = g->adj[w][e];
                if (L[y] == lv + 1) mv_flag(flg, lst, &n, y, MV_SEEN);
            }
        }
        int more = 0;
        for (int64_t i = 0; i < n; i++) {
            int64_t y = lst[i];
            if (!(flg[y] & MV_SEEN) || (flg[y] & (MV_MOVER | MV_SIGSEED)) || L[y] != lv + 1) continue;
            int keep = 0;
            for (int64_t e = 0; e < g->deg[y] && !keep; e++) {
                int64_t z = g->adj[y][e];
                if (L[z] == lv && !(flg[z] & MV_MOVER)) keep = 1;
            }
            flg[y] |= keep ? MV_SIGSEED : MV_MOVER;
            if (!keep) more = 1;
        }
        if (!more) break;
    }
    /* new levels: 1 + the lowest level among non-mover neighbours, then
     * breadth-first among the movers */
    int64_t kmin = INF_LEVEL;
    for (int64_t i = 0; i < n; i++) {
        int64_t w = lst[i];
        if (!(flg[w] & MV_MOVER)) continue;
        int64_t best = INF_LEVEL;
        for (int64_t e = 0; e < g->deg[w]; e++) {
            int64_t z = g->adj[w][e];
            if (!(flg[z] & MV_MOVER) && L[z] != INF_LEVEL && L[z] + 1 < best) best = L[z] + 1;
        }
        L[w] = best;
        if (best < kmin) kmin = best;
    }
    for (int64_t k = kmin; k != INF_LEVEL; k++) {
        int more = 0;
        for (int64_t i = 0; i < n; i++) {
            int64_t w = lst[i];
            if (!(flg[w] & MV_MOVER)) continue;
            if (L[w] != k) {
                if (L[w] > k && L[w] != INF_LEVEL) more = 1;
                continue;
            }
            for (int64_t e = 0; e < g->deg[w]; e++) {
                int64_t z = g->adj[w][e];
                if ((flg[z] & MV_MOVER) && L[z] > k + 1) {
                    L[z] = k + 1;
                    more = 1;
                }
            }
        }
        if (!more) break;
    }
    /* path counts, downwards */
    int64_t lo = INF_LEVEL, hi = -1, ld = L[p];
    int64_t t0 = n;
    for (int64_t i = 0; i < t0; i++) {
        int64_t x = lst[i];
        if (L[x] == INF_LEVEL || !(flg[x] & (MV_MOVER | MV_SIGSEED))) continue;
        if (L[x] < lo) lo = L[x];
        if (L[x] > hi) hi = L[x];
        if (!(flg[x] & MV_MOVER)) continue;
        for (int64_t e = 0; e < g->deg[x]; e++) {
            int64_t y = g->adj[x][e];
            if (L[y] == L[x] + 1 && !(flg[y] & MV_MOVER)) {
                mv_flag(flg, lst, &n, y, MV_SIGSEED);
                if (L[x] + 1 > hi) hi = L[x] + 1;
            }
        }
    }
    for (int64_t lv = lo; lv <= hi; lv++) {
        int64_t t = n;
        for (int64_t i = 0; i < t; i++) {
            int64_t x = lst[i];
            int64_t fx = flg[x];
            if ((fx & MV_DONE) || !(fx & (MV_MOVER | MV_SIGSEED | MV_SIGPROP)) || L[x] != lv) continue;
            int64_t sg = 0;
            for (int64_t e = 0; e < g->deg[x]; e++) {
                int64_t y = g->adj[x][e];
                if (L[y] == lv - 1) sg = (int64_t)((uint64_t)sg + (uint64_t)S[y]);
            }
            int changed = sg != S[x] || (fx & MV_MOVER);
            S[x] = sg;
            flg[x] |= MV_DONE | (changed ? MV_SIGCH : 0);
            if (!changed) continue;
            if (lv > ld) ld = lv;
            for (int64_t e = 0; e < g->deg[x]; e++) {
                int64_t c = g->adj[x][e];
                if (L[c] == lv + 1) {
                    mv_flag(flg, lst, &n, c, MV_SIGPROP);
                    if (lv + 1 > hi) hi = lv + 1;
                }
            }
        }
    }
    /* deltas, upwards */
    mv_flag(flg, lst, &n, p, MV_DEL);
    for (int64_t lv = ld; lv >= 0; lv--) {
        int64_t t = n;
        for (int64_t i = 0; i < t; i++) {
            int64_t x = lst[i];
            int64_t fx = flg[x];
            if (!(fx & (MV_SIGCH | MV_DEL)) || L[x] != lv) continue;
            double acc = 0.0;
            for (int64_t e = 0; e < g->deg[x]; e++) {
                int64_t c = g->adj[x][e];
                if (L[c] == lv + 1) acc += dep(S[x], S[c], D[c]);
            }
            int changed = memcmp(&acc, &D[x], sizeof(double)) != 0;
            if (changed) {
                if (x != r) f->bcpart[x] += acc - D[x];
                D[x] = acc;
            }
            if (!changed && !(fx & MV_SIGCH)) continue;
            for (int64_t e = 0; e < g->deg[x]; e++) {
                int64_t q = g->adj[x][e];
                if (L[q] == lv - 1) mv_flag(flg, lst, &n, q, MV_DEL);
            }
        }
    }
    /* out-of-reach movers; scratch */
    for (int64_t i = 0; i < n; i++) {
        int64_t x = lst[i];
        if (L[x] == INF_LEVEL) {
            S[x] = INF_LEVEL;
            if (D[x] != 0.0) f->bcpart[x] -= D[x];
            D[x] = 0.0;
        }
        flg[x] = 0;
    }
}

void orc_delete(orc_forest *f, const orc_graph *g, int64_t u, int64_t v, int64_t counts[3])
{
    orc_delete_mode(f, g, u, v, counts, ORC_MOVE_REBUILD);
}

void orc_delete_mode(orc_forest *f, const orc_graph *g, int64_t u, int64_t v, int64_t counts[3], int move)
{
    int64_t nv = f->nv;
    counts[0] = counts[1] = counts[2] = 0;
    int64_t *flg = move == ORC_MOVE_IN_PLACE ? (int64_t *)calloc((size_t)nv, sizeof(int64_t)) : NULL;
    int64_t *lst = move == ORC_MOVE_IN_PLACE ? (int64_t *)malloc((size_t)nv * sizeof(int64_t)) : NULL;
    double *old = (double *)malloc((size_t)nv * sizeof(double));
    for (int64_t ri = 0; ri < f->K; ri++) {
        int64_t r = f->roots[ri];
        int64_t *L = f->level + (size_t)ri * nv;
        double *D = f->delta + (size_t)ri * nv;
        if (L[u] == L[v]) {
            counts[0]++;
            continue;
        }
        int64_t s = L[u] > L[v] ? u : v, p = L[u] > L[v] ? v : u;
        int64_t extra = 0;
        for (int64_t e = 0; e < g->deg[s]; e++)
            if (L[g->adj[s][e]] == L[p]) extra++;
        if (extra >= 1) {
            counts[1]++;
            orc_item it;
            memset(&it, 0, sizeof(it));
            fine_update(f, g, ri, s, p, 0, 1, &it);
        } else if (move == ORC_MOVE_IN_PLACE) {
            counts[2]++;
            move_in_place(f, g, ri, s, p, flg, lst);
        } else {
            counts[2]++;
            memcpy(old, D, (size_t)nv * sizeof(double));
            static_root(f, g, ri, ORC_PULL);
            for (int64_t x = 0; x < nv; x++)
                if (x != r && D[x] != old[x]) f->bcpart[x] += D[x] - old[x];
        }
    }
    free(old);
    free(flg);
    free(lst);
    fold_bc(f);
}

/* ------------------------------------------------------------------ coarse */

static void reset_scratch(orc_forest *f, int64_t k)
{
    f->diff[k] = 0;
    f->touched[k] = 0;
    f->ndel[k] = 0.0;
    f->mdelta[k] = 0;
    f->imoved[k] = -1;
    f->nsig[k] = 0;
    f->nlev[k] = 0;
}

/* |gap| == 1: src/streamingbc_internal.c:11-248 */
static void coarse_adjacent(orc_forest *f, const orc_graph *g, int64_t ri, int64_t s, int64_t p,
                            int64_t added, orc_item *it)
{
    int64_t nv = f->nv, r = f->roots[ri];
    int64_t *L = f->level + (size_t)ri * nv, *S = f->sigma + (size_t)ri * nv;
    double *D = f->delta + (size_t)ri * nv;
    int64_t *touched = f->touched, *nsig = f->nsig, *diff = f->diff;
    double *ndel = f->ndel;
    int64_t *qd = f->qd, *qu = f->qu;

    nsig[s] = sig_add(f, S[s], added);
    touched[s] = 1;
    diff[s] = added;
    int64_t nd = 0;
    qd[nd++] = s;
    /* descent (53-119): the frontier bookkeeping of the reference is a FIFO */
    for (int64_t h = 0; h < nd; h++) {
        int64_t c = qd[h];
        int64_t tc1 = touched[c] + 1;
        for (int64_t e = 0; e < g->deg[c]; e++) {
            int64_t k = g->adj[c][e];
            if (L[k] == L[c] + 1) {
                if (touched[k] == 0) {
                    nsig[k] = sig_add(f, nsig[k], S[k]);
                    qd[nd++] = k;
                    touched[k] += tc1;
                }
                nsig[k] = sig_add(f, nsig[k], diff[c]);
                diff[k] = sig_add(f, diff[k], diff[c]);
            }
        }
        it->processed++;
        it->scanned += g->deg[c];
    }
    /* ascent (138-211): QueueDown read backwards merged with QueueUp */
    int64_t di = nd - 1, uh = 0, un = 0;
    for (;;) {
        int64_t c;
        if (di < 0 && uh >= un) break;
        if (di < 0) c = qu[uh++];
        else if (uh < un && !(L[qu[uh]] < L[qd[di]])) c = qu[uh++];
        else c = qd[di--];
        for (int64_t e = 0; e < g->deg[c]; e++) {
            int64_t k = g->adj[c][e];
            if (L[k] == L[c] - 1) {
                if (touched[k] == 0) {
                    ndel[k] += D[k];
                    touched[k] = -1;
                    qu[un++] = k;
                    nsig[k] = sig_add(f, nsig[k], S[k]);
                }
                ndel[k] += dep(nsig[k], nsig[c], ndel[c]);
                if (touched[k] < 0 && (k != p || c != s)) ndel[k] -= dep(S[k], S[c], D[c]);
            }
        }
        it->processed++;
        it->scanned += g->deg[c];
        if (c != r) f->bcpart[c] += ndel[c] - D[c];
    }
    /* commit (213-242) */
    for (int64_t i = 0; i < nd; i++) {
        int64_t k = qd[i];
        D[k] = ndel[k];
        S[k] = nsig[k];
        reset_scratch(f, k);
    }
    for (int64_t i = 0; i < un; i++) {
        int64_t k = qu[i];
        D[k] = ndel[k];
        reset_scratch(f, k);
    }
    it->affected = nd + un;
    it->descent = nd;
    it->moved = 0;
}

/* |gap| >= 2: src/streamingbc_internal.c:251-603 */
static void coarse_move(orc_forest *f, const orc_graph *g, int64_t ri, int64_t s, int64_t p,
                        int64_t prevDist, orc_item *it)
{
    int64_t nv = f->nv, r = f->roots[ri];
    int64_t *L = f->level + (size_t)ri * nv, *S = f->sigma + (size_t)ri * nv;
    double *D = f->delta + (size_t)ri * nv;
    int64_t *touched = f->touched, *nsig = f->nsig, *diff = f->diff, *md = f->mdelta, *im = f->imoved;
    double *ndel = f->ndel;
    int64_t *qd = f->qd, *qu = f->qu, *qs = f->qs;

    int64_t nd = 0, nmoved = 0;
    qd[nd++] = s;
    touched[s] = 1;
    nsig[s] = S[p];
    diff[s] = S[p];
    md[s] = prevDist;
    im[s] = 1;
    /* descent (294-386); levels are rewritten in place as each vertex retires */
    for (int64_t h = 0; h < nd; h++) {
        int64_t c = qd[h];
        int64_t tc1 = touched[c] + 1;
        for (int64_t e = 0; e < g->deg[c]; e++) {
            int64_t k = g->adj[c][e];
            if (touched[k] == 0) {
                int64_t cd = md[c] - (L[c] - L[k] + 1);
                touched[k] = tc1;
                if (cd > 0) {
                    nsig[k] = diff[c];
                    diff[k] = diff[c];
                    md[k] = cd;
                    im[k] = 1;
                    qd[nd++] = k;
                } else if (cd == 0) {
                    nsig[k] = sig_add(f, S[k], diff[c]);
                    diff[k] = sig_add(f, diff[k], diff[c]);
                    md[k] = 0;
                    im[k] = 0;
                    qd[nd++] = k;
                } else {
                    touched[k] = 0;
                }
            } else if (touched[k] == tc1) {
                int64_t cd = md[c] - (L[c] - L[k] + 1);
                if (cd >= 0) {
                    nsig[k] = sig_add(f, nsig[k], diff[c]);
                    diff[k] = sig_add(f, diff[k], diff[c]);
                }
            }
        }
        it->processed++;
        it->scanned += g->deg[c];
        L[c] -= md[c];
        if (im[c] == 1) nmoved++;
    }
    /* The level-ordered drain (392-400) reproduces the descent order: every
     * frontier is one new level and each level list is FIFO. */
    int64_t di = nd - 1, uh = 0, un = 0, sh = 0, sn = 0;
    int64_t depthDown = L[qd[di]], depthUp = -1, depthSame = -1;
    for (;;) {
        int64_t c;
        if (depthDown == -1 && depthSame == -1 && depthUp == -1) break;
        if (depthDown >= depthSame && depthDown >= depthUp) {
            c = qd[di--];
            depthDown = di < 0 ? -1 : L[qd[di]];
        } else if (depthUp >= depthSame && depthUp >= depthDown) {
            c = qu[uh++];
            depthUp = uh >= un ? -1 : L[qu[uh]];
        } else {
            c = qs[sh++];
            depthSame = sh >= sn ? -1 : L[qs[sh]];
        }
        for (int64_t e = 0; e < g->deg[c]; e++) {
            int64_t k = g->adj[c][e];
            if (L[k] == L[c] - 1) {
                if (touched[k] == 0) {
                    ndel[k] = D[k];
                    touched[k] = -1;
                    qu[un] = k;
                    if (depthUp == -1) depthUp = L[qu[uh]];
                    un++;
                    nsig[k] = S[k];
                }
                ndel[k] += dep(nsig[k], nsig[c], ndel[c]);
                if (touched[k] < 0 && (k != p || c != s)) ndel[k] -= dep(S[k], S[c], D[c]);
            } else if (L[k] == L[c] && im[c] == 1 && im[k] < 0) {
                if (touched[k] == 0) {
                    ndel[k] = D[k];
                    touched[k] = -2;
                    qs[sn] = k;
                    if (depthSame == -1) depthSame = L[qs[sh]];
                    sn++;
                    nsig[k] = S[k];
                }
                ndel[k] -= dep(S[k], S[c], D[c]);
            }
        }
        it->processed++;
        it->scanned += g->deg[c];
        if (c != r) f->bcpart[c] += ndel[c] - D[c];
    }
    /* commit (547-591) */
    for (int64_t i = 0; i < nd; i++) {
        int64_t k = qd[i];
        D[k] = ndel[k];
        S[k] = nsig[k];
        reset_scratch(f, k);
    }
    for (int64_t i = 0; i < sn; i++) { int64_t k = qs[i]; D[k] = ndel[k]; reset_scratch(f, k); }
    for (int64_t i = 0; i < un; i++) { int64_t k = qu[i]; D[k] = ndel[k]; reset_scratch(f, k); }
    it->affected = nd + un + sn;
    it->descent = nd;
    it->moved = nmoved;
}

/* ------------------------------------------------------------------ fine */

/* Sequential statement of the FINE update: level-synchronous descent, then a
 * level-by-level ascent in which every vertex pulls from its processed
 * children in adjacency order (src/streamingbc_internal_fine.c:160-299 and
 * 560-999).  s = deeper endpoint, p = shallower, moved = |gap| >= 2.
 * touched[k]: >0 descent (frontier index + 1), -1 ascent-only (U/S). */
static void fine_update(orc_forest *f, const orc_graph *g, int64_t ri, int64_t s, int64_t p,
                        int moved, int del, orc_item *it)
{
    int64_t nv = f->nv, r = f->roots[ri];
    int64_t *L = f->level + (size_t)ri * nv, *S = f->sigma + (size_t)ri * nv;
    double *D = f->delta + (size_t)ri * nv;
    int64_t *touched = f->touched, *nsig = f->nsig, *diff = f->diff, *im = f->imoved, *nlev = f->nlev;
    double *ndel = f->ndel;
    int64_t *qd = f->qd, *qu = f->qu, *qs = f->qs;
    int64_t *ld = f->lvl_d, *ls = f->lvl_s;

    int64_t base = L[p] + 1; /* new level of s */
    int64_t nd = 0, ns = 0, nmoved = 0;
    qd[nd++] = s;
    touched[s] = 1;
    im[s] = moved ? 1 : 0;
    nlev[s] = base;
    /* a deletion (adjacent case) takes sigma[p] paths away from s */
    diff[s] = del ? (int64_t)(0ull - (uint64_t)S[p]) : S[p];
    nsig[s] = moved ? S[p] : sig_add(f, S[s], diff[s]);
    int64_t nf = 0; /* frontiers: qd[ld[j] .. ld[j+1]) */
    ld[0] = 0;
    ls[0] = 0;
    int64_t fs = 0, fe = 1;
    while (fs < fe) {
        int64_t lc = base + nf; /* new level of this frontier */
        for (int64_t h = fs; h < fe; h++) {
            int64_t c = qd[h];
            for (int64_t e = 0; e < g->deg[c]; e++) {
                int64_t k = g->adj[c][e];
                if (touched[k] == 0) {
                    int64_t cd = (L[k] == INF_LEVEL) ? 1 : L[k] - (lc + 1);
                    if (cd >= 0) {
                        touched[k] = nf + 2;
                        im[k] = cd > 0 ? 1 : 0;
                        nlev[k] = lc + 1;
                        diff[k] = diff[c];
                        qd[nd++] = k;
                    } else if (cd == -1 && im[c] == 1) {
                        touched[k] = -1; /* same level as a vertex that moved up */
                        qs[ns++] = k;
                    }
                } else if (touched[k] == nf + 2) {
                    diff[k] = sig_add(f, diff[k], diff[c]);
                }
            }
            it->processed++;
            it->scanned += g->deg[c];
        }
        for (int64_t h = fe; h < nd; h++) {
            int64_t k = qd[h];
            nsig[k] = im[k] == 1 ? diff[k] : sig_add(f, S[k], diff[k]);
        }
        nf++;
        ld[nf] = fe;
        ls[nf] = ns;
        fs = fe;
        fe = nd;
    }
    /* frontier j (0..nf-1) = qd[ld[j]..ld[j+1]), S vertices found while
     * expanding frontier j = qs[ls[j]..ls[j+1]) and sit at level base + j */
    for (int64_t i = 0; i < nd; i++) if (im[qd[i]] == 1) nmoved++;

    int64_t un = 0, uprev = 0; /* U vertices of the current level: qu[uprev..un) */
    for (int64_t lev = base + nf - 1; lev >= 0; lev--) {
        int64_t j = lev - base;
        int64_t d0 = 0, d1 = 0, s0 = 0, s1 = 0;
        if (j >= 0) { d0 = ld[j]; d1 = ld[j + 1]; s0 = ls[j]; s1 = ls[j + 1]; }
        int64_t u0 = uprev, u1 = un;
        if (d1 - d0 + s1 - s0 + u1 - u0 == 0) break;
        uprev = un;
        for (int64_t w = 0; w < (d1 - d0) + (s1 - s0) + (u1 - u0); w++) {
            int64_t k;
            if (w < d1 - d0) k = qd[d0 + w];
            else if (w < d1 - d0 + s1 - s0) k = qs[s0 + w - (d1 - d0)];
            else k = qu[u0 + w - (d1 - d0) - (s1 - s0)];
            int isD = touched[k] > 0;
            int64_t sk = isD ? nsig[k] : S[k];
            double acc = isD ? 0.0 : D[k];
            /* a deletion's p no longer meets s in its list: its pull starts
             * without s's old term */
            if (del && k == p) acc = acc - dep(S[p], S[s], D[s]);
            for (int64_t e = 0; e < g->deg[k]; e++) {
                int64_t c = g->adj[k][e];
                int64_t tcv = touched[c];
                int64_t lcv = tcv > 0 ? nlev[c] : L[c];
                if (tcv != 0 && lcv == lev + 1) {
                    int64_t sc = tcv > 0 ? nsig[c] : S[c];
                    acc += dep(sk, sc, ndel[c]);
                    if (!isD && (k != p || c != s)) acc -= dep(S[k], S[c], D[c]);
                } else if (!isD && tcv > 0 && lcv == lev && im[c] == 1) {
                    acc -= dep(S[k], S[c], D[c]);
                } else if (tcv == 0 && lcv == lev - 1) {
                    touched[c] = -1;
                    qu[un++] = c;
                }
            }
            ndel[k] = acc;
            it->processed++;
            it->scanned += g->deg[k];
            if (k != r) f->bcpart[k] += ndel[k] - D[k];
        }
        /* a deletion's p lost its child s (no scan reaches p through the
         * removed edge): it is updated on the next level */
        if (del && lev == base && touched[p] == 0) {
            touched[p] = -1;
            qu[un++] = p;
        }
    }
    for (int64_t i = 0; i < nd; i++) {
        int64_t k = qd[i];
        L[k] = nlev[k];
        S[k] = nsig[k];
        D[k] = ndel[k];
        reset_scratch(f, k);
    }
    for (int64_t i = 0; i < ns; i++) { int64_t k = qs[i]; D[k] = ndel[k]; reset_scratch(f, k); }
    for (int64_t i = 0; i < un; i++) { int64_t k = qu[i]; D[k] = ndel[k]; reset_scratch(f, k); }
    it->affected = nd + ns + un;
    it->descent = nd;
    it->moved = nmoved;
}

/* ------------------------------------------------------------------ insert */

void orc_insert(orc_forest *f, const orc_graph *g, int64_t u, int64_t v, int mode, orc_item *items)
{
    int64_t nv = f->nv;
    for (int64_t ri = 0; ri < f->K; ri++) {
        int64_t *L = f->level + (size_t)ri * nv, *S = f->sigma + (size_t)ri * nv;
        orc_item tmp, *it = items ? &items[ri] : &tmp;
        memset(it, 0, sizeof(*it));
        int64_t lu = L[u], lv = L[v];
        int64_t diff;
        /* both unreachable: gap 0; one unreachable: it moves up (the
         * reference's LONG_MAX arithmetic gives the same classification) */
        if (lu == INF_LEVEL && lv == INF_LEVEL) diff = 0;
        else if (lu == INF_LEVEL) diff = 2 + (nv + 2);
        else if (lv == INF_LEVEL) diff = -(2 + (nv + 2));
        else diff = lu - lv;
        it->gap = (int32_t)(diff > INT32_MAX ? INT32_MAX : (diff < -INT32_MAX ? -INT32_MAX : diff));
        if (diff == 0) { it->kase = 0; continue; }
        int64_t s = diff > 0 ? u : v, p = diff > 0 ? v : u;
        int64_t ad = diff > 0 ? diff : -diff;
        if (ad == 1) {
            it->kase = 1;
            if (mode == ORC_COARSE) coarse_adjacent(f, g, ri, s, p, S[p], it);
            else fine_update(f, g, ri, s, p, 0, 0, it);
        } else {
            it->kase = 2;
            if (mode == ORC_COARSE) {
                int64_t prev = (L[s] == INF_LEVEL) ? (INF_LEVEL - L[p] - 1) : ad - 1;
                coarse_move(f, g, ri, s, p, prev, it);
            } else {
                fine_update(f, g, ri, s, p, 1, 0, it);
            }
        }
    }
    fold_bc(f);
}

/* ------------------------------------------------------------------ BALANCE */

/* The reference's BALANCE work estimate (src/streamingbc.c:98-122) summed
 * over a window of insertions, for ordering roots in scheduling comparisons.
 * Per root r, from r's levels (levels[r * nv + x], LONG_MAX unreachable) over
 * the CSR graph (rowptr, col): edgesAbove[k] = sum over parents p of
 * (edgesAbove[p] + 1) in increasing level order and edgesBelow[k] = sum over
 * children c of (edgesBelow[c] + 1) in decreasing level order, as the static
 * pass computes them (src/streamingbc_init.c:112-169); then work[r] = sum over
 * the window's insertions (u, v) with different endpoint levels of
 * 2 * edgesBelow[s] + edgesAbove[s], s the deeper endpoint. */
void orc_edge_counts(int64_t nv, const int64_t *rowptr, const int64_t *col, const int64_t *L, int64_t *above,
                     int64_t *below)
{
    int64_t maxl = 0;
    for (int64_t x = 0; x < nv; x++)
        if (L[x] != INF_LEVEL && L[x] > maxl) maxl = L[x];
    int64_t *cnt = (int64_t *)calloc((size_t)(maxl + 2), sizeof(int64_t));
    int64_t *order = (int64_t *)malloc((size_t)nv * sizeof(int64_t));
    for (int64_t x = 0; x < nv; x++)
        if (L[x] != INF_LEVEL) cnt[L[x] + 1]++;
    for (int64_t l = 0; l <= maxl; l++) cnt[l + 1] += cnt[l];
    for (int64_t x = 0; x < nv; x++) {
        above[x] = 0;
        below[x] = 0;
        if (L[x] != INF_LEVEL) order[cnt[L[x]]++] = x;
    }
    const int64_t m = cnt[maxl];
    for (int64_t i = 0; i < m; i++) {
        int64_t k = order[i];
        for (int64_t e = rowptr[k]; e < rowptr[k + 1]; e++) {
            int64_t p = col[e];
            if (L[p] != INF_LEVEL && L[p] == L[k] - 1) above[k] += above[p] + 1;
        }
    }
    for (int64_t i = m - 1; i >= 0; i--) {
        int64_t k = order[i];
        for (int64_t e = rowptr[k]; e < rowptr[k + 1]; e++) {
            int64_t c = col[e];
            if (L[c] != INF_LEVEL && L[c] == L[k] + 1) below[k] += below[c] + 1;
        }
    }
    free(order);
    free(cnt);
}

void orc_balance_work(int64_t nv, const int64_t *rowptr, const int64_t *col, int64_t K, const int64_t *levels,
                      int64_t n, const int64_t *u, const int64_t *v, int64_t *work)
{
    int64_t *above = (int64_t *)malloc((size_t)nv * sizeof(int64_t));
    int64_t *below = (int64_t *)malloc((size_t)nv * sizeof(int64_t));
    for (int64_t r = 0; r < K; r++) {
        const int64_t *L = levels + (size_t)r * (size_t)nv;
        orc_edge_counts(nv, rowptr, col, L, above, below);
        int64_t w = 0;
        for (int64_t i = 0; i < n; i++) {
            int64_t lu = L[u[i]], lv = L[v[i]];
            if (lu == lv) continue;
            int64_t s = (lu == INF_LEVEL || (lv != INF_LEVEL && lu > lv)) ? u[i] : v[i];
            w += 2 * below[s] + above[s];
        }
        work[r] = w;
    }
    free(above);
    free(below);
}
