/*
 * sbc_graphgen.c -- deterministic graph inputs (see include/sbc_graphgen.h).
 *
 * DIMACS10 handling follows the reference's ParseDimacsGraph
 * (src/sbcMain_utils.c:382-422): NV = n + 1 with an isolated phantom vertex 0,
 * line u (1-based) holds u's neighbours in file order.
 */
#include "sbc_graphgen.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

/* ---------------------------------------------------------------- RNG */

typedef struct { uint64_t s[4]; } rng_t;

static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void rng_seed(rng_t *r, uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* xoshiro256** (Blackman & Vigna, 2018) */
static inline uint64_t rng_next(rng_t *r)
{
    uint64_t *s = r->s;
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

/* uniform double in [0, 1) with 53 random bits */
static inline double rng_unif(rng_t *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }

/* uniform integer in [0, n) (Lemire's multiply-shift with rejection) */
static uint64_t rng_below(rng_t *r, uint64_t n)
{
    uint64_t x = rng_next(r);
    __uint128_t m = (__uint128_t)x * n;
    uint64_t l = (uint64_t)m;
    if (l < n) {
        uint64_t t = (0 - n) % n;
        while (l < t) {
            x = rng_next(r);
            m = (__uint128_t)x * n;
            l = (uint64_t)m;
        }
    }
    return (uint64_t)(m >> 64);
}

/* ---------------------------------------------------------------- sort */

static void radix_sort_u64(uint64_t *a, int64_t n)
{
    uint64_t *tmp = (uint64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(uint64_t));
    int64_t *cnt = (int64_t *)malloc(65536 * sizeof(int64_t));
    uint64_t *src = a, *dst = tmp;
    for (int pass = 0; pass < 4; pass++) {
        int sh = pass * 16;
        memset(cnt, 0, 65536 * sizeof(int64_t));
        for (int64_t i = 0; i < n; i++) cnt[(src[i] >> sh) & 0xFFFF]++;
        int64_t s = 0;
        for (int d = 0; d < 65536; d++) { int64_t c = cnt[d]; cnt[d] = s; s += c; }
        for (int64_t i = 0; i < n; i++) dst[cnt[(src[i] >> sh) & 0xFFFF]++] = src[i];
        uint64_t *t = src; src = dst; dst = t;
    }
    /* 4 passes: result is back in a */
    free(tmp);
    free(cnt);
}

/* ---------------------------------------------------------------- graph */

void sbcHostGraphFree(sbcHostGraph *g)
{
    if (!g) return;
    free(g->rowptr);
    free(g->col);
    free(g);
}

/* keys: (u << 32) | v with u, v in [1, n], u != v, both directions present,
 * sorted and unique.  Builds CSR with NV = n + 1. */
static sbcHostGraph *csr_from_sorted_directed(int64_t n, const uint64_t *keys, int64_t nk)
{
    sbcHostGraph *g = (sbcHostGraph *)calloc(1, sizeof(sbcHostGraph));
    g->nv = n + 1;
    g->ne = nk;
    g->rowptr = (int64_t *)calloc((size_t)(g->nv + 1), sizeof(int64_t));
    g->col = (int64_t *)malloc((size_t)(nk > 0 ? nk : 1) * sizeof(int64_t));
    for (int64_t i = 0; i < nk; i++) {
        int64_t u = (int64_t)(keys[i] >> 32);
        g->rowptr[u + 1]++;
        g->col[i] = (int64_t)(keys[i] & 0xFFFFFFFFull);
    }
    for (int64_t v = 0; v < g->nv; v++) g->rowptr[v + 1] += g->rowptr[v];
    return g;
}

/* undirected pairs (1-based, any order) -> simple undirected CSR */
static sbcHostGraph *graph_from_pairs(int64_t n, uint64_t *pairs, int64_t m)
{
    /* normalise to (min,max), drop self loops */
    int64_t w = 0;
    for (int64_t i = 0; i < m; i++) {
        uint64_t a = pairs[i] >> 32, b = pairs[i] & 0xFFFFFFFFull;
        if (a == b) continue;
        if (a > b) { uint64_t t = a; a = b; b = t; }
        pairs[w++] = (a << 32) | b;
    }
    radix_sort_u64(pairs, w);
    int64_t u = 0;
    for (int64_t i = 0; i < w; i++)
        if (i == 0 || pairs[i] != pairs[i - 1]) pairs[u++] = pairs[i];
    uint64_t *dir = (uint64_t *)malloc((size_t)(2 * u > 0 ? 2 * u : 1) * sizeof(uint64_t));
    for (int64_t i = 0; i < u; i++) {
        uint64_t a = pairs[i] >> 32, b = pairs[i] & 0xFFFFFFFFull;
        dir[2 * i] = (a << 32) | b;
        dir[2 * i + 1] = (b << 32) | a;
    }
    radix_sort_u64(dir, 2 * u);
    sbcHostGraph *g = csr_from_sorted_directed(n, dir, 2 * u);
    free(dir);
    return g;
}

sbcHostGraph *sbcGraphFromEdges(int64_t n, int64_t m, const int64_t *eu, const int64_t *ev)
{
    if (n < 0 || n >= (1ll << 32) - 1) return NULL;
    uint64_t *pairs = (uint64_t *)malloc((size_t)(m > 0 ? m : 1) * sizeof(uint64_t));
    for (int64_t i = 0; i < m; i++) {
        if (eu[i] < 1 || eu[i] > n || ev[i] < 1 || ev[i] > n) { free(pairs); return NULL; }
        pairs[i] = ((uint64_t)eu[i] << 32) | (uint64_t)ev[i];
    }
    sbcHostGraph *g = graph_from_pairs(n, pairs, m);
    free(pairs);
    return g;
}

/* component labelling by BFS; returns label per vertex (-1 for none) */
static int64_t *component_labels(const sbcHostGraph *g, int64_t *ncomp, int64_t **sizes_out)
{
    int64_t nv = g->nv;
    int64_t *lab = (int64_t *)malloc((size_t)nv * sizeof(int64_t));
    int64_t *q = (int64_t *)malloc((size_t)nv * sizeof(int64_t));
    int64_t cap = 16, nc = 0;
    int64_t *sizes = (int64_t *)malloc((size_t)cap * sizeof(int64_t));
    for (int64_t v = 0; v < nv; v++) lab[v] = -1;
    for (int64_t s = 1; s < nv; s++) {
        if (lab[s] >= 0) continue;
        if (nc == cap) { cap *= 2; sizes = (int64_t *)realloc(sizes, (size_t)cap * sizeof(int64_t)); }
        int64_t qh = 0, qt = 0;
        q[qt++] = s; lab[s] = nc;
        while (qh < qt) {
            int64_t x = q[qh++];
            for (int64_t e = g->rowptr[x]; e < g->rowptr[x + 1]; e++) {
                int64_t y = g->col[e];
                if (lab[y] < 0) { lab[y] = nc; q[qt++] = y; }
            }
        }
        sizes[nc++] = qt;
    }
    free(q);
    *ncomp = nc;
    *sizes_out = sizes;
    return lab;
}

int64_t sbcLargestComponent(const sbcHostGraph *g, uint8_t *members)
{
    int64_t nc, *sizes;
    int64_t *lab = component_labels(g, &nc, &sizes);
    int64_t best = -1, bsz = 0;
    for (int64_t c = 0; c < nc; c++)
        if (sizes[c] > bsz) { bsz = sizes[c]; best = c; }
    if (members)
        for (int64_t v = 0; v < g->nv; v++) members[v] = (uint8_t)(best >= 0 && lab[v] == best);
    free(lab);
    free(sizes);
    return bsz;
}

/* keep only the largest component, relabelled 1..n' in increasing id order */
static sbcHostGraph *restrict_to_lcc(sbcHostGraph *g)
{
    uint8_t *mem = (uint8_t *)malloc((size_t)g->nv);
    int64_t n2 = sbcLargestComponent(g, mem);
    int64_t *map = (int64_t *)malloc((size_t)g->nv * sizeof(int64_t));
    int64_t next = 1;
    for (int64_t v = 0; v < g->nv; v++) map[v] = mem[v] ? next++ : -1;
    sbcHostGraph *h = (sbcHostGraph *)calloc(1, sizeof(sbcHostGraph));
    h->nv = n2 + 1;
    h->rowptr = (int64_t *)calloc((size_t)(h->nv + 1), sizeof(int64_t));
    int64_t ne = 0;
    for (int64_t v = 0; v < g->nv; v++)
        if (mem[v]) ne += g->rowptr[v + 1] - g->rowptr[v];
    h->ne = ne;
    h->col = (int64_t *)malloc((size_t)(ne > 0 ? ne : 1) * sizeof(int64_t));
    int64_t w = 0;
    for (int64_t v = 0; v < g->nv; v++) {
        if (!mem[v]) continue;
        for (int64_t e = g->rowptr[v]; e < g->rowptr[v + 1]; e++) h->col[w++] = map[g->col[e]];
        h->rowptr[map[v] + 1] = w;
    }
    h->rowptr[0] = 0;
    h->rowptr[1] = 0; /* phantom vertex 0 */
    free(mem);
    free(map);
    sbcHostGraphFree(g);
    return h;
}

/* ---------------------------------------------------------------- generators */

sbcHostGraph *sbcGenRmat(int scale, int64_t draws, double a, double b, double c,
                         uint64_t seed, int keep_lcc)
{
    if (scale < 1 || scale > 30 || draws < 0) return NULL;
    rng_t r;
    rng_seed(&r, seed);
    uint64_t *pairs = (uint64_t *)malloc((size_t)(draws > 0 ? draws : 1) * sizeof(uint64_t));
    double ab = a + b, abc = a + b + c;
    for (int64_t i = 0; i < draws; i++) {
        uint64_t u = 0, v = 0;
        for (int bit = 0; bit < scale; bit++) {
            double p = rng_unif(&r);
            u <<= 1; v <<= 1;
            if (p < a) { /* top-left */ }
            else if (p < ab) { v |= 1; }
            else if (p < abc) { u |= 1; }
            else { u |= 1; v |= 1; }
        }
        pairs[i] = ((u + 1) << 32) | (v + 1);
    }
    sbcHostGraph *g = graph_from_pairs((int64_t)1 << scale, pairs, draws);
    free(pairs);
    return keep_lcc ? restrict_to_lcc(g) : g;
}

sbcHostGraph *sbcGenChungLu(int64_t n, double gamma, double mean_deg, uint64_t seed, int keep_lcc)
{
    if (n < 2 || gamma <= 1.0 || mean_deg <= 0) return NULL;
    rng_t r;
    rng_seed(&r, seed);
    double ex = -1.0 / (gamma - 1.0);
    double *cdf = (double *)malloc((size_t)n * sizeof(double));
    double tot = 0;
    /* i0 = 1 keeps the largest weight finite; weights ~ (i + 1)^(-1/(gamma-1)) */
    for (int64_t i = 0; i < n; i++) {
        double w = pow((double)(i + 1), ex);
        tot += w;
        cdf[i] = tot;
    }
    int64_t m = (int64_t)llround((double)n * mean_deg / 2.0);
    uint64_t *pairs = (uint64_t *)malloc((size_t)(m > 0 ? m : 1) * sizeof(uint64_t));
    for (int64_t i = 0; i < m; i++) {
        uint64_t ends[2];
        for (int k = 0; k < 2; k++) {
            double x = rng_unif(&r) * tot;
            int64_t lo = 0, hi = n - 1;
            while (lo < hi) {
                int64_t mid = (lo + hi) >> 1;
                if (cdf[mid] > x) hi = mid; else lo = mid + 1;
            }
            ends[k] = (uint64_t)lo + 1;
        }
        pairs[i] = (ends[0] << 32) | ends[1];
    }
    free(cdf);
    sbcHostGraph *g = graph_from_pairs(n, pairs, m);
    free(pairs);
    return keep_lcc ? restrict_to_lcc(g) : g;
}

/* ---------------------------------------------------------------- DIMACS */

sbcHostGraph *sbcReadDimacs(const char *path)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return NULL;
    size_t cap = 1 << 16;
    char *line = (char *)malloc(cap);
    int64_t n = -1, m = -1;
    sbcHostGraph *g = NULL;
    int64_t u = 0, w = 0, ecap = 0;
    for (;;) {
        /* read one full line of arbitrary length */
        size_t len = 0;
        int got = 0;
        for (;;) {
            if (!fgets(line + len, (int)(cap - len), fp)) break;
            got = 1;
            len += strlen(line + len);
            if (len > 0 && line[len - 1] == '\n') break;
            if (len + 1 >= cap) { cap *= 2; line = (char *)realloc(line, cap); }
        }
        if (!got) break;
        if (line[0] == '%') continue;
        if (n < 0) {
            if (sscanf(line, "%ld %ld", &n, &m) != 2 || n < 0 || m < 0) break;
            g = (sbcHostGraph *)calloc(1, sizeof(sbcHostGraph));
            g->nv = n + 1;
            g->rowptr = (int64_t *)calloc((size_t)(n + 2), sizeof(int64_t));
            ecap = 2 * m > 0 ? 2 * m : 1;
            g->col = (int64_t *)malloc((size_t)ecap * sizeof(int64_t));
            u = 1;
            continue;
        }
        if (u > n) continue; /* trailing blank lines */
        char *p = line, *end;
        for (;;) {
            long long x = strtoll(p, &end, 10);
            if (end == p) break;
            p = end;
            if (x < 1 || x > n) { sbcHostGraphFree(g); g = NULL; goto done; }
            if (w == ecap) { ecap *= 2; g->col = (int64_t *)realloc(g->col, (size_t)ecap * sizeof(int64_t)); }
            g->col[w++] = (int64_t)x;
        }
        g->rowptr[u + 1] = w;
        u++;
    }
    if (g) {
        for (; u <= n; u++) g->rowptr[u + 1] = w; /* vertices with no line */
        g->ne = w;
    }
done:
    free(line);
    fclose(fp);
    return g;
}

int sbcWriteDimacs(const sbcHostGraph *g, const char *path)
{
    FILE *fp = fopen(path, "w");
    if (!fp) return -1;
    fprintf(fp, "%ld %ld 0\n", (long)(g->nv - 1), (long)(g->ne / 2));
    for (int64_t v = 1; v < g->nv; v++) {
        for (int64_t e = g->rowptr[v]; e < g->rowptr[v + 1]; e++)
            fprintf(fp, e == g->rowptr[v] ? "%ld" : " %ld", (long)g->col[e]);
        fputc('\n', fp);
    }
    return fclose(fp) == 0 ? 0 : -1;
}

/* ---------------------------------------------------------------- streams */

int sbcHostHasEdge(const sbcHostGraph *g, int64_t u, int64_t v)
{
    if (u < 0 || v < 0 || u >= g->nv || v >= g->nv) return 0;
    int64_t du = g->rowptr[u + 1] - g->rowptr[u], dv = g->rowptr[v + 1] - g->rowptr[v];
    if (dv < du) { int64_t t = u; u = v; v = t; }
    for (int64_t e = g->rowptr[u]; e < g->rowptr[u + 1]; e++)
        if (g->col[e] == v) return 1;
    return 0;
}

int sbcPickRoots(const sbcHostGraph *g, int64_t K, uint64_t seed, int from_lcc, int64_t *roots)
{
    if (K == g->nv) {
        for (int64_t i = 0; i < K; i++) roots[i] = i;
        return 0;
    }
    int64_t *cand = (int64_t *)malloc((size_t)g->nv * sizeof(int64_t));
    int64_t nc = 0;
    if (from_lcc) {
        uint8_t *mem = (uint8_t *)malloc((size_t)g->nv);
        sbcLargestComponent(g, mem);
        for (int64_t v = 0; v < g->nv; v++) if (mem[v]) cand[nc++] = v;
        free(mem);
    } else {
        for (int64_t v = 0; v < g->nv; v++) cand[nc++] = v;
    }
    if (K > nc || K < 0) { free(cand); return -1; }
    rng_t r;
    rng_seed(&r, seed);
    for (int64_t i = 0; i < K; i++) {
        int64_t j = i + (int64_t)rng_below(&r, (uint64_t)(nc - i));
        int64_t t = cand[i]; cand[i] = cand[j]; cand[j] = t;
        roots[i] = cand[i];
    }
    free(cand);
    return 0;
}

/* open-addressing set of 64-bit keys (0 is the empty marker; keys are never 0) */
typedef struct { uint64_t *t; int64_t mask; } kset_t;

static void kset_init(kset_t *s, int64_t n)
{
    int64_t c = 16;
    while (c < 4 * n) c <<= 1;
    s->t = (uint64_t *)calloc((size_t)c, sizeof(uint64_t));
    s->mask = c - 1;
}

static int kset_insert(kset_t *s, uint64_t k) /* 1 if newly inserted */
{
    uint64_t h = k * 0x9E3779B97F4A7C15ull;
    int64_t i = (int64_t)(h >> 20) & s->mask;
    while (s->t[i]) {
        if (s->t[i] == k) return 0;
        i = (i + 1) & s->mask;
    }
    s->t[i] = k;
    return 1;
}

int64_t sbcGenInsertions(const sbcHostGraph *g, int64_t count, uint64_t seed, int64_t *u, int64_t *v)
{
    uint8_t *mem = (uint8_t *)malloc((size_t)g->nv);
    int64_t nl = sbcLargestComponent(g, mem);
    int64_t *cand = (int64_t *)malloc((size_t)(nl > 0 ? nl : 1) * sizeof(int64_t));
    int64_t nc = 0;
    for (int64_t x = 0; x < g->nv; x++) if (mem[x]) cand[nc++] = x;
    free(mem);
    if (nc < 2) { free(cand); return 0; }
    rng_t r;
    rng_seed(&r, seed);
    kset_t seen;
    kset_init(&seen, count);
    int64_t made = 0, tries = 0, max_tries = 1000 * (count + 16);
    while (made < count && tries < max_tries) {
        tries++;
        int64_t a = cand[rng_below(&r, (uint64_t)nc)];
        int64_t b = cand[rng_below(&r, (uint64_t)nc)];
        if (a == b) continue;
        if (sbcHostHasEdge(g, a, b)) continue;
        uint64_t lo = (uint64_t)(a < b ? a : b), hi = (uint64_t)(a < b ? b : a);
        if (!kset_insert(&seen, (lo << 32) | hi)) continue;
        u[made] = a;
        v[made] = b;
        made++;
    }
    free(seen.t);
    free(cand);
    return made;
}
