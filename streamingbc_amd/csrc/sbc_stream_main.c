/*
 * sbc_stream -- C driver of the streaming insertion path, written the way a
 * caller of the reference's streamingbc.h drives it: the InsertEdge /
 * updateEdgeNEW loop of src/sbcMain_utils.c:458-574 (stinger_insert_edge in
 * both directions, then insertEdgeStreamingBC timed with CLOCK_MONOTONIC),
 * and the DeleteEdge step of src/sbcMain_utils.c:473-591 (stinger_remove_edge
 * in both directions, then deleteEdgeStreamingBC), against
 * include/streamingbc_hip.h.
 *
 *   sbc_stream -g file.graph | -r SCALE  [-k ROOTS] [-e INSERTIONS] [-s SEED]
 *              [-G COARSE|FINE] [-w WINDOW] [-d DELETIONS] [-o OUT [-t ROOT]]
 *
 * -r generates an undirected R-MAT graph (16 draws per vertex, largest
 * component); -w > 0 applies the insertions with the batched windowed form;
 * -d then deletes the first DELETIONS inserted edges again, newest first.
 * Prints one JSON line: insertions/s, deletions/s and BC checksums (after the
 * static pass and at the end: equal within rounding when every insertion is
 * deleted again).  With the reference's calls (no -w) BC is read from
 * forest->totalBC directly, as the reference's callers do
 * (src/sbcMain_utils.c:488-548).  -o writes, raw and native-endian, the NV
 * doubles of BC, then level, sigma (int64) and delta (double) of root index
 * ROOT (default 0), NV entries each.
 */
#define _POSIX_C_SOURCE 200809L
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "sbc_graphgen.h"
#include "streamingbc_hip.h"

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void usage(void)
{
    fprintf(stderr, "usage: sbc_stream -g file.graph | -r SCALE [-k ROOTS] [-e INSERTIONS] [-s SEED] "
                    "[-G COARSE|FINE] [-w WINDOW] [-d DELETIONS] [-o OUT [-t ROOT]]\n");
    exit(2);
}

int main(int argc, char **argv)
{
    const char *path = NULL, *out = NULL;
    int64_t tree_root = 0;
    int scale = 0, window = 0;
    int64_t NK = 0, E = 64, D = 0;
    uint64_t seed = 1;
    uint32_t gran = FINE;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "-g") && i + 1 < argc) path = argv[++i];
        else if (!strcmp(argv[i], "-r") && i + 1 < argc) scale = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-k") && i + 1 < argc) NK = atoll(argv[++i]);
        else if (!strcmp(argv[i], "-e") && i + 1 < argc) E = atoll(argv[++i]);
        else if (!strcmp(argv[i], "-s") && i + 1 < argc) seed = strtoull(argv[++i], NULL, 10);
        else if (!strcmp(argv[i], "-w") && i + 1 < argc) window = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-d") && i + 1 < argc) D = atoll(argv[++i]);
        else if (!strcmp(argv[i], "-o") && i + 1 < argc) out = argv[++i];
        else if (!strcmp(argv[i], "-t") && i + 1 < argc) tree_root = atoll(argv[++i]);
        else if (!strcmp(argv[i], "-G") && i + 1 < argc) gran = strcmp(argv[++i], "COARSE") ? FINE : COARSE;
        else usage();
    }
    if (!path && scale <= 0) usage();

    sbcHostGraph *hg = path ? sbcReadDimacs(path) : sbcGenRmat(scale, (int64_t)16 << scale, 0.57, 0.19, 0.19, seed, 1);
    if (!hg) {
        fprintf(stderr, "cannot read or generate the graph\n");
        return 1;
    }
    int64_t NV = hg->nv;
    if (NK <= 0 || NK > NV) NK = NV;
    int64_t *roots = (int64_t *)malloc((size_t)NK * sizeof(int64_t));
    int64_t *src = (int64_t *)malloc((size_t)E * sizeof(int64_t));
    int64_t *dst = (int64_t *)malloc((size_t)E * sizeof(int64_t));
    if (sbcPickRoots(hg, NK, seed, 1, roots) != 0) {
        fprintf(stderr, "cannot pick %ld roots\n", (long)NK);
        return 1;
    }
    E = sbcGenInsertions(hg, E, seed + 10, src, dst);

    csrGraph csr;
    csr.NV = hg->nv;
    csr.NE = hg->ne;
    csr.vertexPointerArray = hg->rowptr;
    csr.edgeArray = hg->col;
    struct stinger *G = NULL;
    CreateStingerFromCSR(&csr, &G);

    bcForest *forest = NK == NV ? streamingBCCreateForestExact(NV) : streamingBCCreateForestApproximate(NV, NK, roots);
    extraArraysPerThread **eAPT = streamingBCCreateAuxilary(1, NV);
    double t0 = now_s();
    if (NK == NV) streamingBCInitStaticExact(forest, G, 1, eAPT);
    else streamingBCInitStaticApproximate(forest, G, 1, eAPT, NK, roots);
    double t_static = now_s() - t0;
    /* the static-pass calls leave BC in forest->totalBC */
    double sum0 = 0.0;
    for (int64_t v = 0; v < NV; v++) sum0 += forest->totalBC[v];

    double t_upd = 0.0;
    if (window > 0) {
        sbcBatchConfig cfg;
        memset(&cfg, 0, sizeof(cfg));
        cfg.window = window;
        cfg.block_threads = 1024;
        cfg.use_graph = 1;
        cfg.heavy_roots = -1;
        t0 = now_s();
        if (streamingBCInsertEdgesBatched(forest, G, eAPT, src, dst, E, &cfg, NULL) != SBC_OK) {
            fprintf(stderr, "batched insert failed: %s\n", sbcLastErrorString());
            return 1;
        }
        t_upd = now_s() - t0;
    } else {
        for (int64_t i = 0; i < E; i++) {
            stinger_insert_edge(G, 0, src[i], dst[i], 0, 0);
            stinger_insert_edge(G, 0, dst[i], src[i], 0, 0);
            t0 = now_s();
            insertEdgeStreamingBC(forest, G, src[i], dst[i], roots, NK, NV, 1, eAPT, BALANCE, gran);
            t_upd += now_s() - t0;
        }
    }
    if (D > E) D = E;
    double t_del = 0.0;
    int64_t cases[3] = {0, 0, 0};
    for (int64_t i = D - 1; i >= 0; i--) {
        stinger_remove_edge(G, 0, src[i], dst[i]);
        stinger_remove_edge(G, 0, dst[i], src[i]);
        t0 = now_s();
        StreamingExtraInfo sei = deleteEdgeStreamingBC(forest, G, src[i], dst[i], roots, NK, NV, 1, eAPT, BALANCE, gran);
        t_del += now_s() - t0;
        cases[0] += sei.sameLevel;
        cases[1] += sei.adjacent;
        cases[2] += sei.movement;
    }
    /* the batched form is an addition: it leaves BC on the device until read */
    if (window > 0) streamingBCGetBC(forest, NULL);
    double sum = 0.0, mx = 0.0;
    for (int64_t v = 0; v < NV; v++) {
        sum += forest->totalBC[v];
        if (forest->totalBC[v] > mx) mx = forest->totalBC[v];
    }
    printf("{\"nv\": %ld, \"roots\": %ld, \"insertions\": %ld, \"static_s\": %.4f, \"insertions_per_s\": %.2f, "
           "\"deletions\": %ld, \"deletions_per_s\": %.2f, \"same_level\": %ld, \"adjacent\": %ld, "
           "\"movement\": %ld, \"bc_sum_static\": %.10e, \"bc_sum\": %.10e, \"bc_max\": %.10e}\n",
           (long)NV, (long)NK, (long)E, t_static, E / (t_upd > 0 ? t_upd : 1e-9), (long)D,
           D / (t_del > 0 ? t_del : 1e-9), (long)cases[0], (long)cases[1], (long)cases[2], sum0, sum, mx);
    if (out) {
        if (tree_root < 0 || tree_root >= NK) tree_root = 0;
        int64_t *lv = (int64_t *)malloc((size_t)NV * sizeof(int64_t));
        int64_t *sg = (int64_t *)malloc((size_t)NV * sizeof(int64_t));
        double *dl = (double *)malloc((size_t)NV * sizeof(double));
        FILE *fp = fopen(out, "wb");
        int ok = fp && lv && sg && dl && streamingBCGetTree(forest, tree_root, lv, sg, dl) == SBC_OK &&
                 fwrite(forest->totalBC, sizeof(double), (size_t)NV, fp) == (size_t)NV &&
                 fwrite(lv, sizeof(int64_t), (size_t)NV, fp) == (size_t)NV &&
                 fwrite(sg, sizeof(int64_t), (size_t)NV, fp) == (size_t)NV &&
                 fwrite(dl, sizeof(double), (size_t)NV, fp) == (size_t)NV;
        if (fp && fclose(fp) != 0) ok = 0;
        free(lv);
        free(sg);
        free(dl);
        if (!ok) {
            fprintf(stderr, "cannot write %s\n", out);
            return 1;
        }
    }

    streamingBCDeleteAuxilary(eAPT, 1, NV);
    if (NK == NV) streamingBCDeleteForestExact(&forest);
    else streamingBCDeleteForestApproximate(&forest, NK, roots);
    stinger_free_all(G);
    sbcHostGraphFree(hg);
    free(roots);
    free(src);
    free(dst);
    return 0;
}
