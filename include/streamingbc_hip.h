/*
 * streamingbc_hip.h -- C ABI of the MI355X streaming betweenness-centrality
 * engine (libstreamingbc_hip.so).
 *
 * Drop-in for the insertion path of the reference's public header
 * src/streamingbc.h: the functions in section 1 keep the reference's names,
 * argument order and meaning, so a C caller written against streamingbc.h
 * (e.g. the InsertEdge/updateEdgeNEW loop of src/sbcMain_utils.c:458-574)
 * compiles against this header and runs on the GPU.  What changes underneath:
 *
 *   - `struct stinger` is an opaque device-resident CSR with per-vertex append
 *     slack (int32 neighbour ids, per-slot insertion sequence numbers) plus a
 *     host mirror of the adjacency, instead of STINGER's blocked lists;
 *   - `bcForest` keeps its public `totalBC` / `NV` fields: `totalBC` is a host
 *     mirror of the device BC that the section-1 entry points refresh before
 *     they return (as the reference leaves BC there, src/streamingbc.c:211-218),
 *     and streamingBCGetBC() refreshes after the sbc* and batched forms; the
 *     per-root trees live in HBM as structure-of-arrays (int32 dist, int64
 *     sigma, FP64 delta) and are read with streamingBCGetTree();
 *   - `extraArraysPerThread` is an opaque device work context (HIP stream,
 *     timing events, staging buffers).  Only element 0 of the array is used.
 *
 * Legacy (void / struct-returning) entry points abort with a message on
 * error, as the reference's xmalloc does (src/stinger/xmalloc.c:24-55).  The
 * sbc* entry points of section 2 return an sbcStatus instead;
 * sbcLastErrorString() describes the last failure of the calling thread.
 *
 * Threading: one host thread at a time per forest/graph/context triple;
 * different triples may be driven from different threads.
 */
#ifndef STREAMINGBC_HIP_H
#define STREAMINGBC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ types */

/* src/streamingbc.h:9-22 */
typedef enum { COARSE, FINE } granularity_t;
typedef enum { DELETE, INSERT } operation_t;
typedef enum { NO_BALANCE, BALANCE } lbalance_t;

/* src/streamingbc.h:29-34 */
typedef struct {
    int64_t NV;
    int64_t NE;
    int64_t *edgeArray;
    int64_t *vertexPointerArray;
} csrGraph;

/* src/streamingbc_aux.h:42-48 (insertEdgeStreamingBC returns it zeroed, as the
 * reference does: src/streamingbc.c:220-225, per-root case data is in
 * sbcItemRecord; deleteEdgeStreamingBC fills movement, adjacent and sameLevel
 * with root counts, as src/streamingbc.c:408-414 does) */
typedef struct {
    int64_t movement;
    int64_t adjacent;
    int64_t sameLevel;
    int64_t connectedComponents;
} StreamingExtraInfo;

struct stinger;                          /* device CSR + slack (replaces STINGER) */
typedef struct sbcForestDev sbcForestDev;

/* src/streamingbc_aux.h:35-39, with the tree pointer table replaced by an
 * opaque device handle. */
typedef struct {
    sbcForestDev *dev;
    double *totalBC; /* host mirror: refreshed by the section-1 update and
                        static-pass calls and by streamingBCGetBC() */
    int64_t NV;
} bcForest;
typedef bcForest *bcForestPtr;

typedef struct extraArraysPerThread extraArraysPerThread;

typedef enum {
    SBC_OK = 0,
    SBC_ERR_ARG = 1,       /* bad argument (id out of range, size mismatch, ...) */
    SBC_ERR_HIP = 2,       /* a HIP runtime call failed */
    SBC_ERR_NOMEM = 3,     /* device or host allocation failed */
    SBC_ERR_STATE = 4,     /* call order violated (e.g. update before static pass) */
    SBC_ERR_EXISTS = 5,    /* edge already present or self loop */
    SBC_ERR_LEVELS = 6,    /* BFS depth beyond the forest's level capacity */
    SBC_ERR_BUDGET = 7,    /* admission refused: footprint above the HBM budget */
    SBC_ERR_ABSENT = 8     /* edge to delete is not in the graph */
} sbcStatus;

/* ------------------------------------------- 1. reference entry points */

/* src/streamingbc.h:40 -- one tree per vertex (roots 0..NV-1) */
bcForest *streamingBCCreateForestExact(int64_t NV);
/* src/streamingbc.h:41 -- trees at roots[0..NK-1] (caller-owned array) */
bcForest *streamingBCCreateForestApproximate(int64_t NV, int64_t NK, int64_t *rootArrayForApproximation);
/* src/streamingbc.h:43-44 */
void streamingBCDeleteForestExact(bcForestPtr *deadForest);
void streamingBCDeleteForestApproximate(bcForestPtr *deadForest, int64_t rootArraySize, int64_t *rootArray);

/* src/streamingbc.h:47-48 -- returns threadCount handles; element 0 is the
 * device work context, the others alias it */
extraArraysPerThread **streamingBCCreateAuxilary(int64_t threadCount, int64_t NV);
void streamingBCDeleteAuxilary(extraArraysPerThread **parallelExtra, int64_t threadCount, int64_t NV);

/* src/streamingbc.h:51-52 -- initial Brandes pass on the device */
void streamingBCInitStaticExact(bcForest *forest, struct stinger *stingerGraph, int64_t NT,
                                extraArraysPerThread **auxilary);
void streamingBCInitStaticApproximate(bcForest *forest, struct stinger *stingerGraph, int64_t NT,
                                      extraArraysPerThread **auxilary, int64_t NK,
                                      int64_t *rootArrayForApproximation);

/* src/streamingbc.h:55-56 -- the edge (newU,newV) must already have been
 * added with stinger_insert_edge in both directions.  loadBalancing is
 * accepted and ignored (every order gives identical trees);
 * granularity COARSE runs one wavefront per root, FINE one workgroup per
 * root, and a FINE item whose descent frontier or ascent scan grows large is
 * finished by kernels spanning the whole GPU.  NT is ignored.  forest->totalBC
 * holds the updated BC on return. */
StreamingExtraInfo insertEdgeStreamingBC(bcForest *forest, struct stinger *sStinger, int64_t newU,
                                         int64_t newV, int64_t *rootArrayForApproximation, int64_t NK,
                                         int64_t NV, int64_t NT, extraArraysPerThread **eAPT,
                                         uint32_t loadBalancing, uint32_t granularity);

/* src/streamingbc.h:57-58, src/streamingbc.c:258-416 -- the edge (oldU,oldV)
 * must already have been removed with stinger_remove_edge in both directions,
 * and every inserted edge applied.  Per root: endpoints on the same level
 * leave the tree as it is; the adjacent case (the deeper endpoint keeps a
 * parent) runs the device update with the lost paths subtracted; the movement
 * case rebuilds the root's tree by the static pass over the post-deletion
 * graph.  BC changes by the same amounts (device FP64 atomics, as after an
 * insertion; streamingBCRecomputeBC folds it deterministically).  Returns
 * the root counts of the three cases.  loadBalancing, granularity and NT are
 * ignored. */
StreamingExtraInfo deleteEdgeStreamingBC(bcForest *forest, struct stinger *sStinger, int64_t oldU,
                                         int64_t oldV, int64_t *rootArrayForApproximation, int64_t NK,
                                         int64_t NV, int64_t NT, extraArraysPerThread **eAPT,
                                         uint32_t loadBalancing, uint32_t granularity);

/* src/streamingbc.h:61-62, src/streamingbc.c:56-77 -- connects src to each
 * vertex of adjacencyArray (inserting both directions itself) and updates
 * every root, as one insertion per edge in list order (run as one batched
 * call: the same trees).  Edges already present, self loops and bad ids are
 * skipped (the reference would update a present edge again).  Returns zero
 * counts, as the reference's insertions do. */
StreamingExtraInfo insertVertexStreamingBC(bcForest *forest, struct stinger *sStinger, int64_t src,
                                           int64_t *adjacencyArray, int64_t adjacencySize,
                                           int64_t *rootArrayForApproximation, int64_t NK, int64_t NV,
                                           int64_t NT, extraArraysPerThread **eAPT);
/* src/streamingbc.h:63-65, src/streamingbc.c:229-256 -- removes every edge of
 * src in list order, one deletion update each (deleteEdgeStreamingBC).  The
 * removed neighbours are written to adjacencyArray and their number to
 * *adjacencySize; a positive *adjacencySize on entry is the array's capacity
 * (otherwise it must hold src's degree).  (The reference never advances its
 * output index: it reports 0 and keeps only the last neighbour.)  Returns the
 * summed case counts of the deletions. */
StreamingExtraInfo deleteVertexStreamingBC(bcForest *forest, struct stinger *sStinger, int64_t src,
                                           int64_t *adjacencyArray, int64_t *adjacencySize,
                                           int64_t *rootArrayForApproximation, int64_t NK, int64_t NV,
                                           int64_t NT, extraArraysPerThread **eAPT);

/* src/streamingbc.h:68 -- builds the device graph from a CSR whose
 * vertexPointerArray has NV+1 entries and edgeArray NE entries */
void CreateStingerFromCSR(csrGraph *csr, struct stinger **stingerGraph);
/* src/streamingbc.h:88 -- DIMACS10 file (1-based ids, phantom vertex 0) */
void ParseDimacsGraph(char initial_graph_name[1024], struct stinger **stingerGraph, int64_t *NV,
                      int64_t *NE);
/* src/stinger/stinger.c:778-869 -- appends `to` to `from`'s list.  Returns 1
 * if inserted, 0 if the edge already exists, -1 on error (self loop, bad id).
 * type, weight and timestamp are ignored. */
int stinger_insert_edge(struct stinger *G, int64_t type, int64_t from, int64_t to, int64_t weight,
                        int64_t timestamp);
/* src/stinger/stinger.c:1055-1090 -- removes `to` from `from`'s list, keeping
 * the order of the other neighbours.  Returns 1 if removed, -1 if absent or on
 * error (self loop, bad id).  type is ignored. */
int stinger_remove_edge(struct stinger *G, int64_t type, int64_t from, int64_t to);
/* src/stinger/stinger.c:72-95 */
int64_t stinger_outdegree(struct stinger *G, int64_t v);
/* src/stinger/stinger.c:558-590 -- frees the graph, returns NULL */
struct stinger *stinger_free_all(struct stinger *S);

/* ------------------------------------------------ 2. additions (status) */

/* Status-returning forms of the creation, static and single-insertion calls
 * (the section-1 forms abort on error instead). */
sbcStatus sbcGraphCreate(int64_t NV, const int64_t *vertexPointerArray, const int64_t *edgeArray,
                         struct stinger **out);
/* roots == NULL: exact forest (NK must equal NV, roots 0..NV-1) */
sbcStatus sbcForestCreate(int64_t NV, int64_t NK, const int64_t *roots, bcForest **out);
sbcStatus sbcInitStatic(bcForest *forest, struct stinger *G, extraArraysPerThread **eAPT);
/* Adds (u,v) to the graph in both directions, then runs the update:
 * stinger_insert_edge x2 + insertEdgeStreamingBC as InsertEdge does
 * (src/sbcMain_utils.c:458-471). */
sbcStatus sbcInsertEdge(bcForest *forest, struct stinger *G, extraArraysPerThread **eAPT, int64_t u,
                        int64_t v, uint32_t granularity);
/* Removes (u,v) from the graph in both directions, then runs the deletion
 * update: stinger_remove_edge x2 + deleteEdgeStreamingBC as the reference's
 * driver does (src/sbcMain_utils.c:473-481, 577-591).  SBC_ERR_ABSENT (nothing
 * modified) if the edge is not there.  info (may be NULL) gets the case counts. */
sbcStatus sbcDeleteEdge(bcForest *forest, struct stinger *G, extraArraysPerThread **eAPT, int64_t u,
                        int64_t v, StreamingExtraInfo *info);
/* Status forms of insertVertexStreamingBC / deleteVertexStreamingBC: adj may
 * be NULL for a deletion; more neighbours than cap is SBC_ERR_ARG before
 * anything changes; *n (may be NULL) gets the number removed. */
sbcStatus sbcInsertVertex(bcForest *forest, struct stinger *G, extraArraysPerThread **eAPT, int64_t src,
                          const int64_t *adj, int64_t n);
sbcStatus sbcDeleteVertex(bcForest *forest, struct stinger *G, extraArraysPerThread **eAPT, int64_t src,
                          int64_t *adj, int64_t cap, int64_t *n, StreamingExtraInfo *info);

/* Checkpoint / resume (the reference's only precedent is STINGER's own
 * save/open, src/stinger/stinger.c:1808-2038): sbcCheckpointSave writes the
 * graph (slot layout with slack, insertion sequence numbers, lookup tables)
 * and every root's tree, state bytes and BC to one file after the device is
 * idle; sbcCheckpointLoad rebuilds a graph and forest from it that continue
 * bit-identically.  Every section carries a checksum; a file that is not a
 * checkpoint of this format, is truncated, fails a checksum or holds a graph
 * or tree out of range (list bounds, levels, uncommitted update state) is
 * SBC_ERR_ARG before any device allocation.  Host memory: one root's arrays
 * at a time. */
sbcStatus sbcCheckpointSave(const bcForest *forest, const struct stinger *G, const char *path);
sbcStatus sbcCheckpointLoad(const char *path, bcForest **forest, struct stinger **G);

/* Copies BC (NV doubles) to out (may be NULL) and refreshes forest->totalBC. */
sbcStatus streamingBCGetBC(bcForest *forest, double *out);

/* Per-root readout for parity checks: level and sigma use the reference's
 * sentinel LONG_MAX for unreachable vertices (src/streamingbc_init.c:93-95).
 * rootIndex indexes the root array given at creation. Any output may be NULL. */
sbcStatus streamingBCGetTree(bcForest *forest, int64_t rootIndex, int64_t *level, int64_t *sigma,
                             double *delta);

/* Recomputes BC as sum over roots (in root-array order) of each tree's delta,
 * skipping each tree's own root.  Deterministic: the result depends only on
 * the trees, not on the order in which updates ran. */
sbcStatus streamingBCRecomputeBC(bcForest *forest, extraArraysPerThread **eAPT);

/* Per (root, insertion) record of the last update call. */
typedef struct {
    int32_t root_index;
    int32_t insertion;  /* index within the call's batch */
    int32_t kase;       /* 0: gap 0, 1: |gap| == 1, 2: |gap| >= 2 */
    int32_t gap;        /* dist[u] - dist[v], clamped */
    int64_t processed;  /* vertex visits, descent + ascent */
    int64_t scanned;    /* adjacency entries scanned */
    int64_t affected;   /* distinct vertices touched */
    int64_t moved;      /* vertices whose distance changed */
} sbcItemRecord;

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  sbcInsertEdge with granularity COARSE whose only
 * scratch -- the work-queue counters and the item records -- lives in a fixed
 * per-GPU array of the library (about 640 KiB, held until the process ends):
 * it calls no hipMalloc/hipFree and creates no event.  records (may be NULL)
 * gets the K item records.  Refused before anything changes: more than 16384
 * roots (SBC_ERR_ARG); an endpoint without append slack, or a graph whose
 * device image must be uploaded again, both of which allocate (SBC_ERR_STATE:
 * an allocating call such as sbcInsertEdge handles them); existing edges and
 * self loops (SBC_ERR_EXISTS).  Every call on one GPU shares the array, which
 * keeps the last call's contents: calls that overlap on one GPU (from several
 * host threads) race and must be serialised by the caller. */
sbcStatus sbcInsertEdgeNoAlloc(bcForest *forest, struct stinger *G, extraArraysPerThread **eAPT, int64_t u,
                               int64_t v, sbcItemRecord *records);

/* Batched persistent form: inserts edges (u[i], v[i]), i = 0..count-1, into
 * the graph and the forest.  The batch is cut into windows of `window`
 * insertions; each window is one persistent kernel launch (captured in a HIP
 * graph and replayed) in which every root applies the window's insertions in
 * stream order, each insertion seeing only edges inserted at or before it.
 * rootOrder (optional, K entries): the order in which roots are taken from
 * the device queue; heavyRoots: how many of them (from the front) are run by
 * a whole workgroup, the rest by one wavefront each.  Results do not depend
 * on rootOrder or heavyRoots.
 * records (optional): K * count entries, [insertion][root].
 * Duplicate edges (already in the graph or twice in the batch) and self loops
 * are rejected with SBC_ERR_EXISTS before anything is modified. */
typedef struct {
    int32_t window;         /* insertions per window (>= 1) */
    int32_t block_threads;  /* 64..1024, multiple of 64 (0: default 256) */
    int32_t grid_blocks;    /* persistent grid size (0: one team per root, capped) */
    int32_t use_graph;      /* capture each window launch in a hipGraph */
    int32_t recompute_bc;   /* 1: BC = sum over roots after each window (deterministic) */
    int32_t heavy_roots;    /* roots run in whole-workgroup mode (-1: all) */
    const int32_t *root_order; /* NULL: root-array order */
    int32_t park_min;       /* an item whose descent frontier reaches this many vertices, whose descent
                               scans park_scan adjacency entries, or one of whose ascent levels leaves
                               more than park_scan entries of hubs to scan, is finished by kernels
                               spanning the whole GPU (0: 4096, < 0: never) */
    int32_t use_context_order; /* 1: root order left in the context by streamingBCScheduleWindow
                                  (sbc_costmodel.h); root_order must be NULL */
    int64_t park_scan;      /* adjacency entries left to scan at which a descent or an ascent level is
                               handed to the whole-GPU kernels (0: 512K) */
} sbcBatchConfig;

sbcStatus streamingBCInsertEdgesBatched(bcForest *forest, struct stinger *G, extraArraysPerThread **eAPT,
                                        const int64_t *u, const int64_t *v, int64_t count,
                                        const sbcBatchConfig *cfg, sbcItemRecord *records);

/* Window graphs built and windows launched as graph replays by the batched
 * calls of this context.  The graphs are built node by node, without stream
 * capture, so nothing another host thread does can invalidate one; a build
 * failure is an error of the batched call, never a silent direct launch. */
sbcStatus streamingBCGraphCounts(extraArraysPerThread **eAPT, int64_t *built, int64_t *launched);

/* Number of items of the last streamingBCInsertEdgesBatched call, FINE
 * insertEdgeStreamingBC / sbcInsertEdge call or deletion that were finished
 * by the whole-GPU kernels (sbcBatchConfig.park_min for the batched form; the
 * FINE single insertion and deletions hand items over at a frontier of 4096
 * vertices or 512K adjacency entries to scan). */
int64_t streamingBCLastParked(extraArraysPerThread **eAPT);

/* Per-root work totals of the batched calls since the last
 * streamingBCResetStats: out[r * 8 + f] for root index r < K, f = nonzero
 * items, processed, scanned, affected, moved, largest affected of one item,
 * then two zeros (the sums of the item records' fields, without copying the
 * records). */
sbcStatus streamingBCWorkTotals(extraArraysPerThread **eAPT, int64_t K, int64_t *out);

/* Compute units of the context's device.  With grid_blocks = 0 a batched call
 * runs min(max(heavy roots, light teams), 2 x this) teams at once. */
int32_t streamingBCComputeUnits(extraArraysPerThread **eAPT);

/* Records of the last insertEdgeStreamingBC call (K entries). */
sbcStatus streamingBCLastRecords(extraArraysPerThread **eAPT, sbcItemRecord *out, int64_t K);

/* Kernel timing, accumulated since the last reset, from HIP events recorded
 * on the context's stream around each launch.  kernel: 0 update (a window's
 * graph -- append, queue reset, window kernel -- or a single insertion's or
 * deletion's update kernel), 1 static pass (and deletion rebuilds), 2 BC
 * recompute, 3 edge append, 4 the whole-GPU kernels that finish items handed
 * over by an update launch (one interval per update launch that may hand
 * items over). */
sbcStatus streamingBCKernelStats(extraArraysPerThread **eAPT, int32_t kernel, int64_t *launches,
                                 double *total_ms);
sbcStatus streamingBCResetStats(extraArraysPerThread **eAPT, int32_t enable_timing);

/* Bytes of device memory a forest of NK roots over NV vertices occupies
 * (state, per-root scratch and lists). */
int64_t streamingBCForestFootprint(int64_t NV, int64_t NK);

/* Bytes of device memory one resident graph needs: its forest, the device
 * CSR with append slack and hub hash tables, and the windowed update's
 * scratch (num_cu <= 0: 256 compute units).  Used for admission control
 * (SBC_ERR_BUDGET).  No counterpart in the reference. */
int64_t streamingBCResidentFootprint(const int64_t *rowptr, int64_t NV, int64_t NK, int32_t num_cu);

/* Blocks until all work queued on the context's stream has finished. */
sbcStatus streamingBCSynchronize(extraArraysPerThread **eAPT);

int sbcLastError(void);
const char *sbcLastErrorString(void);

#ifdef __cplusplus
}
#endif

#endif /* STREAMINGBC_HIP_H */
