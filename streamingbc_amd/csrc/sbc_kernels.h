// sbc_kernels.h -- device-side data layout shared by sbc_kernels.hip and the
// host runtime sbc_engine.cpp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Per-vertex adjacency descriptor: slots [start, start + deg) of col/seq;
// the first deg0 slots are visible to every update (the rest carry the
// sequence number of the insertion that appended them).
struct SbcVInfo {
    int32_t start;
    int32_t deg;
    int32_t deg0;
    int32_t pad;
};

// Same layout as sbcItemRecord in include/streamingbc_hip.h.
struct SbcItemRec {
    int32_t root_index;
    int32_t insertion;
    int32_t kase;
    int32_t gap;
    int64_t processed;
    int64_t scanned;
    int64_t affected;
    int64_t moved;
};

// Device-resident descriptor of one insertion window (read by the kernels so
// a captured graph can be replayed with new contents).
struct SbcWinDesc {
    int32_t W;       // insertions in the window
    int32_t seq0;    // sequence number of insertion 0
    int32_t nslots;  // appended slots
    int32_t nvupd;   // vertex-info updates
};

enum { SBC_KERR_LEVELS = 1, SBC_KERR_WIDE = 2 };

// Per-hub term bucket entry of the hub lookup pass (per-workgroup scratch).
#define SBC_BUCKET 64
// term slots of one workgroup's lookup buckets (shared out between the hubs
// of a pass; also the flat hub scan's run space): 256 per hub at 512 hubs
#define SBC_BUCKET_SLOTS (512 * 256)
struct SbcBucketEnt {
    int32_t key;   // position in the hub's adjacency list
    int32_t flag;  // 1: add term, 2: sub term
    double add;
    double sub;
};

#define SBC_TRACE_SLOTS 112
#define SBC_HUBQ_CAP 8192   // hubs per workgroup and pass beyond the LDS queue (global overflow queue)

// Path count and dependency of one vertex in one root's tree, with the
// update scratch, in one 32-byte record: an ascent term reads all four, so
// they arrive in one cache line.  Levels and state bytes stay in dense
// arrays of their own (neighbour scans read only those, and the popular
// low-id vertices of a power-law graph keep them cache-resident).
struct alignas(32) SbcVRec {
    int64_t sigma;    // shortest-path count (LONG_MAX: unreachable)
    double delta;     // dependency
    int64_t dpath;    // update scratch: added paths (the reference's diffPath)
    double ndel;      // update scratch: new dependency
};
static_assert(sizeof(SbcVRec) == 32, "four 8-byte fields");

struct SbcParams {
    const SbcVInfo* vinfo;
    const int32_t* col;
    int64_t col_cap;      // adjacency slots allocated on the device (host-side sizing of wide-level scratch)
    const int32_t* seq;
    const int4* hub;      // per vertex {offset into htab, indexed count, log2 table size, 0} (count 0: none)
    const int2* htab;     // open-addressing tables {neighbour + 1, position} of indexed vertices
    SbcBucketEnt* bucket; // per workgroup: SBC_BUCKET_SLOTS term slots (nullptr: no lookups)
    int32_t nv;
    int32_t K;
    int32_t lcap;
    int32_t heavy;       // queue slots [0, heavy) run by whole-workgroup teams
    int32_t* dist;        // levels, INT32_MAX = unreachable
    uint8_t* sb;          // state bytes, sb_stride bytes per root
    int64_t sb_stride;
    SbcVRec* vr;          // [K][nv] path counts, dependencies, update scratch
    double* bc;
    int32_t* list;
    int32_t* lvl;
    const int32_t* roots;
    const int32_t* order; // queue order (nullptr: identity)
    int32_t* qctr;        // [0] heavy queue head, [1] light queue head
    const SbcWinDesc* win; // nullptr: single insertion (u0, v0), all edges visible
    const int32_t* wu;
    const int32_t* wv;
    int32_t u0;
    int32_t v0;
    SbcItemRec* rec;      // [W][K] or nullptr
    int64_t* wtot;        // [K][8] per-root work totals of the batched form (nonzero items, processed,
                          // scanned, affected, moved, largest affected, -, -), added by the root's own
                          // team or by the wide commit, or nullptr
    int64_t* tprof;       // [W][K][SBC_TRACE_SLOTS] or nullptr (recorded by 1024-thread teams only): timestamps (start, descent end, ascent end,
                          // commit end), hubs looked up, hubs scanned, neighbours of each; ascent time in the
                          // light-vertex passes, the hub lookups, the hub staging and scans; ascent levels,
                          // vertices over the ascent levels; light sub-phases; per ascent level 0..7 (deeper:
                          // 7) its time; team-sorted buckets: staging, prefix, placing and summing time, terms,
                          // buckets; per ascent level 0..7: lookup time, scan time, scanned hubs, candidates;
                          // level-2 light passes: vertices, batched vertices and edges, leftover vertices and
                          // edges, rounds, light time, hub-pass time; lookup pass sub-steps: setup, source
                          // rounds (loads, scans, staging), probes, hubs' inserted edges, 8-term sorts,
                          // wavefront sorts; per ascent level 0..3 (deeper: 3) the time of its phases: light pass, hub set-up,
                          // lookups, hub counters, hub scans, end of level; level-2 light vertices past the batch: with the source
                          // set vertices, chunks, chunks with members, members, time; without it vertices, chunks, time
    int32_t* err;
    int32_t bc_atomic;
    int32_t run_cap;      // terms per wavefront run of the flat hub scan (0: full share of the scratch)
    int32_t* hubq;        // per workgroup: SBC_HUBQ_CAP queued hubs beyond the LDS queue, or nullptr
    int32_t hub_batch;    // hubs per pass (<= the LDS queue; smaller only to test the batching)
    int32_t wide_giant;   // whole-GPU ascent: neighbours from which a vertex gets a workgroup (0: the default;
                          // smaller only to test that path)
    // parked items: a whole-workgroup team whose next descent frontier has at
    // least park_min vertices, whose descent has scanned park_scan adjacency
    // entries, or whose ascent level has hubs with at least park_scan entries
    // that lookups cannot serve, stops, saves the
    // item in park[ri] and appends ri to parked (count in qctr[2]); the host
    // finishes it with the wide kernels and resumes the root at
    // start_item[ri]
    struct SbcWideItem* park;   // [K] or nullptr (never park)
    int32_t* parked;            // [K]
    // a parked root also published in host memory while the launch runs:
    // hpark_root[i] = the i-th parked root, then hpark_flag[i] = 1 (a
    // system-scope release), or nullptr
    int32_t* hpark_root;
    int32_t* hpark_flag;
    int32_t park_min;
    int64_t park_scan;
    const int32_t* start_item;  // [K] first item per root, or nullptr (0)
    int32_t qlen;               // queue slots in use (K, or the roots resumed)
    // static pass over some roots only: root indices sroots[0, *scnt), or
    // every root when sroots is nullptr
    const int32_t* sroots;
    const int32_t* scnt;
    // edge deletion (single-item form): the item's edge has left the graph;
    // s loses p's paths (added paths -sigma[p]), and p, marked as a parent
    // once s's level is done, pulls from its old dependency less s's old term
    int32_t del;
    // in-place deletion, movement case: when set, sbc_delmove_kernel stops
    // after the path counts and leaves, per movement root q, its touched-list
    // length in dm_out[2q] and the highest level of its delta pass in
    // dm_out[2q + 1]; the delta levels and the rest are then done over the
    // whole GPU (sbc_delmove_level_kernel, sbc_delmove_finish_kernel), with
    // the list lengths frozen for each level launch in dm_out[2K + q]
    int32_t* dm_out;
    // whole-GPU delta pass: vertices of a level with WIDE_GIANT or more
    // neighbours, listed as (root number, vertex) pairs after a count in
    // dm_giant[0] (at most dm_giant_cap pairs; past it a wavefront does the
    // vertex), each then summed by a workgroup; or nullptr
    int32_t* dm_giant;
    int32_t dm_giant_cap;
};

// A parked item (one per root slot), in global memory.  cnt/stat have the
// roles of a team's list and work counters, so the same device functions
// serve a team (counters in LDS) and the wide kernels (counters here).
struct SbcWideItem {
    int32_t cnt[4];              // end of the descent list, same-level count, end of the ascent list, unused
    unsigned long long stat[4];  // processed, scanned, moved, unused
    int32_t item, s, p, q;       // window item, deeper and shallower endpoint, visibility sequence
    int32_t base, j, fs, fe;     // new level of s, frontiers done, next frontier list[fs, fe)
    int32_t gap, smoved;
    // phase 0: parked during the descent (j, fs, fe as above).  Phase 1:
    // parked in ascent level lev (its ascent set list[uA, uB)) after the
    // team had processed the level's light vertices; the descent is
    // complete (j frontiers) and the counts hold those light vertices.
    int32_t phase, lev, uA, uB;
};

// Scratch of one whole-GPU level of a parked item: the level's vertices
// past the one-per-lane size, found by the level's first launch and handled
// by its second.
struct SbcWideLists {
    int32_t* ctr;     // [4]: heavy vertices, giant vertices (ascent) or work groups (descent), unused,
                      // unused; zeroed before each level
    int32_t* heavy;   // [2 * nv] ascent: vertex ids; descent: {vertex, its first work group} pairs
    int32_t* aux;     // [aux_cap] ascent: giant vertex ids; descent: heavy-list index of each work group
    int32_t aux_cap;  // at least nv + (adjacency slots) / 512 + 1: a frontier's work groups
};

#define SBC_APPEND_INLINE 8

// Appended adjacency slots and vertex-info updates: either a few entries
// passed by value (single insertions) or a device-resident window
// descriptor (captured window launches).
struct SbcAppend {
    SbcVInfo* vinfo;
    int32_t* col;
    int32_t* seq;
    int32_t nslots, nvupd;
    int32_t slot[SBC_APPEND_INLINE];
    int32_t scol[SBC_APPEND_INLINE];
    int32_t sseq[SBC_APPEND_INLINE];
    int32_t vv[SBC_APPEND_INLINE];
    SbcVInfo vi[SBC_APPEND_INLINE];
    const SbcWinDesc* desc;
    const int32_t* d_slot;
    const int32_t* d_scol;
    const int32_t* d_sseq;
    const int32_t* d_vv;
    const SbcVInfo* d_vi;
};

// Arguments of the cost-model feature kernel (sbc_mlp.hip).
struct SbcFeatArgs {
    const SbcVInfo* vinfo;
    const int32_t* dist;
    const SbcVRec* vr;
    const int32_t* u;
    const int32_t* v;
    const float* hist;
    float* X;
    int32_t nv;
    int32_t K;
    int32_t count;
};

struct sbcCostModel;

extern "C" {
hipError_t sbc_launch_features(const SbcFeatArgs* A, hipStream_t st);
hipError_t sbc_launch_hist_update(const SbcItemRec* rec, int count, int K, float decay, float* hist, hipStream_t st);
hipError_t sbc_cost_schedule(sbcCostModel* m, const float* dX, int n, int K, float light_max, float overhead,
                             float* d_pred, float* d_cost, int32_t* d_order, int32_t* d_heavy, hipStream_t st);
// One descent frontier / ascent level of a parked item over the whole GPU:
// the counters of L are zeroed, then two launches (vertices one per lane;
// the longer ones found by the first).  num_cu sizes the grids.
hipError_t sbc_launch_wide_descent(const SbcParams* P, SbcWideItem* wi, int ri, int fs, int fe, int lc,
                                   const SbcWideLists* L, int num_cu, hipStream_t st);
hipError_t sbc_launch_wide_ascent(const SbcParams* P, SbcWideItem* wi, int ri, const int32_t* here6, int lev,
                                  int hubs_only, const SbcWideLists* L, int num_cu, hipStream_t st);
hipError_t sbc_launch_wide_commit(const SbcParams* P, SbcWideItem* wi, int ri, int nf, int nD, int nUend, int nS,
                                  int grid, hipStream_t st);
hipError_t sbc_launch_window(const SbcParams* P, int nwb, int grid, hipStream_t st);
hipError_t sbc_launch_static(const SbcParams* P, int nwb, int grid, hipStream_t st);
// The static pass with an LDS bitmap of each root's settled levels (same
// trees); only for graphs of at most sbc_static_bm_max_nv() vertices.
hipError_t sbc_launch_static_bm(const SbcParams* P, int grid, hipStream_t st);
int64_t sbc_static_bm_max_nv();
hipError_t sbc_launch_bc_reduce(const SbcParams* P, hipStream_t st);
// Movement-case roots of deleting (P->u0, P->v0), root indices
// P->sroots[0, *P->scnt): each tree updated in place to the static pass's
// tree of the new graph, BC by the changed deltas (P->bc_atomic).
hipError_t sbc_launch_delmove(const SbcParams* P, int grid, hipStream_t st);
// With P->dm_out set: the delta pass of level L of every movement root, over
// the whole GPU (after sbc_launch_delmove, levels from the highest down),
// then the state bytes, out-of-reach movers and scratch.
hipError_t sbc_launch_delmove_level(const SbcParams* P, int L, int grid, hipStream_t st);
hipError_t sbc_launch_delmove_finish(const SbcParams* P, int grid, hipStream_t st);
// BC[v] += sign * sum over root indices list[0, *cnt) of delta_r[v] (r != v).
hipError_t sbc_launch_bc_roots(const SbcParams* P, const int32_t* list, const int32_t* cnt, double sign,
                               hipStream_t st);
// Per root, the dispatch of deleting (u, v) (already removed from the graph):
// out[1..3] count the same-level, adjacent and movement roots; adjacent roots
// are listed at out[4..4+K), movement roots at out[4+K..4+2K).  out[0..3]
// must be zero.
// The static pass of root ri over the whole GPU: reset (fr[0] = list end =
// 1), then per level j the frontier list[fs, fe) reaches level j + 1
// (appending at fr[0]) and the new level list[a, b) is settled; then per
// level L, deepest first, the pull of list[a, b).
hipError_t sbc_launch_root_reset(const SbcParams* P, int ri, int32_t* fr, hipStream_t st);
hipError_t sbc_launch_root_bfs(const SbcParams* P, int ri, int fs, int fe, int j, int32_t* fr, hipStream_t st);
hipError_t sbc_launch_root_settle(const SbcParams* P, int ri, int a, int b, int j, hipStream_t st);
// The ascent launch leaves vertices with more than 16384 neighbours in big
// (count big[0], ids big[1..]); sbc_launch_root_bigsum then computes their
// terms (offsets toff[0..n], nterms = toff[n]) and sums each in order.
hipError_t sbc_launch_root_ascent(const SbcParams* P, int ri, int a, int b, int L, int32_t* big, hipStream_t st);
hipError_t sbc_launch_root_bigsum(const SbcParams* P, int ri, int L, const int32_t* big, const int64_t* toff, int n,
                                  int64_t nterms, double* terms, hipStream_t st);
// Rebuild of several roots over the whole GPU: segment i is root ri's list
// range [a, b) on level lev; the roots' ranges are flattened by prefix sums.
struct SbcRootSeg {
    int32_t ri, a, b, lev;
};
// Resets roots rl[0..n) (fr[i] = 1); then per step kind 0 expands every
// segment's frontier (root i appends at fr[i]), kind 1 settles new levels
// (segment lev = the level before), kind 2 pulls a level's dependencies;
// items = pre[n].
hipError_t sbc_launch_roots_reset(const SbcParams* P, const int32_t* rl, int n, int32_t* fr, hipStream_t st);
hipError_t sbc_launch_roots_level(const SbcParams* P, int kind, const SbcRootSeg* sg, const int64_t* pre, int n,
                                  int64_t items, int32_t* fr, hipStream_t st);
// Marks x as a parent on level lev_parent for parked item wi of root ri (a
// deletion's p once the level of s is done).
hipError_t sbc_launch_wide_mark(const SbcParams* P, struct SbcWideItem* wi, int ri, int32_t x, int lev_parent,
                                hipStream_t st);
hipError_t sbc_launch_delete_classify(const SbcParams* P, int32_t u, int32_t v, int32_t* out, hipStream_t st);
hipError_t sbc_launch_reset(int32_t* q, int n, hipStream_t st);
hipError_t sbc_launch_append(const SbcAppend* A, int maxn, hipStream_t st);
// Device address of this GPU's fixed scratch (the current device's): the
// queue counters (4 int32) and SBC_NOALLOC_MAX_ROOTS item records.  Resolved
// once per device; allocates nothing.
#define SBC_NOALLOC_MAX_ROOTS 16384
hipError_t sbc_noalloc_scratch(int32_t** qctr, SbcItemRec** rec);
// Adds one window's launch sequence to graph g as kernel nodes in a chain:
// append (when maxn_append > 0), queue reset of qctr[0..4), the window kernel
// and, with recompute, the BC reduction.  No stream capture is involved.
hipError_t sbc_graph_window(hipGraph_t g, const SbcAppend* A, int maxn_append, int32_t* qctr, const SbcParams* P,
                            int nwb, int grid, int recompute);
}
