// sbc_engine.cpp -- host runtime behind include/streamingbc_hip.h.
//
// Owns the device graph (CSR + append slack, with a host mirror used for
// duplicate checks, slot assignment and relayout), the per-root forest in HBM,
// and the work context (stream, timing events, window staging, cached HIP
// graphs).  All compute runs in the kernels of sbc_kernels.hip; there is no
// host fallback: every entry point fails if the device calls fail.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <algorithm>
#include <thread>
#include <utility>
#include <vector>

#include "sbc_costmodel.h"
#include "sbc_graphgen.h"
#include "sbc_kernels.h"
#include "streamingbc_hip.h"

// ------------------------------------------------------------ errors

static thread_local int g_last_err = 0;
static thread_local char g_last_msg[512] = "";

static sbcStatus set_err(sbcStatus st, const char* fmt, ...)
{
    g_last_err = st;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_msg, sizeof(g_last_msg), fmt, ap);
    va_end(ap);
    return st;
}

extern "C" int sbcLastError(void) { return g_last_err; }
extern "C" const char* sbcLastErrorString(void) { return g_last_msg; }

#define HIPCHK(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return set_err(SBC_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                           __FILE__, __LINE__);                                         \
    } while (0)

[[noreturn]] static void legacy_abort(const char* where)
{
    fprintf(stderr, "streamingbc_hip: %s failed: %s\n", where, g_last_msg);
    abort();
}

// ------------------------------------------------------------ graph

static const int32_t kNoSeq = 0;

struct stinger {
    int64_t nv = 0;
    // host mirror
    std::vector<int32_t> start, deg, deg0, cap;
    std::vector<int32_t> col, seq;
    // per-vertex host scratch of the batched call, all zero between calls
    std::vector<int32_t> h_cnt;
    std::vector<char> h_mark;
    int64_t total_cap = 0;
    int32_t seq_next = 1;
    // device image
    SbcVInfo* d_vinfo = nullptr;
    int32_t* d_col = nullptr;
    int32_t* d_seq = nullptr;
    int64_t d_cap = 0;   // slots allocated on the device
    bool full_dirty = true;
    uint64_t generation = 0; // bumps when device arrays are reallocated
    // pending per-slot writes and vertex-info updates (legacy path)
    std::vector<int32_t> p_slot, p_col, p_seq, p_v;
    // open-addressing tables over the initial adjacency of high-degree
    // vertices: hub[v] = {offset, count, log2 size, 0}, htab = {neighbour + 1, position}
    std::vector<int4> hub;
    std::vector<int2> htab;
    int4* d_hub = nullptr;
    int2* d_htab = nullptr;
    std::vector<int32_t> p_hub;  // vertices whose table changed since the last upload (edge removals)
};

// (root, vertex) pairs a level of the deletion's whole-GPU delta pass hands
// to one-workgroup-per-vertex sums; a level with more does the rest one
// wavefront per vertex.
static const int32_t kDmGiantCap = 65536;

// Vertices with more initial neighbours than this get a sorted index (the
// update kernel looks neighbours up in it instead of scanning hub lists).
static const int32_t kHubIndexMin = 512;
// Lookup table of a high-degree vertex: 2^logT int2 slots {neighbour + 1,
// position}, at most a quarter used, in buckets of 8 slots (64 bytes, one
// cache line, the table 8-slot aligned); a neighbour takes the first free
// slot of bucket hash(w) or of the buckets after it.  A device lookup reads
// a whole bucket at once and stops at a bucket with a free slot.
static inline int hub_table_log(int64_t d)
{
    int logT = 3;
    while ((1LL << logT) < 4 * d) logT++;
    return logT;
}
static inline void hub_table_put(int2* tab, int logT, int32_t w, int32_t pos)
{
    const uint32_t mask = (1u << (logT - 3)) - 1u;
    uint32_t b = ((uint32_t)w * 0x9E3779B1u) >> (32 - (logT - 3));
    for (;;) {
        int2* bk = tab + ((size_t)b << 3);
        for (int i = 0; i < 8; i++)
            if (bk[i].x == 0) {
                bk[i] = int2{w + 1, pos};
                return;
            }
        b = (b + 1) & mask;
    }
}
static const int64_t kBucketHubs = 512; // HUB_CAP of the kernels
static const int64_t kParkScan = 512 * 1024; // adjacency entries left to scan that hand an ascent level over

static int32_t slack_for(int32_t d) { return 4 + d / 16; }

static sbcStatus graph_init(stinger* G, int64_t nv, const int64_t* rowptr, const int64_t* col)
{
    if (nv < 1 || nv >= INT32_MAX) return set_err(SBC_ERR_ARG, "NV %ld out of range", (long)nv);
    if (rowptr[0] != 0) return set_err(SBC_ERR_ARG, "vertexPointerArray[0] must be 0");
    G->nv = nv;
    G->start.resize(nv);
    G->deg.resize(nv);
    G->deg0.resize(nv);
    G->cap.resize(nv);
    int64_t tot = 0;
    for (int64_t v = 0; v < nv; v++) {
        int64_t d = rowptr[v + 1] - rowptr[v];
        if (d < 0 || d > INT32_MAX / 2) return set_err(SBC_ERR_ARG, "bad degree at vertex %ld", (long)v);
        G->start[v] = (int32_t)tot;
        G->deg[v] = (int32_t)d;
        G->deg0[v] = (int32_t)d;
        G->cap[v] = (int32_t)d + slack_for((int32_t)d);
        tot += G->cap[v];
        if (tot >= INT32_MAX) return set_err(SBC_ERR_ARG, "graph too large for int32 slot ids");
    }
    G->total_cap = tot;
    G->col.assign(tot, 0);
    G->seq.assign(tot, kNoSeq);
    for (int64_t v = 0; v < nv; v++) {
        for (int64_t e = rowptr[v]; e < rowptr[v + 1]; e++) {
            int64_t x = col[e];
            if (x < 0 || x >= nv) return set_err(SBC_ERR_ARG, "edge %ld -> %ld out of range", (long)v, (long)x);
            G->col[G->start[v] + (e - rowptr[v])] = (int32_t)x;
        }
    }
    // lookup tables of high-degree vertices (neighbour -> position in the list)
    G->hub.assign(nv, int4{0, 0, 0, 0});
    G->htab.clear();
    for (int64_t v = 0; v < nv; v++) {
        int32_t d = G->deg[v];
        if (d <= kHubIndexMin || d >= (1 << 22)) continue;
        const int logT = hub_table_log(d);
        const int64_t off = ((int64_t)G->htab.size() + 7) & ~(int64_t)7;
        if (off + (1LL << logT) >= INT32_MAX) break;
        G->htab.resize(off + ((size_t)1 << logT), int2{0, 0});
        for (int32_t i = 0; i < d; i++) hub_table_put(&G->htab[off], logT, G->col[G->start[v] + i], i);
        G->hub[v] = int4{(int32_t)off, d, logT, 0};
    }
    G->full_dirty = true;
    return SBC_OK;
}

// Rebuilds the slot layout with fresh slack for every vertex (when a vertex
// runs out of append room); the device image is re-uploaded on next flush.
static void graph_relayout(stinger* G)
{
    std::vector<int32_t> nstart(G->nv), ncap(G->nv);
    int64_t tot = 0;
    for (int64_t v = 0; v < G->nv; v++) {
        nstart[v] = (int32_t)tot;
        int32_t want = G->deg[v] + slack_for(G->deg[v]);
        ncap[v] = G->cap[v] > want ? G->cap[v] : want;
        tot += ncap[v];
    }
    std::vector<int32_t> ncol(tot, 0), nseq(tot, kNoSeq);
    for (int64_t v = 0; v < G->nv; v++) {
        memcpy(&ncol[nstart[v]], &G->col[G->start[v]], sizeof(int32_t) * G->deg[v]);
        memcpy(&nseq[nstart[v]], &G->seq[G->start[v]], sizeof(int32_t) * G->deg[v]);
    }
    G->start.swap(nstart);
    G->cap.swap(ncap);
    G->col.swap(ncol);
    G->seq.swap(nseq);
    G->total_cap = tot;
    G->full_dirty = true;
    G->p_slot.clear(); G->p_col.clear(); G->p_seq.clear(); G->p_v.clear();
}

static bool graph_has_edge(const stinger* G, int32_t u, int32_t v)
{
    if (G->deg[v] < G->deg[u]) { int32_t t = u; u = v; v = t; }
    const int32_t* a = &G->col[G->start[u]];
    for (int32_t i = 0; i < G->deg[u]; i++)
        if (a[i] == v) return true;
    return false;
}

// Checks a batch of insertions before anything is touched: the first
// insertion (in batch order) that is out of range, a self loop, a repeat of
// an earlier one or an existing edge fails the batch.  The host mirror is
// read at random (a window's endpoints are spread over arrays of hundreds of
// MB), so the passes prefetch what the next pass and the window's host
// staging read: the misses overlap instead of queueing one after another.
static sbcStatus validate_batch(const stinger* G, const int64_t* u, const int64_t* v, int64_t count)
{
    int64_t limit = count, bad_range = -1, bad_loop = -1;
    for (int64_t i = 0; i < count; i++) {
        const int64_t a = u[i], b = v[i];
        if (a < 0 || b < 0 || a >= G->nv || b >= G->nv) { bad_range = i; limit = i; break; }
        if (a == b) { bad_loop = i; limit = i; break; }
        for (int64_t x : {a, b}) {
            __builtin_prefetch(&G->deg[x]);
            __builtin_prefetch(&G->start[x]);
            __builtin_prefetch(&G->cap[x]);
            __builtin_prefetch(&G->deg0[x], 1);
            if ((int64_t)G->h_mark.size() == G->nv) __builtin_prefetch(&G->h_mark[x], 1);
            if ((int64_t)G->h_cnt.size() == G->nv) __builtin_prefetch(&G->h_cnt[x], 1);
        }
    }
    // the smaller endpoint's list is scanned; both lists' ends are written
    for (int64_t i = 0; i < limit; i++) {
        const int32_t a = (int32_t)u[i], b = (int32_t)v[i];
        const int32_t s = G->deg[b] < G->deg[a] ? b : a;
        if ((int64_t)G->start[s] < (int64_t)G->col.size()) __builtin_prefetch(&G->col[G->start[s]]);
        for (int32_t x : {a, b}) {
            const int64_t end = (int64_t)G->start[x] + G->deg[x];
            if (end < (int64_t)G->col.size()) {
                __builtin_prefetch(&G->col[end], 1);
                __builtin_prefetch(&G->seq[end], 1);
            }
        }
    }
    // repeats: sorted (edge, position) pairs; in each run of one edge every
    // position but the first repeats it
    std::vector<std::pair<uint64_t, int64_t>> keys((size_t)limit);
    for (int64_t i = 0; i < limit; i++) {
        const uint64_t a = (uint64_t)u[i], b = (uint64_t)v[i];
        keys[(size_t)i] = {(a < b ? a : b) << 32 | (a < b ? b : a), i};
    }
    std::sort(keys.begin(), keys.end());
    int64_t bad_dup = -1;
    for (size_t k = 1; k < keys.size(); k++)
        if (keys[k].first == keys[k - 1].first && (bad_dup < 0 || keys[k].second < bad_dup)) bad_dup = keys[k].second;
    const int64_t edge_limit = bad_dup >= 0 ? bad_dup : limit;
    int64_t bad_edge = -1;
    for (int64_t i = 0; i < edge_limit; i++)
        if (graph_has_edge(G, (int32_t)u[i], (int32_t)v[i])) { bad_edge = i; break; }
    int64_t bad_exists = bad_dup;
    if (bad_edge >= 0 && (bad_exists < 0 || bad_edge < bad_exists)) bad_exists = bad_edge;
    if (bad_exists >= 0)
        return set_err(SBC_ERR_EXISTS, "insertion %ld (%ld,%ld) duplicates an edge", (long)bad_exists,
                       (long)u[bad_exists], (long)v[bad_exists]);
    if (bad_range >= 0) return set_err(SBC_ERR_ARG, "insertion %ld out of range", (long)bad_range);
    if (bad_loop >= 0) return set_err(SBC_ERR_EXISTS, "insertion %ld is a self loop", (long)bad_loop);
    if ((int64_t)G->seq_next + count >= INT32_MAX) return set_err(SBC_ERR_ARG, "sequence space exhausted");
    return SBC_OK;
}

// Appends v to u's list in the host mirror; returns the slot.
static int32_t graph_push(stinger* G, int32_t u, int32_t v, int32_t seq)
{
    if (G->deg[u] == G->cap[u]) graph_relayout(G);
    int32_t slot = G->start[u] + G->deg[u];
    G->col[slot] = v;
    G->seq[slot] = seq;
    G->deg[u]++;
    return slot;
}

static SbcVInfo vinfo_of(const stinger* G, int32_t v)
{
    SbcVInfo x;
    x.start = G->start[v];
    x.deg = G->deg[v];
    x.deg0 = G->deg0[v];
    x.pad = 0;
    return x;
}

// Sends the tables that edge removals rebuilt (entry and table slots of each
// vertex in p_hub) and waits for the copies, so the host tables may change
// again right after.
static sbcStatus hub_flush(stinger* G, hipStream_t st)
{
    if (G->p_hub.empty() || !G->d_hub) {
        G->p_hub.clear();
        return SBC_OK;
    }
    for (int32_t v : G->p_hub) {
        const int4 hb = G->hub[v];
        HIPCHK(hipMemcpyAsync(G->d_hub + v, &G->hub[v], sizeof(int4), hipMemcpyHostToDevice, st));
        if (hb.z > 0)
            HIPCHK(hipMemcpyAsync(G->d_htab + hb.x, &G->htab[hb.x], sizeof(int2) * ((size_t)1 << hb.z),
                                  hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    G->p_hub.clear();
    return SBC_OK;
}

static sbcStatus graph_upload_full(stinger* G, hipStream_t st)
{
    HIPCHK(hipStreamSynchronize(st));
    if (G->d_cap < G->total_cap || !G->d_vinfo) {
        if (G->d_col) HIPCHK(hipFree(G->d_col));
        if (G->d_seq) HIPCHK(hipFree(G->d_seq));
        if (!G->d_vinfo) HIPCHK(hipMalloc(&G->d_vinfo, sizeof(SbcVInfo) * G->nv));
        int64_t cap = G->total_cap + G->total_cap / 8 + 64;
        HIPCHK(hipMalloc(&G->d_col, sizeof(int32_t) * cap));
        HIPCHK(hipMalloc(&G->d_seq, sizeof(int32_t) * cap));
        G->d_cap = cap;
        G->generation++;
    }
    if (!G->d_hub) {
        HIPCHK(hipMalloc(&G->d_hub, sizeof(int4) * G->nv));
        HIPCHK(hipMalloc(&G->d_htab, sizeof(int2) * (G->htab.size() ? G->htab.size() : 1)));
        HIPCHK(hipMemcpyAsync(G->d_hub, G->hub.data(), sizeof(int4) * G->nv, hipMemcpyHostToDevice, st));
        if (!G->htab.empty())
            HIPCHK(hipMemcpyAsync(G->d_htab, G->htab.data(), sizeof(int2) * G->htab.size(), hipMemcpyHostToDevice,
                                  st));
    }
    std::vector<SbcVInfo> vi(G->nv);
    for (int64_t v = 0; v < G->nv; v++) vi[v] = vinfo_of(G, (int32_t)v);
    HIPCHK(hipMemcpyAsync(G->d_vinfo, vi.data(), sizeof(SbcVInfo) * G->nv, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(G->d_col, G->col.data(), sizeof(int32_t) * G->total_cap, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(G->d_seq, G->seq.data(), sizeof(int32_t) * G->total_cap, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));  // vi is local; the mirrors may change after return
    G->full_dirty = false;
    G->p_slot.clear(); G->p_col.clear(); G->p_seq.clear(); G->p_v.clear();
    return hub_flush(G, st);
}

// Sends pending appends (legacy stinger_insert_edge path) to the device.
static sbcStatus graph_flush(stinger* G, hipStream_t st)
{
    if (G->full_dirty) return graph_upload_full(G, st);
    size_t ns = G->p_slot.size(), nvu = G->p_v.size();
    size_t i = 0, j = 0;
    while (i < ns || j < nvu) {
        SbcAppend A;
        memset(&A, 0, sizeof(A));
        A.vinfo = G->d_vinfo;
        A.col = G->d_col;
        A.seq = G->d_seq;
        while (i < ns && A.nslots < SBC_APPEND_INLINE) {
            A.slot[A.nslots] = G->p_slot[i];
            A.scol[A.nslots] = G->p_col[i];
            A.sseq[A.nslots] = G->p_seq[i];
            A.nslots++;
            i++;
        }
        while (j < nvu && A.nvupd < SBC_APPEND_INLINE) {
            A.vv[A.nvupd] = G->p_v[j];
            A.vi[A.nvupd] = vinfo_of(G, G->p_v[j]);
            A.nvupd++;
            j++;
        }
        HIPCHK(sbc_launch_append(&A, SBC_APPEND_INLINE, st));
    }
    G->p_slot.clear(); G->p_col.clear(); G->p_seq.clear(); G->p_v.clear();
    return hub_flush(G, st);
}

static void graph_free(stinger* G)
{
    if (!G) return;
    if (G->d_vinfo) (void)hipFree(G->d_vinfo);
    if (G->d_col) (void)hipFree(G->d_col);
    if (G->d_seq) (void)hipFree(G->d_seq);
    if (G->d_hub) (void)hipFree(G->d_hub);
    if (G->d_htab) (void)hipFree(G->d_htab);
    delete G;
}

extern "C" void CreateStingerFromCSR(csrGraph* csr, struct stinger** stingerGraph)
{
    stinger* G = new stinger();
    if (!csr || graph_init(G, csr->NV, csr->vertexPointerArray, csr->edgeArray) != SBC_OK) {
        delete G;
        legacy_abort("CreateStingerFromCSR");
    }
    *stingerGraph = G;
}

extern "C" void ParseDimacsGraph(char initial_graph_name[1024], struct stinger** stingerGraph, int64_t* NV,
                                 int64_t* NE)
{
    sbcHostGraph* h = sbcReadDimacs(initial_graph_name);
    if (!h) {
        set_err(SBC_ERR_ARG, "cannot parse DIMACS file %s", initial_graph_name);
        legacy_abort("ParseDimacsGraph");
    }
    stinger* G = new stinger();
    if (graph_init(G, h->nv, h->rowptr, h->col) != SBC_OK) legacy_abort("ParseDimacsGraph");
    *NV = h->nv;
    *NE = h->ne;
    sbcHostGraphFree(h);
    *stingerGraph = G;
}

extern "C" int stinger_insert_edge(struct stinger* G, int64_t type, int64_t from, int64_t to, int64_t weight,
                                   int64_t timestamp)
{
    (void)type; (void)weight; (void)timestamp;
    if (!G || from == to || from < 0 || to < 0 || from >= G->nv || to >= G->nv) return -1;
    const int32_t* a = &G->col[G->start[from]];
    for (int32_t i = 0; i < G->deg[from]; i++)
        if (a[i] == to) return 0;
    bool relaid = G->deg[from] == G->cap[from];
    int32_t slot = graph_push(G, (int32_t)from, (int32_t)to, G->seq_next++);
    G->deg0[from] = G->deg[from]; // visible to every later update
    if (!relaid && !G->full_dirty) {
        G->p_slot.push_back(slot);
        G->p_col.push_back((int32_t)to);
        G->p_seq.push_back(G->seq[slot]);
        G->p_v.push_back((int32_t)from);
    }
    return 1;
}

// Drops position p of v's list, keeping the order of the rest (the
// reference's STINGER leaves a hole there and skips it in traversals, so the
// remaining neighbours are visited in the same order).  A lookup table over
// v's list is rebuilt when p lies in its indexed part.
static void graph_remove_at(stinger* G, int32_t v, int32_t p)
{
    const int32_t s0 = G->start[v], d = G->deg[v];
    for (int32_t i = p; i + 1 < d; i++) {
        G->col[s0 + i] = G->col[s0 + i + 1];
        G->seq[s0 + i] = G->seq[s0 + i + 1];
    }
    G->col[s0 + d - 1] = 0;
    G->seq[s0 + d - 1] = kNoSeq;
    G->deg[v] = d - 1;
    G->deg0[v] = d - 1;  // every remaining neighbour is visible
    if (!G->full_dirty) {
        for (int32_t i = p; i < d; i++) {
            G->p_slot.push_back(s0 + i);
            G->p_col.push_back(G->col[s0 + i]);
            G->p_seq.push_back(G->seq[s0 + i]);
        }
        G->p_v.push_back(v);
    }
    int4& hb = G->hub[v];
    if (hb.y > 0 && p < hb.y) {
        const int32_t n = hb.y - 1;  // indexed entries left, at positions 0..n-1
        const uint32_t mask = (1u << hb.z) - 1u;
        int2* tab = &G->htab[hb.x];
        for (uint32_t i = 0; i <= mask; i++) tab[i] = int2{0, 0};
        for (int32_t i = 0; i < n; i++) hub_table_put(tab, hb.z, G->col[s0 + i], i);
        hb.y = n;
        G->p_hub.push_back(v);
    }
}

extern "C" int stinger_remove_edge(struct stinger* G, int64_t type, int64_t from, int64_t to)
{
    (void)type;
    if (!G || from == to || from < 0 || to < 0 || from >= G->nv || to >= G->nv) return -1;
    const int32_t* a = &G->col[G->start[from]];
    for (int32_t i = 0; i < G->deg[from]; i++)
        if (a[i] == to) {
            graph_remove_at(G, (int32_t)from, i);
            return 1;
        }
    return -1;
}

extern "C" int64_t stinger_outdegree(struct stinger* G, int64_t v)
{
    if (!G || v < 0 || v >= G->nv) return 0;
    return G->deg[v];
}

extern "C" struct stinger* stinger_free_all(struct stinger* S)
{
    graph_free(S);
    return nullptr;
}

// ------------------------------------------------------------ forest

struct sbcForestDev {
    int64_t nv = 0, K = 0;
    int32_t lcap = 0;
    bool exact = false;
    bool initialized = false;
    std::vector<int64_t> roots;
    int32_t* d_roots = nullptr;
    int32_t* d_dist = nullptr;
    uint8_t* d_sb = nullptr;     // state bytes, sb_stride per root
    int64_t sb_stride = 0;
    SbcVRec* d_vr = nullptr;     // [K][nv] path counts, dependencies, update scratch
    double* d_bc = nullptr;
    int32_t* d_list = nullptr;
    int32_t* d_lvl = nullptr;
    int32_t* d_err = nullptr;
    // readouts: the forest's own stream, and the point on the last updating
    // context's stream after which the forest is final
    hipStream_t rs = nullptr;
    hipEvent_t done = nullptr;
    bool done_set = false;
};

static void legacy_refresh_bc(bcForest* forest, const char* where);

// Records that the forest's last enqueued update ends at this point of st.
static sbcStatus forest_mark(sbcForestDev* F, hipStream_t st)
{
    HIPCHK(hipEventRecord(F->done, st));
    F->done_set = true;
    return SBC_OK;
}

// Device -> host copy of forest data, ordered after the forest's last update
// (on the forest's own stream: no device-wide synchronisation, so graphs that
// other threads update meanwhile are not waited for).
static sbcStatus forest_read(sbcForestDev* F, void* dst, const void* src, size_t bytes)
{
    if (F->done_set) HIPCHK(hipStreamWaitEvent(F->rs, F->done, 0));
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, F->rs));
    HIPCHK(hipStreamSynchronize(F->rs));
    return SBC_OK;
}

static sbcStatus forest_write(sbcForestDev* F, void* dst, const void* src, size_t bytes)
{
    if (F->done_set) HIPCHK(hipStreamWaitEvent(F->rs, F->done, 0));
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, F->rs));
    HIPCHK(hipStreamSynchronize(F->rs));
    return SBC_OK;
}

static int32_t lcap_for(int64_t nv) { return (int32_t)(nv + 2 < 4096 ? nv + 2 : 4096); }

extern "C" int64_t streamingBCForestFootprint(int64_t NV, int64_t NK)
{
    int64_t per = 4 + 1 + (int64_t)sizeof(SbcVRec) + 4; // dist, state byte, path record, list
    return NK * NV * per + NK * 2 * (int64_t)lcap_for(NV) * 4 + NV * 8 + NK * 4 + 64;
}

extern "C" int64_t streamingBCResidentFootprint(const int64_t* rowptr, int64_t NV, int64_t NK, int32_t num_cu)
{
    if (!rowptr || NV <= 0 || NK <= 0) return -1;
    if (num_cu <= 0) num_cu = 256;
    int64_t slots = 0, htab = 0;
    for (int64_t v = 0; v < NV; v++) {
        const int64_t d = rowptr[v + 1] - rowptr[v];
        slots += d + slack_for((int32_t)d);
        if (d > kHubIndexMin && d < (1 << 22)) {
            htab += ((int64_t)1 << hub_table_log(d)) + 7;
        }
    }
    int64_t graph = slots * 8 + NV * (int64_t)sizeof(SbcVInfo) + NV * 16 + htab * 8;
    // windowed-update scratch: hub term buckets for up to 4 workgroups per CU,
    // and the lists of the whole-GPU levels of handed-over items
    int64_t scratch = htab ? 4LL * num_cu * SBC_BUCKET_SLOTS * (int64_t)sizeof(SbcBucketEnt) : 0;
    scratch += 4 * (16 + 3 * NV + slots / 512 + 1);
    return streamingBCForestFootprint(NV, NK) + graph + scratch;
}

static void forest_free_dev(sbcForestDev* F)
{
    if (!F) return;
    if (F->done_set) (void)hipEventSynchronize(F->done);
    if (F->rs) (void)hipStreamSynchronize(F->rs);
    void* ptrs[] = {F->d_roots, F->d_dist, F->d_sb, F->d_vr, F->d_bc, F->d_list, F->d_lvl, F->d_err};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (F->done) (void)hipEventDestroy(F->done);
    if (F->rs) (void)hipStreamDestroy(F->rs);
    delete F;
}

template <class T>
static sbcStatus dalloc(T** p, size_t n)
{
    hipError_t e = hipMalloc((void**)p, sizeof(T) * (n ? n : 1));
    if (e != hipSuccess) {
        *p = nullptr;
        return set_err(SBC_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", sizeof(T) * n, hipGetErrorString(e));
    }
    return SBC_OK;
}

static sbcStatus forest_create(int64_t NV, int64_t NK, const int64_t* roots, bool exact, bcForest** out)
{
    if (NV < 1 || NV >= INT32_MAX / 4 || NK < 1 || NK > NV) return set_err(SBC_ERR_ARG, "bad NV/NK");
    sbcForestDev* F = new sbcForestDev();
    F->nv = NV;
    F->K = NK;
    F->lcap = lcap_for(NV);
    F->exact = exact;
    F->roots.resize(NK);
    std::vector<int32_t> r32(NK);
    for (int64_t i = 0; i < NK; i++) {
        int64_t r = exact ? i : roots[i];
        if (r < 0 || r >= NV) {
            forest_free_dev(F);
            return set_err(SBC_ERR_ARG, "root %ld out of range", (long)r);
        }
        F->roots[i] = r;
        r32[i] = (int32_t)r;
    }
    size_t kn = (size_t)NK * (size_t)NV;
    sbcStatus st = SBC_OK;
    F->sb_stride = (NV + 63) / 64 * 64;
    if ((st = dalloc(&F->d_roots, NK)) || (st = dalloc(&F->d_dist, kn)) ||
        (st = dalloc(&F->d_sb, (size_t)NK * (size_t)F->sb_stride)) || (st = dalloc(&F->d_vr, kn)) ||
        (st = dalloc(&F->d_bc, NV)) || (st = dalloc(&F->d_list, kn)) ||
        (st = dalloc(&F->d_lvl, (size_t)NK * 2 * F->lcap)) || (st = dalloc(&F->d_err, 4))) {
        forest_free_dev(F);
        return st;
    }
    hipError_t e = hipStreamCreateWithFlags(&F->rs, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&F->done, hipEventDisableTiming);
    if (e == hipSuccess)
        e = hipMemcpyAsync(F->d_roots, r32.data(), sizeof(int32_t) * NK, hipMemcpyHostToDevice, F->rs);
    if (e == hipSuccess) e = hipMemsetAsync(F->d_sb, 0, (size_t)NK * (size_t)F->sb_stride, F->rs);
    if (e == hipSuccess) e = hipMemsetAsync(F->d_vr, 0, sizeof(SbcVRec) * kn, F->rs);
    if (e == hipSuccess) e = hipMemsetAsync(F->d_bc, 0, sizeof(double) * NV, F->rs);
    if (e == hipSuccess) e = hipMemsetAsync(F->d_err, 0, sizeof(int32_t) * 4, F->rs);
    if (e == hipSuccess) e = hipStreamSynchronize(F->rs);
    if (e != hipSuccess) {
        forest_free_dev(F);
        return set_err(SBC_ERR_HIP, "forest init: %s", hipGetErrorString(e));
    }
    bcForest* f = (bcForest*)calloc(1, sizeof(bcForest));
    f->dev = F;
    f->NV = NV;
    f->totalBC = (double*)calloc((size_t)NV, sizeof(double));
    *out = f;
    return SBC_OK;
}

extern "C" bcForest* streamingBCCreateForestExact(int64_t NV)
{
    bcForest* f = nullptr;
    if (forest_create(NV, NV, nullptr, true, &f) != SBC_OK) legacy_abort("streamingBCCreateForestExact");
    return f;
}

extern "C" bcForest* streamingBCCreateForestApproximate(int64_t NV, int64_t NK, int64_t* roots)
{
    bcForest* f = nullptr;
    if (!roots) {
        set_err(SBC_ERR_ARG, "root array is NULL");
        legacy_abort("streamingBCCreateForestApproximate");
    }
    if (forest_create(NV, NK, roots, false, &f) != SBC_OK)
        legacy_abort("streamingBCCreateForestApproximate");
    return f;
}

static void forest_destroy(bcForestPtr* p)
{
    if (!p || !*p) return;
    forest_free_dev((*p)->dev);
    free((*p)->totalBC);
    free(*p);
    *p = nullptr;
}

extern "C" void streamingBCDeleteForestExact(bcForestPtr* deadForest) { forest_destroy(deadForest); }
extern "C" void streamingBCDeleteForestApproximate(bcForestPtr* deadForest, int64_t rootArraySize, int64_t* rootArray)
{
    (void)rootArraySize; (void)rootArray;
    forest_destroy(deadForest);
}

// ------------------------------------------------------------ context

// Kernel timing classes: the update launch (a window's graph: append, queue
// reset, window kernel; or a single insertion's window kernel), the static
// pass, the BC reduction, edge appends, and the whole-GPU kernels that finish
// handed-over items (with the resumed roots' launches) after an update launch.
enum { KSTAT_UPDATE = 0, KSTAT_STATIC = 1, KSTAT_BC = 2, KSTAT_APPEND = 3, KSTAT_WIDE = 4, KSTAT_N = 5 };

struct WinStage {
    // device window staging and its pinned host twins, one block each with
    // the same layout for a window of W insertions, so one copy moves it:
    // desc at 0 | ints at 64: wu, wv, slot, scol, sseq, vv at stride 4W |
    // vinfo at 64 + 96W.  The views below are set per call (stage_views).
    char* d_base = nullptr;
    SbcWinDesc* d_desc = nullptr;
    int32_t* d_ints = nullptr;
    SbcVInfo* d_vi = nullptr;
    void* h_buf[2] = {nullptr, nullptr};
    hipEvent_t h_done[2] = {nullptr, nullptr};
    int h_cur = 0;
    int cap = 0;                   // max insertions per window
};

struct GraphCache {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    const void* forest = nullptr;
    const void* g = nullptr;
    uint64_t generation = 0;
    int nwb = 0, grid = 0, heavy = 0, recompute = 0, W = 0, has_order = 0, with_rec = 0, park_min = 0;
    int run_cap = 0, hub_batch = 0;  // test hooks baked into the captured arguments
    int64_t park_scan = 0;
};

struct extraArraysPerThread {
    hipStream_t stream = nullptr;
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[KSTAT_N];
    int64_t launches[KSTAT_N] = {0, 0, 0, 0, 0};
    double total_ms[KSTAT_N] = {0, 0, 0, 0, 0};
    int32_t* d_qctr = nullptr;
    SbcItemRec* d_rec = nullptr;
    int64_t rec_cap = 0;       // entries
    int64_t last_rec_n = 0;    // entries written by the last legacy update
    int32_t* d_order = nullptr;
    // parked items: records, parked-root list, resume order and start items
    SbcWideItem* d_park = nullptr;
    int32_t* d_parked = nullptr;
    int32_t* d_order2 = nullptr;
    int32_t* d_start = nullptr;
    int64_t park_cap = 0;
    // finishing parked items while the window launch runs: a second stream
    // (default priority), the resumed root's own queue counters and parked
    // list, and the per-root flags the launch sets in host memory
    hipStream_t stream2 = nullptr;
    int32_t* d_qctr2 = nullptr;
    int32_t* d_parked2 = nullptr;
    // pinned host words read back every call (copies into pageable memory
    // are staged by the runtime and cost tens of microseconds more): the
    // kernel error flags, the scheduler's heavy-root count, a resumed root's
    // parked count, a hand-over's list counters after each level
    int32_t* h_err = nullptr;
    int32_t* h_heavy = nullptr;   // h_err + 1
    int32_t* h_np = nullptr;      // h_err + 2
    int32_t* h_wend = nullptr;    // h_err + 4, two words
    int32_t* h_suv = nullptr;     // pinned [2 * sched_cnt]: a scheduled window's endpoints
    int32_t* h_pflag = nullptr;   // pinned host memory, [K]; device views d_pflag / d_proot
    int32_t* h_proot = nullptr;
    int32_t* d_pflag = nullptr;
    int32_t* d_proot = nullptr;
    std::vector<int32_t> h_start;
    int64_t last_parked = 0;   // items finished by the wide kernels in the last batched call
    // device scheduling (streamingBCScheduleWindow): buffers and a valid-order flag
    int32_t* d_suv = nullptr;
    float* d_shist = nullptr;
    float* d_sX = nullptr;
    float* d_spred = nullptr;
    float* d_scost = nullptr;
    int32_t* d_sheavy = nullptr;
    int64_t sched_items = 0, sched_k = 0, sched_cnt = 0;
    float* d_hist = nullptr;     // per-root history feature kept on the device (streamingBCSetRootHistory)
    int64_t hist_k = 0;
    int64_t last_batch_items = 0;  // insertions of the last window of the last batched call
    bool order_ready = false;
    const void* order_forest = nullptr;  // forest the scheduled order belongs to
    int64_t order_cap = 0;
    int num_cu = 0;
    // optional per-item device timestamps, appended to the file named by
    // SBC_ITEM_TRACE after each batched call (diagnostics)
    int64_t* d_tprof = nullptr;
    int64_t tprof_cap = 0;
    int64_t* d_wtot = nullptr;   // [K][8] per-root work totals since the last stats reset
    // scratch of the parked items' whole-GPU levels: counters, heavy list, aux list
    int32_t* d_wl = nullptr;
    int64_t wl_nv = 0, wl_aux = 0;
    // the deletion's whole-GPU delta pass: a count and kDmGiantCap (root, vertex) pairs
    int32_t* d_dmg = nullptr;
    int64_t wtot_k = 0;
    // per-workgroup term buckets of the hub lookup pass
    SbcBucketEnt* d_bucket = nullptr;
    int64_t bucket_blocks = 0;
    int32_t* d_hubq = nullptr;  // per workgroup: hubs queued beyond the LDS queue
    int64_t hubq_blocks = 0;
    int64_t graphs_built = 0;    // window graphs built (one per change of launch shape or buffers)
    int64_t graph_launches = 0;  // windows launched as a graph replay
    int32_t* d_del = nullptr;   // edge deletion: 4 counts, then the affected root indices
    int64_t del_k = 0;
    // whole-GPU rebuild of one root: a level's big vertices, their term offsets and terms
    int32_t* d_big = nullptr;
    int64_t big_cap = 0;
    int64_t* d_toff = nullptr;
    double* d_terms = nullptr;
    int64_t terms_cap = 0;
    // whole-GPU rebuild of several roots: root list, list ends, segments, prefix
    int32_t* d_mrl = nullptr;
    int32_t* d_mfr = nullptr;
    SbcRootSeg* d_mseg = nullptr;
    int64_t* d_mpre = nullptr;
    int64_t mroot_cap = 0;
    WinStage win;
    GraphCache gc;
    int refs = 1;
};

static sbcStatus ctx_init(extraArraysPerThread* C)
{
    HIPCHK(hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking));
    HIPCHK(hipMalloc(&C->d_qctr, sizeof(int32_t) * 4));
    HIPCHK(hipMemsetAsync(C->d_qctr, 0, sizeof(int32_t) * 4, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    HIPCHK(hipHostMalloc((void**)&C->h_err, 8 * sizeof(int32_t)));
    C->h_heavy = C->h_err + 1;
    C->h_np = C->h_err + 2;
    C->h_wend = C->h_err + 4;
    memset(C->h_err, 0, 8 * sizeof(int32_t));
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    C->num_cu = prop.multiProcessorCount;
    return SBC_OK;
}

static void ctx_free(extraArraysPerThread* C)
{
    if (!C) return;
    if (C->stream) (void)hipStreamSynchronize(C->stream);
    for (auto& v : C->ev)
        for (auto& pr : v) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
    if (C->gc.exec) (void)hipGraphExecDestroy(C->gc.exec);
    if (C->gc.graph) (void)hipGraphDestroy(C->gc.graph);
    if (C->d_qctr) (void)hipFree(C->d_qctr);
    if (C->d_rec) (void)hipFree(C->d_rec);
    if (C->d_order) (void)hipFree(C->d_order);
    for (void* q : {(void*)C->d_park, (void*)C->d_parked, (void*)C->d_order2, (void*)C->d_start})
        if (q) (void)hipFree(q);
    if (C->stream2) (void)hipStreamSynchronize(C->stream2);
    for (void* q : {(void*)C->d_qctr2, (void*)C->d_parked2})
        if (q) (void)hipFree(q);
    for (void* q : {(void*)C->h_pflag, (void*)C->h_proot, (void*)C->h_err, (void*)C->h_suv})
        if (q) (void)hipHostFree(q);
    if (C->stream2) (void)hipStreamDestroy(C->stream2);
    if (C->d_tprof) (void)hipFree(C->d_tprof);
    if (C->d_wtot) (void)hipFree(C->d_wtot);
    if (C->d_wl) (void)hipFree(C->d_wl);
    if (C->d_dmg) (void)hipFree(C->d_dmg);
    if (C->d_bucket) (void)hipFree(C->d_bucket);
    if (C->d_hubq) (void)hipFree(C->d_hubq);
    if (C->d_del) (void)hipFree(C->d_del);
    for (void* q : {(void*)C->d_big, (void*)C->d_toff, (void*)C->d_terms, (void*)C->d_mrl, (void*)C->d_mfr,
                    (void*)C->d_mseg, (void*)C->d_mpre})
        if (q) (void)hipFree(q);
    for (void* q : {(void*)C->d_suv, (void*)C->d_shist, (void*)C->d_sX, (void*)C->d_spred, (void*)C->d_scost,
                    (void*)C->d_sheavy, (void*)C->d_hist})
        if (q) (void)hipFree(q);
    if (C->win.d_base) (void)hipFree(C->win.d_base);
    for (int i = 0; i < 2; i++) {
        if (C->win.h_buf[i]) (void)hipHostFree(C->win.h_buf[i]);
        if (C->win.h_done[i]) (void)hipEventDestroy(C->win.h_done[i]);
    }
    if (C->stream) (void)hipStreamDestroy(C->stream);
    delete C;
}

extern "C" extraArraysPerThread** streamingBCCreateAuxilary(int64_t threadCount, int64_t NV)
{
    (void)NV;
    if (threadCount < 1) threadCount = 1;
    extraArraysPerThread* C = new extraArraysPerThread();
    if (ctx_init(C) != SBC_OK) {
        ctx_free(C);
        legacy_abort("streamingBCCreateAuxilary");
    }
    extraArraysPerThread** arr = (extraArraysPerThread**)calloc((size_t)threadCount, sizeof(*arr));
    for (int64_t i = 0; i < threadCount; i++) arr[i] = C;
    return arr;
}

extern "C" void streamingBCDeleteAuxilary(extraArraysPerThread** parallelExtra, int64_t threadCount, int64_t NV)
{
    (void)threadCount; (void)NV;
    if (!parallelExtra) return;
    ctx_free(parallelExtra[0]);
    free(parallelExtra);
}

static sbcStatus ev_begin(extraArraysPerThread* C, int k, hipEvent_t* e0)
{
    *e0 = nullptr;
    if (!C->timing) return SBC_OK;
    HIPCHK(hipEventCreate(e0));
    HIPCHK(hipEventRecord(*e0, C->stream));
    return SBC_OK;
}

static sbcStatus ev_end(extraArraysPerThread* C, int k, hipEvent_t e0)
{
    C->launches[k]++;
    if (!C->timing || !e0) return SBC_OK;
    hipEvent_t e1;
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e1, C->stream));
    C->ev[k].push_back({e0, e1});
    return SBC_OK;
}

static sbcStatus resolve_events(extraArraysPerThread* C)
{
    HIPCHK(hipStreamSynchronize(C->stream));
    for (int k = 0; k < KSTAT_N; k++) {
        for (auto& pr : C->ev[k]) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, pr.first, pr.second));
            C->total_ms[k] += ms;
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
        C->ev[k].clear();
    }
    return SBC_OK;
}

extern "C" sbcStatus streamingBCKernelStats(extraArraysPerThread** eAPT, int32_t kernel, int64_t* launches,
                                           double* total_ms)
{
    if (!eAPT || !eAPT[0] || kernel < 0 || kernel >= KSTAT_N) return set_err(SBC_ERR_ARG, "bad stats query");
    extraArraysPerThread* C = eAPT[0];
    sbcStatus st = resolve_events(C);
    if (st) return st;
    if (launches) *launches = C->launches[kernel];
    if (total_ms) *total_ms = C->total_ms[kernel];
    return SBC_OK;
}

extern "C" sbcStatus streamingBCResetStats(extraArraysPerThread** eAPT, int32_t enable_timing)
{
    if (!eAPT || !eAPT[0]) return set_err(SBC_ERR_ARG, "null context");
    extraArraysPerThread* C = eAPT[0];
    sbcStatus st = resolve_events(C);
    if (st) return st;
    for (int k = 0; k < KSTAT_N; k++) { C->launches[k] = 0; C->total_ms[k] = 0; }
    C->timing = enable_timing != 0;
    if (C->d_wtot) {
        HIPCHK(hipMemsetAsync(C->d_wtot, 0, sizeof(int64_t) * 8 * (size_t)C->wtot_k, C->stream));
        HIPCHK(hipStreamSynchronize(C->stream));
    }
    return SBC_OK;
}

extern "C" sbcStatus streamingBCSynchronize(extraArraysPerThread** eAPT)
{
    if (!eAPT || !eAPT[0]) return set_err(SBC_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(eAPT[0]->stream));
    return SBC_OK;
}

static sbcStatus ensure_rec(extraArraysPerThread* C, int64_t n)
{
    if (C->rec_cap >= n) return SBC_OK;
    HIPCHK(hipStreamSynchronize(C->stream));
    if (C->d_rec) HIPCHK(hipFree(C->d_rec));
    C->d_rec = nullptr;
    sbcStatus st = dalloc(&C->d_rec, (size_t)n);
    if (st) return st;
    C->rec_cap = n;
    C->gc.with_rec = -1; // invalidate cached graph
    return SBC_OK;
}

// Hub lookup scratch for `grid` workgroups (only graphs with lookup tables).
static sbcStatus ensure_bucket(extraArraysPerThread* C, const stinger* G, int64_t grid)
{
    if (G->htab.empty() || C->bucket_blocks >= grid) return SBC_OK;
    HIPCHK(hipStreamSynchronize(C->stream));
    if (C->d_bucket) HIPCHK(hipFree(C->d_bucket));
    C->d_bucket = nullptr;
    sbcStatus st = dalloc(&C->d_bucket, (size_t)grid * SBC_BUCKET_SLOTS);
    if (st) return st;
    C->bucket_blocks = grid;
    C->gc.with_rec = -1; // cached graph holds the old pointer
    return SBC_OK;
}

static sbcStatus ensure_hubq(extraArraysPerThread* C, int64_t grid)
{
    if (C->hubq_blocks >= grid) return SBC_OK;
    HIPCHK(hipStreamSynchronize(C->stream));
    if (C->d_hubq) HIPCHK(hipFree(C->d_hubq));
    C->d_hubq = nullptr;
    sbcStatus st = dalloc(&C->d_hubq, (size_t)grid * SBC_HUBQ_CAP);
    if (st) return st;
    C->hubq_blocks = grid;
    C->gc.with_rec = -1; // cached graph holds the old pointer
    return SBC_OK;
}

static SbcParams base_params(const sbcForestDev* F, const stinger* G, extraArraysPerThread* C)
{
    SbcParams P;
    memset(&P, 0, sizeof(P));
    P.vinfo = G->d_vinfo;
    P.col = G->d_col;
    P.col_cap = G->d_cap;
    P.seq = G->d_seq;
    P.hub = G->d_hub;
    P.htab = G->d_htab;
    P.bucket = G->htab.empty() ? nullptr : C->d_bucket;
    P.nv = (int32_t)F->nv;
    P.K = (int32_t)F->K;
    P.lcap = F->lcap;
    P.dist = F->d_dist;
    P.sb = F->d_sb;
    P.sb_stride = F->sb_stride;
    P.vr = F->d_vr;
    P.bc = F->d_bc;
    P.list = F->d_list;
    P.lvl = F->d_lvl;
    P.roots = F->d_roots;
    P.qctr = C->d_qctr;
    P.err = F->d_err;
    P.bc_atomic = 1;
    P.hubq = C->d_hubq;
    P.park = nullptr;
    P.parked = nullptr;
    P.park_min = 0;
    P.park_scan = 0;
    P.start_item = nullptr;
    P.qlen = (int32_t)F->K;
    // test hook: a small run capacity forces the flat hub scan's overflow path
    const char* rc = getenv("SBC_HUB_RUN_CAP");
    P.run_cap = rc ? atoi(rc) : 0;
    // test hook: fewer hubs per pass forces the batched hub passes
    const char* hb = getenv("SBC_HUB_BATCH");
    P.hub_batch = hb ? atoi(hb) : 0;
    // test hook: a smaller giant-vertex size of the whole-GPU ascent levels
    const char* wg = getenv("SBC_WIDE_GIANT");
    P.wide_giant = wg ? atoi(wg) : 0;
    return P;
}

// End of an update call on forest F: marks where the forest is final for
// readouts and reports a kernel-side error.
static sbcStatus check_kernel_err(sbcForestDev* F, extraArraysPerThread* C)
{
    sbcStatus st0 = forest_mark(F, C->stream);
    if (st0) return st0;
    // into pinned memory: a copy into pageable memory is staged by the
    // runtime and costs a few tens of microseconds more per call
    HIPCHK(hipMemcpyAsync(C->h_err, F->d_err, sizeof(int32_t), hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    const int32_t err = *C->h_err;
    if (err & (SBC_KERR_LEVELS | SBC_KERR_WIDE)) {
        int32_t z = 0;
        HIPCHK(hipMemcpyAsync(F->d_err, &z, sizeof(int32_t), hipMemcpyHostToDevice, C->stream));
        HIPCHK(hipStreamSynchronize(C->stream));
        if (err & SBC_KERR_WIDE) return set_err(SBC_ERR_STATE, "whole-GPU level scratch overflowed");
        return set_err(SBC_ERR_LEVELS, "BFS depth exceeds the level capacity %d", F->lcap);
    }
    return SBC_OK;
}

// ------------------------------------------------------------ static pass

// The static pass over P's roots (all, or the list in P.sroots), one
// workgroup per root: the LDS-bitmap form when the vertex set fits its
// bitmap, else the per-vertex form.  SBC_STATIC_FORM=vertex forces the
// per-vertex form (both give the same trees).
static hipError_t launch_static(const SbcParams& P, int grid, hipStream_t st)
{
    const char* f = getenv("SBC_STATIC_FORM");
    const bool vertex = f && strcmp(f, "vertex") == 0;
    if (!vertex && P.nv <= sbc_static_bm_max_nv()) return sbc_launch_static_bm(&P, grid, st);
    return sbc_launch_static(&P, 16, grid, st);
}

static sbcStatus run_static(bcForest* forest, stinger* G, extraArraysPerThread* C)
{
    if (!forest || !G || !C) return set_err(SBC_ERR_ARG, "null argument");
    sbcForestDev* F = forest->dev;
    if (F->nv != G->nv) return set_err(SBC_ERR_ARG, "forest NV %ld != graph NV %ld", (long)F->nv, (long)G->nv);
    sbcStatus st = graph_flush(G, C->stream);
    if (st) return st;
    SbcParams P = base_params(F, G, C);
    HIPCHK(sbc_launch_reset(C->d_qctr, 4, C->stream));
    hipEvent_t e0;
    if ((st = ev_begin(C, KSTAT_STATIC, &e0))) return st;
    int grid = (int)(F->K < 4LL * C->num_cu ? F->K : 4LL * C->num_cu);
    HIPCHK(launch_static(P, grid, C->stream));
    if ((st = ev_end(C, KSTAT_STATIC, e0))) return st;
    if ((st = ev_begin(C, KSTAT_BC, &e0))) return st;
    HIPCHK(sbc_launch_bc_reduce(&P, C->stream));
    if ((st = ev_end(C, KSTAT_BC, e0))) return st;
    if ((st = check_kernel_err(F, C))) return st;
    F->initialized = true;
    return SBC_OK;
}

extern "C" void streamingBCInitStaticExact(bcForest* forest, struct stinger* stingerGraph, int64_t NT,
                                           extraArraysPerThread** auxilary)
{
    (void)NT;
    if (!auxilary || run_static(forest, stingerGraph, auxilary[0]) != SBC_OK) legacy_abort("streamingBCInitStaticExact");
    legacy_refresh_bc(forest, "streamingBCInitStaticExact");
}

extern "C" void streamingBCInitStaticApproximate(bcForest* forest, struct stinger* stingerGraph, int64_t NT,
                                                 extraArraysPerThread** auxilary, int64_t NK, int64_t* roots)
{
    (void)NT;
    if (!forest || NK != forest->dev->K) {
        set_err(SBC_ERR_ARG, "NK does not match the forest");
        legacy_abort("streamingBCInitStaticApproximate");
    }
    for (int64_t i = 0; roots && i < NK; i++)
        if (roots[i] != forest->dev->roots[i]) {
            set_err(SBC_ERR_ARG, "root array differs from the one given at forest creation");
            legacy_abort("streamingBCInitStaticApproximate");
        }
    if (!auxilary || run_static(forest, stingerGraph, auxilary[0]) != SBC_OK)
        legacy_abort("streamingBCInitStaticApproximate");
    legacy_refresh_bc(forest, "streamingBCInitStaticApproximate");
}

// ------------------------------------------------------------ plain update

static sbcStatus ensure_park(extraArraysPerThread* C, int64_t K);
static sbcStatus wide_lists(extraArraysPerThread* C, const SbcParams& P, SbcWideLists* L);
static sbcStatus finish_parked_concurrent(sbcForestDev* F, extraArraysPerThread* C, const SbcParams& P, int nwb,
                                          int grid, int nitems, int64_t* total);

static sbcStatus run_insert_plain(bcForest* forest, stinger* G, extraArraysPerThread* C, int64_t u, int64_t v,
                                  uint32_t granularity)
{
    if (!forest || !G || !C) return set_err(SBC_ERR_ARG, "null argument");
    sbcForestDev* F = forest->dev;
    if (!F->initialized) return set_err(SBC_ERR_STATE, "static pass has not run");
    if (u < 0 || v < 0 || u >= F->nv || v >= F->nv || u == v) return set_err(SBC_ERR_ARG, "bad edge (%ld,%ld)", (long)u, (long)v);
    sbcStatus st = graph_flush(G, C->stream);
    if (st) return st;
    if ((st = ensure_rec(C, F->K))) return st;
    if ((st = ensure_bucket(C, G, 4LL * C->num_cu))) return st;
    if ((st = ensure_hubq(C, 4LL * C->num_cu))) return st;
    SbcParams P = base_params(F, G, C);
    P.u0 = (int32_t)u;
    P.v0 = (int32_t)v;
    P.rec = C->d_rec;
    int nwb = 4;
    int grid;
    if (granularity == FINE) {
        nwb = 16;
        P.heavy = (int32_t)F->K;
        grid = (int)(F->K < 2LL * C->num_cu ? F->K : 2LL * C->num_cu);
        // items whose descent or ascent grows large are finished by the
        // whole-GPU level kernels, with the batched form's thresholds
        if ((st = ensure_park(C, F->K))) return st;
        P.park = C->d_park;
        P.parked = C->d_parked;
        P.hpark_flag = C->d_pflag;
        P.hpark_root = C->d_proot;
        P.park_min = 4096;
        P.park_scan = kParkScan;
        const char* ps = getenv("SBC_PARK_SCAN");  // test hooks: the scan and frontier thresholds
        if (ps) P.park_scan = atoll(ps);
        const char* pm = getenv("SBC_PARK_MIN");
        if (pm && atoi(pm) > 0) P.park_min = atoi(pm);
    } else {
        P.heavy = 0;
        int64_t g = (F->K + nwb - 1) / nwb;
        grid = (int)(g < 4LL * C->num_cu ? g : 4LL * C->num_cu);
    }
    if (P.park) {
        // the whole-GPU levels' scratch before the launch; no flag is set now
        SbcWideLists L;
        if ((st = wide_lists(C, P, &L))) return st;
        memset(C->h_pflag, 0, sizeof(int32_t) * F->K);
    }
    HIPCHK(sbc_launch_reset(C->d_qctr, 4, C->stream));
    hipEvent_t e0;
    if ((st = ev_begin(C, KSTAT_UPDATE, &e0))) return st;
    HIPCHK(sbc_launch_window(&P, nwb, grid, C->stream));
    if ((st = ev_end(C, KSTAT_UPDATE, e0))) return st;
    int64_t parked = 0;
    // items the launch hands over are finished while it runs (the one
    // insertion leaves no item to resume)
    if (P.park && (st = finish_parked_concurrent(F, C, P, nwb, grid, 1, &parked))) return st;
    C->last_parked = parked;
    C->last_rec_n = F->K;
    return check_kernel_err(F, C);
}

extern "C" StreamingExtraInfo insertEdgeStreamingBC(bcForest* forest, struct stinger* sStinger, int64_t newU,
                                                    int64_t newV, int64_t* rootArrayForApproximation, int64_t NK,
                                                    int64_t NV, int64_t NT, extraArraysPerThread** eAPT,
                                                    uint32_t loadBalancing, uint32_t granularity)
{
    (void)rootArrayForApproximation; (void)NT; (void)loadBalancing;
    StreamingExtraInfo r = {0, 0, 0, 0};
    if (!forest || NK != forest->dev->K || NV != forest->dev->nv) {
        set_err(SBC_ERR_ARG, "NK/NV do not match the forest");
        legacy_abort("insertEdgeStreamingBC");
    }
    if (!eAPT || run_insert_plain(forest, sStinger, eAPT[0], newU, newV, granularity) != SBC_OK)
        legacy_abort("insertEdgeStreamingBC");
    legacy_refresh_bc(forest, "insertEdgeStreamingBC");
    return r;
}

// ------------------------------------------------------------ edge deletion

// The update after (u, v) left the graph in both directions.  Per root the
// reference dispatches on the endpoints' levels (src/streamingbc.c:258-416):
// same level, nothing to do; the deeper endpoint s keeps another neighbour one
// level up: the adjacent case (removeEdgeWithoutMovementBrandes,
// src/streamingbc_internal.c:607-868), in which no level changes, s and its
// descendants lose sigma[p] paths and dependencies are pulled again up to the
// root -- the update kernel's adjacent case with the added paths negated;
// otherwise the movement case (moveDownTreeBrandes, 880-1401), whose roots
// are rebuilt by the static pass over the post-deletion graph.  BC changes by
// the adjacent roots' atomic updates and by taking the movement roots out
// before their rebuild and back after.  counts: same-level, adjacent and
// movement roots, as the reference's StreamingExtraInfo.
static sbcStatus ensure_park(extraArraysPerThread* C, int64_t K);

// Movement-case roots up to this many per deletion are rebuilt one at a
// time by the whole GPU (one launch per level); more, one workgroup each.
static const int32_t kWideRebuildMax = 8;

static int32_t wide_rebuild_max()
{
    const char* e = getenv("SBC_WIDE_REBUILD_MAX");
    return e ? (int32_t)atoi(e) : kWideRebuildMax;
}

// More movement-case roots than that go to the several-roots whole-GPU
// rebuild on graphs of at least this many adjacency slots (R-MAT scale 20:
// one workgroup rebuilds a root in ~196 ms, the whole GPU in ~6.3 ms, and the
// several-roots form shares each level's launches), else one workgroup per
// root (Chung-Lu 2^15: 3.6 ms for ~41 roots against 5.4 ms level by level,
// whose round trips per level dominate on its many thin levels).
static const int64_t kWideGraphSlots = 4LL << 20;

static int64_t wide_graph_min()
{
    const char* e = getenv("SBC_WIDE_GRAPH_MIN");
    return e ? atoll(e) : kWideGraphSlots;
}

// The static pass of root ri spread over the whole GPU: host-driven, one
// launch per BFS level (the frontier's end read back after each) and per
// ascent level.  fr: one device counter.
static sbcStatus rebuild_root_wide(sbcForestDev* F, const stinger* G, extraArraysPerThread* C, const SbcParams& P,
                                   int32_t ri, int32_t* fr)
{
    // room for every vertex with more than 16384 neighbours
    const int64_t nbig_max = G->total_cap / 16384 + 2;
    if (C->big_cap < nbig_max) {
        HIPCHK(hipStreamSynchronize(C->stream));
        if (C->d_big) HIPCHK(hipFree(C->d_big));
        if (C->d_toff) HIPCHK(hipFree(C->d_toff));
        C->d_big = nullptr;
        C->d_toff = nullptr;
        sbcStatus st;
        if ((st = dalloc(&C->d_big, (size_t)nbig_max + 1)) || (st = dalloc(&C->d_toff, (size_t)nbig_max + 1))) return st;
        C->big_cap = nbig_max;
    }
    HIPCHK(sbc_launch_root_reset(&P, ri, fr, C->stream));
    std::vector<int32_t> lv = {0, 1};
    int32_t j = 0;
    for (;;) {
        const int32_t fs = lv[j], fe = lv[j + 1];
        HIPCHK(sbc_launch_root_bfs(&P, ri, fs, fe, j, fr, C->stream));
        int32_t ne = 0;
        HIPCHK(hipMemcpyAsync(&ne, fr, sizeof(ne), hipMemcpyDeviceToHost, C->stream));
        HIPCHK(hipStreamSynchronize(C->stream));
        if (ne == fe) break;
        HIPCHK(sbc_launch_root_settle(&P, ri, fe, ne, j, C->stream));
        j++;
        if (j + 1 >= F->lcap) return set_err(SBC_ERR_LEVELS, "BFS depth exceeds the level capacity %d", F->lcap);
        lv.push_back(ne);
    }
    // levels 0..j, level L = list[lv[L], lv[L+1]); the deepest has no children.
    // Vertices with more than 16384 neighbours are left by the ascent launch
    // and summed by sbc_launch_root_bigsum.
    for (int32_t L = j - 1; L >= 0; --L) {
        HIPCHK(sbc_launch_reset(C->d_big, 1, C->stream));
        HIPCHK(sbc_launch_root_ascent(&P, ri, lv[L], lv[L + 1], L, C->d_big, C->stream));
        int32_t nb = 0;
        HIPCHK(hipMemcpyAsync(&nb, C->d_big, sizeof(nb), hipMemcpyDeviceToHost, C->stream));
        HIPCHK(hipStreamSynchronize(C->stream));
        if (nb == 0) continue;
        if (nb < 0 || nb > C->big_cap) return set_err(SBC_ERR_STATE, "rebuild: big vertex count %d out of range", nb);
        std::vector<int32_t> ids((size_t)nb);
        HIPCHK(hipMemcpyAsync(ids.data(), C->d_big + 1, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, C->stream));
        HIPCHK(hipStreamSynchronize(C->stream));
        std::vector<int64_t> toff((size_t)nb + 1, 0);
        for (int32_t i = 0; i < nb; i++) {
            if (ids[i] < 0 || ids[i] >= G->nv) return set_err(SBC_ERR_STATE, "rebuild: big vertex out of range");
            toff[i + 1] = toff[i] + G->deg[ids[i]];
        }
        const int64_t T = toff[nb];
        if (C->terms_cap < T) {
            if (C->d_terms) HIPCHK(hipFree(C->d_terms));
            C->d_terms = nullptr;
            sbcStatus st;
            if ((st = dalloc(&C->d_terms, (size_t)T))) return st;
            C->terms_cap = T;
        }
        HIPCHK(hipMemcpyAsync(C->d_toff, toff.data(), sizeof(int64_t) * toff.size(), hipMemcpyHostToDevice, C->stream));
        HIPCHK(sbc_launch_root_bigsum(&P, ri, L, C->d_big, C->d_toff, nb, T, C->d_terms, C->stream));
        HIPCHK(hipStreamSynchronize(C->stream));  // toff is host stack data
    }
    return SBC_OK;
}

// The static pass of roots mr[0..n) together over the whole GPU: every
// launch takes one level of each root still growing (segments flattened by
// prefix sums); the list ends are read back after each BFS level.
static sbcStatus rebuild_roots_wide(sbcForestDev* F, extraArraysPerThread* C, const SbcParams& P,
                                    const std::vector<int32_t>& mr)
{
    const int n = (int)mr.size();
    if (C->mroot_cap < n) {
        HIPCHK(hipStreamSynchronize(C->stream));
        for (void* q : {(void*)C->d_mrl, (void*)C->d_mfr, (void*)C->d_mseg, (void*)C->d_mpre})
            if (q) HIPCHK(hipFree(q));
        C->d_mrl = C->d_mfr = nullptr;
        C->d_mseg = nullptr;
        C->d_mpre = nullptr;
        sbcStatus st;
        if ((st = dalloc(&C->d_mrl, (size_t)n)) || (st = dalloc(&C->d_mfr, (size_t)n)) ||
            (st = dalloc(&C->d_mseg, (size_t)n)) || (st = dalloc(&C->d_mpre, (size_t)n + 1)))
            return st;
        C->mroot_cap = n;
    }
    HIPCHK(hipMemcpyAsync(C->d_mrl, mr.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, C->stream));
    HIPCHK(sbc_launch_roots_reset(&P, C->d_mrl, n, C->d_mfr, C->stream));
    std::vector<std::vector<int32_t>> lv((size_t)n, std::vector<int32_t>{0, 1});
    std::vector<int32_t> j((size_t)n, 0), who, ends;
    std::vector<char> act((size_t)n, 1);
    std::vector<SbcRootSeg> sg, sg2;
    std::vector<int64_t> pre, pre2;
    for (;;) {
        sg.clear();
        who.clear();
        ends.clear();
        pre.assign(1, 0);
        for (int i = 0; i < n; i++) {
            if (!act[i]) continue;
            const int32_t a = lv[i][j[i]], b = lv[i][j[i] + 1];
            sg.push_back(SbcRootSeg{mr[i], a, b, j[i]});
            pre.push_back(pre.back() + (b - a));
            who.push_back(i);
            ends.push_back(b);  // the root's list end: where its next frontier starts
        }
        const int m = (int)sg.size();
        if (m == 0) break;
        HIPCHK(hipMemcpyAsync(C->d_mseg, sg.data(), sizeof(SbcRootSeg) * m, hipMemcpyHostToDevice, C->stream));
        HIPCHK(hipMemcpyAsync(C->d_mpre, pre.data(), sizeof(int64_t) * (m + 1), hipMemcpyHostToDevice, C->stream));
        HIPCHK(hipMemcpyAsync(C->d_mfr, ends.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, C->stream));
        HIPCHK(sbc_launch_roots_level(&P, 0, C->d_mseg, C->d_mpre, m, pre.back(), C->d_mfr, C->stream));
        HIPCHK(hipMemcpyAsync(ends.data(), C->d_mfr, sizeof(int32_t) * m, hipMemcpyDeviceToHost, C->stream));
        HIPCHK(hipStreamSynchronize(C->stream));
        sg2.clear();
        pre2.assign(1, 0);
        for (int s2 = 0; s2 < m; s2++) {
            const int i = who[s2];
            const int32_t fe = sg[s2].b, ne = ends[s2];
            if (ne < fe || ne > F->nv) return set_err(SBC_ERR_STATE, "rebuild: list end %d out of range", ne);
            if (ne == fe) {
                act[i] = 0;
                continue;
            }
            sg2.push_back(SbcRootSeg{mr[i], fe, ne, j[i]});
            pre2.push_back(pre2.back() + (ne - fe));
            j[i]++;
            if (j[i] + 1 >= F->lcap) return set_err(SBC_ERR_LEVELS, "BFS depth exceeds the level capacity %d", F->lcap);
            lv[i].push_back(ne);
        }
        const int m2 = (int)sg2.size();
        if (m2 == 0) continue;
        HIPCHK(hipMemcpyAsync(C->d_mseg, sg2.data(), sizeof(SbcRootSeg) * m2, hipMemcpyHostToDevice, C->stream));
        HIPCHK(hipMemcpyAsync(C->d_mpre, pre2.data(), sizeof(int64_t) * (m2 + 1), hipMemcpyHostToDevice, C->stream));
        HIPCHK(sbc_launch_roots_level(&P, 1, C->d_mseg, C->d_mpre, m2, pre2.back(), nullptr, C->stream));
        HIPCHK(hipStreamSynchronize(C->stream));  // sg2/pre2 are refilled next round
    }
    // ascent: root i's levels 0..j[i], deepest first (its deepest has no children)
    int32_t maxd = 0;
    for (int i = 0; i < n; i++) maxd = j[i] > maxd ? j[i] : maxd;
    for (int32_t step = 0; step < maxd; step++) {
        sg.clear();
        pre.assign(1, 0);
        for (int i = 0; i < n; i++) {
            const int32_t L = j[i] - 1 - step;
            if (L < 0) continue;
            sg.push_back(SbcRootSeg{mr[i], lv[i][L], lv[i][L + 1], L});
            pre.push_back(pre.back() + (lv[i][L + 1] - lv[i][L]));
        }
        const int m = (int)sg.size();
        if (m == 0) continue;
        HIPCHK(hipMemcpyAsync(C->d_mseg, sg.data(), sizeof(SbcRootSeg) * m, hipMemcpyHostToDevice, C->stream));
        HIPCHK(hipMemcpyAsync(C->d_mpre, pre.data(), sizeof(int64_t) * (m + 1), hipMemcpyHostToDevice, C->stream));
        HIPCHK(sbc_launch_roots_level(&P, 2, C->d_mseg, C->d_mpre, m, pre.back(), nullptr, C->stream));
        HIPCHK(hipStreamSynchronize(C->stream));
    }
    return SBC_OK;
}

static sbcStatus run_delete(bcForest* forest, stinger* G, extraArraysPerThread* C, int64_t u, int64_t v,
                            int64_t counts[3])
{
    if (!forest || !G || !C) return set_err(SBC_ERR_ARG, "null argument");
    sbcForestDev* F = forest->dev;
    if (!F->initialized) return set_err(SBC_ERR_STATE, "static pass has not run");
    if (F->nv != G->nv) return set_err(SBC_ERR_ARG, "forest NV %ld != graph NV %ld", (long)F->nv, (long)G->nv);
    if (u < 0 || v < 0 || u >= F->nv || v >= F->nv || u == v) return set_err(SBC_ERR_ARG, "bad edge (%ld,%ld)", (long)u, (long)v);
    sbcStatus st = graph_flush(G, C->stream);
    if (st) return st;
    if (C->del_k < F->K) {
        HIPCHK(hipStreamSynchronize(C->stream));
        if (C->d_del) HIPCHK(hipFree(C->d_del));
        C->d_del = nullptr;
        if ((st = dalloc(&C->d_del, (size_t)(8 + 5 * F->K)))) return st;
        C->del_k = F->K;
    }
    if ((st = ensure_rec(C, F->K))) return st;
    if ((st = ensure_bucket(C, G, 4LL * C->num_cu))) return st;
    if ((st = ensure_hubq(C, 4LL * C->num_cu))) return st;
    if ((st = ensure_park(C, F->K))) return st;
    SbcParams P = base_params(F, G, C);
    C->last_parked = 0;
    HIPCHK(sbc_launch_reset(C->d_del, 4, C->stream));
    HIPCHK(sbc_launch_delete_classify(&P, (int32_t)u, (int32_t)v, C->d_del, C->stream));
    int32_t h[4];
    HIPCHK(hipMemcpyAsync(h, C->d_del, sizeof(h), hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    if (h[1] < 0 || h[2] < 0 || h[3] < 0 || (int64_t)h[1] + h[2] + h[3] != F->K)
        return set_err(SBC_ERR_STATE, "deletion: case counts %d/%d/%d do not add up to %ld roots", h[1], h[2], h[3],
                       (long)F->K);
    hipEvent_t e0;
    if (h[2] > 0) {
        // adjacent roots: the single-item update in deletion mode
        SbcParams Q = P;
        Q.u0 = (int32_t)u;
        Q.v0 = (int32_t)v;
        Q.rec = C->d_rec;
        Q.del = 1;
        Q.order = C->d_del + 4;
        Q.qlen = h[2];
        Q.heavy = h[2];
        // items whose descent or ascent grows large go to the wide kernels,
        // as in the windowed form
        Q.park = C->d_park;
        Q.parked = C->d_parked;
        Q.hpark_flag = C->d_pflag;
        Q.hpark_root = C->d_proot;
        Q.park_min = 4096;
        Q.park_scan = kParkScan;
        const char* ps = getenv("SBC_PARK_SCAN");  // test hook: the scan threshold alone
        if (ps) Q.park_scan = atoll(ps);
        int grid = (int)(h[2] < 2LL * C->num_cu ? h[2] : 2LL * C->num_cu);
        {
            // the whole-GPU levels' scratch before the launch; no flag is set now
            SbcWideLists L;
            if ((st = wide_lists(C, Q, &L))) return st;
            memset(C->h_pflag, 0, sizeof(int32_t) * F->K);
        }
        HIPCHK(sbc_launch_reset(C->d_qctr, 4, C->stream));
        if ((st = ev_begin(C, KSTAT_UPDATE, &e0))) return st;
        HIPCHK(sbc_launch_window(&Q, 16, grid, C->stream));
        if ((st = ev_end(C, KSTAT_UPDATE, e0))) return st;
        // items handed over are finished while the launch runs (one item per
        // root: none is resumed)
        int64_t parked = 0;
        if ((st = finish_parked_concurrent(F, C, Q, 16, grid, 1, &parked))) return st;
        C->last_parked = parked;
    }
    // movement roots: updated in place (BC by the deltas that change), one
    // workgroup each; on graphs of at least the several-roots threshold the
    // delta pass -- the deltas of every ancestor of p, hundreds to thousands
    // of vertices with thousands of neighbours each at R-MAT scale 20, 25 ms
    // a deletion on one workgroup -- runs over the whole GPU, one launch per
    // level.  SBC_DELETE_MOVE (test hooks): in_place (one workgroup through),
    // in_place_wide (the delta pass over the GPU) or rebuild (the static pass
    // rebuilds the roots, leaving BC before the rebuild and returning after
    // it).  The update kernel adds the adjacent roots' changes itself (as
    // insertions do).
    int32_t* mlist = C->d_del + 4 + F->K;
    const char* dm = getenv("SBC_DELETE_MOVE");
    const bool rebuild = dm && strcmp(dm, "rebuild") == 0;
    const bool wide_delta = dm ? strcmp(dm, "in_place_wide") == 0 : G->total_cap >= wide_graph_min();
    if (h[3] > 0 && !rebuild) {
        SbcParams Q = P;
        Q.u0 = (int32_t)u;
        Q.v0 = (int32_t)v;
        Q.sroots = mlist;
        Q.scnt = C->d_del + 3;
        if (wide_delta) {
            Q.dm_out = C->d_del + 8 + 2 * F->K;
            // vertices of many neighbours of a level: one workgroup each
            if (!C->d_dmg && (st = dalloc(&C->d_dmg, (size_t)(1 + 2 * kDmGiantCap)))) return st;
            Q.dm_giant = C->d_dmg;
            Q.dm_giant_cap = kDmGiantCap;
        }
        HIPCHK(sbc_launch_reset(C->d_qctr, 4, C->stream));
        if ((st = ev_begin(C, KSTAT_STATIC, &e0))) return st;
        int grid = (int)(h[3] < 2LL * C->num_cu ? h[3] : 2LL * C->num_cu);
        HIPCHK(sbc_launch_delmove(&Q, grid, C->stream));
        if (wide_delta) {
            std::vector<int32_t> out((size_t)2 * h[3]);
            HIPCHK(hipMemcpyAsync(out.data(), Q.dm_out, sizeof(int32_t) * out.size(), hipMemcpyDeviceToHost,
                                  C->stream));
            HIPCHK(hipStreamSynchronize(C->stream));
            int32_t top = 0;
            for (int32_t q = 0; q < h[3]; q++) top = out[2 * q + 1] > top ? out[2 * q + 1] : top;
            if (top >= F->lcap) return set_err(SBC_ERR_STATE, "deletion: delta pass from level %d", top);
            const int lgrid = (int)C->num_cu;
            for (int32_t L = top; L >= 0; L--) HIPCHK(sbc_launch_delmove_level(&Q, L, lgrid, C->stream));
            HIPCHK(sbc_launch_delmove_finish(&Q, 4 * (int)C->num_cu, C->stream));
        }
        if ((st = ev_end(C, KSTAT_STATIC, e0))) return st;
        C->last_rec_n = 0;
        if ((st = check_kernel_err(F, C))) return st;
        counts[0] = h[1];
        counts[1] = h[2];
        counts[2] = h[3];
        return SBC_OK;
    }
    if (h[3] > 0) HIPCHK(sbc_launch_bc_roots(&P, mlist, C->d_del + 3, -1.0, C->stream));
    if (h[3] > 0 && h[3] <= wide_rebuild_max()) {
        // few movement roots: each rebuilt by the whole GPU
        std::vector<int32_t> mr((size_t)h[3]);
        HIPCHK(hipMemcpyAsync(mr.data(), C->d_del + 4 + F->K, sizeof(int32_t) * mr.size(), hipMemcpyDeviceToHost,
                              C->stream));
        HIPCHK(hipStreamSynchronize(C->stream));
        if ((st = ev_begin(C, KSTAT_STATIC, &e0))) return st;
        for (int32_t ri : mr)
            if ((st = rebuild_root_wide(F, G, C, P, ri, C->d_del + 4 + 2 * F->K))) return st;
        if ((st = ev_end(C, KSTAT_STATIC, e0))) return st;
    } else if (h[3] > 0 && G->total_cap >= wide_graph_min()) {
        // many movement roots: rebuilt together by the whole GPU, level by level
        std::vector<int32_t> mr((size_t)h[3]);
        HIPCHK(hipMemcpyAsync(mr.data(), C->d_del + 4 + F->K, sizeof(int32_t) * mr.size(), hipMemcpyDeviceToHost,
                              C->stream));
        HIPCHK(hipStreamSynchronize(C->stream));
        if ((st = ev_begin(C, KSTAT_STATIC, &e0))) return st;
        if ((st = rebuild_roots_wide(F, C, P, mr))) return st;
        if ((st = ev_end(C, KSTAT_STATIC, e0))) return st;
    } else if (h[3] > 0) {
        // movement roots: rebuilt by the static pass, one workgroup each
        SbcParams Q = P;
        Q.sroots = C->d_del + 4 + F->K;
        Q.scnt = C->d_del + 3;
        HIPCHK(sbc_launch_reset(C->d_qctr, 4, C->stream));
        if ((st = ev_begin(C, KSTAT_STATIC, &e0))) return st;
        int grid = (int)(h[3] < 4LL * C->num_cu ? h[3] : 4LL * C->num_cu);
        HIPCHK(launch_static(Q, grid, C->stream));
        if ((st = ev_end(C, KSTAT_STATIC, e0))) return st;
    }
    if (h[3] > 0) HIPCHK(sbc_launch_bc_roots(&P, mlist, C->d_del + 3, 1.0, C->stream));
    C->last_rec_n = 0;
    if ((st = check_kernel_err(F, C))) return st;
    counts[0] = h[1];
    counts[1] = h[2];
    counts[2] = h[3];
    return SBC_OK;
}

extern "C" StreamingExtraInfo deleteEdgeStreamingBC(bcForest* forest, struct stinger* sStinger, int64_t oldU,
                                                    int64_t oldV, int64_t* rootArrayForApproximation, int64_t NK,
                                                    int64_t NV, int64_t NT, extraArraysPerThread** eAPT,
                                                    uint32_t loadBalancing, uint32_t granularity)
{
    (void)rootArrayForApproximation; (void)NT; (void)loadBalancing; (void)granularity;
    StreamingExtraInfo r = {0, 0, 0, 0};
    if (!forest || NK != forest->dev->K || NV != forest->dev->nv) {
        set_err(SBC_ERR_ARG, "NK/NV do not match the forest");
        legacy_abort("deleteEdgeStreamingBC");
    }
    int64_t c[3];
    if (!eAPT || run_delete(forest, sStinger, eAPT[0], oldU, oldV, c) != SBC_OK) legacy_abort("deleteEdgeStreamingBC");
    legacy_refresh_bc(forest, "deleteEdgeStreamingBC");
    r.sameLevel = c[0];
    r.adjacent = c[1];
    r.movement = c[2];
    return r;
}

extern "C" sbcStatus sbcDeleteEdge(bcForest* forest, struct stinger* G, extraArraysPerThread** eAPT, int64_t u,
                                   int64_t v, StreamingExtraInfo* info)
{
    if (!forest || !G || !eAPT || !eAPT[0]) return set_err(SBC_ERR_ARG, "null argument");
    if (!forest->dev->initialized) return set_err(SBC_ERR_STATE, "static pass has not run");
    if (u < 0 || v < 0 || u >= G->nv || v >= G->nv || u == v) return set_err(SBC_ERR_ARG, "bad edge (%ld,%ld)", (long)u, (long)v);
    if (!graph_has_edge(G, (int32_t)u, (int32_t)v)) return set_err(SBC_ERR_ABSENT, "edge (%ld,%ld) absent", (long)u, (long)v);
    stinger_remove_edge(G, 0, u, v);
    stinger_remove_edge(G, 0, v, u);
    int64_t c[3];
    sbcStatus st = run_delete(forest, G, eAPT[0], u, v, c);
    if (st == SBC_OK && info) {
        info->sameLevel = c[0];
        info->adjacent = c[1];
        info->movement = c[2];
        info->connectedComponents = 0;
    }
    return st;
}

extern "C" sbcStatus streamingBCLastRecords(extraArraysPerThread** eAPT, sbcItemRecord* out, int64_t K)
{
    if (!eAPT || !eAPT[0] || !out) return set_err(SBC_ERR_ARG, "null argument");
    extraArraysPerThread* C = eAPT[0];
    if (K != C->last_rec_n || !C->d_rec) return set_err(SBC_ERR_STATE, "no records of that size");
    static_assert(sizeof(sbcItemRecord) == sizeof(SbcItemRec), "record layout");
    HIPCHK(hipMemcpyAsync(out, C->d_rec, sizeof(SbcItemRec) * K, hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    return SBC_OK;
}

// ------------------------------------------------------------ readouts

extern "C" sbcStatus streamingBCGetBC(bcForest* forest, double* out)
{
    if (!forest) return set_err(SBC_ERR_ARG, "null forest");
    sbcForestDev* F = forest->dev;
    sbcStatus st = forest_read(F, forest->totalBC, F->d_bc, sizeof(double) * F->nv);
    if (st) return st;
    if (out) memcpy(out, forest->totalBC, sizeof(double) * F->nv);
    return SBC_OK;
}

// The reference's update calls leave BC in forest->totalBC, where its callers
// read it (src/streamingbc.c:211-218; src/sbcMain_utils.c:488-548): the
// legacy entry points refresh the host mirror before they return.
static void legacy_refresh_bc(bcForest* forest, const char* where)
{
    if (streamingBCGetBC(forest, nullptr) != SBC_OK) legacy_abort(where);
}

extern "C" sbcStatus streamingBCGetTree(bcForest* forest, int64_t ri, int64_t* level, int64_t* sigma, double* delta)
{
    if (!forest) return set_err(SBC_ERR_ARG, "null forest");
    sbcForestDev* F = forest->dev;
    if (ri < 0 || ri >= F->K) return set_err(SBC_ERR_ARG, "root index out of range");
    size_t off = (size_t)ri * (size_t)F->nv;
    sbcStatus st;
    if (level) {
        std::vector<int32_t> d(F->nv);
        if ((st = forest_read(F, d.data(), F->d_dist + off, sizeof(int32_t) * F->nv))) return st;
        for (int64_t x = 0; x < F->nv; x++) level[x] = d[x] == INT32_MAX ? LONG_MAX : d[x];
    }
    if (sigma || delta) {
        std::vector<SbcVRec> rs((size_t)F->nv);
        if ((st = forest_read(F, rs.data(), F->d_vr + off, sizeof(SbcVRec) * F->nv))) return st;
        for (int64_t x = 0; x < F->nv; x++) {
            if (sigma) sigma[x] = rs[x].sigma;
            if (delta) delta[x] = rs[x].delta;
        }
    }
    return SBC_OK;
}

extern "C" sbcStatus streamingBCRecomputeBC(bcForest* forest, extraArraysPerThread** eAPT)
{
    if (!forest || !eAPT || !eAPT[0]) return set_err(SBC_ERR_ARG, "null argument");
    extraArraysPerThread* C = eAPT[0];
    sbcForestDev* F = forest->dev;
    SbcParams P;
    memset(&P, 0, sizeof(P));
    P.nv = (int32_t)F->nv;
    P.K = (int32_t)F->K;
    P.vr = F->d_vr;
    P.bc = F->d_bc;
    P.roots = F->d_roots;
    hipEvent_t e0;
    sbcStatus st;
    if ((st = ev_begin(C, KSTAT_BC, &e0))) return st;
    HIPCHK(sbc_launch_bc_reduce(&P, C->stream));
    if ((st = ev_end(C, KSTAT_BC, e0))) return st;
    if ((st = forest_mark(F, C->stream))) return st;
    HIPCHK(hipStreamSynchronize(C->stream));
    return SBC_OK;
}

// ------------------------------------------------------------ windowed form

static size_t stage_bytes(int cap)
{
    // desc + 6 int arrays of 2*cap + vinfo array of 4*cap (touched this and last window)
    return sizeof(SbcWinDesc) + sizeof(int32_t) * 6 * 4 * (size_t)cap + sizeof(SbcVInfo) * 4 * (size_t)cap + 64;
}

static sbcStatus ensure_stage(extraArraysPerThread* C, int W)
{
    WinStage& S = C->win;
    if (S.cap >= W) return SBC_OK;
    HIPCHK(hipStreamSynchronize(C->stream));
    if (S.d_base) HIPCHK(hipFree(S.d_base));
    S.d_base = nullptr;
    S.d_desc = nullptr;
    S.d_ints = nullptr;
    S.d_vi = nullptr;
    for (int i = 0; i < 2; i++) {
        if (S.h_buf[i]) HIPCHK(hipHostFree(S.h_buf[i]));
        S.h_buf[i] = nullptr;
        if (!S.h_done[i]) HIPCHK(hipEventCreateWithFlags(&S.h_done[i], hipEventDisableTiming));
    }
    S.cap = W;
    HIPCHK(hipMalloc((void**)&S.d_base, stage_bytes(W)));
    for (int i = 0; i < 2; i++) {
        HIPCHK(hipHostMalloc(&S.h_buf[i], stage_bytes(W)));
        HIPCHK(hipEventRecord(S.h_done[i], C->stream));
    }
    C->gc.W = -1; // invalidate cached graph
    return SBC_OK;
}

static_assert(sizeof(SbcWinDesc) <= 64, "window descriptor exceeds its staging slot");

// Device views of the staging block for windows of W insertions (W <= cap).
static void stage_views(WinStage& S, int W)
{
    S.d_desc = (SbcWinDesc*)S.d_base;
    S.d_ints = (int32_t*)(S.d_base + 64);
    S.d_vi = (SbcVInfo*)(S.d_base + 64 + sizeof(int32_t) * 6 * 4 * (size_t)W);
}

// Bytes from the start of the staging block through nvi vertex records.
static size_t stage_span(int W, size_t nvi)
{
    return 64 + sizeof(int32_t) * 6 * 4 * (size_t)W + sizeof(SbcVInfo) * nvi;
}

static sbcStatus ensure_order(extraArraysPerThread* C, int64_t K)
{
    if (C->order_cap >= K) return SBC_OK;
    HIPCHK(hipStreamSynchronize(C->stream));
    if (C->d_order) HIPCHK(hipFree(C->d_order));
    C->d_order = nullptr;
    sbcStatus st = dalloc(&C->d_order, (size_t)K);
    if (st) return st;
    C->order_cap = K;
    C->gc.has_order = -1;
    C->order_ready = false;
    return SBC_OK;
}

// Enqueues the launches of one window (append, queue reset, window kernel,
// optional BC reduction) on the context stream.
static sbcStatus ensure_park(extraArraysPerThread* C, int64_t K)
{
    if (C->park_cap >= K) return SBC_OK;
    HIPCHK(hipStreamSynchronize(C->stream));
    for (void* q : {(void*)C->d_park, (void*)C->d_parked, (void*)C->d_order2, (void*)C->d_start})
        if (q) HIPCHK(hipFree(q));
    C->d_park = nullptr;
    C->d_parked = C->d_order2 = C->d_start = nullptr;
    sbcStatus st;
    if ((st = dalloc(&C->d_park, (size_t)K)) || (st = dalloc(&C->d_parked, (size_t)K)) ||
        (st = dalloc(&C->d_order2, (size_t)K)) || (st = dalloc(&C->d_start, (size_t)K)))
        return st;
    if (!C->stream2) HIPCHK(hipStreamCreateWithFlags(&C->stream2, hipStreamNonBlocking));
    HIPCHK(hipStreamSynchronize(C->stream2));
    for (void* q : {(void*)C->d_qctr2, (void*)C->d_parked2})
        if (q) HIPCHK(hipFree(q));
    for (void* q : {(void*)C->h_pflag, (void*)C->h_proot})
        if (q) HIPCHK(hipHostFree(q));
    C->d_qctr2 = C->d_parked2 = C->h_pflag = C->h_proot = C->d_pflag = C->d_proot = nullptr;
    if ((st = dalloc(&C->d_qctr2, 4)) || (st = dalloc(&C->d_parked2, (size_t)K))) return st;
    HIPCHK(hipHostMalloc((void**)&C->h_pflag, sizeof(int32_t) * K, hipHostMallocMapped | hipHostMallocCoherent));
    HIPCHK(hipHostMalloc((void**)&C->h_proot, sizeof(int32_t) * K, hipHostMallocMapped | hipHostMallocCoherent));
    HIPCHK(hipHostGetDevicePointer((void**)&C->d_pflag, C->h_pflag, 0));
    HIPCHK(hipHostGetDevicePointer((void**)&C->d_proot, C->h_proot, 0));
    memset(C->h_pflag, 0, sizeof(int32_t) * K);
    C->h_start.assign((size_t)K, 0);
    C->park_cap = K;
    C->gc.with_rec = -1; // cached graph holds the old pointers
    return SBC_OK;
}

// Scratch of the whole-GPU levels: 16 counters, a heavy list of 2 * nv and
// an aux list of nv + slots / 512 + 1 (every work group of a frontier: each
// frontier vertex once, a heavy one with ceil(len / 512) groups).
static sbcStatus wide_lists(extraArraysPerThread* C, const SbcParams& P, SbcWideLists* L)
{
    const int64_t nv = P.nv;
    const int64_t aux = nv + P.col_cap / 512 + 1;
    if (aux > INT32_MAX - 16 - 2 * nv) return set_err(SBC_ERR_ARG, "graph too large for the wide-level scratch");
    if (C->wl_nv < nv || C->wl_aux < aux) {
        HIPCHK(hipStreamSynchronize(C->stream));
        if (C->d_wl) HIPCHK(hipFree(C->d_wl));
        C->d_wl = nullptr;
        C->wl_nv = C->wl_aux = 0;
        sbcStatus st = dalloc(&C->d_wl, (size_t)(16 + 2 * nv + aux));
        if (st) return st;
        C->wl_nv = nv;
        C->wl_aux = aux;
    }
    L->ctr = C->d_wl;
    L->heavy = C->d_wl + 16;
    L->aux = C->d_wl + 16 + 2 * C->wl_nv;
    L->aux_cap = (int32_t)C->wl_aux;
    return SBC_OK;
}

static int wide_grid(const extraArraysPerThread* C, int64_t n)
{
    // one wavefront per vertex, 4 per workgroup, up to 8 workgroups per CU
    int64_t g = (n + 3) / 4;
    int64_t cap = 8LL * (C->num_cu > 0 ? C->num_cu : 256);
    return (int)(g < 1 ? 1 : (g < cap ? g : cap));
}

// Finishes the parked item of root ri: the remaining descent frontiers, the
// ascent levels and the commit, with whole-GPU launches per level (one for
// the light vertices, one lane each, then one for the heavy and giant
// vertices); the level offsets and list counters are read back between
// levels to drive the same loop as a team's.
static sbcStatus wide_ascent_commit(sbcForestDev* F, extraArraysPerThread* C, const SbcParams& P, int32_t ri,
                                    const SbcWideItem& w, const std::vector<int32_t>& lv, int nf, int nD, int nS,
                                    int lev0, int uA, int uB, bool light_counted);

static sbcStatus finish_parked_item(sbcForestDev* F, extraArraysPerThread* C, const SbcParams& P, int32_t ri)
{
    SbcWideItem w;
    SbcWideItem* dw = C->d_park + ri;
    const int lcap = F->lcap;
    std::vector<int32_t> lv(2 * (size_t)lcap);
    int32_t* lvlD = lv.data();
    int32_t* lvlS = lvlD + lcap;
    int32_t* dlvl = F->d_lvl + (size_t)ri * 2 * lcap;
    HIPCHK(hipMemcpyAsync(&w, dw, sizeof(w), hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipMemcpyAsync(lv.data(), dlvl, sizeof(int32_t) * lv.size(), hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    if (w.j < 1 || w.j >= lcap || w.fs < 0 || w.fs > w.fe || w.fe > F->nv || w.item < 0 || w.phase < 0 ||
        w.phase > 1 || (w.phase == 1 && (w.uA < 0 || w.uA > w.uB || w.uB > F->nv || w.lev < 0 || w.lev >= lcap)))
        return set_err(SBC_ERR_STATE, "parked item of root %d is inconsistent", (int)ri);
    // parked in an ascent level: the descent is complete, the team counted
    // the level's light vertices
    if (w.phase == 1) return wide_ascent_commit(F, C, P, ri, w, lv, w.j, w.cnt[0], w.cnt[1], w.lev, w.uA, w.uB, true);
    int j = w.j, fs = w.fs, fe = w.fe;
    int32_t c01[2] = {w.cnt[0], w.cnt[1]};
    SbcWideLists L;
    sbcStatus st = wide_lists(C, P, &L);
    if (st) return st;
    while (fs < fe) {
        HIPCHK(sbc_launch_wide_descent(&P, dw, ri, fs, fe, w.base + j, &L, C->num_cu, C->stream));
        HIPCHK(hipMemcpyAsync(C->h_wend, &dw->cnt[0], sizeof(c01), hipMemcpyDeviceToHost, C->stream));
        HIPCHK(hipStreamSynchronize(C->stream));
        c01[0] = C->h_wend[0];
        c01[1] = C->h_wend[1];
        j++;
        if (j + 1 >= lcap) return set_err(SBC_ERR_LEVELS, "BFS depth exceeds the level capacity %d", lcap);
        if (c01[0] < fe || c01[0] > F->nv || c01[1] < 0 || c01[1] > F->nv)
            return set_err(SBC_ERR_STATE, "parked item of root %d: list counters out of range", (int)ri);
        lvlD[j] = fe;
        lvlS[j] = c01[1];
        fs = fe;
        fe = c01[0];
    }
    const int nf = j, nD = c01[0], nS = c01[1];
    int32_t u = nD;
    HIPCHK(hipMemcpyAsync(dlvl, lv.data(), sizeof(int32_t) * lv.size(), hipMemcpyHostToDevice, C->stream));
    HIPCHK(hipMemcpyAsync(&dw->cnt[2], &u, sizeof(int32_t), hipMemcpyHostToDevice, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));  // lv and u are host stack data
    return wide_ascent_commit(F, C, P, ri, w, lv, nf, nD, nS, w.base + nf - 1, nD, nD, false);
}

// Ascent levels lev0 .. 0 of a parked item (level lev0's ascent set is
// list[uA, uB); the item's counter cnt[2] holds the end of the ascent list),
// then its commit, with the wide kernels.  light_counted: the team counted
// level lev0's light vertices.
static sbcStatus wide_ascent_commit(sbcForestDev* F, extraArraysPerThread* C, const SbcParams& P, int32_t ri,
                                    const SbcWideItem& w, const std::vector<int32_t>& lv, int nf, int nD, int nS,
                                    int lev0, int uA, int uB, bool light_counted)
{
    SbcWideItem* dw = C->d_park + ri;
    const int lcap = F->lcap;
    const int32_t* lvlD = lv.data();
    const int32_t* lvlS = lvlD + lcap;
    int32_t u = 0;
    SbcWideLists L;
    sbcStatus st = wide_lists(C, P, &L);
    if (st) return st;
    for (int lev = lev0; lev >= 0; --lev) {
        const int jj = lev - w.base;
        int32_t h[6] = {0, 0, 0, 0, uA, uB};
        if (jj >= 0) { h[0] = lvlD[jj]; h[1] = lvlD[jj + 1]; h[2] = lvlS[jj]; h[3] = lvlS[jj + 1]; }
        const int64_t tot = (int64_t)(h[1] - h[0]) + (h[3] - h[2]) + (h[5] - h[4]);
        if (tot == 0) break;
        HIPCHK(sbc_launch_wide_ascent(&P, dw, ri, h, lev, light_counted && lev == lev0 ? 1 : 0, &L, C->num_cu,
                                      C->stream));
        // a deletion's p lost its child s: a parent to update once s's level is done
        if (P.del && lev == w.base) HIPCHK(sbc_launch_wide_mark(&P, dw, ri, w.p, lev - 1, C->stream));
        HIPCHK(hipMemcpyAsync(C->h_wend, &dw->cnt[2], sizeof(int32_t), hipMemcpyDeviceToHost, C->stream));
        HIPCHK(hipStreamSynchronize(C->stream));
        u = *C->h_wend;
        if (u < uB || (int64_t)u + nS > F->nv)
            return set_err(SBC_ERR_STATE, "parked item of root %d: ascent list out of range", (int)ri);
        uA = uB;
        uB = u;
    }
    const int64_t ncommit = nD > uB + nS - nD ? nD : uB + nS - nD;
    HIPCHK(sbc_launch_wide_commit(&P, dw, ri, nf, nD, uB, nS, wide_grid(C, ncommit / 64 + 1), C->stream));
    C->h_start[ri] = w.item + 1;
    return SBC_OK;
}

// Points the context's helpers (copies, waits, timing events) at another
// stream while the guard lives.
struct StreamScope {
    extraArraysPerThread* C;
    hipStream_t saved;
    StreamScope(extraArraysPerThread* c, hipStream_t st) : C(c), saved(c->stream) { c->stream = st; }
    ~StreamScope() { C->stream = saved; }
};

// Finishes root ri's parked item, then resumes the root alone at its next
// item, again after every item it parks (on the context's current stream).
// The resumed launch has its own queue counters and parked list and the
// per-workgroup scratch slice past the window launch's grid, so it can run
// beside that launch.
static sbcStatus finish_root_alone(sbcForestDev* F, extraArraysPerThread* C, const SbcParams& P, int nwb, int grid,
                                   int nitems, int32_t ri, int64_t* total)
{
    for (;;) {
        *total += 1;
        sbcStatus st = finish_parked_item(F, C, P, ri);
        if (st) return st;
        if (C->h_start[ri] >= nitems) return SBC_OK;  // the root has no item left
        SbcParams Q = P;
        Q.order = C->d_order2;
        Q.heavy = 1;
        Q.qlen = 1;
        Q.start_item = C->d_start;
        Q.qctr = C->d_qctr2;
        Q.parked = C->d_parked2;
        Q.hpark_flag = nullptr;
        Q.hpark_root = nullptr;
        if (Q.bucket) Q.bucket += (size_t)grid * SBC_BUCKET_SLOTS;
        if (Q.hubq) Q.hubq += (size_t)grid * SBC_HUBQ_CAP;
        HIPCHK(hipMemcpyAsync(C->d_order2, &ri, sizeof(int32_t), hipMemcpyHostToDevice, C->stream));
        HIPCHK(hipMemcpyAsync(C->d_start + ri, &C->h_start[ri], sizeof(int32_t), hipMemcpyHostToDevice, C->stream));
        HIPCHK(sbc_launch_reset(C->d_qctr2, 4, C->stream));
        HIPCHK(sbc_launch_window(&Q, nwb, 1, C->stream));
        HIPCHK(hipMemcpyAsync(C->h_np, C->d_qctr2 + 2, sizeof(int32_t), hipMemcpyDeviceToHost, C->stream));
        // also keeps ri and h_start[ri] unchanged until their uploads ran
        HIPCHK(hipStreamSynchronize(C->stream));
        const int32_t np = *C->h_np;
        if (np == 0) return SBC_OK;
        if (np != 1) return set_err(SBC_ERR_STATE, "resumed root %d parked %d items", (int)ri, (int)np);
    }
}

// While the window launch (nitems items) runs on the context stream, every
// root it parks (published in host memory: C->h_proot[i], then
// C->h_pflag[i]) is finished on the second stream, so the hand-over and the
// root's remaining items use the CUs the launch's other roots have left.
// Returns once the launch has ended and every parked root is done.
static sbcStatus finish_parked_concurrent(sbcForestDev* F, extraArraysPerThread* C, const SbcParams& P, int nwb,
                                          int grid, int nitems, int64_t* total)
{
    sbcStatus err = SBC_OK;
    int handled = 0;
    hipEvent_t e1 = nullptr;
    bool timed = false;
    for (;;) {
        // the launch's state is read first: once it has ended, every flag it
        // set is visible
        const hipError_t q = hipStreamQuery(C->stream);
        const bool ended = q != hipErrorNotReady;
        while (handled < F->K && __atomic_load_n(&C->h_pflag[handled], __ATOMIC_ACQUIRE) != 0) {
            const int32_t ri = __atomic_load_n(&C->h_proot[handled], __ATOMIC_RELAXED);
            handled++;
            if (err) continue;
            if (ri < 0 || ri >= F->K) {
                err = set_err(SBC_ERR_STATE, "parked root %d out of range", (int)ri);
                continue;
            }
            StreamScope on2(C, C->stream2);
            if (!timed) {
                timed = true;
                if ((err = ev_begin(C, KSTAT_WIDE, &e1))) continue;
            }
            err = finish_root_alone(F, C, P, nwb, grid, nitems, ri, total);
        }
        if (ended) {
            if (q != hipSuccess && !err) err = set_err(SBC_ERR_HIP, "window launch: %s", hipGetErrorString(q));
            break;
        }
        // several serving workers may poll at once
        std::this_thread::yield();
    }
    if (timed) {
        StreamScope on2(C, C->stream2);
        const sbcStatus e2 = ev_end(C, KSTAT_WIDE, e1);
        if (!err) err = e2;
    }
    const hipError_t e3 = hipStreamSynchronize(C->stream2);
    if (!err && e3 != hipSuccess) err = set_err(SBC_ERR_HIP, "hand-over stream: %s", hipGetErrorString(e3));
    return err;
}

static sbcStatus enqueue_window(const SbcParams& P, const SbcAppend& A, int nwb, int grid, int W,
                                int recompute, extraArraysPerThread* C, bool timed)
{
    sbcStatus st;
    hipEvent_t e0 = nullptr;
    HIPCHK(sbc_launch_append(&A, 4 * W, C->stream));
    HIPCHK(sbc_launch_reset(C->d_qctr, 4, C->stream));
    if (timed && (st = ev_begin(C, KSTAT_UPDATE, &e0))) return st;
    HIPCHK(sbc_launch_window(&P, nwb, grid, C->stream));
    if (timed && (st = ev_end(C, KSTAT_UPDATE, e0))) return st;
    if (recompute) {
        if (timed && (st = ev_begin(C, KSTAT_BC, &e0))) return st;
        HIPCHK(sbc_launch_bc_reduce(&P, C->stream));
        if (timed && (st = ev_end(C, KSTAT_BC, e0))) return st;
    }
    return SBC_OK;
}

extern "C" sbcStatus streamingBCInsertEdgesBatched(bcForest* forest, struct stinger* G, extraArraysPerThread** eAPT,
                                                   const int64_t* u, const int64_t* v, int64_t count,
                                                   const sbcBatchConfig* cfg, sbcItemRecord* records)
{
    if (!forest || !G || !eAPT || !eAPT[0] || (count > 0 && (!u || !v)) || !cfg)
        return set_err(SBC_ERR_ARG, "null argument");
    extraArraysPerThread* C = eAPT[0];
    sbcForestDev* F = forest->dev;
    if (!F->initialized) return set_err(SBC_ERR_STATE, "static pass has not run");
    if (F->nv != G->nv) return set_err(SBC_ERR_ARG, "forest/graph NV mismatch");
    const int W = cfg->window;
    if (W < 1 || W > (1 << 16)) return set_err(SBC_ERR_ARG, "window %d out of range", W);
    int bt = cfg->block_threads ? cfg->block_threads : 256;
    if (bt != 64 && bt != 128 && bt != 256 && bt != 512 && bt != 1024)
        return set_err(SBC_ERR_ARG, "block_threads must be 64, 128, 256, 512 or 1024");
    const int nwb = bt / 64;
    // validate the whole batch before touching anything
    sbcStatus st;
    if ((st = validate_batch(G, u, v, count))) return st;
    if ((st = graph_flush(G, C->stream))) return st;
    if ((st = ensure_stage(C, W))) return st;
    stage_views(C->win, W);
    if ((st = ensure_rec(C, (int64_t)W * F->K))) return st;
    if ((st = ensure_order(C, F->K))) return st;
    const int park_min = cfg->park_min == 0 ? 4096 : cfg->park_min;
    if (park_min > 0 && nwb > 1 && (st = ensure_park(C, F->K))) return st;
    int heavy = cfg->heavy_roots < 0 || cfg->heavy_roots > F->K ? (int)F->K : cfg->heavy_roots;
    if (cfg->use_context_order && (cfg->root_order || !C->order_ready || C->order_forest != F))
        return set_err(SBC_ERR_ARG, "use_context_order needs a preceding streamingBCScheduleWindow and no root_order");
    const int has_order = cfg->root_order != nullptr || cfg->use_context_order;
    C->order_ready = false;  // a scheduled order serves one batched call
    if (cfg->root_order) {
        std::vector<char> chk(F->K, 0);
        for (int64_t i = 0; i < F->K; i++) {
            int32_t x = cfg->root_order[i];
            if (x < 0 || x >= F->K || chk[x]) return set_err(SBC_ERR_ARG, "root_order is not a permutation");
            chk[x] = 1;
        }
        HIPCHK(hipMemcpyAsync(C->d_order, cfg->root_order, sizeof(int32_t) * F->K, hipMemcpyHostToDevice, C->stream));
        C->order_ready = false;
        HIPCHK(hipStreamSynchronize(C->stream));
    }
    int grid = cfg->grid_blocks;
    if (grid <= 0) {
        int64_t gh = heavy;
        int64_t gl = (F->K - heavy + nwb - 1) / nwb;
        int64_t g = gh > gl ? gh : gl;
        int64_t cap = 2LL * C->num_cu;
        grid = (int)(g < cap ? (g > 0 ? g : 1) : cap);
    }
    // one scratch slice more: a root resumed beside the window launch
    if ((st = ensure_bucket(C, G, grid + 1))) return st;
    if ((st = ensure_hubq(C, grid + 1))) return st;
    // capacity: relayout once up front if any endpoint lacks room for the whole batch
    if ((int64_t)G->h_cnt.size() != G->nv) G->h_cnt.assign((size_t)G->nv, 0);
    if ((int64_t)G->h_mark.size() != G->nv) G->h_mark.assign((size_t)G->nv, 0);
    {
        std::vector<int32_t>& extra = G->h_cnt;
        std::vector<int32_t> ends;
        ends.reserve((size_t)count * 2);
        bool short_room = false;
        for (int64_t i = 0; i < count; i++)
            for (int64_t x : {u[i], v[i]})
                if (extra[x]++ == 0) ends.push_back((int32_t)x);
        for (int32_t x : ends) {
            if (G->deg[x] + extra[x] > G->cap[x]) {
                G->cap[x] = G->deg[x] + extra[x] + slack_for(G->deg[x]);
                short_room = true;
            }
            extra[x] = 0;
        }
        if (short_room) graph_relayout(G);
        if ((st = graph_flush(G, C->stream))) return st;
    }
    const char* trace_path = getenv("SBC_ITEM_TRACE");
    if (trace_path && C->tprof_cap < (int64_t)W * F->K * SBC_TRACE_SLOTS) {
        HIPCHK(hipStreamSynchronize(C->stream));
        if (C->d_tprof) HIPCHK(hipFree(C->d_tprof));
        C->d_tprof = nullptr;
        if ((st = dalloc(&C->d_tprof, (size_t)W * F->K * SBC_TRACE_SLOTS))) return st;
        C->tprof_cap = (int64_t)W * F->K * SBC_TRACE_SLOTS;
        C->gc.with_rec = -1;
    }
    if (C->wtot_k < F->K) {
        HIPCHK(hipStreamSynchronize(C->stream));
        if (C->d_wtot) HIPCHK(hipFree(C->d_wtot));
        C->d_wtot = nullptr;
        if ((st = dalloc(&C->d_wtot, (size_t)8 * F->K))) return st;
        HIPCHK(hipMemsetAsync(C->d_wtot, 0, sizeof(int64_t) * 8 * (size_t)F->K, C->stream));
        C->wtot_k = F->K;
        C->gc.with_rec = -1;  // the captured launch holds the old pointer
    }
    SbcParams P = base_params(F, G, C);
    P.tprof = trace_path ? C->d_tprof : nullptr;
    P.wtot = C->d_wtot;
    P.heavy = heavy;
    P.order = has_order ? C->d_order : nullptr;
    P.win = C->win.d_desc;
    int32_t* di = C->win.d_ints;
    P.wu = di;
    P.wv = di + 4 * W;
    P.rec = C->d_rec;
    P.bc_atomic = cfg->recompute_bc ? 0 : 1;
    if (park_min > 0 && nwb > 1) {
        P.park = C->d_park;
        P.parked = C->d_parked;
        P.hpark_flag = C->d_pflag;
        P.hpark_root = C->d_proot;
        P.park_min = park_min;
        // an ascent level whose hubs leave this many adjacency entries to
        // scan is handed over too: one workgroup scans 50-150 million entries
        // a second, the wide kernels put every hub on a wavefront of its own
        // across the GPU; the rest of the root's window then runs in a launch
        // of its own (headline workload: 2M entries 10.65K insertions/s, 1M
        // 10.89K, 512K 11.49K, 256K 11.26K, 128K 10.59K)
        P.park_scan = cfg->park_scan > 0 ? cfg->park_scan : kParkScan;
        const char* ps = getenv("SBC_PARK_SCAN");  // test hook: the scan threshold alone
        if (ps) P.park_scan = atoll(ps);
    }
    SbcAppend A;
    memset(&A, 0, sizeof(A));
    A.vinfo = G->d_vinfo;
    A.col = G->d_col;
    A.seq = G->d_seq;
    A.desc = C->win.d_desc;
    A.d_slot = di + 8 * W;
    A.d_scol = di + 12 * W;
    A.d_sseq = di + 16 * W;
    A.d_vv = di + 20 * W;
    A.d_vi = C->win.d_vi;

    bool use_graph = cfg->use_graph != 0;
    GraphCache& gc = C->gc;
    bool need_capture = use_graph &&
        (!gc.exec || gc.forest != F || gc.g != G || gc.generation != G->generation || gc.nwb != nwb ||
         gc.grid != grid || gc.heavy != heavy || gc.recompute != cfg->recompute_bc || gc.W != W ||
         gc.has_order != has_order || gc.with_rec != 1 || gc.park_min != P.park_min ||
         gc.run_cap != P.run_cap || gc.hub_batch != P.hub_batch || gc.park_scan != P.park_scan);
    if (need_capture) {
        // the window's launch sequence is built node by node (no stream
        // capture), so work that other host threads enqueue, allocate or
        // synchronise meanwhile cannot invalidate it; a failure is an error
        if (gc.exec) { (void)hipGraphExecDestroy(gc.exec); gc.exec = nullptr; }
        if (gc.graph) { (void)hipGraphDestroy(gc.graph); gc.graph = nullptr; }
        HIPCHK(hipStreamSynchronize(C->stream));
        hipGraph_t g = nullptr;
        HIPCHK(hipGraphCreate(&g, 0));
        hipError_t ee = sbc_graph_window(g, &A, 4 * W, C->d_qctr, &P, nwb, grid, cfg->recompute_bc);
        if (ee == hipSuccess) ee = hipGraphInstantiate(&gc.exec, g, nullptr, nullptr, 0);
        if (ee != hipSuccess) {
            (void)hipGraphDestroy(g);
            gc.exec = nullptr;
            return set_err(SBC_ERR_HIP, "building the window graph failed: %s", hipGetErrorString(ee));
        }
        {
            gc.graph = g;
            C->graphs_built++;
            gc.forest = F; gc.g = G; gc.generation = G->generation; gc.nwb = nwb; gc.grid = grid;
            gc.heavy = heavy; gc.recompute = cfg->recompute_bc; gc.W = W; gc.has_order = has_order;
            gc.with_rec = 1;
            gc.park_min = P.park_min;
            gc.run_cap = P.run_cap;
            gc.hub_batch = P.hub_batch;
            gc.park_scan = P.park_scan;
        }
    }

    if (P.park) {
        // the whole-GPU levels' scratch, allocated now: nothing is allocated
        // or freed while a window launch runs beside the hand-over stream
        SbcWideLists L;
        if ((st = wide_lists(C, P, &L))) return st;
    }
    std::vector<int32_t> prev_touched;
    int64_t nparked = 0, nparked_before = 0;
    std::vector<int32_t> wu(W), wv(W), slot(2 * W), scol(2 * W), sseq(2 * W), vv, touched;
    std::vector<SbcVInfo> vi;
    std::vector<char>& mark = G->h_mark;  // cleared again at the end of each window's loop body
    for (int64_t w0 = 0; w0 < count; w0 += W) {
        const int n = (int)((count - w0) < W ? (count - w0) : W);
        C->last_batch_items = n;  // the device records hold this window's items
        const int32_t seq0 = G->seq_next;
        // deg before this window = always-visible prefix for this window
        touched.clear();
        for (int i = 0; i < n; i++) {
            int32_t a = (int32_t)u[w0 + i], b = (int32_t)v[w0 + i];
            wu[i] = a;
            wv[i] = b;
            for (int32_t x : {a, b})
                if (!mark[x]) { mark[x] = 1; touched.push_back(x); G->deg0[x] = G->deg[x]; }
        }
        for (int32_t x : prev_touched)
            if (!mark[x]) G->deg0[x] = G->deg[x];
        int ns = 0;
        for (int i = 0; i < n; i++) {
            int32_t q = seq0 + i;
            int32_t sa = graph_push(G, wu[i], wv[i], q);
            int32_t sb = graph_push(G, wv[i], wu[i], q);
            slot[ns] = sa; scol[ns] = wv[i]; sseq[ns] = q; ns++;
            slot[ns] = sb; scol[ns] = wu[i]; sseq[ns] = q; ns++;
        }
        G->seq_next += n;
        vv.clear();
        vi.clear();
        for (int32_t x : touched) { vv.push_back(x); vi.push_back(vinfo_of(G, x)); }
        for (int32_t x : prev_touched)
            if (!mark[x]) { vv.push_back(x); vi.push_back(vinfo_of(G, x)); }
        for (int32_t x : touched) mark[x] = 0;
        prev_touched = touched;
        // stage: desc | ints (wu, wv, slot, scol, sseq, vv at stride 4W) | vinfo
        WinStage& S = C->win;
        int b = S.h_cur;
        S.h_cur ^= 1;
        HIPCHK(hipEventSynchronize(S.h_done[b]));
        char* hb = (char*)S.h_buf[b];
        SbcWinDesc* hd = (SbcWinDesc*)hb;
        hd->W = n;
        hd->seq0 = seq0;
        hd->nslots = ns;
        hd->nvupd = (int32_t)vv.size();
        int32_t* hi = (int32_t*)(hb + 64);
        memcpy(hi + 0 * 4 * W, wu.data(), sizeof(int32_t) * n);
        memcpy(hi + 1 * 4 * W, wv.data(), sizeof(int32_t) * n);
        memcpy(hi + 2 * 4 * W, slot.data(), sizeof(int32_t) * ns);
        memcpy(hi + 3 * 4 * W, scol.data(), sizeof(int32_t) * ns);
        memcpy(hi + 4 * 4 * W, sseq.data(), sizeof(int32_t) * ns);
        memcpy(hi + 5 * 4 * W, vv.data(), sizeof(int32_t) * vv.size());
        SbcVInfo* hv = (SbcVInfo*)(hb + 64 + sizeof(int32_t) * 6 * 4 * (size_t)W);
        memcpy(hv, vi.data(), sizeof(SbcVInfo) * vi.size());
        // one copy: the host block has the device block's layout
        HIPCHK(hipMemcpyAsync(S.d_base, hb, stage_span(W, vi.size()), hipMemcpyHostToDevice, C->stream));
        HIPCHK(hipEventRecord(S.h_done[b], C->stream));
        // the last window's launch has ended: no flag is being set
        if (P.park) memset(C->h_pflag, 0, sizeof(int32_t) * F->K);
        if (use_graph) {
            hipEvent_t e0;
            if ((st = ev_begin(C, KSTAT_UPDATE, &e0))) return st;
            HIPCHK(hipGraphLaunch(gc.exec, C->stream));
            C->graph_launches++;
            if ((st = ev_end(C, KSTAT_UPDATE, e0))) return st;
        } else {
            if ((st = enqueue_window(P, A, nwb, grid, W, cfg->recompute_bc, C, true))) return st;
        }
        // items this window hands over are finished while it runs, and
        // before the next window
        if (P.park && (st = finish_parked_concurrent(F, C, P, nwb, grid, n, &nparked))) return st;
        // the captured BC reduction ran before the parked items were finished
        if (nparked > nparked_before && cfg->recompute_bc) HIPCHK(sbc_launch_bc_reduce(&P, C->stream));
        nparked_before = nparked;
        if (trace_path) {
            std::vector<int64_t> tb((size_t)n * F->K * SBC_TRACE_SLOTS);
            HIPCHK(hipMemcpyAsync(tb.data(), C->d_tprof, sizeof(int64_t) * tb.size(), hipMemcpyDeviceToHost, C->stream));
            HIPCHK(hipStreamSynchronize(C->stream));
            FILE* fp = fopen(trace_path, "ab");
            if (fp) {
                fwrite(tb.data(), sizeof(int64_t), tb.size(), fp);
                fclose(fp);
            }
        }
        if (records) {
            // [insertion][root] for this window
            HIPCHK(hipMemcpyAsync(records + (size_t)w0 * F->K, C->d_rec, sizeof(SbcItemRec) * (size_t)n * F->K,
                                  hipMemcpyDeviceToHost, C->stream));
        }
    }
    C->last_parked = nparked;
    // make the last window's appends visible to everyone: one launch with the
    // vertex descriptors staged like a window's (no slots)
    for (int32_t x : prev_touched) G->deg0[x] = G->deg[x];
    if (!prev_touched.empty()) {
        WinStage& S = C->win;
        const int b = S.h_cur;
        S.h_cur ^= 1;
        HIPCHK(hipEventSynchronize(S.h_done[b]));
        char* hb = (char*)S.h_buf[b];
        SbcWinDesc* hd = (SbcWinDesc*)hb;
        memset(hd, 0, sizeof(SbcWinDesc));
        hd->nvupd = (int32_t)prev_touched.size();
        int32_t* hvv = (int32_t*)(hb + 64) + 5 * 4 * W;
        SbcVInfo* hv = (SbcVInfo*)(hb + 64 + sizeof(int32_t) * 6 * 4 * (size_t)W);
        for (size_t i = 0; i < prev_touched.size(); i++) {
            hvv[i] = prev_touched[i];
            hv[i] = vinfo_of(G, prev_touched[i]);
        }
        // one copy of the block (the window's own ints in it are not read by
        // this launch and are staged again by the next window)
        HIPCHK(hipMemcpyAsync(S.d_base, hb, stage_span(W, prev_touched.size()), hipMemcpyHostToDevice, C->stream));
        HIPCHK(hipEventRecord(S.h_done[b], C->stream));
        SbcAppend B;
        memset(&B, 0, sizeof(B));
        B.vinfo = G->d_vinfo;
        B.col = G->d_col;
        B.seq = G->d_seq;
        B.desc = S.d_desc;
        B.d_vv = S.d_ints + 5 * 4 * W;
        B.d_vi = S.d_vi;
        HIPCHK(sbc_launch_append(&B, (int)prev_touched.size(), C->stream));
    }
    return check_kernel_err(F, C);
}

// ------------------------------------------------------------ status forms

extern "C" sbcStatus sbcGraphCreate(int64_t NV, const int64_t* rowptr, const int64_t* col, struct stinger** out)
{
    if (!rowptr || !out || (!col && NV > 0 && rowptr[NV] > 0)) return set_err(SBC_ERR_ARG, "null argument");
    stinger* G = new stinger();
    sbcStatus st = graph_init(G, NV, rowptr, col);
    if (st) {
        delete G;
        return st;
    }
    *out = G;
    return SBC_OK;
}

extern "C" sbcStatus sbcForestCreate(int64_t NV, int64_t NK, const int64_t* roots, bcForest** out)
{
    if (!out) return set_err(SBC_ERR_ARG, "null argument");
    if (!roots && NK != NV) return set_err(SBC_ERR_ARG, "exact forest needs NK == NV");
    return forest_create(NV, NK, roots, roots == nullptr, out);
}

extern "C" sbcStatus sbcInitStatic(bcForest* forest, struct stinger* G, extraArraysPerThread** eAPT)
{
    if (!eAPT || !eAPT[0]) return set_err(SBC_ERR_ARG, "null context");
    return run_static(forest, G, eAPT[0]);
}

extern "C" sbcStatus sbcInsertEdge(bcForest* forest, struct stinger* G, extraArraysPerThread** eAPT, int64_t u,
                                   int64_t v, uint32_t granularity)
{
    if (!forest || !G || !eAPT || !eAPT[0]) return set_err(SBC_ERR_ARG, "null argument");
    if (!forest->dev->initialized) return set_err(SBC_ERR_STATE, "static pass has not run");
    if (u < 0 || v < 0 || u >= G->nv || v >= G->nv) return set_err(SBC_ERR_ARG, "vertex id out of range");
    if (u == v) return set_err(SBC_ERR_EXISTS, "self loop");
    if (graph_has_edge(G, (int32_t)u, (int32_t)v)) return set_err(SBC_ERR_EXISTS, "edge (%ld,%ld) exists", (long)u, (long)v);
    stinger_insert_edge(G, 0, u, v, 0, 0);
    stinger_insert_edge(G, 0, v, u, 0, 0);
    return run_insert_plain(forest, G, eAPT[0], u, v, granularity);
}

// ------------------------------------------------------------ no-allocation form
//
// sbcInsertEdge with granularity COARSE (one wavefront per root), for callers
// that cannot allocate device memory: its only scratch -- the work-queue
// counters and the item records -- lives in the library's fixed per-GPU
// array (sbc_noalloc_scratch), and it refuses, before changing anything,
// whatever would need an allocation: more roots than the array holds, an
// endpoint without append slack (a relayout), a graph whose device image must
// be uploaded again.  No timing events are created.

extern "C" sbcStatus sbcInsertEdgeNoAlloc(bcForest* forest, struct stinger* G, extraArraysPerThread** eAPT, int64_t u,
                                          int64_t v, sbcItemRecord* records)
{
    if (!forest || !G || !eAPT || !eAPT[0]) return set_err(SBC_ERR_ARG, "null argument");
    extraArraysPerThread* C = eAPT[0];
    sbcForestDev* F = forest->dev;
    if (!F->initialized) return set_err(SBC_ERR_STATE, "static pass has not run");
    if (F->nv != G->nv) return set_err(SBC_ERR_ARG, "forest NV %ld != graph NV %ld", (long)F->nv, (long)G->nv);
    if (F->K > SBC_NOALLOC_MAX_ROOTS)
        return set_err(SBC_ERR_ARG, "%ld roots: the fixed scratch holds %d", (long)F->K, SBC_NOALLOC_MAX_ROOTS);
    if (u < 0 || v < 0 || u >= G->nv || v >= G->nv) return set_err(SBC_ERR_ARG, "vertex id out of range");
    if (u == v) return set_err(SBC_ERR_EXISTS, "self loop");
    if (graph_has_edge(G, (int32_t)u, (int32_t)v)) return set_err(SBC_ERR_EXISTS, "edge (%ld,%ld) exists", (long)u, (long)v);
    if (G->full_dirty || !G->d_vinfo)
        return set_err(SBC_ERR_STATE, "the graph's device image must be uploaded again (an allocating call does that)");
    if (G->deg[u] == G->cap[u] || G->deg[v] == G->cap[v])
        return set_err(SBC_ERR_STATE, "no append slack left at vertex %ld: a relayout allocates",
                       (long)(G->deg[u] == G->cap[u] ? u : v));
    int32_t* qctr = nullptr;
    SbcItemRec* rec = nullptr;
    HIPCHK(sbc_noalloc_scratch(&qctr, &rec));
    stinger_insert_edge(G, 0, u, v, 0, 0);
    stinger_insert_edge(G, 0, v, u, 0, 0);
    sbcStatus st = graph_flush(G, C->stream);
    if (st) return st;
    SbcParams P = base_params(F, G, C);
    P.bucket = nullptr;
    P.hubq = nullptr;
    P.qctr = qctr;
    P.rec = rec;
    P.u0 = (int32_t)u;
    P.v0 = (int32_t)v;
    P.heavy = 0;
    const int nwb = 4;
    const int64_t g = (F->K + nwb - 1) / nwb;
    const int grid = (int)(g < 4LL * C->num_cu ? g : 4LL * C->num_cu);
    HIPCHK(sbc_launch_reset(qctr, 4, C->stream));
    HIPCHK(sbc_launch_window(&P, nwb, grid, C->stream));
    static_assert(sizeof(sbcItemRecord) == sizeof(SbcItemRec), "record layout");
    if (records)
        HIPCHK(hipMemcpyAsync(records, rec, sizeof(SbcItemRec) * (size_t)F->K, hipMemcpyDeviceToHost, C->stream));
    return check_kernel_err(F, C);
}

// ------------------------------------------------------------ cost-model features

extern "C" sbcStatus streamingBCWindowFeatures(bcForest* forest, struct stinger* G, extraArraysPerThread** eAPT,
                                               const int64_t* u, const int64_t* v, int64_t count,
                                               const float* root_hist, float* X)
{
    if (!forest || !G || !eAPT || !eAPT[0] || !X || (count > 0 && (!u || !v)) || count < 0 || count > (1 << 20))
        return set_err(SBC_ERR_ARG, "bad argument");
    extraArraysPerThread* C = eAPT[0];
    sbcForestDev* F = forest->dev;
    if (!F->initialized) return set_err(SBC_ERR_STATE, "static pass has not run");
    if (count == 0) return SBC_OK;
    sbcStatus st = graph_flush(G, C->stream);
    if (st) return st;
    std::vector<int32_t> uv(2 * count);
    for (int64_t i = 0; i < count; i++) {
        if (u[i] < 0 || v[i] < 0 || u[i] >= F->nv || v[i] >= F->nv) return set_err(SBC_ERR_ARG, "vertex out of range");
        uv[i] = (int32_t)u[i];
        uv[count + i] = (int32_t)v[i];
    }
    int32_t* d_uv = nullptr;
    float *d_hist = nullptr, *d_X = nullptr;
    const size_t nx = (size_t)count * F->K * SBC_COST_FEATURES;
    if ((st = dalloc(&d_uv, 2 * count)) || (st = dalloc(&d_X, nx))) return st;
    if (root_hist && (st = dalloc(&d_hist, F->K))) return st;
    HIPCHK(hipMemcpyAsync(d_uv, uv.data(), sizeof(int32_t) * 2 * count, hipMemcpyHostToDevice, C->stream));
    if (root_hist) HIPCHK(hipMemcpyAsync(d_hist, root_hist, sizeof(float) * F->K, hipMemcpyHostToDevice, C->stream));
    SbcFeatArgs A;
    A.vinfo = G->d_vinfo;
    A.dist = F->d_dist;
    A.vr = F->d_vr;
    A.u = d_uv;
    A.v = d_uv + count;
    A.hist = root_hist ? d_hist : nullptr;
    A.X = d_X;
    A.nv = (int32_t)F->nv;
    A.K = (int32_t)F->K;
    A.count = (int32_t)count;
    HIPCHK(sbc_launch_features(&A, C->stream));
    HIPCHK(hipMemcpyAsync(X, d_X, sizeof(float) * nx, hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    (void)hipFree(d_uv);
    (void)hipFree(d_X);
    if (d_hist) (void)hipFree(d_hist);
    return SBC_OK;
}

extern "C" sbcStatus streamingBCSetRootHistory(extraArraysPerThread** eAPT, int64_t K, const float* hist)
{
    if (!eAPT || !eAPT[0] || K < 1 || !hist) return set_err(SBC_ERR_ARG, "bad argument");
    extraArraysPerThread* C = eAPT[0];
    if (C->hist_k != K) {
        HIPCHK(hipStreamSynchronize(C->stream));
        if (C->d_hist) HIPCHK(hipFree(C->d_hist));
        C->d_hist = nullptr;
        C->hist_k = 0;
        sbcStatus st = dalloc(&C->d_hist, (size_t)K);
        if (st) return st;
        C->hist_k = K;
    }
    HIPCHK(hipMemcpyAsync(C->d_hist, hist, sizeof(float) * (size_t)K, hipMemcpyHostToDevice, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    return SBC_OK;
}

extern "C" sbcStatus streamingBCUpdateRootHistory(extraArraysPerThread** eAPT, int64_t K, float decay)
{
    if (!eAPT || !eAPT[0] || K < 1) return set_err(SBC_ERR_ARG, "bad argument");
    extraArraysPerThread* C = eAPT[0];
    if (C->hist_k != K) return set_err(SBC_ERR_STATE, "no device root history of %ld roots", (long)K);
    const int64_t n = C->last_batch_items;
    if (n < 1 || !C->d_rec || C->rec_cap < n * K) return set_err(SBC_ERR_STATE, "no batched window to learn from");
    HIPCHK(sbc_launch_hist_update(C->d_rec, (int)n, (int)K, decay, C->d_hist, C->stream));
    return SBC_OK;
}

extern "C" sbcStatus streamingBCGraphCounts(extraArraysPerThread** eAPT, int64_t* built, int64_t* launched)
{
    if (!eAPT || !eAPT[0]) return set_err(SBC_ERR_ARG, "null context");
    if (built) *built = eAPT[0]->graphs_built;
    if (launched) *launched = eAPT[0]->graph_launches;
    return SBC_OK;
}

extern "C" int64_t streamingBCLastParked(extraArraysPerThread** eAPT)
{
    return (eAPT && eAPT[0]) ? eAPT[0]->last_parked : -1;
}

extern "C" sbcStatus streamingBCWorkTotals(extraArraysPerThread** eAPT, int64_t K, int64_t* out)
{
    if (!eAPT || !eAPT[0] || !out || K < 1) return set_err(SBC_ERR_ARG, "bad argument");
    extraArraysPerThread* C = eAPT[0];
    if (K > C->wtot_k) return set_err(SBC_ERR_ARG, "no work totals for %ld roots", (long)K);
    HIPCHK(hipMemcpyAsync(out, C->d_wtot, sizeof(int64_t) * 8 * (size_t)K, hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    return SBC_OK;
}

extern "C" int32_t streamingBCComputeUnits(extraArraysPerThread** eAPT)
{
    return (eAPT && eAPT[0]) ? eAPT[0]->num_cu : -1;
}

// ------------------------------------------------------------ device scheduling

extern "C" sbcStatus streamingBCScheduleWindow(bcForest* forest, struct stinger* G, extraArraysPerThread** eAPT,
                                               sbcCostModel* model, const int64_t* u, const int64_t* v,
                                               int64_t count, const float* root_hist, float light_max,
                                               float item_overhead, int32_t* heavy_out, float* pred_out)
{
    if (!forest || !G || !eAPT || !eAPT[0] || !model || count <= 0 || count > (1 << 16) || !u || !v)
        return set_err(SBC_ERR_ARG, "bad argument");
    extraArraysPerThread* C = eAPT[0];
    sbcForestDev* F = forest->dev;
    if (!F->initialized) return set_err(SBC_ERR_STATE, "static pass has not run");
    if (F->K > 4096) return set_err(SBC_ERR_ARG, "device scheduling supports at most 4096 roots");
    sbcStatus st = graph_flush(G, C->stream);
    if (st) return st;
    if ((st = ensure_order(C, F->K))) return st;
    const int64_t items = count * F->K;
    if (C->sched_items < items || C->sched_k < F->K || C->sched_cnt < count) {
        HIPCHK(hipStreamSynchronize(C->stream));
        for (void* q : {(void*)C->d_suv, (void*)C->d_shist, (void*)C->d_sX, (void*)C->d_spred, (void*)C->d_scost,
                        (void*)C->d_sheavy})
            if (q) HIPCHK(hipFree(q));
        C->d_suv = nullptr; C->d_shist = nullptr; C->d_sX = nullptr; C->d_spred = nullptr; C->d_scost = nullptr;
        C->d_sheavy = nullptr;
        if (C->h_suv) HIPCHK(hipHostFree(C->h_suv));
        C->h_suv = nullptr;
        const int64_t ni = items > C->sched_items ? items : C->sched_items;
        const int64_t nk = F->K > C->sched_k ? F->K : C->sched_k;
        const int64_t nc = count > C->sched_cnt ? count : C->sched_cnt;
        C->sched_items = C->sched_k = C->sched_cnt = 0;  // until every buffer is in place
        if ((st = dalloc(&C->d_suv, (size_t)(2 * nc))) || (st = dalloc(&C->d_shist, (size_t)nk)) ||
            (st = dalloc(&C->d_sX, (size_t)ni * SBC_COST_FEATURES)) || (st = dalloc(&C->d_spred, (size_t)ni)) ||
            (st = dalloc(&C->d_scost, (size_t)nk)) || (st = dalloc(&C->d_sheavy, 1)))
            return st;
        HIPCHK(hipHostMalloc((void**)&C->h_suv, sizeof(int32_t) * 2 * (size_t)nc));
        C->sched_items = ni;
        C->sched_k = nk;
        C->sched_cnt = nc;
    }
    // the last call's copy out of h_suv ended with that call's synchronisation
    int32_t* uv = C->h_suv;
    for (int64_t i = 0; i < count; i++) {
        if (u[i] < 0 || v[i] < 0 || u[i] >= F->nv || v[i] >= F->nv) return set_err(SBC_ERR_ARG, "vertex out of range");
        uv[i] = (int32_t)u[i];
        uv[count + i] = (int32_t)v[i];
    }
    HIPCHK(hipMemcpyAsync(C->d_suv, uv, sizeof(int32_t) * 2 * count, hipMemcpyHostToDevice, C->stream));
    if (root_hist)
        HIPCHK(hipMemcpyAsync(C->d_shist, root_hist, sizeof(float) * F->K, hipMemcpyHostToDevice, C->stream));
    SbcFeatArgs A;
    A.vinfo = G->d_vinfo;
    A.dist = F->d_dist;
    A.vr = F->d_vr;
    A.u = C->d_suv;
    A.v = C->d_suv + count;
    A.hist = root_hist ? C->d_shist : (C->hist_k == F->K ? C->d_hist : nullptr);
    A.X = C->d_sX;
    A.nv = (int32_t)F->nv;
    A.K = (int32_t)F->K;
    A.count = (int32_t)count;
    HIPCHK(sbc_launch_features(&A, C->stream));
    HIPCHK(sbc_cost_schedule(model, C->d_sX, (int)count, (int)F->K, light_max > 0 ? light_max : -1.0f, item_overhead,
                             C->d_spred, C->d_scost, C->d_order, C->d_sheavy, C->stream));
    HIPCHK(hipMemcpyAsync(C->h_heavy, C->d_sheavy, sizeof(int32_t), hipMemcpyDeviceToHost, C->stream));
    if (pred_out)
        HIPCHK(hipMemcpyAsync(pred_out, C->d_spred, sizeof(float) * (size_t)items, hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    const int32_t heavy = *C->h_heavy;
    if (heavy_out) *heavy_out = light_max > 0 ? heavy : -1;
    C->order_ready = true;
    C->order_forest = F;
    return SBC_OK;
}

// ------------------------------------------------------------ vertex updates

// src/streamingbc.c:56-77: connects src to every listed vertex.  The edges not
// yet in the graph (self loops, bad ids and duplicates skipped: the reference
// would run the update again on an edge already there) are inserted in list
// order by one batched call, whose trees equal those of one update per edge.
static sbcStatus insert_vertex(bcForest* forest, stinger* G, extraArraysPerThread** eAPT, int64_t src,
                               const int64_t* adj, int64_t n)
{
    if (!forest || !G || !eAPT || !eAPT[0] || (n > 0 && !adj)) return set_err(SBC_ERR_ARG, "null argument");
    if (src < 0 || src >= G->nv || n < 0) return set_err(SBC_ERR_ARG, "vertex id or list size out of range");
    std::vector<int64_t> u, v;
    std::vector<char> seen;
    for (int64_t d = 0; d < n; d++) {
        const int64_t x = adj[d];
        if (x < 0 || x >= G->nv || x == src || graph_has_edge(G, (int32_t)src, (int32_t)x)) continue;
        if (seen.empty()) seen.assign((size_t)G->nv, 0);
        if (seen[(size_t)x]) continue;
        seen[(size_t)x] = 1;
        u.push_back(src);
        v.push_back(x);
    }
    if (u.empty()) return SBC_OK;
    sbcBatchConfig cfg;
    memset(&cfg, 0, sizeof(cfg));
    cfg.window = u.size() < 128 ? (int32_t)u.size() : 128;
    cfg.block_threads = 1024;
    cfg.heavy_roots = -1;
    return streamingBCInsertEdgesBatched(forest, G, eAPT, u.data(), v.data(), (int64_t)u.size(), &cfg, nullptr);
}

// src/streamingbc.c:229-256: removes every edge of src in list order, one
// deletion update each.  The removed neighbours go to adj (at most cap of
// them; more is an error before anything changes) and their number to *n;
// counts sum the deletions' case counts.  (The reference never advances its
// output index, so it reports 0 neighbours and keeps only the last one.)
static sbcStatus delete_vertex(bcForest* forest, stinger* G, extraArraysPerThread** eAPT, int64_t src,
                               int64_t* adj, int64_t cap, int64_t* n, int64_t counts[3])
{
    if (!forest || !G || !eAPT || !eAPT[0]) return set_err(SBC_ERR_ARG, "null argument");
    if (src < 0 || src >= G->nv) return set_err(SBC_ERR_ARG, "vertex id out of range");
    const std::vector<int32_t> nb(G->col.begin() + G->start[src], G->col.begin() + G->start[src] + G->deg[src]);
    if (adj && (int64_t)nb.size() > cap)
        return set_err(SBC_ERR_ARG, "vertex %ld has %ld neighbours, the array holds %ld", (long)src, (long)nb.size(),
                       (long)cap);
    counts[0] = counts[1] = counts[2] = 0;
    int64_t d = 0;
    for (int32_t x : nb) {
        stinger_remove_edge(G, 0, src, x);
        stinger_remove_edge(G, 0, x, src);
        int64_t c[3];
        sbcStatus st = run_delete(forest, G, eAPT[0], src, x, c);
        if (st) return st;
        for (int i = 0; i < 3; i++) counts[i] += c[i];
        if (adj) adj[d] = x;
        d++;
    }
    if (n) *n = d;
    return SBC_OK;
}

extern "C" StreamingExtraInfo insertVertexStreamingBC(bcForest* forest, struct stinger* sStinger, int64_t src,
                                                      int64_t* adjacencyArray, int64_t adjacencySize,
                                                      int64_t* rootArrayForApproximation, int64_t NK, int64_t NV,
                                                      int64_t NT, extraArraysPerThread** eAPT)
{
    (void)rootArrayForApproximation; (void)NT;
    StreamingExtraInfo r = {0, 0, 0, 0};
    if (!forest || NK != forest->dev->K || NV != forest->dev->nv) {
        set_err(SBC_ERR_ARG, "NK/NV do not match the forest");
        legacy_abort("insertVertexStreamingBC");
    }
    if (insert_vertex(forest, sStinger, eAPT, src, adjacencyArray, adjacencySize) != SBC_OK)
        legacy_abort("insertVertexStreamingBC");
    legacy_refresh_bc(forest, "insertVertexStreamingBC");
    return r;
}

extern "C" StreamingExtraInfo deleteVertexStreamingBC(bcForest* forest, struct stinger* sStinger, int64_t src,
                                                      int64_t* adjacencyArray, int64_t* adjacencySize,
                                                      int64_t* rootArrayForApproximation, int64_t NK, int64_t NV,
                                                      int64_t NT, extraArraysPerThread** eAPT)
{
    (void)rootArrayForApproximation; (void)NT;
    StreamingExtraInfo r = {0, 0, 0, 0};
    if (!forest || !sStinger || NK != forest->dev->K || NV != forest->dev->nv || src < 0 || src >= sStinger->nv) {
        set_err(SBC_ERR_ARG, "NK/NV do not match the forest, or bad vertex");
        legacy_abort("deleteVertexStreamingBC");
    }
    // capacity: *adjacencySize when positive on entry, else the degree
    const int64_t cap = adjacencySize && *adjacencySize > 0 ? *adjacencySize : sStinger->deg[src];
    int64_t c[3];
    if (delete_vertex(forest, sStinger, eAPT, src, adjacencyArray, cap, adjacencySize, c) != SBC_OK)
        legacy_abort("deleteVertexStreamingBC");
    legacy_refresh_bc(forest, "deleteVertexStreamingBC");
    r.sameLevel = c[0];
    r.adjacent = c[1];
    r.movement = c[2];
    return r;
}

extern "C" sbcStatus sbcInsertVertex(bcForest* forest, struct stinger* G, extraArraysPerThread** eAPT, int64_t src,
                                     const int64_t* adj, int64_t n)
{
    if (forest && !forest->dev->initialized) return set_err(SBC_ERR_STATE, "static pass has not run");
    return insert_vertex(forest, G, eAPT, src, adj, n);
}

extern "C" sbcStatus sbcDeleteVertex(bcForest* forest, struct stinger* G, extraArraysPerThread** eAPT, int64_t src,
                                     int64_t* adj, int64_t cap, int64_t* n, StreamingExtraInfo* info)
{
    if (forest && !forest->dev->initialized) return set_err(SBC_ERR_STATE, "static pass has not run");
    int64_t c[3] = {0, 0, 0};
    sbcStatus st = delete_vertex(forest, G, eAPT, src, adj, cap, n, c);
    if (st == SBC_OK && info) {
        info->sameLevel = c[0];
        info->adjacent = c[1];
        info->movement = c[2];
        info->connectedComponents = 0;
    }
    return st;
}

// ------------------------------------------------------------ checkpoint / resume
//
// One file per resident graph and its forest (the reference's only precedent
// is STINGER's own save/open, src/stinger/stinger.c:1808-2038).  Layout,
// little-endian: "SBCCKPT3", int64 {version 3, nv, K, exact, initialized,
// slot count, table slot count, next sequence number}; then sections, each
// followed by the uint64 checksum of its bytes (ckpt_hash): roots[K] (int64);
// the graph's host mirror: start/deg/deg0/cap[nv] (int32), col/seq[slots]
// (int32), hub[nv] (int4), htab[table slots] (int2); per root: dist[nv]
// (int32), state bytes[nv], records[nv] (32 B); BC[nv] (double); then
// "SBCEND03" (version 3: hub tables in 8-slot buckets).  Loading checks the header, the exact file size, every
// section's checksum and the ranges of the graph and the trees (list bounds,
// levels below the level capacity or unreachable, committed state bytes, a
// root at level 0) in a first pass, before any device allocation.

static const char kCkptMagic[8] = {'S', 'B', 'C', 'C', 'K', 'P', 'T', '3'};
static const char kCkptEnd[8] = {'S', 'B', 'C', 'E', 'N', 'D', '0', '3'};

// Streaming 64-bit checksum over a byte sequence: 8-byte little-endian words
// (the tail zero-padded), each xored in, multiplied by the FNV-64 prime and
// folded; the length closes it.
struct CkptHash {
    uint64_t h = 0x9E3779B97F4A7C15ULL, len = 0;
    uint8_t buf[8];
    int nb = 0;
    void word(uint64_t w)
    {
        h ^= w;
        h *= 0x100000001B3ULL;
        h ^= h >> 32;
    }
    void add(const void* p, size_t n)
    {
        const uint8_t* b = (const uint8_t*)p;
        len += n;
        while (nb && n) {
            buf[nb++] = *b++;
            n--;
            if (nb == 8) { uint64_t w; memcpy(&w, buf, 8); word(w); nb = 0; }
        }
        for (; n >= 8; n -= 8, b += 8) { uint64_t w; memcpy(&w, b, 8); word(w); }
        while (n) { buf[nb++] = *b++; n--; }
    }
    uint64_t fin()
    {
        if (nb) {
            memset(buf + nb, 0, (size_t)(8 - nb));
            uint64_t w;
            memcpy(&w, buf, 8);
            word(w);
            nb = 0;
        }
        word(len);
        return h;
    }
};

static int64_t ckpt_size(int64_t nv, int64_t K, int64_t slots, int64_t tslots)
{
    return 8 + 8 * 8 + (8 * K + 8) + (16 * nv + 8 * slots + 16 * nv + 8 * tslots + 8) +
           K * (nv * (4 + 1 + (int64_t)sizeof(SbcVRec)) + 8) + (8 * nv + 8) + 8;
}

extern "C" sbcStatus sbcCheckpointSave(const bcForest* forest, const struct stinger* G, const char* path)
{
    if (!forest || !G || !path) return set_err(SBC_ERR_ARG, "null argument");
    const sbcForestDev* F = forest->dev;
    if (F->nv != G->nv) return set_err(SBC_ERR_ARG, "forest NV %ld != graph NV %ld", (long)F->nv, (long)G->nv);
    sbcForestDev* Fw = forest->dev;
    FILE* fp = fopen(path, "wb");
    if (!fp) return set_err(SBC_ERR_ARG, "cannot open %s for writing", path);
    const int64_t nv = F->nv, K = F->K;
    const int64_t hdr[8] = {3, nv, K, F->exact ? 1 : 0, F->initialized ? 1 : 0, G->total_cap,
                            (int64_t)G->htab.size(), G->seq_next};
    CkptHash ck;
    auto put = [&](const void* p, size_t sz, size_t n) {
        ck.add(p, sz * n);
        return fwrite(p, sz, n, fp) == n;
    };
    auto seal = [&]() {
        const uint64_t c = ck.fin();
        ck = CkptHash();
        return fwrite(&c, 8, 1, fp) == 1;
    };
    bool ok = fwrite(kCkptMagic, 1, 8, fp) == 8 && fwrite(hdr, 8, 8, fp) == 8 &&
              put(F->roots.data(), 8, (size_t)K) && seal();
    for (const std::vector<int32_t>* a : {&G->start, &G->deg, &G->deg0, &G->cap})
        ok = ok && put(a->data(), 4, (size_t)nv);
    ok = ok && put(G->col.data(), 4, (size_t)G->total_cap) && put(G->seq.data(), 4, (size_t)G->total_cap) &&
         put(G->hub.data(), sizeof(int4), (size_t)nv) && put(G->htab.data(), sizeof(int2), G->htab.size()) && seal();
    // per root, through one host buffer of nv records
    std::vector<int32_t> d((size_t)nv);
    std::vector<uint8_t> b((size_t)nv);
    std::vector<SbcVRec> r((size_t)nv);
    for (int64_t ri = 0; ok && ri < K; ri++) {
        const size_t off = (size_t)ri * (size_t)nv;
        if (forest_read(Fw, d.data(), F->d_dist + off, 4 * (size_t)nv) ||
            forest_read(Fw, b.data(), F->d_sb + (size_t)ri * (size_t)F->sb_stride, (size_t)nv) ||
            forest_read(Fw, r.data(), F->d_vr + off, sizeof(SbcVRec) * (size_t)nv)) {
            fclose(fp);
            return set_err(SBC_ERR_HIP, "checkpoint: device copy failed");
        }
        ok = put(d.data(), 4, (size_t)nv) && put(b.data(), 1, (size_t)nv) &&
             put(r.data(), sizeof(SbcVRec), (size_t)nv) && seal();
    }
    std::vector<double> bc((size_t)nv);
    if (ok && forest_read(Fw, bc.data(), F->d_bc, 8 * (size_t)nv)) {
        fclose(fp);
        return set_err(SBC_ERR_HIP, "checkpoint: device copy failed");
    }
    ok = ok && put(bc.data(), 8, (size_t)nv) && seal() && fwrite(kCkptEnd, 1, 8, fp) == 8;
    if (fclose(fp) != 0) ok = false;
    if (!ok) return set_err(SBC_ERR_ARG, "checkpoint: writing %s failed", path);
    return SBC_OK;
}

// Reads n items of sz bytes into p and adds them to the running checksum.
static bool ckpt_get(FILE* fp, CkptHash& ck, void* p, size_t sz, size_t n)
{
    if (fread(p, sz, n, fp) != n) return false;
    ck.add(p, sz * n);
    return true;
}

static bool ckpt_check(FILE* fp, CkptHash& ck)
{
    uint64_t c = 0;
    const bool ok = fread(&c, 8, 1, fp) == 1 && c == ck.fin();
    ck = CkptHash();
    return ok;
}

extern "C" sbcStatus sbcCheckpointLoad(const char* path, bcForest** forest, struct stinger** graph)
{
    if (!path || !forest || !graph) return set_err(SBC_ERR_ARG, "null argument");
    FILE* fp = fopen(path, "rb");
    if (!fp) return set_err(SBC_ERR_ARG, "cannot open %s", path);
    char magic[8];
    int64_t hdr[8];
    bool ok = fread(magic, 1, 8, fp) == 8 && memcmp(magic, kCkptMagic, 8) == 0 && fread(hdr, 8, 8, fp) == 8;
    const int64_t nv = ok ? hdr[1] : 0, K = ok ? hdr[2] : 0, slots = ok ? hdr[5] : 0, tslots = ok ? hdr[6] : 0;
    ok = ok && hdr[0] == 3 && nv >= 1 && nv < INT32_MAX / 4 && K >= 1 && K <= nv && slots >= 0 &&
         slots < INT32_MAX && tslots >= 0 && tslots < INT32_MAX && hdr[7] >= 1 && hdr[7] < INT32_MAX;
    if (ok) {
        // the exact size, before anything is allocated
        if (fseek(fp, 0, SEEK_END) != 0) ok = false;
        const long long sz = ok ? (long long)ftello(fp) : -1;
        ok = ok && sz == ckpt_size(nv, K, slots, tslots) && fseek(fp, 8 + 64, SEEK_SET) == 0;
    }
    if (!ok) {
        fclose(fp);
        return set_err(SBC_ERR_ARG, "%s is not a checkpoint of this format, or is truncated", path);
    }
    std::vector<int64_t> roots((size_t)K);
    stinger* G = new stinger();
    G->nv = nv;
    G->total_cap = slots;
    G->seq_next = (int32_t)hdr[7];
    for (std::vector<int32_t>* a : {&G->start, &G->deg, &G->deg0, &G->cap}) a->resize((size_t)nv);
    G->col.resize((size_t)slots);
    G->seq.resize((size_t)slots);
    G->hub.resize((size_t)nv);
    G->htab.resize((size_t)tslots);
    auto fail = [&](const char* why) {
        fclose(fp);
        delete G;
        return set_err(SBC_ERR_ARG, "%s: %s", path, why);
    };
    CkptHash ck;
    if (!ckpt_get(fp, ck, roots.data(), 8, (size_t)K) || !ckpt_check(fp, ck)) return fail("root section fails its checksum");
    ok = true;
    for (std::vector<int32_t>* a : {&G->start, &G->deg, &G->deg0, &G->cap})
        ok = ok && ckpt_get(fp, ck, a->data(), 4, (size_t)nv);
    ok = ok && ckpt_get(fp, ck, G->col.data(), 4, (size_t)slots) && ckpt_get(fp, ck, G->seq.data(), 4, (size_t)slots) &&
         ckpt_get(fp, ck, G->hub.data(), sizeof(int4), (size_t)nv) &&
         ckpt_get(fp, ck, G->htab.data(), sizeof(int2), (size_t)tslots) && ckpt_check(fp, ck);
    if (!ok) return fail("graph section fails its checksum");
    // the graph arrays must describe lists inside the slots and tables inside htab
    for (int64_t v = 0; ok && v < nv; v++) {
        const int4 h = G->hub[v];
        ok = G->start[v] >= 0 && G->cap[v] >= 0 && G->deg[v] >= 0 && G->deg0[v] >= 0 && G->deg0[v] <= G->deg[v] &&
             G->deg[v] <= G->cap[v] && (int64_t)G->start[v] + G->cap[v] <= slots &&
             (h.y == 0 || (h.z >= 3 && h.z < 31 && h.x >= 0 && (h.x & 7) == 0 &&
                           (int64_t)h.x + (1LL << h.z) <= tslots && h.y <= G->deg[v]));
    }
    for (int64_t i = 0; ok && i < K; i++) ok = roots[(size_t)i] >= 0 && roots[(size_t)i] < nv;
    for (int64_t e = 0; ok && e < slots; e++) ok = G->col[(size_t)e] >= 0 && G->col[(size_t)e] < nv;
    if (!ok) return fail("inconsistent graph in checkpoint");
    // first pass over the trees: checksums and ranges, nothing allocated yet
    const long long trees_at = (long long)ftello(fp);
    const int32_t lcap = lcap_for(nv);
    std::vector<int32_t> d((size_t)nv);
    std::vector<uint8_t> b((size_t)nv);
    std::vector<SbcVRec> r((size_t)nv);
    for (int64_t ri = 0; ri < K; ri++) {
        if (!ckpt_get(fp, ck, d.data(), 4, (size_t)nv) || !ckpt_get(fp, ck, b.data(), 1, (size_t)nv) ||
            !ckpt_get(fp, ck, r.data(), sizeof(SbcVRec), (size_t)nv) || !ckpt_check(fp, ck))
            return fail("a tree section fails its checksum");
        bool good = d[(size_t)roots[(size_t)ri]] == 0;
        for (int64_t x = 0; good && x < nv; x++)
            good = (d[(size_t)x] == INT32_MAX || (d[(size_t)x] >= 0 && d[(size_t)x] < lcap)) && (b[(size_t)x] & 3) == 0 &&
                   r[(size_t)x].dpath == 0 && r[(size_t)x].ndel == 0.0;
        if (!good) return fail("a tree holds a level out of range, a root off level 0 or uncommitted state");
    }
    std::vector<double> bc((size_t)nv);
    if (!ckpt_get(fp, ck, bc.data(), 8, (size_t)nv) || !ckpt_check(fp, ck)) return fail("BC section fails its checksum");
    if (fread(magic, 1, 8, fp) != 8 || memcmp(magic, kCkptEnd, 8) != 0) return fail("end marker missing");
    G->full_dirty = true;
    bcForest* f = nullptr;
    sbcStatus st = forest_create(nv, K, roots.data(), hdr[3] != 0, &f);
    if (st) {
        fclose(fp);
        delete G;
        return st;
    }
    sbcForestDev* F = f->dev;
    ok = fseeko(fp, (off_t)trees_at, SEEK_SET) == 0;
    for (int64_t ri = 0; ok && ri < K; ri++) {
        const size_t off = (size_t)ri * (size_t)nv;
        uint64_t c;
        ok = fread(d.data(), 4, (size_t)nv, fp) == (size_t)nv && fread(b.data(), 1, (size_t)nv, fp) == (size_t)nv &&
             fread(r.data(), sizeof(SbcVRec), (size_t)nv, fp) == (size_t)nv && fread(&c, 8, 1, fp) == 1 &&
             forest_write(F, F->d_dist + off, d.data(), 4 * (size_t)nv) == SBC_OK &&
             forest_write(F, F->d_sb + (size_t)ri * (size_t)F->sb_stride, b.data(), (size_t)nv) == SBC_OK &&
             forest_write(F, F->d_vr + off, r.data(), sizeof(SbcVRec) * (size_t)nv) == SBC_OK;
    }
    fclose(fp);
    memcpy(f->totalBC, bc.data(), 8 * (size_t)nv);
    ok = ok && forest_write(F, F->d_bc, f->totalBC, 8 * (size_t)nv) == SBC_OK;
    if (!ok) {
        forest_destroy(&f);
        delete G;
        return set_err(SBC_ERR_HIP, "%s: uploading the forest failed", path);
    }
    F->initialized = hdr[4] != 0;
    *forest = f;
    *graph = G;
    return SBC_OK;
}
