// sbc_kernels.hip -- gfx950 kernels of the streaming-BC engine.
//
// Per root r the forest holds, in HBM, slices of NV entries: dist (int32
// level, INT32_MAX = unreachable), sb (one state byte, see below; gathered
// for every neighbour while scanning adjacency lists, dense so that popular
// vertices stay cache-resident), vr (one 32-byte record per vertex: sigma
// (int64, LONG_MAX = unreachable), delta (FP64) and the scratch used only
// while an update of that root runs, dpath (int64, the added-path count the
// reference calls diffPath) and ndel (FP64, the new delta); an ascent term
// reads all four from one cache line) and list (int32, the processed
// vertices).  Scratch is cleared on the touched entries only, at commit, as
// src/streamingbc_internal.c:213-242 does.
//
// One update item = (root, inserted edge).  A team (one wavefront, or the
// whole workgroup) runs it:
//   classify  gap = dist[u] - dist[v]  (src/streamingbc.c:98-118)
//   descent   level-synchronous frontier expansion from the deeper endpoint s;
//             frontier j holds exactly the vertices whose new level is
//             dist[p] + 1 + j, so no level re-sort is needed
//             (src/streamingbc_internal.c:294-400); sigma gains are
//             accumulated with int64 atomics (exact in any order)
//   ascent    levels from the deepest descent level up to the root; each
//             vertex pulls from its processed children in adjacency order
//             (src/streamingbc_internal_fine.c:560-999), with the terms
//             summed in that order by walking the wavefront's lanes, so the
//             new delta is bit-identical to the sequential FINE order
//   commit    level/sigma/delta written back, BC updated, scratch cleared.
// Vertices with many neighbours are scanned by every wavefront of a
// multi-wave team; their ascent terms are staged per wave in LDS and summed
// in wave order, which is again adjacency order.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <mutex>
#include <type_traits>

#include "sbc_kernels.h"

#define INF32 0x7fffffff
#define SIGMA_INF 0x7fffffffffffffffll

// ------------------------------------------------------------ primitives

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Per-root arrays are only ever touched by the one workgroup that runs the
// root (kernel boundaries order them across launches), so their atomics and
// coherent loads use workgroup scope and stay in the XCD's L2.  Only data
// shared between workgroups (BC, queue heads, error flag) uses device scope.
#define WG_SCOPE __HIP_MEMORY_SCOPE_WORKGROUP
__device__ __forceinline__ int32_t ld_acq32(const int32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, WG_SCOPE); }
__device__ __forceinline__ int64_t ld_acq64(const int64_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, WG_SCOPE); }
__device__ __forceinline__ double ld_accd(const double* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, WG_SCOPE); }
__device__ __forceinline__ void st_rel32(int32_t* p, int32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, WG_SCOPE); }
__device__ __forceinline__ void st_rel64(int64_t* p, int64_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, WG_SCOPE); }
__device__ __forceinline__ void st_reld(double* p, double v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, WG_SCOPE); }
__device__ __forceinline__ void add64_wg(int64_t* p, int64_t v)
{
    __hip_atomic_fetch_add((unsigned long long*)p, (unsigned long long)v, __ATOMIC_RELAXED, WG_SCOPE);
}
__device__ __forceinline__ int32_t cas32_wg(int32_t* p, int32_t expect, int32_t want)
{
    __hip_atomic_compare_exchange_strong(p, &expect, want, __ATOMIC_RELAXED, __ATOMIC_RELAXED, WG_SCOPE);
    return expect; // the value seen
}

__device__ __forceinline__ double readlane_d(double x, int l)
{
    int2 b = __builtin_bit_cast(int2, x);
    b.x = __builtin_amdgcn_readlane(b.x, l);
    b.y = __builtin_amdgcn_readlane(b.y, l);
    return __builtin_bit_cast(double, b);
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int prefix_in_wave(uint64_t m)
{
    const int l = lane_id();
    return __popcll(l == 0 ? 0ull : (m & (~0ull >> (64 - l))));
}

// acc + the wavefront's terms in lane order, where lane l's term adds `add`
// if c, then subtracts `sub` if s: the terms are packed in order into the
// wavefront's LDS row (a subtraction as the addition of the negated value,
// which IEEE 754 defines it to be) and lane 0 adds them one after another.
// acc must be the same in every lane; so is the result.  Wave-uniform
// control flow only.
// Packing rows of the ordered sums: one row of 128 values per wavefront of
// the workgroup (a single wavefront may use all of them while the others
// wait at a barrier).
__shared__ double sbc_wrow[16][128];

// lane 0's sequential sum acc + v[0] + ... + v[n - 1] over values in LDS
// (loads run ahead of the add chain); the result is that of the plain loop.
__device__ __forceinline__ double lds_chain_sum(double acc, const double* v, int n)
{
    int i = 0;
    for (; i + 8 <= n; i += 8) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = v[i + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = acc + x[u];
    }
    for (; i < n; ++i) acc = acc + v[i];
    return acc;
}

// The same sum for long runs: two register buffers of CH values take turns,
// one being added while the other's loads are in flight (no copies between
// them, so each wait covers only the buffer about to be added).
template <int CH>
__device__ __forceinline__ double lds_chain_sum_pipe(double acc, const double* v, int n)
{
    int i = 0;
    if (n >= 2 * CH) {
        double a[CH], b[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) a[u] = v[u];
        for (i = CH; i + 2 * CH <= n; i += 2 * CH) {
#pragma unroll
            for (int u = 0; u < CH; ++u) b[u] = v[i + u];
#pragma unroll
            for (int u = 0; u < CH; ++u) acc = acc + a[u];
#pragma unroll
            for (int u = 0; u < CH; ++u) a[u] = v[i + CH + u];
#pragma unroll
            for (int u = 0; u < CH; ++u) acc = acc + b[u];
        }
        // a holds v[i - CH, i)
#pragma unroll
        for (int u = 0; u < CH; ++u) acc = acc + a[u];
    }
    for (; i < n; ++i) acc = acc + v[i];
    return acc;
}

__device__ __forceinline__ double wave_ordered_add(double acc, bool c, double add, bool s, double sub)
{
    double (&wrow)[16][128] = sbc_wrow;
    const uint64_t cm = ballot(c), sm = ballot(s);
    if ((cm | sm) == 0) return acc;
    const int lane = lane_id();
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    int pos = __popcll(cm & below) + __popcll(sm & below);
    double* w = wrow[threadIdx.x >> 6];
    if (c) w[pos++] = add;
    if (s) w[pos] = -sub;
    const int n = __popcll(cm) + __popcll(sm);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double r = acc;
    if (lane == 0) r = lds_chain_sum(acc, w, n);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return readlane_d(r, 0);
}

// Appends the lanes with `flag` set to a list whose length lives in LDS.
// Must be called by the whole wavefront.  Returns the slot or -1.
__device__ __forceinline__ int wave_append(bool flag, int* ctr)
{
    uint64_t m = ballot(flag);
    if (m == 0) return -1;
    int base = 0;
    if (lane_id() == 0) base = atomicAdd(ctr, __popcll(m));
    base = __shfl(base, 0);
    return flag ? base + prefix_in_wave(m) : -1;
}

// Reserves n >= 0 consecutive slots per lane from a counter with one atomic
// per wavefront; returns the lane's first slot.  Whole wavefront.
__device__ __forceinline__ int wave_reserve(int n, int* ctr)
{
    const int lane = lane_id();
    int incl = n;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(incl, d);
        if (lane >= d) incl += y;
    }
    const int tot = __shfl(incl, 63);
    int base = 0;
    if (tot > 0 && lane == 0) base = atomicAdd(ctr, tot);
    base = __shfl(base, 0);
    return base + incl - n;
}

// This lane's wavefront in the workgroup, as a wavefront-uniform (scalar)
// value: loops indexed by it, or bounded by other readfirstlane values, are
// uniform loops, so no lane-dependent branch inside can split a wavefront
// around a cross-lane operation.
__device__ __forceinline__ int wave_in_block() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

__device__ __forceinline__ int32_t wave_min32(int32_t x)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const int32_t t = __shfl_xor(x, d);
        x = t < x ? t : x;
    }
    return x;
}

__device__ __forceinline__ int32_t wave_max32(int32_t x)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const int32_t t = __shfl_xor(x, d);
        x = t > x ? t : x;
    }
    return x;
}

__device__ __forceinline__ int64_t wave_sum64(int64_t x)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x = (int64_t)((uint64_t)x + (uint64_t)__shfl_xor(x, d));
    return x;
}

template <int TW>
__device__ __forceinline__ void team_sync()
{
    if constexpr (TW == 1) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    } else {
        __syncthreads();
    }
}

// Per-(root, vertex) state byte, the only per-root array gathered while
// scanning adjacency lists:
//   bit 7    unreachable (level infinite)
//   bits 6:2 level mod 32 (the new level once the vertex is marked)
//   bits 1:0 type: 0 untouched, 1 descent vertex whose level is unchanged,
//            2 descent vertex that moved up, 3 ascent-only vertex (a new
//            parent of a processed vertex, or an unmoved vertex now level
//            with a moved neighbour).
// The full level lives in the int32 dist array, read only for the vertex
// being processed.  Neighbours in a BFS tree differ by at most one level, so
// level mod 32 determines a neighbour's level relative to a known one; the
// one exception, the edge being inserted, is handled explicitly.
#define SB_INF 0x80
__device__ __forceinline__ uint8_t sb_make(int32_t level, int type)
{
    return (uint8_t)(((level & 31) << 2) | type);
}
__device__ __forceinline__ int sb_type(uint8_t b) { return b & 3; }
// level of a finite byte relative to `ref`, valid when it lies in [ref-16, ref+15]
__device__ __forceinline__ int sb_rel(uint8_t b, int32_t ref)
{
    const int d = (((b >> 2) & 31) - (ref & 31)) & 31;
    return d >= 16 ? d - 32 : d;
}

__device__ __forceinline__ uint8_t ld_sb(const uint8_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, WG_SCOPE);
}
__device__ __forceinline__ void st_sb(uint8_t* p, uint8_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, WG_SCOPE);
}

// Compare-and-swap of one state byte through its aligned 32-bit word.
// Returns the byte seen; the swap happened iff that equals `expect`.  Bytes
// only move from type 0 to a marked type once per update, so the retry loop
// ends after at most four rounds.
__device__ __forceinline__ uint8_t cas_sb(uint8_t* p, uint8_t expect, uint8_t want)
{
    unsigned int* w = (unsigned int*)((uintptr_t)p & ~(uintptr_t)3);
    const int sh = (int)((uintptr_t)p & 3) * 8;
    unsigned int old = __hip_atomic_load(w, __ATOMIC_RELAXED, WG_SCOPE);
    for (;;) {
        const uint8_t cur = (uint8_t)(old >> sh);
        if (cur != expect) return cur;
        const unsigned int nw = (old & ~(0xffu << sh)) | ((unsigned int)want << sh);
        unsigned int prev = old;
        __hip_atomic_compare_exchange_strong(w, &prev, nw, __ATOMIC_RELAXED, __ATOMIC_RELAXED, WG_SCOPE);
        if (prev == old) return expect;
        old = prev;
    }
}

// Marks a vertex as an ascent-only parent (type 3) with one atomic OR of
// its type bits: its state byte must have type 0 or 3 and already hold the
// parent level (a neighbour one level up that no descent touched).  True if
// this call set the type.
__device__ __forceinline__ bool mark_sb3(uint8_t* p)
{
    unsigned int* w = (unsigned int*)((uintptr_t)p & ~(uintptr_t)3);
    const int sh = (int)((uintptr_t)p & 3) * 8;
    const unsigned int old = __hip_atomic_fetch_or(w, 3u << sh, __ATOMIC_RELAXED, WG_SCOPE);
    return ((old >> sh) & 3u) == 0u;
}

// Number of leading slots of a vertex visible to an item with sequence q:
// slots below deg0 are always visible; appended slots carry increasing
// sequence numbers, so the visible ones form a prefix.
__device__ __forceinline__ int visible_len(const SbcParams& P, const SbcVInfo& vi, int32_t q)
{
    if (q == INF32 || vi.deg == vi.deg0) return vi.deg;
    int len = vi.deg0;
    for (int e0 = vi.deg0; e0 < vi.deg; e0 += 64) {
        int e = e0 + lane_id();
        bool ok = e < vi.deg && P.seq[vi.start + e] <= q;
        int n = __popcll(ballot(ok));
        len += n;
        if (n < 64) break;
    }
    return len;
}

// Per-thread form of visible_len (no wavefront collectives).
__device__ __forceinline__ int visible_len_thread(const SbcParams& P, const SbcVInfo& vi, int32_t q)
{
    if (q == INF32 || vi.deg == vi.deg0) return vi.deg;
    int len = vi.deg0;
    while (len < vi.deg && P.seq[vi.start + len] <= q) ++len;
    return len;
}

// Exclusive prefix sum over a workgroup of NW wavefronts; `total` receives
// the sum.  tmp: NW + 1 ints of LDS.
template <int NW>
__device__ __forceinline__ int block_exscan(int x, int* tmp, int& total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int v = x;
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(v, d);
        if (lane >= d) v += y;
    }
    if (lane == 63) tmp[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int w = 0; w < NW; ++w) {
            const int t = tmp[w];
            tmp[w] = acc;
            acc += t;
        }
        tmp[NW] = acc;
    }
    __syncthreads();
    const int res = tmp[wave] + v - x;
    total = tmp[NW];
    __syncthreads();
    return res;
}

// Exclusive prefix sums of two values over a workgroup of NW wavefronts in
// one pass (two barriers).  tmp: 2 * (NW + 1) ints of LDS, not read again
// before the caller's next barrier.
template <int NW>
__device__ __forceinline__ int2 block_exscan2(int x, int y, int* tmp, int& total_x, int& total_y)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int vx = x, vy = y;
    for (int d = 1; d < 64; d <<= 1) {
        const int ax = __shfl_up(vx, d), ay = __shfl_up(vy, d);
        if (lane >= d) { vx += ax; vy += ay; }
    }
    if (lane == 63) { tmp[wave] = vx; tmp[NW + 1 + wave] = vy; }
    __syncthreads();
    if (threadIdx.x == 0) {
        int acx = 0, acy = 0;
        for (int w = 0; w < NW; ++w) {
            const int tx = tmp[w], ty = tmp[NW + 1 + w];
            tmp[w] = acx;
            tmp[NW + 1 + w] = acy;
            acx += tx;
            acy += ty;
        }
        tmp[NW] = acx;
        tmp[2 * NW + 1] = acy;
    }
    __syncthreads();
    total_x = tmp[NW];
    total_y = tmp[2 * NW + 1];
    return make_int2(tmp[wave] + vx - x, tmp[NW + 1 + wave] + vy - y);
}

// Three exclusive prefix sums in one pass (two barriers); tmp: 3 * (NW + 1)
// ints of LDS, not read again before the caller's next barrier.
template <int NW>
__device__ __forceinline__ int3 block_exscan3(int x, int y, int z, int* tmp, int& tx, int& ty, int& tz)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int vx = x, vy = y, vz = z;
    for (int d = 1; d < 64; d <<= 1) {
        const int ax = __shfl_up(vx, d), ay = __shfl_up(vy, d), az = __shfl_up(vz, d);
        if (lane >= d) { vx += ax; vy += ay; vz += az; }
    }
    constexpr int S = NW + 1;
    if (lane == 63) { tmp[wave] = vx; tmp[S + wave] = vy; tmp[2 * S + wave] = vz; }
    __syncthreads();
    if (threadIdx.x < 3) {
        int* t = tmp + threadIdx.x * S;
        int acc = 0;
        for (int w = 0; w < NW; ++w) {
            const int v = t[w];
            t[w] = acc;
            acc += v;
        }
        t[NW] = acc;
    }
    __syncthreads();
    tx = tmp[NW];
    ty = tmp[S + NW];
    tz = tmp[2 * S + NW];
    return make_int3(tmp[wave] + vx - x, tmp[S + wave] + vy - y, tmp[2 * S + wave] + vz - z);
}

// (sa / sb) * (d + 1) in FP64 -- the dependency term of
// src/streamingbc_init.c:159-161 and src/streamingbc_internal.c:185-198
__device__ __forceinline__ double dep(int64_t sa, int64_t sb, double d)
{
    return ((double)sa / (double)sb) * (d + 1.0);
}

// Per-team LDS block.
struct TeamShared {
    int cnt[4];                 // 0: |D|, 1: |S|, 2: end of U list, 3: pop result
    unsigned long long stat[4]; // processed, scanned, moved, adjacency entries scanned by the descent
    unsigned long long tr[12];  // trace counters of the item (LDS atomics), copied to the trace at its end
};

// Trace counter i of the item (LDS atomic); wide kernels keep no trace.
template <typename TS>
__device__ __forceinline__ void tr_add(TS& ts, int i, unsigned long long v)
{
    if constexpr (std::is_same<TS, TeamShared>::value) atomicAdd(&ts.tr[i], v);
}

// ------------------------------------------------------------ update item

// Vertices with more visible neighbours than this are scanned by every
// wavefront of a multi-wave team (contiguous chunk ranges per wave).
#ifndef SBC_HUB_MIN
#define SBC_HUB_MIN 512
#endif
constexpr int HUB_MIN = SBC_HUB_MIN;
// 64-neighbour chunks whose loads are issued together by one wavefront.
#ifndef SBC_SCAN_UNROLL
#define SBC_SCAN_UNROLL 2
#endif
constexpr int SCAN_UNROLL = SBC_SCAN_UNROLL;
// chunks whose loads one wavefront issues together when it ascends a single
// vertex on its own
#ifndef SBC_WAVE_UNROLL
#define SBC_WAVE_UNROLL 2
#endif
constexpr int WAVE_UNROLL = SBC_WAVE_UNROLL;
// lowest ascent level whose hubs may be resolved by table lookups (level 0:
// the root itself, whose sources are the processed level-1 vertices)
#ifndef SBC_LOOKUP_MIN_LEV
#define SBC_LOOKUP_MIN_LEV 0
#endif
// (source, hub) lookups whose first table slots one thread loads together

constexpr int HUB_CAP = 512;   // hubs queued per level pass (more: scanned by one wave)
constexpr int RUN_CAP = 64;    // ascent terms per wave and hub kept in LDS

constexpr int PHASH = 2 * HUB_CAP;      // LDS hash: vertex -> number in the level's lookup list
constexpr int SBITS_LOG = 14;           // bits of the membership filter beside the hash
constexpr int SBITS_WORDS = (1 << SBITS_LOG) / 32;
constexpr int TWOHOP_CAP = 1024;
constexpr int PREF_CAP = 128;           // window items a team classifies in one pass        // two-hop set of the root kept for level-3 lookups
// Lookup buckets: a pass's SBC_BUCKET_SLOTS term slots are shared out evenly
// between its lookup hubs, at most TEAM_SORT_CAP per hub; a bucket of up to
// WAVE_SORT_CAP terms is sorted by one wavefront, a larger one by the whole
// team.
constexpr int TEAM_SORT_CAP = 1024;
constexpr int WAVE_SORT_CAP = 256;
static_assert(TEAM_SORT_CAP % 64 == 0 && TEAM_SORT_CAP <= SBC_BUCKET_SLOTS, "bucket shares");
static_assert(SBC_BUCKET_SLOTS / HUB_CAP >= WAVE_SORT_CAP, "every pass can hold wave-sorted buckets");

struct alignas(16) HubShared {
    int n;
    int overflow;
    int pn;
    int np;
    int ns;               // hubs staged for the team scan
    long long svol;       // adjacency entries of the pass's hubs that lookups cannot serve
    int handover;         // the item goes to the wide kernels at the end of this level's hub passes
    int bstride;          // term slots per lookup hub in this pass's buckets
    int n3;               // lookup hubs whose buckets the whole team sorts
    int dup;              // a team-sorted bucket holds a position twice
    int plen[HUB_CAP];    // visible list length of each lookup hub
    int pstart[HUB_CAP];  // first slot of each lookup hub
    int poff[HUB_CAP];    // each lookup hub's table: first slot in htab
    uint8_t plog[HUB_CAP];  // and log2 of its size
    int hlen[HUB_CAP];    // visible list length of each queued hub of the pass
    int hstart[HUB_CAP];  // first slot of each queued hub of the pass
    int hidx[HUB_CAP];    // indexed neighbours of each hub queued in LDS by an ascent level (hub table count)
    uint8_t hasc[HUB_CAP];  // that hub is ascent-only
    int list[HUB_CAP];
    int plist[HUB_CAP];   // ascent-only hubs of the level handled by table lookups
    int bcnt[HUB_CAP];    // terms collected per lookup hub
    int phk[PHASH];       // hash keys (vertex + 1, 0 = empty)
    int phv[PHASH];       // hash values (number in plist)
    int pidx[HUB_CAP];    // queued hub -> number in plist, or -1
    uint32_t sbits[SBITS_WORDS];  // membership filter of phk's keys (sbit_set)
    uint8_t kind[HUB_CAP];
    int ntwo;                 // entries of two (-1: the root's two-hop set is too large)
    int two_ok;               // two/ntwo hold the current root's set (cleared when a root is popped and
                              // when an item's edge ends at the root or at one of its neighbours)
    int32_t two[TWOHOP_CAP];  // the root's two-hop set: parent candidates of level-3 hubs
    // flat scan of the staged hubs: segment i = j + w holds the terms of hub
    // j found by wavefront w, entries [seg_b, seg_e) of w's run
    int seg_j[HUB_CAP + 16];
    int seg_b[HUB_CAP + 16];
    int seg_e[HUB_CAP + 16];
    int wcnt[16];         // terms per wavefront run (above capacity: overflow)
    uint8_t sovf[HUB_CAP];
    int runlen[16];
    // a team scan's terms per wavefront; during a lookup pass (which uses no
    // team scan) entry hi of the flattened rows holds lookup hub hi's old
    // dependency (radd) and sigma (rsub, int64 bits)
    double radd[16][RUN_CAP];
    double rsub[16][RUN_CAP];
    uint8_t rflag[16][RUN_CAP];
    // window items [pf0, pf1) of the team's root, fetched together: their
    // endpoints and two bits each, gap 0 (nothing to do) and an endpoint on
    // level 0 or 1 (the root's two-hop set changes); valid until an item
    // moves vertices (only moves change levels)
    int pf0, pf1;
    int32_t pu[PREF_CAP];
    int32_t pv[PREF_CAP];
    int32_t pdu[PREF_CAP];   // their levels
    int32_t pdv[PREF_CAP];
    uint32_t pzero[PREF_CAP / 32];
    uint32_t plow[PREF_CAP / 32];
};

__device__ __forceinline__ int phash_slot(int32_t v) { return (int)(((uint32_t)v * 0x9E3779B1u) >> 21) & (PHASH - 1); }
// A bit per value of a second hash of the vertex id, set for every member
// of phk: a clear bit rules a vertex out with one LDS load, without the
// probe sequence (one hash measured faster than two or three).
__device__ __forceinline__ uint32_t sbit_of(int32_t v) { return ((uint32_t)v * 0x85EBCA6Bu) >> (32 - SBITS_LOG); }
__device__ __forceinline__ void sbit_set(HubShared& hs, int32_t v)
{
    const uint32_t b = sbit_of(v);
    atomicOr(&hs.sbits[b >> 5], 1u << (b & 31));
}
__device__ __forceinline__ bool sbit_test(const HubShared& hs, int32_t v)
{
    const uint32_t b = sbit_of(v);
    return (hs.sbits[b >> 5] >> (b & 31)) & 1u;
}

// Membership set of an ascent level's possible sources, in hs.phk: every
// vertex that can give a light vertex of the level a term (the processed
// vertices one level down and the descent vertices on the level) or that the
// level may mark as a parent (every vertex one level up).  A neighbour
// outside it contributes nothing, so its state byte is not gathered.
constexpr int SHASH_MAX = PHASH / 2;

__device__ __forceinline__ void shash_insert(HubShared& hs, int32_t v)
{
    sbit_set(hs, v);
    int h = phash_slot(v);
    for (;;) {
        const int old = atomicCAS(&hs.phk[h], 0, v + 1);
        if (old == 0 || old == v + 1) return;
        h = (h + 1) & (PHASH - 1);
    }
}

__device__ __forceinline__ bool shash_has(const HubShared& hs, int32_t v)
{
    if (!sbit_test(hs, v)) return false;
    int h = phash_slot(v);
    for (;;) {
        const int key = hs.phk[h];
        if (key == v + 1) return true;
        if (key == 0) return false;
        h = (h + 1) & (PHASH - 1);
    }
}

// Hubs per hub pass: the LDS queue's capacity (smaller only to test the
// batched passes).
__device__ __forceinline__ int hub_batch(const SbcParams& P)
{
    return (P.hub_batch > 0 && P.hub_batch < HUB_CAP) ? P.hub_batch : HUB_CAP;
}

// Light (non-hub) vertices of one ascent level, batched: every edge of
// every batched vertex is handled by one lane, its term kept at the edge's
// index, then each vertex sums its edges' terms in list order.
constexpr int LIGHT_CAP = 1024;  // vertices per batch (one per thread of a 1024-thread team)
#ifndef SBC_EDGE_CAP
#define SBC_EDGE_CAP 2048
#endif
constexpr int EDGE_CAP = SBC_EDGE_CAP;  // edges per batch

struct alignas(16) LightShared {
    int nb;                       // batched vertices
    int novf;                     // hubs the queues could not take (one wavefront each)
    int scan[52];
    int32_t k[LIGHT_CAP];
    int32_t start[LIGHT_CAP];
    int32_t off[LIGHT_CAP];       // first edge index of the vertex in the batch
    int32_t len[LIGHT_CAP];
    uint8_t isD[LIGHT_CAP];
    int64_t sk_old[LIGHT_CAP];
    int64_t sk_new[LIGHT_CAP];
    double base[LIGHT_CAP];
    int32_t ovf[LIGHT_CAP];       // light vertices of a round past the edge cap keep their slots
    uint8_t eflag[EDGE_CAP];
    double eadd[EDGE_CAP];
    double esub[EDGE_CAP];
};
static_assert(TEAM_SORT_CAP <= LIGHT_CAP && TEAM_SORT_CAP <= EDGE_CAP, "team sort stages a bucket in LightShared");
static_assert(2 * TEAM_SORT_CAP <= 16 * 128, "a team-sorted bucket's values fit the packing rows");
// positions of a team-sorted hub's list: a bitmap in LightShared::base, per-word prefixes in ::ovf
constexpr int TEAM_SORT_BITS = 32 * LIGHT_CAP;
static_assert(TEAM_SORT_BITS <= 64 * LIGHT_CAP, "bitmap fits LightShared::base");

// Per-item constants.
struct Item {
    int32_t* dist;    // levels (unchanged until commit)
    uint8_t* sb;      // state bytes
    SbcVRec* vr;      // the root's path counts, dependencies and scratch
    int32_t* list;
    int nv;
    int32_t q;        // visibility sequence
    int32_t s, p, r;  // deeper endpoint, shallower endpoint, root
};

// One neighbour k (state byte b) of descent vertex c (old level oc, new
// level lc, added paths dc, moved: cmoved): marks k as a child at level
// lc + 1 and adds c's added paths to it, or records it as an unmoved
// same-level neighbour of a moved c.  appD / appS: this call marked k, which
// then joins the next frontier / the same-level list.  Per lane.
__device__ __forceinline__ void descent_edge(const Item& X, int32_t k, uint8_t b, int32_t oc, int32_t lc, int64_t dc,
                                             bool cmoved, bool& appD, bool& appS)
{
    const uint8_t wantStay = sb_make(lc + 1, 1), wantMove = sb_make(lc + 1, 2);
    appD = false;
    appS = false;
    if (sb_type(b) == 0) {
        int64_t cd;
        if (b & SB_INF) cd = 1;  // unreachable until now: moves up
        else cd = (int64_t)oc + sb_rel(b, oc) - (int64_t)(lc + 1);
        if (cd >= 0) {
            const uint8_t want = cd > 0 ? wantMove : wantStay;
            const uint8_t seen = cas_sb(&X.sb[k], b, want);
            appD = (seen == b);
            if (seen == b || seen == want)
                add64_wg(&X.vr[k].dpath, dc);
        } else if (cd == -1 && cmoved) {
            appS = (cas_sb(&X.sb[k], b, sb_make(lc, 3)) == b);
        }
    } else if (b == wantStay || b == wantMove) {
        add64_wg(&X.vr[k].dpath, dc);
    }
}

// Descent over chunks [cb, ce) of vertex c (old level oc, new level lc,
// added paths dc).  Marks children, accumulates their added paths, appends
// them to the next frontier and records unmoved same-level neighbours of a
// moved c.
// TS: the owner of the list counters cnt[0..1] (a team's LDS copy, or a
// parked item's global one).
template <typename TS, int SCAN_UNROLL = ::SCAN_UNROLL>
__device__ __forceinline__ void descent_chunks(const SbcParams& P, const Item& X, int32_t c, const SbcVInfo& vi,
                                               int len, int cb, int ce, int32_t oc, int32_t lc, int64_t dc,
                                               bool cmoved, TS& ts)
{
    const int lane = lane_id();
    const bool c_is_s = (c == X.s);
    for (int ch = cb; ch < ce; ch += SCAN_UNROLL) {
        int32_t kk[SCAN_UNROLL];
        uint8_t bb[SCAN_UNROLL];
        bool ok[SCAN_UNROLL];
#pragma unroll
        for (int u = 0; u < SCAN_UNROLL; ++u) {
            const int e = (ch + u) * 64 + lane;
            ok[u] = (ch + u) < ce && e < len;
            kk[u] = ok[u] ? P.col[vi.start + e] : 0;
            // the inserted edge is the one edge whose ends may be more than
            // one level apart; it never adds a child of s
            if (ok[u] && c_is_s && kk[u] == X.p) ok[u] = false;
        }
#pragma unroll
        for (int u = 0; u < SCAN_UNROLL; ++u) bb[u] = ok[u] ? ld_sb(&X.sb[kk[u]]) : (uint8_t)0;
#pragma unroll
        for (int u = 0; u < SCAN_UNROLL; ++u) {
            bool appD = false, appS = false;
            const int32_t k = kk[u];
            if (ok[u]) descent_edge(X, k, bb[u], oc, lc, dc, cmoved, appD, appS);
            const int pd = wave_append(appD, &ts.cnt[0]);
            if (pd >= 0) X.list[pd] = k;
            const int psl = wave_append(appS, &ts.cnt[1]);
            if (psl >= 0) X.list[X.nv - 1 - psl] = k;
        }
    }
}

// Descent over the hubs queued while expanding one frontier (new level lc of
// their children), as one flat range of 64-neighbour chunks split evenly
// between the wavefronts: per-hub state is read once, in parallel, into LDS
// (the light-vertex arrays are free during the descent), and a wavefront's
// range may span several hubs.  Marks and added paths are exact in any order.
template <int TW>
__device__ __forceinline__ void descent_hubs_flat(const SbcParams& P, const Item& X, int nh, int32_t lc,
                                                  TeamShared& ts, HubShared& hs, LightShared& ls, int twave)
{
    constexpr int NT = TW * 64;
    const int tid = threadIdx.x;
    // thread t stages hubs [t*per, (t+1)*per)
    const int per = (nh + NT - 1) / NT;
    const int h0 = tid * per, h1 = min(h0 + per, nh);
    int mine = 0;
    for (int h = h0; h < h1; ++h) {
        const int32_t c = hs.list[h];
        const SbcVInfo vi = P.vinfo[c];
        const int len = visible_len_thread(P, vi, X.q);
        const bool cm = sb_type(ld_sb(&X.sb[c])) == 2;
        ls.k[h] = c;
        ls.start[h] = vi.start;
        ls.len[h] = len;
        ls.isD[h] = cm ? 1 : 0;
        ls.sk_old[h] = ld_acq64(&X.vr[c].dpath);  // added paths of c
        ls.sk_new[h] = X.dist[c];              // old level of c
        mine += (len + 63) >> 6;
        atomicAdd(&ts.stat[0], 1ull);
        atomicAdd(&ts.stat[1], (unsigned long long)len);
        atomicAdd(&ts.stat[3], (unsigned long long)len);
        if (cm) atomicAdd(&ts.stat[2], 1ull);
    }
    int total;
    int off = block_exscan<TW>(mine, ls.scan, total);
    for (int h = h0; h < h1; ++h) {
        ls.off[h] = off;
        off += (ls.len[h] + 63) >> 6;
    }
    __syncthreads();
    const int cb = (int)(((long)twave * total) / TW), ce = (int)(((long)(twave + 1) * total) / TW);
    if (cb < ce) {
        int lo = 0, hi = nh;  // last hub starting at or before cb
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (ls.off[mid] <= cb) lo = mid;
            else hi = mid;
        }
        for (int h = lo; h < nh; ++h) {
            const int hb = ls.off[h];
            if (hb >= ce) break;
            const int hn = (ls.len[h] + 63) >> 6;
            const int a = cb > hb ? cb - hb : 0;
            const int b = ce - hb < hn ? ce - hb : hn;
            if (a < b) {
                SbcVInfo vi{ls.start[h], 0, 0, 0};
                descent_chunks(P, X, ls.k[h], vi, ls.len[h], a, b, (int32_t)ls.sk_new[h], lc, ls.sk_old[h],
                               ls.isD[h] != 0, ts);
            }
        }
    }
}

// Ascent terms of one 64-neighbour chunk of vertex k at level lev: marks
// unmarked parents (level lev-1) as ascent vertices and returns this lane's
// term: child (add = new dependency of a processed child) and/or hsub (sub =
// old dependency to remove).
struct Term {
    bool child, hsub;
    double add, sub;
};

__device__ __forceinline__ Term ascent_lane(const SbcParams& P, const Item& X, int32_t k, bool isD,
                                            int64_t sk_old, int64_t sk_new, int lev, bool ok, int32_t c,
                                            uint8_t bc, bool& appU)
{
    Term t{false, false, 0.0, 0.0};
    appU = false;
    if (!ok || (bc & SB_INF)) return t;
    const int ctype = sb_type(bc);
    const int rel = sb_rel(bc, lev);
    if (ctype != 0) {
        if (rel == 1) {
            t.child = true;
            const int64_t sc_old = X.vr[c].sigma;
            const int64_t sc_new = ctype == 3 ? sc_old
                : (int64_t)((uint64_t)ld_acq64(&X.vr[c].dpath) + (uint64_t)(ctype == 1 ? sc_old : 0));
            t.add = dep(sk_new, sc_new, ld_accd(&X.vr[c].ndel));
            if (!isD && !(k == X.p && c == X.s)) {
                t.hsub = true;
                t.sub = dep(sk_old, sc_old, X.vr[c].delta);
            }
        } else if (!isD && rel == 0 && ctype == 2) {
            t.hsub = true;
            t.sub = dep(sk_old, X.vr[c].sigma, X.vr[c].delta);
        }
    } else if (lev >= 1 && rel == -1) {
        appU = mark_sb3(&X.sb[c]);
    }
    return t;
}

struct KState {
    bool isD;
    int64_t sk_old, sk_new;
    double base;
};

// Old dependency of ascent-only vertex k, where its pull starts: in an edge
// deletion the shallower endpoint p no longer meets s in its list, so it
// starts without s's old term.  The stored value stays as it is: p's own
// parents subtract p's old term from it.
__device__ __forceinline__ double asc_base(const SbcParams& P, const Item& X, int32_t k)
{
    const double d = X.vr[k].delta;
    if (P.del && k == X.p) return d - dep(X.vr[k].sigma, X.vr[X.s].sigma, X.vr[X.s].delta);
    return d;
}

__device__ __forceinline__ KState kstate(const SbcParams& P, const Item& X, int32_t k)
{
    // the four loads are issued together; the state byte selects which are used
    KState s;
    const uint8_t b = ld_sb(&X.sb[k]);
    const int64_t sg = X.vr[k].sigma;
    const int64_t dp = ld_acq64(&X.vr[k].dpath);
    const double dl = X.vr[k].delta;
    const int kt = sb_type(b);
    s.isD = kt != 3;
    s.sk_old = sg;
    s.sk_new = s.isD ? (int64_t)((uint64_t)dp + (uint64_t)(kt == 1 ? sg : 0)) : sg;
    s.base = s.isD ? 0.0 : ((P.del && k == X.p) ? asc_base(P, X, k) : dl);
    return s;
}

// One wavefront computes the new delta of k, summing the terms in adjacency
// order lane by lane.
template <typename TS, int WAVE_UNROLL = ::WAVE_UNROLL>
__device__ __forceinline__ void ascent_vertex_wave_ks(const SbcParams& P, const Item& X, int32_t k, int32_t start,
                                                      int len, const KState& ks, int lev, TS& ts,
                                                      const HubShared* srcset = nullptr, int64_t* tp = nullptr)
{
    const int lane = lane_id();
    double acc = ks.base;
    const int nch = (len + 63) >> 6;
    // trace: vertices, chunks, chunks with members, members, time (source-set
    // form); vertices, chunks, time (scan form)
    const int64_t tv0 = (tp && lane == 0) ? (int64_t)wall_clock64() : 0;
    if (srcset) {
        if (tp && lane == 0) {
            tr_add(ts, 4, 1ull);
            tr_add(ts, 5, (unsigned long long)nch);
        }
        // neighbour ids of up to 8 chunks in flight, then chunk by chunk:
        // only members of the level's source set can give a term or need a
        // mark, and a chunk without one is skipped
        constexpr int SU = 8;
        for (int ch0 = 0; ch0 < nch; ch0 += SU) {
            int32_t cc[SU];
#pragma unroll
            for (int u = 0; u < SU; ++u) {
                const int e = (ch0 + u) * 64 + lane;
                cc[u] = (ch0 + u) < nch && e < len ? P.col[start + e] : -1;
            }
#pragma unroll
            for (int u = 0; u < SU; ++u) {
                const bool ok = cc[u] >= 0 && shash_has(*srcset, cc[u]);
                const uint64_t bm = ballot(ok);
                if (tp && lane == 0 && bm != 0) {
                    tr_add(ts, 6, 1ull);
                    tr_add(ts, 7, (unsigned long long)__popcll(bm));
                }
                if (bm == 0) continue;
                const uint8_t ww = ok ? ld_sb(&X.sb[cc[u]]) : (uint8_t)0;
                bool appU;
                const Term t = ascent_lane(P, X, k, ks.isD, ks.sk_old, ks.sk_new, lev, ok, cc[u], ww, appU);
                const int pu = wave_append(appU, &ts.cnt[2]);
                if (pu >= 0) X.list[pu] = cc[u];
                acc = wave_ordered_add(acc, t.child, t.add, t.hsub, t.sub);
            }
        }
        if (lane == 0) st_reld(&X.vr[k].ndel, acc);
        if (tp && lane == 0) tr_add(ts, 8, (unsigned long long)((int64_t)wall_clock64() - tv0));
        return;
    }
    if (tp && lane == 0) {
        tr_add(ts, 9, 1ull);
        tr_add(ts, 10, (unsigned long long)nch);
    }
    for (int ch = 0; ch < nch; ch += WAVE_UNROLL) {
        int32_t cc[WAVE_UNROLL];
        bool ok[WAVE_UNROLL];
#pragma unroll
        for (int u = 0; u < WAVE_UNROLL; ++u) {
            const int e = (ch + u) * 64 + lane;
            ok[u] = (ch + u) < nch && e < len;
            cc[u] = ok[u] ? P.col[start + e] : 0;
        }
        if (srcset) {
#pragma unroll
            for (int u = 0; u < WAVE_UNROLL; ++u) ok[u] = ok[u] && shash_has(*srcset, cc[u]);
        }
        uint8_t ww[WAVE_UNROLL];
#pragma unroll
        for (int u = 0; u < WAVE_UNROLL; ++u) ww[u] = ok[u] ? ld_sb(&X.sb[cc[u]]) : (uint8_t)0;
        // every chunk's term loads are issued before the first sum waits on them
        Term t[WAVE_UNROLL];
        bool appU[WAVE_UNROLL];
#pragma unroll
        for (int u = 0; u < WAVE_UNROLL; ++u)
            t[u] = ascent_lane(P, X, k, ks.isD, ks.sk_old, ks.sk_new, lev, ok[u], cc[u], ww[u], appU[u]);
#pragma unroll
        for (int u = 0; u < WAVE_UNROLL; ++u) {
            const int pu = wave_append(appU[u], &ts.cnt[2]);
            if (pu >= 0) X.list[pu] = cc[u];
        }
#pragma unroll
        for (int u = 0; u < WAVE_UNROLL; ++u) acc = wave_ordered_add(acc, t[u].child, t[u].add, t[u].hsub, t[u].sub);
    }
    if (lane == 0) st_reld(&X.vr[k].ndel, acc);
    if (tp && lane == 0) tr_add(ts, 11, (unsigned long long)((int64_t)wall_clock64() - tv0));
}

template <typename TS, int WAVE_UNROLL = ::WAVE_UNROLL>
__device__ __forceinline__ void ascent_vertex_wave(const SbcParams& P, const Item& X, int32_t k, const SbcVInfo& vi, int len,
                                   int lev, TS& ts)
{
    ascent_vertex_wave_ks<TS, WAVE_UNROLL>(P, X, k, vi.start, len, kstate(P, X, k), lev, ts);
}

// The whole team computes the new delta of hub k: wave w scans a contiguous
// range of chunks and keeps its terms in order in LDS; the runs are then
// summed in wave order, which is adjacency order.  A run that outgrows
// RUN_CAP sends the vertex back to the single-wave path.
template <int TW>
__device__ __forceinline__ void ascent_vertex_team(const SbcParams& P, const Item& X, int32_t k, const SbcVInfo& vi, int len,
                                   const KState& ks, int lev, TeamShared& ts, HubShared& hs, int twave)
{
    const int lane = lane_id();
    const int nch = (len + 63) >> 6;
    const int cb = (int)(((long)twave * nch) / TW), ce = (int)(((long)(twave + 1) * nch) / TW);
    int rl = 0;
    for (int ch = cb; ch < ce; ch += SCAN_UNROLL) {
        int32_t cc[SCAN_UNROLL];
        uint8_t ww[SCAN_UNROLL];
        bool ok[SCAN_UNROLL];
#pragma unroll
        for (int u = 0; u < SCAN_UNROLL; ++u) {
            const int e = (ch + u) * 64 + lane;
            ok[u] = (ch + u) < ce && e < len;
            cc[u] = ok[u] ? P.col[vi.start + e] : 0;
        }
#pragma unroll
        for (int u = 0; u < SCAN_UNROLL; ++u) ww[u] = ok[u] ? ld_sb(&X.sb[cc[u]]) : (uint8_t)0;
#pragma unroll
        for (int u = 0; u < SCAN_UNROLL; ++u) {
            bool appU;
            const Term t = ascent_lane(P, X, k, ks.isD, ks.sk_old, ks.sk_new, lev, ok[u], cc[u], ww[u], appU);
            const int pu = wave_append(appU, &ts.cnt[2]);
            if (pu >= 0) X.list[pu] = cc[u];
            const uint64_t m = ballot(t.child || t.hsub);
            const int n = __popcll(m);
            if (rl + n <= RUN_CAP && (t.child || t.hsub)) {
                const int idx = rl + prefix_in_wave(m);
                hs.radd[twave][idx] = t.add;
                hs.rsub[twave][idx] = t.sub;
                hs.rflag[twave][idx] = (uint8_t)((t.child ? 1 : 0) | (t.hsub ? 2 : 0));
            }
            rl += n;
        }
    }
    if (lane == 0) {
        hs.runlen[twave] = rl;
        if (rl > RUN_CAP) atomicOr(&hs.overflow, 1);
    }
    __syncthreads();
    if (twave == 0) {
        if (hs.overflow) {
            ascent_vertex_wave(P, X, k, vi, len, lev, ts);
        } else if (lane == 0) {
            double acc = ks.base;
            for (int w = 0; w < TW; ++w)
                for (int i = 0; i < hs.runlen[w]; ++i) {
                    const uint8_t fl = hs.rflag[w][i];
                    if (fl & 1) acc = acc + hs.radd[w][i];
                    if (fl & 2) acc = acc - hs.rsub[w][i];
                }
            st_reld(&X.vr[k].ndel, acc);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) hs.overflow = 0;
    __syncthreads();
}


// Scan of chunks [a, b) of hub k (state ks): marks parents like the other
// scans and appends the terms, in adjacency order, to the calling
// wavefront's run (entries rl.. ; those past cap are dropped but counted).
#ifndef SBC_FLAT_UNROLL
#define SBC_FLAT_UNROLL 8
#endif
constexpr int FLAT_UNROLL = SBC_FLAT_UNROLL;  // chunks whose loads one wavefront issues together in hub scans

__device__ __forceinline__ int ascent_scan_run(const SbcParams& P, const Item& X, int32_t k, int32_t start, int len,
                                               const KState& ks, int lev, int a, int b, TeamShared& ts,
                                               SbcBucketEnt* run, int rl, int cap)
{
    constexpr int SCAN_UNROLL = FLAT_UNROLL;
    const int lane = lane_id();
    for (int ch = a; ch < b; ch += SCAN_UNROLL) {
        int32_t cc[SCAN_UNROLL];
        uint8_t ww[SCAN_UNROLL];
        bool ok[SCAN_UNROLL];
#pragma unroll
        for (int u = 0; u < SCAN_UNROLL; ++u) {
            const int e = (ch + u) * 64 + lane;
            ok[u] = (ch + u) < b && e < len;
            cc[u] = ok[u] ? P.col[start + e] : 0;
        }
#pragma unroll
        for (int u = 0; u < SCAN_UNROLL; ++u) ww[u] = ok[u] ? ld_sb(&X.sb[cc[u]]) : (uint8_t)0;
#pragma unroll
        for (int u = 0; u < SCAN_UNROLL; ++u) {
            bool appU;
            const Term t = ascent_lane(P, X, k, ks.isD, ks.sk_old, ks.sk_new, lev, ok[u], cc[u], ww[u], appU);
            const int pu = wave_append(appU, &ts.cnt[2]);
            if (pu >= 0) X.list[pu] = cc[u];
            const bool has = t.child || t.hsub;
            const uint64_t m = ballot(has);
            if (has) {
                const int idx = rl + prefix_in_wave(m);
                if (idx < cap) {
                    SbcBucketEnt& e = run[idx];
                    e.key = 0;
                    e.flag = (t.child ? 1 : 0) | (t.hsub ? 2 : 0);
                    e.add = t.add;
                    e.sub = t.sub;
                }
            }
            rl += __popcll(m);
        }
    }
    return rl;
}

// New deltas of the hubs staged in ls[0..nsh) (vertex, list start, visible
// length, state), all scanned together: their chunks form one flat range
// split evenly between the wavefronts, so a wavefront's range may span
// several hubs and a hub several wavefronts.  Each wavefront appends its
// terms in scan order to its own run in the workgroup's scratch; a hub's
// terms are then its segments of consecutive wavefronts' runs, which is
// adjacency order, and one wavefront per hub sums them in that order (the
// order of the single-vertex scans, so the result is bit-identical).  Hubs
// touched by a wavefront whose run overflowed are rescanned by the whole
// team (parent marks are idempotent).
template <int TW>
__device__ __forceinline__ void ascent_hubs_flat(const SbcParams& P, const Item& X, int nsh, int lev, TeamShared& ts,
                                                 HubShared& hs, LightShared& ls, int twave)
{
    constexpr int NT = TW * 64;
    constexpr int share = SBC_BUCKET_SLOTS / TW;  // run stride
    const int cap = (P.run_cap > 0 && P.run_cap < share) ? P.run_cap : share;
    const int tid = threadIdx.x, lane = lane_id();
    SbcBucketEnt* runs = P.bucket + (size_t)blockIdx.x * SBC_BUCKET_SLOTS;
    const int per = (nsh + NT - 1) / NT;
    const int h0 = min(tid * per, nsh), h1 = min(h0 + per, nsh);
    int mine = 0;
    for (int h = h0; h < h1; ++h) mine += (ls.len[h] + 63) >> 6;
    int total;
    int off = block_exscan<TW>(mine, ls.scan, total);
    for (int h = h0; h < h1; ++h) {
        ls.off[h] = off;
        off += (ls.len[h] + 63) >> 6;
    }
    for (int i = tid; i < nsh + TW; i += NT) hs.seg_j[i] = -1;
    for (int h = tid; h < nsh; h += NT) hs.sovf[h] = 0;
    __syncthreads();
    const int cb = (int)(((long)twave * total) / TW), ce = (int)(((long)(twave + 1) * total) / TW);
    int rl = 0;
    if (cb < ce) {
        int lo = 0, hi = nsh;  // last hub starting at or before cb
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (ls.off[mid] <= cb) lo = mid;
            else hi = mid;
        }
        for (int h = lo; h < nsh; ++h) {
            const int hb = ls.off[h];
            if (hb >= ce) break;
            const int hn = (ls.len[h] + 63) >> 6;
            const int a = cb > hb ? cb - hb : 0;
            const int b = ce - hb < hn ? ce - hb : hn;
            if (a < b) {
                const KState kst{ls.isD[h] != 0, ls.sk_old[h], ls.sk_new[h], ls.base[h]};
                const int r0 = rl;
                rl = ascent_scan_run(P, X, ls.k[h], ls.start[h], ls.len[h], kst, lev, a, b, ts,
                                     runs + (size_t)twave * share, rl, cap);
                if (lane == 0) {
                    hs.seg_j[h + twave] = h;
                    hs.seg_b[h + twave] = r0;
                    hs.seg_e[h + twave] = rl;
                }
            }
        }
    }
    if (lane == 0) hs.wcnt[twave] = rl;
    __syncthreads();
    for (int i = tid; i < nsh + TW - 1; i += NT) {
        const int j = hs.seg_j[i];
        if (j >= 0 && hs.wcnt[i - j] > cap) hs.sovf[j] = 1;
    }
    __syncthreads();
    for (int j = twave; j < nsh; j += TW) {
        if (hs.sovf[j]) continue;
        // lane w < TW: wavefront w's segment of hub j (start, length); an
        // inclusive prefix over them numbers the hub's terms in order
        int sbeg = 0, slen = 0;
        if (lane < TW && hs.seg_j[j + lane] == j) {
            sbeg = hs.seg_b[j + lane];
            slen = hs.seg_e[j + lane] - sbeg;
        }
        int pre = slen;
        for (int d = 1; d < TW; d <<= 1) {
            const int y = __shfl_up(pre, d);
            if (lane >= d) pre += y;
        }
        const int T = __shfl(pre, TW - 1);
        double acc = ls.base[j];
        for (int t0 = 0; t0 < T; t0 += 64) {
            const int t = t0 + lane;
            int w = 0;  // segment of term t: segments ending at or before t
            for (int q = 0; q < TW; ++q) w += t >= __shfl(pre, q) ? 1 : 0;
            const int wq = w < TW ? w : TW - 1;
            const int excl = __shfl(pre - slen, wq);
            const int beg = __shfl(sbeg, wq);
            int fl = 0;
            double ad = 0.0, sb2 = 0.0;
            if (t < T) {
                const SbcBucketEnt& e = runs[(size_t)wq * share + beg + (t - excl)];
                fl = e.flag;
                ad = e.add;
                sb2 = e.sub;
            }
            acc = wave_ordered_add(acc, (fl & 1) != 0, ad, (fl & 2) != 0, sb2);
        }
        if (lane == 0) st_reld(&X.vr[ls.k[j]].ndel, acc);
    }
    __syncthreads();
    for (int j = 0; j < nsh; ++j) {
        if (!hs.sovf[j]) continue;
        const KState kst{ls.isD[j] != 0, ls.sk_old[j], ls.sk_new[j], ls.base[j]};
        const SbcVInfo vi{ls.start[j], 0, 0, 0};
        ascent_vertex_team<TW>(P, X, ls.k[j], vi, ls.len[j], kst, lev, ts, hs, twave);
    }
}

// Position of w in the indexed (initial) part of a vertex's adjacency list,
// looked up in the vertex's table, or -1.  A table is 2^logT slots {key + 1,
// position}, at most a quarter used, in buckets of 8 slots (64 bytes, one
// load) that fill in order: w is in bucket hash(w), or in a later one when
// every bucket from hash(w) on is full, so a bucket with a free slot ends
// the search.
__device__ __forceinline__ uint32_t hub_bucket(int32_t w, int logT)
{
    return ((uint32_t)w * 0x9E3779B1u) >> (32 - (logT - 3));
}
// w + 1 among the bucket's 8 keys: its position, -1 (a free slot: absent)
// or -2 (full without it: the next bucket)
__device__ __forceinline__ int bucket_match(const int4& a, const int4& b, const int4& c, const int4& d, int32_t key)
{
    const int k[8] = {a.x, a.z, b.x, b.z, c.x, c.z, d.x, d.z};
    const int p[8] = {a.y, a.w, b.y, b.w, c.y, c.w, d.y, d.w};
    int r = -2;
#pragma unroll
    for (int i = 7; i >= 0; --i) {
        if (k[i] == 0) r = -1;
        if (k[i] == key) r = p[i];
    }
    return r;
}
__device__ __forceinline__ int hub_find_bucket(const int2* __restrict__ tab, int logT, int32_t w, uint32_t b)
{
    const uint32_t mask = (1u << (logT - 3)) - 1u;
    for (;;) {
        const int4* q = reinterpret_cast<const int4*>(tab + ((size_t)b << 3));
        const int r = bucket_match(q[0], q[1], q[2], q[3], w + 1);
        if (r != -2) return r;
        b = (b + 1) & mask;
    }
}
__device__ __forceinline__ int hub_find(const int2* __restrict__ tab, int logT, int32_t w)
{
    return hub_find_bucket(tab, logT, w, hub_bucket(w, logT));
}

// Processed-vertex ranges of one ascent level: descent frontier list[dA..dB),
// same-level list (stored backwards) sl[sA..sB), ascent list list[uA..uB).
struct LevelSets {
    int dA, dB, sA, sB, uA, uB;
    __device__ int size() const { return (dB - dA) + (sB - sA) + (uB - uA); }
    __device__ int32_t at(const Item& X, int i) const
    {
        const int nd = dB - dA, ns = sB - sA;
        if (i < nd) return X.list[dA + i];
        if (i < nd + ns) return X.list[X.nv - 1 - (sA + i - nd)];
        return X.list[uA + i - nd - ns];
    }
};

// Term of processed neighbour w of ascent-only hub k (number hi in the
// level's lookup list) at position pos of k's list, stored in k's bucket.
__device__ __forceinline__ void probe_term(const SbcParams& P, const Item& X, HubShared& hs, SbcBucketEnt* bk,
                                           int32_t k, int hi, int lev, int32_t w, int pos)
{
    bool appU;
    const int64_t sk = X.vr[k].sigma;
    const Term t = ascent_lane(P, X, k, false, sk, sk, lev, true, w, ld_sb(&X.sb[w]), appU);
    if (t.child || t.hsub) {
        const int slot = atomicAdd(&hs.bcnt[hi], 1);
        if (slot < hs.bstride) {
            SbcBucketEnt& e = bk[(size_t)hi * hs.bstride + slot];
            e.key = pos;
            e.flag = (t.child ? 1 : 0) | (t.hsub ? 2 : 0);
            e.add = t.add;
            e.sub = t.sub;
        }
    }
}

static_assert(16 * RUN_CAP >= HUB_CAP, "hub bases and sigmas fit the run rows");
// Source data of a lookup pass kept in LDS by the source round (slot = the
// thread that loaded it): ls.eflag = term kind (0 none, 1 child, 2 moved
// vertex on the level; +4: the source is s), ls.sk_old = old sigma, ls.sk_new
// = new sigma, ls.base = old delta, ls.eadd = new delta.  The hub's sigma is
// hs.rsub[hi] (int64 bits) and its old dependency hs.radd[hi].
__device__ __forceinline__ int64_t probe_hub_sigma(const HubShared& hs, int hi)
{
    return __builtin_bit_cast(int64_t, (&hs.rsub[0][0])[hi]);
}
__device__ __forceinline__ double probe_hub_base(const HubShared& hs, int hi)
{
    return (&hs.radd[0][0])[hi];
}
__device__ __forceinline__ void probe_term_pre(const Item& X, HubShared& hs, const LightShared& ls, SbcBucketEnt* bk,
                                               int hi, int own, int pos)
{
    const int code = ls.eflag[own];
    if ((code & 3) == 0) return;
    const int64_t sk = probe_hub_sigma(hs, hi);
    const bool child = (code & 3) == 1;
    const bool hsub = !child || !(hs.plist[hi] == X.p && (code & 4));
    const double add = child ? dep(sk, ls.sk_new[own], ls.eadd[own]) : 0.0;
    const double sub = hsub ? dep(sk, ls.sk_old[own], ls.base[own]) : 0.0;
    const int slot = atomicAdd(&hs.bcnt[hi], 1);
    if (slot < hs.bstride) {
        SbcBucketEnt& e = bk[(size_t)hi * hs.bstride + slot];
        e.key = pos;
        e.flag = (child ? 1 : 0) | (hsub ? 2 : 0);
        e.add = add;
        e.sub = sub;
    }
}

// Lists x as an ascent-only parent unless it is already marked.  x is one
// level above the level being ascended and of type 0 or 3 (callers: a
// neighbour one level up of a processed vertex, a parent candidate that no
// descent touched, the root, a deletion's shallower endpoint).
__device__ __forceinline__ void mark_parent(const Item& X, TeamShared& ts, int32_t x)
{
    if (mark_sb3(&X.sb[x])) X.list[atomicAdd(&ts.cnt[2], 1)] = x;
}


__device__ __forceinline__ int phash_find(const HubShared& hs, int32_t v)
{
    if (!sbit_test(hs, v)) return -1;
    int h = phash_slot(v);
    for (;;) {
        const int key = hs.phk[h];
        if (key == v + 1) return hs.phv[h];
        if (key == 0) return -1;
        h = (h + 1) & (PHASH - 1);
    }
}

// Source vertices of one lookup pass: the processed vertices one level
// down (candidate children), the descent vertices on the level (candidate
// same-level neighbours) and, on level 2, the root's neighbours (candidate
// parents).
struct ProbeSources {
    LevelSets below;
    int dA, dB;          // descent vertices on the level
    const int32_t* pc;   // parent candidates: every vertex of level lev-1 (level 2: the root's
                         // neighbours; level 3: the root's two-hop set), possibly more
    int nr;              // number of parent candidates
    __device__ int size() const { return below.size() + (dB - dA) + nr; }
};

// The root's two-hop set (neighbours of its neighbours, duplicates kept)
// into hs.two, or hs.ntwo = -1 if it has more than TWOHOP_CAP entries.
// Every level-2 vertex is in it, and it holds no vertex below level 2, so a
// member adjacent to a level-3 vertex is one of that vertex's parents.
template <int TW>
__device__ __forceinline__ void two_hop(const SbcParams& P, const Item& X, int32_t rstart, int len_r, HubShared& hs,
                                        LightShared& ls)
{
    const int tid = threadIdx.x;
    constexpr int NT = TW * 64;
    int lx = 0;
    int32_t xs = 0;
    if (len_r <= NT && tid < len_r) {
        const int32_t x = P.col[rstart + tid];
        const SbcVInfo vi = P.vinfo[x];
        lx = visible_len_thread(P, vi, X.q);
        xs = vi.start;
    }
    int total;
    const int off = block_exscan<TW>(lx, ls.scan, total);
    const bool fits = len_r <= NT && total <= TWOHOP_CAP;
    // every thread copies entries: entry i belongs to the last neighbour
    // whose offset is at most i (a neighbour with no entries shares its
    // offset with the next one)
    if (fits && tid < len_r) {
        ls.off[tid] = off;
        ls.start[tid] = xs;
    }
    __syncthreads();
    if (fits) {
        for (int i = tid; i < total; i += NT) {
            int lo = 0, hi = len_r;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (ls.off[mid] <= i) lo = mid;
                else hi = mid;
            }
            hs.two[i] = P.col[ls.start[lo] + (i - ls.off[lo])];
        }
    }
    if (tid == 0) hs.ntwo = fits ? total : -1;
    __syncthreads();
}

// One lane's entry of a hub's term bucket (lanes >= n hold an empty entry
// with the largest key).
struct BucketLane {
    int key;
    int flag;
    double add, sub;
};

__device__ __forceinline__ BucketLane bucket_lane(const SbcBucketEnt* b, int n)
{
    const int lane = lane_id();
    BucketLane e{INF32, 0, 0.0, 0.0};
    if (lane < n) {
        e.key = b[lane].key;
        e.flag = b[lane].flag;
        e.add = b[lane].add;
        e.sub = b[lane].sub;
    }
    return e;
}

__device__ __forceinline__ double shfl_d(double x, int src)
{
    int2 v = __builtin_bit_cast(int2, x);
    v.x = __shfl(v.x, src);
    v.y = __shfl(v.y, src);
    return __builtin_bit_cast(double, v);
}

// Ascending sort of up to 256 (key, idx) pairs held by a wavefront as
// element e = j * 64 + lane in key[j], idx[j]: a bitonic network whose
// strides below 64 exchange between lanes and whose strides of 64 and 128
// exchange between a lane's registers.  Unused elements carry INF32 keys.
template <int J>
__device__ __forceinline__ void wave_sort_keys(int (&key)[J], int (&idx)[J])
{
    const int lane = lane_id();
#pragma unroll
    for (int k2 = 2; k2 <= 64 * J; k2 <<= 1) {
#pragma unroll
        for (int d = k2 >> 1; d > 0; d >>= 1) {
            if (d >= 64) {
                const int dj = d >> 6;
#pragma unroll
                for (int j = 0; j < J; ++j) {
                    const int pj = j ^ dj;
                    if (pj > j) {
                        const bool up = ((j * 64 + lane) & k2) == 0;
                        if (up ? key[j] > key[pj] : key[j] < key[pj]) {
                            const int tk = key[j], ti = idx[j];
                            key[j] = key[pj]; idx[j] = idx[pj];
                            key[pj] = tk; idx[pj] = ti;
                        }
                    }
                }
            } else {
#pragma unroll
                for (int j = 0; j < J; ++j) {
                    const int pk = __shfl_xor(key[j], d), pi = __shfl_xor(idx[j], d);
                    const bool up = ((j * 64 + lane) & k2) == 0;
                    const bool lower = (lane & d) == 0;
                    const bool take = lower ? ((pk < key[j]) == up) : ((pk > key[j]) == up);
                    if (take) { key[j] = pk; idx[j] = pi; }
                }
            }
        }
    }
}

// base + the bucket's terms in position order: a wavefront bitonic sort of
// (key, lane), then each lane takes the terms of the lane it sorted to.
// base + the terms of a bucket of n <= 64 J entries in key order: J keys
// per lane sorted by one wavefront, then the entries in that order
template <int J>
__device__ __forceinline__ double bucket_sum_sorted(const SbcBucketEnt* b, int n, double base)
{
    const int lane = lane_id();
    int key[J], idx[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int e = j * 64 + lane;
        key[j] = e < n ? b[e].key : INF32;
        idx[j] = e;
    }
    wave_sort_keys<J>(key, idx);
    double acc = base;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        int fl = 0;
        double ad = 0.0, sb2 = 0.0;
        if (key[j] != INF32) {
            const SbcBucketEnt& e = b[idx[j]];
            fl = e.flag;
            ad = e.add;
            sb2 = e.sub;
        }
        acc = wave_ordered_add(acc, (fl & 1) != 0, ad, (fl & 2) != 0, sb2);
    }
    return acc;
}

__device__ __forceinline__ double bucket_sum(const BucketLane& e, double base)
{
    const int lane = lane_id();
    int key = e.key;
    int idx = lane;
    // a network over the first m lanes, m the power of two that holds the
    // terms (the other lanes hold the largest key)
    const uint64_t vm = ballot(key != INF32);
    const int nn = __popcll(vm);
    const int m = nn <= 16 ? 16 : (nn <= 32 ? 32 : 64);
    for (int k2 = 2; k2 <= m; k2 <<= 1) {
        for (int j = k2 >> 1; j > 0; j >>= 1) {
            const int pk = __shfl_xor(key, j), pi = __shfl_xor(idx, j);
            const bool up = (lane & k2) == 0;
            const bool lower = (lane & j) == 0;
            const bool take = lower ? ((pk < key) == up) : ((pk > key) == up);
            if (take) { key = pk; idx = pi; }
        }
    }
    const int fl = __shfl(e.flag, idx);
    const double ad = shfl_d(e.add, idx);
    const double sb2 = shfl_d(e.sub, idx);
    return wave_ordered_add(base, (fl & 1) != 0, ad, (fl & 2) != 0, sb2);
}

// New deltas of the ascent-only hubs hs.plist[0..np) of level lev (1 or 2)
// without scanning their adjacency lists.  Their terms can only come from
// processed vertices one level down (children) and moved descent vertices on
// their own level, and on level 2 their parents are exactly the root's
// neighbours adjacent to them.  Every source vertex w either scans its own
// list for lookup hubs (LDS hash) when that list is short, or is looked up in
// each hub's table; each (w, hub) incidence gives the position of w in the
// hub's list.  Inserted edges of the hubs are checked slot by slot.  Each
// hub's terms are sorted by position and summed in that order, the order of
// the full pull scan, so the result is bit-identical to it.  Hubs with more
// terms than their bucket holds are marked kind 2 and left to the caller,
// which scans them (parent marks are idempotent).
template <int TW>
__device__ __forceinline__ void ascent_hubs_probe(const SbcParams& P, const Item& X, int np, int lev,
                                                  TeamShared& ts, HubShared& hs, LightShared& ls,
                                                  const ProbeSources& src, int64_t* tp)
{
    const int tid = threadIdx.x, lane = lane_id(), twave = tid >> 6;
    constexpr int NT = TW * 64;
    SbcBucketEnt* bk = P.bucket + (size_t)blockIdx.x * SBC_BUCKET_SLOTS;
    int64_t tq = 0;  // trace: sub-step clock
    if (tp && tid == 0) tq = (int64_t)wall_clock64();
    for (int i = tid; i < PHASH; i += NT) hs.phk[i] = 0;
    for (int i = tid; i < SBITS_WORDS; i += NT) hs.sbits[i] = 0;
    for (int i = tid; i < np; i += NT) hs.bcnt[i] = 0;
    if (tid == 0) {
        const int share = (SBC_BUCKET_SLOTS / np) & ~63;
        hs.bstride = share < TEAM_SORT_CAP ? share : TEAM_SORT_CAP;
        hs.n3 = 0;
    }
    __syncthreads();
    for (int hi = tid; hi < np; hi += NT) {
        const int32_t k = hs.plist[hi];
        const int4 hb = P.hub[k];
        hs.poff[hi] = hb.x;
        hs.plog[hi] = (uint8_t)hb.z;
        (&hs.radd[0][0])[hi] = asc_base(P, X, k);
        (&hs.rsub[0][0])[hi] = __builtin_bit_cast(double, X.vr[k].sigma);
        int h = phash_slot(k);
        while (atomicCAS(&hs.phk[h], 0, k + 1) != 0) h = (h + 1) & (PHASH - 1);
        hs.phv[h] = hi;
        sbit_set(hs, k);
    }
    __syncthreads();
    if (tp && tid == 0) { const int64_t x = (int64_t)wall_clock64(); tp[72] += x - tq; tq = x; }
    const int nb = src.below.size(), ndl = src.dB - src.dA;
    const int nsrc = src.size();
    // inserted edges of the hubs (a few each), one thread per hub, beside the
    // first source round (they share no data; its barriers cover them)
    for (int hi = tid; hi < np; hi += NT) {
        const int32_t k = hs.plist[hi];
        const int len = hs.plen[hi];
        const int start = hs.pstart[hi];
        for (int e = P.hub[k].y; e < len; ++e) {
            const int32_t c = P.col[start + e];
            const uint8_t b = ld_sb(&X.sb[c]);
            if (sb_type(b) == 0) {
                if (lev >= 1 && !(b & SB_INF) && sb_rel(b, lev) == -1) mark_parent(X, ts, c);
            } else {
                probe_term(P, X, hs, bk, k, hi, lev, c, e);
            }
        }
    }
    if (lev == 1 && tid == 0) mark_parent(X, ts, X.r);
    for (int r0 = 0; r0 < nsrc; r0 += NT) {
        // one source per thread; short lists are scanned, long ones looked up
        const int i = r0 + tid;
        int32_t w = 0;
        int len = 0;
        int start = 0;
        bool scan_mode = false, table_mode = false;
        if (i < nsrc) {
            if (i < nb) w = src.below.at(X, i);
            else if (i < nb + ndl) w = X.list[src.dA + (i - nb)];
            else w = src.pc[i - nb - ndl];
            const SbcVInfo vi = P.vinfo[w];
            // a parent candidate that is already marked (or processed) can
            // only be left as it is, so it is not looked up at all
            const bool done = i >= nb + ndl && sb_type(ld_sb(&X.sb[w])) != 0;
            len = visible_len_thread(P, vi, X.q);
            start = vi.start;
            scan_mode = !done && len <= 2 * np;
            table_mode = !done && !scan_mode;
            // the source's state, kept in LDS so that the probes load
            // nothing of it (slot = this thread)
            int code = 0;
            if (i < nb + ndl) {
                // a processed vertex one level down or a descent vertex on the
                // level: its term inputs, whichever hub it meets
                const uint8_t bw = ld_sb(&X.sb[w]);
                const int64_t so = X.vr[w].sigma;
                const int64_t dp = ld_acq64(&X.vr[w].dpath);
                const double nd = ld_accd(&X.vr[w].ndel);
                const double od = X.vr[w].delta;
                const int ct = sb_type(bw);
                const int rel = sb_rel(bw, lev);
                if (!(bw & SB_INF) && ct != 0) {
                    if (rel == 1) code = 1;
                    else if (rel == 0 && ct == 2) code = 2;
                }
                if (w == X.s) code |= 4;
                ls.sk_old[tid] = so;
                ls.sk_new[tid] = ct == 3 ? so : (int64_t)((uint64_t)dp + (uint64_t)(ct == 1 ? so : 0));
                ls.base[tid] = od;
                ls.eadd[tid] = nd;
            }
            ls.eflag[tid] = (uint8_t)code;
            ls.k[tid] = w;
            ls.isD[tid] = (i >= nb + ndl) ? 1 : 0;  // 1: parent candidate
        }
        // edge offsets of scan-mode sources, slots of table-mode sources and
        // compact scan-mode slots (number of scan-mode threads before), in
        // thread order
        int ne, nt, nsc;
        const int3 sc = block_exscan3<TW>(scan_mode ? len : 0, table_mode ? 1 : 0, scan_mode ? 1 : 0, ls.scan, ne,
                                          nt, nsc);
        const int eoff = sc.x, toff = sc.y, sslot = sc.z;
        if (scan_mode) {
            ls.start[tid] = start;
            ls.off[tid] = eoff;
            ls.len[sslot] = tid;                    // owner thread of slot
        }
        if (table_mode) ls.ovf[toff] = tid;
        __syncthreads();
        if (tp && tid == 0) { const int64_t x = (int64_t)wall_clock64(); tp[73] += x - tq; tq = x; }
        // edges of scan-mode sources
        for (int e = tid; e < ne; e += NT) {
            int lo = 0, hi = nsc;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (ls.off[ls.len[mid]] <= e) lo = mid;
                else hi = mid;
            }
            const int own = ls.len[lo];
            const int32_t ws = ls.k[own];
            const int32_t c = P.col[ls.start[own] + (e - ls.off[own])];
            const int hi2 = phash_find(hs, c);
            if (hi2 < 0) continue;
            if (ls.isD[own]) {
                mark_parent(X, ts, ws);
            } else {
                const int pos = hub_find(P.htab + hs.poff[hi2], hs.plog[hi2], ws);
                if (pos >= 0) probe_term_pre(X, hs, ls, bk, hi2, own, pos);
            }
        }
        // table-mode sources against every hub, one (source, hub) pair per
        // thread: the whole bucket of the source in the hub's table in one
        // load (64 bytes), the next bucket only if that one is full
        const int npairs = nt * np;
        for (int t = tid; t < npairs; t += NT) {
            const int sj = t / np;
            const int hi2 = t - sj * np;
            const int si = ls.ovf[sj];  // the source's slot
            const int32_t ws = ls.k[si];
            const int lg = hs.plog[hi2];
            const int2* tab = P.htab + hs.poff[hi2];
            const uint32_t b0 = hub_bucket(ws, lg);
            const int4* q = reinterpret_cast<const int4*>(tab + ((size_t)b0 << 3));
            int pos = bucket_match(q[0], q[1], q[2], q[3], ws + 1);
            if (pos == -2) pos = hub_find_bucket(tab, lg, ws, (b0 + 1u) & ((1u << (lg - 3)) - 1u));
            if (pos < 0) continue;
            if (ls.isD[si]) mark_parent(X, ts, ws);
            else probe_term_pre(X, hs, ls, bk, hi2, si, pos);
        }
        __syncthreads();
        if (tp && tid == 0) { const int64_t x = (int64_t)wall_clock64(); tp[74] += x - tq; tq = x; }
    }
    if (nsrc == 0) __syncthreads();
    if (tp && tid == 0) { const int64_t x = (int64_t)wall_clock64(); tp[75] += x - tq; tq = x; }
    if (tp && tid == 0) tp[12] -= (int64_t)wall_clock64();  // trace: sort-and-sum time
    // buckets of at most 8 terms: eight hubs per wavefront, an 8-lane group
    // each (bitonic sort of the group's keys, the group leader sums in key
    // order); larger buckets are marked and done one wavefront per hub
    {
        const int g = lane >> 3, gl = lane & 7, gb = lane & ~7;
        for (int h0 = twave * 8; h0 < np; h0 += TW * 8) {
            const int hi = h0 + g;
            const int n = hi < np ? hs.bcnt[hi] : 0;
            const bool small = hi < np && n <= 8;
            int key = INF32, fl = 0, idx = gl;
            double ad = 0.0, sb2 = 0.0, base = 0.0;
            if (small) {
                if (gl < n) {
                    const SbcBucketEnt& e = bk[(size_t)hi * hs.bstride + gl];
                    key = e.key;
                    fl = e.flag;
                    ad = e.add;
                    sb2 = e.sub;
                }
                if (gl == 0) base = probe_hub_base(hs, hi);
            }
            // the wavefront's largest small bucket (n is the same in a group)
            int nm = small ? n : 0;
            for (int o = 8; o < 64; o <<= 1) nm = max(nm, __shfl_xor(nm, o));
            double acc = base;
            if (nm >= 2) {
                for (int k2 = 2; k2 <= 8; k2 <<= 1) {
                    for (int jj = k2 >> 1; jj > 0; jj >>= 1) {
                        const int pk = __shfl_xor(key, jj), pi = __shfl_xor(idx, jj);
                        const bool up = k2 == 8 || (lane & k2) == 0;
                        const bool lower = (lane & jj) == 0;
                        const bool take = lower ? ((pk < key) == up) : ((pk > key) == up);
                        if (take) { key = pk; idx = pi; }
                    }
                }
                const int src = gb | idx;
                const int sfl = __shfl(fl, src);
                const double sad = shfl_d(ad, src), ssb = shfl_d(sb2, src);
                for (int q = 0; q < nm; ++q) {
                    const int f2 = __shfl(sfl, gb + q);
                    const double a2 = shfl_d(sad, gb + q), b2 = shfl_d(ssb, gb + q);
                    if (gl == 0 && q < n) {
                        if (f2 & 1) acc = acc + a2;
                        if (f2 & 2) acc = acc - b2;
                    }
                }
            } else if (gl == 0 && n == 1) {
                // one term: the group leader's own entry
                if (fl & 1) acc = acc + ad;
                if (fl & 2) acc = acc - sb2;
            }
            if (small && gl == 0) st_reld(&X.vr[hs.plist[hi]].ndel, acc);
            if (hi < np && gl == 0) {
                // kinds: 0 (<= 8 terms, done here), 1 (<= 64, a wavefront), 4 (<= 256,
                // a wavefront's 256-key sort), 3 (more, the team), 2 (overflowed: scanned)
                const uint8_t kd = n > hs.bstride ? 2
                                 : n > WAVE_SORT_CAP ? (hs.plen[hi] <= TEAM_SORT_BITS ? 3 : 2)
                                 : n > SBC_BUCKET ? 4
                                 : (n > 8 ? 1 : 0);
                hs.kind[hi] = kd;
                if (kd == 3) atomicAdd(&hs.n3, 1);
            }
        }
    }
    __syncthreads();
    if (tp && tid == 0) { const int64_t x = (int64_t)wall_clock64(); tp[76] += x - tq; tq = x; }
    // one wavefront per larger bucket (kinds 1 and 4)
    for (int h0 = 0; h0 < np; h0 += 64) {
        const int hl = h0 + lane;
        uint64_t bm = ballot(hl < np && (hs.kind[hl] == 1 || hs.kind[hl] == 4));
        // the wavefronts share out the marked hubs of this group of 64
        int r = 0;
        while (bm) {
            const int l0 = __builtin_ctzll(bm);
            bm &= bm - 1;
            if ((r++ % TW) != twave) continue;
            const int hi = h0 + l0;
            const int n = hs.bcnt[hi];
            const SbcBucketEnt* b = bk + (size_t)hi * hs.bstride;
            double acc;
            if (hs.kind[hi] == 1) {
                acc = bucket_sum(bucket_lane(b, n), probe_hub_base(hs, hi));
            } else {
                acc = bucket_sum_sorted<4>(b, n, probe_hub_base(hs, hi));
            }
            if (lane == 0) st_reld(&X.vr[hs.plist[hi]].ndel, acc);
        }
    }
    __syncthreads();
    if (tp && tid == 0) { const int64_t x = (int64_t)wall_clock64(); tp[77] += x - tq; tq = x; }
    // buckets of more than SBC_BUCKET terms (kind 3), one hub at a time by
    // the whole team: the entries' positions are set in an LDS bitmap over
    // the hub's list, each entry's rank is the number of bits below its own
    // (a prefix of per-word popcounts plus a masked popcount), the entries
    // are placed at their ranks and one wavefront sums them in that order.
    // A position set twice would make two ranks equal: the hub is then
    // handed to the scans (kind 2) instead.
    const int n3 = hs.n3;
    uint32_t* bits = reinterpret_cast<uint32_t*>(ls.base);  // TEAM_SORT_BITS bits
    for (int hi = 0, f = 0; f < n3; ++hi) {
        if (hs.kind[hi] != 3) continue;
        ++f;
        const int n = hs.bcnt[hi];
        const SbcBucketEnt* b = bk + (size_t)hi * hs.bstride;
        const int nw = (hs.plen[hi] + 31) >> 5;
        int64_t tk = 0;
        if (tp && tid == 0) { tk = (int64_t)wall_clock64(); tp[28] += n; tp[29] += 1; }
        for (int i = tid; i < nw; i += NT) bits[i] = 0u;
        if (tid == 0) hs.dup = 0;
        __syncthreads();
        for (int i = tid; i < n; i += NT) {
            const SbcBucketEnt e = b[i];
            ls.off[i] = e.key;
            ls.isD[i] = (uint8_t)e.flag;
            ls.sk_old[i] = __builtin_bit_cast(int64_t, e.add);
            ls.sk_new[i] = __builtin_bit_cast(int64_t, e.sub);
            const uint32_t m = 1u << (e.key & 31);
            if (atomicOr(&bits[e.key >> 5], m) & m) hs.dup = 1;
        }
        __syncthreads();
        if (tp && tid == 0) { const int64_t x = (int64_t)wall_clock64(); tp[24] += x - tk; tk = x; }
        int tot_bits;
        int wpre = 0;
        for (int w0 = 0; w0 < nw; w0 += NT) {
            const int w = w0 + tid;
            const int c = w < nw ? __popc(bits[w]) : 0;
            const int off = block_exscan<TW>(c, ls.scan, tot_bits);
            if (w < nw) ls.ovf[w] = wpre + off;
            wpre += tot_bits;
        }
        __syncthreads();
        if (tp && tid == 0) { const int64_t x = (int64_t)wall_clock64(); tp[25] += x - tk; tk = x; }
        const bool dup = hs.dup != 0;
        if (!dup) {
            for (int i = tid; i < n; i += NT) {
                const int key = ls.off[i];
                const int rank = ls.ovf[key >> 5] + __popc(bits[key >> 5] & ((1u << (key & 31)) - 1u));
                ls.eflag[rank] = ls.isD[i];
                ls.eadd[rank] = __builtin_bit_cast(double, ls.sk_old[i]);
                ls.esub[rank] = __builtin_bit_cast(double, ls.sk_new[i]);
            }
        } else if (tid == 0) {
            hs.kind[hi] = 2;
        }
        __syncthreads();
        if (tp && tid == 0) { const int64_t x = (int64_t)wall_clock64(); tp[26] += x - tk; tk = x; }
        if (!dup && twave == 0) {
            // every value in order into the packing rows (a subtraction as the
            // addition of the negated value, which IEEE 754 defines it to be),
            // then one add chain by lane 0
            const int32_t k = hs.plist[hi];
            double* row = &sbc_wrow[0][0];  // 2048 values >= 2 * TEAM_SORT_CAP
            const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
            int nval = 0;
            for (int e0 = 0; e0 < n; e0 += 64) {
                const int e = e0 + lane;
                const uint8_t fl = e < n ? ls.eflag[e] : (uint8_t)0;
                const uint64_t cm = ballot(fl & 1), sm = ballot(fl & 2);
                int pos = nval + __popcll(cm & below) + __popcll(sm & below);
                if (fl & 1) row[pos++] = ls.eadd[e];
                if (fl & 2) row[pos] = -ls.esub[e];
                nval += __popcll(cm) + __popcll(sm);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane == 0) st_reld(&X.vr[k].ndel, lds_chain_sum(probe_hub_base(hs, hi), row, nval));
        }
        __syncthreads();
        if (tp && tid == 0) tp[27] += (int64_t)wall_clock64() - tk;
    }
    if (tp && tid == 0) tp[12] += (int64_t)wall_clock64();
}

// Inputs of an ascent level's source set (see shash_insert): the processed
// vertices one level down, the descent vertices list[dA..dB) on the level,
// and every vertex one level up (npc of them at pcs; npc == 1 with pcs null:
// the root; npc < 0: unknown, no set).
struct SrcSet {
    LevelSets below;
    int dA, dB;
    const int32_t* pcs;
    int npc;
};
// Light edges of a round from which the source set pays for itself.
#ifndef SBC_SET_MIN_EDGES
#define SBC_SET_MIN_EDGES 4096
#endif
constexpr int SET_MIN_EDGES = SBC_SET_MIN_EDGES;

// All vertices of ascent level `here`, in rounds of one vertex per thread:
// hubs are queued in hs.list for the hub passes; light vertices get a slot
// (their rank in the round) holding their adjacency range and ascent state.
// The first light vertices whose edges fit EDGE_CAP form the batch (one lane
// per edge, terms kept at the edge's index, then each vertex sums its terms
// in list order); the rest of the round's light vertices run one wavefront
// per vertex from their slots while the batch computes its terms.
template <int TW>
__device__ __forceinline__ void ascent_light_level(const SbcParams& P, const Item& X, const LevelSets& here, int tot,
                                                   int lev, TeamShared& ts, HubShared& hs, LightShared& ls,
                                                   int twave, uint32_t& my_proc, uint32_t& my_scan, int64_t* tp,
                                                   const SrcSet& ss)
{
    int64_t tq = 0;  // trace: round setup (13), edge terms and leftovers (14), batch sums (15)
    const int tid = threadIdx.x, lane = lane_id();
    constexpr int NT = TW * 64;
    static_assert(NT <= LIGHT_CAP, "one slot per thread of a round");
    if (tot <= TW) {
        // a small level: one wavefront per vertex, no batch and no block-wide
        // scans (the caller's barrier follows); hubs are queued as in a round
        if (twave < tot) {
            const int32_t k = here.at(X, twave);
            const SbcVInfo vi = P.vinfo[k];
            const int len = visible_len(P, vi, X.q);
            int pos = -1;  // place in the hub queues, or -1 (the wavefront ascends it)
            if (len > HUB_MIN) {
                if (lane == 0) pos = atomicAdd(&hs.n, 1);
                pos = __shfl(pos, 0);
                const int hcap = hub_batch(P);
                if (pos < hcap) {
                    if (lane == 0) {
                        hs.list[pos] = k;
                        hs.hlen[pos] = len;
                        hs.hstart[pos] = vi.start;
                        hs.hidx[pos] = P.hub ? P.hub[k].y : 0;
                        hs.hasc[pos] = sb_type(ld_sb(&X.sb[k])) == 3 ? 1 : 0;
                    }
                } else if (P.hubq && pos < hcap + SBC_HUBQ_CAP) {
                    if (lane == 0) P.hubq[(size_t)blockIdx.x * SBC_HUBQ_CAP + (pos - hcap)] = k;
                } else {
                    pos = -1;  // no queue takes it
                }
            }
            if (pos < 0) {
                ascent_vertex_wave(P, X, k, vi, len, lev, ts);
                my_proc += 1;
                my_scan += (uint32_t)len;
            }
        }
        return;
    }
    bool set_built = false;  // the level's source set, kept for the later rounds
    for (int r0 = 0; r0 < tot; r0 += NT) {
        if (tp && tid == 0) tq = (int64_t)wall_clock64();
        const int t = r0 + tid;
        int32_t k = 0;
        int len = 0;
        SbcVInfo vi{0, 0, 0, 0};
        KState kst{false, 0, 0, 0.0};
        bool light = false;
        if (t < tot) {
            k = here.at(X, t);
            vi = P.vinfo[k];
            kst = kstate(P, X, k);  // loaded with vinfo; unused for hubs
            len = visible_len_thread(P, vi, X.q);
            if (len > HUB_MIN) {
                const int pos = atomicAdd(&hs.n, 1);
                const int hcap = hub_batch(P);
                if (pos < hcap) {
                    // what the first hub pass needs, from this round's loads
                    hs.list[pos] = k;
                    hs.hlen[pos] = len;
                    hs.hstart[pos] = vi.start;
                    hs.hidx[pos] = P.hub ? P.hub[k].y : 0;
                    hs.hasc[pos] = kst.isD ? 0 : 1;
                }
                else if (P.hubq && pos < hcap + SBC_HUBQ_CAP)
                    P.hubq[(size_t)blockIdx.x * SBC_HUBQ_CAP + (pos - hcap)] = k;
                else ls.ovf[atomicAdd(&ls.novf, 1)] = k;
            } else {
                light = true;
            }
        }
        int nl, ne_round;
        const int2 se = block_exscan2<TW>(light ? 1 : 0, light ? len : 0, ls.scan, nl, ne_round);
        const int slot = se.x, eoff = se.y;
        if (light) {
            ls.k[slot] = k;
            ls.start[slot] = vi.start;
            ls.off[slot] = eoff;
            ls.len[slot] = len;
            ls.isD[slot] = kst.isD ? 1 : 0;
            ls.sk_old[slot] = kst.sk_old;
            ls.sk_new[slot] = kst.sk_new;
            ls.base[slot] = kst.base;
            if (eoff + len <= EDGE_CAP) atomicMax(&ls.nb, slot + 1);
        }
        __syncthreads();
        const int nsrc = ss.npc < 0 ? -1 : ss.below.size() + (ss.dB - ss.dA) + ss.npc;
        const bool use_set = set_built || (nsrc >= 0 && nsrc <= SHASH_MAX && ne_round >= SET_MIN_EDGES);
        if (use_set && !set_built) {
            set_built = true;
            for (int i = tid; i < PHASH; i += NT) hs.phk[i] = 0;
            for (int i = tid; i < SBITS_WORDS; i += NT) hs.sbits[i] = 0;
            __syncthreads();
            const int nbl = ss.below.size(), ndl = ss.dB - ss.dA;
            for (int i = tid; i < nsrc; i += NT) {
                int32_t w;
                if (i < nbl) w = ss.below.at(X, i);
                else if (i < nbl + ndl) w = X.list[ss.dA + (i - nbl)];
                else w = ss.pcs ? ss.pcs[i - nbl - ndl] : X.r;
                shash_insert(hs, w);
            }
            __syncthreads();
        }
        if (tp && tid == 0) { const int64_t x = (int64_t)wall_clock64(); tp[13] += x - tq; tq = x; }
        const int nb = ls.nb;
        const int no = ls.novf;
        const int ne = nb > 0 ? ls.off[nb - 1] + ls.len[nb - 1] : 0;
        if (tp && tid == 0 && lev == 2) {
            tp[64] += nl;
            tp[65] += nb;
            tp[66] += ne;
            tp[67] += nl - nb;
            tp[68] += ne_round - ne;
            tp[69] += 1;
        }
        // batch: one lane per edge, two edges per thread in flight
        for (int e0 = tid; e0 < ne; e0 += 2 * NT) {
            int slot2[2], ee[2];
            int32_t c2[2];
            bool ok2[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int e = e0 + u * NT;
                ok2[u] = e < ne;
                ee[u] = e;
                int lo = 0, hi = nb;  // last slot with off <= e
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (ls.off[mid] <= e) lo = mid;
                    else hi = mid;
                }
                slot2[u] = lo;
                c2[u] = ok2[u] ? P.col[ls.start[lo] + (e - ls.off[lo])] : 0;
            }
            bool g2[2];  // neighbour in the level's source set (or no set)
#pragma unroll
            for (int u = 0; u < 2; ++u) g2[u] = ok2[u] && (!use_set || shash_has(hs, c2[u]));
            uint8_t b2[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) b2[u] = g2[u] ? ld_sb(&X.sb[c2[u]]) : (uint8_t)0;
            Term tm[2];
            bool appU[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int lo = slot2[u];
                tm[u] = ascent_lane(P, X, ls.k[lo], ls.isD[lo] != 0, ls.sk_old[lo], ls.sk_new[lo], lev, g2[u], c2[u],
                                    b2[u], appU[u]);
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                if (!ok2[u]) continue;
                if (appU[u]) X.list[atomicAdd(&ts.cnt[2], 1)] = c2[u];
                ls.eflag[ee[u]] = (uint8_t)((tm[u].child ? 1 : 0) | (tm[u].hsub ? 2 : 0));
                ls.eadd[ee[u]] = tm[u].add;
                ls.esub[ee[u]] = tm[u].sub;
            }
        }
        // light vertices past the cap, one wavefront each, from their slots
        for (int sl = nb + twave; sl < nl; sl += TW) {
            KState ks;
            ks.isD = ls.isD[sl] != 0;
            ks.sk_old = ls.sk_old[sl];
            ks.sk_new = ls.sk_new[sl];
            ks.base = ls.base[sl];
            const int l2 = ls.len[sl];
            ascent_vertex_wave_ks(P, X, ls.k[sl], ls.start[sl], l2, ks, lev, ts, use_set ? &hs : nullptr,
                                  lev == 2 ? tp : nullptr);
            my_proc += 1;
            my_scan += (uint32_t)l2;
        }
        // hubs no queue could take
        for (int i = twave; i < no; i += TW) {
            const int32_t kk = ls.ovf[i];
            const SbcVInfo v2 = P.vinfo[kk];
            const int l2 = visible_len(P, v2, X.q);
            ascent_vertex_wave(P, X, kk, v2, l2, lev, ts);
            my_proc += 1;
            my_scan += (uint32_t)l2;
        }
        if (tid == 0) {
            my_proc += (uint32_t)nb;
            my_scan += (uint32_t)ne;
        }
        __syncthreads();
        if (tp && tid == 0) { const int64_t x = (int64_t)wall_clock64(); tp[14] += x - tq; tq = x; }
        // every thread holds nb and no: the counters are free for the next round
        if (tid == 0) { ls.nb = 0; ls.novf = 0; }
        // one wavefront per batched vertex: sum its edges' terms in list order
        for (int sl = twave; sl < nb; sl += TW) {
            const int a = ls.off[sl], l = ls.len[sl];
            double acc = ls.base[sl];
            for (int e0 = 0; e0 < l; e0 += 64) {
                const int e = e0 + lane;
                const uint8_t fl = e < l ? ls.eflag[a + e] : (uint8_t)0;
                const double ad = (fl & 1) ? ls.eadd[a + e] : 0.0;
                const double sb2 = (fl & 2) ? ls.esub[a + e] : 0.0;
                acc = wave_ordered_add(acc, (fl & 1) != 0, ad, (fl & 2) != 0, sb2);
            }
            if (lane == 0) st_reld(&X.vr[ls.k[sl]].ndel, acc);
        }
        __syncthreads();
        if (tp && tid == 0) tp[15] += (int64_t)wall_clock64() - tq;
    }
}

// Routes vertex k of the current pass: multi-wave teams queue hubs in LDS
// (processed afterwards by the whole team); everything else runs on the
// calling wavefront.  Returns true if k was queued.
template <int TW>
__device__ __forceinline__ bool queue_hub(const SbcParams& P, HubShared* hs, int32_t k, int len)
{
    if (TW == 1 || len <= HUB_MIN) return false;
    int pos = 0;
    if (lane_id() == 0) pos = atomicAdd(&hs->n, 1);
    pos = __shfl(pos, 0);
    const int hcap = hub_batch(P);
    if (pos < hcap) {
        if (lane_id() == 0) hs->list[pos] = k;
        return true;
    }
    if (P.hubq && pos < hcap + SBC_HUBQ_CAP) {
        if (lane_id() == 0) P.hubq[(size_t)blockIdx.x * SBC_HUBQ_CAP + (pos - hcap)] = k;
        return true;
    }
    return false;
}

// Number of queued hubs of the current pass, and batch hb..hb+batch of
// them copied into the LDS queue (batch 0 is already there).
__device__ __forceinline__ int hubs_queued(const SbcParams& P, const HubShared& hs)
{
    return min(hs.n, hub_batch(P) + (P.hubq ? SBC_HUBQ_CAP : 0));
}

template <int TW>
__device__ __forceinline__ void load_hub_batch(const SbcParams& P, HubShared& hs, int hb, int nh)
{
    if (hb == 0) return;
    const int32_t* q = P.hubq + (size_t)blockIdx.x * SBC_HUBQ_CAP + (hb - hub_batch(P));
    for (int i = threadIdx.x; i < nh; i += TW * 64) hs.list[i] = q[i];
    __syncthreads();
}

// Adds a finished nonzero item to its root's work totals (only the root's
// team, or the wide commit after it, ever writes them).
__device__ __forceinline__ void add_work_totals(int64_t* t, const unsigned long long* stat, int64_t affected)
{
    t[0] += 1;
    t[1] += (int64_t)stat[0];
    t[2] += (int64_t)stat[1];
    t[3] += affected;
    t[4] += (int64_t)stat[2];
    if (affected > t[5]) t[5] = affected;
}

// Lists root ri as parked (its item saved in P.park[ri]; the caller is the
// team leader, after the team's writes and a barrier) and, when the host
// finishes parked items while the launch runs, publishes it in host memory:
// the root, then its flag with a system-scope release, which makes the
// item's state visible to work launched on another stream.
__device__ __forceinline__ void publish_parked(const SbcParams& P, int ri)
{
    const int pi = atomicAdd(&P.qctr[2], 1);
    P.parked[pi] = ri;
    if (P.hpark_flag) {
        P.hpark_root[pi] = ri;
        __hip_atomic_store(&P.hpark_flag[pi], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Hands the item over to the wide kernels in ascent level lev, after the
// team processed the level's light vertices and looked up its first pass of
// hubs (the light vertices' counts stay in the team's; the wide kernels redo
// the level, counting only its hubs).
template <int TW>
__device__ __forceinline__ void park_in_ascent(const SbcParams& P, int ri, int item, int32_t xs, int32_t xp, int32_t xq,
                                            TeamShared& ts,
                                            int32_t base, int nf, int64_t gap, bool smoved, int lev, int uA, int uB,
                                            uint32_t my_proc, uint32_t my_scan, uint32_t my_moved, bool leader,
                                            int lane)
{
    if (lane == 0) {
        atomicAdd(&ts.stat[0], (unsigned long long)my_proc);
        atomicAdd(&ts.stat[1], (unsigned long long)my_scan);
        atomicAdd(&ts.stat[2], (unsigned long long)my_moved);
    }
    team_sync<TW>();
    if (leader) {
        SbcWideItem& w = P.park[ri];
        w.cnt[0] = ts.cnt[0];
        w.cnt[1] = ts.cnt[1];
        w.cnt[2] = ts.cnt[2];
        w.cnt[3] = 0;
        for (int i = 0; i < 4; ++i) w.stat[i] = ts.stat[i];
        w.item = item;
        w.s = xs;
        w.p = xp;
        w.q = xq;
        w.base = base;
        w.j = nf;
        w.fs = 0;
        w.fe = 0;
        w.gap = (int32_t)(gap > INF32 ? INF32 : (gap < -INF32 ? -INF32 : gap));
        w.smoved = smoved ? 1 : 0;
        w.phase = 1;
        w.lev = lev;
        w.uA = uA;
        w.uB = uB;
        publish_parked(P, ri);
    }
    team_sync<TW>();
}

// Applies window item `item` to root ri.  Returns true if the item was
// parked (its remaining levels are left to the wide kernels).  TRACE: record
// per-item timestamps and counters in P.tprof (a separate instantiation, so
// that the untraced kernel carries none of the trace state).
template <int TW, bool TRACE = false>
__device__ __forceinline__ bool process_item(const SbcParams& P, int ri, int item, TeamShared& ts, HubShared* hs,
                                             LightShared* ls, int twave, int32_t pu = -1, int32_t pv = -1,
                                             int32_t pdu = 0, int32_t pdv = 0)
{
    const int lane = lane_id();
    const bool leader = (twave == 0 && lane == 0);
    const int nv = P.nv;
    const size_t off = (size_t)ri * (size_t)nv;
    Item X;
    X.dist = P.dist + off;
    X.sb = P.sb + (size_t)ri * (size_t)P.sb_stride;
    X.vr = P.vr + off;
    X.list = P.list + off;
    X.nv = nv;
    X.r = P.roots[ri];
    int32_t* lvlD = P.lvl + (size_t)ri * 2 * P.lcap;
    int32_t* lvlS = lvlD + P.lcap;
    // the endpoints and their levels: given (classified in a team pass,
    // no vertex has moved since), or read here
    const int32_t u = pu >= 0 ? pu : (P.win ? P.wu[item] : P.u0);
    const int32_t v = pu >= 0 ? pv : (P.win ? P.wv[item] : P.v0);
    X.q = (P.win == nullptr) ? INF32 : P.win->seq0 + item;

    const int32_t du = pu >= 0 ? pdu : X.dist[u], dv = pu >= 0 ? pdv : X.dist[v];
    // an inserted edge at the root or at one of its neighbours changes the
    // root's two-hop set
    if constexpr (TW > 1)
        if (leader && (du <= 1 || dv <= 1)) hs->two_ok = 0;
    int64_t gap;
    if (du == INF32 && dv == INF32) gap = 0;
    else gap = (int64_t)du - (int64_t)dv;
    SbcItemRec* rec = P.rec ? &P.rec[(size_t)item * P.K + ri] : nullptr;
    int64_t* tp = (TRACE && P.tprof) ? &P.tprof[((size_t)item * P.K + ri) * SBC_TRACE_SLOTS] : nullptr;
    if (tp && leader) {
        const int64_t t0 = (int64_t)wall_clock64();
        tp[0] = t0; tp[1] = t0; tp[2] = t0; tp[3] = t0;
        for (int i = 4; i < SBC_TRACE_SLOTS; ++i) tp[i] = 0;
        for (int i = 0; i < 12; ++i) ts.tr[i] = 0;
    }
    if (gap == 0) {
        if (rec && leader) {
            rec->root_index = ri; rec->insertion = item; rec->kase = 0; rec->gap = 0;
            rec->processed = 0; rec->scanned = 0; rec->affected = 0; rec->moved = 0;
        }
        return false;
    }
    X.s = gap > 0 ? u : v;
    X.p = gap > 0 ? v : u;
    const int64_t agap = gap > 0 ? gap : -gap;
    const bool smoved = agap >= 2;
    const int32_t base = (gap > 0 ? dv : du) + 1; // new level of s (one above p)

    if (leader) {
        st_sb(&X.sb[X.s], sb_make(base, smoved ? 2 : 1));
        const int64_t sp = X.vr[X.p].sigma;
        st_rel64(&X.vr[X.s].dpath, P.del ? (int64_t)(0ull - (uint64_t)sp) : sp);
        X.list[0] = X.s;
        lvlD[0] = 0;
        lvlS[0] = 0;
        ts.cnt[0] = 1;
        ts.cnt[1] = 0;
        ts.cnt[2] = 0;
        ts.stat[0] = 0; ts.stat[1] = 0; ts.stat[2] = 0; ts.stat[3] = 0;
        if (hs) { hs->n = 0; hs->overflow = 0; hs->svol = 0; hs->handover = 0; }
        if (ls) { ls->nb = 0; ls->novf = 0; }
    }
    team_sync<TW>();

    // per-thread work counts of this item (an item scans fewer than 2^32 entries)
    uint32_t my_proc = 0, my_scan = 0, my_moved = 0;

    // ---------------- descent
    int fs = 0, fe = 1, j = 0;
    bool overflow = false;
    while (fs < fe) {
        const int32_t lc = base + j; // new level of frontier j
        for (int idx = fs + twave; idx < fe; idx += TW) {
            const int32_t c = X.list[idx];
            const SbcVInfo vi = P.vinfo[c];
            const int len = visible_len(P, vi, X.q);
            if (queue_hub<TW>(P, hs, c, len)) continue;
            const bool cmoved = sb_type(ld_sb(&X.sb[c])) == 2;
            const int64_t dc = ld_acq64(&X.vr[c].dpath);
            descent_chunks(P, X, c, vi, len, 0, (len + 63) >> 6, X.dist[c], lc, dc, cmoved, ts);
            my_proc += 1;
            my_scan += (uint32_t)len;
            my_moved += cmoved ? 1 : 0;
            if (TW > 1 && P.park && lane == 0) atomicAdd(&ts.stat[3], (unsigned long long)len);
        }
        if constexpr (TW > 1) {
            __syncthreads();
            const int nall = hubs_queued(P, *hs);
            for (int hb = 0; hb < nall; hb += hub_batch(P)) {
                const int nh = min(nall - hb, hub_batch(P));
                load_hub_batch<TW>(P, *hs, hb, nh);
                descent_hubs_flat<TW>(P, X, nh, lc, ts, *hs, *ls, twave);
                __syncthreads();
            }
            if (leader) hs->n = 0;
        }
        team_sync<TW>();
        const int nfe = ts.cnt[0];
        const int nS = ts.cnt[1];
        team_sync<TW>();
        j++;
        if (j + 1 >= P.lcap) { overflow = true; break; }
        // frontier j-1 = list[lvlD[j-1] .. lvlD[j]); S vertices found while
        // expanding it = S list [lvlS[j-1] .. lvlS[j])
        if (leader) { lvlD[j] = fe; lvlS[j] = nS; }
        fs = fe;
        fe = nfe;
        if constexpr (TW > 1) {
            if (P.park && (fe - fs >= P.park_min || (long long)ts.stat[3] >= P.park_scan)) {
                // too wide for one workgroup: save the item, stop this root
                if (lane == 0) {
                    atomicAdd(&ts.stat[0], my_proc);
                    atomicAdd(&ts.stat[1], my_scan);
                    atomicAdd(&ts.stat[2], my_moved);
                }
                team_sync<TW>();
                if (leader) {
                    SbcWideItem& w = P.park[ri];
                    w.cnt[0] = ts.cnt[0];
                    w.cnt[1] = ts.cnt[1];
                    w.cnt[2] = 0;
                    w.cnt[3] = 0;
                    for (int i = 0; i < 4; ++i) w.stat[i] = ts.stat[i];
                    w.item = item;
                    w.s = X.s;
                    w.p = X.p;
                    w.q = X.q;
                    w.base = base;
                    w.j = j;
                    w.fs = fs;
                    w.fe = fe;
                    w.gap = (int32_t)(gap > INF32 ? INF32 : (gap < -INF32 ? -INF32 : gap));
                    w.smoved = smoved ? 1 : 0;
                    w.phase = 0;
                    publish_parked(P, ri);
                }
                team_sync<TW>();
                return true;
            }
        }
    }
    if (overflow) {
        if (leader) atomicOr(P.err, SBC_KERR_LEVELS);
        return false;
    }
    team_sync<TW>();
    if (tp && leader) tp[1] = (int64_t)wall_clock64();
    const int nf = j; // frontiers 0..nf-1
    const int nD = ts.cnt[0];
    if (leader) ts.cnt[2] = nD;
    team_sync<TW>();

    // ---------------- ascent
    int uA = nD, uB = nD;
    LevelSets below{0, 0, 0, 0, 0, 0};
    SbcVInfo vr = P.vinfo[X.r];
    const int len_r = visible_len(P, vr, X.q);
    for (int lev = base + nf - 1; lev >= 0; --lev) {
        const int jj = lev - base;
        int dA = 0, dB = 0, sA = 0, sB = 0;
        if (jj >= 0) { dA = lvlD[jj]; dB = lvlD[jj + 1]; sA = lvlS[jj]; sB = lvlS[jj + 1]; }
        const LevelSets here{dA, dB, sA, sB, uA, uB};
        const int nd = dB - dA, ns = sB - sA, nu = uB - uA;
        const int tot = nd + ns + nu;
        if (tot == 0) break;
        int64_t tl0 = 0, tph = 0;
        if (tp && leader) { tl0 = (int64_t)wall_clock64(); tph = tl0; tp[11] += 1; }
        // trace: phases of the level (light pass, hub set-up, lookups, hub
        // counters, hub scans, end of level), levels 0..3 (deeper: 3)
        auto phase = [&](int ph) {
            if (tp && leader) {
                const int64_t x = (int64_t)wall_clock64();
                tp[80 + 6 * (lev < 3 ? lev : 3) + ph] += x - tph;
                tph = x;
            }
        };
        if constexpr (TW > 1) {
            // levels 1 and 2: every vertex one level up is known (the root, its
            // neighbours), so the level's source set can be built when the
            // light vertices have many edges to scan
            SrcSet ss{below, here.dA, here.dB, nullptr, -1};
            if (lev == 1) ss.npc = 1;
            else if (lev == 2) { ss.pcs = P.col + vr.start; ss.npc = len_r; }
            ascent_light_level<TW>(P, X, here, tot, lev, ts, *hs, *ls, twave, my_proc, my_scan, tp, ss);
        } else {
            for (int w = 0; w < tot; ++w) {
                const int32_t k = here.at(X, w);
                const SbcVInfo vi = P.vinfo[k];
                const int len = visible_len(P, vi, X.q);
                ascent_vertex_wave(P, X, k, vi, len, lev, ts);
                my_proc += 1;
                my_scan += (uint32_t)len;
            }
        }
        if constexpr (TW > 1) {
            __syncthreads();
            if (tp && leader) {
                tp[8] += (int64_t)wall_clock64() - tl0;
                if (lev == 2) tp[70] += (int64_t)wall_clock64() - tl0;
            }
            phase(0);
            const int nall = hubs_queued(P, *hs);
            for (int hb = 0; hb < nall; hb += hub_batch(P)) {
                const int nh = min(nall - hb, hub_batch(P));
                load_hub_batch<TW>(P, *hs, hb, nh);
                int64_t tl1 = 0;
                if (tp && leader) tl1 = (int64_t)wall_clock64();
                // split the batch: ascent-only indexed hubs on levels 0-2 (and 3)
                // whose lists are much longer than the lookup work are
                // handled together by table lookups; the rest are scanned
                if (threadIdx.x == 0) hs->np = 0;
                __syncthreads();
                ProbeSources src;
                src.below = below;
                src.dA = here.dA;
                src.dB = here.dB;
                src.pc = nullptr;
                src.nr = 0;
                if (lev == 2) {
                    src.pc = P.col + vr.start;
                    src.nr = len_r;
                } else if (lev == 3 && P.hub && P.bucket) {
                    // the root's two-hop set, computed once while the edges
                    // that the items add leave it as it is
                    if (!hs->two_ok) {
                        two_hop<TW>(P, X, vr.start, len_r, *hs, *ls);
                        if (threadIdx.x == 0) hs->two_ok = 1;
                    }
                    if (hs->ntwo >= 0) {
                        src.pc = hs->two;
                        src.nr = hs->ntwo;
                    }
                }
                const long cand = (long)src.size();
                // one thread per hub: its descriptor, table count and state (the
                // first pass's were kept in LDS by the light pass; later passes
                // load them in one round trip); its visible length and first
                // slot kept for the lookups, the counters and the scans
                for (int h = threadIdx.x; h < nh; h += TW * 64) {
                    const int32_t k = hs->list[h];
                    int len, hy;
                    bool asc;
                    if (hb == 0) {
                        len = hs->hlen[h];
                        hy = hs->hidx[h];
                        asc = hs->hasc[h] != 0;
                    } else {
                        const SbcVInfo vi = P.vinfo[k];
                        hy = P.hub ? P.hub[k].y : 0;
                        asc = sb_type(ld_sb(&X.sb[k])) == 3;
                        len = visible_len_thread(P, vi, X.q);
                        hs->hlen[h] = len;
                        hs->hstart[h] = vi.start;
                    }
                    bool pk = false;
                    if (P.hub && P.bucket && lev >= SBC_LOOKUP_MIN_LEV && (lev <= 2 || (lev == 3 && src.nr > 0)) && hy > 0 &&
                        asc)
                        pk = cand < (long)hy;
                    atomicAdd(&ts.stat[0], 1ull);
                    atomicAdd(&ts.stat[1], (unsigned long long)len);
                    int hi = -1;
                    if (pk) {
                        hi = atomicAdd(&hs->np, 1);
                        hs->plist[hi] = k;
                        hs->plen[hi] = len;
                        hs->pstart[hi] = hs->hstart[h];
                        hs->kind[hi] = 0;
                    }
                    hs->pidx[h] = hi;
                }
                __syncthreads();
                const int np = hs->np;
                phase(1);
                if (np > 0) ascent_hubs_probe<TW>(P, X, np, lev, ts, *hs, *ls, src, tp);
                phase(2);
                if (tp && leader) {
                    const int64_t t = (int64_t)wall_clock64();
                    tp[9] += t - tl1;
                    tp[32 + (lev < 7 ? lev : 7)] += t - tl1;
                    if (np > 0) tp[56 + (lev < 7 ? lev : 7)] += cand;
                    tl1 = t;
                }
                // counters of every hub of the batch, one thread per hub; hubs
                // not resolved by lookups (not in plist, or their bucket
                // overflowed: kind 2) are staged in LDS with their state for
                // the scans
                if (threadIdx.x == 0) hs->ns = 0;
                __syncthreads();
                for (int h = threadIdx.x; h < nh; h += TW * 64) {
                    const int32_t k = hs->list[h];
                    const int hi = hs->pidx[h];
                    const bool looked_up = hi >= 0 && hs->kind[hi] != 2;
                    const int len = hs->hlen[h];
                    if (tp) {
                        atomicAdd(&ts.tr[looked_up ? 0 : 1], 1ull);
                        atomicAdd(&ts.tr[looked_up ? 2 : 3], (unsigned long long)len);
                    }
                    if (!looked_up) {
                        const int j = atomicAdd(&hs->ns, 1);
                        if (P.park) atomicAdd((unsigned long long*)&hs->svol, (unsigned long long)len);
                        const KState kst = kstate(P, X, k);
                        ls->k[j] = k;
                        ls->start[j] = hs->hstart[h];
                        ls->len[j] = len;
                        ls->isD[j] = kst.isD ? 1 : 0;
                        ls->sk_old[j] = kst.sk_old;
                        ls->sk_new[j] = kst.sk_new;
                        ls->base[j] = kst.base;
                    }
                }
                __syncthreads();
                phase(3);
                if (hb == 0 && P.park && hs->svol >= P.park_scan && hs->n <= hub_batch(P) + (P.hubq ? SBC_HUBQ_CAP : 0)) {
                    // the pass's hubs left to scan are too much for one workgroup:
                    // take back the pass's hub counts (the wide kernels count the
                    // level's hubs) and hand the item over
                    for (int h = threadIdx.x; h < nh; h += TW * 64) {
                        atomicAdd(&ts.stat[0], ~0ull);
                        atomicAdd(&ts.stat[1], (unsigned long long)(-(long long)hs->hlen[h]));
                    }
                    if (leader) hs->handover = 1;
                    break;
                }
                const int nsh = hs->ns;
                if (P.bucket) {
                    if (nsh > 0) ascent_hubs_flat<TW>(P, X, nsh, lev, ts, *hs, *ls, twave);
                } else {
                    for (int j = 0; j < nsh; ++j) {
                        const KState kst{ls->isD[j] != 0, ls->sk_old[j], ls->sk_new[j], ls->base[j]};
                        const SbcVInfo vi{ls->start[j], 0, 0, 0};
                        ascent_vertex_team<TW>(P, X, ls->k[j], vi, ls->len[j], kst, lev, ts, *hs, twave);
                    }
                }
                __syncthreads();
                if (tp && leader) {
                    const int64_t t = (int64_t)wall_clock64();
                    tp[10] += t - tl1;
                    tp[40 + (lev < 7 ? lev : 7)] += t - tl1;
                    tp[48 + (lev < 7 ? lev : 7)] += nsh;
                }
                phase(4);
            }
            if (leader) { hs->n = 0; hs->svol = 0; }
            __syncthreads();
            if (hs->handover) {
                park_in_ascent<TW>(P, ri, item, X.s, X.p, X.q, ts, base, nf, gap, smoved, lev, uA, uB, my_proc,
                                   my_scan, my_moved, leader, lane);
                return true;
            }
        }
        // a deletion's p lost its child s: it is a parent to update even if
        // no scan of this level reaches it (its pull starts from asc_base)
        if (P.del && lev == base && leader) mark_parent(X, ts, X.p);
        team_sync<TW>();
        if (tp && leader) {
            const int ls7 = lev < 7 ? lev : 7;
            tp[16 + ls7] += (int64_t)wall_clock64() - tl0;
            if (lev == 2) tp[71] += (int64_t)wall_clock64() - tl0;
        }
        phase(5);
        below = here;
        uA = uB;
        uB = ts.cnt[2];
        team_sync<TW>();
    }

    // ---------------- commit
    if (tp && leader) tp[2] = (int64_t)wall_clock64();
    const int nUend = ts.cnt[2];
    const int nS = ts.cnt[1];
    const int total = nUend + nS;
    // descent vertices, frontier by frontier (new level base + jf)
    for (int jf = 0; jf < nf; ++jf) {
        const int32_t nl = base + jf;
        for (int w = lvlD[jf] + twave * 64 + lane; w < lvlD[jf + 1]; w += TW * 64) {
            const int32_t k = X.list[w];
            const int ktype = sb_type(ld_sb(&X.sb[k]));
            const double nd = ld_accd(&X.vr[k].ndel);
            const double od = X.vr[k].delta;
            const int64_t so = X.vr[k].sigma;
            X.vr[k].sigma = (int64_t)((uint64_t)ld_acq64(&X.vr[k].dpath) + (uint64_t)(ktype == 1 ? so : 0));
            X.vr[k].delta = nd;
            X.dist[k] = nl;
            st_sb(&X.sb[k], sb_make(nl, 0));
            if (P.bc_atomic && k != X.r) atomicAdd(&P.bc[k], nd - od);
            st_rel64(&X.vr[k].dpath, 0);
            st_reld(&X.vr[k].ndel, 0.0);
        }
    }
    // ascent-only vertices keep their level
    for (int w = twave * 64 + lane; w < total - nD; w += TW * 64) {
        const int32_t k = w < nUend - nD ? X.list[nD + w] : X.list[nv - 1 - (w - (nUend - nD))];
        const uint8_t b = ld_sb(&X.sb[k]);
        const double nd = ld_accd(&X.vr[k].ndel);
        const double od = X.vr[k].delta;
        X.vr[k].delta = nd;
        st_sb(&X.sb[k], (uint8_t)(b & ~3));
        if (P.bc_atomic && k != X.r) atomicAdd(&P.bc[k], nd - od);
        st_reld(&X.vr[k].ndel, 0.0);
    }
    if (lane == 0) {
        atomicAdd(&ts.stat[0], my_proc);
        atomicAdd(&ts.stat[1], my_scan);
        atomicAdd(&ts.stat[2], my_moved);
    }
    team_sync<TW>();
    if (tp && leader) {
        tp[3] = (int64_t)wall_clock64();
        for (int i = 0; i < 4; ++i) tp[4 + i] = (int64_t)ts.tr[i];
        for (int i = 4; i < 12; ++i) tp[100 + i] = (int64_t)ts.tr[i];
    }
    if (rec && leader) {
        rec->root_index = ri;
        rec->insertion = item;
        rec->kase = smoved ? 2 : 1;
        rec->gap = (int32_t)(gap > INF32 ? INF32 : (gap < -INF32 ? -INF32 : gap));
        rec->processed = (int64_t)ts.stat[0];
        rec->scanned = (int64_t)ts.stat[1];
        rec->affected = total;
        rec->moved = (int64_t)ts.stat[2];
    }
    if (P.wtot && leader) add_work_totals(P.wtot + (size_t)ri * 8, ts.stat, total);
    team_sync<TW>();
    return false;
}


// Classifies window items [i0, i0 + PREF_CAP) of root ri for its team: one
// thread per item reads the endpoints and their levels (what process_item
// reads first), so items of gap 0 are passed over without a round trip
// each.  Whole workgroup; the first barrier lets every thread finish
// reading the previous pass.
template <int TW>
__device__ __forceinline__ void classify_items(const SbcParams& P, int ri, int i0, int W, HubShared& hs)
{
    __syncthreads();
    const int n = min(W - i0, PREF_CAP);
    const int32_t* dist = P.dist + (size_t)ri * (size_t)P.nv;
    static_assert(PREF_CAP <= TW * 64, "one thread per item");
    const int t = threadIdx.x;
    bool zero = false, low = false;
    if (t < n) {
        const int32_t u = P.wu[i0 + t], v = P.wv[i0 + t];
        const int32_t du = dist[u], dv = dist[v];
        hs.pu[t] = u;
        hs.pv[t] = v;
        hs.pdu[t] = du;
        hs.pdv[t] = dv;
        zero = (du == INF32 && dv == INF32) || du == dv;
        low = du <= 1 || dv <= 1;
    }
    if (t < PREF_CAP) {
        const uint64_t zb = ballot(zero), lb = ballot(low);
        if (lane_id() == 0) {
            hs.pzero[2 * (t >> 6)] = (uint32_t)zb;
            hs.pzero[2 * (t >> 6) + 1] = (uint32_t)(zb >> 32);
            hs.plow[2 * (t >> 6)] = (uint32_t)lb;
            hs.plow[2 * (t >> 6) + 1] = (uint32_t)(lb >> 32);
        }
    }
    if (t == 0) { hs.pf0 = i0; hs.pf1 = i0 + n; }
    __syncthreads();
}

// Persistent window kernel.  Phase 1: whole-workgroup teams pop root slots
// [0, heavy) of the queue order; phase 2: each wavefront pops slots
// [heavy, K).  A team applies every insertion of the window to its root in
// stream order before popping the next root.
template <int NWB, bool TRACE = false>
__global__ __launch_bounds__(NWB * 64) void sbc_window_kernel(SbcParams P)
{
    __shared__ TeamShared ts[NWB];
    __shared__ HubShared hs;
    __shared__ LightShared ls;
    const int wave = threadIdx.x >> 6;
    const int W = P.win ? P.win->W : 1;
    // phase 1: heavy
    for (;;) {
        if (threadIdx.x == 0) ts[0].cnt[3] = atomicAdd(&P.qctr[0], 1);
        if (threadIdx.x == 0) hs.two_ok = 0;
        __syncthreads();
        const int t = ts[0].cnt[3];
        __syncthreads();
        if (t >= P.heavy || t >= P.qlen) break;
        const int ri = P.order ? P.order[t] : t;
        const int i0 = P.start_item ? P.start_item[ri] : 0;
        if constexpr (NWB == 16 && !TRACE) {
            if (P.win) {
                // items classified a pass at a time; gap-0 items cost no
                // round trip; a pass holds until an item moves vertices
                bool fresh = false;
                for (int i = i0; i < W; ++i) {
                    if (!fresh || i >= hs.pf1) {
                        classify_items<NWB>(P, ri, i, W, hs);
                        fresh = true;
                    }
                    const int b = i - hs.pf0;
                    if ((hs.pzero[b >> 5] >> (b & 31)) & 1u) {
                        if (threadIdx.x == 0) {
                            if ((hs.plow[b >> 5] >> (b & 31)) & 1u) hs.two_ok = 0;
                            if (P.rec) {
                                SbcItemRec* rec = &P.rec[(size_t)i * P.K + ri];
                                rec->root_index = ri; rec->insertion = i; rec->kase = 0; rec->gap = 0;
                                rec->processed = 0; rec->scanned = 0; rec->affected = 0; rec->moved = 0;
                            }
                        }
                        continue;
                    }
                    if (process_item<NWB, TRACE>(P, ri, i, ts[0], &hs, &ls, wave, hs.pu[b], hs.pv[b], hs.pdu[b],
                                                 hs.pdv[b]))
                        break;
                    if (ts[0].stat[2] != 0) fresh = false;
                    __syncthreads();
                }
                continue;
            }
        }
        for (int i = i0; i < W; ++i)
            if (process_item<NWB, TRACE>(P, ri, i, ts[0], &hs, &ls, wave)) break;
    }
    // phase 2: light, one wavefront per root
    for (;;) {
        if (lane_id() == 0) ts[wave].cnt[3] = atomicAdd(&P.qctr[1], 1);
        team_sync<1>();
        const int t = P.heavy + ts[wave].cnt[3];
        team_sync<1>();
        if (t >= P.qlen) break;
        const int ri = P.order ? P.order[t] : t;
        for (int i = P.start_item ? P.start_item[ri] : 0; i < W; ++i)
            process_item<1, TRACE>(P, ri, i, ts[wave], nullptr, nullptr, 0);
    }
}

// ------------------------------------------------------------ parked items

// Chunks of a vertex's list whose loads one wavefront of the wide kernels
// issues together (their registers do not count against the window kernel's)
#ifndef SBC_WIDE_UNROLL
#define SBC_WIDE_UNROLL 8
#endif
constexpr int WIDE_UNROLL = SBC_WIDE_UNROLL;
//
// A parked item is finished by kernels that span the whole GPU, two launches
// per BFS level; the launch boundaries order the levels and make each
// level's writes visible to the next.  Within a level every wavefront takes
// vertices of the level in a grid stride and runs the same per-vertex code
// as a team (marks by CAS, added paths by atomic adds, appends through the
// item's global list counters, delta summed in adjacency order), so the
// result is bit-identical to the team's.

__device__ __forceinline__ Item parked_item(const SbcParams& P, const SbcWideItem& w, int ri)
{
    const size_t off = (size_t)ri * (size_t)P.nv;
    Item X;
    X.dist = P.dist + off;
    X.sb = P.sb + (size_t)ri * (size_t)P.sb_stride;
    X.vr = P.vr + off;
    X.list = P.list + off;
    X.nv = P.nv;
    X.q = w.q;
    X.s = w.s;
    X.p = w.p;
    X.r = P.roots[ri];
    return X;
}

// A deletion's p, marked as a parent on level lev_parent once the parked
// item's level of s is done, and appended to the item's ascent list.
__global__ void sbc_wide_mark_kernel(SbcParams P, SbcWideItem* wi, int ri, int32_t x, int lev_parent)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const Item X = parked_item(P, *wi, ri);
    const uint8_t b = ld_sb(&X.sb[x]);
    if (sb_type(b) == 0 && cas_sb(&X.sb[x], b, sb_make(lev_parent, 3)) == b)
        X.list[atomicAdd(&wi->cnt[2], 1)] = x;
}

// A level's vertices with at most WIDE_LIGHT visible neighbours are handled
// one per lane, the lane walking its list in order (so an ascent vertex's
// sum keeps the list order); longer ones go to a list for the level's second
// launch: a descent vertex's list in work groups of WIDE_GROUP_CH chunks,
// each taken by any wavefront (marks and added paths are exact in any
// order); an ascent vertex one per wavefront, or, from WIDE_GIANT
// neighbours on, one per workgroup whose wavefronts compute the terms while
// one lane adds them in list order.
#ifndef SBC_WIDE_LIGHT
#define SBC_WIDE_LIGHT 16
#endif
#ifndef SBC_WIDE_GIANT
#define SBC_WIDE_GIANT 4096
#endif
constexpr int WIDE_LIGHT = SBC_WIDE_LIGHT;
constexpr int WIDE_GIANT = SBC_WIDE_GIANT;
constexpr int WIDE_GROUP_CH = WIDE_UNROLL;  // one unrolled step of descent_chunks
constexpr int WIDE_LU = 4;                  // neighbours whose loads a lane issues together
static_assert(64 * WIDE_GROUP_CH == 512, "a frontier's work groups fit nv + slots / 512 + 1");

__device__ __forceinline__ void wide_add_stats(SbcWideItem& w, unsigned long long a, unsigned long long b,
                                               unsigned long long c)
{
    const int64_t sa = wave_sum64((int64_t)a), sb = wave_sum64((int64_t)b), sc = wave_sum64((int64_t)c);
    if (lane_id() == 0) {
        if (sa) atomicAdd(&w.stat[0], (unsigned long long)sa);
        if (sb) atomicAdd(&w.stat[1], (unsigned long long)sb);
        if (sc) atomicAdd(&w.stat[2], (unsigned long long)sc);
    }
}

// Appends of one workgroup pass, staged in LDS and moved to the item's lists
// with one global atomic per list: the next frontier from the bottom of
// buf, same-level vertices from the top.  Capacity: a pass appends each
// neighbour at most once.
template <int CAP>
struct WideStage {
    int cnt[4];   // cnt[0]: frontier entries, cnt[1]: same-level entries (the roles of a team's counters)
    int base[2];
    int32_t buf[CAP];
};

// Moves a pass's staged appends to the item's lists and clears the
// counters.  Whole workgroup; ends with a barrier.
template <int NT, int CAP>
__device__ __forceinline__ void wide_flush(const Item& X, SbcWideItem& w, WideStage<CAP>& st)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        st.base[0] = st.cnt[0] ? atomicAdd(&w.cnt[0], st.cnt[0]) : 0;
        st.base[1] = st.cnt[1] ? atomicAdd(&w.cnt[1], st.cnt[1]) : 0;
    }
    __syncthreads();
    const int nd = st.cnt[0], ns = st.cnt[1];
    for (int j = threadIdx.x; j < nd; j += NT) X.list[st.base[0] + j] = st.buf[j];
    for (int j = threadIdx.x; j < ns; j += NT) X.list[X.nv - 1 - (st.base[1] + j)] = st.buf[CAP - 1 - j];
    __syncthreads();
    if (threadIdx.x == 0) { st.cnt[0] = 0; st.cnt[1] = 0; }
    __syncthreads();
}

// Descent of frontier list[fs, fe) (new level lc of the children), first
// launch: every vertex counted, light ones done, heavy ones listed.  The
// workgroup takes 64 * NW frontier vertices a pass.
template <int NW>
__global__ __launch_bounds__(NW * 64) void sbc_wide_descent_light_kernel(SbcParams P, SbcWideItem* wi, int ri, int fs,
                                                                         int fe, int32_t lc, SbcWideLists L)
{
    constexpr int CAP = NW * 64 * WIDE_LIGHT;
    __shared__ WideStage<CAP> st;
    SbcWideItem& w = *wi;
    const Item X = parked_item(P, w, ri);
    const int lane = lane_id();
    const int n = fe - fs;
    if (threadIdx.x == 0) { st.cnt[0] = 0; st.cnt[1] = 0; }
    __syncthreads();
    unsigned long long my_proc = 0, my_scan = 0, my_moved = 0;
    for (int p0 = (int)blockIdx.x * NW * 64; p0 < n; p0 += (int)gridDim.x * NW * 64) {
        const int i = p0 + wave_in_block() * 64 + lane;
        const bool valid = i < n;
        const int32_t c = valid ? X.list[fs + i] : 0;
        SbcVInfo vi{0, 0, 0, 0};
        if (valid) vi = P.vinfo[c];
        const int len = valid ? visible_len_thread(P, vi, X.q) : 0;
        const bool cmoved = valid && sb_type(ld_sb(&X.sb[c])) == 2;
        const int64_t dc = valid ? ld_acq64(&X.vr[c].dpath) : 0;
        const int32_t oc = valid ? X.dist[c] : 0;
        if (valid) {
            my_proc += 1;
            my_scan += (unsigned long long)len;
            my_moved += cmoved ? 1 : 0;
        }
        const bool light = valid && len <= WIDE_LIGHT;
        const bool heavy = valid && !light;
        const int ng = heavy ? (len + 64 * WIDE_GROUP_CH - 1) / (64 * WIDE_GROUP_CH) : 0;
        const int hi = wave_append(heavy, &L.ctr[0]);
        const int g0 = wave_reserve(ng, &L.ctr[1]);
        if (heavy) {
            if (g0 + ng <= L.aux_cap) {
                L.heavy[2 * hi] = c;
                L.heavy[2 * hi + 1] = g0;
                for (int g = 0; g < ng; ++g) L.aux[g0 + g] = hi;
            } else {
                atomicOr(P.err, SBC_KERR_WIDE);
            }
        }
        const bool c_is_s = c == X.s;
        const int m = __builtin_amdgcn_readfirstlane(wave_max32(light ? len : 0));
        for (int j0 = 0; j0 < m; j0 += WIDE_LU) {
            int32_t kk[WIDE_LU];
            bool ok[WIDE_LU];
            uint8_t bb[WIDE_LU];
#pragma unroll
            for (int u = 0; u < WIDE_LU; ++u) {
                ok[u] = light && j0 + u < len;
                kk[u] = ok[u] ? P.col[vi.start + j0 + u] : 0;
                // the inserted edge never adds a child of s
                if (ok[u] && c_is_s && kk[u] == X.p) ok[u] = false;
            }
#pragma unroll
            for (int u = 0; u < WIDE_LU; ++u) bb[u] = ok[u] ? ld_sb(&X.sb[kk[u]]) : (uint8_t)0;
#pragma unroll
            for (int u = 0; u < WIDE_LU; ++u) {
                bool appD = false, appS = false;
                if (ok[u]) descent_edge(X, kk[u], bb[u], oc, lc, dc, cmoved, appD, appS);
                const int pd = wave_append(appD, &st.cnt[0]);
                if (pd >= 0) st.buf[pd] = kk[u];
                const int psl = wave_append(appS, &st.cnt[1]);
                if (psl >= 0) st.buf[CAP - 1 - psl] = kk[u];
            }
        }
        wide_flush<NW * 64>(X, w, st);
    }
    wide_add_stats(w, my_proc, my_scan, my_moved);
}

// Second launch: the heavy vertices' work groups, one per wavefront, NW a
// workgroup pass; descent_chunks appends into the pass's LDS stage (an Item
// whose list is the stage).
template <int NW>
__global__ __launch_bounds__(NW * 64) void sbc_wide_descent_heavy_kernel(SbcParams P, SbcWideItem* wi, int ri,
                                                                         int32_t lc, SbcWideLists L)
{
    constexpr int CAP = NW * 64 * WIDE_GROUP_CH;
    __shared__ WideStage<CAP> st;
    SbcWideItem& w = *wi;
    const Item X = parked_item(P, w, ri);
    Item XS = X;
    XS.list = st.buf;
    XS.nv = CAP;
    const int ngr = L.ctr[1];
    if (ngr > L.aux_cap) return;  // reported by the first launch
    if (threadIdx.x == 0) { st.cnt[0] = 0; st.cnt[1] = 0; }
    __syncthreads();
    for (int p0 = (int)blockIdx.x * NW; p0 < ngr; p0 += (int)gridDim.x * NW) {
        const int g = p0 + wave_in_block();
        if (g < ngr) {
            const int hi = L.aux[g];
            const int32_t c = L.heavy[2 * hi];
            const int g0 = L.heavy[2 * hi + 1];
            const SbcVInfo vi = P.vinfo[c];
            const int len = visible_len(P, vi, X.q);
            const bool cmoved = sb_type(ld_sb(&X.sb[c])) == 2;
            const int64_t dc = ld_acq64(&X.vr[c].dpath);
            const int cb = (g - g0) * WIDE_GROUP_CH;
            const int ce = min(cb + WIDE_GROUP_CH, (len + 63) >> 6);
            descent_chunks<WideStage<CAP>, WIDE_UNROLL>(P, XS, c, vi, len, cb, ce, X.dist[c], lc, dc, cmoved, st);
        }
        wide_flush<NW * 64>(X, w, st);
    }
}

// Ascent of level lev (the processed vertices here: descent frontier,
// same-level set, ascent set), first launch: every vertex counted (only
// those past HUB_MIN with hubs_only: the team counted the level's light
// vertices), light ones done, heavy and giant ones listed.
template <int NW>
__global__ __launch_bounds__(NW * 64) void sbc_wide_ascent_light_kernel(SbcParams P, SbcWideItem* wi, int ri, int dA,
                                                                        int dB, int sA, int sB, int uA, int uB, int lev,
                                                                        int hubs_only, SbcWideLists L)
{
    constexpr int CAP = NW * 64 * WIDE_LIGHT;
    __shared__ WideStage<CAP> st;  // cnt[0]: parents marked in the pass
    SbcWideItem& w = *wi;
    const Item X = parked_item(P, w, ri);
    const LevelSets here{dA, dB, sA, sB, uA, uB};
    const int tot = here.size();
    const int lane = lane_id();
    if (threadIdx.x == 0) st.cnt[0] = 0;
    __syncthreads();
    unsigned long long my_proc = 0, my_scan = 0;
    for (int p0 = (int)blockIdx.x * NW * 64; p0 < tot; p0 += (int)gridDim.x * NW * 64) {
        const int i = p0 + wave_in_block() * 64 + lane;
        const bool valid = i < tot;
        const int32_t k = valid ? here.at(X, i) : 0;
        SbcVInfo vi{0, 0, 0, 0};
        if (valid) vi = P.vinfo[k];
        const int len = valid ? visible_len_thread(P, vi, X.q) : 0;
        if (valid && (!hubs_only || len > HUB_MIN)) {
            my_proc += 1;
            my_scan += (unsigned long long)len;
        }
        const bool light = valid && len <= WIDE_LIGHT;
        const bool giant = valid && !light && len >= (P.wide_giant > 0 ? P.wide_giant : WIDE_GIANT);
        const bool heavy = valid && !light && !giant;
        const int hi = wave_append(heavy, &L.ctr[0]);
        if (hi >= 0) L.heavy[hi] = k;
        const int gi = wave_append(giant, &L.ctr[1]);
        if (gi >= 0) L.aux[gi] = k;
        KState ks{true, 0, 0, 0.0};
        if (light) ks = kstate(P, X, k);
        double acc = ks.base;
        const int m = __builtin_amdgcn_readfirstlane(wave_max32(light ? len : 0));
        for (int j0 = 0; j0 < m; j0 += WIDE_LU) {
            int32_t cc[WIDE_LU];
            bool ok[WIDE_LU];
            uint8_t ww[WIDE_LU];
#pragma unroll
            for (int u = 0; u < WIDE_LU; ++u) {
                ok[u] = light && j0 + u < len;
                cc[u] = ok[u] ? P.col[vi.start + j0 + u] : 0;
            }
#pragma unroll
            for (int u = 0; u < WIDE_LU; ++u) ww[u] = ok[u] ? ld_sb(&X.sb[cc[u]]) : (uint8_t)0;
            Term t[WIDE_LU];
            bool appU[WIDE_LU];
#pragma unroll
            for (int u = 0; u < WIDE_LU; ++u)
                t[u] = ascent_lane(P, X, k, ks.isD, ks.sk_old, ks.sk_new, lev, ok[u], cc[u], ww[u], appU[u]);
#pragma unroll
            for (int u = 0; u < WIDE_LU; ++u) {
                const int pu = wave_append(appU[u], &st.cnt[0]);
                if (pu >= 0) st.buf[pu] = cc[u];
            }
            // the terms in list order: a child's new dependency, then the
            // removal of its old one (the addition of the negated value)
#pragma unroll
            for (int u = 0; u < WIDE_LU; ++u) {
                if (t[u].child) acc = acc + t[u].add;
                if (t[u].hsub) acc = acc + (-t[u].sub);
            }
        }
        if (light) st_reld(&X.vr[k].ndel, acc);
        // the pass's marked parents to the ascent list
        __syncthreads();
        if (threadIdx.x == 0) st.base[0] = st.cnt[0] ? atomicAdd(&w.cnt[2], st.cnt[0]) : 0;
        __syncthreads();
        for (int j = threadIdx.x; j < st.cnt[0]; j += NW * 64) X.list[st.base[0] + j] = st.buf[j];
        __syncthreads();
        if (threadIdx.x == 0) st.cnt[0] = 0;
        __syncthreads();
    }
    wide_add_stats(w, my_proc, my_scan, 0ull);
}

// A giant ascent vertex's list is taken in slabs of GIANT_CH chunks:
// wavefronts 1..GIANT_TW compute one slab's terms (loads, marks, appends)
// and place them in LDS in list order while lane 0 of wavefront 0 adds the
// previous slab's, one after another; two barriers a slab.
#ifndef SBC_GIANT_CHAIN
#define SBC_GIANT_CHAIN 8
#endif
constexpr int GIANT_CHAIN = SBC_GIANT_CHAIN;  // values per register buffer of the adding lane
constexpr int GIANT_TW = 15;
constexpr int GIANT_CPW = 4;                     // chunks per term wavefront per slab
constexpr int GIANT_CH = GIANT_TW * GIANT_CPW;  // 3840 neighbours per slab
static_assert(GIANT_CH <= 64, "a slab's chunk counts are scanned by one wavefront");
struct GiantShared {
    double v[2][GIANT_CH * 128];  // a chunk gives at most two values per neighbour
    int cnt[GIANT_CH];
};

// acc + the terms of a giant vertex's nch 64-neighbour chunks in list order,
// slab by slab: wavefronts 1..GIANT_TW call terms(c0, t) for chunks
// [c0, c0 + GIANT_CPW) (lane l's term of chunk c0 + u in t[u]) and place the
// terms in LDS in list order, while lane 0 of wavefront 0 adds the previous
// slab's.  Whole workgroup of 16 wavefronts; the sum is valid in lane 0 of
// wavefront 0.
template <typename Terms>
__device__ __forceinline__ double giant_sum(double acc, int nch, GiantShared& gs, Terms terms)
{
    const int wave = wave_in_block(), lane = lane_id();
    const int nslab = (nch + GIANT_CH - 1) / GIANT_CH;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    int tprev = 0;  // values of the slab wavefront 0 adds next
    for (int sl = 0; sl <= nslab; ++sl) {
        Term t[GIANT_CPW];
        uint64_t cm[GIANT_CPW], sm[GIANT_CPW];
        if (wave >= 1 && sl < nslab) {
            terms(sl * GIANT_CH + (wave - 1) * GIANT_CPW, t);
#pragma unroll
            for (int u = 0; u < GIANT_CPW; ++u) {
                cm[u] = ballot(t[u].child);
                sm[u] = ballot(t[u].hsub);
                if (lane == 0) gs.cnt[(wave - 1) * GIANT_CPW + u] = __popcll(cm[u]) + __popcll(sm[u]);
            }
        } else if (wave == 0 && sl >= 1 && lane == 0) {
            acc = lds_chain_sum_pipe<GIANT_CHAIN>(acc, gs.v[(sl - 1) & 1], tprev);
        }
        __syncthreads();
        int tcur = 0;
        if (sl < nslab) {
            // chunk offsets within the slab: an exclusive scan of the counts
            const int cn = lane < GIANT_CH ? gs.cnt[lane] : 0;
            int inc = cn;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int y = __shfl_up(inc, d);
                if (lane >= d) inc += y;
            }
            tcur = __shfl(inc, 63);
            if (wave >= 1) {
                double* dst = gs.v[sl & 1];
#pragma unroll
                for (int u = 0; u < GIANT_CPW; ++u) {
                    const int c = (wave - 1) * GIANT_CPW + u;
                    const int off = __shfl(inc - cn, c);
                    int pos = off + __popcll(cm[u] & below) + __popcll(sm[u] & below);
                    if (t[u].child) dst[pos++] = t[u].add;
                    if (t[u].hsub) dst[pos] = -t[u].sub;
                }
            }
        }
        tprev = tcur;
        __syncthreads();
    }
    return acc;
}

__device__ __forceinline__ void giant_ascent(const SbcParams& P, const Item& X, int32_t k, int lev, SbcWideItem& w,
                                             GiantShared& gs)
{
    const int lane = lane_id();
    const SbcVInfo vi = P.vinfo[k];
    const int len = __builtin_amdgcn_readfirstlane(visible_len(P, vi, X.q));
    const KState ks = kstate(P, X, k);
    const int nch = (len + 63) >> 6;
    const double acc = giant_sum(ks.base, nch, gs, [&](int c0, Term (&t)[GIANT_CPW]) {
        int32_t cc[GIANT_CPW];
        bool ok[GIANT_CPW];
        uint8_t ww[GIANT_CPW];
        bool appU[GIANT_CPW];
#pragma unroll
        for (int u = 0; u < GIANT_CPW; ++u) {
            const int e = (c0 + u) * 64 + lane;
            ok[u] = c0 + u < nch && e < len;
            cc[u] = ok[u] ? P.col[vi.start + e] : 0;
        }
#pragma unroll
        for (int u = 0; u < GIANT_CPW; ++u) ww[u] = ok[u] ? ld_sb(&X.sb[cc[u]]) : (uint8_t)0;
#pragma unroll
        for (int u = 0; u < GIANT_CPW; ++u)
            t[u] = ascent_lane(P, X, k, ks.isD, ks.sk_old, ks.sk_new, lev, ok[u], cc[u], ww[u], appU[u]);
#pragma unroll
        for (int u = 0; u < GIANT_CPW; ++u) {
            const int pu = wave_append(appU[u], &w.cnt[2]);
            if (pu >= 0) X.list[pu] = cc[u];
        }
    });
    if (wave_in_block() == 0 && lane == 0) st_reld(&X.vr[k].ndel, acc);
}

// Second launch of an ascent level: giant vertices one per workgroup, then
// heavy ones one per wavefront in a grid stride (the last workgroups first,
// the ones with giants last).
__global__ __launch_bounds__(1024) void sbc_wide_ascent_heavy_kernel(SbcParams P, SbcWideItem* wi, int ri, int lev,
                                                                     SbcWideLists L)
{
    __shared__ GiantShared gs;
    SbcWideItem& w = *wi;
    const Item X = parked_item(P, w, ri);
    const int nh = L.ctr[0], ngi = L.ctr[1];
    for (int gi = (int)blockIdx.x; gi < ngi; gi += (int)gridDim.x) giant_ascent(P, X, L.aux[gi], lev, w, gs);
    const int nwaves = (int)gridDim.x * 16;
    for (int hi = ((int)gridDim.x - 1 - (int)blockIdx.x) * 16 + wave_in_block(); hi < nh; hi += nwaves) {
        const int32_t k = L.heavy[hi];
        const SbcVInfo vi = P.vinfo[k];
        const int len = visible_len(P, vi, X.q);
        ascent_vertex_wave<SbcWideItem, WIDE_UNROLL>(P, X, k, vi, len, lev, w);
    }
}

// Commit of a parked item (the team's commit, over the whole grid), and its
// record.  Descent frontier jf = list[lvlD[jf], lvlD[jf + 1]) gets level
// base + jf; ascent-only vertices keep their level.
template <int NW>
__global__ __launch_bounds__(NW * 64) void sbc_wide_commit_kernel(SbcParams P, SbcWideItem* wi, int ri, int nf, int nD,
                                                                  int nUend, int nS)
{
    SbcWideItem& w = *wi;
    const Item X = parked_item(P, w, ri);
    const int32_t* lvlD = P.lvl + (size_t)ri * 2 * P.lcap;
    const int nthreads = (int)gridDim.x * NW * 64;
    const int gt = (int)blockIdx.x * NW * 64 + (int)threadIdx.x;
    for (int jf = 0; jf < nf; ++jf) {
        const int32_t nl = w.base + jf;
        const int a = lvlD[jf], b = lvlD[jf + 1];
        for (int i = a + gt; i < b; i += nthreads) {
            const int32_t k = X.list[i];
            const int ktype = sb_type(ld_sb(&X.sb[k]));
            const double nd = ld_accd(&X.vr[k].ndel);
            const double od = X.vr[k].delta;
            const int64_t so = X.vr[k].sigma;
            X.vr[k].sigma = (int64_t)((uint64_t)ld_acq64(&X.vr[k].dpath) + (uint64_t)(ktype == 1 ? so : 0));
            X.vr[k].delta = nd;
            X.dist[k] = nl;
            st_sb(&X.sb[k], sb_make(nl, 0));
            if (P.bc_atomic && k != X.r) atomicAdd(&P.bc[k], nd - od);
            st_rel64(&X.vr[k].dpath, 0);
            st_reld(&X.vr[k].ndel, 0.0);
        }
    }
    const int rest = nUend + nS - nD;
    for (int i = gt; i < rest; i += nthreads) {
        const int32_t k = i < nUend - nD ? X.list[nD + i] : X.list[X.nv - 1 - (i - (nUend - nD))];
        const uint8_t b = ld_sb(&X.sb[k]);
        const double nd = ld_accd(&X.vr[k].ndel);
        const double od = X.vr[k].delta;
        X.vr[k].delta = nd;
        st_sb(&X.sb[k], (uint8_t)(b & ~3));
        if (P.bc_atomic && k != X.r) atomicAdd(&P.bc[k], nd - od);
        st_reld(&X.vr[k].ndel, 0.0);
    }
    if (gt == 0 && P.rec) {
        SbcItemRec* rec = &P.rec[(size_t)w.item * P.K + ri];
        rec->root_index = ri;
        rec->insertion = w.item;
        rec->kase = w.smoved ? 2 : 1;
        rec->gap = w.gap;
        rec->processed = (int64_t)w.stat[0];
        rec->scanned = (int64_t)w.stat[1];
        rec->affected = nUend + nS;
        rec->moved = (int64_t)w.stat[2];
    }
    if (gt == 0 && P.wtot) add_work_totals(P.wtot + (size_t)ri * 8, w.stat, nUend + nS);
}

// ------------------------------------------------------------ static pass

// One workgroup per root: BFS computing levels and sigma level by level, then
// each vertex pulls delta from its children in adjacency order.
// (src/streamingbc_init.c:87-183 with the ascent in pull form)
template <int NWB>
__global__ __launch_bounds__(NWB * 64) void sbc_static_kernel(SbcParams P)
{
    __shared__ TeamShared tsh;
    TeamShared& ts = tsh;
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const bool leader = threadIdx.x == 0;
    const int nv = P.nv;
    const int nq = P.sroots ? *P.scnt : P.K;
    for (;;) {
        if (leader) ts.cnt[3] = atomicAdd(&P.qctr[0], 1);
        __syncthreads();
        const int qi = ts.cnt[3];
        __syncthreads();
        if (qi >= nq) break;
        const int ri = P.sroots ? P.sroots[qi] : qi;
        const size_t off = (size_t)ri * (size_t)nv;
        int32_t* dist = P.dist + off;
        uint8_t* sb = P.sb + (size_t)ri * (size_t)P.sb_stride;
        SbcVRec* vr = P.vr + off;
        int32_t* list = P.list + off;
        int32_t* lvl = P.lvl + (size_t)ri * 2 * P.lcap;
        const int32_t r = P.roots[ri];
        for (int x = threadIdx.x; x < nv; x += NWB * 64) {
            st_rel32(&dist[x], INF32);
            sb[x] = SB_INF;
            vr[x].sigma = SIGMA_INF;
            vr[x].delta = 0.0;
        }
        __syncthreads();
        if (leader) {
            st_rel32(&dist[r], 0);
            sb[r] = sb_make(0, 0);
            vr[r].sigma = 1;
            list[0] = r;
            lvl[0] = 0;
            lvl[1] = 1;
            ts.cnt[0] = 1;
        }
        __syncthreads();
        int fs = 0, fe = 1, j = 0;
        bool overflow = false;
        const int32_t unseen = INF32;
        while (fs < fe) {
            const int32_t next = j + 1;
            for (int idx = fs + wave; idx < fe; idx += NWB) {
                const int32_t c = list[idx];
                const int64_t sc = vr[c].sigma;
                const SbcVInfo vi = P.vinfo[c];
                const int len = vi.deg; // the static pass sees the whole graph
                for (int e0 = 0; e0 < len; e0 += 64) {
                    const int e = e0 + lane;
                    bool app = false;
                    int32_t k = 0;
                    if (e < len) {
                        k = P.col[vi.start + e];
                        int32_t wk = ld_acq32(&dist[k]);
                        if (wk == unseen) {
                            const int32_t old = cas32_wg(&dist[k], unseen, next);
                            app = (old == unseen);
                            wk = app ? next : old;
                        }
                        if (wk == next) add64_wg(&vr[k].dpath, sc);
                    }
                    const int pos = wave_append(app, &ts.cnt[0]);
                    if (pos >= 0) list[pos] = k;
                }
            }
            __syncthreads();
            const int nfe = ts.cnt[0];
            __syncthreads();
            for (int x = fe + threadIdx.x; x < nfe; x += NWB * 64) {
                const int32_t k = list[x];
                vr[k].sigma = ld_acq64(&vr[k].dpath);
                sb[k] = sb_make(j + 1, 0);
                st_rel64(&vr[k].dpath, 0);
            }
            j++;
            if (j + 1 >= P.lcap) { overflow = true; break; }
            if (leader) lvl[j + 1] = nfe;
            fs = fe;
            fe = nfe;
            __syncthreads();
        }
        if (overflow) {
            if (leader) atomicOr(P.err, SBC_KERR_LEVELS);
            continue;
        }
        // levels 0..j-1; level L = list[lvl[L] .. lvl[L+1])
        __syncthreads();
        for (int L = j - 1; L >= 0; --L) {
            const int a = lvl[L], b = lvl[L + 1];
            const int32_t childw = L + 1;
            for (int idx = a + wave; idx < b; idx += NWB) {
                const int32_t k = list[idx];
                const int64_t sk = vr[k].sigma;
                const SbcVInfo vi = P.vinfo[k];
                double acc = 0.0;
                for (int e0 = 0; e0 < vi.deg; e0 += 64) {
                    const int e = e0 + lane;
                    bool child = false;
                    double add = 0.0;
                    if (e < vi.deg) {
                        const int32_t c = P.col[vi.start + e];
                        if (ld_acq32(&dist[c]) == childw) {
                            child = true;
                            add = dep(sk, vr[c].sigma, vr[c].delta);
                        }
                    }
                    uint64_t m = ballot(child);
                    while (m) {
                        const int l = __builtin_ctzll(m);
                        m &= m - 1;
                        acc = acc + readlane_d(add, l);
                    }
                }
                if (lane == 0) vr[k].delta = acc;
            }
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------ static pass, LDS level bitmap
//
// The same tree as sbc_static_kernel, for graphs whose vertex set fits an LDS
// bitmap (one bit per vertex).  Two changes of execution, none of arithmetic:
// - the bitmap holds "level <= L" for the level L being expanded (BFS) or
//   ascended; in an undirected graph a neighbour of a level-L vertex sits on
//   level L-1, L or L+1, so a clear bit means "next level" (BFS: the CAS on
//   dist and the added paths) or "child" (ascent), and the five neighbours in
//   six that are neither never touch the per-root arrays in HBM;
// - a wavefront takes 64 vertices of the level at a time, one per lane, and
//   spreads their adjacency entries over its lanes (64 entries a chunk, the
//   owner of each entry found by a binary search over the lanes' degree
//   prefix), so low-degree vertices do not leave lanes idle and 64 vertices'
//   loads are in flight per wavefront instead of one.
// Ascent: a chunk's child terms are packed in lane order into the
// wavefront's LDS row; each owner lane adds its own segment's terms to its
// running sum in order, so every vertex's delta is the sequential sum over
// its children in adjacency order -- the same bits as sbc_static_kernel.
// The order of the frontier lists differs (they are scratch).

// 64-entry chunks whose loads one wavefront issues together
#ifndef SBC_SBM_UNROLL
#define SBC_SBM_UNROLL 4
#endif
constexpr int SBM_UNROLL = SBC_SBM_UNROLL;

struct SbmBatch {
    int32_t v;      // this lane's vertex (0 when past the level's end)
    int32_t base;   // adjacency start of the vertex minus its entries' first index in the batch
    int32_t incl;   // inclusive prefix of the batch's degrees up to this lane
    int32_t excl;   // exclusive prefix
    int64_t sigma;  // the vertex's path count
};

__device__ __forceinline__ SbmBatch sbm_batch(const SbcParams& P, const int32_t* list, const SbcVRec* vr, int idx,
                                              int end)
{
    SbmBatch b;
    const bool have = idx < end;
    b.v = have ? list[idx] : 0;
    int deg = 0, start = 0;
    b.sigma = 0;
    if (have) {
        const SbcVInfo vi = P.vinfo[b.v];
        deg = vi.deg; // the static pass sees the whole graph
        start = vi.start;
        b.sigma = vr[b.v].sigma;
    }
    const int lane = lane_id();
    int incl = deg;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(incl, d);
        if (lane >= d) incl += t;
    }
    b.incl = incl;
    b.excl = incl - deg;
    b.base = start - b.excl;
    return b;
}

// Lane index of the owner of batch entry e (0 <= e < total): the number of
// lanes whose inclusive prefix is at most e.
__device__ __forceinline__ int sbm_owner(int incl, int e)
{
    int c = 0;
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) {
        if (__shfl(incl, c + s - 1) <= e) c += s;
    }
    return c;
}

__device__ __forceinline__ bool sbm_test(const uint32_t* bm, int32_t k) { return (bm[k >> 5] >> (k & 31)) & 1u; }

// this lane's partial sum added to *ctr once per wavefront
__device__ __forceinline__ void sbm_wave_add(unsigned long long* ctr, unsigned long long x)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d);
    if (lane_id() == 0 && x) atomicAdd(ctr, x);
}

// A level is pulled (below) when its unvisited vertices' neighbour entries
// number fewer than this many times its own.
#ifndef SBC_SBM_PULL_FACTOR
#define SBC_SBM_PULL_FACTOR 8
#endif

union SbmRow {
    double d[64];             // ascent: a chunk's child terms in lane order
    unsigned long long s[64]; // pulled BFS level: each lane's vertex's added paths
};

template <int NWB, int BMW>
__global__ __launch_bounds__(NWB * 64) void sbc_static_bm_kernel(SbcParams P)
{
    __shared__ uint32_t bm[BMW];
    __shared__ SbmRow row[NWB];
    __shared__ unsigned long long found[NWB];
    // 0: list end, 1: next batch of the level, 2: next block of vertex ids,
    // 3: popped queue index
    __shared__ int cnt[4];
    // total degree of the level being settled (slot (j+1)&1) and of the one
    // being expanded (slot j&1); the graph's adjacency entries
    __shared__ unsigned long long dsum[2], dall;
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const bool leader = threadIdx.x == 0;
    const int nv = P.nv, nwords = (nv + 31) >> 5;
    const int nq = P.sroots ? *P.scnt : P.K;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    if (leader) dall = 0ull;
    __syncthreads();
    {
        unsigned long long d = 0;
        for (int x = threadIdx.x; x < nv; x += NWB * 64) d += (unsigned long long)P.vinfo[x].deg;
        sbm_wave_add(&dall, d);
    }
    for (;;) {
        if (leader) cnt[3] = atomicAdd(&P.qctr[0], 1);
        __syncthreads();
        const int qi = cnt[3];
        const unsigned long long all_deg = dall;
        __syncthreads();
        if (qi >= nq) break;
        const int ri = P.sroots ? P.sroots[qi] : qi;
        const size_t off = (size_t)ri * (size_t)nv;
        int32_t* dist = P.dist + off;
        uint8_t* sb = P.sb + (size_t)ri * (size_t)P.sb_stride;
        SbcVRec* vr = P.vr + off;
        int32_t* list = P.list + off;
        int32_t* lvl = P.lvl + (size_t)ri * 2 * P.lcap;
        const int32_t r = P.roots[ri];
        for (int x = threadIdx.x; x < nv; x += NWB * 64) {
            st_rel32(&dist[x], INF32);
            sb[x] = SB_INF;
            vr[x].sigma = SIGMA_INF;
            vr[x].delta = 0.0;
        }
        for (int w = threadIdx.x; w < nwords; w += NWB * 64) bm[w] = 0u;
        __syncthreads();
        if (leader) {
            st_rel32(&dist[r], 0);
            sb[r] = sb_make(0, 0);
            vr[r].sigma = 1;
            list[0] = r;
            lvl[0] = 0;
            lvl[1] = 1;
            bm[r >> 5] |= 1u << (r & 31);
            cnt[0] = 1;
            cnt[1] = 0;
            cnt[2] = 0;
            dsum[0] = (unsigned long long)P.vinfo[r].deg;
        }
        __syncthreads();
        // BFS: level j = list[fs, fe); bm = levels <= j.  Three ways to find
        // level j+1, the same levels and path counts from each:
        // - pull, when the unvisited vertices have fewer neighbour entries
        //   than level j: every unvisited vertex (bit clear) scans its list;
        //   a neighbour with its bit set is on level j (one on a lower level
        //   would put the vertex on level <= j), so the vertex is on level
        //   j+1 iff it has one, and its path count is the sum of theirs;
        // - push with stores, when level j has at least nv/8 neighbour
        //   entries: plain stores of the new level and path adds whose
        //   results nobody waits for, then one pass over dist in vertex order
        //   finds and settles the new level;
        // - push with CAS: new vertices are claimed by CAS and appended as
        //   they are found.
        int fs = 0, fe = 1, j = 0;
        bool overflow = false;
        unsigned long long rem = all_deg; // neighbour entries of unvisited vertices
        while (fs < fe) {
            const int32_t next = j + 1;
            const unsigned long long fdeg = dsum[j & 1];
            rem -= fdeg;
            const bool pull = rem < (unsigned long long)SBC_SBM_PULL_FACTOR * fdeg;
            const bool scan = !pull && fdeg * 8ull >= (unsigned long long)nv;
            if (leader) dsum[(j + 1) & 1] = 0ull;
            unsigned long long deg = 0; // this thread's share of level j+1's degree
            if (pull) {
                for (;;) {
                    int x0 = 0;
                    if (lane == 0) x0 = atomicAdd(&cnt[2], 64);
                    x0 = __shfl(x0, 0);
                    if (x0 >= nv) break;
                    const int32_t x = x0 + lane;
                    const bool have = x < nv && !sbm_test(bm, x);
                    int vdeg = 0, start = 0;
                    if (have) {
                        const SbcVInfo vi = P.vinfo[x];
                        vdeg = vi.deg;
                        start = vi.start;
                    }
                    int incl = vdeg;
#pragma unroll
                    for (int d = 1; d < 64; d <<= 1) {
                        const int t = __shfl_up(incl, d);
                        if (lane >= d) incl += t;
                    }
                    const int32_t base = start - (incl - vdeg);
                    const int total = __shfl(incl, 63);
                    if (total == 0) continue;
                    row[wave].s[lane] = 0ull;
                    if (lane == 0) found[wave] = 0ull;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    for (int e0 = 0; e0 < total; e0 += 64 * SBM_UNROLL) {
                        int32_t k[SBM_UNROLL];
                        int o[SBM_UNROLL];
                        bool t[SBM_UNROLL];
#pragma unroll
                        for (int u = 0; u < SBM_UNROLL; ++u) {
                            const int e = e0 + u * 64 + lane;
                            t[u] = e < total;
                            o[u] = sbm_owner(incl, t[u] ? e : total - 1);
                            const int32_t b = __shfl(base, o[u]);
                            k[u] = t[u] ? P.col[b + e] : 0;
                        }
#pragma unroll
                        for (int u = 0; u < SBM_UNROLL; ++u) {
                            if (t[u] && sbm_test(bm, k[u])) {
                                atomicAdd(&row[wave].s[o[u]], (unsigned long long)vr[k[u]].sigma);
                                atomicOr(&found[wave], 1ull << o[u]);
                            }
                        }
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    const bool f = (found[wave] >> lane) & 1ull;
                    if (f) {
                        st_rel32(&dist[x], next);
                        vr[x].sigma = (int64_t)row[wave].s[lane];
                        sb[x] = sb_make(j + 1, 0);
                        deg += (unsigned long long)vdeg;
                    }
                    const int pos = wave_append(f, &cnt[0]);
                    if (pos >= 0) list[pos] = x;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
            } else {
                for (;;) {
                    int b0 = 0;
                    if (lane == 0) b0 = atomicAdd(&cnt[1], 64);
                    b0 = __shfl(b0, 0);
                    if (b0 >= fe) break;
                    const SbmBatch B = sbm_batch(P, list, vr, b0 + lane, fe);
                    const int total = __shfl(B.incl, 63);
                    for (int e0 = 0; e0 < total; e0 += 64 * SBM_UNROLL) {
                        int32_t k[SBM_UNROLL];
                        int64_t sg[SBM_UNROLL];
                        bool t[SBM_UNROLL];
#pragma unroll
                        for (int u = 0; u < SBM_UNROLL; ++u) {
                            const int e = e0 + u * 64 + lane;
                            t[u] = e < total;
                            // every lane takes part in the shuffles
                            const int o = sbm_owner(B.incl, t[u] ? e : total - 1);
                            const int32_t b = __shfl(B.base, o);
                            sg[u] = __shfl(B.sigma, o);
                            k[u] = t[u] ? P.col[b + e] : 0;
                        }
                        if (scan) {
#pragma unroll
                            for (int u = 0; u < SBM_UNROLL; ++u) {
                                if (t[u] && !sbm_test(bm, k[u])) {
                                    // level j+1 or unseen: every writer stores the same level
                                    st_rel32(&dist[k[u]], next);
                                    add64_wg(&vr[k[u]].dpath, sg[u]);
                                }
                            }
                        } else {
#pragma unroll
                            for (int u = 0; u < SBM_UNROLL; ++u) {
                                bool app = false;
                                if (t[u] && !sbm_test(bm, k[u])) {
                                    // level j+1 or unseen: the CAS and the added paths
                                    app = cas32_wg(&dist[k[u]], INF32, next) == INF32;
                                    add64_wg(&vr[k[u]].dpath, sg[u]);
                                }
                                const int pos = wave_append(app, &cnt[0]);
                                if (pos >= 0) list[pos] = k[u];
                            }
                        }
                    }
                }
            }
            __syncthreads();
            if (scan) {
                // the new level in vertex order, settled as it is found
                for (int x0 = wave * 64; x0 < nv; x0 += NWB * 64) {
                    const int32_t x = x0 + lane;
                    const bool m = x < nv && ld_acq32(&dist[x]) == next;
                    if (m) {
                        vr[x].sigma = ld_acq64(&vr[x].dpath);
                        sb[x] = sb_make(j + 1, 0);
                        st_rel64(&vr[x].dpath, 0);
                        atomicOr(&bm[x >> 5], 1u << (x & 31));
                        deg += (unsigned long long)P.vinfo[x].deg;
                    }
                    const int pos = wave_append(m, &cnt[0]);
                    if (pos >= 0) list[pos] = x;
                }
                __syncthreads();
            }
            const int nfe = cnt[0];
            if (pull) {
                for (int x = fe + threadIdx.x; x < nfe; x += NWB * 64) {
                    const int32_t k = list[x];
                    atomicOr(&bm[k >> 5], 1u << (k & 31));
                }
            } else if (!scan) {
                for (int x = fe + threadIdx.x; x < nfe; x += NWB * 64) {
                    const int32_t k = list[x];
                    vr[k].sigma = ld_acq64(&vr[k].dpath);
                    sb[k] = sb_make(j + 1, 0);
                    st_rel64(&vr[k].dpath, 0);
                    atomicOr(&bm[k >> 5], 1u << (k & 31));
                    deg += (unsigned long long)P.vinfo[k].deg;
                }
            }
            sbm_wave_add(&dsum[(j + 1) & 1], deg);
            j++;
            if (j + 1 >= P.lcap) { overflow = true; break; }
            if (leader) {
                lvl[j + 1] = nfe;
                cnt[1] = fe;
                cnt[2] = 0;
            }
            fs = fe;
            fe = nfe;
            __syncthreads();
        }
        if (overflow) {
            if (leader) atomicOr(P.err, SBC_KERR_LEVELS);
            continue;
        }
        // levels 0..j-1, level L = list[lvl[L] .. lvl[L+1]); bm = every level
        for (int L = j - 1; L >= 0; --L) {
            if (L + 1 < j) {
                // bm = levels <= L: take level L+1 out
                const int a1 = lvl[L + 1], b1 = lvl[L + 2];
                for (int x = a1 + threadIdx.x; x < b1; x += NWB * 64) {
                    const int32_t k = list[x];
                    atomicAnd(&bm[k >> 5], ~(1u << (k & 31)));
                }
            }
            const int a = lvl[L], b = lvl[L + 1];
            if (leader) cnt[1] = a;
            __syncthreads();
            for (;;) {
                int b0 = 0;
                if (lane == 0) b0 = atomicAdd(&cnt[1], 64);
                b0 = __shfl(b0, 0);
                if (b0 >= b) break;
                const SbmBatch B = sbm_batch(P, list, vr, b0 + lane, b);
                const int total = __shfl(B.incl, 63);
                double acc = 0.0; // this lane's vertex
                for (int e0 = 0; e0 < total; e0 += 64 * SBM_UNROLL) {
                    int32_t k[SBM_UNROLL];
                    int64_t sk[SBM_UNROLL];
                    bool c[SBM_UNROLL];
#pragma unroll
                    for (int u = 0; u < SBM_UNROLL; ++u) {
                        const int e = e0 + u * 64 + lane;
                        c[u] = e < total;
                        const int o = sbm_owner(B.incl, c[u] ? e : total - 1);
                        const int32_t base = __shfl(B.base, o);
                        sk[u] = __shfl(B.sigma, o);
                        k[u] = c[u] ? P.col[base + e] : 0;
                    }
                    int64_t sc[SBM_UNROLL];
                    double dc[SBM_UNROLL];
#pragma unroll
                    for (int u = 0; u < SBM_UNROLL; ++u) {
                        c[u] = c[u] && !sbm_test(bm, k[u]);
                        sc[u] = 1;
                        dc[u] = 0.0;
                        if (c[u]) {
                            sc[u] = vr[k[u]].sigma;
                            dc[u] = vr[k[u]].delta;
                        }
                    }
#pragma unroll
                    for (int u = 0; u < SBM_UNROLL; ++u) {
                        const int E = e0 + u * 64;
                        if (E >= total) break;
                        const uint64_t cm = ballot(c[u]);
                        if (cm == 0) continue;
                        if (c[u]) row[wave].d[__popcll(cm & below)] = dep(sk[u], sc[u], dc[u]);
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                        // this lane's vertex's entries in the chunk: [s0, s1) - E
                        const int s0 = B.excl > E ? B.excl : E;
                        const int s1 = B.incl < E + 64 ? B.incl : E + 64;
                        if (s0 < s1) {
                            const int ls = s0 - E, le = s1 - E;
                            const uint64_t mlo = ls == 0 ? 0ull : (~0ull >> (64 - ls));
                            const uint64_t mhi = le == 64 ? ~0ull : (~0ull >> (64 - le));
                            const int f = __popcll(cm & mlo), n = __popcll(cm & mhi) - f;
                            acc = lds_chain_sum(acc, &row[wave].d[f], n);
                        }
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    }
                }
                if (b0 + lane < b) vr[B.v].delta = acc;
            }
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------ one root, whole GPU
//
// The static pass of a single root spread over the whole GPU (an edge
// deletion's movement-case roots, when there are few of them): one launch
// per BFS level and per ascent level, wavefronts taking the level's
// vertices in a grid stride, each running the per-vertex code of
// sbc_static_kernel -- the CAS on dist, the added paths, then the pull in
// lane order -- so the tree is bit-identical to that kernel's.  Every
// workgroup writes the root's arrays here, so the atomics use device scope.

// chunks of 64 neighbours whose loads one wavefront issues together
constexpr int ROOT_UNROLL = 8;

__device__ __forceinline__ int32_t cas32_dev(int32_t* p, int32_t expect, int32_t want)
{
    __hip_atomic_compare_exchange_strong(p, &expect, want, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return expect;
}

// Resets root ri's tree to "only the root reached"; fr[0] = 1 (list end).
__global__ void sbc_root_reset_kernel(SbcParams P, int ri, int32_t* fr)
{
    const size_t off = (size_t)ri * (size_t)P.nv;
    const int32_t r = P.roots[ri];
    uint8_t* sb = P.sb + (size_t)ri * (size_t)P.sb_stride;
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < P.nv; x += gridDim.x * blockDim.x) {
        const bool root = x == r;
        P.dist[off + x] = root ? 0 : INF32;
        sb[x] = root ? sb_make(0, 0) : SB_INF;
        P.vr[off + x].sigma = root ? 1 : SIGMA_INF;
        P.vr[off + x].delta = 0.0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        P.list[off] = r;
        fr[0] = 1;
    }
}

// Frontier list[fs, fe) (level j): reaches level j + 1, appending at fr[0].
__global__ void sbc_root_bfs_kernel(SbcParams P, int ri, int fs, int fe, int j, int32_t* fr)
{
    const size_t off = (size_t)ri * (size_t)P.nv;
    int32_t* dist = P.dist + off;
    SbcVRec* vr = P.vr + off;
    int32_t* list = P.list + off;
    const int lane = lane_id();
    const int nw = (gridDim.x * blockDim.x) >> 6;
    const int32_t next = j + 1;
    for (int idx = fs + ((blockIdx.x * blockDim.x + threadIdx.x) >> 6); idx < fe; idx += nw) {
        const int32_t c = list[idx];
        const int64_t sc = vr[c].sigma;
        const SbcVInfo vi = P.vinfo[c];
        // ROOT_UNROLL chunks per round: ids, then levels, then the CASes in
        // flight together; one list reservation for all their appends, then
        // the added paths (no value returned)
        for (int e0 = 0; e0 < vi.deg; e0 += 64 * ROOT_UNROLL) {
            int32_t k[ROOT_UNROLL], wk[ROOT_UNROLL];
            bool app[ROOT_UNROLL];
#pragma unroll
            for (int u = 0; u < ROOT_UNROLL; ++u) {
                const int e = e0 + u * 64 + lane;
                k[u] = e < vi.deg ? P.col[vi.start + e] : -1;
            }
#pragma unroll
            for (int u = 0; u < ROOT_UNROLL; ++u)
                wk[u] = k[u] >= 0 ? __hip_atomic_load(&dist[k[u]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
#pragma unroll
            for (int u = 0; u < ROOT_UNROLL; ++u) {
                app[u] = false;
                if (k[u] >= 0 && wk[u] == INF32) {
                    const int32_t old = cas32_dev(&dist[k[u]], INF32, next);
                    app[u] = old == INF32;
                    wk[u] = app[u] ? next : old;
                }
            }
            uint64_t m[ROOT_UNROLL];
            int tot = 0;
#pragma unroll
            for (int u = 0; u < ROOT_UNROLL; ++u) {
                m[u] = ballot(app[u]);
                tot += __popcll(m[u]);
            }
            if (tot) {
                int base = 0;
                if (lane == 0) base = atomicAdd(&fr[0], tot);
                base = __shfl(base, 0);
#pragma unroll
                for (int u = 0; u < ROOT_UNROLL; ++u) {
                    if (app[u]) list[base + prefix_in_wave(m[u])] = k[u];
                    base += __popcll(m[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < ROOT_UNROLL; ++u)
                if (k[u] >= 0 && wk[u] == next)
                    __hip_atomic_fetch_add((unsigned long long*)&vr[k[u]].dpath, (unsigned long long)sc,
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Level j + 1 = list[a, b): path counts from the added paths, state bytes.
__global__ void sbc_root_settle_kernel(SbcParams P, int ri, int a, int b, int j)
{
    const size_t off = (size_t)ri * (size_t)P.nv;
    uint8_t* sb = P.sb + (size_t)ri * (size_t)P.sb_stride;
    for (int x = a + blockIdx.x * blockDim.x + threadIdx.x; x < b; x += gridDim.x * blockDim.x) {
        const int32_t k = P.list[off + x];
        SbcVRec& r = P.vr[off + k];
        r.sigma = r.dpath;
        r.dpath = 0;
        sb[k] = sb_make(j + 1, 0);
    }
}

// Level L = list[a, b): delta pulled from the children in adjacency order.
// Vertices with more than ROOT_BIG neighbours are left to the two kernels
// below (their positions in the level are appended to big[1..], count big[0]).
constexpr int ROOT_BIG = 16384;

__global__ void sbc_root_ascent_kernel(SbcParams P, int ri, int a, int b, int L, int32_t* big)
{
    const size_t off = (size_t)ri * (size_t)P.nv;
    const int32_t* dist = P.dist + off;
    SbcVRec* vr = P.vr + off;
    const int lane = lane_id();
    const int nw = (gridDim.x * blockDim.x) >> 6;
    const int32_t childw = L + 1;
    for (int idx = a + ((blockIdx.x * blockDim.x + threadIdx.x) >> 6); idx < b; idx += nw) {
        const int32_t k = P.list[off + idx];
        const SbcVInfo vi = P.vinfo[k];
        if (vi.deg > ROOT_BIG) {
            if (lane == 0) big[1 + atomicAdd(&big[0], 1)] = k;
            continue;
        }
        const int64_t sk = vr[k].sigma;
        double acc = 0.0;
        // ROOT_UNROLL chunks per round: ids, levels, then the children's
        // records in flight together; the adds stay in adjacency order
        for (int e0 = 0; e0 < vi.deg; e0 += 64 * ROOT_UNROLL) {
            int32_t c[ROOT_UNROLL];
            bool child[ROOT_UNROLL];
            double add[ROOT_UNROLL];
#pragma unroll
            for (int u = 0; u < ROOT_UNROLL; ++u) {
                const int e = e0 + u * 64 + lane;
                c[u] = e < vi.deg ? P.col[vi.start + e] : -1;
            }
#pragma unroll
            for (int u = 0; u < ROOT_UNROLL; ++u) child[u] = c[u] >= 0 && dist[c[u]] == childw;
#pragma unroll
            for (int u = 0; u < ROOT_UNROLL; ++u) add[u] = child[u] ? dep(sk, vr[c[u]].sigma, vr[c[u]].delta) : 0.0;
#pragma unroll
            for (int u = 0; u < ROOT_UNROLL; ++u) {
                uint64_t m = ballot(child[u]);
                while (m) {
                    const int l = __builtin_ctzll(m);
                    m &= m - 1;
                    acc = acc + readlane_d(add[u], l);
                }
            }
        }
        if (lane == 0) vr[k].delta = acc;
    }
}

// Terms of the big vertices of level L (ids big[1..n], term offsets off[0..n]):
// terms[off[i] + e] = vertex i's term for its e-th neighbour, 0.0 for a
// neighbour that is not a child (adding +0.0 to a non-negative sum that
// starts at +0.0 changes nothing), one thread per term over the whole GPU.
__global__ void sbc_root_terms_kernel(SbcParams P, int ri, int L, const int32_t* big, const int64_t* toff, int n,
                                      double* terms)
{
    const size_t off = (size_t)ri * (size_t)P.nv;
    const int32_t* dist = P.dist + off;
    const SbcVRec* vr = P.vr + off;
    const int64_t T = toff[n];
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < T; t += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = n;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (toff[mid] <= t) lo = mid;
            else hi = mid;
        }
        const int32_t k = big[1 + lo];
        const SbcVInfo vi = P.vinfo[k];
        const int32_t c = P.col[vi.start + (int)(t - toff[lo])];
        terms[t] = dist[c] == L + 1 ? dep(vr[k].sigma, vr[c].sigma, vr[c].delta) : 0.0;
    }
}

// The big vertices' deltas: one wavefront per vertex stages its terms through
// LDS and lane 0 adds them in adjacency order.
__global__ __launch_bounds__(64) void sbc_root_bigsum_kernel(SbcParams P, int ri, const int32_t* big,
                                                             const int64_t* toff, int n, const double* terms)
{
    __shared__ double buf[512];
    const int i = blockIdx.x;
    if (i >= n) return;
    const int lane = threadIdx.x;
    const int64_t t0 = toff[i], t1 = toff[i + 1];
    double acc = 0.0;
    for (int64_t b0 = t0; b0 < t1; b0 += 512) {
        const int m = (int)(t1 - b0 < 512 ? t1 - b0 : 512);
        for (int j = lane; j < m; j += 64) buf[j] = terms[b0 + j];
        __syncthreads();
        if (lane == 0) acc = lds_chain_sum(acc, buf, m);
        __syncthreads();
    }
    if (lane == 0) P.vr[(size_t)ri * (size_t)P.nv + big[1 + i]].delta = acc;
}

// ------------------------------------------------------------ several roots, whole GPU
//
// The same rebuild for n roots at once: every launch takes one level of each
// root, the roots' ranges flattened into one index space by the prefix
// pre[0..n] (segment i: root sg[i].ri, list[sg[i].a, sg[i].b), level
// sg[i].lev), so a level of many small roots fills the GPU as one large level
// does.  Per root the per-vertex code is the single-root kernels', so the
// trees are the same bits; root i appends its next frontier at fr[i].

__device__ __forceinline__ int seg_of(const int64_t* pre, int n, int64_t w)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= w) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void sbc_roots_reset_kernel(SbcParams P, const int32_t* rl, int n, int32_t* fr)
{
    const int64_t tot = (int64_t)n * P.nv;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t / P.nv);
        const int x = (int)(t - (int64_t)i * P.nv);
        const int ri = rl[i];
        const size_t off = (size_t)ri * (size_t)P.nv;
        const int32_t r = P.roots[ri];
        const bool root = x == r;
        P.dist[off + x] = root ? 0 : INF32;
        P.sb[(size_t)ri * (size_t)P.sb_stride + x] = root ? sb_make(0, 0) : SB_INF;
        P.vr[off + x].sigma = root ? 1 : SIGMA_INF;
        P.vr[off + x].delta = 0.0;
        if (x == 0) {
            P.list[off] = r;
            fr[i] = 1;
        }
    }
}

__global__ void sbc_roots_bfs_kernel(SbcParams P, const SbcRootSeg* sg, const int64_t* pre, int n, int32_t* fr)
{
    const int lane = lane_id();
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < pre[n]; w += nw) {
        const int i = seg_of(pre, n, w);
        const SbcRootSeg g = sg[i];
        const size_t off = (size_t)g.ri * (size_t)P.nv;
        int32_t* dist = P.dist + off;
        SbcVRec* vr = P.vr + off;
        int32_t* list = P.list + off;
        const int32_t next = g.lev + 1;
        const int32_t c = list[g.a + (int)(w - pre[i])];
        const int64_t sc = vr[c].sigma;
        const SbcVInfo vi = P.vinfo[c];
        for (int e0 = 0; e0 < vi.deg; e0 += 64 * ROOT_UNROLL) {
            int32_t k[ROOT_UNROLL], wk[ROOT_UNROLL];
            bool app[ROOT_UNROLL];
#pragma unroll
            for (int u = 0; u < ROOT_UNROLL; ++u) {
                const int e = e0 + u * 64 + lane;
                k[u] = e < vi.deg ? P.col[vi.start + e] : -1;
            }
#pragma unroll
            for (int u = 0; u < ROOT_UNROLL; ++u)
                wk[u] = k[u] >= 0 ? __hip_atomic_load(&dist[k[u]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
#pragma unroll
            for (int u = 0; u < ROOT_UNROLL; ++u) {
                app[u] = false;
                if (k[u] >= 0 && wk[u] == INF32) {
                    const int32_t old = cas32_dev(&dist[k[u]], INF32, next);
                    app[u] = old == INF32;
                    wk[u] = app[u] ? next : old;
                }
            }
            uint64_t m[ROOT_UNROLL];
            int tot = 0;
#pragma unroll
            for (int u = 0; u < ROOT_UNROLL; ++u) {
                m[u] = ballot(app[u]);
                tot += __popcll(m[u]);
            }
            if (tot) {
                int base = 0;
                if (lane == 0) base = atomicAdd(&fr[i], tot);
                base = __shfl(base, 0);
#pragma unroll
                for (int u = 0; u < ROOT_UNROLL; ++u) {
                    if (app[u]) list[base + prefix_in_wave(m[u])] = k[u];
                    base += __popcll(m[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < ROOT_UNROLL; ++u)
                if (k[u] >= 0 && wk[u] == next)
                    __hip_atomic_fetch_add((unsigned long long*)&vr[k[u]].dpath, (unsigned long long)sc,
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// new levels: segment i is the new level g.lev + 1 of its root
__global__ void sbc_roots_settle_kernel(SbcParams P, const SbcRootSeg* sg, const int64_t* pre, int n)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < pre[n]; t += (int64_t)gridDim.x * blockDim.x) {
        const int i = seg_of(pre, n, t);
        const SbcRootSeg g = sg[i];
        const size_t off = (size_t)g.ri * (size_t)P.nv;
        const int32_t k = P.list[off + g.a + (int)(t - pre[i])];
        SbcVRec& r = P.vr[off + k];
        r.sigma = r.dpath;
        r.dpath = 0;
        P.sb[(size_t)g.ri * (size_t)P.sb_stride + k] = sb_make(g.lev + 1, 0);
    }
}

__global__ void sbc_roots_ascent_kernel(SbcParams P, const SbcRootSeg* sg, const int64_t* pre, int n)
{
    const int lane = lane_id();
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < pre[n]; w += nw) {
        const int i = seg_of(pre, n, w);
        const SbcRootSeg g = sg[i];
        const size_t off = (size_t)g.ri * (size_t)P.nv;
        const int32_t* dist = P.dist + off;
        SbcVRec* vr = P.vr + off;
        const int32_t childw = g.lev + 1;
        const int32_t k = P.list[off + g.a + (int)(w - pre[i])];
        const int64_t sk = vr[k].sigma;
        const SbcVInfo vi = P.vinfo[k];
        double acc = 0.0;
        for (int e0 = 0; e0 < vi.deg; e0 += 64 * ROOT_UNROLL) {
            int32_t c[ROOT_UNROLL];
            bool child[ROOT_UNROLL];
            double add[ROOT_UNROLL];
#pragma unroll
            for (int u = 0; u < ROOT_UNROLL; ++u) {
                const int e = e0 + u * 64 + lane;
                c[u] = e < vi.deg ? P.col[vi.start + e] : -1;
            }
#pragma unroll
            for (int u = 0; u < ROOT_UNROLL; ++u) child[u] = c[u] >= 0 && dist[c[u]] == childw;
#pragma unroll
            for (int u = 0; u < ROOT_UNROLL; ++u) add[u] = child[u] ? dep(sk, vr[c[u]].sigma, vr[c[u]].delta) : 0.0;
#pragma unroll
            for (int u = 0; u < ROOT_UNROLL; ++u) {
                uint64_t m = ballot(child[u]);
                while (m) {
                    const int l = __builtin_ctzll(m);
                    m &= m - 1;
                    acc = acc + readlane_d(add[u], l);
                }
            }
        }
        if (lane == 0) vr[k].delta = acc;
    }
}

// BC[v] += sign * (sum over the listed roots list[0, *cnt) of delta_r[v],
// r != v): an edge deletion takes the roots it rebuilds out of BC before the
// rebuild (sign -1) and puts them back after (+1).
__global__ void sbc_bc_roots_kernel(SbcParams P, const int32_t* list, const int32_t* cnt, double sign)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= P.nv) return;
    const int n = *cnt;
    double acc = 0.0;
    for (int i = 0; i < n; ++i) {
        const int ri = list[i];
        if (P.roots[ri] == v) continue;
        acc += P.vr[(size_t)ri * (size_t)P.nv + v].delta;
    }
    if (acc != 0.0) P.bc[v] = P.bc[v] + sign * acc;
}

// BC[v] = sum over roots, in root-array order, of delta_r[v] (r != v).
__global__ void sbc_bc_reduce_kernel(SbcParams P)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= P.nv) return;
    double acc = 0.0;
    for (int ri = 0; ri < P.K; ++ri) {
        if (P.roots[ri] == v) continue;
        acc += P.vr[(size_t)ri * (size_t)P.nv + v].delta;
    }
    P.bc[v] = 0.0 + acc;
}

// Edge deletion, one thread per root, after (u, v) left the graph: the
// reference's per-root dispatch (src/streamingbc.c:310-345).  Endpoints on
// the same level (or both unreachable): the edge is on no shortest path and
// the tree stands.  Otherwise the deeper endpoint s keeps its level if
// another neighbour lies one level up (adjacent case: dist unchanged, the
// update kernel runs with del set) or moves down (movement case: the static
// pass rebuilds the tree).
__global__ void sbc_delete_classify_kernel(SbcParams P, int32_t u, int32_t v, int32_t* out)
{
    const int ri = blockIdx.x * blockDim.x + threadIdx.x;
    if (ri >= P.K) return;
    const int32_t* dist = P.dist + (size_t)ri * (size_t)P.nv;
    const int32_t du = dist[u], dv = dist[v];
    if (du == dv) {
        atomicAdd(&out[1], 1);
        return;
    }
    const int32_t s = du > dv ? u : v, p = du > dv ? v : u;
    const int32_t up = dist[p];
    const SbcVInfo vi = P.vinfo[s];
    int extra = 0;
    for (int e = 0; e < vi.deg && !extra; ++e)
        if (dist[P.col[vi.start + e]] == up) extra = 1;
    if (extra) {
        out[4 + atomicAdd(&out[2], 1)] = ri;
    } else {
        out[4 + P.K + atomicAdd(&out[3], 1)] = ri;
    }
}

// ------------------------------------------------------------ deletion, movement case
//
// The movement-case roots of deleting (u0, v0) (already out of the graph),
// one workgroup each (root indices sroots[0, *scnt)), updated in place to
// the tree the static pass would build on the new graph, bit for bit:
// levels and path counts are exact, and a vertex's delta is recomputed --
// by the static pass's own sum over its children in adjacency order --
// exactly when one of its inputs (its level, its path count, its children,
// their path counts or their deltas) changed, so every other delta already
// has the static pass's bits.  The reference moves the subtree down instead
// (moveDownTreeBrandes, src/streamingbc_internal.c:880-1401).
//  A  movers (old levels, level by level from the deeper endpoint s): a
//     child of a mover moves iff every parent it had is a mover; a child
//     that keeps a parent loses one (its path count is recomputed);
//  B  new levels of the movers: 1 + the lowest level among their non-mover
//     neighbours, then a breadth-first pass among the movers (a mover left
//     without any is now out of reach);
//  C  path counts, level by level downwards: movers, vertices that lost or
//     gained a parent, and the children of every vertex whose count changed
//     pull the sum of their parents' counts;
//  D  deltas, level by level upwards: the vertices of C whose count or level
//     changed, p (which lost s), and the parents of every vertex whose
//     delta, count or level changed;
//  E  state bytes, out-of-reach movers, scratch.
// Every vertex touched is listed once in the root's list (flags in the
// dpath scratch word), so the list never outgrows its nv entries.  BC
// changes by new - old for every delta that changes.

#define DM_LIST 1ll     // in the touched list
#define DM_SEEN 2ll     // A: a child of a mover, being decided
#define DM_MOVER 4ll    // level changes
#define DM_SIGSEED 8ll  // lost or gained a parent: path count recomputed
#define DM_SIGPROP 16ll // a parent's path count changed: recomputed
#define DM_SIGCH 32ll   // path count or level changed
#define DM_DEL 64ll     // delta recomputed: a child's term changed
#define DM_DONE 128ll   // path count recomputed

__device__ __forceinline__ int64_t or64_wg(int64_t* p, int64_t v)
{
    return (int64_t)__hip_atomic_fetch_or((unsigned long long*)p, (unsigned long long)v, __ATOMIC_RELAXED, WG_SCOPE);
}

__device__ __forceinline__ int32_t min32_wg(int32_t* p, int32_t v)
{
    return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, WG_SCOPE);
}

// Flags y and lists it the first time it is flagged; the whole wavefront
// calls this once per chunk.
__device__ __forceinline__ void dm_flag(bool want, int32_t y, int64_t flags, SbcVRec* vr, int32_t* list, int* tail)
{
    bool app = false;
    if (want) app = !(or64_wg(&vr[y].dpath, flags | DM_LIST) & DM_LIST);
    const int pos = wave_append(app, tail);
    if (pos >= 0) list[pos] = y;
}

// Vertices with more neighbours than this are handled by the whole workgroup
// in C and D (at most DM_BIGQ a level; more go to single wavefronts): one
// wavefront would need at least 256 rounds of loads for them, while the
// workgroup handles them one at a time.
constexpr int DM_BIG = 16384;
constexpr int DM_BIGQ = 255;
// chunks per wavefront in one segment of a whole-workgroup delta sum
constexpr int DM_SEGU = 4;

// Queues x for the whole-workgroup pass (lane 0 takes the slot); false when
// the queue is full and the calling wavefront handles x itself.
__device__ __forceinline__ bool dm_queue_big(int32_t x, int* big)
{
    int b = 0;
    if (lane_id() == 0) {
        b = atomicAdd(&big[0], 1);
        if (b < DM_BIGQ) big[1 + b] = x;
    }
    return __shfl(b, 0) < DM_BIGQ;
}

// Flags (and lists) the neighbours of a vertex on level lv: chunks w, w +
// step, ... of its list (one wavefront: 0, 1; the whole workgroup: wave,
// NWB).  hi, when given, is raised to lv if any neighbour was flagged.
__device__ __forceinline__ void dm_mark_level(const SbcParams& P, const SbcVInfo& vi, int w, int step, int32_t lv,
                                              int64_t flag, const int32_t* dist, SbcVRec* vr, int32_t* list, int* tail,
                                              int* hi)
{
    bool any = false;
    for (int e0 = w * 64; e0 < vi.deg; e0 += step * 64) {
        const int e = e0 + lane_id();
        int32_t y = 0;
        bool m = false;
        if (e < vi.deg) {
            y = P.col[vi.start + e];
            m = ld_acq32(&dist[y]) == lv;
        }
        any |= m;
        dm_flag(m, y, flag, vr, list, tail);
    }
    if (hi && ballot(any) != 0 && lane_id() == 0) atomicMax(hi, lv);
}

template <int NWB>
__global__ __launch_bounds__(NWB * 64) void sbc_delmove_kernel(SbcParams P)
{
    // 0: touched-list end, 1 and 2: the last round of A and of B that found
    // more work (a round writes its own number: nothing to reset), 3: popped
    // queue index
    __shared__ int cnt[4];
    // 0, 1: level range of pending path counts; 2: highest level of D
    __shared__ int lv[4];
    // whole-workgroup vertices of a level: count, then the vertices
    __shared__ int big[1 + DM_BIGQ];
    __shared__ double dmrow[NWB][DM_SEGU * 64];
    __shared__ int dmcnt[NWB];
    __shared__ unsigned long long bsum;
    __shared__ int bflag;
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const bool leader = threadIdx.x == 0;
    const int nv = P.nv;
    const int nq = *P.scnt;
    for (;;) {
        if (leader) cnt[3] = atomicAdd(&P.qctr[0], 1);
        __syncthreads();
        const int qi = cnt[3];
        __syncthreads();
        if (qi >= nq) break;
        const int ri = P.sroots[qi];
        const size_t off = (size_t)ri * (size_t)nv;
        int32_t* dist = P.dist + off;
        uint8_t* sb = P.sb + (size_t)ri * (size_t)P.sb_stride;
        SbcVRec* vr = P.vr + off;
        int32_t* list = P.list + off;
        const int32_t r = P.roots[ri];
        const int32_t du = ld_acq32(&dist[P.u0]), dv = ld_acq32(&dist[P.v0]);
        const int32_t s = du > dv ? P.u0 : P.v0, p = du > dv ? P.v0 : P.u0;
        const int32_t ls = du > dv ? du : dv, lp = du > dv ? dv : du;
        __syncthreads();
        if (leader) {
            list[0] = s;
            st_rel64(&vr[s].dpath, DM_LIST | DM_SEEN | DM_MOVER);
            cnt[0] = 1;
            cnt[1] = -1;
            cnt[2] = -1;
        }
        __syncthreads();

        // A: movers, old level by old level
        for (int L = ls;; ++L) {
            const int t1 = cnt[0];
            for (int i = wave; i < t1; i += NWB) {
                const int32_t w = list[i];
                if (!(ld_acq64(&vr[w].dpath) & DM_MOVER) || ld_acq32(&dist[w]) != L) continue;
                const SbcVInfo vi = P.vinfo[w];
                for (int e0 = 0; e0 < vi.deg; e0 += 64) {
                    const int e = e0 + lane;
                    int32_t y = 0;
                    bool c = false;
                    if (e < vi.deg) {
                        y = P.col[vi.start + e];
                        c = ld_acq32(&dist[y]) == L + 1;
                    }
                    dm_flag(c, y, DM_SEEN, vr, list, &cnt[0]);
                }
            }
            __syncthreads();
            const int t2 = cnt[0];
            for (int i = wave; i < t2; i += NWB) {
                const int32_t y = list[i];
                const int64_t fy = ld_acq64(&vr[y].dpath);
                if (!(fy & DM_SEEN) || (fy & (DM_MOVER | DM_SIGSEED)) || ld_acq32(&dist[y]) != L + 1) continue;
                const SbcVInfo vi = P.vinfo[y];
                bool keep = false;
                for (int e0 = 0; e0 < vi.deg && !keep; e0 += 64) {
                    const int e = e0 + lane;
                    bool k = false;
                    if (e < vi.deg) {
                        const int32_t z = P.col[vi.start + e];
                        if (ld_acq32(&dist[z]) == L) k = !(ld_acq64(&vr[z].dpath) & DM_MOVER);
                    }
                    keep = ballot(k) != 0;
                }
                if (lane == 0) {
                    or64_wg(&vr[y].dpath, keep ? DM_SIGSEED : DM_MOVER);
                    if (!keep) cnt[1] = L;
                }
            }
            __syncthreads();
            const bool more = cnt[1] == L;
            __syncthreads();
            if (!more) break;
        }

        // B: new levels of the movers
        if (leader) lv[0] = INF32;
        __syncthreads();
        {
            const int t1 = cnt[0];
            for (int i = wave; i < t1; i += NWB) {
                const int32_t w = list[i];
                if (!(ld_acq64(&vr[w].dpath) & DM_MOVER)) continue;
                const SbcVInfo vi = P.vinfo[w];
                int32_t best = INF32;
                for (int e = lane; e < vi.deg; e += 64) {
                    const int32_t z = P.col[vi.start + e];
                    const int32_t dz = ld_acq32(&dist[z]);
                    // a mover's dist may already hold its key: movers are skipped
                    if (dz != INF32 && dz + 1 < best && !(ld_acq64(&vr[z].dpath) & DM_MOVER)) best = dz + 1;
                }
                best = wave_min32(best);
                if (lane == 0) {
                    st_rel32(&dist[w], best);
                    if (best != INF32) atomicMin(&lv[0], best);
                }
            }
        }
        __syncthreads();
        for (int k = lv[0]; k != INF32; ++k) {
            const int t1 = cnt[0];
            for (int i = wave; i < t1; i += NWB) {
                const int32_t w = list[i];
                if (!(ld_acq64(&vr[w].dpath) & DM_MOVER)) continue;
                const int32_t dw = ld_acq32(&dist[w]);
                if (dw != k) {
                    if (dw > k && dw != INF32 && lane == 0) cnt[2] = k;
                    continue;
                }
                const SbcVInfo vi = P.vinfo[w];
                bool relaxed = false;
                for (int e = lane; e < vi.deg; e += 64) {
                    const int32_t z = P.col[vi.start + e];
                    if (ld_acq32(&dist[z]) > k + 1 && (ld_acq64(&vr[z].dpath) & DM_MOVER)) {
                        min32_wg(&dist[z], k + 1);
                        relaxed = true;
                    }
                }
                if (ballot(relaxed) != 0 && lane == 0) cnt[2] = k;
            }
            __syncthreads();
            const bool more = cnt[2] == k;
            __syncthreads();
            if (!more) break;
        }

        // C: path counts, downwards.  Vertices gaining a mover as a parent
        // join the seeds first.
        if (leader) {
            lv[0] = INF32;
            lv[1] = -1;
            lv[2] = lp;
        }
        __syncthreads();
        {
            const int t1 = cnt[0];
            for (int i = wave; i < t1; i += NWB) {
                const int32_t x = list[i];
                const int64_t fx = ld_acq64(&vr[x].dpath);
                const int32_t dx = ld_acq32(&dist[x]);
                if (dx == INF32 || !(fx & (DM_MOVER | DM_SIGSEED))) continue;
                if (lane == 0) {
                    atomicMin(&lv[0], dx);
                    atomicMax(&lv[1], dx);
                }
                if (!(fx & DM_MOVER)) continue;
                const SbcVInfo vi = P.vinfo[x];
                bool any = false;
                for (int e0 = 0; e0 < vi.deg; e0 += 64) {
                    const int e = e0 + lane;
                    int32_t y = 0;
                    bool g = false;
                    if (e < vi.deg) {
                        y = P.col[vi.start + e];
                        g = ld_acq32(&dist[y]) == dx + 1 && !(ld_acq64(&vr[y].dpath) & DM_MOVER);
                    }
                    any |= g;
                    dm_flag(g, y, DM_SIGSEED, vr, list, &cnt[0]);
                }
                if (ballot(any) != 0 && lane == 0) atomicMax(&lv[1], dx + 1);
            }
        }
        __syncthreads();
        for (int L = lv[0];;) {
            const int hi = lv[1];
            if (leader) big[0] = 0;
            __syncthreads();
            if (L > hi) break;
            const int t1 = cnt[0];
            for (int i = wave; i < t1; i += NWB) {
                const int32_t x = list[i];
                const int64_t fx = ld_acq64(&vr[x].dpath);
                if ((fx & DM_DONE) || !(fx & (DM_MOVER | DM_SIGSEED | DM_SIGPROP)) || ld_acq32(&dist[x]) != L) continue;
                const SbcVInfo vi = P.vinfo[x];
                if (vi.deg > DM_BIG && dm_queue_big(x, big)) continue;
                int64_t part = 0;
                for (int e = lane; e < vi.deg; e += 64) {
                    const int32_t y = P.col[vi.start + e];
                    if (ld_acq32(&dist[y]) == L - 1) part = (int64_t)((uint64_t)part + (uint64_t)vr[y].sigma);
                }
                const int64_t sig = wave_sum64(part);
                const bool changed = sig != vr[x].sigma || (fx & DM_MOVER);
                if (lane == 0) {
                    vr[x].sigma = sig;
                    or64_wg(&vr[x].dpath, DM_DONE | (changed ? DM_SIGCH : 0ll));
                    if (changed) atomicMax(&lv[2], L);
                }
                if (changed) dm_mark_level(P, vi, 0, 1, L + 1, DM_SIGPROP, dist, vr, list, &cnt[0], &lv[1]);
            }
            __syncthreads();
            // vertices with many neighbours: the whole workgroup
            const int nb = big[0] < DM_BIGQ ? big[0] : DM_BIGQ;
            for (int bi = 0; bi < nb; ++bi) {
                const int32_t x = big[1 + bi];
                const SbcVInfo vi = P.vinfo[x];
                if (leader) bsum = 0ull;
                __syncthreads();
                int64_t part = 0;
                for (int e = threadIdx.x; e < vi.deg; e += NWB * 64) {
                    const int32_t y = P.col[vi.start + e];
                    if (ld_acq32(&dist[y]) == L - 1) part = (int64_t)((uint64_t)part + (uint64_t)vr[y].sigma);
                }
                part = wave_sum64(part);
                if (lane == 0) atomicAdd(&bsum, (unsigned long long)part);
                __syncthreads();
                if (leader) {
                    const int64_t fx = ld_acq64(&vr[x].dpath);
                    const int64_t sig = (int64_t)bsum;
                    const bool changed = sig != vr[x].sigma || (fx & DM_MOVER);
                    vr[x].sigma = sig;
                    or64_wg(&vr[x].dpath, DM_DONE | (changed ? DM_SIGCH : 0ll));
                    if (changed) atomicMax(&lv[2], L);
                    bflag = changed;
                }
                __syncthreads();
                if (bflag) dm_mark_level(P, vi, wave, NWB, L + 1, DM_SIGPROP, dist, vr, list, &cnt[0], &lv[1]);
                __syncthreads();
            }
            ++L;
        }

        // D: deltas, upwards from the deepest level C changed (p at least).
        // A vertex with more than DM_BIG neighbours is summed by the whole
        // workgroup after the level's others: in segments of NWB x DM_SEGU
        // chunks, each wavefront packs its chunks' child terms in order into
        // its LDS row and one lane adds the rows in order -- the same
        // sequence of additions as one wavefront's sum.
        if (leader) {
            const int64_t fp = or64_wg(&vr[p].dpath, DM_LIST | DM_DEL);
            if (!(fp & DM_LIST)) list[cnt[0]++] = p;
            if (P.dm_out) {
                P.dm_out[2 * qi] = cnt[0];
                P.dm_out[2 * qi + 1] = lv[2];
            }
        }
        if (P.dm_out) {
            // D and E over the whole GPU
            __syncthreads();
            continue;
        }
        for (int L = lv[2]; L >= 0; --L) {
            if (leader) big[0] = 0;
            __syncthreads();
            const int t1 = cnt[0];
            for (int i = wave; i < t1; i += NWB) {
                const int32_t x = list[i];
                const int64_t fx = ld_acq64(&vr[x].dpath);
                if (!(fx & (DM_SIGCH | DM_DEL)) || ld_acq32(&dist[x]) != L) continue;
                const SbcVInfo vi = P.vinfo[x];
                if (vi.deg > DM_BIG && dm_queue_big(x, big)) continue;
                const int64_t sx = vr[x].sigma;
                double acc = 0.0;
                for (int e0 = 0; e0 < vi.deg; e0 += 64 * DM_SEGU) {
                    bool ch[DM_SEGU];
                    double add[DM_SEGU];
#pragma unroll
                    for (int u = 0; u < DM_SEGU; ++u) {
                        const int e = e0 + u * 64 + lane;
                        ch[u] = false;
                        add[u] = 0.0;
                        if (e < vi.deg) {
                            const int32_t c = P.col[vi.start + e];
                            if (ld_acq32(&dist[c]) == L + 1) {
                                ch[u] = true;
                                add[u] = dep(sx, vr[c].sigma, vr[c].delta);
                            }
                        }
                    }
#pragma unroll
                    for (int u = 0; u < DM_SEGU; ++u) acc = wave_ordered_add(acc, ch[u], add[u], false, 0.0);
                }
                const double od = vr[x].delta;
                const bool changed = __double_as_longlong(acc) != __double_as_longlong(od);
                if (lane == 0 && changed) {
                    vr[x].delta = acc;
                    if (P.bc_atomic && x != r) atomicAdd(&P.bc[x], acc - od);
                }
                if (changed || (fx & DM_SIGCH)) dm_mark_level(P, vi, 0, 1, L - 1, DM_DEL, dist, vr, list, &cnt[0], nullptr);
            }
            __syncthreads();
            const int nb = big[0] < DM_BIGQ ? big[0] : DM_BIGQ;
            for (int bi = 0; bi < nb; ++bi) {
                const int32_t x = big[1 + bi];
                const SbcVInfo vi = P.vinfo[x];
                const int64_t sx = vr[x].sigma;
                const int nch = (vi.deg + 63) >> 6;
                double acc = 0.0; // thread 0's
                for (int s0 = 0; s0 < nch; s0 += NWB * DM_SEGU) {
                    bool ch[DM_SEGU];
                    double add[DM_SEGU];
#pragma unroll
                    for (int u = 0; u < DM_SEGU; ++u) {
                        const int e = (s0 + wave * DM_SEGU + u) * 64 + lane;
                        ch[u] = false;
                        add[u] = 0.0;
                        if (e < vi.deg) {
                            const int32_t c = P.col[vi.start + e];
                            if (ld_acq32(&dist[c]) == L + 1) {
                                ch[u] = true;
                                add[u] = dep(sx, vr[c].sigma, vr[c].delta);
                            }
                        }
                    }
                    int n = 0;
#pragma unroll
                    for (int u = 0; u < DM_SEGU; ++u) {
                        const uint64_t m = ballot(ch[u]);
                        if (ch[u]) dmrow[wave][n + __popcll(m & (lane == 0 ? 0ull : (~0ull >> (64 - lane))))] = add[u];
                        n += __popcll(m);
                    }
                    if (lane == 0) dmcnt[wave] = n;
                    __syncthreads();
                    if (leader)
                        for (int w = 0; w < NWB; ++w) acc = lds_chain_sum(acc, dmrow[w], dmcnt[w]);
                    __syncthreads();
                }
                if (leader) {
                    const int64_t fx = ld_acq64(&vr[x].dpath);
                    const double od = vr[x].delta;
                    const bool changed = __double_as_longlong(acc) != __double_as_longlong(od);
                    if (changed) {
                        vr[x].delta = acc;
                        if (P.bc_atomic && x != r) atomicAdd(&P.bc[x], acc - od);
                    }
                    bflag = changed || (fx & DM_SIGCH);
                }
                __syncthreads();
                if (bflag) dm_mark_level(P, vi, wave, NWB, L - 1, DM_DEL, dist, vr, list, &cnt[0], nullptr);
                __syncthreads();
            }
        }

        // E: state bytes, out-of-reach movers, scratch
        {
            const int t1 = cnt[0];
            for (int i = threadIdx.x; i < t1; i += NWB * 64) {
                const int32_t x = list[i];
                const int32_t dx = ld_acq32(&dist[x]);
                if (dx == INF32) {
                    sb[x] = SB_INF;
                    vr[x].sigma = SIGMA_INF;
                    const double od = vr[x].delta;
                    if (od != 0.0) {
                        if (P.bc_atomic) atomicAdd(&P.bc[x], -od);
                        vr[x].delta = 0.0;
                    }
                } else {
                    sb[x] = sb_make(dx, 0);
                }
                st_rel64(&vr[x].dpath, 0);
            }
        }
        __syncthreads();
    }
}

// The delta pass of level L of every movement root over the whole GPU (the
// D step above, spread over every workgroup: one wavefront per vertex, 4
// chunks of loads in flight, device-scope flags and list appends since
// several workgroups work on one root).  Entries appended meanwhile are the
// next level's.
__device__ __forceinline__ int64_t or64_dev(int64_t* p, int64_t v)
{
    return (int64_t)__hip_atomic_fetch_or((unsigned long long*)p, (unsigned long long)v, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(1024) void sbc_delmove_level_kernel(SbcParams P, int L)
{
    const int lane = lane_id();
    const int gw = (int)blockIdx.x * 16 + wave_in_block();
    const int nw = (int)gridDim.x * 16;
    const int nq = *P.scnt;
    const int nv = P.nv;
    const int32_t* lim = P.dm_out + 2 * P.K; // list lengths as the launch began
    for (int q = 0; q < nq; ++q) {
        if (L > P.dm_out[2 * q + 1]) continue;
        // every wavefront takes the same length: entries appended meanwhile
        // belong to the next level, and one read before its writer's store
        // is visible would name a vertex twice
        const int t1 = lim[q];
        const int ri = P.sroots[q];
        const size_t off = (size_t)ri * (size_t)nv;
        const int32_t* dist = P.dist + off;
        SbcVRec* vr = P.vr + off;
        int32_t* list = P.list + off;
        const int32_t r = P.roots[ri];
        for (int i = gw; i < t1; i += nw) {
            const int32_t x = list[i];
            const int64_t fx = vr[x].dpath;
            if (!(fx & (DM_SIGCH | DM_DEL)) || dist[x] != L) continue;
            const SbcVInfo vi = P.vinfo[x];
            // a vertex of many neighbours goes to the level's second launch
            // (one workgroup each), while the list has room
            if (P.dm_giant && vi.deg >= (P.wide_giant > 0 ? P.wide_giant : WIDE_GIANT)) {
                int gp = 0;
                if (lane == 0) gp = atomicAdd(&P.dm_giant[0], 1);
                gp = __builtin_amdgcn_readfirstlane(gp);
                if (gp < P.dm_giant_cap) {
                    if (lane == 0) {
                        P.dm_giant[1 + 2 * gp] = q;
                        P.dm_giant[2 + 2 * gp] = x;
                    }
                    continue;
                }
            }
            const int64_t sx = vr[x].sigma;
            double acc = 0.0;
            for (int e0 = 0; e0 < vi.deg; e0 += 64 * DM_SEGU) {
                bool ch[DM_SEGU];
                double add[DM_SEGU];
#pragma unroll
                for (int u = 0; u < DM_SEGU; ++u) {
                    const int e = e0 + u * 64 + lane;
                    ch[u] = false;
                    add[u] = 0.0;
                    if (e < vi.deg) {
                        const int32_t c = P.col[vi.start + e];
                        if (dist[c] == L + 1) {
                            ch[u] = true;
                            add[u] = dep(sx, vr[c].sigma, vr[c].delta);
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < DM_SEGU; ++u) acc = wave_ordered_add(acc, ch[u], add[u], false, 0.0);
            }
            const double od = vr[x].delta;
            const bool changed = __double_as_longlong(acc) != __double_as_longlong(od);
            if (lane == 0 && changed) {
                vr[x].delta = acc;
                if (P.bc_atomic && x != r) atomicAdd(&P.bc[x], acc - od);
            }
            if (!changed && !(fx & DM_SIGCH)) continue;
            for (int e0 = 0; e0 < vi.deg; e0 += 64) {
                const int e = e0 + lane;
                int32_t y = 0;
                bool app = false;
                if (e < vi.deg) {
                    y = P.col[vi.start + e];
                    if (dist[y] == L - 1) app = !(or64_dev(&vr[y].dpath, DM_DEL | DM_LIST) & DM_LIST);
                }
                const uint64_t m = ballot(app);
                if (m) {
                    int base = 0;
                    if (lane == 0) base = atomicAdd(&P.dm_out[2 * q], __popcll(m));
                    base = __shfl(base, 0);
                    if (app) list[base + prefix_in_wave(m)] = y;
                }
            }
        }
    }
}

// The delta of each listed vertex of level L (the level kernel's rule: the
// sum over children in adjacency order, from zero), one workgroup per
// vertex, then its parents marked and listed as the level kernel does.
__global__ __launch_bounds__(1024) void sbc_delmove_giant_kernel(SbcParams P, int L)
{
    __shared__ GiantShared gs;
    __shared__ int again;
    const int lane = lane_id(), wave = wave_in_block();
    const int ng = min(P.dm_giant[0], P.dm_giant_cap);
    const int nv = P.nv;
    for (int gi = (int)blockIdx.x; gi < ng; gi += (int)gridDim.x) {
        const int q = P.dm_giant[1 + 2 * gi];
        const int32_t x = P.dm_giant[2 + 2 * gi];
        const int ri = P.sroots[q];
        const size_t off = (size_t)ri * (size_t)nv;
        const int32_t* dist = P.dist + off;
        SbcVRec* vr = P.vr + off;
        int32_t* list = P.list + off;
        const SbcVInfo vi = P.vinfo[x];
        const int deg = __builtin_amdgcn_readfirstlane(vi.deg);
        const int nch = (deg + 63) >> 6;
        const int64_t fx = vr[x].dpath;
        const int64_t sx = vr[x].sigma;
        const double acc = giant_sum(0.0, nch, gs, [&](int c0, Term (&t)[GIANT_CPW]) {
            int32_t cc[GIANT_CPW];
            bool ok[GIANT_CPW];
#pragma unroll
            for (int u = 0; u < GIANT_CPW; ++u) {
                const int e = (c0 + u) * 64 + lane;
                ok[u] = c0 + u < nch && e < deg;
                cc[u] = ok[u] ? P.col[vi.start + e] : 0;
            }
#pragma unroll
            for (int u = 0; u < GIANT_CPW; ++u) {
                const bool ch = ok[u] && dist[cc[u]] == L + 1;
                t[u] = Term{ch, false, ch ? dep(sx, vr[cc[u]].sigma, vr[cc[u]].delta) : 0.0, 0.0};
            }
        });
        if (wave == 0 && lane == 0) {
            const double od = vr[x].delta;
            const bool changed = __double_as_longlong(acc) != __double_as_longlong(od);
            if (changed) {
                vr[x].delta = acc;
                if (P.bc_atomic && x != P.roots[ri]) atomicAdd(&P.bc[x], acc - od);
            }
            again = (changed || (fx & DM_SIGCH)) ? 1 : 0;
        }
        __syncthreads();
        if (again) {
            for (int ch = wave; ch < nch; ch += 16) {
                const int e = ch * 64 + lane;
                int32_t y = 0;
                bool app = false;
                if (e < deg) {
                    y = P.col[vi.start + e];
                    if (dist[y] == L - 1) app = !(or64_dev(&vr[y].dpath, DM_DEL | DM_LIST) & DM_LIST);
                }
                const int pos = wave_append(app, &P.dm_out[2 * q]);
                if (pos >= 0) list[pos] = y;
            }
        }
        __syncthreads();
    }
}

// Freezes every movement root's list length for the next level launch.
__global__ void sbc_delmove_snap_kernel(SbcParams P)
{
    const int nq = *P.scnt;
    for (int q = threadIdx.x; q < nq; q += blockDim.x) P.dm_out[2 * P.K + q] = P.dm_out[2 * q];
}

// E of every movement root over the whole GPU: state bytes, out-of-reach
// movers, scratch.
__global__ void sbc_delmove_finish_kernel(SbcParams P)
{
    const int nq = *P.scnt;
    const int nv = P.nv;
    const int gt = (int)(blockIdx.x * blockDim.x + threadIdx.x), nt = (int)(gridDim.x * blockDim.x);
    for (int q = 0; q < nq; ++q) {
        const int t1 = P.dm_out[2 * q];
        const int ri = P.sroots[q];
        const size_t off = (size_t)ri * (size_t)nv;
        const int32_t* dist = P.dist + off;
        uint8_t* sb = P.sb + (size_t)ri * (size_t)P.sb_stride;
        SbcVRec* vr = P.vr + off;
        const int32_t* list = P.list + off;
        for (int i = gt; i < t1; i += nt) {
            const int32_t x = list[i];
            const int32_t dx = dist[x];
            if (dx == INF32) {
                sb[x] = SB_INF;
                vr[x].sigma = SIGMA_INF;
                const double od = vr[x].delta;
                if (od != 0.0) {
                    if (P.bc_atomic) atomicAdd(&P.bc[x], -od);
                    vr[x].delta = 0.0;
                }
            } else {
                sb[x] = sb_make(dx, 0);
            }
            vr[x].dpath = 0;
        }
    }
}

// Clears the work-queue counters (a kernel, not a memset, so it can sit in
// a captured graph).
__global__ void sbc_reset_kernel(int32_t* q, int n)
{
    if ((int)threadIdx.x < n) q[threadIdx.x] = 0;
}

// Applies a batch of appended slots and vertex-info updates.
__global__ void sbc_append_kernel(SbcAppend A)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (A.desc) {
        if (t < A.desc->nslots) {
            A.col[A.d_slot[t]] = A.d_scol[t];
            A.seq[A.d_slot[t]] = A.d_sseq[t];
        }
        if (t < A.desc->nvupd) A.vinfo[A.d_vv[t]] = A.d_vi[t];
    } else {
        if (t < A.nslots) {
            A.col[A.slot[t]] = A.scol[t];
            A.seq[A.slot[t]] = A.sseq[t];
        }
        if (t < A.nvupd) A.vinfo[A.vv[t]] = A.vi[t];
    }
}

// ------------------------------------------------------------ launchers

extern "C" hipError_t sbc_launch_window(const SbcParams* P, int nwb, int grid, hipStream_t st)
{
    // per-item tracing is compiled into the 1024-thread kernel only
    if (P->tprof && nwb == 16) {
        hipLaunchKernelGGL((sbc_window_kernel<16, true>), dim3(grid), dim3(1024), 0, st, *P);
        return hipGetLastError();
    }
    switch (nwb) {
    case 1: hipLaunchKernelGGL(sbc_window_kernel<1>, dim3(grid), dim3(64), 0, st, *P); break;
    case 2: hipLaunchKernelGGL(sbc_window_kernel<2>, dim3(grid), dim3(128), 0, st, *P); break;
    case 4: hipLaunchKernelGGL(sbc_window_kernel<4>, dim3(grid), dim3(256), 0, st, *P); break;
    case 8: hipLaunchKernelGGL(sbc_window_kernel<8>, dim3(grid), dim3(512), 0, st, *P); break;
    case 16: hipLaunchKernelGGL(sbc_window_kernel<16>, dim3(grid), dim3(1024), 0, st, *P); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

constexpr int WIDE_NW = 4;  // wavefronts per workgroup of the wide kernels

extern "C" hipError_t sbc_launch_wide_descent(const SbcParams* P, SbcWideItem* wi, int ri, int fs, int fe, int lc,
                                              const SbcWideLists* L, int num_cu, hipStream_t st)
{
    const int cu = num_cu > 0 ? num_cu : 256;
    const int n = fe - fs;
    int g1 = (n + WIDE_NW * 64 - 1) / (WIDE_NW * 64);
    g1 = g1 < 1 ? 1 : (g1 < 8 * cu ? g1 : 8 * cu);
    hipLaunchKernelGGL(sbc_reset_kernel, dim3(1), dim3(64), 0, st, L->ctr, 4);
    hipLaunchKernelGGL(sbc_wide_descent_light_kernel<WIDE_NW>, dim3(g1), dim3(WIDE_NW * 64), 0, st, *P, wi, ri, fs, fe,
                       lc, *L);
    hipLaunchKernelGGL(sbc_wide_descent_heavy_kernel<WIDE_NW>, dim3(8 * cu), dim3(WIDE_NW * 64), 0, st, *P, wi, ri, lc,
                       *L);
    return hipGetLastError();
}

extern "C" hipError_t sbc_launch_wide_ascent(const SbcParams* P, SbcWideItem* wi, int ri, const int32_t* h, int lev,
                                             int hubs_only, const SbcWideLists* L, int num_cu, hipStream_t st)
{
    const int cu = num_cu > 0 ? num_cu : 256;
    const int tot = (h[1] - h[0]) + (h[3] - h[2]) + (h[5] - h[4]);
    int g1 = (tot + WIDE_NW * 64 - 1) / (WIDE_NW * 64);
    g1 = g1 < 1 ? 1 : (g1 < 8 * cu ? g1 : 8 * cu);
    hipLaunchKernelGGL(sbc_reset_kernel, dim3(1), dim3(64), 0, st, L->ctr, 4);
    hipLaunchKernelGGL(sbc_wide_ascent_light_kernel<WIDE_NW>, dim3(g1), dim3(WIDE_NW * 64), 0, st, *P, wi, ri, h[0],
                       h[1], h[2], h[3], h[4], h[5], lev, hubs_only, *L);
    hipLaunchKernelGGL(sbc_wide_ascent_heavy_kernel, dim3(cu), dim3(1024), 0, st, *P, wi, ri, lev, *L);
    return hipGetLastError();
}

extern "C" hipError_t sbc_launch_wide_commit(const SbcParams* P, SbcWideItem* wi, int ri, int nf, int nD, int nUend,
                                             int nS, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(sbc_wide_commit_kernel<WIDE_NW>, dim3(grid), dim3(WIDE_NW * 64), 0, st, *P, wi, ri, nf, nD, nUend,
                       nS);
    return hipGetLastError();
}

extern "C" hipError_t sbc_launch_static(const SbcParams* P, int nwb, int grid, hipStream_t st)
{
    switch (nwb) {
    case 4: hipLaunchKernelGGL(sbc_static_kernel<4>, dim3(grid), dim3(256), 0, st, *P); break;
    case 8: hipLaunchKernelGGL(sbc_static_kernel<8>, dim3(grid), dim3(512), 0, st, *P); break;
    case 16: hipLaunchKernelGGL(sbc_static_kernel<16>, dim3(grid), dim3(1024), 0, st, *P); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// LDS bitmap sizes (32-bit words) of the two instantiations
constexpr int SBM_WORDS_SMALL = 8448;  // up to 270,336 vertices
constexpr int SBM_WORDS_LARGE = 36864; // up to 1,179,648 vertices

extern "C" int64_t sbc_static_bm_max_nv() { return (int64_t)SBM_WORDS_LARGE * 32; }

extern "C" hipError_t sbc_launch_static_bm(const SbcParams* P, int grid, hipStream_t st)
{
    const int64_t nwords = ((int64_t)P->nv + 31) / 32;
    if (nwords <= SBM_WORDS_SMALL)
        hipLaunchKernelGGL((sbc_static_bm_kernel<16, SBM_WORDS_SMALL>), dim3(grid), dim3(1024), 0, st, *P);
    else if (nwords <= SBM_WORDS_LARGE)
        hipLaunchKernelGGL((sbc_static_bm_kernel<16, SBM_WORDS_LARGE>), dim3(grid), dim3(1024), 0, st, *P);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

extern "C" hipError_t sbc_launch_bc_reduce(const SbcParams* P, hipStream_t st)
{
    const int bs = 256;
    hipLaunchKernelGGL(sbc_bc_reduce_kernel, dim3((P->nv + bs - 1) / bs), dim3(bs), 0, st, *P);
    return hipGetLastError();
}

extern "C" hipError_t sbc_launch_delete_classify(const SbcParams* P, int32_t u, int32_t v, int32_t* out,
                                                 hipStream_t st)
{
    const int bs = 256;
    hipLaunchKernelGGL(sbc_delete_classify_kernel, dim3((P->K + bs - 1) / bs), dim3(bs), 0, st, *P, u, v, out);
    return hipGetLastError();
}

extern "C" hipError_t sbc_launch_delmove(const SbcParams* P, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(sbc_delmove_kernel<16>, dim3(grid), dim3(1024), 0, st, *P);
    return hipGetLastError();
}

extern "C" hipError_t sbc_launch_delmove_level(const SbcParams* P, int L, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(sbc_delmove_snap_kernel, dim3(1), dim3(256), 0, st, *P);
    if (P->dm_giant) hipLaunchKernelGGL(sbc_reset_kernel, dim3(1), dim3(64), 0, st, P->dm_giant, 1);
    hipLaunchKernelGGL(sbc_delmove_level_kernel, dim3(grid), dim3(1024), 0, st, *P, L);
    if (P->dm_giant) hipLaunchKernelGGL(sbc_delmove_giant_kernel, dim3(grid), dim3(1024), 0, st, *P, L);
    return hipGetLastError();
}

extern "C" hipError_t sbc_launch_delmove_finish(const SbcParams* P, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(sbc_delmove_finish_kernel, dim3(grid), dim3(256), 0, st, *P);
    return hipGetLastError();
}

extern "C" hipError_t sbc_launch_wide_mark(const SbcParams* P, SbcWideItem* wi, int ri, int32_t x, int lev_parent,
                                           hipStream_t st)
{
    hipLaunchKernelGGL(sbc_wide_mark_kernel, dim3(1), dim3(64), 0, st, *P, wi, ri, x, lev_parent);
    return hipGetLastError();
}

static int root_grid(long items, int per_block)
{
    const long g = (items + per_block - 1) / per_block;
    return (int)(g < 1 ? 1 : (g > 4096 ? 4096 : g));
}

extern "C" hipError_t sbc_launch_root_reset(const SbcParams* P, int ri, int32_t* fr, hipStream_t st)
{
    hipLaunchKernelGGL(sbc_root_reset_kernel, dim3(root_grid(P->nv, 256)), dim3(256), 0, st, *P, ri, fr);
    return hipGetLastError();
}

extern "C" hipError_t sbc_launch_root_bfs(const SbcParams* P, int ri, int fs, int fe, int j, int32_t* fr,
                                          hipStream_t st)
{
    hipLaunchKernelGGL(sbc_root_bfs_kernel, dim3(root_grid(fe - fs, 4)), dim3(256), 0, st, *P, ri, fs, fe, j, fr);
    return hipGetLastError();
}

extern "C" hipError_t sbc_launch_root_settle(const SbcParams* P, int ri, int a, int b, int j, hipStream_t st)
{
    hipLaunchKernelGGL(sbc_root_settle_kernel, dim3(root_grid(b - a, 256)), dim3(256), 0, st, *P, ri, a, b, j);
    return hipGetLastError();
}

extern "C" hipError_t sbc_launch_root_ascent(const SbcParams* P, int ri, int a, int b, int L, int32_t* big,
                                             hipStream_t st)
{
    hipLaunchKernelGGL(sbc_root_ascent_kernel, dim3(root_grid(b - a, 4)), dim3(256), 0, st, *P, ri, a, b, L, big);
    return hipGetLastError();
}

extern "C" hipError_t sbc_launch_root_bigsum(const SbcParams* P, int ri, int L, const int32_t* big, const int64_t* toff,
                                             int n, int64_t nterms, double* terms, hipStream_t st)
{
    hipLaunchKernelGGL(sbc_root_terms_kernel, dim3(root_grid(nterms, 256)), dim3(256), 0, st, *P, ri, L, big, toff, n,
                       terms);
    hipLaunchKernelGGL(sbc_root_bigsum_kernel, dim3(n), dim3(64), 0, st, *P, ri, big, toff, n, terms);
    return hipGetLastError();
}

extern "C" hipError_t sbc_launch_bc_roots(const SbcParams* P, const int32_t* list, const int32_t* cnt, double sign,
                                          hipStream_t st)
{
    const int bs = 256;
    hipLaunchKernelGGL(sbc_bc_roots_kernel, dim3((P->nv + bs - 1) / bs), dim3(bs), 0, st, *P, list, cnt, sign);
    return hipGetLastError();
}

extern "C" hipError_t sbc_launch_roots_reset(const SbcParams* P, const int32_t* rl, int n, int32_t* fr, hipStream_t st)
{
    hipLaunchKernelGGL(sbc_roots_reset_kernel, dim3(root_grid((long)n * P->nv, 256)), dim3(256), 0, st, *P, rl, n, fr);
    return hipGetLastError();
}

// kind 0: BFS (one wavefront per vertex), 1: settle (one thread), 2: ascent (one wavefront)
extern "C" hipError_t sbc_launch_roots_level(const SbcParams* P, int kind, const SbcRootSeg* sg, const int64_t* pre,
                                             int n, int64_t items, int32_t* fr, hipStream_t st)
{
    if (kind == 0)
        hipLaunchKernelGGL(sbc_roots_bfs_kernel, dim3(root_grid(items, 4)), dim3(256), 0, st, *P, sg, pre, n, fr);
    else if (kind == 1)
        hipLaunchKernelGGL(sbc_roots_settle_kernel, dim3(root_grid(items, 256)), dim3(256), 0, st, *P, sg, pre, n);
    else
        hipLaunchKernelGGL(sbc_roots_ascent_kernel, dim3(root_grid(items, 4)), dim3(256), 0, st, *P, sg, pre, n);
    return hipGetLastError();
}

extern "C" hipError_t sbc_launch_reset(int32_t* q, int n, hipStream_t st)
{
    hipLaunchKernelGGL(sbc_reset_kernel, dim3(1), dim3(64), 0, st, q, n);
    return hipGetLastError();
}

extern "C" hipError_t sbc_launch_append(const SbcAppend* A, int maxn, hipStream_t st)
{
    if (maxn <= 0) return hipSuccess;
    const int bs = 256;
    hipLaunchKernelGGL(sbc_append_kernel, dim3((maxn + bs - 1) / bs), dim3(bs), 0, st, *A);
    return hipGetLastError();
}

// ------------------------------------------------------------ fixed scratch
//
// Device scratch of the entry point that allocates nothing
// (sbcInsertEdgeNoAlloc): the work-queue counters, then one item record per
// root.  One copy per GPU, held until the process ends; every call on a GPU
// shares it, so calls that overlap on one GPU race.  Host code never uses the
// array's name as an address: the device address comes from
// hipGetSymbolAddress, once per device, below.

constexpr size_t SBC_NOALLOC_QCTR_BYTES = 256;
static __device__ __align__(256) unsigned char scratch[SBC_NOALLOC_QCTR_BYTES +
                                                       (size_t)SBC_NOALLOC_MAX_ROOTS * sizeof(SbcItemRec)];

#define SBC_MAX_DEVICES 64
static void* g_scratch_ptr[SBC_MAX_DEVICES];
static int g_scratch_err[SBC_MAX_DEVICES];
static std::once_flag g_scratch_once[SBC_MAX_DEVICES];

extern "C" hipError_t sbc_noalloc_scratch(int32_t** qctr, SbcItemRec** rec)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= SBC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::call_once(g_scratch_once[dev], [dev]() {
        void* p = nullptr;
        // the current device is dev: its own copy of the array
        hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_err[dev] = (int)r;
        g_scratch_ptr[dev] = r == hipSuccess ? p : nullptr;
    });
    if (g_scratch_err[dev] != (int)hipSuccess) return (hipError_t)g_scratch_err[dev];
    unsigned char* base = (unsigned char*)g_scratch_ptr[dev];
    *qctr = (int32_t*)base;
    *rec = (SbcItemRec*)(base + SBC_NOALLOC_QCTR_BYTES);
    return hipSuccess;
}

// ------------------------------------------------------------ window graph
//
// The launch sequence of one window (append, queue reset, window kernel,
// optional BC reduction) built node by node into a HIP graph, with each
// node's arguments copied into the node, so no stream is ever in capture
// mode: allocations, frees and synchronisations that other host threads
// make meanwhile (one per resident graph in the serving loop) cannot
// invalidate it.

static hipError_t add_node(hipGraph_t g, hipGraphNode_t* prev, void* func, dim3 grid, dim3 block, void** args)
{
    hipKernelNodeParams kp;
    memset(&kp, 0, sizeof(kp));
    kp.func = func;
    kp.gridDim = grid;
    kp.blockDim = block;
    kp.kernelParams = args;
    kp.sharedMemBytes = 0;
    kp.extra = nullptr;
    hipGraphNode_t n = nullptr;
    hipError_t e = hipGraphAddKernelNode(&n, g, *prev ? prev : nullptr, *prev ? 1 : 0, &kp);
    if (e == hipSuccess) *prev = n;
    return e;
}

extern "C" hipError_t sbc_graph_window(hipGraph_t g, const SbcAppend* A, int maxn_append, int32_t* qctr,
                                       const SbcParams* P, int nwb, int grid, int recompute)
{
    hipGraphNode_t prev = nullptr;
    hipError_t e;
    SbcAppend a = *A;
    SbcParams p = *P;
    if (maxn_append > 0) {
        void* args[] = {&a};
        e = add_node(g, &prev, (void*)sbc_append_kernel, dim3((maxn_append + 255) / 256), dim3(256), args);
        if (e != hipSuccess) return e;
    }
    {
        int n = 4;
        void* args[] = {&qctr, &n};
        e = add_node(g, &prev, (void*)sbc_reset_kernel, dim3(1), dim3(64), args);
        if (e != hipSuccess) return e;
    }
    {
        void* args[] = {&p};
        void* fn = nullptr;
        if (P->tprof && nwb == 16) fn = (void*)sbc_window_kernel<16, true>;
        else if (nwb == 1) fn = (void*)sbc_window_kernel<1>;
        else if (nwb == 2) fn = (void*)sbc_window_kernel<2>;
        else if (nwb == 4) fn = (void*)sbc_window_kernel<4>;
        else if (nwb == 8) fn = (void*)sbc_window_kernel<8>;
        else if (nwb == 16) fn = (void*)sbc_window_kernel<16>;
        else return hipErrorInvalidValue;
        e = add_node(g, &prev, fn, dim3(grid), dim3(nwb * 64), args);
        if (e != hipSuccess) return e;
    }
    if (recompute) {
        void* args[] = {&p};
        e = add_node(g, &prev, (void*)sbc_bc_reduce_kernel, dim3((P->nv + 255) / 256), dim3(256), args);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
