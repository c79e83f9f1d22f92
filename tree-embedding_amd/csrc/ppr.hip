// ppr.hip -- subset forward push and dynamic insertion update on the device.
//
// Replaces (ppr_computation_store_dynamic.h):
//   nodegree_DenseUndirected_Refresh_PPR_initialization   :2035-2121
//   nodegree_DenseDirected_Refresh_PPR_initialization     :3811-3898
//   nodegree_DenseDirected_Refresh_PPR_initialization_Transpose :4133-4228
//   nodegree_DenseDynamicForwardPush                       :1394-1498
//   nodegree_DenseDirectedDynamicForwardPush[Transpose]    :2945-3030, :3356-3437
//
// The reference runs one FIFO per source on one CPU thread per source range.
// Here all sources advance together in synchronous rounds.  A work item is a
// (vertex v, source group g) pair with a 64-bit mask of the sources of that
// group whose residue at v exceeds rmax; one wavefront pops the item (lane =
// source) and walks v's neighbour list once for all 64 sources, so every
// residue update is a coalesced 64-lane atomic on r[u][g*64 .. g*64+63].
// A residue that crosses rmax upward sets its bit in the next round's mask;
// the first setter of a (u, g) mask appends the item, so each item appears
// once per round.  Only same-address atomics are relied on: a pop is an
// atomic exchange, a relaxation an atomic add, so their order on one address
// decides everything and no cross-address ordering is needed.  A popped value
// at or below rmax (the dynamic update can lower residues) is added back.
//
// Semantics kept from the reference: threshold r > rmax (not degree-scaled,
// positive side only); p += alpha*r, neighbours += (1-alpha)*r/deg; an
// undirected degree-0 vertex (only a source can be one) gets p = alpha*r and
// keeps r; a directed dangling vertex (degree 0 in the push direction) is never
// queued and keeps its residue; the insertion update for new slot j of row i
// (j == 0 skipped): p *= (j+1)/j, r[i] -= p/(j+1)/alpha,
// r[to] += (1-alpha) p/(j+1)/alpha, with the exact ratio (the reference's
// integer (j+1)/j is not reproduced; its n >= 1e7 path, ppr:6518-6522, uses the
// exact ratio too).  Order of pushes differs from the reference's FIFO, so
// reserves agree within the push error bound, not bitwise.
#include "tsvd_ppr.h"

#include <hipcub/hipcub.hpp>
#include <algorithm>

namespace tsvd {

namespace {

constexpr int TPB = 256;
constexpr int WPB = TPB / 64;   // waves per block

__device__ inline float xchg(float *a, float v) { return atomicExch(a, v); }
__device__ inline double xchg(double *a, double v)
{
    unsigned long long o = atomicExch((unsigned long long *)a, (unsigned long long)__double_as_longlong(v));
    return __longlong_as_double((long long)o);
}

__device__ inline void append_item(uint32_t item, uint32_t *list, uint32_t *count, uint32_t cap, uint32_t *overflow)
{
    uint32_t pos = atomicAdd(count, 1u);
    if (pos < cap) list[pos] = item;
    else atomicOr(overflow, 1u);
}

// A residue beyond the push threshold: r > rmax (the reference's positive
// side, ppr:1445), or |r| > rmax when pushing two-sided.  Written without
// short-circuits so that it compiles to compares and masks, not branches.
template <class T>
__device__ inline bool beyond(T x, T rm, int two)
{
    return (x > rm) | ((two != 0) & (x < -rm));
}

// Lane-0 sets `bits` in mask[item]; the first setter appends the item.
__device__ inline void mark(unsigned long long bits, uint32_t item, unsigned long long *mask, uint32_t *list,
                            uint32_t *count, uint32_t cap, uint32_t *overflow)
{
    unsigned long long old = atomicOr(mask + item, bits);
    if (old == 0ull) append_item(item, list, count, cap, overflow);
}

// Staged appends.  Returning atomics on one address retire at about 9e7 per
// second for the whole device, so one global append per item would serialise
// a round of millions of items on the list's counter.  A workgroup instead
// collects its new items in LDS (LDS atomics) and reserves their place in
// the list with a single global atomic when it finishes; past its stage
// capacity it appends directly.
constexpr uint32_t STAGE = 4096;

// The push kernels stage per wavefront: each wave of a workgroup has its own
// LDS stage that only its lane 0 appends to, and the wave drains and flushes
// it by itself, with no workgroup barrier.  A wave can then take several
// items per round without waiting on the slowest wave of its workgroup, and
// a round's list atomics are one per wave per drain.
constexpr int PTPB = 256;
constexpr int PWPB = PTPB / 64;
constexpr uint32_t PSTAGE = 1024;   // items per wave

struct WaveStage {
    uint32_t *items;    // LDS [PSTAGE] of this wave
    uint32_t *n;        // LDS: items staged (written by lane 0 only)
    uint32_t *list;     // the global item list, its counter and capacity
    uint32_t *count;
    uint32_t cap;
    uint32_t *overflow;
};

// Lane 0 only: sets `bits` in mask[item]; the first setter stages the item.
__device__ inline void wave_mark(const WaveStage &ws, unsigned long long bits, uint32_t item,
                                 unsigned long long *mask)
{
    const unsigned long long old = atomicOr(mask + item, bits);
    if (old != 0ull) return;
    const uint32_t pos = *ws.n;
    *ws.n = pos + 1;
    if (pos < PSTAGE) ws.items[pos] = item;
    else append_item(item, ws.list, ws.count, ws.cap, ws.overflow);
}

// Every lane of the wave, at a wave-uniform point: moves the staged items to
// the list once more than `thresh` wait (one list atomic for all of them).
// Lane 0 wrote the stage; a wave's LDS operations complete in order, so the
// other lanes read what it wrote.
__device__ inline void wave_drain(const WaveStage &ws, uint32_t thresh, int lane)
{
    uint32_t n = lane == 0 ? min(*ws.n, PSTAGE) : 0u;
    n = __shfl(n, 0);
    if (n == 0 || n <= thresh) return;
    uint32_t b = lane == 0 ? atomicAdd(ws.count, n) : 0u;
    b = __shfl(b, 0);
    for (uint32_t t = lane; t < n; t += 64) {
        if (b + t < ws.cap) ws.list[b + t] = ws.items[t];
        else atomicOr(ws.overflow, 1u);
    }
    if (lane == 0) *ws.n = 0u;
}

// Every thread of the workgroup, once, after its wave's last mark: the waves'
// remaining staged items go to the list under one list atomic for the whole
// workgroup.  `spos` is LDS for PWPB + 1 counts.
__device__ inline void wave_flush_all(const WaveStage &ws, uint32_t *spos, int lane)
{
    const int wv = threadIdx.x >> 6;
    if (lane == 0) spos[wv] = min(*ws.n, PSTAGE);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < PWPB; w++) {
            const uint32_t c = spos[w];
            spos[w] = tot;
            tot += c;
        }
        spos[PWPB] = tot ? atomicAdd(ws.count, tot) : 0u;
    }
    __syncthreads();
    const uint32_t n = min(*ws.n, PSTAGE), b = spos[PWPB] + spos[wv];
    for (uint32_t t = lane; t < n; t += 64) {
        if (b + t < ws.cap) ws.list[b + t] = ws.items[t];
        else atomicOr(ws.overflow, 1u);
    }
}

// Push counters are spread over STAT_SHARDS copies (by workgroup), for the
// same reason: one address per counter would serialise a workgroup's end.
constexpr int STAT_SHARDS = 32, STAT_FIELDS = 8;

__device__ inline unsigned long long *stat_slot(unsigned long long *stats)
{
    return stats + (blockIdx.x % STAT_SHARDS) * STAT_FIELDS;
}

struct Stage {
    uint32_t *items;    // LDS [lcap]
    uint32_t *n;        // LDS: items staged
    uint32_t *base;     // LDS: reserved list position
    uint32_t *list;     // the global item list, its counter and capacity
    uint32_t *count;
    uint32_t cap;
    uint32_t *overflow;
    uint32_t lcap;      // staged items the LDS holds
};

__device__ inline void stage_mark(const Stage &sg, unsigned long long bits, uint32_t item, unsigned long long *mask)
{
    const unsigned long long old = atomicOr(mask + item, bits);
    if (old != 0ull) return;
    const uint32_t pos = atomicAdd(sg.n, 1u);
    if (pos < sg.lcap) sg.items[pos] = item;
    else append_item(item, sg.list, sg.count, sg.cap, sg.overflow);
}

// Every thread of the workgroup calls this once, after its last stage_mark.
__device__ inline void stage_flush(const Stage &sg)
{
    __syncthreads();
    const uint32_t n = min(*sg.n, sg.lcap);
    if (threadIdx.x == 0) *sg.base = n ? atomicAdd(sg.count, n) : 0u;
    __syncthreads();
    const uint32_t b = *sg.base;
    for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) {
        if (b + t < sg.cap) sg.list[b + t] = sg.items[t];
        else atomicOr(sg.overflow, 1u);
    }
}

// Block-uniform drain of the staged items once more than `thresh` are
// waiting (every thread of the workgroup calls it, at the same point of a
// loop whose trip count is the same for the whole workgroup).
__device__ inline void stage_drain(const Stage &sg, uint32_t thresh)
{
    __syncthreads();
    const uint32_t waiting = *sg.n;
    if (waiting <= thresh) return;
    const uint32_t n = min(waiting, sg.lcap);
    if (threadIdx.x == 0) *sg.base = atomicAdd(sg.count, n);
    __syncthreads();
    const uint32_t b = *sg.base;
    for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) {
        if (b + t < sg.cap) sg.list[b + t] = sg.items[t];
        else atomicOr(sg.overflow, 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) *sg.n = 0u;
    __syncthreads();
}

#define TSVD_STAGE_N(name, lcap_, list_, count_, cap_, overflow_)                             \
    __shared__ uint32_t name##_items[lcap_];                                                 \
    __shared__ uint32_t name##_n, name##_base;                                               \
    if (threadIdx.x == 0) name##_n = 0u;                                                     \
    __syncthreads();                                                                         \
    const Stage name{name##_items, &name##_n, &name##_base, (list_), (count_), (cap_), (overflow_), (lcap_)}
#define TSVD_STAGE(name, list_, count_, cap_, overflow_) TSVD_STAGE_N(name, STAGE, list_, count_, cap_, overflow_)

// Per-wave stages of a PTPB-thread workgroup (the barrier here is the only one).
#define TSVD_WAVE_STAGE(name, list_, count_, cap_, overflow_)                                \
    __shared__ uint32_t name##_items[PWPB][PSTAGE];                                          \
    __shared__ uint32_t name##_n[PWPB];                                                      \
    if ((threadIdx.x & 63) == 0) name##_n[threadIdx.x >> 6] = 0u;                            \
    __syncthreads();                                                                         \
    const WaveStage name{name##_items[threadIdx.x >> 6], &name##_n[threadIdx.x >> 6], (list_), (count_), \
                         (cap_), (overflow_)}

template <class T>
__global__ void k_zero_range(T *__restrict__ P, T *__restrict__ R, int32_t n, int32_t s_pad, int32_t s0, int32_t s1)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t w = s1 - s0;
    if (t >= (int64_t)n * w) return;
    int64_t v = t / w;
    int32_t s = s0 + (int32_t)(t - v * w);
    P[v * s_pad + s] = T(0);
    R[v * s_pad + s] = T(0);
}

// Static branch (ppr:2056-2068): r[src_s] = 1 and queue (src_s, g(s)).
template <class T>
__global__ void k_seed(T *__restrict__ R, const int32_t *__restrict__ src, int32_t s_pad, int32_t ngroups, int32_t s0,
                       int32_t s1, unsigned long long *__restrict__ cm, uint32_t *__restrict__ list,
                       uint32_t *__restrict__ count, uint32_t cap, uint32_t *__restrict__ overflow)
{
    int32_t s = s0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= s1) return;
    int32_t v = src[s];
    R[(int64_t)v * s_pad + s] = T(1);
    uint32_t item = (uint32_t)v * (uint32_t)ngroups + (uint32_t)(s >> 6);
    mark(1ull << (s & 63), item, cm, list, count, cap, overflow);
}

// Dynamic branch: one wave per (grown row i, source group g), lanes = sources.
template <class T>
__global__ void __launch_bounds__(TPB, 8) k_update(T *__restrict__ P, T *__restrict__ R, const int32_t *__restrict__ touched,
                                                int32_t ntouched, const int64_t *__restrict__ rowoff,
                                                const int32_t *__restrict__ adj, const int32_t *__restrict__ deg,
                                                const int32_t *__restrict__ former, int32_t s_pad, int32_t ngroups,
                                                int32_t s0, int32_t s1, double alpha, T rm, int two,
                                                unsigned long long *__restrict__ cm, uint32_t *__restrict__ list,
                                                uint32_t *__restrict__ count, uint32_t cap,
                                                uint32_t *__restrict__ overflow, uint8_t *__restrict__ changed)
{
    TSVD_STAGE(sg, list, count, cap, overflow);
    const int64_t wave = (int64_t)blockIdx.x * WPB + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wave < (int64_t)ntouched * ngroups) {
    const int32_t t = (int32_t)(wave / ngroups);
    const int32_t g = (int32_t)(wave - (int64_t)t * ngroups);
    const int32_t i = touched[t];
    const int32_t s = g * 64 + lane;
    const bool active = s >= s0 && s < s1;
    const int32_t f0 = former[i], d1 = deg[i];
    const int64_t base = rowoff[i];
    T *Pi = P + (int64_t)i * s_pad + s;
    T *Ri = R + (int64_t)i * s_pad + s;
    T pv = active ? *Pi : T(0);
    // The row's own decrements are summed here and applied with one atomic at
    // the end: nothing else lowers r[i], so its final value, and the check of
    // it below, are the same as with one atomic per slot.  The additions to
    // the new neighbours go out U at a time before their results are used.
    // Slot 0 is skipped (ppr:2078-2079).
    constexpr int U = 8;
    double dec = 0.0;
    const int32_t jb = max(f0, 1);
    for (int32_t blk = jb; blk < d1; blk += 64) {
        const int32_t cnt = min(64, d1 - blk);
        const int32_t mine = lane < cnt ? adj[base + blk + lane] : 0;
        for (int32_t q0 = 0; q0 < cnt; q0 += U) {
            int32_t to[U];
            T add[U], old[U];
#pragma unroll
            for (int q = 0; q < U; q++) {
                to[q] = __builtin_amdgcn_readlane(mine, min(q0 + q, 63));
                add[q] = T(0);
                const int32_t j = blk + q0 + q;
                if (active && q0 + q < cnt) {
                    pv = (T)((double)pv * ((double)(j + 1) / (double)j));
                    const double pj = (double)pv / (double)(j + 1) / alpha;
                    dec += pj;
                    add[q] = (T)((1.0 - alpha) * pj);
                }
            }
#pragma unroll
            for (int q = 0; q < U; q++) {
                old[q] = T(0);
                // a source with no reserve at row i adds exactly 0: no atomic (a
                // half-wave with no non-zero add then touches no cache line)
                if (active && q0 + q < cnt && add[q] != T(0))
                    old[q] = atomicAdd(R + (int64_t)to[q] * s_pad + s, add[q]);
            }
#pragma unroll
            for (int q = 0; q < U; q++) {
                if (q0 + q >= cnt) break;
                const bool cross = active & !beyond(old[q], rm, two) & beyond(old[q] + add[q], rm, two);
                const unsigned long long c = __ballot(cross);
                if (c && lane == 0 && deg[to[q]] > 0) stage_mark(sg, c, (uint32_t)to[q] * ngroups + g, cm);
            }
        }
    }
    bool self = false;
    if (active) {
        *Pi = pv;
        const T o = atomicAdd(Ri, (T)(-dec));
        self = beyond((T)(o - (T)dec), rm, two) && d1 > 0;
    }
    unsigned long long c2 = __ballot(self);
    if (c2 && lane == 0) stage_mark(sg, c2, (uint32_t)i * ngroups + g, cm);
    // the row's reserves were rescaled: the block norms re-read this item (a
    // group with no reserve at the row is unchanged, so its sums still hold)
    if (__ballot(active && pv != T(0)) != 0ull && lane == 0) changed[(int64_t)i * ngroups + g] = 1;
    }
    stage_flush(sg);
}

// ---- Insertion update of the n >= 1e7-node path, in batch order ----------
// All_initialization_and_push_final (ppr:6504-6533 forward, :6675-6704
// transpose) walks the batch in order for each source: j = max(former
// degree of `from`, edges of `from` seen so far); j == 0 sets r[from] = 0 and
// skips the edge (Q8); otherwise p *= (j+1)/j, r[from] -= p/(j+1)/alpha,
// r[to] += (1-alpha) p/(j+1)/alpha, in double.  For a row of former degree
// d > 0 the ratios telescope: each of its m new edges moves p0/(d alpha),
// with p0 the reserve before the batch, and p ends at p0 (d+m)/d.  A row of
// former degree 0 has never been pushed along, so p0 = 0 and its edges move
// nothing; what Q8 drops is the residue that row holds when its first edge
// comes, which includes what earlier edges of the batch added to it.  The
// passes below reproduce that order without walking it:
//   rows   (phase 0): r[i] = 0 for every grown row i of former degree 0;
//   adds            : every edge adds (1-alpha) p0[from]/(d alpha) to r[to],
//                     except an edge that precedes the first edge of a
//                     former-degree-0 `to` (its add was dropped with it);
//   rows   (phase 1): rows of former degree d > 0: r -= m p0/(d alpha),
//                     p = p0 (d+m)/d;
//   frontier        : both ends of every edge with r > rmax and an edge to
//                     push along join the push frontier (the reference queues
//                     them and tests r > rmax again when it pops them).
constexpr int32_t NO_EDGE = 0x7F7F7F7F;   // first[] of a row with no new edge

template <class T>
__global__ void k_lane_column(const T *__restrict__ X, int32_t n, int32_t s_pad, int32_t lane, double *__restrict__ out)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) out[v] = (double)X[v * s_pad + lane];
}

__global__ void k_first_edge(const int32_t *__restrict__ efrom, int64_t m, const int32_t *__restrict__ former,
                             int32_t *__restrict__ first, int reset)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const int32_t i = efrom[e];
    if (reset) first[i] = NO_EDGE;
    else if (former[i] == 0) atomicMin(first + i, (int32_t)e);
}

// The edge-order update kernels take EB edges (or rows) per wave and source
// group: lanes 0..EB-1 fetch the block's ids and what decides each one's use
// (clamped in-bounds loads, all in flight together), readlane hands them to
// the wave, and the EB residue loads and atomics are issued together.
constexpr int EB = 8;

__device__ inline int32_t rdl(int32_t v, int l) { return __builtin_amdgcn_readlane(v, l); }

template <class T>
__global__ void __launch_bounds__(TPB) k_edge_rows(T *__restrict__ P, T *__restrict__ R,
                                                   const int32_t *__restrict__ touched, int32_t ntouched,
                                                   const int32_t *__restrict__ deg, const int32_t *__restrict__ former,
                                                   int32_t s_pad, int32_t gcount, int32_t s0, int32_t s1,
                                                   double alpha, int phase)
{
    const int64_t wave = (int64_t)blockIdx.x * WPB + (threadIdx.x >> 6);
    const int64_t nblk = ((int64_t)ntouched + EB - 1) / EB;
    if (wave >= nblk * gcount) return;
    const int lane = threadIdx.x & 63;
    const int64_t tb = wave / gcount;
    const int32_t s = (int32_t)(wave - tb * gcount) * 64 + lane;
    const bool on = s >= s0 && s < s1;
    const int32_t tl = (int32_t)min(tb * EB + min(lane, EB - 1), (int64_t)ntouched - 1);
    const int32_t il = touched[tl];
    const int32_t fl = former[il], dl = deg[il];
    int32_t iq[EB], dq[EB], mq[EB];
    bool use[EB];
#pragma unroll
    for (int q = 0; q < EB; q++) {
        iq[q] = rdl(il, q);
        dq[q] = rdl(fl, q);
        mq[q] = rdl(dl, q) - dq[q];
        use[q] = on && tb * EB + q < ntouched;
    }
    if (phase == 0) {
#pragma unroll
        for (int q = 0; q < EB; q++)
            if (use[q] && dq[q] == 0) R[(int64_t)iq[q] * s_pad + s] = T(0);
        return;
    }
    T p0[EB], r0[EB];
#pragma unroll
    for (int q = 0; q < EB; q++) {
        p0[q] = P[(int64_t)iq[q] * s_pad + s];
        r0[q] = R[(int64_t)iq[q] * s_pad + s];
    }
#pragma unroll
    for (int q = 0; q < EB; q++) {
        if (!use[q] || dq[q] == 0) continue;
        const double c = (double)p0[q] / ((double)dq[q] * alpha);
        R[(int64_t)iq[q] * s_pad + s] = (T)((double)r0[q] - (double)mq[q] * c);
        P[(int64_t)iq[q] * s_pad + s] = (T)((double)p0[q] * ((double)(dq[q] + mq[q]) / (double)dq[q]));
    }
}

template <class T>
__global__ void __launch_bounds__(TPB) k_edge_adds(const T *__restrict__ P, T *__restrict__ R,
                                                   const int32_t *__restrict__ efrom, const int32_t *__restrict__ eto,
                                                   int64_t m, const int32_t *__restrict__ former,
                                                   const int32_t *__restrict__ first, int32_t s_pad, int32_t gcount,
                                                   int32_t s0, int32_t s1, double alpha)
{
    const int64_t wave = (int64_t)blockIdx.x * WPB + (threadIdx.x >> 6);
    const int64_t nblk = (m + EB - 1) / EB;
    if (wave >= nblk * gcount) return;
    const int lane = threadIdx.x & 63;
    const int64_t eb = wave / gcount;
    const int32_t s = (int32_t)(wave - eb * gcount) * 64 + lane;
    const bool on = s >= s0 && s < s1;
    const int64_t el = min(eb * EB + min(lane, EB - 1), m - 1);
    const int32_t fi = efrom[el], ti = eto[el];
    const int32_t df = former[fi], dt = former[ti], ft = first[ti];
    int32_t iq[EB], tq[EB], dq[EB];
    bool use[EB];
#pragma unroll
    for (int q = 0; q < EB; q++) {
        const int64_t e = eb * EB + q;
        iq[q] = rdl(fi, q);
        tq[q] = rdl(ti, q);
        dq[q] = rdl(df, q);
        const int32_t dto = rdl(dt, q), f = rdl(ft, q);
        // the edge moves nothing from a row of former degree 0, and nothing
        // into a former-degree-0 row before that row's first edge
        use[q] = on && e < m && dq[q] > 0 && !(dto == 0 && f != NO_EDGE && e < (int64_t)f);
    }
    T pv[EB];
#pragma unroll
    for (int q = 0; q < EB; q++) pv[q] = P[(int64_t)iq[q] * s_pad + s];
#pragma unroll
    for (int q = 0; q < EB; q++) {
        if (!use[q] || pv[q] == T(0)) continue;   // a zero reserve moves nothing
        const double c = (double)pv[q] / ((double)dq[q] * alpha);
        atomicAdd(R + (int64_t)tq[q] * s_pad + s, (T)((1.0 - alpha) * c));
    }
}

template <class T>
__global__ void __launch_bounds__(TPB) k_edge_frontier(const T *__restrict__ R, const int32_t *__restrict__ efrom,
                                                       const int32_t *__restrict__ eto, int64_t m,
                                                       const int32_t *__restrict__ deg, int32_t s_pad,
                                                       int32_t ngroups, int32_t gcount, int32_t s0, int32_t s1, T rm,
                                                       unsigned long long *__restrict__ cm, uint32_t *__restrict__ list,
                                                       uint32_t *__restrict__ count, uint32_t cap,
                                                       uint32_t *__restrict__ overflow)
{
    TSVD_STAGE(sg, list, count, cap, overflow);
    const int64_t wave = (int64_t)blockIdx.x * WPB + (threadIdx.x >> 6);
    const int64_t ends = 2 * m, nblk = (ends + EB - 1) / EB;
    if (wave < nblk * gcount) {
        const int lane = threadIdx.x & 63;
        const int64_t eb = wave / gcount;
        const int32_t g = (int32_t)(wave - eb * gcount);
        const int32_t s = g * 64 + lane;
        const bool on = s >= s0 && s < s1;
        // end 2e is edge e's source, 2e + 1 its target
        const int64_t xl = min(eb * EB + min(lane, EB - 1), ends - 1);
        const int32_t vl = (xl & 1) ? eto[xl >> 1] : efrom[xl >> 1];
        const int32_t dl = deg[vl];
        int32_t vq[EB];
        T rq[EB];
#pragma unroll
        for (int q = 0; q < EB; q++) {
            vq[q] = rdl(vl, q);
            rq[q] = R[(int64_t)vq[q] * s_pad + s];
        }
#pragma unroll
        for (int q = 0; q < EB; q++) {
            const bool out = on && eb * EB + q < ends && rdl(dl, q) > 0 && rq[q] > rm;
            const unsigned long long c = __ballot(out);
            if (c && lane == 0) stage_mark(sg, c, (uint32_t)vq[q] * (uint32_t)ngroups + (uint32_t)g, cm);
        }
    }
    stage_flush(sg);
}

// Negative residues, which storing the positive entries of the state drops,
// can only sit on the rows an edge-order update lowered: clamping those rows
// after the push leaves the state the full clamp would.
template <class T>
__global__ void __launch_bounds__(TPB) k_clamp_rows(T *__restrict__ R, const int32_t *__restrict__ touched,
                                                    int32_t ntouched, int32_t s_pad)
{
    const int64_t x = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (x >= (int64_t)ntouched * s_pad) return;
    const int64_t t = x / s_pad;
    T *r = R + (int64_t)touched[t] * s_pad + (x - t * s_pad);
    if (*r < T(0)) *r = T(0);
}

// Relax neighbours nb[j0, j1) of a popped (v, g) item: one coalesced 64-lane
// atomic add per neighbour; a lane whose residue crosses rmax upward is
// flagged, and the wave's ballot is OR-ed into the next round's mask of (u, g).
template <class T>
__device__ inline void relax_range(const gptr<const int32_t> __restrict__ nb, int32_t j0, int32_t j1, bool pushing,
                                   T delta, T *__restrict__ Rs, int32_t s_pad, const gptr<const int32_t> __restrict__ deg,
                                   uint32_t G,
                                   int32_t g, T rm, int two, int lane, unsigned long long *__restrict__ nm,
                                   const WaveStage &sg)
{
    // The neighbour ids of 64 consecutive slots arrive in one coalesced load
    // (one id per lane) and are broadcast with readlane, so each relaxation
    // costs one vector memory instruction (the atomic) instead of two.  Every
    // lane of the wave runs this function (j0, j1 are wave-uniform).
    constexpr int U = 8;
    for (int32_t base = j0; base < j1; base += 64) {
        const int32_t cnt = min(64, j1 - base);
        const int32_t mine = lane < cnt ? nb[base + lane] : 0;
        int32_t q0 = 0;
        for (; q0 + U <= cnt; q0 += U) {
            int32_t u[U];
            T o[U];
            bool c[U];
#pragma unroll
            for (int q = 0; q < U; q++) u[q] = __builtin_amdgcn_readlane(mine, q0 + q);
            // all U atomics are issued before any result is used, so a wave
            // keeps U of them in flight
            if (pushing) {
#pragma unroll
                for (int q = 0; q < U; q++) o[q] = atomicAdd(Rs + (int64_t)u[q] * s_pad, delta);
            } else {
#pragma unroll
                for (int q = 0; q < U; q++) o[q] = T(0);
            }
#pragma unroll
            for (int q = 0; q < U; q++)
                c[q] = pushing & !beyond(o[q], rm, two) & beyond(o[q] + delta, rm, two);
#pragma unroll
            for (int q = 0; q < U; q++) {
                const unsigned long long b = __ballot(c[q]);
                if (b && lane == 0 && deg[u[q]] > 0) wave_mark(sg, b, (uint32_t)u[q] * G + g, nm);
            }
        }
        if (q0 < cnt) {
            // the last cnt - q0 < U slots, the same way (slot q0 + q for q < r)
            const int32_t r = cnt - q0;
            int32_t u[U];
            T o[U];
            bool c[U];
#pragma unroll
            for (int q = 0; q < U; q++) u[q] = __builtin_amdgcn_readlane(mine, min(q0 + q, 63));
#pragma unroll
            for (int q = 0; q < U; q++) {
                o[q] = T(0);
                if (pushing && q < r) o[q] = atomicAdd(Rs + (int64_t)u[q] * s_pad, delta);
            }
#pragma unroll
            for (int q = 0; q < U; q++)
                c[q] = pushing & (q < r) & !beyond(o[q], rm, two) & beyond(o[q] + delta, rm, two);
#pragma unroll
            for (int q = 0; q < U; q++) {
                if (q >= r) break;
                const unsigned long long b = __ballot(c[q]);
                if (b && lane == 0 && deg[u[q]] > 0) wave_mark(sg, b, (uint32_t)u[q] * G + g, nm);
            }
        }
        // a block of 64 neighbours stages at most 64 items: keep that much room
        wave_drain(sg, PSTAGE - 64, lane);
    }
}

// Rows longer than HUB_CHUNK neighbours are split: the popping wave relaxes
// the first chunk and hands the rest, with its 64 pushed amounts, to
// k_push_hub (one wave per further chunk), so a hub does not serialise a
// round behind one wavefront.
// Chunk length: on graphs whose rounds are short (few items per wave) a round
// lasts as long as its longest chunk walk, so short chunks win (the 2.7M-node
// directed Chung-Lu graph: 64 -> 29 ms of push per snapshot, 256 -> 34 ms,
// 512 -> 40 ms); graphs with long rounds prefer 4096-8192, which the
// autotuner finds (2.1M-node R-MAT: 8192 is 19% faster than 256).  The hub
// buffers grow as the chunk shrinks, so without a caller setting the first
// push takes 64 or 128 when their worst case fits HUB_AUTO_BYTES, else
// HUB_CHUNK.
constexpr int HUB_CHUNK = 256;    // fallback; TSVD_HUB_CHUNK overrides it per PPR handle
constexpr double HUB_AUTO_BYTES = 4.0 * (1ull << 30);   // budget of the automatic chunk's hub buffers

// One synchronous push round: one wave per pending item pops its lanes and
// relaxes v's neighbours.  The atomic add returns the previous residue, so a
// lane whose residue crosses rmax upward is detected where it happens; the
// wave's ballot of crossing lanes is OR-ed into the next round's mask of
// (u, g) and the first setter appends the item.  Push counters are summed per
// workgroup in LDS before one global update.
//
// The round runs on a fixed grid and reads its item count from the device
// (*count_in, written by the previous round), so the host enqueues several
// rounds without waiting: workgroups stride over the items, and the round
// zeroes the counters that the round after it appends to (*zero_next, and
// the other set of split-row counters, zero_hub[0..1]); *rounds counts the
// rounds that had work.
template <class T>
__global__ void __launch_bounds__(PTPB, 8) k_push_round(T *__restrict__ P, T *__restrict__ R,
                                                    const uint32_t *__restrict__ items,
                                                    const uint32_t *__restrict__ count_in, uint32_t in_cap,
                                                    uint32_t *__restrict__ zero_next,
                                                    uint32_t *__restrict__ zero_hub, uint32_t *__restrict__ rounds,
                                                    unsigned long long *__restrict__ cm,
                                                    unsigned long long *__restrict__ nm, uint32_t *__restrict__ next,
                                                    uint32_t *__restrict__ next_count, uint32_t cap,
                                                    uint32_t *__restrict__ overflow,
                                                    const PushCsr *__restrict__ csr,
                                                    int32_t s_pad, int32_t ngroups, double alpha, T rm, int two,
                                                    int directed, T ride_min, int32_t ride_s0, int32_t ride_s1,
                                                    unsigned long long *__restrict__ stats,
                                                    uint32_t *__restrict__ hub_item, T *__restrict__ hub_delta,
                                                    unsigned long long *__restrict__ hub_chunks,
                                                    uint32_t *__restrict__ hub_counts, uint32_t hub_cap,
                                                    uint32_t chunk_cap, int32_t hub_chunk,
                                                    uint8_t *__restrict__ changed)
{
    // [0] pops, [1] relaxations (lanes x degree), [2] items, [3] row visits,
    // [4] relaxations done here (split rows hand the rest to k_push_hub),
    // [5] 128-byte lines the row's atomics touch (row visits x lines holding
    //     a pushing lane: the L2 atomic units work per line, not per lane)
    const gptr<const int64_t> __restrict__ rowoff = gbl(csr->rowoff);
    const gptr<const int32_t> __restrict__ adj = gbl(csr->adj);
    const gptr<const int32_t> __restrict__ deg = gbl(csr->deg);
    __shared__ unsigned long long sstat[6];
    if (threadIdx.x < 6) sstat[threadIdx.x] = 0ull;
    const uint32_t nitems = min(*count_in, in_cap);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *zero_next = 0u;
        zero_hub[0] = 0u;
        zero_hub[1] = 0u;
        if (nitems > 0) atomicAdd(rounds, 1u);
    }
    TSVD_WAVE_STAGE(sg, next, next_count, cap, overflow);
    const int lane = threadIdx.x & 63;
    // wave-strided: each wave walks its own items, with no workgroup barrier
    const uint32_t nwaves = gridDim.x * PWPB;
    for (uint32_t w = blockIdx.x * PWPB + (threadIdx.x >> 6); w < nitems; w += nwaves) {
    {
        const uint32_t item = items[w];
        const int32_t v = (int32_t)(item / (uint32_t)ngroups);
        const int32_t g = (int32_t)(item - (uint32_t)v * (uint32_t)ngroups);
        const unsigned long long m = cm[item];
        if (lane == 0) cm[item] = 0ull;
        // a popped lane may change its reserve (so may riders, which need a
        // pushing lane): the block norms re-read this item
        if (lane == 0 && m != 0ull) changed[item] = 1;
        const int32_t s = g * 64 + lane;
        const bool active = (m >> lane) & 1ull;
        const int32_t dv = deg[v];
        T *Rv = R + (int64_t)v * s_pad + s;
        T *Pv = P + (int64_t)v * s_pad + s;
        T x = T(0);
        bool pushing = false, again = false;
        if (active) {
            x = xchg(Rv, T(0));
            if (!beyond(x, rm, two)) {
                const T old = atomicAdd(Rv, x);                  // within the threshold: give it back
                again = !beyond(old, rm, two) && beyond(old + x, rm, two) && dv > 0;
            } else if (dv == 0) {
                if (!directed) *Pv = (T)(alpha * (double)x);     // ppr:1431-1441
                atomicAdd(Rv, x);                                // residue kept, not re-queued
            } else {
                pushing = true;
                *Pv = (T)((double)*Pv + alpha * (double)x);
            }
        }
        const unsigned long long ca = __ballot(again);
        if (ca && lane == 0) wave_mark(sg, ca, item, nm);
        unsigned long long pm = __ballot(pushing);
        if (pm != 0ull && ride_min < rm && dv > 0) {
            // The row is walked anyway: lanes whose residue is positive but at or
            // below rmax ride along (same atomics, no extra traffic).  They only
            // lower residues further, so the invariant and r <= rmax still hold.
            bool back = false;
            if (!active && s >= ride_s0 && s < ride_s1 && *Rv > ride_min) {
                x = xchg(Rv, T(0));
                if (x > T(0)) {
                    pushing = true;
                    *Pv = (T)((double)*Pv + alpha * (double)x);
                } else {
                    // a negative push landed between the read and the exchange
                    // (two-sided only): give it back, re-queue on a crossing
                    const T old = atomicAdd(Rv, x);
                    back = !beyond(old, rm, two) && beyond(old + x, rm, two);
                }
            }
            const unsigned long long cb = __ballot(back);
            if (cb && lane == 0) wave_mark(sg, cb, item, nm);
            pm = __ballot(pushing);
        }
        if (pm != 0ull) {
            if (lane == 0) {
                atomicAdd(&sstat[0], (unsigned long long)__popcll(pm));
                atomicAdd(&sstat[1], (unsigned long long)__popcll(pm) * (unsigned long long)dv);
                atomicAdd(&sstat[2], 1ull);
                atomicAdd(&sstat[3], (unsigned long long)dv);
                constexpr int per_line = 128 / (int)sizeof(T);   // lanes per 128-byte line
                unsigned lines = 0;
                for (int q = 0; q < 64; q += per_line)
                    lines += ((pm >> q) & ((1ull << per_line) - 1ull)) != 0ull;
                atomicAdd(&sstat[5], (unsigned long long)lines * (unsigned long long)dv);
            }
            const T delta = pushing ? (T)((1.0 - alpha) * (double)x / (double)dv) : T(0);
            const gptr<const int32_t> nb = adj + rowoff[v];
            int32_t jend = dv;
            if (dv > hub_chunk) {
                const uint32_t more = (uint32_t)((dv - 1) / hub_chunk);   // chunks after the first
                uint32_t slot = 0xFFFFFFFFu, cbase = 0;
                if (lane == 0) {
                    slot = atomicAdd(hub_counts + 0, 1u);
                    if (slot < hub_cap) {
                        cbase = atomicAdd(hub_counts + 1, more);
                        if (cbase + more > chunk_cap) {
                            atomicOr(overflow, 2u);
                            slot = 0xFFFFFFFFu;
                        }
                    } else {
                        slot = 0xFFFFFFFFu;
                    }
                }
                slot = __shfl(slot, 0, 64);
                cbase = __shfl(cbase, 0, 64);
                if (slot != 0xFFFFFFFFu) {
                    if (lane == 0) hub_item[slot] = item;
                    hub_delta[(int64_t)slot * 64 + lane] = delta;
                    for (uint32_t c = lane; c < more; c += 64)
                        hub_chunks[cbase + c] = ((unsigned long long)slot << 32) | (unsigned long long)(c + 1);
                    jend = hub_chunk;
                }
            }
            if (lane == 0) atomicAdd(&sstat[4], (unsigned long long)__popcll(pm) * (unsigned long long)jend);
            relax_range(nb, 0, jend, pushing, delta, R + s, s_pad, deg, (uint32_t)ngroups, g, rm, two, lane, nm, sg);
        }
    }
    wave_drain(sg, PSTAGE / 2, lane);
    }
    __shared__ uint32_t spos[PWPB + 1];
    wave_flush_all(sg, spos, lane);
    if (threadIdx.x < 6 && sstat[threadIdx.x]) atomicAdd(stat_slot(stats) + threadIdx.x, sstat[threadIdx.x]);
}


// Further chunks of split rows (see HUB_CHUNK): one wave per chunk.
template <class T>
__global__ void __launch_bounds__(PTPB) k_push_hub(T *__restrict__ R, const uint32_t *__restrict__ hub_item,
                                                  const T *__restrict__ hub_delta,
                                                  const unsigned long long *__restrict__ hub_chunks,
                                                  const uint32_t *__restrict__ chunk_count, uint32_t chunk_cap,
                                                  unsigned long long *__restrict__ nm, uint32_t *__restrict__ next,
                                                  uint32_t *__restrict__ next_count, uint32_t cap,
                                                  uint32_t *__restrict__ overflow,
                                                  const PushCsr *__restrict__ csr,
                                                  int32_t s_pad, int32_t ngroups, T rm, int two, int32_t hub_chunk)
{
    const gptr<const int64_t> __restrict__ rowoff = gbl(csr->rowoff);
    const gptr<const int32_t> __restrict__ adj = gbl(csr->adj);
    const gptr<const int32_t> __restrict__ deg = gbl(csr->deg);
    TSVD_WAVE_STAGE(sg, next, next_count, cap, overflow);
    const uint32_t nchunks = min(*chunk_count, chunk_cap);
    const int lane = threadIdx.x & 63;
    const uint32_t nwaves = gridDim.x * PWPB;
    for (uint32_t w = blockIdx.x * PWPB + (threadIdx.x >> 6); w < nchunks; w += nwaves) {
    {
        const unsigned long long e = hub_chunks[w];
        const uint32_t slot = (uint32_t)(e >> 32), c = (uint32_t)(e & 0xFFFFFFFFull);
        const uint32_t item = hub_item[slot];
        const int32_t v = (int32_t)(item / (uint32_t)ngroups);
        const int32_t g = (int32_t)(item - (uint32_t)v * (uint32_t)ngroups);
        const T delta = hub_delta[(int64_t)slot * 64 + lane];
        const int32_t dv = deg[v];
        const int32_t j0 = (int32_t)c * hub_chunk, j1 = min(dv, j0 + hub_chunk);
        relax_range(adj + rowoff[v], j0, j1, delta != T(0), delta, R + (int64_t)g * 64 + lane, s_pad, deg,
                    (uint32_t)ngroups, g, rm, two, lane, nm, sg);
    }
    wave_drain(sg, PSTAGE / 2, lane);
    }
    __shared__ uint32_t spos[PWPB + 1];
    wave_flush_all(sg, spos, lane);
}

// ---------------------------------------------------------------- pull rounds
// When the frontier is large, a round runs as pop -> pull: the pop publishes
// each popped lane's pushed amount (1-alpha) r / deg in delta[v] and flags the
// group in pm[v]; every vertex u then sums delta[v][g] over its pull
// neighbours (in-neighbours of the push direction) for each flagged group,
// accumulating in LDS without atomics, and adds the sums to r[u] once.  Hub
// rows are split into chunks of PULL_CHUNK neighbours; a chunk of a split row
// adds with atomics (crossing detected from the returned value), an unsplit
// row is owned by one wave (plain read-modify-write, absolute test).
constexpr int PULL_CHUNK = 2048;

template <class T>
__global__ void __launch_bounds__(TPB) k_pull_pop(T *__restrict__ P, T *__restrict__ R, T *__restrict__ D,
                                                  unsigned long long *__restrict__ pm,
                                                  const uint32_t *__restrict__ items, uint32_t nitems,
                                                  unsigned long long *__restrict__ cm,
                                                  const int32_t *__restrict__ deg, int32_t s_pad, int32_t ngroups,
                                                  double alpha, T rm, int directed,
                                                  unsigned long long *__restrict__ stats)
{
    __shared__ unsigned long long sstat[4];
    if (threadIdx.x < 4) sstat[threadIdx.x] = 0ull;
    __syncthreads();
    const uint32_t w = blockIdx.x * WPB + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w < nitems) {
        const uint32_t item = items[w];
        const int32_t v = (int32_t)(item / (uint32_t)ngroups);
        const int32_t g = (int32_t)(item - (uint32_t)v * (uint32_t)ngroups);
        const unsigned long long m = cm[item];
        if (lane == 0) cm[item] = 0ull;
        const int64_t at = (int64_t)v * s_pad + g * 64 + lane;
        const int32_t dv = deg[v];
        T d = T(0);
        bool pushing = false;
        if ((m >> lane) & 1ull) {
            const T x = R[at];
            if (x > rm) {
                if (dv == 0) {
                    if (!directed) P[at] = (T)(alpha * (double)x);   // ppr:1431-1441, residue kept
                } else {
                    pushing = true;
                    R[at] = T(0);
                    P[at] = (T)((double)P[at] + alpha * (double)x);
                    d = (T)((1.0 - alpha) * (double)x / (double)dv);
                }
            }
        }
        D[at] = d;
        const unsigned long long pmask = __ballot(pushing);
        if (pmask && lane == 0) {
            atomicOr(pm + v, 1ull << g);
            atomicAdd(&sstat[0], (unsigned long long)__popcll(pmask));
            atomicAdd(&sstat[1], (unsigned long long)__popcll(pmask) * (unsigned long long)dv);
            atomicAdd(&sstat[2], 1ull);
            atomicAdd(&sstat[3], (unsigned long long)dv);
        }
    }
    __syncthreads();
    if (threadIdx.x < 4 && sstat[threadIdx.x]) atomicAdd(stat_slot(stats) + threadIdx.x, sstat[threadIdx.x]);
}

template <class T>
__global__ void __launch_bounds__(TPB) k_pull(T *__restrict__ R, const T *__restrict__ D,
                                              const unsigned long long *__restrict__ pm,
                                              const uint32_t *__restrict__ wl_u, const uint32_t *__restrict__ wl_c,
                                              uint32_t nwork, const int64_t *__restrict__ rowoff,
                                              const int32_t *__restrict__ adj, const int32_t *__restrict__ deg_pull,
                                              const int32_t *__restrict__ deg_push, int32_t s_pad, int32_t ngroups,
                                              T rm, unsigned long long *__restrict__ nm, uint32_t *__restrict__ next,
                                              uint32_t *__restrict__ next_count, uint32_t cap,
                                              uint32_t *__restrict__ overflow)
{
    extern __shared__ unsigned char smem[];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    T *acc = (T *)smem + (size_t)wv * ngroups * 64;
    const uint32_t w = blockIdx.x * WPB + wv;
    if (w >= nwork) return;
    const int32_t u = (int32_t)wl_u[w];
    const int32_t c = (int32_t)wl_c[w];
    const int32_t du = deg_pull[u];
    const int32_t j0 = c * PULL_CHUNK, j1 = min(du, j0 + PULL_CHUNK);
    const bool split = du > PULL_CHUNK;
    for (int g = 0; g < ngroups; g++) acc[g * 64 + lane] = T(0);
    const int32_t *nb = adj + rowoff[u];
    unsigned long long uni = 0ull;
    int32_t j = j0;
    for (; j + 4 <= j1; j += 4) {
        const int32_t v0 = nb[j], v1 = nb[j + 1], v2 = nb[j + 2], v3 = nb[j + 3];
        unsigned long long m0 = pm[v0], m1 = pm[v1], m2 = pm[v2], m3 = pm[v3];
        uni |= m0 | m1 | m2 | m3;
        while (m0) { const int g = __ffsll((long long)m0) - 1; m0 &= m0 - 1ull; acc[g * 64 + lane] += D[(int64_t)v0 * s_pad + g * 64 + lane]; }
        while (m1) { const int g = __ffsll((long long)m1) - 1; m1 &= m1 - 1ull; acc[g * 64 + lane] += D[(int64_t)v1 * s_pad + g * 64 + lane]; }
        while (m2) { const int g = __ffsll((long long)m2) - 1; m2 &= m2 - 1ull; acc[g * 64 + lane] += D[(int64_t)v2 * s_pad + g * 64 + lane]; }
        while (m3) { const int g = __ffsll((long long)m3) - 1; m3 &= m3 - 1ull; acc[g * 64 + lane] += D[(int64_t)v3 * s_pad + g * 64 + lane]; }
    }
    for (; j < j1; j++) {
        const int32_t v = nb[j];
        unsigned long long m = pm[v];
        uni |= m;
        while (m) { const int g = __ffsll((long long)m) - 1; m &= m - 1ull; acc[g * 64 + lane] += D[(int64_t)v * s_pad + g * 64 + lane]; }
    }
    const bool can_push = deg_push[u] > 0;
    while (uni) {
        const int g = __ffsll((long long)uni) - 1;
        uni &= uni - 1ull;
        const T a = acc[g * 64 + lane];
        T *r = R + (int64_t)u * s_pad + g * 64 + lane;
        bool cross;
        if (split) {
            const T old = atomicAdd(r, a);
            cross = old <= rm && old + a > rm;
        } else {
            const T nv = *r + a;
            *r = nv;
            cross = nv > rm;
        }
        const unsigned long long b = __ballot(cross);
        if (b && lane == 0 && can_push) mark(b, (uint32_t)u * (uint32_t)ngroups + g, nm, next, next_count, cap, overflow);
    }
}

__global__ void k_pull_clear(unsigned long long *__restrict__ pm, const uint32_t *__restrict__ items, uint32_t nitems,
                             int32_t ngroups)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nitems) pm[items[t] / (uint32_t)ngroups] = 0ull;
}

__global__ void k_chunk_count(const int32_t *__restrict__ deg, int32_t n, int32_t *__restrict__ cnt)
{
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) cnt[v] = (deg[v] + PULL_CHUNK - 1) / PULL_CHUNK;
    if (v == n) cnt[v] = 0;
}

__global__ void k_chunk_fill(const int32_t *__restrict__ cnt, const int32_t *__restrict__ off, int32_t n,
                             uint32_t *__restrict__ wl_u, uint32_t *__restrict__ wl_c)
{
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int32_t o = off[v];
    for (int32_t c = 0; c < cnt[v]; c++) {
        wl_u[o + c] = (uint32_t)v;
        wl_c[o + c] = (uint32_t)c;
    }
}

template <class T>
void refresh_t(tsvd_ppr *s, int mode, int d, int32_t s0, int32_t s1)
{
    tsvd_graph *g = s->g;
    hipStream_t st = g->stream;
    T *P = (T *)s->p(d), *R = (T *)s->r(d);
    const uint32_t cap = (uint32_t)s->F[d].n;
    uint32_t *cnt = s->counters.p;
    if (mode == TSVD_REFRESH_INIT) {
        s->touched_all = true;
        if (s0 == 0 && s1 == s->nsrc) {
            TSVD_HIP(hipMemsetAsync(P, 0, s->P[d].bytes(), st));
            TSVD_HIP(hipMemsetAsync(R, 0, s->R[d].bytes(), st));
            TSVD_HIP(hipMemsetAsync(s->cm[d].p, 0, s->cm[d].bytes(), st));
            TSVD_HIP(hipMemsetAsync(cnt + d, 0, sizeof(uint32_t), st));
        } else {
            const int64_t tot = (int64_t)s->n * (s1 - s0);
            hipLaunchKernelGGL(k_zero_range<T>, dim3((unsigned)ceil_div(tot, TPB)), dim3(TPB), 0, st, P, R, s->n,
                               s->s_pad, s0, s1);
        }
        hipLaunchKernelGGL(k_seed<T>, dim3((unsigned)ceil_div(s1 - s0, TPB)), dim3(TPB), 0, st, R, s->sources.p,
                           s->s_pad, s->ngroups, s0, s1, s->cm[d].p, s->F[d].p, cnt + d, cap, cnt + 3);
    } else {
        const GraphDir &D = g->dir[d];
        if (D.ntouched > 0) {
            const int64_t waves = (int64_t)D.ntouched * s->ngroups;
            // algorithmic bytes per (new slot, source): p RMW, r[from] and r[to] RMW,
            // and the matrix entry of the reference (8 B each in FP32 terms)
            if (g_profile)
                prof_add_work("ppr_update", 4.0 * sizeof(T) * 2.0 * (double)D.appended * (s1 - s0), 0.0);
            PROF("ppr_update", st);
            hipLaunchKernelGGL(k_update<T>, dim3((unsigned)ceil_div(waves, WPB)), dim3(TPB), 0, st, P, R,
                               D.touched.p, D.ntouched, D.rowoff.p, D.adj.p, D.deg.p, D.former.p, s->s_pad,
                               s->ngroups, s0, s1, s->alpha, (T)s->rmax, s->two_sided, s->cm[d].p, s->F[d].p,
                               cnt + d, cap, cnt + 3, s->touched.p);
        }
    }
    TSVD_HIP(hipGetLastError());
    TSVD_HIP(hipStreamSynchronize(st));
}


// The pull adjacency of push direction d: in-edges for the forward push,
// out-edges for the transpose push, the adjacency itself when undirected.
static const GraphDir &pull_dir(tsvd_ppr *s, int d)
{
    return s->g->dir[s->directed ? 1 - d : 0];
}

static void build_worklist(tsvd_ppr *s, int d)
{
    tsvd_graph *g = s->g;
    if (s->wl_version[d] == g->version) return;
    hipStream_t st = g->stream;
    const GraphDir &Q = pull_dir(s, d);
    const int32_t n = g->n;
    DBuf<int32_t> off((size_t)n + 1);
    s->wl_cnt.ensure((size_t)n + 1);
    hipLaunchKernelGGL(k_chunk_count, dim3((unsigned)ceil_div(n + 1, TPB)), dim3(TPB), 0, st, Q.deg.p, n,
                       s->wl_cnt.p);
    size_t tmp = 0;
    TSVD_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, s->wl_cnt.p, off.p, n + 1, st));
    DBuf<uint8_t> t(tmp + 16);
    TSVD_HIP(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, s->wl_cnt.p, off.p, n + 1, st));
    int32_t total = 0;
    TSVD_HIP(hipMemcpyAsync(&total, off.p + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    TSVD_HIP(hipStreamSynchronize(st));
    s->wl_u[d].grow((size_t)std::max(total, 1));
    s->wl_c[d].grow((size_t)std::max(total, 1));
    hipLaunchKernelGGL(k_chunk_fill, dim3((unsigned)ceil_div(n, TPB)), dim3(TPB), 0, st, s->wl_cnt.p, off.p, n,
                       s->wl_u[d].p, s->wl_c[d].p);
    TSVD_HIP(hipGetLastError());
    TSVD_HIP(hipStreamSynchronize(st));
    s->wl_n[d] = total;
    s->wl_version[d] = g->version;
}

// Workgroups of a device-driven round.  The host knows the item count of the
// first round of each batch it enqueues, not of the later ones: the grid
// gives that count 1.5x headroom at up to `items_per_wave` items per wave and
// never less than 8 waves per compute unit; a round that outgrows it strides
// over its items.
static uint32_t cu_count()
{
    static int cus[64] = {};
    int dev = 0;
    TSVD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) dev = 0;
    if (cus[dev] == 0) {
        int c = 0;
        TSVD_HIP(hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev));
        cus[dev] = std::max(c, 1);
    }
    return (uint32_t)cus[dev];
}

// A launch's work-item count (workgroups x PTPB) must stay below 2^32; larger
// rounds stride over their items.
constexpr uint64_t MAX_ROUND_GRID = (1ull << 31) / PTPB;

static uint32_t round_grid(uint32_t count_hint, int32_t per_wave)
{
    // a small round keeps one item per wave up to two full occupancies of
    // the device (64 waves per compute unit): fewer waves would leave it idle
    const int64_t headroom = (int64_t)count_hint + count_hint / 2;
    const int64_t fill = (int64_t)cu_count() * 64;
    per_wave = (int32_t)std::max<int64_t>(1, std::min<int64_t>(per_wave, headroom / fill));
    const uint64_t want = (uint64_t)ceil_div(headroom, (int64_t)PWPB * per_wave);
    return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(want, (uint64_t)cu_count() * 8 / PWPB), MAX_ROUND_GRID);
}

// Rounds until no item is pending.  Push rounds are device-driven: each reads
// its item count from a ring of three device counters (the count it was
// given, the one it appends to, and the one it clears for the round after),
// so the host enqueues several rounds, then reads one count.  Pull rounds
// (off unless TSVD_PULL_FRAC is set) are chosen by the frontier size, so with
// them enabled the host reads the count after every round.
constexpr int ROUNDS_PER_SYNC = 8;

// Push rounds replayed from a captured hipGraph: GRAPH_ROUNDS rounds (round
// and split-row kernels) per replay.  Six is the period of the three-slot
// ring and of the double-buffered masks, lists and split-row counters, so a
// replay starts from the same buffers as the capture did and one graph
// serves every block of a push.  A graph is kept per launch shape (the round
// grid is rounded up to a power of two, so a few graphs cover a push) and
// captured again when any launch argument changes (a buffer grown or moved).
constexpr int GRAPH_ROUNDS = 6;
constexpr size_t GRAPH_CACHE = 16;

static uint32_t graph_grid(uint32_t count_hint, int32_t per_wave)
{
    const uint32_t g = round_grid(count_hint, per_wave);
    uint32_t p = 1;
    while (p < g) p <<= 1;
    return std::min<uint32_t>(p, MAX_ROUND_GRID);
}

template <class T>
void push_t(tsvd_ppr *s, int d, int32_t s0, int32_t s1)
{
    tsvd_graph *g = s->g;
    hipStream_t st = g->stream;
    const GraphDir &D = g->dir[d];
    T *P = (T *)s->p(d), *R = (T *)s->r(d);
    uint32_t *cnt = s->counters.p;
    uint32_t *h = s->h_counters;
    const uint32_t cap = (uint32_t)s->fnext.n;
    TSVD_HIP(hipMemsetAsync(s->stats.p, 0, s->stats.bytes(), st));
    // ring slots cnt[8], cnt[9], split-row sets cnt[4..5] / cnt[6..7], rounds cnt[10]
    TSVD_HIP(hipMemsetAsync(cnt + 4, 0, 7 * sizeof(uint32_t), st));
    // the direction's CSR arrays (an append may have moved them), read by the
    // push kernels through this descriptor so that captured graphs stay valid
    PushCsr *csr = s->csr_ref.p + d;
    s->h_csr_ref[d] = PushCsr{D.rowoff.p, D.adj.p, D.deg.p};
    TSVD_HIP(hipMemcpyAsync(csr, s->h_csr_ref + d, sizeof(PushCsr), hipMemcpyHostToDevice, st));
    TSVD_HIP(hipMemcpyAsync(h, cnt, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    TSVD_HIP(hipStreamSynchronize(st));
    TSVD_REQUIRE(h[3] == 0, "frontier overflow");
    uint32_t count = h[d];
    s->last_pull_rounds[d] = 0;
    uint32_t *ring[3] = {cnt + d, cnt + 8, cnt + 9};
    uint32_t *hubset[2] = {cnt + 4, cnt + 6};
    const bool pull_enabled = !s->two_sided && s->ngroups <= 64 && s->pull_frac > 0.0;
    bool graphs = s->use_graphs != 0 && !pull_enabled;
    if (s->hub_auto) {
        // no setting from the caller: the smallest of 64 and 128 neighbours
        // whose worst-case hub buffers fit in HUB_AUTO_BYTES, else HUB_CHUNK
        // (short chunks shorten the longest walk of a round; measured on the
        // 2.7M-node directed Chung-Lu graph: 64 -> 29 ms of push per snapshot,
        // 128 -> 31 ms, 256 -> 34 ms, 512 -> 40 ms)
        s->hub_auto = false;
        for (const int32_t c : {64, 128}) {
            const int64_t hc = std::min<int64_t>((int64_t)cap, (int64_t)s->ngroups * (ceil_div(D.cap_total, c) + 1));
            if ((double)hc * (4.0 + 8.0 + 64.0 * sizeof(T)) <= HUB_AUTO_BYTES) {
                s->hub_chunk = c;
                break;
            }
        }
    }
    // split rows: at most one chunk list entry per (hub row, group) per hub_chunk
    // neighbours (rounded up to a power of two in graph mode, so that the
    // capacities, which are launch arguments, change rarely)
    uint32_t hub_cap = (uint32_t)std::min<int64_t>((int64_t)cap, (int64_t)s->ngroups *
                                                   (ceil_div(D.cap_total, s->hub_chunk) + 1));
    if (graphs) {
        uint64_t p2 = 1;
        while (p2 < hub_cap) p2 <<= 1;
        hub_cap = (uint32_t)std::min<uint64_t>(p2, cap);
    }
    const uint32_t chunk_cap = hub_cap;
    s->hub_item.grow(std::max<uint32_t>(hub_cap, 1));
    s->hub_delta.grow((size_t)std::max<uint32_t>(hub_cap, 1) * 64 * sizeof(T));
    s->hub_chunks.grow(std::max<uint32_t>(chunk_cap, 1));
    const T ride_min = (T)(std::min(s->ride_frac, 1.0) * s->rmax);
    const uint32_t hub_grid = cu_count() * 32 / PWPB;   // 32 waves per compute unit
    auto launch_round = [&](uint64_t ri, uint32_t grid) {
        uint32_t *in = ring[ri % 3], *next = ring[(ri + 1) % 3], *ahead = ring[(ri + 2) % 3];
        uint32_t *hub = hubset[ri % 2], *hub_ahead = hubset[(ri + 1) % 2];
        hipLaunchKernelGGL(k_push_round<T>, dim3(grid), dim3(PTPB), 0, st, P, R, s->F[d].p, in, cap, ahead,
                           hub_ahead, cnt + 10, s->cm[d].p, s->nm.p, s->fnext.p, next, cap, cnt + 3, csr,
                           s->s_pad, s->ngroups, s->alpha, (T)s->rmax, s->two_sided, s->directed,
                           ride_min, s0, s1, s->stats.p, s->hub_item.p, (T *)s->hub_delta.p, s->hub_chunks.p, hub,
                           hub_cap, chunk_cap, s->hub_chunk, s->touched.p);
    };
    auto launch_hub = [&](uint64_t ri) {
        uint32_t *next = ring[(ri + 1) % 3], *hub = hubset[ri % 2];
        hipLaunchKernelGGL(k_push_hub<T>, dim3(hub_grid), dim3(PTPB), 0, st, R, s->hub_item.p,
                           (const T *)s->hub_delta.p, s->hub_chunks.p, hub + 1, chunk_cap, s->nm.p, s->fnext.p,
                           next, cap, cnt + 3, csr, s->s_pad, s->ngroups, (T)s->rmax, s->two_sided,
                           s->hub_chunk);
    };
    // The graph for this launch shape: found, or captured from the same
    // launches (their host-side buffer swaps come back to the start after
    // GRAPH_ROUNDS rounds).  nullptr when capture is not possible.
    auto block_graph = [&](uint64_t i0, uint32_t grid) -> PushGraph * {
        const bool prof = g_profile;
        auto bits = [](double x) {
            uint64_t u;
            memcpy(&u, &x, sizeof(u));
            return u;
        };
        const std::vector<uint64_t> key = {
            sizeof(T), (uint64_t)d, grid, hub_grid, (uint64_t)P, (uint64_t)R, (uint64_t)s->F[d].p,
            (uint64_t)s->fnext.p, (uint64_t)s->cm[d].p, (uint64_t)s->nm.p, (uint64_t)cnt, cap, hub_cap,
            (uint64_t)s->hub_item.p, (uint64_t)s->hub_delta.p, (uint64_t)s->hub_chunks.p, (uint64_t)s->stats.p,
            (uint64_t)s->touched.p, (uint64_t)csr,
            (uint64_t)s->s_pad, (uint64_t)s->ngroups, bits(s->alpha), bits(s->rmax), (uint64_t)s->two_sided,
            (uint64_t)s->directed, bits((double)ride_min), (uint64_t)s0, (uint64_t)s1, (uint64_t)s->hub_chunk,
            (uint64_t)prof, i0 % GRAPH_ROUNDS};
        for (auto &pg : s->graphs)
            if (pg->key == key) {
                pg->used = ++s->graph_clock;
                return pg.get();
            }
        auto pg = std::make_unique<PushGraph>();
        pg->key = key;
        if (prof) {
            pg->ev.resize(2 * GRAPH_ROUNDS + 1, nullptr);
            for (auto &e : pg->ev)
                if (hipEventCreate(&e) != hipSuccess) {
                    (void)hipGetLastError();
                    return nullptr;
                }
        }
        hipGraph_t graph = nullptr;
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        bool ok = true;
        for (int k = 0; k < GRAPH_ROUNDS; k++) {
            if (prof) ok &= hipEventRecordWithFlags(pg->ev[2 * k], st, hipEventRecordExternal) == hipSuccess;
            launch_round(i0 + k, grid);
            if (prof) ok &= hipEventRecordWithFlags(pg->ev[2 * k + 1], st, hipEventRecordExternal) == hipSuccess;
            launch_hub(i0 + k);
            std::swap(s->cm[d], s->nm);
            std::swap(s->F[d], s->fnext);
        }
        if (prof) ok &= hipEventRecordWithFlags(pg->ev[2 * GRAPH_ROUNDS], st, hipEventRecordExternal) == hipSuccess;
        ok &= hipGetLastError() == hipSuccess;
        const hipError_t ec = hipStreamEndCapture(st, &graph);
        ok &= ec == hipSuccess && graph != nullptr;
        if (ok) ok = hipGraphInstantiate(&pg->exec, graph, nullptr, nullptr, 0) == hipSuccess;
        if (graph) (void)hipGraphDestroy(graph);
        if (!ok) {
            const hipError_t e = hipGetLastError();
            fprintf(stderr, "treesvd: push graph capture failed (grid %u, %s); rounds are launched one by one\n",
                    grid, hipGetErrorString(e != hipSuccess ? e : ec));
            return nullptr;
        }
        s->graph_captures++;
        pg->used = ++s->graph_clock;
        if (s->graphs.size() >= GRAPH_CACHE) {
            auto lru = std::min_element(s->graphs.begin(), s->graphs.end(),
                                        [](const auto &a, const auto &b) { return a->used < b->used; });
            s->graphs.erase(lru);
        }
        s->graphs.push_back(std::move(pg));
        return s->graphs.back().get();
    };
    Timer tm;
    tm.start(st);
    uint64_t i = 0;           // round index (selects the ring slots)
    while (count > 0) {
        PushGraph *pg = nullptr;
        if (graphs) {
            pg = block_graph(i, graph_grid(count, s->items_per_wave));
            if (pg) {
                TSVD_HIP(hipGraphLaunch(pg->exec, st));
                s->graph_launches++;
                i += GRAPH_ROUNDS;
            } else {
                // capture failed: this handle launches its rounds one by one
                // from here on (the ring position need not be a block start)
                graphs = false;
                s->use_graphs = 0;
                s->graph_failures++;
            }
        }
        if (!pg) {
            const uint32_t grid = round_grid(count, s->items_per_wave);
            const int per_sync = pull_enabled ? 1 : ROUNDS_PER_SYNC;
            for (int kk = 0; kk < per_sync; kk++, i++) {
                // pull rounds implement the reference's positive-side threshold only
                const bool pull = pull_enabled && (double)count > s->pull_frac * (double)s->n;
                if (pull) {
                    uint32_t *next = ring[(i + 1) % 3];
                    uint32_t *hub_ahead = hubset[(i + 1) % 2];
                    // per-vertex pushed amounts and group masks of pull rounds, made on first use
                    if (s->pm.n == 0) {
                        s->delta.alloc((size_t)s->n * s->s_pad * s->esize);
                        s->pm.alloc((size_t)s->n);
                        TSVD_HIP(hipMemsetAsync(s->pm.p, 0, s->pm.bytes(), st));
                    }
                    build_worklist(s, d);
                    s->touched_all = true;   // the pull pop does not mark items
                    TSVD_HIP(hipMemsetAsync(next, 0, sizeof(uint32_t), st));
                    TSVD_HIP(hipMemsetAsync(hub_ahead, 0, 2 * sizeof(uint32_t), st));
                    const GraphDir &Q = pull_dir(s, d);
                    {
                        PROF("ppr_pull_pop", st);
                        hipLaunchKernelGGL(k_pull_pop<T>, dim3((unsigned)ceil_div(count, WPB)), dim3(TPB), 0, st, P,
                                           R, (T *)s->delta.p, s->pm.p, s->F[d].p, count, s->cm[d].p, D.deg.p,
                                           s->s_pad, s->ngroups, s->alpha, (T)s->rmax, s->directed, s->stats.p);
                    }
                    {
                        PROF("ppr_pull", st);
                        const size_t shm = (size_t)WPB * s->ngroups * 64 * sizeof(T);
                        hipLaunchKernelGGL(k_pull<T>, dim3((unsigned)ceil_div(s->wl_n[d], WPB)), dim3(TPB), shm, st,
                                           R, (const T *)s->delta.p, s->pm.p, s->wl_u[d].p, s->wl_c[d].p,
                                           (uint32_t)s->wl_n[d], Q.rowoff.p, Q.adj.p, Q.deg.p, D.deg.p, s->s_pad,
                                           s->ngroups, (T)s->rmax, s->nm.p, s->fnext.p, next, cap, cnt + 3);
                    }
                    hipLaunchKernelGGL(k_pull_clear, dim3((unsigned)ceil_div(count, TPB)), dim3(TPB), 0, st,
                                       s->pm.p, s->F[d].p, count, s->ngroups);
                    s->last_pull_rounds[d]++;
                } else {
                    {
                        PROF("ppr_push_round", st);
                        launch_round(i, grid);
                    }
                    {
                        PROF("ppr_push_hub", st);
                        launch_hub(i);
                    }
                }
                TSVD_HIP(hipGetLastError());
                std::swap(s->cm[d], s->nm);
                std::swap(s->F[d], s->fnext);
            }
        }
        // the count of round i (the next one to run) and the overflow flag
        TSVD_HIP(hipMemcpyAsync(h + 2, ring[i % 3], sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        TSVD_HIP(hipMemcpyAsync(h + 3, cnt + 3, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        TSVD_HIP(hipStreamSynchronize(st));
        if (pg && !pg->ev.empty()) {
            // kernel times of the replay, from the graph's event nodes; where
            // they cannot be read the handle's later rounds run one by one,
            // timed by events of their own
            float ms[2 * GRAPH_ROUNDS];
            bool timed = true;
            for (int k = 0; k < 2 * GRAPH_ROUNDS && timed; k++)
                timed = hipEventElapsedTime(&ms[k], pg->ev[k], pg->ev[k + 1]) == hipSuccess;
            if (timed) {
                for (int k = 0; k < 2 * GRAPH_ROUNDS; k++)
                    prof_add_time(k % 2 == 0 ? "ppr_push_round" : "ppr_push_hub", ms[k]);
            } else {
                fprintf(stderr, "treesvd: push graph event times unreadable (%s); rounds are launched one by one\n",
                        hipGetErrorString(hipGetLastError()));
                graphs = false;
                s->use_graphs = 0;
                s->graph_failures++;
            }
        }
        TSVD_REQUIRE(h[3] == 0, "frontier overflow in push round %llu", (unsigned long long)i);
        count = h[2];
    }
    tm.stop(st);
    TSVD_HIP(hipMemsetAsync(cnt + d, 0, sizeof(uint32_t), st));
    TSVD_HIP(hipMemcpyAsync(h + 10, cnt + 10, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    TSVD_HIP(hipMemcpyAsync(s->h_stats, s->stats.p, s->stats.bytes(), hipMemcpyDeviceToHost, st));
    TSVD_HIP(hipStreamSynchronize(st));
    unsigned long long hs[STAT_FIELDS] = {};
    for (int sh = 0; sh < STAT_SHARDS; sh++)
        for (int f = 0; f < STAT_FIELDS; f++) hs[f] += s->h_stats[sh * STAT_FIELDS + f];
    if (g_profile) {
        // algorithmic bytes per relaxation (column id + residue RMW) and per pop
        // (row offsets, residue read/clear, reserve RMW); relaxations of split
        // rows beyond their first chunk belong to k_push_hub (to the pull
        // kernel in pull rounds)
        const double per_relax = sizeof(T) == 8 ? 20.0 : 12.0, per_pop = sizeof(T) == 8 ? 48.0 : 32.0;
        const double own = (double)hs[4], rest = (double)(hs[1] - hs[4]);
        prof_add_work("ppr_push_round", per_relax * own + per_pop * (double)hs[0], 0.0);
        if (rest > 0) prof_add_work(s->last_pull_rounds[d] ? "ppr_pull" : "ppr_push_hub", per_relax * rest, 0.0);
    }
    s->last_rounds[d] = (int64_t)h[10] + s->last_pull_rounds[d];
    s->last_pops[d] = (int64_t)hs[0];
    s->last_relax[d] = (int64_t)hs[1];
    s->last_items[d] = (int64_t)hs[2];
    s->last_wave_relax[d] = (int64_t)hs[3];
    s->last_line_relax[d] = (int64_t)hs[5];
    s->last_ms[d] = tm.ms();
}

}  // namespace

void ppr_refresh(tsvd_ppr *s, int mode, int dir, int32_t s0, int32_t s1)
{
    TSVD_REQUIRE(mode == TSVD_REFRESH_INIT || mode == TSVD_REFRESH_DYNAMIC, "bad refresh mode");
    TSVD_REQUIRE(dir >= 0 && dir < s->ndir, "bad direction");
    TSVD_REQUIRE(s0 >= 0 && s0 <= s1 && s1 <= s->nsrc, "bad source range");
    if (s0 == s1) return;
    if (s->precision == TSVD_FP64) refresh_t<double>(s, mode, dir, s0, s1);
    else refresh_t<float>(s, mode, dir, s0, s1);
}

template <class T>
static void update_edges_t(tsvd_ppr *s, int d, const int32_t *src, const int32_t *dst, int64_t m, int32_t s0,
                           int32_t s1)
{
    tsvd_graph *g = s->g;
    hipStream_t st = g->stream;
    const GraphDir &D = g->dir[d];
    T *P = (T *)s->p(d), *R = (T *)s->r(d);
    const int32_t *efrom = d == 0 ? src : dst, *eto = d == 0 ? dst : src;
    const int32_t gcount = (int32_t)ceil_div(s1, 64);
    uint32_t *cnt = s->counters.p;
    const uint32_t cap = (uint32_t)s->F[d].n;
    if (s->first[d].n == 0) {
        s->first[d].alloc((size_t)s->n);
        TSVD_HIP(hipMemsetAsync(s->first[d].p, 0x7F, s->first[d].bytes(), st));   // NO_EDGE in every byte
    }
    int32_t *first = s->first[d].p;
    if (g_profile)   // per (edge, source): p[from] read, r[to] RMW; per (grown row, source): p and r RMW
        prof_add_work("ppr_update", (double)sizeof(T) * (s1 - s0) * (3.0 * m + 4.0 * D.ntouched), 0.0);
    PROF("ppr_update", st);
    if (m > 0)
        hipLaunchKernelGGL(k_first_edge, dim3((unsigned)ceil_div(m, TPB)), dim3(TPB), 0, st, efrom, m, D.former.p,
                           first, 0);
    const int64_t rw = ceil_div(D.ntouched, EB) * gcount;   // waves of EB rows, EB edges, EB edge ends
    const int64_t ew = ceil_div(m, EB) * gcount, xw = ceil_div(2 * m, EB) * gcount;
    if (rw > 0)
        hipLaunchKernelGGL(k_edge_rows<T>, dim3((unsigned)ceil_div(rw, WPB)), dim3(TPB), 0, st, P, R, D.touched.p,
                           D.ntouched, D.deg.p, D.former.p, s->s_pad, gcount, s0, s1, s->alpha, 0);
    if (m > 0)
        hipLaunchKernelGGL(k_edge_adds<T>, dim3((unsigned)ceil_div(ew, WPB)), dim3(TPB), 0, st, P, R, efrom, eto, m,
                           D.former.p, first, s->s_pad, gcount, s0, s1, s->alpha);
    if (rw > 0)
        hipLaunchKernelGGL(k_edge_rows<T>, dim3((unsigned)ceil_div(rw, WPB)), dim3(TPB), 0, st, P, R, D.touched.p,
                           D.ntouched, D.deg.p, D.former.p, s->s_pad, gcount, s0, s1, s->alpha, 1);
    if (m > 0) {
        hipLaunchKernelGGL(k_edge_frontier<T>, dim3((unsigned)ceil_div(xw, WPB)), dim3(TPB), 0, st, R,
                           efrom, eto, m, D.deg.p, s->s_pad, s->ngroups, gcount, s0, s1, (T)s->rmax, s->cm[d].p,
                           s->F[d].p, cnt + d, cap, cnt + 3);
        hipLaunchKernelGGL(k_first_edge, dim3((unsigned)ceil_div(m, TPB)), dim3(TPB), 0, st, efrom, m, D.former.p,
                           first, 1);
    }
    TSVD_HIP(hipGetLastError());
    TSVD_HIP(hipStreamSynchronize(st));
}

void ppr_update_edges(tsvd_ppr *s, int dir, const int32_t *src, const int32_t *dst, int64_t m, int32_t s0,
                      int32_t s1)
{
    TSVD_REQUIRE(dir >= 0 && dir < s->ndir, "bad direction");
    TSVD_REQUIRE(s0 >= 0 && s0 <= s1 && s1 <= s->nsrc, "bad source range");
    TSVD_REQUIRE(m >= 0 && m < (int64_t)NO_EDGE, "edge batch too large for the edge-order update");
    TSVD_REQUIRE(!s->two_sided, "the edge-order update follows the positive-side threshold");
    if (s0 == s1) return;
    static_assert(NO_EDGE == 0x7F7F7F7F, "the marker is set with a byte fill");
    s->touched_all = true;
    if (s->precision == TSVD_FP64) update_edges_t<double>(s, dir, src, dst, m, s0, s1);
    else update_edges_t<float>(s, dir, src, dst, m, s0, s1);
}

void ppr_save(tsvd_ppr *s, PushSave &v)
{
    hipStream_t st = s->g->stream;
    TSVD_HIP(hipMemcpyAsync(s->h_counters, s->counters.p, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    TSVD_HIP(hipStreamSynchronize(st));
    for (int d = 0; d < s->ndir; d++) {
        v.count[d] = s->h_counters[d];
        v.P[d].alloc(s->P[d].bytes());
        v.R[d].alloc(s->R[d].bytes());
        v.cm[d].alloc(s->cm[d].n);
        v.F[d].alloc(std::max<uint32_t>(v.count[d], 1));
        TSVD_HIP(hipMemcpyAsync(v.P[d].p, s->P[d].p, s->P[d].bytes(), hipMemcpyDeviceToDevice, st));
        TSVD_HIP(hipMemcpyAsync(v.R[d].p, s->R[d].p, s->R[d].bytes(), hipMemcpyDeviceToDevice, st));
        TSVD_HIP(hipMemcpyAsync(v.cm[d].p, s->cm[d].p, s->cm[d].bytes(), hipMemcpyDeviceToDevice, st));
        if (v.count[d])
            TSVD_HIP(hipMemcpyAsync(v.F[d].p, s->F[d].p, sizeof(uint32_t) * v.count[d], hipMemcpyDeviceToDevice, st));
    }
    TSVD_HIP(hipStreamSynchronize(st));
}

void ppr_restore(tsvd_ppr *s, const PushSave &v)
{
    hipStream_t st = s->g->stream;
    s->touched_all = true;
    uint32_t cnt[2] = {v.count[0], v.count[1]};
    for (int d = 0; d < s->ndir; d++) {
        TSVD_HIP(hipMemcpyAsync(s->P[d].p, v.P[d].p, s->P[d].bytes(), hipMemcpyDeviceToDevice, st));
        TSVD_HIP(hipMemcpyAsync(s->R[d].p, v.R[d].p, s->R[d].bytes(), hipMemcpyDeviceToDevice, st));
        TSVD_HIP(hipMemcpyAsync(s->cm[d].p, v.cm[d].p, s->cm[d].bytes(), hipMemcpyDeviceToDevice, st));
        if (v.count[d])
            TSVD_HIP(hipMemcpyAsync(s->F[d].p, v.F[d].p, sizeof(uint32_t) * v.count[d], hipMemcpyDeviceToDevice, st));
        TSVD_HIP(hipMemcpyAsync(s->counters.p + d, &cnt[d], sizeof(uint32_t), hipMemcpyHostToDevice, st));
    }
    TSVD_HIP(hipStreamSynchronize(st));
}

size_t ppr_save_bytes(const tsvd_ppr *s)
{
    size_t b = 0;
    for (int d = 0; d < s->ndir; d++) b += s->P[d].bytes() + s->R[d].bytes() + s->cm[d].bytes() + s->F[d].bytes();
    return b;
}

void ppr_bind_sources(tsvd_ppr *s, const int32_t *sources, int32_t count)
{
    TSVD_REQUIRE(count > 0 && count <= s->s_pad, "source count exceeds the state's width");
    s->sources.ensure((size_t)s->s_pad);
    TSVD_HIP(hipMemcpyAsync(s->sources.p, sources, sizeof(int32_t) * count, hipMemcpyHostToDevice, s->g->stream));
    TSVD_HIP(hipStreamSynchronize(s->g->stream));
    s->h_sources.assign(sources, sources + count);
    s->nsrc = count;
    s->touched_all = true;
}

void ppr_read_lane(tsvd_ppr *s, int dir, int what, int32_t lane, double *out)
{
    TSVD_REQUIRE(dir >= 0 && dir < s->ndir && lane >= 0 && lane < s->s_pad, "bad direction or lane");
    hipStream_t st = s->g->stream;
    DBuf<double> col((size_t)s->n);
    const void *X = what == 0 ? s->p(dir) : s->r(dir);
    if (s->precision == TSVD_FP64)
        hipLaunchKernelGGL(k_lane_column<double>, dim3((unsigned)ceil_div(s->n, TPB)), dim3(TPB), 0, st,
                           (const double *)X, s->n, s->s_pad, lane, col.p);
    else
        hipLaunchKernelGGL(k_lane_column<float>, dim3((unsigned)ceil_div(s->n, TPB)), dim3(TPB), 0, st,
                           (const float *)X, s->n, s->s_pad, lane, col.p);
    TSVD_HIP(hipGetLastError());
    TSVD_HIP(hipMemcpyAsync(out, col.p, sizeof(double) * s->n, hipMemcpyDeviceToHost, st));
    TSVD_HIP(hipStreamSynchronize(st));
}

void ppr_clamp_rows(tsvd_ppr *s, int dir)
{
    hipStream_t st = s->g->stream;
    const GraphDir &D = s->g->dir[dir];
    const int64_t cells = (int64_t)D.ntouched * s->s_pad;
    if (cells == 0) return;
    if (s->precision == TSVD_FP64)
        hipLaunchKernelGGL(k_clamp_rows<double>, dim3((unsigned)ceil_div(cells, TPB)), dim3(TPB), 0, st,
                           (double *)s->R[dir].p, D.touched.p, D.ntouched, s->s_pad);
    else
        hipLaunchKernelGGL(k_clamp_rows<float>, dim3((unsigned)ceil_div(cells, TPB)), dim3(TPB), 0, st,
                           (float *)s->R[dir].p, D.touched.p, D.ntouched, s->s_pad);
    TSVD_HIP(hipGetLastError());
}

void ppr_clear(tsvd_ppr *s, int dir)
{
    hipStream_t st = s->g->stream;
    s->touched_all = true;
    TSVD_HIP(hipMemsetAsync(s->P[dir].p, 0, s->P[dir].bytes(), st));
    TSVD_HIP(hipMemsetAsync(s->R[dir].p, 0, s->R[dir].bytes(), st));
}

void ppr_push(tsvd_ppr *s, int dir, int32_t s0, int32_t s1)
{
    TSVD_REQUIRE(dir >= 0 && dir < s->ndir, "bad direction");
    TSVD_REQUIRE(s0 >= 0 && s0 <= s1 && s1 <= s->nsrc, "bad source range");
    if (s->precision == TSVD_FP64) push_t<double>(s, dir, s0, s1);
    else push_t<float>(s, dir, s0, s1);
}

}  // namespace tsvd

using namespace tsvd;

extern "C" int tsvd_ppr_create(tsvd_graph *g, const int32_t *sources, int32_t nsrc, double alpha, double rmax,
                               int precision, tsvd_ppr **out)
{
    return guard([&] {
        TSVD_REQUIRE(g && out && sources, "null argument");
        TSVD_REQUIRE(nsrc > 0, "empty source set");
        TSVD_REQUIRE(alpha > 0.0 && alpha < 1.0, "alpha must be in (0,1)");
        TSVD_REQUIRE(rmax > 0.0, "rmax must be positive");
        TSVD_REQUIRE(precision == TSVD_FP32 || precision == TSVD_FP64, "bad precision");
        for (int32_t i = 0; i < nsrc; i++)
            TSVD_REQUIRE(sources[i] >= 0 && sources[i] < g->n, "source %d out of range", sources[i]);
        {
            std::vector<int32_t> sorted(sources, sources + nsrc);
            std::sort(sorted.begin(), sorted.end());
            for (int32_t i = 1; i < nsrc; i++) TSVD_REQUIRE(sorted[i] != sorted[i - 1], "duplicate source %d", sorted[i]);
        }
        tsvd_ppr *s = new tsvd_ppr();
        try {
            s->g = g;
            s->n = g->n;
            s->nsrc = nsrc;
            s->s_pad = (int32_t)round_up(nsrc, 64);
            s->ngroups = s->s_pad / 64;
            TSVD_REQUIRE((uint64_t)g->n * (uint64_t)s->ngroups < 0xFFFFFFFFull, "n * groups exceeds 32-bit items");
            s->ndir = g->ndir;
            s->directed = g->directed;
            s->alpha = alpha;
            s->rmax = rmax;
            s->precision = precision;
            s->esize = precision == TSVD_FP64 ? 8 : 4;
            // rounds with more than pull_frac * n items run as pull rounds (0: never)
            if (const char *e = getenv("TSVD_PULL_FRAC")) s->pull_frac = atof(e);
            // lanes above ride_frac * rmax ride along with a pushing item (>= 1: off)
            if (const char *e = getenv("TSVD_RIDE_FRAC")) s->ride_frac = atof(e);
            // rows longer than this many neighbours are split across waves
            s->hub_chunk = HUB_CHUNK;
            if (const char *e = getenv("TSVD_HUB_CHUNK")) {
                s->hub_chunk = std::max(64, atoi(e));
                s->hub_auto = false;
            }
            // 1: push rounds replayed from captured graphs instead of launched one by one
            if (const char *e = getenv("TSVD_PUSH_GRAPHS")) s->use_graphs = atoi(e) != 0;
            // items a wave takes per round (the round grid shrinks by this factor)
            if (const char *e = getenv("TSVD_ITEMS_PER_WAVE")) s->items_per_wave = std::max(1, atoi(e));
            s->h_sources.assign(sources, sources + nsrc);
            s->sources.alloc((size_t)nsrc);
            hipStream_t st = g->stream;
            TSVD_HIP(hipMemcpyAsync(s->sources.p, sources, sizeof(int32_t) * nsrc, hipMemcpyHostToDevice, st));
            const size_t cells = (size_t)g->n * (size_t)s->s_pad;
            const size_t items = (size_t)g->n * (size_t)s->ngroups;
            for (int d = 0; d < s->ndir; d++) {
                s->P[d].alloc(cells * s->esize);
                s->R[d].alloc(cells * s->esize);
                s->cm[d].alloc(items);
                s->F[d].alloc(items);
                TSVD_HIP(hipMemsetAsync(s->P[d].p, 0, s->P[d].bytes(), st));
                TSVD_HIP(hipMemsetAsync(s->R[d].p, 0, s->R[d].bytes(), st));
                TSVD_HIP(hipMemsetAsync(s->cm[d].p, 0, s->cm[d].bytes(), st));
            }
            s->nm.alloc(items);
            s->fnext.alloc(items);
            TSVD_HIP(hipMemsetAsync(s->nm.p, 0, s->nm.bytes(), st));
            s->touched.alloc(items);
            TSVD_HIP(hipMemsetAsync(s->touched.p, 0, s->touched.bytes(), st));
            s->csr_ref.alloc(2);
            TSVD_HIP(hipHostMalloc((void **)&s->h_csr_ref, 2 * sizeof(PushCsr), hipHostMallocDefault));
            s->counters.alloc(16);
            s->stats.alloc((size_t)STAT_SHARDS * STAT_FIELDS);
            TSVD_HIP(hipHostMalloc((void **)&s->h_stats, s->stats.bytes(), hipHostMallocDefault));
            TSVD_HIP(hipMemsetAsync(s->counters.p, 0, s->counters.bytes(), st));
            TSVD_HIP(hipHostMalloc((void **)&s->h_counters, 128, hipHostMallocDefault));
            TSVD_HIP(hipStreamSynchronize(st));
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

extern "C" int tsvd_ppr_destroy(tsvd_ppr *s)
{
    return guard([&] { delete s; });
}

extern "C" int tsvd_ppr_refresh(tsvd_ppr *s, int mode, int dir, int32_t src_begin, int32_t src_end)
{
    return guard([&] {
        TSVD_REQUIRE(s, "null ppr");
        ppr_refresh(s, mode, dir, src_begin, src_end);
    });
}

extern "C" int tsvd_ppr_refresh_edges(tsvd_ppr *s, int dir, const int32_t *src, const int32_t *dst, int64_t m,
                                      int on_device, int32_t src_begin, int32_t src_end)
{
    return guard([&] {
        TSVD_REQUIRE(s && (m == 0 || (src && dst)) && m >= 0, "bad arguments");
        const int32_t *ds = src, *dd = dst;
        DBuf<int32_t> bs, bd;
        if (m > 0 && !on_device) {
            for (int64_t e = 0; e < m; e++)
                TSVD_REQUIRE(src[e] >= 0 && src[e] < s->n && dst[e] >= 0 && dst[e] < s->n,
                             "edge %lld holds a vertex id outside [0, %d)", (long long)e, s->n);
            bs.alloc((size_t)m);
            bd.alloc((size_t)m);
            TSVD_HIP(hipMemcpyAsync(bs.p, src, sizeof(int32_t) * m, hipMemcpyHostToDevice, s->g->stream));
            TSVD_HIP(hipMemcpyAsync(bd.p, dst, sizeof(int32_t) * m, hipMemcpyHostToDevice, s->g->stream));
            ds = bs.p;
            dd = bd.p;
        }
        ppr_update_edges(s, dir, ds, dd, m, src_begin, src_end);
    });
}

extern "C" int tsvd_ppr_get_push_setting(const tsvd_ppr *s, tsvd_push_setting *out)
{
    return guard([&] {
        TSVD_REQUIRE(s && out, "null argument");
        out->hub_chunk = s->hub_chunk;
        out->ride_frac = s->ride_frac;
        out->items_per_wave = s->items_per_wave;
    });
}

extern "C" int tsvd_ppr_set_push_setting(tsvd_ppr *s, const tsvd_push_setting *in)
{
    return guard([&] {
        TSVD_REQUIRE(s && in, "null argument");
        TSVD_REQUIRE(in->hub_chunk >= 64 && in->ride_frac >= 0.0 && in->items_per_wave >= 1 &&
                         in->items_per_wave <= 1024, "bad push setting");
        s->hub_chunk = in->hub_chunk;
        s->hub_auto = false;
        s->ride_frac = in->ride_frac;
        s->items_per_wave = in->items_per_wave;
    });
}

extern "C" int tsvd_ppr_push(tsvd_ppr *s, int dir, int32_t src_begin, int32_t src_end)
{
    return guard([&] {
        TSVD_REQUIRE(s, "null ppr");
        ppr_push(s, dir, src_begin, src_end);
    });
}

extern "C" int tsvd_ppr_read(tsvd_ppr *s, int dir, int what, double *out)
{
    return guard([&] {
        TSVD_REQUIRE(s && out, "null argument");
        TSVD_REQUIRE(dir >= 0 && dir < s->ndir && (what == 0 || what == 1), "bad direction or field");
        const size_t cells = (size_t)s->n * (size_t)s->s_pad;
        std::vector<uint8_t> h(cells * s->esize);
        const void *src = what == 0 ? s->p(dir) : s->r(dir);
        TSVD_HIP(hipMemcpyAsync(h.data(), src, h.size(), hipMemcpyDeviceToHost, s->g->stream));
        TSVD_HIP(hipStreamSynchronize(s->g->stream));
        for (int32_t k = 0; k < s->nsrc; k++) {
            const size_t row = s->out_row.empty() ? (size_t)k : (size_t)s->out_row[(size_t)k];
            for (int32_t v = 0; v < s->n; v++) {
                const size_t c = (size_t)v * s->s_pad + k;
                out[row * s->n + v] =
                    s->esize == 8 ? ((const double *)h.data())[c] : (double)((const float *)h.data())[c];
            }
        }
    });
}

extern "C" int tsvd_ppr_last_push_stats(tsvd_ppr *s, int dir, int64_t *rounds, int64_t *pops, int64_t *relax,
                                        double *kernel_ms)
{
    return guard([&] {
        TSVD_REQUIRE(s && dir >= 0 && dir < s->ndir, "bad ppr or direction");
        if (rounds) *rounds = s->last_rounds[dir];
        if (pops) *pops = s->last_pops[dir];
        if (relax) *relax = s->last_relax[dir];
        if (kernel_ms) *kernel_ms = s->last_ms[dir];
    });
}

extern "C" int tsvd_ppr_set_two_sided(tsvd_ppr *s, int on)
{
    return guard([&] {
        TSVD_REQUIRE(s, "null ppr");
        s->two_sided = on ? 1 : 0;
    });
}

extern "C" int tsvd_ppr_set_graphs(tsvd_ppr *s, int on)
{
    return guard([&] {
        TSVD_REQUIRE(s, "null ppr");
        s->use_graphs = on ? 1 : 0;
        if (!on) s->graphs.clear();
    });
}

extern "C" int tsvd_ppr_graph_stats(const tsvd_ppr *s, int64_t *launches, int64_t *captures, int64_t *failures)
{
    return guard([&] {
        TSVD_REQUIRE(s, "null ppr");
        if (launches) *launches = s->graph_launches;
        if (captures) *captures = s->graph_captures;
        if (failures) *failures = s->graph_failures;
    });
}

namespace tsvd {
namespace {

__device__ inline uint32_t probe_hash(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Each wavefront adds 1 to `iters` rows at hashed indices with one returning
// 64-lane atomic per row (the relaxation's access, without the graph).
// One 64-lane FP32 atomic add per iteration at a hashed 256-byte row; RET:
// the returning form (as the push uses, to see threshold crossings) or the
// non-returning one; only lanes below `lanes` take part (32 lanes = one
// 128-byte line).
template <bool RET>
__global__ void __launch_bounds__(TPB) k_probe_atomic(float *__restrict__ R, uint32_t rows, int32_t iters,
                                                      int32_t lanes, float *__restrict__ sink)
{
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    float acc = 0.f;
    if (lane < lanes) {
        for (int32_t i = 0; i < iters; i++) {
            const uint32_t row = probe_hash(wave * 7919u + (uint32_t)i) % rows;
            if (RET) acc += atomicAdd(R + (size_t)row * 64 + lane, 1.0f);
            else atomicAdd(R + (size_t)row * 64 + lane, 1.0f);
        }
    }
    if (RET && acc < 0.f) sink[0] = acc;   // keeps the returned values live
}

}  // namespace
}  // namespace tsvd

static void probe_atomics(int64_t rows, int32_t iters, int returning, int32_t lanes, double *wave_atomics_per_s)
{
    TSVD_REQUIRE(wave_atomics_per_s && rows >= 1 && rows <= (int64_t)1 << 26 && iters >= 1 && iters <= 4096,
                 "rows in [1, 2^26], iters in [1, 4096]");
    TSVD_REQUIRE(lanes >= 1 && lanes <= 64, "lanes in [1, 64]");
    DBuf<float> R((size_t)rows * 64), sink(1);
    TSVD_HIP(hipMemset(R.p, 0, R.bytes()));
    const int blocks = 8192;
    Timer t;
    double best = 0.0;
    for (int rep = 0; rep < 3; rep++) {
        t.start(nullptr);
        if (returning)
            hipLaunchKernelGGL(k_probe_atomic<true>, dim3(blocks), dim3(TPB), 0, nullptr, R.p, (uint32_t)rows,
                               iters, lanes, sink.p);
        else
            hipLaunchKernelGGL(k_probe_atomic<false>, dim3(blocks), dim3(TPB), 0, nullptr, R.p, (uint32_t)rows,
                               iters, lanes, sink.p);
        TSVD_HIP(hipGetLastError());
        t.stop(nullptr);
        const double ms = t.ms();
        const double waves = (double)blocks * TPB / 64;
        best = std::max(best, waves * iters / (ms * 1e-3));
    }
    *wave_atomics_per_s = best;
}

extern "C" int tsvd_probe_atomic_rate(int64_t rows, int32_t iters, double *wave_atomics_per_s)
{
    return tsvd::guard([&] { probe_atomics(rows, iters, 1, 64, wave_atomics_per_s); });
}

extern "C" int tsvd_probe_atomic_lanes(int64_t rows, int32_t iters, int32_t returning, int32_t lanes,
                                       double *wave_atomics_per_s)
{
    return tsvd::guard([&] { probe_atomics(rows, iters, returning, lanes, wave_atomics_per_s); });
}
