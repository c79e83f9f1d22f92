// tsvd_ppr.h -- device state of the subset PPR (reserves p, residues r).
#pragma once
#include "tsvd_internal.h"
#include "tsvd_graph.h"

#include <memory>

// Layout: vertex-major.  p[v * s_pad + s] is source s's reserve at vertex v
// (s_pad = |S| rounded up to 64), so one 64-lane wavefront covers one vertex
// for 64 consecutive sources ("source group" g = s / 64) and every access of
// the update, the push and the block transform is a contiguous 256/512-byte
// segment.  The pending work of a direction is a list of (vertex, group)
// items plus a 64-bit lane mask per item (cm); a push round pops the list
// and builds the next one in (nm, fnext).
namespace tsvd {
// The CSR arrays of one push direction, read by the push kernels from device
// memory (the arrays move when an append outgrows them; the descriptor does not).
struct PushCsr {
    const int64_t *rowoff;
    const int32_t *adj;
    const int32_t *deg;
};

// A block of push rounds captured as a hipGraph: the kernels of a fixed
// number of rounds with every launch argument frozen.  `key` lists those
// arguments; a replay is valid while they are unchanged.
struct PushGraph {
    std::vector<uint64_t> key;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> ev;   // profiled captures: event record nodes around each kernel
    uint64_t used = 0;
    PushGraph() = default;
    PushGraph(const PushGraph &) = delete;
    PushGraph &operator=(const PushGraph &) = delete;
    ~PushGraph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};
}  // namespace tsvd

struct tsvd_ppr {
    tsvd_graph *g = nullptr;
    int32_t n = 0, nsrc = 0, s_pad = 0, ngroups = 0, ndir = 1, directed = 0;
    double alpha = 0.2, rmax = 1e-7;
    int precision = TSVD_FP32;
    size_t esize = 4;
    tsvd::DBuf<int32_t> sources;
    tsvd::DBuf<uint8_t> P[2], R[2];
    tsvd::DBuf<unsigned long long> cm[2], nm;
    tsvd::DBuf<uint32_t> F[2], fnext;
    tsvd::DBuf<uint32_t> counters;           // [0],[1] pending per dir, [3] overflow, [4..5] and
                                             // [6..7] split rows and their further chunks (two
                                             // sets, alternating rounds), [8],[9] the round-count
                                             // ring with [dir], [10] rounds that had work
    tsvd::DBuf<uint32_t> hub_item;           // split rows of a push round: item,
    tsvd::DBuf<uint8_t> hub_delta;           //   its 64 pushed amounts,
    tsvd::DBuf<unsigned long long> hub_chunks;  //   (slot, chunk) work entries
    tsvd::DBuf<unsigned long long> stats;    // [0] pops, [1] relaxations
    uint32_t *h_counters = nullptr;          // pinned
    unsigned long long *h_stats = nullptr;   // pinned copy of the sharded push counters
    int64_t last_rounds[2] = {0, 0}, last_pops[2] = {0, 0}, last_relax[2] = {0, 0};
    int64_t last_items[2] = {0, 0}, last_wave_relax[2] = {0, 0}, last_line_relax[2] = {0, 0};
    double last_ms[2] = {0, 0};
    int64_t last_pull_rounds[2] = {0, 0};
    // pull rounds: per-vertex pushed amounts, popped-group masks, and the
    // chunked work list over the pull adjacency (rebuilt when the graph changes)
    tsvd::DBuf<uint8_t> delta;               // n * s_pad * esize
    tsvd::DBuf<unsigned long long> pm;       // n
    tsvd::DBuf<uint32_t> wl_u[2], wl_c[2];   // (vertex, chunk) per push direction
    tsvd::DBuf<int32_t> wl_cnt;              // n + 1 scratch
    int64_t wl_n[2] = {0, 0}, wl_version[2] = {-1, -1};
    int two_sided = 0;                       // 1: push |r| > rmax, 0: r > rmax (the reference)
    int32_t hub_chunk = 256;                 // rows longer than this are split across waves
    bool hub_auto = true;                    // chunk chosen at the first push (no caller setting)
    double pull_frac = 0.0;                  // pull when items > pull_frac * n (0: never)
    double ride_frac = 0.5;                  // sub-threshold lanes above ride_frac*rmax ride along (>= 1: off)
    int32_t items_per_wave = 8;              // items per wave of a round's grid (grid-stride beyond)
    std::vector<int32_t> h_sources;
    std::vector<int32_t> out_row;            // tsvd_ppr_read row of lane s (empty: row s)
    tsvd::DBuf<int32_t> first[2];            // edge-order update: first new edge of a row (made on first use)
    // Items whose reserves changed since the block norms last read them: the
    // insertion update and the push set touched[v * ngroups + g] (any
    // direction); touched_all stands for every item (set when p is rewritten
    // wholesale: static refresh, clear, rebind, restore, pull rounds, the
    // edge-order update).
    tsvd::DBuf<uint8_t> touched;             // n * ngroups
    bool touched_all = true;
    // Push rounds replayed from captured graphs (1) or launched one by one
    // (0, the default: measured faster, since the rounds are already enqueued
    // several at a time); captured graphs kept per launch shape.
    int use_graphs = 0;
    std::vector<std::unique_ptr<tsvd::PushGraph>> graphs;
    int64_t graph_launches = 0, graph_captures = 0, graph_failures = 0;
    uint64_t graph_clock = 0;                // last-use order of the cached graphs
    tsvd::DBuf<tsvd::PushCsr> csr_ref;       // per direction, rewritten at every push
    tsvd::PushCsr *h_csr_ref = nullptr;      // pinned staging of csr_ref
    ~tsvd_ppr()
    {
        if (h_counters) (void)hipHostFree(h_counters);
        if (h_stats) (void)hipHostFree(h_stats);
        if (h_csr_ref) (void)hipHostFree(h_csr_ref);
    }
    void *p(int d) { return P[d].p; }
    void *r(int d) { return R[d].p; }
};

namespace tsvd {
void ppr_refresh(tsvd_ppr *s, int mode, int dir, int32_t s0, int32_t s1);
void ppr_push(tsvd_ppr *s, int dir, int32_t s0, int32_t s1);
// Insertion update of the n >= 1e7-node path in batch order (Q8 included);
// src/dst: the appended batch on the device.
void ppr_update_edges(tsvd_ppr *s, int dir, const int32_t *src, const int32_t *dst, int64_t m, int32_t s0,
                      int32_t s1);
// Rebind the state to `count` sources (at most its width s_pad); the caller
// refills P and R.
void ppr_bind_sources(tsvd_ppr *s, const int32_t *sources, int32_t count);
// Zero P and R of a direction.
void ppr_clear(tsvd_ppr *s, int dir);
// Host FP64 copy of one lane's reserves (what 0) or residues (1), n values.
void ppr_read_lane(tsvd_ppr *s, int dir, int what, int32_t lane, double *out);
// r = max(r, 0) over the rows the last graph append grew (the only rows an
// edge-order update can leave negative), every lane.
void ppr_clamp_rows(tsvd_ppr *s, int dir);
// The push's input (p, r, pending masks and lists of every direction), saved
// and put back, so that one push can be timed under several settings from the
// same start.  A push leaves the next-round masks zero, so these suffice.
struct PushSave {
    tsvd::DBuf<uint8_t> P[2], R[2];
    tsvd::DBuf<unsigned long long> cm[2];
    tsvd::DBuf<uint32_t> F[2];
    uint32_t count[2] = {0, 0};
};
void ppr_save(tsvd_ppr *s, PushSave &v);
void ppr_restore(tsvd_ppr *s, const PushSave &v);
size_t ppr_save_bytes(const tsvd_ppr *s);   // an upper bound of what ppr_save allocates
}
