// tsvd_graph.h -- device CSR of the dynamic graph (one per direction).
#pragma once
#include "tsvd_internal.h"

namespace tsvd {

struct GraphDir {
    DBuf<int64_t> rowoff;   // n+1 capacity offsets: row v owns adj[rowoff[v], rowoff[v+1])
    DBuf<int32_t> deg;      // current degree
    DBuf<int32_t> former;   // degree before the last append (former_degree)
    DBuf<int32_t> cnt;      // entries added to each row by the last append
    DBuf<int32_t> touched;  // rows grown by the last append, ascending
    DBuf<int32_t> adj;      // neighbour ids
    int64_t cap_total = 0;
    int64_t appended = 0;   // entries added by the last append (new slots)
    int32_t ntouched = 0;
};

}  // namespace tsvd

struct tsvd_graph {
    int32_t n = 0;
    int directed = 0, ndir = 1;
    double slack = 0.1;
    int64_t m = 0;
    int64_t rebuilds = 0;
    int64_t version = 0;      // bumped by every append
    int64_t release_above = (int64_t)1 << 26;   // batches above this many edges free their staging
    hipStream_t stream = nullptr;
    tsvd::GraphDir dir[2];
    tsvd::DBuf<int32_t> e_src, e_dst, k_in, v_in, k_out, v_out, scan_flag, scan_pos, segstart;
    tsvd::DBuf<uint8_t> cub_tmp;
    tsvd::DBuf<int32_t> scalar;
    void *h_scalar = nullptr;
    int32_t *d_scalar() { return scalar.p; }
    ~tsvd_graph()
    {
        if (h_scalar) (void)hipHostFree(h_scalar);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace tsvd {
void graph_append(tsvd_graph *g, const int32_t *src, const int32_t *dst, int64_t m, int on_device);
// frees the device copy of the last host batch when it exceeded release_above
// edges; callers that read the staged ids (e_src/e_dst) call it afterwards
void graph_release_edges(tsvd_graph *g);
}
