// tsvd_sparse.h -- state of the n >= 1e7-node path (sparse_state pipelines).
#pragma once
#include "tsvd_internal.h"
#include "tsvd_ppr.h"
#include "tsvd_profile.h"

#include <array>
#include <vector>

namespace tsvd {

// Positive entries of one state array (p or r of one direction) for the
// sources of one batch: off[s] .. off[s+1] index (vertex, value) pairs,
// vertices ascending; values in the state's precision (esize bytes).
struct PairList {
    DBuf<int64_t> off;
    DBuf<int32_t> idx;
    DBuf<uint8_t> val;
    int64_t nnz = 0;
    int32_t count = 0;
    size_t esize = 4;
};

// Compaction scratch shared by all lists.
struct SparseWork {
    DBuf<int64_t> cnt, pos;
    DBuf<uint8_t> tmp;
};

// Sources are processed in batches of `gpb` groups of 64: batch b holds
// sources [b * gpb * 64, min(nsrc, (b+1) * gpb * 64)).  `gp` is the dense
// push state of one batch (width gpb * 64), rebound to each batch in turn.
struct SparseState {
    int32_t gpb = 0, nbatch = 0;
    bool resident = false;                         // one batch kept in HBM, no lists
    int32_t keep = 0;                              // rows >= keep are left out of the blocks (Q11)
    tsvd_ppr *gp = nullptr;
    std::vector<std::array<PairList, 4>> lists;   // per batch: p, r of dir 0, p, r of dir 1
    SparseWork work;
    int32_t batch_sources(int32_t b, int32_t nsrc) const
    {
        return std::min(nsrc - b * gpb * 64, gpb * 64);
    }
};

// Store the positive entries of rows [0, count) of the vertex-major state X
// (n x s_pad, esize bytes per value) into L, replacing its contents.
void list_store(SparseWork &w, PairList &L, const void *X, size_t esize, int32_t n, int32_t s_pad, int32_t count,
                hipStream_t st);
// Scatter L into a zeroed state X of width s_pad.
void list_load(const PairList &L, void *X, int32_t s_pad, hipStream_t st);
void list_clear(PairList &L);
// Dense FP64 copy (n values) of source s's list, on the host.
void list_read(const PairList &L, int32_t s, int32_t n, double *out);

}  // namespace tsvd
