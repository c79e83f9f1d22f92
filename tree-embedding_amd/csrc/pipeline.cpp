// pipeline.cpp -- the per-snapshot driver ("plain GPU method").
//
// Replaces the snapshot loop of Tree_u_NC.cpp:507-866 (and Tree_d_NC.cpp:
// 611-1132, Tree_u_LP.cpp:664-925): ingest -> refresh (static before
// dyn_start, dynamic after) -> push -> block transform with the lazy test
// from lazy_start on (update_mat_record, Tree_u_NC.cpp:608-670) -> frPCA of
// each dirty block with its residual norm (Tree_u_NC.cpp:75-226, 679-710) ->
// merges along the dirty paths -> root embedding; a snapshot with no dirty
// block keeps the previous embedding (Tree_u_NC.cpp:713-721).
//
// The reference computes the leaf residual ||U S V^T - A||_F with a dense
// |S| x w_j product (Tree_u_NC.cpp:158-204).  frPCA's U S V^T is the
// orthogonal projection of A onto span(Q V_k), so the same number is
// sqrt(||A||_F^2 - sum_k sigma^2); both terms are already at hand.
#include "tsvd_pipeline.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>

namespace tsvd {

static double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void validate(const tsvd_config &c)
{
    TSVD_REQUIRE(c.n > 0 && c.nsrc > 0 && c.sources, "n, nsrc and sources are required");
    TSVD_REQUIRE(c.alpha > 0 && c.alpha < 1 && c.rmax > 0, "alpha in (0,1) and rmax > 0 required");
    TSVD_REQUIRE(c.precision == TSVD_FP32 || c.precision == TSVD_FP64, "bad precision");
    TSVD_REQUIRE(c.nparts >= 1 && c.nparts <= c.n, "nparts must be in [1, n]");
    TSVD_REQUIRE(c.fanout >= 2, "fan-out must be >= 2");
    TSVD_REQUIRE(c.rank >= 1 && c.rank + 5 <= 512, "rank must be in [1, 507]");
    TSVD_REQUIRE(c.q >= 1, "q must be >= 1");
    TSVD_REQUIRE(c.transform == TSVD_TRANSFORM_NC || c.transform == TSVD_TRANSFORM_LP, "bad transform");
    TSVD_REQUIRE(c.fanout * c.rank <= 4096, "fan-out * rank too large for the merge eigensolve");
    TSVD_REQUIRE(!c.right_embedding || c.transform == TSVD_TRANSFORM_LP,
                 "the right embedding is defined for the LP transform (Tree_u_LP.cpp:291-404)");
    TSVD_REQUIRE(!c.sparse_state || (c.directed && c.transform == TSVD_TRANSFORM_LP),
                 "sparse_state is the directed LP driver's path (Tree_d_LP.cpp:1377-2024)");
    TSVD_REQUIRE(!c.sparse_state || !c.two_sided, "sparse_state keeps the reference's positive-side threshold");
    TSVD_REQUIRE(c.sparse_groups >= 0, "sparse_groups must be >= 0");
    TSVD_REQUIRE(c.dyn_start >= 1 && c.lazy_start >= 1, "dyn_start and lazy_start are 1-based snapshot numbers");
}

// Batches of the sparse state.  When the state of every source fits in 60%
// of the free HBM it stays resident as one batch: no pair lists are written,
// and the negative residues the lists would drop are cleared at the start of
// the next update instead (the same state, since nothing reads them in
// between).  Otherwise, as many groups of 64 sources as 30% of the free HBM
// holds (four FP32 arrays of n x 64 per group plus the push's item lists) are
// pushed at a time, and the rest of the memory is left for the pair lists,
// which on power-law graphs can outgrow the batch state (a hub's residue
// reaches millions of vertices).  sparse_groups > 0 always uses lists.
static void sparse_setup(tsvd_pipeline *p, hipStream_t st)
{
    const tsvd_config &c = p->cfg;
    SparseState &ss = p->sp;
    const int32_t ngroups = (int32_t)ceil_div(c.nsrc, 64);
    int32_t gpb = c.sparse_groups;
    if (gpb <= 0) {
        size_t fr = 0, tot = 0;
        TSVD_HIP(hipMemGetInfo(&fr, &tot));
        const double per = (double)c.n * (4.0 * 64 * 4 + 36.0);
        if (per * ngroups <= 0.6 * (double)fr) {
            ss.resident = true;
            gpb = ngroups;
        } else {
            gpb = (int32_t)std::max(1.0, std::floor(0.3 * (double)fr / per));
        }
    }
    ss.gpb = std::min(gpb, ngroups);
    ss.nbatch = (int32_t)ceil_div(ngroups, ss.gpb);
    ss.keep = 30 * (c.nsrc / 30);
    ss.lists.resize((size_t)ss.nbatch);
    const int32_t first = ss.batch_sources(0, c.nsrc);
    int rc = tsvd_ppr_create(p->g, p->sources.data(), first, c.alpha, c.rmax, TSVD_FP32, &ss.gp);
    if (rc) throw Error{rc};
    (void)st;
}

// V_j = A_j^T U for the listed blocks with the current root U (LP: E = U),
// Tree_u_LP.cpp:1040-1045 / 291-404; each listed block now refers to this
// snapshot's U, and U copies no block refers to any more are released.
static void right_embedding_update(tsvd_pipeline *p, const std::vector<int32_t> &blocks, int32_t it,
                                   hipStream_t st)
{
    if (blocks.empty() || !p->tree.have_root) return;
    RightEmbedding &re = p->re;
    const Blocks &b = p->blocks;
    const int32_t k = p->cfg.rank;
    DBuf<double> &U = re.U[it];
    U.alloc((size_t)b.s_pad * k);
    TSVD_HIP(hipMemcpyAsync(U.p, p->tree.E.p, U.bytes(), hipMemcpyDeviceToDevice, st));
    std::vector<GemmProb> gp;
    for (int32_t j : blocks) {
        re.ver[j] = it;
        if (b.ncols[j] == 0) continue;
        GemmProb g{};
        g.A = b.A.p + (size_t)b.start[j] * b.s_pad * b.esize; g.lda = b.s_pad; g.transA = 1;
        g.B = U.p; g.ldb = b.s_pad;
        g.C = re.V.p + (size_t)b.start[j] * k; g.ldc = b.width[j];
        g.M = b.ncols[j]; g.N = k; g.K = b.nsrc; g.alpha = 1.0; g.beta = 0.0;
        gp.push_back(g);
    }
    const DT T = b.precision == TSVD_FP64 ? DT::F64 : DT::F32;
    p->gemm.run(st, gp, GemmTypes{T, DT::F64, DT::F64, DT::F64}, k > 80, "gemm_right");
    const bool degs = p->cfg.sparse_state != 0;
    if (degs) {
        // ppr:7707 / 7727: score *= outdegree[i] * indegree[j] of this snapshot
        std::vector<int32_t> dout((size_t)b.n), din((size_t)b.n);
        TSVD_HIP(hipMemcpyAsync(dout.data(), p->g->dir[0].deg.p, sizeof(int32_t) * b.n, hipMemcpyDeviceToHost, st));
        TSVD_HIP(hipMemcpyAsync(din.data(), p->g->dir[1].deg.p, sizeof(int32_t) * b.n, hipMemcpyDeviceToHost, st));
        TSVD_HIP(hipStreamSynchronize(st));
        std::vector<double> hs((size_t)b.s_pad, 0.0);
        for (int32_t i = 0; i < b.nsrc; i++) hs[(size_t)i] = (double)dout[(size_t)p->sources[(size_t)i]];
        DBuf<double> &ds = re.dsrc[it];
        ds.alloc((size_t)b.s_pad);
        TSVD_HIP(hipMemcpyAsync(ds.p, hs.data(), ds.bytes(), hipMemcpyHostToDevice, st));
        for (int32_t j : blocks) {
            std::vector<double> hv((size_t)b.width[j]);
            for (int64_t c = 0; c < b.width[j]; c++) hv[(size_t)c] = (double)din[(size_t)(b.start[j] + c)];
            TSVD_HIP(hipMemcpyAsync(re.vdeg.p + b.start[j], hv.data(), sizeof(double) * hv.size(),
                                    hipMemcpyHostToDevice, st));
            TSVD_HIP(hipStreamSynchronize(st));
        }
    }
    std::vector<const double *> ptr(b.nparts, nullptr), dptr(b.nparts, nullptr);
    for (int32_t j = 0; j < b.nparts; j++)
        if (re.ver[j] >= 0) {
            ptr[j] = re.U.at(re.ver[j]).p;
            if (degs) dptr[j] = re.dsrc.at(re.ver[j]).p;
        }
    TSVD_HIP(hipMemcpyAsync(re.ublk.p, ptr.data(), sizeof(const double *) * b.nparts, hipMemcpyHostToDevice, st));
    if (degs)
        TSVD_HIP(hipMemcpyAsync(re.dblk.p, dptr.data(), sizeof(const double *) * b.nparts, hipMemcpyHostToDevice, st));
    TSVD_HIP(hipStreamSynchronize(st));
    for (auto m = re.U.begin(); m != re.U.end();) {
        bool live = false;
        for (int32_t j = 0; j < b.nparts && !live; j++) live = re.ver[j] == m->first;
        if (!live) re.dsrc.erase(m->first);
        m = live ? std::next(m) : re.U.erase(m);
    }
}

// Lazy test (Eqn 2, ppr:2699-2726 / 7204-7226): clean iff resid +
// ||A_new - A_old|| < delta sqrt2 ||A_new||, from lazy_start on
// (Tree_u_NC.cpp:608), or from dyn_start on in the n >= 1e7-node path
// (Tree_d_LP.cpp:1658).  `rebuild` gets the dirty blocks, `factor` those of
// them with a nonzero matrix (an all-zero block is skipped, Tree_u_NC.cpp:
// 683-686, Tree_d_LP.cpp:1786-1790).
static void decide_dirty(const tsvd_config &c, const Blocks &b, int32_t it, std::vector<int32_t> &rebuild,
                         std::vector<int32_t> &factor)
{
    const double thr = c.delta * std::sqrt(2.0);
    const int32_t from = c.sparse_state ? c.dyn_start : c.lazy_start;
    for (int32_t j = 0; j < b.nparts; j++) {
        bool dirty = true;
        if (it >= from) dirty = !(b.resid[j] + std::sqrt(b.e2[j]) < thr * std::sqrt(b.a2[j]));
        if (!dirty) continue;
        rebuild.push_back(j);
        if (b.a2[j] > 0.0) factor.push_back(j);
    }
}

// Refresh, push and block test of the n >= 1e7-node path (Tree_d_LP.cpp:
// 1561-1760, All_initialization_and_push_final at ppr:6188-6832), one batch
// of source groups at a time: load its pair lists into the dense state (or
// start it afresh before dyn_start), update in batch order, push both
// directions, store the positive entries back, and add the batch's rows to
// every block's norms.  sparse_rebuild then rebuilds the dirty blocks from
// the stored reserve lists (the matrix rows of a batch need both directions'
// p).
static void sparse_front(tsvd_pipeline *p, int32_t it, const int32_t *dsrc, const int32_t *ddst, int64_t m,
                         tsvd_snapshot_stats &S)
{
    const tsvd_config &c = p->cfg;
    SparseState &ss = p->sp;
    tsvd_ppr *gp = ss.gp;
    Blocks &b = p->blocks;
    hipStream_t st = p->g->stream;
    const bool fresh = it < c.dyn_start;
    const bool keep_r = it >= c.dyn_start - 1;   // ppr:6298-6300: residues are stored from dyn_start-1 on
    blocks_active(b, p->g, st);
    for (int32_t bt = 0; bt < ss.nbatch; bt++) {
        const int32_t row0 = bt * ss.gpb * 64, cnt = ss.batch_sources(bt, c.nsrc);
        auto &L = ss.lists[(size_t)bt];
        if (!ss.resident) ppr_bind_sources(gp, p->sources.data() + row0, cnt);
        for (int d = 0; d < 2; d++) {
            const double t0 = now_ms();
            if (fresh) {
                ppr_refresh(gp, TSVD_REFRESH_INIT, d, 0, cnt);
            } else {
                if (!ss.resident) {
                    ppr_clear(gp, d);
                    list_load(L[(size_t)(2 * d)], gp->p(d), gp->s_pad, st);
                    list_load(L[(size_t)(2 * d + 1)], gp->r(d), gp->s_pad, st);
                }
                ppr_update_edges(gp, d, dsrc, ddst, m, 0, cnt);
            }
            const double t1 = now_ms();
            ppr_push(gp, d, 0, cnt);
            const double t2p = now_ms();
            // what storing the positive entries drops: a resident state keeps
            // its negative residues, which sit on the rows the update lowered
            if (!fresh && ss.resident) ppr_clamp_rows(gp, d);
            S.ms_refresh += (t1 - t0) + (now_ms() - t2p);
            S.ms_push += t2p - t1;
            S.push_rounds += gp->last_rounds[d];
            S.push_pops += gp->last_pops[d];
            S.push_relax += gp->last_relax[d];
            S.push_items += gp->last_items[d];
            S.push_wave_relax += gp->last_wave_relax[d];
            S.push_line_relax += gp->last_line_relax[d];
        }
        const double t2 = now_ms();
        for (int d = 0; d < 2 && !ss.resident; d++) {
            list_store(ss.work, L[(size_t)(2 * d)], gp->p(d), 4, c.n, gp->s_pad, cnt, st);
            if (keep_r) list_store(ss.work, L[(size_t)(2 * d + 1)], gp->r(d), 4, c.n, gp->s_pad, cnt, st);
            else list_clear(L[(size_t)(2 * d + 1)]);
        }
        blocks_norms_batch(b, (const float *)gp->p(0), (const float *)gp->p(1), gp->s_pad, row0, cnt, ss.keep, st);
        S.ms_blocks += now_ms() - t2;
    }
    const double t3 = now_ms();
    blocks_norms_read(b, st);
    S.ms_blocks += now_ms() - t3;
    S.state_batches = ss.resident ? 0 : ss.nbatch;
    for (const auto &L : ss.lists)
        for (const auto &x : L) S.state_entries += x.nnz;
}

static void sparse_rebuild(tsvd_pipeline *p, const std::vector<int32_t> &rebuild)
{
    const tsvd_config &c = p->cfg;
    SparseState &ss = p->sp;
    tsvd_ppr *gp = ss.gp;
    Blocks &b = p->blocks;
    hipStream_t st = p->g->stream;
    if (!rebuild.empty()) {
        blocks_slots(b, rebuild, st);
        for (int32_t bt = 0; bt < ss.nbatch; bt++) {
            const int32_t row0 = bt * ss.gpb * 64, cnt = ss.batch_sources(bt, c.nsrc);
            if (ss.nbatch > 1) {   // with one batch the state still holds it
                for (int d = 0; d < 2; d++) {
                    ppr_clear(gp, d);
                    list_load(ss.lists[(size_t)bt][(size_t)(2 * d)], gp->p(d), gp->s_pad, st);
                }
            }
            blocks_gather_batch(b, rebuild, (const float *)gp->p(0), (const float *)gp->p(1), gp->s_pad, row0, cnt,
                                ss.keep, st);
        }
        blocks_ncols_read(b, st);
    }
}

// Push lanes by descending degree.  Sources of similar degree push at the
// same vertices in the same rounds more often than sources in subset order,
// so a wavefront's 64 lanes carry more work per atomic (on the batched graph
// 15% fewer wave-level atomics, 11-14% less push time).  Rows of the matrix,
// the tree and the embedding stay in the caller's order through
// Blocks::row_of.  Chosen once, before the first push, while the state is
// being built from scratch; TSVD_SOURCE_ORDER=0 keeps the subset order.
static void order_sources(tsvd_pipeline *p)
{
    p->order_set = true;
    if (const char *e = getenv("TSVD_SOURCE_ORDER"))
        if (atoi(e) == 0) return;
    const tsvd_config &c = p->cfg;
    hipStream_t st = p->g->stream;
    std::vector<int64_t> deg((size_t)c.n, 0);
    std::vector<int32_t> h((size_t)c.n);
    for (int d = 0; d < p->g->ndir; d++) {
        TSVD_HIP(hipMemcpyAsync(h.data(), p->g->dir[d].deg.p, sizeof(int32_t) * c.n, hipMemcpyDeviceToHost, st));
        TSVD_HIP(hipStreamSynchronize(st));
        for (int32_t v = 0; v < c.n; v++) deg[(size_t)v] += h[(size_t)v];
    }
    std::vector<int32_t> row((size_t)c.nsrc);
    for (int32_t i = 0; i < c.nsrc; i++) row[(size_t)i] = i;
    std::stable_sort(row.begin(), row.end(), [&](int32_t a, int32_t b) {
        return deg[(size_t)p->sources[(size_t)a]] > deg[(size_t)p->sources[(size_t)b]];
    });
    std::vector<int32_t> lane_src((size_t)c.nsrc), row_of((size_t)p->blocks.s_pad);
    p->lane_of.assign((size_t)c.nsrc, 0);
    for (int32_t l = 0; l < c.nsrc; l++) {
        lane_src[(size_t)l] = p->sources[(size_t)row[(size_t)l]];
        row_of[(size_t)l] = row[(size_t)l];
        p->lane_of[(size_t)row[(size_t)l]] = l;
    }
    for (int32_t l = c.nsrc; l < p->blocks.s_pad; l++) row_of[(size_t)l] = l;
    ppr_bind_sources(p->ppr, lane_src.data(), c.nsrc);
    p->ppr->out_row.assign(row_of.begin(), row_of.begin() + c.nsrc);
    TSVD_HIP(hipMemcpyAsync(p->blocks.row_of.p, row_of.data(), sizeof(int32_t) * row_of.size(),
                            hipMemcpyHostToDevice, st));
    TSVD_HIP(hipStreamSynchronize(st));
}

// Time this snapshot's push under each candidate setting from the same saved
// state, keep the fastest, and put the state back for the real push.
static void tune_push(tsvd_pipeline *p, const PushTune &t)
{
    tsvd_ppr *q = p->ppr;
    for (int32_t i = 0; i < t.ncand; i++) {
        if (t.cand_ms) t.cand_ms[i] = -1.0;
        TSVD_REQUIRE(t.cand[i].hub_chunk >= 64 && t.cand[i].ride_frac >= 0.0 && t.cand[i].items_per_wave >= 1 &&
                         t.cand[i].items_per_wave <= 1024, "bad push setting %d", i);
    }
    size_t fr = 0, tot = 0;
    TSVD_HIP(hipMemGetInfo(&fr, &tot));
    if (ppr_save_bytes(q) + ((size_t)2 << 30) > fr) return;   // no room to save the state: untuned
    PushSave sv;
    ppr_save(q, sv);
    double best = 0.0;
    tsvd_push_setting pick{q->hub_chunk, q->ride_frac, q->items_per_wave};
    for (int32_t i = 0; i < t.ncand; i++) {
        ppr_restore(q, sv);
        q->hub_chunk = t.cand[i].hub_chunk;
        q->ride_frac = t.cand[i].ride_frac;
        q->items_per_wave = t.cand[i].items_per_wave;
        const double t0 = now_ms();
        for (int d = 0; d < q->ndir; d++) ppr_push(q, d, 0, p->cfg.nsrc);
        const double ms = now_ms() - t0;
        if (t.cand_ms) t.cand_ms[i] = ms;
        if (i == 0 || ms < best) {
            best = ms;
            pick = t.cand[i];
        }
    }
    q->hub_chunk = pick.hub_chunk;
    q->hub_auto = false;
    q->ride_frac = pick.ride_frac;
    q->items_per_wave = pick.items_per_wave;
    ppr_restore(q, sv);
}

bool pipeline_front(tsvd_pipeline *p, const int32_t *src, const int32_t *dst, int64_t m, int on_device,
                    StepPlan &plan, const PushTune *tune)
{
    const tsvd_config &c = p->cfg;
    hipStream_t st = p->g->stream;
    plan = StepPlan{};
    tsvd_snapshot_stats &S = plan.S;
    plan.t0 = now_ms();
    graph_append(p->g, src, dst, m, on_device);
    plan.it = ++p->snapshot;
    S.snapshot = plan.it;
    const double t1 = now_ms();
    S.ms_ingest = t1 - plan.t0;
    // Tree_u_NC.cpp:516-518: snapshots before lazy_start-1 are only ingested
    // (lazy_start 100 turns that off; 200 in the n >= 1e7-node path, Tree_d_LP.cpp:1574)
    if (c.lazy_start != (c.sparse_state ? 200 : 100) && plan.it < c.lazy_start - 1) {
        plan.skipped = true;
        S.reused = 1;
        graph_release_edges(p->g);
        return false;
    }
    Blocks &b = p->blocks;
    if (c.sparse_state) {
        const int32_t *dsrc = on_device ? src : p->g->e_src.p;
        const int32_t *ddst = on_device ? dst : p->g->e_dst.p;
        sparse_front(p, plan.it, dsrc, ddst, m, S);
    } else {
        const int mode = plan.it < c.dyn_start ? TSVD_REFRESH_INIT : TSVD_REFRESH_DYNAMIC;
        if (mode == TSVD_REFRESH_INIT && !p->order_set) order_sources(p);
        for (int d = 0; d < p->ppr->ndir; d++) ppr_refresh(p->ppr, mode, d, 0, c.nsrc);
        if (tune && tune->ncand > 0 && mode == TSVD_REFRESH_DYNAMIC) tune_push(p, *tune);
        const double t2 = now_ms();
        S.ms_refresh = t2 - t1;
        for (int d = 0; d < p->ppr->ndir; d++) {
            ppr_push(p->ppr, d, 0, c.nsrc);
            S.push_rounds += p->ppr->last_rounds[d];
            S.push_pops += p->ppr->last_pops[d];
            S.push_relax += p->ppr->last_relax[d];
            S.push_items += p->ppr->last_items[d];
            S.push_wave_relax += p->ppr->last_wave_relax[d];
            S.push_line_relax += p->ppr->last_line_relax[d];
        }
        const double t3 = now_ms();
        S.ms_push = t3 - t2;
        // only the items whose reserves the update and the push changed are re-read
        blocks_norms_items(b, p->ppr->p(0), p->ppr->ndir > 1 ? p->ppr->p(1) : nullptr, p->g, p->ppr->touched.p,
                           p->ppr->touched_all, 0, b.nparts, st);
        p->ppr->touched_all = false;
        S.norm_items = b.last_items;
        S.ms_blocks = now_ms() - t3;
    }
    graph_release_edges(p->g);
    plan.a2 = b.a2;
    return true;
}

void pipeline_decide(tsvd_pipeline *p, StepPlan &plan)
{
    const double t0 = now_ms();
    Blocks &b = p->blocks;
    decide_dirty(p->cfg, b, plan.it, plan.rebuild, plan.factor);
    if (p->cfg.sparse_state)
        sparse_rebuild(p, plan.rebuild);
    else
        blocks_rebuild(b, plan.rebuild, p->ppr->p(0), p->ppr->ndir > 1 ? p->ppr->p(1) : nullptr, p->g->stream);
    p->last_dirty = plan.factor;
    plan.S.ndirty = (int32_t)plan.factor.size();
    plan.S.ms_blocks += now_ms() - t0;
}

void pipeline_back(tsvd_pipeline *p, StepPlan &plan, hipStream_t st)
{
    const tsvd_config &c = p->cfg;
    tsvd_snapshot_stats &S = plan.S;
    Blocks &b = p->blocks;
    const double t4 = now_ms();
    if (!plan.skipped) {
        const std::vector<int32_t> &factor = plan.factor;
        if (!factor.empty()) {
            std::vector<double> sig2;
            frpca_blocks(p->fw, b, factor, p->tree.level[0].p, p->gemm, *p->eig, c.seed, st, sig2);
            if (plan.it >= c.lazy_start - 1)
                for (size_t i = 0; i < factor.size(); i++) {
                    const int32_t j = factor[i];
                    b.resid[j] = std::sqrt(std::max(0.0, plan.a2[j] - sig2[i]));
                }
        }
        const double t5 = now_ms();
        S.ms_frpca = t5 - t4;
        if (factor.empty() && p->tree.have_root) {
            S.reused = 1;
        } else if (!factor.empty()) {
            S.nmerged = tree_update(p->tree, factor, c.transform, p->gemm, *p->eig, st);
        }
        if (c.right_embedding) right_embedding_update(p, plan.rebuild, plan.it, st);
        S.ms_merge = now_ms() - t5;
    }
    S.ms_total = now_ms() - plan.t0;
}

}  // namespace tsvd

using namespace tsvd;

tsvd_pipeline::~tsvd_pipeline()
{
    delete eig;
    if (ppr) tsvd_ppr_destroy(ppr);
    if (sp.gp) tsvd_ppr_destroy(sp.gp);
    if (g) tsvd_graph_destroy(g);
}

extern "C" int tsvd_pipeline_create(const tsvd_config *cfg, tsvd_pipeline **out)
{
    return guard([&] {
        TSVD_REQUIRE(cfg && out, "null argument");
        validate(*cfg);
        tsvd_pipeline *p = new tsvd_pipeline();
        try {
            p->cfg = *cfg;
            p->sources.assign(cfg->sources, cfg->sources + cfg->nsrc);
            p->cfg.sources = p->sources.data();
            int rc = tsvd_graph_create(cfg->n, cfg->directed, cfg->slack > 0 ? cfg->slack : 0.1, &p->g);
            if (rc) throw Error{rc};
            const int32_t s_pad = (int32_t)round_up(cfg->nsrc, 64);
            if (!cfg->sparse_state) {
                rc = tsvd_ppr_create(p->g, cfg->sources, cfg->nsrc, cfg->alpha, cfg->rmax, cfg->precision, &p->ppr);
                if (rc) throw Error{rc};
                p->ppr->two_sided = cfg->two_sided ? 1 : 0;
            } else {
                std::vector<int32_t> sorted(p->sources);
                std::sort(sorted.begin(), sorted.end());
                for (size_t i = 0; i < sorted.size(); i++) {
                    TSVD_REQUIRE(sorted[i] >= 0 && sorted[i] < cfg->n, "source %d out of range", sorted[i]);
                    TSVD_REQUIRE(i == 0 || sorted[i] != sorted[i - 1], "duplicate source %d", sorted[i]);
                }
            }
            hipStream_t st = p->g->stream;
            blocks_setup(p->blocks, cfg->n, cfg->nparts, cfg->nsrc, s_pad, cfg->precision, cfg->transform,
                         cfg->rmax, p->sources, st);
            frpca_setup(p->fw, p->blocks, cfg->rank, cfg->q);
            tree_setup(p->tree, cfg->nparts, cfg->fanout, cfg->rank, cfg->nsrc, s_pad, p->blocks.esize, st);
            p->tree.merge_bf3 = cfg->merge_bf16x3 != 0;
            if (cfg->sparse_state) sparse_setup(p, st);
            p->eig = new Eig();
            for (int32_t i = 0; i < cfg->nsrc; i++) p->re.row_of[cfg->sources[i]] = i;
            if (cfg->right_embedding) {
                p->re.V.alloc((size_t)cfg->n * cfg->rank);
                TSVD_HIP(hipMemsetAsync(p->re.V.p, 0, p->re.V.bytes(), st));
                p->re.ver.assign(cfg->nparts, -1);
                p->re.ublk.alloc((size_t)cfg->nparts);
                TSVD_HIP(hipMemsetAsync(p->re.ublk.p, 0, p->re.ublk.bytes(), st));
                if (cfg->sparse_state) {
                    p->re.dblk.alloc((size_t)cfg->nparts);
                    p->re.vdeg.alloc((size_t)cfg->n);
                    TSVD_HIP(hipMemsetAsync(p->re.dblk.p, 0, p->re.dblk.bytes(), st));
                    TSVD_HIP(hipMemsetAsync(p->re.vdeg.p, 0, p->re.vdeg.bytes(), st));
                }
                TSVD_HIP(hipStreamSynchronize(st));
            }
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}

extern "C" int tsvd_pipeline_destroy(tsvd_pipeline *p)
{
    return guard([&] { delete p; });
}

extern "C" int tsvd_pipeline_snapshot(tsvd_pipeline *p, const int32_t *src, const int32_t *dst, int64_t m,
                                      int on_device, tsvd_snapshot_stats *stats)
{
    return guard([&] {
        TSVD_REQUIRE(p, "null pipeline");
        StepPlan plan;
        if (pipeline_front(p, src, dst, m, on_device, plan)) pipeline_decide(p, plan);
        pipeline_back(p, plan, p->g->stream);
        if (stats) *stats = plan.S;
    });
}

extern "C" int tsvd_pipeline_snapshot_tuned(tsvd_pipeline *p, const int32_t *src, const int32_t *dst, int64_t m,
                                            int on_device, const tsvd_push_setting *cand, int32_t ncand,
                                            tsvd_push_setting *chosen, double *cand_ms, tsvd_snapshot_stats *stats)
{
    return guard([&] {
        TSVD_REQUIRE(p, "null pipeline");
        TSVD_REQUIRE(ncand >= 0 && (ncand == 0 || cand), "bad candidate list");
        PushTune t;
        t.cand = cand;
        t.ncand = p->cfg.sparse_state ? 0 : ncand;
        t.cand_ms = cand_ms;
        for (int32_t i = 0; i < ncand && cand_ms; i++) cand_ms[i] = -1.0;
        StepPlan plan;
        if (pipeline_front(p, src, dst, m, on_device, plan, &t)) pipeline_decide(p, plan);
        pipeline_back(p, plan, p->g->stream);
        if (chosen) {
            tsvd_ppr *q = p->ppr ? p->ppr : p->sp.gp;
            chosen->hub_chunk = q->hub_chunk;
            chosen->ride_frac = q->ride_frac;
            chosen->items_per_wave = q->items_per_wave;
        }
        if (stats) *stats = plan.S;
    });
}

extern "C" int tsvd_pipeline_embedding(tsvd_pipeline *p, double *E)
{
    return guard([&] {
        TSVD_REQUIRE(p && E, "null argument");
        const int32_t k = p->cfg.rank, ns = p->cfg.nsrc, sp = p->tree.s_pad;
        std::vector<double> h((size_t)sp * k);
        TSVD_HIP(hipMemcpy(h.data(), p->tree.E.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
        // device rows are push lanes; subset row i is lane lane_row(i)
        for (int32_t i = 0; i < ns; i++) {
            const size_t lane = (size_t)p->lane_row(i);
            for (int32_t c = 0; c < k; c++) E[(size_t)i * k + c] = h[(size_t)c * sp + lane];
        }
    });
}

extern "C" int tsvd_pipeline_root_sigma(tsvd_pipeline *p, double *sigma)
{
    return guard([&] {
        TSVD_REQUIRE(p && sigma, "null argument");
        for (int32_t c = 0; c < p->cfg.rank; c++) sigma[c] = p->tree.root_sigma[c];
    });
}

extern "C" tsvd_graph *tsvd_pipeline_graph(tsvd_pipeline *p) { return p ? p->g : nullptr; }
extern "C" tsvd_ppr *tsvd_pipeline_ppr(tsvd_pipeline *p) { return p ? (p->ppr ? p->ppr : p->sp.gp) : nullptr; }

extern "C" int tsvd_pipeline_source_state(tsvd_pipeline *p, int32_t dir, int32_t i, double *reserve,
                                          double *residue)
{
    return guard([&] {
        TSVD_REQUIRE(p, "null pipeline");
        const tsvd_config &c = p->cfg;
        TSVD_REQUIRE(dir >= 0 && dir < (c.directed ? 2 : 1), "bad direction");
        TSVD_REQUIRE(i >= 0 && i < c.nsrc, "source index out of range");
        if (c.sparse_state && !p->sp.resident) {
            const SparseState &ss = p->sp;
            const int32_t bt = i / (ss.gpb * 64), s = i - bt * ss.gpb * 64;
            const auto &L = ss.lists[(size_t)bt];
            if (reserve) list_read(L[(size_t)(2 * dir)], s, c.n, reserve);
            if (residue) list_read(L[(size_t)(2 * dir + 1)], s, c.n, residue);
            return;
        }
        tsvd_ppr *q = p->ppr ? p->ppr : p->sp.gp;   // the dense state, or the resident sparse one
        const int32_t lane = p->lane_of.empty() ? i : p->lane_of[(size_t)i];
        double *out[2] = {reserve, residue};
        for (int k = 0; k < 2; k++) {
            if (!out[k]) continue;
            ppr_read_lane(q, dir, k, lane, out[k]);
            if (c.sparse_state)   // the stored state: positive entries only
                for (int32_t v = 0; v < c.n; v++) out[k][v] = std::max(out[k][v], 0.0);
        }
    });
}

extern "C" int tsvd_pipeline_block_matrix(tsvd_pipeline *p, int32_t j, double *A)
{
    return guard([&] {
        TSVD_REQUIRE(p && A, "null argument");
        const Blocks &b = p->blocks;
        TSVD_REQUIRE(j >= 0 && j < b.nparts, "block out of range");
        const int64_t w = b.width[j], nc = b.ncols[j];
        std::vector<uint8_t> h((size_t)std::max<int64_t>(nc, 1) * b.s_pad * b.esize);
        std::vector<int32_t> cm((size_t)std::max<int64_t>(nc, 1));
        if (nc > 0) {
            TSVD_HIP(hipMemcpy(h.data(), b.A.p + (size_t)b.start[j] * b.s_pad * b.esize, (size_t)nc * b.s_pad * b.esize,
                               hipMemcpyDeviceToHost));
            TSVD_HIP(hipMemcpy(cm.data(), b.colmap.p + b.start[j], sizeof(int32_t) * nc, hipMemcpyDeviceToHost));
        }
        for (size_t t = 0; t < (size_t)w * b.nsrc; t++) A[t] = 0.0;
        for (int64_t c = 0; c < nc; c++) {
            const int64_t col = cm[c] - b.start[j];
            for (int32_t s = 0; s < b.nsrc; s++) {
                const size_t e = (size_t)c * b.s_pad + (size_t)p->lane_row(s);
                A[(size_t)col * b.nsrc + s] = b.esize == 8 ? ((double *)h.data())[e] : (double)((float *)h.data())[e];
            }
        }
    });
}

extern "C" int tsvd_pipeline_block_transform(tsvd_pipeline *p, int32_t j0, int32_t j1, int32_t iter, double delta,
                                             int32_t test, double *a_norm, double *e_norm, int32_t *record)
{
    return guard([&] {
        TSVD_REQUIRE(p, "null pipeline");
        TSVD_REQUIRE(!p->cfg.sparse_state && p->ppr, "the block transform reads the dense device state");
        Blocks &b = p->blocks;
        TSVD_REQUIRE(j0 >= 0 && j0 <= j1 && j1 <= b.nparts, "block range [%d, %d) outside [0, %d)", j0, j1,
                     b.nparts);
        TSVD_REQUIRE(test == 0 || test == 1, "test must be 0 or 1");
        TSVD_REQUIRE(delta >= 0.0, "delta must be non-negative");
        if (j1 == j0) return;
        hipStream_t st = p->g->stream;
        const void *p0 = p->ppr->p(0), *p1 = p->ppr->ndir > 1 ? p->ppr->p(1) : nullptr;
        blocks_norms_items(b, p0, p1, p->g, p->ppr->touched.p, p->ppr->touched_all, j0, j1, st);
        p->ppr->touched_all = false;
        // ppr:2718-2730: clean (-1) when resid + ||A_new - A_old|| < delta sqrt(2) ||A_new||
        const double thr = delta * std::sqrt(2.0);
        std::vector<int32_t> rebuild;
        for (int32_t j = j0; j < j1; j++) {
            const double a = std::sqrt(b.a2[j]), e = std::sqrt(b.e2[j]);
            const bool dirty = test == 0 || !(b.resid[j] + e < thr * a);
            if (a_norm) a_norm[j - j0] = a;
            if (e_norm) e_norm[j - j0] = e;
            if (record) record[j - j0] = dirty ? iter : -1;
            if (dirty) rebuild.push_back(j);
        }
        blocks_rebuild(b, rebuild, p0, p1, st);
    });
}

extern "C" int tsvd_pipeline_block_columns(tsvd_pipeline *p, int32_t j, int32_t *cols, int32_t *count)
{
    return guard([&] {
        TSVD_REQUIRE(p && count, "null argument");
        const Blocks &b = p->blocks;
        TSVD_REQUIRE(j >= 0 && j < b.nparts, "block out of range");
        const int32_t nc = b.ncols[j];
        *count = nc;
        if (cols && nc > 0)
            TSVD_HIP(hipMemcpy(cols, b.colmap.p + b.start[j], sizeof(int32_t) * nc, hipMemcpyDeviceToHost));
    });
}

extern "C" int tsvd_pipeline_block_norms(tsvd_pipeline *p, double *a_norm, double *e_norm, double *resid,
                                         int32_t *dirty)
{
    return guard([&] {
        TSVD_REQUIRE(p, "null pipeline");
        const Blocks &b = p->blocks;
        for (int32_t j = 0; j < b.nparts; j++) {
            if (a_norm) a_norm[j] = std::sqrt(b.a2[j]);
            if (e_norm) e_norm[j] = std::sqrt(b.e2[j]);
            if (resid) resid[j] = b.resid[j];
            if (dirty) dirty[j] = 0;
        }
        if (dirty)
            for (int32_t j : p->last_dirty) dirty[j] = 1;
    });
}

extern "C" int tsvd_pipeline_node(tsvd_pipeline *p, int32_t level, int32_t index, double *out)
{
    return guard([&] {
        TSVD_REQUIRE(p && out, "null argument");
        const Tree &t = p->tree;
        TSVD_REQUIRE(level >= 0 && level < (int32_t)t.sizes.size() - 1, "level out of range");
        TSVD_REQUIRE(index >= 0 && index < t.sizes[level], "node out of range");
        const size_t cnt = (size_t)t.s_pad * t.k;
        std::vector<uint8_t> h(cnt * t.esize);
        TSVD_HIP(hipMemcpy(h.data(), t.level[level].p + (size_t)index * cnt * t.esize, h.size(),
                           hipMemcpyDeviceToHost));
        for (int32_t c = 0; c < t.k; c++)
            for (int32_t s = 0; s < t.nsrc; s++) {
                const size_t e = (size_t)c * t.s_pad + (size_t)p->lane_row(s);
                out[(size_t)c * t.nsrc + s] = t.esize == 8 ? ((double *)h.data())[e] : (double)((float *)h.data())[e];
            }
    });
}

// FP64 Gram A A^T (nsrc x nsrc, row-major) of the current transformed matrix
// over all n columns, accumulated slab by slab on MFMA.
extern "C" int tsvd_pipeline_dense_gram(tsvd_pipeline *p, double *G)
{
    return guard([&] {
        TSVD_REQUIRE(p && G, "null argument");
        TSVD_REQUIRE(p->ppr, "the dense Gram reads the device state (not kept with sparse_state)");
        const Blocks &b = p->blocks;
        hipStream_t st = p->g->stream;
        const int64_t chunk = 65536;
        const int32_t ns = b.nsrc;
        DBuf<double> X((size_t)chunk * b.s_pad), Gd((size_t)ns * ns);
        TSVD_HIP(hipMemsetAsync(Gd.p, 0, Gd.bytes(), st));
        const void *p0 = p->ppr->p(0);
        const void *p1 = p->ppr->ndir > 1 ? p->ppr->p(1) : nullptr;
        for (int64_t v0 = 0; v0 < b.n; v0 += chunk) {
            const int64_t nv = std::min<int64_t>(chunk, b.n - v0);
            transform_chunk(b, p0, p1, X.p, v0, nv, st);
            std::vector<GemmProb> gp(1);
            gp[0] = GemmProb{};
            gp[0].A = X.p; gp[0].lda = b.s_pad; gp[0].transA = 0;
            gp[0].B = X.p; gp[0].ldb = b.s_pad; gp[0].transB = 1;
            gp[0].C = Gd.p; gp[0].ldc = ns;
            gp[0].M = ns; gp[0].N = ns; gp[0].K = (int32_t)nv; gp[0].alpha = 1.0; gp[0].beta = 1.0;
            p->gemm.run(st, gp, GemmTypes{DT::F64, DT::F64, DT::F64, DT::F64}, true, "gemm_dense_gram");
        }
        TSVD_HIP(hipMemcpyAsync(G, Gd.p, Gd.bytes(), hipMemcpyDeviceToHost, st));
        TSVD_HIP(hipStreamSynchronize(st));
    });
}

// Z = A (A^T Y) in FP64 for an nsrc x l block Y, slab by slab: T = X^T Y
// (nv x l), then Z += X T.  Subspace iteration on A A^T without the Gram.
extern "C" int tsvd_pipeline_gram_apply(tsvd_pipeline *p, const double *Y, int32_t l, double *Z)
{
    return guard([&] {
        TSVD_REQUIRE(p && Y && Z, "null argument");
        TSVD_REQUIRE(p->ppr, "gram_apply reads the device state (not kept with sparse_state)");
        const Blocks &b = p->blocks;
        TSVD_REQUIRE(l >= 1 && l <= b.nsrc, "l must be in [1, nsrc]");
        hipStream_t st = p->g->stream;
        const int64_t chunk = 65536;
        const int32_t ns = b.nsrc;
        // row-major host blocks <-> column-major device blocks (ld = nsrc)
        std::vector<double> h((size_t)ns * l);
        for (int32_t s = 0; s < ns; s++)
            for (int32_t c = 0; c < l; c++) h[(size_t)c * ns + s] = Y[(size_t)s * l + c];
        DBuf<double> X((size_t)chunk * b.s_pad), T((size_t)chunk * l), Yd(h.size()), Zd(h.size());
        TSVD_HIP(hipMemcpyAsync(Yd.p, h.data(), Yd.bytes(), hipMemcpyHostToDevice, st));
        TSVD_HIP(hipMemsetAsync(Zd.p, 0, Zd.bytes(), st));
        const void *p0 = p->ppr->p(0);
        const void *p1 = p->ppr->ndir > 1 ? p->ppr->p(1) : nullptr;
        const GemmTypes f64{DT::F64, DT::F64, DT::F64, DT::F64};
        for (int64_t v0 = 0; v0 < b.n; v0 += chunk) {
            const int64_t nv = std::min<int64_t>(chunk, b.n - v0);
            transform_chunk(b, p0, p1, X.p, v0, nv, st);
            std::vector<GemmProb> gp(1);
            gp[0] = GemmProb{};
            gp[0].A = X.p; gp[0].lda = b.s_pad; gp[0].transA = 1;
            gp[0].B = Yd.p; gp[0].ldb = ns;
            gp[0].C = T.p; gp[0].ldc = nv;
            gp[0].M = (int32_t)nv; gp[0].N = l; gp[0].K = ns; gp[0].alpha = 1.0; gp[0].beta = 0.0;
            p->gemm.run(st, gp, f64, true, "gemm_gram_apply");
            gp[0] = GemmProb{};
            gp[0].A = X.p; gp[0].lda = b.s_pad; gp[0].transA = 0;
            gp[0].B = T.p; gp[0].ldb = nv;
            gp[0].C = Zd.p; gp[0].ldc = ns;
            gp[0].M = ns; gp[0].N = l; gp[0].K = (int32_t)nv; gp[0].alpha = 1.0; gp[0].beta = 1.0;
            gp[0].kchunk = 4096;  // split K across workgroups, partial sums added atomically
            p->gemm.run(st, gp, f64, true, "gemm_gram_apply");
        }
        TSVD_HIP(hipMemcpyAsync(h.data(), Zd.p, Zd.bytes(), hipMemcpyDeviceToHost, st));
        TSVD_HIP(hipStreamSynchronize(st));
        for (int32_t s = 0; s < ns; s++)
            for (int32_t c = 0; c < l; c++) Z[(size_t)s * l + c] = h[(size_t)c * ns + s];
    });
}

extern "C" int tsvd_pipeline_right_embedding(tsvd_pipeline *p, int32_t v0, int32_t v1, double *V)
{
    return guard([&] {
        TSVD_REQUIRE(p && V, "null argument");
        TSVD_REQUIRE(p->cfg.right_embedding, "the pipeline was created without right_embedding");
        const Blocks &b = p->blocks;
        TSVD_REQUIRE(v0 >= 0 && v0 <= v1 && v1 <= b.n, "bad vertex range");
        const int32_t k = p->cfg.rank;
        std::vector<int32_t> slot(b.n);
        TSVD_HIP(hipMemcpy(slot.data(), b.slot.p, sizeof(int32_t) * b.n, hipMemcpyDeviceToHost));
        std::vector<double> h((size_t)b.n * k);
        TSVD_HIP(hipMemcpy(h.data(), p->re.V.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
        for (int32_t v = v0; v < v1; v++) {
            const int32_t j = std::min(v / b.common, b.nparts - 1);
            const int32_t c = slot[v];
            for (int32_t t = 0; t < k; t++)
                V[(size_t)(v - v0) * k + t] =
                    (c >= 0 && p->re.ver[j] >= 0) ? h[(size_t)b.start[j] * k + (size_t)t * b.width[j] + c] : 0.0;
        }
    });
}

extern "C" int tsvd_pipeline_lp_scores(tsvd_pipeline *p, const int32_t *src, const int32_t *dst, int64_t m,
                                       double *scores)
{
    return guard([&] {
        TSVD_REQUIRE(p && (m == 0 || (src && dst && scores)) && m >= 0, "bad arguments");
        TSVD_REQUIRE(p->cfg.right_embedding, "the pipeline was created without right_embedding");
        if (m == 0) return;
        std::vector<int32_t> rows((size_t)m);
        for (int64_t e = 0; e < m; e++) {
            auto it = p->re.row_of.find(src[e]);
            TSVD_REQUIRE(it != p->re.row_of.end(), "pair %lld: source %d is not a subset vertex", (long long)e,
                         src[e]);
            TSVD_REQUIRE(dst[e] >= 0 && dst[e] < p->cfg.n, "pair %lld: target %d out of range", (long long)e, dst[e]);
            rows[(size_t)e] = p->lane_row(it->second);   // the root U rows are push lanes
        }
        hipStream_t st = p->g->stream;
        DBuf<int32_t> d_rows((size_t)m), d_dst((size_t)m);
        DBuf<double> d_sc((size_t)m);
        TSVD_HIP(hipMemcpyAsync(d_rows.p, rows.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, st));
        TSVD_HIP(hipMemcpyAsync(d_dst.p, dst, sizeof(int32_t) * m, hipMemcpyHostToDevice, st));
        lp_scores(p->blocks, p->re.V.p, p->re.ublk.p, p->cfg.sparse_state ? p->re.dblk.p : nullptr, p->re.vdeg.p,
                  p->cfg.rank, d_rows.p, d_dst.p, m, d_sc.p, st);
        TSVD_HIP(hipMemcpyAsync(scores, d_sc.p, sizeof(double) * m, hipMemcpyDeviceToHost, st));
        TSVD_HIP(hipStreamSynchronize(st));
    });
}
