// server.cpp -- the served form: continuous batching of insertion batches.
//
// The reference has no served mode: its drivers read one snapshot file per
// loop iteration and run the whole update for it (Tree_u_NC.cpp:507-866).
// Here a server owns one pipeline (all state stays resident in HBM: graph,
// residues, reserves, stored blocks, every tree node) and one worker thread.
// Clients submit insertion batches (requests) at any time.  Each step the
// worker admits the queued requests in arrival order, up to an edge budget
// (from free HBM unless given) and a request cap, appends them as ONE
// snapshot and completes them together with that snapshot's embedding.
// Requests that arrive while a step runs are admitted by the next step, so
// the per-snapshot fixed cost (push rounds over the hub rows, the full
// transform pass, dirty-path merges) is paid once per step instead of once
// per request.  A request cap of 1 is the plain per-snapshot method under
// the same arrival stream.
//
// Steps overlap on the device: a second thread runs the back part of step t
// (frPCA of the dirty blocks, merges, root embedding) on its own stream while
// the worker runs the front part of step t+1 (append, insertion update,
// push, block norms) on the graph's stream.  The worker then waits for step
// t to finish before the dirty test and block rebuild of step t+1, which
// read the residuals and rewrite the blocks step t uses.  With the LP right
// embedding the steps run one after the other (its scores read the degrees
// of their own snapshot).
//
// Appending batches b1..bk as one snapshot gives the same CSR rows as
// appending them one by one (each row gains its neighbours in batch order),
// and the insertion update walks every new slot in order, so the push
// invariant holds exactly either way; the reserves agree within the push
// error bound.  A request's result is the embedding of the state that
// includes every request up to the step's last ticket.
#include "tsvd_pipeline.h"

#include <chrono>
#include <cstdlib>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

namespace tsvd {

struct Request {
    int64_t ticket;
    std::vector<int32_t> src, dst;
    double submit_ms;
};

struct Done {
    int rc = TSVD_OK;
    std::string err;
    tsvd_request_result res{};
    std::shared_ptr<std::vector<double>> E;
};

struct Job {
    std::vector<Request> step;
    int64_t edges = 0;
    double t_start = 0.0;
    StepPlan plan;
    int rc = TSVD_OK;
    std::string err;
};

}  // namespace tsvd

using namespace tsvd;

struct tsvd_server {
    tsvd_pipeline *pipe = nullptr;
    tsvd_server_config scfg{};
    int64_t edge_budget = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request> queue;
    std::map<int64_t, Done> done;
    int64_t next_ticket = 0;
    int64_t completed_below = 0;   // steps complete tickets in submit order
    bool stop = false;
    std::thread worker;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    // back parts: one job at a time, on their own stream, by the finisher thread
    bool overlap = false;
    hipStream_t back_stream = nullptr;
    std::mutex fmu;
    std::condition_variable fcv;
    std::unique_ptr<Job> pending;
    bool back_busy = false, fstop = false;
    std::thread finisher;

    double now_ms() const
    {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    void run();
    void finish_loop();
    void finish(Job &job, hipStream_t st);
    void wait_back_idle()
    {
        std::unique_lock<std::mutex> lk(fmu);
        fcv.wait(lk, [&] { return !back_busy; });
    }
    // Drains the queue, then stops the worker, the finisher, the pipeline and
    // the back stream, in that order; also reached when creation fails part
    // way, with whichever of them had been started.
    ~tsvd_server()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_work.notify_all();
        if (worker.joinable()) worker.join();
        {
            std::lock_guard<std::mutex> lk(fmu);
            fstop = true;
        }
        fcv.notify_all();
        if (finisher.joinable()) finisher.join();
        if (pipe) tsvd_pipeline_destroy(pipe);
        if (back_stream) (void)hipStreamDestroy(back_stream);
    }
};

void tsvd_server::run()
{
    for (;;) {
        std::unique_ptr<Job> job(new Job());
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !queue.empty(); });
            if (queue.empty()) break;   // stop requested and the queue is drained
            while (!queue.empty()) {
                const Request &r = queue.front();
                const int64_t m = (int64_t)r.src.size();
                if (!job->step.empty()) {
                    if (scfg.max_step_requests > 0 && (int32_t)job->step.size() >= scfg.max_step_requests) break;
                    if (job->edges + m > edge_budget) break;
                }
                job->edges += m;
                job->step.push_back(std::move(queue.front()));
                queue.pop_front();
            }
        }
        job->t_start = now_ms();
        std::vector<int32_t> src, dst;
        src.reserve((size_t)job->edges);
        dst.reserve((size_t)job->edges);
        for (const Request &r : job->step) {
            src.insert(src.end(), r.src.begin(), r.src.end());
            dst.insert(dst.end(), r.dst.begin(), r.dst.end());
        }
        bool front = false;
        int rc = guard([&] { front = pipeline_front(pipe, src.data(), dst.data(), job->edges, 0, job->plan); });
        if (overlap) wait_back_idle();   // the previous step's back part is done
        if (rc == TSVD_OK && front) rc = guard([&] { pipeline_decide(pipe, job->plan); });
        job->rc = rc;
        if (rc != TSVD_OK) job->err = tsvd_last_error();
        if (overlap) {
            std::lock_guard<std::mutex> lk(fmu);
            pending = std::move(job);
            back_busy = true;
            fcv.notify_all();
        } else {
            finish(*job, pipe->g->stream);
        }
    }
    if (overlap) wait_back_idle();
}

void tsvd_server::finish_loop()
{
    for (;;) {
        std::unique_ptr<Job> job;
        {
            std::unique_lock<std::mutex> lk(fmu);
            fcv.wait(lk, [&] { return fstop || pending; });
            if (!pending) return;
            job = std::move(pending);
        }
        finish(*job, back_stream);
        {
            std::lock_guard<std::mutex> lk(fmu);
            back_busy = false;
        }
        fcv.notify_all();
    }
}

// Back part of a step, the embedding, and the completion of its requests.
void tsvd_server::finish(Job &job, hipStream_t st)
{
    const int32_t ns = pipe->cfg.nsrc, k = pipe->cfg.rank;
    int rc = job.rc;
    std::string err = job.err;
    if (rc == TSVD_OK) rc = guard([&] { pipeline_back(pipe, job.plan, st); });
    auto E = std::make_shared<std::vector<double>>((size_t)ns * k, 0.0);
    if (rc == TSVD_OK) rc = tsvd_pipeline_embedding(pipe, E->data());
    if (rc != TSVD_OK && err.empty()) err = tsvd_last_error();
    const double t_done = now_ms();
    const tsvd_snapshot_stats &st_ = job.plan.S;
    {
        std::lock_guard<std::mutex> lk(mu);
        for (const Request &r : job.step) {
            Done &d = done[r.ticket];
            d.rc = rc;
            d.err = err;
            d.E = E;
            d.res.ticket = r.ticket;
            d.res.snapshot = st_.snapshot;
            d.res.step_requests = (int32_t)job.step.size();
            d.res.step_edges = job.edges;
            d.res.last_ticket = job.step.back().ticket;
            d.res.submit_ms = r.submit_ms;
            d.res.start_ms = job.t_start;
            d.res.done_ms = t_done;
            d.res.latency_ms = t_done - r.submit_ms;
            d.res.step = st_;
        }
        completed_below = job.step.back().ticket + 1;
    }
    cv_done.notify_all();
}

extern "C" int tsvd_server_create(const tsvd_config *cfg, const tsvd_server_config *scfg, tsvd_server **out)
{
    return guard([&] {
        TSVD_REQUIRE(cfg && out, "null argument");
        tsvd_server_config sc{};
        if (scfg) sc = *scfg;
        TSVD_REQUIRE(sc.max_step_edges >= 0 && sc.max_step_requests >= 0, "negative server limits");
        std::unique_ptr<tsvd_server> s(new tsvd_server());
        s->scfg = sc;
        int rc = tsvd_pipeline_create(cfg, &s->pipe);
        if (rc) throw Error{rc};
        s->overlap = !sc.no_overlap && !cfg->right_embedding;
        if (s->overlap) {
            TSVD_HIP(hipStreamCreateWithFlags(&s->back_stream, hipStreamNonBlocking));
            s->finisher = std::thread([p = s.get()] { p->finish_loop(); });
        }
        // TSVD_FAIL_AFTER_FINISHER (tests only): fail here, with the finisher
        // running and the worker not started, to exercise the destructor's
        // teardown of a half-built server
        TSVD_REQUIRE(!getenv("TSVD_FAIL_AFTER_FINISHER"), "server creation failed after its finisher started");
        if (sc.max_step_edges > 0) {
            s->edge_budget = sc.max_step_edges;
        } else {
            // admission by free HBM: an append stages about 64 B per edge
            // (expanded keys/values, radix-sort buffers, scatter ranks); keep
            // half of what is free for CSR rebuilds and the solver workspaces
            size_t fr = 0, tot = 0;
            TSVD_HIP(hipMemGetInfo(&fr, &tot));
            s->edge_budget = std::max<int64_t>(1, (int64_t)(fr / 2 / 64));
        }
        s->worker = std::thread([p = s.get()] { p->run(); });
        *out = s.release();
    });
}

extern "C" int tsvd_server_destroy(tsvd_server *s)
{
    return guard([&] { delete s; });
}

extern "C" int tsvd_server_submit(tsvd_server *s, const int32_t *src, const int32_t *dst, int64_t m, int64_t *ticket)
{
    return guard([&] {
        TSVD_REQUIRE(s && ticket, "null argument");
        TSVD_REQUIRE(m >= 0 && (m == 0 || (src && dst)), "bad edge arrays");
        const int32_t n = s->pipe->cfg.n;
        for (int64_t e = 0; e < m; e++)
            TSVD_REQUIRE(src[e] >= 0 && src[e] < n && dst[e] >= 0 && dst[e] < n,
                         "edge %lld (%d, %d) out of range [0, %d)", (long long)e, src[e], dst[e], n);
        Request r;
        r.src.assign(src, src + m);
        r.dst.assign(dst, dst + m);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            TSVD_REQUIRE(!s->stop, "server is shutting down");
            r.ticket = s->next_ticket++;
            r.submit_ms = s->now_ms();
            *ticket = r.ticket;
            s->queue.push_back(std::move(r));
        }
        s->cv_work.notify_one();
    });
}

extern "C" int tsvd_server_wait(tsvd_server *s, int64_t ticket, double *E, tsvd_request_result *res)
{
    Done d;
    int rc = guard([&] {
        TSVD_REQUIRE(s, "null server");
        std::unique_lock<std::mutex> lk(s->mu);
        TSVD_REQUIRE(ticket >= 0 && ticket < s->next_ticket, "unknown ticket %lld", (long long)ticket);
        s->cv_done.wait(lk, [&] { return ticket < s->completed_below; });
        auto it = s->done.find(ticket);
        TSVD_REQUIRE(it != s->done.end(), "ticket %lld was already collected", (long long)ticket);
        d = std::move(it->second);
        s->done.erase(it);
    });
    if (rc) return rc;
    if (d.rc) return set_error(d.rc, "%s", d.err.c_str());
    if (res) *res = d.res;
    if (E) std::copy(d.E->begin(), d.E->end(), E);
    return TSVD_OK;
}

extern "C" tsvd_pipeline *tsvd_server_pipeline(tsvd_server *s)
{
    return s ? s->pipe : nullptr;
}

extern "C" int64_t tsvd_server_edge_budget(const tsvd_server *s)
{
    return s ? s->edge_budget : 0;
}
