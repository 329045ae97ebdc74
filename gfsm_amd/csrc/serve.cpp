// serve.cpp -- the serving loop: a resident data graph, queries submitted at
// any time, continuous batching (no reference counterpart; the reference is
// one process per query, main.cu:56-63).
//
// One worker thread owns a BatchEngine.  Between two replays of the captured
// steps it harvests the slots that finished and admits waiting queries in
// arrival order while a slot is free and the arena has room for the query's
// projected intermediate levels; a query that does not fit waits at the head
// of the queue (no overtaking, so no starvation).  Admitted queries join the
// steps already in flight, so their work overlaps on the device with every
// other query's.  Results do not depend on what runs beside them: each slot
// has its own levels, its own scan range and its own totals.
#include "serve.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <chrono>

namespace gfsm {

struct Server {
    Ctx *ctx = nullptr;
    const HostGraph *data = nullptr;
    gfsm_serve_opts opts{};
    std::unique_ptr<BatchEngine> eng;
    std::mutex mu;
    std::condition_variable cv_done, cv_work;
    // A ticket's record: its timestamps and, until taken, its result.  The
    // record is dropped once the result has been taken and the timestamps
    // read; records of taken results whose timestamps are never read are
    // dropped oldest first beyond kTakenKept, so a long-running server holds
    // host memory only for tickets in flight or not yet taken.
    struct Item {
        double t[3] = {0, 0, 0};  // submit, admit, done (seconds on the server clock)
        bool done = false, taken = false, timed = false;
        std::unique_ptr<Result> res;
    };
    static constexpr size_t kTakenKept = 1u << 20;
    std::map<uint64_t, Item> items;
    std::deque<uint64_t> taken_untimed;  // taken tickets whose timestamps were not read yet
    gfsm_serve_stats counters{};        // guarded by mu
    std::deque<std::unique_ptr<BatchJob>> incoming;  // submitted, not yet seen by the worker
    uint64_t next = 1;
    bool stop = false;
    Status fatal;  // a device error ends the loop; later tickets fail with it
    std::thread worker;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    void finish(uint64_t id, Result &r)
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            auto &it = items[id];
            it.res.reset(new Result(std::move(r)));
            it.t[2] = now();
            it.done = true;
            counters.completed++;
        }
        cv_done.notify_all();
    }
};

static void worker_main(Server *s)
{
    // GFSM_BATCH_TRACE: seconds spent admitting and advancing, at the end.
    const bool trace = std::getenv("GFSM_BATCH_TRACE") != nullptr;
    double t_admit = 0, t_advance = 0;
    uint64_t n_advance = 0;
    std::deque<std::unique_ptr<BatchJob>> pending;
    // A job's query stays pinned until its result is recorded.
    auto done = [s](std::unique_ptr<BatchJob> job, Result &r) {
        const HostGraph *q = job->query;
        s->finish(job->id, r);
        job.reset();
        unpin_graph(*q);
    };
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (pending.empty() && s->eng->running() == 0)
                s->cv_work.wait(lk, [s] { return s->stop || !s->incoming.empty(); });
            while (!s->incoming.empty()) {
                pending.push_back(std::move(s->incoming.front()));
                s->incoming.pop_front();
            }
            if (s->stop && pending.empty() && s->eng->running() == 0) {
                if (trace)
                    std::fprintf(stderr, "server worker: admission %.3f s, %llu advances %.3f s\n", t_admit,
                                 (unsigned long long)n_advance, t_advance);
                return;
            }
        }
        // Admission in arrival order while a slot and projected memory are free.
        const double ta = trace ? s->now() : 0;
        while (!pending.empty()) {
            BatchJob &job = *pending.front();
            const uint64_t id = job.id;
            if (job.plan.q == 1 || (job.opts.flags & (GFSM_ROWS | GFSM_SORTED))) {
                // No device steps (one vertex) or rows wanted: single-query path.
                {
                    std::lock_guard<std::mutex> lk(s->mu);
                    s->items[id].t[1] = s->now();
                }
                Result r;
                Status st = match_plain(*s->ctx, *s->data, *job.query, job.opts, r);
                if (!st.good() && st.code != GFSM_ERR_CAP) r.status = st.code;
                done(std::move(pending.front()), r);
                pending.pop_front();
                continue;
            }
            Status st;
            const double t = s->now();
            const uint32_t free_slots = s->eng->free_slots();
            if (!s->eng->try_admit(pending.front(), st)) {
                if (st.good()) {  // waits: for projected memory while a slot is free, or for a slot
                    if (free_slots) pending.front()->waited_memory = true;
                    else pending.front()->waited_slot = true;
                }
                if (!st.good()) {  // can never fit
                    Result r;
                    r.status = st.code;
                    done(std::move(pending.front()), r);
                    pending.pop_front();
                    continue;
                }
                break;
            }
            std::lock_guard<std::mutex> lk(s->mu);
            s->items[id].t[1] = t;
            s->counters.admitted++;
            s->counters.waited_memory += job.waited_memory;
            s->counters.waited_slot += job.waited_slot;
            pending.pop_front();
        }
        if (trace) t_admit += s->now() - ta;
        if (s->eng->running()) {
            const double tv = trace ? s->now() : 0;
            Status st = s->eng->advance(done);
            if (trace) {
                t_advance += s->now() - tv;
                n_advance++;
            }
            if (!st.good()) {
                std::lock_guard<std::mutex> lk(s->mu);
                s->fatal = st;
                s->stop = true;
                break;
            }
        }
    }
    // Device error: fail everything still queued.
    for (auto &j : pending) {
        Result r;
        r.status = GFSM_ERR_DEVICE;
        done(std::move(j), r);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto &kv : s->items)
        if (!kv.second.done) {
            kv.second.res.reset(new Result);
            kv.second.res->status = GFSM_ERR_DEVICE;
            kv.second.done = true;
        }
    s->cv_done.notify_all();
}

// Stop the worker once the submitted work is done, release the engine and
// the data graph, and leave the context; the Server object stays until
// server_stop (its finished results can still be taken).
static void server_halt(void *p)
{
    Server *s = static_cast<Server *>(p);
    if (!s->ctx) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    s->eng.reset();
    {
        std::lock_guard<std::mutex> lk(s->ctx->srv_mu);
        s->ctx->servers.erase(s);
    }
    unpin_graph(*s->data);
    s->ctx = nullptr;
    s->data = nullptr;
}

Status server_start(Ctx &ctx, const HostGraph &data, const gfsm_serve_opts &opts, Server *&out)
{
    out = nullptr;
    std::unique_ptr<Server> s(new Server);
    s->ctx = &ctx;
    s->data = &data;
    s->opts = opts;
    BatchShape sh;
    sh.slots = opts.max_inflight ? opts.max_inflight : ctx.tune.batch_slots;
    sh.items = ctx.tune.batch_items;
    sh.pool_words = ctx.tune.batch_pool_words;
    sh.wgs_per_cu = ctx.tune.batch_wgs_per_cu;
    // Short replays: a finished slot is refilled at the end of its replay.
    sh.steps_per_replay = std::max<uint32_t>(ctx.tune.serve_steps, 1);
    sh.step_budget = ctx.tune.batch_step_budget;
    if (sh.slots > kMaxSlots) return Status::err(GFSM_ERR_ARG, "max_inflight must be at most 1024");
    const uint64_t budget = opts.mem_budget ? opts.mem_budget : ctx.budget;
    s->eng.reset(new BatchEngine(ctx, data));
    Status st = s->eng->init(sh, budget);
    if (!st.good()) return st;
    // One query may take up to a sixteenth of the budget (its projection
    // decides how much below that).
    s->eng->set_max_region(std::max(s->eng->min_region(GFSM_MAXQ), budget / 16));
    pin_graph(data);
    {
        std::lock_guard<std::mutex> lk(ctx.srv_mu);
        ctx.servers.insert(s.get());
        ctx.halt_server = server_halt;
    }
    s->worker = std::thread(worker_main, s.get());
    out = s.release();
    return Status::ok();
}

Status server_submit(Server &s, const HostGraph &query, const gfsm_match_opts *opts, uint64_t &ticket)
{
    {
        std::lock_guard<std::mutex> lk(s.mu);
        if (s.stop || !s.data) return s.fatal.good() ? Status::err(GFSM_ERR_ARG, "server stopped") : s.fatal;
    }
    std::unique_ptr<BatchJob> job(new BatchJob);
    job->query = &query;
    if (opts) job->opts = *opts;
    job->opts.flags |= s.opts.flags;
    Status plan_st = (query.mode == s.data->mode) ? plan_for(*s.data, query, job->opts, job->plan)
                                                   : Status::err(GFSM_ERR_ARG, "query and data graph use different modes");
    // The level model is worked out here, on the submitting thread, so the
    // worker's admission loop only compares sizes.
    if (plan_st.good() && job->plan.q > 1) BatchEngine::model_job(*s.data, *job);
    {
        std::lock_guard<std::mutex> lk(s.mu);
        if (s.stop) return s.fatal.good() ? Status::err(GFSM_ERR_ARG, "server stopped") : s.fatal;
        ticket = s.next++;
        job->id = ticket;
        auto &it = s.items[ticket];
        it.t[0] = s.now();
        if (!plan_st.good()) {
            it.res.reset(new Result);
            it.res->status = plan_st.code;
            it.t[1] = it.t[2] = it.t[0];
            it.done = true;
        } else {
            pin_graph(query);  // released when its result is recorded
            s.incoming.push_back(std::move(job));
        }
        s.counters.submitted++;
    }
    s.cv_work.notify_all();
    return Status::ok();
}

Status server_take(Server &s, uint64_t ticket, bool wait, Result &out)
{
    std::unique_lock<std::mutex> lk(s.mu);
    auto it = s.items.find(ticket);
    if (it == s.items.end()) return Status::err(GFSM_ERR_ARG, "unknown ticket");
    if (it->second.taken) return Status::err(GFSM_ERR_ARG, "the ticket's result was already taken");
    if (!it->second.done) {
        if (!wait) return Status::err(GFSM_ERR_PENDING, "pending");
        s.cv_done.wait(lk, [&] { return it->second.done; });
    }
    Server::Item &item = it->second;
    out = item.res ? std::move(*item.res) : Result{};
    if (!item.res) out.status = GFSM_ERR_DEVICE;
    item.res.reset();
    item.taken = true;
    if (item.timed) {
        s.items.erase(it);
    } else {
        s.taken_untimed.push_back(ticket);
        while (s.taken_untimed.size() > Server::kTakenKept) {
            s.items.erase(s.taken_untimed.front());
            s.taken_untimed.pop_front();
        }
    }
    return Status::ok();
}

Status server_stats(Server &s, gfsm_serve_stats &out)
{
    std::lock_guard<std::mutex> lk(s.mu);
    out = s.counters;
    return Status::ok();
}

Status server_times(Server &s, uint64_t ticket, double t[3])
{
    std::lock_guard<std::mutex> lk(s.mu);
    auto it = s.items.find(ticket);
    if (it == s.items.end()) return Status::err(GFSM_ERR_ARG, "unknown ticket (or its record was dropped)");
    for (int i = 0; i < 3; i++) t[i] = it->second.t[i];
    it->second.timed = true;
    if (it->second.taken) s.items.erase(it);  // its entry in taken_untimed is skipped when evicted
    return Status::ok();
}

void server_stop(Server *s)
{
    server_halt(s);
    delete s;
}

}  // namespace gfsm
