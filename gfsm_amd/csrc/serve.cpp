// serve.cpp -- placeholder serving loop: one worker, FIFO over the plain path.
#include "serve.h"

#include <chrono>

namespace gfsm {

struct Server {
    Ctx *ctx = nullptr;
    const HostGraph *data = nullptr;
    gfsm_serve_opts opts{};
    std::mutex mu;
    std::condition_variable cv;
    struct Item {
        const HostGraph *q;
        double t[3];
        bool done = false;
        Result res;
    };
    std::map<uint64_t, Item> items;
    std::deque<uint64_t> queue;
    uint64_t next = 1;
    bool stop = false;
    std::thread worker;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

static void worker_main(Server *s)
{
    for (;;) {
        uint64_t id;
        const HostGraph *q;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [s] { return s->stop || !s->queue.empty(); });
            if (s->queue.empty()) return;
            id = s->queue.front();
            s->queue.pop_front();
            q = s->items[id].q;
            s->items[id].t[1] = s->now();
        }
        gfsm_match_opts o{};
        o.flags = s->opts.flags;
        Result r;
        Status st = match_plain(*s->ctx, *s->data, *q, o, r);
        if (!st.good()) r.status = st.code;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            auto &it = s->items[id];
            it.res = std::move(r);
            it.t[2] = s->now();
            it.done = true;
        }
        s->cv.notify_all();
    }
}

Status server_start(Ctx &ctx, const HostGraph &data, const gfsm_serve_opts &opts, Server *&out)
{
    Server *s = new Server;
    s->ctx = &ctx;
    s->data = &data;
    s->opts = opts;
    s->worker = std::thread(worker_main, s);
    out = s;
    return Status::ok();
}

Status server_submit(Server &s, const HostGraph &query, uint64_t &ticket)
{
    {
        std::lock_guard<std::mutex> lk(s.mu);
        ticket = s.next++;
        auto &it = s.items[ticket];
        it.q = &query;
        it.t[0] = s.now();
        s.queue.push_back(ticket);
    }
    s.cv.notify_all();
    return Status::ok();
}

Status server_take(Server &s, uint64_t ticket, bool wait, Result &out)
{
    std::unique_lock<std::mutex> lk(s.mu);
    auto it = s.items.find(ticket);
    if (it == s.items.end()) return Status::err(GFSM_ERR_ARG, "unknown ticket");
    if (!it->second.done) {
        if (!wait) return Status::err(GFSM_ERR_PENDING, "pending");
        s.cv.wait(lk, [&] { return it->second.done; });
    }
    out = std::move(it->second.res);
    return Status::ok();
}

Status server_times(Server &s, uint64_t ticket, double t[3])
{
    std::lock_guard<std::mutex> lk(s.mu);
    auto it = s.items.find(ticket);
    if (it == s.items.end()) return Status::err(GFSM_ERR_ARG, "unknown ticket");
    for (int i = 0; i < 3; i++) t[i] = it->second.t[i];
    return Status::ok();
}

void server_stop(Server *s)
{
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    delete s;
}

}  // namespace gfsm
