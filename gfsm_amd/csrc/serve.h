// serve.h -- batched matching and the serving loop (no reference counterpart).
#pragma once

#include <condition_variable>
#include <deque>
#include <map>
#include <thread>

#include "engine.h"

namespace gfsm {

// Many queries at once: every depth step is one launch over the tiles of all
// queries still running.
Status match_batch(Ctx &ctx, const HostGraph &data, const std::vector<const HostGraph *> &queries,
                   const gfsm_match_opts &opts, std::vector<Result> &out);

struct Server;
Status server_start(Ctx &ctx, const HostGraph &data, const gfsm_serve_opts &opts, Server *&out);
Status server_submit(Server &s, const HostGraph &query, uint64_t &ticket);
Status server_take(Server &s, uint64_t ticket, bool wait, Result &out);
Status server_times(Server &s, uint64_t ticket, double t[3]);
void server_stop(Server *s);

}  // namespace gfsm
