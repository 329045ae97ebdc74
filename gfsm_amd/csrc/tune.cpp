// tune.cpp -- the autotuner (no reference counterpart: the reference
// launches fixed shapes, HnGPU.cu:45-47).  It times a caller-chosen sample
// (queries with their windows and caps) under each candidate launch shape
// and keeps the fastest in the context:
//   single-query path: candidates per thread x LDS staging pool x
//                      persistent workgroups per CU,
//   batch path:        the same three for the batch step kernel.
// Every shape computes the same result; only time is compared, as the best
// of kRepeats runs per shape (one run is noisy at millisecond scale).
#include <algorithm>
#include <cstdio>
#include <cstdlib>

#include "serve.h"

namespace gfsm {

namespace {

constexpr int kRepeats = 2;
// A shape replaces the one in use only when it is this much faster, so
// timing noise does not move a well-chosen default.
constexpr double kMargin = 0.03;

struct Sample {
    const HostGraph &data;
    const std::vector<const HostGraph *> &queries;
    const std::vector<gfsm_match_opts> &opts;  // one per query (windows, caps)
};

bool ok_or_capped(gfsm_status s) { return s == GFSM_OK || s == GFSM_ERR_CAP; }

// Device milliseconds of the sample on the single-query path.
Status time_plain(Ctx &ctx, const Sample &s, double &best)
{
    best = 1e300;
    for (int rep = 0; rep < kRepeats; rep++) {
        double ms = 0;
        for (size_t i = 0; i < s.queries.size(); i++) {
            Result r;
            Status st = match_plain(ctx, s.data, *s.queries[i], s.opts[i], r);
            if (!st.good() && st.code != GFSM_ERR_CAP) return st;
            ms += r.stats.device_ms;
        }
        best = std::min(best, ms);
    }
    return Status::ok();
}

Status time_batch(Ctx &ctx, const Sample &s, const BatchShape &shape, double &best)
{
    best = 1e300;
    for (int rep = 0; rep < kRepeats; rep++) {
        std::vector<Result> out;
        double ms = 0;
        Status st = match_batch_shaped(ctx, s.data, s.queries, s.opts, out, shape, &ms);
        if (!st.good()) return st;
        for (auto &r : out)
            if (!ok_or_capped(r.status)) return Status::err(r.status, "autotune sample query failed");
        best = std::min(best, ms);
    }
    return Status::ok();
}

}  // namespace

Status autotune(Ctx &ctx, const HostGraph &data, const std::vector<const HostGraph *> &queries,
                const std::vector<gfsm_match_opts> &opts, uint32_t which)
{
    if (queries.empty() || opts.size() != queries.size()) return Status::err(GFSM_ERR_ARG, "bad autotune sample");
    Sample s{data, queries, opts};
    const uint32_t item_set[] = {2, 4, 8};
    const uint32_t pool_set[] = {4096, 7680, 14336};
    const uint32_t wgs_set[] = {4, 8};
    const bool verbose = std::getenv("GFSM_TUNE_VERBOSE") != nullptr;

    if (which & GFSM_TUNE_PLAIN) {
        const Tuning keep = ctx.tune;
        uint32_t bi = keep.items[0][1], bp = keep.pool_words, bw = keep.wgs_per_cu;
        double warm, best;
        Status st = time_plain(ctx, s, warm);  // first-launch costs out of the comparison
        if (st.good()) st = time_plain(ctx, s, best);  // the shape in use
        if (!st.good()) return st;
        for (uint32_t it : item_set)
            for (uint32_t pw : pool_set)
                for (uint32_t wg : wgs_set) {
                    if (expand_lds_bytes(it, pw) > kMaxDynLds) continue;
                    for (auto &k : ctx.tune.items)
                        for (auto &v : k) v = (uint8_t)it;
                    ctx.tune.pool_words = pw;
                    ctx.tune.wgs_per_cu = wg;
                    double ms;
                    st = time_plain(ctx, s, ms);
                    if (!st.good()) {
                        ctx.tune = keep;
                        return st;
                    }
                    if (verbose) std::fprintf(stderr, "autotune plain items=%u pool=%u wgs=%u: %.3f ms\n", it, pw, wg, ms);
                    if (ms < best * (1 - kMargin)) {
                        best = ms;
                        bi = it;
                        bp = pw;
                        bw = wg;
                    }
                }
        for (auto &k : ctx.tune.items)
            for (auto &v : k) v = (uint8_t)bi;
        ctx.tune.pool_words = bp;
        ctx.tune.wgs_per_cu = bw;
    }
    if (which & GFSM_TUNE_BATCH) {
        BatchShape base;
        base.slots = ctx.tune.batch_slots;
        base.steps_per_replay = ctx.tune.batch_steps;
        base.step_budget = ctx.tune.batch_step_budget;
        BatchShape pick = base;
        pick.items = ctx.tune.batch_items;
        pick.pool_words = ctx.tune.batch_pool_words;
        pick.wgs_per_cu = ctx.tune.batch_wgs_per_cu;
        double warm, best;
        Status st = time_batch(ctx, s, pick, warm);
        if (st.good()) st = time_batch(ctx, s, pick, best);  // the shape in use
        if (!st.good()) return st;
        for (uint32_t it : item_set)
            for (uint32_t pw : pool_set)
                for (uint32_t wg : wgs_set) {
                    if (expand_lds_bytes(it, pw) > kMaxDynLds) continue;
                    BatchShape sh = base;
                    sh.items = it;
                    sh.pool_words = pw;
                    sh.wgs_per_cu = wg;
                    double ms;
                    st = time_batch(ctx, s, sh, ms);
                    if (!st.good()) return st;
                    if (verbose) std::fprintf(stderr, "autotune batch items=%u pool=%u wgs=%u: %.3f ms\n", it, pw, wg, ms);
                    if (ms < best * (1 - kMargin)) {
                        best = ms;
                        pick = sh;
                    }
                }
        ctx.tune.batch_items = pick.items;
        ctx.tune.batch_pool_words = pick.pool_words;
        ctx.tune.batch_wgs_per_cu = pick.wgs_per_cu;
    }
    return Status::ok();
}

}  // namespace gfsm
