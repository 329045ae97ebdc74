// engine.h -- device contexts, resident graphs and the plain matcher driver.
#pragma once

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cstdio>
#include <memory>
#include <mutex>
#include <set>
#include <unordered_map>

#include "host.h"
#include "kernels.h"

namespace gfsm {

struct DevGraph {
    int device = 0;
    const void *ctx = nullptr;
    uint32_t *nbr = nullptr, *dir_ptr = nullptr, *dir = nullptr, *rel_vtx = nullptr, *rel_beg = nullptr,
             *rel_cnt = nullptr;
    uint4 *seg = nullptr;  // segment hash table (see DevGraphView)
    DevGraphView view{};
    size_t bytes = 0;
    // Dense per-relation segment tables ({begin, end} for every vertex), built
    // on first use by the single-query path and kept while the mirror lives.
    std::mutex dense_mu;
    std::unordered_map<uint32_t, uint2 *> dense;
    uint64_t dense_bytes = 0;
};

// Dense segment table of relation `rel` on a resident graph: cached on the
// mirror; built (on stream `s`, waited for) when a step makes at least V/8
// lookups of it, within kDenseMaxBytes per mirror.  Null: use the hash table.
const uint2 *dense_segments(const HostGraph &data, DevGraph &dg, uint32_t rel, uint64_t lookups, hipStream_t s);
constexpr uint64_t kDenseMaxBytes = 4ull << 30;
// The table of `rel` if one was built already, else null (never builds).
const uint2 *dense_segments_cached(DevGraph &dg, uint32_t rel);

// A stream with its own chained-scan workspace and result block.  Kernels of
// one lane run in stream order, so the workspace is never shared by two
// launches at once.
struct Lane {
    hipStream_t stream = nullptr;
    ScanWS ws{};
    Totals *tot_d = nullptr;
    Totals *tot_h = nullptr;  // pinned
    // Fine-grained host words {keep, cnt} that root and inner-depth steps
    // write beside their device totals: the depth loop reads a step's totals
    // after a stream wait, with no copy launch.
    uint64_t *kc_h = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> kev;  // begin/end pairs of timed launches (grown on demand)
    uint64_t launches = 0;  // chained-scan launches issued on this lane
    uint64_t *tile_node = nullptr;  // [2*max_tiles] per-tile frontier nodes (single-query lane only)
    Status init(uint64_t max_tiles, bool tile_index = false);
    // Call before every chained-scan launch: clears the tile words when the
    // 16-bit launch epoch is about to repeat.
    hipError_t before_launch();
    void destroy();
};

// Device arena for intermediate levels and output rows.
struct Arena {
    char *base = nullptr;
    uint64_t bytes = 0;
};

// Tile shapes: candidates per thread (1, 2, 4, 8) per kernel kind and depth
// for the single-query path, the LDS staging pool, workgroups per CU of the
// persistent grids, and the batch step kernel's shape.  Filled with
// defaults, replaced by the autotuner.
struct Tuning {
    uint8_t items[3][GFSM_MAXQ];
    uint32_t pool_words = kPoolWords;
    uint32_t wgs_per_cu = 4;
    uint32_t batch_items = 8;
    uint32_t batch_pool_words = kPoolWords;  // four workgroups per CU, as the step kernel's registers allow
    uint32_t batch_wgs_per_cu = 4;
    uint32_t batch_slots = 256;
    uint32_t batch_steps = 4;  // steps per captured replay of gfsm_match_batch
    uint32_t serve_steps = 2;  // ... of the server (short: a finished slot is refilled at the end of its replay)
    uint64_t batch_step_budget = 1ull << 26;
    Tuning();
};

struct Ctx {
    int device = 0;
    uint64_t budget = 0;
    uint32_t num_cus = 0;
    uint64_t max_tiles = 0;
    Lane lane;
    Arena arena;
    Arena batch_arena;  // level regions of the batched path (grow-only, shared by its engines)
    Tuning tune;
    char *scratch = nullptr;  // this device's allocation-free match array (kernels.hip), not owned
    std::mutex mu;
    std::set<const HostGraph *> graphs;  // data graphs resident on this context
    // Batched-matching engine kept between calls (its arena, slots and
    // captured steps); owned by batch.cpp, dropped with its graph.
    void *batch_cache = nullptr;
    const HostGraph *batch_graph = nullptr;
    void (*batch_cache_free)(void *) = nullptr;
    // Servers running on this context: closing the context stops each of
    // them first (it finishes the work already submitted), so no server
    // outlives the device memory it uses.  Owned by serve.cpp.
    std::mutex srv_mu;
    std::set<void *> servers;
    void (*halt_server)(void *) = nullptr;
    void drop_batch_cache()
    {
        if (batch_cache && batch_cache_free) batch_cache_free(batch_cache);
        batch_cache = nullptr;
        batch_graph = nullptr;
    }
};

Status ctx_open(int device, uint64_t budget, Ctx *&out);
void ctx_close(Ctx *ctx);
Status ensure_arena(Ctx &ctx, uint64_t bytes);
Status ensure_arena(Arena &a, uint64_t bytes);
Status upload_graph(Ctx &ctx, const HostGraph &g, DevGraph *&out);
void release_graph_mirrors(const HostGraph &g);
// Server pins on a graph (see HostGraph::pins).  free_graph_when_unpinned
// frees the graph now, or marks it and leaves it to the last unpin.
void pin_graph(const HostGraph &g);
void unpin_graph(const HostGraph &g);
void free_graph_when_unpinned(HostGraph *g);

// Plain single-query path: host-driven depth loop with chunked expansion.
// With `stream`, rows mode writes each chunk's rows to it (after an output.fs
// header the caller wrote) instead of keeping them; unsorted only.
Status match_plain(Ctx &ctx, const HostGraph &data, const HostGraph &query, const gfsm_match_opts &opts,
                   Result &res, FILE *stream = nullptr);

// gfsm_match in count mode without allocating: the data graph must already
// be resident on ctx; levels, scan workspace and totals live in the device's
// scratch array (kNoAllocScratchBytes); only dense tables built earlier are
// used.  Calls on one device are serialised (noalloc_mutex).
Status match_noalloc(Ctx &ctx, const HostGraph &data, const HostGraph &query, const gfsm_match_opts &opts,
                     Result &res);
std::mutex &noalloc_mutex(int device);

// The query's plan for these options: the reference order, or with
// GFSM_ORDER_COST the data-statistics order (root windows then rejected).
Status plan_for(const HostGraph &data, const HostGraph &query, const gfsm_match_opts &o, Plan &plan);

// Matcher helpers shared with the batched and serving paths.
void root_range(const HostGraph &data, const Plan &plan, uint32_t root_lo, uint32_t root_hi, uint64_t &lo,
                uint64_t &hi);
Status selftest_scan(Ctx &ctx, const uint32_t *vals, uint64_t n, uint64_t *prefix, uint32_t *idx,
                     uint64_t totals[2]);
uint64_t alg_bytes(const Plan &plan, const gfsm_stats &st, uint32_t flags);
// Runs `launches` calibration launches (launch_calibrate) over n words;
// bytes = {HBM bytes read, written} per launch by construction.
Status selftest_traffic(Ctx &ctx, uint32_t mode, uint64_t n, uint32_t launches, uint64_t bytes[2]);

inline Status hip_status(hipError_t e, const char *what)
{
    if (e == hipSuccess) return Status::ok();
    return Status::err(GFSM_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace gfsm

struct gfsm_ctx {
    gfsm::Ctx *c;
};
