// engine.cpp -- contexts, resident graphs and the plain (single-query)
// matcher driver.
//
// Replaces HnSetup::preinit / gpuLaunch and the device-side orchestrator
// hnSolve (HnGPU.cu:847-1063, 455-669).  The reference drives every depth
// from one device thread with dynamic parallelism and, when a level would
// overflow its fixed buffers, either batches on the device (HnGPU.cu:529-630)
// or frees everything and reruns at twice the memory (HnGPU.cu:957-992).
// Here the host drives the depths, and a level that does not fit its buffer
// is expanded in candidate-range chunks, depth first: each chunk's subtree
// runs to the leaves before the next chunk is generated, so memory is bounded
// by the budget whatever the output size, and the result does not depend on
// how the work was chunked.
#include "engine.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>

namespace gfsm {

static std::mutex g_dev_mu;  // guards HostGraph::dev and Ctx::graphs

static double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

#define HIP_TRY(expr)                                                    \
    do {                                                                 \
        hipError_t e_ = (expr);                                          \
        if (e_ != hipSuccess) return hip_status(e_, #expr);              \
    } while (0)

Status Lane::init(uint64_t max_tiles, bool tile_index)
{
    if (tile_index) HIP_TRY(hipMalloc(&tile_node, 2 * max_tiles * sizeof(uint64_t)));
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    ws.max_tiles = max_tiles;
    HIP_TRY(hipMalloc(&ws.wk, max_tiles * sizeof(uint64_t)));
    HIP_TRY(hipMalloc(&ws.wc, max_tiles * sizeof(uint64_t)));
    HIP_TRY(hipMalloc(&ws.ctl, 4 * sizeof(uint32_t)));
    HIP_TRY(hipMemset(ws.wk, 0, max_tiles * sizeof(uint64_t)));
    HIP_TRY(hipMemset(ws.wc, 0, max_tiles * sizeof(uint64_t)));
    HIP_TRY(hipMemset(ws.ctl, 0, 4 * sizeof(uint32_t)));
    HIP_TRY(hipMalloc(&tot_d, sizeof(Totals)));
    HIP_TRY(hipMemset(tot_d, 0, sizeof(Totals)));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&tot_h), sizeof(Totals), hipHostMallocDefault));
    std::memset(tot_h, 0, sizeof(Totals));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&kc_h), 2 * sizeof(uint64_t), hipHostMallocCoherent));
    kc_h[0] = kc_h[1] = 0;
    for (int i = 0; i < 2; i++) HIP_TRY(hipEventCreate(&ev[i]));
    return Status::ok();
}

hipError_t Lane::before_launch()
{
    // Launch L runs with epoch L mod 2^16 (the device counter advances once per
    // launch); a word left by launch L - 2^16 would carry the same epoch.
    hipError_t e = hipSuccess;
    if (launches && (launches % (1ull << kEpochBits)) == 0) e = launch_ws_clear(ws, stream);
    launches++;
    return e;
}

void Lane::destroy()
{
    if (stream) (void)hipStreamSynchronize(stream);
    (void)hipFree(ws.wk);
    (void)hipFree(ws.wc);
    (void)hipFree(ws.ctl);
    (void)hipFree(tile_node);
    tile_node = nullptr;
    (void)hipFree(tot_d);
    if (tot_h) (void)hipHostFree(tot_h);
    if (kc_h) (void)hipHostFree(kc_h);
    kc_h = nullptr;
    for (int i = 0; i < 2; i++)
        if (ev[i]) (void)hipEventDestroy(ev[i]);
    for (hipEvent_t e : kev) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
    *this = Lane{};
}

Tuning::Tuning()
{
    for (auto &k : items)
        for (auto &v : k) v = kItems;
}

Status ctx_open(int device, uint64_t budget, Ctx *&out)
{
    out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return Status::err(GFSM_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return Status::err(GFSM_ERR_ARG, "device index out of range");
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    std::unique_ptr<Ctx> c(new Ctx);
    c->device = device;
    c->budget = budget ? budget : (16ull << 30);
    c->num_cus = (uint32_t)prop.multiProcessorCount;
    c->max_tiles = 1ull << 22;
    Status st = c->lane.init(c->max_tiles, true);
    if (!st.good()) {
        c->lane.destroy();
        return st;
    }
    // The allocation-free match's device array: its address on this device
    // (looked up once per device; the array lives until the process ends).
    hipError_t se = noalloc_scratch(device, c->lane.stream, &c->scratch);
    if (se != hipSuccess) {
        c->lane.destroy();
        return hip_status(se, "device scratch of the allocation-free match");
    }
    out = c.release();
    return Status::ok();
}

static void free_mirror(DevGraph *dg)
{
    (void)hipSetDevice(dg->device);
    for (auto &kv : dg->dense) (void)hipFree(kv.second);
    (void)hipFree(dg->nbr);
    (void)hipFree(dg->dir_ptr);
    (void)hipFree(dg->dir);
    (void)hipFree(dg->rel_vtx);
    (void)hipFree(dg->rel_cnt);
    (void)hipFree(dg->rel_beg);
    (void)hipFree(dg->seg);
    delete dg;
}

void ctx_close(Ctx *ctx)
{
    if (!ctx) return;
    {
        std::vector<void *> running;
        {
            std::lock_guard<std::mutex> lk(ctx->srv_mu);
            running.assign(ctx->servers.begin(), ctx->servers.end());
        }
        for (void *s : running) ctx->halt_server(s);
    }
    (void)hipSetDevice(ctx->device);
    ctx->drop_batch_cache();
    {
        std::lock_guard<std::mutex> lk(g_dev_mu);
        for (const HostGraph *g : ctx->graphs) {
            auto it = g->dev.find(ctx);
            if (it != g->dev.end()) {
                free_mirror(it->second);
                g->dev.erase(it);
            }
        }
        ctx->graphs.clear();
    }
    ctx->lane.destroy();
    if (ctx->arena.base) (void)hipFree(ctx->arena.base);
    if (ctx->batch_arena.base) (void)hipFree(ctx->batch_arena.base);
    delete ctx;
}

void release_graph_mirrors(const HostGraph &g)
{
    std::lock_guard<std::mutex> lk(g_dev_mu);
    for (auto &kv : g.dev) {
        Ctx *c = (Ctx *)const_cast<void *>(kv.first);
        if (c->batch_graph == &g) c->drop_batch_cache();
        c->graphs.erase(&g);
        free_mirror(kv.second);
    }
    g.dev.clear();
}

std::mutex &noalloc_mutex(int device)
{
    static std::mutex mu[64];
    return mu[(unsigned)device % 64];
}

static std::mutex g_pin_mu;  // guards HostGraph::pins and free_pending

void pin_graph(const HostGraph &g)
{
    std::lock_guard<std::mutex> lk(g_pin_mu);
    g.pins++;
}

void unpin_graph(const HostGraph &g)
{
    bool free_now = false;
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        if (g.pins) g.pins--;
        free_now = g.pins == 0 && g.free_pending;
    }
    if (free_now) {
        release_graph_mirrors(g);
        delete &g;
    }
}

void free_graph_when_unpinned(HostGraph *g)
{
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        if (g->pins) {
            g->free_pending = true;
            return;
        }
    }
    release_graph_mirrors(*g);
    delete g;
}

Status ensure_arena(Arena &a, uint64_t bytes)
{
    if (a.bytes >= bytes) return Status::ok();
    if (a.base) {
        (void)hipFree(a.base);
        a = Arena{};
    }
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&a.base), bytes);
    if (e != hipSuccess) {
        a = Arena{};
        return Status::err(GFSM_ERR_NOMEM, "cannot allocate " + std::to_string(bytes) + " device bytes");
    }
    a.bytes = bytes;
    return Status::ok();
}

Status ensure_arena(Ctx &ctx, uint64_t bytes) { return ensure_arena(ctx.arena, bytes); }

template <class T>
static Status dev_copy(T *&dst, const std::vector<T> &src, size_t &bytes)
{
    size_t b = std::max<size_t>(src.size(), 1) * sizeof(T);
    HIP_TRY(hipMalloc(&dst, b));
    if (!src.empty()) HIP_TRY(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    bytes += b;
    return Status::ok();
}

Status upload_graph(Ctx &ctx, const HostGraph &g, DevGraph *&out)
{
    std::lock_guard<std::mutex> lk(g_dev_mu);
    auto it = g.dev.find(&ctx);
    if (it != g.dev.end()) {
        out = it->second;
        return Status::ok();
    }
    HIP_TRY(hipSetDevice(ctx.device));
    DevGraph *dg = new DevGraph;
    dg->device = ctx.device;
    dg->ctx = &ctx;
    std::vector<uint32_t> dir_flat(g.dir.size() * 2);
    for (size_t i = 0; i < g.dir.size(); i++) {
        dir_flat[2 * i] = g.dir[i].rel;
        dir_flat[2 * i + 1] = g.dir[i].beg;
    }
    Status st = dev_copy(dg->nbr, g.nbr, dg->bytes);
    if (st.good()) st = dev_copy(dg->dir_ptr, g.dir_ptr, dg->bytes);
    if (st.good()) st = dev_copy(dg->dir, dir_flat, dg->bytes);
    if (st.good()) st = dev_copy(dg->rel_vtx, g.rel_vtx, dg->bytes);
    if (st.good()) st = dev_copy(dg->rel_cnt, g.rel_cnt, dg->bytes);
    if (st.good()) st = dev_copy(dg->rel_beg, g.rel_beg, dg->bytes);
    if (!st.good()) {
        free_mirror(dg);
        return st;
    }
    // Segment hash table: a power of two at least twice the directory size.
    uint64_t slots = 1024;
    while (slots < 2 * (uint64_t)g.ndir()) slots <<= 1;
    if (slots > (1ull << 32)) {
        free_mirror(dg);
        return Status::err(GFSM_ERR_ARG, "graph has too many (vertex, relation) segments");
    }
    hipError_t he = hipMalloc(&dg->seg, slots * sizeof(uint4));
    if (he == hipSuccess) he = launch_seg_hash_build(dg->dir_ptr, dg->dir, g.V, dg->seg, (uint32_t)(slots - 1), 0);
    if (he == hipSuccess) he = hipDeviceSynchronize();
    if (he != hipSuccess) {
        free_mirror(dg);
        return hip_status(he, "segment hash table");
    }
    dg->bytes += slots * sizeof(uint4);
    dg->view = DevGraphView{dg->nbr, dg->dir_ptr, dg->dir, dg->seg, (uint32_t)(slots - 1), g.V, (uint32_t)g.nent(), 0};
    g.dev[&ctx] = dg;
    ctx.graphs.insert(&g);
    out = dg;
    return Status::ok();
}

Status plan_for(const HostGraph &data, const HostGraph &query, const gfsm_match_opts &o, Plan &plan)
{
    if (!(o.flags & GFSM_ORDER_COST)) return plan_query(query, plan);
    if (o.root_lo || o.root_hi)
        return Status::err(GFSM_ERR_ARG, "root windows refer to the reference order's root; not with GFSM_ORDER_COST");
    return plan_query(query, plan, &data);
}

const uint2 *dense_segments_cached(DevGraph &dg, uint32_t rel)
{
    std::lock_guard<std::mutex> lk(dg.dense_mu);
    auto it = dg.dense.find(rel);
    return it != dg.dense.end() ? it->second : nullptr;
}

const uint2 *dense_segments(const HostGraph &data, DevGraph &dg, uint32_t rel, uint64_t lookups, hipStream_t s)
{
    std::lock_guard<std::mutex> lk(dg.dense_mu);
    auto it = dg.dense.find(rel);
    if (it != dg.dense.end()) return it->second;
    const uint64_t bytes = (uint64_t)data.V * sizeof(uint2);
    if (lookups < data.V / 8 || dg.dense_bytes + bytes > kDenseMaxBytes) return nullptr;
    uint2 *t = nullptr;
    if (hipMalloc(&t, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    uint64_t lo = 0, hi = 0;
    const int64_t k = data.rel_index(rel);
    if (k >= 0) {
        lo = data.rel_off[k];
        hi = data.rel_off[k + 1];
    }
    // Built to completion before it is shared: other streams may read it next.
    if (launch_dense_seg(dg.rel_vtx + lo, dg.rel_beg + lo, dg.rel_cnt + lo, hi - lo, data.V, t, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(t);
        return nullptr;
    }
    dg.dense[rel] = t;
    dg.dense_bytes += bytes;
    return t;
}

void root_range(const HostGraph &data, const Plan &plan, uint32_t root_lo, uint32_t root_hi, uint64_t &lo,
                uint64_t &hi)
{
    lo = hi = 0;
    int64_t k = data.rel_index(plan.dev.root_rel);
    if (k < 0) return;
    lo = data.rel_off[k];
    hi = data.rel_off[k + 1];
    if (root_hi == 0 && root_lo == 0) return;
    if (root_hi == 0) root_hi = 0xFFFFFFFFu;
    const uint32_t *b = data.rel_vtx.data();
    lo = std::lower_bound(b + lo, b + hi, root_lo) - b;
    hi = std::lower_bound(b + lo, b + hi, root_hi) - b;
}

uint64_t alg_bytes(const Plan &plan, const gfsm_stats &st, uint32_t flags)
{
    const uint32_t q = plan.q;
    uint64_t b = 4ull * st.roots;
    for (uint32_t k = 1; k < q; k++) {
        b += 8ull * st.level[k - 1];
        b += 4ull * st.cand[k] * (1ull + plan.dev.depth[k].nchk);
        if (k + 1 < q) b += 8ull * st.level[k];
    }
    if (flags & (GFSM_ROWS | GFSM_SORTED)) b += 4ull * q * st.level[q - 1];
    return b;
}

namespace {

// First chunk of an inner depth of a capped match = cap / kCapStartDiv
// candidates, then the observed fan-out sizes the next: a large first chunk
// tests inner candidates whose children the cap never reaches (measured on
// the power-law hub and middle windows, cap 2^25: 1/8 -> 1/64 cut their time
// by 30% and 15%; 1/128 gained nothing more).
constexpr uint64_t kCapStartDiv = 64;


// Steps of at least this many tiles get their frontier nodes from k_tile_index.
constexpr uint64_t kTileIndexMin = 64;

// Candidates per thread of the lean last-depth kernel (its registers fit 64):
// 8 where the tile's node rows are expected to fit its staging pool, else 4.
constexpr uint32_t kBareItems = 8;
constexpr uint32_t kBareItemsNarrow = 4;

// Width of the lean kernel for a last depth of `cands` candidates over
// `nodes` frontier nodes of `d` placed vertices: 8 candidates per thread when
// an 8-wide tile's nodes (2048 / mean fan-out + 1) keep their rows within the
// pool, else 4, whose tiles cover half as many nodes (measured: 8 is 3% faster
// on the power-law sample; on the road-like grid, fan-out ~4, the rows no
// longer fit at 8 and it is 4% slower).
static uint32_t bare_items(uint64_t cands, uint64_t nodes, uint32_t d)
{
    const double fan = nodes ? (double)cands / (double)nodes : 1.0;
    const double tile_nodes = (double)(kTile * kBareItems) / std::max(fan, 1.0) + 1.0;
    return tile_nodes * d <= (double)kBarePoolWords ? kBareItems : kBareItemsNarrow;
}

// One plain match in flight: level buffers carved from the arena, the lane,
// and the depth-first chunk recursion.  Level k (nodes with k+1 placed
// vertices) holds rows of width k+1, candidate offsets and generator begins.
struct Runner {
    Ctx &ctx;
    const HostGraph &data;
    const DevGraph &dg;
    const Plan &plan;
    uint32_t flags;
    uint64_t max_rows;
    Result &res;
    uint32_t q = 0;
    bool rows_mode = false, digest = false, prof = false;
    uint32_t *vec[GFSM_MAXQ] = {}, *gbeg[GFSM_MAXQ] = {};
    uint64_t *off[GFSM_MAXQ] = {};
    uint64_t cap[GFSM_MAXQ] = {};
    uint64_t chunk[GFSM_MAXQ] = {};
    uint32_t *rows_d = nullptr;
    uint64_t rows_cap = 0;
    uint32_t grid_cap = 0;
    uint64_t max_cand = 0;      // work cap (0 = none)
    uint64_t cand_total = 0;    // candidates launched so far
    uint64_t max_depth_cand = 0;           // deterministic per-depth cap (0 = none)
    uint64_t depth_done[GFSM_MAXQ] = {};   // candidates tested so far per depth (roots: depth 0)
    uint32_t forced_items = 0;  // from gfsm_match_opts.tile (0 = tuned)
    // Rows streamed to an output.fs as each chunk lands (unsorted rows mode):
    // host memory stays one chunk, whatever the result size.
    FILE *stream = nullptr;
    uint64_t streamed = 0, sdg[2] = {0, 0};
    std::vector<uint32_t> stage;

    // The lane (stream, scan workspace, totals); optionally a fixed level
    // arena, with which nothing is allocated (dense tables only if cached).
    Lane &lane;
    const Arena *fixed = nullptr;
    bool no_alloc = false;

    Runner(Ctx &c, const HostGraph &d, const DevGraph &g, const Plan &p, uint32_t f, uint64_t mr, Result &r,
           Lane &l)
        : ctx(c), data(d), dg(g), plan(p), flags(f), max_rows(mr), res(r), lane(l) {}

    // Dense segment table of `rel` for a step of `lookups` lookups: built on
    // demand, or, without allocation, only one an earlier match built.
    const uint2 *dense(uint32_t rel, uint64_t lookups)
    {
        DevGraph &g = const_cast<DevGraph &>(dg);
        return no_alloc ? dense_segments_cached(g, rel) : dense_segments(data, g, rel, lookups, lane.stream);
    }

    // Does the expand of depth d stage the next generator's segment per node?
    // (Only when the next generator is an older ancestor, not the new vertex.)
    uint32_t need_nxt(uint32_t kind, uint32_t d) const
    {
        return (kind == EXPAND_MID && plan.dev.depth[d].nxt_dist > 1) ? 1u : 0u;
    }

    // Candidates per thread for one expand launch: forced, else tuned.
    uint32_t pick_items(uint32_t kind, uint32_t d) const
    {
        return forced_items ? forced_items : ctx.tune.items[kind][d];
    }

    // Candidates per launch: the workspace holds max_tiles tiles, and the
    // launch's next-depth candidate total (at most candidates x max degree)
    // must fit the 46-bit value field of the scan words.
    uint64_t launch_limit() const
    {
        const uint64_t by_ws = lane.ws.max_tiles * kTile;  // enough tiles at any tile size
        const uint64_t by_val = kValueMask / std::max<uint64_t>(1, data.max_degree);
        return std::max<uint64_t>(kTile * 8, std::min(by_ws, by_val));
    }

    uint32_t grid_for(uint64_t ntiles) const
    {
        return (uint32_t)std::min<uint64_t>(ntiles, grid_cap);
    }

    static uint64_t round256(uint64_t b) { return (b + 255) / 256 * 256; }

    Status carve(uint64_t budget)
    {
        q = plan.q;
        rows_mode = flags & (GFSM_ROWS | GFSM_SORTED);
        digest = (flags & GFSM_DIGEST) && !rows_mode;  // rows mode: digest from the host rows
        prof = flags & GFSM_PROFILE;
        grid_cap = std::max<uint32_t>(ctx.num_cus, 1) * ctx.tune.wgs_per_cu;
        const uint64_t min_nodes = 8 * kTile;  // a level must hold one launch tile's survivors
        uint64_t row_bytes = 0;
        if (rows_mode) {
            rows_cap = std::max<uint64_t>(budget / 4 / (4ull * q), min_nodes);
            if (max_rows) rows_cap = std::min<uint64_t>(rows_cap, std::max<uint64_t>(max_rows, min_nodes));
            row_bytes = round256(rows_cap * q * 4);
        }
        if (budget <= row_bytes) return Status::err(GFSM_ERR_NOMEM, "memory budget too small");
        // Equal byte share per level; a node of level k costs 4(k+1) + 12 bytes.
        const uint64_t per_level = (budget - row_bytes) / (q - 1);
        uint64_t total = row_bytes;
        for (uint32_t k = 0; k + 1 < q; k++) {
            const uint64_t node_bytes = 4ull * (k + 1) + 12;
            const uint64_t c = per_level / node_bytes > 256 ? per_level / node_bytes - 256 : 0;
            if (c < min_nodes) return Status::err(GFSM_ERR_NOMEM, "memory budget too small for the query depth");
            cap[k] = c;
            total += round256(c * 4 * (k + 1)) + round256((c + 1) * 8) + round256(c * 4);
            chunk[k + 1] = std::min<uint64_t>(c, launch_limit());
        }
        chunk[q - 1] = launch_limit();
        char *p = nullptr;
        if (fixed) {
            if (total > fixed->bytes)
                return Status::err(GFSM_ERR_NOMEM, "query needs " + std::to_string(total) +
                                                       " level bytes; the fixed arena holds " +
                                                       std::to_string(fixed->bytes));
            p = fixed->base;
        } else {
            Status st = ensure_arena(ctx, total);
            if (!st.good()) return st;
            p = ctx.arena.base;
        }
        auto take = [&p](uint64_t b) {
            char *r = p;
            p += round256(b);
            return r;
        };
        if (rows_mode) rows_d = reinterpret_cast<uint32_t *>(take(rows_cap * q * 4));
        for (uint32_t k = 0; k + 1 < q; k++) {
            vec[k] = reinterpret_cast<uint32_t *>(take(cap[k] * 4 * (k + 1)));
            off[k] = reinterpret_cast<uint64_t *>(take((cap[k] + 1) * 8));
            gbeg[k] = reinterpret_cast<uint32_t *>(take(cap[k] * 4));
        }
        return Status::ok();
    }

    // Under a per-depth cap: the next chunk of depth d sized so its children
    // roughly fill depth d+1's remaining budget, from this chunk's fan-out
    // (cnt children of n candidates).  Only the amount of wasted work
    // changes; the capped result does not depend on chunk sizes.
    uint64_t capped_chunk(uint32_t d, uint64_t n, uint64_t cnt, uint64_t chunk_now, uint64_t tmin) const
    {
        if (d + 1 >= q) return chunk_now;
        const uint64_t used = depth_done[d + 1];
        const uint64_t rem = max_depth_cand > used ? max_depth_cand - used : 0;
        const uint64_t after = rem > cnt ? rem - cnt : 0;
        const uint64_t need = (uint64_t)((double)after * (double)n / (double)std::max<uint64_t>(cnt, 1)) + 1;
        return std::max<uint64_t>(tmin, std::min<uint64_t>(chunk_now, need));
    }

    Status sync_totals()
    {
        HIP_TRY(hipMemcpyAsync(lane.tot_h, lane.tot_d, sizeof(Totals), hipMemcpyDeviceToHost,
                               lane.stream));
        HIP_TRY(hipStreamSynchronize(lane.stream));
        return Status::ok();
    }

    // Host words of the next root / inner step's {keep, cnt}, armed with a
    // value no step writes; null when the lane has none (then the totals are
    // copied).
    uint64_t *arm_step_totals()
    {
        if (!lane.kc_h) return nullptr;
        lane.kc_h[0] = lane.kc_h[1] = ~0ull;
        return lane.kc_h;
    }

    // After a root or non-counting step: take its {keep, cnt} into tot_h,
    // from the host words it wrote or by a copy.  With `early`, the host
    // takes them as soon as the step's last tile has published them, while
    // the step's other workgroups may still be finishing: only for steps whose
    // next use is a launch on the same stream (stream order makes it wait).
    // The stream is queried now and then, so a step that ends without
    // publishing is an error, not a hang.
    Status sync_step_totals(const uint64_t *kc, bool early = false)
    {
        if (!kc) return sync_totals();
        const volatile uint64_t *w = reinterpret_cast<const volatile uint64_t *>(kc);
        if (early) {
            for (uint32_t i = 1; w[0] == ~0ull; i++) {
                if ((i & 255) == 0) {
                    const hipError_t q = hipStreamQuery(lane.stream);
                    if (q == hipSuccess) break;  // ended: the words are final
                    if (q != hipErrorNotReady) return Status::err(GFSM_ERR_DEVICE, hipGetErrorString(q));
                }
                __builtin_ia32_pause();
            }
        } else {
            HIP_TRY(hipStreamSynchronize(lane.stream));
        }
        const uint64_t keep = w[0];
        const uint64_t cnt = w[1];
        if (keep == ~0ull || cnt == ~0ull) return Status::err(GFSM_ERR_DEVICE, "internal: step totals not written");
        lane.tot_h->keep = keep;
        lane.tot_h->cnt = cnt;
        return Status::ok();
    }

    // Per-launch timing (GFSM_PROFILE): an event pair around each matcher
    // launch, read back in batches, so timing adds no host wait between
    // launches beyond those the depth loop makes anyway.
    static constexpr uint32_t kProfPairs = 512;
    uint32_t kn = 0;                     // pairs recorded and not yet read
    bool kexp[kProfPairs] = {};          // pair i timed an expand launch
    Status kbegin()
    {
        if (!prof) return Status::ok();
        std::vector<hipEvent_t> &ev = lane.kev;
        while (ev.size() < 2ull * (kn + 1)) {
            hipEvent_t e = nullptr;
            HIP_TRY(hipEventCreate(&e));
            ev.push_back(e);
        }
        HIP_TRY(hipEventRecord(ev[2 * kn], lane.stream));
        return Status::ok();
    }
    Status kend(bool expand)
    {
        if (!prof) return Status::ok();
        HIP_TRY(hipEventRecord(lane.kev[2 * kn + 1], lane.stream));
        kexp[kn++] = expand;
        return kn == kProfPairs ? prof_flush() : Status::ok();
    }
    Status prof_flush()
    {
        if (kn == 0) return Status::ok();
        const std::vector<hipEvent_t> &ev = lane.kev;
        HIP_TRY(hipEventSynchronize(ev[2 * kn - 1]));
        for (uint32_t i = 0; i < kn; i++) {
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
            res.stats.kernel_ms += ms;
            if (kexp[i]) {
                res.stats.expand_ms += ms;
                res.stats.expand_launches++;
            }
        }
        kn = 0;
        return Status::ok();
    }

    Status run_roots(uint64_t rlo, uint64_t rhi)
    {
        res.stats.roots = rhi - rlo;
        uint64_t step = std::min<uint64_t>(cap[0], launch_limit());
        if (max_depth_cand) step = std::max<uint64_t>(kRootTile, std::min<uint64_t>(step, max_depth_cand / kCapStartDiv));
        for (uint64_t c = rlo; c < rhi;) {
            uint64_t e = std::min(rhi, c + step);
            if (max_depth_cand) {
                if (depth_done[0] >= max_depth_cand) return Status::err(GFSM_ERR_CAP, "per-depth candidate cap reached");
                e = std::min(e, c + (max_depth_cand - depth_done[0]));
            }
            RootArgs a{};
            a.g = dg.view;
            a.io.cand_vtx = dg.rel_vtx + c;
            a.io.cand_beg = dg.rel_beg + c;
            a.io.cand_cnt = dg.rel_cnt + c;
            a.io.n = e - c;
            a.io.out_vec = vec[0];
            a.io.out_off = off[0];
            a.io.out_gbeg = gbeg[0];
            a.io.cap = cap[0];
            a.io.tot = lane.tot_d;
            a.io.host_kc = arm_step_totals();
            a.io.ntiles = (a.io.n + kRootTile - 1) / kRootTile;
            a.nchk = plan.dev.nroot_chk;
            std::memcpy(a.chk_rel, plan.dev.root_chk_rel, sizeof(a.chk_rel));
            a.ws = lane.ws;
            if (a.io.ntiles > lane.ws.max_tiles || e > data.rel_vtx.size())
                return Status::err(GFSM_ERR_ARG, "internal: root launch shape check failed");
            HIP_TRY(lane.before_launch());
            {
                Status ks = kbegin();
                if (!ks.good()) return ks;
            }
            HIP_TRY(launch_root(a, grid_for(a.io.ntiles), lane.stream));
            res.stats.launches++;
            Status st = kend(false);
            if (!st.good()) return st;
            st = sync_step_totals(a.io.host_kc, true);
            if (!st.good()) return st;
            const uint64_t S0 = lane.tot_h->keep, G1 = lane.tot_h->cnt;
            res.stats.level[0] += S0;
            res.stats.cand[0] += e - c;  // root candidates tested
            depth_done[0] += e - c;
            if (max_depth_cand) step = capped_chunk(0, e - c, G1, step, kRootTile);
            if (S0) {
                st = descend(1, 0, G1, S0);
                if (!st.good()) return st;
            }
            c = e;
        }
        return Status::ok();
    }

    // Expand depth d over candidates [g0, g1) of level d-1 (S_prev nodes).
    Status descend(uint32_t d, uint64_t g0, uint64_t g1, uint64_t S_prev)
    {
        const bool last = (d + 1 == q);
        // A counting last depth with nothing to check and no digest runs the
        // lean kernel (at most 4 candidates per thread there: it keeps its
        // registers under 64 for eight waves per SIMD).
        const bool bare = last && !rows_mode && !digest && plan.dev.depth[d].nchk == 0 &&
                          !(forced_items > kBareItems);
        const uint32_t kind =
            last ? (rows_mode ? EXPAND_LAST_ROWS : (bare ? EXPAND_LAST_BARE : EXPAND_LAST_COUNT)) : EXPAND_MID;
        const uint64_t tile_limit = launch_limit();
        const uint32_t items =
            bare ? std::min(pick_items(EXPAND_LAST_COUNT, d), bare_items(g1 - g0, S_prev, d)) : pick_items(kind, d);
        const uint32_t pool_words = bare ? std::min(ctx.tune.pool_words, kBarePoolWords) : ctx.tune.pool_words;
        const uint32_t grid_mul = bare ? 2 : 1;  // twice the resident workgroups
        const uint64_t tc = (uint64_t)kTile * items;
        bool first = true;
        for (uint64_t c = g0; c < g1;) {
            uint64_t e = std::min(g1, c + std::max<uint64_t>(tc, std::min(chunk[d], tile_limit)));
            if (max_depth_cand) {
                // Deterministic cap: depth d tests only its first max_depth_cand
                // candidates in enumeration order; when it has none left, the
                // match stops (every deeper range is finished by then).
                if (depth_done[d] >= max_depth_cand) return Status::err(GFSM_ERR_CAP, "per-depth candidate cap reached");
                e = std::min(e, c + (max_depth_cand - depth_done[d]));
            }
            ExpandArgs a{};
            a.items = items;
            a.pool_words = pool_words;
            a.g = dg.view;
            a.dp = plan.dev.depth[d];
            std::memcpy(a.col, plan.dev.col, sizeof(a.col));
            std::memcpy(a.dcol, plan.dev.dcol, sizeof(a.dcol));
            StepIO &io = a.io;
            io.need_nxt = need_nxt(kind, d);
            if (kind == EXPAND_MID)
                a.dp.nxt_seg = dense(plan.dev.depth[d].nxt_rel, e - c);
            for (uint32_t r = 0; r < std::min<uint32_t>(plan.dev.depth[d].nchk, kDenseChk); r++)
                a.dp.chk_seg[r] = dense(plan.dev.depth[d].chk_rel[r], e - c);
            io.d = d;
            io.q = q;
            io.kind = kind;
            io.digest = (last && digest) ? 1u : 0u;
            io.in_vec = vec[d - 1];
            io.in_off = off[d - 1];
            io.in_gbeg = gbeg[d - 1];
            io.S_prev = S_prev;
            io.g_begin = c;
            io.g_end = e;
            if (!last) {
                io.out_vec = vec[d];
                io.out_off = off[d];
                io.out_gbeg = gbeg[d];
                io.cap = cap[d];
            } else if (rows_mode) {
                io.rows = rows_d;
                io.cap = rows_cap;
            }
            io.tot = lane.tot_d;
            io.host_kc = (kind == EXPAND_MID || kind == EXPAND_LAST_ROWS) ? arm_step_totals() : nullptr;
            io.ntiles = (e - c + tc - 1) / tc;
            a.ws = lane.ws;
            // Launch-shape checks: the workspace holds every tile, the frontier
            // is non-empty and the LDS staging fits.
            if (io.ntiles == 0 || io.ntiles > lane.ws.max_tiles || S_prev == 0 || S_prev > cap[d - 1] ||
                a.dp.gen_dist == 0 || a.dp.gen_dist > d || expand_lds_bytes(items, pool_words) > kMaxDynLds)
                return Status::err(GFSM_ERR_ARG, "internal: expand launch shape check failed");
            if (max_cand && cand_total + (e - c) > max_cand)
                return Status::err(GFSM_ERR_CAP, "candidate cap reached");
            cand_total += e - c;
            // The tiles' frontier nodes, searched by the whole GPU ahead of the
            // step; a step of a few tiles searches inside its tiles instead of
            // paying for another launch.
            if (lane.tile_node && io.ntiles >= kTileIndexMin) {
                HIP_TRY(launch_tile_index(io.in_off, S_prev, c, e, tc, io.ntiles, lane.tile_node, lane.stream));
                io.tile_node = lane.tile_node;
            }
            HIP_TRY(lane.before_launch());
            {
                Status ks = kbegin();
                if (!ks.good()) return ks;
            }
            // Counting kinds deal their tiles by XCD when the grid is a
            // multiple of the XCD count (kernels.hip, TileDeal).
            uint64_t grid = std::min<uint64_t>(io.ntiles, (uint64_t)grid_cap * grid_mul);
            if (kind != EXPAND_MID && kind != EXPAND_LAST_ROWS && grid >= 8) grid &= ~7ull;
            HIP_TRY(launch_expand(a, (uint32_t)grid, lane.stream));
            res.stats.launches++;
            if (!first) res.stats.chunks++;
            first = false;
            Status st = kend(true);
            if (!st.good()) return st;
            if (kind == EXPAND_LAST_COUNT || kind == EXPAND_LAST_BARE) {
                res.stats.cand[d] += e - c;
                depth_done[d] += e - c;
                c = e;
                continue;
            }
            st = sync_step_totals(io.host_kc, kind == EXPAND_MID);
            if (!st.good()) return st;
            const uint64_t keep = lane.tot_h->keep;
            if (keep > e - c) return Status::err(GFSM_ERR_DEVICE, "internal: more nodes kept than candidates");
            const uint64_t limit = (kind == EXPAND_MID) ? cap[d] : rows_cap;
            if (keep > limit) {
                // The level did not fit: discard this launch and halve the range.
                res.stats.overflows++;
                chunk[d] = std::max<uint64_t>(tc, (e - c) / 2);
                continue;
            }
            res.stats.cand[d] += e - c;
            res.stats.level[d] += keep;
            depth_done[d] += e - c;
            if (keep) {
                const double est = (double)limit * (double)(e - c) / (double)keep * 0.7;
                chunk[d] = (uint64_t)std::max<double>((double)tc, std::min<double>(est, (double)tile_limit));
                if (max_depth_cand && kind == EXPAND_MID)
                    chunk[d] = capped_chunk(d, e - c, lane.tot_h->cnt, chunk[d], tc);
            } else {
                chunk[d] = std::min<uint64_t>(chunk[d] * 4, tile_limit);
            }
            if (kind == EXPAND_MID) {
                if (keep) {
                    st = descend(d + 1, 0, lane.tot_h->cnt, keep);
                    if (!st.good()) return st;
                }
            } else {
                if (max_rows && res.rows.size() / q + streamed + keep > max_rows)
                    return Status::err(GFSM_ERR_CAP, "more rows than max_rows");
                if (stream) {
                    stage.resize(keep * q);
                    if (keep)
                        HIP_TRY(hipMemcpy(stage.data(), rows_d, keep * q * sizeof(uint32_t), hipMemcpyDeviceToHost));
                    if (flags & GFSM_DIGEST)
                        for (uint64_t r = 0; r < keep; r++) {
                            uint64_t h[2];
                            row_hash(stage.data() + r * q, q, h);
                            sdg[0] += h[0];
                            sdg[1] += h[1];
                        }
                    if (keep && std::fwrite(stage.data(), sizeof(uint32_t), stage.size(), stream) != stage.size())
                        return Status::err(GFSM_ERR_MISSING_FILE, "short write to output.fs");
                    streamed += keep;
                    c = e;
                    continue;
                }
                size_t at = res.rows.size();
                res.rows.resize(at + keep * q);
                if (keep)
                    HIP_TRY(hipMemcpy(res.rows.data() + at, rows_d, keep * q * sizeof(uint32_t),
                                      hipMemcpyDeviceToHost));
            }
            c = e;
        }
        return Status::ok();
    }
};

}  // namespace

Status selftest_scan(Ctx &ctx, const uint32_t *vals, uint64_t n, uint64_t *prefix, uint32_t *idx, uint64_t totals[2])
{
    std::lock_guard<std::mutex> lk(ctx.mu);
    HIP_TRY(hipSetDevice(ctx.device));
    totals[0] = totals[1] = 0;
    if (n == 0) return Status::ok();
    const uint64_t tc = (uint64_t)kTile * kScanTestItems;
    const uint64_t ntiles = (n + tc - 1) / tc;
    if (ntiles > ctx.max_tiles) return Status::err(GFSM_ERR_ARG, "too many values for one launch");
    uint32_t *dv = nullptr, *di = nullptr;
    uint64_t *dp = nullptr;
    Status st = Status::ok();
    auto cleanup = [&] {
        (void)hipFree(dv);
        (void)hipFree(di);
        (void)hipFree(dp);
    };
    hipError_t e = hipMalloc(&dv, n * 4);
    if (e == hipSuccess) e = hipMalloc(&di, n * 4);
    if (e == hipSuccess) e = hipMalloc(&dp, n * 8);
    if (e == hipSuccess) e = hipMemcpy(dv, vals, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(di, 0xFF, n * 4);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(ntiles, (uint64_t)std::max<uint32_t>(ctx.num_cus, 1) * 8);
    if (e == hipSuccess) e = ctx.lane.before_launch();
    if (e == hipSuccess) e = launch_scan_test(dv, n, dp, di, ctx.lane.ws, ctx.lane.tot_d, grid, ctx.lane.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx.lane.stream);
    if (e == hipSuccess) e = hipMemcpy(prefix, dp, n * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(idx, di, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(ctx.lane.tot_h, ctx.lane.tot_d, sizeof(Totals), hipMemcpyDeviceToHost);
    if (e != hipSuccess) st = hip_status(e, "selftest_scan");
    else {
        totals[0] = ctx.lane.tot_h->keep;
        totals[1] = ctx.lane.tot_h->cnt;
    }
    cleanup();
    return st;
}

namespace {

// A one-vertex query on the device (k_single): the vertices of the root
// window that own a row (out-degree >= 1) and pass the self-loop checks, in
// id order.  With `fixed` (the allocation-free entry point) count mode only.
Status run_single(Ctx &ctx, const HostGraph &data, const DevGraph &dg, const Plan &plan, const gfsm_match_opts &o,
                  Result &res, Lane &L, const Arena *fixed)
{
    uint32_t lo = o.root_lo, hi = (o.root_hi || o.root_lo) ? o.root_hi : 0xFFFFFFFFu;
    if (hi == 0) hi = 0xFFFFFFFFu;
    hi = std::min(hi, data.V);
    lo = std::min(lo, hi);
    const bool rows = o.flags & (GFSM_ROWS | GFSM_SORTED);
    SingleArgs a{};
    a.g = dg.view;
    a.lo = lo;
    a.hi = hi;
    a.nchk = plan.dev.nroot_chk;
    std::memcpy(a.chk_rel, plan.dev.root_chk_rel, sizeof(a.chk_rel));
    a.tot = L.tot_d;
    a.ws = L.ws;
    const uint64_t tc = (uint64_t)kTile * kRootItems;
    a.ntiles = (hi - lo + tc - 1) / tc;
    if (rows) {
        if (fixed) return Status::err(GFSM_ERR_ARG, "rows need an allocation");
        a.cap = hi - lo;
        Status st = ensure_arena(ctx, std::max<uint64_t>(a.cap * 4, 256));
        if (!st.good()) return st;
        a.rows = reinterpret_cast<uint32_t *>(ctx.arena.base);
        if (a.ntiles > L.ws.max_tiles) return Status::err(GFSM_ERR_ARG, "internal: root window too large");
    }
    if (fixed) {
        HIP_TRY(launch_ws_clear(L.ws, L.stream));
        HIP_TRY(launch_ws_reset(L.ws, L.tot_d, L.stream));
    } else {
        HIP_TRY(hipMemsetAsync(L.tot_d, 0, sizeof(Totals), L.stream));
    }
    HIP_TRY(hipEventRecord(L.ev[0], L.stream));
    if (a.ntiles) {
        HIP_TRY(L.before_launch());
        const uint64_t grid = std::min<uint64_t>(a.ntiles, (uint64_t)std::max<uint32_t>(ctx.num_cus, 1) * ctx.tune.wgs_per_cu);
        HIP_TRY(launch_single(a, (uint32_t)grid, L.stream));
        res.stats.launches = 1;
    }
    HIP_TRY(hipEventRecord(L.ev[1], L.stream));
    HIP_TRY(hipMemcpyAsync(L.tot_h, L.tot_d, sizeof(Totals), hipMemcpyDeviceToHost, L.stream));
    HIP_TRY(hipStreamSynchronize(L.stream));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, L.ev[0], L.ev[1]));
    res.stats.device_ms = ms;
    if (rows) {
        res.count = L.tot_h->keep;
        res.rows.resize(res.count);
        if (res.count)
            HIP_TRY(hipMemcpy(res.rows.data(), a.rows, res.count * sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (o.flags & GFSM_DIGEST)
            for (uint32_t x : res.rows) {
                uint64_t h[2];
                row_hash(&x, 1, h);
                res.digest[0] += h[0];
                res.digest[1] += h[1];
            }
    } else {
        res.count = L.tot_h->count;
        res.digest[0] = L.tot_h->dg0;
        res.digest[1] = L.tot_h->dg1;
    }
    res.stats.level[0] = res.count;
    res.stats.roots = res.count;
    res.stats.cand[0] = hi - lo;
    if (o.max_rows && rows && res.count > o.max_rows) return Status::err(GFSM_ERR_CAP, "more rows than max_rows");
    return Status::ok();
}

// One single-query match on `L` with the data graph resident: the depth-first
// chunk walk from the roots, then the result.  With `fixed`, the levels live
// in that arena and nothing is allocated (count mode only).
Status run_match(Ctx &ctx, const HostGraph &data, const DevGraph &dg, const Plan &plan, const gfsm_match_opts &o,
                 Result &res, FILE *stream, Lane &L, const Arena *fixed)
{
    Runner run(ctx, data, dg, plan, o.flags, o.max_rows, res, L);
    run.stream = stream;
    run.fixed = fixed;
    run.no_alloc = fixed != nullptr;
    run.max_cand = o.max_cand;
    run.max_depth_cand = o.max_depth_cand;
    if (o.tile) {
        if (o.tile != 256 && o.tile != 512 && o.tile != 1024 && o.tile != 2048)
            return Status::err(GFSM_ERR_ARG, "tile must be 0, 256, 512, 1024 or 2048 candidates");
        run.forced_items = o.tile / kTile;
    }
    const uint64_t budget = o.mem_budget ? o.mem_budget : ctx.budget;
    Status st = run.carve(fixed ? std::min(budget, fixed->bytes) : budget);
    if (!st.good()) return st;
    // Capped: inner depths start with cap / kCapStartDiv candidates, then
    // follow the observed fan-out.  The counting last depth keeps whole launches: it
    // needs no level buffer, and its ranges are trimmed to the cap exactly.
    if (o.max_depth_cand)
        for (uint32_t k = 0; k + 1 < plan.q; k++)
            run.chunk[k] = std::max<uint64_t>(2048, std::min<uint64_t>(run.chunk[k], o.max_depth_cand / kCapStartDiv));
    uint64_t rlo, rhi;
    root_range(data, plan, o.root_lo, o.root_hi, rlo, rhi);

    if (fixed) {
        // The workspace in the shared array carries the previous call's words:
        // clear the tile words, the control block and the totals (fill
        // kernels, so the sequence holds no memset node).
        HIP_TRY(launch_ws_clear(L.ws, L.stream));
        HIP_TRY(launch_ws_reset(L.ws, L.tot_d, L.stream));
    } else {
        HIP_TRY(hipMemsetAsync(L.tot_d, 0, sizeof(Totals), L.stream));
    }
    HIP_TRY(hipEventRecord(L.ev[0], L.stream));
    st = run.run_roots(rlo, rhi);
    Status capped = Status::ok();
    if (!st.good()) {
        (void)hipStreamSynchronize(L.stream);
        if (st.code != GFSM_ERR_CAP) return st;
        capped = st;  // work cap: report what was done so far
        res.status = GFSM_ERR_CAP;
    }
    HIP_TRY(hipEventRecord(L.ev[1], L.stream));
    st = run.sync_totals();
    if (!st.good()) return st;
    st = run.prof_flush();
    if (!st.good()) return st;
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, L.ev[0], L.ev[1]));
    res.stats.device_ms = ms;
    const uint32_t q = plan.q;
    if (run.rows_mode && stream) {
        res.count = run.streamed;
        res.stats.level[q - 1] = res.count;
        res.digest[0] = run.sdg[0];
        res.digest[1] = run.sdg[1];
    } else if (run.rows_mode) {
        res.count = res.rows.size() / q;
        res.stats.level[q - 1] = res.count;
        if (o.flags & GFSM_DIGEST) {
            uint64_t d0 = 0, d1 = 0;
            for (uint64_t r = 0; r < res.count; r++) {
                uint64_t h[2];
                row_hash(res.rows.data() + r * q, q, h);
                d0 += h[0];
                d1 += h[1];
            }
            res.digest[0] = d0;
            res.digest[1] = d1;
        }
        if (o.flags & GFSM_SORTED) sort_rows(res.rows, q);
    } else {
        res.count = L.tot_h->count;
        res.stats.level[q - 1] = res.count;
        res.digest[0] = L.tot_h->dg0;
        res.digest[1] = L.tot_h->dg1;
    }
    for (uint32_t k = 1; k < q; k++) res.stats.checks[k] = plan.dev.depth[k].nchk;
    res.stats.alg_bytes = alg_bytes(plan, res.stats, o.flags);
    return capped;
}

}  // namespace

Status selftest_traffic(Ctx &ctx, uint32_t mode, uint64_t n, uint32_t launches, uint64_t bytes[2])
{
    if (mode > 2 || n == 0 || (mode >= 1 && (n & (n - 1))) || launches == 0)
        return Status::err(GFSM_ERR_ARG, "mode 0 (stream), 1 or 2 (gather, n a power of two), n > 0, launches > 0");
    std::lock_guard<std::mutex> lk(ctx.mu);
    HIP_TRY(hipSetDevice(ctx.device));
    const uint64_t src_words = mode ? 32 * n : n;
    const uint64_t dst_words = mode == 2 ? 2 * n : n;
    uint32_t *src = nullptr, *dst = nullptr;
    hipError_t e = hipMalloc(&src, src_words * 4);
    if (e == hipSuccess) e = hipMalloc(&dst, dst_words * 4);
    if (e == hipSuccess) e = hipMemsetAsync(src, 0, src_words * 4, ctx.lane.stream);
    for (uint32_t k = 0; k < launches && e == hipSuccess; k++) e = launch_calibrate(mode, src, dst, n, ctx.lane.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx.lane.stream);
    (void)hipFree(src);
    (void)hipFree(dst);
    if (e != hipSuccess) return hip_status(e, "selftest_traffic");
    bytes[0] = mode ? 128 * n : 4 * n;  // read per launch: whole lines for the gathers
    bytes[1] = 4 * dst_words;           // written per launch
    return Status::ok();
}

Status match_plain(Ctx &ctx, const HostGraph &data, const HostGraph &query, const gfsm_match_opts &o, Result &res,
                   FILE *stream)
{
    const double t0 = now_ms();
    res = Result{};
    if (data.mode != query.mode) return Status::err(GFSM_ERR_ARG, "query and data graph use different modes");
    Plan plan;
    Status st = plan_for(data, query, o, plan);
    if (!st.good()) return st;
    res.width = plan.q;
    res.stats.q = plan.q;
    std::lock_guard<std::mutex> lk(ctx.mu);
    HIP_TRY(hipSetDevice(ctx.device));
    DevGraph *dg = nullptr;
    st = upload_graph(ctx, data, dg);
    if (!st.good()) return st;
    if (plan.q == 1) {
        st = run_single(ctx, data, *dg, plan, o, res, ctx.lane, nullptr);  // rows returned, never streamed
        res.stats.host_ms = now_ms() - t0;
        return st;
    }
    st = run_match(ctx, data, *dg, plan, o, res, stream, ctx.lane, nullptr);
    res.stats.host_ms = now_ms() - t0;
    return st;
}

Status match_noalloc(Ctx &ctx, const HostGraph &data, const HostGraph &query, const gfsm_match_opts &o, Result &res)
{
    const double t0 = now_ms();
    res = Result{};
    if (o.flags & ~(GFSM_DIGEST | GFSM_ORDER_COST))
        return Status::err(GFSM_ERR_ARG, "the allocation-free match counts only: flags may hold GFSM_DIGEST and "
                                         "GFSM_ORDER_COST");
    if (o.mem_budget && o.mem_budget > kNoAllocLevelBytes)
        return Status::err(GFSM_ERR_ARG, "mem_budget exceeds the allocation-free match's level scratch (" +
                                             std::to_string(kNoAllocLevelBytes) + " bytes)");
    if (data.mode != query.mode) return Status::err(GFSM_ERR_ARG, "query and data graph use different modes");
    Plan plan;
    Status st = plan_for(data, query, o, plan);
    if (!st.good()) return st;
    res.width = plan.q;
    res.stats.q = plan.q;
    DevGraph *dg = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_dev_mu);
        auto it = data.dev.find(&ctx);
        if (it != data.dev.end()) dg = it->second;
    }
    if (!dg) return Status::err(GFSM_ERR_ARG, "the data graph is not resident on this context (gfsm_graph_upload)");
    if (!ctx.scratch) return Status::err(GFSM_ERR_DEVICE, "no device scratch on this context");
    // Calls on one device share its scratch array: one at a time.
    std::lock_guard<std::mutex> dev_lk(noalloc_mutex(ctx.device));
    std::lock_guard<std::mutex> lk(ctx.mu);
    HIP_TRY(hipSetDevice(ctx.device));
    // The lane: this context's stream, events and pinned totals; the scan
    // workspace, device totals, tile index and levels in the scratch array.
    char *p = ctx.scratch;
    Lane L;
    L.stream = ctx.lane.stream;
    L.ev[0] = ctx.lane.ev[0];
    L.ev[1] = ctx.lane.ev[1];
    L.tot_h = ctx.lane.tot_h;
    L.kc_h = ctx.lane.kc_h;
    L.ws.ctl = reinterpret_cast<uint32_t *>(p);
    L.tot_d = reinterpret_cast<Totals *>(p + 256);
    L.ws.wk = reinterpret_cast<uint64_t *>(p + 512);
    L.ws.wc = L.ws.wk + kNoAllocTiles;
    L.tile_node = L.ws.wc + kNoAllocTiles;
    L.ws.max_tiles = kNoAllocTiles;
    Arena levels;
    levels.base = p + kNoAllocWorkBytes;
    levels.bytes = kNoAllocLevelBytes;
    st = plan.q == 1 ? run_single(ctx, data, *dg, plan, o, res, L, &levels)
                     : run_match(ctx, data, *dg, plan, o, res, nullptr, L, &levels);  // L borrows: nothing to release
    res.stats.host_ms = now_ms() - t0;
    return st;
}

}  // namespace gfsm
