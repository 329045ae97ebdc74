// batch.cpp -- the batched matcher: many queries in flight on one device,
// advanced together one step at a time (no reference counterpart; the
// reference runs one query per process, main.cu:56-63).
//
// Every query admitted to the engine owns a slot (its plan, its level
// buffers and its depth-first chunk walk, all in device memory).  One step is
// a scheduler launch (one workgroup: consume each slot's finished step,
// choose its next one, lay all slots' tiles end to end) followed by one
// persistent step launch that hands out (slot, tile) work items by ticket.
// Steps need no host round trip, so K of them are captured once in a hipGraph
// and replayed; between replays the host harvests finished slots and admits
// waiting queries into them.  The arena is shared by the slots' regions,
// sized per query by a projection of its intermediate levels (admission by
// projected memory); a level that outgrows its region is expanded in smaller
// chunks by the device scheduler, so a projection only affects speed, never
// the result.
#include "serve.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace gfsm {

#define HIP_TRY(expr)                                                    \
    do {                                                                 \
        hipError_t e_ = (expr);                                          \
        if (e_ != hipSuccess) return hip_status(e_, #expr);              \
    } while (0)

static double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

static uint64_t round_up(uint64_t b, uint64_t a) { return (b + a - 1) / a * a; }
static constexpr uint64_t kRegionAlign = 2ull << 20;
static constexpr uint64_t kMinLevelNodes = 8 * kTile;  // >= the largest tile's survivors

// ---------------------------------------------------------------------------
// RegionAlloc
// ---------------------------------------------------------------------------

void RegionAlloc::reset(uint64_t bytes)
{
    free_map_.clear();
    if (bytes) free_map_[0] = bytes;
    free_ = bytes;
}

uint64_t RegionAlloc::take(uint64_t bytes)
{
    for (auto it = free_map_.begin(); it != free_map_.end(); ++it) {
        if (it->second < bytes) continue;
        const uint64_t off = it->first, rest = it->second - bytes;
        free_map_.erase(it);
        if (rest) free_map_[off + bytes] = rest;
        free_ -= bytes;
        return off;
    }
    return UINT64_MAX;
}

void RegionAlloc::give(uint64_t off, uint64_t bytes)
{
    auto it = free_map_.emplace(off, bytes).first;
    free_ += bytes;
    auto nx = std::next(it);
    if (nx != free_map_.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        free_map_.erase(nx);
    }
    if (it != free_map_.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            free_map_.erase(it);
        }
    }
}

uint64_t RegionAlloc::largest() const
{
    uint64_t m = 0;
    for (auto &kv : free_map_) m = std::max(m, kv.second);
    return m;
}

// ---------------------------------------------------------------------------
// memory projection
// ---------------------------------------------------------------------------

static uint64_t node_bytes(uint32_t k) { return 4ull * (k + 1) + 12; }

void BatchEngine::set_max_region(uint64_t b)
{
    max_region_ = std::max<uint64_t>(kRegionAlign, b / kRegionAlign * kRegionAlign);
}

// Smallest region whose equal per-level share still gives every level
// kMinLevelNodes nodes after carve()'s slack.
uint64_t BatchEngine::min_region(uint32_t q) const
{
    if (q < 2) return kRegionAlign;
    const uint64_t per_level = (kMinLevelNodes + 512 + 1) * node_bytes(q - 2) + 1024;
    return round_up(per_level * (q - 1), kRegionAlign);
}

// Model bytes of a query's levels (project_levels at the level layout's
// node sizes, the last level excluded: counting needs no buffer).
double BatchEngine::model_bytes(const Plan &plan, uint64_t nroots, uint64_t g1) const
{
    double nodes[GFSM_MAXQ], cand[GFSM_MAXQ];
    project_levels(data_, plan, nroots, g1, nodes, cand);
    double bytes = 0;
    for (uint32_t k = 0; k + 1 < plan.q; k++) bytes += nodes[k] * node_bytes(k);
    return bytes;
}

// Region bytes a query is given: the model scaled by the calibration learnt
// from finished queries, with headroom, clamped to [min, max] region.
uint64_t BatchEngine::project_bytes(const Plan &plan, uint64_t nroots, uint64_t g1) const
{
    return project_model(plan.q, model_bytes(plan, nroots, g1));
}

uint64_t BatchEngine::project_model(uint32_t q, double model) const
{
    const double b = model * calib_ * 1.25;
    const uint64_t lo = min_region(q);
    const uint64_t hi = std::max(lo, max_region_);
    if (!(b < (double)hi)) return hi;
    return std::max(lo, round_up((uint64_t)b, kRegionAlign));
}

// Calibration: the geometric mean (exponentially weighted) of actual over
// model level bytes of finished, uncapped queries with non-trivial levels.
void BatchEngine::calibrate(double model, const Result &r, uint32_t q)
{
    if (r.status != GFSM_OK || model < 4096) return;
    double actual = 0;
    for (uint32_t k = 0; k + 1 < q; k++) actual += (double)r.stats.level[k] * node_bytes(k);
    if (actual < 4096) return;
    const double ratio = std::min(256.0, std::max(1.0 / 256, actual / model));
    calib_log_ = 0.9 * calib_log_ + 0.1 * std::log(ratio);
    calib_ = std::exp(calib_log_);
    calib_n_++;
}

// ---------------------------------------------------------------------------
// BatchEngine
// ---------------------------------------------------------------------------

BatchEngine::BatchEngine(Ctx &ctx, const HostGraph &data) : ctx_(ctx), data_(data)
{
    trace_ = std::getenv("GFSM_BATCH_TRACE") != nullptr;  // one stderr line per replay
}

BatchEngine::~BatchEngine()
{
    (void)hipSetDevice(ctx_.device);
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    lane_.destroy();
    for (auto &e : ev_)
        if (e) (void)hipEventDestroy(e);
    (void)hipFree(slots_d_);
    (void)hipFree(report_d_);
    (void)hipFree(final_d_);
    (void)hipFree(tile_start_d_);
    (void)hipFree(ctl_d_);
    (void)hipFree(plans_d_);
    if (own_arena_) (void)hipFree(arena_);
    if (stage_h_) (void)hipHostFree(stage_h_);
    if (report_h_) (void)hipHostFree(report_h_);
    if (final_h_) (void)hipHostFree(final_h_);
    if (step_cand_h_) (void)hipHostFree(step_cand_h_);
    if (n_admit_h_) (void)hipHostFree(n_admit_h_);
    if (admit_h_) (void)hipHostFree(admit_h_);
    (void)hipFree(admit_d_);
}

Status BatchEngine::init(const BatchShape &shape, uint64_t arena_bytes, char *arena)
{
    shape_ = shape;
    if (shape_.slots == 0 || shape_.slots > kMaxSlots)
        return Status::err(GFSM_ERR_ARG, "batch slots must be 1..1024");
    if (shape_.items != 1 && shape_.items != 2 && shape_.items != 4 && shape_.items != 8)
        return Status::err(GFSM_ERR_ARG, "batch items must be 1, 2, 4 or 8");
    if (expand_lds_bytes(shape_.items, shape_.pool_words) > kMaxDynLds)
        return Status::err(GFSM_ERR_ARG, "batch LDS pool too large");
    if (shape_.steps_per_replay == 0) shape_.steps_per_replay = 1;
    HIP_TRY(hipSetDevice(ctx_.device));
    Status st = upload_graph(ctx_, data_, dg_);
    if (!st.good()) return st;

    const uint32_t n = shape_.slots;
    const uint64_t tc = (uint64_t)kTile * shape_.items;
    const uint64_t max_tiles = 1ull << 22;
    // The per-slot step size is step_budget / running (set before each
    // replay), so one step needs at most step_budget / tc + n tiles.
    shape_.step_budget = std::max<uint64_t>(tc, std::min<uint64_t>(shape_.step_budget, (max_tiles - n) * tc));
    st = lane_.init(max_tiles, true);
    if (!st.good()) return st;
    HIP_TRY(hipMalloc(&slots_d_, sizeof(BatchSlot) * n));
    HIP_TRY(hipMalloc(&report_d_, sizeof(SlotReport) * n));
    HIP_TRY(hipMalloc(&final_d_, sizeof(SlotFinal) * n));
    HIP_TRY(hipMalloc(&tile_start_d_, sizeof(uint64_t) * (n + 1)));
    HIP_TRY(hipMalloc(&ctl_d_, sizeof(BatchCtl)));
    HIP_TRY(hipMalloc(&plans_d_, sizeof(QueryPlanDev) * n));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&stage_h_), sizeof(BatchSlot) * n, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&report_h_), sizeof(SlotReport) * n, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&final_h_), sizeof(SlotFinal) * n, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&step_cand_h_), sizeof(uint64_t), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&n_admit_h_), sizeof(uint32_t), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&admit_h_), sizeof(AdmitEntry) * n, hipHostMallocDefault));
    HIP_TRY(hipMalloc(&admit_d_, sizeof(AdmitEntry) * n));
    std::memset(stage_h_, 0, sizeof(BatchSlot) * n);
    for (uint32_t s = 0; s < n; s++) {
        stage_h_[s].state = SLOT_FREE;
        stage_h_[s].step = STEP_NONE;
    }
    HIP_TRY(hipMemcpy(slots_d_, stage_h_, sizeof(BatchSlot) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(report_d_, 0, sizeof(SlotReport) * n));
    HIP_TRY(hipMemset(final_d_, 0, sizeof(SlotFinal) * n));
    HIP_TRY(hipMemset(tile_start_d_, 0, sizeof(uint64_t) * (n + 1)));
    HIP_TRY(hipMemset(ctl_d_, 0, sizeof(BatchCtl)));
    if (arena) {
        arena_bytes = arena_bytes / kRegionAlign * kRegionAlign;
        if (arena_bytes < kRegionAlign) return Status::err(GFSM_ERR_NOMEM, "batch arena too small");
        arena_ = arena;
    } else {
        arena_bytes = round_up(std::max<uint64_t>(arena_bytes, kRegionAlign), kRegionAlign);
        HIP_TRY(hipMalloc(&arena_, arena_bytes));
        own_arena_ = true;
    }
    arena_bytes_ = arena_bytes;
    regions_.reset(arena_bytes);
    max_region_ = arena_bytes;
    jobs_.clear();
    jobs_.resize(n);
    for (auto &e : ev_) HIP_TRY(hipEventCreate(&e));

    cfg_ = BatchCfg{};
    cfg_.slots = slots_d_;
    cfg_.plans = plans_d_;
    cfg_.admit = admit_d_;
    cfg_.report = report_d_;
    cfg_.final = final_d_;
    cfg_.tile_start = tile_start_d_;
    cfg_.tile_node = lane_.tile_node;
    cfg_.ctl = ctl_d_;
    cfg_.nslots = n;
    cfg_.items = shape_.items;
    cfg_.pool_words = shape_.pool_words;
    cfg_.max_step_cand = shape_.step_budget;
    cfg_.max_degree = std::max<uint32_t>(data_.max_degree, 1);
    cfg_.g = dg_->view;
    cfg_.ws = lane_.ws;
    grid_ = std::max<uint32_t>(ctx_.num_cus, 1) * std::max<uint32_t>(shape_.wgs_per_cu, 1);
    // The schedule is fresh between replays: replays run [tile index, step,
    // schedule] x K.  One empty step outside capture also sets the step
    // kernel's LDS limit.
    HIP_TRY(launch_batch_schedule(cfg_, lane_.stream));
    HIP_TRY(launch_batch_tile_index(cfg_, lane_.stream));
    HIP_TRY(launch_batch_step(cfg_, grid_, lane_.stream));
    HIP_TRY(launch_batch_schedule(cfg_, lane_.stream));
    HIP_TRY(hipStreamSynchronize(lane_.stream));
    since_clear_ = 1;
    if (shape_.use_graph) return capture();
    return Status::ok();
}

Status BatchEngine::capture()
{
    HIP_TRY(hipStreamBeginCapture(lane_.stream, hipStreamCaptureModeRelaxed));
    hipError_t e = hipSuccess;
    for (uint32_t i = 0; i < shape_.steps_per_replay && e == hipSuccess; i++) {
        e = launch_batch_tile_index(cfg_, lane_.stream);
        if (e == hipSuccess) e = launch_batch_step(cfg_, grid_, lane_.stream);
        if (e == hipSuccess) e = launch_batch_schedule(cfg_, lane_.stream);
    }
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(lane_.stream, &g);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_status(e, "batch step capture");
    }
    HIP_TRY(e2);
    graph_ = g;
    HIP_TRY(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
    return Status::ok();
}

uint32_t BatchEngine::free_slots() const
{
    uint32_t f = 0;
    for (auto &j : jobs_) f += j ? 0u : 1u;
    return f;
}

// Level buffers of a slot inside its region: an equal byte share per level.
void BatchEngine::carve(BatchSlot &s, const Plan &plan, char *base, uint64_t bytes) const
{
    const uint32_t q = plan.q;
    const uint64_t per_level = bytes / (q - 1);
    char *p = base;
    auto take = [&p](uint64_t b) {
        char *r = p;
        p += round_up(b, 256);
        return r;
    };
    for (uint32_t k = 0; k + 1 < q; k++) {
        const uint64_t nb = node_bytes(k);
        const uint64_t c = per_level / nb > 512 ? per_level / nb - 512 : 0;
        s.cap[k] = c;
        s.vec[k] = reinterpret_cast<uint32_t *>(take(c * 4 * (k + 1)));
        s.off[k] = reinterpret_cast<uint64_t *>(take((c + 1) * 8));
        s.gbeg[k] = reinterpret_cast<uint32_t *>(take(c * 4));
    }
}

void BatchEngine::model_job(const HostGraph &data, BatchJob &job)
{
    root_range(data, job.plan, job.opts.root_lo, job.opts.root_hi, job.rlo, job.rhi);
    job.g1 = 0;
    for (uint64_t i = job.rlo; i < job.rhi; i++) job.g1 += data.rel_cnt[i];
    double nodes[GFSM_MAXQ], cand[GFSM_MAXQ];
    project_levels(data, job.plan, job.rhi - job.rlo, job.g1, nodes, cand);
    double bytes = 0;
    for (uint32_t k = 0; k + 1 < job.plan.q; k++) bytes += nodes[k] * node_bytes(k);
    job.model_bytes = bytes;
    job.modeled = true;
}

bool BatchEngine::try_admit(std::unique_ptr<BatchJob> &job, Status &st)
{
    st = Status::ok();
    uint32_t slot = UINT32_MAX;
    for (uint32_t s = 0; s < shape_.slots; s++)
        if (!jobs_[s]) {
            slot = s;
            break;
        }
    if (slot == UINT32_MAX) return false;
    const Plan &plan = job->plan;
    if (!job->modeled) model_job(data_, *job);
    const uint64_t rlo = job->rlo, rhi = job->rhi;
    const uint64_t lo = min_region(plan.q);
    if (lo > arena_bytes_) {
        st = Status::err(GFSM_ERR_NOMEM, "batch arena too small for this query");
        return false;
    }
    // At least an equal share of the arena (so small levels never force tiny
    // chunks), more when the projection asks for it, at most max_region.
    const uint64_t share = arena_bytes_ / shape_.slots / kRegionAlign * kRegionAlign;
    const uint64_t cap = std::min(max_region_, arena_bytes_);
    uint64_t want = std::min(cap, std::max(share, project_model(plan.q, job->model_bytes)));
    want = std::max(lo, want);
    const uint64_t off = regions_.take(want);
    if (off == UINT64_MAX) return false;
    job->region_off = off;
    job->region_bytes = want;
    job->nroots = rhi - rlo;
    job->t_admit = now_ms();

    AdmitEntry &ae = admit_h_[n_admit_];
    std::memset(&ae, 0, sizeof(AdmitEntry));
    ae.slot = slot;
    ae.plan = plan.dev;
    BatchSlot &s = ae.s;
    s.plan = plans_d_ + slot;
    s.root_vtx = dg_->rel_vtx + rlo;
    s.root_beg = dg_->rel_beg + rlo;
    s.root_cnt = dg_->rel_cnt + rlo;
    s.nroots = rhi - rlo;
    carve(s, plan, arena_ + off, want);
    s.max_cand = job->opts.max_cand;
    s.max_depth_cand = job->opts.max_depth_cand;
    s.q = plan.q;
    s.digest = (job->opts.flags & GFSM_DIGEST) ? 1u : 0u;
    s.state = SLOT_NEW;
    s.step = STEP_NONE;
    n_admit_++;  // installed by the next replay's first scheduler launch
    jobs_[slot] = std::move(job);
    running_++;
    return true;
}

Status BatchEngine::replay()
{
    const uint32_t K = shape_.steps_per_replay;
    // Per-slot step size: the step budget shared by the running slots, so a
    // lone query takes large steps and a full batch keeps steps short.
    const uint64_t tc = (uint64_t)kTile * shape_.items;
    const uint64_t want = std::max<uint64_t>(
        tc, std::min<uint64_t>(shape_.step_budget, std::max<uint64_t>(1ull << 20, shape_.step_budget / std::max<uint32_t>(running_, 1))));
    if (want != step_cand_) {
        *step_cand_h_ = want;
        HIP_TRY(hipMemcpyAsync(&ctl_d_->step_cand, step_cand_h_, sizeof(uint64_t), hipMemcpyHostToDevice,
                               lane_.stream));
        step_cand_ = want;
    }
    if (since_clear_ + K >= (1ull << kEpochBits)) {
        HIP_TRY(launch_ws_clear(lane_.ws, lane_.stream));
        since_clear_ = 0;
    }
    if (n_admit_) {
        // One copy of the queued admissions; the scheduler installs them.
        *n_admit_h_ = n_admit_;
        HIP_TRY(hipMemcpyAsync(admit_d_, admit_h_, sizeof(AdmitEntry) * n_admit_, hipMemcpyHostToDevice,
                               lane_.stream));
        HIP_TRY(hipMemcpyAsync(&ctl_d_->n_admit, n_admit_h_, sizeof(uint32_t), hipMemcpyHostToDevice,
                               lane_.stream));
    }
    HIP_TRY(hipEventRecord(ev_[0], lane_.stream));
    if (exec_) {
        HIP_TRY(hipGraphLaunch(exec_, lane_.stream));
    } else {
        for (uint32_t i = 0; i < K; i++) {
            HIP_TRY(launch_batch_tile_index(cfg_, lane_.stream));
            HIP_TRY(launch_batch_step(cfg_, grid_, lane_.stream));
            HIP_TRY(launch_batch_schedule(cfg_, lane_.stream));
        }
    }
    HIP_TRY(hipEventRecord(ev_[1], lane_.stream));
    since_clear_ += K;
    steps_ += K;
    HIP_TRY(hipMemcpyAsync(report_h_, report_d_, sizeof(SlotReport) * shape_.slots, hipMemcpyDeviceToHost,
                           lane_.stream));
    // With the reports, the statistics of slots that finished (written once
    // per finish): a harvest needs nothing more from the device.
    HIP_TRY(hipMemcpyAsync(final_h_, final_d_, sizeof(SlotFinal) * shape_.slots, hipMemcpyDeviceToHost,
                           lane_.stream));
    HIP_TRY(hipStreamSynchronize(lane_.stream));
    n_admit_ = 0;
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
    device_ms_ += ms;
    if (trace_) {
        uint64_t steps_sum = 0;
        for (uint32_t s = 0; s < shape_.slots; s++) steps_sum += report_h_[s].steps;
        BatchCtl c{};
        (void)hipMemcpy(&c, ctl_d_, sizeof(BatchCtl), hipMemcpyDeviceToHost);
        std::fprintf(stderr,
                     "batch replay %llu: running %u step_cand %llu device %.3f ms slot-steps %llu next-step tiles %llu\n",
                     (unsigned long long)(steps_ / K), running_, (unsigned long long)step_cand_, ms,
                     (unsigned long long)steps_sum, (unsigned long long)c.total_tiles);
    }
    return Status::ok();
}

Status BatchEngine::harvest(const std::function<void(std::unique_ptr<BatchJob>, Result &)> &done)
{
    std::vector<uint32_t> fin;
    for (uint32_t s = 0; s < shape_.slots; s++) {
        if (!jobs_[s]) continue;
        const uint32_t st = report_h_[s].state;
        if (st == SLOT_DONE || st == SLOT_CAP || st == SLOT_ERROR) fin.push_back(s);
    }
    if (fin.empty()) return Status::ok();
    const double t = now_ms();
    for (uint32_t s : fin) {
        std::unique_ptr<BatchJob> job = std::move(jobs_[s]);
        const SlotReport &b = report_h_[s];
        const SlotFinal &f = final_h_[s];
        Result r;
        const uint32_t q = job->plan.q;
        r.width = q;
        r.count = b.count;
        r.digest[0] = b.dg0;
        r.digest[1] = b.dg1;
        r.stats.q = q;
        r.stats.launches = b.steps;
        r.stats.overflows = f.overflows;
        r.stats.roots = job->nroots;
        for (uint32_t k = 0; k < q && k < GFSM_MAXQ; k++) {
            r.stats.level[k] = f.level[k];
            r.stats.cand[k] = f.cand[k];
            if (k) r.stats.checks[k] = job->plan.dev.depth[k].nchk;
        }
        r.stats.alg_bytes = alg_bytes(job->plan, r.stats, job->opts.flags);
        r.stats.device_ms = t - job->t_admit;
        if (b.state == SLOT_CAP) r.status = GFSM_ERR_CAP;
        if (b.state == SLOT_ERROR) r.status = GFSM_ERR_ARG;
        regions_.give(job->region_off, job->region_bytes);
        running_--;
        calibrate(job->model_bytes, r, q);
        done(std::move(job), r);
    }
    return Status::ok();
}

Status BatchEngine::advance(const std::function<void(std::unique_ptr<BatchJob>, Result &)> &done)
{
    if (running_ == 0) return Status::ok();
    Status st = replay();
    if (!st.good()) return st;
    return harvest(done);
}

// ---------------------------------------------------------------------------
// gfsm_match_batch
// ---------------------------------------------------------------------------

// The engine a context keeps between batch calls.
struct BatchCache {
    std::unique_ptr<BatchEngine> eng;
    BatchShape shape;
    uint64_t budget = 0;
    bool same(const BatchShape &s, uint64_t b) const
    {
        return b == budget && s.slots == shape.slots && s.items == shape.items && s.pool_words == shape.pool_words &&
               s.wgs_per_cu == shape.wgs_per_cu && s.steps_per_replay == shape.steps_per_replay &&
               s.step_budget == shape.step_budget && s.use_graph == shape.use_graph;
    }
};

static BatchShape shape_from(const Ctx &ctx)
{
    BatchShape sh;
    sh.slots = ctx.tune.batch_slots;
    sh.items = ctx.tune.batch_items;
    sh.pool_words = ctx.tune.batch_pool_words;
    sh.wgs_per_cu = ctx.tune.batch_wgs_per_cu;
    sh.steps_per_replay = ctx.tune.batch_steps;
    sh.step_budget = ctx.tune.batch_step_budget;
    return sh;
}

Status match_batch_shaped(Ctx &ctx, const HostGraph &data, const std::vector<const HostGraph *> &queries,
                          const std::vector<gfsm_match_opts> &opts, std::vector<Result> &out, BatchShape shape,
                          double *device_ms)
{
    const size_t n = queries.size();
    out.assign(n, Result{});
    if (opts.size() != 1 && opts.size() != n) return Status::err(GFSM_ERR_ARG, "opts must have 1 or n entries");
    auto opt = [&](size_t i) -> const gfsm_match_opts & { return opts.size() == 1 ? opts[0] : opts[i]; };
    std::deque<std::unique_ptr<BatchJob>> pending;
    std::vector<size_t> plain;
    for (size_t i = 0; i < n; i++) {
        const gfsm_match_opts &o = opt(i);
        if (data.mode != queries[i]->mode) {
            out[i].status = GFSM_ERR_ARG;
            continue;
        }
        std::unique_ptr<BatchJob> job(new BatchJob);
        job->id = i;
        job->query = queries[i];
        job->opts = o;
        Status st = plan_for(data, *queries[i], o, job->plan);
        if (!st.good()) {
            out[i].status = st.code;
            continue;
        }
        if (job->plan.q == 1 || (o.flags & (GFSM_ROWS | GFSM_SORTED)) || o.tile) {
            plain.push_back(i);
            continue;
        }
        pending.push_back(std::move(job));
    }
    // Single-vertex queries, row output and forced tiles: single-query path.
    for (size_t i : plain) {
        Status st = match_plain(ctx, data, *queries[i], opt(i), out[i]);
        if (!st.good() && st.code != GFSM_ERR_CAP) out[i].status = st.code;
    }
    if (pending.empty()) return Status::ok();

    std::lock_guard<std::mutex> lk(ctx.mu);
    HIP_TRY(hipSetDevice(ctx.device));
    const uint64_t budget = opts[0].mem_budget ? opts[0].mem_budget : ctx.budget;
    // Reuse the context's engine when graph, shape and budget match: its
    // arena, slots and captured steps survive between calls.
    BatchCache *bc = static_cast<BatchCache *>(ctx.batch_cache);
    if (!bc || ctx.batch_graph != &data || !bc->same(shape, budget)) {
        ctx.drop_batch_cache();
        std::unique_ptr<BatchCache> nc(new BatchCache);
        nc->shape = shape;
        nc->budget = budget;
        nc->eng.reset(new BatchEngine(ctx, data));
        Status st = ensure_arena(ctx.batch_arena, budget);
        if (st.good()) st = nc->eng->init(shape, budget, ctx.batch_arena.base);
        if (!st.good()) return st;
        ctx.batch_cache = nc.release();
        ctx.batch_cache_free = [](void *p) { delete static_cast<BatchCache *>(p); };
        ctx.batch_graph = &data;
        bc = static_cast<BatchCache *>(ctx.batch_cache);
    }
    BatchEngine &eng = *bc->eng;
    // The batch's arena is its budget; each query may take up to an equal
    // share of it, so every slot can be filled.
    eng.set_max_region(std::max(eng.min_region(GFSM_MAXQ), budget / std::max<uint32_t>(shape.slots, 1)));
    const double dev0 = eng.device_ms();
    auto done = [&](std::unique_ptr<BatchJob> job, Result &r) { out[job->id] = std::move(r); };
    Status st;
    while (!pending.empty() || eng.running()) {
        while (!pending.empty()) {
            if (!eng.try_admit(pending.front(), st)) break;
            pending.pop_front();
        }
        if (st.good() && !eng.running()) st = Status::err(GFSM_ERR_NOMEM, "batch arena cannot hold the next query");
        if (st.good()) st = eng.advance(done);
        if (!st.good()) {
            ctx.drop_batch_cache();  // slots may still hold work: start afresh next time
            return st;
        }
    }
    if (device_ms) *device_ms = eng.device_ms() - dev0;
    return Status::ok();
}

Status match_batch(Ctx &ctx, const HostGraph &data, const std::vector<const HostGraph *> &queries,
                   const std::vector<gfsm_match_opts> &opts, std::vector<Result> &out)
{
    return match_batch_shaped(ctx, data, queries, opts, out, shape_from(ctx), nullptr);
}

}  // namespace gfsm
