// serve.h -- batched matching and the serving loop (no reference counterpart).
//
// Both run on BatchEngine: a set of device-resident query slots advanced
// together, one step per [scheduler, step-kernel] launch pair, with runs of
// steps captured in a hipGraph and replayed.  The host only admits queries
// into free slots (each with a device memory region sized by a projection of
// its intermediate levels) and harvests finished ones between replays.
#pragma once

#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <thread>

#include "engine.h"

namespace gfsm {

// First-fit allocator of device byte ranges inside one arena (offsets only;
// the host owns the bookkeeping, the arena is one hipMalloc).
class RegionAlloc {
public:
    void reset(uint64_t bytes);
    // offset of a free range of `bytes`, or UINT64_MAX when none fits
    uint64_t take(uint64_t bytes);
    void give(uint64_t off, uint64_t bytes);
    uint64_t free_bytes() const { return free_; }
    uint64_t largest() const;

private:
    std::map<uint64_t, uint64_t> free_map_;  // offset -> size
    uint64_t free_ = 0;
};

// One admitted query as the host sees it.
struct BatchJob {
    uint64_t id = 0;                     // caller's tag (index or ticket)
    const HostGraph *query = nullptr;
    gfsm_match_opts opts{};
    Plan plan;
    uint64_t region_off = 0, region_bytes = 0;
    uint64_t nroots = 0;                 // root candidates of its window
    // The root window's size and the level model, computed ahead of
    // admission (by the submitting thread) when `modeled` is set.
    bool modeled = false;
    uint64_t rlo = 0, rhi = 0, g1 = 0;
    bool waited_memory = false;          // the server held it back for projected memory (a slot was free)
    bool waited_slot = false;            // ... for a free slot
    double model_bytes = 0;  // projected level bytes before calibration
    double t_admit = 0;
};

struct BatchShape {
    uint32_t slots = 256;
    uint32_t items = 4;
    uint32_t pool_words = kPoolWords;
    uint32_t wgs_per_cu = 8;
    uint32_t steps_per_replay = 8;
    uint64_t step_budget = 1ull << 26;  // candidates per step over all running slots
    bool use_graph = true;
};

class BatchEngine {
public:
    BatchEngine(Ctx &ctx, const HostGraph &data);
    ~BatchEngine();
    // Device resources: slots, workspace, an arena of `arena_bytes` for the
    // slots' level regions (`arena` when given, owned by the caller, else
    // allocated here), and the captured step graph.
    Status init(const BatchShape &shape, uint64_t arena_bytes, char *arena = nullptr);

    // Device bytes a query is given: a projection of its level sizes from
    // the data graph's relation statistics, clamped to [min, max] region.
    uint64_t project_bytes(const Plan &plan, uint64_t nroots, uint64_t g1) const;
    double model_bytes(const Plan &plan, uint64_t nroots, uint64_t g1) const;
    uint64_t project_model(uint32_t q, double model) const;
    // Fill job's root window, depth-1 candidates and level model.
    static void model_job(const HostGraph &data, BatchJob &job);
    // Learn the model's error from a finished query (actual / model level bytes).
    void calibrate(double model, const Result &r, uint32_t q);
    double calibration() const { return calib_; }
    uint64_t min_region(uint32_t q) const;
    uint64_t max_region() const { return max_region_; }
    // Largest region a query may take, rounded down to the region alignment
    // so that every region offset keeps it.
    void set_max_region(uint64_t b);

    // Admit a planned job into a free slot when its region fits; false when
    // no slot or no memory is free right now.  The job's region is taken
    // from the arena.
    bool try_admit(std::unique_ptr<BatchJob> &job, Status &st);
    // Run steps until at least one slot finishes or nothing runs; harvested
    // jobs are passed to `done` with their results.
    Status advance(const std::function<void(std::unique_ptr<BatchJob>, Result &)> &done);
    uint32_t running() const { return running_; }
    uint32_t free_slots() const;
    uint64_t steps() const { return steps_; }
    double device_ms() const { return device_ms_; }

private:
    Status capture();
    Status replay();
    Status harvest(const std::function<void(std::unique_ptr<BatchJob>, Result &)> &done);
    void carve(BatchSlot &s, const Plan &plan, char *base, uint64_t bytes) const;

    Ctx &ctx_;
    const HostGraph &data_;
    DevGraph *dg_ = nullptr;
    BatchShape shape_;
    Lane lane_;
    BatchCfg cfg_{};
    BatchSlot *slots_d_ = nullptr;
    BatchSlot *stage_h_ = nullptr;       // pinned staging for admissions
    SlotReport *report_d_ = nullptr, *report_h_ = nullptr;
    SlotFinal *final_d_ = nullptr, *final_h_ = nullptr;
    uint64_t *tile_start_d_ = nullptr;
    BatchCtl *ctl_d_ = nullptr;
    QueryPlanDev *plans_d_ = nullptr;
    char *arena_ = nullptr;
    bool own_arena_ = false;
    uint64_t arena_bytes_ = 0;
    RegionAlloc regions_;
    uint64_t max_region_ = 0;
    std::vector<std::unique_ptr<BatchJob>> jobs_;  // per slot
    uint32_t running_ = 0;
    uint64_t steps_ = 0, since_clear_ = 0;
    uint64_t step_cand_ = 0;             // per-slot step size last written to the device
    double calib_ = 1, calib_log_ = 0;   // projection calibration (see calibrate)
    uint64_t calib_n_ = 0;
    uint64_t *step_cand_h_ = nullptr;    // pinned source of that write
    AdmitEntry *admit_h_ = nullptr, *admit_d_ = nullptr;  // admission queue (pinned, device)
    uint32_t *n_admit_h_ = nullptr;
    uint32_t n_admit_ = 0;               // entries queued since the last replay
    double device_ms_ = 0;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    hipEvent_t ev_[2] = {nullptr, nullptr};
    uint32_t grid_ = 0;
    bool trace_ = false;
};

// Many queries at once (gfsm_match_batch).  opts has 1 or n entries.
Status match_batch(Ctx &ctx, const HostGraph &data, const std::vector<const HostGraph *> &queries,
                   const std::vector<gfsm_match_opts> &opts, std::vector<Result> &out);

// gfsm_match_batch with an explicit engine shape (the autotuner's entry);
// *device_ms = device time of the batched steps.
Status match_batch_shaped(Ctx &ctx, const HostGraph &data, const std::vector<const HostGraph *> &queries,
                          const std::vector<gfsm_match_opts> &opts, std::vector<Result> &out, BatchShape shape,
                          double *device_ms);

// Shape search over the batch engine and the single-query path.
Status autotune(Ctx &ctx, const HostGraph &data, const std::vector<const HostGraph *> &queries,
                const std::vector<gfsm_match_opts> &opts, uint32_t which);

struct Server;
Status server_start(Ctx &ctx, const HostGraph &data, const gfsm_serve_opts &opts, Server *&out);
Status server_submit(Server &s, const HostGraph &query, const gfsm_match_opts *opts, uint64_t &ticket);
Status server_take(Server &s, uint64_t ticket, bool wait, Result &out);
Status server_times(Server &s, uint64_t ticket, double t[3]);
Status server_stats(Server &s, gfsm_serve_stats &out);
void server_stop(Server *s);

}  // namespace gfsm
