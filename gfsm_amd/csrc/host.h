// host.h -- host-side types of libgfsm: graphs in the relation-partitioned
// layout, compiled query plans, results, and the error channel of the C ABI.
#pragma once

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>
#include <mutex>
#include <unordered_map>

#include "gfsm.h"
#include "plan_dev.h"

namespace gfsm {

// Thread-local last-error message behind gfsm_last_error().
void set_error(const std::string &msg);
const char *last_error();

struct Status {
    gfsm_status code = GFSM_OK;
    std::string msg;
    static Status ok() { return {}; }
    static Status err(gfsm_status c, std::string m) { return {c, std::move(m)}; }
    bool good() const { return code == GFSM_OK; }
};

// Directory entry of the relation-partitioned layout: the entries of vertex x
// with relation `rel` are nbr[beg .. next entry's beg).
struct DirEnt {
    uint32_t rel;
    uint32_t beg;
};

struct DevGraph;  // device mirror, owned by a context

// A graph: the input entries as given (t/v/e order), plus the layout.
//   - entries of each vertex are grouped by relation, ascending, and sorted
//     by target id inside a group (one "segment" per (vertex, relation));
//   - dir_ptr[x] .. dir_ptr[x+1] index x's segment directory;
//   - dir has a sentinel {~0u, nent} so that dir[j+1].beg ends segment j;
//   - rel_keys/rel_off/rel_vtx/rel_beg/rel_cnt list, per relation, the
//     vertices that own a segment of that relation, the segment's begin and
//     its length (root candidate lists).
struct HostGraph {
    uint32_t mode = GFSM_MODE_REF;
    uint32_t V = 0;
    std::vector<uint32_t> vlabel, src, dst, elabel;

    std::vector<uint32_t> nbr;
    std::vector<uint32_t> dir_ptr;
    std::vector<DirEnt> dir;
    std::vector<uint32_t> rel_keys;
    std::vector<uint64_t> rel_off;
    std::vector<uint32_t> rel_vtx, rel_beg, rel_cnt;
    uint32_t max_degree = 0;
    uint64_t nent() const { return nbr.size(); }
    uint64_t ndir() const { return dir.empty() ? 0 : dir.size() - 1; }

    // out-degree of x
    uint32_t degree(uint32_t x) const {
        return dir[dir_ptr[x + 1]].beg - dir[dir_ptr[x]].beg;
    }
    // [beg, end) of the (x, rel) segment; beg == end when absent
    void segment(uint32_t x, uint32_t rel, uint32_t &beg, uint32_t &end) const;
    // index of `rel` in rel_keys, or -1
    int64_t rel_index(uint32_t rel) const;

    mutable std::mutex dev_mu;
    mutable std::unordered_map<const void *, DevGraph *> dev;  // ctx -> mirror
    // Running servers that use this graph (as their data graph or as a query
    // in flight): while pinned, gfsm_graph_free only marks it, and the last
    // unpin frees it (pin_graph / unpin_graph / free_graph_when_unpinned).
    mutable uint32_t pins = 0;
    mutable bool free_pending = false;
    mutable std::mutex stats_mu;
    mutable std::vector<uint64_t> rel_entries;  // per rel_keys index, filled on first use
    mutable std::vector<uint64_t> label_count;  // vertices per vertex label, filled on first use
};

// Vertices that could own a segment of `rel`: those with the relation's
// source label in labelled mode, all vertices otherwise.
uint64_t relation_pool(const HostGraph &g, uint32_t rel);

// Vertices owning a segment of `rel` and the entries of those segments
// (false when the graph has no such relation).  Computed once per graph.
bool relation_sources(const HostGraph &g, uint32_t rel, uint64_t &sources, uint64_t &entries);

// Relation of an entry: REF = edge label; VLABEL = (lu << 21) | (lv << 10) | l.
bool relation_of(uint32_t mode, uint32_t lu, uint32_t lv, uint32_t l, uint32_t &rel);

Status build_graph(uint32_t nv, const uint32_t *vlabel, uint64_t ne, const uint32_t *src,
                   const uint32_t *dst, const uint32_t *elabel, uint32_t mode, HostGraph *&out);
Status load_tve(const char *path, uint32_t mode, HostGraph *&out);
Status write_tve(const HostGraph &g, const char *path);
// Versioned binary image of a built graph (input arrays + layout), checked on load.
Status save_bin(const HostGraph &g, const char *path);
Status load_bin(const char *path, HostGraph *&out);

// Compiled query (matching order + per-depth requirements).
struct Plan {
    uint32_t q = 0;
    uint32_t maxval = 0;
    std::vector<uint32_t> col;              // original query id placed at depth d
    std::vector<uint32_t> header, requirements;  // reference layout (HnGPU.cuh:41-58)
    QueryPlanDev dev;                       // what the kernels read
};
// The reference's matching order, or with `by_cost` (a data graph) the
// order chosen from its relation statistics.
Status plan_query(const HostGraph &query, Plan &out, const HostGraph *by_cost = nullptr);
// Relation model of the data graph: mean segment length of `rel` and the
// number of vertices its segments reach.
void rel_model(const HostGraph &data, uint32_t rel, double &mean, double &reach);
// Expected nodes and candidates per depth (q entries each) of a plan whose
// root range has `nroots` roots and `g1` depth-1 candidates.
void project_levels(const HostGraph &data, const Plan &plan, uint64_t nroots, uint64_t g1, double *nodes,
                    double *cand);

// Workload generators (gen.cpp).
Status gen_er(uint32_t n, double avg_degree, uint32_t nvl, uint32_t nel, uint64_t seed, uint32_t mode,
              HostGraph *&out);
Status gen_rmat(uint32_t scale, uint32_t ef, double a, double b, double c, uint32_t nvl, double zs, uint32_t nel,
                uint64_t seed, uint32_t mode, HostGraph *&out);
Status gen_grid(uint32_t rows, uint32_t cols, double del_frac, double diag_frac, uint64_t seed, uint32_t mode,
                HostGraph *&out);
Status gen_query_walk(const HostGraph &data, uint32_t qsize, uint32_t kind, uint64_t seed, HostGraph *&out);

struct Result {
    gfsm_status status = GFSM_OK;
    uint64_t count = 0;
    uint64_t digest[2] = {0, 0};
    uint32_t width = 0;
    std::vector<uint32_t> rows;
    gfsm_stats stats{};
};

void sort_rows(std::vector<uint32_t> &rows, uint32_t width);
uint64_t mix64(uint64_t z);
void row_hash(const uint32_t *row, uint32_t width, uint64_t h[2]);

// Parallel helpers.
unsigned hw_threads();
template <class F>
void parallel_for(uint64_t n, F &&f, unsigned threads = 0);

}  // namespace gfsm

struct gfsm_graph {
    gfsm::HostGraph *g;
};
struct gfsm_result {
    gfsm::Result r;
};

#include "parallel.h"
