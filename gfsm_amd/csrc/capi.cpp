// capi.cpp -- the extern "C" surface declared in include/gfsm.h.  Converts
// internal Status values into status codes plus a thread-local message; no
// entry point exits the process (the reference exits from inside the parser
// and launcher: HnHelper.cuh:293, 313, 348; HnGPU.cu:6-10, 918).
#include <cstring>
#include <new>

#include "engine.h"
#include "serve.h"

using namespace gfsm;

static gfsm_status fail(const Status &st)
{
    set_error(st.msg);
    return st.code;
}

static gfsm_status ok()
{
    set_error("");
    return GFSM_OK;
}

#define GUARD_BEGIN try {
#define GUARD_END                                                              \
    }                                                                          \
    catch (const std::bad_alloc &) {                                           \
        return fail(Status::err(GFSM_ERR_NOMEM, "host allocation failed"));    \
    }                                                                          \
    catch (const std::exception &e) {                                          \
        return fail(Status::err(GFSM_ERR_ARG, e.what()));                      \
    }

extern "C" {

const char *gfsm_last_error(void) { return last_error(); }
const char *gfsm_version(void) { return "gfsm-mi355x 0.1 (gfx950)"; }

gfsm_status gfsm_open(int device, uint64_t mem_budget_bytes, gfsm_ctx **ctx)
{
    GUARD_BEGIN
    if (!ctx) return fail(Status::err(GFSM_ERR_ARG, "null ctx"));
    *ctx = nullptr;
    Ctx *c = nullptr;
    Status st = ctx_open(device, mem_budget_bytes, c);
    if (!st.good()) return fail(st);
    *ctx = new gfsm_ctx{c};
    return ok();
    GUARD_END
}

void gfsm_close(gfsm_ctx *ctx)
{
    if (!ctx) return;
    ctx_close(ctx->c);
    delete ctx;
}

gfsm_status gfsm_graph_load_tve(const char *path, uint32_t mode, gfsm_graph **g)
{
    GUARD_BEGIN
    if (!path || !g) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    *g = nullptr;
    HostGraph *h = nullptr;
    Status st = load_tve(path, mode, h);
    if (!st.good()) return fail(st);
    *g = new gfsm_graph{h};
    return ok();
    GUARD_END
}

gfsm_status gfsm_graph_from_edges(uint32_t nv, const uint32_t *vlabel, uint64_t ne, const uint32_t *src,
                                  const uint32_t *dst, const uint32_t *elabel, uint32_t mode, gfsm_graph **g)
{
    GUARD_BEGIN
    if (!g || (ne && (!src || !dst || !elabel))) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    *g = nullptr;
    HostGraph *h = nullptr;
    Status st = build_graph(nv, vlabel, ne, src, dst, elabel, mode, h);
    if (!st.good()) return fail(st);
    *g = new gfsm_graph{h};
    return ok();
    GUARD_END
}

gfsm_status gfsm_graph_write_tve(const gfsm_graph *g, const char *path)
{
    GUARD_BEGIN
    if (!g || !path) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    Status st = write_tve(*g->g, path);
    return st.good() ? ok() : fail(st);
    GUARD_END
}

gfsm_status gfsm_graph_save_bin(const gfsm_graph *g, const char *path)
{
    GUARD_BEGIN
    if (!g || !path) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    Status st = save_bin(*g->g, path);
    return st.good() ? ok() : fail(st);
    GUARD_END
}

gfsm_status gfsm_graph_load_bin(const char *path, gfsm_graph **g)
{
    GUARD_BEGIN
    if (!path || !g) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    *g = nullptr;
    HostGraph *h = nullptr;
    Status st = load_bin(path, h);
    if (!st.good()) return fail(st);
    *g = new gfsm_graph{h};
    return ok();
    GUARD_END
}

gfsm_status gfsm_graph_edges(const gfsm_graph *g, uint32_t *nv, uint64_t *ne, const uint32_t **vlabel,
                             const uint32_t **src, const uint32_t **dst, const uint32_t **elabel)
{
    if (!g) return fail(Status::err(GFSM_ERR_ARG, "null graph"));
    const HostGraph &h = *g->g;
    if (nv) *nv = h.V;
    if (ne) *ne = h.src.size();
    if (vlabel) *vlabel = h.vlabel.data();
    if (src) *src = h.src.data();
    if (dst) *dst = h.dst.data();
    if (elabel) *elabel = h.elabel.data();
    return ok();
}

gfsm_status gfsm_graph_layout(const gfsm_graph *g, uint32_t *nv, uint64_t *ndir, uint64_t *nent,
                              const uint32_t **dir_ptr, const uint32_t **dir, const uint32_t **nbr)
{
    if (!g) return fail(Status::err(GFSM_ERR_ARG, "null graph"));
    const HostGraph &h = *g->g;
    if (nv) *nv = h.V;
    if (ndir) *ndir = h.ndir();
    if (nent) *nent = h.nent();
    if (dir_ptr) *dir_ptr = h.dir_ptr.data();
    if (dir) *dir = reinterpret_cast<const uint32_t *>(h.dir.data());
    if (nbr) *nbr = h.nbr.data();
    return ok();
}

gfsm_status gfsm_graph_info(const gfsm_graph *g, uint32_t *nv, uint64_t *nent, uint32_t *max_degree,
                            uint32_t *mode)
{
    if (!g) return fail(Status::err(GFSM_ERR_ARG, "null graph"));
    if (nv) *nv = g->g->V;
    if (nent) *nent = g->g->nent();
    if (max_degree) *max_degree = g->g->max_degree;
    if (mode) *mode = g->g->mode;
    return ok();
}

gfsm_status gfsm_graph_upload(gfsm_ctx *ctx, const gfsm_graph *g)
{
    GUARD_BEGIN
    if (!ctx || !g) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    std::lock_guard<std::mutex> lk(ctx->c->mu);
    DevGraph *dg = nullptr;
    Status st = upload_graph(*ctx->c, *g->g, dg);
    return st.good() ? ok() : fail(st);
    GUARD_END
}

void gfsm_graph_free(gfsm_graph *g)
{
    if (!g) return;
    free_graph_when_unpinned(g->g);  // a running server may still use it
    delete g;
}

gfsm_status gfsm_plan_query(const gfsm_graph *query, uint32_t *q, uint32_t *maxval, uint32_t *col,
                            uint32_t *header, uint32_t *requirements)
{
    GUARD_BEGIN
    if (!query) return fail(Status::err(GFSM_ERR_ARG, "null query"));
    Plan p;
    Status st = plan_query(*query->g, p);
    if (!st.good()) return fail(st);
    if (q) *q = p.q;
    if (maxval) *maxval = p.maxval;
    if (col) std::memcpy(col, p.col.data(), p.col.size() * sizeof(uint32_t));
    if (header) std::memcpy(header, p.header.data(), p.header.size() * sizeof(uint32_t));
    if (requirements)
        std::memcpy(requirements, p.requirements.data(), (size_t)2 * p.maxval * p.q * sizeof(uint32_t));
    return ok();
    GUARD_END
}

gfsm_status gfsm_plan_order(const gfsm_graph *data, const gfsm_graph *query, uint32_t flags, uint32_t *col)
{
    GUARD_BEGIN
    if (!data || !query || !col) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    Plan p;
    Status st = plan_query(*query->g, p, (flags & GFSM_ORDER_COST) ? data->g : nullptr);
    if (!st.good()) return fail(st);
    std::memcpy(col, p.col.data(), p.col.size() * sizeof(uint32_t));
    return ok();
    GUARD_END
}

gfsm_status gfsm_query_projection(const gfsm_graph *data, const gfsm_graph *query, const gfsm_match_opts *opts,
                                  double *nodes, double *cand)
{
    GUARD_BEGIN
    if (!data || !query || !nodes || !cand) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    gfsm_match_opts o{};
    if (opts) o = *opts;
    Plan p;
    Status st = plan_for(*data->g, *query->g, o, p);
    if (!st.good()) return fail(st);
    if (p.q < 2) return fail(Status::err(GFSM_ERR_ARG, "a one-vertex query has no levels to project"));
    uint64_t lo, hi, g1 = 0;
    root_range(*data->g, p, o.root_lo, o.root_hi, lo, hi);
    for (uint64_t i = lo; i < hi; i++) g1 += data->g->rel_cnt[i];
    std::vector<double> nd(p.q), cd(p.q);
    project_levels(*data->g, p, hi - lo, g1, nd.data(), cd.data());
    // reported by original query depth order: entry k is depth k
    std::memcpy(nodes, nd.data(), p.q * sizeof(double));
    std::memcpy(cand, cd.data(), p.q * sizeof(double));
    return ok();
    GUARD_END
}

gfsm_status gfsm_query_roots(const gfsm_graph *data, const gfsm_graph *query, uint32_t *ids, uint64_t *n)
{
    GUARD_BEGIN
    if (!data || !query || !n) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    Plan p;
    Status st = plan_query(*query->g, p);
    if (!st.good()) return fail(st);
    const HostGraph &g = *data->g;
    if (p.q == 1) {
        uint64_t k = 0;
        for (uint32_t x = 0; x < g.V; x++)
            if (g.degree(x)) {
                if (ids) ids[k] = x;
                k++;
            }
        *n = k;
        return ok();
    }
    uint64_t lo, hi;
    root_range(g, p, 0, 0, lo, hi);
    *n = hi - lo;
    if (ids && hi > lo) std::memcpy(ids, g.rel_vtx.data() + lo, (hi - lo) * sizeof(uint32_t));
    return ok();
    GUARD_END
}

gfsm_status gfsm_query_window(const gfsm_graph *data, const gfsm_graph *query, uint64_t max_first,
                              uint32_t *root_lo, uint32_t *root_hi, uint64_t *first)
{
    GUARD_BEGIN
    if (!data || !query || !root_lo || !root_hi) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    Plan p;
    Status st = plan_query(*query->g, p);
    if (!st.good()) return fail(st);
    const HostGraph &g = *data->g;
    uint64_t lo, hi, f = 0;
    *root_lo = *root_hi = g.V;  // empty window
    if (p.q > 1) {
        root_range(g, p, 0, 0, lo, hi);
        // Walk down from the highest root id while the first-depth
        // candidate total stays within max_first (at least one root).
        uint64_t k = hi;
        while (k > lo) {
            const uint64_t c = g.rel_cnt[k - 1];
            if (k < hi && f + c > max_first) break;
            f += c;
            k--;
        }
        if (k < hi) *root_lo = g.rel_vtx[k];
    } else {  // one vertex: every vertex with a segment is a root
        uint32_t x = g.V;
        while (x > 0) {
            const uint32_t y = x - 1;
            if (g.degree(y)) {
                if (f > 0 && f >= max_first) break;
                f++;
            }
            x = y;
        }
        *root_lo = x;
    }
    if (first) *first = f;
    return ok();
    GUARD_END
}

gfsm_status gfsm_match(gfsm_ctx *ctx, const gfsm_graph *data, const gfsm_graph *query, const gfsm_match_opts *opts,
                       gfsm_result **r)
{
    GUARD_BEGIN
    if (!ctx || !data || !query || !r) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    *r = nullptr;
    gfsm_match_opts o{};
    if (opts) o = *opts;
    std::unique_ptr<gfsm_result> res(new gfsm_result);
    Status st = match_plain(*ctx->c, *data->g, *query->g, o, res->r);
    if (st.code == GFSM_ERR_CAP && res->r.status == GFSM_ERR_CAP) {
        *r = res.release();  // partial result under the work cap
        return fail(st);
    }
    if (!st.good()) return fail(st);
    *r = res.release();
    return ok();
    GUARD_END
}

gfsm_status gfsm_match_noalloc(gfsm_ctx *ctx, const gfsm_graph *data, const gfsm_graph *query,
                               const gfsm_match_opts *opts, uint64_t *count, uint64_t digest[2])
{
    static_assert(GFSM_NOALLOC_SCRATCH_BYTES == kNoAllocScratchBytes, "scratch size in gfsm.h");
    GUARD_BEGIN
    if (!ctx || !data || !query || !count || !digest) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    gfsm_match_opts o{};
    if (opts) o = *opts;
    Result res;
    Status st = match_noalloc(*ctx->c, *data->g, *query->g, o, res);
    if (st.good() || (st.code == GFSM_ERR_CAP && res.status == GFSM_ERR_CAP)) {
        *count = res.count;
        digest[0] = res.digest[0];
        digest[1] = res.digest[1];
    }
    return st.good() ? ok() : fail(st);
    GUARD_END
}

gfsm_status gfsm_match_write_fs(gfsm_ctx *ctx, const gfsm_graph *data, const gfsm_graph *query,
                                const gfsm_match_opts *opts, const char *path, gfsm_result **r)
{
    GUARD_BEGIN
    if (!ctx || !data || !query || !path || !r) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    *r = nullptr;
    gfsm_match_opts o{};
    if (opts) o = *opts;
    o.flags |= GFSM_ROWS;
    std::unique_ptr<FILE, int (*)(FILE *)> f(std::fopen(path, "wb"), std::fclose);
    if (!f) return fail(Status::err(GFSM_ERR_MISSING_FILE, std::string("cannot write ") + path));
    uint32_t hdr[2] = {0, 0};  // rewritten once the count is known
    if (std::fwrite(hdr, sizeof(uint32_t), 2, f.get()) != 2)
        return fail(Status::err(GFSM_ERR_MISSING_FILE, "short write to output.fs"));
    std::unique_ptr<gfsm_result> res(new gfsm_result);
    // Sorted rows need the whole table; otherwise every chunk is written as it lands.
    Status st = match_plain(*ctx->c, *data->g, *query->g, o, res->r, (o.flags & GFSM_SORTED) ? nullptr : f.get());
    const bool partial = st.code == GFSM_ERR_CAP && res->r.status == GFSM_ERR_CAP;
    if (!st.good() && !partial) return fail(st);
    Result &m = res->r;
    const std::vector<uint32_t> &rows = m.rows;  // one-vertex queries and sorted rows
    if (!rows.empty() && std::fwrite(rows.data(), sizeof(uint32_t), rows.size(), f.get()) != rows.size())
        return fail(Status::err(GFSM_ERR_MISSING_FILE, "short write to output.fs"));
    if (m.count > 0xFFFFFFFFull) return fail(Status::err(GFSM_ERR_CAP, "output.fs holds at most 2^32-1 rows"));
    hdr[0] = (uint32_t)m.count;
    hdr[1] = m.width;
    if (std::fseek(f.get(), 0, SEEK_SET) != 0 || std::fwrite(hdr, sizeof(uint32_t), 2, f.get()) != 2 ||
        std::fflush(f.get()) != 0)
        return fail(Status::err(GFSM_ERR_MISSING_FILE, "short write to output.fs"));
    std::vector<uint32_t>().swap(m.rows);  // the rows live in the file
    *r = res.release();
    return partial ? fail(st) : ok();
    GUARD_END
}

gfsm_status gfsm_result_count(const gfsm_result *r, uint64_t *n)
{
    if (!r || !n) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    *n = r->r.count;
    return ok();
}

gfsm_status gfsm_result_rows(const gfsm_result *r, const uint32_t **rows, uint32_t *width, uint64_t *nrows)
{
    if (!r) return fail(Status::err(GFSM_ERR_ARG, "null result"));
    if (rows) *rows = r->r.rows.data();
    if (width) *width = r->r.width;
    if (nrows) *nrows = r->r.width ? r->r.rows.size() / r->r.width : 0;
    return ok();
}

gfsm_status gfsm_result_digest(const gfsm_result *r, uint64_t out[2])
{
    if (!r || !out) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    out[0] = r->r.digest[0];
    out[1] = r->r.digest[1];
    return ok();
}

gfsm_status gfsm_result_stats(const gfsm_result *r, gfsm_stats *out)
{
    if (!r || !out) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    *out = r->r.stats;
    return ok();
}

gfsm_status gfsm_result_status(const gfsm_result *r)
{
    if (!r) return GFSM_ERR_ARG;
    return r->r.status;
}

gfsm_status gfsm_result_write_fs(const gfsm_result *r, const char *path)
{
    if (!r || !path) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    const Result &res = r->r;
    uint64_t n = res.width ? res.rows.size() / res.width : 0;
    if (n != res.count) return fail(Status::err(GFSM_ERR_ARG, "result has no materialised rows (use GFSM_ROWS)"));
    if (n > 0xFFFFFFFFull) return fail(Status::err(GFSM_ERR_CAP, "output.fs holds at most 2^32-1 rows"));
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(Status::err(GFSM_ERR_MISSING_FILE, std::string("cannot write ") + path));
    uint32_t hdr[2] = {(uint32_t)n, res.width};
    bool good = std::fwrite(hdr, sizeof(uint32_t), 2, f) == 2;
    if (n) good = good && std::fwrite(res.rows.data(), sizeof(uint32_t), res.rows.size(), f) == res.rows.size();
    std::fclose(f);
    return good ? ok() : fail(Status::err(GFSM_ERR_MISSING_FILE, "short write"));
}

void gfsm_result_free(gfsm_result *r) { delete r; }

void gfsm_row_hash(const uint32_t *row, uint32_t width, uint64_t h[2]) { row_hash(row, width, h); }

gfsm_status gfsm_match_batch(gfsm_ctx *ctx, const gfsm_graph *data, const gfsm_graph *const *queries, uint32_t n,
                             const gfsm_match_opts *opts, uint32_t nopts, gfsm_result **results)
{
    GUARD_BEGIN
    if (!ctx || !data || (n && (!queries || !results))) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    if (opts && nopts != 1 && nopts != n) return fail(Status::err(GFSM_ERR_ARG, "nopts must be 1 or n"));
    std::vector<gfsm_match_opts> os;
    if (opts) os.assign(opts, opts + nopts);
    else os.assign(1, gfsm_match_opts{});
    std::vector<const HostGraph *> qs(n);
    for (uint32_t i = 0; i < n; i++) {
        if (!queries[i]) return fail(Status::err(GFSM_ERR_ARG, "null query"));
        qs[i] = queries[i]->g;
        results[i] = nullptr;
    }
    std::vector<Result> rs;
    Status st = match_batch(*ctx->c, *data->g, qs, os, rs);
    if (!st.good()) return fail(st);
    for (uint32_t i = 0; i < n; i++) {
        results[i] = new gfsm_result;
        results[i]->r = std::move(rs[i]);
    }
    return ok();
    GUARD_END
}

gfsm_status gfsm_tuning_get(gfsm_ctx *ctx, gfsm_tuning *t)
{
    GUARD_BEGIN
    if (!ctx || !t) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    const Tuning &u = ctx->c->tune;
    for (int k = 0; k < 3; k++)
        for (uint32_t d = 0; d < GFSM_MAX_Q; d++) t->items[k][d] = u.items[k][d];
    t->pool_words = u.pool_words;
    t->wgs_per_cu = u.wgs_per_cu;
    t->batch_items = u.batch_items;
    t->batch_pool_words = u.batch_pool_words;
    t->batch_wgs_per_cu = u.batch_wgs_per_cu;
    t->batch_slots = u.batch_slots;
    t->batch_steps = u.batch_steps;
    t->serve_steps = u.serve_steps;
    t->batch_step_budget = u.batch_step_budget;
    return ok();
    GUARD_END
}

static bool valid_items(uint32_t v) { return v == 1 || v == 2 || v == 4 || v == 8; }

gfsm_status gfsm_tuning_set(gfsm_ctx *ctx, const gfsm_tuning *t)
{
    GUARD_BEGIN
    if (!ctx || !t) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    for (int k = 0; k < 3; k++)
        for (uint32_t d = 0; d < GFSM_MAX_Q; d++)
            if (!valid_items(t->items[k][d]) || expand_lds_bytes(t->items[k][d], t->pool_words) > kMaxDynLds)
                return fail(Status::err(GFSM_ERR_ARG, "items must be 1, 2, 4 or 8 and fit the LDS with the pool"));
    if (!valid_items(t->batch_items) || expand_lds_bytes(t->batch_items, t->batch_pool_words) > kMaxDynLds)
        return fail(Status::err(GFSM_ERR_ARG, "batch items must be 1, 2, 4 or 8 and fit the LDS with the pool"));
    if (t->wgs_per_cu < 1 || t->wgs_per_cu > 32 || t->batch_wgs_per_cu < 1 || t->batch_wgs_per_cu > 32)
        return fail(Status::err(GFSM_ERR_ARG, "workgroups per CU must be 1..32"));
    if (t->batch_slots < 1 || t->batch_slots > kMaxSlots)
        return fail(Status::err(GFSM_ERR_ARG, "batch slots must be 1..1024"));
    if (t->serve_steps > 64) return fail(Status::err(GFSM_ERR_ARG, "serve_steps must be 0 (default 2) .. 64"));
    if (t->batch_steps < 1 || t->batch_steps > 64)
        return fail(Status::err(GFSM_ERR_ARG, "batch steps per replay must be 1..64"));
    if (t->batch_step_budget < (1ull << 16))
        return fail(Status::err(GFSM_ERR_ARG, "batch step budget must be at least 65536 candidates"));
    std::lock_guard<std::mutex> lk(ctx->c->mu);
    Tuning &u = ctx->c->tune;
    for (int k = 0; k < 3; k++)
        for (uint32_t d = 0; d < GFSM_MAX_Q; d++) u.items[k][d] = (uint8_t)t->items[k][d];
    u.pool_words = t->pool_words;
    u.wgs_per_cu = t->wgs_per_cu;
    u.batch_items = t->batch_items;
    u.batch_pool_words = t->batch_pool_words;
    u.batch_wgs_per_cu = t->batch_wgs_per_cu;
    u.batch_slots = t->batch_slots;
    u.batch_steps = t->batch_steps;
    u.serve_steps = t->serve_steps ? t->serve_steps : 2;
    u.batch_step_budget = t->batch_step_budget;
    return ok();
    GUARD_END
}

gfsm_status gfsm_autotune(gfsm_ctx *ctx, const gfsm_graph *data, const gfsm_graph *const *queries, uint32_t n,
                          const gfsm_match_opts *opts, uint32_t nopts, uint32_t which, gfsm_tuning *t)
{
    GUARD_BEGIN
    if (!ctx || !data || !n || !queries) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    if (opts && nopts != 1 && nopts != n) return fail(Status::err(GFSM_ERR_ARG, "nopts must be 1 or n"));
    std::vector<const HostGraph *> qs(n);
    std::vector<gfsm_match_opts> os(n);
    for (uint32_t i = 0; i < n; i++) {
        if (!queries[i]) return fail(Status::err(GFSM_ERR_ARG, "null query"));
        qs[i] = queries[i]->g;
        if (opts) os[i] = opts[nopts == 1 ? 0 : i];
        os[i].flags &= ~(GFSM_ROWS | GFSM_SORTED | GFSM_PROFILE);  // count mode
        os[i].tile = 0;
    }
    Status st = autotune(*ctx->c, *data->g, qs, os, which);
    if (!st.good()) return fail(st);
    if (t) return gfsm_tuning_get(ctx, t);
    return ok();
    GUARD_END
}

gfsm_status gfsm_server_start(gfsm_ctx *ctx, const gfsm_graph *data, const gfsm_serve_opts *opts, gfsm_server **srv)
{
    GUARD_BEGIN
    if (!ctx || !data || !srv) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    gfsm_serve_opts o{};
    if (opts) o = *opts;
    Server *s = nullptr;
    Status st = server_start(*ctx->c, *data->g, o, s);
    if (!st.good()) return fail(st);
    *srv = reinterpret_cast<gfsm_server *>(s);
    return ok();
    GUARD_END
}

gfsm_status gfsm_server_submit(gfsm_server *srv, const gfsm_graph *query, const gfsm_match_opts *opts,
                               uint64_t *ticket)
{
    GUARD_BEGIN
    if (!srv || !query || !ticket) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    Status st = server_submit(*reinterpret_cast<Server *>(srv), *query->g, opts, *ticket);
    return st.good() ? ok() : fail(st);
    GUARD_END
}

static gfsm_status take_result(Server &s, uint64_t ticket, bool wait, gfsm_result **r)
{
    Result res;
    Status st = server_take(s, ticket, wait, res);
    if (st.code == GFSM_ERR_PENDING) return GFSM_ERR_PENDING;
    if (!st.good()) return fail(st);
    *r = new gfsm_result;
    (*r)->r = std::move(res);
    return ok();
}

gfsm_status gfsm_server_poll(gfsm_server *srv, uint64_t ticket, gfsm_result **r)
{
    GUARD_BEGIN
    if (!srv || !r) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    *r = nullptr;
    return take_result(*reinterpret_cast<Server *>(srv), ticket, false, r);
    GUARD_END
}

gfsm_status gfsm_server_wait(gfsm_server *srv, uint64_t ticket, gfsm_result **r)
{
    GUARD_BEGIN
    if (!srv || !r) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    *r = nullptr;
    return take_result(*reinterpret_cast<Server *>(srv), ticket, true, r);
    GUARD_END
}

gfsm_status gfsm_server_times(gfsm_server *srv, uint64_t ticket, double t[3])
{
    if (!srv || !t) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    Status st = server_times(*reinterpret_cast<Server *>(srv), ticket, t);
    return st.good() ? ok() : fail(st);
}

gfsm_status gfsm_server_stats(gfsm_server *srv, gfsm_serve_stats *out)
{
    if (!srv || !out) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    Status st = server_stats(*reinterpret_cast<Server *>(srv), *out);
    return st.good() ? ok() : fail(st);
}

void gfsm_server_stop(gfsm_server *srv)
{
    if (srv) server_stop(reinterpret_cast<Server *>(srv));
}

gfsm_status gfsm_selftest_scan(gfsm_ctx *ctx, const uint32_t *vals, uint64_t n, uint64_t *prefix, uint32_t *idx,
                               uint64_t totals[2])
{
    GUARD_BEGIN
    if (!ctx || (n && (!vals || !prefix || !idx)) || !totals) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    Status st = selftest_scan(*ctx->c, vals, n, prefix, idx, totals);
    return st.good() ? ok() : fail(st);
    GUARD_END
}

gfsm_status gfsm_selftest_traffic(gfsm_ctx *ctx, uint32_t mode, uint64_t n, uint32_t launches, uint64_t bytes[2])
{
    GUARD_BEGIN
    if (!ctx || !bytes) return fail(Status::err(GFSM_ERR_ARG, "null argument"));
    Status st = selftest_traffic(*ctx->c, mode, n, launches, bytes);
    return st.good() ? ok() : fail(st);
    GUARD_END
}

#define GEN_RETURN(call)                           \
    GUARD_BEGIN                                    \
    if (!g) return fail(Status::err(GFSM_ERR_ARG, "null argument")); \
    *g = nullptr;                                  \
    HostGraph *h = nullptr;                        \
    Status st = (call);                            \
    if (!st.good()) return fail(st);               \
    *g = new gfsm_graph{h};                        \
    return ok();                                   \
    GUARD_END

gfsm_status gfsm_gen_er(uint32_t n, double avg_degree, uint32_t nvlabels, uint32_t nelabels, uint64_t seed,
                        uint32_t mode, gfsm_graph **g)
{
    GEN_RETURN(gen_er(n, avg_degree, nvlabels, nelabels, seed, mode, h))
}

gfsm_status gfsm_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint32_t nvlabels,
                          double zipf_s, uint32_t nelabels, uint64_t seed, uint32_t mode, gfsm_graph **g)
{
    GEN_RETURN(gen_rmat(scale, edge_factor, a, b, c, nvlabels, zipf_s, nelabels, seed, mode, h))
}

gfsm_status gfsm_gen_grid(uint32_t rows, uint32_t cols, double del_frac, double diag_frac, uint64_t seed,
                          uint32_t mode, gfsm_graph **g)
{
    GEN_RETURN(gen_grid(rows, cols, del_frac, diag_frac, seed, mode, h))
}

gfsm_status gfsm_gen_query_walk(const gfsm_graph *data, uint32_t qsize, uint32_t kind, uint64_t seed,
                                gfsm_graph **q)
{
    gfsm_graph **g = q;
    if (!data) return fail(Status::err(GFSM_ERR_ARG, "null data graph"));
    GEN_RETURN(gen_query_walk(*data->g, qsize, kind, seed, h))
}

}  // extern "C"
