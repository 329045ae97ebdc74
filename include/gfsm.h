/*
 * gfsm.h -- C ABI of libgfsm.so, the MI355X-native forest-based subgraph
 * matcher.  Plain pointers and sizes only; every call returns a gfsm_status
 * and never exits the process.
 *
 * What each entry point replaces in the reference (paths relative to the
 * reference repository root).  The reference has no C/FFI surface: its
 * boundary is the process `GFSM <query> <data>` (main.cu:56-63) plus the C++
 * namespace HnSetup (HnGPU.cuh:60-80), so each entry point below names the
 * reference code whose job it takes over:
 *
 *   gfsm_open / gfsm_close      HnSetup::preinit (HnGPU.cu:847-862) + the device
 *                               reset/warm-up of argTest (main.cu:11-20)
 *   gfsm_graph_load_tve         fileParse (HnHelper.cuh:283-499)
 *   gfsm_graph_save/load_bin    (no counterpart) binary graph image that skips
 *                               fileParse's parse + layout on later runs
 *   gfsm_graph_from_edges       fileParse's row staging + CCSR layout
 *                               (HnHelper.cuh:364-478) from in-memory arrays
 *   gfsm_plan_query             HnSetup::preProcessQuery (HnGPU.cu:749-816) +
 *                               CCSR::genereateCCSRRequirements (ccsr.cpp:98-204)
 *   gfsm_match                  HnSetup::solve -> gpuLaunch -> hnSolve
 *                               (HnGPU.cu:710-739, 867-1063, 455-669), with the
 *                               overflow retry (HnGPU.cu:957-992) replaced by
 *                               chunked expansion
 *   gfsm_result_count           the "Function Count" CSV field (HnGPU.cu:999-1000)
 *   gfsm_result_rows            hnWrite + D2H of the table (HnGPU.cu:674-699, 1019-1048)
 *   gfsm_result_write_fs        HnSetup::functionsToFile (HnGPU.cu:1081-1091)
 *   gfsm_result_digest          (no counterpart) order-independent row digest
 *   gfsm_match_batch            (no counterpart) many queries, one persistent launch per depth
 *   gfsm_server_*               (no counterpart) resident-graph serving loop
 *   gfsm_gen_*                  (no counterpart) seeded workload generators
 *
 * Matching semantics (identical to the reference on valid inputs): an
 * embedding is an injective map f from query vertices to data vertices such
 * that, when query vertex n is placed, every query entry n->u to an already
 * placed u has a data entry f(u)->f(n) with the relation of u->n (or of n->u
 * when the query has no u->n entry).  Non-induced; automorphic images count
 * separately.  GFSM_MODE_REF: relation = edge label, vertex labels ignored.
 * GFSM_MODE_VLABEL: relation of u->v = (label(u), label(v), edge label).
 *
 * Threading: a context is used by one host thread at a time; a server is
 * safe for concurrent submit/poll.  Ownership: every handle is released by
 * its own *_free / *_close / *_stop.
 */
#ifndef GFSM_H
#define GFSM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t gfsm_status;
/* Status codes; 1 and 232-234 are the reference's own exit codes. */
#define GFSM_OK 0
#define GFSM_ERR_QUERY 1          /* disconnected or invalid query (ccsr.cpp:126-128, 192-196) */
#define GFSM_ERR_ARG 2            /* bad argument / out-of-domain input                      */
#define GFSM_ERR_NOMEM 3          /* device or host allocation failed, or budget too small   */
#define GFSM_ERR_DEVICE 4         /* HIP runtime error                                       */
#define GFSM_ERR_CAP 5            /* a cap was reached (max_rows, max_cand, max_depth_cand)  */
#define GFSM_ERR_PENDING 6        /* server: result not ready yet                            */
#define GFSM_ERR_MISSING_FILE 232 /* HnHelper.cuh:16 */
#define GFSM_ERR_INVALID_DATA 233 /* HnHelper.cuh:17 */
#define GFSM_ERR_MISSING_DATA 234 /* HnHelper.cuh:18 */

#define GFSM_MODE_REF 0u
#define GFSM_MODE_VLABEL 1u

/* gfsm_match_opts.flags */
#define GFSM_ROWS 1u     /* materialise the embedding table on the host           */
#define GFSM_SORTED 2u   /* ... sorted lexicographically by (f(0), ..., f(q-1))    */
#define GFSM_DIGEST 4u   /* compute the order-independent multiset digest         */
#define GFSM_PROFILE 8u  /* time every kernel launch with HIP events (stats)      */
#define GFSM_ORDER_COST 16u /* matching order chosen from the data graph's relation  */
                            /* statistics instead of the reference's (no reference   */
                            /* counterpart; the same embeddings, in other row order);*/
                            /* root_lo/root_hi are then rejected (GFSM_ERR_ARG)       */

#define GFSM_MAX_Q 32u   /* largest query handled (vertices)                      */

typedef struct gfsm_ctx gfsm_ctx;
typedef struct gfsm_graph gfsm_graph;
typedef struct gfsm_result gfsm_result;
typedef struct gfsm_server gfsm_server;

typedef struct gfsm_match_opts {
    uint32_t flags;
    uint32_t root_lo, root_hi;  /* only embeddings with f(root) in [root_lo, root_hi); 0,0 = all */
    uint32_t tile;              /* candidates per workgroup tile (0 = default)                    */
    uint64_t max_rows;          /* cap on materialised rows (0 = no cap)                          */
    uint64_t mem_budget;        /* device bytes for intermediate levels (0 = context default)     */
    uint64_t max_cand;          /* stop after this many candidates (0 = no cap): the call returns */
                                /* GFSM_ERR_CAP and a partial result whose status says so; where  */
                                /* it stops depends on how the work was chunked                   */
    uint64_t max_depth_cand;    /* deterministic cap (0 = none): every depth (roots included)      */
                                /* tests only its first max_depth_cand candidates in enumeration   */
                                /* order (root id, then segment order at every depth); the match   */
                                /* stops at the first depth that runs out.  The partial count and  */
                                /* digest are the same on every path and chunking; GFSM_ERR_CAP    */
} gfsm_match_opts;

typedef struct gfsm_stats {
    uint32_t q;                 /* query vertices                                            */
    uint32_t launches;          /* matcher kernel launches                                   */
    uint32_t chunks;            /* expand launches beyond one per depth (chunked expansion)  */
    uint32_t overflows;         /* launches discarded because a level overflowed its buffer  */
    uint64_t roots;             /* root candidates read                                      */
    uint64_t level[GFSM_MAX_Q]; /* nodes kept per depth (last depth: embeddings)             */
    uint64_t cand[GFSM_MAX_Q];  /* candidates tested per depth (depth 0: root candidates)     */
    uint32_t checks[GFSM_MAX_Q];/* membership probes per candidate at each depth             */
    uint64_t alg_bytes;         /* algorithmic HBM bytes: 4 per root candidate; per depth k, */
                                /* 8 per frontier node, 4 per candidate and per probe, 8 per */
                                /* kept inner node; rows mode adds 4 q per embedding         */
    double device_ms;           /* first matcher kernel -> result ready (HIP events)         */
    double kernel_ms;           /* sum of matcher kernel durations (GFSM_PROFILE only)       */
    double expand_ms;           /* sum of expand-kernel durations (GFSM_PROFILE only)        */
    uint32_t expand_launches;   /* expand-kernel launches timed in expand_ms                 */
    uint32_t pad;
    double host_ms;             /* wall time of the call                                     */
} gfsm_stats;

/* ---- context ------------------------------------------------------------ */
gfsm_status gfsm_open(int device, uint64_t mem_budget_bytes, gfsm_ctx **ctx);
void        gfsm_close(gfsm_ctx *ctx);
/* Last error message of the calling thread ("" when none). */
const char *gfsm_last_error(void);
/* Library version string. */
const char *gfsm_version(void);

/* ---- graphs (host objects; uploaded to a context on first use, then resident) */
gfsm_status gfsm_graph_load_tve(const char *path, uint32_t mode, gfsm_graph **g);
gfsm_status gfsm_graph_from_edges(uint32_t nv, const uint32_t *vlabel, uint64_t ne,
                                  const uint32_t *src, const uint32_t *dst, const uint32_t *elabel,
                                  uint32_t mode, gfsm_graph **g);
gfsm_status gfsm_graph_write_tve(const gfsm_graph *g, const char *path);
/* Versioned binary image of a built graph: its input arrays and the
 * relation-partitioned layout, so a later run skips the text parse and the
 * layout build that the reference repeats on every start (fileParse,
 * HnHelper.cuh:283-499, parses and builds its CCSR).  No reference
 * counterpart.  Loading checks the magic, version, sizes, a payload checksum
 * and the layout invariants: 232 unreadable file, 233 anything else wrong. */
gfsm_status gfsm_graph_save_bin(const gfsm_graph *g, const char *path);
gfsm_status gfsm_graph_load_bin(const char *path, gfsm_graph **g);
/* Views of the graph's input arrays (valid while g lives). */
gfsm_status gfsm_graph_edges(const gfsm_graph *g, uint32_t *nv, uint64_t *ne, const uint32_t **vlabel,
                             const uint32_t **src, const uint32_t **dst, const uint32_t **elabel);
/* Views of the relation-partitioned layout: dir_ptr[nv+1]; dir[2*(ndir+1)] as
 * {relation, begin} pairs; nbr[nent]. */
gfsm_status gfsm_graph_layout(const gfsm_graph *g, uint32_t *nv, uint64_t *ndir, uint64_t *nent,
                              const uint32_t **dir_ptr, const uint32_t **dir, const uint32_t **nbr);
gfsm_status gfsm_graph_info(const gfsm_graph *g, uint32_t *nv, uint64_t *nent, uint32_t *max_degree,
                            uint32_t *mode);
/* Pre-upload a data graph to a context (optional; gfsm_match does it lazily). */
gfsm_status gfsm_graph_upload(gfsm_ctx *ctx, const gfsm_graph *g);
void        gfsm_graph_free(gfsm_graph *g);

/* ---- query planning (host only) -------------------------------------------
 * Writes the reference's requirement layout for the query: col[q] = original
 * query id placed at depth d; header[3q] = {num, offset, firstQ};
 * requirements[2*maxval*q] = {delta distance, relation}.  Pass NULL arrays to
 * learn q and maxval first. */
gfsm_status gfsm_plan_query(const gfsm_graph *query, uint32_t *q, uint32_t *maxval, uint32_t *col,
                            uint32_t *header, uint32_t *requirements);

/* Root candidates of `query` in `data`: the data vertices (ascending ids) that
 * own a segment of the depth-1 generator relation, i.e. every f(root) that can
 * start an embedding.  Pass ids = NULL to learn n first.  (No counterpart: the
 * reference starts from every row, HnGPU.cu:475-495.) */
gfsm_status gfsm_query_roots(const gfsm_graph *data, const gfsm_graph *query, uint32_t *ids, uint64_t *n);
/* Matching order gfsm_match uses for `flags`: original query ids by depth
 * (col[query size]).  The reference's order, or with GFSM_ORDER_COST the one
 * chosen from data's relation statistics (no reference counterpart). */
gfsm_status gfsm_plan_order(const gfsm_graph *data, const gfsm_graph *query, uint32_t flags, uint32_t *col);
/* Projected size of every depth (query-size entries each, depth order): the
 * expected nodes kept and candidates tested, from data's relation statistics
 * and the plan that `opts` selects (flags, root window).  Host only, no device.
 * The batch engine sizes query regions from it, scaled by a calibration learnt
 * from finished queries.  No reference counterpart (the reference sizes every
 * level at SOLN_LIMIT, HnGPU.cu:849-856). */
gfsm_status gfsm_query_projection(const gfsm_graph *data, const gfsm_graph *query, const gfsm_match_opts *opts,
                                  double *nodes, double *cand);

/* A root window for bounded work (no counterpart): the largest set of the
 * highest-id root candidates whose depth-1 candidates total at most max_first
 * (always at least one root), as [*root_lo, *root_hi) for gfsm_match_opts;
 * *first (may be NULL) = that total.  Deterministic: depends only on the
 * graphs. */
gfsm_status gfsm_query_window(const gfsm_graph *data, const gfsm_graph *query, uint64_t max_first,
                              uint32_t *root_lo, uint32_t *root_hi, uint64_t *first);

/* ---- matching ------------------------------------------------------------ */
gfsm_status gfsm_match(gfsm_ctx *ctx, const gfsm_graph *data, const gfsm_graph *query,
                       const gfsm_match_opts *opts, gfsm_result **r);
gfsm_status gfsm_result_count(const gfsm_result *r, uint64_t *n);
gfsm_status gfsm_result_rows(const gfsm_result *r, const uint32_t **rows, uint32_t *width, uint64_t *nrows);
gfsm_status gfsm_result_digest(const gfsm_result *r, uint64_t out[2]);
gfsm_status gfsm_result_stats(const gfsm_result *r, gfsm_stats *out);
gfsm_status gfsm_result_status(const gfsm_result *r);
gfsm_status gfsm_result_write_fs(const gfsm_result *r, const char *path);
/* Match and write output.fs at `path` (the layout of gfsm_result_write_fs)
 * while matching: unless GFSM_SORTED is set, each chunk's rows go to the file
 * as they land, so host memory holds one chunk whatever the result size (no
 * reference counterpart: the reference copies the whole table first,
 * HnGPU.cu:1019-1048).  *r holds count, digest and stats, no rows. */
gfsm_status gfsm_match_write_fs(gfsm_ctx *ctx, const gfsm_graph *data, const gfsm_graph *query,
                                const gfsm_match_opts *opts, const char *path, gfsm_result **r);
void        gfsm_result_free(gfsm_result *r);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc: gfsm_match in count mode (opts->flags: GFSM_DIGEST,
 * GFSM_ORDER_COST) that allocates nothing.  Its levels, scan workspace and
 * device totals live in a library-owned device array of
 * GFSM_NOALLOC_SCRATCH_BYTES per GPU (64 MiB, held until the process ends;
 * 60 MiB of it for levels): chunked expansion keeps any query inside it, and
 * opts->mem_budget above the level part is rejected (GFSM_ERR_ARG).  `data`
 * must already be resident on ctx (gfsm_graph_upload); dense segment tables
 * are used only where an earlier gfsm_match built them.  Writes the count
 * and digest sums; GFSM_ERR_CAP with the partial values when max_cand or
 * max_depth_cand stops it.  All calls on one GPU share its array: the
 * library runs them one at a time (they would race on it otherwise), and
 * each clears the words it needs. */
#define GFSM_NOALLOC_SCRATCH_BYTES (64ull << 20)
gfsm_status gfsm_match_noalloc(gfsm_ctx *ctx, const gfsm_graph *data, const gfsm_graph *query,
                               const gfsm_match_opts *opts, uint64_t *count, uint64_t digest[2]);

/* Row digest used by GFSM_DIGEST (per seed: mix(sum over columns c of
 * mix(seed ^ (row[c] | c << 32))), mix = the SplitMix64 finaliser): h = sum over rows of two seeded hashes of
 * the row (column order = original query id). */
void        gfsm_row_hash(const uint32_t *row, uint32_t width, uint64_t h[2]);

/* ---- batching (no counterpart): n queries share every step.  Each step is
 * a scheduler launch (one wave per query slot), a tile-index launch and one
 * persistent step launch whose device work queue holds the (query, tile)
 * items of every query in flight; runs of steps
 * are captured in a hipGraph and replayed.  opts holds nopts entries: 1 (the
 * same options for every query) or n (one per query).  Count and digest modes
 * run batched; queries asking for rows run on the single-query path.
 * results[i] is always set (free each); its gfsm_result_status is the query's
 * own status.  The call fails only for invalid arguments or device errors. */
gfsm_status gfsm_match_batch(gfsm_ctx *ctx, const gfsm_graph *data, const gfsm_graph *const *queries,
                             uint32_t n, const gfsm_match_opts *opts, uint32_t nopts, gfsm_result **results);

/* ---- serving: resident data graph, admission by projected memory,
 * continuous batching -------------------------------------------------------- */
typedef struct gfsm_serve_opts {
    uint64_t mem_budget;    /* device bytes shared by in-flight queries (0 = context default) */
    uint32_t max_inflight;  /* queries admitted at once (0 = 256)                              */
    uint32_t flags;         /* GFSM_DIGEST etc., applied to every query                        */
} gfsm_serve_opts;

/* Lifetimes: the server pins `data` and every query in flight.
 * gfsm_graph_free on a pinned graph only marks it; the graph is freed when
 * the server lets go of it (the query's result is recorded, or the server
 * stops).  gfsm_close(ctx) first stops every server still running on ctx:
 * each finishes the work already submitted, its finished results can still
 * be taken, and later submits fail with GFSM_ERR_ARG; gfsm_server_stop must
 * still be called to free the server. */
gfsm_status gfsm_server_start(gfsm_ctx *ctx, const gfsm_graph *data, const gfsm_serve_opts *opts,
                              gfsm_server **srv);
/* opts (may be NULL): this query's flags (added to the server's), root
 * window and work cap; its mem_budget and tile are ignored (the server sizes
 * each query's memory by projection). */
gfsm_status gfsm_server_submit(gfsm_server *srv, const gfsm_graph *query, const gfsm_match_opts *opts,
                               uint64_t *ticket);
/* GFSM_ERR_PENDING while running; on GFSM_OK *r is owned by the caller.  A
 * ticket's result is handed out once: a second poll or wait on it returns
 * GFSM_ERR_ARG. */
gfsm_status gfsm_server_poll(gfsm_server *srv, uint64_t ticket, gfsm_result **r);
gfsm_status gfsm_server_wait(gfsm_server *srv, uint64_t ticket, gfsm_result **r);
/* Per-ticket timestamps (seconds on the server clock): submit, admit, done.
 * A ticket's record is dropped once its result has been taken and its
 * timestamps read (then GFSM_ERR_ARG); of taken results whose timestamps are
 * never read, the server keeps the most recent 2^20 records. */
gfsm_status gfsm_server_times(gfsm_server *srv, uint64_t ticket, double t[3]);
/* Admission counters since the server started.  A query admitted to the
 * batch engine counts in waited_memory when the server held it back at
 * least once because its projected memory did not fit while a slot was free,
 * and in waited_slot when it was held back for a free slot. */
typedef struct gfsm_serve_stats {
    uint64_t submitted;      /* tickets handed out                                   */
    uint64_t admitted;       /* queries run on the batch engine                       */
    uint64_t waited_memory;  /* of those, held back for projected memory              */
    uint64_t waited_slot;    /* of those, held back for a free slot                   */
    uint64_t completed;      /* results recorded (any status)                         */
} gfsm_serve_stats;
gfsm_status gfsm_server_stats(gfsm_server *srv, gfsm_serve_stats *out);
void        gfsm_server_stop(gfsm_server *srv);

/* ---- self test ------------------------------------------------------------
 * (no counterpart) Runs the chained wave64 scan on n u32 values: prefix[i] =
 * sum of vals[0..i) (u64), idx = the indices with vals[i] != 0 in order, and
 * totals[0] = number of non-zero values, totals[1] = sum.  Host arrays. */
gfsm_status gfsm_selftest_scan(gfsm_ctx *ctx, const uint32_t *vals, uint64_t n, uint64_t *prefix,
                               uint32_t *idx, uint64_t totals[2]);
/* (no counterpart) Kernels of known HBM footprint, for scaling a profiler's
 * byte counters: mode 0 streams n words (dst[i] = src[i] + 1: 4n bytes read
 * and written per launch); mode 1 gathers one word from each of n distinct
 * 128-byte lines in a cache-defeating order (n a power of two: 128n bytes of
 * lines read, 4n written per launch); mode 2 gathers words 0 and 16 (both
 * 64-byte halves) of the same lines (128n read, 8n written).  Runs
 * `launches` launches; bytes = {read, written} per launch. */
gfsm_status gfsm_selftest_traffic(gfsm_ctx *ctx, uint32_t mode, uint64_t n, uint32_t launches, uint64_t bytes[2]);

/* ---- tile shapes and the autotuner (no counterpart) ------------------------
 * The kernels' launch shapes, per context.  items[kind][d]: candidates per
 * thread (1, 2, 4 or 8; a tile is 256 x items candidates) of the single-query
 * expand kernel for kind 0 = inner depth, 1 = last depth counting, 2 = last
 * depth writing rows, at depth d.  pool_words: LDS words per workgroup that
 * stage the tile's frontier nodes (checked segments, placed vertices).
 * wgs_per_cu: persistent workgroups per compute unit.  batch_*: the same for
 * the batch step kernel; batch_slots: queries in flight in one batch;
 * batch_steps: steps per captured replay; batch_step_budget: candidates one
 * step processes over all running queries (split evenly among them). */
typedef struct gfsm_tuning {
    uint32_t items[3][GFSM_MAX_Q];
    uint32_t pool_words;
    uint32_t wgs_per_cu;
    uint32_t batch_items;
    uint32_t batch_pool_words;
    uint32_t batch_wgs_per_cu;
    uint32_t batch_slots;
    uint32_t batch_steps;        /* steps per captured replay (1..64)                           */
    uint32_t serve_steps;        /* ... of a server started afterwards (0 = default 2; <= 64)   */
    uint64_t batch_step_budget;  /* candidates of one step over all running queries            */
} gfsm_tuning;
gfsm_status gfsm_tuning_get(gfsm_ctx *ctx, gfsm_tuning *t);
gfsm_status gfsm_tuning_set(gfsm_ctx *ctx, const gfsm_tuning *t);

/* Autotuner: runs the sample -- n queries with their options (nopts = 1:
 * the same for all, or n: root windows and caps per query; NULL = whole
 * range, no caps), count mode -- under each candidate shape, keeps the
 * fastest in ctx and returns it in *t.  which: 1 = single-query shapes, 2 =
 * batch shapes, 3 = both.  The sample should take tens of milliseconds per
 * shape. */
#define GFSM_TUNE_PLAIN 1u
#define GFSM_TUNE_BATCH 2u
gfsm_status gfsm_autotune(gfsm_ctx *ctx, const gfsm_graph *data, const gfsm_graph *const *queries, uint32_t n,
                          const gfsm_match_opts *opts, uint32_t nopts, uint32_t which, gfsm_tuning *t);

/* ---- seeded workload generators (undirected graphs as symmetric entry lists) */
#define GFSM_WALK_DENSE 0u   /* keep every induced edge among the walked vertices */
#define GFSM_WALK_SPARSE 1u  /* walk tree plus one extra induced edge             */
gfsm_status gfsm_gen_er(uint32_t n, double avg_degree, uint32_t nvlabels, uint32_t nelabels,
                        uint64_t seed, uint32_t mode, gfsm_graph **g);
gfsm_status gfsm_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                          uint32_t nvlabels, double zipf_s, uint32_t nelabels, uint64_t seed,
                          uint32_t mode, gfsm_graph **g);
gfsm_status gfsm_gen_grid(uint32_t rows, uint32_t cols, double del_frac, double diag_frac,
                          uint64_t seed, uint32_t mode, gfsm_graph **g);
gfsm_status gfsm_gen_query_walk(const gfsm_graph *data, uint32_t qsize, uint32_t kind, uint64_t seed,
                                gfsm_graph **q);

#ifdef __cplusplus
}
#endif
#endif
