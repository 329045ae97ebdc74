/*
 * oracle.h -- CPU restatement of GFSM's forest-based subgraph matcher.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the GPU product
 * (libgfsm.so) is compared against.  Only the test suite, the smoke check and
 * the CPU-baseline leg of the benchmark may load it.  The product never links,
 * loads or calls anything in this directory.
 *
 * Every function restates the reference algorithm and cites the file:line it
 * follows (paths relative to the reference repository root):
 *
 *   orc_load_tve        HnHelper.cuh:283-381, 191-266  (t/v/e grammar, error codes 232/233/234)
 *   orc_ccsr_build      HnHelper.cuh:383-478, ccsr.cpp:13-16 (degree-sorted CCSR, staggers, remap, row sort)
 *   orc_ccsr_packed     Environment.h:6-7, 69-83; HnHelper.cuh:376, 452-460 ((rel<<24)|offset packing)
 *   orc_plan_build      HnGPU.cu:749-816, ccsr.cpp:98-204 (matching order + requirement lists)
 *   orc_match           HnGPU.cu:76-109 (hnInit), 114-185 (hnGen), 190-262 (hnCheck), 34-52 (reqExists),
 *                       674-699 (hnWrite: CCSR offset -> original id through staggersInfo)
 *   orc_match_canon     the same enumeration in one fixed order under a per-depth candidate cap
 *                       (the product's bounded-work mode; no reference counterpart)
 *   orc_check_sound     authsrc/HnFunction.java:182-194, authsrc/HnGraph.java:28-75
 *   orc_fs_write/read   HnGPU.cu:1081-1091, authsrc/HnFunction.java:21-38, 127-146
 *
 * Pinning (what checks each function independently of this restatement) is
 * listed in tests/test_oracle.py; every function above has at least one check
 * that a plausible slip (wrong sign, dropped term, swapped operand) fails.
 */
#ifndef GFSM_ORACLE_H
#define GFSM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes: the reference's own exit codes (HnHelper.cuh:16-18, ccsr.cpp:127). */
#define ORC_OK 0
#define ORC_ERR_QUERY 1          /* disconnected / invalid query (ccsr.cpp:126-128, 192-196) */
#define ORC_ERR_ARG 2            /* bad argument or out-of-domain input (ids, labels)        */
#define ORC_ERR_NOMEM 3
#define ORC_ERR_CAP 5            /* output does not fit the caller's row cap                 */
#define ORC_ERR_MISSING_FILE 232
#define ORC_ERR_INVALID_DATA 233
#define ORC_ERR_MISSING_DATA 234

/* Relation modes. REF: relation = edge label, vertex labels ignored (HnHelper.cuh:330).
 * VLABEL: relation of u->v = fold(label(u), label(v), edge label). */
#define ORC_MODE_REF 0u
#define ORC_MODE_VLABEL 1u

/* A graph as the t/v/e file states it: directed entries in file order. */
typedef struct orc_edges {
    uint32_t V;          /* header vertex count                          */
    uint64_t E;          /* number of e-lines actually parsed            */
    uint32_t *vlabel;    /* [V] label by v-line id (0 when never stated) */
    uint32_t *src, *dst, *elabel; /* [E]                                 */
} orc_edges;

/* CCSR restated from ccsr.h:23-44 with the 8-bit relation split out of the
 * word so that graphs beyond 2^24 words stay representable. */
typedef struct orc_ccsr {
    uint32_t count;          /* V                                                  */
    uint64_t size;           /* W = sum over rows of (deg + 1)                     */
    uint32_t *tgt;           /* [W] header word = degree; entry = target row offset */
    uint32_t *rel;           /* [W] relation of an entry word; 0 for header words  */
    uint32_t *verticeLocs;   /* [V] row offset of each original vertex             */
    uint32_t *invVerticeLocs;/* [V] original id of each sorted rank                */
    uint32_t maxLen;         /* max row length                                     */
    uint64_t *staggers;      /* [maxLen] words in rows of degree d+1               */
    uint64_t *stagOffset;    /* [maxLen] staggersInfo[d].offset                    */
    uint64_t *stagNorm;      /* [maxLen] staggersInfo[d].normOffset                */
    /* Enumeration aid (not part of the reference layout): each row's entries
     * again, ordered by (relation, target), and a per-row flag (at the header
     * word) set when the row repeats a target. */
    uint32_t *rtgt, *rrel;   /* [W]                                                */
    uint8_t *dup;            /* [W]                                                */
} orc_ccsr;

/* Compiled query: requirements / requirementHeader exactly as the reference
 * lays them out (HnGPU.cuh:41-58), plus the matching order. */
typedef struct orc_plan {
    uint32_t q;              /* query vertex count                                   */
    uint32_t maxValency;     /* max query row length                                 */
    uint32_t *requirements;  /* [2*maxValency*q] {delta dist, relation} pairs         */
    uint32_t *header;        /* [3*q] {num, offset, firstQ}                          */
    uint32_t *order;         /* [q] sorted query index mapped at depth d             */
    uint32_t *col;           /* [q] original query vertex id mapped at depth d       */
} orc_plan;

typedef struct orc_result {
    uint64_t count;          /* number of embeddings                                 */
    uint64_t digest[2];      /* order-independent multiset digest of the rows        */
    uint32_t width;          /* q                                                    */
    uint64_t nrows;          /* rows materialised (0 unless requested)               */
    uint32_t *rows;          /* [nrows*width] row-major, column = original query id  */
    uint64_t *level;         /* [q] survivors per depth (prefix embeddings)          */
} orc_result;

/* Parsing and construction. */
int  orc_load_tve(const char *path, orc_edges **out);
int  orc_from_arrays(uint32_t V, const uint32_t *vlabel, uint64_t E, const uint32_t *src,
                     const uint32_t *dst, const uint32_t *elabel, orc_edges **out);
void orc_edges_free(orc_edges *g);
int  orc_write_tve(const orc_edges *g, const char *path);

/* Relation of every entry under a mode; out[E]. */
int  orc_relations(const orc_edges *g, uint32_t mode, uint32_t *out);

int  orc_ccsr_build(const orc_edges *g, uint32_t mode, orc_ccsr **out);
void orc_ccsr_free(orc_ccsr *c);
/* Reference word packing (rel<<24 | target); ORC_ERR_ARG if W >= 2^24 or rel > 255. */
int  orc_ccsr_packed(const orc_ccsr *c, uint32_t *out_words);
/* Reference id mapping of hnWrite (HnGPU.cu:686-691): CCSR row offset -> original id. */
uint32_t orc_offset_to_id(const orc_ccsr *c, uint32_t offset);

int  orc_plan_build(const orc_ccsr *query, orc_plan **out);
void orc_plan_free(orc_plan *p);

/* Enumerate.  Only roots whose original id lies in [root_lo, root_hi) are
 * expanded.  rows are materialised when max_rows > 0 (ORC_ERR_CAP if more).
 * nthreads <= 0 uses every core. */
int  orc_match(const orc_ccsr *data, const orc_ccsr *query, const orc_plan *plan,
               uint32_t root_lo, uint32_t root_hi, int nthreads, uint64_t max_rows,
               orc_result **out);
/* Work-bounded variant for timing samples: enumeration stops once about
 * max_leaves embeddings have been found (count is then a partial count). */
int  orc_match_capped(const orc_ccsr *data, const orc_ccsr *query, const orc_plan *plan,
                      uint32_t root_lo, uint32_t root_hi, int nthreads, uint64_t max_rows,
                      uint64_t max_leaves, orc_result **out);
void orc_result_free(orc_result *r);

/* Canonical enumeration order: each row's entries again, ordered by
 * (relation, original target id).  Built once per data graph. */
typedef struct orc_canon {
    uint64_t size;           /* W of the CCSR it was built from                    */
    uint32_t *cid;           /* [W] original target ids, per row by (relation, id) */
    uint32_t *coff;          /* [W] the same entries' target row offsets           */
    uint32_t *crel;          /* [W] their relations                                */
} orc_canon;
int  orc_canon_build(const orc_ccsr *data, orc_canon **out);
void orc_canon_free(orc_canon *cn);

/* Sequential enumeration in canonical order (roots by ascending id; at each
 * depth the generator segment by ascending target id) under the deterministic
 * per-depth cap the product offers for bounded work: depth k tests only its
 * first max_depth_cand candidates (0 = no cap), and enumeration stops the
 * first time a depth would test one more (*capped = 1; count and digest are
 * then partial).  Root candidates are the vertices holding an entry of depth
 * 1's generator relation.  cand[q] (optional) receives the candidates tested
 * per depth.  No counterpart in the reference. */
int  orc_match_canon(const orc_ccsr *data, const orc_canon *cn, const orc_ccsr *query,
                     const orc_plan *plan, uint32_t root_lo, uint32_t root_hi,
                     uint64_t max_depth_cand, orc_result **out, int *capped, uint64_t *cand);

/* Convenience: edges -> CCSR -> plan -> match in one call. */
int  orc_solve(const orc_edges *data, const orc_edges *query, uint32_t mode,
               uint32_t root_lo, uint32_t root_hi, int nthreads, uint64_t max_rows,
               orc_result **out);

/* Soundness check restated from HnFunction.isSubgraphIso: returns the index of
 * the first row that maps some query entry (v->e,label) onto a missing data
 * entry, or nrows when every row is sound. */
uint64_t orc_check_sound(const orc_edges *query, const orc_edges *data,
                         const uint32_t *rows, uint64_t nrows, uint32_t width);

/* Row digest (the definition shared with the product). */
void orc_row_hash(const uint32_t *row, uint32_t width, uint64_t h[2]);

/* output.fs: little-endian u32 count, u32 width, count*width u32. */
int  orc_fs_write(const char *path, const uint32_t *rows, uint64_t nrows, uint32_t width);
int  orc_fs_read(const char *path, uint32_t **rows, uint64_t *nrows, uint32_t *width);

void orc_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
