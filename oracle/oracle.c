/*
 * oracle.c -- CPU restatement of GFSM's matcher.  TEST INFRASTRUCTURE ONLY:
 * the checker for the GPU product, never part of it.  See oracle.h for the
 * reference file:line each function follows.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_free(void *p) { free(p); }

/* ------------------------------------------------------------------------ */
/* t/v/e parsing (HnHelper.cuh:191-266 parseValues, 283-381 fileParse)       */
/* ------------------------------------------------------------------------ */

enum { P_COMPLETE, P_FAIL, P_EDGES_FOUND, P_END };
enum { M_HEADER, M_VERTEX, M_EDGE };

/* One line.  Every non-digit character advances the field index, '\n' ends
 * the line (HnHelper.cuh:233-263).  The reference writes past its 4-slot
 * value array on over-long lines; here that is reported as invalid data. */
static int parse_line(const char **str, const char *end, uint64_t *values, int nvalues, int mode)
{
    const char *s = *str;
    if (s >= end) return P_FAIL;
    char ch = *s;
    if (ch == 't') {                                   /* HnHelper.cuh:196-201 */
        if (end - s < 6 || strncmp(s, "t # -1", 6) != 0) return P_FAIL;
        return P_END;
    }
    if (mode == M_VERTEX) {                            /* HnHelper.cuh:212-221 */
        if (ch == 'e') return P_EDGES_FOUND;
        if (ch != 'v') return P_FAIL;
        s += 2;
    } else if (mode == M_EDGE) {                       /* HnHelper.cuh:223-229 */
        if (ch != 'e') return P_FAIL;
        s += 2;
    }
    /* header mode starts at the first digit of the line (HnHelper.cuh:206-209) */
    int cur = 0;
    for (; s < end; s++) {
        ch = *s;
        if (ch >= '0' && ch <= '9') {
            if (cur >= nvalues) return P_FAIL;
            values[cur] = values[cur] * 10u + (uint64_t)(ch - '0');
            if (values[cur] > 0xFFFFFFFFull) return P_FAIL;
        } else {
            cur++;
            if (ch == '\n') { s++; break; }
        }
    }
    *str = s;
    return P_COMPLETE;
}

int orc_load_tve(const char *path, orc_edges **out)
{
    *out = NULL;
    FILE *f = fopen(path, "rb");
    if (!f) return ORC_ERR_MISSING_FILE;               /* HnHelper.cuh:289-294 */
    fseek(f, 0, SEEK_END);
    long size = ftell(f);
    rewind(f);
    char *buf = (char *)malloc((size_t)size + 1);
    if (!buf) { fclose(f); return ORC_ERR_NOMEM; }
    size_t got = fread(buf, 1, (size_t)size, f);
    fclose(f);
    buf[got] = 0;
    const char *s = buf, *end = buf + got;
    int rc = ORC_OK;
    if (got < 6 || strncmp(buf, "t # 0\n", 6) != 0) { free(buf); return ORC_ERR_INVALID_DATA; }
    s += 6;

    uint64_t values[4] = {0, 0, 0, 0};
    if (parse_line(&s, end, values, 4, M_HEADER) != P_COMPLETE) { free(buf); return ORC_ERR_INVALID_DATA; }
    uint64_t V = values[0], E = values[1];
    if (V == 0) { free(buf); return ORC_ERR_INVALID_DATA; }

    orc_edges *g = (orc_edges *)calloc(1, sizeof(orc_edges));
    g->V = (uint32_t)V;
    g->vlabel = (uint32_t *)calloc(V, sizeof(uint32_t));

    /* vertex lines (HnHelper.cuh:333-349): counted; the label is kept here
     * (indexed by the id field) for the label-folding mode. */
    uint64_t vertexSeen = 0;
    int st;
    for (;;) {
        values[0] = values[1] = 0;
        st = parse_line(&s, end, values, 2, M_VERTEX);
        if (st == P_FAIL) { rc = ORC_ERR_INVALID_DATA; goto fail; }
        if (st != P_COMPLETE) break;
        if (values[0] < V) g->vlabel[values[0]] = (uint32_t)values[1];
        vertexSeen++;
    }
    if (vertexSeen < V) { rc = ORC_ERR_MISSING_DATA; goto fail; }

    /* edge lines (HnHelper.cuh:364-381) */
    uint64_t cap = E ? E : 16, n = 0;
    g->src = (uint32_t *)malloc(cap * sizeof(uint32_t));
    g->dst = (uint32_t *)malloc(cap * sizeof(uint32_t));
    g->elabel = (uint32_t *)malloc(cap * sizeof(uint32_t));
    for (;;) {
        values[0] = values[1] = values[2] = 0;
        st = parse_line(&s, end, values, 3, M_EDGE);
        if (st == P_FAIL) { rc = ORC_ERR_INVALID_DATA; goto fail; }
        if (st == P_END) break;
        if (n == cap) {
            cap *= 2;
            g->src = (uint32_t *)realloc(g->src, cap * sizeof(uint32_t));
            g->dst = (uint32_t *)realloc(g->dst, cap * sizeof(uint32_t));
            g->elabel = (uint32_t *)realloc(g->elabel, cap * sizeof(uint32_t));
        }
        g->src[n] = (uint32_t)values[0];
        g->dst[n] = (uint32_t)values[1];
        g->elabel[n] = (uint32_t)values[2];
        n++;
    }
    if (n < E) { rc = ORC_ERR_MISSING_DATA; goto fail; }
    g->E = n;
    free(buf);
    *out = g;
    return ORC_OK;
fail:
    free(buf);
    orc_edges_free(g);
    return rc;
}

int orc_from_arrays(uint32_t V, const uint32_t *vlabel, uint64_t E, const uint32_t *src,
                    const uint32_t *dst, const uint32_t *elabel, orc_edges **out)
{
    orc_edges *g = (orc_edges *)calloc(1, sizeof(orc_edges));
    g->V = V;
    g->E = E;
    g->vlabel = (uint32_t *)calloc(V ? V : 1, sizeof(uint32_t));
    if (vlabel) memcpy(g->vlabel, vlabel, (size_t)V * sizeof(uint32_t));
    size_t b = (size_t)(E ? E : 1) * sizeof(uint32_t);
    g->src = (uint32_t *)malloc(b);
    g->dst = (uint32_t *)malloc(b);
    g->elabel = (uint32_t *)malloc(b);
    if (E) {
        memcpy(g->src, src, E * sizeof(uint32_t));
        memcpy(g->dst, dst, E * sizeof(uint32_t));
        memcpy(g->elabel, elabel, E * sizeof(uint32_t));
    }
    *out = g;
    return ORC_OK;
}

void orc_edges_free(orc_edges *g)
{
    if (!g) return;
    free(g->vlabel); free(g->src); free(g->dst); free(g->elabel);
    free(g);
}

int orc_write_tve(const orc_edges *g, const char *path)
{
    FILE *f = fopen(path, "wb");
    if (!f) return ORC_ERR_MISSING_FILE;
    uint32_t maxv = 0, maxe = 0;
    for (uint32_t i = 0; i < g->V; i++) if (g->vlabel[i] > maxv) maxv = g->vlabel[i];
    for (uint64_t i = 0; i < g->E; i++) if (g->elabel[i] > maxe) maxe = g->elabel[i];
    fprintf(f, "t # 0\n%u %llu %u %u\n", g->V, (unsigned long long)g->E, maxv, maxe);
    for (uint32_t i = 0; i < g->V; i++) fprintf(f, "v %u %u\n", i, g->vlabel[i]);
    for (uint64_t i = 0; i < g->E; i++) fprintf(f, "e %u %u %u\n", g->src[i], g->dst[i], g->elabel[i]);
    fprintf(f, "t # -1");
    fclose(f);
    return ORC_OK;
}

/* Relation of an entry.  REF: the edge label (HnHelper.cuh:376).  VLABEL:
 * fold(label(u), label(v), edge label) = (lu << 21) | (lv << 10) | l, so
 * that requiring equal relations is requiring equal end labels and equal edge
 * labels. */
int orc_relations(const orc_edges *g, uint32_t mode, uint32_t *out)
{
    for (uint64_t i = 0; i < g->E; i++) {
        uint32_t s = g->src[i], d = g->dst[i], l = g->elabel[i];
        if (s >= g->V || d >= g->V || l == 0) return ORC_ERR_ARG;
        if (mode == ORC_MODE_REF) {
            out[i] = l;
        } else {
            uint32_t ls = g->vlabel[s], ld = g->vlabel[d];
            if (ls >= 2048u || ld >= 2048u || l >= 1024u) return ORC_ERR_ARG;
            out[i] = (ls << 21) | (ld << 10) | l;
        }
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* CCSR construction (HnHelper.cuh:383-478)                                 */
/* ------------------------------------------------------------------------ */

typedef struct { uint32_t mapping, size; } order_pair;

/* OrderPair::operator< (HnHelper.cuh:26-31): larger size first, then lower id. */
static int order_cmp(const void *a, const void *b)
{
    const order_pair *x = (const order_pair *)a, *y = (const order_pair *)b;
    if (x->size != y->size) return x->size > y->size ? -1 : 1;
    return (x->mapping > y->mapping) - (x->mapping < y->mapping);
}

typedef struct { uint32_t tgt, rel; } entry_pair;

/* nodeCompare (ccsr.cpp:13-16) orders by target only; equal targets (parallel
 * entries) are ordered by relation so the layout is deterministic. */
static int entry_cmp(const void *a, const void *b)
{
    const entry_pair *x = (const entry_pair *)a, *y = (const entry_pair *)b;
    if (x->tgt != y->tgt) return x->tgt < y->tgt ? -1 : 1;
    return (x->rel > y->rel) - (x->rel < y->rel);
}

int orc_ccsr_build(const orc_edges *g, uint32_t mode, orc_ccsr **out)
{
    *out = NULL;
    uint32_t V = g->V;
    uint64_t E = g->E;
    uint32_t *rel = (uint32_t *)malloc((E ? E : 1) * sizeof(uint32_t));
    int rc = orc_relations(g, mode, rel);
    if (rc) { free(rel); return rc; }

    /* Rows keep file order, like the chained rows of Dynamic2D (HnHelper.cuh:40-174). */
    uint64_t *rowStart = (uint64_t *)calloc((size_t)V + 1, sizeof(uint64_t));
    for (uint64_t i = 0; i < E; i++) rowStart[g->src[i] + 1]++;
    uint32_t maxLen = 0;
    for (uint32_t v = 0; v < V; v++) {
        if (rowStart[v + 1] > 0xFFFFFFFFull) { free(rel); free(rowStart); return ORC_ERR_ARG; }
        if (rowStart[v + 1] > maxLen) maxLen = (uint32_t)rowStart[v + 1];
    }
    for (uint32_t v = 0; v < V; v++) rowStart[v + 1] += rowStart[v];
    uint64_t *fill = (uint64_t *)malloc(((size_t)V + 1) * sizeof(uint64_t));
    memcpy(fill, rowStart, ((size_t)V + 1) * sizeof(uint64_t));
    entry_pair *rows = (entry_pair *)malloc((E ? E : 1) * sizeof(entry_pair));
    for (uint64_t i = 0; i < E; i++) {
        uint64_t p = fill[g->src[i]]++;
        rows[p].tgt = g->dst[i];
        rows[p].rel = rel[i];
    }
    free(fill);
    free(rel);

    /* allocationSize (HnHelper.cuh:383-393) */
    uint64_t W = 0;
    order_pair *ord = (order_pair *)malloc((size_t)V * sizeof(order_pair));
    for (uint32_t v = 0; v < V; v++) {
        uint32_t sz = (uint32_t)(rowStart[v + 1] - rowStart[v]);
        if (sz) W += (uint64_t)sz + 1;
        ord[v].mapping = v;
        ord[v].size = sz;
    }
    if (W >= 0xFFFFFFFFull) { free(ord); free(rows); free(rowStart); return ORC_ERR_ARG; }
    qsort(ord, V, sizeof(order_pair), order_cmp);       /* HnHelper.cuh:413 */

    orc_ccsr *c = (orc_ccsr *)calloc(1, sizeof(orc_ccsr));
    c->count = V;
    c->size = W;
    c->maxLen = maxLen;
    c->tgt = (uint32_t *)malloc((W ? W : 1) * sizeof(uint32_t));
    c->rel = (uint32_t *)calloc(W ? W : 1, sizeof(uint32_t));
    c->verticeLocs = (uint32_t *)calloc(V, sizeof(uint32_t));
    c->invVerticeLocs = (uint32_t *)calloc(V, sizeof(uint32_t));
    size_t ml = maxLen ? maxLen : 1;
    c->staggers = (uint64_t *)calloc(ml, sizeof(uint64_t));
    c->stagOffset = (uint64_t *)calloc(ml, sizeof(uint64_t));
    c->stagNorm = (uint64_t *)calloc(ml, sizeof(uint64_t));
    uint64_t *degreeCounts = (uint64_t *)calloc(ml, sizeof(uint64_t));

    /* Row layout in degree order (HnHelper.cuh:417-436). */
    uint64_t w = 0;
    for (uint32_t i = 0; i < V; i++) {
        uint32_t v = ord[i].mapping;
        uint32_t sz = ord[i].size;
        c->verticeLocs[v] = (uint32_t)w;
        c->invVerticeLocs[i] = v;
        if (sz) {
            c->tgt[w] = sz;
            c->rel[w] = 0;
            for (uint32_t k = 0; k < sz; k++) {
                c->tgt[w + 1 + k] = rows[rowStart[v] + k].tgt;
                c->rel[w + 1 + k] = rows[rowStart[v] + k].rel;
            }
            c->staggers[sz - 1] += (uint64_t)sz + 1;
            degreeCounts[sz - 1]++;
            w += (uint64_t)sz + 1;
        }
    }
    /* Stagger tables, "top heavy" (HnHelper.cuh:443-448). */
    if (maxLen) {
        c->stagOffset[maxLen - 1] = 0;
        c->stagNorm[maxLen - 1] = 0;
        for (int64_t i = (int64_t)maxLen - 2; i >= 0; i--) {
            c->stagOffset[i] = c->stagOffset[i + 1] + c->staggers[i + 1];
            c->stagNorm[i] = c->stagNorm[i + 1] + degreeCounts[i + 1];
        }
    }
    /* Remap targets from vertex ids to row offsets (HnHelper.cuh:452-460). */
    for (uint64_t i = 0; i < W; i += (uint64_t)c->tgt[i] + 1) {
        uint32_t len = c->tgt[i];
        for (uint32_t k = 1; k <= len; k++) c->tgt[i + k] = c->verticeLocs[c->tgt[i + k]];
    }
    /* Per-row sort by target (HnHelper.cuh:469-478). */
    entry_pair *tmp = (entry_pair *)malloc(ml * sizeof(entry_pair));
    for (uint64_t i = 0; i < W; i += (uint64_t)c->tgt[i] + 1) {
        uint32_t len = c->tgt[i];
        for (uint32_t k = 0; k < len; k++) { tmp[k].tgt = c->tgt[i + 1 + k]; tmp[k].rel = c->rel[i + 1 + k]; }
        qsort(tmp, len, sizeof(entry_pair), entry_cmp);
        for (uint32_t k = 0; k < len; k++) { c->tgt[i + 1 + k] = tmp[k].tgt; c->rel[i + 1 + k] = tmp[k].rel; }
    }
    /* Enumeration aid: the same rows ordered by (relation, target), and a
     * flag for rows that repeat a target (parallel entries). */
    c->rtgt = (uint32_t *)calloc(W ? W : 1, sizeof(uint32_t));
    c->rrel = (uint32_t *)calloc(W ? W : 1, sizeof(uint32_t));
    c->dup = (uint8_t *)calloc(W ? W : 1, 1);
    for (uint64_t i = 0; i < W; i += (uint64_t)c->tgt[i] + 1) {
        uint32_t len = c->tgt[i];
        for (uint32_t k = 1; k < len; k++)
            if (c->tgt[i + 1 + k] == c->tgt[i + k]) c->dup[i] = 1;
        for (uint32_t k = 0; k < len; k++) { tmp[k].tgt = c->rel[i + 1 + k]; tmp[k].rel = c->tgt[i + 1 + k]; }
        qsort(tmp, len, sizeof(entry_pair), entry_cmp);  /* by (relation, target) */
        for (uint32_t k = 0; k < len; k++) { c->rrel[i + 1 + k] = tmp[k].tgt; c->rtgt[i + 1 + k] = tmp[k].rel; }
    }
    free(tmp);
    free(degreeCounts);
    free(ord);
    free(rows);
    free(rowStart);
    *out = c;
    return ORC_OK;
}

void orc_ccsr_free(orc_ccsr *c)
{
    if (!c) return;
    free(c->tgt); free(c->rel); free(c->verticeLocs); free(c->invVerticeLocs);
    free(c->staggers); free(c->stagOffset); free(c->stagNorm);
    free(c->rtgt); free(c->rrel); free(c->dup);
    free(c);
}

int orc_ccsr_packed(const orc_ccsr *c, uint32_t *out)
{
    if (c->size >= (1ull << 24)) return ORC_ERR_ARG;
    for (uint64_t i = 0; i < c->size; i++) {
        if (c->rel[i] > 255u) return ORC_ERR_ARG;
        out[i] = (c->rel[i] << 24) | c->tgt[i];
    }
    return ORC_OK;
}

/* Length of the row at offset off.  A vertex with no out-entries has no row:
 * the layout points it one past the last row (offset == size), where the
 * reference reads beyond its segment array; here that row is empty. */
static inline uint32_t row_len(const orc_ccsr *c, uint64_t off)
{
    return off < c->size ? c->tgt[off] : 0u;
}

/* hnWrite's id recovery (HnGPU.cu:684-691).  Vertices without a row share
 * one offset, so their ids cannot be told apart: UINT32_MAX. */
uint32_t orc_offset_to_id(const orc_ccsr *c, uint32_t offset)
{
    uint32_t deg = row_len(c, offset);
    if (deg == 0) return 0xFFFFFFFFu;
    uint64_t rank = c->stagNorm[deg - 1] + (offset - c->stagOffset[deg - 1]) / ((uint64_t)deg + 1);
    return c->invVerticeLocs[rank];
}

/* ------------------------------------------------------------------------ */
/* Query compilation (HnGPU.cu:749-816 preProcessQuery, ccsr.cpp:98-204)    */
/* ------------------------------------------------------------------------ */

typedef struct { uint32_t req, rel; } req_rel;

static int reqrel_cmp(const void *a, const void *b)
{
    const req_rel *x = (const req_rel *)a, *y = (const req_rel *)b;
    if (x->req != y->req) return x->req < y->req ? -1 : 1;
    return (x->rel > y->rel) - (x->rel < y->rel);
}

static int gen_requirements(const orc_ccsr *g, uint32_t *requirements, uint32_t *header,
                            uint32_t maxValency, uint32_t depth, int32_t *mappingData,
                            const uint32_t *localVerticeLocs, uint32_t *chosen)
{
    uint32_t count = g->count;
    uint32_t relativeDepth = 0;
    int found = 0;
    /* ccsr.cpp:114-123: first seen-but-unmapped vertex in sorted order; every
     * mapped vertex ages by one depth. */
    for (uint32_t i = 0; i < count; i++) {
        if (mappingData[i] == 0 && !found) {
            mappingData[i] = 1;
            relativeDepth = i;
            found = 1;
        } else if (mappingData[i] > 0) {
            mappingData[i]++;
        }
    }
    if (!found) return ORC_ERR_QUERY;                 /* ccsr.cpp:126-128 */
    *chosen = relativeDepth;

    uint64_t rowOff = localVerticeLocs[relativeDepth];
    uint32_t deg = row_len(g, rowOff);
    uint32_t reqOffset = depth * maxValency * 2;
    req_rel *reqRels = (req_rel *)calloc(maxValency ? maxValency : 1, sizeof(req_rel));
    uint32_t numOfReqs = 0;
    for (uint32_t k = 0; k < deg; k++) {              /* ccsr.cpp:142-167 */
        uint32_t currentVertex = g->tgt[rowOff + 1 + k];
        uint32_t currentIndex = count;
        for (uint32_t i = 0; i < count; i++)
            if (localVerticeLocs[i] == currentVertex) { currentIndex = i; break; }
        if (currentIndex == count) { free(reqRels); return ORC_ERR_QUERY; }
        if (mappingData[currentIndex] < 0) mappingData[currentIndex] = 0;
        if (mappingData[currentIndex] > 0) {
            uint32_t relation = g->rel[rowOff + 1 + k];
            /* reflected row: prefer the relation of the entry u -> n (ccsr.cpp:152-160) */
            uint32_t rdeg = row_len(g, currentVertex);
            for (uint32_t k2 = 0; k2 < rdeg; k2++) {
                if (g->tgt[currentVertex + 1 + k2] == localVerticeLocs[relativeDepth]) {
                    relation = g->rel[currentVertex + 1 + k2];
                    break;
                }
            }
            reqRels[numOfReqs].req = (uint32_t)(mappingData[currentIndex] - 1);
            reqRels[numOfReqs].rel = relation;
            numOfReqs++;
        }
    }
    uint32_t firstQ = 0;
    if (numOfReqs) {                                  /* ccsr.cpp:169-187: sort, delta-encode */
        qsort(reqRels, numOfReqs, sizeof(req_rel), reqrel_cmp);
        requirements[reqOffset] = reqRels[0].req;
        requirements[reqOffset + 1] = reqRels[0].rel;
        for (uint32_t i = 1; i < numOfReqs; i++) {
            requirements[reqOffset + i * 2] = reqRels[i].req - reqRels[i - 1].req;
            requirements[reqOffset + i * 2 + 1] = reqRels[i].rel;
        }
        firstQ = requirements[reqOffset];
        for (uint32_t i = 0; i < numOfReqs; i++) {
            firstQ = requirements[reqOffset + i * 2];
            if (firstQ) break;
        }
    }
    free(reqRels);
    if ((firstQ > depth && depth) || (depth && !firstQ)) return ORC_ERR_QUERY; /* ccsr.cpp:192-196 */
    header[depth * 3 + 0] = numOfReqs;
    header[depth * 3 + 1] = reqOffset;
    header[depth * 3 + 2] = firstQ;
    return ORC_OK;
}

int orc_plan_build(const orc_ccsr *q, orc_plan **out)
{
    *out = NULL;
    uint32_t count = q->count;
    /* getMaxValency (ccsr.cpp:206-219) */
    uint32_t maxValency = 0;
    for (uint64_t i = 0; i < q->size; i += (uint64_t)q->tgt[i] + 1)
        if (q->tgt[i] > maxValency) maxValency = q->tgt[i];
    if (count == 0 || maxValency == 0) return ORC_ERR_QUERY;

    /* Every query vertex needs a row, or the reference's offset walk
     * (ccsr.cpp:103-108) reads past the array. */
    uint32_t *localVerticeLocs = (uint32_t *)calloc(count, sizeof(uint32_t));
    uint64_t off = 0;
    for (uint32_t i = 0; i < count; i++) {
        if (off >= q->size) { free(localVerticeLocs); return ORC_ERR_QUERY; }
        localVerticeLocs[i] = (uint32_t)off;
        off += (uint64_t)q->tgt[off] + 1;
    }

    orc_plan *p = (orc_plan *)calloc(1, sizeof(orc_plan));
    p->q = count;
    p->maxValency = maxValency;
    p->requirements = (uint32_t *)calloc((size_t)2 * maxValency * count, sizeof(uint32_t));
    p->header = (uint32_t *)calloc((size_t)3 * count, sizeof(uint32_t));
    p->order = (uint32_t *)calloc(count, sizeof(uint32_t));
    p->col = (uint32_t *)calloc(count, sizeof(uint32_t));
    int32_t *mappingData = (int32_t *)malloc(count * sizeof(int32_t));
    for (uint32_t i = 0; i < count; i++) mappingData[i] = -1;

    /* Root: the first row of maximum length, a word offset used as an index
     * (HnGPU.cu:758-768); in degree-sorted order that is always 0. */
    uint32_t degVal = q->tgt[0];
    uint64_t loc = 0;
    for (uint64_t i = 0; i < q->size; i += (uint64_t)q->tgt[i] + 1)
        if (degVal < q->tgt[i]) { degVal = q->tgt[i]; loc = i; }
    if (loc >= count) { free(mappingData); free(localVerticeLocs); orc_plan_free(p); return ORC_ERR_QUERY; }
    mappingData[loc] = 0;

    int rc = ORC_OK;
    for (uint32_t depth = 0; depth < count && rc == ORC_OK; depth++) {
        uint32_t chosen = 0;
        rc = gen_requirements(q, p->requirements, p->header, maxValency, depth, mappingData,
                              localVerticeLocs, &chosen);
        p->order[depth] = chosen;
        p->col[depth] = q->invVerticeLocs[chosen];
    }
    /* HnGPU.cu:807-813 */
    for (uint64_t i = 0; rc == ORC_OK && i < (uint64_t)maxValency * count; i++)
        if (p->requirements[i * 2] > count) rc = ORC_ERR_QUERY;
    free(mappingData);
    free(localVerticeLocs);
    if (rc) { orc_plan_free(p); return rc; }
    *out = p;
    return ORC_OK;
}

void orc_plan_free(orc_plan *p)
{
    if (!p) return;
    free(p->requirements); free(p->header); free(p->order); free(p->col);
    free(p);
}

/* ------------------------------------------------------------------------ */
/* Enumeration                                                              */
/* ------------------------------------------------------------------------ */

/* Row digest: a multiset hash, summed over rows mod 2^64 with two seeds.
 * Per seed: h = mix(sum over columns c of mix(seed ^ (row[c] | c << 32))),
 * where mix is the SplitMix64 finaliser. */
static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

void orc_row_hash(const uint32_t *row, uint32_t width, uint64_t h[2])
{
    uint64_t a = 0, b = 0;
    for (uint32_t c = 0; c < width; c++) {
        uint64_t x = (uint64_t)row[c] | ((uint64_t)c << 32);
        a += mix64(0x9E3779B97F4A7C15ull ^ x);
        b += mix64(0xD1B54A32D192ED03ull ^ x);
    }
    h[0] = mix64(a);
    h[1] = mix64(b);
}

/* reqExists (HnGPU.cu:34-52): is there an entry (rel, vertex) in the row at
 * rowOff?  Rows are sorted by target, so the target is found by binary
 * search and the equal-target run is scanned for the relation. */
static inline int req_exists(const orc_ccsr *g, uint32_t rowOff, uint32_t rel, uint32_t vertex)
{
    uint32_t len = row_len(g, rowOff);
    if (len == 0) return 0;
    const uint32_t *t = g->tgt + rowOff + 1;
    const uint32_t *r = g->rel + rowOff + 1;
    uint32_t lo = 0, hi = len;
    while (lo < hi) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if (t[mid] < vertex) lo = mid + 1; else hi = mid;
    }
    for (; lo < len && t[lo] == vertex; lo++)
        if (r[lo] == rel) return 1;
    return 0;
}

typedef struct {
    uint64_t count, d0, d1;
    uint64_t *level;
    uint32_t *rows;
    uint64_t nrows, cap;
    int overflow;
} thread_acc;

typedef struct {
    const orc_ccsr *g;
    const orc_plan *p;
    const orc_ccsr *qg;
    uint64_t max_rows;
    uint64_t max_leaves;     /* stop after about this many leaves (0 = never) */
    uint64_t *leaves;        /* shared leaf counter, flushed every 4096 leaves */
    int *stop;
} match_ctx;

static void emit_leaf(const match_ctx *mc, const uint32_t *m, thread_acc *acc)
{
    uint32_t q = mc->p->q;
    uint32_t row[64];
    for (uint32_t d = 0; d < q; d++) row[mc->p->col[d]] = orc_offset_to_id(mc->g, m[d]);
    uint64_t h[2];
    orc_row_hash(row, q, h);
    acc->count++;
    if (mc->max_leaves && (acc->count & 4095) == 0) {
        uint64_t n;
#pragma omp atomic capture
        n = *mc->leaves += 4096;
        if (n >= mc->max_leaves) {
#pragma omp atomic write
            *mc->stop = 1;
        }
    }
    acc->d0 += h[0];
    acc->d1 += h[1];
    if (mc->max_rows) {
        if (acc->nrows == acc->cap) {
            uint64_t nc = acc->cap ? acc->cap * 2 : 1024;
            uint32_t *nr = (uint32_t *)realloc(acc->rows, nc * q * sizeof(uint32_t));
            if (!nr) { acc->overflow = 1; return; }
            acc->rows = nr;
            acc->cap = nc;
        }
        memcpy(acc->rows + acc->nrows * q, row, q * sizeof(uint32_t));
        acc->nrows++;
    }
}

/* One candidate c at depth d: hnCheck rejects it when it equals any
 * ancestor (HnGPU.cu:223-231) or misses a requirement (243-257). */
static inline int check_candidate(const match_ctx *mc, uint32_t d, const uint32_t *m, uint32_t c)
{
    const orc_ccsr *g = mc->g;
    const orc_plan *p = mc->p;
    uint32_t num = p->header[d * 3 + 0];
    uint32_t roff = p->header[d * 3 + 1];
    for (uint32_t k = 0; k < d; k++)
        if (m[k] == c) return 0;
    uint32_t cur = d;
    for (uint32_t j = 0; j < num; j++) {
        cur -= p->requirements[roff + 2 * j];
        uint32_t anc = (cur == d) ? c : m[cur];
        if (!req_exists(g, anc, p->requirements[roff + 2 * j + 1], c)) return 0;
    }
    return 1;
}

/* Depth d >= 1: hnGen takes every entry of the row of the ancestor firstQ
 * levels up (HnGPU.cu:130-175) and hnCheck filters them.  A candidate that
 * passes carries, in that row, an entry with the relation of the first
 * requirement at distance firstQ.  When the row repeats no target, that entry
 * is the one that produced the candidate, so only the entries with that
 * relation (a contiguous run of the (relation, target)-ordered copy) can pass,
 * and iterating just them yields exactly the same embeddings; rows that repeat
 * a target are iterated in full, as hnGen does. */
static void expand(const match_ctx *mc, uint32_t d, uint32_t *m, thread_acc *acc)
{
    const orc_ccsr *g = mc->g;
    const orc_plan *p = mc->p;
    uint32_t q = p->q;
    if (d == q) { emit_leaf(mc, m, acc); return; }
    uint32_t num = p->header[d * 3 + 0];
    uint32_t roff = p->header[d * 3 + 1];
    uint32_t firstQ = p->header[d * 3 + 2];
    uint32_t parent = m[d - firstQ];
    uint32_t len = row_len(g, parent);
    if (len == 0) return;
    if (!g->dup[parent]) {
        uint32_t cur = d, grel = 0;
        int found = 0;
        for (uint32_t j = 0; j < num && !found; j++) {
            cur -= p->requirements[roff + 2 * j];
            if (d - cur == firstQ) { grel = p->requirements[roff + 2 * j + 1]; found = 1; }
        }
        if (found) {
            const uint32_t *rr = g->rrel + parent + 1, *rt = g->rtgt + parent + 1;
            uint32_t lo = 0, hi = len;
            while (lo < hi) {
                uint32_t mid = lo + ((hi - lo) >> 1);
                if (rr[mid] < grel) lo = mid + 1; else hi = mid;
            }
            for (uint32_t e = lo; e < len && rr[e] == grel; e++) {
                if (mc->max_leaves && *(volatile int *)mc->stop) return;
                uint32_t c = rt[e];
                if (!check_candidate(mc, d, m, c)) continue;
                m[d] = c;
                acc->level[d]++;
                expand(mc, d + 1, m, acc);
            }
            return;
        }
    }
    for (uint32_t e = 0; e < len; e++) {
        if (mc->max_leaves && *(volatile int *)mc->stop) return;
        uint32_t c = g->tgt[parent + 1 + e];
        if (!check_candidate(mc, d, m, c)) continue;
        m[d] = c;
        acc->level[d]++;
        expand(mc, d + 1, m, acc);
    }
}

static int row_cmp(const void *a, const void *b, void *w)
{
    uint32_t width = *(const uint32_t *)w;
    const uint32_t *x = (const uint32_t *)a, *y = (const uint32_t *)b;
    for (uint32_t i = 0; i < width; i++)
        if (x[i] != y[i]) return x[i] < y[i] ? -1 : 1;
    return 0;
}

int orc_match(const orc_ccsr *g, const orc_ccsr *qg, const orc_plan *p,
              uint32_t root_lo, uint32_t root_hi, int nthreads, uint64_t max_rows,
              orc_result **out)
{
    return orc_match_capped(g, qg, p, root_lo, root_hi, nthreads, max_rows, 0, out);
}

int orc_match_capped(const orc_ccsr *g, const orc_ccsr *qg, const orc_plan *p,
                     uint32_t root_lo, uint32_t root_hi, int nthreads, uint64_t max_rows,
                     uint64_t max_leaves, orc_result **out)
{
    *out = NULL;
    uint32_t q = p->q;
    if (q == 0 || q > 64) return ORC_ERR_QUERY;
    /* Root candidates: one per data row, in offset order (HnGPU.cu:475-495). */
    uint64_t nroots = 0;
    for (uint64_t i = 0; i < g->size; i += (uint64_t)g->tgt[i] + 1) nroots++;
    uint32_t *roots = (uint32_t *)malloc((nroots ? nroots : 1) * sizeof(uint32_t));
    nroots = 0;
    for (uint64_t i = 0; i < g->size; i += (uint64_t)g->tgt[i] + 1) roots[nroots++] = (uint32_t)i;

    int nt = 1;
#ifdef _OPENMP
    nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    (void)nthreads;
#endif
    thread_acc *accs = (thread_acc *)calloc((size_t)nt, sizeof(thread_acc));
    for (int t = 0; t < nt; t++) accs[t].level = (uint64_t *)calloc(q, sizeof(uint64_t));
    uint64_t leaves = 0;
    int stop = 0;
    match_ctx mc = {g, p, qg, max_rows, max_leaves, &leaves, &stop};

#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
#endif
    for (uint64_t r = 0; r < nroots; r++) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        thread_acc *acc = &accs[tid];
        if (max_leaves && *(volatile int *)&stop) continue;
        uint32_t x = roots[r];
        uint32_t id = orc_offset_to_id(g, x);
        if (id < root_lo || id >= root_hi) continue;
        /* hnInit (HnGPU.cu:95-101): depth-0 requirements are self loops. */
        uint32_t num = p->header[0], roff = p->header[1];
        int ok = 1;
        for (uint32_t j = 0; j < num && ok; j++)
            if (!req_exists(g, x, p->requirements[roff + 2 * j + 1], x)) ok = 0;
        if (!ok) continue;
        uint32_t m[64];
        m[0] = x;
        acc->level[0]++;
        expand(&mc, 1, m, acc);
    }

    orc_result *res = (orc_result *)calloc(1, sizeof(orc_result));
    res->width = q;
    res->level = (uint64_t *)calloc(q, sizeof(uint64_t));
    int overflow = 0;
    uint64_t total_rows = 0;
    for (int t = 0; t < nt; t++) {
        res->count += accs[t].count;
        res->digest[0] += accs[t].d0;
        res->digest[1] += accs[t].d1;
        for (uint32_t d = 0; d < q; d++) res->level[d] += accs[t].level[d];
        total_rows += accs[t].nrows;
        overflow |= accs[t].overflow;
    }
    int rc = ORC_OK;
    if (max_rows) {
        if (overflow || total_rows > max_rows) {
            rc = ORC_ERR_CAP;
        } else {
            res->rows = (uint32_t *)malloc((total_rows ? total_rows : 1) * q * sizeof(uint32_t));
            uint64_t at = 0;
            for (int t = 0; t < nt; t++) {
                if (accs[t].nrows) memcpy(res->rows + at * q, accs[t].rows, accs[t].nrows * q * sizeof(uint32_t));
                at += accs[t].nrows;
            }
            res->nrows = total_rows;
            qsort_r(res->rows, total_rows, q * sizeof(uint32_t), row_cmp, &q);
        }
    }
    for (int t = 0; t < nt; t++) { free(accs[t].level); free(accs[t].rows); }
    free(accs);
    free(roots);
    if (rc) { orc_result_free(res); return rc; }
    *out = res;
    return ORC_OK;
}

void orc_result_free(orc_result *r)
{
    if (!r) return;
    free(r->rows); free(r->level);
    free(r);
}

/* ------------------------------------------------------------------------ */
/* Canonical order and the deterministic per-depth cap                       */
/* ------------------------------------------------------------------------ */

/* The same enumeration as orc_match, in one fixed order: roots by ascending
 * original id, and at every depth the generator's entries of the generator
 * relation by ascending original target id.  The product's bounded-work mode
 * (a per-depth candidate cap, no reference counterpart) is defined over this
 * order: depth k tests only its first max_depth_cand candidates, and the
 * match stops the first time any depth would test one more. */

typedef struct { uint32_t rel, id, off; } canon_ent;

static int canon_cmp(const void *a, const void *b)
{
    const canon_ent *x = (const canon_ent *)a, *y = (const canon_ent *)b;
    if (x->rel != y->rel) return x->rel < y->rel ? -1 : 1;
    if (x->id != y->id) return x->id < y->id ? -1 : 1;
    return (x->off > y->off) - (x->off < y->off);
}

int orc_canon_build(const orc_ccsr *g, orc_canon **out)
{
    *out = NULL;
    orc_canon *cn = (orc_canon *)calloc(1, sizeof(orc_canon));
    uint64_t W = g->size ? g->size : 1;
    cn->size = g->size;
    cn->cid = (uint32_t *)calloc(W, sizeof(uint32_t));
    cn->coff = (uint32_t *)calloc(W, sizeof(uint32_t));
    cn->crel = (uint32_t *)calloc(W, sizeof(uint32_t));
    uint64_t nrows = 0;
    for (uint64_t i = 0; i < g->size; i += (uint64_t)g->tgt[i] + 1) nrows++;
    uint64_t *rows = (uint64_t *)malloc((nrows ? nrows : 1) * sizeof(uint64_t));
    nrows = 0;
    for (uint64_t i = 0; i < g->size; i += (uint64_t)g->tgt[i] + 1) rows[nrows++] = i;
    if (!cn->cid || !cn->coff || !cn->crel || !rows) {
        free(rows);
        orc_canon_free(cn);
        return ORC_ERR_NOMEM;
    }
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        canon_ent *buf = NULL;
        uint32_t bcap = 0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1024)
#endif
        for (uint64_t r = 0; r < nrows; r++) {
            uint64_t o = rows[r];
            uint32_t len = g->tgt[o];
            if (len > bcap) {
                bcap = len;
                buf = (canon_ent *)realloc(buf, (size_t)bcap * sizeof(canon_ent));
            }
            for (uint32_t e = 0; e < len; e++) {
                buf[e].rel = g->rel[o + 1 + e];
                buf[e].off = g->tgt[o + 1 + e];
                buf[e].id = orc_offset_to_id(g, buf[e].off);
            }
            qsort(buf, len, sizeof(canon_ent), canon_cmp);
            for (uint32_t e = 0; e < len; e++) {
                cn->crel[o + 1 + e] = buf[e].rel;
                cn->cid[o + 1 + e] = buf[e].id;
                cn->coff[o + 1 + e] = buf[e].off;
            }
        }
        free(buf);
    }
    free(rows);
    *out = cn;
    return ORC_OK;
}

void orc_canon_free(orc_canon *cn)
{
    if (!cn) return;
    free(cn->cid); free(cn->coff); free(cn->crel);
    free(cn);
}

/* First entry of relation rel in the (relation, id)-ordered copy of the row
 * at offset o (len entries), and the length of that run. */
static uint32_t canon_run(const orc_canon *cn, uint64_t o, uint32_t len, uint32_t rel, uint32_t *n)
{
    const uint32_t *r = cn->crel + o + 1;
    uint32_t lo = 0, hi = len;
    while (lo < hi) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if (r[mid] < rel) lo = mid + 1; else hi = mid;
    }
    uint32_t e = lo;
    while (e < len && r[e] == rel) e++;
    *n = e - lo;
    return lo;
}

typedef struct {
    match_ctx mc;
    const orc_canon *cn;
    uint64_t cap;             /* candidates per depth (0 = no cap) */
    uint64_t cand[64];        /* candidates tested per depth (roots at depth 0) */
    int stop;
    thread_acc acc;
} canon_ctx;

/* Relation of the generator of depth d: the requirement at distance firstQ. */
static uint32_t gen_rel(const orc_plan *p, uint32_t d)
{
    uint32_t num = p->header[d * 3 + 0], roff = p->header[d * 3 + 1], firstQ = p->header[d * 3 + 2];
    uint32_t cur = d;
    for (uint32_t j = 0; j < num; j++) {
        cur -= p->requirements[roff + 2 * j];
        if (d - cur == firstQ) return p->requirements[roff + 2 * j + 1];
    }
    return 0;
}

static void canon_expand(canon_ctx *c, uint32_t d, uint32_t *m)
{
    const orc_ccsr *g = c->mc.g;
    const orc_plan *p = c->mc.p;
    if (d == p->q) { emit_leaf(&c->mc, m, &c->acc); return; }
    uint32_t parent = m[d - p->header[d * 3 + 2]];
    uint32_t len = row_len(g, parent), n = 0;
    if (len == 0) return;
    uint32_t b = canon_run(c->cn, parent, len, gen_rel(p, d), &n);
    for (uint32_t e = b; e < b + n; e++) {
        if (c->cap && c->cand[d] == c->cap) { c->stop = 1; return; }
        c->cand[d]++;
        uint32_t x = c->cn->coff[parent + 1 + e];
        if (!check_candidate(&c->mc, d, m, x)) continue;
        m[d] = x;
        c->acc.level[d]++;
        canon_expand(c, d + 1, m);
        if (c->stop) return;
    }
}

int orc_match_canon(const orc_ccsr *g, const orc_canon *cn, const orc_ccsr *qg, const orc_plan *p,
                    uint32_t root_lo, uint32_t root_hi, uint64_t max_depth_cand,
                    orc_result **out, int *capped, uint64_t *cand)
{
    *out = NULL;
    *capped = 0;
    uint32_t q = p->q;
    if (q == 0 || q > 64) return ORC_ERR_QUERY;
    if (cn->size != g->size) return ORC_ERR_ARG;
    canon_ctx *c = (canon_ctx *)calloc(1, sizeof(canon_ctx));
    c->mc.g = g;
    c->mc.p = p;
    c->mc.qg = qg;
    c->cn = cn;
    c->cap = max_depth_cand;
    c->acc.level = (uint64_t *)calloc(q, sizeof(uint64_t));
    /* Root candidates (HnGPU.cu:475-495 takes every row): vertices with a row
     * that holds an entry of depth 1's generator relation -- the others have
     * no child and so no embedding -- in ascending id order. */
    uint32_t r1 = q > 1 ? gen_rel(p, 1) : 0;
    uint32_t num = p->header[0], roff = p->header[1];
    uint32_t hi = root_hi < g->count ? root_hi : g->count;
    for (uint32_t x = root_lo; x < hi && !c->stop; x++) {
        uint32_t off = g->verticeLocs[x];
        uint32_t len = row_len(g, off), n = 0;
        if (len == 0) continue;
        if (q > 1) {
            canon_run(cn, off, len, r1, &n);
            if (n == 0) continue;
        }
        if (c->cap && c->cand[0] == c->cap) { c->stop = 1; break; }
        c->cand[0]++;
        int ok = 1;  /* hnInit (HnGPU.cu:95-101): depth-0 requirements are self loops */
        for (uint32_t j = 0; j < num && ok; j++)
            if (!req_exists(g, off, p->requirements[roff + 2 * j + 1], off)) ok = 0;
        if (!ok) continue;
        uint32_t m[64];
        m[0] = off;
        c->acc.level[0]++;
        canon_expand(c, 1, m);
    }
    orc_result *res = (orc_result *)calloc(1, sizeof(orc_result));
    res->width = q;
    res->count = c->acc.count;
    res->digest[0] = c->acc.d0;
    res->digest[1] = c->acc.d1;
    res->level = c->acc.level;
    *capped = c->stop;
    if (cand) memcpy(cand, c->cand, q * sizeof(uint64_t));
    free(c);
    *out = res;
    return ORC_OK;
}

int orc_solve(const orc_edges *data, const orc_edges *query, uint32_t mode,
              uint32_t root_lo, uint32_t root_hi, int nthreads, uint64_t max_rows,
              orc_result **out)
{
    *out = NULL;
    orc_ccsr *dg = NULL, *qg = NULL;
    orc_plan *p = NULL;
    int rc = orc_ccsr_build(data, mode, &dg);
    if (!rc) rc = orc_ccsr_build(query, mode, &qg);
    if (!rc) rc = orc_plan_build(qg, &p);
    if (!rc) rc = orc_match(dg, qg, p, root_lo, root_hi, nthreads, max_rows, out);
    orc_plan_free(p);
    orc_ccsr_free(qg);
    orc_ccsr_free(dg);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Soundness check (HnFunction.java:182-194) and output.fs                  */
/* ------------------------------------------------------------------------ */

typedef struct { uint32_t s, d, l; } trip;

static int trip_cmp(const void *a, const void *b)
{
    const trip *x = (const trip *)a, *y = (const trip *)b;
    if (x->s != y->s) return x->s < y->s ? -1 : 1;
    if (x->d != y->d) return x->d < y->d ? -1 : 1;
    return (x->l > y->l) - (x->l < y->l);
}

uint64_t orc_check_sound(const orc_edges *query, const orc_edges *data,
                         const uint32_t *rows, uint64_t nrows, uint32_t width)
{
    trip *t = (trip *)malloc((data->E ? data->E : 1) * sizeof(trip));
    for (uint64_t i = 0; i < data->E; i++) { t[i].s = data->src[i]; t[i].d = data->dst[i]; t[i].l = data->elabel[i]; }
    qsort(t, data->E, sizeof(trip), trip_cmp);
    uint64_t bad = nrows;
    for (uint64_t r = 0; r < nrows && bad == nrows; r++) {
        const uint32_t *f = rows + r * width;
        for (uint64_t e = 0; e < query->E; e++) {
            if (query->src[e] >= width || query->dst[e] >= width) { bad = r; break; }
            trip key = {f[query->src[e]], f[query->dst[e]], query->elabel[e]};
            if (!bsearch(&key, t, data->E, sizeof(trip), trip_cmp)) { bad = r; break; }
        }
    }
    free(t);
    return bad;
}

int orc_fs_write(const char *path, const uint32_t *rows, uint64_t nrows, uint32_t width)
{
    if (nrows > 0xFFFFFFFFull) return ORC_ERR_CAP;
    FILE *f = fopen(path, "wb");
    if (!f) return ORC_ERR_MISSING_FILE;
    uint32_t hdr[2] = {(uint32_t)nrows, width};
    fwrite(hdr, sizeof(uint32_t), 2, f);
    fwrite(rows, sizeof(uint32_t), nrows * width, f);
    fclose(f);
    return ORC_OK;
}

int orc_fs_read(const char *path, uint32_t **rows, uint64_t *nrows, uint32_t *width)
{
    FILE *f = fopen(path, "rb");
    if (!f) return ORC_ERR_MISSING_FILE;
    uint32_t hdr[2];
    if (fread(hdr, sizeof(uint32_t), 2, f) != 2) { fclose(f); return ORC_ERR_INVALID_DATA; }
    uint64_t n = (uint64_t)hdr[0] * hdr[1];
    uint32_t *r = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    if (fread(r, sizeof(uint32_t), n, f) != n) { free(r); fclose(f); return ORC_ERR_INVALID_DATA; }
    fclose(f);
    *rows = r;
    *nrows = hdr[0];
    *width = hdr[1];
    return ORC_OK;
}
