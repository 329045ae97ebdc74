// plan_dev.h -- plain-old-data query plan read by the kernels.  Shared by the
// host planner and the HIP code; no host or device library types inside.
#pragma once

#include <stdint.h>

#define GFSM_MAXQ 32
#define GFSM_MAXCHK 31

// Work at depth d >= 1 (the query vertex placed d-th):
//   candidates = segment (f(anc at depth d - gen_dist), gen_rel);
//   a candidate w survives when it differs from every placed vertex and, for
//   each check j, w lies in segment (f(anc at depth d - chk_dist[j]), chk_rel[j])
//   (chk_dist 0 = w's own segment: a query self loop);
//   nxt_* is the generator of depth d+1 (its segment length is the number of
//   children a survivor will have; survivors with none are dropped).
struct DepthPlanDev {
    uint32_t gen_dist;
    uint32_t gen_rel;
    uint32_t nchk;
    uint32_t nxt_dist;
    uint32_t nxt_rel;
    uint32_t chk_dist[GFSM_MAXCHK];
    uint32_t chk_rel[GFSM_MAXCHK];
    // Dense segment tables ({begin, end} per data vertex, see DevGraph) of
    // nxt_rel and of the first four chk_rel, filled in per run where one
    // exists; null: the segment hash table.
    const void *nxt_seg;
    const void *chk_seg[4];
};

struct QueryPlanDev {
    uint32_t q;
    uint32_t root_rel;                   // generator relation of depth 1
    uint32_t nroot_chk;                  // self-loop relations of the root
    uint32_t root_chk_rel[GFSM_MAXCHK];
    uint32_t col[GFSM_MAXQ];             // original query id placed at depth d
    uint32_t dcol[GFSM_MAXQ];            // depth whose vertex is column c
    DepthPlanDev depth[GFSM_MAXQ];
};
