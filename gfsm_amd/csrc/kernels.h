// kernels.h -- host-visible description of the matcher kernels: argument
// blocks and launch wrappers (kernels.hip).  Plain structs only, so host
// translation units need no device headers.
#pragma once

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#include "plan_dev.h"

namespace gfsm {

// Device view of a resident data graph (see HostGraph for the layout).
struct DevGraphView {
    const uint32_t *nbr;      // [nent]   targets, grouped per (vertex, relation), ascending
    const uint32_t *dir_ptr;  // [V+1]    directory range of each vertex
    const uint32_t *dir;      // [2*(ndir+1)] {relation, begin} pairs + sentinel
    const uint32_t *rel_vtx;  // [nidx]   vertices owning a segment, grouped per relation
    const uint32_t *rel_cnt;  // [nidx]   ... and that segment's length
    uint32_t V;
    uint32_t nent;            // entries in nbr (bounds guard)
};

// Chained-scan workspace of one launch lane (one stream).  Each tile
// publishes two self-describing 64-bit words -- kept nodes and next-depth
// candidates -- packed as [63:62] status (0 none, 1 aggregate, 2 inclusive
// prefix), [61:46] launch epoch, [45:0] value.  One word carries status and
// value together, so relaxed agent-scope atomics suffice: no release/acquire
// fences, which on a multi-XCD part would write back / invalidate L2.
// ctl = {ticket, done, epoch, unused}.
constexpr uint32_t kEpochBits = 16;
constexpr uint64_t kValueMask = (1ull << 46) - 1;
struct ScanWS {
    uint64_t *wk;     // [max_tiles] kept-node words
    uint64_t *wc;     // [max_tiles] candidate words
    uint32_t *ctl;    // [4]
    uint64_t max_tiles;
};

// Results of a launch, written on the device, read by the host afterwards.
struct Totals {
    uint64_t keep;    // nodes kept (last level: embeddings when rows are written)
    uint64_t cnt;     // candidates of the next depth
    uint64_t count;   // last level, count-only: embeddings
    uint64_t dg0, dg1;// digest sums
    uint64_t pad[3];
};

struct LevelPtrs {
    const uint32_t *vtx[GFSM_MAXQ];
    const uint32_t *par[GFSM_MAXQ];
};

enum ExpandKind : uint32_t {
    EXPAND_MID = 0,        // keep survivors with children, write M-nodes + next offsets
    EXPAND_LAST_COUNT = 1, // last depth: count (and digest)
    EXPAND_LAST_ROWS = 2,  // last depth: write rows in original ids (and digest)
};

struct ExpandArgs {
    DevGraphView g;
    DepthPlanDev dp;
    uint32_t d, q;
    uint32_t kind;
    uint32_t digest;
    uint32_t col[GFSM_MAXQ];   // original query id of depth k
    uint32_t dcol[GFSM_MAXQ];  // depth of column c
    const uint64_t *off_prev;  // [S_prev+1] exclusive candidate offsets of level d-1
    uint64_t S_prev;
    uint64_t g_begin, g_end;   // candidate range of this launch
    LevelPtrs lv;              // levels 0 .. d-1
    uint32_t *out_vtx, *out_par;
    uint64_t *out_off;         // [cap+1]
    uint32_t *rows;            // [cap*q] (EXPAND_LAST_ROWS)
    uint64_t cap;
    ScanWS ws;
    Totals *tot;
    uint64_t ntiles;
};

struct RootArgs {
    DevGraphView g;
    const uint32_t *cand_vtx;  // root candidates (a slice of rel_vtx)
    const uint32_t *cand_cnt;  // their depth-1 generator segment lengths
    uint64_t n;
    uint32_t nchk;             // self-loop relations of the root
    uint32_t chk_rel[GFSM_MAXCHK];
    uint32_t *out_vtx;
    uint64_t *out_off;
    uint64_t cap;
    ScanWS ws;
    Totals *tot;
    uint64_t ntiles;
};

constexpr uint32_t kTile = 256;   // threads per workgroup
constexpr uint32_t kItems = 2;    // candidates per thread per tile
constexpr uint32_t kTileCand = kTile * kItems;  // candidates per tile

// LDS bytes the expand kernel needs for depth d with `nchk` ancestor checks.
size_t expand_lds_bytes(uint32_t d, uint32_t nchk);

hipError_t launch_root(const RootArgs &a, uint32_t grid, hipStream_t s);
hipError_t launch_expand(const ExpandArgs &a, uint32_t grid, hipStream_t s);

// Chained-scan self test: out[i] = exclusive prefix of vals (u64), and the
// indices i with vals[i] != 0 compacted in order into idx; totals in tot.
hipError_t launch_scan_test(const uint32_t *vals, uint64_t n, uint64_t *out, uint32_t *idx,
                            ScanWS ws, Totals *tot, uint32_t grid, hipStream_t s);

// Zero the control block of a workspace (device kernel; safe under capture).
hipError_t launch_ws_reset(ScanWS ws, Totals *tot, hipStream_t s);
// Zero every tile word (needed once per 2^kEpochBits launches on a lane).
hipError_t launch_ws_clear(ScanWS ws, hipStream_t s);

}  // namespace gfsm
