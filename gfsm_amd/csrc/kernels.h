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

enum ExpandKind : uint32_t {
    EXPAND_MID = 0,        // keep survivors with children, write the next level
    EXPAND_LAST_COUNT = 1, // last depth: count (and digest)
    EXPAND_LAST_ROWS = 2,  // last depth: write rows in original ids (and digest)
};

// A frontier level holding nodes with k+1 placed vertices (depths 0..k):
//   vec[i*(k+1) + j] = data vertex placed at depth j by node i,
//   off[i] = exclusive prefix of the nodes' next-depth candidate counts
//            (off[S] = total), strictly increasing (every node has >= 1),
//   gbeg[i] = begin (in nbr) of the node's next-depth generator segment.
// The next depth's candidate g of node i is nbr[gbeg[i] + (g - off[i])].
struct ExpandArgs {
    DevGraphView g;
    DepthPlanDev dp;
    uint32_t d, q;
    uint32_t kind;
    uint32_t digest;
    uint32_t col[GFSM_MAXQ];   // original query id of depth k
    uint32_t dcol[GFSM_MAXQ];  // depth of column c
    const uint32_t *in_vec;    // level d-1 (width d)
    const uint64_t *in_off;    // [S_prev+1]
    const uint32_t *in_gbeg;   // [S_prev]
    uint64_t S_prev;
    uint64_t g_begin, g_end;   // candidate range of this launch
    uint32_t *out_vec;         // level d (width d+1), EXPAND_MID
    uint64_t *out_off;         // [cap+1]
    uint32_t *out_gbeg;        // [cap]
    uint32_t *rows;            // [cap*q], EXPAND_LAST_ROWS
    uint64_t cap;
    ScanWS ws;
    Totals *tot;
    uint64_t ntiles;
    uint32_t items;            // candidates per thread: 1, 2, 4 or 8 (tile = 256 * items)
    uint32_t need_nxt;         // stage the next generator per node (MID, generator not the new vertex)
};

struct RootArgs {
    DevGraphView g;
    const uint32_t *cand_vtx;  // root candidates (a slice of the relation index)
    const uint32_t *cand_beg;  // their depth-1 generator segment begin
    const uint32_t *cand_cnt;  // ... and length
    uint64_t n;
    uint32_t nchk;             // self-loop relations of the root
    uint32_t chk_rel[GFSM_MAXCHK];
    uint32_t *out_vec;         // level 0 (width 1)
    uint64_t *out_off;
    uint32_t *out_gbeg;
    uint64_t cap;
    ScanWS ws;
    Totals *tot;
    uint64_t ntiles;
};

constexpr uint32_t kTile = 256;     // threads per workgroup
constexpr uint32_t kItems = 8;      // default candidates per thread per tile
constexpr uint32_t kRootItems = 4;  // root level
constexpr uint32_t kRootTile = kTile * kRootItems;

// LDS bytes the expand kernel needs with `nchk` ancestor checks, per-node
// next-generator staging (`need_nxt`) and `items` candidates per thread.
size_t expand_lds_bytes(uint32_t nchk, uint32_t need_nxt, uint32_t items);

hipError_t launch_root(const RootArgs &a, uint32_t grid, hipStream_t s);
hipError_t launch_expand(const ExpandArgs &a, uint32_t grid, hipStream_t s);

// Chained-scan self test: out[i] = exclusive prefix of vals (u64), and the
// indices i with vals[i] != 0 compacted in order into idx; totals in tot.
hipError_t launch_scan_test(const uint32_t *vals, uint64_t n, uint64_t *out, uint32_t *idx,
                            ScanWS ws, Totals *tot, uint32_t grid, hipStream_t s);
constexpr uint32_t kScanTestItems = 2;

// Zero the control block of a workspace (device kernel; safe under capture).
hipError_t launch_ws_reset(ScanWS ws, Totals *tot, hipStream_t s);
// Zero every tile word (needed once per 2^kEpochBits launches on a lane).
hipError_t launch_ws_clear(ScanWS ws, hipStream_t s);

}  // namespace gfsm
