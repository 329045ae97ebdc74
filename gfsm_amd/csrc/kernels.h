// kernels.h -- host-visible description of the matcher kernels: argument
// blocks and launch wrappers (kernels.hip).  Plain structs only, so host
// translation units need no device headers.
#pragma once

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#include "plan_dev.h"

namespace gfsm {

// Device view of a resident data graph (see HostGraph for the layout).
// Segments are found through `seg`, an open-addressing hash table keyed by
// (vertex, relation): slot = {vertex, relation, begin, end}, vertex ~0u =
// empty, at most half full, linear probing from seg_hash(vertex, relation).
// One probe usually suffices where a directory search takes log2(relations
// of the vertex) dependent loads.
constexpr uint32_t kSegEmpty = 0xFFFFFFFFu;
struct DevGraphView {
    const uint32_t *nbr;      // [nent]   targets, grouped per (vertex, relation), ascending
    const uint32_t *dir_ptr;  // [V+1]    directory range of each vertex
    const uint32_t *dir;      // [2*(ndir+1)] {relation, begin} pairs + sentinel
    const uint4 *seg;         // [seg_mask+1] segment hash table
    uint32_t seg_mask;
    uint32_t V;
    uint32_t nent;            // entries in nbr (bounds guard)
    uint32_t pad;
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

// Results of one step, written on the device, read by the host afterwards.
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
    EXPAND_LAST_BARE = 3,  // last depth, count only, no checks and no digest: a lean kernel
                           // that fits twice the resident workgroups of the others
};
constexpr uint32_t kBarePoolWords = 2048;  // LDS pool of the lean last-depth kernel

// Data movement of one expand step over a candidate range of depth d.
// A frontier level holding nodes with k+1 placed vertices (depths 0..k):
//   vec[i*(k+1) + j] = data vertex placed at depth j by node i,
//   off[i] = exclusive prefix of the nodes' next-depth candidate counts
//            (off[S] = total), strictly increasing (every node has >= 1),
//   gbeg[i] = begin (in nbr) of the node's next-depth generator segment.
// The next depth's candidate g of node i is nbr[gbeg[i] + (g - off[i])].
struct StepIO {
    const uint32_t *in_vec;    // level d-1 (width d)
    const uint64_t *in_off;    // [S_prev+1]
    const uint32_t *in_gbeg;   // [S_prev]
    uint64_t S_prev;
    uint64_t g_begin, g_end;   // candidate range of this step
    uint32_t *out_vec;         // level d (width d+1), EXPAND_MID
    uint64_t *out_off;         // [cap+1]
    uint32_t *out_gbeg;        // [cap]
    uint32_t *rows;            // [cap*q], EXPAND_LAST_ROWS
    uint64_t cap;
    Totals *tot;
    uint64_t ntiles;
    uint32_t d, q;
    uint32_t kind;
    uint32_t digest;
    uint32_t need_nxt;         // stage the next generator per node (MID, generator not the new vertex)
    uint32_t pad;
    const uint64_t *tile_node; // [2*ntiles] first and last frontier node of each tile (k_tile_index),
                               // or null: each tile searches in_off itself
    uint64_t *host_kc;         // or null: fine-grained host words that also receive {keep, cnt},
                               // so the host reads a step's totals without a copy
};
constexpr uint32_t kDenseChk = 4;  // checked relations per depth with a dense-table slot (DepthPlanDev)

// Root step: the candidates of depth 0 (a slice of the relation index).
struct RootIO {
    const uint32_t *cand_vtx;  // root candidates
    const uint32_t *cand_beg;  // their depth-1 generator segment begin
    const uint32_t *cand_cnt;  // ... and length
    uint64_t n;
    uint32_t *out_vec;         // level 0 (width 1)
    uint64_t *out_off;
    uint32_t *out_gbeg;
    uint64_t cap;
    Totals *tot;
    uint64_t ntiles;
    uint64_t *host_kc;         // as StepIO::host_kc
};

struct ExpandArgs {
    DevGraphView g;
    DepthPlanDev dp;
    uint32_t col[GFSM_MAXQ];   // original query id of depth k
    uint32_t dcol[GFSM_MAXQ];  // depth of column c
    StepIO io;
    ScanWS ws;
    uint32_t items;            // candidates per thread: 1, 2, 4 or 8 (tile = 256 * items)
    uint32_t pool_words;       // LDS words for per-node staging (checked segments, rows)
};

struct RootArgs {
    DevGraphView g;
    RootIO io;
    uint32_t nchk;             // self-loop relations of the root
    uint32_t chk_rel[GFSM_MAXCHK];
    ScanWS ws;
};

// One-vertex query (the reference's root level with no edges to expand,
// HnGPU.cu:475-495): every vertex id in [lo, hi) that has an out-entry and
// carries each self-loop relation.  Rows mode writes the kept ids in id
// order (chained scan) to rows[0, cap); count mode sums count and digest.
struct SingleArgs {
    DevGraphView g;
    uint32_t lo, hi;
    uint32_t nchk;
    uint32_t chk_rel[GFSM_MAXCHK];
    uint32_t *rows;            // rows mode: [cap] kept ids, else null
    uint64_t cap;
    Totals *tot;               // keep (rows mode) or count and digest
    uint64_t ntiles;
    ScanWS ws;
};
hipError_t launch_single(const SingleArgs &a, uint32_t grid, hipStream_t s);

constexpr uint32_t kTile = 256;     // threads per workgroup
constexpr uint32_t kItems = 8;      // default candidates per thread per tile
constexpr uint32_t kRootItems = 4;  // root level (single-query path)
constexpr uint32_t kRootTile = kTile * kRootItems;
constexpr uint32_t kPoolWords = 7680;  // default staging pool (30 KB: four workgroups fit a CU's LDS)

// Dynamic LDS bytes of the expand and batch-step kernels for `items`
// candidates per thread and a staging pool of `pool_words` words.
size_t expand_lds_bytes(uint32_t items, uint32_t pool_words);
constexpr size_t kMaxDynLds = 150 * 1024;

hipError_t launch_root(const RootArgs &a, uint32_t grid, hipStream_t s);
hipError_t launch_expand(const ExpandArgs &a, uint32_t grid, hipStream_t s);
// Dense segment table of one relation: out[v] = {begin, end} of v's segment
// of that relation ({0, 0} when v has none).  vtx/beg/cnt are the
// relation's slice of the relation index; out has V entries.
hipError_t launch_dense_seg(const uint32_t *vtx, const uint32_t *beg, const uint32_t *cnt, uint64_t n, uint32_t V,
                            uint2 *out, hipStream_t s);
// Frontier nodes of every tile of an expand step, searched ahead of it with
// the whole GPU: out[2t] = the node holding candidate g_begin + t * tc,
// out[2t+1] = the node holding the tile's last candidate (offsets must be
// strictly increasing, as every level's are).  A tile then reads two words
// instead of walking the offsets with dependent loads.
hipError_t launch_tile_index(const uint64_t *off, uint64_t S, uint64_t g_begin, uint64_t g_end, uint64_t tc,
                             uint64_t ntiles, uint64_t *out, hipStream_t s);

// ---------------------------------------------------------------------------
// Batched matching: many queries share every step.  Each query owns a slot in
// device memory; a scheduler kernel advances every slot by one step of its
// own depth-first chunk walk and lays the slots' tiles end to end; one
// persistent step kernel then works through all tiles of all slots, handed
// out by ticket (a device work queue of (slot, tile) items).  A slot's chained
// scan spans only its own tiles.  No host round trip is needed between
// steps, so runs of steps are captured in a hipGraph and replayed.
enum SlotState : uint32_t {
    SLOT_FREE = 0,      // unused
    SLOT_NEW = 1,       // admitted by the host; the root step is scheduled next
    SLOT_RUN = 2,       // running
    SLOT_DONE = 3,      // finished: count and digest are final
    SLOT_CAP = 4,       // the work cap was reached: partial count
    SLOT_ERROR = 5,     // a step shape check failed (see error)
};
constexpr uint32_t STEP_NONE = 7;
constexpr uint32_t STEP_ROOT = 8;
constexpr uint32_t kMaxSlots = 1024;

// Per-slot summary the host polls (one small copy for all slots).
struct SlotReport {
    uint32_t state;
    uint32_t steps;
    uint64_t count, dg0, dg1;
};

// Statistics of a finished slot, written by the scheduler when the slot
// finishes and copied with the reports, so a harvest needs no copy of the
// slot table.
struct SlotFinal {
    uint32_t overflows, pad;
    uint64_t level[GFSM_MAXQ], cand[GFSM_MAXQ];
};

struct BatchSlot {
    // Written by the host at admission, read-only on the device.
    const QueryPlanDev *plan;     // device copy of the query plan
    const uint32_t *root_vtx, *root_beg, *root_cnt;  // root candidates
    uint64_t nroots;
    uint32_t *vec[GFSM_MAXQ];     // level buffers (the slot's arena region)
    uint64_t *off[GFSM_MAXQ];
    uint32_t *gbeg[GFSM_MAXQ];
    uint64_t cap[GFSM_MAXQ];
    uint64_t max_cand;            // work cap (0 = none)
    uint64_t max_depth_cand;      // deterministic per-depth cap (0 = none)
    uint32_t q, digest;
    uint32_t state;               // SLOT_NEW at admission
    uint32_t error;
    // Depth-first chunk walk, maintained by the scheduler.
    uint32_t step;                // STEP_NONE, STEP_ROOT or an ExpandKind
    uint32_t d;                   // depth of the current step
    uint64_t root_next;           // next root candidate
    uint64_t lo[GFSM_MAXQ], hi[GFSM_MAXQ];  // remaining candidate range per depth
    uint64_t S[GFSM_MAXQ];        // nodes in level k
    uint64_t chunk[GFSM_MAXQ];    // candidates per step per depth (chunk[0]: roots)
    uint64_t cur_lo, cur_hi;      // range of the scheduled step
    uint64_t tile0;               // global index of its first tile
    uint64_t cand_done;
    uint64_t count, dg0, dg1;     // result (embeddings, digest sums)
    uint64_t level[GFSM_MAXQ], cand[GFSM_MAXQ];  // statistics
    uint32_t steps, overflows;
    Totals tot;                   // totals of the scheduled step
    // Arguments of the scheduled step (written by the scheduler).
    StepIO io;
    RootIO rio;
};

struct BatchCtl {
    uint64_t total_tiles;         // tiles of the scheduled step, all slots
    uint64_t step_cand;           // candidates of one slot in one step (set by the host
                                  // between replays; 0 = max_step_cand)
    uint32_t running;             // slots in SLOT_NEW or SLOT_RUN after scheduling
    uint32_t n_admit;             // admission entries waiting (set by the host, cleared by the scheduler)
    uint32_t sched_done;          // scheduler workgroups finished (the last one lays out the tiles
                                  // and resets it)
    uint32_t pad;
};

// One admission, written by the host into a contiguous queue between
// replays; the scheduler installs it into its slot.
struct AdmitEntry {
    uint32_t slot;
    uint32_t pad[3];
    BatchSlot s;
    QueryPlanDev plan;
};

// Batch configuration shared by the scheduler and the step kernel.
struct BatchCfg {
    BatchSlot *slots;
    QueryPlanDev *plans;          // [nslots], slot s's plan (BatchSlot::plan points here)
    const AdmitEntry *admit;      // [nslots] admission queue
    SlotReport *report;           // [nslots]
    SlotFinal *final;             // [nslots], written when a slot finishes
    uint64_t *tile_start;         // [nslots + 1]
    uint64_t *tile_node;          // [2 * tiles of a step] frontier nodes of every expand tile (k_batch_tile_index)
    BatchCtl *ctl;
    uint32_t nslots;
    uint32_t items;               // candidates per thread (tile = 256 * items)
    uint32_t pool_words;
    uint32_t pad;
    uint64_t max_step_cand;       // upper bound of BatchCtl::step_cand (workspace size)
    uint64_t max_degree;          // of the data graph (scan value bound)
    DevGraphView g;
    ScanWS ws;
};

// One batch step: the scheduler (one workgroup) consumes every slot's last
// step, schedules its next one and lays out the tiles; the step kernel
// (`grid` persistent workgroups) runs them.  Both read only `cfg` and device
// memory, so the pair can be captured and replayed.
hipError_t launch_batch_schedule(const BatchCfg &cfg, hipStream_t s);
hipError_t launch_batch_step(const BatchCfg &cfg, uint32_t grid, hipStream_t s);
// Between the scheduler and the step kernel: the frontier nodes of every
// expand tile of the scheduled step, searched by the whole GPU (one wave per
// tile), so the step's tiles read two words instead of searching.
hipError_t launch_batch_tile_index(const BatchCfg &cfg, hipStream_t s);

// Chained-scan self test: out[i] = exclusive prefix of vals (u64), and the
// indices i with vals[i] != 0 compacted in order into idx; totals in tot.
hipError_t launch_scan_test(const uint32_t *vals, uint64_t n, uint64_t *out, uint32_t *idx,
                            ScanWS ws, Totals *tot, uint32_t grid, hipStream_t s);
constexpr uint32_t kScanTestItems = 2;

// Build the segment hash table of a graph: fill `slots` (mask+1 entries)
// with empty slots, then insert every directory entry.
hipError_t launch_seg_hash_build(const uint32_t *dir_ptr, const uint32_t *dir, uint32_t V, uint4 *slots,
                                 uint32_t mask, hipStream_t s);

// Device scratch of the allocation-free match: a file-scope array of
// kNoAllocScratchBytes in kernels.hip, one per GPU, held until the process
// ends.  Layout: control block (256 B), totals (256 B), the scan words of
// kNoAllocTiles tiles (two arrays), the tile index (2 words per tile), then
// from kNoAllocWorkBytes on the level arena of kNoAllocLevelBytes.
constexpr uint64_t kNoAllocScratchBytes = 64ull << 20;
constexpr uint64_t kNoAllocTiles = 1ull << 16;
constexpr uint64_t kNoAllocWorkBytes = 4ull << 20;
constexpr uint64_t kNoAllocLevelBytes = kNoAllocScratchBytes - kNoAllocWorkBytes;
static_assert(512 + 32 * kNoAllocTiles <= kNoAllocWorkBytes, "scratch workspace layout");
// Device address of `device`'s scratch array: looked up with
// hipGetSymbolAddress on the first call for a device (that device made
// current; its workspace part cleared on `s`), then cached.
hipError_t noalloc_scratch(int device, hipStream_t s, char **ptr);

// Traffic calibration (kernels of known footprint).  mode 0: dst[i] =
// src[i] + 1 over n words (n words read and written); mode 1: one word from
// each of n distinct 128-byte lines of src (n a power of two, src holding 32n
// words) written to dst[0..n); mode 2: words 0 and 16 of each of those lines
// (both 64-byte halves) written to dst[0..2n).
hipError_t launch_calibrate(uint32_t mode, const uint32_t *src, uint32_t *dst, uint64_t n, hipStream_t s);

// Zero the control block of a workspace (device kernel; safe under capture).
hipError_t launch_ws_reset(ScanWS ws, Totals *tot, hipStream_t s);
// Zero every tile word (needed once per 2^kEpochBits launches on a lane).
hipError_t launch_ws_clear(ScanWS ws, hipStream_t s);

}  // namespace gfsm
