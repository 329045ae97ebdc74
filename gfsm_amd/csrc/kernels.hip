// kernels.hip -- CDNA4 (gfx950) kernels of the forest-based matcher.
//
// A step expands one candidate range of one depth.  A workgroup of 256
// threads (4 wave64s) owns a tile of 256 * I consecutive candidates of the
// step (I = 1, 2, 4 or 8 per thread, striped so that loads stay coalesced):
//   1. one wave finds the frontier nodes whose candidate ranges cover the
//      tile (a 32-ary search per tile end over their exclusive offsets) and
//      marks where each node's candidates start in an LDS bitmap; the nodes'
//      placed vertices, checked segments and next generators are staged in an
//      LDS pool when they fit,
//   2. each candidate finds its node from the bitmap (mbcnt of a wave-uniform
//      word), reads itself coalesced from the node's generator segment, is
//      checked for injectivity against the node's placed vertices and
//      searched in the segments of the checked ancestors -- the I searches of
//      a thread advance in lockstep, so each thread keeps I loads in flight,
//   3. survivors are compacted with wave64 ballot/mbcnt counts and DPP prefix
//      sums of their next-depth candidate counts, chained across workgroups
//      by a decoupled look-back, so that the next level -- rows, candidate
//      offsets and generator segment begins -- comes out of the same pass.
// This replaces hnGen + hnCheck + CUB ExclusiveSum + hnRemove + the degree
// scan of one reference depth (HnGPU.cu:114-262, 309-349, 371-439).
//
// The single-query kernels (k_root, k_expand) run one step per launch.  The
// batch kernels run one step of every admitted query per launch: a scheduler
// kernel advances each query's depth-first chunk walk on the device, and a
// persistent step kernel takes (query, tile) work items by ticket.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kernels.h"

namespace gfsm {

#define DEV __device__ __forceinline__

// Pointers into the global address space.  Buffer pointers that a kernel
// loads from memory (the batch slots) are generic, and accesses through
// generic pointers become flat instructions; these keep them global.
template <class T>
using gptr_t = __attribute__((address_space(1))) T *;
template <class T>
DEV gptr_t<T> gp(T *p)
{
    return (gptr_t<T>)p;
}

constexpr uint32_t kWaves = kTile / 64;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));  // one 16-byte store
constexpr uint32_t kNodeLk = 4;  // node-setup lookups a thread keeps in flight
constexpr uint32_t kStageLd = 8; // row words a thread loads before storing them to LDS

// ---------------------------------------------------------------------------
// wave64 primitives
// ---------------------------------------------------------------------------

DEV uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// A copy of a workgroup-uniform struct read from memory, every word through
// readfirstlane: the copy lives in scalar registers, where the compiler would
// otherwise keep the loaded words in vector registers (and spill them).
template <class T>
DEV T uniform_copy(const T &src)
{
    static_assert(sizeof(T) % 4 == 0, "words");
    T out;
    const uint32_t *a = reinterpret_cast<const uint32_t *>(&src);
    uint32_t *b = reinterpret_cast<uint32_t *>(&out);
#pragma unroll
    for (uint32_t i = 0; i < sizeof(T) / 4; i++) b[i] = (uint32_t)__builtin_amdgcn_readfirstlane((int)a[i]);
    return out;
}

// Number of set bits of `mask` in lanes below this one.
DEV uint32_t mask_prefix(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Inclusive prefix sum over the wave with DPP moves: Hillis-Steele shifts of
// 1, 2, 4, 8 lanes inside each 16-lane row, then row_bcast:15 adds each row's
// last lane into the next row and row_bcast:31 adds lane 31 into rows 2-3.
// Disabled or out-of-row source lanes contribute the `old` operand, 0.
DEV uint32_t wave_incl_scan(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}

DEV uint64_t wave_sum64(uint64_t v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

DEV uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// ---------------------------------------------------------------------------
// graph access, I lookups per thread in lockstep
// ---------------------------------------------------------------------------

// Hash of a (vertex, relation) key (the "lowbias32" integer mix).
DEV uint32_t seg_hash(uint32_t x, uint32_t rel)
{
    uint32_t h = x * 0x9E3779B1u + rel * 0x85EBCA6Bu;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}

// seg[k] = [beg, end) of the (x[k], rel) segment, {0, 0} when x[k] has none
// or act[k] is false: linear probing in the segment hash table, the I
// lookups issuing their probes together.
template <uint32_t I>
DEV void seg_lookup_r(const DevGraphView &g, const uint32_t (&x)[I], const uint32_t (&rel)[I], const bool (&act)[I],
                      uint2 (&seg)[I])
{
    uint32_t p[I];
    bool pend[I];
#pragma unroll
    for (uint32_t k = 0; k < I; k++) {
        pend[k] = act[k];
        p[k] = seg_hash(x[k], rel[k]) & g.seg_mask;
        seg[k] = make_uint2(0u, 0u);
    }
    for (;;) {
        bool more = false;
#pragma unroll
        for (uint32_t k = 0; k < I; k++) more |= pend[k];
        if (!more) break;
        uint4 e[I];
#pragma unroll
        for (uint32_t k = 0; k < I; k++) e[k] = pend[k] ? g.seg[p[k]] : make_uint4(kSegEmpty, 0u, 0u, 0u);
#pragma unroll
        for (uint32_t k = 0; k < I; k++) {
            if (!pend[k]) continue;
            if (e[k].x == x[k] && e[k].y == rel[k]) {
                seg[k] = make_uint2(e[k].z, e[k].w);
                pend[k] = false;
            } else if (e[k].x == kSegEmpty) {
                pend[k] = false;  // absent
            } else {
                p[k] = (p[k] + 1) & g.seg_mask;
            }
        }
    }
}

// The same I lookups, all with one relation.
template <uint32_t I>
DEV void seg_lookup_n(const DevGraphView &g, const uint32_t (&x)[I], uint32_t rel, const bool (&act)[I],
                      uint2 (&seg)[I])
{
    uint32_t r[I];
#pragma unroll
    for (uint32_t k = 0; k < I; k++) r[k] = rel;
    seg_lookup_r<I>(g, x, r, act, seg);
}

// keep[k] &= (w[k] is in the sorted slice nbr[seg[k].x, seg[k].y)): the last
// element <= w[k] by halving, then one compare; I searches in lockstep.
template <uint32_t I>
DEV void seg_contains_n(const uint32_t *__restrict__ nbr, const uint2 (&seg)[I], const uint32_t (&w)[I],
                        bool (&keep)[I])
{
    uint32_t lo[I], n[I];
#pragma unroll
    for (uint32_t k = 0; k < I; k++) {
        lo[k] = seg[k].x;
        n[k] = keep[k] ? seg[k].y - seg[k].x : 0u;
    }
    for (;;) {
        bool more = false;
#pragma unroll
        for (uint32_t k = 0; k < I; k++) more |= n[k] > 1;
        if (!more) break;
        uint32_t v[I];
#pragma unroll
        for (uint32_t k = 0; k < I; k++) v[k] = (n[k] > 1) ? nbr[lo[k] + (n[k] >> 1)] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < I; k++) {
            if (n[k] > 1) {
                const uint32_t h = n[k] >> 1;
                if (v[k] <= w[k]) lo[k] += h;
                n[k] -= h;
            }
        }
    }
    uint32_t v[I];
#pragma unroll
    for (uint32_t k = 0; k < I; k++) v[k] = n[k] ? nbr[lo[k]] : 0u;
#pragma unroll
    for (uint32_t k = 0; k < I; k++) keep[k] = (n[k] != 0) && v[k] == w[k];
}

// ---------------------------------------------------------------------------
// chained scan across workgroups (decoupled look-back over packed words)
// ---------------------------------------------------------------------------

DEV uint64_t ld_word(const uint64_t *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
DEV void st_word(uint64_t *p, uint64_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
DEV uint64_t pack(uint64_t st, uint32_t ep, uint64_t v)
{
    return (st << 62) | ((uint64_t)(ep & ((1u << kEpochBits) - 1)) << 46) | (v & kValueMask);
}

// Exclusive prefix of tile t over one word array, where the scan starts at
// tile t0 (the first tile of the step's query), computed by all 64 lanes of
// one wave: each pass reads the words of the 64 preceding tiles, waits until
// every tile up to the nearest inclusive prefix has published, and sums.
DEV uint64_t lookback(const uint64_t *words, uint32_t ep, uint64_t t, uint64_t t0)
{
    const uint32_t lane = lane_id();
    const uint64_t epm = (uint64_t)(ep & ((1u << kEpochBits) - 1));
    uint64_t sum = 0;
    int64_t base = (int64_t)t - 1;
    for (;;) {
        const int64_t idx = base - (int64_t)lane;
        uint64_t w = 0, pm;
        uint32_t st;
        for (;;) {
            st = 2u;  // before the first tile the inclusive prefix is 0
            w = 0;
            if (idx >= (int64_t)t0) {
                w = ld_word(words + idx);
                st = (((w >> 46) & ((1u << kEpochBits) - 1)) == epm) ? (uint32_t)(w >> 62) : 0u;
            }
            pm = __ballot(st == 2u);
            const uint64_t xm = __ballot(st == 0u);
            const uint64_t upto = pm ? (pm ^ (pm - 1)) : ~0ull;
            if (!(xm & upto)) break;
            // The longest back-off (about 8,000 clocks) between polls: each
            // poll is an agent-scope read of up to 64 words that goes past
            // the L2, and a waiting tile holds up no other tile, since its
            // aggregate is already published (measured: polling after
            // s_sleep(1) made the power-law sample 5% slower; a short first
            // retry and then the longest gained nothing over the longest alone).
            __builtin_amdgcn_s_sleep(127);
        }
        const uint64_t upto = pm ? (pm ^ (pm - 1)) : ~0ull;
        const uint64_t v = ((upto >> lane) & 1ull) ? (w & kValueMask) : 0ull;
        sum += wave_sum64(v);
        if (pm) break;
        base -= 64;
    }
    return sum;
}

// Tile tickets: tiles are handed out in launch order, so a tile only waits on
// tiles already owned by running workgroups.  Each workgroup holds one
// prefetched ticket (in LDS) so the atomic's latency overlaps work.  The
// prefetch is issued once the current tile no longer waits on anyone --
// right after its look-back, or at tile start for kinds without one -- so a
// claimed tile is never left idle behind a whole tile's work: a tile's
// predecessors are always running or about to start, and the look-back
// does not queue behind them.
DEV void tickets_init(const ScanWS &ws, uint32_t *s_next)
{
    if (threadIdx.x == 0) *s_next = atomicAdd(ws.ctl + 0, 1u);
}

DEV uint64_t tickets_take(uint32_t *s_next)
{
    __syncthreads();
    return *s_next;  // overwritten only after a later barrier (ticket_prefetch)
}

DEV void ticket_prefetch(const ScanWS &ws, uint32_t *s_next)
{
    if (threadIdx.x == 0) *s_next = atomicAdd(ws.ctl + 0, 1u);
}

template <uint32_t I>
struct ScanSmem {
    uint32_t wk[I][kWaves];
    uint64_t wc[I][kWaves];
    uint64_t ex[2];   // the tile's exclusive prefix (kept, candidates)
    uint64_t tot[2];  // ... and inclusive
};

// Publish this tile's totals (the per-(item, wave) counts already in sm) and
// learn its exclusive prefix: wave 0 chains with the preceding tiles of the
// scan that starts at tile t0.
template <uint32_t I>
DEV void tile_prefix(uint64_t t, uint64_t t0, const ScanWS &ws, uint32_t ep, ScanSmem<I> &sm, uint32_t *s_next)
{
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    if (wave == 0) {
        uint64_t ak = 0, ac = 0;
#pragma unroll
        for (uint32_t k = 0; k < I; k++)
#pragma unroll
            for (uint32_t w = 0; w < kWaves; w++) {
                ak += sm.wk[k][w];
                ac += sm.wc[k][w];
            }
        uint64_t ek = 0, ec = 0;
        if (t != t0) {
            if (lane == 0) {
                st_word(ws.wk + t, pack(1, ep, ak));
                st_word(ws.wc + t, pack(1, ep, ac));
            }
            ek = lookback(ws.wk, ep, t, t0);
            ec = lookback(ws.wc, ep, t, t0);
        }
        if (lane == 0) {
            st_word(ws.wk + t, pack(2, ep, ek + ak));
            st_word(ws.wc + t, pack(2, ep, ec + ac));
            sm.ex[0] = ek;
            sm.ex[1] = ec;
            sm.tot[0] = ek + ak;
            sm.tot[1] = ec + ac;
        }
    }
    __syncthreads();
    ticket_prefetch(ws, s_next);  // this tile waits on nobody any more
}

// Tile-wide exclusive prefix of (keep, cnt) in candidate order (item k of
// thread x is tile position k * 256 + x), continued across the tiles of one
// scan.  Returns each item's kept index and candidate offset and the totals
// through this tile.
template <uint32_t I>
DEV void tile_scan(const bool (&keep)[I], const uint32_t (&cnt)[I], uint64_t t, uint64_t t0, const ScanWS &ws,
                   uint32_t ep, ScanSmem<I> &sm, uint32_t *s_next, uint64_t (&my_k)[I], uint64_t (&my_c)[I],
                   uint64_t &run_k, uint64_t &run_c)
{
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    uint32_t kpre[I], cpre[I];
#pragma unroll
    for (uint32_t k = 0; k < I; k++) {
        const uint64_t m = __ballot(keep[k]);
        kpre[k] = mask_prefix(m);
        const uint32_t cinc = wave_incl_scan(cnt[k]);
        cpre[k] = cinc - cnt[k];
        if (lane == 63) {
            sm.wk[k][wave] = (uint32_t)__popcll(m);
            sm.wc[k][wave] = cinc;
        }
    }
    __syncthreads();
    tile_prefix<I>(t, t0, ws, ep, sm, s_next);
    uint64_t bk = sm.ex[0], bc = sm.ex[1];
#pragma unroll
    for (uint32_t k = 0; k < I; k++) {
        uint64_t ok_ = bk, oc_ = bc;
#pragma unroll
        for (uint32_t w = 0; w < kWaves; w++) {
            if (w < wave) {
                ok_ += sm.wk[k][w];
                oc_ += sm.wc[k][w];
            }
            bk += sm.wk[k][w];
            bc += sm.wc[k][w];
        }
        my_k[k] = ok_ + kpre[k];
        my_c[k] = oc_ + cpre[k];
    }
    run_k = sm.tot[0];
    run_c = sm.tot[1];
}

DEV uint32_t read_epoch(const ScanWS &ws)
{
    return __hip_atomic_load(ws.ctl + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The last workgroup to finish resets the ticket and advances the epoch.
DEV void finish_launch(const ScanWS &ws)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        uint32_t prev = atomicAdd(ws.ctl + 1, 1u);
        if (prev == gridDim.x - 1) {
            __hip_atomic_store(ws.ctl + 0, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(ws.ctl + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(ws.ctl + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Sum (count, digest0, digest1) over the workgroup and add it to `tot`: one
// set of atomics per workgroup, not per wave.
DEV void flush_sums(uint64_t cnt, uint64_t h0, uint64_t h1, Totals *tot_ptr, uint64_t (*s_red)[kWaves])
{
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    cnt = wave_sum64(cnt);
    h0 = wave_sum64(h0);
    h1 = wave_sum64(h1);
    if (lane == 0) {
        s_red[0][wave] = cnt;
        s_red[1][wave] = h0;
        s_red[2][wave] = h1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t c = 0, a = 0, b = 0;
        for (uint32_t w = 0; w < kWaves; w++) {
            c += s_red[0][w];
            a += s_red[1][w];
            b += s_red[2][w];
        }
        const auto tot = gp(tot_ptr);
        if (c) __hip_atomic_fetch_add(&tot->count, c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (a) __hip_atomic_fetch_add(&tot->dg0, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (b) __hip_atomic_fetch_add(&tot->dg1, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// One per-node lookup of a tile's staging (a checked ancestor's segment or
// the next generator's): the row column whose vertex is looked up (~0u: none,
// the check is the candidate's own), its relation and its dense table.
struct NodeLookup {
    uint32_t col, rel;
    const void *tab;
};

// Statically allocated LDS of one tile loop.
template <uint32_t I>
struct TileShared {
    ScanSmem<I> sm;
    uint64_t tile;
    uint64_t node[2];
    uint64_t red[3][kWaves];
    uint32_t slot;
    uint32_t next;  // prefetched tile ticket
    NodeLookup lk[GFSM_MAXCHK + 1];  // the depth's lookups, read per lane from LDS
};

// ---------------------------------------------------------------------------
// root level (hnInit, HnGPU.cu:76-109)
// ---------------------------------------------------------------------------

// A step's {keep, cnt} in the host's words: cnt first, then keep behind a
// system-scope fence, so a host that sees keep sees cnt too (it polls keep).
DEV void publish_kc(uint64_t *kc, uint64_t keep, uint64_t cnt)
{
    __hip_atomic_store(kc + 1, cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(kc, keep, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Root tile `lt` of a root step (global tile gt, the step's scan starting at
// gt0): keep candidates that carry every self-loop relation of the root.
template <uint32_t I, class CP>
DEV void root_tile(const DevGraphView &g, const RootIO &io, uint32_t nchk, CP chk_rel, uint64_t lt,
                   uint64_t gt, uint64_t gt0, const ScanWS &ws, uint32_t ep, ScanSmem<I> &sm, uint32_t *s_next)
{
    bool keep[I];
    uint32_t v[I], cnt[I];
#pragma unroll
    for (uint32_t k = 0; k < I; k++) {
        const uint64_t c = lt * (kTile * I) + k * kTile + threadIdx.x;
        keep[k] = false;
        v[k] = 0;
        cnt[k] = 0;
        if (c < io.n) {
            v[k] = gp(io.cand_vtx)[c];
            cnt[k] = gp(io.cand_cnt)[c];
            keep[k] = cnt[k] > 0;
        }
    }
    for (uint32_t r = 0; r < nchk; r++) {
        uint2 sg[I];
        seg_lookup_n<I>(g, v, chk_rel[r], keep, sg);
        seg_contains_n<I>(g.nbr, sg, v, keep);
    }
#pragma unroll
    for (uint32_t k = 0; k < I; k++)
        if (!keep[k]) cnt[k] = 0;
    uint64_t kk[I], off[I], rk, rc;
    tile_scan<I>(keep, cnt, gt, gt0, ws, ep, sm, s_next, kk, off, rk, rc);
#pragma unroll
    for (uint32_t k = 0; k < I; k++) {
        if (keep[k] && kk[k] < io.cap) {
            const uint64_t c = lt * (kTile * I) + k * kTile + threadIdx.x;
            gp(io.out_vec)[kk[k]] = v[k];
            gp(io.out_off)[kk[k]] = off[k];
            gp(io.out_gbeg)[kk[k]] = gp(io.cand_beg)[c];
        }
    }
    if (lt == io.ntiles - 1 && threadIdx.x == 0) {
        gp(io.tot)->keep = rk;
        gp(io.tot)->cnt = rc;
        if (rk <= io.cap) gp(io.out_off)[rk] = rc;
        if (io.host_kc) publish_kc(io.host_kc, rk, rc);
    }
}

template <uint32_t I>
__global__ __launch_bounds__(kTile) void k_root(RootArgs a)
{
    __shared__ TileShared<I> sh;
    const uint32_t ep = read_epoch(a.ws);
    tickets_init(a.ws, &sh.next);
    for (;;) {
        const uint64_t t = tickets_take(&sh.next);
        if (t >= a.io.ntiles) break;
        root_tile<I>(a.g, a.io, a.nchk, a.chk_rel, t, t, 0, a.ws, ep, sh.sm, &sh.next);
    }
    finish_launch(a.ws);
}

// One-vertex queries: a vertex qualifies when its directory is non-empty
// (it owns a segment, so it has an out-entry) and it lies in its own segment
// of every self-loop relation.  The row hash of a one-column row {x} is
// mix(mix(seed ^ x)) per seed.
template <bool ROWS>
__global__ __launch_bounds__(kTile) void k_single(SingleArgs a)
{
    constexpr uint32_t I = kRootItems;
    __shared__ TileShared<I> sh;
    const uint32_t ep = read_epoch(a.ws);
    uint64_t acc = 0, h0 = 0, h1 = 0;
    if (ROWS) tickets_init(a.ws, &sh.next);
    for (uint64_t t = ROWS ? 0 : blockIdx.x;; t += ROWS ? 0 : gridDim.x) {
        if (ROWS) t = tickets_take(&sh.next);
        if (t >= a.ntiles) break;
        bool keep[I];
        uint32_t x[I], zero[I];
#pragma unroll
        for (uint32_t k = 0; k < I; k++) {
            const uint64_t v = (uint64_t)a.lo + t * (kTile * I) + k * kTile + threadIdx.x;
            x[k] = (uint32_t)v;
            zero[k] = 0;
            keep[k] = v < a.hi && a.g.dir_ptr[v + 1] > a.g.dir_ptr[v];
        }
        for (uint32_t r = 0; r < a.nchk; r++) {
            uint2 sg[I];
            seg_lookup_n<I>(a.g, x, a.chk_rel[r], keep, sg);
            seg_contains_n<I>(a.g.nbr, sg, x, keep);
        }
        if (ROWS) {
            uint64_t kk[I], off[I], rk, rc;
            tile_scan<I>(keep, zero, t, 0, a.ws, ep, sh.sm, &sh.next, kk, off, rk, rc);
#pragma unroll
            for (uint32_t k = 0; k < I; k++)
                if (keep[k] && kk[k] < a.cap) gp(a.rows)[kk[k]] = x[k];
            if (t == a.ntiles - 1 && threadIdx.x == 0) gp(a.tot)->keep = rk;
        } else {
#pragma unroll
            for (uint32_t k = 0; k < I; k++) {
                if (!keep[k]) continue;
                acc++;
                h0 += mix64(mix64(0x9E3779B97F4A7C15ull ^ (uint64_t)x[k]));
                h1 += mix64(mix64(0xD1B54A32D192ED03ull ^ (uint64_t)x[k]));
            }
        }
    }
    if (!ROWS) flush_sums(acc, h0, h1, a.tot, sh.red);
    finish_launch(a.ws);
}

// ---------------------------------------------------------------------------
// one depth: expand + filter + compact + next level
// ---------------------------------------------------------------------------

// Largest i in [0, S) with off[i] <= key, found by half a wave (32 lanes,
// `sub` = lane within the half) in 32-ary steps; off[0] = 0 <= key.
template <class P>
DEV uint64_t frontier_search(P off, uint64_t S, uint64_t key, uint32_t sub)
{
    uint64_t lo = 0, hi = S;
    while (hi - lo > 1) {
        const uint64_t n = hi - lo;
        const uint64_t stride = (n + 31) / 32;
        const uint64_t pos = lo + (uint64_t)sub * stride;
        const bool ok = pos < hi && off[pos] <= key;
        const uint64_t m = __ballot(ok);
        const uint32_t half = (uint32_t)((lane_id() < 32) ? (m & 0xFFFFFFFFull) : (m >> 32));
        const uint32_t top = 31u - __builtin_clz(half | 1u);  // lane 0 of the half always qualifies
        lo = lo + (uint64_t)top * stride;
        hi = (lo + stride < hi) ? lo + stride : hi;
    }
    return lo;
}

// Column c of the node row at `row` (c < d), or `own` (the candidate itself)
// when c == d.  The row word is read for every lane, at column 0 where it is
// not wanted, and the value is selected afterwards: a select between the two
// sources' addresses would put the candidates' register array in scratch
// memory and make every such read a flat load.
template <class RP>
DEV uint32_t pick_row_or(RP rows, uint32_t row, uint32_t c, uint32_t d, uint32_t own)
{
    const uint32_t v = rows[row + (c < d ? c : 0u)];
    return c < d ? v : own;
}

// Expand tile `lt` of a step (global tile gt; the step's scan starts at gt0).
// `dyn` is the dynamic LDS: the node-start bitmap, the per-position generator
// base and a pool of pool_words words that stages, per frontier node of the
// tile, the checked ancestor segments, the next generator segment and the
// placed vertices, as far as they fit.  LAST_COUNT adds its embeddings and
// digest to the thread's accumulators; the other kinds write their output.
template <uint32_t KIND, uint32_t I, class DP, class CP>
DEV void expand_tile(const DevGraphView &g, const DP &dp, CP col, CP dcol,
                     const StepIO &io, uint64_t lt, uint64_t gt, uint64_t gt0, const ScanWS &ws, uint32_t ep,
                     uint32_t pool_words, TileShared<I> &sh, unsigned char *dyn, uint64_t &acc_cnt,
                     uint64_t &acc_h0, uint64_t &acc_h1, bool tickets = true)
{
    constexpr uint32_t T = kTile, TC = kTile * I, NW = TC / 64;
    constexpr bool BARE = KIND == EXPAND_LAST_BARE;  // counting, no checks, no digest
    constexpr bool COUNTING = KIND == EXPAND_LAST_COUNT || BARE;
    const uint32_t d = io.d;
    // The thread index passes through an empty asm, so the compiler sees a
    // new value per tile: the LDS offsets and addresses derived from it are
    // recomputed in each tile (a few VALU ops) instead of being hoisted out
    // of the tile loop and spilled (measured: the inner kernel's scratch
    // fell from 112 to 32 bytes per lane, the lean kernel's from 36 to 0).
    uint32_t tid_ = threadIdx.x;
    __asm__ volatile("" : "+v"(tid_));
    const uint32_t tid = tid_;
    const uint32_t lane = tid & 63u, wave = tid >> 6;  // wave64: lanes are consecutive threads
    uint64_t *s_bits = reinterpret_cast<uint64_t *>(dyn);            // [NW] node start positions
    uint32_t *s_seg = reinterpret_cast<uint32_t *>(s_bits + NW);     // [TC] nbr index of tile position 0
    uint32_t *s_pool = s_seg + TC;                                   // [pool_words]

    const uint64_t c0 = io.g_begin + lt * TC;
    const uint64_t c1 = (c0 + TC < io.g_end) ? c0 + TC : io.g_end;
    const uint32_t n = (uint32_t)(c1 - c0);
    const auto in_off = gp(io.in_off);
    const auto in_gbeg = gp(io.in_gbeg);
    const auto out_vec = gp(io.out_vec);
    const auto out_off = gp(io.out_off);
    const auto out_gbeg = gp(io.out_gbeg);
    const auto out_rows = gp(io.rows);
    const auto tot = gp(io.tot);

    const uint32_t nchk = BARE ? 0u : dp.nchk;
    const uint32_t need_nxt = (KIND == EXPAND_MID) ? io.need_nxt : 0u;
    // The depth's lookups, one LDS entry each, so the staging loop below
    // selects its lookup per lane from LDS rather than indexing the plan
    // (kernel arguments or global memory) with a per-lane index.
    if (!BARE && tid < nchk + need_nxt) {
        NodeLookup e;
        if (tid < nchk) {
            const uint32_t dist = dp.chk_dist[tid];
            e.col = dist ? d - dist : ~0u;  // self checks: per candidate
            e.rel = dp.chk_rel[tid];
            e.tab = tid < kDenseChk ? dp.chk_seg[tid] : nullptr;
        } else {
            e.col = d + 1 - dp.nxt_dist;
            e.rel = dp.nxt_rel;
            e.tab = dp.nxt_seg;
        }
        sh.lk[tid] = e;
    }
    // Frontier nodes covering [c0, c1): offsets are strictly increasing.
    if (tid < NW) s_bits[tid] = 0;
    if (io.tile_node) {
        if (tid < 2) sh.node[tid] = gp(io.tile_node)[2 * lt + tid];
    } else if (tid < 64) {
        const uint32_t sub = tid & 31;
        const uint64_t key = (tid < 32) ? c0 : c1 - 1;
        const uint64_t i = frontier_search(in_off, io.S_prev, key, sub);
        if (sub == 0) sh.node[tid >> 5] = i;
    }
    __syncthreads();
    if (KIND == EXPAND_LAST_COUNT && tickets) ticket_prefetch(ws, &sh.next);  // no look-back in this kind
    const uint64_t ilo = sh.node[0];
    const uint32_t nf = (uint32_t)(sh.node[1] - ilo + 1);
    // The rows come first: every candidate reads its node's row (d words,
    // twice on inner depths), but a node's segments save lookups only for
    // its candidates after the first.
    const uint32_t meta_words = 2 * nf * nchk + (need_nxt ? 2 * nf : 0);
    const uint32_t row_words = (nf * d + 1) & ~1u;  // keeps the segments 8-byte aligned
    const bool staged = row_words <= pool_words;
    const bool stage_meta = meta_words + (staged ? row_words : 0) <= pool_words;
    uint32_t *s_rows = s_pool;                                                   // [nf][d]
    uint2 *s_chk = reinterpret_cast<uint2 *>(s_pool + (staged ? row_words : 0));  // [nchk][nf]
    uint2 *s_nxt = s_chk + nchk * nf;                                            // [nf]
    // Digest of a counting last depth: each node's placed columns contribute
    // the same terms to all its candidates' row hashes, so their sums are
    // formed once per node (two u64 per node, after the staged data).
    const uint32_t pool_used = (staged ? row_words : 0u) + (stage_meta ? meta_words : 0u);
    const bool part_ok = KIND == EXPAND_LAST_COUNT && io.digest && pool_used + 4 * nf <= pool_words;
    uint64_t *s_part = reinterpret_cast<uint64_t *>(s_pool + pool_used);         // [nf][2]
    const auto in_rows = gp(io.in_vec) + ilo * d;       // the tile's node rows, contiguous

    // The rows' words, kStageLd loads in flight per thread before their LDS
    // stores (one load, wait, store per word would chain round trips).
    if (staged) {
        const uint32_t nw = nf * d;
        for (uint32_t u0 = tid; u0 < nw; u0 += T * kStageLd) {
            uint32_t v[kStageLd];
#pragma unroll
            for (uint32_t b = 0; b < kStageLd; b++) {
                const uint32_t u = u0 + b * T;
                v[b] = u < nw ? in_rows[u] : 0u;
            }
#pragma unroll
            for (uint32_t b = 0; b < kStageLd; b++) {
                const uint32_t u = u0 + b * T;
                if (u < nw) s_rows[u] = v[b];
            }
        }
    }
    // Per frontier node: start bit, generator base (both words loaded before
    // either is used).
    for (uint32_t j = tid; j < nf; j += T) {
        const uint64_t i = ilo + j;
        const uint64_t o = in_off[i];
        const uint32_t gbv = in_gbeg[i];
        const uint32_t first = (o > c0) ? (uint32_t)(o - c0) : 0u;
        atomicOr((unsigned long long *)&s_bits[first >> 6], 1ull << (first & 63));
        s_seg[j] = gbv + (uint32_t)(c0 - o);
        if (part_ok) {
            uint64_t p0 = 0, p1 = 0;
            for (uint32_t c = 0; c < io.q; c++) {
                const uint32_t dc = dcol[c];
                if (dc == d) continue;
                const uint64_t xv = (uint64_t)in_rows[(uint64_t)j * d + dc] | ((uint64_t)c << 32);
                p0 += mix64(0x9E3779B97F4A7C15ull ^ xv);
                p1 += mix64(0xD1B54A32D192ED03ull ^ xv);
            }
            s_part[2 * j] = p0;
            s_part[2 * j + 1] = p1;
        }
    }
    // Staged segments: every (node, lookup) pair of the tile -- the node's
    // checked ancestors and its next generator -- striped over the threads
    // and resolved kNodeLk at a time in lockstep, so a node's lookups do not
    // run one after another.
    if (!BARE && stage_meta) {  // the lean kind has no lookups to stage
        constexpr uint32_t L = kNodeLk;
        const uint32_t nl = nchk + need_nxt;
        const uint32_t total = nf * nl;
        for (uint32_t u0 = tid; u0 < total; u0 += T * L) {
            uint32_t x[L], rel[L], dst[L];
            bool act[L];
            const void *tab[L];
#pragma unroll
            for (uint32_t k = 0; k < L; k++) {
                const uint32_t u = u0 + k * T;
                act[k] = false;
                x[k] = rel[k] = dst[k] = 0;
                tab[k] = nullptr;
                if (u < total) {
                    const uint32_t j = u / nl, r = u - j * nl;
                    // row column of the vertex whose segment is looked up
                    // (a check's ancestor at distance chk_dist, or the next
                    // generator's at distance nxt_dist - 1 from this depth)
                    const NodeLookup e = sh.lk[r];
                    act[k] = e.col != ~0u;
                    rel[k] = e.rel;
                    dst[k] = r * nf + j;  // s_chk[r][j], or s_nxt[j] for r == nchk
                    tab[k] = e.tab;
                    if (act[k]) x[k] = in_rows[(uint64_t)j * d + e.col];
                }
            }
            // dense tables where the relation has one, hash lookups for the rest
            bool hact[L];
            uint64_t dv[L];
#pragma unroll
            for (uint32_t k = 0; k < L; k++) {
                hact[k] = act[k] && !tab[k];
                dv[k] = (act[k] && tab[k]) ? gp(reinterpret_cast<const uint64_t *>(tab[k]))[x[k]] : 0ull;
            }
            uint2 sg[L];
            seg_lookup_r<L>(g, x, rel, hact, sg);
#pragma unroll
            for (uint32_t k = 0; k < L; k++)
                if (act[k]) s_chk[dst[k]] = tab[k] ? make_uint2((uint32_t)dv[k], (uint32_t)(dv[k] >> 32)) : sg[k];
        }
    }
    __syncthreads();

    // Bitmap words of this wave (word k * 4 + wave covers its 64 positions
    // of item k) and the set bits before each of them.
    uint32_t wpre[I];
    uint64_t wbits[I];
    if constexpr (I > 4 && !BARE) {
        // Every wave sums all words: at 8 candidates per thread the scalar
        // form below costs the general kernel more in register allocation than
        // it saves (measured on the road-like grid: 52.7-54.5 against 51.7 ms).
        uint32_t acc = 0;
#pragma unroll
        for (uint32_t k = 0; k < I; k++)
#pragma unroll
            for (uint32_t w = 0; w < kWaves; w++) {
                const uint64_t b = s_bits[k * kWaves + w];
                if (w == wave) {
                    wpre[k] = acc;
                    wbits[k] = b;
                }
                acc += (uint32_t)__popcll(b);
            }
    } else {
        // Lane l (< NW) holds bitmap word l; one wave scan of their popcounts
        // gives every word's prefix, and this wave's word of each item (word
        // k * kWaves + wave) is read back from its lane into scalar registers.
        static_assert(NW <= 64, "one bitmap word per lane");
        const uint64_t b = lane < NW ? s_bits[lane] : 0ull;
        const uint32_t pc = (uint32_t)__popcll(b);
        const uint32_t exc = wave_incl_scan(pc) - pc;
        const uint32_t wv = __builtin_amdgcn_readfirstlane(wave);
#pragma unroll
        for (uint32_t k = 0; k < I; k++) {
            const uint32_t at = k * kWaves + wv;
            // readlane returns int: through uint32_t, so the low word is not sign-extended
            wpre[k] = (uint32_t)__builtin_amdgcn_readlane(exc, at);
            wbits[k] = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((uint32_t)(b >> 32), at) << 32) |
                       (uint64_t)(uint32_t)__builtin_amdgcn_readlane((uint32_t)b, at);
        }
    }

    // Phase 1: node, generator position, candidate read (all items).
    uint32_t jj[I], w[I], cnt[I], gb[I];
    bool keep[I];
#pragma unroll
    for (uint32_t k = 0; k < I; k++) {
        const uint32_t x = k * T + tid;
        const bool valid = x < n;
        const uint32_t j = wpre[k] + mask_prefix(wbits[k]) + (uint32_t)((wbits[k] >> lane) & 1ull) - 1u;
        jj[k] = valid ? j : 0u;
        const uint32_t pos = s_seg[jj[k]] + x;
        keep[k] = valid && pos < g.nent;  // offsets and segments always agree; never read past the array
        w[k] = keep[k] ? g.nbr[pos] : 0u;
        cnt[k] = 0;
        gb[k] = 0;
    }
    // Phases 2-4 and the write-out read the placed vertices either from the
    // LDS copy or, when the tile's nodes did not fit, from the level itself.
    // Two inlined copies keep every access in its own address space: a
    // pointer that could be either would turn the LDS reads into flat loads,
    // which take the vector-memory path.
    auto tail = [&](auto rows) {
        // Phase 2: injectivity against the node's placed vertices, one placed
        // column at a time for all items, so I reads are in flight together.
        for (uint32_t kk = 0; kk < d; kk++) {
            uint32_t pv[I];
#pragma unroll
            for (uint32_t k = 0; k < I; k++) pv[k] = rows[jj[k] * d + kk];
#pragma unroll
            for (uint32_t k = 0; k < I; k++) keep[k] &= pv[k] != w[k];
        }
        // Phase 3: membership in every checked segment, I searches in lockstep.
        for (uint32_t r = 0; r < nchk; r++) {
            const uint32_t dist = dp.chk_dist[r];
            uint2 sg[I];
            if (dist && stage_meta) {
#pragma unroll
                for (uint32_t k = 0; k < I; k++) sg[k] = s_chk[r * nf + jj[k]];
            } else {
                uint32_t x[I];
#pragma unroll
                for (uint32_t k = 0; k < I; k++) x[k] = pick_row_or(rows, jj[k] * d, dist ? d - dist : d, d, w[k]);
                const void *ct = r < kDenseChk ? dp.chk_seg[r] : nullptr;
                if (ct) {
                    const auto tab = gp(reinterpret_cast<const uint64_t *>(ct));
#pragma unroll
                    for (uint32_t k = 0; k < I; k++) {
                        const uint64_t v = keep[k] ? tab[x[k]] : 0ull;
                        sg[k] = make_uint2((uint32_t)v, (uint32_t)(v >> 32));
                    }
                } else {
                    seg_lookup_n<I>(g, x, dp.chk_rel[r], keep, sg);
                }
            }
            seg_contains_n<I>(g.nbr, sg, w, keep);
        }

        if (KIND == EXPAND_MID) {
            // Phase 4: the survivor's next generator segment; no children, no node.
            uint2 ns[I];
            if (need_nxt && stage_meta) {
#pragma unroll
                for (uint32_t k = 0; k < I; k++) ns[k] = s_nxt[jj[k]];
            } else if (dp.nxt_seg) {
                // one read of the relation's dense table ({begin, end} words)
                const auto tab = gp(reinterpret_cast<const uint64_t *>(dp.nxt_seg));
#pragma unroll
                for (uint32_t k = 0; k < I; k++) {
                    const uint32_t x = pick_row_or(rows, jj[k] * d, need_nxt ? d + 1 - dp.nxt_dist : d, d, w[k]);
                    const uint64_t v = keep[k] ? tab[x] : 0ull;
                    ns[k] = make_uint2((uint32_t)v, (uint32_t)(v >> 32));
                }
            } else {
                uint32_t x[I];
#pragma unroll
                for (uint32_t k = 0; k < I; k++)
                    x[k] = pick_row_or(rows, jj[k] * d, need_nxt ? d + 1 - dp.nxt_dist : d, d, w[k]);
                seg_lookup_n<I>(g, x, dp.nxt_rel, keep, ns);
            }
#pragma unroll
            for (uint32_t k = 0; k < I; k++) {
                if (!keep[k]) continue;
                cnt[k] = ns[k].y - ns[k].x;
                gb[k] = ns[k].x;
                keep[k] = cnt[k] > 0;
            }
        }

        if (!COUNTING) {
            // Compaction: wave-level ballot/mbcnt and DPP scans, then the
            // tile's prefix from the look-back, then each item's slot.
            ScanSmem<I> &sm = sh.sm;
            uint32_t kpre[I], cpre[I];
#pragma unroll
            for (uint32_t k = 0; k < I; k++) {
                const uint64_t m = __ballot(keep[k]);
                kpre[k] = mask_prefix(m);
                const uint32_t cinc = wave_incl_scan(cnt[k]);
                cpre[k] = cinc - cnt[k];
                if (lane == 63) {
                    sm.wk[k][wave] = (uint32_t)__popcll(m);
                    sm.wc[k][wave] = cinc;
                }
            }
            __syncthreads();
            tile_prefix<I>(gt, gt0, ws, ep, sm, &sh.next);
            uint64_t bk = sm.ex[0], bc = sm.ex[1];
            uint64_t idx[I];
            bool put[I];
            // MID rows of one wave and item are consecutive in the next
            // level: staged in this wave's share of s_seg (free after phase
            // 1) and written out as whole words by all lanes, instead of
            // each lane storing its row with a stride of d + 1 words.
            const uint32_t row_w = d + 1;
            const bool coalesce = KIND == EXPAND_MID && row_w * 64 <= TC / kWaves;
            uint32_t *wst = s_seg + wave * (TC / kWaves);
#pragma unroll
            for (uint32_t k = 0; k < I; k++) {
                uint64_t ok_ = bk, oc_ = bc;
#pragma unroll
                for (uint32_t ww = 0; ww < kWaves; ww++) {
                    if (ww < wave) {
                        ok_ += sm.wk[k][ww];
                        oc_ += sm.wc[k][ww];
                    }
                    bk += sm.wk[k][ww];
                    bc += sm.wc[k][ww];
                }
                idx[k] = ok_ + kpre[k];
                put[k] = keep[k] && idx[k] < io.cap;
                if (KIND == EXPAND_MID && put[k]) {
                    out_off[idx[k]] = oc_ + cpre[k];
                    out_gbeg[idx[k]] = gb[k];
                } else if (KIND == EXPAND_LAST_ROWS && put[k]) {
                    out_rows[idx[k] * io.q + col[d]] = w[k];
                }
                if (KIND == EXPAND_MID && coalesce) {
                    if (put[k]) {
                        uint32_t *sr = wst + kpre[k] * row_w;
                        for (uint32_t kk = 0; kk < d; kk++) sr[kk] = rows[jj[k] * d + kk];
                        sr[d] = w[k];
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    // rows [ok_, ok_ + nrow) of the level, capped at its capacity
                    uint64_t nrow = (uint64_t)__popcll(__ballot(keep[k]));
                    nrow = ok_ >= io.cap ? 0 : (nrow < io.cap - ok_ ? nrow : io.cap - ok_);
                    const uint32_t words = (uint32_t)nrow * row_w;  // at most 64 * I
                    const auto dst = out_vec + ok_ * row_w;
                    // 16-byte stores for the body (a quarter of the store
                    // instructions; the level is 256-byte aligned), single
                    // words for the head before the first 16-byte boundary
                    // and for the tail.
                    const uint32_t head = (uint32_t)((4u - ((ok_ * row_w) & 3u)) & 3u);
                    const uint32_t h = head < words ? head : words;
                    const uint32_t nq = (words - h) >> 2;
                    const uint32_t t0 = h + 4 * nq;
                    const uint32_t hv = lane < h ? wst[lane] : 0u;
                    const uint32_t tv = lane < words - t0 ? wst[t0 + lane] : 0u;
                    u32x4 bv[(I + 3) / 4];
#pragma unroll
                    for (uint32_t i = 0; i < (I + 3) / 4; i++) {
                        const uint32_t c = lane + 64 * i;  // nq <= 16 * I groups
                        const uint32_t *sw = wst + h + 4 * c;
                        bv[i] = c < nq ? u32x4{sw[0], sw[1], sw[2], sw[3]} : u32x4{0u, 0u, 0u, 0u};
                    }
                    if (lane < h) dst[lane] = hv;
                    if (lane < words - t0) dst[t0 + lane] = tv;
                    const auto dst4 = (gptr_t<u32x4>)(dst + h);
#pragma unroll
                    for (uint32_t i = 0; i < (I + 3) / 4; i++)
                        if (lane + 64 * i < nq) dst4[lane + 64 * i] = bv[i];
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                } else if (KIND == EXPAND_MID && put[k]) {
                    out_vec[idx[k] * row_w + d] = w[k];
                }
            }
            // Otherwise the placed vertices, one column at a time for all items.
            if (!(KIND == EXPAND_MID && coalesce)) {
                for (uint32_t kk = 0; kk < d; kk++) {
                    const uint32_t ck = (KIND == EXPAND_MID) ? kk : col[kk];
                    const uint64_t wd = (KIND == EXPAND_MID) ? d + 1 : io.q;
                    const auto out = (KIND == EXPAND_MID) ? out_vec : out_rows;
#pragma unroll
                    for (uint32_t k = 0; k < I; k++)
                        if (put[k]) out[idx[k] * wd + ck] = rows[jj[k] * d + kk];
                }
            }
            if (lt == io.ntiles - 1 && tid == 0) {
                tot->keep = sm.tot[0];
                tot->cnt = sm.tot[1];
                if (io.host_kc) publish_kc(io.host_kc, sm.tot[0], sm.tot[1]);
                if (KIND == EXPAND_MID && sm.tot[0] <= io.cap) out_off[sm.tot[0]] = sm.tot[1];
            }
        } else {
#pragma unroll
            for (uint32_t k = 0; k < I; k++) acc_cnt += keep[k] ? 1u : 0u;
        }
        if (KIND != EXPAND_MID && !BARE && io.digest && part_ok) {
            const uint64_t cw = (uint64_t)col[d] << 32;
#pragma unroll
            for (uint32_t k = 0; k < I; k++) {
                if (!keep[k]) continue;
                const uint64_t xv = (uint64_t)w[k] | cw;
                acc_h0 += mix64(s_part[2 * jj[k]] + mix64(0x9E3779B97F4A7C15ull ^ xv));
                acc_h1 += mix64(s_part[2 * jj[k] + 1] + mix64(0xD1B54A32D192ED03ull ^ xv));
            }
        } else if (KIND != EXPAND_MID && !BARE && io.digest) {
#pragma unroll
            for (uint32_t k = 0; k < I; k++) {
                if (!keep[k]) continue;
                uint64_t x0 = 0, x1 = 0;
                for (uint32_t c = 0; c < io.q; c++) {
                    const uint32_t dc = dcol[c];
                    const uint32_t val = pick_row_or(rows, jj[k] * d, dc, d, w[k]);
                    const uint64_t xv = (uint64_t)val | ((uint64_t)c << 32);
                    x0 += mix64(0x9E3779B97F4A7C15ull ^ xv);
                    x1 += mix64(0xD1B54A32D192ED03ull ^ xv);
                }
                acc_h0 += mix64(x0);
                acc_h1 += mix64(x1);
            }
        }
    };
    if (staged)
        tail(s_rows);
    else
        tail(in_rows);
}


// Tiles of the kinds without a look-back, dealt so that runs of kXcdRun
// consecutive tiles go to workgroups of one XCD: consecutive tiles expand
// neighbouring nodes, often from the same generator segment, so the lines
// one tile fetches into its XCD's L2 serve the next.  Workgroups are placed
// on the 8 XCDs round-robin by workgroup index; run r of the launch belongs
// to XCD r mod 8, and the workgroups of that XCD take its tiles in order.
// Every tile is taken once for any grid; grids that are not a multiple of 8
// deal plain round-robin.
constexpr uint32_t kXcds = 8;
constexpr uint64_t kXcdRun = 16;

struct TileDeal {
    uint64_t s, step, x;
    bool xcd;
    DEV TileDeal()
    {
        xcd = (gridDim.x % kXcds) == 0;
        x = blockIdx.x % kXcds;
        s = xcd ? blockIdx.x / kXcds : blockIdx.x;
        step = xcd ? gridDim.x / kXcds : gridDim.x;
    }
    // tile of this workgroup's s-th share (increasing in s)
    DEV uint64_t tile() const { return xcd ? ((s / kXcdRun) * kXcds + x) * kXcdRun + (s % kXcdRun) : s; }
    DEV void next() { s += step; }
};

// Four waves per SIMD (at most 128 VGPRs): the kernel is bound by the latency
// of dependent loads, and the extra resident workgroup hides more of it than
// the few registers spilled at 8 candidates per thread cost (measured).
template <uint32_t KIND, uint32_t I>
__global__ __launch_bounds__(kTile) __attribute__((amdgpu_waves_per_eu(4))) void k_expand(ExpandArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ TileShared<I> sh;
    const uint32_t ep = read_epoch(a.ws);
    uint64_t acc_cnt = 0, acc_h0 = 0, acc_h1 = 0;
    if (KIND == EXPAND_LAST_COUNT) {
        // No tile waits on another, so tiles are dealt without tickets: a shared
        // ticket counter that every tile passes through caps this kind's
        // rate (measured: 1.7x faster without it).
        for (TileDeal dl; dl.tile() < a.io.ntiles; dl.next()) {
            const uint64_t t = dl.tile();
            __syncthreads();  // the previous tile's readers of the tile LDS are done
            expand_tile<KIND, I>(a.g, a.dp, a.col, a.dcol, a.io, t, t, 0, a.ws, ep, a.pool_words, sh, smem, acc_cnt,
                                 acc_h0, acc_h1, false);
        }
    } else {
        tickets_init(a.ws, &sh.next);
        for (;;) {
            const uint64_t t = tickets_take(&sh.next);
            if (t >= a.io.ntiles) break;
            expand_tile<KIND, I>(a.g, a.dp, a.col, a.dcol, a.io, t, t, 0, a.ws, ep, a.pool_words, sh, smem, acc_cnt,
                                 acc_h0, acc_h1);
        }
    }
    if (KIND != EXPAND_MID) flush_sums(acc_cnt, acc_h0, acc_h1, a.io.tot, sh.red);
    finish_launch(a.ws);
}

// The lean last-depth kernel: counting only, no checks, no digest, a small
// staging pool.  Its tile body needs few registers, so eight waves per SIMD
// (twice the others) hide the latency of each tile's dependent loads.
template <uint32_t I>
__global__ __launch_bounds__(kTile) __attribute__((amdgpu_waves_per_eu(8))) void k_expand_bare(ExpandArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ TileShared<I> sh;
    const uint32_t ep = read_epoch(a.ws);
    uint64_t acc_cnt = 0, acc_h0 = 0, acc_h1 = 0;
    for (TileDeal dl; dl.tile() < a.io.ntiles; dl.next()) {  // dealt by XCD, as for LAST_COUNT
        const uint64_t t = dl.tile();
        __syncthreads();
        expand_tile<EXPAND_LAST_BARE, I>(a.g, a.dp, a.col, a.dcol, a.io, t, t, 0, a.ws, ep, a.pool_words, sh, smem,
                                         acc_cnt, acc_h0, acc_h1, false);
    }
    flush_sums(acc_cnt, 0, 0, a.io.tot, sh.red);
    finish_launch(a.ws);
}

// ---------------------------------------------------------------------------
// batched matching: device scheduler + persistent step kernel
// ---------------------------------------------------------------------------

DEV uint64_t umin64(uint64_t a, uint64_t b) { return a < b ? a : b; }
DEV uint64_t umax64(uint64_t a, uint64_t b) { return a > b ? a : b; }

// Under a per-depth cap, the next chunk of depth d should generate about as
// many depth-(d+1) candidates as that depth has budget left once this
// chunk's `cnt` children are tested: estimated from this chunk's fan-out
// (cnt children from `n` candidates).  Fewer candidates wasted beyond the
// cap; the result does not depend on chunk sizes.
DEV uint64_t capped_chunk(const BatchSlot &B, uint32_t d, uint64_t n, uint64_t cnt, uint64_t chunk, uint64_t TC,
                          uint64_t limit)
{
    if (d + 1 >= B.q) return chunk;
    const uint64_t used = B.cand[d + 1];
    const uint64_t rem = B.max_depth_cand > used ? B.max_depth_cand - used : 0;
    const uint64_t after = rem > cnt ? rem - cnt : 0;
    const uint64_t need = after / umax64(cnt, 1) * n + after % umax64(cnt, 1) * n / umax64(cnt, 1) + 1;
    return umax64(TC, umin64(umin64(chunk, need), limit));
}

// Consume the finished step of slot B (depth-first chunk walk), then
// schedule its next step.  B may be a copy of the slot; `tot` is the slot's
// own totals word block, which the step writes.  Returns the new step's tile count (0 when the slot
// has no step).  Mirrors the single-query driver's recursion: a level that
// overflows its buffer is retried on half the range; otherwise the chunk is
// re-estimated from the observed selectivity.
DEV uint32_t schedule_slot(const BatchCfg &cfg, BatchSlot &B, Totals *tot)
{
    const uint64_t TC = (uint64_t)kTile * cfg.items;
    const uint64_t by_val = kValueMask / umax64(cfg.max_degree, 1);
    const uint64_t step = cfg.ctl->step_cand ? umin64(cfg.ctl->step_cand, cfg.max_step_cand) : cfg.max_step_cand;
    const uint64_t limit = umax64(TC, umin64(step, by_val));
    const uint32_t q = B.q;
    uint32_t state = B.state;
    if (state == SLOT_NEW) {
        B.root_next = 0;
        for (uint32_t k = 0; k < GFSM_MAXQ; k++) {
            B.lo[k] = B.hi[k] = B.S[k] = 0;
            B.level[k] = B.cand[k] = 0;
            B.chunk[k] = (k + 1 < q) ? umin64(B.cap[k], limit) : limit;
        }
        B.chunk[0] = umax64(TC, umin64(B.cap[0], limit));
        if (B.max_depth_cand)  // capped: start small, grow from the observed fan-out
            for (uint32_t k = 0; k < GFSM_MAXQ; k++) B.chunk[k] = umax64(TC, umin64(B.chunk[k], B.max_depth_cand / 64));
        B.cand_done = 0;
        B.count = B.dg0 = B.dg1 = 0;
        B.steps = B.overflows = 0;
        B.error = 0;
        B.step = STEP_NONE;
        state = SLOT_RUN;
    } else if (state == SLOT_RUN) {
        const Totals t = B.tot;
        const uint64_t b = B.cur_lo, e = B.cur_hi;
        const uint32_t d = B.d;
        if (B.step == STEP_ROOT) {
            B.root_next = e;
            B.cand[0] += e - b;
            B.level[0] += t.keep;
            if (t.keep) {
                B.S[0] = t.keep;
                B.lo[1] = 0;
                B.hi[1] = t.cnt;
            }
            if (B.max_depth_cand) B.chunk[0] = capped_chunk(B, 0, e - b, t.cnt, B.chunk[0], TC, limit);
            B.steps++;
        } else if (B.step == EXPAND_MID) {
            B.steps++;
            if (t.keep > B.cap[d]) {
                B.overflows++;
                B.chunk[d] = umax64(TC, (e - b) / 2);
            } else {
                B.lo[d] = e;
                B.cand[d] += e - b;
                B.level[d] += t.keep;
                if (t.keep) {
                    const uint64_t est = B.cap[d] / 10 * 7 * (e - b) / t.keep;
                    B.chunk[d] = umax64(TC, umin64(est, limit));
                    if (B.max_depth_cand) B.chunk[d] = capped_chunk(B, d, e - b, t.cnt, B.chunk[d], TC, limit);
                    B.S[d] = t.keep;
                    B.lo[d + 1] = 0;
                    B.hi[d + 1] = t.cnt;
                } else {
                    B.chunk[d] = umin64(B.chunk[d] * 4, limit);
                }
            }
        } else if (B.step == EXPAND_LAST_COUNT) {
            B.steps++;
            B.lo[d] = e;
            B.cand[d] += e - b;
            B.level[d] += t.count;
            B.count += t.count;
            B.dg0 += t.dg0;
            B.dg1 += t.dg1;
        }
    } else {
        B.step = STEP_NONE;
        return 0;
    }

    // Next step: the deepest depth with candidates left, else more roots.
    // Every depth's range is read at once (independent loads in flight
    // together) rather than one depth at a time: the scheduler is one
    // workgroup, and its dependent loads are on the critical path of every
    // step.
    int32_t d = 0;
    {
        bool left[GFSM_MAXQ];
#pragma unroll
        for (uint32_t k = 1; k < GFSM_MAXQ; k++) left[k] = k < q && B.lo[k] < B.hi[k];
#pragma unroll
        for (uint32_t k = 1; k < GFSM_MAXQ; k++)
            if (left[k]) d = (int32_t)k;
    }
    uint32_t ntiles = 0;
    B.tot = Totals{};
    // Deterministic cap: depth k tests only its first max_depth_cand
    // candidates (B.cand[k], roots included); the next step's depth has none
    // left -> the match stops (every deeper range is finished by then).
    const uint32_t dk = (d >= 1) ? (uint32_t)d : 0u;
    if ((d >= 1 || B.root_next < B.nroots) && B.max_depth_cand && B.cand[dk] >= B.max_depth_cand) {
        B.state = SLOT_CAP;
        B.step = STEP_NONE;
        return 0;
    }
    if (d >= 1) {
        const uint32_t kind = (d == (int32_t)q - 1) ? (uint32_t)EXPAND_LAST_COUNT : (uint32_t)EXPAND_MID;
        const uint64_t b = B.lo[d];
        uint64_t e = umin64(B.hi[d], b + umax64(TC, umin64(B.chunk[d], limit)));
        if (B.max_depth_cand) e = umin64(e, b + (B.max_depth_cand - B.cand[d]));
        const DepthPlanDev &dp = B.plan->depth[d];
        if (B.S[d - 1] == 0 || B.S[d - 1] > B.cap[d - 1] || dp.gen_dist == 0 || dp.gen_dist > (uint32_t)d) {
            B.state = SLOT_ERROR;
            B.error = 1;
            B.step = STEP_NONE;
            return 0;
        }
        if (B.max_cand && B.cand_done + (e - b) > B.max_cand) {
            B.state = SLOT_CAP;
            B.step = STEP_NONE;
            return 0;
        }
        B.cand_done += e - b;
        StepIO io{};
        io.in_vec = B.vec[d - 1];
        io.in_off = B.off[d - 1];
        io.in_gbeg = B.gbeg[d - 1];
        io.S_prev = B.S[d - 1];
        io.g_begin = b;
        io.g_end = e;
        if (kind == EXPAND_MID) {
            io.out_vec = B.vec[d];
            io.out_off = B.off[d];
            io.out_gbeg = B.gbeg[d];
            io.cap = B.cap[d];
            io.need_nxt = dp.nxt_dist > 1 ? 1u : 0u;
        }
        io.tot = tot;
        io.ntiles = (e - b + TC - 1) / TC;
        io.d = (uint32_t)d;
        io.q = q;
        io.kind = kind;
        io.digest = (kind == EXPAND_LAST_COUNT) ? B.digest : 0u;
        B.io = io;
        B.step = kind;
        B.d = (uint32_t)d;
        B.cur_lo = b;
        B.cur_hi = e;
        ntiles = (uint32_t)io.ntiles;
    } else if (B.root_next < B.nroots) {
        const uint64_t b = B.root_next;
        uint64_t e = umin64(B.nroots, b + B.chunk[0]);
        if (B.max_depth_cand) e = umin64(e, b + (B.max_depth_cand - B.cand[0]));
        RootIO r{};
        r.cand_vtx = B.root_vtx + b;
        r.cand_beg = B.root_beg + b;
        r.cand_cnt = B.root_cnt + b;
        r.n = e - b;
        r.out_vec = B.vec[0];
        r.out_off = B.off[0];
        r.out_gbeg = B.gbeg[0];
        r.cap = B.cap[0];
        r.tot = tot;
        r.ntiles = (e - b + TC - 1) / TC;
        B.rio = r;
        B.step = STEP_ROOT;
        B.d = 0;
        B.cur_lo = b;
        B.cur_hi = e;
        ntiles = (uint32_t)r.ntiles;
    } else {
        state = SLOT_DONE;
        B.step = STEP_NONE;
    }
    B.state = state;
    return ntiles;
}

// N words copied by the 64 lanes of one wave, each lane issuing 16 loads
// before their stores (a slot is 748 words, a plan 2,594).
template <uint32_t N, class SP, class DP>
DEV void wave_copy_words(DP dst, SP src, uint32_t lane)
{
    constexpr uint32_t kBatch = 16;
    for (uint32_t w0 = 0; w0 < N; w0 += 64 * kBatch) {
        uint32_t v[kBatch];
#pragma unroll
        for (uint32_t i = 0; i < kBatch; i++) {
            const uint32_t w = w0 + i * 64 + lane;
            v[i] = w < N ? src[w] : 0u;
        }
#pragma unroll
        for (uint32_t i = 0; i < kBatch; i++) {
            const uint32_t w = w0 + i * 64 + lane;
            if (w < N) dst[w] = v[i];
        }
    }
}

// Orders a wave's LDS writes before its other lanes' reads.
DEV void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One wave per slot: install the slot's queued admission, copy the slot into
// LDS (its words in flight together), schedule it there -- the scheduling
// logic's chain of dependent reads and writes then runs at LDS latency, not
// at that of global memory -- and write it back.  The last workgroup to
// finish lays the slots' tiles end to end.
constexpr uint32_t kSlotWords = sizeof(BatchSlot) / 4;
constexpr uint32_t kPlanWords = sizeof(QueryPlanDev) / 4;
// tile_start[s] between a slot's scheduling and the layout: its tile count,
// whether it is running, and whether its step reads the tile index.
constexpr uint64_t kSchedRun = 1ull << 63, kSchedIndexed = 1ull << 62;

__global__ __launch_bounds__(kTile) void k_batch_schedule(BatchCfg cfg)
{
    static_assert(sizeof(BatchSlot) % 8 == 0 && sizeof(QueryPlanDev) % 4 == 0, "word copies");
    __shared__ __align__(16) uint64_t s_slot[kWaves][kSlotWords / 2];
    __shared__ uint32_t s_wt[kWaves], s_wr[kWaves];
    __shared__ uint32_t s_last;
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    const uint32_t s = blockIdx.x * kWaves + wave;
    const auto slots = gp(cfg.slots);
    const auto tile_start = gp(cfg.tile_start);
    if (s < cfg.nslots) {
        uint32_t *lw = reinterpret_cast<uint32_t *>(s_slot[wave]);
        // this slot's queued admission, if any (its entry is found by a ballot)
        const uint32_t n_admit = gp(cfg.ctl)->n_admit;
        const uint32_t ne = n_admit < cfg.nslots ? n_admit : cfg.nslots;
        const auto admit = gp(cfg.admit);
        uint32_t found = ~0u;
        for (uint32_t e0 = 0; e0 < ne; e0 += 64) {
            const uint32_t e = e0 + lane;
            const uint64_t m = __ballot(e < ne && admit[e].slot == s);
            if (m) {
                found = e0 + (uint32_t)__builtin_ctzll(m);
                break;
            }
        }
        if (found != ~0u) {
            wave_copy_words<kSlotWords>(lw, reinterpret_cast<gptr_t<const uint32_t>>(&admit[found].s), lane);
            wave_copy_words<kPlanWords>(reinterpret_cast<gptr_t<uint32_t>>(gp(cfg.plans) + s),
                                        reinterpret_cast<gptr_t<const uint32_t>>(&admit[found].plan), lane);
        } else {
            wave_copy_words<kSlotWords>(lw, reinterpret_cast<gptr_t<const uint32_t>>(slots + s), lane);
        }
        wave_sync_lds();
        BatchSlot &B = *reinterpret_cast<BatchSlot *>(lw);
        uint32_t fin = 0;
        if (lane == 0) {
            const uint32_t before = B.state;
            const uint32_t nt = schedule_slot(cfg, B, &cfg.slots[s].tot);
            fin = (B.state >= SLOT_DONE && before < SLOT_DONE) ? 1u : 0u;  // just finished: its statistics
            SlotReport &rp = cfg.report[s];
            rp.state = B.state;
            rp.steps = B.steps;
            rp.count = B.count;
            rp.dg0 = B.dg0;
            rp.dg1 = B.dg1;
            const bool indexed = nt && (B.step == EXPAND_MID || B.step == EXPAND_LAST_COUNT);
            tile_start[s] = (uint64_t)nt | (B.state == SLOT_RUN ? kSchedRun : 0ull) | (indexed ? kSchedIndexed : 0ull);
        }
        fin = __builtin_amdgcn_readfirstlane(fin);
        wave_sync_lds();
        if (fin) {
            SlotFinal &f = cfg.final[s];
            if (lane == 0) f.overflows = B.overflows;
            if (lane < GFSM_MAXQ) {
                f.level[lane] = B.level[lane];
                f.cand[lane] = B.cand[lane];
            }
        }
        wave_copy_words<kSlotWords>(reinterpret_cast<gptr_t<uint32_t>>(slots + s), lw, lane);
    }
    // The last workgroup to finish lays the tiles out: every other one has
    // written its slots (released by its count, acquired here).
    __syncthreads();
    if (tid == 0) {
        const uint32_t done =
            __hip_atomic_fetch_add(&gp(cfg.ctl)->sched_done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (done + 1 == gridDim.x) ? 1u : 0u;
    }
    __syncthreads();
    if (!s_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const uint32_t per = (cfg.nslots + kTile - 1) / kTile;  // <= kMaxSlots / kTile
    const uint32_t s0 = tid * per;
    uint64_t packed[kMaxSlots / kTile];
    uint32_t my_tiles = 0, my_run = 0;
#pragma unroll
    for (uint32_t u = 0; u < kMaxSlots / kTile; u++) {
        packed[u] = (u < per && s0 + u < cfg.nslots) ? tile_start[s0 + u] : 0ull;
        my_tiles += (uint32_t)packed[u];
        my_run += (packed[u] & kSchedRun) ? 1u : 0u;
    }
    const uint32_t inc = wave_incl_scan(my_tiles);
    const uint32_t rinc = wave_incl_scan(my_run);
    if (lane == 63) {
        s_wt[wave] = inc;
        s_wr[wave] = rinc;
    }
    __syncthreads();
    uint32_t base = inc - my_tiles, total = 0, running = 0;
    for (uint32_t w = 0; w < kWaves; w++) {
        if (w < wave) base += s_wt[w];
        total += s_wt[w];
        running += s_wr[w];
    }
#pragma unroll
    for (uint32_t u = 0; u < kMaxSlots / kTile; u++) {
        const uint32_t sl = s0 + u;
        if (u < per && sl < cfg.nslots) {
            const uint32_t nt = (uint32_t)packed[u];
            tile_start[sl] = base;
            slots[sl].tile0 = base;
            // an expand step reads its tiles' frontier nodes from the index
            if (packed[u] & kSchedIndexed) slots[sl].io.tile_node = cfg.tile_node + 2 * base;
            base += nt;
        }
    }
    if (tid == 0) {
        tile_start[cfg.nslots] = total;
        gp(cfg.ctl)->total_tiles = total;
        gp(cfg.ctl)->running = running;
        gp(cfg.ctl)->n_admit = 0;     // every workgroup has read it
        gp(cfg.ctl)->sched_done = 0;  // for the next launch
    }
}

// Slot owning global tile t: the last s in [0, nslots) with start[s] <= t
// (empty slots share their start with the next slot), by one wave in 64-ary
// steps.
DEV uint32_t slot_search(const uint64_t *__restrict__ start, uint32_t nslots, uint64_t t)
{
    uint32_t lo = 0, hi = nslots;
    while (hi - lo > 1) {
        const uint32_t n = hi - lo;
        const uint32_t stride = (n + 63) / 64;
        const uint32_t pos = lo + lane_id() * stride;
        const bool ok = pos < hi && start[pos] <= t;
        const uint64_t m = __ballot(ok);
        const uint32_t top = 63u - (uint32_t)__builtin_clzll(m | 1ull);
        lo = lo + top * stride;
        hi = (lo + stride < hi) ? lo + stride : hi;
    }
    return lo;
}

// One wave per tile of the scheduled step: the owning slot, then (expand
// steps) the nodes holding the tile's first and last candidates, one
// half-wave each, as k_tile_index does for the single-query path.
__global__ __launch_bounds__(kTile) void k_batch_tile_index(BatchCfg cfg)
{
    const uint64_t total = cfg.ctl->total_tiles;
    const uint64_t TC = (uint64_t)kTile * cfg.items;
    const uint32_t lane = lane_id();
    const uint64_t nwaves = ((uint64_t)gridDim.x * kTile) >> 6;
    for (uint64_t t = ((uint64_t)blockIdx.x * kTile + threadIdx.x) >> 6; t < total; t += nwaves) {
        const uint32_t s = slot_search(cfg.tile_start, cfg.nslots, t);
        const BatchSlot &B = cfg.slots[s];
        if (B.step != EXPAND_MID && B.step != EXPAND_LAST_COUNT) continue;
        const uint64_t lt = t - B.tile0;
        const uint64_t c0 = B.io.g_begin + lt * TC;
        const uint64_t c1 = (c0 + TC < B.io.g_end) ? c0 + TC : B.io.g_end;
        const uint64_t key = (lane < 32) ? c0 : c1 - 1;
        const uint64_t i = frontier_search(gp(B.io.in_off), B.io.S_prev, key, lane & 31);
        if ((lane & 31) == 0) gp(cfg.tile_node)[2 * t + (lane >> 5)] = i;
    }
}

// Four waves per SIMD, as k_expand: at 8 candidates per thread the step
// kernel would take 207 VGPRs (two waves per SIMD); capped at 128 it spills,
// but with a 7,680-word pool four workgroups fit a CU and the batched call
// runs 15% faster (measured: 48.0 against 56.6 ms on the 256-query batch).
template <uint32_t I>
__global__ __launch_bounds__(kTile) __attribute__((amdgpu_waves_per_eu(4))) void k_batch_step(BatchCfg cfg)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ TileShared<I> sh;
    const uint64_t total = cfg.ctl->total_tiles;
    const uint32_t ep = read_epoch(cfg.ws);
    tickets_init(cfg.ws, &sh.next);
    for (;;) {
        const uint64_t t = tickets_take(&sh.next);
        if (t >= total) break;
        if (threadIdx.x < 64) {
            const uint32_t s = slot_search(cfg.tile_start, cfg.nslots, t);
            if (threadIdx.x == 0) sh.slot = s;
        }
        __syncthreads();
        const uint32_t s = __builtin_amdgcn_readfirstlane(sh.slot);
        const BatchSlot &B = cfg.slots[s];
        const uint32_t step = B.step;
        const uint64_t t0 = B.tile0;
        const auto plan = gp(B.plan);  // loaded from memory: keep its reads global
        if (step == STEP_ROOT) {
            root_tile<I>(cfg.g, B.rio, plan->nroot_chk, plan->root_chk_rel, t - t0, t, t0, cfg.ws, ep, sh.sm, &sh.next);
        } else if (step == EXPAND_MID) {
            // the slot's step descriptor in scalar registers (measured: the
            // 256-query batch 36.3 against 39.3 ms)
            uint64_t c = 0, h0 = 0, h1 = 0;
            const StepIO io = uniform_copy(B.io);
            expand_tile<EXPAND_MID, I>(cfg.g, plan->depth[B.d], plan->col, plan->dcol, io, t - t0, t, t0, cfg.ws,
                                       ep, cfg.pool_words, sh, smem, c, h0, h1);
        } else {
            uint64_t c = 0, h0 = 0, h1 = 0;
            const StepIO io = uniform_copy(B.io);
            expand_tile<EXPAND_LAST_COUNT, I>(cfg.g, plan->depth[B.d], plan->col, plan->dcol, io, t - t0, t, t0,
                                              cfg.ws, ep, cfg.pool_words, sh, smem, c, h0, h1);
            flush_sums(c, h0, h1, io.tot, sh.red);
        }
    }
    finish_launch(cfg.ws);
}

// ---------------------------------------------------------------------------
// chained-scan self test and workspace maintenance
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kTile) void k_scan_test(const uint32_t *vals, uint64_t n, uint64_t *out,
                                                    uint32_t *idx, ScanWS ws, Totals *tot, uint64_t ntiles)
{
    constexpr uint32_t I = kScanTestItems;
    __shared__ TileShared<I> sh;
    const uint32_t ep = read_epoch(ws);
    tickets_init(ws, &sh.next);
    for (;;) {
        const uint64_t t = tickets_take(&sh.next);
        if (t >= ntiles) break;
        bool keep[I];
        uint32_t v[I];
#pragma unroll
        for (uint32_t k = 0; k < I; k++) {
            const uint64_t i = t * (kTile * I) + k * kTile + threadIdx.x;
            v[k] = i < n ? vals[i] : 0u;
            keep[k] = v[k] != 0;
        }
        uint64_t kx[I], off[I], rk, rc;
        tile_scan<I>(keep, v, t, 0, ws, ep, sh.sm, &sh.next, kx, off, rk, rc);
#pragma unroll
        for (uint32_t k = 0; k < I; k++) {
            const uint64_t i = t * (kTile * I) + k * kTile + threadIdx.x;
            if (i < n) out[i] = off[k];
            if (keep[k]) idx[kx[k]] = (uint32_t)i;
        }
        if (t == ntiles - 1 && threadIdx.x == 0) {
            tot->keep = rk;
            tot->cnt = rc;
        }
    }
    finish_launch(ws);
}

__global__ void k_dense_seg(const uint32_t *vtx, const uint32_t *beg, const uint32_t *cnt, uint64_t n, uint2 *out)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[vtx[i]] = make_uint2(beg[i], beg[i] + cnt[i]);
}

// One half-wave per search key: key 2t is tile t's first candidate, key 2t+1
// its last.
// One search per tile boundary: key t < ntiles is tile t's first candidate,
// key ntiles the step's last.  Every frontier node has at least one child
// (offsets strictly increase), so the candidate before a boundary belongs to
// the boundary's node when that node starts earlier, else to the node before.
__global__ __launch_bounds__(kTile) void k_tile_index(const uint64_t *off, uint64_t S, uint64_t g_begin,
                                                      uint64_t g_end, uint64_t tc, uint64_t ntiles, uint64_t *out)
{
    const uint64_t t = ((uint64_t)blockIdx.x * kTile + threadIdx.x) >> 5;
    if (t > ntiles) return;  // whole half-waves leave together (32 | kTile)
    const uint64_t key = (t < ntiles) ? g_begin + t * tc : g_end - 1;
    const uint64_t i = frontier_search(off, S, key, threadIdx.x & 31);
    if ((threadIdx.x & 31) != 0) return;
    if (t == ntiles) {
        out[2 * t - 1] = i;  // the last tile's last node
        return;
    }
    out[2 * t] = i;
    if (t > 0) out[2 * t - 1] = (off[i] < key || i == 0) ? i : i - 1;
}

__global__ void k_seg_fill(uint4 *slots, uint32_t mask)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= mask;
         i += (uint64_t)gridDim.x * blockDim.x)
        slots[i] = make_uint4(kSegEmpty, 0u, 0u, 0u);
}

// One thread per vertex inserts its directory entries; a slot is claimed by
// a compare-and-swap on its vertex word (keys are unique, and the table is
// at most half full, so every insertion finds a free slot).
__global__ void k_seg_insert(const uint32_t *dir_ptr, const uint32_t *dir, uint32_t V, uint4 *slots, uint32_t mask)
{
    for (uint64_t xi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; xi < V;
         xi += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)xi;
        const uint32_t jb = dir_ptr[x], je = dir_ptr[x + 1];
        for (uint32_t j = jb; j < je; j++) {
            const uint32_t rel = dir[2 * j], beg = dir[2 * j + 1], end = dir[2 * j + 3];
            uint32_t p = seg_hash(x, rel) & mask;
            for (;;) {
                uint32_t *key = reinterpret_cast<uint32_t *>(slots + p);
                if (atomicCAS(key, kSegEmpty, x) == kSegEmpty) {
                    key[1] = rel;
                    key[2] = beg;
                    key[3] = end;
                    break;
                }
                p = (p + 1) & mask;
            }
        }
    }
}

__global__ void k_ws_reset(ScanWS ws, Totals *tot)
{
    if (threadIdx.x < 4) ws.ctl[threadIdx.x] = 0u;
    if (tot && threadIdx.x < sizeof(Totals) / 8) reinterpret_cast<uint64_t *>(tot)[threadIdx.x] = 0ull;
}

__global__ void k_ws_clear(ScanWS ws)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ws.max_tiles;
         i += (uint64_t)gridDim.x * blockDim.x) {
        ws.wk[i] = 0;
        ws.wc[i] = 0;
    }
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------

size_t expand_lds_bytes(uint32_t items, uint32_t pool_words)
{
    const size_t TC = (size_t)kTile * items;
    return TC / 64 * 8 + TC * 4 + (size_t)pool_words * 4;  // bitmap, generator bases, pool
}

// Raise a kernel's dynamic LDS limit once (before any launch or capture).
template <class K>
static hipError_t allow_lds(K kernel, bool &done)
{
    if (done) return hipSuccess;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxDynLds);
    if (e == hipSuccess) done = true;
    return e;
}

template <uint32_t KIND, uint32_t I>
static hipError_t launch_expand_t(const ExpandArgs &a, uint32_t grid, hipStream_t s)
{
    static bool attr = false;
    const size_t lds = expand_lds_bytes(I, a.pool_words);
    if (lds > kMaxDynLds) return hipErrorInvalidValue;
    hipError_t e = allow_lds(&k_expand<KIND, I>, attr);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_expand<KIND, I>), dim3(grid), dim3(kTile), lds, s, a);
    return hipGetLastError();
}

template <uint32_t I>
static hipError_t launch_expand_bare_t(const ExpandArgs &a, uint32_t grid, hipStream_t s)
{
    static bool attr = false;
    const size_t lds = expand_lds_bytes(I, a.pool_words);
    if (lds > kMaxDynLds) return hipErrorInvalidValue;
    hipError_t e = allow_lds(&k_expand_bare<I>, attr);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_expand_bare<I>), dim3(grid), dim3(kTile), lds, s, a);
    return hipGetLastError();
}

static hipError_t launch_expand_bare(const ExpandArgs &a, uint32_t grid, hipStream_t s)
{
    switch (a.items) {
    case 1: return launch_expand_bare_t<1>(a, grid, s);
    case 2: return launch_expand_bare_t<2>(a, grid, s);
    case 4: return launch_expand_bare_t<4>(a, grid, s);
    case 8: return launch_expand_bare_t<8>(a, grid, s);
    default: return hipErrorInvalidValue;
    }
}

template <uint32_t KIND>
static hipError_t launch_expand_k(const ExpandArgs &a, uint32_t grid, hipStream_t s)
{
    switch (a.items) {
    case 1: return launch_expand_t<KIND, 1>(a, grid, s);
    case 2: return launch_expand_t<KIND, 2>(a, grid, s);
    case 4: return launch_expand_t<KIND, 4>(a, grid, s);
    case 8: return launch_expand_t<KIND, 8>(a, grid, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_expand(const ExpandArgs &a, uint32_t grid, hipStream_t s)
{
    switch (a.io.kind) {
    case EXPAND_MID: return launch_expand_k<EXPAND_MID>(a, grid, s);
    case EXPAND_LAST_COUNT: return launch_expand_k<EXPAND_LAST_COUNT>(a, grid, s);
    case EXPAND_LAST_ROWS: return launch_expand_k<EXPAND_LAST_ROWS>(a, grid, s);
    case EXPAND_LAST_BARE: return launch_expand_bare(a, grid, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_dense_seg(const uint32_t *vtx, const uint32_t *beg, const uint32_t *cnt, uint64_t n, uint32_t V,
                            uint2 *out, hipStream_t s)
{
    hipError_t e = hipMemsetAsync(out, 0, (size_t)V * sizeof(uint2), s);
    if (e != hipSuccess || n == 0) return e;
    const uint64_t grid = (n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192;
    hipLaunchKernelGGL(k_dense_seg, dim3((uint32_t)grid), dim3(256), 0, s, vtx, beg, cnt, n, out);
    return hipGetLastError();
}

hipError_t launch_tile_index(const uint64_t *off, uint64_t S, uint64_t g_begin, uint64_t g_end, uint64_t tc,
                             uint64_t ntiles, uint64_t *out, hipStream_t s)
{
    if (ntiles == 0 || S == 0 || g_end <= g_begin) return hipErrorInvalidValue;
    const uint64_t keys_per_wg = kTile / 32;
    const uint64_t grid = (ntiles + keys_per_wg) / keys_per_wg;  // ntiles + 1 keys
    hipLaunchKernelGGL(k_tile_index, dim3((uint32_t)grid), dim3(kTile), 0, s, off, S, g_begin, g_end, tc, ntiles, out);
    return hipGetLastError();
}

hipError_t launch_single(const SingleArgs &a, uint32_t grid, hipStream_t s)
{
    if (a.rows)
        hipLaunchKernelGGL(k_single<true>, dim3(grid), dim3(kTile), 0, s, a);
    else
        hipLaunchKernelGGL(k_single<false>, dim3(grid), dim3(kTile), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_root(const RootArgs &a, uint32_t grid, hipStream_t s)
{
    hipLaunchKernelGGL((k_root<kRootItems>), dim3(grid), dim3(kTile), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_batch_schedule(const BatchCfg &cfg, hipStream_t s)
{
    if (cfg.nslots == 0 || cfg.nslots > kMaxSlots) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_batch_schedule, dim3((cfg.nslots + kWaves - 1) / kWaves), dim3(kTile), 0, s, cfg);
    return hipGetLastError();
}

template <uint32_t I>
static hipError_t launch_batch_step_t(const BatchCfg &cfg, uint32_t grid, hipStream_t s)
{
    static bool attr = false;
    const size_t lds = expand_lds_bytes(I, cfg.pool_words);
    if (lds > kMaxDynLds) return hipErrorInvalidValue;
    hipError_t e = allow_lds(&k_batch_step<I>, attr);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_batch_step<I>), dim3(grid), dim3(kTile), lds, s, cfg);
    return hipGetLastError();
}

hipError_t launch_batch_tile_index(const BatchCfg &cfg, hipStream_t s)
{
    if (cfg.nslots == 0 || cfg.nslots > kMaxSlots || !cfg.tile_node) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_batch_tile_index, dim3(2048), dim3(kTile), 0, s, cfg);
    return hipGetLastError();
}

hipError_t launch_batch_step(const BatchCfg &cfg, uint32_t grid, hipStream_t s)
{
    if (cfg.nslots == 0 || cfg.nslots > kMaxSlots) return hipErrorInvalidValue;
    switch (cfg.items) {
    case 1: return launch_batch_step_t<1>(cfg, grid, s);
    case 2: return launch_batch_step_t<2>(cfg, grid, s);
    case 4: return launch_batch_step_t<4>(cfg, grid, s);
    case 8: return launch_batch_step_t<8>(cfg, grid, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_scan_test(const uint32_t *vals, uint64_t n, uint64_t *out, uint32_t *idx, ScanWS ws,
                            Totals *tot, uint32_t grid, hipStream_t s)
{
    const uint64_t tc = (uint64_t)kTile * kScanTestItems;
    const uint64_t ntiles = (n + tc - 1) / tc;
    hipLaunchKernelGGL(k_scan_test, dim3(grid), dim3(kTile), 0, s, vals, n, out, idx, ws, tot, ntiles);
    return hipGetLastError();
}

hipError_t launch_seg_hash_build(const uint32_t *dir_ptr, const uint32_t *dir, uint32_t V, uint4 *slots,
                                 uint32_t mask, hipStream_t s)
{
    hipLaunchKernelGGL(k_seg_fill, dim3(2048), dim3(256), 0, s, slots, mask);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_seg_insert, dim3(4096), dim3(256), 0, s, dir_ptr, dir, V, slots, mask);
    return hipGetLastError();
}

// Traffic calibration kernels of known HBM footprint (for scaling the byte
// counters of a profiler run).  Stream: dst[i] = src[i] + 1 over n words.
// Gather: dst[i] = src[32 * line(i)] for i < n, line(i) = i * odd mod n (n a
// power of two): one word from each of n distinct 128-byte lines, in an order
// that defeats the caches, written contiguously.
__global__ void k_cal_stream(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i] + 1u;
}

__global__ void k_cal_gather(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t line = (i * 0x9E3779B97F4A7C15ull) & (n - 1);
        dst[i] = src[line * 32];
    }
}

// Both 64-byte halves of each line: dst[2i], dst[2i+1] = words 0 and 16 of
// line(i).  Against k_cal_gather this tells whether a scattered read moves a
// whole 128-byte line or only the 64-byte half it touches.
__global__ void k_cal_gather2(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t line = (i * 0x9E3779B97F4A7C15ull) & (n - 1);
        dst[2 * i] = src[line * 32];
        dst[2 * i + 1] = src[line * 32 + 16];
    }
}

hipError_t launch_calibrate(uint32_t mode, const uint32_t *src, uint32_t *dst, uint64_t n, hipStream_t s)
{
    if (mode > 2 || (mode >= 1 && (n & (n - 1)))) return hipErrorInvalidValue;
    const uint32_t grid = 8192;
    if (mode == 0)
        hipLaunchKernelGGL(k_cal_stream, dim3(grid), dim3(256), 0, s, src, dst, n);
    else if (mode == 1)
        hipLaunchKernelGGL(k_cal_gather, dim3(grid), dim3(256), 0, s, src, dst, n);
    else
        hipLaunchKernelGGL(k_cal_gather2, dim3(grid), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

// The allocation-free match's scratch.  In host code `scratch` names a
// host-side stand-in, not device memory: every use goes through the address
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch(int device, hipStream_t s, char **ptr)
{
    static std::mutex mu;
    static char *addr[64] = {};
    *ptr = nullptr;
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[device]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess) e = hipSetDevice(device);
        if (e != hipSuccess) return e;
        char *p = nullptr;
        e = hipGetSymbolAddress(reinterpret_cast<void **>(&p), HIP_SYMBOL(scratch));
        if (e == hipSuccess && !p) e = hipErrorInvalidSymbol;
        if (e == hipSuccess) {
            // Zero the workspace words once: the control block, totals, tile
            // words and tile index (each call clears what it uses again).
            ScanWS ws{reinterpret_cast<uint64_t *>(p + 512), reinterpret_cast<uint64_t *>(p + 512) + kNoAllocTiles,
                      reinterpret_cast<uint32_t *>(p), kNoAllocTiles};
            hipLaunchKernelGGL(k_ws_clear, dim3(1024), dim3(256), 0, s, ws);
            e = hipGetLastError();
            if (e == hipSuccess) {
                hipLaunchKernelGGL(k_ws_reset, dim3(1), dim3(64), 0, s, ws, reinterpret_cast<Totals *>(p + 256));
                e = hipGetLastError();
            }
            if (e == hipSuccess) e = hipStreamSynchronize(s);
        }
        (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        addr[device] = p;
    }
    *ptr = addr[device];
    return hipSuccess;
}

hipError_t launch_ws_reset(ScanWS ws, Totals *tot, hipStream_t s)
{
    hipLaunchKernelGGL(k_ws_reset, dim3(1), dim3(64), 0, s, ws, tot);
    return hipGetLastError();
}

hipError_t launch_ws_clear(ScanWS ws, hipStream_t s)
{
    hipLaunchKernelGGL(k_ws_clear, dim3(1024), dim3(256), 0, s, ws);
    return hipGetLastError();
}

}  // namespace gfsm
