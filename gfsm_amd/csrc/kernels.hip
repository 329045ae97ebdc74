// kernels.hip -- CDNA4 (gfx950) kernels of the forest-based matcher.
//
// One launch per (depth, candidate range).  A workgroup of 256 threads
// (4 wave64s) owns a tile of 256 * I consecutive candidates of the depth
// (I = 1, 2, 4 or 8 per thread, striped so that loads stay coalesced):
//   1. one wave finds the frontier nodes whose candidate ranges cover the
//      tile (a 32-ary search per tile end over their exclusive offsets) and
//      marks where each node's candidates start in an LDS bitmap,
//   2. each candidate finds its node from the bitmap (mbcnt of a wave-uniform
//      word), reads itself coalesced from the node's generator segment, is
//      checked for injectivity against the node's placed vertices (a row of
//      the previous level, read through the cache) and binary-searched in the
//      segments of the checked ancestors (staged per node in LDS),
//   3. survivors are compacted with wave64 ballot/mbcnt counts and DPP prefix
//      sums of their next-depth candidate counts, chained across workgroups
//      by a decoupled look-back, so that the next level -- rows, candidate
//      offsets and generator segment begins -- comes out of the same pass.
// This replaces hnGen + hnCheck + CUB ExclusiveSum + hnRemove + the degree
// scan of one reference depth (HnGPU.cu:114-262, 309-349, 371-439).
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace gfsm {

#define DEV __device__ __forceinline__

constexpr uint32_t kWaves = kTile / 64;

// ---------------------------------------------------------------------------
// wave64 primitives
// ---------------------------------------------------------------------------

DEV uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

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
// graph access
// ---------------------------------------------------------------------------

// [beg, end) of the (x, rel) segment, {0, 0} when x has none.
DEV uint2 seg_lookup(const DevGraphView &g, uint32_t x, uint32_t rel)
{
    const uint2 *dir = reinterpret_cast<const uint2 *>(g.dir);
    uint32_t lo = g.dir_ptr[x], hi = g.dir_ptr[x + 1];
    uint32_t n = hi - lo;
    while (n > 0) {
        uint32_t h = n >> 1;
        if (dir[lo + h].x < rel) {
            lo += h + 1;
            n -= h + 1;
        } else {
            n = h;
        }
    }
    if (lo < hi) {
        uint2 e = dir[lo];
        if (e.x == rel) return make_uint2(e.y, dir[lo + 1].y);
    }
    return make_uint2(0u, 0u);
}

// Is w in the sorted slice nbr[beg, end)?
DEV bool seg_contains(const uint32_t *__restrict__ nbr, uint32_t beg, uint32_t end, uint32_t w)
{
    uint32_t lo = beg, n = end - beg;
    while (n > 0) {
        uint32_t h = n >> 1;
        if (nbr[lo + h] < w) {
            lo += h + 1;
            n -= h + 1;
        } else {
            n = h;
        }
    }
    return lo < end && nbr[lo] == w;
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

// Exclusive prefix of tile t over one word array, computed by all 64 lanes of
// one wave: each pass reads the words of the 64 preceding tiles, waits until
// every tile up to the nearest inclusive prefix has published, and sums.
DEV uint64_t lookback(const uint64_t *words, uint32_t ep, uint64_t t)
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
            st = 2u;  // before tile 0 the inclusive prefix is 0
            w = 0;
            if (idx >= 0) {
                w = ld_word(words + idx);
                st = (((w >> 46) & ((1u << kEpochBits) - 1)) == epm) ? (uint32_t)(w >> 62) : 0u;
            }
            pm = __ballot(st == 2u);
            const uint64_t xm = __ballot(st == 0u);
            const uint64_t upto = pm ? (pm ^ (pm - 1)) : ~0ull;
            if (!(xm & upto)) break;
            __builtin_amdgcn_s_sleep(1);
        }
        const uint64_t upto = pm ? (pm ^ (pm - 1)) : ~0ull;
        const uint64_t v = ((upto >> lane) & 1ull) ? (w & kValueMask) : 0ull;
        sum += wave_sum64(v);
        if (pm) break;
        base -= 64;
    }
    return sum;
}

template <uint32_t I>
struct ScanSmem {
    uint32_t wk[I][kWaves];
    uint64_t wc[I][kWaves];
    uint64_t ex[2];
};

// Tile-wide exclusive prefix of (keep, cnt) in candidate order (item k of
// thread x is tile position k * 256 + x), continued across tiles.  Returns
// each item's kept index and candidate offset and the totals through this
// tile.
template <uint32_t I>
DEV void tile_scan(const bool (&keep)[I], const uint32_t (&cnt)[I], uint64_t t, const ScanWS &ws, uint32_t ep,
                   ScanSmem<I> &sm, uint64_t (&my_k)[I], uint64_t (&my_c)[I], uint64_t &run_k, uint64_t &run_c)
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
    uint64_t ak = 0, ac = 0, offk[I], offc[I];
#pragma unroll
    for (uint32_t k = 0; k < I; k++) {
#pragma unroll
        for (uint32_t w = 0; w < kWaves; w++) {
            if (w == wave) {
                offk[k] = ak;
                offc[k] = ac;
            }
            ak += sm.wk[k][w];
            ac += sm.wc[k][w];
        }
    }
    if (wave == 0) {
        uint64_t ek = 0, ec = 0;
        if (t != 0) {
            if (lane == 0) {
                st_word(ws.wk + t, pack(1, ep, ak));
                st_word(ws.wc + t, pack(1, ep, ac));
            }
            ek = lookback(ws.wk, ep, t);
            ec = lookback(ws.wc, ep, t);
        }
        if (lane == 0) {
            st_word(ws.wk + t, pack(2, ep, ek + ak));
            st_word(ws.wc + t, pack(2, ep, ec + ac));
            sm.ex[0] = ek;
            sm.ex[1] = ec;
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < I; k++) {
        my_k[k] = sm.ex[0] + offk[k] + kpre[k];
        my_c[k] = sm.ex[1] + offc[k] + cpre[k];
    }
    run_k = sm.ex[0] + ak;
    run_c = sm.ex[1] + ac;
}

DEV uint32_t read_epoch(const ScanWS &ws)
{
    return __hip_atomic_load(ws.ctl + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Tile tickets: tiles are handed out in launch order, so a tile only waits on
// tiles already owned by running workgroups (a prefetched ticket is always
// processed next by its owner, after its current tile).  Thread 0 keeps one
// ticket in flight so the atomic's latency overlaps the current tile.
DEV void tickets_init(const ScanWS &ws, uint32_t &next)
{
    if (threadIdx.x == 0) next = atomicAdd(ws.ctl + 0, 1u);
}

DEV uint64_t tickets_take(const ScanWS &ws, uint32_t &next, uint64_t *s_tile, uint64_t ntiles)
{
    if (threadIdx.x == 0) *s_tile = next;
    __syncthreads();
    const uint64_t t = *s_tile;
    if (threadIdx.x == 0 && t < ntiles) next = atomicAdd(ws.ctl + 0, 1u);
    return t;
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

// Sum (count, digest0, digest1) over the workgroup and add it to the launch
// totals: one set of atomics per workgroup, not per wave.
DEV void flush_sums(uint64_t cnt, uint64_t h0, uint64_t h1, Totals *tot, uint64_t (*s_red)[kWaves])
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
        if (c) atomicAdd((unsigned long long *)&tot->count, (unsigned long long)c);
        if (a) atomicAdd((unsigned long long *)&tot->dg0, (unsigned long long)a);
        if (b) atomicAdd((unsigned long long *)&tot->dg1, (unsigned long long)b);
    }
}

// ---------------------------------------------------------------------------
// root level (hnInit, HnGPU.cu:76-109)
// ---------------------------------------------------------------------------

template <uint32_t I>
__global__ __launch_bounds__(kTile) void k_root(RootArgs a)
{
    __shared__ ScanSmem<I> sm;
    __shared__ uint64_t s_tile;
    const uint32_t ep = read_epoch(a.ws);
    uint32_t next = 0;
    tickets_init(a.ws, next);
    for (;;) {
        const uint64_t t = tickets_take(a.ws, next, &s_tile, a.ntiles);
        if (t >= a.ntiles) break;
        bool keep[I];
        uint32_t v[I], cnt[I];
#pragma unroll
        for (uint32_t k = 0; k < I; k++) {
            const uint64_t c = t * (kTile * I) + k * kTile + threadIdx.x;
            keep[k] = false;
            v[k] = 0;
            cnt[k] = 0;
            if (c < a.n) {
                v[k] = a.cand_vtx[c];
                cnt[k] = a.cand_cnt[c];
                bool ok = cnt[k] > 0;
                for (uint32_t r = 0; r < a.nchk && ok; r++) {
                    uint2 sg = seg_lookup(a.g, v[k], a.chk_rel[r]);
                    ok = seg_contains(a.g.nbr, sg.x, sg.y, v[k]);
                }
                keep[k] = ok;
                if (!ok) cnt[k] = 0;
            }
        }
        uint64_t kk[I], off[I], rk, rc;
        tile_scan<I>(keep, cnt, t, a.ws, ep, sm, kk, off, rk, rc);
#pragma unroll
        for (uint32_t k = 0; k < I; k++) {
            if (keep[k] && kk[k] < a.cap) {
                const uint64_t c = t * (kTile * I) + k * kTile + threadIdx.x;
                a.out_vec[kk[k]] = v[k];
                a.out_off[kk[k]] = off[k];
                a.out_gbeg[kk[k]] = a.cand_beg[c];
            }
        }
        if (t == a.ntiles - 1 && threadIdx.x == 0) {
            a.tot->keep = rk;
            a.tot->cnt = rc;
            if (rk <= a.cap) a.out_off[rk] = rc;
        }
    }
    finish_launch(a.ws);
}

// ---------------------------------------------------------------------------
// one depth: expand + filter + compact + next level
// ---------------------------------------------------------------------------

// Largest i in [0, S) with off[i] <= key, found by half a wave (32 lanes,
// `sub` = lane within the half) in 32-ary steps; off[0] = 0 <= key.
DEV uint64_t frontier_search(const uint64_t *__restrict__ off, uint64_t S, uint64_t key, uint32_t sub)
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

template <uint32_t KIND, uint32_t I>
__global__ __launch_bounds__(kTile) void k_expand(ExpandArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ ScanSmem<I> sm;
    __shared__ uint64_t s_tile;
    __shared__ uint64_t s_node[2];
    __shared__ uint64_t s_red[3][kWaves];
    constexpr uint32_t T = kTile, TC = kTile * I, NW = TC / 64;
    const uint32_t d = a.d;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = lane_id(), wave = tid >> 6;

    // LDS staging, sized for TC frontier nodes (a tile never spans more).
    // s_bits marks the tile positions where a node's candidates start: the
    // node of position x is (set bits at positions <= x) - 1.
    uint64_t *s_bits = reinterpret_cast<uint64_t *>(smem);                // [NW]
    uint32_t *s_seg = reinterpret_cast<uint32_t *>(s_bits + NW);          // [TC] nbr index of tile position 0
    uint2 *s_chk = reinterpret_cast<uint2 *>(s_seg + TC);                 // [nchk][TC] checked segments
    uint2 *s_nxt = s_chk + a.dp.nchk * TC;                                // [TC] next generator (need_nxt)

    const uint32_t ep = read_epoch(a.ws);
    uint64_t acc_cnt = 0, acc_h0 = 0, acc_h1 = 0;
    uint32_t next = 0;
    tickets_init(a.ws, next);
    for (;;) {
        const uint64_t t = tickets_take(a.ws, next, &s_tile, a.ntiles);
        if (t >= a.ntiles) break;
        const uint64_t c0 = a.g_begin + t * TC;
        const uint64_t c1 = (c0 + TC < a.g_end) ? c0 + TC : a.g_end;
        const uint32_t n = (uint32_t)(c1 - c0);

        // Frontier nodes covering [c0, c1): offsets are strictly increasing.
        if (tid < NW) s_bits[tid] = 0;
        if (tid < 64) {
            const uint32_t sub = tid & 31;
            const uint64_t key = (tid < 32) ? c0 : c1 - 1;
            const uint64_t i = frontier_search(a.in_off, a.S_prev, key, sub);
            if (sub == 0) s_node[tid >> 5] = i;
        }
        __syncthreads();
        const uint64_t ilo = s_node[0];
        const uint32_t nf = (uint32_t)(s_node[1] - ilo + 1);

        // Per frontier node: start bit, generator base, checked segments.
        for (uint32_t j = tid; j < nf; j += T) {
            const uint64_t i = ilo + j;
            const uint64_t o = a.in_off[i];
            const uint32_t first = (o > c0) ? (uint32_t)(o - c0) : 0u;
            atomicOr((unsigned long long *)&s_bits[first >> 6], 1ull << (first & 63));
            s_seg[j] = a.in_gbeg[i] + (uint32_t)(c0 - o);
            const uint32_t *av = a.in_vec + i * d;
            for (uint32_t r = 0; r < a.dp.nchk; r++) {
                const uint32_t dist = a.dp.chk_dist[r];
                if (dist) s_chk[r * TC + j] = seg_lookup(a.g, av[d - dist], a.dp.chk_rel[r]);
            }
            if (KIND == EXPAND_MID && a.need_nxt) s_nxt[j] = seg_lookup(a.g, av[d + 1 - a.dp.nxt_dist], a.dp.nxt_rel);
        }
        __syncthreads();

        bool keep[I];
        uint32_t cnt[I], w[I], gb[I];
        uint64_t node[I];
#pragma unroll
        for (uint32_t k = 0; k < I; k++) {
            const uint32_t x = k * T + tid;
            keep[k] = false;
            cnt[k] = 0;
            w[k] = 0;
            gb[k] = 0;
            node[k] = ilo;
            if (x < n) {
                // word k * 4 + wave covers this wave's 64 positions of item k
                const uint32_t wi = k * kWaves + wave;
                const uint64_t bits = s_bits[wi];
                uint32_t pre = 0;
                for (uint32_t u = 0; u < wi; u++) pre += (uint32_t)__popcll(s_bits[u]);
                const uint32_t j = pre + mask_prefix(bits) + (uint32_t)((bits >> lane) & 1ull) - 1u;
                node[k] = ilo + j;
                const uint32_t pos = s_seg[j] + x;
                bool ok = pos < a.g.nent;  // offsets and segments always agree; never read past the array
                const uint32_t wv = ok ? a.g.nbr[pos] : 0u;
                w[k] = wv;
                const uint32_t *av = a.in_vec + node[k] * d;
                for (uint32_t kk = 0; kk < d; kk++) ok &= (av[kk] != wv);
                for (uint32_t r = 0; r < a.dp.nchk && ok; r++) {
                    const uint2 sg = a.dp.chk_dist[r] ? s_chk[r * TC + j] : seg_lookup(a.g, wv, a.dp.chk_rel[r]);
                    ok = seg_contains(a.g.nbr, sg.x, sg.y, wv);
                }
                if (KIND == EXPAND_MID) {
                    if (ok) {
                        const uint2 ns = a.need_nxt ? s_nxt[j] : seg_lookup(a.g, wv, a.dp.nxt_rel);
                        cnt[k] = ns.y - ns.x;
                        gb[k] = ns.x;
                        keep[k] = cnt[k] > 0;
                    }
                } else {
                    keep[k] = ok;
                }
            }
        }

        if (KIND == EXPAND_MID) {
            uint64_t kx[I], off[I], rk, rc;
            tile_scan<I>(keep, cnt, t, a.ws, ep, sm, kx, off, rk, rc);
            const uint32_t wd = d + 1;
#pragma unroll
            for (uint32_t k = 0; k < I; k++) {
                if (keep[k] && kx[k] < a.cap) {
                    const uint32_t *av = a.in_vec + node[k] * d;
                    uint32_t *ov = a.out_vec + kx[k] * wd;
                    for (uint32_t kk = 0; kk < d; kk++) ov[kk] = av[kk];
                    ov[d] = w[k];
                    a.out_off[kx[k]] = off[k];
                    a.out_gbeg[kx[k]] = gb[k];
                }
            }
            if (t == a.ntiles - 1 && tid == 0) {
                a.tot->keep = rk;
                a.tot->cnt = rc;
                if (rk <= a.cap) a.out_off[rk] = rc;
            }
        } else {
            if (KIND == EXPAND_LAST_ROWS) {
                uint64_t kx[I], off[I], rk, rc;
                const uint32_t zero[I] = {};
                tile_scan<I>(keep, zero, t, a.ws, ep, sm, kx, off, rk, rc);
#pragma unroll
                for (uint32_t k = 0; k < I; k++) {
                    if (keep[k] && kx[k] < a.cap) {
                        const uint32_t *av = a.in_vec + node[k] * d;
                        uint32_t *row = a.rows + kx[k] * a.q;
                        for (uint32_t kk = 0; kk < d; kk++) row[a.col[kk]] = av[kk];
                        row[a.col[d]] = w[k];
                    }
                }
                if (t == a.ntiles - 1 && tid == 0) a.tot->keep = rk;
            } else {
#pragma unroll
                for (uint32_t k = 0; k < I; k++) acc_cnt += keep[k] ? 1u : 0u;
            }
            if (a.digest) {
#pragma unroll
                for (uint32_t k = 0; k < I; k++) {
                    if (!keep[k]) continue;
                    const uint32_t *av = a.in_vec + node[k] * d;
                    uint64_t x0 = 0x9E3779B97F4A7C15ull, x1 = 0xD1B54A32D192ED03ull;
                    for (uint32_t c = 0; c < a.q; c++) {
                        const uint32_t dc = a.dcol[c];
                        const uint32_t val = (dc == d) ? w[k] : av[dc];
                        const uint64_t xv = (uint64_t)val | ((uint64_t)c << 32);
                        x0 = mix64(x0 ^ xv);
                        x1 = mix64(x1 ^ xv);
                    }
                    acc_h0 += x0;
                    acc_h1 += x1;
                }
            }
            __syncthreads();  // LDS staging is rewritten by the next tile
        }
    }
    if (KIND != EXPAND_MID) flush_sums(acc_cnt, acc_h0, acc_h1, a.tot, s_red);
    finish_launch(a.ws);
}

// ---------------------------------------------------------------------------
// chained-scan self test and workspace maintenance
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kTile) void k_scan_test(const uint32_t *vals, uint64_t n, uint64_t *out,
                                                    uint32_t *idx, ScanWS ws, Totals *tot, uint64_t ntiles)
{
    constexpr uint32_t I = kScanTestItems;
    __shared__ ScanSmem<I> sm;
    __shared__ uint64_t s_tile;
    const uint32_t ep = read_epoch(ws);
    uint32_t next = 0;
    tickets_init(ws, next);
    for (;;) {
        const uint64_t t = tickets_take(ws, next, &s_tile, ntiles);
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
        tile_scan<I>(keep, v, t, ws, ep, sm, kx, off, rk, rc);
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

size_t expand_lds_bytes(uint32_t nchk, uint32_t need_nxt, uint32_t items)
{
    const size_t TC = (size_t)kTile * items;
    size_t b = TC / 64 * 8;                 // s_bits
    b += TC * 4;                            // s_seg
    b += (size_t)nchk * TC * 8;             // s_chk
    if (need_nxt) b += TC * 8;              // s_nxt
    return b;
}

template <uint32_t KIND, uint32_t I>
static hipError_t launch_expand_t(const ExpandArgs &a, uint32_t grid, hipStream_t s)
{
    const size_t lds = expand_lds_bytes(a.dp.nchk, a.need_nxt, I);
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_expand<KIND, I>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    hipLaunchKernelGGL((k_expand<KIND, I>), dim3(grid), dim3(kTile), lds, s, a);
    return hipGetLastError();
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
    switch (a.kind) {
    case EXPAND_MID: return launch_expand_k<EXPAND_MID>(a, grid, s);
    case EXPAND_LAST_COUNT: return launch_expand_k<EXPAND_LAST_COUNT>(a, grid, s);
    default: return launch_expand_k<EXPAND_LAST_ROWS>(a, grid, s);
    }
}

hipError_t launch_root(const RootArgs &a, uint32_t grid, hipStream_t s)
{
    hipLaunchKernelGGL((k_root<kRootItems>), dim3(grid), dim3(kTile), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_scan_test(const uint32_t *vals, uint64_t n, uint64_t *out, uint32_t *idx, ScanWS ws,
                            Totals *tot, uint32_t grid, hipStream_t s)
{
    const uint64_t tc = (uint64_t)kTile * kScanTestItems;
    const uint64_t ntiles = (n + tc - 1) / tc;
    hipLaunchKernelGGL(k_scan_test, dim3(grid), dim3(kTile), 0, s, vals, n, out, idx, ws, tot, ntiles);
    return hipGetLastError();
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
