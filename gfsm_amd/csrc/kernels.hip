// kernels.hip -- CDNA4 (gfx950) kernels of the forest-based matcher.
//
// One launch per (depth, candidate range).  A workgroup of 256 threads
// (4 wave64s) owns a tile of kTileCand consecutive candidates of the depth
// (kItems per thread, striped so that loads stay coalesced):
//   1. one wave finds the frontier M-nodes whose candidate ranges cover the
//      tile (a 32-ary search per tile end over their exclusive offsets),
//   2. per frontier node the workgroup stages in LDS the placed vertices (one
//      walk up the forest of parent pointers), the generator segment's base
//      and the bounds of every checked ancestor segment,
//   3. every candidate is read coalesced from its generator segment, checked
//      for injectivity against the staged ancestors and binary-searched in
//      each checked segment,
//   4. survivors are compacted with wave64 ballot/mbcnt counts and DPP prefix
//      sums of their next-depth candidate counts, chained across workgroups
//      by a decoupled look-back, so that the next depth's offsets come out of
//      the same pass.
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

struct ScanSmem {
    uint32_t wk[kItems][kWaves];
    uint64_t wc[kItems][kWaves];
    uint64_t ex[2];
};

// Tile-wide exclusive prefix of (keep, cnt) in candidate order (item k of
// thread x is tile position k * 256 + x), continued across tiles.  Returns
// each item's kept index and candidate offset and the totals through this
// tile.
DEV void tile_scan(const bool (&keep)[kItems], const uint32_t (&cnt)[kItems], uint64_t t, const ScanWS &ws,
                   uint32_t ep, ScanSmem &sm, uint64_t (&my_k)[kItems], uint64_t (&my_c)[kItems], uint64_t &run_k,
                   uint64_t &run_c)
{
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    uint32_t kpre[kItems], cpre[kItems];
#pragma unroll
    for (uint32_t k = 0; k < kItems; k++) {
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
    uint64_t ak = 0, ac = 0, offk[kItems], offc[kItems];
#pragma unroll
    for (uint32_t k = 0; k < kItems; k++) {
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
    for (uint32_t k = 0; k < kItems; k++) {
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

// Tile ticket: tiles are handed out in launch order, so a tile only waits on
// tiles already owned by running workgroups.
DEV uint64_t next_tile(const ScanWS &ws, uint64_t *s_tile)
{
    if (threadIdx.x == 0) *s_tile = atomicAdd(ws.ctl + 0, 1u);
    __syncthreads();
    return *s_tile;
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

__global__ __launch_bounds__(kTile) void k_root(RootArgs a)
{
    __shared__ ScanSmem sm;
    __shared__ uint64_t s_tile;
    const uint32_t ep = read_epoch(a.ws);
    for (;;) {
        const uint64_t t = next_tile(a.ws, &s_tile);
        if (t >= a.ntiles) break;
        bool keep[kItems];
        uint32_t v[kItems], cnt[kItems];
#pragma unroll
        for (uint32_t k = 0; k < kItems; k++) {
            const uint64_t c = t * kTileCand + k * kTile + threadIdx.x;
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
        uint64_t kk[kItems], off[kItems], rk, rc;
        tile_scan(keep, cnt, t, a.ws, ep, sm, kk, off, rk, rc);
#pragma unroll
        for (uint32_t k = 0; k < kItems; k++) {
            if (keep[k] && kk[k] < a.cap) {
                a.out_vtx[kk[k]] = v[k];
                a.out_off[kk[k]] = off[k];
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
// one depth: expand + filter + compact + next-depth offsets
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

DEV size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

template <uint32_t KIND>
__global__ __launch_bounds__(kTile) void k_expand(ExpandArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ ScanSmem sm;
    __shared__ uint64_t s_tile;
    __shared__ uint64_t s_node[2];
    __shared__ uint64_t s_red[3][kWaves];
    constexpr uint32_t T = kTile, TC = kTileCand;
    const uint32_t d = a.d;
    const uint32_t tid = threadIdx.x;

    // LDS staging, sized for TC frontier nodes (a tile never spans more).
    uint32_t *s_first = reinterpret_cast<uint32_t *>(smem);                          // [TC+1]
    uint32_t *s_seg = reinterpret_cast<uint32_t *>(smem + align16((TC + 1) * 4));    // [TC]
    uint2 *s_chk = reinterpret_cast<uint2 *>(s_seg + TC);                            // [nchk][TC]
    uint32_t *s_anc = reinterpret_cast<uint32_t *>(s_chk + a.dp.nchk * TC);          // [d][TC]
    uint32_t *s_nxt = s_anc + d * TC;                                                // [TC]

    const uint32_t ep = read_epoch(a.ws);
    uint64_t acc_cnt = 0, acc_h0 = 0, acc_h1 = 0;
    for (;;) {
        const uint64_t t = next_tile(a.ws, &s_tile);
        if (t >= a.ntiles) break;
        const uint64_t c0 = a.g_begin + t * TC;
        const uint64_t c1 = (c0 + TC < a.g_end) ? c0 + TC : a.g_end;
        const uint32_t n = (uint32_t)(c1 - c0);

        // Frontier nodes covering [c0, c1): offsets are strictly increasing
        // (every kept node has at least one candidate).
        if (tid < 64) {
            const uint32_t sub = tid & 31;
            const uint64_t key = (tid < 32) ? c0 : c1 - 1;
            const uint64_t i = frontier_search(a.off_prev, a.S_prev, key, sub);
            if (sub == 0) s_node[tid >> 5] = i;
        }
        __syncthreads();
        const uint64_t ilo = s_node[0];
        const uint32_t nf = (uint32_t)(s_node[1] - ilo + 1);

        // Per frontier node: placed vertices, generator base, checked segments.
        for (uint32_t j = tid; j < nf; j += T) {
            const uint64_t i = ilo + j;
            const uint64_t o = a.off_prev[i];
            s_first[j] = (o > c0) ? (uint32_t)(o - c0) : 0u;
            uint32_t v = a.lv.vtx[d - 1][i];
            s_anc[(d - 1) * TC + j] = v;
            uint32_t p = (d >= 2) ? a.lv.par[d - 1][i] : 0u;
            for (int k = (int)d - 2; k >= 0; k--) {
                v = a.lv.vtx[k][p];
                s_anc[k * TC + j] = v;
                if (k > 0) p = a.lv.par[k][p];
            }
            const uint32_t gp = s_anc[(d - a.dp.gen_dist) * TC + j];
            const uint2 sg = seg_lookup(a.g, gp, a.dp.gen_rel);
            // nbr index of tile position x inside this node's range: base + x
            s_seg[j] = sg.x + (uint32_t)(c0 - o);
            for (uint32_t r = 0; r < a.dp.nchk; r++) {
                const uint32_t dist = a.dp.chk_dist[r];
                if (dist) s_chk[r * TC + j] = seg_lookup(a.g, s_anc[(d - dist) * TC + j], a.dp.chk_rel[r]);
            }
            if (KIND == EXPAND_MID && a.dp.nxt_dist > 1) {
                const uint2 ns = seg_lookup(a.g, s_anc[(d + 1 - a.dp.nxt_dist) * TC + j], a.dp.nxt_rel);
                s_nxt[j] = ns.y - ns.x;
            }
        }
        if (tid == 0) s_first[nf] = n;
        __syncthreads();

        bool keep[kItems];
        uint32_t cnt[kItems], w[kItems], jj[kItems];
#pragma unroll
        for (uint32_t k = 0; k < kItems; k++) {
            const uint32_t x = k * T + tid;
            keep[k] = false;
            cnt[k] = 0;
            w[k] = 0;
            jj[k] = 0;
            if (x < n) {
                uint32_t lo = 0, len = nf;
                while (len > 1) {
                    const uint32_t h = len >> 1;
                    if (s_first[lo + h] <= x) {
                        lo += h;
                        len -= h;
                    } else {
                        len = h;
                    }
                }
                const uint32_t j = lo;
                jj[k] = j;
                const uint32_t pos = s_seg[j] + x;
                bool ok = pos < a.g.nent;  // offsets and segments always agree; never read past the array
                const uint32_t wv = ok ? a.g.nbr[pos] : 0u;
                w[k] = wv;
                for (uint32_t kk = 0; kk < d; kk++) ok &= (s_anc[kk * TC + j] != wv);
                for (uint32_t r = 0; r < a.dp.nchk && ok; r++) {
                    const uint2 sg = a.dp.chk_dist[r] ? s_chk[r * TC + j] : seg_lookup(a.g, wv, a.dp.chk_rel[r]);
                    ok = seg_contains(a.g.nbr, sg.x, sg.y, wv);
                }
                if (KIND == EXPAND_MID) {
                    if (ok) {
                        uint32_t c;
                        if (a.dp.nxt_dist == 1) {
                            const uint2 ns = seg_lookup(a.g, wv, a.dp.nxt_rel);
                            c = ns.y - ns.x;
                        } else {
                            c = s_nxt[j];
                        }
                        cnt[k] = c;
                        keep[k] = c > 0;
                    }
                } else {
                    keep[k] = ok;
                }
            }
        }

        if (KIND == EXPAND_MID) {
            uint64_t kx[kItems], off[kItems], rk, rc;
            tile_scan(keep, cnt, t, a.ws, ep, sm, kx, off, rk, rc);
#pragma unroll
            for (uint32_t k = 0; k < kItems; k++) {
                if (keep[k] && kx[k] < a.cap) {
                    a.out_vtx[kx[k]] = w[k];
                    a.out_par[kx[k]] = (uint32_t)(ilo + jj[k]);
                    a.out_off[kx[k]] = off[k];
                }
            }
            if (t == a.ntiles - 1 && tid == 0) {
                a.tot->keep = rk;
                a.tot->cnt = rc;
                if (rk <= a.cap) a.out_off[rk] = rc;
            }
        } else {
            if (KIND == EXPAND_LAST_ROWS) {
                uint64_t kx[kItems], off[kItems], rk, rc;
                const uint32_t zero[kItems] = {};
                tile_scan(keep, zero, t, a.ws, ep, sm, kx, off, rk, rc);
#pragma unroll
                for (uint32_t k = 0; k < kItems; k++) {
                    if (keep[k] && kx[k] < a.cap) {
                        uint32_t *row = a.rows + kx[k] * a.q;
                        for (uint32_t kk = 0; kk < d; kk++) row[a.col[kk]] = s_anc[kk * TC + jj[k]];
                        row[a.col[d]] = w[k];
                    }
                }
                if (t == a.ntiles - 1 && tid == 0) a.tot->keep = rk;
            } else {
#pragma unroll
                for (uint32_t k = 0; k < kItems; k++) acc_cnt += keep[k] ? 1u : 0u;
            }
            if (a.digest) {
#pragma unroll
                for (uint32_t k = 0; k < kItems; k++) {
                    if (!keep[k]) continue;
                    uint64_t x0 = 0x9E3779B97F4A7C15ull, x1 = 0xD1B54A32D192ED03ull;
                    for (uint32_t c = 0; c < a.q; c++) {
                        const uint32_t dc = a.dcol[c];
                        const uint32_t val = (dc == d) ? w[k] : s_anc[dc * TC + jj[k]];
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
    __shared__ ScanSmem sm;
    __shared__ uint64_t s_tile;
    const uint32_t ep = read_epoch(ws);
    for (;;) {
        const uint64_t t = next_tile(ws, &s_tile);
        if (t >= ntiles) break;
        bool keep[kItems];
        uint32_t v[kItems];
#pragma unroll
        for (uint32_t k = 0; k < kItems; k++) {
            const uint64_t i = t * kTileCand + k * kTile + threadIdx.x;
            v[k] = i < n ? vals[i] : 0u;
            keep[k] = v[k] != 0;
        }
        uint64_t kx[kItems], off[kItems], rk, rc;
        tile_scan(keep, v, t, ws, ep, sm, kx, off, rk, rc);
#pragma unroll
        for (uint32_t k = 0; k < kItems; k++) {
            const uint64_t i = t * kTileCand + k * kTile + threadIdx.x;
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

size_t expand_lds_bytes(uint32_t d, uint32_t nchk)
{
    const size_t TC = kTileCand;
    size_t b = ((TC + 1) * 4 + 15) & ~size_t(15);  // s_first
    b += TC * 4;                                   // s_seg
    b += (size_t)nchk * TC * 8;                    // s_chk
    b += (size_t)d * TC * 4;                       // s_anc
    b += TC * 4;                                   // s_nxt
    return b;
}

template <uint32_t KIND>
static hipError_t launch_expand_t(const ExpandArgs &a, uint32_t grid, hipStream_t s)
{
    const size_t lds = expand_lds_bytes(a.d, a.dp.nchk);
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_expand<KIND>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    hipLaunchKernelGGL(k_expand<KIND>, dim3(grid), dim3(kTile), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_expand(const ExpandArgs &a, uint32_t grid, hipStream_t s)
{
    switch (a.kind) {
    case EXPAND_MID: return launch_expand_t<EXPAND_MID>(a, grid, s);
    case EXPAND_LAST_COUNT: return launch_expand_t<EXPAND_LAST_COUNT>(a, grid, s);
    default: return launch_expand_t<EXPAND_LAST_ROWS>(a, grid, s);
    }
}

hipError_t launch_root(const RootArgs &a, uint32_t grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_root, dim3(grid), dim3(kTile), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_scan_test(const uint32_t *vals, uint64_t n, uint64_t *out, uint32_t *idx, ScanWS ws,
                            Totals *tot, uint32_t grid, hipStream_t s)
{
    const uint64_t ntiles = (n + kTileCand - 1) / kTileCand;
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
