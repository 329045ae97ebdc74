// kernels.hip -- CDNA4 (gfx950) kernels of the forest-based matcher.
//
// One launch per (depth, candidate range).  A workgroup of 256 threads
// (4 wave64s) owns a tile of 256 consecutive candidates of the depth:
//   1. it finds the frontier M-nodes whose candidate ranges cover the tile
//      (binary search over their exclusive offsets),
//   2. stages, per frontier node, the placed vertices (one walk up the forest
//      of parent pointers), the generator's segment base and every checked
//      ancestor's segment bounds in LDS,
//   3. gives each thread one candidate: a coalesced read of the generator
//      segment, injectivity against the staged ancestors, and a binary search
//      of each checked segment,
//   4. compacts the survivors with a wave64 ballot/mbcnt count and a DPP
//      prefix sum of their next-depth candidate counts, chained across
//      workgroups by a decoupled look-back, so that the next depth's offsets
//      come out of the same pass (no separate scan kernels).
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
// chained scan across workgroups (decoupled look-back)
// ---------------------------------------------------------------------------

DEV uint32_t ld_acq(const uint32_t *p) { return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT); }
DEV void st_rel(uint32_t *p, uint32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT); }
DEV uint64_t ld_rlx(const uint64_t *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
DEV void st_rlx(uint64_t *p, uint64_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Exclusive (kept, candidates) prefix of tile t, computed by all 64 lanes of
// one wave: each pass reads the status of the 64 preceding tiles, waits until
// every tile up to the nearest inclusive prefix has published, and sums.
DEV void lookback(const ScanWS &ws, uint32_t tag, uint64_t t, uint64_t &ek, uint64_t &ec)
{
    const uint32_t lane = lane_id();
    uint64_t sk = 0, sc = 0;
    int64_t base = (int64_t)t - 1;
    for (;;) {
        const int64_t idx = base - (int64_t)lane;
        uint32_t st;
        uint64_t pm;
        for (;;) {
            st = 2u;  // before tile 0 the inclusive prefix is 0
            if (idx >= 0) {
                uint32_t f = ld_acq(ws.flag + idx);
                st = ((f & ~3u) == tag) ? (f & 3u) : 0u;
            }
            pm = __ballot(st == 2u);
            uint64_t xm = __ballot(st == 0u);
            uint64_t upto = pm ? (pm ^ (pm - 1)) : ~0ull;
            if (!(xm & upto)) break;
            __builtin_amdgcn_s_sleep(1);
        }
        uint64_t upto = pm ? (pm ^ (pm - 1)) : ~0ull;
        uint64_t vk = 0, vc = 0;
        if (((upto >> lane) & 1ull) && idx >= 0) {
            const Pair64 *src = (st == 2u) ? ws.inc : ws.agg;
            vk = ld_rlx(&src[idx].k);
            vc = ld_rlx(&src[idx].c);
        }
        sk += wave_sum64(vk);
        sc += wave_sum64(vc);
        if (pm) break;
        base -= 64;
    }
    ek = sk;
    ec = sc;
}

struct ScanSmem {
    uint32_t wk[kWaves];
    uint64_t wc[kWaves];
    uint64_t ex[2];
};

// Tile-wide exclusive prefix of (keep, cnt) continued across tiles.
// Returns this thread's kept index and candidate offset, and the running
// totals through this tile.
DEV void tile_scan(bool keep, uint32_t cnt, uint64_t t, const ScanWS &ws, uint32_t tag, ScanSmem &sm,
                   uint64_t &my_k, uint64_t &my_c, uint64_t &run_k, uint64_t &run_c)
{
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    const uint64_t m = __ballot(keep);
    const uint32_t kpre = mask_prefix(m);
    const uint32_t cinc = wave_incl_scan(cnt);
    if (lane == 63) {
        sm.wk[wave] = (uint32_t)__popcll(m);
        sm.wc[wave] = cinc;
    }
    __syncthreads();
    uint64_t wk_off = 0, wc_off = 0, tk = 0, tc = 0;
#pragma unroll
    for (uint32_t w = 0; w < kWaves; w++) {
        if (w < wave) {
            wk_off += sm.wk[w];
            wc_off += sm.wc[w];
        }
        tk += sm.wk[w];
        tc += sm.wc[w];
    }
    if (wave == 0) {
        uint64_t ek = 0, ec = 0;
        if (t != 0) {
            if (lane == 0) {
                st_rlx(&ws.agg[t].k, tk);
                st_rlx(&ws.agg[t].c, tc);
                st_rel(ws.flag + t, tag | 1u);
            }
            lookback(ws, tag, t, ek, ec);
        }
        if (lane == 0) {
            st_rlx(&ws.inc[t].k, ek + tk);
            st_rlx(&ws.inc[t].c, ec + tc);
            st_rel(ws.flag + t, tag | 2u);
            sm.ex[0] = ek;
            sm.ex[1] = ec;
        }
    }
    __syncthreads();
    my_k = sm.ex[0] + wk_off + kpre;
    my_c = sm.ex[1] + wc_off + (cinc - cnt);
    run_k = sm.ex[0] + tk;
    run_c = sm.ex[1] + tc;
}

DEV uint32_t read_tag(const ScanWS &ws)
{
    uint32_t e = __hip_atomic_load(ws.ctl + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return (e & 0x3FFFFFFFu) << 2;
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
            __hip_atomic_fetch_add(ws.ctl + 2, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ---------------------------------------------------------------------------
// root level (hnInit, HnGPU.cu:76-109)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kTile) void k_root(RootArgs a)
{
    __shared__ ScanSmem sm;
    __shared__ uint64_t s_tile;
    const uint32_t tag = read_tag(a.ws);
    for (;;) {
        const uint64_t t = next_tile(a.ws, &s_tile);
        if (t >= a.ntiles) break;
        const uint64_t c = t * kTile + threadIdx.x;
        bool keep = false;
        uint32_t v = 0, cnt = 0;
        if (c < a.n) {
            v = a.cand_vtx[c];
            cnt = a.cand_cnt[c];
            keep = cnt > 0;
            for (uint32_t r = 0; r < a.nchk && keep; r++) {
                uint2 sg = seg_lookup(a.g, v, a.chk_rel[r]);
                keep = seg_contains(a.g.nbr, sg.x, sg.y, v);
            }
            if (!keep) cnt = 0;
        }
        uint64_t k, off, rk, rc;
        tile_scan(keep, cnt, t, a.ws, tag, sm, k, off, rk, rc);
        if (keep && k < a.cap) {
            a.out_vtx[k] = v;
            a.out_par[k] = 0u;
            a.out_off[k] = off;
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

DEV size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

template <uint32_t KIND>
__global__ __launch_bounds__(kTile) void k_expand(ExpandArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ ScanSmem sm;
    __shared__ uint64_t s_tile;
    __shared__ uint64_t s_node[2];
    constexpr uint32_t T = kTile;
    const uint32_t d = a.d;
    const uint32_t tid = threadIdx.x;

    uint32_t *s_first = reinterpret_cast<uint32_t *>(smem);                     // [T+1]
    uint64_t *s_seg = reinterpret_cast<uint64_t *>(smem + align16((T + 1) * 4)); // [T]
    uint2 *s_chk = reinterpret_cast<uint2 *>(s_seg + T);                        // [nchk][T]
    uint32_t *s_anc = reinterpret_cast<uint32_t *>(s_chk + a.dp.nchk * T);      // [d][T]
    uint32_t *s_nxt = s_anc + d * T;                                            // [T]

    const uint32_t tag = read_tag(a.ws);
    for (;;) {
        const uint64_t t = next_tile(a.ws, &s_tile);
        if (t >= a.ntiles) break;
        const uint64_t c0 = a.g_begin + t * T;
        const uint64_t c1 = (c0 + T < a.g_end) ? c0 + T : a.g_end;
        const uint32_t n = (uint32_t)(c1 - c0);

        // Frontier nodes covering [c0, c1): last node starting at or before
        // each end (offsets are strictly increasing: every kept node has
        // at least one candidate).
        if (tid < 2) {
            const uint64_t key = tid ? c1 - 1 : c0;
            uint64_t lo = 0, len = a.S_prev;
            while (len > 1) {
                uint64_t h = len >> 1;
                if (a.off_prev[lo + h] <= key) {
                    lo += h;
                    len -= h;
                } else {
                    len = h;
                }
            }
            s_node[tid] = lo;
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
            s_anc[(d - 1) * T + j] = v;
            uint32_t p = (d >= 2) ? a.lv.par[d - 1][i] : 0u;
            for (int k = (int)d - 2; k >= 0; k--) {
                v = a.lv.vtx[k][p];
                s_anc[k * T + j] = v;
                if (k > 0) p = a.lv.par[k][p];
            }
            const uint32_t gp = s_anc[(d - a.dp.gen_dist) * T + j];
            const uint2 sg = seg_lookup(a.g, gp, a.dp.gen_rel);
            s_seg[j] = (uint64_t)sg.x - o;
            for (uint32_t r = 0; r < a.dp.nchk; r++) {
                const uint32_t dist = a.dp.chk_dist[r];
                if (dist) s_chk[r * T + j] = seg_lookup(a.g, s_anc[(d - dist) * T + j], a.dp.chk_rel[r]);
            }
            if (KIND == EXPAND_MID && a.dp.nxt_dist > 1) {
                const uint2 ns = seg_lookup(a.g, s_anc[(d + 1 - a.dp.nxt_dist) * T + j], a.dp.nxt_rel);
                s_nxt[j] = ns.y - ns.x;
            }
        }
        if (tid == 0) s_first[nf] = n;
        __syncthreads();

        bool keep = false;
        uint32_t cnt = 0, w = 0, j = 0;
        if (tid < n) {
            uint32_t lo = 0, len = nf;
            while (len > 1) {
                uint32_t h = len >> 1;
                if (s_first[lo + h] <= tid) {
                    lo += h;
                    len -= h;
                } else {
                    len = h;
                }
            }
            j = lo;
            const uint64_t pos = s_seg[j] + c0 + tid;
            bool ok = pos < a.g.nent;  // offsets and segments always agree; never read past the array
            w = ok ? a.g.nbr[pos] : 0u;
            for (uint32_t k = 0; k < d; k++) ok &= (s_anc[k * T + j] != w);
            for (uint32_t r = 0; r < a.dp.nchk && ok; r++) {
                const uint2 sg = a.dp.chk_dist[r] ? s_chk[r * T + j] : seg_lookup(a.g, w, a.dp.chk_rel[r]);
                ok = seg_contains(a.g.nbr, sg.x, sg.y, w);
            }
            if (KIND == EXPAND_MID) {
                if (ok) {
                    if (a.dp.nxt_dist == 1) {
                        const uint2 ns = seg_lookup(a.g, w, a.dp.nxt_rel);
                        cnt = ns.y - ns.x;
                    } else {
                        cnt = s_nxt[j];
                    }
                    keep = cnt > 0;
                }
            } else {
                keep = ok;
            }
        }

        if (KIND == EXPAND_MID) {
            uint64_t k, off, rk, rc;
            tile_scan(keep, keep ? cnt : 0u, t, a.ws, tag, sm, k, off, rk, rc);
            if (keep && k < a.cap) {
                a.out_vtx[k] = w;
                a.out_par[k] = (uint32_t)(ilo + j);
                a.out_off[k] = off;
            }
            if (t == a.ntiles - 1 && tid == 0) {
                a.tot->keep = rk;
                a.tot->cnt = rc;
                if (rk <= a.cap) a.out_off[rk] = rc;
            }
        } else {
            if (KIND == EXPAND_LAST_ROWS) {
                uint64_t k, off, rk, rc;
                tile_scan(keep, 0u, t, a.ws, tag, sm, k, off, rk, rc);
                if (keep && k < a.cap) {
                    uint32_t *row = a.rows + k * a.q;
                    for (uint32_t kk = 0; kk < d; kk++) row[a.col[kk]] = s_anc[kk * T + j];
                    row[a.col[d]] = w;
                }
                if (t == a.ntiles - 1 && tid == 0) a.tot->keep = rk;
            } else {
                const uint64_t m = __ballot(keep);
                if (lane_id() == 0 && m) atomicAdd((unsigned long long *)&a.tot->count, (unsigned long long)__popcll(m));
            }
            if (a.digest) {
                uint64_t h0 = 0, h1 = 0;
                if (keep) {
                    uint64_t x0 = 0x9E3779B97F4A7C15ull, x1 = 0xD1B54A32D192ED03ull;
                    for (uint32_t c = 0; c < a.q; c++) {
                        const uint32_t dc = a.dcol[c];
                        const uint32_t val = (dc == d) ? w : s_anc[dc * T + j];
                        const uint64_t x = (uint64_t)val | ((uint64_t)c << 32);
                        x0 = mix64(x0 ^ x);
                        x1 = mix64(x1 ^ x);
                    }
                    h0 = x0;
                    h1 = x1;
                }
                h0 = wave_sum64(h0);
                h1 = wave_sum64(h1);
                if (lane_id() == 0) {
                    atomicAdd((unsigned long long *)&a.tot->dg0, (unsigned long long)h0);
                    atomicAdd((unsigned long long *)&a.tot->dg1, (unsigned long long)h1);
                }
            }
            __syncthreads();  // LDS staging is rewritten by the next tile
        }
    }
    finish_launch(a.ws);
}

// ---------------------------------------------------------------------------
// chained-scan self test and workspace reset
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kTile) void k_scan_test(const uint32_t *vals, uint64_t n, uint64_t *out,
                                                    uint32_t *idx, ScanWS ws, Totals *tot, uint64_t ntiles)
{
    __shared__ ScanSmem sm;
    __shared__ uint64_t s_tile;
    const uint32_t tag = read_tag(ws);
    for (;;) {
        const uint64_t t = next_tile(ws, &s_tile);
        if (t >= ntiles) break;
        const uint64_t i = t * kTile + threadIdx.x;
        const uint32_t v = i < n ? vals[i] : 0u;
        uint64_t k, off, rk, rc;
        tile_scan(v != 0, v, t, ws, tag, sm, k, off, rk, rc);
        if (i < n) out[i] = off;
        if (v) idx[k] = (uint32_t)i;
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

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------

size_t expand_lds_bytes(uint32_t d, uint32_t nchk)
{
    const size_t T = kTile;
    size_t b = ((T + 1) * 4 + 15) & ~size_t(15);
    b += T * 8;              // s_seg
    b += (size_t)nchk * T * 8; // s_chk
    b += (size_t)d * T * 4;  // s_anc
    b += T * 4;              // s_nxt
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
    const uint64_t ntiles = (n + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_scan_test, dim3(grid), dim3(kTile), 0, s, vals, n, out, idx, ws, tot, ntiles);
    return hipGetLastError();
}

hipError_t launch_ws_reset(ScanWS ws, Totals *tot, hipStream_t s)
{
    hipLaunchKernelGGL(k_ws_reset, dim3(1), dim3(64), 0, s, ws, tot);
    return hipGetLastError();
}

}  // namespace gfsm
