// graph.cpp -- t/v/e parsing and the relation-partitioned layout.
//
// Replaces fileParse (HnHelper.cuh:283-499).  The reference packs an 8-bit
// relation and a 24-bit row offset into one word and sorts vertices by
// degree; here every vertex keeps its original id, entries are grouped per
// (vertex, relation) into sorted segments, and a per-vertex directory finds a
// segment, so a kernel streams only the candidates with the right relation
// and answers "is w a neighbour of x under relation r" by binary search.
#include "host.h"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <memory>
#include <thread>

namespace gfsm {

static thread_local std::string t_error;
void set_error(const std::string &msg) { t_error = msg; }
const char *last_error() { return t_error.c_str(); }

uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Row hash of GFSM_DIGEST: per seed, the finalised sum over columns of
// mix64(seed ^ (row[c] | c << 32)).  A sum inside each row lets a kernel add
// a node's placed columns once for all its candidates; the final mix keeps
// rows apart (sums over rows of unmixed column terms would not).
void row_hash(const uint32_t *row, uint32_t width, uint64_t h[2])
{
    uint64_t a = 0, b = 0;
    for (uint32_t c = 0; c < width; c++) {
        const uint64_t x = (uint64_t)row[c] | ((uint64_t)c << 32);
        a += mix64(0x9E3779B97F4A7C15ull ^ x);
        b += mix64(0xD1B54A32D192ED03ull ^ x);
    }
    h[0] = mix64(a);
    h[1] = mix64(b);
}

bool relation_of(uint32_t mode, uint32_t lu, uint32_t lv, uint32_t l, uint32_t &rel)
{
    if (l == 0) return false;  // label 0 is never remapped by the reference (HnHelper.cuh:455)
    if (mode == GFSM_MODE_REF) {
        rel = l;
        return true;
    }
    if (lu >= 2048u || lv >= 2048u || l >= 1024u) return false;
    rel = (lu << 21) | (lv << 10) | l;
    return true;
}

void HostGraph::segment(uint32_t x, uint32_t rel, uint32_t &beg, uint32_t &end) const
{
    beg = end = 0;
    if (x >= V) return;
    const DirEnt *lo = dir.data() + dir_ptr[x], *hi = dir.data() + dir_ptr[x + 1];
    const DirEnt *it = std::lower_bound(lo, hi, rel, [](const DirEnt &e, uint32_t r) { return e.rel < r; });
    if (it != hi && it->rel == rel) {
        beg = it->beg;
        end = (it + 1)->beg;
    }
}

bool relation_sources(const HostGraph &g, uint32_t rel, uint64_t &sources, uint64_t &entries)
{
    const int64_t k = g.rel_index(rel);
    if (k < 0) return false;
    {
        std::lock_guard<std::mutex> lk(g.stats_mu);
        if (g.rel_entries.size() != g.rel_keys.size()) {
            std::vector<uint64_t> e(g.rel_keys.size(), 0);
            for (size_t r = 0; r < e.size(); r++)
                for (uint64_t i = g.rel_off[r]; i < g.rel_off[r + 1]; i++) e[r] += g.rel_cnt[i];
            g.rel_entries.swap(e);
        }
        entries = g.rel_entries[k];
    }
    sources = g.rel_off[k + 1] - g.rel_off[k];
    return true;
}

uint64_t relation_pool(const HostGraph &g, uint32_t rel)
{
    if (g.mode != GFSM_MODE_VLABEL) return g.V;
    const uint32_t lu = rel >> 21;
    std::lock_guard<std::mutex> lk(g.stats_mu);
    if (g.label_count.empty()) {
        g.label_count.assign(2048, 0);
        for (uint32_t x = 0; x < g.V; x++)
            if (g.vlabel[x] < 2048) g.label_count[g.vlabel[x]]++;
    }
    return lu < g.label_count.size() ? g.label_count[lu] : 0;
}

int64_t HostGraph::rel_index(uint32_t rel) const
{
    auto it = std::lower_bound(rel_keys.begin(), rel_keys.end(), rel);
    if (it == rel_keys.end() || *it != rel) return -1;
    return it - rel_keys.begin();
}

Status build_graph(uint32_t nv, const uint32_t *vlabel, uint64_t ne, const uint32_t *src,
                   const uint32_t *dst, const uint32_t *elabel, uint32_t mode, HostGraph *&out)
{
    out = nullptr;
    if (mode != GFSM_MODE_REF && mode != GFSM_MODE_VLABEL)
        return Status::err(GFSM_ERR_ARG, "unknown mode");
    if (nv == 0) return Status::err(GFSM_ERR_ARG, "graph has no vertices");
    if (ne >= 0xFFFFFFFFull)
        return Status::err(GFSM_ERR_ARG, "more than 2^32-1 entries: segment offsets are 32-bit");
    std::unique_ptr<HostGraph> g(new HostGraph);
    g->mode = mode;
    g->V = nv;
    g->vlabel.assign(nv, 0);
    if (vlabel) std::memcpy(g->vlabel.data(), vlabel, sizeof(uint32_t) * nv);
    g->src.assign(src, src + ne);
    g->dst.assign(dst, dst + ne);
    g->elabel.assign(elabel, elabel + ne);

    // Relation per entry and validation.
    std::vector<uint64_t> key(ne);
    std::atomic<int> bad{0};
    std::atomic<uint64_t> bad_at{0};
    parallel_for(ne, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            uint32_t s = src[i], d = dst[i], rel;
            if (s >= nv || d >= nv) { bad = 1; bad_at = i; continue; }
            if (!relation_of(mode, g->vlabel[s], g->vlabel[d], elabel[i], rel)) { bad = 2; bad_at = i; continue; }
            key[i] = ((uint64_t)rel << 32) | d;
        }
    });
    if (bad == 1)
        return Status::err(GFSM_ERR_ARG, "entry " + std::to_string(bad_at.load()) + " names a vertex >= V");
    if (bad == 2)
        return Status::err(GFSM_ERR_ARG, "entry " + std::to_string(bad_at.load()) +
                                             ": edge label must be >= 1 (and < 1024, vertex labels < 2048, in vlabel mode)");

    // Counting sort of entries by source.
    std::vector<std::atomic<uint32_t>> deg(nv);
    for (auto &d : deg) d.store(0, std::memory_order_relaxed);
    parallel_for(ne, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) deg[src[i]].fetch_add(1, std::memory_order_relaxed);
    });
    std::vector<uint64_t> rowp(nv + 1, 0);
    uint32_t maxdeg = 0;
    for (uint32_t v = 0; v < nv; v++) {
        uint32_t d = deg[v].load(std::memory_order_relaxed);
        rowp[v + 1] = rowp[v] + d;
        maxdeg = std::max(maxdeg, d);
    }
    if (maxdeg >= (1u << 26))
        return Status::err(GFSM_ERR_ARG, "vertex degree >= 2^26 is not supported (wave-level u32 scans)");
    g->max_degree = maxdeg;
    std::vector<uint64_t> sorted(ne);
    {
        std::vector<std::atomic<uint64_t>> fill(nv);
        for (uint32_t v = 0; v < nv; v++) fill[v].store(rowp[v], std::memory_order_relaxed);
        parallel_for(ne, [&](uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; i++) sorted[fill[src[i]].fetch_add(1, std::memory_order_relaxed)] = key[i];
        });
    }
    std::vector<uint64_t>().swap(key);

    // Sort every row by (relation, target) and count its segments.
    std::vector<uint32_t> nseg(nv, 0);
    {
        std::atomic<uint64_t> next{0};
        const uint64_t grain = 4096;
        unsigned nt = hw_threads();
        std::vector<std::thread> ts;
        for (unsigned t = 0; t < nt; t++) {
            ts.emplace_back([&] {
                for (;;) {
                    uint64_t b = next.fetch_add(grain);
                    if (b >= nv) break;
                    uint64_t e = std::min<uint64_t>(nv, b + grain);
                    for (uint64_t v = b; v < e; v++) {
                        uint64_t *r0 = sorted.data() + rowp[v], *r1 = sorted.data() + rowp[v + 1];
                        if (!std::is_sorted(r0, r1)) std::sort(r0, r1);
                        uint32_t n = 0;
                        for (uint64_t *p = r0; p < r1; p++)
                            if (p == r0 || (p[0] >> 32) != (p[-1] >> 32)) n++;
                        nseg[v] = n;
                    }
                }
            });
        }
        for (auto &t : ts) t.join();
    }
    g->dir_ptr.assign(nv + 1, 0);
    for (uint32_t v = 0; v < nv; v++) g->dir_ptr[v + 1] = g->dir_ptr[v] + nseg[v];
    uint64_t nd = g->dir_ptr[nv];
    g->dir.resize(nd + 1);
    g->nbr.resize(ne);
    parallel_for(nv, [&](uint64_t b, uint64_t e) {
        for (uint64_t v = b; v < e; v++) {
            uint32_t j = g->dir_ptr[v];
            for (uint64_t p = rowp[v]; p < rowp[v + 1]; p++) {
                g->nbr[p] = (uint32_t)sorted[p];
                if (p == rowp[v] || (sorted[p] >> 32) != (sorted[p - 1] >> 32))
                    g->dir[j++] = DirEnt{(uint32_t)(sorted[p] >> 32), (uint32_t)p};
            }
        }
    });
    g->dir[nd] = DirEnt{0xFFFFFFFFu, (uint32_t)ne};
    std::vector<uint64_t>().swap(sorted);

    // Relation index: for each relation, the vertices owning a segment of it
    // (ascending id) and the segment length.
    {
        std::vector<uint32_t> keys;
        {
            unsigned nt = nd < 65536 ? 1u : hw_threads();
            std::vector<std::vector<uint32_t>> part(nt);
            std::vector<std::thread> ts;
            for (unsigned t = 0; t < nt; t++) ts.emplace_back([&, t] {
                uint64_t b = nd * t / nt, e = nd * (t + 1) / nt;
                auto &loc = part[t];
                for (uint64_t j = b; j < e; j++) loc.push_back(g->dir[j].rel);
                std::sort(loc.begin(), loc.end());
                loc.erase(std::unique(loc.begin(), loc.end()), loc.end());
            });
            for (auto &t : ts) t.join();
            for (auto &p : part) keys.insert(keys.end(), p.begin(), p.end());
            std::sort(keys.begin(), keys.end());
            keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
        }
        size_t nk = keys.size();
        unsigned nt = hw_threads();
        if (nv < 4096) nt = 1;
        std::vector<uint64_t> cnt((size_t)nt * nk, 0);
        auto block = [&](unsigned t, uint64_t &b, uint64_t &e) { b = (uint64_t)nv * t / nt; e = (uint64_t)nv * (t + 1) / nt; };
        std::vector<uint32_t> kidx(nd);
        {
            std::vector<std::thread> ts;
            for (unsigned t = 0; t < nt; t++) ts.emplace_back([&, t] {
                uint64_t b, e;
                block(t, b, e);
                for (uint64_t v = b; v < e; v++)
                    for (uint32_t j = g->dir_ptr[v]; j < g->dir_ptr[v + 1]; j++) {
                        uint32_t k = (uint32_t)(std::lower_bound(keys.begin(), keys.end(), g->dir[j].rel) - keys.begin());
                        kidx[j] = k;
                        cnt[(size_t)k * nt + t]++;
                    }
            });
            for (auto &t : ts) t.join();
        }
        g->rel_keys = keys;
        g->rel_off.assign(nk + 1, 0);
        uint64_t acc = 0;
        for (size_t k = 0; k < nk; k++) {
            g->rel_off[k] = acc;
            for (unsigned t = 0; t < nt; t++) {
                uint64_t c = cnt[k * nt + t];
                cnt[k * nt + t] = acc;
                acc += c;
            }
        }
        g->rel_off[nk] = acc;
        g->rel_vtx.resize(acc);
        g->rel_cnt.resize(acc);
        g->rel_beg.resize(acc);
        std::vector<std::thread> ts;
        for (unsigned t = 0; t < nt; t++) ts.emplace_back([&, t] {
            uint64_t b, e;
            block(t, b, e);
            for (uint64_t v = b; v < e; v++)
                for (uint32_t j = g->dir_ptr[v]; j < g->dir_ptr[v + 1]; j++) {
                    uint64_t p = cnt[(size_t)kidx[j] * nt + t]++;
                    g->rel_vtx[p] = (uint32_t)v;
                    g->rel_cnt[p] = g->dir[j + 1].beg - g->dir[j].beg;
                    g->rel_beg[p] = g->dir[j].beg;
                }
        });
        for (auto &t : ts) t.join();
    }
    out = g.release();
    return Status::ok();
}

// ---------------------------------------------------------------------------
// t/v/e text (README.md:7-25).  The grammar is the reference's
// (HnHelper.cuh:191-266, 312-381): "t # 0\n", a header line "V E Lv Le",
// "v id label" lines, "e src dst label" lines, "t # -1".  Inside a line every
// non-digit character ends a field and '\n' ends the line.  Fewer v-lines than
// V or fewer e-lines than E is "missing data" (234); anything else malformed is
// "invalid data" (233); an unreadable file is 232.
// ---------------------------------------------------------------------------

namespace {
struct LineReader {
    const char *p, *end;
    // Parse up to n fields of the line at p (after the 2-char tag when tagged).
    bool fields(uint64_t *vals, int n, bool tagged)
    {
        for (int i = 0; i < n; i++) vals[i] = 0;
        if (tagged) p += 2;
        int cur = 0;
        while (p < end) {
            char ch = *p++;
            if (ch >= '0' && ch <= '9') {
                if (cur >= n) return false;
                vals[cur] = vals[cur] * 10 + (uint64_t)(ch - '0');
                if (vals[cur] > 0xFFFFFFFFull) return false;
            } else {
                cur++;
                if (ch == '\n') break;
            }
        }
        return true;
    }
    bool at_end_tag() const { return end - p >= 6 && std::memcmp(p, "t # -1", 6) == 0; }
};
}  // namespace

Status load_tve(const char *path, uint32_t mode, HostGraph *&out)
{
    out = nullptr;
    FILE *f = std::fopen(path, "rb");
    if (!f) return Status::err(GFSM_ERR_MISSING_FILE, std::string("missing file ") + path);
    std::fseek(f, 0, SEEK_END);
    long size = std::ftell(f);
    std::rewind(f);
    std::vector<char> buf((size_t)std::max(size, 0L));
    size_t got = size > 0 ? std::fread(buf.data(), 1, (size_t)size, f) : 0;
    std::fclose(f);
    if (got < 6 || std::memcmp(buf.data(), "t # 0\n", 6) != 0)
        return Status::err(GFSM_ERR_INVALID_DATA, "file does not start with 't # 0'");
    LineReader r{buf.data() + 6, buf.data() + got};
    uint64_t hdr[4];
    if (r.p >= r.end || !r.fields(hdr, 4, false)) return Status::err(GFSM_ERR_INVALID_DATA, "bad header line");
    uint64_t V = hdr[0], E = hdr[1];
    if (V == 0 || V > 0xFFFFFFFFull) return Status::err(GFSM_ERR_INVALID_DATA, "bad vertex count");

    std::vector<uint32_t> vlabel(V, 0), src, dst, el;
    src.reserve(E); dst.reserve(E); el.reserve(E);
    uint64_t vseen = 0;
    uint64_t vals[3];
    for (;;) {
        if (r.p >= r.end) return Status::err(GFSM_ERR_INVALID_DATA, "missing 't # -1'");
        char ch = *r.p;
        if (ch == 't') {
            if (!r.at_end_tag()) return Status::err(GFSM_ERR_INVALID_DATA, "bad 't' line");
            break;
        }
        if (ch == 'e') break;
        if (ch != 'v' || !r.fields(vals, 2, true)) return Status::err(GFSM_ERR_INVALID_DATA, "bad vertex line");
        if (vals[0] < V) vlabel[vals[0]] = (uint32_t)vals[1];
        vseen++;
    }
    if (vseen < V) return Status::err(GFSM_ERR_MISSING_DATA, "fewer vertex lines than the header states");
    for (;;) {
        if (r.p >= r.end) return Status::err(GFSM_ERR_INVALID_DATA, "missing 't # -1'");
        char ch = *r.p;
        if (ch == 't') {
            if (!r.at_end_tag()) return Status::err(GFSM_ERR_INVALID_DATA, "bad 't' line");
            break;
        }
        if (ch != 'e' || !r.fields(vals, 3, true)) return Status::err(GFSM_ERR_INVALID_DATA, "bad edge line");
        src.push_back((uint32_t)vals[0]);
        dst.push_back((uint32_t)vals[1]);
        el.push_back((uint32_t)vals[2]);
    }
    if (src.size() < E) return Status::err(GFSM_ERR_MISSING_DATA, "fewer edge lines than the header states");
    return build_graph((uint32_t)V, vlabel.data(), src.size(), src.data(), dst.data(), el.data(), mode, out);
}

Status write_tve(const HostGraph &g, const char *path)
{
    FILE *f = std::fopen(path, "wb");
    if (!f) return Status::err(GFSM_ERR_MISSING_FILE, std::string("cannot write ") + path);
    uint32_t maxv = 0, maxe = 0;
    for (uint32_t x : g.vlabel) maxv = std::max(maxv, x);
    for (uint32_t x : g.elabel) maxe = std::max(maxe, x);
    std::fprintf(f, "t # 0\n%u %zu %u %u\n", g.V, g.src.size(), maxv, maxe);
    std::vector<char> line(64);
    for (uint32_t i = 0; i < g.V; i++) std::fprintf(f, "v %u %u\n", i, g.vlabel[i]);
    for (size_t i = 0; i < g.src.size(); i++) std::fprintf(f, "e %u %u %u\n", g.src[i], g.dst[i], g.elabel[i]);
    std::fprintf(f, "t # -1");
    std::fclose(f);
    return Status::ok();
}

// Lexicographic row sort: LSD over columns with a stable counting pass per
// 16-bit digit would touch the rows many times; a comparison sort on row
// indices keeps it simple and is only used for host-side SORTED output.
void sort_rows(std::vector<uint32_t> &rows, uint32_t width)
{
    if (!width) return;
    size_t n = rows.size() / width;
    if (n > 0xFFFFFFFFull) return;  // callers cap host-side sorting far below this
    std::vector<uint32_t> idx(n);
    for (size_t i = 0; i < n; i++) idx[i] = (uint32_t)i;
    const uint32_t *R = rows.data();
    auto less = [R, width](uint32_t a, uint32_t b) {
        const uint32_t *x = R + (size_t)a * width, *y = R + (size_t)b * width;
        for (uint32_t c = 0; c < width; c++)
            if (x[c] != y[c]) return x[c] < y[c];
        return false;
    };
    unsigned nt = hw_threads();
    if (n < 100000) nt = 1;
    // sort nt blocks in parallel, then merge pairwise
    std::vector<size_t> bounds(nt + 1);
    for (unsigned t = 0; t <= nt; t++) bounds[t] = n * t / nt;
    {
        std::vector<std::thread> ts;
        for (unsigned t = 0; t < nt; t++)
            ts.emplace_back([&, t] { std::sort(idx.begin() + bounds[t], idx.begin() + bounds[t + 1], less); });
        for (auto &t : ts) t.join();
    }
    for (size_t w = 1; w < nt; w *= 2) {
        std::vector<std::thread> ts;
        for (size_t t = 0; t + w < nt; t += 2 * w) {
            size_t b = bounds[t], m = bounds[t + w], e = bounds[std::min<size_t>(t + 2 * w, nt)];
            ts.emplace_back([&, b, m, e] { std::inplace_merge(idx.begin() + b, idx.begin() + m, idx.begin() + e, less); });
        }
        for (auto &t : ts) t.join();
    }
    std::vector<uint32_t> out(rows.size());
    parallel_for(n, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++)
            std::memcpy(out.data() + i * width, R + (size_t)idx[i] * width, width * sizeof(uint32_t));
    });
    rows.swap(out);
}

}  // namespace gfsm
