// gen.cpp -- seeded workload generators: Erdos-Renyi, R-MAT and damaged
// grids (undirected, simple, written as symmetric entry lists), and query
// extraction by random walk.  Every random draw is a pure function of
// (seed, index), so a graph is identical however many threads build it.
#include "host.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <memory>
#include <set>
#include <thread>
#include <unordered_set>

namespace gfsm {

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        s += 0x9E3779B97F4A7C15ull;
        return mix64(s);
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
};

inline double hash_uni(uint64_t seed, uint64_t i)
{
    return (double)(mix64(seed ^ mix64(i + 0x632BE59BD9B4E019ull)) >> 11) * (1.0 / 9007199254740992.0);
}

// P(label = k) proportional to k^-s, k = 1..n.
struct Zipf {
    std::vector<double> cdf;
    Zipf(uint32_t n, double s)
    {
        cdf.resize(n);
        double acc = 0;
        for (uint32_t k = 1; k <= n; k++) {
            acc += std::pow((double)k, -s);
            cdf[k - 1] = acc;
        }
        for (auto &c : cdf) c /= acc;
    }
    uint32_t pick(double u) const
    {
        size_t k = std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin();
        if (k >= cdf.size()) k = cdf.size() - 1;
        return 1u + (uint32_t)k;
    }
};

inline uint32_t edge_label(uint64_t seed, uint32_t u, uint32_t v, uint32_t nel)
{
    if (nel <= 1) return 1;
    uint64_t a = std::min(u, v), b = std::max(u, v);
    return 1u + (uint32_t)(mix64(seed ^ 0xE7037ED1A0B428DBull ^ ((a << 32) | b)) % nel);
}

// Undirected simple graph from a list of (u, v) pairs (any order, duplicates
// and self loops allowed): symmetrise, deduplicate, label, build.
Status finish_undirected(uint32_t n, std::vector<uint64_t> &pairs, const std::vector<uint32_t> &vlabel,
                         uint32_t nel, uint64_t seed, uint32_t mode, HostGraph *&out)
{
    // counting sort by source of both directions
    std::vector<std::atomic<uint32_t>> deg(n);
    for (auto &d : deg) d.store(0, std::memory_order_relaxed);
    parallel_for(pairs.size(), [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            uint32_t u = (uint32_t)(pairs[i] >> 32), v = (uint32_t)pairs[i];
            if (u == v) continue;
            deg[u].fetch_add(1, std::memory_order_relaxed);
            deg[v].fetch_add(1, std::memory_order_relaxed);
        }
    });
    std::vector<uint64_t> rowp(n + 1, 0);
    for (uint32_t v = 0; v < n; v++) rowp[v + 1] = rowp[v] + deg[v].load(std::memory_order_relaxed);
    std::vector<uint32_t> adj(rowp[n]);
    {
        std::vector<std::atomic<uint64_t>> fill(n);
        for (uint32_t v = 0; v < n; v++) fill[v].store(rowp[v], std::memory_order_relaxed);
        parallel_for(pairs.size(), [&](uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; i++) {
                uint32_t u = (uint32_t)(pairs[i] >> 32), v = (uint32_t)pairs[i];
                if (u == v) continue;
                adj[fill[u].fetch_add(1, std::memory_order_relaxed)] = v;
                adj[fill[v].fetch_add(1, std::memory_order_relaxed)] = u;
            }
        });
    }
    std::vector<uint64_t>().swap(pairs);
    std::vector<uint32_t> ndeg(n, 0);
    parallel_for(n, [&](uint64_t b, uint64_t e) {
        for (uint64_t v = b; v < e; v++) {
            uint32_t *r0 = adj.data() + rowp[v], *r1 = adj.data() + rowp[v + 1];
            std::sort(r0, r1);
            ndeg[v] = (uint32_t)(std::unique(r0, r1) - r0);
        }
    });
    std::vector<uint64_t> outp(n + 1, 0);
    for (uint32_t v = 0; v < n; v++) outp[v + 1] = outp[v] + ndeg[v];
    std::vector<uint32_t> src(outp[n]), dst(outp[n]), el(outp[n]);
    parallel_for(n, [&](uint64_t b, uint64_t e) {
        for (uint64_t v = b; v < e; v++)
            for (uint32_t k = 0; k < ndeg[v]; k++) {
                uint64_t p = outp[v] + k;
                src[p] = (uint32_t)v;
                dst[p] = adj[rowp[v] + k];
                el[p] = edge_label(seed, (uint32_t)v, dst[p], nel);
            }
    });
    std::vector<uint32_t>().swap(adj);
    Status st = build_graph(n, vlabel.data(), src.size(), src.data(), dst.data(), el.data(), mode, out);
    return st;
}

std::vector<uint32_t> uniform_labels(uint32_t n, uint32_t nvl, uint64_t seed)
{
    std::vector<uint32_t> l(n, 1);
    if (nvl > 1)
        parallel_for(n, [&](uint64_t b, uint64_t e) {
            for (uint64_t v = b; v < e; v++) l[v] = 1u + (uint32_t)(mix64(seed ^ 0x5851F42D4C957F2Dull ^ v) % nvl);
        });
    return l;
}

std::vector<uint32_t> zipf_labels(uint32_t n, uint32_t nvl, double s, uint64_t seed)
{
    std::vector<uint32_t> l(n, 1);
    if (nvl > 1) {
        Zipf z(nvl, s);
        parallel_for(n, [&](uint64_t b, uint64_t e) {
            for (uint64_t v = b; v < e; v++) l[v] = z.pick(hash_uni(seed ^ 0x2545F4914F6CDD1Dull, v));
        });
    }
    return l;
}

}  // namespace

Status gen_er(uint32_t n, double avg_degree, uint32_t nvl, uint32_t nel, uint64_t seed, uint32_t mode,
              HostGraph *&out)
{
    if (n < 2 || avg_degree < 0) return Status::err(GFSM_ERR_ARG, "bad ER parameters");
    const double p = std::min(1.0, avg_degree / (double)(n - 1));
    std::vector<uint64_t> pairs;
    // Geometric skipping over the n(n-1)/2 pairs (Batagelj & Brandes 2005).
    Rng r(seed);
    if (p > 0) {
        const double lq = std::log(1.0 - p);
        int64_t v = 1, w = -1;
        while (v < (int64_t)n) {
            double u = r.uni();
            w += 1 + (p >= 1.0 ? 0 : (int64_t)std::floor(std::log(1.0 - u) / lq));
            while (w >= v && v < (int64_t)n) {
                w -= v;
                v++;
            }
            if (v < (int64_t)n) pairs.push_back(((uint64_t)v << 32) | (uint64_t)w);
        }
    }
    return finish_undirected(n, pairs, uniform_labels(n, nvl, seed), nel, seed, mode, out);
}

Status gen_rmat(uint32_t scale, uint32_t ef, double a, double b, double c, uint32_t nvl, double zs, uint32_t nel,
                uint64_t seed, uint32_t mode, HostGraph *&out)
{
    if (scale < 1 || scale > 31 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return Status::err(GFSM_ERR_ARG, "bad R-MAT parameters");
    const uint32_t n = 1u << scale;
    const uint64_t m = (uint64_t)n * ef;
    std::vector<uint64_t> pairs(m);
    const double ab = a + b, abc = a + b + c;
    parallel_for(m, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; i++) {
            Rng r(mix64(seed ^ (i * 0xD6E8FEB86659FD93ull)));
            uint32_t u = 0, v = 0;
            for (uint32_t l = 0; l < scale; l++) {
                double x = r.uni();
                uint32_t bu = x >= ab, bv = (x >= a && x < ab) || x >= abc;
                u = (u << 1) | bu;
                v = (v << 1) | bv;
            }
            pairs[i] = ((uint64_t)u << 32) | v;
        }
    });
    return finish_undirected(n, pairs, zipf_labels(n, nvl, zs, seed), nel, seed, mode, out);
}

Status gen_grid(uint32_t rows, uint32_t cols, double del_frac, double diag_frac, uint64_t seed, uint32_t mode,
                HostGraph *&out)
{
    if (rows < 2 || cols < 2 || del_frac < 0 || del_frac > 1 || diag_frac < 0)
        return Status::err(GFSM_ERR_ARG, "bad grid parameters");
    const uint64_t n64 = (uint64_t)rows * cols;
    if (n64 >= 0xFFFFFFFFull) return Status::err(GFSM_ERR_ARG, "grid too large");
    const uint32_t n = (uint32_t)n64;
    std::vector<uint64_t> edges;
    edges.reserve(2 * n64);
    for (uint32_t r = 0; r < rows; r++)
        for (uint32_t c = 0; c < cols; c++) {
            uint32_t id = r * cols + c;
            if (c + 1 < cols) edges.push_back(((uint64_t)id << 32) | (id + 1));
            if (r + 1 < rows) edges.push_back(((uint64_t)id << 32) | (id + cols));
        }
    const uint64_t e0 = edges.size();
    const uint64_t ndel = (uint64_t)std::llround(del_frac * (double)e0);
    Rng r(seed);
    // partial Fisher-Yates: the first ndel positions become the deleted set
    for (uint64_t i = 0; i < ndel; i++) std::swap(edges[i], edges[i + r.below(e0 - i)]);
    std::vector<uint64_t> keep(edges.begin() + ndel, edges.end());
    const uint64_t ndiag = (uint64_t)std::llround(diag_frac * (double)e0);
    std::unordered_set<uint64_t> diag;
    diag.reserve(ndiag * 2);
    while (diag.size() < ndiag) {
        uint32_t rr = (uint32_t)r.below(rows - 1), cc = (uint32_t)r.below(cols);
        bool down_right = r.next() & 1;
        if (down_right && cc + 1 >= cols) continue;
        if (!down_right && cc == 0) continue;
        uint32_t u = rr * cols + cc, v = (rr + 1) * cols + (down_right ? cc + 1 : cc - 1);
        uint64_t key = ((uint64_t)std::min(u, v) << 32) | std::max(u, v);
        if (diag.insert(key).second) keep.push_back(key);
    }
    std::vector<uint32_t> vl(n, 1);
    return finish_undirected(n, keep, vl, 1, seed, mode, out);
}

// Edge label of the data entry x -> y (0 when absent).
static uint32_t entry_label(const HostGraph &g, uint32_t x, uint32_t y)
{
    for (uint32_t j = g.dir_ptr[x]; j < g.dir_ptr[x + 1]; j++) {
        const uint32_t *b = g.nbr.data() + g.dir[j].beg, *e = g.nbr.data() + g.dir[j + 1].beg;
        if (std::binary_search(b, e, y)) return g.mode == GFSM_MODE_REF ? g.dir[j].rel : (g.dir[j].rel & 1023u);
    }
    return 0;
}

Status gen_query_walk(const HostGraph &data, uint32_t qsize, uint32_t kind, uint64_t seed, HostGraph *&out)
{
    if (qsize < 1 || qsize > GFSM_MAXQ) return Status::err(GFSM_ERR_ARG, "query size must be 1..32");
    Rng r(seed);
    std::vector<uint32_t> chosen;
    std::vector<std::pair<uint32_t, uint32_t>> tree;
    for (int attempt = 0; attempt < 10000; attempt++) {
        uint32_t start = (uint32_t)r.below(data.V);
        if (data.degree(start) == 0) continue;
        chosen.assign(1, start);
        tree.clear();
        uint32_t cur = start, stall = 0;
        while (chosen.size() < qsize && stall <= 100) {
            const uint32_t b = data.dir[data.dir_ptr[cur]].beg, e = data.dir[data.dir_ptr[cur + 1]].beg;
            const uint32_t nb = data.nbr[b + (uint32_t)r.below(e - b)];
            auto it = std::find(chosen.begin(), chosen.end(), nb);
            if (it == chosen.end()) {
                tree.push_back({(uint32_t)(std::find(chosen.begin(), chosen.end(), cur) - chosen.begin()),
                                (uint32_t)chosen.size()});
                chosen.push_back(nb);
                stall = 0;
            } else {
                stall++;
            }
            cur = nb;
        }
        if (chosen.size() == qsize) break;
    }
    if (chosen.size() != qsize) return Status::err(GFSM_ERR_ARG, "random walk found no component of that size");
    struct E { uint32_t i, j, l; };
    std::vector<E> und;
    if (kind == GFSM_WALK_DENSE) {
        for (uint32_t i = 0; i < qsize; i++)
            for (uint32_t j = i + 1; j < qsize; j++) {
                uint32_t l = entry_label(data, chosen[i], chosen[j]);
                if (l) und.push_back({i, j, l});
            }
    } else {
        std::vector<E> extra;
        std::set<std::pair<uint32_t, uint32_t>> in_tree;
        for (auto &t : tree) {
            uint32_t i = std::min(t.first, t.second), j = std::max(t.first, t.second);
            in_tree.insert({i, j});
            und.push_back({i, j, entry_label(data, chosen[i], chosen[j])});
        }
        for (uint32_t i = 0; i < qsize; i++)
            for (uint32_t j = i + 1; j < qsize; j++) {
                if (in_tree.count({i, j})) continue;
                uint32_t l = entry_label(data, chosen[i], chosen[j]);
                if (l) extra.push_back({i, j, l});
            }
        if (!extra.empty()) und.push_back(extra[r.below(extra.size())]);
    }
    std::vector<uint32_t> src, dst, el, vl(qsize);
    for (uint32_t i = 0; i < qsize; i++) vl[i] = data.vlabel[chosen[i]];
    for (auto &e : und) {
        src.push_back(e.i); dst.push_back(e.j); el.push_back(e.l);
        src.push_back(e.j); dst.push_back(e.i); el.push_back(e.l);
    }
    return build_graph(qsize, vl.data(), src.size(), src.data(), dst.data(), el.data(), data.mode, out);
}

}  // namespace gfsm
