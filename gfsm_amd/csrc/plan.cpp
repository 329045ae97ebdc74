// plan.cpp -- query compilation: matching order and per-depth requirements.
//
// Keeps the reference's order (HnGPU.cu:749-773 preProcessQuery,
// ccsr.cpp:98-204 genereateCCSRRequirements) so per-depth work is comparable:
//   * query vertices ranked by out-degree descending, then id ascending
//     (the CCSR row order, HnHelper.cuh:21-31, 413);
//   * depth 0 places rank 0; each later depth places the lowest-ranked vertex
//     that is an out-neighbour of a placed vertex and is not yet placed;
//   * when n is placed at depth d, each entry n->u to a placed u (u == n for a
//     self loop) becomes a requirement (distance d - depth(u), relation of
//     u->n when the query has it, else of n->u), sorted by distance;
//   * firstQ = smallest non-zero distance; its ancestor generates candidates.
// The reference generates from every entry of that ancestor's row and then
// checks all requirements; here the generator's requirement selects the
// segment that is streamed, and the remaining requirements become checks.
//
// With a data graph given (GFSM_ORDER_COST, no reference counterpart) the
// order is chosen from the data graph's relation statistics instead: the
// root is the vertex whose cheapest out-relation has the fewest source
// vertices, and each later depth places the reachable vertex with the
// smallest expected number of children per node (mean generator segment
// length times the pass rate of its checks).  Requirements are derived the
// same way for either order; the set of embeddings does not depend on it.
#include "host.h"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace gfsm {

// Mean segment length of `rel` over every vertex that could own one (those
// without a segment count as length 0) and the number of vertices its
// segments can reach (the reverse relation's sources in labelled mode).
void rel_model(const HostGraph &data, uint32_t rel, double &mean, double &reach)
{
    mean = 0;
    reach = 1;
    uint64_t src, ent;
    if (!relation_sources(data, rel, src, ent) || src == 0) return;
    mean = (double)ent / (double)std::max<uint64_t>(src, relation_pool(data, rel));
    uint32_t rev = rel;
    if (data.mode == GFSM_MODE_VLABEL) rev = ((rel >> 10) & 0x7FF) << 21 | (rel >> 21) << 10 | (rel & 0x3FF);
    uint64_t rsrc = src, rent;
    if (!relation_sources(data, rev, rsrc, rent)) rsrc = src;
    reach = std::max<double>(1.0, (double)rsrc);
}

// Expected level sizes: a candidate from a segment of relation r passes a
// check of relation c with probability min(1, mean length of c / vertices c
// reaches); each survivor has (mean length of the next generator) children.
void project_levels(const HostGraph &data, const Plan &plan, uint64_t nroots, uint64_t g1, double *nodes,
                    double *cand)
{
    const uint32_t q = plan.q;
    double S = (double)nroots, G = (double)g1;
    for (uint32_t k = 0; k < q; k++) nodes[k] = cand[k] = 0;
    if (!q) return;
    nodes[0] = S;
    cand[0] = S;
    for (uint32_t d = 1; d < q; d++) {
        const DepthPlanDev &dp = plan.dev.depth[d];
        cand[d] = G;
        double p = 1;
        for (uint32_t r = 0; r < dp.nchk; r++) {
            double m, reach;
            rel_model(data, dp.chk_rel[r], m, reach);
            p *= std::min(1.0, m / reach);
        }
        S = std::min(G * p, 1e30);
        nodes[d] = S;
        double m, reach;
        rel_model(data, dp.nxt_rel, m, reach);
        G = std::min(S * m, 1e30);
    }
}

Status plan_query(const HostGraph &qg, Plan &out, const HostGraph *by_cost)
{
    out = Plan{};
    uint32_t q = qg.V;
    if (q == 0) return Status::err(GFSM_ERR_QUERY, "empty query");
    if (q > GFSM_MAXQ) return Status::err(GFSM_ERR_QUERY, "query larger than 32 vertices");
    std::vector<uint32_t> deg(q);
    uint32_t maxval = 0;
    for (uint32_t v = 0; v < q; v++) {
        deg[v] = qg.degree(v);
        maxval = std::max(maxval, deg[v]);
    }
    // Every query vertex needs a CCSR row in the reference (its offset walk,
    // ccsr.cpp:103-108, assumes one); a lone vertex has none.
    if (q >= 2)
        for (uint32_t v = 0; v < q; v++)
            if (!deg[v]) return Status::err(GFSM_ERR_QUERY, "query vertex " + std::to_string(v) + " has no out-entries");
    std::vector<uint32_t> rank(q);
    std::iota(rank.begin(), rank.end(), 0u);
    std::sort(rank.begin(), rank.end(), [&](uint32_t a, uint32_t b) {
        if (deg[a] != deg[b]) return deg[a] > deg[b];
        return a < b;
    });

    out.q = q;
    out.maxval = maxval;
    out.col.assign(q, 0);
    out.header.assign(3 * q, 0);
    out.requirements.assign((size_t)2 * std::max(maxval, 1u) * q, 0);
    std::memset(&out.dev, 0, sizeof(out.dev));
    out.dev.q = q;

    std::vector<int32_t> placed(q, -1);
    std::vector<char> seen(q, 0);
    struct Req { uint32_t dist, rel; };
    // Requirements of vertex n if it is placed at depth d: each entry n->u to
    // a placed u (the relation of u->n when the query has it, else of n->u),
    // by distance.  Placing n itself makes its self loops distance 0.
    auto reqs_of = [&](uint32_t n, uint32_t d) {
        std::vector<Req> reqs;
        for (uint32_t j = qg.dir_ptr[n]; j < qg.dir_ptr[n + 1]; j++) {
            uint32_t r = qg.dir[j].rel;
            for (uint32_t p = qg.dir[j].beg; p < qg.dir[j + 1].beg; p++) {
                uint32_t u = qg.nbr[p];
                if (placed[u] < 0 && u != n) continue;
                const uint32_t du = (u == n) ? d : (uint32_t)placed[u];
                uint32_t rel = r;
                for (uint32_t k = qg.dir_ptr[u]; k < qg.dir_ptr[u + 1]; k++) {
                    const uint32_t *b = qg.nbr.data() + qg.dir[k].beg, *e = qg.nbr.data() + qg.dir[k + 1].beg;
                    if (std::binary_search(b, e, (uint32_t)n)) { rel = qg.dir[k].rel; break; }
                }
                reqs.push_back({d - du, rel});
            }
        }
        std::sort(reqs.begin(), reqs.end(), [](const Req &a, const Req &b) {
            return a.dist != b.dist ? a.dist < b.dist : a.rel < b.rel;
        });
        return reqs;
    };
    // Next vertex in the reference order: the lowest-ranked reachable one.
    auto next_reference = [&](uint32_t d) -> int64_t {
        if (d == 0) return rank[0];
        for (uint32_t i = 0; i < q; i++)
            if (seen[rank[i]] && placed[rank[i]] < 0) return rank[i];
        return -1;
    };
    // Next vertex by expected cost on the data graph (ties: rank order).
    auto next_by_cost = [&](uint32_t d) -> int64_t {
        int64_t best = -1;
        double best_cost = 0;
        size_t best_reqs = 0;
        for (uint32_t i = 0; i < q; i++) {
            const uint32_t n = rank[i];
            if (placed[n] >= 0 || (d && !seen[n])) continue;
            double cost;
            size_t nreq = 0;
            if (d == 0) {
                cost = 1e300;  // fewest roots: the cheapest out-relation's sources
                for (uint32_t j = qg.dir_ptr[n]; j < qg.dir_ptr[n + 1]; j++) {
                    uint64_t src = 0, ent;
                    if (!relation_sources(*by_cost, qg.dir[j].rel, src, ent)) src = 0;
                    cost = std::min(cost, (double)src);
                }
            } else {
                const std::vector<Req> reqs = reqs_of(n, d);
                bool gen = false;
                cost = 1;
                for (auto &r : reqs) {
                    double mean, reach;
                    rel_model(*by_cost, r.rel, mean, reach);
                    if (!gen && r.dist) {
                        cost *= mean;  // children per node: the generator segment
                        gen = true;
                    } else {
                        cost *= std::min(1.0, mean / reach);  // pass rate of a check
                    }
                }
                if (!gen) continue;  // no placed out-neighbour to generate from
                nreq = reqs.size();
            }
            if (best < 0 || cost < best_cost || (cost == best_cost && nreq > best_reqs)) {
                best = n;
                best_cost = cost;
                best_reqs = nreq;
            }
        }
        return best;
    };
    for (uint32_t d = 0; d < q; d++) {
        const int64_t n = by_cost ? next_by_cost(d) : next_reference(d);
        if (n < 0) return Status::err(GFSM_ERR_QUERY, "query is not connected through out-entries");
        const std::vector<Req> reqs = reqs_of((uint32_t)n, d);
        placed[n] = (int32_t)d;
        out.col[d] = (uint32_t)n;
        for (uint32_t p = qg.dir[qg.dir_ptr[n]].beg; p < qg.dir[qg.dir_ptr[n + 1]].beg; p++) seen[qg.nbr[p]] = 1;
        uint32_t firstQ = 0;
        for (auto &r : reqs)
            if (r.dist) { firstQ = r.dist; break; }
        if (d && (!firstQ || firstQ > d))
            return Status::err(GFSM_ERR_QUERY, "query compile failure at depth " + std::to_string(d));
        uint32_t off = d * std::max(maxval, 1u) * 2;
        out.header[3 * d + 0] = (uint32_t)reqs.size();
        out.header[3 * d + 1] = off;
        out.header[3 * d + 2] = firstQ;
        for (size_t i = 0; i < reqs.size(); i++) {
            out.requirements[off + 2 * i] = i ? reqs[i].dist - reqs[i - 1].dist : reqs[i].dist;
            out.requirements[off + 2 * i + 1] = reqs[i].rel;
        }
        // Device plan.
        if (d == 0) {
            for (auto &r : reqs) {
                if (out.dev.nroot_chk >= GFSM_MAXCHK) return Status::err(GFSM_ERR_QUERY, "too many requirements");
                out.dev.root_chk_rel[out.dev.nroot_chk++] = r.rel;
            }
        } else {
            DepthPlanDev &dp = out.dev.depth[d];
            bool gen_taken = false;
            for (auto &r : reqs) {
                if (!gen_taken && r.dist == firstQ) {
                    dp.gen_dist = r.dist;
                    dp.gen_rel = r.rel;
                    gen_taken = true;
                    continue;
                }
                if (dp.nchk >= GFSM_MAXCHK) return Status::err(GFSM_ERR_QUERY, "too many requirements");
                dp.chk_dist[dp.nchk] = r.dist;
                dp.chk_rel[dp.nchk] = r.rel;
                dp.nchk++;
            }
        }
    }
    for (uint32_t d = 0; d < q; d++) {
        out.dev.col[d] = out.col[d];
        out.dev.dcol[out.col[d]] = d;
        if (d + 1 < q) {
            out.dev.depth[d].nxt_dist = out.dev.depth[d + 1].gen_dist;
            out.dev.depth[d].nxt_rel = out.dev.depth[d + 1].gen_rel;
        }
    }
    if (q >= 2) out.dev.root_rel = out.dev.depth[1].gen_rel;
    return Status::ok();
}

}  // namespace gfsm
