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
#include "host.h"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace gfsm {

Status plan_query(const HostGraph &qg, Plan &out)
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
    seen[rank[0]] = 1;
    struct Req { uint32_t dist, rel; };
    for (uint32_t d = 0; d < q; d++) {
        int64_t n = -1;
        for (uint32_t i = 0; i < q; i++)
            if (seen[rank[i]] && placed[rank[i]] < 0) { n = rank[i]; break; }
        if (n < 0) return Status::err(GFSM_ERR_QUERY, "query is not connected through out-entries");
        placed[n] = (int32_t)d;
        out.col[d] = (uint32_t)n;
        std::vector<Req> reqs;
        for (uint32_t j = qg.dir_ptr[n]; j < qg.dir_ptr[n + 1]; j++) {
            uint32_t r = qg.dir[j].rel;
            for (uint32_t p = qg.dir[j].beg; p < qg.dir[j + 1].beg; p++) {
                uint32_t u = qg.nbr[p];
                seen[u] = 1;
                if (placed[u] < 0) continue;
                uint32_t rel = r;
                for (uint32_t k = qg.dir_ptr[u]; k < qg.dir_ptr[u + 1]; k++) {
                    const uint32_t *b = qg.nbr.data() + qg.dir[k].beg, *e = qg.nbr.data() + qg.dir[k + 1].beg;
                    if (std::binary_search(b, e, (uint32_t)n)) { rel = qg.dir[k].rel; break; }
                }
                reqs.push_back({d - (uint32_t)placed[u], rel});
            }
        }
        std::sort(reqs.begin(), reqs.end(), [](const Req &a, const Req &b) {
            return a.dist != b.dist ? a.dist < b.dist : a.rel < b.rel;
        });
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
