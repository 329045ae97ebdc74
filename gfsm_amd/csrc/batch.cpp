// batch.cpp -- placeholder: sequential batches over the plain path.
#include "serve.h"

namespace gfsm {

Status match_batch(Ctx &ctx, const HostGraph &data, const std::vector<const HostGraph *> &queries,
                   const gfsm_match_opts &opts, std::vector<Result> &out)
{
    out.assign(queries.size(), Result{});
    for (size_t i = 0; i < queries.size(); i++) {
        Status st = match_plain(ctx, data, *queries[i], opts, out[i]);
        if (!st.good()) out[i].status = st.code;
    }
    return Status::ok();
}

}  // namespace gfsm
