// Candidate filters (inc/filter.h:1049-1087), drop-in over the C ABI: the NLF, GQL, CFL,
// CECI and DAF kernels run on the device, then FilterParamConstruct's host arrays and
// aux-edge tables are filled from the device candidate space.  All return false when some
// C(u) is empty, as the reference does (Filter_CECI also where the reference ends the process
// on an empty set, inc/filter.h:700-703); candidates / candidates_count then hold the
// filtered sets and the membership arrays and tables stay unbuilt.
#pragma once

#include "api.h"

namespace sgm {
inline bool run_filter(const Graph& data, const Query& query, CandidateParam& can, int kind)
{
    if (&can.data() != &data || &can.query() != &query)
        throw Error(SGM_EINVAL, "CandidateParam was built for another data graph or query");
    std::lock_guard<std::mutex> lk(device_mutex());
    const int rc = sgm_filter(can.device(), kind);
    if (rc != SGM_OK && rc != SGM_EEMPTY) check(rc, "filter");
    if (rc == SGM_OK) check(sgm_build_tables(can.device()), "tables");
    can.fetch_host_arrays(rc == SGM_OK);
    return rc == SGM_OK;
}
}  // namespace sgm

inline bool Filter_NLF(const Graph& data, const Query& query, CandidateParam& can)
{
    return sgm::run_filter(data, query, can, SGM_FILTER_NLF);
}

inline bool Filter_GQL(const Graph& data, const Query& query, CandidateParam& can)
{
    return sgm::run_filter(data, query, can, SGM_FILTER_GQL);
}

inline bool Filter_CFL(const Graph& data, const Query& query, CandidateParam& can)
{
    return sgm::run_filter(data, query, can, SGM_FILTER_CFL);
}

inline bool Filter_CECI(const Graph& data, const Query& query, CandidateParam& can)
{
    return sgm::run_filter(data, query, can, SGM_FILTER_CECI);
}

inline bool Filter_DAF(const Graph& data, const Query& query, CandidateParam& can)
{
    return sgm::run_filter(data, query, can, SGM_FILTER_DAF);
}
