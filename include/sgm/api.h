// Carriers of the filter -> order -> enumerate calls (inc/api.h:9-227), drop-in over the
// C ABI.
//
// CandidateParam owns a device candidate space (sgm_query) plus the reference's host
// arrays, which the filter calls fill from the device: candidates[u] (ascending ids),
// candidates_count[u], candidates_exsist[u][v] and edge_matrix[a][b] for every query
// edge (the buildTables CSR, inc/table.h:7-137).  Host arrays follow the reference's
// ownership (new[]/delete[], inc/api.h:81-112); Config owns nothing (inc/api.h:206-209).
#pragma once

#include "graph.h"

class CandidateParam {
public:
    VertexID** candidates = nullptr;
    uint32_t* candidates_count = nullptr;
    Edges*** edge_matrix = nullptr;
    uint32_t _qv_cnt = 0;
    bool** candidates_exsist = nullptr;

    CandidateParam(const Graph& data, const Query& query) : _data(&data), _query(&query)
    {
        _qv_cnt = query.getVertexCnt();
        const uint32_t V = data.getVertexCnt();
        const uint32_t cap = std::max<uint32_t>(data.getMaxLabelFreq(), 1);
        candidates_count = new uint32_t[_qv_cnt]();
        candidates = new VertexID*[_qv_cnt];
        candidates_exsist = new bool*[_qv_cnt];
        edge_matrix = new Edges**[_qv_cnt];
        for (uint32_t i = 0; i < _qv_cnt; ++i) {
            candidates[i] = new VertexID[cap];
            candidates_exsist[i] = new bool[V ? V : 1]();
            edge_matrix[i] = new Edges*[_qv_cnt];
            std::fill(edge_matrix[i], edge_matrix[i] + _qv_cnt, nullptr);
        }
        sgm_graph* g = data.device();
        std::lock_guard<std::mutex> lk(sgm::device_mutex());
        sgm_query* q = nullptr;
        sgm::check(sgm_query_create(g, query.offsets(), query.neighbours(), _qv_cnt, &q), "query create");
        _dev = q;
        if (!query.unlabeled()) sgm::check(sgm_query_set_labels(q, query.getVertexLabels(), _qv_cnt), "query labels");
    }
    CandidateParam(const CandidateParam&) = delete;
    CandidateParam& operator=(const CandidateParam&) = delete;

    ~CandidateParam()
    {
        for (uint32_t i = 0; i < _qv_cnt; ++i) {
            delete[] candidates[i];
            delete[] candidates_exsist[i];
            for (uint32_t j = 0; j < _qv_cnt; ++j) delete edge_matrix[i][j];
            delete[] edge_matrix[i];
        }
        delete[] candidates;
        delete[] candidates_exsist;
        delete[] edge_matrix;
        delete[] candidates_count;
        if (_dev) {
            std::lock_guard<std::mutex> lk(sgm::device_mutex());
            sgm_query_destroy(_dev);
        }
    }

    sgm_query* device() const { return _dev; }
    const Graph& data() const { return *_data; }
    const Query& query() const { return *_query; }

    // Filter_* body after the device filter: fetch the candidate space into the host arrays.
    // With `construct` (the filter succeeded) also candidates_exsist and the edge tables,
    // as FilterParamConstruct does (inc/filter.h:60-78); the reference skips that step when
    // a candidate set is empty.  Called with device_mutex held.
    void fetch_host_arrays(bool construct)
    {
        const uint32_t n = _qv_cnt, V = _data->getVertexCnt();
        sgm::check(sgm_cand_counts(_dev, candidates_count), "candidate counts");
        for (uint32_t u = 0; u < n; u++)
            if (candidates_count[u])
                sgm::check(sgm_cand_download(_dev, u, candidates[u], candidates_count[u]), "candidate download");
        if (!construct) return;
        std::vector<uint32_t> words(V ? V : 1);
        sgm::check(sgm_cand_exist_download(_dev, words.data(), V), "candidate membership");
        for (uint32_t u = 0; u < n; u++)
            for (uint32_t v = 0; v < V; v++) candidates_exsist[u][v] = (words[v] >> u) & 1u;
        for (uint32_t a = 0; a < n; a++) {
            uint32_t cnt = 0;
            const VertexID* nb = _query->getNeb(a, cnt);
            for (uint32_t i = 0; i < cnt; i++) {
                const uint32_t b = nb[i];
                uint64_t e = 0;
                uint32_t md = 0;
                sgm::check(sgm_table_info(_dev, a, b, &e, &md), "table info");
                Edges* t = new Edges();
                t->_v_cnt = candidates_count[a];
                t->_e_cnt = (uint32_t)e;
                t->_max_degree = md;
                t->_offset = new uint32_t[(size_t)t->_v_cnt + 1];
                t->_edge = new uint32_t[e ? e : 1];
                t->_edge_v = new uint32_t[e ? e : 1];
                sgm::check(sgm_table_download(_dev, a, b, t->_offset, t->_edge, t->_edge_v), "table download");
                delete edge_matrix[a][b];
                edge_matrix[a][b] = t;
            }
        }
    }

private:
    const Graph* _data;
    const Query* _query;
    sgm_query* _dev = nullptr;
};

// PlanIR (inc/code_gen.h:682-847) and its ops (VertexSetIR, :573-680): the intersection-cache
// plan exactly as the library builds and runs it (sgm_plan_export_labeled), with the
// reference's accessors.  Op ids are positions in getTotSetOpsArr(); INVALID marks no parent
// and no iteration set (depth 0).
class VertexSetIR {
public:
    static constexpr uint32_t INVALID = 0xFFFFFFFFu;
    VertexSetIR() = default;
    VertexSetIR(uint32_t vid, uint32_t parent, uint32_t depth, uint32_t v_depth, LabelID label,
                std::vector<VertexID> edge_idx)
        : _vid(vid), _parent_id(parent), _depth(depth), _v_depth(v_depth), _v_label(label), _edge_idx(std::move(edge_idx))
    {
    }
    const std::vector<VertexID> getEdge(const Order& order) const
    {
        std::vector<VertexID> e(_edge_idx.size());
        for (size_t i = 0; i < _edge_idx.size(); i++) e[i] = order[_edge_idx[i]];
        return e;
    }
    const std::vector<VertexID>& getEdgeIdx() const { return _edge_idx; }
    uint32_t getVSetID() const { return _vid; }
    VertexID getDepth() const { return _depth; }
    uint32_t getParentVSetID() const { return _parent_id; }
    uint32_t getVDepth() const { return _v_depth; }
    LabelID getLabel() const { return _v_label; }
    bool hasEdge(VertexID v) const { return std::find(_edge_idx.begin(), _edge_idx.end(), v) != _edge_idx.end(); }

private:
    uint32_t _vid{INVALID}, _parent_id{INVALID}, _depth{INVALID}, _v_depth{INVALID};
    LabelID _v_label{INVALID};
    std::vector<VertexID> _edge_idx;
};

class PlanIR {
public:
    PlanIR() = default;
    // Throws sgm::Error(SGM_EORDER) where the reference exits (a disconnected order,
    // inc/code_gen.h:743-746).
    PlanIR(const Order& order, const Query& query) : _order(order)
    {
        const uint32_t n = query.getVertexCnt();
        if (order.size() != n) throw sgm::Error(SGM_EORDER, "order length differs from the query size");
        std::vector<uint32_t> depth(512), parent(512), vdepth(512), emask(512), iter(32);
        uint32_t n_ops = 0;
        sgm::check(sgm_plan_export_labeled(query.offsets(), query.neighbours(),
                                           query.unlabeled() ? nullptr : query.getVertexLabels(), n, order.data(),
                                           &n_ops, depth.data(), parent.data(), vdepth.data(), emask.data(),
                                           iter.data()),
                   "plan");
        _setOps.assign(n, {});
        for (uint32_t i = 0; i < n_ops; i++) {
            std::vector<VertexID> e;
            for (uint32_t b = 0; b < 32; b++)
                if ((emask[i] >> b) & 1u) e.push_back(b);
            const LabelID lab = query.unlabeled() ? 0 : query.getVertexLabels()[order[vdepth[i]]];
            _ops.emplace_back(i, parent[i], depth[i], vdepth[i], lab, std::move(e));
            _setOps.at(depth[i]).push_back(i);
        }
        _iterSet.assign(iter.begin(), iter.begin() + n);
    }
    const std::vector<VertexID>& getOrder() const { return _order; }
    const std::vector<std::vector<uint32_t>>& getSetOps() const { return _setOps; }
    const std::vector<uint32_t>& getIterVSet() const { return _iterSet; }
    uint32_t getIterVSetAt(uint32_t depth) const { return _iterSet[depth]; }
    const std::vector<VertexSetIR>& getTotSetOpsArr() const { return _ops; }
    uint32_t getTotOp() const { return (uint32_t)_ops.size(); }
    const std::vector<VertexSetIR> getSetOpsArray() const { return _ops; }
    bool isLastOp(uint32_t cur_dep) const { return cur_dep == _iterSet.size() - 2; }

private:
    std::vector<VertexID> _order;
    std::vector<VertexSetIR> _ops;
    std::vector<std::vector<uint32_t>> _setOps;
    std::vector<uint32_t> _iterSet;
};

// ReuseParam (inc/api.h:167-182): owns its plan (and restrictions) as the reference does.
// The accelerated search builds the same plan on the device side of the enumeration call;
// this one is for callers that read reuse->plan.
class ReuseParam {
public:
    Restriction* rests{nullptr};
    PlanIR* plan{nullptr};
    Order order;
    ReuseParam() = default;
    ReuseParam(Restriction* r, PlanIR* p) : rests(r), plan(p) {}
    ReuseParam(const Query& query, const Order& ord) : order(ord) { plan = new PlanIR(ord, query); }
    ReuseParam(const ReuseParam&) = delete;
    ReuseParam& operator=(const ReuseParam&) = delete;
    ~ReuseParam()
    {
        delete plan;
        delete rests;
    }
};

// Config (inc/api.h:184-227).  Check() throws where the reference exits.
class Config {
public:
    bool useFilter{false};
    std::string filterType{"gql"};
    std::string orderType{"gql"};
    bool useCache{false};
    bool usePrune{false};
    CandidateParam* can{nullptr};
    ReuseParam* reuse{nullptr};
    bool useDAFFailingPrune{false};
    bool useGuPFailingPrune{false};
    bool useGUPConflictDetection{false};
    bool useBICEConflictDetection{false};
    bool useVEQAutomorphismPrune{false};
    int gupConflictDetectionSize{3};

    void Check() const
    {
        if (useFilter && can == nullptr) throw sgm::Error(SGM_EINVAL, "config error, use filter but canParam is nullptr");
        if (useCache && reuse == nullptr) throw sgm::Error(SGM_EINVAL, "config error, use cache but reuseParam is nullptr");
    }
};
