// Data and query graphs (inc/graph.h:150-495) and auxiliary edge tables (Edges,
// inc/graph.h:13-49) for the drop-in C++ surface.
//
// A Graph keeps the reference's CSR on the host (sorted neighbour lists) and uploads itself
// to HBM on first use, labels included; it stays resident until the Graph is destroyed.
// Labels reach the device filters (labelled NLF / GQL) and the plan; the data graph's
// labels are fixed once it is on the device.
#pragma once

#include <algorithm>
#include <fstream>
#include <map>
#include <sstream>

#include "common.h"

class Edges {
public:
    uint32_t* _offset = nullptr;
    uint32_t* _edge = nullptr;    // candidate index k into C(b)
    uint32_t* _edge_v = nullptr;  // the same entries as vertex ids
    uint32_t _v_cnt = 0;
    uint32_t _e_cnt = 0;
    uint32_t _max_degree = 0;

    Edges() = default;
    Edges(const Edges&) = delete;
    Edges& operator=(const Edges&) = delete;
    ~Edges()
    {
        delete[] _offset;
        delete[] _edge;
        delete[] _edge_v;
    }
    const uint32_t* getNeb(const uint32_t vertex_id, uint32_t& cnt)
    {
        cnt = _offset[vertex_id + 1] - _offset[vertex_id];
        return _edge + _offset[vertex_id];
    }
    const uint32_t* getNeb_V(const uint32_t vertex_id, uint32_t& cnt)
    {
        cnt = _offset[vertex_id + 1] - _offset[vertex_id];
        return _edge_v + _offset[vertex_id];
    }
};

class Graph {
public:
    std::string path;

    Graph() = default;
    Graph(const Graph&) = delete;
    Graph& operator=(const Graph&) = delete;
    virtual ~Graph() { release_device(); }

    // Reference text format: "t V E", then "v id label degree" lines, then "e u v" lines
    // (inc/graph.h:190-266).  Degrees must match the edge list (the reference sizes its
    // CSR from them); a mismatch, a self loop or a duplicate edge throws.
    virtual void Load(const std::string& file_path)
    {
        std::ifstream in(file_path);
        if (!in.is_open()) throw sgm::Error(SGM_EINVAL, "can't open " + file_path);
        path = file_path;
        char type = 0;
        uint64_t V = 0, E = 0;
        if (!(in >> type >> V >> E) || type != 't') throw sgm::Error(SGM_EINVAL, file_path + ": missing 't V E' header");
        std::vector<uint32_t> decl(V, 0);
        std::vector<LabelID> labels(V, 0);
        std::vector<std::pair<uint32_t, uint32_t>> edges;
        edges.reserve(E);
        while (in >> type) {
            if (type == 'v') {
                uint64_t id, lab, d;
                in >> id >> lab >> d;
                if (id >= V) throw sgm::Error(SGM_EINVAL, file_path + ": vertex id out of range");
                labels[id] = (LabelID)lab;
                decl[id] = (uint32_t)d;
            } else if (type == 'e') {
                uint64_t a, b;
                in >> a >> b;
                if (a >= V || b >= V || a == b) throw sgm::Error(SGM_EINVAL, file_path + ": bad edge");
                edges.emplace_back((uint32_t)a, (uint32_t)b);
            } else {
                std::string rest;
                std::getline(in, rest);
            }
        }
        if (edges.size() != E) throw sgm::Error(SGM_EINVAL, file_path + ": edge count differs from the header");
        AssignEdges((uint32_t)V, edges);
        for (uint32_t v = 0; v < _v_cnt; v++)
            if (getVertexDegree(v) != decl[v])
                throw sgm::Error(SGM_EINVAL, file_path + ": declared degree differs from the edge list");
        _labels = std::move(labels);
        count_labels();
    }

    // Binary CSR cache written by subgraphmatching_amd.graph.save_csr (little-endian):
    // "SGMCSR01", u64 V, u64 A (= 2E), u64 flags (bit 0: u32 labels[V] follow), u32 off[V+1],
    // u32 adj[A].  No reference counterpart: a cache of the text format above.
    void LoadCSR(const std::string& file_path)
    {
        std::ifstream in(file_path, std::ios::binary);
        if (!in.is_open()) throw sgm::Error(SGM_EINVAL, "can't open " + file_path);
        char magic[8];
        uint64_t head[3];
        if (!in.read(magic, 8) || std::string(magic, 8) != "SGMCSR01" ||
            !in.read(reinterpret_cast<char*>(head), sizeof head))
            throw sgm::Error(SGM_EINVAL, file_path + ": not an sgm binary CSR");
        const uint64_t V = head[0], A = head[1];
        if (V >= 0xFFFFFFFFull || A >= 0x80000000ull || (A & 1))
            throw sgm::Error(SGM_EINVAL, file_path + ": sizes out of range");
        std::vector<uint32_t> off(V + 1), adj(A);
        if (!in.read(reinterpret_cast<char*>(off.data()), sizeof(uint32_t) * (V + 1)) ||
            !in.read(reinterpret_cast<char*>(adj.data()), sizeof(uint32_t) * A))
            throw sgm::Error(SGM_EINVAL, file_path + ": truncated");
        std::vector<LabelID> labels(V, 0);
        if (head[2] & 1) {
            std::vector<uint32_t> lab(V);
            if (!in.read(reinterpret_cast<char*>(lab.data()), sizeof(uint32_t) * V))
                throw sgm::Error(SGM_EINVAL, file_path + ": truncated labels");
            for (uint64_t v = 0; v < V; v++) labels[v] = (LabelID)lab[v];
        }
        if (off[0] != 0 || off[V] != A) throw sgm::Error(SGM_EINVAL, file_path + ": offsets are not a CSR");
        uint32_t maxd = 0;
        for (uint64_t v = 0; v < V; v++) {
            if (off[v + 1] < off[v]) throw sgm::Error(SGM_EINVAL, file_path + ": offsets are not a CSR");
            for (uint32_t p = off[v]; p < off[v + 1]; p++)
                if (adj[p] >= V || (p > off[v] && adj[p] <= adj[p - 1]))
                    throw sgm::Error(SGM_EINVAL, file_path + ": neighbour lists must be sorted, simple, in range");
            maxd = std::max(maxd, off[v + 1] - off[v]);
        }
        release_device();
        path = file_path;
        _v_cnt = (uint32_t)V;
        _offset = std::move(off);
        _neb = std::move(adj);
        _e_cnt = (uint32_t)(A / 2);
        _max_degree = maxd;
        _labels = std::move(labels);
        count_labels();
    }

    // Programmatic construction from an undirected edge list (each edge once).
    void AssignEdges(uint32_t V, const std::vector<std::pair<uint32_t, uint32_t>>& edges)
    {
        release_device();
        _v_cnt = V;
        _offset.assign((size_t)V + 1, 0);
        for (auto& e : edges) { _offset[e.first + 1]++; _offset[e.second + 1]++; }
        for (uint32_t v = 0; v < V; v++) _offset[v + 1] += _offset[v];
        _neb.assign(_offset[V], 0);
        std::vector<uint32_t> fill(_offset.begin(), _offset.end() - 1);
        for (auto& e : edges) { _neb[fill[e.first]++] = e.second; _neb[fill[e.second]++] = e.first; }
        _max_degree = 0;
        for (uint32_t v = 0; v < V; v++) {
            std::sort(_neb.begin() + _offset[v], _neb.begin() + _offset[v + 1]);
            for (uint32_t p = _offset[v] + 1; p < _offset[v + 1]; p++)
                if (_neb[p] == _neb[p - 1]) throw sgm::Error(SGM_EINVAL, "duplicate edge");
            _max_degree = std::max(_max_degree, _offset[v + 1] - _offset[v]);
        }
        _e_cnt = (uint32_t)edges.size();
        _labels.assign(V, 0);
        count_labels();
    }

    uint32_t getVertexCnt() const { return _v_cnt; }
    uint32_t getVertexDegree(const VertexID id) const { return _offset[id + 1] - _offset[id]; }
    uint32_t getEdgeCnt() const { return _e_cnt; }
    uint32_t getLabelCnt() const { return _label_cnt; }
    uint32_t getMaxDgree() const { return _max_degree; }
    uint32_t getMaxLabelFreq() const { return _max_label_freq; }
    uint32_t getLabelFreq(LabelID label) const
    {
        auto it = _labels_freq.find(label);
        return it == _labels_freq.end() ? 0 : it->second;
    }
    const VertexID* getNeb(VertexID id, uint32_t& neb_count) const
    {
        neb_count = _offset[id + 1] - _offset[id];
        return _neb.data() + _offset[id];
    }
    LabelID getVertexLabel(VertexID id) const { return _labels[id]; }
    const LabelID* getVertexLabels() const { return _labels.data(); }
    bool hasEdge(VertexID u, VertexID v) const
    {
        return std::binary_search(_neb.begin() + _offset[u], _neb.begin() + _offset[u + 1], v);
    }
    const uint32_t* offsets() const { return _offset.data(); }
    const uint32_t* neighbours() const { return _neb.data(); }
    bool unlabeled() const { return _label_cnt <= 1 && getLabelFreq(0) == _v_cnt; }

    // The resident device copy (uploaded on first call).
    sgm_graph* device() const
    {
        std::lock_guard<std::mutex> lk(sgm::device_mutex());
        if (!_dev) {
            sgm_graph* g = nullptr;
            sgm::check(sgm_graph_upload(sgm::context(), _offset.data(), _neb.data(), _v_cnt, _neb.size(), &g),
                       "graph upload");
            _dev = g;
            if (!unlabeled()) sgm::check(sgm_graph_set_labels(g, _labels.data(), _v_cnt), "graph labels");
        }
        return _dev;
    }

protected:
    uint32_t _v_cnt = 0, _e_cnt = 0, _label_cnt = 0, _max_degree = 0, _max_label_freq = 0;
    std::vector<uint32_t> _offset{0};
    std::vector<VertexID> _neb;
    std::vector<LabelID> _labels;
    std::map<LabelID, uint32_t> _labels_freq;
    mutable sgm_graph* _dev = nullptr;

    void count_labels()
    {
        _labels_freq.clear();
        LabelID max_label = 0;
        for (LabelID l : _labels) { _labels_freq[l]++; max_label = std::max(max_label, l); }
        _label_cnt = _v_cnt ? std::max<uint32_t>((uint32_t)_labels_freq.size(), max_label + 1) : 0;
        _max_label_freq = 0;
        for (auto& e : _labels_freq) _max_label_freq = std::max(_max_label_freq, e.second);
    }
    void release_device()
    {
        if (_dev) {
            std::lock_guard<std::mutex> lk(sgm::device_mutex());
            sgm_graph_destroy(_dev);
            _dev = nullptr;
        }
    }
};

// Query graph (inc/graph.h:382-495): same format, at most 32 vertices on this path.
class Query : public Graph {
public:
    void Load(const std::string& file_path) override
    {
        Graph::Load(file_path);
        if (getVertexCnt() > 32) throw sgm::Error(SGM_EINVAL, file_path + ": queries have at most 32 vertices");
    }
};

// A vertex of a query's BFS tree (inc/graph.h:497-517).
class TreeNode {
public:
    VertexID _id = 0;
    uint32_t _level = 0;
    VertexID _parent = 0;
    std::vector<VertexID> _children;
    size_t _estimated_embeddings_num = 0;
};

// BFS (inc/common_ops.h:318-350) of a query graph (at most 32 vertices): tree gets a new
// TreeNode[vertex count] (the caller deletes it, as the reference's callers do); children in
// the order they are discovered, i.e. the adjacency order of their parent.
inline void BFS(const Graph& graph, VertexID start_vertex, TreeNode*& tree, Order& bfs_order)
{
    const uint32_t n = graph.getVertexCnt();
    if (n > 32) throw sgm::Error(SGM_EINVAL, "BFS: queries have at most 32 vertices");
    std::vector<uint32_t> order(n), parent(n), level(n);
    const int got = sgm_query_bfs(graph.offsets(), graph.neighbours(), n, start_vertex, order.data(), parent.data(),
                                  level.data());
    if (got < 0) throw sgm::Error(got, "BFS");
    tree = new TreeNode[n];
    bfs_order.assign(order.begin(), order.begin() + got);
    for (uint32_t i = 0; i < (uint32_t)got; i++) {
        const VertexID u = order[i];
        tree[u]._id = u;
        tree[u]._level = level[u];
        tree[u]._parent = u == start_vertex ? 0 : parent[u];
        uint32_t cnt = 0;
        const VertexID* nb = graph.getNeb(u, cnt);
        for (uint32_t k = 0; k < cnt; k++)
            if (nb[k] != start_vertex && parent[nb[k]] == u && level[nb[k]] == level[u] + 1) tree[u]._children.push_back(nb[k]);
    }
}
