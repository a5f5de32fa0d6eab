// std::sort over records ordered like the reference's VertexSetIR::operator<
// (inc/code_gen.h:633-650): depth, label (always 0), size, then lexicographic edge
// indices.  Test infrastructure: it lets tests compare the oracle's restated sort with
// the library routine itself on the same inputs.
#include <algorithm>
#include <cstdint>
#include <vector>

namespace {
struct Rec {
    uint32_t depth, size, payload;
    std::vector<uint32_t> e;
    bool operator<(const Rec &r) const {
        if (this == &r) return false;
        if (depth != r.depth) return depth < r.depth;
        if (e.size() != r.e.size()) return e.size() < r.e.size();
        for (size_t i = 0; i < e.size(); i++)
            if (e[i] != r.e[i]) return e[i] < r.e[i];
        return false;
    }
};
}  // namespace

extern "C" void probe_std_sort(long n, const uint32_t *depth, const uint32_t *size,
                               const uint32_t *elems, uint32_t *perm_out) {
    std::vector<Rec> v(n);
    for (long i = 0; i < n; i++) {
        v[i].depth = depth[i];
        v[i].size = size[i];
        v[i].payload = (uint32_t)i;
        v[i].e.assign(elems + i * 32, elems + i * 32 + size[i]);
    }
    std::sort(v.begin(), v.end());
    for (long i = 0; i < n; i++) perm_out[i] = v[i].payload;
}

// std::priority_queue over (vertex, key) pairs with the comparator of the reference's
// CFLStartVertex (inc/filter.h:437-441: l.second < r.second, a max-heap on the key): push
// (v, keys[v]) for v < count, pop while more than `keep` remain, then pop the rest into
// out[] (the order in which CFLStartVertex visits them).
#include <queue>
#include <utility>

extern "C" void probe_std_heap(unsigned count, const double *keys, unsigned keep, uint32_t *out) {
    auto cmp = [](std::pair<uint32_t, double> l, std::pair<uint32_t, double> r) { return l.second < r.second; };
    std::priority_queue<std::pair<uint32_t, double>, std::vector<std::pair<uint32_t, double>>, decltype(cmp)> q(cmp);
    for (unsigned v = 0; v < count; v++) q.push(std::make_pair((uint32_t)v, keys[v]));
    while (q.size() > keep) q.pop();
    unsigned k = 0;
    while (!q.empty()) {
        out[k++] = q.top().first;
        q.pop();
    }
}
