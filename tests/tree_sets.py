"""Set-level restatements of the BFS-tree filters and the QSI / CFL / GuP orders (pure
Python, small inputs), written independently of oracle/sgm_oracle_tree.c to pin it.

The oracle follows the reference's array shape (candidate arrays with INVALID holes,
flag counting over the pivots' candidates, frontier invalidation).  Here every step is a
statement about sets: "v survives iff N(v) meets C(p) for every pivot p", "a parent
candidate with no NLF(u) neighbour is dropped", with Python sets, dicts and heapq.  The
core numbers come from repeated peeling for each k, the BFS tree from a deque, and the one
libstdc++-specific tie order (CFLStartVertex's priority queue) from std::priority_queue
itself through oracle/libsortprobe.so.
"""
from __future__ import annotations

import heapq
from collections import deque

import oracle as O
from brute import adjsets, nlf_by_definition


def qnbrs(q):
    return [q.getNeb(u).tolist() for u in range(q.getVertexCnt())]


def core_numbers(q):
    """Core number of each vertex: the largest k such that the vertex survives repeated
    removal of vertices with fewer than k remaining neighbours."""
    n = q.getVertexCnt()
    N = [set(x) for x in qnbrs(q)]
    core = [0] * n
    for k in range(1, n + 1):
        alive = set(range(n))
        changed = True
        while changed:
            changed = False
            for v in list(alive):
                if len(N[v] & alive) < k:
                    alive.discard(v)
                    changed = True
        for v in alive:
            core[v] = k
    return core


def bfs_tree(q, s):
    N = qnbrs(q)
    order, parent, level, children = [s], {s: 0}, {s: 0}, {v: [] for v in range(q.getVertexCnt())}
    dq = deque([s])
    seen = {s}
    while dq:
        u = dq.popleft()
        for w in N[u]:
            if w not in seen:
                seen.add(w)
                dq.append(w)
                order.append(w)
                parent[w] = u
                level[w] = level[u] + 1
                children[u].append(w)
    return order, parent, level, children


def prev_follow(q, order):
    pos = {v: i for i, v in enumerate(order)}
    N = qnbrs(q)
    prev = {v: sorted([w for w in N[v] if pos[w] < pos[v]], key=pos.get) for v in order}
    fol = {v: sorted([w for w in N[v] if pos[w] > pos[v]], key=pos.get) for v in order}
    return prev, fol


def label_freq(g, label):
    lab = g.label_array().tolist()
    return sum(1 for x in lab if x == label)


def max_label_freq(g):
    lab = g.label_array().tolist()
    return max(lab.count(x) for x in set(lab))


def cfl_start(g, q):
    n = q.getVertexCnt()
    core = core_numbers(q)
    klen = sum(1 for c in core if c > 1)
    ql = q.label_array().tolist()
    cand = [u for u in range(n) if klen == 0 or core[u] > 1]
    keys = [label_freq(g, ql[u]) / q.getVertexDegree(u) for u in cand]
    visit = O.heap_pop_order(keys, 3, use_std=True)  # std::priority_queue's own tie order
    nlf = nlf_by_definition(g, q)
    best, start = float(max_label_freq(g) + 1), 0
    for i in visit:
        u = cand[i]
        s = len(nlf[u]) / q.getVertexDegree(u)
        if s < best:
            best, start = s, u
    return start


def ceci_daf_start(g, q, daf):
    nlf = nlf_by_definition(g, q)
    best, start = float(g.getVertexCnt()), 0
    for u in range(q.getVertexCnt()):
        d = q.getVertexDegree(u)
        if daf and d <= 1:
            continue
        if len(nlf[u]) / d < best:
            best, start = len(nlf[u]) / d, u
    return start


def _keep(A, cand_u, pivots, C):
    return {v for v in cand_u if all(A[v] & C[p] for p in pivots)}


def cfl_sets(g, q):
    A = adjsets(g)
    nlf = [set(x) for x in nlf_by_definition(g, q)]
    order, parent, level, children = bfs_tree(q, cfl_start(g, q))
    prev, fol = prev_follow(q, order)
    C = {order[0]: set(nlf[order[0]])}
    for u in order[1:]:
        C[u] = set()
    for lv in range(1, max(level.values()) + 1):
        here = [u for u in order if level[u] == lv]
        for u in here:
            C[u] = _keep(A, nlf[u], prev[u], C)
        for u in reversed(here):
            same = [w for w in fol[u] if level[w] == lv]
            if same:
                C[u] = _keep(A, C[u], same, C)
    for lv in range(max(level.values()) - 1, -1, -1):
        for u in [x for x in order if level[x] == lv]:
            deeper = [w for w in fol[u] if level[w] > lv]
            if deeper:
                C[u] = _keep(A, C[u], deeper, C)
    out = [sorted(C[u]) for u in range(q.getVertexCnt())]
    return all(out), out


def daf_sets(g, q, passes=3):
    A = adjsets(g)
    nlf = [set(x) for x in nlf_by_definition(g, q)]
    if not all(nlf):
        return False, None
    order, parent, level, children = bfs_tree(q, ceci_daf_start(g, q, True))
    prev, fol = prev_follow(q, order)
    C = {u: set(nlf[u]) for u in range(q.getVertexCnt())}
    for k in range(passes):
        seq = order[1:] if k % 2 == 0 else list(reversed(order[:-1]))
        for u in seq:
            C[u] = _keep(A, C[u], prev[u] if k % 2 == 0 else fol[u], C)
    out = [sorted(C[u]) for u in range(q.getVertexCnt())]
    return all(out), out


def ceci_sets(g, q):
    """CECI over sets: C(u) = NLF(u) vertices next to a live candidate of u's tree parent;
    a parent candidate without an NLF(u) neighbour dies; finally a candidate of x dies if
    some later neighbour u of x has no NLF(u) vertex next to it."""
    A = adjsets(g)
    nlf = [set(x) for x in nlf_by_definition(g, q)]
    order, parent, level, children = bfs_tree(q, ceci_daf_start(g, q, False))
    prev, fol = prev_follow(q, order)
    C = {order[0]: set(nlf[order[0]])}
    if not C[order[0]]:
        return False, None
    for u in order[1:]:
        p = parent[u]
        C[u] = {v for v in nlf[u] if A[v] & C[p]}
        C[p] = {v for v in C[p] if A[v] & nlf[u]}
        if not C[u]:
            return False, None
    for x in order:
        C[x] = {v for v in C[x] if all(A[v] & nlf[u] for u in fol[x])}
    out = [sorted(C[u]) for u in range(q.getVertexCnt())]
    return all(out), out


# ------------------------------------------------------------------------------ orders
def edge_counts(g, q, cands):
    A = adjsets(g)
    N = qnbrs(q)
    return {(a, b): sum(len(A[v] & set(cands[b])) for v in cands[a]) for a in range(q.getVertexCnt()) for b in N[a]}


def qsi_order(g, q, cands):
    n = q.getVertexCnt()
    e = edge_counts(g, q, cands)
    deg = [q.getVertexDegree(u) for u in range(n)]
    N = qnbrs(q)
    # start edge: fewest entries, then smaller degree sum, first seen; higher-degree end first
    best = None
    for a in range(n):
        for b in N[a]:
            key = (e[a, b], deg[a] + deg[b])
            if best is None or key < best[0]:
                best = (key, (b, a) if deg[a] < deg[b] else (a, b))
    order = list(best[1])
    vis = set(order)
    while len(order) < n:
        pick = None
        for a in range(n):
            if a not in vis:
                continue
            for b in N[a]:
                if b in vis:
                    continue
                adjs = sum(1 for w in N[b] if w in vis)
                cur = (e[a, b], adjs, -deg[b])
                # strictly better: fewer entries; then fewer visited neighbours; then larger degree
                if pick is None or cur < pick[0]:
                    pick = (cur, b)
        order.append(pick[1])
        vis.add(pick[1])
    return order


def tables(g, q, cands):
    A = adjsets(g)
    N = qnbrs(q)
    T = {}
    for a in range(q.getVertexCnt()):
        for b in N[a]:
            pos = {v: k for k, v in enumerate(cands[b])}
            T[a, b] = [sorted(pos[w] for w in A[v] if w in pos) for v in cands[a]]
    return T


def gup_order(g, q, cands):
    n = q.getVertexCnt()
    N = qnbrs(q)
    deg = [len(x) for x in N]
    if max(deg) <= 2:
        return None
    core = core_numbers(q)
    nlf = nlf_by_definition(g, q)
    nlfcnt = []
    for u in range(n):  # _NLF_Filter stops at the first empty set
        nlfcnt.append(len(nlf[u]))
        if not nlf[u]:
            nlfcnt += [0] * (n - u - 1)
            break
    maxk = max(core)
    cand = [u for u in range(n) if core[u] == maxk]
    root = cand[0]
    for k in range(maxk, 0, -1):
        ps = []
        for u in cand:
            cd = sum(1 for w in N[u] if core[w] >= k)
            if cd:
                ps.append((u, nlfcnt[u] / cd))
        if not ps:
            continue
        minp = min(p for _, p in ps)
        mins = [u for u, p in ps if abs(p - minp) <= 1e-12 * max(1.0, abs(p), abs(minp))]
        if len(mins) == 1 or k == 1:
            root = mins[0]
            break
    bfs_order, parent, level, children = bfs_tree(q, root)
    cnt = [len(c) for c in cands]
    e = edge_counts(g, q, cands)
    w = {}
    for v in range(n):
        for x in N[v]:
            if core[v] >= 2 and core[x] >= 2 and bfs_order[v] < bfs_order[x]:
                w[v, x] = w[x, v] = e[v, x] / cnt[v] / cnt[x]
    order, ordered = [root], {root}
    queued, sel = set(), [1.0] * n
    heap = []
    v = root
    while True:
        for x in N[v]:
            if core[x] >= 2 and x not in ordered:
                queued.add(x)
                sel[x] *= w.get((v, x), 0.0)
        for x in N[v]:
            if core[x] >= 2 and x not in ordered:
                c = sel[x] * cnt[x]
                for y in N[x]:
                    if y in queued and y not in ordered:
                        c *= w.get((x, y), 0.0) * sel[y]
                heapq.heappush(heap, (c, x))
        nxt = None
        while heap:
            _, v = heapq.heappop(heap)
            if v not in ordered:
                nxt = v
                break
        if nxt is None:
            break
        order.append(nxt)
        ordered.add(nxt)
    T = tables(g, q, cands)
    nembeds = [0.0] * n
    M = 1 << 64

    def count_embed(u):
        npaths = [1] * cnt[u]
        for c in N[u]:
            if c in ordered or parent.get(c) != u or c == root:
                continue
            sub = count_embed(c)
            npsum = 0
            for iv in range(cnt[u]):
                s = sum(sub[k] for k in T[u, c][iv]) % M
                npaths[iv] = (npaths[iv] * s) % M
                npsum = (npsum + s) % M
            nembeds[c] = float(npsum) / cnt[u]  # (double)npsum first, as in C
        return npaths

    heap = []
    for u in range(n):
        if u in ordered and any(x not in ordered for x in N[u]):
            count_embed(u)
            for x in N[u]:
                if x not in ordered:
                    heapq.heappush(heap, (0, nembeds[x], x))
    while heap:
        d, _, u = heapq.heappop(heap)
        assert u not in ordered
        order.append(u)
        ordered.add(u)
        for x in N[u]:
            if x not in ordered:
                heapq.heappush(heap, (d - 1, nembeds[x], x))
    return order


def cfl_order(g, q, cands):
    n = q.getVertexCnt()
    N = qnbrs(q)
    deg = [len(x) for x in N]
    core = core_numbers(q)
    cnt = [len(c) for c in cands]
    T = tables(g, q, cands)
    order_b, parent, level, children = bfs_tree(q, cfl_start(g, q))
    root = order_b[0]

    def core_paths(u, cur, out):
        cur = cur + [u]
        kids = [c for c in children[u] if core[c] > 1]
        for c in kids:
            core_paths(c, cur, out)
        if not kids:
            out.append(cur)

    def tree_paths(u, cur, out):
        cur = cur + [u]
        kids = [c for c in children[u] if deg[c] > 1]
        for c in kids:
            tree_paths(c, cur, out)
        if not kids and len(cur) > 1:
            out.append(cur)

    def estimate(path):
        M = 1 << 64
        a, b = path[-2], path[-1]
        child = [len(r) for r in T[a, b]]
        est = [0] * (len(path) - 1)
        est[-1] = sum(child) % M
        for i in range(len(path) - 2, 0, -1):
            s, e = path[i - 1], path[i]
            par = [sum(child[k] for k in row) % M for row in T[s, e]]
            est[i - 1] = sum(par) % M
            child = par
        return est

    cps, forests = [], []
    if core[root] > 1:
        core_paths(root, [], cps)
        for u in range(n):
            if core[u] > 1:
                tp = []
                tree_paths(u, [], tp)
                if tp:
                    forests.append(tp)
    else:
        tp = []
        tree_paths(root, [], tp)
        if tp:
            forests.append(tp)
    order, vis = [root], {root}

    def root_idx(path, first_zero_ok):
        for j, v in enumerate(path):
            if v not in vis:
                return (j if j == 0 else j - 1) if first_zero_ok else j - 1
        return 0

    if cps:
        ests = [estimate(p) for p in cps]
        ntree = []
        for p in cps:
            nt = deg[p[0]] - len(children[p[0]]) + sum(deg[v] - len(children[v]) - 1 for v in p[1:])
            ntree.append(nt + 1)
        sel = min(range(len(cps)), key=lambda i: (float(ests[i][0]) / ntree[i], i))
        while True:
            for v in cps[sel][1:]:
                if v not in vis:
                    order.append(v)
                    vis.add(v)
            del cps[sel], ests[sel]
            if not cps:
                break
            vals = []
            for i, p in enumerate(cps):
                r = root_idx(p, False)
                vals.append((float(ests[i][r]) / cnt[p[r]], i))
            sel = min(vals)[1]
    for tp in forests:
        ests = [estimate(p) for p in tp]
        while tp:
            vals = []
            for i, p in enumerate(tp):
                r = root_idx(p, True)
                vals.append((float(ests[i][r]) / cnt[p[r]], i))
            sel = min(vals)[1]
            for v in tp[sel]:
                if v not in vis:
                    order.append(v)
                    vis.add(v)
            del tp[sel], ests[sel]
    leaves = [u for u in range(n) if deg[u] == 1]
    while leaves:
        sel = min(range(len(leaves)), key=lambda i: (cnt[leaves[i]], i))
        if leaves[sel] not in vis:
            order.append(leaves[sel])
            vis.add(leaves[sel])
        del leaves[sel]
    return order
