"""Graph input: the reference text format and the binary CSR cache beside it.

The text loader follows the reference's parse rules (inc/graph.h:190-266); the binary CSR
(subgraphmatching_amd.graph.save_csr / load_csr, and Graph::LoadCSR in include/sgm/graph.h)
has no reference counterpart and is checked here against the text path it caches: the same
CSR, the same SHA-256, and damaged files rejected instead of half-read."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from subgraphmatching_amd.generators import chung_lu, rmat, with_labels
from subgraphmatching_amd.graph import (CSR_MAGIC, Graph, Query, load_cached, load_csr, load_text, save_csr,
                                        save_text)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "subgraphmatching_amd")


def _same(a, b):
    return (a.getVertexCnt() == b.getVertexCnt() and np.array_equal(a.off, b.off)
            and np.array_equal(a.adj, b.adj))


@pytest.mark.parametrize("make", [lambda: chung_lu(2000, 2.5, 7.4, seed=3), lambda: rmat(11, 8, seed=4)])
def test_text_csr_round_trip(tmp_path, make):
    g = make()
    tp = str(tmp_path / "g.txt")
    save_text(g, tp)
    gt = load_text(tp)
    assert _same(g, gt)
    cp = str(tmp_path / "g.csr")
    m = save_csr(gt, cp, {"generator": "test"})
    assert m["sha256"] == g.digest() and m["vertices"] == g.getVertexCnt() and m["edges"] == g.getEdgeCnt()
    gc = load_csr(cp)
    assert _same(g, gc) and gc.digest() == g.digest()
    # fixed layout: 32-byte header, then off and adj as little-endian u32
    raw = open(cp, "rb").read()
    assert raw[:8] == CSR_MAGIC
    assert len(raw) == 32 + 4 * (g.getVertexCnt() + 1) + 4 * g.adj.size
    assert np.array_equal(np.frombuffer(raw[32:32 + 4 * (g.getVertexCnt() + 1)], dtype="<u4"), g.off)


def test_empty_and_isolated_vertices(tmp_path):
    g = Graph.from_edges(5, np.zeros((0, 2), dtype=np.int64))
    cp = str(tmp_path / "e.csr")
    save_csr(g, cp)
    ge = load_csr(cp)
    assert ge.getVertexCnt() == 5 and ge.getEdgeCnt() == 0
    q = Query.from_edges(3, [(0, 1), (1, 2)])
    save_csr(q, cp)
    assert isinstance(load_csr(cp, cls=Query), Query)


def _damaged(tmp_path, mutate):
    g = chung_lu(500, 2.5, 6.0, seed=5)
    cp = str(tmp_path / "d.csr")
    save_csr(g, cp)
    raw = bytearray(open(cp, "rb").read())
    raw = mutate(raw, g)
    open(cp, "wb").write(bytes(raw))
    return cp


def test_damaged_files_are_rejected(tmp_path):
    V_OFF = 32  # off[0]
    cases = {
        "magic": lambda r, g: b"XGMCSR01" + r[8:],
        "truncated": lambda r, g: r[:-4],
        # a flipped neighbour id changes the digest (or breaks sortedness/range)
        "digest": lambda r, g: r[:-4] + (int.from_bytes(r[-4:], "little") ^ 1).to_bytes(4, "little"),
        "offsets": lambda r, g: r[:V_OFF] + (7).to_bytes(4, "little") + r[V_OFF + 4:],
    }
    for name, mut in cases.items():
        cp = _damaged(tmp_path, mut)
        with pytest.raises(ValueError):
            load_csr(cp)
    # without a manifest the structural checks still catch unsorted lists
    g = chung_lu(500, 2.5, 6.0, seed=5)
    cp = str(tmp_path / "u.csr")
    save_csr(g, cp)
    os.remove(cp + ".json")
    raw = bytearray(open(cp, "rb").read())
    V = g.getVertexCnt()
    v = int(np.argmax(g.degrees()))
    p = 32 + 4 * (V + 1) + 4 * int(g.off[v])
    a, b = raw[p:p + 4], raw[p + 4:p + 8]
    raw[p:p + 4], raw[p + 4:p + 8] = b, a
    open(cp, "wb").write(bytes(raw))
    with pytest.raises(ValueError):
        load_csr(cp)


def test_labelled_round_trip(tmp_path):
    g = with_labels(chung_lu(800, 2.5, 6.0, seed=7), 5, seed=8, skew=1.0)
    assert g.is_labelled()
    tp, cp = str(tmp_path / "l.txt"), str(tmp_path / "l.csr")
    save_text(g, tp)
    gt = load_text(tp)
    assert _same(g, gt) and np.array_equal(gt.labels, g.labels) and gt.digest() == g.digest()
    m = save_csr(gt, cp)
    raw = open(cp, "rb").read()
    V = g.getVertexCnt()
    assert int.from_bytes(raw[24:32], "little") == 1
    assert np.array_equal(np.frombuffer(raw[-4 * V:], dtype="<u4"), g.labels)
    gc = load_csr(cp)
    assert np.array_equal(gc.labels, g.labels) and gc.digest() == m["sha256"]
    # labels are part of the digest: a changed label is caught
    bad = bytearray(raw)
    bad[-4] ^= 1
    open(cp, "wb").write(bytes(bad))
    with pytest.raises(ValueError):
        load_csr(cp)
    # all-zero labels are the unlabeled graph
    z = Graph(g.off, g.adj, labels=np.zeros(V, dtype=np.uint32))
    assert not z.is_labelled() and z.digest() == chung_lu(800, 2.5, 6.0, seed=7).digest()


def test_load_cached_writes_then_reuses(tmp_path):
    g = chung_lu(1500, 2.5, 7.4, seed=6)
    tp = str(tmp_path / "c.txt")
    save_text(g, tp)
    g1 = load_cached(tp)
    cp = tp + ".csr"
    assert os.path.exists(cp) and os.path.exists(cp + ".json")
    m1 = os.path.getmtime(cp)
    g2 = load_cached(tp)
    assert _same(g1, g2) and _same(g, g2) and os.path.getmtime(cp) == m1
    # a damaged cache is rebuilt from the text file
    with open(cp, "r+b") as f:
        f.write(b"garbage!")
    g3 = load_cached(tp)
    assert _same(g, g3) and open(cp, "rb").read(8) == CSR_MAGIC
    # a text file replaced by another graph with an older modification time (cp -p, tar,
    # rsync -t) is noticed by its digest, not trusted by its mtime
    h = chung_lu(1500, 2.5, 7.4, seed=8)
    save_text(h, tp)
    old = os.path.getmtime(cp) - 3600
    os.utime(tp, (old, old))
    g4 = load_cached(tp)
    assert _same(h, g4) and not _same(g, g4)


CPP = r'''
#include <cstdio>
#include "sgm/graph.h"
static unsigned long long fold(const Graph& g) {
    unsigned long long h = 1469598103934665603ull;
    for (uint32_t v = 0; v < g.getVertexCnt(); v++) {
        uint32_t n; const VertexID* nb = g.getNeb(v, n);
        h = (h ^ n) * 1099511628211ull;
        for (uint32_t i = 0; i < n; i++) h = (h ^ nb[i]) * 1099511628211ull;
    }
    return h;
}
static unsigned long long labs(const Graph& g) {
    unsigned long long h = 0;
    for (uint32_t v = 0; v < g.getVertexCnt(); v++) h = h * 31 + g.getVertexLabel(v);
    return h;
}
int main(int argc, char** argv) {
    Graph a, b;
    a.Load(argv[1]);
    b.LoadCSR(argv[2]);
    std::printf("%u %u %u %llu %llu\n", a.getVertexCnt(), a.getEdgeCnt(), a.getMaxDgree(), fold(a), labs(a));
    std::printf("%u %u %u %llu %llu\n", b.getVertexCnt(), b.getEdgeCnt(), b.getMaxDgree(), fold(b), labs(b));
    try { Graph c; c.LoadCSR(argv[3]); std::printf("accepted\n"); }
    catch (const std::exception& e) { std::printf("rejected\n"); }
    return 0;
}
'''


def test_cpp_loadcsr_matches_text_load(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    lib = os.path.join(PKG, "libsgm_hip.so")
    if not os.path.exists(lib):
        pytest.skip("libsgm_hip.so not built")
    g = with_labels(rmat(10, 8, seed=9), 4, seed=10)
    tp, cp, bad = str(tmp_path / "g.txt"), str(tmp_path / "g.csr"), str(tmp_path / "bad.csr")
    save_text(g, tp)
    save_csr(g, cp)
    raw = open(cp, "rb").read()
    open(bad, "wb").write(raw[:-8])
    src = tmp_path / "io.cpp"
    src.write_text(CPP)
    exe = tmp_path / "io"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        "-L", PKG, "-lsgm_hip", "-Wl,-rpath," + PKG, "-pthread"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe), tp, cp, bad], check=True, capture_output=True, text=True).stdout.split("\n")
    assert out[0] == out[1]
    V, E, D, _, h = out[0].split()
    assert int(h) != 0
    assert (int(V), int(E), int(D)) == (g.getVertexCnt(), g.getEdgeCnt(), g.getMaxDgree())
    assert out[2] == "rejected"
