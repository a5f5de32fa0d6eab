"""GPU parity at the R-MAT benchmark workloads' own sizes, against oracle results in
tests/golden/rmat_fixtures.json (written by tests/golden/make_rmat_fixtures.py; the oracle's
GQL takes seconds to minutes per query on these graphs, so it is not rerun here).

  rmat20    R-MAT 2^20, 2*10^7 edges; one replayed batch of 64 queries of the 1,000-query mix
            (every 16th by size, 4..32 vertices, dense and sparse); cap 10^7
  rmat20 learned the same batch with learned orders (a cost model trained here on node counts the
            enumerator logs): counts equal the plain run's and the oracle's
  rmat22    R-MAT 2^22, average degree 32; 16 queries of the serving stream served on a
            Poisson arrival stream; cap 10^6

Candidate sets are compared bit for bit through their SHA-256 digests; counts are
min(#embeddings, cap), which no order changes."""
import json
import os

import numpy as np
import pytest

import oracle as O
import subgraphmatching_amd as sgm
from subgraphmatching_amd.generators import query_set, rmat

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rmat_fixtures.json")


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def fix():
    with open(FIX) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def rmat20_in(fix):
    say("rmat20: building R-MAT 2^20")
    g = rmat(20, 20_000_000, seed=3)
    qs = query_set(g, 1000, lambda i, r: int(r.integers(4, 33)), lambda i, r: i % 2 == 0, seed=33)
    F = fix["rmat20"]
    sub = [qs[r["id"]] for r in F["queries"]]
    runner = sgm.Batch(g, slots=64)
    yield g, qs, sub, F, runner
    runner.close()


def _check_sets(g, sub, recs):
    cans = [sgm.CandidateParam(g, q) for q in sub]
    oks = sgm.filter_batch(cans)
    for can, ok, rec in zip(cans, oks, recs):
        assert ok == rec["ok"], rec["id"]
        assert [int(x) for x in can.candidates_count] == rec["cand_counts"], rec["id"]
        assert O.cand_digest(can.candidates) == rec["cand_digest"], rec["id"]
        assert sgm.Order_RI(can.query) == rec["order"], rec["id"]
        can.close()


def _check_counts(res, ok, recs, cap):
    checked = 0
    for r, good, rec in zip(res, ok, recs):
        assert bool(good) == rec["ok"], rec["id"]
        if rec["count"] is None:  # the oracle ran out of time on this one
            continue
        assert (r.ans if good else 0) == rec["count"], rec["id"]
        assert r.capped == (rec["count"] == cap), rec["id"]
        checked += 1
    return checked


def test_rmat20_candidate_sets(rmat20_in):
    g, qs, sub, F, runner = rmat20_in
    say("rmat20: candidate digests of", len(sub), "queries")
    _check_sets(g, sub, F["queries"])


def test_rmat20_replayed_batch(rmat20_in):
    g, qs, sub, F, runner = rmat20_in
    ok, cnt, _ = runner.filter(sub)
    for i, rec in enumerate(F["queries"]):
        assert list(cnt[i, :rec["n"]]) == rec["cand_counts"], rec["id"]
    for rep in range(2):  # the captured enumeration graph, replayed
        res = runner.enumerate(cap=F["cap"])
        checked = _check_counts(res, ok, F["queries"], F["cap"])
        say("rmat20 replay", rep, "checked", checked, "of", len(sub))
        assert checked >= len(sub) - 4


def test_rmat20_learned_orders_keep_counts(rmat20_in):
    """Learned orders (cost model trained on the enumerator's node counts for a few
    training queries of the same mix) change no count."""
    from subgraphmatching_amd.learned import collect_training, train_model
    g, qs, sub, F, runner = rmat20_in
    train = query_set(g, 128, lambda i, r: int(r.integers(4, 33)), lambda i, r: i % 2 == 0, seed=44)
    items, labels, info = collect_training(runner, train, n_samples=2, cap=F["cap"], seed=1, time_limit_s=5.0,
                                           label="nodes")
    say("rmat20 learned: trained on", len(items), "triples")
    model, losses = train_model(items, labels, seed=7, epochs=10)
    assert np.isfinite(losses).all()
    ok, _, _ = runner.filter(sub)
    orders, pred, pred_ri, used = runner.rank_orders(model, n_samples=32, seed=5, margin=0.5)
    say("rmat20 learned: learned orders used for", int(np.sum(used)), "queries")
    res = runner.enumerate(orders=orders, cap=F["cap"])
    assert _check_counts(res, ok, F["queries"], F["cap"]) >= len(sub) - 4
    ri = runner.enumerate(cap=F["cap"])
    assert [r.ans for r in res] == [r.ans for r in ri]
    model.close()


def test_rmat22_serving(fix):
    from subgraphmatching_amd.serving import poisson_arrivals, serve
    F = fix["rmat22"]
    say("rmat22: building R-MAT 2^22")
    g = rmat(22, 1 << 26, seed=5)
    qs = query_set(g, 1200, lambda i, r: int(r.integers(8, 33)), lambda i, r: i % 2 == 0, seed=55)
    sub = [qs[r["id"]] for r in F["queries"]]
    _check_sets(g, sub, F["queries"])
    runner = sgm.Batch(g, slots=8)
    out = serve(runner, sub, poisson_arrivals(len(sub), 200.0, seed=77), cap=F["cap"])
    say("rmat22: served", out.completed, "in", out.batches, "batches")
    checked = 0
    for c, rec in zip(out.counts, F["queries"]):
        if rec["count"] is None:
            continue
        assert c == rec["count"], rec["id"]
        checked += 1
    assert checked >= len(sub) - 2
    runner.close()
