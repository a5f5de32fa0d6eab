"""Learned-order cost model.

CPU: the seeded order sampler gives connected permutations deterministically, and the
model's input features of a hand-checked query are what their definition says.
GPU: the BF16-MFMA forward pass and gradients equal a plain PyTorch fp32 restatement of
the same network with the same parameters (forward within BF16 rounding; gradients by
cosine similarity and relative norm), and training lowers the loss."""
import math

import numpy as np
import pytest

from brute import query_from_edges, random_connected_query
from subgraphmatching_amd import _lib as L
from subgraphmatching_amd.learned import CostModel, sample_order

# parameter layout (sgm_costmodel_host.cpp): W0 32x64, b0, 3 x (W 192x64, b), w_o, b_o
OW0, OB0, OL1, LAYER = 0, 2048, 2112, 192 * 64 + 64
OWO, OBO = OL1 + 3 * LAYER, OL1 + 3 * LAYER + 64


def _connected(q, order):
    seen = {order[0]}
    for u in order[1:]:
        if not any(int(w) in seen for w in q.getNeb(u)):
            return False
        seen.add(u)
    return True


def test_order_sampler_connected_and_seeded():
    rng = np.random.default_rng(5)
    for t in range(50):
        q = random_connected_query(int(rng.integers(2, 33)), int(rng.integers(0, 20)), rng)
        o1, o2 = sample_order(q, 1000 + t), sample_order(q, 1000 + t)
        assert o1 == o2
        assert sorted(o1) == list(range(q.getVertexCnt()))
        assert _connected(q, o1)
    q = random_connected_query(12, 10, rng)
    assert len({tuple(sample_order(q, s)) for s in range(40)}) > 10  # the seed matters


def test_features_of_a_path_query():
    # path 0-1-2-3 in order 1,0,2,3; |C| = 3, 7, 15, 0; degree sums 6, 14, 30, 0
    q = query_from_edges(4, [(0, 1), (1, 2), (2, 3)])
    item = (q, [1, 0, 2, 3], np.array([3, 7, 15, 0]), np.array([6, 14, 30, 0]))
    f, back, fwd = CostModel.features(item)
    # row i = order position; bits are positions
    assert list(back[:4]) == [0, 0b1, 0b1, 0b100] and list(fwd[:4]) == [0b110, 0, 0b1000, 0]
    assert not back[4:].any() and not f[4:].any() and not f[:, 8:].any()
    deg = [2, 1, 2, 1]          # degrees of vertices 1, 0, 2, 3
    cnt = [7, 3, 15, 0]
    mean = [2.0, 2.0, 2.0, 0.0]
    nb = [0, 1, 1, 1]
    for i in range(4):
        exp = [deg[i] / 8, math.log2(1 + cnt[i]) / 16, math.log2(1 + mean[i]) / 16, i / 32,
               nb[i] / deg[i], nb[i] / 8, 1.0 if i == 0 else 0.0, 4 / 32]
        assert np.allclose(f[i, :8], exp, atol=1e-6), (i, f[i, :8], exp)


# ------------------------------------------------------------------------------ GPU
def _items(count, seed):
    rng = np.random.default_rng(seed)
    items = []
    for _ in range(count):
        n = int(rng.integers(2, 33))
        q = random_connected_query(n, int(rng.integers(0, 2 * n)), rng)
        order = sample_order(q, int(rng.integers(1 << 40)))
        counts = rng.integers(1, 1 << 20, size=n)
        degsum = counts * rng.integers(1, 200, size=n)
        items.append((q, order, counts, degsum))
    return items


def _torch_model(params, items, bf16):
    import torch
    P = torch.tensor(params, dtype=torch.float32, requires_grad=True)
    feats, Bm, Fm, valid, nrows = [], [], [], [], []
    for it in items:
        f, b, w = CostModel.features(it)
        feats.append(f)
        bm = np.zeros((32, 32), np.float32)
        fm = np.zeros((32, 32), np.float32)
        for i in range(32):
            jb = [j for j in range(32) if (int(b[i]) >> j) & 1]
            jf = [j for j in range(32) if (int(w[i]) >> j) & 1]
            for j in jb:
                bm[i, j] = 1.0 / len(jb)
            for j in jf:
                fm[i, j] = 1.0 / len(jf)
        Bm.append(bm)
        Fm.append(fm)
        n = it[0].getVertexCnt()
        valid.append([1.0 if i < n else 0.0 for i in range(32)])
        nrows.append(n)
    F = torch.tensor(np.stack(feats))
    Bm, Fm = torch.tensor(np.stack(Bm)), torch.tensor(np.stack(Fm))
    V = torch.tensor(valid).unsqueeze(-1)
    N = torch.tensor(nrows, dtype=torch.float32)
    r = (lambda x: x.to(torch.bfloat16).float()) if bf16 else (lambda x: x)
    W0, b0 = P[OW0:OB0].view(32, 64), P[OB0:OB0 + 64]
    H = torch.relu(r(F) @ r(W0) + b0) * V
    for l in range(3):
        off = OL1 + l * LAYER
        W, b = P[off:off + 192 * 64].view(192, 64), P[off + 192 * 64:off + LAYER]
        Hb = r(H)
        X = torch.cat([Hb, r(Bm @ Hb), r(Fm @ Hb)], dim=-1)
        H = torch.relu(X @ r(W) + b) * V
    y = (H.sum(1) / N.unsqueeze(-1)) @ P[OWO:OBO] + P[OBO]
    return P, y




@pytest.mark.gpu
def test_predict_matches_torch_reference():
    m = CostModel(seed=11)
    items = _items(48, seed=1)
    got = m.predict(items)
    _, ref = _torch_model(m.params(), items, bf16=True)
    ref = ref.detach().numpy()
    print("max abs diff", float(np.abs(got - ref).max()), "range", float(ref.min()), float(ref.max()), flush=True)
    assert np.allclose(got, ref, rtol=1e-2, atol=1e-2)
    # and against the unrounded fp32 network, within BF16 operand error
    _, ref32 = _torch_model(m.params(), items, bf16=False)
    assert np.allclose(got, ref32.detach().numpy(), rtol=5e-2, atol=5e-2)


@pytest.mark.gpu
def test_gradient_matches_torch_autograd():
    import torch
    m = CostModel(seed=12)
    items = _items(40, seed=2)
    targets = np.random.default_rng(3).uniform(5, 25, size=len(items)).astype(np.float32)
    g, loss = m.grad(items, targets)
    P, y = _torch_model(m.params(), items, bf16=False)
    ref_loss = torch.mean((y - torch.tensor(targets)) ** 2)
    ref_loss.backward()
    rg = P.grad.numpy()
    print("loss", loss, float(ref_loss), flush=True)
    assert abs(loss - float(ref_loss)) <= 2e-2 * (1 + float(ref_loss))
    blocks = {"W0": (OW0, OW0 + 8 * 64), "b0": (OB0, OB0 + 64), "wo": (OWO, OBO), "bo": (OBO, OBO + 1)}
    for l in range(3):
        off = OL1 + l * LAYER
        blocks[f"W{l + 1}"] = (off, off + 192 * 64)
        blocks[f"b{l + 1}"] = (off + 192 * 64, off + LAYER)
    for name, (a, b) in blocks.items():
        x, yv = g[a:b].astype(np.float64), rg[a:b].astype(np.float64)
        nx, ny = np.linalg.norm(x), np.linalg.norm(yv)
        cos = float(x @ yv / (nx * ny)) if nx > 0 and ny > 0 else 1.0
        print(name, "cos", cos, "rel", abs(nx - ny) / max(ny, 1e-12), flush=True)
        if ny < 1e-6:
            assert nx < 1e-4, name
            continue
        assert cos > 0.99, name
        assert abs(nx - ny) <= 5e-2 * ny, name
    assert not g[OW0 + 8 * 64:OB0].any()  # zero feature columns get no gradient


@pytest.mark.gpu
def test_training_lowers_loss():
    teacher = CostModel(seed=21)
    items = _items(600, seed=4)
    targets = teacher.predict(items) * 4.0 + 10.0
    m = CostModel(seed=22)
    losses = m.train(items, targets, epochs=40, minibatch=64, lr=3e-3, seed=5)
    print("loss", losses[0], losses[-1], flush=True)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.25 * losses[0]
    t = m.timing()
    assert t["train_ms"] > 0 and t["train_flops"] > 0


@pytest.mark.gpu
def test_device_ranking_matches_host_features_and_keeps_counts():
    import subgraphmatching_amd as sgm
    from subgraphmatching_amd.generators import chung_lu, query_set
    g = chung_lu(3000, 2.3, 8.0, seed=13)
    qs = query_set(g, 24, lambda i, r: int(r.integers(4, 20)), lambda i, r: bool(i % 2), seed=131)
    runner = sgm.Batch(g, slots=24)
    ok, cnt, dsum = runner.filter(qs)
    m = CostModel(seed=31)
    orders, pbest, pri, used = runner.rank_orders(m, n_samples=16, seed=7, margin=-100.0)  # always take the best
    ri_items = [(q, sgm.Order_RI(q), cnt[i], dsum[i]) for i, q in enumerate(qs)]
    host = m.predict(ri_items)
    print("max |device - host| RI score", float(np.abs(host - pri).max()), flush=True)
    assert np.allclose(pri, host, atol=2e-2, rtol=1e-2)  # device features = host features
    assert np.all(pbest <= pri + 1e-6)
    # the chosen order's score is the model's score of that order (inputs, scores and the
    # choice are ordered on one stream)
    best_items = [(q, orders[i], cnt[i], dsum[i]) for i, q in enumerate(qs)]
    assert np.allclose(pbest, m.predict(best_items), atol=2e-2, rtol=1e-2)
    for q, o in zip(qs, orders):
        assert sorted(o) == list(range(q.getVertexCnt())) and _connected(q, o)
    ri_res = runner.enumerate(cap=50000)
    learned_res = runner.enumerate(orders=orders, cap=50000)
    assert [r.ans for r in learned_res] == [r.ans for r in ri_res]
    # with a prohibitive margin every query keeps RI
    orders2, _, _, used2 = runner.rank_orders(m, n_samples=16, seed=7, margin=1e9)
    assert not used2.any() and orders2 == [sgm.Order_RI(q) for q in qs]
