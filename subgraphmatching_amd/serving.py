"""Online serving loop: queries arrive over time and are answered in dynamic batches by the
batched runner against one resident data graph.

Scheduling: whenever the GPU is free, every query that has arrived (up to the runner's slot
count) forms the next batch -- filter (NLF + GQL), order (RI, or the learned cost model with
RI as fallback), enumerate (cache + filter + failing sets, per-query cap).  A query's latency
is its batch's completion time minus its arrival time.  The slot count is admitted by
device memory: per-slot candidate state plus the per-wave DFS stacks of the persistent
enumeration must fit a stated budget.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

__all__ = ["poisson_arrivals", "admit_slots", "ServeResult", "serve"]


def poisson_arrivals(count: int, rate_per_s: float, seed: int) -> np.ndarray:
    """Arrival times (seconds from 0) of a Poisson stream: exponential gaps of mean 1/rate."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.cumsum(rng.exponential(1.0 / rate_per_s, size=count))


# per slot: split block (2 + 2*64 + 4*2048 u32) and chunk records (37 fields x 2048 u64)
SPLIT_SLOT_BYTES = 4 * (2 + 2 * 64 + 4 * 2048) + 8 * 37 * 2048


def admit_slots(num_vertices: int, max_degree: int, mem_budget_bytes: int, waves: int, max_slots: int = 64):
    """Largest slot count whose device state fits the budget together with the DFS stacks.

    Per slot: membership words by id plus the GQL scratch (rank-space membership, touch /
    dirty marks and NLF-test failures, 4 words per data vertex, and three summary words per 4
    vertices) and the last query vertex's candidate bitmap (one bit per data vertex), i.e.
    under 24 bytes per data vertex, plus descriptors and the depth-1 split block and chunk
    records (SPLIT_SLOT_BYTES); per wave: the set-op stack the enumeration reserves
    (min(8 * max_degree + 1024, 2^20) words).  Returns (slots, stack_bytes, per_slot_bytes)."""
    per_slot = 24 * num_vertices + (1 << 14) + SPLIT_SLOT_BYTES
    stack = waves * 4 * min(8 * max_degree + 1024, 1 << 20)
    free = mem_budget_bytes - stack
    if free < per_slot:
        raise ValueError(f"memory budget {mem_budget_bytes} leaves no room for one query slot "
                         f"(stacks {stack} B, slot {per_slot} B)")
    return int(min(max_slots, free // per_slot)), stack, per_slot


@dataclass
class ServeResult:
    completed: int
    elapsed_s: float
    queries_per_s: float
    embeddings_per_s: float
    latency_ms: dict
    batches: int
    mean_batch: float
    counts: list = field(default_factory=list)
    learned_used: int = 0
    device_ms: dict = field(default_factory=dict)


def serve(runner, queries, arrivals, cap: int, order_fn=None, enum_opts=None, clock=time.perf_counter,
          sleep=time.sleep) -> ServeResult:
    """Serve ``queries`` arriving at ``arrivals`` (seconds, ascending) with ``runner`` (a
    ``Batch``).  ``order_fn(runner)`` returns (orders, learned flags) for the loaded batch, or
    None for RI orders."""
    n = len(queries)
    arrivals = np.asarray(arrivals, dtype=np.float64)
    assert arrivals.shape == (n,) and np.all(np.diff(arrivals) >= 0), "arrivals must be ascending"
    opts = dict(enum_opts or {})
    queue = deque()
    lat = np.zeros(n)
    counts = [0] * n
    nxt = done = batches = learned = 0
    filt_ms = enum_ms = 0.0
    emb = 0
    t0 = clock()
    while done < n:
        now = clock() - t0
        while nxt < n and arrivals[nxt] <= now:
            queue.append(nxt)
            nxt += 1
        if not queue:
            sleep(max(0.0, min(arrivals[nxt] - now, 0.001)))
            continue
        ids = [queue.popleft() for _ in range(min(len(queue), runner.slots))]
        part = [queries[i] for i in ids]
        ok, _, _ = runner.filter(part)
        orders = None
        if order_fn is not None:
            orders, flags = order_fn(runner)
            learned += int(np.sum(flags))
        res = runner.enumerate(orders=orders, cap=cap, **opts)
        f_ms, e_ms = runner.last_ms()
        filt_ms += f_ms
        enum_ms += e_ms
        t_done = clock() - t0
        for i, good, r in zip(ids, ok, res):
            lat[i] = t_done - arrivals[i]
            counts[i] = r.ans if good else 0
            emb += counts[i]
        done += len(ids)
        batches += 1
    elapsed = clock() - t0
    span = max(elapsed - float(arrivals[0]), 1e-9)
    return ServeResult(completed=n, elapsed_s=elapsed, queries_per_s=n / span, embeddings_per_s=emb / span,
                       latency_ms={"p50": 1e3 * float(np.percentile(lat, 50)),
                                   "p99": 1e3 * float(np.percentile(lat, 99)),
                                   "mean": 1e3 * float(lat.mean()), "max": 1e3 * float(lat.max())},
                       batches=batches, mean_batch=n / max(batches, 1), counts=counts, learned_used=learned,
                       device_ms={"filter": filt_ms, "enumerate": enum_ms})
