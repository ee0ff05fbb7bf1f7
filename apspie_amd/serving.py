"""Served-workload driver: a device-resident pool of seeded RMAT graphs, a
Poisson stream of BFS / SSSP requests, and the same stream through plain
one-at-a-time launches for comparison.

Workload (all draws from one MT19937 stream, seed 5555):
  pool     16 directed RMAT graphs, scales 16,16,17,17,18,18,19,19,20,20,21,21,
           22,22,22,22, edge factor 16, weights 1..64 (seed 5000 + 100*i +
           scale, weight seed + 1,000,000)
  request  kind = BFS if draw < 2^31 else SSSP; graph = (draw * 16) >> 32;
           source = (draw * m) >> 32; inter-arrival = -ln(1 - draw / 2^32) / rate
"""
from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import REQ_BFS, REQ_SSSP, Context, DeviceCsr, HostCsr, Server, now, rmat

POOL_SCALES = [16, 16, 17, 17, 18, 18, 19, 19, 20, 20, 21, 21, 22, 22, 22, 22]


@dataclass
class Req:
    kind: int
    graph: int
    src: int
    arrival: float  # seconds from stream start


def mt19937_stream(seed: int, n: int) -> np.ndarray:
    """n genrand_int32 draws for init_genrand(seed) (numpy's RandomState uses
    the same seeding and generator)."""
    return np.random.RandomState(seed).randint(0, 2**32, size=n, dtype=np.uint64).astype(np.uint64)


def build_pool(scales: List[int] = POOL_SCALES) -> List[HostCsr]:
    pool = []
    for i, s in enumerate(scales):
        seed = 5000 + 100 * i + s
        pool.append(rmat(s, 16, seed, undirected=False, weight_seed=seed + 1_000_000))
    return pool


def make_stream(pool_m: List[int], n: int = 20000, rate: float = 2000.0, seed: int = 5555) -> List[Req]:
    d = mt19937_stream(seed, 4 * n).reshape(n, 4)
    reqs, t = [], 0.0
    for i in range(n):
        kind = REQ_BFS if int(d[i, 0]) < 2**31 else REQ_SSSP
        g = int((int(d[i, 1]) * len(pool_m)) >> 32)
        src = int((int(d[i, 2]) * pool_m[g]) >> 32)
        t += -math.log(1.0 - int(d[i, 3]) / 2**32) / rate
        reqs.append(Req(kind, g, src, t))
    return reqs


def run_served(server: Server, reqs: List[Req], open_loop: bool, outs: Optional[dict] = None) -> dict:
    """Submit the stream (at its arrival times when open_loop, else all at
    once), wait for every ticket, return the server's report."""
    server.reset_stats()
    tickets = [0] * len(reqs)
    t0 = now() + 0.05

    def submitter():
        delay = t0 - now()  # the stream starts at t0 in both modes
        if delay > 0:
            time.sleep(delay)
        for i, r in enumerate(reqs):
            arr = t0 + r.arrival if open_loop else t0
            if open_loop:
                delay = arr - now()
                if delay > 0:
                    time.sleep(delay)
            out = outs.get(i) if outs else None
            tickets[i] = server.submit(r.kind, r.graph, r.src, out, arr)

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    server.drain()
    rep = server.report()
    rep["mode"] = "open-loop" if open_loop else "saturation"
    return rep


def run_plain(ctx: Context, graphs: List[DeviceCsr], reqs: List[Req], open_loop: bool,
              degrees: Optional[List[np.ndarray]] = None) -> dict:
    """The same requests one at a time through apspie_bfs / apspie_sssp.
    With `degrees` (each graph's out-degrees), a request whose source has no
    out-edges is answered without a launch, as the server answers it on the
    host, so the two paths do the same work per request."""
    bufs = [ctx.alloc(4 * g.m) for g in graphs]
    lat = []
    skipped = 0
    t0 = now() + 0.05
    for r in reqs:
        arr = t0 + r.arrival if open_loop else t0
        if open_loop:
            delay = arr - now()
            if delay > 0:
                time.sleep(delay)
        g = graphs[r.graph]
        if degrees is not None and degrees[r.graph][r.src] == 0:
            skipped += 1
        elif r.kind == REQ_BFS:
            ctx.bfs_into(g, r.src, bufs[r.graph])
        else:
            ctx.sssp_into(g, r.src, bufs[r.graph])
        lat.append((now() - arr) * 1e3)
    t_end = now()
    for b in bufs:
        b.free()
    lat = np.sort(np.array(lat))
    first = t0 + (reqs[0].arrival if open_loop else 0.0)
    return {
        "mode": "open-loop" if open_loop else "saturation",
        "completed": len(reqs),
        "answered_without_launch": skipped,
        "p50_ms": float(lat[int(0.50 * (len(lat) - 1) + 0.5)]),
        "p99_ms": float(lat[int(0.99 * (len(lat) - 1) + 0.5)]),
        "qps": len(reqs) / (t_end - first),
    }
