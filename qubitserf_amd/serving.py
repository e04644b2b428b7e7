"""Serving loop over the batched engine: mixed distance requests arrive as a Poisson
process, the server admits what has arrived (up to a device-memory budget and a batch
size), solves it in one batched call, and records each request's latency.

The arrival clock is virtual and the service is real: arrivals are drawn in advance; when
the server is idle it jumps the clock to the next arrival, takes every request that has
arrived by then, runs the batch on the device and advances the clock by the batch's
measured wall time. Latencies and device load are thus those of the real engine under that
arrival process, reproducible from the seed. The rate is tuned so that the device is busy
for the target fraction of the time (busy = sum of batch times / makespan).

Request mix: half classical codes (n ~ U{64..256}, rate ~ U[0.3, 0.6], Bernoulli(1/2) H,
full rank), half bivariate-bicycle CSS codes (l in {6, 9, 12, 15}, m = 6, A = x^3 + y + y^2,
B = y^3 + x + x^2), min distance over both sides. Each request carries a work budget in
codewords; a request that would exceed it returns its bounds unproven.
"""
from __future__ import annotations

import time

import numpy as np

from . import codes
from .distfind import api as df


def make_requests(count, seed=0, classical_cap=int(1e10), css_cap=int(1e13)):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        if rng.random() < 0.5:
            H = codes.random_rate_code(rng, 64, 256)
            out.append(dict(kind="classical", n=H.shape[1], problem=("classical", H), work_cap=classical_cap))
        else:
            l = int(rng.choice([6, 9, 12, 15]))
            Hx, Hz = codes.bb_family(l)
            out.append(dict(kind=f"bb{l}", n=Hx.shape[1], problem=("css", Hx, Hz, "min"), work_cap=css_cap))
    return out


def device_bytes(req):
    """Device memory a request's solve holds (generators, residual tables, cached plans):
    a bound from n alone, K <= n, up to n / 8 information sets."""
    n = req["n"]
    st = (n + 63) // 64
    sets = max(2, n // 8)
    return sets * n * (st * 8 + 8) + 64 * 1024


def serve(requests, rate, seed=1, max_batch=64, mem_budget=1 << 30):
    """Run the loop; returns per-request latencies and load statistics."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(requests)))
    clock = 0.0
    busy = 0.0
    nxt = 0
    queue = []
    lat = np.zeros(len(requests))
    proven = np.zeros(len(requests), bool)
    batches = []
    while nxt < len(requests) or queue:
        if not queue:
            clock = max(clock, arrivals[nxt])
        while nxt < len(requests) and arrivals[nxt] <= clock:
            queue.append(nxt)
            nxt += 1
        take, mem = [], 0
        for r in queue:   # admission: first come, within the batch size and memory budget
            b = device_bytes(requests[r])
            if take and (len(take) >= max_batch or mem + b > mem_budget):
                break
            take.append(r)
            mem += b
        queue = queue[len(take):]
        t0 = time.perf_counter()
        res = df.batch_distance([{"problem": requests[r]["problem"], "work_cap": requests[r]["work_cap"]} for r in take])
        dt = time.perf_counter() - t0
        clock += dt
        busy += dt
        batches.append((len(take), dt))
        for r, x in zip(take, res):
            lat[r] = clock - arrivals[r]
            proven[r] = bool(x.proven)
    makespan = clock - arrivals[0] if len(requests) else 0.0
    return dict(p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                mean_ms=float(lat.mean() * 1e3), throughput_rps=len(requests) / max(makespan, 1e-9),
                busy=busy / max(makespan, 1e-9), batches=len(batches),
                mean_batch=float(np.mean([b for b, _ in batches])) if batches else 0.0,
                proven=int(proven.sum()), requests=len(requests), rate_rps=rate)


def calibrate_rate(requests, target_busy=0.7, seed=1, rounds=3):
    """Arrival rate for the target device load: start from the measured mean service time
    of single requests, then rescale by the observed load."""
    sample = requests[: min(16, len(requests))]
    t0 = time.perf_counter()
    for r in sample:
        df.batch_distance([{"problem": r["problem"], "work_cap": r["work_cap"]}])
    per = (time.perf_counter() - t0) / len(sample)
    rate = target_busy / per
    best = None
    for _ in range(rounds):
        stats = serve(requests, rate, seed)
        if best is None or abs(stats["busy"] - target_busy) < abs(best[1]["busy"] - target_busy):
            best = (rate, stats)
        if abs(stats["busy"] - target_busy) < 0.05:
            break
        # busy = rate x service / (1 + ...): damped step, since batching shortens service as load grows
        rate *= (target_busy / max(stats["busy"], 1e-3)) ** 1.5
    return best
