"""Serving loop over the batched engine: mixed distance requests arrive as a Poisson
process, the server admits what has arrived (within a device-memory budget) into a
continuous batch, and records each request's latency.

Arrivals are drawn in advance from the seed. Two clocks:
* live (the default): the wall clock. A request counts as arrived once its arrival time
  has passed since the start; the server admits arrived requests into a continuous batch
  (distfind.Session), runs one round at a time on the device, and sleeps until the next
  arrival when it has nothing to do. Latency = wall time from arrival to the poll that
  returns the result.
* virtual: the same loop, with the clock advanced by each admission's and round's measured
  wall time and jumped to the next arrival when idle, so the figures are reproducible and
  free of host scheduling noise.
The rate is tuned so that the device is busy for the target fraction of the time
(busy = time spent in admissions and rounds / makespan).

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
    """Device memory a request's solve will hold before it has been built (generators,
    residual tables, cached plans): a bound from n alone, K <= n, up to n / 8 information
    sets. Once built, the session's measured bytes replace it."""
    n = req["n"]
    st = (n + 63) // 64
    sets = max(2, n // 8)
    return sets * n * (st * 8 + 8) + 64 * 1024


def serve(requests, rate, seed=1, max_active=256, mem_budget=1 << 30, keep_results=False, live=True):
    """Run the loop with continuous batching; returns latency and load statistics.

    Between rounds the server admits every request that has arrived, in arrival order,
    while device memory stays within the budget (min(mem_budget, half the device's free
    memory at start, hipMemGetInfo)) and at most `max_active` are in flight. The memory
    in use is measured: the session's solve arenas as the engine holds them
    (Session.device_bytes), plus the estimate of requests admitted in this round whose
    solves are not built yet. A request that alone exceeds the budget is admitted only
    onto an idle device. Each round enumerates one level of every solve in flight; a solve
    leaves as soon as it finishes, so small requests do not wait for large ones.
    keep_results: also return each request's (distance, lower_bound, proven, levels)
    under "results", in request order. live: wall clock (True) or virtual clock."""
    from collections import deque

    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(requests)))
    sess = df.Session()
    free, _total = df.device_memory()
    budget = min(mem_budget, free // 2)
    busy = 0.0
    if live:
        t_start = time.perf_counter()

        def now():
            return time.perf_counter() - t_start

        def spend(dt):   # wall time passes by itself
            pass

        def idle_until(t):
            time.sleep(max(0.0, t - now()))
    else:
        vclock = [0.0]

        def now():
            return vclock[0]

        def spend(dt):
            vclock[0] += dt

        def idle_until(t):
            vclock[0] = max(vclock[0], t)

    nxt = 0
    queue = deque()
    peak = 0
    owner = {}                       # session id -> request index
    lat = np.zeros(len(requests))
    proven = np.zeros(len(requests), bool)
    results = [None] * len(requests)
    done = rounds = admitted_batches = 0
    while done < len(requests):
        while nxt < len(requests) and arrivals[nxt] <= now():
            queue.append(nxt)
            nxt += 1
        take = []
        held = sess.device_bytes      # measured arenas of the solves in flight
        pending = 0                   # estimates of this round's admissions
        while queue and len(owner) + len(take) < max_active:
            b = device_bytes(requests[queue[0]])
            if held + pending + b > budget and (owner or take):
                break
            r = queue.popleft()
            take.append(r)
            pending += b
        if take:
            t0 = time.perf_counter()
            ids = sess.add([{"problem": requests[r]["problem"], "work_cap": requests[r]["work_cap"]} for r in take])
            dt = time.perf_counter() - t0
            spend(dt)
            busy += dt
            admitted_batches += 1
            for i, r in zip(ids, take):
                owner[i] = r
        if sess.active > 0:
            t0 = time.perf_counter()
            sess.step()
            dt = time.perf_counter() - t0
            peak = max(peak, sess.device_bytes)
            spend(dt)
            busy += dt
            rounds += 1
        for i, res in sess.poll():
            r = owner.pop(i)
            lat[r] = now() - arrivals[r]
            proven[r] = bool(res.proven)
            results[r] = (res.distance, res.lower_bound, res.proven, res.levels)
            done += 1
        if not owner and not queue and nxt < len(requests):
            idle_until(arrivals[nxt])   # nothing to do before the next arrival
    end = now()
    sess.close()
    makespan = end - arrivals[0] if len(requests) else 0.0
    out = dict(p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
               mean_ms=float(lat.mean() * 1e3), throughput_rps=len(requests) / max(makespan, 1e-9),
               busy=busy / max(makespan, 1e-9), rounds=rounds, admissions=admitted_batches,
               proven=int(proven.sum()), requests=len(requests), rate_rps=rate,
               device_bytes_peak=int(peak), mem_budget=int(budget), clock="wall" if live else "virtual")
    if keep_results:
        out["results"] = results
    return out


def calibrate_rate(requests, target_busy=0.7, seed=1, rounds=3, live=True):
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
        stats = serve(requests, rate, seed, live=live)
        if best is None or abs(stats["busy"] - target_busy) < abs(best[1]["busy"] - target_busy):
            best = (rate, stats)
        if abs(stats["busy"] - target_busy) < 0.05:
            break
        # busy = rate x service / (1 + ...): damped step, since batching shortens service as load grows
        rate *= (target_busy / max(stats["busy"], 1e-3)) ** 1.5
    return best
