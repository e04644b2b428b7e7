"""Benchmark: exact distance of the [[144,12,12]] gross code (bivariate bicycle l=12, m=6,
A = x^3 + y + y^2, B = y^3 + x + x^2) by Brouwer–Zimmermann on one MI355X.

One step = one complete certified solve, css_distance(Hx, Hz, which="min") through the
distfind C ABI: both CSS sides, every weight level until the lower bound meets the upper
bound (d = 1..10 per side, 2.93e12 codewords). The metric is codewords per second of the
whole solve (the reference's unit of work, C(K,d) x active information sets per level,
summed over levels and sides), with the time-to-certify alongside.

Prints ONE JSON line. Fields beyond the common contract:
  time_to_certify_s   mean wall time of one solve
  roofline            of the enumeration kernel that holds most of the device time, with
                      every kernel's figures under "kernels" (a matrix-test level's products
                      with a list too short for the matrix tiles run on the VALU kernel and
                      count there). bz_mfma (int8 matrix-core hot
                      test): achieved = 64 int8 ops (32 multiply-adds) x its codewords / its
                      device time, peak = 5.0 POP/s dense int8. bz_f4 (fp4 products,
                      32x32x64 / 16x16x128 / two chained 32x32x64): 128 / 256 / 256 ops per
                      codeword, peak 10.1 POP/s dense fp4. bz_tiled (VALU hot test):
                      achieved = sum over levels of (the variant's hot-loop lane-ops per
                      codeword x codewords) / device time (HIP events on its stream), peak =
                      256 CU x 4 SIMD x 16 lanes x 2.4 GHz = 39.3 T lane-ops/s.
                      five_op_model_TOPs restates the rate with 5 ops per codeword (a 64-bit
                      test)
  cpu_baseline        the reference's own CPU enumerator (oracle/_ref) on one level of the
                      same workload, on this host's cores
  config1             time-to-certify of the [64,32] random code of seed 0 on the device and
                      by the reference's CPU bz on this host (median of repeated solves)
  collect             collect-mode throughput (codeaut_bz_collect, gross Z-side generator,
                      subset sizes 1..8, weights 1..12): codewords/s, rows, device sort time
  batched_learned     plain (one code per call, natural information sets) vs batched +
                      learned (one call, trained scorer's proposals) on 256 random codes
                      (n 64..160, rate 0.3..0.6) and 16 quasi-cyclic codes (L 20..40), under
                      a per-code work cap; the elimination kernel's HBM rate; the serving
                      loop's p50/p99 latency and throughput at 70% device load (wall clock;
                      the same rate on the reproducible virtual clock beside it)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12     # int32 lane-ops/s, gfx950
OPS_PER_CODEWORD = 5                                # 2 xor + 2 bcnt + 1 compare (64-bit test)
# Matrix-core hot tests by the trace's `mfma` field: kernel, ops per codeword (2 per
# multiply-add: 32 int8 dimensions, or 64 / 128 fp4 ones), dense MI355X peak in TOP/s.
MATRIX = {1: ("bz_mfma", 64, 5000.0), 2: ("bz_f4_32x32x64", 128, 10100.0), 3: ("bz_f4_16x16x128", 256, 10100.0),
          4: ("bz_f4_2x32x32x64", 256, 10100.0), 5: ("bz_f4_packed_32x32x64", 64, 10100.0)}
TRAFFIC_FILES = {"bz_mfma": ROOT / "profiles" / "bz_mfma_gross_traffic.json",
                 "bz_f4_packed_32x32x64": ROOT / "profiles" / "bz_f4_packed_gross_traffic.json",
                 "bz_tiled": ROOT / "profiles" / "bz_tiled_gross_traffic.json"}


def traffic_bytes(kernel: str):
    """HBM bytes per launch of the kernel's largest gross dispatch, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-kernel scaled), or None."""
    try:
        return json.loads(TRAFFIC_FILES[kernel].read_text())["bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None


def host_cpus():
    """The host's CPUs as this process sees them: os.cpu_count() (nproc), the affinity
    mask, a cgroup CPU quota if one is set, and the CPU model. `usable` = every core the
    process may run on = min(affinity, quota)."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = nproc
    quota = None
    try:
        q, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) / int(period)))
    except (OSError, ValueError):
        pass
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    usable = min(affinity, quota) if quota else affinity
    return {"nproc": nproc, "affinity": affinity, "cgroup_quota": quota, "usable": usable, "model": model}


def cpu_baseline(threads: int):
    """The reference CPU enumerator on the gross Z-side level d=8 (C(78,8) = 2.3e10
    codewords, the full-rank information set, bound 12). Falls back to the oracle port
    when the reference library was not built."""
    from qubitserf_amd import codes
    from tests import oracle as O

    Hx, Hz = codes.gross()
    gamma, ranks, check, _ = O.side_plan(1, Hx, Hz)
    d = 8
    kind = "reference" if O.ref_available() else "port"
    fn = O.ref_enumerate_plan if kind == "reference" else O.enumerate_plan
    t0 = time.perf_counter()
    got = fn(gamma, Hx.shape[1], d, check=check, num_gamma=1, current_best=12, threads=threads)
    dt = time.perf_counter() - t0
    cw = math.comb(gamma.shape[1], d)
    return {
        "value": cw / dt,
        "unit": "codewords/s",
        "cores": threads,
        "kind": kind,
        "sample": f"gross Z side, weight level d={d} of the full-rank information set "
                  f"(C(78,{d}) = {cw:.3e} codewords, bound 12), {dt:.2f} s, result {got}",
    }


def config1_gpu(reps=20):
    """Time-to-certify on the device of the reference's CPU-runnable case: the [64, 32]
    random code of seed 0 (32 x 64 Bernoulli(1/2) parity check, full rank)."""
    from qubitserf_amd import codes
    from qubitserf_amd import distfind as df

    H = codes.config1_parity(0)
    df.classical_distance(H, backend="gpu")
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        d = df.classical_distance(H, backend="gpu")
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return {"code": "[64,32] random, seed 0", "distance": d, "gpu_time_to_certify_s": ts[len(ts) // 2]}


def config1_cpu(threads, reps=5):
    """The reference's CPU bz (oracle/_ref, threads=all given) on the same code; the oracle
    port when the reference library was not built."""
    from qubitserf_amd import codes
    from tests import oracle as O

    H = codes.config1_parity(0)
    kind = "reference" if O.ref_available() else "port"
    fn = (lambda: O.ref_classical_distance(H, threads)) if kind == "reference" else (lambda: O.classical_distance(H, threads))
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        r = fn()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return {"cpu_kind": kind, "cpu_threads": threads, "cpu_distance": int(r.distance), "cpu_time_to_certify_s": ts[len(ts) // 2]}


def _certified(rs):
    return sum(1 for r in rs if r.proven)


def batched_learned(log, c4_cap=int(1e11)):
    """Plain vs batched+learned time-to-certify on the batched (256 random codes) and learned
    (quasi-cyclic evaluation codes) workloads, the elimination phase's HBM rate, and the
    serving loop. Plain = one code per call with the natural information sets; batched +
    learned = one call for the whole set with the trained scorer's proposals (scorer
    inference included). Every code carries the same work cap in both forms; a code over
    the cap is reported unproven (the full sets need ~1e19 codewords). c4_cap: the
    quasi-cyclic codes' cap (1e11 by default; at 1e13 more of them certify and the run
    takes ~12 s)."""
    import numpy as np
    from qubitserf_amd import codes, engine, serving
    from qubitserf_amd import distfind as df
    from qubitserf_amd.proposal import MPScorer, choose_orders

    wfile = ROOT / "qubitserf_amd" / "proposal" / "weights.npz"
    scorer = MPScorer.load(wfile) if wfile.exists() else None
    out = {"scorer": str(wfile.relative_to(ROOT)) if scorer else None}

    def compare(name, Hs, cap, cands):
        # untimed warm-up of each form at its full size, so that every timed run is in the
        # steady state (pinned staging, device pools and solve arenas already grown)
        for H in Hs:
            df.batch_distance([("classical", H)], work_cap=cap)
        df.batch_distance(choose_orders(Hs, scorer, candidates=cands), work_cap=cap)
        df.batch_distance(choose_orders(Hs, None, candidates=cands), work_cap=cap)
        t0 = time.perf_counter()
        plain = [df.batch_distance([("classical", H)], work_cap=cap)[0] for H in Hs]
        t_plain = time.perf_counter() - t0
        t0 = time.perf_counter()
        learned = df.batch_distance(choose_orders(Hs, scorer, candidates=cands), work_cap=cap)
        t_bl = time.perf_counter() - t0
        st = df.last_batch_stats()
        t0 = time.perf_counter()
        rand = df.batch_distance(choose_orders(Hs, None, candidates=cands), work_cap=cap)
        t_rand = time.perf_counter() - t0
        agree = all(a.distance == b.distance for a, b in zip(plain, learned) if a.proven and b.proven)
        r = {"codes": len(Hs), "work_cap": cap, "candidates": cands,
             "plain_sequential_s": t_plain, "batched_learned_s": t_bl, "batched_random_s": t_rand,
             "speedup": t_plain / t_bl, "certified_plain": _certified(plain), "certified_learned": _certified(learned),
             "certified_random": _certified(rand), "distances_agree": agree,
             "levels_plain": sum(x.levels for x in plain), "levels_learned": sum(x.levels for x in learned),
             "batched_rounds": st["rounds"], "batched_codewords": st["codewords"],
             "batched_device_ms": st["device_ms"],
             "note": "every code carries the work cap; a code over it stops unproven, so for "
                     "it the time is time-to-cap, not time-to-certify (certifying every code "
                     "of these sets is infeasible: an estimated 1.6e19 codewords for the 256 "
                     "random codes, 1e13-1e23 per quasi-cyclic code with L >= 28)"}
        log(f"{name}: {json.dumps(r)}")
        return r

    c3 = codes.config3_batch(0, 256)
    out["batched_c3"] = compare("batched", c3, int(1e10), 8)
    c4 = codes.config4_eval(0, 16, 20, 40)
    out["learned_c4"] = compare("learned", c4, int(c4_cap), 16)
    if scorer is not None:
        out["learned_c4"]["proposal_ceiling"] = proposal_ceiling(c4, scorer, 16)
    out["elimination"] = elimination(c3)
    reqs = serving.make_requests(96, seed=0)
    rate, stats = serving.calibrate_rate(reqs, 0.7, rounds=3)   # wall clock
    out["serving"] = stats
    log(f"serving: {json.dumps(stats)}")
    out["serving_virtual_clock"] = serving.serve(reqs, rate, live=False)   # same rate, reproducible clock
    return out


def proposal_ceiling(Hs, scorer, cands, optimum_orders=256, seed=5):
    """Predicted BZ work (set ranks from the device peel, device upper bound) of the natural
    peel, of the best of `cands` orders sampled from the trained scores, of the best of
    `cands` uniform orders, and of the best of `optimum_orders` uniform orders (the
    empirical optimum): geometric-mean ratios over the codes. When uniform proposals
    already reach the optimum, no scorer can gain over them."""
    import numpy as np
    from qubitserf_amd import distfind as df
    from qubitserf_amd.proposal.fit import family_work, sample_orders

    ub = [r.distance for r in df.batch_distance([("classical", H) for H in Hs], work_cap=1)]
    rng = np.random.default_rng(seed)
    orders = []
    for H, s in zip(Hs, scorer.scores(Hs)):
        n = H.shape[1]
        orders.append([np.arange(n, dtype=np.int32)] + sample_orders(s, cands, rng)
                      + sample_orders(np.zeros(n, np.float32), optimum_orders, rng))
    work = family_work(Hs, orders, ub)
    nat = np.array([w[0] for w in work])
    lrn = np.array([min([w[0]] + w[1:1 + cands]) for w in work])
    uni = np.array([min([w[0]] + w[1 + cands:1 + 2 * cands]) for w in work])
    opt = np.array([min(w) for w in work])
    gm = lambda xs: float(np.exp(np.mean(np.log(xs))))
    return {"candidates": cands, "optimum_orders": optimum_orders,
            "gain_over_natural_learned": gm(nat / lrn), "gain_over_natural_uniform": gm(nat / uni),
            "learned_over_optimum": gm(lrn / opt), "uniform_over_optimum": gm(uni / opt),
            "codes_with_deficient_natural_peel": int((nat > opt).sum()),
            "codes_uniform_at_optimum": int((uni == opt).sum()), "codes": len(Hs)}


def collect_figure(p=8, keep=12, reps=3):
    """Collect mode: codeaut_bz_collect of the gross code's Z-side generator in systematic
    form (ker Hx in reduced row echelon form, K = 78, n = 144), subset sizes 1..p, weights
    1..keep: one device pass over the p levels, device sort and dedup."""
    from qubitserf_amd import codes, engine
    from qubitserf_amd.codeaut import bz_collect

    Hx, _ = codes.gross()
    G = codes.gf2_rref(codes.gf2_kernel(Hx))
    bz_collect(G, p, keep)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        rows, combos = bz_collect(G, p, keep)
        ts.append(time.perf_counter() - t0)
    st = engine.last_stats()
    cs = engine.last_collect_stats()
    t = min(ts)
    return {"code": "gross Z side, systematic generator K=78 n=144", "p": p, "keep": keep,
            "codewords": st["codewords"], "combos": combos, "rows": int(rows.shape[0]),
            "appended": cs["appended"], "wall_s": t, "codewords_per_s": st["codewords"] / t,
            "kernel_ms": st["kernel_ms"], "sort_ms": cs["sort_ms"],
            "device_codewords_per_s": st["codewords"] / (st["kernel_ms"] * 1e-3) if st["kernel_ms"] > 0 else None}


def elimination(Hs=None, reps=1):
    """The elimination phase: information sets of the 256 batched-workload codes in one
    info_sets launch. bytes = generators read once + systematic generators written; GB/s
    over the kernel's device time (HIP events), against the 8 TB/s HBM peak."""
    from qubitserf_amd import codes, engine

    Hs = Hs if Hs is not None else codes.config3_batch(0, 256)
    probs = [dict(gen=engine.pack_rows(codes.gf2_kernel(H)), n=H.shape[1], set_cap=8) for H in Hs]
    for _ in range(reps):
        engine.eliminate([dict(p) for p in probs], what=engine.ELIM_SETS)
    es = engine.last_elim_stats()
    gbps = es["sets_bytes"] / (es["sets_ms"] * 1e-3) / 1e9 if es["sets_ms"] > 0 else 0.0
    return {"jobs": es["jobs"], "bytes": es["sets_bytes"], "kernel_ms": es["sets_ms"],
            "GBps": gbps, "peak_GBps": 8000.0, "frac": gbps / 8000.0, "bound": "hbm",
            "traffic": elimination_traffic()}


def elimination_traffic():
    """HBM bytes of the info_sets launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (copy-kernel scaled), or None."""
    try:
        return json.loads((ROOT / "profiles" / "info_sets_c3_traffic.json").read_text())["bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the batched / learned / serving figures")
    ap.add_argument("--c4-cap", type=float, default=1e11, help="work cap per quasi-cyclic code (learned figure)")
    ap.add_argument("--elimination-only", action="store_true",
                    help="only the elimination phase of the 256 batched-workload codes (for counter runs)")
    args = ap.parse_args()
    if args.elimination_only:
        print(json.dumps({"elimination": elimination(reps=3)}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(a level's combination range would be split across GPUs with one 4-byte MIN "
              "all-reduce per level; that path is not built)", file=sys.stderr)
        sys.exit(2)

    from qubitserf_amd import codes
    from qubitserf_amd import distfind as df

    Hx, Hz = codes.gross()
    if "gpu" not in df.available_backends():
        raise SystemExit("gfx950 engine unavailable")

    def solve():
        r = df.api.css_distance_raw(Hx, Hz, "bz", "M", "gpu")
        return r, df.last_trace()

    for _ in range(args.warmup):
        solve()

    walls, codewords, tiled_cw, tiled_ms, tiled_ops = [], 0, 0, 0.0, 0.0
    per = {}          # kernel -> [codewords, device ms, algorithmic ops]
    result = None
    for _ in range(args.steps):
        t0 = time.perf_counter()
        r, trace = solve()
        walls.append(time.perf_counter() - t0)
        result = r
        codewords += sum(t["codewords"] for t in trace)
        for t in trace:
            if t["kernel"] == 2:
                tiled_cw += t["codewords"]
                tiled_ms += t["kernel_ms"]
                tiled_ops += t["ops_per_codeword"] * t["codewords"]
                parts = []
                if t["mfma"]:      # matrix-core hot test: its products are the work
                    name, ops, _ = MATRIX[t["mfma"]]
                    cw, ms = t["codewords"] - t["side_codewords"], t["kernel_ms"] - t["side_ms"]
                    parts.append((name, cw, ms, ops))
                    if t["side_codewords"]:   # its thin products, on the VALU kernel
                        parts.append(("bz_tiled", t["side_codewords"], t["side_ms"], t["side_ops_per_codeword"]))
                else:              # VALU hot test: the variant's hot-loop lane-ops
                    parts.append(("bz_tiled", t["codewords"], t["kernel_ms"], t["ops_per_codeword"]))
                for name, cw, ms, ops in parts:
                    k = per.setdefault(name, [0, 0.0, 0.0])
                    k[0] += cw
                    k[1] += ms
                    k[2] += ops * cw
    total_s = sum(walls)
    value = codewords / total_s
    achieved = tiled_ops / (tiled_ms * 1e-3) / 1e12 if tiled_ms > 0 else 0.0
    five_op = OPS_PER_CODEWORD * tiled_cw / (tiled_ms * 1e-3) / 1e12 if tiled_ms > 0 else 0.0
    peaks = {name: pk for name, _, pk in MATRIX.values()}
    peaks["bz_tiled"] = VALU_PEAK_TOPS
    kernels = {name: {"codewords_per_s": cw / (ms * 1e-3), "ms_per_step": ms / args.steps,
                      "TOPs": o / (ms * 1e-3) / 1e12, "peak": peaks[name],
                      "frac": o / (ms * 1e-3) / 1e12 / peaks[name]}
               for name, (cw, ms, o) in per.items() if ms > 0}
    # the roofline of the kernel that holds most of the device time
    rf_kernel = max(kernels, key=lambda n: kernels[n]["ms_per_step"]) if kernels else "bz_tiled"
    rf = kernels.get(rf_kernel, {"TOPs": 0.0, "peak": VALU_PEAK_TOPS})
    rf_achieved, rf_peak = rf["TOPs"], rf["peak"]
    rf_bound = "valu" if rf_kernel == "bz_tiled" else "mfma"

    line = {
        "metric": "codewords/s + time-to-certify exact d (plain vs batched+learned, 1 GPU)",
        "value": value,
        "unit": "codewords/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total_s / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "css_distance of the [[144,12,12]] gross code (BB l=12 m=6), which=min, "
                               "plain BZ, full certification", "n": 144, "K_per_side": 78,
                   "codewords_per_step": codewords // args.steps},
        "time_to_certify_s": total_s / args.steps,
        "distance": result.distance,
        "proven": bool(result.proven),
        "levels": result.levels,
        "roofline": {
            "bound": rf_bound,
            "achieved": rf_achieved,
            "peak": rf_peak,
            "unit": "TOP/s",
            "frac": rf_achieved / rf_peak,
            "traffic": traffic_bytes(rf_kernel),
            "kernel": rf_kernel,
            "kernels": kernels,
            "valu_TOPs_all_tiled": achieved,
            "ops_per_codeword": tiled_ops / tiled_cw if tiled_cw else None,
            "kernel_ms_per_step": tiled_ms / args.steps,
            "tiled_codewords_per_s": tiled_cw / (tiled_ms * 1e-3) if tiled_ms > 0 else None,
            "five_op_model_TOPs": five_op,
        },
    }
    if not args.no_extras:
        line["batched_learned"] = batched_learned(lambda m: print(m, file=sys.stderr, flush=True), int(args.c4_cap))
        line["config1"] = config1_gpu()
        line["collect"] = collect_figure()
    if not args.no_cpu_baseline:
        host = host_cpus()
        threads = host["usable"]
        line["cpu_baseline"] = cpu_baseline(threads)
        line["cpu_baseline"]["host"] = host
        if "config1" in line:
            line["config1"].update(config1_cpu(threads))
            line["config1"]["cpu_host"] = host
            line["config1"]["distances_agree"] = line["config1"]["distance"] == line["config1"]["cpu_distance"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
