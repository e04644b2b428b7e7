"""Writes tests/golden/reference_configs.json: the reference CPU library's own answers for
the seeded workloads of the batched, learned-proposal and serving benchmarks, for every
code the reference certifies within a fixed CPU budget.

* c3: codes.config3_batch(0, 256)  (n ~ U{64..160}, rate ~ U[0.3, 0.6], Bernoulli(1/2) H)
* c4: codes.config4_eval(0, 16)    (quasi-cyclic, L ~ U{20..40}, 3 x 6 circulants of weight 3)
* c5: the classical requests of serving.make_requests(96, seed=0)  (n ~ U{64..256}); its
      bivariate-bicycle requests (l = 6, 9, 12, 15) are already pinned by
      reference_cases.json (bb_l6, bb_l9) and reference_heavy.json (bb_l12, bb_l15).

Each code runs in its own process (the reference library, oracle/_ref, backend "cpu",
`--threads` threads) under a wall-clock limit of `--budget` seconds; a code that does not
finish is recorded as {"timeout": budget} and carries no answer. Codes are identified by
their index in the generator's sequence and the SHA-256 of their packed matrix, so the
fixture stays small and a change in the generator is caught. Run where the reference
tree is (this container):

    make -f oracle/ref.mk REF=/root/reference
    python tests/golden/make_golden_configs.py [--budget 30] [--threads 2] [--jobs 4]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from qubitserf_amd import codes  # noqa: E402

OUT = HERE / "reference_configs.json"


def workload(name):
    """The parity-check matrices of one workload, in generator order."""
    if name == "c3":
        return codes.config3_batch(0, 256)
    if name == "c4":
        return codes.config4_eval(0, 16)
    if name == "c5":
        from qubitserf_amd.serving import make_requests

        return [r["problem"][1] for r in make_requests(96, seed=0) if r["kind"] == "classical"]
    raise KeyError(name)


def matrix_sha(H) -> str:
    H = np.ascontiguousarray(np.asarray(H, np.uint8) & 1)
    return hashlib.sha256(np.array(H.shape, np.int64).tobytes() + np.packbits(H, axis=1).tobytes()).hexdigest()


def one(name, index, threads):
    """Child process: the reference's classical_distance of one code, as a JSON line."""
    from tests import oracle as O

    H = workload(name)[index]
    t0 = time.time()
    r = O.ref_classical_distance(H, threads=threads)
    print(json.dumps(dict(distance=r.distance, lower_bound=r.lower_bound, proven=r.proven, levels=r.levels,
                          seconds=round(time.time() - t0, 3))), flush=True)


def run(name, index, budget, threads):
    cmd = [sys.executable, str(Path(__file__).resolve()), "--one", name, str(index), "--threads", str(threads)]
    try:
        p = subprocess.run(cmd, capture_output=True, text=True, timeout=budget)
    except subprocess.TimeoutExpired:
        return dict(timeout=budget)
    if p.returncode != 0:
        raise RuntimeError(f"{name}[{index}]: {p.stderr[-2000:]}")
    return json.loads(p.stdout.strip().splitlines()[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--one", nargs=2)
    ap.add_argument("--budget", type=float, default=30.0)
    ap.add_argument("--threads", type=int, default=2)
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--only", default="c3,c4,c5")
    ap.add_argument("--indices", default=None,
                    help="comma-separated codes to (re)run, e.g. the smallest quasi-cyclic ones with a larger budget; "
                         "the workload's other entries are kept")
    a = ap.parse_args()
    if a.one:
        one(a.one[0], int(a.one[1]), a.threads)
        return
    from tests import oracle as O

    if not O.ref_available():
        raise SystemExit("oracle/_ref/libqubitserf_ref.so missing: make -f oracle/ref.mk REF=/root/reference")
    data = json.loads(OUT.read_text()) if OUT.exists() else {}
    data["meta"] = dict(source="reference CPU library (oracle/_ref), backend cpu, one process per code",
                        budget_s=a.budget, threads=a.threads, rng="numpy default_rng (PCG64)")
    for name in a.only.split(","):
        Hs = workload(name)
        t0 = time.time()
        todo = [int(i) for i in a.indices.split(",")] if a.indices else list(range(len(Hs)))
        with ThreadPoolExecutor(a.jobs) as ex:
            res = dict(zip(todo, ex.map(lambda i: run(name, i, a.budget, a.threads), todo)))
        old = {e["index"]: e for e in data.get(name, [])}
        recs = []
        for i, H in enumerate(Hs):
            if i in res:
                recs.append(dict(index=i, n=int(H.shape[1]), m=int(H.shape[0]), sha256=matrix_sha(H), ref=res[i],
                                 budget_s=a.budget, threads=a.threads))
            else:
                recs.append(old[i])
        data[name] = recs
        done = sum(1 for r in res.values() if "timeout" not in r)
        print(f"{name}: {done} of {len(todo)} certified within {a.budget:.0f} s, {time.time() - t0:.0f} s", flush=True)
        OUT.write_text(json.dumps(data, indent=1))


if __name__ == "__main__":
    main()
