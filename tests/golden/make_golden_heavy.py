"""Writes tests/golden/reference_heavy.json: the reference CPU library's own certified
distances for the heavy BB instances (the gross code, l = 12, and l = 15), which take
about an hour each on 6 cores. Same conventions as make_golden.py.

    python tests/golden/make_golden_heavy.py [threads]
"""
from __future__ import annotations

import json
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from qubitserf_amd import codes  # noqa: E402
from tests import oracle as O  # noqa: E402
from tests.golden.fixtures import enc  # noqa: E402


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 0
    out = HERE / "reference_heavy.json"
    data = json.loads(out.read_text()) if out.exists() else {"meta": {"source": "reference CPU library (oracle/_ref)"}, "css": []}
    done = {c["name"] for c in data["css"]}
    for ell in (12, 15):
        name = f"bb_l{ell}"
        if name in done:
            continue
        Hx, Hz = codes.bb_family(ell)
        t0 = time.time()
        r = O.ref_css_distance(Hx, Hz, "M", threads=threads)
        data["css"].append(dict(name=name, Hx=enc(Hx), Hz=enc(Hz), threads=threads, seconds=time.time() - t0,
                                ref={"M": dict(distance=r.distance, lower_bound=r.lower_bound, proven=r.proven,
                                               levels=r.levels)}))
        out.write_text(json.dumps(data, indent=1))
        print(name, r.distance, r.proven, f"{time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main()
