"""The engine's kernel-choice autotuner: measures, on the current gfx950 device, the
constants of the cost model that picks a hot-test variant for every weight level
(choose_variant in csrc/enum_hip.hip), and writes them to csrc/tuning.hpp.

    python -m qubitserf_amd.tune [--write] [--out PATH]

Measured trials, all through the plugin-level C ABI with the variant forced
(QSA_TILED_VARIANT):
  * cost per codeword of each VALU variant and of the int8, fp4 and chained fp4 matrix
    tests, on one
    large level (K = 80, d = 8, C(80, 8) = 2.9e10 codewords) of a random systematic
    generator whose threshold lets almost nothing pass the hot test;
  * cost per hot-test pass: the same level with a looser bound, so that ~20% of the test
    units pass (the difference in time over the difference in passes);
  * launch constant: a small level (d = 4, 1.6e6 codewords) minus its per-codeword cost;
  * the structure factor's elasticity: the gross code's hot-test passes over their
    fair-coin expectation at d = 3..10 (int8 test forced), fitted as factor ~ p^-e over the
    levels in the tail (fewer than half the groups would pass on fair coins), p being the
    fair per-codeword probability of passing the 30-bit test.
The table is committed; the library compiles it in. Results never depend on it.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time
from pathlib import Path

import numpy as np

from . import engine

HERE = Path(__file__).resolve().parent
VALU = [(1, 1, 32), (1, 2, 16), (1, 2, 24), (1, 2, 32), (2, 2, 16), (2, 4, 12), (3, 4, 12), (4, 4, 12)]
NONE = 1 << 30


def cdf_below(bits: int, thr: int) -> float:
    """P(Binomial(bits, 1/2) < thr)."""
    if thr <= 0:
        return 0.0
    if thr > bits:
        return 1.0
    return sum(math.comb(bits, k) for k in range(thr)) / 2.0**bits


def unit_and_bits(v):
    fd, xd, rpl, mm = v
    if mm == 1:
        return 8192.0, 30
    if mm == 2:
        return 8192.0, 60
    if mm == 4:
        return 4096.0, 124
    if mm == 5:
        return 16384.0, 30
    return 64.0 * rpl, 32 * fd


def fair_hits(v, cw, thr, res):
    unit, bits = unit_and_bits(v)
    return cw / unit * min(1.0, unit * cdf_below(min(bits, res), thr))


def systematic(rng, K, stride):
    gamma = rng.integers(0, 2**64, size=(1, K, stride), dtype=np.uint64)
    cols = rng.choice(stride * 64, size=K, replace=False)
    for r, c in enumerate(cols):
        w, b = divmod(int(c), 64)
        gamma[0, :, w] &= ~np.uint64(1 << b)
        gamma[0, r, w] |= np.uint64(1 << b)
    return gamma


class Runner:
    def __init__(self, gamma, n):
        self.gamma, self.n = gamma, n
        self.key = engine.new_key()

    def run(self, d, best, v, reps=3):
        fd, xd, rpl, mm = v
        os.environ["QSA_TILED_VARIANT"] = f"{fd},{xd},{rpl},{mm}"
        try:
            ts = []
            with engine.kernel_policy("tiled"):
                for _ in range(reps):
                    engine.enumerate_level(self.gamma, self.n, d, current_best=best, key=self.key)
                    st = engine.last_stats()
                    assert (st["fd"], st["mfma"]) == (fd, mm), (v, st)
                    ts.append(st["kernel_ms"] * 1e-3)
        finally:
            del os.environ["QSA_TILED_VARIANT"]
        return min(ts), st["medium_hits"]


def loose_thr(v, cw, res, target=0.2):
    """The smallest threshold at which about `target` of the variant's test units pass."""
    unit, bits = unit_and_bits(v)
    for thr in range(1, min(bits, res) + 1):
        if unit * cdf_below(min(bits, res), thr) >= target:
            return thr
    return min(bits, res)


def gross_growth(log):
    from . import codes
    from .distfind import api

    Hx, Hz = codes.gross()
    os.environ["QSA_TILED_VARIANT"] = "0,4,4,1"
    os.environ["QSA_MATRIX_MIN_WORK"] = "0"
    try:
        api.css_distance_raw(Hx, Hz, "bz", "M", "gpu")
        tr = api.last_trace()
    finally:
        del os.environ["QSA_TILED_VARIANT"], os.environ["QSA_MATRIX_MIN_WORK"]
    factors, pts = {}, []
    for t in tr:
        if t["side"] != "Z" or t["mfma"] != 1:
            continue
        cw, d = t["codewords"], t["d"]
        p = cdf_below(30, 12 - d)   # every row is flagged: threshold = bound 12 - d
        e = fair_hits((0, 4, 4, 1), cw, 12 - d, 66)
        if e >= 8:
            factors[d] = t["medium_hits"] / e
            if 8192 * p <= 0.5:
                pts.append((math.log(p), math.log(max(factors[d], 1.0))))
    if len(pts) >= 2:
        mx = sum(x for x, _ in pts) / len(pts)
        my = sum(y for _, y in pts) / len(pts)
        el = -sum((x - mx) * (y - my) for x, y in pts) / sum((x - mx) ** 2 for x, _ in pts)
    else:
        el = 0.2
    log(f"gross structure factors (int8, Z side): {json.dumps({d: round(f, 2) for d, f in factors.items()})}; "
        f"elasticity {el:.3f} over {len(pts)} tail levels")
    return el, factors


def measure(log=print):
    if not engine.available():
        raise SystemExit("gfx950 device unavailable")
    rng = np.random.default_rng(2024)
    K, stride, d = 80, 3, 8
    n = stride * 64
    res = n - K
    run = Runner(systematic(rng, K, stride), n)
    cw = math.comb(K, d)
    tight = d + 2
    variants = [(f, x, r, 0) for f, x, r in VALU] + [(0, 4, 4, 1), (0, 4, 4, 2), (0, 4, 4, 4), (0, 4, 4, 5)]
    out = {"level": f"K={K} n={n} d={d} random systematic, {cw:.3e} codewords", "variants": {}}
    small_d = 4
    small_cw = math.comb(K, small_d)
    for v in variants:
        t_lo, h_lo = run.run(d, tight, v)
        thr = min(loose_thr(v, cw, res), 18)   # bound <= 26: below the level's lightest codeword (~29)
        t_hi, h_hi = run.run(d, d + thr, v)
        t_small, _ = run.run(small_d, small_d + 2, v)
        hit = (t_hi - t_lo) / (h_hi - h_lo) if h_hi > h_lo + 100 else None
        cost = t_lo / cw
        launch = max(0.0, t_small - small_cw * cost)
        rec = dict(cost=cost, hit=hit, launch=launch, t_tight=t_lo, hits_tight=h_lo, thr_loose=thr, t_loose=t_hi,
                   hits_loose=h_hi, t_small=t_small)
        out["variants"][",".join(map(str, v))] = rec
        log(f"variant {v}: {json.dumps(rec)}")
    el, factors = gross_growth(log)
    out["elasticity"] = el
    out["gross_factors"] = factors
    return out


def header(m, device):
    V = m["variants"]
    valu = [V[f"{f},{x},{r},0"] for f, x, r in VALU]
    i8, f4, f4x2, f4pk = V["0,4,4,1"], V["0,4,4,2"], V["0,4,4,4"], V["0,4,4,5"]
    hits = [r["hit"] for r in valu if r["hit"]]
    valu_hit_default = sorted(hits)[len(hits) // 2] if hits else 1.5e-9

    def fmt(x):
        return f"{x:.3e}"

    launch_valu = sorted(r["launch"] for r in valu)[len(valu) // 2]
    launch_mm = max(i8["launch"], f4["launch"], f4x2["launch"], f4pk["launch"])
    lines = [
        "// Kernel-choice constants of the enumeration engine, measured on an MI355X.",
        "// Regenerate with:  python -m qubitserf_amd.tune --write   (on a gfx950 device)",
        f"// Measured on {device}, {time.strftime('%Y-%m-%d')}: {m['level']}.",
        "#pragma once",
        "",
        "namespace qsa {",
        "namespace tuning {",
        "",
        "// Seconds of full-chip device time per codeword on a large level whose hot test rarely",
        "// passes, for the VALU variants in run_tiled's table order (fd, xd, rpl): <1,1,32>",
        "// <1,2,16> <1,2,24> <1,2,32> <2,2,16> <2,4,12> <3,4,12> <4,4,12>.",
        "constexpr double kValuCost[8] = {" + ", ".join(fmt(r["cost"]) for r in valu) + "};",
        "// Seconds per hot-test pass (a stream step entering the medium test) of those variants.",
        "constexpr double kValuHit[8] = {" + ", ".join(fmt(r["hit"] or valu_hit_default) for r in valu) + "};",
        "// int8 and fp4 (32x32x64) matrix-core tests: seconds per codeword, and per group of 8192",
        "// codewords sent to the exact tests.",
        f"constexpr double kInt8Cost = {fmt(i8['cost'])};",
        f"constexpr double kInt8Hit = {fmt(i8['hit'] or 1.8e-10)};",
        f"constexpr double kF4Cost = {fmt(f4['cost'])};",
        f"constexpr double kF4Hit = {fmt(f4['hit'] or i8['hit'] or 1.8e-10)};",
        "// Two chained fp4 32x32x64 products (124 test bits): per codeword, and per group of 4096.",
        f"constexpr double kF4x2Cost = {fmt(f4x2['cost'])};",
        f"constexpr double kF4x2Hit = {fmt(f4x2['hit'] or f4['hit'] or 1.8e-10)};",
        "// Packed fp4 32x32x64 products (30 test bits, two lane elements per row): per codeword, and",
        "// per group of 16384.",
        f"constexpr double kF4PkCost = {fmt(f4pk['cost'])};",
        f"constexpr double kF4PkHit = {fmt(f4pk['hit'] or i8['hit'] or 1.8e-10)};",
        "// Fixed seconds per level launch (set-up, tail) of the VALU and matrix-core kernels.",
        f"constexpr double kValuLaunch = {fmt(launch_valu)};",
        f"constexpr double kMatrixLaunch = {fmt(launch_mm)};",
        "// Structure factor (measured hot-test passes / fair-coin expectation): its elasticity with",
        "// respect to the fair per-codeword pass probability (factor ~ p^-kElasticity), fitted on",
        "// the gross code's Z side (int8 test), and the value assumed before a solve has measured it.",
        f"constexpr double kElasticity = {m['elasticity']:.3f};",
        "constexpr double kPriorFactor = 1.0;",
        "",
        "}  // namespace tuning",
        "}  // namespace qsa",
        "",
    ]
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--write", action="store_true", help="write csrc/tuning.hpp")
    ap.add_argument("--out", default=None, help="also write the header here")
    ap.add_argument("--json", default=None, help="write the raw measurements here")
    a = ap.parse_args()
    m = measure(lambda s: print(s, flush=True))
    from . import _native  # noqa: F401

    device = os.environ.get("QSA_DEVICE_NAME", "gfx950")
    h = header(m, device)
    if a.json:
        Path(a.json).write_text(json.dumps(m, indent=1))
    if a.out:
        Path(a.out).write_text(h)
    if a.write:
        (HERE / "csrc" / "tuning.hpp").write_text(h)
    print(h)


if __name__ == "__main__":
    main()
