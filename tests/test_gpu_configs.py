"""The batched, learned-proposal and serving workloads on the device, against the
reference's own answers (tests/golden/reference_configs.json, written by
tests/golden/make_golden_configs.py from the reference CPU library within a per-code CPU
budget; a code over that budget carries no answer).

* Batched (256 random codes, n ~ U{64..160}, rate ~ U[0.3, 0.6], seed 0): one
  batch_distance call for all 256 under the benchmark's work cap equals 256 one-code calls
  (distance, lower bound, proven, levels); every code the reference certified is either
  certified with the reference's distance or bracketed (lower bound <= d <= upper bound);
  uncapped, on the sound level schedule, every such code is certified with the reference's
  distance, and equals the single-code entry point. One reference certificate (code 32:
  12) is refuted by a codeword of weight 11 in a level the reference's schedule skips
  (tests/test_cpu_soundness.py); there the expected distance is 11.
* Learned proposal (16 quasi-cyclic codes, L ~ U{20..40}, seed 0): the trained scorer's
  proposals (qubitserf_amd/proposal/weights.npz), uniform proposals and the natural
  information sets give identical distances wherever certified, and the reference's where
  it certified.
* Serving (96 mixed requests, seed 0): every request's result equals its one-request call;
  bivariate-bicycle requests equal reference_cases.json / reference_heavy.json.
"""
import hashlib
from pathlib import Path

import numpy as np
import pytest

from tests.golden.fixtures import load

pytestmark = pytest.mark.gpu
HERE = Path(__file__).resolve().parent
CONF = load("reference_configs.json") if (HERE / "golden" / "reference_configs.json").exists() else {}
CAP_C3 = int(1e10)
CAP_C4 = int(1e11)


def _sha(H):
    H = np.ascontiguousarray(np.asarray(H, np.uint8) & 1)
    return hashlib.sha256(np.array(H.shape, np.int64).tobytes() + np.packbits(H, axis=1).tobytes()).hexdigest()


def _fixture(name, Hs):
    recs = CONF.get(name)
    if not recs:
        pytest.skip(f"no {name} fixture (run tests/golden/make_golden_configs.py)")
    assert len(recs) == len(Hs)
    for rec, H in zip(recs, Hs):
        assert rec["sha256"] == _sha(H), f"{name}[{rec['index']}]: the code generator changed"
    return recs


def _key(r):
    return (r.distance, r.lower_bound, r.proven, r.levels)


# Reference certificates refuted by an explicit lighter codeword (the reference's level
# schedule skips a rank-deficient information set's first K - r levels but counts it in the
# bound; tests/test_cpu_soundness.py exhibits the codeword): (workload, index) -> distance.
REFUTED = {("c3", 32): 11}


def _against_reference(name, recs, results):
    """Certified: the reference's distance (or, where a lighter codeword refutes the
    reference's certificate, that codeword's weight). Capped: lower bound <= d <= upper."""
    checked = 0
    for rec, r in zip(recs, results):
        ref = rec["ref"]
        if "timeout" in ref:
            continue
        checked += 1
        assert ref["proven"] == 1
        true_d = REFUTED.get((name, rec["index"]), ref["distance"])
        if r.proven:
            assert r.distance == true_d, (name, rec["index"], _key(r), ref)
        else:
            assert r.lower_bound <= true_d <= r.distance, (name, rec["index"], _key(r), ref)
    return checked


def test_c3_batched_equals_single_and_reference(hip):
    from qubitserf_amd import codes
    from qubitserf_amd import distfind as df

    Hs = codes.config3_batch(0, 256)
    recs = _fixture("c3", Hs)
    print("c3: one batched call, 256 codes", flush=True)
    batched = df.batch_distance([("classical", H) for H in Hs], work_cap=CAP_C3)
    print("c3: 256 one-code calls", flush=True)
    single = [df.batch_distance([("classical", H)], work_cap=CAP_C3)[0] for H in Hs]
    for i, (a, b) in enumerate(zip(batched, single)):
        assert _key(a) == _key(b), (i, _key(a), _key(b))
    n = _against_reference("c3", recs, batched)
    assert n >= 100, f"only {n} reference-certified codes"
    # the reference-certified codes without a cap, on the sound level schedule: all
    # certified, the reference's distance (the refuted certificate: the lighter codeword's
    # weight), and equal to the single-code entry point
    idx = [rec["index"] for rec in recs if "timeout" not in rec["ref"]]
    print(f"c3: {len(idx)} reference-certified codes uncapped, sound schedule", flush=True)
    with df.sound_bounds():
        full = df.batch_distance([("classical", Hs[i]) for i in idx])
        for i, r in zip(idx, full):
            want = REFUTED.get(("c3", i), recs[i]["ref"]["distance"])
            assert r.proven and r.distance == want, (i, _key(r), recs[i]["ref"])
        for i in idx[:24] + [32]:
            one = df.api.classical_distance_raw(Hs[i], "bz", "gpu")
            assert (one.distance, one.proven) == (REFUTED.get(("c3", i), recs[i]["ref"]["distance"]), 1)


def test_c4_learned_uniform_and_natural_agree(hip):
    from qubitserf_amd import codes
    from qubitserf_amd import distfind as df
    from qubitserf_amd.proposal import MPScorer, choose_orders

    w = Path(__file__).resolve().parent.parent / "qubitserf_amd" / "proposal" / "weights.npz"
    assert w.exists(), "trained weights missing"
    scorer = MPScorer.load(w)
    Hs = codes.config4_eval(0, 16, 20, 40)
    print("c4: natural, learned, uniform", flush=True)
    plain = df.batch_distance([("classical", H) for H in Hs], work_cap=CAP_C4)
    learned = df.batch_distance(choose_orders(Hs, scorer, candidates=16), work_cap=CAP_C4)
    uniform = df.batch_distance(choose_orders(Hs, None, candidates=16), work_cap=CAP_C4)
    for i, (a, b, c) in enumerate(zip(plain, learned, uniform)):
        ds = {x.distance for x in (a, b, c) if x.proven}
        assert len(ds) <= 1, (i, _key(a), _key(b), _key(c))
        lo = max(x.lower_bound for x in (a, b, c))
        hi = min(x.distance for x in (a, b, c))
        assert lo <= hi, (i, _key(a), _key(b), _key(c))
    assert sum(x.proven for x in learned) >= 1
    recs = CONF.get("c4")
    if recs:
        for res in (plain, learned, uniform):
            _against_reference("c4", recs, res)


def test_c5_serving_equals_single_calls_and_reference(hip):
    from qubitserf_amd import distfind as df
    from qubitserf_amd import serving
    from tests.golden.fixtures import dec

    reqs = serving.make_requests(96, seed=0)
    print("c5: serving 96 requests", flush=True)
    stats = serving.serve(reqs, rate=40.0, seed=1, keep_results=True)
    res = stats["results"]
    assert all(r is not None for r in res)
    print("c5: one-request calls", flush=True)
    for i, rq in enumerate(reqs):
        one = df.batch_distance([{"problem": rq["problem"], "work_cap": rq["work_cap"]}])[0]
        assert res[i] == _key(one), (i, rq["kind"], res[i], _key(one))
    gold = {c["name"]: c for c in load()["css"]}
    gold.update({c["name"]: c for c in load("reference_heavy.json")["css"]})
    for i, rq in enumerate(reqs):
        if rq["kind"].startswith("bb"):
            case = gold[f"bb_l{rq['kind'][2:]}"]
            Hx, Hz = dec(case["Hx"]), dec(case["Hz"])
            assert (Hx == rq["problem"][1]).all() and (Hz == rq["problem"][2]).all()
            want = case["ref"]["M"]
            assert (res[i][0], res[i][2]) == (want["distance"], want["proven"]), (i, rq["kind"], res[i], want)
    recs = CONF.get("c5")
    if recs:
        cls = [i for i, rq in enumerate(reqs) if rq["kind"] == "classical"]
        assert len(cls) == len(recs)

        class R:
            def __init__(self, t):
                self.distance, self.lower_bound, self.proven, self.levels = t

        _against_reference("c5", recs, [R(res[i]) for i in cls])
