"""The learned proposal on the device: the BF16/FP32 MFMA GEMM against numpy, the scorer's
forward and backward against the CPU oracle (oracle/mp_scorer.py, same BF16 rounding
points; tolerances below), an Adam step, training, and proposals in batched solves (the
distance never changes; a proposal only swaps the information sets)."""
import importlib.util
from pathlib import Path

import numpy as np
import pytest

from qubitserf_amd import codes
from tests import oracle as O

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
_spec = importlib.util.spec_from_file_location("mp_scorer", ROOT / "oracle" / "mp_scorer.py")
MO = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(MO)


@pytest.fixture(scope="module")
def mp_dev(hip):
    """The scorer's kernels and device arrays (the library's own HIP runtime)."""
    from qubitserf_amd.proposal.devarray import BF16, F32, DevArray
    return DevArray, BF16, F32


@pytest.mark.parametrize("M,N,K,tA,tB,batch", [(16, 16, 16, 0, 0, 1), (37, 70, 45, 0, 0, 3), (64, 1, 64, 0, 0, 2),
                                               (33, 64, 129, 1, 0, 2), (50, 29, 64, 0, 1, 1), (20, 17, 300, 1, 1, 2),
                                               (1, 64, 1, 0, 1, 1)])
def test_gemm_matches_numpy(mp_dev, M, N, K, tA, tB, batch):
    DevArray, BF16, F32 = mp_dev
    from qubitserf_amd.proposal.model import Dev
    dv = Dev()
    rng = np.random.default_rng(M * 1000 + N * 10 + K)
    A = MO.bf16(rng.standard_normal((batch, K, M) if tA else (batch, M, K)).astype(np.float32))
    B = MO.bf16(rng.standard_normal((batch, N, K) if tB else (batch, K, N)).astype(np.float32))
    bias = rng.standard_normal(N).astype(np.float32)
    C0 = rng.standard_normal((batch, M, N)).astype(np.float32)
    opA = np.swapaxes(A, 1, 2) if tA else A
    opB = np.swapaxes(B, 1, 2) if tB else B
    ref = opA.astype(np.float64) @ opB.astype(np.float64)
    dA, dB = DevArray.from_numpy(A, BF16), DevArray.from_numpy(B, BF16)
    C = DevArray.from_numpy(C0)
    dv.gemm(dA, dB, C, M, N, K, batch=batch, transA=tA, transB=tB, sA=M * K, sB=K * N, sC=M * N,
            bias=DevArray.from_numpy(bias), accumulate=1)
    want = ref + C0 + bias
    got = C.numpy()
    assert np.allclose(got, want, rtol=1e-5, atol=1e-4 * np.sqrt(K))
    Cb = DevArray((batch, M, N), BF16)
    dv.gemm(dA, dB, Cb, M, N, K, batch=batch, transA=tA, transB=tB, sA=M * K, sB=K * N, sC=M * N, relu=1)
    gb = Cb.numpy()
    wb = MO.bf16(np.maximum(ref, 0).astype(np.float32))
    assert np.allclose(gb, wb, rtol=2 ** -7, atol=1e-3)


def test_device_packing_equals_host_packing(mp_dev):
    """The scorer's inputs made on the device (qsf_mp_pack) equal the host packing
    (pack_batch_bits, unseeded) bit for bit, on ragged codes with empty rows and columns,
    a zero matrix and the 256-code random set; scores from either are the same."""
    from qubitserf_amd.proposal.model import Dev, MPScorer, pack_batch_bits
    rng = np.random.default_rng(8)
    H_gap = codes.random_rate_code(rng, 30, 70).copy()
    H_gap[:, 5] = 0
    H_gap[3, :] = 0
    sets = [codes.config3_batch(5, 12) + [np.eye(3, 7, dtype=np.uint8), np.zeros((2, 5), np.uint8), H_gap],
            codes.config3_batch(0, 256), [codes.quasi_cyclic(rng, L) for L in (20, 33, 40)]]
    dv = Dev()
    for Hs in sets:
        X, Ac, Av, mask = pack_batch_bits(Hs)
        dX, dAc, dAv, dmask = dv.pack(Hs)
        from qubitserf_amd.proposal.devarray import bf16_bits
        assert np.array_equal(dmask, mask)
        assert np.array_equal(dAc.bits(), Ac) and np.array_equal(dAv.bits(), Av)
        assert np.array_equal(dX.bits(), bf16_bits(X))
    Hs = sets[0]
    sc = MPScorer.load(ROOT / "qubitserf_amd" / "proposal" / "weights.npz")
    host = [np.asarray(x) for x in sc.forward(*pack_batch_bits(Hs)[:3])[0].numpy()]
    dev = sc.scores(Hs)
    for b, H in enumerate(Hs):
        assert np.array_equal(dev[b], host[b][: H.shape[1]])


def _scorer_inputs(seed=0, count=5):
    rng = np.random.default_rng(seed)
    Hs = [codes.random_rate_code(rng, 20, 60) for _ in range(count - 1)] + [codes.quasi_cyclic(rng, 7)]
    from qubitserf_amd.proposal.model import pack_batch
    return Hs, pack_batch(Hs)


def test_forward_backward_match_oracle(mp_dev):
    from qubitserf_amd.proposal.model import MPScorer, init_params
    Hs, (X, Ac, Av, mask) = _scorer_inputs()
    params = init_params(3)
    sc = MPScorer(params)
    s, cache = sc.forward(X, Ac, Av)
    s_ref, c_ref = MO.forward(params, X, Ac, Av, bf=True)
    s_dev = s.numpy()
    assert np.abs(s_dev - s_ref).max() <= 2e-2 * (1 + np.abs(s_ref).max())
    rng = np.random.default_rng(4)
    T = (rng.random(mask.shape) < 0.4).astype(np.float32)
    loss, ds = sc.loss_grad(s, T, mask)
    l_ref, ds_ref = MO.loss_grad(s_ref, T, mask)
    assert abs(loss.item() - l_ref) <= 1e-2 * (1 + abs(l_ref))
    g = sc.backward(cache, ds)
    g_ref = MO.backward(c_ref, ds_ref, bf=True)
    for k in MO.param_names():
        a = g[k].numpy().reshape(-1).astype(np.float64)
        b = np.asarray(g_ref[k], np.float64).reshape(-1)
        scale = np.linalg.norm(b) + 1e-12
        rel = np.linalg.norm(a - b) / scale
        cos = float(a @ b / (np.linalg.norm(a) * scale + 1e-30))
        assert rel < 5e-2 and (cos > 0.998 or scale < 1e-7), (k, rel, cos)


def test_adam_step(mp_dev):
    DevArray, BF16, F32 = mp_dev
    from qubitserf_amd.proposal.model import MPScorer, init_params
    p = init_params(1)
    sc = MPScorer(p)
    rng = np.random.default_rng(2)
    gh = {k: rng.standard_normal(v.shape).astype(np.float32) for k, v in p.items()}
    sc.adam({k: DevArray.from_numpy(v) for k, v in gh.items()}, lr=1e-2)
    for k, v in p.items():
        gk = gh[k]
        m = 0.1 * gk
        vv = 0.001 * gk * gk
        want = v - 1e-2 * (m / 0.1) / (np.sqrt(vv / 0.001) + 1e-8)
        assert np.allclose(sc.w[k].numpy(), want, rtol=1e-5, atol=1e-6), k
        if k in sc.wb:
            assert np.allclose(sc.wb[k].numpy(), MO.bf16(want.astype(np.float32)), rtol=2 ** -7, atol=1e-7)


def test_training_lowers_loss(mp_dev):
    from qubitserf_amd.proposal.model import MPScorer
    from qubitserf_amd.proposal.train import train
    rng = np.random.default_rng(9)
    Hs = [codes.quasi_cyclic(rng, int(rng.integers(6, 12))) for _ in range(24)]
    targets = []
    for H in Hs:   # a learnable target: columns of the first circulant block
        y = np.zeros(H.shape[1], np.float32)
        y[: H.shape[1] // 6] = 1
        targets.append(y)
    sc = MPScorer(seed=0)
    hist = train(sc, Hs, targets, epochs=12, batch=8, lr=3e-3)
    assert hist[-1] < 0.8 * hist[0], hist


def test_proposals_keep_distances(hip, mp_dev):
    from qubitserf_amd import distfind as df
    from qubitserf_amd.proposal import choose_orders
    from qubitserf_amd.proposal.model import MPScorer
    rng = np.random.default_rng(31)
    Hs = [codes.quasi_cyclic(rng, L) for L in (5, 6, 7)] + [codes.random_rate_code(rng, 40, 70) for _ in range(3)]
    want = [O.classical_distance(H).distance for H in Hs]
    plain = df.batch_distance([("classical", H) for H in Hs])
    rand = df.batch_distance(choose_orders(Hs, None, candidates=6))
    learned = df.batch_distance(choose_orders(Hs, MPScorer(seed=5), candidates=6))
    for w, a, b, c in zip(want, plain, rand, learned):
        assert a.distance == b.distance == c.distance == w
        assert a.proven and b.proven and c.proven


def test_proposal_fixes_a_rank_deficient_peel(hip):
    """A code whose natural-order peel leaves its second set rank-deficient: a proposal's
    order restores the rank, so the lower bound climbs faster and fewer levels run."""
    from qubitserf_amd import distfind as df
    from qubitserf_amd.proposal import choose_orders, predicted_work
    rng = np.random.default_rng(0)
    for _ in range(40):
        H = codes.quasi_cyclic(rng, int(rng.integers(10, 15)))
        info = O.problem_info(0, H)
        K, n = info["K"], H.shape[1]
        if len(info["ranks"]) >= 2 and info["ranks"][1] < min(K, n - K):
            break
    else:
        pytest.skip("no rank-deficient natural peel among the sampled codes")
    d = O.classical_distance(H).distance
    plain = df.batch_distance([("classical", H)])[0]
    prop = df.batch_distance(choose_orders([H], None, candidates=16))[0]
    assert plain.distance == prop.distance == d and plain.proven and prop.proven
    assert prop.levels <= plain.levels


def test_trained_scorer_against_uniform_on_held_out_quasi_cyclic(hip):
    """What the trained network gains on held-out quasi-cyclic codes (3 x 6 weight-3
    circulants, L 20..40, a seed the training set never drew), priced by the BZ work the
    information sets predict (their ranks and the device's upper bound):

    * over the natural peel: the proposal fixes every rank-deficient natural peel;
    * over uniform proposals at the same number of candidates: nothing, because there is
      nothing left to gain. The best of 256 uniform orders (the empirical optimum of the
      rank profile) is reached by 16 uniform candidates on every code, so no scorer can
      do better on this family; the test pins that ceiling and that the trained scorer
      reaches it too (never worse than uniform)."""
    from qubitserf_amd import distfind as df
    from qubitserf_amd.proposal import MPScorer
    from qubitserf_amd.proposal.fit import family_work, sample_orders

    w = ROOT / "qubitserf_amd" / "proposal" / "weights.npz"
    assert w.exists(), "trained weights missing"
    scorer = MPScorer.load(w)
    Hs = codes.config4_eval(7, 16, 20, 40)
    ub = [r.distance for r in df.batch_distance([("classical", H) for H in Hs], work_cap=1)]
    rng = np.random.default_rng(5)
    cand = 16
    learned = scorer.scores(Hs)
    orders = []
    for H, s in zip(Hs, learned):
        n = H.shape[1]
        orders.append([np.arange(n, dtype=np.int32)] + sample_orders(s, cand, rng)
                      + sample_orders(np.zeros(n, np.float32), 256, rng))
    work = family_work(Hs, orders, ub)
    gm = lambda xs: float(np.exp(np.mean(np.log(xs))))
    nat = np.array([wk[0] for wk in work])
    lrn = np.array([min([wk[0]] + wk[1:1 + cand]) for wk in work])
    uni = np.array([min([wk[0]] + wk[1 + cand:1 + 2 * cand]) for wk in work])
    opt = np.array([min(wk) for wk in work])
    print(f"held-out QC: gain over natural learned {gm(nat / lrn):.2f}x uniform {gm(nat / uni):.2f}x; "
          f"learned / optimum {gm(lrn / opt):.3f}, uniform / optimum {gm(uni / opt):.3f}; "
          f"codes with a deficient natural peel: {int((nat > opt).sum())}", flush=True)
    assert np.all(lrn <= nat) and np.all(uni <= nat)
    assert np.all(uni == opt), "uniform proposals left rank headroom on this family"
    assert np.all(lrn == opt), "the trained scorer missed the rank optimum that uniform proposals reach"
    assert (nat > opt).sum() >= 1 and gm(nat / lrn) > 1.0
