"""The message-passing scorer on the device. Weights: FP32 masters + BF16 copies; every
GEMM runs on qsf_mp_gemm (MFMA, BF16 in, FP32 accumulate); tensors are DevArrays on the
library's own HIP runtime, kernels run on the null stream.

Network (per code with parity check H, m checks x n columns, hidden width D = 64):
    Ac = H / deg_c,  Av = H^T / deg_v
    h0 = relu(X Win + bin)
    l = 1..3:  cm = Ac h;  cg = relu(cm Wc + bc);  va = Av cg;  h = relu(h Ws + va Wn + b)
    s = h wo + bo                      one logit per column
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from .. import _native
from .devarray import BF16, F32, I32, U8, DevArray

D = 64           # hidden width
F = 16           # input features per column (padded)
LAYERS = 3


def _lib():
    L = _native.load()
    if not getattr(L, "_mp_bound", False):
        from ctypes import c_int, c_longlong, c_float, c_void_p, POINTER

        class Desc(ctypes.Structure):
            _fields_ = [("A", c_void_p), ("B", c_void_p), ("C", c_void_p), ("bias", c_void_p),
                        ("batch", c_int), ("M", c_int), ("N", c_int), ("K", c_int), ("lda", c_int),
                        ("ldb", c_int), ("ldc", c_int), ("sA", c_longlong), ("sB", c_longlong),
                        ("sC", c_longlong), ("transA", c_int), ("transB", c_int), ("out_bf16", c_int),
                        ("accumulate", c_int), ("relu", c_int)]
        L._Desc = Desc
        L.qsf_mp_gemm.argtypes = [POINTER(Desc), c_void_p]
        L.qsf_mp_to_bf16.argtypes = [c_void_p, c_void_p, c_longlong, c_void_p]
        L.qsf_mp_pack.argtypes = [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p]
        L.qsf_mp_relu_bwd.argtypes = [c_void_p, c_void_p, c_longlong, c_void_p]
        L.qsf_mp_colsum.argtypes = [c_void_p, c_void_p, c_longlong, c_int, c_int, c_void_p]
        L.qsf_mp_bce.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_longlong, c_float, c_void_p]
        L.qsf_mp_adam.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_longlong, c_float, c_float,
                                  c_float, c_float, c_int, c_void_p]
        for f in ("qsf_mp_gemm", "qsf_mp_pack", "qsf_mp_to_bf16", "qsf_mp_relu_bwd", "qsf_mp_colsum", "qsf_mp_bce", "qsf_mp_adam"):
            getattr(L, f).restype = c_int
        L._mp_bound = True
    return L


def _ok(rc, what):
    if rc:
        raise RuntimeError(f"{what} failed (rc={rc})")


class Dev:
    """Thin wrapper of the kernels (null stream)."""

    stream = None

    def __init__(self):
        from .. import engine
        if not engine.available():
            raise RuntimeError("the learned proposal needs a gfx950 device")
        self.L = _lib()

    def gemm(self, A, B, C, M, N, K, *, batch=1, transA=0, transB=0, lda=None, ldb=None, ldc=None,
             sA=0, sB=0, sC=0, bias=None, relu=0, accumulate=0):
        assert A.dtype == BF16 and B.dtype == BF16
        out_bf16 = 1 if C.dtype == BF16 else 0
        lda = lda if lda is not None else (M if transA else K)
        ldb = ldb if ldb is not None else (K if transB else N)
        ldc = ldc if ldc is not None else N
        # bounds of the last batch's last element, checked on the host before the launch
        need_a = (batch - 1) * sA + (K - 1) * lda + M if transA else (batch - 1) * sA + (M - 1) * lda + K
        need_b = (batch - 1) * sB + (N - 1) * ldb + K if transB else (batch - 1) * sB + (K - 1) * ldb + N
        need_c = (batch - 1) * sC + (M - 1) * ldc + N
        assert need_a <= A.numel and need_b <= B.numel and need_c <= C.numel, "GEMM operand out of bounds"
        assert bias is None or (bias.dtype == F32 and bias.numel >= N)
        d = self.L._Desc(A.ptr, B.ptr, C.ptr, bias.ptr if bias is not None else None,
                         batch, M, N, K, lda, ldb, ldc, sA, sB, sC, transA, transB, out_bf16, accumulate, relu)
        _ok(self.L.qsf_mp_gemm(ctypes.byref(d), self.stream), "qsf_mp_gemm")

    def pack(self, Hs):
        """pack_batch_bits(Hs) (unseeded) computed on the device from the padded 0/1 checks:
        X, Ac, Av as bfloat16 DevArrays, and the column mask (numpy)."""
        B = len(Hs)
        m = max(h.shape[0] for h in Hs)
        n = max(h.shape[1] for h in Hs)
        Hb = np.zeros((B, m, n), np.uint8)
        hms = np.zeros(B, np.int32)
        hns = np.zeros(B, np.int32)
        for b, H in enumerate(Hs):
            H = np.asarray(H)
            Hb[b, : H.shape[0], : H.shape[1]] = H.astype(np.uint8, copy=False) & 1
            hms[b], hns[b] = H.shape
        from .devarray import bf16_bits
        inv = bf16_bits(np.float32(1.0) / np.arange(1, max(m, n) + 1, dtype=np.float32))
        noise = np.random.default_rng(0).standard_normal((n, 4)).astype(np.float32) * 0.5
        dH, dm, dn = DevArray.from_raw(Hb, U8), DevArray.from_raw(hms, I32), DevArray.from_raw(hns, I32)
        dinv, dnoise = DevArray.from_raw(inv, BF16), DevArray.from_raw(noise, F32)
        X = DevArray((B, n, F), BF16, zero=False)
        Ac = DevArray((B, m, n), BF16, zero=False)
        Av = DevArray((B, n, m), BF16, zero=False)
        _ok(self.L.qsf_mp_pack(dH.ptr, dm.ptr, dn.ptr, B, m, n, dinv.ptr, dnoise.ptr, Ac.ptr, Av.ptr, X.ptr,
                               self.stream), "qsf_mp_pack")
        mask = (np.arange(n)[None, :] < hns[:, None]).astype(np.float32)
        return X, Ac, Av, mask

    def bf16(self, x):
        y = DevArray(x.shape, BF16, zero=False)
        _ok(self.L.qsf_mp_to_bf16(x.ptr, y.ptr, x.numel, self.stream), "qsf_mp_to_bf16")
        return y

    def relu_bwd(self, y, dy):
        assert y.numel == dy.numel and y.dtype == BF16 and dy.dtype == F32
        _ok(self.L.qsf_mp_relu_bwd(y.ptr, dy.ptr, dy.numel, self.stream), "qsf_mp_relu_bwd")

    def colsum(self, x, out, accumulate=0):
        cols = out.numel
        assert x.numel % cols == 0
        _ok(self.L.qsf_mp_colsum(x.ptr, out.ptr, x.numel // cols, cols, accumulate, self.stream), "qsf_mp_colsum")


def features(H: np.ndarray, seed: int = 0) -> np.ndarray:
    """n x F column features: constant, normalised degree, degree spread of the column's
    checks, and random symmetry-breaking values (seeded per code)."""
    H = (np.asarray(H, np.uint8) & 1).astype(np.float32)
    m, n = H.shape
    dv = H.sum(0)
    dc = H.sum(1)
    X = np.zeros((n, F), np.float32)
    X[:, 0] = 1.0
    X[:, 1] = dv / max(1.0, dv.max())
    mean_dc = (H * dc[:, None]).sum(0) / np.maximum(dv, 1)
    X[:, 2] = mean_dc / max(1.0, dc.max())
    X[:, 3] = m / n
    X[:, 4:8] = np.random.default_rng(seed).standard_normal((n, 4)).astype(np.float32) * 0.5
    return X


def pack_batch(Hs, seeds=None):
    """Pad codes to common (m, n): X (B, n, F), Ac (B, m, n), Av (B, n, m), mask (B, n)."""
    B = len(Hs)
    m = max(h.shape[0] for h in Hs)
    n = max(h.shape[1] for h in Hs)
    X = np.zeros((B, n, F), np.float32)
    Ac = np.zeros((B, m, n), np.float32)
    Av = np.zeros((B, n, m), np.float32)
    mask = np.zeros((B, n), np.float32)
    for b, H in enumerate(Hs):
        H = (np.asarray(H, np.uint8) & 1).astype(np.float32)
        hm, hn = H.shape
        X[b, :hn] = features(H, 0 if seeds is None else seeds[b])
        dc = np.maximum(H.sum(1, keepdims=True), 1)
        dv = np.maximum(H.sum(0, keepdims=True), 1)
        Ac[b, :hm, :hn] = H / dc
        Av[b, :hn, :hm] = (H / dv).T
        mask[b, :hn] = 1
    return X, Ac, Av, mask


def pack_batch_bits(Hs, seeds=None):
    """pack_batch computed for the whole batch at once, with the two averaging operators
    built directly as bfloat16 bit patterns (uint16): an entry of Ac is 1 / deg(check) and
    of Av 1 / deg(column), so both come from one table of the patterns of 1 / k. Equal, bit
    for bit, to pack_batch with bf16_bits applied to its float32 operators (the degree sums
    are small integers, exact in float32 in any summation order)."""
    from .devarray import bf16_bits

    B = len(Hs)
    m = max(h.shape[0] for h in Hs)
    n = max(h.shape[1] for h in Hs)
    Hb = np.zeros((B, m, n), np.uint8)
    hms = np.zeros(B, np.int64)
    hns = np.zeros(B, np.int64)
    for b, H in enumerate(Hs):
        H = np.asarray(H)
        Hb[b, : H.shape[0], : H.shape[1]] = H.astype(np.uint8, copy=False) & 1
        hms[b], hns[b] = H.shape
    mask = (np.arange(n)[None, :] < hns[:, None]).astype(np.float32)
    dc_raw = Hb.sum(2, dtype=np.uint16).astype(np.float32)          # (B, m)
    dv_raw = Hb.sum(1, dtype=np.uint16).astype(np.float32)          # (B, n)
    inv = bf16_bits(np.float32(1.0) / np.arange(1, max(m, n) + 1, dtype=np.float32))
    pc = inv[np.maximum(dc_raw, 1).astype(np.int64) - 1]          # (B, m) pattern of 1 / deg_c
    pv = inv[np.maximum(dv_raw, 1).astype(np.int64) - 1]          # (B, n) pattern of 1 / deg_v
    # H is 0 / 1, so H x pattern is the pattern or 0: Ac = H / deg_c, Av = H^T / deg_v
    Ac = Hb * pc[:, :, None]
    Av = np.multiply(Hb.transpose(0, 2, 1), pv[:, :, None], dtype=np.uint16)   # strided read, one pass
    X = np.zeros((B, n, F), np.float32)
    X[:, :, 0] = mask
    X[:, :, 1] = dv_raw / np.maximum(np.float32(1.0), dv_raw.max(1, keepdims=True))
    mean_dc = np.matmul(dc_raw[:, None, :], Hb, dtype=np.float32)[:, 0, :] / np.maximum(dv_raw, np.float32(1.0))
    X[:, :, 2] = mean_dc / np.maximum(np.float32(1.0), dc_raw.max(1, keepdims=True))
    X[:, :, 3] = (hms / hns).astype(np.float32)[:, None] * mask
    if seeds is None:   # every code draws from seed 0: the first n_b rows of one draw
        noise = np.random.default_rng(0).standard_normal((n, 4)).astype(np.float32) * 0.5
        for b in range(B):
            X[b, : hns[b], 4:8] = noise[: hns[b]]
    else:
        for b in range(B):
            X[b, : hns[b], 4:8] = np.random.default_rng(seeds[b]).standard_normal(
                (int(hns[b]), 4)).astype(np.float32) * 0.5
    return X, Ac, Av, mask


def param_shapes():
    shp = {"Win": (F, D), "bin": (D,)}
    for l in range(LAYERS):
        shp.update({f"Wc{l}": (D, D), f"bc{l}": (D,), f"Ws{l}": (D, D), f"Wn{l}": (D, D), f"b{l}": (D,)})
    shp.update({"wo": (D, 1), "bo": (1,)})
    return shp


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    out = {}
    for k, s in param_shapes().items():
        if k.startswith(("W", "w")):
            out[k] = (rng.standard_normal(s) * math.sqrt(2.0 / s[0])).astype(np.float32)
        else:
            out[k] = np.zeros(s, np.float32)
    return out


class MPScorer:
    """Device-resident scorer: forward, backward and Adam steps on the HIP kernels."""

    def __init__(self, params=None, seed=0):
        self.dev = Dev()
        p = params if params is not None else init_params(seed)
        self.w = {k: DevArray.from_numpy(np.asarray(v, np.float32)) for k, v in p.items()}
        self.wb = {k: self.dev.bf16(v) for k, v in self.w.items() if k.startswith(("W", "w"))}
        self.m = {k: DevArray(v.shape) for k, v in self.w.items()}
        self.v = {k: DevArray(v.shape) for k, v in self.w.items()}
        self.step_no = 0

    def params(self):
        return {k: v.numpy().copy() for k, v in self.w.items()}

    # ---------------------------------------------------------------- forward
    def forward(self, X, Ac, Av):
        dv = self.dev
        B, n, _ = X.shape
        m = Ac.shape[1]
        c = {"B": B, "n": n, "m": m}
        # inputs on the host (float32, or the operators as bfloat16 patterns from
        # pack_batch_bits), or already on the device (Dev.pack)
        c["X"] = X if isinstance(X, DevArray) else DevArray.from_numpy(X, BF16)
        if isinstance(Ac, DevArray):
            c["Ac"], c["Av"] = Ac, Av
        else:
            c["Ac"] = DevArray.from_bits(Ac) if Ac.dtype == np.uint16 else DevArray.from_numpy(Ac, BF16)
            c["Av"] = DevArray.from_bits(Av) if Av.dtype == np.uint16 else DevArray.from_numpy(Av, BF16)
        h = DevArray((B, n, D), BF16, zero=False)
        dv.gemm(c["X"], self.wb["Win"], h, B * n, D, F, bias=self.w["bin"], relu=1)
        c["h0"] = h
        for l in range(LAYERS):
            cm = DevArray((B, m, D), BF16, zero=False)
            dv.gemm(c["Ac"], h, cm, m, D, n, batch=B, sA=m * n, sB=n * D, sC=m * D)
            cg = DevArray((B, m, D), BF16, zero=False)
            dv.gemm(cm, self.wb[f"Wc{l}"], cg, B * m, D, D, bias=self.w[f"bc{l}"], relu=1)
            va = DevArray((B, n, D), BF16, zero=False)
            dv.gemm(c["Av"], cg, va, n, D, m, batch=B, sA=n * m, sB=m * D, sC=n * D)
            pre = DevArray((B, n, D), F32, zero=False)
            dv.gemm(h, self.wb[f"Ws{l}"], pre, B * n, D, D)
            dv.gemm(va, self.wb[f"Wn{l}"], pre, B * n, D, D, bias=self.w[f"b{l}"], relu=1, accumulate=1)
            hn = dv.bf16(pre)
            c[f"cm{l}"], c[f"cg{l}"], c[f"va{l}"], c[f"h{l + 1}"] = cm, cg, va, hn
            h = hn
        s = DevArray((B, n), F32, zero=False)
        dv.gemm(h, self.wb["wo"], s, B * n, 1, D, bias=self.w["bo"])
        c["s"] = s
        return s, c

    # ---------------------------------------------------------------- backward
    def backward(self, c, ds):
        """Gradients (FP32, on the device) of every parameter given d loss / d logits."""
        dv = self.dev
        B, n, m = c["B"], c["n"], c["m"]
        g = {k: DevArray(v.shape) for k, v in self.w.items()}
        dsq = dv.bf16(ds)
        hL = c[f"h{LAYERS}"]
        dv.gemm(hL, dsq, g["wo"], D, 1, B * n, transA=1)
        dv.colsum(ds, g["bo"])
        dh = DevArray((B, n, D), F32, zero=False)
        dv.gemm(dsq, self.wb["wo"], dh, B * n, D, 1, transB=1)
        for l in reversed(range(LAYERS)):
            hn, hp = c[f"h{l + 1}"], c[f"h{l}"]
            dpre = dh
            dv.relu_bwd(hn, dpre)
            dv.colsum(dpre, g[f"b{l}"])
            dpq = dv.bf16(dpre)
            dv.gemm(hp, dpq, g[f"Ws{l}"], D, D, B * n, transA=1)
            dv.gemm(c[f"va{l}"], dpq, g[f"Wn{l}"], D, D, B * n, transA=1)
            dh_prev = DevArray((B, n, D), F32, zero=False)
            dv.gemm(dpq, self.wb[f"Ws{l}"], dh_prev, B * n, D, D, transB=1)
            dva = DevArray((B, n, D), BF16, zero=False)
            dv.gemm(dpq, self.wb[f"Wn{l}"], dva, B * n, D, D, transB=1)
            dcg = DevArray((B, m, D), F32, zero=False)
            dv.gemm(c["Av"], dva, dcg, m, D, n, batch=B, transA=1, lda=m, sA=n * m, sB=n * D, sC=m * D)
            dv.relu_bwd(c[f"cg{l}"], dcg)
            dv.colsum(dcg, g[f"bc{l}"])
            dcq = dv.bf16(dcg)
            dv.gemm(c[f"cm{l}"], dcq, g[f"Wc{l}"], D, D, B * m, transA=1)
            dcm_in = DevArray((B, m, D), BF16, zero=False)
            dv.gemm(dcq, self.wb[f"Wc{l}"], dcm_in, B * m, D, D, transB=1)
            dv.gemm(c["Ac"], dcm_in, dh_prev, n, D, m, batch=B, transA=1, lda=n, sA=m * n, sB=m * D, sC=n * D,
                    accumulate=1)
            dh = dh_prev
        dv.relu_bwd(c["h0"], dh)
        dv.colsum(dh, g["bin"])
        dv.gemm(c["X"], dv.bf16(dh), g["Win"], F, D, B * n, transA=1)
        return g

    def loss_grad(self, s, target, mask):
        t = DevArray.from_numpy(target)
        mk = DevArray.from_numpy(mask)
        ds = DevArray(s.shape, F32, zero=False)
        loss = DevArray((1,), F32)
        cnt = max(1.0, float(np.asarray(mask).sum()))
        _ok(self.dev.L.qsf_mp_bce(s.ptr, t.ptr, mk.ptr, ds.ptr, loss.ptr, s.numel, 1.0 / cnt, self.dev.stream),
            "qsf_mp_bce")
        return loss, ds

    def adam(self, g, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_no += 1
        for k, w in self.w.items():
            wb = self.wb.get(k)
            _ok(self.dev.L.qsf_mp_adam(w.ptr, g[k].ptr, self.m[k].ptr, self.v[k].ptr,
                                       wb.ptr if wb is not None else None, w.numel, lr, b1, b2, eps,
                                       self.step_no, self.dev.stream), "qsf_mp_adam")

    def scores(self, Hs, seeds=None):
        """Column logits of each code (numpy, length n each). Unseeded, the inputs are made on
        the device (Dev.pack, equal to pack_batch_bits); with seeds, on the host."""
        X, Ac, Av, mask = self.dev.pack(Hs) if seeds is None else pack_batch_bits(Hs, seeds)
        s, _ = self.forward(X, Ac, Av)
        s = s.numpy()
        return [s[b, : H.shape[1]].copy() for b, H in enumerate(Hs)]

    def save(self, path):
        np.savez(path, **self.params())

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            return cls({k: z[k] for k in z.files})
