"""Learned schedule selector for SpMM and SDDMM.

A 3-layer MLP (hidden 64, fp32) maps cheap sparsity features — log sizes,
the row-length histogram (computed on the device by armsp_row_stats) and the
dense width — to the schedule/tile configuration to launch.  It is trained
on timings from the exhaustive autotuner below, with the library's own HIP
kernels doing every forward, backward and Adam step (armsp_sgemm_f32,
armsp_softmax_xent, armsp_relu_backward, armsp_colsum_f32,
armsp_adam_step), and runs inference in armsp_mlp_infer.  The autotuner is
the fallback when the network is unsure and the yardstick of its regret.

The reference has no counterpart: it hard-codes an N threshold
(gcn/pytorch-custom/spmm_kernel.cu:437-457).

    python -m armsparse_amd.selector --matrices 2000 --out <dir>
builds the corpus, autotunes, trains on 80%, reports held-out regret.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import queue
import sys
import threading
import time
from typing import Optional

import numpy as np
import torch

from . import _lib, data, kernels
from ._lib import Config, lib

NBINS = 12
NFEAT = 5 + NBINS
WIDTHS = (1, 32, 64, 128, 256)
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_weights.json")

# Class 0 of each list is the library's built-in heuristic (AUTO): "the method
# without the network".
SPMM_CONFIGS = [None,
                dict(schedule="row"), dict(schedule="row", group=16),
                dict(schedule="nnz", chunk=256, unroll=8),
                dict(schedule="merge", group=64, chunk=128, unroll=16),
                dict(schedule="merge", group=64, chunk=256, unroll=4),
                dict(schedule="merge", group=64, chunk=256, unroll=16),
                dict(schedule="merge", group=64, chunk=512, unroll=8),
                dict(schedule="merge", chunk=1024, unroll=8),
                dict(schedule="merge", group=32, chunk=512, unroll=8),
                dict(schedule="merge", group=32, vec=2, chunk=256, unroll=4),
                dict(schedule="merge", group=32, vec=2, chunk=512, unroll=8)]
SDDMM_CONFIGS = ([None] + [dict(schedule="merge", group=g, chunk=c) for g in (4, 8, 16) for c in (128, 256, 512)]
                 + [dict(schedule="nnz", group=g, chunk=256) for g in (8, 16)]
                 + [dict(schedule="merge", group=8, chunk=c, unroll=2) for c in (128, 256)])  # column halves
CONFIGS = {"spmm": SPMM_CONFIGS, "sddmm": SDDMM_CONFIGS}


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


# --------------------------------------------------------------- features
def row_stats(rowptr: torch.Tensor, m: int):
    """(hist[NBINS], max row length) from the device kernel armsp_row_stats."""
    out = torch.empty(NBINS + 1, dtype=torch.int32, device=rowptr.device)
    _lib.call("armsp_row_stats", m, _ptr(rowptr), NBINS, _ptr(out), ctypes.c_void_p(out.data_ptr() + 4 * NBINS),
              _stream())
    h = out.cpu().numpy()
    return h[:NBINS].astype(np.int64), int(h[NBINS])


def features(m: int, nnz: int, hist: np.ndarray, maxlen: int, n: int) -> np.ndarray:
    f = np.empty(NFEAT, np.float32)
    f[0] = math.log2(max(m, 1))
    f[1] = math.log2(nnz + 1)
    f[2] = math.log2(nnz / max(m, 1) + 1)
    f[3] = math.log2(maxlen + 1)
    f[4] = math.log2(max(n, 1))
    f[5:] = hist / max(m, 1)
    return f


# ------------------------------------------------------------------- MLP
class MLP:
    """F → H → H → C, ReLU, trained with libarmsparse kernels on the device."""

    def __init__(self, nfeat: int, nclass: int, hidden: int = 64, seed: int = 0, device="cuda"):
        self.F, self.H, self.C = nfeat, hidden, nclass
        self.dev = torch.device(device)
        r = np.random.default_rng(seed)

        def he(fan_in, shape):
            a = math.sqrt(6.0 / fan_in)
            return r.uniform(-a, a, shape).astype(np.float32)

        init = {"W1": he(nfeat, (nfeat, hidden)), "b1": np.zeros(hidden, np.float32),
                "W2": he(hidden, (hidden, hidden)), "b2": np.zeros(hidden, np.float32),
                "W3": he(hidden, (hidden, nclass)), "b3": np.zeros(nclass, np.float32)}
        self.p = {k: torch.from_numpy(v).to(self.dev) for k, v in init.items()}
        self.mu = torch.zeros(nfeat, device=self.dev)
        self.sd = torch.ones(nfeat, device=self.dev)

    # forward of a standardised batch; returns logits and the activations
    def forward(self, xs: torch.Tensor):
        B = xs.shape[0]
        p, s = self.p, _stream()
        h1 = torch.empty(B, self.H, device=self.dev)
        h2 = torch.empty(B, self.H, device=self.dev)
        z3 = torch.empty(B, self.C, device=self.dev)
        _lib.call("armsp_sgemm_f32", B, self.H, self.F, _ptr(xs), self.F, 1, _ptr(p["W1"]), self.H, 1, _ptr(h1),
                  self.H, 0.0, _ptr(p["b1"]), 1, s)
        _lib.call("armsp_sgemm_f32", B, self.H, self.H, _ptr(h1), self.H, 1, _ptr(p["W2"]), self.H, 1, _ptr(h2),
                  self.H, 0.0, _ptr(p["b2"]), 1, s)
        _lib.call("armsp_sgemm_f32", B, self.C, self.H, _ptr(h2), self.H, 1, _ptr(p["W3"]), self.C, 1, _ptr(z3),
                  self.C, 0.0, _ptr(p["b3"]), 0, s)
        return z3, (h1, h2)

    def loss_and_grads(self, xs: torch.Tensor, q: torch.Tensor):
        """Soft-target cross-entropy and parameter gradients, all on libarmsparse kernels."""
        B = xs.shape[0]
        H, C, F = self.H, self.C, self.F
        s = _stream()
        z3, (h1, h2) = self.forward(xs)
        d3 = torch.empty(B, C, device=self.dev)
        loss = torch.empty(1, device=self.dev)
        _lib.call("armsp_softmax_xent", B, C, _ptr(z3), None, _ptr(q), _ptr(d3), _ptr(loss), s)
        g = {k: torch.empty_like(v) for k, v in self.p.items()}
        sg = lambda M, N, K, A, sam, sak, Bm, sbk, sbn, Cm, ldc: _lib.call(  # noqa: E731
            "armsp_sgemm_f32", M, N, K, _ptr(A), sam, sak, _ptr(Bm), sbk, sbn, _ptr(Cm), ldc, 0.0, None, 0, s)
        sg(H, C, B, h2, 1, H, d3, C, 1, g["W3"], C)          # h2ᵀ·d3
        _lib.call("armsp_colsum_f32", B, C, _ptr(d3), _ptr(g["b3"]), s)
        d2 = torch.empty(B, H, device=self.dev)
        sg(B, H, C, d3, C, 1, self.p["W3"], 1, C, d2, H)     # d3·W3ᵀ
        _lib.call("armsp_relu_backward", B * H, _ptr(h2), _ptr(d2), s)
        sg(H, H, B, h1, 1, H, d2, H, 1, g["W2"], H)
        _lib.call("armsp_colsum_f32", B, H, _ptr(d2), _ptr(g["b2"]), s)
        d1 = torch.empty(B, H, device=self.dev)
        sg(B, H, H, d2, H, 1, self.p["W2"], 1, H, d1, H)
        _lib.call("armsp_relu_backward", B * H, _ptr(h1), _ptr(d1), s)
        sg(F, H, B, xs, 1, F, d1, H, 1, g["W1"], H)
        _lib.call("armsp_colsum_f32", B, H, _ptr(d1), _ptr(g["b1"]), s)
        return loss, g

    def fit(self, x: np.ndarray, q: np.ndarray, steps: int = 3000, lr: float = 3e-3, wd: float = 1e-4,
            log_every: int = 0):
        mu = x.mean(0)
        sd = x.std(0) + 1e-3
        self.mu = torch.from_numpy(mu.astype(np.float32)).to(self.dev)
        self.sd = torch.from_numpy(sd.astype(np.float32)).to(self.dev)
        xs = torch.from_numpy(((x - mu) / sd).astype(np.float32)).to(self.dev).contiguous()
        qt = torch.from_numpy(q.astype(np.float32)).to(self.dev).contiguous()
        m = {k: torch.zeros_like(v) for k, v in self.p.items()}
        v = {k: torch.zeros_like(v) for k, v in self.p.items()}
        hist = []
        for step in range(1, steps + 1):
            loss, g = self.loss_and_grads(xs, qt)
            for k in self.p:
                _lib.call("armsp_adam_step", self.p[k].numel(), _ptr(self.p[k]), _ptr(g[k]), _ptr(m[k]), _ptr(v[k]),
                          lr, 0.9, 0.999, 1e-8, wd, step, _stream())
            if log_every and (step % log_every == 0 or step == 1):
                hist.append((step, float(loss)))
        return hist

    def predict(self, x: np.ndarray, logits: bool = False):
        """Classes (and optionally logits) from the HIP inference kernel on raw features."""
        n = x.shape[0]
        xd = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(self.dev)
        cls = torch.empty(n, dtype=torch.int32, device=self.dev)
        lg = torch.empty(n, self.C, device=self.dev)
        p = self.p
        _lib.call("armsp_mlp_infer", n, self.F, self.H, self.C, _ptr(xd), _ptr(self.mu), _ptr(self.sd), _ptr(p["W1"]),
                  _ptr(p["b1"]), _ptr(p["W2"]), _ptr(p["b2"]), _ptr(p["W3"]), _ptr(p["b3"]), _ptr(lg), _ptr(cls),
                  _stream())
        c = cls.cpu().numpy()
        return (c, lg.cpu().numpy()) if logits else c

    def state(self) -> dict:
        st = {k: v.cpu().numpy().tolist() for k, v in self.p.items()}
        st.update(mu=self.mu.cpu().numpy().tolist(), sd=self.sd.cpu().numpy().tolist(), F=self.F, H=self.H, C=self.C)
        return st

    @classmethod
    def from_state(cls, st: dict, device="cuda") -> "MLP":
        net = cls(st["F"], st["C"], st["H"], device=device)
        for k in ("W1", "b1", "W2", "b2", "W3", "b3"):
            net.p[k] = torch.tensor(st[k], dtype=torch.float32, device=net.dev)
        net.mu = torch.tensor(st["mu"], dtype=torch.float32, device=net.dev)
        net.sd = torch.tensor(st["sd"], dtype=torch.float32, device=net.dev)
        return net


def soft_targets(times: np.ndarray, tau: float = 0.03) -> np.ndarray:
    """q_c ∝ exp(-(t_c / t_best - 1) / tau); unsupported configs (inf) get 0."""
    best = np.min(times, axis=1, keepdims=True)
    z = np.where(np.isfinite(times), -(times / best - 1.0) / tau, -np.inf)
    z -= z.max(axis=1, keepdims=True)
    e = np.exp(z)
    return e / e.sum(axis=1, keepdims=True)


# ------------------------------------------------------------- autotuner
class Problem:
    """Device operands of one matrix's timing problems (dense operands cached per width)."""

    def __init__(self, rp, ci, va, m, k, nnz):
        self.rp, self.ci, self.va, self.m, self.k, self.nnz = rp, ci, va, m, k, nnz
        self._dense: dict = {}

    def _operands(self, op: str, n: int):
        key = (op, n)
        if key not in self._dense:
            dev = self.rp.device
            if op == "spmm":
                B = torch.empty(self.k, n, device=dev)
                kernels.fill_uniform_(B, 1, 16)
                self._dense[key] = (B, torch.empty(self.m, n, device=dev))
            else:
                U, V = torch.empty(self.m, n, device=dev), torch.empty(self.k, n, device=dev)
                kernels.fill_uniform_(U, 2, 16)
                kernels.fill_uniform_(V, 3, 16)
                self._dense[key] = (U, V, torch.empty(max(self.nnz, 1), device=dev))
        return self._dense[key]

    def launcher(self, op: str, n: int, cfg):
        c = Config.make(**cfg) if cfg else None
        cref = ctypes.byref(c) if c is not None else None
        s = _stream()
        ops = self._operands(op, n)
        if op == "spmm":
            B, C = ops
            args = (self.m, self.k, n, self.nnz, _ptr(self.rp), _ptr(self.ci), _ptr(self.va), _ptr(B), n, _ptr(C), n,
                    0, cref, s)
            fn = lib.armsp_spmm_f32
        else:
            U, V, out = ops
            args = (self.m, self.k, n, self.nnz, _ptr(self.rp), _ptr(self.ci), _ptr(self.va), _ptr(U), n, _ptr(V), n,
                    _ptr(out), cref, s)
            fn = lib.armsp_sddmm_csr_f32

        def run():
            st = fn(*args)
            if st != 0:
                raise _lib.ArmspError(op, st)
            return c, ops

        return run

    def release(self):
        self._dense.clear()


def time_launch(run, target_us: float = 300.0, max_reps: int = 20) -> float:
    """Mean device time (µs) of back-to-back launches bracketed by events."""
    run()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    run()
    e1.record()
    e1.synchronize()
    t1 = e0.elapsed_time(e1) * 1e3
    reps = int(min(max_reps, max(2, target_us / max(t1, 1.0))))
    e0.record()
    for _ in range(reps):
        run()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) * 1e3 / reps


def autotune(prob: Problem, op: str, n: int, configs=None) -> np.ndarray:
    configs = CONFIGS[op] if configs is None else configs
    t = np.full(len(configs), np.inf)
    for i, cfg in enumerate(configs):
        try:
            t[i] = time_launch(prob.launcher(op, n, cfg))
        except _lib.ArmspError:
            pass
    return t


# ----------------------------------------------------------------- corpus
def corpus_specs(n: int, seed: int = 2024, nnz_cap: int = 1 << 22, big_every: int = 20) -> list:
    """Seeded corpus; every big_every-th matrix is graph-scale (2^18–2^20 rows, up
    to 2^24 nonzeros, uniform or R-MAT) so the regime where B far exceeds the L2
    is represented."""
    r = np.random.default_rng(seed)
    kinds = ["uniform", "rmat", "banded", "blockdiag"]
    specs = []
    for i in range(n):
        if big_every and i % big_every == big_every - 1:
            kind = ["uniform", "rmat"][(i // big_every) % 2]
            m = 1 << int(r.integers(18, 21))
            avg = float(np.exp(r.uniform(math.log(4), math.log(32))))
            avg = min(avg, (1 << 24) / m)
            param = float(r.uniform(0.45, 0.75)) if kind == "rmat" else 0.0
            specs.append(dict(kind=kind, m=m, avg=avg, param=param, seed=seed * 100003 + i))
            continue
        kind = kinds[i % 4]
        m = 1 << int(r.integers(12, 22))
        avg = float(np.exp(r.uniform(math.log(2), math.log(256))))
        param = 0.0
        if kind == "rmat":
            param = float(r.uniform(0.45, 0.75))
        elif kind == "banded":
            param = float(int(np.exp(r.uniform(math.log(8), math.log(512)))))
            avg = min(avg, 2 * param + 1)
        elif kind == "blockdiag":
            param = float(max(1, round(avg / 32)))
            avg = 32 * param
        if m * avg > nnz_cap:
            if kind == "blockdiag":  # ≥ 32 nonzeros per row: fewer blocks, then fewer rows
                param = float(max(1, nnz_cap // (32 * m)))
                avg = 32 * param
                while m * avg > nnz_cap:
                    m //= 2
            else:
                avg = nnz_cap / m
        specs.append(dict(kind=kind, m=m, avg=avg, param=param, seed=seed * 100003 + i))
    return specs


def _pack(feats, times, mids) -> dict:
    return {op: dict(x=np.array(feats[op]), t=np.array(times[op]), mid=np.array(mids[op])) for op in feats}


def save_dataset(path: str, ds: dict) -> None:
    np.savez_compressed(path, **{f"{op}_{k}": v for op, d in ds.items() for k, v in d.items()})


def build_dataset(specs: list, widths=WIDTHS, log=print, checkpoint: str = "") -> dict:
    """Autotune every (matrix, width) for SpMM and SDDMM; matrices are generated
    on the host by a background thread while the device times the previous one."""
    q: "queue.Queue" = queue.Queue(maxsize=4)

    def producer():
        for sp in specs:
            q.put((sp, data.corpus_csr(sp["kind"], sp["m"], sp["avg"], sp["param"], sp["seed"])))
        q.put(None)

    threading.Thread(target=producer, daemon=True).start()
    feats = {"spmm": [], "sddmm": []}
    times = {"spmm": [], "sddmm": []}
    mids = {"spmm": [], "sddmm": []}
    t0 = time.time()
    for idx in range(len(specs)):
        item = q.get()
        if item is None:
            break
        sp, a = item
        dev = torch.device("cuda")
        rp = torch.from_numpy(a.rowptr).to(dev)
        ci = torch.from_numpy(a.colind).to(dev)
        va = torch.from_numpy(a.values).to(dev)
        prob = Problem(rp, ci, va, a.shape[0], a.shape[1], a.nnz)
        hist, mx = row_stats(rp, a.shape[0])
        for op in ("spmm", "sddmm"):
            for n in widths:
                feats[op].append(features(a.shape[0], a.nnz, hist, mx, n))
                times[op].append(autotune(prob, op, n))
                mids[op].append(idx)
            prob.release()
        del prob, rp, ci, va
        if (idx + 1) % 25 == 0 or idx + 1 == len(specs):
            log(f"[selector] {idx + 1}/{len(specs)} matrices autotuned, {time.time() - t0:.0f}s", flush=True)
        if checkpoint and (idx + 1) % 100 == 0:
            save_dataset(checkpoint, _pack(feats, times, mids))
    return _pack(feats, times, mids)


# ------------------------------------------------------------ evaluation
def regret_stats(t_sel: np.ndarray, t_best: np.ndarray) -> dict:
    r = t_sel / t_best - 1.0
    return dict(mean=float(r.mean()), median=float(np.median(r)), p90=float(np.percentile(r, 90)),
                max=float(r.max()), total=float(t_sel.sum() / t_best.sum() - 1.0))


def train_eval(ds: dict, seed: int = 7, steps: int = 3000, p_min: float = 0.5, log=print):
    """Train one MLP per op on 80% of the matrices; report held-out regret."""
    report, nets = {}, {}
    for op, d in ds.items():
        mids = np.unique(d["mid"])
        r = np.random.default_rng(seed)
        test_m = set(r.permutation(mids)[: max(1, len(mids) // 5)].tolist())
        te = np.array([m in test_m for m in d["mid"]])
        ok = np.isfinite(d["t"]).any(axis=1)
        tr, te = ok & ~te, ok & te
        x, t = d["x"], np.where(np.isfinite(d["t"]), d["t"], np.inf)
        net = MLP(NFEAT, t.shape[1], seed=seed)
        hist = net.fit(x[tr], soft_targets(t[tr]), steps=steps, log_every=max(1, steps // 10))
        cls, lg = net.predict(x[te], logits=True)
        best = t[te].min(axis=1)
        sel = t[te][np.arange(te.sum()), cls]
        pr = np.exp(lg - lg.max(1, keepdims=True))
        pr /= pr.sum(1, keepdims=True)
        unsure = pr.max(1) < p_min
        sel_fb = np.where(unsure, best, sel)
        fixed = int(np.argmin(np.where(np.isfinite(t[tr]), t[tr], 1e30).sum(0)))
        report[op] = {
            "train_samples": int(tr.sum()), "test_samples": int(te.sum()),
            "loss_curve": hist,
            "mlp": regret_stats(sel, best),
            "mlp_with_autotune_fallback": dict(regret_stats(sel_fb, best), p_min=p_min,
                                               fallback_rate=float(unsure.mean())),
            "heuristic_auto": regret_stats(t[te][:, 0], best),
            "best_fixed_config": dict(regret_stats(t[te][:, fixed], best), config=CONFIGS[op][fixed]),
            "mlp_accuracy": float((cls == t[te].argmin(1)).mean()),
        }
        nets[op] = net
        log(f"[selector] {op}: mlp regret {report[op]['mlp']}, heuristic {report[op]['heuristic_auto']}")
    return report, nets


# ---------------------------------------------------------------- runtime
class Selector:
    """kernels.set_selector hook: features on the device, MLP inference in HIP,
    autotune fallback when the network's top probability is below p_min."""

    def __init__(self, weights: Optional[dict] = None, p_min: float = 0.0):
        if weights is None:
            with open(WEIGHTS_PATH) as f:
                weights = json.load(f)
        self.nets = {op: MLP.from_state(st) for op, st in weights.items() if op in CONFIGS}
        self.p_min = p_min
        self.cache: dict = {}

    def __call__(self, op, rowptr, m, nnz, n):
        if op not in self.nets or not torch.is_tensor(rowptr) or not rowptr.is_cuda:
            return None
        hist, mx = row_stats(rowptr, m)
        f = features(m, nnz, hist, mx, n)
        key = (op, f.tobytes())
        if key not in self.cache:
            cls, lg = self.nets[op].predict(f[None, :], logits=True)
            self.cache[key] = int(cls[0])
        cfg = CONFIGS[op][self.cache[key]]
        return None if cfg is None else Config.make(**cfg)


def install(p_min: float = 0.0) -> Selector:
    sel = Selector(p_min=p_min)
    kernels.set_selector(sel)
    return sel


NATIVE_OPS = {"spmm": 0, "sddmm": 1}  # ARMSP_SELECTOR_SPMM / _SDDMM


def install_native(weights: Optional[dict] = None) -> None:
    """Install the networks in the library itself (armsp_selector_set): every AUTO
    SpMM / SDDMM call, from Python, the C ABI or the command-line drivers, then runs
    the configuration the network picks (features and inference on the device, one
    read-back per matrix, cached)."""
    if weights is None:
        with open(WEIGHTS_PATH) as f:
            weights = json.load(f)
    for op, code in NATIVE_OPS.items():
        st = weights[op]
        F, H, C = int(st["F"]), int(st["H"]), int(st["C"])
        arrs = [np.ascontiguousarray(np.asarray(st[k], np.float32).reshape(-1))
                for k in ("mu", "sd", "W1", "b1", "W2", "b2", "W3", "b3")]
        classes = (Config * C)(*[Config.make(**c) if c else Config() for c in CONFIGS[op]])
        _lib.call("armsp_selector_set", code, F, H, C, *[ctypes.c_void_p(a.ctypes.data) for a in arrs], classes)


def uninstall_native() -> None:
    for code in NATIVE_OPS.values():
        _lib.call("armsp_selector_clear", code)


def native_choice(op: str, rowptr: torch.Tensor, m: int, nnz: int, n: int, element_bytes: int = 4) -> Optional[dict]:
    """The configuration the library's AUTO runs for this matrix and width (None: the heuristic)."""
    out = Config()
    _lib.call("armsp_selector_choose", NATIVE_OPS[op], m, nnz, n, element_bytes, _ptr(rowptr), ctypes.byref(out),
              _stream())
    return None if not any((out.schedule, out.group, out.vec, out.chunk, out.unroll)) else out.as_dict()


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--matrices", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--out", default="selector_out")
    ap.add_argument("--dataset", default="", help="reuse a saved dataset .npz instead of autotuning")
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    if a.dataset:
        z = np.load(a.dataset)
        ds = {op: dict(x=z[f"{op}_x"], t=z[f"{op}_t"], mid=z[f"{op}_mid"]) for op in ("spmm", "sddmm")}
    else:
        path = os.path.join(a.out, "dataset.npz")
        ds = build_dataset(corpus_specs(a.matrices, a.seed), log=lambda *x, **k: print(*x, **k, file=sys.stderr),
                           checkpoint=path)
        save_dataset(path, ds)
    report, nets = train_eval(ds, steps=a.steps, log=lambda *x: print(*x, file=sys.stderr, flush=True))
    report["configs"] = {op: c for op, c in CONFIGS.items()}
    with open(os.path.join(a.out, "report.json"), "w") as f:
        json.dump(report, f, indent=1)
    with open(os.path.join(a.out, "selector_weights.json"), "w") as f:
        json.dump({op: n.state() for op, n in nets.items()}, f)
    print(json.dumps({op: report[op]["mlp"] for op in ("spmm", "sddmm")}))


if __name__ == "__main__":
    main()
