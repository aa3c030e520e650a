"""Benchmark of the hot path on one MI355X.

Headline (the JSON `value`): FP32 CSR SpMM C = A·B on the 1,048,576² R-MAT
matrix (a=.57, b=c=.19, d=.05; 2^24 distinct nonzeros, int32 CSR, values
U(-1,1)) with a dense 1,048,576×128 fp32 B, inputs resident in HBM.  GB/s
counts algorithmic bytes: 4(M+1) + 8·nnz + 4·K·N + 4·M·N (each array once).

Also reported in the same line (`extras`): SDDMM (Kd 32/128) and SpMV on the
same matrix, bf16 / fp8-e4m3 B SpMM, and the GCN fwd+bwd ms/epoch fp32 vs
bf16; `roofline` from HIP events on the kernel's own stream; `cpu_baseline`
from the oracle's restatement of the reference CPU kernel (rowBaseSpmm3,
OpenMP) on this host's cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--only spmm]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def log(*a):
    print(*a, file=sys.stderr, flush=True)


ROOT = os.path.dirname(os.path.abspath(__file__))
# Sources of the headline SpMM kernel: a committed counter summary is used only while
# they are unchanged (its stamp is their hash).
SPMM_SOURCES = ["armsparse_amd/csrc/spmm.hip", "armsparse_amd/csrc/partition.hip", "armsparse_amd/csrc/common.hpp",
                "include/armsp.h"]
GEMM_SOURCES = ["armsparse_amd/csrc/gemm.hip", "armsparse_amd/csrc/common.hpp", "include/armsp.h"]


def source_stamp(paths):
    """Hash of the sources' code: C and C++ comments and whitespace are left out, so
    an edit of the documentation alone keeps a counter summary current."""
    import hashlib
    import re
    h = hashlib.sha256()
    for p in paths:
        with open(os.path.join(ROOT, p), encoding="utf-8") as f:
            text = f.read()
        text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
        text = re.sub(r"//[^\n]*", " ", text)
        h.update(" ".join(text.split()).encode())
    return h.hexdigest()[:16]


def stamped_profile(name, sources):
    """A committed counter summary (profiles/<name>) if its stamp matches the current
    sources, else None (stale figures are not reported)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    rec = json.load(open(path))
    return rec if rec.get("source_stamp") == source_stamp(sources) else None


def sample_rows(rowptr, n_random=4096, n_longest=64, seed=1):
    """Sorted row ids: the longest rows and seeded random ones (FP64 checks on a sample)."""
    import numpy as np
    M = len(rowptr) - 1
    longest = np.argsort(np.diff(rowptr))[-min(n_longest, M):]
    rnd = np.random.default_rng(seed).choice(M, min(n_random, M), replace=False)
    return np.unique(np.concatenate([longest, rnd]))


def sub_csr(rowptr, colind, val, rows):
    import numpy as np
    lens = (rowptr[rows + 1] - rowptr[rows]).astype(np.int64)
    rp = np.zeros(len(rows) + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    idx = np.concatenate([np.arange(rowptr[r], rowptr[r + 1]) for r in rows])
    return rp, colind[idx], val[idx]


def spmm_rows_err(rowptr, colind, val, Bh, got_rows, rows):
    """Max norm-wise relative error per row of got_rows (= C[rows]) against the FP64
    product of those rows of A with Bh (scipy in float64; independent of the kernels)."""
    import numpy as np
    import scipy.sparse as sp
    srp, sci, sva = sub_csr(rowptr, colind, val, rows)
    uc, inv = np.unique(sci, return_inverse=True)
    sub = sp.csr_matrix((sva.astype(np.float64), inv.astype(np.int64), srp.astype(np.int64)), shape=(len(rows), len(uc)))
    ref = sub @ Bh[uc].astype(np.float64)
    num = np.linalg.norm(got_rows.astype(np.float64) - ref, axis=1)
    den = np.linalg.norm(ref, axis=1)
    return float(np.max(np.where(den > 0, num / np.where(den > 0, den, 1), num)))


def spmv_rows_err(rowptr, colind, val, xh, got_rows, rows):
    """Max |y - y64| / Σ|a·x| over the sampled rows (the dot product's own scale)."""
    import numpy as np
    import scipy.sparse as sp
    srp, sci, sva = sub_csr(rowptr, colind, val, rows)
    sub = sp.csr_matrix((sva.astype(np.float64), sci.astype(np.int64), srp.astype(np.int64)), shape=(len(rows), len(xh)))
    ref = sub @ xh.astype(np.float64)
    scale = abs(sub) @ np.abs(xh.astype(np.float64))
    return float(np.max(np.abs(got_rows - ref) / np.maximum(scale, 1e-300)))


def sddmm_rows_err(rowptr, colind, val, Uh, Vh, got, rows):
    """Max |out - out64| / Σ|a·u·v| over the nonzeros of the sampled rows."""
    import numpy as np
    worst = 0.0
    for r in rows:
        lo, hi = int(rowptr[r]), int(rowptr[r + 1])
        if lo == hi:
            continue
        terms = val[lo:hi, None].astype(np.float64) * Uh[r].astype(np.float64) * Vh[colind[lo:hi]].astype(np.float64)
        ref, scale = terms.sum(1), np.abs(terms).sum(1)
        worst = max(worst, float(np.max(np.abs(got[lo:hi] - ref) / np.maximum(scale, 1e-300))))
    return worst


def host_info():
    """Cores this process may use and what the host has (physical cores from sysfs)."""
    cores = set()
    base = "/sys/devices/system/cpu"
    for d in os.listdir(base):
        if d.startswith("cpu") and d[3:].isdigit():
            try:
                pkg = open(f"{base}/{d}/topology/physical_package_id").read().strip()
                core = open(f"{base}/{d}/topology/core_id").read().strip()
                cores.add((pkg, core))
            except OSError:
                pass
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()
        quota = None if q == "max" else round(int(q) / int(per), 2)
    except (OSError, ValueError):
        pass
    return {"physical_cores": len(cores) or None, "logical_cpus": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)), "cgroup_cpu_limit": quota}


_HOST = None  # host_info() as main() found it at start


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--only", default="", help="comma list: spmm,pitched,sddmm,spmv,mixed,widths,shards,small,gcn,gcn_scaled,cpu (default all); "
                        "regimes (column-placement variants of the headline matrix), probe (the B-row gathers "
                        "alone on those column patterns), pitches (B at several row pitches) and spmv_variants "
                        "(SpMV schedules) run only when named")
    p.add_argument("--schedule", default="", help="override the SpMM schedule (row|nnz|merge)")
    p.add_argument("--chunk", type=int, default=0)
    p.add_argument("--unroll", type=int, default=0)
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--copy-probe", action="store_true",
                   help="also run a plain 512 MiB device copy (calibrates rocprof FETCH_SIZE/WRITE_SIZE)")
    return p.parse_args()


def cpu_baselines(args, A, B, headline_bytes):
    """The oracle's restatements of the reference CPU kernels (OpenMP, C) timed on this
    host: rowBaseSpmm3 on the headline workload (the reported value), plus
    sparse_dense_vector on the same matrix's blocked CSR, sparse_matrix_mul3 on the
    `./SpMM 4096 4096 4096` case and a PubMed-size GCN training epoch through the
    restated CPU op.  Threads are bound spread over cores; every call is timed, and
    the mean and median reported."""
    import numpy as np

    import oracle
    from armsparse_amd import data

    threads = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    budget = args.cpu_seconds

    def calls(fn, min_calls=3, max_calls=30, seconds=budget):
        fn()  # warm-up
        ts = []
        t0 = time.perf_counter()
        while len(ts) < max_calls and (len(ts) < min_calls or time.perf_counter() - t0 < seconds):
            t = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t)
        return ts

    def stats(ts, nbytes=None):
        import statistics
        d = {"calls": len(ts), "mean_ms": round(statistics.fmean(ts) * 1e3, 2),
             "median_ms": round(statistics.median(ts) * 1e3, 2)}
        if nbytes:
            d["GBps_mean"] = round(nbytes / statistics.fmean(ts) / 1e9, 2)
        return d

    M, N = A.shape[0], B.shape[1]
    Bh = B.cpu().numpy()
    Ch = np.zeros((M, N), np.float32)
    ts = calls(lambda: oracle.rowBaseSpmm3(A.rowptr, A.colind, A.values, Bh, Ch, threads=threads))
    head = stats(ts, headline_bytes)
    others = {}
    # sparse_dense_vector: y += A·x on the blocked CSR of the same matrix
    v = data.csr_to_vcsr(A)
    x = data.host_uniform(v.shape[1], data.CFG_SEED + 1, data.STREAM_DENSE + 1)
    y = np.zeros(M, np.float32)
    ts = calls(lambda: oracle.sparse_dense_vector(x, y, v.rowptr, v.colind, v.values, threads=threads),
               seconds=budget / 4)
    others["sparse_dense_vector_rmat1m"] = stats(ts, 4 * (M + 1) + 8 * v.nnz + 4 * v.shape[1] + 8 * M)
    del v
    # sparse_matrix_mul3: the literal ./SpMM 4096 4096 4096 product
    As = data.uniform_csr(4096, 4096, 167_772, data.CFG_SEED + 0)
    Bs = data.uniform_csr(4096, 4096, 167_772, data.CFG_SEED + 5)
    Cs = np.zeros((4096, 4096), np.float32)
    a3, b3 = (As.rowptr, As.colind, As.values), (Bs.rowptr, Bs.colind, Bs.values)
    ts = calls(lambda: oracle.sparse_matrix_mul3(a3, b3, Cs, threads=threads), seconds=budget / 4)
    others["sparse_matrix_mul3_4096"] = stats(ts, 2 * (4 * 4097 + 8 * 167_772) + 8 * 4096 * 4096)
    others["gcn_pubmed_epoch"] = stats(_cpu_gcn_epochs(oracle, threads, budget / 4))
    return {"value": head["GBps_mean"], "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"full 1M R-MAT N=128, rowBaseSpmm3 (C += A·B, accumulate bytes), 1 warm-up + "
                      f"{head['calls']} timed calls, mean {head['mean_ms']} ms, median {head['median_ms']} ms",
            "median_GBps": round(headline_bytes / (head["median_ms"] / 1e3) / 1e9, 2),
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
            "host": _HOST or host_info(), "others": others}


def _cpu_gcn_epochs(oracle, threads, seconds):
    """Times of PubMed-size training epochs on the CPU: the reference's GCNConv maths
    (gcn/pytorch-custom/op.py:102-124) with torch's CPU ops and the restated CPU op
    (csr_spmm, spmm.cpp:31-71) for the aggregation and its transpose in backward."""
    import numpy as np
    import torch
    import torch.nn.functional as F

    from armsparse_amd import data

    g = data.gcn_graph()
    n = g.n
    cp, ri, vt = oracle.csr_to_csc(g.rowptr, g.colind, g.value_csr, n)

    class Agg(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x):
            return torch.from_numpy(oracle.csr_spmm(g.rowptr, g.colind, g.value_csr, x.detach().numpy(), threads))

        @staticmethod
        def backward(ctx, gy):
            return torch.from_numpy(oracle.csr_spmm(cp, ri, vt, gy.contiguous().numpy(), threads))

    r = np.random.default_rng(0)
    xd = np.where(r.random((n, 500)) < 0.01, r.random((n, 500)), 0.0).astype(np.float32)
    xd /= np.maximum(xd.sum(1, keepdims=True), 1e-30)
    x = torch.from_numpy(xd)
    y = torch.from_numpy(r.integers(0, 3, n))
    train = torch.from_numpy(r.choice(n, 60, replace=False))
    deg = torch.from_numpy(np.diff(g.rowptr).astype(np.float32))
    norm = (1 / torch.sqrt(deg)).unsqueeze(1)
    torch.manual_seed(0)

    def glorot(a, b):
        lim = (6.0 / (a + b)) ** 0.5
        return torch.empty(a, b).uniform_(-lim, lim).requires_grad_(True)

    W1, W2 = glorot(500, 64), glorot(64, 3)
    b1, b2 = torch.zeros(64, requires_grad=True), torch.zeros(3, requires_grad=True)
    opt = torch.optim.Adam([dict(params=[W1, b1], weight_decay=5e-4), dict(params=[W2, b2], weight_decay=0)], lr=0.01)
    torch.set_num_threads(threads)

    def epoch():
        opt.zero_grad()
        h = F.relu(Agg.apply((x @ W1) * norm) * norm + b1)
        h = F.dropout(h, 0.5, training=True)
        out = F.log_softmax(Agg.apply((h @ W2) * norm) * norm + b2, dim=1)
        F.nll_loss(out[train], y[train]).backward()
        opt.step()

    for _ in range(3):
        epoch()
    ts = []
    t0 = time.perf_counter()
    while len(ts) < 50 and (len(ts) < 5 or time.perf_counter() - t0 < seconds):
        t = time.perf_counter()
        epoch()
        ts.append(time.perf_counter() - t)
    return ts


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1 (no multi-GPU path)"}))
        sys.exit(2)
    # the CPU baseline's OpenMP threads spread over the cores (read when an OpenMP
    # runtime starts, so set before anything loads one)
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    os.environ.setdefault("OMP_PLACES", "cores")
    global _HOST
    _HOST = host_info()  # before an OpenMP runtime binds this thread to one core
    import numpy as np
    import torch

    from armsparse_amd import _lib, data, kernels

    only = set(filter(None, args.only.split(","))) or {"spmm", "pitched", "sddmm", "spmv", "mixed", "widths", "shards",
                                                        "small", "gcn", "gcn_scaled", "cpu"}
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream(dev)

    t0 = time.time()
    A = data.rmat_1m()
    log(f"[bench] R-MAT generated in {time.time() - t0:.1f}s: M={A.shape[0]} nnz={A.nnz}")
    M, K = A.shape
    nnz = A.nnz
    rp = torch.from_numpy(A.rowptr).to(dev)
    ci = torch.from_numpy(A.colind).to(dev)
    va = torch.from_numpy(A.values).to(dev)
    N = 128
    B = torch.empty((K, N), device=dev)
    kernels.fill_uniform_(B, data.CFG_SEED + 2, data.STREAM_DENSE)
    C = torch.empty((M, N), device=dev)

    cfg = None
    if args.schedule or args.chunk or args.unroll:
        cfg = _lib.Config.make(args.schedule or "merge", chunk=args.chunk, unroll=args.unroll)

    def timed(fn, steps, warmup, families):
        """Wall seconds of `steps` calls; with `families`, the library's HIP-event
        kernel times too (recording those events costs host time per launch, so it
        is switched on only when asked for)."""
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        _lib.lib.armsp_timing_reset()
        _lib.lib.armsp_timing_enable(1 if families else 0)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        _lib.lib.armsp_timing_enable(0)
        fam = {f: _lib.kernel_time(f) for f in families}
        return el, fam

    def spmm_bytes(sb, scales=False, accumulate=False):
        return 4 * (M + 1) + 8 * nnz + sb * K * N + (4 * K if scales else 0) + 4 * M * N * (2 if accumulate else 1)

    if args.copy_probe:  # a copy kernel of known size: 512 MiB read + 512 MiB written per launch
        src = torch.empty(128 << 20, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # ------------------------------------------------------------ SpMM fp32
    fn = lambda: kernels.spmm(rp, ci, va, B, out=C, config=cfg)  # noqa: E731
    el, fam = timed(fn, args.steps, args.warmup, ["spmm", "spmm_fixup"])
    nbytes = spmm_bytes(4)
    value = nbytes * args.steps / el / 1e9
    k_ms, k_n = fam["spmm"]
    k_avg = k_ms / max(k_n, 1) / 1e3
    achieved = nbytes / k_avg / 1e9
    f_ms, f_n = fam["spmm_fixup"]
    log(f"[bench] spmm fp32: {value:.1f} GB/s end-to-end, main kernel {k_avg * 1e6:.1f} us "
        f"({achieved:.1f} GB/s), fixup {f_ms / max(f_n, 1) * 1e3:.1f} us")
    traffic = None
    trec = stamped_profile("spmm_rmat1m_n128_traffic.json", SPMM_SOURCES) if cfg is None else None
    if trec is not None:
        traffic = trec.get("hbm_bytes_per_launch")
    else:
        log("[bench] no counter summary for the current SpMM sources: roofline.traffic is null")
    extras = {"spmm_main_kernel_us": round(k_avg * 1e6, 2),
              "spmm_fixup_kernel_us": round(f_ms / max(f_n, 1) * 1e3, 2),
              "gflops": round(2.0 * nnz * N * args.steps / el / 1e9, 1)}
    steps_x = max(10, args.steps // 2)
    # error of every timed op against FP64 (scipy/numpy in float64) on a sample of rows:
    # 4,096 seeded random rows plus the 16 longest
    rows = sample_rows(A.rowptr, 4096, 16)
    rows_t = torch.from_numpy(rows).to(dev).long()
    errs = {}
    Bh = B.cpu().numpy()
    errs["spmm_fp32_n128"] = spmm_rows_err(A.rowptr, A.colind, A.values, Bh, C[rows_t].cpu().numpy(), rows)
    log(f"[bench] spmm fp32 max row rel err vs FP64 (sampled rows): {errs['spmm_fp32_n128']:.2e}")
    if cfg is None:  # the same call with the learned selector deciding AUTO inside the library
        from armsparse_amd import selector

        selector.install_native()
        chosen = selector.native_choice("spmm", rp, M, nnz, N)
        el_s, fam_s = timed(fn, max(10, args.steps // 2), 2, ["spmm"])
        selector.uninstall_native()
        extras["spmm_native_selector_config"] = chosen or "heuristic"
        extras["spmm_native_selector_main_kernel_us"] = round(fam_s["spmm"][0] / max(fam_s["spmm"][1], 1) * 1e3, 2)
        extras["spmm_native_selector_GBps"] = round(nbytes * max(10, args.steps // 2) / el_s / 1e9, 1)

    if cfg is None and "pitched" in only:  # the same B at the gather pitch (kernels.row_pitch: 640 bytes a row)
        pb = kernels.row_pitch(N, 4)
        Bp = torch.empty((K, pb), device=dev)[:, :N]
        Bp.copy_(B)
        Cp = torch.empty_like(C)
        el_p, fam_p = timed(lambda: kernels.spmm(rp, ci, va, Bp, out=Cp), max(10, args.steps // 2), 2,
                            ["spmm", "spmm_fixup"])
        kernels.spmm(rp, ci, va, B, out=C)  # AUTO picks the same kernel for both layouts: the same bits
        extras["spmm_pitched_B_bitwise_equal"] = bool(torch.equal(Cp, C))
        errs["spmm_fp32_n128_pitched_B"] = spmm_rows_err(A.rowptr, A.colind, A.values, Bh, Cp[rows_t].cpu().numpy(),
                                                         rows)
        extras["spmm_pitch_bytes"] = pb * 4
        extras["spmm_pitched_B_main_kernel_us"] = round(fam_p["spmm"][0] / max(fam_p["spmm"][1], 1) * 1e3, 2)
        extras["spmm_pitched_B_GBps"] = round(nbytes * max(10, args.steps // 2) / el_p / 1e9, 1)
        log(f"[bench] spmm fp32 with B rows at a {pb * 4}-byte pitch: main kernel "
            f"{extras['spmm_pitched_B_main_kernel_us']} us, {extras['spmm_pitched_B_GBps']} GB/s end-to-end")
        del Bp, Cp

    if "mixed" in only:
        Bb = kernels.to_bf16(B)
        el_b, fam_b = timed(lambda: kernels.spmm(rp, ci, va, Bb, out=C, config=cfg), steps_x, 3, ["spmm"])
        q, sc = kernels.quantize_fp8_rows(B)
        el_q, fam_q = timed(lambda: kernels.spmm(rp, ci, va, q, dense_scale=sc, out=C, config=cfg), steps_x, 3,
                            ["spmm"])
        extras["spmm_bf16_GBps"] = round(spmm_bytes(2) * steps_x / el_b / 1e9, 1)
        extras["spmm_fp8_GBps"] = round(spmm_bytes(1, scales=True) * steps_x / el_q / 1e9, 1)
        extras["spmm_bf16_us"] = round(el_b / steps_x * 1e6, 1)
        extras["spmm_fp8_us"] = round(el_q / steps_x * 1e6, 1)
        kernels.spmm(rp, ci, va, Bb, out=C, config=cfg)
        errs["spmm_bf16_n128_vs_dequantised"] = spmm_rows_err(A.rowptr, A.colind, A.values, Bb.float().cpu().numpy(),
                                                              C[rows_t].cpu().numpy(), rows)
        errs["spmm_bf16_n128_vs_fp32_B"] = spmm_rows_err(A.rowptr, A.colind, A.values, Bh, C[rows_t].cpu().numpy(),
                                                         rows)
        kernels.spmm(rp, ci, va, q, dense_scale=sc, out=C, config=cfg)
        errs["spmm_fp8_n128_vs_dequantised"] = spmm_rows_err(A.rowptr, A.colind, A.values,
                                                             kernels.dequantize_fp8_rows(q, sc).cpu().numpy(),
                                                             C[rows_t].cpu().numpy(), rows)
        errs["spmm_fp8_n128_vs_fp32_B"] = spmm_rows_err(A.rowptr, A.colind, A.values, Bh, C[rows_t].cpu().numpy(),
                                                        rows)
        del Bb, q, sc
    if "regimes" in only:  # not in the default set: what bounds the gathers (see the counter summaries)
        rng = np.random.default_rng(0)
        c64 = A.colind.astype(np.int64)
        variants = {"rmat": A.colind, "rmat_mod64k": (c64 % 65536).astype(np.int32),
                    "uniform": rng.integers(0, K, nnz).astype(np.int32),
                    "rmat_oddmul": ((c64 * 2654435761) & (K - 1)).astype(np.int32)}
        row_of = np.repeat(np.arange(M, dtype=np.int64), np.diff(A.rowptr))
        for name, cv in variants.items():
            order = np.argsort(row_of * K + cv, kind="stable")  # keep every row's columns sorted
            vci = torch.from_numpy(np.ascontiguousarray(cv[order])).to(dev)
            vva = torch.from_numpy(np.ascontiguousarray(A.values[order])).to(dev)
            el_r, fam_r = timed(lambda: kernels.spmm(rp, vci, vva, B, out=C), max(2, args.steps // 5), 1, ["spmm"])
            extras[f"regime_{name}_kernel_us"] = round(fam_r["spmm"][0] / max(fam_r["spmm"][1], 1) * 1e3, 1)
            log(f"[bench] regime {name}: {extras[f'regime_{name}_kernel_us']} us")
            del vci, vva
    if "probe" in only:  # not in the default set: the B-row gathers alone (armsp_probe_gather_f32)
        rng = np.random.default_rng(0)
        c64 = A.colind.astype(np.int64)
        pats = {"rmat": A.colind, "rmat_oddmul": ((c64 * 2654435761) & (K - 1)).astype(np.int32),
                "uniform": rng.integers(0, K, nnz).astype(np.int32), "rmat_mod64k": (c64 % 65536).astype(np.int32)}
        pout = torch.empty(((nnz + 511) // 512 + 4) * 64 * 8, device=dev)
        for pname, pidx in pats.items():
            di = torch.from_numpy(np.ascontiguousarray(pidx)).to(dev)
            for grp, vec in ((32, 2), (16, 4), (64, 2)):
                call = lambda: _lib.call("armsp_probe_gather_f32", nnz, kernels._ptr(di), kernels._ptr(B), K, N,  # noqa: E731
                                         N, grp, vec, kernels._ptr(pout), kernels._stream(dev))
                el_p, fam_p = timed(call, 10, 2, ["probe"])
                us = fam_p["probe"][0] / max(fam_p["probe"][1], 1) * 1e3
                extras[f"probe_{pname}_g{grp}v{vec}_us"] = round(us, 1)
                extras[f"probe_{pname}_g{grp}v{vec}_TBps"] = round(nnz * N * 4 / us / 1e6, 2)
                log(f"[bench] gather probe {pname} group {grp} vec {vec}: {us:.1f} us, {nnz * N * 4 / us / 1e6:.2f} TB/s")
            if pname in ("rmat", "rmat_oddmul"):  # the same gathers with B rows at the gather pitch
                pb = kernels.row_pitch(N, 4)
                Bp = torch.empty((K, pb), device=dev)
                callp = lambda: _lib.call("armsp_probe_gather_f32", nnz, kernels._ptr(di), kernels._ptr(Bp), K, pb,  # noqa: E731
                                          N, 32, 2, kernels._ptr(pout), kernels._stream(dev))
                _, fam_p = timed(callp, 10, 2, ["probe"])
                us = fam_p["probe"][0] / max(fam_p["probe"][1], 1) * 1e3
                extras[f"probe_{pname}_g32v2_pitch{pb * 4}_us"] = round(us, 1)
                log(f"[bench] gather probe {pname} group 32 vec 2, {pb * 4}-byte pitch: {us:.1f} us")
                del Bp
            del di
    if "pitches" in only:  # not in the default set: the same B values at several row pitches
        for dt, n2, pitches in (("fp32", 128, (128, 144, 160, 192, 256)), ("fp32", 64, (64, 96)),
                                ("fp32", 256, (256, 288)), ("bf16", 256, (256, 320))):
            esz = 4 if dt == "fp32" else 2
            C2 = torch.empty((M, n2), device=dev)
            for pe in pitches:
                Bq = torch.empty((K, pe), device=dev)
                kernels.fill_uniform_(Bq, data.CFG_SEED + 2, data.STREAM_DENSE)
                if dt == "bf16":
                    Bq = kernels.to_bf16(Bq)
                Bv = Bq[:, :n2]
                _, fam_q = timed(lambda: kernels.spmm(rp, ci, va, Bv, out=C2), 8, 2, ["spmm"])
                us = fam_q["spmm"][0] / max(fam_q["spmm"][1], 1) * 1e3
                extras[f"pitch_spmm_{dt}_n{n2}_{pe * esz}B_us"] = round(us, 1)
                log(f"[bench] spmm {dt} N={n2} ({n2 * esz}-byte rows) at a {pe * esz}-byte pitch: main kernel {us:.1f} us")
                del Bq, Bv
            del C2
    if "widths" in only:  # the same matrix at N = 64 and 256 (fp32 and bf16)
        for n2 in (64, 256):
            B2 = torch.empty((K, n2), device=dev)
            kernels.fill_uniform_(B2, data.CFG_SEED + 2, data.STREAM_DENSE)
            C2 = torch.empty((M, n2), device=dev)
            for nm, BB, sb in (("fp32", B2, 4), ("bf16", kernels.to_bf16(B2), 2)):
                el_w, _ = timed(lambda: kernels.spmm(rp, ci, va, BB, out=C2), 10, 2, [])
                byt = 4 * (M + 1) + 8 * nnz + sb * K * n2 + 4 * M * n2
                extras[f"spmm_n{n2}_{nm}_us"] = round(el_w / 10 * 1e6, 1)
                extras[f"spmm_n{n2}_{nm}_GBps"] = round(byt * 10 / el_w / 1e9, 1)
                ref_b = (BB.float() if nm == "bf16" else BB).cpu().numpy()
                errs[f"spmm_{nm}_n{n2}"] = spmm_rows_err(A.rowptr, A.colind, A.values, ref_b,
                                                         C2[rows_t].cpu().numpy(), rows)
            del B2, C2
    if "shards" in only:  # one GPU's share of an 8-way row split: heaviest and median 131,072-row blocks
        srows = M // 8
        per = [int(A.rowptr[(i + 1) * srows] - A.rowptr[i * srows]) for i in range(8)]
        order = sorted(range(8), key=lambda i: per[i])
        for tag, i in (("heavy", order[-1]), ("median", order[3])):
            lo, hi = int(A.rowptr[i * srows]), int(A.rowptr[(i + 1) * srows])
            srp = (rp[i * srows:(i + 1) * srows + 1] - lo).contiguous()
            sci, sva = ci[lo:hi], va[lo:hi]
            Cs = C[i * srows:(i + 1) * srows]
            el_s, _ = timed(lambda: kernels.spmm(srp, sci, sva, B, out=Cs), 10, 2, [])
            extras[f"shard_{tag}_nnz"] = hi - lo
            extras[f"shard_{tag}_us"] = round(el_s / 10 * 1e6, 1)
        extras["shard_rows"] = srows
    if "small" in only:  # the ./SpMM 4096 case and the 65,536² 0.1% case (latency-bound, reported anyway)
        As, Bs, xs = data.spmm_small()
        srp, sci, sva, sB, sx = (torch.from_numpy(np.ascontiguousarray(t)).to(dev) for t in
                                 (As.rowptr, As.colind, As.values, Bs, xs))
        sC = torch.empty((4096, 64), device=dev)
        sy = torch.empty(4096, device=dev)
        call_a = lambda: kernels.spmm(srp, sci, sva, sB, out=sC)  # noqa: E731
        call_b = lambda: kernels.spmv(srp, sci, sva, sx, out=sy)  # noqa: E731
        el_a, _ = timed(call_a, 200, 10, [])
        el_b, _ = timed(call_b, 200, 10, [])
        extras["cfg0_spmm_us"] = round(el_a / 200 * 1e6, 2)  # per Python call (host-bound at this size)
        extras["cfg0_spmv_us"] = round(el_b / 200 * 1e6, 2)
        all_s = np.arange(4096)
        errs["cfg0_spmm_n64"] = spmm_rows_err(As.rowptr, As.colind, As.values, Bs, sC.cpu().numpy(), all_s)
        errs["cfg0_spmv"] = spmv_rows_err(As.rowptr, As.colind, As.values, xs, sy.cpu().numpy(), all_s)
        # the same call captured once in a HIP graph and replayed: the device-side cost
        kernels.workspace_reserve(0)
        for nm, fn in (("spmm", call_a), ("spmv", call_b)):
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            g = torch.cuda.CUDAGraph()
            with torch.cuda.stream(side):
                with torch.cuda.graph(g, stream=side):
                    fn()
            torch.cuda.current_stream().wait_stream(side)
            el_g, _ = timed(g.replay, 200, 10, [])
            extras[f"cfg0_{nm}_graph_us"] = round(el_g / 200 * 1e6, 2)
        # the literal `./SpMM 4096 4096 4096 32` product: dense C (+)= A·B, A and B
        # 4096² at 1% in CSR (sparse_matrix_mul3)
        Bs2 = data.uniform_csr(4096, 4096, 167_772, data.CFG_SEED + 5)
        ga = tuple(torch.from_numpy(t).to(dev) for t in (As.rowptr, As.colind, As.values))
        gb = tuple(torch.from_numpy(t).to(dev) for t in (Bs2.rowptr, Bs2.colind, Bs2.values))
        gc = torch.zeros((4096, 4096), device=dev)
        el_g2, fam_g2 = timed(lambda: kernels.spgemm_dense(ga, gb, gc), 50, 5, ["spgemm"])
        kg = fam_g2["spgemm"][0] / max(fam_g2["spgemm"][1], 1) * 1e-3
        gbytes = 2 * (4 * 4097 + 8 * 167_772) + 8 * 4096 * 4096
        extras["spgemm_4096_kernel_us"] = round(kg * 1e6, 1)
        extras["spgemm_4096_GBps"] = round(gbytes / kg / 1e9, 1)
        import scipy.sparse as sps
        gcalls = 5 + 50  # warm-up and timed calls all accumulate into gc: C = 55·A·B
        a64 = sps.csr_matrix((As.values.astype(np.float64), As.colind, As.rowptr), shape=(4096, 4096))
        b64 = sps.csr_matrix((Bs2.values.astype(np.float64), Bs2.colind, Bs2.rowptr), shape=(4096, 4096))
        ref_g = (a64 @ b64).toarray() * gcalls
        got_g = gc.cpu().numpy().astype(np.float64)
        den = np.linalg.norm(ref_g, axis=1)
        errs["spgemm_4096"] = float(np.max(np.linalg.norm(got_g - ref_g, axis=1) / np.where(den > 0, den, 1)))
        del gc
        A6 = data.uniform_65k()
        r6, c6, v6 = (torch.from_numpy(t).to(dev) for t in (A6.rowptr, A6.colind, A6.values))
        x6 = torch.empty(65536, device=dev)
        kernels.fill_uniform_(x6, data.CFG_SEED + 1, data.STREAM_DENSE + 1)
        y6 = torch.empty(65536, device=dev)
        el_c, _ = timed(lambda: kernels.spmv(r6, c6, v6, x6, out=y6), 100, 5, [])
        extras["spmv_65k_us"] = round(el_c / 100 * 1e6, 2)
        rows6 = sample_rows(A6.rowptr, 4096, 16)
        errs["spmv_65k"] = spmv_rows_err(A6.rowptr, A6.colind, A6.values, x6.cpu().numpy(),
                                         y6.cpu().numpy()[rows6], rows6)
        for kd in (32, 128):
            U6 = torch.empty((65536, kd), device=dev)
            kernels.fill_normal_(U6, data.CFG_SEED + 1, data.STREAM_DENSE + 2)
            el_d, _ = timed(lambda: kernels.sddmm(r6, c6, U6, U6, scale=v6), 50, 5, [])
            extras[f"sddmm_65k_kd{kd}_us"] = round(el_d / 50 * 1e6, 2)
            o6 = kernels.sddmm(r6, c6, U6, U6, scale=v6).cpu().numpy()
            U6h = U6.cpu().numpy()
            errs[f"sddmm_65k_kd{kd}"] = sddmm_rows_err(A6.rowptr, A6.colind, A6.values, U6h, U6h, o6, rows6)
    if "spmv_variants" in only:  # not in the default set: SpMV schedules side by side
        x = torch.empty(K, device=dev)
        kernels.fill_uniform_(x, data.CFG_SEED + 1, data.STREAM_DENSE + 1)
        y = torch.empty(M, device=dev)
        for name, vcfg in (("row", "row"), ("merge_ipt4", dict(schedule="merge", chunk=256)),
                           ("merge_ipt8", dict(schedule="merge", chunk=512)),
                           ("merge_ipt16", dict(schedule="merge", chunk=1024))):
            n_v = 10 if name == "row" else 40
            el_v, fam_v = timed(lambda: kernels.spmv(rp, ci, va, x, out=y, config=vcfg), n_v, 5,
                                ["spmv", "merge_partition", "spmv_fixup"])
            per = {f: round(ms / max(c, 1) * 1e3, 1) for f, (ms, c) in fam_v.items()}
            extras[f"spmv_{name}_us"] = round(el_v / n_v * 1e6, 1)
            log(f"[bench] spmv {name}: {el_v / n_v * 1e6:.1f} us/call, kernels {per}")
        del x, y

    if "spmv" in only:
        x = torch.empty(K, device=dev)
        kernels.fill_uniform_(x, data.CFG_SEED + 1, data.STREAM_DENSE + 1)
        y = torch.empty(M, device=dev)
        # wall time per call with the library's timing events off, then the kernels
        # on their own (events on) in a second pass
        el_v, _ = timed(lambda: kernels.spmv(rp, ci, va, x, out=y), steps_x * 4, 5, [])
        _, fam_v = timed(lambda: kernels.spmv(rp, ci, va, x, out=y), steps_x * 4, 2,
                         ["spmv", "merge_partition", "spmv_fixup"])
        vb = 4 * (M + 1) + 8 * nnz + 4 * K + 4 * M
        extras["spmv_GBps"] = round(vb * steps_x * 4 / el_v / 1e9, 1)
        extras["spmv_us"] = round(el_v / (steps_x * 4) * 1e6, 1)
        xh = x.cpu().numpy()
        errs["spmv"] = spmv_rows_err(A.rowptr, A.colind, A.values, xh, y.cpu().numpy()[rows], rows)
        extras["spmv_kernel_us"] = round(fam_v["spmv"][0] / max(fam_v["spmv"][1], 1) * 1e3, 1)
        extras["spmv_partition_fixup_us"] = round(sum(fam_v[f][0] / max(fam_v[f][1], 1) * 1e3
                                                      for f in ("merge_partition", "spmv_fixup")), 1)
        # the reference's vector-aware (zero-padded 1×4 blocks) CSR SpMV, y += A·x; the
        # blocked arrays built on the host and on the device (same arrays: tests)
        t0 = time.perf_counter()
        Av = data.csr_to_vcsr(A)
        extras["vcsr_convert_host_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        conv = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            dv = kernels.csr_to_vcsr(rp, ci, va, K)
            torch.cuda.synchronize()
            conv.append(time.perf_counter() - t0)
            del dv
        extras["vcsr_convert_device_ms"] = round(min(conv) * 1e3, 2)
        rv, cv, vv = (torch.from_numpy(t).to(dev) for t in (Av.rowptr, Av.colind, Av.values))
        xv = torch.zeros(Av.shape[1], device=dev)
        xv[:K] = x
        y.zero_()
        kernels.vcsr_spmv(rv, cv, vv, xv, y)
        errs["vcsr_spmv"] = spmv_rows_err(Av.rowptr, Av.colind, Av.values, xv.cpu().numpy(), y.cpu().numpy()[rows],
                                          rows)
        # bit-exact: the whole call (listing the medium and long rows, then the walkers)
        # with timing events off, then the walkers' kernel on its own
        el_w, _ = timed(lambda: kernels.vcsr_spmv(rv, cv, vv, xv, y), steps_x * 2, 5, [])
        _, fam_w = timed(lambda: kernels.vcsr_spmv(rv, cv, vv, xv, y), steps_x * 2, 2, ["spmv_vcsr"])
        pv = int(cv.numel())
        wb = 4 * (M + 1) + 8 * pv + 4 * K + 8 * M
        kw = fam_w["spmv_vcsr"][0] / max(fam_w["spmv_vcsr"][1], 1) * 1e-3
        extras["vcsr_spmv_padded_nnz"] = pv
        extras["vcsr_spmv_us"] = round(el_w / (steps_x * 2) * 1e6, 1)  # wall per call
        extras["vcsr_spmv_kernel_us"] = round(kw * 1e6, 1)
        extras["vcsr_spmv_GBps"] = round(wb * steps_x * 2 / el_w / 1e9, 1)
        el_wb, _ = timed(lambda: kernels.vcsr_spmv(rv, cv, vv, xv, y, balanced=True), steps_x * 2, 5, [])
        extras["vcsr_spmv_balanced_us"] = round(el_wb / (steps_x * 2) * 1e6, 1)  # merge path, wall per call
        extras["vcsr_spmv_balanced_GBps"] = round(wb * steps_x * 2 / el_wb / 1e9, 1)
        del rv, cv, vv, xv, Av
    if "sddmm" in only:
        for kd in (32, 128):
            U = torch.empty((M, kd), device=dev)
            V = torch.empty((K, kd), device=dev)
            kernels.fill_normal_(U, data.CFG_SEED + 1, data.STREAM_DENSE + 2)
            kernels.fill_normal_(V, data.CFG_SEED + 1, data.STREAM_DENSE + 3)
            el_s, fam_s = timed(lambda: kernels.sddmm(rp, ci, U, V, scale=va), steps_x, 3, ["sddmm"])
            sb = 4 * (M + 1) + 12 * nnz + 2 * 4 * M * kd
            Uh, Vh = U.cpu().numpy(), V.cpu().numpy()
            errs[f"sddmm_kd{kd}"] = sddmm_rows_err(A.rowptr, A.colind, A.values, Uh, Vh,
                                                   kernels.sddmm(rp, ci, U, V, scale=va).cpu().numpy(), rows)
            extras[f"sddmm_kd{kd}_GBps"] = round(sb * steps_x / el_s / 1e9, 1)
            extras[f"sddmm_kd{kd}_us"] = round(el_s / steps_x * 1e6, 1)
            # bf16 operands (the GCN's dA: features in bf16), fp32 accumulation
            Ub, Vb = U.to(torch.bfloat16), V.to(torch.bfloat16)
            el_b, _ = timed(lambda: kernels.sddmm(rp, ci, Ub, Vb, scale=va), steps_x, 3, [])
            sbb = 4 * (M + 1) + 12 * nnz + 2 * 2 * M * kd
            extras[f"sddmm_kd{kd}_bf16_GBps"] = round(sbb * steps_x / el_b / 1e9, 1)
            extras[f"sddmm_kd{kd}_bf16_us"] = round(el_b / steps_x * 1e6, 1)
            errs[f"sddmm_kd{kd}_bf16"] = sddmm_rows_err(A.rowptr, A.colind, A.values, Ub.float().cpu().numpy(),
                                                        Vb.float().cpu().numpy(),
                                                        kernels.sddmm(rp, ci, Ub, Vb, scale=va).cpu().numpy(), rows)
            del Ub, Vb
            if kd == 128:  # dot products split into column halves over the XCDs (a selector candidate)
                split = _lib.Config.make("merge", unroll=2)
                el_s, _ = timed(lambda: kernels.sddmm(rp, ci, U, V, scale=va, config=split), steps_x, 3, [])
                extras["sddmm_kd128_split_GBps"] = round(sb * steps_x / el_s / 1e9, 1)
                extras["sddmm_kd128_split_us"] = round(el_s / steps_x * 1e6, 1)
                errs["sddmm_kd128_split"] = sddmm_rows_err(
                    A.rowptr, A.colind, A.values, Uh, Vh, kernels.sddmm(rp, ci, U, V, scale=va, config=split).cpu().numpy(),
                    rows)
                # U and V rows at the gather pitch (kernels.row_pitch), the same values
                pk = kernels.row_pitch(kd, 4)
                Up = torch.empty((M, pk), device=dev)[:, :kd]
                Vp = torch.empty((K, pk), device=dev)[:, :kd]
                Up.copy_(U)
                Vp.copy_(V)
                for tag, cf in (("", None), ("_split", split)):
                    el_p, _ = timed(lambda: kernels.sddmm(rp, ci, Up, Vp, scale=va, config=cf), steps_x, 3, [])
                    extras[f"sddmm_kd128{tag}_pitched_us"] = round(el_p / steps_x * 1e6, 1)
                    errs[f"sddmm_kd128{tag}_pitched"] = sddmm_rows_err(
                        A.rowptr, A.colind, A.values, Uh, Vh,
                        kernels.sddmm(rp, ci, Up, Vp, scale=va, config=cf).cpu().numpy(), rows)
                log(f"[bench] sddmm Kd=128 with U, V rows at a {pk * 4}-byte pitch: "
                    f"{extras['sddmm_kd128_pitched_us']} us, split {extras['sddmm_kd128_split_pitched_us']} us")
                del Up, Vp
            del U, V
    if "gcn" in only:
        from armsparse_amd import gcn

        for prec in ("fp32", "bf16"):
            torch.manual_seed(0)
            task = gcn.make_task(precision=prec)
            el_g, _ = timed(lambda: gcn.train_step(task, sync=False), 30, 5, [])
            extras[f"gcn_{prec}_ms_per_epoch"] = round(el_g / 30 * 1e3, 3)
            del task
        extras["gcn_bf16_speedup"] = round(extras["gcn_fp32_ms_per_epoch"] / extras["gcn_bf16_ms_per_epoch"], 3)
        # each precision's step captured once in a HIP graph and replayed (one launch per
        # epoch): the device-side cost of an epoch, like for like
        for prec in ("fp32", "bf16"):
            torch.manual_seed(0)
            task = gcn.make_task(precision=prec, capturable=True)
            step = gcn.GraphedStep(task)
            el_g, _ = timed(step.replay, 30, 5, [])
            extras[f"gcn_{prec}_graph_ms_per_epoch"] = round(el_g / 30 * 1e3, 3)
            del step, task
        extras["gcn_bf16_graph_speedup"] = round(
            extras["gcn_fp32_graph_ms_per_epoch"] / extras["gcn_bf16_graph_ms_per_epoch"], 3)
    if "gcn_scaled" in only:
        from armsparse_amd import gcn

        t0 = time.time()
        base = gcn.make_task(precision="fp32", **gcn.SCALED)
        log(f"[bench] scaled GCN graph ({base.nnz} nnz) built in {time.time() - t0:.1f}s")
        # fp32 and bf16 with the reference's layer order (x·W, then aggregate), then bf16
        # with the first layer aggregating its 100 input features before x·W (order "auto")
        for prec, order, tag in (("fp32", "transform_first", "fp32"), ("bf16", "transform_first", "bf16"),
                                 ("bf16", "auto", "bf16_auto_order")):
            torch.manual_seed(0)
            task = gcn.retarget(base, prec, order=order)
            el_g, fam_g = timed(lambda: gcn.train_step(task, sync=False), 10, 3, ["spmm", "gemm"])
            extras[f"gcn_scaled_{tag}_ms_per_epoch"] = round(el_g / 10 * 1e3, 2)
            extras[f"gcn_scaled_{tag}_spmm_ms"] = round(fam_g["spmm"][0] / 10, 2)
            log(f"[bench] scaled GCN {tag}: {el_g / 10 * 1e3:.1f} ms/epoch, SpMM {fam_g['spmm'][0] / 10:.1f} ms")
            del task
        del base
        torch.cuda.empty_cache()
        extras["gcn_scaled_bf16_speedup"] = round(
            extras["gcn_scaled_fp32_ms_per_epoch"] / extras["gcn_scaled_bf16_ms_per_epoch"], 3)

    cpu = None
    if "cpu" in only:
        cpu = cpu_baselines(args, A, B, spmm_bytes(4, accumulate=True))
        log(f"[bench] cpu baseline {cpu}")

    mf = stamped_profile("gemm_mfma.json", GEMM_SOURCES)
    if mf is not None:  # MFMA utilisation of the GCN's x·W GEMMs (counter summary of these sources)
        extras["mfma"] = mf.get("kernels")
    extras["err_vs_fp64"] = {k: float(f"{v:.3g}") for k, v in errs.items()}

    out = {
        "metric": "FP32 CSR SpMM GB/s (algorithmic bytes) on 1M x 1M R-MAT, N=128",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded R-MAT, U(-1,1) values and B)",
        "config": {"workload": "FP32 CSR SpMM C=A*B, A = 1048576^2 R-MAT (a=.57,b=c=.19,d=.05) with 2^24 nnz, "
                               "B 1048576x128 fp32", "M": M, "K": K, "N": N, "nnz": nnz,
                   "schedule": (cfg.as_dict() if cfg else "auto"), "algorithmic_bytes": nbytes,
                   "hbm_pct_of_peak": round(100 * value / HBM_PEAK_GBS, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "SpMM main kernel of the AUTO schedule (k_spmm_slice: merge path, half a wave per "
                               "nonzero, 256-byte column slices)" if cfg is None else "SpMM main kernel",
                     "avg_launch_us": round(k_avg * 1e6, 2)},
        "cpu_baseline": cpu,
        "extras": extras,
    }
    del stream
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
