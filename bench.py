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


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--only", default="", help="comma list: spmm,sddmm,spmv,mixed,widths,shards,small,gcn,gcn_scaled,cpu (default all); "
                        "regimes (column-placement variants of the headline matrix) runs only when named")
    p.add_argument("--schedule", default="", help="override the SpMM schedule (row|nnz|merge)")
    p.add_argument("--chunk", type=int, default=0)
    p.add_argument("--unroll", type=int, default=0)
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--copy-probe", action="store_true",
                   help="also run a plain 512 MiB device copy (calibrates rocprof FETCH_SIZE/WRITE_SIZE)")
    return p.parse_args()


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1 (no multi-GPU path)"}))
        sys.exit(2)
    import numpy as np
    import torch

    from armsparse_amd import _lib, data, kernels

    only = set(filter(None, args.only.split(","))) or {"spmm", "sddmm", "spmv", "mixed", "widths", "shards", "small",
                                                        "gcn", "gcn_scaled", "cpu"}
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
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        _lib.lib.armsp_timing_reset()
        _lib.lib.armsp_timing_enable(1)
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
    tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "spmm_rmat1m_n128_traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")
    extras = {"spmm_main_kernel_us": round(k_avg * 1e6, 2),
              "spmm_fixup_kernel_us": round(f_ms / max(f_n, 1) * 1e3, 2),
              "gflops": round(2.0 * nnz * N * args.steps / el / 1e9, 1)}
    steps_x = max(10, args.steps // 2)

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
        del Bb, q, sc
    if "regimes" in only:  # not in the default set: what bounds the gathers (see the counter summaries)
        rng = np.random.default_rng(0)
        c64 = A.colind.astype(np.int64)
        variants = {"rmat": A.colind, "rmat_mod64k": (c64 % 65536).astype(np.int32),
                    "uniform": rng.integers(0, K, nnz).astype(np.int32),
                    "rmat_oddmul": ((c64 * 2654435761) & (K - 1)).astype(np.int32)}
        rows = np.repeat(np.arange(M, dtype=np.int64), np.diff(A.rowptr))
        for name, cv in variants.items():
            order = np.argsort(rows * K + cv, kind="stable")  # keep every row's columns sorted
            vci = torch.from_numpy(np.ascontiguousarray(cv[order])).to(dev)
            vva = torch.from_numpy(np.ascontiguousarray(A.values[order])).to(dev)
            el_r, fam_r = timed(lambda: kernels.spmm(rp, vci, vva, B, out=C), max(2, args.steps // 5), 1, ["spmm"])
            extras[f"regime_{name}_kernel_us"] = round(fam_r["spmm"][0] / max(fam_r["spmm"][1], 1) * 1e3, 1)
            log(f"[bench] regime {name}: {extras[f'regime_{name}_kernel_us']} us")
            del vci, vva
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
            del B2, C2
    if "shards" in only:  # one GPU's share of an 8-way row split: heaviest and median 131,072-row blocks
        rows = M // 8
        per = [int(A.rowptr[(i + 1) * rows] - A.rowptr[i * rows]) for i in range(8)]
        order = sorted(range(8), key=lambda i: per[i])
        for tag, i in (("heavy", order[-1]), ("median", order[3])):
            lo, hi = int(A.rowptr[i * rows]), int(A.rowptr[(i + 1) * rows])
            srp = (rp[i * rows:(i + 1) * rows + 1] - lo).contiguous()
            sci, sva = ci[lo:hi], va[lo:hi]
            Cs = C[i * rows:(i + 1) * rows]
            el_s, _ = timed(lambda: kernels.spmm(srp, sci, sva, B, out=Cs), 10, 2, [])
            extras[f"shard_{tag}_nnz"] = hi - lo
            extras[f"shard_{tag}_us"] = round(el_s / 10 * 1e6, 1)
        extras["shard_rows"] = rows
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
        del gc
        A6 = data.uniform_65k()
        r6, c6, v6 = (torch.from_numpy(t).to(dev) for t in (A6.rowptr, A6.colind, A6.values))
        x6 = torch.empty(65536, device=dev)
        kernels.fill_uniform_(x6, data.CFG_SEED + 1, data.STREAM_DENSE + 1)
        y6 = torch.empty(65536, device=dev)
        el_c, _ = timed(lambda: kernels.spmv(r6, c6, v6, x6, out=y6), 100, 5, [])
        extras["spmv_65k_us"] = round(el_c / 100 * 1e6, 2)
        for kd in (32, 128):
            U6 = torch.empty((65536, kd), device=dev)
            kernels.fill_normal_(U6, data.CFG_SEED + 1, data.STREAM_DENSE + 2)
            el_d, _ = timed(lambda: kernels.sddmm(r6, c6, U6, U6, scale=v6), 50, 5, [])
            extras[f"sddmm_65k_kd{kd}_us"] = round(el_d / 50 * 1e6, 2)
    if "spmv" in only:
        x = torch.empty(K, device=dev)
        kernels.fill_uniform_(x, data.CFG_SEED + 1, data.STREAM_DENSE + 1)
        y = torch.empty(M, device=dev)
        el_v, fam_v = timed(lambda: kernels.spmv(rp, ci, va, x, out=y), steps_x * 4, 5, ["spmv"])
        vb = 4 * (M + 1) + 8 * nnz + 4 * K + 4 * M
        extras["spmv_GBps"] = round(vb * steps_x * 4 / el_v / 1e9, 1)
        extras["spmv_us"] = round(el_v / (steps_x * 4) * 1e6, 1)
        extras["spmv_kernel_us"] = round(fam_v["spmv"][0] / max(fam_v["spmv"][1], 1) * 1e3, 1)
        # the reference's vector-aware (zero-padded 1×4 blocks) CSR SpMV, y += A·x
        Av = data.csr_to_vcsr(A)
        rv, cv, vv = (torch.from_numpy(t).to(dev) for t in (Av.rowptr, Av.colind, Av.values))
        xv = torch.zeros(Av.shape[1], device=dev)
        xv[:K] = x
        el_w, fam_w = timed(lambda: kernels.vcsr_spmv(rv, cv, vv, xv, y), steps_x * 2, 5, ["spmv_vcsr"])
        pv = int(cv.numel())
        wb = 4 * (M + 1) + 8 * pv + 4 * K + 8 * M
        kw = fam_w["spmv_vcsr"][0] / max(fam_w["spmv_vcsr"][1], 1) * 1e-3
        extras["vcsr_spmv_padded_nnz"] = pv
        extras["vcsr_spmv_kernel_us"] = round(kw * 1e6, 1)  # bit-exact: one 4-lane group per row
        extras["vcsr_spmv_GBps"] = round(wb / kw / 1e9, 1)
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
            extras[f"sddmm_kd{kd}_GBps"] = round(sb * steps_x / el_s / 1e9, 1)
            extras[f"sddmm_kd{kd}_us"] = round(el_s / steps_x * 1e6, 1)
            # bf16 operands (the GCN's dA: features in bf16), fp32 accumulation
            Ub, Vb = U.to(torch.bfloat16), V.to(torch.bfloat16)
            el_b, _ = timed(lambda: kernels.sddmm(rp, ci, Ub, Vb, scale=va), steps_x, 3, [])
            sbb = 4 * (M + 1) + 12 * nnz + 2 * 2 * M * kd
            extras[f"sddmm_kd{kd}_bf16_GBps"] = round(sbb * steps_x / el_b / 1e9, 1)
            extras[f"sddmm_kd{kd}_bf16_us"] = round(el_b / steps_x * 1e6, 1)
            del Ub, Vb
            if kd == 128:  # dot products split into column halves over the XCDs (a selector candidate)
                split = _lib.Config.make("merge", unroll=2)
                el_s, _ = timed(lambda: kernels.sddmm(rp, ci, U, V, scale=va, config=split), steps_x, 3, [])
                extras["sddmm_kd128_split_GBps"] = round(sb * steps_x / el_s / 1e9, 1)
                extras["sddmm_kd128_split_us"] = round(el_s / steps_x * 1e6, 1)
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
        # the bf16 step captured once in a HIP graph and replayed (one launch per epoch)
        torch.manual_seed(0)
        task = gcn.make_task(precision="bf16", capturable=True)
        step = gcn.GraphedStep(task)
        el_g, _ = timed(step.replay, 30, 5, [])
        extras["gcn_bf16_graph_ms_per_epoch"] = round(el_g / 30 * 1e3, 3)
        del step, task
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
        import oracle

        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        Bh = B.cpu().numpy()
        Ch = np.zeros((M, N), np.float32)
        oracle.rowBaseSpmm3(A.rowptr, A.colind, A.values, Bh, Ch, threads=threads)  # warm-up
        iters, t = 0, time.perf_counter()
        while True:
            oracle.rowBaseSpmm3(A.rowptr, A.colind, A.values, Bh, Ch, threads=threads)
            iters += 1
            if time.perf_counter() - t >= args.cpu_seconds or iters >= 30:
                break
        el_c = time.perf_counter() - t
        cpu = {"value": round(spmm_bytes(4, accumulate=True) * iters / el_c / 1e9, 2), "unit": "GB/s",
               "cores": threads, "kind": "port",
               "sample": f"full 1M R-MAT N=128, rowBaseSpmm3 (C += A·B, accumulate bytes), 1 warm-up + {iters} "
                         f"timed calls, {el_c / iters * 1e3:.0f} ms/call"}
        log(f"[bench] cpu baseline {cpu}")

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
