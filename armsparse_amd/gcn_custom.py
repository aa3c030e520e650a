"""Training script of gcn/pytorch-custom/gcn_custom.py on the MI355X:

    python -m armsparse_amd.gcn_custom --n-hidden=64 [--precision bf16] [--graphed]

Same network (GCNConv → ReLU → dropout → GCNConv → log_softmax, gcn_custom.py:72-90),
optimizer (Adam, lr 0.01, weight decay 5e-4 on the first layer only, :116-119),
200-epoch loop of one training step and one evaluation per epoch with the
reference's per-epoch line (:133-152).  Planetoid cannot be downloaded here, so the
graph is the seeded PubMed-shaped power-law graph (19,717 nodes, 88,648 directed
edges + self loops, 500 TF-IDF-like features, 3 classes) or, with --mtx, a Matrix
Market graph with synthetic features; the split is Planetoid's PubMed split sizes
(20 training nodes per class, 500 validation, 1,000 test), seeded.  Instead of the
reference's profiler table the script ends with the train step's time per epoch
(median and mean after 3 warm-up epochs) and the step's summed kernel time.
"""
from __future__ import annotations

import argparse
import ctypes
import statistics
import time

import numpy as np
import torch
import torch.nn.functional as F

from . import _lib, data, gcn


def parse(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--use_gdc", action="store_true", help="accepted for compatibility; GDC is not applied")
    p.add_argument("--n-hidden", type=int, default=64, help="number of hidden features")
    p.add_argument("--epochs", type=int, default=200)
    p.add_argument("--precision", choices=["fp32", "bf16"], default="fp32")
    p.add_argument("--order", choices=["transform_first", "aggregate_first", "auto"], default="transform_first")
    p.add_argument("--graphed", action="store_true", help="replay the training step from a HIP graph")
    p.add_argument("--mtx", default="", help="Matrix Market graph instead of the synthetic PubMed-shaped one")
    p.add_argument("--features", type=int, default=500)
    p.add_argument("--classes", type=int, default=3)
    p.add_argument("--seed", type=int, default=data.CFG_SEED + 3)
    p.add_argument("--optimizer", choices=("native", "torch"), default="native",
                   help="Adam on the library's kernels (native) or torch.optim.Adam (torch)")
    return p.parse_args(argv)


def _mtx_task(path, precision, hidden, features, classes, seed, capturable, order, optimizer):
    """GCNTask on a Matrix Market graph: symmetric adjacency (mirrored, stored entries
    only, weights dropped as gcn_custom.py's all-ones COO does) plus self loops."""
    a = data.read_matrix_market(path)
    if a.shape[0] != a.shape[1]:
        raise SystemExit(f"{path}: a graph needs a square adjacency matrix")
    n = a.shape[0]
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(a.rowptr))
    cols = a.colind.astype(np.int64)
    keep = rows != cols
    rows = np.concatenate([rows[keep], np.arange(n)])
    cols = np.concatenate([cols[keep], np.arange(n)])
    order_ = np.lexsort((cols, rows))
    rows, cols = rows[order_], cols[order_]
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, rows + 1, 1)
    g = data.GCNGraph(n, np.cumsum(rp).astype(np.int32), cols.astype(np.int32), np.ones(cols.size, np.float32),
                      None, None, None)
    g.colptr, g.rowind, g.value_csc = g.rowptr.copy(), g.colind.copy(), g.value_csr.copy()
    return gcn.make_task(precision, seed=seed, n=n, features=features, hidden=hidden, classes=classes,
                         capturable=capturable, order=order, graph=g, optimizer=optimizer)


def splits(n, classes, seed):
    """Planetoid-PubMed-sized split: 20·classes training, 500 validation, 1000 test nodes."""
    r = np.random.default_rng(seed + 1)
    perm = r.permutation(n)
    ntr = min(20 * classes, n // 3)
    nva = min(500, (n - ntr) // 2)
    nte = min(1000, n - ntr - nva)
    return perm[:ntr], perm[ntr:ntr + nva], perm[ntr + nva:ntr + nva + nte]


def main(argv=None) -> int:
    args = parse(argv)
    if not torch.cuda.is_available():
        raise SystemExit("gcn_custom: needs a ROCm device")
    capturable = args.graphed
    if args.mtx:
        task = _mtx_task(args.mtx, args.precision, args.n_hidden, args.features, args.classes, args.seed, capturable,
                         args.order, args.optimizer)
    else:
        task = gcn.make_task(args.precision, seed=args.seed, hidden=args.n_hidden, features=args.features,
                             classes=args.classes, capturable=capturable, order=args.order, optimizer=args.optimizer)
    n = task.x.shape[0]
    tr, va, te = splits(n, args.classes, args.seed)
    dev = task.x.device
    idx = {k: torch.from_numpy(v).to(dev) for k, v in (("train", tr), ("val", va), ("test", te))}
    task.train_idx = idx["train"]
    task.y_train = task.y.index_select(0, task.train_idx)
    step = gcn.GraphedStep(task, warmup=3) if args.graphed else None

    @torch.no_grad()
    def test():
        task.model.eval()
        logits = task.model(task.x, task.graph)
        accs = []
        for k in ("train", "val", "test"):
            pred = logits.index_select(0, idx[k]).max(1)[1]
            accs.append(pred.eq(task.y.index_select(0, idx[k])).sum().item() / idx[k].numel())
        return accs

    times = []
    _lib.lib.armsp_timing_enable(0)
    best_val = test_acc = 0.0
    for epoch in range(1, args.epochs + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if step is not None:
            step.replay()
        else:
            gcn.train_step(task, sync=False)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        train_acc, val_acc, tmp_test = test()
        if val_acc > best_val:
            best_val, test_acc = val_acc, tmp_test
        print("Epoch: {:03d}, Train: {:.4f}, Val: {:.4f}, Test: {:.4f}".format(epoch, train_acc, best_val, test_acc),
              flush=True)
    steady = times[3:] if len(times) > 3 else times
    kern = _kernel_ms_per_step(task, step)
    print(f"train step: median {statistics.median(steady) * 1e3:.3f} ms/epoch, mean "
          f"{statistics.fmean(steady) * 1e3:.3f} ms/epoch over {len(steady)} epochs after 3 warm-up "
          f"({args.precision}, {'HIP graph' if step else 'eager'}, {n} nodes, {task.nnz} edges)")
    if kern is not None:
        print(f"train step time in the library's SpMM, SDDMM, GEMM and transpose kernels: {kern:.3f} ms/epoch")
    return 0


def _kernel_ms_per_step(task, step, reps: int = 5):
    """Summed event time of the library's kernels in one training step (eager steps
    only: timing events are not recorded inside a captured graph)."""
    if step is not None:
        return None
    _lib.lib.armsp_timing_enable(1)
    _lib.lib.armsp_timing_reset()
    for _ in range(reps):
        gcn.train_step(task, sync=False)
    torch.cuda.synchronize()
    total = 0.0
    for fam in ("spmm", "spmm_fixup", "merge_partition", "sddmm", "gemm", "transpose"):
        ms, cnt = ctypes.c_double(0), ctypes.c_int64(0)
        _lib.lib.armsp_timing_read(fam.encode(), ctypes.byref(ms), ctypes.byref(cnt))
        total += ms.value
    _lib.lib.armsp_timing_enable(0)
    return total / reps


if __name__ == "__main__":
    raise SystemExit(main())
