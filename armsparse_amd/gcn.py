"""The 2-layer GCN of gcn/pytorch-custom/gcn_custom.py:72-126 on synthetic
PubMed-shaped data, for the training-step benchmark and tests.

Planetoid data cannot be downloaded here, so the graph is a seeded power-law
graph of the same size (19,717 nodes, 44,324 undirected pairs + self loops)
and the features are TF-IDF-like: 500 columns, ~1% nonzero U(0,1) entries,
row-normalised as T.NormalizeFeatures does (gcn_custom.py:21).  Labels are
random over 3 classes; 60 training nodes as in Planetoid's PubMed split.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch
import torch.nn.functional as F

from . import data, kernels, optim
from .op import GCNConv


class Net(torch.nn.Module):
    """conv1 → ReLU → dropout → conv2 → log_softmax (gcn_custom.py:72-90)."""

    def __init__(self, in_feats: int, n_hidden: int, n_classes: int, precision: str = "fp32", dropout: float = 0.5,
                 order: str = "transform_first"):
        super().__init__()
        self.conv1 = GCNConv(in_feats, n_hidden, cached=True, normalize=True, precision=precision, order=order)
        self.conv2 = GCNConv(n_hidden, n_classes, cached=True, normalize=True, precision=precision, order=order)
        self.reg_params = self.conv1.parameters()
        self.non_reg_params = self.conv2.parameters()
        self.p = dropout

    def forward(self, x, g):
        if self.conv1.fuses_activation():  # bias, ReLU, dropout and the bf16 cast in one pass
            x = self.conv1.forward_act(x, *g, p=self.p if self.training else 0.0)
        else:
            x = F.relu(self.conv1(x, *g))
            x = F.dropout(x, p=self.p, training=self.training)
        x = self.conv2(x, *g)
        return F.log_softmax(x, dim=1)


@dataclass
class GCNTask:
    model: Net
    optimizer: torch.optim.Optimizer
    x: torch.Tensor
    graph: tuple
    y: torch.Tensor
    train_mask: torch.Tensor
    precision: str
    nnz: int
    train_idx: torch.Tensor = None  # positions of train_mask (index_select needs no host sync)
    y_train: torch.Tensor = None

    def __post_init__(self):
        if self.train_idx is None:
            self.train_idx = self.train_mask.nonzero().squeeze(1)
            self.y_train = self.y.index_select(0, self.train_idx)


def _adam(model: Net, capturable: bool, optimizer: str = "native"):
    """Adam over the two parameter groups of gcn_custom.py:116-119.  "native": the
    library's kernels (armsparse_amd.optim.Adam, capturable, one call for all four
    tensors); "torch": torch.optim.Adam, fused on the GPU."""
    groups = [dict(params=model.reg_params, weight_decay=5e-4), dict(params=model.non_reg_params, weight_decay=0)]
    if optimizer == "native":
        return optim.Adam(groups, lr=0.01)
    if optimizer == "torch":
        fused = next(model.parameters()).is_cuda
        return torch.optim.Adam(groups, lr=0.01, capturable=capturable, fused=fused)
    raise ValueError(f"optimizer must be 'native' or 'torch', got {optimizer!r}")


def line_pitched(x: torch.Tensor) -> torch.Tensor:
    """x [n×f] as a view with rows at kernels.row_pitch (the padding zero): rows of a
    line or more start on 128-byte lines, so the aggregation's gathers of a row touch
    no more lines than the row needs, and span an odd number of lines, which the
    MI355X's memory system serves faster than a pitch that is a multiple of 256 bytes."""
    n, f = x.shape
    fp = kernels.row_pitch(f, x.element_size())
    if fp == f and x.is_contiguous():
        return x
    buf = torch.zeros((n, fp), dtype=x.dtype, device=x.device)
    buf[:, :f] = x
    return buf[:, :f]


def make_task(precision: str = "fp32", seed: int = data.CFG_SEED + 3, n: int = 19_717, pairs: int = 44_324,
              features: int = 500, hidden: int = 64, classes: int = 3, feat_density: float = 0.01,
              device: str = "cuda", dropout: float = 0.5, capturable: bool = False,
              order: str = "transform_first", graph: "data.GCNGraph" = None, optimizer: str = "native") -> GCNTask:
    """Synthetic features, labels and training nodes on the seeded power-law graph of
    n nodes and `pairs` undirected edges (or on `graph`, e.g. a Matrix Market file's).
    `optimizer`: "native" (the library's Adam, always capturable) or "torch"
    (torch.optim.Adam, capturable when `capturable`)."""
    g = graph if graph is not None else data.gcn_graph(n, pairs, seed)
    n = g.n
    r = np.random.default_rng(seed)
    x = np.zeros((n, features), np.float32)
    for lo in range(0, n, 1 << 18):  # in row blocks to bound host temporaries
        hi = min(n, lo + (1 << 18))
        blk = r.random((hi - lo, features), dtype=np.float32)
        x[lo:hi] = np.where(r.random((hi - lo, features), dtype=np.float32) < feat_density, blk, 0.0)
    rs = x.sum(1, keepdims=True)
    np.divide(x, rs, out=x, where=rs > 0)  # NormalizeFeatures
    y = r.integers(0, classes, n)
    train = np.zeros(n, bool)
    train[r.choice(n, 20 * classes, replace=False)] = True
    dev = torch.device(device)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    xt = t(x)
    if precision == "bf16":
        xt = line_pitched(xt.to(torch.bfloat16))
    graph = (t(g.rowptr), t(g.colind), t(g.colptr), t(g.rowind), t(g.value_csr), t(g.value_csc))
    model = Net(features, hidden, classes, precision, dropout, order).to(dev)
    return GCNTask(model, _adam(model, capturable, optimizer), xt, graph, t(y), t(train), precision,
                   int(g.colind.size))


SCALED = dict(n=2_400_000, pairs=62_000_000, features=100, hidden=256, feat_density=0.1)


def retarget(task: GCNTask, precision: str, capturable: bool = False, order: str = "transform_first",
             optimizer: str = None) -> GCNTask:
    """A fresh model and optimizer of `precision` (and layer `order`, which needs
    bf16's fused aggregation) on the same graph, features and labels; the
    optimizer of the task's kind unless `optimizer` names one."""
    if optimizer is None:
        optimizer = "native" if isinstance(task.optimizer, optim.Adam) else "torch"
    x = task.x.float() if precision == "fp32" else line_pitched(task.x.to(torch.bfloat16))
    dev = x.device
    model = Net(x.shape[1], task.model.conv1.out_channels, task.model.conv2.out_channels, precision,
                task.model.p, order).to(dev)
    return GCNTask(model, _adam(model, capturable, optimizer), x, task.graph, task.y, task.train_mask, precision,
                   task.nnz, task.train_idx, task.y_train)


def _step(task: GCNTask) -> torch.Tensor:
    task.model.train()
    task.optimizer.zero_grad(set_to_none=True)
    out = task.model(task.x, task.graph)
    loss = F.nll_loss(out.index_select(0, task.train_idx), task.y_train)
    loss.backward()
    task.optimizer.step()
    return loss.detach()


def train_step(task: GCNTask, sync: bool = True):
    """forward + nll loss over the training nodes + backward + Adam step
    (gcn_custom.py:122-126)."""
    loss = _step(task)
    return float(loss) if sync else loss


class GraphedStep:
    """The training step captured once in a HIP graph and replayed: one launch
    per epoch instead of one per kernel, for graphs small enough that an epoch
    is launch-bound (PubMed size).  The task's optimizer must be capturable
    (the native Adam is; torch's needs make_task(..., capturable=True)).
    Warm-up steps run eagerly on a side stream (they fill the GCN layers'
    caches and the library's workspace high-water mark), the library's capture
    arena is reserved, then one step is captured; replay() runs it again,
    loss() reads its last loss."""

    def __init__(self, task: GCNTask, warmup: int = 3):
        if not task.optimizer.defaults.get("capturable"):
            raise ValueError("GraphedStep needs a capturable optimizer (the native Adam, or "
                             "make_task(..., capturable=True))")
        self.task = task
        dev = task.x.device
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):
            for _ in range(warmup):
                _step(task)
        torch.cuda.current_stream(dev).wait_stream(side)
        torch.cuda.synchronize(dev)
        kernels.workspace_reserve(0, dev)
        task.optimizer.zero_grad(set_to_none=True)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._loss = _step(task)

    def replay(self) -> None:
        self.graph.replay()

    def loss(self) -> float:
        return float(self._loss)
