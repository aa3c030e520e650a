"""The shipped selector weights against the recorded MI355X autotuner timings
(tests/golden/selector_dataset.npz, written by `python -m armsparse_amd.selector
--matrices 2000`), evaluated on the CPU: the network's forward pass in FP64
(oracle.mlp) on the matrices held out of its training (the same split rule as
selector.train_eval), so the regret claim can be checked without a GPU."""
import json
import os

import numpy as np
import pytest

import oracle.mlp as omlp
from armsparse_amd import selector
from conftest import GOLDEN

DATASET = os.path.join(GOLDEN, "selector_dataset.npz")


def held_out(mids, seed=7):
    u = np.unique(mids)
    r = np.random.default_rng(seed)
    test = set(r.permutation(u)[: max(1, len(u) // 5)].tolist())
    return np.array([m in test for m in mids])


@pytest.mark.parametrize("op", ["spmm", "sddmm"])
def test_shipped_weights_regret_on_held_out_matrices(op):
    z = np.load(DATASET)
    x, t, mid = z[f"{op}_x"], z[f"{op}_t"], z[f"{op}_mid"]
    with open(selector.WEIGHTS_PATH) as f:
        st = json.load(f)[op]
    p = {k: np.array(st[k], np.float64) for k in ("W1", "b1", "W2", "b2", "W3", "b3")}
    te = held_out(mid) & np.isfinite(t).any(axis=1)
    xs = (x[te] - np.array(st["mu"])) / np.array(st["sd"])
    logits, _ = omlp.forward(p, xs)
    cls = logits.argmax(1)
    tt = np.where(np.isfinite(t[te]), t[te], np.inf)
    best = tt.min(1)
    sel = tt[np.arange(len(cls)), cls]
    total = sel.sum() / best.sum() - 1
    heuristic = tt[:, 0].sum() / best.sum() - 1  # class 0: the library's AUTO choice
    print(f"{op}: held-out total regret {total:.4f} (heuristic AUTO {heuristic:.4f}) over {te.sum()} samples")
    assert len(selector.CONFIGS[op]) == t.shape[1]
    assert total <= 0.10
    assert total < heuristic
