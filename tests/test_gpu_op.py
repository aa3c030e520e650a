"""GPU tests of the drop-in Python op (spmm.csr_spmm, SPMMFunction, GCNConv) and
the GCN training step: forward against the restated CPU op, backward against
the analytic Aᵀ·G on the restated CPU kernels, finite differences, and the
edge-weight gradient (SDDMM) against its closed form."""
import numpy as np
import pytest
import torch

import oracle
from armsparse_amd import data, gcn, kernels, spmm
from armsparse_amd.op import GCNConv, SPMMFunction

pytestmark = pytest.mark.gpu


def dev(*arrays):
    return [None if a is None else torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrays]


@pytest.fixture(scope="module")
def graph512():
    """A 512-node power-law subgraph with self loops and random positive weights."""
    g = data.gcn_graph(512, 1500, seed=5)
    w = np.random.default_rng(5).uniform(0.1, 1.0, g.colind.size).astype(np.float32)
    cp, ri, wt = oracle.csr_to_csc(g.rowptr, g.colind, w, 512)
    return g, w, cp, ri, wt


def test_csr_spmm_dropin_matches_cpu_op(graph512):
    g, w, *_ = graph512
    x = np.random.default_rng(1).standard_normal((512, 64)).astype(np.float32)
    out = spmm.csr_spmm(*dev(g.rowptr, g.colind, w, x))
    ref = oracle.csr_spmm(g.rowptr, g.colind, w, x)
    assert out.device.type == "cuda" and out.dtype == torch.float32
    assert np.max(oracle.row_rel_err(out.cpu().numpy(), ref)) <= 1e-5
    # CPU tensors in, CPU tensor out (the reference op's contract), computed on the GPU
    out_cpu = spmm.csr_spmm(*(torch.from_numpy(a) for a in (g.rowptr, g.colind, w, x)))
    assert out_cpu.device.type == "cpu" and torch.equal(out_cpu, out.cpu())


def test_spmm_function_backward(graph512):
    g, w, cp, ri, wt = graph512
    rp, ci, d_w, d_cp, d_ri, d_wt = dev(g.rowptr, g.colind, w, cp, ri, wt)
    x = torch.randn(512, 16, device="cuda", requires_grad=True)
    G = torch.randn(512, 16, device="cuda")
    out = SPMMFunction.apply(rp, ci, d_cp, d_ri, x, d_w, d_wt)
    out.backward(G)
    # analytic: Aᵀ·G with the restated CPU kernel on the transposed CSR
    ref = oracle.csr_spmm(cp, ri, wt, G.cpu().numpy())
    assert np.max(oracle.row_rel_err(x.grad.cpu().numpy(), ref)) <= 1e-5
    # missing CSC weights raises like the reference
    x2 = torch.randn(512, 16, device="cuda", requires_grad=True)
    with pytest.raises(RuntimeError, match="edge values in both"):
        SPMMFunction.apply(rp, ci, d_cp, d_ri, x2, d_w, None).sum().backward()


def test_spmm_function_finite_differences(graph512):
    """Central differences of L = <SPMM(x), R> w.r.t. x and the edge weights (linear: exact up to rounding)."""
    g, w, cp, ri, wt = graph512
    rp, ci, d_cp, d_ri, d_wt = dev(g.rowptr, g.colind, cp, ri, wt)
    R = torch.randn(512, 8, device="cuda", dtype=torch.float32)
    x = torch.randn(512, 8, device="cuda", requires_grad=True)
    ew = torch.from_numpy(w).cuda().requires_grad_(True)
    loss = (SPMMFunction.apply(rp, ci, d_cp, d_ri, x, ew, d_wt) * R).sum()
    loss.backward()
    r = np.random.default_rng(0)
    eps = 0.25
    with torch.no_grad():
        for _ in range(8):
            i, j = int(r.integers(512)), int(r.integers(8))
            xp, xm = x.detach().clone(), x.detach().clone()
            xp[i, j] += eps
            xm[i, j] -= eps
            fd = ((SPMMFunction.apply(rp, ci, d_cp, d_ri, xp, ew, d_wt) * R).sum()
                  - (SPMMFunction.apply(rp, ci, d_cp, d_ri, xm, ew, d_wt) * R).sum()) / (2 * eps)
            assert abs(float(fd) - float(x.grad[i, j])) <= 1e-3 * (1 + abs(float(fd)))
            e = int(r.integers(w.size))
            wp, wm = ew.detach().clone(), ew.detach().clone()
            wp[e] += eps
            wm[e] -= eps
            fd = ((SPMMFunction.apply(rp, ci, d_cp, d_ri, x.detach(), wp, d_wt) * R).sum()
                  - (SPMMFunction.apply(rp, ci, d_cp, d_ri, x.detach(), wm, d_wt) * R).sum()) / (2 * eps)
            assert abs(float(fd) - float(ew.grad[e])) <= 1e-3 * (1 + abs(float(fd)))
    # closed form dA_e = <R[row e], x[col e]>
    ref = oracle.sddmm_f64(g.rowptr, g.colind, R.cpu().numpy(), x.detach().cpu().numpy())
    assert np.allclose(ew.grad.cpu().numpy(), ref, rtol=1e-5, atol=1e-5)


def _gcnconv_f64(x, W, b, rowptr, colind, colptr, val):
    """The reference GCNConv maths (op.py:112-149) in float64 on the CPU."""
    indeg = np.diff(rowptr).astype(np.float64)
    outdeg = np.diff(colptr).astype(np.float64)
    h = (x.astype(np.float64) @ W.astype(np.float64)) / np.sqrt(outdeg)[:, None]
    agg = oracle.densify(rowptr, colind, val, len(rowptr) - 1).astype(np.float64) @ h
    return agg / np.sqrt(indeg)[:, None] + b.astype(np.float64)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_gcnconv_forward(graph512, precision):
    g, *_ = graph512
    rp, ci, v, cp, ri, vc = dev(g.rowptr, g.colind, g.value_csr, g.colptr, g.rowind, g.value_csc)
    torch.manual_seed(0)
    conv = GCNConv(40, 16, precision=precision).cuda()
    with torch.no_grad():
        conv.bias.uniform_(-0.1, 0.1)
    x = torch.rand(512, 40, device="cuda")
    out = conv(x, rp, ci, cp, ri, v, vc).detach().cpu().numpy()
    xin = x.cpu().numpy()
    if precision == "bf16":
        xin = x.to(torch.bfloat16).float().cpu().numpy()
    ref = _gcnconv_f64(xin, conv.weight.detach().cpu().numpy(), conv.bias.detach().cpu().numpy(), g.rowptr,
                       g.colind, g.colptr, g.value_csr)
    err = np.max(oracle.row_rel_err(out, ref))
    print(f"GCNConv {precision}: max row rel err vs FP64 = {err:.2e}")
    assert err <= (1e-5 if precision == "fp32" else 1e-2)


@pytest.mark.parametrize("order", ["transform_first", "aggregate_first"])
@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_fused_gcnconv_edge_weight_gradient(graph512, precision, order):
    """Learnable CSR edge weights get the same gradient through the fused aggregation
    (coefficients folded once per graph) as through the unfused SPMMFunction path, whose
    dA is the SDDMM closed form; bf16 features within 1e-2."""
    g, w, cp, ri, wt = graph512
    rp, ci, d_cp, d_ri, d_wt = dev(g.rowptr, g.colind, cp, ri, wt)
    torch.manual_seed(3)
    x = torch.rand(512, 24, device="cuda")
    R = torch.randn(512, 16, device="cuda")
    grads = {}
    for fused in (False, True):
        torch.manual_seed(4)
        prec = precision if fused else "fp32"
        conv = GCNConv(24, 16, precision=prec, fused=fused, order=order if fused else "transform_first").cuda()
        with torch.no_grad():
            conv.bias.uniform_(-0.1, 0.1)
        ew = torch.from_numpy(w).cuda().requires_grad_(True)
        (conv(x, rp, ci, d_cp, d_ri, ew, d_wt) * R).sum().backward()
        grads[fused] = ew.grad.detach().cpu().numpy().astype(np.float64)
    a, b = grads[True], grads[False]
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    print(f"edge-weight grad fused {precision}/{order} vs unfused fp32: rel err {rel:.2e}")
    assert rel <= (1e-5 if precision == "fp32" else 1e-2)


def test_fused_activation_layer_edge_weight_gradient(graph512):
    """forward_act (fused bf16 hidden layer: aggregation + bias/ReLU/dropout epilogue) also
    returns the edge-weight gradient: equal to the fp32 unfused layer's with dropout off."""
    g, w, cp, ri, wt = graph512
    rp, ci, d_cp, d_ri, d_wt = dev(g.rowptr, g.colind, cp, ri, wt)
    torch.manual_seed(3)
    x = torch.rand(512, 24, device="cuda")
    R = torch.randn(512, 16, device="cuda")
    torch.manual_seed(4)
    ref = GCNConv(24, 16, precision="fp32", fused=False).cuda()
    torch.manual_seed(4)
    conv = GCNConv(24, 16, precision="bf16").cuda()
    for c in (ref, conv):
        with torch.no_grad():
            c.bias.fill_(0.05)
    ew0 = torch.from_numpy(w).cuda().requires_grad_(True)
    (torch.relu(ref(x, rp, ci, d_cp, d_ri, ew0, d_wt)) * R).sum().backward()
    ew1 = torch.from_numpy(w).cuda().requires_grad_(True)
    (conv.forward_act(x, rp, ci, d_cp, d_ri, ew1, d_wt, p=0.0).float() * R).sum().backward()
    a, b = ew1.grad.double().cpu().numpy(), ew0.grad.double().cpu().numpy()
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    print(f"edge-weight grad of the fused activation layer vs fp32: rel err {rel:.2e}")
    assert rel <= 2e-2


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_gcn_training_step(precision):
    """2-layer GCN on the PubMed-shaped synthetic graph: loss decreases, grads match FP32 run."""
    torch.manual_seed(0)
    task = gcn.make_task(precision=precision, seed=data.CFG_SEED + 3)
    losses = [gcn.train_step(task) for _ in range(5)]
    print(f"{precision} losses: {losses}")
    assert all(np.isfinite(losses)) and losses[-1] < losses[0]


def test_fused_gcn_layer_matches_unfused(graph512):
    """The fused aggregation (one SpMM with scales and bias) equals the reference's
    four-op GCNConv in value and in every gradient."""
    g, *_ = graph512
    gt = dev(g.rowptr, g.colind, g.colptr, g.rowind, g.value_csr, g.value_csc)
    torch.manual_seed(1)
    plain = GCNConv(40, 24).cuda()
    fused = GCNConv(40, 24, fused=True).cuda()
    fused.load_state_dict(plain.state_dict())
    with torch.no_grad():
        plain.bias.uniform_(-0.2, 0.2)
        fused.bias.copy_(plain.bias)
    x = torch.rand(512, 40, device="cuda")
    G = torch.randn(512, 24, device="cuda")
    outs, grads = [], []
    for conv in (plain, fused):
        y = conv(x, *gt[:4], *gt[4:])
        y.backward(G)
        outs.append(y.detach().cpu().numpy())
        grads.append((conv.weight.grad.cpu().numpy(), conv.bias.grad.cpu().numpy()))
    assert np.max(oracle.row_rel_err(outs[1], outs[0])) <= 1e-5
    for a, b in zip(grads[0], grads[1]):
        assert np.max(np.abs(a - b)) <= 1e-5 * (np.max(np.abs(a)) + 1e-12)


def test_narrow_layer_over_column_blocks(graph512, monkeypatch):
    """A fused 3-class layer whose aggregations run over column blocks (forced to 4
    here; the library splits only graphs whose gathered rows outgrow the L2) gives
    the one-product layer's output and gradients."""
    g, *_ = graph512
    gt = dev(g.rowptr, g.colind, g.colptr, g.rowind)
    torch.manual_seed(2)
    one = GCNConv(40, 3, cached=True, fused=True).cuda()
    blk = GCNConv(40, 3, cached=True, fused=True).cuda()
    blk.load_state_dict(one.state_dict())
    with torch.no_grad():
        one.bias.uniform_(-0.2, 0.2)
        blk.bias.copy_(one.bias)
    x = torch.rand(512, 40, device="cuda")
    G = torch.randn(512, 3, device="cuda")
    res = []
    for conv, nb in ((one, 1), (blk, 4)):
        monkeypatch.setattr(kernels, "gather_blocks", lambda *a, nb=nb, **k: nb)
        xi = x.clone().requires_grad_(True)
        y = conv(xi, *gt)
        y.backward(G)
        res.append((y.detach().cpu().numpy(), xi.grad.cpu().numpy(), conv.weight.grad.cpu().numpy(),
                    conv.bias.grad.cpu().numpy()))
    assert one.cached_blocks is None and blk.cached_blocks is not None
    assert np.max(oracle.row_rel_err(res[1][0], res[0][0])) <= 1e-5
    for a, b in zip(res[0][1:], res[1][1:]):
        assert np.max(np.abs(a - b)) <= 1e-5 * (np.max(np.abs(a)) + 1e-12)


def _gcn_dense_fp64(g):
    """Â = diag(1/sqrt(in-degree counts))·A·diag(1/sqrt(out-degree counts)), dense fp64."""
    n = g.rowptr.size - 1
    a = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(g.rowptr))
    np.add.at(a, (rows, g.colind), g.value_csr.astype(np.float64))
    in_n = 1 / np.sqrt(np.diff(g.rowptr).astype(np.float64))
    out_n = 1 / np.sqrt(np.diff(g.colptr).astype(np.float64))
    return in_n[:, None] * a * out_n[None, :]


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("bf16", 1e-2)])
def test_aggregate_first_layer_matches_fp64(graph512, precision, tol):
    """order="aggregate_first": y = (Â·x)·W + b, dW = (Â·x)ᵀ·G, db = Σ G, dx = Âᵀ·G·Wᵀ
    against the dense fp64 closed forms; "auto" picks it only when in < out."""
    g, *_ = graph512
    gt = dev(g.rowptr, g.colind, g.colptr, g.rowind, g.value_csr, g.value_csc)
    torch.manual_seed(2)
    conv = GCNConv(40, 96, fused=True, precision=precision, order="aggregate_first").cuda()
    with torch.no_grad():
        conv.bias.uniform_(-0.2, 0.2)
    assert conv.aggregates_first()
    assert GCNConv(40, 96, precision="bf16", order="auto").aggregates_first()
    assert not GCNConv(96, 40, precision="bf16", order="auto").aggregates_first()
    with pytest.raises(ValueError):
        GCNConv(40, 96, order="aggregate_first")  # needs the fused aggregation
    x = torch.rand(512, 40, device="cuda", requires_grad=True)
    G = torch.randn(512, 96, device="cuda")
    y = conv(x, *gt[:4], *gt[4:])
    y.backward(G)
    ah = _gcn_dense_fp64(g)
    xd, gd = x.detach().double().cpu().numpy(), G.double().cpu().numpy()
    wd, bd = conv.weight.detach().double().cpu().numpy(), conv.bias.detach().double().cpu().numpy()
    z = ah @ xd
    ref_y = z @ wd + bd
    assert y.dtype == torch.float32
    assert np.max(oracle.row_rel_err(y.detach().cpu().numpy(), ref_y)) <= tol

    def rel(a, b):
        return np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(conv.weight.grad.cpu().numpy(), z.T @ gd) <= tol
    assert rel(conv.bias.grad.cpu().numpy(), gd.sum(0)) <= tol
    assert rel(x.grad.cpu().numpy(), ah.T @ gd @ wd.T) <= tol


def test_aggregate_first_training_matches_transform_first():
    """A bf16 model whose first layer aggregates first trains like the reference
    order: same initial weights, five steps, losses within bf16 rounding."""
    torch.manual_seed(0)
    a = gcn.make_task("bf16", n=4000, pairs=9000, features=32, hidden=96, dropout=0.0)
    b = gcn.retarget(a, "bf16", order="auto")
    assert b.model.conv1.aggregates_first() and not b.model.conv2.aggregates_first()
    b.model.load_state_dict(a.model.state_dict())
    la = [gcn.train_step(a) for _ in range(5)]
    lb = [gcn.train_step(b) for _ in range(5)]
    assert lb[-1] < lb[0]
    assert max(abs(p - q) for p, q in zip(la, lb)) <= 1e-2 * max(abs(q) for q in la)


def _bits(mask, n):
    """uint8 [M, n/8] mask → bool [M, n] (bit j of byte k is column 8k + j)."""
    m = mask.cpu().numpy()
    return np.unpackbits(m, axis=1, bitorder="little")[:, :n].astype(bool)


@pytest.mark.parametrize("n", [256, 24])
def test_bias_relu_dropout_kernels(n):
    """h = bf16(keep·relu(y + b)/(1-p)) bit for bit against torch on the kernel's
    own mask; the mask is ReLU's mask with about p of it dropped; the seed
    advances; the backward gives bf16(mask·g/(1-p)) bit for bit and the bias
    gradient within 1e-5 of FP64, identically on a second call."""
    print(f"[case] bias_relu_dropout N={n}", flush=True)
    torch.manual_seed(3)
    m = 3001
    y = torch.randn(m, n, device="cuda")
    b = torch.randn(n, device="cuda") * 0.1
    v = y + b
    # p = 0: plain bias + ReLU
    h0, mask0 = kernels.bias_relu_dropout(y, b, 0.0, None)
    assert torch.equal(torch.from_numpy(_bits(mask0, n)), (v > 0).cpu())
    assert torch.equal(h0, torch.relu(v).to(torch.bfloat16))
    p, scale = 0.5, 2.0
    seed = torch.tensor([12345], dtype=torch.int64, device="cuda")
    h, mask = kernels.bias_relu_dropout(y, b, p, seed)
    keep = torch.from_numpy(_bits(mask, n)).cuda()
    assert not (keep & (v <= 0)).any()
    frac = keep.sum().item() / (v > 0).sum().item()
    assert abs(frac - (1 - p)) < 0.02, frac
    assert torch.equal(h, torch.where(keep, v * scale, torch.zeros_like(v)).to(torch.bfloat16))
    assert seed.item() == 12346
    h2, mask2 = kernels.bias_relu_dropout(y, b, p, seed)
    assert not torch.equal(mask2, mask) and seed.item() == 12347
    # backward
    g = torch.randn(m, n, device="cuda").to(torch.bfloat16)
    gy, db = kernels.bias_relu_dropout_backward(g, mask, p, True)
    ref = torch.where(keep, g.float() * scale, torch.zeros_like(v))
    assert torch.equal(gy, ref.to(torch.bfloat16))
    db64 = ref.double().sum(0)
    assert ((db.double() - db64).abs().max() / db64.abs().max()).item() <= 1e-5
    gy2, db2 = kernels.bias_relu_dropout_backward(g, mask, p, True)
    assert torch.equal(gy2, gy) and torch.equal(db2, db)
    # the same gradient written at the gather pitch
    gyp, dbp = kernels.bias_relu_dropout_backward(g, mask, p, True, pitched=True)
    assert gyp.stride(0) == kernels.row_pitch(n, 2) and torch.equal(gyp, gy) and torch.equal(dbp, db)
    # no rows: a zero bias gradient; N not a multiple of 8 is refused
    _, dbz = kernels.bias_relu_dropout_backward(g[:0], mask[:0], p, True)
    assert torch.equal(dbz, torch.zeros(n, device="cuda"))
    with pytest.raises(RuntimeError):
        kernels.bias_relu_dropout(torch.randn(4, 12, device="cuda"), None, 0.0, None)


@pytest.mark.parametrize("fin,fout", [(40, 96), (96, 40)])
def test_fused_activation_layer_matches_fp64(graph512, fin, fout):
    """GCNConv.forward_act with p = 0 — aggregating first (40 → 96) or after x·W
    (96 → 40) — equals relu(Â·x·W + b) within bf16 rounding (1e-2 of FP64), and
    its weight and bias gradients match the FP64 closed forms through the
    forward's ReLU mask."""
    g, *_ = graph512
    gt = dev(g.rowptr, g.colind, g.colptr, g.rowind, g.value_csr, g.value_csc)
    torch.manual_seed(4)
    conv = GCNConv(fin, fout, precision="bf16", order="auto").cuda()
    with torch.no_grad():
        conv.bias.uniform_(-0.2, 0.2)
    assert conv.fuses_activation() and conv.aggregates_first() == (fin < fout)
    x = torch.rand(512, fin, device="cuda")
    h = conv.forward_act(x, *gt[:4], *gt[4:], p=0.0)
    assert h.dtype == torch.bfloat16
    G = torch.randn(512, fout, device="cuda").to(torch.bfloat16)
    h.backward(G)
    ah = _gcn_dense_fp64(g)
    z = ah @ x.double().cpu().numpy()
    wd, bd = conv.weight.detach().double().cpu().numpy(), conv.bias.detach().double().cpu().numpy()
    pre = z @ wd + bd
    ref = np.maximum(pre, 0)
    assert np.max(oracle.row_rel_err(h.detach().float().cpu().numpy(), ref)) <= 1e-2
    # the backward follows the forward's own ReLU mask (h > 0): where the fp64
    # pre-activation is within bf16 rounding of 0 the two may disagree
    live = h.detach().float().cpu().numpy() > 0
    assert np.mean(live != (pre > 0)) < 0.01
    gd = G.double().cpu().numpy() * live

    def rel(a, b):
        return np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(conv.weight.grad.cpu().numpy(), z.T @ gd) <= 1e-2
    assert rel(conv.bias.grad.cpu().numpy(), gd.sum(0)) <= 1e-2


def test_graphed_gcn_step_with_fused_activation():
    """The bf16 step with a first layer that aggregates first and fuses its
    epilogue, captured in a HIP graph: with dropout 0 every replay equals an
    eager step of an identical model; with dropout 0.5 replays draw fresh masks
    (the losses differ) and stay finite."""
    print("[case] graphed GCN step, fused activation", flush=True)
    torch.manual_seed(0)
    a = gcn.make_task("bf16", n=4000, pairs=9000, features=32, hidden=96, dropout=0.0, capturable=True)
    a = gcn.retarget(a, "bf16", capturable=True, order="auto")
    assert a.model.conv1.fuses_activation()
    b = gcn.retarget(a, "bf16", capturable=True, order="auto")
    b.model.load_state_dict(a.model.state_dict())
    step = gcn.GraphedStep(a, warmup=3)
    for _ in range(3):
        gcn.train_step(b)
    for i in range(4):
        step.replay()
        eager_loss = gcn.train_step(b)
        torch.cuda.synchronize()
        assert abs(step.loss() - eager_loss) <= 1e-6 * max(1.0, abs(eager_loss)), i
    c = gcn.retarget(a, "bf16", capturable=True, order="auto")
    c.model.p = 0.5
    step_c = gcn.GraphedStep(c, warmup=2)
    s0 = c.model.conv1._seed.item()
    losses = []
    for _ in range(4):
        step_c.replay()
        losses.append(step_c.loss())
    assert c.model.conv1._seed.item() == s0 + 4  # one fresh mask per replay
    assert all(np.isfinite(losses)) and len(set(losses)) == 4, losses


def test_graphed_gcn_step_matches_eager():
    """The bf16 training step captured in a HIP graph: five replays, each
    followed by an eager step of an identically initialised model and a
    synchronisation, must give the eager step's loss and parameters every time
    (dropout 0: the graph's and the eager step's random streams differ)."""
    print("[case] graphed GCN step vs eager", flush=True)
    torch.manual_seed(0)
    a = gcn.make_task("bf16", n=4000, pairs=9000, features=120, hidden=32, dropout=0.0, capturable=True)
    b = gcn.retarget(a, "bf16", capturable=True)
    b.model.load_state_dict(a.model.state_dict())
    step = gcn.GraphedStep(a, warmup=3)  # three eager steps of `a`, then the capture
    for _ in range(3):
        gcn.train_step(b)
    for i in range(5):
        step.replay()
        eager_loss = gcn.train_step(b)
        torch.cuda.synchronize()
        assert abs(step.loss() - eager_loss) <= 1e-6 * max(1.0, abs(eager_loss)), i
        for (na, pa), (nb, pb) in zip(a.model.named_parameters(), b.model.named_parameters()):
            d = (pa - pb).abs().max().item()
            assert d <= 1e-6 * max(1.0, pb.abs().max().item()), (i, na, d)


def test_workspace_reserve_keeps_growing():
    r1 = kernels.workspace_reserve(1 << 20)
    r2 = kernels.workspace_reserve(0)
    r3 = kernels.workspace_reserve(3 << 20)
    assert r1 >= 1 << 20 and r2 == r1 and r3 >= 3 << 20


def test_library_calls_replay_in_a_hip_graph():
    """Merge-path SpMM (partition + carries), a bf16-output reduction GEMM (two
    workspaces held at once) and the column sum, captured in one HIP graph:
    new inputs are copied in before each of four replays, with eager library work
    and a synchronisation between replays, and every replay must equal the eager
    calls on the same inputs bit for bit."""
    print("[case] library calls in a HIP graph", flush=True)
    from armsparse_amd import kernels as K
    a = data.rmat_csr(13, 1 << 16, seed=3)
    rp, ci, va = dev(a.rowptr, a.colind, a.values)
    r = torch.Generator(device="cuda").manual_seed(1)
    X = torch.empty((a.shape[1], 64), device="cuda")
    GA = torch.empty((96, 20000), device="cuda")
    GB = torch.empty((20000, 48), device="cuda")
    T = torch.empty((9000, 40), device="cuda")

    def fill():
        for t in (X, GA, GB, T):
            t.normal_(generator=r)

    def work():
        y = K.spmm(rp, ci, va, X, config=dict(schedule="merge"))
        g = K.gemm_bf16(GA.to(torch.bfloat16), GB.to(torch.bfloat16), out_dtype=torch.bfloat16)
        c = K.colsum(T)
        return y, g, c

    fill()
    work()  # eager warm-up: the workspace high-water mark
    torch.cuda.synchronize()
    K.workspace_reserve(0)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph, stream=side):
            static = work()
    torch.cuda.current_stream().wait_stream(side)
    for i in range(4):
        fill()
        graph.replay()
        torch.cuda.synchronize()
        eager = work()
        K.spmm(rp, ci, va, X)  # unrelated eager library work between replays
        torch.cuda.synchronize()
        for s_, e_ in zip(static, eager):
            assert torch.equal(s_, e_), i


@pytest.mark.parametrize("m,n", [(3001, 256), (1000, 100), (77, 64), (5, 3)])
def test_to_bf16_pitched_rows(m, n):
    """The pitched fp32 → bf16 conversion equals the flat one bit for bit, from a
    contiguous or a strided source, with rows at row_pitch."""
    print(f"[case] to_bf16 pitched {m}x{n}", flush=True)
    torch.manual_seed(m + n)
    x = torch.randn(m, n + 5, device="cuda")[:, :n]
    y = kernels.to_bf16(x, pitched=True)
    assert y.shape == (m, n) and y.stride() == (kernels.row_pitch(n, 2), 1)
    assert torch.equal(y, kernels.to_bf16(x.contiguous()))
    assert torch.equal(y, x.to(torch.bfloat16))

