"""GPU tests of the library's Adam (armsp_adam_f32 through armsparse_amd.optim.Adam):
the first step against its closed form, several steps against torch.optim.Adam
(the same update in fp32 on the same grads), parameters without gradients, lists
longer than one call takes, a step captured in a HIP graph, and the GCN training
step with either optimizer."""
import numpy as np
import pytest
import torch

from armsparse_amd import _lib, gcn, optim

pytestmark = pytest.mark.gpu

SHAPES = [(500, 64), (64,), (64, 3), (3,), (1,), (1025,), (257, 3)]


def _params(seed, shapes=SHAPES):
    g = torch.Generator(device="cuda").manual_seed(seed)
    return [torch.randn(s, device="cuda", generator=g) for s in shapes]


def _twins(seed, shapes=SHAPES, wd=(5e-4, 0.0), lr=0.01):
    """The same parameters under the library's Adam and under torch.optim.Adam,
    split into two groups with different weight decay."""
    base = _params(seed, shapes)
    a = [torch.nn.Parameter(t.clone()) for t in base]
    b = [torch.nn.Parameter(t.clone()) for t in base]
    h = len(base) // 2
    oa = optim.Adam([dict(params=a[:h], weight_decay=wd[0]), dict(params=a[h:], weight_decay=wd[1])], lr=lr)
    ob = torch.optim.Adam([dict(params=b[:h], weight_decay=wd[0]), dict(params=b[h:], weight_decay=wd[1])], lr=lr,
                          foreach=False, fused=False)
    return a, oa, b, ob


def _set_grads(a, b, seed, scale=1.0):
    g = torch.Generator(device="cuda").manual_seed(seed)
    for pa, pb in zip(a, b):
        gr = torch.randn(pa.shape, device="cuda", generator=g) * scale
        pa.grad = gr.clone()
        pb.grad = gr.clone()


def _close(a, oa, b, ob, steps):
    for pa, pb in zip(a, b):
        sa, sb = oa.state[pa], ob.state[pb]
        assert float(sa["step"]) == float(sb["step"]) == steps
        if pa.numel() == 0:
            continue
        # the first moment sums gradients of either sign (and here of magnitudes 0.1 to
        # 10): where they cancel, the two roundings differ by ulps of the terms, not of
        # the sum, so the bound scales with the tensor's largest moment
        ma = sb["exp_avg"].abs().max().item()
        torch.testing.assert_close(sa["exp_avg"], sb["exp_avg"], rtol=1e-5, atol=1e-6 * ma)
        torch.testing.assert_close(sa["exp_avg_sq"], sb["exp_avg_sq"], rtol=1e-5, atol=1e-9)
        # each step moves a parameter by at most about lr = 0.01; the two updates round
        # differently (fused multiply-adds, torch's lerp form of the first moment)
        assert (pa.detach() - pb.detach()).abs().max().item() <= 1e-6


def test_adam_first_step_closed_form():
    """At t = 1 the bias corrections cancel the moments' (1 - beta) factors:
    param -= lr * g' / (|g'| + eps) with g' = grad + wd * param (float64 on the host),
    exp_avg = (1 - beta1) g', exp_avg_sq = (1 - beta2) g'^2."""
    print("[case] adam first step, closed form", flush=True)
    torch.manual_seed(0)
    p0 = _params(3)
    ps = [torch.nn.Parameter(t.clone()) for t in p0]
    opt = optim.Adam([dict(params=ps[:3], weight_decay=1e-2), dict(params=ps[3:], weight_decay=0.0)], lr=0.05,
                     eps=1e-3)
    g = _params(4)
    for p, gr in zip(ps, g):
        p.grad = gr.clone()
    opt.step()
    for i, (p, t0, gr) in enumerate(zip(ps, p0, g)):
        wd = 1e-2 if i < 3 else 0.0
        x, gd = t0.double().cpu().numpy(), gr.double().cpu().numpy()
        gg = gd + wd * x
        want = x - 0.05 * gg / (np.abs(gg) + 1e-3)
        got = p.detach().double().cpu().numpy()
        assert np.max(np.abs(got - want)) <= 1e-6, i
        st = opt.state[p]
        assert float(st["step"]) == 1.0
        np.testing.assert_allclose(st["exp_avg"].double().cpu().numpy(), 0.1 * gg, rtol=1e-5, atol=1e-9)
        np.testing.assert_allclose(st["exp_avg_sq"].double().cpu().numpy(), 1e-3 * gg * gg, rtol=1e-5, atol=1e-12)


@pytest.mark.parametrize("betas,eps", [((0.9, 0.999), 1e-8), ((0.5, 0.9), 1e-3)])
def test_adam_matches_torch_adam(betas, eps):
    print(f"[case] adam vs torch.optim.Adam betas={betas} eps={eps}", flush=True)
    a, oa, b, ob = _twins(11)
    for o in (oa, ob):
        for grp in o.param_groups:
            grp["betas"], grp["eps"] = betas, eps
    for t in range(1, 7):
        _set_grads(a, b, 100 + t, scale=10.0 ** (t % 3 - 1))
        oa.step()
        ob.step()
        torch.cuda.synchronize()
        _close(a, oa, b, ob, t)


def test_adam_skips_parameters_without_gradients():
    """A parameter without a gradient keeps its value, moments and counter; its
    counter then lags the others, and later steps use it, as torch's does."""
    print("[case] adam, a parameter without gradient", flush=True)
    a, oa, b, ob = _twins(21)
    for t in range(1, 5):
        _set_grads(a, b, 200 + t)
        if t in (2, 3):
            a[1].grad = None
            b[1].grad = None
        oa.step()
        ob.step()
    torch.cuda.synchronize()
    assert float(oa.state[a[1]]["step"]) == float(ob.state[b[1]]["step"]) == 2
    for pa, pb in zip(a, b):
        assert float(oa.state[pa]["step"]) == float(ob.state[pb]["step"])
        assert (pa.detach() - pb.detach()).abs().max().item() <= 1e-6


def test_adam_over_more_tensors_than_one_call_takes():
    n = _lib.ADAM_MAX_TENSORS + 5
    print(f"[case] adam over {n} tensors", flush=True)
    shapes = [(7 * k + 1,) for k in range(n)]
    a, oa, b, ob = _twins(31, shapes=shapes)
    for t in range(1, 4):
        _set_grads(a, b, 300 + t)
        oa.step()
        ob.step()
        torch.cuda.synchronize()
        _close(a, oa, b, ob, t)


def test_adam_after_load_state_dict():
    """load_state_dict replaces the moment and counter tensors: the next steps must
    update the new ones (the cached argument array is rebuilt), as torch's do."""
    print("[case] adam after load_state_dict", flush=True)
    a, oa, b, ob = _twins(51)
    for t in (1, 2):
        _set_grads(a, b, 600 + t)
        oa.step()
        ob.step()
    oa.load_state_dict(oa.state_dict())
    ob.load_state_dict(ob.state_dict())
    for t in (3, 4):
        _set_grads(a, b, 600 + t)
        oa.step()
        ob.step()
        torch.cuda.synchronize()
        _close(a, oa, b, ob, t)


def test_adam_step_in_a_graph():
    """optim.Adam.step() captured in a HIP graph: every replay (new gradients
    copied in, eager work and a synchronisation in between) equals an eager step
    of an identical optimizer."""
    print("[case] adam step captured in a graph", flush=True)
    a, oa, b, _ = _twins(41)
    h = len(b) // 2
    ob = optim.Adam([dict(params=b[:h], weight_decay=5e-4), dict(params=b[h:], weight_decay=0.0)], lr=0.01)
    grads = [torch.zeros_like(p) for p in a]
    for p, g in zip(a, grads):
        p.grad = g
    _set_grads(a, b, 400)
    for p, g in zip(a, grads):  # the graph reads the gradients from these buffers
        g.copy_(p.grad)
        p.grad = g
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        oa.step()  # creates the state outside the capture
    torch.cuda.current_stream().wait_stream(side)
    ob.step()
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        oa.step()
    for r in range(4):
        gen = torch.Generator(device="cuda").manual_seed(500 + r)
        for p, q, g in zip(a, b, grads):
            new = torch.randn(p.shape, device="cuda", generator=gen)
            g.copy_(new)
            q.grad = new.clone()
        graph.replay()
        ob.step()
        torch.cuda.synchronize()
        for pa, pb in zip(a, b):
            assert float(oa.state[pa]["step"]) == float(ob.state[pb]["step"]) == r + 2
            assert torch.equal(pa.detach(), pb.detach()), r
            assert torch.equal(oa.state[pa]["exp_avg_sq"], ob.state[pb]["exp_avg_sq"]), r


def test_adam_with_empty_tensors():
    """Zero-element parameters advance their counters (as torch's do) and launch
    no update workgroups; the tensors around them are updated as usual."""
    print("[case] adam with empty tensors", flush=True)
    shapes = [(0,), (300,), (0, 5), (7,)]
    a, oa, b, ob = _twins(61, shapes=shapes)
    for t in (1, 2):
        _set_grads(a, b, 700 + t)
        oa.step()
        ob.step()
        torch.cuda.synchronize()
        _close(a, oa, b, ob, t)


def test_adam_rejects_what_it_cannot_update():
    p = torch.nn.Parameter(torch.zeros(4, device="cuda", dtype=torch.float64))
    p.grad = torch.ones_like(p)
    with pytest.raises(RuntimeError, match="float32"):
        optim.Adam([p]).step()
    with pytest.raises(ValueError):
        optim.Adam([p], betas=(1.0, 0.999))


def test_gcn_training_with_native_and_torch_adam():
    """Five fp32 GCN steps (dropout 0) from the same initial weights with the
    library's Adam and with torch.optim.Adam: the same losses and weights."""
    print("[case] GCN training, native Adam vs torch Adam", flush=True)
    tasks = []
    for kind in ("native", "torch"):
        torch.manual_seed(0)
        tasks.append(gcn.make_task("fp32", n=4000, pairs=9000, features=32, hidden=96, dropout=0.0, optimizer=kind))
    a, b = tasks
    assert isinstance(a.optimizer, optim.Adam) and not isinstance(b.optimizer, optim.Adam)
    b.model.load_state_dict(a.model.state_dict())
    la = [gcn.train_step(a) for _ in range(5)]
    lb = [gcn.train_step(b) for _ in range(5)]
    assert la[-1] < la[0]
    assert max(abs(p - q) for p, q in zip(la, lb)) <= 1e-5 * max(abs(q) for q in lb), (la, lb)
    for (na, pa), (_, pb) in zip(a.model.named_parameters(), b.model.named_parameters()):
        d = (pa - pb).abs().max().item()
        assert d <= 1e-5, (na, d)
