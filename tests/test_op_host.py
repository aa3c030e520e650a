"""Host-side checks of the Python op's layer options (no GPU calls)."""
import pytest

from armsparse_amd.op import GCNConv


def test_gcnconv_order_rules():
    assert not GCNConv(100, 256).aggregates_first()  # fp32 default: the reference's order
    assert not GCNConv(100, 256, precision="bf16").aggregates_first()  # order defaults to transform_first
    assert GCNConv(100, 256, precision="bf16", order="auto").aggregates_first()
    assert not GCNConv(500, 64, precision="bf16", order="auto").aggregates_first()
    assert not GCNConv(64, 64, precision="bf16", order="auto").aggregates_first()
    assert GCNConv(500, 64, fused=True, order="aggregate_first").aggregates_first()
    with pytest.raises(ValueError, match="fused"):
        GCNConv(100, 256, order="auto")  # fp32 without the fused aggregation
    with pytest.raises(ValueError, match="order"):
        GCNConv(100, 256, precision="bf16", order="sideways")


def test_fused_activation_rules():
    assert GCNConv(500, 64, precision="bf16").fuses_activation()  # transform-first hidden layer
    assert GCNConv(100, 256, precision="bf16", order="auto").fuses_activation()  # aggregate-first
    assert not GCNConv(64, 3, precision="bf16").fuses_activation()  # 3 outputs: no 8-column groups
    assert not GCNConv(500, 64).fuses_activation()  # fp32 keeps the reference's separate ops
    # the fused backward's bias-gradient column sums hold at most 2048 columns
    assert GCNConv(16, 2048, precision="bf16").fuses_activation()
    assert not GCNConv(16, 2056, precision="bf16").fuses_activation()


def test_line_pitched_features():
    import torch
    from armsparse_amd.gcn import line_pitched
    x = torch.arange(6 * 100, dtype=torch.float32).reshape(6, 100).to(torch.bfloat16)
    v = line_pitched(x)
    # 200-byte rows: two lines, padded to an odd number of lines (three, 384 bytes)
    assert v.shape == (6, 100) and v.stride() == (192, 1) and torch.equal(v, x)
    assert v.untyped_storage().nbytes() == 6 * 192 * 2
    base = v.untyped_storage()
    pad = torch.tensor([], dtype=torch.bfloat16).set_(base, 0, (6, 192), (192, 1))[:, 100:]
    assert not pad.any()  # the padding is zero
    y = torch.zeros(3, 64, dtype=torch.bfloat16)
    assert line_pitched(y) is y  # exactly one 128-byte line a row


def test_row_pitch_rule():
    """Rows shorter than a line stay packed; longer rows start on 128-byte lines and
    span an odd number of them."""
    from armsparse_amd.kernels import row_pitch
    cases = {(3, 2): 3, (3, 4): 3, (31, 4): 31, (64, 2): 64, (32, 4): 32,   # packed or one line
             (100, 2): 192, (128, 2): 192, (64, 4): 96,                      # 2 lines -> 3
             (128, 4): 160, (256, 2): 320,                                   # 4 lines -> 5
             (256, 4): 288, (512, 2): 576, (160, 4): 160, (100, 4): 160,     # 8 -> 9; 5 stays; 400 B -> 4 -> 5 lines
             (0, 4): 0}
    for (n, esz), want in cases.items():
        assert row_pitch(n, esz) == want, (n, esz)
        p = row_pitch(n, esz)
        assert p >= n
        if n * esz >= 128:
            assert (p * esz) % 128 == 0 and ((p * esz) // 128 == 1 or (p * esz) // 128 % 2 == 1)


def test_adam_refuses_host_tensors_and_bad_hyper_parameters():
    """The library's Adam has no CPU path: a step over host tensors raises before any
    library call; hyper-parameters outside torch.optim.Adam's ranges are rejected."""
    import torch

    from armsparse_amd import optim

    p = torch.nn.Parameter(torch.zeros(4))
    p.grad = torch.ones(4)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        optim.Adam([p]).step()
    for kw in (dict(lr=-1.0), dict(eps=-1e-8), dict(betas=(0.9, 1.0)), dict(betas=(-0.1, 0.9)),
               dict(weight_decay=-1e-4)):
        with pytest.raises(ValueError):
            optim.Adam([p], **kw)
    q = torch.nn.Parameter(torch.zeros(3))  # no gradient: skipped, nothing to do
    optim.Adam([q]).step()
    assert optim.Adam([p]).defaults["capturable"]
