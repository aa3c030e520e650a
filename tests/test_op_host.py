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
    assert v.shape == (6, 100) and v.stride() == (128, 1) and torch.equal(v, x)
    assert v.untyped_storage().nbytes() == 6 * 128 * 2  # rows on 128-byte lines, zero padding
    y = torch.zeros(3, 64, dtype=torch.bfloat16)
    assert line_pitched(y) is y  # already 128 bytes a row
