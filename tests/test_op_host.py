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
