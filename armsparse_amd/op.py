"""Drop-in for gcn/pytorch-custom/op.py: ``SPMMFunction`` and ``GCNConv``
with the reference's signatures, running on libarmsparse.

Differences from the reference, all deliberate:
- the forward output is zero-initialised (spmm.cpp:41 leaves it undefined);
- backward with value-less edges computes Aᵀ·G with all-ones values (the
  reference returns a stale global, op.py:21,32);
- the edge-weight gradient, returned as None by the reference (op.py:31),
  is computed by SDDMM when the CSR edge weights require grad;
- no "[I] Treat edge weight as no_grad." print per backward call;
- ``precision="bf16"`` (extension) stores the dense features in bfloat16:
  x·W runs on the MFMA GEMM and the aggregation reads bf16 features,
  accumulating in fp32.
"""
from __future__ import annotations

import math
from typing import Optional

import torch
from torch.nn import Parameter

from . import kernels


def glorot(t: Optional[torch.Tensor]) -> None:
    """Uniform(-a, a), a = sqrt(6 / (fan_in + fan_out)) — the init op.py:98 takes from PyG."""
    if t is not None:
        a = math.sqrt(6.0 / (t.size(-2) + t.size(-1)))
        with torch.no_grad():
            t.uniform_(-a, a)


def zeros(t: Optional[torch.Tensor]) -> None:
    if t is not None:
        with torch.no_grad():
            t.zero_()


class SPMMFunction(torch.autograd.Function):
    """out = A·feat with A given by (rowptr, colind, edge_weight_csr); backward
    dfeat = Aᵀ·grad via (colptr, rowind, edge_weight_csc) (op.py:8-37)."""

    @staticmethod
    def forward(ctx, rowptr, colind, colptr, rowind, feat, edge_weight_csr=None, edge_weight_csc=None):
        out = kernels.spmm(rowptr, colind, edge_weight_csr, feat)
        ctx.backward_csc = (rowptr, colind, colptr, rowind, feat, edge_weight_csr, edge_weight_csc)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        rowptr, colind, colptr, rowind, feat, ew_csr, ew_csc = ctx.backward_csc
        if ew_csr is not None and ew_csc is None:
            raise RuntimeError(
                "Backward of SPMM require edge values in both src-first and dst-first order. "
                "Call with SPMMFunction.apply(rowptr, colind, colptr, rowind, in_feat, "
                "edge_value_row_first, edge_value_col_first)")
        grad = grad_out.contiguous()
        grad_feat = grad_ew = None
        if ctx.needs_input_grad[4]:
            g = grad if feat.dtype == torch.float32 else grad.to(feat.dtype)
            grad_feat = kernels.spmm(colptr, rowind, ew_csc, g).to(feat.dtype)
        if ctx.needs_input_grad[5] and ew_csr is not None:
            x = feat if feat.dtype == torch.float32 else feat.float()
            grad_ew = kernels.sddmm(rowptr, colind, grad.float(), x.contiguous())
        return None, None, None, None, grad_feat, grad_ew, None


class SPMMMaxFunction(torch.autograd.Function):
    """out[r] = max(init, max over row r's edges of w[e]·feat[colind[e]]) per
    column: the max reducer of the reference's DGL path
    (gcn/dgl-custom/binary_reduce_max.cu:151-166, init -10000 as its
    max_init()).  Backward routes each output's gradient to the feature row
    that won (no edge-weight gradient)."""

    @staticmethod
    def forward(ctx, rowptr, colind, feat, edge_weight=None, init=kernels.MAX_INIT):
        out, arg = kernels.spmm_max(rowptr, colind, edge_weight, feat, init=init, return_argmax=True)
        ctx.save_for_backward(colind, arg)
        ctx.edge_weight = edge_weight
        ctx.num_src = feat.shape[0]
        return out

    @staticmethod
    def backward(ctx, grad_out):
        colind, arg = ctx.saved_tensors
        gf = None
        if ctx.needs_input_grad[2]:
            gf = kernels.spmm_max_backward(colind, ctx.edge_weight, grad_out, arg, ctx.num_src)
        return None, None, gf, None, None


def _rows(x: torch.Tensor) -> torch.Tensor:
    """x with unit column stride; a row pitch wider than the row (feature rows kept on
    128-byte lines) is kept, since every kernel here takes the pitch as an argument."""
    return x if x.dim() == 2 and x.stride(1) == 1 else x.contiguous()


class GCNAggregate(torch.autograd.Function):
    """y = diag(in_norm)·A·diag(out_norm)·h + bias in one fused SpMM (armsp_spmm_ex);
    backward dh = diag(out_norm)·Aᵀ·diag(in_norm)·dy in one SpMM on (colptr, rowind).
    The column scales are folded into per-edge coefficients once per graph:
    coef_fwd[e] = w_csr[e]·out_norm[colind[e]], coef_bwd[e] = w_csc[e]·in_norm[rowind[e]]
    (the same products the kernel would form per nonzero).  h may be bfloat16
    (fp32 accumulation; y and the gradients are fp32)."""

    @staticmethod
    def forward(ctx, h, rowptr, colind, colptr, rowind, coef_fwd, coef_bwd, in_norm, out_norm, bias, blocks=None):
        # blocks: the graph's column blocks (forward, backward) for narrow features,
        # whose rows outgrow the L2 (kernels.spmm_blocks); None: one product
        if blocks is not None:
            y = kernels.spmm_blocks(blocks[0], h, row_scale=in_norm, bias=bias)
        else:
            y = kernels.spmm(rowptr, colind, coef_fwd, h, row_scale=in_norm, bias=bias)
        ctx.blocks = None if blocks is None else blocks[1]
        ctx.graph = (colptr, rowind, coef_bwd, out_norm, h.dtype)
        # learnable edge weights reach coef_fwd (= w_csr·out_norm[colind]); its gradient
        # needs h and the forward pattern
        ctx.edge_grad = (h, rowptr, colind, in_norm) if ctx.needs_input_grad[5] else None
        return y

    @staticmethod
    def backward(ctx, gy):
        colptr, rowind, coef_bwd, out_norm, hdt = ctx.graph
        g = gy.contiguous()
        dh = dcoef = None
        if ctx.needs_input_grad[0]:
            gb = kernels.to_bf16(g.float(), pitched=True) if hdt == torch.bfloat16 else g
            if ctx.blocks is not None:
                dh = kernels.spmm_blocks(ctx.blocks, gb, row_scale=out_norm).to(hdt)
            else:
                dh = kernels.spmm(colptr, rowind, coef_bwd, gb, row_scale=out_norm).to(hdt)
        if ctx.edge_grad is not None:
            dcoef = _edge_coef_grad(g, *ctx.edge_grad)
        dbias = kernels.colsum(gy.float()) if ctx.needs_input_grad[9] else None
        return dh, None, None, None, None, dcoef, None, None, None, dbias, None


def _edge_coef_grad(gy, h, rowptr, colind, in_norm):
    """d loss / d coef[e] of y = diag(in_norm)·A(coef)·h: in_norm[r]·<gy[r], h[c]> over the
    pattern of A (an SDDMM, fp32).  Autograd carries it back to the edge weights through
    the coefficients' products (the SDDMM the unfused SPMMFunction path runs for dA)."""
    g = gy.float()
    if in_norm is not None:
        g = g * in_norm.reshape(-1, 1)
    return kernels.sddmm(rowptr, colind, g.contiguous(), h.float().contiguous())


def gcn_edge_coefficients(colind, rowind, w_csr, w_csc, in_norm, out_norm):
    """Per-edge coefficients of the fused GCN aggregation (see GCNAggregate)."""
    if w_csr is not None and w_csc is None:
        raise RuntimeError("GCNAggregate backward needs the edge values in CSC order as well")
    cf = out_norm[colind.long()] if out_norm is not None else torch.ones(colind.numel(), device=colind.device)
    cb = in_norm[rowind.long()] if in_norm is not None else torch.ones(rowind.numel(), device=rowind.device)
    if w_csr is not None:
        cf = cf * w_csr
        cb = cb * w_csc
    return cf.contiguous(), cb.contiguous()


class _BF16Linear(torch.autograd.Function):
    """y = x·W on the MFMA GEMM with W float32 master weights [F×H] and x [M×F]
    in bfloat16, or float32 (rounded to bfloat16 here, and given a float32
    gradient back, so no conversion round trip sits at the autograd boundary).
    y is bfloat16 when out_bf16 (written so by the GEMM), else float32."""

    @staticmethod
    def forward(ctx, x, weight, out_bf16=False):
        xb = x if x.dtype == torch.bfloat16 else kernels.to_bf16(x.contiguous())
        wb = kernels.to_bf16(weight.detach())
        ctx.save_for_backward(xb, wb)
        ctx.x_dtype = x.dtype
        if out_bf16:  # bf16 features for the aggregation's gathers: rows at the gather pitch
            y = kernels.empty_rows(xb.shape[0], wb.shape[1], torch.bfloat16, xb.device)
            return kernels.gemm_bf16(xb, wb, out=y, out_dtype=torch.bfloat16)
        return kernels.gemm_bf16(xb, wb)

    @staticmethod
    def backward(ctx, gy):
        xb, wb = ctx.saved_tensors
        # rows on 16 bytes: both GEMMs read them with 16-byte loads
        gyb = gy.contiguous() if gy.dtype == torch.bfloat16 else kernels.to_bf16(_rows(gy), rows16=True)
        gx = gw = None
        if ctx.needs_input_grad[0]:
            gx = kernels.gemm_bf16(gyb, wb.t(), out_dtype=ctx.x_dtype)  # [M×H]·[H×F]
        if ctx.needs_input_grad[1]:
            gw = kernels.gemm_bf16(xb.t(), gyb)  # [F×M]·[M×H]
        return gx, gw, None


class _LinearBiasReluDropout(torch.autograd.Function):
    """h = bf16(dropout_p(relu(z·W + b))) — x·W on the MFMA GEMM (fp32 out), then
    the bias, ReLU, dropout and bf16 cast in one pass (armsp_bias_relu_dropout_bf16);
    backward: one pass gives the bf16 gradient of z·W + b and the bias gradient,
    then dW = zᵀ·gy (and dz = gy·Wᵀ when z needs it) on the GEMM."""

    @staticmethod
    def forward(ctx, z, weight, bias, p, seed):
        zb = z if z.dtype == torch.bfloat16 else kernels.to_bf16(_rows(z), rows16=True)  # 16-byte GEMM loads
        wb = kernels.to_bf16(weight.detach())
        y = kernels.gemm_bf16(zb, wb)
        h, mask = kernels.bias_relu_dropout(y, None if bias is None else bias.detach(), p, seed)
        ctx.save_for_backward(zb, wb, mask)
        ctx.p = p
        ctx.z_dtype = z.dtype
        ctx.has_bias = bias is not None
        return h

    @staticmethod
    def backward(ctx, gh):
        zb, wb, mask = ctx.saved_tensors
        gyb, dbias = kernels.bias_relu_dropout_backward(gh.to(torch.bfloat16), mask, ctx.p,
                                                        ctx.has_bias and ctx.needs_input_grad[2])
        gz = gw = None
        if ctx.needs_input_grad[0]:
            gz = kernels.gemm_bf16(gyb, wb.t(), out_dtype=ctx.z_dtype)
        if ctx.needs_input_grad[1]:
            gw = kernels.gemm_bf16(zb.t(), gyb)
        return gz, gw, dbias, None, None


class _AggregateBiasReluDropout(torch.autograd.Function):
    """h = bf16(dropout_p(relu(diag(in_norm)·A·diag(out_norm)·x + bias))) for bf16
    x: the fused SpMM (GCNAggregate's forward without its bias), then bias, ReLU,
    dropout and the cast in one pass; backward: one pass gives the bf16 gradient
    of the pre-activation and the bias gradient, which feed the Aᵀ SpMM directly
    (no fp32 round trip of the gradient)."""

    @staticmethod
    def forward(ctx, h, rowptr, colind, colptr, rowind, coef_fwd, coef_bwd, in_norm, out_norm, bias, p, seed):
        y = kernels.spmm(rowptr, colind, coef_fwd, h, row_scale=in_norm)
        out, mask = kernels.bias_relu_dropout(y, None if bias is None else bias.detach(), p, seed)
        ctx.save_for_backward(mask)
        ctx.graph = (colptr, rowind, coef_bwd, out_norm, h.dtype)
        ctx.edge_grad = (h, rowptr, colind, in_norm) if ctx.needs_input_grad[5] else None
        ctx.p = p
        ctx.has_bias = bias is not None
        return out

    @staticmethod
    def backward(ctx, gout):
        (mask,) = ctx.saved_tensors
        colptr, rowind, coef_bwd, out_norm, hdt = ctx.graph
        gyb, dbias = kernels.bias_relu_dropout_backward(gout.to(torch.bfloat16), mask, ctx.p,
                                                        ctx.has_bias and ctx.needs_input_grad[9],
                                                        pitched=ctx.needs_input_grad[0])
        dh = dcoef = None
        if ctx.needs_input_grad[0]:
            dh = kernels.spmm(colptr, rowind, coef_bwd, gyb, row_scale=out_norm).to(hdt)
        if ctx.edge_grad is not None:
            dcoef = _edge_coef_grad(gyb, *ctx.edge_grad)
        return dh, None, None, None, None, dcoef, None, None, None, dbias, None, None


class _Sub:
    """Context of one of two Functions run as a single autograd node: it holds what
    the inner forward saves and the input-gradient flags its backward reads."""

    def __init__(self, needs_input_grad):
        self.needs_input_grad = needs_input_grad
        self.saved_tensors = ()

    def save_for_backward(self, *tensors):
        self.saved_tensors = tensors


class _LinearAggregate(torch.autograd.Function):
    """GCNAggregate(x·W) as one autograd node (the fused bf16 layer in the
    reference's order): _BF16Linear's forward, then GCNAggregate's; the backward
    runs theirs in reverse.  One node instead of two saves the host time of a
    Python backward call per layer in an eager step."""

    @staticmethod
    def forward(ctx, x, weight, out_bf16, rowptr, colind, colptr, rowind, coef_fwd, coef_bwd, in_norm, out_norm,
                bias, blocks):
        n = ctx.needs_input_grad
        ctx.lin = _Sub((n[0], n[1], False))
        h = _BF16Linear.forward(ctx.lin, x, weight, out_bf16)
        h = h if out_bf16 else h.contiguous()
        ctx.agg = _Sub((n[0] or n[1], False, False, False, False, n[7], False, False, False, n[11], False))
        return GCNAggregate.forward(ctx.agg, h, rowptr, colind, colptr, rowind, coef_fwd, coef_bwd, in_norm,
                                    out_norm, bias, blocks)

    @staticmethod
    def backward(ctx, gy):
        dh, _, _, _, _, dcoef, _, _, _, dbias, _ = GCNAggregate.backward(ctx.agg, gy)
        gx = gw = None
        if dh is not None:
            gx, gw, _ = _BF16Linear.backward(ctx.lin, dh)
        return gx, gw, None, None, None, None, None, dcoef, None, None, None, dbias, None


class _LinearAggregateAct(torch.autograd.Function):
    """_AggregateBiasReluDropout(x·W) as one autograd node (the fused hidden layer
    in the reference's order): _BF16Linear's forward with bf16 output, then the
    aggregation with its bias, ReLU, dropout and cast; the backward runs theirs in
    reverse."""

    @staticmethod
    def forward(ctx, x, weight, rowptr, colind, colptr, rowind, coef_fwd, coef_bwd, in_norm, out_norm, bias, p,
                seed):
        n = ctx.needs_input_grad
        ctx.lin = _Sub((n[0], n[1], False))
        h = _BF16Linear.forward(ctx.lin, x, weight, True)
        ctx.agg = _Sub((n[0] or n[1], False, False, False, False, n[6], False, False, False, n[10], False, False))
        return _AggregateBiasReluDropout.forward(ctx.agg, h, rowptr, colind, colptr, rowind, coef_fwd, coef_bwd,
                                                 in_norm, out_norm, bias, p, seed)

    @staticmethod
    def backward(ctx, gout):
        dh, _, _, _, _, dcoef, _, _, _, dbias, _, _ = _AggregateBiasReluDropout.backward(ctx.agg, gout)
        gx = gw = None
        if dh is not None:
            gx, gw, _ = _BF16Linear.backward(ctx.lin, dh)
        return gx, gw, None, None, None, None, dcoef, None, None, None, dbias, None, None


class GCNConv(torch.nn.Module):
    """GCN layer of gcn/pytorch-custom/op.py:78-153: x·W, scale by
    1/sqrt(out-degree) (counts from colptr), aggregate with SPMMFunction,
    scale by 1/sqrt(in-degree) (counts from rowptr), add bias."""

    def __init__(self, in_channels, out_channels, improved=False, cached=False, bias=True, normalize=True,
                 precision: str = "fp32", fused: Optional[bool] = None, order: str = "transform_first", **kwargs):
        super().__init__()
        if precision not in ("fp32", "bf16"):
            raise ValueError("precision must be 'fp32' or 'bf16'")
        if order not in ("transform_first", "aggregate_first", "auto"):
            raise ValueError("order must be 'transform_first', 'aggregate_first' or 'auto'")
        # fused: normalisation, aggregation and bias in one SpMM launch (default for bf16;
        # the fp32 default keeps the reference's separate tensor ops)
        self.fused = (precision == "bf16") if fused is None else bool(fused)
        # order (fused layers): Â·(x·W) as the reference computes it, or (Â·x)·W, which
        # gathers in_channels instead of out_channels per edge and, when x needs no
        # gradient, runs no SpMM in the backward pass; "auto" aggregates first when
        # in_channels < out_channels
        if order != "transform_first" and not self.fused:
            raise ValueError("order other than 'transform_first' needs the fused aggregation")
        self.order = order
        self.in_channels = in_channels
        self.out_channels = out_channels
        self.improved = improved
        self.cached = cached
        self.normalize = normalize
        self.precision = precision
        self.weight = Parameter(torch.empty(in_channels, out_channels))
        if bias:
            self.bias = Parameter(torch.empty(out_channels))
        else:
            self.register_parameter("bias", None)
        self.reset_parameters()

    def reset_parameters(self):
        glorot(self.weight)
        zeros(self.bias)
        self.cached_result = None
        self.cached_num_edges = None
        self.cached_coef = None
        self.cached_blocks = None  # (graph key, (forward blocks, backward blocks))

    @staticmethod
    def in_deg_sqrt(indptr):
        return (1 / torch.sqrt((indptr[1:] - indptr[:-1]).float())).unsqueeze(dim=1)

    @staticmethod
    def out_deg_sqrt(indptr):
        return (1 / torch.sqrt((indptr[1:] - indptr[:-1]).float())).unsqueeze(dim=1)

    def fuses_activation(self) -> bool:
        """forward_act available: a fused bf16 layer with a multiple of 8 outputs."""
        # (the backward pass's column sums of the bias gradient hold at most 2048 columns)
        return self.precision == "bf16" and self.fused and self.out_channels % 8 == 0 and self.out_channels <= 2048

    def forward_act(self, x, rowptr, colind, colptr, rowind, edge_weight_csr=None, edge_weight_csc=None,
                    p: float = 0.0):
        """dropout_p(relu(self(x, ...))) as a bfloat16 tensor, with the bias, ReLU,
        dropout and cast fused into one pass after the layer's last product (a
        hidden layer of Net; fuses_activation())."""
        if not self.fuses_activation():
            raise RuntimeError("forward_act needs a fused bf16 layer with a multiple of 8 outputs, at most 2048")
        in_n, out_n, coefs = self._norms_and_coefs(rowptr, colind, colptr, rowind, edge_weight_csr,
                                                   edge_weight_csc)
        seed = None
        if p > 0.0:
            if getattr(self, "_seed", None) is None or self._seed.device != x.device:
                self._seed = torch.randint(0, 2**62, (1,), dtype=torch.int64).to(x.device)
            seed = self._seed
        if not self.aggregates_first():  # x·W to bf16 features for the gathers, then the aggregation
            return _LinearAggregateAct.apply(_rows(x), self.weight, rowptr, colind, colptr, rowind, coefs[0],
                                             coefs[1], in_n, out_n, self.bias, float(p), seed)
        h = x
        if h.dtype != torch.bfloat16:
            h = h.to(torch.bfloat16) if h.requires_grad else kernels.to_bf16(h.contiguous())
        z = GCNAggregate.apply(_rows(h), rowptr, colind, colptr, rowind, coefs[0], coefs[1], in_n, out_n, None)
        return _LinearBiasReluDropout.apply(z, self.weight, self.bias, float(p), seed)

    def aggregates_first(self) -> bool:
        return self.fused and (self.order == "aggregate_first" or
                               (self.order == "auto" and self.in_channels < self.out_channels))

    def _norms_and_coefs(self, rowptr, colind, colptr, rowind, edge_weight_csr, edge_weight_csc):
        """Degree norms (1/sqrt of the counts, cached with ``cached``) and the fused
        aggregation's per-edge coefficients."""
        if not self.cached or self.cached_result is None:
            if self.normalize:
                self.cached_result = self.in_deg_sqrt(rowptr), self.out_deg_sqrt(colptr)
            else:
                self.cached_result = None, None
        in_deg_norm, out_deg_norm = self.cached_result
        in_n = in_deg_norm.reshape(-1).contiguous() if self.normalize else None
        out_n = out_deg_norm.reshape(-1).contiguous() if self.normalize else None
        coefs = self.cached_coef if self.cached else None
        if coefs is None:
            coefs = gcn_edge_coefficients(colind, rowind, edge_weight_csr, edge_weight_csc, in_n, out_n)
            if self.cached:
                self.cached_coef = coefs
        return in_n, out_n, coefs

    def _column_blocks(self, h, rowptr, colind, colptr, rowind, coefs):
        """Column blocks of the graph for an aggregation of narrow features (at most 4
        per node) whose rows outgrow the L2, cached with the coefficients (cached=True
        and coefficients that need no gradient); None: one product per aggregation."""
        if h.dim() != 2:
            return None
        return self._column_blocks_for(h.shape[0], h.shape[1], h.element_size(), h.is_cuda, rowptr, colind,
                                       colptr, rowind, coefs)

    def _column_blocks_for(self, rows, cols, esize, is_cuda, rowptr, colind, colptr, rowind, coefs):
        """_column_blocks for features of rows × cols elements of esize bytes."""
        if not (self.cached and cols <= 4 and is_cuda):
            return None
        if coefs[0].requires_grad or coefs[1].requires_grad:
            return None
        row_bytes = cols * esize
        nb = kernels.gather_blocks(rows, row_bytes)
        if nb < 2:
            return None
        key = (rowptr.data_ptr(), colind.data_ptr(), colptr.data_ptr(), rowind.data_ptr(), nb)
        if self.cached_blocks is None or self.cached_blocks[0] != key:
            m = rowptr.numel() - 1
            fwd = kernels.column_blocks(rowptr, colind, coefs[0], rows, nb)
            bwd = kernels.column_blocks(colptr, rowind, coefs[1], m, kernels.gather_blocks(m, row_bytes))
            self.cached_blocks = (key, (fwd, bwd))
        return self.cached_blocks[1]

    def _forward_aggregate_first(self, x, rowptr, colind, colptr, rowind, edge_weight_csr, edge_weight_csc):
        """(diag(in_norm)·A·diag(out_norm)·x)·W + bias: the SpMM gathers in_channels per
        edge; x·W's rounding moves after the aggregation (bf16: x and the aggregate are
        rounded to bfloat16, the sums and products stay fp32)."""
        in_n, out_n, coefs = self._norms_and_coefs(rowptr, colind, colptr, rowind, edge_weight_csr,
                                                   edge_weight_csc)
        h = x
        if self.precision == "bf16" and h.dtype != torch.bfloat16:
            h = h.to(torch.bfloat16) if h.requires_grad else kernels.to_bf16(h.contiguous())
        z = GCNAggregate.apply(_rows(h), rowptr, colind, colptr, rowind, coefs[0], coefs[1], in_n, out_n, None)
        y = _BF16Linear.apply(z, self.weight, False) if self.precision == "bf16" else torch.matmul(z, self.weight)
        return y + self.bias if self.bias is not None else y

    def forward(self, x, rowptr, colind, colptr, rowind, edge_weight_csr=None, edge_weight_csc=None):
        if self.aggregates_first():
            return self._forward_aggregate_first(x, rowptr, colind, colptr, rowind, edge_weight_csr,
                                                 edge_weight_csc)
        if self.precision == "bf16":
            # bf16 features for the aggregation straight from the GEMM (fused path with
            # more than 8 output channels); fp32 otherwise
            out_bf16 = self.fused and self.out_channels > 8
            if self.fused:  # x·W and the aggregation as one autograd node
                in_n, out_n, coefs = self._norms_and_coefs(rowptr, colind, colptr, rowind, edge_weight_csr,
                                                           edge_weight_csc)
                blocks = self._column_blocks_for(x.shape[0], self.out_channels, 2 if out_bf16 else 4, x.is_cuda,
                                                 rowptr, colind, colptr, rowind, coefs)
                return _LinearAggregate.apply(_rows(x), self.weight, out_bf16, rowptr, colind, colptr, rowind,
                                              coefs[0], coefs[1], in_n, out_n, self.bias, blocks)
            x = _BF16Linear.apply(_rows(x), self.weight, out_bf16)
        else:
            x = torch.matmul(x, self.weight)
        if self.fused:
            # bf16 features halve the gather bytes; a handful of output channels gather
            # next to nothing, so those stay fp32 (no conversions either way)
            in_n, out_n, coefs = self._norms_and_coefs(rowptr, colind, colptr, rowind, edge_weight_csr,
                                                       edge_weight_csc)
            x = x.contiguous()
            return GCNAggregate.apply(x, rowptr, colind, colptr, rowind, coefs[0], coefs[1], in_n, out_n,
                                      self.bias, self._column_blocks(x, rowptr, colind, colptr, rowind, coefs))
        if not self.cached or self.cached_result is None:
            if self.normalize:
                in_deg_norm = self.in_deg_sqrt(rowptr)
                out_deg_norm = self.out_deg_sqrt(colptr)
            else:
                in_deg_norm = torch.ones(rowptr.shape[0] - 1, 1, dtype=x.dtype, device=x.device)
                out_deg_norm = torch.ones(colptr.shape[0] - 1, 1, dtype=x.dtype, device=x.device)
            self.cached_result = in_deg_norm, out_deg_norm
        in_deg_norm, out_deg_norm = self.cached_result
        if self.normalize:
            x = x * out_deg_norm
        if self.precision == "bf16":
            x = x.to(torch.bfloat16)
        aggr_out = SPMMFunction.apply(rowptr, colind, colptr, rowind, x.contiguous(), edge_weight_csr,
                                      edge_weight_csc)
        if self.normalize:
            aggr_out = aggr_out * in_deg_norm
        if self.bias is not None:
            aggr_out = aggr_out + self.bias
        return aggr_out

    def __repr__(self):
        return "{}({}, {})".format(self.__class__.__name__, self.in_channels, self.out_channels)
