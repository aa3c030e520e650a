"""Torch-tensor front end of libarmsparse: every operator is a HIP kernel of
the library launched on torch's current stream through the C ABI.

PyTorch is used only for device memory and streams.  Tensors on the CPU are
staged to the current ROCm device and the result is copied back (the same
contract as the reference's CPU op, gcn/pytorch-custom/spmm.cpp:31-47);
without a ROCm device every operator raises.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Optional, Union

import torch

from . import _lib
from ._lib import Config, call

ConfigLike = Union[None, str, dict, Config]

# Optional schedule selector: f(op, rowptr(host-or-device), n_rows, nnz, n_dense) -> Config | None.
_selector: Optional[Callable[..., Optional[Config]]] = None


def set_selector(fn: Optional[Callable[..., Optional[Config]]]) -> None:
    """Install (or clear) the automatic schedule selector used when config is None."""
    global _selector
    _selector = fn


def _config(cfg: ConfigLike, op: str, rowptr, m: int, nnz: int, n: int):
    if cfg is None and _selector is not None:
        cfg = _selector(op, rowptr, m, nnz, n)
    if cfg is None:
        return None
    if isinstance(cfg, Config):
        return ctypes.byref(cfg)
    if isinstance(cfg, str):
        return ctypes.byref(Config.make(cfg))
    if isinstance(cfg, dict):
        return ctypes.byref(Config.make(**cfg))
    raise TypeError(f"config must be None, a schedule name, a dict or a Config, got {type(cfg)}")


def _require(cond: bool, msg: str) -> None:
    if not cond:
        raise RuntimeError(msg)


def _device() -> torch.device:
    _require(torch.cuda.is_available(), "armsparse_amd: no ROCm device available (no CPU fallback)")
    return torch.device("cuda", torch.cuda.current_device())


def _on_device(t: Optional[torch.Tensor], dev: torch.device) -> Optional[torch.Tensor]:
    if t is None:
        return None
    return t if t.device == dev else t.to(dev)


def _ptr(t: Optional[torch.Tensor]):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def _stream(dev: torch.device):
    """torch's current stream on `dev` as a raw handle (the direct accessor when this
    torch has it: about 5 us less host time per call than torch.cuda.current_stream)."""
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    if _raw_stream is not None:
        return ctypes.c_void_p(_raw_stream(idx))
    return ctypes.c_void_p(torch.cuda.current_stream(idx).cuda_stream)


def _check_csr(fn: str, rowptr: torch.Tensor, colind: torch.Tensor, values: Optional[torch.Tensor]) -> None:
    _require(rowptr.dtype == torch.int32, f"{fn}: rowptr must be int32")
    _require(colind.dtype == torch.int32, f"{fn}: colind must be int32")
    _require(rowptr.dim() == 1 and rowptr.numel() >= 1, f"{fn}: rowptr must be 1-D with M+1 entries")
    _require(rowptr.is_contiguous() and colind.is_contiguous(), f"{fn}: index arrays must be contiguous")
    if values is not None:
        _require(values.dtype == torch.float32, f"{fn}: values must be float32")
        _require(values.is_contiguous(), f"{fn}: values must be contiguous")
        _require(values.numel() == colind.numel(), f"{fn}: values and colind differ in length")


# ----------------------------------------------------------------- SpMM
def spmm(rowptr: torch.Tensor, colind: torch.Tensor, values: Optional[torch.Tensor], dense: torch.Tensor, *,
         dense_scale: Optional[torch.Tensor] = None, out: Optional[torch.Tensor] = None, accumulate: bool = False,
         trunc4: bool = False, ncols: Optional[int] = None, config: ConfigLike = None,
         col_scale: Optional[torch.Tensor] = None, row_scale: Optional[torch.Tensor] = None,
         bias: Optional[torch.Tensor] = None, relu: bool = False) -> torch.Tensor:
    """out[M×N] (=, or += with accumulate) A·dense.

    dense: float32, bfloat16, or fp8-e4m3 (torch.float8_e4m3fn or raw uint8
    bytes) with ``dense_scale`` holding one float32 scale per row.  values=None
    means all-ones edge values.  ``ncols`` restricts the update to the first
    ncols columns (the reference's ``irregular``/``rowBaseSpmm`` tails).
    Fused epilogue (armsp_spmm_ex): out[r] = row_scale[r]·Σ val·col_scale[c]·dense[c]
    + bias, then ReLU.
    """
    fn = "spmm"
    _check_csr(fn, rowptr, colind, values)
    _require(dense.dim() == 2, f"{fn}: dense must be 2-D")
    _require(dense.stride(1) == 1, f"{fn}: dense rows must be contiguous")
    m = rowptr.numel() - 1
    k, n = dense.shape
    nnz = colind.numel()
    host = not dense.is_cuda
    dev = _device() if host else dense.device
    rp, ci, va = _on_device(rowptr, dev), _on_device(colind, dev), _on_device(values, dev)
    d = _on_device(dense, dev)
    if out is None:
        _require(not accumulate, f"{fn}: accumulate needs an output tensor")
        c = torch.empty((m, n), dtype=torch.float32, device=dev)
    else:
        _require(out.shape == (m, n) and out.dtype == torch.float32, f"{fn}: out must be float32 [M, N]")
        _require(out.stride(1) == 1, f"{fn}: out rows must be contiguous")
        c = _on_device(out, dev)
    flags = (_lib.ACCUMULATE if accumulate else 0) | (_lib.TRUNC4 if trunc4 else 0)
    nc = n if ncols is None else int(ncols)
    _require(0 <= nc <= n, f"{fn}: ncols out of range")
    cfg = _config(config, "spmm", rowptr, m, nnz, nc)
    s = _stream(dev)
    if col_scale is not None or row_scale is not None or bias is not None or relu:
        ep_t = []
        for name, t, n_expect in (("col_scale", col_scale, k), ("row_scale", row_scale, m), ("bias", bias, nc)):
            if t is not None:
                _require(t.dtype == torch.float32 and t.numel() == n_expect, f"{fn}: {name} must be float32[{n_expect}]")
                t = _on_device(t.contiguous(), dev)
            ep_t.append(t)
        ep = _lib.Epilogue(*(None if t is None else t.data_ptr() for t in ep_t), 1 if relu else 0)
        if d.dtype == torch.float32:
            dt, sc = _lib.DTYPE_F32, None
        elif d.dtype == torch.bfloat16:
            dt, sc = _lib.DTYPE_BF16, None
        elif d.dtype in (torch.uint8, torch.float8_e4m3fn):
            _require(dense_scale is not None, f"{fn}: fp8 dense needs dense_scale")
            dt, sc = _lib.DTYPE_FP8_E4M3, _on_device(dense_scale.contiguous(), dev)
        else:
            raise RuntimeError(f"{fn}: unsupported dense dtype {d.dtype}")
        call("armsp_spmm_ex", m, k, nc, nnz, _ptr(rp), _ptr(ci), _ptr(va), dt, _ptr(d), _ptr(sc), d.stride(0),
             _ptr(c), c.stride(0), flags, ctypes.byref(ep), cfg, s)
    elif d.dtype == torch.float32:
        call("armsp_spmm_f32", m, k, nc, nnz, _ptr(rp), _ptr(ci), _ptr(va), _ptr(d), d.stride(0), _ptr(c),
             c.stride(0), flags, cfg, s)
    elif d.dtype == torch.bfloat16:
        call("armsp_spmm_bf16", m, k, nc, nnz, _ptr(rp), _ptr(ci), _ptr(va), _ptr(d), d.stride(0), _ptr(c),
             c.stride(0), flags, cfg, s)
    elif d.dtype in (torch.uint8, torch.float8_e4m3fn):
        _require(dense_scale is not None and dense_scale.dtype == torch.float32 and dense_scale.numel() == k,
                 f"{fn}: fp8 dense needs a float32 per-row scale of length K")
        sc = _on_device(dense_scale.contiguous(), dev)
        call("armsp_spmm_fp8", m, k, nc, nnz, _ptr(rp), _ptr(ci), _ptr(va), _ptr(d), _ptr(sc), d.stride(0),
             _ptr(c), c.stride(0), flags, cfg, s)
    else:
        raise RuntimeError(f"{fn}: unsupported dense dtype {d.dtype}")
    if host:
        c = c.cpu()
        if out is not None:
            out.copy_(c)
            return out
    return c


# ----------------------------------------------------------------- SpMV
def spmv(rowptr: torch.Tensor, colind: torch.Tensor, values: Optional[torch.Tensor], x: torch.Tensor, *,
         out: Optional[torch.Tensor] = None, accumulate: bool = False, config: ConfigLike = None) -> torch.Tensor:
    """y[M] (=, or += with accumulate) A·x on plain CSR."""
    fn = "spmv"
    _check_csr(fn, rowptr, colind, values)
    _require(x.dim() == 1 and x.dtype == torch.float32 and x.is_contiguous(), f"{fn}: x must be contiguous float32")
    m, nnz = rowptr.numel() - 1, colind.numel()
    host = not x.is_cuda
    dev = _device() if host else x.device
    rp, ci, va, xd = (_on_device(t, dev) for t in (rowptr, colind, values, x))
    if out is None:
        _require(not accumulate, f"{fn}: accumulate needs an output tensor")
        y = torch.empty(m, dtype=torch.float32, device=dev)
    else:
        _require(out.shape == (m,) and out.dtype == torch.float32 and out.is_contiguous(), f"{fn}: bad out")
        y = _on_device(out, dev)
    cfg = _config(config, "spmv", rowptr, m, nnz, 1)
    call("armsp_spmv_f32", m, x.numel(), nnz, _ptr(rp), _ptr(ci), _ptr(va), _ptr(xd), _ptr(y),
         _lib.ACCUMULATE if accumulate else 0, cfg, _stream(dev))
    if host:
        y = y.cpu()
        if out is not None:
            out.copy_(y)
            return out
    return y


def vcsr_spmv(rowptr: torch.Tensor, colind: torch.Tensor, values: torch.Tensor, x: torch.Tensor, y: torch.Tensor, *,
              fused: bool = False, balanced: bool = False) -> torch.Tensor:
    """y += A·x on the zero-padded 1×4-blocked CSR (sparse_dense_vector, SpMV.cpp:155-179). In place.
    Bit-exact with the reference's order by default; ``balanced`` runs the merge-path
    schedule over the same arrays (skewed rows, within 1e-5 per row)."""
    _require(not (fused and balanced), "vcsr_spmv: balanced runs the unfused products only")
    fn = "vcsr_spmv"
    _check_csr(fn, rowptr, colind, values)
    _require(values is not None, f"{fn}: values required")
    _require(x.dtype == torch.float32 and y.dtype == torch.float32 and x.is_contiguous() and y.is_contiguous(),
             f"{fn}: x and y must be contiguous float32")
    m, nnz = rowptr.numel() - 1, colind.numel()
    _require(y.numel() == m, f"{fn}: y must have M entries")
    host = not y.is_cuda
    dev = _device() if host else y.device
    rp, ci, va, xd, yd = (_on_device(t, dev) for t in (rowptr, colind, values, x, y))
    call("armsp_spmv_vcsr_f32", m, nnz, _ptr(rp), _ptr(ci), _ptr(va), _ptr(xd), _ptr(yd),
         (_lib.FUSED_MLA if fused else 0) | (_lib.VCSR_BALANCED if balanced else 0), _stream(dev))
    if host:
        y.copy_(yd.cpu())
    return y


def csr_to_vcsr(rowptr: torch.Tensor, colind: torch.Tensor, values: Optional[torch.Tensor], num_cols: int, *,
                threshold: float = 0.0):
    """(rowptr, colind, values) of the zero-padded 1×4-blocked CSR of
    toVectorAwareCSR (SpMV.cpp:15-38), built on the device from canonical CSR;
    ``threshold`` > 0 stores a block only when some |v| > threshold
    (numaAwareSpMV.cpp:48).  The blocked matrix has ceil(K/4)·4 columns."""
    fn = "csr_to_vcsr"
    _check_csr(fn, rowptr, colind, values)
    m, nnz = rowptr.numel() - 1, colind.numel()
    host = not rowptr.is_cuda
    dev = _device() if host else rowptr.device
    rp, ci, va = (_on_device(t, dev) for t in (rowptr, colind, values))
    out_rp = torch.empty(m + 1, dtype=torch.int32, device=dev)
    padded = ctypes.c_int64(0)
    call("armsp_csr_to_vcsr_device", m, num_cols, nnz, _ptr(rp), _ptr(ci), _ptr(va), float(threshold),
         _ptr(out_rp), None, None, 0, ctypes.byref(padded), _stream(dev))
    n = padded.value
    out_ci = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    out_va = torch.empty(max(n, 1), dtype=torch.float32, device=dev)
    call("armsp_csr_to_vcsr_device", m, num_cols, nnz, _ptr(rp), _ptr(ci), _ptr(va), float(threshold),
         _ptr(out_rp), _ptr(out_ci), _ptr(out_va), out_ci.numel(), ctypes.byref(padded), _stream(dev))
    res = (out_rp, out_ci[:n], out_va[:n])
    return tuple(t.cpu() for t in res) if host else res


def column_blocks(rowptr: torch.Tensor, colind: torch.Tensor, values: Optional[torch.Tensor], num_cols: int,
                  nblocks: int):
    """[(rowptr_x, colind_x, values_x)]: the matrix split into nblocks blocks of
    columns (armsp_csr_column_split), each a CSR matrix over all rows; device tensors."""
    fn = "column_blocks"
    _check_csr(fn, rowptr, colind, values)
    _require(rowptr.is_cuda, f"{fn}: needs device tensors")
    m, nnz = rowptr.numel() - 1, colind.numel()
    dev = rowptr.device
    out_rp = torch.empty(nblocks * (m + 1), dtype=torch.int32, device=dev)
    out_ci = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    out_va = torch.empty(max(nnz, 1), dtype=torch.float32, device=dev) if values is not None else None
    starts = (ctypes.c_int64 * (nblocks + 1))()
    call("armsp_csr_column_split", m, num_cols, nnz, _ptr(rowptr), _ptr(colind), _ptr(values), nblocks,
         _ptr(out_rp), _ptr(out_ci), _ptr(out_va), starts, _stream(dev))
    return [(out_rp[x * (m + 1):(x + 1) * (m + 1)], out_ci[starts[x]:starts[x + 1]],
             None if out_va is None else out_va[starts[x]:starts[x + 1]]) for x in range(nblocks)]


def gather_blocks(num_cols: int, row_bytes: int, budget: int = 4 << 20) -> int:
    """Column blocks for a product whose gathered rows of B (num_cols rows of
    row_bytes) outgrow one XCD's 4 MiB L2: the blocks' rows of B fit the budget
    (1: no split).  Measured on the MI355X, N=3 fp32 over a 2.4M-node graph with
    126M edges: 1986 us in one product, 1347 us in 4 column blocks, 1338 in 8."""
    n = -(-num_cols * row_bytes // budget)
    return max(1, min(63, n))


def spmm_blocks(blocks, dense: torch.Tensor, *, row_scale: Optional[torch.Tensor] = None,
                bias: Optional[torch.Tensor] = None, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """out = row_scale·(Σ_x A_x·dense) + bias over column blocks of A (column_blocks),
    block after block: the first writes out (bias included), the others add to it."""
    first = True
    for rp, ci, va in blocks:
        if not first and ci.numel() == 0:
            continue
        out = spmm(rp, ci, va, dense, out=out, accumulate=not first, row_scale=row_scale,
                   bias=bias if first else None)
        first = False
    return out


# ---------------------------------------------------------------- SDDMM
def sddmm(rows: torch.Tensor, colind: torch.Tensor, U: torch.Tensor, V: torch.Tensor, *,
          scale: Optional[torch.Tensor] = None, coo: bool = False, config: ConfigLike = None) -> torch.Tensor:
    """out[e] = scale[e]·<U[row(e)], V[colind[e]]>; rows is rowptr (CSR) or rowind (COO)."""
    fn = "sddmm"
    _require(rows.dtype == torch.int32 and colind.dtype == torch.int32, f"{fn}: indices must be int32")
    _require(rows.is_contiguous() and colind.is_contiguous(), f"{fn}: indices must be contiguous")
    _require(U.dim() == 2 and V.dim() == 2 and U.shape[1] == V.shape[1], f"{fn}: U, V must be 2-D with equal width")
    _require(U.dtype == V.dtype and U.dtype in (torch.float32, torch.bfloat16), f"{fn}: U, V must be float32 or bf16")
    _require(U.stride(1) == 1 and V.stride(1) == 1, f"{fn}: U, V rows must be contiguous")
    nnz, kd = colind.numel(), U.shape[1]
    m = U.shape[0]
    if coo:
        _require(rows.numel() == nnz, f"{fn}: rowind and colind differ in length")
    else:
        _require(rows.numel() == m + 1, f"{fn}: rowptr must have U.shape[0]+1 entries")
    if scale is not None:
        _require(scale.dtype == torch.float32 and scale.numel() == nnz and scale.is_contiguous(), f"{fn}: bad scale")
    host = not U.is_cuda
    dev = _device() if host else U.device
    rd, ci, ud, vd, sd = (_on_device(t, dev) for t in (rows, colind, U, V, scale))
    out = torch.empty(nnz, dtype=torch.float32, device=dev)
    s = _stream(dev)
    if coo:
        _require(U.dtype == torch.float32, f"{fn}: COO form takes float32")
        call("armsp_sddmm_coo_f32", m, V.shape[0], kd, nnz, _ptr(rd), _ptr(ci), _ptr(sd), _ptr(ud), ud.stride(0),
             _ptr(vd), vd.stride(0), _ptr(out), s)
    else:
        cfg = _config(config, "sddmm", rows, m, nnz, kd)
        name = "armsp_sddmm_csr_f32" if U.dtype == torch.float32 else "armsp_sddmm_csr_bf16"
        call(name, m, V.shape[0], kd, nnz, _ptr(rd), _ptr(ci), _ptr(sd), _ptr(ud), ud.stride(0), _ptr(vd),
             vd.stride(0), _ptr(out), cfg, s)
    return out.cpu() if host else out


# ------------------------------------------------------------ SpGEMM
def spgemm_dense(a: tuple, b: tuple, c: torch.Tensor) -> torch.Tensor:
    """c[M×N] += A·B with A = (rowptr, colind, values) M×K and B K×N in CSR. In place."""
    fn = "spgemm_dense"
    (rp1, ci1, v1), (rp2, ci2, v2) = a, b
    _check_csr(fn, rp1, ci1, v1)
    _check_csr(fn, rp2, ci2, v2)
    _require(c.dim() == 2 and c.dtype == torch.float32 and c.stride(1) == 1, f"{fn}: c must be float32 2-D")
    m, n = c.shape
    k = rp2.numel() - 1
    _require(rp1.numel() == m + 1, f"{fn}: A has {rp1.numel() - 1} rows, c has {m}")
    host = not c.is_cuda
    dev = _device() if host else c.device
    t = [_on_device(x, dev) for x in (rp1, ci1, v1, rp2, ci2, v2, c)]
    call("armsp_spgemm_dense_f32", m, k, n, ci1.numel(), *(_ptr(x) for x in t[:3]), ci2.numel(),
         *(_ptr(x) for x in t[3:6]), _ptr(t[6]), t[6].stride(0), _stream(dev))
    if host:
        c.copy_(t[6].cpu())
    return c


# --------------------------------------------------------- transpose
def csr_transpose(rowptr: torch.Tensor, colind: torch.Tensor, values: Optional[torch.Tensor], num_cols: int):
    """(colptr, rowind, values_t): CSC of A, stable (rows ascend within a column)."""
    fn = "csr_transpose"
    _check_csr(fn, rowptr, colind, values)
    m, nnz = rowptr.numel() - 1, colind.numel()
    host = not rowptr.is_cuda
    dev = _device() if host else rowptr.device
    rp, ci, va = (_on_device(t, dev) for t in (rowptr, colind, values))
    colptr = torch.empty(num_cols + 1, dtype=torch.int32, device=dev)
    rowind = torch.empty(nnz, dtype=torch.int32, device=dev)
    vt = torch.empty(nnz, dtype=torch.float32, device=dev) if values is not None else None
    call("armsp_csr_transpose", m, num_cols, nnz, _ptr(rp), _ptr(ci), _ptr(va), _ptr(colptr), _ptr(rowind), _ptr(vt),
         _stream(dev))
    res = (colptr, rowind, vt)
    return tuple(None if x is None else x.cpu() for x in res) if host else res


def spmm_datareuse(rowptr: torch.Tensor, colind: torch.Tensor, values: Optional[torch.Tensor], flag: torch.Tensor,
                   dense: torch.Tensor, *, out: Optional[torch.Tensor] = None, accumulate: bool = False) -> torch.Tensor:
    """A·dense on the 2-row DataReuse format (data.csr_to_datareuse): entry p of
    row m goes to output row m + flag[p] (DataReuseSpmm)."""
    fn = "spmm_datareuse"
    _require(rowptr.dtype == torch.int32 and colind.dtype == torch.int32, f"{fn}: indices must be int32")
    _require(flag.dtype == torch.uint8 and flag.numel() == colind.numel(), f"{fn}: flag must be uint8, one per entry")
    _require(dense.is_cuda and dense.dim() == 2 and dense.dtype == torch.float32, f"{fn}: dense must be a 2-D float32 "
             "device tensor")
    m, nnz = rowptr.numel() - 1, colind.numel()
    k, n = dense.shape
    d = dense.contiguous()
    if out is None:
        _require(not accumulate, f"{fn}: accumulate needs an output tensor")
        out = torch.empty((m, n), device=dense.device)
    _require(out.shape == (m, n) and out.is_contiguous(), f"{fn}: out must be a contiguous [M, N] tensor")
    call("armsp_spmm_datareuse_f32", m, k, n, nnz, _ptr(rowptr), _ptr(colind), _ptr(values), _ptr(flag), _ptr(d), n,
         _ptr(out), n, _lib.ACCUMULATE if accumulate else 0, _stream(dense.device))
    return out


MAX_INIT = -10000.0  # max_init() of gcn/dgl-custom/binary_reduce_max.cu:22-24


def spmm_max(rowptr: torch.Tensor, colind: torch.Tensor, values: Optional[torch.Tensor], dense: torch.Tensor,
             init: float = MAX_INIT, return_argmax: bool = False):
    """out[r, j] = max(init, max_p values[p]·dense[colind[p], j]) over row r's
    nonzeros (values None: dense rows as they are), the DGL max reducer; with
    return_argmax also the winning nonzero positions (int32, -1 where init won)."""
    fn = "spmm_max"
    _check_csr(fn, rowptr, colind, values)
    _require(dense.dim() == 2 and dense.dtype == torch.float32, f"{fn}: dense must be 2-D float32")
    _require(dense.is_cuda and rowptr.is_cuda, f"{fn}: operands must be device tensors")
    m, nnz = rowptr.numel() - 1, colind.numel()
    k, n = dense.shape
    d = dense.contiguous()
    out = torch.empty((m, n), device=dense.device)
    arg = torch.empty((m, n), dtype=torch.int32, device=dense.device) if return_argmax else None
    call("armsp_spmm_max_f32", m, k, n, nnz, _ptr(rowptr), _ptr(colind), _ptr(values), _ptr(d), n, float(init),
         _ptr(out), n, _ptr(arg), _stream(dense.device))
    return (out, arg) if return_argmax else out


def spmm_max_backward(colind: torch.Tensor, values: Optional[torch.Tensor], grad_out: torch.Tensor,
                      argmax: torch.Tensor, num_src: int) -> torch.Tensor:
    """d dense of spmm_max: grad_out[r, j]·values[p] added to row colind[p] for p = argmax[r, j]."""
    fn = "spmm_max_backward"
    _require(argmax.shape == grad_out.shape and argmax.dtype == torch.int32, f"{fn}: argmax must match grad_out")
    m, n = grad_out.shape
    g = grad_out.contiguous().float()
    db = torch.zeros((num_src, n), device=grad_out.device)
    call("armsp_spmm_max_backward_f32", m, num_src, n, colind.numel(), _ptr(colind), _ptr(values), _ptr(g), n,
         _ptr(argmax.contiguous()), _ptr(db), n, _stream(grad_out.device))
    return db


def coo_to_csr(rows: torch.Tensor, cols: torch.Tensor, values: Optional[torch.Tensor], num_rows: int,
               num_cols: int, drop_zeros: bool = False):
    """(rowptr, colind, values): canonical CSR of unordered COO triplets built on
    the device — (row, column) order, duplicates summed in input order, zero
    entries dropped when drop_zeros (toCSR's `!= 0` rule)."""
    fn = "coo_to_csr"
    _require(rows.dtype == torch.int32 and cols.dtype == torch.int32, f"{fn}: indices must be int32")
    _require(rows.dim() == 1 and rows.shape == cols.shape, f"{fn}: rows and cols must be 1-D of one length")
    _require(values is None or (values.dtype == torch.float32 and values.shape == rows.shape),
             f"{fn}: values must be float32 of the same length")
    nnz = rows.numel()
    host = not rows.is_cuda
    dev = _device() if host else rows.device
    r, c, v = (None if t is None else _on_device(t, dev).contiguous() for t in (rows, cols, values))
    rowptr = torch.empty(num_rows + 1, dtype=torch.int32, device=dev)
    colind = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    val = torch.empty(max(nnz, 1), dtype=torch.float32, device=dev) if values is not None else None
    nout = torch.empty(1, dtype=torch.int32, device=dev)
    call("armsp_coo_to_csr", num_rows, num_cols, nnz, _ptr(r), _ptr(c), _ptr(v), int(bool(drop_zeros)),
         _ptr(rowptr), _ptr(colind), _ptr(val), _ptr(nout), _stream(dev))
    n = int(nout.item())
    res = (rowptr, colind[:n], val[:n] if val is not None else None)
    return tuple(None if x is None else x.cpu() for x in res) if host else res


# ------------------------------------------------------- dense helpers
def gemm_bf16(a: torch.Tensor, b: torch.Tensor, *, out: Optional[torch.Tensor] = None,
              accumulate: bool = False, out_dtype: torch.dtype = torch.float32) -> torch.Tensor:
    """out (=|+=) a @ b for 2-D bfloat16 tensors of any strides on MFMA with fp32
    accumulation; the result is float32, or bfloat16 (rounded) with out_dtype."""
    fn = "gemm_bf16"
    _require(a.dim() == 2 and b.dim() == 2 and a.shape[1] == b.shape[0], f"{fn}: shape mismatch")
    _require(a.dtype == torch.bfloat16 and b.dtype == torch.bfloat16, f"{fn}: operands must be bfloat16")
    _require(a.is_cuda and b.is_cuda, f"{fn}: operands must be device tensors")
    _require(out_dtype in (torch.float32, torch.bfloat16), f"{fn}: out_dtype must be float32 or bfloat16")
    m, kd = a.shape
    n = b.shape[1]
    if out is None:
        _require(not accumulate, f"{fn}: accumulate needs an output tensor")
        out = torch.empty((m, n), dtype=out_dtype, device=a.device)
    _require(out.dtype == out_dtype and out.shape == (m, n) and out.stride(1) == 1, f"{fn}: bad out")
    _require(not (accumulate and out_dtype == torch.bfloat16), f"{fn}: accumulate needs a float32 output")
    call("armsp_gemm_bf16_ex", m, n, kd, _ptr(a), a.stride(0), a.stride(1), _ptr(b), b.stride(0), b.stride(1),
         _ptr(out), _lib.DTYPE_BF16 if out_dtype == torch.bfloat16 else _lib.DTYPE_F32, out.stride(0),
         1 if accumulate else 0,
         _stream(a.device))
    return out


def bias_relu_dropout(y: torch.Tensor, bias: Optional[torch.Tensor], p: float,
                      seed: Optional[torch.Tensor]) -> tuple:
    """(h, mask): h = bf16(dropout_p(relu(y + bias))) for a 2-D float32 device
    tensor y [M×N] (N % 8 == 0, unit column stride), mask one bit per element
    (uint8 [M, N/8]: ReLU passed and kept).  ``seed`` is a one-element int64
    device tensor that the call advances when p > 0 (armsp_bias_relu_dropout_bf16)."""
    fn = "bias_relu_dropout"
    _require(y.is_cuda and y.dtype == torch.float32 and y.dim() == 2 and y.stride(1) == 1, f"{fn}: needs a 2-D "
             "float32 device tensor with unit column stride")
    m, n = y.shape
    _require(n % 8 == 0, f"{fn}: the column count must be a multiple of 8")
    _require(0.0 <= p < 1.0, f"{fn}: p must be in [0, 1)")
    if bias is not None:
        _require(bias.is_cuda and bias.dtype == torch.float32 and bias.numel() == n, f"{fn}: bad bias")
        bias = bias.contiguous()
    if p > 0.0:
        _require(seed is not None and seed.is_cuda and seed.dtype == torch.int64 and seed.numel() == 1,
                 f"{fn}: dropout needs a one-element int64 device seed tensor")
    h = torch.empty((m, n), dtype=torch.bfloat16, device=y.device)
    mask = torch.empty((m, n // 8), dtype=torch.uint8, device=y.device)
    call("armsp_bias_relu_dropout_bf16", m, n, _ptr(y), y.stride(0), _ptr(bias), float(p),
         _ptr(seed) if p > 0.0 else None, _ptr(h), _ptr(mask), _stream(y.device))
    return h, mask


def bias_relu_dropout_backward(g: torch.Tensor, mask: torch.Tensor, p: float, want_dbias: bool,
                               pitched: bool = False) -> tuple:
    """(gy, dbias) of bias_relu_dropout for the bf16 gradient g of h: gy =
    bf16(mask ? g/(1-p) : 0), dbias = the fp32 column sums of the same values
    (deterministic), or None.  ``pitched``: gy's rows at row_pitch (for an SpMM
    that gathers them)."""
    fn = "bias_relu_dropout_backward"
    g = g.contiguous()
    _require(g.is_cuda and g.dtype == torch.bfloat16 and g.dim() == 2, f"{fn}: needs a 2-D bfloat16 device tensor")
    m, n = g.shape
    _require(mask.dtype == torch.uint8 and mask.shape == (m, n // 8) and mask.is_contiguous(), f"{fn}: bad mask")
    gy = empty_rows(m, n, torch.bfloat16, g.device) if pitched else torch.empty_like(g)
    dbias = torch.empty(n, device=g.device) if want_dbias else None
    call("armsp_bias_relu_dropout_backward_bf16", m, n, _ptr(g), _ptr(mask), float(p), _ptr(gy), gy.stride(0),
         _ptr(dbias), _stream(g.device))
    return gy, dbias


def colsum(x: torch.Tensor) -> torch.Tensor:
    """x.sum(0) for a 2-D float32 device tensor, deterministic (armsp_colsum_f32)."""
    _require(x.is_cuda and x.dtype == torch.float32 and x.dim() == 2, "colsum: needs a 2-D float32 device tensor")
    x = x.contiguous()
    out = torch.empty(x.shape[1], device=x.device)
    call("armsp_colsum_f32", x.shape[0], x.shape[1], _ptr(x), _ptr(out), _stream(x.device))
    return out


def workspace_reserve(nbytes: int = 0, device: Optional[torch.device] = None) -> int:
    """Reserve the capture arena of `device` (the current one by default): at
    least nbytes and at least the largest workspace used so far."""
    r = ctypes.c_size_t(0)
    with torch.cuda.device(device if device is not None else torch.cuda.current_device()):
        call("armsp_workspace_reserve", ctypes.c_size_t(nbytes), ctypes.byref(r))
    return r.value


def to_bf16(x: torch.Tensor, pitched: bool = False, rows16: bool = False) -> torch.Tensor:
    """float32 → bfloat16 (round to nearest even) with the library kernel;
    ``pitched`` (2-D x with unit column stride): the result's rows at row_pitch,
    for an SpMM that gathers them; ``rows16``: rows padded to a multiple of 16
    bytes, so the GEMM reads them with 16-byte loads."""
    _require(x.is_cuda and x.dtype == torch.float32, "to_bf16: needs a float32 device tensor")
    if pitched or rows16:
        _require(x.dim() == 2 and x.stride(1) == 1, "to_bf16: pitched output needs a 2-D x with unit column stride")
        m, n = x.shape
        if pitched:
            y = empty_rows(m, n, torch.bfloat16, x.device)
        else:
            p = (n + 7) // 8 * 8
            y = torch.empty((m, p), dtype=torch.bfloat16, device=x.device)[:, :n] if p != n else \
                torch.empty((m, n), dtype=torch.bfloat16, device=x.device)
        call("armsp_convert_rows_f32_to_bf16", m, n, _ptr(x), x.stride(0), _ptr(y), y.stride(0), _stream(x.device))
        return y
    x = x.contiguous()
    y = torch.empty(x.shape, dtype=torch.bfloat16, device=x.device)
    call("armsp_convert_f32_to_bf16", x.numel(), _ptr(x), _ptr(y), _stream(x.device))
    return y


def row_pitch(n: int, element_size: int) -> int:
    """Row pitch (in elements) for an operand whose rows an SpMM or SDDMM gathers.
    Rows shorter than a 128-byte line stay packed (several share a line).  Longer
    rows start on 128-byte lines, and a row of two or more lines spans an odd
    number of them.  Measured on the MI355X (1M R-MAT, 2^24 nnz): rows on a pitch
    that is a multiple of 256 bytes gather 11-24% slower (fp32 N=128: 1011 us at
    512 bytes, 841 us at 640; fp32 N=256: 2240 us at 1024 bytes, 1728 us at 1152;
    bf16 N=256: 1091 us at 512 bytes, 968 us at 640)."""
    _require(n >= 0 and element_size in (1, 2, 4, 8), "row_pitch: bad arguments")
    if n * element_size < 128:
        return n
    lines = (n * element_size + 127) // 128
    if lines >= 2 and lines % 2 == 0:
        lines += 1
    return lines * 128 // element_size


def empty_rows(m: int, n: int, dtype: torch.dtype, device) -> torch.Tensor:
    """An uninitialised [m×n] view with rows at row_pitch (the padding is never read)."""
    esz = dtype.itemsize
    p = row_pitch(n, esz)
    if p == n:
        return torch.empty((m, n), dtype=dtype, device=device)
    return torch.empty((m, p), dtype=dtype, device=device)[:, :n]


def quantize_fp8_rows(b: torch.Tensor):
    """(bytes uint8 [K×N], scale float32 [K]) — per-row OCP e4m3 quantisation."""
    _require(b.is_cuda and b.dtype == torch.float32 and b.dim() == 2 and b.stride(1) == 1,
             "quantize_fp8_rows: needs a float32 device matrix with contiguous rows")
    k, n = b.shape
    q = torch.empty((k, n), dtype=torch.uint8, device=b.device)
    sc = torch.empty(k, dtype=torch.float32, device=b.device)
    if k and n:
        _require(b.stride(0) == n, "quantize_fp8_rows: rows must be packed")
        call("armsp_quantize_fp8_rows", k, n, _ptr(b), b.stride(0), _ptr(q), _ptr(sc), _stream(b.device))
    return q, sc


def dequantize_fp8_rows(q: torch.Tensor, scale: Optional[torch.Tensor] = None) -> torch.Tensor:
    """float32 [K×N] = e4m3(q)·scale[:, None] with the device's packed fp8 conversion."""
    _require(q.is_cuda and q.dtype == torch.uint8 and q.dim() == 2 and q.stride(1) == 1,
             "dequantize_fp8_rows: needs a uint8 device matrix with contiguous rows")
    k, n = q.shape
    out = torch.empty((k, n), dtype=torch.float32, device=q.device)
    if scale is not None:
        _require(scale.is_cuda and scale.dtype == torch.float32 and scale.numel() == k and scale.is_contiguous(),
                 "dequantize_fp8_rows: bad scale")
    if k and n:
        call("armsp_dequantize_fp8_rows", k, n, _ptr(q), q.stride(0), _ptr(scale), _ptr(out), n, _stream(q.device))
    return out


def fill_uniform_(t: torch.Tensor, seed: int, stream_id: int, lo: float = -1.0, hi: float = 1.0) -> torch.Tensor:
    _require(t.is_cuda and t.dtype == torch.float32 and t.is_contiguous(), "fill_uniform_: float32 device tensor")
    call("armsp_fill_uniform_f32", t.numel(), seed, stream_id, lo, hi, _ptr(t), _stream(t.device))
    return t


def fill_normal_(t: torch.Tensor, seed: int, stream_id: int) -> torch.Tensor:
    _require(t.is_cuda and t.dtype == torch.float32 and t.is_contiguous(), "fill_normal_: float32 device tensor")
    call("armsp_fill_normal_f32", t.numel(), seed, stream_id, _ptr(t), _stream(t.device))
    return t
